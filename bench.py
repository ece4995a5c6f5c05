"""Benchmark of the simulation step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c2|c0]
                    [--cpu-steps C] [--no-cpu]
    python bench.py --workload c3 [--c3-nets 512] [--c3-solo 32]   (batched sweep)
    python bench.py --workload c4 [--c4-runs 4096] [--c4-cands 1000000]   (rate surrogate)
    python bench.py --workload mlp [--steps 2000]   (surrogate training kernels alone)
    python bench.py --copy-probe              (calibration copy kernel for PMC runs)
    python bench.py --pmc-summary DIR [--kernel REGEX]

A "step" is one simulation cycle of the whole network (the reference's
SGSim loop body: LIF update, spike compaction, delayed delivery, synaptic
accumulation).  The default workload is the standard feed-forward network
at ~100 Hz (64 x 10,000 LIF, ~6.3e8 synapses, delays 1..6 steps, dt 0.1 ms),
built from the seeded generator of bsim_amd.netspec and uploaded before the
timed region.  The network is first pre-rolled a fixed number of cycles
(--preroll, default 1,000 = 0.1 s simulated, independent of --warmup) so that
the timed cycles are in the configured firing regime rather than the cold
start; then W cycles run untimed and exactly K cycles are timed between two
stream synchronisations.  One JSON line is printed:

  value        synaptic events (deliveries) per second over the K cycles
  roofline     the step kernel's algorithmic bytes (8 B per delivered event:
               fp32 weight + int32 target, as the reference stores them) per
               launch over its mean launch duration, measured with HIP events
               recorded by the library around each launch on the simulation's
               stream during --roofline-steps (200) single launches right
               after the pre-roll -- the same cycles in every run, whatever K
               and W; peak = 8 TB/s (MI355X HBM3E).  The cycle's second kernel
               (active-row lists) is timed the same way and reported beside
               it.  Then W warm-up cycles, then the K timed ones.
  cpu_baseline the oracle's event-driven CSR form (oracle/stsim.c) on the
               host's cores, timed over the same cycles of the same network
               that the GPU timed (the first --cpu-steps of them when K is
               larger); the deliveries it counts over those cycles are
               compared with the GPU's ("events_match")

Rasters are not recorded in the timed region (recording costs one 8-byte
store per 64 neurons per cycle); bit-exactness is established by the gpu
tests on the same code path.

roofline.traffic is the step kernel's HBM bytes per launch, measured by two
rocprofv3 counter runs (FETCH_SIZE, then WRITE_SIZE) of this script, each
scaled by what the same counter reads for --copy-probe, a plain copy kernel of
known size; --pmc-summary prints the per-dispatch averages of such a run.  The
figures are read from profiles/step_kernel_<workload>_traffic.txt when that
file exists for the same workload and tile and was measured at a load within
10 % of this run's deliveries per launch (null otherwise).  `--steps 0` runs
the pre-roll and the roofline pass only (the counter runs: their last
--roofline-steps step-kernel dispatches are the roofline cycles).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from bsim_amd import api, netspec  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVENT = 8


def host_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    mem_gb = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal:"):
                mem_gb = round(int(line.split()[1]) / 2 ** 20, 1)
                break
    except OSError:
        pass
    return dict(cpu=model, nproc=os.cpu_count(), machine=platform.machine(), mem_gib=mem_gb,
                affinity=len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None)


def make_workload(name):
    """(host network, description): built by the library's seeded generator."""
    if name == "c1":
        return api.Network.standard(**netspec.C1), \
            "standard feed-forward ~100 Hz: 64 x 10,000 LIF_brian, ~1,000 syn/neuron, " \
            "delays 1..6, dt 0.1 ms"
    if name == "c2":
        return api.Network.standard(**netspec.C2), \
            "high-activity stress: 100 x 10,000 LIF_brian, ~2,000 syn/neuron, delays 1..50, " \
            "dt 0.1 ms"
    if name == "c0":
        return api.Network.from_spec(netspec.c0()), \
            "plain method: 2 x 1,000 LIF_brian, 1e6 all-to-all synapses"
    if name.startswith("c1x"):   # c1x<npop>: the C1 generator with fewer populations
        k = int(name[3:])
        return api.Network.standard(**{**netspec.C1, "npop": k}), \
            f"standard feed-forward, {k} x 10,000 LIF_brian"
    raise SystemExit(f"unknown workload {name}")


def cpu_reference(workload, steps):
    """The reference's own CPU simulator (STSim, float build, compiled from
    its sources into oracle/_ref) timed on a bounded sample: the workload's
    generator with 2 populations (its cost is O(synapses) per cycle, so the
    full network is out of reach), events counted by the oracle over the same
    cycles.  None when oracle/_ref is not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402
    import ref_io  # noqa: E402
    if ref_io.ref_binary() is None or workload not in ("c1", "c2"):
        return None
    params = netspec.C1 if workload == "c1" else netspec.C2
    spec = api.Network.standard(**{**params, "npop": 2}).export_spec()
    secs = ref_io.time_ref(spec, netspec.DT, steps)
    ev = oracle.run(spec, netspec.DT, steps).events
    return {"value": ev / secs, "unit": "synaptic events/s", "cores": 1, "kind": "reference",
            "sample": f"reference STSim::run (oracle/_ref, float) on the same generator with 2 "
                      f"populations ({len(spec.pre)} synapses), cycles 0..{steps}: {ev} events "
                      f"in {secs:.2f}s ({secs / steps * 1e3:.1f} ms per cycle: every synapse "
                      f"is visited every cycle)"}


def cpu_baseline(spec, start, steps, threads):
    """The oracle's CSR form over cycles [0, start + steps), timed over
    [start, start + steps): (events/s, events in the window, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402  (the oracle is only the checker / baseline)
    r = oracle.run(spec, netspec.DT, start + steps, mode="csr", tile=2048, threads=threads,
                   raster_cap=1 << 26, time_from=start)
    return r.events_from / r.wall_s, r.events_from, r.wall_s


COPY_PROBE_BYTES = 1 << 30


def copy_probe():
    """A plain device copy of known size: 1 GiB read + 1 GiB written, 5 times."""
    import torch
    n = COPY_PROBE_BYTES // 4
    src = torch.rand(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": COPY_PROBE_BYTES,
                      "copy_probe_bytes_written": COPY_PROBE_BYTES, "dispatches": 5}))


def pmc_summary(path, kernel_re, last=0):
    """Per-dispatch mean of every counter in rocprofv3 counter_collection.csv
    files under path, over the dispatches whose kernel name matches (the last
    `last` of them by dispatch id, if last > 0)."""
    rx = re.compile(kernel_re)
    vals = {}
    files = glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if not rx.search(row.get("Kernel_Name", "")):
                    continue
                key = (fn, int(row.get("Dispatch_Id", 0)))
                name = row["Counter_Name"]
                vals.setdefault(name, {}).setdefault(key, 0.0)
                vals[name][key] += float(row["Counter_Value"])
    out = {}
    for name, v in vals.items():
        keys = sorted(v, key=lambda k: k[1])
        if last > 0:
            keys = keys[-last:]
        out[name] = {"dispatches": len(keys),
                     "mean_per_dispatch": sum(v[k] for k in keys) / max(len(keys), 1)}
    print(json.dumps({"files": len(files), "kernel": kernel_re, "last": last, "counters": out}))


def read_traffic(workload, tile, events_per_launch):
    """(HBM bytes per launch of the step kernel, deliveries per launch of the
    profiled cycles) from the committed counter summary, or (None, profiled
    deliveries) when the profile's tile differs or its load is not within 10 %
    of this run's events_per_launch (traffic depends on the firing regime)."""
    fn = os.path.join(ROOT, "profiles", f"step_kernel_{workload}_traffic.txt")
    if not os.path.exists(fn):
        return None, None
    kv = {}
    for line in open(fn):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        prof_ev = float(kv["events_per_launch"]) if "events_per_launch" in kv else \
            float(kv["algorithmic_bytes_per_launch"]) / BYTES_PER_EVENT
        if int(kv["tile"]) != tile:
            return None, prof_ev
        if not (abs(prof_ev - events_per_launch) <= 0.1 * max(events_per_launch, 1.0)):
            return None, prof_ev
        return float(kv["hbm_bytes_per_launch"]), prof_ev
    except (KeyError, ValueError):
        return None, None


def bench_sweep(a):
    """Batched parameter sweep: networks/s of one batched simulation of all
    sweep networks versus the same networks simulated one at a time (timed on
    a random sample), each sampled network's raster checked bit for bit
    between the two."""
    from bsim_amd import sweep
    size = (a.c3_min_size, a.c3_max_size)
    pts = sweep.sample(a.c3_nets, seed=a.c3_seed, size=size)
    steps = a.c3_steps
    nsyn_est = sum(p.fan_in * sum(p.sizes[1:]) + p.drive_fan_in * p.sizes[0] for p in pts)
    print(f"[bench] c3: {len(pts)} networks, {sum(p.n for p in pts)} neurons, ~{nsyn_est:.3g} "
          f"synapses, {steps} cycles each", file=sys.stderr, flush=True)
    t0 = time.time()
    # the tuning batch: the first networks, up to --c3-tune-nets of them and
    # --c3-tune-syn synapses (the whole sweep at the default sizes)
    ntune, tsyn = 0, 0
    for p in pts[:a.c3_tune_nets]:
        ps = p.fan_in * sum(p.sizes[1:]) + p.drive_fan_in * p.sizes[0]
        if ntune and tsyn + ps > a.c3_tune_syn:
            break
        ntune += 1
        tsyn += ps
    (tile, rank_every), tt = sweep.autotune(pts[:ntune])
    print(f"[bench] autotune batch: {ntune} networks, ~{tsyn:.3g} synapses", file=sys.stderr, flush=True)
    print(f"[bench] autotuned tile {tile}, tiles ranked every {rank_every} cycles (0: never), "
          f"from {tt} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    rng = np.random.default_rng(a.c3_seed + 1)
    sample = sorted(rng.choice(len(pts), size=min(a.c3_solo, len(pts)), replace=False).tolist())
    # batched: consecutive networks in simulations of at most --c3-sim-syn
    # synapses (one simulation unless the sweep exceeds the device memory),
    # each built in parts of at most --c3-batch-syn synapses joined on the
    # device (bsim_sim_create_chains), timed without recording; then the
    # sampled networks again with the raster for the check
    t0 = time.time()
    os.environ.setdefault("BSIM_PROGRESS", "1")     # one line per part of the build
    groups, cur, cur_syn = [], [], 0
    for p in pts:
        ps = p.fan_in * sum(p.sizes[1:]) + p.drive_fan_in * p.sizes[0]
        if cur and cur_syn + ps > a.c3_sim_syn:
            groups.append(cur)
            cur, cur_syn = [], 0
        cur.append(p)
        cur_syn += ps
    groups.append(cur)
    secs_b, events, ntiles, nsyn = 0.0, 0, 0, 0
    for gi, grp in enumerate(groups):
        _, sb, info = sweep.run_batch_parts(grp, steps, tile=tile, record=False,
                                            part_synapses=int(a.c3_batch_syn),
                                            rank_every=rank_every)
        secs_b += sb
        events += info["events"]
        ntiles += info["ntiles"]
        nsyn += info["s"]
        print(f"[bench] batched simulation {gi + 1}/{len(groups)}: {len(grp)} networks, "
              f"{info['s']} synapses, {sb:.3f}s of stepping", file=sys.stderr, flush=True)
    print(f"[bench] batched: {secs_b:.3f}s for {len(pts)} networks in {len(groups)} "
          f"simulation(s) ({ntiles} tiles, {nsyn} synapses, {events} events; built and run in "
          f"{time.time() - t0:.0f}s)", file=sys.stderr, flush=True)
    rast, _, _ = sweep.run_batch([pts[i] for i in sample], steps, tile=tile, record=True,
                                 rank_every=rank_every)
    secs_s, exact = 0.0, 0
    for j, i in enumerate(sample):
        _, dt_solo = sweep.run_solo(pts[i], steps, tile=tile, record=False)
        secs_s += dt_solo
        solo, _ = sweep.run_solo(pts[i], steps, tile=tile, record=True)
        exact += int(all(np.array_equal(rast[j][c], solo[c]) for c in range(steps)))
        print(f"[bench] solo {j + 1}/{len(sample)}: network {i}, {dt_solo:.3f}s, bit-exact so far "
              f"{exact}/{j + 1}", file=sys.stderr, flush=True)
    nps_b = len(pts) / secs_b
    nps_s = len(sample) / secs_s
    print(f"[bench] one at a time: {secs_s:.3f}s for {len(sample)} networks; bit-exact "
          f"{exact}/{len(sample)}", file=sys.stderr, flush=True)
    out = {
        "metric": "sweep networks/s (batched)",
        "value": nps_b,
        "unit": "networks/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": secs_b / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sweep sampler)",
        "config": {"workload": f"batched sweep: {len(pts)} chains of 2-8 LIF_brian populations, "
                               f"sizes log-uniform [{size[0]}, {size[1]}], fan-in U[100, 1000], "
                               f"w0 U[0.3, 1.5], inhibitory U[0.1, 0.4], drive U[5, 50] Hz, "
                               f"delays 1..U{{6..20}}, dt 0.1 ms, {steps} cycles",
                   "tile": tile, "rank_every": rank_every, "autotune_networks": ntune,
                   "autotune": {f"tile {k[0]}, rank every {k[1]}": v for k, v in tt.items()},
                   "simulations": len(groups), "synapses": nsyn,
                   "parallelism": "single GPU"},
        "one_at_a_time_networks_per_s": nps_s,
        "speedup": nps_b / nps_s,
        "bit_exact_networks": f"{exact}/{len(sample)}",
        "synaptic_events_per_s_batched": events / secs_b,
    }
    print(json.dumps(out), flush=True)


def bench_surrogate(a):
    """Rate surrogate: simulate the training runs (batched), train the MLP on
    the GPU (bf16 MFMA kernels), screen candidates, and let the simulator
    judge the top picks against the same number of random picks."""
    from bsim_amd import surrogate as S
    from bsim_amd import sweep

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    size = (a.c4_min_size, a.c4_max_size)
    rng = np.random.default_rng(a.c4_seed)
    train_pts = sweep.sample_table(a.c4_runs, seed=a.c4_seed, size=size).points()
    t0 = time.perf_counter()
    rates, sim_secs = S.simulate_rates(train_pts, log=log)
    gen_secs = time.perf_counter() - t0
    log(f"training data: {len(train_pts)} runs, {gen_secs:.1f}s ({sim_secs:.1f}s stepping)")
    X, Y, idx = S.dataset(train_pts, rates, size=size)
    if a.c4_save:
        np.savez_compressed(a.c4_save, X=X, Y=Y, idx=idx)
    # runs split 80 / 10 / 10: training, early stopping, reported held-out error
    part = rng.random(len(train_pts))[idx[:, 0]]
    tr, va, te = part < 0.8, (part >= 0.8) & (part < 0.9), part >= 0.9
    batch = 4096
    m = S.MLP(max_batch=batch, seed=a.c4_seed)
    # warm-up (first launches), then a fresh model for the timed training
    S.fit(m, X[tr], Y[tr], epochs=1, batch=batch)
    m.close()
    m = S.MLP(max_batch=batch, seed=a.c4_seed)
    losses, train_secs, fit_info = S.fit(m, X[tr], Y[tr], epochs=a.c4_epochs, batch=batch,
                                         lr=1e-3, lr_final=1e-4, seed=a.c4_seed,
                                         val=(X[va], Y[va]), val_every=25)
    steps = len(losses)
    flops = S.flops_per_step(m.dims, batch, 8, S.N_BINS)
    tflops = flops * steps / train_secs / 1e12
    val_rmse = S.rmse_hz(m, X[te], Y[te])
    yt = Y[te].astype(np.float64) * S.RATE_SCALE
    pt = m.predict(X[te]).astype(np.float64) * S.RATE_SCALE
    # reference point: the error of a flat prediction at each held-out
    # population's own exact mean rate (what 1 ms spike-timing noise and the
    # onset transient leave to a predictor of the mean alone)
    flat_own_mean = float(np.sqrt(np.mean((yt - yt.mean(axis=1, keepdims=True)) ** 2)))
    mt, mp = yt.mean(axis=1), pt.mean(axis=1)
    mean_rmse = float(np.sqrt(np.mean((mp - mt) ** 2)))
    mean_within = float(np.mean(np.abs(mp - mt) <= 0.1 * np.maximum(mt, 1e-9)))
    log(f"trained {steps} steps in {train_secs:.2f}s ({tflops:.1f} TFLOP/s), loss "
        f"{losses[0]:.4f} -> {losses[-1]:.4f}, kept epoch {fit_info['best_epoch']}, "
        f"held-out rate RMSE {val_rmse:.2f} Hz")
    last = np.array([S.last_pop_mean(r[-1]) for r in rates])
    target = float(np.quantile(last, 0.9))
    cands = sweep.sample_table(a.c4_cands, seed=a.c4_seed + 1, size=size)
    top, pred, pred_secs = S.screen(m, cands, target, top_frac=0.01, size=size)
    nrows = int(cands.npop.sum())
    log(f"screened {len(cands)} candidates ({nrows} rows) in {pred_secs:.2f}s; target {target:.1f} Hz")
    resim = top if a.c4_resim <= 0 or len(top) <= a.c4_resim else \
        np.sort(rng.choice(top, a.c4_resim, replace=False))
    rnd = np.sort(rng.choice(len(cands), len(resim), replace=False))
    picks = sorted(set(resim.tolist()) | set(rnd.tolist()))
    sim, resim_secs = S.simulate_rates(cands.points(picks), log=log)
    simr = dict(zip(picks, sim))
    js = S.judge(resim, pred, simr, target)
    jr = S.judge(rnd, pred, simr, target)
    log(f"surrogate picks {js}, random picks {jr}")
    out = {
        "metric": "surrogate training samples/s",
        "value": steps * batch / train_secs,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": train_secs / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulation, fp32 master weights)",
        "data": "rates simulated by the library from seeded sweep networks",
        "config": {"workload": f"rate surrogate: MLP 8-512x4-200, batch {batch}; {len(train_pts)} "
                               f"training runs of 2-8-population chains, sizes log-uniform "
                               f"[{size[0]}, {size[1]}], 200 ms, 1 ms bins; {len(cands)} "
                               f"candidates, top 1 % ({len(top)}), {len(resim)} of them "
                               f"re-simulated beside {len(rnd)} random candidates",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None,
                     "note": "whole training step (all kernels, launch gaps included)"},
        "train_loss_first_last": [float(losses[0]), float(losses[-1])],
        "heldout_rate_rmse_hz": val_rmse,
        "heldout_rmse_flat_own_mean_hz": flat_own_mean,
        "heldout_mean_rate_rmse_hz": mean_rmse,
        "heldout_mean_rate_within_10pct": mean_within,
        "early_stopping": fit_info,
        "screen_rows_per_s": nrows / pred_secs,
        "target_hz": target,
        "surrogate_picks": js,
        "random_picks": jr,
        "hit_rate_gain": js["hit_rate"] / max(jr["hit_rate"], 1e-9),
        "data_gen_seconds": gen_secs,
        "resim_seconds": resim_secs,
    }
    print(json.dumps(out), flush=True)


def bench_mlp(a):
    """The surrogate's training step alone (gather, forward, loss, backward,
    Adam) at batch 4096 on a fixed synthetic regression set of the surrogate's
    shape (8 inputs, 200 outputs)."""
    from bsim_amd import surrogate as S
    batch, n = 4096, 65536
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (n, 8)).astype(np.float32)
    t = np.linspace(0, 1, S.N_BINS, dtype=np.float32)
    y = (np.sin(3 * x[:, :1] + 2 * t[None, :]) * x[:, 1:2] + 0.5 * x[:, 2:3] ** 2).astype(np.float32)
    m = S.MLP(max_batch=batch, seed=1)
    m.set_data(x, y)
    rows = rng.integers(0, n, (a.warmup + a.steps, batch)).astype(np.int32)
    m.train(rows[:a.warmup], batch, 1e-3)
    t0 = time.perf_counter()
    losses = m.train(rows[a.warmup:], batch, 1e-3)
    secs = time.perf_counter() - t0
    flops = S.flops_per_step(m.dims, batch, 8, S.N_BINS)
    tflops = flops * a.steps / secs / 1e12
    out = {
        "metric": "surrogate training samples/s", "value": a.steps * batch / secs,
        "unit": "samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": secs / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulation, fp32 master weights)",
        "data": "synthetic regression set (8 -> 200)",
        "config": {"workload": f"surrogate training step, MLP 8-512x4-200, batch {batch}",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None,
                     "note": "whole training step (all kernels, launch gaps included)"},
        "loss_first_last": [float(losses[0]), float(losses[-1])],
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--preroll", type=int, default=1000,
                    help="cycles run before the warm-up to reach the firing regime (setup)")
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=1000,
                    help="largest number of timed cycles the CPU baseline repeats")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ref-steps", type=int, default=100,
                    help="cycles of the reference CPU simulator's timed sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-steps", type=int, default=200)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--pmc-summary", default=None)
    ap.add_argument("--kernel", default="bsim_step_kernel")
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--c3-nets", type=int, default=512)
    ap.add_argument("--c3-solo", type=int, default=32)
    ap.add_argument("--c3-tune-nets", type=int, default=512)
    ap.add_argument("--c3-tune-syn", type=float, default=1.5e9,
                    help="synapses of the autotuner's probe batch")
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--c3-min-size", type=int, default=300)
    ap.add_argument("--c3-max-size", type=int, default=3000)
    ap.add_argument("--c3-seed", type=int, default=7)
    ap.add_argument("--c3-batch-syn", type=float, default=1.5e9,
                    help="synapses per part of the batched simulation's host build")
    ap.add_argument("--c3-sim-syn", type=float, default=2.5e9,
                    help="synapses per batched simulation (device and host memory)")
    ap.add_argument("--c4-runs", type=int, default=4096)
    ap.add_argument("--c4-cands", type=int, default=1000000)
    ap.add_argument("--c4-resim", type=int, default=1024)
    ap.add_argument("--c4-epochs", type=int, default=300)
    ap.add_argument("--c4-min-size", type=int, default=100)
    ap.add_argument("--c4-max-size", type=int, default=1000)
    ap.add_argument("--c4-seed", type=int, default=11)
    ap.add_argument("--c4-save", default="", help="save the simulated training set (npz)")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe()
        return
    if a.pmc_summary:
        pmc_summary(a.pmc_summary, a.kernel, a.last)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU spike "
              "exchange is not part of this build.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "c3":
        bench_sweep(a)
        return
    if a.workload == "c4":
        bench_surrogate(a)
        return
    if a.workload == "mlp":
        bench_mlp(a)
        return

    t0 = time.time()
    net, desc = make_workload(a.workload)
    n_neurons, n_syn = net.n, net.s
    spec = None if a.no_cpu else net.export_spec()
    t_gen = time.time() - t0
    print(f"[bench] workload {a.workload}: n={n_neurons} s={n_syn} (generated in {t_gen:.1f}s)",
          file=sys.stderr, flush=True)
    t0 = time.time()
    sim = api.Simulation(net, netspec.DT, tile=a.tile)
    del net
    t_build = time.time() - t0
    info = sim.info()
    print(f"[bench] built+uploaded in {t_build:.1f}s: {info}", file=sys.stderr, flush=True)

    t0 = time.time()
    sim.step(a.preroll)
    sim.sync()
    ev_pre = sim.events()
    print(f"[bench] pre-rolled {a.preroll} cycles in {time.time() - t0:.2f}s ({ev_pre} events)",
          file=sys.stderr, flush=True)
    # live kernel timing at a fixed point of the run (cycles preroll ..
    # preroll + R, whatever K and W are): HIP events around each kernel of R
    # single launches, recorded by the library on the simulation's stream
    R = a.roofline_steps
    evr0 = sim.events()
    step_us, rows_us = sim.time_steps(R)
    evr1 = sim.events()
    mean_ms = step_us * 1e-3
    ev_per_launch = (evr1 - evr0) / R
    achieved = ev_per_launch * BYTES_PER_EVENT / (mean_ms * 1e-3) / 1e9
    print(f"[bench] step kernel mean {step_us:.2f} us, row-list kernel mean {rows_us:.2f} us "
          f"over {R} launches (cycles {a.preroll}..{a.preroll + R}), {ev_per_launch:.0f} "
          f"events/launch", file=sys.stderr, flush=True)
    traffic, traffic_ev = read_traffic(a.workload, info["tile"], ev_per_launch)
    if a.steps == 0:            # a profiling run of the pre-roll and the roofline pass only
        sim.close()
        print(json.dumps({"roofline_only": True, "mean_launch_us": step_us,
                          "events_per_launch": ev_per_launch}), flush=True)
        return
    sim.step(a.warmup)
    sim.sync()
    start = a.preroll + R + a.warmup
    # the first min(K, --cpu-steps) timed launches are the CPU baseline's
    # window; their deliveries are read at a stream synchronisation inside the
    # timed region (one extra host wait when K is larger).  Launch L (counted
    # from 1) performs the delivery phase of cycle L - 2 and the update of
    # cycle L - 1, so timed launches start+1 .. start+K deliver cycles
    # start-1 .. start+K-2: the oracle times and counts those cycles.
    k_cpu = 0 if a.no_cpu else min(a.steps, a.cpu_steps)
    ev0 = sim.events()
    t0 = time.perf_counter()
    ev_cpu_window = None
    if 0 < k_cpu < a.steps:
        sim.step(k_cpu)
        sim.sync()
        ev_cpu_window = sim.events() - ev0
        sim.step(a.steps - k_cpu)
    else:
        sim.step(a.steps)
    sim.sync()
    t1 = time.perf_counter()
    ev1 = sim.events()
    wall = t1 - t0
    events = ev1 - ev0
    if ev_cpu_window is None:
        ev_cpu_window = events
    value = events / wall
    print(f"[bench] timed {a.steps} cycles (cycles {start}..{start + a.steps}) in {wall:.4f}s, "
          f"{events} events ({events / max(a.steps, 1):.4g} per cycle)", file=sys.stderr,
          flush=True)

    sim.close()

    cpu = None
    cpu_ref = None
    if not a.no_cpu:
        threads = a.cpu_threads or min(os.cpu_count() or 1, 16)
        c0 = max(start - 1, 0)
        v, ev, secs = cpu_baseline(spec, c0, k_cpu, threads)
        cpu = {"value": v, "unit": "synaptic events/s", "cores": threads, "kind": "port",
               "sample": f"oracle CSR form (OpenMP), cycles {c0}..{c0 + k_cpu} of the same "
                         f"network = the deliveries of the GPU's first {k_cpu} timed launches "
                         f"({ev} events in {secs:.3f}s; cycles 0..{c0} run untimed first)",
               "events_match": bool(ev == ev_cpu_window), "gpu_over_cpu": value / v,
               "host": host_info()}
        print(f"[bench] cpu baseline {v:.3g} events/s on {threads} threads over the same "
              f"{k_cpu} cycles; events {ev} (GPU {ev_cpu_window})", file=sys.stderr, flush=True)
        cpu_ref = cpu_reference(a.workload, a.ref_steps)
        if cpu_ref:
            print(f"[bench] reference CPU path {cpu_ref['value']:.3g} events/s ({cpu_ref['sample']})",
                  file=sys.stderr, flush=True)

    out = {
        "metric": "synaptic events/s",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": desc, "name": a.workload, "neurons": n_neurons,
                   "synapses": n_syn, "dt_s": netspec.DT, "tile": info["tile"],
                   "preroll": a.preroll, "roofline_cycles": [a.preroll, a.preroll + R],
                   "timed_cycles": [start, start + a.steps],
                   "timed_delivery_cycles": [start - 1, start + a.steps - 1],
                   "events_per_cycle": events / max(a.steps, 1),
                   "simulated_s": a.steps * netspec.DT, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_profile_events_per_launch": traffic_ev,
                     "kernel": "bsim_step_kernel", "mean_launch_us": step_us,
                     "events_per_launch": ev_per_launch,
                     "row_list_kernel_us": rows_us},
        "cpu_baseline": cpu,
        "cpu_reference": cpu_ref,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
