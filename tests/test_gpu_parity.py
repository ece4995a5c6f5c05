"""Parity of the HIP step with the reference (golden vectors from the
reference's own CPU simulator) and with the oracle, through the C ABI.

Bar: rasters equal as ordered id lists per cycle; V traces, final state and
per-neuron spike counts equal bit for bit; delivered-event counts equal.
"""
import sys

import numpy as np
import pytest

import oracle
from bsim_amd import api, netspec

import golden_io

pytestmark = pytest.mark.gpu
DT = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_gpu(spec, steps, rec_ids, tile=0):
    net = api.Network.from_spec(spec)
    sim = api.Simulation(net, DT, tile=tile, rec_steps=steps, rec_ids=rec_ids)
    if len(rec_ids):
        sim.enable_state_trace()
    sim.step(steps)
    sim.finish()
    sim.sync()
    counts, ids, _ = sim.raster()
    out = dict(raster_count=counts, raster_ids=ids, rec_v=sim.trace(), events=sim.events(),
               info=sim.info(), **sim.state())
    if len(rec_ids):
        out.update({"rec_" + k: v for k, v in sim.state_trace().items()})
    sim.close()
    return out


STATE_TRACES = ("rec_ie", "rec_ii", "rec_in_e", "rec_in_i")


def assert_same(g, o, what):
    for k in STATE_TRACES:
        if k in g:
            assert np.array_equal(g[k].view(np.uint32), getattr(o, k).view(np.uint32)), \
                f"{what}: {k}"
    np.testing.assert_array_equal(g["raster_count"], o.raster_count, err_msg=f"{what}: counts")
    np.testing.assert_array_equal(g["raster_ids"], o.raster_ids, err_msg=f"{what}: ids")
    assert np.array_equal(g["rec_v"].view(np.uint32), o.rec_v.view(np.uint32)), f"{what}: V"
    lif = slice(None)
    assert np.array_equal(g["v"].view(np.uint32)[lif], o.final_v.view(np.uint32)[lif]), what
    assert np.array_equal(g["ie"].view(np.uint32), o.final_ie.view(np.uint32)), what
    assert np.array_equal(g["ii"].view(np.uint32), o.final_ii.view(np.uint32)), what
    np.testing.assert_array_equal(g["refrac"], o.final_refrac, err_msg=what)
    np.testing.assert_array_equal(g["fire_count"], o.fire_count, err_msg=what)
    assert g["events"] == o.events, (what, g["events"], o.events)


GOLDEN_F32 = [n for n in golden_io.case_names() if n.endswith("_f32")]


@pytest.mark.parametrize("tile", [0, 192])
@pytest.mark.parametrize("name", GOLDEN_F32)
def test_golden_cases(name, tile):
    """GPU == reference CPU build, bit for bit."""
    meta, spec, dtype, g = golden_io.load(name)
    say(f"[gpu] golden {name} tile={tile} n={spec.n} s={spec.s} steps={meta['steps']}")
    # shuffled creation orders included: each target's inputs are summed in
    # its own creation order
    r = run_gpu(spec, meta["steps"], meta["rec_ids"], tile)
    np.testing.assert_array_equal(r["raster_count"], g["raster_count"])
    np.testing.assert_array_equal(r["raster_ids"], g["raster_ids"])
    assert np.array_equal(r["rec_v"].view(np.uint32), g["rec_v_bits"])
    for k in STATE_TRACES:   # i_E / i_I after the update, the input it consumed
        assert np.array_equal(r[k].view(np.uint32), g[k + "_bits"]), k
    np.testing.assert_array_equal(r["fire_count"], g["fire_count"])
    lif = spec.kind == netspec.LIF
    assert np.array_equal(r["v"].view(np.uint32)[lif], g["consts_bits"][lif][:, 8])
    o = oracle.run(spec, DT, meta["steps"], rec_ids=meta["rec_ids"])
    assert r["events"] == o.events


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("tile", [192, 384, 576, 0])
def test_random_networks_vs_oracle(seed, tile):
    spec = netspec.random_net(100 + seed, n_lif=300, n_arr=20, n_bern=20, p_conn=0.05,
                              dmax=1 + 7 * seed, allow_dup=seed % 2 == 0, spike_prob=0.1)
    steps = 400
    rec = list(range(0, spec.n, 7))
    say(f"[gpu] random seed={seed} tile={tile} n={spec.n} s={spec.s}")
    r = run_gpu(spec, steps, rec, tile)
    o = oracle.run(spec, DT, steps, rec_ids=rec)
    assert o.raster_ids.size > 100
    assert_same(r, o, f"random{seed}")


def test_max_delay_63():
    spec = netspec.standard(npop=3, n=300, fanout=40, dmax=63, w0=4.0, drive_rate_hz=300.0,
                            n_drive=100, drive_w=14.0, seed=5, i_offset=(12.0, 14.5))
    say(f"[gpu] dmax63 n={spec.n} s={spec.s}")
    r = run_gpu(spec, 500, [150, 450, 750], 192)
    o = oracle.run(spec, DT, 500, rec_ids=[150, 450, 750])
    assert r["info"]["dmax"] == 63
    assert_same(r, o, "dmax63")


def test_standard_medium_vs_csr_oracle():
    """A C1-shaped network at 1/40 scale (16 x 1000, fan-out 250, delays
    1..6, Bernoulli drive): GPU vs the oracle's CSR form on 8 threads."""
    spec = netspec.standard(npop=16, n=1000, fanout=250, w0=2.0, inh_frac=0.25, dmax=6,
                            drive_rate_hz=500.0, n_drive=500, drive_w=10.0, seed=11,
                            i_offset=(10.0, 14.9))
    steps = 1000
    rec = [500, 1500, 9000, 16499]
    say(f"[gpu] standard medium n={spec.n} s={spec.s}")
    r = run_gpu(spec, steps, rec)
    o = oracle.run(spec, DT, steps, rec_ids=rec, mode="csr", tile=512, threads=8)
    assert o.events > 1_000_000
    assert_same(r, o, "standard-medium")


@pytest.mark.parametrize("tile", [0, 192])
def test_stress_shape_vs_csr_oracle(tile):
    """The high-activity stress shape at 1/100 scale (10 x 1000, fan-out 200
    with w0 x 10 to keep the recurrent input of fan-out 2,000, delays 1..50,
    no refractory period): activity saturates (about 30 % of the neurons fire
    every cycle at the end), so much of the 50-deep history is live.  GPU at
    the automatic tile (768 for delays over 24 steps) and at 192 vs the
    oracle's CSR form."""
    spec = netspec.c2(npop=10, n=1000, fanout=200, n_drive=1000, seed=21, w0=0.45 * 1.3 * 10)
    steps = 600
    rec = [1000, 5000, 10999]
    say(f"[gpu] stress shape n={spec.n} s={spec.s} tile={tile}")
    r = run_gpu(spec, steps, rec, tile)
    o = oracle.run(spec, DT, steps, rec_ids=rec, mode="csr", tile=512, threads=8)
    assert r["info"]["dmax"] == 50
    assert r["info"]["tile"] == (768 if tile == 0 else tile)
    assert o.events > 50_000_000
    assert_same(r, o, "stress-shape")


@pytest.mark.parametrize("seed", range(3))
def test_input_logs_vs_oracle(seed):
    """FireInfo "Y" (summed LIF input) and "X" (delay slots visited) from the
    GPU vs the oracle, bit for bit: random networks with array and Bernoulli
    sources and mixed delays, and a standard network."""
    if seed < 2:
        spec = netspec.random_net(seed + 40, n_lif=300, n_arr=20, n_bern=20, p_conn=0.1,
                                  dmax=3 + 8 * seed, spike_prob=0.05, steps=300)
    else:
        spec = netspec.standard(npop=4, n=500, fanout=80, w0=3.0, drive_rate_hz=400.0,
                                n_drive=100, seed=7, i_offset=(12.0, 14.9))
    steps = 300
    net = api.Network.from_spec(spec)
    sim = api.Simulation(net, DT, tile=192)
    sim.step(10)                       # graph-free cycles first, then the log turned on
    sim.enable_input_log()
    sim.step(steps - 10)
    sim.finish()
    sim.sync()
    y_late, x_late = sim.input_log()
    sim.close()
    sim = api.Simulation(net, DT, tile=192)
    sim.enable_input_log()
    sim.step(steps)
    sim.finish()
    sim.sync()
    y, x = sim.input_log()
    sim.close()
    o = oracle.run(spec, DT, steps)
    assert np.array_equal(y.view(np.uint32), o.input_sum.view(np.uint32))
    np.testing.assert_array_equal(x, o.slot_count)
    assert x.sum() > 0 and np.count_nonzero(y) > 0
    assert x_late.sum() < x.sum()      # counting starts where it is turned on


def test_run_twice_identical():
    spec = netspec.standard(npop=4, n=500, fanout=80, w0=3.0, drive_rate_hz=400.0, n_drive=200,
                            drive_w=12.0, seed=3, i_offset=(11.0, 14.0))
    a = run_gpu(spec, 300, [10, 900])
    b = run_gpu(spec, 300, [10, 900])
    for k in ("raster_ids", "fire_count"):
        np.testing.assert_array_equal(a[k], b[k])
    assert np.array_equal(a["rec_v"].view(np.uint32), b["rec_v"].view(np.uint32))
    assert a["events"] == b["events"]


def test_step_in_chunks_equals_one_call():
    spec = netspec.standard(npop=3, n=400, fanout=60, w0=3.0, drive_rate_hz=400.0, n_drive=100,
                            drive_w=12.0, seed=9, i_offset=(11.0, 14.0))
    net = api.Network.from_spec(spec)
    s1 = api.Simulation(net, DT, rec_steps=200, rec_ids=[5, 700])
    for n in (1, 2, 3, 50, 144):
        s1.step(n)
    s1.sync()
    s2 = api.Simulation(net, DT, rec_steps=200, rec_ids=[5, 700])
    s2.step(200)
    s2.sync()
    np.testing.assert_array_equal(s1.raster()[1], s2.raster()[1])
    assert np.array_equal(s1.trace().view(np.uint32), s2.trace().view(np.uint32))


def test_graph_replays_match_single_launches():
    """Runs of 64 cycles are replayed from a captured graph; single-cycle calls
    launch one kernel each.  Graph replays interleaved with single launches,
    host reads and synchronisations give the oracle's result bit for bit, and
    the run after a read continues where it stopped."""
    spec = netspec.standard(npop=3, n=400, fanout=60, w0=3.0, drive_rate_hz=400.0, n_drive=100,
                            drive_w=12.0, seed=11, i_offset=(11.0, 14.0))
    steps = 64 * 3 + 37
    o = oracle.run(spec, DT, steps, rec_ids=[5, 700])
    net = api.Network.from_spec(spec)
    s1 = api.Simulation(net, DT, rec_steps=steps, rec_ids=[5, 700])
    done = 0
    for n in (64, 1, 128, 3, 33):      # graph, eager, 2 graphs, eager, eager
        s1.step(n)
        done += n
        s1.sync()
        c, ids, _ = s1.raster()
        off = int(o.raster_count[:done].sum())
        np.testing.assert_array_equal(c[:done], o.raster_count[:done])
        np.testing.assert_array_equal(ids, o.raster_ids[:off])
    assert done == steps
    s1.finish()
    s1.sync()
    assert s1.events() == o.events
    assert np.array_equal(s1.trace().view(np.uint32), o.rec_v.view(np.uint32))
    s1.close()


def dense_net(n_src=700, n_tgt=300, p_fire=0.9, dmax=3, seed=4):
    """All-to-all from busy Bernoulli sources, every (pre, post) pair created
    twice in a row: ~600 active pres per cycle, rows of 2 x 256 events, so
    the item, segment and boundary-table passes each overflow their staging
    capacity several times per cycle."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = netspec.Builder()
    b.population(np.zeros((n_src, netspec.LIF_NPARAM)), kind=netspec.BERNOULLI,
                 bern_p=np.uint32(int(p_fire * 2 ** 32)))
    rows = np.tile(netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 2e-4, 5e-3, 8e-3, 15.0, 0.0),
                   (n_tgt, 1))
    rows[:, 9] = rng.uniform(10.0, 14.0, n_tgt)
    b.population(rows)
    pre = np.repeat(np.arange(n_src), 2 * n_tgt)
    post = n_src + np.tile(np.repeat(np.arange(n_tgt), 2), n_src)
    m = len(pre)
    w = np.where(rng.random(m) < 0.3, -rng.uniform(0, 0.02, m), rng.uniform(0, 0.01, m))
    k = rng.integers(1, dmax + 1, n_src)            # one delay per pre: rows stay whole
    d = np.repeat(k, 2 * n_tgt).astype(np.float32) * np.float32(DT)
    b.connect(pre, post, w.astype(np.float32), d, (w < 0).astype(np.int8))
    spec = b.build()
    spec.seed = seed
    return spec


@pytest.mark.parametrize("tile", [192, 384, 768, 1536])
def test_dense_overflow_paths(tile):
    spec = dense_net()
    steps = 60
    rec = [700, 850, 999]
    say(f"[gpu] dense n={spec.n} s={spec.s} tile={tile}")
    r = run_gpu(spec, steps, rec, tile)
    o = oracle.run(spec, DT, steps, rec_ids=rec, mode="csr", tile=256, threads=8)
    assert o.events > steps * 100_000
    assert_same(r, o, "dense")


def _edge_net(kind):
    b = netspec.Builder()
    lif = netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 2e-3, 5e-3, 8e-3, 1.0, 0.0)
    if kind == "no_synapses":
        rows = np.tile(lif, (300, 1))
        rows[:, 9] = np.linspace(0.5, 4.0, 300)          # tonic firing above 1.0, no input
        b.population(rows)
    elif kind == "one_neuron":
        r = lif.copy()
        r[9] = 5.0
        b.population(r[None, :])
        b.connect([0], [0], np.float32([0.2]), np.float32([3 * DT]))    # self-synapse
    elif kind == "sources_only":
        rng = np.random.default_rng(5)
        sp = [np.sort(rng.choice(200, size=20, replace=False)) for _ in range(7)]
        b.population(np.zeros((7, netspec.LIF_NPARAM)), kind=netspec.ARRAY, spikes=sp)
        b.population(np.zeros((5, netspec.LIF_NPARAM)), kind=netspec.BERNOULLI,
                     bern_p=np.uint32(int(0.1 * 2 ** 32)))
        pre = np.repeat(np.arange(12), 12)
        post = np.tile(np.arange(12), 12)
        b.connect(pre, post, np.float32(np.linspace(-1, 1, 144)), np.float32(np.full(144, 2 * DT)))
    elif kind == "all_inhibitory":
        rows = np.tile(lif, (250, 1))
        rows[:50, 9] = np.linspace(2.0, 4.0, 50)
        rows[50:, 9] = np.linspace(1.0, 1.8, 200)
        b.population(rows[:50])
        b.population(rows[50:])
        rng = np.random.default_rng(6)
        pre = np.repeat(np.arange(50), 40)
        post = 50 + np.concatenate([np.sort(rng.choice(200, 40, replace=False)) for _ in range(50)])
        b.connect(pre, post, np.float32(-rng.uniform(0, 0.3, len(pre))),
                  np.float32(rng.integers(1, 5, len(pre)) * DT))
    return b.build()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["no_synapses", "one_neuron", "sources_only", "all_inhibitory"])
@pytest.mark.parametrize("tile", [0, 1536])
def test_edge_networks_vs_oracle(kind, tile):
    """Networks without synapses, of one neuron (with a self-synapse), of
    sources only, and with only inhibitory synapses; a tile much larger than
    the network: raster, V, state and the input logs bit-exact."""
    spec = _edge_net(kind)
    steps = 200
    rec = [0, spec.n - 1]
    say(f"[gpu] edge {kind} n={spec.n} s={spec.s} tile={tile}")
    net = api.Network.from_spec(spec)
    sim = api.Simulation(net, DT, tile=tile, rec_steps=steps, rec_ids=rec)
    sim.enable_input_log()
    sim.step(steps)
    sim.finish()
    sim.sync()
    counts, ids, _ = sim.raster()
    r = dict(raster_count=counts, raster_ids=ids, rec_v=sim.trace(), events=sim.events(),
             info=sim.info(), **sim.state())
    y, x = sim.input_log()
    sim.close()
    o = oracle.run(spec, DT, steps, rec_ids=rec)
    assert_same(r, o, kind)
    assert np.array_equal(y.view(np.uint32), o.input_sum.view(np.uint32)), kind
    np.testing.assert_array_equal(x, o.slot_count, err_msg=kind)
    if kind == "no_synapses":
        assert r["events"] == 0 and x.sum() == 0 and len(ids) > 0


@pytest.mark.gpu
def test_device_allocation_failure_is_reported():
    """A recording window larger than the device (2^31 - 1 cycles of a
    10,000-neuron raster, 2.7 TB): bsim_sim_create reports BSIM_E_NOMEM's
    failure instead of crashing and frees what it allocated; the device then
    runs the next simulation as usual (bit-exact against the oracle)."""
    b = netspec.Builder()
    rows = np.tile(netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 2e-3, 5e-3, 8e-3, 1.0, 0.0), (10000, 1))
    b.population(rows)
    big = api.Network.from_spec(b.build())
    with pytest.raises(api.BsimError, match="out of memory"):
        api.Simulation(big, DT, rec_steps=2 ** 31 - 1)
    spec = _edge_net("all_inhibitory")
    r = run_gpu(spec, 200, [0, spec.n - 1])
    assert_same(r, oracle.run(spec, DT, 200, rec_ids=[0, spec.n - 1]), "after a failed create")


@pytest.mark.gpu
def test_zero_steps_and_finish():
    spec = _edge_net("one_neuron")
    sim = api.Simulation(api.Network.from_spec(spec), DT, rec_steps=10)
    sim.step(0)
    sim.finish()
    sim.sync()
    assert sim.events() == 0
    counts, ids, _ = sim.raster()
    assert len(counts) == 0 and len(ids) == 0
    sim.close()


@pytest.mark.parametrize("dt", [5e-5, 2e-4])
def test_other_time_steps_vs_oracle(dt):
    """dt other than 0.1 ms: LIF constants, refractory steps and delay steps
    all follow dt (delays are whole multiples of the new dt)."""
    spec = netspec.random_net(77, n_lif=200, n_arr=10, n_bern=10, p_conn=0.15, dmax=7,
                              spike_prob=0.05, steps=400)
    spec.d = (np.rint(np.asarray(spec.d, np.float64) / DT) * dt).astype(np.float32)
    steps = 400
    net = api.Network.from_spec(spec)
    sim = api.Simulation(net, dt, rec_steps=steps, rec_ids=[20, 150])
    sim.step(steps)
    sim.finish()
    sim.sync()
    counts, ids, _ = sim.raster()
    r = dict(raster_count=counts, raster_ids=ids, rec_v=sim.trace(), events=sim.events(),
             info=sim.info(), **sim.state())
    sim.close()
    o = oracle.run(spec, dt, steps, rec_ids=[20, 150])
    assert len(ids) > 50
    assert_same(r, o, f"dt={dt}")


def test_two_simulations_interleaved():
    """Simulations share no device state: two stepped in alternation give the
    rasters of each run alone."""
    a = netspec.standard(npop=3, n=400, fanout=60, w0=3.0, drive_rate_hz=400.0, n_drive=100,
                         seed=31, i_offset=(12.0, 14.9))
    b = netspec.random_net(32, n_lif=300, n_arr=10, n_bern=10, p_conn=0.1, dmax=9,
                           spike_prob=0.05, steps=300)
    steps = 300
    alone = [run_gpu(s, steps, [1]) for s in (a, b)]
    sims = [api.Simulation(api.Network.from_spec(s), DT, rec_steps=steps, rec_ids=[1])
            for s in (a, b)]
    for c in range(0, steps, 50):
        for sim in sims:
            sim.step(50)
    for sim, ref in zip(sims, alone):
        sim.finish()
        sim.sync()
        counts, ids, _ = sim.raster()
        np.testing.assert_array_equal(counts, ref["raster_count"])
        np.testing.assert_array_equal(ids, ref["raster_ids"])
        assert sim.events() == ref["events"]
        sim.close()
