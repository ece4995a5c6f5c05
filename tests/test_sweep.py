"""Batched sweeps (bsim_amd/sweep.py): many networks in one simulation, each
bit-exact against itself simulated alone, and against the oracle."""
import numpy as np
import pytest

import oracle
from bsim_amd import api, netspec, sweep

DT = netspec.DT

SMALL = dict(npop=(2, 4), size=(150, 1500), fan_in=(20, 120), dmax=(6, 20), drive_fan_in=20)


def test_chain_is_the_same_alone_and_in_a_batch():
    """A chain appended after other chains is the same network, neuron for
    neuron and synapse for synapse (ids shifted), as the chain built alone."""
    pts = sweep.sample(4, seed=3, **SMALL)
    b = sweep.build(pts)
    full = b.net.export_spec()
    for k, p in enumerate(pts):
        alone = api.Network()
        assert alone.add_chain(**p.kwargs()) == 0
        a = alone.export_spec()
        f0, n = int(b.first[k]), p.n
        assert np.array_equal(full.kind[f0:f0 + n], a.kind)
        assert np.array_equal(full.lif[f0:f0 + n], a.lif)
        assert np.array_equal(full.bern_p[f0:f0 + n], a.bern_p)
        sel = (full.pre >= f0) & (full.pre < f0 + n)
        assert np.array_equal(full.pre[sel] - f0, a.pre)
        assert np.array_equal(full.post[sel] - f0, a.post)
        assert np.array_equal(full.w[sel].view(np.uint32), a.w.view(np.uint32))
        assert np.array_equal(full.d[sel].view(np.uint32), a.d.view(np.uint32))
        assert np.array_equal(full.stype[sel], a.stype)


def test_batched_generator_equals_chain_by_chain():
    """bsim_net_gen_chains builds the same network as bsim_net_gen_chain called
    once per chain, appended after an existing chain."""
    pts = sweep.sample(6, seed=4, **SMALL)
    a, b = api.Network(), api.Network()
    a.add_chain(**pts[0].kwargs())
    b.add_chain(**pts[0].kwargs())
    fa = [a.add_chain(**p.kwargs()) for p in pts[1:]]
    fb = b.add_chains([p.kwargs() for p in pts[1:]])
    assert list(fa) == list(fb)
    sa, sb = a.export_spec(), b.export_spec()
    assert sa.seed == sb.seed
    for f in ("kind", "lif", "bern_p", "pre", "post", "w", "d", "stype"):
        x, y = np.ascontiguousarray(getattr(sa, f)), np.ascontiguousarray(getattr(sb, f))
        assert x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), f


def test_chain_respects_fan_in_and_pre_order():
    p = sweep.sample(1, seed=5, **SMALL)[0]
    net = api.Network()
    net.add_chain(**p.kwargs())
    s = net.export_spec()
    _, ds = oracle.synapse_reset(s, DT)
    assert oracle.csr_order_ok(s, ds)
    start = p.n_drive
    for i in range(1, len(p.sizes)):
        lo, hi = start + sum(p.sizes[:i]), start + sum(p.sizes[:i + 1])
        fin = np.bincount(s.post[(s.post >= lo) & (s.post < hi)] - lo, minlength=hi - lo)
        assert abs(fin.mean() - p.fan_in) <= 0.5 * p.fan_in + 1


def test_split_raster():
    counts = np.array([3, 0, 2], np.int32)
    ids = np.array([1, 5, 12, 4, 11], np.int32)
    r = sweep.split_raster(counts, ids, np.array([0, 4, 10]), np.array([4, 6, 5]))
    assert [list(x) for x in r[0]] == [[1], [], []]
    assert [list(x) for x in r[1]] == [[1], [], [0]]
    assert [list(x) for x in r[2]] == [[2], [], [1]]


def test_sample_is_seeded():
    a = sweep.sample(5, seed=9)
    b = sweep.sample(5, seed=9)
    assert [p.kwargs() for p in a] == [p.kwargs() for p in b]
    for p in a:
        assert 2 <= len(p.sizes) <= 8 and 100 <= p.fan_in <= 1000 and 6 <= p.dmax <= 20


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 384])
def test_batch_rasters_equal_solo_rasters(tile):
    steps = 300
    pts = sweep.sample(10, seed=11, **SMALL)
    rasters, _, info = sweep.run_batch(pts, steps, tile=tile)
    total = 0
    for k, p in enumerate(pts):
        solo, _ = sweep.run_solo(p, steps, tile=tile)
        assert len(solo) == steps
        for c in range(steps):
            np.testing.assert_array_equal(rasters[k][c], solo[c], err_msg=f"network {k} cycle {c}")
        total += sum(len(x) for x in solo)
    assert total > 0


@pytest.mark.gpu
def test_tile_schedule_does_not_change_results():
    """The tile schedule (tiles ranked heaviest first every r cycles, or
    never) only reorders the workgroups: every network's raster and the
    deliveries are the same for r = 0, 1, 3 and the default, and switching
    the schedule between calls continues the run bit for bit."""
    steps = 300
    pts = sweep.sample(10, seed=12, **SMALL)
    ref, _, info = sweep.run_batch(pts, steps, tile=192)
    for r in (0, 1, 3):
        got, _, inf = sweep.run_batch(pts, steps, tile=192, rank_every=r)
        assert inf["events"] == info["events"], r
        for k in range(len(pts)):
            for c in range(steps):
                np.testing.assert_array_equal(got[k][c], ref[k][c], err_msg=f"r={r} network {k} cycle {c}")
    b = sweep.build(pts)
    sim = api.Simulation(b.net, netspec.DT, tile=192, rec_steps=steps)
    for r in (0, 5, 1, 0, 4):
        sim.set_schedule(r)
        sim.step(60)
    sim.finish()
    sim.sync()
    counts, ids, _ = sim.raster()
    got = sweep.split_raster(counts, ids, b.first, b.sizes)
    assert sim.events() == info["events"]
    sim.close()
    for k in range(len(pts)):
        for c in range(steps):
            np.testing.assert_array_equal(got[k][c], ref[k][c], err_msg=f"switched, network {k} cycle {c}")


@pytest.mark.gpu
def test_autotune_times_every_candidate():
    pts = sweep.sample(6, seed=13, **SMALL)
    (tile, r), times = sweep.autotune(pts, steps=64, preroll=64, tiles=(192, 384), ranks=(0, 4))
    assert set(times) == {(t, x) for t in (192, 384) for x in (0, 4)}
    assert (tile, r) in times and times[(tile, r)] == min(times.values())


@pytest.mark.gpu
def test_sweep_network_alone_matches_oracle():
    steps = 250
    for p in sweep.sample(2, seed=13, **SMALL):
        net = api.Network()
        net.add_chain(**p.kwargs())
        spec = net.export_spec()
        solo, _ = sweep.run_solo(p, steps)
        o = oracle.run(spec, DT, steps)
        np.testing.assert_array_equal(np.array([len(x) for x in solo]), o.raster_count)
        np.testing.assert_array_equal(np.concatenate(solo), o.raster_ids)


@pytest.mark.gpu
@pytest.mark.timeout(400)
def test_full_size_sweep_batch_equals_solo():
    """Networks drawn from the sweep's full distribution (2-8 populations of
    1,000-50,000 neurons, fan-in 100-1,000, delays up to 6..20 steps), 2,000
    cycles (0.2 s) in one batched simulation; a sample of them run alone is
    bit for bit the same, and one of those equals the oracle."""
    steps = 2000
    pts = sweep.sample(40, seed=17)
    nsyn = sum(p.fan_in * sum(p.sizes[1:]) + p.drive_fan_in * p.sizes[0] for p in pts)
    print(f"[gpu] full-size sweep batch: {len(pts)} networks, {sum(p.n for p in pts)} neurons, "
          f"~{nsyn:.3g} synapses", flush=True)
    rasters, secs, info = sweep.run_batch(pts, steps)
    print(f"[gpu] batch stepped in {secs:.2f}s, tile {info['tile']}", flush=True)
    rng = np.random.default_rng(3)
    pick = sorted(rng.choice(len(pts), 5, replace=False).tolist())
    spikes = 0
    for k in pick:
        solo, _ = sweep.run_solo(pts[k], steps)
        for c in range(steps):
            np.testing.assert_array_equal(rasters[k][c], solo[c], err_msg=f"network {k} cycle {c}")
        spikes += sum(len(x) for x in solo)
    assert spikes > 1000
    k = min(pick, key=lambda i: pts[i].n)
    net = api.Network()
    net.add_chain(**pts[k].kwargs())
    o = oracle.run(net.export_spec(), DT, steps, mode="csr", tile=2048, threads=8)
    np.testing.assert_array_equal(np.array([len(x) for x in rasters[k]]), o.raster_count)
    np.testing.assert_array_equal(np.concatenate(rasters[k]), o.raster_ids)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 192])
def test_batch_built_in_parts_equals_one_build(tile):
    """A batch laid out in parts joined on the device (bsim_sim_create_chains,
    here forced into several parts, each padded to whole tiles) gives every
    network the raster of the batch built in one host network, and the same
    deliveries."""
    steps = 400
    pts = sweep.sample(12, seed=19, **SMALL)
    one, _, info1 = sweep.run_batch(pts, steps, tile=tile)
    per = max(p.fan_in * sum(p.sizes[1:]) + p.drive_fan_in * p.sizes[0] for p in pts)
    parts, _, info2 = sweep.run_batch_parts(pts, steps, tile=tile, part_synapses=2 * per)
    assert info2["n"] >= info1["n"]           # inert padding between the parts
    assert info2["s"] == info1["s"] and info2["events"] == info1["events"]
    for k in range(len(pts)):
        for c in range(steps):
            np.testing.assert_array_equal(parts[k][c], one[k][c], err_msg=f"network {k} cycle {c}")
    assert sum(len(x) for r in one for x in r) > 0
