"""Parity at the workloads' full sizes, through the C ABI.

* The standard network (64 x 10,000 LIF_brian, 6.31e8 synapses, delays 1..6):
  1,100 cycles -- past the cold start into the ~100 Hz regime -- against the
  oracle's tile form on the host's cores: raster (ordered ids per cycle), V
  traces, final state and delivered events, bit for bit.
* The high-activity stress network (100 x 10,000, ~2e9 synapses, delays
  1..50): its first cycles bit for bit against the oracle, then through the
  ~2,000 Hz regime every cycle's deliveries against the count that the GPU's
  own raster implies (the sum, over the spikes of the D previous cycles, of
  the lengths of their (pre, delay) rows), computed on the host.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from bsim_amd import api, netspec

pytestmark = pytest.mark.gpu
DT = netspec.DT


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def run_both(net, spec, steps, rec, time_from=0):
    sim = api.Simulation(net, DT, rec_steps=steps, rec_ids=rec)
    sim.step(steps)
    sim.finish()
    sim.sync()
    counts, ids, _ = sim.raster()
    g = dict(counts=counts, ids=ids, trace=sim.trace(), events=sim.events(), **sim.state())
    info = sim.info()
    sim.close()
    say(f"[full] GPU done: {len(ids)} spikes, {g['events']} events, tile {info['tile']}")
    o = oracle.run(spec, DT, steps, rec_ids=rec, mode="csr", tile=2048, threads=threads(),
                   raster_cap=max(1 << 20, 2 * len(ids) + 1024))
    say(f"[full] oracle done: {len(o.raster_ids)} spikes, {o.events} events")
    return g, o


def assert_same(g, o, what):
    np.testing.assert_array_equal(g["counts"], o.raster_count, err_msg=f"{what}: counts")
    np.testing.assert_array_equal(g["ids"], o.raster_ids, err_msg=f"{what}: ids")
    assert np.array_equal(g["trace"].view(np.uint32), o.rec_v.view(np.uint32)), f"{what}: V trace"
    for a, b in (("v", o.final_v), ("ie", o.final_ie), ("ii", o.final_ii)):
        assert np.array_equal(g[a].view(np.uint32), b.view(np.uint32)), f"{what}: final {a}"
    np.testing.assert_array_equal(g["refrac"], o.final_refrac, err_msg=f"{what}: refrac")
    np.testing.assert_array_equal(g["fire_count"], o.fire_count, err_msg=f"{what}: spike counts")
    assert g["events"] == o.events, (what, g["events"], o.events)


@pytest.mark.timeout(400)
def test_standard_network_full_size_vs_oracle():
    """The 640k-neuron standard network for 1,100 cycles (0.11 s), bit for bit."""
    net = api.Network.standard(**netspec.C1)
    say(f"[full] C1 n={net.n} s={net.s}")
    spec = net.export_spec()
    steps = 1100
    rec = [10000, 10001, 25000, 333333, 640000, 649999]
    g, o = run_both(net, spec, steps, rec)
    del net, spec
    assert_same(g, o, "C1")
    # the last 100 cycles are in the ~100 Hz regime: about 1 % of the
    # 650,000 neurons fire per cycle
    late = g["counts"][-100:].mean()
    say(f"[full] C1 spikes per cycle over the last 100 cycles: {late:.0f}; "
        f"events {g['events']}")
    assert 3000 < late < 15000


def row_lengths(spec, dmax, chunk=1 << 27):
    """Synapses per (pre, delay steps) row, [n, dmax + 1], from the creation-
    order arrays (integral delays: round(d / dt) in float, as the reference's
    StaticSynapse::reset computes it)."""
    n = spec.n
    out = np.zeros(n * (dmax + 1), np.int64)
    dt = np.float32(DT)
    for a in range(0, spec.s, chunk):
        pre = spec.pre[a:a + chunk].astype(np.int64)
        ds = np.rint(spec.d[a:a + chunk].astype(np.float32) / dt).astype(np.int64)
        out += np.bincount(pre * (dmax + 1) + ds, minlength=out.size)
    return out.reshape(n, dmax + 1)


@pytest.mark.timeout(900)
def test_stress_network_full_size():
    """The ~2,000 Hz stress network (100 x 10,000 LIF_brian, 2.19e9 synapses,
    delays 1..50): the first 150 cycles bit for bit against the oracle, then
    1,150 more cycles through the saturated regime, each cycle's deliveries
    equal to the sum over its active (pre, delay) rows of their lengths,
    computed on the host from the GPU's own raster."""
    net = api.Network.standard(**netspec.C2)
    say(f"[full] C2 n={net.n} s={net.s}")
    assert net.s > 2 ** 31        # every 64-bit offset path is in use
    spec = net.export_spec()
    k0, steps, D = 150, 1300, netspec.C2["dmax"]
    rec = [5000, 10000, 10001, 55555, 500000, 1009999]
    sim = api.Simulation(net, DT, rec_steps=steps, rec_ids=rec)
    del net
    info = sim.info()
    say(f"[full] C2 layout: {info}")
    ev = np.zeros(steps + 1, np.int64)          # deliveries after each launch
    for t in range(steps):
        sim.step(1)
        ev[t + 1] = sim.events()
        if t + 1 == k0:
            st = sim.state()
            tr = sim.trace()
        if (t + 1) % 100 == 0:
            say(f"[full] C2 launch {t + 1}: {ev[t + 1] - ev[t]} events")
    counts, ids, per_cycle = sim.raster()
    sim.close()
    say(f"[full] C2 GPU done: {len(ids)} spikes, {ev[-1]} events")
    # the saturated regime: ~20 % of the LIF neurons fire per cycle (2,000 Hz)
    lif_frac = counts[-100:].mean() / 1e6
    say(f"[full] C2 firing fraction over the last 100 cycles: {lif_frac:.3f}")
    assert 0.12 < lif_frac < 0.25
    # deliveries of phase (ii) of cycle c (performed by launch c + 1)
    rl = row_lengths(spec, D)
    for c in range(k0, steps - 1):
        want = 0
        for d in range(1, D + 1):
            if c - d >= 0 and len(per_cycle[c - d]):
                want += int(rl[per_cycle[c - d], d].sum())
        assert ev[c + 2] - ev[c + 1] == want, (c, ev[c + 2] - ev[c + 1], want)
    del rl
    say("[full] C2 per-cycle deliveries match the raster")
    o = oracle.run(spec, DT, k0, rec_ids=rec, mode="csr", tile=2048, threads=threads(),
                   raster_cap=1 << 26)
    say(f"[full] C2 oracle done: {len(o.raster_ids)} spikes, {o.events} events")
    off = int(counts[:k0].sum())
    np.testing.assert_array_equal(counts[:k0], o.raster_count)
    np.testing.assert_array_equal(ids[:off], o.raster_ids)
    assert np.array_equal(tr[:k0].view(np.uint32), o.rec_v.view(np.uint32))
    for a, b in (("v", o.final_v), ("ie", o.final_ie), ("ii", o.final_ii)):
        assert np.array_equal(st[a].view(np.uint32), b.view(np.uint32)), a
    np.testing.assert_array_equal(st["refrac"], o.final_refrac)
    np.testing.assert_array_equal(st["fire_count"], o.fire_count)
    assert ev[k0 + 1] == o.events, (ev[k0 + 1], o.events)
