"""Poisson sources (Poisson_spikes, src/neuron/poisson/PoissonNeuron.cpp:13-63).

The reference draws its intervals from a clock-seeded engine, so no run of it
can be matched spike for spike; what it fixes is the rule: first spike at
int(startTime/dt) + X, every later one 1 + X cycles after the previous, none at
or after int(endTime/dt), X ~ Poisson(rate).  Here the draws are seeded
(inverse CDF at a counter-based hash), and

* CPU: the oracle's inverse-CDF tables against scipy's Poisson CDF (an
  independent implementation), the intervals' distribution against the
  Poisson pmf (chi-square), and the rule's closed forms (rate 0 fires every
  cycle of [start, end); nothing outside it);
* GPU: the library's rasters, V and counts equal the oracle's bit for bit on
  networks that Poisson sources drive.
"""
import ctypes as C
import sys

import numpy as np
import pytest
from scipy import stats

import oracle
from bsim_amd import netspec

DT = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def poisson_table(rate):
    lib = oracle.lib()
    lib.bso_poisson_table.restype = C.POINTER(C.c_uint32)
    lib.bso_poisson_table.argtypes = [C.c_double]
    p = lib.bso_poisson_table(float(rate))
    J = p[0]
    tab = np.array([p[1 + j] for j in range(J)], np.uint64)
    C.CDLL(None).free(C.cast(p, C.c_void_p))
    return tab


def source_net(rates, starts, durs, n_lif=0, seed=5, fan=0, w=0.0):
    b = netspec.Builder()
    m = len(rates)
    b.population(np.zeros((m, netspec.LIF_NPARAM)), kind=netspec.POISSON,
                 pois=(np.asarray(rates), np.asarray(starts), np.asarray(durs)))
    if n_lif:
        rng = np.random.Generator(np.random.PCG64(seed))
        rows = np.tile(netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 2e-4, 5e-3, 8e-3, 15.0, 0.0),
                       (n_lif, 1))
        rows[:, 9] = rng.uniform(12.0, 14.5, n_lif)
        b.population(rows)
        pre = np.repeat(np.arange(m), fan)
        post = m + rng.integers(0, n_lif, m * fan)
        wts = np.where(rng.random(m * fan) < 0.25, -w * rng.random(m * fan),
                       w * rng.random(m * fan)).astype(np.float32)
        d = rng.integers(1, 5, m).repeat(fan).astype(np.float32) * np.float32(DT)
        b.connect(pre, post, wts, d, (wts < 0).astype(np.int8))
    spec = b.build()
    spec.seed = seed
    return spec


def spike_cycles(res, n):
    """Per neuron, the ascending cycles it fired at."""
    out = [[] for _ in range(n)]
    pos = 0
    for t, c in enumerate(res.raster_count):
        for i in res.raster_ids[pos:pos + c]:
            out[i].append(t)
        pos += c
    return out


@pytest.mark.parametrize("rate", [0.0, 0.3, 7.5, 120.0, 3000.0])
def test_table_matches_scipy_cdf(rate):
    tab = poisson_table(rate)
    J = len(tab)
    j = np.arange(J - 1)
    ref = np.floor(stats.poisson.cdf(j, rate) * 2.0 ** 32) if rate > 0 else np.full(J - 1, 2.0 ** 32)
    ref = np.minimum(ref, 2.0 ** 32 - 1)
    assert np.all(np.abs(tab[:-1].astype(np.float64) - ref) <= 2.0 + 1e-7 * ref), rate
    assert np.all(np.diff(tab.astype(np.int64)) >= 0)
    # the tail the table drops is below 2^-32
    assert stats.poisson.sf(J - 2, max(rate, 1e-300)) < 2.0 ** -32 or rate == 0.0


@pytest.mark.parametrize("rate", [2.5, 40.0])
def test_intervals_are_one_plus_poisson(rate):
    m = 400
    steps = 6000
    spec = source_net([rate] * m, [0.0] * m, [10.0] * m)
    res = oracle.run(spec, DT, steps)
    gaps = np.concatenate([np.diff(np.array(c)) - 1 for c in spike_cycles(res, m)])
    assert len(gaps) > 20000
    assert gaps.min() >= 0
    # chi-square against the Poisson pmf over bins with enough expected mass
    lo, hi = int(stats.poisson.ppf(1e-4, rate)), int(stats.poisson.ppf(1 - 1e-4, rate))
    edges = np.arange(lo, hi + 2)
    obs = np.array([np.sum(gaps < lo)] + [np.sum(gaps == k) for k in edges[:-1]] +
                   [np.sum(gaps > hi)], np.float64)
    p = np.concatenate([[stats.poisson.cdf(lo - 1, rate)], stats.poisson.pmf(edges[:-1], rate),
                        [stats.poisson.sf(hi, rate)]])
    exp = p * len(gaps)
    keep = exp >= 5
    chi2 = np.sum((obs[keep] - exp[keep]) ** 2 / exp[keep])
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, (chi2, keep.sum())
    assert abs(gaps.mean() - rate) < 5 * np.sqrt(rate / len(gaps)) + 1e-9


def test_rate_zero_fires_every_cycle_of_its_window():
    """X = 0 (but for the one uniform value 2^32 - 1): a spike at every
    cycle c with int(start/dt) <= c < int(end/dt), none outside."""
    starts = np.array([0.0, 0.0013, 0.0057, 0.002], np.float32)
    durs = np.array([0.0, 0.0021, 0.001, 0.05], np.float32)
    steps = 200
    res = oracle.run(source_net([0.0] * 4, starts, durs), DT, steps)
    got = spike_cycles(res, 4)
    for i in range(4):
        c0 = int(np.float32(starts[i]) / np.float32(DT))
        c1 = int(np.float32(np.float32(starts[i]) + np.float32(durs[i])) / np.float32(DT))
        assert got[i] == list(range(c0, min(c1, steps))), i


def test_first_spike_after_start_and_none_after_end():
    m = 300
    rng = np.random.default_rng(3)
    starts = rng.uniform(0.0, 0.02, m).astype(np.float32)
    durs = rng.uniform(0.0, 0.03, m).astype(np.float32)
    rates = rng.uniform(0.0, 30.0, m).astype(np.float32)
    steps = 600
    res = oracle.run(source_net(rates, starts, durs), DT, steps)
    for i, c in enumerate(spike_cycles(res, m)):
        c0 = int(starts[i] / np.float32(DT))
        c1 = int(np.float32(starts[i] + durs[i]) / np.float32(DT))
        if c:
            assert c[0] >= c0 and c[-1] < c1, i
            assert np.all(np.diff(c) >= 1)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 384])
def test_poisson_driven_network_vs_oracle(tile):
    from test_gpu_parity import assert_same, run_gpu
    rng = np.random.default_rng(11)
    m = 500
    rates = rng.choice(np.array([0.0, 4.0, 25.0, 150.0], np.float32), m)
    starts = rng.uniform(0.0, 0.01, m).astype(np.float32)
    durs = rng.uniform(0.005, 0.05, m).astype(np.float32)
    spec = source_net(rates, starts, durs, n_lif=700, fan=40, w=4.0)
    steps = 500
    rec = [0, 3, 520, 900, 1199]
    say(f"[gpu] poisson n={spec.n} s={spec.s} tile={tile}")
    g = run_gpu(spec, steps, rec, tile)
    o = oracle.run(spec, DT, steps, rec_ids=rec)
    assert o.fire_count[:m].sum() > 1000 and o.fire_count[m:].sum() > 0 and o.events > 1000
    assert_same(g, o, "poisson")


def test_library_rejects_bad_poisson_arguments():
    """Host-side checks of bsim_net_add_poisson (no GPU involved)."""
    from bsim_amd import api
    net = api.Network()
    for rate, start, dur in [(-1.0, 0.0, 1.0), (2e6, 0.0, 1.0), (1.0, -0.1, 1.0),
                             (1.0, 0.0, -1.0), (float("nan"), 0.0, 1.0)]:
        with pytest.raises(Exception):
            net.add_poisson([rate], [start], [dur], 1)
    assert net.add_poisson([3.0, 0.0], [0.0, 0.1], [1.0, 0.0], 1) == 0
    assert net.n == 2
