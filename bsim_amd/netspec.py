"""Flat description of a BSim network and seeded generators for the workloads.

A :class:`NetSpec` is the network exactly as BSim's ``Network`` would hold it
after the user's ``createPopulation``/``connect`` calls: neurons in creation
order (population after population), synapses in ``connect`` call order.  Real
values are kept in float64 as written by the user; a float build rounds them
to float32 once (``(float)x``), exactly as passing a double literal to a
``real`` parameter does in the reference.

Per-neuron LIF row (``LIF_NPARAM`` values), the arguments of
``LIF_brian(LIFENeuron(v_init, v_rest, v_reset, cm, tau_m, tau_refrac,
tau_syn_E, tau_syn_I, v_thresh, i_offset), tau_syn_E, tau_syn_I)``
(src/neuron/lif/LIFENeuron.h:9, src/base/CompositeNeuron.h:21):

    0 v_init  1 v_rest  2 v_reset  3 cm  4 tau_m  5 tau_refrac
    6 tau_syn_E  7 tau_syn_I  8 v_thresh  9 i_offset
    10 composite tau_syn_E  11 composite tau_syn_I

Neuron kinds: ``LIF`` (LIF_brian), ``ARRAY`` (Array_spikes: fires at the
listed cycles, src/neuron/array/ArrayNeuron.cpp:67-80), ``BERNOULLI``
(a seeded per-cycle Bernoulli source) and ``POISSON`` (Poisson_spikes with the
reference's CPU semantics, src/neuron/poisson/PoissonNeuron.cpp:36-63: first
spike at int(start/dt) + X, then every 1 + X cycles before int(end/dt), X ~
Poisson(rate) drawn from seeded hashes instead of the reference's
clock-seeded engine, PoissonNeuron.cpp:25-26).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

LIF, ARRAY, BERNOULLI, POISSON = 0, 1, 2, 3
LIF_NPARAM = 12


@dataclass
class NetSpec:
    kind: np.ndarray                 # int8 [n]
    lif: np.ndarray                  # float64 [n, 12]
    pre: np.ndarray                  # int32 [s]
    post: np.ndarray                 # int32 [s]
    w: np.ndarray                    # float32/64 [s]
    d: np.ndarray                    # float32/64 [s] (seconds)
    stype: np.ndarray | None = None  # int8 [s], 0 excitatory / 1 inhibitory
    tau: np.ndarray | None = None    # [s], 0 -> destination's composite tau
    arr_off: np.ndarray | None = None  # int64 [n+1]
    arr_t: np.ndarray | None = None    # int32 spike cycles
    bern_p: np.ndarray | None = None   # uint32 [n] threshold of the source hash
    seed: int = 0
    pois: np.ndarray | None = None     # float [n, 3] rate, startTime, duration (Poisson rows)
    pops: list = field(default_factory=list)  # [(start, count)] per population

    @property
    def n(self) -> int:
        return int(len(self.kind))

    @property
    def s(self) -> int:
        return int(len(self.pre))


def lif_row(v_init, v_rest, v_reset, cm, tau_m, tau_refrac, tau_syn_E, tau_syn_I,
            v_thresh, i_offset, comp_tau_E=None, comp_tau_I=None):
    if comp_tau_E is None:
        comp_tau_E = tau_syn_E
    if comp_tau_I is None:
        comp_tau_I = tau_syn_I
    return np.array([v_init, v_rest, v_reset, cm, tau_m, tau_refrac, tau_syn_E, tau_syn_I,
                     v_thresh, i_offset, comp_tau_E, comp_tau_I], np.float64)


class Builder:
    """Incremental construction in the reference's call order."""

    def __init__(self):
        self.kinds = []
        self.lifs = []
        self.arr = []           # per neuron list of cycles (ARRAY) or None
        self.bern = []          # per neuron threshold
        self.pois = []          # per population [cnt, 3] Poisson parameters or None
        self.pre, self.post, self.w, self.d, self.stype, self.tau = [], [], [], [], [], []
        self.pops = []
        self.n = 0

    def population(self, rows, kind=LIF, spikes=None, bern_p=None, pois=None):
        """pois: (rate, startTime, duration), each a scalar or one per neuron,
        for kind POISSON."""
        rows = np.atleast_2d(np.asarray(rows, np.float64))
        cnt = rows.shape[0]
        self.pois.append(None if pois is None else np.stack(
            [np.broadcast_to(np.asarray(x, np.float64), (cnt,)) for x in pois], axis=1))
        start = self.n
        self.kinds.append(np.full(cnt, kind, np.int8))
        self.lifs.append(rows)
        for i in range(cnt):
            self.arr.append(None if spikes is None else np.asarray(spikes[i], np.int32))
        self.bern.append(np.zeros(cnt, np.uint32) if bern_p is None
                         else np.broadcast_to(np.asarray(bern_p, np.uint32), (cnt,)).copy())
        self.pops.append((start, cnt))
        self.n += cnt
        return len(self.pops) - 1

    def connect(self, pre, post, w, d, stype=None, tau=None):
        pre = np.asarray(pre, np.int32)
        self.pre.append(pre)
        self.post.append(np.asarray(post, np.int32))
        self.w.append(np.asarray(w))
        self.d.append(np.asarray(d))
        self.stype.append(np.zeros(len(pre), np.int8) if stype is None
                          else np.asarray(stype, np.int8))
        self.tau.append(np.zeros(len(pre)) if tau is None else np.asarray(tau))

    def connect_all_to_all(self, psrc, pdst, w, d, stype=None):
        """``Network::connect(pSrc, pDst, w, d, type, size)``: synapse i joins
        src i / dstSize to dst i % dstSize (src/net/Network.h:176-209)."""
        s0, sn = self.pops[psrc]
        d0, dn = self.pops[pdst]
        i = np.arange(sn * dn, dtype=np.int64)
        self.connect(s0 + i // dn, d0 + i % dn, w, d, stype)

    def build(self, wdtype=np.float32) -> NetSpec:
        kind = np.concatenate(self.kinds) if self.kinds else np.zeros(0, np.int8)
        lif = np.concatenate(self.lifs) if self.lifs else np.zeros((0, LIF_NPARAM))
        if any(a is not None for a in self.arr):
            lens = np.array([0 if a is None else len(a) for a in self.arr], np.int64)
            arr_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            arr_t = (np.concatenate([a for a in self.arr if a is not None]).astype(np.int32)
                     if lens.sum() else np.zeros(0, np.int32))
        else:
            arr_off, arr_t = None, None
        cat = lambda xs, dt: (np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt))  # noqa
        pois = None
        if any(p is not None for p in self.pois):
            pois = np.concatenate([np.zeros((len(k), 3)) if p is None else p
                                   for p, k in zip(self.pois, self.kinds)])
        return NetSpec(kind=kind, lif=lif, pre=cat(self.pre, np.int32),
                       post=cat(self.post, np.int32), w=cat(self.w, wdtype),
                       d=cat(self.d, wdtype), stype=cat(self.stype, np.int8),
                       tau=cat(self.tau, wdtype), arr_off=arr_off, arr_t=arr_t,
                       bern_p=np.concatenate(self.bern) if self.bern else None,
                       pois=pois, pops=list(self.pops))


# ---- the reference's own unit networks (test/gpu/lif_unit{1..4}_test.cpp) ----

def lif_unit(k: int, wdtype=np.float32) -> NetSpec:
    """3 populations x 10 LIF neurons, all-to-all p0->p1->p2, as in
    test/gpu/lif_unit{1,2,3,4}_test.cpp (dt = 1e-4, run 0.1 s)."""
    N = 10
    refr = {4: (2e-4, 1e-4, 0.0)}.get(k, (0.0, 0.0, 0.0))
    b = Builder()
    b.population(np.tile(lif_row(0.0, 0.0, 0.0, 1.0e-1, 50.0e-3, refr[0], 1.0, 1.0, 15.0e-3,
                                 100.0e-1, 1.0, 1.0), (N, 1)))
    b.population(np.tile(lif_row(0.0, 0.0, 0.0, 1.0e-1, 50.0e-3, refr[1], 1.0, 1.0, 15.0e-3,
                                 0.0e-3, 1.0, 1.0), (N, 1)))
    b.population(np.tile(lif_row(0.0, 0.0, 0.0, 1.0e-1, 50.0e-3, refr[2], 1.0, 1.0, 15.0e-3,
                                 0.0e-3, 1.0, 1.0), (N, 1)))
    if k == 2:
        i = np.arange(N * N, dtype=np.float64)
        w0 = 1e-4 + i / float(N * N) * 2e-3
        w1 = 2.1e-3 - i / float(N * N) * 2e-3
    else:
        w0 = np.full(N * N, 1e-5)
        w1 = np.full(N * N, 2e-5)
    d0 = np.full(N * N, 2e-4 if k == 3 else 1e-4)
    d1 = np.full(N * N, 1e-4)
    # the reference stores these into `real` arrays before connect()
    cast = (lambda x: x.astype(np.float32)) if wdtype == np.float32 else (lambda x: x)
    b.connect_all_to_all(0, 1, cast(w0), cast(d0))
    b.connect_all_to_all(1, 2, cast(w1), cast(d1))
    return b.build(wdtype)


# ---- synthetic workloads -------------------------------------------------------

DT = 1e-4


def c0(seed: int = 1234, n: int = 1000, tau_syn_E=5e-3, tau_syn_I=10e-3) -> NetSpec:
    """Plain method, small: 2 x n LIF (tau_m 20 ms, tau_refrac 2 ms, v_rest =
    v_reset = 0, v_thresh 15 mV, R = 1), i_offset ~ U[14.5, 16.5), all-to-all
    p0->p1 with w ~ U[0, 0.2) and integer delays U{1..6} steps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tau_m = 20e-3
    b = Builder()
    for _ in range(2):
        ioff = rng.uniform(14.5, 16.5, n).astype(np.float32).astype(np.float64)
        rows = np.tile(lif_row(0.0, 0.0, 0.0, tau_m, tau_m, 2e-3, tau_syn_E, tau_syn_I, 15.0,
                               0.0), (n, 1))
        rows[:, 9] = ioff
        b.population(rows)
    w = rng.uniform(0.0, 0.2, n * n).astype(np.float32)
    k = rng.integers(1, 7, n * n)
    d = k.astype(np.float32) * np.float32(DT)
    b.connect_all_to_all(0, 1, w, d)
    return b.build()


# ---- counter-based hash shared with the C++ generator (csrc/gen.cpp) ---------

_U64 = np.uint64


def _mix64(z):
    with np.errstate(over="ignore"):
        z = z + _U64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> _U64(30))) * _U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> _U64(27))) * _U64(0x94D049BB133111EB)
    return z ^ (z >> _U64(31))


def _h3(seed, stream, counter):
    with np.errstate(over="ignore"):
        a = _mix64(_U64(seed) ^ (_U64(stream) * _U64(0xD1B54A32D192ED03)))
    return _mix64(a ^ np.asarray(counter, _U64))


def _u01(x):
    return (np.asarray(x, _U64) >> _U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


TAG_TGT, TAG_INH, TAG_W, TAG_D, TAG_IOFF, TAG_REM, TAG_REMDST = 1, 2, 3, 4, 5, 6, 7


def standard(npop=4, n=1000, fanout=100, w0=1.0, inh_frac=0.25, dmax=6, remote_frac=0.2,
             drive_rate_hz=0.0, n_drive=None, drive_w=None, drive_fanout=None, seed=1,
             tau_refrac=2e-3, i_offset=(0.0, 0.0), tau_syn_E=5e-3, tau_syn_I=10e-3,
             v_thresh=15.0):
    """Feed-forward chain of ``npop`` LIF populations of ``n`` neurons; numpy
    restatement of the library's generator (bsim_net_gen_standard,
    csrc/gen.cpp), used for small networks and to pin that generator.

    Population i projects to i+1; a ``remote_frac`` share of populations also
    project to one uniformly chosen remote population (not i, not i+1) and
    split their ``fanout`` out-synapses per neuron evenly between the two
    projections.  Within a projection, pre neuron i draws one target in each
    of f disjoint integer strata of the destination (distinct targets).
    Excitatory weights ~ U[0, 0.7 w0), a share ``inh_frac`` inhibitory with
    weights ~ U[-0.5 w0, 0); integer delays U{1..dmax} steps.  LIF neurons:
    tau_m 20 ms, R = 1, v_rest = v_reset = 0, i_offset ~ U[i_offset).
    Population 0 is driven by ``n_drive`` Bernoulli sources firing at
    ``drive_rate_hz`` (``drive_fanout`` targets each, weight ``drive_w``).
    All draws come from a counter-based hash of ``seed``.  Synapses are
    created source population by source population, pre-major, so every
    target's inputs are created in ascending pre order.
    """
    f32 = lambda x: float(np.float32(x))  # noqa: E731  (the C struct holds floats)
    w0, inh_frac, remote_frac = f32(w0), f32(inh_frac), f32(remote_frac)
    drive_rate_hz, tau_refrac, dtf = f32(drive_rate_hz), f32(tau_refrac), f32(DT)
    if n_drive is None:
        n_drive = n
    if drive_w is None:
        drive_w = 0.7 * w0
    drive_w = f32(drive_w)
    if drive_fanout is None:
        drive_fanout = fanout
    lo_i, hi_i = f32(i_offset[0]), f32(i_offset[1])
    kinds, rows_l, bern_l, pops = [], [], [], []
    nn = 0
    drive = drive_rate_hz > 0 and n_drive > 0
    if drive:
        pr = min(1.0, drive_rate_hz * dtf)
        thr = np.uint32(min(4294967295.0, np.round(pr * 4294967296.0)))
        kinds.append(np.full(n_drive, BERNOULLI, np.int8))
        rows = np.zeros((n_drive, LIF_NPARAM))
        rows[:, 10:12] = f32(1e-3)
        rows_l.append(rows)
        bern_l.append(np.full(n_drive, thr, np.uint32))
        pops.append((0, n_drive))
        nn = n_drive
    first_lif = len(pops)
    g = np.arange(nn, nn + npop * n, dtype=np.uint64)
    ioff = (np.zeros(len(g), np.float32) if not hi_i > lo_i else
            (lo_i + (hi_i - lo_i) * _u01(_h3(seed, TAG_IOFF << 32, g))).astype(np.float32))
    rows = np.tile(lif_row(0.0, 0.0, 0.0, 20e-3, 20e-3, tau_refrac, tau_syn_E, tau_syn_I,
                           v_thresh, 0.0), (npop * n, 1)).astype(np.float32).astype(np.float64)
    rows[:, 9] = ioff
    kinds.append(np.full(npop * n, LIF, np.int8))
    rows_l.append(rows)
    bern_l.append(np.zeros(npop * n, np.uint32))
    for i in range(npop):
        pops.append((nn + i * n, n))
    remote = [-1] * npop
    for i in range(npop):
        if npop > 2 and _u01(_h3(seed, TAG_REM << 32, i)) < remote_frac:
            nch = npop - (2 if i + 1 < npop else 1)
            c = min(int(_u01(_h3(seed, TAG_REMDST << 32, i)) * nch), nch - 1)
            remote[i] = [j for j in range(npop) if j != i and j != i + 1][c]
    proj = []
    if drive:
        proj.append((0, first_lif, min(drive_fanout, n)))
    for i in range(npop):
        src = first_lif + i
        if i + 1 < npop:
            proj.append((src, src + 1, min(fanout - fanout // 2 if remote[i] >= 0 else fanout, n)))
        if remote[i] >= 0:
            proj.append((src, first_lif + remote[i], min(fanout // 2, n)))
    total = sum(pops[a][1] * f for a, _, f in proj)
    pre = np.empty(total, np.int32)
    post = np.empty(total, np.int32)
    w = np.empty(total, np.float32)
    d = np.empty(total, np.float32)
    st = np.zeros(total, np.int8)
    o = 0
    for pi, (src, dst, f) in enumerate(proj):
        s0, sn = pops[src]
        d0, dn = pops[dst]
        m = sn * f
        c = np.arange(m, dtype=np.uint64)
        k = np.tile(np.arange(f, dtype=np.int64), sn)
        lo = (k * dn) // f
        hi = ((k + 1) * dn) // f
        off = lo + (_u01(_h3(seed, (TAG_TGT << 32) | pi, c)) * (hi - lo).astype(np.float64)
                    ).astype(np.int64)
        off = np.minimum(off, hi - 1)
        pre[o:o + m] = np.repeat(np.arange(s0, s0 + sn, dtype=np.int32), f)
        post[o:o + m] = d0 + off
        if drive and src == 0:
            w[o:o + m] = np.float32(drive_w)
        else:
            inh = _u01(_h3(seed, (TAG_INH << 32) | pi, c)) < inh_frac
            u = _u01(_h3(seed, (TAG_W << 32) | pi, c))
            w[o:o + m] = np.where(inh, -0.5 * w0 * u, 0.7 * w0 * u).astype(np.float32)
            st[o:o + m] = inh
        dk = np.minimum(1 + (_u01(_h3(seed, (TAG_D << 32) | pi, c)) * dmax).astype(np.int64),
                        dmax)
        d[o:o + m] = dk.astype(np.float32) * np.float32(dtf)
        o += m
    return NetSpec(kind=np.concatenate(kinds), lif=np.concatenate(rows_l), pre=pre, post=post,
                   w=w, d=d, stype=st, tau=None, bern_p=np.concatenate(bern_l), seed=seed,
                   pops=pops)


# The standard feed-forward workload at ~100 Hz: 64 x 10,000 LIF, ~1,000
# out-synapses per neuron, 20 % remote projections, 25 % inhibitory, delays
# 1..6 steps; neurons carry i_offset ~ U[14.5, 16.5) (as the plain method)
# and population 0 receives a Bernoulli drive.
C1 = dict(npop=64, n=10000, fanout=1000, w0=0.45, inh_frac=0.25, dmax=6, remote_frac=0.2,
          drive_rate_hz=20.0, n_drive=10000, drive_w=1.0, drive_fanout=100, seed=1,
          i_offset=(14.5, 16.5))
# High-activity stress workload at ~2,000 Hz: 100 x 10,000 LIF, 2,200
# out-synapses per neuron (2.19e9 synapses: past 2^31, so that every 64-bit
# offset of the layout and the kernels is in use), w0 x 1.3, 20 % inhibitory,
# delays 1..50 steps.
# The refractory period of 0.35 ms (refrac_time = int(0.35e-3 / 1e-4) = 3
# steps) makes a saturated neuron fire every fifth step, i.e. 2,000 Hz; the
# chain's populations saturate one after another within the first ~1,000
# cycles (without a refractory period they fire every other step, 5,000 Hz).
C2 = dict(npop=100, n=10000, fanout=2200, w0=0.45 * 1.3, inh_frac=0.2, dmax=50,
          remote_frac=0.2, drive_rate_hz=20.0, n_drive=10000, drive_w=1.0, drive_fanout=100,
          seed=2, i_offset=(14.5, 16.5), tau_refrac=0.35e-3)


def c1(**over):
    return standard(**{**C1, **over})


def c2(**over):
    return standard(**{**C2, **over})


def random_net(seed, n_lif=40, n_arr=5, n_bern=3, p_conn=0.2, dmax=5, allow_dup=False,
               shuffle=False, spike_prob=0.05, steps=200):
    """Small random network for fuzzing: LIF + array + Bernoulli sources with
    mixed-sign weights and random integer delays.  ``shuffle`` permutes the
    creation order of synapses (the exact oracle must follow it)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = Builder()
    if n_arr:
        sp = [np.sort(rng.choice(steps, size=rng.integers(0, steps // 4 + 1), replace=False))
              for _ in range(n_arr)]
        b.population(np.zeros((n_arr, LIF_NPARAM)), kind=ARRAY, spikes=sp)
    if n_bern:
        b.population(np.zeros((n_bern, LIF_NPARAM)), kind=BERNOULLI,
                     bern_p=np.uint32(int(spike_prob * 2 ** 32)))
    rows = np.tile(lif_row(0.0, -0.5, 0.0, 0.02, 0.02, 1e-3, 5e-3, 8e-3, 1.0, 0.0), (n_lif, 1))
    rows[:, 9] = rng.uniform(0.5, 1.6, n_lif)
    rows[:, 5] = rng.choice([0.0, 1e-4, 2e-4, 5e-4], n_lif)
    b.population(rows)
    n = b.n
    pre, post = np.nonzero(rng.random((n, n)) < p_conn)
    # no synapse may target a source neuron (they ignore input anyway)
    keep = b_kind(b)[post] == LIF
    pre, post = pre[keep], post[keep]
    if allow_dup:
        extra = rng.integers(0, len(pre), len(pre) // 5 + 1) if len(pre) else np.zeros(0, int)
        pre = np.concatenate([pre, pre[extra]])
        post = np.concatenate([post, post[extra]])
    m = len(pre)
    w = np.where(rng.random(m) < 0.3, -rng.uniform(0, 0.4, m), rng.uniform(0, 0.3, m))
    k = rng.integers(1, dmax + 1, m)
    # creation order: pre-major, then target, duplicates in ascending delay
    order = np.lexsort((k, post, pre))
    pre, post, w, k = pre[order], post[order], w[order], k[order]
    if shuffle and m:
        perm = rng.permutation(m)
        pre, post, w, k = pre[perm], post[perm], w[perm], k[perm]
    b.connect(pre, post, w.astype(np.float32), k.astype(np.float32) * np.float32(DT),
              (w < 0).astype(np.int8))
    spec = b.build()
    spec.seed = seed * 7919 + 17
    return spec


def b_kind(b: Builder):
    return np.concatenate(b.kinds)


def standard_test(depth=8, n=100, w=(0.7, 0.5, 0.6, 0.3), who=3) -> NetSpec:
    """The reference's benchmark program test/gpu/standard_test.cpp:55-138 at
    a chosen depth and population size, with its own neuron parameters
    (v_thresh -54 mV, v_reset -60 mV, c_m 0.2 nF, v_rest -74 mV, tau_m 10 ms,
    no refractory period, tau_syn 2 / 0.6 ms, i_offset 2 nA on g[1] only),
    weights (1e-9 * w_k / n, computed in float as there), 0.1 ms delays and
    call order:

    * g[1..depth] created first; for i = 2..depth, g[i-1] -> g[i] all-to-all
      and, when i % who == 0, g[1] -> g[i] (so those targets are fed by a
      later population first, then by an earlier one);
    * then for i = 1, 5, 9, ... (i + 1 <= depth) a side population p[i]:
      g[i] -> p[i] (w3) and p[i] -> g[i+1] (-w4, Inhibitory).

    README.md:87-102 names w = (0.7, 0.5, 0.6, 0.3) with who = 6 the 100 Hz
    regime of depth x 1000 neurons."""
    f = np.float32
    fthreshold, freset, c_m, v_rest = f(-54e-3), f(-60e-3), f(0.2e-9), f(-74e-3)
    tau_syn_e, tau_syn_i, tau_m, i_off, frefr, fv = f(2e-3), f(0.6e-3), f(10e-3), f(2e-9), f(0), \
        f(-74e-3)

    def row(ioff):
        return lif_row(fv, v_rest, freset, c_m, tau_m, frefr, tau_syn_e, tau_syn_i, fthreshold,
                       ioff, tau_syn_e, tau_syn_i)

    w1, w2, w3, w4 = (f(x) for x in w)
    nn = n * n
    weight6 = np.full(nn, f(1e-9) * w1 / f(n), np.float32)
    weight6_30 = np.full(nn, f(1e-9) * w2 / f(n), np.float32)
    weight5 = np.full(nn, f(1e-9) * w3 / f(n), np.float32)
    weight3 = np.full(nn, f(-1e-9) * w4 / f(n), np.float32)
    delay = np.full(nn, f(0.1e-3), np.float32)
    b = Builder()
    g = [None, b.population(np.tile(row(i_off), (n, 1)))]
    for _ in range(2, depth + 1):
        g.append(b.population(np.tile(row(f(0)), (n, 1))))
    for i in range(2, depth + 1):
        b.connect_all_to_all(g[i - 1], g[i], weight6, delay)
        if i % who == 0:
            b.connect_all_to_all(g[1], g[i], weight6_30, delay)
    i = 1
    while i + 1 <= depth:
        p = b.population(np.tile(row(f(0)), (n, 1)))
        b.connect_all_to_all(g[i], p, weight5, delay)
        b.connect_all_to_all(p, g[i + 1], weight3, delay, np.ones(nn, np.int8))
        i += 4
    return b.build()


def nonint_net(seed=1, n_lif=80, n_arr=12, p_conn=0.2, dmax=6, steps=300) -> NetSpec:
    """Small network whose delays are not whole steps, to exercise
    StaticSynapse::reset's rescale w *= expf(-(delay - dt * round(delay /
    dt)) / tau) (src/synapse/static/StaticSynapse.cpp:43-54) and the tau
    choice of CompositeNeuron::createSynapse (src/base/CompositeNeuron.h:
    62-79): an explicit per-synapse tau on a third of the synapses, else the
    destination's composite tau_syn_E or tau_syn_I by synapse type.  Types
    are drawn independently of the weight's sign (the type only picks the
    tau; the sign picks the E or I accumulator), composite taus differ from
    the neuron's own and between E and I, and delay fractions include
    exact halves (round() takes them away from zero).  Creation order is
    pre-major with ascending delays."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = Builder()
    sp = [np.sort(rng.choice(steps, size=rng.integers(steps // 10, steps // 4 + 1),
                             replace=False)) for _ in range(n_arr)]
    b.population(np.zeros((n_arr, LIF_NPARAM)), kind=ARRAY, spikes=sp)
    rows = np.tile(lif_row(0.0, -0.5, 0.0, 0.02, 0.02, 5e-4, 5e-3, 8e-3, 1.0, 0.0,
                           3e-3, 12e-3), (n_lif, 1))
    rows[:, 9] = rng.uniform(0.6, 1.5, n_lif)
    rows[n_lif // 2:, 10] = 7e-3       # a second population-like block with other taus
    rows[n_lif // 2:, 11] = 1.5e-3
    b.population(rows)
    n = b.n
    pre, post = np.nonzero(rng.random((n, n)) < p_conn)
    keep = post >= n_arr
    pre, post = pre[keep], post[keep]
    m = len(pre)
    k = rng.integers(1, dmax + 1, m)
    frac = rng.choice(np.array([0.0, 0.5, -0.3, 0.27, 0.41, -0.45, 0.12]), m)
    frac[k == 1] = np.abs(frac[k == 1])          # keep round(delay / dt) >= 1
    order = np.lexsort((k + frac, post, pre))
    pre, post, k, frac = pre[order], post[order], k[order], frac[order]
    w = np.where(rng.random(m) < 0.3, -rng.uniform(0, 0.5, m), rng.uniform(0, 0.4, m))
    stype = (rng.random(m) < 0.4).astype(np.int8)
    tau = np.where(rng.random(m) < 1 / 3, rng.uniform(1e-3, 10e-3, m), 0.0)
    d = ((k + frac) * DT).astype(np.float32)
    b.connect(pre, post, w.astype(np.float32), d, stype, tau.astype(np.float32))
    spec = b.build()
    spec.seed = seed
    return spec
