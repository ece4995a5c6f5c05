"""Batched parameter sweeps: many independent networks advanced together.

A sweep point is one feed-forward chain network (bsim_net_gen_chain).  A
batch appends all of its networks to one host network -- disjoint neuron id
ranges, no synapse between networks, every draw keyed by network-local
indices -- and simulates them with one step launch per cycle (graph-replayed
by bsim_sim_step).  Each network's raster is the slice of the batch raster at
its id range, shifted to local ids, and is bit for bit the raster of the same
network simulated alone: a target's inputs, their order and the Bernoulli
draws do not depend on the other networks.

Sweep ranges follow the batched-sweep workload: 2-8 populations of
log-uniform size, uniform fan-in, w0, inhibitory fraction, Poisson drive rate
and maximum delay; see ``C3``.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import api, netspec

# parameter ranges of the batched-sweep workload (512 networks, 0.2 s each)
C3 = dict(n_nets=512, npop=(2, 8), size=(1000, 50000), fan_in=(100, 1000), w0=(0.3, 1.5),
          inh_frac=(0.1, 0.4), drive_hz=(5.0, 50.0), dmax=(6, 20), steps=2000)


@dataclass
class Point:
    """One sweep point (the arguments of Network.add_chain)."""
    sizes: list
    fan_in: int
    w0: float
    inh_frac: float
    dmax: int
    drive_rate_hz: float
    n_drive: int
    drive_fan_in: int
    drive_w: float
    seed: int

    def kwargs(self):
        return dict(sizes=self.sizes, fan_in=self.fan_in, w0=self.w0, inh_frac=self.inh_frac,
                    dmax=self.dmax, drive_rate_hz=self.drive_rate_hz, n_drive=self.n_drive,
                    drive_fan_in=self.drive_fan_in, drive_w=self.drive_w, seed=self.seed)

    @property
    def n(self):
        return self.n_drive + int(sum(self.sizes))

    def features(self):
        """The 6 swept parameters (for the surrogate): w0, inh_frac, drive
        rate, dmax, fan-in, mean population size."""
        return np.array([self.w0, self.inh_frac, self.drive_rate_hz, self.dmax, self.fan_in,
                         float(np.mean(self.sizes))], np.float64)


@dataclass
class Table:
    """n sweep points as arrays (one row per point; sizes zero past npop)."""
    npop: np.ndarray
    sizes: np.ndarray
    fan_in: np.ndarray
    w0: np.ndarray
    inh_frac: np.ndarray
    dmax: np.ndarray
    drive_hz: np.ndarray
    n_drive: np.ndarray
    seed: np.ndarray
    drive_fan_in: int
    drive_w: float

    def __len__(self):
        return len(self.npop)

    def point(self, i):
        k = int(self.npop[i])
        return Point(sizes=[int(x) for x in self.sizes[i, :k]], fan_in=int(self.fan_in[i]),
                     w0=float(self.w0[i]), inh_frac=float(self.inh_frac[i]),
                     dmax=int(self.dmax[i]), drive_rate_hz=float(self.drive_hz[i]),
                     n_drive=int(self.n_drive[i]), drive_fan_in=self.drive_fan_in,
                     drive_w=self.drive_w, seed=int(self.seed[i]))

    def points(self, idx=None):
        idx = range(len(self)) if idx is None else idx
        return [self.point(int(i)) for i in idx]


def sample_table(n_nets, seed=0, npop=C3["npop"], size=C3["size"], fan_in=C3["fan_in"],
                 w0=C3["w0"], inh_frac=C3["inh_frac"], drive_hz=C3["drive_hz"], dmax=C3["dmax"],
                 n_drive=None, drive_fan_in=50, drive_w=2.0):
    """n_nets sweep points drawn from the given ranges (seeded): population
    count and maximum delay uniform integers, sizes log-uniform, fan-in
    uniform integer, w0 / inhibitory fraction / drive rate uniform (fp32);
    the drive population has the first population's size unless n_drive is
    given."""
    rng = np.random.default_rng(seed)
    n = int(n_nets)
    k = rng.integers(npop[0], npop[1] + 1, n)
    sizes = np.exp(rng.uniform(np.log(size[0]), np.log(size[1]), (n, 8))).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    sizes[np.arange(8)[None, :] >= k[:, None]] = 0
    fi = rng.integers(fan_in[0], fan_in[1] + 1, n)
    w = rng.uniform(w0[0], w0[1], n).astype(np.float32)
    inh = rng.uniform(inh_frac[0], inh_frac[1], n).astype(np.float32)
    dm = rng.integers(dmax[0], dmax[1] + 1, n)
    dr = rng.uniform(drive_hz[0], drive_hz[1], n).astype(np.float32)
    sd = rng.integers(1, 2 ** 62, n)
    nd = sizes[:, 0].copy() if n_drive is None else np.full(n, int(n_drive), np.int64)
    return Table(npop=k, sizes=sizes, fan_in=fi, w0=w, inh_frac=inh, dmax=dm, drive_hz=dr,
                 n_drive=nd, seed=sd, drive_fan_in=int(drive_fan_in), drive_w=float(drive_w))


def sample(n_nets, seed=0, **ranges):
    """n_nets sweep points (sample_table's draws) as Point objects."""
    return sample_table(n_nets, seed, **ranges).points()


@dataclass
class Batch:
    """A host network holding several sweep networks, and where each starts."""
    net: api.Network
    points: list
    first: np.ndarray = field(default=None)

    @property
    def sizes(self):
        return np.array([p.n for p in self.points], np.int64)


def build(points):
    """All points appended to one host network."""
    net = api.Network()
    first = net.add_chains([p.kwargs() for p in points])
    return Batch(net=net, points=list(points), first=first)


def split_raster(counts, ids, first, sizes):
    """Per network: list over cycles of ascending local fired ids."""
    off = np.concatenate([[0], np.cumsum(counts)])
    out = [[] for _ in first]
    order = np.argsort(first)
    starts = first[order]
    for c in range(len(counts)):
        cyc = ids[off[c]:off[c + 1]]
        b = np.searchsorted(starts, cyc, side="right") - 1
        for j in np.unique(b):
            k = order[j]
            sel = cyc[b == j]
            out[k].append((c, sel - first[k]))
    res = []
    for k in range(len(first)):
        per = [np.zeros(0, np.int32) for _ in range(len(counts))]
        for c, sel in out[k]:
            per[c] = sel.astype(np.int32)
        res.append(per)
    return res


def run_batch(points, steps, tile=0, record=True, rank_every=None):
    """Simulates the points together (with the tile schedule rank_every, or
    the default); returns (per-network rasters or None, seconds of the
    stepping, simulation info)."""
    b = build(points)
    sim = api.Simulation(b.net, netspec.DT, tile=tile, rec_steps=steps if record else 0)
    if rank_every is not None:
        sim.set_schedule(rank_every)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.finish()
    sim.sync()
    secs = time.perf_counter() - t0
    info = sim.info()
    rasters = None
    if record:
        counts, ids, _ = sim.raster()
        rasters = split_raster(counts, ids, b.first, b.sizes)
    ev = sim.events()
    sim.close()
    return rasters, secs, dict(info, events=ev)


def run_batch_parts(points, steps, tile=0, record=True, part_synapses=int(1.5e9), rank_every=None):
    """run_batch for sweeps too large for one host build: the networks are
    generated, laid out and uploaded in parts of at most part_synapses
    synapses and joined on the device into one simulation
    (bsim_sim_create_chains).  Returns (per-network rasters or None, seconds
    of the stepping, simulation info)."""
    sim, first = api.Simulation.from_chains([p.kwargs() for p in points], netspec.DT, tile=tile,
                                            rec_steps=steps if record else 0,
                                            part_synapses=part_synapses)
    if rank_every is not None:
        sim.set_schedule(rank_every)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.finish()
    sim.sync()
    secs = time.perf_counter() - t0
    info = sim.info()
    rasters = None
    if record:
        counts, ids, _ = sim.raster()
        rasters = split_raster(counts, ids, first, np.array([p.n for p in points], np.int64))
    ev = sim.events()
    sim.close()
    return rasters, secs, dict(info, events=ev)


def run_solo(point, steps, tile=0, record=True):
    """Simulates one point alone; returns (raster or None, seconds of stepping)."""
    net = api.Network()
    net.add_chain(**point.kwargs())
    sim = api.Simulation(net, netspec.DT, tile=tile, rec_steps=steps if record else 0)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.finish()
    sim.sync()
    secs = time.perf_counter() - t0
    raster = sim.raster()[2] if record else None
    sim.close()
    return raster, secs


def autotune(points, steps=128, preroll=256, tiles=(192, 384, 768), ranks=(0, 1, 2, 4)):
    """The project's schedule autotuner for a batch: the tile size (targets
    per workgroup) and the tile schedule (tiles ranked heaviest first every
    r cycles, 0 = never, Simulation.set_schedule) with the fastest run of
    `steps` cycles of the batch, timed after `preroll` cycles so that the
    probe sees the swept regime rather than the cold start (no recording;
    results do not depend on either choice).  Returns ((tile, rank_every),
    {(tile, rank_every): seconds})."""
    times = {}
    b = build(points)
    for t in tiles:
        sim = api.Simulation(b.net, netspec.DT, tile=t)
        sim.step(preroll)
        sim.sync()
        for r in ranks:
            sim.set_schedule(r)
            sim.step(64)                      # graph capture, first ranking
            sim.sync()
            t0 = time.perf_counter()
            sim.step(steps)
            sim.sync()
            times[(t, r)] = time.perf_counter() - t0
        sim.close()
    return min(times, key=times.get), times


def autotune_tile(points, steps=128, preroll=256, tiles=(192, 384, 768)):
    """The tile size alone (the default schedule): (best tile, {tile: seconds})."""
    best, times = autotune(points, steps, preroll, tiles, ranks=(2,))
    return best[0], {t: s for (t, _), s in times.items()}
