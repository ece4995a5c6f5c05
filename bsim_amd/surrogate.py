"""Rate surrogate for sweep screening (include/bsim_mlp.h over ctypes).

A ReLU MLP (8 inputs, 4 hidden layers of 512, 200 outputs) is trained on the
GPU by the library's own bf16 MFMA kernels to predict a population's 1 ms
binned firing-rate trajectory over 200 ms from 8 features: the 6 sweep
parameters of its network (w0, inhibitory fraction, drive rate, maximum
delay, fan-in, mean population size), its index in the chain and its size.
The training set comes from the simulator: batched sweep runs (sweep.py)
recorded and binned per population.

Screening: the surrogate scores many candidate sweep points, the best are
re-simulated, and the simulator's rates judge both the pick and the
surrogate's error (``screen``).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import api, netspec, sweep

P = C.c_void_p
DT = float(netspec.DT)
BIN_STEPS = 10          # 1 ms bins at dt = 0.1 ms
N_BINS = 200            # 200 ms
RATE_SCALE = 50.0       # network output = rate / RATE_SCALE (Hz)

_SIGS = {
    "bsim_mlp_create": (P, [P]),
    "bsim_mlp_destroy": (None, [P]),
    "bsim_mlp_num_params": (C.c_int64, [P]),
    "bsim_mlp_layout": (C.c_int32, [P, P, C.c_int32]),
    "bsim_mlp_get_params": (C.c_int32, [P, P]),
    "bsim_mlp_set_params": (C.c_int32, [P, P]),
    "bsim_mlp_set_data": (C.c_int32, [P, P, P, C.c_int64]),
    "bsim_mlp_train": (C.c_int32, [P, P, C.c_int32, C.c_int32, C.c_float, P]),
    "bsim_mlp_train_noalloc": (C.c_int32, [P, P, C.c_int32, C.c_int32, C.c_float, P]),
    "bsim_mlp_gradients": (C.c_int32, [P, P, P, C.c_int32, P, P]),
    "bsim_mlp_predict": (C.c_int32, [P, P, C.c_int64, P]),
}
_BOUND = False


def lib():
    global _BOUND
    L = api.lib()
    if not _BOUND:
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _BOUND = True
    return L


class MLPConfig(C.Structure):
    _fields_ = [("in_dim", C.c_int32), ("hidden", C.c_int32), ("layers", C.c_int32),
                ("out_dim", C.c_int32), ("max_batch", C.c_int32), ("device", C.c_int32),
                ("seed", C.c_uint64), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("pad", C.c_float)]


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


def _check(rc, what):
    if rc < 0:
        api._err(rc, what)
    return rc


class MLP:
    """A bsim_mlp handle: fp32 master weights, bf16 GEMM copies, Adam state."""

    def __init__(self, in_dim=8, hidden=512, layers=4, out_dim=N_BINS, max_batch=4096, seed=1,
                 beta1=0.9, beta2=0.999, eps=1e-8, device=0):
        self.cfg = MLPConfig(in_dim, hidden, layers, out_dim, max_batch, device, seed, beta1,
                             beta2, eps, 0.0)
        self._h = lib().bsim_mlp_create(C.byref(self.cfg))
        if not self._h:
            api._err(-1, "bsim_mlp_create")
        dims = np.zeros(16, np.int32)
        nl = _check(lib().bsim_mlp_layout(self._h, dims.ctypes.data, 16), "bsim_mlp_layout")
        self.dims = [int(x) for x in dims[:nl + 1]]
        self.num_params = int(lib().bsim_mlp_num_params(self._h))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().bsim_mlp_destroy(h)
            self._h = None

    __del__ = close

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        _check(lib().bsim_mlp_get_params(self._h, p.ctypes.data), "bsim_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.num_params
        _check(lib().bsim_mlp_set_params(self._h, p.ctypes.data), "bsim_mlp_set_params")

    def set_data(self, x, y):
        x, y = _f32(x), _f32(y)
        assert x.shape[1] == self.cfg.in_dim and y.shape[1] == self.cfg.out_dim
        assert x.shape[0] == y.shape[0]
        self._data = (x, y)
        _check(lib().bsim_mlp_set_data(self._h, x.ctypes.data, y.ctypes.data, x.shape[0]),
               "bsim_mlp_set_data")

    def train(self, rows, batch, lr):
        """Adam steps on the uploaded set; rows[steps][batch]. Returns losses."""
        rows = np.ascontiguousarray(np.asarray(rows, np.int32))
        steps = rows.shape[0]
        loss = np.zeros(max(steps, 1), np.float32)
        _check(lib().bsim_mlp_train(self._h, rows.ctypes.data, steps, batch, float(lr),
                                    loss.ctypes.data), "bsim_mlp_train")
        return loss[:steps]

    def train_noalloc(self, rows, batch, lr):
        """train() through bsim_mlp_train_noalloc: no device allocation per
        call (fixed per-GPU scratch); round_up(steps, 64) + steps * batch must
        fit BSIM_MLP_NOALLOC_WORDS."""
        rows = np.ascontiguousarray(np.asarray(rows, np.int32))
        steps = rows.shape[0]
        loss = np.zeros(max(steps, 1), np.float32)
        _check(lib().bsim_mlp_train_noalloc(self._h, rows.ctypes.data, steps, batch, float(lr),
                                            loss.ctypes.data), "bsim_mlp_train_noalloc")
        return loss[:steps]

    def gradients(self, x, y):
        x, y = _f32(x), _f32(y)
        g = np.zeros(self.num_params, np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().bsim_mlp_gradients(self._h, x.ctypes.data, y.ctypes.data, x.shape[0],
                                        g.ctypes.data, loss.ctypes.data), "bsim_mlp_gradients")
        return g, float(loss[0])

    def predict(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.cfg.out_dim), np.float32)
        _check(lib().bsim_mlp_predict(self._h, x.ctypes.data, x.shape[0], y.ctypes.data),
               "bsim_mlp_predict")
        return y


def flops_per_step(dims, batch, in_dim, out_dim):
    """Algorithmic flops of one training step at the unpadded widths:
    forward 2BKN per layer, backward 4BKN (the input layer's input gradient
    is not computed: 2BKN)."""
    w = [in_dim] + dims[1:-1] + [out_dim]
    f = 0
    for l in range(len(w) - 1):
        kn = w[l] * w[l + 1]
        f += 2 * batch * kn * (3 if l > 0 else 2)
    return f


def flops_forward(dims, rows, in_dim, out_dim):
    w = [in_dim] + dims[1:-1] + [out_dim]
    return sum(2 * rows * w[l] * w[l + 1] for l in range(len(w) - 1))


# ---- features -------------------------------------------------------------

def _lognorm(s, size):
    lo, hi = np.log(size[0]), np.log(size[1])
    return (np.log(np.maximum(s, 1.0)) - 0.5 * (lo + hi)) / (0.5 * (hi - lo))


def features(points, size=sweep.C3["size"]):
    """One row per (point, population): the 8 normalised inputs.  Also
    returns the (point, population) index of each row.  points: a list of
    sweep.Point or a sweep.Table."""
    if not isinstance(points, sweep.Table):
        tab = sweep.Table(npop=np.array([len(p.sizes) for p in points], np.int64),
                          sizes=np.array([list(p.sizes) + [0] * (8 - len(p.sizes)) for p in points],
                                         np.int64).reshape(-1, 8),
                          fan_in=np.array([p.fan_in for p in points]),
                          w0=np.array([p.w0 for p in points]),
                          inh_frac=np.array([p.inh_frac for p in points]),
                          dmax=np.array([p.dmax for p in points]),
                          drive_hz=np.array([p.drive_rate_hz for p in points]),
                          n_drive=np.array([p.n_drive for p in points]),
                          seed=np.zeros(len(points), np.int64), drive_fan_in=0, drive_w=0.0)
    else:
        tab = points
    k = tab.npop.astype(np.int64)
    pt = np.repeat(np.arange(len(k)), k)
    pop = np.arange(len(pt)) - np.repeat(np.cumsum(k) - k, k)
    mean_size = tab.sizes.sum(axis=1) / k
    X = np.stack([(tab.w0[pt] - 0.9) / 0.6, (tab.inh_frac[pt] - 0.25) / 0.15,
                  (tab.drive_hz[pt] - 27.5) / 22.5, (tab.dmax[pt] - 13.0) / 7.0,
                  (tab.fan_in[pt] - 550.0) / 450.0, _lognorm(mean_size[pt], size),
                  (pop - 3.5) / 3.5, _lognorm(tab.sizes[pt, pop].astype(np.float64), size)],
                 axis=1).astype(np.float32)
    return X, np.stack([pt, pop], axis=1).astype(np.int64)


# ---- training data from the simulator --------------------------------------

def binned_rates(counts, ids, pop_lo, pop_n, bin_steps=BIN_STEPS):
    """Per population [lo, lo + n): firing rate (Hz) per bin of bin_steps
    cycles, from a raster (spikes per cycle, ascending ids per cycle)."""
    nb = len(counts) // bin_steps
    cyc = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    order = np.argsort(pop_lo)
    lo_sorted = np.asarray(pop_lo)[order]
    j = np.searchsorted(lo_sorted, ids, side="right") - 1
    inside = (j >= 0) & (ids < lo_sorted[np.maximum(j, 0)] + np.asarray(pop_n)[order][np.maximum(j, 0)])
    keep = inside & (cyc < nb * bin_steps)
    pop = order[j[keep]]
    h = np.bincount(pop * nb + cyc[keep] // bin_steps, minlength=len(pop_lo) * nb)
    h = h.reshape(len(pop_lo), nb).astype(np.float64)
    return h / (np.asarray(pop_n, np.float64)[:, None] * bin_steps * DT)


def simulate_rates(points, steps=N_BINS * BIN_STEPS, tile=0, max_synapses=4e8, log=None):
    """Simulates the points in batches of at most ~max_synapses synapses and
    returns, per point, its populations' binned rates [npop][steps / 10] (Hz)
    and the simulation seconds."""
    out = [None] * len(points)
    secs = 0.0
    i = 0
    while i < len(points):
        j, syn = i, 0.0
        while j < len(points) and (j == i or syn + _synapses(points[j]) <= max_synapses):
            syn += _synapses(points[j])
            j += 1
        chunk = points[i:j]
        b = sweep.build(chunk)
        sim = api.Simulation(b.net, DT, tile=tile, rec_steps=steps)
        t0 = time.perf_counter()
        sim.step(steps)
        sim.finish()
        sim.sync()
        secs += time.perf_counter() - t0
        counts, ids, _ = sim.raster()
        sim.close()
        lo, n, owner = [], [], []
        for k, p in enumerate(chunk):
            base = int(b.first[k]) + p.n_drive
            for s in p.sizes:
                lo.append(base)
                n.append(s)
                owner.append(k)
                base += s
        r = binned_rates(counts, ids, np.asarray(lo, np.int64), np.asarray(n, np.int64))
        owner = np.asarray(owner)
        for k in range(len(chunk)):
            out[i + k] = r[owner == k]
        if log:
            log(f"simulated {j}/{len(points)} networks ({secs:.1f}s)")
        i = j
    return out, secs


def _synapses(p):
    return float(p.fan_in) * sum(p.sizes[1:]) + float(p.drive_fan_in) * p.sizes[0]


def dataset(points, rates, size=sweep.C3["size"]):
    """(X[rows][8], Y[rows][200] in rate / RATE_SCALE, (point, population) per row)."""
    X, idx = features(points, size)
    Y = np.stack([rates[k][i] for k, i in idx]).astype(np.float32) / RATE_SCALE
    return X, Y, idx


def rmse_hz(mlp, X, Y):
    """Rate RMSE (Hz) of the model on (X, Y)."""
    d = (mlp.predict(X) - Y) * RATE_SCALE
    return float(np.sqrt(np.mean(d.astype(np.float64) ** 2)))


def fit(mlp, X, Y, epochs, batch=4096, lr=1e-3, seed=0, lr_final=None, val=None, val_every=50):
    """Trains on (X, Y) for the given epochs of shuffled minibatches (the last
    partial batch of an epoch is dropped); lr decays geometrically to
    lr_final.  With val = (Xv, Yv), the rate RMSE on it is checked every
    val_every epochs and the best parameters are kept (early stopping).
    Returns (per-step losses, seconds of training, {"best_epoch", "val_rmse_hz"})."""
    mlp.set_data(X, Y)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    per = n // batch
    if per == 0:
        raise ValueError(f"need at least one batch of {batch} rows, have {n}")
    losses, secs = [], 0.0
    lr_final = lr if lr_final is None else lr_final
    best = (np.inf, None, -1)
    for e in range(epochs):
        rows = rng.permutation(n)[:per * batch].reshape(per, batch).astype(np.int32)
        rate = lr * (lr_final / lr) ** (e / max(1, epochs - 1))
        t0 = time.perf_counter()
        losses.append(mlp.train(rows, batch, rate))
        secs += time.perf_counter() - t0
        if val is not None and ((e + 1) % val_every == 0 or e + 1 == epochs):
            r = rmse_hz(mlp, *val)
            if r < best[0]:
                best = (r, mlp.get_params(), e + 1)
    info = {"best_epoch": epochs, "val_rmse_hz": None}
    if val is not None:
        mlp.set_params(best[1])
        info = {"best_epoch": best[2], "val_rmse_hz": best[0]}
    return np.concatenate(losses), secs, info


# ---- screening -------------------------------------------------------------

def last_pop_mean(rates_pred_or_sim, bins=slice(N_BINS // 2, N_BINS)):
    return float(np.mean(rates_pred_or_sim[bins]))


def screen(mlp, candidates, target_hz, top_frac=0.01, size=sweep.C3["size"]):
    """Scores every candidate (a list of points or a sweep.Table) by the
    surrogate's predicted mean rate of its last population over the last
    100 ms against target_hz.  Returns (indices of the best top_frac,
    function k -> predicted rates [npop][200] (Hz) of candidate k, seconds of
    prediction)."""
    X, idx = features(candidates, size)
    t0 = time.perf_counter()
    Yp = mlp.predict(X)
    secs = time.perf_counter() - t0
    k = np.bincount(idx[:, 0], minlength=len(candidates))
    start = np.cumsum(k) - k
    last = start + k - 1
    score = np.abs(Yp[last, N_BINS // 2:].mean(axis=1) * RATE_SCALE - target_hz)
    ntop = max(1, int(round(top_frac * len(candidates))))
    top = np.argsort(score, kind="stable")[:ntop]

    def pred(c):
        return Yp[start[c]:start[c] + k[c]].astype(np.float64) * RATE_SCALE

    return top, pred, secs


def judge(picks, pred, sim_rates, target_hz):
    """Simulator verdict on picked candidates (pred: screen's function,
    sim_rates: candidate -> simulated rates): surrogate rate RMSE (Hz, every
    population and bin), the fraction of populations whose predicted mean
    rate is within 10 % of the simulated one, and the fraction of picks whose
    simulated last-population rate is within 10 % of target_hz."""
    err, within, hit = [], [], []
    for k in picks:
        ps, ss = pred(k), sim_rates[k]
        err.append((ps - ss).ravel())
        pm, sm = ps.mean(axis=1), ss.mean(axis=1)
        within.extend(list(np.abs(pm - sm) <= 0.1 * np.maximum(sm, 1e-9)))
        hit.append(abs(last_pop_mean(ss[-1]) - target_hz) <= 0.1 * target_hz)
    e = np.concatenate(err)
    return dict(rmse_hz=float(np.sqrt(np.mean(e * e))), within_10pct=float(np.mean(within)),
                hit_rate=float(np.mean(hit)))
