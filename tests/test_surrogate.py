"""Rate surrogate: the oracle pinned against torch (float64) and finite
differences; the GPU's bf16 MFMA forward, gradients and Adam against the
oracle; training, and screening judged by the simulator."""
import os
import re

import numpy as np
import pytest

import mlp_oracle as mo
from bsim_amd import surrogate, sweep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rand_params(dims, in_dim, out_dim, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for l in range(len(dims) - 1):
        n, k = dims[l + 1], dims[l]
        fan = in_dim if l == 0 else k
        W = rng.uniform(-1, 1, (n, k)) * np.sqrt(6.0 / fan) * scale
        b = rng.uniform(-0.1, 0.1, n)
        if l == 0:
            W[:, in_dim:] = 0
        if l == len(dims) - 2:
            W[out_dim:] = 0
            b[out_dim:] = 0
        Ws.append(W)
        bs.append(b)
    return mo.pack(Ws, bs)


# ---- oracle pinning (CPU) ---------------------------------------------------

def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 0.0,
                  3.0e38, 1e-40], np.float32)
    want = np.array([0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x0000, 0x7F62, 0x0001], np.uint32)
    got = mo.bf16(x).view(np.uint32) >> 16
    assert np.array_equal(got, want), [hex(v) for v in got]


def _torch_model(flat, dims):
    import torch
    Ws, bs = mo.unpack(flat, dims)
    Wt = [torch.tensor(W, dtype=torch.float64, requires_grad=True) for W in Ws]
    bt = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for b in bs]
    return Wt, bt


def _torch_forward(Wt, bt, x, dims):
    import torch
    h = torch.zeros((x.shape[0], dims[0]), dtype=torch.float64)
    h[:, :x.shape[1]] = torch.tensor(x, dtype=torch.float64)
    for l, (W, b) in enumerate(zip(Wt, bt)):
        h = torch.nn.functional.linear(h, W, b)
        if l + 1 < len(Wt):
            h = torch.relu(h)
    return h


def test_oracle_matches_torch_autograd():
    import torch
    in_dim, out_dim = 8, 50
    dims = mo.dims_of(in_dim, 64, 3, out_dim)
    flat = rand_params(dims, in_dim, out_dim, 1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(32, in_dim))
    y = rng.normal(size=(32, out_dim))
    loss, g = mo.loss_and_grad(flat, dims, x, y, emulate=False)
    Wt, bt = _torch_model(flat, dims)
    out = _torch_forward(Wt, bt, x, dims)
    tl = torch.mean((out[:, :out_dim] - torch.tensor(y)) ** 2)
    tl.backward()
    tg = np.concatenate([np.concatenate([W.grad.numpy().ravel(), b.grad.numpy().ravel()])
                         for W, b in zip(Wt, bt)])
    assert abs(loss - tl.item()) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(g, tg, rtol=1e-10, atol=1e-14)
    o, _ = mo.forward(flat, dims, x, emulate=False)
    np.testing.assert_allclose(o, out.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_oracle_gradient_finite_differences():
    in_dim, out_dim = 8, 20
    dims = mo.dims_of(in_dim, 64, 2, out_dim)
    flat = rand_params(dims, in_dim, out_dim, 3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(16, in_dim))
    y = rng.normal(size=(16, out_dim))
    _, g = mo.loss_and_grad(flat, dims, x, y, emulate=False)
    nz = np.flatnonzero(np.abs(g) > 1e-6)
    for i in rng.choice(nz, 40, replace=False):
        h = 1e-6
        a, b = flat.copy(), flat.copy()
        a[i] += h
        b[i] -= h
        fd = (mo.loss_and_grad(a, dims, x, y, emulate=False)[0] -
              mo.loss_and_grad(b, dims, x, y, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])), (i, fd, g[i])


def test_oracle_adam_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=300)
    grads = [rng.normal(size=300) * s for s in (1.0, 0.01, 3.0)]
    tp = torch.tensor(p0.copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(300), np.zeros(300)
    for t, g in enumerate(grads, 1):
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = mo.adam(p, m, v, g, t, 1e-3, f32=False)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_adam_first_step_closed_form():
    """First step: m_hat = g and v_hat = g^2 exactly (fp32 betas), so the
    step is lr * g / (|g| + eps)."""
    g = np.array([1e-3, -2.0, 5.0, 0.0, 1e-5], np.float32)
    p0 = np.zeros(5, np.float32)
    p, _, _ = mo.adam(p0, np.zeros(5, np.float32), np.zeros(5, np.float32), g, 1, 1e-3, eps=1e-12)
    step = p0 - p
    np.testing.assert_allclose(step[[0, 1, 2, 4]], np.float32(1e-3) * np.sign(g[[0, 1, 2, 4]]),
                               rtol=2e-6)
    assert step[3] == 0


def test_emulation_stays_near_exact():
    in_dim, out_dim = 8, 200
    dims = mo.dims_of(in_dim, 512, 4, out_dim)
    flat = rand_params(dims, in_dim, out_dim, 6)
    x = np.random.default_rng(7).normal(size=(64, in_dim))
    a, _ = mo.forward(flat, dims, x, emulate=True)
    b, _ = mo.forward(flat, dims, x, emulate=False)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert 1e-5 < rel < 3e-2


def test_flops_per_step():
    dims = mo.dims_of(8, 512, 4, 200)
    f = surrogate.flops_per_step(dims, 4096, 8, 200)
    fwd = 2 * 4096 * (8 * 512 + 3 * 512 * 512 + 512 * 200)
    assert f == 3 * fwd - 2 * 4096 * 8 * 512
    assert surrogate.flops_forward(dims, 4096, 8, 200) == fwd


# ---- host side (CPU) ---------------------------------------------------------

def test_library_exports_mlp_header_symbols():
    src = open(os.path.join(ROOT, "include", "bsim_mlp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(bsim_mlp_[a-z0-9_]+)\s*\(", src)))
    L = surrogate.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(surrogate._SIGS)


@pytest.mark.parametrize("bad", [dict(hidden=100), dict(layers=0), dict(in_dim=65),
                                 dict(max_batch=100), dict(beta1=1.0)])
def test_mlp_rejects_bad_config(bad):
    kw = dict(in_dim=8, hidden=512, layers=4, out_dim=200, max_batch=4096)
    kw.update(bad)
    with pytest.raises(Exception, match="invalid MLP config"):
        surrogate.MLP(**kw)


def test_binned_rates():
    # two populations [10, 14) and [20, 22); 20 cycles, bins of 10
    counts = np.array([2, 0, 1] + [0] * 7 + [3] + [0] * 9, np.int32)
    ids = np.array([10, 21, 13, 5, 11, 20], np.int32)
    r = surrogate.binned_rates(counts, ids, np.array([10, 20]), np.array([4, 2]), bin_steps=10)
    dt = surrogate.DT
    np.testing.assert_allclose(r[0], np.array([2, 1]) / (4 * 10 * dt))
    np.testing.assert_allclose(r[1], np.array([1, 1]) / (2 * 10 * dt))


def test_features_one_row_per_population():
    pts = sweep.sample(5, seed=1)
    X, idx = surrogate.features(pts)
    assert X.shape == (sum(len(p.sizes) for p in pts), 8)
    assert np.all(np.abs(X) <= 1.0 + 1e-6)
    assert [tuple(r) for r in idx[:len(pts[0].sizes)]] == [(0, i) for i in range(len(pts[0].sizes))]


# ---- GPU ---------------------------------------------------------------------

def _model(**kw):
    a = dict(in_dim=8, hidden=512, layers=4, out_dim=200, max_batch=1024, seed=3)
    a.update(kw)
    return surrogate.MLP(**a)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [1000, 4096])   # 1000: not a multiple of 64
def test_gpu_forward_matches_oracle(batch):
    m = _model(max_batch=max(batch, 1024))
    flat = m.get_params()
    rng = np.random.default_rng(8)
    x = rng.normal(size=(batch, 8)).astype(np.float32)
    y = m.predict(x)
    ref, _ = mo.forward(flat, m.dims, x, emulate=True)
    assert y.shape == (batch, 200)
    assert _rel(y, ref[:, :200]) < 2e-3
    # per-sample: a transposed or shifted output tile would fail here
    np.testing.assert_allclose(y, ref[:, :200], rtol=2e-2, atol=2e-2 * np.abs(ref).max())
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [512, 4096])    # 4096: the training batch
def test_gpu_gradients_match_oracle(batch):
    m = _model(max_batch=max(batch, 1024))
    flat = m.get_params()
    rng = np.random.default_rng(9)
    x = rng.normal(size=(batch, 8)).astype(np.float32)
    y = rng.normal(size=(batch, 200)).astype(np.float32)
    g, loss = m.gradients(x, y)
    rl, rg = mo.loss_and_grad(flat, m.dims, x, y, emulate=True)
    assert abs(loss - rl) <= 1e-4 * rl
    Wg, bg = mo.unpack(g, m.dims)
    Wr, br = mo.unpack(rg, m.dims)
    for l in range(len(Wg)):
        assert _rel(Wg[l], Wr[l]) < 1e-2, ("W", l, _rel(Wg[l], Wr[l]))
        assert _rel(bg[l], br[l]) < 1e-2, ("b", l, _rel(bg[l], br[l]))
    # the gradient is deterministic
    g2, loss2 = m.gradients(x, y)
    assert np.array_equal(g, g2) and loss == loss2
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [512, 4096])
def test_gpu_adam_step_matches_oracle(batch):
    m = _model(max_batch=max(batch, 1024))
    p0 = m.get_params()
    rng = np.random.default_rng(10)
    x = rng.normal(size=(batch, 8)).astype(np.float32)
    y = rng.normal(size=(batch, 200)).astype(np.float32)
    g, loss = m.gradients(x, y)
    m.set_data(x, y)
    losses = m.train(np.arange(batch, dtype=np.int32)[None, :], batch, 1e-3)
    assert losses[0] == loss
    p1 = m.get_params()
    want, _, _ = mo.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, 1e-3)
    np.testing.assert_allclose(p1, want, rtol=1e-6, atol=1e-9)
    m.close()


@pytest.mark.gpu
def test_gpu_training_is_deterministic_and_fits():
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, size=(4096, 8)).astype(np.float32)
    t = np.linspace(0, 1, 200, dtype=np.float32)
    y = (np.sin(3 * x[:, :1] + 2 * t[None, :]) * x[:, 1:2] + 0.5 * x[:, 2:3] ** 2).astype(np.float32)
    out = []
    for _ in range(2):
        m = _model(max_batch=1024)
        losses, _, _ = surrogate.fit(m, x, y, epochs=40, batch=1024, lr=1e-3, seed=0)
        out.append((losses, m.get_params()))
        m.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    losses = out[0][0]
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])


SMALL = dict(npop=(2, 4), size=(150, 600), fan_in=(20, 120), dmax=(6, 20), drive_fan_in=20)


@pytest.mark.gpu
def test_gpu_screening_beats_random_picks():
    """Train on simulated rates, screen candidates for a rare last-population
    rate, and let the simulator judge: the surrogate's picks hit the target
    more often than the same number of random picks."""
    size = SMALL["size"]
    train_pts = sweep.sample(300, seed=21, **SMALL)
    rates, _ = surrogate.simulate_rates(train_pts)
    X, Y, _ = surrogate.dataset(train_pts, rates, size=size)
    m = _model(max_batch=512, seed=5)
    losses, _, _ = surrogate.fit(m, X, Y, epochs=150, batch=256, lr=1e-3, seed=1, lr_final=1e-4)
    assert losses[-1] < 0.5 * losses[0]
    last = np.array([surrogate.last_pop_mean(r[-1]) for r in rates])
    target = float(np.quantile(last, 0.9))
    cands = sweep.sample(2000, seed=22, **SMALL)
    top, pred, _ = surrogate.screen(m, cands, target, top_frac=0.05, size=size)
    rnd = np.random.default_rng(23).choice(len(cands), len(top), replace=False)
    picks = sorted(set(top.tolist()) | set(rnd.tolist()))
    sim, _ = surrogate.simulate_rates([cands[k] for k in picks])
    sim_by = dict(zip(picks, sim))
    simr = [sim_by.get(k) for k in range(len(cands))]
    js = surrogate.judge(top, pred, simr, target)
    jr = surrogate.judge(rnd, pred, simr, target)
    print("surrogate", js, "random", jr, "target", target)
    # 100 picks each: a 10-point margin is about 2 standard errors
    assert js["hit_rate"] > jr["hit_rate"] + 0.10, (js, jr)
    m.close()
