"""CPU oracle of the rate surrogate (TEST INFRASTRUCTURE).

Only tests/ and bench.py's checks import this module; the library's
surrogate (bsim_amd/csrc/mlp.hip) never calls it.

The reference has no learned component, so there is no reference file to
follow: this is the textbook back-propagation of a fully connected ReLU
network under a mean-squared-error loss, and Adam (Kingma & Ba, "Adam: a
method for stochastic optimization", ICLR 2015, Algorithm 1, bias corrections
applied to both moments).  Pinned in tests/test_surrogate.py against torch's
CPU autograd and torch.optim.Adam in float64 (independent implementations),
central finite differences, and Adam's first-step closed form
(|step| = lr wherever the gradient is nonzero and eps is negligible).

Parameters use the library's flat padded layout: for each linear layer l the
weights W_l[dims[l + 1]][dims[l]] row-major, then the bias b_l[dims[l + 1]].

Two modes:
* ``emulate=False``: float64 throughout, weights taken as given;
* ``emulate=True``: the GPU's rounding points -- bf16 (round to nearest even)
  inputs, weights, hidden activations and activation gradients; fp32 biases,
  pre-activations, output, loss gradient and gradient sums -- with float64
  products and sums in between (the GPU sums in fp32 in another order, so the
  two agree to a tolerance, not bit for bit).
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def dims_of(in_dim, hidden, layers, out_dim):
    """Padded widths: 64 inputs, hidden x layers, out_dim rounded up to 64."""
    return [64] + [hidden] * layers + [(out_dim + 63) // 64 * 64]


def unpack(flat, dims):
    Ws, bs, off = [], [], 0
    for l in range(len(dims) - 1):
        n, k = dims[l + 1], dims[l]
        Ws.append(np.asarray(flat[off:off + n * k]).reshape(n, k))
        off += n * k
        bs.append(np.asarray(flat[off:off + n]))
        off += n
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in zip(Ws, bs)])


def forward(flat, dims, x, emulate=True):
    """Returns (out[B][dims[-1]], activations [input, hidden 1..L])."""
    Ws, bs = unpack(flat, dims)
    B = x.shape[0]
    X = np.zeros((B, dims[0]), np.float64)
    X[:, :x.shape[1]] = bf16(x) if emulate else x
    H = X
    acts = [H]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Wl = bf16(W).astype(np.float64) if emulate else np.asarray(W, np.float64)
        Z = H @ Wl.T
        if emulate:
            Z = (Z.astype(np.float32) + np.asarray(b, np.float32)).astype(np.float64)
        else:
            Z = Z + b
        if l + 1 < len(Ws):
            H = np.maximum(Z, 0.0)
            if emulate:
                H = bf16(H).astype(np.float64)
            acts.append(H)
        else:
            return Z, acts


def loss_and_grad(flat, dims, x, y, emulate=True):
    """Mean squared error over the batch and the y.shape[1] outputs, and its
    gradient in the flat layout."""
    Ws, bs = unpack(flat, dims)
    out, acts = forward(flat, dims, x, emulate)
    B, nout = y.shape
    if emulate:
        diff = out[:, :nout].astype(np.float32) - np.asarray(y, np.float32)
        scale = np.float32(2.0) / (np.float32(B) * np.float32(nout))
        dz = np.zeros((B, dims[-1]), np.float64)
        dz[:, :nout] = bf16(scale * diff)
        loss = float(np.sum(diff.astype(np.float64) ** 2) / (B * nout))
    else:
        diff = out[:, :nout] - y
        dz = np.zeros((B, dims[-1]), np.float64)
        dz[:, :nout] = 2.0 * diff / (B * nout)
        loss = float(np.sum(diff ** 2) / (B * nout))
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gW[l] = dz.T @ acts[l]
        gb[l] = dz.sum(axis=0)
        if l == 0:
            break
        Wl = bf16(Ws[l]).astype(np.float64) if emulate else np.asarray(Ws[l], np.float64)
        dh = dz @ Wl
        if emulate:
            dh = bf16(dh).astype(np.float64)
        dz = np.where(acts[l] > 0, dh, 0.0)
    return loss, pack(gW, gb)


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8, f32=True):
    """One Adam step; returns (p, m, v).  f32: the GPU's fp32 operation order."""
    if f32:
        f = np.float32
        p, m, v, g = (np.asarray(a, np.float32) for a in (p, m, v, g))
        m = f(b1) * m + (f(1) - f(b1)) * g
        v = f(b2) * v + (f(1) - f(b2)) * (g * g)
        # bias corrections from the fp32 betas the moments use (in double)
        c1 = f(1.0 - float(f(b1)) ** t)
        c2 = f(1.0 - float(f(b2)) ** t)
        p = p - f(lr) * ((m / c1) / (np.sqrt(v / c2) + f(eps)))
        return p, m, v
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** t)
    vhat = v / (1 - b2 ** t)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v
