"""bsim_mlp_train_noalloc (include/bsim_mlp.h): the training entry point
that allocates no device memory per call, against bsim_mlp_train on an
identical model -- the same losses and parameters, bit for bit -- on two calls
in a row, the second with a smaller input; and its size limit."""
import numpy as np
import pytest

from bsim_amd import api, surrogate

pytestmark = pytest.mark.gpu

WORDS = 1 << 20     # BSIM_MLP_NOALLOC_WORDS


def _pair(seed):
    kw = dict(in_dim=8, hidden=512, layers=4, out_dim=200, max_batch=4096, seed=seed)
    return surrogate.MLP(**kw), surrogate.MLP(**kw)


def test_noalloc_training_equals_training():
    rng = np.random.default_rng(31)
    n = 20000
    x = rng.uniform(-1, 1, (n, 8)).astype(np.float32)
    y = rng.normal(size=(n, 200)).astype(np.float32)
    a, b = _pair(7)
    assert np.array_equal(a.get_params(), b.get_params())
    a.set_data(x, y)
    b.set_data(x, y)
    # call 1: 40 steps at batch 4096; call 2 (smaller): 3 steps at batch 512
    for steps, batch, lr in ((40, 4096, 1e-3), (3, 512, 5e-4)):
        rows = rng.integers(0, n, (steps, batch)).astype(np.int32)
        la = a.train(rows, batch, lr)
        lb = b.train_noalloc(rows, batch, lr)
        print(f"[gpu] noalloc steps={steps} batch={batch} loss {la[0]:.5f} -> {la[-1]:.5f}",
              flush=True)
        assert np.array_equal(la.view(np.uint32), lb.view(np.uint32))
        assert np.array_equal(a.get_params().view(np.uint32), b.get_params().view(np.uint32))
    assert la[-1] < 2.0
    a.close()
    b.close()


def test_noalloc_rejects_more_than_its_scratch():
    a, _b = _pair(3)
    _b.close()
    rng = np.random.default_rng(2)
    a.set_data(rng.uniform(-1, 1, (1000, 8)).astype(np.float32),
               rng.normal(size=(1000, 200)).astype(np.float32))
    steps = WORDS // 4096                      # 256: 256 + 256 * 4096 > 2^20
    rows = np.zeros((steps, 4096), np.int32)
    with pytest.raises(api.BsimError, match="exceeds the scratch"):
        a.train_noalloc(rows, 4096, 1e-3)
    p0 = a.get_params()
    ok = a.train_noalloc(rows[:255], 4096, 1e-3)   # 256 + 255 * 4096 <= 2^20
    assert len(ok) == 255 and not np.array_equal(a.get_params(), p0)
    a.close()
