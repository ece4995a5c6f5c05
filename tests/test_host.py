"""Host-side checks of the product library (no GPU calls): it loads, exports
every symbol include/bsim_amd.h declares, computes the reference's constants
bit for bit, lays synapses out in (pre, delay, creation) order per tile, and
rejects networks the GPU step cannot run with a clear error."""
import os
import re

import numpy as np
import pytest

import oracle
from bsim_amd import api, netspec

import golden_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DT = 1e-4


def header_functions():
    src = open(os.path.join(ROOT, "include", "bsim_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bsim_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    L = api.lib()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(api._SIGS), set(names) ^ set(api._SIGS)
    assert L.bsim_abi_version() == 1


@pytest.mark.parametrize("name", [n for n in golden_io.case_names() if n.endswith("_f32")])
def test_constants_match_reference(name):
    meta, spec, dtype, g = golden_io.load(name)
    net = api.Network.from_spec(spec)
    c9, rt = net.lif_constants(meta["dt"])
    lif = spec.kind == netspec.LIF
    assert np.array_equal(c9[:, :8].copy().view(np.uint32)[lif], g["consts_bits"][lif][:, :8])
    np.testing.assert_array_equal(rt[lif], g["refrac_time"][lif])


def test_layout_is_pre_major_within_groups():
    """Synapses of a group of tiles are stored in (pre, delay, tile, part,
    creation) order: each target's inputs keep their (pre, delay, creation)
    order and the group's tiles share the cache lines of every row."""
    spec = netspec.random_net(5, n_lif=700, p_conn=0.03, dmax=7, allow_dup=True)
    net = api.Network.from_spec(spec)
    tile = 192
    L = net.layout(DT, tile)
    w_ref, ds = oracle.synapse_reset(spec, DT)
    ts = L["tile_syn"]
    k = 0
    ngroups = 0
    while k < L["ntiles"]:
        k1 = k
        while k1 < L["ntiles"] and ts[k1] == ts[k]:
            k1 += 1
        a, b = ts[k], (ts[k1] if k1 < L["ntiles"] else L["s"])
        sel = np.nonzero((spec.post >= k * tile) & (spec.post < k1 * tile))[0]   # creation order
        tl = spec.post[sel] // tile
        part = (spec.post[sel] - tl * tile) // (tile // 3)
        order = sel[np.lexsort((sel, part, tl, ds[sel], spec.pre[sel]))]
        assert b - a == len(order)
        np.testing.assert_array_equal(L["syn_tgt"][a:b], spec.post[order] % tile)
        assert np.array_equal(L["syn_w"][a:b].view(np.uint32), w_ref[order].view(np.uint32))
        ngroups += 1
        k = k1
    assert ngroups >= 1


def _segments(pre, post, ds):
    """Per synapse (creation order): the index of the run of ascending (pre,
    delay) order it belongs to among its target's inputs."""
    last, seg = {}, {}
    out = np.zeros(len(pre), np.int64)
    for k in range(len(pre)):
        j, key = int(post[k]), (int(pre[k]), int(ds[k]))
        if j in last and key < last[j]:
            seg[j] = seg.get(j, 0) + 1
        last[j] = key
        out[k] = seg.get(j, 0)
    return out


@pytest.mark.parametrize("which", ["standard_test", "shuffled"])
def test_layout_orders_each_target_by_creation_runs(which):
    """Networks whose targets are fed by a later population first and an
    earlier one second (test/gpu/standard_test.cpp:116-121), or in any
    creation order: the group block is sorted by (run of ascending (pre,
    delay) order in the target's creation sequence, pre, delay, tile, part,
    creation), which makes each target's stream order its creation order."""
    spec = (netspec.standard_test(depth=7, n=60, who=3) if which == "standard_test" else
            netspec.random_net(3, shuffle=True, n_lif=60, p_conn=0.25))
    net = api.Network.from_spec(spec)
    tile = 192
    L = net.layout(DT, tile)
    w_ref, ds = oracle.synapse_reset(spec, DT)
    seg = _segments(spec.pre, spec.post, ds)
    if which == "standard_test":
        assert seg.max() == 1
    else:
        assert seg.max() >= 3
    ts = L["tile_syn"]
    k = 0
    while k < L["ntiles"]:
        k1 = k
        while k1 < L["ntiles"] and ts[k1] == ts[k]:
            k1 += 1
        a, b = ts[k], (ts[k1] if k1 < L["ntiles"] else L["s"])
        sel = np.nonzero((spec.post >= k * tile) & (spec.post < k1 * tile))[0]
        tl = spec.post[sel] // tile
        part = (spec.post[sel] - tl * tile) // (tile // 3)
        # the ranges of one segment (ascending pres) come before those of the
        # next
        order = sel[np.lexsort((sel, part, tl, ds[sel], spec.pre[sel], seg[sel]))]
        assert b - a == len(order)
        np.testing.assert_array_equal(L["syn_tgt"][a:b], spec.post[order] % tile)
        assert np.array_equal(L["syn_w"][a:b].view(np.uint32), w_ref[order].view(np.uint32))
        # each target's inputs, in stream order, are its creation order
        for j in np.unique(spec.post[sel]):
            mine = order[spec.post[order] == j]
            assert np.all(np.diff(mine) > 0)
        k = k1


def test_layout_single_tile_matches_reference_csr():
    """test/lif_unit1_test.py:105-106: p_dst = [10..19] x 10 + [20..29] x 10."""
    net = api.Network.from_spec(netspec.lif_unit(1))
    L = net.layout(DT, 192)
    assert L["ntiles"] == 1 and L["dmax"] == 1
    np.testing.assert_array_equal(L["syn_tgt"], list(range(10, 20)) * 10 + list(range(20, 30)) * 10)
    np.testing.assert_array_equal(L["syn_w"], np.float32([1e-5] * 100 + [2e-5] * 100))


def _two_neuron(delay_steps, w=0.5, tau=None, dexact=True):
    b = netspec.Builder()
    b.population(np.tile(netspec.lif_row(0, 0, 0, 0.02, 0.02, 0, 5e-3, 5e-3, 1.0, 0.5), (2, 1)))
    d = (np.float32(delay_steps) * np.float32(DT)) if dexact else np.float32(delay_steps * DT * 1.3)
    b.connect([0], [1], np.float32([w]), np.float32([d]), tau=None if tau is None else [tau])
    return b.build()


@pytest.mark.parametrize("steps,msg", [(0, "outside 1..63"), (64, "outside 1..63")])
def test_rejects_unsupported_delay(steps, msg):
    net = api.Network.from_spec(_two_neuron(steps))
    with pytest.raises(api.BsimError, match=msg):
        net.layout(DT, 192)


def test_rejects_nonintegral_delay_without_tau():
    spec = _two_neuron(2, dexact=False)
    spec.tau = np.zeros(1, np.float32)
    spec.lif[:, 10:12] = 0.0   # composite taus zero: the reference asserts
    net = api.Network.from_spec(spec)
    with pytest.raises(api.BsimError, match="not a whole number of steps"):
        net.layout(DT, 192)


def test_nonintegral_delay_rescales_weight_like_reference():
    spec = _two_neuron(2, dexact=False)
    net = api.Network.from_spec(spec)
    L = net.layout(DT, 192)
    w, ds = oracle.synapse_reset(spec, DT)
    assert L["syn_w"][0].view(np.uint32) == w[0].view(np.uint32)
    assert L["dmax"] == ds[0]


def test_rejects_more_creation_runs_than_the_layout_keeps():
    """A target whose 40 inputs were created in descending pre order has 40
    runs of ascending (pre, delay) order; the layout keeps 32 per target."""
    b = netspec.Builder()
    b.population(np.tile(netspec.lif_row(0, 0, 0, 0.02, 0.02, 0, 5e-3, 5e-3, 1.0, 0.5), (41, 1)))
    pre = np.arange(39, -1, -1)
    b.connect(pre, np.full(40, 40), np.full(40, 0.1, np.float32), np.full(40, DT, np.float32))
    net = api.Network.from_spec(b.build())
    with pytest.raises(api.BsimError, match="at most 32 such runs per target"):
        net.layout(DT, 192)
    # 32 runs are accepted
    b2 = netspec.Builder()
    b2.population(np.tile(netspec.lif_row(0, 0, 0, 0.02, 0.02, 0, 5e-3, 5e-3, 1.0, 0.5), (41, 1)))
    b2.connect(pre[8:], np.full(32, 40), np.full(32, 0.1, np.float32), np.full(32, DT, np.float32))
    api.Network.from_spec(b2.build()).layout(DT, 192)


def test_bad_arguments():
    net = api.Network()
    with pytest.raises(api.BsimError):
        net.add_array(np.array([1, 2], np.int64), np.array([0], np.int32))
    with pytest.raises(api.BsimError, match="multiple of 192"):
        api.Network.from_spec(netspec.lif_unit(1)).layout(DT, 100)
    with pytest.raises(api.BsimError, match="outside the network"):
        n2 = api.Network.from_spec(netspec.lif_unit(1))
        n2.connect([0], [999], [0.1], [1e-4])
        n2.layout(DT, 192)


def test_run_cycles_rounds_halves_away_from_zero():
    """Simulation.run's cycle count is C's round() of the float quotient
    (SingleGPUSimulator.cu:34), checked against libm's roundf itself on
    quotients at and around x.5 step boundaries."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.roundf.restype = ctypes.c_float
    libm.roundf.argtypes = [ctypes.c_float]
    dt = np.float32(1e-4)
    rng = np.random.default_rng(5)
    ks = np.concatenate([np.arange(0, 2000), rng.integers(0, 200000, 4000)])
    Ts = []
    for k in ks:
        q = np.float32(k + 0.5)
        t = q * dt
        Ts += [t, np.nextafter(t, np.float32(0)), np.nextafter(t, np.float32(1))]
    Ts += [np.float32(0.45), np.float32(0.1), np.float32(1.0), np.float32(0.0)]
    mism = 0
    for t in Ts:
        q = np.float32(np.float32(t) / dt)
        want = int(libm.roundf(float(q)))
        mism += api.run_cycles(t, dt) != want
    assert mism == 0
    # the case numpy's round-half-to-even gets wrong: 4.5 steps -> 5 cycles
    assert api.run_cycles(np.float32(4.5) * dt / np.float32(1.0), dt) == \
        int(libm.roundf(float(np.float32(np.float32(4.5) * dt) / dt)))


def test_device_cycle_quotient_is_exact():
    """The step kernel reads the cycle as count / (nslots / 8) from a residue's
    completion counter without an integer division: truncate(count * inv +
    inv / 2) with inv = 8 / nslots in float64, as the library computes it on
    the host (csrc/capi.cpp) and the kernel on the device (csrc/step.hip,
    compiled without contraction, so the product and the sum round
    separately, as numpy's do).  Exact for every count of the form t * per +
    r, 0 <= r < per, over the grids and cycles the kernel allows (per up to
    2^18 - 1 workgroups per residue, t up to 2^31 - 1), checked at both ends
    of each interval and inside it."""
    rng = np.random.default_rng(11)
    pers = [1, 2, 3, 7, 8, 171, 438, 439, 1000, 4096, 12711, 65537, (1 << 17) + 3, (1 << 18) - 1]
    ts = np.concatenate([np.arange(0, 64), [2**31 - 1, 2**31 - 2, 2**30, 2**24 + 1],
                         rng.integers(0, 2**31 - 1, 2000)]).astype(np.int64)
    for per in pers:
        nslots = 8 * per
        inv = np.float64(8.0) / np.float64(nslots)
        half = np.float64(0.5) * inv
        rs = np.unique(np.array([0, 1, per // 2, max(per - 2, 0), per - 1], dtype=np.int64))
        rs = rs[rs < per]
        c = (ts[:, None] * per + rs[None, :]).ravel()
        assert c.max() < 2**53
        got = np.trunc(c.astype(np.float64) * inv + half).astype(np.int64)
        want = np.repeat(ts, len(rs))
        assert np.array_equal(got, want), per


def test_host_layout_check_under_sanitizers():
    """tests/cpp/host_layout_check.cpp, built from the library's host sources
    under AddressSanitizer and UBSan: networks laid out in parts and joined
    equal the same network laid out at once, array for array, and arbitrary
    creation orders lay out without a memory error."""
    import shutil
    import subprocess
    cpp = os.path.join(ROOT, "tests", "cpp")
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    r = subprocess.run(["make", "-s", "-C", cpp, "host_layout_check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([os.path.join(cpp, "host_layout_check")], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0 and "host_layout_check ok" in r.stdout, r.stdout + r.stderr
