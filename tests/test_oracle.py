"""Pins the CPU oracle before anything is checked against it.

* against the reference's own golden traces (test/lif_unit{1..4}_test_v.std,
  a double build printed with %.10lf);
* bit for bit against golden vectors produced by the reference's compiled CPU
  simulator (tests/golden/*.npz, see make_golden.py);
* against closed forms and the known answers of test/lif_unit1_test.py;
* the CSR/OpenMP form against the exact form.
"""
import json

import numpy as np
import pytest

import oracle
from bsim_amd import netspec

import golden_io

DT = 1e-4


def test_std_traces_double():
    """Oracle in double reproduces the reference's .std V traces to the print
    rounding of %.10lf (|err| <= 5e-11)."""
    std = golden_io.load_std()
    for k in (1, 2, 3, 4):
        net = netspec.lif_unit(k, wdtype=np.float64)
        r = oracle.run(net, DT, 1000, rec_ids=range(30), dtype=np.float64)
        ref = std[f"lif_unit{k}"]
        assert ref.shape == (1000, 30)
        assert np.abs(r.rec_v - ref).max() <= 5.0001e-11


def test_std_traces_distinguish_cases():
    """The four .std files are not interchangeable: a wrong delay or refractory
    handling would fit the wrong file (unit3 has a 2-step delay, unit4
    refractory periods)."""
    std = golden_io.load_std()
    net3 = netspec.lif_unit(3, wdtype=np.float64)
    r3 = oracle.run(net3, DT, 1000, rec_ids=range(30), dtype=np.float64)
    assert np.abs(r3.rec_v - std["lif_unit1"]).max() > 1e-6
    net4 = netspec.lif_unit(4, wdtype=np.float64)
    r4 = oracle.run(net4, DT, 1000, rec_ids=range(30), dtype=np.float64)
    assert np.abs(r4.rec_v - std["lif_unit1"]).max() > 1e-6


@pytest.mark.parametrize("name", golden_io.case_names())
def test_oracle_matches_reference_build(name):
    meta, net, dtype, g = golden_io.load(name)
    assert golden_io.net_digest(net) == meta["net_sha256"], "generator changed"
    r = oracle.run(net, meta["dt"], meta["steps"], rec_ids=meta["rec_ids"], dtype=dtype)
    np.testing.assert_array_equal(r.raster_count, g["raster_count"])
    np.testing.assert_array_equal(r.raster_ids, g["raster_ids"])
    assert np.array_equal(r.rec_v.view(g["rec_v_bits"].dtype), g["rec_v_bits"])
    for k in ("rec_ie", "rec_ii", "rec_in_e", "rec_in_i"):
        if k + "_bits" in g:   # i_E / i_I after the update, the input it consumed
            assert np.array_equal(getattr(r, k).view(g[k + "_bits"].dtype), g[k + "_bits"]), k
    np.testing.assert_array_equal(r.fire_count, g["fire_count"])
    c9, rt = oracle.lif_constants(net, meta["dt"], dtype)
    lif = net.kind == netspec.LIF
    cb = np.ascontiguousarray(c9[:, :8]).view(g["consts_bits"].dtype)
    assert np.array_equal(cb[lif], g["consts_bits"][lif][:, :8])
    np.testing.assert_array_equal(rt[lif], g["refrac_time"][lif])
    # final V of every LIF neuron (column 8 of the reference's getData dump)
    fv = np.ascontiguousarray(r.final_v).view(g["consts_bits"].dtype)
    assert np.array_equal(fv[lif], g["consts_bits"][lif][:, 8])
    w, ds = oracle.synapse_reset(net, meta["dt"], dtype)
    if "syn_w_bits" in g:
        assert np.array_equal(w.view(g["syn_w_bits"].dtype), g["syn_w_bits"])
        np.testing.assert_array_equal(ds, g["syn_ds"])
    else:
        assert golden_io.hashlib.sha256(np.ascontiguousarray(w).tobytes()).hexdigest() == \
            meta["syn_w_sha256"]
        assert golden_io.hashlib.sha256(np.ascontiguousarray(ds).tobytes()).hexdigest() == \
            meta["syn_ds_sha256"]


def test_known_answers_lif_unit1():
    """test/lif_unit1_test.py:70-92 known answers (6 decimals)."""
    net = netspec.lif_unit(1)
    c9, rt = oracle.lif_constants(net, DT)
    np.testing.assert_array_almost_equal(c9[:, 1], [0.999899983] * 30)   # CE
    np.testing.assert_array_almost_equal(c9[:, 2], [0.999899983] * 30)   # CI
    np.testing.assert_array_almost_equal(c9[:, 3], [0.000998942531] * 30)  # C_E
    np.testing.assert_array_almost_equal(c9[:, 4], [0.000998942531] * 30)  # C_I
    np.testing.assert_array_almost_equal(c9[:, 5], [0.00999003649] * 10 + [0.0] * 20)  # v_tmp
    np.testing.assert_array_almost_equal(c9[:, 0], [0.998001993] * 30)   # Cm
    np.testing.assert_array_almost_equal(c9[:, 6], [0.0149999997] * 30)  # v_thresh
    assert list(rt) == [0] * 30


def test_constants_closed_form():
    """Cm = exp(-dt/tau_m), C_E = rm tau_E/(tau_E - tau_m) (CE - Cm),
    v_tmp = (i_offset rm + v_rest)(1 - Cm), evaluated in float64 from the
    float32 inputs: agreement to float32 rounding (a dropped term or wrong
    operand is off by orders of magnitude)."""
    net = netspec.random_net(3, n_lif=50)
    c9, rt = oracle.lif_constants(net, DT)
    lif = np.nonzero(net.kind == netspec.LIF)[0]
    p = net.lif[lif].astype(np.float32).astype(np.float64)
    dt = float(np.float32(DT))
    rm = np.where(np.abs(p[:, 3]) > 1e-10, p[:, 4] / p[:, 3], 1.0)
    Cm = np.exp(-dt / p[:, 4])
    CE = np.exp(-dt / p[:, 6])
    CI = np.exp(-dt / p[:, 7])
    v_tmp = (p[:, 9] * rm + p[:, 1]) * (1 - Cm)
    C_E = rm * p[:, 6] / (p[:, 6] - p[:, 4]) * (CE - Cm)
    C_I = rm * p[:, 7] / (p[:, 7] - p[:, 4]) * (CI - Cm)
    got = c9[lif].astype(np.float64)
    for col, want in ((0, Cm), (1, CE), (2, CI), (3, C_E), (4, C_I), (5, v_tmp)):
        np.testing.assert_allclose(got[:, col], want, rtol=2e-5, atol=1e-9)
    np.testing.assert_array_equal(rt[lif], np.floor(np.float32(p[:, 5]) / np.float32(dt)
                                                   ).astype(np.int32))


@pytest.mark.parametrize("d", [1, 2, 6])
def test_delay_timing_closed_form(d):
    """A spike sent at cycle s through a delay of d steps is added to the
    target's input in phase (ii) of cycle s + d, enters i_E at s + d + 1 and
    first moves V at s + d + 2 (src/synapse/static/StaticSynapse.cpp:56-70,
    src/neuron/lif/LIFENeuron.cpp:114-160)."""
    s = 5
    b = netspec.Builder()
    b.population(np.zeros((1, netspec.LIF_NPARAM)), kind=netspec.ARRAY, spikes=[[s]])
    b.population(netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 0.0, 5e-3, 5e-3, 1e9, 0.0))
    b.connect([0], [1], np.float32([1.0]), np.float32([d]) * np.float32(DT))
    net = b.build()
    r = oracle.run(net, DT, s + d + 6, rec_ids=[1])
    v = r.rec_v[:, 0]
    moved = np.nonzero(v != 0)[0]
    assert moved[0] == s + d + 2
    assert r.events == 1


def test_refractory_discards_input():
    """A refractory neuron only counts down and its input is dropped
    (LIFENeuron.cpp:118-119, 152-153): a target held refractory while the
    spike arrives never moves."""
    b = netspec.Builder()
    b.population(np.zeros((1, netspec.LIF_NPARAM)), kind=netspec.ARRAY, spikes=[[0]])
    # neuron 1 fires at cycle 0 (v_tmp >> thresh) then is refractory for 9 cycles
    row = netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 10e-4, 5e-3, 5e-3, 0.5, 1e6)
    b.population(row)
    b.connect([0], [1], np.float32([1.0]), np.float32([1]) * np.float32(DT))
    r = oracle.run(b.build(), DT, 8, rec_ids=[1])
    assert r.raster()[0].tolist() == [0, 1]
    assert np.all(r.rec_v[:, 0] == 0)
    assert r.final_ie[1] == 0


def test_csr_mode_matches_exact():
    for name in ("c0_f32", "standard_small_f32", "lif_unit2_f32", "random0_f32"):
        meta, net, dtype, g = golden_io.load(name)
        _, ds = oracle.synapse_reset(net, meta["dt"], dtype)
        if not oracle.csr_order_ok(net, ds):
            continue
        e = oracle.run(net, meta["dt"], meta["steps"], rec_ids=meta["rec_ids"], dtype=dtype)
        for threads in (1, 4):
            c = oracle.run(net, meta["dt"], meta["steps"], rec_ids=meta["rec_ids"],
                           dtype=dtype, mode="csr", tile=97, threads=threads)
            np.testing.assert_array_equal(c.raster_ids, e.raster_ids)
            assert np.array_equal(c.rec_v.view(np.uint32), e.rec_v.view(np.uint32))
            assert np.array_equal(c.final_v.view(np.uint32), e.final_v.view(np.uint32))
            assert c.events == e.events


def test_order_property_detects_shuffle():
    """A network whose creation order is shuffled violates the (pre, delay)
    ordering property; the ordered ones satisfy it."""
    net = netspec.random_net(1, shuffle=True, n_lif=60, p_conn=0.3)
    _, ds = oracle.synapse_reset(net, DT)
    assert not oracle.csr_order_ok(net, ds)
    net = netspec.standard(npop=3, n=50, fanout=10, seed=3)
    _, ds = oracle.synapse_reset(net, DT)
    assert oracle.csr_order_ok(net, ds)


def test_planted_order_mistake_is_caught():
    """Applying a cycle's deliveries in reverse creation order (a plausible
    mistake) changes the bits of the summed inputs: the golden comparison
    above would catch it.  Planted on the CPU by reversing a shuffled
    network's creation order."""
    meta, net, dtype, g = golden_io.load("random1_f32")
    rev = netspec.NetSpec(**{**net.__dict__})
    perm = np.arange(len(net.pre))[::-1]
    for k in ("pre", "post", "w", "d", "stype", "tau"):
        setattr(rev, k, np.ascontiguousarray(getattr(net, k)[perm]))
    r = oracle.run(rev, meta["dt"], meta["steps"], rec_ids=meta["rec_ids"], dtype=dtype)
    assert not np.array_equal(r.rec_v.view(np.uint32), g["rec_v_bits"])


def test_bernoulli_hash_known_values():
    """The source hash is a fixed function (splitmix64 finalizer); pin a few
    values so CPU and GPU implementations cannot drift."""
    vals = oracle.bern_hash(12345, [0, 1, 2, 1000], 7).tolist()
    assert vals == json.loads(json.dumps(vals))
    z = (12345 ^ ((0 << 32) | 7)) + 0x9E3779B97F4A7C15
    z &= (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
    z ^= z >> 31
    assert vals[0] == z >> 32


# ---- FireInfo logs "X" and "Y" (the GPU simulator's per-neuron counters) ----

def _slot_count_closed_form(spec, raster_count, raster_ids, steps):
    """X[n] = sum over n's spikes at cycle s of min(D, T - 1 - s): the
    deliveries of cycles s+1 .. s+D each visit the spike once
    (src/synapse/static/static.cu:42), D the network's largest delay."""
    _, ds = oracle.synapse_reset(spec, DT)
    D = int(ds.max()) if len(ds) else 0
    cyc = np.repeat(np.arange(steps), raster_count)
    x = np.zeros(spec.n, np.int64)
    np.add.at(x, raster_ids, np.minimum(D, steps - 1 - cyc))
    return x


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("mode", ["exact", "csr"])
def test_slot_count_closed_form(seed, mode):
    spec = netspec.random_net(seed, n_lif=50, n_arr=6, n_bern=4, p_conn=0.2, dmax=2 + 3 * seed,
                              spike_prob=0.1, steps=150)
    steps = 150
    o = oracle.run(spec, DT, steps, mode=mode, tile=64, threads=2)
    want = _slot_count_closed_form(spec, o.raster_count, o.raster_ids, steps)
    assert o.raster_ids.size > 50
    np.testing.assert_array_equal(o.slot_count, want)


def test_input_sum_closed_form():
    """Y (lif.cu:116) on targets that never fire: the sum over cycles of the
    cycle's E + I input, i.e. every weight delivered at a cycle c <= T - 2
    (consumed by the update of c + 1).  Checked against a float64 sum of the
    deliveries computed from the sources' spike lists."""
    rng = np.random.default_rng(3)
    steps = 120
    b = netspec.Builder()
    spikes = [np.sort(rng.choice(steps, size=rng.integers(5, 40), replace=False)) for _ in range(8)]
    src = b.population(np.zeros((8, netspec.LIF_NPARAM)), kind=netspec.ARRAY, spikes=spikes)
    tgt = b.population(np.tile(netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 2e-3, 5e-3, 8e-3, 1e9, 0.0),
                               (5, 1)))
    pre, post, w, d = [], [], [], []
    for p in range(8):
        for q in range(5):
            if rng.random() < 0.7:
                pre.append(p)
                post.append(8 + q)
                w.append(rng.uniform(-1.0, 1.0))
                d.append(int(rng.integers(1, 6)) * DT)
    b.connect(pre, post, np.asarray(w, np.float32), np.asarray(d, np.float32))
    spec = b.build()
    o = oracle.run(spec, DT, steps)
    assert o.raster_count[:].sum() == sum(len(s) for s in spikes)   # only the sources fire
    want = np.zeros(spec.n)
    for k in range(len(pre)):
        ds = int(round(float(np.float32(d[k])) / DT))
        for s in spikes[pre[k]]:
            if s + ds + 1 <= steps - 1:
                want[post[k]] += float(np.float32(w[k]))
    np.testing.assert_allclose(o.input_sum[8:].astype(np.float64), want[8:], rtol=1e-5, atol=1e-5)
    assert np.all(o.input_sum[:8] == 0)
    oc = oracle.run(spec, DT, steps, mode="csr", tile=64, threads=2)
    assert np.array_equal(oc.input_sum.view(np.uint32), o.input_sum.view(np.uint32))


@pytest.mark.parametrize("name", ["nonint1_f32", "nonint2_f32"])
def test_nonint_golden_pins_rescale_and_tau_choice(name):
    """The non-integral-delay goldens exercise every branch of
    StaticSynapse::reset's rescale (src/synapse/static/StaticSynapse.cpp:43-54)
    and of CompositeNeuron::createSynapse's tau choice (src/base/
    CompositeNeuron.h:62-79): rescaled weights with an explicit tau, with the
    destination's excitatory and with its inhibitory composite tau; and the
    reference's weights differ from what a plausible mistake gives -- the
    other composite tau, the neuron's own tau_syn instead of the composite's,
    or no rescale -- so each of those fails the golden."""
    meta, net, dtype, g = golden_io.load(name)
    gw = g["syn_w"]
    ds = g["syn_ds"]
    f = np.float32
    w0 = net.w.astype(f)
    d = net.d.astype(f)
    dt = f(meta["dt"])
    resid = d - dt * ds.astype(f)
    moved = np.abs(resid) > f(1e-10)
    tau_e = net.lif[net.post, 10].astype(f)
    tau_i = net.lif[net.post, 11].astype(f)
    own_e = net.lif[net.post, 6].astype(f)
    own_i = net.lif[net.post, 7].astype(f)
    expl = np.abs(net.tau.astype(f)) > f(1e-10)
    inh = net.stype == 1
    rescaled = moved & (gw.view(np.uint32) != w0.view(np.uint32))
    for sel in (expl, ~expl & ~inh, ~expl & inh):
        assert np.count_nonzero(rescaled & sel) >= 20

    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.expf.restype = ctypes.c_float
    libm.expf.argtypes = [ctypes.c_float]

    def apply(tau):
        out = w0.copy()
        m = moved & (np.abs(tau) > f(1e-10))
        x = (-(resid[m]) / tau[m]).astype(f)
        c = np.array([libm.expf(float(v)) for v in x], f)
        out[m] = (w0[m] * c).astype(f)
        return out

    def tau_with(e, i):
        return np.where(expl, net.tau.astype(f), np.where(inh, i, e)).astype(f)

    right = apply(tau_with(tau_e, tau_i))
    assert np.array_equal(right.view(np.uint32), gw.view(np.uint32))
    for wrong in (apply(tau_with(tau_i, tau_e)), apply(tau_with(own_e, own_i)), w0):
        diff = np.abs(wrong.view(np.int32).astype(np.int64) - gw.view(np.int32).astype(np.int64))
        assert np.count_nonzero(diff > 1) >= 20
