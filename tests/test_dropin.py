"""The reference's C++ user API (include/BSim.h) over the C ABI.

A BSim-style program (tests/cpp/dropin_run.cpp: createPopulation, addNeuron,
connect with weight/delay arrays, SGSim(dt).run(T, FireInfo)) is compiled
against the header; on the GPU it runs the C0 network and its output files
(GSim.log raster, g_v.data trace, FireInfo "count") are compared with the
golden vectors recorded from the reference's own CPU simulator.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import golden_io
import oracle
from bsim_amd import netspec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_run.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_run")


def test_header_compiles_standalone(tmp_path):
    """include/BSim.h is self-contained C++11 and the program builds warning-free."""
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_stsim_is_a_compile_error(tmp_path):
    """STSim (the reference's CPU simulator) is not provided: a program that
    constructs one fails to compile with a message naming SGSim, while a
    program that only names the type still compiles."""
    bad = tmp_path / "st.cpp"
    bad.write_text("""#include "BSim.h"
int main() { Network net; STSim st(&net, 1e-4f); return 0; }
""")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(bad)], capture_output=True, text=True)
    assert r.returncode != 0
    assert "use SGSim" in r.stderr
    ok = tmp_path / "ok.cpp"
    ok.write_text("""#include "BSim.h"
int main() { Network net; SGSim sg(&net, 1e-4f); STSim *unused = 0; (void)unused; return 0; }
""")
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(ok)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_poisson_population_compiles(tmp_path):
    """Poisson_spikes populations (PoissonNeuron(rate, refract, startTime,
    duration), src/neuron/poisson/PoissonNeuron.cpp:13) build through the
    header, seeded or not."""
    src = tmp_path / "pois.cpp"
    src.write_text("""#include "BSim.h"
int main()
{
    Network net;
    Population<Poisson_spikes> *drive = net.createPopulation(
        100, Poisson_spikes(PoissonNeuron(40.0f, 0.0f, 0.0f, 0.5f), 1e-3f, 1e-3f));
    Poisson_spikes seeded(PoissonNeuron(5.0f, 0.0f, 0.01f, 0.1f), 1e-3f, 1e-3f);
    seeded.setSeed(7);
    Population<Poisson_spikes> *more = net.createPopulation(10, seeded);
    more->addNeuron(seeded);
    Population<LIF_brian> *p = net.createPopulation(
        50, LIF_brian(LIFENeuron(0, 0, 0, 1e-1f, 50e-3f, 0, 1e-3f, 1e-3f, 15e-3f, 0), 1e-3f, 1e-3f));
    net.connect(drive, p, 0.1f, 1e-4f, Excitatory);
    return drive->getNum() == 100 ? 0 : 1;
}
""")
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def write_input(path, spec, pops, projections, dt, T, monitor):
    """pops: [(first, n)], projections: [(src_pop, dst_pop, w, d)] all-to-all."""
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(pops)))
        for first, n in pops:
            row = np.asarray(spec.lif[first], np.float32)
            f.write(struct.pack("<i", n))
            f.write(row.tobytes())
            f.write(np.asarray(spec.lif[first:first + n, 9], np.float32).tobytes())
        f.write(struct.pack("<i", len(projections)))
        for s, d, w, dl in projections:
            f.write(struct.pack("<ii", s, d))
            f.write(np.asarray(w, np.float32).tobytes())
            f.write(np.asarray(dl, np.float32).tobytes())
        f.write(struct.pack("<ff", dt, T))
        f.write(struct.pack("<i", len(monitor)))
        f.write(np.asarray(monitor, np.int32).tobytes())


@pytest.mark.gpu
def test_dropin_program_matches_reference_c0(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_run is not built (run __graft_entry__.build())")
    meta, spec, dtype, gold = golden_io.load("c0_f32")
    n = 1000
    # the all-to-all projection p0 -> p1 in creation order k = iSrc * n + iDst
    k = np.arange(n * n)
    assert np.array_equal(spec.pre, k // n) and np.array_equal(spec.post, n + k % n)
    steps = len(gold["raster_count"])
    dt = np.float32(netspec.DT)
    T = np.float32(steps * float(dt))
    assert int(np.round(np.float32(T / dt))) == steps
    rec = list(meta["rec_ids"])
    inp = tmp_path / "c0.bin"
    write_input(inp, spec, [(0, n), (n, n)], [(0, 1, spec.w, spec.d)], dt, T, rec)
    r = subprocess.run([BIN, str(inp), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = (tmp_path / "GSim.log").read_text().split("\n")[:steps]
    counts = np.array([len(l.split()) for l in lines])
    ids = np.array([int(x) for l in lines for x in l.split()], np.int32)
    np.testing.assert_array_equal(counts, gold["raster_count"])
    np.testing.assert_array_equal(ids, gold["raster_ids"])
    fc = np.array((tmp_path / "GFire.count").read_text().split(), np.int32)
    np.testing.assert_array_equal(fc, gold["fire_count"])
    gv = np.loadtxt(tmp_path / "g_v.data", ndmin=2)
    # g_v.data prints 10 decimals, as the reference's does
    np.testing.assert_allclose(gv, gold["rec_v"].astype(np.float64), rtol=0, atol=5.1e-11)
    info = (tmp_path / "GSim.info").read_text()
    assert "NeuronNum: 2000, SynapseNum: 1000000" in info
    # FireInfo "Y" and "X" against the oracle (bit for bit)
    o = oracle.run(spec, netspec.DT, steps)
    y = np.array([float.fromhex(v) for v in (tmp_path / "input_Y.txt").read_text().split()],
                 np.float32)
    x = np.array((tmp_path / "input_X.txt").read_text().split(), np.int32)
    assert np.array_equal(y.view(np.uint32), o.input_sum.view(np.uint32))
    np.testing.assert_array_equal(x, o.slot_count)
    assert x.sum() > 0 and np.count_nonzero(y) > 0
    # setLogData(true): the reference's -DLOG_DATA files for every LIF neuron
    # (src/sim/SingleGPUSimulator.cu:150-232), against the reference's own
    # values for the traced neurons and the oracle's for all of them
    logd = tmp_path / "logdata"
    logd.mkdir()
    r = subprocess.run([BIN, str(inp), str(logd), "logdata"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = (logd / "GSim.log").read_text().split("\n")[:steps]
    ids2 = np.array([int(x) for l in lines for x in l.split()], np.int32)
    np.testing.assert_array_equal(ids2, gold["raster_ids"])
    allrec = list(range(2 * n))
    o2 = oracle.run(spec, netspec.DT, steps, rec_ids=allrec)
    for fname, gkey, okey in (("g_v.data", "rec_v", "rec_v"), ("g_ie.data", "rec_ie", "rec_ie"),
                              ("g_ii.data", "rec_ii", "rec_ii"),
                              ("g_input_e.data", "rec_in_e", "rec_in_e"),
                              ("g_input_i.data", "rec_in_i", "rec_in_i")):
        a = np.loadtxt(logd / fname, ndmin=2)
        assert a.shape == (steps, 2 * n), (fname, a.shape)
        np.testing.assert_allclose(a, getattr(o2, okey).astype(np.float64), rtol=0, atol=5.1e-11,
                                   err_msg=fname)
        np.testing.assert_allclose(a[:, rec], gold[gkey].astype(np.float64), rtol=0,
                                   atol=5.1e-11, err_msg=fname)
    # run(T) writes GFire.count by itself
    plain = tmp_path / "plain"
    plain.mkdir()
    r = subprocess.run([BIN, str(inp), str(plain), "plain"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    fc2 = np.array((plain / "GFire.count").read_text().split(), np.int32)
    np.testing.assert_array_equal(fc2, gold["fire_count"])


@pytest.mark.gpu
def test_dropin_poisson_drive_matches_oracle(tmp_path):
    """A BSim program with a Poisson_spikes drive population (seeded through
    PoissonNeuron::setSeed) driving two LIF populations, run through
    include/BSim.h: its raster and GFire.count equal the oracle's."""
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_run is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(21)
    n0, n1, m = 300, 200, 120
    rate, start, dur, seed = 20.0, 0.002, 0.03, 9
    lif = netspec.lif_row(0.0, 0.0, 0.0, 0.02, 0.02, 2e-4, 5e-3, 8e-3, 15.0, 0.0)
    b = netspec.Builder()
    for cnt in (n0, n1):
        rows = np.tile(lif, (cnt, 1))
        rows[:, 9] = rng.uniform(10.0, 14.0, cnt).astype(np.float32)
        b.population(rows)
    prow = np.zeros(netspec.LIF_NPARAM)
    prow[10] = prow[11] = 1e-3                     # the composite's default taus
    b.population(np.tile(prow, (m, 1)), kind=netspec.POISSON, pois=(rate, start, dur))
    w01 = rng.uniform(-0.3, 0.6, n0 * n1).astype(np.float32)
    d01 = (rng.integers(1, 4, n0 * n1) * np.float32(netspec.DT)).astype(np.float32)
    wp = rng.uniform(0.0, 2.0, m * n0).astype(np.float32)
    dp = (rng.integers(1, 3, m * n0) * np.float32(netspec.DT)).astype(np.float32)
    k = np.arange(n0 * n1)
    b.connect(k // n1, n0 + k % n1, w01, d01, np.zeros(n0 * n1, np.int8))   # type NULL
    k = np.arange(m * n0)
    b.connect(n0 + n1 + k // n0, k % n0, wp, dp, np.zeros(m * n0, np.int8))
    spec = b.build()
    spec.seed = seed
    steps = 400
    dt = np.float32(netspec.DT)
    T = np.float32(steps * float(dt))
    assert int(np.round(np.float32(T / dt))) == steps
    inp = tmp_path / "pois.bin"
    write_input(inp, spec, [(0, n0), (n0, n1)], [(0, 1, w01, d01)], dt, T, [])
    with open(inp, "ab") as f:
        f.write(struct.pack("<i", 1))
        f.write(struct.pack("<i", m))
        f.write(np.array([rate, start, dur], np.float32).tobytes())
        f.write(struct.pack("<Q", seed))
        f.write(struct.pack("<i", 0))
        f.write(wp.tobytes())
        f.write(dp.tobytes())
    r = subprocess.run([BIN, str(inp), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    o = oracle.run(spec, netspec.DT, steps)
    lines = (tmp_path / "GSim.log").read_text().split("\n")[:steps]
    counts = np.array([len(l.split()) for l in lines])
    ids = np.array([int(x) for l in lines for x in l.split()], np.int32)
    np.testing.assert_array_equal(counts, o.raster_count)
    np.testing.assert_array_equal(ids, o.raster_ids)
    fc = np.array((tmp_path / "GFire.count").read_text().split(), np.int32)
    np.testing.assert_array_equal(fc, o.fire_count)
    assert fc[n0 + n1:].sum() > 0 and fc[:n0].sum() > 0


STD_BIN = os.path.join(ROOT, "tests", "cpp", "standard_run")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["standard_test_f32", "standard_test_hi_f32"])
def test_standard_topology_program_matches_reference(tmp_path, name):
    """tests/cpp/standard_run.cpp builds the reference's benchmark topology
    (test/gpu/standard_test.cpp:55-138: g[i-1] -> g[i] connected before
    g[1] -> g[i], inhibitory side populations) through include/BSim.h;
    on the GPU its GSim.log raster, GFire.count and g_v.data equal the
    golden vectors recorded from the reference's own CPU simulator."""
    if not os.path.exists(STD_BIN):
        pytest.fail("tests/cpp/standard_run is not built (run __graft_entry__.build())")
    meta, spec, dtype, gold = golden_io.load(name)
    kw = meta["kwargs"]
    steps = meta["steps"]
    T = np.float32(steps * float(np.float32(netspec.DT)))
    rec = [str(i) for i in meta["rec_ids"]]
    args = [STD_BIN, str(kw["depth"]), str(kw["n"]), *[repr(float(x)) for x in kw["w"]],
            str(kw["who"]), repr(float(T)), str(tmp_path), *rec]
    print("[gpu] standard_run", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = (tmp_path / "GSim.log").read_text().split("\n")[:steps]
    counts = np.array([len(l.split()) for l in lines])
    ids = np.array([int(x) for l in lines for x in l.split()], np.int32)
    np.testing.assert_array_equal(counts, gold["raster_count"])
    np.testing.assert_array_equal(ids, gold["raster_ids"])
    fc = np.array((tmp_path / "GFire.count").read_text().split(), np.int32)
    np.testing.assert_array_equal(fc, gold["fire_count"])
    gv = np.loadtxt(tmp_path / "g_v.data", ndmin=2)
    np.testing.assert_allclose(gv, gold["rec_v"].astype(np.float64), rtol=0, atol=5.1e-11)
    assert len(ids) > 1000


def test_truncated_input_fails_cleanly(tmp_path):
    """A truncated network file makes the program stop with a message and a
    non-zero status before any device work (the reference prints and
    returns -1 on file errors)."""
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/dropin_run is not built")
    meta, spec, dtype, gold = golden_io.load("c0_f32")
    n = 1000
    inp = tmp_path / "c0.bin"
    write_input(inp, spec, [(0, n), (n, n)], [(0, 1, spec.w, spec.d)], np.float32(1e-4),
                np.float32(0.01), [0])
    data = inp.read_bytes()
    for cut in (3, len(data) // 2, len(data) - 5):
        bad = tmp_path / f"cut{cut}.bin"
        bad.write_bytes(data[:cut])
        r = subprocess.run([BIN, str(bad), str(tmp_path)], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "short read" in r.stderr, (cut, r.returncode, r.stderr)
    r = subprocess.run([BIN, str(tmp_path / "missing.bin"), str(tmp_path)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
