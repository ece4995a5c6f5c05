"""File exchange with oracle/_ref/ref_stsim_{f32,f64} (TEST INFRASTRUCTURE).

``write_input`` serialises a network (NetSpec layout) for the reference
harness; ``run_ref`` runs the harness built by oracle/build_ref.sh and parses
its output.  Bernoulli sources are handed to the reference as Array_spikes
neurons whose cycle lists are the cycles the source hash selects (both fire at
most once per cycle, src/neuron/array/ArrayNeuron.cpp:67-80).
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


def ref_binary(dtype=np.float32):
    name = "ref_stsim_f32" if np.dtype(dtype) == np.float32 else "ref_stsim_f64"
    path = os.path.join(REF_DIR, name)
    return path if os.path.exists(path) else None


def _bernoulli_as_arrays(net, steps):
    """kind/arr_off/arr_t with every Bernoulli source replaced by the array
    source that fires at the same cycles."""
    import oracle  # noqa: WPS433 (same directory)
    kind = np.asarray(net.kind, np.int8).copy()
    n = len(kind)
    lists = []
    for i in range(n):
        if kind[i] == 1:
            lists.append(np.asarray(net.arr_t[net.arr_off[i]:net.arr_off[i + 1]], np.int32))
        elif kind[i] == 2:
            h = np.array([oracle.lib().bso_hash(int(net.seed), i, t) for t in range(steps)],
                         np.uint64)
            lists.append(np.nonzero(h < int(net.bern_p[i]))[0].astype(np.int32))
            kind[i] = 1
        else:
            lists.append(np.zeros(0, np.int32))
    lens = np.array([len(x) for x in lists], np.int64)
    arr_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    arr_t = np.concatenate(lists).astype(np.int32) if lens.sum() else np.zeros(0, np.int32)
    return kind, arr_off, arr_t


def write_input(path, net, dt, steps, rec_ids, dtype=np.float32):
    dtype = np.dtype(dtype)
    kind, arr_off, arr_t = _bernoulli_as_arrays(net, steps)
    n, s = len(kind), len(net.pre)
    rec = np.asarray(rec_ids, np.int32)
    tau = np.zeros(s) if net.tau is None else np.asarray(net.tau)
    stype = np.zeros(s, np.int8) if net.stype is None else np.asarray(net.stype, np.int8)
    with open(path, "wb") as f:
        f.write(b"BSNT")
        f.write(np.array([1, dtype.itemsize], np.uint32).tobytes())
        f.write(np.array([n], np.int32).tobytes())
        f.write(np.array([s], np.int64).tobytes())
        f.write(np.array([steps], np.int32).tobytes())
        f.write(np.array([dt], dtype).tobytes())
        f.write(np.array([len(rec)], np.int32).tobytes())
        f.write(rec.tobytes())
        f.write(kind.tobytes())
        f.write(np.asarray(net.lif, np.float64).astype(dtype).tobytes())
        f.write(arr_off.tobytes())
        f.write(np.array([len(arr_t)], np.int64).tobytes())
        f.write(arr_t.tobytes())
        f.write(np.asarray(net.pre, np.int32).tobytes())
        f.write(np.asarray(net.post, np.int32).tobytes())
        f.write(np.asarray(net.w).astype(dtype).tobytes())
        f.write(np.asarray(net.d).astype(dtype).tobytes())
        f.write(stype.tobytes())
        f.write(tau.astype(dtype).tobytes())


def read_output(path, n, s, steps, n_rec, dtype=np.float32):
    dtype = np.dtype(dtype)
    buf = open(path, "rb").read()
    o = 0

    def take(dt, cnt):
        nonlocal o
        a = np.frombuffer(buf, dt, cnt, o)
        o += a.nbytes
        return a.copy()

    res = {}
    res["raster_count"] = take(np.int32, steps)
    ln = int(take(np.int64, 1)[0])
    res["raster_ids"] = take(np.int32, ln)
    # V, the E / I input the update consumed, i_E, i_I (LIFENeuron.cpp:144-148)
    for key in ("rec_v", "rec_in_e", "rec_in_i", "rec_ie", "rec_ii"):
        res[key] = take(dtype, steps * n_rec).reshape(steps, n_rec)
    rec = np.dtype([("c", dtype, 9), ("rt", np.int32)])
    cr = take(rec, n)
    res["consts"] = cr["c"]
    res["refrac_time"] = cr["rt"]
    sr = take(np.dtype([("w", dtype), ("ds", np.int32)]), s)
    res["syn_w"] = sr["w"]
    res["syn_ds"] = sr["ds"]
    res["fire_count"] = take(np.int32, n)
    assert o == len(buf), (o, len(buf))
    return res


def run_ref(net, dt, steps, rec_ids=(), dtype=np.float32):
    exe = ref_binary(dtype)
    if exe is None:
        raise FileNotFoundError("oracle/_ref not built (oracle/build_ref.sh)")
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bsnt")
        fout = os.path.join(d, "out.bsout")
        write_input(fin, net, dt, steps, rec_ids, dtype)
        subprocess.run([exe, fin, fout], check=True, cwd=d, stdout=subprocess.DEVNULL)
        return read_output(fout, len(net.kind), len(net.pre), steps, len(rec_ids), dtype)


def time_ref(net, dt, steps, dtype=np.float32):
    """Wall seconds of the reference STSim::run over ``steps`` cycles."""
    import json
    exe = ref_binary(dtype)
    if exe is None:
        raise FileNotFoundError("oracle/_ref not built (oracle/build_ref.sh)")
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bsnt")
        write_input(fin, net, dt, steps, (), dtype)
        p = subprocess.run([exe, "--time", fin], check=True, cwd=d, capture_output=True,
                           text=True)
        line = [ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1]
        return json.loads(line)["seconds"]
