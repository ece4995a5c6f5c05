"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module.  It drives ``oracle/liboracle.so`` (built from stsim.c by
oracle/Makefile), a plain-C restatement of BSim's single-threaded CPU simulator:

* LIF constants   -- src/neuron/lif/LIFENeuron.cpp:68-112
* LIF step        -- src/neuron/lif/LIFENeuron.cpp:114-160
* E/I input split -- src/neuron/lif/LIFENeuron.cpp:52-61
* synapse reset   -- src/synapse/static/StaticSynapse.cpp:43-54
* step order      -- src/net/Network.cpp:308-330
* array source    -- src/neuron/array/ArrayNeuron.cpp:67-80
* Poisson source  -- src/neuron/poisson/PoissonNeuron.cpp:36-63 (seeded draws)

A network is passed as an object with numpy attributes (``kind``, ``lif``,
``arr_off``, ``arr_t``, ``bern_p``, ``seed``, ``pre``, ``post``, ``w``, ``d``,
``stype``, ``tau``, optional ``pois``), the layout of ``bsim_amd.netspec.NetSpec``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile liboracle.so in place (gcc); returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        for suf in ("f32", "f64"):
            real = C.c_float if suf == "f32" else C.c_double
            net_t, out_t = _structs(suf)
            f = getattr(_LIB, f"bso_run_{suf}")
            f.argtypes = [C.POINTER(net_t), real, C.c_int32, C.POINTER(out_t)]
            f.restype = C.c_int
            f = getattr(_LIB, f"bso_run_csr_{suf}")
            f.argtypes = [C.POINTER(net_t), real, C.c_int32, C.c_int32, C.c_int32,
                          C.POINTER(out_t)]
            f.restype = C.c_int
            f = getattr(_LIB, f"bso_lif_constants_{suf}")
            f.argtypes = [C.POINTER(net_t), real, C.c_void_p, C.c_void_p]
            f.restype = C.c_int
            f = getattr(_LIB, f"bso_synapse_reset_{suf}")
            f.argtypes = [C.POINTER(net_t), real, C.c_void_p, C.c_void_p]
            f.restype = C.c_int
        _LIB.bso_hash.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        _LIB.bso_hash.restype = C.c_uint32
        _LIB.bso_csr_order_ok.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        _LIB.bso_csr_order_ok.restype = C.c_int
    return _LIB


_STRUCTS = {}


def _structs(suf):
    if suf in _STRUCTS:
        return _STRUCTS[suf]
    real = C.c_float if suf == "f32" else C.c_double
    P = C.c_void_p

    class Net(C.Structure):
        _fields_ = [("n", C.c_int32), ("kind", P), ("lif", P), ("arr_off", P),
                    ("arr_t", P), ("bern_p", P), ("seed", C.c_uint64), ("s", C.c_int64),
                    ("pre", P), ("post", P), ("w", P), ("d", P), ("stype", P), ("tau", P),
                    ("pois", P)]

    class Out(C.Structure):
        _fields_ = [("raster_count", P), ("raster_ids", P), ("raster_cap", C.c_int64),
                    ("raster_len", C.c_int64), ("n_rec", C.c_int32), ("rec_ids", P),
                    ("rec_v", P), ("fire_count", P), ("final_v", P), ("final_ie", P),
                    ("final_ii", P), ("final_refrac", P), ("events", C.c_int64),
                    ("time_from", C.c_int32), ("wall_s", C.c_double),
                    ("events_from", C.c_int64), ("input_sum", P), ("slot_count", P),
                    ("rec_ie", P), ("rec_ii", P), ("rec_in_e", P), ("rec_in_i", P)]

    del real
    _STRUCTS[suf] = (Net, Out)
    return Net, Out


def _ptr(a):
    return None if a is None else a.ctypes.data


class _NetArgs:
    """Holds the converted arrays alive for the duration of a call."""

    def __init__(self, net, dtype):
        self.dtype = np.dtype(dtype)
        n = len(net.kind)
        self.kind = np.ascontiguousarray(net.kind, dtype=np.int8)
        self.lif = np.ascontiguousarray(np.asarray(net.lif, dtype=np.float64).astype(dtype))
        self.arr_off = (None if net.arr_off is None
                        else np.ascontiguousarray(net.arr_off, dtype=np.int64))
        self.arr_t = (None if net.arr_t is None
                      else np.ascontiguousarray(net.arr_t, dtype=np.int32))
        self.bern_p = (np.zeros(n, np.uint32) if net.bern_p is None
                       else np.ascontiguousarray(net.bern_p, dtype=np.uint32))
        self.pre = np.ascontiguousarray(net.pre, dtype=np.int32)
        self.post = np.ascontiguousarray(net.post, dtype=np.int32)
        self.w = np.ascontiguousarray(np.asarray(net.w).astype(dtype))
        self.d = np.ascontiguousarray(np.asarray(net.d).astype(dtype))
        self.stype = (None if net.stype is None
                      else np.ascontiguousarray(net.stype, dtype=np.int8))
        self.tau = (None if net.tau is None
                    else np.ascontiguousarray(np.asarray(net.tau).astype(dtype)))
        pois = getattr(net, "pois", None)
        self.pois = (None if pois is None
                     else np.ascontiguousarray(np.asarray(pois, np.float64).astype(dtype)))
        suf = "f32" if self.dtype == np.float32 else "f64"
        Net, _ = _structs(suf)
        self.c = Net(n, _ptr(self.kind), _ptr(self.lif), _ptr(self.arr_off), _ptr(self.arr_t),
                     _ptr(self.bern_p), int(net.seed), len(self.pre), _ptr(self.pre),
                     _ptr(self.post), _ptr(self.w), _ptr(self.d), _ptr(self.stype),
                     _ptr(self.tau), _ptr(self.pois))


@dataclass
class OracleResult:
    raster_count: np.ndarray     # [steps] fired per cycle
    raster_ids: np.ndarray       # concatenated ascending ids
    rec_v: np.ndarray            # [steps, n_rec]
    fire_count: np.ndarray       # [n]
    final_v: np.ndarray
    final_ie: np.ndarray
    final_ii: np.ndarray
    final_refrac: np.ndarray
    events: int
    input_sum: np.ndarray = None   # [n] FireInfo "Y": LIF input E + I summed over the cycles
                                   # it neither was refractory nor fired (lif.cu:116)
    slot_count: np.ndarray = None  # [n] FireInfo "X": delay slots of its spikes visited by
                                   # the deliveries (static.cu:42, :82, :106)
    wall_s: float = 0.0        # seconds spent in cycles >= time_from
    events_from: int = 0       # deliveries in cycles >= time_from
    rec_ie: np.ndarray = None    # [steps, n_rec] i_E after each cycle's update
    rec_ii: np.ndarray = None    # [steps, n_rec] i_I after each cycle's update
    rec_in_e: np.ndarray = None  # [steps, n_rec] E input each cycle's update consumed
    rec_in_i: np.ndarray = None  # [steps, n_rec] I input each cycle's update consumed

    def raster(self):
        """List of per-cycle ascending fired-id arrays."""
        off = np.concatenate([[0], np.cumsum(self.raster_count)])
        return [self.raster_ids[off[i]:off[i + 1]] for i in range(len(self.raster_count))]


def run(net, dt, steps, rec_ids=(), dtype=np.float32, mode="exact", tile=4096, threads=1,
        raster_cap=None, time_from=-1):
    """Run the oracle for ``steps`` cycles.  mode: "exact" (creation-order
    delivery) or "csr" (tile/OpenMP form, needs the ordering property)."""
    L = lib()
    a = _NetArgs(net, dtype)
    suf = "f32" if a.dtype == np.float32 else "f64"
    _, Out = _structs(suf)
    n = len(a.kind)
    if raster_cap is None:
        raster_cap = max(1024, n * min(steps, 64))
    rec = np.ascontiguousarray(np.asarray(rec_ids, dtype=np.int32))
    res = dict(raster_count=np.zeros(steps, np.int32),
               raster_ids=np.zeros(raster_cap, np.int32),
               rec_v=np.zeros((steps, len(rec)), a.dtype),
               rec_ie=np.zeros((steps, len(rec)), a.dtype),
               rec_ii=np.zeros((steps, len(rec)), a.dtype),
               rec_in_e=np.zeros((steps, len(rec)), a.dtype),
               rec_in_i=np.zeros((steps, len(rec)), a.dtype),
               fire_count=np.zeros(n, np.int32),
               final_v=np.zeros(n, a.dtype), final_ie=np.zeros(n, a.dtype),
               final_ii=np.zeros(n, a.dtype), final_refrac=np.zeros(n, np.int32),
               input_sum=np.zeros(n, a.dtype), slot_count=np.zeros(n, np.int32))
    while True:
        out = Out(_ptr(res["raster_count"]), _ptr(res["raster_ids"]), raster_cap, 0,
                  len(rec), _ptr(rec) if len(rec) else None,
                  _ptr(res["rec_v"]) if len(rec) else None, _ptr(res["fire_count"]),
                  _ptr(res["final_v"]), _ptr(res["final_ie"]), _ptr(res["final_ii"]),
                  _ptr(res["final_refrac"]), 0, int(time_from), 0.0, 0,
                  _ptr(res["input_sum"]), _ptr(res["slot_count"]),
                  *[_ptr(res[k]) if len(rec) else None
                    for k in ("rec_ie", "rec_ii", "rec_in_e", "rec_in_i")])
        dt_c = float(np.asarray(dt, a.dtype))
        if mode == "exact":
            rc = getattr(L, f"bso_run_{suf}")(C.byref(a.c), dt_c, steps, C.byref(out))
        else:
            rc = getattr(L, f"bso_run_csr_{suf}")(C.byref(a.c), dt_c, steps, tile, threads,
                                                  C.byref(out))
        if rc == -5:
            raster_cap *= 4
            res["raster_ids"] = np.zeros(raster_cap, np.int32)
            continue
        if rc != 0:
            raise RuntimeError(f"oracle run failed: rc={rc}")
        break
    res["raster_ids"] = res["raster_ids"][:out.raster_len].copy()
    return OracleResult(events=int(out.events), wall_s=float(out.wall_s),
                        events_from=int(out.events_from), **res)


def lif_constants(net, dt, dtype=np.float32):
    """Per-neuron [Cm, CE, CI, C_E, C_I, v_tmp, v_thresh, v_reset, v_init] and
    refrac_time, as LIFENeuron::reset computes them."""
    L = lib()
    a = _NetArgs(net, dtype)
    suf = "f32" if a.dtype == np.float32 else "f64"
    n = len(a.kind)
    out9 = np.zeros((n, 9), a.dtype)
    rt = np.zeros(n, np.int32)
    dt_c = float(np.asarray(dt, a.dtype))
    rc = getattr(L, f"bso_lif_constants_{suf}")(C.byref(a.c), dt_c, out9.ctypes.data,
                                                rt.ctypes.data)
    if rc:
        raise RuntimeError(f"oracle constants failed: rc={rc}")
    return out9, rt


def synapse_reset(net, dt, dtype=np.float32):
    """(weights after StaticSynapse::reset, delay_steps)."""
    L = lib()
    a = _NetArgs(net, dtype)
    suf = "f32" if a.dtype == np.float32 else "f64"
    w = np.zeros(len(a.pre), a.dtype)
    ds = np.zeros(len(a.pre), np.int32)
    dt_c = float(np.asarray(dt, a.dtype))
    rc = getattr(L, f"bso_synapse_reset_{suf}")(C.byref(a.c), dt_c, w.ctypes.data,
                                                ds.ctypes.data)
    if rc:
        raise RuntimeError(f"oracle synapse reset failed: rc={rc}")
    return w, ds


def bern_hash(seed, ids, t):
    L = lib()
    return np.array([L.bso_hash(seed, int(i), int(t)) for i in np.atleast_1d(ids)], np.uint32)


def csr_order_ok(net, delay_steps):
    L = lib()
    pre = np.ascontiguousarray(net.pre, np.int32)
    post = np.ascontiguousarray(net.post, np.int32)
    ds = np.ascontiguousarray(delay_steps, np.int32)
    return bool(L.bso_csr_order_ok(len(net.kind), len(pre), pre.ctypes.data, post.ctypes.data,
                                   ds.ctypes.data))
