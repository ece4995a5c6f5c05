"""Python front end of libbsim_amd.so (the C ABI of include/bsim_amd.h).

Mirrors the reference's user-facing flow -- build a Network, then
``SGSim(&net, dt).run(T)`` (README.md:33-57, src/sim/SingleGPUSimulator.cu:31)
-- for tests and benchmarks:

    net = Network.from_spec(spec)          # or add_lif / connect calls
    sim = Simulation(net, dt, rec_steps=T, rec_ids=[0, 1000])
    sim.run(T)                             # round(T/dt) cycles + final phase (ii)
    ids_per_cycle = sim.raster()

The library must be built (``python -m bsim_amd.build`` or
``__graft_entry__.build()``); importing it never falls back to any other
implementation.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import netspec

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbsim_amd.so")
_LIB = None

# symbols declared in include/bsim_amd.h, with (restype, argtypes)
P = C.c_void_p
_SIGS = {
    "bsim_abi_version": (C.c_int32, []),
    "bsim_last_error": (C.c_char_p, []),
    "bsim_net_create": (P, []),
    "bsim_net_destroy": (None, [P]),
    "bsim_net_add_lif": (C.c_int32, [P, C.c_int32, P]),
    "bsim_net_add_array": (C.c_int32, [P, C.c_int32, P, P, P]),
    "bsim_net_add_bernoulli": (C.c_int32, [P, C.c_int32, P, C.c_uint64]),
    "bsim_net_add_bernoulli_keyed": (C.c_int32, [P, C.c_int32, P, C.c_uint64, C.c_uint32]),
    "bsim_net_add_poisson": (C.c_int32, [P, C.c_int32, P, P, P, P, C.c_uint64]),
    "bsim_net_gen_chain": (C.c_int32, [P, P]),
    "bsim_net_gen_chains": (C.c_int32, [P, P, C.c_int32, P]),
    "bsim_net_connect": (C.c_int64, [P, C.c_int64, P, P, P, P, P, P]),
    "bsim_net_num_neurons": (C.c_int32, [P]),
    "bsim_net_num_synapses": (C.c_int64, [P]),
    "bsim_sim_create": (P, [P, C.c_float, P]),
    "bsim_sim_step": (C.c_int32, [P, C.c_int32]),
    "bsim_sim_finish": (C.c_int32, [P]),
    "bsim_sim_sync": (C.c_int32, [P]),
    "bsim_sim_cycle": (C.c_int32, [P]),
    "bsim_sim_read_raster": (C.c_int64, [P, P, P, C.c_int64, P]),
    "bsim_sim_read_trace": (C.c_int32, [P, P]),
    "bsim_sim_enable_state_trace": (C.c_int32, [P]),
    "bsim_sim_read_state_trace": (C.c_int32, [P, P, P, P, P]),
    "bsim_sim_read_state": (C.c_int32, [P, P, P, P, P, P]),
    "bsim_sim_enable_input_log": (C.c_int32, [P]),
    "bsim_sim_read_input_log": (C.c_int32, [P, P, P]),
    "bsim_sim_events": (C.c_int64, [P]),
    "bsim_sim_time_steps": (C.c_int32, [P, C.c_int32, P, P]),
    "bsim_sim_get_info": (C.c_int32, [P, P]),
    "bsim_sim_stream": (P, [P]),
    "bsim_sim_set_schedule": (C.c_int32, [P, C.c_int32]),
    "bsim_sim_destroy": (None, [P]),
    "bsim_net_gen_standard": (C.c_int64, [P, P]),
    "bsim_net_kinds": (C.c_int32, [P, P]),
    "bsim_sim_create_chains": (P, [P, C.c_int32, C.c_float, P, C.c_int64, P]),
    "bsim_net_export_neurons": (C.c_int32, [P, P, P, P, P]),
    "bsim_net_export_synapses": (C.c_int64, [P, C.c_int64, C.c_int64, P, P, P, P, P]),
    "bsim_net_lif_constants": (C.c_int32, [P, C.c_float, P, P]),
    "bsim_net_layout": (C.c_int32, [P, C.c_float, C.c_int32, P, P, P, P, P, P]),
}


class BsimError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BsimError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _err(rc, what):
    msg = lib().bsim_last_error()
    raise BsimError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _p(a):
    return None if a is None else a.ctypes.data


class SimConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("tile", C.c_int32), ("rec_steps", C.c_int32),
                ("n_rec", C.c_int32), ("rec_ids", P)]


class SimInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("tile", C.c_int32), ("ntiles", C.c_int32),
                ("dmax", C.c_int32), ("s", C.c_int64), ("device_bytes", C.c_int64),
                ("n_src_entries", C.c_int64), ("off_entries", C.c_int64)]


class GenParams(C.Structure):
    _fields_ = [("npop", C.c_int32), ("n", C.c_int32), ("fanout", C.c_int32),
                ("w0", C.c_float), ("inh_frac", C.c_float), ("dmax", C.c_int32),
                ("remote_frac", C.c_float), ("drive_rate_hz", C.c_float),
                ("n_drive", C.c_int32), ("drive_w", C.c_float), ("drive_fanout", C.c_int32),
                ("seed", C.c_uint64), ("tau_refrac", C.c_float), ("ioff_lo", C.c_float),
                ("ioff_hi", C.c_float), ("tau_syn_E", C.c_float), ("tau_syn_I", C.c_float),
                ("v_thresh", C.c_float), ("dt", C.c_float)]


class ChainParams(C.Structure):
    _fields_ = [("npop", C.c_int32), ("sizes", C.c_int32 * 8), ("fan_in", C.c_int32),
                ("w0", C.c_float), ("inh_frac", C.c_float), ("dmax", C.c_int32),
                ("drive_rate_hz", C.c_float), ("n_drive", C.c_int32), ("drive_fan_in", C.c_int32),
                ("drive_w", C.c_float), ("seed", C.c_uint64), ("tau_refrac", C.c_float),
                ("ioff_lo", C.c_float), ("ioff_hi", C.c_float), ("tau_syn_E", C.c_float),
                ("tau_syn_I", C.c_float), ("v_thresh", C.c_float), ("dt", C.c_float)]


class Network:
    """Host network (bsim_network).  Neuron ids are creation order."""

    def __init__(self):
        self._h = lib().bsim_net_create()
        if not self._h:
            _err(-3, "bsim_net_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().bsim_net_destroy(h)
            self._h = None

    @property
    def n(self):
        return lib().bsim_net_num_neurons(self._h)

    @property
    def s(self):
        return lib().bsim_net_num_synapses(self._h)

    @staticmethod
    def chain_params(sizes, fan_in, w0, inh_frac, dmax, drive_rate_hz, n_drive, drive_fan_in,
                     drive_w, seed, tau_refrac=2e-3, i_offset=(14.5, 16.5), tau_syn_E=5e-3,
                     tau_syn_I=10e-3, v_thresh=15.0, dt=netspec.DT, out=None):
        """A bsim_chain_params (filled into out if given)."""
        cp = ChainParams() if out is None else out
        cp.npop = len(sizes)
        for i, x in enumerate(sizes):
            cp.sizes[i] = int(x)
        cp.fan_in, cp.w0, cp.inh_frac, cp.dmax = int(fan_in), w0, inh_frac, int(dmax)
        cp.drive_rate_hz, cp.n_drive, cp.drive_fan_in = drive_rate_hz, int(n_drive), int(drive_fan_in)
        cp.drive_w, cp.seed, cp.tau_refrac = drive_w, int(seed), tau_refrac
        cp.ioff_lo, cp.ioff_hi = i_offset
        cp.tau_syn_E, cp.tau_syn_I, cp.v_thresh, cp.dt = tau_syn_E, tau_syn_I, v_thresh, dt
        return cp

    def add_chains(self, kwargs_list):
        """Appends several sweep networks at once (bsim_net_gen_chains: the
        same networks as add_chain in order); returns their first ids."""
        n = len(kwargs_list)
        arr = (ChainParams * max(n, 1))()
        for i, kw in enumerate(kwargs_list):
            self.chain_params(out=arr[i], **kw)
        first = np.zeros(max(n, 1), np.int32)
        r = lib().bsim_net_gen_chains(self._h, arr, n, _p(first))
        if r < 0:
            _err(r, "bsim_net_gen_chains")
        return first[:n].astype(np.int64)

    def add_chain(self, sizes, fan_in, w0, inh_frac, dmax, drive_rate_hz, n_drive, drive_fan_in,
                  drive_w, seed, tau_refrac=2e-3, i_offset=(14.5, 16.5), tau_syn_E=5e-3,
                  tau_syn_I=10e-3, v_thresh=15.0, dt=netspec.DT):
        """Appends one sweep network (bsim_net_gen_chain); returns its first id."""
        cp = ChainParams()
        cp.npop = len(sizes)
        for i, x in enumerate(sizes):
            cp.sizes[i] = int(x)
        cp.fan_in, cp.w0, cp.inh_frac, cp.dmax = int(fan_in), w0, inh_frac, int(dmax)
        cp.drive_rate_hz, cp.n_drive, cp.drive_fan_in = drive_rate_hz, int(n_drive), int(drive_fan_in)
        cp.drive_w, cp.seed, cp.tau_refrac = drive_w, int(seed), tau_refrac
        cp.ioff_lo, cp.ioff_hi = i_offset
        cp.tau_syn_E, cp.tau_syn_I, cp.v_thresh, cp.dt = tau_syn_E, tau_syn_I, v_thresh, dt
        r = lib().bsim_net_gen_chain(self._h, C.byref(cp))
        if r < 0:
            _err(r, "bsim_net_gen_chain")
        return r

    def add_lif(self, params):
        p = np.ascontiguousarray(np.atleast_2d(params), np.float32)
        assert p.shape[1] == netspec.LIF_NPARAM
        r = lib().bsim_net_add_lif(self._h, p.shape[0], _p(p))
        if r < 0:
            _err(r, "bsim_net_add_lif")
        return r

    def add_array(self, offsets, cycles, taus=None):
        off = np.ascontiguousarray(offsets, np.int64)
        cyc = np.ascontiguousarray(cycles, np.int32)
        tz = None if taus is None else np.ascontiguousarray(taus, np.float32)
        r = lib().bsim_net_add_array(self._h, len(off) - 1, _p(off), _p(cyc), _p(tz))
        if r < 0:
            _err(r, "bsim_net_add_array")
        return r

    def add_bernoulli(self, p, seed):
        pp = np.ascontiguousarray(p, np.uint32)
        r = lib().bsim_net_add_bernoulli(self._h, len(pp), _p(pp), int(seed))
        if r < 0:
            _err(r, "bsim_net_add_bernoulli")
        return r

    def add_poisson(self, rate, start, duration, seed, taus=None):
        """Poisson_spikes(rate, refract, startTime, duration) sources with
        seeded draws (src/neuron/poisson/PoissonNeuron.cpp:13-63); rate,
        start, duration per neuron; taus [n, 2] composite tau_syn_E/I."""
        r_ = np.ascontiguousarray(rate, np.float32)
        n = len(r_)
        s_ = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.float32), (n,)))
        d_ = np.ascontiguousarray(np.broadcast_to(np.asarray(duration, np.float32), (n,)))
        t_ = None if taus is None else np.ascontiguousarray(taus, np.float32).reshape(n, 2)
        r = lib().bsim_net_add_poisson(self._h, n, _p(r_), _p(s_), _p(d_), _p(t_), int(seed))
        if r < 0:
            _err(r, "bsim_net_add_poisson")
        return r

    def connect(self, pre, post, w, d, stype=None, tau=None):
        pre = np.ascontiguousarray(pre, np.int32)
        post = np.ascontiguousarray(post, np.int32)
        w = np.ascontiguousarray(w, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        st = None if stype is None else np.ascontiguousarray(stype, np.int8)
        tz = None if tau is None else np.ascontiguousarray(tau, np.float32)
        r = lib().bsim_net_connect(self._h, len(pre), _p(pre), _p(post), _p(w), _p(d), _p(st),
                                   _p(tz))
        if r < 0:
            _err(r, "bsim_net_connect")
        return r

    @classmethod
    def from_spec(cls, spec: netspec.NetSpec, chunk: int = 1 << 26) -> "Network":
        """Adds the spec's neurons in creation order (runs of one kind) and its
        synapses in creation order."""
        net = cls()
        kind = np.asarray(spec.kind)
        i = 0
        n = len(kind)
        while i < n:
            j = i
            while j < n and kind[j] == kind[i]:
                j += 1
            if kind[i] == netspec.LIF:
                net.add_lif(np.asarray(spec.lif[i:j], np.float64).astype(np.float32))
            elif kind[i] == netspec.ARRAY:
                off = np.asarray(spec.arr_off[i:j + 1], np.int64)
                cyc = np.asarray(spec.arr_t[off[0]:off[-1]], np.int32)
                taus = np.asarray(spec.lif[i:j, 10:12], np.float32)
                net.add_array(off - off[0], cyc, taus)
            elif kind[i] == netspec.POISSON:
                pp = np.asarray(spec.pois[i:j], np.float64).astype(np.float32)
                net.add_poisson(pp[:, 0], pp[:, 1], pp[:, 2], spec.seed,
                                np.asarray(spec.lif[i:j, 10:12], np.float32))
            else:
                net.add_bernoulli(np.asarray(spec.bern_p[i:j], np.uint32), spec.seed)
            i = j
        s = len(spec.pre)
        for a in range(0, s, chunk):
            b = min(s, a + chunk)
            net.connect(spec.pre[a:b], spec.post[a:b], np.asarray(spec.w[a:b], np.float32),
                        np.asarray(spec.d[a:b], np.float32),
                        None if spec.stype is None else spec.stype[a:b],
                        None if spec.tau is None or not np.any(spec.tau) else
                        np.asarray(spec.tau[a:b], np.float32))
        return net

    @classmethod
    def standard(cls, npop=4, n=1000, fanout=100, w0=1.0, inh_frac=0.25, dmax=6,
                 remote_frac=0.2, drive_rate_hz=0.0, n_drive=None, drive_w=None,
                 drive_fanout=None, seed=1, tau_refrac=2e-3, i_offset=(0.0, 0.0),
                 tau_syn_E=5e-3, tau_syn_I=10e-3, v_thresh=15.0, dt=netspec.DT):
        """The standard feed-forward workload built by the library's generator
        (same parameters and same network as netspec.standard)."""
        net = cls()
        p = GenParams(npop, n, fanout, w0, inh_frac, dmax, remote_frac, drive_rate_hz,
                      n if n_drive is None else n_drive,
                      0.7 * float(np.float32(w0)) if drive_w is None else drive_w,
                      fanout if drive_fanout is None else drive_fanout, seed, tau_refrac,
                      i_offset[0], i_offset[1], tau_syn_E, tau_syn_I, v_thresh, dt)
        r = lib().bsim_net_gen_standard(net._h, C.byref(p))
        if r < 0:
            _err(r, "bsim_net_gen_standard")
        return net

    def export_spec(self) -> netspec.NetSpec:
        """The network's arrays (creation order) as a NetSpec."""
        n, s = self.n, self.s
        kind = np.zeros(n, np.int8)
        lif = np.zeros((n, netspec.LIF_NPARAM), np.float32)
        bern = np.zeros(n, np.uint32)
        seed = C.c_uint64()
        r = lib().bsim_net_export_neurons(self._h, _p(kind), _p(lif), _p(bern), C.byref(seed))
        if r < 0:
            _err(r, "bsim_net_export_neurons")
        pre, post = np.zeros(s, np.int32), np.zeros(s, np.int32)
        w, d = np.zeros(s, np.float32), np.zeros(s, np.float32)
        st = np.zeros(s, np.int8)
        r = lib().bsim_net_export_synapses(self._h, 0, s, _p(pre), _p(post), _p(w), _p(d),
                                           _p(st))
        if r < 0:
            _err(r, "bsim_net_export_synapses")
        if np.any(kind == netspec.ARRAY):
            raise BsimError("export_spec does not carry array-source spike lists")
        return netspec.NetSpec(kind=kind, lif=lif.astype(np.float64), pre=pre, post=post, w=w,
                               d=d, stype=st, tau=None, bern_p=bern, seed=int(seed.value))

    def lif_constants(self, dt):
        n = self.n
        out = np.zeros((n, 9), np.float32)
        rt = np.zeros(n, np.int32)
        r = lib().bsim_net_lif_constants(self._h, float(dt), _p(out), _p(rt))
        if r < 0:
            _err(r, "bsim_net_lif_constants")
        return out, rt

    def layout(self, dt, tile):
        nt = C.c_int32()
        s = C.c_int64()
        dm = C.c_int32()
        r = lib().bsim_net_layout(self._h, float(dt), tile, C.byref(nt), C.byref(s),
                                  C.byref(dm), None, None, None)
        if r < 0:
            _err(r, "bsim_net_layout")
        ts = np.zeros(nt.value + 1, np.int64)
        tg = np.zeros(s.value, np.uint16)
        w = np.zeros(s.value, np.float32)
        r = lib().bsim_net_layout(self._h, float(dt), tile, None, None, None, _p(ts), _p(tg),
                                  _p(w))
        if r < 0:
            _err(r, "bsim_net_layout")
        return dict(ntiles=nt.value, s=s.value, dmax=dm.value, tile_syn=ts, syn_tgt=tg, syn_w=w)


def run_cycles(T, dt) -> int:
    """round(T/dt) as SingleGPUSimulator::run computes it
    (src/sim/SingleGPUSimulator.cu:34): the quotient in float, then C's
    round(), which takes halves away from zero (numpy's round takes them to
    even).  The float32 quotient plus 0.5 is exact in double."""
    q = float(np.float32(T) / np.float32(dt))
    r = math.floor(abs(q) + 0.5)
    return int(r if q >= 0 else -r)


class Simulation:
    """Device simulation (bsim_sim): SGSim's run loop on one GPU."""

    def __init__(self, net: Network, dt, tile=0, rec_steps=0, rec_ids=(), device=0):
        self.dt = np.float32(dt)
        self._rec = np.ascontiguousarray(np.asarray(rec_ids, np.int32))
        cfg = SimConfig(device, tile, rec_steps, len(self._rec),
                        _p(self._rec) if len(self._rec) else None)
        self._h = lib().bsim_sim_create(net._h, float(self.dt), C.byref(cfg))
        if not self._h:
            _err(-1, "bsim_sim_create")
        self.n = net.n
        self.n_rec = len(self._rec)
        self.rec_steps = rec_steps

    @classmethod
    def from_chains(cls, kwargs_list, dt, tile=0, rec_steps=0, part_synapses=int(1.5e9),
                    device=0):
        """One simulation of several sweep networks built part by part
        (bsim_sim_create_chains); returns (simulation, first ids)."""
        n = len(kwargs_list)
        arr = (ChainParams * max(n, 1))()
        for i, kw in enumerate(kwargs_list):
            Network.chain_params(out=arr[i], **kw)
        first = np.zeros(max(n, 1), np.int32)
        self = cls.__new__(cls)
        self.dt = np.float32(dt)
        self._rec = np.zeros(0, np.int32)
        cfg = SimConfig(device, tile, rec_steps, 0, None)
        self._h = lib().bsim_sim_create_chains(arr, n, float(self.dt), C.byref(cfg),
                                               int(part_synapses), _p(first))
        if not self._h:
            _err(-1, "bsim_sim_create_chains")
        info = self.info()
        self.n = info["n"]
        self.n_rec = 0
        self.rec_steps = rec_steps
        return self, first[:n].astype(np.int64)

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().bsim_sim_destroy(h)
            self._h = None

    __del__ = close

    def step(self, n=1):
        r = lib().bsim_sim_step(self._h, int(n))
        if r < 0:
            _err(r, "bsim_sim_step")

    def finish(self):
        r = lib().bsim_sim_finish(self._h)
        if r < 0:
            _err(r, "bsim_sim_finish")

    def sync(self):
        r = lib().bsim_sim_sync(self._h)
        if r < 0:
            _err(r, "bsim_sim_sync")

    def run(self, T):
        """SingleGPUSimulator::run(T): round(T/dt) cycles (computed in float,
        src/sim/SingleGPUSimulator.cu:34), then the last cycle's phase (ii)."""
        cycles = run_cycles(T, self.dt)
        self.step(cycles)
        self.finish()
        self.sync()
        return cycles

    @property
    def cycle(self):
        return lib().bsim_sim_cycle(self._h)

    def raster(self):
        """List of ascending fired-id arrays, one per recorded cycle."""
        steps = min(self.cycle, self.rec_steps)
        counts = np.zeros(steps, np.int32)
        need = C.c_int64()
        r = lib().bsim_sim_read_raster(self._h, _p(counts), None, 0, C.byref(need))
        ids = np.zeros(max(1, need.value), np.int32)
        r = lib().bsim_sim_read_raster(self._h, _p(counts), _p(ids), len(ids), C.byref(need))
        if r < 0:
            _err(r, "bsim_sim_read_raster")
        ids = ids[:r]
        off = np.concatenate([[0], np.cumsum(counts)])
        return counts, ids, [ids[off[i]:off[i + 1]] for i in range(steps)]

    def trace(self):
        steps = min(self.cycle, self.rec_steps)
        v = np.zeros((steps, self.n_rec), np.float32)
        r = lib().bsim_sim_read_trace(self._h, _p(v))
        if r < 0:
            _err(r, "bsim_sim_read_trace")
        return v

    def enable_state_trace(self):
        """i_E / i_I after the update and the E / I input it consumed, per
        traced neuron and cycle (bsim_sim_enable_state_trace)."""
        r = lib().bsim_sim_enable_state_trace(self._h)
        if r < 0:
            _err(r, "bsim_sim_enable_state_trace")

    def state_trace(self):
        steps = min(self.cycle, self.rec_steps)
        out = {k: np.zeros((steps, self.n_rec), np.float32) for k in ("ie", "ii", "in_e", "in_i")}
        r = lib().bsim_sim_read_state_trace(self._h, _p(out["ie"]), _p(out["ii"]),
                                            _p(out["in_e"]), _p(out["in_i"]))
        if r < 0:
            _err(r, "bsim_sim_read_state_trace")
        return out

    def state(self):
        n = self.n
        v, ie, ii = (np.zeros(n, np.float32) for _ in range(3))
        rf, fc = np.zeros(n, np.int32), np.zeros(n, np.int32)
        r = lib().bsim_sim_read_state(self._h, _p(v), _p(ie), _p(ii), _p(rf), _p(fc))
        if r < 0:
            _err(r, "bsim_sim_read_state")
        return dict(v=v, ie=ie, ii=ii, refrac=rf, fire_count=fc)

    def time_steps(self, n):
        """Advances n cycles one launch at a time with HIP events around each
        kernel; returns the mean (step kernel, row-list kernel) durations in us."""
        a, b = C.c_double(), C.c_double()
        r = lib().bsim_sim_time_steps(self._h, int(n), C.byref(a), C.byref(b))
        if r < 0:
            _err(r, "bsim_sim_time_steps")
        return a.value, b.value

    def enable_input_log(self):
        """FireInfo "Y" / "X" logs from the next cycle on (bsim_sim_enable_input_log)."""
        r = lib().bsim_sim_enable_input_log(self._h)
        if r < 0:
            _err(r, "bsim_sim_enable_input_log")

    def input_log(self):
        """(Y: summed LIF input per neuron, X: delay slots visited per neuron)."""
        y = np.zeros(self.n, np.float32)
        x = np.zeros(self.n, np.int32)
        r = lib().bsim_sim_read_input_log(self._h, _p(y), _p(x))
        if r < 0:
            _err(r, "bsim_sim_read_input_log")
        return y, x

    def events(self):
        r = lib().bsim_sim_events(self._h)
        if r < 0:
            _err(r, "bsim_sim_events")
        return int(r)

    def info(self):
        i = SimInfo()
        r = lib().bsim_sim_get_info(self._h, C.byref(i))
        if r < 0:
            _err(r, "bsim_sim_get_info")
        return {k: getattr(i, k) for k, _ in SimInfo._fields_}

    def stream(self):
        return lib().bsim_sim_stream(self._h)

    def set_schedule(self, rank_every):
        """Rank the tiles heaviest first every rank_every cycles (0: never);
        results do not depend on it (bsim_sim_set_schedule)."""
        r = lib().bsim_sim_set_schedule(self._h, int(rank_every))
        if r < 0:
            _err(r, "bsim_sim_set_schedule")
