"""Writes the golden vectors under tests/golden/ from the reference's own CPU
simulator (oracle/_ref/ref_stsim_*, built from the BSim sources by
oracle/build_ref.sh).  Run in the build container, where the reference tree
exists:

    python oracle/build_ref.sh && python tests/golden/make_golden.py [case ...]

Each case is an .npz (no pickles) holding the case's generator arguments, a
SHA-256 of the generated network arrays (so a changed generator is caught),
and the reference outputs: per-cycle raster (counts + ascending ids), the V,
i_E, i_I and consumed E / I input traces of the recorded neurons (raw bits;
the input and i_E / i_I traces in cases written since they were added), LIF constants after reset (raw
bits), synapse weights and delay steps after reset (arrays for small cases,
digests for large ones) and per-neuron spike counts.

The reference's own V traces test/lif_unit{1..4}_test_v.std (a double build,
printed with %.10lf) are copied as data into lif_unit_std.npz.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from bsim_amd import netspec  # noqa: E402
import ref_io  # noqa: E402

REF = os.environ.get("BSIM_REFERENCE", "/root/reference")
DT = 1e-4

# (name, generator, kwargs, steps, rec_ids, dtype)
CASES = [
    *[(f"lif_unit{k}_f32", "lif_unit", {"k": k}, 1000, list(range(30)), "float32")
      for k in (1, 2, 3, 4)],
    *[(f"lif_unit{k}_f64", "lif_unit", {"k": k, "wdtype": "float64"}, 1000, list(range(30)),
       "float64") for k in (1, 2, 3, 4)],
    *[(f"random{s}_f32", "random_net",
       {"seed": s, "shuffle": s % 2 == 1, "allow_dup": s % 3 == 2, "n_lif": 60, "p_conn": 0.25,
        "spike_prob": 0.08}, 300, list(range(0, 68, 3)), "float32") for s in range(8)],
    ("c0_f32", "c0", {"seed": 1234}, 1000, [0, 1000], "float32"),
    ("standard_small_f32", "standard",
     {"npop": 4, "n": 400, "fanout": 60, "w0": 3.0, "inh_frac": 0.25, "dmax": 6,
      "drive_rate_hz": 400.0, "n_drive": 200, "drive_w": 12.0, "seed": 7,
      "i_offset": [10.0, 14.0]}, 600, [200, 600, 1000, 1400, 1799], "float32"),
    # the reference's benchmark topology (test/gpu/standard_test.cpp): targets
    # fed by g[i-1] first and g[1] second, inhibitory side populations
    ("standard_test_f32", "standard_test",
     {"depth": 10, "n": 100, "w": [0.7, 0.5, 0.6, 0.3], "who": 3}, 1000,
     [0, 150, 250, 520, 980, 1000, 1120], "float32"),
    ("standard_test_hi_f32", "standard_test",
     {"depth": 6, "n": 120, "w": [1.3, 1.0, 2.0, 1.0], "who": 2}, 600,
     [0, 130, 250, 610, 720], "float32"),
    # delays that are not whole steps: StaticSynapse::reset's rescale and the
    # explicit / composite E / I tau choice
    *[(f"nonint{s}_f32", "nonint_net", {"seed": s}, 300, list(range(12, 92, 5)), "float32")
      for s in (1, 2)],
    ("nonint1_f64", "nonint_net", {"seed": 1}, 300, list(range(12, 92, 5)), "float64"),
]


def make_net(gen, kwargs):
    kw = dict(kwargs)
    if "w" in kw:
        kw["w"] = tuple(kw["w"])
    if "wdtype" in kw:
        kw["wdtype"] = np.dtype(kw["wdtype"]).type
    if "i_offset" in kw:
        kw["i_offset"] = tuple(kw["i_offset"])
    return getattr(netspec, gen)(**kw)


def net_digest(net) -> str:
    h = hashlib.sha256()
    for a in (net.kind, net.lif, net.pre, net.post, net.w, net.d, net.stype, net.tau,
              net.arr_off, net.arr_t, net.bern_p):
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    h.update(str(int(net.seed)).encode())
    return h.hexdigest()


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if ref_io.ref_binary(np.float32) is None:
        raise SystemExit("build oracle/_ref first: oracle/build_ref.sh")
    only = set(sys.argv[1:])     # case names to (re)write; all when none are given
    for name, gen, kw, steps, rec, dts in CASES:
        if only and name not in only:
            continue
        dtype = np.dtype(dts)
        net = make_net(gen, kw)
        r = ref_io.run_ref(net, DT, steps, rec, dtype)
        small = len(net.pre) <= 200000
        ui = np.uint32 if dtype == np.float32 else np.uint64
        meta = dict(name=name, generator=gen, kwargs=kw, steps=steps, rec_ids=rec, dtype=dts,
                    dt=DT, net_sha256=net_digest(net), syn_w_sha256=digest(r["syn_w"]),
                    syn_ds_sha256=digest(r["syn_ds"]), n=int(len(net.kind)),
                    s=int(len(net.pre)))
        arrays = dict(meta=np.frombuffer(json.dumps(meta).encode(), np.uint8),
                      raster_count=r["raster_count"], raster_ids=r["raster_ids"],
                      rec_v_bits=np.ascontiguousarray(r["rec_v"]).view(ui),
                      **{k + "_bits": np.ascontiguousarray(r[k]).view(ui)
                         for k in ("rec_ie", "rec_ii", "rec_in_e", "rec_in_i")},
                      consts_bits=np.ascontiguousarray(r["consts"]).view(ui),
                      refrac_time=r["refrac_time"], fire_count=r["fire_count"])
        if small:
            arrays["syn_w_bits"] = np.ascontiguousarray(r["syn_w"]).view(ui)
            arrays["syn_ds"] = r["syn_ds"]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, "spikes", len(r["raster_ids"]), "synapses", len(net.pre))
    if only:
        return
    std = {}
    for k in (1, 2, 3, 4):
        std[f"lif_unit{k}"] = np.loadtxt(os.path.join(REF, "test", f"lif_unit{k}_test_v.std"))
    np.savez_compressed(os.path.join(HERE, "lif_unit_std.npz"), **std)
    print("lif_unit_std.npz")


if __name__ == "__main__":
    main()
