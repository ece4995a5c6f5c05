"""Loads the golden vectors of tests/golden/ (written by make_golden.py from
the reference's own CPU simulator) and regenerates each case's network."""
import glob
import hashlib
import json
import os

import numpy as np

from bsim_amd import netspec

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not p.endswith("lif_unit_std.npz"))


def net_digest(net) -> str:
    h = hashlib.sha256()
    for a in (net.kind, net.lif, net.pre, net.post, net.w, net.d, net.stype, net.tau,
              net.arr_off, net.arr_t, net.bern_p):
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    h.update(str(int(net.seed)).encode())
    return h.hexdigest()


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    kw = dict(meta["kwargs"])
    if "w" in kw:
        kw["w"] = tuple(kw["w"])
    if "wdtype" in kw:
        kw["wdtype"] = np.dtype(kw["wdtype"]).type
    if "i_offset" in kw:
        kw["i_offset"] = tuple(kw["i_offset"])
    net = getattr(netspec, meta["generator"])(**kw)
    dtype = np.dtype(meta["dtype"])
    data = {k: z[k] for k in z.files if k != "meta"}
    fl = np.float32 if dtype == np.float32 else np.float64
    data["rec_v"] = data["rec_v_bits"].view(fl)
    for k in ("rec_ie", "rec_ii", "rec_in_e", "rec_in_i"):
        if k + "_bits" in data:
            data[k] = data[k + "_bits"].view(fl)
    data["consts"] = data["consts_bits"].view(fl)
    if "syn_w_bits" in data:
        data["syn_w"] = data["syn_w_bits"].view(fl)
    return meta, net, dtype, data


def load_std():
    z = np.load(os.path.join(GOLDEN, "lif_unit_std.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}
