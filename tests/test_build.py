"""The HIP kernels as compiled for gfx950: no private (scratch) memory in any
instantiation -- a spilled accumulator array once cost the larger tiles a
third of their speed -- and no scalar-memory stores."""
import os
import re
import subprocess

import pytest

from bsim_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _device_asm(src, tmp_path):
    out = tmp_path / (os.path.basename(src) + ".s")
    cmd = [build.hipcc(), *[f for f in build.flags() if f not in ("-fopenmp", "-fPIC")],
           "--offload-device-only", "-S", src, "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


@pytest.mark.parametrize("src", ["step.hip", "mlp.hip"])
def test_kernels_use_no_scratch(src, tmp_path):
    asm = _device_asm(os.path.join(ROOT, "bsim_amd", "csrc", src), tmp_path)
    names = re.findall(r"^\s+\.name:\s+(\S+)", asm, flags=re.M)
    priv = [int(x) for x in re.findall(r"^\s+\.private_segment_fixed_size:\s+(\d+)", asm, flags=re.M)]
    assert len(priv) >= 2 and len(priv) == len(names)
    assert all(p == 0 for p in priv), [(n, p) for n, p in zip(names, priv) if p]
    assert not re.search(r"^\s+s_(store|buffer_store|atomic)", asm, flags=re.M)
