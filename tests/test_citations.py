"""Every reference citation (file:line or file:line-line) in the product and
oracle sources names a line range that exists in the reference tree.  Skipped
where the reference tree is absent."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("BSIM_REFERENCE", "/root/reference")
SCAN = ["oracle", "include", os.path.join("bsim_amd", "csrc"), "bsim_amd", "tests",
        "bench.py", "__graft_entry__.py", "DESIGN.md", "INTEGRATION.md"]
EXT = (".c", ".cpp", ".h", ".hip", ".inc", ".py", ".md", ".sh")
CITE = re.compile(r"((?:[\w.-]+/)*[\w.-]+\.(?:cpp|cu|h|py|c|txt|std|cmake|md)):(\d+)(?:-(\d+))?")


def own_files():
    """This repository's files: repository-relative paths and file names."""
    names, paths = set(), set()
    for dp, dn, fn in os.walk(ROOT):
        dn[:] = [d for d in dn if d not in (".git", "_work", "gpurun_out", "__pycache__")]
        for f in fn:
            names.add(f)
            paths.add(os.path.relpath(os.path.join(dp, f), ROOT).replace(os.sep, "/"))
    return names, paths


def is_own(path, names, paths):
    """A cited path that names a file of this repository (by its path, a tail
    of it, or its file name)."""
    return (path in paths or os.path.basename(path) in names or
            any(p.endswith("/" + path) for p in paths))


def files():
    for s in SCAN:
        p = os.path.join(ROOT, s)
        if os.path.isfile(p):
            yield p
            continue
        for dp, dn, fn in os.walk(p):
            dn[:] = [d for d in dn if d not in ("_ref", "__pycache__", "_obj", "golden")]
            for f in fn:
                if f.endswith(EXT):
                    yield os.path.join(dp, f)
            if s == "bsim_amd":
                break


def ref_index():
    idx = {}
    for dp, dn, fn in os.walk(REF):
        dn[:] = [d for d in dn if not d.startswith(".")]
        for f in fn:
            idx.setdefault(f, []).append(os.path.join(dp, f))
    return idx


def resolve(path, idx):
    """The reference file a cited path names (a path relative to the tree, or
    a unique tail of one)."""
    full = os.path.join(REF, path)
    if os.path.isfile(full):
        return [full]
    cands = [c for c in idx.get(os.path.basename(path), [])
             if c.replace(os.sep, "/").endswith("/" + path)]
    return cands


def citations():
    out = []
    names, paths = own_files()
    for fn in sorted(set(files())):
        if os.path.abspath(fn) == os.path.abspath(__file__):
            continue
        for ln, line in enumerate(open(fn, errors="replace"), 1):
            for m in CITE.finditer(line):
                path = m.group(1)
                if is_own(path, names, paths):
                    continue
                out.append((os.path.relpath(fn, ROOT), ln, path, int(m.group(2)),
                            int(m.group(3) or m.group(2))))
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_every_reference_citation_exists():
    idx = ref_index()
    cites = citations()
    assert len(cites) > 100
    bad = []
    for where, ln, path, a, b in cites:
        cands = resolve(path, idx)
        if not cands:
            bad.append(f"{where}:{ln}: {path} not found in the reference")
            continue
        ok = False
        for c in cands:
            with open(c, errors="replace") as f:
                n = sum(1 for _ in f)
            ok = ok or (1 <= a <= b <= n)
        if not ok:
            bad.append(f"{where}:{ln}: {path}:{a}-{b} beyond the file's end")
    assert not bad, "\n".join(bad)
