"""Pins of the oracle that the fixtures do not cover by themselves, and a check
of the reference citations in the sources (CPU only).

* lbref_norm2 is compared bit for bit with the NORM2 intrinsic of the compiler
  that builds the reference (oracle/norm2_dump.f90) on grids of the vectors the
  reference's geometry builders hand to it (module_geometry.f90:229-237,270).
  The plain sqrt(sum x^2) must differ from it somewhere on the same grids, so
  the comparison can tell the two algorithms apart.
* Every `name.f90:N-M` citation in the product sources, the oracle, the tests
  and the documents lies inside the cited reference file.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import lbref

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
NORM2_DUMP = os.path.join(REPO, "oracle", "_ref", "norm2_dump")


def _grid3(lo, hi, step):
    k = np.arange(lo, hi + 1, step, dtype=np.float64) / 2.0
    c, b, a = np.meshgrid(k, k, k, indexing="ij")  # first component fastest
    return np.stack([a.ravel(), b.ravel(), c.ravel()], axis=1)


def _grid2(lo, hi):
    k = np.arange(lo, hi + 1, dtype=np.float64) / 2.0
    b, a = np.meshgrid(k, k, indexing="ij")
    return np.stack([a.ravel(), b.ravel()], axis=1)


def _norm2_rows(v):
    v = np.ascontiguousarray(v)
    out = np.empty(v.shape[0])
    lbref.lib().lbref_norm2_rows(v.ctypes.data, v.shape[1], v.shape[0], out.ctypes.data)
    return out


@pytest.mark.skipif(not os.path.exists(NORM2_DUMP), reason="oracle/_ref/norm2_dump not built (make -C oracle ref)")
def test_norm2_matches_the_compiler(tmp_path):
    out = tmp_path / "norm2.bin"
    subprocess.run([NORM2_DUMP, str(out)], check=True)
    got = np.fromfile(out, dtype=np.float64)
    grids = [_grid3(-64, 64, 1), _grid3(-1200, 1200, 25), _grid2(-600, 600)]
    assert got.size == sum(g.shape[0] for g in grids)
    off, plain_differs = 0, 0
    for g in grids:
        want = got[off:off + g.shape[0]]
        off += g.shape[0]
        mine = _norm2_rows(g)
        bad = np.flatnonzero(mine.view(np.uint64) != want.view(np.uint64))
        assert bad.size == 0, f"{bad.size} differences, first at {g[bad[0]]}: {mine[bad[0]]!r} vs {want[bad[0]]!r}"
        plain = np.sqrt((g * g).sum(axis=1))
        plain_differs += int(np.count_nonzero(plain != want))
    assert plain_differs > 0, "the grids do not separate NORM2's algorithm from sqrt(sum x^2)"


_CITE = re.compile(r"([A-Za-z_][A-Za-z0-9_]*\.f90):(\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)")
_SCANNED = ("include", "laboetie_amd", "oracle", "tests", "bench.py", "__graft_entry__.py", "DESIGN.md",
            "INTEGRATION.md", "README.md")


def _scanned_files():
    for entry in _SCANNED:
        p = os.path.join(REPO, entry)
        if os.path.isfile(p):
            yield p
            continue
        for root, dirs, files in os.walk(p):
            dirs[:] = [d for d in dirs if not d.startswith(("_", ".")) and d != "__pycache__"]
            for f in files:
                if f.endswith((".h", ".hip", ".cpp", ".c", ".inc", ".py", ".f90", ".md", "Makefile")):
                    yield os.path.join(root, f)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "src")), reason="reference sources not present")
def test_reference_citations_in_range():
    lengths = {}
    for root, _, files in os.walk(REFERENCE):
        for f in files:
            if f.endswith(".f90"):
                with open(os.path.join(root, f), errors="replace") as fh:
                    n = sum(1 for _ in fh)
                lengths[f] = max(lengths.get(f, 0), n)
    checked, bad = 0, []
    for path in _scanned_files():
        with open(path, errors="replace") as fh:
            for ln, line in enumerate(fh, 1):
                for m in _CITE.finditer(line):
                    name = m.group(1)
                    if name not in lengths:  # one of this repository's own .f90 files
                        continue
                    for part in m.group(2).split(","):
                        a, _, b = part.partition("-")
                        a, b = int(a), int(b or a)
                        checked += 1
                        if not 1 <= a <= b <= lengths[name]:
                            bad.append(f"{os.path.relpath(path, REPO)}:{ln}: {m.group(0)} ({name} has {lengths[name]})")
    assert checked > 100
    assert not bad, "\n".join(bad)
