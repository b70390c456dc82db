"""The library's C++ readers of lb.in and geom.in (lbx_lbin_*,
lbx_geom_in_parse) and its supercell.xsf writer (lbx_supercell_xsf), host
code with no device calls.  The readers against the Python reader
(laboetie_amd/lbin.py) and the geometries the reference program built from
the same files (the fixtures' nature arrays, dumped from its node%nature by
oracle/ref_dump.f90): every tag of every fixture's lb.in, the reference
grammar's corner cases, and every fixture's geom.in.  The writer against
the supercell.xsf the reference program writes for every fixture's
geometry, byte for byte."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from golden_cases import load
from laboetie_amd import lbx
from laboetie_amd.lbin import LbIn, LbInError, read_geom_in

FIXTURES = ["bcc12_ads", "chromat_2disks", "cyl9_tau08", "random_3d", "shipped_geom", "slit_4x4x34"]
DP = ["relaxation_time", "initialSolventDensity", "target_error", "tracer_Db", "tracer_ka", "tracer_kd"]
INT = ["lx", "ly", "lz", "geometryLabel", "print_frequency", "maximum_moment_propagation_steps"]


def tags_of(text):
    return sorted({ln.split()[0] for ln in LbIn(text).lines})


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_lb_in_same_values(name):
    text = str(load(name)["lb_in"])
    py, cc = LbIn(text), lbx.LbInNative(text)
    for tag in tags_of(text):
        assert cc.char(tag) == py.char(tag), tag
        if tag == "f_ext":
            assert cc.dp3(tag) == py.dp3(tag)
        elif tag in DP:
            assert cc.dp(tag) == py.dp(tag), tag
        elif tag in INT:
            assert cc.int(tag) == py.int(tag), tag
    assert cc.dp("no_such_tag", 7.5) == 7.5 and cc.int("no_such_tag") is None


def test_failure_cases_lb_in():
    for text in load("failures")["lb_in"]:
        text = str(text)
        py, cc = LbIn(text), lbx.LbInNative(text)
        for tag in tags_of(text):
            assert cc.char(tag) == py.char(tag), (text, tag)


GRAMMAR = """  # a comment line
lx = 12   # trailing comment
lxy = 99
lx = 13
tau	= 0.8
relaxation_time = 8.0d-1
f_ext = 1e-6, 0.0 ,-2.5D-3 / ignored
flag1 = T
flag2 = .false.
flag3 = .TRUE.
name = 'bcc'
steps = 1.0e3
bad = abc
short = 1.0 2.0
#hidden = 5
empty =
"""


def test_grammar_corner_cases_match_python_reader():
    py, cc = LbIn(GRAMMAR), lbx.LbInNative(GRAMMAR)
    assert cc.int("lx") == py.int("lx") == 12           # first match wins, comment cut
    assert cc.int("lxy") == py.int("lxy") == 99         # a longer tag is another tag
    assert cc.dp("tau") is None and not py.has("tau")  # a tab is not the blank a tag needs after it
    assert cc.dp("relaxation_time") == py.dp("relaxation_time") == 0.8  # d exponent
    assert cc.dp3("f_ext") == py.dp3("f_ext") == [1e-6, 0.0, -2.5e-3]   # commas, '/' ends the record
    assert cc.log("flag1") is py.log("flag1") is True
    assert cc.log("flag2") is py.log("flag2") is False
    assert cc.log("flag3") is py.log("flag3") is True
    assert cc.char("name") == py.char("name") == "bcc"
    assert cc.int("steps") == py.int("steps") == 1000   # a real read into an integer
    assert cc.int("hidden") is None and not py.has("hidden")
    assert cc.char("empty") == py.char("empty") == ""
    for getter in ("dp", "int"):
        with pytest.raises(lbx.LbxError) as e:
            getattr(cc, getter)("bad")
        assert e.value.status == lbx.ERR_ARG
        with pytest.raises((LbInError, ValueError)):
            getattr(py, getter)("bad")
    with pytest.raises(lbx.LbxError):
        cc.dp3("short")
    with pytest.raises(LbInError):
        py.dp3("short")
    with pytest.raises(lbx.LbxError):
        cc.log("name")


@pytest.mark.parametrize("name", [n for n in FIXTURES if n in ("chromat_2disks", "random_3d", "shipped_geom")])
def test_geom_in_builds_the_reference_geometry(name):
    g = load(name)
    dims = [int(v) for v in g["dims"]]
    text = str(g["geom_in"])
    nat, ns = lbx.geom_in_nature(text, *dims)
    assert np.array_equal(nat, g["nature"].astype(np.int8).ravel())
    pts = read_geom_in(text)
    assert ns == len({tuple(p) for p in pts.tolist()})
    assert np.array_equal(nat, lbx.build_geometry(0, *dims, pts))


def test_geom_in_edge_cases():
    text = "1 1 1\n\n  2 1 1  # a\n2,1,1\n1 1\n"
    nat, ns = lbx.geom_in_nature(text, 2, 1, 1)
    assert list(nat) == [1, 1] and ns == 2
    for bad in ("3 1 1\n", "0 1 1\n", "1 2 1\n", "1 1 -1\n", "a b c\n"):
        with pytest.raises(lbx.LbxError) as e:
            lbx.geom_in_nature(bad, 2, 1, 1)
        assert e.value.status == lbx.ERR_ARG
    nat, ns = lbx.geom_in_nature("", 2, 2, 2)
    assert not nat.any() and ns == 0


REF_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "laboetie")


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference program not built (needs the reference sources)")
@pytest.mark.parametrize("name", FIXTURES)
def test_supercell_xsf_same_bytes_as_the_reference(name):
    """The reference run with no force and no moment propagation (it stops
    at t = 4) writes output/supercell.xsf before anything else; the
    library's writer gives the same bytes from the same nature and
    interfacial arrays."""
    g = load(name)
    dims = [int(v) for v in g["dims"]]
    lb_in = re.sub(r"(?m)^f_ext = .*$", "f_ext = 0.0 0.0 0.0", str(g["lb_in"]))
    lb_in = re.sub(r"(?m)^maximum_moment_propagation_steps = .*$", "maximum_moment_propagation_steps = 0", lb_in)
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "lb.in"), "w").write(lb_in)
        if str(g["geom_in"]):
            open(os.path.join(d, "geom.in"), "w").write(str(g["geom_in"]))
        p = subprocess.run([REF_BIN], cwd=d, env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True,
                           text=True, timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        ref = open(os.path.join(d, "output", "supercell.xsf")).read()
    got = lbx.supercell_xsf(*dims, g["nature"], g["interfacial"])
    assert got == ref


def test_supercell_xsf_refusals():
    nat = np.zeros(8, np.int8)
    with pytest.raises(lbx.LbxError):
        lbx.supercell_xsf(2, 2, 2, np.full(8, 2, np.int8), nat)   # undocumented nature
    with pytest.raises(lbx.LbxError):
        lbx.supercell_xsf(1, 1, 950001, np.zeros(950001, np.int8), np.zeros(950001, np.int8))  # coordinate 950000
