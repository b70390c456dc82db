"""Pins of the FP32-storage restatements (oracle/lbshift.c), precision forms
the reference does not have: FP32-shift (FP32 deviations, FP32 collision)
and FP32-store (the same storage, collision and j_old in FP64).  Their
results against the FP64 restatement (pinned to the reference's own outputs)
within each form's error bounds, exact invariants (rest state, conservation
of sum_r (P + Pa)) and the bulk closed form D_eff = Db.

The bounds below are the measured accuracy of the forms on these inputs
with margin; the weakest case is the strong Poiseuille flow of the 4x4x34
slit (Peclet ~25), where the flux error of the FP32-shift steady state
(~2.5e-5 relative; FP32-store ~3.9e-6) carries into D_eff through Taylor
dispersion.
"""
import numpy as np
import pytest

import lbref
from golden_cases import Case

CASES = ["slit_4x4x34", "bcc12_ads", "cyl9_tau08", "random_3d"]


@pytest.fixture(scope="module")
def runs():
    out = {}
    for name in CASES:
        c = Case(name)
        a = lbref.equilibrate(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target)
        b = lbref.equilibrate_shift(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, max_steps=4 * a["steps"])
        s = lbref.equilibrate_shift(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, max_steps=4 * a["steps"],
                                    arith="fp64")
        out[name] = (c, a, b, s)
    return out


@pytest.mark.parametrize("name", CASES)
def test_shift_flow_within_fp32_bounds_of_fp64(runs, name):
    c, a, b, _ = runs[name]
    assert b["status"] == lbref.OK
    assert b["exit_reason"] in (lbref.EXIT_CONVERGED, lbref.EXIT_FLOOR, lbref.EXIT_CAP)
    jm = np.abs(a["j"]).max()
    assert np.abs(a["j"] - b["j"]).max() <= 5e-5 * jm
    assert np.abs(a["rho"] - b["rho"]).max() <= 1e-7
    assert np.abs(a["n"] - b["n"]).max() <= 1e-7
    # the l2err trace follows the FP64 one until the FP32 floor
    k = min(len(a["l2err"]), len(b["l2err"]), 100)
    big = a["l2err"][:k] > 1e-6
    assert np.all(np.abs(a["l2err"][:k][big] - b["l2err"][:k][big]) <= 1e-3 * a["l2err"][:k][big])


@pytest.mark.parametrize("name", CASES)
def test_shift_mp_within_bounds_and_conserving(runs, name):
    c, a, b, _ = runs[name]
    itf = lbref.interfacial(*c.dims, c.nature)
    nt = lbref.tracer_population(c.dims, c.nature, a["rho"], a["j"], c.fext)
    mp = lbref.MP(c.dims, c.nature, itf, nt, a["rho"], c.Db, c.ka, c.kd)
    st, v = mp.run(c.mp_steps)
    assert st == lbref.OK
    st2, v2, P, Pa = lbref.mp_shift(c.dims, c.nature, itf, b["rho"], b["j"], c.fext, c.Db, c.ka, c.kd, c.mp_steps)
    assert st2 == lbref.OK
    d, d2 = lbref.deff(v), lbref.deff(v2)
    assert np.abs(d - d2).max() <= 1e-4 * np.abs(d).max()
    # the FP32-rounded propagator conserves sum_r (P + Pa) up to FP64 rounding
    _, _, P1, Pa1 = lbref.mp_shift(c.dims, c.nature, itf, b["rho"], b["j"], c.fext, c.Db, c.ka, c.kd, 0)
    for a_ in range(3):
        s1 = P1[a_::3].sum() + Pa1[a_::3].sum()
        s = P[a_::3].sum() + Pa[a_::3].sum()
        scale = np.abs(P1[a_::3]).sum()
        assert abs(s - s1) <= 1e-12 * scale


def test_shift_rest_state_is_a_fixed_point():
    """No force, rest start: every deviation stays exactly 0."""
    _, nat = lbref.geometry(3, 7, 7, 7)
    dims = (7, 7, 7)
    b = lbref.equilibrate_shift(dims, nat, 0.8, 1.3, [0, 0, 0], 1e-12, max_steps=20)
    a0 = lbref.tables()[1]
    fluid = nat == 0
    n = b["n"].reshape(19, -1)
    for l in range(19):
        assert np.all(n[l][fluid] == a0[l] * 1.3)
    assert np.all(b["j"] == 0.0) and np.all(b["rho"][fluid] == 1.3)


def test_shift_bulk_deff_equals_db():
    """Bulk lattice at rest: D_eff = Db exactly in exact arithmetic; the FP32
    rounding of the scattering probabilities shifts it by < 1e-7 relative."""
    dims = (6, 6, 6)
    nat = np.zeros(216, np.int8)
    itf = np.zeros(216, np.int8)
    rho = np.ones(216)
    j = np.zeros((3, 216))
    for Db in (0.01, 0.1):
        st, v, _, _ = lbref.mp_shift(dims, nat, itf, rho, j, [0, 0, 0], Db, 0.0, 0.0, 20)
        assert st == lbref.OK
        d = lbref.deff(v)
        assert np.all(np.abs(d - Db) <= 1e-7 * Db)


def test_shift_floor_exit_on_strong_flow(runs):
    """The strong slit flow reaches the FP32 rounding floor (l2err cycling
    above 1e-10): the floor test ends it instead of running forever."""
    c, a, b, _ = runs["slit_4x4x34"]
    assert b["exit_reason"] == lbref.EXIT_FLOOR
    assert b["l2err"][-1] > c.target


@pytest.mark.parametrize("name", CASES)
def test_store_flow_closer_to_fp64_than_shift(runs, name):
    """FP32-store: the same FP32 storage with the collision in FP64 puts the
    steady flux within 1e-5 of max|j| on every case (3.9e-6 on the slit) and
    nearer the FP64 flux than the FP32-shift form."""
    c, a, b, s = runs[name]
    assert s["status"] == lbref.OK
    assert s["exit_reason"] in (lbref.EXIT_CONVERGED, lbref.EXIT_FLOOR, lbref.EXIT_CAP)
    jm = np.abs(a["j"]).max()
    err_store, err_shift = np.abs(a["j"] - s["j"]).max(), np.abs(a["j"] - b["j"]).max()
    assert err_store <= 1e-5 * jm
    assert err_store < err_shift
    assert np.abs(a["rho"] - s["rho"]).max() <= 1e-8
    assert np.abs(a["n"] - s["n"]).max() <= 1e-8


@pytest.mark.parametrize("name", CASES)
def test_store_deff_within_1e5(runs, name):
    """D_eff of the FP32-store flow (FP32 probabilities, FP64 P) within 1e-5
    relative of the FP64 path on every case, the slit included."""
    c, a, _, s = runs[name]
    itf = lbref.interfacial(*c.dims, c.nature)
    nt = lbref.tracer_population(c.dims, c.nature, a["rho"], a["j"], c.fext)
    mp = lbref.MP(c.dims, c.nature, itf, nt, a["rho"], c.Db, c.ka, c.kd)
    st, v = mp.run(c.mp_steps)
    st2, v2, _, _ = lbref.mp_shift(c.dims, c.nature, itf, s["rho"], s["j"], c.fext, c.Db, c.ka, c.kd, c.mp_steps)
    assert st == st2 == lbref.OK
    d, d2 = lbref.deff(v), lbref.deff(v2)
    assert np.abs(d - d2).max() <= 1e-5 * np.abs(d).max()


def test_store_rest_state_is_a_fixed_point():
    _, nat = lbref.geometry(3, 7, 7, 7)
    dims = (7, 7, 7)
    s = lbref.equilibrate_shift(dims, nat, 0.8, 1.3, [0, 0, 0], 1e-12, max_steps=20, arith="fp64")
    a0 = lbref.tables()[1]
    fluid = nat == 0
    n = s["n"].reshape(19, -1)
    for l in range(19):
        assert np.all(n[l][fluid] == a0[l] * 1.3)
    assert np.all(s["j"] == 0.0) and np.all(s["rho"][fluid] == 1.3)
