"""Pins the CPU oracle (oracle/lbref.c) against what the reference program
itself wrote (tests/golden, from make_golden.py) and against closed forms.

Bit-exact comparisons use np.array_equal (which treats -0.0 == +0.0); the
reference's list-directed text output (l2err.dat, vacf.dat) is read back
exactly because flang prints the shortest round-trip decimal.
"""
import numpy as np
import pytest

import lbref
from golden_cases import CONV_CASES, FULL_CASES, Case, load


def test_tables_match_reference_dump():
    g = load("geometries")
    c, a0, a1, a2, inv = lbref.tables()
    assert np.array_equal(c, g["tables_c"])
    assert np.array_equal(inv + 1, g["tables_inv"])
    assert np.array_equal(np.stack([a0, a1, a2]), g["tables_a"])
    # module_lbmodel.f90:127-147 closed forms
    assert a0.sum() == 1.0
    assert list(a2[[0, 1, 7]]) == [1.5, 0.25, 0.125]
    assert all(np.array_equal(c[inv[l]], -c[l]) for l in range(19))


def test_geometries_match_reference_over_many_sizes():
    g = load("geometries")
    n = 0
    for key in g.files if hasattr(g, "files") else list(g):
        if key.startswith("tables"):
            continue
        kind, lx, ly, lz = key.split("_")
        lx, ly, lz = int(lx), int(ly), int(lz)
        label = {"bcc": 3, "cyl": 2, "slit": 1}[kind]
        st, nat = lbref.geometry(label, lx, ly, lz)
        ref = g[key]
        if ref.size == 0:  # reference stopped: "All nodes are solid"
            assert st == lbref.BAD_ARG, key
            continue
        assert st == 0, key
        assert np.array_equal(nat, ref[0]), key
        assert np.array_equal(lbref.interfacial(lx, ly, lz, nat), ref[1]), key
        n += 1
    assert n > 100


def test_norm2_is_not_plain_euclid_but_matches_on_exact_cases():
    # the scaled form reproduces exact Pythagorean values ...
    assert lbref.norm2([3.0, 4.0]) == 5.0
    assert lbref.norm2([0.0, 0.0, 0.0]) == 0.0
    # ... and differs from sqrt(sum x^2) in the last bit on some geometry inputs
    diffs = 0
    for i in range(1, 40):
        for j in range(1, 40):
            v = np.array([i - 20.5, j - 20.5, 7.5])
            diffs += lbref.norm2(v) != np.sqrt((v * v).sum())
    assert diffs > 0


def test_collide_matches_reference_on_random_inputs():
    g = load("collide_random")
    st, n = lbref.collide(tuple(g["dims"]), g["nature"], float(g["tau"]), g["n_in"], g["rho"], g["j"], g["f"])
    assert st == 0
    assert np.array_equal(n, g["n_out"])
    # solid nodes are untouched (the where-mask of module_collision.f90:88)
    N = g["nature"].size
    solid = np.tile(g["nature"] == 1, 19)
    assert np.array_equal(n[solid], g["n_in"][solid])
    assert N == 60


@pytest.mark.parametrize("name", FULL_CASES)
def test_equilibration_matches_reference_bitwise(name):
    c = Case(name)
    st, nat = lbref.geometry(c.label, *c.dims, c.solid_ijk)
    assert st == 0 and np.array_equal(nat, c.nature)
    assert np.array_equal(lbref.interfacial(*c.dims, nat), c.interfacial)
    r = lbref.equilibrate(c.dims, nat, c.tau, c.rho0, c.fext, c.target)
    assert r["status"] == 0 and r["exit_reason"] == lbref.EXIT_CONVERGED
    assert r["steps"] == int(c.g["l2err_step"][-1])
    assert np.array_equal(r["l2err"], c.g["l2err"])
    assert np.array_equal(r["n"], c.g["n"])
    assert np.array_equal(r["rho"], c.g["rho"])
    assert np.array_equal(r["j"], c.j)


@pytest.mark.parametrize("name", FULL_CASES)
def test_moment_propagation_matches_reference_bitwise(name):
    c = Case(name)
    nt = lbref.tracer_population(c.dims, c.nature, c.g["rho"], c.j, c.fext)
    assert np.array_equal(nt, c.g["tracer_n"])
    mp = lbref.MP(c.dims, c.nature, c.interfacial, nt, c.g["rho"], c.Db, c.ka, c.kd)
    assert mp.status == 0
    st, vacf = mp.run(c.mp_steps)
    assert st == 0
    assert np.array_equal(vacf, c.g["vacf"])


@pytest.mark.parametrize("name", CONV_CASES)
def test_moment_propagation_stops_where_the_reference_converges(name):
    """maximum_moment_propagation_steps = -1: the restatement's convergence
    test (module_moment_propagation.f90:284-288) first passes at the step
    where the reference stopped, with the same VACF at every step."""
    c = Case(name)
    assert c.mp_steps < 0
    r = lbref.equilibrate(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target)
    assert r["steps"] == int(c.g["l2err_step"][-1])
    assert np.array_equal(r["j"], c.j) and np.array_equal(r["rho"], c.g["rho"])
    nt = lbref.tracer_population(c.dims, c.nature, c.g["rho"], c.j, c.fext)
    mp = lbref.MP(c.dims, c.nature, c.interfacial, nt, c.g["rho"], c.Db, c.ka, c.kd)
    vacf = [mp.vacf0]
    while True:
        st, v = mp.step()
        assert st == 0
        vacf.append(v)
        if mp.converged(v):
            break
        assert mp.it < 10 * int(c.g["vacf_step"][-1]), "no convergence"
    assert mp.it == int(c.g["vacf_step"][-1])
    assert np.array_equal(np.array(vacf), c.g["vacf"])
    # the test is the reference's: both limits, every component, and it > 2
    last = np.abs(c.g["vacf"][-1])
    assert np.all(last < 1e-12) and not np.all(np.abs(c.g["vacf"][-2]) < np.float32(1e-12))


def test_failures_match_reference_error_stops():
    g = load("failures")
    rows = {str(n): i for i, n in enumerate(g["names"])}
    # relaxation_time < 0.5 -> module_collision.f90:39
    assert "relaxation_time must be > 0.5" in str(g["message"][rows["fail_tau"]])
    st, nat = lbref.geometry(1, 3, 3, 6)
    r = lbref.equilibrate((3, 3, 6), nat, 0.4, 1.0, [1e-5, 0, 0], 1e-10)
    assert r["status"] == lbref.BAD_TAU
    # negative population -> equilibration.f90:248, after l2err_lines steps
    i = rows["fail_negpop"]
    assert "population n(x,y,z,vel) < 0" in str(g["message"][i])
    st, nat = lbref.geometry(1, 2, 2, 8)
    r = lbref.equilibrate((2, 2, 8), nat, 1.0, 1.0, [0.5, 0, 0], 1e-10)
    assert r["status"] == lbref.NEG_POP
    assert r["steps"] == int(g["l2err_lines"][i]) + 1
    # tracer_Db = 0.3: rest probability 1 - 4 Db < 0 -> module_moment_propagation.f90:257
    assert "restpart is negative" in str(g["message"][rows["fail_restpart"]])
    st, nat = lbref.geometry(1, 3, 3, 6)
    r = lbref.equilibrate((3, 3, 6), nat, 1.0, 1.0, [0, 0, 0], 1e-10)
    nt = lbref.tracer_population((3, 3, 6), nat, r["rho"], r["j"], [0, 0, 0])
    mp = lbref.MP((3, 3, 6), nat, lbref.interfacial(3, 3, 6, nat), nt, r["rho"], 0.3)
    st, _ = mp.step()
    assert st == lbref.RESTPART


# ---------------------------------------------------------------- closed forms

def test_rest_state_stage_one_exits_at_t3_for_any_geometry():
    for label, dims in [(1, (3, 2, 7)), (3, (9, 9, 9)), (2, (7, 7, 2))]:
        st, nat = lbref.geometry(label, *dims)
        r = lbref.equilibrate(dims, nat, 0.9, 1.0, [1e-5, 2e-6, 0.0], 1e-10, max_steps=4)
        assert list(r["l2err"][:3]) == [0.0, 0.0, 0.0]
        assert r["l2err"][3] > 0.0  # the force acts from t = 4


def test_poiseuille_second_difference_slit():
    """Steady plane Poiseuille: u(z+1) - 2u(z) + u(z-1) = -F/(rho nu),
    nu = (tau - 1/2)/3, at interior nodes (the 4x4x34 slit, F = 1e-5, tau 1: -6e-5)."""
    c = Case("slit_4x4x34")
    lx, ly, lz = c.dims
    ux = (c.g["jx"] / np.where(c.g["rho"] > 0, c.g["rho"], 1)).reshape(lz, ly, lx)[:, 0, 0]
    d2 = ux[3:lz - 3] - 2 * ux[2:lz - 4] + ux[1:lz - 5]
    nu = (c.tau - 0.5) / 3
    assert np.allclose(d2, -c.fext[0] / nu, rtol=2e-4)


def test_bulk_tracer_diffuses_at_Db():
    """All-fluid periodic box at rest: VACF(0) = 2 Db, P(r,1) = 0, so D = Db."""
    dims = (4, 3, 5)
    st, nat = lbref.geometry(-1, *dims)
    r = lbref.equilibrate(dims, nat, 1.0, 1.0, [0, 0, 0], 1e-10)
    nt = lbref.tracer_population(dims, nat, r["rho"], r["j"], [0, 0, 0])
    mp = lbref.MP(dims, nat, lbref.interfacial(*dims, nat), nt, r["rho"], 0.01)
    assert np.allclose(mp.vacf0, 0.02, rtol=1e-13)
    assert np.all(mp.P == 0.0)
    st, vacf = mp.run(5)
    d = lbref.deff(vacf)
    assert np.allclose(d, 0.01, rtol=1e-13)


def test_slit_transverse_vacf0_closed_form():
    """The 4x4x34 slit (no adsorption): D_y(0) = Db (N_bulk + 5/6 N_int)/N_f, so
    VACF_y(0) = 2 Db (480 + 32*5/6)/512."""
    c = Case("slit_4x4x34")
    assert np.isclose(c.g["vacf"][0, 1], 2 * 0.01 * (480 + 32 * 5 / 6) / 512, rtol=1e-13)


def test_probability_is_conserved_with_adsorption():
    c = Case("bcc12_ads")
    nt = lbref.tracer_population(c.dims, c.nature, c.g["rho"], c.j, c.fext)
    mp = lbref.MP(c.dims, c.nature, c.interfacial, nt, c.g["rho"], c.Db, c.ka, c.kd)
    assert mp.ads == 1
    total0 = mp.P.reshape(-1, 3).sum(0)
    for _ in range(50):
        mp.step()
    total = (mp.P + mp.Pa).reshape(-1, 3).sum(0)
    assert np.allclose(total, total0, rtol=0, atol=1e-14 * np.abs(mp.P).sum())


def test_deff_brute_force_and_trapezoid_weight():
    rs = np.random.RandomState(0)
    v = rs.normal(size=(101, 3)) * 1e-3
    d = lbref.deff(v)
    ref = 0.5 * v[0] + np.array([sum(float(x) for x in v[1:, a]) for a in range(3)])
    assert np.allclose(d, ref, rtol=1e-14)
