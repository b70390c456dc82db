"""The FP32-storage precision forms on the GPU: FP32-shift
(Lattice(precision=FP32_SHIFT): FP32 deviations, FP32 collision) and
FP32-store (precision=FP32_STORE: the same storage, collision in FP64).

Bar: bit-identical to their C restatements in oracle/lbshift.c (per-node
arithmetic in the same order): the l2err trace, exit step and reason, and
the FP64 n, rho, j built from the FP32 deviations; P and Pa of the moment
propagation; VACF to 1e-12 of |VACF(0)| (summation order differs).  Against
the FP64 path (the reference's results) each form's own error bounds hold.
In stage 2 both forms measure l2err every 64 steps; the trace holds NaN at
the other steps, in the library and the restatement alike.
"""
import sys

import numpy as np
import pytest

import lbref
from golden_cases import Case
from laboetie_amd import lbx

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SHIFT = lbx.PREC_FP32_SHIFT
STORE = lbx.PREC_FP32_STORE
ARITH = {SHIFT: "fp32", STORE: "fp64"}


@pytest.mark.parametrize("prec", [SHIFT, STORE])
@pytest.mark.parametrize("name,cap", [("bcc12_ads", 0), ("random_3d", 1600), ("cyl9_tau08", 1400),
                                      ("slit_4x4x34", 12000)])
def test_shift_matches_its_restatement(name, cap, prec):
    c = Case(name)
    say("fp32 storage", prec, name, c.dims, cap)
    ref = lbref.equilibrate_shift(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, max_steps=cap,
                                  arith=ARITH[prec])
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=prec) as L:
        L.init_rest(c.rho0)
        r = L.equilibrate(c.fext, c.target, max_steps=cap)
        assert (r["steps"], r["exit_reason"]) == (ref["steps"], ref["exit_reason"])
        assert np.array_equal(r["l2err"], ref["l2err"], equal_nan=True)
        n, rho, j = L.get_state()
        assert np.array_equal(n, ref["n"]) and np.array_equal(rho, ref["rho"]) and np.array_equal(j, ref["j"])
        steps = min(c.mp_steps, 300)
        v0 = L.mp_init(c.fext, c.Db, c.ka, c.kd)
        m = L.mp_run(steps)
        itf = lbref.interfacial(*c.dims, c.nature)
        st, v, P, Pa = lbref.mp_shift(c.dims, c.nature, itf, ref["rho"], ref["j"], c.fext, c.Db, c.ka, c.kd, steps)
        assert st == lbref.OK
        scale = np.abs(v[0]).max()
        assert np.all(np.abs(v0 - v[0]) <= 1e-12 * scale)
        assert np.all(np.abs(m["vacf"] - v[1:]) <= 1e-12 * scale)
        Pg, Pag = L.mp_get()
        assert np.array_equal(Pg, P) and np.array_equal(Pag, Pa)


@pytest.mark.parametrize("prec", [SHIFT, STORE])
@pytest.mark.parametrize("name", ["bcc12_ads", "random_3d", "cyl9_tau08"])
def test_shift_error_against_fp64(name, prec):
    """FP32-shift and FP32-store vs FP64 on the GPU: flux within 1e-5 of
    max|j| (FP32-store: 2e-6) plus 2e-10, the state difference the stop
    tolerance of 1e-10 on max|dj| leaves (the weak flow of bcc12_ads, max|j|
    4e-6, stops 66 steps apart), D_eff within 1e-6 relative."""
    c = Case(name)
    out = []
    for prec_ in (lbx.PREC_FP64, prec):
        with lbx.Lattice(c.dims, c.nature, c.tau, precision=prec_) as L:
            L.init_rest(c.rho0)
            r = L.equilibrate(c.fext, c.target, max_steps=3000)
            _, rho, j = L.get_state()
            L.mp_init(c.fext, c.Db, c.ka, c.kd)
            L.mp_run(c.mp_steps)
            out.append((r, rho, j, L.mp_deff()))
    (r64, rho64, j64, d64), (r32, rho32, j32, d32) = out
    say(name, "steps", r64["steps"], r32["steps"], "exit", r32["exit_reason"])
    assert np.abs(j32 - j64).max() <= (1e-5 if prec == SHIFT else 2e-6) * np.abs(j64).max() + 2e-10
    assert np.abs(rho32 - rho64).max() <= 1e-8
    assert np.all(np.abs(d32 - d64) <= 1e-6 * np.abs(d64).max())


def _disk(n):
    """1 x n x n, periodic, one solid disk of radius n/5 in the middle"""
    y, z = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    return (((y - n / 2) ** 2 + (z - n / 2) ** 2) <= (n / 5) ** 2).astype(np.int8).ravel()


@pytest.mark.parametrize("prec", [SHIFT, STORE])
@pytest.mark.parametrize("case", ["floor", "box_tau1", "box_tau07", "box_tau3"])
def test_fp32_forms_floor_and_periodic_box_match_restatement(case, prec):
    """Bit-identical to the restatement on two more kinds of run: a floor
    exit (the 1x20x20 disk at F = 1e-4 with target 1e-13, which FP32 storage
    cannot reach), and a periodic box that accelerates without end for 2000
    steps at tau = 1, 0.7 (1/tau inexact in FP32) and 3 (1/tau < 1/2): the
    FP32 relaxation coefficients with A + B = 1 are the restatement's."""
    if case == "floor":
        dims, nat, tau, F, target, cap = (1, 20, 20), _disk(20), 1.0, [0.0, 1e-4, 0.0], 1e-13, 40000
    else:
        tau = {"box_tau1": 1.0, "box_tau07": 0.7, "box_tau3": 3.0}[case]
        dims, nat, F, target, cap = (4, 4, 4), np.zeros(64, np.int8), [1e-5, 0.0, 0.0], 1e-30, 2000
    say("fp32 storage", prec, case)
    ref = lbref.equilibrate_shift(dims, nat, tau, 1.0, F, target, max_steps=cap, arith=ARITH[prec])
    with lbx.Lattice(dims, nat, tau, precision=prec) as L:
        L.init_rest(1.0)
        r = L.equilibrate(F, target, max_steps=cap)
        assert (r["steps"], r["exit_reason"]) == (ref["steps"], ref["exit_reason"])
        assert r["exit_reason"] == (lbx.EXIT_FLOOR if case == "floor" else lbx.EXIT_CAP)
        assert np.array_equal(r["l2err"], ref["l2err"], equal_nan=True)
        n, rho, j = L.get_state()
        assert np.array_equal(n, ref["n"]) and np.array_equal(rho, ref["rho"]) and np.array_equal(j, ref["j"])


@pytest.mark.parametrize("prec", [SHIFT, STORE])
@pytest.mark.parametrize("every", [1, 16])
def test_check_interval_matches_restatement(every, prec):
    """lbx_set_check_interval(1): l2err of every step, as the FP64 form;
    16: every 16th step.  Bit-identical to the restatement with the same
    interval (trace, exit, fields), single lattice and batch."""
    c = Case("cyl9_tau08")
    say("check interval", every, prec)
    ref = lbref.equilibrate_shift(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, max_steps=1400,
                                  arith=ARITH[prec], check_every=every)
    tr = ref["l2err"]
    assert np.isnan(tr).any() == (every > 1)
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=prec) as L:
        L.set_check_interval(every)
        L.init_rest(c.rho0)
        r = L.equilibrate(c.fext, c.target, max_steps=1400)
        assert (r["steps"], r["exit_reason"]) == (ref["steps"], ref["exit_reason"])
        assert np.array_equal(r["l2err"], tr, equal_nan=True)
        n, rho, j = L.get_state()
        assert np.array_equal(n, ref["n"]) and np.array_equal(rho, ref["rho"]) and np.array_equal(j, ref["j"])
    nf = int((c.nature == 0).sum())
    with lbx.Batch((nf + 255) // 256 * 256 + 4096, int(np.prod(c.dims)) + 4096, 2, precision=prec) as B:
        B.set_check_interval(every)
        iid = B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, 10, max_lb_steps=1400)
        while B.round(37, 37) > 0:
            pass
        g = B.result(iid, vacf_rows=11, flow=True)
        assert B.info(iid)["lb_steps"] == ref["steps"]
        assert np.array_equal(g["rho"], ref["rho"]) and np.array_equal(g["j"], ref["j"])


def test_check_interval_refusals():
    c = Case("bcc12_ads")
    with lbx.Lattice(c.dims, c.nature, c.tau) as L:  # FP64: per-step l2err only
        L.set_check_interval(1)
        with pytest.raises(lbx.LbxError):
            L.set_check_interval(64)
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=SHIFT) as L:
        for bad in (0, 3, 8192):
            with pytest.raises(lbx.LbxError):
                L.set_check_interval(bad)


def test_shift_batched_tracers_and_floor_window():
    c = Case("bcc12_ads")
    ka = [c.ka, 0.0]
    kd = [c.kd, 0.0]
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=SHIFT) as L:
        L.init_rest(c.rho0)
        L.equilibrate(c.fext, c.target)
        single = []
        for t in range(2):
            L.mp_init(c.fext, c.Db, ka[t], kd[t])
            single.append(L.mp_run(50)["vacf"])
        L.mp_init_batch(c.fext, c.Db, ka, kd)
        vb = L.mp_run(50)["vacf"]
        for t in range(2):
            assert np.array_equal(vb[:, t, :], single[t])
        L.set_floor_window(0)
        L.set_floor_window(500)
    with lbx.Lattice(c.dims, c.nature, c.tau) as L:
        L.set_floor_window(0)
        with pytest.raises(lbx.LbxError) as e:
            L.set_floor_window(10)
        assert e.value.status == lbx.ERR_ARG
