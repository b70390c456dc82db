"""The FP32-storage precision forms on the GPU: FP32-shift
(Lattice(precision=FP32_SHIFT): FP32 deviations, FP32 collision) and
FP32-store (precision=FP32_STORE: the same storage, collision in FP64).

Bar: bit-identical to their C restatements in oracle/lbshift.c (per-node
arithmetic in the same order): the l2err trace, exit step and reason, and
the FP64 n, rho, j built from the FP32 deviations; P and Pa of the moment
propagation; VACF to 1e-12 of |VACF(0)| (summation order differs).  Against
the FP64 path (the reference's results) each form's own error bounds hold.
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
        assert np.array_equal(r["l2err"], ref["l2err"])
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
    max|j| (FP32-store: 2e-6), D_eff within 1e-6 relative on these weak
    flows."""
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
    assert np.abs(j32 - j64).max() <= (1e-5 if prec == SHIFT else 2e-6) * np.abs(j64).max()
    assert np.abs(rho32 - rho64).max() <= 1e-8
    assert np.all(np.abs(d32 - d64) <= 1e-6 * np.abs(d64).max())


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
