"""The 64-bit addressing form of the step kernels gives the same bits as the
32-bit buffer-offset form.

The library addresses the population and moment arrays through 32-bit
buffer offsets while 19*Np*8 < 4 GiB and through 64-bit pointers above that.
Two checks:
- on small lattices, the 64-bit form forced on (lbx_set_addressing,
  lbxb_set_addressing) against the default form, single lattice and batch,
  FP64, FP32-shift and FP32-store: bit-identical populations, density, flux,
  step count, VACF(t), P and D_eff;
- a lattice large enough to select the 64-bit form by itself: the
  reference's random_3d case (7x6x5, periodic) repeated 50 x 60 x 60 times
  (350x360x300, 28.4 M fluid nodes, 19*Np*8 = 4.32e9 bytes).  Every copy
  then runs the small lattice's dynamics exactly, so the flux and density
  are the reference's fixture tiled, bit for bit, the l2err trace and exit
  step are the fixture's, and VACF(t) agrees with the fixture to 1e-12 of
  |VACF(0)| (the normalisation 1/Pstat and the sum order differ).
"""
import sys

import numpy as np
import pytest

from golden_cases import Case
from laboetie_amd import lbx

pytestmark = pytest.mark.gpu

CASES = ["bcc12_ads", "random_3d", "cyl9_tau08", "chromat_2disks"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve(c, prec, wide, mp_steps=80):
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=prec) as L:
        assert L.addressing == 0
        if wide:
            L.set_addressing(1)
            assert L.addressing == 1
        L.init_rest(c.rho0)
        r = L.equilibrate(c.fext, c.target, max_steps=20000)
        n, rho, j = L.get_state()
        v0 = L.mp_init(c.fext, c.Db, c.ka, c.kd)
        m = L.mp_run(mp_steps)
        P, Pa = L.mp_get()
        d = L.mp_deff()
        kd = c.kd if c.kd > 0 else 0.01
        vb0 = L.mp_init_batch(c.fext, c.Db, [0.0, kd, 2.0 * kd], [kd] * 3)
        mb = L.mp_run(mp_steps // 2)
        Pb, Pab = L.mp_get(2)
        return dict(steps=r["steps"], exit_reason=r["exit_reason"], l2err=r["l2err"], n=n, rho=rho, j=j,
                    vacf=np.vstack([v0, m["vacf"]]), P=P, Pa=Pa, deff=d, vb0=vb0, vb=mb["vacf"], Pb=Pb, Pab=Pab,
                    deffb=L.mp_deff())


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT, lbx.PREC_FP32_STORE])
@pytest.mark.parametrize("name", CASES)
def test_forced_64bit_addressing_same_bits(name, prec):
    c = Case(name)
    say("64-bit addressing vs buffer offsets", name, prec)
    a = solve(c, prec, False)
    b = solve(c, prec, True)
    for k in a:  # the FP32 forms' l2err trace is NaN between check steps
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_batch_forced_64bit_addressing_same_bits():
    cases = [Case(n) for n in CASES]
    mp_steps = 60

    def run(wide):
        cap = sum(((int((c.nature == 0).sum()) + 255) // 256) * 256 for c in cases) + 4096
        box = sum(c.N for c in cases) + 4096
        with lbx.Batch(cap, box, len(cases) + 1) as B:
            assert B.addressing == 0
            if wide:
                B.set_addressing(1)
                assert B.addressing == 1
            ids = [B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, mp_steps,
                         max_lb_steps=20000) for c in cases]
            assert None not in ids
            rounds = 0
            while B.round(64, 32) > 0:
                rounds += 1
                assert rounds < 2000
            out = []
            for iid in ids:
                res = B.result(iid, vacf_rows=mp_steps + 1, flow=True)
                res.update(B.info(iid))
                out.append(res)
            return out

    say("batch: 64-bit addressing vs buffer offsets")
    for c, a, b in zip(cases, run(False), run(True)):
        assert (a["lb_steps"], a["exit_reason"]) == (b["lb_steps"], b["exit_reason"]), c.name
        for k in ("rho", "j", "vacf", "deff"):
            assert np.array_equal(a[k], b[k]), (c.name, k)


def test_lattice_past_4gib_selects_64bit_form_and_matches_reference():
    c = Case("random_3d")
    reps = (50, 60, 60)  # copies along x, y, z
    lx, ly, lz = c.dims
    dims = (lx * reps[0], ly * reps[1], lz * reps[2])
    nat = np.tile(c.nature.reshape(lz, ly, lx), (reps[2], reps[1], reps[0])).ravel()
    say("tiled random_3d", dims, "fluid", int((nat == 0).sum()))
    with lbx.Lattice(dims, nat, c.tau) as L:
        assert L.fluid_nodes == 28440000
        assert L.addressing == 1  # 19 * 28440000 * 8 bytes > 4 GiB
        L.init_rest(c.rho0)
        r = L.equilibrate(c.fext, c.target)
        say("tiled random_3d: LB steps", r["steps"])
        assert r["exit_reason"] == lbx.EXIT_CONVERGED
        assert r["steps"] == int(c.g["l2err_step"][-1])
        assert np.array_equal(r["l2err"], c.g["l2err"])
        _, rho, j = L.get_state(populations=False)
        shape = (reps[2], lz, reps[1], ly, reps[0], lx)
        small = (lz, ly, lx)
        assert np.all(rho.reshape(shape) == c.g["rho"].reshape(small)[None, :, None, :, None, :])
        for a in range(3):
            assert np.all(j[a].reshape(shape) == c.j[a].reshape(small)[None, :, None, :, None, :]), a
        del rho, j
        v0 = L.mp_init(c.fext, c.Db, c.ka, c.kd)
        ref = c.g["vacf"]
        scale = np.abs(ref[0]).max()
        m = L.mp_run(c.mp_steps)
        say("tiled random_3d: MP steps", m["steps"])
        assert np.all(np.abs(v0 - ref[0]) <= 1e-12 * scale)
        assert np.all(np.abs(m["vacf"] - ref[1:]) <= 1e-12 * scale), np.abs(m["vacf"] - ref[1:]).max()
