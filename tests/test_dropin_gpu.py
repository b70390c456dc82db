"""The Fortran drop-in (laboetie_amd/fortran: the reference program with its
module_collision, equilibration and moment_propagation replaced by the lbx
shims) against the reference program itself, on the same lb.in: every output
file the two write must agree -- the LB outputs (l2err.dat, the profiles, the
2-D flux field, supercell.xsf) character for character, vacf.dat to 1e-12
relative to VACF(0).  Both binaries are built in the container that holds
the reference sources (oracle/Makefile, laboetie_amd/fortran/Makefile)."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from golden_cases import GOLDEN, load

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(REPO, "oracle", "_ref", "laboetie")
GPU_BIN = os.path.join(REPO, "laboetie_amd", "fortran", "_build", "laboetie_gpu")

TEXT_OUTPUTS = ["l2err.dat", "mass-flux_profile_along_x.dat", "mass-flux_profile_along_y.dat",
                "mass-flux_profile_along_z.dat", "mean-density_profile_along_x.dat",
                "mean-density_profile_along_y.dat", "mean-density_profile_along_z.dat",
                "mass-flux_field_2d_at_x.eq.1.dat", "supercell.xsf"]


def run_case(exe, lb_in, geom_in, d):
    open(os.path.join(d, "lb.in"), "w").write(lb_in)
    if geom_in:
        open(os.path.join(d, "geom.in"), "w").write(geom_in)
    p = subprocess.run([exe], cwd=d, env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True,
                       text=True, timeout=600)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.skipif(not (os.path.exists(REF_BIN) and os.path.exists(GPU_BIN)),
                    reason="drop-in or reference binary not built (needs the reference sources at build time)")
@pytest.mark.parametrize("name,mp_steps", [("slit_4x4x34", 600), ("bcc12_ads", 400), ("chromat_2disks", 200)])
def test_dropin_matches_reference_program(name, mp_steps):
    g = load(name)
    lb_in = str(g["lb_in"]).replace(
        [ln for ln in str(g["lb_in"]).splitlines() if ln.startswith("maximum_moment_propagation_steps")][0],
        f"maximum_moment_propagation_steps = {mp_steps}")
    print("drop-in case", name, flush=True)
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        rc_r, out_r = run_case(REF_BIN, lb_in, str(g["geom_in"]), dr)
        rc_g, out_g = run_case(GPU_BIN, lb_in, str(g["geom_in"]), dg)
        assert rc_r == 0, out_r
        assert rc_g == 0, out_g
        for f in TEXT_OUTPUTS:
            a = open(os.path.join(dr, "output", f)).read()
            b = open(os.path.join(dg, "output", f)).read()
            assert a == b, f
        vr = np.loadtxt(os.path.join(dr, "output", "vacf.dat"), comments="#")
        vg = np.loadtxt(os.path.join(dg, "output", "vacf.dat"), comments="#")
        assert vr.shape == vg.shape == (mp_steps + 1, 4)
        assert np.array_equal(vr[:, 0], vg[:, 0])
        assert np.all(np.abs(vr[:, 1:] - vg[:, 1:]) <= 1e-12 * np.abs(vr[0, 1:]).max())


BCC32_LB_IN = """lx = 32
ly = 32
lz = 32
geometryLabel = 3
relaxation_time = 1.0
initialSolventDensity = 1.0
f_ext = 0.0 0.0 1e-6
target_error = 1e-10
maximum_moment_propagation_steps = 200
tracer_Db = 0.01
tracer_ka = 0.027182818284590453
tracer_kd = 0.01
"""


@pytest.mark.skipif(not (os.path.exists(REF_BIN) and os.path.exists(GPU_BIN)),
                    reason="drop-in or reference binary not built (needs the reference sources at build time)")
def test_dropin_matches_reference_program_bcc32():
    """The bench's geometry and forcing at 32^3 (BCC spheres, F = 1e-6
    along z, an adsorbing tracer at dF = -1): 43 tiles of fluid nodes
    instead of the fixtures' few, same outputs as the reference program."""
    print("drop-in case bcc32", flush=True)
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        rc_r, out_r = run_case(REF_BIN, BCC32_LB_IN, "", dr)
        rc_g, out_g = run_case(GPU_BIN, BCC32_LB_IN, "", dg)
        assert rc_r == 0, out_r
        assert rc_g == 0, out_g
        for f in TEXT_OUTPUTS:
            assert open(os.path.join(dr, "output", f)).read() == open(os.path.join(dg, "output", f)).read(), f
        vr = np.loadtxt(os.path.join(dr, "output", "vacf.dat"), comments="#")
        vg = np.loadtxt(os.path.join(dg, "output", "vacf.dat"), comments="#")
        assert vr.shape == vg.shape == (201, 4)
        assert np.all(np.abs(vr[:, 1:] - vg[:, 1:]) <= 1e-12 * np.abs(vr[0, 1:]).max())


@pytest.mark.skipif(not (os.path.exists(REF_BIN) and os.path.exists(GPU_BIN)),
                    reason="drop-in or reference binary not built (needs the reference sources at build time)")
@pytest.mark.parametrize("key", ["fp32_shift", "fp32_store"])
def test_dropin_fp32_precision_keys(key):
    """`gpu_precision = fp32_shift` / `fp32_store` in lb.in runs the same
    program in that FP32-storage form: VACF within 1e-6 of VACF(0) of the
    reference program's (FP64), and an unknown value stops with a message."""
    g = load("bcc12_ads")
    mp_steps = 300
    lb_in = str(g["lb_in"]).replace(
        [ln for ln in str(g["lb_in"]).splitlines() if ln.startswith("maximum_moment_propagation_steps")][0],
        f"maximum_moment_propagation_steps = {mp_steps}")
    print("drop-in key", key, flush=True)
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg, \
            tempfile.TemporaryDirectory() as db:
        rc_r, out_r = run_case(REF_BIN, lb_in, str(g["geom_in"]), dr)
        rc_g, out_g = run_case(GPU_BIN, lb_in + f"\ngpu_precision = {key}\n", str(g["geom_in"]), dg)
        assert rc_r == 0 and rc_g == 0, out_g
        vr = np.loadtxt(os.path.join(dr, "output", "vacf.dat"), comments="#")
        vg = np.loadtxt(os.path.join(dg, "output", "vacf.dat"), comments="#")
        assert vr.shape == vg.shape
        scale = np.abs(vr[0, 1:]).max()
        assert np.all(np.abs(vg[:, 1:] - vr[:, 1:]) <= 1e-6 * scale)
        # l2err.dat: every stage-1 step, then only the measured steps (every 64th
        # after the force switch at t = 3), all finite
        l2 = np.loadtxt(os.path.join(dg, "output", "l2err.dat"))
        steps = l2[:, 0].astype(int)
        assert list(steps[:3]) == [1, 2, 3] and np.all(np.isfinite(l2[:, 1]))
        assert np.all((steps[3:] - 3) % 64 == 0) and np.all(np.diff(steps) > 0)
        rc_b, out_b = run_case(GPU_BIN, lb_in + "\ngpu_precision = fp16\n", str(g["geom_in"]), db)
        assert "gpu_precision must be fp64, fp32_shift or fp32_store" in out_b


@pytest.mark.skipif(not (os.path.exists(REF_BIN) and os.path.exists(GPU_BIN)),
                    reason="drop-in or reference binary not built (needs the reference sources at build time)")
@pytest.mark.parametrize("name", ["slit_conv_ads", "slit_conv"])
def test_dropin_stops_where_the_reference_converges(name):
    """maximum_moment_propagation_steps = -1: the drop-in's propagate()
    reports convergence at the reference's step (module_moment_propagation.f90
    :284-288), so both programs write the same number of vacf.dat lines."""
    g = load(name)
    print("drop-in convergence case", name, flush=True)
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        rc_r, out_r = run_case(REF_BIN, str(g["lb_in"]), "", dr)
        rc_g, out_g = run_case(GPU_BIN, str(g["lb_in"]), "", dg)
        assert rc_r == 0, out_r
        assert rc_g == 0, out_g
        for f in TEXT_OUTPUTS:
            assert open(os.path.join(dr, "output", f)).read() == open(os.path.join(dg, "output", f)).read(), f
        vr = np.loadtxt(os.path.join(dr, "output", "vacf.dat"), comments="#")
        vg = np.loadtxt(os.path.join(dg, "output", "vacf.dat"), comments="#")
        assert vr.shape == vg.shape == (int(g["vacf_step"][-1]) + 1, 4)
        assert np.array_equal(vr[:, 0], vg[:, 0])
        assert np.all(np.abs(vr[:, 1:] - vg[:, 1:]) <= 1e-12 * np.abs(vr[0, 1:]).max())


@pytest.mark.skipif(not (os.path.exists(REF_BIN) and os.path.exists(GPU_BIN)),
                    reason="drop-in or reference binary not built (needs the reference sources at build time)")
@pytest.mark.parametrize("name", ["slit_4x4x34", "bcc12_ads"])
def test_dropin_periodic_profiles_and_total_mass_flux(name):
    """print_files_frequency = 50 and write_total_mass_flux = T: the profiles
    at every 50th step (equilibration.f90:154-176) and output/
    total_mass_flux.dat (:125-128,257-261) match the reference program's byte
    for byte, as do the stdout step lines (print_frequency = 7)."""
    g = load(name)
    lb_in = str(g["lb_in"]).replace(
        [ln for ln in str(g["lb_in"]).splitlines() if ln.startswith("maximum_moment_propagation_steps")][0],
        "maximum_moment_propagation_steps = 20") + \
        "print_files_frequency = 50\nwrite_total_mass_flux = T\nprint_frequency = 7\n"
    print("drop-in outputs case", name, flush=True)
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        rc_r, out_r = run_case(REF_BIN, lb_in, str(g["geom_in"]), dr)
        rc_g, out_g = run_case(GPU_BIN, lb_in, str(g["geom_in"]), dg)
        assert rc_r == 0, out_r
        assert rc_g == 0, out_g
        for f in TEXT_OUTPUTS + ["total_mass_flux.dat"]:
            assert open(os.path.join(dr, "output", f)).read() == open(os.path.join(dg, "output", f)).read(), f
        # several profile blocks were written
        assert open(os.path.join(dg, "output", "mass-flux_profile_along_z.dat")).read().count("# timestep") > 2
        step_lines = lambda out: [ln for ln in out.splitlines() if "(target" in ln]
        assert step_lines(out_r) == step_lines(out_g) and len(step_lines(out_g)) > 3
