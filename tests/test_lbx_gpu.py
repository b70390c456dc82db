"""Parity of the HIP path (liblbx.so through its C-ABI) with the reference.

Bar: FP64 populations, density, flux, the per-step l2err trace and the exit
step are bit-identical to the reference's (its golden fixtures, or the C
restatement oracle/lbref.c pinned to them); per-node moment-propagation
fields P are bit-identical to the restatement; VACF(t), a sum over all fluid
nodes that the GPU reduces in a different order, agrees to 1e-12 relative
to the scale sum_r |P(r,t) u*(r)| (bounded here by |VACF(0)|, which
dominates every later term); D_eff to 1e-12 relative.
"""
import sys

import numpy as np
import pytest

import lbref
from golden_cases import FULL_CASES, Case, load
from laboetie_amd import lbx

pytestmark = pytest.mark.gpu

VACF_RTOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def vacf_close(got, ref, scale):
    return np.all(np.abs(got - ref) <= VACF_RTOL * scale)


def test_collide_matches_reference_on_random_inputs():
    g = load("collide_random")
    say("collide_random", tuple(g["dims"]))
    n = lbx.collide(tuple(g["dims"]), g["nature"], float(g["tau"]), g["n_in"], g["rho"], g["j"], g["f"])
    assert np.array_equal(n, g["n_out"])


@pytest.mark.parametrize("name", FULL_CASES)
def test_equilibration_and_mp_match_reference(name):
    c = Case(name)
    say("case", name, c.dims)
    nat = lbx.build_geometry(c.label, *c.dims, c.solid_ijk)
    assert np.array_equal(nat, c.nature)
    with lbx.Lattice(c.dims, nat, c.tau) as L:
        L.init_rest(c.rho0)
        r = L.equilibrate(c.fext, c.target)
        assert r["exit_reason"] == lbx.EXIT_CONVERGED
        assert r["steps"] == int(c.g["l2err_step"][-1])
        assert np.array_equal(r["l2err"], c.g["l2err"])
        n, rho, j = L.get_state()
        assert np.array_equal(n, c.g["n"])
        assert np.array_equal(rho, c.g["rho"])
        assert np.array_equal(j, c.j)
        # moment propagation
        v0 = L.mp_init(c.fext, c.Db, c.ka, c.kd)
        ref = c.g["vacf"]
        scale = np.abs(ref[0]).max()
        assert vacf_close(v0, ref[0], scale)
        m = L.mp_run(c.mp_steps)
        assert m["steps"] == c.mp_steps
        assert vacf_close(m["vacf"], ref[1:], scale), np.abs(m["vacf"] - ref[1:]).max()
        # per-node P against the restatement after the same number of steps
        nt = lbref.tracer_population(c.dims, nat, c.g["rho"], c.j, c.fext)
        mp = lbref.MP(c.dims, nat, c.interfacial, nt, c.g["rho"], c.Db, c.ka, c.kd)
        for _ in range(c.mp_steps):
            mp.step()
        P, Pa = L.mp_get()
        assert np.array_equal(P, mp.P)
        if mp.ads:
            assert np.array_equal(Pa, mp.Pa)
        d = L.mp_deff()
        dref = lbref.deff(ref)
        assert np.all(np.abs(d - dref) <= 1e-12 * np.abs(dref).max())


def test_mp_per_step_calls_equal_one_run():
    """propagate() called once per step (the reference's interface) gives the
    same VACF as one multi-step run."""
    c = Case("bcc12_ads")
    with lbx.Lattice(c.dims, c.nature, c.tau) as L:
        L.init_rest(c.rho0)
        L.equilibrate(c.fext, c.target)
        L.mp_init(c.fext, c.Db, c.ka, c.kd)
        one = np.concatenate([L.mp_run(1)["vacf"] for _ in range(40)])
        L.mp_init(c.fext, c.Db, c.ka, c.kd)
        many = L.mp_run(40)["vacf"]
        assert np.array_equal(one, many)


def test_mp_from_host_flow_equals_resident_flow():
    c = Case("cyl9_tau08")
    with lbx.Lattice(c.dims, c.nature, c.tau) as L:
        L.init_rest(c.rho0)
        L.equilibrate(c.fext, c.target)
        v_res = L.mp_init(c.fext, c.Db, c.ka, c.kd)
        a = L.mp_run(50)["vacf"]
        v_host = L.mp_init(c.fext, c.Db, c.ka, c.kd, flow=(c.g["rho"], c.j))
        b = L.mp_run(50)["vacf"]
        assert np.array_equal(v_res, v_host) and np.array_equal(a, b)


def test_mp_batch_equals_single_tracer_runs():
    """Three tracers propagated together (lbx_mp_init_batch) give, tracer by
    tracer, the same bits as three single-tracer runs: VACF, P, Pa, D_eff."""
    c = Case("bcc12_ads")
    ka = [c.ka, 0.0, lbx.ka_kd_from_dF(-2.0)[0]]
    kd = [c.kd, 0.0, lbx.KD_DEFAULT]
    steps = 60
    say("batch", c.dims, ka, kd)
    with lbx.Lattice(c.dims, c.nature, c.tau) as L:
        L.init_rest(c.rho0)
        L.equilibrate(c.fext, c.target)
        single = []
        for t in range(3):
            v0 = L.mp_init(c.fext, c.Db, ka[t], kd[t])
            v = L.mp_run(steps)["vacf"]
            P, Pa = L.mp_get()
            single.append((v0, v, P, Pa, L.mp_deff()))
        v0b = L.mp_init_batch(c.fext, c.Db, ka, kd)
        assert L.mp_tracers == 3
        vb = L.mp_run(steps)["vacf"]
        assert vb.shape == (steps, 3, 3)
        db = L.mp_deff()
        for t in range(3):
            P, Pa = L.mp_get(t)
            s0, sv, sP, sPa, sd = single[t]
            assert np.array_equal(v0b[t], s0)
            assert np.array_equal(vb[:, t, :], sv)
            assert np.array_equal(P, sP) and np.array_equal(Pa, sPa)
            assert np.array_equal(db[t], sd)
        # one call per step gives the same as one run
        L.mp_init_batch(c.fext, c.Db, ka, kd)
        one = np.concatenate([L.mp_run(1)["vacf"] for _ in range(steps)])
        assert np.array_equal(one, vb)
        # the convergence test is per tracer: a batch cannot stop on it
        r = L.mp_run(5, stop_on_convergence=True, check=False)
        assert r["status"] == lbx.ERR_ARG
        for bad in ([], [0.0] * 5):
            with pytest.raises(lbx.LbxError) as e:
                L.mp_init_batch(c.fext, c.Db, bad, bad)
            assert e.value.status == lbx.ERR_ARG
        # any tracer whose rest probability drops below epsilon stops the batch
        L.mp_init_batch(c.fext, c.Db, [0.0, 2.0], [0.0, 0.01])
        assert L.mp_run(3, check=False)["status"] == lbx.ERR_RESTPART


def test_error_statuses():
    nat = lbx.build_geometry(1, 3, 3, 6)
    with pytest.raises(lbx.LbxError) as e:
        lbx.Lattice((3, 3, 6), nat, tau=0.4)
    assert e.value.status == lbx.ERR_TAU
    g = load("failures")
    rows = {str(n): i for i, n in enumerate(g["names"])}
    nat = lbx.build_geometry(1, 2, 2, 8)
    with lbx.Lattice((2, 2, 8), nat, 1.0) as L:
        L.init_rest(1.0)
        r = L.equilibrate([0.5, 0, 0], 1e-10, check=False)
        assert r["status"] == lbx.ERR_NEG_POP
        assert r["steps"] == int(g["l2err_lines"][rows["fail_negpop"]]) + 1
    nat = lbx.build_geometry(1, 3, 3, 6)
    with lbx.Lattice((3, 3, 6), nat, 1.0) as L:
        L.init_rest(1.0)
        L.equilibrate([0, 0, 0], 1e-10)
        L.mp_init([0, 0, 0], 0.3)
        r = L.mp_run(10, check=False)
        assert r["status"] == lbx.ERR_RESTPART


@pytest.mark.parametrize("label,dims,tau,fext", [
    (3, (40, 40, 40), 1.0, [0.0, 0.0, 1e-6]),
    (2, (33, 33, 20), 0.8, [0.0, 2e-6, 1e-5]),
    (1, (24, 17, 30), 0.65, [1e-5, 0.0, 0.0]),
])
def test_capped_runs_match_restatement_at_larger_sizes(label, dims, tau, fext):
    say("capped", label, dims)
    nat = lbx.build_geometry(label, *dims)
    steps = 150
    ref = lbref.equilibrate(dims, nat, tau, 1.0, fext, 1e-10, max_steps=steps)
    with lbx.Lattice(dims, nat, tau) as L:
        L.init_rest(1.0)
        r = L.equilibrate(fext, 1e-10, max_steps=steps)
        assert r["steps"] == steps and r["exit_reason"] == lbx.EXIT_CAP
        assert np.array_equal(r["l2err"], ref["l2err"])
        n, rho, j = L.get_state()
        assert np.array_equal(n, ref["n"])
        assert np.array_equal(rho, ref["rho"])
        assert np.array_equal(j, ref["j"])
        # MP on the capped flow, 30 steps, against the restatement
        itf = lbref.interfacial(*dims, nat)
        L.mp_init(fext, 0.01, 0.02, 0.01)
        m = L.mp_run(30)
        nt = lbref.tracer_population(dims, nat, rho, j, fext)
        mp = lbref.MP(dims, nat, itf, nt, rho, 0.01, 0.02, 0.01)
        st, vref = mp.run(30)
        scale = np.abs(vref[0]).max()
        assert vacf_close(m["vacf"], vref[1:], scale)
        P, Pa = L.mp_get()
        assert np.array_equal(P, mp.P)


@pytest.mark.parametrize("dims", [(1, 1, 9), (2, 3, 4), (1, 6, 5), (2, 2, 2), (5, 1, 3), (3, 2, 1)])
def test_degenerate_axes_match_restatement(dims):
    """Axes of length 1 (a node is its own neighbour) and 2 (x+1 and x-1 are
    the same node) on seeded random media, against the restatement: LB and
    MP with adsorption, bit for bit (VACF to 1e-12 of |VACF(0)|)."""
    say("degenerate axes", dims)
    rng = np.random.default_rng(sum(dims) * 7 + dims[0])
    N = int(np.prod(dims))
    nat = (rng.random(N) < 0.25).astype(np.int8)
    nat[rng.integers(N)] = 0  # at least one fluid node
    fext = [1e-5, -2e-5, 3e-5]
    steps = 60
    ref = lbref.equilibrate(dims, nat, 0.9, 1.0, fext, 1e-12, max_steps=steps)
    with lbx.Lattice(dims, nat, 0.9) as L:
        L.init_rest(1.0)
        r = L.equilibrate(fext, 1e-12, max_steps=steps)
        assert (r["steps"], r["exit_reason"]) == (ref["steps"], ref["exit_reason"])
        assert np.array_equal(r["l2err"], ref["l2err"])
        n, rho, j = L.get_state()
        assert np.array_equal(n, ref["n"]) and np.array_equal(rho, ref["rho"]) and np.array_equal(j, ref["j"])
        itf = lbref.interfacial(*dims, nat)
        L.mp_init(fext, 0.01, 0.02, 0.01)
        m = L.mp_run(25)
        nt = lbref.tracer_population(dims, nat, rho, j, fext)
        mp = lbref.MP(dims, nat, itf, nt, rho, 0.01, 0.02, 0.01)
        st, vref = mp.run(25)
        scale = np.abs(vref[0]).max()
        assert vacf_close(m["vacf"], vref[1:], scale)
        P, Pa = L.mp_get()
        assert np.array_equal(P, mp.P)
        if mp.ads:
            assert np.array_equal(Pa, mp.Pa)


def test_lattices_the_library_refuses():
    """No fluid node, or a box of 2^29 nodes or more: LBX_ERR_ARG, before any
    device memory is taken."""
    with pytest.raises(lbx.LbxError) as e:
        lbx.Lattice((3, 4, 5), np.ones(60, np.int8), 1.0)
    assert e.value.status == lbx.ERR_ARG
    big = np.zeros(1 << 29, np.int8)
    with pytest.raises(lbx.LbxError) as e:
        lbx.Lattice((1 << 10, 1 << 10, 1 << 9), big, 1.0)
    assert e.value.status == lbx.ERR_ARG


def test_full_size_invariants_bcc128():
    """128^3 BCC at full size: mass is conserved to rounding while the force
    acts, stage 1 ends at t = 3, and the tracer probability (free + adsorbed)
    is conserved by moment propagation."""
    dims = (128, 128, 128)
    say("bcc128 invariants")
    nat = lbx.build_geometry(3, *dims)
    with lbx.Lattice(dims, nat, 1.0) as L:
        nf = L.fluid_nodes
        assert nf == int((nat == 0).sum())
        L.init_rest(1.0)
        r = L.equilibrate([0.0, 0.0, 1e-6], 1e-10, max_steps=400)
        assert list(r["l2err"][:3]) == [0.0, 0.0, 0.0] and r["l2err"][3] > 0
        n, rho, j = L.get_state()
        assert abs(rho.sum() - nf) <= 1e-12 * nf
        assert np.all(n >= 0)
        L.mp_init([0.0, 0.0, 1e-6], 0.01, 0.01 * np.e, 0.01)
        P0, _ = L.mp_get()
        tot0 = P0.reshape(-1, 3).sum(0)
        L.mp_run(200)
        P, Pa = L.mp_get()
        tot = (P + Pa).reshape(-1, 3).sum(0)
        assert np.allclose(tot, tot0, rtol=0, atol=1e-13 * np.abs(P0).sum())


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT])
def test_equilibration_in_pieces_equals_one_run(prec):
    """lbx_equilibrate_begin / _advance with pauses at steps 1, 2, 3, 4, 50,
    the force switch and beyond give the one-shot run's bits (steps, exit,
    l2err trace, state); lbx_flow at a pause is the oracle's state after
    that many steps, and the flux sums are the sums of its flux."""
    c = Case("bcc12_ads")
    say("pieces", prec)
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=prec) as L:
        L.init_rest(c.rho0)
        one = L.equilibrate(c.fext, c.target)
        n1, rho1, j1 = L.get_state()
        L.init_rest(c.rho0)
        L.equilibrate_begin(c.fext, c.target, flux_sums=True)
        flows = {}
        for stop in (1, 2, 3, 4, 50, 51, 333):
            r = L.equilibrate_advance(stop)
            if r["exit_reason"]:
                break
            assert r["steps"] == stop
            flows[stop] = L.flow()
            with pytest.raises(lbx.LbxError) as e:  # the state is extracted at the run's end only
                L.get_state()
            assert e.value.status == lbx.ERR_STATE
        r = L.equilibrate_advance()
        assert (r["steps"], r["exit_reason"]) == (one["steps"], one["exit_reason"])
        assert np.array_equal(L.l2err_trace(), one["l2err"], equal_nan=True)
        n2, rho2, j2 = L.get_state()
        assert np.array_equal(n2, n1) and np.array_equal(rho2, rho1) and np.array_equal(j2, j1)
        sums = L.flux_sums()
        assert sums.shape == (one["steps"], 3)
    # the flux sums are of the step's own flux: FP64, or FP32 in the
    # FP32-shift form (within FP32 rounding of the FP64 flux extracted from
    # its deviations)
    rtol = 1e-12 if prec == lbx.PREC_FP64 else 1e-6
    for stop, (rho, j) in flows.items():
        if prec == lbx.PREC_FP64:
            ref = lbref.equilibrate(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, max_steps=stop)
        else:
            ref = lbref.equilibrate_shift(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, max_steps=stop)
        assert np.array_equal(rho, ref["rho"]) and np.array_equal(j, ref["j"]), stop
        assert np.allclose(sums[stop - 1], ref["j"].sum(axis=1), rtol=rtol, atol=1e-15 * c.N)
    assert np.allclose(sums[-1], j1.sum(axis=1), rtol=rtol, atol=1e-15 * c.N)


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT])
@pytest.mark.parametrize("geom", ["bcc64", "random48", "slit_4x4x34"])
def test_mp_coefficient_forms_give_the_same_bits(geom, prec):
    """The moment-propagation step with the incoming scattering
    probabilities recomputed from the neighbours' density and flow (the FP64
    default), read from the 18 stored per node (the FP32 default), and the
    10 of the z-directions stored with the others recomputed: three adsorbing
    tracers on a developing flow, VACF, P, Pa and D_eff bit-identical."""
    if geom == "bcc64":
        dims, nat, fext, tau = (64, 64, 64), lbx.build_geometry(3, 64, 64, 64), [0.0, 0.0, 1e-6], 1.0
    elif geom == "random48":
        dims, fext, tau = (48, 48, 48), [1e-6, 0.0, 0.0], 0.8
        nat = lbx.geometry_random(*dims, 8.0, 0.6, 20261202)
    else:
        dims, nat, fext, tau = (4, 4, 34), lbx.build_geometry(1, 4, 4, 34), [1e-5, 0.0, 0.0], 1.0
    say("mp coefficient forms", geom, prec)
    ka = [lbx.ka_kd_from_dF(d)[0] for d in (0.0, -1.0, -2.0)]
    kd = [lbx.KD_DEFAULT] * 3
    out = []
    with lbx.Lattice(dims, nat, tau, precision=prec) as L:
        L.init_rest(1.0)
        L.equilibrate(fext, 1e-10, max_steps=300)
        forms = (lbx.MP_STORED, lbx.MP_RECOMPUTED, lbx.MP_MIXED)
        for k, form in enumerate(forms):
            L.set_mp_coefficients(form)
            v0 = L.mp_init_batch(fext, 0.01, ka, kd)
            L.set_mp_coefficients(forms[(k + 1) % 3])  # from the next init only
            v = L.mp_run(150)["vacf"]
            fields = [L.mp_get(t) for t in range(3)]
            out.append((v0, v, fields, L.mp_deff()))
    (a0, va, fa, da) = out[0]
    for (b0, vb, fb, db) in out[1:]:
        assert np.array_equal(a0, b0) and np.array_equal(va, vb) and np.array_equal(da, db)
        for (Pa_, Paa), (Pb, Pab) in zip(fa, fb):
            assert np.array_equal(Pa_, Pb) and np.array_equal(Paa, Pab)


@pytest.mark.parametrize("name", ["bcc32", "slit_conv"])
def test_mp_run_start_wait_equals_mp_run(name):
    """lbx_mp_run_start / lbx_mp_run_wait in chunks of 256, 256 and the rest
    (the drop-in's propagate pattern): the same VACF, P, Pa and D_eff bits as
    one lbx_mp_run, the same convergence step with the test on, and the other
    MP calls refused while a chunk is in flight."""
    if name == "bcc32":
        dims, nat, fext, tau, conv = (32, 32, 32), lbx.build_geometry(3, 32, 32, 32), [0.0, 0.0, 1e-6], 1.0, False
        Db, ka, kd, steps = 0.01, lbx.ka_kd_from_dF(-1.0)[0], lbx.KD_DEFAULT, 600
        rho0, target = 1.0, 1e-10
    else:
        c = Case(name)
        dims, nat, fext, tau, conv = c.dims, c.nature, c.fext, c.tau, True
        Db, ka, kd, steps, rho0, target = c.Db, c.ka, c.kd, 10 * int(c.g["vacf_step"][-1]), c.rho0, c.target
    say("mp start/wait", name, dims)
    with lbx.Lattice(dims, nat, tau) as L:
        L.init_rest(rho0)
        L.equilibrate(fext, target, max_steps=2000)
        L.mp_init(fext, Db, ka, kd)
        ref = L.mp_run(steps, stop_on_convergence=conv)
        P0, Pa0 = L.mp_get()
        d0 = L.mp_deff()
        L.mp_init(fext, Db, ka, kd)
        got, done, converged = [], 0, False
        L.mp_run_start(min(256, steps), stop_on_convergence=conv)
        with pytest.raises(lbx.LbxError):  # a chunk is in flight
            L.mp_deff()
        while True:
            r = L.mp_run_wait()
            got.append(r["vacf"])
            done += r["steps"]
            converged = r["converged"]
            if converged or done >= steps:
                break
            L.mp_run_start(min(256, steps - done), stop_on_convergence=conv)
        assert done == ref["steps"] and converged == ref["converged"]
        assert np.array_equal(np.concatenate(got), ref["vacf"])
        P1, Pa1 = L.mp_get()
        assert np.array_equal(P0, P1) and np.array_equal(Pa0, Pa1) and np.array_equal(d0, L.mp_deff())
        if conv:  # after the convergence step a begun chunk runs nothing
            L.mp_run_start(5, stop_on_convergence=True)
            assert L.mp_run_wait()["steps"] == 0


@pytest.mark.gpu
def test_bandwidth_helpers_measure():
    """The two measured ceilings come back as plausible rates (GB/s)."""
    for g in (lbx.copy_bandwidth(64 << 20, 2), lbx.planes_bandwidth(19, 1 << 20, 2),
              lbx.planes_bandwidth(22, 1 << 20, 2)):
        assert 10.0 < g < 20000.0


@pytest.mark.gpu
@pytest.mark.parametrize("form", [lbx.MP_STORED, lbx.MP_RECOMPUTED, lbx.MP_MIXED])
@pytest.mark.parametrize("dims", [(63, 6, 7), (65, 5, 6), (130, 3, 4), (2, 10, 11), (1, 30, 31)])
def test_mp_lane_reuse_shapes_match_restatement(dims, form):
    """The MP step's reuse of the z-neighbours of the lanes beside (directions
    11..14 through wave lane moves and edge-lane records) on rows that end
    inside a wave, cross wave and block boundaries, are longer than the LDS
    halo, two nodes long (x+1 and x-1 the same node) and one node long (no
    reuse): three adsorbing tracers in each coefficient form, P, Pa bit for bit
    and VACF to 1e-12 of |VACF(0)| against the restatement run per tracer."""
    say("lane reuse shapes", dims, form)
    rng = np.random.default_rng(dims[0] * 1000 + dims[1] * 10 + dims[2])
    N = int(np.prod(dims))
    nat = (rng.random(N) < 0.2).astype(np.int8)
    nat[rng.integers(N)] = 0
    fext = [2e-5, -1e-5, 3e-5]
    ka = [lbx.ka_kd_from_dF(d)[0] for d in (0.0, -1.0, -2.0)]
    kd = [lbx.KD_DEFAULT] * 3
    with lbx.Lattice(dims, nat, 0.9) as L:
        L.init_rest(1.0)
        L.equilibrate(fext, 1e-12, max_steps=40)
        n, rho, j = L.get_state()
        L.set_mp_coefficients(form)
        L.mp_init_batch(fext, 0.01, ka, kd)
        m = L.mp_run(30)
        got = [L.mp_get(t) for t in range(3)]
    itf = lbref.interfacial(*dims, nat)
    nt = lbref.tracer_population(dims, nat, rho, j, fext)
    for t in range(3):
        mp = lbref.MP(dims, nat, itf, nt, rho, 0.01, ka[t], kd[t])
        st, vref = mp.run(30)
        assert vacf_close(m["vacf"][:, t, :], vref[1:], np.abs(vref[0]).max())
        P, Pa = got[t]
        assert np.array_equal(P, mp.P)
        if mp.ads:
            assert np.array_equal(Pa, mp.Pa)


@pytest.mark.parametrize("prec,side,porosity", [(lbx.PREC_FP64, 112, 0.75), (lbx.PREC_FP32_STORE, 128, 0.9)])
def test_streaming_store_lattices_match_restatement(prec, side, porosity):
    """Lattices whose two population sets exceed 256 MiB (the size at which
    the LB step writes its populations with the streaming cache hint) under
    32-bit buffer offsets give the restatement's populations, density, flux
    and l2err trace bit for bit after 8 capped steps on a random medium
    (every step but the first reads what the previous one streamed out)."""
    dims = (side,) * 3
    say("streaming stores", prec, dims)
    nat = lbx.geometry_random(*dims, 6.0, porosity, 20261203)
    nf = int((nat == 0).sum())
    width = 8 if prec == lbx.PREC_FP64 else 4
    assert 2 * 19 * ((nf + 63) // 64 * 64) * width > 256 << 20
    fext, steps = [2e-6, -1e-6, 3e-6], 8
    if prec == lbx.PREC_FP64:
        ref = lbref.equilibrate(dims, nat, 0.8, 1.0, fext, 1e-30, max_steps=steps)
    else:
        ref = lbref.equilibrate_shift(dims, nat, 0.8, 1.0, fext, 1e-30, max_steps=steps, arith="fp64")
    with lbx.Lattice(dims, nat, 0.8, precision=prec) as L:
        assert L.addressing == 0
        L.init_rest(1.0)
        r = L.equilibrate(fext, 1e-30, max_steps=steps)
        assert (r["steps"], r["exit_reason"]) == (ref["steps"], ref["exit_reason"])
        assert np.array_equal(r["l2err"], ref["l2err"], equal_nan=True)
        n, rho, j = L.get_state()
    assert np.array_equal(n, ref["n"]) and np.array_equal(rho, ref["rho"]) and np.array_equal(j, ref["j"])
