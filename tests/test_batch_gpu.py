"""The batch engine (lbxb_*): instances advanced together in one packed
fluid-node list give the same results as the single-lattice path, alone or
in a mix (the equivalence a served request relies on).

Bar: with 256-node blocks, bit-identical LB step counts, exit reasons, flow
fields, VACF(t) and D_eff; with 128-node blocks, identical flow fields and
VACF within 1e-12 of |VACF(0)| (the per-block VACF partials group nodes
differently).
"""
import sys

import numpy as np
import pytest

from golden_cases import Case
from laboetie_amd import lbx

pytestmark = pytest.mark.gpu

CASES = ["bcc12_ads", "cyl9_tau08", "random_3d", "chromat_2disks", "slit_4x4x34"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def single(c, prec, mp_steps):
    with lbx.Lattice(c.dims, c.nature, c.tau, precision=prec) as L:
        L.init_rest(c.rho0)
        r = L.equilibrate(c.fext, c.target, max_steps=20000)
        _, rho, j = L.get_state()
        v0 = L.mp_init(c.fext, c.Db, c.ka, c.kd)
        m = L.mp_run(mp_steps)
        return dict(steps=r["steps"], exit_reason=r["exit_reason"], rho=rho, j=j,
                    vacf=np.vstack([v0, m["vacf"]]), deff=L.mp_deff())


def run_batch(cases, prec, mp_steps, block=256, graphs=True, lb_round=64, mp_round=64):
    cap = sum(((int((c.nature == 0).sum()) + 255) // 256) * 256 for c in cases) + 4096
    box = sum(int(np.prod(c.dims)) for c in cases) + 4096
    with lbx.Batch(cap, box, len(cases) + 2, precision=prec) as B:
        B.set_block(block, block)
        B.set_graphs(graphs)
        ids = [B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, mp_steps,
                     max_lb_steps=20000) for c in cases]
        assert None not in ids
        rounds = 0
        while B.round(lb_round, mp_round) > 0:
            rounds += 1
            assert rounds < 2000
        out = []
        for iid in ids:
            inf = B.info(iid)
            assert inf["phase"] == lbx.PHASE_DONE, inf
            res = B.result(iid, vacf_rows=mp_steps + 1, flow=True)
            res.update(inf)
            out.append(res)
        return out, B.stats()


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT, lbx.PREC_FP32_STORE])
def test_batch_matches_single_lattice_bitwise(prec):
    cases = [Case(n) for n in CASES]
    mp_steps = 150
    say("batch vs single", prec, [c.name for c in cases])
    got, stats = run_batch(cases, prec, mp_steps)
    for c, g in zip(cases, got):
        s = single(c, prec, mp_steps)
        say(c.name, "steps", s["steps"], g["lb_steps"], "exit", s["exit_reason"], g["exit_reason"])
        assert (g["lb_steps"], g["exit_reason"]) == (s["steps"], s["exit_reason"])
        assert np.array_equal(g["rho"], s["rho"]) and np.array_equal(g["j"], s["j"])
        assert np.array_equal(g["vacf"], s["vacf"])
        assert np.array_equal(g["deff"], s["deff"])
    assert stats["lb_fluid_updates"] > 0 and stats["mp_fluid_updates"] > 0


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT, lbx.PREC_FP32_STORE])
def test_batch_block128_no_graphs_and_mixed_rounds(prec):
    """128-node blocks, per-step launches, rounds of 37 LB / 23 MP steps: the
    FP32 forms' l2err check steps (every 64th after the force switch) fall
    inside rounds and on their last steps."""
    cases = [Case(n) for n in ["bcc12_ads", "random_3d", "cyl9_tau08"]]
    mp_steps = 100
    ref, _ = run_batch(cases, prec, mp_steps)
    got, _ = run_batch(cases, prec, mp_steps, block=128, graphs=False, lb_round=37, mp_round=23)
    for c, r, g in zip(cases, ref, got):
        assert g["lb_steps"] == r["lb_steps"]
        assert np.array_equal(g["rho"], r["rho"]) and np.array_equal(g["j"], r["j"])
        scale = np.abs(r["vacf"][0]).max()
        assert np.all(np.abs(g["vacf"] - r["vacf"]) <= 1e-12 * scale)


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT, lbx.PREC_FP32_STORE])
@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_batch_block64(prec, mode):
    """64-node blocks (one wave, four blocks per 256-node tile; a point of the
    autotuner's search), in graph-replayed and persistent launches: the LB
    fields and exit steps bit-identical to 256-node blocks, VACF within 1e-12
    (its partials sum per block)."""
    cases = [Case(n) for n in ["bcc12_ads", "random_3d", "cyl9_tau08"]]
    mp_steps = 100
    ref, _ = run_batch(cases, prec, mp_steps)
    if mode == "graph":
        got, _ = run_batch(cases, prec, mp_steps, block=64)
    else:
        cap = sum(((int((c.nature == 0).sum()) + 255) // 256) * 256 for c in cases) + 4096
        box = sum(int(np.prod(c.dims)) for c in cases) + 4096
        with lbx.Batch(cap, box, len(cases) + 2, precision=prec) as B:
            B.set_block(64, 64)
            B.set_launch_mode(lbx.LAUNCH_PERSISTENT)
            ids = [B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, mp_steps,
                         max_lb_steps=20000) for c in cases]
            rounds = 0
            while B.round(64, 64) > 0:
                rounds += 1
                assert rounds < 2000
            got = []
            for iid in ids:
                g = B.result(iid, vacf_rows=mp_steps + 1, flow=True)
                g.update(B.info(iid))
                got.append(g)
    for c, r, g in zip(cases, ref, got):
        assert g["lb_steps"] == r["lb_steps"]
        assert np.array_equal(g["rho"], r["rho"]) and np.array_equal(g["j"], r["j"])
        scale = np.abs(r["vacf"][0]).max()
        assert np.all(np.abs(g["vacf"] - r["vacf"]) <= 1e-12 * scale)


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT, lbx.PREC_FP32_STORE])
def test_batch_rounds_of_many_lengths_same_bits(prec):
    """Graph-replayed rounds of 1, 2, ..., 90 steps (more distinct round
    lengths than the captured-graph cache keeps, so it starts over) give the
    same bits as rounds of 64, in every precision form."""
    cases = [Case(n) for n in ["bcc12_ads", "random_3d"]]
    mp_steps = 120
    ref, _ = run_batch(cases, prec, mp_steps)
    cap = sum(((int((c.nature == 0).sum()) + 255) // 256) * 256 for c in cases) + 4096
    box = sum(int(np.prod(c.dims)) for c in cases) + 4096
    with lbx.Batch(cap, box, len(cases) + 1, precision=prec) as B:
        ids = [B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, mp_steps,
                     max_lb_steps=20000) for c in cases]
        k = 0
        while B.round(1 + k % 90, 1 + k % 90) > 0:
            k += 1
            assert k < 5000
        for c, iid, r in zip(cases, ids, ref):
            g = B.result(iid, vacf_rows=mp_steps + 1, flow=True)
            assert B.info(iid)["lb_steps"] == r["lb_steps"]
            assert np.array_equal(g["rho"], r["rho"]) and np.array_equal(g["j"], r["j"])
            assert np.array_equal(g["vacf"], r["vacf"]) and np.array_equal(g["deff"], r["deff"])


def test_batch_admission_autotune_and_slot_reuse():
    c = Case("bcc12_ads")
    nf = int((c.nature == 0).sum())
    tiles = (nf + 255) // 256
    box = int(np.prod(c.dims))
    with lbx.Batch(2 * tiles * 256, 2 * box, 3) as B:
        a = B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, 80)
        b = B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, 80)
        assert a is not None and b is not None
        # no room for a third lattice of this size: admission refuses it
        assert B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, 80) is None
        B.round(16, 0)  # stage 1 ends at t = 3
        B.round(16, 0)  # force switch, then stage 2
        assert B.info(a)["phase"] == lbx.PHASE_LB
        t = B.autotune(3)
        assert t["lb_block"] in (64, 128, 256)
        assert all(t["ms"][k] > 0 for k in ("lb64", "lb128", "lb256", "lb_graph", "lb_persistent"))
        assert t["ms"]["mp64"] == 0 and t["ms"]["mp_graph"] == 0  # no MP phase yet
        assert t["mode"] in (lbx.LAUNCH_STREAM, lbx.LAUNCH_GRAPH, lbx.LAUNCH_PERSISTENT)
        while B.round(64, 64) > 0:
            pass
        ra = B.result(a, vacf_rows=81, flow=True)
        B.remove(a)
        # the freed slot and nodes take a new instance, which gives the same results
        a2 = B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, 80)
        assert a2 is not None
        while B.round(64, 64) > 0:
            pass
        r2 = B.result(a2, vacf_rows=81, flow=True)
        assert np.array_equal(ra["j"], r2["j"]) and np.array_equal(ra["rho"], r2["rho"])
        scale = np.abs(ra["vacf"][0]).max()
        assert np.all(np.abs(ra["vacf"] - r2["vacf"]) <= 1e-12 * scale)


def test_served_requests_equal_alone():
    """Requests arriving while others run (continuous batching) give the same
    D_eff, bit for bit, as each request solved alone on the single-lattice
    path."""
    from laboetie_amd import serving
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(6):
        N = int(rng.integers(8, 17))
        kind = k % 3
        nat = lbx.build_geometry(3 if kind == 0 else 2, N, N, N) if kind < 2 else lbx.geometry_random(N, N, N, 2.0,
                                                                                                    0.6, k + 1)
        dF = float(rng.uniform(-3, 0))
        reqs.append(dict(dims=(N, N, N), nature=nat, tau=1.0, rho0=1.0, fext=(0.0, 0.0, 2e-6), target=1e-9, Db=0.01,
                         ka=0.01 * np.exp(-dF), kd=0.01, mp_steps=120, max_lb=3000))
    arrivals = np.array([0.0, 0.0, 0.002, 0.004, 0.006, 0.008])
    cap, box = serving.capacity_for(reqs[:3])
    s = serving.serve(reqs, arrivals, lbx.PREC_FP64, cap, box, max_instances=4, lb_round=32, mp_round=32)
    assert np.all(s["status"] == 0) and np.all(np.isfinite(s["latency"]))
    for r, k in zip(reqs, range(len(reqs))):
        with lbx.Lattice(r["dims"], r["nature"], r["tau"]) as L:
            L.init_rest(r["rho0"])
            L.equilibrate(r["fext"], r["target"], max_steps=r["max_lb"])
            L.mp_init(r["fext"], r["Db"], r["ka"], r["kd"])
            L.mp_run(r["mp_steps"])
            assert np.array_equal(L.mp_deff(), s["deff"][k]), k


@pytest.mark.parametrize("prec", [lbx.PREC_FP64, lbx.PREC_FP32_SHIFT, lbx.PREC_FP32_STORE])
def test_batch_persistent_launches_same_bits(prec):
    """A round's steps as one cooperative launch per phase (grid-wide barrier
    between steps) give the same bits as per-step launches."""
    cases = [Case(n) for n in ["bcc12_ads", "random_3d", "cyl9_tau08", "slit_4x4x34"]]
    mp_steps = 90
    ref, _ = run_batch(cases, prec, mp_steps)
    cap = sum(((int((c.nature == 0).sum()) + 255) // 256) * 256 for c in cases) + 4096
    box = sum(int(np.prod(c.dims)) for c in cases) + 4096
    with lbx.Batch(cap, box, len(cases) + 2, precision=prec) as B:
        B.set_launch_mode(lbx.LAUNCH_PERSISTENT)
        assert B.launch_mode == lbx.LAUNCH_PERSISTENT
        ids = [B.add(c.dims, c.nature, c.tau, c.rho0, c.fext, c.target, c.Db, c.ka, c.kd, mp_steps,
                     max_lb_steps=20000) for c in cases]
        while B.round(64, 64) > 0:
            pass
        for c, iid, r in zip(cases, ids, ref):
            inf = B.info(iid)
            g = B.result(iid, vacf_rows=mp_steps + 1, flow=True)
            say("persistent", c.name, inf["lb_steps"], r["lb_steps"])
            assert (inf["lb_steps"], inf["exit_reason"]) == (r["lb_steps"], r["exit_reason"])
            assert np.array_equal(g["rho"], r["rho"]) and np.array_equal(g["j"], r["j"])
            assert np.array_equal(g["vacf"], r["vacf"]) and np.array_equal(g["deff"], r["deff"])
