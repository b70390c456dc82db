"""Write the golden fixtures of tests/golden from the reference program.

Runs the reference (maxlevesque/laboetie, compiled by oracle/Makefile into
oracle/_ref/) on small inputs at OMP_NUM_THREADS=1 and stores what it writes
-- its text outputs (l2err.dat, vacf.dat) and the binary dumps of its module
arrays made by oracle/ref_dump.f90 -- as compressed .npz fixtures.  Only this
container has the reference sources needed to build oracle/_ref; the fixtures
are committed so that tests run anywhere.

    python tests/golden/make_golden.py          # all cases
    python tests/golden/make_golden.py slit_4x4x34
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")
TUTO = "/root/reference/tuto"
SHIPPED_GEOM = "/root/reference/geom.in"


def lb_in(d):
    return "".join(f"{k} = {v}\n" for k, v in d.items())


# Flow + tracer cases run through `ref_dump full`.  Keys are lb.in tags.
FULL = {
    # the CPU-runnable case: 4x4x34 slit, tau 1, F = 1e-5 along x, neutral tracer
    "slit_4x4x34": dict(lb=dict(lx=4, ly=4, lz=34, geometryLabel=1, relaxation_time=1.0,
                              initialSolventDensity=1.0, f_ext="1e-5 0.0 0.0", target_error=1e-10,
                              maximum_moment_propagation_steps=5000, tracer_Db=0.01)),
    # BCC of touching spheres (label 3), flow along z, adsorbing tracer (dF = -1 kT)
    "bcc12_ads": dict(lb=dict(lx=12, ly=12, lz=12, geometryLabel=3, relaxation_time=1.0,
                              initialSolventDensity=1.0, f_ext="0.0 0.0 1e-6", target_error=1e-10,
                              maximum_moment_propagation_steps=400, tracer_Db=0.01,
                              tracer_ka=0.01 * np.exp(1.0), tracer_kd=0.01)),
    # cylinder along z (label 2), tau != 1, strong adsorption
    "cyl9_tau08": dict(lb=dict(lx=9, ly=9, lz=5, geometryLabel=2, relaxation_time=0.8,
                               initialSolventDensity=1.0, f_ext="0.0 0.0 1e-5", target_error=1e-11,
                               maximum_moment_propagation_steps=300, tracer_Db=0.02,
                               tracer_ka=0.05, tracer_kd=0.01)),
    # chromatography geometry shipped with the reference (label 0), flow along y
    "chromat_2disks": dict(lb=dict(lx=1, ly=50, lz=50, geometryLabel=0, relaxation_time=1.0,
                                   initialSolventDensity=1.0, f_ext="0.0 1e-5 0.0", target_error=1e-10,
                                   maximum_moment_propagation_steps=300, tracer_Db=0.01,
                                   tracer_ka=0.01 * np.exp(2.0), tracer_kd=0.01),
                           geom_file=os.path.join(TUTO, "geom.in_chromat_2disks-dia10-1x50x50_v1")),
    # seeded random solids, force with three non-zero components, tau 0.7, density 1.3
    "random_3d": dict(lb=dict(lx=7, ly=6, lz=5, geometryLabel=0, relaxation_time=0.7,
                              initialSolventDensity=1.3, f_ext="2e-5 -1e-5 3e-5", target_error=1e-12,
                              maximum_moment_propagation_steps=200, tracer_Db=0.03,
                              tracer_ka=0.03, tracer_kd=0.02),
                      geom_random=dict(seed=10, frac=0.25)),
    # the reference's own geom.in (label 0) in a 10^3 box, no force: stage 2 exits at t=4
    "shipped_geom": dict(lb=dict(lx=10, ly=10, lz=10, geometryLabel=0, f_ext="0.0 0.0 0.0",
                                 maximum_moment_propagation_steps=100, tracer_Db=0.05),
                         geom_file=SHIPPED_GEOM),
}

# Moment propagation run until the reference's convergence test stops it
# (maximum_moment_propagation_steps < 0, drop_tracers.f90:40; the test at
# module_moment_propagation.f90:284-288): no-flow slits, whose VACF decays
# below 1e-12 within a few thousand steps.
CONV = {
    "slit_conv_ads": dict(lb=dict(lx=4, ly=4, lz=8, geometryLabel=1, f_ext="0.0 0.0 0.0", target_error=1e-10,
                                  maximum_moment_propagation_steps=-1, tracer_Db=0.1,
                                  tracer_ka=0.02, tracer_kd=0.01)),
    "slit_conv": dict(lb=dict(lx=5, ly=4, lz=7, geometryLabel=1, f_ext="0.0 0.0 0.0", target_error=1e-10,
                              maximum_moment_propagation_steps=-1, tracer_Db=0.15)),
}

# Reference error stops (section "runtime assertions" of the reference).
FAIL = {
    "fail_tau": dict(lx=3, ly=3, lz=6, geometryLabel=1, relaxation_time=0.4, f_ext="1e-5 0 0"),
    "fail_negpop": dict(lx=2, ly=2, lz=8, geometryLabel=1, f_ext="0.5 0.0 0.0"),
    "fail_restpart": dict(lx=3, ly=3, lz=6, geometryLabel=1, f_ext="0 0 0",
                          maximum_moment_propagation_steps=10, tracer_Db=0.3),
}


def run(cwd, args):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    p = subprocess.run(args, cwd=cwd, env=env, capture_output=True, text=True)
    return p.returncode, p.stdout + p.stderr


def read_eq_state(path):
    b = open(path, "rb").read()
    lx, ly, lz = (int(v) for v in np.frombuffer(b[:12], "<i4"))
    N = lx * ly * lz
    a = np.frombuffer(b[12:], "<f8")
    assert a.size == 23 * N
    return dict(dims=np.array([lx, ly, lz], np.int32), n=a[:19 * N].copy(), rho=a[19 * N:20 * N].copy(),
                jx=a[20 * N:21 * N].copy(), jy=a[21 * N:22 * N].copy(), jz=a[22 * N:23 * N].copy())


def read_geom(path):
    b = open(path, "rb").read()
    lx, ly, lz = (int(v) for v in np.frombuffer(b[:12], "<i4"))
    N = lx * ly * lz
    g = np.frombuffer(b[12:], np.int8)
    return np.array([lx, ly, lz], np.int32), g[:N].copy(), g[N:2 * N].copy()


def random_geom(lx, ly, lz, seed, frac):
    rs = np.random.RandomState(seed)
    pts = [(i, j, k) for k in range(1, lz + 1) for j in range(1, ly + 1) for i in range(1, lx + 1)
           if rs.rand() < frac]
    return "".join(f"{i} {j} {k}\n" for i, j, k in pts)


def make_full(name, spec):
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "lb.in"), "w").write(lb_in(spec["lb"]))
        if "geom_file" in spec:
            open(os.path.join(d, "geom.in"), "w").write(open(spec["geom_file"]).read())
        if "geom_random" in spec:
            lb = spec["lb"]
            open(os.path.join(d, "geom.in"), "w").write(
                random_geom(lb["lx"], lb["ly"], lb["lz"], **spec["geom_random"]))
        rc, out = run(d, [os.path.join(REF, "ref_dump"), "full"])
        assert rc == 0, out
        fx = read_eq_state(os.path.join(d, "output", "eq_state.bin"))
        l2 = np.loadtxt(os.path.join(d, "output", "l2err.dat"), ndmin=2)
        vacf = np.loadtxt(os.path.join(d, "output", "vacf.dat"), comments="#", ndmin=2)
        nt = np.frombuffer(open(os.path.join(d, "output", "tracer_n.bin"), "rb").read(), "<f8").copy()
        rc, _ = run(d, [os.path.join(REF, "ref_dump"), "geom"])
        assert rc == 0
        _, nature, interfacial = read_geom(os.path.join(d, "output", "geom.bin"))
        geom_in = open(os.path.join(d, "geom.in")).read() if os.path.exists(os.path.join(d, "geom.in")) else ""
        np.savez_compressed(os.path.join(HERE, name + ".npz"), lb_in=lb_in(spec["lb"]), geom_in=geom_in,
                            nature=nature, interfacial=interfacial,
                            l2err_step=l2[:, 0].astype(np.int64), l2err=l2[:, 1],
                            vacf_step=vacf[:, 0].astype(np.int64), vacf=vacf[:, 1:4].copy(),
                            tracer_n=nt, **fx)
        print(name, "steps", int(l2[-1, 0]), "mp", int(vacf[-1, 0]))


def make_fail():
    res = {}
    for name, lb in FAIL.items():
        with tempfile.TemporaryDirectory() as d:
            open(os.path.join(d, "lb.in"), "w").write(lb_in(lb))
            rc, out = run(d, [os.path.join(REF, "laboetie")])
            l2p = os.path.join(d, "output", "l2err.dat")
            nl2 = len(open(l2p).read().splitlines()) if os.path.exists(l2p) else -1
            lines = [ln.strip() for ln in out.splitlines()
                     if ln.strip() and not ln.startswith("IEEE arithmetic exceptions")]
            msg = lines[-1]
            res[name] = (lb_in(lb), rc, msg, nl2)
            print(name, rc, msg, nl2)
    np.savez_compressed(os.path.join(HERE, "failures.npz"),
                        names=np.array(list(res)), lb_in=np.array([v[0] for v in res.values()]),
                        returncode=np.array([v[1] for v in res.values()]),
                        message=np.array([v[2] for v in res.values()]),
                        l2err_lines=np.array([v[3] for v in res.values()]))


def make_geometries():
    """node%nature and node%isInterfacial for BCC (label 3) and cylinder (label 2)
    boxes of many sizes, plus the D3Q19 tables."""
    out = {}
    with tempfile.TemporaryDirectory() as d:
        cases = [("bcc", n, n, n, 3) for n in range(2, 34)] + \
                [("cyl", n, n, lz, 2) for n in range(3, 41) for lz in (1, 2)] + \
                [("slit", 3, 2, n, 1) for n in range(2, 6)]
        for kind, lx, ly, lz, label in cases:
            open(os.path.join(d, "lb.in"), "w").write(lb_in(dict(lx=lx, ly=ly, lz=lz, geometryLabel=label)))
            rc, o = run(d, [os.path.join(REF, "ref_dump"), "geom"])
            if rc != 0:
                # supercell_definition.f90:88-90 stops on an all-solid box
                assert "All nodes are solid" in o, o
                out[f"{kind}_{lx}_{ly}_{lz}"] = np.zeros((0,), np.int8)
                continue
            dims, nat, itf = read_geom(os.path.join(d, "output", "geom.bin"))
            out[f"{kind}_{lx}_{ly}_{lz}"] = np.stack([nat, itf])
        rc, o = run(d, [os.path.join(REF, "ref_dump"), "tables"])
        assert rc == 0, o
        b = open(os.path.join(d, "output", "tables.bin"), "rb").read()
        out["tables_c"] = np.frombuffer(b[:228], "<i4").reshape(19, 3).copy()
        out["tables_inv"] = np.frombuffer(b[228:304], "<i4").copy()
        out["tables_a"] = np.frombuffer(b[304:], "<f8").reshape(3, 19).copy()
    np.savez_compressed(os.path.join(HERE, "geometries.npz"), **out)
    print("geometries", len(out))


def make_collide():
    """One reference collide() call on seeded random inputs (5x4x3 box with
    random solids, tau 0.7, all force components non-zero)."""
    lx, ly, lz = 5, 4, 3
    N = lx * ly * lz
    rs = np.random.RandomState(11)
    with tempfile.TemporaryDirectory() as d:
        geom = random_geom(lx, ly, lz, 3, 0.3)
        open(os.path.join(d, "geom.in"), "w").write(geom)
        open(os.path.join(d, "lb.in"), "w").write(lb_in(dict(lx=lx, ly=ly, lz=lz, geometryLabel=0,
                                                              relaxation_time=0.7)))
        n = rs.uniform(0.001, 0.1, 19 * N)
        rho = rs.uniform(0.8, 1.2, N)
        j = rs.uniform(-0.02, 0.02, 3 * N)
        f = rs.uniform(-1e-3, 1e-3, 3 * N)
        with open(os.path.join(d, "collide_in.bin"), "wb") as fh:
            fh.write(n.astype("<f8").tobytes())
            fh.write(np.concatenate([rho, j, f]).astype("<f8").tobytes())
        rc, o = run(d, [os.path.join(REF, "ref_dump"), "collide"])
        assert rc == 0, o
        nout = np.frombuffer(open(os.path.join(d, "collide_out.bin"), "rb").read(), "<f8").copy()
        rc, o = run(d, [os.path.join(REF, "ref_dump"), "geom"])
        _, nat, _ = read_geom(os.path.join(d, "output", "geom.bin"))
    np.savez_compressed(os.path.join(HERE, "collide_random.npz"), dims=np.array([lx, ly, lz], np.int32),
                        tau=0.7, nature=nat, n_in=n, rho=rho, j=j.reshape(3, N), f=f.reshape(3, N), n_out=nout)
    print("collide_random")


def main(argv):
    want = set(argv[1:])
    for name, spec in list(FULL.items()) + list(CONV.items()):
        if not want or name in want:
            make_full(name, spec)
    if not want or "failures" in want:
        make_fail()
    if not want or "geometries" in want:
        make_geometries()
    if not want or "collide_random" in want:
        make_collide()


if __name__ == "__main__":
    main(sys.argv)
