"""Benchmark of the laboetie hot path on one MI355X.

Workload (one bench "step" = one full solve, inputs resident in HBM):
BCC cell of touching solid spheres (geometryLabel = 3), lattice 128^3,
periodic, density 1, tau 1, f_ext = 1e-6 along z; D3Q19 equilibration until
max|j - j_old| <= 1e-10 (the reference's two-stage loop), then for each
adsorption free energy dF in {0, -1, -2} kT (kd = 0.01, ka = kd exp(-dF))
10000 moment-propagation steps of a tracer with Db = 0.01, giving VACF and
D_eff.  FP64 throughout.  The three tracers share the flow, so by default one
kernel propagates them together (--mp-one-at-a-time: one after the other);
either way every tracer's results are bit-identical to its own run.

metric: lattice-node updates per second over LB and MP steps together
(lx*ly*lz per step, the usual MLUPS convention), whole job.  The JSON line
also carries the roofline of the dominant kernel (algorithmic HBM bytes per
launch / measured mean launch duration, HIP events on the library's stream)
and the reference program timed on this machine's host cores.  Two measured
ceilings go beside it: the plain streaming copy ("copy_ceiling_gbs") and the
LB step's own streams without its gathers ("soa_planes_ceiling": 22 FP64
planes of the 256^3 medium's fluid nodes in and out), which "roofline_lb"
reads its traffic against as "frac_of_soa_planes_ceiling".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-fp32-shift] [--no-large]
                    [--no-batch] [--no-serving]

The same job also runs in the two FP32-storage precision forms (FP32-shift:
FP32 collision; FP32-store: FP64 collision); their throughput, rooflines and
D_eff error against the FP64 run go under "fp32_shift" and "fp32_store".  The
128^3 job's LB working set (~0.24 GB) fits the 256 MB MALL, so "roofline_lb"
is read on the 256^3 random medium (3 GB of FP64 populations), whose
shortened precision run ("precision_256": the error of every output of the
FP32 forms against FP64; "roofline_256": both kernels in all three forms)
runs outside the timed region.  The other configurations run beside it,
also outside the timed region: the 4x4x34 slit ("slit_4x4x34"), the
1024-instance batch ("batch_1024", FP64, FP32-shift and FP32-store, each
against the FP64 batch and against the plain FP64 path on a 16-instance
sample), a bounded serving stream ("serving": p50/p99 latency and MLUPS at
0.5x, 1x and 2x of saturation, FP32-shift; its saturation requests also on the
plain FP64 path, for the rate and D_eff comparison), and, in "cpu_baseline",
the reference program on a bounded sample of the job and of each of them.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from laboetie_amd import lbx  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
LB_BYTES_PER_FLUID = 19 * 8 * 2                 # read + write of 19 FP64 populations
LB_FLUX_BYTES_PER_FLUID = 3 * 8 * 2             # read + write of the flux kept for the FP64 stop test
MP_COEF_BYTES_PER_FLUID = 19 * 8 + 8            # coefficient source (n, rho), shared by the tracers of a batch
MP_P_BYTES_PER_FLUID = 2 * 3 * 8                # P read + write, per tracer
MP_BYTES_PER_ADS_INTERFACIAL = 2 * 3 * 8        # Pa read + write, per adsorbing tracer
LB32_BYTES_PER_FLUID = 19 * 4 * 2               # FP32-shift: read + write of 19 FP32 deviations
MP32_COEF_BYTES_PER_FLUID = 19 * 4 + 4          # FP32-shift: coefficient source in FP32

WORKLOAD = dict(label=3, dims=(128, 128, 128), tau=1.0, rho0=1.0, fext=(0.0, 0.0, 1e-6), target=1e-10,
                Db=0.01, kd=0.01, dF=(0.0, -1.0, -2.0), mp_steps=10000)
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_bcc128.txt")


def tracer_params(w):
    """ka, kd per dF (ka = kd exp(-dF))."""
    ka = [lbx.ka_kd_from_dF(dF, w["kd"])[0] for dF in w["dF"]]
    return ka, [w["kd"]] * len(ka)


def solve(L, w, batch=True):
    """One full solve; returns (lb_steps, MP tracer-steps, D_eff per dF).

    With batch (the default) the tracers of the dF sweep, which share the
    flow and Db, are propagated together by one kernel (each tracer's results
    are bit-identical to its own run); otherwise one after the other."""
    L.init_rest(w["rho0"])
    r = L.equilibrate(w["fext"], w["target"], trace_cap=1)
    ka, kd = tracer_params(w)
    if batch:
        L.mp_init_batch(w["fext"], w["Db"], ka, kd)
        m = L.mp_run(w["mp_steps"])
        mp_steps = m["steps"] * len(ka)
        deffs = L.mp_deff().tolist()
    else:
        deffs, mp_steps = [], 0
        for a, d in zip(ka, kd):
            L.mp_init(w["fext"], w["Db"], a, d)
            m = L.mp_run(w["mp_steps"])
            mp_steps += m["steps"]
            deffs.append(L.mp_deff().tolist())
    L.sync()
    return r["steps"], mp_steps, deffs


def fp32_leg(w, nat, args, nf, nfi, prec=lbx.PREC_FP32_SHIFT):
    """The same job in an FP32-storage precision form (populations as FP32
    deviations from a0*rho0, FP32 MP probabilities; FP64 moments, P, VACF and
    D_eff), timed the same way: FP32-shift collides in FP32, FP32-store in
    FP64."""
    with lbx.Lattice(w["dims"], nat, w["tau"], precision=prec) as L:
        for _ in range(args.warmup):
            solve(L, w, not args.mp_one_at_a_time)
        L.timing(True)
        L.sync()
        t0 = time.perf_counter()
        results = [solve(L, w, not args.mp_one_at_a_time) for _ in range(args.steps)]
        L.sync()
        dt = time.perf_counter() - t0
        lb_ms, lb_n = L.kernel_time(lbx.KERNEL_LB)
        mp_ms, mp_n = L.kernel_time(lbx.KERNEL_MP)
    lx, ly, lz = w["dims"]
    N = lx * ly * lz
    lb_steps, mp_steps, deffs = results[-1]
    T = 1 if args.mp_one_at_a_time else len(w["dF"])
    lb_gbs = LB32_BYTES_PER_FLUID * nf / (lb_ms / lb_n / 1e3) / 1e9
    mp_bytes = (MP32_COEF_BYTES_PER_FLUID * nf + MP_P_BYTES_PER_FLUID * nf * T
                + MP_BYTES_PER_ADS_INTERFACIAL * nfi * T)
    mp_gbs = mp_bytes / (mp_ms / mp_n / 1e3) / 1e9
    print(f"bench: fp32 leg (precision {prec}) done", file=sys.stderr, flush=True)
    return {"value": round(N * sum(r[0] + r[1] for r in results) / dt / 1e6, 2), "unit": "MLUPS",
            "ms_per_step": round(dt / args.steps * 1e3, 3), "lb_steps": lb_steps, "mp_steps": mp_steps,
            "roofline_lb": {"achieved": round(lb_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(lb_gbs / HBM_PEAK_GBS, 4), "bytes_per_fluid_node": LB32_BYTES_PER_FLUID,
                            "avg_launch_us": round(lb_ms / lb_n * 1e3, 2)},
            "roofline_mp": {"achieved": round(mp_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(mp_gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": round(mp_bytes),
                            "avg_launch_us": round(mp_ms / mp_n * 1e3, 2)},
            "deff": deffs}


LARGE = dict(dims=(256, 256, 256), sigma=8.0, porosity=0.6, seed=20261202, tau=0.8, fext=(1e-6, 0.0, 0.0),
             Db=0.01, kd=0.01, dF=-1.0, lb_steps=2000, mp_steps=1000)


def large_leg():
    """The precision configuration, shortened: the 256^3 Gaussian random
    medium (correlation length 8, porosity 0.6), tau 0.8, F = 1e-6 along x;
    LARGE["lb_steps"] LB steps from rest (the force on from t = 4) and
    LARGE["mp_steps"] MP steps of one tracer (dF = -1 kT), in FP64,
    FP32-shift and FP32-store (laboetie_amd.precision.error_table): the error
    of every output of the FP32 forms against FP64, and the two kernels'
    rooflines on a working set far above the 256 MB MALL (kernel times from
    HIP events)."""
    from laboetie_amd import precision

    g = LARGE
    nat = lbx.geometry_random(*g["dims"], g["sigma"], g["porosity"], g["seed"])
    lx, ly, lz = g["dims"]
    itf = lbx.detect_interfacial(lx, ly, lz, nat)
    nfi = int(((nat == 0) & (itf == 1)).sum())
    tab = precision.error_table(lx, g["lb_steps"], g["mp_steps"], g["seed"], nat=nat, keep_runs=True)
    runs = tab.pop("_runs")
    out = {"workload": tab["workload"],
           "error_table": {k: v for k, v in tab.items() if k.startswith(("errors_vs", "deff", "permeability"))}}
    for name, lbb, coef in (("fp64", LB_BYTES_PER_FLUID, MP_COEF_BYTES_PER_FLUID),
                            ("fp32_shift", LB32_BYTES_PER_FLUID, MP32_COEF_BYTES_PER_FLUID),
                            ("fp32_store", LB32_BYTES_PER_FLUID, MP32_COEF_BYTES_PER_FLUID)):
        r = runs[name]
        nf = r["fluid"]
        lb_s, mp_s = r["lb_kernel_us"] / 1e6, r["mp_kernel_us"] / 1e6
        mp_bytes = (coef + MP_P_BYTES_PER_FLUID) * nf + MP_BYTES_PER_ADS_INTERFACIAL * nfi
        out[name] = {
            "fluid_nodes": nf,
            "roofline_lb": {"bound": "hbm", "achieved": round(lbb * nf / lb_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(lbb * nf / lb_s / 1e9 / HBM_PEAK_GBS, 4),
                            "kernel": "k_lb_step", "bytes_per_launch": lbb * nf, "avg_launch_us": round(lb_s * 1e6, 1)},
            "roofline_mp": {"bound": "hbm", "achieved": round(mp_bytes / mp_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(mp_bytes / mp_s / 1e9 / HBM_PEAK_GBS, 4),
                            "kernel": "k_mp_step", "bytes_per_launch": round(mp_bytes),
                            "avg_launch_us": round(mp_s * 1e6, 1)},
            "lb_mlups": round(lx * ly * lz / lb_s / 1e6, 1), "mp_mlups": round(lx * ly * lz / mp_s / 1e6, 1)}
    print("bench: 256^3 leg done", file=sys.stderr, flush=True)
    return out


def serving_leg(n_requests=16, seed=20261202):
    """The serving configuration, bounded: n_requests requests per rate
    (N^3 with N log-uniform in [32, 192]; slit, cylinder along z, BCC spheres
    or Gaussian random medium of porosity U[0.4, 0.8]; F log-uniform in
    [1e-7, 1e-5]; target 1e-9, LB capped at 10000 steps; 2000 adsorbing MP
    steps, dF U[-3, 0] kT) in the FP32-shift form, continuous batching with
    HBM-budget admission.  Saturation throughput from a closed batch of the
    first half of the requests; then Poisson arrivals at 0.5x, 1x and 2x of
    it, latency from arrival to result."""
    from laboetie_amd import serving

    rng = np.random.default_rng(seed)
    reqs = [serving.serving_request(rng) for _ in range(n_requests)]
    prec = lbx.PREC_FP32_SHIFT
    cap_nodes, cap_box = 48 << 20, 64 << 20
    sat = serving.serve(reqs[:n_requests // 2], np.zeros(n_requests // 2), prec, cap_nodes, cap_box)
    rate0 = (n_requests // 2) / sat["wall_s"]
    out = {"workload": f"{n_requests} requests per rate: N^3 (N log-uniform [32,192]), slit/cylinder/BCC/random, "
                       "F log-uniform [1e-7,1e-5], target 1e-9 (cap 10000), 2000 MP steps, dF U[-3,0], FP32-shift, "
                       "48 M-node HBM budget",
           "saturation_req_per_s": round(rate0, 3), "saturation_mlups": round(sat["mlups"], 1)}
    # the same requests on the plain FP64 path, one after the other: the aggregate rate the
    # batched FP32-shift form is compared with, and each request's D_eff against which its
    # served D_eff is read (each form stopped by its own rule)
    plain = serving.run_one_at_a_time(reqs[:n_requests // 2], lbx.PREC_FP64)
    d64 = np.array(plain["deff"])
    d32 = np.array(sat["deff"])
    err = np.abs(d32 - d64).max(axis=1) / np.abs(d64).max(axis=1)
    out["plain_fp64_one_at_a_time"] = {"requests": plain["requests"], "mlups": round(plain["mlups"], 1),
                                       "wall_s": round(plain["wall_s"], 2)}
    out["saturation_vs_plain_fp64"] = round(sat["mlups"] / plain["mlups"], 3)
    out["deff_rel_error_vs_plain_fp64"] = {"requests": int(err.size), "median": float(f"{np.median(err):.3e}"),
                                           "max": float(f"{err.max():.3e}"),
                                           "count_above_1e-5": int((err > 1e-5).sum()),
                                           "kinds": [r["kind"] for r in reqs[:n_requests // 2]],
                                           "per_request": [float(f"{e:.3e}") for e in err]}
    for f in (0.5, 1.0, 2.0):
        rate = f * rate0
        arr = np.cumsum(np.random.default_rng(seed + int(f * 1000)).exponential(1.0 / rate, n_requests))
        s = serving.serve(reqs, arr, prec, cap_nodes, cap_box)
        lat = s["latency"][np.isfinite(s["latency"])]
        out[f"rate_{f:g}x"] = {"req_per_s": round(rate, 3), "p50_s": round(float(np.percentile(lat, 50)), 3),
                               "p99_s": round(float(np.percentile(lat, 99)), 3), "mlups": round(s["mlups"], 1),
                               "wall_s": round(s["wall_s"], 2), "failed": int((s["status"] != 0).sum())}
        print(f"bench: serving rate {f:g}x done", file=sys.stderr, flush=True)
    return out


def slit_leg():
    """The CPU-runnable configuration on the GPU: slit 4x4x34 (walls at z = 1
    and 34), tau 1, F = 1e-5 along x, equilibrated to max|dj| <= 1e-10 (the
    step count the reference's own loop takes), then 5000 MP steps of a
    neutral tracer (Db = 0.01).  A 512-fluid-node lattice is launch-bound:
    the figure is the whole job's wall rate."""
    dims = (4, 4, 34)
    nat = lbx.build_geometry(1, *dims)
    with lbx.Lattice(dims, nat, 1.0) as L:
        L.sync()
        t0 = time.perf_counter()
        L.init_rest(1.0)
        r = L.equilibrate((1e-5, 0.0, 0.0), 1e-10, trace_cap=1)
        L.mp_init((1e-5, 0.0, 0.0), 0.01)
        m = L.mp_run(5000)
        L.sync()
        dt = time.perf_counter() - t0
        d = L.mp_deff()
    N = int(np.prod(dims))
    return {"lattice": list(dims), "lb_steps": r["steps"], "mp_steps": m["steps"],
            "mlups": round(N * (r["steps"] + m["steps"]) / dt / 1e6, 2), "seconds": round(dt, 3),
            "deff": d.tolist()}


def batch_leg(n=1024):
    """The batching configuration: n chromatography-style 2-D instances
    (1x50x50 / 1x100x100, 1-8 disks of radius 5, F = 1e-5 along y, dF
    uniform in [-3, 0] kT, 2000 MP steps each) admitted at once into the
    batch engine and run to completion, FP64 and FP32-shift."""
    from laboetie_amd import serving

    reqs = serving.chromatography_requests(n)
    out = {"workload": f"{n} chromatography 2-D instances (1x50x50 / 1x100x100, 1-8 disks r=5), tau 1, "
                       "F=1e-5 along y, target 1e-10 (cap 20000), 2000 MP steps, dF U[-3,0]"}
    deffs = {}
    conv = None
    for prec, name in ((lbx.PREC_FP64, "fp64"), (lbx.PREC_FP32_SHIFT, "fp32_shift"),
                       (lbx.PREC_FP32_STORE, "fp32_store")):
        r = serving.run_batch(reqs, prec)
        deffs[name] = np.array([x["deff"] for x in r["results"]])
        if conv is None:  # instances whose FP64 run converged (the others stop at the step cap)
            conv = np.array([x["exit_reason"] == lbx.EXIT_CONVERGED for x in r["results"]])
            steps64 = [x["lb_steps"] for x in r["results"]]
        out[name] = {"mlups": round(r["mlups"], 1), "mflups": round(r["mflups"], 1), "wall_s": round(r["wall_s"], 2),
                     "lb_kernel_mlups": round(r["lb_kernel_mlups"], 1), "mp_kernel_mlups": round(r["mp_kernel_mlups"], 1)}
    # precision: each instance run in the FP32 forms for exactly the LB steps its FP64 run took
    # (no stop test), so that the two differ by their arithmetic only and not by where each
    # form's stop rule ends a still-settling flow
    for prec, name in ((lbx.PREC_FP32_SHIFT, "fp32_shift"), (lbx.PREC_FP32_STORE, "fp32_store")):
        r = serving.run_batch(reqs, prec, lb_steps=steps64)
        ds = np.array([x["deff"] for x in r["results"]])
        err = np.abs(ds - deffs["fp64"]).max(axis=1) / np.abs(deffs["fp64"]).max(axis=1)
        out[f"deff_rel_error_{name}_same_lb_steps"] = {
            "median": float(f"{np.median(err):.3e}"), "max": float(f"{err.max():.3e}"),
            "count_above_1e-5": int((err > 1e-5).sum())}
    for name in ("fp32_shift", "fp32_store"):
        err = np.abs(deffs[name] - deffs["fp64"]).max(axis=1) / np.abs(deffs["fp64"]).max(axis=1)
        out[f"deff_rel_error_{name}"] = {"median": float(f"{np.median(err):.3e}"), "max": float(f"{err.max():.3e}"),
                                         "count_above_1e-5": int((err > 1e-5).sum())}
        ec = err[conv]
        out[f"deff_rel_error_{name}_where_fp64_converged"] = {
            "instances": int(conv.sum()), "median": float(f"{np.median(ec):.3e}"), "max": float(f"{ec.max():.3e}"),
            "count_above_1e-5": int((ec > 1e-5).sum())}
        out[f"{name}_vs_fp64"] = round(out[name]["mlups"] / out["fp64"]["mlups"], 3)
    # the plain FP64 path, one instance after the other, on a sample of 16 instances: its
    # aggregate rate against the batched forms'
    plain = serving.run_one_at_a_time(reqs[:16], lbx.PREC_FP64)
    out["plain_fp64_one_at_a_time"] = {"instances": plain["requests"], "mlups": round(plain["mlups"], 1),
                                       "wall_s": round(plain["wall_s"], 2)}
    for name in ("fp64", "fp32_shift", "fp32_store"):
        out[f"{name}_vs_plain_fp64"] = round(out[name]["mlups"] / plain["mlups"], 2)
    print("bench: batch leg done", file=sys.stderr, flush=True)
    return out


def read_traffic():
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), as committed in profiles/traffic_bcc128.txt lines
    `<kernel> <bytes_per_launch>`; None when absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return {}
    out = {}
    for ln in open(TRAFFIC_FILE):
        ln = ln.split("#")[0].split()
        if len(ln) == 2:
            out[ln[0]] = float(ln[1])
    return out


def lb_in_text(w, fext, mp_steps, ka, print_frequency=1):
    lx, ly, lz = w["dims"]
    return (f"lx = {lx}\nly = {ly}\nlz = {lz}\ngeometryLabel = {w.get('label', 0)}\n"
            f"relaxation_time = {w['tau']}\ninitialSolventDensity = {w['rho0']}\n"
            f"f_ext = {fext[0]} {fext[1]} {fext[2]}\ntarget_error = {w['target']}\n"
            f"print_frequency = {print_frequency}\nmaximum_moment_propagation_steps = {mp_steps}\n"
            f"tracer_Db = {w['Db']}\ntracer_ka = {ka}\ntracer_kd = {w['kd']}\n")


def _run_on_pty(exe, d, threads, stop, max_wait):
    """Run the reference in directory d with its stdout on a pseudo-terminal
    (so the Fortran runtime writes each line as it is printed) and return the
    (arrival time, line) pairs; stop(lines) ends the run early."""
    import pty
    import select

    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    master, slave = pty.openpty()
    p = subprocess.Popen([exe], cwd=d, env=env, stdout=slave, stderr=slave, stdin=subprocess.DEVNULL)
    os.close(slave)
    lines, buf, t0 = [], b"", time.time()
    try:
        while time.time() - t0 < max_wait:
            r, _, _ = select.select([master], [], [], 0.5)
            if r:
                try:
                    chunk = os.read(master, 65536)
                except OSError:
                    break
                if not chunk:
                    break
                buf += chunk
                now = time.time()
                while b"\n" in buf:
                    ln, buf = buf.split(b"\n", 1)
                    lines.append((now, ln.decode(errors="replace").strip()))
                if stop(lines):
                    break
            elif p.poll() is not None:
                break
    finally:
        if p.poll() is None:
            p.kill()
        p.wait()
        os.close(master)
    return lines


def _lb_steps_per_second(exe, d, threads, window, min_steps=0):
    """The equilibration prints `t l2err (target ...)` every step with
    print_frequency = 1 (equilibration.f90:149); time those lines over at
    least `window` seconds and `min_steps` steps."""
    def steps(lines):
        return [(t, int(l.split()[0])) for t, l in lines if l.endswith(")") and "(target" in l]

    def stop(lines):
        s = steps(lines)
        return len(s) >= 2 and s[-1][0] - s[0][0] >= window and s[-1][1] - s[0][1] >= min_steps

    s = steps(_run_on_pty(exe, d, threads, stop, max_wait=600))
    if len(s) < 2 or s[-1][1] == s[0][1]:
        return None
    return (s[-1][1] - s[0][1]) / (s[-1][0] - s[0][0])


def _mp_steps_per_second(exe, d, threads, nsteps):
    """Time nsteps moment-propagation steps: from the `0 VACF(0)` line printed
    at the end of moment_propagation::init (module_moment_propagation.f90:139)
    to the closing status line of print_tail (module_io.f90:64)."""
    lines = _run_on_pty(exe, d, threads, lambda ls: any("status OK" in l for _, l in ls), max_wait=600)
    t_init = next((t for t, l in lines if l.split()[:1] == ["0"] and len(l.split()) == 4), None)
    t_end = next((t for t, l in lines if "status OK" in l), None)
    if t_init is None or t_end is None or t_end <= t_init:
        return None
    return nsteps / (t_end - t_init)


def geom_in_text(dims, nat):
    """geom.in for geometryLabel 0: one 1-based `i j k` line per solid node
    (module_geometry.f90:380-426)."""
    lx, ly, lz = dims
    z, y, x = np.nonzero(np.asarray(nat).reshape(lz, ly, lx) != 0)
    ijk = np.stack([x + 1, y + 1, z + 1], axis=1)
    return "\n".join(f"{a} {b} {c}" for a, b, c in ijk.tolist()) + "\n"


def _ref_rates(exe, threads, w, nat, lb_window, mp_sample):
    """The reference's LB steps/s (per-step stdout lines over lb_window
    seconds) and MP steps/s (mp_sample steps on the flow at rest) on one
    lattice; geometryLabel 0 lattices get their geom.in from nat."""
    geom = geom_in_text(w["dims"], nat) if w.get("label", 0) == 0 else None

    def setup(d, text):
        open(os.path.join(d, "lb.in"), "w").write(text)
        if geom is not None:
            open(os.path.join(d, "geom.in"), "w").write(geom)

    with tempfile.TemporaryDirectory() as d:
        setup(d, lb_in_text(w, w["fext"], 0, 0.0))
        lb_rate = _lb_steps_per_second(exe, d, threads, lb_window)
    with tempfile.TemporaryDirectory() as d:
        setup(d, lb_in_text(w, (0.0, 0.0, 0.0), mp_sample, w.get("ka", 0.0)))
        mp_rate = _mp_steps_per_second(exe, d, threads, mp_sample)
    return lb_rate, mp_rate


def dropin_leg(w, mp_steps=2000):
    """The drop-in binary (laboetie_amd/fortran: the reference program with
    its hot-path modules replaced by the lbx shims) on the bench lattice with
    one tracer (dF = -1 kT): its moment-propagation rate, timed like the
    reference's (from the `0 VACF(0)` line to the closing status line, on a
    pseudo-terminal), beside lbx_mp_run's rate on the same lattice and
    tracer in this process.  The drop-in writes vacf.dat every step, as the
    reference does."""
    exe = os.path.join(HERE, "laboetie_amd", "fortran", "_build", "laboetie_gpu")
    if not os.path.exists(exe):
        return None
    ka = w["kd"] * float(np.exp(1.0))
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "lb.in"), "w").write(lb_in_text(w, w["fext"], mp_steps, ka, print_frequency=1000))
        rate = _mp_steps_per_second(exe, d, 1, mp_steps)
    lx, ly, lz = w["dims"]
    nat = lbx.build_geometry(w["label"], lx, ly, lz)
    with lbx.Lattice(w["dims"], nat, w["tau"]) as L:
        L.init_rest(w["rho0"])
        L.equilibrate(w["fext"], w["target"], trace_cap=1)
        L.mp_init(w["fext"], w["Db"], ka, w["kd"])
        L.mp_run(50)
        L.sync()
        t0 = time.perf_counter()
        L.mp_run(mp_steps)
        L.sync()
        lib_rate = mp_steps / (time.perf_counter() - t0)
    print("bench: drop-in leg done", file=sys.stderr, flush=True)
    return {"workload": f"BCC 128^3, one tracer (dF=-1 kT), {mp_steps} MP steps",
            "dropin_mp_steps_per_s": round(rate, 1) if rate else None,
            "lbx_mp_run_steps_per_s": round(lib_rate, 1),
            "dropin_vs_lbx_mp_run": round(rate / lib_rate, 3) if rate else None}


def cpu_configs(exe, threads):
    """The reference on a bounded sample of each other configuration, on the
    same host cores: node-update rates of its LB and MP phases (lattice nodes
    x steps / s), which the GPU figures of those configurations are read
    beside (the slit and 256^3 legs of this bench, the batch and serving
    tools).  Returns {config: {...}}."""
    from laboetie_amd import serving

    out = {}
    # slit 4x4x34, tau 1, F = 1e-5 along x, neutral tracer
    w1 = dict(label=1, dims=(4, 4, 34), tau=1.0, rho0=1.0, fext=(1e-5, 0.0, 0.0), target=1e-10, Db=0.01,
              kd=0.0, ka=0.0)
    jobs = [("slit_4x4x34", w1, None, 3.0, 2000)]
    # the 256^3 Gaussian random medium (tau 0.8, F = 1e-6 along x, dF = -1 kT), sampled by its
    # 128^3 corner: the reference spends minutes reading a 256^3 geom.in and setting it up,
    # and its per-node step cost does not depend on the lattice size
    g = LARGE
    lx3, ly3, lz3 = g["dims"]
    nat3 = lbx.geometry_random(*g["dims"], g["sigma"], g["porosity"], g["seed"]).reshape(lz3, ly3, lx3)
    nat3 = np.ascontiguousarray(nat3[:lz3 // 2, :ly3 // 2, :lx3 // 2]).ravel()
    w3 = dict(label=0, dims=(lx3 // 2, ly3 // 2, lz3 // 2), tau=g["tau"], rho0=1.0, fext=g["fext"], target=1e-10,
              Db=g["Db"], kd=g["kd"], ka=lbx.ka_kd_from_dF(g["dF"], g["kd"])[0])
    jobs.append(("grf_256_corner_128", w3, nat3, 8.0, 10))
    # chromatography instances: one 1x50x50 and one 1x100x100 of the batch workload
    for r in serving.chromatography_requests(2):
        wr = dict(label=0, dims=r["dims"], tau=r["tau"], rho0=r["rho0"], fext=r["fext"], target=r["target"],
                  Db=r["Db"], kd=r["kd"], ka=r["ka"])
        jobs.append((f"chromatography_1x{r['dims'][1]}x{r['dims'][2]}", wr, r["nature"], 2.0, 500))
    # two requests of the serving stream (its first draws)
    rng = np.random.default_rng(20261202)
    for k in range(2):
        r = serving.serving_request(rng)
        wr = dict(label=0, dims=r["dims"], tau=r["tau"], rho0=r["rho0"], fext=r["fext"], target=r["target"],
                  Db=r["Db"], kd=r["kd"], ka=r["ka"])
        jobs.append((f"serving_{r['kind']}_{r['dims'][0]}", wr, r["nature"], 5.0, 20))
    for name, w, nat, window, mp_sample in jobs:
        print(f"cpu_baseline: reference on {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        lb_rate, mp_rate = _ref_rates(exe, threads, w, nat, window, mp_sample)
        N = int(np.prod(w["dims"]))
        out[name] = {"lattice": list(w["dims"]),
                     "lb_mlups": round(N * lb_rate / 1e6, 4) if lb_rate else None,
                     "mp_mlups": round(N * mp_rate / 1e6, 4) if mp_rate else None,
                     "sample": f"LB steps over {window:.0f} s, {mp_sample} MP steps",
                     "wall_s": round(time.perf_counter() - t0, 1)}
    return out


def host_threads():
    """The host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where the environment sets it (one GPU's share of a
    shared host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(w, lb_steps, mp_steps_total, window=5.0, lb_sample=200, mp_sample=200, other_configs=True):
    """The reference program (oracle/_ref/laboetie_fast: flang -O3 -fopenmp
    -funroll-loops, the reference Makefile's flags) on this machine's host
    cores (host_threads()), timed on a bounded sample of the same workload:
    the LB step rate from its per-step stdout lines over at least lb_sample
    steps of the 128^3 BCC run, the MP step rate from a run of mp_sample
    steps of the dF = -1 tracer on a flow at rest (f_ext = 0, so
    equilibration ends at t = 4; the MP cost does not depend on the flow
    values).  Whole-job MLUPS = N (LB + MP steps) / (LB/rate + MP/rate)."""
    exe = os.path.join(HERE, "oracle", "_ref", "laboetie_fast")
    if not os.path.exists(exe):
        return None
    threads = host_threads()
    N = int(np.prod(w["dims"]))
    print("cpu_baseline: timing the reference LB phase", file=sys.stderr, flush=True)
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "lb.in"), "w").write(lb_in_text(w, w["fext"], 0, 0.0))
        lb_rate = _lb_steps_per_second(exe, d, threads, window, lb_sample)
    print(f"cpu_baseline: LB {lb_rate} steps/s; timing the reference MP phase", file=sys.stderr, flush=True)
    with tempfile.TemporaryDirectory() as d:
        ka = w["kd"] * np.exp(1.0)
        open(os.path.join(d, "lb.in"), "w").write(lb_in_text(w, (0.0, 0.0, 0.0), mp_sample, ka))
        mp_rate = _mp_steps_per_second(exe, d, threads, mp_sample)
    print(f"cpu_baseline: MP {mp_rate} steps/s", file=sys.stderr, flush=True)
    if not lb_rate or not mp_rate:
        return None
    t_job = lb_steps / lb_rate + mp_steps_total / mp_rate
    print("cpu_baseline: the reference's LB phase on one core", file=sys.stderr, flush=True)
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "lb.in"), "w").write(lb_in_text(w, w["fext"], 0, 0.0))
        lb_rate1 = _lb_steps_per_second(exe, d, 1, 12.0)
    configs = cpu_configs(exe, threads) if other_configs else None
    return {"value": round(N * (lb_steps + mp_steps_total) / t_job / 1e6, 4), "unit": "MLUPS",
            "cores": threads, "kind": "reference",
            "sample": (f"reference laboetie (flang -O3 -fopenmp -funroll-loops), OMP_NUM_THREADS={threads} (the "
                       f"host cores this process may use): LB {lb_rate:.3f} steps/s over >= {lb_sample} steps of "
                       f"the 128^3 BCC run, MP {mp_rate:.3f} steps/s over {mp_sample} steps (dF=-1 tracer), "
                       f"combined over this job's {lb_steps} LB + {mp_steps_total} MP steps"),
            "lb_mlups": round(N * lb_rate / 1e6, 4), "mp_mlups": round(N * mp_rate / 1e6, 4),
            "lb_mlups_1_core": round(N * lb_rate1 / 1e6, 4) if lb_rate1 else None,
            "other_configs": configs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed full solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32-shift", action="store_true", help="skip the FP32-shift and FP32-store legs")
    ap.add_argument("--mp-one-at-a-time", action="store_true",
                    help="propagate the dF tracers one after the other instead of together")
    ap.add_argument("--no-large", action="store_true", help="skip the 256^3 roofline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 1024-instance batch leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (Poisson stream) leg")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the Fortran drop-in's MP timing (a child process: profiled runs leave it out)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain-copy bandwidth probe (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        print(json.dumps({"copy_gbs": round(lbx.copy_bandwidth(1 << 30, 20), 1), "bytes_each_way": 1 << 30}))
        return
    if args.gpus != 1:
        print(f"bench.py: the bench runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    w = WORKLOAD
    lx, ly, lz = w["dims"]
    N = lx * ly * lz
    nat = lbx.build_geometry(w["label"], lx, ly, lz)
    itf = lbx.detect_interfacial(lx, ly, lz, nat)
    nfi = int(((nat == 0) & (itf == 1)).sum())
    with lbx.Lattice(w["dims"], nat, w["tau"], precision=lbx.PREC_FP64) as L:
        nf = L.fluid_nodes
        for _ in range(args.warmup):
            solve(L, w, not args.mp_one_at_a_time)
            print("bench: warmup solve done", file=sys.stderr, flush=True)
        L.timing(True)
        L.sync()
        t0 = time.perf_counter()
        results = [solve(L, w, not args.mp_one_at_a_time) for _ in range(args.steps)]
        L.sync()
        dt = time.perf_counter() - t0
        lb_ms, lb_n = L.kernel_time(lbx.KERNEL_LB)
        mp_ms, mp_n = L.kernel_time(lbx.KERNEL_MP)
        # Darcy permeability of the converged flow, k = nu <j_z>_box / F_z (lattice units)
        _, _, jf = L.get_state(populations=False)
        permeability = (w["tau"] - 0.5) / 3.0 * float(jf[2][nat == 0].sum()) / N / w["fext"][2]
        del jf
    shift = None if args.no_fp32_shift else fp32_leg(w, nat, args, nf, nfi, lbx.PREC_FP32_SHIFT)
    store = None if args.no_fp32_shift else fp32_leg(w, nat, args, nf, nfi, lbx.PREC_FP32_STORE)
    large = None if args.no_large else large_leg()
    slit = slit_leg()
    batch = None if args.no_batch else batch_leg()
    served = None if args.no_serving else serving_leg()
    dropin = None if args.no_dropin else dropin_leg(w)
    copy_gbs = lbx.copy_bandwidth(2 << 30, 20)  # measured ceiling, outside the timed region
    # the ceiling of the LB step's own streams (19 population planes and 3 flux planes in and
    # out, no gathers) at the 256^3 medium's fluid-node count
    planes_nodes = large["fp64"]["fluid_nodes"] if large is not None else 10066330
    planes_gbs = lbx.planes_bandwidth(22, planes_nodes, 20)
    lb_steps, mp_steps, deffs = results[-1]
    updates = N * sum(r[0] + r[1] for r in results)
    value = updates / dt / 1e6
    lb_avg_s = lb_ms / lb_n / 1e3
    mp_avg_s = mp_ms / mp_n / 1e3
    lb_bytes = LB_BYTES_PER_FLUID * nf
    # adsorption is on for a tracer when K = ka/kd > eps (module_moment_propagation.f90:46-56)
    ka, kd = tracer_params(w)
    ads = [kd_ > np.finfo(float).eps and ka_ / kd_ > np.finfo(float).eps for ka_, kd_ in zip(ka, kd)]
    per_launch = 1 if args.mp_one_at_a_time else len(ka)  # tracers advanced by one launch
    n_ads = sum(ads) / len(ads) * per_launch
    mp_bytes = (MP_COEF_BYTES_PER_FLUID * nf + MP_P_BYTES_PER_FLUID * nf * per_launch
                + MP_BYTES_PER_ADS_INTERFACIAL * nfi * n_ads)
    lb_gbs = lb_bytes / lb_avg_s / 1e9
    mp_gbs = mp_bytes / mp_avg_s / 1e9
    traffic = read_traffic()
    dominant = "lb_step" if lb_ms >= mp_ms else "mp_step"
    roof_lb = {"bound": "hbm", "achieved": round(lb_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(lb_gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get("k_lb_step"),
               "kernel": "k_lb_step", "bytes_per_launch": lb_bytes, "avg_launch_us": round(lb_avg_s * 1e6, 2)}
    # the same launch priced per tracer-step, as the reference reads its coefficients once per
    # tracer: 208 B per fluid node and tracer (+48 B Pa at adsorbing interfacial nodes)
    mp_bytes_tr = (MP_COEF_BYTES_PER_FLUID + MP_P_BYTES_PER_FLUID) * nf * per_launch \
        + MP_BYTES_PER_ADS_INTERFACIAL * nfi * n_ads
    roof_mp = {"bound": "hbm", "achieved": round(mp_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(mp_gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get("k_mp_step"),
               "kernel": "k_mp_step", "tracers_per_launch": per_launch, "bytes_per_launch": round(mp_bytes),
               "avg_launch_us": round(mp_avg_s * 1e6, 2),
               "per_tracer_step_bytes": round(mp_bytes_tr),
               "per_tracer_step_achieved": round(mp_bytes_tr / mp_avg_s / 1e9, 1)}
    line = {
        "metric": "D3Q19+MP MLUPS & HBM roofline %, FP64 vs FP32-shift; p50/p99 solve latency",
        "value": round(value, 2), "unit": "MLUPS", "mflups": round(value * nf / N, 2),
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "BCC touching spheres 128^3 (geometryLabel 3), FP64: equilibrate to "
                               "max|dj|<=1e-10 then 10000 MP steps of 3 adsorbing tracers (dF 0,-1,-2 kT) "
                               + ("one after the other" if args.mp_one_at_a_time else "propagated together"),
                   "lattice": [lx, ly, lz], "fluid_nodes": nf, "lb_steps": lb_steps, "mp_steps": mp_steps,
                   "precision": "fp64"},
        "roofline": roof_lb if dominant == "lb_step" else roof_mp,
        # the LB kernel's roofline is read on the 256^3 medium: the 128^3 job's LB
        # working set (~0.24 GB) sits in the 256 MB MALL
        "roofline_lb": (dict(large["fp64"]["roofline_lb"], workload="256^3 random medium (FP64)",
                             # populations and the flux kept for the stop test (304 + 48 B per
                             # fluid node) against the same streams without gathers
                             frac_of_soa_planes_ceiling=round(
                                 (LB_BYTES_PER_FLUID + LB_FLUX_BYTES_PER_FLUID) * planes_nodes
                                 / (large["fp64"]["roofline_lb"]["avg_launch_us"] * 1e-6) / 1e9 / planes_gbs, 4))
                        if large is not None else None),
        "roofline_lb_bcc128_mall_resident": roof_lb, "roofline_mp": roof_mp,
        "copy_ceiling_gbs": round(copy_gbs, 1),
        "soa_planes_ceiling": {"planes": 22, "nodes": planes_nodes, "gbs": round(planes_gbs, 1)},
        "lb_mlups": round(N * lb_steps * args.steps / (lb_ms / 1e3) / 1e6, 1),
        "mp_mlups": round(N * mp_steps * args.steps / (mp_ms / 1e3) / 1e6, 1),
        "mp_tracers_together": per_launch,
        "deff": deffs,
        "permeability": permeability,
    }
    for name, leg in (("fp32_shift", shift), ("fp32_store", store)):
        if leg is not None:
            d64 = np.array(deffs)
            d32 = np.array(leg.pop("deff"))
            leg["deff_rel_error_vs_fp64"] = float(f"{np.abs(d32 - d64).max() / np.abs(d64).max():.3e}")
            leg["vs_fp64_value"] = round(leg["value"] / value, 3)
            line[name] = leg
    line["slit_4x4x34"] = slit
    if batch is not None:
        line["batch_1024"] = batch
    if large is not None:
        line["precision_256"] = large.pop("error_table")
        line["roofline_256"] = large
    if served is not None:
        line["serving"] = served
    if dropin is not None:
        line["dropin_mp"] = dropin
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, lb_steps, mp_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
