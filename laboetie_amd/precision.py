"""FP64 vs the FP32-storage forms on one workload: per-output error table and
throughput.  FP32-shift stores populations as FP32 deviations g = n - a0 rho0
and collides in FP32; FP32-store keeps the same storage and collides in FP64.

Default workload: a 256^3 Gaussian random porous medium (correlation length
8 nodes, seed 20261202, thresholded to porosity 0.6, every solid cluster
kept), tau = 0.8, f_ext = 1e-6 along x, 10000 flow steps from rest (the
force switched on after the stage-1 convergence at t = 3), then 5000
moment-propagation steps of a tracer with Db = 0.01 adsorbing at
dF = -1 kT (kd = 0.01, ka = kd e).  All forms run the same steps on the same
medium; the FP64 form is the reference's arithmetic (bit-identical to it at
every size the tests cover), so the differences are the FP32 forms' errors.

    python -m laboetie_amd.precision [--size 256] [--flow-steps 10000] [--mp-steps 5000]

Prints one JSON object: the error of every output (populations n, density,
flux j, permeability, VACF(t), D_eff) of each FP32 form relative to its FP64
scale, and the MLUPS / HBM roofline of every form.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import lbx

HBM_PEAK_GBS = 8000.0
LB_BYTES = {lbx.PREC_FP64: 19 * 8 * 2, lbx.PREC_FP32_SHIFT: 19 * 4 * 2,
            lbx.PREC_FP32_STORE: 19 * 4 * 2}  # populations read + written per fluid node per step


def run(prec, dims, nat, tau, fext, flow_steps, mp_steps, Db, ka, kd):
    N = int(np.prod(dims))
    with lbx.Lattice(dims, nat, tau, precision=prec) as L:
        if prec != lbx.PREC_FP64:
            L.set_floor_window(0)  # a fixed step count, like the FP64 run
        L.init_rest(1.0)
        L.timing(True)
        t0 = time.perf_counter()
        r = L.equilibrate(fext, 0.0, max_steps=flow_steps, trace_cap=1)
        L.sync()
        t_lb = time.perf_counter() - t0
        n, rho, j = L.get_state()
        L.mp_init(fext, Db, ka, kd)
        t0 = time.perf_counter()
        m = L.mp_run(mp_steps)
        L.sync()
        t_mp = time.perf_counter() - t0
        d = L.mp_deff()
        v0 = L.mp_init(fext, Db, ka, kd)  # VACF(0) of the same flow (mp_run's trace starts at t = 1)
        lb_ms, lb_n = L.kernel_time(lbx.KERNEL_LB)
        mp_ms, mp_n = L.kernel_time(lbx.KERNEL_MP)
        nf = L.fluid_nodes
    vacf = np.vstack([v0, m["vacf"]])
    lb_avg = lb_ms / lb_n / 1e3
    gbs = LB_BYTES[prec] * nf / lb_avg / 1e9
    return dict(steps=r["steps"], n=n, rho=rho, j=j, vacf=vacf, deff=d,
                lb_mlups=N * r["steps"] / t_lb / 1e6, mp_mlups=N * mp_steps / t_mp / 1e6,
                lb_kernel_us=lb_avg * 1e6, mp_kernel_us=mp_ms / mp_n * 1e3,
                lb_roofline={"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_fluid_node": LB_BYTES[prec]},
                fluid=nf)


def rel(a, b):
    """max |a - b| / max |b|"""
    s = float(np.abs(b).max())
    return float(np.abs(a - b).max()) / s if s > 0 else float(np.abs(a - b).max())


def error_table(size=256, flow_steps=10000, mp_steps=5000, seed=20261202, nat=None, keep_runs=False):
    """Run the three forms on the Gaussian random medium of `size`^3 and
    return the report (the JSON object main prints); with keep_runs, also
    the per-form results under "_runs" (fluid node counts, kernel times)."""
    L_ = size
    dims = (L_, L_, L_)
    tau, fext, Db, kd = 0.8, (1e-6, 0.0, 0.0), 0.01, lbx.KD_DEFAULT
    ka = lbx.ka_kd_from_dF(-1.0, kd)[0]
    t0 = time.perf_counter()
    if nat is None:
        nat = lbx.geometry_random(*dims, 8.0, 0.6, seed)
    print(f"precision: medium {dims} built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    res = {}
    forms = (("fp64", lbx.PREC_FP64), ("fp32_shift", lbx.PREC_FP32_SHIFT), ("fp32_store", lbx.PREC_FP32_STORE))
    for name, prec in forms:
        res[name] = run(prec, dims, nat, tau, fext, flow_steps, mp_steps, Db, ka, kd)
        print(f"precision: {name} done", file=sys.stderr, flush=True)
    a = res["fp64"]
    fluid = nat == 0
    nu = (tau - 0.5) / 3.0
    N = fluid.size

    def perm(r):  # Darcy permeability from the superficial velocity, k = nu <j_x>_box / F_x
        return nu * float(r["j"][0][fluid].sum()) / N / fext[0]

    def errors(b):
        return {
            "n": rel(b["n"], a["n"]),
            "rho": rel(b["rho"][fluid], a["rho"][fluid]),
            "j": rel(b["j"][:, fluid], a["j"][:, fluid]),
            "permeability": abs(perm(b) - perm(a)) / abs(perm(a)),
            "vacf": rel(b["vacf"], a["vacf"]),
            "deff": float(np.abs(b["deff"] - a["deff"]).max() / np.abs(a["deff"]).max()),
        }

    out = {
        "workload": f"Gaussian random medium {L_}^3 (sigma 8, porosity 0.6, seed {seed}), tau 0.8, "
                    f"F=(1e-6,0,0), {flow_steps} flow + {mp_steps} MP steps, dF=-1 kT",
        "fluid_nodes": a["fluid"],
        "error_scale": "max|x32 - x64| / max|x64| per output (D_eff: per component, relative to max |D|)",
        "deff_fp64": a["deff"].tolist(), "permeability_fp64": perm(a),
    }
    for name in ("fp32_shift", "fp32_store"):
        out[f"errors_vs_fp64_{name}"] = {k: float(f"{v:.3e}") for k, v in errors(res[name]).items()}
        out[f"deff_{name}"] = res[name]["deff"].tolist()
        out[f"permeability_{name}"] = perm(res[name])
    for name, _ in forms:
        r = res[name]
        out[name] = {k: (round(v, 2) if isinstance(v, float) else v)
                     for k, v in r.items() if k in ("lb_mlups", "mp_mlups", "lb_kernel_us", "mp_kernel_us",
                                                    "lb_roofline")}
    for name in ("fp32_shift", "fp32_store"):
        out[f"lb_speedup_{name}"] = round(res[name]["lb_mlups"] / a["lb_mlups"], 3)
    if keep_runs:
        out["_runs"] = {name: {k: res[name][k] for k in ("fluid", "lb_kernel_us", "mp_kernel_us", "lb_mlups",
                                                         "mp_mlups")} for name, _ in forms}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--flow-steps", type=int, default=10000)
    ap.add_argument("--mp-steps", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=20261202)
    args = ap.parse_args(argv)
    out = error_table(args.size, args.flow_steps, args.mp_steps, args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
