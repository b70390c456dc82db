"""Batched and served solves on the batch engine (lbxb_*): workload
generators, a continuous-batching serving loop, and their measurements.

Two workloads:

* batch -- chromatography-style 2D instances, a 50/50 mix of 1x50x50 and
  1x100x100, each with 1-8 solid disks of radius 5 at seeded non-overlapping
  centres (lbx_geometry_disks), periodic, tau 1, f_ext = 1e-5 along y,
  equilibrated to max|dj| <= 1e-10 (capped), then moment propagation of a
  tracer with Db = 0.01 and dF uniform in [-3, 0] kT (kd = 0.01,
  ka = kd exp(-dF)).  All instances are admitted at once and advanced
  together.

* serve -- a Poisson stream of requests; each draws a cubic lattice
  N^3 with N log-uniform in [32, 192], a geometry (slit, cylinder along z,
  BCC spheres, Gaussian random medium of porosity U[0.4, 0.8]), f_ext
  log-uniform in [1e-7, 1e-5] along the open axis, target 1e-9, and 2000
  adsorbing MP steps with dF uniform in [-3, 0] kT.  Requests that arrive while
  others run are admitted as soon as their nodes fit (continuous batching);
  latency is arrival -> result.

    python -m laboetie_amd.serving batch [--instances 1024] [--precision fp64|fp32_shift|fp32_store|both|all]
    python -m laboetie_amd.serving serve [--requests 48] [--rates 0.5,1,2] [--hold 5]
                                         [--capacity-nodes N] [--capacity-box N]
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import lbx

KD = lbx.KD_DEFAULT


def _prec(name):
    return {"fp64": lbx.PREC_FP64, "fp32": lbx.PREC_FP32_SHIFT, "fp32_shift": lbx.PREC_FP32_SHIFT,
            "fp32_store": lbx.PREC_FP32_STORE}[name]


# ------------------------------------------------------------------ workloads
def chromatography_requests(n, seed=20261202, lb_cap=20000, mp_steps=2000):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        L = 50 if k % 2 == 0 else 100
        nd = int(rng.integers(1, 9))
        nat, _ = lbx.geometry_disks(1, L, L, nd, 5.0, int(rng.integers(1, 2 ** 62)))
        dF = float(rng.uniform(-3.0, 0.0))
        reqs.append(dict(dims=(1, L, L), nature=nat, tau=1.0, rho0=1.0, fext=(0.0, 1e-5, 0.0), target=1e-10,
                         Db=0.01, ka=KD * math.exp(-dF), kd=KD, mp_steps=mp_steps, max_lb=lb_cap,
                         kind=f"disks{nd}", dF=dF))
    return reqs


def serving_request(rng, lb_cap=10000, mp_steps=2000):
    N = int(round(math.exp(rng.uniform(math.log(32), math.log(192)))))
    kind = ["slit", "cylinder", "bcc", "random"][int(rng.integers(0, 4))]
    F = math.exp(rng.uniform(math.log(1e-7), math.log(1e-5)))
    if kind == "slit":
        nat, fext = lbx.build_geometry(1, N, N, N), (F, 0.0, 0.0)
    elif kind == "cylinder":
        nat, fext = lbx.build_geometry(2, N, N, N), (0.0, 0.0, F)
    elif kind == "bcc":
        nat, fext = lbx.build_geometry(3, N, N, N), (0.0, 0.0, F)
    else:
        phi = float(rng.uniform(0.4, 0.8))
        nat, fext = lbx.geometry_random(N, N, N, 4.0, phi, int(rng.integers(1, 2 ** 62))), (F, 0.0, 0.0)
    dF = float(rng.uniform(-3.0, 0.0))
    return dict(dims=(N, N, N), nature=nat, tau=1.0, rho0=1.0, fext=fext, target=1e-9, Db=0.01,
                ka=KD * math.exp(-dF), kd=KD, mp_steps=mp_steps, max_lb=lb_cap, kind=kind, dF=dF)


def capacity_for(reqs, slack=1.0):
    nodes = sum(((int((r["nature"] == 0).sum()) + 255) // 256) * 256 for r in reqs)
    box = sum(int(np.prod(r["dims"])) for r in reqs)
    return int(nodes * slack) + 256, int(box * slack) + 256


def _add(B, r):
    return B.add(r["dims"], r["nature"], r["tau"], r["rho0"], r["fext"], r["target"], r["Db"], r["ka"], r["kd"],
                 r["mp_steps"], max_lb_steps=r["max_lb"])


# ------------------------------------------------------------------ batch
def run_batch(reqs, precision, lb_round=128, mp_round=128, autotune=True, launch=None, lb_steps=None):
    """All requests admitted at once and advanced together; returns per
    request D_eff, LB steps and exit reason, plus throughput.  launch (one of
    lbx.LAUNCH_*) fixes the launch mode and turns the autotuner off.
    lb_steps (one count per request): each request takes exactly that many
    LB steps (no stop test: target 0, no floor test), so that another form's
    run of the same requests can be compared step for step."""
    cap, box = capacity_for(reqs)
    with lbx.Batch(cap, box, len(reqs), precision=precision) as B:
        if launch is not None:
            B.set_launch_mode(launch)
            autotune = False
        if lb_steps is not None:
            if precision != lbx.PREC_FP64:
                B.set_floor_window(0)
            reqs = [dict(r, target=0.0, max_lb=int(n)) for r, n in zip(reqs, lb_steps)]
        ids = [_add(B, r) for r in reqs]
        assert None not in ids, "capacity"
        tuned = []
        t0 = time.perf_counter()
        rounds = 0
        mp_prev, mp_tuned = 0, 0
        while True:
            active = B.round(lb_round, mp_round)
            rounds += 1
            if autotune and rounds == 2:
                tuned.append(B.autotune(5))  # LB phase
            if autotune:
                # MP work of this round; tuned again whenever it has more than doubled
                # since the last tuning (instances of mixed sizes enter MP at different
                # times, and the block size that suits the first small ones does not
                # suit the large ones)
                mp_now = B.stats()["mp_fluid_updates"]
                d, mp_prev = mp_now - mp_prev, mp_now
                if d > 2 * mp_tuned:
                    tuned.append(B.autotune(5))
                    mp_tuned = d
            if active == 0:
                break
        wall = time.perf_counter() - t0
        st = B.stats()
        out = []
        for iid in ids:
            inf = B.info(iid)
            res = B.result(iid)
            out.append(dict(lb_steps=inf["lb_steps"], exit_reason=inf["exit_reason"], status=inf["status"],
                            phase=inf["phase"], deff=res["deff"].tolist()))
    box_updates = st["lb_box_updates"] + st["mp_box_updates"]
    fluid_updates = st["lb_fluid_updates"] + st["mp_fluid_updates"]
    return dict(results=out, wall_s=wall, rounds=rounds, tuned=tuned,
                mlups=box_updates / wall / 1e6, mflups=fluid_updates / wall / 1e6,
                lb_kernel_mlups=st["lb_box_updates"] / (st["lb_ms"] / 1e3) / 1e6 if st["lb_ms"] else 0.0,
                mp_kernel_mlups=st["mp_box_updates"] / (st["mp_ms"] / 1e3) / 1e6 if st["mp_ms"] else 0.0,
                stats=st)


def run_one_at_a_time(reqs, precision):
    """The plain single-lattice path, one request after the other; returns
    the aggregate rate and each request's D_eff."""
    box = 0
    deff = []
    t0 = time.perf_counter()
    for r in reqs:
        with lbx.Lattice(r["dims"], r["nature"], r["tau"], precision=precision) as L:
            L.init_rest(r["rho0"])
            e = L.equilibrate(r["fext"], r["target"], max_steps=r["max_lb"], trace_cap=1)
            L.mp_init(r["fext"], r["Db"], r["ka"], r["kd"])
            m = L.mp_run(r["mp_steps"])
            deff.append(L.mp_deff().tolist())
            box += int(np.prod(r["dims"])) * (e["steps"] + m["steps"])
    wall = time.perf_counter() - t0
    return dict(wall_s=wall, mlups=box / wall / 1e6, requests=len(reqs), deff=deff)


# ------------------------------------------------------------------ serving
def serve(reqs, arrivals, precision, capacity_nodes, capacity_box, max_instances=256, lb_round=64, mp_round=64,
          hold=5.0):
    """Continuous batching: requests arrive at `arrivals` (s, from start);
    each is admitted as soon as its nodes fit, runs alongside the others, and
    is removed when done.  Returns per-request latency (s)."""
    order = np.argsort(arrivals)
    pending = list(order)
    queue = []
    running = {}
    lat = np.full(len(reqs), np.nan)
    status = np.zeros(len(reqs), np.int32)
    deff = np.full((len(reqs), 3), np.nan)
    with lbx.Batch(capacity_nodes, capacity_box, max_instances, precision=precision) as B:
        t0 = time.perf_counter()
        while pending or queue or running:
            now = time.perf_counter() - t0
            while pending and arrivals[pending[0]] <= now:
                queue.append(pending.pop(0))
            # admission in arrival order as HBM allows; a request that does not
            # fit yet keeps its place while later ones that fit go ahead
            # (backfilling), unless it has waited longer than `hold` seconds
            waiting = []
            for k in queue:
                if waiting and now - arrivals[waiting[0]] > hold:
                    waiting.append(k)
                    continue
                iid = _add(B, reqs[k])
                if iid is None:
                    if not running:  # larger than the whole batch: refused
                        status[k] = lbx.ERR_NOMEM
                        lat[k] = time.perf_counter() - t0 - arrivals[k]
                        continue
                    waiting.append(k)
                else:
                    running[iid] = k
            queue = waiting
            if not running:
                if pending and not queue:
                    time.sleep(max(0.0, min(0.01, arrivals[pending[0]] - now)))
                continue
            B.round(lb_round, mp_round)
            done_t = time.perf_counter() - t0
            for iid in list(running):
                inf = B.info(iid)
                if inf["phase"] in (lbx.PHASE_DONE, lbx.PHASE_FAILED):
                    k = running.pop(iid)
                    lat[k] = done_t - arrivals[k]
                    status[k] = inf["status"]
                    deff[k] = B.result(iid)["deff"]
                    B.remove(iid)
        wall = time.perf_counter() - t0
        st = B.stats()
    return dict(latency=lat, status=status, deff=deff, wall_s=wall,
                mlups=(st["lb_box_updates"] + st["mp_box_updates"]) / wall / 1e6, stats=st)


def _percentiles(x):
    x = np.asarray(x)
    x = x[np.isfinite(x)]
    return {"p50_s": round(float(np.percentile(x, 50)), 3), "p99_s": round(float(np.percentile(x, 99)), 3),
            "mean_s": round(float(x.mean()), 3), "max_s": round(float(x.max()), 3)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["batch", "serve"])
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--precision", default="all",
                    help="fp64, fp32_shift, fp32_store, both (fp64 + fp32_shift) or all; serve uses fp32_shift "
                         "for both/all")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rates", default="0.5,1,2")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--single-sample", type=int, default=16, help="requests timed on the plain FP64 path")
    ap.add_argument("--hold", type=float, default=5.0,
                    help="seconds a request that does not fit may be overtaken by later ones (0: first come, "
                         "first served)")
    ap.add_argument("--capacity-nodes", type=int, default=48 << 20, help="serve: fluid-node budget of the batch")
    ap.add_argument("--capacity-box", type=int, default=64 << 20, help="serve: lattice-node budget of the batch")
    ap.add_argument("--launch", choices=["auto", "graph", "stream", "persistent"], default="auto",
                    help="batch: launch mode (auto: the autotuner picks)")
    args = ap.parse_args(argv)
    launch = {"auto": None, "graph": lbx.LAUNCH_GRAPH, "stream": lbx.LAUNCH_STREAM,
              "persistent": lbx.LAUNCH_PERSISTENT}[args.launch]
    if args.mode == "batch":
        reqs = chromatography_requests(args.instances, args.seed)
        out = {"workload": f"{args.instances} chromatography 2D instances (1x50x50 / 1x100x100, 1-8 disks r=5), "
                           "tau 1, f_ext 1e-5 along y, target 1e-10 (cap 20000), 2000 MP steps, dF U[-3,0]"}
        precs = {"both": ["fp64", "fp32_shift"], "all": ["fp64", "fp32_shift", "fp32_store"]}.get(
            args.precision, [args.precision])
        res = {}
        for p in precs:
            r = run_batch(reqs, _prec(p), launch=launch)
            res[p] = r
            out[p] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.items()
                      if k in ("wall_s", "rounds", "tuned", "mlups", "mflups", "lb_kernel_mlups", "mp_kernel_mlups",
                               "stats")}
            print(f"serving: batch {p} done in {r['wall_s']:.1f} s", file=sys.stderr, flush=True)
        for p in ("fp32_shift", "fp32_store"):
            if p in res and "fp64" in res:
                # precision: the same requests, each with the FP64 run's LB step count
                steps64 = [x["lb_steps"] for x in res["fp64"]["results"]]
                same = run_batch(reqs, _prec(p), launch=launch, lb_steps=steps64)
                d64 = np.array([x["deff"] for x in res["fp64"]["results"]])
                ds = np.array([x["deff"] for x in same["results"]])
                rel_s = np.abs(ds - d64).max(axis=1) / np.abs(d64).max(axis=1)
                out[f"deff_rel_error_{p}_same_lb_steps"] = {
                    "max": float(f"{rel_s.max():.3e}"), "median": float(f"{np.median(rel_s):.3e}"),
                    "count_above_1e-5": int((rel_s > 1e-5).sum())}
                d32 = np.array([x["deff"] for x in res[p]["results"]])
                rel = np.abs(d32 - d64).max(axis=1) / np.abs(d64).max(axis=1)
                out[f"deff_rel_error_{p}"] = {"max": float(f"{rel.max():.3e}"), "median": float(f"{np.median(rel):.3e}"),
                                              "count_above_1e-5": int((rel > 1e-5).sum())}
                # the instances whose FP64 run converged (the others stop at the step cap,
                # where the FP32 forms' floor exit can take a different step)
                conv = np.array([x["exit_reason"] == lbx.EXIT_CONVERGED for x in res["fp64"]["results"]])
                rc = rel[conv]
                out[f"deff_rel_error_{p}_where_fp64_converged"] = {
                    "instances": int(conv.sum()), "max": float(f"{rc.max():.3e}"),
                    "median": float(f"{np.median(rc):.3e}"), "count_above_1e-5": int((rc > 1e-5).sum())}
                out[f"speedup_{p}_batch_vs_fp64_batch"] = round(res[p]["mlups"] / res["fp64"]["mlups"], 3)
        if args.single_sample > 0:
            one = run_one_at_a_time(reqs[:args.single_sample], lbx.PREC_FP64)
            out["plain_fp64_one_at_a_time"] = {k: round(v, 2) if isinstance(v, float) else v
                                               for k, v in one.items() if k != "deff"}
            for p in ("fp32_shift", "fp32_store"):
                if p in res:
                    out[f"speedup_{p}_batch_vs_plain_fp64"] = round(res[p]["mlups"] / one["mlups"], 2)
        print(json.dumps(out), flush=True)
        return
    # serve
    rng = np.random.default_rng(args.seed)
    reqs = [serving_request(rng) for _ in range(args.requests)]
    prec = _prec("fp32_shift" if args.precision in ("both", "all") else args.precision)
    # HBM budget (default 48 M fluid nodes, ~20 GB in the FP32-shift form: room
    # for several of the largest requests at once)
    cap_nodes, cap_box = args.capacity_nodes, args.capacity_box
    # saturation throughput: a closed batch of the first requests, all admitted at t = 0
    sat_reqs = reqs[:12]
    t = serve(sat_reqs, np.zeros(len(sat_reqs)), prec, cap_nodes, cap_box, hold=args.hold)
    sat_rate = len(sat_reqs) / t["wall_s"]
    out = {"workload": f"{args.requests} requests per rate: N^3, N log-uniform in [32,192], slit/cylinder/BCC/"
                       "random(phi U[0.4,0.8]), f_ext log-uniform [1e-7,1e-5], target 1e-9 (cap 10000 LB steps), "
                       "2000 MP steps, dF U[-3,0]",
           "precision": {lbx.PREC_FP64: "fp64", lbx.PREC_FP32_SHIFT: "fp32_shift",
                         lbx.PREC_FP32_STORE: "fp32_store"}[prec],
           "capacity_nodes": cap_nodes, "capacity_box": cap_box, "hold_s": args.hold,
           "saturation_req_per_s": round(sat_rate, 3),
           "saturation_mlups": round(t["mlups"], 1)}
    for f in [float(x) for x in args.rates.split(",")]:
        rate = f * sat_rate
        arr = np.cumsum(np.random.default_rng(args.seed + int(f * 1000)).exponential(1.0 / rate, len(reqs)))
        s = serve(reqs, arr, prec, cap_nodes, cap_box, hold=args.hold)
        out[f"rate_{f:g}x"] = dict(_percentiles(s["latency"]), req_per_s=round(rate, 3), wall_s=round(s["wall_s"], 1),
                                   mlups=round(s["mlups"], 1), failed=int((s["status"] != 0).sum()))
        print(f"serving: rate {f:g}x done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
