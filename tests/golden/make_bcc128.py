"""Write tests/golden/bcc128_full.npz: the benchmark's lattice at its full size
(BCC of touching spheres, 128^3, tau 1, F = 1e-6 along z, target 1e-10),
solved by the oracle (oracle/lbref.c, bit-identical to the reference program
on every fixture of tests/golden): the exit step, the whole l2err trace, a
SHA-256 of the final n, rho and j, and for the three tracers of the benchmark
(Db 0.01, kd 0.01, ka = kd exp(-dF), dF = 0, -1, -2 kT) the first 500 VACF
rows, D_eff over them and a SHA-256 of P and Pa after 500 steps.  The oracle
needs about an hour of one CPU core for it.

    python tests/golden/make_bcc128.py
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import lbref  # noqa: E402

DIMS = (128, 128, 128)
TAU, RHO0, FEXT, TARGET = 1.0, 1.0, [0.0, 0.0, 1e-6], 1e-10
DB, KD, DFS, MP_STEPS = 0.01, 0.01, (0.0, -1.0, -2.0), 500


def digest(*arrays):
    """SHA-256 of the arrays' float64 bytes, -0.0 read as +0.0."""
    h = hashlib.sha256()
    for a in arrays:
        h.update((np.ascontiguousarray(a, np.float64) + 0.0).tobytes())
    return h.hexdigest()


def main():
    t0 = time.time()
    st, nat = lbref.geometry(3, *DIMS)
    assert st == 0
    r = lbref.equilibrate(DIMS, nat, TAU, RHO0, FEXT, TARGET)
    assert r["status"] == 0 and r["exit_reason"] == lbref.EXIT_CONVERGED
    print("equilibrated in", r["steps"], "steps", round(time.time() - t0), "s", flush=True)
    itf = lbref.interfacial(*DIMS, nat)
    nt = lbref.tracer_population(DIMS, nat, r["rho"], r["j"], FEXT)
    out = dict(dims=np.array(DIMS, np.int32), exit_step=r["steps"], l2err=r["l2err"],
               state_sha256=digest(r["n"], r["rho"], r["j"]), dF=np.array(DFS), mp_steps=MP_STEPS)
    vac, deff, psha = [], [], []
    for dF in DFS:
        ka = KD * np.exp(-dF)
        mp = lbref.MP(DIMS, nat, itf, nt, r["rho"], DB, ka, KD)
        st, v = mp.run(MP_STEPS)
        assert st == 0
        vac.append(v)
        deff.append(lbref.deff(v))
        psha.append(digest(mp.P, mp.Pa))
        print("tracer dF", dF, "done", round(time.time() - t0), "s", flush=True)
    out.update(vacf=np.array(vac), deff=np.array(deff), P_sha256=np.array(psha))
    np.savez_compressed(os.path.join(HERE, "bcc128_full.npz"), **out)
    print("wrote bcc128_full.npz", flush=True)


if __name__ == "__main__":
    main()
