"""Test infrastructure, not product code: numpy bindings of liblbref.so, the
scalar C restatement of the reference path (lbref.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker.

Arrays use the reference's Fortran column-major layouts flattened to 1-D
(x fastest); see lbref.h.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbref.so")
_lib = None

OK, NEG_POP, RESTPART, BAD_TAU, BAD_ARG = 0, 1, 2, 3, 5
EXIT_CONVERGED, EXIT_CAP, EXIT_FLOOR = 1, 2, 3

_p = ctypes.c_void_p
_i32, _i64, _f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built; run `make -C oracle lbref`")
        L = ctypes.CDLL(_LIB_PATH)
        L.lbref_tables.argtypes = [_p] * 5
        L.lbref_norm2.argtypes = [_p, ctypes.c_int]
        L.lbref_norm2.restype = _f64
        L.lbref_norm2_rows.argtypes = [_p, ctypes.c_int, _i64, _p]
        L.lbref_geometry.argtypes = [_i32] * 4 + [_p, _i64, _p]
        L.lbref_interfacial.argtypes = [_i32] * 3 + [_p, _p]
        L.lbref_collide.argtypes = [_i32] * 3 + [_p, _f64] + [_p] * 8
        L.lbref_equilibrate.argtypes = [_i32] * 3 + [_p, _f64, _f64, _p, _f64, _i64] + [_p] * 6 + [_i64, _p, _p]
        L.lbref_tracer_population.argtypes = [_i32] * 3 + [_p] * 7
        L.lbref_mp_init.argtypes = [_i32] * 3 + [_p] * 4 + [_f64] * 3 + [_p] * 3
        L.lbref_mp_step.argtypes = [_i32] * 3 + [_p] * 4 + [_f64] * 3 + [_i32] + [_p] * 5
        L.lbref_mp_converged.argtypes = [_i64, _p, _i32, _i32, _i32, _f64]
        L.lbref_deff.argtypes = [_p, _i64, _p]
        L.lbshift_equilibrate.argtypes = [_i32] * 3 + [_p, _f64, _f64, _p, _f64, _i64, _i64, _i64] + [_p] * 6 + [_i64, _p, _p]
        L.lbstore_equilibrate.argtypes = L.lbshift_equilibrate.argtypes
        L.lbshift_mp.argtypes = [_i32] * 3 + [_p] * 7 + [_f64] * 3 + [_i64] + [_p] * 3
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def tables():
    c = np.zeros(57, np.int32)
    a0, a1, a2 = np.zeros(19), np.zeros(19), np.zeros(19)
    inv = np.zeros(19, np.int32)
    lib().lbref_tables(_ptr(c), _ptr(a0), _ptr(a1), _ptr(a2), _ptr(inv))
    return c.reshape(19, 3), a0, a1, a2, inv


def norm2(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().lbref_norm2(_ptr(v), v.size)


def geometry(label, lx, ly, lz, solid_ijk=None):
    nat = np.zeros(lx * ly * lz, np.int8)
    pts = np.ascontiguousarray(np.zeros((0, 3), np.int32) if solid_ijk is None else solid_ijk, np.int32)
    st = lib().lbref_geometry(label, lx, ly, lz, _ptr(pts), pts.size // 3, _ptr(nat))
    return st, nat


def interfacial(lx, ly, lz, nature):
    itf = np.zeros(lx * ly * lz, np.int8)
    lib().lbref_interfacial(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)), _ptr(itf))
    return itf


def collide(dims, nature, tau, n, rho, j, f):
    lx, ly, lz = dims
    n = np.array(n, np.float64)
    args = [np.ascontiguousarray(a, np.float64) for a in (rho, j[0], j[1], j[2], f[0], f[1], f[2])]
    st = lib().lbref_collide(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)), tau, _ptr(n),
                             *[_ptr(a) for a in args])
    return st, n


def equilibrate(dims, nature, tau, rho0, fext, target_error, max_steps=0, trace_cap=1 << 20):
    lx, ly, lz = dims
    N = lx * ly * lz
    n = np.zeros(19 * N)
    rho, jx, jy, jz = (np.zeros(N) for _ in range(4))
    tr = np.zeros(trace_cap)
    sd, er = _i64(), _i32()
    fe = np.ascontiguousarray(fext, np.float64)
    st = lib().lbref_equilibrate(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)), tau, rho0, _ptr(fe),
                                 target_error, max_steps, _ptr(n), _ptr(rho), _ptr(jx), _ptr(jy), _ptr(jz),
                                 _ptr(tr), trace_cap, ctypes.byref(sd), ctypes.byref(er))
    steps = sd.value
    return dict(status=st, steps=steps, exit_reason=er.value, n=n, rho=rho, j=np.stack([jx, jy, jz]),
                l2err=tr[:min(steps, trace_cap)].copy())


def tracer_population(dims, nature, rho, j, fext):
    lx, ly, lz = dims
    nt = np.zeros(19 * lx * ly * lz)
    fe = np.ascontiguousarray(fext, np.float64)
    js = [np.ascontiguousarray(j[a], np.float64) for a in range(3)]
    lib().lbref_tracer_population(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)),
                                  _ptr(np.ascontiguousarray(rho, np.float64)), *[_ptr(a) for a in js],
                                  _ptr(fe), _ptr(nt))
    return nt


class MP:
    """Moment propagation state (module_moment_propagation.f90) on the CPU."""

    def __init__(self, dims, nature, interfacial, nt, rho, Db, ka=0.0, kd=0.0):
        self.dims = tuple(int(v) for v in dims)
        lx, ly, lz = self.dims
        N = lx * ly * lz
        self.nature = np.ascontiguousarray(nature, np.int8)
        self.interfacial = np.ascontiguousarray(interfacial, np.int8)
        self.nt = np.ascontiguousarray(nt, np.float64)
        self.rho = np.ascontiguousarray(rho, np.float64)
        self.Db, self.ka, self.kd = float(Db), float(ka), float(kd)
        self.P = np.zeros(3 * N)
        self.Pa = np.zeros(3 * N)
        self.vacf0 = np.zeros(3)
        ads = _i32()
        self.status = lib().lbref_mp_init(lx, ly, lz, _ptr(self.nature), _ptr(self.interfacial), _ptr(self.nt),
                                          _ptr(self.rho), self.Db, self.ka, self.kd, _ptr(self.P),
                                          _ptr(self.vacf0), ctypes.byref(ads))
        self.ads = ads.value
        self.it = 0

    def step(self):
        lx, ly, lz = self.dims
        N = lx * ly * lz
        Pn, Pan, v = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3)
        st = lib().lbref_mp_step(lx, ly, lz, _ptr(self.nature), _ptr(self.interfacial), _ptr(self.nt),
                                 _ptr(self.rho), self.Db, self.ka, self.kd, self.ads, _ptr(self.P),
                                 _ptr(self.Pa), _ptr(Pn), _ptr(Pan), _ptr(v))
        self.P = Pn
        if self.ads:
            self.Pa = Pan
        self.it += 1
        return st, v

    def converged(self, vacf):
        lx, ly, lz = self.dims
        v = np.ascontiguousarray(vacf, np.float64)
        return bool(lib().lbref_mp_converged(self.it, _ptr(v), lx, ly, lz, self.Db))

    def run(self, nsteps):
        out = np.zeros((nsteps + 1, 3))
        out[0] = self.vacf0
        for t in range(1, nsteps + 1):
            st, v = self.step()
            if st:
                return st, out[:t]
            out[t] = v
        return OK, out


def deff(vacf):
    v = np.ascontiguousarray(vacf, np.float64)
    d = np.zeros(3)
    lib().lbref_deff(_ptr(v), v.shape[0] - 1, _ptr(d))
    return d


# ---------------------------------------------------------------- FP32-shift
def equilibrate_shift(dims, nature, tau, rho0, fext, target_error, max_steps=0, floor_window=1000,
                      trace_cap=1 << 20, arith="fp32", check_every=64):
    """lbshift_equilibrate (arith="fp32": the FP32-shift form) or
    lbstore_equilibrate (arith="fp64": the FP32-store form, collision in
    FP64): equilibrate with FP32 deviation storage (same outputs, FP64 n /
    rho / j built from the FP32 deviations).  Stage 2 measures l2err every
    check_every steps (the library's kCheckEvery = 64); the l2err trace is NaN
    at the other steps."""
    lx, ly, lz = dims
    N = lx * ly * lz
    n = np.zeros(19 * N)
    rho, jx, jy, jz = (np.zeros(N) for _ in range(4))
    tr = np.zeros(trace_cap)
    sd, er = _i64(), _i32()
    fe = np.ascontiguousarray(fext, np.float64)
    fn = lib().lbshift_equilibrate if arith == "fp32" else lib().lbstore_equilibrate
    st = fn(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)), tau, rho0, _ptr(fe), target_error, max_steps,
            floor_window, check_every, _ptr(n), _ptr(rho), _ptr(jx), _ptr(jy), _ptr(jz), _ptr(tr), trace_cap, ctypes.byref(sd),
            ctypes.byref(er))
    steps = sd.value
    nl2 = steps - 1 if st == NEG_POP else steps
    return dict(status=st, steps=steps, exit_reason=er.value, n=n, rho=rho, j=np.stack([jx, jy, jz]),
                l2err=tr[:min(nl2, trace_cap)].copy())


def mp_shift(dims, nature, interfacial, rho, j, fext, Db, ka, kd, nsteps):
    """lbshift_mp: FP32-shift moment propagation; returns (status, vacf
    (nsteps+1, 3), P, Pa)."""
    lx, ly, lz = dims
    N = lx * ly * lz
    P, Pa = np.zeros(3 * N), np.zeros(3 * N)
    v = np.zeros(3 * (nsteps + 1))
    js = [np.ascontiguousarray(j[a], np.float64) for a in range(3)]
    fe = np.ascontiguousarray(fext, np.float64)
    st = lib().lbshift_mp(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)),
                          _ptr(np.ascontiguousarray(interfacial, np.int8)),
                          _ptr(np.ascontiguousarray(rho, np.float64)), *[_ptr(a) for a in js], _ptr(fe),
                          float(Db), float(ka), float(kd), int(nsteps), _ptr(P), _ptr(Pa), _ptr(v))
    return st, v.reshape(-1, 3), P, Pa
