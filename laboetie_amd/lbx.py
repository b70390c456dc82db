"""ctypes binding of liblbx.so, the HIP library behind include/lbx.h.

Host-side mirror of the reference interface for the hot path: `Lattice`
wraps one device-resident lattice (an lbx_ctx) with the operations the
reference's Fortran routines perform (equilibration, moment propagation
init/propagate, collide), returning numpy arrays in the reference's
Fortran-order layouts.  The library must be built (make -C
laboetie_amd/csrc); there is no fallback: if liblbx.so is missing or no HIP
device is present, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbx.so")

OK, ERR_NEG_POP, ERR_RESTPART, ERR_TAU, ERR_NOMEM, ERR_ARG, ERR_DEVICE, ERR_STATE, ERR_SIZE = range(9)
NOALLOC_MAX_NODES = 320000
PREC_FP64, PREC_FP32_SHIFT, PREC_FP32_STORE = 0, 1, 2
EXIT_CONVERGED, EXIT_CAP, EXIT_FLOOR = 1, 2, 3
KERNEL_LB, KERNEL_MP = 0, 1
MP_AUTO, MP_STORED, MP_RECOMPUTED, MP_MIXED = -1, 0, 1, 2
AUTOTUNE_TIMES = 10  # LBXB_AUTOTUNE_TIMES

_p = ctypes.c_void_p
_i32, _i64, _f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_lib = None

# every exported symbol and its (restype, argtypes)
SIGNATURES = {
    "lbx_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "lbx_version": (ctypes.c_int, []),
    "lbx_d3q19_tables": (ctypes.c_int, [_p] * 5),
    "lbx_build_geometry": (ctypes.c_int, [_i32] * 4 + [_p, _i64, _p]),
    "lbx_detect_interfacial": (ctypes.c_int, [_i32] * 3 + [_p, _p]),
    "lbx_gaussian_field": (ctypes.c_int, [_i32] * 3 + [_f64, ctypes.c_uint64, _p]),
    "lbx_geometry_random": (ctypes.c_int, [_i32] * 3 + [_f64, _f64, ctypes.c_uint64, _p]),
    "lbx_geometry_disks": (ctypes.c_int, [_i32] * 4 + [_f64, ctypes.c_uint64, _p, _p]),
    "lbx_lbin_parse": (ctypes.c_int, [ctypes.c_char_p, _i64, _p]),
    "lbx_lbin_free": (ctypes.c_int, [_p]),
    "lbx_lbin_dp": (ctypes.c_int, [_p, ctypes.c_char_p, _i32, _p, _p]),
    "lbx_lbin_int": (ctypes.c_int, [_p, ctypes.c_char_p, _i32, _p, _p]),
    "lbx_lbin_log": (ctypes.c_int, [_p, ctypes.c_char_p, _p, _p]),
    "lbx_lbin_char": (ctypes.c_int, [_p, ctypes.c_char_p, _p, _i64, _p]),
    "lbx_geom_in_parse": (ctypes.c_int, [ctypes.c_char_p, _i64, _i32, _i32, _i32, _p, _p]),
    "lbx_supercell_xsf": (ctypes.c_int, [_i32, _i32, _i32, _p, _p, _p, _i64, _p]),
    "lbx_footprint_bytes": (_i64, [_i32] * 5),
    "lbx_create": (ctypes.c_int, [_i32] * 3 + [_p, _f64, _i32, _p]),
    "lbx_destroy": (ctypes.c_int, [_p]),
    "lbx_init_rest": (ctypes.c_int, [_p, _f64]),
    "lbx_set_state": (ctypes.c_int, [_p] * 6),
    "lbx_equilibrate": (ctypes.c_int, [_p, _f64, _p, _i64, _p, _i64, _p, _p]),
    "lbx_get_state": (ctypes.c_int, [_p] * 6),
    "lbx_equilibrate_begin": (ctypes.c_int, [_p, _f64, _p, _i64, _i32]),
    "lbx_equilibrate_advance": (ctypes.c_int, [_p, _i64, _p, _p]),
    "lbx_flux_sums": (ctypes.c_int, [_p, _i64, _i64, _p, _p]),
    "lbx_flow": (ctypes.c_int, [_p] * 5),
    "lbx_l2err_trace": (ctypes.c_int, [_p, _i64, _i64, _p, _p]),
    "lbx_mp_init": (ctypes.c_int, [_p, _p, _f64, _f64, _f64, _p]),
    "lbx_mp_init_from_flow": (ctypes.c_int, [_p] * 6 + [_f64] * 3 + [_p]),
    "lbx_mp_init_batch": (ctypes.c_int, [_p, _p, _f64, _i32, _p, _p, _p]),
    "lbx_mp_tracers": (ctypes.c_int, [_p, _p]),
    "lbx_mp_run": (ctypes.c_int, [_p, _i64, _i32, _p, _p, _p]),
    "lbx_mp_run_start": (ctypes.c_int, [_p, _i64, _i32]),
    "lbx_mp_run_wait": (ctypes.c_int, [_p, _p, _p, _p]),
    "lbx_mp_get": (ctypes.c_int, [_p, _p, _p]),
    "lbx_mp_get_tracer": (ctypes.c_int, [_p, _i32, _p, _p]),
    "lbx_mp_deff": (ctypes.c_int, [_p, _p]),
    "lbx_mp_deff_batch": (ctypes.c_int, [_p, _p]),
    "lbx_mp_free": (ctypes.c_int, [_p]),
    "lbx_collide": (ctypes.c_int, [_i32] * 3 + [_p, _f64] + [_p] * 8),
    "lbx_collide_noalloc": (ctypes.c_int, [_i32] * 3 + [_p, _f64] + [_p] * 8),
    "lbx_timing_enable": (ctypes.c_int, [_p, _i32]),
    "lbx_kernel_time": (ctypes.c_int, [_p, _i32, _p, _p]),
    "lbx_set_floor_window": (ctypes.c_int, [_p, _i64]),
    "lbx_set_check_interval": (ctypes.c_int, [_p, _i32]),
    "lbx_set_addressing": (ctypes.c_int, [_p, _i32]),
    "lbx_addressing": (ctypes.c_int, [_p]),
    "lbx_set_mp_coefficients": (ctypes.c_int, [_p, _i32]),
    "lbx_fluid_nodes": (_i64, [_p]),
    "lbx_sync": (ctypes.c_int, [_p]),
    "lbxb_footprint_bytes": (_i64, [_i32, _i64, _i64, _i32]),
    "lbxb_create": (ctypes.c_int, [_i32, _i64, _i64, _i32, _p]),
    "lbxb_destroy": (ctypes.c_int, [_p]),
    "lbxb_add": (ctypes.c_int, [_p, _i32, _i32, _i32, _p, _f64, _f64, _p, _f64, _i64, _f64, _f64, _f64, _i64, _p]),
    "lbxb_remove": (ctypes.c_int, [_p, _i32]),
    "lbxb_round": (ctypes.c_int, [_p, _i32, _i32, _p]),
    "lbxb_info": (ctypes.c_int, [_p, _i32, _p, _p, _p, _p, _p]),
    "lbxb_result": (ctypes.c_int, [_p, _i32, _p, _p, _i64, _p, _p, _p, _p]),
    "lbxb_set_block": (ctypes.c_int, [_p, _i32, _i32]),
    "lbxb_set_check_interval": (ctypes.c_int, [_p, _i32]),
    "lbxb_set_floor_window": (ctypes.c_int, [_p, _i64]),
    "lbxb_autotune": (ctypes.c_int, [_p, _i32, _p, _p, _p]),
    "lbxb_set_launch_mode": (ctypes.c_int, [_p, _i32]),
    "lbxb_set_addressing": (ctypes.c_int, [_p, _i32]),
    "lbxb_addressing": (ctypes.c_int, [_p]),
    "lbxb_launch_mode": (ctypes.c_int, [_p]),
    "lbxb_stats": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p]),
    "lbxb_free_tiles": (_i64, [_p]),
    "lbx_copy_bandwidth": (ctypes.c_int, [_i64, _i32, _p]),
    "lbx_planes_bandwidth": (ctypes.c_int, [_i32, _i64, _i32, _p]),
}


class LbxError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = lib().lbx_status_string(status).decode()
        super().__init__(f"{where}: lbx status {status}: {msg}" if where else f"lbx status {status}: {msg}")


def lib():
    """Load liblbx.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C laboetie_amd/csrc` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _check(st, where):
    if st != OK:
        raise LbxError(st, where)


def d3q19_tables():
    c = np.zeros(57, np.int32)
    a0, a1, a2 = np.zeros(19), np.zeros(19), np.zeros(19)
    inv = np.zeros(19, np.int32)
    _check(lib().lbx_d3q19_tables(_ptr(c), _ptr(a0), _ptr(a1), _ptr(a2), _ptr(inv)), "d3q19_tables")
    return c.reshape(19, 3), a0, a1, a2, inv


def build_geometry(label, lx, ly, lz, solid_ijk=None):
    """node%nature of supercell_definition (labels -1, 0, 1, 2, 3)."""
    nat = np.zeros(lx * ly * lz, np.int8)
    pts = np.ascontiguousarray(np.zeros((0, 3), np.int32) if solid_ijk is None else solid_ijk, np.int32)
    _check(lib().lbx_build_geometry(label, lx, ly, lz, _ptr(pts), pts.size // 3, _ptr(nat)), "build_geometry")
    return nat


def detect_interfacial(lx, ly, lz, nature):
    itf = np.zeros(lx * ly * lz, np.int8)
    _check(lib().lbx_detect_interfacial(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)), _ptr(itf)),
           "detect_interfacial")
    return itf


def gaussian_field(lx, ly, lz, sigma, seed):
    """Periodic Gaussian-filtered white noise (lbx_gaussian_field), (lx,ly,lz) x fastest."""
    f = np.zeros(lx * ly * lz)
    _check(lib().lbx_gaussian_field(lx, ly, lz, float(sigma), int(seed), _ptr(f)), "gaussian_field")
    return f


def geometry_random(lx, ly, lz, sigma=8.0, porosity=0.6, seed=20261202):
    """Gaussian random porous medium with exactly round(porosity*N) fluid nodes."""
    nat = np.zeros(lx * ly * lz, np.int8)
    _check(lib().lbx_geometry_random(lx, ly, lz, float(sigma), float(porosity), int(seed), _ptr(nat)),
           "geometry_random")
    return nat


def geometry_disks(lx, ly, lz, ndisks, radius=5.0, seed=1):
    """Seeded non-overlapping disks (spheres) of the chromatography inputs;
    returns (nature, centres (ndisks, 3), 1-based)."""
    nat = np.zeros(lx * ly * lz, np.int8)
    c = np.zeros(3 * max(ndisks, 1))
    _check(lib().lbx_geometry_disks(lx, ly, lz, int(ndisks), float(radius), int(seed), _ptr(nat), _ptr(c)),
           "geometry_disks")
    return nat, c[:3 * ndisks].reshape(-1, 3)


class LbInNative:
    """lb.in through the library's C++ reader (lbx_lbin_*), the reference's
    grammar; getters return the default when the tag is absent."""

    def __init__(self, text):
        raw = text.encode() if isinstance(text, str) else bytes(text)
        h = _p()
        _check(lib().lbx_lbin_parse(raw, len(raw), ctypes.byref(h)), "lbin_parse")
        self._h = h

    def close(self):
        if self._h:
            lib().lbx_lbin_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, fn, tag, count, ctype, default):
        out = (ctype * count)()
        found = _i32()
        _check(fn(self._h, tag.encode(), count, out, ctypes.byref(found)), f"lbin {tag}")
        if not found.value:
            return default
        return out[0] if count == 1 else list(out)

    def dp(self, tag, default=None):
        return self._get(lib().lbx_lbin_dp, tag, 1, _f64, default)

    def dp3(self, tag, default=None):
        return self._get(lib().lbx_lbin_dp, tag, 3, _f64, default)

    def int(self, tag, default=None):
        return self._get(lib().lbx_lbin_int, tag, 1, _i64, default)

    def int3(self, tag, default=None):
        return self._get(lib().lbx_lbin_int, tag, 3, _i64, default)

    def log(self, tag, default=None):
        out, found = _i32(), _i32()
        _check(lib().lbx_lbin_log(self._h, tag.encode(), ctypes.byref(out), ctypes.byref(found)), f"lbin {tag}")
        return bool(out.value) if found.value else default

    def char(self, tag, default=None):
        buf = ctypes.create_string_buffer(128)
        found = _i32()
        _check(lib().lbx_lbin_char(self._h, tag.encode(), buf, 128, ctypes.byref(found)), f"lbin {tag}")
        return buf.value.decode() if found.value else default

    def has(self, tag):
        return self.char(tag) is not None


def geom_in_nature(text, lx, ly, lz):
    """geom.in -> nature (lx,ly,lz, x fastest) through lbx_geom_in_parse;
    returns (nature, number of distinct solid nodes)."""
    raw = text.encode() if isinstance(text, str) else bytes(text)
    nat = np.zeros(lx * ly * lz, np.int8)
    ns = _i64()
    _check(lib().lbx_geom_in_parse(raw, len(raw), lx, ly, lz, _ptr(nat), ctypes.byref(ns)), "geom_in_parse")
    return nat, ns.value


def supercell_xsf(lx, ly, lz, nature, interfacial):
    """The text of output/supercell.xsf (lbx_supercell_xsf)."""
    nat = np.ascontiguousarray(nature, np.int8)
    itf = np.ascontiguousarray(interfacial, np.int8)
    n = _i64()
    _check(lib().lbx_supercell_xsf(lx, ly, lz, _ptr(nat), _ptr(itf), None, 0, ctypes.byref(n)), "supercell_xsf")
    buf = ctypes.create_string_buffer(n.value)
    _check(lib().lbx_supercell_xsf(lx, ly, lz, _ptr(nat), _ptr(itf), buf, n.value, ctypes.byref(n)),
           "supercell_xsf")
    return buf.raw[:n.value].decode()


def footprint_bytes(lx, ly, lz, precision=PREC_FP64, with_mp=1):
    """HBM bytes of a context; with_mp = number of tracers (True = 1)."""
    return int(lib().lbx_footprint_bytes(lx, ly, lz, precision, int(with_mp)))


KD_DEFAULT = 0.01


def ka_kd_from_dF(dF, kd=KD_DEFAULT):
    """Adsorption/desorption probabilities per step for an adsorption free
    energy dF (in kT).  The reference takes ka and kd directly
    (module_moment_propagation.f90:41-56, K = ka/kd); here kd is fixed
    (default 0.01) and K = exp(-dF), so ka = kd exp(-dF) (dF in [-3, 0]
    gives ka in [0.01, 0.201])."""
    return kd * float(np.exp(-dF)), kd


def copy_bandwidth(bytes_=2 << 30, iters=20):
    """Read+write GB/s of a plain streaming copy on the device."""
    g = _f64()
    _check(lib().lbx_copy_bandwidth(int(bytes_), int(iters), ctypes.byref(g)), "copy_bandwidth")
    return g.value


def planes_bandwidth(planes=19, nodes=10066330, iters=20):
    """Read+write GB/s of streaming `planes` FP64 planes of `nodes` values
    (the LB step's structure-of-arrays streams without its gathers)."""
    g = _f64()
    _check(lib().lbx_planes_bandwidth(int(planes), int(nodes), int(iters), ctypes.byref(g)), "planes_bandwidth")
    return g.value


def collide(dims, nature, tau, n, density, j, f, noalloc=False):
    """module_collision::collide on host arrays; returns the new n.
    noalloc: through lbx_collide_noalloc (no device allocation; at most
    NOALLOC_MAX_NODES nodes)."""
    lx, ly, lz = dims
    n = np.array(n, np.float64)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (density, j[0], j[1], j[2], f[0], f[1], f[2])]
    fn = lib().lbx_collide_noalloc if noalloc else lib().lbx_collide
    _check(fn(lx, ly, lz, _ptr(np.ascontiguousarray(nature, np.int8)), tau, _ptr(n), *[_ptr(a) for a in arrs]),
           "collide_noalloc" if noalloc else "collide")
    return n


class Lattice:
    """One lattice resident in HBM (an lbx_ctx)."""

    def __init__(self, dims, nature, tau=1.0, precision=PREC_FP64):
        self._batch = False
        self.dims = tuple(int(v) for v in dims)
        lx, ly, lz = self.dims
        self.N = lx * ly * lz
        self.nature = np.ascontiguousarray(nature, np.int8)
        assert self.nature.size == self.N
        h = _p()
        _check(lib().lbx_create(lx, ly, lz, _ptr(self.nature), float(tau), precision, ctypes.byref(h)), "create")
        self._h = h
        self.tau = tau

    def close(self):
        if self._h is not None and self._h.value:
            lib().lbx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_floor_window(self, steps):
        """FP32-shift form: end a stage when l2err set no new minimum for
        `steps` steps (0: off; default 1000 for FP32-shift contexts)."""
        _check(lib().lbx_set_floor_window(self._h, int(steps)), "set_floor_window")

    def set_check_interval(self, steps):
        """FP32-storage forms: steps between stage-2 l2err measurements (a
        power of two; default 64; 1 measures every step)."""
        _check(lib().lbx_set_check_interval(self._h, int(steps)), "set_check_interval")

    def set_addressing(self, wide):
        """1: the 64-bit addressing form of the step kernels on any size (the
        library selects it by itself above 4 GiB of populations); 0: the
        size's choice."""
        _check(lib().lbx_set_addressing(self._h, int(wide)), "set_addressing")

    @property
    def addressing(self):
        return int(lib().lbx_addressing(self._h))

    def set_mp_coefficients(self, form):
        """MP_RECOMPUTED: the MP step recomputes the incoming scattering
        probabilities from each node's density and flow; MP_STORED: keeps them
        (18 per node); MP_MIXED: keeps the 10 with c_z != 0 and recomputes the
        others; MP_AUTO (default): recomputed for several FP64 tracers, stored
        otherwise.  Same bits; from the next mp_init."""
        _check(lib().lbx_set_mp_coefficients(self._h, int(form)), "set_mp_coefficients")

    @property
    def fluid_nodes(self):
        return int(lib().lbx_fluid_nodes(self._h))

    def init_rest(self, rho0=1.0):
        _check(lib().lbx_init_rest(self._h, float(rho0)), "init_rest")

    def set_state(self, n, rho, j):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (n, rho, j[0], j[1], j[2])]
        _check(lib().lbx_set_state(self._h, *[_ptr(a) for a in arrs]), "set_state")

    def equilibrate(self, fext, target_error=1e-10, max_steps=0, trace_cap=1 << 20, check=True):
        fe = np.ascontiguousarray(fext, np.float64)
        tr = np.zeros(trace_cap)
        sd, er = _i64(), _i32()
        st = lib().lbx_equilibrate(self._h, float(target_error), _ptr(fe), int(max_steps), _ptr(tr), trace_cap,
                                   ctypes.byref(sd), ctypes.byref(er))
        if check and st not in (OK,):
            raise LbxError(st, "equilibrate")
        steps = sd.value
        nl2 = steps - 1 if st == ERR_NEG_POP else steps
        return dict(status=st, steps=steps, exit_reason=er.value, l2err=tr[:min(nl2, trace_cap)].copy())

    def equilibrate_begin(self, fext, target_error=1e-10, max_steps=0, flux_sums=False):
        """Start the equilibration loop; lbx_equilibrate_advance runs it on."""
        fe = np.ascontiguousarray(fext, np.float64)
        _check(lib().lbx_equilibrate_begin(self._h, float(target_error), _ptr(fe), int(max_steps),
                                           1 if flux_sums else 0), "equilibrate_begin")

    def equilibrate_advance(self, until_step=(1 << 63) - 1, check=True):
        """Run the begun equilibration to step until_step or its end;
        exit_reason 0 while it goes on."""
        sd, er = _i64(), _i32()
        st = lib().lbx_equilibrate_advance(self._h, int(until_step), ctypes.byref(sd), ctypes.byref(er))
        if check and st != OK:
            raise LbxError(st, "equilibrate_advance")
        return dict(status=st, steps=sd.value, exit_reason=er.value)

    def flux_sums(self):
        """sum_r j(r, t) for every step taken, (steps, 3)."""
        n = _i64()
        _check(lib().lbx_flux_sums(self._h, 1, 0, None, ctypes.byref(n)), "flux_sums")
        out = np.zeros(3 * max(n.value, 1))
        if n.value:
            _check(lib().lbx_flux_sums(self._h, 1, n.value, _ptr(out), None), "flux_sums")
        return out[:3 * n.value].reshape(-1, 3)

    def flow(self):
        """(rho, j) after the last step taken (also during a paused run)."""
        rho, jx, jy, jz = (np.zeros(self.N) for _ in range(4))
        _check(lib().lbx_flow(self._h, _ptr(rho), _ptr(jx), _ptr(jy), _ptr(jz)), "flow")
        return rho, np.stack([jx, jy, jz])

    def l2err_trace(self):
        """Every l2err of the last equilibrate call."""
        n = _i64()
        _check(lib().lbx_l2err_trace(self._h, 1, 0, None, ctypes.byref(n)), "l2err_trace")
        out = np.zeros(n.value)
        if n.value:
            _check(lib().lbx_l2err_trace(self._h, 1, n.value, _ptr(out), None), "l2err_trace")
        return out

    def get_state(self, populations=True):
        """(n, rho, j) on the full box; n is None with populations=False."""
        n = np.zeros(19 * self.N) if populations else None
        rho, jx, jy, jz = (np.zeros(self.N) for _ in range(4))
        _check(lib().lbx_get_state(self._h, _ptr(n) if populations else None, _ptr(rho), _ptr(jx), _ptr(jy),
                                   _ptr(jz)), "get_state")
        return n, rho, np.stack([jx, jy, jz])

    def mp_init(self, fext, Db, ka=0.0, kd=0.0, flow=None):
        fe = np.ascontiguousarray(fext, np.float64)
        v0 = np.zeros(3)
        if flow is None:
            _check(lib().lbx_mp_init(self._h, _ptr(fe), float(Db), float(ka), float(kd), _ptr(v0)), "mp_init")
        else:
            rho, j = flow
            arrs = [np.ascontiguousarray(a, np.float64) for a in (rho, j[0], j[1], j[2])]
            _check(lib().lbx_mp_init_from_flow(self._h, *[_ptr(a) for a in arrs], _ptr(fe), float(Db), float(ka),
                                               float(kd), _ptr(v0)), "mp_init_from_flow")
        self._batch = False
        return v0

    def mp_init_batch(self, fext, Db, ka, kd):
        """Tracers with the given ka[t], kd[t] (up to 4) propagated together;
        returns VACF(0) as (ntracers, 3)."""
        fe = np.ascontiguousarray(fext, np.float64)
        ka = np.ascontiguousarray(ka, np.float64).ravel()
        kd = np.ascontiguousarray(kd, np.float64).ravel()
        if ka.size != kd.size:
            raise ValueError("ka and kd need one value per tracer")
        v0 = np.zeros(3 * max(ka.size, 1))
        _check(lib().lbx_mp_init_batch(self._h, _ptr(fe), float(Db), int(ka.size), _ptr(ka), _ptr(kd), _ptr(v0)),
               "mp_init_batch")
        self._batch = True
        return v0.reshape(-1, 3)

    @property
    def mp_tracers(self):
        n = _i32()
        _check(lib().lbx_mp_tracers(self._h, ctypes.byref(n)), "mp_tracers")
        return n.value

    def mp_run(self, nsteps, stop_on_convergence=False, check=True):
        """VACF per step: (steps, 3) after mp_init, (steps, ntracers, 3)
        after mp_init_batch."""
        nt = max(self.mp_tracers, 1)
        tr = np.zeros(3 * nt * max(nsteps, 1))
        sd, cv = _i64(), _i32()
        st = lib().lbx_mp_run(self._h, int(nsteps), 1 if stop_on_convergence else 0, _ptr(tr), ctypes.byref(sd),
                              ctypes.byref(cv))
        if check and st != OK:
            raise LbxError(st, "mp_run")
        v = tr[:3 * nt * sd.value]
        v = v.reshape(-1, nt, 3) if self._batch else v.reshape(-1, 3)
        return dict(status=st, steps=sd.value, converged=bool(cv.value), vacf=v)

    def mp_run_start(self, nsteps, stop_on_convergence=False):
        """lbx_mp_run_start: launch 1..512 steps and return at once."""
        _check(lib().lbx_mp_run_start(self._h, int(nsteps), 1 if stop_on_convergence else 0), "mp_run_start")

    def mp_run_wait(self):
        """lbx_mp_run_wait: the begun steps' result, as mp_run returns it."""
        nt = max(self.mp_tracers, 1)
        tr = np.zeros(3 * nt * 512)
        sd, cv = _i64(), _i32()
        _check(lib().lbx_mp_run_wait(self._h, _ptr(tr), ctypes.byref(sd), ctypes.byref(cv)), "mp_run_wait")
        v = tr[:3 * nt * sd.value]
        v = v.reshape(-1, nt, 3) if self._batch else v.reshape(-1, 3)
        return dict(status=OK, steps=sd.value, converged=bool(cv.value), vacf=v)

    def mp_get(self, tracer=0):
        P, Pa = np.zeros(3 * self.N), np.zeros(3 * self.N)
        _check(lib().lbx_mp_get_tracer(self._h, int(tracer), _ptr(P), _ptr(Pa)), "mp_get_tracer")
        return P, Pa

    def mp_deff(self):
        """D_eff: (3,) after mp_init, (ntracers, 3) after mp_init_batch."""
        if self._batch:
            d = np.zeros(3 * self.mp_tracers)
            _check(lib().lbx_mp_deff_batch(self._h, _ptr(d)), "mp_deff_batch")
            return d.reshape(-1, 3)
        d = np.zeros(3)
        _check(lib().lbx_mp_deff(self._h, _ptr(d)), "mp_deff")
        return d

    def mp_free(self):
        _check(lib().lbx_mp_free(self._h), "mp_free")

    def timing(self, on=True):
        _check(lib().lbx_timing_enable(self._h, 1 if on else 0), "timing_enable")

    def kernel_time(self, which):
        ms, n = _f64(), _i64()
        _check(lib().lbx_kernel_time(self._h, which, ctypes.byref(ms), ctypes.byref(n)), "kernel_time")
        return ms.value, n.value

    def sync(self):
        _check(lib().lbx_sync(self._h), "sync")


# ------------------------------------------------------------- batch engine
PHASE_LB, PHASE_SWITCH, PHASE_LB_DONE, PHASE_MP, PHASE_DONE, PHASE_FAILED = 1, 2, 3, 4, 5, 6
LAUNCH_STREAM, LAUNCH_GRAPH, LAUNCH_PERSISTENT = 0, 1, 2


class Batch:
    """Many lattices resident in HBM, advanced together (an lbxb)."""

    def __init__(self, capacity_nodes, capacity_box, max_instances, precision=PREC_FP64):
        h = _p()
        _check(lib().lbxb_create(precision, int(capacity_nodes), int(capacity_box), int(max_instances),
                                 ctypes.byref(h)), "lbxb_create")
        self._h = h
        self.precision = precision
        self.dims = {}

    @staticmethod
    def footprint_bytes(capacity_nodes, capacity_box, max_instances, precision=PREC_FP64):
        return int(lib().lbxb_footprint_bytes(precision, int(capacity_nodes), int(capacity_box), int(max_instances)))

    def close(self):
        if self._h is not None and self._h.value:
            lib().lbxb_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, dims, nature, tau, rho0, fext, target_error, Db, ka, kd, mp_steps, max_lb_steps=0):
        """Admit an instance; returns its id, or None when it does not fit now."""
        lx, ly, lz = (int(v) for v in dims)
        nat = np.ascontiguousarray(nature, np.int8)
        fe = np.ascontiguousarray(fext, np.float64)
        i = _i32()
        st = lib().lbxb_add(self._h, lx, ly, lz, _ptr(nat), float(tau), float(rho0), _ptr(fe), float(target_error),
                            int(max_lb_steps), float(Db), float(ka), float(kd), int(mp_steps), ctypes.byref(i))
        if st == ERR_NOMEM:
            return None
        _check(st, "lbxb_add")
        self.dims[i.value] = (lx, ly, lz)
        return i.value

    def remove(self, iid):
        _check(lib().lbxb_remove(self._h, int(iid)), "lbxb_remove")
        self.dims.pop(iid, None)

    def round(self, lb_steps=64, mp_steps=64):
        n = _i32()
        _check(lib().lbxb_round(self._h, int(lb_steps), int(mp_steps), ctypes.byref(n)), "lbxb_round")
        return n.value

    def info(self, iid):
        ph, st, er = _i32(), _i32(), _i32()
        lb, mp = _i64(), _i64()
        _check(lib().lbxb_info(self._h, int(iid), ctypes.byref(ph), ctypes.byref(st), ctypes.byref(lb),
                               ctypes.byref(er), ctypes.byref(mp)), "lbxb_info")
        return dict(phase=ph.value, status=st.value, lb_steps=lb.value, exit_reason=er.value, mp_steps=mp.value)

    def result(self, iid, vacf_rows=0, flow=False):
        d = np.zeros(3)
        v = np.zeros(3 * max(vacf_rows, 1))
        lx, ly, lz = self.dims[iid]
        N = lx * ly * lz
        fl = [np.zeros(N) for _ in range(4)] if flow else [None] * 4
        _check(lib().lbxb_result(self._h, int(iid), _ptr(d), _ptr(v) if vacf_rows else None, int(vacf_rows),
                                 *[_ptr(a) for a in fl]), "lbxb_result")
        out = dict(deff=d, vacf=v[:3 * vacf_rows].reshape(-1, 3))
        if flow:
            out["rho"], out["j"] = fl[0], np.stack(fl[1:])
        return out

    def set_block(self, lb_block, mp_block):
        _check(lib().lbxb_set_block(self._h, int(lb_block), int(mp_block)), "lbxb_set_block")

    def set_floor_window(self, steps):
        """lbx_set_floor_window for the instances added from now on."""
        _check(lib().lbxb_set_floor_window(self._h, int(steps)), "lbxb_set_floor_window")

    def set_check_interval(self, steps):
        """lbxb_set_check_interval: the FP32-storage forms' stage-2 l2err interval."""
        _check(lib().lbxb_set_check_interval(self._h, int(steps)), "lbxb_set_check_interval")

    def autotune(self, launches=10):
        """lbxb_autotune: block size per phase (64/128/256 nodes) and launch
        mode (per-step, graph, persistent); the chosen point and every timing
        (mean ms per step)."""
        lb, mp = _i32(), _i32()
        t = np.zeros(AUTOTUNE_TIMES)
        _check(lib().lbxb_autotune(self._h, int(launches), ctypes.byref(lb), ctypes.byref(mp), _ptr(t)),
               "lbxb_autotune")
        keys = ("lb64", "lb128", "lb256", "mp64", "mp128", "mp256", "lb_graph", "mp_graph", "lb_persistent",
                "mp_persistent")
        return dict(lb_block=lb.value, mp_block=mp.value, mode=self.launch_mode,
                    ms={k: float(v) for k, v in zip(keys, t)})

    def set_launch_mode(self, mode):
        _check(lib().lbxb_set_launch_mode(self._h, int(mode)), "lbxb_set_launch_mode")

    def set_addressing(self, wide):
        """As Lattice.set_addressing, for the batch's step kernels."""
        _check(lib().lbxb_set_addressing(self._h, int(wide)), "lbxb_set_addressing")

    @property
    def addressing(self):
        return int(lib().lbxb_addressing(self._h))

    @property
    def launch_mode(self):
        return int(lib().lbxb_launch_mode(self._h))

    def set_graphs(self, on):
        self.set_launch_mode(LAUNCH_GRAPH if on else LAUNCH_STREAM)

    def stats(self):
        lbms, mpms = _f64(), _f64()
        a, b_, c, d = _i64(), _i64(), _i64(), _i64()
        _check(lib().lbxb_stats(self._h, ctypes.byref(lbms), ctypes.byref(a), ctypes.byref(b_), ctypes.byref(mpms),
                                ctypes.byref(c), ctypes.byref(d)), "lbxb_stats")
        return dict(lb_ms=lbms.value, lb_box_updates=a.value, lb_fluid_updates=b_.value, mp_ms=mpms.value,
                    mp_box_updates=c.value, mp_fluid_updates=d.value)

    def free_tiles(self):
        return int(lib().lbxb_free_tiles(self._h))
