"""Host-check solver (test infrastructure): the product's kernel bodies and RK3W driver
(dopamine_amd/csrc/solver_impl.h) compiled for the CPU, so CPU tests can compare them with the
oracle. Never used by the product."""
import ctypes
import os

import numpy as np

from dopamine_amd.solver import MeshDesc, Params, STATS_COLUMNS, build_abi

_P = ctypes.POINTER
_d, _i64 = ctypes.c_double, ctypes.c_int64
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhostcheck.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C tests/hostcheck`")
        L = ctypes.CDLL(LIB_PATH)
        L.hc_solver_create.restype = ctypes.c_void_p
        L.hc_solver_create.argtypes = [_P(MeshDesc), _P(Params)]
        L.hc_solver_destroy.argtypes = [ctypes.c_void_p]
        L.hc_last_error.restype = ctypes.c_char_p
        L.hc_solver_field.restype = _P(_d)
        L.hc_solver_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _P(_i64)]
        L.hc_solver_apply_all_bc.argtypes = [ctypes.c_void_p]
        L.hc_solver_project_once.argtypes = [ctypes.c_void_p, _d]
        L.hc_solver_advance.argtypes = [ctypes.c_void_p, _d, _P(_d)]
        L.hc_solver_stats.argtypes = [ctypes.c_void_p, _d * 8]
        L.hc_solver_time.restype = _d
        L.hc_solver_time.argtypes = [ctypes.c_void_p]
        L.hc_solver_counters.argtypes = [ctypes.c_void_p, _i64 * 4]
        L.hc_solver_nonortho_info.argtypes = [ctypes.c_void_p, ctypes.c_int32 * 3]
        L.hc_solver_proj_counters.argtypes = [ctypes.c_void_p, _i64 * 2]
        L.hc_solver_stats_csv.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _d, _P(_d), ctypes.c_int]
        L.hc_solver_cell_divergence.argtypes = [ctypes.c_void_p, _d * 2]
        L.hc_solver_write_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.hc_solver_read_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.hc_solver_state_is_finite.argtypes = [ctypes.c_void_p]
        L.hc_solver_retries.restype = _i64
        L.hc_solver_retries.argtypes = [ctypes.c_void_p]
        L.hc_solver_error.restype = ctypes.c_char_p
        L.hc_solver_error.argtypes = [ctypes.c_void_p]
        _pi = _P(ctypes.c_int)
        L.hc_amg_solve.argtypes = [ctypes.c_int, ctypes.c_int, _pi, _P(_d), _P(_d), _P(_d), _P(_d), _d, _d,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(_d), _pi, _P(_d), _pi]
        L.hc_amg_vcycle.argtypes = [ctypes.c_int, ctypes.c_int, _pi, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d), _pi,
                                    _pi]
        L.hc_cg_solve_retry.argtypes = [ctypes.c_int, ctypes.c_int, _pi, _P(_d), _P(_d), _P(_d), _P(_d), _d, _d,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _pi, _P(_d), _pi, _pi]
        L.hc_coef_roundtrip.argtypes = [_P(ctypes.c_float), _P(ctypes.c_uint16), _P(ctypes.c_float), ctypes.c_int]
        _lib = L
    return _lib


def coef_roundtrip(values):
    """The multigrid's binary16 coupling storage: (half bits, value read back)."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    h = np.zeros(v.size, dtype=np.uint16)
    out = np.zeros(v.size, dtype=np.float32)
    lib().hc_coef_roundtrip(v.ctypes.data_as(_P(ctypes.c_float)), h.ctypes.data_as(_P(ctypes.c_uint16)),
                            out.ctypes.data_as(_P(ctypes.c_float)), v.size)
    return h, out


class HostSolver:
    def __init__(self, mesh, params, precision=0):
        params.validate()
        self.params = params
        self.md, self.pr, self._keep = build_abi(mesh, params, precision)
        self.h = lib().hc_solver_create(ctypes.byref(self.md), ctypes.byref(self.pr))
        if not self.h:
            raise RuntimeError(lib().hc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().hc_solver_destroy(self.h)
            self.h = None

    def array(self, name):
        n = _i64(0)
        p = lib().hc_solver_field(self.h, name.encode(), ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def get(self, name):
        return self.array(name).copy()

    @property
    def U(self):
        return self.array("U").reshape(-1, 3)

    @property
    def p(self):
        return self.array("p")

    def set_velocity(self, U):
        self.array("U")[:] = np.asarray(U, dtype=np.float64).reshape(-1)
        lib().hc_solver_apply_all_bc(self.h)

    def set_temperature(self, T):
        self.array("T")[:] = np.asarray(T, dtype=np.float64).reshape(-1)
        lib().hc_solver_apply_all_bc(self.h)

    def set_scalar(self, k, C):
        self.array(f"C{k}")[:] = np.asarray(C, dtype=np.float64).reshape(-1)
        lib().hc_solver_apply_all_bc(self.h)

    def scalar(self, k):
        return self.array(f"C{k}").copy()

    def cell_divergence(self):
        out = (_d * 2)()
        lib().hc_solver_cell_divergence(self.h, out)
        return {"max_abs_divU_cell": out[0], "volavg_abs_divU_cell": out[1]}

    def stats_csv(self, path, step, dt, E0=0.0, append=True):
        e0 = _d(E0)
        if lib().hc_solver_stats_csv(self.h, str(path).encode(), step, dt, ctypes.byref(e0), int(append)) != 0:
            raise OSError(f"cannot write {path}")
        return e0.value

    def write_checkpoint(self, directory, step, rank=0):
        if lib().hc_solver_write_checkpoint(self.h, str(directory).encode(), step, rank) != 0:
            raise OSError(lib().hc_last_error().decode())

    def read_checkpoint(self, directory, step, rank=0):
        rc = lib().hc_solver_read_checkpoint(self.h, str(directory).encode(), step, rank)
        if rc < 0:
            raise OSError(lib().hc_last_error().decode())
        return rc

    def project_once(self, dt):
        return lib().hc_solver_project_once(self.h, dt)

    def advance(self, dt=None):
        used = _d(0.0)
        rc = lib().hc_solver_advance(self.h, self.params.dt if dt is None else dt, ctypes.byref(used))
        if rc != 0:
            raise RuntimeError(f"advance failed {rc}: {lib().hc_solver_error(self.h).decode()}")
        return used.value

    def state_is_finite(self):
        return lib().hc_solver_state_is_finite(self.h) == 1

    @property
    def retries(self):
        return lib().hc_solver_retries(self.h)

    @property
    def time(self):
        return lib().hc_solver_time(self.h)

    def stats(self):
        out = (_d * 8)()
        lib().hc_solver_stats(self.h, out)
        return dict(zip(STATS_COLUMNS, list(out)))

    def nonortho_info(self):
        out = (ctypes.c_int32 * 3)()
        lib().hc_solver_nonortho_info(self.h, out)
        return {"skip_nonortho": out[0], "skip_grad_p": out[1], "n_nonortho_corr": out[2]}

    def proj_counters(self):
        out = (_i64 * 2)()
        lib().hc_solver_proj_counters(self.h, out)
        return {"projections": out[0], "iterations": out[1]}

    def counters(self):
        out = (_i64 * 4)()
        lib().hc_solver_counters(self.h, out)
        return {"solves": out[0], "iterations": out[1], "last": out[2], "sweeps": out[3]}


def cg_solve_retry(nbr, coef, diag, b, x0, rtol, atol, maxit, nullspace=False, single=False, mixed=False):
    """The product CG with the reference's failure handling (solve_with_retry) on an ELL operator
    (nbr/coef as [slot][row] arrays); returns (rc, x, its, resid, total_its, retried)."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    w, n = nbr.shape
    its, total, retried = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    resid = _d(0.0)
    pi, pd = _P(ctypes.c_int), _P(_d)
    rc = lib().hc_cg_solve_retry(n, w, nbr.ctypes.data_as(pi), coef.ctypes.data_as(pd), diag.ctypes.data_as(pd),
                                 b.ctypes.data_as(pd), x.ctypes.data_as(pd), rtol, atol, maxit, int(nullspace),
                                 int(single), int(mixed), ctypes.byref(its), ctypes.byref(resid),
                                 ctypes.byref(total), ctypes.byref(retried))
    return rc, x, its.value, resid.value, total.value, retried.value


# theta, omega, alpha, coarse_max, theta0 (< 0: off): the product's defaults except a coarsest
# level of 400 cells (the product's 1500), so the small test operators keep several levels
AMG_DEFAULTS = (0.08, 0.8, 1.8, 400, -1.0)


def amg_solve(nbr, coef, diag, b, rtol, atol=1e-50, maxit=2000, nullspace=False, mixed=False, cfg=AMG_DEFAULTS):
    """The product's flexible CG preconditioned by the aggregation multigrid (host-check build);
    returns (rc, x, its, resid, level sizes, refinement sweeps)."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    w, n = nbr.shape
    its = ctypes.c_int(0)
    resid = _d(0.0)
    info = (ctypes.c_int * 16)()
    c = (_d * 5)(*(tuple(cfg) + (-1.0,) * (5 - len(cfg))))
    pi, pd = _P(ctypes.c_int), _P(_d)
    rc = lib().hc_amg_solve(n, w, nbr.ctypes.data_as(pi), coef.ctypes.data_as(pd), diag.ctypes.data_as(pd),
                            b.ctypes.data_as(pd), x.ctypes.data_as(pd), rtol, atol, maxit, int(nullspace), int(mixed),
                            c, ctypes.byref(its), ctypes.byref(resid), info)
    if rc == -100:
        raise RuntimeError(lib().hc_last_error().decode())
    return rc, x, its.value, resid.value, list(info[1:1 + info[0]]), info[15]


def ell_off16(nbr):
    """The 16-bit offset form of an ELL neighbour list (host-check build): (W, base[W], d[W, n],
    src[W, n]), W = 0 when the offsets do not fit."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    w, n = nbr.shape
    base = (ctypes.c_int * 8)()
    d = np.zeros((8, n), dtype=np.int16)
    src = np.zeros((8, n), dtype=np.int32)
    L = lib()
    L.hc_ell_off16.argtypes = [ctypes.c_int, ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), ctypes.c_void_p,
                               _P(ctypes.c_int)]
    W = L.hc_ell_off16(n, w, nbr.ctypes.data_as(_P(ctypes.c_int)), base, d.ctypes.data, src.ctypes.data_as(_P(ctypes.c_int)))
    return W, np.array(base[:W]), d.reshape(-1)[: W * n].reshape(W, n), src.reshape(-1)[: W * n].reshape(W, n)


def amg_vcycle(nbr, coef, diag, r, cfg=AMG_DEFAULTS):
    """One V-cycle z = M r (host-check build); returns (z, level-0 aggregates, level sizes)."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    w, n = nbr.shape
    agg = np.zeros(n, dtype=np.int32)
    info = (ctypes.c_int * 16)()
    c = (_d * 5)(*(tuple(cfg) + (-1.0,) * (5 - len(cfg))))
    pi, pd = _P(ctypes.c_int), _P(_d)
    rc = lib().hc_amg_vcycle(n, w, nbr.ctypes.data_as(pi), coef.ctypes.data_as(pd), diag.ctypes.data_as(pd),
                             r.ctypes.data_as(pd), z.ctypes.data_as(pd), c, agg.ctypes.data_as(pi), info)
    if rc != 0:
        raise RuntimeError(lib().hc_last_error().decode())
    return z, agg, list(info[1:1 + info[0]])
