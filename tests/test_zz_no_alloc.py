"""fvm_petsc_solve_pressure_noalloc, the allocation-free pressure solve, against the allocating
entry point fvm_petsc_solve_pressure with the same Jacobi CG: identical iterates (bit for bit) on
two calls in a row, the second with a smaller operator, and a clear error beyond its maximum size.
Device buffers for rhs and solution come from the HIP runtime directly (the caller's memory)."""
import ctypes

import numpy as np
import pytest

import oracle
from dopamine_amd import cases
from dopamine_amd._lib import lib

pytestmark = pytest.mark.gpu
_P = ctypes.POINTER
_ARGS = [ctypes.c_int] * 3 + [_P(ctypes.c_int)] * 5 + [_P(ctypes.c_double), ctypes.c_void_p, _P(ctypes.c_double)] + \
    [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P(ctypes.c_int), _P(ctypes.c_double),
     ctypes.c_int, ctypes.c_void_p, _P(ctypes.c_int), _P(ctypes.c_double)]


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime, for the test's own device buffers."""

    def __init__(self):
        self.L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.L.hipMalloc.argtypes = [_P(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.empty(n)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _operator(name):
    case = cases.regression_cases()[name]
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    return (mesh.n_cells, o.int_array("ps_owner").copy(), o.int_array("ps_nb").copy(), o.int_array("ps_skip").copy(),
            o.array("ps_coeff").copy(), o.array("ps_diag").copy())


def _solve(fn, n, own, nb, skip, coeff, diag, b_dev, x_dev, rtol, n_dir):
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    l2g, fo, fnl, sk = i32(np.arange(1, n + 1)), i32(own + 1), i32(nb + 1), i32(skip)
    dg, cf = f64(diag), f64(coeff)
    drows, dvals = i32(np.zeros(max(1, n_dir))), f64(np.zeros(max(1, n_dir)))
    its, res = ctypes.c_int(0), ctypes.c_double(0.0)
    ip = lambda a: a.ctypes.data_as(_P(ctypes.c_int))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(_P(ctypes.c_double))  # noqa: E731
    fn.argtypes = _ARGS
    fn.restype = ctypes.c_int
    rc = fn(n, n, len(own), ip(l2g), ip(fo), ip(fnl), ip(fnl), ip(sk), dp(dg), b_dev, dp(cf), 1, 1000, rtol, 1e-300,
            ip(drows), dp(dvals), n_dir, x_dev, ctypes.byref(its), ctypes.byref(res))
    return rc, its.value, res.value


def test_noalloc_solve_matches_the_allocating_entry_point():
    L = lib()
    hip = Hip()
    L.fvm_petsc_set_pressure_solver.argtypes = [ctypes.c_char_p]
    L.fvm_petsc_insert_option.argtypes = [ctypes.c_char_p]
    L.fvm_petsc_set_io_on_device.argtypes = [ctypes.c_int]
    L.dopamine_backend_last_error.restype = ctypes.c_char_p
    assert L.fvm_petsc_backend_initialize() == 0
    L.fvm_petsc_insert_option(b"-ksp_type cg")
    L.fvm_petsc_set_pressure_solver(b"jacobi")  # the recurrence the allocation-free path runs
    L.fvm_petsc_set_io_on_device(1)
    try:
        for name, n_dir in (("tgv_skew", 0), ("hydro", 1)):  # 4096 then 1728 cells
            n, own, nb, skip, coeff, diag = _operator(name)
            b = np.random.default_rng(17).standard_normal(n)
            if n_dir == 0:
                b -= b.mean()
            print(f"[gpu] no-alloc solve on {name}: {n} cells", flush=True)
            b_dev, xr_dev, xn_dev = hip.upload(b), hip.upload(np.zeros(n)), hip.upload(np.full(n, 7.0))
            try:
                rc1, its1, res1 = _solve(L.fvm_petsc_solve_pressure, n, own, nb, skip, coeff, diag, b_dev, xr_dev,
                                         1e-10, n_dir)
                # a stale initial guess in the solution buffer is the caller's x0: start both from zero
                hip.free(xn_dev)
                xn_dev = hip.upload(np.zeros(n))
                rc2, its2, res2 = _solve(L.fvm_petsc_solve_pressure_noalloc, n, own, nb, skip, coeff, diag, b_dev,
                                         xn_dev, 1e-10, n_dir)
                xr, xn = hip.download(xr_dev, n), hip.download(xn_dev, n)
            finally:
                for p in (b_dev, xr_dev, xn_dev):
                    hip.free(p)
            print(f"[gpu] {name}: allocating rc {rc1} its {its1} resid {res1:.3e}; no-alloc rc {rc2} its {its2} "
                  f"resid {res2:.3e}; max |dx| {np.abs(xr - xn).max():.1e}", flush=True)
            assert rc1 == 0 and rc2 == 0
            assert its1 == its2 and res1 == res2
            assert np.array_equal(xr, xn)
            assert res2 <= 1e-10 * np.linalg.norm(b) * 1.001
        # beyond the scratch: refused with its reason, before any device work
        n_big = 131073
        z = np.zeros(1)
        rc = _solve(L.fvm_petsc_solve_pressure_noalloc, n_big, np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64),
                    np.zeros(1, dtype=np.int64), z, np.ones(n_big), None, None, 1e-8, 0)[0]
        assert rc == -61 and b"131072" in L.dopamine_backend_last_error()
    finally:
        L.fvm_petsc_set_io_on_device(0)
        L.fvm_petsc_set_pressure_solver(b"boomeramg")
        L.fvm_petsc_backend_finalize()
