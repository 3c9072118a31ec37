"""The Fortran boundary on the GPU.

- tests/fortran/pressure_dropin.x declares the pressure backend with the reference's own bind(C)
  interface block (src/solver/mod_pressure_solver.f90:23-74) and calls fvm_petsc_solve_pressure
  from Fortran; its answer must equal, bit for bit, the same call made through ctypes, and solve
  the system (oracle CG / sparse direct solve).
- tests/fortran/libdopamine_ftest.so calls the drop-in modules (mod_gradient, mod_advection,
  mod_viscous, mod_vreman, mod_amd, mod_wall_model, FVSolver) with the reference's argument lists;
  each routine's result must equal the direct C entry point's bit for bit and the oracle's single
  routine on the same fields to FP64 round-off (relative L2 below 1e-12; the oracle sums faces in
  face order, the kernels per cell).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from dopamine_amd import cases, params as P
from dopamine_amd._lib import lib
from dopamine_amd.solver import FVSolver
from helpers import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN = os.path.join(HERE, "fortran", "pressure_dropin.x")
FTEST = os.path.join(HERE, "fortran", "libdopamine_ftest.so")
_P = ctypes.POINTER
_vp = ctypes.c_void_p
TOL = 1e-12


def _ftest():
    lib()  # the product library first, so the harness binds the same instance
    L = ctypes.CDLL(FTEST)
    ci, cd = ctypes.c_int, ctypes.c_double
    for name, args in {
        "ft_gradient_vector": [_vp, ci, ci, _vp, _vp, _vp],
        "ft_gradient_scalar": [_vp, ci, ci, _vp, _vp, _vp],
        "ft_advection_vector": [_vp, ci, _vp, _vp, _vp, _vp, _vp],
        "ft_viscous_vector": [_vp, cd, _vp, _vp, _vp, _vp, _vp],
        "ft_advection_scalar": [_vp, ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "ft_viscous_scalar": [_vp, _vp, cd, _vp, cd, _vp, _vp, _vp],
        "ft_sgs": [_vp, ci, _vp, _vp, _vp, cd, cd, ci, cd, _vp],
        "ft_wall_model": [_vp, _vp, cd, cd, cd, cd, cd, _vp],
        "ft_solver_run": [_vp, ci, cd, _vp, _vp, _P(cd)],
    }.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ci
    return L


def _a(x):
    return None if x is None else x.ctypes.data


def _c_op(name, *args):
    """The direct C entry point (include/dopamine_hip.h) with numpy arrays / scalars."""
    L = lib()
    L.dopamine_last_error.restype = ctypes.c_char_p
    f = getattr(L, name)
    f.restype = ctypes.c_int
    conv = []
    for a in args:
        if isinstance(a, _vp):
            conv.append(a)
        elif isinstance(a, np.ndarray) or a is None:
            conv.append(_vp(_a(a)))
        elif isinstance(a, float):
            conv.append(ctypes.c_double(a))
        else:
            conv.append(ctypes.c_int(a))
    rc = f(*conv)
    assert rc == 0, L.dopamine_last_error()
    return rc


class _State:
    """A case advanced one step by the oracle, with noise on U, its closures evaluated, and the
    GPU solver of the same mesh (its handle carries the device mesh)."""

    def __init__(self, case, noise=0.05, seed=11, dt_stab=0.0):
        self.case = case
        self.mesh = case.make_mesh()
        self.o = oracle.Solver(self.mesh, case.params)
        P.initialize(self.o, case, self.mesh)
        self.o.advance()
        rng = np.random.default_rng(seed)
        U = self.o.U.reshape(-1, 3)
        scale = max(np.abs(U).max(), 1.0)
        self.o.set_velocity(U + noise * scale * rng.standard_normal(U.shape))
        self.o.eval_closures(dt_stab)
        self.g = FVSolver(self.mesh, case.params)
        self.h = _vp(self.g.h)
        self.nc = self.mesh.n_cells
        self.U = self.o.array("U").copy()
        self.Ub = self.o.array("U_bnd").copy()
        self.nb = max(1, self.Ub.size // 3)
        self.ff = self.o.array("U_face").copy()
        self.gU = self.o.array("grad_U").copy()
        self.nus = self.o.array("nu_sgs").copy()
        self.phi = rng.standard_normal(self.nc)
        self.pb = rng.standard_normal(self.nb)


STATES = {}


def _state(name):
    if name not in STATES:
        if name == "channel":
            STATES[name] = _State(cases.small_channel(16, 12, 8))
        elif name == "tets":
            STATES[name] = _State(cases.cavity_tets(5), noise=0.2)
        else:
            STATES[name] = _State(cases.regression_cases()[name])
    return STATES[name]


MESHES = ["channel", "tgv_skew", "tets"]


@pytest.mark.parametrize("gg", [0, 1])
@pytest.mark.parametrize("name", MESHES)
def test_fortran_gradients(name, gg):
    s = _state(name)
    print(f"[gpu] fortran gradients {name} gg={gg}", flush=True)
    F = _ftest()
    its = s.case.params.gg_iterations if gg else 1
    g = np.zeros(9 * s.nc)
    assert F.ft_gradient_vector(s.h, gg, its, _a(s.U), _a(s.Ub), _a(g)) == 0
    ref = s.o.op_gradient_vector(gg=bool(gg), iterations=its)
    assert rel_l2(ref, g) < TOL
    c = np.zeros(9 * s.nc)
    if gg:
        _c_op("dopamine_op_gg_gradient_vector", s.h, s.U, s.Ub, None, c)
        for _ in range(2, its + 1):
            _c_op("dopamine_op_gg_gradient_vector", s.h, s.U, s.Ub, c.copy(), c)
    else:
        _c_op("dopamine_op_ls_gradient_vector", s.h, s.U, s.Ub, c)
    assert np.array_equal(c, g)
    gs = np.zeros(3 * s.nc)
    assert F.ft_gradient_scalar(s.h, gg, its, _a(s.phi), _a(s.pb), _a(gs)) == 0
    refs = s.o.op_gradient_scalar(s.phi, s.pb, gg=bool(gg), iterations=its)
    assert rel_l2(refs, gs) < TOL


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("name", MESHES)
def test_fortran_advection_vector(name, scheme):
    s = _state(name)
    print(f"[gpu] fortran advection_rhs_vector {name} scheme {scheme}", flush=True)
    F = _ftest()
    rhs = np.full(3 * s.nc, 7.0)  # overwritten
    assert F.ft_advection_vector(s.h, scheme, _a(s.U), _a(s.Ub), _a(s.ff), _a(s.gU), _a(rhs)) == 0
    ref = s.o.op_advection_vector(scheme)
    assert rel_l2(ref, rhs) < TOL
    c = np.zeros(3 * s.nc)
    _c_op("dopamine_op_advection_rhs_vector", s.h, scheme, s.U, s.Ub, s.ff, s.gU, c)
    assert np.array_equal(c, rhs)


@pytest.mark.parametrize("name", MESHES)
def test_fortran_viscous_vector(name):
    s = _state(name)
    print(f"[gpu] fortran viscous_rhs_vector {name}", flush=True)
    F = _ftest()
    base = np.random.default_rng(3).standard_normal(3 * s.nc)
    rhs = base.copy()
    assert F.ft_viscous_vector(s.h, s.case.params.nu, _a(s.U), _a(s.nus), _a(s.gU), _a(s.Ub), _a(rhs)) == 0
    ref = s.o.op_viscous_vector()
    assert rel_l2(ref, rhs - base) < 1e-11  # adds to rhs: the difference carries the base's rounding
    assert rel_l2(ref + base, rhs) < TOL


@pytest.mark.parametrize("scheme", [1, 2, 3, 5])
@pytest.mark.parametrize("name", MESHES)
def test_fortran_scalar_transport(name, scheme):
    s = _state(name)
    print(f"[gpu] fortran scalar advection/viscous {name} scheme {scheme}", flush=True)
    F = _ftest()
    gphi = s.o.op_gradient_scalar(s.phi, s.pb)
    rhs = np.full(s.nc, -3.0)
    assert F.ft_advection_scalar(s.h, scheme, _a(s.phi), _a(s.U), _a(s.pb), _a(s.Ub), _a(s.ff), _a(gphi),
                                 _a(rhs)) == 0
    ref = s.o.op_advection_scalar(scheme, s.phi, s.pb, gphi)
    assert rel_l2(ref, rhs) < TOL
    alpha, prt = 2.0e-3, 0.85
    rv = np.zeros(s.nc)
    assert F.ft_viscous_scalar(s.h, _a(s.phi), alpha, _a(s.nus), prt, _a(gphi), _a(s.pb), _a(rv)) == 0
    refv = s.o.op_viscous_scalar(s.phi, s.pb, gphi, alpha, prt)
    assert rel_l2(refv, rv) < TOL


@pytest.mark.parametrize("model", [1, 2, 3, 4])
@pytest.mark.parametrize("name", MESHES)
def test_fortran_sgs(name, model):
    s = _state(name)
    print(f"[gpu] fortran sgs model {model} {name}", flush=True)
    F = _ftest()
    coef = {1: 0.025, 2: 0.1, 3: 1.0 / 3.0, 4: 0.0}[model]
    nu = s.case.params.nu
    for dt_stab in (0.0, 1.0e-3):
        nus = np.zeros(s.nc)
        assert F.ft_sgs(s.h, model, _a(s.U), _a(s.Ub), _a(s.gU), coef, nu, 1, dt_stab, _a(nus)) == 0
        ref = s.o.op_sgs(model, coef, True, dt_stab)
        assert rel_l2(ref, nus) < TOL, dt_stab
    if model in (1, 3):  # without grad_U_ext the routines take the LS gradient of U first
        nus2 = np.zeros(s.nc)
        assert F.ft_sgs(s.h, model, _a(s.U), _a(s.Ub), None, coef, nu, 1, 0.0, _a(nus2)) == 0
        g_ls = s.o.op_gradient_vector()
        s.o.array("grad_U")[:] = g_ls
        ref2 = s.o.op_sgs(model, coef, True, 0.0)
        s.o.array("grad_U")[:] = s.gU
        assert rel_l2(ref2, nus2) < TOL


def test_fortran_wall_model():
    s = _state("channel")
    print("[gpu] fortran apply_wall_model_bc channel", flush=True)
    F = _ftest()
    tau = np.zeros(3 * s.nb)
    assert F.ft_wall_model(s.h, _a(s.U), s.case.params.nu, -1.0, 0.0, 0.0, 0.0, _a(tau)) == 0
    ref = s.o.op_wall_model()
    assert np.abs(ref).max() > 0.0
    assert rel_l2(ref, tau) < TOL


@pytest.mark.parametrize("t_filt,tf_fac", [(0.02, 0.0), (0.0, 2.0)])
def test_fortran_wall_model_sampling_height_and_filter(t_filt, tf_fac):
    """apply_wall_model_bc with h_wm, t_filt / tf_fac and dt: the first call seeds the filter, the
    second moves it once towards the new sampled velocity; both against the oracle's routine."""
    from helpers import wallmodel_case
    case = wallmodel_case(0.12, t_filt, tf_fac)
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    P.initialize(o, case, mesh)
    o.advance()
    g = FVSolver(mesh, case.params)
    print(f"[gpu] fortran apply_wall_model_bc h_wm 0.12 T {t_filt} f {tf_fac}", flush=True)
    F = _ftest()
    nb = o.array("tau_w_bnd").size // 3
    dt, nu = case.params.dt, case.params.nu
    U1 = o.array("U").copy()
    U2 = U1 * (1.0 + 0.1 * np.random.default_rng(9).standard_normal(U1.size))
    f = oracle.Solver(mesh, case.params)  # a fresh filter (the stepped oracle's is seeded)
    for U in (U1, U2):
        tau = np.zeros(3 * nb)
        assert F.ft_wall_model(_vp(g.h), _a(U), nu, 0.12, t_filt, dt, tf_fac, _a(tau)) == 0
        f.array("U")[:] = U
        ref = f.op_wall_model(dt)
        assert np.abs(ref).max() > 0.0
        assert rel_l2(ref, tau) < TOL


def test_fortran_fvsolver_advance_matches_python_driver():
    """FVSolver%advance from Fortran and the Python driver over the same C ABI give the same state."""
    case = cases.regression_cases()["tgv_skew"]
    mesh = case.make_mesh()
    a, b = FVSolver(mesh, case.params), FVSolver(mesh, case.params)
    for s in (a, b):
        P.initialize(s, case, mesh)
    print("[gpu] fortran FVSolver%advance tgv_skew 3 steps", flush=True)
    F = _ftest()
    U = np.zeros(3 * mesh.n_cells)
    p = np.zeros(mesh.n_cells)
    t = ctypes.c_double(0.0)
    assert F.ft_solver_run(_vp(a.h), 3, case.params.dt, _a(U), _a(p), ctypes.byref(t)) == 0
    for _ in range(3):
        b.advance()
    assert np.array_equal(U, b.U.reshape(-1))
    assert np.array_equal(p, b.p.reshape(-1))
    assert t.value == b.time


# ---------------------------------------------------------------- pressure backend from Fortran

def _system(name):
    case = cases.regression_cases()[name]
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    return (mesh.n_cells, o.int_array("ps_owner").copy(), o.int_array("ps_nb").copy(),
            o.int_array("ps_skip").copy(), o.array("ps_coeff").copy(), o.array("ps_diag").copy())


def _write_input(path, n, own, nb, skip, coeff, diag, b, x0, selector, option, rtol, atol, maxit, n_dir):
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).tobytes()  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).tobytes()  # noqa: E731
    nf = len(own)
    with open(path, "wb") as fh:
        fh.write(i32([n, nf, n_dir, 1, maxit]))
        fh.write(f64([rtol, atol]))
        fh.write(i32([len(selector)]) + selector.encode())
        fh.write(i32([len(option)]) + option.encode())
        fh.write(i32(np.arange(1, n + 1)) + i32(own + 1) + i32(nb + 1) + i32(nb + 1) + i32(skip))
        fh.write(f64(diag) + f64(b) + f64(coeff))
        fh.write(i32(np.zeros(max(1, n_dir))))
        fh.write(f64(np.zeros(max(1, n_dir))) + f64(x0))


def _read_output(path, n):
    raw = open(path, "rb").read()
    ints = np.frombuffer(raw[:20], dtype=np.int32)
    resid = np.frombuffer(raw[20:28], dtype=np.float64)[0]
    x = np.frombuffer(raw[28:28 + 8 * n], dtype=np.float64).copy()
    return dict(available=int(ints[0]), initialize=int(ints[1]), rc=int(ints[2]), its=int(ints[3]),
                finalize=int(ints[4]), resid=float(resid), x=x)


def _ctypes_solve(n, own, nb, skip, coeff, diag, b, x0, selector, option, rtol, atol, maxit, n_dir):
    L = lib()
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    arrs = [i32(np.arange(1, n + 1)), i32(own + 1), i32(nb + 1), i32(nb + 1), i32(skip), f64(diag), f64(b),
            f64(coeff), i32(np.zeros(max(1, n_dir))), f64(np.zeros(max(1, n_dir)))]
    x = f64(x0).copy()
    its, res = ctypes.c_int(0), ctypes.c_double(0.0)
    L.fvm_petsc_backend_initialize()
    L.fvm_petsc_set_pressure_solver.argtypes = [ctypes.c_char_p]
    L.fvm_petsc_set_pressure_solver(selector.encode())
    if option:
        L.fvm_petsc_insert_option.argtypes = [ctypes.c_char_p]
        L.fvm_petsc_insert_option(option.encode())
    L.fvm_petsc_solve_pressure.restype = ctypes.c_int
    ip = lambda a: a.ctypes.data_as(_P(ctypes.c_int))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(_P(ctypes.c_double))  # noqa: E731
    L.fvm_petsc_solve_pressure.argtypes = [ctypes.c_int] * 3 + [_P(ctypes.c_int)] * 5 + [_P(ctypes.c_double)] * 3 + \
        [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P(ctypes.c_int), _P(ctypes.c_double),
         ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int), _P(ctypes.c_double)]
    rc = L.fvm_petsc_solve_pressure(n, n, len(own), ip(arrs[0]), ip(arrs[1]), ip(arrs[2]), ip(arrs[3]), ip(arrs[4]),
                                    dp(arrs[5]), dp(arrs[6]), dp(arrs[7]), 1, maxit, rtol, atol, ip(arrs[8]),
                                    dp(arrs[9]), n_dir, dp(x), ctypes.byref(its), ctypes.byref(res))
    L.fvm_petsc_backend_finalize()
    return rc, x, its.value, res.value


# options persist in the backend (as in PETSc's options database): every case names its CG type
@pytest.mark.parametrize("selector,option", [("jacobi", "-ksp_type cg"), ("jacobi", "-ksp_type pipecg"),
                                             ("boomeramg", "-ksp_type cg"), ("mixed-multigrid", "-ksp_type cg")])
@pytest.mark.parametrize("name", ["tgv_skew", "hydrowall"])
def test_fortran_caller_of_fvm_petsc_solve_pressure(name, selector, option, tmp_path):
    n, own, nb, skip, coeff, diag = _system(name)
    keep = skip == 0
    A = sp.coo_matrix((np.concatenate([-coeff[keep], -coeff[keep]]),
                       (np.concatenate([own[keep], nb[keep]]), np.concatenate([nb[keep], own[keep]]))),
                      shape=(n, n)).tocsr() + sp.diags(diag)
    n_dir = int(np.sum(np.abs(np.asarray(A.sum(axis=1)).ravel()) > 1e-12 * diag))
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n)
    if n_dir == 0:
        b -= b.mean()
    x0 = np.zeros(n)
    rtol, atol, maxit = 1e-10, 1e-300, 2000
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, n, own, nb, skip, coeff, diag, b, x0, selector, option, rtol, atol, maxit, n_dir)
    print(f"[gpu] fortran pressure_dropin {name} {selector} {option!r}: {n} cells", flush=True)
    proc = subprocess.run([DROPIN, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(proc.stdout.strip(), proc.stderr.strip()[-400:], flush=True)
    assert proc.returncode == 0
    r = _read_output(out, n)
    assert r["available"] == 1 and r["initialize"] == 0 and r["finalize"] == 0
    assert r["rc"] == 0 and r["its"] > 0
    res_true = np.linalg.norm(b - A @ r["x"])
    assert abs(r["resid"] - res_true) <= 1e-8 * np.linalg.norm(b)
    assert res_true <= 10 * rtol * np.linalg.norm(b)
    rc, x, its, res = _ctypes_solve(n, own, nb, skip, coeff, diag, b, x0, selector, option, rtol, atol, maxit, n_dir)
    assert rc == 0 and its == r["its"]
    assert np.array_equal(x, r["x"]) and res == r["resid"]
    if n_dir:
        xd = spla.spsolve(A.tocsc(), b)
    else:
        Af = A.tolil()
        Af[0, :] = 0.0
        Af[0, 0] = 1.0
        bd = b.copy()
        bd[0] = 0.0
        xd = spla.spsolve(Af.tocsc(), bd)
        xd -= xd.mean()
        x = r["x"] - r["x"].mean()
        assert rel_l2(xd, x) < 1e-6
        return
    assert rel_l2(xd, r["x"]) < 1e-6
