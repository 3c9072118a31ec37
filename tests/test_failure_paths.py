"""Failure handling of the pressure solve and the state check, on the CPU (host-check build of the
product's CG and driver):

- a solve that stops at its iteration limit is repeated from x = 0 with four times the limit, then
  soft-accepted when the true residual is below 1e4 rtol or 1e3 atol, or failed with
  -(1000 + KSP_DIVERGED_ITS) = -997 (src/solver/pressure_backend_petsc.c:878-937);
- solver_state_is_finite (src/solver/mod_fv_solver.f90:3925-3945) flags NaN/Inf in U or p.
"""
import numpy as np
import pytest

from dopamine_amd import cases, params as P
from hostcheck import HostSolver, cg_solve_retry


def poisson_ell(n):
    """7-point Laplacian on an n^3 box with a Dirichlet diagonal shift on the x = 0 face (SPD):
    ELL slots [slot][row]."""
    N = n ** 3
    nbr = -np.ones((6, N), dtype=np.int32)
    coef = np.zeros((6, N))
    diag = np.zeros(N)
    idx = np.arange(N).reshape(n, n, n)
    k = 0
    for ax in range(3):
        for sh in (-1, 1):
            nb = np.roll(idx, sh, axis=ax)
            valid = np.ones_like(idx, dtype=bool)
            sl = [slice(None)] * 3
            sl[ax] = 0 if sh == 1 else n - 1
            valid[tuple(sl)] = False
            nbr[k][valid.reshape(-1)] = nb[valid].reshape(-1)
            coef[k][valid.reshape(-1)] = 1.0
            diag += valid.reshape(-1)
            k += 1
    diag[idx[0].reshape(-1)] += 2.0  # Dirichlet face
    return nbr, coef, diag


def dense(nbr, coef, diag):
    N = diag.size
    A = np.diag(diag)
    for k in range(nbr.shape[0]):
        for i in range(N):
            if nbr[k, i] >= 0:
                A[i, nbr[k, i]] -= coef[k, i]
    return A


@pytest.mark.parametrize("single", [False, True])
def test_converged_solve_is_not_retried(single):
    nbr, coef, diag = poisson_ell(8)
    b = np.random.default_rng(1).standard_normal(diag.size)
    rc, x, its, res, total, retried = cg_solve_retry(nbr, coef, diag, b, np.zeros_like(b), 1e-8, 1e-50, 2000,
                                                     single=single)
    assert rc == 0 and retried == 0 and its == total and 0 < its < 200
    assert np.linalg.norm(b - dense(nbr, coef, diag) @ x) == pytest.approx(res, rel=1e-6)
    assert res <= 1e-8 * np.linalg.norm(b)


@pytest.mark.parametrize("single", [False, True])
def test_limit_hit_then_retry_from_zero_with_four_times_the_limit(single):
    nbr, coef, diag = poisson_ell(8)
    b = np.random.default_rng(2).standard_normal(diag.size)
    _, _, need, _, _, _ = cg_solve_retry(nbr, coef, diag, b, np.zeros_like(b), 1e-8, 1e-50, 2000, single=single)
    maxit = need // 2  # the first solve stops at the limit, 4 x the limit is enough
    # a garbage initial guess: the retry must start from x = 0, not from the failed iterate
    x0 = np.full_like(b, 1.0e6)
    rc, x, its, res, total, retried = cg_solve_retry(nbr, coef, diag, b, x0, 1e-8, 1e-50, maxit, single=single)
    assert rc == 0 and retried == 1
    assert its == need  # the retry from zero repeats the plain solve exactly
    assert total == maxit + need
    assert res <= 1e-8 * np.linalg.norm(b)


def test_retry_that_also_stops_at_its_limit_fails_with_minus_997():
    nbr, coef, diag = poisson_ell(8)
    b = np.random.default_rng(3).standard_normal(diag.size)
    rc, x, its, res, total, retried = cg_solve_retry(nbr, coef, diag, b, np.zeros_like(b), 1e-8, 1e-50, 3)
    assert retried == 1 and its == 12 and total == 15
    assert res > 1e-4  # neither 1e4 rtol nor 1e3 atol
    assert rc == -(1000 + (-3))


def test_retry_that_stops_near_the_tolerance_is_soft_accepted():
    nbr, coef, diag = poisson_ell(8)
    b = 1.0e-7 * np.random.default_rng(4).standard_normal(diag.size)  # ||b|| ~ 2e-6 < 1e4 rtol = 1e-4
    rc, x, its, res, total, retried = cg_solve_retry(nbr, coef, diag, b, np.zeros_like(b), 1e-8, 1e-50, 3)
    assert retried == 1 and res > 1e-8 * np.linalg.norm(b)  # not converged ...
    assert res < 1e-8 * 1.0e4 and rc == 0  # ... but accepted


def test_mixed_precision_solve_takes_the_same_retry():
    nbr, coef, diag = poisson_ell(8)
    b = np.random.default_rng(5).standard_normal(diag.size)
    rc, x, its, res, total, retried = cg_solve_retry(nbr, coef, diag, b, np.zeros_like(b), 1e-8, 1e-50, 20,
                                                     mixed=True)
    assert retried == 1 and rc == 0 and res <= 1e-8 * np.linalg.norm(b)


def test_state_is_finite_flags_nan_and_inf():
    case = cases.small_channel(8, 6, 4)
    mesh = case.make_mesh()
    s = HostSolver(mesh, case.params)
    P.initialize(s, case, mesh)
    assert s.state_is_finite()
    U = s.U.copy()
    U[5, 1] = np.nan
    s.array("U")[:] = U.reshape(-1)
    assert not s.state_is_finite()
    U[5, 1] = 0.0
    s.array("U")[:] = U.reshape(-1)
    assert s.state_is_finite()
    s.array("p")[3] = np.inf
    assert not s.state_is_finite()


def test_learned_with_mixed_precision_is_refused_before_any_device_call():
    from dopamine_amd.solver import check_precision
    p = cases.small_channel(8, 6, 4).params
    p.precond = "LEARNED"
    with pytest.raises(ValueError, match="learned"):
        check_precision(p, 1)
    check_precision(p, 0)
    p.precond = "JACOBI"
    check_precision(p, 1)
    with pytest.raises(ValueError):
        check_precision(p, 2)
