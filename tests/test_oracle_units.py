"""Pins for the oracle functions the reference's regression CSVs leave open (the regression cases
run neither Vreman, Smagorinsky, the log-law wall model nor Green-Gauss gradients), checked
against what the mathematics fixes: closed forms of the published models on fields whose
discrete gradient is exact, a published law inverted, and a direct sparse solve."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from dopamine_amd import cases, mesh as M
from dopamine_amd.params import SolverParams

WALLS = ("noSlip",) * 6
N = 10
H = 1.0 / N


def _params(**kw):
    base = dict(nu=1e-3, advection_scheme="CDS", cell_projection="LS", wall_model="none", use_sgs=True,
                sgs_model="VREMAN", c_vreman=0.07, gradient_scheme="LEAST_SQUARES", p_tol=1e-10)
    base.update(kw)
    return SolverParams(**base)


def _closures(params, A, mesh=None):
    """Oracle gradients / nu_t of the linear field u = A x on a uniform walled box."""
    mesh = mesh or M.uniform_box(N, N, N, 1.0, 1.0, 1.0, group_types=WALLS)
    o = oracle.Solver(mesh, params)
    x = mesh.cell_centroid.reshape(-1, 3)
    o.set_velocity(x @ np.asarray(A).T)
    o.eval_closures(1e-12)  # stability cap far above the model value
    return mesh, o


def _inner(mesh, layers):
    """Cells at least `layers` cells away from every wall."""
    x = mesh.cell_centroid.reshape(-1, 3)
    lo, hi = (layers + 0.5) * H - 1e-9, 1.0 - (layers + 0.5) * H + 1e-9
    return np.all((x >= lo) & (x <= hi), axis=1)


def _grad(o, mesh):
    # grad_U[9c + 3j + i] = d u_j / d x_i, so [c, j, i] = A_ji for u = A x
    return o.array("grad_U").reshape(-1, 3, 3)


@pytest.mark.parametrize("scheme,iters", [("LEAST_SQUARES", 1), ("GREEN_GAUSS", 1), ("GREEN_GAUSS", 3)])
def test_gradients_exact_for_linear_fields(scheme, iters):
    A = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [-0.4, 0.9, 1.5]])  # u_i = A_ij x_j
    mesh, o = _closures(_params(gradient_scheme=scheme, gg_iterations=iters, use_sgs=False), A)
    g = _grad(o, mesh)[_inner(mesh, 1)]
    assert np.abs(g - A).max() < 1e-12


def test_least_squares_gradient_exact_on_skewed_cells():
    A = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [-0.4, 0.9, 1.5]])
    mesh = M.skewed_box(8, 0.06, group_types=WALLS)
    o = oracle.Solver(mesh, _params(use_sgs=False))
    x = mesh.cell_centroid.reshape(-1, 3)
    o.set_velocity(x @ A.T)
    o.eval_closures(1e-12)
    # walls carry u = 0 at the boundary, so only cells away from the boundary see a linear field
    lo = x.min(axis=0) + 2.0 / 8
    hi = x.max(axis=0) - 2.0 / 8
    inner = np.all((x > lo) & (x < hi), axis=1)
    assert inner.sum() > 20
    assert np.abs(_grad(o, mesh)[inner] - A).max() < 1e-11


def test_vreman_vanishes_in_pure_shear():
    """Vreman (2004): B_beta = 0 for a one-dimensional shear, so nu_t = 0 (laminar shear)."""
    mesh, o = _closures(_params(), [[0, 2.5, 0], [0, 0, 0], [0, 0, 0]])
    nu = o.array("nu_sgs")[_inner(mesh, 3)]
    assert np.abs(nu).max() < 1e-14


@pytest.mark.parametrize("a", [1.0, -3.0])
def test_vreman_axisymmetric_strain_closed_form(a):
    """u = (a x, -a y/2, -a z/2) on cubes of side h (filter tensor h^2 I):
    nu_t = c sqrt(B_beta / alpha:alpha) = c h^2 |a| sqrt(3/8)."""
    c = 0.07
    mesh, o = _closures(_params(c_vreman=c), np.diag([a, -a / 2, -a / 2]))
    nu = o.array("nu_sgs")[_inner(mesh, 3)]
    assert np.allclose(nu, c * H * H * abs(a) * math.sqrt(3.0 / 8.0), rtol=1e-12)
    # filter tensor of an interior cube cell
    G = o.array("Gf").reshape(-1, 3, 3)[_inner(mesh, 3)]
    assert np.allclose(G, H * H * np.eye(3), rtol=1e-12, atol=1e-18)


def test_vreman_general_gradient_anisotropic_cells():
    """Vreman (2004, Phys. Fluids 16:3670) eqs. 5-7 for a general, non-symmetric gradient on
    rectangular cells hx x hy x hz: alpha_ij = du_j/dx_i, beta = alpha^T G alpha,
    nu_t = c sqrt(B_beta / alpha:alpha), B_beta the sum of beta's principal 2x2 minors. The filter
    tensor G is proportional to sum_faces A dx dx^T = V/2 diag(h) with det G = V^2, so
    G = V^(1/3) diag(hx, hy, hz). With anisotropic G a transposed alpha changes B_beta."""
    c = 0.07
    n = np.array([12, 8, 10])
    h = 1.0 / n
    V = h.prod()
    mesh = M.uniform_box(*n, 1.0, 1.0, 1.0, group_types=WALLS)
    A = np.array([[0.2, 1.7, -0.4], [-0.3, 0.5, 0.9], [1.1, -0.6, -0.7]])  # u = A x, traceless
    G = V ** (1.0 / 3.0) * np.diag(h)

    def vreman(alpha):
        beta = alpha.T @ G @ alpha
        Bb = (beta[0, 0] * beta[1, 1] - beta[0, 1] ** 2 + beta[0, 0] * beta[2, 2] - beta[0, 2] ** 2
              + beta[1, 1] * beta[2, 2] - beta[1, 2] ** 2)
        return c * math.sqrt(Bb / np.sum(alpha * alpha))

    expect = vreman(A.T)
    assert abs(vreman(A) - expect) > 1e-3 * expect
    o = oracle.Solver(mesh, _params(c_vreman=c))
    x = mesh.cell_centroid.reshape(-1, 3)
    o.set_velocity(x @ A.T)
    o.eval_closures(1e-12)
    inner = np.all((x > 3 * h) & (x < 1 - 3 * h), axis=1)
    assert inner.sum() == np.prod(n - 6)
    assert np.allclose(o.array("Gf").reshape(-1, 3, 3)[inner], G, rtol=1e-12, atol=1e-18)
    assert np.allclose(o.array("nu_sgs")[inner], expect, rtol=1e-12)


def test_vreman_stability_cap():
    """nu_t <= 0.4 V^(2/3) / dt_stab."""
    mesh = M.uniform_box(N, N, N, 1.0, 1.0, 1.0, group_types=WALLS)
    o = oracle.Solver(mesh, _params(c_vreman=0.07))
    x = mesh.cell_centroid.reshape(-1, 3)
    o.set_velocity(x @ np.diag([1e3, -5e2, -5e2]).T)
    o.eval_closures(1.0)
    assert np.allclose(o.array("nu_sgs")[_inner(mesh, 3)], 0.4 * H * H, rtol=1e-12)


def test_smagorinsky_shear_closed_form():
    """u = (g y, 0, 0): |S| = sqrt(2 S:S) = g, nu_t = (C_s h)^2 g with h = det(G)^(1/6)."""
    g, cs = 2.5, 0.1
    mesh, o = _closures(_params(sgs_model="SMAGORINSKY", cs_smagorinsky=cs), [[0, g, 0], [0, 0, 0], [0, 0, 0]])
    nu = o.array("nu_sgs")[_inner(mesh, 3)]
    assert np.allclose(nu, (cs * H) ** 2 * g, rtol=1e-12)


KAPPA, B = 0.41, 5.2  # log-law constants of the reference (src/mod_parameters.f90)


def _reichardt(yp):
    """Reichardt (1951) composite law with C chosen so that u+ -> ln(y+)/kappa + B."""
    c = B - math.log(KAPPA) / KAPPA
    return math.log(1 + KAPPA * yp) / KAPPA + c * (1 - math.exp(-yp / 11) - yp / 11 * math.exp(-0.33 * yp))


@pytest.mark.parametrize("yplus", [0.5, 5.0, 30.0, 300.0, 3000.0])
def test_wall_model_inverts_reichardt_law(yplus):
    nu, utau = 1.0 / 395.0, 0.8
    y = yplus * nu / utau
    u = utau * _reichardt(yplus)
    tau = oracle.wall_tau(u, y, nu)
    assert math.isclose(tau, utau * utau, rel_tol=1e-7)


def test_wall_model_limits():
    nu = 1e-3
    # viscous sublayer: tau -> nu u / y
    assert math.isclose(oracle.wall_tau(1e-4, 1e-3, nu), nu * 1e-4 / 1e-3, rel_tol=1e-3)
    # log layer: u+ = ln(y+)/kappa + B at large y+
    utau, yp = 0.05, 2e4
    u = utau * (math.log(yp) / KAPPA + B)
    assert math.isclose(oracle.wall_tau(u, yp * nu / utau, nu), utau ** 2, rel_tol=2e-3)
    assert oracle.wall_tau(0.0, 0.1, nu) == 0.0


def _pressure_operator(name):
    case = cases.regression_cases()[name]
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    own, nb = o.int_array("ps_owner").copy(), o.int_array("ps_nb").copy()
    coeff, diag = o.array("ps_coeff").copy(), o.array("ps_diag").copy()
    n = mesh.n_cells
    A = sp.coo_matrix((np.concatenate([diag, -coeff, -coeff]),
                       (np.concatenate([np.arange(n), own, nb]), np.concatenate([np.arange(n), nb, own]))),
                      shape=(n, n)).tocsr()
    return own, nb, coeff, diag, A


@pytest.mark.parametrize("name", ["hydro", "tgv_skew"])
def test_cg_jacobi_matches_direct_solve(name):
    own, nb, coeff, diag, A = _pressure_operator(name)
    n = len(diag)
    nullspace = np.allclose(np.asarray(A.sum(axis=1)).ravel(), 0.0, atol=1e-12 * diag.max())
    b = np.random.default_rng(3).standard_normal(n)
    if nullspace:
        b -= b.mean()
    x, its, res, rc = oracle.cg_jacobi(own, nb, coeff, diag, b, np.zeros(n), 1e-12, 1e-300, 5000, nullspace)
    assert rc == 0 and res <= 1e-12 * np.linalg.norm(b) * 1.0001
    assert math.isclose(res, np.linalg.norm(b - A @ x), rel_tol=1e-3)  # round-off of a 1e-11 residual
    if nullspace:
        e0 = sp.coo_matrix((np.ones(1), ([0], [0])), shape=(n, n))
        xd = spla.spsolve((A + e0).tocsc(), b)
        xd, x = xd - xd.mean(), x - x.mean()
    else:
        xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_pressure_operator_is_a_symmetric_m_matrix():
    """Row sums vanish except on rows touching a fixed-pressure outlet, where they are the outlet
    coefficients A/d (hydro has outlets at z0 and zL)."""
    own, nb, coeff, diag, A = _pressure_operator("hydro")
    assert np.all(coeff >= 0) and np.all(diag > 0)
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.all(rs >= -1e-12 * diag.max())
    assert 0 < np.count_nonzero(rs > 1e-9) < len(diag)


# Known answers of the reference's tests/unit/test_amd.f90:17-50 (C = 1/3, tolerance 1e-12).
C_AMD = 1.0 / 3.0


@pytest.mark.parametrize("label,alpha,G,c,want", [
    ("pure shear", [[0, 0, 0], [3, 0, 0], [0, 0, 0]], np.eye(3), C_AMD, 0.0),
    ("solid-body rotation", [[0, 2, 0], [-2, 0, 0], [0, 0, 0]], np.eye(3), C_AMD, 0.0),
    ("isotropic filter strain", np.diag([1.0, 1.0, -2.0]), np.eye(3), C_AMD, C_AMD),
    ("anisotropic filter strain", np.diag([1.0, 1.0, -2.0]), np.diag([4.0, 1.0, 1.0]), C_AMD, 0.5 * C_AMD),
    ("backscatter clipped", -np.diag([1.0, 1.0, -2.0]), np.eye(3), C_AMD, 0.0),
    ("zero gradient", np.zeros((3, 3)), np.eye(3), C_AMD, 0.0),
    ("constant scaling", np.diag([5.0, 5.0, -10.0]), np.eye(3), 2 * C_AMD, 2 * C_AMD * 5.0),
])
def test_amd_point_known_answers(label, alpha, G, c, want):
    nu = oracle.amd_point(alpha, G, c)
    assert abs(nu - want) <= 1e-12 * max(1.0, abs(want)), label


def test_amd_on_mesh_axisymmetric_compression():
    """u = (a x, -a y/2, -a z/2), a < 0, on cubes (G = h^2 I): beta = h^2 alpha^2,
    -beta:alpha = -(3/4) h^2 a^3, alpha:alpha = (3/2) a^2, nu_t = -C h^2 a / 2."""
    a = -2.0
    mesh, o = _closures(_params(sgs_model="AMD", c_amd=C_AMD, wall_model="none"), np.diag([a, -a / 2, -a / 2]))
    nu = o.array("nu_sgs")[_inner(mesh, 3)]
    assert np.allclose(nu, -C_AMD * H * H * a / 2, rtol=1e-12)
    # stretching along x (a > 0) is backscatter for AMD: clipped to zero
    mesh, o = _closures(_params(sgs_model="AMD", c_amd=C_AMD, wall_model="none"), np.diag([-a, a / 2, a / 2]))
    assert np.abs(o.array("nu_sgs")[_inner(mesh, 3)]).max() == 0.0


def test_threaded_cg_is_the_same_jacobi_cg():
    """The oracle's OpenMP CG (the CPU baseline's) against its sequential CG on a regression
    operator: the same iteration count to within one and the same solution to the tolerance."""
    import oracle
    case = cases.regression_cases()["hydrowall"]
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    own, nb, skip = o.int_array("ps_owner").copy(), o.int_array("ps_nb").copy(), o.int_array("ps_skip").copy()
    coeff, diag = o.array("ps_coeff").copy(), o.array("ps_diag").copy()
    keep = skip == 0
    b = np.random.default_rng(9).standard_normal(mesh.n_cells)
    b -= b.mean()
    x1, it1, r1, rc1 = oracle.cg_jacobi(own[keep], nb[keep], coeff[keep], diag, b, np.zeros_like(b), 1e-10, 1e-300,
                                        2000, True)
    try:
        oracle.set_threads(4)
        x4, it4, r4, rc4 = oracle.cg_jacobi(own[keep], nb[keep], coeff[keep], diag, b, np.zeros_like(b), 1e-10,
                                            1e-300, 2000, True)
    finally:
        oracle.set_threads(1)
    assert rc1 == rc4 == 0 and abs(it1 - it4) <= 1
    assert np.linalg.norm(x4 - x1) <= 1e-8 * np.linalg.norm(x1)
