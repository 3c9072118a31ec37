"""The dynamic Smagorinsky model (sgs_model 'DYNAMIC_SMAG', src/sgs/mod_dynamic_smagorinsky.f90):
the oracle pinned by closed forms, the product's host build against the oracle.

For a linear velocity u_j = a_j + sum_i B_ij x_i on a uniform grid, every cell whose test-filter
stencils stay inside the box (three layers from the walls) sees
  L_ij = (2 h^2 / 7) (B^T B)_ij (the seven-cell top-hat of a linear field), S^ = S,
  M_ij = -6 Delta^2 |S| S_ij (ALPHA_SQ = 4),
so nu_t = max(0, <L M> / <M M>) Delta^2 |S| = max(0, -(h^2 / 21) sum(B^T B : S) / sum(S : S)),
independent of Delta; a planted factor (the 2 of M, the 4 of ALPHA_SQ, a filter weight) breaks it."""
import numpy as np
import pytest

import oracle
from dopamine_amd import cases, mesh as M
from dopamine_amd.params import SolverParams
from helpers import rel_l2, run_case
from hostcheck import HostSolver

N = 10
H = 1.0 / N


def _linear(B):
    mesh = M.uniform_box(N, N, N, 1.0, 1.0, 1.0, group_types=("noSlip",) * 6)
    p = SolverParams(nu=1e-3, use_sgs=True, sgs_model="DYNAMIC_SMAG", wall_model="none", cell_projection="LS")
    o = oracle.Solver(mesh, p)
    a = np.array([0.3, -0.2, 0.1])
    xc = mesh.cell_centroid.reshape(-1, 3)
    xf = mesh.face_centre.reshape(-1, 3)[mesh.n_int_faces:]
    o.array("U")[:] = (a + xc @ B).reshape(-1)
    o.array("U_bnd")[:] = (a + xf @ B).reshape(-1)
    o.array("grad_U")[:] = o.op_gradient_vector()
    return mesh, o, xc


def _deep(xc):
    lo, hi = 3 * H - 1e-9, 1.0 - 3 * H + 1e-9
    return np.all((xc > lo) & (xc < hi), axis=1)


@pytest.mark.parametrize("B", [
    -0.7 * np.eye(3),
    np.array([[-1.1, 0.4, 0.0], [0.2, -0.5, 0.3], [0.0, -0.6, -0.2]]),
    np.array([[-0.8, 1.5, 0.2], [-0.3, 0.1, 0.9], [0.4, 0.0, -0.9]]),
])
def test_linear_field_closed_form(B):
    mesh, o, xc = _linear(B)
    nu = o.op_sgs(4, 0.0, False, 1e-12)[_deep(xc)]
    S = 0.5 * (B + B.T)
    want = max(0.0, -(H * H / 21.0) * np.sum((B.T @ B) * S) / np.sum(S * S))
    assert want > 0.0
    assert np.abs(nu - want).max() < 1e-12 * want


def test_backscatter_is_clipped():
    B = 0.6 * np.eye(3)  # sum(B^T B : S) > 0: a negative Lilly ratio
    mesh, o, xc = _linear(B)
    nu = o.op_sgs(4, 0.0, False, 1e-12)
    assert np.all(nu[_deep(xc)] == 0.0)
    assert np.all(nu >= 0.0)


def test_stability_cap():
    B = -0.7 * np.eye(3)
    mesh, o, xc = _linear(B)
    G = o.array("Gf").reshape(-1, 3, 3)
    d2 = np.linalg.det(G) ** (1.0 / 3.0)
    dt = 1.0e3 * H * H  # cap 0.4 Delta^2 / dt below the model value
    nu = o.op_sgs(4, 0.0, False, dt)
    assert np.allclose(nu[_deep(xc)], np.minimum(0.7 * H * H / 21.0, 0.4 * d2[_deep(xc)] / dt), rtol=1e-12)


def _channel_dynamic():
    case = cases.small_channel(16, 12, 8)
    case.params.sgs_model = "DYNAMIC_SMAG"
    case.params.p_tol = 1e-12
    return case


def test_host_dynamic_smagorinsky_matches_oracle():
    case = _channel_dynamic()
    mesh = case.make_mesh()
    h, o = HostSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (h, o):
        run_case(s, case, mesh, nsteps=3, stats_every=3)
    nu = o.array("nu_sgs")
    assert nu.max() > 0.0
    for name in ("U", "p", "nu_sgs"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name
