"""The oracle's single-operator routines (om_op_*), which the Fortran-module GPU tests compare
against, pinned on the CPU: each equals what the oracle's own step pipeline computes for the same
state (gradient, SGS viscosity and wall traction of eval_closures), and the closed forms a plausible
mistake would break: least-squares and Green-Gauss gradients of a linear field, zero advection
of a uniform field by a divergence-free flux, zero diffusion of a uniform field."""
import numpy as np
import pytest

import oracle
from dopamine_amd import cases, mesh as M, params as P


def _state(case, noise=0.05):
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    P.initialize(o, case, mesh)
    o.advance()
    U = o.U.reshape(-1, 3)
    rng = np.random.default_rng(2)
    o.set_velocity(U + noise * max(np.abs(U).max(), 1.0) * rng.standard_normal(U.shape))
    o.eval_closures(1.0e-3)
    return mesh, o


@pytest.fixture(scope="module")
def channel():
    return _state(cases.small_channel(8, 8, 6))


def test_ops_equal_the_step_pipeline(channel):
    mesh, o = channel
    assert np.array_equal(o.op_gradient_vector(gg=False), o.array("grad_U"))  # LS is the case's scheme
    p = o.params
    assert np.array_equal(o.op_sgs(1, p.c_vreman, False, 1.0e-3), o.array("nu_sgs"))  # Vreman
    tau = o.op_wall_model()
    assert np.abs(tau).max() > 0.0
    assert np.array_equal(tau, o.array("tau_w_bnd"))


def test_ops_leave_the_solver_state_alone(channel):
    mesh, o = channel
    before = {k: o.array(k).copy() for k in ("nu_sgs", "grad_U", "tau_w_bnd", "U")}
    o.op_sgs(3, 1.0 / 3.0, True, 0.0)
    o.op_sgs(2, 0.1, False, 0.0)
    o.op_wall_model()
    o.op_advection_vector(5)
    for k, v in before.items():
        assert np.array_equal(v, o.array(k)), k


@pytest.mark.parametrize("gg", [False, True])
def test_gradient_of_a_linear_field_is_exact(gg):
    """U = a + B x with boundary values from the same linear field: the LS gradient is exact on the
    jittered tetrahedra of the all-wall cavity, the Green-Gauss one (exact face values on a uniform
    hex grid) on every cell of a uniform walled box, boundary cells included."""
    if gg:
        case = cases.regression_cases()["hydrowall"]
        mesh = M.uniform_box(6, 5, 4, 1.0, 1.0, 1.0, group_types=("noSlip",) * 6)
    else:
        case = cases.cavity_tets(4)
        mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    B = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.5, 0.9, 1.1]])  # U_j = a_j + sum_i B[i, j] x_i
    a = np.array([0.2, -0.1, 0.4])
    xc = mesh.cell_centroid.reshape(-1, 3)
    xf = mesh.face_centre.reshape(-1, 3)[mesh.n_int_faces:]
    o.array("U")[:] = (a + xc @ B).reshape(-1)
    o.array("U_bnd")[:] = (a + xf @ B).reshape(-1)
    g = o.op_gradient_vector(gg=gg, iterations=2 if gg else 1).reshape(-1, 3, 3)  # g[c, j, i] = dU_j/dx_i
    assert np.abs(g - B.T).max() < 1e-10 * np.abs(B).max()
    c = np.array([0.4, -0.8, 2.5])
    gs = o.op_gradient_scalar(1.5 + xc @ c, 1.5 + xf @ c, gg=gg, iterations=2 if gg else 1).reshape(-1, 3)
    assert np.abs(gs - c).max() < 1e-10 * np.abs(c).max()


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_uniform_field_has_no_advection_or_diffusion(scheme):
    """Periodic skewed box, uniform U and a divergence-free face flux (each cell's net flux 0):
    every scheme reconstructs the uniform value, so the advection rhs vanishes; the viscous term of
    a uniform field with a zero gradient vanishes."""
    case = cases.regression_cases()["tgv_skew"]
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    u0 = np.array([0.7, -0.3, 0.2])
    o.array("U")[:] = np.tile(u0, mesh.n_cells)
    o.array("U_bnd")[:] = np.tile(u0, o.array("U_bnd").size // 3)
    o.array("grad_U")[:] = 0.0
    sf = mesh.face_sf.reshape(-1, 3)
    o.array("U_face")[:] = sf @ u0
    net = np.zeros(mesh.n_cells)
    own, nb = mesh.face_owner, mesh.face_neighbour
    np.add.at(net, own, sf @ u0)
    inner = nb >= 0
    np.add.at(net, nb[inner], -(sf @ u0)[inner])
    assert np.abs(net).max() < 1e-12
    r = o.op_advection_vector(scheme)
    assert np.abs(r).max() < 1e-9 * np.abs(u0).max() ** 2 / mesh.cell_volume.min() ** (1 / 3)
    o.array("nu_sgs")[:] = 0.01
    v = o.op_viscous_vector()
    assert np.abs(v).max() < 1e-9
