"""Pins of the oracle's convective outlet, backflow traction, wall-model sampling height and time
filter, and per-patch MOST roughness (src/bc/mod_bc.f90:681-818, src/sgs/mod_wall_model.f90:
126-313, src/abl/mod_abl_wall.f90:249-262).

- outlet_dn_u: the reference's own unit test (tests/unit/test_outlet_bc.f90) mirrored: second
  order in space on an orthogonal offset (the one-sided formula is first order), at least first
  order and more accurate on a skewed one, and third order in time under the Wray RK3 weights;
- the outlet in a run: a convective outlet holds a uniform through-flow exactly; the traction
  never raises the outlet pressure, vanishes where the flow leaves and acts where it enters;
- the sampling walk: the cell and distance it must reach on a uniform grid;
- the filter: the recursion it must follow between steps (once per step), fixed and adaptive;
- the roughness: each wall patch's faces follow the MOST solve with that patch's lengths.
"""
import math

import numpy as np
import pytest

import oracle
from dopamine_amd import cases, mesh as M, params as P
from dopamine_amd.params import SolverParams

RK_A = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)  # Wray (1990) low-storage RK3
RK_B = (0.0, -17.0 / 60.0, -5.0 / 12.0)
RK_C = (0.0, 8.0 / 15.0, 2.0 / 3.0)


def _fld(x):
    return math.sin(1.3 * x[0] + 0.4 * x[1]) + 0.5 * math.cos(0.7 * x[2] + x[0])


def _grd(x):
    return np.array([1.3 * math.cos(1.3 * x[0] + 0.4 * x[1]) - 0.5 * math.sin(0.7 * x[2] + x[0]),
                     0.4 * math.cos(1.3 * x[0] + 0.4 * x[1]), -0.35 * math.sin(0.7 * x[2] + x[0])])


@pytest.mark.parametrize("skew", [False, True])
def test_outlet_face_derivative_order(skew):
    xo = np.array([0.3, 0.2, 0.1])
    nrm = np.array([1.0, 0.0, 0.0])
    e_new, e_old = [], []
    for i in range(4):
        h = 0.4 / 2 ** i
        xb = xo + h * np.array([0.5, 0.0, 0.0])
        if skew:
            xb = xb + h * np.array([0.0, 0.15, -0.1])
        dl = xb - xo
        len_d = np.linalg.norm(dl)
        dl = dl / len_d
        nd = nrm @ dl
        tn = nrm - nd * dl
        exact = nrm @ _grd(xb)
        dn = oracle.outlet_dn_u(_fld(xb), _fld(xo), _grd(xo) @ dl, _grd(xo) @ tn, len_d, nd)
        e_new.append(abs(dn - exact))
        e_old.append(abs(nd * (_fld(xb) - _fld(xo)) / len_d + _grd(xo) @ tn - exact))
    o_new = math.log(e_new[2] / e_new[3]) / math.log(2.0)
    o_old = math.log(e_old[2] / e_old[3]) / math.log(2.0)
    if skew:
        assert o_new > 0.9 and e_new[3] < e_old[3]
    else:
        assert o_new > 1.9 and o_old < 1.1


def test_outlet_boundary_state_is_third_order_in_time():
    uc, wn, tend, h = 1.3, 3.0, 1.0, 0.2

    def integrate(nstep):
        dt = tend / nstep
        ub = math.sin(wn * 0.5 * h)
        gprev = 0.0
        for istep in range(nstep):
            for k in range(3):
                t = istep * dt + RK_C[k] * dt
                g = -uc * oracle.outlet_dn_u(ub, math.sin(wn * (-uc * t)), wn * math.cos(wn * (-uc * t)), 0.0,
                                             0.5 * h, 1.0)
                ub = ub + dt * (RK_A[k] * g + RK_B[k] * gprev)
                gprev = g
        return ub

    ref = integrate(4096)
    e = [abs(integrate(16 * 2 ** i) - ref) for i in range(4)]
    order = math.log(e[2] / e[3]) / math.log(2.0)
    assert 2.8 < order < 3.3


def _outlet_box(beta=None, mode="convective", n=8, uz=0.4):
    """Uniform 8^3 box, periodic in x and y, inlet at z0 with velocity (0, 0, uz), outlet at zL."""
    mesh = M.uniform_box(n, n, n, 1.0, 1.0, 1.0,
                         group_types=("periodic", "periodic", "periodic", "periodic", "inlet", "outlet"))
    p = SolverParams(nu=1e-3, advection_scheme="CDS", use_sgs=False, wall_model="none", dt=5e-3, cfl_max=0.5,
                     p_tol=1e-12, cell_projection="LS", patch_U={"z0": (0.0, 0.0, uz)}, patch_p={"zL": 0.0},
                     patch_outlet_mode={"zL": mode})
    if beta is not None:
        p.patch_backflow_beta = {"zL": beta}
    return mesh, p


def _outlet_faces(mesh, name):
    q = list(mesh.patch_names).index(name)
    return np.arange(mesh.patch_start[q], mesh.patch_start[q] + mesh.patch_nfaces[q])


def test_convective_outlet_holds_a_uniform_through_flow():
    mesh, p = _outlet_box()
    o = oracle.Solver(mesh, p)
    o.set_velocity(np.tile([0.0, 0.0, 0.4], (mesh.n_cells, 1)))
    o.project_once(p.dt)
    for _ in range(5):
        o.advance()
    ub = o.array("U_bnd").reshape(-1, 3)[_outlet_faces(mesh, "zL")]
    assert np.abs(ub - [0.0, 0.0, 0.4]).max() < 1e-12
    assert np.abs(o.U - [0.0, 0.0, 0.4]).max() < 1e-12
    uc, uref = o.outlet_speeds()[list(mesh.patch_names).index("zL")]
    assert uc == pytest.approx(0.4, rel=1e-12) and uref == pytest.approx(0.4, rel=1e-12)


def test_convective_outlet_advects_the_boundary_state():
    """A sheared outflow: the convective faces' velocity moves off the extrapolated value by the
    RK3 integral of -U_c n.grad u, and the run stays divergence-free."""
    mesh, p = _outlet_box()
    o = oracle.Solver(mesh, p)
    e = oracle.Solver(mesh, SolverParams(**{**p.__dict__, "patch_outlet_mode": {}}))
    x = mesh.cell_centroid.reshape(-1, 3)
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = 0.2 * np.sin(2 * np.pi * x[:, 1]) * x[:, 2]
    U[:, 2] = 0.4
    for s in (o, e):
        s.set_velocity(U)
        s.project_once(p.dt)
        for _ in range(4):
            s.advance()
    f = _outlet_faces(mesh, "zL")
    d = np.abs(o.array("U_bnd").reshape(-1, 3)[f] - e.array("U_bnd").reshape(-1, 3)[f]).max()
    assert d > 1e-6
    assert o.stats()["max_abs_divU"] < 1e-9


def test_backflow_traction_acts_only_on_entering_flow():
    mesh, p = _outlet_box(beta=1.0, mode="extrapolate")
    o = oracle.Solver(mesh, p)
    x = mesh.cell_centroid.reshape(-1, 3)
    U = np.zeros((mesh.n_cells, 3))
    U[:, 2] = 0.4 + 0.6 * np.sin(2 * np.pi * x[:, 0])  # re-enters through part of the outlet
    o.set_velocity(U)
    o.project_once(p.dt)
    o.advance()
    f = _outlet_faces(mesh, "zL")
    pb = o.array("outlet_pb")[f]
    own = mesh.face_owner[mesh.n_int_faces + f]
    u = o.U.reshape(-1, 3)[own]
    un = u[:, 2]  # outward normal +z
    leaving = un > 0.3
    entering = un < -0.05
    assert leaving.any() and entering.any()
    assert np.all(pb <= 0.0)  # p_val = 0; the traction only lowers it
    assert np.abs(pb[leaving]).max() < 1e-8
    # entering: S0 > (1 + tanh 2) / 2 = 0.98, so p_b ~ -beta |u|^2 / 2, with u the last stage's
    # predictor velocity (the step's final u differs from it by the last projection's correction)
    u2 = np.sum(u[entering] ** 2, axis=1)
    assert np.all(pb[entering] < -0.35 * u2) and np.all(pb[entering] > -0.7 * u2)
    # the outlet's Dirichlet pressure is the traction value
    assert np.abs(o.array("p_bnd")[f] - pb).max() < 1e-12


def test_wall_model_sampling_walk_reaches_the_height():
    """Uniform 6 x 10 x 4 box, walls at y = 0 and 1 (dy = 0.1): h_wm = 0.22 walks from the
    first cell (y 0.05) to the third (y 0.25); h_wm = 0.05 stays; h_wm beyond the half height stops
    after 20 steps at most, inside the box."""
    mesh = M.uniform_box(6, 10, 4, 1.0, 1.0, 1.0,
                         group_types=("periodic", "periodic", "wallModel", "wallModel", "periodic", "periodic"))
    for h, want in ((0.22, 0.25), (0.05, 0.05), (0.31, 0.35)):
        p = SolverParams(nu=1e-3, use_sgs=False, wall_model="loglaw", wall_model_height=h, cell_projection="LS")
        o = oracle.Solver(mesh, p)
        f = _outlet_faces(mesh, "y0")
        for b in f[:5]:
            c, y = o.wall_sample(int(b))
            fg = mesh.n_int_faces + b
            assert y == pytest.approx(want, abs=1e-12)
            xc = mesh.cell_centroid.reshape(-1, 3)[c]
            xf = mesh.face_centre.reshape(-1, 3)[fg]
            assert xc[1] == pytest.approx(want, abs=1e-12)
            assert abs(xc[0] - xf[0]) < 1e-12 and abs(xc[2] - xf[2]) < 1e-12


def _filter_case(**kw):
    case = cases.small_channel(8, 8, 6)
    for k, v in kw.items():
        setattr(case.params, k, v)
    return case


def _tangential(o, mesh, b):
    fg = mesh.n_int_faces + b
    n = mesh.face_normal.reshape(-1, 3)[fg]
    c, _ = o.wall_sample(int(b))
    u = o.U.reshape(-1, 3)[c]
    return u - (u @ n) * n


@pytest.mark.parametrize("mode", ["fixed", "adaptive"])
def test_wall_model_time_filter_recursion(mode):
    """Between steps the filter moves once: u_f += eps (u_par - u_f) with the stage-1 velocity,
    eps = dt / (T + dt) (fixed) or min(1, dt kappa u_tau(|u_f|) / (f y)) (adaptive); the first
    filtered call seeds u_f."""
    kw = {"wall_model_filter_time": 0.05} if mode == "fixed" else {"wall_model_filter_tf_fac": 2.0}
    case = _filter_case(**kw)
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    P.initialize(o, case, mesh)
    dt = case.params.dt
    walls = np.concatenate([_outlet_faces(mesh, "bottom_wall"), _outlet_faces(mesh, "top_wall")])
    ready, _ = o.wall_filter_state()
    assert not ready
    o.advance()
    ready, uf = o.wall_filter_state()
    assert ready
    for step in range(3):
        expect = {}
        for b in walls[::7]:
            up = _tangential(o, mesh, b)
            prev = uf[b].copy()
            if mode == "fixed":
                eps = dt / (0.05 + dt)
            else:
                _, y = o.wall_sample(int(b))
                ut = math.sqrt(oracle.wall_tau(np.linalg.norm(prev), y, case.params.nu))
                eps = min(1.0, dt * 0.41 * ut / (2.0 * y)) if ut > 1e-30 else 1.0
            expect[b] = prev + eps * (up - prev)
        o.advance()
        _, uf = o.wall_filter_state()
        for b, v in expect.items():
            assert np.abs(uf[b] - v).max() < 1e-14 * max(1.0, np.abs(v).max())


def test_wall_model_filter_sets_the_traction():
    """tau_w follows the filtered velocity: with a very long time constant it stays at the value
    of the seeding call, while the unfiltered traction moves."""
    case = _filter_case(wall_model_filter_time=1e12)
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    u = oracle.Solver(mesh, _filter_case().params)
    for s in (o, u):
        P.initialize(s, case, mesh)
        s.advance()
    t1 = o.array("tau_w_bnd").copy()
    u1 = u.array("tau_w_bnd").copy()
    for s in (o, u):
        s.advance()
        s.advance()
    assert np.abs(o.array("tau_w_bnd") - t1).max() < 1e-12 * np.abs(t1).max()
    assert np.abs(u.array("tau_w_bnd") - u1).max() > 1e-6 * np.abs(u1).max()


def test_per_patch_roughness():
    case = cases.most_channel(8, 8, 6, dT=1.0, z0_m=1e-3)
    case.params.patch_z0_m = {"top_wall": 2e-2}
    case.params.patch_z0_h = {"top_wall": -1.0, "bottom_wall": 5e-5}
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    P.initialize(o, case, mesh)
    o.eval_closures(case.params.dt)
    tau = o.array("tau_w_bnd").reshape(-1, 3)
    U = o.U.reshape(-1, 3)
    T = o.array("T")
    L = o.array("L_bnd")
    for name, z0m, z0h in (("bottom_wall", 1e-3, 5e-5), ("top_wall", 2e-2, 2e-2 * math.exp(-2.0))):
        for b in _outlet_faces(mesh, name)[::5]:
            fg = mesh.n_int_faces + b
            own = mesh.face_owner[fg]
            n = mesh.face_normal.reshape(-1, 3)[fg]
            up = U[own] - (U[own] @ n) * n
            y = abs((mesh.face_centre.reshape(-1, 3)[fg] - mesh.cell_centroid.reshape(-1, 3)[own]) @ n) + 1e-30
            _, yw = o.wall_sample(int(b))
            us, _, _ = oracle.most_iterate(np.linalg.norm(up), T[own], case.params.T_surface, yw, z0m, z0h,
                                           case.params.L_obukhov)
            assert np.linalg.norm(tau[b]) == pytest.approx(us * us, rel=1e-12)
            assert abs(yw - y) < 1e-12
    assert np.isfinite(L).all()
