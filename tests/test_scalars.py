"""Passive scalars (n_scalars > 0): transport by the scalar scheme, diffusion with each scalar's
diffusivity and turbulent Schmidt number, per-patch BC modes.
Reference: src/solver/mod_fv_solver.f90:1719-1743 (H_C), :1335-1347 (predictor), :1003-1077
(gradients), src/bc/mod_bc.f90:582-679 (apply_passive_scalar_bc), src/mod_scalar.f90:13-15, 63-73.

Oracle pins: a diffusion mode decays by the RK3 polynomial of the discrete 7-point eigenvalue
(exact to rounding); the scalar total is conserved on a periodic box under a swirling flow for
every scheme; a scalar with the temperature's diffusivity, Schmidt number, scheme and Dirichlet
walls is the temperature field bit for bit (the temperature path is pinned separately); the
Dirichlet, Neumann (with its oscillating flux) and zero-gradient boundary values follow the
reference formulas. The product (host-check build) against the oracle.
"""
import numpy as np
import pytest

import oracle
from dopamine_amd import cases, mesh as M
from dopamine_amd.params import SolverParams
from helpers import rel_l2, run_case
from hostcheck import HostSolver


def _rk3_amp(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


def _periodic(n, **kw):
    mesh = M.uniform_box(n, n, n, 1.0, 1.0, 1.0)
    prm = SolverParams(nu=1e-2, advection_scheme="KEP", use_sgs=False, cell_projection="LS", dt=1e-3, cfl_max=0.9,
                       p_tol=1e-12, rhie_chow_dual=False, **kw)
    return mesh, prm


def test_scalar_diffusion_mode_decays_at_the_discrete_rate():
    n, k, D = 16, 2, 0.03
    mesh, prm = _periodic(n, n_scalars=2, scalar_scheme="CDS", scalar_diffusivity=(D, 2 * D, 0, 0))
    o = oracle.Solver(mesh, prm)
    x = mesh.cell_centroid[:, 0]
    C0 = np.cos(2 * np.pi * k * x)
    o.set_velocity(np.zeros((mesh.n_cells, 3)))
    o.set_scalar(0, C0)
    o.set_scalar(1, C0)
    for _ in range(10):
        o.advance()
    h = 1.0 / n
    for s, d in ((0, D), (1, 2 * D)):
        lam = -d * (2 / h ** 2) * (1 - np.cos(2 * np.pi * k * h))
        assert np.abs(o.scalar(s) - C0 * _rk3_amp(lam * prm.dt) ** 10).max() < 1e-13


@pytest.mark.parametrize("scheme", ["CDS", "QUICK", "VANLEER_TVD", "KOREN_TVD"])
def test_scalar_total_is_conserved_on_a_periodic_box(scheme):
    mesh, prm = _periodic(8, n_scalars=1, scalar_scheme=scheme, scalar_diffusivity=(1e-3, 0, 0, 0))
    prm.dt = 5e-3
    o = oracle.Solver(mesh, prm)
    from dopamine_amd.params import taylor_green_ic
    o.set_velocity(taylor_green_ic(mesh))
    o.project_once(prm.dt)
    rng = np.random.default_rng(8)
    C0 = rng.random(mesh.n_cells)
    o.set_scalar(0, C0)
    vol = mesh.cell_volume
    for _ in range(5):
        o.advance()
    assert abs(np.sum(o.scalar(0) * vol) - np.sum(C0 * vol)) < 1e-13 * np.sum(C0 * vol)
    assert not np.allclose(o.scalar(0), C0)


@pytest.mark.parametrize("scheme", ["CDS", "VANLEER_TVD"])
def test_scalar_with_the_temperature_settings_is_the_temperature(scheme):
    """Rayleigh-Benard cell (hot bottom, cold top, periodic sides); the scalar gets the same
    diffusivity, Schmidt number, scheme and wall values: C equals T exactly."""
    case = cases.rayleigh_benard(6, 6, 6)
    case.make_mesh = lambda: M.rb_box(6, 6, 6, lx=1.0, ly=1.0, lz=1.0)
    p = case.params
    p.temperature_scheme = scheme
    p.n_scalars = 1
    p.scalar_scheme = scheme
    p.scalar_diffusivity = (p.alpha_T, 0, 0, 0)
    p.scalar_pr_t = (p.Pr_t, 1, 1, 1)
    p.patch_scalar = {"bottom_wall": [{"mode": "dirichlet", "value": p.patch_T["bottom_wall"]}],
                      "top_wall": [{"mode": "dirichlet", "value": p.patch_T["top_wall"]}]}
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, p)
    run_case(o, case, mesh, nsteps=1, stats_every=1)  # initial T and U
    o.set_scalar(0, o.array("T").copy())
    for _ in range(4):
        o.advance()
    assert np.array_equal(o.scalar(0), o.array("T"))
    assert np.ptp(o.scalar(0)) > 0.0


def test_scalar_boundary_values_follow_the_modes():
    mesh = M.uniform_box(4, 4, 4, 1.0, 1.0, 1.0,
                         group_types=("periodic", "periodic", "noSlip", "noSlip", "noSlip", "noSlip"))
    D, prt, q, om, ph = 0.2, 0.85, 3.0, 5.0, 0.3
    prm = SolverParams(nu=1e-2, advection_scheme="CDS", use_sgs=False, cell_projection="LS", dt=1e-3,
                       rhie_chow_dual=False, n_scalars=1, scalar_diffusivity=(D, 0, 0, 0), Pr_t=prt,
                       patch_scalar={"y0": [{"mode": "dirichlet", "value": 7.0}],
                                     "yL": [{"mode": "neumann", "flux": q}],
                                     "z0": [{"mode": "flux", "flux": q, "omega": om, "phase": ph}]})
    o = oracle.Solver(mesh, prm)
    C = np.arange(mesh.n_cells, dtype=np.float64)
    o.set_scalar(0, C)
    Cb = o.array("C0_bnd")
    nu_t = o.array("nu_sgs")
    for p_idx, name in enumerate(mesh.patch_names):
        for k in range(mesh.patch_nfaces[p_idx]):
            b = mesh.patch_start[p_idx] + k
            f = mesh.n_int_faces + b
            c = mesh.face_owner[f]
            dcf = mesh.face_dcf[f]
            af = D + nu_t[c] / prt
            if name == "y0":
                want = 7.0
            elif name == "yL":
                want = C[c] + q * dcf / af
            elif name == "z0":
                want = C[c] + q * np.cos(om * o.time + ph) * dcf / af
            elif name in ("x0", "xL"):
                continue
            else:
                want = C[c]
            assert Cb[b] == pytest.approx(want, rel=1e-15, abs=1e-15), (name, b)


def scalar_channel(scheme="VANLEER_TVD"):
    """Small channel with two passive scalars (van Leer TVD by default): Dirichlet walls for the
    first, Neumann walls (one flux oscillating) for the second, under KEP + Vreman + log-law walls."""
    c = cases.small_channel(12, 10, 8)
    p = c.params
    p.p_tol = 1e-12
    p.n_scalars = 2
    p.scalar_scheme = scheme
    p.scalar_diffusivity = (1e-3, 5e-4, 0, 0)
    p.scalar_pr_t = (0.7, 0.9, 1, 1)
    p.patch_scalar = {"bottom_wall": [{"mode": "dirichlet", "value": 1.0}, {"mode": "neumann", "flux": 0.2}],
                      "top_wall": [{"mode": "dirichlet", "value": 0.0},
                                   {"mode": "neumann", "flux": -0.1, "omega": 30.0, "phase": 0.5}]}
    return c


@pytest.mark.parametrize("scheme", ["VANLEER_TVD", "CDS", "QUICK", "KOREN_TVD"])
def test_host_scalars_match_oracle(scheme):
    """Two scalars sharing a scheme run the multi-field pass (face fluxes once per face, then the
    per-cell sums); every scheme against the oracle's per-field statement."""
    case = scalar_channel(scheme)
    mesh = case.make_mesh()
    h, o = HostSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    y = mesh.cell_centroid[:, 1]
    for s in (h, o):
        run_case(s, case, mesh, nsteps=1, stats_every=1)
        s.set_scalar(0, 1.0 - y / 2.0)
        s.set_scalar(1, np.sin(3.0 * y))
        for _ in range(3):
            s.advance()
    for k in range(2):
        assert rel_l2(o.scalar(k), h.scalar(k)) < 1e-10, k
    assert rel_l2(o.U, h.U) < 1e-10


def test_host_full_stack_matches_oracle():
    """The full-stack workload at 16 x 16 x 8 with three obstacles: MOST walls on the ground and
    the obstacles, AMD, temperature, two scalars, Coriolis."""
    case = cases.full_stack(16, 16, 8, n_obstacles=3, seed=5)
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    h, o = HostSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (h, o):
        run_case(s, case, mesh, nsteps=4, stats_every=4)
    assert rel_l2(o.U, h.U) < 1e-10 and rel_l2(o.p, h.p) < 1e-10
    assert max(rel_l2(o.scalar(k), h.scalar(k)) for k in range(2)) < 1e-10
