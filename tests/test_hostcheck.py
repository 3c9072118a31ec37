"""The product's kernel bodies and RK3W driver (dopamine_amd/csrc/solver_impl.h, kernels.h,
cg_impl.h) compiled for the CPU (tests/hostcheck) against the oracle: same seeded inputs, relative
L2 of u, p, nu_t and wall traction within 1e-10 after several steps. The GPU tests repeat this on
the device; here it runs without one."""
import numpy as np
import pytest

import oracle
from dopamine_amd import cases, mesh as M
from helpers import compare_stats, load_reference_csv, rel_l2, run_case
from hostcheck import HostSolver



def fresh(name):
    """A new copy of a regression case (tests change its parameters)."""
    return cases.regression_cases()[name]


def _pair(case, mesh, nsteps):
    h, o = HostSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (h, o):
        run_case(s, case, mesh, nsteps=nsteps, stats_every=nsteps)
    return h, o


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("name", ["tgv_skew", "tgv_skew_chk", "hydrowall", "hydro", "pdrop"])
def test_host_kernels_match_oracle_fields(name, variant):
    case = fresh(name)
    case.params.p_tol = 1e-12
    case.params.cg_variant = variant
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 4)
    assert rel_l2(o.U, h.U) < 1e-10
    assert rel_l2(o.p, h.p) < 1e-10
    assert h.nonortho_info() == o.nonortho_info()
    assert h.counters()["iterations"] == o.counters()["iterations"]


def test_host_kernels_match_reference_csv():
    case = fresh("tgv_skew")
    mesh = case.make_mesh()
    _, fails = compare_stats(load_reference_csv("tgv_skew"), run_case(HostSolver(mesh, case.params), case, mesh))
    assert not fails, fails[:5]


@pytest.mark.parametrize("gradient", ["LEAST_SQUARES", "GREEN_GAUSS"])
def test_host_channel_closures_match_oracle(gradient):
    case = cases.small_channel(16, 12, 8)
    case.params.p_tol = 1e-12
    case.params.gradient_scheme = gradient
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 3)
    for name in ("U", "p", "nu_sgs", "tau_w_bnd", "grad_U"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name


@pytest.mark.parametrize("wall_gradient", [True, False])
def test_host_channel_amd_matches_oracle(wall_gradient):
    """AMD with and without the log-law wall-cell gradient on the stretched wall-model channel."""
    case = cases.small_channel(16, 12, 8)
    case.params.p_tol = 1e-12
    case.params.sgs_model = "AMD"
    case.params.amd_wall_gradient = wall_gradient
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 3)
    assert o.array("nu_sgs").max() > 0
    for name in ("U", "p", "nu_sgs"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name


@pytest.mark.parametrize("sgs", ["VREMAN", "SMAGORINSKY", "AMD"])
def test_host_tgv_sgs_matches_oracle(sgs):
    case = cases.tgv32(12)
    case.params.p_tol = 1e-12
    case.params.sgs_model = sgs
    case.extra["noise"] = 1e-4
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 3)
    assert rel_l2(o.U, h.U) < 1e-10
    assert rel_l2(o.array("nu_sgs"), h.get("nu_sgs")) < 1e-10


def test_host_rotational_off_and_cds_match_oracle():
    case = fresh("pdrop")
    case.params.p_tol = 1e-12
    case.params.use_rotational_projection = False
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 3)
    assert rel_l2(o.U, h.U) < 1e-10 and rel_l2(o.p, h.p) < 1e-10


@pytest.mark.parametrize("name", ["tgv_skew", "hydrowall"])
def test_host_mixed_precision_matches_oracle_to_solver_tolerance(name):
    case = fresh(name)
    case.params.p_tol = 1e-8
    mesh = case.make_mesh()
    h, o = HostSolver(mesh, case.params, precision=1), oracle.Solver(mesh, case.params)
    for s in (h, o):
        run_case(s, case, mesh, nsteps=3, stats_every=3)
    assert rel_l2(o.U, h.U) < 1e-7 and rel_l2(o.p, h.p) < 1e-6
    assert h.counters()["sweeps"] > 0
    assert h.stats()["max_abs_divU"] < max(10 * o.stats()["max_abs_divU"], 1e-10)


@pytest.mark.parametrize("scheme", ["QUICK", "VANLEER_TVD", "KOREN_TVD"])
@pytest.mark.parametrize("name", ["tgv_skew", "pdrop"])
def test_host_reconstruction_schemes_match_oracle(scheme, name):
    case = fresh(name)
    case.params.p_tol = 1e-12
    case.params.advection_scheme = scheme
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 3)
    assert rel_l2(o.U, h.U) < 1e-10 and rel_l2(o.p, h.p) < 1e-10


def test_host_tet_cavity_matches_oracle():
    """Jittered, shuffled tetrahedra (refinement-interface skew > 0.15: up to 8 correctors),
    Koren TVD, AMD SGS, moving lid."""
    case = cases.cavity_tets(6)
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    assert mesh.face_skew.max() > 0.15
    h, o = _pair(case, mesh, 4)
    assert h.nonortho_info() == o.nonortho_info()
    for name in ("U", "p", "nu_sgs"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name


@pytest.mark.parametrize("dT", [2.0, -2.0])
def test_host_most_channel_matches_oracle(dT):
    """MOST rough walls (stable and unstable surface) with Boussinesq temperature and the wall heat
    flux in the temperature equation."""
    case = cases.most_channel(16, 12, 8, dT=dT)
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    h, o = _pair(case, mesh, 3)
    for name in ("U", "p", "T", "nu_sgs", "tau_w_bnd", "q_w_bnd", "L_bnd"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name
    qw = o.array("q_w_bnd")
    assert np.count_nonzero(qw) > 0 and np.all(np.sign(qw[qw != 0]) == -np.sign(dT))


def _duct_tets_exact():
    """Jittered shuffled tets with a velocity inlet (fixed outflow bconst), an outlet (pressure
    force and compact outlet flux term) and walls, under the exact cell projection."""
    from dopamine_amd import mesh as M
    from dopamine_amd.params import SolverParams
    mesh = M.tet_box(4, lx=2.0, jitter=0.15, shuffle_seed=2,
                     group_types=("inlet", "outlet", "noSlip", "noSlip", "slip", "slip"))
    prm = SolverParams(nu=1e-2, advection_scheme="CDS", use_sgs=False, dt=5e-3, cfl_max=0.5, p_tol=1e-12,
                       p_max_iter=1000, rhie_chow_dual=False, cell_projection="EXACT", pressure_min_cos_theta=0.0,
                       patch_U={"x0": (1.0, 0.0, 0.0)}, patch_p={"xL": 0.0})
    return cases.Case("duct_tets_exact", lambda: mesh, prm, ic="uniform", ic_uref=1.0), mesh


@pytest.mark.parametrize("name", ["tgv_skew_exact", "hydro_exact", "duct_tets_exact"])
def test_host_exact_projection_matches_oracle(name):
    """cell_projection = 'EXACT' (project_exact, ProjBoundary ... ProjCorrect) against the oracle:
    fields within 1e-10 and the same number of projections and flexible-PCG iterations."""
    if name == "duct_tets_exact":
        case, mesh = _duct_tets_exact()
    else:
        case = fresh(name)
        mesh = case.make_mesh()
    case.params.p_tol = 1e-12
    h, o = _pair(case, mesh, 4)
    assert rel_l2(o.U, h.U) < 1e-10
    assert rel_l2(o.p, h.p) < 1e-10
    assert h.proj_counters() == o.proj_counters()
    assert h.proj_counters()["iterations"] > h.proj_counters()["projections"]


def test_rcm_renumbered_meshes():
    """Renumbering changes only summation orders where every least-squares stencil is complete:
    the skewed hex box gives the same flow cell by cell. On tets the stencil is capped at 32 cells
    taken in the cell's vertex order (src/mesh/mod_mesh.f90:513-560), so the numbering selects
    the stencil and the flows differ; there the product still matches the oracle on the
    renumbered mesh."""
    c = fresh("tgv_skew")
    c.params.p_tol = 1e-12
    e = M.skewed_box_elements(16, 0.08)
    ma, mb = M.Mesh(e), M.Mesh(M.rcm_elements(e))
    res = []
    for m in (ma, mb):
        o = oracle.Solver(m, c.params)
        run_case(o, c, m, nsteps=3, stats_every=3)
        res.append(o.U.copy())
    ka, kb = np.lexsort(ma.cell_centroid.T), np.lexsort(mb.cell_centroid.T)
    assert rel_l2(res[0][ka], res[1][kb]) < 1e-12
    t = cases.cavity_tets(5, rcm=True)
    t.params.p_tol = 1e-12
    mt = t.make_mesh()
    h, o = _pair(t, mt, 3)
    assert rel_l2(o.U, h.U) < 1e-10
