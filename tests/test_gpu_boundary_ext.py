"""The convective outlet, backflow traction, wall-model sampling height and time filter, per-patch
MOST roughness and the filter's checkpoint record on the GPU against the oracle: relative L2 within
1e-10 after several steps (the plain-FP64 bar)."""
import numpy as np
import pytest

import oracle
from dopamine_amd.solver import FVSolver
from helpers import backflow_setup, rel_l2, roughness_case, run_backflow, run_case, wallmodel_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,beta", [("convective", None), ("extrapolate", 1.0), ("convective", 0.0)])
def test_gpu_outlet_matches_oracle(mode, beta):
    mesh, p, U = backflow_setup(10, mode=mode, beta=beta)
    print(f"[gpu] outlet {mode} beta {beta}: {mesh.n_cells} cells", flush=True)
    g = run_backflow(FVSolver(mesh, p), p, U, 6)
    o = run_backflow(oracle.Solver(mesh, p), p, U, 6)
    du, dp = rel_l2(o.U, g.U), rel_l2(o.p, g.p)
    print(f"[gpu] outlet {mode}: rel-L2 U {du:.2e} p {dp:.2e}", flush=True)
    assert du < 1e-10 and dp < 1e-10
    for name in ("U_bnd", "outlet_pb", "p_bnd"):
        a = o.array(name)
        assert np.abs(a - g.get(name)[: a.size]).max() <= 1e-10 * max(np.abs(a).max(), 1e-30), name
    assert g.stats()["max_abs_divU"] < 1e-9


@pytest.mark.parametrize("height,t_filt,tf_fac", [(0.12, 0.0, 0.0), (0.0, 0.02, 0.0), (0.12, 0.0, 2.0)])
def test_gpu_wall_model_sampling_and_filter_match_oracle(height, t_filt, tf_fac):
    case = wallmodel_case(height, t_filt, tf_fac)
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    print(f"[gpu] wall model h {height} T {t_filt} f {tf_fac}", flush=True)
    g, o = FVSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (g, o):
        run_case(s, case, mesh, nsteps=4, stats_every=4)
    for name in ("U", "p", "tau_w_bnd"):
        a = o.array(name)
        d = rel_l2(a, g.get(name)[: a.size])
        print(f"[gpu]   {name}: {d:.2e}", flush=True)
        assert d < 1e-10, name
    if t_filt or tf_fac:
        _, uf = o.wall_filter_state()
        assert rel_l2(uf.reshape(-1), g.get("wm_filt")[: uf.size]) < 1e-10


def test_gpu_per_patch_roughness_matches_oracle():
    case = roughness_case()
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    print("[gpu] per-patch MOST roughness", flush=True)
    g, o = FVSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (g, o):
        run_case(s, case, mesh, nsteps=3, stats_every=3)
    for name in ("U", "p", "T", "tau_w_bnd", "q_w_bnd"):
        a = o.array(name)
        assert rel_l2(a, g.get(name)[: a.size]) < 1e-10, name


def test_gpu_filter_state_survives_a_restart(tmp_path):
    case = wallmodel_case(0.12, 0.02, 0.0)
    mesh = case.make_mesh()
    print("[gpu] wall filter checkpoint restart", flush=True)
    a = FVSolver(mesh, case.params)
    run_case(a, case, mesh, nsteps=3, stats_every=3)
    a.write_checkpoint(tmp_path, 3)
    for _ in range(2):
        a.advance()
    b = FVSolver(mesh, case.params)
    assert b.read_checkpoint(tmp_path, 3) == 3
    for _ in range(2):
        b.advance()
    assert np.array_equal(a.get("wm_filt"), b.get("wm_filt"))
    assert rel_l2(a.U, b.U) < 1e-12
