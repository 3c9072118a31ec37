"""The product's convective outlet, backflow traction, wall-model sampling height and time filter,
per-patch MOST roughness and the filter's checkpoint record (dopamine_amd/csrc, compiled for the
CPU by tests/hostcheck) against the oracle: relative L2 within 1e-10 after several steps."""
import numpy as np
import pytest

import oracle
from helpers import backflow_setup, rel_l2, roughness_case, run_backflow, run_case, wallmodel_case
from hostcheck import HostSolver


@pytest.mark.parametrize("mode,beta", [("convective", None), ("extrapolate", 1.0), ("convective", 0.0)])
def test_host_outlet_matches_oracle(mode, beta):
    mesh, p, U = backflow_setup(8, mode=mode, beta=beta)
    h = run_backflow(HostSolver(mesh, p), p, U, 5)
    o = run_backflow(oracle.Solver(mesh, p), p, U, 5)
    assert rel_l2(o.U, h.U) < 1e-10 and rel_l2(o.p, h.p) < 1e-10
    for name in ("U_bnd", "outlet_pb", "p_bnd"):
        a = o.array(name)
        assert np.abs(a - h.get(name)[: a.size]).max() <= 1e-10 * max(np.abs(a).max(), 1e-30), name
    assert h.stats()["max_abs_divU"] < 1e-9


@pytest.mark.parametrize("height,t_filt,tf_fac", [(0.12, 0.0, 0.0), (0.0, 0.02, 0.0), (0.12, 0.0, 2.0)])
def test_host_wall_model_sampling_and_filter_match_oracle(height, t_filt, tf_fac):
    case = wallmodel_case(height, t_filt, tf_fac)
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    h, o = HostSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (h, o):
        run_case(s, case, mesh, nsteps=4, stats_every=4)
    for name in ("U", "p", "tau_w_bnd"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name
    if t_filt or tf_fac:
        ready, uf = o.wall_filter_state()
        assert ready
        assert rel_l2(uf.reshape(-1), h.get("wm_filt")[: uf.size]) < 1e-10


def test_host_per_patch_roughness_matches_oracle():
    case = roughness_case()
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    h, o = HostSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (h, o):
        run_case(s, case, mesh, nsteps=3, stats_every=3)
    for name in ("U", "p", "T", "tau_w_bnd", "q_w_bnd"):
        a = o.array(name)
        assert rel_l2(a, h.get(name)[: a.size]) < 1e-10, name


def test_host_filter_state_survives_a_restart(tmp_path):
    """Write after 3 steps, read into a fresh solver, 2 more steps: equal to 5 straight steps (the
    filter record, header bit 8 of mod_checkpoint_io.f90:54-60, carries the filtered velocity)."""
    case = wallmodel_case(0.12, 0.02, 0.0)
    mesh = case.make_mesh()
    a = HostSolver(mesh, case.params)
    run_case(a, case, mesh, nsteps=3, stats_every=3)
    a.write_checkpoint(tmp_path, 3)
    for _ in range(2):
        a.advance()
    b = HostSolver(mesh, case.params)
    assert b.read_checkpoint(tmp_path, 3) == 3
    for _ in range(2):
        b.advance()
    assert np.array_equal(a.get("wm_filt"), b.get("wm_filt"))
    assert rel_l2(a.U, b.U) < 1e-12
