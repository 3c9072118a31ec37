"""The dynamic Smagorinsky model (sgs_model 'DYNAMIC_SMAG') on the GPU against the oracle: the
stretched channel and the jittered tetrahedra, relative L2 of u, p and nu_t within 1e-10 after
three steps."""
import pytest

import oracle
from dopamine_amd import cases
from dopamine_amd.solver import FVSolver
from helpers import rel_l2, run_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["channel", "tets"])
def test_gpu_dynamic_smagorinsky_matches_oracle(name):
    case = cases.small_channel(16, 12, 8) if name == "channel" else cases.cavity_tets(5)
    case.params.sgs_model = "DYNAMIC_SMAG"
    case.params.use_sgs = True
    case.params.p_tol = 1e-12
    mesh = case.make_mesh()
    print(f"[gpu] dynamic Smagorinsky {name}: {mesh.n_cells} cells", flush=True)
    g, o = FVSolver(mesh, case.params), oracle.Solver(mesh, case.params)
    for s in (g, o):
        run_case(s, case, mesh, nsteps=3, stats_every=3)
    assert o.array("nu_sgs").max() > 0.0
    for fld in ("U", "p", "nu_sgs"):
        a = o.array(fld)
        d = rel_l2(a, g.get(fld)[: a.size])
        print(f"[gpu]   {fld}: {d:.2e}", flush=True)
        assert d < 1e-10, fld
