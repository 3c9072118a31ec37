"""The reference's published scalar-advection study (docs/assets/convergence/advection.json, under
tests/golden/convergence/) run on the GPU: the final profiles at N = 100 and the errors at
N = 25 and 50 for CDS, QUICK, KOREN_TVD and VANLEER_TVD agree with the reference's numbers to
rounding, as the oracle's do (tests/test_advection_study.py)."""
import numpy as np
import pytest

from dopamine_amd.solver import FVSolver
from test_advection_study import REF, SCHEMES, bar_profile, exact_gauss

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", SCHEMES)
def test_gpu_profiles_match_the_reference_study(scheme):
    print(f"[gpu] advection study profile N=100 {scheme}", flush=True)
    T = bar_profile(FVSolver, 100, scheme, "scalar_pulse")
    d = np.abs(T - np.array(REF["profiles"]["schemes"][scheme])).max()
    print(f"[gpu] {scheme}: max |T - reference| {d:.2e}", flush=True)
    assert d < 1e-9


@pytest.mark.parametrize("scheme", SCHEMES)
def test_gpu_errors_match_the_reference_convergence_study(scheme):
    for n in (25, 50):
        print(f"[gpu] advection study N={n} {scheme}", flush=True)
        err = bar_profile(FVSolver, n, scheme, "scalar_gauss") - exact_gauss(n)
        row = [r for r in REF["convergence"][scheme] if r["N"] == n][0]
        L2 = float(np.sqrt(np.mean(err ** 2)))
        print(f"[gpu] {scheme} N={n}: L2 {L2:.12e} reference {row['L2']:.12e}", flush=True)
        assert abs(L2 - row["L2"]) <= 1e-9 * row["L2"]
        assert abs(float(np.abs(err).max()) - row["Linf"]) <= 1e-9 * row["Linf"]
