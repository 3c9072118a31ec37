"""Pins the oracle against the reference's own regression references
(tests/regression/reference/<case>_np1.csv, copied under tests/golden/regression/) with the
reference's comparison rule (tests/regression/compare_stats.py: rtol 1e-6 on time, E, Umax, Ub,
tauw_mean, nu_sgs_mean; face-flux divergence below 1e-9). Case inputs:
tests/regression/cases/<case>/case.input; meshes: tests/regression/mk_box_mesh.py."""
import pytest

import oracle
from dopamine_amd import cases
from helpers import compare_stats, load_reference_csv, run_case



def fresh(name):
    """A new copy of a regression case (tests change its parameters)."""
    return cases.regression_cases()[name]


@pytest.mark.parametrize("name", ["tgv", "tgv_skew", "tgv_skew_chk", "hydrowall", "hydro", "pdrop"])
def test_oracle_matches_reference_csv(name):
    case = fresh(name)
    mesh = case.make_mesh()
    rows = run_case(oracle.Solver(mesh, case.params), case, mesh)
    ref = load_reference_csv(name)
    assert set(rows) <= set(ref)
    worst, fails = compare_stats(ref, rows)
    assert not fails, "\n".join(fails[:10])
    # the restatement is far tighter than the reference's own tolerance
    assert max(worst.values()) < 1e-8, worst


def test_planted_sign_error_is_caught():
    """A wrong body-force sign (a plausible mistake) must fail the hydrowall pin."""
    case = fresh("hydrowall")
    case.params.f_const = tuple(-f for f in case.params.f_const)
    mesh = case.make_mesh()
    rows = run_case(oracle.Solver(mesh, case.params), case, mesh)
    _, fails = compare_stats(load_reference_csv("hydrowall"), rows)
    assert fails
