"""Shared test helpers: running a case on any engine (oracle, host-check, GPU) and the
reference's stats comparison rule (tests/regression/compare_stats.py)."""
import csv
import os

import numpy as np

from dopamine_amd import params as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COMPARED = ("time", "E", "Umax", "Ub", "tauw_mean", "nu_sgs_mean")


def load_reference_csv(case_name, np_ranks=1):
    rows = {}
    with open(os.path.join(GOLDEN, "regression", f"{case_name}_np{np_ranks}.csv"), newline="") as fh:
        for r in csv.DictReader(fh):
            rows[int(r["step"])] = {k: float(v) for k, v in r.items() if k != "step"}
    return rows


def run_case(solver, case, mesh, nsteps=None, stats_every=None):
    """Initial condition + projections, then nsteps solver_advance calls; stats rows by step."""
    P.initialize(solver, case, mesh)
    rows = {}
    every = stats_every or case.stats_every
    for step in range(1, (nsteps or case.nsteps) + 1):
        solver.advance()
        if step % every == 0:
            st = solver.stats()
            st["time"] = solver.time
            rows[step] = st
    return rows


def compare_stats(ref, test, rtol=1e-6, atol=1e-12, div_max=1e-9):
    """compare_stats.py rule; returns (worst relative error per column, failures)."""
    worst, fails = {}, []
    for step, r in ref.items():
        if step not in test:
            continue
        for col in COMPARED:
            a, b = r[col], test[step][col]
            err = abs(a - b) / max(abs(a), atol / rtol)
            worst[col] = max(worst.get(col, 0.0), err)
            if err > rtol:
                fails.append(f"step {step} {col}: ref={a:.10e} test={b:.10e} rel={err:.2e}")
        if test[step]["max_abs_divU"] > div_max:
            fails.append(f"step {step} max_abs_divU={test[step]['max_abs_divU']:.3e}")
    return worst, fails


def rank_spread(case_name):
    """Per-column relative difference between the reference's one- and two-rank runs of a case,
    with compare_stats' normalisation: how much the column moves when only the pressure
    preconditioner's parallel path changes."""
    a, b = load_reference_csv(case_name, 1), load_reference_csv(case_name, 2)
    worst = {}
    for step, r in a.items():
        if step not in b:
            continue
        for col in COMPARED:
            worst[col] = max(worst.get(col, 0.0), abs(r[col] - b[step][col]) / max(abs(r[col]), 1e-12 / 1e-6))
    return worst


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300))
