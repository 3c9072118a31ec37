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


# ---- boundary extensions: convective outlet with backflow traction, wall-model sampling height
# and time filter, per-patch MOST roughness (src/bc/mod_bc.f90:681-818,
# src/sgs/mod_wall_model.f90:126-313, src/abl/mod_abl_wall.f90:249-262)

def backflow_setup(n=10, amp=0.04, mode="convective", beta=None):
    """Skewed n^3 box, periodic in x and y, inlet (0, 0, 0.4) at z0, outlet at zL; the initial
    axial velocity 0.4 + 0.6 sin(2 pi x) re-enters through part of the outlet, plus a shear
    u = 0.2 sin(2 pi y) z that the convective outlet carries out. Returns (mesh, params, U0)."""
    from dopamine_amd import mesh as M
    from dopamine_amd.params import SolverParams
    mesh = M.skewed_box(n, amp, group_types=("periodic", "periodic", "periodic", "periodic", "inlet", "outlet"))
    p = SolverParams(nu=2e-3, advection_scheme="KEP", use_sgs=True, sgs_model="VREMAN", wall_model="none",
                     dt=4e-3, cfl_max=0.5, p_tol=1e-12, n_nonortho_corr=2, rhie_chow_dual=True,
                     cell_projection="LS", patch_U={"z0": (0.0, 0.0, 0.4)}, patch_p={"zL": 0.0},
                     patch_outlet_mode={"zL": mode})
    if beta is not None:
        p.patch_backflow_beta = {"zL": beta}
    x = mesh.cell_centroid.reshape(-1, 3)
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = 0.2 * np.sin(2 * np.pi * x[:, 1]) * x[:, 2]
    U[:, 2] = 0.4 + 0.6 * np.sin(2 * np.pi * x[:, 0])
    return mesh, p, U


def run_backflow(solver, params, U, nsteps):
    solver.set_velocity(U)
    solver.project_once(params.dt)
    for _ in range(nsteps):
        solver.advance()
    return solver


def wallmodel_case(height=0.0, t_filt=0.0, tf_fac=0.0):
    """The 16 x 12 x 8 channel with the log-law wall model sampled `height` above the walls and
    filtered in time."""
    from dopamine_amd import cases
    case = cases.small_channel(16, 12, 8)
    case.params.wall_model_height = height
    case.params.wall_model_filter_time = t_filt
    case.params.wall_model_filter_tf_fac = tf_fac
    return case


def roughness_case():
    """The MOST channel with its own roughness on each wall: z0_m 4e-3 on the top wall (z0_h from
    kB^-1), z0_h 5e-5 on the bottom one."""
    from dopamine_amd import cases
    case = cases.most_channel(16, 12, 8, dT=1.0, z0_m=1e-3)
    case.params.patch_z0_m = {"top_wall": 4e-3}
    case.params.patch_z0_h = {"top_wall": -1.0, "bottom_wall": 5e-5}
    return case
