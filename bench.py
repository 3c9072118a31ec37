#!/usr/bin/env python3
"""Throughput of the RK3W fractional step on one MI355X: cell-updates/s per RK3 step.

Workload: turbulent channel 2pi x 2 x pi, 128 x 96 x 64 = 786,432 hex cells, tanh wall-normal
stretching (factor 2), periodic in x and z, Re_tau 395 (nu = 1/395, f_x = 1), KEP advection,
Vreman SGS, log-law wall model (kappa 0.41, B 5.2), laminar parabolic profile with a seeded 10%
uniform perturbation, pressure to relative residual 1e-8, fixed dt = 1e-3.

A step is one solver_advance (three RK3W stages with their pressure solves) with every field
resident in HBM. By default the measured path is the mixed-precision pressure solve (FP32 CG
sweeps inside FP64 iterative refinement, same FP64 residual test) and the plain FP64 Jacobi-CG
path is timed the same way beside it (`plain_fp64`, `speedup_vs_plain_fp64`); `--precond
learned` measures flexible CG with the learned preconditioner instead. `value` = cells x timed
steps / wall time of the timed steps. The roofline entry
is for the dominant kernel (the CG iteration pass), from HIP events that its dispatches stamp on
the solver's stream over a further run of the same number of steps; the CPU baseline is the FP64
C restatement of the same algorithm run single-threaded on one step of the same workload.
`--workload tets` measures the unstructured-gather workload instead: the lid-driven cavity on
998,250 jittered, shuffled tets renumbered by reverse Cuthill-McKee (Koren TVD, AMD SGS,
non-orthogonal correction).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s per RK3 step at p-residual 1e-8: learned/mixed vs plain FP64"


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("channel", "tets"), default="channel",
                    help="channel: the turbulent channel (the line's default workload); tets: the lid-driven "
                         "cavity on 6 x 55^3 jittered, shuffled tets renumbered by reverse Cuthill-McKee")
    ap.add_argument("--mesh", type=int, nargs=3, default=(128, 96, 64), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed",
                    help="mixed: FP32 CG sweeps inside FP64 iterative refinement (the default measured path); "
                         "fp64: the plain method")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 run that the line reports beside a mixed/learned measurement")
    ap.add_argument("--cg-variant", type=int, choices=(0, 1), default=1,
                    help="pressure CG recurrence: 0 two passes per iteration (KSPCG's), 1 single reduction")
    ap.add_argument("--precond", choices=("jacobi", "learned"), default="jacobi",
                    help="learned: the warm-up steps' Jacobi solves train the message-passing network, the "
                         "timed steps run flexible CG with it")
    ap.add_argument("--learn-iters", type=int, default=300, help="Adam steps of the learned preconditioner")
    ap.add_argument("--cpu-steps", type=int, default=1, help="RK3 steps of the CPU baseline (0 = skip)")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without the per-launch event pairs (no roofline entry)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cg_pass_bytes(mesh, variant, mixed):
    """Algorithmic HBM bytes of the CG's first pass per iteration: one coefficient and two int32
    indices per operator face, and per cell:
    FP64 two-pass recurrence: diag, z, p_old read, p, q written (5 x 8 B);
    FP64 single-reduction recurrence (the whole iteration): diag, p, s, x, r, w read and p, s, x,
    r, w written (11 x 8 B);
    mixed precision (FP32 single-reduction sweep): 1/diag, p, s, r, w and the correction e read
    and p, s, r, w, e written, all FP32 (11 x 4 B); FP32 coefficients."""
    nfp = mesh.n_int_faces + int(np.count_nonzero(mesh.periodic_pair >= 0)) // 2
    if mixed:
        return 44 * mesh.n_cells + 12 * nfp, nfp
    per_cell = 88 if variant == 1 else 40
    return per_cell * mesh.n_cells + 16 * nfp, nfp


def measured_traffic(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the
    same kernel and workload (profiles/), or None."""
    import glob
    for path in glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "*.md")):
        for line in open(path):
            line = line.strip()
            if line.startswith("TRAFFIC "):
                d = json.loads(line[len("TRAFFIC "):])
                if d.get("kernel") == kernel and d.get("workload") == workload:
                    return d["bytes_per_launch"]
    return None


def cpu_baseline(mesh, case, U, p, steps):
    import oracle
    o = oracle.Solver(mesh, case.params)
    o.array("p")[:] = p
    o.set_velocity(U)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.advance()
    dt = time.perf_counter() - t0
    its = o.counters()["iterations"]
    return {"value": mesh.n_cells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} RK3 step(s) of the same {mesh.n_cells}-cell workload from the GPU state after the "
                      f"timed steps; FP64 C restatement, one thread; {its} CG iterations; {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the path's kernels at one GPU's share; "
                                   "--gpus must be 1"}))
        sys.exit(2)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from dopamine_amd import cases, params as P
    from dopamine_amd.solver import FVSolver

    nx, ny, nz = args.mesh
    tets = args.workload == "tets"
    case = cases.cavity_tets(55, rcm=True) if tets else cases.channel(nx, ny, nz)
    case.params.cg_variant = args.cg_variant
    learned = args.precond == "learned"
    if learned:
        case.params.precond = "LEARNED"
        case.params.learn_collect = max(1, 3 * args.warmup)
        case.params.learn_iters = args.learn_iters
    t0 = time.perf_counter()
    mesh = case.make_mesh()
    log(f"mesh {case.name}: {mesh.n_cells} cells, {mesh.n_faces} faces ({time.perf_counter() - t0:.1f} s)")
    width = 4 if mesh.max_cell_faces <= 4 else 6 if mesh.max_cell_faces <= 6 else 8 if mesh.max_cell_faces <= 8 else 0
    wl_key = "tets 6x55^3 rcm" if tets else f"channel {nx}x{ny}x{nz}"

    def timed_run(params, precision, tag):
        solver = FVSolver(mesh, params, precision=precision)
        P.initialize(solver, case, mesh)
        solver.sync()
        log(f"[{tag}] initial projections done")
        for i in range(args.warmup):
            solver.advance()
            log(f"[{tag}] warmup step {i + 1}/{args.warmup}: {solver.counters()['last']} CG iterations in the last solve")
        solver.sync()
        if params.precond.upper() == "LEARNED":
            log(f"[{tag}] learned preconditioner after warm-up: {solver.learned_info()}")
        it0 = solver.counters()["iterations"]
        solver.sync()
        t0 = time.perf_counter()
        for i in range(args.steps):
            solver.advance()
        solver.sync()
        elapsed = time.perf_counter() - t0
        its = solver.counters()["iterations"] - it0
        log(f"[{tag}] timed {args.steps} steps in {elapsed:.3f} s, {its} CG iterations")
        return solver, elapsed, its

    mixed_run = args.precision == "mixed" and not learned  # the learned FCG runs in FP64
    plain = None
    if (mixed_run or learned) and not args.no_plain:
        import dataclasses
        ps, pe, pi = timed_run(dataclasses.replace(case.params, precond="JACOBI"), 0, "plain FP64")
        plain = {"value": mesh.n_cells * args.steps / pe, "ms_per_step": 1e3 * pe / args.steps,
                 "cg_iterations_per_step": pi / args.steps}
        ps.close()
    solver, elapsed, its = timed_run(case.params, 1 if mixed_run else 0,
                                     "learned" if learned else ("mixed" if mixed_run else "plain FP64"))

    # Kernel durations: the same number of further steps with a dispatch-stamped event pair on
    # every launch of the timed kernel classes (the stamps add a few microseconds between
    # launches, so these steps are not part of the throughput figure).
    kernels = ("cg_dir", "cg_upd", "explicit", "ls_grad")
    ktime = {k: (0, 0.0) for k in kernels}
    ktimed_ms = None
    if not args.no_kernel_timing:
        for k in kernels:
            solver.kernel_timing(k, True)
        t1 = time.perf_counter()
        for i in range(args.steps):
            solver.advance()
        solver.sync()
        ktimed_ms = 1e3 * (time.perf_counter() - t1) / args.steps
        ktime = {k: solver.kernel_time(k) for k in kernels}
        for k in kernels:
            solver.kernel_timing(k, False)
        log(f"kernel-timing pass: {ktimed_ms:.1f} ms per step")

    nc = mesh.n_cells
    mixed = mixed_run
    bytes_dir, nfp = cg_pass_bytes(mesh, args.cg_variant, mixed)
    info = solver.learned_info() if learned else None
    timed_learned = learned and info is not None and info["phase"] != "off"
    if timed_learned:  # hidden layer: h read and written (32 bfloat16 per cell), FP32 edge weight + index per face side
        bytes_dir = 128 * nc + 16 * nfp
    n_dir, ms_dir = ktime["cg_dir"]
    avg_s = ms_dir / max(n_dir, 1) / 1e3
    achieved = bytes_dir / avg_s / 1e9 if avg_s > 0 else 0.0
    st = solver.stats()
    out = {
        "metric": METRIC,
        "value": nc * args.steps / elapsed,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else ("f64+bf16" if learned else "f64"),
        "data": ("synthetic: fluid at rest, moving lid, on a generated seeded jittered mesh" if tets else
                 "synthetic: seeded 10% uniform perturbation of the laminar profile on a generated mesh"),
        "config": {"workload": (f"lid-driven cavity Re 3200 on {nc} jittered shuffled tets renumbered by reverse "
                                f"Cuthill-McKee ({nfp} pressure faces), Koren TVD + AMD, Jasak non-orthogonal "
                                "correction, Jacobi CG to 1e-8, " if tets else
                                f"turbulent channel {nx}x{ny}x{nz} ({nc} hex cells, {nfp} pressure faces), "
                                "Re_tau 395, KEP + Vreman + log-law wall model, Jacobi CG to 1e-8, ")
                               + ("FP32 CG sweeps in FP64 iterative refinement" if mixed else "plain FP64")
                               + (", flexible CG with the learned message-passing preconditioner (Jacobi CG "
                                  "fallback)" if learned else ""),
                   "precision": "mixed" if mixed else "fp64", "cg_variant": args.cg_variant, "precond": args.precond,
                   "dt": case.params.dt},
        "roofline": {"bound": "hbm",
                     "kernel": ("learned preconditioner hidden layer k_tile<1,mp::L1<6>> (MFMA bf16)" if timed_learned else
                                f"CG FP32 sweep iteration k_pass<6,6,cgk::Fused32T<{width}>>" if mixed else
                                f"CG iteration pass k_pass<6,6,cgk::FusedT<{width}>>" if args.cg_variant == 1
                                else "CG direction/operator pass k_pass<4,1,cgk::Dir>"),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (measured_traffic(f"Fused32T<{width}>" if mixed else f"FusedT<{width}>", wl_key)
                                 if (not timed_learned and args.cg_variant == 1) else None),
                     "bytes_per_launch": bytes_dir, "avg_launch_us": avg_s * 1e6,
                     "launches": n_dir},
        "kernels_us": {k: (v[1] * 1e3 / v[0] if v[0] else None) for k, v in ktime.items()},
        "kernel_timing_pass_ms_per_step": ktimed_ms,
        "cg_iterations_per_step": its / args.steps,

        "state": {"E": st["E"], "Ub": st["Ub"], "max_abs_divU": st["max_abs_divU"]},
        "cpu_baseline": None,
    }
    if plain is not None:
        out["plain_fp64"] = plain
        out["speedup_vs_plain_fp64"] = out["value"] / plain["value"]
    if learned:
        info = solver.learned_info()
        loss = solver.learned_loss()
        out["learned"] = {**info, "loss_first": float(loss[:10].mean()) if loss.size else None,
                          "loss_last": float(loss[-10:].mean()) if loss.size else None}
    # plain copy kernel of known size: achievable HBM bandwidth, and the calibration of the traffic
    # counters (FETCH_SIZE / WRITE_SIZE) for the same process
    import ctypes
    from dopamine_amd._lib import lib as _native
    L = _native()
    L.dopamine_hbm_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ncopy = 8 * nc
    cms = ctypes.c_double(0.0)
    if L.dopamine_hbm_copy_probe(ncopy, 20, ctypes.byref(cms)) == 0 and cms.value > 0:
        out["hbm_copy"] = {"bytes_per_launch": 16 * ncopy, "avg_launch_us": cms.value * 1e3,
                           "GBs": 16 * ncopy / (cms.value * 1e-3) / 1e9}
    if args.cpu_steps > 0:
        U, p = solver.U, solver.p
        log("CPU baseline (oracle restatement, one thread) ...")
        out["cpu_baseline"] = cpu_baseline(mesh, case, U, p, args.cpu_steps)
        log(f"CPU baseline: {out['cpu_baseline']['sample']}")
    solver.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
