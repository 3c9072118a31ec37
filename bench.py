#!/usr/bin/env python3
"""Throughput of the RK3W fractional step on one MI355X: cell-updates/s per RK3 step.

Default workload (the full stack): a 1000 x 1000 x 200 m box meshed 250 x 210 x 160 with 40 seeded
cuboid obstacles (heights 20-120 m) whose cells are removed (7,943,987 hex cells), periodic sides,
slip top, MOST rough walls (z0 = 0.1 m) over a surface 2 K warmer than the air, AMD SGS, Boussinesq
temperature and two passive scalars by van Leer TVD, KEP momentum, Coriolis f = 1e-4, a constant
pressure gradient, uniform 5 m/s start with seeded temperature noise, dt = 0.2 s, pressure to
relative residual 1e-8.

A step is one solver_advance (three RK3W stages with their pressure solves) with every field
resident in HBM. The measured path is mixed precision with the aggregation-multigrid
preconditioner: FP32 flexible-CG sweeps preconditioned by the FP32 multigrid V-cycle inside FP64
iterative refinement, the same FP64 residual test. The plain FP64 path (the reference's plain
method: Jacobi-preconditioned CG, FP64 throughout) is timed the same way in the same invocation
(`plain_fp64`, `speedup_vs_plain_fp64`). `value` = cells x timed steps / wall time of the timed
steps. The roofline entry is for the dominant kernel of the measured path, from HIP events that its
dispatches stamp on the solver's stream over a further run of the same number of steps; the CPU
baseline is the FP64 C restatement of the same algorithm (Jacobi CG on OpenMP threads) on a
bounded sample of the workload.

`--workload channel` measures the turbulent channel 128 x 96 x 64 (Re_tau 395, KEP + Vreman +
log-law wall model); `--workload tets` the lid-driven cavity on 998,250 jittered, shuffled tets
renumbered by reverse Cuthill-McKee (Koren TVD, AMD SGS, non-orthogonal correction).
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s per RK3 step at p-residual 1e-8: learned/mixed vs plain FP64"
WORKLOAD_KEYS = {"full": "full stack 250x210x160", "channel": "channel 128x96x64", "tets": "tets 6x55^3 rcm",
                 "rb": "rb 256x256x32"}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("full", "channel", "tets", "rb"), default="full",
                    help="full: the full-stack obstacle box (the line's workload); channel: the turbulent channel; "
                         "tets: the lid-driven cavity on 6 x 55^3 jittered, shuffled, RCM-renumbered tets; "
                         "rb: Rayleigh-Benard convection 256 x 256 x 32 (the learned preconditioner's configuration)")
    ap.add_argument("--mesh", type=int, nargs=3, default=None, metavar=("NX", "NY", "NZ"),
                    help="mesh of the full-stack or channel workload (defaults 250 210 160 and 128 96 64)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed",
                    help="mixed: FP32 CG sweeps inside FP64 iterative refinement (the measured path); fp64: FP64 CG")
    ap.add_argument("--precond", choices=("multigrid", "jacobi", "learned"), default="multigrid",
                    help="multigrid: FP32 aggregation-multigrid V-cycle in flexible CG (the measured path); jacobi: "
                         "Jacobi CG; learned: flexible CG with the learned message-passing preconditioner (FP64 "
                         "only; the warm-up steps' Jacobi solves train it)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 Jacobi-CG run that the line reports beside the measured path")
    ap.add_argument("--cg-variant", type=int, choices=(0, 1), default=1,
                    help="Jacobi CG recurrence: 0 two passes per iteration (KSPCG's), 1 single reduction")
    ap.add_argument("--learn-iters", type=int, default=300, help="Adam steps of the learned preconditioner")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0: the host's cores, at most 16; -1: skip it)")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without the per-launch event pairs (no roofline entry)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def n_op_faces(mesh):
    """Faces of the pressure operator: interior faces plus one per periodic pair."""
    return mesh.n_int_faces + int(np.count_nonzero(mesh.periodic_pair >= 0)) // 2


def cg_pass_bytes(mesh, variant, mixed):
    """Algorithmic HBM bytes of the Jacobi CG's first pass per iteration: one coefficient and two
    int32 indices per operator face, and per cell:
    FP64 two-pass recurrence: diag, z, p_old read, p, q written (5 x 8 B);
    FP64 single-reduction recurrence (the whole iteration): diag, p, s, x, r, w read and p, s, x,
    r, w written (11 x 8 B);
    mixed precision (FP32 single-reduction sweep): 1/diag, p, s, r, w and the correction e read
    and p, s, r, w, e written, all FP32 (11 x 4 B); FP32 coefficients."""
    nfp = n_op_faces(mesh)
    if mixed:
        return 44 * mesh.n_cells + 12 * nfp, nfp
    per_cell = 88 if variant == 1 else 40
    return per_cell * mesh.n_cells + 16 * nfp, nfp


def multigrid_bytes(mesh, mixed):
    """Algorithmic HBM bytes per launch of the multigrid path's finest-level kernels (FP32 operator
    values and vectors when mixed, FP64 vectors otherwise), with nf = operator faces:
    cg_dir, the flexible CG's direction pass FcgDir32T: one FP32 coefficient and two int32 indices
      per face (12 nf); per cell the diagonal, z and the previous direction read, the direction
      and q = A p written (5 x 4 B); FP64 (FcgDirT): 16 nf + 5 x 8 B per cell;
    cg_upd, FcgUpd32: e, r read and written, p, q read (6 x 4 B per cell); FP64 (FcgUpd): x, r read
      and written, r_old written, p, q read (7 x 8 B);
    mg_down, the pre-smoothing pass Down0: the column-scaled binary16 coupling of each face side
      and two int32 indices (12 nf); per cell b, 1/diag read, x, r written (4 x 4 B; b FP64 in
      FP64 mode);
    mg_up, the post-smoothing pass Up0: the row-scaled binary16 coupling of each face side and two
      int32 indices (12 nf); per cell x, 1/diag, r, q read, z written (5 x 4 B; r, r_old, z FP64 in
      FP64 mode)."""
    nf, nc = n_op_faces(mesh), mesh.n_cells
    v = 4 if mixed else 8
    return {"cg_dir": 12 * nf + 20 * nc if mixed else 16 * nf + 40 * nc,
            "cg_upd": 24 * nc if mixed else 56 * nc,
            "mg_down": 12 * nf + (v + 12) * nc,
            "mg_up": 12 * nf + (8 + 3 * v) * nc}


def hbm_used_gb():
    """Device memory in use on the current GPU (hipMemGetInfo: total - free), GB."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None
    return (total.value - free.value) / 1e9


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


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count(), model


def cpu_baseline(case_fn, threads, sample_mesh, full_mesh_cells):
    """The oracle (FP64 C restatement of the reference algorithm, Jacobi CG to 1e-8 on OpenMP
    threads, the rest of the step sequential) on a bounded sample of the workload: the same case
    generated at a reduced mesh, the initial projections, one warm-up step and one timed step."""
    import oracle
    from dopamine_amd import params as P
    case = case_fn(*sample_mesh)
    mesh = case.make_mesh()
    oracle.set_threads(threads)
    try:
        o = oracle.Solver(mesh, case.params)
        P.initialize(o, case, mesh)
        o.advance()
        c0 = o.counters()["iterations"]
        t0 = time.perf_counter()
        o.advance()
        dt = time.perf_counter() - t0
        its = o.counters()["iterations"] - c0
    finally:
        oracle.set_threads(1)
    ncpu, model = host_info()
    return {"value": mesh.n_cells / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"the same workload generated at {sample_mesh[0]}x{sample_mesh[1]}x{sample_mesh[2]} "
                      f"({mesh.n_cells} cells, against {full_mesh_cells} on the GPU): initial projections and one "
                      f"warm-up step, then one timed RK3 step of the FP64 C restatement with its Jacobi CG on "
                      f"{threads} OpenMP threads ({its} CG iterations in the step; {dt:.1f} s)",
            "cg_iterations_per_step": its, "host_cpus": ncpu, "host_model": model}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the path's kernels at one GPU's share; "
                                   "--gpus must be 1"}))
        sys.exit(2)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from dopamine_amd import cases, params as P
    from dopamine_amd.solver import FVSolver

    wl = args.workload
    if wl == "full":
        dims = tuple(args.mesh) if args.mesh else (250, 210, 160)
        case_fn = cases.full_stack
        case = case_fn(*dims)
    elif wl == "channel":
        dims = tuple(args.mesh) if args.mesh else (128, 96, 64)
        case_fn = cases.channel
        case = case_fn(*dims)
    elif wl == "rb":
        dims = tuple(args.mesh) if args.mesh else (256, 256, 32)
        case_fn = cases.rayleigh_benard
        case = case_fn(*dims)
    else:
        dims = None
        case_fn = None
        case = cases.cavity_tets(55, rcm=True)
    case.params.cg_variant = args.cg_variant
    learned = args.precond == "learned"
    multigrid = args.precond == "multigrid"
    if learned:
        case.params.precond = "LEARNED"
        case.params.learn_collect = max(1, 3 * args.warmup)
        case.params.learn_iters = args.learn_iters
    elif multigrid:
        case.params.precond = "MULTIGRID"
    mixed = args.precision == "mixed" and not learned  # the learned flexible CG runs in FP64
    t0 = time.perf_counter()
    mesh = case.make_mesh()
    nc = mesh.n_cells
    log(f"mesh {case.name}: {nc} cells, {mesh.n_faces} faces ({time.perf_counter() - t0:.1f} s)")
    width = 4 if mesh.max_cell_faces <= 4 else 6 if mesh.max_cell_faces <= 6 else 8 if mesh.max_cell_faces <= 8 else 0
    wl_key = WORKLOAD_KEYS[wl] if (wl == "tets" or args.mesh is None) else f"{wl} {dims[0]}x{dims[1]}x{dims[2]}"

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
        c0 = solver.counters()
        solver.sync()
        t0 = time.perf_counter()
        for i in range(args.steps):
            solver.advance()
        solver.sync()
        elapsed = time.perf_counter() - t0
        c1 = solver.counters()
        its, solves = c1["iterations"] - c0["iterations"], c1["solves"] - c0["solves"]
        if not solver.state_is_finite():
            raise FloatingPointError(f"[{tag}] non-finite velocity or pressure after the timed steps")
        log(f"[{tag}] timed {args.steps} steps in {elapsed:.3f} s, {its} CG iterations in {solves} solves")
        return solver, elapsed, its, solves

    plain = None
    if (mixed or learned or multigrid) and not args.no_plain:
        ps, pe, pi, pn = timed_run(dataclasses.replace(case.params, precond="JACOBI"), 0, "plain FP64")
        st = ps.stats()
        plain = {"value": nc * args.steps / pe, "ms_per_step": 1e3 * pe / args.steps,
                 "cg_iterations_per_step": pi / args.steps, "cg_iterations_per_solve": pi / max(pn, 1),
                 "max_abs_divU": st["max_abs_divU"], "max_abs_cell_net_flux": ps.cell_net_flux_max()}
        ps.close()
    tag = ("mixed" if mixed else "FP64") + (" + multigrid" if multigrid else " + learned" if learned else " Jacobi")
    solver, elapsed, its, solves = timed_run(case.params, 1 if mixed else 0, tag)

    # Kernel durations: the same number of further steps with a dispatch-stamped event pair on
    # every launch of the timed kernel classes (the stamps add a few microseconds between
    # launches, so these steps are not part of the throughput figure).
    kernels = ("cg_dir", "cg_upd", "explicit", "ls_grad", "mg_down", "mg_up")
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

    # the dominant kernel of the pressure path and its algorithmic bytes per launch
    info = solver.learned_info() if learned else None
    timed_learned = learned and info is not None and info["phase"] != "off"
    if multigrid:
        model = multigrid_bytes(mesh, mixed)
        vt = "float" if mixed else "double"
        names = {"cg_dir": f"flexible-CG direction pass k_pass<5,1,cgk::FcgDir32T<{width}>>" if mixed else
                 f"flexible-CG direction pass k_pass<5,1,cgk::FcgDirT<{width}>>",
                 "cg_upd": "flexible-CG update pass k_pass<1,1,cgk::FcgUpd32>" if mixed else
                 "flexible-CG update pass k_pass<1,1,cgk::FcgUpd>",
                 "mg_down": f"multigrid pre-smoothing k_pass<1,1,amgk::Down0<{vt},{width}>>",
                 "mg_up": f"multigrid post-smoothing k_pass<1,5,amgk::Up0<{vt},{width}>>"}
        traffic_names = {"cg_dir": f"FcgDir32T<{width}>", "cg_upd": "FcgUpd32", "mg_down": f"Down0<float, {width}>",
                         "mg_up": f"Up0<float, {width}>"}
        dom = max(model, key=lambda k: ktime[k][1])
        bytes_dom, kname, tname = model[dom], names[dom], traffic_names[dom] if mixed else None
    else:
        bytes_dom, _ = cg_pass_bytes(mesh, args.cg_variant, mixed)
        if timed_learned:  # hidden layer: h read and written (32 bfloat16 per cell), FP32 weight + index per face side
            bytes_dom = 128 * nc + 16 * n_op_faces(mesh)
        dom = "cg_dir"
        kname = ("learned preconditioner hidden layer k_tile<1,mp::L1<6>> (MFMA bf16)" if timed_learned else
                 f"CG FP32 sweep iteration k_pass<6,6,cgk::Fused32T<{width}>>" if mixed else
                 f"CG iteration pass k_pass<6,6,cgk::FusedT<{width}>>" if args.cg_variant == 1
                 else "CG direction/operator pass k_pass<4,1,cgk::Dir>")
        tname = (f"Fused32T<{width}>" if mixed else f"FusedT<{width}>") if (not timed_learned and
                                                                             args.cg_variant == 1) else None
    n_dom, ms_dom = ktime[dom]
    avg_s = ms_dom / max(n_dom, 1) / 1e3
    achieved = bytes_dom / avg_s / 1e9 if avg_s > 0 else 0.0
    st = solver.stats()
    if wl == "full":
        workload_text = (f"full stack: 1000 x 1000 x 200 m box {dims[0]}x{dims[1]}x{dims[2]} with 40 seeded cuboid "
                         f"obstacles removed ({nc} hex cells), MOST walls, AMD SGS, temperature + 2 passive scalars "
                         "(van Leer TVD), Coriolis, pressure to 1e-8")
    elif wl == "channel":
        workload_text = (f"turbulent channel {dims[0]}x{dims[1]}x{dims[2]} ({nc} hex cells), Re_tau 395, KEP + "
                         "Vreman + log-law wall model, pressure to 1e-8")
    elif wl == "rb":
        workload_text = (f"Rayleigh-Benard convection 4 x 4 x 1 box {dims[0]}x{dims[1]}x{dims[2]} ({nc} wall-clustered "
                         "hex cells), Ra 1e7, Pr 0.7, Boussinesq, KEP + van Leer TVD temperature, pressure to 1e-8")
    else:
        workload_text = (f"lid-driven cavity Re 3200 on {nc} jittered shuffled tets renumbered by reverse "
                         "Cuthill-McKee, Koren TVD + AMD, Jasak non-orthogonal correction, pressure to 1e-8")
    if multigrid:
        path_text = ("FP32 flexible-CG sweeps preconditioned by the FP32 aggregation-multigrid V-cycle in FP64 "
                     "iterative refinement" if mixed else
                     "FP64 flexible CG preconditioned by the FP32 aggregation-multigrid V-cycle")
    elif learned:
        path_text = "flexible CG with the learned message-passing preconditioner (Jacobi CG fallback)"
    else:
        path_text = "FP32 Jacobi-CG sweeps in FP64 iterative refinement" if mixed else "plain FP64 Jacobi CG"
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
        "data": ("synthetic: seeded obstacles and temperature noise on a generated mesh" if wl == "full" else
                 "synthetic: fluid at rest, moving lid, on a generated seeded jittered mesh" if wl == "tets" else
                 "synthetic: fluid at rest with seeded 1e-3 temperature noise on a generated mesh" if wl == "rb" else
                 "synthetic: seeded 10% uniform perturbation of the laminar profile on a generated mesh"),
        "config": {"workload": workload_text + "; " + path_text, "precision": "mixed" if mixed else "fp64",
                   "precond": args.precond, "cg_variant": args.cg_variant, "dt": case.params.dt},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(tname, wl_key) if tname else None,
                     "bytes_per_launch": bytes_dom, "avg_launch_us": avg_s * 1e6, "launches": n_dom},
        "kernels_us": {k: (v[1] * 1e3 / v[0] if v[0] else None) for k, v in ktime.items()},
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in ktime.items()},
        "kernel_timing_pass_ms_per_step": ktimed_ms,
        "cg_iterations_per_step": its / args.steps,
        "cg_iterations_per_solve": its / max(solves, 1),
        "state": {"E": st["E"], "Ub": st["Ub"], "max_abs_divU": st["max_abs_divU"],
                  "max_abs_cell_net_flux": solver.cell_net_flux_max()},
        "cpu_baseline": None,
    }
    if multigrid:
        out["multigrid"] = solver.multigrid_info()
    out["hbm_used_gb"] = hbm_used_gb()
    if plain is not None:
        out["plain_fp64"] = plain
        out["speedup_vs_plain_fp64"] = out["value"] / plain["value"]
    if learned:
        loss = solver.learned_loss()
        out["learned"] = {**solver.learned_info(), "loss_first": float(loss[:10].mean()) if loss.size else None,
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
    solver.close()
    if args.cpu_threads >= 0 and case_fn is not None:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = (100, 84, 64) if wl == "full" else (64, 48, 32)
        log(f"CPU baseline (oracle restatement, {threads} threads) on {sample} ...")
        out["cpu_baseline"] = cpu_baseline(case_fn, threads, sample, nc)
        log(f"CPU baseline: {out['cpu_baseline']['sample']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
