/* Pressure-backend C ABI, symbol for symbol the one src/solver/pressure_backend_petsc.c exports,
 * implemented by libdopamine_hip.so with a device-resident Jacobi-preconditioned CG on gfx950.
 * src/solver/mod_pressure_solver.f90:23-74 binds these with bind(C) interfaces and links unchanged.
 *
 * Conventions kept from the reference: all index arrays are 1-based (Fortran); the operator is
 * A_ii = diagonal[i] (non-finite or |.| < 1e-20 -> 1), A_on = A_no = -face_coeff[f] for every face
 * with skip_face[f] == 0 and face_neighbor_global[f] > 0 (zero or non-finite coefficients dropped);
 * the constant null space is projected out iff no Dirichlet row exists; the solve context is cached
 * and rebuilt only when (n_cells_local, n_cells_global, n_faces, reference_cell) change.
 */
#ifndef FVM_PRESSURE_BACKEND_H
#define FVM_PRESSURE_BACKEND_H

#ifdef __cplusplus
extern "C" {
#endif

/* 1: a usable HIP device and the backend are present (pressure_backend_petsc.c:555-561). */
int fvm_petsc_backend_available(void);
/* 0 on success, negative on error (pressure_backend_petsc.c:567-616). */
int fvm_petsc_backend_initialize(void);
/* Releases the cached solve context; 0 on success (pressure_backend_petsc.c:618-646). */
int fvm_petsc_backend_finalize(void);
/* Preconditioner selector (pressure_backend_petsc.c:534-539). The reference's algebraic-multigrid
 * names "boomeramg" (its default) and "gamg" run the flexible CG preconditioned by this backend's
 * FP32 aggregation-multigrid V-cycle; "gmres" and "jacobi" the Jacobi CG; "mixed-multigrid" FP32
 * multigrid-preconditioned flexible-CG sweeps inside FP64 iterative refinement; "mixed" selects
 * FP32-operator Jacobi-CG sweeps inside FP64 iterative refinement (same stopping test on the FP64
 * residual); "fcg-learned" selects the learned preconditioner (dopamine_mpnn.h): the first solves
 * run Jacobi CG and keep residual snapshots, the network is then trained on them and later solves
 * run flexible CG with it, falling back to Jacobi CG when a solve fails or needs more iterations.
 * Any other name makes fvm_petsc_solve_pressure return -11. A change of selector rebuilds the
 * solve context. */
void fvm_petsc_set_pressure_solver(const char *solver_name);
/* PETSc-style option string (pressure_backend_petsc.c:543-551); "-ksp_max_it N" and
 * "-ksp_type pipecg|groppcg" (the one-pass single-reduction CG recurrence; default "cg", two
 * passes) are understood, and with "fcg-learned" "-learn_collect N" (Jacobi solves before
 * training, default 600), "-learn_capacity N" (128), "-learn_iters N" (300), "-learn_batch N" (4),
 * "-learn_lr X" (3e-3), "-learn_seed N" (1); other options are recorded and ignored. */
void fvm_petsc_insert_option(const char *opt_str);
/* Node-local rank used to pick the HIP device (local_rank % device count) at initialize
 * (pressure_backend_petsc.c:563-565). */
void fvm_petsc_set_device_local_rank(int local_rank);
/* 1: rhs/solution may be device pointers (pressure_backend_petsc.c:35-41). */
int fvm_petsc_device_io_supported(void);
/* When on != 0, rhs and solution of the next solves are device pointers (pressure_backend_petsc.c:22-24). */
void fvm_petsc_set_io_on_device(int on);
/* Solve A x = rhs (pressure_backend_petsc.c:648-1034). solution is the initial guess in and the
 * answer out; *residual_norm is the true ||rhs - A x||_2. A solve that stops at its iteration limit
 * is repeated from x = 0 with four times the limit (Jacobi CG; the reference retries with block
 * Jacobi, :878-916). Returns 0 when converged (or soft-accepted after the retry: residual
 * < 1e4*rtol or < 1e3*atol), -(1000+reason) when not (-997, KSP_DIVERGED_ITS), -11 for an unknown
 * selector, other negative values for set-up and copy errors. */
int fvm_petsc_solve_pressure(int n_cells_local, int n_cells_global, int n_faces, const int *local_to_global,
                             const int *face_owner, const int *face_neighbor_local, const int *face_neighbor_global,
                             const int *skip_face, const double *diagonal, const double *rhs,
                             const double *face_coeff, int reference_cell, int max_iterations, double rtol,
                             double atol, const int *dirichlet_rows, const double *dirichlet_values,
                             int n_dirichlet_rows, double *solution, int *iteration_count, double *residual_norm);
/* Whether MPI may receive device pointers: DOPAMINE_GPU_AWARE_MPI=0/1 override, else 0
 * (src/mpi_gpu_support.c:13-21). */
int dopamine_mpi_gpu_aware(void);
/* Profiler range markers (src/io/nvtx_shim.c:7-8), emitted as ROCTx ranges. */
void dopamine_nvtx_push(const char *name);
void dopamine_nvtx_pop(void);

#ifdef __cplusplus
}
#endif
#endif
