/* Learned pressure preconditioner on its own operator (libdopamine_hip.so).
 *
 * No counterpart in the reference: its pressure solve is PETSc KSPCG with BoomerAMG/GAMG
 * (src/solver/pressure_backend_petsc.c:375-465). This interface exposes the message-passing
 * network that the solver's precond = 1 mode (dopamine_hip.h) uses inside flexible CG, so that
 * its forward pass, gradients, optimizer and solves can be checked and driven directly.
 *
 * Network: 3 message-passing layers of 32 channels over the operator's graph, bfloat16 weights
 * and layer inputs on the matrix cores with FP32 accumulation; z = (r + s o) / diag, where o is
 * the network's per-cell readout and s = RMS(r). Parameters are one FP32 vector of
 * dopamine_mpnn_nparams() values in the order Wa0 (32x3), Wb0 (32x3), b0 (32), Wa1 (32x32),
 * Wb1, b1, Wa2, Wb2, b2, wo (32), row-major.
 *
 * Operator: n rows in ELL form, slot-major (entry [k*n + c] is slot k of row c), nbr -1 for an
 * empty slot; A_cc = diag[c], A_c,nbr = -coef (the reference's assembly convention,
 * pressure_backend_petsc.c:296-365). Host arrays, copied. Returns: 0 ok, negative on error
 * (dopamine_mpnn_last_error()).
 */
#ifndef DOPAMINE_MPNN_H
#define DOPAMINE_MPNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *dopamine_mpnn_last_error(void);
int dopamine_mpnn_nparams(void);
/* capacity: residual samples kept for training; seed: initial weights and sample draws.
 * The initial network has a zero readout, i.e. it is the Jacobi preconditioner. */
void *dopamine_mpnn_create(int n, int w, const int32_t *nbr, const double *coef, const double *diag, int capacity,
                           int seed);
void dopamine_mpnn_free(void *h);
/* set != 0: load params (FP32 master copy, repacked for the matrix cores); 0: read them. */
int dopamine_mpnn_params(void *h, float *params, int set);
/* z = M(r). acts (or NULL): the stored activations h1, h2, h3 as bfloat16 bit patterns,
 * [3][n][32]. */
int dopamine_mpnn_forward(void *h, const double *r, double *z, uint16_t *acts);
/* Training loss ||A z - r||^2 / ||r||^2 of one residual and its gradient (dopamine_mpnn_nparams()
 * values); the gradient is also left as the pending one for dopamine_mpnn_adam. */
int dopamine_mpnn_loss_grad(void *h, const double *r, double *loss, double *grad);
/* As dopamine_mpnn_loss_grad, also returning the activations the gradient was taken at. */
int dopamine_mpnn_loss_grad_acts(void *h, const double *r, double *loss, double *grad, uint16_t *acts);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the pending gradient divided by count. */
int dopamine_mpnn_adam(void *h, double lr, int count);
/* Keep r as a training sample; returns the number kept. */
int dopamine_mpnn_add_sample(void *h, const double *r);
int dopamine_mpnn_samples(void *h);
/* iters Adam steps on batch samples each; loss_hist (or NULL) receives each step's mean loss. */
int dopamine_mpnn_train(void *h, int iters, int batch, double lr, double *loss_hist);
/* Solve A x = b (x: initial guess in, solution out) to ||b' - A x|| <= max(rtol ||b'||, atol):
 * mode 0 Jacobi CG, 1 flexible CG with the network, 2 Jacobi CG keeping residual snapshots as
 * training samples. nullspace != 0 projects the constant null space. *resid is the true
 * residual norm. Returns 0 when converged. */
int dopamine_mpnn_solve(void *h, int mode, const double *b, double *x, double rtol, double atol, int maxit,
                        int nullspace, int *its, double *resid);

#ifdef __cplusplus
}
#endif
#endif
