"""NumPy statement of the aggregation-multigrid preconditioner (test infrastructure only: the
product's HIP/host code never imports this module).

The reference has no multigrid of its own: it hands the pressure system to PETSc with hypre
BoomerAMG or GAMG (src/solver/pressure_backend_petsc.c:375-465). This module states the method
that replaces them, from its published description, so the product's implementation can be
checked against an independent one:

- greedy aggregation by strong connections (Vanek, Mandel & Brezina, Computing 56 (1996)
  179-196): j is strongly connected to i when coef_ij >= theta sqrt(d_i d_j); phase 1 roots an
  aggregate at a cell whose strong neighbourhood is entirely unaggregated, phase 2 attaches the
  remaining cells to the aggregate of their strongest aggregated strong neighbour, phase 3 groups
  what is left with its unaggregated neighbours; cells are visited in index order;
- when the first aggregation leaves fewer than min_ratio0 cells per aggregate (line aggregates of
  strongly anisotropic cells), it is redone with the small threshold theta0 and the cycle uses
  the over-correction alpha_aggr;
- piecewise-constant prolongation P, Galerkin coarse operators A_c = P^T A P;
- the coarsest operator inverted densely, each singular connected component shifted by
  (mean |diagonal| / component size) 1 1^T;
- a V(1,1) cycle: x = w D^-1 b, r = b - A x, restrict, recurse, x += alpha P x_c, one more damped
  Jacobi step y = x + w D^-1 (b - A x).

Everything here is float64 (the product runs the cycle in FP32).
"""
import numpy as np
import scipy.sparse as sp


def ell_to_csr(nbr, coef, diag):
    """A (diag on the diagonal, -coef off it) from ELL slot-major arrays [slot][row]."""
    w, n = nbr.shape
    rows = np.repeat(np.arange(n)[None, :], w, axis=0)
    m = nbr >= 0
    A = sp.coo_matrix((-coef[m], (rows[m], nbr[m])), shape=(n, n)).tocsr()
    return (A + sp.diags(diag)).tocsr()


def aggregate(nbr, coef, diag, theta):
    """Aggregate index of every cell (int array) and the aggregate count."""
    w, n = nbr.shape
    agg = -np.ones(n, dtype=np.int64)

    def strong(i):
        out = []
        for k in range(w):
            j = nbr[k, i]
            if j >= 0 and j != i and coef[k, i] >= theta * np.sqrt(diag[i] * diag[j]):
                out.append((j, coef[k, i]))
        return out

    S = [strong(i) for i in range(n)]
    na = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        if any(agg[j] >= 0 for j, _ in S[i]):
            continue
        agg[i] = na
        for j, _ in S[i]:
            agg[j] = na
        na += 1
    join = -np.ones(n, dtype=np.int64)
    for i in range(n):
        if agg[i] >= 0:
            continue
        best = -1.0
        for j, c in S[i]:
            if agg[j] >= 0 and c > best:
                best, join[i] = c, agg[j]
    sel = (agg < 0) & (join >= 0)
    agg[sel] = join[sel]
    for i in range(n):
        if agg[i] >= 0:
            continue
        agg[i] = na
        for k in range(w):
            j = nbr[k, i]
            if j >= 0 and agg[j] < 0:
                agg[j] = na
        na += 1
    return agg, na


def prolongation(agg, na):
    n = agg.size
    return sp.coo_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, na)).tocsr()


def csr_to_ell(A):
    """ELL arrays (nbr, coef, diag) of a CSR operator with -coef off the diagonal."""
    A = A.tocsr()
    n = A.shape[0]
    diag = A.diagonal().copy()
    off = (A - sp.diags(diag)).tocoo()
    order = np.lexsort((off.col, off.row))
    r, c, v = off.row[order], off.col[order], off.data[order]
    keep = v != 0.0
    r, c, v = r[keep], c[keep], v[keep]
    cnt = np.bincount(r, minlength=n)
    w = max(1, int(cnt.max()) if cnt.size else 1)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    slot = np.arange(r.size) - start[r]
    nbr = -np.ones((w, n), dtype=np.int64)
    coef = np.zeros((w, n))
    nbr[slot, r] = c
    coef[slot, r] = -v
    return nbr, coef, diag


def dense_inverse(A):
    """Inverse of the coarsest operator with every singular connected component shifted."""
    A = A.toarray() if sp.issparse(A) else np.array(A, dtype=float)
    n = A.shape[0]
    d = np.abs(np.diag(A))
    dmean = d.sum() / n if d.sum() > 0 else 1.0
    ncomp, lab = sp.csgraph.connected_components(sp.csr_matrix(A != 0), directed=False)
    rs = A.sum(axis=1)
    B = A.copy()
    for c in range(ncomp):
        idx = np.flatnonzero(lab == c)
        if np.all(np.abs(rs[idx]) <= 1e-10 * np.abs(np.diag(A)[idx])):
            B[np.ix_(idx, idx)] += dmean / idx.size
    return np.linalg.inv(B)


class Hierarchy:
    """The multigrid levels of an ELL operator (theta, omega, alpha, coarse_max as the product's
    AmgCfg)."""

    def __init__(self, nbr, coef, diag, theta=0.08, omega=0.8, alpha=1.8, coarse_max=1500, theta0=0.02,
                 min_ratio0=3.5, alpha_aggr=2.4, max_levels=16):
        self.omega, self.alpha = omega, alpha
        self.aggressive = False
        self.A, self.P, self.aggs = [], [], []
        nbr, coef, diag = np.asarray(nbr), np.asarray(coef, dtype=float), np.asarray(diag, dtype=float)
        A = ell_to_csr(nbr, coef, diag)
        lev = 0
        while True:
            self.A.append(A)
            n = A.shape[0]
            if (n <= coarse_max and lev > 0) or lev + 1 >= max_levels:
                self.inv = dense_inverse(A)
                break
            agg, na = aggregate(nbr, coef, diag, theta)
            if lev == 0 and theta0 >= 0 and n < min_ratio0 * na:
                # aggressive first coarsening (line aggregates of anisotropic cells)
                agg, na = aggregate(nbr, coef, diag, theta0)
                self.alpha, self.aggressive = alpha_aggr, True
            if na * 3 > n * 2 and n > 16:
                agg, na = aggregate(nbr, coef, diag, 0.0)
            P = prolongation(agg, na)
            self.P.append(P)
            self.aggs.append(agg)
            A = (P.T @ A @ P).tocsr()
            nbr, coef, diag = csr_to_ell(A)
            lev += 1

    @property
    def sizes(self):
        return [A.shape[0] for A in self.A]

    def _dinv(self, A):
        d = A.diagonal()
        dmax = np.abs(d).max() if d.size else 0.0
        out = np.zeros_like(d)
        ok = d > 1e-12 * dmax
        out[ok] = 1.0 / d[ok]
        return out

    def vcycle(self, b, lev=0):
        A = self.A[lev]
        if lev == len(self.A) - 1:
            return self.inv @ b
        di = self._dinv(A)
        x = self.omega * di * b
        r = b - A @ x
        xc = self.vcycle(self.P[lev].T @ r, lev + 1)
        x = x + self.alpha * (self.P[lev] @ xc)
        return x + self.omega * di * (b - A @ x)


def fcg(A, b, M, rtol=1e-8, maxit=2000, nullspace=False):
    """Flexible CG (beta = max(0, z.(r - r_old) / (z.r)_old), the flexible beta of
    src/solver/mod_cell_projection.f90:273-281) with preconditioner M; returns (x, iterations)."""
    b = np.array(b, dtype=float)
    if nullspace:
        b -= b.mean()
    x = np.zeros_like(b)
    r = b.copy()
    bn = np.linalg.norm(b)
    p = None
    rz_old = None
    r_old = r.copy()
    for k in range(maxit + 1):
        if np.linalg.norm(r) <= rtol * bn:
            return x, k
        z = M(r)
        if nullspace:
            z -= z.mean()
        zr = z @ r
        beta = 0.0 if k == 0 else max(0.0, (zr - z @ r_old) / rz_old)
        rz_old = zr
        p = z if k == 0 else z + beta * p
        q = A @ p
        a = zr / (p @ q)
        x += a * p
        r_old = r.copy()
        r -= a * q
    return x, maxit
