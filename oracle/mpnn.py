"""ORACLE — test infrastructure only.

NumPy statement of the message-passing preconditioner network (forward, loss, backward) that the
HIP kernels in dopamine_amd/csrc/mpnn.h implement. It checks the device forward pass and the
device gradients; nothing in dopamine_amd imports it.

Network (3 message-passing layers, 32 channels, tanh), on the pressure operator's graph:
  features   x_c = [r_c / s, (r_c / s) (dbar / d_c), log(d_c / dbar)],  s = RMS(r), dbar = mean(d)
  edges      a_cj = coef_cj / sqrt(d_c d_j)  (symmetric)
  layer l    h_l,c = tanh(Wa_l h_(l-1),c + Wb_l sum_j a_cj h_(l-1),j + b_l),  h_0 = x
  readout    o_c = w_o . h_3,c
  output     z_c = (r_c + s o_c) / d_c       (o = 0 gives the Jacobi preconditioner)
Loss of one residual sample r: L = ||A z - r||^2 / ||r||^2 (A symmetric: the residual left after
one preconditioned correction).
"""
import numpy as np

NCH = 32
NIN = 3


def init_weights(seed=0, scale=0.1):
    rng = np.random.default_rng(seed)
    w = {}
    dims = [NIN, NCH, NCH, NCH]
    for l in range(3):
        w[f"Wa{l}"] = scale * rng.standard_normal((NCH, dims[l]))
        w[f"Wb{l}"] = scale * rng.standard_normal((NCH, dims[l]))
        w[f"b{l}"] = np.zeros(NCH)
    w["wo"] = scale * rng.standard_normal(NCH)
    return w


class Graph:
    """The operator in ELL form (slot-major, -1 = empty) with its diagonal."""

    def __init__(self, nbr, coef, diag):
        self.nbr = np.asarray(nbr)          # (w, n)
        self.coef = np.asarray(coef)        # (w, n)
        self.diag = np.asarray(diag)        # (n,)
        d = self.diag
        valid = self.nbr >= 0
        j = np.where(valid, self.nbr, 0)
        self.a = np.where(valid, self.coef / np.sqrt(d[None, :] * d[j]), 0.0)
        self.j = j
        self.valid = valid

    def aggregate(self, h):
        """sum_j a_cj h_j for h of shape (n, k)."""
        return np.einsum("sn,snk->nk", self.a, h[self.j])

    def apply(self, v):
        """A v with A_cc = d_c, A_cj = -coef_cj."""
        return self.diag * v - np.einsum("sn,sn->n", np.where(self.valid, self.coef, 0.0), v[self.j])


def forward(w, g, r):
    """Returns z and the cache for backward."""
    s = np.sqrt(np.mean(r * r))
    dbar = np.mean(g.diag)
    rh = r / s
    x = np.stack([rh, rh * dbar / g.diag, np.log(g.diag / dbar)], axis=1)
    hs, ms = [x], []
    h = x
    for l in range(3):
        m = g.aggregate(h)
        h = np.tanh(h @ w[f"Wa{l}"].T + m @ w[f"Wb{l}"].T + w[f"b{l}"])
        ms.append(m)
        hs.append(h)
    o = h @ w["wo"]
    z = (r + s * o) / g.diag
    return z, dict(s=s, hs=hs, ms=ms, o=o, r=r)


def loss(w, g, r):
    z, cache = forward(w, g, r)
    res = g.apply(z) - r
    return float(res @ res / (r @ r)), z, res, cache


def backward(w, g, r):
    """Loss and its gradient with respect to every weight."""
    L, z, res, cache = loss(w, g, r)
    gz = 2.0 * g.apply(res) / (r @ r)          # A symmetric
    go = gz * cache["s"] / g.diag
    hs, ms = cache["hs"], cache["ms"]
    grads = {"wo": hs[3].T @ go}
    gh = np.outer(go, w["wo"])
    for l in (2, 1, 0):
        ga = gh * (1.0 - hs[l + 1] ** 2)
        grads[f"Wa{l}"] = ga.T @ hs[l]
        grads[f"Wb{l}"] = ga.T @ ms[l]
        grads[f"b{l}"] = ga.sum(axis=0)
        if l > 0:
            gb = ga @ w[f"Wb{l}"]
            gh = ga @ w[f"Wa{l}"] + g.aggregate(gb)   # a symmetric: the transpose aggregation
    return L, grads
