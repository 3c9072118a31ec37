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

The device form (forward_bf16 / backward_from) rounds what the matrix cores read to bfloat16: the
weights, each layer's inputs [h, m] (m summed first, then rounded) and the stored activations; the
bias, the accumulation, tanh and the readout are FP32 on the device (float64 here). Its gradients
are those of the rounded network with the rounded weights treated as the weights (the FP32 master
copy receives them unchanged). The parameters are one flat vector in the order of PARAM_ORDER.
"""
import numpy as np

NCH = 32
NIN = 3


PARAM_ORDER = ("Wa0", "Wb0", "b0", "Wa1", "Wb1", "b1", "Wa2", "Wb2", "b2", "wo")


def param_shapes():
    dims = [NIN, NCH, NCH]
    out = {}
    for l in range(3):
        out[f"Wa{l}"] = (NCH, dims[l])
        out[f"Wb{l}"] = (NCH, dims[l])
        out[f"b{l}"] = (NCH,)
    out["wo"] = (NCH,)
    return out


N_PARAMS = sum(int(np.prod(v)) for v in param_shapes().values())


def flatten(w):
    return np.concatenate([np.asarray(w[k], dtype=np.float64).reshape(-1) for k in PARAM_ORDER])


def unflatten(flat):
    flat = np.asarray(flat, dtype=np.float64)
    out, o = {}, 0
    for k in PARAM_ORDER:
        shp = param_shapes()[k]
        m = int(np.prod(shp))
        out[k] = flat[o:o + m].reshape(shp).copy()
        o += m
    return out


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


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


def features(g, r):
    """Input features and their aggregate, before rounding."""
    s = np.sqrt(np.mean(r * r))
    dbar = np.mean(g.diag)
    rh = r / s
    x = np.stack([rh, rh * dbar / g.diag, np.log(g.diag / dbar)], axis=1)
    return s, x, g.aggregate(x)


def forward_bf16(w, g, r, rnd=bf16):
    """Device-form forward pass: z and the stored activations h1, h2, h3 (bfloat16 values).
    rnd replaces the rounding (the identity gives the float64 network)."""
    bf = rnd
    s, x, mx = features(g, r)
    wb = {k: (bf(v) if not k.startswith("b") else v) for k, v in w.items()}
    hin, min_ = bf(x), bf(mx)
    hs = []
    for l in range(3):
        pre = hin @ wb[f"Wa{l}"].T + min_ @ wb[f"Wb{l}"].T + wb[f"b{l}"]
        h = bf(np.tanh(pre))
        hs.append(h)
        if l < 2:
            hin, min_ = h, bf(g.aggregate(h))
    o = hs[2] @ wb["wo"]
    z = (r + s * o) / g.diag
    return z, hs


def backward_from(w, g, r, hs, rnd=bf16):
    """Loss and weight gradients of the device-form network, given its stored activations."""
    bf = rnd
    s, x, mx = features(g, r)
    wb = {k: (bf(v) if not k.startswith("b") else v) for k, v in w.items()}
    ins = [(bf(x), bf(mx)), (hs[0], bf(g.aggregate(hs[0]))), (hs[1], bf(g.aggregate(hs[1])))]
    o = hs[2] @ wb["wo"]
    z = (r + s * o) / g.diag
    rr = r @ r
    y = g.apply(z) - r
    L = float(y @ y / rr)
    go = 2.0 * g.apply(y) / rr * s / g.diag
    grads = {"wo": hs[2].T @ go}
    gh = np.outer(go, wb["wo"])
    for l in (2, 1, 0):
        ga = gh * (1.0 - hs[l] ** 2)
        hin, min_ = ins[l]
        grads[f"Wa{l}"] = ga.T @ hin
        grads[f"Wb{l}"] = ga.T @ min_
        grads[f"b{l}"] = ga.sum(axis=0)
        if l > 0:
            gh = ga @ wb[f"Wa{l}"] + g.aggregate(ga @ wb[f"Wb{l}"])
    return L, grads


def adam_step(flat, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1) of the flat parameter vector in place."""
    m[:] = beta1 * m + (1.0 - beta1) * grad
    v[:] = beta2 * v + (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    flat[:] = flat - lr * mh / (np.sqrt(vh) + eps)
