"""The NumPy statement of the learned preconditioner (oracle/mpnn.py): its analytic gradient
against central finite differences, and the untrained-network limit (o = 0 is Jacobi)."""
import numpy as np
import pytest

from oracle import mpnn


def _graph(n=5, seed=1):
    """A small random SPD graph Laplacian plus diagonal shift, in ELL form."""
    rng = np.random.default_rng(seed)
    nn = n * n
    nbr = -np.ones((4, nn), dtype=np.int64)
    coef = np.zeros((4, nn))
    for c in range(nn):
        i, j = divmod(c, n)
        for s, (di, dj) in enumerate(((1, 0), (-1, 0), (0, 1), (0, -1))):
            ii, jj = i + di, j + dj
            if 0 <= ii < n and 0 <= jj < n:
                nbr[s, c] = ii * n + jj
    w_edge = {}
    for c in range(nn):
        for s in range(4):
            j = nbr[s, c]
            if j >= 0:
                key = (min(c, j), max(c, j))
                if key not in w_edge:
                    w_edge[key] = rng.uniform(0.5, 2.0)
                coef[s, c] = w_edge[key]
    diag = coef.sum(axis=0) + rng.uniform(0.01, 0.2, nn)
    return mpnn.Graph(nbr, coef, diag)


def test_gradient_matches_finite_differences():
    g = _graph()
    w = mpnn.init_weights(seed=3, scale=0.3)
    for k in w:
        if k.startswith("b"):
            w[k] = 0.1 * np.random.default_rng(5).standard_normal(w[k].shape)
    r = np.random.default_rng(7).standard_normal(g.diag.size)
    L, grads = mpnn.backward(w, g, r)
    rng = np.random.default_rng(11)
    for name, val in w.items():
        idx = tuple(rng.integers(0, d) for d in val.shape)
        eps = 1e-6
        wp = {k: v.copy() for k, v in w.items()}
        wm = {k: v.copy() for k, v in w.items()}
        wp[name][idx] += eps
        wm[name][idx] -= eps
        fd = (mpnn.loss(wp, g, r)[0] - mpnn.loss(wm, g, r)[0]) / (2 * eps)
        assert abs(fd - grads[name][idx]) <= 1e-6 * max(1.0, abs(fd)), name


def test_zero_readout_is_jacobi():
    g = _graph()
    w = mpnn.init_weights()
    w["wo"][:] = 0.0
    r = np.random.default_rng(2).standard_normal(g.diag.size)
    z, _ = mpnn.forward(w, g, r)
    assert np.allclose(z, r / g.diag, rtol=1e-15)


def test_output_scales_with_the_residual():
    g = _graph()
    w = mpnn.init_weights(seed=4, scale=0.5)
    r = np.random.default_rng(3).standard_normal(g.diag.size)
    z1, _ = mpnn.forward(w, g, r)
    z2, _ = mpnn.forward(w, g, 7.5 * r)
    assert np.allclose(z2, 7.5 * z1, rtol=1e-12)


def test_device_form_reduces_to_the_float64_network_without_rounding():
    g = _graph()
    w = mpnn.init_weights(seed=6, scale=0.4)
    r = np.random.default_rng(8).standard_normal(g.diag.size)
    ident = lambda a: np.asarray(a, dtype=np.float64)
    z, hs = mpnn.forward_bf16(w, g, r, rnd=ident)
    z0, _ = mpnn.forward(w, g, r)
    assert np.allclose(z, z0, rtol=1e-13, atol=1e-15)
    L, grads = mpnn.backward_from(w, g, r, hs, rnd=ident)
    L0, grads0 = mpnn.backward(w, g, r)
    assert abs(L - L0) <= 1e-13 * L0
    for k in grads0:
        assert np.allclose(grads[k], grads0[k], rtol=1e-11, atol=1e-14), k


def test_bf16_rounding_and_flat_layout():
    v = np.array([1.0, 1.00390625, 1.01171875, -3.0e-3, 65504.0])
    # ties go to the even mantissa; 8 significant bits: relative error at most 2^-8
    assert list(mpnn.bf16(v)[:3]) == [1.0, 1.0, 1.015625]
    assert 0 < abs(mpnn.bf16(v)[3] + 3.0e-3) <= 2.0 ** -8 * 3.0e-3
    w = mpnn.init_weights(seed=2)
    flat = mpnn.flatten(w)
    assert flat.size == mpnn.N_PARAMS == 4416
    back = mpnn.unflatten(flat)
    assert all(np.array_equal(back[k], w[k]) for k in w)


def test_adam_matches_the_published_update():
    flat = np.array([1.0, -2.0])
    m, v = np.zeros(2), np.zeros(2)
    g = np.array([0.5, -0.25])
    mpnn.adam_step(flat, g, m, v, 1, 1e-2)
    # at t = 1 the bias-corrected step is lr * g / (|g| + eps): a signed step of lr
    assert np.allclose(flat, [1.0 - 1e-2, -2.0 + 1e-2], rtol=0, atol=1e-9)
