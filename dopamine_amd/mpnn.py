"""The learned pressure preconditioner on its own operator (include/dopamine_mpnn.h).

A 3-layer, 32-channel message-passing network over the operator's graph, with bfloat16 weights on
the matrix cores, applied inside flexible CG; trained on residuals of Jacobi CG solves with the
library's own forward, backward and Adam kernels. The operator is given in ELL form (slot-major
[w][n] neighbour indices, -1 = empty, coefficients with A_cj = -coef, diagonal).
"""
import ctypes

import numpy as np

from ._lib import lib

_P = ctypes.POINTER
_d, _i32, _f32 = ctypes.c_double, ctypes.c_int32, ctypes.c_float
_u16 = ctypes.c_uint16

PARAM_ORDER = ("Wa0", "Wb0", "b0", "Wa1", "Wb1", "b1", "Wa2", "Wb2", "b2", "wo")
NCH = 32


def _bind(L):
    if getattr(L, "_mpnn_bound", False):
        return L
    L.dopamine_mpnn_last_error.restype = ctypes.c_char_p
    L.dopamine_mpnn_create.restype = ctypes.c_void_p
    L.dopamine_mpnn_create.argtypes = [ctypes.c_int, ctypes.c_int, _P(_i32), _P(_d), _P(_d), ctypes.c_int,
                                       ctypes.c_int]
    L.dopamine_mpnn_free.argtypes = [ctypes.c_void_p]
    L.dopamine_mpnn_params.argtypes = [ctypes.c_void_p, _P(_f32), ctypes.c_int]
    L.dopamine_mpnn_forward.argtypes = [ctypes.c_void_p, _P(_d), _P(_d), _P(_u16)]
    L.dopamine_mpnn_loss_grad.argtypes = [ctypes.c_void_p, _P(_d), _P(_d), _P(_d)]
    L.dopamine_mpnn_loss_grad_acts.argtypes = [ctypes.c_void_p, _P(_d), _P(_d), _P(_d), _P(_u16)]
    L.dopamine_mpnn_adam.argtypes = [ctypes.c_void_p, _d, ctypes.c_int]
    L.dopamine_mpnn_add_sample.argtypes = [ctypes.c_void_p, _P(_d)]
    L.dopamine_mpnn_samples.argtypes = [ctypes.c_void_p]
    L.dopamine_mpnn_train.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _d, _P(_d)]
    L.dopamine_mpnn_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, _P(_d), _P(_d), _d, _d, ctypes.c_int,
                                      ctypes.c_int, _P(ctypes.c_int), _P(_d)]
    L._mpnn_bound = True
    return L


def bf16_to_float(bits):
    """bfloat16 bit patterns (uint16) to float64."""
    b = np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def _dp(a):
    return a.ctypes.data_as(_P(_d))


class LearnedPreconditioner:
    """The network and its flexible-CG solves on one operator, on the current HIP device."""

    def __init__(self, nbr, coef, diag, capacity=64, seed=1):
        self.L = _bind(lib())
        nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        if nbr.ndim != 2 or coef.shape != nbr.shape or diag.shape != (nbr.shape[1],):
            raise ValueError("expected nbr and coef of shape (w, n) and diag of shape (n,)")
        self.w, self.n = nbr.shape
        self.nparams = self.L.dopamine_mpnn_nparams()
        self.h = self.L.dopamine_mpnn_create(self.n, self.w, nbr.ctypes.data_as(_P(_i32)), _dp(coef), _dp(diag),
                                             capacity, seed)
        if not self.h:
            raise RuntimeError("dopamine_mpnn_create failed: " + self._err())

    def _err(self):
        return self.L.dopamine_mpnn_last_error().decode()

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError(f"{what} failed ({rc}): {self._err()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.dopamine_mpnn_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        out = np.zeros(self.nparams, dtype=np.float32)
        self._check(self.L.dopamine_mpnn_params(self.h, out.ctypes.data_as(_P(_f32)), 0), "params")
        return out

    @params.setter
    def params(self, flat):
        a = np.ascontiguousarray(flat, dtype=np.float32).reshape(-1)
        if a.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        self._check(self.L.dopamine_mpnn_params(self.h, a.ctypes.data_as(_P(_f32)), 1), "params")

    def _acts(self, bits):
        return [bf16_to_float(bits[i]).reshape(self.n, NCH) for i in range(3)]

    def forward(self, r, activations=False):
        """z = M(r); with activations=True also the stored h1, h2, h3."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.n)
        bits = np.zeros((3, self.n * NCH), dtype=np.uint16) if activations else None
        self._check(self.L.dopamine_mpnn_forward(self.h, _dp(r), _dp(z),
                                                 bits.ctypes.data_as(_P(_u16)) if activations else None), "forward")
        return (z, self._acts(bits)) if activations else z

    def loss_grad(self, r, activations=False):
        r = np.ascontiguousarray(r, dtype=np.float64)
        loss, g = _d(0.0), np.zeros(self.nparams)
        if activations:
            bits = np.zeros((3, self.n * NCH), dtype=np.uint16)
            self._check(self.L.dopamine_mpnn_loss_grad_acts(self.h, _dp(r), ctypes.byref(loss), _dp(g),
                                                            bits.ctypes.data_as(_P(_u16))), "loss_grad")
            return loss.value, g, self._acts(bits)
        self._check(self.L.dopamine_mpnn_loss_grad(self.h, _dp(r), ctypes.byref(loss), _dp(g)), "loss_grad")
        return loss.value, g

    def adam(self, lr, count=1):
        self._check(self.L.dopamine_mpnn_adam(self.h, lr, count), "adam")

    def add_sample(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        return self._check(self.L.dopamine_mpnn_add_sample(self.h, _dp(r)), "add_sample")

    @property
    def samples(self):
        return self.L.dopamine_mpnn_samples(self.h)

    def train(self, iters, batch=4, lr=3e-3):
        hist = np.zeros(max(1, iters))
        self._check(self.L.dopamine_mpnn_train(self.h, iters, batch, lr, _dp(hist)), "train")
        return hist[:iters]

    def solve(self, b, x0=None, mode="learned", rtol=1e-8, atol=0.0, maxit=5000, nullspace=False):
        """Solve A x = b; mode "jacobi", "learned" or "collect" (Jacobi, keeping residuals).
        Returns (x, iterations, true residual norm, converged)."""
        m = {"jacobi": 0, "learned": 1, "collect": 2}[mode]
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        its, res = ctypes.c_int(0), _d(0.0)
        rc = self._check(self.L.dopamine_mpnn_solve(self.h, m, _dp(b), _dp(x), rtol, atol, maxit, int(nullspace),
                                                    ctypes.byref(its), ctypes.byref(res)), "solve")
        return x, its.value, res.value, rc == 0


def unflatten(flat):
    """Parameter vector -> dict of arrays (shapes of PARAM_ORDER)."""
    shapes = {"Wa0": (32, 3), "Wb0": (32, 3), "b0": (32,), "Wa1": (32, 32), "Wb1": (32, 32), "b1": (32,),
              "Wa2": (32, 32), "Wb2": (32, 32), "b2": (32,), "wo": (32,)}
    out, o = {}, 0
    flat = np.asarray(flat, dtype=np.float64)
    for k in PARAM_ORDER:
        m = int(np.prod(shapes[k]))
        out[k] = flat[o:o + m].reshape(shapes[k]).copy()
        o += m
    return out
