"""CPU oracle of the warm-start network -- TEST INFRASTRUCTURE.

Float64 numpy forward and backward of the MLP the product trains on the BF16
matrix cores (snoo.py_amd/mlp.py), with the same rounding points: the inputs,
weights and stored activations are rounded to bfloat16 exactly where the
device kernels round them, so the remaining differences are the FP32
accumulation order.  The backward itself is pinned by central finite
differences of the unrounded float64 network (tests/test_mlp_oracle.py).
"""
import numpy as np
from scipy.special import erf


def bf16(a):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(Ws, bs, X, rounding=True):
    """Returns Y and the cache (A list incl. input, Z list)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    A = [r(X)]
    Z = []
    L = len(Ws)
    for l in range(L):
        z = A[-1] @ r(Ws[l]) + np.asarray(bs[l], dtype=np.float64)
        if l < L - 1:
            if rounding:
                z = z.astype(np.float32).astype(np.float64)
            Z.append(z)
            A.append(r(gelu(z)))
        else:
            Y = z
    return Y, (A, Z)


def backward(Ws, cache, dY, rounding=True):
    """dW, db for dL/dY = dY (rounded to bfloat16 where the kernels round it)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    A, Z = cache
    L = len(Ws)
    gW, gb = [None] * L, [None] * L
    dz_f = np.asarray(dY, dtype=np.float64)
    dz_b = r(dY)
    for l in reversed(range(L)):
        gW[l] = A[l].T @ dz_b
        gb[l] = dz_f.sum(axis=0)
        if l > 0:
            d = (dz_b @ r(Ws[l]).T) * gelu_grad(Z[l - 1])
            dz_f = d
            dz_b = r(d)
    return gW, gb


def mse(Y, T):
    d = Y - T
    return float(np.mean(d * d)), 2.0 * d / d.size
