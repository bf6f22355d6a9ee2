"""CPU ORACLE of the MLP surrogate (test infrastructure only: imported by tests/ as the
checker of libaero_lbm.so's aero_mlp_* kernels, never by the product path).

FP64 numpy forward and backward of the network the library trains: widths
d0 -> ... -> dL, GELU(x) = x Phi(x) (erf form) after every hidden layer, mean-squared
error over batch x dL.  The reference has no learned part, so there is no reference
line to follow; this restates the mathematics:

    z_l = a_l W_l^T + b_l,  a_{l+1} = gelu(z_l),  y = z_{L-1}
    L = mean((y - t)^2)
    dz_{L-1} = 2 (y - t) / (B dL)
    dW_l = dz_l^T a_l,  db_l = sum_m dz_l,  dz_{l-1} = (dz_l W_l) * gelu'(z_{l-1})

With ``bf16=True`` it rounds to BF16 (nearest even) at the points where the GPU
kernels store BF16 - layer inputs, weights used as MFMA operands, back-propagated
errors - and keeps everything else in FP64, so the comparison with the GPU isolates
FP32 accumulation order.  Pinned by central finite differences of the FP64 loss
(tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

SQRT1_2 = 1.0 / math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x: np.ndarray) -> np.ndarray:
    return 0.5 * x * (1.0 + _erf(x * SQRT1_2))


def gelu_grad(x: np.ndarray) -> np.ndarray:
    return 0.5 * (1.0 + _erf(x * SQRT1_2)) + x * INV_SQRT_2PI * np.exp(-0.5 * x * x)


def unpack(params: np.ndarray, dims: Sequence[int]) -> List[Tuple[np.ndarray, np.ndarray]]:
    """One member's flat parameters -> [(W_l (d_{l+1} x d_l), b_l)]."""
    out, o = [], 0
    for l in range(len(dims) - 1):
        n, k = dims[l + 1], dims[l]
        W = params[o:o + n * k].reshape(n, k)
        o += n * k
        b = params[o:o + n]
        o += n
        out.append((np.asarray(W, np.float64), np.asarray(b, np.float64)))
    assert o == params.size
    return out


def param_count(dims: Sequence[int]) -> int:
    return sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(len(dims) - 1))


def forward(params: np.ndarray, dims: Sequence[int], x: np.ndarray, bf: bool = False):
    """Returns (y, acts, pres) for one member."""
    layers = unpack(params, dims)
    q = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    a = q(x)
    acts, pres = [], []
    for l, (W, b) in enumerate(layers):
        acts.append(a)
        z = a @ q(W).T + b
        pres.append(z)
        if l + 1 < len(layers):
            a = q(gelu(z))
    return pres[-1], acts, pres


def loss_grad(params: np.ndarray, dims: Sequence[int], x: np.ndarray, t: np.ndarray, bf: bool = False):
    """Returns (loss, flat gradient) for one member."""
    layers = unpack(params, dims)
    q = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    y, acts, pres = forward(params, dims, x, bf)
    B, dl = y.shape
    d = y - t
    loss = float(np.mean(d * d))
    dz = q(2.0 * d / (B * dl))
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dW = dz.T @ acts[l]
        db = dz.sum(0)
        grads[l] = (dW, db)
        if l > 0:
            dz = q((dz @ q(W)) * gelu_grad(pres[l - 1]))
    flat = np.concatenate([np.concatenate([g[0].reshape(-1), g[1]]) for g in grads])
    return loss, flat
