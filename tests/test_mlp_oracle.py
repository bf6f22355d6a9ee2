"""Pins the MLP oracle (oracle/mlp_oracle.py): GELU against erf closed forms, BF16
rounding against torch's independent conversion, and the analytic gradient against
central finite differences of the FP64 loss."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def test_gelu_closed_forms():
    x = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    want = np.array([v * 0.5 * (1 + math.erf(v / math.sqrt(2))) for v in x])
    assert np.allclose(mo.gelu(x), want, rtol=0, atol=1e-15)
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.allclose(mo.gelu_grad(x), fd, atol=1e-8)
    assert mo.gelu_grad(np.array([0.0]))[0] == 0.5


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000),
                        [1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0]]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), want)
    assert mo.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0              # tie -> even
    assert mo.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6  # tie -> even (up)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    dims = (5, 7, 6, 3)
    P = mo.param_count(dims)
    p = rng.standard_normal(P) * 0.5
    x = rng.standard_normal((9, 5))
    t = rng.standard_normal((9, 3))
    loss, g = mo.loss_grad(p, dims, x, t)
    h = 1e-6
    fd = np.empty(P)
    for i in range(P):
        e = np.zeros(P)
        e[i] = h
        fd[i] = (mo.loss_grad(p + e, dims, x, t)[0] - mo.loss_grad(p - e, dims, x, t)[0]) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())
    # a dropped term (bias gradient, GELU derivative) fails this check
    assert np.abs(g).min() >= 0 and np.abs(fd).max() > 1e-3


def test_bf16_emulation_stays_near_fp64():
    rng = np.random.default_rng(5)
    dims = (16, 64, 64, 4)
    p = rng.standard_normal(mo.param_count(dims)) * 0.2
    x = rng.standard_normal((32, 16))
    t = rng.standard_normal((32, 4))
    l64, g64 = mo.loss_grad(p, dims, x, t)
    lbf, gbf = mo.loss_grad(p, dims, x, t, bf=True)
    assert abs(lbf - l64) / l64 < 2e-2
    assert np.linalg.norm(gbf - g64) / np.linalg.norm(g64) < 3e-2
    assert not np.array_equal(gbf, g64)
