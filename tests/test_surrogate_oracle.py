"""Pins of the CPU surrogate restatement (oracle/surrogate_ref.py) that checks the GPU kernels:
finite differences of the float64 network, the bf16 rounding against known values, shapes."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as sr  # noqa: E402


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert sr.bf16(1.0 + 2.0 ** -8) == 1.0
    assert sr.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie, rounds to even mantissa
    assert sr.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert sr.bf16(-3.0) == -3.0
    assert sr.bf16(np.pi) == 3.140625


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    shapes = sr.layer_shapes(3, 4, 6, 2, 5)
    n = sum(m * k + m for m, k in shapes)
    p = rng.standard_normal(n) * 0.5
    B = rng.standard_normal((4, 3))
    x = rng.random((7, 3))
    t = rng.standard_normal((7, 5))
    w = rng.random(7) + 0.5
    loss, g = sr.loss_and_grad(p, B, x, t, w, shapes, emulate=False)
    idx = rng.choice(n, 40, replace=False)
    for i in idx:
        e = np.zeros(n)
        e[i] = 1e-6
        fd = (sr.loss_and_grad(p + e, B, x, t, w, shapes, emulate=False)[0] -
              sr.loss_and_grad(p - e, B, x, t, w, shapes, emulate=False)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_emulated_network_is_close_to_float64():
    rng = np.random.default_rng(1)
    shapes = sr.layer_shapes(6, 32, 64, 3, 64)
    n = sum(m * k + m for m, k in shapes)
    p = np.concatenate([np.concatenate([rng.standard_normal(m * k) * np.sqrt(2 / k), np.zeros(m)]) for m, k in shapes])
    B = rng.standard_normal((32, 6))
    x = rng.random((16, 6))
    y64 = sr.forward(p, B, x, shapes, emulate=False)
    y16 = sr.forward(p, B, x, shapes, emulate=True)
    err = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < err < 5e-2  # bf16 operands: percent-level, never exact
    assert len(p) == n
