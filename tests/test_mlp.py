# SPDX-License-Identifier: Apache-2.0
"""The network oracle (oracle/mlp.py) pinned without a GPU: its exact-mode
gradient equals central finite differences of its loss in float64 for every
parameter block, so a dropped term, a transposed operand or a wrong GELU
derivative fails; and its BF16 rounding agrees with the bit-level rounding
the kernels use."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp  # noqa: E402


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mlp.init_params(rng).astype(np.float64)
    p[mlp.P_B1:mlp.P_W2] = rng.normal(0, 0.1, mlp.HID)
    p[mlp.P_B3:mlp.P_B3 + 3] = rng.normal(0, 0.1, 3)
    x = np.zeros((37, 16))
    x[:, :14] = rng.normal(size=(37, 14))
    t = rng.normal(size=(37, 4))
    scale = 1.0 / (3 * 37)
    _, g = mlp.loss_and_grad(x, t, p, scale, bf16_mode=False)
    # Every block, a sample of its entries.
    blocks = [(mlp.P_W1, 14 * 64), (mlp.P_B1, 64), (mlp.P_W2, 4096), (mlp.P_B2, 64),
              (mlp.P_W3, 1024), (mlp.P_B3, 3)]
    h = 1e-6
    for start, size in blocks:
        idx = rng.choice(size, size=min(size, 24), replace=False) + start
        if start == mlp.P_W3:
            idx = [mlp.P_W3 + f * 16 + c for f, c in zip(rng.integers(0, 64, 24), rng.integers(0, 3, 24))]
        for i in idx:
            pp = p.copy()
            pm = p.copy()
            pp[i] += h
            pm[i] -= h
            lp, _ = mlp.loss_and_grad(x, t, pp, scale, bf16_mode=False)
            lm, _ = mlp.loss_and_grad(x, t, pm, scale, bf16_mode=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (start, i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.0078125, 1.001953125, 1.005859375, -2.5, 3.0e-3],
                    dtype=np.float32)
    got = mlp.bf16(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0); 1 + 3*2^-9
    # rounds up to 1 + 2^-7.
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125
    assert got[3] == 1.0 and got[4] == 1.0078125 and got[5] == -2.5
    assert abs(got[6] - 3.0e-3) <= 2.0 ** -8 * 3.0e-3  # unit roundoff of an 8-bit significand
    assert np.array_equal(mlp.from_bits(mlp.bf16_bits(vals)), got)
