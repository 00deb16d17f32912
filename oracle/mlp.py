# SPDX-License-Identifier: Apache-2.0
"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the learned warm start's
network (nepheline_amd/csrc/mlp_hip.hip), used by tests/ as the checker of
the GPU kernels and never by the product.

The network: 16 inputs (14 contact features, zero-padded) -> 64 -> 64 -> 16
outputs (3 used), GELU (the erf form), row-major weights with the input
index first, one FP32 parameter vector:
    W1 [16][64], b1 [64], W2 [64][64], b2 [64], W3 [64][16], b3 [16].

With bf16=True the oracle rounds where the kernels round (weights, the GELU
outputs, the loss gradients fed to the matrix cores) and accumulates in
float64, so it differs from the kernels only by their FP32 accumulation
order.  With bf16=False it is the exact network, whose gradient
tests/test_mlp.py pins against central finite differences.

Pinned: the exact gradient against finite differences (tests/test_mlp.py);
the rounding mode against the kernels on the GPU (tests/test_gpu_mlp.py).
"""
import numpy as np
from scipy.special import erf

IN, HID, OUT, NOUT = 16, 64, 16, 3
P_W1 = 0
P_B1 = P_W1 + IN * HID
P_W2 = P_B1 + HID
P_B2 = P_W2 + HID * HID
P_W3 = P_B2 + HID
P_B3 = P_W3 + HID * OUT
P_COUNT = P_B3 + OUT


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), back as float64."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """The bfloat16 bit patterns of float32 values (nearest even)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[P_W1:P_B1].reshape(IN, HID), p[P_B1:P_W2], p[P_W2:P_B2].reshape(HID, HID),
            p[P_B2:P_W3], p[P_W3:P_B3].reshape(HID, OUT), p[P_B3:P_COUNT])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def init_params(rng):
    """He-style initial weights (the zero rows and columns kept zero)."""
    p = np.zeros(P_COUNT, dtype=np.float32)
    w1 = rng.normal(0.0, np.sqrt(2.0 / 14.0), (IN, HID))
    w1[14:, :] = 0.0
    p[P_W1:P_B1] = w1.ravel()
    p[P_W2:P_B2] = rng.normal(0.0, np.sqrt(2.0 / HID), (HID, HID)).ravel()
    w3 = rng.normal(0.0, np.sqrt(1.0 / HID), (HID, OUT))
    w3[:, NOUT:] = 0.0
    p[P_W3:P_B3] = w3.ravel()
    return p


def forward(x, p, bf16_mode=True):
    """x: (n, 16) float inputs (already BF16 values in bf16 mode).  Returns
    y (n, 16) and the intermediates."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ r(w1) + b1
    h1 = r(gelu(z1))
    z2 = h1 @ r(w2) + b2
    h2 = r(gelu(z2))
    y = h2 @ r(w3) + b3
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(x, t, p, scale, bf16_mode=True):
    """The loss scale * sum over rows and outputs 0..2 of (y - t)^2 and its
    gradient in the parameter vector's layout (W1 rows 14, 15 and W3 columns
    3..15 masked as the kernels mask them)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    y, (x, z1, h1, z2, h2) = forward(x, p, bf16_mode)
    t = np.asarray(t, dtype=np.float64)
    e = np.zeros_like(y)
    e[:, :NOUT] = y[:, :NOUT] - t[:, :NOUT]
    loss = scale * float(np.sum(e * e))
    dy = 2.0 * e * scale
    dyb = r(dy)
    g_w3 = h2.T @ dyb
    g_b3 = dy.sum(axis=0)
    dz2 = (dyb @ r(w3).T) * gelu_grad(z2)
    dz2b = r(dz2)
    g_w2 = h1.T @ dz2b
    g_b2 = dz2.sum(axis=0)
    dz1 = (dz2b @ r(w2).T) * gelu_grad(z1)
    g_w1 = x.T @ r(dz1)
    g_b1 = dz1.sum(axis=0)
    g_w1[14:, :] = 0.0
    g_w3[:, NOUT:] = 0.0
    g_b3[NOUT:] = 0.0
    g = np.concatenate([g_w1.ravel(), g_b1, g_w2.ravel(), g_b2, g_w3.ravel(), g_b3])
    return loss, g
