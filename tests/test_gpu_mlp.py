# SPDX-License-Identifier: Apache-2.0
"""The learned warm start's kernels on the GPU against the CPU oracle
(oracle/mlp.py): the forward pass and the loss gradient, through the C ABI,
on seeded inputs; and the network learns (Adam on its own gradients).
Tolerances: the kernels and the oracle round the same values to BF16; they
differ by FP32 versus float64 accumulation and by the kernels' GELU
approximation (within 5e-7 of the erf form, tested below), so a GELU
output can land one BF16 step apart.  Forward outputs agree
within 1e-2 of the output scale, gradients within 2e-2 relative (L2) — a
wrong lane layout or a dropped term is off by order one."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp  # noqa: E402

from nepheline_amd import learned as L  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(rng, n):
    x = np.zeros((n, 16), dtype=np.float32)
    x[:, :14] = rng.normal(size=(n, 14))
    bits = mlp.bf16_bits(x)
    t = np.zeros((n, 4), dtype=np.float32)
    t[:, :3] = rng.normal(size=(n, 3))
    p = mlp.init_params(rng)
    p[mlp.P_B1:mlp.P_W2] = rng.normal(0, 0.1, 64)
    p[mlp.P_B2:mlp.P_W3] = rng.normal(0, 0.1, 64)
    p[mlp.P_B3:mlp.P_B3 + 3] = rng.normal(0, 0.1, 3)
    return bits, mlp.from_bits(bits), t, p


@pytest.mark.parametrize("n", [1, 17, 1000, 4099, 300001])
def test_forward_matches_oracle(gpu_api, n):
    print(f"mlp forward n={n}", flush=True)
    rng = np.random.default_rng(n)
    bits, xf, _, p = _case(rng, n)
    net = L.Network(gpu_api, p)
    y = net.forward(L.DeviceArray(gpu_api, bits))
    want, _ = mlp.forward(xf, p, bf16_mode=True)
    err = np.abs(y[:, :3] - want[:, :3]).max()
    assert err <= 1e-2 * (np.abs(want[:, :3]).max() + 1.0), err
    assert np.all(y[:, 3] == 0.0)


@pytest.mark.parametrize("n", [16, 333, 4096])
def test_gradient_matches_oracle(gpu_api, n):
    print(f"mlp gradient n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    bits, xf, t, p = _case(rng, n)
    net = L.Network(gpu_api, p)
    scale = 1.0 / (3.0 * n)
    net.gradient(L.DeviceArray(gpu_api, bits), L.DeviceArray(gpu_api, t), scale)
    g = net.grad.numpy().astype(np.float64)
    loss = float(net.loss.numpy()[0])
    want_loss, want = mlp.loss_and_grad(xf, t, p, scale, bf16_mode=True)
    assert abs(loss - want_loss) <= 1e-3 * want_loss
    for start, stop in ((mlp.P_W1, mlp.P_B1), (mlp.P_B1, mlp.P_W2), (mlp.P_W2, mlp.P_B2),
                        (mlp.P_B2, mlp.P_W3), (mlp.P_W3, mlp.P_B3), (mlp.P_B3, mlp.P_COUNT)):
        a, b = g[start:stop], want[start:stop]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        assert rel <= 2e-2, (start, rel)


def test_the_network_learns_a_teacher(gpu_api):
    """Adam on the kernels' gradients fits another network's outputs: the
    loss falls tenfold."""
    print("mlp training on a teacher network", flush=True)
    rng = np.random.default_rng(7)
    n = 1 << 15
    x = np.zeros((n, 16), dtype=np.float32)
    x[:, :14] = rng.normal(size=(n, 14))
    bits = mlp.bf16_bits(x)
    teacher = mlp.init_params(rng)
    t = np.zeros((n, 4), dtype=np.float32)
    t[:, :3] = mlp.forward(mlp.from_bits(bits), teacher, bf16_mode=False)[0][:, :3]
    net = L.Network(gpu_api, mlp.init_params(np.random.default_rng(8)))

    def mean_loss():
        y = net.forward(L.DeviceArray(gpu_api, bits))
        return float(np.mean((y[:, :3] - t[:, :3]) ** 2))

    before = mean_loss()
    order = np.random.default_rng(0)
    batch = 2048
    hist = []
    for _ in range(6):
        perm = order.permutation(n)
        for s in range(0, n, batch):
            idx = perm[s:s + batch]
            net.gradient(L.DeviceArray(gpu_api, bits[idx]), L.DeviceArray(gpu_api, t[idx]),
                         1.0 / (3.0 * batch))
            net.adam(3e-3)
        hist.append(mean_loss())
    assert hist[-1] < 0.1 * before, (before, hist)


def test_gelu_approximation_bound(gpu_api):
    """The kernels' GELU (x Phi(x), Phi by Hastings' rational approximation,
    Abramowitz and Stegun 26.2.17) and its derivative, on the device over
    [-12, 12]: within 5e-7 (gelu) and 6e-7 (derivative) times max(1, |x|)
    of the exact erf form (float64), and within 1e-6 times max(1, |x|) of
    the same through the device's erff and expf. A wrong coefficient or a
    dropped term misses by 1e-3 or more."""
    from scipy.special import erf
    print("mlp gelu bound", flush=True)
    x = np.linspace(-12.0, 12.0, 1_000_001).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, -0.0, 1e-30, -1e-30, 40.0, -40.0])])
    f = L._bind(gpu_api)
    dx = L.DeviceArray(gpu_api, x)
    out = L.DeviceArray(gpu_api, shape=(len(x), 4), dtype=np.float32)
    assert f.mlp_gelu(dx.ptr, len(x), out.ptr, None) == 0
    got = out.numpy().astype(np.float64)
    xd = x.astype(np.float64)
    cdf = 0.5 * (1.0 + erf(xd / np.sqrt(2.0)))
    gelu = xd * cdf
    grad = cdf + xd * np.exp(-0.5 * xd * xd) / np.sqrt(2.0 * np.pi)
    scale = np.maximum(1.0, np.abs(xd))
    assert np.all(np.isfinite(got))
    assert (np.abs(got[:, 0] - gelu) / scale).max() <= 5e-7
    assert (np.abs(got[:, 1] - grad) / scale).max() <= 6e-7
    assert (np.abs(got[:, 0] - got[:, 2]) / scale).max() <= 1e-6
    assert (np.abs(got[:, 1] - got[:, 3]) / scale).max() <= 1e-6
