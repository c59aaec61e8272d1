"""CPU pins of the U-Net oracle (oracle/unet_oracle.py): its convolution
against torch's CPU conv3d in FP64, its backward pass against central finite
differences."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_conv_matches_torch_conv3d_fp64():
    torch = pytest.importorskip("torch")
    r = np.random.default_rng(0)
    x = r.standard_normal((2, 6, 5, 7, 3))
    W = r.standard_normal((27, 4, 3))
    b = r.standard_normal(4)
    ours = U.conv(x, W, b)
    wt = np.zeros((4, 3, 3, 3, 3))
    for t in range(27):
        wt[:, :, t // 9, (t // 3) % 3, t % 3] = W[t]
    ref = torch.nn.functional.conv3d(torch.from_numpy(x).permute(0, 4, 1, 2, 3),
                                     torch.from_numpy(wt), torch.from_numpy(b), padding=1)
    np.testing.assert_allclose(ours, ref.permute(0, 2, 3, 4, 1).numpy(), rtol=1e-12, atol=1e-12)


def test_conv_adjoints():
    r = np.random.default_rng(1)
    x = r.standard_normal((1, 4, 6, 5, 3))
    W = r.standard_normal((27, 2, 3))
    gy = r.standard_normal((1, 4, 6, 5, 2))
    lhs = (U.conv(x, W) * gy).sum()
    assert np.isclose(lhs, (x * U.conv_input_grad(gy, W)).sum(), rtol=1e-12)
    gw, _ = U.conv_weight_grad(x, gy)
    assert np.isclose(lhs, (W * gw).sum(), rtol=1e-12)
    p = r.standard_normal((1, 4, 6, 8, 2))
    q = r.standard_normal((1, 2, 3, 4, 2))
    assert np.isclose((U.pool(p) * q).sum(), (p * U.pool_bwd(q)).sum(), rtol=1e-12)
    assert np.isclose((U.up(q) * p).sum(), (q * U.up_bwd(p)).sum(), rtol=1e-12)


def test_unet_backward_matches_finite_differences():
    r = np.random.default_rng(2)
    lay, n = U.layout()
    theta = r.standard_normal(n) * 0.1
    for k, (o, s) in lay.items():  # He scaling keeps activations O(1)
        if k.startswith("w") and len(s) == 3:
            theta[o:o + int(np.prod(s))] *= 10 * np.sqrt(2.0 / (27 * s[2]))
    o, _ = lay["wh"]
    theta[o:o + 16] = r.standard_normal(16) * 0.5
    x = r.standard_normal((1, 8, 8, 8))
    t = r.standard_normal((1, 8, 8, 8))
    loss, g, _ = U.loss_grad(theta, x, t)
    idx = []
    for k in ("w0", "b0", "w1", "w3", "w5", "b5", "w6", "w8", "w9", "b9", "wh", "bh"):
        o, s = lay[k]
        sz = int(np.prod(s))
        idx += list(o + r.choice(sz, min(2, sz), replace=False))
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (U.loss_grad(tp, x, t)[0] - U.loss_grad(tm, x, t)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
