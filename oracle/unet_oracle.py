"""FP64 numpy restatement of the learned preconditioner's U-Net
(sep_code_amd/csrc/unet.hip) -- TEST INFRASTRUCTURE ONLY.

Same parameter layout as the library (flat FP32 vector):
  first conv 1->16:  W [27][16], b [16]
  9 convs (cin->cout) e1b 16->16, e2a 16->32, e2b 32->32, b1 32->64,
  b2 64->64, d2a 96->32, d2b 32->32, d1a 48->16, d1b 16->16:
                     W [27][cout][cin], b [cout]
  head 1x1x1 16->1:  W [16], b [1]
Tap t = 9 (dz+1) + 3 (dy+1) + (dx+1).  Activations are channels-last
[n][z][y][x][c]; zero padding; ReLU; 2x2x2 average pooling; nearest x2
upsampling; decoder inputs are [upsampled, skip] concatenations; out =
x + head(d1b).

Pinned by tests/test_unet_oracle.py: the convolution against
torch.nn.functional.conv3d in FP64 on the CPU (an independent library), and
the backward pass against central finite differences.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16),
          (16, 16)]


def layout():
    off = 0
    lay = {}
    lay["w0"] = (off, (27, 16)); off += 27 * 16
    lay["b0"] = (off, (16,)); off += 16
    for i, (ci, co) in enumerate(LAYERS):
        lay[f"w{i + 1}"] = (off, (27, co, ci)); off += 27 * co * ci
        lay[f"b{i + 1}"] = (off, (co,)); off += co
    lay["wh"] = (off, (16,)); off += 16
    lay["bh"] = (off, (1,)); off += 1
    return lay, off


def unpack(theta):
    lay, n = layout()
    assert theta.size == n
    return {k: theta[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for k, (o, s) in lay.items()}


def pack(p):
    lay, n = layout()
    out = np.zeros(n)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = p[k].ravel()
    return out


def _taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def _shift(xp, dz, dy, dx, D, H, W):
    return xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]


def conv(x, W, b=None):
    """x [n,D,H,W,ci], W [27][co][ci] -> [n,D,H,W,co]."""
    n, D, H, Wd, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, D, H, Wd, W.shape[1]))
    for t, dz, dy, dx in _taps():
        out += _shift(xp, dz, dy, dx, D, H, Wd) @ W[t].T
    if b is not None:
        out += b
    return out


def conv_input_grad(gy, W):
    """Adjoint of conv in x: gx = sum_t shift_{-t}(gy @ W[t])."""
    n, D, H, Wd, co = gy.shape
    gp = np.pad(gy, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    gx = np.zeros((n, D, H, Wd, W.shape[2]))
    for t, dz, dy, dx in _taps():
        gx += _shift(gp, -dz, -dy, -dx, D, H, Wd) @ W[t]
    return gx


def conv_weight_grad(x, gy):
    n, D, H, Wd, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    g = np.zeros((27, gy.shape[-1], ci))
    G = gy.reshape(-1, gy.shape[-1])
    for t, dz, dy, dx in _taps():
        g[t] = G.T @ _shift(xp, dz, dy, dx, D, H, Wd).reshape(-1, ci)
    return g, G.sum(0)


def pool(x):
    n, D, H, W, c = x.shape
    return x.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).mean(axis=(2, 4, 6))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(g):
    n, D, H, W, c = g.shape
    return g.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).sum(axis=(2, 4, 6))


def relu(x):
    return np.maximum(x, 0.0)


def forward(theta, x):
    """x [n,D,H,W] -> (out [n,D,H,W], cache)."""
    p = unpack(np.asarray(theta, np.float64))
    x0 = np.asarray(x, np.float64)[..., None]
    c = {}
    w0 = p["w0"][:, :, None]  # [27][16][1]
    c["e1a"] = relu(conv(x0, w0, p["b0"]))
    c["e1b"] = relu(conv(c["e1a"], p["w1"], p["b1"]))
    c["p1"] = pool(c["e1b"])
    c["e2a"] = relu(conv(c["p1"], p["w2"], p["b2"]))
    c["e2b"] = relu(conv(c["e2a"], p["w3"], p["b3"]))
    c["p2"] = pool(c["e2b"])
    c["b1"] = relu(conv(c["p2"], p["w4"], p["b4"]))
    c["b2"] = relu(conv(c["b1"], p["w5"], p["b5"]))
    c["cat2"] = np.concatenate([up(c["b2"]), c["e2b"]], -1)
    c["d2a"] = relu(conv(c["cat2"], p["w6"], p["b6"]))
    c["d2b"] = relu(conv(c["d2a"], p["w7"], p["b7"]))
    c["cat1"] = np.concatenate([up(c["d2b"]), c["e1b"]], -1)
    c["d1a"] = relu(conv(c["cat1"], p["w8"], p["b8"]))
    c["d1b"] = relu(conv(c["d1a"], p["w9"], p["b9"]))
    out = x0[..., 0] + c["d1b"] @ p["wh"] + p["bh"][0]
    c["x0"] = x0
    return out, c, p


def loss_grad(theta, x, t):
    """MSE loss mean((out - t)^2) and its gradient w.r.t. theta (FP64)."""
    out, c, p = forward(theta, x)
    n = out.size
    r = out - t
    loss = float((r * r).sum() / n)
    g_out = 2.0 * r / n
    g = {}
    g["wh"] = c["d1b"].reshape(-1, 16).T @ g_out.ravel()
    g["bh"] = np.array([g_out.sum()])
    gy = g_out[..., None] * p["wh"]

    def layer(i, gy, y, xin, need_x=True):
        gpre = gy * (y > 0)
        g[f"w{i}"], g[f"b{i}"] = conv_weight_grad(xin, gpre)
        return conv_input_grad(gpre, p[f"w{i}"]) if need_x else None

    g_d1a = layer(9, gy, c["d1b"], c["d1a"])
    g_cat1 = layer(8, g_d1a, c["d1a"], c["cat1"])
    g_d2b = up_bwd(g_cat1[..., :32])
    g_e1b = g_cat1[..., 32:].copy()
    g_d2a = layer(7, g_d2b, c["d2b"], c["d2a"])
    g_cat2 = layer(6, g_d2a, c["d2a"], c["cat2"])
    g_b2 = up_bwd(g_cat2[..., :64])
    g_e2b = g_cat2[..., 64:].copy()
    g_b1 = layer(5, g_b2, c["b2"], c["b1"])
    g_p2 = layer(4, g_b1, c["b1"], c["p2"])
    g_e2b += pool_bwd(g_p2)
    g_e2a = layer(3, g_e2b, c["e2b"], c["e2a"])
    g_p1 = layer(2, g_e2a, c["e2a"], c["p1"])
    g_e1b += pool_bwd(g_p1)
    g_e1a = layer(1, g_e1b, c["e1b"], c["e1a"])
    gpre = g_e1a * (c["e1a"] > 0)
    gw, gb = conv_weight_grad(c["x0"], gpre)
    g["w0"], g["b0"] = gw[:, :, 0], gb
    return loss, pack(g), out
