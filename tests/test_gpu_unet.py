"""GPU tests of the learned preconditioner: BF16-MFMA U-Net forward and
gradients against the FP64 oracle (oracle/unet_oracle.py), training, and use
inside flexible CG."""
import sys

import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, file=sys.stderr, flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def net(dev):
    from sep_code_amd.unet import UNet
    dev.create_space((10.0, 10.0, 10.0), 4.0, 0.0, (16, 16, 16))
    u = UNet(dev, batch=2, patch=64, seed=3)
    yield u
    u.close()


def test_param_layout_matches_oracle(net):
    assert net.nparams == U.layout()[1]


def _theta(net, head=0.5, seed=0):
    theta = net.get_params().astype(np.float64)
    lay, _ = U.layout()
    r = np.random.default_rng(seed)
    o, _ = lay["wh"]
    theta[o:o + 16] = r.standard_normal(16) * head
    for k, (o, s) in lay.items():  # non-zero biases exercise the bias paths
        if k.startswith("b") and k != "bh":
            theta[o:o + int(np.prod(s))] = r.standard_normal(int(np.prod(s))) * 0.05
    return theta


def test_forward_matches_fp64_oracle(net):
    announce("unet forward vs oracle")
    theta = _theta(net)
    net.set_params(theta)
    r = np.random.default_rng(1)
    x = r.standard_normal((2, 64, 64, 64)).astype(np.float32)
    y = net.forward(x)
    yo, _, _ = U.forward(net.get_params(), x)
    # compare the network part (out - x); BF16 weights and activations
    e = rel(y - x, yo - x)
    announce(f"  forward rel err {e:.3e}")
    assert e < 3e-2


def test_gradients_match_fp64_oracle(net):
    announce("unet gradients vs oracle")
    theta = _theta(net, seed=2)
    net.set_params(theta)
    r = np.random.default_rng(2)
    x = r.standard_normal((1, 64, 64, 64)).astype(np.float32)
    t = r.standard_normal((1, 64, 64, 64)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    lo, go, _ = U.loss_grad(net.get_params(), x, t)
    assert abs(loss - lo) / lo < 1e-2
    lay, _ = U.layout()
    for k, (o, s) in lay.items():
        n = int(np.prod(s))
        e = rel(g[o:o + n], go[o:o + n])
        announce(f"  grad {k}: rel err {e:.3e}")
        assert e < 8e-2, k


def test_training_reduces_loss(net):
    announce("unet training")
    from sep_code_amd.unet import UNet
    u = UNet(net.dev, batch=2, patch=64, seed=11)
    r = np.random.default_rng(5)
    x = r.standard_normal((2, 64, 64, 64)).astype(np.float32)
    # target: a smoothed and rescaled copy (learnable by 3x3x3 convolutions)
    t = 0.5 * x.copy()
    for ax in (1, 2, 3):
        t = (t + np.roll(t, 1, ax) + np.roll(t, -1, ax)) / 3.0
    u._stage(u.xbuf, x)
    u._stage(u.tbuf, t)
    losses = [u.train_step_device(2, 1e-3) for _ in range(60)]
    announce(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.3 * losses[0]
    u.close()


def test_apply_covers_volume_and_is_scale_equivariant(net):
    from sep_code_amd.operator import DeviceVector
    net.set_params(_theta(net, head=0.2))
    dims = (100, 90, 70)
    r = np.random.default_rng(7)
    g = r.standard_normal(dims[::-1]).astype(np.float32)
    gv = DeviceVector.from_array(net.dev, g)
    zv = DeviceVector(net.dev, g.size)
    net.apply(gv, zv, dims, overlap=8)
    z = zv.numpy().reshape(g.shape)
    assert np.isfinite(z).all() and np.count_nonzero(z) == z.size
    # z = rms(g) * net(g / rms(g)): a power-of-two factor passes through the
    # FP64 rms and the FP32 scalings exactly, so the BF16 network sees the
    # same input bits; another factor changes last bits of g / rms, which the
    # BF16 input rounding can amplify to one BF16 ulp (2^-8) per value.
    for f, tol in ((4.0, 1e-6), (3.0, 1e-2)):
        gv2 = DeviceVector.from_array(net.dev, f * g)
        zv2 = DeviceVector(net.dev, g.size)
        net.apply(gv2, zv2, dims, overlap=8)
        z2 = zv2.numpy().astype(np.float64)
        err = np.linalg.norm(z2 - f * z.ravel()) / np.linalg.norm(f * z.ravel())
        assert err < tol, (f, err)
    # the tile with patch origin (40, 40, 40) equals rms * net(patch / rms)
    # on its central 48^3 (clipped to the volume), zeros outside the volume;
    # the scalings are the library's FP32 ones so the BF16 inputs agree
    rms = np.sqrt(float(np.dot(g.ravel().astype(np.float64), g.ravel().astype(np.float64)))
                  / g.size)
    patch = np.zeros((64, 64, 64), np.float32)
    src = g[40:104, 40:104, 40:104]
    patch[:src.shape[0], :src.shape[1], :src.shape[2]] = src * np.float32(1.0 / rms)
    y = net.forward(patch[None])[0] * np.float32(rms)
    got, want = z[48:70, 48:90, 48:96], y[8:30, 8:50, 8:56]
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-2
    np.testing.assert_allclose(got, want, rtol=2e-2, atol=2e-2 * np.abs(want).max())


def test_split_preconditioner_with_identity_network_is_the_diagonal(net):
    """With a zero head the residual network is the identity, so the
    preconditioner callback returns W^2 g with a split weight W, and g
    without one (the sign and RMS scalings cancel)."""
    from sep_code_amd.operator import PRECOND_FN, DeviceVector
    net.set_params(_theta(net, head=0.0))
    dims = (100, 90, 70)
    r = np.random.default_rng(8)
    g = r.standard_normal(dims[::-1]).astype(np.float32).ravel()
    W = r.uniform(0.1, 1.0, g.size).astype(np.float32)
    gv = DeviceVector.from_array(net.dev, g)
    wv = DeviceVector.from_array(net.dev, W)
    zv = DeviceVector(net.dev, g.size)
    _, addr, user = net.precond(dims, overlap=8, split_weight=wv)
    fn = PRECOND_FN(addr)
    assert fn(user, gv.h, zv.h) == 0
    want = W.astype(np.float64) ** 2 * g
    np.testing.assert_allclose(zv.numpy(), want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    _, addr, user = net.precond(dims, overlap=8)
    assert fn(user, gv.h, zv.h) == 0
    np.testing.assert_allclose(zv.numpy(), g, rtol=1e-5, atol=1e-6 * np.abs(g).max())


def test_split_guard_replaces_a_fixed_direction(net):
    """Guard mode 0: a network whose output ignores its input (whole
    network, head weights zero, a large head bias: net(x) = x + b) gives
    nearly the same direction on every call; the guard then returns the
    diagonal direction W^2 g."""
    from sep_code_amd.operator import PRECOND_FN, DeviceVector
    from oracle import unet_oracle as U
    theta = _theta(net, head=0.0)
    lay, _ = U.layout()
    o, _ = lay["bh"]
    theta[o] = 50.0
    net.set_params(theta)
    net.set_odd(False)
    net.set_guard_mode(0)
    dims = (40, 36, 30)
    r = np.random.default_rng(9)
    W = r.uniform(0.2, 1.0, int(np.prod(dims))).astype(np.float32)
    wv = DeviceVector.from_array(net.dev, W)
    zv = DeviceVector(net.dev, W.size)
    try:
        _, addr, user = net.precond(dims, overlap=8, split_weight=wv)
        fn = PRECOND_FN(addr)
        outs = []
        for k in range(3):
            g = r.standard_normal(W.size).astype(np.float32)
            gv = DeviceVector.from_array(net.dev, g)
            assert fn(user, gv.h, zv.h) == 0
            outs.append((g, zv.numpy()))
        assert net.guard_count() >= 2
        g, z = outs[-1]
        want = W.astype(np.float64) ** 2 * g
        np.testing.assert_allclose(z, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    finally:
        net.set_guard_mode(1)
        net.set_odd(True)
        net.set_params(_theta(net))


def test_split_guard_keeps_the_new_part_of_the_correction(net):
    """Guard mode 1 against its rule restated in numpy.  The network's raw
    split outputs z_k come from the same calls with the guard's thresholds
    out of reach (mode 0, repeat cos 1, min cos 0).  With c_k = z_k - W^2 g_k
    (the network's correction), mode 1 returns W^2 g_k + c_k, less the
    projection of c_k on c_{k-1} when their cos exceeds 0.8, and W^2 g_k when
    that is no descent direction or has cos < 0.05 with g_k.  The network
    here has a large fixed head bias (a repeating part) and a head that
    follows its input (a new part)."""
    from sep_code_amd.operator import PRECOND_FN, DeviceVector
    from oracle import unet_oracle as U
    theta = _theta(net, head=0.5, seed=4)
    lay, _ = U.layout()
    o, _ = lay["bh"]
    theta[o] = 10.0
    net.set_params(theta)
    net.set_odd(False)
    dims = (40, 36, 30)
    r = np.random.default_rng(11)
    W = r.uniform(0.2, 1.0, int(np.prod(dims))).astype(np.float32)
    gs = [r.standard_normal(W.size).astype(np.float32) for _ in range(4)]
    wv = DeviceVector.from_array(net.dev, W)
    zv = DeviceVector(net.dev, W.size)

    def run(mode, rep, mn):
        net.set_guard(rep, mn)
        net.set_guard_mode(mode)
        _, addr, user = net.precond(dims, overlap=8, split_weight=wv)
        fn = PRECOND_FN(addr)
        out = []
        for g in gs:
            gv = DeviceVector.from_array(net.dev, g)
            assert fn(user, gv.h, zv.h) == 0
            out.append(zv.numpy().astype(np.float64))
        return out

    try:
        raw = run(0, 1.0, 0.0)
        assert net.guard_count() == 0
        got = run(1, 0.8, 0.05)
        Wd = W.astype(np.float64)
        prev = None
        projected = 0
        for g, zr, z in zip(gs, raw, got):
            d = Wd * Wd * g
            c = zr - d
            cn = c.copy()
            if prev is not None:
                cp, pp, cc = c @ prev, prev @ prev, c @ c
                if cp > 0.8 * np.sqrt(cc * pp):
                    cn = c - (cp / pp) * prev
                    projected += 1
            prev = c
            want = d + cn
            zg = want @ g
            if not zg > 0 or zg < 0.05 * np.sqrt((want @ want) * (g @ g)):
                want = d
            err = np.linalg.norm(z - want) / np.linalg.norm(want)
            assert err < 1e-5, err
        assert projected >= 1  # the fixed bias makes the corrections repeat
        assert net.guard_count() >= projected
    finally:
        net.set_guard(0.8, 0.05)
        net.set_odd(True)
        net.set_params(_theta(net))
