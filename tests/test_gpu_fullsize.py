"""Full-size GPU checks through properties that hold at any size.

The oracle finishes in seconds only on small grids, so at the production size
(512^3) and at the largest grid the library accepts (just under 2^31 cells)
the HIP path is checked through properties instead:

* locality: with the sponge off, a field confined to a small box evolves
  exactly as the same box on a small grid, so a window of the largest grid
  is compared with the FP64 oracle (wavefield tolerance 1e-5);
* the adjoint dot-product test of Born modelling and RTM imaging (1e-4);
* linearity in the data;
* checkpoint recompute against the FP32 store of every snapshot (bit-exact).
"""
import sys

import numpy as np
import pytest

from oracle import born_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def announce(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def bigdev():
    """A context of its own, closed (and its device memory freed) at the end
    of the module."""
    from sep_code_amd.device import BornDevice
    d = BornDevice(0)
    yield d
    d.close()


def test_grid_of_2_31_cells_refused(bigdev):
    """Cell indices are 32-bit: a grid of 2^31 cells is refused before any
    allocation, and the context stays usable."""
    from sep_code_amd._lib import BrtmError
    with pytest.raises(BrtmError, match="2\\^31"):
        bigdev.create_space((10.0, 10.0, 10.0), 0.0, 0.0, (1024, 1024, 2048))
    bigdev.create_space((10.0, 10.0, 10.0), 0.0, 0.0, (32, 32, 32))


def test_largest_grid_window_matches_oracle(bigdev):
    """1024 x 1024 x 2047 cells (2^31 - 1048576): a random 16^3 box in the
    far corner, where cell indices exceed 2.1e9, advanced 3 steps with the
    sponge off.  The field spreads 4 cells per step and stays clear of the
    zero boundary, so the 64^3 corner window equals the oracle on a 64^3 grid
    holding the same box, and nothing outside the window is touched."""
    n = (1024, 1024, 2047)
    m = 64
    d = (10.0, 10.0, 10.0)
    vdt2 = (2000.0 * 0.001) ** 2
    nsteps = 3
    announce(f"largest grid {n}: {n[0] * n[1] * n[2]} cells")
    r = np.random.default_rng(31)
    box0 = r.standard_normal((16, 16, 16))
    box1 = r.standard_normal((16, 16, 16))

    small_cur = np.zeros((m, m, m))
    small_old = np.zeros((m, m, m))
    small_cur[24:40, 24:40, 24:40] = box0
    small_old[24:40, 24:40, 24:40] = box1
    P = O.Problem((m, m, m), d, np.full(m ** 3, vdt2), 0.0, 0.0)
    o_cur, o_old = P.propagate(0, nsteps, small_cur.ravel(), small_old.ravel())
    o_cur = o_cur.reshape(m, m, m)
    # the support stays inside the 64^3 window with room to spare
    assert np.abs(o_cur[:12]).max() == 0 and np.abs(o_cur[52:]).max() == 0

    bigdev.create_space(d, 0.0, 0.0, n)
    V = np.full((n[2], n[1], n[0]), vdt2, np.float32)
    bigdev.set_velocity(V)
    del V
    cur = np.zeros((n[2], n[1], n[0]), np.float32)
    old = np.zeros((n[2], n[1], n[0]), np.float32)
    w = (slice(n[2] - m, n[2]), slice(n[1] - m, n[1]), slice(n[0] - m, n[0]))
    cur[w] = small_cur
    old[w] = small_old
    announce("largest grid: propagating")
    g_cur, g_old = bigdev.propagate(0, nsteps, cur, old)
    del cur, old
    win_cur = g_cur[w].copy()
    win_old = g_old[w].copy()
    nz = np.count_nonzero(g_cur)
    del g_cur, g_old
    bigdev.create_space(d, 0.0, 0.0, (32, 32, 32))  # frees the large buffers
    assert rel(win_cur, o_cur) < 1e-5
    assert rel(win_old, o_old.reshape(m, m, m)) < 1e-5
    assert nz == np.count_nonzero(win_cur)


@pytest.fixture(scope="module")
def shot512(bigdev):
    """The 512^3 production shot (layered velocity with anomalies, 32-cell
    sponge, 512^2 surface receivers), 60 steps, set up on the device."""
    from sep_code_amd import configs
    w = configs.rtm512(nt=60)
    shot = w.shots[0]
    sinc = np.zeros((1, 8), np.float32)
    sinc[0, 3] = 1
    bigdev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    bigdev.set_velocity(w.V)
    bigdev.set_imaging_scale(w.cimg)
    bigdev.set_sinc_table(0, sinc)
    bigdev.set_sinc_table(1, sinc)
    bigdev.set_source(shot.src_locs, shot.src_traces)
    bigdev.set_receivers(shot.rec_locs, None, ntr=w.ntr, nx=shot.nrx, ny=shot.nry)
    bigdev.select_shot(0)
    return w, shot


def _image(dev, w, shot, d):
    dev.set_receivers(shot.rec_locs, d, nx=shot.nrx, ny=shot.nry)
    dev.zero_image()
    dev.rtm_adjoint(w.jt, w.nt, w.jimg)
    return dev.get_image()


def test_rtm512_dot_product_linearity_and_checkpoints(bigdev, shot512):
    """At 512^3: <F m, d> = <m, F^T d> within 1e-4 (FP64 sums); the image of
    2d is twice the image of d (scaling by 2 is exact in binary floating
    point, up to rounding in the denormal range); and the image from
    checkpoints every 16 steps with segment recompute has the same bits as
    the image from FP32 snapshots of every step."""
    w, shot = shot512
    r = np.random.default_rng(512)
    m = np.zeros((w.n[2], w.n[1], w.n[0]), np.float32)
    m[4:-4, 4:-4, 4:-4] = r.standard_normal((w.n[2] - 8, w.n[1] - 8, w.n[0] - 8),
                                            dtype=np.float32)
    announce("512^3: source propagation (FP32 store) and Born modelling")
    bigdev.source_prop(w.jt, w.nt, w.jimg, store="f32")
    bigdev.set_model(m)
    bigdev.rtm_forward(w.jt, w.nt, w.jimg)
    Fm = bigdev.get_receivers().astype(np.float64)
    d = r.standard_normal(Fm.shape).astype(np.float32)
    announce("512^3: adjoint imaging of d and 2d")
    img = _image(bigdev, w, shot, d)
    img2 = _image(bigdev, w, shot, 2 * d)
    lhs = float((Fm * d).sum())
    rhs = float((m.astype(np.float64) * img).sum())
    announce(f"512^3 dot test: {lhs:.9e} {rhs:.9e} rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert np.abs(img).max() > 0
    assert abs(lhs - rhs) / abs(lhs) < 1e-4
    assert rel(img2, 2.0 * img.astype(np.float64)) < 1e-12

    announce("512^3: checkpoints every 16 steps")
    bigdev.release_store(0)
    bigdev.source_prop(w.jt, w.nt, w.jimg, store="ckpt", ckpt=16)
    img_ck = _image(bigdev, w, shot, d)
    bigdev.release_store(0)
    np.testing.assert_array_equal(img_ck, img)
