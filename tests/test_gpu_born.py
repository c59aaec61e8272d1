"""GPU parity of the HIP Born path against the FP64 oracle (through the C ABI).

Tolerances: wavefields within 1e-5 relative L2 of
the FP64 oracle; images and adjoint dot-product tests within 1e-4."""
import sys

import numpy as np
import pytest

from oracle import born_oracle as O
from sep_code_amd import geometry as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def announce(msg):
    print(msg, file=sys.stderr, flush=True)


def random_field(r, n, scale=1.0):
    n1, n2, n3 = n
    p = np.zeros((n3, n2, n1))
    p[4:-4, 4:-4, 4:-4] = r.standard_normal((n3 - 8, n2 - 8, n1 - 8)) * scale
    return p


def velocity(r, n, d, dt=0.001, vlo=1500.0, vhi=3500.0):
    n1, n2, n3 = n
    v = r.uniform(vlo, vhi, (n3, n2, n1))
    return (v * dt) ** 2


@pytest.mark.parametrize("n", [(70, 36, 40), (128, 24, 30), (64, 64, 20)])
@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("ty", [8, 16, 32])
def test_step_parity(dev, n, form, ty):
    announce(f"step parity n={n} form={form} ty={ty}")
    r = np.random.default_rng(11)
    d = (10.0, 11.0, 12.0)
    V = velocity(r, n, d)
    bc_a, bc_b = 7.0, 0.03
    dev.set_tuning(ty, 2048)
    dev.create_space(d, bc_a, bc_b, n)
    dev.set_velocity(V)
    P = O.Problem(n, d, V, bc_a, bc_b)
    p1, p0 = random_field(r, n), random_field(r, n)
    # one step: only FP32 rounding separates the two
    g_cur, g_old = dev.propagate(form, 1, p1, p0)
    o_cur, o_old = P.propagate(form, 1, p1.ravel(), p0.ravel())
    assert rel(g_cur, o_cur) < 1e-6
    np.testing.assert_array_equal(g_old.ravel(), p1.ravel().astype(np.float32))
    # 60 steps
    g_cur, g_old = dev.propagate(form, 60, p1, p0)
    o_cur, o_old = P.propagate(form, 60, p1.ravel(), p0.ravel())
    assert rel(g_cur, o_cur) < 1e-5
    assert rel(g_old, o_old) < 1e-5
    # the zero halo is never written
    U = np.zeros(g_cur.shape, bool)
    U[4:-4, 4:-4, 4:-4] = True
    assert np.all(g_cur[~U] == 0)
    dev.set_tuning(0, 2048)


# (cells per lane, halo ring[, load lookahead of the plain ring step])
VARIANTS = [(2, 0), (4, 0), (1, 1), (4, 1, 1), (4, 1, 2)]
VARIANT_IDS = ["x2", "x4", "ring", "x4ring", "x4ring_pf2"]


def set_variant(dev, v):
    dev.set_x_per_lane(v[0])
    dev.set_halo_ring(v[1])
    dev.set_prefetch(v[2] if len(v) > 2 else 0)


@pytest.mark.parametrize("v", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("n", [(70, 36, 40), (130, 20, 24), (64, 64, 20), (71, 30, 24),
                               (260, 20, 24), (36, 40, 30)])
@pytest.mark.parametrize("ty", [8, 16, 32])
def test_step_parity_variants(dev, n, ty, v):
    """Forward-form step with two or four cells per lane (128- or 256-wide
    tiles, 8- or 16-byte accesses) or on the six-plane halo ring: oracle
    parity and agreement with the default kernel.  Grids whose n1 the width
    does not divide (70 and 130 for four cells, 71 for both) run the widest
    kernel they allow."""
    announce(f"step parity variant={v} n={n} ty={ty}")
    r = np.random.default_rng(12)
    d = (10.0, 11.0, 12.0)
    V = velocity(r, n, d)
    dev.set_tuning(ty, 2048)
    dev.set_x_per_lane(1)
    dev.create_space(d, 7.0, 0.03, n)
    dev.set_velocity(V)
    P = O.Problem(n, d, V, 7.0, 0.03)
    p1, p0 = random_field(r, n), random_field(r, n)
    g1, g1o = dev.propagate(0, 60, p1, p0)
    set_variant(dev, v)
    try:
        c2, _ = dev.propagate(0, 1, p1, p0)
        g2, g2o = dev.propagate(0, 60, p1, p0)
    finally:
        set_variant(dev, (0, -1))
        dev.set_tuning(0, 2048)
    o1, _ = P.propagate(0, 1, p1.ravel(), p0.ravel())
    o_cur, o_old = P.propagate(0, 60, p1.ravel(), p0.ravel())
    assert rel(c2, o1) < 1e-6
    assert rel(g2, o_cur) < 1e-5 and rel(g2o, o_old) < 1e-5
    assert rel(g2, g1) < 1e-6 and rel(g2o, g1o) < 1e-6
    U = np.zeros(g2.shape, bool)
    U[4:-4, 4:-4, 4:-4] = True
    assert np.all(g2[~U] == 0)


def make_shot(dev, n=(48, 44, 40), nt=120, jt=1, jimg=1, seed=3, cimg=1e-3, bc_a=8.0,
              bc_b=0.02, d=(10.0, 10.0, 10.0), rec_step=3):
    """Same shot on the device and in the oracle."""
    r = np.random.default_rng(seed)
    n1, n2, n3 = n
    dt = 0.001
    V = velocity(r, n, d, dt)
    ntr = (nt - 1) // jt + 9
    w = G.wavelet_derivative(G.ricker(25.0, dt * jt, ntr - 8), dt * jt)
    slocs, strc = G.source_points(n1, n2, n1 / 2 + 0.3, n2 / 2 - 0.4, 9.5, w, ntr)
    rlocs, nx, ny = G.surface_receivers(n, 4, 3.0, step=rec_step, origin=1)
    sinc = O.sinc_table(jt).astype(np.float32)
    dev.create_space(d, bc_a, bc_b, n)
    dev.set_velocity(V)
    dev.set_imaging_scale(cimg)
    dev.set_sinc_table(0, sinc)
    dev.set_sinc_table(1, sinc)
    dev.set_source(slocs, strc)
    dev.set_receivers(rlocs, None, ntr=ntr, nx=nx, ny=ny)
    # inactive receivers (outside the update set) are dropped on both sides
    rl = rlocs.copy()
    i1, i2, i3 = rl % n1, (rl // n1) % n2, rl // (n1 * n2)
    inside = (i1 >= 4) & (i1 < n1 - 4) & (i2 >= 4) & (i2 < n2 - 4) & (i3 >= 4) & (i3 < n3 - 4)
    rl[~inside] = -1
    P = O.Problem(n, d, V, bc_a, bc_b, src_locs=slocs, src_tr=strc.astype(np.float64),
                  sinc_s=sinc.astype(np.float64), rec_locs=rl, ntr_d=ntr,
                  sinc_d=sinc.astype(np.float64), jt=jt, nt=nt, jimg=jimg, cimg=cimg)
    return P, r


@pytest.mark.parametrize("n", [(64, 40, 36), (260, 20, 24)])
def test_two_slice_lookahead_is_bit_identical(dev, n):
    """The load lookahead changes only when loads are issued, not the
    arithmetic: the plain ring step, and the source propagation with the
    source injected inside the step, give the same bits with one and with
    two slices of lookahead (short z chunks included)."""
    announce(f"lookahead bit-exact n={n}")
    P, r = make_shot(dev, n=n, nt=80)
    p1, p0 = random_field(r, n), random_field(r, n)
    out = {}
    for pf in (1, 2):
        set_variant(dev, (4, 1, pf))
        try:
            a, b = dev.propagate(0, 30, p1, p0)
            s0, s1 = dev.source_prop(1, P.nt, 1, store="none", get_last=True)
        finally:
            set_variant(dev, (0, -1))
        out[pf] = (a, b, s0, s1)
    assert np.abs(out[1][2]).max() > 0
    for x, y in zip(out[1], out[2]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("store", ["f32", "bf16", "f16"])
def test_stream_loads_are_bit_identical(dev, store):
    """Streaming (nontemporal) loads change only the cache policy: the plain
    ring step, Born modelling (fused Born source, receivers extracted in the
    step) and the adjoint (fused imaging, receivers injected in the step)
    give the same bits with the hint on and off."""
    announce(f"stream loads bit-exact store={store}")
    n = (64, 40, 36)
    P, r = make_shot(dev, n=n, nt=60, jimg=2)
    p1, p0 = random_field(r, n), random_field(r, n)
    m = random_field(r, n)
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    out = {}
    for mode in (0, 1):
        set_variant(dev, (4, 1))
        dev.set_stream_loads(mode)
        try:
            a, b = dev.propagate(0, 30, p1, p0)
            dev.source_prop(1, P.nt, 2, store=store)
            dev.set_model(m)
            dev.rtm_forward(1, P.nt, 2)
            D = dev.get_receivers()
            dev.set_receivers(P.rlocs, d)
            dev.zero_image()
            dev.rtm_adjoint(1, P.nt, 2)
            out[mode] = (a, b, D, dev.get_image())
        finally:
            dev.set_stream_loads(-1)
            set_variant(dev, (0, -1))
    assert np.abs(out[0][3]).max() > 0 and np.abs(out[0][2]).max() > 0
    for x, y in zip(out[0], out[1]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("jt,jimg", [(1, 1), (2, 3)])
def test_source_prop_parity(dev, jt, jimg):
    announce(f"source_prop parity jt={jt} jimg={jimg}")
    P, r = make_shot(dev, jt=jt, jimg=jimg)
    p0, p1 = dev.source_prop(jt, P.nt, jimg, store="f32", get_last=True)
    snaps, last2 = P.source_prop()
    assert rel(p1, last2[1]) < 1e-5 and rel(p0, last2[0]) < 1e-5
    for k in (0, P.nimg // 2, P.nimg - 1):
        assert rel(dev.get_snapshot(k), snaps[k]) < 1e-5, k


@pytest.mark.parametrize("direct", [False, True])
@pytest.mark.parametrize("store", ["f32", "none", "ckpt"])
def test_born_forward_parity(dev, store, direct):
    """Born modelling against the oracle's transpose-form recursion, both on
    the default path (forward-form kernel on u = V.w.q) and on the direct
    transpose-form kernel."""
    announce(f"born_forward parity store={store} direct_lv={direct}")
    P, r = make_shot(dev, jimg=2)
    dev.set_born_path(direct)
    dev.source_prop(1, P.nt, 2, store=store, ckpt=17)
    m = random_field(r, P.n)
    dev.set_model(m)
    dev.rtm_forward(1, P.nt, 2)
    dev.set_born_path(False)
    D = dev.get_receivers()
    snaps, _ = P.source_prop()
    Do = P.born_forward(m.ravel(), snaps)
    assert rel(D, Do) < 1e-5


@pytest.mark.parametrize("v", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("store", ["f32", "bf16", "f16"])
def test_actions_variants(dev, store, v):
    """Born modelling (fused Born source) and the adjoint (fused imaging) on the
    two- and four-cells-per-lane kernels and the halo ring agree with the
    default kernel, and with the oracle for FP32 snapshots."""
    announce(f"born/adjoint variant={v} store={store}")
    P, r = make_shot(dev, n=(52, 44, 40), jimg=2)
    m = random_field(r, P.n)
    snaps, _ = P.source_prop()
    Do = P.born_forward(m.ravel(), snaps)
    out = {}
    for k in ((1, 0), v):
        set_variant(dev, k)
        try:
            dev.source_prop(1, P.nt, 2, store=store)
            dev.set_model(m)
            dev.rtm_forward(1, P.nt, 2)
            D = dev.get_receivers()
            dev.set_receivers(P.rlocs, Do)
            dev.zero_image()
            dev.rtm_adjoint(1, P.nt, 2)
            out[k] = (D, dev.get_image())
        finally:
            set_variant(dev, (0, -1))
    base = out[(1, 0)]
    assert rel(out[v][0], base[0]) < 1e-6
    assert rel(out[v][1], base[1]) < 1e-6
    if store == "f32":
        assert rel(out[v][0], Do) < 1e-5
        assert rel(out[v][1], P.rtm_adjoint(Do, snaps)) < 1e-4


@pytest.mark.parametrize("variant", [(4, 1), (0, -1)], ids=["x4ring", "small_grid"])
@pytest.mark.parametrize("jt,store", [(1, "f32"), (2, "f32"), (1, "bf16")])
def test_receiver_fusion(dev, jt, store, variant):
    """Receivers (and the source points) injected / extracted inside the step
    -- the four-cell ring kernel, or at this size with the automatic shape the
    one-cell kernel's 8-row small-grid tiles: traces bit-identical to the
    separate extraction kernel, images to FP32 rounding of the separate
    injection kernel, both against the oracle."""
    announce(f"receiver fusion jt={jt} store={store} variant={variant}")
    P, r = make_shot(dev, n=(52, 44, 40), jt=jt, jimg=2)
    m = random_field(r, P.n)
    snaps, _ = P.source_prop()
    Do = P.born_forward(m.ravel(), snaps)
    out = {}
    set_variant(dev, variant)
    try:
        for fuse in (0, 1):
            dev.set_receiver_fusion(fuse)
            dev.set_receivers(P.rlocs, None, ntr=P.ntr_d, nx=len(P.rlocs), ny=1)
            dev.source_prop(jt, P.nt, 2, store=store)
            dev.set_model(m)
            dev.rtm_forward(jt, P.nt, 2)
            D = dev.get_receivers()
            dev.set_receivers(P.rlocs, Do)
            dev.zero_image()
            dev.rtm_adjoint(jt, P.nt, 2)
            out[fuse] = (D, dev.get_image())
    finally:
        dev.set_receiver_fusion(1)
        set_variant(dev, (0, -1))
    np.testing.assert_array_equal(out[1][0], out[0][0])
    assert rel(out[1][1], out[0][1]) < 1e-6
    if store == "f32":
        assert rel(out[1][0], Do) < 1e-5
        assert rel(out[1][1], P.rtm_adjoint(Do, snaps)) < 1e-4


@pytest.mark.parametrize("variant", [(4, 1), (0, -1)], ids=["x4ring", "small_grid"])
@pytest.mark.parametrize("layout", ["duplicate", "two_planes"])
def test_receivers_off_one_plane_fall_back(dev, layout, variant):
    """Receivers sharing a cell, or on two depths, cannot go through the
    per-plane receiver map: the separate kernels run, and results still match
    the oracle."""
    announce(f"receiver fallback {layout} variant={variant}")
    P, r = make_shot(dev, n=(52, 44, 40), jimg=2)
    n1, n2 = P.n[0], P.n[1]
    rl = P.rlocs.copy()
    act = np.flatnonzero(rl >= 0)
    if layout == "duplicate":
        rl[act[1]] = rl[act[0]]
    else:
        rl[act[::2]] += n1 * n2 * 3
    Q = O.Problem(P.n, (10.0, 10.0, 10.0), P.V.reshape(P.n[::-1]), 8.0, 0.02, src_locs=P.slocs,
                  src_tr=P.strc, rec_locs=rl, ntr_d=P.ntr_d, jt=1, nt=P.nt, jimg=2, cimg=1e-3)
    m = random_field(r, P.n)
    snaps, _ = Q.source_prop()
    Do = Q.born_forward(m.ravel(), snaps)
    set_variant(dev, variant)
    try:
        dev.set_receivers(rl, None, ntr=P.ntr_d, nx=len(rl), ny=1)
        dev.source_prop(1, P.nt, 2, store="f32")
        dev.set_model(m)
        dev.rtm_forward(1, P.nt, 2)
        D = dev.get_receivers()
        dev.set_receivers(rl, Do)
        dev.zero_image()
        dev.rtm_adjoint(1, P.nt, 2)
        img = dev.get_image()
    finally:
        set_variant(dev, (0, -1))
    assert rel(D, Do) < 1e-5
    assert rel(img, Q.rtm_adjoint(Do, snaps)) < 1e-4


@pytest.mark.parametrize("store,bits", [("bf16", 8), ("f16", 11)])
def test_snapshot_codec_roundtrip(dev, store, bits):
    """Per-z-slice scaled 16-bit snapshots: every value within half an ulp of
    its slice maximum (2^-(bits+1) relative), zero slices stay zero."""
    announce(f"codec roundtrip {store}")
    P, r = make_shot(dev, nt=60)
    dev.source_prop(1, P.nt, 1, store="f32")
    ref = [dev.get_snapshot(k) for k in (0, 10, 30, 59)]
    dev.source_prop(1, P.nt, 1, store=store)
    for k, s32 in zip((0, 10, 30, 59), ref):
        s16 = dev.get_snapshot(k)
        assert np.isfinite(s16).all()
        smax = np.abs(s32).max(axis=(1, 2), keepdims=True)
        err = np.abs(s16 - s32)
        assert np.all(err <= smax * 2.0 ** -(bits + 1) * 1.01 + 1e-30), k
        assert np.all(s16[smax[:, 0, 0] == 0] == 0)


@pytest.mark.parametrize("store", ["f32", "ckpt"])
def test_rtm_adjoint_parity(dev, store):
    announce(f"rtm_adjoint parity store={store}")
    P, r = make_shot(dev, jimg=2)
    dev.source_prop(1, P.nt, 2, store=store, ckpt=13)
    snaps, _ = P.source_prop()
    m = random_field(r, P.n)
    Dobs = P.born_forward(m.ravel(), snaps)
    dev.set_receivers(P.rlocs, Dobs)
    dev.zero_image()
    dev.rtm_adjoint(1, P.nt, 2)
    img = dev.get_image()
    io = P.rtm_adjoint(Dobs, snaps)
    assert rel(img, io) < 1e-4


@pytest.mark.parametrize("ckpt", [2, 3, 7, 13, 200])
def test_checkpoint_adjoint_is_bit_identical_to_full_store(dev, ckpt):
    """Checkpoints written by the propagation itself and a segment recompute
    that starts from them give the same bits as storing every snapshot in
    FP32 (the same arithmetic on the same inputs), and leave the last two
    source fields as the plain propagation does."""
    announce(f"checkpoint bit-exact ckpt={ckpt}")
    P, r = make_shot(dev, nt=120, jimg=1)
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    d[P.rlocs < 0] = 0
    out = {}
    for store in ("f32", "ckpt"):
        p0, p1 = dev.source_prop(1, P.nt, 1, store=store, ckpt=ckpt, get_last=True)
        dev.set_receivers(P.rlocs, d)
        dev.zero_image()
        dev.rtm_adjoint(1, P.nt, 1)
        out[store] = (dev.get_image(), p0, p1)
    for a, b in zip(out["ckpt"], out["f32"]):
        np.testing.assert_array_equal(a, b)
    assert np.abs(out["f32"][0]).max() > 0


@pytest.mark.parametrize("store", ["f32", "bf16", "f16", "ckpt"])
def test_dot_product_gpu(dev, store):
    """<F m, d> = <m, F^T d> within 1e-4 with the FP32 GPU operator, for every
    source-wavefield store (both directions read the same snapshots)."""
    announce(f"dot test store={store}")
    P, r = make_shot(dev, n=(40, 36, 44), nt=100, jimg=1 if store != "f16" else 2)
    jimg = P.jimg
    dev.source_prop(1, P.nt, jimg, store=store, ckpt=9)
    m = random_field(r, P.n)
    dev.set_model(m)
    dev.rtm_forward(1, P.nt, jimg)
    Fm = dev.get_receivers().astype(np.float64)
    d = r.standard_normal(Fm.shape)
    d[P.rlocs < 0] = 0
    dev.set_receivers(P.rlocs, d)
    dev.zero_image()
    dev.rtm_adjoint(1, P.nt, jimg)
    Ftd = dev.get_image().astype(np.float64)
    lhs, rhs = (Fm * d).sum(), (m * Ftd).sum()
    assert abs(lhs - rhs) / abs(lhs) < 1e-4, (lhs, rhs)


@pytest.mark.parametrize("store,tol", [("bf16", 2e-2), ("f16", 2e-3)])
def test_compressed_store_image_error(dev, store, tol):
    """Image from BF16 / FP16 snapshots (per-z-slice FP32 scale) against the
    FP64 oracle with exact snapshots: error bounded by the format."""
    announce(f"compressed image error store={store}")
    P, r = make_shot(dev, jimg=4)
    dev.source_prop(1, P.nt, 4, store=store)
    snaps, _ = P.source_prop()
    m = random_field(r, P.n)
    Dobs = P.born_forward(m.ravel(), snaps)
    dev.set_receivers(P.rlocs, Dobs)
    dev.zero_image()
    dev.rtm_adjoint(1, P.nt, 4)
    e = rel(dev.get_image(), P.rtm_adjoint(Dobs, snaps))
    announce(f"  image rel err {store} = {e:.3e}")
    assert e < tol


def test_cfl_violation_refused(dev):
    from sep_code_amd._lib import BrtmError
    n = (32, 32, 32)
    dev.create_space((10.0, 10.0, 10.0), 4.0, 0.01, n)
    V = np.full((32, 32, 32), (5000 * 0.001) ** 2, np.float32)  # v dt / d = 0.5
    with pytest.raises(BrtmError, match="stability"):
        dev.set_velocity(V)
    dev.set_velocity(V * 0.8)  # 0.447 < 0.4529


def test_adjoint_requires_store(dev):
    from sep_code_amd._lib import BrtmError
    P, r = make_shot(dev, nt=20)
    dev.source_prop(1, 20, 1, store="none")
    with pytest.raises(BrtmError, match="stored source"):
        dev.rtm_adjoint(1, 20, 1)


@pytest.mark.parametrize("ckpt", [2, 7])
def test_checkpoints_after_propagating_fields_with_halos(dev, ckpt):
    """brtm_propagate takes whole caller fields, halos included, into the
    context's work buffers; a checkpointed source propagation afterwards
    writes those buffers out of place and must not read the stale halos.
    The image and last fields equal the full FP32 store's bit for bit."""
    announce(f"checkpoints after nonzero-halo propagate ckpt={ckpt}")
    P, r = make_shot(dev, nt=60, jimg=1)
    shape = (P.n[2], P.n[1], P.n[0])
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    d[P.rlocs < 0] = 0
    out = {}
    for store in ("f32", "ckpt"):
        # every cell nonzero, the halo too
        dev.propagate(0, 2, r.uniform(1, 2, shape), r.uniform(1, 2, shape))
        p0, p1 = dev.source_prop(1, P.nt, 1, store=store, ckpt=ckpt, get_last=True)
        dev.set_receivers(P.rlocs, d)
        dev.zero_image()
        dev.rtm_adjoint(1, P.nt, 1)
        out[store] = (dev.get_image(), p0, p1)
    for a, b in zip(out["ckpt"], out["f32"]):
        np.testing.assert_array_equal(a, b)
    assert np.abs(out["f32"][0]).max() > 0


@pytest.mark.parametrize("store", ["f32", "f16", "ckpt", "none"])
def test_graph_launch_is_bit_identical(dev, store):
    """Source propagation, Born modelling and the adjoint launched as graphs
    of their captured kernels give the same bits as launched one by one:
    three rounds in a row, with eager work in between, each checked."""
    announce(f"graph launch store={store}")
    P, r = make_shot(dev, n=(52, 44, 40), nt=60, jimg=2)
    shape = (P.n[2], P.n[1], P.n[0])
    m = random_field(r, P.n)
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    d[P.rlocs < 0] = 0

    def run():
        dev.set_receivers(P.rlocs, None, ntr=P.ntr_d, nx=len(P.rlocs), ny=1)
        dev.source_prop(1, P.nt, 2, store=store, ckpt=7)
        dev.set_model(m)
        dev.rtm_forward(1, P.nt, 2)
        D = dev.get_receivers()
        if store == "none":
            return (D,)
        dev.set_receivers(P.rlocs, d)
        dev.zero_image()
        dev.rtm_adjoint(1, P.nt, 2)
        return D, dev.get_image()

    dev.set_graphs(0)
    ref = run()
    assert np.abs(ref[0]).max() > 0
    dev.set_graphs(1)
    try:
        for k in range(3):
            got = run()
            for a, b in zip(got, ref):
                np.testing.assert_array_equal(a, b)
            dev.propagate(0, 2, r.uniform(1, 2, shape), r.uniform(1, 2, shape))  # eager work
    finally:
        dev.set_graphs(-1)
