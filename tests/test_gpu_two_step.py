"""Two source-field steps per pass (temporal blocking, brtm_set_two_step)
against the steps one by one on the four-cell ring kernel, whose arithmetic
each step of the pass repeats operation for operation: the source snapshots,
the last two fields, the checkpoints, the recomputed segments and the image
are the same bits with the pass on and off.  Grids with partial 64 x 32
tiles, several z chunks, odd and even nt, jt = 1 and 2 (the second step's
own source sample and taps)."""
import numpy as np
import pytest

from test_gpu_born import announce, make_shot, rel, set_variant

pytestmark = pytest.mark.gpu

SHAPES = [(64, 40, 36), (132, 52, 44), (200, 100, 30)]


def with_mode(dev, mode, fn):
    set_variant(dev, (4, 1))
    dev.set_two_step(mode)
    try:
        return fn()
    finally:
        dev.set_two_step(0)
        set_variant(dev, (0, -1))


@pytest.mark.parametrize("n", SHAPES)
@pytest.mark.parametrize("nt,jt", [(61, 1), (80, 2)])
def test_pass_gives_the_single_steps_bits(dev, n, nt, jt):
    announce(f"two-step pass: source field n={n} nt={nt} jt={jt}")
    P, r = make_shot(dev, n=n, nt=nt, jt=jt)

    def run():
        p0, p1 = dev.source_prop(jt, nt, 1, store="f32", get_last=True)
        return [p0, p1] + [dev.get_snapshot(k) for k in range(nt)]

    one = with_mode(dev, 0, run)
    two = with_mode(dev, 1, run)
    assert np.abs(one[1]).max() > 0
    for k, (a, b) in enumerate(zip(one, two)):
        np.testing.assert_array_equal(a, b, err_msg=f"field {k}")


@pytest.mark.parametrize("ckpt", [2, 5, 7, 50])
@pytest.mark.parametrize("nt", [101, 120])
def test_pass_checkpoints_and_recompute(dev, ckpt, nt):
    """Checkpoints written by the pass, segments recomputed by it (segments
    that start at odd steps begin with one single step), imaging every step:
    the image and the last two fields match the single steps' bits, and the
    FP32 store of every step."""
    announce(f"two-step pass: checkpoints ckpt={ckpt} nt={nt}")
    n = (132, 52, 44)
    P, r = make_shot(dev, n=n, nt=nt)
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    d[P.rlocs < 0] = 0

    def run(store):
        def f():
            p0, p1 = dev.source_prop(1, nt, 1, store=store, ckpt=ckpt, get_last=True)
            dev.set_receivers(P.rlocs, d)
            dev.zero_image()
            dev.rtm_adjoint(1, nt, 1)
            return dev.get_image(), p0, p1
        return f

    ref = with_mode(dev, 0, run("ckpt"))
    assert np.abs(ref[0]).max() > 0
    for store in ("ckpt", "f32"):
        got = with_mode(dev, 1, run(store))
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("store,jimg", [("f16", 2), ("bf16", 3), ("f32", 3)])
def test_pass_with_strided_snapshots(dev, store, jimg):
    """Snapshots every jimg steps: the pass's first or second output is
    encoded (16-bit) or written in place (FP32) as the step's would be;
    Born modelling and the adjoint read them."""
    announce(f"two-step pass: snapshots store={store} jimg={jimg}")
    n = (64, 40, 36)
    nt = 75
    P, r = make_shot(dev, n=n, nt=nt, jimg=jimg)
    m = r.standard_normal((n[2], n[1], n[0]))
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    d[P.rlocs < 0] = 0

    def run():
        dev.source_prop(1, nt, jimg, store=store)
        dev.set_model(m)
        dev.rtm_forward(1, nt, jimg)
        D = dev.get_receivers()
        dev.set_receivers(P.rlocs, d)
        dev.zero_image()
        dev.rtm_adjoint(1, nt, jimg)
        return D, dev.get_image()

    one = with_mode(dev, 0, run)
    two = with_mode(dev, 1, run)
    assert np.abs(one[1]).max() > 0
    for a, b in zip(one, two):
        np.testing.assert_array_equal(a, b)


def test_pass_source_field_matches_the_oracle(dev):
    """Independently of the single steps: the pass's source field against
    the FP64 oracle (jt = 2, snapshots every 3 steps)."""
    announce("two-step pass: oracle parity")
    P, r = make_shot(dev, n=(132, 52, 44), jt=2, jimg=3)

    def run():
        p0, p1 = dev.source_prop(2, P.nt, 3, store="f32", get_last=True)
        return p0, p1, [dev.get_snapshot(k) for k in (0, P.nimg // 2, P.nimg - 1)]

    p0, p1, snaps = with_mode(dev, 1, run)
    ref, last2 = P.source_prop()
    assert rel(p1, last2[1]) < 1e-5 and rel(p0, last2[0]) < 1e-5
    for s, k in zip(snaps, (0, P.nimg // 2, P.nimg - 1)):
        assert rel(s, ref[k]) < 1e-5, k


def test_pass_timing_entry(dev):
    announce("two-step pass: timing entry")
    make_shot(dev, n=(132, 52, 44), nt=20)
    ms = dev.time_step(2, 0, 0, 3)
    assert ms > 0
