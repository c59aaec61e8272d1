"""Pins of the CPU oracle (oracle/) -- each against something other than the
oracle's own formula: closed forms, library routines, the reference's own code
built from its sources (oracle/_ref), adjoint identities or brute force."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import born_oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


def small_problem(n=(20, 22, 24), d=(10.0, 12.0, 14.0), bc_a=6, bc_b=0.02, seed=0, prec=64,
                  **kw):
    r = rng(seed)
    n1, n2, n3 = n
    V = r.uniform(0.05, 0.12, (n3, n2, n1)) * min(d) ** 2 / 4.0
    return O.Problem(n, d, V, bc_a, bc_b, prec=prec, **kw), r


# --- A1 coefficients -------------------------------------------------------
def test_coeffs_are_fornberg_weights():
    """Born/wave_fkernel.3d8o.cu:26-27 equals the textbook 8th-order second
    derivative weights {-205/72, 8/5, -1/5, 8/315, -1/560}/h^2 per axis."""
    d = (3.0, 5.0, 7.0)
    c = O.coeffs(*d)
    fw = [8 / 5, -1 / 5, 8 / 315, -1 / 560]
    assert np.isclose(c[0], sum(-205 / 72 / h ** 2 for h in d), rtol=1e-14)
    for k in range(4):
        for a in range(3):
            assert np.isclose(c[1 + 3 * k + a], fw[k] / d[a] ** 2, rtol=1e-14)


def test_1d_weights_exact_on_polynomials():
    """The 1-D weights differentiate x^p exactly for p <= 9 (order 8)."""
    w = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    x0, h = 0.3, 0.1
    xs = x0 + h * np.arange(-4, 5)
    for p in range(10):
        approx = (w * xs ** p).sum() / h ** 2
        exact = p * (p - 1) * x0 ** (p - 2) if p >= 2 else 0.0
        assert abs(approx - exact) < 1e-8 * max(1, abs(exact)), p
    approx = (w * xs ** 10).sum() / h ** 2
    assert abs(approx - 90 * x0 ** 8) > 1e-6  # degree 10 is not exact


def test_stencil_laplacian_of_polynomial_anisotropic():
    """With w = 1, V = 1, p_old = 0 the VL step returns L p + 2p; for a
    quartic in each axis L8 p is the exact Laplacian (catches axis swaps)."""
    n, d = (18, 19, 20), (2.0, 3.0, 5.0)
    P = O.Problem(n, d, np.ones((n[2], n[1], n[0])), 0.0, 0.0)
    i3, i2, i1 = np.meshgrid(*[np.arange(k, dtype=np.float64) for k in n[::-1]], indexing="ij")
    x, y, z = i1 * d[0], i2 * d[1], i3 * d[2]
    p = 0.5 * x ** 4 + 2 * y ** 3 - 0.25 * z ** 4 + x * y * z + 3 * y ** 2 * z
    lap = 6 * x ** 2 + 12 * y - 3 * z ** 2 + 6 * z
    pn = P.step_vl(p.ravel(), np.zeros(p.size)).reshape(p.shape)
    U = (slice(4, -4),) * 3
    np.testing.assert_allclose((pn - 2 * p)[U], lap[U], rtol=1e-9, atol=1e-6)
    # outside the update set nothing is written
    mask = np.ones(p.shape, bool)
    mask[U] = False
    assert np.all(pn[mask] == 0)


def test_stencil_symbol_on_plane_wave():
    """L8 cos(k.x) = sigma(k) cos(k.x) at interior points, sigma the discrete
    symbol sum_a (c0_a + 2 sum_k c_ak cos(k theta_a))."""
    n, d = (24, 24, 24), (10.0, 10.0, 10.0)
    P = O.Problem(n, d, np.ones((24, 24, 24)), 0.0, 0.0)
    th = (0.7, 1.3, 2.1)
    i3, i2, i1 = np.meshgrid(*[np.arange(24.0)] * 3, indexing="ij")
    p = np.cos(th[0] * i1 + th[1] * i2 + th[2] * i3)
    fw = [8 / 5, -1 / 5, 8 / 315, -1 / 560]
    sig = sum((-205 / 72 + 2 * sum(fw[k] * math.cos((k + 1) * t) for k in range(4))) / 100.0
              for t in th)
    pn = P.step_vl(p.ravel(), np.zeros(p.size)).reshape(p.shape)
    U = (slice(4, -4),) * 3
    np.testing.assert_allclose((pn - 2 * p)[U], sig * p[U], atol=1e-12)


# --- A4 sponge -------------------------------------------------------------
def test_sponge_table_closed_form():
    w = O.sponge_table(50.0, 0.0005)
    assert len(w) == 50
    for e in (0, 10, 49):
        assert np.isclose(w[e], math.exp(-0.0005 * (50 - e)), rtol=1e-15)
    assert np.all(np.diff(w) > 0) and np.all(w < 1)


def test_sponge_edge_distance_matches_reference_formula():
    """w on a ones field (V = 0, p_old = 0: p' = 2w).  Hand-evaluated cells of
    Born/wave_fkernel.3d8o.cu:194: e = min(i-4, n-4-i) per axis, so the low
    face's first updated cell has e = 0 and the high face's last one e = 1."""
    n = (20, 21, 22)
    a, b = 5.0, 0.1
    P = O.Problem(n, (1.0, 1.0, 1.0), np.zeros((22, 21, 20)), a, b)
    p = np.ones(20 * 21 * 22)
    w = P.step_vl(p, np.zeros_like(p)).reshape(22, 21, 20) / 2
    ex = lambda e: math.exp(-b * (a - e)) if e < a else 1.0
    assert np.isclose(w[11, 10, 4], ex(0))          # i1 = 4
    assert np.isclose(w[11, 10, 15], ex(1))         # i1 = n1-5 -> n1-4-i1 = 1
    assert np.isclose(w[11, 4, 10], ex(0))          # i2 = 4
    assert np.isclose(w[17, 10, 10], ex(1))         # i3 = n3-5
    assert np.isclose(w[6, 7, 8], ex(2))            # min(4, 3, 2) = 2
    assert np.isclose(w[11, 10, 10], 1.0)           # deep interior


# --- A2/A3: the LV step is the exact transpose of the VL step ---------------
def _step_matrix(P, form):
    n1, n2, n3 = P.n
    N = n1 * n2 * n3
    U = np.zeros((n3, n2, n1), bool)
    U[4:-4, 4:-4, 4:-4] = True
    idx = np.flatnonzero(U.ravel())
    A = np.zeros((len(idx), len(idx)))
    B = np.zeros((len(idx), len(idx)))
    for j, cj in enumerate(idx):
        e = np.zeros(N)
        e[cj] = 1.0
        f = P.step_vl if form == 0 else P.step_lv
        A[:, j] = f(e, np.zeros(N))[idx]
        B[:, j] = -f(np.zeros(N), e)[idx]
    return A, B


def test_lv_step_is_transpose_of_vl_step():
    P, _ = small_problem(n=(10, 11, 12), d=(1.0, 1.0, 1.0), bc_a=3.0, bc_b=0.2)
    P.V[:] = np.random.default_rng(3).uniform(0.01, 0.08, P.V.size)
    A_vl, B_vl = _step_matrix(P, 0)
    A_lv, B_lv = _step_matrix(P, 1)
    assert np.abs(A_vl - A_vl.T).max() > 1e-3  # VL alone is not symmetric
    np.testing.assert_allclose(A_lv, A_vl.T, atol=1e-14)
    np.testing.assert_allclose(B_lv, B_vl.T, atol=1e-15)
    assert np.allclose(np.diag(B_vl), np.diag(B_vl)) and np.all(np.diag(B_vl) <= 1)


def test_undamped_leapfrog_is_time_reversible():
    """Without sponge and sources p' = V L p + 2p - p_old is exactly
    reversible: stepping back from (p_N, p_{N-1}) recovers (p_0, p_1)."""
    P, r = small_problem(bc_a=0.0, bc_b=0.0)
    p0 = np.zeros(P.ncell)
    p1 = np.zeros(P.ncell)
    U = np.zeros(P.ncell, bool).reshape(24, 22, 20)
    U[4:-4, 4:-4, 4:-4] = True
    p1[U.ravel()] = r.standard_normal(U.sum())
    p0[U.ravel()] = r.standard_normal(U.sum())
    cur, old = P.propagate(0, 40, p1, p0)   # cur = p_41, old = p_40
    back_cur, back_old = P.propagate(0, 40, old, cur)
    np.testing.assert_allclose(back_cur, p0, atol=1e-9)
    np.testing.assert_allclose(back_old, p1, atol=1e-9)


def test_leapfrog_stability_limit_from_the_weights():
    """The 3-D 8th-order leapfrog limit.  The stencil's largest eigenvalue
    magnitude is its checkerboard (Nyquist) symbol
    lam = -(c0 + 2 sum_{k,a} (-1)^k c_{k,a}), and p' = 2p - p_old + V L p is
    stable while V lam <= 4.  For equal spacing that is v dt / d = 0.4529,
    the limit the library refuses V beyond.  The undamped oracle leapfrog
    started from a checkerboard stays bounded over 300 steps at 0.98 of the
    limit and blows up at 1.02 (unequal spacing)."""
    def lam_of(d):
        c = O.coeffs(*d)
        return -(c[0] + 2 * sum(c[1 + 3 * (k - 1) + a] * (-1) ** k
                                for k in range(1, 5) for a in range(3)))

    assert abs(math.sqrt(4 / lam_of((10.0, 10.0, 10.0))) / 10.0 - 0.45286) < 1e-5
    d = (10.0, 12.0, 14.0)
    lam = lam_of(d)
    n = (24, 26, 28)
    i = np.indices((n[2], n[1], n[0]))
    inside = np.zeros((n[2], n[1], n[0]))
    inside[4:-4, 4:-4, 4:-4] = 1
    p0 = ((-1.0) ** (i[0] + i[1] + i[2]) * inside).ravel()
    peak = {}
    for f in (0.98, 1.02):
        P = O.Problem(n, d, np.full((n[2], n[1], n[0]), f * 4 / lam), 0.0, 0.0)
        cur, _ = P.propagate(0, 300, p0.copy(), p0.copy())
        peak[f] = np.abs(cur).max()
    assert peak[0.98] < 50 and peak[1.02] > 1e6


def test_undamped_leapfrog_conserves_discrete_energy():
    """Without the sponge the VL step is p' = 2p - p_old + V L p with L
    symmetric on the update set, which conserves
    E = <V^-1 (p' - p), p' - p> - <L p', p> exactly (variable V, unequal
    spacing): a non-symmetric tap, damping, or V applied inside L breaks it."""
    r = rng(3)
    n, d = (20, 22, 24), (10.0, 12.0, 14.0)
    shape = (n[2], n[1], n[0])
    V = r.uniform(0.05, 0.12, shape) * min(d) ** 2 / 4.0
    P = O.Problem(n, d, V, 0.0, 0.0)
    P1 = O.Problem(n, d, np.ones(shape), 0.0, 0.0)
    inside = np.zeros(shape)
    inside[4:-4, 4:-4, 4:-4] = 1
    inside = inside.ravel()
    Vi = np.where(inside > 0, 1.0 / V.ravel(), 0.0)

    def lap(p):  # V = 1, no sponge: step(p, 0) = L p + 2p
        return P1.step_vl(p, np.zeros_like(p)) - 2 * p

    p_old = r.standard_normal(inside.size) * inside
    p = r.standard_normal(inside.size) * inside
    E = []
    for _ in range(100):
        pn = P.step_vl(p, p_old)
        E.append((pn - p) @ (Vi * (pn - p)) - lap(pn) @ p)
        p_old, p = p, pn
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 1e-12 * abs(E[0])


def test_sponge_dissipates_energy():
    P, r = small_problem(n=(32, 32, 32), bc_a=8.0, bc_b=0.05)
    p = np.zeros(P.ncell)
    p.reshape(32, 32, 32)[16, 16, 16] = 1.0
    cur, old = P.propagate(0, 300, p, np.zeros(P.ncell))
    P0, _ = small_problem(n=(32, 32, 32), bc_a=0.0, bc_b=0.0)
    c0, o0 = P0.propagate(0, 300, p, np.zeros(P.ncell))
    assert np.linalg.norm(cur) < 0.5 * np.linalg.norm(c0)


# --- A5-A7 injection / extraction ------------------------------------------
def test_inject_extract_adjoint_pair():
    r = rng(5)
    N, npt, ntr = 500, 7, 30
    locs = r.choice(N, npt, replace=False).astype(np.int32)
    s8 = r.standard_normal(8)
    p = r.standard_normal(N)
    tr = r.standard_normal((npt, ntr))
    L = O.lib(64)
    inj = np.zeros(N)
    L.orc_inject(npt, locs.ctypes.data, tr.ctypes.data, ntr, s8.ctypes.data, 5, 1.0,
                 inj.ctypes.data)
    ext = np.zeros((npt, ntr))
    L.orc_extract(npt, locs.ctypes.data, ext.ctypes.data, ntr, s8.ctypes.data, 5, p.ctypes.data)
    assert np.isclose(inj @ p, (ext * tr).sum(), rtol=1e-13)
    # closed form for one point
    assert np.isclose(inj[locs[0]], (s8 * tr[0, 5:13]).sum())


# --- Born forward / RTM adjoint ----------------------------------------------
def born_case(n=(24, 26, 28), nt=60, jt=1, jimg=1, seed=1, prec=64, bc_a=6.0, bc_b=0.03):
    r = rng(seed)
    n1, n2, n3 = n
    d = (10.0, 10.0, 10.0)
    r.uniform(1500, 3000, (n3, n2, n1))  # keeps the seeded stream of the cases below
    V = (r.uniform(1500, 3000, (n3, n2, n1)) * 0.001) ** 2
    ntr = (nt - 1) // jt + 9
    w = O.wavelet_derivative(np.sin(np.linspace(0, 6, ntr - 8)) * np.hanning(ntr - 8), 0.001)
    slocs, str_ = O.source_geometry(n1, n2, 9.3, n1 / 2 + 0.4, n2 / 2 - 0.2, w, ntr)
    xs, ys = np.meshgrid(np.arange(5, n1 - 5, 3), np.arange(5, n2 - 5, 4))
    rl = O.receiver_index(n1, n2, xs.ravel() + 0.5, ys.ravel(), np.full(xs.size, 6.2))
    sinc = O.sinc_table(jt) if jt > 1 else None
    P = O.Problem(n, d, V, bc_a, bc_b, src_locs=slocs, src_tr=str_, sinc_s=sinc,
                  rec_locs=rl.astype(np.int32), ntr_d=ntr, sinc_d=sinc, jt=jt, nt=nt, jimg=jimg,
                  cimg=1e-3, prec=prec)
    return P, r


@pytest.mark.parametrize("jt,jimg", [(1, 1), (2, 3)])
def test_born_dot_product_fp64(jt, jimg):
    """<F m, d> = <m, F^T d> to FP64 round-off (my_operator::dot_test,
    Born/my_operator.cpp:14-76, at 1e-12 instead of 1e-4)."""
    P, r = born_case(jt=jt, jimg=jimg)
    snaps, _ = P.source_prop()
    m = r.standard_normal(P.ncell)
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    Fm = P.born_forward(m, snaps)
    Ftd = P.rtm_adjoint(d, snaps)
    lhs, rhs = (Fm * d).sum(), (m * Ftd).sum()
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)
    assert abs(lhs) > 0


def test_composite_dot_product_two_steps_per_sample():
    """The multi-shot operator (rtm_zero_op's composition D_t W B D_z E mute)
    with two propagation steps per data sample is adjoint to its transpose,
    and its sinc tables are the operator's own (delta at phase 0, the
    half-sample interpolator at phase 1)."""
    P, r = born_case(jt=2, nt=41)
    n = P.n
    shots = [(P.slocs, P.strc, P.rlocs), (P.slocs + 2, P.strc, P.rlocs)]
    box = (n[0] - 10, n[1] - 10, n[2] - 12, 5, 5, 6)
    mute = np.ones((box[2], box[1], box[0]))
    mute[:2] = 0
    C = O.CompositeBorn(n, (10.0, 10.0, 10.0), P.V.reshape(n[2], n[1], n[0]), 6.0, 0.03, shots,
                        21, 0.002, jt=2, jimg=2, cimg=1e-3, model_box=box, mute=mute)
    s = C.probs[0].sinc_d
    assert abs(s[0, 3] - 1) < 1e-12 and np.allclose(s[1], s[1][::-1])
    m = r.standard_normal(box[0] * box[1] * box[2])
    d = r.standard_normal((2, len(P.rlocs), 21))
    Fm = C.forward(m)
    lhs, rhs = (Fm * d).sum(), (m * C.adjoint(d).ravel()).sum()
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)
    assert abs(lhs) > 0


def test_born_forward_live_source_equals_snapshots():
    P, r = born_case(jimg=2)
    snaps, _ = P.source_prop()
    m = r.standard_normal(P.ncell)
    np.testing.assert_array_equal(P.born_forward(m, snaps), P.born_forward(m, None))


def test_adjoint_checkpoint_equals_snapshots():
    P, r = born_case(nt=50, jimg=2)
    snaps, _ = P.source_prop()
    d = r.standard_normal((len(P.rlocs), P.ntr_d))
    a = P.rtm_adjoint(d, snaps)
    for ck in (1, 7, 50, 64):
        np.testing.assert_array_equal(P.rtm_adjoint(d, None, ckpt=ck), a)


def test_born_is_linear_in_model():
    P, r = born_case(nt=40)
    snaps, _ = P.source_prop()
    m1, m2 = r.standard_normal(P.ncell), r.standard_normal(P.ncell)
    np.testing.assert_allclose(P.born_forward(2 * m1 - 3 * m2, snaps),
                               2 * P.born_forward(m1, snaps) - 3 * P.born_forward(m2, snaps),
                               rtol=1e-10, atol=1e-12)


def test_fp32_build_agrees_with_fp64():
    P64, r = born_case(nt=40)
    P32, _ = born_case(nt=40, prec=32)
    s64, l64 = P64.source_prop()
    s32, l32 = P32.source_prop()
    rel = np.linalg.norm(l32 - l64) / np.linalg.norm(l64)
    assert rel < 1e-5, rel


def test_fp32_denormal_flush_is_invisible_and_restored():
    """The CPU baseline times the FP32 build with denormals flushed: the
    shot's image moves by less than 1e-6, and turning the flush off again
    restores gradual underflow on the calling thread."""
    P32, r = born_case(nt=80, prec=32)
    D = r.standard_normal((len(P32.rlocs), P32.ntr_d)).astype(np.float32)
    img0 = P32.rtm_adjoint(D, None, ckpt=16)
    tiny = np.array([2e-38], np.float32)  # normal; half of it is not
    P32.L.orc_flush_denormals(1)
    try:
        assert (tiny * np.float32(0.5))[0] == 0.0
        img1 = P32.rtm_adjoint(D, None, ckpt=16)
    finally:
        P32.L.orc_flush_denormals(0)
    assert (tiny * np.float32(0.5))[0] > 0.0
    assert np.linalg.norm(img1 - img0) <= 1e-6 * np.linalg.norm(img0)


# --- time recursion: closed forms with V = 0 and no sponge ------------------
# With V = 0 and w = 1 both step forms reduce to p' = 2p - p_old at every
# cell, so every cell runs its own scalar recursion from zero fields.  A
# forcing f(m) added after step m then gives p^n = sum_{m<n} (n-m) f(m), the
# double cumulative sum of f; the adjoint's backward recursion forced after
# each step gives lambda^n = sum_{m>=n-1} (m-n+2) h(m).  The forcings below
# follow the reference's own indexing, not the oracle's code:
#   * step it reads sample id = it/jt with sinc row ii = it%jt
#     (Born/gpu_funcs_3d.cu:228-229, 398-400, 598-601);
#   * the step runs first and the injection follows it (:247-250);
#   * source trace sample t+4 holds the wavelet sample t
#     (Born/source_func_3d.cpp:127), so with the zero-shift row (tap 3) the
#     source enters one sample late;
#   * Born modelling adds c s^{it+1} m to q^{it+1} when it % jt == 0 and
#     extracts the traces from that q^{it+1} at samples id+k (:423-462);
#   * the adjoint injects the data after the step and images the injected
#     lambda^{it+1} when ii == 0 (:626-667).
# The sinc rows for jt = 2 are the reference's own sinc_bob output
# (tests/golden/ref_outputs.json).  A one-step shift of any of these indices,
# inject-before-step, or a lost or extra step fails one of the tests.
def _ref_sinc(jt):
    with open(GOLDEN) as f:
        return np.array(json.load(f)["sinc"][str(jt)], np.float64)


def _zero_velocity_case(jt, nt, jimg, seed=11):
    """A 14^3 grid with V = 0, no sponge, a 9-point source around (6.3, 6.6)
    on plane 5, receivers on two source cells and one cell the source never
    reaches."""
    r = rng(seed)
    n = (14, 14, 14)
    nt_data = (nt - 1) // jt + 1
    ntr = nt_data + 8
    w = r.standard_normal(nt_data)
    slocs, str_ = O.source_geometry(n[0], n[1], 6.0, 6.3, 6.6, w, ntr)
    rl = np.array([slocs[4], slocs[0], 8 + 8 * 14 + 8 * 196], np.int32)
    sinc = _ref_sinc(jt) if jt > 1 else None
    P = O.Problem(n, (10.0, 10.0, 10.0), np.zeros((14, 14, 14)), 0.0, 0.0, src_locs=slocs,
                  src_tr=str_, sinc_s=sinc, rec_locs=rl, ntr_d=ntr, sinc_d=sinc, jt=jt, nt=nt,
                  jimg=jimg, cimg=0.37)
    scale = np.array([math.exp(-(6.0 - 5) ** 2 - (6.3 - i1) ** 2 - (6.6 - i2) ** 2)
                      for i2 in (5, 6, 7) for i1 in (5, 6, 7)])
    S = sinc if sinc is not None else np.eye(8)[3][None, :]
    return P, r, w, scale, S


def _forcing(tr, S, jt, nt):
    """The injected value after step it: sum_k S[it%jt][k] tr[it/jt + k]."""
    return np.array([S[it % jt] @ tr[it // jt:it // jt + 8] for it in range(nt)])


def _source_field(w, scale_j, S, jt, nt):
    """s^{it+1}, it = 0..nt-1, at one source point of weight scale_j: the
    double cumulative sum of the forcing built from trace[t+4] = w[t]."""
    tr = np.zeros(len(w) + 8)
    tr[4:4 + len(w)] = scale_j * w
    return np.cumsum(np.cumsum(_forcing(tr, S, jt, nt)))


def test_source_prop_one_sample_delay_jt1():
    """jt = 1: s^n = scale * sum_{t<=n-2} (n-1-t) w[t], so s^1 = 0 and
    s^2 = scale * w[0] at every source point (the +4 trace offset read
    through tap 3 of the zero-shift row)."""
    nt = 40
    P, _, w, scale, S = _zero_velocity_case(1, nt, 1)
    snaps, last2 = P.source_prop()
    cw = np.cumsum(np.cumsum(w))
    for j in (0, 4, 8):
        got = snaps[:, P.slocs[j]]            # snaps[it] = s^{it+1}
        want = np.concatenate([[0.0], scale[j] * cw[:nt - 1]])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
        assert got[0] == 0.0 and np.isclose(got[1], scale[j] * w[0], rtol=1e-14)
    # last two fields: s^{nt-1}, s^{nt}
    np.testing.assert_allclose(last2[:, P.slocs[4]], snaps[-2:, P.slocs[4]], rtol=0, atol=0)
    # nothing outside the nine source cells moves (V = 0)
    other = np.ones(P.ncell, bool)
    other[P.slocs] = False
    assert np.all(snaps[:, other] == 0)


@pytest.mark.parametrize("jt,jimg", [(2, 2), (2, 3)])
def test_source_prop_closed_form_two_steps_per_sample(jt, jimg):
    """jt = 2 with the reference's two-phase sinc rows: s^{it+1} is the double
    cumulative sum of f(it) = S[it%2] . trace[it/2 : it/2+8]; snapshots are
    s^{it+1} for it % jimg == 0."""
    nt = 2 * 25 + 1
    P, _, w, scale, S = _zero_velocity_case(jt, nt, jimg)
    snaps, last2 = P.source_prop()
    for j in (1, 4, 6):
        want = _source_field(w, scale[j], S, jt, nt)
        np.testing.assert_allclose(snaps[:, P.slocs[j]], want[::jimg], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(last2[:, P.slocs[j]], want[-2:], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("jt,jimg", [(1, 1), (2, 2), (2, 3)])
@pytest.mark.parametrize("live", [False, True])
def test_born_forward_closed_form(jt, jimg, live):
    """q^{it+1} = 2q^it - q^{it-1} + c s^{it+1} m [it % jimg == 0], then
    D[id+k] += S_d[ii][k] q^{it+1}: at a receiver on a source cell the trace
    is the sinc-spread double cumulative sum of c m s^{it+1}; a receiver the
    source never reaches records zero.  Holds with stored snapshots and with
    the source propagated alongside."""
    nt = 31 if jt == 1 else 2 * 15 + 1
    P, r, w, scale, S = _zero_velocity_case(jt, nt, jimg)
    m = r.standard_normal(P.ncell)
    snaps = None if live else P.source_prop()[0]
    D = P.born_forward(m, snaps)
    for rec, j in ((0, 4), (1, 0)):
        x = P.slocs[j]
        s = _source_field(w, scale[j], S, jt, nt)             # s^{it+1}
        g = np.where(np.arange(nt) % jimg == 0, P.cimg * m[x] * s, 0.0)
        q = np.cumsum(np.cumsum(g))                           # q^{it+1}
        want = np.zeros(P.ntr_d)
        for it in range(nt):
            want[it // jt:it // jt + 8] += S[it % jt] * q[it]
        np.testing.assert_allclose(D[rec], want, rtol=1e-11, atol=1e-14)
    assert np.all(D[2] == 0)


@pytest.mark.parametrize("jt,jimg", [(1, 1), (2, 2), (2, 3)])
def test_rtm_adjoint_closed_form(jt, jimg):
    """lambda^{it+1} = 2 lambda^{it+2} - lambda^{it+3} + sum_k S_d[ii][k]
    D[id+k] from zero fields at nt+1, nt+2, and img += c s^{it+1}
    lambda^{it+1} for it % jimg == 0: at a receiver on a source cell the image
    is c sum s^{it+1} lambda^{it+1}, with lambda the reversed double
    cumulative sum of the injected data.  The checkpointed adjoint agrees."""
    nt = 31 if jt == 1 else 2 * 15 + 1
    P, r, w, scale, S = _zero_velocity_case(jt, nt, jimg)
    D = r.standard_normal((3, P.ntr_d))
    snaps, _ = P.source_prop()
    img = P.rtm_adjoint(D, snaps)
    for rec, j in ((0, 4), (1, 0)):
        x = P.slocs[j]
        h = _forcing(D[rec], S, jt, nt)                       # injected after step it
        lam = np.cumsum(np.cumsum(h[::-1]))[::-1]             # lambda^{it+1}
        s = _source_field(w, scale[j], S, jt, nt)
        want = P.cimg * sum(s[it] * lam[it] for it in range(0, nt, jimg))
        assert np.isclose(img[x], want, rtol=1e-11), (img[x], want)
    # the receiver the source never reaches images nothing; nor do cells
    # without a receiver
    mask = np.ones(P.ncell, bool)
    mask[P.slocs[[4, 0]]] = False
    assert np.all(img[mask] == 0)
    np.testing.assert_array_equal(P.rtm_adjoint(D, None, ckpt=4), img)


# --- host-side operators ---------------------------------------------------
def test_dz_dt_window_are_adjoint_pairs():
    r = rng(7)
    m = r.standard_normal((9, 5, 6))
    d = r.standard_normal((9, 5, 6))
    assert np.isclose((O.dz_forward(m) * d).sum(), (m * O.dz_adjoint(d)).sum(), rtol=1e-13)
    tr = r.standard_normal((4, 3, 40))
    dd = r.standard_normal((4, 3, 40))
    assert np.isclose((O.dt_forward(tr, 0.004) * dd).sum(), (tr * O.dt_adjoint(dd, 0.004)).sum(),
                      rtol=1e-13)
    wm = r.standard_normal((4, 48))
    wd = r.standard_normal((4, 40))
    assert np.isclose((O.window_forward(wm, 40) * wd).sum(),
                      (wm * O.window_adjoint(wd, 48)).sum(), rtol=1e-13)


def test_dz_and_dt_closed_forms():
    z = np.arange(10.0)[:, None, None] * np.ones((10, 2, 3))
    dz = O.dz_forward(z)   # d[i] = m[i-1] - m[i] = -1 for 1 <= i <= n-2
    assert np.allclose(dz[1:-1], -1.0) and np.allclose(dz[0], 0.0) and np.allclose(dz[-1], 8.0)
    t = np.arange(20.0) * 0.5
    dt = O.dt_forward(t, 0.5)  # (m[t-1]-m[t])/dt = -1, first sample -m[0]/dt = 0
    assert np.allclose(dt[1:], -1.0) and dt[0] == 0.0


def test_window_hand_example():
    """Born/map_data_3d.cpp:44-76: data sample t of each trace is sample t+4
    of that trace's ntb-long device trace; the adjoint puts it back there
    and leaves the first 4 and the last ntb-4-nt samples zero."""
    m = np.arange(20.0).reshape(2, 10)  # two traces, ntb = 10
    np.testing.assert_array_equal(O.window_forward(m, 5), [[4, 5, 6, 7, 8],
                                                           [14, 15, 16, 17, 18]])
    np.testing.assert_array_equal(O.window_adjoint([[1.0, 2.0], [3.0, 4.0]], 7),
                                  [[0, 0, 0, 0, 1, 2, 0], [0, 0, 0, 0, 3, 4, 0]])


# --- A14 interpolation tables ----------------------------------------------
def test_levinson_matches_scipy_toeplitz():
    r = rng(2)
    col = np.array([4.0, 1.5, 0.7, 0.2, 0.1, 0.05, 0.02, 0.01])
    g = r.standard_normal(8)
    np.testing.assert_allclose(O.levinson(col, g), scipy.linalg.solve_toeplitz(col, g),
                               rtol=1e-10)


def test_sinc_zero_shift_is_identity_and_interpolates():
    t = O.sinc_table(4)
    np.testing.assert_allclose(t[0], np.eye(8)[3], atol=1e-9)
    # interpolates a low-frequency sinusoid at fractional shift d
    for i, dfr in enumerate((0.25, 0.5, 0.75)):
        x = np.arange(8) - 3.0
        f = 0.05
        y = np.sin(2 * np.pi * f * x)
        assert abs((t[i + 1] * y).sum() - np.sin(2 * np.pi * f * dfr)) < 2e-3


@pytest.mark.parametrize("ntab", [1, 2, 4, 5])
def test_sinc_matches_reference_sinc_bob(ntab):
    """Against Born/sinc_bob.cpp built from the reference sources: the FP64
    table agrees to 1e-4 (measured gap <= 4.2e-5 at ntab = 5; the reference
    sums its Toeplitz terms and runs its Levinson vectors in float)."""
    ref = O.ref_lib("ref_sinc")
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes as C
    out = np.zeros((ntab, 8), np.float32)
    ref.ref_sinc_table(ntab, 8, C.c_void_p(out.ctypes.data))
    np.testing.assert_allclose(O.sinc_table(ntab), out, atol=1e-4)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_outputs.json")


def test_oracle_matches_reference_fixture():
    """Against tests/golden/ref_outputs.json, outputs of the reference's own
    Born/sinc_bob.cpp and Born/cgstep.cpp on seeded inputs (written by
    tests/golden/make_ref_fixtures.py from oracle/_ref): runs where the
    reference sources are absent.  Sinc tables for 1-8 phases to 1e-4; cgstep
    models and residuals to float precision."""
    with open(GOLDEN) as f:
        g = json.load(f)
    assert set(g["sinc"]) == {"1", "2", "3", "4", "5", "8"}
    for ntab, t in g["sinc"].items():
        np.testing.assert_allclose(O.sinc_table(int(ntab)), np.array(t), atol=1e-4)
    assert len(g["cgstep"]) == 3
    for c in g["cgstep"]:
        A, d = np.array(c["A"], np.float64), np.array(c["d"], np.float64)
        m64, res64 = O.cgstep_solve(A, d, c["niter"])
        np.testing.assert_allclose(np.array(c["m"]), m64, rtol=2e-4, atol=2e-5)
        np.testing.assert_allclose(np.array(c["resid"]), res64, rtol=2e-4)


# --- A24/A25 CG --------------------------------------------------------------
def test_cgstep_converges_to_lstsq():
    r = rng(4)
    A = r.standard_normal((40, 12))
    d = r.standard_normal(40)
    m, res = O.cgstep_solve(A, d, 12)
    np.testing.assert_allclose(m, np.linalg.lstsq(A, d, rcond=None)[0], rtol=1e-7, atol=1e-9)
    assert all(res[i + 1] <= res[i] * (1 + 1e-12) for i in range(len(res) - 1))


def test_cgstep_matches_reference_cgstep():
    ref = O.ref_lib("ref_cgstep")
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes as C
    r = rng(6)
    A = r.standard_normal((30, 10)).astype(np.float32)
    d = r.standard_normal(30).astype(np.float32)
    m = np.zeros(10, np.float32)
    res = np.zeros(8)
    ref.ref_cgstep_solve.restype = C.c_int
    n = ref.ref_cgstep_solve(30, 10, C.c_void_p(A.ctypes.data), C.c_void_p(d.ctypes.data), 8,
                             C.c_void_p(m.ctypes.data), C.c_void_p(res.ctypes.data))
    assert n == 8
    m64, res64 = O.cgstep_solve(A.astype(np.float64), d.astype(np.float64), 8)
    np.testing.assert_allclose(m, m64, rtol=2e-4, atol=2e-5)
    np.testing.assert_allclose(res, res64, rtol=2e-4)


def test_fcg_history1_is_cgstep():
    r = rng(21)
    A = r.standard_normal((50, 15))
    d = r.standard_normal(50)
    m1, res1 = O.cgstep_solve(A, d, 10)
    m2, res2, _ = O.fcg_solve(A, d, 10, history=1)
    np.testing.assert_allclose(res2, res1, rtol=1e-10)
    np.testing.assert_allclose(m2, m1, rtol=1e-9, atol=1e-12)


def test_fcg_fixed_preconditioner_keeps_cg_iterates():
    """With a fixed SPD preconditioner the extra history coefficients vanish
    (the steps are already A-conjugate): any history gives the PCG iterates."""
    r = rng(22)
    A = r.standard_normal((60, 20)) * np.linspace(0.2, 3.0, 20)
    d = r.standard_normal(60)
    w = r.uniform(0.5, 2.0, 20)
    _, res1, _ = O.fcg_solve(A, d, 8, history=1, precond=lambda g, it: w * g)
    _, res4, _ = O.fcg_solve(A, d, 8, history=4, precond=lambda g, it: w * g)
    np.testing.assert_allclose(res4, res1, rtol=1e-8)


def test_fcg_full_history_minimises_over_all_directions():
    """With a preconditioner that changes every iteration and a history at
    least as long as the run, iterate k minimises |A m - d| over the span of
    every direction used so far (checked by lstsq on that span); truncated
    to one step (cgstep) it does no better."""
    r = rng(23)
    A = r.standard_normal((70, 25))
    d = r.standard_normal(70)
    P = [r.uniform(0.1, 3.0, 25) for _ in range(12)]
    pre = lambda g, it: P[it] * g  # noqa: E731
    _, res, dirs = O.fcg_solve(A, d, 12, history=12, precond=pre)
    for k in range(1, 13):
        Z = np.stack(dirs[:k], axis=1)
        c = np.linalg.lstsq(A @ Z, d, rcond=None)[0]
        best = np.sum((A @ Z @ c - d) ** 2)
        assert abs(res[k - 1] - best) <= 1e-9 * (d @ d), k
    _, res1, _ = O.fcg_solve(A, d, 12, history=1, precond=pre)
    assert res[-1] <= res1[-1] * (1 + 1e-9)
    # and with n directions it reaches the least-squares solution
    m, _, _ = O.fcg_solve(A, d, 25, history=25, precond=lambda g, it: P[it % 12] * g)
    np.testing.assert_allclose(m, np.linalg.lstsq(A, d, rcond=None)[0], rtol=1e-6, atol=1e-9)


# --- solver forms (simple / reg / prec with weight and starting model) ---------
def solver_case(seed=21, nd=24, nm=10):
    r = rng(seed)
    L = r.standard_normal((nd, nm))
    d = r.standard_normal(nd)
    W = np.diag(r.uniform(0.5, 2.0, nd))
    R = np.eye(nm) - np.eye(nm, k=1)          # first difference (nm x nm)
    P = np.diag(r.uniform(0.2, 3.0, nm))
    m0 = r.standard_normal(nm)
    return L, d, W, R, P, m0


def test_lin_solve_forms_reach_their_closed_forms():
    """After as many iterations as unknowns (FP64, exact CG termination up to
    round-off) each form equals its normal-equation solution; a dropped eps,
    a transposed weight or the model in place of the update fails here."""
    L, d, W, R, P, m0 = solver_case()
    eps = 0.7
    r = d - L @ m0
    WL = W @ L
    m, _ = O.lin_solve("simple", L, d, 30, W=W, m0=m0)
    np.testing.assert_allclose(m, np.linalg.lstsq(WL, W @ r, rcond=None)[0], rtol=1e-8, atol=1e-10)
    m, _ = O.lin_solve("reg", L, d, 30, aux=R, eps=eps, W=W, m0=m0)
    closed = np.linalg.solve(WL.T @ WL + eps ** 2 * R.T @ R, WL.T @ (W @ r))
    np.testing.assert_allclose(m, closed, rtol=1e-8, atol=1e-10)
    m, _ = O.lin_solve("prec", L, d, 30, aux=P, eps=eps, W=W)
    WLP = WL @ P
    p = np.linalg.solve(WLP.T @ WLP + eps ** 2 * np.eye(10), WLP.T @ (W @ d))
    np.testing.assert_allclose(m, P @ p, rtol=1e-8, atol=1e-10)
    # a banded weight is not its own transpose (a diagonal one is): applied
    # transposed it gives another weighted solution
    Wb = W + 0.4 * np.eye(W.shape[0], k=1)
    m, _ = O.lin_solve("simple", L, d, 30, W=Wb, m0=m0)
    np.testing.assert_allclose(m, np.linalg.lstsq(Wb @ L, Wb @ r, rcond=None)[0], rtol=1e-8,
                               atol=1e-10)
    # eps = 0 and identity weight reduce to plain cgstep
    m1, res1 = O.lin_solve("reg", L, d, 6, aux=R, eps=0.0)
    m2, res2 = O.cgstep_solve(L, d, 6)
    np.testing.assert_allclose(m1, m2, rtol=1e-10, atol=1e-12)


def test_lin_solve_accepts_operator_pairs():
    L, d, W, R, P, m0 = solver_case(seed=5)
    a, _ = O.lin_solve("reg", L, d, 4, aux=R, eps=0.3, W=W, m0=m0)
    pair = lambda M: ((lambda x: M @ x), (lambda y: M.T @ y))  # noqa: E731
    b, _ = O.lin_solve("reg", pair(L), d, 4, aux=pair(R), eps=0.3, W=pair(W), m0=m0)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)


def test_tmute_hand_example():
    """t0 = 0.1 s, vmute = 2000 m/s, receivers at x = 100 + 50 i m, 4 ms
    samples from ot = 0: mute times 0.15, 0.175, 0.2 s -> samples 37, 43, 50
    (truncated in FP32: 37.5 -> 37, 43.75 -> 43, 0.2/0.004 -> 50); a mute
    time before ot starts at 0, one past the trace keeps nothing."""
    s = O.tmute_start(100, 0.0, 0.004, 3, 100.0, 50.0, 0.1, 2000.0)
    assert list(s) == [37, 43, 50]
    # the reference's float arithmetic (Born/tmute.cpp:63-64, float axis o/d,
    # t_0 and vmute): 30 m / 1500 m/s = 0.02 s is sample 5 exactly, but in
    # FP32 0.02f / 0.004f = 4.9999995 truncates to 4 (FP64 would give 5)
    assert list(O.tmute_start(100, 0.0, 0.004, 4, 0.0, 10.0, 0.0, 1500.0)) == [0, 1, 3, 4]
    assert list(O.tmute_start(10, 0.5, 0.004, 2, 0.0, 10.0, 0.1, 2000.0)) == [0, 0]
    assert list(O.tmute_start(10, 0.0, 0.004, 1, 0.0, 10.0, 5.0, 2000.0)) == [10]
    x = np.arange(2 * 3 * 100, dtype=np.float64).reshape(2, 3, 100)
    y = O.tmute_apply(x, s)
    assert (y[:, 0, :37] == 0).all() and (y[:, 0, 37:] == x[:, 0, 37:]).all()
    # a 0/1 mask: self-adjoint and idempotent
    z = rng(2).standard_normal(x.shape)
    assert np.isclose((y * z).sum(), (x * O.tmute_apply(z, s)).sum())
    np.testing.assert_array_equal(O.tmute_apply(y, s), y)


# --- geometry ------------------------------------------------------------------
def test_source_geometry_hand_example():
    locs, tr = O.source_geometry(10, 12, 5.0, 4.0, 6.0, np.array([1.0, 2.0]), 8)
    # plane iz-1 = 4, columns 3..5 x rows 5..7, centre weight exp(-1) (zn = 1)
    assert locs[4] == 4 + 6 * 10 + 4 * 120
    assert locs[0] == 3 + 5 * 10 + 4 * 120
    assert np.isclose(tr[4, 4], math.exp(-1.0)) and np.isclose(tr[4, 5], 2 * math.exp(-1.0))
    assert np.isclose(tr[0, 4], math.exp(-3.0))
    assert np.all(tr[:, :4] == 0)


def test_wavelet_derivative_hand_example():
    """Born/source_func_3d.cpp:9-13: w'[t] = (w[t+1] - w[t]) / dt, and the
    last sample 0 (a forward difference: the first output sees w[1])."""
    np.testing.assert_array_equal(O.wavelet_derivative([1.0, 3.0, 6.0, 10.0], 0.5),
                                  [4.0, 6.0, 8.0, 0.0])


def test_receiver_index_hand_example():
    assert O.receiver_index(10, 12, [2.7], [3.1], [5.9])[0] == 2 + 3 * 10 + (5 + 4) * 120


def test_inactive_receivers_hand_example():
    """A receiver outside the update set [4, n-4) on any axis is marked -1.
    The first and last updated cells of each axis are kept and their outer
    neighbours dropped; the axes have different lengths, so a swapped axis
    or a wrong stride fails."""
    n = (11, 13, 15)  # update sets [4, 7), [4, 9), [4, 11)
    centre = (5, 6, 7)

    def loc(i):
        return i[0] + n[0] * (i[1] + n[1] * i[2])

    keep = {centre: True, (4, 4, 4): True, (6, 8, 10): True}
    for ax in range(3):
        for i, kept in ((3, False), (4, True), (n[ax] - 5, True), (n[ax] - 4, False)):
            c = list(centre)
            c[ax] = i
            keep[tuple(c)] = kept
    locs = [loc(k) for k in keep]
    n1, n2, n3 = n
    got = O.mask_inactive(locs, (n1, n2, n3))
    assert got.dtype == np.int32
    for (k, kept), l, g in zip(keep.items(), locs, got):
        assert g == (l if kept else -1), k


@pytest.mark.parametrize("n", [0, 1, 1000, 100003])
def test_dot_accumulates_products_in_fp64(n):
    """orc_dot (Born/hypercube_float.cpp:123-132) adds the products in FP64 in
    both builds. A product of two FP32 values is exact in FP64, so the FP32
    build returns the FP64 build's bits on the same values, and both are
    within n ulp-sized steps of the exact sum; a product or a sum rounded to
    FP32 would miss both."""
    r = rng(31)
    a = r.standard_normal(n).astype(np.float32)
    b = r.standard_normal(n).astype(np.float32)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    d32 = O.lib(32).orc_dot(n, a.ctypes.data, b.ctypes.data)
    d64 = O.lib(64).orc_dot(n, a64.ctypes.data, b64.ctypes.data)
    prods = a64 * b64
    exact = math.fsum(prods.tolist())
    assert d32 == d64
    assert abs(d64 - exact) <= n * 2.0 ** -53 * float(np.abs(prods).sum())
    if n > 1:
        # the same sum accumulated in FP32 is far outside that bound
        s = np.float32(0)
        for p in prods[:1000].astype(np.float32):
            s = np.float32(s + p)
        part = math.fsum(prods[:1000].tolist())
        assert abs(float(s) - part) > min(n, 1000) * 2.0 ** -53 * float(np.abs(prods[:1000]).sum())


# --- FP32 over long runs: the error does not accumulate, and the summed ----
# --- (Henrici) form of the update does not reduce it ------------------------
def _fp32_runs(n=36, nt=1000, jimg=100, bc_a=8.0, bc_b=0.02):
    """One point source in a heterogeneous 36^3 box with a sponge, advanced
    nt steps three ways: the FP64 oracle, and in numpy FP32 both the plain
    VL update p' = w(V L p + 2p) - w^2 p_old (the GPU's, 16 B per cell) and
    its summed form (Henrici), which carries d = p - p_old:
    d' = w V L p - (1-w)^2 p + w^2 d, p' = p + d' (20 B per cell).
    Returns the two FP32 errors at every jimg-th step relative to the FP64
    field's largest norm, and that field's norm at the same steps."""
    d = (10.0, 10.0, 10.0)
    r = rng(9)
    V = ((2000.0 * 0.001) ** 2 * (1 + 0.2 * r.random((n, n, n)))).astype(np.float32)
    t = np.arange(nt + 8) * 0.001
    a = (np.pi * 25.0 * (t - 0.06)) ** 2
    tr = np.zeros((1, nt + 8))
    tr[0, 4:] = ((1 - 2 * a) * np.exp(-a))[:nt + 4]
    c3 = n // 2
    loc = np.array([c3 + c3 * n + c3 * n * n], np.int32)
    P = O.Problem((n, n, n), d, V.astype(np.float64), bc_a, bc_b, src_locs=loc, src_tr=tr,
                  nt=nt, jt=1, jimg=jimg)
    s64, _ = P.source_prop()
    c = O.coeffs(*d).astype(np.float32)
    wt = O.sponge_table(bc_a, bc_b)
    i = np.arange(n)
    e1 = np.minimum(i - 4, n - 4 - i)
    E = np.minimum(np.minimum(e1[:, None, None], e1[None, :, None]), e1[None, None, :])
    Wf = np.where((E >= 0) & (E < len(wt)), wt[np.clip(E, 0, len(wt) - 1)], 1.0)
    U = (slice(4, -4),) * 3
    Vu, Wu = V[U], Wf.astype(np.float32)[U]

    def lap(p):
        acc = c[0] * p[U]
        for k in range(1, 5):
            acc = acc + c[1 + 3 * (k - 1)] * (p[4:-4, 4:-4, 4 + k:n - 4 + k] + p[4:-4, 4:-4, 4 - k:n - 4 - k])
            acc = acc + c[2 + 3 * (k - 1)] * (p[4:-4, 4 + k:n - 4 + k, 4:-4] + p[4:-4, 4 - k:n - 4 - k, 4:-4])
            acc = acc + c[3 + 3 * (k - 1)] * (p[4 + k:n - 4 + k, 4:-4, 4:-4] + p[4 - k:n - 4 - k, 4:-4, 4:-4])
        return acc

    z = np.zeros((n, n, n), np.float32)
    p, po, q, dl = z.copy(), z.copy(), z.copy(), z.copy()
    ip = np.unravel_index(loc[0], (n, n, n))
    e_plain, e_sum = [], []
    one, two = np.float32(1), np.float32(2)
    for it in range(nt):
        f = np.float32(tr[0, it + 3])
        pn = z.copy()
        pn[U] = Wu * (Vu * lap(p) + two * p[U]) - (Wu * Wu) * po[U]
        pn[ip] += f
        po, p = p, pn
        dn = z.copy()
        dn[U] = Wu * Vu * lap(q) - (one - Wu) ** 2 * q[U] + (Wu * Wu) * dl[U]
        dn[ip] += f
        q, dl = q + dn, dn
        if it % jimg == 0:
            ref = s64[it // jimg].reshape(n, n, n)
            e_plain.append(np.linalg.norm(p - ref))
            e_sum.append(np.linalg.norm(q - ref))
    norms = np.array([np.linalg.norm(s) for s in s64])
    peak = norms.max()
    return np.array(e_plain) / peak, np.array(e_sum) / peak, norms / peak


def test_fp32_error_does_not_accumulate_and_summing_does_not_help():
    """The FP32 error of the leapfrog, measured against the FP64 oracle and
    relative to the field's largest norm, peaks while the wavefield is
    energetic (a few 1e-6) and then falls as the field leaves through the
    sponge: it does not grow with the number of steps.  Its ratio to the
    decaying current norm does grow, which is what a per-step relative error
    shows.  The summed form, whose point is to stop rounding from being
    integrated twice, leaves the error where it was (within 30 %), so the
    GPU keeps the plain 16 B/cell update."""
    ep, es, norms = _fp32_runs()
    assert ep.max() < 1e-5 and es.max() < 1e-5
    k = int(np.argmax(ep))
    assert norms[k] > 1e-2                      # the peak error is in the energetic phase
    assert ep[-1] < 0.1 * ep.max() and es[-1] < 0.1 * es.max()
    assert norms[-1] < 1e-3                      # the field itself has decayed by 1000x
    assert ep[-1] / norms[-1] > 10 * ep[1] / norms[1]   # relative to itself, it grows
    assert 0.7 * ep.max() < es.max() < 1.3 * ep.max()
