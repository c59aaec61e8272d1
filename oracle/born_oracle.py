"""CPU oracle for the Born acoustic RTM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker.  The product
(``sep_code_amd``) never imports it.

It wraps the C restatement in ``born_oracle.c`` (FP64 ``liborc64.so``, FP32
OpenMP ``liborc32.so``) and restates, in numpy FP64, the small host-side
operators that surround the propagation in the reference:

* ``dz_forward/dz_adjoint``   -- Born/laplac_3d.cpp:46-65 / :4-22 (a first
  difference along depth despite its name);
* ``dt_forward/dt_adjoint``   -- Born/deriv_3d.cpp:25-46 / :4-23, restated
  within each trace (the reference's duplicated first-sample term and its
  write into the next trace are not reproduced);
* ``window_forward/adjoint``  -- Born/map_data_3d.cpp:64-76 / :44-56;
* ``sinc_table``              -- Born/sinc_bob.cpp:4-114 (Hale's
  least-squares tapered sinc, Levinson solve) in FP64;
* ``cgstep``                  -- Born/cgstep.cpp:14-58;
* ``fcg_solve``               -- flexible CG with truncated history (cgstep's
  plane search at history 1);
* ``lin_solve``               -- simple_solver / reg_solver / prec_solver,
  Born/lin_solver.cpp:7-53,79-155 (stacked systems of Born/combo_oper.cpp);
* ``tmute_start/tmute_apply`` -- Born/tmute.cpp:4-76;
* ``source_geometry``, ``wavelet_derivative`` -- Born/source_func_3d.cpp:
  100-133 and 9-13;
* ``receiver_index``, ``mask_inactive`` -- Born/map_data_3d.cpp:4-32, and
  receivers outside the update set marked inactive;
* ``CompositeBorn``           -- the multi-shot operator D_t W B D_z E of
  Born/rtm_zero_op_3d.cpp:87-293 over the C propagation.

Every function here is pinned by tests/test_oracle.py against a closed form,
a hand example, a library routine (scipy Toeplitz solve, numpy lstsq), the
reference's own code built from its sources into oracle/_ref (sinc_bob,
cgstep), or an adjoint identity; tests/test_oracle_planted.py checks that
plausible mistakes planted in copies of the oracle fail those pins.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrcProb(C.Structure):
    """Mirror of ``orc_prob`` in born_oracle.c."""

    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("nw", C.c_int),
        ("c", C.c_void_p), ("wtab", C.c_void_p), ("V", C.c_void_p),
        ("nsrc", C.c_int), ("slocs", C.c_void_p), ("strc", C.c_void_p),
        ("ntr_s", C.c_int), ("sinc_s", C.c_void_p),
        ("nrec", C.c_int), ("rlocs", C.c_void_p), ("ntr_d", C.c_int),
        ("sinc_d", C.c_void_p),
        ("jt", C.c_int), ("nt", C.c_int), ("jimg", C.c_int), ("cimg", C.c_double),
    ]


_LIBS: dict = {}


def build() -> None:
    """Compile the oracle libraries (and oracle/_ref when the reference is
    present in this container)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/Born"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib(prec: int = 64) -> C.CDLL:
    """Load liborc64.so (prec=64) or liborc32.so (prec=32)."""
    if prec in _LIBS:
        return _LIBS[prec]
    path = os.path.join(HERE, f"liborc{prec}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    sfx = "_f64" if prec == 64 else "_f32"
    for name in ("orc_step_vl", "orc_step_lv", "orc_source_prop", "orc_born_forward",
                 "orc_rtm_adjoint", "orc_propagate", "orc_coeffs", "orc_sponge_table",
                 "orc_inject", "orc_extract", "orc_dot", "orc_flush_denormals"):
        setattr(L, name, getattr(L, name + sfx))
    P, I, R = C.c_void_p, C.c_int, (C.c_double if prec == 64 else C.c_float)
    argtypes = {
        "orc_step_vl": [P, P, P, P],
        "orc_step_lv": [P, P, P, P, P],
        "orc_source_prop": [P, P, P],
        "orc_born_forward": [P, P, P, P],
        "orc_rtm_adjoint": [P, P, P, I, P],
        "orc_propagate": [P, I, I, P, P],
        "orc_coeffs": [C.c_double, C.c_double, C.c_double, P],
        "orc_sponge_table": [C.c_double, C.c_double, I, P],
        "orc_inject": [I, P, P, I, P, I, R, P],
        "orc_extract": [I, P, P, I, P, I, P],
        "orc_dot": [C.c_long, P, P],
    }
    for name, at in argtypes.items():
        getattr(L, name).argtypes = at
        getattr(L, name).restype = C.c_int
    L.orc_dot.restype = C.c_double
    L.orc_flush_denormals.argtypes = [I]
    L.orc_flush_denormals.restype = None
    L.orc_coeffs.restype = None
    _LIBS[prec] = L
    return L


def ref_lib(name: str):
    """oracle/_ref/lib<name>.so built from the reference's own sources, or
    None when it has not been built (the reference is absent)."""
    path = os.path.join(HERE, "_ref", f"lib{name}.so")
    return C.CDLL(path) if os.path.exists(path) else None


def _ptr(a):
    return None if a is None else a.ctypes.data


# --------------------------------------------------------------------------
# Problem set-up (mirrors of the product's host-side set-up, restated here so
# the oracle does not depend on the product).
# --------------------------------------------------------------------------
def coeffs(d1, d2, d3, prec=64):
    """A1: 13 stencil weights (Born/wave_fkernel.3d8o.cu:26-27)."""
    dt = np.float64 if prec == 64 else np.float32
    c = np.zeros(13, dt)
    lib(prec).orc_coeffs(d1, d2, d3, _ptr(c))
    return c


def sponge_table(a, b, prec=64):
    """A4: w(e) = exp(-b (a-e)) for e < ceil(a) (Born/wave_fkernel.3d8o.cu:201)."""
    dt = np.float64 if prec == 64 else np.float32
    w = np.zeros(max(1, int(math.ceil(a))) + 1, dt)
    nw = lib(prec).orc_sponge_table(a, b, len(w), _ptr(w))
    return w[:nw].copy()


class Problem:
    """Holds the arrays of one shot for the C oracle and builds OrcProb."""

    def __init__(self, n, d, V, bc_a, bc_b, src_locs=None, src_tr=None, sinc_s=None,
                 rec_locs=None, ntr_d=8, sinc_d=None, jt=1, nt=1, jimg=1, cimg=1e-6,
                 prec=64):
        self.prec = prec
        self.dt = np.float64 if prec == 64 else np.float32
        dt = self.dt
        self.n = tuple(int(x) for x in n)  # (n1, n2, n3)
        self.c = coeffs(*d, prec=prec)
        self.wtab = sponge_table(bc_a, bc_b, prec=prec)
        if len(self.wtab) == 0:
            self.wtab = np.ones(1, dt)
            self.nw = 0
        else:
            self.nw = len(self.wtab)
        self.V = np.ascontiguousarray(V, dtype=dt).reshape(-1)
        e8 = np.zeros((jt, 8), dt)
        e8[:, 3] = 1
        self.slocs = np.ascontiguousarray(src_locs if src_locs is not None else np.zeros(0), np.int32)
        self.strc = np.ascontiguousarray(src_tr if src_tr is not None else np.zeros((0, 8)), dt)
        self.sinc_s = np.ascontiguousarray(sinc_s if sinc_s is not None else e8, dt)
        self.rlocs = np.ascontiguousarray(rec_locs if rec_locs is not None else np.zeros(0), np.int32)
        self.sinc_d = np.ascontiguousarray(sinc_d if sinc_d is not None else e8, dt)
        self.ntr_s = self.strc.shape[-1] if self.strc.size else 8
        self.ntr_d = int(ntr_d)
        self.jt, self.nt, self.jimg, self.cimg = int(jt), int(nt), int(jimg), float(cimg)
        self.s = OrcProb(self.n[0], self.n[1], self.n[2], self.nw,
                         _ptr(self.c), _ptr(self.wtab), _ptr(self.V),
                         len(self.slocs), _ptr(self.slocs), _ptr(self.strc), self.ntr_s,
                         _ptr(self.sinc_s), len(self.rlocs), _ptr(self.rlocs), self.ntr_d,
                         _ptr(self.sinc_d), self.jt, self.nt, self.jimg, self.cimg)
        self.L = lib(prec)

    @property
    def ncell(self):
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def nimg(self):
        return (self.nt - 1) // self.jimg + 1

    def arr(self, *shape):
        return np.zeros(shape, self.dt)

    def step_vl(self, p, po):
        pn = np.array(po, self.dt, copy=True)
        self.L.orc_step_vl(C.byref(self.s), _ptr(np.ascontiguousarray(p, self.dt)), _ptr(pn), _ptr(pn))
        return pn

    def step_lv(self, p, po):
        pn = np.array(po, self.dt, copy=True)
        u = self.arr(self.ncell)
        self.L.orc_step_lv(C.byref(self.s), _ptr(np.ascontiguousarray(p, self.dt)), _ptr(pn), _ptr(pn), _ptr(u))
        return pn

    def propagate(self, form, nsteps, p_cur, p_old):
        a = np.array(p_cur, self.dt, copy=True).reshape(-1)
        b = np.array(p_old, self.dt, copy=True).reshape(-1)
        rc = self.L.orc_propagate(C.byref(self.s), int(form), int(nsteps), _ptr(a), _ptr(b))
        assert rc == 0
        return a, b

    def source_prop(self, want_snaps=True):
        snaps = self.arr(self.nimg, self.ncell) if want_snaps else None
        last2 = self.arr(2, self.ncell)
        rc = self.L.orc_source_prop(C.byref(self.s), _ptr(snaps), _ptr(last2))
        assert rc == 0
        return snaps, last2

    def born_forward(self, mt, snaps=None):
        D = self.arr(len(self.rlocs), self.ntr_d)
        mt = np.ascontiguousarray(mt, self.dt).reshape(-1)
        snaps = None if snaps is None else np.ascontiguousarray(snaps, self.dt)
        rc = self.L.orc_born_forward(C.byref(self.s), _ptr(mt), _ptr(snaps), _ptr(D))
        assert rc == 0
        return D

    def rtm_adjoint(self, D, snaps=None, ckpt=0):
        img = self.arr(self.ncell)
        D = np.ascontiguousarray(D, self.dt)
        snaps = None if snaps is None else np.ascontiguousarray(snaps, self.dt)
        rc = self.L.orc_rtm_adjoint(C.byref(self.s), _ptr(D), _ptr(snaps), int(ckpt), _ptr(img))
        assert rc == 0
        return img


# --------------------------------------------------------------------------
# Host-side linear operators around the propagation (FP64 numpy).
# Volumes are numpy arrays of shape (n3, n2, n1); traces (..., nt).
# --------------------------------------------------------------------------
def dz_forward(m):
    """Born/laplac_3d.cpp:46-65: d[i3+1] += m[i3]; d[i3] -= m[i3] for i3 < n3-1."""
    m = np.asarray(m, np.float64)
    d = np.zeros_like(m)
    d[1:] += m[:-1]
    d[:-1] -= m[:-1]
    return d


def dz_adjoint(d):
    """Born/laplac_3d.cpp:4-22: m[i3] += d[i3+1] - d[i3] for i3 < n3-1."""
    d = np.asarray(d, np.float64)
    m = np.zeros_like(d)
    m[:-1] = d[1:] - d[:-1]
    return m


def dt_forward(m, dt):
    """Born/deriv_3d.cpp:25-46 within each trace: d[t] = (m[t-1] - m[t]) / dt,
    m[-1] = 0 (the reference's sign convention)."""
    m = np.asarray(m, np.float64)
    d = -m / dt
    d[..., 1:] += m[..., :-1] / dt
    return d


def dt_adjoint(d, dt):
    """Born/deriv_3d.cpp:4-23 within each trace: m[t] = (d[t+1] - d[t]) / dt,
    d[n] = 0."""
    d = np.asarray(d, np.float64)
    m = -d / dt
    m[..., :-1] += d[..., 1:] / dt
    return m


def window_forward(m, nt):
    """Born/map_data_3d.cpp:64-76: d[t] = m[t+4], t < nt."""
    return np.array(m[..., 4:4 + nt], np.float64)


def window_adjoint(d, ntb):
    """Born/map_data_3d.cpp:44-56: m[t+4] += d[t]."""
    d = np.asarray(d, np.float64)
    m = np.zeros(d.shape[:-1] + (ntb,), np.float64)
    m[..., 4:4 + d.shape[-1]] = d
    return m


def sinc_table(ntab, nsinc=8):
    """Born/sinc_bob.cpp:4-114 in FP64: table[i] = mksinc(d = i/ntab).

    The Toeplitz right-hand side sums cos over s = j*ds, j = 0..2*lsinc-1
    (the reference accumulates s in float, which makes its trip count
    rounding-dependent; here the intended 2*lsinc terms are used)."""
    out = np.zeros((ntab, nsinc), np.float64)
    pi2 = 2 * 3.141592654
    snat = 0.5 * (0.066 + 0.265 * math.log(nsinc))
    ds = snat / (nsinc * 2 - 1)
    svals = ds * np.arange(2 * nsinc)
    for it in range(ntab):
        d = it / ntab
        eta = nsinc // 2 - 1.0 + d
        j = np.arange(nsinc)
        b = np.cos(pi2 * np.outer(j, svals)).sum(1)
        c = np.cos(pi2 * np.outer(eta - j, svals)).sum(1)
        out[it] = levinson(b, c)
    return out


def levinson(r, g):
    """Born/sinc_bob.cpp:82-114 (toep): solve the symmetric Toeplitz system
    with first column r for right-hand side g, by Levinson recursion."""
    m = len(r)
    a = np.zeros(m)
    f = np.zeros(m)
    a[0] = 1.0
    v = r[0]
    f[0] = g[0] / r[0]
    for j in range(1, m):
        e = sum(a[i] * r[j - i] for i in range(j))
        c = e / v
        v -= e * c
        jh = j // 2
        for i in range(jh + 1):
            bot = a[j - i] - c * a[i]
            a[i] -= c * a[j - i]
            a[j - i] = bot
        w = sum(f[i] * r[j - i] for i in range(j))
        c = (g[j] - w) / v
        for i in range(j + 1):
            f[i] += c * a[j - i]
    return f


def cgstep_solve(A, d, niter):
    """Born/cgstep.cpp:14-58 driven as Born/lin_solver.cpp:115-143 with m0 = 0
    (rr = -d, Born/lin_solver.cpp:95-98), in FP64 with dense A.
    Returns (m, [rr.rr after each iteration])."""
    A = np.asarray(A, np.float64)
    m = np.zeros(A.shape[1])
    rr = -np.asarray(d, np.float64).copy()
    s = np.zeros_like(m)
    ss = np.zeros_like(rr)
    res = []
    for it in range(niter):
        g = A.T @ rr
        gg = A @ g
        if it == 0:
            ggg = gg @ gg
            if abs(ggg) < 1e-20:
                break
            alfa, beta = -(gg @ rr) / ggg, 0.0
        else:
            ggg, sss, ggs = gg @ gg, ss @ ss, gg @ ss
            if abs(ggg) < 1e-20 or abs(sss) < 1e-20:
                break
            det = ggg * sss * (1.0 - ggs / ggg * ggs / sss)
            ggr, ssr = -(gg @ rr), -(ss @ rr)
            alfa = (sss * ggr - ggs * ssr) / det
            beta = (-ggs * ggr + ggg * ssr) / det
        s = beta * s + alfa * g
        ss = beta * ss + alfa * gg
        m += s
        rr += ss
        res.append(rr @ rr)
    return m, res


def fcg_solve(A, d, niter, history, precond=None):
    """Flexible CG with truncated history on min |A m - d| from m = 0, in
    FP64 with dense A (or a pair of callables (forward, adjoint)).

    Each iteration takes the gradient g = A^T rr, the direction z = P(g, it)
    (g itself when P is None or when z.g <= 0), and minimises
    |rr + a A z + sum_i b_i S_i| over a and the coefficients of the last
    `history` steps (S_i their data-space images).  history = 1 is cgstep's
    plane search (Born/cgstep.cpp:28-53).  Returns (m, [rr.rr after each
    iteration], [directions used])."""
    if callable(A) or isinstance(A, tuple):
        fwd, adj = A
    else:
        Ad = np.asarray(A, np.float64)
        fwd, adj = (lambda x: Ad @ x), (lambda y: Ad.T @ y)
    rr = -np.asarray(d, np.float64).ravel().copy()
    m = None
    hs, hS, res, dirs = [], [], [], []
    for it in range(niter):
        g = adj(rr)
        if m is None:
            m = np.zeros_like(g)
        z = g
        if precond is not None:
            zz = np.asarray(precond(g, it), np.float64)
            if np.isfinite(zz).all() and zz @ g > 0:
                z = zz
        G = fwd(z)
        B = [G] + hS
        M = np.array([[bi @ bj for bj in B] for bi in B])
        rhs = -np.array([bi @ rr for bi in B])
        if M[0, 0] < 1e-20:
            break
        x = np.linalg.solve(M, rhs)
        s = x[0] * z + sum(xi * si for xi, si in zip(x[1:], hs))
        S = x[0] * G + sum(xi * Si for xi, Si in zip(x[1:], hS))
        m = m + s
        rr = rr + S
        hs = ([s] + hs)[:history]
        hS = ([S] + hS)[:history]
        res.append(rr @ rr)
        dirs.append(z)
    return m, res, dirs


def _ops(A):
    """(forward, adjoint) callables of a dense matrix or a callable pair."""
    if callable(A) or isinstance(A, tuple):
        return A
    Ad = np.asarray(A, np.float64)
    return (lambda x: Ad @ x), (lambda y: Ad.T @ y)


def lin_solve(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    """simple_solver / reg_solver / prec_solver (Born/lin_solver.cpp:7-53,
    79-101,115-155) in FP64, on operators given as dense matrices or
    (forward, adjoint) pairs:

    * "simple": min |W(L m - d)|;
    * "reg":    min |W(L m - d)|^2 + |eps R m|^2, R = aux (col_op of W L and
                the scaled R, Born/lin_solver.cpp:39-53);
    * "prec":   m = P p, min |W(L P p - d)|^2 + |eps p|^2, P = aux (col_op of
                the chain W L P and eps I, Born/lin_solver.cpp:21-37).

    With m0 the solver works on the update: the initial residual is
    W(L m0 - d) (init_rr, Born/lin_solver.cpp:79-101) and the result is the
    update.  The iteration is fcg_solve on the stacked system with the
    stacked right-hand side [W(d - L m0); 0].  Returns (model, [rr.rr])."""
    Lf, La = _ops(L)
    Wf, Wa = _ops(W) if W is not None else ((lambda y: y), (lambda y: y))
    d = np.asarray(d, np.float64).ravel()
    r = d - Lf(np.asarray(m0, np.float64).ravel()) if m0 is not None else d
    rhs = Wf(r)
    if kind == "simple":
        fwd = lambda m: Wf(Lf(m))  # noqa: E731
        adj = lambda y: La(Wa(y))  # noqa: E731
        m, res, _ = fcg_solve((fwd, adj), rhs, niter, history)
        return m, res
    Af, Aa = _ops(aux)
    nr = rhs.size
    if kind == "reg":
        def fwd(m):
            return np.concatenate([Wf(Lf(m)), eps * Af(m)])

        def adj(y):
            return La(Wa(y[:nr])) + eps * Aa(y[nr:])

        nm = La(Wa(np.zeros(nr))).size
        nextra = Af(np.zeros(nm)).size
        m, res, _ = fcg_solve((fwd, adj), np.concatenate([rhs, np.zeros(nextra)]), niter,
                              history)
        return m, res
    if kind == "prec":
        def fwd(p):
            return np.concatenate([Wf(Lf(Af(p))), eps * p])

        def adj(y):
            p = Aa(La(Wa(y[:nr])))
            return p + eps * y[nr:]

        npd = Aa(La(Wa(np.zeros(nr)))).size
        p, res, _ = fcg_solve((fwd, adj), np.concatenate([rhs, np.zeros(npd)]), niter, history)
        return Af(p), res
    raise ValueError(kind)


def tmute_start(nt, ot, dt, nrx, orx, drx, t0, vmute):
    """First kept sample of each receiver column of tmute (Born/tmute.cpp:
    28-29, 63-64): float t0 + (o2 + d2*i2)/vmute, truncated (t - o1)/d1, clamped to
    [0, nt] (the reference indexes from a negative start when the mute time
    precedes o1)."""
    f32 = np.float32
    out = np.zeros(nrx, np.int64)
    for i in range(nrx):
        t = f32(t0) + (f32(orx) + f32(drx) * f32(i)) / f32(vmute)
        f = (f32(t) - f32(ot)) / f32(dt)
        out[i] = 0 if f <= 0 else (nt if f >= nt else int(f))
    return out


def tmute_apply(x, start):
    """tmute forward (= adjoint) on data [..., nrx, nt]: zero the samples
    before start[rx]."""
    x = np.asarray(x, np.float64)
    nt = x.shape[-1]
    keep = np.arange(nt)[None, :] >= np.asarray(start)[:, None]
    return x * keep


def source_geometry(n1, n2, fz, fx, fy, wavelet_deriv, nts):
    """Born/source_func_3d.cpp:100-133: 9 points on the plane iz-1 around
    (ix, iy) with weights exp(-(dz^2+dx^2+dy^2)); trace sample t+4 holds
    weight * wavelet'(t).  Returns (locs[9], traces[9][nts])."""
    ix, iy, iz = int(fx), int(fy), int(fz)
    locs = np.zeros(9, np.int32)
    tr = np.zeros((9, nts), np.float64)
    i = 0
    i3 = iz - 1
    zn = fz - i3
    for i2 in range(iy - 1, iy + 2):
        yn = fy - i2
        for i1 in range(ix - 1, ix + 2):
            xn = fx - i1
            locs[i] = i1 + i2 * n1 + i3 * n1 * n2
            sc = math.exp(-zn * zn - xn * xn - yn * yn)
            nw = len(wavelet_deriv)
            tr[i, 4:4 + nw] = sc * np.asarray(wavelet_deriv, np.float64)
            i += 1
    return locs, tr


def receiver_index(n1, n2, sx, sy, sz):
    """Born/map_data_3d.cpp:17-29: ix + iy*n1 + (int(s_z)+4)*n1*n2."""
    ix = np.asarray(sx).astype(np.int64)
    iy = np.asarray(sy).astype(np.int64)
    iz = np.asarray(sz).astype(np.int64) + 4
    return (ix + iy * n1 + iz * n1 * n2).astype(np.int64)


def wavelet_derivative(w, dt):
    """Born/source_func_3d.cpp:10-14: forward difference / dt, last sample 0."""
    w = np.asarray(w, np.float64)
    out = np.zeros_like(w)
    out[:-1] = (w[1:] - w[:-1]) / dt
    return out


def mask_inactive(locs, n):
    """Receivers outside the update set read a field that is always zero:
    mark them -1 (they then neither extract nor inject)."""
    n1, n2, n3 = n
    rl = np.asarray(locs, np.int64).copy()
    i1, i2, i3 = rl % n1, (rl // n1) % n2, rl // (n1 * n2)
    inside = (i1 >= 4) & (i1 < n1 - 4) & (i2 >= 4) & (i2 < n2 - 4) & (i3 >= 4) & (i3 < n3 - 4)
    rl[~inside] = -1
    return rl.astype(np.int32)


class CompositeBorn:
    """FP64 restatement of rtm_zero_op::forward / adjoint
    (Born/rtm_zero_op_3d.cpp:87-293) over all shots:
    L m = D_t W B D_z E (mute . m), data [shot][rec][nt_data]."""

    def __init__(self, n, d, V, bc_a, bc_b, shots, nt_data, dt_data, jt=1, jimg=1, cimg=1e-6,
                 model_box=None, mute=None):
        self.n = tuple(n)
        n1, n2, n3 = n
        self.box = model_box or (n1, n2, n3, 0, 0, 0)
        self.mute = None if mute is None else np.asarray(mute, np.float64)
        self.nt_data, self.dt = nt_data, dt_data
        nt = (nt_data - 1) * jt + 1
        self.ntr = nt_data + 8
        self.probs, self.snaps = [], []
        # the operator's own jt-phase tables for the source and the data
        # (Born/fd_prop_3d.cpp:28-32, Born/rtm_zero_op_3d.cpp:32-36)
        sinc = sinc_table(jt) if jt > 1 else None
        for (sl, st, rl) in shots:
            P = Problem(n, d, V, bc_a, bc_b, src_locs=sl, src_tr=np.asarray(st, np.float64),
                        sinc_s=sinc, rec_locs=mask_inactive(rl, n), ntr_d=self.ntr,
                        sinc_d=sinc, jt=jt, nt=nt, jimg=jimg, cimg=cimg)
            self.probs.append(P)
            self.snaps.append(P.source_prop()[0])

    def _embed(self, m):
        m1, m2, m3, o1, o2, o3 = self.box
        n1, n2, n3 = self.n
        m = np.asarray(m, np.float64).reshape(m3, m2, m1)
        if self.mute is not None:
            m = m * self.mute.reshape(m3, m2, m1)
        comp = np.zeros((n3, n2, n1))
        comp[o3:o3 + m3, o2:o2 + m2, o1:o1 + m1] = m
        return comp

    def forward(self, m):
        mt = dz_forward(self._embed(m)).ravel()
        out = []
        for P, S in zip(self.probs, self.snaps):
            D = P.born_forward(mt, S)
            out.append(dt_forward(window_forward(D, self.nt_data), self.dt))
        return np.stack(out)

    def adjoint(self, data):
        n1, n2, n3 = self.n
        img = np.zeros(n1 * n2 * n3)
        for s, (P, S) in enumerate(zip(self.probs, self.snaps)):
            tr = window_adjoint(dt_adjoint(data[s], self.dt), self.ntr)
            img += P.rtm_adjoint(tr, S)
        tmp = dz_adjoint(img.reshape(n3, n2, n1))
        m1, m2, m3, o1, o2, o3 = self.box
        m = tmp[o3:o3 + m3, o2:o2 + m2, o1:o1 + m1].copy()
        if self.mute is not None:
            m *= self.mute.reshape(m3, m2, m1)
        return m
