"""GPU tests of the composable operators and the three solver forms
(include/born_solver.h) against the FP64 oracle: dense caller-defined
operators through callbacks, tmute, the combinators' adjointness, and
reg_solver / prec_solver / simple_solver with a weight and a starting model
on the Born operator."""
import sys

import numpy as np
import pytest

from oracle import born_oracle as O
from sep_code_amd import configs
from sep_code_amd.operator import (BornOp, CallbackOp, ColOp, DeviceVector, DiagOp, IOp,
                                   PrecSolver, RegSolver, RowOp, RtmZeroOp, SimpleSolver,
                                   TMute, WeightOp, cg_solve, cgstep)

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, file=sys.stderr, flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dense_op(dev, M):
    """A dense matrix as a caller-defined operator (host multiply)."""
    M = np.asarray(M, np.float64)

    def fwd(add, x, y):
        v = M @ x.numpy().astype(np.float64)
        y.upload((y.numpy() + v) if add else v)

    def adj(add, y, x):
        v = M.T @ y.numpy().astype(np.float64)
        x.upload((x.numpy() + v) if add else v)

    return CallbackOp(dev, M.shape[1], M.shape[0], fwd, adj)


def small_problem(r, nd=40, nm=16):
    L = r.standard_normal((nd, nm))
    d = r.standard_normal(nd)
    w = r.uniform(0.5, 2.0, nd)
    R = np.eye(nm) - np.eye(nm, k=1)
    p = r.uniform(0.2, 3.0, nm)
    m0 = r.standard_normal(nm)
    return L, d, w, R, p, m0


def test_dense_solver_forms_match_oracle(dev):
    """Every solver form on small dense operators through the device solver
    (FP32 vectors, FP64 scalars) against the FP64 oracle: after 5 steps, and
    after 30 by continuing the same solver, against the closed form (FP32 CG
    loses exact termination at 16 = unknowns: a numpy FP32 restatement of
    the prec case is 3.4e-3 from it there and 1.6e-7 at 30)."""
    announce("dense solver forms")
    dev.create_space((10.0, 10.0, 10.0), 4.0, 0.0, (16, 16, 16))
    r = np.random.default_rng(4)
    L, d, w, R, p, m0 = small_problem(r)
    Lop, Rop = dense_op(dev, L), dense_op(dev, R)
    wv, pv = DeviceVector.from_array(dev, w), DeviceVector.from_array(dev, p)
    W, P = WeightOp(wv), WeightOp(pv)
    dv, m0v = DeviceVector.from_array(dev, d), DeviceVector.from_array(dev, m0)
    eps = 0.5
    cases = [
        ("simple", lambda: SimpleSolver(cgstep(), dv, Lop, wt=W, m0=m0v), {}),
        ("reg", lambda: RegSolver(cgstep(), dv, Lop, Rop, eps, wt=W, m0=m0v),
         {"aux": R, "eps": eps}),
        ("prec", lambda: PrecSolver(cgstep(), dv, Lop, P, eps, wt=W), {"aux": np.diag(p),
                                                                         "eps": eps}),
    ]
    for kind, make, kw in cases:
        m0k = None if kind == "prec" else m0
        s = make()
        s.solve(5)
        mo, _ = O.lin_solve(kind, L, d, 5, W=np.diag(w), m0=m0k, **kw)
        e5 = rel(s.return_model().numpy(), mo)
        s.solve(25)  # continues: 30 iterations in all
        mc, _ = O.lin_solve(kind, L, d, 40, W=np.diag(w), m0=m0k, **kw)
        e30 = rel(s.return_model().numpy(), mc)
        announce(f"  {kind}: 5 it {e5:.2e}, 30 it vs closed form {e30:.2e}")
        assert e5 < 1e-4 and e30 < 1e-4, kind
        s.close()


def test_update_model_and_continuation(dev):
    dev.create_space((10.0, 10.0, 10.0), 4.0, 0.0, (16, 16, 16))
    r = np.random.default_rng(8)
    L, d, w, R, p, m0 = small_problem(r)
    Lop = dense_op(dev, L)
    dv, m0v = DeviceVector.from_array(dev, d), DeviceVector.from_array(dev, m0)
    a = SimpleSolver(cgstep(), dv, Lop, m0=m0v)
    a.solve(2)
    a.solve(3)
    b = SimpleSolver(cgstep(), dv, Lop, m0=m0v)
    res = b.solve(5)
    np.testing.assert_allclose(a.return_model().numpy(), b.return_model().numpy(), rtol=1e-6,
                               atol=1e-6)
    # the model is the update: |L (m0 + update) - d| = rel_resid |L m0 - d|
    full = b.update_model(m0v.clone()).numpy().astype(np.float64)
    assert np.linalg.norm(L @ full - d) == pytest.approx(
        res.rel_resid * np.linalg.norm(L @ m0 - d), rel=1e-3)


def test_tmute_matches_oracle(dev):
    dev.create_space((10.0, 10.0, 10.0), 4.0, 0.0, (16, 16, 16))
    nt, nrx, nouter = 300, 7, 6
    geo = dict(nt=nt, ot=-0.02, dt=0.002, nrx=nrx, orx=-150.0, drx=75.0)
    T = TMute(dev, 0.05, 1800.0, nouter=nouter, **geo)
    start = O.tmute_start(nt, geo["ot"], geo["dt"], nrx, geo["orx"], geo["drx"], 0.05, 1800.0)
    x = np.random.default_rng(1).standard_normal((nouter, nrx, nt)).astype(np.float32)
    xv = DeviceVector.from_array(dev, x)
    yv = T.range_vector()
    T.forward(False, xv, yv)
    np.testing.assert_array_equal(yv.numpy().reshape(x.shape), O.tmute_apply(x, start))
    T.adjoint(True, xv, yv)  # x + mute(y) = x + mute(x)
    np.testing.assert_array_equal(xv.numpy().reshape(x.shape),
                                  x + O.tmute_apply(x, start).astype(np.float32))
    assert start.min() == 0 and start.max() < nt and len(set(start)) > 3  # clamp at 0 too


def small_workload(nt=60, n_side=32):
    return configs.lsrtm(seed=5, n_side=n_side, nt=nt, nshots=(2, 1), rec_step=3, store="f32")


def born_setup(dev, w):
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    n1, n2, n3 = w.n
    nb = w.nbound
    mbox = (n1 - 2 * nb, n2 - 2 * nb, n3 - 2 * nb, nb, nb, nb)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox)
    op.prepare()
    shots = [(s.src_locs, s.src_traces, s.rec_locs) for s in w.shots]
    orc = O.CompositeBorn(w.n, w.d, w.V.astype(np.float64), w.bc_a, w.bc_b, shots, w.nt, w.dt,
                          jt=w.jt, jimg=w.jimg, cimg=w.cimg, model_box=mbox)
    return op, orc


def test_combinator_dot_tests(dev):
    announce("combinator dot tests")
    w = small_workload()
    op, _ = born_setup(dev, w)
    B = BornOp(op)
    nm, nd = op.model_size, op.data_size
    wv = DeviceVector(dev, nm).random(3)
    T = TMute(dev, 0.02, 2000.0, nt=w.nt, ot=0.0, dt=w.dt, nrx=op.nrx, orx=0.0, drx=37.5,
              nouter=op.nry * len(w.shots))
    for name, A in [("born", B), ("col(born, 0.3 I)", ColOp(B, IOp(dev, nm).set_scale(0.3))),
                    ("row(weight, born, tmute)", RowOp(WeightOp(wv), B, T)),
                    ("diag(born, I)", DiagOp(B, IOp(dev, 5))),
                    ("2.5 born", BornOp(op).set_scale(2.5))]:
        e0, e1 = A.dot_test()
        announce(f"  {name}: {e0:.2e} {e1:.2e}")
        assert e0 < 1e-4 and e1 < 1e-4, name
    # the stacked parts land where the layout says
    x = DeviceVector(dev, nm).random(9)
    y = DeviceVector(dev, nd + nm)
    ColOp(B, IOp(dev, nm).set_scale(0.3)).forward(False, x, y)
    ref = op.data_vector()
    op.forward(False, x, ref)
    yy = y.numpy()
    np.testing.assert_array_equal(yy[:nd], ref.numpy())
    np.testing.assert_allclose(yy[nd:], 0.3 * x.numpy(), rtol=1e-7)


def test_born_reg_and_prec_solvers_match_oracle(dev):
    """reg_solver (eps times the identity), prec_solver (a diagonal P) and a
    tmuted simple_solver from a starting model, three cgstep
    iterations each, against the FP64 oracle on the composite operator."""
    announce("born reg / prec / weighted solvers")
    w = small_workload()
    op, orc = born_setup(dev, w)
    nm = op.model_size
    r = np.random.default_rng(12)
    m_true = r.standard_normal(nm)
    d = orc.forward(m_true)
    dshape = d.shape
    dv = op.data_vector(d)
    Lp = (lambda m: orc.forward(m).ravel(), lambda y: orc.adjoint(y.reshape(dshape)).ravel())
    scale = np.sqrt(np.mean(d ** 2) / np.mean(m_true ** 2))
    eps = 0.1 * scale
    # reg
    s = RegSolver(cgstep(), dv, op, IOp(dev, nm), eps)
    s.solve(3)
    mo, _ = O.lin_solve("reg", Lp, d, 3, aux=np.eye(nm), eps=eps)
    e = rel(s.return_model().numpy(), mo)
    announce(f"  reg: {e:.2e}")
    assert e < 1e-3
    # prec with a diagonal P
    pw = r.uniform(0.5, 2.0, nm)
    P = WeightOp(DeviceVector.from_array(dev, pw))
    s = PrecSolver(cgstep(), dv, op, P, eps)
    s.solve(3)
    mo, _ = O.lin_solve("prec", Lp, d, 3, aux=np.diag(pw), eps=eps)
    e = rel(s.return_model().numpy(), mo)
    announce(f"  prec: {e:.2e}")
    assert e < 1e-3
    # tmuted data weight and a starting model
    start = O.tmute_start(w.nt, 0.0, w.dt, op.nrx, 0.0, 5.0, 0.03, 2500.0)
    T = TMute(dev, 0.03, 2500.0, nt=w.nt, ot=0.0, dt=w.dt, nrx=op.nrx, orx=0.0, drx=5.0,
              nouter=op.nry * len(w.shots))
    m0 = 0.5 * m_true
    s = SimpleSolver(cgstep(), dv, op, wt=T, m0=op.model_vector(m0))
    s.solve(3)
    Wp = (lambda y: O.tmute_apply(y.reshape(-1, op.nrx, w.nt), start).ravel(),
          lambda y: O.tmute_apply(y.reshape(-1, op.nrx, w.nt), start).ravel())
    mo, _ = O.lin_solve("simple", Lp, d, 3, W=Wp, m0=m0)
    e = rel(s.return_model().numpy(), mo)
    announce(f"  tmute + m0: {e:.2e}")
    assert e < 1e-3


def test_cg_solve_entry_point_equals_simple_solver(dev):
    w = small_workload(nt=50, n_side=32)
    op, orc = born_setup(dev, w)
    d = orc.forward(np.random.default_rng(2).standard_normal(op.model_size))
    dv = op.data_vector(d)
    m = op.model_vector()
    a = cg_solve(op, dv, m, 3)
    s = SimpleSolver(cgstep(), dv, op)
    b = s.solve(3)
    np.testing.assert_array_equal(m.numpy(), s.return_model().numpy())
    np.testing.assert_array_equal(a.resid_hist, b.resid_hist)
