"""GPU tests of the operator / solver layer (rtm_zero_op, my_vector, cgstep)
against the FP64 oracle's composite operator and numpy CG."""
import sys

import numpy as np
import pytest

from oracle import born_oracle as O
from sep_code_amd import configs
from sep_code_amd.operator import DeviceVector, RtmZeroOp, SimpleSolver, cgstep

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, file=sys.stderr, flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def small_workload(nt=90, n_side=40, store="f32", jimg=1):
    w = configs.lsrtm(seed=5, n_side=n_side, nt=nt, nshots=(2, 1), rec_step=3, jimg=jimg,
                      store=store)
    return w


def setup(dev, w, box=True, mute=True):
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    n1, n2, n3 = w.n
    nb = w.nbound
    mbox = (n1 - 2 * nb, n2 - 2 * nb, n3 - 2 * nb, nb, nb, nb) if box else None
    mu = None
    if mute:
        m1, m2, m3 = (mbox or (n1, n2, n3, 0, 0, 0))[:3]
        mu = np.ones((m3, m2, m1), np.float32)
        mu[:3] = 0.0  # mute the top of the model (the reference's mute volume)
    return mbox, mu


def oracle_for(w, mbox, mu):
    shots = [(s.src_locs, s.src_traces, s.rec_locs) for s in w.shots]
    return O.CompositeBorn(w.n, w.d, w.V.astype(np.float64), w.bc_a, w.bc_b, shots, w.nt,
                           w.dt, jt=w.jt, jimg=w.jimg, cimg=w.cimg, model_box=mbox, mute=mu)


@pytest.mark.parametrize("store", ["f32", "bf16", "ckpt"])
def test_operator_dot_test(dev, store):
    announce(f"operator dot test store={store}")
    w = small_workload(store=store, jimg=2 if store == "bf16" else 1)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store=store, ckpt=20, model_box=mbox, mute=mu)
    op.prepare()
    e0, e1 = op.dot_test(verb=True)
    announce(f"  dot test errors {e0:.2e} {e1:.2e}")
    assert e0 < 1e-4 and e1 < 1e-4
    op.close()


def test_operator_matches_oracle(dev):
    announce("operator forward/adjoint vs oracle")
    w = small_workload()
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    orc = oracle_for(w, mbox, mu)
    r = np.random.default_rng(0)
    m = r.standard_normal(op.model_size).astype(np.float32)
    mv = op.model_vector(m)
    dv = op.data_vector()
    op.forward(False, mv, dv)
    Do = orc.forward(m)
    assert rel(dv.numpy(), Do) < 1e-5
    d = r.standard_normal(op.data_size).astype(np.float32)
    dv.upload(d)
    op.adjoint(False, mv, dv)
    mo = orc.adjoint(d.reshape(Do.shape))
    assert rel(mv.numpy(), mo) < 1e-4
    # add = true accumulates
    op.adjoint(True, mv, dv)
    assert rel(mv.numpy(), 2 * mo) < 1e-4
    op.close()


def test_vector_ops(dev):
    dev.create_space((10.0, 10.0, 10.0), 4.0, 0.0, (16, 16, 16))
    r = np.random.default_rng(1)
    a = r.standard_normal(100003).astype(np.float32)
    b = r.standard_normal(100003).astype(np.float32)
    va, vb = DeviceVector.from_array(dev, a), DeviceVector.from_array(dev, b)
    assert abs(va.dot(vb) - float(a.astype(np.float64) @ b.astype(np.float64))) < 1e-9 * 1e5
    va.scale_add(2.0, vb, -3.0)
    np.testing.assert_allclose(va.numpy(), 2 * a - 3 * b, rtol=1e-6, atol=1e-6)
    vc = va.clone().add(vb)
    np.testing.assert_allclose(vc.numpy(), 2 * a - 2 * b, rtol=1e-6, atol=1e-6)
    vr = DeviceVector(dev, 1000).random(7)
    x = vr.numpy()
    assert x.min() >= -0.5 and x.max() < 0.5 and abs(x.mean()) < 0.05
    np.testing.assert_array_equal(DeviceVector(dev, 1000).random(7).numpy(), x)


def test_cg_matches_numpy_cgstep(dev):
    """Five device cgstep iterations against the FP64 restatement of
    Born/cgstep.cpp driven on the oracle's composite operator."""
    announce("cg vs oracle cgstep")
    w = small_workload(nt=70, n_side=36)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    orc = oracle_for(w, mbox, mu)
    m_true = (w.refl[mbox[5]:mbox[5] + mbox[2], mbox[4]:mbox[4] + mbox[1],
                     mbox[3]:mbox[3] + mbox[0]]).astype(np.float64)
    d = orc.forward(m_true)
    dv = op.data_vector(d)
    res = SimpleSolver(cgstep(), dv, op).solve(5)
    # oracle CG in FP64 on the same operator
    mo = np.zeros(m_true.size)
    rr = -d.ravel().copy()
    s = np.zeros_like(mo)
    ss = np.zeros_like(rr)
    hist = [1.0]
    for it in range(5):
        g = orc.adjoint(rr.reshape(d.shape)).ravel()
        gg = orc.forward(g).ravel()
        if it == 0:
            alfa, beta = -(gg @ rr) / (gg @ gg), 0.0
        else:
            a, b_, c = gg @ gg, ss @ ss, gg @ ss
            det = a * b_ * (1 - c / a * c / b_)
            gr, sr = -(gg @ rr), -(ss @ rr)
            alfa, beta = (b_ * gr - c * sr) / det, (-c * gr + a * sr) / det
        s = beta * s + alfa * g
        ss = beta * ss + alfa * gg
        mo += s
        rr += ss
        hist.append(np.linalg.norm(rr) / np.linalg.norm(d))
    announce(f"  device {res.resid_hist} oracle {hist}")
    np.testing.assert_allclose(res.resid_hist, hist, rtol=2e-3)
    assert all(np.diff(res.resid_hist) <= 1e-7)
    op.close()


def test_cg_model_matches_numpy_cgstep(dev):
    """Model after three iterations, device vs FP64 oracle CG."""
    w = small_workload(nt=60, n_side=32)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    orc = oracle_for(w, mbox, mu)
    r = np.random.default_rng(9)
    m_true = r.standard_normal(op.model_size)
    d = orc.forward(m_true)
    solver = SimpleSolver(cgstep(), op.data_vector(d), op)
    solver.solve(3)
    mo = np.zeros(m_true.size)
    rr = -d.ravel().copy()
    s = np.zeros_like(mo)
    ss = np.zeros_like(rr)
    for it in range(3):
        g = orc.adjoint(rr.reshape(d.shape)).ravel()
        gg = orc.forward(g).ravel()
        if it == 0:
            alfa, beta = -(gg @ rr) / (gg @ gg), 0.0
        else:
            a, b_, c = gg @ gg, ss @ ss, gg @ ss
            det = a * b_ * (1 - c / a * c / b_)
            gr, sr = -(gg @ rr), -(ss @ rr)
            alfa, beta = (b_ * gr - c * sr) / det, (-c * gr + a * sr) / det
        s = beta * s + alfa * g
        ss = beta * ss + alfa * gg
        mo += s
        rr += ss
    assert rel(solver.return_model().numpy(), mo) < 1e-3
    op.close()


def test_cg_offset_judges_noise_free_misfit(dev):
    """With offset e the solver reports |L m - (d - e)| / |d - e| (checked
    against a fresh application of the operator)."""
    w = small_workload(nt=60, n_side=32)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    r = np.random.default_rng(4)
    mv = op.model_vector(r.standard_normal(op.model_size))
    dv = op.data_vector()
    op.forward(False, mv, dv)
    clean = dv.numpy().astype(np.float64)
    e = (0.05 * np.sqrt((clean ** 2).mean()) * r.standard_normal(clean.size)).astype(np.float32)
    dv.upload(clean + e)
    ev = op.data_vector(e)
    solver = SimpleSolver(cgstep(), dv, op)
    res = solver.solve(4, offset=ev)
    lm = op.data_vector()
    op.forward(False, solver.return_model(), lm)
    ref = np.linalg.norm(lm.numpy().astype(np.float64) - clean) / np.linalg.norm(clean)
    assert abs(res.rel_resid - ref) < 1e-4 * max(ref, 1e-3) + 1e-6
    assert res.resid_hist[0] == pytest.approx(np.linalg.norm(clean + e - e) /
                                              np.linalg.norm(clean), rel=1e-5)
    op.close()


def test_cg_zero_data_breaks_down_cleanly(dev):
    w = small_workload(nt=40, n_side=32)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    res = SimpleSolver(cgstep(), op.data_vector(), op).solve(3)
    assert res.breakdown and res.iterations == 0
    op.close()


def test_cg_residual_stop_and_bad_preconditioner_fallback(dev):
    announce("cg stop + preconditioner fallback")
    w = small_workload(nt=60, n_side=32)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    mv = op.model_vector(np.random.default_rng(3).standard_normal(op.model_size))
    dv = op.data_vector()
    op.forward(False, mv, dv)
    full = SimpleSolver(cgstep(), dv, op).solve(4)
    tol = full.resid_hist[2] * 1.0001
    stopped = SimpleSolver(cgstep(), dv, op).solve(10, rel_tol=tol)
    assert stopped.converged and stopped.iterations == 2

    from sep_code_amd import _lib

    def sabotage(g_h, z_h):  # z = -g: not a descent direction -> fallback to g
        lib = _lib.load()
        lib.brtm_vec_copy(z_h, g_h)
        lib.brtm_vec_scale(z_h, -1.0)
        return 0

    bad = SimpleSolver(cgstep(), dv, op, precond=sabotage).solve(4)
    assert bad.fallbacks == 4
    np.testing.assert_allclose(bad.resid_hist, full.resid_hist, rtol=1e-5)
    op.close()


@pytest.mark.parametrize("history,niter", [(1, 6), (3, 6), (16, 12)])
def test_flexible_cg_matches_oracle(dev, history, niter):
    """Flexible CG with a preconditioner that changes every iteration (a
    different positive diagonal each time): device residual history and model
    against the FP64 oracle's fcg_solve on the same composite operator.  With
    history 16 over 12 iterations the search holds up to 11 earlier steps."""
    announce(f"flexible cg history={history} niter={niter}")
    w = small_workload(nt=60, n_side=32)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    orc = oracle_for(w, mbox, mu)
    r = np.random.default_rng(17)
    m_true = r.standard_normal(op.model_size)
    d = orc.forward(m_true)
    # to about unit norm by a power of two (exact in FP32 and FP64): the
    # solvers stop when |A z|^2 < 1e-20, which the unscaled data (|d| ~ 8e-3)
    # reach after 10 iterations
    d *= 2.0 ** -np.round(np.log2(np.linalg.norm(d)))
    mshape = (mbox[2], mbox[1], mbox[0])
    P = [r.uniform(0.3, 3.0, op.model_size) for _ in range(niter)]
    calls = []

    def pre(g_h, z_h):
        it = len(calls)
        calls.append(it)
        g = DeviceVector.view(dev, g_h, op.model_size).numpy()
        DeviceVector.view(dev, z_h, op.model_size).upload(P[it] * g)
        return 0

    solver = SimpleSolver(cgstep(), op.data_vector(d), op, precond=pre, history=history)
    res = solver.solve(niter)
    fwd = lambda x: orc.forward(x.reshape(mshape)).ravel()  # noqa: E731
    adj = lambda y: orc.adjoint(y.reshape(d.shape)).ravel()  # noqa: E731
    mo, ro, _ = O.fcg_solve((fwd, adj), d, niter, history, precond=lambda g, it: P[it] * g)
    hist = [1.0] + [np.sqrt(x) / np.linalg.norm(d) for x in ro]
    announce(f"  device {np.round(res.resid_hist, 5)} oracle {np.round(hist, 5)}")
    assert res.fallbacks == 0 and res.iterations == niter
    np.testing.assert_allclose(res.resid_hist, hist, rtol=2e-3)
    assert rel(solver.return_model().numpy(), mo) < 2e-3
    op.close()


def test_flexible_cg_history_is_cg_without_preconditioner(dev):
    """Without a preconditioner the history terms vanish: flexible CG with
    four steps follows cgstep."""
    w = small_workload(nt=60, n_side=32)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store="f32", model_box=mbox, mute=mu)
    op.prepare()
    mv = op.model_vector(np.random.default_rng(4).standard_normal(op.model_size))
    dv = op.data_vector()
    op.forward(False, mv, dv)
    a = SimpleSolver(cgstep(), dv, op).solve(6)
    b = SimpleSolver(cgstep(), dv, op, history=4).solve(6)
    np.testing.assert_allclose(b.resid_hist, a.resid_hist, rtol=1e-4)
    with pytest.raises(RuntimeError):
        SimpleSolver(cgstep(), dv, op, history=99).solve(2)
    op.close()


@pytest.mark.parametrize("store,tol", [("f32", 1e-5), ("bf16", 1e-2)])
def test_illumination_matches_oracle(dev, store, tol):
    """sum over shots and snapshots of s^2, windowed to the model box and
    muted, against the oracle's FP64 snapshots."""
    announce(f"illumination store={store}")
    w = small_workload(nt=60, n_side=32, store=store, jimg=2 if store == "bf16" else 1)
    mbox, mu = setup(dev, w)
    op = RtmZeroOp(dev, w, store=store, model_box=mbox, mute=mu)
    op.prepare()
    orc = oracle_for(w, mbox, mu)
    acc = np.zeros(int(np.prod(w.n)))
    for S in orc.snaps:
        for s in S:
            acc += np.asarray(s, np.float64).ravel() ** 2
    m1, m2, m3, o1, o2, o3 = mbox
    n1, n2, n3 = w.n
    ref = acc.reshape(n3, n2, n1)[o3:o3 + m3, o2:o2 + m2, o1:o1 + m1] * mu
    got = op.illumination().numpy().reshape(m3, m2, m1)
    assert rel(got, ref) < tol
    op.close()
