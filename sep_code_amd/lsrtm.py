"""Least-squares RTM driver (the RTM_inv workflow, Born/RTM_inv.cpp:14-64).

``run_lsrtm`` builds the operator for a workload, synthesises observed data
from the true reflectivity plus Gaussian noise, and solves with plain CG
(cgstep) or with a preconditioner (flexible CG: the preconditioned gradient
enters cgstep's plane search), timing from the start of the solver loop to
the first iteration whose judged residual is at or below ``rel_tol``.  With
the noise vector e passed, the judged residual is the misfit to the
noise-free data, |L m - (d - e)| / |d - e|; without it, |L m - d| / |d|.
Both are recorded every iteration (the second as ``data_resid_hist``), and
both are re-evaluated at the end with a fresh application of the plain FP32
operator and FP64 norms.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import configs
from .device import BornDevice
from .operator import DeviceVector, RtmZeroOp, SimpleSolver, cgstep


@dataclasses.dataclass
class LsrtmResult:
    name: str
    iterations: int
    converged: bool
    seconds: float
    rel_resid: float
    rel_resid_recomputed: float
    resid_hist: list
    time_hist: list
    fallbacks: int
    prepare_seconds: float
    noise_floor: float
    # |L m - d| / |d| against the noisy data as given (the residual
    # definition of the LSRTM metric), running and recomputed at the end;
    # rel_resid above is the misfit to the noise-free data when noise is
    # passed, else the same quantity
    data_rel_resid: float = float("nan")
    data_rel_resid_recomputed: float = float("nan")
    data_resid_hist: list = dataclasses.field(default_factory=list)


def model_box(w):
    n1, n2, n3 = w.n
    nb = w.nbound
    return (n1 - 2 * nb, n2 - 2 * nb, n3 - 2 * nb, nb, nb, nb)


def crop(w, vol):
    m1, m2, m3, o1, o2, o3 = model_box(w)
    return np.ascontiguousarray(vol[o3:o3 + m3, o2:o2 + m2, o1:o1 + m1])


def model_mute(w):
    """The reference's mute volume (Born/image_rtm_3d.cpp:65-70): zero over
    the shallow near field, one below."""
    m1, m2, m3 = model_box(w)[:3]
    mute = np.ones((m3, m2, m1), np.float32)
    mute[:w.mute_top] = 0.0
    return mute


def setup_operator(dev: BornDevice, w, store=None, ckpt=None):
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    op = RtmZeroOp(dev, w, store=store, ckpt=ckpt, model_box=model_box(w), mute=model_mute(w))
    t0 = time.perf_counter()
    op.prepare()
    dev.sync()
    return op, time.perf_counter() - t0


def observed_data(op, w, noise_frac=0.01, seed=1234, keep_noise=False):
    """d = L m_true + e with Gaussian e of sigma = noise_frac * RMS(L m_true)
    (|e| / |L m_true| = noise_frac).  Returns (d, floor[, e]) with floor =
    |e| / |d|, the misfit of the noise-free data itself."""
    mt = op.model_vector(crop(w, w.refl))
    d = op.data_vector()
    op.forward(False, mt, d)
    clean = d.numpy()
    rng = configs.SplitMix64(seed)
    sigma = noise_frac * float(np.sqrt((clean.astype(np.float64) ** 2).mean()))
    noise = (rng.normal_array(clean.size) * sigma).astype(np.float32)
    d.upload(clean + noise)
    floor = float(np.linalg.norm(noise.astype(np.float64)) /
                  np.linalg.norm((clean + noise).astype(np.float64)))
    if keep_noise:
        return d, floor, op.data_vector(noise)
    return d, floor


def illumination_weight(op, eps=1e-2, power=1.0):
    """Diagonal w = (I + eps max I)^-power of the source illumination I
    (sum over shots and stored snapshots of s^2, on the model grid) as a
    device vector.  power = 1 is the classical illumination-compensation
    preconditioner; power = 1/2 its square root, for split use."""
    illum = op.illumination()
    a = illum.numpy().astype(np.float64)
    illum.close()
    wv = (a + eps * a.max()) ** -power
    return DeviceVector.from_array(op.dev, (wv / wv.max()).astype(np.float32))


def diagonal_precond(dev, wv):
    """Preconditioner callback z = w . g (device vectors, no host copies)."""
    n = wv.n

    def pre(g_h, z_h):
        DeviceVector.view(dev, z_h, n).copy_from(DeviceVector.view(dev, g_h, n)).mult(wv)
        return 0
    return pre


def run_lsrtm(dev, w, op, d, niter=60, rel_tol=1e-2, precond=None, name="cg", verbose=False,
              floor=0.0, prepare_seconds=0.0, noise=None, history=0):
    """Solve and judge.  With noise e given, the residual judged (for the
    stop and the final check) is the misfit to the noise-free data d - e.
    history >= 2 selects flexible CG with that many previous steps."""
    solver = SimpleSolver(cgstep(), d, op, precond=precond, verbose=verbose, history=history)
    res = solver.solve(niter, rel_tol=rel_tol, offset=noise)
    # judge: fresh plain-operator residuals with FP64 norms, against the
    # noise-free data (when the noise is known) and against d itself
    m = solver.return_model()
    Lm = op.data_vector()
    op.forward(False, m, Lm)
    r = Lm.clone().scale_add(1.0, d, -1.0)
    rel_d = float(np.sqrt(r.dot(r) / d.dot(d)))
    rel = rel_d
    if noise is not None:
        ref = d.clone().scale_add(1.0, noise, -1.0)
        r.copy_from(Lm).scale_add(1.0, ref, -1.0)
        rel = float(np.sqrt(r.dot(r) / ref.dot(ref)))
        ref.close()
    r.close()
    Lm.close()
    return LsrtmResult(name, res.iterations, res.converged, float(res.time_hist[-1]),
                       res.rel_resid, rel, [float(x) for x in res.resid_hist],
                       [float(x) for x in res.time_hist], res.fallbacks, prepare_seconds,
                       floor, float(res.data_rel_resid), rel_d,
                       [float(x) for x in res.data_resid_hist]), m
