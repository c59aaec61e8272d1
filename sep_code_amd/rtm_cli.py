"""Command-line programs on SEP datasets, the MI355X counterparts of the
reference's Born/RTM_fwd.cpp (Born modelling), Born/RTM2_3d.cpp (RTM image)
and Born/RTM_inv.cpp (LSRTM with cgstep):

    python -m sep_code_amd.rtm_cli fwd velocity=v.H wavelet=w.H data=d.H image=m.H dataout=o.H
    python -m sep_code_amd.rtm_cli rtm velocity=v.H wavelet=w.H data=d.H image=img.H
    python -m sep_code_amd.rtm_cli inv velocity=v.H wavelet=w.H data=d.H image=img.H niter=20

Datasets and their axes, as the reference reads them:
  velocity  v(x, y, z) in m/s, the model grid          (Born/vel_fd_3d.cpp:3-12)
  wavelet   w(t), sampled at the data interval         (Born/source_func_3d.cpp:4-19)
  data      d(t, rx, ry, sx, sy): receivers on axes 2-3, shots on axes 4-5
            (Born/data_rtm_3d.cpp:4-77); fwd reads only its axes
  image     the model / image on the velocity grid      (Born/image_rtm_3d.cpp:4-16)
  mute      optional, velocity grid, multiplies the model (Born/image_rtm_3d.cpp:65-70)

Parameters (the reference's names and defaults): src_depth=0 rec_depth=0
bc_a=50 bc_b=0.0005 nbound=bc_a (Born/fd_prop_3d.cpp:37-42), niter=5 (inv).
aper= and n_gpus= are accepted and have no effect: the reference stores the
aperture without using it (Born/source_func_3d.cpp:37-63), and one MI355X
holds the whole survey.  Additional: store=f32|bf16|f16|ckpt (source-wavefield
store; ckpt for fwd and rtm, f16 for inv by default), ckpt=, jimg= (imaging
cadence, default jt),
rel_tol= (inv: stop at that relative residual), precond=none|diag|learned
(inv; illum_eps= the illumination floor), history= (inv, flexible-CG steps,
8 by default with precond=learned; each holds one data-space vector),
eps= (inv: reg_solver with eps times the identity), tmute= vmute= (inv: the
tmute data weight, mute before tmute + x/vmute), device= (GPU index).

The time step follows Born/fd_prop_3d.cpp:7-25 (the data interval divided
into jt equal steps), with the 3-D 8th-order leapfrog limit 0.45 dmin/vmax
in place of the reference's 0.5 dmin/vmax, which can exceed it.  The model is
padded by nbound + 4 cells (and the reference's extra depth padding), each
axis rounded up to a multiple of 16 (Born/source_func_3d.cpp:37-63); the
velocity outside the model repeats the nearest model sample
(Born/vel_fd_3d.cpp:42-56) without the reference's random perturbation.
"""
from __future__ import annotations

import dataclasses
import math
import sys
import time

import numpy as np

from . import configs, geometry as G, sep

FAT = 4
COURANT = 0.45  # below the 8th-order 3-D leapfrog limit 0.4529


@dataclasses.dataclass
class SepProblem:
    workload: configs.Workload
    model_axes: list          # velocity / image axes (x, y, z)
    data_axes: list           # (t, rx, ry, sx, sy)
    model_box: tuple          # (m1, m2, m3, o1, o2, o3) on the compute grid
    mute: np.ndarray | None   # (m3, m2, m1) or None


def time_step(dt_data, vmax, dmin, courant=COURANT):
    """(dt, jt): the data interval split into jt steps of at most
    courant * dmin / vmax (Born/fd_prop_3d.cpp:7-25)."""
    jt = int(math.ceil(dt_data / (courant * dmin / vmax) - 1e-9))
    jt = max(jt, 1)
    return dt_data / jt, jt


def pad_velocity(vel, vaxes, n, nbound):
    """Velocity on the compute grid (origin -d*nbound on each axis): the
    nearest-lower model sample, clamped (Born/vel_fd_3d.cpp:42-56)."""
    idx = []
    for a, (ax, nn) in enumerate(zip(vaxes, n)):
        x = -ax.d * nbound + ax.d * np.arange(nn)
        i = np.floor((x - ax.o) / ax.d + 1e-6).astype(np.int64)
        idx.append(np.clip(i, 0, ax.n - 1))
    i1, i2, i3 = idx
    return np.ascontiguousarray(vel[np.ix_(i3, i2, i1)], dtype=np.float32)


def problem_from_sep(p: sep.Params, store: str) -> SepProblem:
    vel, vaxes = sep.read(p.get_string("velocity"))
    if len(vaxes) != 3:
        raise sep.SepError("velocity must be 3-D (x, y, z)")
    wav, waxes = sep.read(p.get_string("wavelet"))
    dh = sep.read_header(p.get_string("data"))
    daxes = sep.header_axes(dh, ndim=5)
    at, arx, ary, asx, asy = daxes

    bc_a = p.get_float("bc_a", 50.0)
    bc_b = p.get_float("bc_b", 0.0005)
    nb = p.get_int("nbound", int(bc_a))
    src_depth = p.get_float("src_depth", 0.0)
    rec_depth = p.get_float("rec_depth", 0.0)

    m = (vaxes[0].n, vaxes[1].n, vaxes[2].n)
    d = (vaxes[0].d, vaxes[1].d, vaxes[2].d)
    n, nbound = G.compute_size(m, nb, fat=FAT, extra_z=True)
    dt, jt = time_step(at.d, float(vel.max()), min(d))
    if abs(waxes[0].d - at.d) > 1e-6 * at.d:
        # resample the wavelet to the data interval (linear)
        t_w = waxes[0].o + waxes[0].d * np.arange(waxes[0].n)
        wav = np.interp(at.o + at.d * np.arange(at.n), t_w, np.asarray(wav, np.float64).ravel())
    wder = G.wavelet_derivative(np.asarray(wav, np.float64).ravel(), at.d)
    ntr = at.n + 8

    def frac(pos, a):
        return (np.asarray(pos, np.float64) - (-d[a] * nbound)) / d[a]

    # receivers: a fixed spread on data axes 2-3 (Born/data_rtm_3d.cpp:65-73)
    ry, rx = np.meshgrid(ary.coords(), arx.coords(), indexing="ij")
    rlocs = G.receiver_locs(n[0], n[1], frac(rx.ravel(), 0), frac(ry.ravel(), 1),
                            np.full(rx.size, frac(rec_depth, 2)))
    shots = []
    for sy in asy.coords():
        for sx in asx.coords():
            slocs, strc = G.source_points(n[0], n[1], float(frac(sx, 0)), float(frac(sy, 1)),
                                          float(frac(src_depth, 2)), wder, ntr)
            shots.append(configs.Shot(slocs, strc, rlocs, arx.n, ary.n))
    velc = pad_velocity(vel, vaxes, n, nbound)
    jimg = p.get_int("jimg", jt)
    w = configs.Workload("sep", n, d, dt, at.n, jt, jimg, bc_a, bc_b, nbound, velc,
                         np.zeros((n[2], n[1], n[0]), np.float32), shots, ntr, store=store,
                         ckpt=p.get_int("ckpt", 50))
    off = [int(round((vaxes[a].o + d[a] * nbound) / d[a])) for a in range(3)]
    box = (m[0], m[1], m[2], off[0], off[1], off[2])
    mute = None
    if p.has("mute"):
        mv, maxes = sep.read(p.get_string("mute"))
        if mv.shape != (m[2], m[1], m[0]):
            raise sep.SepError("mute must be on the velocity grid")
        mute = mv.astype(np.float32)
    return SepProblem(w, vaxes, daxes, box, mute)


def make_operator(dev, sp: SepProblem):
    """The Born operator of the survey with its source fields prepared (it
    installs its own jt-phase sinc tables)."""
    from .operator import RtmZeroOp
    w = sp.workload
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    op = RtmZeroOp(dev, w, model_box=sp.model_box, mute=sp.mute, nt_data=w.nt)
    op.prepare()
    return op


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in ("fwd", "rtm", "inv"):
        print(__doc__, file=sys.stderr)
        return 2
    prog, p = argv[0], sep.Params(argv[1:])
    from .device import BornDevice
    from .operator import IOp, RegSolver, SimpleSolver, TMute, cgstep
    # inv: FP16 snapshots, the same 2 bytes per cell as BF16 and 6.5x closer
    # to FP64 (image 8.9e-5 against 5.8e-4 at 256^3, nt 4000)
    store = p.get_string("store", "f16" if prog == "inv" else "ckpt")
    t0 = time.perf_counter()
    sp = problem_from_sep(p, store)
    w = sp.workload
    _log(f"{prog}: compute grid {w.n}, dt {w.dt:.6g} s (jt {w.jt}), {len(w.shots)} shots, "
         f"{len(w.shots[0].rec_locs)} receivers, store {store}")
    dev = BornDevice(p.get_int("device", 0))
    try:
        op = make_operator(dev, sp)
        _log(f"source fields prepared in {time.perf_counter() - t0:.1f} s")
        mshape = op.model_shape
        dshape = op.data_shape
        if prog == "fwd":
            model, maxes = sep.read(p.get_string("image"))
            if model.shape != mshape:
                raise sep.SepError(f"image {model.shape} is not on the velocity grid {mshape}")
            mv = op.model_vector(model)
            dv = op.data_vector()
            op.forward(False, mv, dv)
            sep.write(p.get_string("dataout"), dv.numpy(), sp.data_axes)
        elif prog == "rtm":
            data, _ = sep.read(p.get_string("data"))
            if data.size != int(np.prod(dshape)):
                raise sep.SepError("data size does not match its axes")
            dv = op.data_vector(data)
            mv = op.model_vector()
            op.adjoint(False, mv, dv)
            sep.write(p.get_string("image"), mv.numpy(), sp.model_axes)
        else:
            data, _ = sep.read(p.get_string("data"))
            dv = op.data_vector(data)
            precond, history = None, p.get_int("history", 0)
            kind = p.get_string("precond", "none")
            if kind in ("diag", "learned"):
                from . import lsrtm
                mbox = sp.model_box
                if kind == "diag":
                    wv = lsrtm.illumination_weight(op, p.get_float("illum_eps", 0.01))
                    precond = lsrtm.diagonal_precond(dev, wv)
                else:
                    import os
                    from .unet import UNet
                    net = UNet(dev, batch=8, patch=64, seed=1)
                    net.load(p.get_string("weights", os.path.join(
                        os.path.dirname(os.path.abspath(__file__)), "weights",
                        "precond_split_e0.01.npy")))
                    W = lsrtm.illumination_weight(op, p.get_float("illum_eps", 0.01), power=0.5)
                    precond = net.precond(mbox[:3], overlap=8, split_weight=W)
                    history = history or 8
            wt = None
            if p.has("tmute"):
                # the data weight of the reference's commented-out tmute
                # (Born/RTM_fwd.cpp:50): tmute= t0, vmute= velocity
                at, arx = sp.data_axes[0], sp.data_axes[1]
                wt = TMute(dev, p.get_float("tmute"), p.get_float("vmute", 1500.0), nt=at.n,
                           ot=at.o, dt=at.d, nrx=arx.n, orx=arx.o, drx=arx.d,
                           nouter=op.nry * len(w.shots))
            eps = p.get_float("eps", 0.0)
            if eps > 0:
                # reg_solver with i_op: damped least squares
                solver = RegSolver(cgstep(), dv, op, IOp(dev, op.model_size), eps, wt=wt,
                                   precond=precond, verbose=True, history=history)
            else:
                solver = SimpleSolver(cgstep(), dv, op, precond=precond, verbose=True,
                                      history=history, wt=wt)
            res = solver.solve(p.get_int("niter", 5), rel_tol=p.get_float("rel_tol", 0.0))
            _log(f"{res.iterations} iterations, relative residual {res.rel_resid:.4e}, "
                 f"{res.seconds:.1f} s")
            sep.write(p.get_string("image"), solver.return_model().numpy(), sp.model_axes)
        _log(f"done in {time.perf_counter() - t0:.1f} s")
    finally:
        dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
