"""LSRTM time-to-residual: plain CG against learned-preconditioned flexible CG.

    python -m sep_code_amd.lsrtm_bench --n 256 --nt 2500 [--diag-eps 0.01]
    python -m sep_code_amd.lsrtm_bench --n 256 --nt 2500 --op-timing 2   # operator only

All solvers run on the same operator (FP16 source snapshots, resident in
HBM before the clock starts), from m = 0, until the misfit to the noise-free
data |L m - d_clean| / |d_clean| <= --tol (see lsrtm.py), and are judged by a
fresh application of the plain FP32 operator with FP64 norms:
  * plain CG (cgstep), the baseline;
  * optionally CG preconditioned by the illumination diagonal alone
    (--diag-eps), to show what the network adds to it;
  * flexible CG (--history steps) with the learned preconditioner
    z = W P(W g), W = (I + eps max I)^-1/2 from the source illumination and
    P the trained U-Net (shipped weights by default).
Training is excluded from the clock; preconditioner inference (and the
illumination, computed once from the stored snapshots) is included.  Prints
one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import _lib, configs, lsrtm
from .device import BornDevice

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "precond_split_e0.01.npy")


def _entry(r, extra=None):
    """The JSON record of one solver run: iterations and seconds to the
    tolerance on the judged (noise-free) misfit, and both residual
    definitions at the stop, running and recomputed with the plain FP32
    operator and FP64 norms."""
    out = {"iterations": r.iterations, "converged": r.converged, "seconds": r.seconds,
           "judged": r.rel_resid_recomputed, "data_rel_resid": r.data_rel_resid_recomputed,
           "hist": r.resid_hist[::5], "data_hist": r.data_resid_hist[::5]}
    tol_hit = [i for i, x in enumerate(r.data_resid_hist) if x <= 1e-2]
    out["data_resid_to_1e-2"] = (
        {"iterations": tol_hit[0], "seconds": r.time_hist[tol_hit[0]]} if tol_hit else None)
    out.update(extra or {})
    return out


def compare(dev, w, op, d, e, floor, *, niter=400, tol=1e-2, weights=DEFAULT_WEIGHTS, batch=8,
            even=False, guard="0.8,0.05", history=8, diag_eps="", split_eps=0.01,
            skip_plain=False, quiet=True, log=None, guard_mode=1):
    """Plain CG, optionally illumination-diagonal PCG for each eps in the
    comma list diag_eps, and learned flexible CG with the network in
    `weights` ('' skips it), each from m = 0 on the prepared operator op
    with data d = L m_true + e.  Returns the comparison as a dict."""
    log = log or (lambda m: None)
    store = {v: k for k, v in _lib.STORE_NAMES.items()}[op.desc.store]
    out = {"workload": w.name, "grid": list(w.n), "nt": w.nt, "shots": len(w.shots),
           "store": store, "jimg": w.jimg, "tol": tol, "noise_rel": floor,
           "stop_rule": "misfit to the noise-free data |L m - (d - e)| / |d - e| <= tol; "
                        "data_rel_resid is |L m - d| / |d| against the noisy data"}
    if not skip_plain:
        r, m = lsrtm.run_lsrtm(dev, w, op, d, niter=niter, rel_tol=tol, name="cg", noise=e,
                               verbose=not quiet)
        m.close()
        out["cg"] = _entry(r)
        log(f"plain CG: {r.iterations} it, {r.seconds:.1f}s, judged "
            f"{r.rel_resid_recomputed:.4e}, |Lm-d|/|d| {r.data_rel_resid_recomputed:.4e}")
    for eps in [float(v) for v in diag_eps.split(",") if v]:
        t0 = time.perf_counter()
        wv = lsrtm.illumination_weight(op, eps)
        tw = time.perf_counter() - t0
        r, m = lsrtm.run_lsrtm(dev, w, op, d, niter=niter, rel_tol=tol, name="diag", noise=e,
                               precond=lsrtm.diagonal_precond(dev, wv), verbose=not quiet)
        m.close()
        out[f"diag_{eps:g}"] = _entry(r, {"seconds": r.seconds + tw})
        log(f"illumination-diagonal PCG eps={eps:g}: {r.iterations} it, {r.seconds + tw:.1f}s, "
            f"judged {r.rel_resid_recomputed:.4e}")
        wv.close()
    if weights:
        from .unet import UNet
        net = UNet(dev, batch=batch, patch=64, seed=1)
        net.load(weights)
        net.set_odd(not even)
        rep, mn = (float(v) for v in guard.split(","))
        net.set_guard(rep, mn)
        net.set_guard_mode(guard_mode)
        dims = lsrtm.model_box(w)[:3]
        t0 = time.perf_counter()
        W = lsrtm.illumination_weight(op, split_eps, power=0.5) if split_eps > 0 else None
        tw = time.perf_counter() - t0
        pre = net.precond(dims, overlap=8, split_weight=W)
        r, m = lsrtm.run_lsrtm(dev, w, op, d, niter=niter, rel_tol=tol, name="fpcg", noise=e,
                               precond=pre, history=history, verbose=not quiet)
        m.close()
        out["fpcg"] = _entry(r, {"history": history, "split_eps": split_eps,
                                 "seconds": r.seconds + tw, "illumination_seconds": tw,
                                 "fallbacks": r.fallbacks,
                                 "guard_diagonal_steps": net.guard_count(), "odd": not even,
                                 "guard": guard, "guard_mode": guard_mode})
        log(f"learned FPCG: {r.iterations} it, {r.seconds + tw:.1f}s, "
            f"judged {r.rel_resid_recomputed:.4e}, fallbacks {r.fallbacks}, "
            f"guard {net.guard_count()}")
        net.close()
        if W is not None:
            W.close()
        if "cg" in out and out["cg"]["converged"] and r.converged:
            out["speedup"] = out["cg"]["seconds"] / out["fpcg"]["seconds"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nt", type=int, default=2500)
    ap.add_argument("--seed", type=int, default=77)
    ap.add_argument("--niter", type=int, default=400)
    ap.add_argument("--tol", type=float, default=1e-2)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS,
                    help="trained network ('' to skip the learned run)")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--even", action="store_true",
                    help="use the whole network, not its odd part")
    ap.add_argument("--guard-mode", type=int, default=1,
                    help="1: the diagonal's direction plus the new part of the network's "
                         "correction; 0: a repeating network direction is replaced by the "
                         "diagonal's")
    ap.add_argument("--guard", default="0.8,0.05",
                    help="repeat_cos,min_cos of the split preconditioner's guard")
    ap.add_argument("--quiet", action="store_true",
                    help="no per-iteration progress lines on stderr")
    ap.add_argument("--history", type=int, default=8,
                    help="steps kept by the flexible CG of the learned run (each holds "
                         "one data-space vector: 1.31 GB at 256^3)")
    ap.add_argument("--diag-eps", default="",
                    help="comma list: also run CG preconditioned by the illumination "
                         "diagonal (I + eps max I)^-1 for each eps")
    ap.add_argument("--split-eps", type=float, default=0.01,
                    help="> 0: the learned run is z = W P(W g), W = (I + eps max I)^-1/2 "
                         "(the network trained with the same split)")
    ap.add_argument("--store", default="",
                    help="source-wavefield store of the operator (default: the workload's)")
    ap.add_argument("--op-timing", type=int, default=0,
                    help="> 0: only time that many Born forward / adjoint applications of "
                         "the operator (wall clock per shot-step) and exit")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731

    t0 = time.perf_counter()
    w = configs.lsrtm(seed=a.seed, n_side=a.n, nt=a.nt)
    dev = BornDevice(0)
    op, tprep = lsrtm.setup_operator(dev, w, store=a.store or None)
    if a.op_timing > 0:
        m, d = op.model_vector().random(1), op.data_vector()
        steps = len(w.shots) * a.nt
        res = {"workload": w.name, "grid": list(w.n), "nt": a.nt, "shots": len(w.shots),
               "store": a.store or w.store, "jimg": w.jimg, "prepare_s": tprep,
               "forward_us": [], "adjoint_us": []}
        for _ in range(a.op_timing):
            dev.sync()
            t0 = time.perf_counter()
            op.forward(False, m, d)
            dev.sync()
            t1 = time.perf_counter()
            op.adjoint(False, m, d)
            dev.sync()
            t2 = time.perf_counter()
            res["forward_us"].append((t1 - t0) / steps * 1e6)
            res["adjoint_us"].append((t2 - t1) / steps * 1e6)
        print(json.dumps(res), flush=True)
        dev.close()
        return
    d, floor, e = lsrtm.observed_data(op, w, keep_noise=True)
    log(f"{w.name}: {len(w.shots)} shots, nt={a.nt}, prepared in {tprep:.1f}s, "
        f"noise |e|/|d| = {floor:.4f}, setup {time.perf_counter() - t0:.1f}s")
    out = compare(dev, w, op, d, e, floor, niter=a.niter, tol=a.tol, weights=a.weights,
                  batch=a.batch, even=a.even, guard=a.guard, history=a.history,
                  diag_eps=a.diag_eps, split_eps=a.split_eps, skip_plain=a.skip_plain,
                  quiet=a.quiet, log=log, guard_mode=a.guard_mode)
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
