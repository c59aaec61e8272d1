"""LSRTM time-to-residual: plain CG against learned-preconditioned flexible CG.

    python -m sep_code_amd.lsrtm_bench --n 256 --nt 2500 [--diag-eps 0.01]
    python -m sep_code_amd.lsrtm_bench --n 256 --nt 2500 --op-timing 2   # operator only

All solvers run on the same operator (BF16 source snapshots, resident in
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

import numpy as np

from . import configs, lsrtm
from .device import BornDevice

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nt", type=int, default=2500)
    ap.add_argument("--seed", type=int, default=77)
    ap.add_argument("--niter", type=int, default=400)
    ap.add_argument("--tol", type=float, default=1e-2)
    ap.add_argument("--weights", default=os.path.join(HERE, "weights", "precond_split_e0.01.npy"),
                    help="trained network ('' to skip the learned run)")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--even", action="store_true",
                    help="use the whole network, not its odd part")
    ap.add_argument("--guard", default="0.8,0.05",
                    help="repeat_cos,min_cos of the split preconditioner's guard")
    ap.add_argument("--quiet", action="store_true",
                    help="no per-iteration progress lines on stderr")
    ap.add_argument("--history", type=int, default=4,
                    help="steps kept by the flexible CG of the learned run")
    ap.add_argument("--diag-eps", default="",
                    help="comma list: also run CG preconditioned by the illumination "
                         "diagonal (I + eps max I)^-1 for each eps")
    ap.add_argument("--split-eps", type=float, default=0.01,
                    help="> 0: the learned run is z = W P(W g), W = (I + eps max I)^-1/2 "
                         "(the network trained with the same split)")
    ap.add_argument("--op-timing", type=int, default=0,
                    help="> 0: only time that many Born forward / adjoint applications of "
                         "the operator (wall clock per shot-step) and exit")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731

    t0 = time.perf_counter()
    w = configs.lsrtm(seed=a.seed, n_side=a.n, nt=a.nt)
    dev = BornDevice(0)
    op, tprep = lsrtm.setup_operator(dev, w)
    if a.op_timing > 0:
        m, d = op.model_vector().random(1), op.data_vector()
        steps = len(w.shots) * a.nt
        res = {"workload": w.name, "grid": list(w.n), "nt": a.nt, "shots": len(w.shots),
               "store": w.store, "jimg": w.jimg, "prepare_s": tprep, "forward_us": [],
               "adjoint_us": []}
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
    out = {"workload": w.name, "grid": list(w.n), "nt": a.nt, "shots": len(w.shots),
           "store": w.store, "jimg": w.jimg, "tol": a.tol, "noise_rel": floor}
    if not a.skip_plain:
        r, _ = lsrtm.run_lsrtm(dev, w, op, d, niter=a.niter, rel_tol=a.tol, name="cg",
                               noise=e, verbose=not a.quiet)
        out["cg"] = {"iterations": r.iterations, "converged": r.converged,
                     "seconds": r.seconds, "judged": r.rel_resid_recomputed,
                     "hist": r.resid_hist[::5]}
        log(f"plain CG: {r.iterations} it, {r.seconds:.1f}s, judged {r.rel_resid_recomputed:.4e}")
    for eps in [float(v) for v in a.diag_eps.split(",") if v]:
        t0 = time.perf_counter()
        wv = lsrtm.illumination_weight(op, eps)
        tw = time.perf_counter() - t0
        r, _ = lsrtm.run_lsrtm(dev, w, op, d, niter=a.niter, rel_tol=a.tol, name="diag",
                               noise=e, precond=lsrtm.diagonal_precond(dev, wv),
                               verbose=not a.quiet)
        out[f"diag_{eps:g}"] = {"iterations": r.iterations, "converged": r.converged,
                                "seconds": r.seconds + tw, "judged": r.rel_resid_recomputed,
                                "hist": r.resid_hist[::5]}
        log(f"illumination-diagonal PCG eps={eps:g}: {r.iterations} it, {r.seconds + tw:.1f}s, "
            f"judged {r.rel_resid_recomputed:.4e}")
        wv.close()
    if a.weights:
        from .unet import UNet
        net = UNet(dev, batch=a.batch, patch=64, seed=1)
        net.load(a.weights)
        net.set_odd(not a.even)
        rep, mn = (float(v) for v in a.guard.split(","))
        net.set_guard(rep, mn)
        dims = lsrtm.model_box(w)[:3]
        t0 = time.perf_counter()
        W = lsrtm.illumination_weight(op, a.split_eps, power=0.5) if a.split_eps > 0 else None
        tw = time.perf_counter() - t0
        pre = net.precond(dims, overlap=8, split_weight=W)
        r, _ = lsrtm.run_lsrtm(dev, w, op, d, niter=a.niter, rel_tol=a.tol, name="fpcg",
                               noise=e, precond=pre, history=a.history, verbose=not a.quiet)
        out["fpcg"] = {"history": a.history, "split_eps": a.split_eps,
                       "iterations": r.iterations, "converged": r.converged,
                       "seconds": r.seconds + tw, "illumination_seconds": tw,
                       "judged": r.rel_resid_recomputed,
                       "fallbacks": r.fallbacks, "guard_diagonal_steps": net.guard_count(),
                       "odd": not a.even, "guard": a.guard, "hist": r.resid_hist[::5]}
        log(f"learned FPCG: {r.iterations} it, {r.seconds + tw:.1f}s, "
            f"judged {r.rel_resid_recomputed:.4e}, fallbacks {r.fallbacks}")
        if "cg" in out and out["cg"]["converged"] and r.converged:
            out["speedup"] = out["cg"]["seconds"] / out["fpcg"]["seconds"]
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
