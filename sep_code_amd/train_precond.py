"""Train the learned image-domain preconditioner.

Training pairs are made by this library on seeded random models (layered,
dipping and Gaussian-anomaly velocities with sparse reflectivity, the
acquisition of the LSRTM workload: 8 shots on a 2 x 4 surface grid):
  input  = RTM image       L^T d
  target = LSRTM image     20 iterations of cgstep on |L m - d|
each normalised by its own RMS (the CG plane search fixes the step scale, so
only the direction of P(g) matters).  The network is trained with its own
forward/backward kernels and Adam on random 64^3 patches with x/y flips.

    python -m sep_code_amd.train_precond --pairs 96 --out weights.npy
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs, lsrtm
from .device import BornDevice
from .operator import DeviceVector
from .unet import UNet


def make_pair(dev, seed, n_side=128, nt=1300, cg_iters=20, noise=0.01):
    w = configs.lsrtm(seed=seed, n_side=n_side, nt=nt, anomalies=int(seed % 4))
    op, _ = lsrtm.setup_operator(dev, w)
    d, _ = lsrtm.observed_data(op, w, noise_frac=noise, seed=seed + 99)
    rtm = op.model_vector()
    op.adjoint(False, rtm, d)
    res, m = lsrtm.run_lsrtm(dev, w, op, d, niter=cg_iters, rel_tol=0.0)
    x, y = rtm.numpy(), m.numpy()
    dims = lsrtm.model_box(w)[:3]
    op.close()
    return x, y, dims, res.rel_resid


def normalise(a):
    r = float(np.sqrt((a.astype(np.float64) ** 2).mean()))
    return (a / r).astype(np.float32) if r > 0 else a


def train(dev, pairs, steps=2000, batch=4, lr=1e-3, seed=0, log=print):
    net = UNet(dev, batch=batch, patch=64, seed=seed + 1)
    rng = np.random.default_rng(seed)
    vols = []
    for x, y, dims in pairs:
        vols.append((DeviceVector.from_array(dev, normalise(x)),
                     DeviceVector.from_array(dev, normalise(y)), dims))
    P = 64
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        cur_lr = lr * (0.5 if it > 0.6 * steps else 1.0) * (0.5 if it > 0.85 * steps else 1.0)
        for k in range(batch):
            xv, yv, dims = vols[rng.integers(len(vols))]
            o = [int(rng.integers(-8, max(-7, dims[a] - P + 9))) for a in range(3)]
            net.extract(xv, dims, o, 1.0, "x", k)
            net.extract(yv, dims, o, 1.0, "t", k)
        loss = net.train_step_device(batch, cur_lr)
        hist.append(loss)
        if it % 100 == 0 or it == steps - 1:
            log(f"step {it} loss {np.mean(hist[-100:]):.5f} ({time.perf_counter() - t0:.1f}s)")
    return net, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=96)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--nt", type=int, default=1300)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--out", default="precond_unet.npy")
    ap.add_argument("--save-pairs", default="")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    dev = BornDevice(0)
    pairs = []
    t0 = time.perf_counter()
    for i in range(a.pairs):
        x, y, dims, rr = make_pair(dev, a.seed0 + i, n_side=a.n, nt=a.nt)
        pairs.append((x, y, dims))
        log(f"pair {i + 1}/{a.pairs} residual after 20 CG {rr:.3f} "
            f"({time.perf_counter() - t0:.1f}s)")
    if a.save_pairs:
        np.savez(a.save_pairs, **{f"x{i}": p[0] for i, p in enumerate(pairs)},
                 **{f"y{i}": p[1] for i, p in enumerate(pairs)})
    tgen = time.perf_counter() - t0
    net, hist = train(dev, pairs, steps=a.steps, log=log)
    net.save(a.out)
    print(json.dumps({"pairs": a.pairs, "gen_seconds": tgen, "steps": a.steps,
                      "loss_first100": float(np.mean(hist[:100])),
                      "loss_last100": float(np.mean(hist[-100:]))}), flush=True)


if __name__ == "__main__":
    main()
