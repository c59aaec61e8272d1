"""Image error of each source-wavefield store against the FP64 oracle, on the
precision variant of the 512^3 production shot: the same physics (layered
1500-4500 m/s plus three Gaussian anomalies, d = 12.5 m, Ricker 15 Hz) on a
256^3 grid, one shot, nt = 2000, imaging every 4 steps.

Run in two stages, so the GPU does not wait for the FP64 oracle (about a
quarter of an hour on 8 CPU cores):

    python tests/precision_256.py oracle ORACLE.npy      # CPU: FP64 image
    python tests/precision_256.py gpu ORACLE.npy         # GPU: every store vs it

The receiver data are seeded synthetic traces (Ricker pulses at random
times and amplitudes), generated identically by both stages, so nothing but
the image crosses between them.  The GPU stage prints one JSON line with the
relative L2 error of the f32 / bf16 / f16 / ckpt images.  This file is not a
pytest module (no test_ prefix); the small-grid versions of the same checks
are tests/test_gpu_born.py::test_compressed_store_image_error.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from sep_code_amd import configs  # noqa: E402

JIMG = 4


def workload(n_side, nt):
    return configs.rtm512(n_side=n_side, nt=nt, rec_step=2, store="f32", jimg=JIMG)


def synthetic_data(nrec, ntr, dt, seed=7, npulse=6, f0=15.0):
    """[nrec][ntr] traces: npulse Ricker pulses per trace."""
    rng = configs.SplitMix64(seed)
    tmax = (ntr - 8) * dt
    t0 = 0.1 + rng.uniform_array(nrec * npulse).reshape(nrec, npulse) * (tmax - 0.2)
    amp = 2.0 * rng.uniform_array(nrec * npulse).reshape(nrec, npulse) - 1.0
    t = np.arange(ntr) * dt
    out = np.zeros((nrec, ntr))
    for k in range(npulse):
        tau = np.pi * f0 * (t[None, :] - t0[:, k:k + 1])
        out += amp[:, k:k + 1] * (1.0 - 2.0 * tau ** 2) * np.exp(-tau ** 2)
    return out


def stage_oracle(a):
    from oracle import born_oracle as O
    w = workload(a.n, a.nt)
    s = w.shots[0]
    D = synthetic_data(len(s.rec_locs), w.ntr, w.dt)
    sinc = O.sinc_table(1)
    P = O.Problem(w.n, w.d, w.V.astype(np.float64), w.bc_a, w.bc_b, src_locs=s.src_locs,
                  src_tr=s.src_traces.astype(np.float64), sinc_s=sinc, rec_locs=s.rec_locs,
                  ntr_d=w.ntr, sinc_d=sinc, jt=1, nt=w.nt, jimg=JIMG, cimg=w.cimg)
    t0 = time.perf_counter()
    img = P.rtm_adjoint(D, None, ckpt=int(np.ceil(np.sqrt(w.nt))))
    np.save(a.path, img.astype(np.float32).reshape(w.n[2], w.n[1], w.n[0]))
    print(json.dumps({"stage": "oracle", "grid": list(w.n), "nt": w.nt,
                      "seconds": time.perf_counter() - t0,
                      "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}))


def stage_gpu(a):
    from sep_code_amd.device import BornDevice
    w = workload(a.n, a.nt)
    s = w.shots[0]
    ref = np.load(a.path).ravel()
    D = synthetic_data(len(s.rec_locs), w.ntr, w.dt).astype(np.float32)
    sinc = np.zeros((1, 8), np.float32)
    sinc[0, 3] = 1
    dev = BornDevice(0)
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    dev.set_sinc_table(0, sinc)
    dev.set_sinc_table(1, sinc)
    dev.set_source(s.src_locs, s.src_traces)
    dev.set_receivers(s.rec_locs, D, ntr=w.ntr, nx=s.nrx, ny=s.nry)
    out = {"grid": list(w.n), "nt": w.nt, "jimg": JIMG, "errors": {}, "vs_f32_store": {}}
    f32img = None
    for store in ("f32", "bf16", "f16", "ckpt"):
        dev.source_prop(1, w.nt, JIMG, store=store, ckpt=50)
        dev.zero_image()
        dev.rtm_adjoint(1, w.nt, JIMG)
        img = dev.get_image().astype(np.float64).ravel()
        err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
        out["errors"][store] = err
        # the store's own error: against the same FP32 arithmetic with FP32 snapshots
        if f32img is None:
            f32img = img
        out["vs_f32_store"][store] = float(np.linalg.norm(img - f32img) / np.linalg.norm(f32img))
        print(f"{store}: relative L2 image error {err:.3e} (vs FP32 store "
              f"{out['vs_f32_store'][store]:.3e})", file=sys.stderr, flush=True)
        dev.release_store(0)
    print(json.dumps(out), flush=True)
    dev.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("stage", choices=("oracle", "gpu"))
    ap.add_argument("path")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nt", type=int, default=2000)
    a = ap.parse_args()
    (stage_oracle if a.stage == "oracle" else stage_gpu)(a)


if __name__ == "__main__":
    main()
