"""Kernel-level timing driver for profiling (rocprofv3) at a fixed size.

    python -m sep_code_amd.profile_stencil --n 512 --reps 10

runs, on one context: a plain 16-byte streaming copy of --copy-gib GiB (the
calibration kernel whose FETCH_SIZE / WRITE_SIZE per byte scale the other
kernels' counters), a kernel with the plain step's read/write mix (three
n^3 FP32 streams read and summed into a fourth, with a plain and with a
nontemporal store), then --reps launches each of the stencil variants
  VL plain step, VL + fused imaging (FP32 snapshot), VL + fused Born source
  (BF16 snapshot), VL + fused imaging (BF16 snapshot), and the pass of two
  plain VL steps (vl_pair; its cells-per-lane and ring settings are fixed)
and prints their HIP-event mean durations and algorithmic bandwidths, for each
cells-per-lane setting in --xl (1: 64-wide tiles; 2: 128-wide tiles with
8-byte accesses; 4: 256-wide tiles with 16-byte accesses; names get an
"_x2" / "_x4" suffix), halo-ring setting in --ring (1 on, suffix "_ring";
0 off, suffix "_noring"; -1 the library's automatic choice, no suffix) and
load lookahead in --prefetch (1 or 2 slices, suffix "_pf1" / "_pf2"; 0
automatic, no suffix), streaming loads in --stream (1 on, suffix "_stream";
0 off, suffix "_nostream"; -1 automatic, no suffix).  With --ty, every workgroup target in --tb is timed
(suffix "_ty<rows>_tb<target>").
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from .device import BornDevice

# (name, form, act, fmt, algorithmic bytes per cell); form 2 is the pass of
# two plain steps (read p, p_old, V; write two fields), "gpts" counting both
VARIANTS = [("vl", 0, 0, 0, 16), ("vl_img_f32", 0, 1, 1, 28), ("vl_born_bf16", 0, 2, 2, 22),
            ("vl_img_bf16", 0, 1, 2, 26), ("vl_pair", 2, 0, 0, 20)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--copy-gib", type=float, default=1.0)
    ap.add_argument("--ty", type=int, default=0,
                    help="tile rows (0: the library's per-variant choice)")
    ap.add_argument("--tb", default="2048",
                    help="comma list of workgroup targets (used only with --ty)")
    ap.add_argument("--variants", default="all")
    ap.add_argument("--xl", default="1,2")
    ap.add_argument("--ring", default="-1", help="comma list of halo-ring settings (-1, 0, 1)")
    ap.add_argument("--stream", default="-1",
                    help="comma list of streaming-load settings of the ring step (-1 auto, 0, 1)")
    ap.add_argument("--prefetch", default="0",
                    help="comma list of load lookaheads of the plain ring step (0 auto, 1, 2)")
    a = ap.parse_args()
    n = a.n
    dev = BornDevice(0)
    tbs = [int(v) for v in a.tb.split(",")] if a.ty else [2048]
    dev.set_tuning(a.ty, tbs[0])
    dev.create_space((12.5, 12.5, 12.5), 32.0, 0.001, (n, n, n))
    dev.set_velocity(np.full((n, n, n), (3000 * 0.001) ** 2, np.float32))
    out = {"n": n, "reps": a.reps}
    nbytes = a.copy_gib * 2 ** 30
    ms = dev.copy_bench(nbytes, a.reps)
    out["copy"] = {"bytes": nbytes, "ms": ms, "GBps": 2 * nbytes / ms / 1e6}
    # the plain step's mix without its arithmetic: three n^3 FP32 streams
    # read, one written (plain and nontemporal store)
    fb = 4.0 * n ** 3
    for nt in (0, 1):
        ms = dev.stream_bench(3, fb, nt_store=bool(nt), nrep=a.reps)
        out[f"read3_write1{'_ntstore' if nt else ''}"] = {"bytes": 4 * fb, "ms": ms,
                                                          "GBps": 4 * fb / ms / 1e6}
    for ring in [int(v) for v in a.ring.split(",")]:
        dev.set_halo_ring(ring)
        for xl in [int(v) for v in a.xl.split(",")]:
            dev.set_x_per_lane(xl)
            for pf, sl in [(int(v), int(u)) for v in a.prefetch.split(",")
                           for u in a.stream.split(",")]:
                dev.set_prefetch(pf)
                dev.set_stream_loads(sl)
                for tb in tbs:
                    dev.set_tuning(a.ty, tb)
                    for name, form, act, fmt, B in VARIANTS:
                        if a.variants != "all" and name not in a.variants.split(","):
                            continue
                        ms = dev.time_step(form, act, fmt, a.reps)
                        key = (name + (f"_x{xl}" if xl != 1 else "") +
                               {1: "_ring", 0: "_noring", -1: ""}[ring] +
                               (f"_pf{pf}" if pf else "") +
                               ({-1: "", 0: "_nostream", 1: "_stream"}[sl]) +
                               (f"_ty{a.ty}_tb{tb}" if a.ty else ""))
                        out[key] = {"ms": ms, "alg_GBps": B * n ** 3 / ms / 1e6,
                                    "gpts": (2 if form == 2 else 1) * n ** 3 / ms / 1e6}
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
