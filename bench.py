"""Benchmark of the Born RTM hot path on one MI355X.

A step is one full RTM shot of the 512^3 production workload (configs.rtm512:
512^3 FP32 grid, d = 12.5 m, nt = 4000): source propagation with FP32
checkpoints every 50 steps, then the adjoint (receiver-field) propagation
with segment recompute of the source field and fused zero-lag imaging.
Inputs (velocity, source, Born-modelled receiver data) are resident in HBM
before the timed region.

value = RTM Gpts/s = 2 * n1*n2*n3 * nt / time per shot (the two wavefields
an RTM must advance; the checkpoint recompute is overhead, not credited).

roofline: the dominant kernel is the fused VL stencil step (k_step_x4r at
512^3: four cells per lane on the six-plane halo ring, streaming loads of
p_old and V, no imaging; it runs
for the source propagation and the recompute).  Its launches inside the timed region are bracketed by
HIP events on the library stream; achieved = 16 B per cell (read p_n,
p_{n-1}, V; write p_{n+1}) * n1*n2*n3 / mean launch duration.  traffic is
the HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
(scaled by a 1 GiB copy's counters) for the default configuration.

cpu_baseline: the FP32 OpenMP CPU restatement (oracle/liborc32.so) of the
same stencil step at 512^3, timed on the host cores: the median of three
samples of ~3.3 s each.  With --workload rtm128 (128^3, nt = 1000) the
whole shot runs on the CPU instead, three times, and the line also carries
the relative difference of the CPU image from the GPU image.

lsrtm: the second half of the metric, on the LSRTM workload (lsrtm256:
256^3, 8 shots, nt = 2500, FP16 source snapshots every 4 steps, data with
1 % Gaussian noise).  The default run adds --lsrtm-iters iterations each of
plain CG (cgstep) and of learned flexible CG (shipped network, guard and
illumination inside the clock): seconds per iteration and both residuals
after them, |L m - d| / |d| against the noisy data and the misfit to the
noise-free data.  --workload lsrtm256 runs plain CG, the illumination
diagonal and learned flexible CG to the 1e-2 tolerance instead and reports
the time to it (value = learned seconds, lower is better).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 16.0
# HBM bytes per launch of the plain VL step (k_step_x4r, 512^3) from rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel calibration (2.0 / 1.0):
# profiles/stencil512_traffic.md.  Used for the default configuration only.
TRAFFIC_X4R_512 = (2.0 * 794832 + 508032) * 1024.0
# The library's ring step reads the data it uses once with nontemporal loads
# when a step moves more than this many bytes (16 per cell).
STREAM_MIN_STEP_BYTES = 512 * 2 ** 20
# ... and launches its propagation loops as graphs on grids of at most this
# many cells (brtm_set_graphs, automatic).
GRAPH_MAX_CELLS = 1 << 22


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_device(w, ty, target, xl, graphs=-1):
    from sep_code_amd.device import BornDevice
    dev = BornDevice(0)
    dev.set_tuning(ty, target)
    dev.set_x_per_lane(xl)
    dev.set_graphs(graphs)
    shot = w.shots[0]
    sinc = np.zeros((1, 8), np.float32)
    sinc[0, 3] = 1
    dev.create_space(w.d, w.bc_a, w.bc_b, w.n)
    dev.set_velocity(w.V)
    dev.set_imaging_scale(w.cimg)
    dev.set_sinc_table(0, sinc)
    dev.set_sinc_table(1, sinc)
    dev.set_source(shot.src_locs, shot.src_traces)
    dev.set_receivers(shot.rec_locs, None, ntr=w.ntr, nx=shot.nrx, ny=shot.nry)
    # observed data: Born modelling of the true reflectivity (source field
    # propagated alongside), left resident in the receiver buffer.
    dev.select_shot(0)
    dev.source_prop(w.jt, w.nt, w.jimg, store="none")
    dev.set_model(w.refl)
    dev.rtm_forward(w.jt, w.nt, w.jimg)
    dev.sync()
    return dev


def rtm_shot(dev, w):
    dev.source_prop(w.jt, w.nt, w.jimg, store=w.store, ckpt=w.ckpt)
    dev.zero_image()
    dev.rtm_adjoint(w.jt, w.nt, w.jimg)


def cells_per_lane(n1, want):
    """The forward-form kernel the library runs for this n1: cells per lane
    asked for (1, 2, 4) where n1 allows it, or automatically (0) the widest
    whose tiles keep at least 3/4 of the lanes busy."""
    if want == 0:
        want = 1
        for k in (4, 2):
            w = 64 * k
            if n1 % k == 0 and 4 * n1 >= 3 * (-(-n1 // w) * w):
                want = k
                break
    if want == 4 and n1 % 4 == 0:
        return 4
    if want >= 2 and n1 % 2 == 0:
        return 2
    return 1


def plain_step_kernel(n, want):
    """(cells per lane, tile rows) of the plain forward step the library
    runs on grid n = (n1, n2, n3) with automatic tiles: the default shape,
    or 64-wide 8-row tiles when that shape would leave some of MI355X's 256
    compute units without a workgroup."""
    xl = cells_per_lane(n[0], want)
    ty, tb = {4: (16, 256), 2: (16, 2048), 1: (32, 1024)}[xl]
    if want == 0:
        tiles = -(-n[0] // (64 * xl)) * -(-n[1] // ty)
        nz = n[2] - 8
        zc = min(max(-(-tiles * nz // tb), 8), nz)
        if tiles * -(-nz // zc) < 256:
            return 1, 8
    return xl, ty


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads():
    """Host threads for the CPU restatement; set before libgomp loads."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    return threads


def cpu_baseline(w, seconds=10.0, runs=3):
    """FP32 OpenMP restatement of the VL step at the workload's size: the
    median per-step rate of `runs` timed samples of seconds/runs each."""
    threads = _omp_threads()
    from oracle import born_oracle as O  # loads libgomp after the settings above
    P = O.Problem(w.n, w.d, w.V, w.bc_a, w.bc_b, prec=32)
    rng = np.random.default_rng(0)
    cur = np.zeros(w.ncell, np.float32)
    old = np.zeros(w.ncell, np.float32)
    v = cur.reshape(w.n[2], w.n[1], w.n[0])
    v[4:-4, 4:-4, 4:-4] = rng.standard_normal((w.n[2] - 8, w.n[1] - 8, w.n[0] - 8))
    import ctypes as C
    P.L.orc_flush_denormals(1)
    try:
        P.L.orc_propagate(C.byref(P.s), 0, 1, cur.ctypes.data, old.ctypes.data)  # warm-up
        rates, total_steps, total_s = [], 0, 0.0
        for _ in range(runs):
            steps, t0 = 0, time.perf_counter()
            while True:
                P.L.orc_propagate(C.byref(P.s), 0, 1, cur.ctypes.data, old.ctypes.data)
                steps += 1
                el = time.perf_counter() - t0
                if (el >= seconds / runs and steps >= 2) or steps >= 1000:
                    break
            rates.append(w.ncell * steps / el / 1e9)
            total_steps += steps
            total_s += el
    finally:
        P.L.orc_flush_denormals(0)
    gpts = float(np.median(rates))
    # A checkpointed shot advances three fields per time step (source,
    # recompute, receiver) and credits two; imaging only adds to that.
    props = 3 if w.store == "ckpt" else 2
    shot_s = props * w.ncell * w.nt / (gpts * 1e9)
    return {"value": round(gpts, 4), "unit": "Gpts/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"median of {runs} samples ({total_steps} FP32 VL stencil steps with "
                      f"fused sponge in all, {total_s:.1f} s) of the {w.n[0]}x{w.n[1]}x"
                      f"{w.n[2]} grid on {threads} OpenMP threads (OMP_PROC_BIND="
                      f"{os.environ['OMP_PROC_BIND']}, OMP_PLACES={os.environ['OMP_PLACES']}, denormals "
                      f"flushed); per-step rate, not the full {w.nt}-step shot",
            "extrapolated_shot": {
                "seconds": round(shot_s, 1),
                "rtm_gpts": round(2.0 * w.ncell * w.nt / shot_s / 1e9, 4),
                "basis": f"extrapolated, not measured: {props} stencil steps per time step "
                         f"at the sampled rate over nt = {w.nt}, imaging not counted (a lower "
                         "bound on the CPU shot time)"}}


def cpu_full_shot(w, D, img_gpu, runs=3):
    """The whole RTM shot on the FP32 OpenMP restatement, the workload's own
    algorithm: FP32 snapshots of every imaging step, or checkpoints with
    segment recompute.  Same source, receivers and Born data as the GPU run;
    reports the median RTM Gpts/s of `runs` shots and the relative
    difference of the CPU image from the GPU image."""
    threads = _omp_threads()
    from oracle import born_oracle as O  # loads libgomp after the settings above
    shot = w.shots[0]
    P = O.Problem(w.n, w.d, w.V, w.bc_a, w.bc_b, src_locs=shot.src_locs,
                  src_tr=shot.src_traces, rec_locs=shot.rec_locs, ntr_d=w.ntr, jt=w.jt,
                  nt=w.nt, jimg=w.jimg, cimg=w.cimg, prec=32)
    D = np.ascontiguousarray(D, np.float32).reshape(len(shot.rec_locs), w.ntr)
    times, img = [], None
    P.L.orc_flush_denormals(1)
    try:
        for _ in range(runs):
            t0 = time.perf_counter()
            if w.store == "f32":
                snaps, _ = P.source_prop(True)
                img = P.rtm_adjoint(D, snaps)
                del snaps
            else:
                img = P.rtm_adjoint(D, None, ckpt=w.ckpt)
            times.append(time.perf_counter() - t0)
    finally:
        P.L.orc_flush_denormals(0)
    el = float(np.median(times))
    g = np.asarray(img_gpu, np.float64).ravel()
    diff = float(np.linalg.norm(np.asarray(img, np.float64) - g) / max(np.linalg.norm(g), 1e-300))
    algo = ("FP32 snapshots of every imaging step" if w.store == "f32"
            else f"checkpoints every {w.ckpt} steps and segment recompute")
    return {"value": round(2.0 * w.ncell * w.nt / el / 1e9, 4), "unit": "Gpts/s",
            "cores": threads, "kind": "port", "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"the full {w.nt}-step RTM shot of the {w.n[0]}x{w.n[1]}x{w.n[2]} "
                      f"grid ({algo}), median of {runs} shots ({el:.2f} s each) on {threads} "
                      f"OpenMP threads (OMP_PROC_BIND={os.environ['OMP_PROC_BIND']}, "
                      f"OMP_PLACES={os.environ['OMP_PLACES']}, denormals flushed); same metric as "
                      "value",
            "image_rel_diff_vs_gpu": float(f"{diff:.3e}")}


LSRTM_WORKLOAD = ("lsrtm256: LSRTM, 256x256x256 FP32, 8 shots, nt=2500, FP16 source "
                  "snapshots every 4 steps, Born data + 1 % Gaussian noise")


def _lsrtm_setup(log, graphs=-1):
    from sep_code_amd import configs, lsrtm
    from sep_code_amd.device import BornDevice
    t0 = time.perf_counter()
    w = configs.lsrtm()
    dev = BornDevice(0)
    dev.set_graphs(graphs)
    op, tprep = lsrtm.setup_operator(dev, w)
    d, floor, e = lsrtm.observed_data(op, w, keep_noise=True)
    log(f"lsrtm256 set up in {time.perf_counter() - t0:.1f}s (source fields {tprep:.1f}s), "
        f"noise |e|/|d| = {floor:.4f}")
    return w, dev, op, d, floor, e


def lsrtm_slice(iters, log):
    """`iters` iterations each of plain CG and learned flexible CG on the
    LSRTM workload from m = 0: seconds per iteration (the solver loop's own
    clock, preconditioner inference and guard included; the illumination
    weight, computed once, reported beside it) and both residuals after
    `iters` iterations, recomputed with the plain FP32 operator and FP64
    norms."""
    from sep_code_amd import lsrtm
    from sep_code_amd.lsrtm_bench import DEFAULT_WEIGHTS
    from sep_code_amd.unet import UNet
    w, dev, op, d, floor, e = _lsrtm_setup(log)

    def rec(r, extra=None):
        out = {"iterations": r.iterations, "s_per_iter": round(r.seconds / r.iterations, 4),
               "data_rel_resid": round(r.data_rel_resid_recomputed, 6),
               "misfit_noise_free": round(r.rel_resid_recomputed, 6)}
        out.update(extra or {})
        return out

    res = {"workload": LSRTM_WORKLOAD, "noise_rel": round(floor, 6),
           "residuals": "data_rel_resid = |L m - d| / |d| (noisy data as given); "
                        "misfit_noise_free = |L m - (d - e)| / |d - e|",
           "time_to_tolerance": "python bench.py --workload lsrtm256"}
    r, m = lsrtm.run_lsrtm(dev, w, op, d, niter=iters, rel_tol=0.0, name="cg", noise=e)
    m.close()
    res["cg"] = rec(r)
    log(f"lsrtm plain CG: {iters} it, {r.seconds / iters:.3f} s/it, |Lm-d|/|d| "
        f"{r.data_rel_resid_recomputed:.4f}, noise-free misfit {r.rel_resid_recomputed:.4f}")
    net = UNet(dev, batch=8, patch=64, seed=1)
    net.load(DEFAULT_WEIGHTS)
    net.set_odd(True)
    net.set_guard(0.8, 0.05)
    net.set_guard_mode(1)
    t0 = time.perf_counter()
    W = lsrtm.illumination_weight(op, 0.01, power=0.5)
    tw = time.perf_counter() - t0
    pre = net.precond(lsrtm.model_box(w)[:3], overlap=8, split_weight=W)
    r, m = lsrtm.run_lsrtm(dev, w, op, d, niter=iters, rel_tol=0.0, name="fpcg", noise=e,
                           precond=pre, history=8)
    m.close()
    res["fpcg"] = rec(r, {"illumination_s": round(tw, 3), "history": 8, "guard_mode": 1,
                          "guard_diagonal_steps": net.guard_count()})
    log(f"lsrtm learned FCG: {iters} it, {r.seconds / iters:.3f} s/it, |Lm-d|/|d| "
        f"{r.data_rel_resid_recomputed:.4f}, noise-free misfit {r.rel_resid_recomputed:.4f}")
    net.close()
    W.close()
    dev.close()
    return res


def lsrtm_main(a):
    """--workload lsrtm256: plain CG, the illumination diagonal and learned
    flexible CG to the tolerance; the line's value is the learned run's
    time to it.  The roofline object covers the FP16 imaging step, the
    adjoint's dominant kernel (26 algorithmic bytes per cell)."""
    from sep_code_amd.lsrtm_bench import DEFAULT_WEIGHTS, compare
    a.lsrtm_weights = a.lsrtm_weights or DEFAULT_WEIGHTS
    w, dev, op, d, floor, e = _lsrtm_setup(log, a.graphs)
    t0 = time.perf_counter()
    solvers = set(a.lsrtm_solvers.split(","))
    out = compare(dev, w, op, d, e, floor, niter=a.lsrtm_niter, tol=1e-2,
                  diag_eps="0.01" if "diag" in solvers else "",
                  weights=a.lsrtm_weights if "fpcg" in solvers else "",
                  skip_plain="cg" not in solvers, quiet=False, log=log)
    el = time.perf_counter() - t0
    # the imaging step alone, 200 launches between HIP events on the library
    # stream (the solves launch it millions of times)
    kcount = 200
    k_ms = dev.time_step(0, 1, 3, kcount)
    achieved = 26.0 * w.ncell / (k_ms * 1e-3) / 1e9
    fp = out.get("fpcg", {})
    cpu = None
    if not a.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(w, a.cpu_seconds)
        cpu.pop("extrapolated_shot", None)
        cpu["sample"] += ("; one CG iteration advances 2 fields per shot (the source "
                          "field is stored: the scattered field in Born modelling, then the "
                          "receiver field in the adjoint), 8 shots")
        it_s = 2 * len(w.shots) * w.ncell * w.nt / (cpu["value"] * 1e9)
        cpu["extrapolated_cg_iteration_s"] = {
            "seconds": round(it_s, 1),
            "basis": "extrapolated, not measured: the sampled per-step rate over 2 x 8 x "
                     f"{w.nt} steps, imaging and the Born source not counted"}
    line = {
        "metric": "RTM Gpts/s (% HBM roofline); LSRTM time-to-1e-2 residual, learned vs plain CG",
        "value": round(fp.get("seconds", float("nan")), 2), "unit": "s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(el * 1e3, 1), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": LSRTM_WORKLOAD, "grid": list(w.n), "nt": w.nt,
                   "shots": len(w.shots), "store": w.store, "graphs": a.graphs,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "brtm::k_step_x4r VL step with fused imaging from FP16 "
                               "snapshots (26 B per cell)",
                     "kernel_ms": round(k_ms, 4), "launches": kcount},
        "cpu_baseline": cpu, "lsrtm": out,
    }
    print(json.dumps(line), flush=True)
    dev.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rtm512")
    ap.add_argument("--nt", type=int, default=0, help="override nt (testing only)")
    ap.add_argument("--ty", type=int, default=0, help="tile rows (0 = per-variant default)")
    ap.add_argument("--target-blocks", type=int, default=2048)
    ap.add_argument("--x-per-lane", type=int, default=0, choices=(0, 1, 2, 4),
                    help="cells per lane of the forward-form step kernel (0: automatic, "
                         "4 at 512^3)")
    ap.add_argument("--graphs", type=int, default=-1, choices=(-1, 0, 1),
                    help="propagation loops as graphs: -1 automatic (grids up to 4 M cells), "
                         "0 off, 1 on")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the whole shot on the CPU restatement (the default for "
                         "rtm128; at 512^3 it takes hours)")
    ap.add_argument("--traffic-bytes", type=float, default=0.0,
                    help="HBM bytes per stencil launch measured with rocprofv3 --pmc")
    ap.add_argument("--lsrtm-iters", type=int, default=5,
                    help="iterations of plain CG and of learned flexible CG on lsrtm256 "
                         "added to the RTM line (0: none)")
    ap.add_argument("--lsrtm-niter", type=int, default=400,
                    help="iteration limit of each solver with --workload lsrtm256")
    ap.add_argument("--lsrtm-solvers", default="cg,diag,fpcg",
                    help="with --workload lsrtm256: which of plain CG (cg), the illumination "
                         "diagonal (diag) and learned flexible CG (fpcg) to run")
    ap.add_argument("--lsrtm-weights", default="",
                    help="network parameters of the learned run (default: the shipped ones)")
    a = ap.parse_args()

    if a.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU: the whole shot domain "
                                   "sits in one MI355X's HBM; multi-GPU sharding is by shot "
                                   "and is not implemented here", "n_gpus": a.gpus}))
        sys.exit(2)

    if a.workload == "lsrtm256":
        lsrtm_main(a)
        return

    from sep_code_amd import configs
    t0 = time.perf_counter()
    kw = {"nt": a.nt} if a.nt else {}
    w = configs.WORKLOADS[a.workload](**kw)
    log(f"workload {w.name} n={w.n} nt={w.nt} generated in {time.perf_counter() - t0:.1f}s")
    dev = setup_device(w, a.ty, a.target_blocks, a.x_per_lane, a.graphs)
    log(f"setup done ({dev.device_bytes() / 2**30:.1f} GiB device) "
        f"{time.perf_counter() - t0:.1f}s")

    for i in range(a.warmup):
        rtm_shot(dev, w)
        dev.sync()
        log(f"warmup {i + 1}/{a.warmup} done")

    # Per-launch HIP events around the plain step bracket every launch of the
    # timed shots -- unless the propagation loops run as graphs (small grids),
    # which per-launch events would turn back into launch-by-launch work: then
    # one more shot, untimed and launched eagerly, gives the kernel's duration.
    graphs = a.graphs == 1 or (a.graphs == -1 and w.ncell <= GRAPH_MAX_CELLS)
    if not graphs:
        dev.kernel_timing(True, 0, 0, 0)
    dev.sync()
    t1 = time.perf_counter()
    for i in range(a.steps):
        rtm_shot(dev, w)
        dev.sync()
        log(f"step {i + 1}/{a.steps} done at {time.perf_counter() - t1:.2f}s")
    dev.sync()
    el = time.perf_counter() - t1
    if graphs:
        dev.kernel_timing(True, 0, 0, 0)
        rtm_shot(dev, w)
        dev.sync()
    ktot, kcount = dev.kernel_timing_result()
    dev.kernel_timing(False)
    img = dev.get_image()
    finite = bool(np.isfinite(img).all())

    ncell = w.ncell
    ms_step = el / a.steps * 1e3
    value = 2.0 * ncell * w.nt * a.steps / el / 1e9
    k_ms = ktot / max(kcount, 1)
    achieved = BYTES_PER_CELL * ncell / (k_ms * 1e-3) / 1e9
    traffic = a.traffic_bytes or None
    default_cfg = w.n == (512, 512, 512) and a.x_per_lane in (0, 4) and a.ty == 0
    if traffic is None and default_cfg:
        traffic = TRAFFIC_X4R_512
    kxl, kty = plain_step_kernel(w.n, a.x_per_lane) if a.ty == 0 else (
        cells_per_lane(w.n[0], a.x_per_lane), a.ty)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": ("--traffic-bytes of this run" if a.traffic_bytes else
                               "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch of "
                               "this kernel at this shape, from a separate counter run of the "
                               "same build (profiles/stencil512_traffic.md), not measured in "
                               "this run" if traffic else "not measured for this shape"),
            "kernel": {1: "brtm::k_step (VL, fused sponge)",
                       2: "brtm::k_step_x2 (VL, 2 cells/lane, fused sponge)",
                       4: "brtm::k_step_x4r (VL, 4 cells/lane, six-plane halo ring, fused sponge)"
                       }[kxl] + f", {kty}-row tiles" + (
                           " + streaming loads" if kxl == 4 and 16 * ncell > STREAM_MIN_STEP_BYTES
                           else ""),
            "kernel_ms": round(k_ms, 4), "launches": kcount,
            "kernel_timing": ("HIP events around every launch in the timed shots" if not graphs
                              else "HIP events around every launch of one extra, eagerly "
                                   "launched shot (the timed shots ran as graphs)"),
            "stencil_gpts": round(ncell / (k_ms * 1e-3) / 1e9, 2)}
    cpu = None
    if not a.no_cpu:
        log("cpu baseline ...")
        if a.cpu_full or w.name == "rtm128":
            cpu = cpu_full_shot(w, dev.get_receivers(), img)
        else:
            cpu = cpu_baseline(w, a.cpu_seconds)
    out = {
        "metric": "RTM Gpts/s (% HBM roofline); LSRTM time-to-1e-2 residual, learned vs plain CG",
        "value": round(value, 3), "unit": "Gpts/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_step, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{w.name}: RTM shot, {w.n[0]}x{w.n[1]}x{w.n[2]} FP32, "
                               f"nt={w.nt}, " + (f"source checkpoints every {w.ckpt} steps"
                                                 if w.store == "ckpt" else
                                                 f"{w.store} source snapshots") +
                               ", imaging every step",
                   "grid": list(w.n), "nt": w.nt, "store": w.store, "ckpt": w.ckpt,
                   "graphs": graphs, "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "image_finite": finite,
        # the same shot under other counts of cell-updates per time step:
        # value credits the two wavefields of an RTM (source and receiver);
        # one field per time step is n^3 nt / t; all propagated steps adds the
        # checkpoint recompute (three fields per time step)
        "rates_gpts": {"rtm_two_fields": round(value, 3),
                       "one_field_per_time_step": round(value / 2, 3),
                       "all_propagated_steps": round(value * (1.5 if w.store == "ckpt" else 1.0),
                                                     3)},
    }
    dev.close()
    if a.lsrtm_iters > 0 and w.name == "rtm512":
        out["lsrtm"] = lsrtm_slice(a.lsrtm_iters, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
