/*
 * born_rtm.h -- C ABI of the MI355X Born acoustic RTM library
 * (sep_code_amd/lib/libborn_rtm.so).
 *
 * This is the drop-in boundary for the reference's device layer
 * Born/gpu_funcs_3d.h:5-27.  Every entry point below names the reference
 * function it replaces.  Differences from the reference, by design:
 *   - state lives in an explicit context (the reference keeps file-scope
 *     globals and __constant__ symbols, Born/gpu_finite_3d.h:31-81);
 *   - every call returns an int status (BRTM_OK = 0) and brtm_last_error()
 *     describes the last failure of the calling thread (the reference
 *     returns void and mostly ignores errors, Born/gpu_funcs_3d.cu:47-62);
 *   - device buffers are owned by the context and freed only by
 *     brtm_destroy / brtm_release_store (the reference frees traces and the
 *     source inside each propagation, Born/gpu_funcs_3d.cu:312-313,500-503);
 *   - one GPU holds the whole domain: the slab split and P2P halos of the
 *     reference (Born/gpu_funcs_3d.cu:143,256-269) are plain in-HBM padding.
 * A context must be used by one host thread at a time.
 *
 * Host arrays are caller-owned, FP32, little endian.  Volumes are
 * [n3][n2][n1] with n1 fastest on the padded compute grid; the stencil
 * radius is 4 and cells within 4 of a face are never updated.  Receiver
 * traces cross the boundary in the reference layout [ny][nx][ntr] (time
 * fastest); inside the library they are stored [ntr][nrec] so a time step
 * touches one contiguous row.
 *
 * The reference's exact signatures (setup_cuda, source_prop, rtm_forward,
 * rtm_adjoint, ...) are provided over a process-wide default context by
 * gpu_funcs_3d.h in this directory.
 */
#ifndef BORN_RTM_H
#define BORN_RTM_H

#ifdef __cplusplus
extern "C" {
#endif

#define BRTM_OK 0
#define BRTM_EINVAL -1     /* invalid argument                          */
#define BRTM_EHIP -2       /* HIP runtime error                         */
#define BRTM_ENOMEM -3     /* device or host allocation failed          */
#define BRTM_ESTATE -4     /* call order violated (e.g. no velocity)    */
#define BRTM_EUNSTABLE -5  /* velocity violates the leapfrog CFL limit  */

/* Source-wavefield policies for brtm_source_prop. */
#define BRTM_STORE_NONE 0  /* keep nothing (last two fields only)             */
#define BRTM_STORE_F32 1   /* snapshot every jimg steps, FP32                 */
#define BRTM_STORE_BF16 2  /* snapshot every jimg steps, BF16 + per-z scale   */
#define BRTM_STORE_F16 3   /* snapshot every jimg steps, FP16 + per-z scale   */
#define BRTM_STORE_CKPT 4  /* FP32 checkpoints every ckpt steps, recompute    */

typedef struct brtm_ctx brtm_ctx;

/* Message describing the last failing call on this thread. */
const char *brtm_last_error(void);

/* Library version string (no counterpart in the reference). */
const char *brtm_version(void);

/* Replaces setup_cuda (Born/gpu_funcs_3d.cu:10-45): bind a context to one
 * GPU (no peer access: one GPU holds the domain). */
int brtm_create(int device, brtm_ctx **out);
/* Frees the context's device memory.  Vectors, operators, solvers and
 * networks created on it (born_op.h, born_solver.h, born_unet.h) may be
 * destroyed before or after it: each holds a reference to the context's
 * host record (stream, events), which goes with the last of them.  Between
 * brtm_destroy and their own destroy call, every other call on them returns
 * BRTM_ESTATE.  Destroying a context twice is an error. */
int brtm_destroy(brtm_ctx *ctx);

/* Replaces create_gpu_space (Born/gpu_funcs_3d.cu:777-803): grid spacing,
 * sponge parameters (w = exp(-bc_b (bc_a - e)) for edge distance e < bc_a,
 * Born/wave_fkernel.3d8o.cu:194-205) and compute-grid size.  Allocates the
 * wavefields.  Stencil weights are formed in double as
 * Born/wave_fkernel.3d8o.cu:26-27. */
int brtm_create_space(brtm_ctx *ctx, float d1, float d2, float d3, float bc_a, float bc_b,
                      int n1, int n2, int n3);

/* Replaces transfer_vel_func1 (Born/gpu_funcs_3d.cu:761-767): V = v^2 dt^2
 * on the compute grid.  Returns BRTM_EUNSTABLE when max V exceeds the 3-D
 * 8th-order leapfrog limit (the reference's 0.5 dmin/vmax rule at
 * Born/fd_prop_3d.cpp:13 does not guarantee it). */
int brtm_set_velocity(brtm_ctx *ctx, const float *V);

/* Replaces transfer_sinc_table_s / _d (Born/gpu_funcs_3d.cu:714-733):
 * which = 0 source table, 1 receiver table; table is [ntab][nsinc], nsinc
 * must be 8; ntab = steps per data sample (jt). */
int brtm_set_sinc_table(brtm_ctx *ctx, int which, int nsinc, int ntab, const float *table);

/* Replaces transfer_source_func (Born/gpu_funcs_3d.cu:735-745): npts
 * injection cells (linear index on the compute grid, inside the update
 * set) and their traces [npts][ntr]. */
int brtm_set_source(brtm_ctx *ctx, int npts, int ntr, const int *locs, const float *vals);

/* Replaces transfer_receiver_func (Born/gpu_funcs_3d.cu:747-758): nx*ny
 * receiver cells (locs < 0, or cells outside the update set, are inactive)
 * and their traces [ny][nx][ntr] (vals may be NULL: zero traces). */
int brtm_set_receivers(brtm_ctx *ctx, int nx, int ny, int ntr, const int *locs,
                       const float *vals);
/* Download the receiver traces in the layout [ny][nx][ntr] (the copy back
 * at the end of rtm_forward, Born/gpu_funcs_3d.cu:489). */
int brtm_get_receivers(brtm_ctx *ctx, float *vals);

/* Born model (compute grid, the img argument of rtm_forward,
 * Born/gpu_funcs_3d.cu:317,381) and image accumulator (the img output of
 * rtm_adjoint, :509,695). */
int brtm_set_model(brtm_ctx *ctx, const float *m);
int brtm_get_image(brtm_ctx *ctx, float *img);
int brtm_zero_image(brtm_ctx *ctx);

/* Source-wavefield stores are kept per shot in HBM; subsequent propagation
 * calls use the store of the selected shot.  They replace the last two
 * source fields that rtm_zero_op::create_source_fields writes per shot to
 * its source_fields file (Born/rtm_zero_op_3d.cpp:70-76). */
int brtm_select_shot(brtm_ctx *ctx, int ishot);
int brtm_release_store(brtm_ctx *ctx, int ishot);

/* Replaces source_prop (Born/gpu_funcs_3d.cu:133-315): nt steps of the
 * source field s^{it+1} = w.(V.L s^it + 2 s^it) - w^2 s^{it-1} plus
 * injection (dir = +1, sinc row it % jt at sample it / jt).  store selects
 * what is kept for the selected shot: snapshots s^{it+1} for it % jimg == 0
 * (FP32/BF16/FP16), or FP32 checkpoints every ckpt steps.  p0, p1 (may be
 * NULL) receive s^{nt-1} and s^{nt}. */
int brtm_source_prop(brtm_ctx *ctx, int jt, int nt, int jimg, int store, int ckpt, float *p0,
                     float *p1);

/* Replaces rtm_forward (Born/gpu_funcs_3d.cu:317-505): Born modelling of
 * the model set by brtm_set_model into the receiver traces (zeroed first
 * unless accumulate).  Scattered field q^{it+1} = L(V.w.q^it) + 2w q^it -
 * w^2 q^{it-1} + cimg s^{it+1}.m (Born source when it % jimg == 0), then
 * extraction at sample it/jt + k with receiver sinc row it % jt.  The source
 * field comes from the selected shot's store, or is propagated alongside
 * when the store is BRTM_STORE_NONE / BRTM_STORE_CKPT. */
int brtm_rtm_forward(brtm_ctx *ctx, int jt, int nt, int jimg, int accumulate);

/* Replaces rtm_adjoint (Born/gpu_funcs_3d.cu:509-712): the exact transpose
 * of brtm_rtm_forward.  Receiver field lambda^{it+1} = w.(V.L lambda^{it+2}
 * + 2 lambda^{it+2}) - w^2 lambda^{it+3} + injected traces, image +=
 * cimg s^{it+1}.lambda^{it+1} when it % jimg == 0.  The image accumulator
 * is added to (brtm_zero_image clears it).  Requires a store for the
 * selected shot (snapshots or checkpoints). */
int brtm_rtm_adjoint(brtm_ctx *ctx, int jt, int nt, int jimg);

/* Imaging scale (1e-6 in the reference, Born/gpu_funcs_3d.cu:117,128). */
int brtm_set_imaging_scale(brtm_ctx *ctx, float cimg);

/* ---- Diagnostics, timing and tuning (tests and benchmarks): no
 * counterpart in the reference, and none changes a result's meaning. ---- */

/* Decode snapshot k of the selected shot's store into host FP32. */
int brtm_get_snapshot(brtm_ctx *ctx, int k, float *out);

/* Plain propagation of host fields (tests and timing): nsteps of the VL
 * (form 0) or LV (form 1) step, no sources; p_cur/p_old are replaced by the
 * newest and previous fields. */
int brtm_propagate(brtm_ctx *ctx, int form, int nsteps, float *p_cur, float *p_old);

/* Kernel timing: run nrep steps of one stencil variant on the context's
 * resident fields between two HIP events on the library stream; avg_ms
 * receives the mean duration.  form 0 = VL, 1 = LV; act 0 = none,
 * 1 = fused imaging (VL only), 2 = fused Born source; fmt 0..3 = snapshot
 * format (none, FP32, BF16, FP16; none exactly when act is none). */
int brtm_time_step(brtm_ctx *ctx, int form, int act, int fmt, int nrep, double *avg_ms);

/* Born modelling path: 0 (default) advances u = V.w.q with the forward-form
 * kernel (the transpose-form step equals D^-1 A_VL D, D = V.w); 1 runs the
 * transpose-form kernel on q directly (cross-check). */
int brtm_set_born_path(brtm_ctx *ctx, int direct_lv);

/* Per-launch timing of one stencil variant (form, act, fmt as above): while
 * enabled, every launch of that variant is bracketed by HIP events on the
 * library stream; the result is the summed duration and launch count since
 * the last brtm_kernel_timing call. */
int brtm_kernel_timing(brtm_ctx *ctx, int enable, int form, int act, int fmt);
int brtm_kernel_timing_result(brtm_ctx *ctx, double *total_ms, int *count);

/* Tile rows per workgroup (8, 16 or 32; 0 = chosen per step variant, the
 * default) and target workgroup count (used when ty_tile != 0). */
int brtm_set_tuning(brtm_ctx *ctx, int ty_tile, int target_blocks);

/* Cells per lane along x for VL-form steps: 1 (64-wide tiles, 4-byte
 * accesses), 2 (128-wide tiles, 8-byte accesses; n1 even) or 4 (256-wide
 * tiles, 16-byte accesses; n1 % 4 == 0); 0 (the default) picks the widest
 * whose tiles keep at least 3/4 of the lanes busy.  A grid that does not
 * allow the width asked for uses the widest one it allows.  Results agree
 * to FP32 rounding. */
int brtm_set_x_per_lane(brtm_ctx *ctx, int cells);

/* One- and four-cell VL steps staged through a ring of six LDS planes whose
 * halos are loaded together with the z-queue (on = 1), or through the
 * two-plane buffer that loads each halo one slice ahead (on = 0); -1, the
 * default, uses the ring for the four-cell kernel (3 % faster at 512^3) and
 * not for the one-cell kernel.  Same results. */
int brtm_set_halo_ring(brtm_ctx *ctx, int on);

/* Receivers and sources inside the step (no counterpart in the reference):
 * when the active receivers share one depth plane with at most one receiver
 * per cell, the four-cell ring step injects the data (RTM adjoint) or
 * extracts it (Born modelling) on that plane itself instead of a separate
 * launch per step; the source points likewise.  on = 1 (the default) allows
 * it, 0 keeps the separate kernels.  The same results, to FP32 rounding of
 * the imaging term. */
int brtm_set_receiver_fusion(brtm_ctx *ctx, int on);

/* Global-load lookahead of the plain four-cell ring step, in z slices: 1
 * (inputs of slice z+1 loaded while slice z computes), 2 (two slices ahead,
 * 36 more registers per lane) or 0 (the default: the faster of the two on
 * MI355X).  No counterpart in the reference; the same results. */
int brtm_set_prefetch(brtm_ctx *ctx, int slices);

/* Streaming (nontemporal) loads of p_old, V, 16-bit snapshots and the
 * model, which the four-cell ring step reads once: 1 on, 0 off, -1 (the
 * default) on when a step moves more than 512 MB, twice the Infinity Cache
 * (512^3 yes, 256^3 no).  The image and FP32 snapshots stream at every
 * size.  No counterpart in the reference; the same results. */
int brtm_set_stream_loads(brtm_ctx *ctx, int mode);

/* Launch each propagation loop (source propagation, Born modelling, the
 * adjoint) as one graph of its captured kernel launches instead of launch by
 * launch: 1 on, 0 off, -1 (the default) on for grids of at most 4 M cells,
 * where the time between dependent launches is a sizeable part of a step
 * (128^3).  Per-launch kernel timing (brtm_kernel_timing) launches eagerly.
 * No counterpart in the reference; the same results. */
int brtm_set_graphs(brtm_ctx *ctx, int mode);

/* Advance the source field two plain steps per pass (temporal blocking):
 * the source propagation and the checkpoint recompute run steps (it, it+1),
 * it even, as one kernel that reads p^n, p^{n-1} and V once and writes
 * p^{n+1} and p^{n+2}, 20 bytes per cell for two steps instead of 32.  The
 * source points are injected inside the pass, so it applies when they lie
 * on one plane, one per cell (else the steps run one by one).  1 on, 0 off
 * (the default), -1 automatic.  No counterpart in the reference; each step
 * gives the same bits as the single step.  brtm_kernel_timing and
 * brtm_time_step take form 2 for this pass. */
int brtm_set_two_step(brtm_ctx *ctx, int mode);

/* Calibration: nrep launches of a plain streaming copy of `bytes` bytes
 * (16-byte vector loads and stores) between two device buffers; avg_ms
 * receives the mean launch duration.  Its rocprof FETCH_SIZE / WRITE_SIZE
 * scale the counters of the other kernels to HBM bytes.  No counterpart in
 * the reference. */
int brtm_copy_bench(brtm_ctx *ctx, double bytes, int nrep, double *avg_ms);

/* Calibration: nrep launches of out = sum of nin (1..3) input streams of
 * `bytes` bytes each (16-byte vector loads and stores; nt_store 1: the
 * store is nontemporal, as in the stencil steps); avg_ms receives the mean
 * launch duration.  nin = 3 is the plain step's mix (read p, p_old, V;
 * write p_new) without its arithmetic, halos or LDS.  No counterpart in the
 * reference. */
int brtm_stream_bench(brtm_ctx *ctx, int nin, double bytes, int nt_store, int nrep,
                      double *avg_ms);

/* FP64 inner product sum_i (double)a[i] * (double)b[i] of two device FP32
 * arrays of n elements (n >= 0), into the host double *out, on the
 * context's stream; the same kernels and the same bits as brtm_vec_dot
 * (born_op.h).  No counterpart in the reference: this is brtm_vec_dot for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc, hipFree or allocating routine: its scratch, the 1024 partial
 * sums and the result (8,200 bytes), is a static device array of the
 * library, one per GPU, held until the process ends.  Its use does not grow
 * with n, so every n >= 0 fits; a negative n is refused.  Calls that
 * overlap on one GPU (several host threads, or contexts of one device on
 * different streams) share that array and race: serialise them. */
int brtm_dot_noalloc(brtm_ctx *ctx, const float *a, const float *b, long n, double *out);

/* Device synchronisation of the library stream (no counterpart in the
 * reference, whose calls return after their last copy). */
int brtm_sync(brtm_ctx *ctx);

/* Bytes of device memory the context currently holds (no counterpart in
 * the reference). */
int brtm_device_bytes(brtm_ctx *ctx, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* BORN_RTM_H */
