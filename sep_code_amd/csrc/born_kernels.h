// Device-side parameter blocks and launchers for the Born RTM kernels
// (gfx950).  Host code in born_context.hip includes this header.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace brtm {

// Storage format of a source-wavefield snapshot.
enum SnapFmt { FMT_NONE = 0, FMT_F32 = 1, FMT_BF16 = 2, FMT_F16 = 3 };
// Step form: VL = V.L(p) (Born/wave_fkernel.3d8o.cu:31), LV = L(V.p) (:109).
enum StepForm { FORM_VL = 0, FORM_LV = 1 };
// What the step does with the snapshot s besides the update:
//   ACT_IMG  img += cimg . s . p_new        (zero-lag imaging, VL only)
//   ACT_BORN p_new += cimg . s . m          (Born secondary source)
enum StepAct { ACT_NONE = 0, ACT_IMG = 1, ACT_BORN = 2 };
// Receivers handled inside the step (the four-cell ring kernel, and the
// one-cell kernel's 8-row small-grid shape):
//   REC_INJECT   pn += sum_k rsinc[k] rtr(r, rid+k) before imaging and the store
//   REC_EXTRACT  rtr(r, rid+k) += rsinc[k] . pn . (rscale[r] or 1)
// for the cells of plane rz whose rmap entry is a receiver index r >= 0.
enum RecMode { REC_NONE = 0, REC_INJECT = 1, REC_EXTRACT = 2 };

constexpr int kWMax = 256;  // longest sponge table the kernels accept

// One time step of the fused stencil:
//   VL: pn = w.(V.L p + 2p) - w^2.po
//   LV: pn = L(V.w.p) + 2w.p - w^2.po
// followed by the action (StepAct).
// s is the source snapshot (format fmt, per-slice scale sscale for 16-bit).
struct StepParams {
  int n1, n2, n3, nw;
  int ntx, nty, ntz, zc;
  float c[13];
  const float *p;
  const float *po;
  float *pn;            // may alias po
  const float *V;
  const float *wtab;    // nw sponge weights
  float *img;           // VL: image accumulator
  const void *snap;     // source snapshot
  const float *sscale;  // [n3] slice scales (16-bit formats), else null
  const float *m;       // LV: Born model
  float cimg;
  // receivers (RecMode): rmap [n2][n1] point index or -1 on plane rz; the
  // trace sample (r, t) at rtr[r * rsp + t * rst], time index rid, taps rsinc
  // (the data receivers: rsp 1, rst nrec; the source points: rsp ntr, rst 1)
  int rz, nrec, rid;
  long rsp, rst;
  // slices of global-load lookahead of the four-cell ring step (1, or 2 for
  // the plain step)
  int pf;
  // 1: the four-cell ring step reads p_old, V, the snapshot and the
  // image/model with nontemporal (streaming) loads
  int stream;
  const int *rmap;
  float *rtr;
  const float *rscale;
  float rsinc[8];
  // two-step pass (launch_step2): p^{n+1} goes to pn, p^{n+2} to pn2; the
  // second step's receiver time index and taps
  float *pn2;
  int rid2;
  float rsinc2[8];
};

// rec (RecMode): the receivers or source points of one plane inside the
// step; VL form with 8-row tiles (the small-grid shape) only.
hipError_t launch_step(int form, int act, int fmt, StepParams P, int ty_tile, hipStream_t st,
                       int rec = REC_NONE);
// The VL step with a six-plane LDS ring (halos loaded with the z-queue).
hipError_t launch_step_ring(int act, int fmt, StepParams P, int ty_tile, hipStream_t st);
// The VL step with two consecutive x per lane (128-wide tiles, n1 even).
hipError_t launch_step_x2(int act, int fmt, StepParams P, int ty_tile, hipStream_t st);
// The VL step with four consecutive x per lane (256-wide tiles, n1 % 4 == 0),
// on the two-plane buffer or (ring) the six-plane halo ring; rec (RecMode,
// ring only) handles the receivers inside the step.
hipError_t launch_step_x4(int act, int fmt, StepParams P, int ty_tile, bool ring,
                          hipStream_t st, int rec = REC_NONE);
// Two consecutive plain VL steps in one pass: p^{n+1} = step(p = p^n,
// po = p^{n-1}) into pn and p^{n+2} = step(p^{n+1}, p^n) into pn2, with the
// sponge, and (rec = REC_INJECT) the source points of plane rz injected after
// each step (time indices rid, rid2; taps rsinc, rsinc2).  64 x 32 output
// tiles; the first step also runs on the tile's 4-cell rim.  pn and pn2 must
// not alias p, po or each other (other tiles read p and po around their
// own).  n1 % 4 == 0.  Fills the tile geometry itself.
hipError_t launch_step2(StepParams P, int target_blocks, hipStream_t st, int rec = REC_NONE);
// Tile / chunk geometry chosen for a grid (fills ntx..zc of P); xw is the
// tile width in x (64, 128 or 256 for the 1-, 2- and 4-cell kernels).
void plan_step(StepParams &P, int ty_tile, int target_blocks, int xw = 64);

// p[loc_j] += scale * sum_k sinc8[k] tr(j, id+k), tr(j,t) = tr[j*sp + t*st].
// When img != null: img[loc_j] += cimg * s[loc_j] * (added amount).
struct InjectParams {
  int npts;
  const int *locs;
  const float *tr;
  long sp, st;
  int id;
  float sinc8[8];
  float scale;
  float *p;
  float *img;
  const void *snap;
  const float *sscale;
  int fmt;
  int n12;
  float cimg;
};
hipError_t launch_inject(const InjectParams &P, hipStream_t st);

// tr(j, id+k) += sinc8[k] * p[loc_j]  (device trace layout [t][rec]).
// With rscale != null the extracted value is p[loc_j] * rscale[j].
hipError_t launch_extract(int npts, const int *locs, float *tr, long st_t, int id,
                          const float sinc8[8], const float *p, const float *rscale,
                          hipStream_t st);
// out[i] = V[i] * w(i) * in[i] over the compute grid (w the sponge weight,
// 0 outside the update set); the diagonal D with A_LV = D^-1 A_VL D.
hipError_t launch_vw_scale(const float *V, const float *wtab, int nw, const float *in,
                           float *out, int n1, int n2, int n3, hipStream_t st);
// rs[j] = 1 / (V w)[loc_j] (0 for inactive receivers).
hipError_t launch_receiver_scale(const float *V, const float *wtab, int nw, const int *locs,
                                 int npts, float *rs, int n1, int n2, int n3, hipStream_t st);

// Snapshot encoding: per-slice max |s| then quantise to fmt.
hipError_t launch_encode(int fmt, const float *s, void *dst, float *sscale, int n12, int n3,
                         hipStream_t st);
// Decode one snapshot to FP32 (tests, diagnostics).
hipError_t launch_decode(int fmt, const void *src, const float *sscale, float *dst, int n12,
                         int n3, hipStream_t st);

// acc += s^2 over one snapshot (illumination).
hipError_t launch_square_accumulate(int fmt, const void *src, const float *sscale, float *acc,
                                   int n12, int n3, hipStream_t st);

// FP64 reductions / vector kernels.
constexpr int kDotBlocks = 1024;
hipError_t launch_dot(const float *a, const float *b, long n, double *partials,
                      double *out, hipStream_t st);
hipError_t launch_axpby(float *y, const float *x, double a, double b, long n,
                        hipStream_t st);  // y = a*y + b*x
hipError_t launch_mul(float *y, const float *x, long n, hipStream_t st);  // y *= x
hipError_t launch_fill(float *y, float v, long n, hipStream_t st);
// Zero the 4-cell halo (outside the update set) of nfields fields of
// [n3][n2][n1] floats placed stride floats apart.
hipError_t launch_zero_halo(float *p, long stride, int nfields, int n1, int n2, int n3,
                            hipStream_t st);

// Volume operators (compute grid [n3][n2][n1]).
hipError_t launch_dz(int adj, const float *in, float *out, int n12, int n3, int add,
                     hipStream_t st);
// Embed a model-grid volume into the compute grid at offset (o1,o2,o3)
// (adj = 0), or accumulate the compute-grid window back (adj = 1).
hipError_t launch_embed(int adj, const float *model, float *comp, int m1, int m2, int m3,
                        int n1, int n2, int n3, int o1, int o2, int o3, hipStream_t st);
// Traces: device [ntb][nrec] <-> data [nrec][nt] with window (drop 4) and
// time derivative d[t] = (w[t-1] - w[t]) / dt.
hipError_t launch_trace_to_data(const float *tr, float *data, int nrec, int ntb, int nt,
                                float inv_dt, int add, hipStream_t st);
hipError_t launch_data_to_trace(const float *data, float *tr, int nrec, int ntb, int nt,
                                float inv_dt, hipStream_t st);
// Streaming copy of n16 16-byte words (calibration of bandwidth counters).
hipError_t launch_copy16(const void *src, void *dst, long n16, hipStream_t st);
// out = a (+ b (+ c)): nin (1..3) streams of n16 16-byte words read, one
// written (nontemporal store when nt_store); bandwidth calibration for a
// kernel's read/write mix.
hipError_t launch_sum16(int nin, bool nt_store, const void *a, const void *b, const void *c,
                        void *out, long n16, hipStream_t st);
// Plain transpose [rows][cols] -> [cols][rows].
hipError_t launch_transpose(const float *in, float *out, long rows, long cols,
                            hipStream_t st);

}  // namespace brtm
