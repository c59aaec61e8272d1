// Internal context of the Born RTM library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "born_kernels.h"

struct brtm_ctx;

namespace brtm {

void set_error(const std::string &msg);

#define BRTM_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::brtm::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return BRTM_EHIP;                                                             \
    }                                                                               \
  } while (0)

#define BRTM_REQUIRE(cond, code, msg)   \
  do {                                  \
    if (!(cond)) {                      \
      ::brtm::set_error(msg);           \
      return (code);                    \
    }                                   \
  } while (0)

// Device allocation owned by a context.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  int alloc(size_t n);
  void release();
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// One shot's source wavefield.
struct Store {
  int kind = 0;    // BRTM_STORE_*
  int nt = 0, jt = 1, jimg = 1, ckpt = 0, count = 0;
  DevBuf data;     // snapshots [count][ncell] or checkpoints [count][2][ncell]
  DevBuf scale;    // [count][n3] slice scales (16-bit formats)
  // checkpoints and FP32 snapshots: every field slot's halo is zero (the
  // steps writing into the slots never touch it), so a later propagation of
  // the same kind can reuse the buffer as it is
  bool halo_zero = false;
  // the operator (brtm_op) whose brtm_op_prepare made this store, or null:
  // operators on one context share its per-shot stores, and one releases or
  // reads only the stores it made
  const void *owner = nullptr;
  void release() {
    data.release();
    scale.release();
    halo_zero = false;
    owner = nullptr;
  }
};

}  // namespace brtm

struct brtm_ctx {
  // Lifetime.  Every vector, operator, solver and network created on the
  // context holds one reference (brtm::CtxHold) and brtm_create holds one.
  // brtm_destroy frees the context's device memory at once and drops its
  // reference; the host record (stream, events) stays until the last holder
  // is destroyed, so destroying a child after its context is safe.
  std::atomic<int> refs{1};
  bool destroyed = false;
  int device = 0;
  int ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  bool have_space = false, have_vel = false;
  int n1 = 0, n2 = 0, n3 = 0;
  long ncell = 0;
  float d[3] = {0, 0, 0};
  float bc_a = 0, bc_b = 0;
  float coeffs[13];
  int nw = 0;
  std::vector<float> wtab;
  float cimg = 1e-6f;
  int ty_tile = 0, target_blocks = 2048;  // ty_tile 0: chosen per step variant
  int x_per_lane = 0;  // VL steps: 0 automatic, 1, 2 (n1 even) or 4 (n1 % 4 == 0)
  int halo_ring = -1;  // six-plane LDS ring: -1 automatic (on for 4 cells per lane), 0, 1
  bool rec_fusion = true;  // receivers inside the four-cell ring step where they allow it
  int prefetch = 0;        // load lookahead of the plain four-cell ring step: 0 auto, 1, 2
  int stream_loads = -1;   // streaming loads in the four-cell ring step: -1 auto, 0, 1
  int graphs = -1;         // propagation loops launched as graphs: -1 auto, 0, 1
  // plain source-field steps two per pass (launch_step2): -1 auto, 0, 1
  int two_step = 0;
  // executable graphs launched and not yet known to be finished: freed once
  // the stream has been synchronised
  std::vector<hipGraphExec_t> graph_execs;

  brtm::DevBuf d_wtab, V, fa, fb, fc, fd, img, mdl, seg;
  brtm::DevBuf mvw, rscale;  // V.w.m and 1/(V.w) at receivers (Born path on u = V.w.q)
  bool born_direct_lv = false;
  int seg_slots = 0;
  brtm::DevBuf partials, dotv;
  // this device's copy of the allocation-free entry point's scratch array
  // (its device address, from hipGetSymbolAddress at brtm_create)
  double *noalloc_scratch = nullptr;

  // interpolation tables [jt][8]
  std::vector<float> sinc[2];

  // source geometry and traces
  int nsrc = 0, ntr_s = 0;
  brtm::DevBuf slocs, strc;
  // receivers
  int nrx = 0, nry = 0, nrec = 0, ntr_d = 0;
  brtm::DevBuf rlocs, traces;  // traces [ntr_d][nrec]
  std::vector<int> h_rlocs;
  // When the active receivers share one plane (rec_z) with one receiver per
  // cell, rmap [n2][n1] holds each cell's receiver index (-1: none) and the
  // step kernel injects / extracts them itself.
  bool rec_plane = false;
  int rec_z = -1;
  brtm::DevBuf rmap;
  // the same for the source points (the source propagation's injection)
  bool src_plane = false;
  int src_z = -1;
  brtm::DevBuf smap;

  std::vector<brtm::Store> stores;
  int shot = 0;

  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // Per-launch timing of one (form, fmt) stencil variant: an event pair is
  // recorded around every such launch while enabled.
  bool ktime_on = false;
  int ktime_form = 0, ktime_act = 0, ktime_fmt = 0;
  std::vector<hipEvent_t> ktime_ev;  // 2 per launch
  size_t ktime_used = 0;
};

namespace brtm {

// Reference counting of a context's host record (see brtm_ctx::refs).
void ctx_retain(brtm_ctx *c);
void ctx_release(brtm_ctx *c);

// One reference to a context, held by an object created on it.  Declared as
// the first member of such an object, so it is released after the object's
// destructor and its other members (device buffers) are done.
struct CtxHold {
  brtm_ctx *c = nullptr;
  CtxHold() = default;
  CtxHold(const CtxHold &) = delete;
  CtxHold &operator=(const CtxHold &) = delete;
  ~CtxHold() { reset(nullptr); }
  void reset(brtm_ctx *ctx) {
    if (ctx) ctx_retain(ctx);
    if (c) ctx_release(c);
    c = ctx;
  }
};

// BRTM_ESTATE unless the context is live (brtm_destroy not yet called).
#define BRTM_LIVE(ctx)                                                            \
  BRTM_REQUIRE((ctx) && !(ctx)->destroyed, BRTM_ESTATE,                          \
               "the context this object was created on has been destroyed")

// Internal propagation entry points (used by the C ABI and the operator).
int source_prop(brtm_ctx *c, int jt, int nt, int jimg, int store, int ckpt, float *p0,
                float *p1);
int born_forward(brtm_ctx *c, int jt, int nt, int jimg, bool accumulate);
int rtm_adjoint(brtm_ctx *c, int jt, int nt, int jimg);
int dot(brtm_ctx *c, const float *a, const float *b, long n, double *out);
// acc += sum over the current shot's stored snapshots of s^2 (compute grid).
int illumination(brtm_ctx *c, float *acc);
Store &cur_store(brtm_ctx *c);
// rec (RecMode, or REC_SOURCE for the source points): receivers handled by
// the step at time index it (jt steps per sample, rscale for extraction)
// when *fused comes back true; the caller runs inject_data / extract_data /
// inject_source otherwise.
constexpr int REC_SOURCE = 3;
int step(brtm_ctx *c, int form, int act, int fmt, const float *p, const float *po, float *pn,
         const void *snap, const float *sscale, float *img_or_null, const float *m_or_null,
         int rec = REC_NONE, int it = 0, int jt = 1, const float *rscale = nullptr,
         bool *fused = nullptr);

}  // namespace brtm

// A float vector in HBM (the C ABI's brtm_vec).  ptr is the storage: buf's
// allocation for an owning vector, or a slice of another vector's storage
// for a view (a part of a stacked vector), which owns nothing.
struct brtm_vec {
  brtm::CtxHold hold;  // owning vectors only; views hold nothing
  brtm_ctx *ctx = nullptr;
  long n = 0;
  float *ptr = nullptr;
  brtm::DevBuf buf;
  float *p() const { return ptr; }
};

struct brtm_op;

namespace brtm {
// BRTM_EINVAL with a message unless v is a vector of n elements.
int check_vec(const brtm_vec *v, long n, const char *what);
// The context a Born operator propagates on.
brtm_ctx *op_context(const brtm_op *op);
}  // namespace brtm
