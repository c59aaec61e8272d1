// Context, propagation loops and C ABI of the Born RTM library.
//
// The time loops restate Born/gpu_funcs_3d.cu: source_prop (:133-315),
// rtm_forward (:317-505) and rtm_adjoint (:509-712), on one GPU, one
// stream, with the sponge, imaging and Born source fused into the stencil
// launch and no per-step device synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/born_rtm.h"
#include "born_ctx.h"
#include "step_tables.h"

// Scratch of the allocation-free FP64 dot product (brtm_dot_noalloc): the
// kDotBlocks per-block partial sums, then the result.  One copy per GPU,
// held until the process ends.  In host code `scratch` is only a stand-in
// for the symbol: its address on a device comes from hipGetSymbolAddress
// with that device current (kept in brtm_ctx::noalloc_scratch).
constexpr int kNoallocDoubles = brtm::kDotBlocks + 1;  // 8,200 bytes
static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace brtm {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int DevBuf::alloc(size_t n) {
  if (ptr && bytes >= n) return BRTM_OK;
  release();
  if (n == 0) return BRTM_OK;
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e));
    return BRTM_ENOMEM;
  }
  bytes = n;
  return BRTM_OK;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

Store &cur_store(brtm_ctx *c) {
  if ((int)c->stores.size() <= c->shot) c->stores.resize(c->shot + 1);
  return c->stores[c->shot];
}

static int fmt_of_store(int kind) {
  switch (kind) {
    case BRTM_STORE_F32: return FMT_F32;
    case BRTM_STORE_BF16: return FMT_BF16;
    case BRTM_STORE_F16: return FMT_F16;
    default: return FMT_NONE;
  }
}

static size_t fmt_bytes(int fmt) { return fmt == FMT_F32 ? 4 : 2; }

int illumination(brtm_ctx *c, float *acc) {
  Store &S = cur_store(c);
  const int sfmt = fmt_of_store(S.kind);
  BRTM_REQUIRE(sfmt != FMT_NONE && S.count > 0, BRTM_ESTATE,
               "illumination needs a snapshot store (f32, bf16 or f16) for the shot");
  for (int k = 0; k < S.count; ++k) {
    const void *snap = S.data.as<char>() + fmt_bytes(sfmt) * (size_t)c->ncell * k;
    const float *ss = sfmt == FMT_F32 ? nullptr : S.scale.as<float>() + (size_t)c->n3 * k;
    BRTM_HIP(launch_square_accumulate(sfmt, snap, ss, acc, c->n1 * c->n2, c->n3, c->stream));
  }
  return BRTM_OK;
}

// Host-to-device copy ordered on the library stream: work already queued
// may still read the destination (the stream is non-blocking, so a plain
// hipMemcpy on the null stream would not wait for it), and the host buffer
// is pageable, so the copy completes before returning.
static hipError_t h2d(brtm_ctx *c, void *dst, const void *src, size_t bytes) {
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

// Load lookahead of the plain four-cell ring step when left automatic.
constexpr int kAutoPrefetch = 1;
// Streaming (nontemporal) loads pay off once a step's working set (16 B per
// cell: p, p_old, V, p_new) is well beyond the 256 MB Infinity Cache: 6 %
// faster at 512^3.  Below that the fields stay cache-resident between
// launches, which the hint would forfeit: 11 % slower at 256^3.
constexpr double kStreamMinStepBytes = 512.0 * 1024 * 1024;

#define BRTM_TRY(x)              \
  do {                           \
    int rc_ = (x);               \
    if (rc_ != BRTM_OK) return rc_; \
  } while (0)

// Event pair around a timed launch (brtm_kernel_timing).
static int ktime_begin(brtm_ctx *c, bool timed) {
  if (!timed) return BRTM_OK;
  if (c->ktime_used + 2 > c->ktime_ev.size()) {
    for (int k = 0; k < 256; ++k) {
      hipEvent_t e;
      BRTM_HIP(hipEventCreate(&e));
      c->ktime_ev.push_back(e);
    }
  }
  BRTM_HIP(hipEventRecord(c->ktime_ev[c->ktime_used], c->stream));
  return BRTM_OK;
}

static int ktime_end(brtm_ctx *c, bool timed) {
  if (!timed) return BRTM_OK;
  BRTM_HIP(hipEventRecord(c->ktime_ev[c->ktime_used + 1], c->stream));
  c->ktime_used += 2;
  return BRTM_OK;
}

int step(brtm_ctx *c, int form, int act, int fmt, const float *p, const float *po, float *pn,
         const void *snap, const float *sscale, float *img, const float *m, int rec, int it,
         int jt, const float *rscale, bool *fused) {
  if (fused) *fused = false;
  StepParams P{};
  P.n1 = c->n1;
  P.n2 = c->n2;
  P.n3 = c->n3;
  P.nw = c->nw;
  std::memcpy(P.c, c->coeffs, sizeof(P.c));
  P.p = p;
  P.po = po;
  P.pn = pn;
  P.V = c->V.as<float>();
  P.wtab = c->d_wtab.as<float>();
  P.img = img;
  P.snap = snap;
  P.sscale = sscale;
  P.m = m;
  P.cimg = c->cimg;
  P.pf = c->prefetch == 0 ? kAutoPrefetch : c->prefetch;
  P.stream = c->stream_loads < 0 ? (16.0 * (double)c->ncell > kStreamMinStepBytes)
                                 : c->stream_loads != 0;
  // Tile rows / workgroup target per variant (measured on MI355X at 512^3):
  // plain steps prefer tall tiles, 16-bit snapshot variants short ones.
  int ty = c->ty_tile, tb = c->target_blocks;
  // cells per lane: 4 or 2 where asked for and n1 allows, else 1; automatic
  // (0): the widest whose tiles keep at least 3/4 of the lanes busy
  int xl = 1;
  if (form == FORM_VL) {
    int want = c->x_per_lane;
    if (want == 0) {
      want = 1;
      for (int k : {4, 2}) {
        const int w = 64 * k, used = ((c->n1 + w - 1) / w) * w;
        if (c->n1 % k == 0 && 4 * c->n1 >= 3 * used) { want = k; break; }
      }
    }
    if (want == 4 && c->n1 % 4 == 0) xl = 4;
    else if (want >= 2 && c->n1 % 2 == 0) xl = 2;
  }
  // halo ring: automatic (-1) means on for the 4-cell kernel
  bool ring = c->halo_ring < 0 ? xl == 4 : c->halo_ring != 0;
  if (ty == 0) {
    if (xl == 4) { ty = (ring && act == ACT_IMG && fmt == FMT_F32) ? 8 : 16; tb = 256; }
    else if (xl == 2) { ty = 16; tb = 2048; }
    else if (act == ACT_NONE) { ty = 32; tb = 1024; }
    else if (fmt == FMT_F32) { ty = 16; tb = 8192; }
    else { ty = 8; tb = 8192; }
  }
  if (xl == 4 && ty == 32) ty = 16;  // the 4-cell kernel's tiles are 8 or 16 rows
  plan_step(P, ty, tb, 64 * xl);
  // Small grids: when the automatic shape leaves compute units without a
  // workgroup, 64-wide 8-row tiles give several times as many (128^3: 480
  // workgroups instead of 120, the plain step 10.7 us instead of 12.8).
  if (c->ty_tile == 0 && c->x_per_lane == 0 && c->halo_ring < 0 &&
      (long)P.ntx * P.nty * P.ntz < c->ncu) {
    xl = 1;
    ring = false;
    ty = 8;
    tb = 2048;
    plan_step(P, ty, tb, 64);
  }
  const bool timed =
      c->ktime_on && form == c->ktime_form && act == c->ktime_act && fmt == c->ktime_fmt;
  BRTM_TRY(ktime_begin(c, timed));
  // receivers inside the step: the four-cell ring kernel, or the one-cell
  // kernel's 8-row small-grid shape; one receiver plane
  const bool src = rec == REC_SOURCE;
  const bool rec_kernel = (xl == 4 && ring) || (xl == 1 && !ring && ty == 8);
  const bool in_step = rec != REC_NONE && fused && c->rec_fusion && rec_kernel &&
                       form == FORM_VL &&
                       (src ? c->src_plane && c->nsrc > 0 : c->rec_plane && c->nrec > 0);
  if (in_step) {
    P.rz = src ? c->src_z : c->rec_z;
    P.nrec = src ? c->nsrc : c->nrec;
    P.rsp = src ? c->ntr_s : 1;
    P.rst = src ? 1 : c->nrec;
    P.rid = it / jt;
    P.rmap = src ? c->smap.as<int>() : c->rmap.as<int>();
    P.rtr = src ? c->strc.as<float>() : c->traces.as<float>();
    P.rscale = rec == REC_EXTRACT ? rscale : nullptr;
    std::memcpy(P.rsinc, &c->sinc[src ? 0 : 1][8 * (it % jt)], sizeof(P.rsinc));
  }
  const int kmode = !in_step ? REC_NONE : (src ? REC_INJECT : rec);
  if (xl == 1 && form == FORM_VL && ring)
    BRTM_HIP(launch_step_ring(act, fmt, P, ty, c->stream));
  else if (xl == 4)
    BRTM_HIP(launch_step_x4(act, fmt, P, ty, ring, c->stream, kmode));
  else if (xl == 2)
    BRTM_HIP(launch_step_x2(act, fmt, P, ty, c->stream));
  else
    BRTM_HIP(launch_step(form, act, fmt, P, ty, c->stream, kmode));
  BRTM_TRY(ktime_end(c, timed));
  if (fused) *fused = in_step;
  return BRTM_OK;
}

// Whether the source field's plain steps may run two per pass: the mode
// allows it, n1 % 4 == 0, and the source points (if any) can be injected
// inside the pass (one plane, one per cell).  Automatic: grids whose 64 x 32
// tiles, cut into z chunks of at least 64 slices, still give every compute
// unit a workgroup (the first step's 8 extra slices per chunk then cost at
// most an eighth).
static bool pair_ok(const brtm_ctx *c) {
  if (c->two_step == 0 || c->n1 % 4 != 0) return false;
  if (c->nsrc > 0 && !(c->src_plane && c->rec_fusion)) return false;
  if (c->two_step < 0) {
    const long tiles = (long)((c->n1 + 63) / 64) * ((c->n2 + 31) / 32);
    if (tiles * (c->n3 - 8) / 64 < c->ncu) return false;
  }
  return true;
}

// Source-field steps it and it+1 in one pass: pn1 = s^{it+1} from
// (p, po) = (s^it, s^{it-1}), pn2 = s^{it+2}, with the source points of steps
// it and it+1 injected inside (Born/gpu_funcs_3d.cu:227-250 per step).
// pn1 and pn2 are distinct from p and po and from each other.
// src = false: the pass alone, without the source points (timing).
static int step_pair(brtm_ctx *c, const float *p, const float *po, float *pn1, float *pn2,
                     int it, int jt, bool src = true) {
  StepParams P{};
  P.n1 = c->n1;
  P.n2 = c->n2;
  P.n3 = c->n3;
  P.nw = c->nw;
  std::memcpy(P.c, c->coeffs, sizeof(P.c));
  P.p = p;
  P.po = po;
  P.pn = pn1;
  P.pn2 = pn2;
  P.V = c->V.as<float>();
  P.wtab = c->d_wtab.as<float>();
  P.stream = c->stream_loads < 0 ? (16.0 * (double)c->ncell > kStreamMinStepBytes)
                                 : c->stream_loads != 0;
  int rec = REC_NONE;
  if (src && c->nsrc > 0) {
    BRTM_REQUIRE(c->src_plane && c->rec_fusion, BRTM_ESTATE,
                 "two steps per pass need the source points on one plane");
    rec = REC_INJECT;
    P.rz = c->src_z;
    P.nrec = c->nsrc;
    P.rsp = c->ntr_s;
    P.rst = 1;
    P.rid = it / jt;
    P.rid2 = (it + 1) / jt;
    P.rmap = c->smap.as<int>();
    P.rtr = c->strc.as<float>();
    std::memcpy(P.rsinc, &c->sinc[0][8 * (it % jt)], sizeof(P.rsinc));
    std::memcpy(P.rsinc2, &c->sinc[0][8 * ((it + 1) % jt)], sizeof(P.rsinc2));
  }
  const bool timed = c->ktime_on && c->ktime_form == 2;
  BRTM_TRY(ktime_begin(c, timed));
  BRTM_HIP(launch_step2(P, c->ncu, c->stream, rec));
  BRTM_TRY(ktime_end(c, timed));
  return BRTM_OK;
}

static int inject_source(brtm_ctx *c, int it, int jt, float *p) {
  if (c->nsrc == 0) return BRTM_OK;
  InjectParams I{};
  I.npts = c->nsrc;
  I.locs = c->slocs.as<int>();
  I.tr = c->strc.as<float>();
  I.sp = c->ntr_s;
  I.st = 1;
  I.id = it / jt;
  std::memcpy(I.sinc8, &c->sinc[0][8 * (it % jt)], sizeof(I.sinc8));
  I.scale = 1.f;
  I.p = p;
  I.n12 = c->n1 * c->n2;
  BRTM_HIP(launch_inject(I, c->stream));
  return BRTM_OK;
}

static int inject_data(brtm_ctx *c, int it, int jt, float *p, const void *snap, int fmt,
                       const float *sscale, float *img) {
  if (c->nrec == 0) return BRTM_OK;
  InjectParams I{};
  I.npts = c->nrec;
  I.locs = c->rlocs.as<int>();
  I.tr = c->traces.as<float>();
  I.sp = 1;
  I.st = c->nrec;
  I.id = it / jt;
  std::memcpy(I.sinc8, &c->sinc[1][8 * (it % jt)], sizeof(I.sinc8));
  I.scale = 1.f;
  I.p = p;
  I.img = img;
  I.snap = snap;
  I.sscale = sscale;
  I.fmt = fmt;
  I.n12 = c->n1 * c->n2;
  I.cimg = c->cimg;
  BRTM_HIP(launch_inject(I, c->stream));
  return BRTM_OK;
}

static int extract_data(brtm_ctx *c, int it, int jt, const float *p, const float *rscale) {
  if (c->nrec == 0) return BRTM_OK;
  BRTM_HIP(launch_extract(c->nrec, c->rlocs.as<int>(), c->traces.as<float>(), c->nrec, it / jt,
                          &c->sinc[1][8 * (it % jt)], p, rscale, c->stream));
  return BRTM_OK;
}

static int check_ready(brtm_ctx *c, int jt, int nt, int jimg) {
  BRTM_LIVE(c);
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  BRTM_REQUIRE(c->have_vel, BRTM_ESTATE, "brtm_set_velocity has not been called");
  BRTM_REQUIRE(jt >= 1 && nt >= 1 && jimg >= 1, BRTM_EINVAL, "jt, nt and jimg must be >= 1");
  BRTM_REQUIRE((int)c->sinc[0].size() >= 8 * jt && (int)c->sinc[1].size() >= 8 * jt,
               BRTM_ESTATE, "interpolation tables have fewer than jt rows");
  return BRTM_OK;
}

// ---------------------------------------------------------------------------
// Graph launch of a propagation loop.  On small grids a step is short, and
// the time between two dependent launches (about 4 us at 128^3, against 12-16
// us steps) is a fifth of the loop.  A loop whose launches are captured into
// one graph and launched together avoids most of it.  Everything a loop
// enqueues is a kernel launch: no memset, copy, event or synchronisation.
// ---------------------------------------------------------------------------
constexpr long kGraphMaxCells = 1L << 22;  // automatic: grids of up to 4 M cells

static bool graphs_for(const brtm_ctx *c) {
  if (c->ktime_on) return false;  // per-launch timing events need eager launches
  if (c->graphs >= 0) return c->graphs != 0;
  return c->ncell <= kGraphMaxCells;
}

// Free the executable graphs of finished launches (the stream is idle).
static void free_graphs(brtm_ctx *c) {
  for (hipGraphExec_t x : c->graph_execs) (void)hipGraphExecDestroy(x);
  c->graph_execs.clear();
}

template <class F>
static int run_loop(brtm_ctx *c, F &&body) {
  if (!graphs_for(c)) return body();
  if (c->graph_execs.size() >= 64) {  // bound what is kept: wait, then free
    BRTM_HIP(hipStreamSynchronize(c->stream));
    free_graphs(c);
  }
  BRTM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != BRTM_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc != BRTM_OK) return rc;
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return BRTM_EHIP;
  }
  hipGraphExec_t x = nullptr;
  hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 == hipSuccess) e2 = hipGraphLaunch(x, c->stream);
  if (x) c->graph_execs.push_back(x);
  if (e2 != hipSuccess) {
    set_error(std::string("graph launch of a propagation loop: ") + hipGetErrorString(e2));
    return BRTM_EHIP;
  }
  return BRTM_OK;
}

int source_prop(brtm_ctx *c, int jt, int nt, int jimg, int kind, int ckpt, float *p0,
                float *p1) {
  BRTM_TRY(check_ready(c, jt, nt, jimg));
  BRTM_REQUIRE(kind >= BRTM_STORE_NONE && kind <= BRTM_STORE_CKPT, BRTM_EINVAL,
               "unknown store kind");
  BRTM_REQUIRE(kind != BRTM_STORE_CKPT || ckpt >= 1, BRTM_EINVAL, "ckpt must be >= 1");
  BRTM_REQUIRE(c->nsrc == 0 || (nt - 1) / jt + 8 <= c->ntr_s, BRTM_EINVAL,
               "source traces shorter than (nt-1)/jt + 8 samples");
  Store &S = cur_store(c);
  S.owner = nullptr;  // a caller that prepares it for an operator claims it afterwards
  const size_t fbytes = sizeof(float) * (size_t)c->ncell;
  const int fmt = fmt_of_store(kind);
  // Checkpoints written straight by the steps (ckpt >= 2) need zero halos:
  // a checkpoint buffer that already has them is kept, anything else is
  // allocated afresh.
  const bool ck_direct = kind == BRTM_STORE_CKPT && ckpt >= 2;
  const int ck_count = kind == BRTM_STORE_CKPT ? (nt + ckpt - 1) / ckpt : 0;
  // FP32 snapshots are written by the steps too: the same reuse applies
  const bool direct = ck_direct || fmt == FMT_F32;
  const size_t need = ck_direct ? 2 * fbytes * (size_t)ck_count
                                : fbytes * (size_t)((nt - 1) / jimg + 1);
  const bool keep = direct && S.kind == kind && S.halo_zero && S.data.bytes >= need;
  if (!keep) S.release();
  S.kind = kind;
  S.nt = nt;
  S.jt = jt;
  S.jimg = jimg;
  S.ckpt = ckpt;
  S.count = 0;
  if (fmt != FMT_NONE) {
    S.count = (nt - 1) / jimg + 1;
    BRTM_TRY(S.data.alloc(fmt_bytes(fmt) * (size_t)c->ncell * S.count));
    if (fmt != FMT_F32) BRTM_TRY(S.scale.alloc(sizeof(float) * (size_t)c->n3 * S.count));
    // FP32 snapshots are written by the steps themselves and read back as
    // the next step's input, so their halos must be zero (once per buffer)
    if (fmt == FMT_F32 && !keep) {
      BRTM_HIP(launch_zero_halo(S.data.as<float>(), c->ncell, (int)(S.data.bytes / fbytes),
                                c->n1, c->n2, c->n3, c->stream));
      S.halo_zero = true;
    }
  } else if (kind == BRTM_STORE_CKPT) {
    S.count = ck_count;
    BRTM_TRY(S.data.alloc(2 * fbytes * S.count));
  }
  float *const wx = c->fb.as<float>(), *const wy = c->fa.as<float>();
  // two more work buffers for steps run two per pass, whose outputs may not
  // overwrite their inputs (the receiver-field buffers, idle here)
  float *const wz = c->fc.as<float>(), *const ww = c->fd.as<float>();
  const bool pairs = pair_ok(c) && !(kind == BRTM_STORE_CKPT && !ck_direct);
  float *cur = wx, *old = wy;  // s^it, s^{it-1}
  // Checkpoint pair K = (s^{K ckpt}, s^{K ckpt - 1}).  With ckpt >= 2 the
  // two steps before each checkpoint write straight into its pair, and the
  // steps that read a pair write out of place into a work buffer.
  auto ck_cur = [&](int K) {
    return reinterpret_cast<float *>(S.data.as<char>() + 2 * fbytes * (size_t)K);
  };
  if (ck_direct) {
    // the steps write only the update set: the pairs' halos must be zero
    // like the work buffers', and pair 0 is (s^0, s^-1) = 0.  The work
    // buffers are written out of place here without a full clear, so their
    // halos are cleared too (other calls may have filled them).
    BRTM_HIP(hipMemsetAsync(ck_cur(0), 0, 2 * fbytes, c->stream));
    BRTM_HIP(launch_zero_halo(wx, c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
    BRTM_HIP(launch_zero_halo(wy, c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
    if (!keep) {
      const size_t nf = S.data.bytes / fbytes;  // whole buffer, also beyond count pairs
      BRTM_HIP(launch_zero_halo(ck_cur(0) + 2 * c->ncell, c->ncell, (int)nf - 2, c->n1, c->n2,
                                c->n3, c->stream));
      S.halo_zero = true;
    }
    cur = ck_cur(0);
    old = ck_cur(0) + c->ncell;
  } else {
    BRTM_HIP(hipMemsetAsync(cur, 0, fbytes, c->stream));
    BRTM_HIP(hipMemsetAsync(old, 0, fbytes, c->stream));
  }
  // the pass writes only the update set of its outputs
  if (pairs) {
    BRTM_HIP(launch_zero_halo(wz, c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
    BRTM_HIP(launch_zero_halo(ww, c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
  }
  const int n12 = c->n1 * c->n2;
  // where the output of step it must go (a checkpoint half or an FP32
  // snapshot), or null: any free work buffer
  auto fixed_out = [&](int it) -> float * {
    if (ck_direct) {
      if ((it + 1) % ckpt == 0 && (it + 1) / ckpt < S.count) return ck_cur((it + 1) / ckpt);
      if ((it + 2) % ckpt == 0 && (it + 2) / ckpt < S.count)
        return ck_cur((it + 2) / ckpt) + c->ncell;
    } else if (fmt == FMT_F32 && it % jimg == 0) {
      return S.data.as<float>() + (size_t)c->ncell * (it / jimg);
    }
    return nullptr;
  };
  auto free_work = [&](const float *a, const float *b, const float *d) -> float * {
    for (float *w : {wx, wy, wz, ww})
      if (w != a && w != b && w != d) return w;
    return nullptr;
  };
  auto encode = [&](int it, const float *out) -> int {
    if (fmt != FMT_NONE && fmt != FMT_F32 && it % jimg == 0) {
      const int k = it / jimg;
      BRTM_HIP(launch_encode(fmt, out, S.data.as<char>() + 2 * (size_t)c->ncell * k,
                             S.scale.as<float>() + (size_t)c->n3 * k, n12, c->n3, c->stream));
    }
    return BRTM_OK;
  };
  auto loop = [&]() -> int {
  for (int it = 0; it < nt; ++it) {
    if (pairs && it % 2 == 0 && it + 1 < nt) {
      float *o1 = fixed_out(it);
      if (!o1) o1 = free_work(cur, old, nullptr);
      float *o2 = fixed_out(it + 1);
      if (!o2) o2 = free_work(cur, old, o1);
      BRTM_TRY(step_pair(c, cur, old, o1, o2, it, jt));
      BRTM_TRY(encode(it, o1));
      BRTM_TRY(encode(it + 1, o2));
      old = o1;
      cur = o2;
      ++it;
      continue;
    }
    if (kind == BRTM_STORE_CKPT && !ck_direct && it % ckpt == 0) {
      char *base = S.data.as<char>() + 2 * fbytes * (size_t)(it / ckpt);
      BRTM_HIP(hipMemcpyAsync(base, cur, fbytes, hipMemcpyDeviceToDevice, c->stream));
      BRTM_HIP(hipMemcpyAsync(base + fbytes, old, fbytes, hipMemcpyDeviceToDevice, c->stream));
    }
    float *out = old;  // in place over s^{it-1}
    if (ck_direct) {
      if ((it + 1) % ckpt == 0 && (it + 1) / ckpt < S.count)
        out = ck_cur((it + 1) / ckpt);                 // s^{it+1} opens checkpoint K
      else if ((it + 2) % ckpt == 0 && (it + 2) / ckpt < S.count)
        out = ck_cur((it + 2) / ckpt) + c->ncell;      // s^{it+1} = s^{K ckpt - 1}
      else if (old != wx && old != wy && old != wz && old != ww)
        out = free_work(cur, old, nullptr);            // s^{it-1} is a checkpoint: keep it
    } else if (fmt == FMT_F32) {
      if (it % jimg == 0)
        out = S.data.as<float>() + (size_t)c->ncell * (it / jimg);  // snapshot k itself
      else if (old != wx && old != wy && old != wz && old != ww)
        out = free_work(cur, old, nullptr);            // s^{it-1} is a snapshot: keep it
    }
    bool fused = false;
    BRTM_TRY(step(c, FORM_VL, ACT_NONE, FMT_NONE, cur, old, out, nullptr, nullptr, nullptr,
                  nullptr, REC_SOURCE, it, jt, nullptr, &fused));
    if (!fused) BRTM_TRY(inject_source(c, it, jt, out));
    BRTM_TRY(encode(it, out));  // (FP32: the step wrote snapshot k)
    old = cur;
    cur = out;
  }
  return BRTM_OK;
  };
  // (checkpoints every step copy the pair each step: that loop runs eagerly)
  BRTM_TRY(kind == BRTM_STORE_CKPT && !ck_direct ? loop() : run_loop(c, loop));
  if (p0) BRTM_HIP(hipMemcpyAsync(p0, old, fbytes, hipMemcpyDeviceToHost, c->stream));
  if (p1) BRTM_HIP(hipMemcpyAsync(p1, cur, fbytes, hipMemcpyDeviceToHost, c->stream));
  if (p0 || p1) BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

int born_forward(brtm_ctx *c, int jt, int nt, int jimg, bool accumulate) {
  BRTM_TRY(check_ready(c, jt, nt, jimg));
  BRTM_REQUIRE(c->nrec == 0 || (nt - 1) / jt + 8 <= c->ntr_d, BRTM_EINVAL,
               "receiver traces shorter than (nt-1)/jt + 8 samples");
  Store &S = cur_store(c);
  const int sfmt = fmt_of_store(S.kind);
  const bool live = sfmt == FMT_NONE;
  if (!live) {
    BRTM_REQUIRE(S.nt == nt && S.jimg == jimg && S.jt == jt, BRTM_ESTATE,
                 "stored source wavefield was made with other nt/jt/jimg");
  }
  const size_t fbytes = sizeof(float) * (size_t)c->ncell;
  if (!accumulate && c->nrec)
    BRTM_HIP(hipMemsetAsync(c->traces.ptr, 0, sizeof(float) * (size_t)c->nrec * c->ntr_d,
                            c->stream));
  // The transpose-form step is a diagonal similarity of the forward form:
  // A_LV = D^-1 A_VL D with D = V.w (and the w^2 term commutes with D).  So
  // the scattered field q is advanced as u = D q with the VL kernel: the
  // Born source becomes cimg.s.(D m) and receivers read u / D.  The direct
  // LV kernel path is kept (born_direct_lv) as a cross-check.
  const bool direct = c->born_direct_lv;
  const float *mdl = c->mdl.as<float>();
  const float *rscale = nullptr;
  if (!direct) {
    BRTM_TRY(c->mvw.alloc(fbytes));
    BRTM_HIP(launch_vw_scale(c->V.as<float>(), c->d_wtab.as<float>(), c->nw, mdl,
                             c->mvw.as<float>(), c->n1, c->n2, c->n3, c->stream));
    mdl = c->mvw.as<float>();
    if (c->nrec) {
      BRTM_TRY(c->rscale.alloc(sizeof(float) * (size_t)c->nrec));
      BRTM_HIP(launch_receiver_scale(c->V.as<float>(), c->d_wtab.as<float>(), c->nw,
                                     c->rlocs.as<int>(), c->nrec, c->rscale.as<float>(), c->n1,
                                     c->n2, c->n3, c->stream));
      rscale = c->rscale.as<float>();
    }
  }
  const int qform = direct ? FORM_LV : FORM_VL;
  float *qc = c->fd.as<float>(), *qo = c->fc.as<float>();
  float *sc = c->fb.as<float>(), *so = c->fa.as<float>();
  BRTM_HIP(hipMemsetAsync(qc, 0, fbytes, c->stream));
  BRTM_HIP(hipMemsetAsync(qo, 0, fbytes, c->stream));
  if (live) {
    BRTM_HIP(hipMemsetAsync(sc, 0, fbytes, c->stream));
    BRTM_HIP(hipMemsetAsync(so, 0, fbytes, c->stream));
  }
  BRTM_TRY(run_loop(c, [&]() -> int {
  for (int it = 0; it < nt; ++it) {
    bool fused = false;
    if (live) {
      bool sfused = false;
      BRTM_TRY(step(c, FORM_VL, ACT_NONE, FMT_NONE, sc, so, so, nullptr, nullptr, nullptr,
                    nullptr, REC_SOURCE, it, jt, nullptr, &sfused));
      if (!sfused) BRTM_TRY(inject_source(c, it, jt, so));
    }
    if (it % jimg == 0) {
      const int k = it / jimg;
      const void *snap;
      const float *ss = nullptr;
      int f;
      if (live) {
        snap = so;
        f = FMT_F32;
      } else {
        f = sfmt;
        snap = S.data.as<char>() + fmt_bytes(f) * (size_t)c->ncell * k;
        if (f != FMT_F32) ss = S.scale.as<float>() + (size_t)c->n3 * k;
      }
      BRTM_TRY(step(c, qform, ACT_BORN, f, qc, qo, qo, snap, ss, nullptr, mdl, REC_EXTRACT, it,
                    jt, rscale, &fused));
    } else {
      BRTM_TRY(step(c, qform, ACT_NONE, FMT_NONE, qc, qo, qo, nullptr, nullptr, nullptr,
                    nullptr, REC_EXTRACT, it, jt, rscale, &fused));
    }
    if (!fused) BRTM_TRY(extract_data(c, it, jt, qo, rscale));
    std::swap(qc, qo);
    if (live) std::swap(sc, so);
  }
  return BRTM_OK;
  }));
  return BRTM_OK;
}

int rtm_adjoint(brtm_ctx *c, int jt, int nt, int jimg) {
  BRTM_TRY(check_ready(c, jt, nt, jimg));
  BRTM_REQUIRE(c->nrec == 0 || (nt - 1) / jt + 8 <= c->ntr_d, BRTM_EINVAL,
               "receiver traces shorter than (nt-1)/jt + 8 samples");
  Store &S = cur_store(c);
  BRTM_REQUIRE(S.kind != BRTM_STORE_NONE, BRTM_ESTATE,
               "rtm_adjoint needs a stored source wavefield (brtm_source_prop with a store)");
  BRTM_REQUIRE(S.nt == nt && S.jimg == jimg && S.jt == jt, BRTM_ESTATE,
               "stored source wavefield was made with other nt/jt/jimg");
  const size_t fbytes = sizeof(float) * (size_t)c->ncell;
  float *lc = c->fd.as<float>(), *lo = c->fc.as<float>();
  BRTM_HIP(hipMemsetAsync(lc, 0, fbytes, c->stream));
  BRTM_HIP(hipMemsetAsync(lo, 0, fbytes, c->stream));
  float *img = c->img.as<float>();
  const int sfmt = fmt_of_store(S.kind);
  if (sfmt != FMT_NONE) {
    return run_loop(c, [&]() -> int {
    for (int it = nt - 1; it >= 0; --it) {
      if (it % jimg == 0) {
        const int k = it / jimg;
        const void *snap = S.data.as<char>() + fmt_bytes(sfmt) * (size_t)c->ncell * k;
        const float *ss = sfmt == FMT_F32 ? nullptr : S.scale.as<float>() + (size_t)c->n3 * k;
        bool fused = false;
        BRTM_TRY(step(c, FORM_VL, ACT_IMG, sfmt, lc, lo, lo, snap, ss, img, nullptr, REC_INJECT,
                      it, jt, nullptr, &fused));
        if (!fused) BRTM_TRY(inject_data(c, it, jt, lo, snap, sfmt, ss, img));
      } else {
        bool fused = false;
        BRTM_TRY(step(c, FORM_VL, ACT_NONE, FMT_NONE, lc, lo, lo, nullptr, nullptr, nullptr,
                      nullptr, REC_INJECT, it, jt, nullptr, &fused));
        if (!fused) BRTM_TRY(inject_data(c, it, jt, lo, nullptr, FMT_NONE, nullptr, nullptr));
      }
      std::swap(lc, lo);
    }
    return BRTM_OK;
    });
  }
  // Checkpoints: recompute each segment of the source field out of place
  // into seg slots 0..ckpt-1 (slot j holds s^{k0+1+j}), starting from the
  // checkpoint pair itself (no copy), then step the receiver field back
  // through the segment.
  const int ck = S.ckpt;
  if (c->seg_slots < ck) {
    // the recompute writes only the update set: keep the slots' halos zero
    BRTM_TRY(c->seg.alloc(fbytes * (size_t)ck));
    BRTM_HIP(launch_zero_halo(c->seg.as<float>(), c->ncell, ck, c->n1, c->n2, c->n3, c->stream));
    c->seg_slots = ck;
  }
  auto slot = [&](int j) { return c->seg.as<float>() + (size_t)c->ncell * j; };
  // the recompute pairs the steps as the source propagation did (it even)
  const bool pairs = pair_ok(c);
  return run_loop(c, [&]() -> int {
  for (int cidx = S.count - 1; cidx >= 0; --cidx) {
    const int k0 = cidx * ck, k1 = std::min(k0 + ck, nt);
    const float *ck_p = reinterpret_cast<const float *>(S.data.as<char>() +
                                                        2 * fbytes * (size_t)cidx);  // s^{k0}
    const float *ck_po = ck_p + c->ncell;                                            // s^{k0-1}
    for (int it = k0; it < k1; ++it) {
      const int j = it - k0;
      const float *p = j == 0 ? ck_p : slot(j - 1);
      const float *po = j == 0 ? ck_po : (j == 1 ? ck_p : slot(j - 2));
      if (pairs && it % 2 == 0 && it + 1 < k1) {
        BRTM_TRY(step_pair(c, p, po, slot(j), slot(j + 1), it, jt));
        ++it;
        continue;
      }
      bool fused = false;
      BRTM_TRY(step(c, FORM_VL, ACT_NONE, FMT_NONE, p, po, slot(j), nullptr, nullptr, nullptr,
                    nullptr, REC_SOURCE, it, jt, nullptr, &fused));
      if (!fused) BRTM_TRY(inject_source(c, it, jt, slot(j)));
    }
    for (int it = k1 - 1; it >= k0; --it) {
      if (it % jimg == 0) {
        const float *snap = slot(it - k0);
        bool fused = false;
        BRTM_TRY(step(c, FORM_VL, ACT_IMG, FMT_F32, lc, lo, lo, snap, nullptr, img, nullptr,
                      REC_INJECT, it, jt, nullptr, &fused));
        if (!fused) BRTM_TRY(inject_data(c, it, jt, lo, snap, FMT_F32, nullptr, img));
      } else {
        bool fused = false;
        BRTM_TRY(step(c, FORM_VL, ACT_NONE, FMT_NONE, lc, lo, lo, nullptr, nullptr, nullptr,
                      nullptr, REC_INJECT, it, jt, nullptr, &fused));
        if (!fused) BRTM_TRY(inject_data(c, it, jt, lo, nullptr, FMT_NONE, nullptr, nullptr));
      }
      std::swap(lc, lo);
    }
  }
  return BRTM_OK;
  });
}

int dot(brtm_ctx *c, const float *a, const float *b, long n, double *out) {
  BRTM_TRY(c->partials.alloc(sizeof(double) * kDotBlocks));
  BRTM_TRY(c->dotv.alloc(sizeof(double)));
  BRTM_HIP(launch_dot(a, b, n, c->partials.as<double>(), c->dotv.as<double>(), c->stream));
  BRTM_HIP(hipMemcpyAsync(out, c->dotv.ptr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};  // address of `scratch` on each device

// The address of `scratch` on `device`, which must be the current device:
// looked up once per device and kept.
int noalloc_scratch(int device, double **out) {
  BRTM_REQUIRE(device >= 0 && device < kMaxDevices, BRTM_EINVAL,
               "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    BRTM_REQUIRE(e == hipSuccess && p, BRTM_EHIP,
                 std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(p);
  }
  *out = g_scratch[device];
  return BRTM_OK;
}
}  // namespace

void ctx_retain(brtm_ctx *c) { c->refs.fetch_add(1); }

void ctx_release(brtm_ctx *c) {
  if (c->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_graphs(c);
  for (hipEvent_t e : c->ktime_ev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // namespace brtm

using namespace brtm;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *brtm_last_error(void) { return g_err.c_str(); }
const char *brtm_version(void) { return "sep_code_amd born_rtm 0.1 (gfx950)"; }

int brtm_create(int device, brtm_ctx **out) {
  BRTM_REQUIRE(out, BRTM_EINVAL, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  BRTM_HIP(hipGetDeviceCount(&ndev));
  BRTM_REQUIRE(device >= 0 && device < ndev, BRTM_EINVAL, "no such device");
  BRTM_HIP(hipSetDevice(device));
  brtm_ctx *c = new brtm_ctx();
  c->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      ncu > 0)
    c->ncu = ncu;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    set_error("stream/event creation failed");
    return BRTM_EHIP;
  }
  // the allocation-free entry point's scratch on this device (current above)
  const int rc = noalloc_scratch(device, &c->noalloc_scratch);
  if (rc != BRTM_OK) {
    ctx_release(c);
    return rc;
  }
  *out = c;
  return BRTM_OK;
}

int brtm_destroy(brtm_ctx *c) {
  if (!c) return BRTM_OK;
  BRTM_REQUIRE(!c->destroyed, BRTM_ESTATE, "brtm_destroy called twice on one context");
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  // The device memory goes now; objects still alive on the context keep
  // only its host record (stream, events) until they are destroyed.
  c->stores.clear();
  for (DevBuf *b : {&c->d_wtab, &c->V, &c->fa, &c->fb, &c->fc, &c->fd, &c->img, &c->mdl,
                    &c->seg, &c->mvw, &c->rscale, &c->partials, &c->dotv, &c->slocs, &c->strc,
                    &c->rlocs, &c->rmap, &c->smap, &c->traces})
    b->release();
  c->have_space = c->have_vel = false;
  c->destroyed = true;
  ctx_release(c);
  return BRTM_OK;
}

int brtm_create_space(brtm_ctx *c, float d1, float d2, float d3, float bc_a, float bc_b, int n1,
                      int n2, int n3) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(n1 >= 9 && n2 >= 9 && n3 >= 9, BRTM_EINVAL, "each axis needs at least 9 cells");
  BRTM_REQUIRE((double)n1 * n2 * n3 < 2147483647.0, BRTM_EINVAL,
               "grid exceeds 2^31 cells (32-bit cell indexing)");
  BRTM_REQUIRE(d1 > 0 && d2 > 0 && d3 > 0, BRTM_EINVAL, "grid spacing must be positive");
  BRTM_REQUIRE(bc_a >= 0 && bc_a <= (float)kWMax && bc_b >= 0, BRTM_EINVAL,
               "sponge parameters out of range (0 <= bc_a <= 256, bc_b >= 0)");
  BRTM_HIP(hipSetDevice(c->device));
  BRTM_HIP(hipStreamSynchronize(c->stream));  // buffers below are reallocated
  c->n1 = n1;
  c->n2 = n2;
  c->n3 = n3;
  c->ncell = (long)n1 * n2 * n3;
  c->d[0] = d1;
  c->d[1] = d2;
  c->d[2] = d3;
  c->bc_a = bc_a;
  c->bc_b = bc_b;
  stencil_coeffs(d1, d2, d3, c->coeffs);
  sponge_weights(bc_a, bc_b, &c->nw, &c->wtab);
  const size_t fbytes = sizeof(float) * (size_t)c->ncell;
  BRTM_TRY(c->d_wtab.alloc(sizeof(float) * c->wtab.size()));
  BRTM_HIP(h2d(c, c->d_wtab.ptr, c->wtab.data(), sizeof(float) * c->wtab.size()));
  for (DevBuf *b : {&c->V, &c->fa, &c->fb, &c->fc, &c->fd, &c->img, &c->mdl}) {
    b->release();
    BRTM_TRY(b->alloc(fbytes));
    BRTM_HIP(hipMemsetAsync(b->ptr, 0, fbytes, c->stream));
  }
  c->seg.release();
  c->seg_slots = 0;
  for (auto &s : c->stores) {
    s.release();
    s.kind = BRTM_STORE_NONE;
  }
  // Source points and receivers index the previous grid: a new grid starts
  // with none (brtm_set_source / brtm_set_receivers give them again).
  c->nsrc = 0;
  c->src_plane = false;
  c->src_z = -1;
  c->nrec = c->nrx = c->nry = 0;
  c->rec_plane = false;
  c->rec_z = -1;
  c->h_rlocs.clear();
  BRTM_HIP(hipStreamSynchronize(c->stream));
  c->have_space = true;
  c->have_vel = false;
  if (c->sinc[0].empty()) c->sinc[0] = {0, 0, 0, 1, 0, 0, 0, 0};
  if (c->sinc[1].empty()) c->sinc[1] = {0, 0, 0, 1, 0, 0, 0, 0};
  return BRTM_OK;
}

int brtm_set_velocity(brtm_ctx *c, const float *V) {
  BRTM_REQUIRE(c && V, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  // Leapfrog limit: max V * lambda_max(-L8) <= 4, lambda_max(-L8) = sum_a
  // C_a (1025/576 + 2 (1 + 1/8 + 1/63 + 1/896)) (symbol of L8 at pi).
  // V must be positive on the update set: Born modelling advances
  // u = V.w.q, which determines q only where V.w > 0.
  double vmax = 0;
  for (long i = 0; i < c->ncell; ++i) {
    const double v = V[i];
    BRTM_REQUIRE(std::isfinite(v) && v >= 0, BRTM_EINVAL, "V must be finite and >= 0");
    vmax = std::max(vmax, v);
  }
  for (int i3 = 4; i3 < c->n3 - 4; ++i3)
    for (int i2 = 4; i2 < c->n2 - 4; ++i2) {
      const float *row = V + ((long)i3 * c->n2 + i2) * c->n1;
      for (int i1 = 4; i1 < c->n1 - 4; ++i1)
        BRTM_REQUIRE(row[i1] > 0.f, BRTM_EINVAL, "V must be > 0 on the update set");
    }
  const double S = 1025.0 / 576.0 + 2.0 * (1.0 + 1.0 / 8 + 1.0 / 63 + 1.0 / 896);
  double lam = 0;
  for (int a = 0; a < 3; ++a) lam += 8.0 / (5.0 * (double)c->d[a] * c->d[a]) * S;
  BRTM_REQUIRE(vmax * lam <= 4.0, BRTM_EUNSTABLE,
               "velocity violates the 3-D 8th-order leapfrog stability limit "
               "(max v dt / d above 0.4529 for equal spacing)");
  BRTM_HIP(hipSetDevice(c->device));
  BRTM_HIP(h2d(c, c->V.ptr, V, sizeof(float) * (size_t)c->ncell));
  c->have_vel = true;
  return BRTM_OK;
}

int brtm_set_sinc_table(brtm_ctx *c, int which, int nsinc, int ntab, const float *table) {
  BRTM_REQUIRE(c && table, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(which == 0 || which == 1, BRTM_EINVAL, "which must be 0 (source) or 1 (data)");
  BRTM_REQUIRE(nsinc == 8, BRTM_EINVAL, "only 8-tap interpolation tables are supported");
  BRTM_REQUIRE(ntab >= 1, BRTM_EINVAL, "ntab must be >= 1");
  c->sinc[which].assign(table, table + 8 * ntab);
  return BRTM_OK;
}

// When the active points (loc >= 0) all lie in one depth plane with at most
// one per cell, *ok is set, *z names the plane and map [n2][n1] holds each
// cell's point index (-1: none); otherwise *ok is cleared.
static int plane_map(brtm_ctx *c, const int *locs, int n, bool *ok, int *z, DevBuf *map) {
  const long n12 = (long)c->n1 * c->n2;
  std::vector<int> m(n12, -1);
  int pz = -1;
  bool good = true;
  for (int j = 0; j < n && good; ++j) {
    const long l = locs[j];
    if (l < 0) continue;
    const int lz = (int)(l / n12);
    if (pz < 0) pz = lz;
    if (lz != pz || m[l % n12] >= 0) good = false;
    else m[l % n12] = j;
  }
  *ok = good && pz >= 0;
  *z = *ok ? pz : -1;
  if (*ok) {
    BRTM_TRY(map->alloc(sizeof(int) * (size_t)n12));
    BRTM_HIP(h2d(c, map->ptr, m.data(), sizeof(int) * (size_t)n12));
  }
  return BRTM_OK;
}

static bool in_update_set(const brtm_ctx *c, long loc) {
  if (loc < 0 || loc >= c->ncell) return false;
  const int i1 = (int)(loc % c->n1), i2 = (int)((loc / c->n1) % c->n2),
            i3 = (int)(loc / ((long)c->n1 * c->n2));
  return i1 >= 4 && i1 < c->n1 - 4 && i2 >= 4 && i2 < c->n2 - 4 && i3 >= 4 && i3 < c->n3 - 4;
}

int brtm_set_source(brtm_ctx *c, int npts, int ntr, const int *locs, const float *vals) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  BRTM_REQUIRE(npts >= 0 && ntr >= 1 && (npts == 0 || (locs && vals)), BRTM_EINVAL,
               "bad source arguments");
  for (int j = 0; j < npts; ++j)
    BRTM_REQUIRE(in_update_set(c, locs[j]), BRTM_EINVAL,
                 "source point " + std::to_string(j) + " outside the update set");
  BRTM_HIP(hipSetDevice(c->device));
  // queued work may still read the current source buffers
  BRTM_HIP(hipStreamSynchronize(c->stream));
  c->nsrc = npts;
  c->ntr_s = ntr;
  if (npts == 0) return BRTM_OK;
  BRTM_TRY(c->slocs.alloc(sizeof(int) * npts));
  BRTM_TRY(c->strc.alloc(sizeof(float) * (size_t)npts * ntr));
  BRTM_HIP(h2d(c, c->slocs.ptr, locs, sizeof(int) * npts));
  BRTM_HIP(h2d(c, c->strc.ptr, vals, sizeof(float) * (size_t)npts * ntr));
  BRTM_TRY(plane_map(c, locs, npts, &c->src_plane, &c->src_z, &c->smap));
  return BRTM_OK;
}

int brtm_set_receivers(brtm_ctx *c, int nx, int ny, int ntr, const int *locs, const float *vals) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  BRTM_REQUIRE(nx >= 0 && ny >= 0 && ntr >= 8 && (nx * ny == 0 || locs), BRTM_EINVAL,
               "bad receiver arguments");
  BRTM_HIP(hipSetDevice(c->device));
  // queued work may still read the current receiver buffers
  BRTM_HIP(hipStreamSynchronize(c->stream));
  const int nrec = nx * ny;
  c->nrx = nx;
  c->nry = ny;
  c->nrec = nrec;
  c->ntr_d = ntr;
  if (nrec == 0) return BRTM_OK;
  // Receivers outside the update set read a field that is always zero and
  // would write into cells no step updates: mark them inactive.
  c->h_rlocs.assign(locs, locs + nrec);
  for (auto &l : c->h_rlocs)
    if (!in_update_set(c, l)) l = -1;
  BRTM_TRY(c->rlocs.alloc(sizeof(int) * nrec));
  BRTM_HIP(h2d(c, c->rlocs.ptr, c->h_rlocs.data(), sizeof(int) * nrec));
  // One plane, one receiver per cell: the step kernel can take them.
  BRTM_TRY(plane_map(c, c->h_rlocs.data(), nrec, &c->rec_plane, &c->rec_z, &c->rmap));
  const size_t tb = sizeof(float) * (size_t)nrec * ntr;
  BRTM_TRY(c->traces.alloc(tb));
  if (!vals) {
    BRTM_HIP(hipMemsetAsync(c->traces.ptr, 0, tb, c->stream));
  } else {
    DevBuf tmp;
    BRTM_TRY(tmp.alloc(tb));
    BRTM_HIP(h2d(c, tmp.ptr, vals, tb));
    BRTM_HIP(launch_transpose(tmp.as<float>(), c->traces.as<float>(), nrec, ntr, c->stream));
    BRTM_HIP(hipStreamSynchronize(c->stream));
    tmp.release();
  }
  BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

int brtm_get_receivers(brtm_ctx *c, float *vals) {
  BRTM_REQUIRE(c && vals, BRTM_EINVAL, "NULL argument");
  if (c->nrec == 0) return BRTM_OK;
  BRTM_HIP(hipSetDevice(c->device));
  const size_t tb = sizeof(float) * (size_t)c->nrec * c->ntr_d;
  DevBuf tmp;
  BRTM_TRY(tmp.alloc(tb));
  BRTM_HIP(launch_transpose(c->traces.as<float>(), tmp.as<float>(), c->ntr_d, c->nrec, c->stream));
  BRTM_HIP(hipMemcpyAsync(vals, tmp.ptr, tb, hipMemcpyDeviceToHost, c->stream));
  BRTM_HIP(hipStreamSynchronize(c->stream));
  tmp.release();
  return BRTM_OK;
}

int brtm_set_model(brtm_ctx *c, const float *m) {
  BRTM_REQUIRE(c && m, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  BRTM_HIP(hipSetDevice(c->device));
  BRTM_HIP(h2d(c, c->mdl.ptr, m, sizeof(float) * (size_t)c->ncell));
  return BRTM_OK;
}

int brtm_get_image(brtm_ctx *c, float *img) {
  BRTM_REQUIRE(c && img, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  BRTM_HIP(hipSetDevice(c->device));
  BRTM_HIP(hipMemcpyAsync(img, c->img.ptr, sizeof(float) * (size_t)c->ncell,
                          hipMemcpyDeviceToHost, c->stream));
  BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

int brtm_zero_image(brtm_ctx *c) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(c->have_space, BRTM_ESTATE, "brtm_create_space has not been called");
  BRTM_HIP(hipMemsetAsync(c->img.ptr, 0, sizeof(float) * (size_t)c->ncell, c->stream));
  return BRTM_OK;
}

int brtm_select_shot(brtm_ctx *c, int ishot) {
  BRTM_REQUIRE(c && ishot >= 0, BRTM_EINVAL, "bad shot index");
  c->shot = ishot;
  cur_store(c);
  return BRTM_OK;
}

int brtm_release_store(brtm_ctx *c, int ishot) {
  BRTM_REQUIRE(c && ishot >= 0, BRTM_EINVAL, "bad shot index");
  if (ishot < (int)c->stores.size()) {
    c->stores[ishot].release();
    c->stores[ishot].kind = BRTM_STORE_NONE;
  }
  return BRTM_OK;
}

int brtm_source_prop(brtm_ctx *c, int jt, int nt, int jimg, int store, int ckpt, float *p0,
                     float *p1) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_HIP(hipSetDevice(c->device));
  return source_prop(c, jt, nt, jimg, store, ckpt, p0, p1);
}

int brtm_rtm_forward(brtm_ctx *c, int jt, int nt, int jimg, int accumulate) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_HIP(hipSetDevice(c->device));
  return born_forward(c, jt, nt, jimg, accumulate != 0);
}

int brtm_rtm_adjoint(brtm_ctx *c, int jt, int nt, int jimg) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_HIP(hipSetDevice(c->device));
  return rtm_adjoint(c, jt, nt, jimg);
}

int brtm_set_imaging_scale(brtm_ctx *c, float cimg) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  c->cimg = cimg;
  return BRTM_OK;
}

int brtm_get_snapshot(brtm_ctx *c, int k, float *out) {
  BRTM_REQUIRE(c && out, BRTM_EINVAL, "NULL argument");
  Store &S = cur_store(c);
  const int fmt = fmt_of_store(S.kind);
  BRTM_REQUIRE(fmt != FMT_NONE, BRTM_ESTATE, "selected shot has no snapshot store");
  BRTM_REQUIRE(k >= 0 && k < S.count, BRTM_EINVAL, "snapshot index out of range");
  BRTM_HIP(hipSetDevice(c->device));
  const size_t fbytes = sizeof(float) * (size_t)c->ncell;
  if (fmt == FMT_F32) {
    BRTM_HIP(hipMemcpyAsync(out, S.data.as<char>() + fbytes * (size_t)k, fbytes,
                            hipMemcpyDeviceToHost, c->stream));
  } else {
    DevBuf tmp;
    BRTM_TRY(tmp.alloc(fbytes));
    BRTM_HIP(launch_decode(fmt, S.data.as<char>() + 2 * (size_t)c->ncell * k,
                           S.scale.as<float>() + (size_t)c->n3 * k, tmp.as<float>(),
                           c->n1 * c->n2, c->n3, c->stream));
    BRTM_HIP(hipMemcpyAsync(out, tmp.ptr, fbytes, hipMemcpyDeviceToHost, c->stream));
    BRTM_HIP(hipStreamSynchronize(c->stream));
  }
  BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

int brtm_propagate(brtm_ctx *c, int form, int nsteps, float *p_cur, float *p_old) {
  BRTM_REQUIRE(c && p_cur && p_old, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(form == 0 || form == 1, BRTM_EINVAL, "form must be 0 (VL) or 1 (LV)");
  BRTM_REQUIRE(c->have_space && c->have_vel, BRTM_ESTATE, "space/velocity not set");
  BRTM_HIP(hipSetDevice(c->device));
  const size_t fbytes = sizeof(float) * (size_t)c->ncell;
  float *cur = c->fb.as<float>(), *old = c->fa.as<float>();
  BRTM_HIP(hipMemcpyAsync(cur, p_cur, fbytes, hipMemcpyHostToDevice, c->stream));
  BRTM_HIP(hipMemcpyAsync(old, p_old, fbytes, hipMemcpyHostToDevice, c->stream));
  for (int s = 0; s < nsteps; ++s) {
    BRTM_TRY(
        step(c, form, ACT_NONE, FMT_NONE, cur, old, old, nullptr, nullptr, nullptr, nullptr));
    std::swap(cur, old);
  }
  BRTM_HIP(hipMemcpyAsync(p_cur, cur, fbytes, hipMemcpyDeviceToHost, c->stream));
  BRTM_HIP(hipMemcpyAsync(p_old, old, fbytes, hipMemcpyDeviceToHost, c->stream));
  // the caller's fields may have had nonzero halos: leave the work buffers
  // with zero halos, as every propagation expects of them
  BRTM_HIP(launch_zero_halo(cur, c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
  BRTM_HIP(launch_zero_halo(old, c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
  BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

int brtm_time_step(brtm_ctx *c, int form, int act, int fmt, int nrep, double *avg_ms) {
  BRTM_REQUIRE(c && avg_ms && nrep >= 1, BRTM_EINVAL, "bad arguments");
  BRTM_REQUIRE(form >= 0 && form <= 2, BRTM_EINVAL, "form must be 0, 1 or 2");
  BRTM_REQUIRE(form != 2 || (act == ACT_NONE && c->n1 % 4 == 0), BRTM_EINVAL,
               "the two-step pass is plain and needs n1 % 4 == 0");
  BRTM_REQUIRE(act >= ACT_NONE && act <= ACT_BORN, BRTM_EINVAL, "act must be 0..2");
  BRTM_REQUIRE(fmt >= FMT_NONE && fmt <= FMT_F16, BRTM_EINVAL, "fmt must be 0..3");
  BRTM_REQUIRE((act == ACT_NONE) == (fmt == FMT_NONE), BRTM_EINVAL,
               "an action needs a snapshot format and vice versa");
  BRTM_REQUIRE(c->have_space && c->have_vel, BRTM_ESTATE, "space/velocity not set");
  BRTM_HIP(hipSetDevice(c->device));
  DevBuf sc;
  BRTM_TRY(sc.alloc(sizeof(float) * c->n3));
  BRTM_HIP(launch_fill(sc.as<float>(), 1.f, c->n3, c->stream));
  float *cur = c->fb.as<float>(), *old = c->fa.as<float>();
  const void *snap = c->fd.ptr;  // any resident field serves as the snapshot
  float *img = act == ACT_IMG ? c->img.as<float>() : nullptr;
  const float *m = act == ACT_BORN ? c->mdl.as<float>() : nullptr;
  if (form == 2) {
    // two steps per launch over four buffers: (p, po) -> (pn1, pn2)
    float *b[4] = {cur, old, c->fc.as<float>(), c->fd.as<float>()};
    BRTM_HIP(launch_zero_halo(b[2], c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
    BRTM_HIP(launch_zero_halo(b[3], c->ncell, 1, c->n1, c->n2, c->n3, c->stream));
    BRTM_TRY(step_pair(c, b[0], b[1], b[2], b[3], 0, 1, false));
    BRTM_HIP(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nrep; ++s) {
      BRTM_TRY(step_pair(c, b[3], b[2], b[1], b[0], 0, 1, false));
      std::swap(b[0], b[3]);
      std::swap(b[1], b[2]);
    }
  } else {
  // warm-up launch, then the timed run
  BRTM_TRY(step(c, form, act, fmt, cur, old, old, snap, sc.as<float>(), img, m));
  BRTM_HIP(hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < nrep; ++s) {
    BRTM_TRY(step(c, form, act, fmt, cur, old, old, snap, sc.as<float>(), img, m));
    std::swap(cur, old);
  }
  }
  BRTM_HIP(hipEventRecord(c->ev1, c->stream));
  BRTM_HIP(hipEventSynchronize(c->ev1));
  float ms = 0;
  BRTM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *avg_ms = ms / nrep;
  sc.release();
  return BRTM_OK;
}

int brtm_set_tuning(brtm_ctx *c, int ty_tile, int target_blocks) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(ty_tile == 0 || ty_tile == 8 || ty_tile == 16 || ty_tile == 32, BRTM_EINVAL,
               "ty_tile must be 0 (automatic), 8, 16 or 32");
  BRTM_REQUIRE(target_blocks >= 1, BRTM_EINVAL, "target_blocks must be >= 1");
  c->ty_tile = ty_tile;
  c->target_blocks = target_blocks;
  return BRTM_OK;
}

int brtm_set_stream_loads(brtm_ctx *c, int mode) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(mode >= -1 && mode <= 1, BRTM_EINVAL, "stream loads: -1 (automatic), 0 or 1");
  c->stream_loads = mode;
  return BRTM_OK;
}

int brtm_set_prefetch(brtm_ctx *c, int slices) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(slices >= 0 && slices <= 2, BRTM_EINVAL, "prefetch: 0 (automatic), 1 or 2 slices");
  c->prefetch = slices;
  return BRTM_OK;
}

int brtm_set_receiver_fusion(brtm_ctx *c, int on) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  c->rec_fusion = on != 0;
  return BRTM_OK;
}

int brtm_set_halo_ring(brtm_ctx *c, int on) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(on >= -1 && on <= 1, BRTM_EINVAL, "halo ring: -1 (automatic), 0 or 1");
  c->halo_ring = on;
  return BRTM_OK;
}

int brtm_set_x_per_lane(brtm_ctx *c, int cells) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_REQUIRE(cells == 0 || cells == 1 || cells == 2 || cells == 4, BRTM_EINVAL,
               "cells per lane must be 0 (automatic), 1, 2 or 4");
  c->x_per_lane = cells;
  return BRTM_OK;
}

int brtm_copy_bench(brtm_ctx *c, double bytes, int nrep, double *avg_ms) {
  BRTM_REQUIRE(c && avg_ms && nrep >= 1 && bytes >= 16, BRTM_EINVAL, "bad arguments");
  BRTM_HIP(hipSetDevice(c->device));
  const long n16 = (long)(bytes / 16);
  DevBuf a, b;
  BRTM_TRY(a.alloc(16 * (size_t)n16));
  BRTM_TRY(b.alloc(16 * (size_t)n16));
  BRTM_HIP(hipMemsetAsync(a.ptr, 0, 16 * (size_t)n16, c->stream));
  BRTM_HIP(launch_copy16(a.ptr, b.ptr, n16, c->stream));  // warm-up
  BRTM_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < nrep; ++r) BRTM_HIP(launch_copy16(a.ptr, b.ptr, n16, c->stream));
  BRTM_HIP(hipEventRecord(c->ev1, c->stream));
  BRTM_HIP(hipEventSynchronize(c->ev1));
  float ms = 0;
  BRTM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *avg_ms = ms / nrep;
  return BRTM_OK;
}

int brtm_stream_bench(brtm_ctx *c, int nin, double bytes, int nt_store, int nrep,
                      double *avg_ms) {
  BRTM_REQUIRE(c && avg_ms && nrep >= 1 && bytes >= 16 && nin >= 1 && nin <= 3, BRTM_EINVAL,
               "bad arguments");
  BRTM_HIP(hipSetDevice(c->device));
  const long n16 = (long)(bytes / 16);
  DevBuf in[3], out;
  for (int k = 0; k < nin; ++k) {
    BRTM_TRY(in[k].alloc(16 * (size_t)n16));
    BRTM_HIP(hipMemsetAsync(in[k].ptr, 0, 16 * (size_t)n16, c->stream));
  }
  BRTM_TRY(out.alloc(16 * (size_t)n16));
  auto run = [&]() -> hipError_t {
    return launch_sum16(nin, nt_store != 0, in[0].ptr, in[nin > 1 ? 1 : 0].ptr,
                        in[nin > 2 ? 2 : 0].ptr, out.ptr, n16, c->stream);
  };
  BRTM_HIP(run());  // warm-up
  BRTM_HIP(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < nrep; ++r) BRTM_HIP(run());
  BRTM_HIP(hipEventRecord(c->ev1, c->stream));
  BRTM_HIP(hipEventSynchronize(c->ev1));
  float ms = 0;
  BRTM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *avg_ms = ms / nrep;
  return BRTM_OK;
}

int brtm_set_born_path(brtm_ctx *c, int direct_lv) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  c->born_direct_lv = direct_lv != 0;
  return BRTM_OK;
}

int brtm_kernel_timing(brtm_ctx *c, int enable, int form, int act, int fmt) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_HIP(hipStreamSynchronize(c->stream));
  c->ktime_on = enable != 0;
  c->ktime_form = form;
  c->ktime_act = act;
  c->ktime_fmt = fmt;
  c->ktime_used = 0;
  return BRTM_OK;
}

int brtm_kernel_timing_result(brtm_ctx *c, double *total_ms, int *count) {
  BRTM_REQUIRE(c && total_ms && count, BRTM_EINVAL, "NULL argument");
  BRTM_HIP(hipStreamSynchronize(c->stream));
  double t = 0;
  for (size_t k = 0; k + 1 < c->ktime_used; k += 2) {
    float ms = 0;
    BRTM_HIP(hipEventElapsedTime(&ms, c->ktime_ev[k], c->ktime_ev[k + 1]));
    t += ms;
  }
  *total_ms = t;
  *count = (int)(c->ktime_used / 2);
  return BRTM_OK;
}

int brtm_dot_noalloc(brtm_ctx *c, const float *a, const float *b, long n, double *out) {
  BRTM_REQUIRE(c && out && (n == 0 || (a && b)), BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(c);
  BRTM_REQUIRE(n >= 0, BRTM_EINVAL, "n must be >= 0");
  BRTM_REQUIRE(c->noalloc_scratch, BRTM_ESTATE, "the context has no scratch address");
  static_assert(kDotBlocks + 1 <= kNoallocDoubles, "the dot's partial sums exceed the scratch");
  BRTM_HIP(hipSetDevice(c->device));
  // the same two kernels as brtm_vec_dot, on the static scratch: every
  // partial sum and the result are written before they are read, so
  // nothing needs clearing between calls
  double *partials = c->noalloc_scratch, *result = c->noalloc_scratch + kDotBlocks;
  BRTM_HIP(launch_dot(a, b, n, partials, result, c->stream));
  BRTM_HIP(hipMemcpyAsync(out, result, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BRTM_HIP(hipStreamSynchronize(c->stream));
  return BRTM_OK;
}

int brtm_sync(brtm_ctx *c) {
  BRTM_REQUIRE(c, BRTM_EINVAL, "ctx is NULL");
  BRTM_HIP(hipStreamSynchronize(c->stream));
  free_graphs(c);
  return BRTM_OK;
}

int brtm_set_two_step(brtm_ctx *c, int mode) {
  BRTM_REQUIRE(c && mode >= -1 && mode <= 1, BRTM_EINVAL, "mode must be -1, 0 or 1");
  c->two_step = mode;
  return BRTM_OK;
}

int brtm_set_graphs(brtm_ctx *c, int mode) {
  BRTM_REQUIRE(c && mode >= -1 && mode <= 1, BRTM_EINVAL, "mode must be -1, 0 or 1");
  c->graphs = mode;
  return BRTM_OK;
}

int brtm_device_bytes(brtm_ctx *c, double *bytes) {
  BRTM_REQUIRE(c && bytes, BRTM_EINVAL, "NULL argument");
  double b = 0;
  for (DevBuf *x : {&c->d_wtab, &c->V, &c->fa, &c->fb, &c->fc, &c->fd, &c->img, &c->mdl,
                    &c->seg, &c->partials, &c->dotv, &c->slocs, &c->strc, &c->rlocs, &c->rmap, &c->smap,
                    &c->traces, &c->mvw, &c->rscale})
    b += (double)x->bytes;
  for (auto &s : c->stores) b += (double)s.data.bytes + (double)s.scale.bytes;
  *bytes = b;
  return BRTM_OK;
}

}  // extern "C"
