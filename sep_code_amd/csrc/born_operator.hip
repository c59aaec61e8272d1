// Operator and solver layer: HBM-resident vectors, the multi-shot Born
// operator (rtm_zero_op) and the cgstep / simple_solver CG driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/born_op.h"
#include "../../include/born_rtm.h"
#include "born_ctx.h"

struct brtm_op {
  brtm::CtxHold hold;
  brtm_ctx *ctx = nullptr;
  brtm_op_desc d{};
  std::vector<int> src_locs, rec_locs;
  std::vector<float> src_traces, mute;
  std::vector<float> sinc;   // [jt][8] interpolation table for injection and extraction
  brtm::DevBuf d_mute;       // [m3][m2][m1]
  brtm::DevBuf work;         // compute-grid scratch
  int nt_fd = 0, ntr_d = 0;
  bool prepared = false;
  long nrec() const { return (long)d.nrx * d.nry; }
  long model_size() const { return (long)d.m1 * d.m2 * d.m3; }
  long data_size() const { return (long)d.nshots * nrec() * d.nt_data; }
};

namespace brtm {

int check_vec(const brtm_vec *v, long n, const char *what) {
  BRTM_REQUIRE(v && v->p(), BRTM_EINVAL, std::string(what) + " is NULL");
  BRTM_LIVE(v->ctx);
  BRTM_REQUIRE(v->n == n, BRTM_EINVAL,
               std::string(what) + " has " + std::to_string(v->n) + " elements, expected " +
                   std::to_string(n));
  return BRTM_OK;
}

brtm_ctx *op_context(const brtm_op *op) { return op ? op->ctx : nullptr; }

namespace {

__global__ void k_random(float *v, long n, unsigned long long seed) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    unsigned long long z = seed + (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v[i] = (float)((double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  }
}

__global__ void k_add(float *y, const float *__restrict__ x, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] += x[i];
}

int grid_of(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

// Levinson solve of the symmetric Toeplitz system (first column r) for g,
// the recursion of Born/sinc_bob.cpp:82-114, in double.
void levinson(int m, const double *r, const double *g, double *f) {
  std::vector<double> a(m, 0.0);
  a[0] = 1.0;
  double v = r[0];
  f[0] = g[0] / r[0];
  for (int j = 1; j < m; ++j) {
    double e = 0;
    a[j] = 0;
    f[j] = 0;
    for (int i = 0; i < j; ++i) e += a[i] * r[j - i];
    double c = e / v;
    v -= e * c;
    for (int i = 0; i <= j / 2; ++i) {
      const double bot = a[j - i] - c * a[i];
      a[i] -= c * a[j - i];
      a[j - i] = bot;
    }
    double w = 0;
    for (int i = 0; i < j; ++i) w += f[i] * r[j - i];
    c = (g[j] - w) / v;
    for (int i = 0; i <= j; ++i) f[i] += c * a[j - i];
  }
}


#define OP_TRY(x)                     \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != BRTM_OK) return rc_;   \
  } while (0)

// Shot geometry into the context (source + receivers, zero traces).
// Whether shot s's stored source wavefield on the context is the one this
// operator's brtm_op_prepare made.
bool owns_store(const brtm_op *op, int s) {
  const brtm_ctx *c = op->ctx;
  return s < (int)c->stores.size() && c->stores[s].owner == op;
}

int check_store(const brtm_op *op, int s) {
  if (owns_store(op, s)) return BRTM_OK;
  set_error("the stored source wavefield of shot " + std::to_string(s) +
            " was made by another operator on this context (or released): call "
            "brtm_op_prepare again");
  return BRTM_ESTATE;
}

int load_shot(brtm_op *op, int s) {
  brtm_ctx *c = op->ctx;
  const brtm_op_desc &d = op->d;
  OP_TRY(brtm_set_source(c, d.nsrc, d.ntr_s, op->src_locs.data() + (size_t)s * d.nsrc,
                         op->src_traces.data() + (size_t)s * d.nsrc * d.ntr_s));
  OP_TRY(brtm_set_receivers(c, d.nrx, d.nry, op->ntr_d, op->rec_locs.data() + (size_t)s * op->nrec(),
                            nullptr));
  // The operator's own jt-phase tables for the source and the data
  // (Born/fd_prop_3d.cpp:28-32, Born/rtm_zero_op_3d.cpp:32-36): installed
  // for every shot because the context may be shared with other work.
  OP_TRY(brtm_set_sinc_table(c, 0, 8, d.jt, op->sinc.data()));
  OP_TRY(brtm_set_sinc_table(c, 1, 8, d.jt, op->sinc.data()));
  OP_TRY(brtm_select_shot(c, s));
  return BRTM_OK;
}

}  // namespace
}  // namespace brtm

using namespace brtm;

extern "C" {

// ---------------------------------------------------------------- vectors
int brtm_vec_create(brtm_ctx *ctx, long n, brtm_vec **out) {
  BRTM_REQUIRE(ctx && out && n > 0, BRTM_EINVAL, "bad vector arguments");
  BRTM_LIVE(ctx);
  BRTM_HIP(hipSetDevice(ctx->device));
  brtm_vec *v = new brtm_vec();
  v->hold.reset(ctx);
  v->ctx = ctx;
  v->n = n;
  int rc = v->buf.alloc(sizeof(float) * (size_t)n);
  if (rc != BRTM_OK) {
    delete v;
    return rc;
  }
  v->ptr = v->buf.as<float>();
  hipError_t e = hipMemsetAsync(v->p(), 0, sizeof(float) * (size_t)n, ctx->stream);
  if (e != hipSuccess) {
    delete v;
    set_error(hipGetErrorString(e));
    return BRTM_EHIP;
  }
  *out = v;
  return BRTM_OK;
}

int brtm_vec_destroy(brtm_vec *v) {
  if (!v) return BRTM_OK;
  // the context's host record outlives brtm_destroy while v holds it
  (void)hipSetDevice(v->ctx->device);
  (void)hipStreamSynchronize(v->ctx->stream);
  delete v;
  return BRTM_OK;
}

long brtm_vec_size(const brtm_vec *v) { return v ? v->n : 0; }
void *brtm_vec_data(brtm_vec *v) { return v ? v->p() : nullptr; }

int brtm_vec_upload(brtm_vec *v, const float *host) {
  BRTM_REQUIRE(v && host, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(v->ctx);
  BRTM_HIP(hipMemcpyAsync(v->p(), host, sizeof(float) * (size_t)v->n, hipMemcpyHostToDevice,
                          v->ctx->stream));
  BRTM_HIP(hipStreamSynchronize(v->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_download(const brtm_vec *v, float *host) {
  BRTM_REQUIRE(v && host, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(v->ctx);
  BRTM_HIP(hipMemcpyAsync(host, v->p(), sizeof(float) * (size_t)v->n, hipMemcpyDeviceToHost,
                          v->ctx->stream));
  BRTM_HIP(hipStreamSynchronize(v->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_zero(brtm_vec *v) {
  BRTM_REQUIRE(v, BRTM_EINVAL, "NULL vector");
  BRTM_LIVE(v->ctx);
  BRTM_HIP(hipMemsetAsync(v->p(), 0, sizeof(float) * (size_t)v->n, v->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_copy(brtm_vec *dst, const brtm_vec *src) {
  BRTM_REQUIRE(dst && src && dst->n == src->n, BRTM_EINVAL, "vector size mismatch");
  BRTM_LIVE(dst->ctx);
  BRTM_HIP(hipMemcpyAsync(dst->p(), src->p(), sizeof(float) * (size_t)src->n,
                          hipMemcpyDeviceToDevice, dst->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_scale(brtm_vec *v, double s) {
  BRTM_REQUIRE(v, BRTM_EINVAL, "NULL vector");
  BRTM_LIVE(v->ctx);
  BRTM_HIP(launch_axpby(v->p(), v->p(), s, 0.0, v->n, v->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_scale_add(brtm_vec *v, double a, const brtm_vec *x, double b) {
  BRTM_REQUIRE(v && x && v->n == x->n, BRTM_EINVAL, "vector size mismatch");
  BRTM_LIVE(v->ctx);
  BRTM_HIP(launch_axpby(v->p(), x->p(), a, b, v->n, v->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_add(brtm_vec *v, const brtm_vec *x) {
  BRTM_REQUIRE(v && x && v->n == x->n, BRTM_EINVAL, "vector size mismatch");
  BRTM_LIVE(v->ctx);
  k_add<<<grid_of(v->n), 256, 0, v->ctx->stream>>>(v->p(), x->p(), v->n);
  BRTM_HIP(hipGetLastError());
  return BRTM_OK;
}

int brtm_vec_mult(brtm_vec *v, const brtm_vec *x) {
  BRTM_REQUIRE(v && x && v->n == x->n, BRTM_EINVAL, "vector size mismatch");
  BRTM_LIVE(v->ctx);
  BRTM_HIP(launch_mul(v->p(), x->p(), v->n, v->ctx->stream));
  return BRTM_OK;
}

int brtm_vec_dot(const brtm_vec *a, const brtm_vec *b, double *out) {
  BRTM_REQUIRE(a && b && out && a->n == b->n, BRTM_EINVAL, "vector size mismatch");
  BRTM_LIVE(a->ctx);
  return brtm::dot(a->ctx, a->p(), b->p(), a->n, out);
}

int brtm_vec_random(brtm_vec *v, unsigned long long seed) {
  BRTM_REQUIRE(v, BRTM_EINVAL, "NULL vector");
  BRTM_LIVE(v->ctx);
  k_random<<<grid_of(v->n), 256, 0, v->ctx->stream>>>(v->p(), v->n, seed);
  BRTM_HIP(hipGetLastError());
  return BRTM_OK;
}

// --------------------------------------------------------------- operator
int brtm_op_create(brtm_ctx *ctx, const brtm_op_desc *desc, brtm_op **out) {
  BRTM_REQUIRE(ctx && desc && out, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(ctx);
  BRTM_REQUIRE(ctx->have_space && ctx->have_vel, BRTM_ESTATE,
               "the context needs brtm_create_space and brtm_set_velocity first");
  const brtm_op_desc &d = *desc;
  BRTM_REQUIRE(d.m1 > 0 && d.m2 > 0 && d.m3 > 0 && d.nshots > 0 && d.nsrc >= 0 &&
                   d.nrx > 0 && d.nry > 0 && d.nt_data > 1 && d.jt >= 1 && d.jimg >= 1 &&
                   d.dt_data > 0 && d.ntr_s >= 1,
               BRTM_EINVAL, "bad operator description");
  BRTM_REQUIRE(d.o1 >= 0 && d.o2 >= 0 && d.o3 >= 0 && d.o1 + d.m1 <= ctx->n1 &&
                   d.o2 + d.m2 <= ctx->n2 && d.o3 + d.m3 <= ctx->n3,
               BRTM_EINVAL, "model grid does not fit in the compute grid at its origin");
  BRTM_REQUIRE(d.store >= BRTM_STORE_F32 && d.store <= BRTM_STORE_CKPT, BRTM_EINVAL,
               "the operator needs a source-wavefield store (F32, BF16, F16 or CKPT)");
  BRTM_REQUIRE(d.store != BRTM_STORE_CKPT || d.ckpt >= 1, BRTM_EINVAL, "ckpt must be >= 1");
  BRTM_REQUIRE(d.nsrc == 0 || (d.src_locs && d.src_traces), BRTM_EINVAL, "missing sources");
  BRTM_REQUIRE(d.rec_locs, BRTM_EINVAL, "missing receivers");
  brtm_op *op = new brtm_op();
  op->hold.reset(ctx);
  op->ctx = ctx;
  op->d = d;
  op->nt_fd = (d.nt_data - 1) * d.jt + 1;  // Born/rtm_zero_op_3d.cpp:21
  op->ntr_d = d.nt_data + 8;               // Born/rtm_zero_op_3d.cpp:99
  op->src_locs.assign(d.src_locs, d.src_locs + (size_t)d.nshots * d.nsrc);
  op->src_traces.assign(d.src_traces, d.src_traces + (size_t)d.nshots * d.nsrc * d.ntr_s);
  op->rec_locs.assign(d.rec_locs, d.rec_locs + (size_t)d.nshots * op->nrec());
  op->d.src_locs = nullptr;
  op->d.src_traces = nullptr;
  op->d.rec_locs = nullptr;
  op->d.mute = nullptr;
  op->sinc.assign((size_t)8 * d.jt, 0.f);
  if (brtm_sinc_table(d.jt, 8, op->sinc.data()) != BRTM_OK) {
    delete op;
    return BRTM_EINVAL;
  }
  if (d.mute) {
    const size_t mb = sizeof(float) * (size_t)op->model_size();
    int rc = op->d_mute.alloc(mb);
    if (rc == BRTM_OK && hipMemcpy(op->d_mute.ptr, d.mute, mb, hipMemcpyHostToDevice) != hipSuccess)
      rc = BRTM_EHIP;
    if (rc != BRTM_OK) {
      delete op;
      set_error("mute upload failed");
      return rc;
    }
  }
  if (op->work.alloc(sizeof(float) * (size_t)ctx->ncell) != BRTM_OK) {
    delete op;
    return BRTM_ENOMEM;
  }
  *out = op;
  return BRTM_OK;
}

int brtm_op_destroy(brtm_op *op) {
  if (!op) return BRTM_OK;
  (void)hipSetDevice(op->ctx->device);
  (void)hipStreamSynchronize(op->ctx->stream);
  // a destroyed context has already freed the shots' stores; stores that
  // another operator on the context has prepared since are not this one's
  if (!op->ctx->destroyed)
    for (int s = 0; s < op->d.nshots; ++s)
      if (owns_store(op, s)) brtm_release_store(op->ctx, s);
  delete op;
  return BRTM_OK;
}

long brtm_op_model_size(const brtm_op *op) { return op ? op->model_size() : 0; }
long brtm_op_data_size(const brtm_op *op) { return op ? op->data_size() : 0; }

int brtm_op_prepare(brtm_op *op) {
  BRTM_REQUIRE(op, BRTM_EINVAL, "NULL operator");
  BRTM_LIVE(op->ctx);
  BRTM_HIP(hipSetDevice(op->ctx->device));
  for (int s = 0; s < op->d.nshots; ++s) {
    OP_TRY(load_shot(op, s));
    OP_TRY(brtm_source_prop(op->ctx, op->d.jt, op->nt_fd, op->d.jimg, op->d.store, op->d.ckpt,
                            nullptr, nullptr));
    op->ctx->stores[s].owner = op;
  }
  op->prepared = true;
  return BRTM_OK;
}

int brtm_op_forward(brtm_op *op, int add, const brtm_vec *model, brtm_vec *data) {
  BRTM_REQUIRE(op, BRTM_EINVAL, "NULL operator");
  BRTM_LIVE(op->ctx);
  BRTM_REQUIRE(op->prepared, BRTM_ESTATE, "brtm_op_prepare has not been called");
  OP_TRY(check_vec(model, op->model_size(), "model"));
  OP_TRY(check_vec(data, op->data_size(), "data"));
  brtm_ctx *c = op->ctx;
  const brtm_op_desc &d = op->d;
  BRTM_HIP(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  if (!add) BRTM_HIP(hipMemsetAsync(data->p(), 0, sizeof(float) * (size_t)data->n, st));
  // m~ = D_z E (mute . m): the same for every shot.
  float *w = op->work.as<float>();
  const float *msrc = model->p();
  if (op->d_mute.ptr) {
    BRTM_HIP(hipMemcpyAsync(w, model->p(), sizeof(float) * (size_t)model->n,
                            hipMemcpyDeviceToDevice, st));
    BRTM_HIP(launch_mul(w, op->d_mute.as<float>(), model->n, st));
    msrc = w;
  }
  BRTM_HIP(launch_embed(0, msrc, c->img.as<float>(), d.m1, d.m2, d.m3, c->n1, c->n2, c->n3, d.o1,
                        d.o2, d.o3, st));
  BRTM_HIP(launch_dz(0, c->img.as<float>(), c->mdl.as<float>(), c->n1 * c->n2, c->n3, 0, st));
  const float inv_dt = 1.f / d.dt_data;
  const long nrec = op->nrec();
  for (int s = 0; s < d.nshots; ++s) {
    OP_TRY(check_store(op, s));
    OP_TRY(load_shot(op, s));
    OP_TRY(born_forward(c, d.jt, op->nt_fd, d.jimg, false));
    BRTM_HIP(launch_trace_to_data(c->traces.as<float>(),
                                  data->p() + (size_t)s * nrec * d.nt_data, (int)nrec, op->ntr_d,
                                  d.nt_data, inv_dt, 1, st));
  }
  return BRTM_OK;
}

int brtm_op_adjoint(brtm_op *op, int add, brtm_vec *model, const brtm_vec *data) {
  BRTM_REQUIRE(op, BRTM_EINVAL, "NULL operator");
  BRTM_LIVE(op->ctx);
  BRTM_REQUIRE(op->prepared, BRTM_ESTATE, "brtm_op_prepare has not been called");
  OP_TRY(check_vec(model, op->model_size(), "model"));
  OP_TRY(check_vec(data, op->data_size(), "data"));
  brtm_ctx *c = op->ctx;
  const brtm_op_desc &d = op->d;
  BRTM_HIP(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  if (!add) BRTM_HIP(hipMemsetAsync(model->p(), 0, sizeof(float) * (size_t)model->n, st));
  BRTM_HIP(hipMemsetAsync(c->img.ptr, 0, sizeof(float) * (size_t)c->ncell, st));
  const float inv_dt = 1.f / d.dt_data;
  const long nrec = op->nrec();
  for (int s = 0; s < d.nshots; ++s) {
    OP_TRY(check_store(op, s));
    OP_TRY(load_shot(op, s));
    BRTM_HIP(launch_data_to_trace(data->p() + (size_t)s * nrec * d.nt_data, c->traces.as<float>(),
                                  (int)nrec, op->ntr_d, d.nt_data, inv_dt, st));
    OP_TRY(rtm_adjoint(c, d.jt, op->nt_fd, d.jimg));
  }
  // model += mute . E^T D_z^T (sum of shot images)
  float *w = op->work.as<float>();
  BRTM_HIP(launch_dz(1, c->img.as<float>(), w, c->n1 * c->n2, c->n3, 0, st));
  if (op->d_mute.ptr) {
    // accumulate into a zeroed model-size staging area inside c->mdl
    float *stage = c->mdl.as<float>();
    BRTM_HIP(hipMemsetAsync(stage, 0, sizeof(float) * (size_t)model->n, st));
    BRTM_HIP(launch_embed(1, stage, w, d.m1, d.m2, d.m3, c->n1, c->n2, c->n3, d.o1, d.o2, d.o3,
                          st));
    BRTM_HIP(launch_mul(stage, op->d_mute.as<float>(), model->n, st));
    k_add<<<grid_of(model->n), 256, 0, st>>>(model->p(), stage, model->n);
    BRTM_HIP(hipGetLastError());
  } else {
    BRTM_HIP(launch_embed(1, model->p(), w, d.m1, d.m2, d.m3, c->n1, c->n2, c->n3, d.o1, d.o2,
                          d.o3, st));
  }
  return BRTM_OK;
}

int brtm_op_illumination(brtm_op *op, brtm_vec *out) {
  BRTM_REQUIRE(op, BRTM_EINVAL, "NULL operator");
  BRTM_LIVE(op->ctx);
  BRTM_REQUIRE(op->prepared, BRTM_ESTATE, "brtm_op_prepare has not been called");
  OP_TRY(check_vec(out, op->model_size(), "out"));
  brtm_ctx *c = op->ctx;
  const brtm_op_desc &d = op->d;
  BRTM_HIP(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  BRTM_HIP(hipMemsetAsync(out->p(), 0, sizeof(float) * (size_t)out->n, st));
  float *acc = op->work.as<float>();
  BRTM_HIP(hipMemsetAsync(acc, 0, sizeof(float) * (size_t)c->ncell, st));
  for (int s = 0; s < d.nshots; ++s) {
    OP_TRY(check_store(op, s));
    OP_TRY(brtm_select_shot(c, s));
    OP_TRY(illumination(c, acc));
  }
  BRTM_HIP(launch_embed(1, out->p(), acc, d.m1, d.m2, d.m3, c->n1, c->n2, c->n3, d.o1, d.o2,
                        d.o3, st));
  if (op->d_mute.ptr) BRTM_HIP(launch_mul(out->p(), op->d_mute.as<float>(), out->n, st));
  return BRTM_OK;
}

int brtm_op_dot_test(brtm_op *op, unsigned long long seed, double err[2]) {
  BRTM_REQUIRE(op && err, BRTM_EINVAL, "NULL argument");
  brtm_ctx *c = op->ctx;
  brtm_vec *m1 = nullptr, *m2 = nullptr, *d1 = nullptr, *d2 = nullptr;
  int rc = brtm_vec_create(c, op->model_size(), &m1);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, op->model_size(), &m2);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, op->data_size(), &d1);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, op->data_size(), &d2);
  double dot1 = 0, dot2 = 0, dot3 = 0, dot4 = 0;
  // Born/my_operator.cpp:14-76: forward(false), adjoint(false), then both
  // again with add = true on top of the first results.
  if (rc == BRTM_OK) rc = brtm_vec_random(m1, seed);
  if (rc == BRTM_OK) rc = brtm_vec_random(d2, seed ^ 0x5DEECE66Dull);
  if (rc == BRTM_OK) rc = brtm_op_forward(op, 0, m1, d1);
  if (rc == BRTM_OK) rc = brtm_vec_dot(d1, d2, &dot1);
  if (rc == BRTM_OK) rc = brtm_op_adjoint(op, 0, m2, d2);
  if (rc == BRTM_OK) rc = brtm_vec_dot(m2, m1, &dot2);
  if (rc == BRTM_OK) rc = brtm_op_forward(op, 1, m1, d1);
  if (rc == BRTM_OK) rc = brtm_op_adjoint(op, 1, m2, d2);
  if (rc == BRTM_OK) rc = brtm_vec_dot(m1, m2, &dot3);
  if (rc == BRTM_OK) rc = brtm_vec_dot(d1, d2, &dot4);
  brtm_vec_destroy(m1);
  brtm_vec_destroy(m2);
  brtm_vec_destroy(d1);
  brtm_vec_destroy(d2);
  if (rc != BRTM_OK) return rc;
  BRTM_REQUIRE(std::fabs(dot1) > 1e-30 && std::fabs(dot3) > 1e-30, BRTM_ESTATE,
               "dot product suspiciously small");
  err[0] = std::fabs((dot1 - dot2) / dot1);
  err[1] = std::fabs((dot3 - dot4) / dot3);
  return BRTM_OK;
}

int brtm_sinc_table(int ntab, int nsinc, float *table) {
  BRTM_REQUIRE(table && ntab >= 1 && nsinc >= 2 && nsinc % 2 == 0 && nsinc <= 32, BRTM_EINVAL,
               "bad sinc table arguments");
  // Born/sinc_bob.cpp:43-80 (mksinc): least-squares spectral matching of a
  // tapered sinc, Toeplitz system from cosine sums over 2*nsinc samples of
  // the band [0, snat], solved by Levinson.
  const double pi2 = 2.0 * 3.141592654;
  const double snat = 0.5 * (0.066 + 0.265 * std::log((double)nsinc));
  const double ds = snat / (2.0 * nsinc - 1.0);
  std::vector<double> b(nsinc), cc(nsinc), f(nsinc);
  for (int it = 0; it < ntab; ++it) {
    const double dfrac = (double)it / ntab;
    const double eta = nsinc / 2 - 1.0 + dfrac;
    for (int j = 0; j < nsinc; ++j) {
      b[j] = cc[j] = 0;
      for (int k = 0; k < 2 * nsinc; ++k) {
        const double sv = k * ds;
        b[j] += std::cos(pi2 * sv * j);
        cc[j] += std::cos(pi2 * sv * (eta - j));
      }
    }
    levinson(nsinc, b.data(), cc.data(), f.data());
    for (int k = 0; k < nsinc; ++k) table[it * nsinc + k] = (float)f[k];
    if (it == 0) {
      // zero shift: the interpolator is the identity (tap nsinc/2-1)
      for (int k = 0; k < nsinc; ++k) table[k] = k == nsinc / 2 - 1 ? 1.f : 0.f;
    }
  }
  return BRTM_OK;
}

}  // extern "C"
