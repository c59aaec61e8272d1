// Learned image-domain preconditioner: a 3-level 3-D U-Net trained and run
// with the MFMA convolutions of conv3d.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/born_unet.h"
#include "born_ctx.h"
#include "conv3d.h"

using namespace brtm;

namespace {

struct LayerDef {
  int cin, cout;
};
// e1b, e2a, e2b, b1, b2, d2a, d2b, d1a, d1b (the first layer 1->16 and the
// 1x1x1 head are separate).
const LayerDef kConv[9] = {{16, 16}, {16, 32}, {32, 32}, {32, 64}, {64, 64},
                           {96, 32}, {32, 32}, {48, 16}, {16, 16}};
constexpr int C1 = 16, C2 = 32, C3 = 64;

int pad32(int c) { return (c + 31) / 32 * 32; }

__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ t, float *g, long n,
                      double *partials) {
  double s = 0;
  const float sc = 2.f / (float)n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const float d = y[i] - t[i];
    g[i] = sc * d;
    s += (double)d * d;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void k_sum(const double *p, int n, double *out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_extract(const float *__restrict__ vol, int m1, int m2, int m3, int o1, int o2,
                          int o3, float scale, float *dst, int P) {
  const long n = (long)P * P * P;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int a = (int)(i % P), b = (int)((i / P) % P), c = (int)(i / ((long)P * P));
    const int x = o1 + a, y = o2 + b, z = o3 + c;
    float v = 0.f;
    if (x >= 0 && x < m1 && y >= 0 && y < m2 && z >= 0 && z < m3)
      v = vol[((long)z * m2 + y) * m1 + x];
    dst[i] = scale * v;
  }
}

// Central (P - 2 ov)^3 of a patch into the volume, scaled.
__global__ void k_insert(const float *__restrict__ src, int P, int ov, float *vol, int m1, int m2,
                         int m3, int o1, int o2, int o3, float scale) {
  const int S = P - 2 * ov;
  const long n = (long)S * S * S;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int a = (int)(i % S) + ov, b = (int)((i / S) % S) + ov, c = (int)(i / ((long)S * S)) + ov;
    const int x = o1 + a, y = o2 + b, z = o3 + c;
    if (x >= 0 && x < m1 && y >= 0 && y < m2 && z >= 0 && z < m3)
      vol[((long)z * m2 + y) * m1 + x] = scale * src[((long)c * P + b) * P + a];
  }
}

// As k_insert, of the odd part (net(x) - net(-x)) / 2 of two outputs.
__global__ void k_insert_odd(const float *__restrict__ sp, const float *__restrict__ sm, int P,
                             int ov, float *vol, int m1, int m2, int m3, int o1, int o2, int o3,
                             float scale) {
  const int S = P - 2 * ov;
  const long n = (long)S * S * S;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int a = (int)(i % S) + ov, b = (int)((i / S) % S) + ov, c = (int)(i / ((long)S * S)) + ov;
    const int x = o1 + a, y = o2 + b, z = o3 + c;
    if (x >= 0 && x < m1 && y >= 0 && y < m2 && z >= 0 && z < m3) {
      const long j = ((long)c * P + b) * P + a;
      vol[((long)z * m2 + y) * m1 + x] = 0.5f * scale * (sp[j] - sm[j]);
    }
  }
}

struct Rng {
  unsigned long long s;
  unsigned long long next() {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); }
};

}  // namespace

struct brtm_unet {
  CtxHold hold;
  brtm_ctx *ctx = nullptr;
  int B = 0, P = 0;
  long P3 = 0;
  long nparams = 0;
  long w0 = 0, b0 = 0, wl[9] = {}, bl[9] = {}, wh = 0, bh = 0;
  DevBuf params, grads, am, av;
  DevBuf wbf[9], wtf[9], gpad[9];
  DevBuf e1a, cat1, p1, e2a, cat2, p2, b1, b2, d2a, d2b, d1a, d1b, out;
  DevBuf g_out, dy_d1b, dy_d1a, dy_cat1, dy_d2b, dy_d2a, dy_cat2, dy_b2, dy_b1, dy_p2, dy_e2a,
      dy_p1, dy_e1a, dpre;
  DevBuf partials, lossv, xpatch, ypatch;
  int step = 0;
  int gm1 = 0, gm2 = 0, gm3 = 0, gov = 8;
  const brtm_vec *split_w = nullptr;  // split weight W of the preconditioner (not owned)
  brtm_vec *split_tmp = nullptr;      // W . g
  brtm_vec *znet_prev = nullptr;      // the network's previous output (guard)
  brtm_vec *split_d = nullptr;        // W^2 g (guard mode 1)
  bool have_prev = false;
  double guard_repeat = 0.8, guard_min = 0.05;
  int guard_mode = 1;                 // 1: W^2 g + new part of the correction; 0: repeat -> W^2 g
  int guard_uses = 0;                 // iterations that took W^2 g since the weight was set
  bool odd = true;                    // preconditioner uses the odd part of the network
  float *prm() const { return params.as<float>(); }
};

namespace {

#define U_TRY(x)                    \
  do {                              \
    int rc_ = (x);                  \
    if (rc_ != BRTM_OK) return rc_; \
  } while (0)

Act act(const DevBuf &b, int c, int cs, int c0 = 0) {
  return Act{static_cast<uint16_t *>(b.ptr) + c0, c, cs};
}

int repack(brtm_unet *u) {
  hipStream_t st = u->ctx->stream;
  for (int l = 0; l < 9; ++l) {
    BRTM_HIP(pack_weights(u->prm() + u->wl[l], u->wbf[l].as<uint16_t>(), kConv[l].cin,
                          kConv[l].cout, st));
    BRTM_HIP(conv3d_flip_weights(u->wbf[l].as<uint16_t>(), u->wtf[l].as<uint16_t>(),
                                 kConv[l].cin, kConv[l].cout, st));
  }
  return BRTM_OK;
}

int forward(brtm_unet *u, const float *x, int nb) {
  hipStream_t st = u->ctx->stream;
  const int P = u->P;
  const Vol v1{nb, P, P, P}, v2{nb, P / 2, P / 2, P / 2}, v3{nb, P / 4, P / 4, P / 4};
  const float *pr = u->prm();
  auto W = [&](int l) { return u->wbf[l].as<uint16_t>(); };
  auto Bv = [&](int l) { return pr + u->bl[l]; };
  BRTM_HIP(conv3d_first_fwd(x, pr + u->w0, pr + u->b0, act(u->e1a, C1, C1), v1, st));
  BRTM_HIP(conv3d_fwd(act(u->e1a, C1, C1), W(0), Bv(0), act(u->cat1, C1, 48, 32), v1, true,
                      false, st));
  BRTM_HIP(avgpool_fwd(act(u->cat1, C1, 48, 32), act(u->p1, C1, C1), v1, st));
  BRTM_HIP(conv3d_fwd(act(u->p1, C1, C1), W(1), Bv(1), act(u->e2a, C2, C2), v2, true, false, st));
  BRTM_HIP(conv3d_fwd(act(u->e2a, C2, C2), W(2), Bv(2), act(u->cat2, C2, 96, 64), v2, true, false,
                      st));
  BRTM_HIP(avgpool_fwd(act(u->cat2, C2, 96, 64), act(u->p2, C2, C2), v2, st));
  BRTM_HIP(conv3d_fwd(act(u->p2, C2, C2), W(3), Bv(3), act(u->b1, C3, C3), v3, true, false, st));
  BRTM_HIP(conv3d_fwd(act(u->b1, C3, C3), W(4), Bv(4), act(u->b2, C3, C3), v3, true, false, st));
  BRTM_HIP(upsample_fwd(act(u->b2, C3, C3), act(u->cat2, C3, 96, 0), v3, st));
  BRTM_HIP(conv3d_fwd(act(u->cat2, 96, 96), W(5), Bv(5), act(u->d2a, C2, C2), v2, true, false, st));
  BRTM_HIP(conv3d_fwd(act(u->d2a, C2, C2), W(6), Bv(6), act(u->d2b, C2, C2), v2, true, false, st));
  BRTM_HIP(upsample_fwd(act(u->d2b, C2, C2), act(u->cat1, C2, 48, 0), v2, st));
  BRTM_HIP(conv3d_fwd(act(u->cat1, 48, 48), W(7), Bv(7), act(u->d1a, C1, C1), v1, true, false, st));
  BRTM_HIP(conv3d_fwd(act(u->d1a, C1, C1), W(8), Bv(8), act(u->d1b, C1, C1), v1, true, false, st));
  BRTM_HIP(head_fwd(act(u->d1b, C1, C1), pr + u->wh, pr + u->bh, x, u->out.as<float>(), v1, st));
  return BRTM_OK;
}

// Backward of one MFMA layer l: given the FP32 gradient of its (ReLU) output
// gy (stride gcs) and its output activation y, accumulate its parameter
// gradients from input x and (if gx) write the input gradient gx (FP32,
// all cin channels, stride cin).
int layer_bwd(brtm_unet *u, int l, const float *gy, int gcs, const Act &y, const Act &x, float *gx,
              const Vol &v) {
  hipStream_t st = u->ctx->stream;
  const int cin = kConv[l].cin, cout = kConv[l].cout;
  Act dp = act(u->dpre, cout, cout);
  BRTM_HIP(relu_bwd(gy, gcs, y, dp, v, st));
  BRTM_HIP(conv3d_wgrad(x, dp, u->gpad[l].as<float>(), u->grads.as<float>() + u->bl[l], v, st));
  if (gx) {
    Act out{gx, cin, cin};
    BRTM_HIP(conv3d_fwd(dp, u->wtf[l].as<uint16_t>(), nullptr, out, v, false, true, st));
  }
  return BRTM_OK;
}

int backward(brtm_unet *u, const float *x, int nb) {
  hipStream_t st = u->ctx->stream;
  const int P = u->P;
  const Vol v1{nb, P, P, P}, v2{nb, P / 2, P / 2, P / 2}, v3{nb, P / 4, P / 4, P / 4};
  const long n1 = (long)nb * v1.spatial(), n2 = (long)nb * v2.spatial(),
             n3 = (long)nb * v3.spatial();
  float *G = u->grads.as<float>();
  const float *pr = u->prm();
  BRTM_HIP(hipMemsetAsync(G, 0, sizeof(float) * u->nparams, st));
  for (int l = 0; l < 9; ++l)
    BRTM_HIP(hipMemsetAsync(u->gpad[l].ptr, 0, u->gpad[l].bytes, st));
  BRTM_HIP(hipMemsetAsync(u->dy_d2b.ptr, 0, sizeof(float) * n2 * C2, st));
  BRTM_HIP(hipMemsetAsync(u->dy_b2.ptr, 0, sizeof(float) * n3 * C3, st));
  // head
  BRTM_HIP(head_bwd(act(u->d1b, C1, C1), pr + u->wh, u->g_out.as<float>(),
                    u->dy_d1b.as<float>(), C1, G + u->wh, G + u->bh, v1, st));
  // d1b, d1a
  U_TRY(layer_bwd(u, 8, u->dy_d1b.as<float>(), C1, act(u->d1b, C1, C1), act(u->d1a, C1, C1),
                  u->dy_d1a.as<float>(), v1));
  U_TRY(layer_bwd(u, 7, u->dy_d1a.as<float>(), C1, act(u->d1a, C1, C1), act(u->cat1, 48, 48),
                  u->dy_cat1.as<float>(), v1));
  // up1: cat1[0:32] came from d2b
  BRTM_HIP(upsample_bwd(u->dy_cat1.as<float>(), 48, u->dy_d2b.as<float>(), C2, C2, v2, st));
  U_TRY(layer_bwd(u, 6, u->dy_d2b.as<float>(), C2, act(u->d2b, C2, C2), act(u->d2a, C2, C2),
                  u->dy_d2a.as<float>(), v2));
  U_TRY(layer_bwd(u, 5, u->dy_d2a.as<float>(), C2, act(u->d2a, C2, C2), act(u->cat2, 96, 96),
                  u->dy_cat2.as<float>(), v2));
  // up2: cat2[0:64] came from b2
  BRTM_HIP(upsample_bwd(u->dy_cat2.as<float>(), 96, u->dy_b2.as<float>(), C3, C3, v3, st));
  U_TRY(layer_bwd(u, 4, u->dy_b2.as<float>(), C3, act(u->b2, C3, C3), act(u->b1, C3, C3),
                  u->dy_b1.as<float>(), v3));
  U_TRY(layer_bwd(u, 3, u->dy_b1.as<float>(), C3, act(u->b1, C3, C3), act(u->p2, C2, C2),
                  u->dy_p2.as<float>(), v3));
  // e2b (cat2[64:96]) gets the skip gradient plus the pooled path
  BRTM_HIP(avgpool_bwd(u->dy_p2.as<float>(), C2, u->dy_cat2.as<float>() + 64, 96, C2, v2, st));
  U_TRY(layer_bwd(u, 2, u->dy_cat2.as<float>() + 64, 96, act(u->cat2, C2, 96, 64),
                  act(u->e2a, C2, C2), u->dy_e2a.as<float>(), v2));
  U_TRY(layer_bwd(u, 1, u->dy_e2a.as<float>(), C2, act(u->e2a, C2, C2), act(u->p1, C1, C1),
                  u->dy_p1.as<float>(), v2));
  BRTM_HIP(avgpool_bwd(u->dy_p1.as<float>(), C1, u->dy_cat1.as<float>() + 32, 48, C1, v1, st));
  U_TRY(layer_bwd(u, 0, u->dy_cat1.as<float>() + 32, 48, act(u->cat1, C1, 48, 32),
                  act(u->e1a, C1, C1), u->dy_e1a.as<float>(), v1));
  // first layer: ReLU backward then weight gradient (no input gradient)
  Act dp = act(u->dpre, C1, C1);
  BRTM_HIP(relu_bwd(u->dy_e1a.as<float>(), C1, act(u->e1a, C1, C1), dp, v1, st));
  BRTM_HIP(conv3d_first_wgrad(x, dp, G + u->w0, G + u->b0, v1, st));
  for (int l = 0; l < 9; ++l)
    BRTM_HIP(unpack_grad(u->gpad[l].as<float>(), G + u->wl[l], kConv[l].cin, kConv[l].cout, st));
  (void)n1;
  return BRTM_OK;
}

int loss_grad(brtm_unet *u, const float *x, const float *t, int nb, double *loss) {
  U_TRY(forward(u, x, nb));
  const long n = (long)nb * u->P3;
  k_mse<<<256, 256, 0, u->ctx->stream>>>(u->out.as<float>(), t, u->g_out.as<float>(), n,
                                         u->partials.as<double>());
  k_sum<<<1, 256, 0, u->ctx->stream>>>(u->partials.as<double>(), 256, u->lossv.as<double>());
  BRTM_HIP(hipGetLastError());
  U_TRY(backward(u, x, nb));
  double s = 0;
  BRTM_HIP(hipMemcpyAsync(&s, u->lossv.ptr, sizeof(double), hipMemcpyDeviceToHost,
                          u->ctx->stream));
  BRTM_HIP(hipStreamSynchronize(u->ctx->stream));
  if (loss) *loss = s / (double)n;
  return BRTM_OK;
}

}  // namespace

extern "C" {

int brtm_unet_create(brtm_ctx *ctx, int batch, int patch, unsigned long long seed,
                     brtm_unet **out) {
  BRTM_REQUIRE(ctx && out && batch >= 1 && patch >= 64 && patch % 64 == 0, BRTM_EINVAL,
               "bad U-Net arguments (patch must be a multiple of 64)");
  BRTM_LIVE(ctx);
  BRTM_HIP(hipSetDevice(ctx->device));
  brtm_unet *u = new brtm_unet();
  u->hold.reset(ctx);
  u->ctx = ctx;
  u->B = batch;
  u->odd = batch >= 2;
  u->P = patch;
  u->P3 = (long)patch * patch * patch;
  long off = 0;
  u->w0 = off; off += 27 * C1;
  u->b0 = off; off += C1;
  for (int l = 0; l < 9; ++l) {
    u->wl[l] = off; off += 27L * kConv[l].cin * kConv[l].cout;
    u->bl[l] = off; off += kConv[l].cout;
  }
  u->wh = off; off += C1;
  u->bh = off; off += 1;
  u->nparams = off;
  // He-normal convolution weights, zero biases, a near-zero head.
  std::vector<float> h(off, 0.f);
  Rng rng{seed};
  for (int i = 0; i < 27 * C1; ++i) h[u->w0 + i] = (float)(rng.normal() * std::sqrt(2.0 / 27.0));
  for (int l = 0; l < 9; ++l) {
    const double sd = std::sqrt(2.0 / (27.0 * kConv[l].cin));
    for (long i = 0; i < 27L * kConv[l].cin * kConv[l].cout; ++i)
      h[u->wl[l] + i] = (float)(rng.normal() * sd);
  }
  for (int i = 0; i < C1; ++i) h[u->wh + i] = (float)(rng.normal() * 1e-3);
  const long B = batch, P3 = u->P3, P3b = P3 / 8, P3c = P3 / 64;
  const size_t bf = 2, f4 = 4;
  int rc = BRTM_OK;
  auto A = [&](DevBuf &b, size_t bytes) {
    if (rc == BRTM_OK) rc = b.alloc(bytes);
  };
  A(u->params, f4 * off);
  A(u->grads, f4 * off);
  A(u->am, f4 * off);
  A(u->av, f4 * off);
  for (int l = 0; l < 9; ++l) {
    A(u->wbf[l], bf * 27 * kConv[l].cout * pad32(kConv[l].cin));
    A(u->wtf[l], bf * 27 * kConv[l].cin * pad32(kConv[l].cout));
    A(u->gpad[l], f4 * 27 * kConv[l].cout * pad32(kConv[l].cin));
  }
  A(u->e1a, bf * B * P3 * C1);
  A(u->cat1, bf * B * P3 * 48);
  A(u->p1, bf * B * P3b * C1);
  A(u->e2a, bf * B * P3b * C2);
  A(u->cat2, bf * B * P3b * 96);
  A(u->p2, bf * B * P3c * C2);
  A(u->b1, bf * B * P3c * C3);
  A(u->b2, bf * B * P3c * C3);
  A(u->d2a, bf * B * P3b * C2);
  A(u->d2b, bf * B * P3b * C2);
  A(u->d1a, bf * B * P3 * C1);
  A(u->d1b, bf * B * P3 * C1);
  A(u->out, f4 * B * P3);
  A(u->g_out, f4 * B * P3);
  A(u->dy_d1b, f4 * B * P3 * C1);
  A(u->dy_d1a, f4 * B * P3 * C1);
  A(u->dy_cat1, f4 * B * P3 * 48);
  A(u->dy_d2b, f4 * B * P3b * C2);
  A(u->dy_d2a, f4 * B * P3b * C2);
  A(u->dy_cat2, f4 * B * P3b * 96);
  A(u->dy_b2, f4 * B * P3c * C3);
  A(u->dy_b1, f4 * B * P3c * C3);
  A(u->dy_p2, f4 * B * P3c * C2);
  A(u->dy_e2a, f4 * B * P3b * C2);
  A(u->dy_p1, f4 * B * P3b * C1);
  A(u->dy_e1a, f4 * B * P3 * C1);
  A(u->dpre, bf * B * P3 * C1);
  A(u->partials, sizeof(double) * 256);
  A(u->lossv, sizeof(double));
  A(u->xpatch, f4 * B * P3);
  A(u->ypatch, f4 * B * P3);
  if (rc == BRTM_OK &&
      (hipMemcpyAsync(u->params.ptr, h.data(), f4 * off, hipMemcpyHostToDevice, ctx->stream) !=
           hipSuccess ||
       hipMemsetAsync(u->am.ptr, 0, f4 * off, ctx->stream) != hipSuccess ||
       hipMemsetAsync(u->av.ptr, 0, f4 * off, ctx->stream) != hipSuccess))
    rc = BRTM_EHIP;
  if (rc == BRTM_OK) rc = repack(u);
  if (rc == BRTM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = BRTM_EHIP;
  if (rc != BRTM_OK) {
    delete u;
    return rc;
  }
  *out = u;
  return BRTM_OK;
}

int brtm_unet_destroy(brtm_unet *u) {
  if (!u) return BRTM_OK;
  (void)hipSetDevice(u->ctx->device);
  (void)hipStreamSynchronize(u->ctx->stream);
  brtm_vec_destroy(u->split_tmp);
  brtm_vec_destroy(u->znet_prev);
  brtm_vec_destroy(u->split_d);
  delete u;
  return BRTM_OK;
}

long brtm_unet_num_params(const brtm_unet *u) { return u ? u->nparams : 0; }

int brtm_unet_get_params(brtm_unet *u, float *host) {
  BRTM_REQUIRE(u && host, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(u->ctx);
  BRTM_HIP(hipMemcpyAsync(host, u->params.ptr, sizeof(float) * u->nparams, hipMemcpyDeviceToHost,
                          u->ctx->stream));
  BRTM_HIP(hipStreamSynchronize(u->ctx->stream));
  return BRTM_OK;
}

int brtm_unet_set_params(brtm_unet *u, const float *host) {
  BRTM_REQUIRE(u && host, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(u->ctx);
  BRTM_HIP(hipStreamSynchronize(u->ctx->stream));
  BRTM_HIP(hipMemcpyAsync(u->params.ptr, host, sizeof(float) * u->nparams, hipMemcpyHostToDevice,
                          u->ctx->stream));
  U_TRY(repack(u));
  BRTM_HIP(hipStreamSynchronize(u->ctx->stream));
  return BRTM_OK;
}

int brtm_unet_forward(brtm_unet *u, const float *x, float *y, int nb) {
  BRTM_REQUIRE(u && x && y && nb >= 1 && nb <= u->B, BRTM_EINVAL, "bad forward arguments");
  BRTM_LIVE(u->ctx);
  U_TRY(forward(u, x, nb));
  BRTM_HIP(hipMemcpyAsync(y, u->out.ptr, sizeof(float) * nb * u->P3, hipMemcpyDeviceToDevice,
                          u->ctx->stream));
  return BRTM_OK;
}

int brtm_unet_loss_grad(brtm_unet *u, const float *x, const float *t, int nb, double *loss,
                        float *grad, int grad_on_host) {
  BRTM_REQUIRE(u && x && t && nb >= 1 && nb <= u->B, BRTM_EINVAL, "bad arguments");
  BRTM_LIVE(u->ctx);
  U_TRY(loss_grad(u, x, t, nb, loss));
  if (grad) {
    BRTM_HIP(hipMemcpyAsync(grad, u->grads.ptr, sizeof(float) * u->nparams,
                            grad_on_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                            u->ctx->stream));
    BRTM_HIP(hipStreamSynchronize(u->ctx->stream));
  }
  return BRTM_OK;
}

int brtm_unet_train_step(brtm_unet *u, const float *x, const float *t, int nb, float lr,
                         double *loss) {
  BRTM_REQUIRE(u && x && t && nb >= 1 && nb <= u->B, BRTM_EINVAL, "bad arguments");
  BRTM_LIVE(u->ctx);
  U_TRY(loss_grad(u, x, t, nb, loss));
  u->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - std::pow(b1, (float)u->step), bc2 = 1.f - std::pow(b2, (float)u->step);
  BRTM_HIP(adam_step(u->prm(), u->grads.as<float>(), u->am.as<float>(), u->av.as<float>(),
                     u->nparams, lr, b1, b2, 1e-8f, bc1, bc2, 0.f, u->ctx->stream));
  U_TRY(repack(u));
  return BRTM_OK;
}

int brtm_unet_extract(brtm_unet *u, const float *vol, int m1, int m2, int m3, int o1, int o2,
                      int o3, float scale, float *dst) {
  BRTM_REQUIRE(u && vol && dst, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(u->ctx);
  k_extract<<<1024, 256, 0, u->ctx->stream>>>(vol, m1, m2, m3, o1, o2, o3, scale, dst, u->P);
  BRTM_HIP(hipGetLastError());
  return BRTM_OK;
}

// z = sign . rms(g) . net(sign . g / rms(g)), patch by patch.
// odd: z = sign . rms . (net(x) - net(-x)) / 2 with x = sign . g / rms, the
// odd part of the network (two evaluations per patch).
static int apply_signed(brtm_unet *u, const brtm_vec *g, brtm_vec *z, int m1, int m2, int m3, int ov,
                 float sign, bool odd = false) {
  BRTM_REQUIRE(u && g && z, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(!odd || u->B >= 2, BRTM_EINVAL, "the odd part needs a batch of at least 2");
  BRTM_REQUIRE((long)m1 * m2 * m3 == brtm_vec_size(g) && brtm_vec_size(z) == brtm_vec_size(g),
               BRTM_EINVAL, "vector size does not match the model grid");
  BRTM_REQUIRE(ov >= 0 && 2 * ov < u->P, BRTM_EINVAL, "overlap too large for the patch");
  double gg = 0;
  U_TRY(brtm_vec_dot(g, g, &gg));
  const long n = brtm_vec_size(g);
  const double rms = std::sqrt(gg / (double)n);
  BRTM_HIP(hipMemsetAsync(brtm_vec_data(z), 0, sizeof(float) * n, u->ctx->stream));
  if (rms == 0) return BRTM_OK;
  const int S = u->P - 2 * ov;
  const int k1 = (m1 + S - 1) / S, k2 = (m2 + S - 1) / S, k3 = (m3 + S - 1) / S;
  std::vector<int> orig;
  for (int c = 0; c < k3; ++c)
    for (int b = 0; b < k2; ++b)
      for (int a = 0; a < k1; ++a) {
        orig.push_back(a * S - ov);
        orig.push_back(b * S - ov);
        orig.push_back(c * S - ov);
      }
  const int np = (int)orig.size() / 3;
  const float *gp = static_cast<const float *>(brtm_vec_data(const_cast<brtm_vec *>(g)));
  float *zp = static_cast<float *>(brtm_vec_data(z));
  hipStream_t st = u->ctx->stream;
  const int per = odd ? u->B / 2 : u->B;  // patches per forward pass
  for (int p0 = 0; p0 < np; p0 += per) {
    const int nb = std::min(per, np - p0);
    for (int k = 0; k < nb; ++k) {
      const int *o = &orig[3 * (p0 + k)];
      k_extract<<<1024, 256, 0, st>>>(gp, m1, m2, m3, o[0], o[1], o[2],
                                      sign * (float)(1.0 / rms),
                                      u->xpatch.as<float>() + k * u->P3, u->P);
      if (odd)
        k_extract<<<1024, 256, 0, st>>>(gp, m1, m2, m3, o[0], o[1], o[2],
                                        -sign * (float)(1.0 / rms),
                                        u->xpatch.as<float>() + (nb + k) * u->P3, u->P);
    }
    BRTM_HIP(hipGetLastError());
    U_TRY(forward(u, u->xpatch.as<float>(), odd ? 2 * nb : nb));
    for (int k = 0; k < nb; ++k) {
      const int *o = &orig[3 * (p0 + k)];
      const float *out = u->out.as<float>();
      if (odd)
        k_insert_odd<<<1024, 256, 0, st>>>(out + k * u->P3, out + (nb + k) * u->P3, u->P, ov, zp,
                                           m1, m2, m3, o[0], o[1], o[2], sign * (float)rms);
      else
        k_insert<<<1024, 256, 0, st>>>(out + k * u->P3, u->P, ov, zp, m1, m2, m3, o[0], o[1],
                                       o[2], sign * (float)rms);
    }
    BRTM_HIP(hipGetLastError());
  }
  return BRTM_OK;
}

int brtm_unet_apply(brtm_unet *u, const brtm_vec *g, brtm_vec *z, int m1, int m2, int m3,
                    int ov) {
  BRTM_REQUIRE(u, BRTM_EINVAL, "NULL network");
  BRTM_LIVE(u->ctx);
  return apply_signed(u, g, z, m1, m2, m3, ov, 1.f);
}

int brtm_unet_set_grid(brtm_unet *u, int m1, int m2, int m3, int overlap) {
  BRTM_REQUIRE(u && m1 > 0 && m2 > 0 && m3 > 0 && overlap >= 0 && 2 * overlap < u->P,
               BRTM_EINVAL, "bad grid");
  BRTM_LIVE(u->ctx);
  u->gm1 = m1;
  u->gm2 = m2;
  u->gm3 = m3;
  u->gov = overlap;
  u->have_prev = false;
  return BRTM_OK;
}

int brtm_unet_set_split_weight(brtm_unet *u, const brtm_vec *w) {
  BRTM_REQUIRE(u, BRTM_EINVAL, "NULL network");
  BRTM_LIVE(u->ctx);
  if (w && (!u->split_tmp || brtm_vec_size(u->split_tmp) != brtm_vec_size(w))) {
    brtm_vec_destroy(u->split_tmp);
    brtm_vec_destroy(u->znet_prev);
    brtm_vec_destroy(u->split_d);
    u->split_tmp = u->znet_prev = u->split_d = nullptr;
    U_TRY(brtm_vec_create(u->ctx, brtm_vec_size(w), &u->split_tmp));
    U_TRY(brtm_vec_create(u->ctx, brtm_vec_size(w), &u->znet_prev));
    U_TRY(brtm_vec_create(u->ctx, brtm_vec_size(w), &u->split_d));
  }
  u->split_w = w;
  u->have_prev = false;
  u->guard_uses = 0;
  return BRTM_OK;
}

int brtm_unet_set_guard(brtm_unet *u, double repeat_cos, double min_cos) {
  BRTM_REQUIRE(u && repeat_cos > 0 && repeat_cos <= 1 && min_cos >= 0 && min_cos < 1, BRTM_EINVAL,
               "guard thresholds out of range");
  u->guard_repeat = repeat_cos;
  u->guard_min = min_cos;
  return BRTM_OK;
}

int brtm_unet_guard_count(const brtm_unet *u) { return u ? u->guard_uses : 0; }

int brtm_unet_set_guard_mode(brtm_unet *u, int mode) {
  BRTM_REQUIRE(u && (mode == 0 || mode == 1), BRTM_EINVAL, "guard mode must be 0 or 1");
  u->guard_mode = mode;
  u->have_prev = false;
  return BRTM_OK;
}

int brtm_unet_precond(void *user, const brtm_vec *g, brtm_vec *z) {
  brtm_unet *u = static_cast<brtm_unet *>(user);
  if (!u || u->gm1 == 0) {
    set_error("brtm_unet_precond: call brtm_unet_set_grid first");
    return BRTM_ESTATE;
  }
  BRTM_LIVE(u->ctx);
  if (u->split_w) {
    // z = W . (-s net(-W g / s)), s = rms(W g)
    BRTM_REQUIRE(brtm_vec_size(u->split_w) == brtm_vec_size(g), BRTM_EINVAL,
                 "split weight does not match the model grid");
    U_TRY(brtm_vec_copy(u->split_tmp, g));
    U_TRY(brtm_vec_mult(u->split_tmp, u->split_w));
    U_TRY(apply_signed(u, u->split_tmp, z, u->gm1, u->gm2, u->gm3, u->gov, -1.f, u->odd));
    U_TRY(brtm_vec_mult(z, u->split_w));
    if (u->guard_mode == 1) {
      // z = W^2 g + c with c the network's correction z_net - W^2 g; when c
      // nearly repeats the previous correction, only its component
      // orthogonal to that one is kept
      brtm_vec *d = u->split_d;
      U_TRY(brtm_vec_copy(d, u->split_tmp));
      U_TRY(brtm_vec_mult(d, u->split_w));           // W^2 g
      U_TRY(brtm_vec_scale_add(z, 1.0, d, -1.0));    // c
      double cc = 0, cp = 0, pp = 0;
      U_TRY(brtm_vec_dot(z, z, &cc));
      bool rep = false;
      if (u->have_prev) {
        U_TRY(brtm_vec_dot(z, u->znet_prev, &cp));
        U_TRY(brtm_vec_dot(u->znet_prev, u->znet_prev, &pp));
        rep = pp > 0 && cc > 0 && cp > u->guard_repeat * std::sqrt(cc * pp);
      }
      U_TRY(brtm_vec_copy(u->split_tmp, z));         // c as produced, for the next call
      if (rep) U_TRY(brtm_vec_scale_add(z, 1.0, u->znet_prev, -cp / pp));
      std::swap(u->split_tmp, u->znet_prev);
      u->have_prev = true;
      U_TRY(brtm_vec_add(z, d));
      double zg = 0, zz = 0, gg = 0;
      U_TRY(brtm_vec_dot(z, g, &zg));
      U_TRY(brtm_vec_dot(z, z, &zz));
      U_TRY(brtm_vec_dot(g, g, &gg));
      const bool bad = !(zg > 0) || !std::isfinite(zg) || !std::isfinite(zz) ||
                       zg < u->guard_min * std::sqrt(zz * gg);
      if (rep || bad) ++u->guard_uses;
      if (bad) U_TRY(brtm_vec_copy(z, d));
      return BRTM_OK;
    }
    // Guard: use the diagonal W^2 g (what the split gives for an identity
    // network) instead of the network's direction when that direction
    //  * is not a descent direction or not finite;
    //  * is nearly orthogonal to the gradient (cos < guard_min); or
    //  * nearly repeats the network's previous output (cos > guard_repeat).
    // The last two mark a network that has stopped following its input:
    // on gradients unlike its training data its output settles on one fixed
    // direction, which the search over previous steps has already used, and
    // the solver stalls.
    double zg = 0, zz = 0, gg = 0, zp = 0, pp = 0;
    U_TRY(brtm_vec_dot(z, g, &zg));
    U_TRY(brtm_vec_dot(z, z, &zz));
    U_TRY(brtm_vec_dot(g, g, &gg));
    bool use_diag = !(zg > 0) || !std::isfinite(zg) || !std::isfinite(zz) ||
                    zg < u->guard_min * std::sqrt(zz * gg);
    if (u->have_prev) {
      U_TRY(brtm_vec_dot(z, u->znet_prev, &zp));
      U_TRY(brtm_vec_dot(u->znet_prev, u->znet_prev, &pp));
      if (pp > 0 && zz > 0 && zp > u->guard_repeat * std::sqrt(zz * pp)) use_diag = true;
    }
    U_TRY(brtm_vec_copy(u->znet_prev, z));
    u->have_prev = true;
    if (use_diag) {
      ++u->guard_uses;
      U_TRY(brtm_vec_copy(z, u->split_tmp));  // W g
      U_TRY(brtm_vec_mult(z, u->split_w));    // W^2 g
    }
    return BRTM_OK;
  }
  // The solver's gradient is g = L^T (L m - d) = -(RTM image of the data
  // residual); the network maps RTM images to LSRTM images, so it is given -g
  // and its output is negated back: z = -rms . net(-g / rms).
  return apply_signed(u, g, z, u->gm1, u->gm2, u->gm3, u->gov, -1.f, u->odd);
}

int brtm_unet_set_odd(brtm_unet *u, int on) {
  BRTM_REQUIRE(u, BRTM_EINVAL, "NULL network");
  BRTM_REQUIRE(!on || u->B >= 2, BRTM_EINVAL, "the odd part needs a batch of at least 2");
  u->odd = on != 0;
  return BRTM_OK;
}

}  // extern "C"
