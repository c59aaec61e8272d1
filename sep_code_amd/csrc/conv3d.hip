// 3-D convolutions of the learned preconditioner on gfx950 MFMA.
//
// Forward (and input-gradient) convolutions are implicit GEMMs: M = output
// channels, N = 16 consecutive x positions (of YT rows) per wave, K = the 27
// taps x input channels, one v_mfma_f32_16x16x32_bf16 per (16-channel output
// tile, row, tap, 32-channel chunk) and one v_mfma_f32_16x16x16_bf16 per
// 16-channel remainder.  Operand lanes: lane l holds A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16] (i < 8; K = 16: 4(l/16)+i, i < 4), and receives
// D[4(l/16)+j][l%16] (j < 4) (checked on the device against a host product).  In channels-last layout
// the 8 K values of a lane are 8 consecutive channels of one position: one
// 16-byte load.  The weight gradient is a GEMM over positions (K) staged
// through LDS transposed so both operands are again 16-byte lane loads.
#include "conv3d.h"

namespace brtm {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
  const uint4 u = *reinterpret_cast<const uint4 *>(p);
  return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// four bf16 values (8 bytes) as the K = 16 MFMA's operand type
typedef short i16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ i16x4 ld4h(const uint16_t *p) {
  const uint2 u = *reinterpret_cast<const uint2 *>(p);
  return __builtin_bit_cast(i16x4, u);
}
__device__ __forceinline__ i16x4 zero4h() { return __builtin_bit_cast(i16x4, make_uint2(0, 0)); }

__host__ __device__ inline int cpad(int c) { return (c + 31) / 32 * 32; }
int grid1(long n, int bs = 256) {
  long g = (n + bs - 1) / bs;
  return (int)(g < 1 ? 1 : (g > 65535 * 4 ? 65535 * 4 : g));
}

// ------------------------------------------------------------- forward GEMM
// A wave computes 16 x positions of YT consecutive rows at one z, all output
// channels.  K runs tap by tap: for each (dz, dx) and 32-channel chunk the
// wave loads the YT + 2 input rows y0-1 .. y0+YT once (a lane's 8 K values
// are 8 consecutive channels of one position: one 16-byte load) and the
// three dy taps use them, so an input row is loaded once for three MFMAs of
// each row tile.  The weight operand of a tap is loaded once for the YT row
// tiles.  A 16-channel remainder (CIN = 16, 48) runs on the K = 16 MFMA
// (v_mfma_f32_16x16x16_bf16: 4 K values per lane, one 8-byte load) instead
// of being padded to 32.
template <int CIN, int COUT, int YT, bool RELU, bool OUTF32>
__global__ void __launch_bounds__(256)
k_conv_fwd(const uint16_t *__restrict__ x, int xcs, const uint16_t *__restrict__ W,
           const float *__restrict__ bias, void *y, int ycs, int N, int D, int H, int Wd) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts");
  constexpr int CIN_P = (CIN + 31) / 32 * 32;
  constexpr int MT = COUT / 16;
  constexpr int NC = CIN / 32;          // 32-channel chunks
  constexpr bool R16 = CIN % 32 == 16;  // and a 16-channel remainder
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wt = Wd >> 4, hg = H / YT;
  // XCD-aware order: workgroup b runs on XCD b % 8, so XCD k takes a
  // contiguous run of tiles (x fastest, then rows, then z).  Neighbouring
  // rows and planes, whose inputs the 27 taps share, then meet in that XCD's
  // L2 instead of being fetched again by another XCD.
  const int nb = gridDim.x, q8 = nb >> 3, r8 = nb & 7;
  const int bid = blockIdx.x, xcd = bid & 7, jj = bid >> 3;
  const int blk = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const long tile = (long)blk * 4 + wave;
  if (tile >= (long)N * D * hg * wt) return;  // whole wave; no barriers below
  const int xt = (int)(tile % wt);
  long r = tile / wt;
  const int y0 = (int)(r % hg) * YT;
  r /= hg;
  const int zz = (int)(r % D);
  const int nn = (int)(r / D);
  const int x0 = xt * 16;
  const int col = lane & 15, kq = lane >> 4;
  f32x4 acc[MT][YT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int t = 0; t < YT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int dz = 0; dz < 3; ++dz) {
    const int z2 = zz + dz - 1;
    if (z2 < 0 || z2 >= D) continue;  // the same for the whole wave
    const long zb = ((long)nn * D + z2) * H;
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
      const int x2 = x0 + col + dx - 1;
      const bool okx = x2 >= 0 && x2 < Wd;
#pragma unroll
      for (int cg = 0; cg < NC; ++cg) {
        const int c = cg * 32 + 8 * kq;
        bf16x8 bv[YT + 2];
#pragma unroll
        for (int rr = 0; rr < YT + 2; ++rr) {
          const int y2 = y0 + rr - 1;
          bv[rr] = okx && y2 >= 0 && y2 < H ? ld8(x + ((zb + y2) * Wd + x2) * xcs + c) : zero8();
        }
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          const int tap = dz * 9 + dy * 3 + dx;
          bf16x8 a[MT];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
            a[mt] = ld8(W + ((long)tap * COUT + mt * 16 + col) * CIN_P + c);
#pragma unroll
          for (int t = 0; t < YT; ++t)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
              acc[mt][t] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], bv[t + dy], acc[mt][t], 0, 0, 0);
        }
      }
      if constexpr (R16) {
        const int c = NC * 32 + 4 * kq;
        i16x4 bv[YT + 2];
#pragma unroll
        for (int rr = 0; rr < YT + 2; ++rr) {
          const int y2 = y0 + rr - 1;
          bv[rr] = okx && y2 >= 0 && y2 < H ? ld4h(x + ((zb + y2) * Wd + x2) * xcs + c) : zero4h();
        }
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          const int tap = dz * 9 + dy * 3 + dx;
          i16x4 a[MT];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
            a[mt] = ld4h(W + ((long)tap * COUT + mt * 16 + col) * CIN_P + c);
#pragma unroll
          for (int t = 0; t < YT; ++t)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
              acc[mt][t] =
                  __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], bv[t + dy], acc[mt][t], 0, 0, 0);
        }
      }
    }
  }
  // a lane's four channels 4kq .. 4kq+3 of a 16-channel tile are adjacent:
  // one 8-byte (BF16) or 16-byte (FP32) store when the channel stride
  // keeps them aligned
  const bool vec = (ycs & 3) == 0;
#pragma unroll
  for (int t = 0; t < YT; ++t) {
    const long opos = (((long)nn * D + zz) * H + y0 + t) * Wd + x0 + col;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int c0 = mt * 16 + 4 * kq;
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = acc[mt][t][j];
        if constexpr (!OUTF32) {
          v[j] += bias[c0 + j];
          if (RELU) v[j] = fmaxf(v[j], 0.f);
        }
      }
      if constexpr (OUTF32) {
        float *o = static_cast<float *>(y) + opos * ycs + c0;
        if (vec) {
          *reinterpret_cast<float4 *>(o) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) o[j] = v[j];
        }
      } else {
        uint16_t *o = static_cast<uint16_t *>(y) + opos * ycs + c0;
        if (vec) {
          *reinterpret_cast<uint2 *>(o) =
              make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                         (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) o[j] = f2bf(v[j]);
        }
      }
    }
  }
}

// ------------------------------------------------------------- weight GEMM
template <int CIN_P, int COUT>
__global__ void __launch_bounds__(256)
k_conv_wgrad(const uint16_t *__restrict__ x, int xcs, int cin, const uint16_t *__restrict__ dy,
             int dycs, float *dW, int N, int D, int H, int Wd, long chunk) {
  constexpr int MT = COUT / 16, NTC = CIN_P / 16, T = MT * NTC, TPW = (T + 3) / 4;
  constexpr int LP = 40;  // 32 positions + pad, rows stay 16-byte aligned
  __shared__ __attribute__((aligned(16))) uint16_t As[COUT][LP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[CIN_P][LP];
  const int tap = blockIdx.y;
  const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const long npos = (long)N * D * H * Wd;
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = p0 + chunk < npos ? p0 + chunk : npos;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  f32x4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long p = p0; p < p1; p += 32) {
    for (int i = tid; i < 32 * COUT; i += 256) {
      const int pp = i / COUT, co = i % COUT;
      const long q = p + pp;
      As[co][pp] = q < p1 ? dy[q * dycs + co] : (uint16_t)0;
    }
    for (int i = tid; i < 32 * CIN_P; i += 256) {
      const int pp = i / CIN_P, ci = i % CIN_P;
      const long q = p + pp;
      uint16_t v = 0;
      if (q < p1 && ci < cin) {
        const int xx = (int)(q % Wd);
        long r = q / Wd;
        const int yv = (int)(r % H);
        r /= H;
        const int zv = (int)(r % D);
        const int z2 = zv + dz, y2 = yv + dyy, x2 = xx + dx;
        if (z2 >= 0 && z2 < D && y2 >= 0 && y2 < H && x2 >= 0 && x2 < Wd)
          v = x[(q + ((long)dz * H + dyy) * Wd + dx) * xcs + ci];
      }
      Bs[ci][pp] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
      const int t = wave + 4 * k;
      if (t < T) {
        const int mt = t / NTC, nt = t % NTC;
        const bf16x8 a = ld8(&As[mt * 16 + col][8 * kq]);
        const bf16x8 b = ld8(&Bs[nt * 16 + col][8 * kq]);
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    const int t = wave + 4 * k;
    if (t < T) {
      const int mt = t / NTC, nt = t % NTC;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co = mt * 16 + 4 * kq + j, ci = nt * 16 + col;
        if (ci < cin) atomicAdd(dW + ((long)tap * COUT + co) * CIN_P + ci, acc[k][j]);
      }
    }
  }
}

__global__ void k_bias_grad(const uint16_t *__restrict__ dy, int dycs, int c, long npos,
                            float *db) {
  // thread t accumulates channel t % c over positions t / c, stride blockDim/c
  __shared__ float red[256];
  const int per = 256 / c;  // positions per block pass (c divides 256)
  const int ch = threadIdx.x % c, sub = threadIdx.x / c;
  float s = 0.f;
  if (sub < per)
    for (long p = (long)blockIdx.x * per + sub; p < npos; p += (long)gridDim.x * per)
      s += bf2f(dy[p * dycs + ch]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < c) {
    float t = 0.f;
    for (int k = 0; k < per; ++k) t += red[k * c + threadIdx.x];
    atomicAdd(db + threadIdx.x, t);
  }
}

__global__ void k_flip(const uint16_t *W, uint16_t *Wt, int cin, int cout) {
  // W [27][cout][cin_p] -> Wt [27][cin][cout_p]
  const int cin_p = cpad(cin), cout_p = cpad(cout);
  const long n = 27L * cin * cout_p;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int co = (int)(i % cout_p);
    const int ci = (int)((i / cout_p) % cin);
    const int tap = (int)(i / ((long)cout_p * cin));
    Wt[i] = co < cout ? W[((long)(26 - tap) * cout + co) * cin_p + ci] : (uint16_t)0;
  }
}

__global__ void k_pack(const float *Wf, uint16_t *W, int cin, int cout) {
  const int cin_p = cpad(cin);
  const long n = 27L * cout * cin_p;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % cin_p);
    const long r = i / cin_p;  // tap*cout + co
    W[i] = ci < cin ? f2bf(Wf[r * cin + ci]) : (uint16_t)0;
  }
}

__global__ void k_unpack(const float *gp, float *g, int cin, int cout) {
  const int cin_p = cpad(cin);
  const long n = 27L * cout * cin;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % cin);
    const long r = i / cin;
    g[i] = gp[r * cin_p + ci];
  }
}

// ------------------------------------------------------------- first layer
__global__ void k_first_fwd(const float *__restrict__ x, const float *__restrict__ W,
                            const float *__restrict__ b, uint16_t *y, int ycs, int cout, int N,
                            int D, int H, int Wd) {
  const long npos = (long)N * D * H * Wd;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npos;
       p += (long)gridDim.x * blockDim.x) {
    const int xx = (int)(p % Wd);
    long r = p / Wd;
    const int yv = (int)(r % H);
    r /= H;
    const int zv = (int)(r % D);
    float nb[27];
#pragma unroll
    for (int tap = 0; tap < 27; ++tap) {
      const int z2 = zv + tap / 9 - 1, y2 = yv + (tap / 3) % 3 - 1, x2 = xx + tap % 3 - 1;
      const bool ok = z2 >= 0 && z2 < D && y2 >= 0 && y2 < H && x2 >= 0 && x2 < Wd;
      nb[tap] = ok ? x[p + ((long)(tap / 9 - 1) * H + (tap / 3) % 3 - 1) * Wd + tap % 3 - 1] : 0.f;
    }
    if (cout % 8 == 0 && ycs % 8 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0) {
      // 8 channels per 16-byte store
      for (int c0 = 0; c0 < cout; c0 += 8) {
        uint32_t w[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          float s0 = b[c0 + 2 * h], s1 = b[c0 + 2 * h + 1];
#pragma unroll
          for (int tap = 0; tap < 27; ++tap) {
            s0 = fmaf(W[tap * cout + c0 + 2 * h], nb[tap], s0);
            s1 = fmaf(W[tap * cout + c0 + 2 * h + 1], nb[tap], s1);
          }
          w[h] = (uint32_t)f2bf(fmaxf(s0, 0.f)) | ((uint32_t)f2bf(fmaxf(s1, 0.f)) << 16);
        }
        *reinterpret_cast<uint4 *>(y + p * ycs + c0) = make_uint4(w[0], w[1], w[2], w[3]);
      }
    } else {
      for (int co = 0; co < cout; ++co) {
        float s = b[co];
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) s = fmaf(W[tap * cout + co], nb[tap], s);
        y[p * ycs + co] = f2bf(fmaxf(s, 0.f));
      }
    }
  }
}

// Four consecutive x positions per thread (Wd % 4 == 0, 16-byte aligned x and
// y, cout % 8 == 0): each of the nine (z, y) rows of the 3x3x3 neighbourhood
// is one 16-byte load plus its two end cells, a quarter of the loads of one
// position per thread.  The taps and their order are k_first_fwd's: the same
// arithmetic for every output.
__global__ void k_first_fwd4(const float *__restrict__ x, const float *__restrict__ W,
                             const float *__restrict__ b, uint16_t *y, int ycs, int cout, int N,
                             int D, int H, int Wd) {
  const int wq = Wd >> 2;
  const long nq = (long)N * D * H * wq;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += (long)gridDim.x * blockDim.x) {
    const int x0 = (int)(q % wq) * 4;
    long r = q / wq;
    const int yv = (int)(r % H);
    r /= H;
    const int zv = (int)(r % D);
    const long nn = r / D;
    // nb[dz * 3 + dy][k]: the input at (zv + dz - 1, yv + dy - 1, x0 + k - 1)
    float nb[9][6];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int z2 = zv + t / 3 - 1, y2 = yv + t % 3 - 1;
      float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
      float lo = 0.f, hi = 0.f;
      if (z2 >= 0 && z2 < D && y2 >= 0 && y2 < H) {
        const float *row = x + ((nn * D + z2) * H + y2) * Wd;
        c = *reinterpret_cast<const float4 *>(row + x0);
        if (x0 > 0) lo = row[x0 - 1];
        if (x0 + 4 < Wd) hi = row[x0 + 4];
      }
      nb[t][0] = lo;
      nb[t][1] = c.x;
      nb[t][2] = c.y;
      nb[t][3] = c.z;
      nb[t][4] = c.w;
      nb[t][5] = hi;
    }
    const long p0 = 4 * q;  // the position of x0
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      for (int c0 = 0; c0 < cout; c0 += 8) {
        uint32_t w[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          float s0 = b[c0 + 2 * h], s1 = b[c0 + 2 * h + 1];
#pragma unroll
          for (int tap = 0; tap < 27; ++tap) {
            const float v = nb[tap / 3][j + tap % 3];
            s0 = fmaf(W[tap * cout + c0 + 2 * h], v, s0);
            s1 = fmaf(W[tap * cout + c0 + 2 * h + 1], v, s1);
          }
          w[h] = (uint32_t)f2bf(fmaxf(s0, 0.f)) | ((uint32_t)f2bf(fmaxf(s1, 0.f)) << 16);
        }
        *reinterpret_cast<uint4 *>(y + (p0 + j) * ycs + c0) = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
  }
}

__global__ void k_first_wgrad(const float *__restrict__ x, const uint16_t *__restrict__ dy,
                              int dycs, int cout, float *dW, float *db, int N, int D, int H,
                              int Wd) {
  // blockIdx.y = tap (27) or 27 for the bias; thread accumulates all channels
  const int tap = blockIdx.y;
  const long npos = (long)N * D * H * Wd;
  float acc[64];
  for (int c = 0; c < cout; ++c) acc[c] = 0.f;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npos;
       p += (long)gridDim.x * blockDim.x) {
    float xv = 1.f;
    if (tap < 27) {
      const int xx = (int)(p % Wd);
      long r = p / Wd;
      const int yv = (int)(r % H);
      r /= H;
      const int zv = (int)(r % D);
      const int z2 = zv + tap / 9 - 1, y2 = yv + (tap / 3) % 3 - 1, x2 = xx + tap % 3 - 1;
      const bool ok = z2 >= 0 && z2 < D && y2 >= 0 && y2 < H && x2 >= 0 && x2 < Wd;
      xv = ok ? x[p + ((long)(tap / 9 - 1) * H + (tap / 3) % 3 - 1) * Wd + tap % 3 - 1] : 0.f;
    }
    for (int c = 0; c < cout; ++c) acc[c] = fmaf(bf2f(dy[p * dycs + c]), xv, acc[c]);
  }
  __shared__ float red[256];
  for (int c = 0; c < cout; ++c) {
    red[threadIdx.x] = acc[c];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(tap < 27 ? dW + tap * cout + c : db + c, red[0]);
    __syncthreads();
  }
}

// -------------------------------------------------------------------- head
__global__ void k_head_fwd(const uint16_t *__restrict__ y, int ycs, int c,
                           const float *__restrict__ W, const float *__restrict__ b,
                           const float *__restrict__ x0, float *out, long npos) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npos;
       p += (long)gridDim.x * blockDim.x) {
    float s = b[0] + x0[p];
    if (c % 8 == 0 && ycs % 8 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0) {
      // 8 channels per 16-byte load
      for (int k0 = 0; k0 < c; k0 += 8) {
        const uint4 u = *reinterpret_cast<const uint4 *>(y + p * ycs + k0);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          s = fmaf(W[k0 + 2 * h], __uint_as_float(w[h] << 16), s);
          s = fmaf(W[k0 + 2 * h + 1], __uint_as_float(w[h] & 0xffff0000u), s);
        }
      }
    } else {
      for (int k = 0; k < c; ++k) s = fmaf(W[k], bf2f(y[p * ycs + k]), s);
    }
    out[p] = s;
  }
}

__global__ void k_head_bwd(const uint16_t *__restrict__ y, int ycs, int c,
                           const float *__restrict__ W, const float *__restrict__ g, float *dy,
                           int dycs, float *dW, float *db, long npos) {
  float acc[17];
  for (int k = 0; k <= c; ++k) acc[k] = 0.f;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npos;
       p += (long)gridDim.x * blockDim.x) {
    const float gv = g[p];
    for (int k = 0; k < c; ++k) {
      dy[p * dycs + k] = gv * W[k];
      acc[k] = fmaf(gv, bf2f(y[p * ycs + k]), acc[k]);
    }
    acc[c] += gv;
  }
  __shared__ float red[256];
  for (int k = 0; k <= c; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(k < c ? dW + k : db, red[0]);
    __syncthreads();
  }
}

// ----------------------------------------------------------- elementwise
__global__ void k_relu_bwd(const float *__restrict__ dy, int dycs,
                           const uint16_t *__restrict__ y, int ycs, uint16_t *dpre, int dcs,
                           int c, long npos) {
  const long n = npos * c;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long p = i / c;
    const int k = (int)(i % c);
    const float g = bf2f(y[p * ycs + k]) > 0.f ? dy[p * dycs + k] : 0.f;
    dpre[p * dcs + k] = f2bf(g);
  }
}

__global__ void k_pool_fwd(const uint16_t *__restrict__ x, int xcs, uint16_t *y, int ycs, int c,
                           int N, int D, int H, int Wd) {
  const int d2 = D / 2, h2 = H / 2, w2 = Wd / 2;
  const long n = (long)N * d2 * h2 * w2 * c;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % c);
    long r = i / c;
    const int xx = (int)(r % w2);
    r /= w2;
    const int yv = (int)(r % h2);
    r /= h2;
    const int zv = (int)(r % d2);
    const int nn = (int)(r / d2);
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
      const long q = (((long)nn * D + 2 * zv + (o >> 2)) * H + 2 * yv + ((o >> 1) & 1)) * Wd +
                     2 * xx + (o & 1);
      s += bf2f(x[q * xcs + k]);
    }
    y[(i / c) * ycs + k] = f2bf(0.125f * s);
  }
}

__global__ void k_pool_bwd(const float *__restrict__ dy, int dycs, float *dx, int dxcs, int c,
                           int N, int D, int H, int Wd) {
  const long n = (long)N * D * H * Wd * c;  // over fine positions
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % c);
    long r = i / c;
    const int xx = (int)(r % Wd);
    long t = r / Wd;
    const int yv = (int)(t % H);
    t /= H;
    const int zv = (int)(t % D);
    const int nn = (int)(t / D);
    const long q = (((long)nn * (D / 2) + zv / 2) * (H / 2) + yv / 2) * (Wd / 2) + xx / 2;
    dx[r * dxcs + k] += 0.125f * dy[q * dycs + k];
  }
}

__global__ void k_up_fwd(const uint16_t *__restrict__ x, int xcs, uint16_t *y, int ycs, int c,
                         int N, int D, int H, int Wd) {
  // v = coarse shape; fine = 2x
  const long n = (long)N * 8 * D * H * Wd * c;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % c);
    long r = i / c;
    const int xx = (int)(r % (2 * Wd));
    long t = r / (2 * Wd);
    const int yv = (int)(t % (2 * H));
    t /= 2 * H;
    const int zv = (int)(t % (2 * D));
    const int nn = (int)(t / (2 * D));
    const long q = (((long)nn * D + zv / 2) * H + yv / 2) * Wd + xx / 2;
    y[r * ycs + k] = x[q * xcs + k];
  }
}

// The same with 16-byte moves of 8 channels (c, both strides and both
// slice offsets multiples of 8 channels).
__global__ void k_up_fwd8(const uint16_t *__restrict__ x, int xcs, uint16_t *y, int ycs, int c8,
                          int N, int D, int H, int Wd) {
  const long n = (long)N * 8 * D * H * Wd * c8;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % c8) * 8;
    long r = i / c8;
    const int xx = (int)(r % (2 * Wd));
    long t = r / (2 * Wd);
    const int yv = (int)(t % (2 * H));
    t /= 2 * H;
    const int zv = (int)(t % (2 * D));
    const int nn = (int)(t / (2 * D));
    const long q = (((long)nn * D + zv / 2) * H + yv / 2) * Wd + xx / 2;
    *reinterpret_cast<uint4 *>(y + r * ycs + k) =
        *reinterpret_cast<const uint4 *>(x + q * xcs + k);
  }
}

__global__ void k_up_bwd(const float *__restrict__ dy, int dycs, float *dx, int dxcs, int c,
                         int N, int D, int H, int Wd) {
  const long n = (long)N * D * H * Wd * c;  // over coarse positions
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % c);
    long r = i / c;
    const int xx = (int)(r % Wd);
    long t = r / Wd;
    const int yv = (int)(t % H);
    t /= H;
    const int zv = (int)(t % D);
    const int nn = (int)(t / D);
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
      const long q = (((long)nn * 2 * D + 2 * zv + (o >> 2)) * 2 * H + 2 * yv + ((o >> 1) & 1)) *
                         2 * Wd + 2 * xx + (o & 1);
      s += dy[q * dycs + k];
    }
    dx[r * dxcs + k] += s;
  }
}

__global__ void k_add_ch(const float *__restrict__ src, int scs, float *dst, int dcs, int c,
                         long npos) {
  const long n = npos * c;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long p = i / c;
    const int k = (int)(i % c);
    dst[p * dcs + k] += src[p * scs + k];
  }
}

__global__ void k_adam(float *p, const float *__restrict__ g, float *m, float *v, long n,
                       float lr, float b1, float b2, float eps, float bc1, float bc2, float wd) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + wd * p[i]);
  }
}

template <int CIN, int COUT, int YT>
hipError_t fwd_y(const Act &x, const uint16_t *W, const float *b, Act y, Vol v, bool relu,
                 bool f32, hipStream_t st) {
  const long tiles = (long)v.n * v.d * (v.h / YT) * (v.w / 16);
  const int grid = (int)((tiles + 3) / 4);
  if (f32)
    k_conv_fwd<CIN, COUT, YT, false, true><<<grid, 256, 0, st>>>(
        (const uint16_t *)x.p, x.cs, W, b, y.p, y.cs, v.n, v.d, v.h, v.w);
  else if (relu)
    k_conv_fwd<CIN, COUT, YT, true, false><<<grid, 256, 0, st>>>(
        (const uint16_t *)x.p, x.cs, W, b, y.p, y.cs, v.n, v.d, v.h, v.w);
  else
    k_conv_fwd<CIN, COUT, YT, false, false><<<grid, 256, 0, st>>>(
        (const uint16_t *)x.p, x.cs, W, b, y.p, y.cs, v.n, v.d, v.h, v.w);
  return hipGetLastError();
}

// Four rows per wave where that still gives 1024 workgroups, else one; eight
// for the 16-output-channel layers (one output tile: the rows are the
// wave's only independent MFMA chains).
template <int CIN, int COUT>
hipError_t fwd_t(const Act &x, const uint16_t *W, const float *b, Act y, Vol v, bool relu,
                 bool f32, hipStream_t st) {
  const long blocks4 = (long)v.n * v.d * (v.h / 4) * (v.w / 16) / 4;
  if constexpr (COUT == 16) {
    if (v.h % 8 == 0 && blocks4 >= 2048) return fwd_y<CIN, COUT, 8>(x, W, b, y, v, relu, f32, st);
  }
  if (v.h % 4 == 0 && blocks4 >= 1024) return fwd_y<CIN, COUT, 4>(x, W, b, y, v, relu, f32, st);
  return fwd_y<CIN, COUT, 1>(x, W, b, y, v, relu, f32, st);
}

template <int CIN_P, int COUT>
hipError_t wgrad_t(const Act &x, const Act &dy, float *dW, Vol v, hipStream_t st) {
  const long npos = (long)v.n * v.spatial();
  const long chunk = 2048;
  const dim3 grid((unsigned)((npos + chunk - 1) / chunk), 27);
  k_conv_wgrad<CIN_P, COUT><<<grid, 256, 0, st>>>((const uint16_t *)x.p, x.cs, x.c,
                                                  (const uint16_t *)dy.p, dy.cs, dW, v.n, v.d,
                                                  v.h, v.w, chunk);
  return hipGetLastError();
}

}  // namespace

// The (cin_pad, cout) pairs of the network's weight gradients (for the
// forward convolutions and their transposes).
#define BRTM_CONV_PAIRS(X) \
  X(32, 16) X(32, 32) X(32, 64) X(64, 64) X(96, 32) X(64, 16) X(64, 32) X(32, 96) X(32, 48)

// The (cin, cout) pairs of the forward convolutions (16->16, 16->32, 32->32,
// 32->64, 64->64, 96->32, 48->16) and of the transposed ones (input
// gradients: 32->16, 64->32, 32->96, 16->48).
#define BRTM_CONV_FWD_PAIRS(X)                                                          \
  X(16, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(96, 32) X(48, 16) X(32, 16) X(64, 32) \
  X(32, 96) X(16, 48)

hipError_t conv3d_fwd(const Act &x, const uint16_t *W, const float *bias, Act y, Vol v,
                      bool relu, bool out_f32, hipStream_t st) {
  if (v.w % 16 || x.c % 16 || x.cs % 8 || y.c % 16) return hipErrorInvalidValue;
#define X(CI, CO) \
  if (x.c == CI && y.c == CO) return fwd_t<CI, CO>(x, W, bias, y, v, relu, out_f32, st);
  BRTM_CONV_FWD_PAIRS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t conv3d_wgrad(const Act &x, const Act &dy, float *dW, float *db, Vol v,
                        hipStream_t st) {
  if (x.c % 8 || x.cs % 8 || dy.c % 16 || 256 % dy.c) return hipErrorInvalidValue;
  const int cp = cpad(x.c);
  hipError_t e = hipErrorInvalidValue;
#define X(CP, CO) \
  if (cp == CP && dy.c == CO) e = wgrad_t<CP, CO>(x, dy, dW, v, st);
  BRTM_CONV_PAIRS(X)
#undef X
  if (e != hipSuccess) return e;
  const long npos = (long)v.n * v.spatial();
  k_bias_grad<<<512, 256, 0, st>>>((const uint16_t *)dy.p, dy.cs, dy.c, npos, db);
  return hipGetLastError();
}

hipError_t conv3d_flip_weights(const uint16_t *W, uint16_t *Wt, int cin, int cout,
                               hipStream_t st) {
  k_flip<<<grid1(27L * cin * cpad(cout)), 256, 0, st>>>(W, Wt, cin, cout);
  return hipGetLastError();
}

hipError_t pack_weights(const float *Wf, uint16_t *W, int cin, int cout, hipStream_t st) {
  k_pack<<<grid1(27L * cout * cpad(cin)), 256, 0, st>>>(Wf, W, cin, cout);
  return hipGetLastError();
}

hipError_t unpack_grad(const float *gp, float *g, int cin, int cout, hipStream_t st) {
  k_unpack<<<grid1(27L * cout * cin), 256, 0, st>>>(gp, g, cin, cout);
  return hipGetLastError();
}

hipError_t conv3d_first_fwd(const float *x, const float *W, const float *b, Act y, Vol v,
                            hipStream_t st) {
  const bool quad = v.w % 4 == 0 && y.c % 8 == 0 && y.cs % 8 == 0 &&
                    reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(y.p) % 16 == 0;
  if (quad)
    k_first_fwd4<<<grid1((long)v.n * v.spatial() / 4), 256, 0, st>>>(
        x, W, b, (uint16_t *)y.p, y.cs, y.c, v.n, v.d, v.h, v.w);
  else
    k_first_fwd<<<grid1((long)v.n * v.spatial()), 256, 0, st>>>(x, W, b, (uint16_t *)y.p, y.cs,
                                                                 y.c, v.n, v.d, v.h, v.w);
  return hipGetLastError();
}

hipError_t conv3d_first_wgrad(const float *x, const Act &dy, float *dW, float *db, Vol v,
                              hipStream_t st) {
  if (dy.c > 64) return hipErrorInvalidValue;
  k_first_wgrad<<<dim3(256, 28), 256, 0, st>>>(x, (const uint16_t *)dy.p, dy.cs, dy.c, dW, db,
                                               v.n, v.d, v.h, v.w);
  return hipGetLastError();
}

hipError_t head_fwd(const Act &y, const float *W, const float *b, const float *x0, float *out,
                    Vol v, hipStream_t st) {
  const long npos = (long)v.n * v.spatial();
  k_head_fwd<<<grid1(npos), 256, 0, st>>>((const uint16_t *)y.p, y.cs, y.c, W, b, x0, out, npos);
  return hipGetLastError();
}

hipError_t head_bwd(const Act &y, const float *W, const float *g, float *dy, int dy_cs,
                    float *dW, float *db, Vol v, hipStream_t st) {
  if (y.c > 16) return hipErrorInvalidValue;
  const long npos = (long)v.n * v.spatial();
  k_head_bwd<<<512, 256, 0, st>>>((const uint16_t *)y.p, y.cs, y.c, W, g, dy, dy_cs, dW, db,
                                  npos);
  return hipGetLastError();
}

hipError_t relu_bwd(const float *dy, int dy_cs, const Act &y, Act dpre, Vol v, hipStream_t st) {
  const long npos = (long)v.n * v.spatial();
  k_relu_bwd<<<grid1(npos * y.c), 256, 0, st>>>(dy, dy_cs, (const uint16_t *)y.p, y.cs,
                                                (uint16_t *)dpre.p, dpre.cs, y.c, npos);
  return hipGetLastError();
}

hipError_t avgpool_fwd(const Act &x, Act y, Vol v, hipStream_t st) {
  const long n = (long)v.n * (v.d / 2) * (v.h / 2) * (v.w / 2) * x.c;
  k_pool_fwd<<<grid1(n), 256, 0, st>>>((const uint16_t *)x.p, x.cs, (uint16_t *)y.p, y.cs, x.c,
                                       v.n, v.d, v.h, v.w);
  return hipGetLastError();
}

hipError_t avgpool_bwd(const float *dy, int dy_cs, float *dx, int dx_cs, int c, Vol v,
                       hipStream_t st) {
  k_pool_bwd<<<grid1((long)v.n * v.spatial() * c), 256, 0, st>>>(dy, dy_cs, dx, dx_cs, c, v.n,
                                                                   v.d, v.h, v.w);
  return hipGetLastError();
}

hipError_t upsample_fwd(const Act &x, Act y, Vol v, hipStream_t st) {
  const bool vec = x.c % 8 == 0 && x.cs % 8 == 0 && y.cs % 8 == 0 &&
                   reinterpret_cast<uintptr_t>(x.p) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(y.p) % 16 == 0;
  if (vec) {
    k_up_fwd8<<<grid1((long)v.n * 8 * v.spatial() * (x.c / 8)), 256, 0, st>>>(
        (const uint16_t *)x.p, x.cs, (uint16_t *)y.p, y.cs, x.c / 8, v.n, v.d, v.h, v.w);
    return hipGetLastError();
  }
  k_up_fwd<<<grid1((long)v.n * 8 * v.spatial() * x.c), 256, 0, st>>>(
      (const uint16_t *)x.p, x.cs, (uint16_t *)y.p, y.cs, x.c, v.n, v.d, v.h, v.w);
  return hipGetLastError();
}

hipError_t upsample_bwd(const float *dy, int dy_cs, float *dx, int dx_cs, int c, Vol v,
                        hipStream_t st) {
  k_up_bwd<<<grid1((long)v.n * v.spatial() * c), 256, 0, st>>>(dy, dy_cs, dx, dx_cs, c, v.n,
                                                                 v.d, v.h, v.w);
  return hipGetLastError();
}

hipError_t add_channels(const float *src, int scs, float *dst, int dcs, int c, long npos,
                        hipStream_t st) {
  k_add_ch<<<grid1(npos * c), 256, 0, st>>>(src, scs, dst, dcs, c, npos);
  return hipGetLastError();
}

hipError_t adam_step(float *p, const float *g, float *m, float *v, long n, float lr, float b1,
                     float b2, float eps, float bc1, float bc2, float wd, hipStream_t st) {
  k_adam<<<grid1(n), 256, 0, st>>>(p, g, m, v, n, lr, b1, b2, eps, bc1, bc2, wd);
  return hipGetLastError();
}

}  // namespace brtm
