// Born acoustic RTM kernels for gfx950 (MI355X).
//
// The stencil is a 2.5D sweep: a workgroup owns a 64 x TY column tile of the
// x-y plane (x fastest, tiles aligned to 64 floats so each wave row is one
// 256-byte coalesced access) and streams a chunk of the slow axis z.  Each
// thread keeps a 9-deep register queue of its column's z taps; the x-y plane
// of the current slice is staged in LDS with a 4-cell halo (double-buffered,
// one barrier per slice).  The next slice's loads are issued before the
// barrier so HBM latency overlaps the current slice's arithmetic.  Workgroup
// ids are remapped so each XCD (dispatched round-robin) owns a contiguous
// run of tiles: neighbouring tiles, whose halos overlap, share that XCD's L2.
//
// Reference kernels restated (semantics, not code):
//   wave_kernel      Born/wave_fkernel.3d8o.cu:31-106   (FORM_VL)
//   wave_kernel_adj  Born/wave_fkernel.3d8o.cu:109-181  (FORM_LV)
//   damp_kernel      Born/wave_fkernel.3d8o.cu:183-210  (fused, table weights)
//   img_kernel       Born/gpu_funcs_3d.cu:111-120       (fused into FORM_VL)
//   img_add_kernel   Born/gpu_funcs_3d.cu:122-131       (fused into FORM_LV)
//   new_src_inject / new_data_inject  Born/gpu_funcs_3d.cu:64-93  (k_inject)
//   new_data_extract Born/gpu_funcs_3d.cu:95-109        (k_extract)
#include "born_kernels.h"

#include <hip/hip_fp16.h>

#include <type_traits>

namespace brtm {

namespace {

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int FMT>
__device__ __forceinline__ float snap_at(const void *s, const float *sscale, int i, int z) {
  if constexpr (FMT == FMT_F32) {
    return static_cast<const float *>(s)[i];
  } else if constexpr (FMT == FMT_BF16) {
    return bf16_to_f32(static_cast<const uint16_t *>(s)[i]) * sscale[z];
  } else if constexpr (FMT == FMT_F16) {
    return __half2float(static_cast<const __half *>(s)[i]) * sscale[z];
  } else {
    return 0.f;
  }
}

__device__ __forceinline__ float snap_dyn(int fmt, const void *s, const float *sscale, int i,
                                          int z) {
  switch (fmt) {
    case FMT_F32: return snap_at<FMT_F32>(s, sscale, i, z);
    case FMT_BF16: return snap_at<FMT_BF16>(s, sscale, i, z);
    case FMT_F16: return snap_at<FMT_F16>(s, sscale, i, z);
    default: return 0.f;
  }
}

// Edge distance of Born/wave_fkernel.3d8o.cu:194 along one axis.
__device__ __forceinline__ int edist(int i, int n) { return min(i - 4, n - 4 - i); }

// ---------------------------------------------------------------------------
// The fused stencil step.
// ---------------------------------------------------------------------------
template <int FORM, int TY, int RPT, int FMT, int ACT, int REC = REC_NONE>
__global__ void __launch_bounds__(64 * (TY / RPT))
k_step(const StepParams P) {
  static_assert((ACT == ACT_NONE) == (FMT == FMT_NONE), "an action needs a snapshot format");
  static_assert(REC == REC_NONE || FORM == FORM_VL, "receivers inside the VL step only");
  static_assert(!(REC == REC_INJECT && ACT == ACT_BORN) && !(REC == REC_EXTRACT && ACT == ACT_IMG),
                "injection goes with the adjoint, extraction with Born modelling");
  constexpr int NWV = TY / RPT;  // waves per workgroup
  constexpr int HY = 8 / NWV;    // y-halo rows loaded per wave
  static_assert(8 % NWV == 0, "waves must divide the 8 halo rows");
  constexpr int LW = 72;         // LDS row: 4 + 64 + 4
  __shared__ float plane[2][TY + 8][LW];
  __shared__ float wsh[kWMax];

  const int lane = threadIdx.x;
  const int wy = threadIdx.y;
  const int n1 = P.n1, n2 = P.n2, n3 = P.n3, nw = P.nw;
  const int s3 = n1 * n2;
  for (int e = wy * 64 + lane; e < nw; e += 64 * NWV) wsh[e] = P.wtab[e];

  // XCD-aware tile order: workgroup b runs on XCD b % 8; give each XCD a
  // contiguous range of tiles (x fastest, then y, then z chunk).
  const int nb = P.ntx * P.nty * P.ntz;
  const int bid = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = bid & 7, jj = bid >> 3;
  const int t = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const int tx = t % P.ntx;
  const int ty = (t / P.ntx) % P.nty;
  const int tz = t / (P.ntx * P.nty);

  const int x = tx * 64 + lane;
  const int y0 = ty * TY;
  const int yb = y0 + wy * RPT;
  const int z0 = 4 + tz * P.zc;
  const int z1 = min(z0 + P.zc, n3 - 4);
  const bool xmem = x < n1;
  const bool xin = (x >= 4) && (x < n1 - 4);
  const int ex = edist(x, n1);

  // Staged value at (xx, y, z): p for VL; V.w.p for LV.
  auto stage = [&](int xx, int y, int z, bool ok, float &raw) -> float {
    if (!ok) { raw = 0.f; return 0.f; }
    const int i = z * s3 + y * n1 + xx;
    const float v = P.p[i];
    raw = v;
    if constexpr (FORM == FORM_LV) {
      const int e = min(min(edist(xx, n1), edist(y, n2)), edist(z, n3));
      // p is zero outside the update set, so any weight works there.
      const float w = (e >= 0 && e < nw) ? P.wtab[e] : 1.f;
      return P.V[i] * w * v;
    } else {
      return v;
    }
  };

  // ---- prologue: register queue for z0-4 .. z0+4 -------------------------
  float q[RPT][9];
  float rq[RPT][5];  // LV: raw p at z .. z+4
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = yb + r;
    const bool ok = xmem && y < n2;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      float raw;
      q[r][k] = stage(x, y, z0 - 4 + k, ok, raw);
      if constexpr (FORM == FORM_LV) {
        if (k >= 4) rq[r][k - 4] = raw;
      }
    }
  }

  // Halo coordinates (fixed per thread).
  int hyy[HY], hyrow[HY];
#pragma unroll
  for (int j = 0; j < HY; ++j) {
    const int h = wy * HY + j;
    hyy[j] = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
    hyrow[j] = h < 4 ? h : TY + h;
  }
  const bool hxl = lane < 8;
  const int hxx = lane < 4 ? tx * 64 - 4 + lane : tx * 64 + 60 + lane;
  const int hxcol = lane < 4 ? lane : 64 + lane;
  const bool hxok = hxl && hxx >= 0 && hxx < n1;

  float hy[HY], hx[RPT];
  float cpo[RPT], cV[RPT], cS[RPT], cX[RPT];  // cX: img (VL) or m (LV)

  auto load_halo = [&](int z, float *ohy, float *ohx) {
#pragma unroll
    for (int j = 0; j < HY; ++j) {
      float raw;
      ohy[j] = stage(x, hyy[j], z, xmem && hyy[j] >= 0 && hyy[j] < n2, raw);
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      float raw;
      ohx[r] = stage(hxx, yb + r, z, hxok && (yb + r) < n2, raw);
    }
  };
  auto load_center = [&](int z, float *opo, float *oV, float *oS, float *oX) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const bool ok = xmem && y < n2;
      const int i = z * s3 + y * n1 + x;
      opo[r] = ok ? P.po[i] : 0.f;
      if constexpr (FORM == FORM_VL) oV[r] = ok ? P.V[i] : 0.f;
      if constexpr (ACT != ACT_NONE) {
        oS[r] = ok ? snap_at<FMT>(P.snap, P.sscale, i, z) : 0.f;
        oX[r] = ok ? (ACT == ACT_IMG ? P.img[i] : P.m[i]) : 0.f;
      }
    }
  };

  load_halo(z0, hy, hx);
  load_center(z0, cpo, cV, cS, cX);

  const float c0 = P.c[0];
  for (int z = z0; z < z1; ++z) {
    float(*pl)[LW] = plane[z & 1];
#pragma unroll
    for (int r = 0; r < RPT; ++r) pl[4 + wy * RPT + r][lane + 4] = q[r][4];
#pragma unroll
    for (int j = 0; j < HY; ++j) pl[hyrow[j]][lane + 4] = hy[j];
    if (hxl) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) pl[4 + wy * RPT + r][hxcol] = hx[r];
    }

    // Prefetch the next slice before the barrier.
    const int zn = z + 1;
    float nq[RPT], nraw[RPT], nhy[HY], nhx[RPT];
    float npo[RPT], nV[RPT], nS[RPT], nX[RPT];
    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int y = yb + r;
        nq[r] = stage(x, y, zn + 4, xmem && y < n2, nraw[r]);
      }
      load_halo(zn, nhy, nhx);
      load_center(zn, npo, nV, nS, nX);
    }
    __syncthreads();

    const int ez = edist(z, n3);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const int row = 4 + wy * RPT + r;
      const int col = lane + 4;
      const float cc = q[r][4];
      float acc = c0 * cc;
#pragma unroll
      for (int k = 1; k <= 4; ++k) {
        acc = fmaf(P.c[1 + 3 * (k - 1)], pl[row][col - k] + pl[row][col + k], acc);
        acc = fmaf(P.c[2 + 3 * (k - 1)], pl[row - k][col] + pl[row + k][col], acc);
        acc = fmaf(P.c[3 + 3 * (k - 1)], q[r][4 - k] + q[r][4 + k], acc);
      }
      const bool in = xin && y >= 4 && y < n2 - 4;
      if (in) {
        const int e = min(min(ex, edist(y, n2)), ez);
        const float w = e < nw ? wsh[e] : 1.f;
        const int i = z * s3 + y * n1 + x;
        float pn;
        if constexpr (FORM == FORM_VL) {
          pn = w * fmaf(cV[r], acc, 2.f * cc) - (w * w) * cpo[r];
        } else {
          pn = acc + 2.f * w * rq[r][0] - (w * w) * cpo[r];
        }
        if constexpr (ACT == ACT_BORN) pn = fmaf(P.cimg * cS[r], cX[r], pn);
        if constexpr (REC != REC_NONE) {
          // the receiver plane: k_inject / k_extract's arithmetic for the
          // (at most one) receiver of this cell, before imaging and the store
          if (z == P.rz) {
            const int ri = P.rmap[y * n1 + x];
            if (ri >= 0) {
              if constexpr (REC == REC_INJECT) {
                float a = 0.f;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                  if (P.rsinc[k] != 0.f)
                    a = fmaf(P.rsinc[k], P.rtr[ri * P.rsp + (P.rid + k) * P.rst], a);
                pn += a;
              } else {
                const float v = P.rscale ? pn * P.rscale[ri] : pn;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                  if (P.rsinc[k] != 0.f) P.rtr[ri * P.rsp + (P.rid + k) * P.rst] += v * P.rsinc[k];
              }
            }
          }
        }
        // The new field and the image are not read again in this launch:
        // stream them past the caches.
        __builtin_nontemporal_store(pn, P.pn + i);
        if constexpr (ACT == ACT_IMG)
          __builtin_nontemporal_store(fmaf(P.cimg * cS[r], pn, cX[r]), P.img + i);
      }
    }

    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[r][k] = q[r][k + 1];
        q[r][8] = nq[r];
        if constexpr (FORM == FORM_LV) {
#pragma unroll
          for (int k = 0; k < 4; ++k) rq[r][k] = rq[r][k + 1];
          rq[r][4] = nraw[r];
        }
        cpo[r] = npo[r];
        cV[r] = nV[r];
        cS[r] = nS[r];
        cX[r] = nX[r];
        hx[r] = nhx[r];
      }
#pragma unroll
      for (int j = 0; j < HY; ++j) hy[j] = nhy[j];
    }
  }
}

// ---------------------------------------------------------------------------
// The VL step with an LDS ring of six x-y planes.  Each tile loads the halo
// of slice z + 5 in the same iteration as its own (and its neighbours') z-queue
// load of that slice, so a halo line is fetched while the neighbouring tile's
// copy of it is still in L2; the plane is used four iterations later.  Slot
// (z + 4) % 6 is written at iteration z, slot z % 6 read: the slot written was
// last read at iteration z - 2, before the barrier every wave has passed, so
// one barrier per slice still suffices.
// ---------------------------------------------------------------------------
template <int TY, int RPT, int FMT, int ACT>
__global__ void __launch_bounds__(64 * (TY / RPT))
k_step_ring(const StepParams P) {
  static_assert((ACT == ACT_NONE) == (FMT == FMT_NONE), "an action needs a snapshot format");
  constexpr int NWV = TY / RPT;
  constexpr int HY = 8 / NWV;
  static_assert(8 % NWV == 0, "waves must divide the 8 halo rows");
  constexpr int LW = 72;
  constexpr int NR = 6;
  __shared__ float ring[NR][TY + 8][LW];
  __shared__ float wsh[kWMax];

  const int lane = threadIdx.x;
  const int wy = threadIdx.y;
  const int n1 = P.n1, n2 = P.n2, n3 = P.n3, nw = P.nw;
  const int s3 = n1 * n2;
  for (int e = wy * 64 + lane; e < nw; e += 64 * NWV) wsh[e] = P.wtab[e];

  const int nb = P.ntx * P.nty * P.ntz;
  const int bid = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = bid & 7, jj = bid >> 3;
  const int t = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const int tx = t % P.ntx;
  const int ty = (t / P.ntx) % P.nty;
  const int tz = t / (P.ntx * P.nty);

  const int x = tx * 64 + lane;
  const int y0 = ty * TY;
  const int yb = y0 + wy * RPT;
  const int z0 = 4 + tz * P.zc;
  const int z1 = min(z0 + P.zc, n3 - 4);
  const bool xmem = x < n1;
  const bool xin = (x >= 4) && (x < n1 - 4);
  const int ex = edist(x, n1);

  auto ld = [&](const float *a, int i, bool ok) -> float { return ok ? a[i] : 0.f; };

  float q[RPT][9];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = yb + r;
    const bool ok = xmem && y < n2;
#pragma unroll
    for (int k = 0; k < 9; ++k) q[r][k] = ld(P.p, (z0 - 4 + k) * s3 + y * n1 + x, ok);
  }

  int hyy[HY], hyrow[HY];
#pragma unroll
  for (int j = 0; j < HY; ++j) {
    const int h = wy * HY + j;
    hyy[j] = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
    hyrow[j] = h < 4 ? h : TY + h;
  }
  const bool hxl = lane < 8;
  const int hxx = lane < 4 ? tx * 64 - 4 + lane : tx * 64 + 60 + lane;
  const int hxcol = lane < 4 ? lane : 64 + lane;
  const bool hxok = hxl && hxx >= 0 && hxx < n1;

  auto load_halo = [&](int z, float *ohy, float *ohx) {
#pragma unroll
    for (int j = 0; j < HY; ++j)
      ohy[j] = ld(P.p, z * s3 + hyy[j] * n1 + x, xmem && hyy[j] >= 0 && hyy[j] < n2);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      ohx[r] = ld(P.p, z * s3 + (yb + r) * n1 + hxx, hxok && (yb + r) < n2);
  };
  // the plane of slice z: centre values c[r], halos hy / hx
  auto put_plane = [&](int z, const float *c, const float *hy, const float *hx) {
    float(*pl)[LW] = ring[z % NR];
#pragma unroll
    for (int r = 0; r < RPT; ++r) pl[4 + wy * RPT + r][lane + 4] = c[r];
#pragma unroll
    for (int j = 0; j < HY; ++j) pl[hyrow[j]][lane + 4] = hy[j];
    if (hxl) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) pl[4 + wy * RPT + r][hxcol] = hx[r];
    }
  };
  auto load_center = [&](int z, float *opo, float *oV, float *oS, float *oX) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const bool ok = xmem && y < n2;
      const int i = z * s3 + y * n1 + x;
      opo[r] = ld(P.po, i, ok);
      oV[r] = ld(P.V, i, ok);
      if constexpr (ACT != ACT_NONE) {
        oS[r] = ok ? snap_at<FMT>(P.snap, P.sscale, i, z) : 0.f;
        oX[r] = ok ? (ACT == ACT_IMG ? P.img[i] : P.m[i]) : 0.f;
      }
    }
  };

  // prologue: planes z0 .. z0+3 into the ring; the halo of z0+4 waits in hy/hx
  float hy[HY], hx[RPT];
  {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float th[HY], tx_[RPT], c[RPT];
      load_halo(z0 + k, th, tx_);
#pragma unroll
      for (int r = 0; r < RPT; ++r) c[r] = q[r][4 + k];
      put_plane(z0 + k, c, th, tx_);
    }
    load_halo(z0 + 4, hy, hx);
  }
  float cpo[RPT], cV[RPT], cS[RPT], cX[RPT];
  load_center(z0, cpo, cV, cS, cX);

  const float c0 = P.c[0];
  for (int z = z0; z < z1; ++z) {
    {
      float c[RPT];
#pragma unroll
      for (int r = 0; r < RPT; ++r) c[r] = q[r][8];
      put_plane(z + 4, c, hy, hx);
    }
    const int zn = z + 1;
    float nq[RPT], nhy[HY], nhx[RPT];
    float npo[RPT], nV[RPT], nS[RPT], nX[RPT];
    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int y = yb + r;
        nq[r] = ld(P.p, (zn + 4) * s3 + y * n1 + x, xmem && y < n2);
      }
      load_halo(zn + 4, nhy, nhx);
      load_center(zn, npo, nV, nS, nX);
    }
    __syncthreads();

    float(*pl)[LW] = ring[z % NR];
    const int ez = edist(z, n3);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const int row = 4 + wy * RPT + r;
      const int col = lane + 4;
      const float cc = q[r][4];
      float acc = c0 * cc;
#pragma unroll
      for (int k = 1; k <= 4; ++k) {
        acc = fmaf(P.c[1 + 3 * (k - 1)], pl[row][col - k] + pl[row][col + k], acc);
        acc = fmaf(P.c[2 + 3 * (k - 1)], pl[row - k][col] + pl[row + k][col], acc);
        acc = fmaf(P.c[3 + 3 * (k - 1)], q[r][4 - k] + q[r][4 + k], acc);
      }
      if (xin && y >= 4 && y < n2 - 4) {
        const int e = min(min(ex, edist(y, n2)), ez);
        const float w = e < nw ? wsh[e] : 1.f;
        const int i = z * s3 + y * n1 + x;
        float pn = w * fmaf(cV[r], acc, 2.f * cc) - (w * w) * cpo[r];
        if constexpr (ACT == ACT_BORN) pn = fmaf(P.cimg * cS[r], cX[r], pn);
        __builtin_nontemporal_store(pn, P.pn + i);
        if constexpr (ACT == ACT_IMG)
          __builtin_nontemporal_store(fmaf(P.cimg * cS[r], pn, cX[r]), P.img + i);
      }
    }

    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[r][k] = q[r][k + 1];
        q[r][8] = nq[r];
        cpo[r] = npo[r];
        cV[r] = nV[r];
        cS[r] = nS[r];
        cX[r] = nX[r];
        hx[r] = nhx[r];
      }
#pragma unroll
      for (int j = 0; j < HY; ++j) hy[j] = nhy[j];
    }
  }
}

// ---------------------------------------------------------------------------
// The VL step with two consecutive x per lane (n1 even): a wave row covers
// 128 floats with 8-byte loads and stores, and the x taps of the two cells
// come from four 8-byte LDS reads instead of sixteen 4-byte ones.
// ---------------------------------------------------------------------------
template <int FMT>
__device__ __forceinline__ float2 snap2_at(const void *s, const float *sscale, int i, int z) {
  if constexpr (FMT == FMT_F32) {
    return *reinterpret_cast<const float2 *>(static_cast<const float *>(s) + i);
  } else if constexpr (FMT == FMT_BF16) {
    const uint32_t u = *reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(s) + i);
    const float sc = sscale[z];
    return make_float2(bf16_to_f32((uint16_t)(u & 0xffffu)) * sc, bf16_to_f32((uint16_t)(u >> 16)) * sc);
  } else if constexpr (FMT == FMT_F16) {
    const __half2 h = *reinterpret_cast<const __half2 *>(static_cast<const __half *>(s) + i);
    const float2 f = __half22float2(h);
    const float sc = sscale[z];
    return make_float2(f.x * sc, f.y * sc);
  } else {
    return make_float2(0.f, 0.f);
  }
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stream_store2(float *dst, float a, float b) {
  f32x2 v;
  v.x = a;
  v.y = b;
  __builtin_nontemporal_store(v, reinterpret_cast<f32x2 *>(dst));
}

template <int TY, int RPT, int FMT, int ACT>
__global__ void __launch_bounds__(64 * (TY / RPT))
k_step_x2(const StepParams P) {
  static_assert((ACT == ACT_NONE) == (FMT == FMT_NONE), "an action needs a snapshot format");
  constexpr int NWV = TY / RPT;
  constexpr int HY = 8 / NWV;
  static_assert(8 % NWV == 0, "waves must divide the 8 halo rows");
  constexpr int LW = 136;  // 4 + 128 + 4
  __shared__ __attribute__((aligned(16))) float plane[2][TY + 8][LW];
  __shared__ float wsh[kWMax];

  const int lane = threadIdx.x;
  const int wy = threadIdx.y;
  const int n1 = P.n1, n2 = P.n2, n3 = P.n3, nw = P.nw;
  const int s3 = n1 * n2;
  for (int e = wy * 64 + lane; e < nw; e += 64 * NWV) wsh[e] = P.wtab[e];

  const int nb = P.ntx * P.nty * P.ntz;
  const int bid = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = bid & 7, jj = bid >> 3;
  const int t = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const int tx = t % P.ntx;
  const int ty = (t / P.ntx) % P.nty;
  const int tz = t / (P.ntx * P.nty);

  const int x = tx * 128 + 2 * lane;  // cells x and x + 1 (x even, n1 even)
  const int y0 = ty * TY;
  const int yb = y0 + wy * RPT;
  const int z0 = 4 + tz * P.zc;
  const int z1 = min(z0 + P.zc, n3 - 4);
  const bool xmem = x < n1;
  const bool xin = (x >= 4) && (x < n1 - 4);  // same for x + 1 since both are even-aligned
  const int ex0 = edist(x, n1), ex1 = edist(x + 1, n1);

  auto ld2 = [&](const float *a, int i, bool ok) -> float2 {
    return ok ? *reinterpret_cast<const float2 *>(a + i) : make_float2(0.f, 0.f);
  };

  float2 q[RPT][9];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = yb + r;
    const bool ok = xmem && y < n2;
#pragma unroll
    for (int k = 0; k < 9; ++k) q[r][k] = ld2(P.p, (z0 - 4 + k) * s3 + y * n1 + x, ok);
  }

  int hyy[HY], hyrow[HY];
#pragma unroll
  for (int j = 0; j < HY; ++j) {
    const int h = wy * HY + j;
    hyy[j] = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
    hyrow[j] = h < 4 ? h : TY + h;
  }
  // x halo: lanes 0,1 load columns -4..-1, lanes 2,3 load 128..131 (8 bytes each)
  const bool hxl = lane < 4;
  const int hxx = lane < 2 ? tx * 128 - 4 + 2 * lane : tx * 128 + 128 + 2 * (lane - 2);
  const int hxcol = lane < 2 ? 2 * lane : 132 + 2 * (lane - 2);
  const bool hxok = hxl && hxx >= 0 && hxx < n1;

  float2 hy[HY], hx[RPT];
  float2 cpo[RPT], cV[RPT], cS[RPT], cX[RPT];

  auto load_halo = [&](int z, float2 *ohy, float2 *ohx) {
#pragma unroll
    for (int j = 0; j < HY; ++j)
      ohy[j] = ld2(P.p, z * s3 + hyy[j] * n1 + x, xmem && hyy[j] >= 0 && hyy[j] < n2);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      ohx[r] = ld2(P.p, z * s3 + (yb + r) * n1 + hxx, hxok && (yb + r) < n2);
  };
  auto load_center = [&](int z, float2 *opo, float2 *oV, float2 *oS, float2 *oX) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const bool ok = xmem && y < n2;
      const int i = z * s3 + y * n1 + x;
      opo[r] = ld2(P.po, i, ok);
      oV[r] = ld2(P.V, i, ok);
      if constexpr (ACT != ACT_NONE) {
        oS[r] = ok ? snap2_at<FMT>(P.snap, P.sscale, i, z) : make_float2(0.f, 0.f);
        oX[r] = ld2(ACT == ACT_IMG ? P.img : P.m, i, ok);
      }
    }
  };

  load_halo(z0, hy, hx);
  load_center(z0, cpo, cV, cS, cX);

  const float c0 = P.c[0];
  for (int z = z0; z < z1; ++z) {
    float(*pl)[LW] = plane[z & 1];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      *reinterpret_cast<float2 *>(&pl[4 + wy * RPT + r][4 + 2 * lane]) = q[r][4];
#pragma unroll
    for (int j = 0; j < HY; ++j) *reinterpret_cast<float2 *>(&pl[hyrow[j]][4 + 2 * lane]) = hy[j];
    if (hxl) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
        *reinterpret_cast<float2 *>(&pl[4 + wy * RPT + r][hxcol]) = hx[r];
    }

    const int zn = z + 1;
    float2 nq[RPT], nhy[HY], nhx[RPT];
    float2 npo[RPT], nV[RPT], nS[RPT], nX[RPT];
    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int y = yb + r;
        nq[r] = ld2(P.p, (zn + 4) * s3 + y * n1 + x, xmem && y < n2);
      }
      load_halo(zn, nhy, nhx);
      load_center(zn, npo, nV, nS, nX);
    }
    __syncthreads();

    const int ez = edist(z, n3);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const int row = 4 + wy * RPT + r;
      const int col = 4 + 2 * lane;
      const float2 cc = q[r][4];
      const float2 L0 = *reinterpret_cast<const float2 *>(&pl[row][col - 4]);  // x-4, x-3
      const float2 L1 = *reinterpret_cast<const float2 *>(&pl[row][col - 2]);  // x-2, x-1
      const float2 R0 = *reinterpret_cast<const float2 *>(&pl[row][col + 2]);  // x+2, x+3
      const float2 R1 = *reinterpret_cast<const float2 *>(&pl[row][col + 4]);  // x+4, x+5
      // axis-1 taps k = 1..4 for cell a (x) and cell b (x + 1)
      const float xa[4] = {cc.y + L1.y, R0.x + L1.x, R0.y + L0.y, R1.x + L0.x};
      const float xb[4] = {R0.x + cc.x, R0.y + L1.y, R1.x + L1.x, R1.y + L0.y};
      float acc_a = c0 * cc.x, acc_b = c0 * cc.y;
#pragma unroll
      for (int k = 1; k <= 4; ++k) {
        const float2 up = *reinterpret_cast<const float2 *>(&pl[row - k][col]);
        const float2 dn = *reinterpret_cast<const float2 *>(&pl[row + k][col]);
        const float c1 = P.c[1 + 3 * (k - 1)], c2 = P.c[2 + 3 * (k - 1)], c3 = P.c[3 + 3 * (k - 1)];
        acc_a = fmaf(c1, xa[k - 1], acc_a);
        acc_b = fmaf(c1, xb[k - 1], acc_b);
        acc_a = fmaf(c2, up.x + dn.x, acc_a);
        acc_b = fmaf(c2, up.y + dn.y, acc_b);
        acc_a = fmaf(c3, q[r][4 - k].x + q[r][4 + k].x, acc_a);
        acc_b = fmaf(c3, q[r][4 - k].y + q[r][4 + k].y, acc_b);
      }
      if (xin && y >= 4 && y < n2 - 4) {
        const int eyz = min(edist(y, n2), ez);
        const int e0 = min(ex0, eyz), e1 = min(ex1, eyz);
        const float w0 = e0 < nw ? wsh[e0] : 1.f, w1 = e1 < nw ? wsh[e1] : 1.f;
        const int i = z * s3 + y * n1 + x;
        float2 pn;
        pn.x = w0 * fmaf(cV[r].x, acc_a, 2.f * cc.x) - (w0 * w0) * cpo[r].x;
        pn.y = w1 * fmaf(cV[r].y, acc_b, 2.f * cc.y) - (w1 * w1) * cpo[r].y;
        if constexpr (ACT == ACT_BORN) {
          pn.x = fmaf(P.cimg * cS[r].x, cX[r].x, pn.x);
          pn.y = fmaf(P.cimg * cS[r].y, cX[r].y, pn.y);
        }
        stream_store2(P.pn + i, pn.x, pn.y);
        if constexpr (ACT == ACT_IMG) {
          stream_store2(P.img + i, fmaf(P.cimg * cS[r].x, pn.x, cX[r].x),
                        fmaf(P.cimg * cS[r].y, pn.y, cX[r].y));
        }
      }
    }

    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[r][k] = q[r][k + 1];
        q[r][8] = nq[r];
        cpo[r] = npo[r];
        cV[r] = nV[r];
        cS[r] = nS[r];
        cX[r] = nX[r];
        hx[r] = nhx[r];
      }
#pragma unroll
      for (int j = 0; j < HY; ++j) hy[j] = nhy[j];
    }
  }
}

// ---------------------------------------------------------------------------
// The VL step with four consecutive x per lane (n1 % 4 == 0): a wave row
// covers 256 floats with 16-byte loads and stores, 16-bit snapshots are read
// 8 bytes per lane, and the x taps of the four cells come from two 16-byte
// LDS reads (the centre values are the lane's own registers).
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Streaming loads: data a step reads exactly once and no other workgroup
// touches (p_old, V, the snapshot, the image or model).  The nontemporal
// hint keeps these lines from displacing, in L2, the field lines that the
// neighbouring tiles are about to read as their halos.
__device__ __forceinline__ f32x4 ld4_stream(const float *a, int i) {
  return __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(a + i));
}

// Four consecutive snapshot values decoded to FP32; STREAM selects the
// streaming (nontemporal) load.
template <int FMT, bool STREAM = false>
__device__ __forceinline__ f32x4 snap4_at(const void *s, const float *sscale, int i, int z) {
  f32x4 r = {0.f, 0.f, 0.f, 0.f};
  if constexpr (FMT == FMT_F32) {
    const f32x4 *a = reinterpret_cast<const f32x4 *>(static_cast<const float *>(s) + i);
    if constexpr (STREAM) r = __builtin_nontemporal_load(a);
    else r = *a;
  } else {
    // 16-bit codes: four of them in 8 bytes
    const u32x2 *a = reinterpret_cast<const u32x2 *>(static_cast<const uint16_t *>(s) + i);
    u32x2 u;
    if constexpr (STREAM) u = __builtin_nontemporal_load(a);
    else u = *a;
    const float sc = sscale[z];
    if constexpr (FMT == FMT_BF16) {
      r.x = __uint_as_float(u.x << 16) * sc;
      r.y = __uint_as_float(u.x & 0xffff0000u) * sc;
      r.z = __uint_as_float(u.y << 16) * sc;
      r.w = __uint_as_float(u.y & 0xffff0000u) * sc;
    } else if constexpr (FMT == FMT_F16) {
      const float2 lo = __half22float2(__builtin_bit_cast(__half2, (unsigned int)u.x));
      const float2 hi = __half22float2(__builtin_bit_cast(__half2, (unsigned int)u.y));
      r.x = lo.x * sc;
      r.y = lo.y * sc;
      r.z = hi.x * sc;
      r.w = hi.y * sc;
    }
  }
  return r;
}

template <int TY, int RPT, int FMT, int ACT>
__global__ void __launch_bounds__(64 * (TY / RPT))
k_step_x4(const StepParams P) {
  static_assert((ACT == ACT_NONE) == (FMT == FMT_NONE), "an action needs a snapshot format");
  constexpr int NWV = TY / RPT;
  constexpr int HY = 8 / NWV;
  static_assert(8 % NWV == 0, "waves must divide the 8 halo rows");
  constexpr int LW = 264;  // 4 + 256 + 4
  __shared__ __attribute__((aligned(16))) float plane[2][TY + 8][LW];
  __shared__ float wsh[kWMax];

  const int lane = threadIdx.x;
  const int wy = threadIdx.y;
  const int n1 = P.n1, n2 = P.n2, n3 = P.n3, nw = P.nw;
  const int s3 = n1 * n2;
  for (int e = wy * 64 + lane; e < nw; e += 64 * NWV) wsh[e] = P.wtab[e];

  const int nb = P.ntx * P.nty * P.ntz;
  const int bid = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = bid & 7, jj = bid >> 3;
  const int t = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const int tx = t % P.ntx;
  const int ty = (t / P.ntx) % P.nty;
  const int tz = t / (P.ntx * P.nty);

  const int x = tx * 256 + 4 * lane;  // cells x .. x + 3 (x and n1 multiples of 4)
  const int y0 = ty * TY;
  const int yb = y0 + wy * RPT;
  const int z0 = 4 + tz * P.zc;
  const int z1 = min(z0 + P.zc, n3 - 4);
  const bool xmem = x < n1;
  // the four cells are all inside [4, n1 - 4) or all outside it
  const bool xin = (x >= 4) && (x < n1 - 4);
  int ex[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) ex[j] = edist(x + j, n1);

  auto ld4 = [&](const float *a, int i, bool ok) -> f32x4 {
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (ok) v = *reinterpret_cast<const f32x4 *>(a + i);
    return v;
  };

  f32x4 q[RPT][9];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = yb + r;
    const bool ok = xmem && y < n2;
#pragma unroll
    for (int k = 0; k < 9; ++k) q[r][k] = ld4(P.p, (z0 - 4 + k) * s3 + y * n1 + x, ok);
  }

  int hyy[HY], hyrow[HY];
#pragma unroll
  for (int j = 0; j < HY; ++j) {
    const int h = wy * HY + j;
    hyy[j] = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
    hyrow[j] = h < 4 ? h : TY + h;
  }
  // x halo: lane 0 loads columns -4..-1, lane 1 loads 256..259 (16 bytes each)
  const bool hxl = lane < 2;
  const int hxx = lane == 0 ? tx * 256 - 4 : tx * 256 + 256;
  const int hxcol = lane == 0 ? 0 : 260;
  const bool hxok = hxl && hxx >= 0 && hxx < n1;

  f32x4 hy[HY], hx[RPT];
  f32x4 cpo[RPT], cV[RPT], cS[RPT], cX[RPT];

  auto load_halo = [&](int z, f32x4 *ohy, f32x4 *ohx) {
#pragma unroll
    for (int j = 0; j < HY; ++j)
      ohy[j] = ld4(P.p, z * s3 + hyy[j] * n1 + x, xmem && hyy[j] >= 0 && hyy[j] < n2);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      ohx[r] = ld4(P.p, z * s3 + (yb + r) * n1 + hxx, hxok && (yb + r) < n2);
  };
  auto load_center = [&](int z, f32x4 *opo, f32x4 *oV, f32x4 *oS, f32x4 *oX) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const bool ok = xmem && y < n2;
      const int i = z * s3 + y * n1 + x;
      opo[r] = ld4(P.po, i, ok);
      oV[r] = ld4(P.V, i, ok);
      if constexpr (ACT != ACT_NONE) {
        const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
        oS[r] = ok ? snap4_at<FMT>(P.snap, P.sscale, i, z) : zero;
        oX[r] = ld4(ACT == ACT_IMG ? P.img : P.m, i, ok);
      }
    }
  };

  load_halo(z0, hy, hx);
  load_center(z0, cpo, cV, cS, cX);

  const float c0 = P.c[0];
  for (int z = z0; z < z1; ++z) {
    float(*pl)[LW] = plane[z & 1];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      *reinterpret_cast<f32x4 *>(&pl[4 + wy * RPT + r][4 + 4 * lane]) = q[r][4];
#pragma unroll
    for (int j = 0; j < HY; ++j) *reinterpret_cast<f32x4 *>(&pl[hyrow[j]][4 + 4 * lane]) = hy[j];
    if (hxl) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
        *reinterpret_cast<f32x4 *>(&pl[4 + wy * RPT + r][hxcol]) = hx[r];
    }

    const int zn = z + 1;
    f32x4 nq[RPT], nhy[HY], nhx[RPT];
    f32x4 npo[RPT], nV[RPT], nS[RPT], nX[RPT];
    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int y = yb + r;
        nq[r] = ld4(P.p, (zn + 4) * s3 + y * n1 + x, xmem && y < n2);
      }
      load_halo(zn, nhy, nhx);
      load_center(zn, npo, nV, nS, nX);
    }
    __syncthreads();

    const int ez = edist(z, n3);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const int row = 4 + wy * RPT + r;
      const int col = 4 + 4 * lane;
      const f32x4 C = q[r][4];
      const f32x4 L = *reinterpret_cast<const f32x4 *>(&pl[row][col - 4]);
      const f32x4 R = *reinterpret_cast<const f32x4 *>(&pl[row][col + 4]);
      const float a[12] = {L.x, L.y, L.z, L.w, C.x, C.y, C.z, C.w, R.x, R.y, R.z, R.w};
      float acc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = c0 * a[4 + j];
#pragma unroll
      for (int k = 1; k <= 4; ++k) {
        const f32x4 up = *reinterpret_cast<const f32x4 *>(&pl[row - k][col]);
        const f32x4 dn = *reinterpret_cast<const f32x4 *>(&pl[row + k][col]);
        const float c1 = P.c[1 + 3 * (k - 1)], c2 = P.c[2 + 3 * (k - 1)], c3 = P.c[3 + 3 * (k - 1)];
        const f32x4 ud = up + dn;
        const f32x4 zz = q[r][4 - k] + q[r][4 + k];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[j] = fmaf(c1, a[4 + j - k] + a[4 + j + k], acc[j]);
          acc[j] = fmaf(c2, ud[j], acc[j]);
          acc[j] = fmaf(c3, zz[j], acc[j]);
        }
      }
      if (xin && y >= 4 && y < n2 - 4) {
        const int eyz = min(edist(y, n2), ez);
        const int i = z * s3 + y * n1 + x;
        f32x4 pn;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = min(ex[j], eyz);
          const float w = e < nw ? wsh[e] : 1.f;
          float v = w * fmaf(cV[r][j], acc[j], 2.f * C[j]) - (w * w) * cpo[r][j];
          if constexpr (ACT == ACT_BORN) v = fmaf(P.cimg * cS[r][j], cX[r][j], v);
          pn[j] = v;
        }
        __builtin_nontemporal_store(pn, reinterpret_cast<f32x4 *>(P.pn + i));
        if constexpr (ACT == ACT_IMG) {
          f32x4 im;
#pragma unroll
          for (int j = 0; j < 4; ++j) im[j] = fmaf(P.cimg * cS[r][j], pn[j], cX[r][j]);
          __builtin_nontemporal_store(im, reinterpret_cast<f32x4 *>(P.img + i));
        }
      }
    }

    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[r][k] = q[r][k + 1];
        q[r][8] = nq[r];
        cpo[r] = npo[r];
        cV[r] = nV[r];
        cS[r] = nS[r];
        cX[r] = nX[r];
        hx[r] = nhx[r];
      }
#pragma unroll
      for (int j = 0; j < HY; ++j) hy[j] = nhy[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Four cells per lane on the six-plane LDS ring (the two ideas above
// together): 16-byte accesses, and each halo loaded in the same iteration as
// the neighbouring tiles' z-queue loads of that slice.  16-row tiles take
// 149 KB of LDS, which the one workgroup per CU that the registers allow
// can have.
// ---------------------------------------------------------------------------
template <int TY, int RPT, int FMT, int ACT, int REC = REC_NONE, int PF = 1, bool STREAM = false>
__global__ void __launch_bounds__(64 * (TY / RPT))
k_step_x4r(const StepParams P) {
  static_assert((ACT == ACT_NONE) == (FMT == FMT_NONE), "an action needs a snapshot format");
  static_assert(PF == 1 || (PF == 2 && ACT == ACT_NONE),
                "two slices of prefetch only for the plain step (registers)");
  static_assert(!(REC == REC_INJECT && ACT == ACT_BORN) && !(REC == REC_EXTRACT && ACT == ACT_IMG),
                "injection goes with the adjoint, extraction with Born modelling");
  constexpr int NWV = TY / RPT;
  constexpr int HY = 8 / NWV;
  static_assert(8 % NWV == 0, "waves must divide the 8 halo rows");
  constexpr int LW = 264;
  constexpr int NR = 6;
  __shared__ __attribute__((aligned(16))) float ring[NR][TY + 8][LW];
  __shared__ float wsh[kWMax];

  const int lane = threadIdx.x;
  const int wy = threadIdx.y;
  const int n1 = P.n1, n2 = P.n2, n3 = P.n3, nw = P.nw;
  const int s3 = n1 * n2;
  for (int e = wy * 64 + lane; e < nw; e += 64 * NWV) wsh[e] = P.wtab[e];

  const int nb = P.ntx * P.nty * P.ntz;
  const int bid = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = bid & 7, jj = bid >> 3;
  const int t = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const int tx = t % P.ntx;
  const int ty = (t / P.ntx) % P.nty;
  const int tz = t / (P.ntx * P.nty);

  const int x = tx * 256 + 4 * lane;
  const int y0 = ty * TY;
  const int yb = y0 + wy * RPT;
  const int z0 = 4 + tz * P.zc;
  const int z1 = min(z0 + P.zc, n3 - 4);
  const bool xmem = x < n1;
  const bool xin = (x >= 4) && (x < n1 - 4);
  int ex[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) ex[j] = edist(x + j, n1);

  auto ld4 = [&](const float *a, int i, bool ok) -> f32x4 {
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (ok) v = *reinterpret_cast<const f32x4 *>(a + i);
    return v;
  };
  // the lane's own streams (read once, by this tile only)
  auto ldc = [&](const float *a, int i, bool ok) -> f32x4 {
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (ok) {
      if constexpr (STREAM) v = ld4_stream(a, i);
      else v = *reinterpret_cast<const f32x4 *>(a + i);
    }
    return v;
  };

  f32x4 q[RPT][9];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = yb + r;
    const bool ok = xmem && y < n2;
#pragma unroll
    for (int k = 0; k < 9; ++k) q[r][k] = ld4(P.p, (z0 - 4 + k) * s3 + y * n1 + x, ok);
  }

  int hyy[HY], hyrow[HY];
#pragma unroll
  for (int j = 0; j < HY; ++j) {
    const int h = wy * HY + j;
    hyy[j] = h < 4 ? y0 - 4 + h : y0 + TY + (h - 4);
    hyrow[j] = h < 4 ? h : TY + h;
  }
  const bool hxl = lane < 2;
  const int hxx = lane == 0 ? tx * 256 - 4 : tx * 256 + 256;
  const int hxcol = lane == 0 ? 0 : 260;
  const bool hxok = hxl && hxx >= 0 && hxx < n1;

  auto load_halo = [&](int z, f32x4 *ohy, f32x4 *ohx) {
#pragma unroll
    for (int j = 0; j < HY; ++j)
      ohy[j] = ld4(P.p, z * s3 + hyy[j] * n1 + x, xmem && hyy[j] >= 0 && hyy[j] < n2);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      ohx[r] = ld4(P.p, z * s3 + (yb + r) * n1 + hxx, hxok && (yb + r) < n2);
  };
  auto put_plane = [&](int z, const f32x4 *c, const f32x4 *hy, const f32x4 *hx) {
    float(*pl)[LW] = ring[z % NR];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      *reinterpret_cast<f32x4 *>(&pl[4 + wy * RPT + r][4 + 4 * lane]) = c[r];
#pragma unroll
    for (int j = 0; j < HY; ++j) *reinterpret_cast<f32x4 *>(&pl[hyrow[j]][4 + 4 * lane]) = hy[j];
    if (hxl) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) *reinterpret_cast<f32x4 *>(&pl[4 + wy * RPT + r][hxcol]) = hx[r];
    }
  };
  auto load_center = [&](int z, f32x4 *opo, f32x4 *oV, f32x4 *oS, f32x4 *oX) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const bool ok = xmem && y < n2;
      const int i = z * s3 + y * n1 + x;
      opo[r] = ldc(P.po, i, ok);
      oV[r] = ldc(P.V, i, ok);
      if constexpr (ACT != ACT_NONE) {
        const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
        // an FP32 snapshot streams at every size (64.6 against 70.4 us for
        // FP32 imaging at 256^3); 16-bit ones only with STREAM (BF16 imaging
        // at 256^3: 77.4 against 74.6 us)
        oS[r] = ok ? snap4_at<FMT, STREAM || FMT == FMT_F32>(P.snap, P.sscale, i, z) : zero;
        // the image streams at every size (FP32 / BF16 imaging at 256^3:
        // 62.3 / 71.6 against 65.2 / 74.6 us); the model only with STREAM
        // (the Born step at 256^3: 66.0 against 64.5 us)
        const float *xs = ACT == ACT_IMG ? P.img : P.m;
        if constexpr (STREAM || ACT == ACT_IMG) oX[r] = ok ? ld4_stream(xs, i) : zero;
        else oX[r] = ld4(xs, i, ok);
      }
    }
  };

  // prologue: planes z0 .. z0+3 into the ring; the halo of z0+4 waits in hy/hx
  f32x4 hy[HY], hx[RPT];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f32x4 th[HY], tx_[RPT], c[RPT];
    load_halo(z0 + k, th, tx_);
#pragma unroll
    for (int r = 0; r < RPT; ++r) c[r] = q[r][4 + k];
    put_plane(z0 + k, c, th, tx_);
  }
  load_halo(z0 + 4, hy, hx);
  f32x4 cpo[RPT], cV[RPT], cS[RPT], cX[RPT];
  load_center(z0, cpo, cV, cS, cX);
  // PF = 2: the inputs of slice z+1 were loaded one iteration earlier and
  // wait in the p* registers while the loads of slice z+2 are in flight
  f32x4 pq[RPT], phy[HY], phx[RPT], ppo[RPT], pV[RPT], pS[RPT], pX[RPT];
  if constexpr (PF == 2) {
    if (z0 + 1 < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int y = yb + r;
        pq[r] = ld4(P.p, (z0 + 5) * s3 + y * n1 + x, xmem && y < n2);
      }
      load_halo(z0 + 5, phy, phx);
      load_center(z0 + 1, ppo, pV, pS, pX);
    }
  }

  const float c0 = P.c[0];
  for (int z = z0; z < z1; ++z) {
    {
      f32x4 c[RPT];
#pragma unroll
      for (int r = 0; r < RPT; ++r) c[r] = q[r][8];
      put_plane(z + 4, c, hy, hx);
    }
    const int zn = z + 1;
    const int zf = z + PF;  // the slice whose inputs are loaded now
    f32x4 nq[RPT], nhy[HY], nhx[RPT];
    f32x4 npo[RPT], nV[RPT], nS[RPT], nX[RPT];
    if (zf < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int y = yb + r;
        nq[r] = ld4(P.p, (zf + 4) * s3 + y * n1 + x, xmem && y < n2);
      }
      load_halo(zf + 4, nhy, nhx);
      load_center(zf, npo, nV, nS, nX);
    }
    __syncthreads();

    float(*pl)[LW] = ring[z % NR];
    const int ez = edist(z, n3);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int y = yb + r;
      const int row = 4 + wy * RPT + r;
      const int col = 4 + 4 * lane;
      const f32x4 C = q[r][4];
      const f32x4 L = *reinterpret_cast<const f32x4 *>(&pl[row][col - 4]);
      const f32x4 R = *reinterpret_cast<const f32x4 *>(&pl[row][col + 4]);
      const float a[12] = {L.x, L.y, L.z, L.w, C.x, C.y, C.z, C.w, R.x, R.y, R.z, R.w};
      float acc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = c0 * a[4 + j];
#pragma unroll
      for (int k = 1; k <= 4; ++k) {
        const f32x4 up = *reinterpret_cast<const f32x4 *>(&pl[row - k][col]);
        const f32x4 dn = *reinterpret_cast<const f32x4 *>(&pl[row + k][col]);
        const float c1 = P.c[1 + 3 * (k - 1)], c2 = P.c[2 + 3 * (k - 1)], c3 = P.c[3 + 3 * (k - 1)];
        const f32x4 ud = up + dn;
        const f32x4 zz = q[r][4 - k] + q[r][4 + k];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[j] = fmaf(c1, a[4 + j - k] + a[4 + j + k], acc[j]);
          acc[j] = fmaf(c2, ud[j], acc[j]);
          acc[j] = fmaf(c3, zz[j], acc[j]);
        }
      }
      if (xin && y >= 4 && y < n2 - 4) {
        const int eyz = min(edist(y, n2), ez);
        const int i = z * s3 + y * n1 + x;
        f32x4 pn;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = min(ex[j], eyz);
          const float w = e < nw ? wsh[e] : 1.f;
          float v = w * fmaf(cV[r][j], acc[j], 2.f * C[j]) - (w * w) * cpo[r][j];
          if constexpr (ACT == ACT_BORN) v = fmaf(P.cimg * cS[r][j], cX[r][j], v);
          pn[j] = v;
        }
        if constexpr (REC != REC_NONE) {
          // the receiver plane: the same arithmetic as k_inject / k_extract
          // for the (at most one) receiver of each cell
          if (z == P.rz) {
            const int4 rr = *reinterpret_cast<const int4 *>(P.rmap + y * n1 + x);
            const int ri[4] = {rr.x, rr.y, rr.z, rr.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              if (ri[j] < 0) continue;
              if constexpr (REC == REC_INJECT) {
                float acc = 0.f;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                  if (P.rsinc[k] != 0.f)
                    acc = fmaf(P.rsinc[k], P.rtr[ri[j] * P.rsp + (P.rid + k) * P.rst], acc);
                pn[j] += acc;
              } else {
                const float v = P.rscale ? pn[j] * P.rscale[ri[j]] : pn[j];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                  if (P.rsinc[k] != 0.f) P.rtr[ri[j] * P.rsp + (P.rid + k) * P.rst] += v * P.rsinc[k];
              }
            }
          }
        }
        __builtin_nontemporal_store(pn, reinterpret_cast<f32x4 *>(P.pn + i));
        if constexpr (ACT == ACT_IMG) {
          f32x4 im;
#pragma unroll
          for (int j = 0; j < 4; ++j) im[j] = fmaf(P.cimg * cS[r][j], pn[j], cX[r][j]);
          __builtin_nontemporal_store(im, reinterpret_cast<f32x4 *>(P.img + i));
        }
      }
    }

    if (zn < z1) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[r][k] = q[r][k + 1];
        if constexpr (PF == 2) {
          q[r][8] = pq[r];
          cpo[r] = ppo[r];
          cV[r] = pV[r];
          hx[r] = phx[r];
          pq[r] = nq[r];
          ppo[r] = npo[r];
          pV[r] = nV[r];
          phx[r] = nhx[r];
        } else {
          q[r][8] = nq[r];
          cpo[r] = npo[r];
          cV[r] = nV[r];
          cS[r] = nS[r];
          cX[r] = nX[r];
          hx[r] = nhx[r];
        }
      }
#pragma unroll
      for (int j = 0; j < HY; ++j) {
        if constexpr (PF == 2) {
          hy[j] = phy[j];
          phy[j] = nhy[j];
        } else {
          hy[j] = nhy[j];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Two time steps in one pass (temporal blocking of the plain VL step).
//
// A workgroup owns a 64 x 32 output tile of the x-y plane and streams z.  It
// reads p^n on the tile plus an 8-cell rim, p^{n-1} and V on the tile plus a
// 4-cell rim, and computes p^{n+1} = step(p^n, p^{n-1}) on that 72 x 40
// region at plane z, then p^{n+2} = step(p^{n+1}, p^n) on the tile at plane
// z - 4, where all nine z taps of p^{n+1} exist.  Both fields are written;
// nothing of p^{n+1} goes through memory between the steps, so a pair of
// steps moves 20 bytes per cell (read p^n, p^{n-1}, V; write p^{n+1},
// p^{n+2}) instead of 32.
//
// Threads: 512, one 4-cell group of the output tile each (16 groups per row,
// 32 rows).  Lanes 0..25 of each wave also own one group of the first
// step's rim (whose first step they compute), lanes 26..55 one group of the
// input rim (which they only load).  Each thread keeps p^n of its groups and V of
// its tile group for ten planes in ten-slot register buffers; the z loop is
// unrolled ten times, so every slot index is a compile-time constant and the
// buffers turn without register moves (a chunk's iteration count is rounded
// up to a multiple of ten; the extra iterations store nothing).  Plane z of
// p^n goes from the registers into a two-plane LDS buffer over the 80 x 48
// input region for the first step's x-y taps; p^{n+1} goes into a ten-plane
// LDS ring over the 72 x 40 first-step region, which gives the second step
// its x-y taps at z - 4 and its z taps at z - 8 .. z (30 + 113 KB).  One
// barrier per iteration.  The source points' contributions are computed
// once per launch for the thread's cells.
//
// The per-cell arithmetic is the four-cell ring step's, operation for
// operation (centre, then for each distance the x, y and z pairs), so each
// of the two steps gives the single step's bits.
// ---------------------------------------------------------------------------
template <int REC, bool STREAM>
__global__ void __launch_bounds__(512)
k_step2_x4(const StepParams P) {
  constexpr int TXV = 16, TY = 32;           // output tile: 16 groups of 4 x 32 rows
  constexpr int W1 = TXV + 2, H1 = TY + 8;   // first-step region, in groups / rows
  constexpr int W0 = TXV + 4, H0 = TY + 16;  // input region
  constexpr int NT = TXV * TY;
  constexpr int NS = 10;  // register slots, p^{n+1} ring planes, loop unroll
  constexpr int NRIM1 = W1 * H1 - TXV * TY;  // 208
  constexpr int NRIM0 = W0 * H0 - W1 * H1;   // 240
  __shared__ f32x4 s0[2][H0][W0];
  __shared__ f32x4 s1[NS][H1][W1];
  __shared__ float wsh[kWMax];

  const int tid = threadIdx.x;
  const int n1 = P.n1, n2 = P.n2, n3 = P.n3, nw = P.nw;
  const size_t s3 = (size_t)n1 * n2;
  for (int e = tid; e < nw; e += NT) wsh[e] = P.wtab[e];

  const int nb = P.ntx * P.nty * P.ntz;
  const int bid = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = bid & 7, jj = bid >> 3;
  const int t = (xcd < r8) ? xcd * (q8 + 1) + jj : r8 * (q8 + 1) + (xcd - r8) * q8 + jj;
  const int tx = t % P.ntx;
  const int ty = (t / P.ntx) % P.nty;
  const int tz = t / (P.ntx * P.nty);
  const int x0 = tx * 4 * TXV, y0 = ty * TY;
  const int z0 = 4 + tz * P.zc;
  const int z1 = min(z0 + P.zc, n3 - 4);

  // A: the thread's group of the output tile
  const int ro = tid / TXV, co = tid % TXV;
  const int xa = x0 + 4 * co, ya = y0 + ro;
  // X: a group of the first step's rim (B: rows 0-3 and 36-39 of that
  // region, then its left and right columns) or of the input rim (C: rows
  // 0-3 and 44-47 of the input region, then its side columns)
  // (every wave holds 26 rim groups of the first step and 30 of the input)
  constexpr int NWV = NT / 64, PB = NRIM1 / NWV, PC = NRIM0 / NWV;
  static_assert(PB * NWV == NRIM1 && PC * NWV == NRIM0 && PB + PC <= 64, "rim split");
  const int wv = tid >> 6, ln = tid & 63;
  const bool hasB = ln < PB;
  const bool hasC = ln >= PB && ln < PB + PC;
  int rx, cx, xx, yx;   // LDS position in the input region, grid position
  int rb = 0, cb = 0;   // B: position in the first-step region
  if (hasB) {
    const int ib = wv * PB + ln;
    if (ib < 8 * W1) {
      const int rr = ib / W1;
      rb = rr < 4 ? rr : TY + rr;
      cb = ib % W1;
    } else {
      const int u = ib - 8 * W1;
      rb = 4 + (u >> 1);
      cb = (u & 1) ? W1 - 1 : 0;
    }
    rx = rb + 4;
    cx = cb + 1;
  } else {
    const int uc = hasC ? wv * PC + ln - PB : 0;
    if (uc < 8 * W0) {
      const int rr = uc / W0;
      rx = rr < 4 ? rr : TY + 8 + rr;
      cx = uc % W0;
    } else {
      const int u = uc - 8 * W0;
      rx = 4 + (u >> 1);
      cx = (u & 1) ? W0 - 1 : 0;
    }
  }
  xx = x0 - 8 + 4 * cx;
  yx = y0 - 8 + rx;

  // in the grid (may be loaded) / in the update set (computed and kept)
  const bool hasX = hasB || hasC;
  const bool mA = xa < n1 && ya < n2;
  const bool mX = hasX && xx >= 0 && xx < n1 && yx >= 0 && yx < n2;
  const bool uA = xa >= 4 && xa < n1 - 4 && ya >= 4 && ya < n2 - 4;
  const bool uB = hasB && xx >= 4 && xx < n1 - 4 && yx >= 4 && yx < n2 - 4;
  const int oA = mA ? ya * n1 + xa : 0;
  const int oX = mX ? yx * n1 + xx : 0;
  // the sponge's x-y edge distance per cell (clamped at 0: cells outside the
  // update set are computed and discarded)
  int exyA[4], exyB[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    exyA[j] = max(min(edist(xa + j, n1), edist(ya, n2)), 0);
    exyB[j] = max(min(edist(xx + j, n1), edist(yx, n2)), 0);
  }

  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  auto ldp = [&](int z, int off, bool m) -> f32x4 {
    f32x4 v = zero;
    if (m && z >= 0 && z < n3) v = *reinterpret_cast<const f32x4 *>(P.p + (size_t)z * s3 + off);
    return v;
  };
  // p^{n-1} and V: read by the first step of this tile only where the cell
  // is in the update set
  auto ldc = [&](const float *a, int z, int off, bool m) -> f32x4 {
    f32x4 v = zero;
    if (m && z >= 4 && z < n3 - 4) {
      if constexpr (STREAM) v = ld4_stream(a + (size_t)z * s3, off);
      else v = *reinterpret_cast<const f32x4 *>(a + (size_t)z * s3 + off);
    }
    return v;
  };
  auto put0 = [&](int z, const f32x4 &a, const f32x4 &x) {
    f32x4(*pl)[W0] = s0[z % 2];
    pl[ro + 8][co + 2] = a;
    if (hasX) pl[rx][cx] = x;
  };
  // One VL update of the 4-cell group at region position (row, col) of the
  // plane pl (width W); q[k] is the group's value at z offset k - 4; p_old
  // po, V v, the sponge weight of edge distance min(exy, ez).
  const float c0 = P.c[0];
  auto update = [&](const f32x4 *pl, int W, int row, int col, const f32x4 *q, const f32x4 &po,
                    const f32x4 &v, const int *exy, int ez) -> f32x4 {
    const f32x4 C = q[4];
    const f32x4 L = pl[row * W + col - 1];
    const f32x4 R = pl[row * W + col + 1];
    const float a[12] = {L.x, L.y, L.z, L.w, C.x, C.y, C.z, C.w, R.x, R.y, R.z, R.w};
    float acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = c0 * a[4 + j];
#pragma unroll
    for (int k = 1; k <= 4; ++k) {
      const f32x4 up = pl[(row - k) * W + col];
      const f32x4 dn = pl[(row + k) * W + col];
      const float c1 = P.c[1 + 3 * (k - 1)], c2 = P.c[2 + 3 * (k - 1)], c3 = P.c[3 + 3 * (k - 1)];
      const f32x4 ud = up + dn;
      const f32x4 zz = q[4 - k] + q[4 + k];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[j] = fmaf(c1, a[4 + j - k] + a[4 + j + k], acc[j]);
        acc[j] = fmaf(c2, ud[j], acc[j]);
        acc[j] = fmaf(c3, zz[j], acc[j]);
      }
    }
    f32x4 pn;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = min(exy[j], ez);
      const float w = e < nw ? wsh[e] : 1.f;
      pn[j] = w * fmaf(v[j], acc[j], 2.f * C[j]) - (w * w) * po[j];
    }
    return pn;
  };
  // The source points of plane rz: k_inject's sum for the cells of a group
  // at time index rid with taps rs, and which cells hold a point.
  auto source_sum = [&](int x, int y, int rid, const float *rs, f32x4 &add, int &mask) {
    const int4 rr = *reinterpret_cast<const int4 *>(P.rmap + y * n1 + x);
    const int ri[4] = {rr.x, rr.y, rr.z, rr.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float acc = 0.f;
      if (ri[j] >= 0) {
        mask |= 1 << j;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (rs[k] != 0.f) acc = fmaf(rs[k], P.rtr[ri[j] * P.rsp + (rid + k) * P.rst], acc);
      }
      add[j] = acc;
    }
  };
  f32x4 addA1 = zero, addB1 = zero, addA2 = zero;  // step 1 (tile, rim), step 2
  int mskA = 0, mskB = 0;
  if constexpr (REC == REC_INJECT) {
    if (uA) {
      source_sum(xa, ya, P.rid, P.rsinc, addA1, mskA);
      int m2 = 0;
      source_sum(xa, ya, P.rid2, P.rsinc2, addA2, m2);
    }
    if (uB) source_sum(xx, yx, P.rid, P.rsinc, addB1, mskB);
  }
  auto inject = [&](f32x4 &pn, const f32x4 &add, int mask) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (mask & (1 << j)) pn[j] += add[j];
  };

  // Iterations ts, ts+1, ...: the first step at plane z, the second at z - 4.
  // Plane p's values sit in slot (p - ts) % NS of qa (p^n, tile group), qx
  // (p^n, rim group) and vd (V, tile group), p_old and the rim group's V in
  // slot (p - ts) % 2, and p^{n+1} of plane p in ring plane p % NS.
  const int ts = z0 - 4, te = z1 + 4;
  f32x4 qa[NS], qx[NS], vd[NS], poA[2], poB[2], vB[2];
#pragma unroll
  for (int k = 0; k < NS; ++k) vd[k] = zero;
  // planes ts-4 .. ts+4 (slots 6..9, 0..4)
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    qa[(k + 6) % NS] = ldp(ts - 4 + k, oA, mA);
    qx[(k + 6) % NS] = ldp(ts - 4 + k, oX, mX);
  }
  poA[0] = ldc(P.po, ts, oA, uA);
  poB[0] = ldc(P.po, ts, oX, uB);
  vd[0] = ldc(P.V, ts, oA, uA);
  vB[0] = ldc(P.V, ts, oX, uB);

  auto iteration = [&](int z, auto uc) {
    constexpr int u = decltype(uc)::value;
    constexpr int cur = u % 2, nxt = (u + 1) % 2;
    auto sl = [](int k) constexpr { return ((k % NS) + NS) % NS; };
    put0(z, qa[sl(u)], qx[sl(u)]);
    // loads for the next iteration (out-of-grid planes load nothing)
    qa[sl(u + 5)] = ldp(z + 5, oA, mA);
    qx[sl(u + 5)] = ldp(z + 5, oX, mX);
    poA[nxt] = ldc(P.po, z + 1, oA, uA);
    vd[sl(u + 1)] = ldc(P.V, z + 1, oA, uA);
    poB[nxt] = ldc(P.po, z + 1, oX, uB);
    vB[nxt] = ldc(P.V, z + 1, oX, uB);
    __syncthreads();

    // first step at plane z over the tile and its rim (computed everywhere,
    // kept in the update set)
    const bool zin = z >= 4 && z < n3 - 4;
    const int ez = max(edist(z, n3), 0);
    const f32x4 *pl0 = &s0[z % 2][0][0];
    f32x4(*r1)[W1] = s1[z % NS];
    f32x4 za[9], zx[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      za[k] = qa[sl(u - 4 + k)];
      zx[k] = qx[sl(u - 4 + k)];
    }
    f32x4 pA = update(pl0, W0, ro + 8, co + 2, za, poA[cur], vd[sl(u)], exyA, ez);
    if constexpr (REC == REC_INJECT)
      if (z == P.rz) inject(pA, addA1, mskA);
    if (!(zin && uA)) pA = zero;
    r1[ro + 4][co + 1] = pA;
    if (hasB) {
      f32x4 pB = update(pl0, W0, rx, cx, zx, poB[cur], vB[cur], exyB, ez);
      if constexpr (REC == REC_INJECT)
        if (z == P.rz) inject(pB, addB1, mskB);
      if (!(zin && uB)) pB = zero;
      r1[rb][cb] = pB;
    }

    // second step at plane z - 4 over the tile: z taps of p^{n+1} at
    // z - 8 .. z from the ring (plane z is the thread's own pA)
    const int z2 = z - 4;
    f32x4 z1q[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) z1q[k] = s1[(z + 2 + k) % NS][ro + 4][co + 1];
    z1q[8] = pA;
    const f32x4 *pl1 = &s1[(z + 6) % NS][0][0];
    f32x4 p2 = update(pl1, W1, ro + 4, co + 1, z1q, qa[sl(u - 4)], vd[sl(u - 4)], exyA,
                      max(edist(z2, n3), 0));
    if (z2 >= z0 && z2 < z1 && uA) {
      if constexpr (REC == REC_INJECT)
        if (z2 == P.rz) inject(p2, addA2, mskA);
      const size_t i = (size_t)z2 * s3 + oA;
      __builtin_nontemporal_store(z1q[4], reinterpret_cast<f32x4 *>(P.pn + i));
      __builtin_nontemporal_store(p2, reinterpret_cast<f32x4 *>(P.pn2 + i));
    }
  };

  const int nit = ((te - ts + NS - 1) / NS) * NS;
  for (int zt = ts; zt < ts + nit; zt += NS) {
    iteration(zt + 0, std::integral_constant<int, 0>{});
    iteration(zt + 1, std::integral_constant<int, 1>{});
    iteration(zt + 2, std::integral_constant<int, 2>{});
    iteration(zt + 3, std::integral_constant<int, 3>{});
    iteration(zt + 4, std::integral_constant<int, 4>{});
    iteration(zt + 5, std::integral_constant<int, 5>{});
    iteration(zt + 6, std::integral_constant<int, 6>{});
    iteration(zt + 7, std::integral_constant<int, 7>{});
    iteration(zt + 8, std::integral_constant<int, 8>{});
    iteration(zt + 9, std::integral_constant<int, 9>{});
  }
}

// ---------------------------------------------------------------------------
// Injection / extraction (one thread per point).
// ---------------------------------------------------------------------------
__global__ void k_inject(const InjectParams P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.npts) return;
  const int loc = P.locs[j];
  if (loc < 0) return;
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (P.sinc8[k] != 0.f) acc = fmaf(P.sinc8[k], P.tr[j * P.sp + (long)(P.id + k) * P.st], acc);
  }
  const float add = P.scale * acc;
  atomicAdd(P.p + loc, add);
  if (P.img) {
    const float s = snap_dyn(P.fmt, P.snap, P.sscale, loc, loc / P.n12);
    atomicAdd(P.img + loc, P.cimg * s * add);
  }
}

__global__ void k_extract(int npts, const int *__restrict__ locs, float *tr, long st_t, int id,
                          float s0, float s1, float s2, float s3, float s4, float s5, float s6,
                          float s7, const float *__restrict__ p, const float *__restrict__ rs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  const int loc = locs[j];
  if (loc < 0) return;
  const float v = rs ? p[loc] * rs[j] : p[loc];
  const float s[8] = {s0, s1, s2, s3, s4, s5, s6, s7};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (s[k] != 0.f) tr[(long)(id + k) * st_t + j] += v * s[k];
  }
}

// Sponge weight of a cell, 0 outside the update set.
__device__ __forceinline__ float cell_weight(const float *wtab, int nw, int i1, int i2, int i3,
                                             int n1, int n2, int n3) {
  const int e = min(min(edist(i1, n1), edist(i2, n2)), edist(i3, n3));
  if (e < 0) return 0.f;
  return e < nw ? wtab[e] : 1.f;
}

__global__ void k_vw_scale(const float *__restrict__ V, const float *__restrict__ wtab, int nw,
                           const float *__restrict__ in, float *out, int n1, int n2, int n3) {
  const long n = (long)n1 * n2 * n3;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int i1 = (int)(i % n1), i2 = (int)((i / n1) % n2), i3 = (int)(i / ((long)n1 * n2));
    out[i] = V[i] * cell_weight(wtab, nw, i1, i2, i3, n1, n2, n3) * in[i];
  }
}

__global__ void k_receiver_scale(const float *__restrict__ V, const float *__restrict__ wtab,
                                 int nw, const int *__restrict__ locs, int npts, float *rs, int n1,
                                 int n2, int n3) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  const int loc = locs[j];
  float v = 0.f;
  if (loc >= 0) {
    const int i1 = loc % n1, i2 = (loc / n1) % n2, i3 = loc / (n1 * n2);
    const float d = V[loc] * cell_weight(wtab, nw, i1, i2, i3, n1, n2, n3);
    v = d > 0.f ? 1.f / d : 0.f;
  }
  rs[j] = v;
}

// ---------------------------------------------------------------------------
// Snapshot codec: scale = max |s| over each z slice; q = s / scale.
// ---------------------------------------------------------------------------
// 256 threads (4 waves) per block; blockIdx.y is the slice.  A slice whose
// length is a multiple of 4 is read 16 bytes per lane (hipMalloc'ed fields
// are 16-byte aligned, so every slice is).  The maximum is exact, so the
// result does not depend on the order of the reduction.
__global__ void __launch_bounds__(256) k_slice_absmax(const float *__restrict__ s,
                                                      unsigned int *smax, int n12) {
  const int z = blockIdx.y;
  const float *sl = s + (long)z * n12;
  const int stride = gridDim.x * blockDim.x;
  float m = 0.f;
  if ((n12 & 3) == 0) {
    const f32x4 *v = reinterpret_cast<const f32x4 *>(sl);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n12 >> 2); i += stride) {
      const f32x4 a = v[i];
      m = fmaxf(m, fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))));
    }
  } else {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n12; i += stride)
      m = fmaxf(m, fabsf(sl[i]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  // Non-negative floats order like their bit patterns.
  if (threadIdx.x == 0)
    atomicMax(smax + z, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

template <int FMT>
__global__ void k_encode(const float *__restrict__ s, void *dst, const float *__restrict__ sscale,
                         int n12) {
  const int z = blockIdx.y;
  const float sc = sscale[z];
  // A slice whose maximum is below 1e-30 (including denormals, whose
  // reciprocal overflows) is stored as zeros.
  const float inv = sc >= 1e-30f ? 1.f / sc : 0.f;
  const long base = (long)z * n12;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n12; i += gridDim.x * blockDim.x) {
    const float v = s[base + i] * inv;
    if constexpr (FMT == FMT_BF16)
      static_cast<uint16_t *>(dst)[base + i] = f32_to_bf16_rne(v);
    else
      static_cast<__half *>(dst)[base + i] = __float2half_rn(v);
  }
}

__global__ void k_decode(int fmt, const void *src, const float *sscale, float *dst, int n12) {
  const int z = blockIdx.y;
  const long base = (long)z * n12;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n12; i += gridDim.x * blockDim.x)
    dst[base + i] = snap_dyn(fmt, src, sscale, (int)(base + i), z);
}

__global__ void k_square_acc(int fmt, const void *src, const float *sscale, float *acc, int n12) {
  const int z = blockIdx.y;
  const long base = (long)z * n12;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n12; i += gridDim.x * blockDim.x) {
    const float v = snap_dyn(fmt, src, sscale, (int)(base + i), z);
    acc[base + i] = fmaf(v, v, acc[base + i]);
  }
}

// ---------------------------------------------------------------------------
// FP64 dot product: fixed grid, per-block partials, single-block finish, so
// the result is deterministic run to run.
// ---------------------------------------------------------------------------
__global__ void k_dot_partial(const float *__restrict__ a, const float *__restrict__ b, long n,
                              double *partials) {
  double acc = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void k_dot_final(const double *partials, int np, double *out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partials[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_axpby(float *y, const float *__restrict__ x, double a, double b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = (float)(a * (double)y[i] + b * (double)x[i]);
}

__global__ void k_mul(float *y, const float *__restrict__ x, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] *= x[i];
}

// Zero the cells outside the update set (the 4-cell halo every step leaves
// untouched) of nfields fields of [n3][n2][n1] floats, stride apart: one
// block row per plane (blockIdx.y), one field per blockIdx.z.
__global__ void k_zero_halo(float *p, long stride, int n1, int n2, int n3) {
  const int z = blockIdx.y;
  float *f = p + (long)blockIdx.z * stride + (long)z * n1 * n2;
  const bool slab = z < 4 || z >= n3 - 4;
  const int rows = 8 * n1;
  const int total = slab ? n1 * n2 : rows + (n2 - 8) * 8;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int y, x;
    if (slab) {
      y = t / n1;
      x = t - y * n1;
    } else if (t < rows) {
      const int r = t / n1;
      x = t - r * n1;
      y = r < 4 ? r : n2 - 8 + r;
    } else {
      const int e = t - rows;
      y = 4 + e / 8;
      const int k = e - (e / 8) * 8;
      x = k < 4 ? k : n1 - 8 + k;
    }
    f[(long)y * n1 + x] = 0.f;
  }
}

__global__ void k_fill(float *y, float v, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    y[i] = v;
}

// D_z (Born/laplac_3d.cpp:46-65 forward, :4-22 adjoint) on [n3][n12].
__global__ void k_dz(int adj, const float *__restrict__ in, float *out, int n12, int n3,
                     int add) {
  const long n = (long)n12 * n3;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i / n12);
    float v;
    if (!adj) {  // d[z] = m[z-1] (z >= 1) - m[z] (z <= n3-2)
      v = (z >= 1 ? in[i - n12] : 0.f) - (z <= n3 - 2 ? in[i] : 0.f);
    } else {     // m[z] = d[z+1] - d[z] for z <= n3-2
      v = z <= n3 - 2 ? in[i + n12] - in[i] : 0.f;
    }
    out[i] = add ? out[i] + v : v;
  }
}

__global__ void k_embed(int adj, const float *model, float *comp, int m1, int m2, int m3, int n1,
                        int n2, int n3, int o1, int o2, int o3) {
  const long n = (long)n1 * n2 * n3;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int i1 = (int)(i % n1);
    const int i2 = (int)((i / n1) % n2);
    const int i3 = (int)(i / ((long)n1 * n2));
    const int j1 = i1 - o1, j2 = i2 - o2, j3 = i3 - o3;
    const bool ok = j1 >= 0 && j1 < m1 && j2 >= 0 && j2 < m2 && j3 >= 0 && j3 < m3;
    const long j = ((long)j3 * m2 + j2) * m1 + j1;
    if (!adj) {
      comp[i] = ok ? model[j] : 0.f;
    } else if (ok) {
      const_cast<float *>(model)[j] += comp[i];
    }
  }
}

// Device traces [ntb][nrec] -> data [nrec][nt]:
//   w[t] = tr[t+4];  d[t] = (w[t-1] - w[t]) * inv_dt   (w[-1] = 0).
// 32x32 LDS tiles so both sides are coalesced.
__global__ void k_trace_to_data(const float *__restrict__ tr, float *data, int nrec, int ntb,
                                int nt, float inv_dt, int add) {
  __shared__ float tile[33][33];
  const int r0 = blockIdx.x * 32, t0 = blockIdx.y * 32;
  const int tx = threadIdx.x, tyy = threadIdx.y;  // 32 x 8
  // load w[t0-1 .. t0+31] for 32 receivers
  for (int k = tyy; k < 33; k += 8) {
    const int t = t0 - 1 + k;
    const int r = r0 + tx;
    float v = 0.f;
    if (r < nrec && t >= 0 && t < nt) v = tr[(long)(t + 4) * nrec + r];
    tile[k][tx] = v;
  }
  __syncthreads();
  for (int k = tyy; k < 32; k += 8) {
    const int r = r0 + k;
    const int t = t0 + tx;
    if (r < nrec && t < nt) {
      const float d = (tile[tx][k] - tile[tx + 1][k]) * inv_dt;
      float *o = data + (long)r * nt + t;
      *o = add ? *o + d : d;
    }
  }
}

// Adjoint of k_trace_to_data: tr[t+4] = (d[t+1] - d[t]) * inv_dt (d[nt] = 0),
// other samples of tr zero.
__global__ void k_data_to_trace(const float *__restrict__ data, float *tr, int nrec, int ntb,
                                int nt, float inv_dt) {
  __shared__ float tile[33][33];
  const int r0 = blockIdx.x * 32, t0 = blockIdx.y * 32;  // t0 indexes trace samples
  const int tx = threadIdx.x, tyy = threadIdx.y;
  // data samples t = t0-4 .. t0+28 (33 values) for 32 receivers
  for (int k = tyy; k < 32; k += 8) {
    const int r = r0 + k;
    for (int kk = tx; kk < 33; kk += 32) {
      const int t = t0 - 4 + kk;
      float v = 0.f;
      if (r < nrec && t >= 0 && t < nt) v = data[(long)r * nt + t];
      tile[kk][k] = v;
    }
  }
  __syncthreads();
  for (int k = tyy; k < 32; k += 8) {
    const int ts = t0 + k;  // trace sample
    const int r = r0 + tx;
    if (r < nrec && ts < ntb) {
      const int t = ts - 4;
      float v = 0.f;
      if (t >= 0 && t < nt) v = (tile[k + 1][tx] - tile[k][tx]) * inv_dt;
      tr[(long)ts * nrec + r] = v;
    }
  }
}

// One 16-byte element per thread: short-lived waves keep every CU's memory
// pipeline full (a grid-stride loop waits for each store before its next
// load, measured 20 % slower on MI355X).
__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// out = sum of NIN input streams, 16 bytes per thread per stream (the
// stencil step's read/write mix for NIN = 3); NT: nontemporal store.
template <int NIN, bool NT>
__global__ void k_sum16(const f32x4 *__restrict__ a, const f32x4 *__restrict__ b,
                        const f32x4 *__restrict__ c, f32x4 *out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f32x4 v = a[i];
  if constexpr (NIN >= 2) v += b[i];
  if constexpr (NIN >= 3) v += c[i];
  if constexpr (NT) __builtin_nontemporal_store(v, out + i);
  else out[i] = v;
}

__global__ void k_transpose(const float *__restrict__ in, float *out, long rows, long cols) {
  __shared__ float tile[32][33];
  const long c0 = (long)blockIdx.x * 32, r0 = (long)blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const long r = r0 + k, c = c0 + threadIdx.x;
    if (r < rows && c < cols) tile[k][threadIdx.x] = in[r * cols + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const long c = c0 + k, r = r0 + threadIdx.x;
    if (r < rows && c < cols) out[c * rows + r] = tile[threadIdx.x][k];
  }
}

int grid_for(long n, int bs) {
  long g = (n + bs - 1) / bs;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

template <int TY, int RPT, int ACT>
hipError_t launch_x2_a(int fmt, const StepParams &P, hipStream_t st) {
  const dim3 block(64, TY / RPT);
  const dim3 grid(P.ntx * P.nty * P.ntz);
  switch (fmt) {
    case FMT_F32: k_step_x2<TY, RPT, FMT_F32, ACT><<<grid, block, 0, st>>>(P); break;
    case FMT_BF16: k_step_x2<TY, RPT, FMT_BF16, ACT><<<grid, block, 0, st>>>(P); break;
    case FMT_F16: k_step_x2<TY, RPT, FMT_F16, ACT><<<grid, block, 0, st>>>(P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int TY, int RPT>
hipError_t launch_x2_t(int act, int fmt, const StepParams &P, hipStream_t st) {
  if (act == ACT_NONE) {
    if (fmt != FMT_NONE) return hipErrorInvalidValue;
    k_step_x2<TY, RPT, FMT_NONE, ACT_NONE><<<dim3(P.ntx * P.nty * P.ntz), dim3(64, TY / RPT), 0,
                                             st>>>(P);
    return hipGetLastError();
  }
  if (act == ACT_BORN) return launch_x2_a<TY, RPT, ACT_BORN>(fmt, P, st);
  if (act == ACT_IMG) return launch_x2_a<TY, RPT, ACT_IMG>(fmt, P, st);
  return hipErrorInvalidValue;
}

template <int TY, int RPT, int ACT>
hipError_t launch_ring_a(int fmt, const StepParams &P, hipStream_t st) {
  const dim3 block(64, TY / RPT);
  const dim3 grid(P.ntx * P.nty * P.ntz);
  switch (fmt) {
    case FMT_F32: k_step_ring<TY, RPT, FMT_F32, ACT><<<grid, block, 0, st>>>(P); break;
    case FMT_BF16: k_step_ring<TY, RPT, FMT_BF16, ACT><<<grid, block, 0, st>>>(P); break;
    case FMT_F16: k_step_ring<TY, RPT, FMT_F16, ACT><<<grid, block, 0, st>>>(P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int TY, int RPT>
hipError_t launch_ring_t(int act, int fmt, const StepParams &P, hipStream_t st) {
  if (act == ACT_NONE) {
    if (fmt != FMT_NONE) return hipErrorInvalidValue;
    k_step_ring<TY, RPT, FMT_NONE, ACT_NONE><<<dim3(P.ntx * P.nty * P.ntz), dim3(64, TY / RPT),
                                               0, st>>>(P);
    return hipGetLastError();
  }
  if (act == ACT_BORN) return launch_ring_a<TY, RPT, ACT_BORN>(fmt, P, st);
  if (act == ACT_IMG) return launch_ring_a<TY, RPT, ACT_IMG>(fmt, P, st);
  return hipErrorInvalidValue;
}

// The ring step with its receiver mode, lookahead and load policy.
template <int TY, int RPT, int FMT, int ACT, bool STREAM>
hipError_t launch_x4r(const StepParams &P, hipStream_t st, int rec, dim3 grid, dim3 block) {
  if constexpr (ACT == ACT_NONE) {
    if (P.pf == 2 && rec == REC_NONE) {
      k_step_x4r<TY, RPT, FMT, ACT, REC_NONE, 2, STREAM><<<grid, block, 0, st>>>(P);
      return hipGetLastError();
    }
    if (P.pf == 2 && rec == REC_INJECT) {
      k_step_x4r<TY, RPT, FMT, ACT, REC_INJECT, 2, STREAM><<<grid, block, 0, st>>>(P);
      return hipGetLastError();
    }
  }
  if (rec == REC_NONE) {
    k_step_x4r<TY, RPT, FMT, ACT, REC_NONE, 1, STREAM><<<grid, block, 0, st>>>(P);
  } else if (rec == REC_INJECT) {
    if constexpr (ACT != ACT_BORN)
      k_step_x4r<TY, RPT, FMT, ACT, REC_INJECT, 1, STREAM><<<grid, block, 0, st>>>(P);
    else
      return hipErrorInvalidValue;
  } else if (rec == REC_EXTRACT) {
    if constexpr (ACT != ACT_IMG)
      k_step_x4r<TY, RPT, FMT, ACT, REC_EXTRACT, 1, STREAM><<<grid, block, 0, st>>>(P);
    else
      return hipErrorInvalidValue;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool RING, int TY, int RPT, int FMT, int ACT>
hipError_t launch_x4_kernel(const StepParams &P, hipStream_t st, int rec) {
  const dim3 block(64, TY / RPT);
  const dim3 grid(P.ntx * P.nty * P.ntz);
  if constexpr (RING) {
    return P.stream ? launch_x4r<TY, RPT, FMT, ACT, true>(P, st, rec, grid, block)
                    : launch_x4r<TY, RPT, FMT, ACT, false>(P, st, rec, grid, block);
  } else {
    if (rec != REC_NONE) return hipErrorInvalidValue;
    k_step_x4<TY, RPT, FMT, ACT><<<grid, block, 0, st>>>(P);
  }
  return hipGetLastError();
}

template <bool RING, int TY, int RPT, int ACT>
hipError_t launch_x4_a(int fmt, const StepParams &P, hipStream_t st, int rec) {
  switch (fmt) {
    case FMT_F32: return launch_x4_kernel<RING, TY, RPT, FMT_F32, ACT>(P, st, rec);
    case FMT_BF16: return launch_x4_kernel<RING, TY, RPT, FMT_BF16, ACT>(P, st, rec);
    case FMT_F16: return launch_x4_kernel<RING, TY, RPT, FMT_F16, ACT>(P, st, rec);
    default: return hipErrorInvalidValue;
  }
}

template <bool RING, int TY, int RPT>
hipError_t launch_x4_t(int act, int fmt, const StepParams &P, hipStream_t st, int rec) {
  if (act == ACT_NONE) {
    if (fmt != FMT_NONE) return hipErrorInvalidValue;
    return launch_x4_kernel<RING, TY, RPT, FMT_NONE, ACT_NONE>(P, st, rec);
  }
  if (act == ACT_BORN) return launch_x4_a<RING, TY, RPT, ACT_BORN>(fmt, P, st, rec);
  if (act == ACT_IMG) return launch_x4_a<RING, TY, RPT, ACT_IMG>(fmt, P, st, rec);
  return hipErrorInvalidValue;
}

template <int FORM, int TY, int RPT, int ACT, int REC = REC_NONE>
hipError_t launch_step_a(int fmt, const StepParams &P, hipStream_t st) {
  const dim3 block(64, TY / RPT);
  const dim3 grid(P.ntx * P.ntz * P.nty);
  switch (fmt) {
    case FMT_F32: k_step<FORM, TY, RPT, FMT_F32, ACT, REC><<<grid, block, 0, st>>>(P); break;
    case FMT_BF16: k_step<FORM, TY, RPT, FMT_BF16, ACT, REC><<<grid, block, 0, st>>>(P); break;
    case FMT_F16: k_step<FORM, TY, RPT, FMT_F16, ACT, REC><<<grid, block, 0, st>>>(P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int FORM, int TY, int RPT>
hipError_t launch_step_t(int act, int fmt, const StepParams &P, hipStream_t st) {
  if (act == ACT_NONE) {
    if (fmt != FMT_NONE) return hipErrorInvalidValue;
    k_step<FORM, TY, RPT, FMT_NONE, ACT_NONE><<<dim3(P.ntx * P.nty * P.ntz), dim3(64, TY / RPT),
                                                0, st>>>(P);
    return hipGetLastError();
  }
  if (act == ACT_BORN) return launch_step_a<FORM, TY, RPT, ACT_BORN>(fmt, P, st);
  if constexpr (FORM == FORM_VL) {
    if (act == ACT_IMG) return launch_step_a<FORM, TY, RPT, ACT_IMG>(fmt, P, st);
  }
  return hipErrorInvalidValue;
}

// The small-grid shape (VL, 8-row tiles) with the receivers or the source
// points of one plane inside the step.
template <int REC>
hipError_t launch_step_rec_t(int act, int fmt, const StepParams &P, hipStream_t st) {
  if (act == ACT_NONE) {
    if (fmt != FMT_NONE) return hipErrorInvalidValue;
    k_step<FORM_VL, 8, 2, FMT_NONE, ACT_NONE, REC><<<dim3(P.ntx * P.nty * P.ntz), dim3(64, 4), 0,
                                                    st>>>(P);
    return hipGetLastError();
  }
  if constexpr (REC == REC_EXTRACT) {
    if (act == ACT_BORN) return launch_step_a<FORM_VL, 8, 2, ACT_BORN, REC>(fmt, P, st);
  } else {
    if (act == ACT_IMG) return launch_step_a<FORM_VL, 8, 2, ACT_IMG, REC>(fmt, P, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_step_x2(int act, int fmt, StepParams P, int ty_tile, hipStream_t st) {
  if (P.nw > kWMax || P.n3 < 9 || P.n2 < 9 || P.n1 < 10 || (P.n1 & 1)) return hipErrorInvalidValue;
  if (ty_tile == 32) return launch_x2_t<32, 4>(act, fmt, P, st);
  if (ty_tile == 8) return launch_x2_t<8, 2>(act, fmt, P, st);
  return launch_x2_t<16, 2>(act, fmt, P, st);
}

hipError_t launch_step_ring(int act, int fmt, StepParams P, int ty_tile, hipStream_t st) {
  if (P.nw > kWMax || P.n3 < 9 || P.n2 < 9 || P.n1 < 9) return hipErrorInvalidValue;
  if (ty_tile == 32) return launch_ring_t<32, 4>(act, fmt, P, st);
  if (ty_tile == 8) return launch_ring_t<8, 2>(act, fmt, P, st);
  return launch_ring_t<16, 4>(act, fmt, P, st);
}

hipError_t launch_step_x4(int act, int fmt, StepParams P, int ty_tile, bool ring,
                          hipStream_t st, int rec) {
  if (P.nw > kWMax || P.n3 < 9 || P.n2 < 9 || P.n1 < 12 || (P.n1 & 3)) return hipErrorInvalidValue;
  if (rec != REC_NONE && (!P.rmap || !P.rtr || P.nrec <= 0 || P.rz < 4 || P.rz >= P.n3 - 4))
    return hipErrorInvalidValue;
  // 8-row tiles with one row per thread, or 16-row tiles with two (ty_tile
  // 16 or 32); four cells per lane leave no registers for taller tiles.
  if (ring) {
    if (ty_tile == 8) return launch_x4_t<true, 8, 1>(act, fmt, P, st, rec);
    return launch_x4_t<true, 16, 2>(act, fmt, P, st, rec);
  }
  if (ty_tile == 8) return launch_x4_t<false, 8, 1>(act, fmt, P, st, rec);
  return launch_x4_t<false, 16, 2>(act, fmt, P, st, rec);
}

template <int REC>
hipError_t launch_step2_r(const StepParams &P, hipStream_t st) {
  const dim3 grid(P.ntx * P.nty * P.ntz), block(512);
  if (P.stream) k_step2_x4<REC, true><<<grid, block, 0, st>>>(P);
  else k_step2_x4<REC, false><<<grid, block, 0, st>>>(P);
  return hipGetLastError();
}

hipError_t launch_step2(StepParams P, int target_blocks, hipStream_t st, int rec) {
  if (P.nw > kWMax || P.n3 < 9 || P.n2 < 9 || P.n1 < 12 || (P.n1 & 3)) return hipErrorInvalidValue;
  if ((long)P.n1 * P.n2 > 0x7fffffffL) return hipErrorInvalidValue;  // in-plane int offsets
  if (!P.p || !P.po || !P.pn || !P.pn2 || !P.V || P.pn == P.p || P.pn == P.po ||
      P.pn2 == P.p || P.pn2 == P.po || P.pn == P.pn2)
    return hipErrorInvalidValue;
  if (rec != REC_NONE &&
      (rec != REC_INJECT || !P.rmap || !P.rtr || P.nrec <= 0 || P.rz < 4 || P.rz >= P.n3 - 4))
    return hipErrorInvalidValue;
  plan_step(P, 32, target_blocks, 64);
  if (rec == REC_INJECT) return launch_step2_r<REC_INJECT>(P, st);
  return launch_step2_r<REC_NONE>(P, st);
}

void plan_step(StepParams &P, int ty_tile, int target_blocks, int xw) {
  P.ntx = (P.n1 + xw - 1) / xw;
  P.nty = (P.n2 + ty_tile - 1) / ty_tile;
  const int nz = P.n3 - 8;
  const long tiles = (long)P.ntx * P.nty;
  long zc = (tiles * nz + target_blocks - 1) / target_blocks;
  if (zc < 8) zc = 8;
  if (zc > nz) zc = nz;
  P.zc = (int)zc;
  P.ntz = (nz + P.zc - 1) / P.zc;
}

hipError_t launch_step(int form, int act, int fmt, StepParams P, int ty_tile, hipStream_t st,
                       int rec) {
  if (P.nw > kWMax || P.n3 < 9 || P.n2 < 9 || P.n1 < 9) return hipErrorInvalidValue;
  if (rec != REC_NONE) {
    if (form != FORM_VL || ty_tile != 8 || !P.rmap || !P.rtr || P.nrec <= 0 || P.rz < 4 ||
        P.rz >= P.n3 - 4)
      return hipErrorInvalidValue;
    if (rec == REC_INJECT) return launch_step_rec_t<REC_INJECT>(act, fmt, P, st);
    if (rec == REC_EXTRACT) return launch_step_rec_t<REC_EXTRACT>(act, fmt, P, st);
    return hipErrorInvalidValue;
  }
  if (form == FORM_VL) {
    if (ty_tile == 32) return launch_step_t<FORM_VL, 32, 4>(act, fmt, P, st);
    if (ty_tile == 8) return launch_step_t<FORM_VL, 8, 2>(act, fmt, P, st);
    return launch_step_t<FORM_VL, 16, 4>(act, fmt, P, st);
  }
  if (ty_tile == 32) return launch_step_t<FORM_LV, 32, 4>(act, fmt, P, st);
  if (ty_tile == 8) return launch_step_t<FORM_LV, 8, 2>(act, fmt, P, st);
  return launch_step_t<FORM_LV, 16, 4>(act, fmt, P, st);
}

hipError_t launch_inject(const InjectParams &P, hipStream_t st) {
  if (P.npts <= 0) return hipSuccess;
  k_inject<<<(P.npts + 255) / 256, 256, 0, st>>>(P);
  return hipGetLastError();
}

hipError_t launch_extract(int npts, const int *locs, float *tr, long st_t, int id,
                          const float s[8], const float *p, const float *rscale,
                          hipStream_t st) {
  if (npts <= 0) return hipSuccess;
  k_extract<<<(npts + 255) / 256, 256, 0, st>>>(npts, locs, tr, st_t, id, s[0], s[1], s[2], s[3],
                                                s[4], s[5], s[6], s[7], p, rscale);
  return hipGetLastError();
}

hipError_t launch_vw_scale(const float *V, const float *wtab, int nw, const float *in, float *out,
                           int n1, int n2, int n3, hipStream_t st) {
  k_vw_scale<<<grid_for((long)n1 * n2 * n3, 256), 256, 0, st>>>(V, wtab, nw, in, out, n1, n2, n3);
  return hipGetLastError();
}

hipError_t launch_receiver_scale(const float *V, const float *wtab, int nw, const int *locs,
                                 int npts, float *rs, int n1, int n2, int n3, hipStream_t st) {
  if (npts <= 0) return hipSuccess;
  k_receiver_scale<<<(npts + 255) / 256, 256, 0, st>>>(V, wtab, nw, locs, npts, rs, n1, n2, n3);
  return hipGetLastError();
}

hipError_t launch_encode(int fmt, const float *s, void *dst, float *sscale, int n12, int n3,
                         hipStream_t st) {
  // the per-slice maxima start at zero (a fill kernel, not a memset: the
  // propagation loops that call this may be captured into a graph)
  k_fill<<<grid_for(n3, 256), 256, 0, st>>>(sscale, 0.f, n3);
  // absmax: 16 values per lane per pass
  int ga = (n12 + 4095) / 4096;
  if (ga > 64) ga = 64;
  k_slice_absmax<<<dim3(ga, n3), 256, 0, st>>>(s, reinterpret_cast<unsigned int *>(sscale), n12);
  int gx = (n12 + 255) / 256;
  if (gx > 64) gx = 64;
  if (fmt == FMT_BF16)
    k_encode<FMT_BF16><<<dim3(gx, n3), 256, 0, st>>>(s, dst, sscale, n12);
  else if (fmt == FMT_F16)
    k_encode<FMT_F16><<<dim3(gx, n3), 256, 0, st>>>(s, dst, sscale, n12);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_decode(int fmt, const void *src, const float *sscale, float *dst, int n12,
                         int n3, hipStream_t st) {
  int gx = (n12 + 255) / 256;
  if (gx > 64) gx = 64;
  k_decode<<<dim3(gx, n3), 256, 0, st>>>(fmt, src, sscale, dst, n12);
  return hipGetLastError();
}

hipError_t launch_square_accumulate(int fmt, const void *src, const float *sscale, float *acc,
                                   int n12, int n3, hipStream_t st) {
  int gx = (n12 + 255) / 256;
  if (gx > 64) gx = 64;
  k_square_acc<<<dim3(gx, n3), 256, 0, st>>>(fmt, src, sscale, acc, n12);
  return hipGetLastError();
}

hipError_t launch_dot(const float *a, const float *b, long n, double *partials, double *out,
                      hipStream_t st) {
  k_dot_partial<<<kDotBlocks, 256, 0, st>>>(a, b, n, partials);
  k_dot_final<<<1, 256, 0, st>>>(partials, kDotBlocks, out);
  return hipGetLastError();
}

hipError_t launch_axpby(float *y, const float *x, double a, double b, long n, hipStream_t st) {
  k_axpby<<<grid_for(n, 256), 256, 0, st>>>(y, x, a, b, n);
  return hipGetLastError();
}

hipError_t launch_mul(float *y, const float *x, long n, hipStream_t st) {
  k_mul<<<grid_for(n, 256), 256, 0, st>>>(y, x, n);
  return hipGetLastError();
}

hipError_t launch_zero_halo(float *p, long stride, int nfields, int n1, int n2, int n3,
                            hipStream_t st) {
  if (nfields <= 0) return hipSuccess;
  if (n1 < 9 || n2 < 9 || n3 < 9 || n3 > 65535 || nfields > 65535) return hipErrorInvalidValue;
  const int gx = (8 * n1 + 8 * n2 + 255) / 256;
  k_zero_halo<<<dim3(gx, n3, nfields), 256, 0, st>>>(p, stride, n1, n2, n3);
  return hipGetLastError();
}

hipError_t launch_fill(float *y, float v, long n, hipStream_t st) {
  k_fill<<<grid_for(n, 256), 256, 0, st>>>(y, v, n);
  return hipGetLastError();
}

hipError_t launch_dz(int adj, const float *in, float *out, int n12, int n3, int add,
                     hipStream_t st) {
  k_dz<<<grid_for((long)n12 * n3, 256), 256, 0, st>>>(adj, in, out, n12, n3, add);
  return hipGetLastError();
}

hipError_t launch_embed(int adj, const float *model, float *comp, int m1, int m2, int m3, int n1,
                        int n2, int n3, int o1, int o2, int o3, hipStream_t st) {
  k_embed<<<grid_for((long)n1 * n2 * n3, 256), 256, 0, st>>>(adj, model, comp, m1, m2, m3, n1,
                                                             n2, n3, o1, o2, o3);
  return hipGetLastError();
}

hipError_t launch_trace_to_data(const float *tr, float *data, int nrec, int ntb, int nt,
                                float inv_dt, int add, hipStream_t st) {
  dim3 grid((nrec + 31) / 32, (nt + 31) / 32);
  k_trace_to_data<<<grid, dim3(32, 8), 0, st>>>(tr, data, nrec, ntb, nt, inv_dt, add);
  return hipGetLastError();
}

hipError_t launch_data_to_trace(const float *data, float *tr, int nrec, int ntb, int nt,
                                float inv_dt, hipStream_t st) {
  dim3 grid((nrec + 31) / 32, (ntb + 31) / 32);
  k_data_to_trace<<<grid, dim3(32, 8), 0, st>>>(data, tr, nrec, ntb, nt, inv_dt);
  return hipGetLastError();
}

hipError_t launch_copy16(const void *src, void *dst, long n16, hipStream_t st) {
  const long blocks = (n16 + 255) / 256;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  k_copy16<<<dim3((unsigned)blocks), 256, 0, st>>>(static_cast<const uint4 *>(src),
                                                   static_cast<uint4 *>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_sum16(int nin, bool nt_store, const void *a, const void *b, const void *c,
                        void *out, long n16, hipStream_t st) {
  const long blocks = (n16 + 255) / 256;
  if (blocks > 0x7fffffffL || nin < 1 || nin > 3) return hipErrorInvalidValue;
  const dim3 g((unsigned)blocks), t(256);
  const f32x4 *A = static_cast<const f32x4 *>(a), *B = static_cast<const f32x4 *>(b),
              *C = static_cast<const f32x4 *>(c);
  f32x4 *O = static_cast<f32x4 *>(out);
#define SUM16(N)                                                      \
  if (nin == N) {                                                     \
    if (nt_store) k_sum16<N, true><<<g, t, 0, st>>>(A, B, C, O, n16);  \
    else k_sum16<N, false><<<g, t, 0, st>>>(A, B, C, O, n16);          \
  }
  SUM16(1) SUM16(2) SUM16(3)
#undef SUM16
  return hipGetLastError();
}

hipError_t launch_transpose(const float *in, float *out, long rows, long cols, hipStream_t st) {
  dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
  k_transpose<<<grid, dim3(32, 8), 0, st>>>(in, out, rows, cols);
  return hipGetLastError();
}

}  // namespace brtm
