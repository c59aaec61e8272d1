/*
 * born_oracle.c -- CPU restatement of the Born acoustic RTM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the libraries built from it.  The product path (sep_code_amd) never
 * links or calls it and fails loudly when its HIP library is missing.
 *
 * It is compiled twice by oracle/Makefile:
 *   ORACLE_REAL=double, suffix _f64 : the FP64 oracle every GPU result is
 *                                     compared against;
 *   ORACLE_REAL=float + OpenMP, _f32: the FP32 CPU restatement that bench.py
 *                                     times as the CPU baseline.
 *
 * Layout: float arrays [n3][n2][n1], n1 fastest; stencil radius 4.  The
 * update set U is [4, n-4) on every axis; cells outside U are never written
 * and stay zero (the padding the reference offsets its pointers past,
 * Born/gpu_funcs_3d.cu:781).
 *
 * Time recursion (one shot), fields start at zero:
 *   VL step   p' = w.(V.L p + 2p) - w^2.p_old          (wave_kernel form,
 *             Born/wave_fkernel.3d8o.cu:86-103, damping of :183-210 moved to
 *             the output so the step needs w only at the centre cell)
 *   LV step   p' = L(V.w.p) + 2w.p - w^2.p_old          (wave_kernel_adj form,
 *             Born/wave_fkernel.3d8o.cu:164-178, damping on the input)
 * The LV step is the exact transpose of the VL step, so Born modelling
 * (scattered field in LV form, as Born/gpu_funcs_3d.cu:423) and RTM
 * (receiver field in VL form, as Born/gpu_funcs_3d.cu:626) form an exact
 * adjoint pair when both read the same source snapshots.
 *
 * Parity pins (tests/test_oracle.py): coefficients against the closed-form
 * Fornberg weights and polynomial exactness; the LV step against the
 * explicit transpose of the VL step (dense matrices on tiny grids); the
 * undamped leapfrog against exact time reversal; inject/extract and
 * forward/adjoint against dot-product tests; a 1-D plane wave against the
 * discrete dispersion relation; the FP32 build against the FP64 build; and
 * the time indexing of source propagation, Born modelling and the adjoint
 * (sample it/jt, sinc row it%jt, step before injection, the +4 trace offset,
 * the Born source from s^{it+1}) against the closed-form solution of the
 * V = 0 undamped recursion, p^n = sum_{m<n} (n-m) f(m).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
#ifndef ORACLE_SUFFIX
#define ORACLE_SUFFIX _f64
#endif

typedef ORACLE_REAL real;
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, ORACLE_SUFFIX)

#define RAD 4

/* One shot of the Born problem on the padded compute grid.  Pointers are
 * host arrays owned by the caller; every `real` array uses the precision
 * this library was built with. */
typedef struct {
  int n1, n2, n3;
  int nw;               /* sponge table length (weights for e < nw)        */
  const real *c;        /* 13 stencil weights, Born/wave_fkernel.3d8o.cu:27 */
  const real *wtab;     /* sponge weights w(e), e = 0..nw-1                 */
  const real *V;        /* v^2 dt^2 on the compute grid                     */
  int nsrc;             /* source points (9 in the reference)               */
  const int *slocs;     /* linear cell index of each source point           */
  const real *strc;     /* source traces [nsrc][ntr_s], time fastest        */
  int ntr_s;
  const real *sinc_s;   /* [jt][8] source interpolation table               */
  int nrec;
  const int *rlocs;     /* receiver cell index, <0 = inactive receiver      */
  int ntr_d;            /* receiver trace length (nt_data + 8)              */
  const real *sinc_d;   /* [jt][8] receiver interpolation table             */
  int jt;               /* FD steps per data sample                         */
  int nt;               /* FD steps                                         */
  int jimg;             /* imaging / Born-source cadence in FD steps        */
  double cimg;          /* imaging scale (1e-6 in the reference)            */
} orc_prob;

/* ---- A1: stencil weights.  Born/wave_fkernel.3d8o.cu:26-27, evaluated in
 * double as Born/gpu_funcs_3d.cu:783 does.  Order: c0, then for k=1..4 the
 * three axis weights (axis1, axis2, axis3). */
void FN(orc_coeffs)(double d1, double d2, double d3, real *c) {
  double C1 = 8.0 / (5.0 * d1 * d1), C2 = 8.0 / (5.0 * d2 * d2),
         C3 = 8.0 / (5.0 * d3 * d3);
  static const double f[4] = {1.0, -1.0 / 8.0, 1.0 / 63.0, -1.0 / 896.0};
  c[0] = (real)(-1025.0 / 576.0 * (C1 + C2 + C3));
  for (int k = 0; k < 4; k++) {
    c[1 + 3 * k] = (real)(C1 * f[k]);
    c[2 + 3 * k] = (real)(C2 * f[k]);
    c[3 + 3 * k] = (real)(C3 * f[k]);
  }
}

/* ---- A4: sponge weights.  Born/wave_fkernel.3d8o.cu:194,201-202:
 * w = exp(-b (a - e)) where that is < 1, else 1; e is the distance (in
 * cells) to the nearest face minus 4.  Returned for e = 0..nw-1 with
 * nw = ceil(a); w = 1 for e >= nw. */
int FN(orc_sponge_table)(double a, double b, int nwmax, real *w) {
  int nw = (int)ceil(a);
  if (nw < 0) nw = 0;
  if (nw > nwmax) return -1;
  for (int e = 0; e < nw; e++) {
    double v = exp(-b * (a - (double)e));
    w[e] = (real)(v < 1.0 ? v : 1.0);
  }
  return nw;
}

static inline int imin(int a, int b) { return a < b ? a : b; }

/* Edge distance of Born/wave_fkernel.3d8o.cu:194 (single-GPU branch). */
static inline int edge_dist(int i1, int i2, int i3, int n1, int n2, int n3) {
  int e = imin(i1 - RAD, n1 - RAD - i1);
  e = imin(e, imin(i2 - RAD, n2 - RAD - i2));
  e = imin(e, imin(i3 - RAD, n3 - RAD - i3));
  return e;
}

static inline real weight(const orc_prob *P, int e) {
  return (e >= 0 && e < P->nw) ? P->wtab[e] : (real)1;
}

/* 8th-order Laplacian at linear index i (A.2). */
static inline real lap8(const real *p, const real *c, long i, long s2, long s3) {
  real acc = c[0] * p[i];
  for (int k = 1; k <= 4; k++) {
    acc += c[1 + 3 * (k - 1)] * (p[i + k] + p[i - k]);
    acc += c[2 + 3 * (k - 1)] * (p[i + k * s2] + p[i - k * s2]);
    acc += c[3 + 3 * (k - 1)] * (p[i + k * s3] + p[i - k * s3]);
  }
  return acc;
}

/* ---- A2 + A4 fused: VL step, output damping.  pn may alias po. */
void FN(orc_step_vl)(const orc_prob *P, const real *p, const real *po, real *pn) {
  const int n1 = P->n1, n2 = P->n2, n3 = P->n3;
  const long s2 = n1, s3 = (long)n1 * n2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i3 = RAD; i3 < n3 - RAD; i3++)
    for (int i2 = RAD; i2 < n2 - RAD; i2++) {
      long base = i3 * s3 + i2 * s2;
      for (int i1 = RAD; i1 < n1 - RAD; i1++) {
        long i = base + i1;
        real w = weight(P, edge_dist(i1, i2, i3, n1, n2, n3));
        real l = lap8(p, P->c, i, s2, s3);
        pn[i] = w * (P->V[i] * l + 2 * p[i]) - w * w * po[i];
      }
    }
}

/* ---- A3 + A4 fused: LV step, input damping.  u is a workspace of n cells;
 * pn may alias po. */
void FN(orc_step_lv)(const orc_prob *P, const real *p, const real *po, real *pn,
                     real *u) {
  const int n1 = P->n1, n2 = P->n2, n3 = P->n3;
  const long s2 = n1, s3 = (long)n1 * n2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i3 = 0; i3 < n3; i3++)
    for (int i2 = 0; i2 < n2; i2++) {
      long base = i3 * s3 + i2 * s2;
      for (int i1 = 0; i1 < n1; i1++) {
        long i = base + i1;
        int e = edge_dist(i1, i2, i3, n1, n2, n3);
        u[i] = (e >= 0) ? P->V[i] * weight(P, e) * p[i] : (real)0;
      }
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i3 = RAD; i3 < n3 - RAD; i3++)
    for (int i2 = RAD; i2 < n2 - RAD; i2++) {
      long base = i3 * s3 + i2 * s2;
      for (int i1 = RAD; i1 < n1 - RAD; i1++) {
        long i = base + i1;
        real w = weight(P, edge_dist(i1, i2, i3, n1, n2, n3));
        real l = lap8(u, P->c, i, s2, s3);
        pn[i] = l + 2 * w * p[i] - w * w * po[i];
      }
    }
}

/* ---- A5/A7: p[loc_j] += scale * sum_k sinc8[k] tr[j][id+k].
 * Born/gpu_funcs_3d.cu:64-75 (source, scale = dir) and :77-93 (data). */
void FN(orc_inject)(int npts, const int *locs, const real *tr, int ntr,
                    const real *sinc8, int id, real scale, real *p) {
  for (int j = 0; j < npts; j++) {
    if (locs[j] < 0) continue;
    real acc = 0;
    for (int k = 0; k < 8; k++) acc += sinc8[k] * tr[(long)j * ntr + id + k];
    p[locs[j]] += scale * acc;
  }
}

/* ---- A6: tr[j][id+k] += sinc8[k] p[loc_j].  Born/gpu_funcs_3d.cu:95-109. */
void FN(orc_extract)(int npts, const int *locs, real *tr, int ntr,
                     const real *sinc8, int id, const real *p) {
  for (int j = 0; j < npts; j++) {
    if (locs[j] < 0) continue;
    real v = p[locs[j]];
    for (int k = 0; k < 8; k++) tr[(long)j * ntr + id + k] += v * sinc8[k];
  }
}

/* acc += c * a .* b over n cells: A9 imaging (Born/gpu_funcs_3d.cu:111-120)
 * and A8 Born secondary source (:122-131). */
void FN(orc_xcorr)(long n, real c, const real *a, const real *b, real *acc) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) acc[i] += c * a[i] * b[i];
}

static real *zalloc(long n) { return (real *)calloc((size_t)n, sizeof(real)); }

/* One source-field step: s' = VL step, then injection (dir = +1).
 * Born/gpu_funcs_3d.cu:249-250 with id/ii of :228-229. */
static void source_step(const orc_prob *P, int it, const real *cur, real *old_new) {
  int id = it / P->jt, ii = it % P->jt;
  FN(orc_step_vl)(P, cur, old_new, old_new);
  FN(orc_inject)(P->nsrc, P->slocs, P->strc, P->ntr_s, P->sinc_s + 8 * ii, id,
                 (real)1, old_new);
}

/* ---- A10: source propagation, Born/gpu_funcs_3d.cu:133-315.
 * Runs nt steps.  snaps (nullable) receives s^{it+1} for it % jimg == 0,
 * snapshot k = it / jimg; last2 (nullable) receives s^{nt-1}, s^{nt}. */
int FN(orc_source_prop)(const orc_prob *P, real *snaps, real *last2) {
  long n = (long)P->n1 * P->n2 * P->n3;
  real *a = zalloc(n), *b = zalloc(n);
  if (!a || !b) { free(a); free(b); return -1; }
  for (int it = 0; it < P->nt; it++) {
    source_step(P, it, b, a); /* a <- s^{it+1} */
    if (snaps && it % P->jimg == 0)
      memcpy(snaps + (long)(it / P->jimg) * n, a, n * sizeof(real));
    real *t = a; a = b; b = t;
  }
  if (last2) {
    memcpy(last2, a, n * sizeof(real));
    memcpy(last2 + n, b, n * sizeof(real));
  }
  free(a); free(b);
  return 0;
}

/* ---- A11: Born modelling, Born/gpu_funcs_3d.cu:317-505.
 * mt is the (already differentiated, muted) model on the compute grid.
 * snaps (nullable): source snapshots at imaging cadence; when NULL the
 * source field is propagated alongside.  D [nrec][ntr_d] is accumulated. */
int FN(orc_born_forward)(const orc_prob *P, const real *mt, const real *snaps,
                         real *D) {
  long n = (long)P->n1 * P->n2 * P->n3;
  real *qa = zalloc(n), *qb = zalloc(n), *u = zalloc(n);
  real *sa = snaps ? NULL : zalloc(n), *sb = snaps ? NULL : zalloc(n);
  if (!qa || !qb || !u || (!snaps && (!sa || !sb))) {
    free(qa); free(qb); free(u); free(sa); free(sb);
    return -1;
  }
  for (int it = 0; it < P->nt; it++) {
    int id = it / P->jt, ii = it % P->jt;
    const real *s = NULL;
    if (!snaps) { source_step(P, it, sb, sa); s = sa; }
    FN(orc_step_lv)(P, qb, qa, qa, u); /* qa <- q^{it+1} */
    if (it % P->jimg == 0) {
      if (snaps) s = snaps + (long)(it / P->jimg) * n;
      FN(orc_xcorr)(n, (real)P->cimg, s, mt, qa);
    }
    FN(orc_extract)(P->nrec, P->rlocs, D, P->ntr_d, P->sinc_d + 8 * ii, id, qa);
    real *t = qa; qa = qb; qb = t;
    if (!snaps) { t = sa; sa = sb; sb = t; }
  }
  free(qa); free(qb); free(u); free(sa); free(sb);
  return 0;
}

/* Adjoint steps it = k1-1 .. k0 over a window of source snapshots
 * seg[(it - k0)/..] (imaging steps only; seg_of(it) gives the pointer). */
static void adjoint_window(const orc_prob *P, int k0, int k1, const real *D,
                           real *const *segp, real **la, real **lb, real *img) {
  long n = (long)P->n1 * P->n2 * P->n3;
  for (int it = k1 - 1; it >= k0; it--) {
    int id = it / P->jt, ii = it % P->jt;
    FN(orc_step_vl)(P, *lb, *la, *la); /* la <- lambda^{it+1} */
    FN(orc_inject)(P->nrec, P->rlocs, D, P->ntr_d, P->sinc_d + 8 * ii, id,
                   (real)1, *la);
    if (it % P->jimg == 0) FN(orc_xcorr)(n, (real)P->cimg, segp[it - k0], *la, img);
    real *t = *la; *la = *lb; *lb = t;
  }
}

/* ---- A12: RTM adjoint (exact transpose of orc_born_forward),
 * Born/gpu_funcs_3d.cu:509-712 step order: step, data injection, imaging.
 * snaps non-NULL: uses them.  snaps NULL: checkpoints the source field every
 * ckpt steps and recomputes each segment.  img [n] is accumulated. */
int FN(orc_rtm_adjoint)(const orc_prob *P, const real *D, const real *snaps,
                        int ckpt, real *img) {
  long n = (long)P->n1 * P->n2 * P->n3;
  real *la = zalloc(n), *lb = zalloc(n);
  if (!la || !lb) { free(la); free(lb); return -1; }
  if (snaps) {
    real **segp = (real **)calloc((size_t)P->nt, sizeof(real *));
    for (int it = 0; it < P->nt; it++)
      if (it % P->jimg == 0) segp[it] = (real *)snaps + (long)(it / P->jimg) * n;
    adjoint_window(P, 0, P->nt, D, segp, &la, &lb, img);
    free(segp);
  } else {
    if (ckpt < 1) ckpt = 1;
    int nck = (P->nt + ckpt - 1) / ckpt;
    real *ck = zalloc(2 * n * (long)nck); /* (s^{k0}, s^{k0-1}) per segment */
    real *seg = zalloc(n * (long)ckpt);
    real **segp = (real **)calloc((size_t)ckpt, sizeof(real *));
    real *sa = zalloc(n), *sb = zalloc(n);
    if (!ck || !seg || !segp || !sa || !sb) {
      free(ck); free(seg); free(segp); free(sa); free(sb); free(la); free(lb);
      return -1;
    }
    for (int it = 0; it < P->nt; it++) {
      if (it % ckpt == 0) {
        long c = it / ckpt;
        memcpy(ck + 2 * c * n, sb, n * sizeof(real));     /* s^{it}   */
        memcpy(ck + (2 * c + 1) * n, sa, n * sizeof(real)); /* s^{it-1} */
      }
      source_step(P, it, sb, sa);
      real *t = sa; sa = sb; sb = t;
    }
    for (int c = nck - 1; c >= 0; c--) {
      int k0 = c * ckpt, k1 = k0 + ckpt < P->nt ? k0 + ckpt : P->nt;
      memcpy(sb, ck + 2L * c * n, n * sizeof(real));
      memcpy(sa, ck + (2L * c + 1) * n, n * sizeof(real));
      for (int it = k0; it < k1; it++) {
        source_step(P, it, sb, sa);
        memcpy(seg + (long)(it - k0) * n, sa, n * sizeof(real));
        segp[it - k0] = seg + (long)(it - k0) * n;
        real *t = sa; sa = sb; sb = t;
      }
      adjoint_window(P, k0, k1, D, segp, &la, &lb, img);
    }
    free(ck); free(seg); free(segp); free(sa); free(sb);
  }
  free(la); free(lb);
  return 0;
}

/* Step n_steps of the plain propagation (no source), for the CPU baseline
 * timing and the reversibility/energy pins.  form 0 = VL, 1 = LV. */
int FN(orc_propagate)(const orc_prob *P, int form, int n_steps, real *p_cur,
                      real *p_old) {
  long n = (long)P->n1 * P->n2 * P->n3;
  real *u = form ? zalloc(n) : NULL;
  if (form && !u) return -1;
  real *a = p_old, *b = p_cur;
  for (int s = 0; s < n_steps; s++) {
    if (form) FN(orc_step_lv)(P, b, a, a, u);
    else FN(orc_step_vl)(P, b, a, a);
    real *t = a; a = b; b = t;
  }
  /* leave the newest field in p_cur, the previous one in p_old */
  if (b != p_cur) {
    for (long i = 0; i < n; i++) { real t = p_cur[i]; p_cur[i] = p_old[i]; p_old[i] = t; }
  }
  free(u);
  return 0;
}

/* ---- A26: FP64 dot product, sequential as Born/hypercube_float.cpp:123-132. */
double FN(orc_dot)(long n, const real *a, const real *b) {
  double s = 0;
  for (long i = 0; i < n; i++) s += (double)a[i] * (double)b[i];
  return s;
}

/* Flush denormal results and operands to zero (x86 MXCSR FTZ and DAZ) on
 * the calling thread and on every thread of the OpenMP team, or restore
 * gradual underflow.  The tails ahead of a wavefront sweep through the
 * denormal range, where x86 arithmetic runs tens of times slower; values
 * below 1.2e-38 (FP32) change nothing a 1e-5 comparison can see.  Only the
 * timing of the FP32 build turns it on; parity runs leave it off. */
void FN(orc_flush_denormals)(int on) {
  const unsigned bits = 0x8040u; /* FTZ (bit 15) | DAZ (bit 6) */
#pragma omp parallel
  {
    unsigned c = _mm_getcsr();
    _mm_setcsr(on ? (c | bits) : (c & ~bits));
  }
  unsigned c = _mm_getcsr();
  _mm_setcsr(on ? (c | bits) : (c & ~bits));
}
