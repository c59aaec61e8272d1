/*
 * born_op.h -- operator and solver layer of the MI355X Born RTM library.
 *
 * Replaces the reference's C++ inversion layer with HBM-resident vectors:
 *   brtm_vec  -- my_vector / hypercube_float (Born/my_vector.h:10-41,
 *                Born/hypercube_float.cpp): FP32 storage, FP64 dot products
 *                (Born/hypercube_float.cpp:123-132) reduced on the GPU;
 *   brtm_op   -- rtm_zero_op (Born/rtm_zero_op_3d.h:9-38,
 *                Born/rtm_zero_op_3d.cpp): the Born operator over all shots,
 *                forward(add, model, data) / adjoint(add, model, data);
 *   brtm_cg_solve -- simple_solver + cgstep (Born/lin_solver.cpp:7-17,
 *                79-101,115-143; Born/cgstep.cpp:14-58) with a relative
 *                data-residual stop and an optional preconditioner applied
 *                to the gradient (flexible: cgstep's plane search keeps the
 *                residual non-increasing for any direction).
 *
 * Vectors: the model is a volume on the model grid [m3][m2][m1]; the data
 * vector is [nshots][nry][nrx][nt] (time fastest), the layout of the
 * reference's data_rtm_3d axes (t, rx, ry, shot).  The operator is
 *   L m = D_t W B(V) D_z E (mute . m)   per shot, summed into the data,
 * E the embedding of the model grid into the compute grid at the model
 * origin (image_rtm_3d::extract_sub, Born/image_rtm_3d.cpp:52-91), D_z the
 * depth difference (Born/laplac_3d.cpp:46-65), B the Born modelling of
 * brtm_rtm_forward, W the window dropping 4 pad samples
 * (Born/map_data_3d.cpp:64-76) and D_t the time difference
 * (Born/deriv_3d.cpp:25-46, within each trace).  adjoint is its exact
 * transpose.
 */
#ifndef BORN_OP_H
#define BORN_OP_H

#include "born_rtm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct brtm_vec brtm_vec;
typedef struct brtm_op brtm_op;

/* ---- vectors (my_vector) ---------------------------------------------- */
int brtm_vec_create(brtm_ctx *ctx, long n, brtm_vec **out);
int brtm_vec_destroy(brtm_vec *v);
long brtm_vec_size(const brtm_vec *v);
/* Device address of the vector's FP32 storage (for kernels of the caller;
 * hypercube_float::vals, Born/hypercube_float.h). */
void *brtm_vec_data(brtm_vec *v);
/* Host <-> device copies of the values (no counterpart: the reference's
 * vectors live in host memory). */
int brtm_vec_upload(brtm_vec *v, const float *host);
int brtm_vec_download(const brtm_vec *v, float *host);
int brtm_vec_zero(brtm_vec *v);                                 /* set_val(0)   */
int brtm_vec_copy(brtm_vec *dst, const brtm_vec *src);           /* clone_vec    */
int brtm_vec_scale(brtm_vec *v, double s);                       /* scale        */
/* v = a v + b x  (my_vector::scale_add, Born/hypercube_float.cpp:212-218) */
int brtm_vec_scale_add(brtm_vec *v, double a, const brtm_vec *x, double b);
int brtm_vec_add(brtm_vec *v, const brtm_vec *x);                /* add          */
int brtm_vec_mult(brtm_vec *v, const brtm_vec *x);               /* mult         */
/* FP64 inner product, deterministic (fixed-grid reduction): my_vector::dot
 * (Born/hypercube_float.cpp:123-132). */
int brtm_vec_dot(const brtm_vec *a, const brtm_vec *b, double *out);
/* Uniform values in [-0.5, 0.5) from SplitMix64(seed)
 * (my_vector::random, Born/hypercube_float.cpp:118-122, with a seeded
 * generator instead of rand()). */
int brtm_vec_random(brtm_vec *v, unsigned long long seed);

/* ---- the Born operator (rtm_zero_op) ---------------------------------- */
typedef struct {
  int m1, m2, m3;           /* model (image) grid                               */
  int o1, o2, o3;           /* model origin inside the compute grid             */
  int nshots;
  int nsrc;                 /* source points per shot (9 in the reference)      */
  const int *src_locs;      /* [nshots][nsrc] compute-grid cells                */
  const float *src_traces;  /* [nshots][nsrc][ntr_s]                            */
  int ntr_s;
  int nrx, nry;             /* receivers per shot                               */
  const int *rec_locs;      /* [nshots][nry*nrx] compute-grid cells             */
  int nt_data;              /* samples per data trace                           */
  float dt_data;            /* data sampling (D_t divides by it)                */
  int jt;                   /* FD steps per data sample (nt = (nt_data-1)jt+1)  */
  int jimg;                 /* imaging / Born-source cadence in FD steps        */
  int store;                /* BRTM_STORE_* for the source wavefields           */
  int ckpt;                 /* checkpoint interval for BRTM_STORE_CKPT          */
  const float *mute;        /* [m3][m2][m1] multiplier, or NULL                 */
} brtm_op_desc;

/* Replaces rtm_zero_op::rtm_zero_op (Born/rtm_zero_op_3d.cpp:8-37): copies
 * the geometry; the context must already hold the space and velocity. */
int brtm_op_create(brtm_ctx *ctx, const brtm_op_desc *desc, brtm_op **out);
/* Replaces ~rtm_zero_op (Born/rtm_zero_op_3d.h:20): releases the per-shot
 * stores this operator's brtm_op_prepare made (not those another operator
 * on the context has prepared since). */
int brtm_op_destroy(brtm_op *op);
/* Replaces create_source_fields (Born/rtm_zero_op_3d.cpp:38-85): propagate
 * and store every shot's source wavefield.  The stores are the context's,
 * one per shot index: after another operator on the same context prepares
 * a shot, this operator's forward, adjoint and illumination return
 * BRTM_ESTATE until it prepares again. */
int brtm_op_prepare(brtm_op *op);
long brtm_op_model_size(const brtm_op *op);
long brtm_op_data_size(const brtm_op *op);
/* rtm_zero_op::forward / adjoint (Born/rtm_zero_op_3d.cpp:87-293);
 * add = 0 zeroes the output first. */
int brtm_op_forward(brtm_op *op, int add, const brtm_vec *model, brtm_vec *data);
int brtm_op_adjoint(brtm_op *op, int add, brtm_vec *model, const brtm_vec *data);
/* my_operator::dot_test (Born/my_operator.cpp:14-76) with seeded random
 * vectors: err[0] = |<Lm,d>-<m,L'd>|/|<Lm,d>| for add = false, err[1] for
 * add = true.  Returns BRTM_OK; the caller compares against 1e-4. */
int brtm_op_dot_test(brtm_op *op, unsigned long long seed, double err[2]);
/* Source illumination on the model grid: out = mute . E^T sum_shots sum_k
 * s_k^2 over the stored source snapshots (snapshot stores only; no
 * counterpart in the reference, which has no preconditioner).  Used to
 * balance image amplitudes for preconditioning. */
int brtm_op_illumination(brtm_op *op, brtm_vec *out);

/* ---- CG (simple_solver + cgstep) -------------------------------------- */
/* Preconditioner: z = P(g) on device model vectors; return 0 on success. */
typedef int (*brtm_precond_fn)(void *user, const brtm_vec *g, brtm_vec *z);

typedef struct {
  int niter;                /* maximum iterations (lin_solver::solve(niter))    */
  double rel_tol;           /* stop when the judged residual <= rel_tol (0: never) */
  brtm_precond_fn precond;  /* NULL: plain cgstep                               */
  void *precond_user;
  int verbose;
  /* Optional data-space offset e (NULL: none).  The judged residual is then
   * |rr + e| / |d - e| instead of |rr| / |d|: with e the additive noise of
   * synthetic data d = L m_true + e, it is the relative misfit to the
   * noise-free data (|L m - d_clean| / |d_clean|).  The iteration itself
   * still minimises |L m - d|. */
  const brtm_vec *offset;
  /* Steps kept for the search (0 or 1: cgstep's plane search over the new
   * direction and the previous step, Born/cgstep.cpp:28-53).  h >= 2
   * (at most BRTM_CG_MAX_HISTORY): flexible CG with truncated history --
   * each step minimises |rr + L(a z + sum_i b_i s_i)| over the new
   * direction z and the last h steps s_i, an (h+1)x(h+1) FP64 Gram solve.
   * With a fixed SPD preconditioner this is CG again; with a preconditioner
   * that changes between iterations (a network) it keeps the search
   * directions L-conjugate to the last h steps. */
  int history;
  /* Optional (NULL: none), niter+1 entries: |rr| / |d| before the first and
   * after every iteration, the relative residual of the data as given.
   * Without an offset it equals the judged residual; with one it is the
   * residual of the noisy data, |L m - d| / |d|. */
  double *data_resid_hist;
} brtm_cg_params;

#define BRTM_CG_MAX_HISTORY 16

typedef struct {
  int iterations;           /* steps taken                                      */
  int fallbacks;            /* iterations where P(g) was not a descent direction
                               and g was used instead                           */
  int converged;            /* rel residual reached rel_tol                     */
  int breakdown;            /* cgstep breakdown (gg.gg or ss.ss < 1e-20)        */
  double seconds;           /* wall-clock of the iteration loop                 */
  double rel_resid;         /* final judged residual (|rr|/|d| without offset)   */
  double data_rel_resid;    /* final |rr|/|d|                                   */
} brtm_cg_stats;

/* Solve min |L m - d| from m = 0 (Born/lin_solver.cpp:79-101, rr = -d).
 * model receives m.  resid_hist (nullable, niter+1 entries) receives
 * |rr|/|d| before the first and after every iteration; time_hist (nullable,
 * niter+1 entries) the elapsed seconds at the same points. */
int brtm_cg_solve(brtm_op *op, const brtm_vec *data, brtm_vec *model,
                  const brtm_cg_params *params, double *resid_hist, double *time_hist,
                  brtm_cg_stats *stats);

/* Hale's 8-tap least-squares sinc tables (Born/sinc_bob.cpp:4-114) in FP64,
 * rounded to FP32: table[i][k] for fractional shift i/ntab. */
int brtm_sinc_table(int ntab, int nsinc, float *table);

#ifdef __cplusplus
}
#endif
#endif /* BORN_OP_H */
