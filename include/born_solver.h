/*
 * born_solver.h -- composable linear operators and the reference's three
 * least-squares solvers on HBM-resident vectors.
 *
 * Replaces the reference's generic inversion layer:
 *   brtm_linop  -- my_operator (Born/my_operator.h:7-36) and its kinds:
 *                  rtm_zero_op, i_op (Born/i_op.cpp:4-28), tmute
 *                  (Born/tmute.cpp:4-76), row_op / col_op / diag_op
 *                  (Born/combo_oper.cpp:1-282), and caller-defined
 *                  operators through callbacks;
 *   brtm_solver -- simple_solver, reg_solver and prec_solver
 *                  (Born/lin_solver.cpp:7-53,79-101,115-155) driven by
 *                  cgstep (Born/cgstep.cpp:14-58) or flexible CG.
 *
 * Stacked vectors (the reference's super_vector, Born/super_vector.cpp)
 * are one contiguous brtm_vec holding the parts back to back, so a
 * dot product or an update over the whole stack is one kernel.
 *
 * Combinators borrow their parts: destroy a composite before its parts,
 * and the parts before the brtm_op / vectors they refer to.
 */
#ifndef BORN_SOLVER_H
#define BORN_SOLVER_H

#include "born_op.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct brtm_linop brtm_linop;

/* A caller-defined operator direction on device vectors: the forward
 * callback computes out = (add ? out : 0) + A in (in: domain, out: range),
 * the adjoint one out = (add ? out : 0) + A' in (in: range, out: domain).
 * Return 0 on success. */
typedef int (*brtm_apply_fn)(void *user, int add, const brtm_vec *in, brtm_vec *out);

/* rtm_zero_op as an operator (Born/rtm_zero_op_3d.h:9-38). */
int brtm_linop_born(brtm_op *op, brtm_linop **out);
/* i_op (Born/i_op.cpp:4-28): out = add*out + scale*in on n elements. */
int brtm_linop_identity(brtm_ctx *ctx, long n, brtm_linop **out);
/* Pointwise weight: out = add*out + scale*(w . in); self-adjoint.  The
 * weight vector is borrowed.  No class of its own in the reference, whose
 * data weight (wt) is any my_operator (Born/my_operator.h:7-36); this is
 * the diagonal one the preconditioned and weighted solves use. */
int brtm_linop_diagonal(const brtm_vec *w, brtm_linop **out);
/* tmute (Born/tmute.cpp:4-76): data [nouter][nrx][nt] (t fastest; nouter =
 * nry*nshots) keep sample it of trace rx when it >= (int)((t0 + (orx +
 * drx*rx)/vmute - ot)/dt) (the first kept sample clamped to 0) and zero it
 * otherwise; self-adjoint. */
int brtm_linop_tmute(brtm_ctx *ctx, int nt, float ot, float dt, int nrx, float orx, float drx,
                     long nouter, float t0, float vmute, brtm_linop **out);
int brtm_linop_callback(brtm_ctx *ctx, long ndomain, long nrange, brtm_apply_fn forward,
                        brtm_apply_fn adjoint, void *user, brtm_linop **out);
/* row_op (Born/combo_oper.cpp:176-235,257-282): ops[n-1] o ... o ops[0]; the
 * range of each must be the domain of the next. */
int brtm_linop_chain(int n, brtm_linop *const *ops, brtm_linop **out);
/* col_op (Born/combo_oper.cpp:1-99): one domain, ranges stacked:
 * [ops[0] m; ops[1] m; ...]. */
int brtm_linop_column(int n, brtm_linop *const *ops, brtm_linop **out);
/* diag_op (Born/combo_oper.cpp:103-175): block diagonal, domains and
 * ranges stacked. */
int brtm_linop_blockdiag(int n, brtm_linop *const *ops, brtm_linop **out);
/* my_operator::set_scale: every kind multiplies its output (forward and
 * adjoint) by s.  The reference's i_op is the kind that honours it there. */
int brtm_linop_set_scale(brtm_linop *A, double s);
long brtm_linop_domain_size(const brtm_linop *A);
long brtm_linop_range_size(const brtm_linop *A);
int brtm_linop_forward(brtm_linop *A, int add, const brtm_vec *model, brtm_vec *data);
int brtm_linop_adjoint(brtm_linop *A, int add, brtm_vec *model, const brtm_vec *data);
/* my_operator::dot_test (Born/my_operator.cpp:14-76), seeded. */
int brtm_linop_dot_test(brtm_linop *A, unsigned long long seed, double err[2]);
int brtm_linop_destroy(brtm_linop *A);

/* ---- solvers (lin_solver) ---------------------------------------------- */
#define BRTM_SOLVER_SIMPLE 0  /* min |W(L m - d)|                        */
#define BRTM_SOLVER_REG 1     /* min |W(L m - d)|^2 + |eps R m|^2         */
#define BRTM_SOLVER_PREC 2    /* m = P p, min |W(L P p - d)|^2 + |eps p|^2 */

typedef struct brtm_solver brtm_solver;

/* simple_solver / reg_solver / prec_solver constructors
 * (Born/lin_solver.cpp:7-53).  op: L.  aux: R (REG) or P (PREC), NULL for
 * SIMPLE.  wt: data weight W or NULL.  m0: starting model or NULL; the
 * solver then works on the update (rr = W(L m0 - d), init_rr,
 * Born/lin_solver.cpp:79-101) and return_model gives the update, as in the
 * reference.  REG calls set_scale(eps) on R (Born/lin_solver.cpp:45); PREC
 * appends eps times the identity on P's domain (Born/lin_solver.cpp:26-30).
 * The operators and vectors are borrowed. */
int brtm_solver_create(int kind, brtm_linop *op, const brtm_vec *data, brtm_linop *aux,
                       double eps, brtm_linop *wt, const brtm_vec *m0, brtm_solver **out);
/* lin_solver::solve (Born/lin_solver.cpp:115-143): niter more iterations
 * from the current state (a second call continues where the first ended).
 * The judged residual is |rr + e|/|rr0 + e| (rr0 the initial residual, e
 * the optional offset on the stacked range); resid_hist / time_hist as for
 * brtm_cg_solve. */
int brtm_solver_solve(brtm_solver *s, const brtm_cg_params *params, double *resid_hist,
                      double *time_hist, brtm_cg_stats *stats);
/* lin_solver::return_model (m) / prec_solver::return_model (P m) into
 * model (Born/lin_solver.cpp:144-155). */
int brtm_solver_return_model(brtm_solver *s, brtm_vec *model);
/* update_model: model += m (or P m) (Born/lin_solver.cpp:145,152-154). */
int brtm_solver_update_model(brtm_solver *s, brtm_vec *model);
/* The solver-domain iterate m and the stacked residual rr (borrowed). */
brtm_vec *brtm_solver_iterate(brtm_solver *s);
brtm_vec *brtm_solver_residual(brtm_solver *s);
int brtm_solver_destroy(brtm_solver *s);

#ifdef __cplusplus
}
#endif
#endif /* BORN_SOLVER_H */
