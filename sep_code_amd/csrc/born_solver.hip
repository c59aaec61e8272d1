// Composable linear operators (my_operator, i_op, tmute, row_op, col_op,
// diag_op) and the least-squares solvers (simple_solver, reg_solver,
// prec_solver with cgstep or flexible CG) on HBM-resident vectors.
//
// Stacked vectors are single brtm_vec allocations; an operator of a stack
// sees its part through a view (a brtm_vec whose ptr points into the
// stack), so whole-stack dots and updates stay one kernel each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/born_solver.h"
#include "born_ctx.h"
#include "gram_solve.h"
#include "tmute_start.h"

struct brtm_linop {
  brtm::CtxHold hold;
  enum Kind { BORN, IDENTITY, DIAG, TMUTE, CALLBACK, CHAIN, COLUMN, BLOCKDIAG };
  Kind kind = IDENTITY;
  brtm_ctx *ctx = nullptr;
  long nd = 0, nr = 0;  // domain and range sizes
  double scale = 1.0;
  brtm_op *op = nullptr;          // BORN
  const brtm_vec *w = nullptr;    // DIAG
  brtm::DevBuf tstart;            // TMUTE: first kept sample of each receiver column
  int nt = 0, nrx = 0;            // TMUTE
  long ntrace = 0;                // TMUTE: traces = nouter * nrx
  brtm_apply_fn fwd = nullptr, adj = nullptr;  // CALLBACK
  void *user = nullptr;
  std::vector<brtm_linop *> ops;  // CHAIN / COLUMN / BLOCKDIAG (borrowed)
  std::vector<long> doff, roff;   // part offsets in stacked domain / range
  std::vector<brtm_vec *> tmp;    // CHAIN: range of ops[i] for i < n-1
  brtm_vec *sd = nullptr, *sr = nullptr;  // scaled accumulation scratch

  ~brtm_linop() {
    for (brtm_vec *v : tmp) brtm_vec_destroy(v);
    brtm_vec_destroy(sd);
    brtm_vec_destroy(sr);
  }
};

struct brtm_solver {
  brtm::CtxHold hold;
  brtm_ctx *ctx = nullptr;
  int kind = BRTM_SOLVER_SIMPLE;
  brtm_linop *oper = nullptr;          // the stacked system
  brtm_linop *pop = nullptr;           // P (prec_solver)
  std::vector<brtm_linop *> owned;     // composites built by the constructor
  brtm_vec *m = nullptr, *rr = nullptr, *g = nullptr, *gg = nullptr, *s = nullptr,
           *ss = nullptr, *z = nullptr;
  std::vector<brtm_vec *> hs, hS;      // flexible-CG history, newest first
  int H = 0, nh = 0;
  bool first = true;                   // cgstep::forget (Born/cgstep.cpp:16-26)
  double rr0_rr0 = 0;                  // |rr0|^2
  const brtm_vec *offset = nullptr;    // offset of the first judged solve
  double jnorm = 0;                    // |rr0 + offset|

  ~brtm_solver() {
    for (brtm_vec *v : {m, rr, g, gg, s, ss, z}) brtm_vec_destroy(v);
    for (brtm_vec *v : hs) brtm_vec_destroy(v);
    for (brtm_vec *v : hS) brtm_vec_destroy(v);
    for (auto it = owned.rbegin(); it != owned.rend(); ++it) delete *it;
  }
};

namespace brtm {
namespace {

#define LS_TRY(x)                   \
  do {                              \
    int rc_ = (x);                  \
    if (rc_ != BRTM_OK) return rc_; \
  } while (0)

int grid_n(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

// out = (add ? out : 0) + s * (w ? w . in : in); out is not read when add = 0.
__global__ void k_scaled(float *out, const float *__restrict__ in, const float *__restrict__ w,
                         float s, int add, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    float v = s * (w ? w[i] * in[i] : in[i]);
    out[i] = add ? out[i] + v : v;
  }
}

// tmute on [trace][nt]: one block row per trace (strided over gridDim.y),
// samples along x.  Sample it of trace tr is kept when it >= start[tr % nrx].
__global__ void k_tmute(float *out, const float *__restrict__ in, const int *__restrict__ start,
                        int nt, int nrx, long ntrace, float s, int add) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nt) return;
  for (long tr = blockIdx.y; tr < ntrace; tr += gridDim.y) {
    const long i = tr * nt + it;
    const float v = it >= start[tr % nrx] ? s * in[i] : 0.f;
    out[i] = add ? out[i] + v : v;
  }
}

// A part [off, off + n) of a stacked vector.
struct View {
  brtm_vec v;
  View(const brtm_vec *base, long off, long n) {
    v.ctx = base->ctx;
    v.n = n;
    v.ptr = base->p() + off;
  }
};

int apply(brtm_linop *A, bool adjoint, int add, const brtm_vec *in, brtm_vec *out);

// The operator without its scale (elementwise kinds take it in the kernel).
int apply_raw(brtm_linop *A, bool adjoint, int add, const brtm_vec *in, brtm_vec *out,
              float s) {
  hipStream_t st = A->ctx->stream;
  switch (A->kind) {
    case brtm_linop::BORN:
      return adjoint ? brtm_op_adjoint(A->op, add, out, in) : brtm_op_forward(A->op, add, in, out);
    case brtm_linop::IDENTITY:
    case brtm_linop::DIAG:
      k_scaled<<<grid_n(out->n), 256, 0, st>>>(out->p(), in->p(),
                                               A->kind == brtm_linop::DIAG ? A->w->p() : nullptr,
                                               s, add, out->n);
      BRTM_HIP(hipGetLastError());
      return BRTM_OK;
    case brtm_linop::TMUTE: {
      dim3 grid((unsigned)((A->nt + 255) / 256), (unsigned)std::min<long>(A->ntrace, 65535));
      k_tmute<<<grid, 256, 0, st>>>(out->p(), in->p(), A->tstart.as<int>(), A->nt, A->nrx,
                                    A->ntrace, s, add);
      BRTM_HIP(hipGetLastError());
      return BRTM_OK;
    }
    case brtm_linop::CALLBACK: {
      brtm_apply_fn f = adjoint ? A->adj : A->fwd;
      BRTM_REQUIRE(f, BRTM_ESTATE, "operator has no such callback");
      if (f(A->user, add, in, out) != 0) {
        set_error(adjoint ? "adjoint callback failed" : "forward callback failed");
        return BRTM_ESTATE;
      }
      return BRTM_OK;
    }
    case brtm_linop::CHAIN: {
      // row_op::forward / adjoint (Born/combo_oper.cpp:176-235)
      const int n = (int)A->ops.size();
      if (!adjoint) {
        const brtm_vec *x = in;
        for (int i = 0; i < n; ++i) {
          brtm_vec *y = i == n - 1 ? out : A->tmp[i];
          LS_TRY(apply(A->ops[i], false, i == n - 1 ? add : 0, x, y));
          x = y;
        }
      } else {
        const brtm_vec *y = in;
        for (int i = n - 1; i >= 0; --i) {
          brtm_vec *x = i == 0 ? out : A->tmp[i - 1];
          LS_TRY(apply(A->ops[i], true, i == 0 ? add : 0, y, x));
          y = x;
        }
      }
      return BRTM_OK;
    }
    case brtm_linop::COLUMN: {
      // col_op::forward / adjoint (Born/combo_oper.cpp:56-99): the adjoint
      // sums the parts' adjoints into the one model.
      const int n = (int)A->ops.size();
      for (int i = 0; i < n; ++i) {
        brtm_linop *B = A->ops[i];
        if (!adjoint) {
          View part(out, A->roff[i], B->nr);
          LS_TRY(apply(B, false, add, in, &part.v));
        } else {
          View part(in, A->roff[i], B->nr);
          LS_TRY(apply(B, true, i == 0 ? add : 1, &part.v, out));
        }
      }
      return BRTM_OK;
    }
    case brtm_linop::BLOCKDIAG: {
      // diag_op::forward / adjoint (Born/combo_oper.cpp:142-175)
      const int n = (int)A->ops.size();
      for (int i = 0; i < n; ++i) {
        brtm_linop *B = A->ops[i];
        View dm(adjoint ? out : in, A->doff[i], B->nd);
        View rg(adjoint ? in : out, A->roff[i], B->nr);
        if (!adjoint)
          LS_TRY(apply(B, false, add, &dm.v, &rg.v));
        else
          LS_TRY(apply(B, true, add, &rg.v, &dm.v));
      }
      return BRTM_OK;
    }
  }
  set_error("unknown operator kind");
  return BRTM_EINVAL;
}

int apply(brtm_linop *A, bool adjoint, int add, const brtm_vec *in, brtm_vec *out) {
  const bool elementwise = A->kind == brtm_linop::IDENTITY || A->kind == brtm_linop::DIAG ||
                           A->kind == brtm_linop::TMUTE;
  if (elementwise || A->scale == 1.0) return apply_raw(A, adjoint, add, in, out, (float)A->scale);
  if (!add) {
    LS_TRY(apply_raw(A, adjoint, 0, in, out, 1.f));
    return brtm_vec_scale(out, A->scale);
  }
  brtm_vec *&t = adjoint ? A->sd : A->sr;
  if (!t) LS_TRY(brtm_vec_create(A->ctx, out->n, &t));
  LS_TRY(apply_raw(A, adjoint, 0, in, t, 1.f));
  return brtm_vec_scale_add(out, 1.0, t, A->scale);
}

int new_linop(brtm_ctx *ctx, brtm_linop::Kind kind, long nd, long nr, brtm_linop **out) {
  BRTM_REQUIRE(ctx && out, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(nd > 0 && nr > 0, BRTM_EINVAL, "operator sizes must be positive");
  BRTM_LIVE(ctx);
  brtm_linop *A = new brtm_linop();
  A->hold.reset(ctx);
  A->kind = kind;
  A->ctx = ctx;
  A->nd = nd;
  A->nr = nr;
  *out = A;
  return BRTM_OK;
}

int check_parts(int n, brtm_linop *const *ops) {
  BRTM_REQUIRE(n >= 1 && ops, BRTM_EINVAL, "a combination needs at least one operator");
  for (int i = 0; i < n; ++i) {
    BRTM_REQUIRE(ops[i], BRTM_EINVAL, "NULL operator in a combination");
    BRTM_REQUIRE(ops[i]->ctx == ops[0]->ctx, BRTM_EINVAL,
                 "operators of a combination must share a context");
  }
  return BRTM_OK;
}

// One iteration of lin_solver::solve (Born/lin_solver.cpp:123-141):
// g = A' rr, direction = P(g) (or g), gg = A direction, then the step.
// Returns BRTM_OK with *stop set on a breakdown.
int iterate(brtm_solver *S, const brtm_cg_params *prm, brtm_cg_stats *st, double *alfa_out,
            double *beta_out, bool *stop) {
  const double small = 1e-20;
  brtm_linop *A = S->oper;
  LS_TRY(apply(A, true, 0, S->rr, S->g));
  brtm_vec *dir = S->g;
  if (prm->precond) {
    if (!S->z) LS_TRY(brtm_vec_create(S->ctx, A->nd, &S->z));
    if (prm->precond(prm->precond_user, S->g, S->z) != 0) {
      set_error("preconditioner failed");
      return BRTM_ESTATE;
    }
    double zg = 0;
    LS_TRY(brtm_vec_dot(S->z, S->g, &zg));
    if (zg > 0 && std::isfinite(zg)) dir = S->z;  // descent direction for |Am-d|^2/2 is -g
    else ++st->fallbacks;
  }
  LS_TRY(apply(A, false, 0, dir, S->gg));
  double alfa = 0, beta = 0;
  if (S->H) {
    // Flexible CG: minimise |rr + a gg + sum_i b_i hS_i| over (a, b).
    const int K = S->nh + 1;
    const brtm_vec *B[BRTM_CG_MAX_HISTORY + 1];
    B[0] = S->gg;
    for (int i = 0; i < S->nh; ++i) B[i + 1] = S->hS[i];
    double M[(BRTM_CG_MAX_HISTORY + 1) * (BRTM_CG_MAX_HISTORY + 1)], bv[BRTM_CG_MAX_HISTORY + 1],
        x[BRTM_CG_MAX_HISTORY + 1];
    for (int a = 0; a < K; ++a) {
      for (int b = 0; b <= a; ++b) {
        LS_TRY(brtm_vec_dot(B[a], B[b], &M[a * K + b]));
        M[b * K + a] = M[a * K + b];
      }
      LS_TRY(brtm_vec_dot(B[a], S->rr, &bv[a]));
      bv[a] = -bv[a];
    }
    if (std::fabs(M[0]) < small || gram_solve(K, M, bv, x) == 0) {
      *stop = true;
      return BRTM_OK;
    }
    // new step into the scratch pair (s, ss)
    LS_TRY(brtm_vec_copy(S->s, dir));
    LS_TRY(brtm_vec_scale(S->s, x[0]));
    LS_TRY(brtm_vec_copy(S->ss, S->gg));
    LS_TRY(brtm_vec_scale(S->ss, x[0]));
    for (int i = 0; i < S->nh; ++i) {
      if (x[i + 1] == 0) continue;
      LS_TRY(brtm_vec_scale_add(S->s, 1.0, S->hs[i], x[i + 1]));
      LS_TRY(brtm_vec_scale_add(S->ss, 1.0, S->hS[i], x[i + 1]));
    }
    LS_TRY(brtm_vec_add(S->m, S->s));
    LS_TRY(brtm_vec_add(S->rr, S->ss));
    // push: the new step becomes hs[0]; the dropped (or unused) slot
    // becomes the next scratch pair
    const int slot = S->nh < S->H ? S->nh : S->H - 1;
    std::swap(S->s, S->hs[slot]);
    std::swap(S->ss, S->hS[slot]);
    std::rotate(S->hs.begin(), S->hs.begin() + slot, S->hs.begin() + slot + 1);
    std::rotate(S->hS.begin(), S->hS.begin() + slot, S->hS.begin() + slot + 1);
    S->nh = std::min(S->nh + 1, S->H);
    alfa = x[0];
    beta = S->nh > 1 ? x[1] : 0.0;
  } else {
    // cgstep::step_it (Born/cgstep.cpp:14-58)
    double gg_gg = 0;
    LS_TRY(brtm_vec_dot(S->gg, S->gg, &gg_gg));
    if (S->first) {
      if (std::fabs(gg_gg) < small) {
        *stop = true;
        return BRTM_OK;
      }
      double gg_rr = 0;
      LS_TRY(brtm_vec_dot(S->gg, S->rr, &gg_rr));
      beta = 0;
      alfa = -gg_rr / gg_gg;
    } else {
      double ss_ss = 0, gg_ss = 0, gg_rr = 0, ss_rr = 0;
      LS_TRY(brtm_vec_dot(S->ss, S->ss, &ss_ss));
      LS_TRY(brtm_vec_dot(S->gg, S->ss, &gg_ss));
      LS_TRY(brtm_vec_dot(S->gg, S->rr, &gg_rr));
      LS_TRY(brtm_vec_dot(S->ss, S->rr, &ss_rr));
      if (std::fabs(gg_gg) < small || std::fabs(ss_ss) < small) {
        *stop = true;
        return BRTM_OK;
      }
      const double determ = gg_gg * ss_ss * (1.0 - gg_ss / gg_gg * gg_ss / ss_ss);
      gg_rr = -gg_rr;
      ss_rr = -ss_rr;
      alfa = (ss_ss * gg_rr - gg_ss * ss_rr) / determ;
      beta = (-gg_ss * gg_rr + gg_gg * ss_rr) / determ;
    }
    LS_TRY(brtm_vec_scale_add(S->s, beta, dir, alfa));      // s  = beta s  + alfa g
    LS_TRY(brtm_vec_scale_add(S->ss, beta, S->gg, alfa));   // ss = beta ss + alfa gg
    LS_TRY(brtm_vec_add(S->m, S->s));                       // m += s
    LS_TRY(brtm_vec_add(S->rr, S->ss));                     // rr += ss
  }
  S->first = false;
  *alfa_out = alfa;
  *beta_out = beta;
  return BRTM_OK;
}

// History buffers for h >= 2 (flexible CG); a change of h restarts them.
int set_history(brtm_solver *S, int history) {
  const int H = history >= 2 ? history : 0;
  if (H == S->H) return BRTM_OK;
  for (brtm_vec *v : S->hs) brtm_vec_destroy(v);
  for (brtm_vec *v : S->hS) brtm_vec_destroy(v);
  S->hs.assign(H, nullptr);
  S->hS.assign(H, nullptr);
  S->H = H;
  S->nh = 0;
  for (int i = 0; i < H; ++i) {
    LS_TRY(brtm_vec_create(S->ctx, S->oper->nd, &S->hs[i]));
    LS_TRY(brtm_vec_create(S->ctx, S->oper->nr, &S->hS[i]));
  }
  return BRTM_OK;
}

}  // namespace
}  // namespace brtm

using namespace brtm;

extern "C" {

// ------------------------------------------------------------- operators
int brtm_linop_born(brtm_op *op, brtm_linop **out) {
  BRTM_REQUIRE(op && out, BRTM_EINVAL, "NULL argument");
  const long nm = brtm_op_model_size(op), nd = brtm_op_data_size(op);
  LS_TRY(new_linop(op_context(op), brtm_linop::BORN, nm, nd, out));
  (*out)->op = op;
  return BRTM_OK;
}

int brtm_linop_identity(brtm_ctx *ctx, long n, brtm_linop **out) {
  return new_linop(ctx, brtm_linop::IDENTITY, n, n, out);
}

int brtm_linop_diagonal(const brtm_vec *w, brtm_linop **out) {
  BRTM_REQUIRE(w && w->p(), BRTM_EINVAL, "NULL weight");
  LS_TRY(new_linop(w->ctx, brtm_linop::DIAG, w->n, w->n, out));
  (*out)->w = w;
  return BRTM_OK;
}

int brtm_linop_tmute(brtm_ctx *ctx, int nt, float ot, float dt, int nrx, float orx, float drx,
                     long nouter, float t0, float vmute, brtm_linop **out) {
  BRTM_REQUIRE(ctx && out, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(nt > 0 && nrx > 0 && nouter > 0 && dt > 0 && vmute != 0, BRTM_EINVAL,
               "bad tmute geometry");
  std::vector<int> start(nrx);
  tmute_starts(nt, ot, dt, nrx, orx, drx, t0, vmute, start.data());
  brtm_linop *A = nullptr;
  LS_TRY(new_linop(ctx, brtm_linop::TMUTE, nouter * nrx * (long)nt, nouter * nrx * (long)nt, &A));
  A->nt = nt;
  A->nrx = nrx;
  A->ntrace = nouter * nrx;
  int rc = A->tstart.alloc(sizeof(int) * (size_t)nrx);
  if (rc == BRTM_OK &&
      hipMemcpy(A->tstart.ptr, start.data(), sizeof(int) * nrx, hipMemcpyHostToDevice) !=
          hipSuccess) {
    set_error("tmute table upload failed");
    rc = BRTM_EHIP;
  }
  if (rc != BRTM_OK) {
    delete A;
    return rc;
  }
  *out = A;
  return BRTM_OK;
}

int brtm_linop_callback(brtm_ctx *ctx, long ndomain, long nrange, brtm_apply_fn forward,
                        brtm_apply_fn adjoint, void *user, brtm_linop **out) {
  BRTM_REQUIRE(forward && adjoint, BRTM_EINVAL, "a callback operator needs both directions");
  LS_TRY(new_linop(ctx, brtm_linop::CALLBACK, ndomain, nrange, out));
  (*out)->fwd = forward;
  (*out)->adj = adjoint;
  (*out)->user = user;
  return BRTM_OK;
}

int brtm_linop_chain(int n, brtm_linop *const *ops, brtm_linop **out) {
  LS_TRY(check_parts(n, ops));
  for (int i = 1; i < n; ++i)
    BRTM_REQUIRE(ops[i]->nd == ops[i - 1]->nr, BRTM_EINVAL,
                 "chain: the range of each operator must be the domain of the next");
  brtm_linop *A = nullptr;
  LS_TRY(new_linop(ops[0]->ctx, brtm_linop::CHAIN, ops[0]->nd, ops[n - 1]->nr, &A));
  A->ops.assign(ops, ops + n);
  A->tmp.assign(n - 1, nullptr);
  for (int i = 0; i < n - 1; ++i) {
    const int rc = brtm_vec_create(A->ctx, ops[i]->nr, &A->tmp[i]);
    if (rc != BRTM_OK) {
      delete A;
      return rc;
    }
  }
  *out = A;
  return BRTM_OK;
}

int brtm_linop_column(int n, brtm_linop *const *ops, brtm_linop **out) {
  LS_TRY(check_parts(n, ops));
  long nr = 0;
  std::vector<long> roff(n);
  for (int i = 0; i < n; ++i) {
    BRTM_REQUIRE(ops[i]->nd == ops[0]->nd, BRTM_EINVAL, "column: the domains must match");
    roff[i] = nr;
    nr += ops[i]->nr;
  }
  LS_TRY(new_linop(ops[0]->ctx, brtm_linop::COLUMN, ops[0]->nd, nr, out));
  (*out)->ops.assign(ops, ops + n);
  (*out)->roff = roff;
  return BRTM_OK;
}

int brtm_linop_blockdiag(int n, brtm_linop *const *ops, brtm_linop **out) {
  LS_TRY(check_parts(n, ops));
  long nd = 0, nr = 0;
  std::vector<long> doff(n), roff(n);
  for (int i = 0; i < n; ++i) {
    doff[i] = nd;
    roff[i] = nr;
    nd += ops[i]->nd;
    nr += ops[i]->nr;
  }
  LS_TRY(new_linop(ops[0]->ctx, brtm_linop::BLOCKDIAG, nd, nr, out));
  (*out)->ops.assign(ops, ops + n);
  (*out)->doff = doff;
  (*out)->roff = roff;
  return BRTM_OK;
}

int brtm_linop_set_scale(brtm_linop *A, double s) {
  BRTM_REQUIRE(A && std::isfinite(s), BRTM_EINVAL, "bad scale");
  A->scale = s;
  return BRTM_OK;
}

long brtm_linop_domain_size(const brtm_linop *A) { return A ? A->nd : 0; }
long brtm_linop_range_size(const brtm_linop *A) { return A ? A->nr : 0; }

int brtm_linop_forward(brtm_linop *A, int add, const brtm_vec *model, brtm_vec *data) {
  BRTM_REQUIRE(A, BRTM_EINVAL, "NULL operator");
  LS_TRY(check_vec(model, A->nd, "model"));
  LS_TRY(check_vec(data, A->nr, "data"));
  BRTM_HIP(hipSetDevice(A->ctx->device));
  return apply(A, false, add ? 1 : 0, model, data);
}

int brtm_linop_adjoint(brtm_linop *A, int add, brtm_vec *model, const brtm_vec *data) {
  BRTM_REQUIRE(A, BRTM_EINVAL, "NULL operator");
  LS_TRY(check_vec(model, A->nd, "model"));
  LS_TRY(check_vec(data, A->nr, "data"));
  BRTM_HIP(hipSetDevice(A->ctx->device));
  return apply(A, true, add ? 1 : 0, data, model);
}

int brtm_linop_dot_test(brtm_linop *A, unsigned long long seed, double err[2]) {
  BRTM_REQUIRE(A && err, BRTM_EINVAL, "NULL argument");
  brtm_vec *m1 = nullptr, *m2 = nullptr, *d1 = nullptr, *d2 = nullptr;
  int rc = brtm_vec_create(A->ctx, A->nd, &m1);
  if (rc == BRTM_OK) rc = brtm_vec_create(A->ctx, A->nd, &m2);
  if (rc == BRTM_OK) rc = brtm_vec_create(A->ctx, A->nr, &d1);
  if (rc == BRTM_OK) rc = brtm_vec_create(A->ctx, A->nr, &d2);
  double dot1 = 0, dot2 = 0, dot3 = 0, dot4 = 0;
  // Born/my_operator.cpp:14-76
  if (rc == BRTM_OK) rc = brtm_vec_random(m1, seed);
  if (rc == BRTM_OK) rc = brtm_vec_random(d2, seed ^ 0x5DEECE66Dull);
  if (rc == BRTM_OK) rc = brtm_linop_forward(A, 0, m1, d1);
  if (rc == BRTM_OK) rc = brtm_vec_dot(d1, d2, &dot1);
  if (rc == BRTM_OK) rc = brtm_linop_adjoint(A, 0, m2, d2);
  if (rc == BRTM_OK) rc = brtm_vec_dot(m2, m1, &dot2);
  if (rc == BRTM_OK) rc = brtm_linop_forward(A, 1, m1, d1);
  if (rc == BRTM_OK) rc = brtm_linop_adjoint(A, 1, m2, d2);
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

int brtm_linop_destroy(brtm_linop *A) {
  if (!A) return BRTM_OK;
  (void)hipSetDevice(A->ctx->device);
  (void)hipStreamSynchronize(A->ctx->stream);
  delete A;
  return BRTM_OK;
}

// --------------------------------------------------------------- solvers
int brtm_solver_create(int kind, brtm_linop *op, const brtm_vec *data, brtm_linop *aux,
                       double eps, brtm_linop *wt, const brtm_vec *m0, brtm_solver **out) {
  BRTM_REQUIRE(op && data && out, BRTM_EINVAL, "NULL argument");
  BRTM_REQUIRE(kind == BRTM_SOLVER_SIMPLE || kind == BRTM_SOLVER_REG || kind == BRTM_SOLVER_PREC,
               BRTM_EINVAL, "unknown solver kind");
  BRTM_REQUIRE(kind == BRTM_SOLVER_SIMPLE || aux, BRTM_EINVAL,
               "reg_solver / prec_solver need their second operator");
  BRTM_REQUIRE(std::isfinite(eps), BRTM_EINVAL, "eps must be finite");
  LS_TRY(check_vec(data, op->nr, "data"));
  if (wt)
    BRTM_REQUIRE(wt->nd == op->nr, BRTM_EINVAL, "wt and op don't match");  // lin_solver.cpp:83
  if (m0) LS_TRY(check_vec(m0, op->nd, "m0"));
  if (kind == BRTM_SOLVER_REG)
    BRTM_REQUIRE(aux->nd == op->nd, BRTM_EINVAL, "the regularisation must act on the model");
  if (kind == BRTM_SOLVER_PREC)
    BRTM_REQUIRE(aux->nr == op->nd, BRTM_EINVAL, "the preconditioner's range must be the model");
  brtm_ctx *c = op->ctx;
  BRTM_LIVE(c);
  BRTM_HIP(hipSetDevice(c->device));
  brtm_solver *S = new brtm_solver();
  S->hold.reset(c);
  S->ctx = c;
  S->kind = kind;
  auto fail = [&](int rc) {
    delete S;
    return rc;
  };
  int rc = BRTM_OK;
  // create_wt_op (Born/lin_solver.cpp:70-78): W o L
  brtm_linop *wt_op = op;
  if (wt) {
    brtm_linop *parts[2] = {op, wt};
    if ((rc = brtm_linop_chain(2, parts, &wt_op)) != BRTM_OK) return fail(rc);
    S->owned.push_back(wt_op);
  }
  // init_rr (Born/lin_solver.cpp:79-101): rtemp = W(L m0 - d) (the intended
  // form; the reference writes the m0 term into an unset member there).
  const long nres = wt ? wt->nr : op->nr;
  brtm_vec *rtemp = nullptr, *dtemp = nullptr;
  if ((rc = brtm_vec_create(c, nres, &rtemp)) != BRTM_OK) return fail(rc);
  if (!wt) {
    rc = brtm_vec_copy(rtemp, data);
    if (rc == BRTM_OK && m0) rc = brtm_vec_scale(rtemp, -1.0);
    if (rc == BRTM_OK && m0) rc = brtm_linop_forward(op, 1, m0, rtemp);
    if (rc == BRTM_OK && !m0) rc = brtm_vec_scale(rtemp, -1.0);
  } else {
    rc = brtm_vec_create(c, op->nr, &dtemp);
    if (rc == BRTM_OK) rc = brtm_vec_copy(dtemp, data);
    if (rc == BRTM_OK) rc = brtm_vec_scale(dtemp, -1.0);
    if (rc == BRTM_OK && m0) rc = brtm_linop_forward(op, 1, m0, dtemp);
    if (rc == BRTM_OK) rc = brtm_linop_forward(wt, 0, dtemp, rtemp);
  }
  brtm_vec_destroy(dtemp);
  if (rc != BRTM_OK) {
    brtm_vec_destroy(rtemp);
    return fail(rc);
  }
  long extra = 0;  // rows appended to the residual (zero at the start)
  if (kind == BRTM_SOLVER_SIMPLE) {
    S->oper = wt_op;
  } else if (kind == BRTM_SOLVER_REG) {
    // reg_solver (Born/lin_solver.cpp:39-53): [W L; eps R]
    rc = brtm_linop_set_scale(aux, eps);
    brtm_linop *parts[2] = {wt_op, aux};
    if (rc == BRTM_OK) rc = brtm_linop_column(2, parts, &S->oper);
    if (rc == BRTM_OK) S->owned.push_back(S->oper);
    extra = aux->nr;
  } else {
    // prec_solver (Born/lin_solver.cpp:21-37): [W L P; eps I] on P's domain
    brtm_linop *iop = nullptr, *chain = nullptr;
    rc = brtm_linop_identity(c, aux->nd, &iop);
    if (rc == BRTM_OK) S->owned.push_back(iop);
    if (rc == BRTM_OK) rc = brtm_linop_set_scale(iop, eps);
    brtm_linop *cparts[2] = {aux, wt_op};
    if (rc == BRTM_OK) rc = brtm_linop_chain(2, cparts, &chain);
    if (rc == BRTM_OK) S->owned.push_back(chain);
    brtm_linop *parts[2] = {chain, iop};
    if (rc == BRTM_OK) rc = brtm_linop_column(2, parts, &S->oper);
    if (rc == BRTM_OK) S->owned.push_back(S->oper);
    S->pop = aux;
    extra = aux->nd;
  }
  if (rc != BRTM_OK) {
    brtm_vec_destroy(rtemp);
    return fail(rc);
  }
  // range_vec: rr = [rtemp; 0] (Born/combo_oper.cpp:237-243)
  if (extra == 0) {
    S->rr = rtemp;
  } else {
    rc = brtm_vec_create(c, S->oper->nr, &S->rr);
    if (rc == BRTM_OK) {
      View head(S->rr, 0, nres);
      rc = brtm_vec_copy(&head.v, rtemp);
    }
    brtm_vec_destroy(rtemp);
    if (rc != BRTM_OK) return fail(rc);
  }
  // create_solver_vecs (Born/lin_solver.cpp:62-69)
  rc = brtm_vec_create(c, S->oper->nd, &S->m);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, S->oper->nd, &S->g);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, S->oper->nd, &S->s);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, S->oper->nr, &S->gg);
  if (rc == BRTM_OK) rc = brtm_vec_create(c, S->oper->nr, &S->ss);
  if (rc == BRTM_OK) rc = brtm_vec_dot(S->rr, S->rr, &S->rr0_rr0);
  if (rc != BRTM_OK) return fail(rc);
  *out = S;
  return BRTM_OK;
}

int brtm_solver_solve(brtm_solver *S, const brtm_cg_params *prm, double *resid_hist,
                      double *time_hist, brtm_cg_stats *stats) {
  BRTM_REQUIRE(S && prm, BRTM_EINVAL, "NULL argument");
  BRTM_LIVE(S->ctx);
  BRTM_REQUIRE(prm->niter >= 0, BRTM_EINVAL, "niter must be >= 0");
  BRTM_REQUIRE(prm->history >= 0 && prm->history <= BRTM_CG_MAX_HISTORY, BRTM_EINVAL,
               "history must be in [0, BRTM_CG_MAX_HISTORY]");
  brtm_ctx *c = S->ctx;
  BRTM_HIP(hipSetDevice(c->device));
  LS_TRY(set_history(S, prm->history));
  brtm_cg_stats st{};
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  // Judged residual |rr + e| / |rr0 + e|; rr0 + e is fixed by the first
  // solve that passes e.
  const brtm_vec *e = prm->offset;
  double ee = 0;
  if (e) {
    LS_TRY(check_vec(e, S->oper->nr, "offset"));
    if (S->offset != e) {
      BRTM_REQUIRE(S->first, BRTM_ESTATE,
                   "the offset must be given from the first solve of this solver");
      double re = 0;
      LS_TRY(brtm_vec_dot(e, e, &ee));
      LS_TRY(brtm_vec_dot(S->rr, e, &re));
      S->offset = e;
      S->jnorm = std::sqrt(std::max(S->rr0_rr0 + 2 * re + ee, 0.0));
    }
    LS_TRY(brtm_vec_dot(e, e, &ee));
  }
  const double dnorm = std::sqrt(S->rr0_rr0);
  auto judged = [&](double rr_rr, double *out) -> int {
    if (!e) {
      *out = dnorm > 0 ? std::sqrt(rr_rr) / dnorm : 0.0;
      return BRTM_OK;
    }
    double re = 0;
    const int r2 = brtm_vec_dot(S->rr, e, &re);
    *out = S->jnorm > 0 ? std::sqrt(std::max(rr_rr + 2 * re + ee, 0.0)) / S->jnorm : 0.0;
    return r2;
  };
  // |rr| / |d|: the residual of the data as given (noisy when e is set)
  auto data_rel = [&](double rr_rr) { return dnorm > 0 ? std::sqrt(rr_rr) / dnorm : 0.0; };
  double rr_rr = 0;
  LS_TRY(brtm_vec_dot(S->rr, S->rr, &rr_rr));
  LS_TRY(judged(rr_rr, &st.rel_resid));
  st.data_rel_resid = data_rel(rr_rr);
  if (resid_hist) resid_hist[0] = st.rel_resid;
  if (prm->data_resid_hist) prm->data_resid_hist[0] = st.data_rel_resid;
  if (time_hist) time_hist[0] = 0.0;
  int rc = BRTM_OK;
  for (int iter = 0; iter < prm->niter; ++iter) {
    if (dnorm == 0) {
      st.breakdown = 1;
      break;
    }
    double alfa = 0, beta = 0;
    bool stop = false;
    rc = iterate(S, prm, &st, &alfa, &beta, &stop);
    if (rc != BRTM_OK) break;
    if (stop) {
      st.breakdown = 1;
      break;
    }
    st.iterations = iter + 1;
    rc = brtm_vec_dot(S->rr, S->rr, &rr_rr);
    if (rc == BRTM_OK) rc = judged(rr_rr, &st.rel_resid);
    if (rc != BRTM_OK) break;
    st.data_rel_resid = data_rel(rr_rr);
    if (resid_hist) resid_hist[iter + 1] = st.rel_resid;
    if (prm->data_resid_hist) prm->data_resid_hist[iter + 1] = st.data_rel_resid;
    if (time_hist) time_hist[iter + 1] = elapsed();
    if (prm->verbose)
      std::fprintf(stderr, "cg iter %d alpha=%g beta=%g rel_resid=%.6e t=%.2fs\n", iter, alfa,
                   beta, st.rel_resid, elapsed());
    if (prm->rel_tol > 0 && st.rel_resid <= prm->rel_tol) {
      st.converged = 1;
      break;
    }
  }
  if (rc == BRTM_OK) {
    const hipError_t he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) {
      set_error(hipGetErrorString(he));
      rc = BRTM_EHIP;
    }
  }
  st.seconds = elapsed();
  if (stats) *stats = st;
  return rc;
}

int brtm_solver_return_model(brtm_solver *S, brtm_vec *model) {
  BRTM_REQUIRE(S, BRTM_EINVAL, "NULL solver");
  if (S->pop) return brtm_linop_forward(S->pop, 0, S->m, model);
  LS_TRY(check_vec(model, S->m->n, "model"));
  return brtm_vec_copy(model, S->m);
}

int brtm_solver_update_model(brtm_solver *S, brtm_vec *model) {
  BRTM_REQUIRE(S, BRTM_EINVAL, "NULL solver");
  if (S->pop) return brtm_linop_forward(S->pop, 1, S->m, model);
  LS_TRY(check_vec(model, S->m->n, "model"));
  return brtm_vec_add(model, S->m);
}

brtm_vec *brtm_solver_iterate(brtm_solver *S) { return S ? S->m : nullptr; }
brtm_vec *brtm_solver_residual(brtm_solver *S) { return S ? S->rr : nullptr; }

int brtm_solver_destroy(brtm_solver *S) {
  if (!S) return BRTM_OK;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  delete S;
  return BRTM_OK;
}

// ---------------------------------------------------- simple CG on L m = d
int brtm_cg_solve(brtm_op *op, const brtm_vec *data, brtm_vec *model,
                  const brtm_cg_params *prm, double *resid_hist, double *time_hist,
                  brtm_cg_stats *stats) {
  BRTM_REQUIRE(op && data && model && prm, BRTM_EINVAL, "NULL argument");
  LS_TRY(check_vec(model, brtm_op_model_size(op), "model"));
  LS_TRY(check_vec(data, brtm_op_data_size(op), "data"));
  brtm_linop *L = nullptr;
  brtm_solver *S = nullptr;
  LS_TRY(brtm_linop_born(op, &L));
  int rc = brtm_solver_create(BRTM_SOLVER_SIMPLE, L, data, nullptr, 0.0, nullptr, nullptr, &S);
  if (rc == BRTM_OK) rc = brtm_solver_solve(S, prm, resid_hist, time_hist, stats);
  const int rc2 = S ? brtm_solver_return_model(S, model) : BRTM_OK;
  if (rc == BRTM_OK) rc = rc2;
  brtm_solver_destroy(S);
  brtm_linop_destroy(L);
  return rc;
}

}  // extern "C"
