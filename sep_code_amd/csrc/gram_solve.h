// Host-side Gram solve of the flexible CG (brtm_solver_solve with history
// h >= 2): plain C++, no device code, so that it can also be built and
// checked on its own.
#pragma once
#include <cmath>

#include "../../include/born_op.h"

namespace brtm {

// Solve the K x K symmetric system M x = b (FP64) by Cholesky, skipping
// basis vectors whose pivot falls below tol times their own norm (nearly
// dependent on the ones before them; a vector of no length is one): their
// coefficient is 0.  Returns the number of basis vectors kept (0 when none
// is; the flexible CG stops earlier when its new direction has no length).
static int gram_solve(int K, const double *M, const double *b, double *x, double tol = 1e-9) {
  double Lm[BRTM_CG_MAX_HISTORY + 1][BRTM_CG_MAX_HISTORY + 1] = {};
  bool keep[BRTM_CG_MAX_HISTORY + 1] = {};
  int kept = 0;
  for (int i = 0; i < K; ++i) {
    double d = M[i * K + i];
    for (int k = 0; k < i; ++k)
      if (keep[k]) d -= Lm[i][k] * Lm[i][k];
    if (!(d > tol * M[i * K + i]) || !(M[i * K + i] > 0)) {
      keep[i] = false;
      continue;
    }
    keep[i] = true;
    ++kept;
    Lm[i][i] = std::sqrt(d);
    for (int j = i + 1; j < K; ++j) {
      double v = M[j * K + i];
      for (int k = 0; k < i; ++k)
        if (keep[k]) v -= Lm[j][k] * Lm[i][k];
      Lm[j][i] = v / Lm[i][i];
    }
  }
  double y[BRTM_CG_MAX_HISTORY + 1] = {};
  for (int i = 0; i < K; ++i) {
    if (!keep[i]) continue;
    double v = b[i];
    for (int k = 0; k < i; ++k)
      if (keep[k]) v -= Lm[i][k] * y[k];
    y[i] = v / Lm[i][i];
  }
  for (int i = K - 1; i >= 0; --i) {
    x[i] = 0;
    if (!keep[i]) continue;
    double v = y[i];
    for (int k = i + 1; k < K; ++k)
      if (keep[k]) v -= Lm[k][i] * x[k];
    x[i] = v / Lm[i][i];
  }
  return kept;
}

}  // namespace brtm
