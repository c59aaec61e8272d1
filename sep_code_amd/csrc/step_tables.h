// Host-side tables of the time step, set up by brtm_create_space: the
// stencil weights and the sponge weights.  Plain C++, so that they can also
// be built and checked on their own.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace brtm {

// Born/wave_fkernel.3d8o.cu:26-27 evaluated in double: c[0] the centre,
// c[1 + 3k + a] the weight of distance k+1 along axis a.
static void stencil_coeffs(float d1, float d2, float d3, float *coeffs) {
  const double C[3] = {8.0 / (5.0 * (double)d1 * d1), 8.0 / (5.0 * (double)d2 * d2),
                       8.0 / (5.0 * (double)d3 * d3)};
  const double f[4] = {1.0, -1.0 / 8.0, 1.0 / 63.0, -1.0 / 896.0};
  coeffs[0] = (float)(-1025.0 / 576.0 * (C[0] + C[1] + C[2]));
  for (int k = 0; k < 4; ++k)
    for (int a = 0; a < 3; ++a) coeffs[1 + 3 * k + a] = (float)(C[a] * f[k]);
}

// Sponge weights in FP64 (Born/wave_fkernel.3d8o.cu:201-202): w(e) for edge
// distances e < nw = ceil(bc_a), 1 beyond (the table holds at least one).
static void sponge_weights(float bc_a, float bc_b, int *nw, std::vector<float> *wtab) {
  *nw = (int)std::ceil((double)bc_a);
  wtab->assign(std::max(*nw, 1), 1.f);
  for (int e = 0; e < *nw; ++e) {
    double w = std::exp(-(double)bc_b * ((double)bc_a - e));
    (*wtab)[e] = (float)(w < 1.0 ? w : 1.0);
  }
}

}  // namespace brtm
