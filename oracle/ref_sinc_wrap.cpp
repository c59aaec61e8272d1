// C wrapper (our own code) around the reference's Born/sinc_bob.cpp, compiled
// from the reference sources in place by oracle/Makefile into
// oracle/_ref/libref_sinc.so.  Test infrastructure only: it pins the
// interpolation tables the product computes.
#include "sinc_bob.h"

extern "C" int ref_sinc_table(int ntab, int nsinc, float *out) {
  sinc_bob s(ntab, nsinc);
  for (int i = 0; i < ntab; i++)
    for (int k = 0; k < nsinc; k++) out[i * nsinc + k] = s.table[i][k];
  return 0;
}
