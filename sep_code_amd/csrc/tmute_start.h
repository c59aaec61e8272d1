// Host-side first kept sample of each receiver column of tmute
// (brtm_linop_tmute): plain C++, so that it can also be built and checked on
// its own.
#pragma once

namespace brtm {

// Born/tmute.cpp:28-29,63-64: float t0 = t_0 + (o2 + d2*i2)/vmute; the first
// kept sample is (int)((t0 - o1)/d1) (truncation), clamped to [0, nt].
// One rounding per statement (no fused multiply-add), as the reference's
// host build computes it.
static void tmute_starts(int nt, float ot, float dt, int nrx, float orx, float drx, float t0,
                         float vmute, int *start) {
  for (int i = 0; i < nrx; ++i) {
    const float x = drx * (float)i;
    const float xo = orx + x;
    const float q = xo / vmute;
    const float t = t0 + q;
    const float u = t - ot;
    const float f = u / dt;
    start[i] = f <= 0.f ? 0 : (f >= (float)nt ? nt : (int)f);
  }
}

}  // namespace brtm
