// The reference's device ABI (Born/gpu_funcs_3d.h:5-27) over a default
// context of this library; see include/gpu_funcs_3d.h for the semantics.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/born_rtm.h"
#include "../../include/gpu_funcs_3d.h"

namespace {

brtm_ctx *g_ctx = nullptr;
int g_n[3] = {0, 0, 0};
int g_nrx = 0, g_nry = 0, g_ntr = 0;

void fail(const char *where) {
  std::fprintf(stderr, "%s: %s\n", where, brtm_last_error());
  std::exit(-1);
}

void ok(int rc, const char *where) {
  if (rc != BRTM_OK) fail(where);
}

brtm_ctx *ctx() {
  if (!g_ctx) ok(brtm_create(0, &g_ctx), "setup_cuda");
  return g_ctx;
}

}  // namespace

extern "C" {

void setup_cuda(int ngpus, int argc, char **argv) {
  (void)ngpus;
  (void)argc;
  (void)argv;
  std::fprintf(stderr, "%s: one GPU holds the whole domain\n", brtm_version());
  ctx();
}

void create_gpu_space(float d1, float d2, float d3, float bc_a, float bc_b, float bc_b_y, int nx,
                      int ny, int nz) {
  (void)bc_b_y;
  ok(brtm_create_space(ctx(), d1, d2, d3, bc_a, bc_b, nx, ny, nz), "create_gpu_space");
  g_n[0] = nx;
  g_n[1] = ny;
  g_n[2] = nz;
}

void transfer_vel_func1(int nx, int ny, int nz, float *vel) {
  if (nx != g_n[0] || ny != g_n[1] || nz != g_n[2]) {
    std::fprintf(stderr, "transfer_vel_func1: size differs from create_gpu_space\n");
    std::exit(-1);
  }
  ok(brtm_set_velocity(ctx(), vel), "transfer_vel_func1");
}

void transfer_vel_func2(int nx, int ny, int nz, float *vel) {
  (void)nx;
  (void)ny;
  (void)nz;
  (void)vel;
}

static void sinc_table(int which, int nsinc, int ns, float **tables) {
  std::vector<float> flat((size_t)nsinc * ns);
  for (int i = 0; i < ns; ++i)
    for (int k = 0; k < nsinc; ++k) flat[(size_t)i * nsinc + k] = tables[i][k];
  ok(brtm_set_sinc_table(ctx(), which, nsinc, ns, flat.data()), "transfer_sinc_table");
}

void transfer_sinc_table_s(int nsinc, int ns, float **tables) { sinc_table(0, nsinc, ns, tables); }
void transfer_sinc_table_d(int nsinc, int ns, float **tabled) { sinc_table(1, nsinc, ns, tabled); }

void transfer_source_func(int npts, int nt, int *locs, float *vals) {
  ok(brtm_set_source(ctx(), npts, nt, locs, vals), "transfer_source_func");
}

void transfer_receiver_func(int nx, int ny, int nt, int *locs, float *vals) {
  ok(brtm_set_receivers(ctx(), nx, ny, nt, locs, vals), "transfer_receiver_func");
  g_nrx = nx;
  g_nry = ny;
  g_ntr = nt;
}

void source_prop(int n1, int n2, int n3, bool damp, bool get_last, float *p0, float *p1, int jt,
                 int npts, int nt) {
  (void)n1;
  (void)n2;
  (void)n3;
  (void)damp;
  (void)npts;
  // the reference runs it = 0 .. nt (Born/gpu_funcs_3d.cu:227)
  ok(brtm_source_prop(ctx(), jt, nt + 1, jt, BRTM_STORE_NONE, 0, get_last ? p0 : nullptr,
                      get_last ? p1 : nullptr),
     "source_prop");
}

void rtm_forward(int n1, int n2, int n3, int jt, float *img, float *dat, int npts_src, int nt,
                 int nt_big, int nx, int ny) {
  (void)n1;
  (void)n2;
  (void)n3;
  (void)npts_src;
  if (nx != g_nrx || ny != g_nry || nt_big != g_ntr) {
    std::fprintf(stderr, "rtm_forward: receiver geometry differs from transfer_receiver_func\n");
    std::exit(-1);
  }
  brtm_ctx *c = ctx();
  ok(brtm_select_shot(c, 0), "rtm_forward");
  ok(brtm_release_store(c, 0), "rtm_forward");  // source propagated alongside
  ok(brtm_set_model(c, img), "rtm_forward");
  ok(brtm_rtm_forward(c, jt, nt, jt, 1), "rtm_forward");
  ok(brtm_get_receivers(c, dat), "rtm_forward");
}

void rtm_adjoint(int n1, int n2, int n3, int jt, float *p0_s_cpu, float *p1_s_cpu, float *img,
                 int npts, int nt) {
  (void)n1;
  (void)n2;
  (void)n3;
  (void)p0_s_cpu;
  (void)p1_s_cpu;
  (void)npts;
  brtm_ctx *c = ctx();
  const int ck = std::max(1, (int)std::ceil(std::sqrt((double)nt)));
  ok(brtm_select_shot(c, 0), "rtm_adjoint");
  ok(brtm_source_prop(c, jt, nt, jt, BRTM_STORE_CKPT, ck, nullptr, nullptr), "rtm_adjoint");
  ok(brtm_zero_image(c), "rtm_adjoint");
  ok(brtm_rtm_adjoint(c, jt, nt, jt), "rtm_adjoint");
  ok(brtm_get_image(c, img), "rtm_adjoint");
  ok(brtm_release_store(c, 0), "rtm_adjoint");
}

}  // extern "C"
