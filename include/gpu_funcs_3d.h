/*
 * gpu_funcs_3d.h -- the reference's device ABI, Born/gpu_funcs_3d.h:5-27,
 * provided by the MI355X library over one process-wide default context, so
 * the reference's host code (Born/rtm_zero_op_3d.cpp, Born/fd_prop_3d.cpp)
 * links against it unchanged.
 *
 * Semantics follow the library (born_rtm.h); differences from the
 * reference's CUDA implementation:
 *   - setup_cuda uses one GPU whatever ngpus says (0, the reference's
 *     default that divides by zero at Born/gpu_funcs_3d.cu:143, is accepted);
 *   - the sponge is applied to every wavefield inside the stencil step, and
 *     source_prop's damp flag (unused by the reference) is ignored;
 *   - rtm_adjoint is the exact transpose of rtm_forward: it recomputes the
 *     source wavefield from FP32 checkpoints instead of reversing it in time
 *     from the two final snapshots, so p0_s_cpu / p1_s_cpu are not read;
 *   - rtm_forward accumulates into the traces given to
 *     transfer_receiver_func (the reference's `+=` extraction) and returns
 *     them in dat; rtm_adjoint returns this shot's image in img;
 *   - transfer_vel_func2 (unused by the reference's kernels) is accepted and
 *     ignored;
 *   - errors print brtm_last_error() to stderr and exit(-1), as the
 *     reference's process_error does (Born/gpu_funcs_3d.cu:47-62).
 */
#ifndef GPU_FUNCS_3D_H
#define GPU_FUNCS_3D_H 1

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

void setup_cuda(int ngpus, int argc, char **argv);
void source_prop(int n1, int n2, int n3, bool damp, bool get_last, float *p0, float *p1, int jt,
                 int npts, int nt);
void rtm_adjoint(int n1, int n2, int n3, int jt, float *p0_s_cpu, float *p1_s_cpu, float *img,
                 int npts, int nt);
void rtm_forward(int n1, int n2, int n3, int jt, float *img, float *dat, int npts_src, int nt,
                 int nt_big, int nx, int ny);
void transfer_sinc_table_s(int nsinc, int ns, float **tables);
void transfer_sinc_table_d(int nsinc, int ns, float **tabled);
void transfer_source_func(int npts, int nt, int *locs, float *vals);
void transfer_receiver_func(int nx, int ny, int nt, int *locs, float *vals);
void transfer_vel_func1(int nx, int ny, int nz, float *vel);
void transfer_vel_func2(int nx, int ny, int nz, float *vel);
void create_gpu_space(float d1, float d2, float d3, float bc_a, float bc_b, float bc_b_y, int nx,
                      int ny, int nz);

#ifdef __cplusplus
}
#endif
#endif /* GPU_FUNCS_3D_H */
