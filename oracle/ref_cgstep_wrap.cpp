// C wrapper (our own code) that drives the reference's Born/cgstep.cpp on a
// small dense least-squares problem, compiled from the reference sources in
// place by oracle/Makefile into oracle/_ref/libref_cgstep.so.  Test
// infrastructure only: it pins the product's CG step.
//
// The vector type below implements the reference's abstract my_vector
// interface (Born/my_vector.h:10-41) over a float array with FP64 dot
// products, the semantics of Born/hypercube_float.cpp:123-132,212-223.  The
// iteration loop restates Born/lin_solver.cpp:115-143 (simple_solver with
// m0 = 0, Born/lin_solver.cpp:79-101: rr = -d).
#include <cstring>
#include <vector>
#include "cgstep.h"

namespace {
class arr_vec : public my_vector {
 public:
  explicit arr_vec(size_t n) : v(n, 0.f) { name = "arr_vec"; }
  std::vector<float> v;
  my_vector *clone_vec() override { arr_vec *o = new arr_vec(v.size()); o->v = v; return o; }
  my_vector *clone_space() override { return new arr_vec(v.size()); }
  void set_val(double val) override { for (auto &x : v) x = (float)val; }
  void scale(const double s) override { for (auto &x : v) x = (float)(x * s); }
  void scale_add(const double me, my_vector *o, const double oth) override {
    arr_vec *b = (arr_vec *)o;
    for (size_t i = 0; i < v.size(); i++) v[i] = (float)(v[i] * me + oth * b->v[i]);
  }
  void add(my_vector *o) override {
    arr_vec *b = (arr_vec *)o;
    for (size_t i = 0; i < v.size(); i++) v[i] += b->v[i];
  }
  double dot(my_vector *o) override {
    arr_vec *b = (arr_vec *)o;
    double s = 0;
    for (size_t i = 0; i < v.size(); i++) s += (double)v[i] * (double)b->v[i];
    return s;
  }
  void info(char *, int) override {}
};
}  // namespace

// A is row-major [nrow][ncol]; solves min |A m - d| with niter cgstep
// iterations.  resid[i] = rr.rr after iteration i.  Returns iterations done.
extern "C" int ref_cgstep_solve(int nrow, int ncol, const float *A,
                                const float *d, int niter, float *m_out,
                                double *resid) {
  arr_vec m(ncol), g(ncol), rr(nrow), gg(nrow);
  for (int i = 0; i < nrow; i++) rr.v[i] = -d[i];
  cgstep st;
  st.alloc_step(&m, &rr);
  int it = 0;
  for (; it < niter; it++) {
    for (int j = 0; j < ncol; j++) {
      float s = 0;
      for (int i = 0; i < nrow; i++) s += A[(size_t)i * ncol + j] * rr.v[i];
      g.v[j] = s;
    }
    for (int i = 0; i < nrow; i++) {
      float s = 0;
      for (int j = 0; j < ncol; j++) s += A[(size_t)i * ncol + j] * g.v[j];
      gg.v[i] = s;
    }
    double val = 0;
    if (!st.step_it(it, &m, &g, &rr, &gg, &val)) break;
    resid[it] = val;
  }
  std::memcpy(m_out, m.v.data(), sizeof(float) * ncol);
  return it;
}
