// Launchers of the gfx950 kernels (ugmg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ugmg_common.h"

namespace ugmg {

// device view of a SELL-64 layout (see ugmg_internal.h)
template <typename TV>
struct DSell {
  int nslices;
  const int *soff;
  const int *rows;
  const int *col;
  const TV *val;
};

struct BaseArgs {
  int n;
  int maxiter;
  int failed;
  double damp;
  double abslimit;
  double red;
  const double *inv;
};

template <typename TV, typename TX>
void launch_spmv(hipStream_t st, const DSell<TV> &A, int slice0, int nsl, const TX *x,
                 TX *y, TX *c, int mode);
void launch_ir_update(hipStream_t st, const DSell<double> &A, const float *c32, double *r,
                      double *x, float *r32, const uint8_t *vclass, double *part, int *nparts);
template <typename T>
void launch_jac(hipStream_t st, int n, const T *b, const T *diag, const uint8_t *vclass,
                T damp, T *t);
template <typename TV, typename T>
void launch_gs_color(hipStream_t st, const DSell<TV> &L, int slice0, int nsl, const T *b,
                     const T *diag, const uint8_t *vclass, T *v);
template <typename TV, typename T>
void launch_gs_defect(hipStream_t st, const DSell<TV> &U, const T *v, T *b,
                      const uint8_t *vclass, T damp, T *c, int cmode, T *xout);
template <typename T>
void launch_restrict(hipStream_t st, int nc, const int *ptr, const int *idx, const T *w,
                     const uint32_t *cskip, const uint8_t *cvnclass, const T *bf, T damp,
                     T *bc);
template <typename T>
void launch_interp(hipStream_t st, int nf, const int *ptr, const int *idx, const T *w,
                   const uint32_t *fskip, const T *cc, T damp, T *t, T *c);
int reduce_parts(int n);
template <typename T>
void launch_dot(hipStream_t st, int n, const T *x, const T *y, const uint8_t *vclass,
                double *part, double *out, int do_sqrt);
void launch_reduce_final(hipStream_t st, int nparts, const double *part, double *out,
                         int do_sqrt);
template <typename T>
void launch_fill(hipStream_t st, int n, T *x, T val);
template <typename TD, typename TS>
void launch_convert(hipStream_t st, int n, const TS *s, TD *d);
template <typename T>
void launch_add(hipStream_t st, int n, T *x, const T *y);
template <typename T>
void launch_base_solve(hipStream_t st, const BaseArgs &ba, const DSell<double> &A,
                       const uint8_t *vclass, T *b, T *c, int cmode);

}  // namespace ugmg
