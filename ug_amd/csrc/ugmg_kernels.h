// Launchers and device-side descriptors of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ugmg_common.h"

namespace ugmg {

constexpr int kMaxLevels = 32;
constexpr int kMaxColors = 32;
constexpr int kCoarseThreads = 1024;

// device view of a SELL-64 layout (see ugmg_internal.h)
template <typename TV>
struct DSell {
  int nslices;
  const int *soff;
  const int *rows;
  const int *col;
  const TV *val;
  const int *sdoff;   // per slice: offset into dtab (uniform slice) or -1
  const int *dtab;    // shared column-offset lists of uniform slices
};

// base-level solver settings: ls($m,$red,$abslimit) over ex($damp)
struct BaseArgs {
  int n;
  int maxiter;
  int failed;
  double damp;
  double abslimit;
  double red;
  const double *inv;   // n x n, row-major
};

// one level as seen by the fused coarse-cycle kernel
template <typename T>
struct CLevel {
  int n;
  int ncolors;
  int color_slices[kMaxColors + 1];
  DSell<T> A, L, U, P, PT;   // P: interpolation from l-1, PT: restriction to l-1
  const T *diag;
  const uint8_t *vclass, *vnclass;
  const uint32_t *skip;
  T *c, *b, *t;
};

struct CycleParams {
  int nu1, nu2, gamma, smoother;
  double damp, lmgc_damp;
};

template <typename T>
struct CoarseArgs {
  const CLevel<T> *lv;   // device array indexed by level
  int bl, top;
  CycleParams cp;
  BaseArgs base;
  DSell<double> baseA;   // FP64 copy of the base-level matrix
  const uint8_t *base_vclass;
  T *c_top, *b_top;
  int cmode_top;         // 0: c := correction, 1: c += correction
};

template <typename TV, typename TX>
void launch_spmv(hipStream_t st, const DSell<TV> &A, int slice0, int nsl, const TX *x, TX *y,
                 TX *c, int mode, bool nt);
void launch_ir_update(hipStream_t st, const DSell<double> &A, const float *c32, double *r,
                      double *x, float *r32, const uint8_t *vclass, double *part, int *nparts,
                      bool nt);
template <typename T>
void launch_jac(hipStream_t st, int n, const T *b, const T *diag, const uint8_t *vclass, T damp,
                T *t);
template <typename TV, typename T>
void launch_gs_color(hipStream_t st, const DSell<TV> &L, int slice0, int nsl, const T *b,
                     const T *diag, const uint8_t *vclass, T *v, bool nt);
template <typename TV, typename T>
void launch_gs_defect(hipStream_t st, const DSell<TV> &U, const T *v, T *b, const uint8_t *vclass,
                      T damp, T *c, int cmode, T *xout, bool nt);
// returns the number of blocks (= partial sums written when part != null)
template <typename T>
int launch_restrict(hipStream_t st, const DSell<T> &PT, const uint32_t *cskip,
                    const uint8_t *cvnclass, const uint8_t *cvclass, const T *bf, T damp, T *bc,
                    double *part, bool nt);
template <typename T>
void launch_base_gemv(hipStream_t st, int n, const double *inv, const T *b, T *c, T *xout,
                      int cmode, double damp, double abslimit, double red, const double *part,
                      int nparts, double *first, int it);
template <typename T>
int launch_dot_part(hipStream_t st, int n, const T *x, const uint8_t *vclass, double *part);
template <typename T>
void launch_interp(hipStream_t st, const DSell<T> &P, const uint32_t *fskip, const T *cc, T damp,
                   T *t, T *c, bool nt);
int reduce_parts(int n);
template <typename T>
void launch_dot(hipStream_t st, int n, const T *x, const T *y, const uint8_t *vclass,
                double *part, double *out, int do_sqrt);
void launch_reduce_final(hipStream_t st, int nparts, const double *part, double *out, int do_sqrt);
template <typename T>
void launch_fill(hipStream_t st, int n, T *x, T val);
template <typename TD, typename TS>
void launch_convert(hipStream_t st, int n, const TS *s, TD *d);
template <typename T>
void launch_add(hipStream_t st, int n, T *x, const T *y);
template <typename T>
void launch_coarse_cycle(hipStream_t st, const CoarseArgs<T> &args, int base_n);

}  // namespace ugmg
