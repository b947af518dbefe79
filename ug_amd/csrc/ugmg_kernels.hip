// CDNA4 (gfx950) kernels of the multigrid solve path.
//
// All matrix kernels read SELL-64 slabs: one wavefront owns 64 consecutive
// slots (rows) of a slice and walks the slice's entries column-major, so
// every value/column load of a wavefront is one contiguous 256/512-byte
// stretch; the x gathers of neighbouring rows of a stencil hit the same
// cache lines.  Blocks are 4 wavefronts; block ids are remapped so that each
// of the 8 XCDs (round-robin dispatch) walks one contiguous range of slices
// and keeps its share of x in its own L2.  No MFMA: these are sparse
// contractions at ~0.15 flop/byte, bound by HBM bandwidth.
#include <hip/hip_runtime.h>

#include "ugmg_kernels.h"

namespace ugmg {

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int nx = 8;
  const int xcd = bid % nx, idx = bid / nx;
  const int per = nb / nx, rem = nb % nx;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

template <typename T>
__device__ __forceinline__ T ntload(const T *p) {
  return __builtin_nontemporal_load(p);
}

template <typename T> struct Acc { using type = T; };

// ------------------------------------------------------------------- SpMV
// MODE: 0 y -= A x, 1 y = A x, 2 y -= A x and c += x, 3 y -= A x and c = x
template <typename TV, typename TX, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(DSell<TV> A, int slice0,
                                                  int nsl, const TX *__restrict__ x,
                                                  TX *__restrict__ y,
                                                  TX *__restrict__ c) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int s = slice0 + b * (kBlock / kSlice) + (threadIdx.x >> 6);
  if (s >= slice0 + nsl) return;
  const int lane = threadIdx.x & 63;
  const int r = A.rows[s * kSlice + lane];
  const int e0 = A.soff[s];
  const int w = (A.soff[s + 1] - e0) >> 6;
  const TV *v = A.val + e0 + lane;
  const int *ci = A.col + e0 + lane;
  TX sum = 0;
  int k = 0;
  for (; k + 4 <= w; k += 4) {
    const TV a0 = ntload(v), a1 = ntload(v + 64), a2 = ntload(v + 128), a3 = ntload(v + 192);
    const int c0 = ntload(ci), c1 = ntload(ci + 64), c2 = ntload(ci + 128), c3 = ntload(ci + 192);
    sum += (TX)a0 * x[c0];
    sum += (TX)a1 * x[c1];
    sum += (TX)a2 * x[c2];
    sum += (TX)a3 * x[c3];
    v += 256;
    ci += 256;
  }
  for (; k < w; ++k) {
    sum += (TX)ntload(v) * x[ntload(ci)];
    v += 64;
    ci += 64;
  }
  if (r < 0) return;
  if (MODE == 1) {
    y[r] = sum;
  } else {
    y[r] -= sum;
    if (MODE == 2) c[r] += x[r];
    if (MODE == 3) c[r] = x[r];
  }
}

// FP64 defect of the mixed-precision refinement step:
//   r -= A c32 (FP64 values, FP64 sums);  x += c32;  r32 = (float) r;
//   per-block partial sum of r^2 over NEW_DEFECT rows.
__global__ __launch_bounds__(kBlock) void k_ir_update(DSell<double> A, int nsl,
                                                       const float *__restrict__ c32,
                                                       double *__restrict__ r,
                                                       double *__restrict__ x,
                                                       float *__restrict__ r32,
                                                       const uint8_t *__restrict__ vclass,
                                                       double *__restrict__ part) {
  __shared__ double red[kBlock / 64];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int s = b * (kBlock / kSlice) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double sq = 0.0;
  if (s < nsl) {
    const int rr = A.rows[s * kSlice + lane];
    const int e0 = A.soff[s];
    const int w = (A.soff[s + 1] - e0) >> 6;
    const double *v = A.val + e0 + lane;
    const int *ci = A.col + e0 + lane;
    double sum = 0.0;
    for (int k = 0; k < w; ++k) sum += ntload(v + 64 * k) * (double)c32[ntload(ci + 64 * k)];
    if (rr >= 0) {
      double nr = r[rr] - sum;
      r[rr] = nr;
      x[rr] += (double)c32[rr];
      r32[rr] = (float)nr;
      if (vclass[rr] >= 2) sq = nr * nr;
    }
  }
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if (lane == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

// ----------------------------------------------------------------- Jacobi
// l_jac + dscalx: t = damp * b / diag, 0 for VCLASS < 3
template <typename T>
__global__ __launch_bounds__(kBlock) void k_jac(int n, const T *__restrict__ b,
                                                 const T *__restrict__ diag,
                                                 const uint8_t *__restrict__ vclass,
                                                 T damp, T *__restrict__ t) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  t[i] = vclass[i] < 3 ? (T)0 : (b[i] / diag[i]) * damp;
}

// ------------------------------------------------------ multicolour GS
// forward substitution of one colour: v_r = (b_r - sum_{lower} a v) / a_rr
template <typename TV, typename T>
__global__ __launch_bounds__(kBlock) void k_gs_color(DSell<TV> L, int slice0, int nsl,
                                                      const T *__restrict__ b,
                                                      const T *__restrict__ diag,
                                                      const uint8_t *__restrict__ vclass,
                                                      T *__restrict__ v) {
  const int bb = xcd_remap(blockIdx.x, gridDim.x);
  const int s = slice0 + bb * (kBlock / kSlice) + (threadIdx.x >> 6);
  if (s >= slice0 + nsl) return;
  const int lane = threadIdx.x & 63;
  const int r = L.rows[s * kSlice + lane];
  const int e0 = L.soff[s];
  const int w = (L.soff[s + 1] - e0) >> 6;
  const TV *a = L.val + e0 + lane;
  const int *ci = L.col + e0 + lane;
  T sum = 0;
  for (int k = 0; k < w; ++k) sum += (T)ntload(a + 64 * k) * v[ntload(ci + 64 * k)];
  if (r < 0) return;
  v[r] = vclass[r] < 3 ? (T)0 : (b[r] - sum) / diag[r];
}

// defect after the sweep, b -= A (damp v), using (D+L) v = b on active rows:
//   active   b <- b - damp (b + U v)
//   inactive b <- b - damp (A v)      (v_r = 0, U holds the whole row)
// and the correction: CMODE 0 c = damp v, 1 c += damp v; xout = damp v.
template <typename TV, typename T, int CMODE>
__global__ __launch_bounds__(kBlock) void k_gs_defect(DSell<TV> U, int nsl,
                                                       const T *__restrict__ v,
                                                       T *__restrict__ b,
                                                       const uint8_t *__restrict__ vclass,
                                                       T damp, T *__restrict__ c,
                                                       T *__restrict__ xout) {
  const int bb = xcd_remap(blockIdx.x, gridDim.x);
  const int s = bb * (kBlock / kSlice) + (threadIdx.x >> 6);
  if (s >= nsl) return;
  const int lane = threadIdx.x & 63;
  const int r = U.rows[s * kSlice + lane];
  const int e0 = U.soff[s];
  const int w = (U.soff[s + 1] - e0) >> 6;
  const TV *a = U.val + e0 + lane;
  const int *ci = U.col + e0 + lane;
  T sum = 0;
  for (int k = 0; k < w; ++k) sum += (T)ntload(a + 64 * k) * v[ntload(ci + 64 * k)];
  if (r < 0) return;
  const T br = b[r];
  b[r] = vclass[r] >= 3 ? br - damp * (br + sum) : br - damp * sum;
  const T t = damp * v[r];
  if (c) {
    if (CMODE == 0) c[r] = t;
    else c[r] += t;
  }
  if (xout) xout[r] = t;
}

// --------------------------------------------------------------- transfer
// StandardRestrictNodeVector as a gather over P^T rows
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(int nc, const int *__restrict__ ptr,
                                                      const int *__restrict__ idx,
                                                      const T *__restrict__ w,
                                                      const uint32_t *__restrict__ cskip,
                                                      const uint8_t *__restrict__ cvnclass,
                                                      const T *__restrict__ bf, T damp,
                                                      T *__restrict__ bc) {
  const int I = blockIdx.x * kBlock + threadIdx.x;
  if (I >= nc) return;
  T acc = cvnclass[I] >= 2 ? (T)0 : bc[I];
  if (!(cskip[I] & 1u))
    for (int q = ptr[I]; q < ptr[I + 1]; ++q) acc += w[q] * (damp * bf[idx[q]]);
  bc[I] = acc;
}

// StandardIntCorNodeVector (+ dadd): t = damp P c_c on non-Dirichlet rows
template <typename T, int CADD>
__global__ __launch_bounds__(kBlock) void k_interp(int nf, const int *__restrict__ ptr,
                                                    const int *__restrict__ idx,
                                                    const T *__restrict__ w,
                                                    const uint32_t *__restrict__ fskip,
                                                    const T *__restrict__ cc, T damp,
                                                    T *__restrict__ t, T *__restrict__ c) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= nf) return;
  T acc = 0;
  if (!(fskip[v] & 1u))
    for (int q = ptr[v]; q < ptr[v + 1]; ++q) acc += (w[q] * damp) * cc[idx[q]];
  t[v] = acc;
  if (CADD) c[v] += acc;
}

// -------------------------------------------------------------- reductions
template <typename T, int DOT>
__global__ __launch_bounds__(kBlock) void k_dot_part(int n, const T *__restrict__ x,
                                                      const T *__restrict__ y,
                                                      const uint8_t *__restrict__ vclass,
                                                      double *__restrict__ part) {
  __shared__ double red[kBlock / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (vclass && vclass[i] < 2) continue;
    const double xi = (double)x[i];
    s += DOT ? xi * (double)y[i] : xi * xi;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_reduce_final(int nparts, const double *__restrict__ part,
                                                        double *__restrict__ out, int do_sqrt) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 1024) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < 16; ++i) t += red[i];
    out[0] = do_sqrt ? sqrt(t) : t;
  }
}

// ------------------------------------------------------------------ BLAS-1
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int n, T *x, T val) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = val;
}
template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const TS *__restrict__ s, TD *__restrict__ d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) d[i] = (TD)s[i];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add(int n, T *__restrict__ x, const T *__restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] += y[i];
}

// ------------------------------------------------------------- base solve
// ls($m, $red, $abslimit) over ex on the base level, one workgroup:
//   first = |b|; converged = first < abslimit
//   repeat m: x = damp * inv b; b -= A x; c += x; test |b|
// The dense inverse equals the band LU solve of EXApplyLUDOUBLE; every
// product runs in FP64 whatever T is.
template <typename T, int CMODE>
__global__ __launch_bounds__(1024) void k_base_solve(BaseArgs ba, DSell<double> A,
                                                      const uint8_t *__restrict__ vclass,
                                                      T *__restrict__ b, T *__restrict__ c) {
  extern __shared__ double sh[];
  const int n = ba.n;
  double *bl = sh;          // n
  double *xl = sh + n;      // n
  double *cl = sh + 2 * n;  // n
  __shared__ double red[16];
  __shared__ int conv;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += 1024) {
    bl[i] = (double)b[i];
    cl[i] = CMODE ? (double)c[i] : 0.0;
  }
  __syncthreads();
  auto norm = [&]() -> double {
    double s = 0.0;
    for (int i = tid; i < n; i += 1024)
      if (vclass[i] >= 2) s += bl[i] * bl[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < 16; ++i) t += red[i];
    return sqrt(t);
  };
  const double first = norm();
  double target = first * ba.red;
  if (target == 0.0) target = ba.red;
  if (tid == 0) conv = fabs(first) < fabs(ba.abslimit);
  __syncthreads();
  for (int it = 0; it < ba.maxiter; ++it) {
    if (conv) break;
    if (ba.failed) break;
    for (int i = tid; i < n; i += 1024) {
      const double *row = ba.inv + (size_t)i * n;
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += row[j] * bl[j];
      xl[i] = s * ba.damp;
    }
    __syncthreads();
    const int nslot = A.nslices * kSlice;
    for (int slot = tid; slot < nslot; slot += 1024) {
      const int r = A.rows[slot];
      if (r < 0) continue;
      const int s = slot >> 6, lane = slot & 63;
      const int e0 = A.soff[s], w = (A.soff[s + 1] - e0) >> 6;
      double sum = 0.0;
      for (int k = 0; k < w; ++k) sum += A.val[e0 + 64 * k + lane] * xl[A.col[e0 + 64 * k + lane]];
      bl[r] -= sum;
    }
    __syncthreads();
    for (int i = tid; i < n; i += 1024) cl[i] += xl[i];
    const double last = norm();
    if (tid == 0) conv = fabs(last) < fabs(ba.abslimit) || fabs(last) < fabs(target);
    __syncthreads();
  }
  for (int i = tid; i < n; i += 1024) {
    b[i] = (T)bl[i];
    c[i] = (T)cl[i];
  }
}

// ----------------------------------------------------------------- launchers
static inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }
static inline int nblk_slices(int nsl) { return (nsl + kBlock / kSlice - 1) / (kBlock / kSlice); }

template <typename TV, typename TX>
void launch_spmv(hipStream_t st, const DSell<TV> &A, int slice0, int nsl, const TX *x,
                 TX *y, TX *c, int mode) {
  if (nsl <= 0) return;
  dim3 g(nblk_slices(nsl)), blk(kBlock);
  switch (mode) {
    case 0: k_spmv<TV, TX, 0><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c); break;
    case 1: k_spmv<TV, TX, 1><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c); break;
    case 2: k_spmv<TV, TX, 2><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c); break;
    default: k_spmv<TV, TX, 3><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c); break;
  }
}
template void launch_spmv<double, double>(hipStream_t, const DSell<double> &, int, int,
                                          const double *, double *, double *, int);
template void launch_spmv<float, float>(hipStream_t, const DSell<float> &, int, int,
                                        const float *, float *, float *, int);

void launch_ir_update(hipStream_t st, const DSell<double> &A, const float *c32, double *r,
                      double *x, float *r32, const uint8_t *vclass, double *part, int *nparts) {
  const int g = nblk_slices(A.nslices);
  *nparts = g;
  if (g == 0) return;
  k_ir_update<<<g, kBlock, 0, st>>>(A, A.nslices, c32, r, x, r32, vclass, part);
}

template <typename T>
void launch_jac(hipStream_t st, int n, const T *b, const T *diag, const uint8_t *vclass,
                T damp, T *t) {
  if (n <= 0) return;
  k_jac<T><<<nblk(n), kBlock, 0, st>>>(n, b, diag, vclass, damp, t);
}
template void launch_jac<double>(hipStream_t, int, const double *, const double *,
                                 const uint8_t *, double, double *);
template void launch_jac<float>(hipStream_t, int, const float *, const float *,
                                const uint8_t *, float, float *);

template <typename TV, typename T>
void launch_gs_color(hipStream_t st, const DSell<TV> &L, int slice0, int nsl, const T *b,
                     const T *diag, const uint8_t *vclass, T *v) {
  if (nsl <= 0) return;
  k_gs_color<TV, T><<<nblk_slices(nsl), kBlock, 0, st>>>(L, slice0, nsl, b, diag, vclass, v);
}
template void launch_gs_color<double, double>(hipStream_t, const DSell<double> &, int, int,
                                              const double *, const double *,
                                              const uint8_t *, double *);
template void launch_gs_color<float, float>(hipStream_t, const DSell<float> &, int, int,
                                            const float *, const float *, const uint8_t *,
                                            float *);

template <typename TV, typename T>
void launch_gs_defect(hipStream_t st, const DSell<TV> &U, const T *v, T *b,
                      const uint8_t *vclass, T damp, T *c, int cmode, T *xout) {
  if (U.nslices <= 0) return;
  dim3 g(nblk_slices(U.nslices));
  if (cmode == 0)
    k_gs_defect<TV, T, 0><<<g, kBlock, 0, st>>>(U, U.nslices, v, b, vclass, damp, c, xout);
  else
    k_gs_defect<TV, T, 1><<<g, kBlock, 0, st>>>(U, U.nslices, v, b, vclass, damp, c, xout);
}
template void launch_gs_defect<double, double>(hipStream_t, const DSell<double> &,
                                               const double *, double *, const uint8_t *,
                                               double, double *, int, double *);
template void launch_gs_defect<float, float>(hipStream_t, const DSell<float> &, const float *,
                                             float *, const uint8_t *, float, float *, int,
                                             float *);

template <typename T>
void launch_restrict(hipStream_t st, int nc, const int *ptr, const int *idx, const T *w,
                     const uint32_t *cskip, const uint8_t *cvnclass, const T *bf, T damp,
                     T *bc) {
  if (nc <= 0) return;
  k_restrict<T><<<nblk(nc), kBlock, 0, st>>>(nc, ptr, idx, w, cskip, cvnclass, bf, damp, bc);
}
template void launch_restrict<double>(hipStream_t, int, const int *, const int *,
                                      const double *, const uint32_t *, const uint8_t *,
                                      const double *, double, double *);
template void launch_restrict<float>(hipStream_t, int, const int *, const int *, const float *,
                                     const uint32_t *, const uint8_t *, const float *, float,
                                     float *);

template <typename T>
void launch_interp(hipStream_t st, int nf, const int *ptr, const int *idx, const T *w,
                   const uint32_t *fskip, const T *cc, T damp, T *t, T *c) {
  if (nf <= 0) return;
  if (c)
    k_interp<T, 1><<<nblk(nf), kBlock, 0, st>>>(nf, ptr, idx, w, fskip, cc, damp, t, c);
  else
    k_interp<T, 0><<<nblk(nf), kBlock, 0, st>>>(nf, ptr, idx, w, fskip, cc, damp, t, c);
}
template void launch_interp<double>(hipStream_t, int, const int *, const int *, const double *,
                                    const uint32_t *, const double *, double, double *,
                                    double *);
template void launch_interp<float>(hipStream_t, int, const int *, const int *, const float *,
                                   const uint32_t *, const float *, float, float *, float *);

int reduce_parts(int n) {
  int g = nblk(n);
  return g < 1 ? 1 : (g > 1024 ? 1024 : g);
}

template <typename T>
void launch_dot(hipStream_t st, int n, const T *x, const T *y, const uint8_t *vclass,
                double *part, double *out, int do_sqrt) {
  const int g = reduce_parts(n);
  if (y)
    k_dot_part<T, 1><<<g, kBlock, 0, st>>>(n, x, y, vclass, part);
  else
    k_dot_part<T, 0><<<g, kBlock, 0, st>>>(n, x, y, vclass, part);
  k_reduce_final<<<1, 1024, 0, st>>>(g, part, out, do_sqrt);
}
template void launch_dot<double>(hipStream_t, int, const double *, const double *,
                                 const uint8_t *, double *, double *, int);
template void launch_dot<float>(hipStream_t, int, const float *, const float *,
                                const uint8_t *, double *, double *, int);

void launch_reduce_final(hipStream_t st, int nparts, const double *part, double *out,
                         int do_sqrt) {
  k_reduce_final<<<1, 1024, 0, st>>>(nparts, part, out, do_sqrt);
}

template <typename T>
void launch_fill(hipStream_t st, int n, T *x, T val) {
  if (n > 0) k_fill<T><<<nblk(n), kBlock, 0, st>>>(n, x, val);
}
template void launch_fill<double>(hipStream_t, int, double *, double);
template void launch_fill<float>(hipStream_t, int, float *, float);

template <typename TD, typename TS>
void launch_convert(hipStream_t st, int n, const TS *s, TD *d) {
  if (n > 0) k_convert<TD, TS><<<nblk(n), kBlock, 0, st>>>(n, s, d);
}
template void launch_convert<float, double>(hipStream_t, int, const double *, float *);
template void launch_convert<double, float>(hipStream_t, int, const float *, double *);

template <typename T>
void launch_add(hipStream_t st, int n, T *x, const T *y) {
  if (n > 0) k_add<T><<<nblk(n), kBlock, 0, st>>>(n, x, y);
}
template void launch_add<double>(hipStream_t, int, double *, const double *);
template void launch_add<float>(hipStream_t, int, float *, const float *);

template <typename T>
void launch_base_solve(hipStream_t st, const BaseArgs &ba, const DSell<double> &A,
                       const uint8_t *vclass, T *b, T *c, int cmode) {
  const size_t shmem = 3 * (size_t)ba.n * sizeof(double);
  if (cmode)
    k_base_solve<T, 1><<<1, 1024, shmem, st>>>(ba, A, vclass, b, c);
  else
    k_base_solve<T, 0><<<1, 1024, shmem, st>>>(ba, A, vclass, b, c);
}
template void launch_base_solve<double>(hipStream_t, const BaseArgs &, const DSell<double> &,
                                        const uint8_t *, double *, double *, int);
template void launch_base_solve<float>(hipStream_t, const BaseArgs &, const DSell<double> &,
                                       const uint8_t *, float *, float *, int);

}  // namespace ugmg
