// Base-level set-up on the device: the band LU of EXPreProcess
// (EXDecomposeMatrixDOUBLE, gm/evm.cc:1400-1418: no pivoting, a zero pivot
// fails) and the explicit inverse, one column per unit vector pushed through
// EXApplyLUDOUBLE (gm/evm.cc:1493-1510).
//
// Every value is formed by the same operations in the same order as the host
// loops of build_base, with products and differences rounded separately (no
// contraction into fused multiply-adds), so the factors and the inverse are
// bit-identical to the host's.
//
// Band storage is UG's EX_MAT (gm/evm.h:392): entry (i, j) at
// band[2 bw i + j], |i - j| <= bw, rows and columns in band positions (BFS
// order, b consecutive scalar unknowns per vector).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ugmg_kernels.h"

namespace ugmg {

namespace {

__device__ __forceinline__ double &bandref(double *m, int bw, int i, int j) {
  return m[2 * (size_t)bw * i + j];
}

// Pivot step i (i < n - 1): a_jk <- a_jk - (a_ji / a_ii) a_ik for j, k in
// (i, min(i + bw, n - 1)].  The same launch stores the multipliers of step
// i - 1 into column i - 1, which no thread of step i reads.  Step n - 1 only
// stores the last multipliers.  A zero pivot sets *failed and leaves the
// band as it is (the host loop stops there).
__global__ __launch_bounds__(256) void k_band_lu_step(double *__restrict__ band, int bw, int n,
                                                      int i, int *__restrict__ failed) {
#pragma clang fp contract(off)
  if (*failed) return;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n - 1) {
    const double d = bandref(band, bw, i, i);
    if (d == 0.0) {
      if (idx == 0) *failed = 1;
      return;
    }
    const int jmax = min(i + bw, n - 1);
    const int m = jmax - i;
    if (idx < (long long)m * m) {
      const int j = i + 1 + (int)(idx / m);
      const int k = i + 1 + (int)(idx % m);
      const double f = bandref(band, bw, j, i) / d;
      const double p = f * bandref(band, bw, i, k);
      bandref(band, bw, j, k) = bandref(band, bw, j, k) - p;
    }
  }
  if (i > 0) {
    const int q = i - 1;
    const int mp = min(q + bw, n - 1) - q;
    if (idx < mp) {
      const int j = q + 1 + (int)idx;
      bandref(band, bw, j, q) = bandref(band, bw, j, q) / bandref(band, bw, q, q);
    }
  }
}

// Columns c0 .. c0 + 63 of the inverse in band positions, one lane per
// column c (the unit vector e_c): X[k][c], row stride ldx.  Forward
// substitution from row c0 + 1 (rows below a lane's own c hold zeros, so
// the extra terms of lanes with c > c0 subtract zeros and change nothing),
// then backward substitution over every row, as EXApplyLUDOUBLE.
__global__ __launch_bounds__(64) void k_band_inverse(const double *__restrict__ band, int bw, int n,
                                                     double *__restrict__ X, size_t ldx,
                                                     const int *__restrict__ failed) {
#pragma clang fp contract(off)
  if (*failed) return;
  const int c0 = blockIdx.x * 64;
  const int c = c0 + (int)threadIdx.x;
  double *x = X + c;   // X has ldx >= c0 + 64 columns
  for (int k = 0; k < n; ++k) x[(size_t)k * ldx] = k == c ? 1.0 : 0.0;
  for (int i = c0 + 1; i < n; ++i) {
    const double *__restrict__ Li = band + 2 * (size_t)bw * i;
    double acc = x[(size_t)i * ldx];
    for (int k = max(i - bw, c0); k < i; ++k) {
      const double p = Li[k] * x[(size_t)k * ldx];
      acc = acc - p;
    }
    x[(size_t)i * ldx] = acc;
  }
  for (int i = n - 1; i >= 0; --i) {
    const double *__restrict__ Ui = band + 2 * (size_t)bw * i;
    const int kmax = min(i + bw, n - 1);
    double acc = x[(size_t)i * ldx];
    for (int k = i + 1; k <= kmax; ++k) {
      const double p = Ui[k] * x[(size_t)k * ldx];
      acc = acc - p;
    }
    x[(size_t)i * ldx] = acc / Ui[i];
  }
}

// inv[r][u] = X[pos[r]][pos[u]] (list order rows and columns)
__global__ __launch_bounds__(256) void k_band_inverse_out(const double *__restrict__ X, size_t ldx,
                                                          const int *__restrict__ pos, int n,
                                                          double *__restrict__ inv,
                                                          const int *__restrict__ failed) {
  if (*failed) return;
  const int u = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (u >= n) return;
  inv[(size_t)r * n + u] = X[(size_t)pos[r] * ldx + pos[u]];
}

}  // namespace

void launch_base_factor(hipStream_t st, double *band, int n, int bw, const int *pos, double *X,
                        size_t ldx, double *inv, int *failed) {
  if (n <= 0) return;
  for (int i = 0; i < n; ++i) {
    const int m = i < n - 1 ? std::min(i + bw, n - 1) - i : 0;
    const long long work = std::max<long long>({(long long)m * m, (long long)bw, 1LL});
    const unsigned nb = (unsigned)((work + 255) / 256);
    k_band_lu_step<<<nb, 256, 0, st>>>(band, bw, n, i, failed);
  }
  k_band_inverse<<<(n + 63) / 64, 64, 0, st>>>(band, bw, n, X, ldx, failed);
  k_band_inverse_out<<<dim3((n + 255) / 256, n), 256, 0, st>>>(X, ldx, pos, n, inv, failed);
}

}  // namespace ugmg
