// Base-level set-up on the device: the band LU of EXPreProcess
// (EXDecomposeMatrixDOUBLE, gm/evm.cc:1400-1418, or EXDecomposeMatrixFLOAT,
// 1355-1373, for ex $f: no pivoting, a zero pivot fails) and the explicit
// inverse, one column per unit vector pushed through EXApplyLUDOUBLE
// (gm/evm.cc:1493-1510; EXApplyLUFLOAT, 1448-1466, is the same double
// arithmetic on the FLOAT factors).
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
// R = float: EXDecomposeMatrixFLOAT (gm/evm.cc:1355-1373) on a band of
// FLOAT values held as doubles: every operation in float arithmetic, the
// results stored back exactly.
template <typename R>
__global__ __launch_bounds__(256) void k_band_lu_step(double *__restrict__ band, int bw, int n,
                                                      int i, int *__restrict__ failed) {
#pragma clang fp contract(off)
  if (*failed) return;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n - 1) {
    const R d = (R)bandref(band, bw, i, i);
    if (d == (R)0) {
      if (idx == 0) *failed = 1;
      return;
    }
    const int jmax = min(i + bw, n - 1);
    const int m = jmax - i;
    if (idx < (long long)m * m) {
      const int j = i + 1 + (int)(idx / m);
      const int k = i + 1 + (int)(idx % m);
      const R f = (R)bandref(band, bw, j, i) / d;
      const R p = f * (R)bandref(band, bw, i, k);
      bandref(band, bw, j, k) = (double)((R)bandref(band, bw, j, k) - p);
    }
  }
  if (i > 0) {
    const int q = i - 1;
    const int mp = min(q + bw, n - 1) - q;
    if (idx < mp) {
      const int j = q + 1 + (int)idx;
      bandref(band, bw, j, q) = (double)((R)bandref(band, bw, j, q) / (R)bandref(band, bw, q, q));
    }
  }
}

// acc - sum_{k0 <= k < k1} row[k] x[k ldx], subtracting term by term in
// ascending k (each product rounded, then the difference), in chunks of
// 64 terms.  The band row is wave-uniform: lane j loads row[k + j] in one
// coalesced load and each term takes its factor with readlane, while the
// 64 column entries of a chunk are loaded before any of them is used, so a
// chunk costs one memory round trip; only the subtractions form the
// dependent chain.
template <bool TAIL>
__device__ __forceinline__ double band_chunk_sub(int llo, int lhi, const double *__restrict__ x,
                                                 size_t ldx, int k, int k1, double acc) {
#pragma clang fp contract(off)
  constexpr int kCh = 64;
  double xv[kCh];
#pragma unroll
  for (int j = 0; j < kCh; ++j) xv[j] = x[(size_t)(TAIL ? min(k + j, k1 - 1) : k + j) * ldx];
  // keep the scheduler from sinking each load next to its use
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int j = 0; j < kCh; ++j) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane(llo, j);
    const long long hi = __builtin_amdgcn_readlane(lhi, j);
    const double lj = __longlong_as_double((hi << 32) | lo);
    const double p = lj * xv[j];
    const double t = acc - p;
    // terms past k1 (last chunk only) are discarded by a select, not a
    // branch, which would let the compiler sink each load into its block
    acc = (!TAIL || k + j < k1) ? t : acc;
  }
  return acc;
}

__device__ __forceinline__ double band_row_sub(const double *__restrict__ row,
                                               const double *__restrict__ x, size_t ldx, int k0,
                                               int k1, double acc) {
  constexpr int kCh = 64;
  const int lane = threadIdx.x & 63;
  int k = k0;
  for (; k + kCh <= k1; k += kCh) {
    const long long lb = __double_as_longlong(row[k + lane]);
    acc = band_chunk_sub<false>((int)lb, (int)(lb >> 32), x, ldx, k, k1, acc);
  }
  if (k < k1) {
    const long long lb = __double_as_longlong(row[min(k + lane, k1 - 1)]);
    acc = band_chunk_sub<true>((int)lb, (int)(lb >> 32), x, ldx, k, k1, acc);
  }
  return acc;
}

// Columns c0 .. c0 + 63 of the inverse in band positions, one lane per
// column c (the unit vector e_c).  The 64 columns of one wave are stored
// row by row in one contiguous panel (entry (k, c) at
// X[(c / 64) n 64 + 64 k + c % 64]), so the rows a substitution step reads
// are a few consecutive pages rather than one page per row.  Forward
// substitution from row c0 + 1 (rows below a lane's own c hold zeros, so
// the extra terms of lanes with c > c0 subtract zeros and change nothing),
// then backward substitution over every row, as EXApplyLUDOUBLE.
__global__ __launch_bounds__(64) void k_band_inverse(const double *__restrict__ band, int bw, int n,
                                                     double *__restrict__ X,
                                                     const int *__restrict__ failed) {
#pragma clang fp contract(off)
  if (*failed) return;
  const int c0 = blockIdx.x * 64;
  const int c = c0 + (int)threadIdx.x;
  const size_t ldx = 64;
  double *x = X + (size_t)blockIdx.x * n * 64 + threadIdx.x;
  for (int k = 0; k < n; ++k) x[(size_t)k * ldx] = k == c ? 1.0 : 0.0;
  for (int i = c0 + 1; i < n; ++i)
    x[(size_t)i * ldx] = band_row_sub(band + 2 * (size_t)bw * i, x, ldx, max(i - bw, c0), i,
                                      x[(size_t)i * ldx]);
  for (int i = n - 1; i >= 0; --i) {
    const double *__restrict__ Ui = band + 2 * (size_t)bw * i;
    const double acc = band_row_sub(Ui, x, ldx, i + 1, min(i + bw, n - 1) + 1, x[(size_t)i * ldx]);
    x[(size_t)i * ldx] = acc / Ui[i];
  }
}

// inv[r][u] = entry (pos[r], pos[u]) of the panels (list order rows and
// columns)
__global__ __launch_bounds__(256) void k_band_inverse_out(const double *__restrict__ X,
                                                          const int *__restrict__ pos, int n,
                                                          double *__restrict__ inv,
                                                          const int *__restrict__ failed) {
  if (*failed) return;
  const int u = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (u >= n) return;
  const int q = pos[u];
  inv[(size_t)r * n + u] = X[(size_t)(q >> 6) * n * 64 + (size_t)pos[r] * 64 + (q & 63)];
}

}  // namespace

void launch_base_factor(hipStream_t st, double *band, int n, int bw, const int *pos, double *X,
                        double *inv, int *failed, bool fmode) {
  if (n <= 0) return;
  for (int i = 0; i < n; ++i) {
    const int m = i < n - 1 ? std::min(i + bw, n - 1) - i : 0;
    const long long work = std::max<long long>({(long long)m * m, (long long)bw, 1LL});
    const unsigned nb = (unsigned)((work + 255) / 256);
    if (fmode) k_band_lu_step<float><<<nb, 256, 0, st>>>(band, bw, n, i, failed);
    else k_band_lu_step<double><<<nb, 256, 0, st>>>(band, bw, n, i, failed);
  }
  k_band_inverse<<<(n + 63) / 64, 64, 0, st>>>(band, bw, n, X, failed);
  k_band_inverse_out<<<dim3((n + 255) / 256, n), 256, 0, st>>>(X, pos, n, inv, failed);
}

}  // namespace ugmg
