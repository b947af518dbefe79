// Device-side assembly of one level from element lists: the application's
// element loop with AddVlistMValues / AddVlistVValues
// (np/udm/disctools.cc:536-555, 841-950) followed by
// AssembleDirichletBoundary (disctools.cc:1837-1875), producing the
// flattened level that ugmg_set_level takes (ConvertMatrix layout,
// disctools.cc:2061-2128: each row starts with its diagonal block).
//
// Pattern: vector v couples with every vector sharing an element with it
// (ElementElementCreateConnection, gm/algebra.cc:2409-2475); after the
// diagonal, a row lists its columns in ascending order.  Values: every
// entry is 0 plus the element contributions in ascending element order
// (each a separately rounded product scale * kref when the elements share
// one reference matrix), then the Dirichlet rows: for every VECSKIP bit j of
// v, row j of every block of the row is zeroed, the diagonal (j, j) set to
// 1 and rhs_j to the Dirichlet value.
//
// One thread owns one row: it sorts its vector's incidences (element,
// corner) by element, merges the corner lists of those elements into the
// row's sorted column list, and accumulates the element blocks in order,
// so no atomics touch a value.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>
#include <vector>

#include "ugmg_kernels.h"

namespace ugmg {

namespace {

constexpr int kAsmMaxInc = 96;   // elements at one vector
constexpr int kAsmMaxRow = 96;   // vectors in one matrix row

enum { kErrIndex = 1, kErrInc = 2, kErrRow = 4 };

__global__ __launch_bounds__(256) void k_asm_check(const int *__restrict__ elem, long long total,
                                                   int n, int *__restrict__ flag) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= total) return;
  const int v = elem[q];
  if (v < 0 || v >= n) atomicOr(flag, kErrIndex);
}

__global__ __launch_bounds__(256) void k_asm_inc_count(const int *__restrict__ elem, long long total,
                                                       int *__restrict__ icnt) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= total) return;
  atomicAdd(&icnt[elem[q]], 1);
}

// incidence q = e * nc + i of corner i of element e, in arbitrary order
__global__ __launch_bounds__(256) void k_asm_inc_fill(const int *__restrict__ elem, long long total,
                                                      const int *__restrict__ iptr,
                                                      int *__restrict__ icur, int *__restrict__ inc) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= total) return;
  const int v = elem[q];
  const int p = atomicAdd(&icur[v], 1);
  inc[iptr[v] + p] = (int)q;
}

// Sorts v's incidences ascending (element order) in place and builds the
// sorted list of the vectors sharing an element with v (v included).
// Returns the list length, or sets err.
__device__ int asm_row_list(int v, const int *__restrict__ iptr, int *__restrict__ inc,
                            const int *__restrict__ elem, int nc, int *cand, int &err) {
  const int q0 = iptr[v], q1 = iptr[v + 1];
  if (q1 - q0 > kAsmMaxInc) {
    err |= kErrInc;
    return 0;
  }
  for (int q = q0 + 1; q < q1; ++q) {   // insertion sort (a few entries)
    const int t = inc[q];
    int p = q - 1;
    while (p >= q0 && inc[p] > t) {
      inc[p + 1] = inc[p];
      --p;
    }
    inc[p + 1] = t;
  }
  int m = 1;
  cand[0] = v;
  for (int q = q0; q < q1; ++q) {
    const int e = inc[q] / nc;
    for (int j = 0; j < nc; ++j) {
      const int w = elem[(size_t)e * nc + j];
      int p = m;
      while (p > 0 && cand[p - 1] > w) --p;
      if (p > 0 && cand[p - 1] == w) continue;
      if (m == kAsmMaxRow) {
        err |= kErrRow;
        return 0;
      }
      for (int t = m; t > p; --t) cand[t] = cand[t - 1];
      cand[p] = w;
      ++m;
    }
  }
  return m;
}

__global__ __launch_bounds__(256) void k_asm_row_count(int n, const int *__restrict__ iptr,
                                                       int *__restrict__ inc,
                                                       const int *__restrict__ elem, int nc,
                                                       int *__restrict__ cnt,
                                                       int *__restrict__ flag) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  int cand[kAsmMaxRow];
  int err = 0;
  const int m = asm_row_list(v, iptr, inc, elem, nc, cand, err);
  if (err) atomicOr(flag, err);
  cnt[v] = err ? 1 : m;
}

__global__ __launch_bounds__(256) void k_asm_row_fill(AsmInput in, const int *__restrict__ iptr,
                                                      int *__restrict__ inc,
                                                      const int *__restrict__ rp,
                                                      int *__restrict__ ci,
                                                      double *__restrict__ vals,
                                                      double *__restrict__ rhs) {
#pragma clang fp contract(off)
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= in.n) return;
  const int nc = in.nc, b = in.b, bb = b * b, mrow = nc * b;
  int cand[kAsmMaxRow];
  int err = 0;
  const int m = asm_row_list(v, iptr, inc, in.elem, nc, cand, err);
  if (err) return;   // reported by k_asm_row_count
  int vpos = 0;
  while (cand[vpos] != v) ++vpos;
  const int r0 = rp[v];
  ci[r0] = v;
  for (int t = 0, k = 1; t < m; ++t)
    if (t != vpos) ci[r0 + k++] = cand[t];
  double *__restrict__ row = vals + (size_t)r0 * bb;
  for (int t = 0; t < m * bb; ++t) row[t] = 0.0;
  double fv[3] = {0.0, 0.0, 0.0};
  for (int q = iptr[v]; q < iptr[v + 1]; ++q) {
    const int e = inc[q] / nc, i = inc[q] % nc;
    const double *__restrict__ K =
        in.emat ? in.emat + (size_t)e * mrow * mrow : in.kref;
    const double s = in.emat ? 1.0 : in.scale[e];
    for (int j = 0; j < nc; ++j) {
      const int w = in.elem[(size_t)e * nc + j];
      int lo = 0, hi = m - 1;   // w is in cand
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cand[mid] < w) lo = mid + 1;
        else hi = mid;
      }
      const int pos = lo == vpos ? 0 : (lo < vpos ? lo + 1 : lo);
      double *__restrict__ blk = row + (size_t)pos * bb;
      for (int k = 0; k < b; ++k)
        for (int l = 0; l < b; ++l) {
          const double a = K[(size_t)(i * b + k) * mrow + j * b + l];
          if (in.emat) {
            blk[k * b + l] = blk[k * b + l] + a;
          } else {
            const double p = s * a;
            blk[k * b + l] = blk[k * b + l] + p;
          }
        }
    }
    if (in.evec)
      for (int k = 0; k < b; ++k) fv[k] = fv[k] + in.evec[((size_t)e * nc + i) * b + k];
  }
  const uint32_t sk = in.skip ? in.skip[v] : 0u;
  for (int j = 0; j < b; ++j) {
    if (!((sk >> j) & 1u)) continue;
    for (int t = 0; t < m; ++t)
      for (int c = 0; c < b; ++c) row[(size_t)t * bb + j * b + c] = 0.0;
    row[j * b + j] = 1.0;
    fv[j] = in.dval ? in.dval[(size_t)v * b + j] : 0.0;
  }
  if (rhs)
    for (int k = 0; k < b; ++k) rhs[(size_t)v * b + k] = fv[k];
}

template <typename T>
hipError_t dmalloc(T **p, size_t n) {
  return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

inline unsigned nblk(long long n) { return (unsigned)std::max<long long>(1, (n + 255) / 256); }

}  // namespace

int assemble_device(hipStream_t st, const AsmInput &in, double *rhs, std::vector<int> &rp_h,
                    hvec<int> &ci_h, hvec<double> &vals_h, std::string &err, AsmKeep *keep) {
  const int n = in.n;
  const long long total = (long long)in.nelem * in.nc;
  const int bb = in.b * in.b;
  int *icnt = nullptr, *iptr = nullptr, *icur = nullptr, *inc = nullptr, *cnt = nullptr,
      *rp = nullptr, *ci = nullptr, *flag = nullptr;
  double *vals = nullptr;
  void *tmp = nullptr;
  auto release = [&] {
    for (void *p : {(void *)icnt, (void *)iptr, (void *)icur, (void *)inc, (void *)cnt, (void *)rp,
                    (void *)ci, (void *)flag, (void *)vals, tmp})
      hipFree(p);
  };
  auto hip_fail = [&](hipError_t e, const char *what) {
    release();
    err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? -4 : -2;
  };
  hipError_t e = hipSuccess;
  if ((e = dmalloc(&icnt, n + 1)) || (e = dmalloc(&iptr, n + 1)) || (e = dmalloc(&icur, n)) ||
      (e = dmalloc(&inc, total)) || (e = dmalloc(&cnt, n + 1)) || (e = dmalloc(&rp, n + 1)) ||
      (e = dmalloc(&flag, 1)))
    return hip_fail(e, "hipMalloc (assembly)");
  size_t tmp_bytes = 0;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, icnt, iptr, n + 1, st)) ||
      (e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1))))
    return hip_fail(e, "scan set-up");
  if ((e = hipMemsetAsync(flag, 0, sizeof(int), st)) ||
      (e = hipMemsetAsync(icnt, 0, sizeof(int) * (n + 1), st)) ||
      (e = hipMemsetAsync(icur, 0, sizeof(int) * std::max(n, 1), st)) ||
      (e = hipMemsetAsync(cnt, 0, sizeof(int) * (n + 1), st)))
    return hip_fail(e, "hipMemsetAsync");
  // corner indices first: nothing below indexes with an unchecked value
  k_asm_check<<<nblk(total), 256, 0, st>>>(in.elem, total, n, flag);
  int hflag = 0;
  if ((e = hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st)) ||
      (e = hipStreamSynchronize(st)))
    return hip_fail(e, "assembly check");
  if (hflag) {
    release();
    err = "element corner index out of range";
    return -1;
  }
  k_asm_inc_count<<<nblk(total), 256, 0, st>>>(in.elem, total, icnt);
  if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, icnt, iptr, n + 1, st)))
    return hip_fail(e, "scan");
  k_asm_inc_fill<<<nblk(total), 256, 0, st>>>(in.elem, total, iptr, icur, inc);
  k_asm_row_count<<<nblk(n), 256, 0, st>>>(n, iptr, inc, in.elem, in.nc, cnt, flag);
  if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, rp, n + 1, st)))
    return hip_fail(e, "scan");
  int nnzb = 0;
  if ((e = hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st)) ||
      (e = hipMemcpyAsync(&nnzb, rp + n, sizeof(int), hipMemcpyDeviceToHost, st)) ||
      (e = hipStreamSynchronize(st)))
    return hip_fail(e, "assembly pattern");
  if (hflag) {
    release();
    err = (hflag & kErrInc) ? "more than 96 elements meet at one vector"
                            : "a matrix row couples more than 96 vectors";
    return -1;
  }
  if ((e = dmalloc(&ci, nnzb)) || (e = dmalloc(&vals, (size_t)nnzb * bb)))
    return hip_fail(e, "hipMalloc (assembly)");
  k_asm_row_fill<<<nblk(n), 256, 0, st>>>(in, iptr, inc, rp, ci, vals, rhs);
  if ((e = hipGetLastError())) return hip_fail(e, "assembly kernels");
  try {
    rp_h.resize((size_t)n + 1);
    ci_h.resize(nnzb);
    vals_h.resize((size_t)nnzb * bb);
  } catch (const std::bad_alloc &) {
    release();
    err = "host allocation (assembly)";
    return -4;
  }
  if ((e = hipMemcpyAsync(rp_h.data(), rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st)) ||
      (e = hipMemcpyAsync(ci_h.data(), ci, sizeof(int) * (size_t)nnzb, hipMemcpyDeviceToHost, st)) ||
      (e = hipMemcpyAsync(vals_h.data(), vals, sizeof(double) * (size_t)nnzb * bb,
                          hipMemcpyDeviceToHost, st)) ||
      (e = hipStreamSynchronize(st)))
    return hip_fail(e, "assembly copy-out");
  if (keep) {   // the flattened arrays stay on the device for the layout build
    keep->rp = rp;
    keep->ci = ci;
    keep->vals = vals;
    rp = ci = nullptr;
    vals = nullptr;
  }
  release();
  return 0;
}

}  // namespace ugmg
