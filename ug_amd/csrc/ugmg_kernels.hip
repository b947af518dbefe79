// CDNA4 (gfx950) kernels of the multigrid solve path.
//
// Matrix layout: SELL-64 slabs (ugmg_internal.h).  One wavefront owns the 64
// slots of a slice and walks the slice's entries column-major, so every
// value/column load of a wavefront is one contiguous 256/512-byte stretch and
// the x gathers of neighbouring stencil rows share cache lines.  Row sums are
// formed in the entries' list order (diagonal first), i.e. UG's order.
// Loads are issued in chunks of 8 entries (clamped to the row width, so the
// chunk needs no branch) to keep eight independent memory requests in flight
// per lane: small levels are latency bound, not bandwidth bound.
// Streams over the large levels use non-temporal loads so that the matrix
// does not evict the vectors from L2; small levels use cached loads.
// Blocks are 4 wavefronts; block ids are remapped so that each of the 8 XCDs
// (round-robin dispatch) walks one contiguous range of slices and keeps its
// share of x in its own L2.  Batched launches (several right-hand sides,
// column index = second grid coordinate) remap slice-major: an XCD takes a
// range of slice blocks and runs every column of one block before the next,
// so that block's matrix entries are read from HBM once and then hit in L2.  No MFMA anywhere: these are sparse contractions
// at ~0.15 flop/byte, bounded by HBM bandwidth.
#include <hip/hip_runtime.h>

#include "ugmg_kernels.h"

#include "../../include/ugmg.h"

namespace ugmg {

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int nx = 8;
  const int xcd = bid % nx, idx = bid / nx;
  const int per = nb / nx, rem = nb % nx;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

template <int NT, typename T>
__device__ __forceinline__ T ld(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// sum_{k<w} val[64 k] * x[col_k] in entry order, in chunks of CH entries:
// the loads of a chunk are clamped to the row width (no branch, so they all
// issue before the first product) and a row of at most CH entries costs one
// round of loads (two for explicit columns).  CH is chosen per launch from
// the widest slice.
// Uniform slice (sdo >= 0): col_k = rb + dtab[sdo + k] (wave-uniform offsets
// in scalar registers), so the x gathers issue together with the value loads;
// explicit slice: col_k = ci[64 k].  TA is the accumulation type.
template <int CH, int NT, typename TV, typename TX, typename TA>
__device__ __forceinline__ TA sell_dot(const TV *__restrict__ v, const int *__restrict__ ci,
                                       int w, const TX *__restrict__ x, int sdo,
                                       const int *__restrict__ dtab, int rb) {
  TA sum = 0;
  if (sdo >= 0) {
    const int *__restrict__ dt = dtab + sdo;
    for (int k = 0; k < w; k += CH) {
      TV a[CH];
      TX xv[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int kk = (k + j < w) ? k + j : w - 1;
        a[j] = ld<NT>(v + 64 * kk);
        xv[j] = x[rb + dt[kk]];
      }
#pragma unroll
      for (int j = 0; j < CH; ++j)
        if (k + j < w) sum += (TA)a[j] * (TA)xv[j];
    }
    return sum;
  }
  for (int k = 0; k < w; k += CH) {
    TV a[CH];
    int c[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int kk = (k + j < w) ? k + j : w - 1;
      a[j] = ld<NT>(v + 64 * kk);
      c[j] = ld<NT>(ci + 64 * kk);
    }
    TX xv[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) xv[j] = x[c[j]];
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (k + j < w) sum += (TA)a[j] * (TA)xv[j];
  }
  return sum;
}

// b x b blocks (b = 3): out_i = sum_k sum_j A_k[i][j] x[b col_k + j] over the
// row's entries in order, entry k's block in b*b planes of pst values;
// chunks of 2 entries (18 values and 6 gathers in flight per lane).
template <int B, int NT, typename TV, typename TX, typename TA>
__device__ __forceinline__ void sell_dot_blk(const TV *__restrict__ v, size_t pst,
                                             const int *__restrict__ ci, int w,
                                             const TX *__restrict__ x, int sdo,
                                             const int *__restrict__ dtab, int rb, TA (&out)[B]) {
  constexpr int CH = 2;
#pragma unroll
  for (int i = 0; i < B; ++i) out[i] = 0;
  for (int k = 0; k < w; k += CH) {
    TV a[CH][B * B];
    TX xv[CH][B];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int kk = (k + j < w) ? k + j : w - 1;
      const int c = sdo >= 0 ? rb + dtab[sdo + kk] : ld<NT>(ci + 64 * kk);
#pragma unroll
      for (int q = 0; q < B * B; ++q) a[j][q] = ld<NT>(v + q * pst + 64 * kk);
#pragma unroll
      for (int i = 0; i < B; ++i) xv[j][i] = x[(size_t)B * c + i];
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (k + j < w) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
          TA t = 0;
#pragma unroll
          for (int jj = 0; jj < B; ++jj) t += (TA)a[j][i * B + jj] * (TA)xv[j][jj];
          out[i] += t;
        }
      }
  }
}

// scalar weights on vectors of B components (the transfer): out_i =
// sum_k w_k x[B col_k + i]
template <int B, int NT, typename TV, typename TX, typename TA>
__device__ __forceinline__ void sell_dot_diag(const TV *__restrict__ v, const int *__restrict__ ci,
                                              int w, const TX *__restrict__ x, int sdo,
                                              const int *__restrict__ dtab, int rb, TA (&out)[B]) {
  constexpr int CH = 4;
#pragma unroll
  for (int i = 0; i < B; ++i) out[i] = 0;
  for (int k = 0; k < w; k += CH) {
    TV a[CH];
    TX xv[CH][B];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int kk = (k + j < w) ? k + j : w - 1;
      const int c = sdo >= 0 ? rb + dtab[sdo + kk] : ld<NT>(ci + 64 * kk);
      a[j] = ld<NT>(v + 64 * kk);
#pragma unroll
      for (int i = 0; i < B; ++i) xv[j][i] = x[(size_t)B * c + i];
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (k + j < w) {
#pragma unroll
        for (int i = 0; i < B; ++i) out[i] += (TA)a[j] * (TA)xv[j][i];
      }
  }
}

// SolveSmallBlock for n = 3 (np/algebra/block.cc:131-142): elimination
// without pivoting, in UG's operation order; m row-major
template <typename T>
__device__ __forceinline__ void solve3(const T (&m)[9], const T (&r)[3], T (&s)[3]) {
  const T M3div0 = m[3] / m[0];
  const T M6div0 = m[6] / m[0];
  const T aux = (m[7] - M6div0 * m[1]) / (m[4] - M3div0 * m[1]);
  s[2] = (r[2] - M6div0 * r[0] - aux * (r[1] - M3div0 * r[0])) /
         (m[8] - M6div0 * m[2] - aux * (m[5] - M3div0 * m[2]));
  s[1] = (r[1] - m[3] / m[0] * r[0] - (m[5] - M3div0 * m[2]) * s[2]) / (m[4] - M3div0 * m[1]);
  s[0] = (r[0] - m[1] * s[1] - m[2] * s[2]) / m[0];
}

// SolveSmallBlock (np/algebra/block.cc:119-205) for n = 2 (the inverse
// determinant times the cofactor products, UG's operation order), n = 3,
// and n >= 4: Gauss elimination in which row i is exchanged with the row of
// largest |entry| in column i only where |pivot| < 1e-10 (block.cc:148-186),
// then back substitution.  Blocks UG would refuse (n = 2: det = 0; n >= 4:
// |pivot| < SMALL_D 1e-10 after the exchange) are refused at preprocess
// (k_block_check), so the solve here never meets them.  Fully unrolled: the
// row exchange selects by compile-time index and the block stays in registers.
template <int B, typename T>
__device__ __forceinline__ void solve_small(const T (&m)[B * B], const T (&r)[B], T (&s)[B]) {
  if constexpr (B == 2) {
    T det = m[0] * m[3] - m[1] * m[2];
    det = (T)1 / det;
    s[0] = (r[0] * m[3] - r[1] * m[1]) * det;
    s[1] = (r[1] * m[0] - r[0] * m[2]) * det;
  } else if constexpr (B == 3) {
    solve3<T>(m, r, s);
  } else {
    T M[B * B], rh[B];
#pragma unroll
    for (int q = 0; q < B * B; ++q) M[q] = m[q];
#pragma unroll
    for (int q = 0; q < B; ++q) rh[q] = r[q];
#pragma unroll
    for (int i = 0; i < B - 1; ++i) {
      T piv = M[i * B + i];
      if (fabs(piv) < (T)1e-10) {
        int pr = i;
#pragma unroll
        for (int j = i + 1; j < B; ++j)
          if (fabs(M[j * B + i]) > fabs(piv)) {
            piv = M[j * B + i];
            pr = j;
          }
#pragma unroll
        for (int j = i + 1; j < B; ++j)
          if (j == pr) {
#pragma unroll
            for (int k = i; k < B; ++k) {
              const T f = M[i * B + k];
              M[i * B + k] = M[j * B + k];
              M[j * B + k] = f;
            }
            const T f = rh[i];
            rh[i] = rh[j];
            rh[j] = f;
          }
      }
#pragma unroll
      for (int j = i + 1; j < B; ++j) {
        const T f = M[j * B + i] / piv;
#pragma unroll
        for (int k = i + 1; k < B; ++k) M[j * B + k] -= M[i * B + k] * f;
        rh[j] -= rh[i] * f;
      }
    }
#pragma unroll
    for (int i = B - 1; i >= 0; --i) {
#pragma unroll
      for (int j = i + 1; j < B; ++j) rh[i] -= M[i * B + j] * s[j];
      s[i] = rh[i] / M[i * B + i];
    }
  }
}

// block sum of B doubles per thread; thread 0 gets the totals
template <int B>
__device__ __forceinline__ void block_sum_b(double (&v)[B], double (*red)[kBlock / 64]) {
#pragma unroll
  for (int i = 0; i < B; ++i) {
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_down(v[i], off, 64);
    if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < B; ++i) {
      double t = 0.0;
      for (int k = 0; k < kBlock / 64; ++k) t += red[i][k];
      v[i] = t;
    }
}

// Chunked round-robin: XCD x runs chunks x, x + 8, x + 16, ... of kChunk
// consecutive blocks (the tail beyond whole rounds of 8 chunks keeps its
// place).  For launches whose work per slice changes along the slice range
// -- the defect over every colour, where rows of the first colour hold all
// upper entries and rows of the last none; the interpolation, whose fine
// rows hold 1, 2, 4 or 8 coarse corners by colour -- every XCD gets a share
// of each region, while the blocks of a chunk still share their L2.
__device__ __forceinline__ int xcd_chunked(int bid, int nb) {
  constexpr int kChunk = 32;
  const int nfull = nb / (8 * kChunk) * (8 * kChunk);
  if (bid >= nfull) return bid;
  const int x = bid % 8, k = bid / 8;
  return ((k / kChunk) * 8 + x) * kChunk + k % kChunk;
}

// Colour-balanced spatial map for a launch over every colour of a
// colour-grouped slab (the GS defect): XCD x (hardware block h runs on XCD
// h % 8) takes the x-th eighth of each colour's blocks, so every XCD gets
// the same share of each colour's work (their rows' widths differ) and its
// gathers stay in one spatial eighth of the level, whose vector fits its
// L2.  cb[k]: first block of colour k, cb[n] = all blocks.  -1: no block.
__device__ __forceinline__ int xcd_colour_block(int h, int n, const int *cb) {
  const int x = h & 7;
  int j = h >> 3;
  for (int k = 0; k < n; ++k) {
    const int nk = cb[k + 1] - cb[k];
    const int lo = (x * nk) >> 3, hi = ((x + 1) * nk) >> 3;
    if (j < hi - lo) return cb[k] + lo + j;
    j -= hi - lo;
  }
  return -1;
}

// Slice block and column of this workgroup.  Single column: the XCD-aware
// remap of blockIdx.x (contiguous range per XCD, or chunked round-robin
// with CHUNKED).  Several columns: the linear workgroup id is remapped over
// the whole grid and split slice-block-major.
template <bool CHUNKED = false>
__device__ __forceinline__ void batch_map(int &sb, int &col) {
  if (gridDim.y == 1) {
    sb = CHUNKED ? xcd_chunked(blockIdx.x, gridDim.x) : xcd_remap(blockIdx.x, gridDim.x);
    col = 0;
    return;
  }
  const int nb = gridDim.x * gridDim.y;
  const int id = blockIdx.y * gridDim.x + blockIdx.x;
  const int lr = CHUNKED ? xcd_chunked(id, nb) : xcd_remap(id, nb);
  sb = lr / gridDim.y;
  col = lr % gridDim.y;
}

// the slice block of a launch over a slab: the colour map of a whole
// colour-grouped slab with one column, else batch_map; -1: no block
template <bool CHUNKED, typename TV>
__device__ __forceinline__ void slab_map(const DSell<TV> &S, int slice0, int nsl, int &sb, int &col) {
  if (S.ncb > 0 && gridDim.y == 1 && slice0 == 0 && nsl == S.nslices) {
    sb = xcd_colour_block(blockIdx.x, S.ncb, S.cb);
    col = 0;
    return;
  }
  batch_map<CHUNKED>(sb, col);
}

// slice s (wave-uniform, in scalar registers), lane, row r (-1 = padding),
// entry base e0, width w, uniform-offset list sdo, index base rb, slice
// block sb_, column col_
#define SLICE_PROLOGUE(S, slice0, nsl) SLICE_PROLOGUE_MAP(S, slice0, nsl, false)
#define SLICE_PROLOGUE_MAP(S, slice0, nsl, CHUNKED)                                      \
  int sb_, col_;                                                                         \
  slab_map<CHUNKED>(S, slice0, nsl, sb_, col_);                                          \
  if (sb_ < 0) return;                                                                   \
  const int s = __builtin_amdgcn_readfirstlane((slice0) + sb_ * (kBlock / kSlice) +      \
                                               (int)(threadIdx.x >> 6));                 \
  if (s >= (slice0) + (nsl)) return;                                                     \
  const int lane = threadIdx.x & 63;                                                     \
  const int r = (S).rows[s * kSlice + lane];                                             \
  const int e0 = (S).soff[s];                                                            \
  const int w = ((S).soff[s + 1] - e0) >> 6;                                             \
  const int sdo = (S).sdoff[s];                                                          \
  const int rb = r >= 0 ? r : (S).rows[s * kSlice];

// same for a block of BS threads (BS / 64 slices)
#define SLICE_PROLOGUE_BS(S, slice0, nsl, BS)                                            \
  int sb_, col_;                                                                         \
  batch_map<false>(sb_, col_);                                                           \
  const int s = __builtin_amdgcn_readfirstlane((slice0) + sb_ * ((BS) / kSlice) +        \
                                               (int)(threadIdx.x >> 6));                 \
  if (s >= (slice0) + (nsl)) return;                                                     \
  const int lane = threadIdx.x & 63;                                                     \
  const int r = (S).rows[s * kSlice + lane];                                             \
  const int e0 = (S).soff[s];                                                            \
  const int w = ((S).soff[s + 1] - e0) >> 6;                                             \
  const int sdo = (S).sdoff[s];                                                          \
  const int rb = r >= 0 ? r : (S).rows[s * kSlice];

// same, for kernels that end in a block reduction (no early return)
#define SLICE_PROLOGUE_RED(S, nsl)                                                       \
  int sb_, col_;                                                                         \
  slab_map<false>(S, 0, nsl, sb_, col_);                                                 \
  if (sb_ < 0) return;   /* the whole workgroup */                                       \
  const int s = __builtin_amdgcn_readfirstlane(sb_ * (kBlock / kSlice) +                 \
                                               (int)(threadIdx.x >> 6));                 \
  const int lane = threadIdx.x & 63;                                                     \
  const bool sok = s < (nsl);                                                            \
  const int r = sok ? (S).rows[s * kSlice + lane] : -1;                                  \
  const int e0 = sok ? (S).soff[s] : 0;                                                  \
  const int w = sok ? ((S).soff[s + 1] - e0) >> 6 : 0;                                   \
  const int sdo = sok ? (S).sdoff[s] : -1;                                               \
  const int rb = r >= 0 ? r : (sok ? (S).rows[s * kSlice] : 0);

// a slab's values for the workgroup's column (members: one matrix per column)
#define MV(S) ((S).val + (S).mst * (size_t)col_)

// block sum of one double per thread; thread 0 gets the total
__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) t += red[i];
  return t;
}

// ------------------------------------------------------------------- SpMV
// MODE: 0 y -= A x, 1 y = A x, 2 y -= A x and c += x, 3 y -= A x and c = x
template <int MW, typename TV, typename TX, typename TY, int MODE, int NT>
__global__ __launch_bounds__(kBlock) void k_spmv(DSell<TV> A, int slice0, int nsl,
                                                  const TX *__restrict__ x, TY *__restrict__ y,
                                                  TY *__restrict__ c, size_t vs) {
  SLICE_PROLOGUE(A, slice0, nsl)
  const size_t o = (size_t)col_ * vs;
  x += o;
  y += o;
  const TY sum = sell_dot<MW, NT, TV, TX, TY>(MV(A) + e0 + lane, A.col + e0 + lane, w, x, sdo, A.dtab, rb);
  if (r < 0) return;
  if (MODE == 1) {
    y[r] = sum;
  } else {
    y[r] -= sum;
    if (MODE == 2) c[o + r] += (TY)x[r];
    if (MODE == 3) c[o + r] = (TY)x[r];
  }
}

// y -= A x over a GS level's colour-grouped rows, and for the rows of colour
// 0 (slices below ncs0) also the next GS sweep's colour 0, u_r = y_r / a_rr
// (colour 0 has no lower neighbours; every row active): the separate sweep
// would compute (y_r - 0) / a_rr, the same value.
template <int MW, typename TV, typename T, int NT>
__global__ __launch_bounds__(kBlock) void k_spmv_c0(DSell<TV> A, int nsl, const T *__restrict__ x,
                                                     T *__restrict__ y, const T *__restrict__ diag,
                                                     T *__restrict__ u, int ncs0, size_t vs) {
  SLICE_PROLOGUE(A, 0, nsl)
  const size_t o = (size_t)col_ * vs;
  x += o;
  y += o;
  const T sum = sell_dot<MW, NT, TV, T, T>(MV(A) + e0 + lane, A.col + e0 + lane, w, x, sdo, A.dtab, rb);
  if (r < 0) return;
  const T yn = y[r] - sum;
  y[r] = yn;
  if (s < ncs0) u[o + r] = yn / diag[A.dms * col_ + r];
}

// FP64 defect of a mixed-precision refinement step:
//   r -= A c32 (FP64 values and sums); x += c32; r32 = (float) r;
//   per-block partial sums of r^2 over NEW_DEFECT rows.
template <int MW, int NT>
__global__ __launch_bounds__(kBlock) void k_ir_update(DSell<double> A, int nsl,
                                                       const float *__restrict__ c32,
                                                       double *__restrict__ rv,
                                                       double *__restrict__ x,
                                                       float *__restrict__ r32,
                                                       const uint8_t *__restrict__ vclass,
                                                       double *__restrict__ part, size_t vs) {
  __shared__ double red[kBlock / 64];
  SLICE_PROLOGUE_RED(A, nsl)
  const size_t o = (size_t)col_ * vs;
  double sq = 0.0;
  if (sok) {
    const double sum =
        sell_dot<MW, NT, double, float, double>(MV(A) + e0 + lane, A.col + e0 + lane, w, c32 + o, sdo, A.dtab, rb);
    if (r >= 0) {
      const double nr = rv[o + r] - sum;
      rv[o + r] = nr;
      x[o + r] += (double)c32[o + r];
      r32[o + r] = (float)nr;
      if (vclass[r] >= 2) sq = nr * nr;
    }
  }
  const double t = block_sum(sq, red);
  if (threadIdx.x == 0) part[(size_t)col_ * gridDim.x + sb_] = t;
}

// ----------------------------------------------------------------- Jacobi
// l_jac + dscalx: t = damp * b / diag, 0 for VCLASS < 3
template <typename T>
__global__ __launch_bounds__(kBlock) void k_jac(int n, const T *__restrict__ b,
                                                 const T *__restrict__ diag,
                                                 const uint8_t *__restrict__ vclass, T damp,
                                                 T *__restrict__ t, size_t vs, size_t dms) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const size_t o = (size_t)blockIdx.y * vs;
  t[o + i] = vclass[i] < 3 ? (T)0 : (b[o + i] / diag[dms * blockIdx.y + i]) * damp;
}

// ------------------------------------------------------ multicolour GS
// forward substitution of one colour: v_r = (b_r - sum_{lower} a v) / a_rr
template <int MW, typename TV, typename T, int NT, int BS = kBlock>
__global__ __launch_bounds__(BS) void k_gs_color(DSell<TV> L, int slice0, int nsl,
                                                  const T *__restrict__ b,
                                                  const T *__restrict__ diag,
                                                  const uint8_t *__restrict__ vclass,
                                                  T *__restrict__ v, size_t vs) {
  SLICE_PROLOGUE_BS(L, slice0, nsl, BS)
  const size_t o = (size_t)col_ * vs;
  v += o;
  // the row's own data load alongside the entries
  T br = 0, dr = 1;
  int act = 0;
  if (r >= 0) {
    br = b[o + r];
    dr = diag[L.dms * col_ + r];
    act = vclass[r] >= 3;
  }
  const T sum = sell_dot<MW, NT, TV, T, T>(MV(L) + e0 + lane, L.col + e0 + lane, w, v, sdo, L.dtab, rb);
  if (r < 0) return;
  v[r] = act ? (br - sum) / dr : (T)0;
}

// Step s+1 of the sweep with step s's defect folded in (levels whose rows
// are all active).  When colour k runs, v still holds step s at the higher
// colours and step s+1 at the lower ones, so one pass over the row gives
//   b_r <- b_r - damp (b_r + (U v_s)_r)      (the defect of step s)
//   c_r (+)= damp v_s,r                      (CMODE 0: =, 1: +=)
//   v_r <- (b_r - (L v_{s+1})_r) / a_rr      (the sweep of step s+1)
// with the same operations, in the same order, as the separate kernels.
template <int CHU, int CHL, typename TV, typename T, int CMODE, int NT, int BS = kBlock>
__global__ __launch_bounds__(BS) void k_gs_color_fused(DSell<TV> L, DSell<TV> U, int slice0,
                                                        int nsl, T *__restrict__ b,
                                                        const T *__restrict__ diag, T damp,
                                                        T *__restrict__ v,
                                                        T *__restrict__ c, size_t vs) {
  SLICE_PROLOGUE_BS(L, slice0, nsl, BS)
  const size_t o = (size_t)col_ * vs;
  v += o;
  const int eu = U.soff[s];
  const int wu = (U.soff[s + 1] - eu) >> 6;
  const int sdu = U.sdoff[s];
  T br = 0, dr = 1, vo = 0, cr = 0;
  if (r >= 0) {
    br = b[o + r];
    dr = diag[L.dms * col_ + r];
    vo = v[r];
    if (CMODE == 1) cr = c[o + r];
  }
  const T su = sell_dot<CHU, NT, TV, T, T>(MV(U) + eu + lane, U.col + eu + lane, wu, v, sdu, U.dtab, rb);
  const T sl = sell_dot<CHL, NT, TV, T, T>(MV(L) + e0 + lane, L.col + e0 + lane, w, v, sdo, L.dtab, rb);
  if (r < 0) return;
  const T bn = br - damp * (br + su);
  b[o + r] = bn;
  const T t = damp * vo;
  c[o + r] = CMODE == 0 ? t : cr + t;
  v[r] = (bn - sl) / dr;
}

// The same step with the row's two products on the two wavefronts of one
// workgroup per slice: wavefront 0 forms U v_s (then b, c and v), wavefront
// 1 forms L v_{s+1} and passes it through LDS.  Both read v only at other
// colours, and each product runs in the same order as above, so the results
// are the same bits with half the dependent load rounds per wavefront.
template <int CHU, int CHL, typename TV, typename T, int CMODE, int NT>
__global__ __launch_bounds__(2 * kSlice) void k_gs_color_fused2(DSell<TV> L, DSell<TV> U,
                                                                 int slice0, int nsl,
                                                                 T *__restrict__ b,
                                                                 const T *__restrict__ diag,
                                                                 T damp, T *__restrict__ v,
                                                                 T *__restrict__ c, size_t vs) {
  __shared__ T lsum[kSlice];
  int sb_, col_;
  batch_map<false>(sb_, col_);
  const int s = __builtin_amdgcn_readfirstlane(slice0 + sb_);
  if (s >= slice0 + nsl) return;   // the whole workgroup
  const int half = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int r = L.rows[s * kSlice + lane];
  const int rb = r >= 0 ? r : L.rows[s * kSlice];
  const size_t o = (size_t)col_ * vs;
  v += o;
  T br = 0, dr = 1, vo = 0, cr = 0, part;
  if (half == 0) {
    if (r >= 0) {
      br = b[o + r];
      dr = diag[L.dms * col_ + r];
      vo = v[r];
      if (CMODE == 1) cr = c[o + r];
    }
    const int eu = U.soff[s];
    part = sell_dot<CHU, NT, TV, T, T>(MV(U) + eu + lane, U.col + eu + lane, (U.soff[s + 1] - eu) >> 6,
                                       v, U.sdoff[s], U.dtab, rb);
  } else {
    const int e0 = L.soff[s];
    part = sell_dot<CHL, NT, TV, T, T>(MV(L) + e0 + lane, L.col + e0 + lane, (L.soff[s + 1] - e0) >> 6,
                                       v, L.sdoff[s], L.dtab, rb);
    lsum[lane] = part;
  }
  __syncthreads();
  if (half == 1 || r < 0) return;
  const T bn = br - damp * (br + part);
  b[o + r] = bn;
  const T t = damp * vo;
  c[o + r] = CMODE == 0 ? t : cr + t;
  v[r] = (bn - lsum[lane]) / dr;
}

// defect after the sweep, b -= A (damp v), with (D+L) v = b on active rows:
//   active   b <- b - damp (b + U v)
//   inactive b <- b - damp (A v)      (v_r = 0; U holds the whole row)
// and the correction: CMODE 0 c = damp v, 1 c += damp v; xout = damp v.
template <int MW, typename TV, typename T, int CMODE, int NT>
__global__ __launch_bounds__(kBlock) void k_gs_defect(DSell<TV> U, int nsl,
                                                       const T *__restrict__ v,
                                                       T *__restrict__ b,
                                                       const uint8_t *__restrict__ vclass,
                                                       T damp, T *__restrict__ c,
                                                       T *__restrict__ xout, size_t vs) {
  SLICE_PROLOGUE_MAP(U, 0, nsl, true)
  const size_t o = (size_t)col_ * vs;
  v += o;
  T br = 0, vr = 0, cr = 0;
  int act = 0;
  if (r >= 0) {
    br = b[o + r];
    vr = v[r];
    act = vclass[r] >= 3;
    if (CMODE == 1 && c) cr = c[o + r];
  }
  const T sum = sell_dot<MW, NT, TV, T, T>(MV(U) + e0 + lane, U.col + e0 + lane, w, v, sdo, U.dtab, rb);
  if (r < 0) return;
  b[o + r] = act ? br - damp * (br + sum) : br - damp * sum;
  const T t = damp * vr;
  if (c) {
    if (CMODE == 0) c[o + r] = t;
    else c[o + r] = cr + t;
  }
  if (xout) xout[o + r] = t;
}

// --------------------------------------------------------------- transfer
// StandardRestrictNodeVector as a gather over the slabs of P^T:
//   bc_I = (VNCLASS_I >= 2 ? 0 : bc_I) + (Dirichlet_I ? 0 : damp sum w b_f)
// With `part`, also the per-block partial sums of the new bc^2 over NEW_DEFECT
// vectors (the base solver's first LinearResiduum).
template <int MW, typename T, int NT>
__global__ __launch_bounds__(kBlock) void k_restrict(DSell<T> PT, int nsl,
                                                      const uint32_t *__restrict__ cskip,
                                                      const uint8_t *__restrict__ cvnclass,
                                                      const uint8_t *__restrict__ cvclass,
                                                      const T *__restrict__ bf, T damp,
                                                      T *__restrict__ bc,
                                                      double *__restrict__ part, size_t vsf,
                                                      size_t vsc,
                                                      const uint8_t *__restrict__ cc0,
                                                      const T *__restrict__ cdiag,
                                                      T *__restrict__ ct,
                                                      const int *__restrict__ cpos,
                                                      T *__restrict__ cbc, size_t cdms) {
  __shared__ double red[kBlock / 64];
  SLICE_PROLOGUE_RED(PT, nsl)
  bf += (size_t)col_ * vsf;
  bc += (size_t)col_ * vsc;
  double sq = 0.0;
  if (sok) {
    const T sum = sell_dot<MW, NT, T, T, T>(MV(PT) + e0 + lane, PT.col + e0 + lane, w, bf, sdo, PT.dtab, rb);
    if (r >= 0) {
      const T base = cvnclass[r] >= 2 ? (T)0 : bc[r];
      const T v = (cskip[r] & 1u) ? base : base + damp * sum;
      bc[r] = v;
      // the coarse pre-smoother's colour-0 sweep: no lower neighbours, so
      // its value is (b_r - 0) / a_rr
      if (ct && cc0[r]) ct[(size_t)col_ * vsc + r] = v / cdiag[cdms * col_ + r];
      // b in the base inverse's column order
      if (cpos) {
        const int pj = cpos[r];
        if (pj >= 0) cbc[(size_t)col_ * vsc + pj] = v;
      }
      if (part && cvclass[r] >= 2) sq = (double)v * (double)v;
    }
  }
  if (!part) return;
  const double t = block_sum(sq, red);
  if (threadIdx.x == 0) part[(size_t)col_ * gridDim.x + sb_] = t;
}

// StandardIntCorNodeVector (+ dadd): t = damp P c_c on non-Dirichlet rows
template <int MW, typename T, int CADD, int NT>
__global__ __launch_bounds__(kBlock) void k_interp(DSell<T> P, int nsl,
                                                    const uint32_t *__restrict__ fskip,
                                                    const T *__restrict__ cc, T damp,
                                                    T *__restrict__ t, T *__restrict__ c,
                                                    size_t vsf, size_t vsc) {
  SLICE_PROLOGUE_MAP(P, 0, nsl, true)
  cc += (size_t)col_ * vsc;
  const size_t o = (size_t)col_ * vsf;
  const T sum = sell_dot<MW, NT, T, T, T>(MV(P) + e0 + lane, P.col + e0 + lane, w, cc, sdo, P.dtab, rb);
  if (r < 0) return;
  const T tv = (fskip[r] & 1u) ? (T)0 : damp * sum;
  t[o + r] = tv;
  if (CADD) c[o + r] += tv;
}

// -------------------------------------------------------------- reductions
template <typename T, int DOT>
__global__ __launch_bounds__(kBlock) void k_dot_part(int n, const T *__restrict__ x,
                                                      const T *__restrict__ y,
                                                      const uint8_t *__restrict__ vclass,
                                                      double *__restrict__ part, size_t vs) {
  __shared__ double red[kBlock / 64];
  const size_t o = (size_t)blockIdx.y * vs;
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (vclass && vclass[i] < 2) continue;
    const double xi = (double)x[o + i];
    s += DOT ? xi * (double)y[o + i] : xi * xi;
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_reduce_final(int nparts,
                                                        const double *__restrict__ part,
                                                        double *__restrict__ out, int do_sqrt) {
  __shared__ double red[16];
  part += (size_t)blockIdx.x * nparts;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 1024) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < 16; ++i) t += red[i];
    out[blockIdx.x] = do_sqrt ? sqrt(t) : t;
  }
}

// ------------------------------------------------------------------ BLAS-1
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int n, T *x, T val, size_t vs) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[(size_t)blockIdx.y * vs + i] = val;
}
template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const TS *__restrict__ s,
                                                     TD *__restrict__ d, size_t vss, size_t vsd) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) d[(size_t)blockIdx.y * vsd + i] = (TD)s[(size_t)blockIdx.y * vss + i];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add(int n, T *__restrict__ x, const T *__restrict__ y,
                                                 size_t vs) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const size_t o = (size_t)blockIdx.y * vs;
  if (i < n) x[o + i] += y[o + i];
}

// class byte of the FINE_GRID_DOF flag (gm/algebra.cc:2908-2909): 3 where
// VCLASS >= 2 and VNCLASS <= 1, else 0, for the class-filtered reductions
__global__ __launch_bounds__(kBlock) void k_fgd_class(int n, const uint8_t *__restrict__ vc,
                                                       const uint8_t *__restrict__ vnc,
                                                       uint8_t *__restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (vc[i] >= 2 && vnc[i] <= 1) ? 3 : 0;
}

// dst = src on the rows whose class byte is >= 2 (scalar entry e, row e / b)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_rows(int ne, int b, const uint8_t *__restrict__ cls,
                                                       const T *__restrict__ src,
                                                       T *__restrict__ dst) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < ne && cls[e / b] >= 2) dst[e] = src[e];
}

// ================================================= 2x2 .. 4x4 blocks
// The same operations on vectors of B = 2, 3 or 4 components with B x B matrix
// blocks (linear elasticity, coupled systems); diag holds the diagonal
// block of every row.

template <int B, typename TV, typename TX, typename TY, int MODE, int NT>
__global__ __launch_bounds__(kBlock) void k_spmv_bk(DSell<TV> A, int slice0, int nsl,
                                                     const TX *__restrict__ x,
                                                     TY *__restrict__ y, TY *__restrict__ c,
                                                     size_t vs) {
  SLICE_PROLOGUE(A, slice0, nsl)
  const size_t o = (size_t)col_ * vs;
  x += o;
  y += o;
  TY sum[B];
  sell_dot_blk<B, NT, TV, TX, TY>(MV(A) + e0 + lane, A.pst, A.col + e0 + lane, w, x, sdo, A.dtab, rb, sum);
  if (r < 0) return;
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const size_t q = (size_t)B * r + i;
    if (MODE == 1) {
      y[q] = sum[i];
    } else {
      y[q] -= sum[i];
      if (MODE == 2) c[o + q] += (TY)x[q];
      if (MODE == 3) c[o + q] = (TY)x[q];
    }
  }
}

template <int B, int NT>
__global__ __launch_bounds__(kBlock) void k_ir_update_bk(DSell<double> A, int nsl,
                                                          const float *__restrict__ c32,
                                                          double *__restrict__ rv,
                                                          double *__restrict__ x,
                                                          float *__restrict__ r32,
                                                          const uint8_t *__restrict__ vclass,
                                                          double *__restrict__ part, size_t vs) {
  __shared__ double red[B][kBlock / 64];
  SLICE_PROLOGUE_RED(A, nsl)
  const size_t o = (size_t)col_ * vs;
  double sq[B] = {};
  if (sok) {
    double sum[B];
    sell_dot_blk<B, NT, double, float, double>(MV(A) + e0 + lane, A.pst, A.col + e0 + lane, w,
                                               c32 + o, sdo, A.dtab, rb, sum);
    if (r >= 0)
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const size_t q = o + (size_t)B * r + i;
        const double nr = rv[q] - sum[i];
        rv[q] = nr;
        x[q] += (double)c32[q];
        r32[q] = (float)nr;
        if (vclass[r] >= 2) sq[i] = nr * nr;
      }
  }
  block_sum_b<B>(sq, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < B; ++i) part[((size_t)col_ * B + i) * gridDim.x + sb_] = sq[i];
}

template <int B, typename T>
__global__ __launch_bounds__(kBlock) void k_jac_bk(int n, const T *__restrict__ b,
                                                    const T *__restrict__ diag,
                                                    const uint8_t *__restrict__ vclass, T damp,
                                                    T *__restrict__ t, size_t vs, size_t dms) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const size_t o = (size_t)blockIdx.y * vs + B * (size_t)i;
  diag += dms * blockIdx.y;
  if (vclass[i] < 3) {
#pragma unroll
    for (int q = 0; q < B; ++q) t[o + q] = (T)0;
    return;
  }
  T m[B * B], r[B], x[B];
#pragma unroll
  for (int q = 0; q < B * B; ++q) m[q] = diag[B * B * (size_t)i + q];
#pragma unroll
  for (int q = 0; q < B; ++q) r[q] = b[o + q];
  solve_small<B, T>(m, r, x);
#pragma unroll
  for (int q = 0; q < B; ++q) t[o + q] = x[q] * damp;
}

template <int B, typename TV, typename T, int NT, int BS = kBlock>
__global__ __launch_bounds__(BS) void k_gs_color_bk(DSell<TV> L, int slice0, int nsl,
                                                         const T *__restrict__ b,
                                                         const T *__restrict__ diag,
                                                         const uint8_t *__restrict__ vclass,
                                                         T *__restrict__ v, size_t vs) {
  SLICE_PROLOGUE_BS(L, slice0, nsl, BS)
  const size_t o = (size_t)col_ * vs;
  v += o;
  T sum[B];
  sell_dot_blk<B, NT, TV, T, T>(MV(L) + e0 + lane, L.pst, L.col + e0 + lane, w, v, sdo, L.dtab, rb, sum);
  if (r < 0) return;
  T x[B] = {};
  if (vclass[r] >= 3) {
    T m[B * B], rr[B];
#pragma unroll
    for (int q = 0; q < B * B; ++q) m[q] = diag[L.dms * col_ + B * B * (size_t)r + q];
#pragma unroll
    for (int q = 0; q < B; ++q) rr[q] = b[o + (size_t)B * r + q] - sum[q];
    solve_small<B, T>(m, rr, x);
  }
#pragma unroll
  for (int q = 0; q < B; ++q) v[(size_t)B * r + q] = x[q];
}

template <int B, typename TV, typename T, int CMODE, int NT>
__global__ __launch_bounds__(kBlock) void k_gs_defect_bk(DSell<TV> U, int nsl,
                                                          const T *__restrict__ v,
                                                          T *__restrict__ b,
                                                          const uint8_t *__restrict__ vclass,
                                                          T damp, T *__restrict__ c,
                                                          T *__restrict__ xout, size_t vs) {
  SLICE_PROLOGUE_MAP(U, 0, nsl, true)
  const size_t o = (size_t)col_ * vs;
  v += o;
  T sum[B];
  sell_dot_blk<B, NT, TV, T, T>(MV(U) + e0 + lane, U.pst, U.col + e0 + lane, w, v, sdo, U.dtab, rb, sum);
  if (r < 0) return;
  const bool act = vclass[r] >= 3;
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const size_t k = (size_t)B * r + q;
    const T br = b[o + k];
    b[o + k] = act ? br - damp * (br + sum[q]) : br - damp * sum[q];
    const T t = damp * v[k];
    if (c) {
      if (CMODE == 0) c[o + k] = t;
      else c[o + k] += t;
    }
    if (xout) xout[o + k] = t;
  }
}

template <int B, typename TV, typename T, int CMODE, int NT, int BS = kBlock>
__global__ __launch_bounds__(BS) void k_gs_color_fused_bk(DSell<TV> L, DSell<TV> U,
                                                               int slice0, int nsl,
                                                               T *__restrict__ b,
                                                               const T *__restrict__ diag, T damp,
                                                               T *__restrict__ v,
                                                               T *__restrict__ c, size_t vs) {
  SLICE_PROLOGUE_BS(L, slice0, nsl, BS)
  const size_t o = (size_t)col_ * vs;
  v += o;
  const int eu = U.soff[s];
  const int wu = (U.soff[s + 1] - eu) >> 6;
  const int sdu = U.sdoff[s];
  T su[B], sl[B];
  sell_dot_blk<B, NT, TV, T, T>(MV(U) + eu + lane, U.pst, U.col + eu + lane, wu, v, sdu, U.dtab, rb, su);
  sell_dot_blk<B, NT, TV, T, T>(MV(L) + e0 + lane, L.pst, L.col + e0 + lane, w, v, sdo, L.dtab, rb, sl);
  if (r < 0) return;
  T m[B * B], rr[B], x[B];
#pragma unroll
  for (int q = 0; q < B * B; ++q) m[q] = diag[L.dms * col_ + B * B * (size_t)r + q];
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const size_t k = (size_t)B * r + q;
    const T br = b[o + k];
    const T bn = br - damp * (br + su[q]);
    b[o + k] = bn;
    const T t = damp * v[k];
    c[o + k] = CMODE == 0 ? t : c[o + k] + t;
    rr[q] = bn - sl[q];
  }
  solve_small<B, T>(m, rr, x);
#pragma unroll
  for (int q = 0; q < B; ++q) v[(size_t)B * r + q] = x[q];
}

// MAT: RestrictByMatrix (transgrid.cc:1160-1232), a b x b block M per entry
// (coarse component i, fine component j): b_c,i += sum_j M_ij b_f,j on the
// components whose VECSKIP bit is clear, and $damp scales only the coarse
// rows of VNCLASS >= 2 after the sum; otherwise the scalar weights of
// StandardRestrictNodeVector (every term times damp)
template <int B, typename T, int NT, bool MAT = false>
__global__ __launch_bounds__(kBlock) void k_restrict_bk(DSell<T> PT, int nsl,
                                                         const uint32_t *__restrict__ cskip,
                                                         const uint8_t *__restrict__ cvnclass,
                                                         const uint8_t *__restrict__ cvclass,
                                                         const T *__restrict__ bf, T damp,
                                                         T *__restrict__ bc,
                                                         double *__restrict__ part, size_t vsf,
                                                         size_t vsc) {
  __shared__ double red[B][kBlock / 64];
  SLICE_PROLOGUE_RED(PT, nsl)
  bf += (size_t)col_ * vsf;
  bc += (size_t)col_ * vsc;
  double sq[B] = {};
  if (sok) {
    T sum[B];
    if constexpr (MAT)
      sell_dot_blk<B, NT, T, T, T>(MV(PT) + e0 + lane, PT.pst, PT.col + e0 + lane, w, bf, sdo, PT.dtab, rb, sum);
    else
      sell_dot_diag<B, NT, T, T, T>(MV(PT) + e0 + lane, PT.col + e0 + lane, w, bf, sdo, PT.dtab, rb, sum);
    if (r >= 0) {
      const bool zero = cvnclass[r] >= 2;
      const uint32_t sk = cskip[r];
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const size_t k = (size_t)B * r + q;
        const T base = zero ? (T)0 : bc[k];
        const T v = ((sk >> q) & 1u) ? base
                    : MAT           ? (zero ? damp * sum[q] : base + sum[q])
                                    : base + damp * sum[q];
        bc[k] = v;
        if (part && cvclass[r] >= 2) sq[q] = (double)v * (double)v;
      }
    }
  }
  if (!part) return;
  block_sum_b<B>(sq, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < B; ++i) part[((size_t)col_ * B + i) * gridDim.x + sb_] = sq[i];
}

// MAT: InterpolateCorrectionByMatrix (transgrid.cc:1340-1380), t_v,i =
// damp sum_j M_ji c_j over the entries (the slab holds M^T per entry) on the
// components whose VECSKIP bit is clear; otherwise scalar weights
template <int B, typename T, int CADD, int NT, bool MAT = false>
__global__ __launch_bounds__(kBlock) void k_interp_bk(DSell<T> P, int nsl,
                                                       const uint32_t *__restrict__ fskip,
                                                       const T *__restrict__ cc, T damp,
                                                       T *__restrict__ t, T *__restrict__ c,
                                                       size_t vsf, size_t vsc) {
  SLICE_PROLOGUE_MAP(P, 0, nsl, true)
  cc += (size_t)col_ * vsc;
  const size_t o = (size_t)col_ * vsf;
  T sum[B];
  if constexpr (MAT)
    sell_dot_blk<B, NT, T, T, T>(MV(P) + e0 + lane, P.pst, P.col + e0 + lane, w, cc, sdo, P.dtab, rb, sum);
  else
    sell_dot_diag<B, NT, T, T, T>(MV(P) + e0 + lane, P.col + e0 + lane, w, cc, sdo, P.dtab, rb, sum);
  if (r < 0) return;
  const uint32_t sk = fskip[r];
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const size_t k = o + (size_t)B * r + q;
    const T tv = ((sk >> q) & 1u) ? (T)0 : damp * sum[q];
    t[k] = tv;
    if (CADD) c[k] += tv;
  }
}

// The blocks SolveSmallBlock refuses (np/algebra/block.cc:119-205; l_jac
// and l_lgs then fail with NUM_SMALL_DIAG): n = 2 with det = 0, n >= 4 where
// the elimination meets |pivot| < SMALL_D 1e-10 after the row exchange
// (SMALL_D = DBL_EPSILON * 10, low/architecture.h:27,34).  n = 1 and n = 3
// divide unchecked in UG and are not checked here.  One thread per vector of
// class >= 3 (the ones the smoothers solve); *bad = the smallest failing
// row.  FP64 without contraction: UG's products and differences.
template <int B>
__global__ __launch_bounds__(kBlock) void k_block_check(int n, const double *__restrict__ diag,
                                                         const uint8_t *__restrict__ vclass,
                                                         int *__restrict__ bad) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || vclass[i] < 3) return;
  double M[B * B];
#pragma unroll
  for (int q = 0; q < B * B; ++q) M[q] = diag[B * B * (size_t)i + q];
  bool fails = false;
  if constexpr (B == 2) {
    fails = M[0] * M[3] - M[1] * M[2] == 0.0;
  } else {
    constexpr double kSmallD = 2.2204460492503131e-16 * 10;
#pragma unroll
    for (int c = 0; c < B - 1; ++c) {
      double piv = M[c * B + c];
      if (fabs(piv) < 1e-10) {
        int pr = c;
#pragma unroll
        for (int j = c + 1; j < B; ++j)
          if (fabs(M[j * B + c]) > fabs(piv)) {
            piv = M[j * B + c];
            pr = j;
          }
        if (fabs(piv) < kSmallD * 1e-10) fails = true;
#pragma unroll
        for (int j = c + 1; j < B; ++j)
          if (j == pr)
#pragma unroll
            for (int k = c; k < B; ++k) {
              const double f = M[c * B + k];
              M[c * B + k] = M[j * B + k];
              M[j * B + k] = f;
            }
      }
      if (fails) break;
#pragma unroll
      for (int j = c + 1; j < B; ++j) {
        const double f = M[j * B + c] / piv;
#pragma unroll
        for (int k = c + 1; k < B; ++k) M[j * B + k] -= M[c * B + k] * f;
      }
    }
  }
  if (fails) atomicMin(bad, i);
}

void launch_block_check(hipStream_t st, int n, int b, const double *diag, const uint8_t *vclass,
                        int *bad) {
  if (n <= 0) return;
  const int g = (n + kBlock - 1) / kBlock;
  if (b == 2)
    k_block_check<2><<<g, kBlock, 0, st>>>(n, diag, vclass, bad);
  else if (b == 4)
    k_block_check<4><<<g, kBlock, 0, st>>>(n, diag, vclass, bad);
}

// partial sums per component: part[(column * 3 + i) * gridDim.x + block]
template <int B, typename T, int DOT>
__global__ __launch_bounds__(kBlock) void k_dot_part_bk(int n, const T *__restrict__ x,
                                                         const T *__restrict__ y,
                                                         const uint8_t *__restrict__ vclass,
                                                         double *__restrict__ part, size_t vs) {
  __shared__ double red[B][kBlock / 64];
  const size_t o = (size_t)blockIdx.y * vs;
  double s[B] = {};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (vclass && vclass[i] < 2) continue;
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const double xi = (double)x[o + (size_t)B * i + q];
      s[q] += DOT ? xi * (double)y[o + (size_t)B * i + q] : xi * xi;
    }
  }
  block_sum_b<B>(s, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < B; ++q) part[((size_t)blockIdx.y * B + q) * gridDim.x + blockIdx.x] = s[q];
}

// partial sums per (type, component) of vectors with several types (dnrm2x,
// ddot: a[aoff[VTYPE(v)] + i], ugblas.cc:3104-3176): part[(column * nc + k)
// * gridDim.x + block], k = vt[row] * b + component, nc <= UGMG_MAXB
template <typename T, int DOT>
__global__ __launch_bounds__(kBlock) void k_dot_part_typed(int n, int b, int nc, const T *__restrict__ x,
                                                            const T *__restrict__ y,
                                                            const uint8_t *__restrict__ vclass,
                                                            const uint8_t *__restrict__ vt,
                                                            double *__restrict__ part, size_t vs) {
  __shared__ double red[UGMG_MAXB][kBlock / 64];
  const size_t o = (size_t)blockIdx.y * vs;
  double s[UGMG_MAXB] = {};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    if (vclass && vclass[i] < 2) continue;
    const int k0 = vt[i] * b;
    for (int q = 0; q < b; ++q) {
      const double xi = (double)x[o + (size_t)b * i + q];
      const double v = DOT ? xi * (double)y[o + (size_t)b * i + q] : xi * xi;
#pragma unroll
      for (int k = 0; k < UGMG_MAXB; ++k)
        if (k == k0 + q) s[k] += v;
    }
  }
  block_sum_b<UGMG_MAXB>(s, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < nc; ++k) part[((size_t)blockIdx.y * nc + k) * gridDim.x + blockIdx.x] = s[k];
}

// lane's share of row . b for the base GEMV: four independent accumulators
// keep four loads of the row in flight per lane (the wave sums the shares)
// (rr's columns are in the order cp: column j multiplies b[cp[j]]; the
// first ncols of them are used).  The inverse is read once per cycle and
// streams with non-temporal loads, which keeps the caches for the levels'
// vectors and matrices (measured: mixed 129^3 solve 46.6 -> 44.7 ms).
template <typename T, typename TI>
__device__ __forceinline__ double base_row_share(const TI *__restrict__ rr,
                                                 const T *__restrict__ bb,
                                                 const int *__restrict__ cp, int ncols, int lane) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int j = lane;
  for (; j + 192 < ncols; j += 256) {
    a0 += (double)ld<1>(rr + j) * (double)bb[cp[j]];
    a1 += (double)ld<1>(rr + j + 64) * (double)bb[cp[j + 64]];
    a2 += (double)ld<1>(rr + j + 128) * (double)bb[cp[j + 128]];
    a3 += (double)ld<1>(rr + j + 192) * (double)bb[cp[j + 192]];
  }
  for (; j < ncols; j += 64) a0 += (double)ld<1>(rr + j) * (double)bb[cp[j]];
  return (a0 + a1) + (a2 + a3);
}

// This lane's share of row `row` of the inverse times b.  The inverse holds
// only the rows that are not unit rows (rowmap[row] = their index; -1: a unit
// row, from a Dirichlet row of the matrix, whose product is b[row] itself),
// n columns each, in the order cp (the non-unit unknowns first); the product
// runs over the first ncols columns (ncols < n when b is known to vanish on
// the others).
template <typename T, typename TI>
__device__ __forceinline__ double base_row_or_unit(const TI *__restrict__ inv,
                                                   const int *__restrict__ rowmap,
                                                   const int *__restrict__ cp, int ncols, int row,
                                                   const T *__restrict__ bb, int n, int lane) {
  const int k = rowmap[row];
  if (k < 0) return lane == 0 ? (double)bb[row] : 0.0;
  return base_row_share<T, TI>(inv + (size_t)k * n, bb, cp, ncols, lane);
}

// ls over ex on the base level with per-component stop tests (sc_cmp:
// every one of the B norm entries below its limit: the components, or the
// (type, component) pairs of several vector types); inv over scalar unknowns
template <int B, typename T, typename TI>
__global__ __launch_bounds__(kBlock) void k_base_gemv_bk(int n3, const TI *__restrict__ inv,
                                                          const int *__restrict__ rowmap,
                                                          const int *__restrict__ cp, int ncols,
                                                          const T *__restrict__ b,
                                                          T *__restrict__ c, T *__restrict__ xout,
                                                          int cmode, double damp, double abslimit,
                                                          double red, const double *__restrict__ part,
                                                          int nparts, double *__restrict__ first,
                                                          int it, size_t vs, size_t ims, int idents,
                                                          Batch::PDamp pd) {
  const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int cj = blockIdx.y;
  inv += ims * cj;
  const size_t o = (size_t)cj * vs;
  // the row product first, while the norms below are reduced
  double x = row < n3 ? base_row_or_unit<T, TI>(inv, rowmap, cp, ncols, row, b + o, n3, lane) : 0.0;
  double cur[B];
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const double *p = part + ((size_t)cj * B + q) * nparts;
    double s = 0.0;
    for (int i = lane; i < nparts; i += 64) s += p[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    cur[q] = sqrt(s);
  }
  bool below_abs = true, below_red = true;
  double target[B];
  if (it == 0 && row == 0 && lane == 0)
    for (int q = 0; q < B; ++q) first[(size_t)cj * B + q] = cur[q];
  for (int q = 0; q < B; ++q) {
    target[q] = it == 0 ? 0.0 : first[(size_t)cj * B + q] * red;
    if (target[q] == 0.0) target[q] = red;
  }
  if (idents < 0) {
    for (int q = 0; q < B; ++q) {
      below_abs &= fabs(cur[q]) < fabs(abslimit);
      below_red &= fabs(cur[q]) < fabs(target[q]);
    }
  } else {
    // sc_cmp with VD_IDENT (npscan.cc:1033-1048): per group leader i, the
    // group's sum of squares below the limits' (ident of entry q in bits 4q..)
    for (int i = 0; i < B; ++i) {
      if (((idents >> (4 * i)) & 15) != i) continue;
      double sx = 0.0, sa = 0.0, st = 0.0;
      for (int j = 0; j < B; ++j)
        if (((idents >> (4 * j)) & 15) == i) {
          sx += cur[j] * cur[j];
          sa += abslimit * abslimit;
          st += target[j] * target[j];
        }
      below_abs &= sx < sa;
      below_red &= sx < st;
    }
  }
  if (it == 0) below_red = false;
  const bool conv = below_abs || below_red;
  if (row >= n3) return;
  if (conv) {
    x = 0.0;
  } else {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    x *= pd.vb ? pd.d[row % pd.vb] : damp;
  }
  if (lane == 0) {
    const T xv = (T)x;
    if (!conv || cmode == 0) c[o + row] = cmode ? c[o + row] + xv : xv;
    if (xout) xout[o + row] = xv;
  }
}

// ================================================================ Galerkin
// AssembleGalerkinByMatrix (np/algebra/transgrid.cc:1575-1700) as a gather:
// one thread per coarse row I walks the fine vectors v that interpolate
// from I (fine list order), their matrix rows and the interpolation rows of
// the columns, and adds (a wi) wj to the coarse entry (I, J) of the given
// pattern (the diagonal when (I, J) is missing).  For a fixed coarse entry
// the contributions arrive in UG's scatter order (v, then m, then J).
template <int BB>
__global__ __launch_bounds__(kBlock) void k_galerkin(
    int nc, const int *__restrict__ ptp, const int *__restrict__ pti,
    const double *__restrict__ ptw, const int *__restrict__ frp, const int *__restrict__ fci,
    const double *__restrict__ fa, const int *__restrict__ pp, const int *__restrict__ pi,
    const double *__restrict__ pw, const int *__restrict__ crp, const int *__restrict__ cci,
    double *__restrict__ ca) {
  const int I = blockIdx.x * kBlock + threadIdx.x;
  if (I >= nc) return;
  const int c0 = crp[I], c1 = crp[I + 1];
  for (size_t k = (size_t)c0 * BB; k < (size_t)c1 * BB; ++k) ca[k] = 0.0;
  for (int q = ptp[I]; q < ptp[I + 1]; ++q) {
    const int v = pti[q];
    const double wi = ptw[q];
    for (int m = frp[v]; m < frp[v + 1]; ++m) {
      const int w = fci[m];
      for (int r = pp[w]; r < pp[w + 1]; ++r) {
        const int J = pi[r];
        const double wj = pw[r];
        int slot = c0;
        for (int k = c0; k < c1; ++k)
          if (cci[k] == J) {
            slot = k;
            break;
          }
#pragma unroll
        for (int e = 0; e < BB; ++e)
          ca[(size_t)slot * BB + e] += (fa[(size_t)m * BB + e] * wi) * wj;
      }
    }
  }
}

void launch_galerkin(hipStream_t st, int nc, int bb, const int *ptp, const int *pti,
                     const double *ptw, const int *frp, const int *fci, const double *fa,
                     const int *pp, const int *pi, const double *pw, const int *crp,
                     const int *cci, double *ca) {
  if (nc <= 0) return;
  if (bb == 16)
    k_galerkin<16><<<(nc + kBlock - 1) / kBlock, kBlock, 0, st>>>(nc, ptp, pti, ptw, frp, fci, fa, pp, pi, pw, crp, cci, ca);
  else if (bb == 9)
    k_galerkin<9><<<(nc + kBlock - 1) / kBlock, kBlock, 0, st>>>(nc, ptp, pti, ptw, frp, fci, fa, pp, pi, pw, crp, cci, ca);
  else if (bb == 4)
    k_galerkin<4><<<(nc + kBlock - 1) / kBlock, kBlock, 0, st>>>(nc, ptp, pti, ptw, frp, fci, fa, pp, pi, pw, crp, cci, ca);
  else
    k_galerkin<1><<<(nc + kBlock - 1) / kBlock, kBlock, 0, st>>>(nc, ptp, pti, ptw, frp, fci, fa, pp, pi, pw, crp, cci, ca);
}

// ==================================================== fused coarse cycle
// One workgroup runs Lmgc on every level from `top` down to the base level:
// each operation is a loop over rows or slabs followed by a workgroup
// barrier, so the levels below a few thousand vectors cost one launch per
// cycle instead of dozens of latency-bound launches.

__device__ __forceinline__ double wg_sum(double s, double *red) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kCoarseThreads / 64; ++i) t += red[i];
  return t;
}

template <typename TV, typename TX, class Epi>
__device__ __forceinline__ void wg_sell(const DSell<TV> &S, int s0, int s1,
                                        const TX *__restrict__ x, Epi epi) {
  const int nslot = (s1 - s0) * kSlice;
  for (int slot = threadIdx.x; slot < nslot; slot += kCoarseThreads) {
    const int s = __builtin_amdgcn_readfirstlane(s0 + (slot >> 6));
    const int lane = slot & 63;
    const int r = S.rows[s * kSlice + lane];
    const int e0 = S.soff[s], w = (S.soff[s + 1] - e0) >> 6;
    const int sdo = S.sdoff[s];
    const int rb = r >= 0 ? r : S.rows[s * kSlice];
    const TX sum = sell_dot<8, 0, TV, TX, TX>(S.val + e0 + lane, S.col + e0 + lane, w, x, sdo, S.dtab, rb);
    if (r >= 0) epi(r, sum);
  }
  __syncthreads();
}

// ls over ex on the base level: first = |b|; converged = first < abslimit;
// repeat m: x = damp inv b; b -= A x; c += x; stop on abslimit or red.
template <typename T>
__device__ void wg_base_solve(const BaseArgs &ba, const DSell<double> &A,
                              const uint8_t *__restrict__ vclass, T *__restrict__ b,
                              T *__restrict__ c, int cmode, double *sh, double *red) {
  const int n = ba.n;
  double *bl = sh, *xl = sh + n, *cl = sh + 2 * n;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += kCoarseThreads) {
    bl[i] = (double)b[i];
    cl[i] = cmode ? (double)c[i] : 0.0;
  }
  __syncthreads();
  auto norm = [&]() {
    double s = 0.0;
    for (int i = tid; i < n; i += kCoarseThreads)
      if (vclass[i] >= 2) s += bl[i] * bl[i];
    return sqrt(wg_sum(s, red));
  };
  const double first = norm();
  double target = first * ba.red;
  if (target == 0.0) target = ba.red;
  bool conv = fabs(first) < fabs(ba.abslimit);
  for (int it = 0; it < ba.maxiter && !conv && !ba.failed; ++it) {
    // x = damp * inv b: one wavefront per row, lanes across the columns
    const int wave = tid >> 6, lane = tid & 63;
    for (int i = wave; i < n; i += kCoarseThreads / 64) {
      const double *row = ba.inv + (size_t)i * n;
      double s = 0.0;
      for (int j = lane; j < n; j += 64) s += row[j] * bl[j];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
      if (lane == 0) xl[i] = s * ba.damp;
    }
    __syncthreads();
    wg_sell(A, 0, A.nslices, xl, [&](int r, double s) { bl[r] -= s; });
    for (int i = tid; i < n; i += kCoarseThreads) cl[i] += xl[i];
    const double last = norm();
    conv = fabs(last) < fabs(ba.abslimit) || fabs(last) < fabs(target);
  }
  __syncthreads();
  for (int i = tid; i < n; i += kCoarseThreads) {
    b[i] = (T)bl[i];
    c[i] = (T)cl[i];
  }
  __syncthreads();
}

template <typename T>
__device__ void wg_smooth(const CLevel<T> &L, const CycleParams &cp, T *__restrict__ c,
                          T *__restrict__ b, int cmode) {
  const T damp = (T)cp.damp;
  T *__restrict__ t = L.t;
  if (cp.smoother == 0) {
    for (int i = threadIdx.x; i < L.n; i += kCoarseThreads)
      t[i] = L.vclass[i] < 3 ? (T)0 : (b[i] / L.diag[i]) * damp;
    __syncthreads();
    wg_sell(L.A, 0, L.A.nslices, t, [&](int r, T s) {
      b[r] -= s;
      if (cmode) c[r] += t[r];
      else c[r] = t[r];
    });
  } else {
    for (int k = 0; k < L.ncolors; ++k)
      wg_sell(L.L, L.color_slices[k], L.color_slices[k + 1], t, [&](int r, T s) {
        t[r] = L.vclass[r] < 3 ? (T)0 : (b[r] - s) / L.diag[r];
      });
    wg_sell(L.U, 0, L.U.nslices, t, [&](int r, T s) {
      const T br = b[r];
      b[r] = L.vclass[r] >= 3 ? br - damp * (br + s) : br - damp * s;
      const T tt = damp * t[r];
      if (cmode) c[r] += tt;
      else c[r] = tt;
    });
  }
}

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_coarse_cycle(CoarseArgs<T> a) {
  extern __shared__ double sh[];
  __shared__ double red[kCoarseThreads / 64];
  const CycleParams cp = a.cp;
  int phase[kMaxLevels], gcount[kMaxLevels], cm[kMaxLevels];
  int l = a.top;
  phase[l] = 0;
  cm[l] = a.cmode_top;
  while (l <= a.top) {
    const CLevel<T> &L = a.lv[l];
    T *c = l == a.top ? a.c_top : L.c;
    T *b = l == a.top ? a.b_top : L.b;
    if (l <= a.bl) {
      wg_base_solve<T>(a.base, a.baseA, a.base_vclass, b, c, cm[l], sh, red);
      ++l;
      continue;
    }
    const CLevel<T> &C = a.lv[l - 1];
    if (phase[l] == 0) {
      for (int i = 0; i < cp.nu1; ++i) wg_smooth<T>(L, cp, c, b, i == 0 ? cm[l] : 1);
      if (cp.nu1 == 0 && cm[l] == 0) {
        for (int i = threadIdx.x; i < L.n; i += kCoarseThreads) c[i] = (T)0;
        __syncthreads();
      }
      wg_sell(L.PT, 0, L.PT.nslices, b, [&](int I, T s) {
        const T base = C.vnclass[I] >= 2 ? (T)0 : C.b[I];
        C.b[I] = (C.skip[I] & 1u) ? base : base + s;
      });
      gcount[l] = 0;
      phase[l] = 1;
      if (cp.gamma > 0) {
        --l;
        phase[l] = 0;
        cm[l] = 0;
        continue;
      }
      for (int i = threadIdx.x; i < C.n; i += kCoarseThreads) C.c[i] = (T)0;
      __syncthreads();
    } else {
      if (++gcount[l] < cp.gamma) {
        --l;
        phase[l] = 0;
        cm[l] = 1;
        continue;
      }
    }
    const T ld = (T)cp.lmgc_damp;
    wg_sell(L.P, 0, L.P.nslices, C.c, [&](int v, T s) {
      const T tv = (L.skip[v] & 1u) ? (T)0 : ld * s;
      L.t[v] = tv;
      c[v] += tv;
    });
    wg_sell(L.A, 0, L.A.nslices, L.t, [&](int r, T s) { b[r] -= s; });
    for (int i = 0; i < cp.nu2; ++i) wg_smooth<T>(L, cp, c, b, 1);
    if (l == a.top) break;
    ++l;
  }
}

// ----------------------------------------------------------------- launchers
int reduce_parts(int n);
static inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }
static inline int nblk_slices(int nsl) { return (nsl + kBlock / kSlice - 1) / (kBlock / kSlice); }
static inline dim3 grid2(int gx, int nrhs) { return dim3(gx, nrhs < 1 ? 1 : nrhs); }
// grid of a launch over slices [slice0, slice0 + nsl) of S: with the colour
// map (whole colour-grouped slab, one column), 8 x the busiest XCD's blocks
template <typename TV>
static dim3 slab_grid(const DSell<TV> &S, int slice0, int nsl, int nrhs) {
  if (S.ncb > 0 && nrhs <= 1 && slice0 == 0 && nsl == S.nslices) {
    int most = 0;
    for (int x = 0; x < 8; ++x) {
      int tot = 0;
      for (int k = 0; k < S.ncb; ++k) {
        const int nk = S.cb[k + 1] - S.cb[k];
        tot += ((x + 1) * nk >> 3) - (x * nk >> 3);
      }
      most = std::max(most, tot);
    }
    return dim3(8 * most, 1);
  }
  return grid2(nblk_slices(nsl), nrhs);
}

// widest row of a launch -> chunk MW (one round of loads up to 28 entries)
#define W_DISPATCH(wmax, ...)     \
  do {                            \
    if ((wmax) <= 8 || (wmax) > 28) { \
      constexpr int MWV = 8;      \
      __VA_ARGS__;                \
    } else if ((wmax) <= 16) {    \
      constexpr int MWV = 16;     \
      __VA_ARGS__;                \
    } else {                      \
      constexpr int MWV = 28;     \
      __VA_ARGS__;                \
    }                             \
  } while (0)

// block vectors: BK = 2, 3 or 4 components
#define BK_DISPATCH(BB, ...)   \
  do {                         \
    if ((BB) == 3) {           \
      constexpr int BK = 3;    \
      __VA_ARGS__;             \
    } else if ((BB) == 4) {    \
      constexpr int BK = 4;    \
      __VA_ARGS__;             \
    } else {                   \
      constexpr int BK = 2;    \
      __VA_ARGS__;             \
    }                          \
  } while (0)

#define NT_DISPATCH(nt, ...) \
  do {                       \
    if (nt) {                \
      constexpr int NTV = 1; \
      __VA_ARGS__;           \
    } else {                 \
      constexpr int NTV = 0; \
      __VA_ARGS__;           \
    }                        \
  } while (0)

template <typename TV, typename TX, typename TY>
void launch_spmv(hipStream_t st, const DSell<TV> &A, int slice0, int nsl, const TX *x, TY *y,
                 TY *c, int mode, bool nt, Batch bt) {
  if (nsl <= 0) return;
  const dim3 g = slab_grid(A, slice0, nsl, bt.nrhs), blk(kBlock);
  if (bt.b > 1) {
    BK_DISPATCH(bt.b, NT_DISPATCH(nt, switch (mode) {
      case 0: k_spmv_bk<BK, TV, TX, TY, 0, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
      case 1: k_spmv_bk<BK, TV, TX, TY, 1, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
      case 2: k_spmv_bk<BK, TV, TX, TY, 2, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
      default: k_spmv_bk<BK, TV, TX, TY, 3, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
    }));
    return;
  }
  W_DISPATCH(A.wmax, NT_DISPATCH(nt, switch (mode) {
    case 0: k_spmv<MWV, TV, TX, TY, 0, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
    case 1: k_spmv<MWV, TV, TX, TY, 1, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
    case 2: k_spmv<MWV, TV, TX, TY, 2, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
    default: k_spmv<MWV, TV, TX, TY, 3, NTV><<<g, blk, 0, st>>>(A, slice0, nsl, x, y, c, bt.vs); break;
  }));
}
template void launch_spmv<double, double, double>(hipStream_t, const DSell<double> &, int, int,
                                                  const double *, double *, double *, int, bool,
                                                  Batch);
template void launch_spmv<float, float, float>(hipStream_t, const DSell<float> &, int, int,
                                               const float *, float *, float *, int, bool, Batch);

template <typename TV, typename T>
void launch_spmv_c0(hipStream_t st, const DSell<TV> &A, const T *x, T *y, const T *diag, T *u,
                    int ncs0, bool nt, Batch bt) {
  if (A.nslices <= 0) return;
  const dim3 g = slab_grid(A, 0, A.nslices, bt.nrhs);
  W_DISPATCH(A.wmax, NT_DISPATCH(nt, k_spmv_c0<MWV, TV, T, NTV><<<g, kBlock, 0, st>>>(
                                         A, A.nslices, x, y, diag, u, ncs0, bt.vs)));
}
template void launch_spmv_c0<double, double>(hipStream_t, const DSell<double> &, const double *,
                                             double *, const double *, double *, int, bool, Batch);
template void launch_spmv_c0<float, float>(hipStream_t, const DSell<float> &, const float *, float *,
                                           const float *, float *, int, bool, Batch);

void launch_ir_update(hipStream_t st, const DSell<double> &A, const float *c32, double *r,
                      double *x, float *r32, const uint8_t *vclass, double *part, int *nparts,
                      bool nt, Batch bt) {
  const int g = nblk_slices(A.nslices);   // partial sums: one per slice block
  *nparts = g;
  if (g == 0) return;
  if (bt.b > 1) {
    BK_DISPATCH(bt.b, NT_DISPATCH(nt, k_ir_update_bk<BK, NTV><<<grid2(g, bt.nrhs), kBlock, 0, st>>>(
                                          A, A.nslices, c32, r, x, r32, vclass, part, bt.vs)));
    return;
  }
  W_DISPATCH(A.wmax, NT_DISPATCH(nt, k_ir_update<MWV, NTV><<<slab_grid(A, 0, A.nslices, bt.nrhs), kBlock, 0, st>>>(
                                         A, A.nslices, c32, r, x, r32, vclass, part, bt.vs)));
}

template <typename T>
void launch_jac(hipStream_t st, int n, const T *b, const T *diag, const uint8_t *vclass, T damp,
                T *t, Batch bt) {
  if (n <= 0) return;
  if (bt.b > 1)
    BK_DISPATCH(bt.b, k_jac_bk<BK, T><<<grid2(nblk(n), bt.nrhs), kBlock, 0, st>>>(n, b, diag, vclass, damp, t,
                                                                                 bt.vs, bt.dms));
  else k_jac<T><<<grid2(nblk(n), bt.nrhs), kBlock, 0, st>>>(n, b, diag, vclass, damp, t, bt.vs, bt.dms);
}
template void launch_jac<double>(hipStream_t, int, const double *, const double *,
                                 const uint8_t *, double, double *, Batch);
template void launch_jac<float>(hipStream_t, int, const float *, const float *, const uint8_t *,
                                float, float *, Batch);

// Chunk of a colour sweep's row product.  FP64: the colour's widest row
// rounded up to 8/16/24/28 entries, so a row costs one round of loads with
// few clamped repeats (one launch is one colour, whose rows have nearly one
// width); measured 1.7 % faster than chunks of 8 over the 129^3 FP64 solve.
// FP32: chunks of 8 (the wider chunks measured 3-4 % slower over the mixed
// solve, in the plain and the fused sweeps alike).
template <typename T>
static int gs_chunk(int w) {
  if (!std::is_same_v<T, double>) return 8;
  return w <= 8 ? 8 : w <= 16 ? 16 : w <= 24 ? 24 : 28;
}

// Colour sweeps run in workgroups of one wavefront (one slice): measured on
// the 129^3 solves against 128/256/512/1024 threads, the FP64 solve 4 % and
// the mixed one 2 % faster than with 256, the 65^3 elasticity solves (3x3
// blocks) 6-8 % (a colour's launch is short, and single-wave groups start
// and retire independently).  The launches over a whole level (SpMV, GS
// defect, interpolation) gain nothing from it (measured) and keep 256.
constexpr int kSweepBlock = 64;
#define GSB_DISPATCH(NSL, ...)                                                  \
  {                                                                             \
    constexpr int BSV = kSweepBlock;                                            \
    const dim3 g = grid2(((NSL) + BSV / kSlice - 1) / (BSV / kSlice), bt.nrhs); \
    __VA_ARGS__;                                                                \
  }

template <typename TV, typename T>
void launch_gs_color(hipStream_t st, const DSell<TV> &L, int slice0, int nsl, int wmax, const T *b,
                     const T *diag, const uint8_t *vclass, T *v, bool nt, Batch bt) {
  if (nsl <= 0) return;
  if (bt.b > 1) {
    BK_DISPATCH(bt.b, GSB_DISPATCH(nsl, NT_DISPATCH(nt, k_gs_color_bk<BK, TV, T, NTV, BSV><<<g, BSV, 0, st>>>(
                                                            L, slice0, nsl, b, diag, vclass, v, bt.vs))));
    return;
  }
  if constexpr (!std::is_same_v<T, double>) {
    GSB_DISPATCH(nsl, NT_DISPATCH(nt, k_gs_color<8, TV, T, NTV, BSV><<<g, BSV, 0, st>>>(L, slice0, nsl, b, diag, vclass, v, bt.vs)));
  } else {
    switch (gs_chunk<T>(wmax)) {
      case 28:
        GSB_DISPATCH(nsl, NT_DISPATCH(nt, k_gs_color<28, TV, T, NTV, BSV><<<g, BSV, 0, st>>>(L, slice0, nsl, b, diag, vclass, v, bt.vs)));
        break;
      case 24:
        GSB_DISPATCH(nsl, NT_DISPATCH(nt, k_gs_color<24, TV, T, NTV, BSV><<<g, BSV, 0, st>>>(L, slice0, nsl, b, diag, vclass, v, bt.vs)));
        break;
      case 16:
        GSB_DISPATCH(nsl, NT_DISPATCH(nt, k_gs_color<16, TV, T, NTV, BSV><<<g, BSV, 0, st>>>(L, slice0, nsl, b, diag, vclass, v, bt.vs)));
        break;
      default:
        GSB_DISPATCH(nsl, NT_DISPATCH(nt, k_gs_color<8, TV, T, NTV, BSV><<<g, BSV, 0, st>>>(L, slice0, nsl, b, diag, vclass, v, bt.vs)));
    }
  }
}
template void launch_gs_color<double, double>(hipStream_t, const DSell<double> &, int, int, int,
                                              const double *, const double *, const uint8_t *,
                                              double *, bool, Batch);
template void launch_gs_color<float, float>(hipStream_t, const DSell<float> &, int, int, int,
                                            const float *, const float *, const uint8_t *,
                                            float *, bool, Batch);

// Colours of at most this many slices run the folded sweep on two
// wavefronts per slice (launches there are bound by the chain of dependent
// loads: -0.2 to -0.7 us each on 33^3 and 65^3 levels); wider ones on one
// (on 4096+ slices the second wavefront costs more than it hides: +0.1 to
// +1.8 us per launch at 129^3).
constexpr int kSplitMaxSlices = 1024;

template <int CHU, int CHL, typename TV, typename T, int NTV>
void fused_launch(hipStream_t st, int nrhs, const DSell<TV> &L, const DSell<TV> &U, int slice0, int nsl,
                  T *b, const T *diag, T damp, T *v, T *c, int cmode, size_t vs, bool split) {
  const Batch bt{nrhs, vs, 1};
  if (split && nsl <= kSplitMaxSlices) {
    const dim3 g = grid2(nsl, nrhs);
    if (cmode == 0)
      k_gs_color_fused2<CHU, CHL, TV, T, 0, NTV><<<g, 2 * kSlice, 0, st>>>(L, U, slice0, nsl, b, diag, damp, v, c, vs);
    else
      k_gs_color_fused2<CHU, CHL, TV, T, 1, NTV><<<g, 2 * kSlice, 0, st>>>(L, U, slice0, nsl, b, diag, damp, v, c, vs);
    return;
  }
  GSB_DISPATCH(nsl, if (cmode == 0)
    k_gs_color_fused<CHU, CHL, TV, T, 0, NTV, BSV><<<g, BSV, 0, st>>>(L, U, slice0, nsl, b, diag, damp, v, c, vs);
  else
    k_gs_color_fused<CHU, CHL, TV, T, 1, NTV, BSV><<<g, BSV, 0, st>>>(L, U, slice0, nsl, b, diag, damp, v, c, vs));
}

template <typename TV, typename T>
void launch_gs_color_fused(hipStream_t st, const DSell<TV> &L, const DSell<TV> &U, int slice0,
                           int nsl, int wmax_l, int wmax_u, T *b, const T *diag, T damp, T *v,
                           T *c, int cmode, bool nt, Batch bt, bool split) {
  if (nsl <= 0) return;
  if (bt.b > 1) {
    BK_DISPATCH(bt.b, GSB_DISPATCH(nsl, NT_DISPATCH(nt, if (cmode == 0) k_gs_color_fused_bk<BK, TV, T, 0, NTV, BSV><<<g, BSV, 0, st>>>(
                        L, U, slice0, nsl, b, diag, damp, v, c, bt.vs);
                else k_gs_color_fused_bk<BK, TV, T, 1, NTV, BSV><<<g, BSV, 0, st>>>(
                        L, U, slice0, nsl, b, diag, damp, v, c, bt.vs))));
    return;
  }
  // one chunk per part sized to the colour's widest row (a colour's upper
  // and lower widths add up to the off-diagonal row, so one of the two is
  // narrow): pairs with both parts wider than 8 fall back to 16 x 16
  if constexpr (!std::is_same_v<T, double>) {
    NT_DISPATCH(nt, (fused_launch<8, 8, TV, T, NTV>(st, bt.nrhs, L, U, slice0, nsl, b, diag, damp, v, c, cmode,
                                                    bt.vs, split)));
  } else {
    int cu = gs_chunk<T>(wmax_u), cl = gs_chunk<T>(wmax_l);
    if (cu > 8 && cl > 8) cu = cl = 16;
    const int code = cu * 100 + cl;
#define FUSED_CASE(A, B)                                                                      \
  case A * 100 + B:                                                                           \
    NT_DISPATCH(nt, (fused_launch<A, B, TV, T, NTV>(st, bt.nrhs, L, U, slice0, nsl, b, diag, damp, v, c, \
                                                    cmode, bt.vs, split)));                           \
    break;
    switch (code) {
      FUSED_CASE(28, 8)
      FUSED_CASE(24, 8)
      FUSED_CASE(16, 8)
      FUSED_CASE(8, 16)
      FUSED_CASE(8, 24)
      FUSED_CASE(8, 28)
      FUSED_CASE(16, 16)
      default:
        NT_DISPATCH(nt, (fused_launch<8, 8, TV, T, NTV>(st, bt.nrhs, L, U, slice0, nsl, b, diag, damp, v, c, cmode,
                                                        bt.vs, split)));
    }
#undef FUSED_CASE
  }
}
template void launch_gs_color_fused<double, double>(hipStream_t, const DSell<double> &,
                                                    const DSell<double> &, int, int, int, int,
                                                    double *, const double *, double, double *,
                                                    double *, int, bool, Batch, bool);
template void launch_gs_color_fused<float, float>(hipStream_t, const DSell<float> &,
                                                  const DSell<float> &, int, int, int, int, float *,
                                                  const float *, float, float *, float *, int,
                                                  bool, Batch, bool);

template <typename TV, typename T>
void launch_gs_defect(hipStream_t st, const DSell<TV> &U, const T *v, T *b, const uint8_t *vclass,
                      T damp, T *c, int cmode, T *xout, bool nt, Batch bt) {
  if (U.nslices <= 0) return;
  const dim3 g = slab_grid(U, 0, U.nslices, bt.nrhs);
  if (bt.b > 1) {
    BK_DISPATCH(bt.b, NT_DISPATCH(nt, if (cmode == 0) k_gs_defect_bk<BK, TV, T, 0, NTV><<<g, kBlock, 0, st>>>(
                        U, U.nslices, v, b, vclass, damp, c, xout, bt.vs);
                else k_gs_defect_bk<BK, TV, T, 1, NTV><<<g, kBlock, 0, st>>>(
                        U, U.nslices, v, b, vclass, damp, c, xout, bt.vs)));
    return;
  }
  // chunks of 8 entries: the upper widths of the colours run from the whole
  // off-diagonal row down to 0, and a wider chunk mostly loads clamped
  // repeats on the narrow colours (measured slower for 28 and 16)
  NT_DISPATCH(nt, if (cmode == 0) k_gs_defect<8, TV, T, 0, NTV><<<g, kBlock, 0, st>>>(
                      U, U.nslices, v, b, vclass, damp, c, xout, bt.vs);
              else k_gs_defect<8, TV, T, 1, NTV><<<g, kBlock, 0, st>>>(
                      U, U.nslices, v, b, vclass, damp, c, xout, bt.vs));
}
template void launch_gs_defect<double, double>(hipStream_t, const DSell<double> &,
                                               const double *, double *, const uint8_t *, double,
                                               double *, int, double *, bool, Batch);
template void launch_gs_defect<float, float>(hipStream_t, const DSell<float> &, const float *,
                                             float *, const uint8_t *, float, float *, int,
                                             float *, bool, Batch);

template <typename T>
int launch_restrict(hipStream_t st, const DSell<T> &PT, const uint32_t *cskip,
                    const uint8_t *cvnclass, const uint8_t *cvclass, const T *bf, T damp, T *bc,
                    double *part, bool nt, int nrhs, size_t vsf, size_t vsc, int b,
                    const uint8_t *cc0, const T *cdiag, T *ct, const int *cpos, T *cbc,
                    size_t cdms) {
  if (PT.nslices <= 0) return 0;
  const int g = nblk_slices(PT.nslices);
  if (b > 1 && PT.bb > 1) {
    BK_DISPATCH(b, NT_DISPATCH(nt, k_restrict_bk<BK, T, NTV, true><<<grid2(g, nrhs), kBlock, 0, st>>>(
                                       PT, PT.nslices, cskip, cvnclass, cvclass, bf, damp, bc, part, vsf, vsc)));
    return g;
  }
  if (b > 1) {
    BK_DISPATCH(b, NT_DISPATCH(nt, k_restrict_bk<BK, T, NTV><<<grid2(g, nrhs), kBlock, 0, st>>>(
                                       PT, PT.nslices, cskip, cvnclass, cvclass, bf, damp, bc, part, vsf, vsc)));
    return g;
  }
  W_DISPATCH(PT.wmax, NT_DISPATCH(nt, k_restrict<MWV, T, NTV><<<grid2(g, nrhs), kBlock, 0, st>>>(
                                          PT, PT.nslices, cskip, cvnclass, cvclass, bf, damp, bc,
                                          part, vsf, vsc, cc0, cdiag, ct, cpos, cbc, cdms)));
  return g;
}
template int launch_restrict<double>(hipStream_t, const DSell<double> &, const uint32_t *,
                                     const uint8_t *, const uint8_t *, const double *, double,
                                     double *, double *, bool, int, size_t, size_t, int,
                                     const uint8_t *, const double *, double *, const int *,
                                     double *, size_t);
template int launch_restrict<float>(hipStream_t, const DSell<float> &, const uint32_t *,
                                    const uint8_t *, const uint8_t *, const float *, float,
                                    float *, double *, bool, int, size_t, size_t, int,
                                    const uint8_t *, const float *, float *, const int *, float *,
                                    size_t);

// --------------------------------------------------- base solve (GEMV)
// A quarter of a row of the inverse times b: columns j0 + 256 m (j0 = 64 q +
// lane for wavefront q of the row's workgroup), four streams in flight; b
// read as bb[cp[j]], or (COL) as bc[j] when it is already in column order
template <bool COL, typename T, typename TI>
__device__ __forceinline__ double base_row_quarter(const TI *__restrict__ rr,
                                                   const T *__restrict__ bb,
                                                   const int *__restrict__ cp, int ncols, int j0) {
  auto bj = [&](int j) { return COL ? (double)bb[j] : (double)bb[cp[j]]; };
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int j = j0;
  for (; j + 768 < ncols; j += 1024) {
    a0 += (double)ld<1>(rr + j) * bj(j);
    a1 += (double)ld<1>(rr + j + 256) * bj(j + 256);
    a2 += (double)ld<1>(rr + j + 512) * bj(j + 512);
    a3 += (double)ld<1>(rr + j + 768) * bj(j + 768);
  }
  for (; j < ncols; j += 256) a0 += (double)ld<1>(rr + j) * bj(j);
  return (a0 + a1) + (a2 + a3);
}


// ls over ex on a scalar base level:
//   cur = |b| from the block partials of the defect norm;
//   converged: cur < abslimit (it == 0), cur < abslimit or cur < red*first;
//   x = converged ? 0 : damp * inv b;  c (+)= x;  xout = x.
// One workgroup (4 wavefronts) per row, each wavefront a quarter of the
// row's columns (one wavefront per row kept too few loads in flight: <= 3.2
// TB/s on a 3375 x 3375 inverse); wavefront 0 also reduces the stop test's
// norm, then adds the four shares and writes.
template <typename T, typename TI>
__global__ __launch_bounds__(kBlock) void k_base_gemv(int n, const TI *__restrict__ inv,
                                                       const int *__restrict__ rowmap,
                                                       const int *__restrict__ cp, int ncols,
                                                       const T *__restrict__ b,
                                                       const T *__restrict__ bc,
                                                       T *__restrict__ c, T *__restrict__ xout,
                                                       int cmode, double damp, double abslimit,
                                                       double red, const double *__restrict__ part,
                                                       int nparts, double *__restrict__ first,
                                                       int it, size_t vs, size_t ims) {
  __shared__ double wsum[kBlock / 64];
  const int row = blockIdx.x;
  const int q = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int cj = blockIdx.y;
  inv += ims * cj;
  const size_t o = (size_t)cj * vs;
  part += (size_t)cj * nparts;
  const int k = rowmap[row];
  double x;
  if (k < 0) x = (q == 0 && lane == 0) ? (double)b[o + row] : 0.0;   // unit row
  else if (bc) x = base_row_quarter<true, T, TI>(inv + (size_t)k * n, bc + o, cp, ncols, 64 * q + lane);
  else x = base_row_quarter<false, T, TI>(inv + (size_t)k * n, b + o, cp, ncols, 64 * q + lane);
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if (lane == 0) wsum[q] = x;
  // the stop test's norm, reduced while the other wavefronts finish
  double s = 0.0;
  if (q == 0) {
    for (int i = lane; i < nparts; i += 64) s += part[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  }
  __syncthreads();
  if (q != 0) return;
  const double cur = sqrt(s);
  bool conv;
  if (it == 0) {
    conv = fabs(cur) < fabs(abslimit);
    if (row == 0 && lane == 0) first[cj] = cur;
  } else {
    double target = first[cj] * red;
    if (target == 0.0) target = red;
    conv = fabs(cur) < fabs(abslimit) || fabs(cur) < fabs(target);
  }
  x = conv ? 0.0 : damp * ((wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
  if (lane == 0) {
    const T xv = (T)x;
    if (!conv || cmode == 0) c[o + row] = cmode ? c[o + row] + xv : xv;
    if (xout) xout[o + row] = xv;
  }
}

template <typename T, typename TI>
void launch_base_gemv(hipStream_t st, int n, const TI *inv, const int *rowmap, const int *cp,
                      int ncols, const T *b, const T *bc, T *c, T *xout, int cmode, double damp,
                      double abslimit, double red, const double *part, int nparts, double *first,
                      int it, Batch bt) {
  if (n <= 0) return;
  // the stop tests read nc norm entries per column (components, or
  // (type, component) pairs); n counts scalar unknowns
  if (bt.nc() > 1) {   // one wavefront per row
    const int g = (n * 64 + kBlock - 1) / kBlock;
    BK_DISPATCH(bt.nc(), k_base_gemv_bk<BK, T, TI><<<grid2(g, bt.nrhs), kBlock, 0, st>>>(
                          n, inv, rowmap, cp, ncols, b, c, xout, cmode, damp, abslimit, red, part, nparts,
                          first, it, bt.vs, bt.ims, bt.idents, bt.pdamp));
  } else {           // one workgroup per row
    k_base_gemv<T, TI><<<grid2(n, bt.nrhs), kBlock, 0, st>>>(n, inv, rowmap, cp, ncols, b, bc, c, xout,
                                                             cmode, damp, abslimit, red, part,
                                                             nparts, first, it, bt.vs, bt.ims);
  }
}
// ls over jac on the base level, after the Jacobi step formed x: the stop
// test of k_base_gemv_bk on the current defect's partial norms (nc entries:
// components, or (type, component) pairs; grouped by VD_IDENT when idents
// >= 0), then x = 0 once it holds, and c (+)= x.  n scalar unknowns.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_base_gate(int n, int nc, T *__restrict__ x, T *__restrict__ c,
                                                       int cmode, double abslimit, double red,
                                                       const double *__restrict__ part, int nparts,
                                                       double *__restrict__ first, int it, size_t vs,
                                                       int idents) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int cj = blockIdx.y;
  double cur[UGMG_MAXB] = {}, target[UGMG_MAXB] = {};
  for (int q = 0; q < nc; ++q) {
    const double *p = part + ((size_t)cj * nc + q) * nparts;
    double s = 0.0;
    for (int k = lane; k < nparts; k += 64) s += p[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    cur[q] = sqrt(s);
  }
  if (it == 0 && i == 0)
    for (int q = 0; q < nc; ++q) first[(size_t)cj * nc + q] = cur[q];
  for (int q = 0; q < nc; ++q) {
    target[q] = it == 0 ? 0.0 : first[(size_t)cj * nc + q] * red;
    if (target[q] == 0.0) target[q] = red;
  }
  bool below_abs = true, below_red = true;
  if (idents < 0) {
    for (int q = 0; q < nc; ++q) {
      below_abs &= fabs(cur[q]) < fabs(abslimit);
      below_red &= fabs(cur[q]) < fabs(target[q]);
    }
  } else {   // sc_cmp with VD_IDENT (npscan.cc:1033-1048)
    for (int g = 0; g < nc; ++g) {
      if (((idents >> (4 * g)) & 15) != g) continue;
      double sx = 0.0, sa = 0.0, st = 0.0;
      for (int j = 0; j < nc; ++j)
        if (((idents >> (4 * j)) & 15) == g) {
          sx += cur[j] * cur[j];
          sa += abslimit * abslimit;
          st += target[j] * target[j];
        }
      below_abs &= sx < sa;
      below_red &= sx < st;
    }
  }
  if (it == 0) below_red = false;
  const bool conv = below_abs || below_red;
  if (i >= n) return;
  const size_t k = (size_t)cj * vs + i;
  const T xv = conv ? (T)0 : x[k];
  x[k] = xv;
  if (!conv || cmode == 0) c[k] = cmode ? c[k] + xv : xv;
}
template <typename T>
void launch_base_gate(hipStream_t st, int n, T *x, T *c, int cmode, double abslimit, double red,
                      const double *part, int nparts, int it, double *first, Batch bt) {
  if (n <= 0) return;
  k_base_gate<T><<<grid2(nblk(n), bt.nrhs), kBlock, 0, st>>>(n, bt.nc(), x, c, cmode, abslimit, red, part,
                                                              nparts, first, it, bt.vs, bt.idents);
}
template void launch_base_gate<double>(hipStream_t, int, double *, double *, int, double, double,
                                       const double *, int, int, double *, Batch);
template void launch_base_gate<float>(hipStream_t, int, float *, float *, int, double, double,
                                      const double *, int, int, double *, Batch);

template void launch_base_gemv<double, double>(hipStream_t, int, const double *, const int *,
                                               const int *, int, const double *, const double *, double *,
                                               double *, int, double, double, double,
                                               const double *, int, double *, int, Batch);
template void launch_base_gemv<float, double>(hipStream_t, int, const double *, const int *,
                                              const int *, int, const float *, const float *, float *, float *, int,
                                              double, double, double, const double *, int, double *,
                                              int, Batch);
template void launch_base_gemv<float, float>(hipStream_t, int, const float *, const int *,
                                             const int *, int, const float *, const float *, float *, float *, int,
                                             double, double, double, const double *, int, double *,
                                             int, Batch);

// members: d64[j] = src[map[j]] (0 where map[j] < 0) and its FP32 copy
__global__ __launch_bounds__(kBlock) void k_member_gather(size_t n, const int *__restrict__ map,
                                                           const double *__restrict__ src,
                                                           double *__restrict__ d64,
                                                           float *__restrict__ d32) {
  const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const int k = map[j];
  const double v = k >= 0 ? src[k] : 0.0;
  d64[j] = v;
  d32[j] = (float)v;
}
void launch_member_gather(hipStream_t st, size_t n, const int *map, const double *src, double *d64,
                          float *d32) {
  if (n > 0)
    k_member_gather<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(n, map, src, d64, d32);
}

// entry (act[k], cp[j]) of the n x n inverse -> (k, j) of the FP64 and FP32
// copies (rows act, columns in the order cp)
__global__ __launch_bounds__(kBlock) void k_base_compact(int nact, int n, const int *__restrict__ act,
                                                          const int *__restrict__ cp,
                                                          const double *__restrict__ inv,
                                                          double *__restrict__ c64,
                                                          float *__restrict__ c32) {
  const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (size_t)nact * n) return;
  const int k = (int)(e / (size_t)n), j = (int)(e % (size_t)n);
  const double v = inv[(size_t)act[k] * n + cp[j]];
  c64[e] = v;
  c32[e] = (float)v;
}
void launch_base_compact(hipStream_t st, int nact, int n, const int *act, const int *cp,
                         const double *inv, double *c64, float *c32) {
  const size_t ne = (size_t)nact * n;
  if (ne > 0)
    k_base_compact<<<(unsigned)((ne + kBlock - 1) / kBlock), kBlock, 0, st>>>(nact, n, act, cp, inv, c64, c32);
}

template <typename T>
int launch_dot_part(hipStream_t st, int n, const T *x, const uint8_t *vclass, double *part,
                    Batch bt) {
  const int g = reduce_parts(n);
  if (bt.vt)
    k_dot_part_typed<T, 0><<<grid2(g, bt.nrhs), kBlock, 0, st>>>(n, bt.b, bt.nc(), x, nullptr, vclass, bt.vt,
                                                                 part, bt.vs);
  else if (bt.b > 1)
    BK_DISPATCH(bt.b, k_dot_part_bk<BK, T, 0><<<grid2(g, bt.nrhs), kBlock, 0, st>>>(n, x, nullptr, vclass, part,
                                                                                   bt.vs));
  else k_dot_part<T, 0><<<grid2(g, bt.nrhs), kBlock, 0, st>>>(n, x, nullptr, vclass, part, bt.vs);
  return g;
}

template <typename T>
void launch_interp(hipStream_t st, const DSell<T> &P, const uint32_t *fskip, const T *cc, T damp,
                   T *t, T *c, bool nt, int nrhs, size_t vsf, size_t vsc, int b) {
  if (P.nslices <= 0) return;
  const dim3 g = grid2(nblk_slices(P.nslices), nrhs);
  if (b > 1 && P.bb > 1) {
    BK_DISPATCH(b, NT_DISPATCH(nt, if (c) k_interp_bk<BK, T, 1, NTV, true><<<g, kBlock, 0, st>>>(P, P.nslices, fskip, cc,
                                                                                       damp, t, c, vsf, vsc);
                               else k_interp_bk<BK, T, 0, NTV, true><<<g, kBlock, 0, st>>>(P, P.nslices, fskip, cc,
                                                                                    damp, t, c, vsf, vsc)));
    return;
  }
  if (b > 1) {
    BK_DISPATCH(b, NT_DISPATCH(nt, if (c) k_interp_bk<BK, T, 1, NTV><<<g, kBlock, 0, st>>>(P, P.nslices, fskip, cc,
                                                                                       damp, t, c, vsf, vsc);
                               else k_interp_bk<BK, T, 0, NTV><<<g, kBlock, 0, st>>>(P, P.nslices, fskip, cc,
                                                                                    damp, t, c, vsf, vsc)));
    return;
  }
  // interpolation rows hold 1, 2, 4 or 8 coarse corners: chunks of 4
  // entries load fewer clamped repeats than 8 (measured faster)
  if (P.wmax <= 8) {
    NT_DISPATCH(nt, if (c) k_interp<4, T, 1, NTV><<<g, kBlock, 0, st>>>(P, P.nslices, fskip, cc, damp, t, c, vsf, vsc);
                else k_interp<4, T, 0, NTV><<<g, kBlock, 0, st>>>(P, P.nslices, fskip, cc, damp, t, c, vsf, vsc));
    return;
  }
  W_DISPATCH(P.wmax, NT_DISPATCH(nt, if (c) k_interp<MWV, T, 1, NTV><<<g, kBlock, 0, st>>>(
                                                P, P.nslices, fskip, cc, damp, t, c, vsf, vsc);
                                 else k_interp<MWV, T, 0, NTV><<<g, kBlock, 0, st>>>(
                                                P, P.nslices, fskip, cc, damp, t, c, vsf, vsc)));
}
template void launch_interp<double>(hipStream_t, const DSell<double> &, const uint32_t *,
                                    const double *, double, double *, double *, bool, int,
                                    size_t, size_t, int);
template void launch_interp<float>(hipStream_t, const DSell<float> &, const uint32_t *,
                                   const float *, float, float *, float *, bool, int, size_t,
                                   size_t, int);

int reduce_parts(int n) {
  int g = nblk(n);
  return g < 1 ? 1 : (g > 1024 ? 1024 : g);
}

template <typename T>
void launch_dot(hipStream_t st, int n, const T *x, const T *y, const uint8_t *vclass,
                double *part, double *out, int do_sqrt, Batch bt) {
  const int g = reduce_parts(n);
  const dim3 gg = grid2(g, bt.nrhs);
  if (bt.vt) {
    if (y) k_dot_part_typed<T, 1><<<gg, kBlock, 0, st>>>(n, bt.b, bt.nc(), x, y, vclass, bt.vt, part, bt.vs);
    else k_dot_part_typed<T, 0><<<gg, kBlock, 0, st>>>(n, bt.b, bt.nc(), x, y, vclass, bt.vt, part, bt.vs);
    k_reduce_final<<<gg.y * bt.nc(), 1024, 0, st>>>(g, part, out, do_sqrt);
    return;
  }
  if (bt.b > 1) {
    BK_DISPATCH(bt.b, if (y) k_dot_part_bk<BK, T, 1><<<gg, kBlock, 0, st>>>(n, x, y, vclass, part, bt.vs);
                      else k_dot_part_bk<BK, T, 0><<<gg, kBlock, 0, st>>>(n, x, y, vclass, part, bt.vs));
    k_reduce_final<<<gg.y * bt.b, 1024, 0, st>>>(g, part, out, do_sqrt);
    return;
  }
  if (y) k_dot_part<T, 1><<<gg, kBlock, 0, st>>>(n, x, y, vclass, part, bt.vs);
  else k_dot_part<T, 0><<<gg, kBlock, 0, st>>>(n, x, y, vclass, part, bt.vs);
  k_reduce_final<<<gg.y, 1024, 0, st>>>(g, part, out, do_sqrt);
}
template void launch_dot<double>(hipStream_t, int, const double *, const double *,
                                 const uint8_t *, double *, double *, int, Batch);
template void launch_dot<float>(hipStream_t, int, const float *, const float *, const uint8_t *,
                                double *, double *, int, Batch);
template int launch_dot_part<double>(hipStream_t, int, const double *, const uint8_t *, double *,
                                     Batch);
template int launch_dot_part<float>(hipStream_t, int, const float *, const uint8_t *, double *,
                                    Batch);

void launch_reduce_final(hipStream_t st, int nparts, const double *part, double *out, int do_sqrt,
                         int nrhs) {
  k_reduce_final<<<nrhs < 1 ? 1 : nrhs, 1024, 0, st>>>(nparts, part, out, do_sqrt);
}

__global__ __launch_bounds__(kBlock) void k_iskip(int n, const uint8_t *__restrict__ vt,
                                                   const uint32_t *__restrict__ skip,
                                                   uint32_t *__restrict__ iskip) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) iskip[i] = vt[i] == 1 ? 0u : skip[i];
}
void launch_iskip(hipStream_t st, int n, const uint8_t *vt, const uint32_t *skip, uint32_t *iskip) {
  if (n > 0) k_iskip<<<nblk(n), kBlock, 0, st>>>(n, vt, skip, iskip);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_comp(int n, int b, Batch::PDamp pd, T *__restrict__ t,
                                                        T *__restrict__ c, int cmode, size_t vs,
                                                        const uint8_t *__restrict__ cls, int clsmin) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n * b) return;
  if (cls && cls[i / b] < clsmin) return;
  const size_t k = (size_t)blockIdx.y * vs + i;
  const T v = t[k] * (T)pd.d[i % b];
  t[k] = v;
  if (c) c[k] = cmode == 0 ? v : c[k] + v;
}
template <typename T>
void launch_scale_comp(hipStream_t st, int n, int b, const double *d, T *t, T *c, int cmode, Batch bt,
                       const uint8_t *cls, int clsmin) {
  if (n <= 0) return;
  Batch::PDamp pd;
  for (int q = 0; q < b && q < 4; ++q) pd.d[q] = d[q];
  pd.vb = b;
  k_scale_comp<T><<<grid2(nblk(n * b), bt.nrhs), kBlock, 0, st>>>(n, b, pd, t, c, cmode, bt.vs, cls, clsmin);
}
template void launch_scale_comp<double>(hipStream_t, int, int, const double *, double *, double *, int, Batch,
                                        const uint8_t *, int);
template void launch_scale_comp<float>(hipStream_t, int, int, const double *, float *, float *, int, Batch,
                                       const uint8_t *, int);

__global__ __launch_bounds__(kBlock) void k_clear_dirichlet(int n, int b, const int *__restrict__ perm,
                                                             const uint32_t *__restrict__ skip,
                                                             double *__restrict__ x) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const uint32_t sk = skip[p];
  if (!sk) return;
  const size_t v = perm ? (size_t)perm[p] : (size_t)p;
  for (int j = 0; j < b; ++j)
    if ((sk >> j) & 1u) x[v * b + j] = 0.0;
}
void launch_clear_dirichlet(hipStream_t st, int n, int b, const int *perm, const uint32_t *skip, double *x) {
  if (n > 0) k_clear_dirichlet<<<nblk(n), kBlock, 0, st>>>(n, b, perm, skip, x);
}

template <typename T>
void launch_fill(hipStream_t st, int n, T *x, T val, Batch bt) {
  n *= bt.b;
  if (n > 0) k_fill<T><<<grid2(nblk(n), bt.nrhs), kBlock, 0, st>>>(n, x, val, bt.vs);
}
template void launch_fill<double>(hipStream_t, int, double *, double, Batch);
template void launch_fill<float>(hipStream_t, int, float *, float, Batch);

template <typename TD, typename TS>
void launch_convert(hipStream_t st, int n, const TS *s, TD *d, int nrhs, size_t vss, size_t vsd) {
  if (n > 0) k_convert<TD, TS><<<grid2(nblk(n), nrhs), kBlock, 0, st>>>(n, s, d, vss, vsd);
}
template void launch_convert<float, double>(hipStream_t, int, const double *, float *, int, size_t,
                                            size_t);
template void launch_convert<double, float>(hipStream_t, int, const float *, double *, int, size_t,
                                            size_t);
template void launch_convert<double, double>(hipStream_t, int, const double *, double *, int,
                                             size_t, size_t);
template void launch_convert<float, float>(hipStream_t, int, const float *, float *, int, size_t,
                                           size_t);

template <typename T>
void launch_add(hipStream_t st, int n, T *x, const T *y, Batch bt) {
  n *= bt.b;
  if (n > 0) k_add<T><<<grid2(nblk(n), bt.nrhs), kBlock, 0, st>>>(n, x, y, bt.vs);
}
template void launch_add<double>(hipStream_t, int, double *, const double *, Batch);

void launch_fgd_class(hipStream_t st, int n, const uint8_t *vc, const uint8_t *vnc, uint8_t *out) {
  if (n > 0) k_fgd_class<<<nblk(n), kBlock, 0, st>>>(n, vc, vnc, out);
}
void launch_copy_rows(hipStream_t st, int n, int b, const uint8_t *cls, const double *src,
                      double *dst) {
  if (n > 0) k_copy_rows<double><<<nblk(n * b), kBlock, 0, st>>>(n * b, b, cls, src, dst);
}
template void launch_add<float>(hipStream_t, int, float *, const float *, Batch);

template <typename TD, typename TS, int SCATTER>
__global__ __launch_bounds__(kBlock) void k_permute(int n, const int *__restrict__ perm,
                                                     const TS *__restrict__ s, TD *__restrict__ d,
                                                     size_t vss, size_t vsd, int b) {
  const int e = blockIdx.x * kBlock + threadIdx.x;   // scalar entry: vector e / b
  if (e >= n * b) return;
  const int p = e, q = perm[e / b] * b + e % b;
  s += (size_t)blockIdx.y * vss;
  d += (size_t)blockIdx.y * vsd;
  if (SCATTER) d[q] = (TD)s[p];
  else d[p] = (TD)s[q];
}

template <typename TD, typename TS>
void launch_gather(hipStream_t st, int n, const int *perm, const TS *s, TD *d, int nrhs,
                   size_t vss, size_t vsd, int b) {
  if (n > 0)
    k_permute<TD, TS, 0><<<grid2(nblk(n * b), nrhs), kBlock, 0, st>>>(n, perm, s, d, vss, vsd, b);
}
template <typename TD, typename TS>
void launch_scatter(hipStream_t st, int n, const int *perm, const TS *s, TD *d, int nrhs,
                    size_t vss, size_t vsd, int b) {
  if (n > 0)
    k_permute<TD, TS, 1><<<grid2(nblk(n * b), nrhs), kBlock, 0, st>>>(n, perm, s, d, vss, vsd, b);
}
#define PERM_INST(TD, TS)                                                                      \
  template void launch_gather<TD, TS>(hipStream_t, int, const int *, const TS *, TD *, int,    \
                                      size_t, size_t, int);                                    \
  template void launch_scatter<TD, TS>(hipStream_t, int, const int *, const TS *, TD *, int,   \
                                       size_t, size_t, int);
PERM_INST(double, double)
PERM_INST(float, double)
PERM_INST(double, float)
PERM_INST(float, float)
#undef PERM_INST

template <typename T>
void launch_coarse_cycle(hipStream_t st, const CoarseArgs<T> &args, int base_n) {
  const size_t shmem = 3 * (size_t)base_n * sizeof(double);
  k_coarse_cycle<T><<<1, kCoarseThreads, shmem, st>>>(args);
}
template void launch_coarse_cycle<double>(hipStream_t, const CoarseArgs<double> &, int);
template void launch_coarse_cycle<float>(hipStream_t, const CoarseArgs<float> &, int);

// ============================================================ BCGSSolver
// Right-preconditioned BiCGStab of np/procs/ls.cc:1864-2062 for several
// right-hand sides at once.  Every column runs its own recurrence: the
// scalars live in BcgsState on the device and change only in
// k_bcgs_scalar (one workgroup per column), so a whole iteration is
// stream-ordered device work with no host round trip.  UG's operation order
// is kept per entry: p = ((beta p) + b) + (-beta omega) v, x = x + alpha q,
// s = b + (-alpha) v, x = x + omega q, b = s + (-omega) t.

int bcgs_nparts(int n) { return reduce_parts(n); }
int bcgs_spmv_nparts(int nslices) { return nblk_slices(nslices) < 1 ? 1 : nblk_slices(nslices); }

// sc_eq (np/udm/npscan.cc:1095-1104) for one component
__device__ __forceinline__ bool sc_eq1(double x, double y, double ac) {
  return !(x < 0.0 || y < 0.0 || fabs(x - y) > ac * sqrt(x * y));
}

__device__ __forceinline__ void record_defect(BcgsState &S, double *defects, int maxdef, int cj,
                                              double d) {
  if (defects && S.nd < maxdef) defects[(size_t)cj * maxdef + S.nd] = d;
  ++S.nd;
}

__global__ __launch_bounds__(1024) void k_bcgs_scalar(int op, int nparts,
                                                       const double *__restrict__ part,
                                                       BcgsState *__restrict__ st, int restart,
                                                       double abslimit, double red,
                                                       double *__restrict__ defects, int maxdef) {
  __shared__ double sh[2][16];
  const int cj = blockIdx.x;
  const double *p0 = part + (size_t)(cj * 2) * nparts;
  const double *p1 = p0 + nparts;
  // RHO after a restart reads b.b (k = 0): r was just set to b
  const bool two = op == BCGS_OMEGA;
  const double *pa = (op == BCGS_RHO && !restart) ? p1 : p0;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 1024) {
    a += pa[i];
    if (two) b += p1[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  a = 0.0;
  b = 0.0;
  for (int i = 0; i < 16; ++i) {
    a += sh[0][i];
    b += sh[1][i];
  }
  BcgsState S = st[cj];
  switch (op) {
    case BCGS_FIRST: {
      const double first = sqrt(a);
      S = BcgsState{};
      S.first = S.last = first;
      S.target = first * red;   // sc_mul_check
      if (S.target == 0.0) S.target = red;
      S.old = -1.0;
      S.abslimit = abslimit;
      S.conv = fabs(first) < fabs(abslimit);
      S.active = !S.conv;
      record_defect(S, defects, maxdef, cj, first);
      break;
    }
    case BCGS_RHO:
      if (!S.active) return;
      if (restart) S.alpha = S.rho = S.omega = 1.0;
      S.rho_new = a;
      if (S.rho != 0.0 && S.omega != 0.0) S.beta = S.rho_new * S.alpha / S.rho / S.omega;
      break;
    case BCGS_ALPHA:
      if (!S.active) return;
      S.alpha = a != 0.0 ? S.rho_new / a : a;
      break;
    case BCGS_RES1:
      if (!S.active) return;
      S.last = sqrt(a);
      ++S.nit;
      record_defect(S, defects, maxdef, cj, S.last);
      if (fabs(S.last) < fabs(S.abslimit) || fabs(S.last) < fabs(S.target)) {
        S.half = 1;
        S.conv = 1;
      }
      break;
    case BCGS_OMEGA:
      if (!S.active || S.half) return;
      S.omega = a != 0.0 ? b / a : b;   // a = (t,t), b = (s,t)
      break;
    default: {  // BCGS_RES2
      if (!S.active) return;
      if (S.half) {
        S.active = 0;
        break;
      }
      S.rho = S.rho_new;
      S.last = sqrt(a);
      record_defect(S, defects, maxdef, cj, S.last);
      ++S.nit;
      if (fabs(S.last) < fabs(S.abslimit) || fabs(S.last) < fabs(S.target)) {
        S.conv = 1;
        S.active = 0;
        break;
      }
      if (sc_eq1(S.last, S.old, 1e-4)) ++S.eq_count;
      else S.eq_count = 0;
      S.old = S.last;
      if (S.eq_count > 4) {
        S.conv = 0;
        S.active = 0;
      }
      break;
    }
  }
  st[cj] = S;
}

void launch_bcgs_scalar(hipStream_t st, int op, int nparts, const double *part, BcgsState *s,
                        int nrhs, int restart, double abslimit, double red, double *defects,
                        int maxdef) {
  k_bcgs_scalar<<<nrhs, 1024, 0, st>>>(op, nparts, part, s, restart, abslimit, red, defects,
                                       maxdef);
}

__global__ __launch_bounds__(kBlock) void k_bcgs_p(int n, const double *__restrict__ b,
                                                    double *__restrict__ r, double *__restrict__ v,
                                                    double *__restrict__ p,
                                                    double *__restrict__ w64,
                                                    float *__restrict__ w32,
                                                    const BcgsState *__restrict__ st, int restart,
                                                    size_t vs) {
  const int cj = blockIdx.y;
  if (!st[cj].active) return;
  const double beta = st[cj].beta, mbo = -beta * st[cj].omega;
  const size_t o = (size_t)cj * vs;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double bi = b[o + i];
    double pi, vi;
    if (restart) {
      r[o + i] = bi;
      v[o + i] = 0.0;
      pi = 0.0;
      vi = 0.0;
    } else {
      pi = p[o + i];
      vi = v[o + i];
    }
    double t = beta * pi;
    t = t + bi;
    t = t + mbo * vi;
    p[o + i] = t;
    if (w64) w64[o + i] = t;
    if (w32) w32[o + i] = (float)t;
  }
}

void launch_bcgs_p(hipStream_t st, int n, const double *b, double *r, double *v, double *p,
                   double *w64, float *w32, const BcgsState *s, int restart, Batch bt) {
  k_bcgs_p<<<grid2(bcgs_nparts(n), bt.nrhs), kBlock, 0, st>>>(n, b, r, v, p, w64, w32, s, restart,
                                                              bt.vs);
}

template <int MW, typename TQ, int NT>
__global__ __launch_bounds__(kBlock) void k_bcgs_spmv(DSell<double> A, int nsl,
                                                       const TQ *__restrict__ q,
                                                       double *__restrict__ y,
                                                       const double *__restrict__ z,
                                                       const uint8_t *__restrict__ vclass,
                                                       double *__restrict__ part, int yy,
                                                       int skip_half,
                                                       const BcgsState *__restrict__ st,
                                                       size_t vs) {
  __shared__ double red[2][kBlock / 64];
  SLICE_PROLOGUE_RED(A, nsl)
  if (!st[col_].active || (skip_half && st[col_].half)) return;   // uniform per block
  const size_t o = (size_t)col_ * vs;
  double yz = 0.0, y2 = 0.0;
  if (sok) {
    const double sum =
        sell_dot<MW, NT, double, TQ, double>(MV(A) + e0 + lane, A.col + e0 + lane, w, q + o, sdo, A.dtab, rb);
    if (r >= 0) {
      y[o + r] = sum;
      if (vclass[r] >= 2) {
        yz = sum * z[o + r];
        y2 = sum * sum;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    yz += __shfl_down(yz, off, 64);
    y2 += __shfl_down(y2, off, 64);
  }
  if (lane == 0) {
    red[0][threadIdx.x >> 6] = yz;
    red[1][threadIdx.x >> 6] = y2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b2 = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) {
      a += red[0][i];
      b2 += red[1][i];
    }
    // OMEGA wants (t,t) in k = 0 and (s,t) in k = 1; ALPHA wants (v,r) in k = 0
    double *pp = part + (size_t)(col_ * 2) * gridDim.x;
    if (yy) {
      pp[sb_] = b2;
      pp[gridDim.x + sb_] = a;
    } else {
      pp[sb_] = a;
    }
  }
}

template <typename TQ>
void launch_bcgs_spmv(hipStream_t st, const DSell<double> &A, const TQ *q, double *y,
                      const double *z, const uint8_t *vclass, double *part, int yy, int skip_half, const BcgsState *s, bool nt, Batch bt) {
  if (A.nslices <= 0) return;
  const dim3 g = grid2(nblk_slices(A.nslices), bt.nrhs);
  DSell<double> Ap = A;
  Ap.ncb = 0;   // its partial sums are strided by the grid: the plain map
  W_DISPATCH(A.wmax, NT_DISPATCH(nt, k_bcgs_spmv<MWV, TQ, NTV><<<g, kBlock, 0, st>>>(
                                         Ap, A.nslices, q, y, z, vclass, part, yy, skip_half, s,
                                         bt.vs)));
}
template void launch_bcgs_spmv<double>(hipStream_t, const DSell<double> &, const double *,
                                       double *, const double *, const uint8_t *, double *, int,
                                       int, const BcgsState *, bool, Batch);
template void launch_bcgs_spmv<float>(hipStream_t, const DSell<double> &, const float *, double *,
                                      const double *, const uint8_t *, double *, int, int,
                                      const BcgsState *, bool, Batch);

template <typename TQ>
__global__ __launch_bounds__(kBlock) void k_bcgs_s(int n, const double *__restrict__ b,
                                                    const double *__restrict__ v,
                                                    const TQ *__restrict__ q,
                                                    double *__restrict__ x,
                                                    double *__restrict__ sv,
                                                    double *__restrict__ w64,
                                                    float *__restrict__ w32,
                                                    const uint8_t *__restrict__ vclass,
                                                    double *__restrict__ part,
                                                    const BcgsState *__restrict__ st, size_t vs) {
  __shared__ double red[kBlock / 64];
  const int cj = blockIdx.y;
  if (!st[cj].active) return;   // uniform per block
  const double alpha = st[cj].alpha, malpha = -alpha;
  const size_t o = (size_t)cj * vs;
  double s2 = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[o + i] = x[o + i] + alpha * (double)q[o + i];
    const double si = b[o + i] + malpha * v[o + i];
    sv[o + i] = si;
    if (w64) w64[o + i] = si;
    if (w32) w32[o + i] = (float)si;
    if (vclass[i] >= 2) s2 += si * si;
  }
  const double t = block_sum(s2, red);
  if (threadIdx.x == 0) part[(size_t)(cj * 2) * gridDim.x + blockIdx.x] = t;
}

template <typename TQ>
void launch_bcgs_s(hipStream_t st, int n, const double *b, const double *v, const TQ *q,
                   double *x, double *sv, double *w64, float *w32, const uint8_t *vclass,
                   double *part, const BcgsState *s, Batch bt) {
  k_bcgs_s<TQ><<<grid2(bcgs_nparts(n), bt.nrhs), kBlock, 0, st>>>(n, b, v, q, x, sv, w64, w32,
                                                                  vclass, part, s, bt.vs);
}
template void launch_bcgs_s<double>(hipStream_t, int, const double *, const double *,
                                    const double *, double *, double *, double *, float *,
                                    const uint8_t *, double *, const BcgsState *, Batch);
template void launch_bcgs_s<float>(hipStream_t, int, const double *, const double *,
                                   const float *, double *, double *, double *, float *,
                                   const uint8_t *, double *, const BcgsState *, Batch);

template <typename TQ>
__global__ __launch_bounds__(kBlock) void k_bcgs_b(int n, const double *__restrict__ sv,
                                                    const double *__restrict__ t,
                                                    const TQ *__restrict__ q,
                                                    const double *__restrict__ r,
                                                    double *__restrict__ x,
                                                    double *__restrict__ b,
                                                    const uint8_t *__restrict__ vclass,
                                                    double *__restrict__ part,
                                                    const BcgsState *__restrict__ st, size_t vs) {
  __shared__ double red[2][kBlock / 64];
  const int cj = blockIdx.y;
  if (!st[cj].active) return;   // uniform per block
  const bool half = st[cj].half;
  const double omega = st[cj].omega, momega = -omega;
  const size_t o = (size_t)cj * vs;
  double b2 = 0.0, br = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double bi;
    if (half) {
      bi = sv[o + i];
    } else {
      x[o + i] = x[o + i] + omega * (double)q[o + i];
      bi = sv[o + i] + momega * t[o + i];
    }
    b[o + i] = bi;
    if (vclass[i] >= 2) {
      b2 += bi * bi;
      br += bi * r[o + i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    b2 += __shfl_down(b2, off, 64);
    br += __shfl_down(br, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = b2;
    red[1][threadIdx.x >> 6] = br;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) {
      a += red[0][i];
      c += red[1][i];
    }
    double *pp = part + (size_t)(cj * 2) * gridDim.x;
    pp[blockIdx.x] = a;
    pp[gridDim.x + blockIdx.x] = c;
  }
}

template <typename TQ>
void launch_bcgs_b(hipStream_t st, int n, const double *sv, const double *t, const TQ *q,
                   const double *r, double *x, double *b, const uint8_t *vclass, double *part,
                   const BcgsState *s, Batch bt) {
  k_bcgs_b<TQ><<<grid2(bcgs_nparts(n), bt.nrhs), kBlock, 0, st>>>(n, sv, t, q, r, x, b, vclass,
                                                                  part, s, bt.vs);
}
template void launch_bcgs_b<double>(hipStream_t, int, const double *, const double *,
                                    const double *, const double *, double *, double *,
                                    const uint8_t *, double *, const BcgsState *, Batch);
template void launch_bcgs_b<float>(hipStream_t, int, const double *, const double *,
                                   const float *, const double *, double *, double *,
                                   const uint8_t *, double *, const BcgsState *, Batch);

__global__ __launch_bounds__(kBlock) void k_bcgs_norm(int n, const double *__restrict__ x,
                                                       const uint8_t *__restrict__ vclass,
                                                       double *__restrict__ part, size_t vs) {
  __shared__ double red[kBlock / 64];
  const int cj = blockIdx.y;
  const size_t o = (size_t)cj * vs;
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    if (vclass[i] >= 2) s += x[o + i] * x[o + i];
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) {
    double *pp = part + (size_t)(cj * 2) * gridDim.x;
    pp[blockIdx.x] = t;
    pp[gridDim.x + blockIdx.x] = t;
  }
}

void launch_bcgs_norm(hipStream_t st, int n, const double *x, const uint8_t *vclass, double *part,
                      Batch bt) {
  k_bcgs_norm<<<grid2(bcgs_nparts(n), bt.nrhs), kBlock, 0, st>>>(n, x, vclass, part, bt.vs);
}

}  // namespace ugmg
