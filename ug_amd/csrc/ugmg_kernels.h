// Launchers and device-side descriptors of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

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
  int wmax;           // widest slice (entries per row)
  size_t pst;         // entries per value plane (b x b blocks: b*b planes)
  // members (one matrix per column of a batched launch, same pattern):
  // column j reads its values at val + j * mst and the level's diagonal
  // blocks at diag + j * dms; 0 = one matrix for every column
  size_t mst = 0;
  size_t dms = 0;
  // colour-grouped slab of a scalar GS level (A, U): launches over the whole
  // slab with one column give XCD x the x-th eighth of every colour's blocks
  // (blocks of kBlock / 64 slices; cb[k] = first block of colour k, cb[ncb]
  // = all blocks); ncb = 0: the plain XCD remap
  int ncb = 0;
  int cb[kMaxColors + 1] = {};
  // value planes per entry: 1 (scalar weights), or b*b (a b x b block per
  // entry: matrices, and the transfer slabs of an interpolation matrix)
  int bb = 1;
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

// Batched launches: every vector argument holds `nrhs` columns, column j at
// offset j * vs (vs = the level's column stride, >= n); the column index of a
// workgroup is its second grid coordinate.  nrhs = 1 is the plain case.
// b = components per vector (1, or 3 with 3x3 blocks): vector v's
// components are consecutive (UG's layout), column strides count scalars.
struct Batch {
  int nrhs;
  size_t vs;
  int b = 1;
  // members: column j's diagonal blocks at diag + j * dms and its base
  // inverse at inv + j * ims (0: shared by every column)
  size_t dms = 0;
  size_t ims = 0;
  // vector types (VTYPE per row, nt of them): the norms and dot products
  // have nt * b entries per column, type-major; null: one type
  const uint8_t *vt = nullptr;
  int nt = 1;
  int nc() const { return vt ? nt * b : b; }
  // the stop tests' VD_IDENT grouping of the norm entries (4 bits per
  // entry, entry q at bits 4q..4q+3); -1: NO_IDENT
  int idents = -1;
  // ex $damp per component for the base solve (unknown r takes
  // pdamp.d[r % pdamp.vb]); pdamp.vb = 0: the one damping factor
  struct PDamp {
    double d[4] = {};
    int vb = 0;
  } pdamp;
};

// t[v b + q] *= d[q] on n vectors of b components, then (c non-null) c =
// t (cmode 0) or c += t (cmode 1): dscalx with a damping factor per
// component (np/algebra/ugblas.cc:2510-2528) after a step or an
// interpolation formed with damping 1; cls: only the rows whose class is
// at least clsmin (the coarse rows a restriction wrote)
template <typename T>
void launch_scale_comp(hipStream_t st, int n, int b, const double *d, T *t, T *c, int cmode, Batch bt,
                       const uint8_t *cls = nullptr, int clsmin = 0);

// per right-hand side state of the batched BCGSSolver (np/procs/ls.cc:
// 1864-2062); written only by the scalar kernels, read by the vector kernels
struct BcgsState {
  double rho, omega, alpha, beta, rho_new;
  double first, target, last, old;
  double abslimit;
  int eq_count;   // consecutive sc_eq(last, old, 1e-4) hits
  int active;     // still iterating
  int half;       // converged after the first half step of this iteration
  int conv;
  int nit;        // LinearResiduum calls after the first defect
  int nd;         // defects recorded
};

enum BcgsOp { BCGS_FIRST = 0, BCGS_RHO = 1, BCGS_ALPHA = 2, BCGS_RES1 = 3, BCGS_OMEGA = 4, BCGS_RES2 = 5 };

template <typename TV, typename TX, typename TY>
void launch_spmv(hipStream_t st, const DSell<TV> &A, int slice0, int nsl, const TX *x, TY *y,
                 TY *c, int mode, bool nt, Batch bt);
void launch_ir_update(hipStream_t st, const DSell<double> &A, const float *c32, double *r,
                      double *x, float *r32, const uint8_t *vclass, double *part, int *nparts,
                      bool nt, Batch bt);
template <typename T>
void launch_jac(hipStream_t st, int n, const T *b, const T *diag, const uint8_t *vclass, T damp,
                T *t, Batch bt);
template <typename TV, typename T>
void launch_gs_color(hipStream_t st, const DSell<TV> &L, int slice0, int nsl, int wmax, const T *b,
                     const T *diag, const uint8_t *vclass, T *v, bool nt, Batch bt);
// sweep of one colour with the previous step's defect folded in (all rows
// active); cmode is the correction mode of the previous step; split (scalar
// rows): the upper and lower row products on two wavefronts per slice
template <typename TV, typename T>
void launch_gs_color_fused(hipStream_t st, const DSell<TV> &L, const DSell<TV> &U, int slice0,
                           int nsl, int wmax_l, int wmax_u, T *b, const T *diag, T damp, T *v,
                           T *c, int cmode, bool nt, Batch bt, bool split);
// y -= A x on a GS level's colour-grouped rows, plus u = y / diag on the rows
// of colour 0 (the next sweep's first colour; scalar, every row active)
template <typename TV, typename T>
void launch_spmv_c0(hipStream_t st, const DSell<TV> &A, const T *x, T *y, const T *diag, T *u,
                    int ncs0, bool nt, Batch bt);
template <typename TV, typename T>
void launch_gs_defect(hipStream_t st, const DSell<TV> &U, const T *v, T *b, const uint8_t *vclass,
                      T damp, T *c, int cmode, T *xout, bool nt, Batch bt);
// returns the number of blocks per column (= partial sums per column when
// part != null; column j's partials start at part + j * return value).
// With ct (scalar only) also ct_r = bc_r / cdiag_r on the coarse rows whose
// cc0 flag is set: the coarse GS pre-smoother's colour-0 sweep.  With cpos
// (scalar only) also cbc[cpos_r] = bc_r where cpos_r >= 0: b in the base
// inverse's column order.
template <typename T>
int launch_restrict(hipStream_t st, const DSell<T> &PT, const uint32_t *cskip,
                    const uint8_t *cvnclass, const uint8_t *cvclass, const T *bf, T damp, T *bc,
                    double *part, bool nt, int nrhs, size_t vsf, size_t vsc, int b,
                    const uint8_t *cc0 = nullptr, const T *cdiag = nullptr, T *ct = nullptr,
                    const int *cpos = nullptr, T *cbc = nullptr, size_t cdms = 0);
// Element lists of one level on the device (all pointers device memory):
// nelem elements of nc corners, elem[e*nc + i] = vector of corner i; element
// matrices (nc b) x (nc b) row-major over corner-major scalar unknowns,
// emat (per element) or scale[e] * kref (emat null); evec (nc b per
// element) or null; skip = VECSKIP per vector or null; dval = Dirichlet
// values (n b) or null (0).
struct AsmInput {
  int n, b, nelem, nc;
  const int *elem;
  const double *emat, *kref, *scale, *evec;
  const uint32_t *skip;
  const double *dval;
};
// Assembles the level on the device and copies the CSR (diagonal block
// first, then ascending columns) to the host vectors; rhs (device, n b) or
// null.  Returns 0 or a negative UGMG_E_* code with err set.
// keep (optional): the flattened level's device arrays, handed to the
// caller (hipFree) instead of freed
struct AsmKeep {
  int *rp = nullptr, *ci = nullptr;
  double *vals = nullptr;
};
int assemble_device(hipStream_t st, const AsmInput &in, double *rhs, std::vector<int> &rp_h,
                    hvec<int> &ci_h, hvec<double> &vals_h, std::string &err,
                    AsmKeep *keep = nullptr);

// Band LU (no pivoting) of the base level and its explicit inverse, on the
// device: band (n rows, bandwidth bw, UG's EX_MAT layout) is factored in
// place, X (n rounded up to 64, times n doubles) receives the inverse in
// band positions as panels of 64 columns, inv (n x n) the inverse in list
// order: inv[r][u] = entry (pos[r], pos[u]).  *failed (0 on entry) is set
// on a zero pivot.
void launch_base_factor(hipStream_t st, double *band, int n, int bw, const int *pos, double *X,
                        double *inv, int *failed, bool fmode = false);
// inv: the rows of the dense inverse that are not unit rows (rowmap[r]: the
// row's index there, -1 for a unit row), columns in the order cp, in FP64
// (TI = double) or the FP32 copy for the FP32 cycle (TI = float); the
// products run over the first ncols columns and are summed in FP64.  bc
// (scalar rows; may be null): b already in column order, bc_j = b[cp[j]]
template <typename T, typename TI>
void launch_base_gemv(hipStream_t st, int n, const TI *inv, const int *rowmap, const int *cp,
                      int ncols, const T *b, const T *bc, T *c, T *xout, int cmode, double damp,
                      double abslimit, double red, const double *part, int nparts, double *first,
                      int it, Batch bt);
// ls over jac on the base level: after the Jacobi step's x (n scalar
// unknowns), the base stop test on the defect's partial norms, x = 0 once
// it holds, c (+)= x (cmode 0 / 1); first: the first defect's norms
template <typename T>
void launch_base_gate(hipStream_t st, int n, T *x, T *c, int cmode, double abslimit, double red,
                      const double *part, int nparts, int it, double *first, Batch bt);
// members: d64[j] = src[map[j]] (0 where map[j] < 0), d32[j] = (float) d64[j]
void launch_member_gather(hipStream_t st, size_t n, const int *map, const double *src, double *d64,
                          float *d32);
// rows act[0..nact), columns in the order cp, of the n x n inverse into
// compact FP64 / FP32 copies
void launch_base_compact(hipStream_t st, int nact, int n, const int *act, const int *cp,
                         const double *inv, double *c64, float *c32);
// columns of the partial sums: one per component of every right-hand side
// (column j * b + i), nparts each
template <typename T>
int launch_dot_part(hipStream_t st, int n, const T *x, const uint8_t *vclass, double *part,
                    Batch bt);
template <typename T>
void launch_interp(hipStream_t st, const DSell<T> &P, const uint32_t *fskip, const T *cc, T damp,
                   T *t, T *c, bool nt, int nrhs, size_t vsf, size_t vsc, int b);
int reduce_parts(int n);
template <typename T>
void launch_dot(hipStream_t st, int n, const T *x, const T *y, const uint8_t *vclass,
                double *part, double *out, int do_sqrt, Batch bt);
// out[j] = sum of column j's nparts partials (sqrt if do_sqrt)
void launch_reduce_final(hipStream_t st, int nparts, const double *part, double *out, int do_sqrt,
                         int nrhs);
template <typename T>
void launch_fill(hipStream_t st, int n, T *x, T val, Batch bt);
// the VECSKIP bits the interpolation reads: an edge row's (vt 1) own bits
// are not read (StandardIntCorNodeVector's edge loop), iskip 0
void launch_iskip(hipStream_t st, int n, const uint8_t *vt, const uint32_t *skip, uint32_t *iskip);
// ClearDirichletValues on a list-order vector of n vectors of b components:
// x[v b + j] = 0 where bit j of VECSKIP is set (skip in device order, perm:
// device position -> list index, null: list order)
void launch_clear_dirichlet(hipStream_t st, int n, int b, const int *perm, const uint32_t *skip, double *x);
template <typename TD, typename TS>
void launch_convert(hipStream_t st, int n, const TS *s, TD *d, int nrhs, size_t vss, size_t vsd);
template <typename T>
void launch_add(hipStream_t st, int n, T *x, const T *y, Batch bt);
// out[i] = 3 for FINE_GRID_DOF vectors (VCLASS >= 2, VNCLASS <= 1), else 0
void launch_fgd_class(hipStream_t st, int n, const uint8_t *vc, const uint8_t *vnc, uint8_t *out);
// dst = src on the n vectors (b components) whose class byte is >= 2
void launch_copy_rows(hipStream_t st, int n, int b, const uint8_t *cls, const double *src,
                      double *dst);
// list order <-> device order of a level: gather d[p] = s[perm[p]],
// scatter d[perm[p]] = s[p]; column strides vss (source), vsd (destination)
template <typename TD, typename TS>
void launch_gather(hipStream_t st, int n, const int *perm, const TS *s, TD *d, int nrhs,
                   size_t vss, size_t vsd, int b);
template <typename TD, typename TS>
void launch_scatter(hipStream_t st, int n, const int *perm, const TS *s, TD *d, int nrhs,
                    size_t vss, size_t vsd, int b);
template <typename T>
void launch_coarse_cycle(hipStream_t st, const CoarseArgs<T> &args, int base_n);
// coarse values ca (pattern crp/cci, bb values per entry) = P^T A P; pt*:
// every fine vector that interpolates from coarse row I, in list order
// *bad = min(*bad, first vector of class >= 3 whose diagonal block
// SolveSmallBlock refuses); b = 2 and 4 (UG checks no other size)
void launch_block_check(hipStream_t st, int n, int b, const double *diag, const uint8_t *vclass,
                        int *bad);
void launch_galerkin(hipStream_t st, int nc, int bb, const int *ptp, const int *pti,
                     const double *ptw, const int *frp, const int *fci, const double *fa,
                     const int *pp, const int *pi, const double *pw, const int *crp,
                     const int *cci, double *ca);

// ---- batched BCGSSolver pieces (FP64 outer vectors; TQ = type of q, the
// preconditioned direction: double, or float when the cycle runs in FP32)
// partial sums per column: two per block, layout part[(j * 2 + k) * nparts + blk]
int bcgs_nparts(int n);
void launch_bcgs_scalar(hipStream_t st, int op, int nparts, const double *part, BcgsState *s,
                        int nrhs, int restart, double abslimit, double red, double *defects,
                        int maxdef);
// restart: r = b, v = 0, p = 0 first.  p = beta p + b - beta omega v;
// then the cycle's input: w64 = p or w32 = (float) p (whichever is non-null)
void launch_bcgs_p(hipStream_t st, int n, const double *b, double *r, double *v, double *p,
                   double *w64, float *w32, const BcgsState *s, int restart, Batch bt);
// y = A q on active columns (skip_half: and not converged at the half step),
// with the partial sums of y.z (k = 0) and y.y (k = 1, when yy) over NEW_DEFECT rows
template <typename TQ>
void launch_bcgs_spmv(hipStream_t st, const DSell<double> &A, const TQ *q, double *y,
                      const double *z, const uint8_t *vclass, double *part, int yy, int skip_half, const BcgsState *s, bool nt, Batch bt);
int bcgs_spmv_nparts(int nslices);
// x += alpha q; s = b - alpha v; cycle input w = s; partials of s.s (k = 0)
template <typename TQ>
void launch_bcgs_s(hipStream_t st, int n, const double *b, const double *v, const TQ *q,
                   double *x, double *sv, double *w64, float *w32, const uint8_t *vclass,
                   double *part, const BcgsState *s, Batch bt);
// half-converged: b = s; else x += omega q; b = s - omega t.  partials of
// b.b (k = 0) and b.r (k = 1)
template <typename TQ>
void launch_bcgs_b(hipStream_t st, int n, const double *sv, const double *t, const TQ *q,
                   const double *r, double *x, double *b, const uint8_t *vclass, double *part,
                   const BcgsState *s, Batch bt);
// partials of x.x (k = 0, and k = 1) for the first defect
void launch_bcgs_norm(hipStream_t st, int n, const double *x, const uint8_t *vclass, double *part,
                      Batch bt);

}  // namespace ugmg
