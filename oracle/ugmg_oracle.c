/*
 * ugmg_oracle.c -- CPU restatement of UG's linear multigrid solve path.
 *
 * TEST INFRASTRUCTURE ONLY (see ugmg_oracle.h): the checker for the GPU
 * library and the CPU baseline of bench.py.  Nothing in the product links it.
 *
 * Restated from the reference (file:line relative to the UG tree):
 *   dmatmul_minus / dmatmul   np/algebra/ugblas.cc:3997-4042, 3782-3890,
 *                             matloop.ct:75-99, MATMUL_* ugblas.h:161-213
 *   dnrm2x / ddot             np/algebra/ugblas.cc:3156-3176, 3104-3122
 *   SolveSmallBlock           np/algebra/block.cc:104-205
 *   l_jac                     np/algebra/ugiter.cc:271-335
 *   l_setindex + l_lgs        np/algebra/ugiter.cc:169-186, 412-615
 *   Smoother                  np/procs/iter.cc:817-842
 *   StandardRestrict...       np/algebra/transgrid.cc:117-215
 *   StandardIntCor...         np/algebra/transgrid.cc:235-336
 *   EXPreProcess/EXSmoother   np/procs/iter.cc:8685-9071
 *   band LU                   gm/evm.cc:1400-1418, 1493-1510, evm.h:392
 *   Lmgc                      np/procs/iter.cc:7741-7949
 *   LinearSolver/LSUpdate     np/procs/ls.cc:637-749, 869-875
 *   LinearDefect/Residuum     np/procs/ls.cc:562-593
 *   sc_cmp / sc_mul_check     np/udm/npscan.cc:1027-1057, 1151-1161
 * orc_ls_lmgc_solve takes a uniformly refined hierarchy: FULLREFINELEVEL is
 * the top level, so ON_SURFACE operations touch the top level's NEW_DEFECT
 * (VCLASS >= 2) vectors only (vecloop.ct:22-39).  orc_ls_lmgc_solve_surface
 * takes a locally refined one: the surface then reaches down to
 * FULLREFINELEVEL (FINE_GRID_DOF vectors of the lower levels).
 */
#include "ugmg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ACTIVE_CLASS 3
#define NEWDEF_CLASS 2
#define EDGE_TYPE 1

/* VTYPE of vector v and the first VEC_SCALAR entry of its type */
static int vtype_of(const orc_level *L, int v) { return L->vtype ? L->vtype[v] : 0; }
static int aoff(const orc_level *L, int v) { return vtype_of(L, v) * L->b; }
/* VD_NCOMP: entries of a VEC_SCALAR */
static int ncomp_of(const orc_level *L) { return (L->ntypes > 0 ? L->ntypes : 1) * L->b; }

/* ---------------------------------------------------------------- ugblas */

void orc_dmatmul_minus(const orc_level *L, double *x, const double *y) {
  const int b = L->b, bb = b * b;
  for (int v = 0; v < L->n; ++v) {
    double s[ORC_MAXB] = {0.0};
    for (int m = L->rp[v]; m < L->rp[v + 1]; ++m) {
      const double *M = L->a + (size_t)m * bb;
      const double *w = y + (size_t)L->ci[m] * b;
      for (int i = 0; i < b; ++i) {
        double t = 0.0;
        for (int j = 0; j < b; ++j) t += M[i * b + j] * w[j];
        s[i] += t;
      }
    }
    for (int i = 0; i < b; ++i) x[(size_t)v * b + i] -= s[i];
  }
}

void orc_dmatmul(const orc_level *L, double *x, const double *y) {
  const int b = L->b;
  for (size_t i = 0; i < (size_t)L->n * b; ++i) x[i] = 0.0;
  /* x = 0; x += M y  (matmode.ct:76-78) */
  const int bb = b * b;
  for (int v = 0; v < L->n; ++v) {
    double s[ORC_MAXB] = {0.0};
    for (int m = L->rp[v]; m < L->rp[v + 1]; ++m) {
      const double *M = L->a + (size_t)m * bb;
      const double *w = y + (size_t)L->ci[m] * b;
      for (int i = 0; i < b; ++i) {
        double t = 0.0;
        for (int j = 0; j < b; ++j) t += M[i * b + j] * w[j];
        s[i] += t;
      }
    }
    for (int i = 0; i < b; ++i) x[(size_t)v * b + i] += s[i];
  }
}

/* dnrm2x (ugblas.cc:3156-3176): a[aoff[VTYPE(v)] + i] += x_i^2, each entry
 * summed in list order (vecmode.ct visits one type at a time) */
void orc_dnrm2x(const orc_level *L, const double *x, double *out) {
  const int b = L->b, nc = ncomp_of(L);
  double a[ORC_MAXB] = {0.0};
  for (int v = 0; v < L->n; ++v)
    for (int i = 0; i < b; ++i) {
      double s = x[(size_t)v * b + i];
      a[aoff(L, v) + i] += s * s;
    }
  for (int i = 0; i < nc; ++i) out[i] = sqrt(a[i]);
}

void orc_ddot(const orc_level *L, const double *x, const double *y,
              double *out) {
  const int b = L->b, nc = ncomp_of(L);
  double a[ORC_MAXB] = {0.0};
  for (int v = 0; v < L->n; ++v)
    for (int i = 0; i < b; ++i)
      a[aoff(L, v) + i] += x[(size_t)v * b + i] * y[(size_t)v * b + i];
  for (int i = 0; i < nc; ++i) out[i] = a[i];
}

/* surface norm: top level, NEW_DEFECT vectors */
static void surface_nrm2x(const orc_level *L, const double *x, double *out) {
  const int b = L->b, nc = ncomp_of(L);
  double a[ORC_MAXB] = {0.0};
  for (int v = 0; v < L->n; ++v) {
    if (L->vclass[v] < NEWDEF_CLASS) continue;
    for (int i = 0; i < b; ++i) {
      double s = x[(size_t)v * b + i];
      a[aoff(L, v) + i] += s * s;
    }
  }
  for (int i = 0; i < nc; ++i) out[i] = sqrt(a[i]);
}

static void surface_dmatmul_minus(const orc_level *L, double *x,
                                  const double *y) {
  const int b = L->b, bb = b * b;
  for (int v = 0; v < L->n; ++v) {
    if (L->vclass[v] < NEWDEF_CLASS) continue;
    double s[ORC_MAXB] = {0.0};
    for (int m = L->rp[v]; m < L->rp[v + 1]; ++m) {
      const double *M = L->a + (size_t)m * bb;
      const double *w = y + (size_t)L->ci[m] * b;
      for (int i = 0; i < b; ++i) {
        double t = 0.0;
        for (int j = 0; j < b; ++j) t += M[i * b + j] * w[j];
        s[i] += t;
      }
    }
    for (int i = 0; i < b; ++i) x[(size_t)v * b + i] -= s[i];
  }
}

/* dscalx (ugblas.cc:2510-2528): x_i *= a[aoff[VTYPE(v)] + i] */
static void dscalx(const orc_level *L, double *x, const double *a) {
  const int b = L->b;
  for (int v = 0; v < L->n; ++v)
    for (int i = 0; i < b; ++i) x[(size_t)v * b + i] *= a[aoff(L, v) + i];
}

static void dadd(size_t len, double *x, const double *y) {
  for (size_t i = 0; i < len; ++i) x[i] += y[i];
}

/* sc_cmp (np/udm/npscan.cc:1027-1057): 1 iff |x_i| < |y_i| for every
 * component; with VD_IDENT (ident != NULL) iff, for every group leader i
 * (ident[i] == i), the sum of x_j^2 over the group (ident[j] == i) is below
 * that of y_j^2 */
static int sc_cmp(int b, const double *x, const double *y, const int *ident) {
  if (ident) {
    for (int i = 0; i < b; ++i) {
      if (ident[i] != i) continue;
      double sx = 0.0, sy = 0.0;
      for (int j = 0; j < b; ++j)
        if (ident[j] == i) {
          sx += x[j] * x[j];
          sy += y[j] * y[j];
        }
      if (sx >= sy) return 0;
    }
    return 1;
  }
  for (int i = 0; i < b; ++i)
    if (fabs(x[i]) >= fabs(y[i])) return 0;
  return 1;
}
#define SC_IDENT(p) ((p)->has_ident ? (p)->ident : NULL)

/* ---------------------------------------------------------- small blocks */

/* MAX_SINGLE_VEC_COMP (np/udm/udm.h:62): UG refuses blocks of 40 and more */
#define ORC_MAX_SINGLE_VEC_COMP 40
/* SMALL_D = DBL_EPSILON * SMALL_FAC (low/architecture.h:27,34) */
#define ORC_SMALL_D (2.2204460492503131e-16 * 10)

int orc_solve_small_block(int n, const double *mat, const double *rhs_in,
                          double *sol) {
  double rhs[ORC_MAX_SINGLE_VEC_COMP], M[ORC_MAX_SINGLE_VEC_COMP * ORC_MAX_SINGLE_VEC_COMP],
      X[ORC_MAX_SINGLE_VEC_COMP];
  if (n >= ORC_MAX_SINGLE_VEC_COMP) return 1;
  for (int i = 0; i < n; ++i) rhs[i] = rhs_in[i];
  switch (n) {
    case 1:
      sol[0] = rhs[0] / mat[0];
      return 0;
    case 2: {
      double det = mat[0] * mat[3] - mat[1] * mat[2];
      if (det == 0.0) return 1;
      det = 1.0 / det;
      sol[0] = (rhs[0] * mat[3] - rhs[1] * mat[1]) * det;
      sol[1] = (rhs[1] * mat[0] - rhs[0] * mat[2]) * det;
      return 0;
    }
    case 3: {
      /* explicit elimination without pivoting (block.cc:131-142) */
      double M3div0 = mat[3] / mat[0];
      double M6div0 = mat[6] / mat[0];
      double aux = (mat[7] - M6div0 * mat[1]) / (mat[4] - M3div0 * mat[1]);
      sol[2] = (rhs[2] - M6div0 * rhs[0] - aux * (rhs[1] - M3div0 * rhs[0])) /
               (mat[8] - M6div0 * mat[2] - aux * (mat[5] - M3div0 * mat[2]));
      sol[1] = (rhs[1] - mat[3] / mat[0] * rhs[0] -
                (mat[5] - M3div0 * mat[2]) * sol[2]) /
               (mat[4] - M3div0 * mat[1]);
      sol[0] = (rhs[0] - mat[1] * sol[1] - mat[2] * sol[2]) / mat[0];
      return 0;
    }
    default: {
      for (int i = 0; i < n * n; ++i) M[i] = mat[i];
      for (int i = 0; i < n - 1; ++i) {
        double piv = M[i * n + i];
        if (fabs(piv) < 1e-10) {
          int pivrow = i;
          for (int j = i + 1; j < n; ++j)
            if (fabs(M[j * n + i]) > fabs(piv)) { piv = M[j * n + i]; pivrow = j; }
          if (fabs(piv) < ORC_SMALL_D * 1e-10) return 1;
          for (int k = i; k < n; ++k) {
            double f = M[i * n + k]; M[i * n + k] = M[pivrow * n + k]; M[pivrow * n + k] = f;
          }
          double f = rhs[i]; rhs[i] = rhs[pivrow]; rhs[pivrow] = f;
        }
        for (int j = i + 1; j < n; ++j) {
          double f = M[j * n + i] / piv;
          for (int k = i + 1; k < n; ++k) M[j * n + k] -= M[i * n + k] * f;
          rhs[j] -= rhs[i] * f;
        }
      }
      for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) rhs[i] -= M[i * n + j] * X[j];
        X[i] = rhs[i] / M[i * n + i];
      }
      for (int i = 0; i < n; ++i) sol[i] = X[i];
      return 0;
    }
  }
}

/* ------------------------------------------------------------- smoothers */

int orc_l_jac(const orc_level *L, double *v, const double *d) {
  const int b = L->b, bb = b * b;
  for (int i = 0; i < L->n; ++i) {
    double *vi = v + (size_t)i * b;
    if (L->vclass[i] < ACTIVE_CLASS) {
      for (int k = 0; k < b; ++k) vi[k] = 0.0;
      continue;
    }
    const double *D = L->a + (size_t)L->rp[i] * bb; /* VSTART = diagonal */
    if (b == 1) {
      vi[0] = d[i] / D[0];
    } else if (orc_solve_small_block(b, D, d + (size_t)i * b, vi)) {
      return 1;
    }
  }
  return 0;
}

static const int *g_rank_for_sort;
static int cmp_rank(const void *pa, const void *pb) {
  int a = *(const int *)pa, b = *(const int *)pb;
  int ra = g_rank_for_sort[a], rb = g_rank_for_sort[b];
  if (ra != rb) return ra < rb ? -1 : 1;
  return a < b ? -1 : (a > b);
}

int orc_l_lgs(const orc_level *L, const int *rank, double *v, const double *d) {
  const int b = L->b, bb = b * b;
  /* vectors are visited in increasing VINDEX: list order for UG's
   * l_setindex numbering, increasing rank otherwise */
  int *order = (int *)malloc(sizeof(int) * (size_t)(L->n > 0 ? L->n : 1));
  for (int i = 0; i < L->n; ++i) order[i] = i;
  if (rank) {
    g_rank_for_sort = rank;
    qsort(order, (size_t)L->n, sizeof(int), cmp_rank);
  }
  for (int oi = 0; oi < L->n; ++oi) {
    const int i = order[oi];
    double *vi = v + (size_t)i * b;
    if (L->vclass[i] < ACTIVE_CLASS) {
      for (int k = 0; k < b; ++k) vi[k] = 0.0;
      continue;
    }
    const long myindex = rank ? rank[i] : i;
    if (b == 1) {
      double sum = 0.0;
      for (int m = L->rp[i] + 1; m < L->rp[i + 1]; ++m) {
        int w = L->ci[m];
        long wi = rank ? rank[w] : w;
        if (wi < myindex && L->vclass[w] >= ACTIVE_CLASS) sum += L->a[m] * v[w];
      }
      vi[0] = (d[i] - sum) / L->a[L->rp[i]];
    } else {
      double s[ORC_MAXB];
      for (int k = 0; k < b; ++k) s[k] = d[(size_t)i * b + k];
      double acc[ORC_MAXB] = {0.0};
      for (int m = L->rp[i] + 1; m < L->rp[i + 1]; ++m) {
        int w = L->ci[m];
        long wi = rank ? rank[w] : w;
        if (!(L->vclass[w] >= ACTIVE_CLASS && myindex > wi)) continue;
        const double *M = L->a + (size_t)m * bb;
        for (int r = 0; r < b; ++r) {
          double t = 0.0;
          for (int c = 0; c < b; ++c) t += M[r * b + c] * v[(size_t)w * b + c];
          acc[r] += t;
        }
      }
      for (int k = 0; k < b; ++k) s[k] -= acc[k];
      if (orc_solve_small_block(b, L->a + (size_t)L->rp[i] * bb, s, vi)) {
        free(order);
        return 1;
      }
    }
  }
  free(order);
  return 0;
}

int orc_greedy_color(const orc_level *L, int *color) {
  int ncol = 0;
  int cap = 64;
  int *mark = (int *)calloc((size_t)cap, sizeof(int));
  for (int i = 0; i < L->n; ++i) color[i] = -1;
  for (int i = 0; i < L->n; ++i) {
    int deg = L->rp[i + 1] - L->rp[i];
    if (deg + 1 > cap) {
      cap = deg + 1;
      mark = (int *)realloc(mark, (size_t)cap * sizeof(int));
    }
    for (int k = 0; k < cap; ++k) mark[k] = 0;
    for (int m = L->rp[i]; m < L->rp[i + 1]; ++m) {
      int w = L->ci[m];
      if (w == i) continue;
      int c = color[w];
      if (c >= 0 && c < cap) mark[c] = 1;
    }
    int c = 0;
    while (c < cap && mark[c]) ++c;
    color[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(mark);
  return ncol;
}

int orc_smoother(const orc_level *L, int kind, const int *rank,
                 const double *damp, double *x, double *b) {
  int err = (kind == ORC_JAC) ? orc_l_jac(L, x, b) : orc_l_lgs(L, rank, x, b);
  if (err) return err;
  dscalx(L, x, damp);
  orc_dmatmul_minus(L, b, x);
  return 0;
}

/* -------------------------------------------------------------- transfer */

/* CheckDamp (transgrid.cc): 1 iff some entry differs from 1 */
static int check_damp(int n, const double *damp) {
  for (int i = 0; i < n; ++i)
    if (damp[i] != 1.0) return 1;
  return 0;
}

/* RestrictByMatrix_General with flag 0 (transgrid.cc:1160-1232; no CRITBIT
 * is set here): reset the coarse rows of VNCLASS >= 2, add M b_f over the
 * fine vectors of VCLASS >= 2 on the coarse components whose VECSKIP bit is
 * clear, then scale the coarse rows of VNCLASS >= 2 by damp */
static void restrict_by_matrix(const orc_level *F, const orc_level *C, double *to_c,
                               const double *from_f, const double *damp) {
  const int b = F->b, bb = b * b;
  for (int w = 0; w < C->n; ++w)
    if (C->vnclass[w] >= NEWDEF_CLASS)
      for (int i = 0; i < b; ++i) to_c[(size_t)w * b + i] = 0.0;
  for (int v = 0; v < F->n; ++v) {
    if (F->vclass[v] < NEWDEF_CLASS) continue;
    const double *vp = from_f + (size_t)v * b;
    for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
      const int w = F->pi[q];
      const double *M = F->pblk + (size_t)q * bb;
      const uint32_t vecskip = C->skip[w];
      for (int i = 0; i < b; ++i) {
        if (vecskip & (1u << i)) continue;
        double sum = 0.0;
        for (int j = 0; j < b; ++j) sum += M[i * b + j] * vp[j];
        to_c[(size_t)w * b + i] += sum;
      }
    }
  }
  if (check_damp(ncomp_of(C), damp))
    for (int w = 0; w < C->n; ++w)
      if (C->vnclass[w] >= NEWDEF_CLASS)
        for (int i = 0; i < b; ++i) to_c[(size_t)w * b + i] *= damp[aoff(C, w) + i];
}

/* InterpolateCorrectionByMatrix_General with flag 0 (transgrid.cc:1300-
 * 1380): dset(to, 0); t_v,i += sum_j M_ji c_j on the fine components whose
 * VECSKIP bit is clear; dscalx(damp) when some damp differs from 1 */
static void interpolate_by_matrix(const orc_level *F, double *to_f, const double *from_c,
                                  const double *damp) {
  const int b = F->b, bb = b * b;
  for (size_t i = 0; i < (size_t)F->n * b; ++i) to_f[i] = 0.0;
  for (int v = 0; v < F->n; ++v) {
    const uint32_t vecskip = F->skip[v];
    double *vp = to_f + (size_t)v * b;
    for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
      const double *M = F->pblk + (size_t)q * bb;
      const double *wp = from_c + (size_t)F->pi[q] * b;
      for (int i = 0; i < b; ++i) {
        if (vecskip & (1u << i)) continue;
        double sum = 0.0;
        for (int j = 0; j < b; ++j) sum += M[j * b + i] * wp[j];
        vp[i] += sum;
      }
    }
  }
  if (check_damp(ncomp_of(F), damp)) dscalx(F, to_f, damp);
}

void orc_restrict(const orc_level *F, const orc_level *C, double *to_c,
                  const double *from_f, const double *damp) {
  if (F->pblk) {
    restrict_by_matrix(F, C, to_c, from_f, damp);
    return;
  }
  const int b = F->b;
  for (int v = 0; v < C->n; ++v)
    if (C->vnclass[v] >= NEWDEF_CLASS)
      for (int i = 0; i < b; ++i) to_c[(size_t)v * b + i] = 0.0;
  for (int v = 0; v < F->n; ++v) {
    if (vtype_of(F, v) != 0) continue;
    if (F->vclass[v] < NEWDEF_CLASS) continue;
    double s[ORC_MAXB];
    for (int i = 0; i < b; ++i) s[i] = damp[i] * from_f[(size_t)v * b + i];
    for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
      int vc = F->pi[q];
      double c = F->pw[q];
      uint32_t vecskip = C->skip[vc];
      for (int j = 0; j < b; ++j)
        if (!(vecskip & (1u << j))) to_c[(size_t)vc * b + j] += c * s[j];
    }
  }
  /* edge vectors (transgrid.cc:188-212): 0.5 damp_j b_j to the father
   * object of each of the edge's nodes, in the row's order */
  for (int v = 0; v < F->n; ++v) {
    if (vtype_of(F, v) != EDGE_TYPE) continue;
    for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
      int vc = F->pi[q];
      uint32_t vecskip = C->skip[vc];
      for (int j = 0; j < b; ++j)
        if (!(vecskip & (1u << j)))
          to_c[(size_t)vc * b + j] += F->pw[q] * damp[j] * from_f[(size_t)v * b + j];
    }
  }
}

void orc_interpolate(const orc_level *F, const orc_level *C, double *to_f,
                     const double *from_c, const double *damp) {
  const int b = F->b;
  (void)C;
  if (F->pblk) {
    interpolate_by_matrix(F, to_f, from_c, damp);
    return;
  }
  for (size_t i = 0; i < (size_t)F->n * b; ++i) to_f[i] = 0.0;
  for (int v = 0; v < F->n; ++v) {
    if (vtype_of(F, v) != 0) continue;
    uint32_t vecskip = F->skip[v];
    int skip = 1;
    for (int i = 0; i < b; ++i)
      if (!(vecskip & (1u << i))) skip = 0;
    if (skip) continue;
    for (int j = 0; j < b; ++j) {
      if (vecskip & (1u << j)) continue;
      for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
        double c = F->pw[q];
        if (c != 0.0)
          to_f[(size_t)v * b + j] += c * damp[j] * from_c[(size_t)F->pi[q] * b + j];
      }
    }
  }
  /* edge vectors (transgrid.cc:317-336): += 0.5 damp_j c_j of each father
   * object whose VECSKIP bit j is clear */
  for (int v = 0; v < F->n; ++v) {
    if (vtype_of(F, v) != EDGE_TYPE) continue;
    for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
      int vc = F->pi[q];
      uint32_t vecskip = C->skip[vc];
      for (int j = 0; j < b; ++j)
        if (!(vecskip & (1u << j)))
          to_f[(size_t)v * b + j] += F->pw[q] * damp[j] * from_c[(size_t)vc * b + j];
    }
  }
}

/* -------------------------------------------------------------------- ex */

struct orc_ex {
  int nv;         /* unknowns (vectors * b)             */
  int nvec;
  int b;
  int bw;
  int failed;
  int *order;     /* new vector list order               */
  int *index;     /* VINDEX per vector (component offset) */
  double *mat;    /* band, EX_MAT(m,bw,i,j) = m[2*bw*i+j] */
  double *vec;
};

#define EX_MAT(m, bw, i, j) ((m)[2 * (size_t)(bw) * (i) + (j)])

int orc_ex_create(const orc_level *L, int reorder, orc_ex **out) {
  orc_ex *ex = (orc_ex *)calloc(1, sizeof(orc_ex));
  const int n = L->n, b = L->b;
  ex->nvec = n;
  ex->b = b;
  ex->order = (int *)malloc(sizeof(int) * (size_t)n);
  ex->index = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) ex->order[i] = i;
  if (reorder && n > 0) {
    /* breadth-first search from the first vector, then a second one from
     * the last vector reached; the second visit order is the new list. */
    int *fifo = (int *)malloc(sizeof(int) * (size_t)n);
    char *used = (char *)calloc((size_t)n, 1);
    int head = 0, tail = 0, v = 0;
    fifo[tail++] = 0;
    used[0] = 1;
    while (head < tail) {
      v = fifo[head++];
      for (int m = L->rp[v] + 1; m < L->rp[v + 1]; ++m) {
        int w = L->ci[m];
        if (!used[w]) { fifo[tail++] = w; used[w] = 1; }
      }
    }
    head = tail = 0;
    fifo[tail++] = v;
    used[v] = 0;
    int i = 0;
    while (head < tail) {
      int u = fifo[head++];
      ex->order[i++] = u;
      for (int m = L->rp[u] + 1; m < L->rp[u + 1]; ++m) {
        int w = L->ci[m];
        if (used[w]) { fifo[tail++] = w; used[w] = 0; }
      }
    }
    free(fifo);
    free(used);
    if (i != n) { /* disconnected pattern: UG asserts here */
      orc_ex_free(ex);
      return 2;
    }
  }
  /* VINDEX and bandwidth (iter.cc:8813-8862) */
  for (int k = 0; k < n; ++k) ex->index[ex->order[k]] = k * b;
  int bw = 0;
  for (int v = 0; v < n; ++v)
    for (int m = L->rp[v] + 1; m < L->rp[v + 1]; ++m) {
      int k = ex->index[v] - ex->index[L->ci[m]];
      if (k < 0) k = -k;
      if (k > bw) bw = k;
    }
  ex->bw = b == 1 ? bw : bw + b - 1;
  ex->nv = n * b;
  bw = ex->bw;
  const size_t mem = (size_t)ex->nv * (2 * (size_t)bw + 1);
  ex->mat = (double *)calloc(mem ? mem : 1, sizeof(double));
  ex->vec = (double *)calloc((size_t)ex->nv + 1, sizeof(double));
  /* EXCopyMatrixDOUBLE (iter.cc:8518-8595) */
  for (int v = 0; v < n; ++v)
    for (int m = L->rp[v]; m < L->rp[v + 1]; ++m) {
      int ri = ex->index[v], cj = ex->index[L->ci[m]];
      for (int i = 0; i < b; ++i)
        for (int j = 0; j < b; ++j)
          EX_MAT(ex->mat, bw, ri + i, cj + j) = L->a[(size_t)m * b * b + i * b + j];
    }
  /* EXDecomposeMatrixDOUBLE (evm.cc:1400-1418) */
  const int nv = ex->nv;
  for (int i = 0; i < nv - 1; ++i) {
    double d = EX_MAT(ex->mat, bw, i, i);
    if (d == 0.0) { ex->failed = 1; break; }
    int jmax = i + bw < nv - 1 ? i + bw : nv - 1;
    for (int j = i + 1; j <= jmax; ++j) {
      double f = EX_MAT(ex->mat, bw, j, i) / d;
      EX_MAT(ex->mat, bw, j, i) = f;
      for (int k = i + 1; k <= jmax; ++k)
        EX_MAT(ex->mat, bw, j, k) -= f * EX_MAT(ex->mat, bw, i, k);
    }
  }
  *out = ex;
  return ex->failed;
}

void orc_ex_free(orc_ex *ex) {
  if (!ex) return;
  free(ex->order);
  free(ex->index);
  free(ex->mat);
  free(ex->vec);
  free(ex);
}

int orc_ex_bandwidth(const orc_ex *ex) { return ex->bw; }
const int *orc_ex_order(const orc_ex *ex) { return ex->order; }

int orc_ex_smoother(const orc_ex *ex, const orc_level *L, const double *damp,
                    double *x, double *b) {
  const int nb = L->b;
  if (ex->failed) {
    for (size_t i = 0; i < (size_t)L->n * nb; ++i) x[i] = 0.0;
    return 1;
  }
  const int n = ex->nv, bw = ex->bw;
  double *Vec = ex->vec;
  for (int k = 0; k < ex->nvec; ++k) {
    int v = ex->order[k];
    for (int j = 0; j < nb; ++j) Vec[k * nb + j] = b[(size_t)v * nb + j];
  }
  /* EXApplyLUDOUBLE (evm.cc:1493-1510) */
  for (int i = 1; i < n; ++i)
    for (int j = (i - bw > 0 ? i - bw : 0); j < i; ++j)
      Vec[i] -= EX_MAT(ex->mat, bw, i, j) * Vec[j];
  for (int i = n - 1; i >= 0; --i) {
    int jmax = i + bw < n - 1 ? i + bw : n - 1;
    for (int j = i + 1; j <= jmax; ++j) Vec[i] -= EX_MAT(ex->mat, bw, i, j) * Vec[j];
    Vec[i] /= EX_MAT(ex->mat, bw, i, i);
  }
  for (int k = 0; k < ex->nvec; ++k) {
    int v = ex->order[k];
    for (int j = 0; j < nb; ++j) x[(size_t)v * nb + j] = Vec[k * nb + j];
  }
  dscalx(L, x, damp);
  orc_dmatmul_minus(L, b, x);
  return 0;
}

/* ------------------------------------------------------------------ lmgc */

typedef struct {
  const orc_level *lv;
  int top;
  const orc_params *p;
  double **c, **b, **t;
  int **rank;       /* per level ordering for the GS smoothers */
  orc_ex *ex;
  double *cbase;    /* base solver correction (np_ls->c)       */
} orc_mg;

/* LinearSolver (ls.cc:637-749) with Iter = ex, called by Lmgc on the base
 * level after BaseSolver->Residuum. */
static int base_solver(orc_mg *mg, int level, const double *first) {
  const orc_level *L = &mg->lv[level];
  const orc_params *p = mg->p;
  const int b = L->b, nc = ncomp_of(L);
  double defect2reach[ORC_MAXB], last[ORC_MAXB];
  for (int i = 0; i < nc; ++i) {
    defect2reach[i] = first[i] * p->base_red[i];
    if (defect2reach[i] == 0.0) defect2reach[i] = p->base_red[i];
    last[i] = first[i];
  }
  int converged = sc_cmp(nc, first, p->base_abslimit, SC_IDENT(p));
  for (int i = 0; i < p->base_maxiter; ++i) {
    if (converged) break;
    memset(mg->cbase, 0, sizeof(double) * (size_t)L->n * b);
    if (p->base_iter == 1) { /* ls over jac: Smoother with l_jac */
      if (orc_smoother(L, ORC_JAC, NULL, p->base_damp, mg->cbase, mg->b[level])) return 1;
    } else if (orc_ex_smoother(mg->ex, L, p->base_damp, mg->cbase, mg->b[level])) {
      return 1;
    }
    dadd((size_t)L->n * b, mg->c[level], mg->cbase); /* LSUpdate */
    surface_nrm2x(L, mg->b[level], last);             /* LinearResiduum */
    if (sc_cmp(nc, last, p->base_abslimit, SC_IDENT(p)) || sc_cmp(nc, last, defect2reach, SC_IDENT(p))) {
      converged = 1;
      break;
    }
  }
  return 0;
}

static int lmgc(orc_mg *mg, int level) {
  const orc_params *p = mg->p;
  const orc_level *L = &mg->lv[level];
  const size_t len = (size_t)L->n * L->b;
  if (level <= p->baselevel) {
    double first[ORC_MAXB];
    surface_nrm2x(L, mg->b[level], first); /* BaseSolver->Residuum */
    return base_solver(mg, level, first);
  }
  double *t = mg->t[level];
  for (int i = 0; i < p->nu1; ++i) {
    if (orc_smoother(L, p->smoother, p->smoother == ORC_GS ? NULL : mg->rank[level], p->damp, t,
                     mg->b[level]))
      return 1;
    dadd(len, mg->c[level], t);
  }
  double one[ORC_MAXB]; /* Factor_One */
  for (int i = 0; i < ORC_MAXB; ++i) one[i] = 1.0;
  orc_restrict(L, &mg->lv[level - 1], mg->b[level - 1], mg->b[level], one);
  memset(mg->c[level - 1], 0, sizeof(double) * (size_t)mg->lv[level - 1].n * L->b);
  for (int g = 0; g < p->gamma; ++g)
    if (lmgc(mg, level - 1)) return 1;
  orc_interpolate(L, &mg->lv[level - 1], t, mg->c[level - 1], p->lmgc_damp);
  dadd(len, mg->c[level], t);
  orc_dmatmul_minus(L, mg->b[level], t);
  /* $S pre post base: the post object smooths after the correction */
  const int post = p->post_set ? p->post_smoother : p->smoother;
  const double *pdamp = p->post_set ? p->post_damp : p->damp;
  for (int i = 0; i < p->nu2; ++i) {
    if (orc_smoother(L, post, post == ORC_GS ? NULL : mg->rank[level], pdamp, t, mg->b[level]))
      return 1;
    dadd(len, mg->c[level], t);
  }
  return 0;
}

static int mg_setup(orc_mg *mg, const orc_level *levels, int top,
                    const orc_params *p) {
  memset(mg, 0, sizeof(*mg));
  mg->lv = levels;
  mg->top = top;
  mg->p = p;
  mg->c = (double **)calloc((size_t)top + 1, sizeof(double *));
  mg->b = (double **)calloc((size_t)top + 1, sizeof(double *));
  mg->t = (double **)calloc((size_t)top + 1, sizeof(double *));
  mg->rank = (int **)calloc((size_t)top + 1, sizeof(int *));
  for (int l = 0; l <= top; ++l) {
    size_t len = (size_t)levels[l].n * levels[l].b;
    mg->c[l] = (double *)calloc(len + 1, sizeof(double));
    mg->b[l] = (double *)calloc(len + 1, sizeof(double));
    mg->t[l] = (double *)calloc(len + 1, sizeof(double));
    if ((p->smoother == ORC_MCGS || (p->post_set && p->post_smoother == ORC_MCGS)) && l > p->baselevel) {
      int *color = (int *)malloc(sizeof(int) * (size_t)levels[l].n);
      orc_greedy_color(&levels[l], color);
      mg->rank[l] = (int *)malloc(sizeof(int) * (size_t)levels[l].n);
      for (int v = 0; v < levels[l].n; ++v)
        mg->rank[l][v] = color[v] * levels[l].n + v;
      free(color);
    }
  }
  int bl = p->baselevel < top ? p->baselevel : top;
  mg->cbase = (double *)calloc((size_t)levels[bl].n * levels[bl].b + 1, sizeof(double));
  return orc_ex_create(&levels[bl], p->ex_reorder, &mg->ex) ? 1 : 0;
}

static void mg_free(orc_mg *mg) {
  for (int l = 0; l <= mg->top; ++l) {
    free(mg->c[l]);
    free(mg->b[l]);
    free(mg->t[l]);
    free(mg->rank[l]);
  }
  free(mg->c);
  free(mg->b);
  free(mg->t);
  free(mg->rank);
  free(mg->cbase);
  orc_ex_free(mg->ex);
}

int orc_lmgc_iter(const orc_level *levels, int top, const orc_params *p,
                  double *c, double *b) {
  orc_mg mg;
  int err = mg_setup(&mg, levels, top, p);
  if (!err) {
    size_t len = (size_t)levels[top].n * levels[top].b;
    memcpy(mg.b[top], b, len * sizeof(double));
    memset(mg.c[top], 0, len * sizeof(double));
    err = lmgc(&mg, top);
    memcpy(c, mg.c[top], len * sizeof(double));
    memcpy(b, mg.b[top], len * sizeof(double));
  }
  mg_free(&mg);
  return err;
}

int orc_ls_lmgc_solve(const orc_level *levels, int top, const orc_params *p,
                      double *x, double *b, int do_defect,
                      const double *abslimit, const double *reduction,
                      int maxiter, double *defects, int *nit, int *converged) {
  orc_mg mg;
  const orc_level *L = &levels[top];
  const int nb = ncomp_of(L);                      /* norm entries */
  const size_t len = (size_t)L->n * L->b;
  int err = mg_setup(&mg, levels, top, p);
  if (err) { mg_free(&mg); return err; }
  if (do_defect) surface_dmatmul_minus(L, b, x); /* LinearDefect */
  double first[ORC_MAXB], last[ORC_MAXB], defect2reach[ORC_MAXB];
  surface_nrm2x(L, b, first);                       /* LinearResiduum */
  for (int i = 0; i < nb; ++i) {
    defects[i] = first[i];
    defect2reach[i] = first[i] * reduction[i];     /* sc_mul_check */
    if (defect2reach[i] == 0.0) defect2reach[i] = reduction[i];
  }
  int conv = sc_cmp(nb, first, abslimit, SC_IDENT(p));
  int it = 0;
  int i;
  for (i = 0; i < maxiter; ++i) {
    if (conv) break;
    memcpy(mg.b[top], b, len * sizeof(double));
    memset(mg.c[top], 0, len * sizeof(double));     /* dset(c,0) */
    if (lmgc(&mg, top)) { err = 1; break; }
    memcpy(b, mg.b[top], len * sizeof(double));
    dadd(len, x, mg.c[top]);                        /* LSUpdate */
    surface_nrm2x(L, b, last);
    for (int k = 0; k < nb; ++k) defects[(size_t)(i + 1) * nb + k] = last[k];
    if (sc_cmp(nb, last, abslimit, SC_IDENT(p)) || sc_cmp(nb, last, defect2reach, SC_IDENT(p))) {
      conv = 1;
      it = i + 1;
      break;
    }
  }
  if (!conv) it = i;
  *nit = it;
  *converged = conv;
  mg_free(&mg);
  return err;
}

/* ------------------------------------------------ composite surface */

/* FINE_GRID_DOF (SetSurfaceClasses, gm/algebra.cc:2908-2909) */
static int fine_grid_dof(const orc_level *L, int v) {
  return L->vclass[v] >= NEWDEF_CLASS && L->vnclass[v] <= 1;
}

/* FULLREFINELEVEL as SetSurfaceClasses leaves it (algebra.cc:2885-2916):
 * the lowest level holding a FINE_GRID_DOF vector, the top level if none */
int orc_fullrefinelevel(const orc_level *levels, int top) {
  int f = top;
  for (int l = top; l >= 0; --l)
    for (int v = 0; v < levels[l].n; ++v)
      if (fine_grid_dof(&levels[l], v)) { f = l; break; }
  return f;
}

/* row v of level l on the surface of levels frl..tl (vecloop.ct:22-39,
 * matloop.ct:22-50): FINE_GRID_DOF below tl, NEW_DEFECT on tl */
static int on_surface(const orc_level *L, int l, int tl, int v) {
  return l < tl ? fine_grid_dof(L, v) : L->vclass[v] >= NEWDEF_CLASS;
}

/* dmatmul_minus ON_SURFACE (ugblas.cc:3997-4042 in the T_SURFACE_MODE loop
 * of matmode.ct): level by level from FULLREFINELEVEL, x_l -= A_l y_l */
static void composite_dmatmul_minus(const orc_level *levels, int frl, int tl,
                                    double **x, double *const *y) {
  for (int l = frl; l <= tl; ++l) {
    const orc_level *L = &levels[l];
    const int b = L->b, bb = b * b;
    for (int v = 0; v < L->n; ++v) {
      if (!on_surface(L, l, tl, v)) continue;
      double s[ORC_MAXB] = {0.0};
      for (int m = L->rp[v]; m < L->rp[v + 1]; ++m) {
        const double *M = L->a + (size_t)m * bb;
        const double *w = y[l] + (size_t)L->ci[m] * b;
        for (int i = 0; i < b; ++i) {
          double t = 0.0;
          for (int j = 0; j < b; ++j) t += M[i * b + j] * w[j];
          s[i] += t;
        }
      }
      for (int i = 0; i < b; ++i) x[l][(size_t)v * b + i] -= s[i];
    }
  }
}

/* dnrm2x ON_SURFACE (ugblas.cc:3156-3176): one running sum per (type,
 * component) over the surface rows, level by level, then sqrt */
static void composite_nrm2x(const orc_level *levels, int frl, int tl,
                            double *const *x, double *out) {
  const int b = levels[tl].b, nc = ncomp_of(&levels[tl]);
  double a[ORC_MAXB] = {0.0};
  for (int l = frl; l <= tl; ++l) {
    const orc_level *L = &levels[l];
    for (int v = 0; v < L->n; ++v) {
      if (!on_surface(L, l, tl, v)) continue;
      for (int i = 0; i < b; ++i) {
        double s = x[l][(size_t)v * b + i];
        a[aoff(L, v) + i] += s * s;
      }
    }
  }
  for (int i = 0; i < nc; ++i) out[i] = sqrt(a[i]);
}

int orc_ls_lmgc_solve_surface(const orc_level *levels, int top, const orc_params *p,
                              double **x, double **b, int do_defect,
                              const double *abslimit, const double *reduction,
                              int maxiter, double *defects, int *nit, int *converged) {
  orc_mg mg;
  const int vb = levels[top].b;              /* components per vector */
  const int nb = ncomp_of(&levels[top]);      /* norm entries */
  const int bl = p->baselevel < top ? p->baselevel : top;
  const int frl = orc_fullrefinelevel(levels, top);
  if (frl < bl) return 3;   /* surface rows below the base level */
  int err = mg_setup(&mg, levels, top, p);
  if (err) { mg_free(&mg); return err; }
  /* the defect recurrence runs in the coarse levels' b too */
  for (int l = bl; l <= top; ++l)
    memcpy(mg.b[l], b[l], sizeof(double) * (size_t)levels[l].n * vb);
  if (do_defect) composite_dmatmul_minus(levels, frl, top, mg.b, x); /* LinearDefect */
  double first[ORC_MAXB], last[ORC_MAXB], defect2reach[ORC_MAXB];
  composite_nrm2x(levels, frl, top, mg.b, first);                    /* LinearResiduum */
  for (int i = 0; i < nb; ++i) {
    defects[i] = first[i];
    defect2reach[i] = first[i] * reduction[i];
    if (defect2reach[i] == 0.0) defect2reach[i] = reduction[i];
  }
  int conv = sc_cmp(nb, first, abslimit, SC_IDENT(p));
  int it = 0;
  int i;
  for (i = 0; i < maxiter; ++i) {
    if (conv) break;
    memset(mg.c[top], 0, sizeof(double) * (size_t)levels[top].n * vb); /* dset(c,0) */
    if (lmgc(&mg, top)) { err = 1; break; }
    for (int l = bl; l <= top; ++l)                 /* LSUpdate: baselevel..level */
      dadd((size_t)levels[l].n * vb, x[l], mg.c[l]);
    composite_nrm2x(levels, frl, top, mg.b, last);
    for (int k = 0; k < nb; ++k) defects[(size_t)(i + 1) * nb + k] = last[k];
    if (sc_cmp(nb, last, abslimit, SC_IDENT(p)) || sc_cmp(nb, last, defect2reach, SC_IDENT(p))) {
      conv = 1;
      it = i + 1;
      break;
    }
  }
  if (!conv) it = i;
  for (int l = bl; l <= top; ++l)
    memcpy(b[l], mg.b[l], sizeof(double) * (size_t)levels[l].n * vb);
  *nit = it;
  *converged = conv;
  mg_free(&mg);
  return err;
}

/* ------------------------------------------------------------- BiCGStab */

static double ddotw_surface(const orc_level *L, const double *x, const double *y) {
  /* ddotw (ugblas.cc:3023-3044) with unit weights, ON_SURFACE */
  const int b = L->b;
  double a[ORC_MAXB] = {0.0};
  for (int v = 0; v < L->n; ++v) {
    if (L->vclass[v] < NEWDEF_CLASS) continue;
    for (int i = 0; i < b; ++i) a[i] += x[(size_t)v * b + i] * y[(size_t)v * b + i];
  }
  double s = 0.0;
  for (int i = 0; i < b; ++i) s += a[i];
  return s;
}

/* sc_eq (npscan.cc:1095-1104) */
static int sc_eq(int b, const double *x, const double *y, double ac) {
  for (int i = 0; i < b; ++i)
    if (x[i] < 0.0 || y[i] < 0.0 || fabs(x[i] - y[i]) > ac * sqrt(x[i] * y[i])) return 0;
  return 1;
}

/* BCGSSolver (np/procs/ls.cc:1864-2062) with Iter = lmgc (or none).
 * defects receives the first defect and every LinearResiduum of the loop
 * (two per iteration), ndef their count. */
int orc_bcgs_solve(const orc_level *levels, int top, const orc_params *p, int use_iter,
                   double *x, double *b, int do_defect, const double *abslimit,
                   const double *reduction, int maxiter, int restart, double *defects,
                   int maxdef, int *ndef, int *nit, int *converged) {
  orc_mg mg;
  const orc_level *L = &levels[top];
  const int nb = L->b;
  const size_t len = (size_t)L->n * nb;
  int err = mg_setup(&mg, levels, top, p);
  if (err) { mg_free(&mg); return err; }
  double *r = (double *)calloc(len + 1, sizeof(double));
  double *pv = (double *)calloc(len + 1, sizeof(double));
  double *v = (double *)calloc(len + 1, sizeof(double));
  double *sv = (double *)calloc(len + 1, sizeof(double));
  double *t = (double *)calloc(len + 1, sizeof(double));
  double *q = (double *)calloc(len + 1, sizeof(double));
  double first[ORC_MAXB], last[ORC_MAXB], old[ORC_MAXB], d2r[ORC_MAXB];
  int nd = 0, its = 0, conv = 0;
  if (do_defect) surface_dmatmul_minus(L, b, x);
  surface_nrm2x(L, b, first);
  for (int i = 0; i < nb; ++i) {
    last[i] = first[i];
    old[i] = -1.0;
    d2r[i] = first[i] * reduction[i];
    if (d2r[i] == 0.0) d2r[i] = reduction[i];
    if (nd < maxdef) defects[nd * nb + i] = first[i];
  }
  ++nd;
  conv = sc_cmp(nb, first, abslimit, SC_IDENT(p));
  double alpha = 0.0, rho = 0.0, omega = 0.0, rho_new = 0.0, beta = 0.0, tt = 0.0;
  int need_restart = 1, eq_count = 0;
  /* Iter(q, z): q = lmgc correction of defect z (z left untouched) */
  #define BCGS_ITER(Z)                                                  \
    do {                                                                \
      memcpy(mg.b[top], (Z), len * sizeof(double));                     \
      memset(mg.c[top], 0, len * sizeof(double));                       \
      if (lmgc(&mg, top)) { err = 1; goto done; }                       \
      memcpy(q, mg.c[top], len * sizeof(double));                       \
    } while (0)
  for (int i = 0; i < maxiter; ++i) {
    if (conv) break;
    if ((restart > 0 && i % restart == 0) || need_restart) {
      memset(pv, 0, len * sizeof(double));
      memset(v, 0, len * sizeof(double));
      memcpy(r, b, len * sizeof(double));
      alpha = rho = omega = 1.0;
      need_restart = 0;
    }
    rho_new = ddotw_surface(L, b, r);
    if (rho != 0.0 && omega != 0.0) beta = rho_new * alpha / rho / omega;
    for (size_t k = 0; k < len; ++k) pv[k] *= beta;                    /* dscal */
    for (size_t k = 0; k < len; ++k) pv[k] += b[k];                    /* dadd  */
    for (size_t k = 0; k < len; ++k) pv[k] += -beta * omega * v[k];   /* daxpy */
    if (use_iter) {
      BCGS_ITER(pv);
      orc_dmatmul(L, v, q);   /* v = A q (all rows = surface here) */
      alpha = ddotw_surface(L, v, r);
      if (alpha != 0.0) alpha = rho_new / alpha;
      for (size_t k = 0; k < len; ++k) x[k] += alpha * q[k];
    } else {
      orc_dmatmul(L, v, pv);
      alpha = ddotw_surface(L, v, r);
      if (alpha != 0.0) alpha = rho_new / alpha;
      for (size_t k = 0; k < len; ++k) x[k] += alpha * pv[k];
    }
    ++its;
    for (size_t k = 0; k < len; ++k) sv[k] = b[k] + -alpha * v[k];
    surface_nrm2x(L, sv, last);
    if (nd < maxdef) for (int c = 0; c < nb; ++c) defects[nd * nb + c] = last[c];
    ++nd;
    if (sc_cmp(nb, last, abslimit, SC_IDENT(p)) || sc_cmp(nb, last, d2r, SC_IDENT(p))) {
      memcpy(b, sv, len * sizeof(double));
      conv = 1;
      break;
    }
    if (use_iter) BCGS_ITER(sv);
    else memcpy(q, sv, len * sizeof(double));
    orc_dmatmul(L, t, q);
    tt = ddotw_surface(L, t, t);
    omega = ddotw_surface(L, sv, t);
    if (tt != 0.0) omega /= tt;
    for (size_t k = 0; k < len; ++k) x[k] += omega * q[k];
    for (size_t k = 0; k < len; ++k) b[k] = sv[k] + -omega * t[k];
    rho = rho_new;
    surface_nrm2x(L, b, last);
    if (nd < maxdef) for (int c = 0; c < nb; ++c) defects[nd * nb + c] = last[c];
    ++nd;
    ++its;
    if (sc_cmp(nb, last, abslimit, SC_IDENT(p)) || sc_cmp(nb, last, d2r, SC_IDENT(p))) { conv = 1; break; }
    if (sc_eq(nb, last, old, 1e-4)) ++eq_count;
    else eq_count = 0;
    for (int c = 0; c < nb; ++c) old[c] = last[c];
    if (eq_count > 4) { conv = 0; break; }
  }
  #undef BCGS_ITER
done:
  *ndef = nd;
  *nit = its;
  *converged = conv;
  free(r); free(pv); free(v); free(sv); free(t); free(q);
  mg_free(&mg);
  return err;
}

/* -------------------------------------------------------------- Galerkin */

/* AssembleGalerkinByMatrix (np/algebra/transgrid.cc:1575-1700) with the
 * interpolation of F as the interpolation matrices: for every fine vector v
 * (list order), every MATRIX m = (v, w) of its row (diagonal first), every
 * interpolation entry (I, wi) of v and (J, wj) of w:
 *   A_c[I][J] += (a_m wi) wj      (blocks: componentwise)
 * into the pattern of C; a product whose connection is missing from the
 * pattern goes to the diagonal of I (the branch taken when
 * CreateExtraConnection fails).  ac (nnzb_C * b * b) is cleared first. */
void orc_galerkin(const orc_level *F, const orc_level *C, double *ac) {
  const int b = F->b, bb = b * b;
  const size_t nz = (size_t)C->rp[C->n] * bb;
  for (size_t i = 0; i < nz; ++i) ac[i] = 0.0;
  for (int v = 0; v < F->n; ++v)
    for (int m = F->rp[v]; m < F->rp[v + 1]; ++m) {
      const int w = F->ci[m];
      const double *M = F->a + (size_t)m * bb;
      for (int q = F->pp[v]; q < F->pp[v + 1]; ++q) {
        const int I = F->pi[q];
        const double wi = F->pw[q];
        for (int r = F->pp[w]; r < F->pp[w + 1]; ++r) {
          const int J = F->pi[r];
          const double wj = F->pw[r];
          int slot = -1;
          for (int k = C->rp[I]; k < C->rp[I + 1]; ++k)
            if (C->ci[k] == J) { slot = k; break; }
          if (slot < 0) slot = C->rp[I];   /* diagonal first */
          double *T = ac + (size_t)slot * bb;
          for (int e = 0; e < bb; ++e) T[e] += (M[e] * wi) * wj;
        }
      }
    }
}
