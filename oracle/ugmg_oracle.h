/*
 * ugmg_oracle.h -- CPU restatement of UG 3.12.1's linear multigrid solve
 * path, for TESTS ONLY.
 *
 * This is test infrastructure: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker (or the timed
 * CPU baseline).  The product library (libugmg) never links or calls it.
 *
 * Every function restates the UG routine named in its comment, in UG's loop
 * order (vector list order, row lists diagonal first), in FP64, single
 * threaded.  Data come in ConvertMatrix layout (np/udm/disctools.cc:
 * 2061-2128): per level n vectors of b components, row_ptr/col_idx over
 * block entries with the diagonal block first in each row, b x b row-major
 * blocks; vector flags VECSKIP / VCLASS / VNCLASS; the standard node
 * interpolation from level-1 as CSR rows of (coarse corner, GNs weight).
 *
 * Parity pins (tests/test_oracle_*.py): dense numpy products, closed-form
 * Q1 transfer weights and invariants, scipy triangular and dense solves,
 * and the textbook multigrid error-propagation operator
 * E = S^nu2 (I - P (I - E_c) A_c^{-1} R A) S^nu1.
 */
#ifndef UGMG_ORACLE_H
#define UGMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXB 4

/* Vector types (VTYPE): 0 node vectors, 1 edge vectors (the two the
 * standard transfer serves, transgrid.cc:461-503).  Every type carries b
 * components; norms, dot products and VEC_SCALAR parameters have one entry
 * per (type, component), type-major (VD_OFFSET, np/udm/udm.h:105). */
typedef struct {
  int n;                  /* vectors on the level                         */
  int b;                  /* components per vector (1 .. ORC_MAXB)        */
  const int *rp;          /* n+1                                          */
  const int *ci;          /* nnzb; diagonal first in every row            */
  const double *a;        /* nnzb*b*b                                     */
  const uint32_t *skip;   /* VECSKIP bit j = component j Dirichlet        */
  const uint8_t *vclass;  /* VCLASS                                       */
  const uint8_t *vnclass; /* VNCLASS                                      */
  const int *pp;          /* interpolation from level-1 (NULL on level 0) */
  const int *pi;          /* node rows: father corners with GNs weights;  */
  const double *pw;       /* edge rows: the father objects of the edge's  */
                          /* two nodes (NFATHEREDGE / NFATHER), 0.5 each  */
  const uint8_t *vtype;   /* VTYPE per vector, NULL: all node vectors     */
  int ntypes;             /* types with components (0 read as 1); the     */
                          /* norms have ntypes * b entries                */
  const double *pblk;     /* interpolation matrix (transfer $M): a b x b  */
                          /* block per entry of pp/pi, [i*b+j] = coarse   */
                          /* component i, fine component j; NULL: pw      */
} orc_level;

enum { ORC_JAC = 0, ORC_GS = 1, ORC_MCGS = 2 };

typedef struct {
  int nu1, nu2, gamma, baselevel;
  int smoother;                  /* ORC_JAC, ORC_GS (l_lgs), ORC_MCGS     */
  double damp[ORC_MAXB];         /* smoother $damp                         */
  double lmgc_damp[ORC_MAXB];    /* lmgc $damp (interpolation damping)     */
  double base_damp[ORC_MAXB];    /* ex $damp                               */
  double base_abslimit[ORC_MAXB];/* base ls $abslimit                      */
  double base_red[ORC_MAXB];     /* base ls $red                           */
  int base_maxiter;              /* base ls $m                             */
  int ex_reorder;                /* ex $o                                  */
  int has_ident;                 /* VD_NID != NO_IDENT                     */
  int ident[ORC_MAXB];           /* VD_IDENT per norm entry (sc_cmp)       */
  int post_set;                  /* 0: post-smoother = pre-smoother        */
  int post_smoother;             /* the post-smoother's kind and $damp     */
  double post_damp[ORC_MAXB];
  int base_iter;                 /* the base ls's $I: 0 ex, 1 jac ($damp =
                                    base_damp)                             */
} orc_params;

/* ugblas (np/algebra/ugblas.cc) on one level, ALL_VECTORS */
void orc_dmatmul_minus(const orc_level *L, double *x, const double *y);
void orc_dmatmul(const orc_level *L, double *x, const double *y);
void orc_dnrm2x(const orc_level *L, const double *x, double *out /* ntypes * b */);
void orc_ddot(const orc_level *L, const double *x, const double *y, double *out);

/* SolveSmallBlock (np/algebra/block.cc:104-205); 0 ok, 1 singular */
int orc_solve_small_block(int n, const double *mat, const double *rhs,
                          double *sol);

/* l_jac (np/algebra/ugiter.cc:271-335) */
int orc_l_jac(const orc_level *L, double *v, const double *d);
/* l_lgs (np/algebra/ugiter.cc:412-615) with VINDEX replaced by rank[]:
 * rank = list position gives UG's lexicographic GS (l_setindex,
 * ugiter.cc:169-186); rank = colour*n + position gives the multicolour
 * ordering of the GPU smoother.  rank == NULL means list position. */
int orc_l_lgs(const orc_level *L, const int *rank, double *v, const double *d);
/* greedy first-fit colouring in list order over the matrix pattern;
 * returns the number of colours */
int orc_greedy_color(const orc_level *L, int *color);

/* Smoother (np/procs/iter.cc:817-842): Step, x *= damp, b -= A x */
int orc_smoother(const orc_level *L, int kind, const int *rank,
                 const double *damp, double *x, double *b);

/* StandardRestrictNodeVector / StandardIntCorNodeVector
 * (np/algebra/transgrid.cc:117-215, 235-336): the node loop, then the edge
 * loop over the edge vectors (no VCLASS test; each father object's VECSKIP
 * bits mask its term; the fine edge vector's own bits are not read).
 * With F->pblk: RestrictByMatrix / InterpolateCorrectionByMatrix
 * (transgrid.cc:1160-1232, 1300-1380) over the interpolation matrix */
void orc_restrict(const orc_level *F, const orc_level *C, double *to_c,
                  const double *from_f, const double *damp);
void orc_interpolate(const orc_level *F, const orc_level *C, double *to_f,
                     const double *from_c, const double *damp);

/* ex band LU (np/procs/iter.cc:8685-9071, gm/evm.cc:1400-1510) */
typedef struct orc_ex orc_ex;
int orc_ex_create(const orc_level *L, int reorder, orc_ex **out); /* 0 ok, 1 zero pivot */
void orc_ex_free(orc_ex *ex);
int orc_ex_bandwidth(const orc_ex *ex);
const int *orc_ex_order(const orc_ex *ex);       /* new list order        */
/* EXSmoother: x = damp * LU^{-1} b; b -= A x */
int orc_ex_smoother(const orc_ex *ex, const orc_level *L, const double *damp,
                    double *x, double *b);

/* The ls + lmgc + smoother + transfer + ex solve (np/procs/ls.cc:637-749,
 * np/procs/iter.cc:7741-7949).  levels[0..top].  x,b are level `top`
 * vectors; b is the right-hand side on entry (the defect recurrence runs
 * in it).  do_defect: apply LinearDefect (b -= A x) first.
 * defects receives (maxiter+1)*b norms: first defect then one per
 * iteration; *nit the iteration count, *converged the flag.
 * Returns 0 or a nonzero error. */
int orc_ls_lmgc_solve(const orc_level *levels, int top, const orc_params *p,
                      double *x, double *b, int do_defect,
                      const double *abslimit, const double *reduction,
                      int maxiter, double *defects, int *nit, int *converged);

/* FULLREFINELEVEL of SetSurfaceClasses (gm/algebra.cc:2885-2916): the
 * lowest level with a FINE_GRID_DOF vector (VCLASS >= 2, VNCLASS <= 1), the
 * top level if there is none. */
int orc_fullrefinelevel(const orc_level *levels, int top);
/* The same solve on a locally refined hierarchy: x[l], b[l] for every level
 * l in baselevel..top (entries below may be NULL).  LinearDefect and
 * LinearResiduum run ON_SURFACE from FULLREFINELEVEL (FINE_GRID_DOF rows on
 * the lower levels, NEW_DEFECT rows on top; vecloop.ct / matloop.ct), the
 * coarse levels' b carry the defect recurrence from call to call, and
 * LSUpdate adds c into x on baselevel..top (ls.cc:869-875).  Returns 3 when
 * FULLREFINELEVEL lies below the base level. */
int orc_ls_lmgc_solve_surface(const orc_level *levels, int top, const orc_params *p,
                              double **x, double **b, int do_defect,
                              const double *abslimit, const double *reduction,
                              int maxiter, double *defects, int *nit, int *converged);
/* One lmgc iteration on `top` (NP_ITER semantics): c overwritten with the
 * correction, b := b - A c.  Work vectors for the coarse levels are
 * allocated internally. */
int orc_lmgc_iter(const orc_level *levels, int top, const orc_params *p,
                  double *c, double *b);

/* BCGSSolver (np/procs/ls.cc:1864-2062), right-preconditioned by one
 * lmgc cycle per application (use_iter) or unpreconditioned.  defects gets
 * the first defect and each LinearResiduum of the loop (up to maxdef). */
int orc_bcgs_solve(const orc_level *levels, int top, const orc_params *p, int use_iter,
                   double *x, double *b, int do_defect, const double *abslimit,
                   const double *reduction, int maxiter, int restart, double *defects,
                   int maxdef, int *ndef, int *nit, int *converged);

/* AssembleGalerkinByMatrix (np/algebra/transgrid.cc:1575-1700): coarse
 * values A_c = P^T A_f P (scalar interpolation weights, blocks
 * componentwise) in C's pattern, missing connections lumped to the
 * diagonal; ac has nnzb_C * b * b doubles. */
void orc_galerkin(const orc_level *F, const orc_level *C, double *ac);

#ifdef __cplusplus
}
#endif
#endif
