/*
 * ugmg.h -- C ABI of the MI355X multigrid solve path (libugmg.so).
 *
 * Replaces, behind UG's numproc interface, the CPU solve path
 *   lmgc      Lmgc / LmgcPreProcess / LmgcPostProcess   np/procs/iter.cc:7707-7984
 *   jac, gs   Smoother + JacobiStep / GSStep            np/procs/iter.cc:817-1087
 *   transfer  RestrictDefect / InterpolateCorrection     np/procs/transfer.cc:724-773
 *   ex        EXPreProcess / EXSmoother (band LU)        np/procs/iter.cc:8685-9071
 *   ls        LinearDefect / LinearResiduum / LinearSolver np/procs/ls.cc:562-749
 *   ugblas    dmatmul_minus / dmatmul / dnrm2x / ddot    np/algebra/ugblas.cc:2291-4042
 * A UG numproc class flattens the MATRIX lists of each
 * level in ConvertMatrix order (np/udm/disctools.cc:2061-2128) and hands the
 * arrays over here; everything after that runs on the GPU.
 *
 * Conventions
 *  - Every function returns 0 (UGMG_OK) or a negative UGMG_E_* code; the
 *    message is available from ugmg_last_error().  No C++ exception crosses
 *    this boundary.
 *  - Vectors are `n*b` doubles, component-fastest, in the level's vector
 *    list order.  Functions whose name ends in _host take host pointers and
 *    copy; all others take DEVICE pointers and are stream ordered on the
 *    context's stream (ugmg_set_stream), returning before the GPU finishes
 *    unless they report a norm.
 *  - Arrays passed in are caller owned and copied; device memory belongs to
 *    the context.
 *  - One context per host thread; calls on one context are not re-entrant.
 */
#ifndef UGMG_H
#define UGMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UGMG_MAXB 4

enum {
  UGMG_OK = 0,
  UGMG_E_ARG = -1,       /* invalid argument                               */
  UGMG_E_HIP = -2,       /* HIP runtime error (message has the HIP string) */
  UGMG_E_STATE = -3,     /* call order (e.g. solve before preprocess)      */
  UGMG_E_NOMEM = -4,     /* device or host allocation failed               */
  UGMG_E_SINGULAR = -5,  /* zero pivot in the coarse LU / singular block   */
  UGMG_E_BUDGET = -6,    /* request does not fit the HBM budget            */
  UGMG_E_NODEV = -7      /* no GPU                                          */
};

enum { UGMG_SMOOTH_JAC = 0, UGMG_SMOOTH_GS = 1 };   /* GS = multicolour GS  */
enum { UGMG_FP64 = 0, UGMG_MIXED = 1 };
enum { UGMG_BASE_EX = 0, UGMG_BASE_JAC = 1 };   /* the base ls's $I            */

/* lmgc / smoother / base-solver settings, named after the numproc options. */
typedef struct {
  int nu1, nu2, gamma, baselevel;   /* lmgc $n1 $n2 $g $b                    */
  int smoother;                     /* UGMG_SMOOTH_JAC / UGMG_SMOOTH_GS      */
  double damp[UGMG_MAXB];           /* smoother $damp                         */
  double lmgc_damp[UGMG_MAXB];      /* lmgc $damp (interpolation damping)     */
  double base_damp[UGMG_MAXB];      /* ex $damp                               */
  double base_abslimit[UGMG_MAXB];  /* base ls $abslimit                      */
  double base_red[UGMG_MAXB];       /* base ls $red                           */
  int base_maxiter;                 /* base ls $m                             */
  int ex_reorder;                   /* ex $o (BFS band reorder)               */
  int precision;                    /* UGMG_FP64 or UGMG_MIXED                */
  int inner_cycles;                 /* MIXED: FP32 V-cycles per FP64 step     */
  int use_graph;                    /* replay the cycle from a hipGraph       */
  int ex_float;                     /* ex $f: band LU in FLOAT arithmetic on
                                       FLOAT entries (EXDecomposeMatrixFLOAT),
                                       applied in DOUBLE (EXApplyLUFLOAT)     */
  int inner_adapt;                  /* MIXED: a refinement step runs fewer than
                                       inner_cycles FP32 cycles when the rate
                                       of the previous steps predicts that
                                       fewer reach the stop test (default 1) */
  int base_inv32;                   /* MIXED: the FP32 cycle's base solve reads
                                       an FP32 copy of the dense coarse inverse
                                       (default 0: the FP64 inverse, sums in
                                       FP64)                                   */
  int post_set;                     /* 0 (default): lmgc's post-smoother is its
                                       pre-smoother; 1: post_smoother and
                                       post_damp (lmgc $S pre post base with
                                       two different smoother objects)        */
  int post_smoother;                /* UGMG_SMOOTH_JAC / UGMG_SMOOTH_GS       */
  double post_damp[UGMG_MAXB];      /* post-smoother $damp                    */
  int base_iter;                    /* the base ls's iteration: UGMG_BASE_EX
                                       (default: ex, the dense inverse) or
                                       UGMG_BASE_JAC (jac with base_damp as
                                       its $damp, base_maxiter steps, the
                                       base stop test after each)            */
} ugmg_params;

/* LRESULT (np/procs/ls.h:71-77) */
typedef struct {
  int error_code;
  int converged;
  double first_defect[UGMG_MAXB];
  double last_defect[UGMG_MAXB];
  int number_of_linear_iterations;
} ugmg_lresult;

/* called after every iteration with the per-component defect (DoPCR hook) */
typedef void (*ugmg_defect_cb)(void *user, int iteration, int ncomp,
                               const double *defect);

typedef struct ugmg_ctx ugmg_ctx;

/* defaults of the numprocs: n1=n2=g=1, b=0, damp 1, base ls $m 1 $red 1e-12
 * $abslimit 1e-10 (ABS_LIMIT, ls.cc:66), ex $o 1, FP64, no graph. */
void ugmg_params_default(ugmg_params *p);

int ugmg_create(int device, ugmg_ctx **out);
int ugmg_destroy(ugmg_ctx *ctx);
const char *ugmg_last_error(const ugmg_ctx *ctx);
int ugmg_set_stream(ugmg_ctx *ctx, void *hip_stream);  /* NULL: own stream */
int ugmg_sync(ugmg_ctx *ctx);

/* ---- hierarchy (LmgcPreProcess input) ----------------------------------
 * ugmg_set_nlevels first, then one ugmg_set_level per level (0 = coarsest)
 * and one ugmg_set_transfer per level >= 1, then ugmg_set_params and
 * ugmg_preprocess.  row_ptr/col_idx/vals: ConvertMatrix order, diagonal
 * block first in each row, vals = nnzb*b*b row-major blocks.  skip =
 * VECSKIP bits, vclass/vnclass = VCLASS/VNCLASS (gm/gm.h:2200-2288).
 * Interpolation rows: coarse corners of the father element with their GNs
 * weights (StandardIntCorNodeVector, transgrid.cc:235-336). */
/* b: components per vector, 1 .. UGMG_MAXB (b x b blocks: Jacobi and GS
 * with SolveSmallBlock's n = 2 and n = 3 forms and, for n = 4, its Gauss
 * elimination with row pivoting where |pivot| < 1e-10; per-component
 * Dirichlet bits,
 * norms and stop tests; damping factors per component (dscalx: the
 * smoother's x, lmgc's interpolated correction and ex's x scaled per
 * component, ugmg_restrict damping each fine term or, over an
 * interpolation matrix, the reset coarse rows); one base $abslimit / $red
 * for all components; bcgs and the server take b = 1) */
int ugmg_set_nlevels(ugmg_ctx *ctx, int nlevels, int b);
int ugmg_set_level(ugmg_ctx *ctx, int level, int n, const int *row_ptr,
                   const int *col_idx, const double *vals,
                   const uint32_t *skip, const uint8_t *vclass,
                   const uint8_t *vnclass);
int ugmg_set_transfer(ugmg_ctx *ctx, int level, const int *p_ptr,
                      const int *p_idx, const double *p_w);
/* The interpolation matrix of `level` (UG's transfer $M: RestrictByMatrix /
 * InterpolateCorrectionByMatrix, transgrid.cc:1160-1380): per fine vector the
 * same rows as ugmg_set_transfer, each entry a b x b block (b*b doubles,
 * p_blocks[q*b*b + i*b + j] = coarse component i, fine component j: MVALUE
 * of the IMAT entry).  Restriction b_c,i += sum_j M_ij b_f,j over fine
 * vectors of VCLASS >= 2 on the coarse components whose VECSKIP bit is
 * clear (coarse rows of VNCLASS >= 2 reset first, $damp scaling them after
 * the sum); interpolation t_f,i = $damp sum_j M_ji c_j on the fine
 * components whose VECSKIP bit is clear.  Replaces the level's weights;
 * ugmg_set_transfer returns it to the standard transfer.  Levels with a
 * matrix are laid out on the host; served groups and ugmg_galerkin take
 * scalar weights.  b = 1: the scalar kernels (a $damp other than 1 in
 * ugmg_restrict scales each term). */
int ugmg_set_transfer_matrix(ugmg_ctx *ctx, int level, const int *p_ptr, const int *p_idx,
                             const double *p_blocks);
/* Vector types of a level's vectors (VTYPE, np/udm/udm.h:105-108): 0 node
 * vectors, 1 edge vectors, each with the b components of the hierarchy;
 * NULL (or never called): node vectors only.  Call after ugmg_set_level
 * (which resets the level to node vectors).  With edge vectors present:
 *  - norms, dot products and the VEC_SCALARs of the stop tests
 *    (abslimit, reduction, ugmg_lresult's defects, the per-iteration
 *    callback) take one entry per (type, component), type-major (VD_OFFSET),
 *    at most UGMG_MAXB of them; a damping factor per component applies
 *    alike to both types, one base $abslimit / $red to every entry;
 *  - the transfer rows of an edge vector name the father objects of the
 *    edge's two nodes (the father node's vector for a corner's son, the
 *    father edge's vector for a midnode) with weight 1/2 each: the edge
 *    loops of StandardRestrictNodeVector / StandardIntCorNodeVector
 *    (transgrid.cc:188-212, 317-336), which read no VCLASS of the edge
 *    vector in the restriction, mask each father object's term by its own
 *    VECSKIP bits in the interpolation and not by the edge vector's; a
 *    father object with only some components Dirichlet is refused (b > 1);
 *  - ugmg_solve and ugmg_iter run; the surface solve, bcgs and the server
 *    take node vectors only (UGMG_E_ARG).
 * UG's standard transfer serves node and edge vectors only (element and
 * side vectors: "not implemented", transgrid.cc:474-503). */
int ugmg_set_vtypes(ugmg_ctx *ctx, int level, const uint8_t *vtype);
/* AssembleDirichletBoundary (np/udm/disctools.cc:1837-1875) on a set
 * level: for every component j with VECSKIP bit j, b[v*b+j] = x[v*b+j]
 * (host arrays, list order), row j of the vector's diagonal block becomes
 * the unit row and row j of its off-diagonal blocks zero (columns are not
 * touched, so the matrix turns nonsymmetric, as in UG).  A change to the
 * matrix drops the level's layouts: ugmg_preprocess again. */
int ugmg_assemble_dirichlet(ugmg_ctx *ctx, int level, const double *x, double *b);
/* ClearDirichletValues (np/udm/disctools.cc:1801-1817): x[v*b+j] = 0 where
 * VECSKIP bit j is set; x a list-order device vector of a prepared level. */
int ugmg_clear_dirichlet(ugmg_ctx *ctx, int level, double *x);
/* VD_IDENT of the solution descriptor (np/udm/udm.h:113): ident[i] for
 * each of the n norm entries, the leader of its group (ident[ident[i]] ==
 * ident[i]).  The stop tests of ugmg_solve, ugmg_solve_surface and the
 * base solve then compare per group the sum of squares of the defects with
 * that of the limits (sc_cmp, np/udm/npscan.cc:1027-1057); n must equal
 * ugmg_norm_entries at the solve.  NULL or n <= 0: NO_IDENT (every entry on
 * its own, the default). */
int ugmg_set_ident(ugmg_ctx *ctx, const int *ident, int n);
/* VD_NCOMP of the prepared hierarchy: the entries of its norms, dot
 * products and VEC_SCALARs (vector types x b); UGMG_E_STATE before
 * ugmg_preprocess. */
int ugmg_norm_entries(const ugmg_ctx *ctx);
int ugmg_set_params(ugmg_ctx *ctx, const ugmg_params *p);
/* LmgcPreProcess (iter.cc:7707-7739): device layouts of levels bl..level,
 * colouring (GS), coarse band LU at the base level (EXPreProcess). */
int ugmg_preprocess(ugmg_ctx *ctx, int level);
/* Where ugmg_preprocess builds the level layouts (colouring, GS device
 * order, permuted levels, SELL slabs of A/L/U/P/P^T) -- the part of
 * LmgcPreProcess's flatten (np/udm/disctools.cc:2061-2128) that runs over
 * every entry: UGMG_SETUP_DEVICE (default) on the GPU from the level's CSR,
 * UGMG_SETUP_HOST on the host, then uploaded.  Both give the same layout,
 * value for value (ugmg_layout_digest).  The base level's band LU is
 * ordered on the host either way.  Environment UGMG_SETUP=host|device
 * overrides the choice. */
enum { UGMG_SETUP_HOST = 0, UGMG_SETUP_DEVICE = 1 };
int ugmg_set_setup(ugmg_ctx *ctx, int where);
/* 64-bit FNV-1a digests of every array of a prepared level's device layout
 * (slabs A, L, U, P, P^T: offsets, slot rows, columns, uniform-slice
 * offsets, offset lists, FP64 and FP32 values; diagonal blocks FP64/FP32,
 * classes, VECSKIP, colour-0 flags, device order) and one of its host-side
 * description (colours, colour slices, widths, byte model); an absent array
 * digests to 0.  Returns the number of digests (out may hold nout). */
int ugmg_layout_digest(ugmg_ctx *ctx, int level, unsigned long long *out, int nout);
int ugmg_postprocess(ugmg_ctx *ctx);         /* LmgcPostProcess            */
int ugmg_level_size(const ugmg_ctx *ctx, int level, int *n, int *nnzb);
/* components per vector of the hierarchy (ugmg_set_nlevels' b), or < 0 */
int ugmg_block(const ugmg_ctx *ctx);
int ugmg_num_colors(const ugmg_ctx *ctx, int level);
/* greedy colour of every vector of a level (host array of n ints) */
int ugmg_get_colors(ugmg_ctx *ctx, int level, int *colors);
/* decompress the device matrix layout back to CSR (host arrays sized like
 * the input); checks that the device index arrays carry UG's pattern */
int ugmg_get_pattern(ugmg_ctx *ctx, int level, int *row_ptr, int *col_idx,
                     double *vals);

/* ---- ugblas on one level, ALL_VECTORS, device pointers ----------------- */
int ugmg_dmatmul_minus(ugmg_ctx *ctx, int level, double *x, const double *y); /* x -= A y */
int ugmg_dmatmul(ugmg_ctx *ctx, int level, double *x, const double *y);       /* x  = A y */
int ugmg_smatmul_minus(ugmg_ctx *ctx, int level, float *x, const float *y);   /* FP32 copy of A */
/* dnrm2x / ddot: one value per component (out_host holds b doubles) */
int ugmg_dnrm2x(ugmg_ctx *ctx, int level, const double *x, double *out_host);
int ugmg_ddot(ugmg_ctx *ctx, int level, const double *x, const double *y,
              double *out_host);
/* ugmg_dnrm2x that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory at call time (tight
 * serving loops, harnesses that forbid hipMalloc).  The partial sums and
 * the norms live in a static device array of the library (1024 partials per
 * component, 3 components, plus the norms: 3,136 doubles, 24.5 KB per GPU,
 * held until the process ends) instead of the context's reduction buffers;
 * same kernels, same result bit for bit.  UGMG_E_ARG if a level would need
 * more partial sums than the array holds.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (other streams, other
 * contexts, other host threads) race and must be serialised by the caller. */
int ugmg_dnrm2x_noalloc(ugmg_ctx *ctx, int level, const double *x, double *out_host);

/* ---- smoothers: Smoother (iter.cc:817-842) = Step; x *= damp; b -= A x -- */
int ugmg_jacobi_step(ugmg_ctx *ctx, int level, double *x, double *b);
int ugmg_gs_step(ugmg_ctx *ctx, int level, double *x, double *b);

/* ---- transfer (transgrid.cc:117-336); level = the fine level ---------- */
int ugmg_restrict(ugmg_ctx *ctx, int level, double *to_coarse,
                  const double *from_fine, const double *damp);
int ugmg_interpolate(ugmg_ctx *ctx, int level, double *to_fine,
                     const double *from_coarse, const double *damp);

/* ---- base solver: ls($m,$red,$abslimit) over ex, on the base level ----- */
int ugmg_base_solve(ugmg_ctx *ctx, double *c, double *b);

/* ---- NP_ITER: one lmgc cycle on `level`: c := correction, b -= A c ----- */
int ugmg_iter(ugmg_ctx *ctx, int level, double *c, double *b);
int ugmg_iter_host(ugmg_ctx *ctx, int level, double *c, double *b);

/* ---- NP_LINEAR_SOLVER: ls with lmgc (ls.cc:562-749) ---------------------
 * do_defect: LinearDefect (b -= A x) before the first residuum.  b holds the
 * defect on return, x the solution.  cb may be NULL. */
int ugmg_solve(ugmg_ctx *ctx, int level, double *x, double *b,
               const double *abslimit, const double *reduction, int maxiter,
               int do_defect, ugmg_lresult *res, ugmg_defect_cb cb, void *user);
int ugmg_solve_host(ugmg_ctx *ctx, int level, double *x, double *b,
                    const double *abslimit, const double *reduction,
                    int maxiter, int do_defect, ugmg_lresult *res,
                    ugmg_defect_cb cb, void *user);

/* The last ugmg_solve on this context: V-cycles it ran (FP32 cycles in the
 * mixed form) and the algorithmic HBM bytes of every kernel it launched, from
 * the same layout-true byte model as ugmg_time_kernel (every value slot of a
 * slab, column indices of explicit slices only, each vector once per pass).
 * No counterpart in UG; for rooflines of whole solves. */
int ugmg_solve_stats(const ugmg_ctx *ctx, long long *cycles, double *bytes);

/* ---- NP_LINEAR_SOLVER on a locally refined hierarchy --------------------
 * FULLREFINELEVEL as SetSurfaceClasses (gm/algebra.cc:2885-2916) leaves it:
 * the lowest preprocessed level with a FINE_GRID_DOF vector (VCLASS >= 2,
 * VNCLASS <= 1), the top level when there is none (uniform refinement). */
int ugmg_fullrefinelevel(const ugmg_ctx *ctx);
/* the base level the last ugmg_preprocess settled on (lmgc's $b resolved:
 * min(baselevel, level), or the deepest level with at most -baselevel
 * vectors), or < 0 before it */
int ugmg_base_level(const ugmg_ctx *ctx);
/* LinearSolver (ls.cc:637-749) when the surface spans several levels:
 * x[l], b[l] for every level l in baselevel..level (n_l * b doubles, list
 * order; entries below baselevel are not read).  LinearDefect / Residuum
 * run ON_SURFACE from FULLREFINELEVEL (FINE_GRID_DOF rows below `level`,
 * NEW_DEFECT rows on it: matloop.ct / vecloop.ct), the coarse levels' b keep
 * their defect rows from iteration to iteration, and LSUpdate adds c into x
 * on baselevel..level (ls.cc:869-875).  FP64 (precision UGMG_FP64), one
 * column; UGMG_E_ARG when FULLREFINELEVEL lies below the base level.  On a
 * uniformly refined hierarchy this is ugmg_solve.  _host: host arrays. */
int ugmg_solve_surface(ugmg_ctx *ctx, int level, double *const *x, double *const *b,
                       const double *abslimit, const double *reduction, int maxiter,
                       int do_defect, ugmg_lresult *res, ugmg_defect_cb cb, void *user);
int ugmg_solve_surface_host(ugmg_ctx *ctx, int level, double *const *x, double *const *b,
                            const double *abslimit, const double *reduction, int maxiter,
                            int do_defect, ugmg_lresult *res, ugmg_defect_cb cb, void *user);

/* ---- several right-hand sides per launch --------------------------------
 * UG solves one right-hand side per call; these entry points run nrhs
 * independent columns through the same kernels (one launch per operation,
 * the column index in the grid), so the matrix of each level is read once
 * per operation for all columns.  Column j of a vector argument starts at
 * j * ld (ld >= n), device memory.  Every column follows exactly the
 * recurrence of a single-column call.
 * ugmg_set_batch: columns the level vectors are laid out for (default 1);
 *   changing it frees the device hierarchy (call ugmg_preprocess again).
 * ugmg_iter_batch: NP_ITER on each column (lmgc, iter.cc:7741-7949). */
int ugmg_set_batch(ugmg_ctx *ctx, int nrhs);
int ugmg_iter_batch(ugmg_ctx *ctx, int level, int nrhs, double *c, double *b,
                    int ld);

/* ---- NP_LINEAR_SOLVER bcgs: BCGSSolver (ls.cc:1864-2062) ----------------
 * Right-preconditioned BiCGStab with the lmgc cycle as $I (use_iter; FP32
 * cycle when the precision is UGMG_MIXED, FP64 vectors, dots and norms
 * throughout), restart every $R iterations (restart > 0), unit weights $w,
 * stop on abslimit / reduction per column or after 5 near-equal defects
 * (sc_eq 1e-4).  nrhs columns at once (see above); res has nrhs entries
 * (number_of_linear_iterations counts both half steps, as UG does).
 * defects (host, may be NULL): column j's defect sequence -- the first and
 * every LinearResiduum -- at defects[j * maxdef ...]. */
int ugmg_bcgs_solve(ugmg_ctx *ctx, int level, int nrhs, double *x, double *b,
                    int ld, const double *abslimit, const double *reduction,
                    int maxiter, int restart, int use_iter, int do_defect,
                    ugmg_lresult *res, double *defects, int maxdef);
int ugmg_bcgs_solve_host(ugmg_ctx *ctx, int level, int nrhs, double *x,
                         double *b, int ld, const double *abslimit,
                         const double *reduction, int maxiter, int restart,
                         int use_iter, int do_defect, ugmg_lresult *res,
                         double *defects, int maxdef);

/* ---- Galerkin coarse matrices: AssembleGalerkinByMatrix ------------------
 * (np/algebra/transgrid.cc:1575-1700) on the GPU.  The values of level-1
 * become P^T A P from level's matrix and its interpolation (scalar weights;
 * 3x3 blocks componentwise), in the pattern set for level-1; a product
 * whose connection is missing from that pattern is added to the diagonal
 * (UG's branch when CreateExtraConnection fails).  Replaces level-1's
 * values in the context; call ugmg_preprocess afterwards.  Applied from the
 * top down it builds a rediscretisation-free hierarchy. */
int ugmg_galerkin(ugmg_ctx *ctx, int level);

/* ---- device-side assembly of a level from element lists -----------------
 * The application's element loop with AddVlistMValues / AddVlistVValues
 * (np/udm/disctools.cc:536-555, 841-950) and AssembleDirichletBoundary
 * (disctools.cc:1837-1875) run on the GPU, in place of the host flatten
 * (ConvertMatrix, disctools.cc:2061-2128) behind ugmg_set_level: the level
 * is then set as ugmg_set_level would set it.  Every array in
 * ugmg_elements is DEVICE memory.  Vector v couples with every vector
 * sharing an element with it; a row lists its diagonal block, then the
 * coupled vectors in ascending order; every entry is 0 plus its element
 * contributions in element order (scale[e]*kref rounded before the add).
 * Element matrices are (nc b) x (nc b), row-major over the corners'
 * components (corner-major), as AddVlistMValues reads them.  Dirichlet:
 * for each VECSKIP bit j of v, row j of every block of v's row is zeroed,
 * the diagonal (j, j) set to 1 and rhs_j to dirichlet[v b + j] (0 when
 * NULL).  rhs (device, n b doubles, or NULL) receives the element vectors'
 * sums in element order.  At most 96 elements meet at a vector and 96
 * vectors share a row (UGMG_E_ARG otherwise, as for a corner index outside
 * [0, n)). */
typedef struct {
  int n;                      /* vectors of the level */
  int nelem, nc;              /* elements, corners per element */
  const int *elem;            /* nelem*nc vector indices */
  const double *emat;         /* nelem*(nc*b)^2 element matrices, or NULL */
  const double *kref;         /* (nc*b)^2 reference matrix when emat is NULL */
  const double *scale;        /* nelem factors of kref when emat is NULL */
  const double *evec;         /* nelem*nc*b element vectors, or NULL */
  const uint32_t *skip;       /* n VECSKIP masks, or NULL */
  const double *dirichlet;    /* n*b Dirichlet values, or NULL */
  const uint8_t *vclass;      /* n VCLASS, or NULL (3) */
  const uint8_t *vnclass;     /* n VNCLASS, or NULL (0) */
} ugmg_elements;
int ugmg_assemble_level(ugmg_ctx *ctx, int level, const ugmg_elements *el,
                        double *rhs);
/* The level as the context holds it (list order): with NULL arrays only
 * *n and *nnzb; else also row_ptr[n+1], col_idx[nnzb], vals[nnzb*b*b] and
 * skip[n] (host arrays; any of them may be NULL). */
int ugmg_get_level(ugmg_ctx *ctx, int level, int *n, int *nnzb, int *row_ptr,
                   int *col_idx, double *vals, uint32_t *skip);

/* ---- host-only setup routines (no GPU needed) ---------------------------
 * The pieces of PreProcess that run on the host, callable on their own.
 * ugmg_color: greedy first-fit colouring in list order over the symmetrised
 *   pattern (the multicolour GS ordering); returns the colour count or < 0.
 * ugmg_coarse_inverse: EXPreProcess (iter.cc:8685-8970) -- optional BFS
 *   band reorder, band LU without pivoting (evm.cc:1400-1418) -- followed by
 *   LU solves of the unit vectors: inv (n*n, row-major, list order), order
 *   (new list order), bandwidth.  UGMG_E_SINGULAR on a zero pivot. */
int ugmg_color(int n, const int *row_ptr, const int *col_idx, int *colors);
/* ugmg_color on the device of ctx (host arrays in and out): the same
 * colouring, in rounds over the vectors whose lower neighbours are all
 * coloured; *rounds (may be NULL) receives the number of rounds.  Returns
 * the colour count or < 0. */
int ugmg_color_device(ugmg_ctx *ctx, int n, const int *row_ptr, const int *col_idx,
                      int *colors, int *rounds);
int ugmg_coarse_inverse(int n, const int *row_ptr, const int *col_idx,
                        const double *vals, int reorder, double *inv,
                        int *order, int *bandwidth);

/* ---- serving: continuous batching of independent solves ----------------
 * No counterpart in UG (its ugd daemon is a remote UI, ugd.cc:528-608).
 * Each request is a whole hierarchy + rhs (scalar, FP64 solves, one
 * smoother class and $damp before and after the coarse correction, the ex
 * base solve).  Requests
 * with the same hierarchy (patterns, transfers, classes, Dirichlet bits)
 * and parameters share a resident group: one device hierarchy holding up
 * to max_slots matrices of that pattern, one per column of its vectors,
 * whose layout worker threads build once.  ugmg_server_step admits queued
 * requests in submission order into free slots of idle groups (a request
 * whose group is still being laid out holds back none) while the total of
 * admitted requests stays within max_slots; a group is set up on the device
 * when the HBM budget allows, releasing idle groups least recently used
 * first.  Each step launches the next ls iteration of every idle group as
 * one batched launch sequence over its occupied columns (a captured graph
 * per column count), so concurrent solves of one shape share every kernel
 * launch and groups overlap through their streams.  Results equal those of
 * a stand-alone ugmg_solve of the request, bit for bit. */
typedef struct {
  int n;
  const int *row_ptr, *col_idx;
  const double *vals;
  const uint32_t *skip;
  const uint8_t *vclass, *vnclass;
  const int *p_ptr, *p_idx;        /* interpolation from level-1 (level >= 1) */
  const double *p_w;
} ugmg_level_desc;

enum { UGMG_REQ_QUEUED = 0, UGMG_REQ_ACTIVE = 1, UGMG_REQ_DONE = 2, UGMG_REQ_FAILED = 3 };

typedef struct ugmg_server ugmg_server;
int ugmg_server_create(int device, double hbm_budget_bytes, int max_slots,
                       ugmg_server **out);
int ugmg_server_destroy(ugmg_server *s);
const char *ugmg_server_last_error(const ugmg_server *s);
int ugmg_server_submit(ugmg_server *s, int nlevels, const ugmg_level_desc *levels,
                       const ugmg_params *p, const double *rhs, double red,
                       double abslimit, int maxiter, int *id);
int ugmg_server_step(ugmg_server *s, int *nactive, int *nqueued, int *ndone);
/* resident groups (hierarchies on the device), group iterations launched
 * (each one batched launch sequence over its slots) and the solve
 * iterations they carried: columns / launches = solves per launch */
int ugmg_server_stats(const ugmg_server *s, int *groups, long long *launches,
                      long long *columns);
/* times[3]: submit, admission, completion in seconds since creation */
int ugmg_server_status(ugmg_server *s, int id, int *state, ugmg_lresult *res,
                       double *times);
int ugmg_server_result(ugmg_server *s, int id, double *x, int n);

/* ---- measurement --------------------------------------------------------
 * kind: 0 dmatmul_minus FP64, 1 smatmul_minus FP32, 2 jacobi step FP64,
 *       3 gs step FP64, 4 restrict, 5 interpolate, 6 one cycle (current
 *       precision), 7 gs step FP32, 8 gs step FP64 as a cycle runs the
 *       steps after the first (8 colour sweeps with the previous step's
 *       defect folded in; scalar levels whose rows are all active).  Runs
 *       `reps` back-to-back launches on the context stream between two
 *       hipEvents and returns the mean duration in ms, plus the algorithmic
 *       bytes of one launch as the data are laid out: every value slot of
 *       the slabs (padding and the zero phantoms of uniform slices
 *       included), column indices of explicit slices only, the slot-to-row
 *       maps, and each vector once per pass (twice when read and written). */
int ugmg_time_kernel(ugmg_ctx *ctx, int level, int kind, int reps,
                     double *ms_per_launch, double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
