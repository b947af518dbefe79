/*
 * ugmg_problems.h -- synthetic multigrid problems in UG's per-level layout.
 *
 * UG ships no discretisation: an application assembles one stiffness
 * matrix per grid level into the MATRIX lists of the multigrid
 * (gm/gm.h:601-674), applies Dirichlet rows with AssembleDirichletBoundary
 * (np/udm/disctools.cc:1837-1875) and lets the standard transfer
 * interpolate with father-element shape functions (np/algebra/transgrid.cc:
 * 117-336, gm/shapes.cc:748-819).  This library is that application for
 * the solve-path workloads: uniformly refined Q1 grids on the unit square
 * and cube, assembled per level, flattened in ConvertMatrix order
 * (np/udm/disctools.cc:2061-2128: vector list order, diagonal block first
 * in each row).
 *
 * It is host-only C++ and has no GPU dependency.  All arrays returned are
 * owned by the problem object and live until ugmg_problem_free().
 */
#ifndef UGMG_PROBLEMS_H
#define UGMG_PROBLEMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ugmg_problem ugmg_problem;

/* kappa_mode */
#define UGMG_KAPPA_ONE   0   /* -Laplace u = f                                  */
#define UGMG_KAPPA_JUMP  1   /* kappa per finest element = 10^U[-2,2]            */

/* Uniformly refined Q1 hierarchy.
 *   dim          2 (bilinear quads) or 3 (trilinear hexes)
 *   coarse_cells cells per direction on level 0 (2 for the 2x2 quad grid,
 *                1 for the single hex)
 *   nref         number of uniform refinements; levels 0..nref
 *   kappa_mode   UGMG_KAPPA_ONE or UGMG_KAPPA_JUMP
 *   seed         SplitMix64 seed: kappa of every finest element (x fastest),
 *                then the nodal f of every finest node, both drawn in order
 * Coarse-level matrices integrate the finest-level coefficient exactly
 * (element-local Galerkin products), so every level is the FE stiffness
 * matrix of the same bilinear form.  Zero Dirichlet data on the whole
 * boundary; rhs = consistent Q1 mass matrix times nodal f, zero on
 * Dirichlet nodes.  Returns 0 or a negative error code. */
int ugmg_problem_q1(int dim, int coarse_cells, int nref, int kappa_mode,
                    uint64_t seed, ugmg_problem **out);

/* 2D convection-diffusion -eps Laplace u + b.grad u = f, b = (cos t, sin t),
 * streamline-upwind Petrov-Galerkin P1 triangles: 2x2 coarse squares split
 * by the (0,0)-(1,1) diagonal (8 triangles) refined nref times (level l has
 * (2^(l+1)+1)^2 nodes), every level rediscretised, zero Dirichlet boundary.
 * theta = NaN draws t ~ U[0, 2 pi) first from the SplitMix64 stream; then f
 * nodal U[-1,1] on the finest level.  Nonsymmetric. */
int ugmg_problem_p1_cd(int nref, double eps, double theta, uint64_t seed,
                       ugmg_problem **out);
double ugmg_problem_theta(const ugmg_problem *p);

/* 3D linear elasticity -div(2 mu eps(u) + lambda div u I) = f, P1 on the
 * Kuhn split of the unit cube (6 tetrahedra) refined nref times by red
 * refinement (level l: (2^l+1)^3 nodes, 15-block stencil), 3x3 blocks
 * (b = 3, displacement components x, y, z), E and nu given, clamped
 * (all components Dirichlet) on the face x = 0, traction-free elsewhere,
 * every level rediscretised.  Body force f ~ U[-1,1]^3 per finest node
 * from the SplitMix64 stream (node order, components fastest); rhs =
 * consistent P1 mass times f. */
int ugmg_problem_elasticity(int nref, double E, double nu, uint64_t seed,
                            ugmg_problem **out);

/* 3D P1 linear elasticity (as above: b = 3, E, nu, clamped on x = 0) on a
 * locally refined tetrahedral hierarchy: the 6 Kuhn tetrahedra of the unit
 * cube refined red nuni times, then depth times only toward the corner
 * (0,0,0) (step k marks the regular elements inside [0, 2^-k]^3), closed
 * green as UG closes tetrahedra without a complete rule set (a centre node
 * and one son per side triangle), every other element copied.  Vector
 * classes, list order and interpolation as UG's refinement and
 * SetSurfaceClasses give them; every level's own matrix.  No load: the
 * right-hand sides are per level (the surface solve's).  nuni + depth <= 9. */
int ugmg_problem_elasticity_local(int nuni, int depth, double E, double nu,
                                  ugmg_problem **out);

/* The tetrahedral mesh of a level of ugmg_problem_elasticity_local: node
 * coordinates (3 per vector, list order), 4 corners per element in UG's
 * corner order and the element kind (0 red, 1 green, 2 copy).  -1 for
 * the other generators. */
int ugmg_problem_mesh(const ugmg_problem *p, int level, int *nnodes, const double **xyz,
                      int *nelem, const int **tets, const uint8_t **kind);

void ugmg_problem_free(ugmg_problem *p);

int ugmg_problem_nlevels(const ugmg_problem *p);     /* nref + 1            */
int ugmg_problem_block(const ugmg_problem *p);       /* components per node */

/* Level l matrix in ConvertMatrix order: n block rows, row_ptr[n+1],
 * col_idx[nnzb], vals[nnzb*b*b] (row-major b x b blocks); per vector the
 * VECSKIP bitmask, VCLASS and VNCLASS (gm/gm.h:2200-2208, 2282-2288). */
int ugmg_problem_level(const ugmg_problem *p, int level, int *n, int *nnzb,
                       const int **row_ptr, const int **col_idx,
                       const double **vals, const uint32_t **skip,
                       const uint8_t **vclass, const uint8_t **vnclass);

/* Interpolation from level-1 to level (level >= 1) as CSR over fine
 * vectors: row v lists the coarse corners of v's father element with their
 * nonzero GNs weights, in element corner order (a corner node has the one
 * entry (father, 1.0)). */
int ugmg_problem_transfer(const ugmg_problem *p, int level, int *nnz,
                          const int **p_ptr, const int **p_idx,
                          const double **p_w);

/* Finest-level right-hand side (n*b doubles) and grid geometry. */
int ugmg_problem_rhs(const ugmg_problem *p, const double **b);
int ugmg_problem_grid(const ugmg_problem *p, int level, int *dim,
                      int *nodes_per_dir);

/* Element data of the finest level of the generators above, as an element
 * loop hands it to AddVlistMValues (np/udm/disctools.cc:841): nelem
 * elements of nc corners, elem[e*nc + i] = vector of corner i; element
 * matrix (nc b) x (nc b), row-major over corner-major scalar unknowns, either
 * emat[e] (full) or scale[e] * kref (one reference matrix; emat NULL);
 * skip = the VECSKIP masks of the level.  Same arithmetic and SplitMix64
 * draws as the generators, so assembling in element order reproduces their
 * finest level exactly.  Free with ugmg_element_data_free. */
typedef struct ugmg_element_data {
  int n, b, nelem, nc;
  int *elem;
  double *emat;           /* nelem (nc b)^2, or NULL */
  double *scale, *kref;   /* nelem and (nc b)^2, or NULL */
  uint32_t *skip;         /* n */
} ugmg_element_data;
int ugmg_problem_q1_elements(int dim, int coarse_cells, int nref, int kappa_mode,
                             uint64_t seed, ugmg_element_data **out);
int ugmg_problem_p1_cd_elements(int nref, double eps, double theta, uint64_t seed,
                                ugmg_element_data **out);
int ugmg_problem_elasticity_elements(int nref, double E, double nu, uint64_t seed,
                                     ugmg_element_data **out);
void ugmg_element_data_free(ugmg_element_data *d);

/* SplitMix64 stream shared with the tests: fills out[i] = a+(b-a)*u_i. */
void ugmg_splitmix_uniform(uint64_t seed, int64_t count, double a, double b,
                           double *out);

#ifdef __cplusplus
}
#endif
#endif
