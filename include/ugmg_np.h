/*
 * ugmg_np.h -- numproc layer over libugmg: UG's npcreate / npinit /
 * npexecute for the solve path, in C++ behind a C ABI.
 *
 * Mirrors the numproc interface the GPU path replaces (np/udm/numproc.cc,
 * np/procs/ls.cc, np/procs/iter.cc): objects are created by class name,
 * configured with UG's option strings ("$m 100 $red 1e-10 $I mgc ..."),
 * and executed with UG's flags ("$i $d $r $s $p").  The classes and the
 * options they read (names, defaults, and the status npinit returns) follow
 * UG's Init functions:
 *   ls        LinearSolverInit + NPLinearSolverInit   ls.cc:771-797, 382-407
 *             $m (required), $I, $red (required), $abslimit (ABS_LIMIT 1e-10),
 *             $x $b $A, $display no|red|full (ReadArgvDisplay, default no)
 *   bcgs      BCGSInit                                ls.cc:1750-1800
 *             $m (required), $R (0; < 0 not active), $I, $red, $abslimit,
 *             $weight (1; only unit weights are supported), $display
 *   lmgc      LmgcInit                                iter.cc:7613-7656
 *             $S pre post base, $T transfer, $g 1, $n1 1, $n2 1, $b 0
 *             (negative: deepest level with at most -b vectors), $damp 1
 *   jac, gs   SmootherInit                            iter.cc:763-790
 *             $damp 1   (gs is the multicolour-ordered l_lgs of the GPU)
 *   ex        EXInit                                  iter.cc:8248-8260
 *             $o 1 (BFS band reorder), $f (FLOAT band LU), $damp 1
 *   transfer  the transfer the hierarchy was given: the standard one
 *             (ugmg_set_transfer), or $M's interpolation matrices
 *             (ugmg_set_transfer_matrix); no options read here
 * $damp, $red and $abslimit are VEC_SCALARs (sc_read, npscan.cc:1211-1278):
 * one number for every entry, or per vector type a type letter (n node,
 * k edge) and the values of its components ("$damp n0.6:0.7:0.8",
 * "$red n1e-8|k1e-6"); the damping factors of node and edge vectors must
 * agree per component, and the base ls takes one $red and $abslimit.
 * GPU-only options of lmgc: $precision fp64|mixed, $inner k (FP32 cycles
 * per FP64 step), $adapt 0|1 (a step runs fewer cycles when the rate
 * predicts the stop test met; default 1), $graph 0|1; of ls/bcgs: $nrhs k
 * (bcgs columns).
 *
 * The pre- and post-smoother of one lmgc may be different jac / gs objects
 * with their own $damp (ugmg_params.post_set).  Differences a script sees:
 * the smoothers must be jac or gs; the base solver must be an ls over ex
 * or jac (ugmg_params.base_iter);
 * vectors are host arrays
 * registered by name.  npinit returns UGMG_NP_NOT_ACTIVE for a
 * configuration the GPU path cannot run, with the reason in
 * ugmg_np_last_error.
 */
#ifndef UGMG_NP_H
#define UGMG_NP_H

#include "ugmg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* numproc states (np/udm/numproc.h:54-57) */
enum { UGMG_NP_NOT_INIT = 0, UGMG_NP_NOT_ACTIVE = 1, UGMG_NP_ACTIVE = 2, UGMG_NP_EXECUTABLE = 3 };

typedef struct ugmg_np_env ugmg_np_env;

/* An environment of numprocs over one context whose hierarchy has been set
 * (ugmg_set_nlevels / set_level / set_transfer).  ctx may be NULL to
 * create and configure objects without a GPU (npexecute then fails). */
int ugmg_np_env_create(ugmg_ctx *ctx, ugmg_np_env **out);
int ugmg_np_env_destroy(ugmg_np_env *env);
const char *ugmg_np_last_error(const ugmg_np_env *env);
/* CURRENTLEVEL: the level npexecute works on (default: the finest set) */
int ugmg_np_set_level(ugmg_np_env *env, int level);
/* a named vector ($x sol, $b rhs): caller-owned host array of n*b doubles
 * per column, `ncols` columns at stride n*b */
int ugmg_np_set_vector(ugmg_np_env *env, const char *name, double *data, int ncols);

/* the named vector's array (n_l*b doubles) on a coarser level l: on a
 * locally refined hierarchy (ugmg_fullrefinelevel below the current level)
 * ls works ON_SURFACE on x and b of every level from its base level up
 * (ugmg_solve_surface), as UG's LinearSolver does on the levels' VECTOR
 * lists; the finest level's array is the one ugmg_np_set_vector registers */
int ugmg_np_set_level_vector(ugmg_np_env *env, const char *name, int level, double *data);
/* npcreate <name> $c <class>: 0 or a negative code */
int ugmg_np_create(ugmg_np_env *env, const char *name, const char *cls);
/* npinit <name> <options>: returns the numproc state (UGMG_NP_*) or < 0 */
int ugmg_np_init(ugmg_np_env *env, const char *name, const char *options);
/* npexecute <name> <flags>: 0 or a negative code */
int ugmg_np_execute(ugmg_np_env *env, const char *name, const char *flags);
/* The text of the linear solvers' displays since the last call: the head
 * line, DoPCR's per-iteration defect and rate lines, the averages and the
 * "LS  :" / "BCGS:" line, in UG's formats (np/udm/pcr.cc:263-606,
 * ls.cc:666-744 and 1896-2060; TSOLVE/TIT in seconds of wall clock).  A UG
 * plugin hands it to UserWrite.  Valid until the next call on env. */
const char *ugmg_np_output(ugmg_np_env *env);

/* the LRESULT of the last Residuum / Solver of a linear solver (column j) */
int ugmg_np_result(const ugmg_np_env *env, const char *name, int col, ugmg_lresult *res);
/* the parameters a linear solver object hands to libugmg (for inspection) */
int ugmg_np_params(const ugmg_np_env *env, const char *name, ugmg_params *p);

#ifdef __cplusplus
}
#endif
#endif
