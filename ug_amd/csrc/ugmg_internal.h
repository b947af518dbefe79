// Internal data structures of libugmg (host setup + device state).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "ugmg_common.h"

namespace ugmg {

// Host copy of one level as it crosses the C ABI (ConvertMatrix layout).
struct HostLevel {
  int n = 0, nnzb = 0;
  int b = 1;   // components per vector; a holds b*b values per entry
  std::vector<int> rp;
  hvec<int> ci;
  hvec<double> a;
  std::vector<uint32_t> skip;
  std::vector<uint8_t> vc, vnc;
  bool has_p = false;
  std::vector<int> pp, pi;
  std::vector<double> pw;
  // vector types (VTYPE: 0 node, 1 edge; empty: all node vectors)
  std::vector<uint8_t> vt;
  // the interpolation weights the layouts use when they differ from pw
  // (edge rows: 0 for a father object whose components are all Dirichlet)
  std::vector<double> pwf;
  const std::vector<double> &p_weights() const { return pwf.empty() ? pw : pwf; }
  // an interpolation matrix (RestrictByMatrix / InterpolateCorrectionByMatrix):
  // a b x b block per entry of pp/pi, M[i*b+j] = coarse component i, fine
  // component j (MVALUE of the IMAT entry); empty: the scalar weights pw
  std::vector<double> pblk;
  // P^T takes node rows of VCLASS >= 2 and every edge row of the standard
  // transfer (transgrid.cc:153, 188-212); with a matrix, every row of VCLASS
  // >= 2 (transgrid.cc:1180)
  bool restricts(int v) const { return vc[v] >= 2 || (pblk.empty() && !vt.empty() && vt[v] == 1); }
};

// SELL-64 layout of a set of rows, built on the host.
//   slice s covers slots [64 s, 64 s + 64); slot -> row via rows[] (-1 pad)
//   entry (s, k, lane) at soff[s] + 64 k + lane, k < width(s)
struct HostSell {
  int nrows = 0;         // rows covered (slots that are not padding)
  int nslices = 0;
  std::vector<int> soff;  // nslices + 1, entry offsets (multiples of 64)
  std::vector<int> rows;  // nslices * 64
  hvec<int> col;
  // bb planes of soff[nslices] values: component q of entry e at q*nent + e
  hvec<double> val;
  int bb = 1;
  bool identity_rows = false;
  // uniform-delta slices: every row of the slice has the column offsets
  // dtab[sdoff[s] + k] (k < width), entries missing in a row are stored as
  // value 0 ("phantoms", col = -1 in `col`); sdoff[s] = -1 for an explicit
  // slice, whose columns are read from `col`.
  std::vector<int> sdoff;
  std::vector<int> dtab;
};

// Host-side products of LmgcPreProcess for one level.
struct HostPrep {
  // device order of the level's vectors: perm[p] = list index of device
  // position p (empty: list order); color in device order
  std::vector<int> perm;
  int window = 0;
  std::vector<int> color_ug;   // colouring in list order
  HostSell A;                // all rows, list order
  std::vector<double> diag;  // diagonal block (b*b, row-major) of every row
  // multicolour Gauss-Seidel
  int ncolors = 0;
  std::vector<int> color;          // per vector
  std::vector<int> color_slices;   // ncolors + 1, slice ranges in L
  HostSell L;                      // rows grouped by colour, lower-colour entries
  HostSell U;                      // same rows, higher-colour entries (all for inactive rows)
  // restriction as a gather: coarse rows of P^T from fine rows with VCLASS >= 2
  std::vector<int> pt_ptr, pt_idx;
  std::vector<double> pt_w;
  HostSell P;    // interpolation rows (fine vectors)
  HostSell PT;   // restriction rows (coarse vectors)
};

// Base level: band LU as in EXPreProcess, then the explicit inverse.
struct HostBase {
  int n = 0;                   // scalar unknowns (vectors * b)
  int bw = 0;
  bool failed = false;
  std::vector<int> order;      // BFS list order
  std::vector<int> pos;        // scalar unknown (list order) -> band position
  std::vector<double> band;    // EX_MAT band, n (2 bw + 1) + 1 (before factoring)
  bool fmode = false;          // ex $f: FLOAT entries, LU in FLOAT arithmetic
  std::vector<double> inv;     // n x n row-major, list order (vector-major);
                               // empty until factor_base_host (or factored
                               // on the device)
};

// filters for build_sell
enum class Part { All, Lower, Upper };

// ---- the same products built on the device (ugmg_devsetup.hip) ----------
// A SELL-64 slab in device memory (the arrays of HostSell; values in bb
// planes of nent) with host copies of its per-slice offsets.
struct DevSellParts {
  int nrows = 0, nslices = 0, bb = 1;
  size_t nent = 0, ndtab = 0;
  int *soff = nullptr, *rows = nullptr, *col = nullptr, *sdoff = nullptr, *dtab = nullptr;
  double *val = nullptr;
  std::vector<int> soff_h, sdoff_h;
};
// One level laid out on the device: what build_prep (after greedy_color,
// choose_gs_window, gs_device_order and permute_level on a GS level) gives
// on the host, value for value.
struct DevPrep {
  int n = 0, nnzb = 0, window = 0, ncolors = 0;
  std::vector<int> perm_h;            // GS levels: device position -> list index
  std::vector<int> color_ug_h;        // GS levels: colouring in list order
  std::vector<int> color_slices;      // ncolors + 1
  std::vector<int> color_rows;        // vectors per colour
  int *perm = nullptr;                // device copy of perm_h
  int *pos = nullptr;                 // list index -> device position (caller frees)
  DevSellParts A, L, U, P, PT;
  double *diag = nullptr;             // n b*b
  uint8_t *vc = nullptr, *vnc = nullptr, *c0 = nullptr;   // device order
  uint8_t *vt = nullptr;                                   // device order (typed levels)
  uint32_t *skip = nullptr;
  int p_nnz = 0, pt_nnz = 0;
  bool all_active = false;
};
// Final device arrays are allocated through `keep` (which registers them
// with their owner); returns 0 or a UGMG error code.
struct DevKeep {
  void *user = nullptr;
  int (*alloc)(void *user, void **p, size_t bytes) = nullptr;
};
// Lays out level H on the device: coarse_n vectors on level l-1 (0: none),
// coarse_pos its list index -> device position map (device; null: list
// order), gs: a multicolour GS level, force_win >= 0: the device-order
// window instead of the chosen one.  Returns 0 or a UGMG error code (err
// says why).
// drp/dci/da (optional): the level's CSR already on the device (equal to
// H's), used instead of an upload.
int build_level_device(void *stream, const HostLevel &H, int coarse_n, const int *coarse_pos,
                       bool gs, int force_win, const DevKeep &keep, DevPrep &out,
                       std::string &err, const int *drp = nullptr, const int *dci = nullptr,
                       const double *da = nullptr);
// greedy_color on the device (the same colouring): colors (device, n ints)
// receives it, *rounds the frontier rounds it took; returns the colour count
// or < 0.
int color_device(void *stream, int n, const int *rp, const int *ci, int *colors, int *rounds,
                 std::string &err);

void build_sell(const HostLevel &L, const std::vector<int> *rows, Part part,
                const std::vector<int> *color, HostSell &out);
void build_sell_csr(int n, const int *rp, const int *ci, const double *a, HostSell &out, int bb = 1);
int greedy_color(const HostLevel &L, std::vector<int> &color);
void to_float(const double *in, size_t n, hvec<float> &out);
// L (and its transfer) in device order: row p of out is row perm[p] of L,
// columns through pos; interpolation columns through cpos (coarse list
// index -> coarse device position; null: identity)
void permute_level(const HostLevel &L, const std::vector<int> &perm, const std::vector<int> &pos,
                   const std::vector<int> *cpos, HostLevel &out);
// device order of a multicolour-GS level: list order stably regrouped by
// colour inside windows of `window` consecutive vectors (one grid line of a
// lexicographic lattice), so each colour's vectors are contiguous runs and
// the column offsets of a stencil stay uniform along a run.
void gs_device_order(const std::vector<int> &color, int ncolors, int window,
                     std::vector<int> &perm);
// window for gs_device_order: the candidate (grid-line lengths guessed from
// the stencil offsets, or the whole level) whose colour runs give the most
// uniform-offset slices on a sample
int choose_gs_window(const HostLevel &L, const std::vector<int> &color, int ncolors);
// color: the level's colouring in the row order of L (null: Jacobi level)
void build_prep(const HostLevel &L, const HostLevel *coarse, const std::vector<int> *color,
                int ncolors, HostPrep &out);
// EXPreProcess's order (BFS unless reorder is 0), bandwidth and band
void build_base_band(const HostLevel &L, int reorder, HostBase &out, bool fmode = false);
// band LU without pivoting and the inverse on the host (frees the band)
void factor_base_host(HostBase &out);
void build_base(const HostLevel &L, int reorder, HostBase &out, bool fmode = false);

}  // namespace ugmg
