// C ABI of libugmg: context, device hierarchy, lmgc / ls drivers.
//
// The drivers issue the same operation sequence as UG's numprocs:
//   Lmgc            np/procs/iter.cc:7741-7949
//   Smoother        np/procs/iter.cc:817-842 (JacobiStep 917, GSStep 1039)
//   LinearSolver    np/procs/ls.cc:637-749 (+ LinearDefect 562, Residuum 577)
//   base ls over ex np/procs/ls.cc:637-749, np/procs/iter.cc:8972-9071
// with three liberties that change no reported number:
//   * the top-level correction c of a linear-solver iteration is accumulated
//     straight into x (dset(c,0) + Lmgc + dadd(x,c) == Lmgc into x, up to
//     the rounding of the additions; the defect recurrence never reads x);
//   * the coarse-level dset(c,0) before the recursion is folded into the
//     first kernel that writes c on that level (it stores instead of adds);
//   * the ls stop test runs on the host after a one-double device-to-host
//     copy per iteration, as UG's LinearSolver does after dnrm2x.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ugmg.h"
#include "ugmg_internal.h"
#include "ugmg_kernels.h"

// Scratch of ugmg_dnrm2x_noalloc, the entry point that allocates nothing:
// the reduction's partial sums (at most 1024 blocks per component, the cap
// of reduce_parts) followed by the final norms.  Every GPU the library runs
// on holds its own copy until the process ends.  Host code never uses the
// array's name: its device address is resolved once per device (when the
// first context on that device is created) and only that address is passed
// to kernels and copies.
constexpr int kScratchParts = 1024 * UGMG_MAXB;
constexpr int kScratchDoubles = kScratchParts + 64;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kScratchMaxDevices] = {};
double *g_scratch_ptr[kScratchMaxDevices] = {};
hipError_t g_scratch_err[kScratchMaxDevices] = {};

// with `device` current: the device address of this GPU's scratch (once)
void resolve_scratch(int device) {
  if (device < 0 || device >= kScratchMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? static_cast<double *>(ptr) : nullptr;
  g_scratch_done[device] = true;
}

}  // namespace

using namespace ugmg;

namespace {

struct DevSellBuf {
  int nslices = 0;
  size_t nent = 0;
  int nuniform = 0;
  size_t nent_explicit = 0;     // entries of slices that read column indices
  size_t ndtab = 0;             // shared offset-list length
  int bb = 1;                   // value planes (b*b)
  int *soff = nullptr, *rows = nullptr, *col = nullptr, *sdoff = nullptr, *dtab = nullptr;
  double *v64 = nullptr;
  float *v32 = nullptr;
  int wmax = 0;                 // widest slice
  std::vector<int> width_h;     // per-slice width (host)
  // prefix sums over the slices (host): value slots, and the slots of
  // explicit slices (those that read column indices)
  std::vector<double> slot_pre, expl_pre;
  // Bytes one pass over slices [s0, s1) reads as laid out: every value slot
  // (padding lanes and zero phantoms of uniform slices included, sv bytes
  // per value), column indices only of explicit slices, the slot->row map
  // and soff/sdoff per slice.  Vectors not included.  O(1): the solver adds
  // these up after every iteration.
  double bytes_range(int s0, int s1, int sv) const {
    if (s1 <= s0) return 0.0;
    const double e = slot_pre[s1] - slot_pre[s0], ex = expl_pre[s1] - expl_pre[s0];
    return e * bb * sv + 4.0 * ex + (4.0 * 64.0 + 8.0) * (s1 - s0);
  }
  // the whole slab, with its shared offset lists
  double bytes(int sv) const { return bytes_range(0, nslices, sv) + 4.0 * (double)ndtab; }
  int max_width(int s0, int s1) const {
    int m = 0;
    for (int s = s0; s < s1; ++s) m = std::max(m, width_h[s]);
    return m;
  }
  // members: values of member m at v64/v32 + m * mst; the level's diagonal
  // blocks of member m at diag + m * dms (0: one matrix)
  size_t mst = 0, dms = 0;
  size_t nval = 0;              // values of one member (bb planes of nent)
  // colour map of a colour-grouped slab (DSell::ncb, cb)
  int ncb = 0;
  int cb[kMaxColors + 1] = {};
  template <typename TV>
  DSell<TV> view(const TV *v) const {
    DSell<TV> d{nslices, soff, rows, col, v, sdoff, dtab, wmax, nent, mst, dms};
    d.ncb = ncb;
    d.bb = bb;
    for (int k = 0; k <= ncb && k <= kMaxColors; ++k) d.cb[k] = cb[k];
    return d;
  }
  DSell<double> d64() const { return view<double>(v64); }
  DSell<float> d32() const { return view<float>(v32); }
  template <typename T>
  DSell<T> d() const {
    if constexpr (std::is_same_v<T, double>) return d64();
    else return d32();
  }
};

struct DevLevel {
  int n = 0, nnzb = 0;
  DevSellBuf A, L, U;
  int ncolors = 0;
  std::vector<int> color_slices, color, color_wmax, color_wmax_u;
  std::vector<int> color_rows;   // vectors per colour
  bool all_active = false;   // every vector has VCLASS >= 3
  double *diag64 = nullptr;
  float *diag32 = nullptr;
  uint8_t *vclass = nullptr, *vnclass = nullptr;
  uint8_t *c0 = nullptr;   // GS levels: 1 on the rows of colour 0 (device order)
  uint32_t *skip = nullptr;
  // several vector types: VTYPE per row, and the VECSKIP bits the
  // interpolation reads (an edge row's own bits are not read: 0); a single
  // type: vt null, iskip = skip
  uint8_t *vt = nullptr;
  uint32_t *iskip = nullptr;
  int ntypes = 1;
  int nc() const { return vt ? ntypes * nb : nb; }   // norm entries
  int p_nnz = 0, pt_nnz = 0;
  DevSellBuf P, PT;   // interpolation from / restriction to level-1
  bool nt = false;    // stream the matrix with non-temporal loads
  bool nt32 = false;  // the same for the FP32 copy
  template <typename T> bool ntv() const { if constexpr (std::is_same_v<T, double>) return nt; else return nt32; }
  int nb = 1;         // components per vector (block size)
  size_t vs = 0;      // column stride of the level vectors (scalars; batched launches)
  // device order (GS levels): perm_h[p] = list index of device position p;
  // empty / null: list order
  std::vector<int> perm_h;
  int *perm = nullptr;
  double *c = nullptr, *b = nullptr, *t = nullptr;
  float *c32 = nullptr, *b32 = nullptr, *t32 = nullptr;
  // the post-smoother's iterate when its colour-0 sweep runs inside the
  // SpMV after the interpolation (which still reads t)
  double *u = nullptr;
  float *u32 = nullptr;
  // surface solve (allocated on first use): FINE_GRID_DOF class bytes, the
  // level's x in device order, its b-component squared sums
  uint8_t *fgd = nullptr;
  double *xs = nullptr;
  float *acc32 = nullptr;   // mixed surface solve: a step's summed FP32 corrections
  // members: for every value slot of the A, L and U slabs and of the
  // diagonal blocks, the index of its value in the level's list-order CSR
  // values (-1: padding or phantom); diag member stride
  int *map_a = nullptr, *map_l = nullptr, *map_u = nullptr, *map_d = nullptr;
  size_t dms = 0;

  template <typename T> T *cv() const { if constexpr (std::is_same_v<T, double>) return c; else return c32; }
  template <typename T> T *bv() const { if constexpr (std::is_same_v<T, double>) return b; else return b32; }
  template <typename T> T *tv() const { if constexpr (std::is_same_v<T, double>) return t; else return t32; }
  template <typename T> T *uv() const { if constexpr (std::is_same_v<T, double>) return u; else return u32; }
  template <typename T> const T *diag() const { if constexpr (std::is_same_v<T, double>) return diag64; else return diag32; }
};

}  // namespace

struct ugmg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  int nlevels = 0, b = 1;
  // vector types of the prepared hierarchy and the entries of its norms
  // and VEC_SCALARs (ntypes * b)
  int ntypes = 1, nc = 1;
  // VD_IDENT of the solution descriptor (sc_cmp's grouping), nident entries
  // (0: NO_IDENT)
  int nident = 0;
  int ident[UGMG_MAXB] = {};
  std::vector<HostLevel> host;
  std::vector<char> level_set, transfer_set;
  ugmg_params p{};
  bool params_set = false;
  bool prepared = false;
  int top = -1, bl = 0;
  int frl = -1;   // FULLREFINELEVEL of the levels set
  std::vector<DevLevel> dev;
  std::vector<void *> allocs;
  double dev_bytes = 0.0;
  HostBase base;
  double *base_inv = nullptr;
  // the inverse's rows that are not unit rows, FP64 and FP32 (the FP32
  // cycle's base solve), and each row's index among them (-1: unit row)
  double *base_cinv = nullptr;
  float *base_cinv32 = nullptr;
  int *base_rowmap = nullptr;
  // column order of the compact inverse: the unknowns whose defect can be
  // nonzero after the restriction first (base_ncyc of them), then the
  // Dirichlet unknowns the restriction sets to zero (VECSKIP component of a
  // unit row with VNCLASS >= 2), whose columns a cycle's base solve skips
  int *base_colperm = nullptr;
  int base_ncyc = 0;
  int base_nact = 0;   // rows of the compact inverse
  // the last ugmg_solve: V-cycles launched (FP32 cycles in the mixed form)
  // and the algorithmic bytes of every kernel it launched
  long long stat_cycles = 0;
  double stat_bytes = 0.0;
  // position of each base unknown among the first base_ncyc columns (-1:
  // none), for the restriction that writes b in column order
  int *base_cpos = nullptr;
  double *part = nullptr;
  int part_cap = 0;
  double *red = nullptr;      // device scalars
  double *red_host = nullptr; // pinned
  // cached graph of one solver iteration
  // (mixed: one graph per number of FP32 cycles, index = cycles)
  static constexpr int kMaxGraphs = 65;
  hipGraphExec_t gexec[kMaxGraphs] = {};
  int gkind = -1, glevel = -1;
  const void *gx = nullptr, *gb = nullptr;
  // members: `members` matrices of one pattern, one per column of the
  // level vectors (batched launches read column j's values, diagonal
  // blocks and base inverse); set through member_set
  // levels flattened on the device by ugmg_assemble_level: their CSR kept
  // there for the next device layout build (freed by it, or when the level
  // is set again)
  struct DevCsr {
    int *rp = nullptr, *ci = nullptr;
    double *a = nullptr;
  };
  std::vector<DevCsr> dcsr;
  int members = 1;
  // a serving group's context: every level carries member planes and the
  // index maps that fill them (also with one member)
  bool grouped = false;
  // level layouts built on the device (ugmg_set_setup; default) or on the
  // host
  bool setup_device = true;
  size_t base_ims = 0;          // values of one member's compact base inverse
  int *base_act = nullptr;      // rows of the inverse kept in the compact copy
  double *stage = nullptr;      // staging of one level's member values
  double *base_stage = nullptr; // staging of a member's full base inverse
  // cached graphs of one surface-solve iteration (as gexec), and the
  // per-level squared surface norms (device, pinned host copy)
  hipGraphExec_t sgexec[kMaxGraphs] = {};
  int sgkind = -1, sglevel = -1;
  const void *sgx = nullptr, *sgb = nullptr;
  double *surf_sq = nullptr;
  double *surf_sq_host = nullptr;
  int surf_sq_host_cap = 0;
  // levels bl..fuse_top run inside one fused coarse-cycle kernel
  int fuse_top = -1;
  CLevel<double> *clev64 = nullptr;
  CLevel<float> *clev32 = nullptr;
  // base solver: partial sums of the base defect norm, first norm
  double *base_part = nullptr;
  double *base_first = nullptr;
  int base_nparts = 0;
  bool base_b_dead = false;   // every base vector has VNCLASS >= 2
  // batched launches: level vectors hold nrhs_cap columns; nb = columns of
  // the work being enqueued
  int nrhs_cap = 1, nb = 1;
  bool fuse_steps = true;   // fold GS defects into the next sweep (UGMG_GS_FUSE=0: off)
  // folded sweeps with the upper and lower row products on two wavefronts
  // per slice (UGMG_GS_SPLIT=0: one wavefront per slice)
  bool gs_split = true;
  // BCGSSolver work (allocated on first use, freed with the hierarchy)
  struct Bcgs {
    int nrhs = 0, maxdef = 0, npart = 0;
    double *x = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *v = nullptr, *s = nullptr,
           *t = nullptr, *partA = nullptr, *partB = nullptr, *defects = nullptr;
    BcgsState *state = nullptr;
    hipGraphExec_t g[2] = {nullptr, nullptr};   // iteration without / with restart
    int glevel = -1, gnrhs = 0, gprec = -1, giter = -1, gk = -1, gmaxdef = -1;
  } bc;
  BcgsState *state_host = nullptr;   // pinned
  int state_host_cap = 0;
};

namespace {

int fail(ugmg_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(ctx, UGMG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(ugmg_ctx *ctx, T **out, size_t n) {
  *out = nullptr;
  if (n == 0) n = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) return fail(ctx, UGMG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocs.push_back(p);
  ctx->dev_bytes += (double)n * sizeof(T);
  *out = (T *)p;
  return 0;
}

template <typename T>
int upload(ugmg_ctx *ctx, T **out, const T *src, size_t n) {
  int rc = dalloc(ctx, out, n);
  if (rc) return rc;
  if (n) HIPCHK(hipMemcpy(*out, src, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// host-side slice metadata of a slab from its entry offsets and offset-list
// starts (per slice)
void sell_meta(DevSellBuf &D, int nslices, const int *soff, const int *sdoff) {
  D.nslices = nslices;
  D.width_h.resize(nslices);
  D.wmax = 0;
  D.nuniform = 0;
  D.nent_explicit = 0;
  D.slot_pre.assign(nslices + 1, 0.0);
  D.expl_pre.assign(nslices + 1, 0.0);
  for (int q = 0; q < nslices; ++q) {
    D.width_h[q] = (soff[q + 1] - soff[q]) / kSlice;
    D.wmax = std::max(D.wmax, D.width_h[q]);
    const size_t slots = (size_t)(soff[q + 1] - soff[q]);
    D.nuniform += sdoff[q] >= 0;
    if (sdoff[q] < 0) D.nent_explicit += slots;
    D.slot_pre[q + 1] = D.slot_pre[q] + (double)slots;
    D.expl_pre[q + 1] = D.expl_pre[q] + (sdoff[q] < 0 ? (double)slots : 0.0);
  }
}

// members > 0: `members` value planes filled later through the index maps
// (the slab values are those maps), else the values themselves
int upload_sell(ugmg_ctx *ctx, const HostSell &S, DevSellBuf &D, int members = 0) {
  sell_meta(D, S.nslices, S.soff.data(), S.sdoff.data());
  D.nent = S.col.size();
  int rc;
  if ((rc = upload(ctx, &D.soff, S.soff.data(), S.soff.size()))) return rc;
  if ((rc = upload(ctx, &D.rows, S.rows.data(), S.rows.size()))) return rc;
  if ((rc = upload(ctx, &D.col, S.col.data(), S.col.size()))) return rc;
  if ((rc = upload(ctx, &D.sdoff, S.sdoff.data(), S.sdoff.size()))) return rc;
  if ((rc = upload(ctx, &D.dtab, S.dtab.data(), S.dtab.size()))) return rc;
  D.ndtab = S.dtab.size();
  D.bb = S.bb;
  D.nval = S.val.size();
  if (members > 0) {   // the values are the index maps; member_set fills the planes
    if ((rc = dalloc(ctx, &D.v64, D.nval * members)) || (rc = dalloc(ctx, &D.v32, D.nval * members)))
      return rc;
    HIPCHK(hipMemset(D.v64, 0, sizeof(double) * D.nval * members));
    HIPCHK(hipMemset(D.v32, 0, sizeof(float) * D.nval * members));
    D.mst = D.nval;
    return 0;
  }
  if ((rc = upload(ctx, &D.v64, S.val.data(), D.nval))) return rc;
  hvec<float> v32;
  to_float(S.val.data(), S.val.size(), v32);
  if ((rc = upload(ctx, &D.v32, v32.data(), v32.size()))) return rc;
  D.mst = 0;
  return 0;
}

// a slab laid out on the device (build_level_device): its arrays are the
// context's already; the FP32 copy of the values is converted here
int adopt_sell(ugmg_ctx *ctx, const DevSellParts &S, DevSellBuf &D) {
  sell_meta(D, S.nslices, S.soff_h.data(), S.sdoff_h.data());
  D.nent = S.nent;
  D.soff = S.soff;
  D.rows = S.rows;
  D.col = S.col;
  D.sdoff = S.sdoff;
  D.dtab = S.dtab;
  D.ndtab = S.ndtab;
  D.bb = S.bb;
  D.nval = S.nent * S.bb;
  D.v64 = S.val;
  int rc;
  if ((rc = dalloc(ctx, &D.v32, D.nval))) return rc;
  if (D.nval) launch_convert<float, double>(ctx->stream, (int)D.nval, D.v64, D.v32, 1, 0, 0);
  HIPCHK(hipGetLastError());
  D.mst = 0;
  return 0;
}

void drop_dcsr(ugmg_ctx *ctx, int level) {
  if (level < 0 || level >= (int)ctx->dcsr.size()) return;
  auto &d = ctx->dcsr[level];
  (void)hipFree(d.rp);
  (void)hipFree(d.ci);
  (void)hipFree(d.a);
  d = ugmg_ctx::DevCsr();
}

void drop_bcgs_graphs(ugmg_ctx *ctx) {
  for (auto &g : ctx->bc.g) {
    if (g) hipGraphExecDestroy(g);
    g = nullptr;
  }
  ctx->bc.glevel = -1;
}

void drop_graph(ugmg_ctx *ctx);

void free_device(ugmg_ctx *ctx) {
  drop_graph(ctx);
  ctx->bc = ugmg_ctx::Bcgs();
  ctx->surf_sq = nullptr;
  ctx->base_ims = 0;
  ctx->base_act = nullptr;
  ctx->stage = nullptr;
  ctx->base_stage = nullptr;
  for (void *p : ctx->allocs) hipFree(p);
  ctx->allocs.clear();
  ctx->dev_bytes = 0.0;
  ctx->dev.clear();
  ctx->base_inv = nullptr;
  ctx->base_cinv = nullptr;
  ctx->base_cinv32 = nullptr;
  ctx->base_rowmap = nullptr;
  ctx->base_colperm = nullptr;
  ctx->base_ncyc = 0;
  ctx->base_nact = 0;
  ctx->base_cpos = nullptr;
  ctx->clev64 = nullptr;
  ctx->clev32 = nullptr;
  ctx->fuse_top = -1;
  ctx->base_part = nullptr;
  ctx->base_first = nullptr;
  ctx->part = nullptr;
  ctx->red = nullptr;
  ctx->prepared = false;
  ctx->top = -1;
}

void drop_graph(ugmg_ctx *ctx) {
  for (hipGraphExec_t &g : ctx->gexec)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
  ctx->gkind = -1;
  for (hipGraphExec_t &g : ctx->sgexec)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
  ctx->sgkind = -1;
  drop_bcgs_graphs(ctx);
}

inline int packed_ident(const ugmg_ctx *ctx) {
  if (ctx->nident <= 0) return -1;
  int v = 0;
  for (int i = 0; i < ctx->nident; ++i) v |= ctx->ident[i] << (4 * i);
  return v;
}
inline Batch batch(const ugmg_ctx *ctx, int l) {
  const DevLevel &d = ctx->dev[l];
  Batch bt{ctx->nb, d.vs, d.nb, d.dms, ctx->base_ims, d.vt, d.ntypes};
  bt.idents = packed_ident(ctx);
  return bt;
}
inline Batch B1(const DevLevel &d) { return Batch{1, d.vs, d.nb, 0, 0, d.vt, d.ntypes}; }

// the caller's vectors (list order, column stride ld) to the level's device
// order and back; a plain (converting) copy on levels kept in list order
template <typename TD, typename TS>
void to_dev(ugmg_ctx *ctx, int l, const TS *src, size_t lds, TD *dst, int nrhs = 1) {
  const DevLevel &d = ctx->dev[l];
  if (d.perm) launch_gather<TD, TS>(ctx->stream, d.n, d.perm, src, dst, nrhs, lds, d.vs, d.nb);
  else launch_convert<TD, TS>(ctx->stream, d.n * d.nb, src, dst, nrhs, lds, d.vs);
}
template <typename TD, typename TS>
void to_user(ugmg_ctx *ctx, int l, const TS *src, TD *dst, size_t ldd, int nrhs = 1) {
  const DevLevel &d = ctx->dev[l];
  if (d.perm) launch_scatter<TD, TS>(ctx->stream, d.n, d.perm, src, dst, nrhs, d.vs, ldd, d.nb);
  else launch_convert<TD, TS>(ctx->stream, d.n * d.nb, src, dst, nrhs, d.vs, ldd);
}

// ------------------------------------------------------------ cycle drivers

// lmgc's pre- (post = false) and post-smoother: class and $damp
inline int sm_kind(const ugmg_ctx *ctx, bool post) {
  return post && ctx->p.post_set ? ctx->p.post_smoother : ctx->p.smoother;
}
inline const double *sm_damp(const ugmg_ctx *ctx, bool post) {
  return post && ctx->p.post_set ? ctx->p.post_damp : ctx->p.damp;
}
// a GS layout (colours, device order) when either smoother is GS
inline bool uses_gs(const ugmg_params &p) {
  return p.smoother == UGMG_SMOOTH_GS || (p.post_set && p.post_smoother == UGMG_SMOOTH_GS);
}
// pre- and post-smoother differ in class or damping
inline bool split_smoothers(const ugmg_params &p) {
  if (!p.post_set) return false;
  if (p.post_smoother != p.smoother) return true;
  for (int i = 0; i < UGMG_MAXB; ++i)
    if (p.post_damp[i] != p.damp[i]) return true;
  return false;
}

// the base solve cannot run: ex's band LU met a zero pivot (ls over jac
// does not use it)
inline bool base_singular(const ugmg_ctx *ctx) {
  return ctx->base.failed && ctx->p.base_iter == UGMG_BASE_EX;
}

// damping factors that differ between the components of a block
inline bool per_comp(const ugmg_ctx *ctx, const double *d) {
  for (int i = 1; i < ctx->b; ++i)
    if (d[i] != d[0]) return true;
  return false;
}
// ... or between the norm entries ((vector type, component) pairs)
inline bool per_comp_entries(const ugmg_ctx *ctx, const double *d) {
  for (int i = 1; i < ctx->nc; ++i)
    if (d[i] != d[0]) return true;
  return false;
}

template <typename T>
void smooth_step(ugmg_ctx *ctx, int l, T *t, T *b, T *c, int cmode, T *xout, int k0 = 0,
                 bool post = false) {
  DevLevel &d = ctx->dev[l];
  hipStream_t st = ctx->stream;
  const int kind = sm_kind(ctx, post);
  const double *dv = sm_damp(ctx, post);
  const T damp = (T)dv[0];
  const Batch bt = batch(ctx, l);
  const int mode = cmode < 0 ? 0 : (cmode == 0 ? 3 : 2);
  if (per_comp(ctx, dv)) {
    // Smoother (np/procs/iter.cc:817-842, dscalx at 836) with $damp per component: the step's x,
    // x *= damp per component, then b -= A x (and c (+)= x)
    if (kind == UGMG_SMOOTH_JAC) {
      launch_jac<T>(st, d.n, b, d.diag<T>(), d.vclass, (T)1, t, bt);
    } else {
      for (int k = k0; k < d.ncolors; ++k)
        launch_gs_color<T, T>(st, d.L.d<T>(), d.color_slices[k],
                              d.color_slices[k + 1] - d.color_slices[k], d.color_wmax[k], b,
                              d.diag<T>(), d.vclass, t, d.ntv<T>(), bt);
    }
    launch_scale_comp<T>(st, d.n, d.nb, dv, t, (T *)nullptr, 0, bt);
    launch_spmv<T, T, T>(st, d.A.d<T>(), 0, d.A.nslices, t, b, c, mode, d.ntv<T>(), bt);
    if (xout)   // an error here is left for the caller's hipGetLastError check
      for (int r = 0; r < bt.nrhs; ++r)
        (void)(hipMemcpyAsync(xout + r * bt.vs, t + r * bt.vs, sizeof(T) * d.n * d.nb,
                                hipMemcpyDeviceToDevice, st));
    return;
  }
  if (kind == UGMG_SMOOTH_JAC) {
    launch_jac<T>(st, d.n, b, d.diag<T>(), d.vclass, damp, t, bt);
    launch_spmv<T, T, T>(st, d.A.d<T>(), 0, d.A.nslices, t, b, c, mode, d.ntv<T>(), bt);
  } else {
    for (int k = k0; k < d.ncolors; ++k)
      launch_gs_color<T, T>(st, d.L.d<T>(), d.color_slices[k],
                            d.color_slices[k + 1] - d.color_slices[k], d.color_wmax[k], b,
                            d.diag<T>(), d.vclass, t, d.ntv<T>(), bt);
    launch_gs_defect<T, T>(st, d.U.d<T>(), t, b, d.vclass, damp, cmode < 0 ? nullptr : c,
                           cmode < 0 ? 1 : cmode, xout, d.ntv<T>(), bt);
  }
}

// nsteps smoother steps (cmode: correction mode of the first).  On a GS
// level whose rows are all active, the defect of every step but the last is
// folded into the next step's colour sweeps.  k0 = 1: the first step's
// colour-0 sweep has already written t (by the kernel that produced b).
template <typename T>
void smooth(ugmg_ctx *ctx, int l, T *t, T *b, T *c, int cmode, int nsteps, int k0 = 0,
            bool post = false) {
  DevLevel &d = ctx->dev[l];
  if (sm_kind(ctx, post) != UGMG_SMOOTH_GS || !d.all_active || nsteps < 2 || !ctx->fuse_steps ||
      per_comp(ctx, sm_damp(ctx, post))) {
    for (int i = 0; i < nsteps; ++i)
      smooth_step<T>(ctx, l, t, b, c, i == 0 ? cmode : 1, nullptr, i == 0 ? k0 : 0, post);
    return;
  }
  hipStream_t st = ctx->stream;
  const T damp = (T)sm_damp(ctx, post)[0];
  const Batch bt = batch(ctx, l);
  for (int k = k0; k < d.ncolors; ++k)
    launch_gs_color<T, T>(st, d.L.d<T>(), d.color_slices[k],
                          d.color_slices[k + 1] - d.color_slices[k], d.color_wmax[k], b,
                          d.diag<T>(), d.vclass, t, d.ntv<T>(), bt);
  for (int i = 1; i < nsteps; ++i)
    for (int k = 0; k < d.ncolors; ++k)
      launch_gs_color_fused<T, T>(st, d.L.d<T>(), d.U.d<T>(), d.color_slices[k],
                                  d.color_slices[k + 1] - d.color_slices[k], d.color_wmax[k],
                                  d.color_wmax_u[k], b, d.diag<T>(), damp, t, c,
                                  i == 1 ? cmode : 1, d.ntv<T>(), bt, ctx->gs_split);
  launch_gs_defect<T, T>(st, d.U.d<T>(), t, b, d.vclass, damp, c, 1, nullptr, d.ntv<T>(), bt);
}

// A GS level whose rows are all active (scalar): the colour-0 sweep of a
// smoother step (no lower neighbours: v_r = b_r / a_rr) can run inside the
// kernel that produces b, bit-identical to the separate sweep
bool fold_colour0(const ugmg_ctx *ctx, int l, int nsteps, bool post = false) {
  const DevLevel &d = ctx->dev[l];
  return sm_kind(ctx, post) == UGMG_SMOOTH_GS && d.all_active && d.nb == 1 && ctx->fuse_steps &&
         nsteps >= 1 && d.ncolors > 0 && d.c0;
}

// Post-smoothing after the coarse-grid correction t: b -= A t also sweeps
// colour 0 (A's rows are grouped by colour, so colour 0 is its first
// color_slices[1] slices) into u, since t is still read; the steps then run
// with u as the iterate
template <typename T>
void post_smooth_folded(ugmg_ctx *ctx, int l, const T *t, T *b, T *c, int nsteps) {
  DevLevel &d = ctx->dev[l];
  T *u = d.uv<T>();
  launch_spmv_c0<T, T>(ctx->stream, d.A.d<T>(), t, b, d.diag<T>(), u, d.color_slices[1],
                       d.ntv<T>(), batch(ctx, l));
  smooth<T>(ctx, l, u, b, c, 1, nsteps, 1, true);
}

BaseArgs base_args(ugmg_ctx *ctx) {
  BaseArgs ba;
  ba.n = ctx->dev[ctx->bl].n;
  ba.maxiter = ctx->p.base_maxiter;
  ba.failed = ctx->base.failed ? 1 : 0;
  ba.damp = ctx->p.base_damp[0];
  ba.abslimit = ctx->p.base_abslimit[0];
  ba.red = ctx->p.base_red[0];
  ba.inv = ctx->base_inv;
  return ba;
}

// Lmgc on levels top..bl in one workgroup (levels at most a few thousand
// vectors, where a separate launch per operation would be latency bound)
template <typename T>
void coarse_cycle(ugmg_ctx *ctx, int top, T *c, T *b, int cmode) {
  CoarseArgs<T> a;
  if constexpr (std::is_same_v<T, double>) a.lv = ctx->clev64;
  else a.lv = ctx->clev32;
  a.bl = ctx->bl;
  a.top = top;
  const ugmg_params &p = ctx->p;
  a.cp = CycleParams{p.nu1, p.nu2, p.gamma, p.smoother, p.damp[0], p.lmgc_damp[0]};
  a.base = base_args(ctx);
  a.baseA = ctx->dev[ctx->bl].A.d64();
  a.base_vclass = ctx->dev[ctx->bl].vclass;
  a.c_top = c;
  a.b_top = b;
  a.cmode_top = cmode;
  launch_coarse_cycle<T>(ctx->stream, a, ctx->dev[ctx->bl].n);
}

// ls over ex on the base level: `maxiter` GEMVs with the dense inverse.  The
// first defect norm comes from the restriction's partial sums (have_parts)
// or from a separate reduction.  Inside a cycle the base-level defect b is
// dead once the coarse correction is known (the next restriction overwrites
// every entry when all base vectors have VNCLASS >= 2), so b -= A x is only
// formed when a later iteration or the caller needs it.
// restricted: b comes from the restriction inside a cycle (zero on the
// skipped Dirichlet unknowns, and b -= A x keeps it so)
template <typename T>
void base_solve(ugmg_ctx *ctx, T *c, T *b, int cmode, bool have_parts, bool need_b,
                bool restricted = false, const T *bc = nullptr) {
  DevLevel &d = ctx->dev[ctx->bl];
  hipStream_t st = ctx->stream;
  const ugmg_params &p = ctx->p;
  Batch bt = batch(ctx, ctx->bl);
  if (per_comp(ctx, p.base_damp)) {   // ex $damp per component
    for (int q = 0; q < d.nb && q < 4; ++q) bt.pdamp.d[q] = p.base_damp[q];
    bt.pdamp.vb = d.nb;
  }
  int np = have_parts ? ctx->base_nparts
                      : launch_dot_part<T>(st, d.n, b, d.vclass, ctx->base_part, bt);
  T *x = d.tv<T>();
  const int m = std::max(0, p.base_maxiter);
  if (m == 0 && cmode == 0) launch_fill<T>(st, d.n, c, (T)0, bt);
  if (p.base_iter == UGMG_BASE_JAC) {
    // ls over jac (LinearSolver np/procs/ls.cc:637-749 with Smoother iter.cc:817-842):
    // x = damp D^-1 b; x = 0 once the defect before it meets the stop test
    // (every later step is then a no-op, as the loop's break); c (+)= x;
    // b -= A x; the next defect's partial norms
    const bool pc = per_comp(ctx, p.base_damp);
    for (int it = 0; it < m; ++it) {
      launch_jac<T>(st, d.n, b, d.diag<T>(), d.vclass, (T)(pc ? 1.0 : p.base_damp[0]), x, bt);
      if (pc) launch_scale_comp<T>(st, d.n, d.nb, p.base_damp, x, (T *)nullptr, 0, bt);
      launch_base_gate<T>(st, d.n * d.nb, x, c, it == 0 ? cmode : 1, p.base_abslimit[0], p.base_red[0],
                          ctx->base_part, np, it, ctx->base_first, bt);
      launch_spmv<T, T, T>(st, d.A.d<T>(), 0, d.A.nslices, x, b, (T *)nullptr, 0, d.ntv<T>(), bt);
      if (it + 1 < m) np = launch_dot_part<T>(st, d.n, b, d.vclass, ctx->base_part, bt);
    }
    return;
  }
  // the FP32 cycle reads the FP64 inverse (sums in FP64) unless base_inv32
  // asks for its FP32 copy
  const bool inv64 = std::is_same_v<T, double> || !p.base_inv32;
  for (int it = 0; it < m; ++it) {
    if (!ctx->base.failed) {
      const int ncols = restricted ? ctx->base_ncyc : d.n * d.nb;
      if (inv64)
        launch_base_gemv<T, double>(st, d.n * d.nb, ctx->base_cinv, ctx->base_rowmap,
                                    ctx->base_colperm, ncols, b, it == 0 ? bc : nullptr, c, x,
                                    it == 0 ? cmode : 1, p.base_damp[0], p.base_abslimit[0],
                                    p.base_red[0], ctx->base_part, np, ctx->base_first, it, bt);
      else
        launch_base_gemv<T, float>(st, d.n * d.nb, ctx->base_cinv32, ctx->base_rowmap,
                                   ctx->base_colperm, ncols, b, it == 0 ? bc : nullptr, c, x,
                                   it == 0 ? cmode : 1, p.base_damp[0], p.base_abslimit[0],
                                   p.base_red[0], ctx->base_part, np, ctx->base_first, it, bt);
    }
    const bool more = it + 1 < m;
    if (more || need_b || !ctx->base_b_dead)
      launch_spmv<T, T, T>(st, d.A.d<T>(), 0, d.A.nslices, x, b, (T *)nullptr, 0, d.ntv<T>(), bt);
    if (more) np = launch_dot_part<T>(st, d.n, b, d.vclass, ctx->base_part, bt);
  }
}

// Lmgc on level l; cmode 0: c := correction, 1: c += correction
template <typename T>
void lmgc(ugmg_ctx *ctx, int l, T *c, T *b, int cmode, bool c0_done = false) {
  const ugmg_params &p = ctx->p;
  hipStream_t st = ctx->stream;
  if (l <= ctx->fuse_top && ctx->nb == 1) {
    coarse_cycle<T>(ctx, l, c, b, cmode);
    return;
  }
  if (l <= ctx->bl) {
    base_solve<T>(ctx, c, b, cmode, false, true);
    return;
  }
  DevLevel &d = ctx->dev[l];
  DevLevel &dc = ctx->dev[l - 1];
  T *t = d.tv<T>();
  const int nb = ctx->nb;
  smooth<T>(ctx, l, t, b, c, cmode, p.nu1, c0_done ? 1 : 0);
  if (p.nu1 == 0 && cmode == 0) launch_fill<T>(st, d.n, c, (T)0, batch(ctx, l));
  // RestrictDefect(level, b, b, Factor_One); on a coarse level that runs
  // lmgc below (not the base solve or the fused coarse cycle) also its
  // first pre-smoothing sweep's colour 0
  const bool to_base = l - 1 == ctx->bl && (ctx->fuse_top < ctx->bl || nb > 1);
  const bool pre0 = l - 1 > ctx->bl && l - 1 > ctx->fuse_top && p.gamma >= 1 &&
                    fold_colour0(ctx, l - 1, p.nu1);
  // onto the base level (scalar) also b in the inverse's column order, into
  // the base level's spare vector u, for the first base iteration.  With
  // several vector types the base solve forms its own per-type norms.
  const bool typed = dc.vt != nullptr;
  const bool bcol = to_base && nb == 1 && ctx->base_cpos && !ctx->base.failed && !typed &&
                    p.base_iter == UGMG_BASE_EX;
  const int np = launch_restrict<T>(st, d.PT.d<T>(), dc.skip, dc.vnclass, dc.vclass, b, (T)1,
                                    dc.bv<T>(), to_base && !typed ? ctx->base_part : nullptr, d.ntv<T>(), nb,
                                    d.vs, dc.vs, d.nb, pre0 ? dc.c0 : nullptr, dc.diag<T>(),
                                    pre0 ? dc.tv<T>() : nullptr, bcol ? ctx->base_cpos : nullptr,
                                    dc.uv<T>(), dc.dms);
  if (p.gamma <= 0) launch_fill<T>(st, dc.n, dc.cv<T>(), (T)0, batch(ctx, l - 1));
  for (int g = 0; g < p.gamma; ++g) {
    if (to_base) {
      // the first visit uses the restriction's partial norms; after a
      // base solve b has changed, so later visits recompute them
      if (g == 0) ctx->base_nparts = np;
      // a later visit (gamma > 1) reads the base defect the earlier one leaves
      base_solve<T>(ctx, dc.cv<T>(), dc.bv<T>(), g == 0 ? 0 : 1, g == 0 && !typed, g + 1 < p.gamma, true,
                    g == 0 && bcol ? dc.uv<T>() : nullptr);
    } else {
      lmgc<T>(ctx, l - 1, dc.cv<T>(), dc.bv<T>(), g == 0 ? 0 : 1, g == 0 && pre0);
    }
  }
  // InterpolateCorrection(level, t, c, damp); dadd(c, t); dmatmul_minus(b, A, t)
  if (per_comp(ctx, p.lmgc_damp)) {   // damp per component: dscalx after the sum
    launch_interp<T>(st, d.P.d<T>(), d.iskip, dc.cv<T>(), (T)1, t, (T *)nullptr, d.ntv<T>(), nb, d.vs,
                     dc.vs, d.nb);
    launch_scale_comp<T>(st, d.n, d.nb, p.lmgc_damp, t, c, 1, batch(ctx, l));
  } else {
    launch_interp<T>(st, d.P.d<T>(), d.iskip, dc.cv<T>(), (T)p.lmgc_damp[0], t, c, d.ntv<T>(), nb, d.vs,
                     dc.vs, d.nb);
  }
  if (fold_colour0(ctx, l, p.nu2, true)) {
    post_smooth_folded<T>(ctx, l, t, b, c, p.nu2);
    return;
  }
  launch_spmv<T, T, T>(st, d.A.d<T>(), 0, d.A.nslices, t, b, (T *)nullptr, 0, d.ntv<T>(),
                       batch(ctx, l));
  smooth<T>(ctx, l, t, b, c, 1, p.nu2, 0, true);
}

int check_ready(ugmg_ctx *ctx, int level) {
  if (!ctx) return UGMG_E_ARG;
  if (!ctx->prepared) return fail(ctx, UGMG_E_STATE, "ugmg_preprocess has not been called");
  if (level < ctx->bl || level > ctx->top)
    return fail(ctx, UGMG_E_ARG, "level outside the preprocessed range");
  return 0;
}

int surface_norm(ugmg_ctx *ctx, int level, const double *x, double *out) {
  DevLevel &d = ctx->dev[level];
  launch_dot<double>(ctx->stream, d.n, x, nullptr, d.vclass, ctx->part, ctx->red, 1, B1(d));
  HIPCHK(hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double) * d.nc(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < d.nc(); ++i) out[i] = ctx->red_host[i];
  return 0;
}

// one iteration of the linear solver, enqueued on the stream; leaves the
// defect norm in ctx->red[0]
void enqueue_iteration(ugmg_ctx *ctx, int level, double *x, double *b, int cycles) {
  DevLevel &d = ctx->dev[level];
  if (ctx->p.precision == UGMG_MIXED) {
    const int k = std::max(1, cycles);
    for (int j = 0; j < k; ++j) lmgc<float>(ctx, level, d.c32, d.b32, j == 0 ? 0 : 1);
    int np = 0;
    launch_ir_update(ctx->stream, d.A.d64(), d.c32, b, x, d.b32, d.vclass, ctx->part, &np, d.nt,
                     B1(d));
    // several vector types: the norms per (type, component) of the FP64 defect
    if (d.vt) launch_dot<double>(ctx->stream, d.n, b, nullptr, d.vclass, ctx->part, ctx->red, 1, B1(d));
    else launch_reduce_final(ctx->stream, np, ctx->part, ctx->red, 1, d.nb);
  } else {
    lmgc<double>(ctx, level, x, b, 1);
    launch_dot<double>(ctx->stream, d.n, b, nullptr, d.vclass, ctx->part, ctx->red, 1, B1(d));
  }
}

bool can_capture(ugmg_ctx *ctx) { return ctx->p.use_graph && ctx->stream != nullptr; }

// enqueue one solver iteration (from the cached graph when enabled) and the
// asynchronous copy of its defect norm to ctx->red_host[0]
// cycles: FP32 cycles of a mixed step (ignored in FP64)
int launch_iteration(ugmg_ctx *ctx, int level, double *x, double *b, int cycles) {
  const bool mixed = ctx->p.precision == UGMG_MIXED;
  cycles = mixed ? std::max(1, cycles) : 1;
  if (!can_capture(ctx) || cycles >= ugmg_ctx::kMaxGraphs) {
    enqueue_iteration(ctx, level, x, b, cycles);
    HIPCHK(hipGetLastError());
  } else {
    const int kind = ctx->p.precision;
    if (ctx->gkind != kind || ctx->glevel != level || ctx->gx != x || ctx->gb != b) {
      drop_graph(ctx);
      ctx->gkind = kind;
      ctx->glevel = level;
      ctx->gx = x;
      ctx->gb = b;
    }
    hipGraphExec_t &ge = ctx->gexec[cycles];
    if (!ge) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      enqueue_iteration(ctx, level, x, b, cycles);
      HIPCHK(hipStreamEndCapture(ctx->stream, &g));
      hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        ge = nullptr;
        return fail(ctx, UGMG_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    HIPCHK(hipGraphLaunch(ge, ctx->stream));
  }
  HIPCHK(hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double) * ctx->nc, hipMemcpyDeviceToHost,
                        ctx->stream));
  return 0;
}

int run_iteration(ugmg_ctx *ctx, int level, double *x, double *b, int cycles) {
  int rc = launch_iteration(ctx, level, x, b, cycles);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

inline bool sc_cmp1(double x, double y) { return std::fabs(x) < std::fabs(y); }

// sc_cmp (np/udm/npscan.cc:1027-1057) over n entries: |x_i| < |y_i| for
// every i, or with VD_IDENT, per group leader i (ident[i] == i) the group's
// sum of x_j^2 below its sum of y_j^2
bool sc_cmp_vd(const ugmg_ctx *ctx, int n, const double *x, const double *y) {
  if (ctx->nident == n) {
    for (int i = 0; i < n; ++i) {
      if (ctx->ident[i] != i) continue;
      double sx = 0.0, sy = 0.0;
      for (int j = 0; j < n; ++j)
        if (ctx->ident[j] == i) {
          sx += x[j] * x[j];
          sy += y[j] * y[j];
        }
      if (sx >= sy) return false;
    }
    return true;
  }
  for (int i = 0; i < n; ++i)
    if (!sc_cmp1(x[i], y[i])) return false;
  return true;
}

// ------------------------------------------------------------ byte model
// Algorithmic bytes of each pass as the data are laid out, the figures the
// roofline fractions divide by: slab bytes (DevSellBuf::bytes), every vector
// a pass touches counted once per pass (a gathered vector once, one read
// and written twice), flags one byte per vector.  sv = bytes per value and
// vector entry (8 FP64, 4 FP32).
enum class Op { Spmv, SpmvC, SpmvC0, Jac, Restrict, Interp };

double vbytes(const DevLevel &d, int sv) { return (double)d.n * d.nb * sv; }

double op_bytes(const ugmg_ctx *ctx, int l, Op op, int sv) {
  const DevLevel &d = ctx->dev[l];
  const double v = vbytes(d, sv);
  switch (op) {
    case Op::Spmv: return d.A.bytes(sv) + 3.0 * v;    // x gathered, y read and written
    case Op::SpmvC: return d.A.bytes(sv) + 5.0 * v;   // and c += x
    case Op::SpmvC0: {                                // colour-0 rows: diagonal read, u written
      const double r0 = d.color_rows.empty() ? 0.0 : d.color_rows[0];
      return d.A.bytes(sv) + 3.0 * v + r0 * d.nb * sv * (d.nb + 1);
    }
    case Op::Jac: return (double)d.n * (d.nb * sv * (2.0 + d.nb) + 1.0);   // b, diagonal block, class, t
    case Op::Restrict: {                               // b_f gathered, b_c written; skip, VNCLASS, VCLASS
      const DevLevel &dc = ctx->dev[l - 1];
      return d.PT.bytes(sv) + v + vbytes(dc, sv) + 6.0 * dc.n;
    }
    case Op::Interp: {                                 // c_c gathered, t written, c read and written; skip
      const DevLevel &dc = ctx->dev[l - 1];
      return d.P.bytes(sv) + vbytes(dc, sv) + 3.0 * v + 4.0 * d.n;
    }
  }
  return 0.0;
}

// colour sweeps k0.. of one GS step: L slab of those colours; b, diagonal
// block and class read and v written per row; v gathered
double sweep_bytes(const DevLevel &d, int sv, int k0) {
  if (d.ncolors == 0) return 0.0;
  double rows = 0.0;
  for (int k = k0; k < d.ncolors; ++k) rows += d.color_rows[k];
  return d.L.bytes_range(d.color_slices[k0], d.color_slices[d.ncolors], sv) +
         rows * (d.nb * sv * (2.0 + d.nb) + 1.0) + vbytes(d, sv);
}
// the defect after the sweeps: U slab; b read and written, v, class, c
// (read and written, or written), v gathered
double defect_bytes(const DevLevel &d, int sv, bool c_rmw) {
  return d.U.bytes(sv) + vbytes(d, sv) * (5.0 + (c_rmw ? 1.0 : 0.0)) + d.n;
}
// a step after the first with the previous defect folded in: L and U
// slabs; b read and written, diagonal block, old and new v, c; v gathered
double fused_bytes(const DevLevel &d, int sv, bool c_rmw) {
  return d.L.bytes(sv) + d.U.bytes(sv) +
         (double)d.n * d.nb * sv * (2.0 + d.nb + 3.0 + (c_rmw ? 2.0 : 1.0));
}
double gs_step_bytes(const ugmg_ctx *ctx, int l, int sv, bool fused) {
  const DevLevel &d = ctx->dev[l];
  return fused ? fused_bytes(d, sv, true) : sweep_bytes(d, sv, 0) + defect_bytes(d, sv, true);
}

// nsteps smoother steps as smooth() issues them
double smooth_bytes(const ugmg_ctx *ctx, int l, int sv, int cmode, int nsteps, int k0, bool post = false) {
  const DevLevel &d = ctx->dev[l];
  double by = 0.0;
  if (sm_kind(ctx, post) == UGMG_SMOOTH_JAC) {
    for (int i = 0; i < nsteps; ++i)
      by += op_bytes(ctx, l, Op::Jac, sv) + d.A.bytes(sv) + vbytes(d, sv) * ((i == 0 && cmode == 0) ? 4.0 : 5.0);
    return by;
  }
  if (!d.all_active || nsteps < 2 || !ctx->fuse_steps || per_comp(ctx, sm_damp(ctx, post))) {
    for (int i = 0; i < nsteps; ++i)
      by += sweep_bytes(d, sv, i == 0 ? k0 : 0) + defect_bytes(d, sv, (i == 0 ? cmode : 1) == 1);
    return by;
  }
  by = sweep_bytes(d, sv, k0);
  for (int i = 1; i < nsteps; ++i) by += fused_bytes(d, sv, (i == 1 ? cmode : 1) == 1);
  return by + defect_bytes(d, sv, true);
}

// ls over ex on the base level (m GEMVs over the compact inverse, whose
// entries are FP64 or the FP32 copy), with the defects between them
double base_bytes(const ugmg_ctx *ctx, int sv, bool restricted, bool need_b) {
  const DevLevel &d = ctx->dev[ctx->bl];
  const int m = std::max(0, ctx->p.base_maxiter);
  const int nsc = d.n * d.nb;
  const int si = (sv == 4 && ctx->p.base_inv32) ? 4 : 8;
  const double ncols = restricted ? ctx->base_ncyc : nsc;
  double by = 0.0;
  if (ctx->p.base_iter == UGMG_BASE_JAC) {   // step, gate (x, c), b -= A x, norms
    for (int it = 0; it < m; ++it)
      by += op_bytes(ctx, ctx->bl, Op::Jac, sv) + vbytes(d, sv) * 3.0 + op_bytes(ctx, ctx->bl, Op::Spmv, sv) +
            (it + 1 < m ? vbytes(d, sv) + d.n : 0.0);
    return by;
  }
  for (int it = 0; it < m; ++it) {
    if (!ctx->base.failed)
      by += (double)ctx->base_nact * ncols * si + ncols * (sv + 4.0) + 4.0 * nsc + 2.0 * nsc * sv;
    const bool more = it + 1 < m;
    if (more || need_b || !ctx->base_b_dead) by += op_bytes(ctx, ctx->bl, Op::Spmv, sv);
    if (more) by += vbytes(d, sv) + d.n;
  }
  return by;
}

// one lmgc cycle as lmgc() issues it (the fused coarse-cycle kernel, an
// opt-in for tiny hierarchies, is not modelled and counts 0)
double lmgc_bytes(const ugmg_ctx *ctx, int l, int sv, int cmode, bool c0_done) {
  const ugmg_params &p = ctx->p;
  if (l <= ctx->fuse_top) return 0.0;
  if (l <= ctx->bl) return base_bytes(ctx, sv, false, true) + vbytes(ctx->dev[l], sv) + ctx->dev[l].n;
  const DevLevel &d = ctx->dev[l], &dc = ctx->dev[l - 1];
  double by = smooth_bytes(ctx, l, sv, cmode, p.nu1, c0_done ? 1 : 0);
  if (p.nu1 == 0 && cmode == 0) by += vbytes(d, sv);
  const bool to_base = l - 1 == ctx->bl && ctx->fuse_top < ctx->bl;
  const bool pre0 = l - 1 > ctx->bl && l - 1 > ctx->fuse_top && p.gamma >= 1 && fold_colour0(ctx, l - 1, p.nu1);
  by += op_bytes(ctx, l, Op::Restrict, sv);
  if (pre0) by += dc.n + (double)dc.color_rows[0] * 2.0 * sv;
  if (to_base && ctx->nb == 1 && ctx->base_cpos && !ctx->base.failed && !dc.vt &&
      p.base_iter == UGMG_BASE_EX)
    by += 4.0 * dc.n + (double)ctx->base_ncyc * sv;
  if (p.gamma <= 0) by += vbytes(dc, sv);
  for (int g = 0; g < p.gamma; ++g)
    by += to_base ? base_bytes(ctx, sv, true, g + 1 < p.gamma)
                  : lmgc_bytes(ctx, l - 1, sv, g == 0 ? 0 : 1, g == 0 && pre0);
  by += op_bytes(ctx, l, Op::Interp, sv);
  if (fold_colour0(ctx, l, p.nu2, true))
    return by + op_bytes(ctx, l, Op::SpmvC0, sv) + smooth_bytes(ctx, l, sv, 1, p.nu2, 1, true);
  return by + op_bytes(ctx, l, Op::Spmv, sv) + smooth_bytes(ctx, l, sv, 1, p.nu2, 0, true);
}

double cycle_bytes(const ugmg_ctx *ctx, int l, int sv) { return lmgc_bytes(ctx, l, sv, 0, false); }

// one iteration of the linear solver (enqueue_iteration): the FP64 cycle
// and the defect norm, or `cycles` FP32 cycles and the FP64 refinement
// update (r -= A c32, x += c32, r32 = r, norm partials)
double iteration_bytes(const ugmg_ctx *ctx, int l, int cycles) {
  const DevLevel &d = ctx->dev[l];
  if (ctx->p.precision != UGMG_MIXED) return lmgc_bytes(ctx, l, 8, 1, false) + vbytes(d, 8) + d.n;
  double by = 0.0;
  for (int j = 0; j < cycles; ++j) by += lmgc_bytes(ctx, l, 4, j == 0 ? 0 : 1, false);
  return by + d.A.bytes(8) + (double)d.n * d.nb * (4.0 + 4.0 + 16.0 + 16.0 + 4.0) + d.n;
}

// ------------------------------------------------------------ BCGSSolver
// Work of BCGSSolver (np/procs/ls.cc:1864-2062) on nb columns of level
// `level`: vectors x, b, r, p, v, s, t in the level's column stride; the
// cycle's input and output are the level's own c/b vectors (FP64) or
// c32/b32 (FP32 cycle), so q lives in c (c32).
int bcgs_alloc(ugmg_ctx *ctx, int level, int nrhs, int maxdef) {
  auto &w = ctx->bc;
  const DevLevel &d = ctx->dev[level];
  const int np = std::max(bcgs_nparts(d.n), bcgs_spmv_nparts(d.A.nslices));
  if (w.x && w.nrhs >= nrhs && w.maxdef >= maxdef && w.npart >= np) return 0;
  drop_bcgs_graphs(ctx);
  // earlier buffers stay in ctx->allocs until the hierarchy is freed
  const int nr = std::max(nrhs, w.nrhs);
  const int md = std::max(std::max(maxdef, w.maxdef), 1);
  const size_t nv = d.vs * (size_t)nr;
  int rc;
  if ((rc = dalloc(ctx, &w.x, nv)) || (rc = dalloc(ctx, &w.b, nv)) || (rc = dalloc(ctx, &w.r, nv)) ||
      (rc = dalloc(ctx, &w.p, nv)) || (rc = dalloc(ctx, &w.v, nv)) || (rc = dalloc(ctx, &w.s, nv)) ||
      (rc = dalloc(ctx, &w.t, nv)) || (rc = dalloc(ctx, &w.partA, (size_t)2 * np * nr)) ||
      (rc = dalloc(ctx, &w.partB, (size_t)2 * np * nr)) ||
      (rc = dalloc(ctx, &w.defects, (size_t)md * nr)) || (rc = dalloc(ctx, &w.state, nr)))
    return rc;
  for (double *v : {w.x, w.b, w.r, w.p, w.v, w.s, w.t}) HIPCHK(hipMemset(v, 0, sizeof(double) * nv));
  w.nrhs = nr;
  w.maxdef = md;
  w.npart = np;
  if (ctx->state_host_cap < nr) {
    if (ctx->state_host) hipHostFree(ctx->state_host);
    ctx->state_host = nullptr;
    ctx->state_host_cap = 0;
    HIPCHK(hipHostMalloc((void **)&ctx->state_host, sizeof(BcgsState) * nr, hipHostMallocDefault));
    ctx->state_host_cap = nr;
  }
  return 0;
}

// the preconditioner Iter(q, w): c := cycle correction of the input held
// in the level's b (b32), which the cycle overwrites
void bcgs_precond(ugmg_ctx *ctx, int level, bool mixed) {
  DevLevel &d = ctx->dev[level];
  if (mixed) {
    const int k = std::max(1, ctx->p.inner_cycles);
    for (int j = 0; j < k; ++j) lmgc<float>(ctx, level, d.c32, d.b32, j == 0 ? 0 : 1);
  } else {
    lmgc<double>(ctx, level, d.c, d.b, 0);
  }
}

// one loop pass of BCGSSolver on every active column
void enqueue_bcgs_iteration(ugmg_ctx *ctx, int level, bool use_iter, int restart,
                            double abslimit, double red, int maxdef) {
  auto &w = ctx->bc;
  DevLevel &d = ctx->dev[level];
  hipStream_t st = ctx->stream;
  const Batch bt = batch(ctx, level);
  const int nb = ctx->nb, n = d.n;
  const bool mixed = use_iter && ctx->p.precision == UGMG_MIXED;
  const int npv = bcgs_nparts(n), nps = bcgs_spmv_nparts(d.A.nslices);
  const DSell<double> A = d.A.d64();
  // no preconditioner: q = p and q = s, held in c
  double *w64 = !use_iter ? d.c : (mixed ? nullptr : d.b);
  float *w32 = mixed ? d.b32 : nullptr;
  launch_bcgs_scalar(st, BCGS_RHO, npv, w.partA, w.state, nb, restart, abslimit, red, w.defects, maxdef);
  launch_bcgs_p(st, n, w.b, w.r, w.v, w.p, w64, w32, w.state, restart, bt);
  if (use_iter) bcgs_precond(ctx, level, mixed);
  if (mixed) launch_bcgs_spmv<float>(st, A, d.c32, w.v, w.r, d.vclass, w.partB, 0, 0, w.state, d.nt, bt);
  else launch_bcgs_spmv<double>(st, A, d.c, w.v, w.r, d.vclass, w.partB, 0, 0, w.state, d.nt, bt);
  launch_bcgs_scalar(st, BCGS_ALPHA, nps, w.partB, w.state, nb, restart, abslimit, red, w.defects, maxdef);
  if (mixed) launch_bcgs_s<float>(st, n, w.b, w.v, d.c32, w.x, w.s, w64, w32, d.vclass, w.partB, w.state, bt);
  else launch_bcgs_s<double>(st, n, w.b, w.v, d.c, w.x, w.s, w64, w32, d.vclass, w.partB, w.state, bt);
  launch_bcgs_scalar(st, BCGS_RES1, npv, w.partB, w.state, nb, restart, abslimit, red, w.defects, maxdef);
  if (use_iter) bcgs_precond(ctx, level, mixed);
  if (mixed) launch_bcgs_spmv<float>(st, A, d.c32, w.t, w.s, d.vclass, w.partB, 1, 1, w.state, d.nt, bt);
  else launch_bcgs_spmv<double>(st, A, d.c, w.t, w.s, d.vclass, w.partB, 1, 1, w.state, d.nt, bt);
  launch_bcgs_scalar(st, BCGS_OMEGA, nps, w.partB, w.state, nb, restart, abslimit, red, w.defects, maxdef);
  if (mixed) launch_bcgs_b<float>(st, n, w.s, w.t, d.c32, w.r, w.x, w.b, d.vclass, w.partA, w.state, bt);
  else launch_bcgs_b<double>(st, n, w.s, w.t, d.c, w.r, w.x, w.b, d.vclass, w.partA, w.state, bt);
  launch_bcgs_scalar(st, BCGS_RES2, npv, w.partA, w.state, nb, restart, abslimit, red, w.defects, maxdef);
}

int launch_bcgs_iteration(ugmg_ctx *ctx, int level, bool use_iter, int restart, double abslimit,
                          double red, int maxdef) {
  auto &w = ctx->bc;
  if (!can_capture(ctx)) {
    enqueue_bcgs_iteration(ctx, level, use_iter, restart, abslimit, red, maxdef);
    HIPCHK(hipGetLastError());
    return 0;
  }
  const int k = ctx->p.inner_cycles;
  if (w.glevel != level || w.gnrhs != ctx->nb || w.gprec != ctx->p.precision ||
      w.giter != (int)use_iter || w.gk != k || w.gmaxdef != maxdef) {
    drop_bcgs_graphs(ctx);
    w.gmaxdef = maxdef;
    w.glevel = level;
    w.gnrhs = ctx->nb;
    w.gprec = ctx->p.precision;
    w.giter = use_iter;
    w.gk = k;
  }
  hipGraphExec_t &ge = w.g[restart ? 1 : 0];
  if (!ge) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_bcgs_iteration(ctx, level, use_iter, restart, abslimit, red, maxdef);
    HIPCHK(hipStreamEndCapture(ctx->stream, &g));
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      ge = nullptr;
      return fail(ctx, UGMG_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  HIPCHK(hipGraphLaunch(ge, ctx->stream));
  return 0;
}

}  // namespace

// Host half of LmgcPreProcess (CPU only, no HIP call; safe on a worker
// thread while the context is not otherwise used).
// multiply-adds of the base factorisation (about n bw (bw + n)) above which
// it runs on the device: about a second of host time
constexpr double kBaseDeviceWork = 2.0e9;

// FULLREFINELEVEL of the levels set up to `level` (SetSurfaceClasses,
// gm/algebra.cc:2885-2916): the lowest level holding a FINE_GRID_DOF vector
// (VCLASS >= 2, VNCLASS <= 1); `level` when there is none
static int host_fullrefinelevel(const ugmg_ctx *ctx, int level) {
  int frl = level;
  for (int l = level; l >= 0; --l) {
    if (!ctx->level_set[l]) continue;
    const HostLevel &H = ctx->host[l];
    for (size_t v = 0; v < H.vc.size() && v < H.vnc.size(); ++v)
      if (H.vc[v] >= 2 && H.vnc[v] <= 1) {
        frl = l;
        break;
      }
  }
  return frl;
}

// lmgc's $b: a level, or (negative) LmgcInit's choice (np/procs/iter.cc:
// 7643-7648): from FULLREFINELEVEL down, the first level with at most
// -baselevel vectors (level 0 if none); never above `level`
static int resolve_baselevel(const ugmg_ctx *ctx, int level) {
  int bl = ctx->p.baselevel;
  if (bl < 0) {
    int i;
    for (i = host_fullrefinelevel(ctx, level); i > 0; --i)
      if (ctx->level_set[i] && ctx->host[i].n <= -ctx->p.baselevel) break;
    bl = i;
  }
  return std::min(bl, level);
}

struct PrepState {
  int level = -1, bl = 0, fuse_rows = 0;
  int ntypes = 1;   // vector types of the hierarchy
  bool mats = false;   // some level has an interpolation matrix
  double nt_bytes = 64.0e6;
  // layouts built on the device (build_level_device) in prep_device, or on
  // the host here (hp, plv)
  bool device = false;
  int force_win = -1;
  std::vector<DevPrep> dp;
  std::vector<HostPrep> hp;
  std::vector<HostLevel> plv;   // levels in device order (GS levels)
  HostBase base;
  // members: per level, the CSR index of every slab value (see DevLevel)
  std::vector<std::vector<int>> map_a, map_l, map_u, map_d;
};

// the layout of `hp` built from an index-valued copy of the level (value k
// + 1 at CSR value k): the builders only move values, so each slab value
// slot then holds its source index (0 for padding and phantoms)
void index_maps(const HostPrep &idx, std::vector<int> &ma, std::vector<int> &ml,
                std::vector<int> &mu, std::vector<int> &md) {
  auto conv = [](const auto &src, std::vector<int> &dst) {
    dst.resize(src.size());
    for (size_t q = 0; q < src.size(); ++q) dst[q] = (int)src[q] - 1;
  };
  conv(idx.A.val, ma);
  conv(idx.L.val, ml);
  conv(idx.U.val, mu);
  conv(idx.diag, md);
}

int prep_host(ugmg_ctx *ctx, int level, PrepState &ps) {
  if (ctx->nlevels <= 0) return fail(ctx, UGMG_E_STATE, "ugmg_set_nlevels first");
  if (level < 0 || level >= ctx->nlevels) return fail(ctx, UGMG_E_ARG, "level out of range");
  if (!ctx->params_set) return fail(ctx, UGMG_E_STATE, "ugmg_set_params first");
  const ugmg_params &p = ctx->p;
  const int bl = resolve_baselevel(ctx, level);
  for (int l = bl; l <= level; ++l) {
    if (!ctx->level_set[l]) return fail(ctx, UGMG_E_STATE, "level " + std::to_string(l) + " not set");
    if (l > bl && !ctx->transfer_set[l])
      return fail(ctx, UGMG_E_STATE, "transfer of level " + std::to_string(l) + " not set");
  }
  if ((long long)ctx->host[bl].n * ctx->b > 8192)
    return fail(ctx, UGMG_E_ARG, "base level has more than 8192 unknowns; choose a coarser baselevel");
  // vector types (node and edge vectors with b components each): norms,
  // dot products and VEC_SCALARs take one entry per (type, component)
  int nt = 1;
  for (int l = bl; l <= level; ++l)
    for (uint8_t t : ctx->host[l].vt) nt = std::max(nt, t + 1);
  ps.ntypes = nt;
  if (nt > 1) {
    if (nt * ctx->b > UGMG_MAXB)
      return fail(ctx, UGMG_E_ARG, "vector types x components exceed UGMG_MAXB norm entries");
    if (ctx->grouped) return fail(ctx, UGMG_E_ARG, "served contexts take one vector type");
    // damping factors per component apply alike to every vector type (the
    // kernels scale entry j of each vector by damp[j]); one base limit
    for (int i = 1; i < nt * ctx->b; ++i)
      if (p.base_abslimit[i] != p.base_abslimit[0] || p.base_red[i] != p.base_red[0])
        return fail(ctx, UGMG_E_ARG, "one base $abslimit / $red for every vector type and component");
  }
  // interpolation matrices: laid out on the host (the device build takes
  // scalar weights); not in served groups
  bool mats = false;
  for (int l = bl + 1; l <= level; ++l) mats |= !ctx->host[l].pblk.empty();
  if (mats && ctx->grouped) return fail(ctx, UGMG_E_ARG, "served contexts take the standard transfer");
  ps.mats = mats;
  try {
    const uint32_t full = ctx->b >= 32 ? ~0u : (1u << ctx->b) - 1u;
    for (int l = bl; l <= level; ++l) {
      HostLevel &F = ctx->host[l];
      F.pwf.clear();
      if (nt > 1 && F.vt.empty()) F.vt.assign(F.n, 0);   // node vectors
      // the matrix transfer has no edge branch: its entries cover every type
      if (l == bl || F.vt.empty() || !F.has_p || !F.pblk.empty()) continue;
      // StandardIntCorNodeVector's edge loop adds a father object's value
      // only where that object's VECSKIP bit is clear (transgrid.cc:317-336):
      // an object Dirichlet in every component contributes nothing (weight
      // 0); the restriction skips the same components at the coarse row
      const HostLevel &Cl = ctx->host[l - 1];
      std::vector<double> w;
      for (int v = 0; v < F.n; ++v) {
        if (F.vt[v] != 1) continue;
        for (int q = F.pp[v]; q < F.pp[v + 1]; ++q) {
          const uint32_t m = Cl.skip[F.pi[q]] & full;
          if (m == 0) continue;
          if (m != full)
            return fail(ctx, UGMG_E_ARG, "an edge vector's father object is Dirichlet in some but not all components");
          if (w.empty()) w = F.pw;
          w[q] = 0.0;
        }
      }
      F.pwf.swap(w);
    }
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation during preprocess");
  }
  const char *env_nt = std::getenv("UGMG_NT_BYTES");
  ps.nt_bytes = env_nt ? std::atof(env_nt) : 64.0e6;
  const char *env_gsf = std::getenv("UGMG_GS_FUSE");
  ctx->fuse_steps = !(env_gsf && std::atoi(env_gsf) == 0);
  const char *env_split = std::getenv("UGMG_GS_SPLIT");
  ctx->gs_split = !(env_split && std::atoi(env_split) == 0);
  const char *env_fuse = std::getenv("UGMG_FUSE_ROWS");
  ps.fuse_rows = env_fuse ? std::atoi(env_fuse) : 0;
  ps.level = level;
  ps.bl = bl;
  // UGMG_GS_WINDOW: window of the GS device order (unset: chosen per level;
  // 1: list order)
  const char *env_win = std::getenv("UGMG_GS_WINDOW");
  const int force_win = env_win ? std::atoi(env_win) : -1;
  ps.force_win = force_win;
  // where the level layouts are built: UGMG_SETUP=host|device overrides the
  // context's choice; a context with several members keeps the host build
  // (its layouts carry index maps)
  ps.device = ctx->setup_device && !ctx->grouped;
  if (const char *env_setup = std::getenv("UGMG_SETUP")) {
    if (!std::strcmp(env_setup, "host")) ps.device = false;
    else if (!std::strcmp(env_setup, "device")) ps.device = !ctx->grouped;
  }
  if (ps.mats) ps.device = false;
  try {
    ps.hp.assign(level + 1, HostPrep());
    std::vector<HostLevel> &plv = ps.plv;          // levels in device order
    plv.assign(level + 1, HostLevel());
    std::vector<std::vector<int>> pos(level + 1);   // list index -> device position
    // members: the layouts are built from index-valued copies of the levels
    // (value k + 1 at CSR value k); the builders only move values, so every
    // slab value slot then holds its source index, and each member's values
    // are gathered through those maps (member_set)
    std::vector<HostLevel> idx;
    if (ctx->grouped) {
      idx.resize(level + 1);
      for (int l = bl; l <= level; ++l) {
        idx[l] = ctx->host[l];
        for (size_t q = 0; q < idx[l].a.size(); ++q) idx[l].a[q] = (double)(q + 1);
      }
    }
    auto src = [&](int l) -> const HostLevel & { return idx.empty() ? ctx->host[l] : idx[l]; };
    for (int l = bl; l <= level && !ps.device; ++l) {
      const bool gs = uses_gs(p) && l > bl;
      HostPrep &hp = ps.hp[l];
      const bool cperm = l > bl && !pos[l - 1].empty();
      const HostLevel *coarse = l > bl ? (cperm ? &plv[l - 1] : &src(l - 1)) : nullptr;
      if (!gs) {
        build_prep(src(l), coarse, nullptr, 0, hp);
        continue;
      }
      const HostLevel &H = src(l);
      std::vector<int> col;
      const int nc = greedy_color(H, col);
      const int win = force_win >= 0 ? force_win : choose_gs_window(H, col, nc);
      hp.window = win;
      gs_device_order(col, nc, win, hp.perm);
      pos[l].resize(H.n);
      for (int q = 0; q < H.n; ++q) pos[l][hp.perm[q]] = q;
      permute_level(H, hp.perm, pos[l], cperm ? &pos[l - 1] : nullptr, plv[l]);
      std::vector<int> colp(H.n);
      for (int q = 0; q < H.n; ++q) colp[q] = col[hp.perm[q]];
      build_prep(plv[l], coarse, &colp, nc, hp);
      hp.color_ug = col;
    }
    if (ctx->grouped) {
      ps.map_a.assign(level + 1, {});
      ps.map_l.assign(level + 1, {});
      ps.map_u.assign(level + 1, {});
      ps.map_d.assign(level + 1, {});
      for (int l = bl; l <= level; ++l)
        index_maps(ps.hp[l], ps.map_a[l], ps.map_l[l], ps.map_u[l], ps.map_d[l]);
    }
    // EXPreProcess: order and band here.  The band LU and the inverse of a
    // large base level run on the device in prep_device; a small one is
    // factored here, where the server overlaps it with other requests'
    // device work (UGMG_BASE_HOST=1: always here, 0: always on the device).
    build_base_band(ctx->host[bl], p.ex_reorder, ps.base, p.ex_float != 0);
    const char *env_bh = std::getenv("UGMG_BASE_HOST");
    const int base_mode = env_bh ? std::atoi(env_bh) : -1;
    const double nb = ps.base.n, bwb = ps.base.bw;
    if (base_mode == 1 || (base_mode < 0 && nb * bwb * (bwb + nb) < kBaseDeviceWork))
      factor_base_host(ps.base);
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation during preprocess");
  }
  return UGMG_OK;
}

// The band LU and inverse of the base level on the device (bit-identical to
// factor_base_host); sets B.failed on a zero pivot.
int factor_base_device(ugmg_ctx *ctx, HostBase &B) {
  const int n = B.n;
  int rc;
  if ((rc = dalloc(ctx, &ctx->base_inv, (size_t)n * n))) return rc;
  if (n == 0) return 0;
  const size_t ldx = ((size_t)n + 63) / 64 * 64;
  double *band = nullptr, *X = nullptr;
  int *pos = nullptr, *flag = nullptr;
  auto release = [&] {
    hipFree(band);
    hipFree(X);
    hipFree(pos);
    hipFree(flag);
  };
  hipError_t e = hipMalloc((void **)&band, B.band.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&X, (size_t)n * ldx * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&pos, (size_t)n * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&flag, sizeof(int));
  if (e != hipSuccess) {
    release();
    return fail(ctx, UGMG_E_NOMEM, std::string("hipMalloc (base factorisation): ") + hipGetErrorString(e));
  }
  const int zero = 0;
  e = hipMemcpyAsync(band, B.band.data(), B.band.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pos, B.pos.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(flag, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    launch_base_factor(ctx->stream, band, n, B.bw, pos, X, ctx->base_inv, flag, B.fmode);
    e = hipGetLastError();
  }
  int failed = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&failed, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release();
  if (e != hipSuccess)
    return fail(ctx, UGMG_E_HIP, std::string("base factorisation: ") + hipGetErrorString(e));
  B.failed = failed != 0;
  std::vector<double>().swap(B.band);
  return 0;
}

// the layouts of levels bl..level built on the device into ps.dp
int build_layouts_device(ugmg_ctx *ctx, PrepState &ps) {
  const int level = ps.level, bl = ps.bl;
  ps.dp.assign(level + 1, DevPrep());
  DevKeep keep;
  keep.user = ctx;
  keep.alloc = [](void *u, void **p, size_t bytes) {
    return dalloc((ugmg_ctx *)u, (char **)p, bytes);
  };
  int rc = 0;
  for (int l = bl; l <= level && !rc; ++l) {
    const bool gs = uses_gs(ctx->p) && l > bl;
    const int coarse_n = l > bl ? ctx->host[l - 1].n : 0;
    const int *coarse_pos = l > bl ? ps.dp[l - 1].pos : nullptr;
    std::string err;
    const ugmg_ctx::DevCsr dc = l < (int)ctx->dcsr.size() ? ctx->dcsr[l] : ugmg_ctx::DevCsr();
    rc = build_level_device(ctx->stream, ctx->host[l], coarse_n, coarse_pos, gs, ps.force_win, keep,
                            ps.dp[l], err, dc.rp, dc.ci, dc.a);
    if (!rc) drop_dcsr(ctx, l);   // laid out: the host copy serves later builds
    if (rc) rc = fail(ctx, rc, "level " + std::to_string(l) + " layout: " + err);
    if (l > bl && ps.dp[l - 1].pos) {
      (void)hipFree(ps.dp[l - 1].pos);
      ps.dp[l - 1].pos = nullptr;
    }
  }
  for (DevPrep &d : ps.dp)
    if (d.pos) {
      (void)hipFree(d.pos);
      d.pos = nullptr;
    }
  return rc;
}

// DevLevel fields of a level laid out on the device
int adopt_level(ugmg_ctx *ctx, DevPrep &dp, bool has_p, DevLevel &D) {
  int rc;
  D.n = dp.n;
  D.nb = ctx->b;
  D.nnzb = dp.nnzb;
  if (dp.perm) {
    D.perm_h = std::move(dp.perm_h);
    D.perm = dp.perm;
  }
  if ((rc = adopt_sell(ctx, dp.A, D.A))) return rc;
  const size_t nd = (size_t)dp.n * D.nb * D.nb;
  D.diag64 = dp.diag;
  if ((rc = dalloc(ctx, &D.diag32, nd))) return rc;
  launch_convert<float, double>(ctx->stream, (int)nd, D.diag64, D.diag32, 1, 0, 0);
  HIPCHK(hipGetLastError());
  D.dms = 0;
  D.vclass = dp.vc;
  D.vnclass = dp.vnc;
  D.skip = dp.skip;
  D.vt = dp.vt;
  if (dp.ncolors > 0) {
    D.ncolors = dp.ncolors;
    D.color_slices = dp.color_slices;
    D.color = std::move(dp.color_ug_h);
    if ((rc = adopt_sell(ctx, dp.L, D.L)) || (rc = adopt_sell(ctx, dp.U, D.U))) return rc;
    D.c0 = dp.c0;
    D.color_rows = dp.color_rows;
    D.all_active = dp.all_active;
    D.color_wmax.resize(D.ncolors);
    D.color_wmax_u.resize(D.ncolors);
    for (int k = 0; k < D.ncolors; ++k) {
      D.color_wmax[k] = D.L.max_width(D.color_slices[k], D.color_slices[k + 1]);
      D.color_wmax_u[k] = D.U.max_width(D.color_slices[k], D.color_slices[k + 1]);
    }
  }
  if (has_p) {
    D.p_nnz = dp.p_nnz;
    D.pt_nnz = dp.pt_nnz;
    if ((rc = adopt_sell(ctx, dp.P, D.P)) || (rc = adopt_sell(ctx, dp.PT, D.PT))) return rc;
  }
  return 0;
}

// Device half: builds the level layouts on the device (ps.device) or
// uploads those prep_host built, then the base inverse and the work vectors.
int prep_device(ugmg_ctx *ctx, PrepState &ps) {
  free_device(ctx);
  const int level = ps.level, bl = ps.bl;
  try {
    ctx->dev.assign(level + 1, DevLevel());
    ctx->bl = bl;
    int rc;
    int max_slices = 0;
    if (ps.device && (rc = build_layouts_device(ctx, ps))) {
      free_device(ctx);
      return rc;
    }
    for (int l = bl; l <= level; ++l) {
      DevLevel &D = ctx->dev[l];
      if (ps.device) {
        if ((rc = adopt_level(ctx, ps.dp[l], l > bl, D))) return rc;
        max_slices = std::max(max_slices, D.A.nslices);
      } else {
      HostPrep &hp = ps.hp[l];
      const bool permuted = !hp.perm.empty();
      const HostLevel &H = permuted ? ps.plv[l] : ctx->host[l];
      if (permuted) {
        D.perm_h = hp.perm;
        if ((rc = upload(ctx, &D.perm, hp.perm.data(), hp.perm.size()))) return rc;
      }
      D.n = H.n;
      D.nb = ctx->b;
      D.nnzb = H.nnzb;
      max_slices = std::max(max_slices, hp.A.nslices);
      const bool grp = ctx->grouped;
      const int mb = grp ? ctx->members : 0;
      if ((rc = upload_sell(ctx, hp.A, D.A, mb))) return rc;
      if (grp) {   // diagonal blocks 1 until member_set fills a member's
        if ((rc = dalloc(ctx, &D.diag64, hp.diag.size() * mb)) ||
            (rc = dalloc(ctx, &D.diag32, hp.diag.size() * mb)))
          return rc;
        launch_fill<double>(ctx->stream, (int)(hp.diag.size() * mb), D.diag64, 1.0, Batch{1, 0, 1});
        launch_fill<float>(ctx->stream, (int)(hp.diag.size() * mb), D.diag32, 1.0f, Batch{1, 0, 1});
        HIPCHK(hipGetLastError());
      } else {
        if ((rc = upload(ctx, &D.diag64, hp.diag.data(), hp.diag.size()))) return rc;
        hvec<float> d32;
        to_float(hp.diag.data(), hp.diag.size(), d32);
        if ((rc = upload(ctx, &D.diag32, d32.data(), d32.size()))) return rc;
      }
      D.dms = grp ? hp.diag.size() : 0;
      if (grp &&
          ((rc = upload(ctx, &D.map_a, ps.map_a[l].data(), ps.map_a[l].size())) ||
           (rc = upload(ctx, &D.map_d, ps.map_d[l].data(), ps.map_d[l].size()))))
        return rc;
      if ((rc = upload(ctx, &D.vclass, H.vc.data(), H.vc.size()))) return rc;
      if ((rc = upload(ctx, &D.vnclass, H.vnc.data(), H.vnc.size()))) return rc;
      if ((rc = upload(ctx, &D.skip, H.skip.data(), H.skip.size()))) return rc;
      if (!H.vt.empty() && (rc = upload(ctx, &D.vt, H.vt.data(), H.vt.size()))) return rc;
      if (hp.ncolors > 0) {
        D.ncolors = hp.ncolors;
        D.color_slices = hp.color_slices;
        D.color = hp.color_ug;
        if ((rc = upload_sell(ctx, hp.L, D.L, mb))) return rc;
        if ((rc = upload_sell(ctx, hp.U, D.U, mb))) return rc;
        if (grp &&
            ((rc = upload(ctx, &D.map_l, ps.map_l[l].data(), ps.map_l[l].size())) ||
             (rc = upload(ctx, &D.map_u, ps.map_u[l].data(), ps.map_u[l].size()))))
          return rc;
        std::vector<uint8_t> c0(H.n);
        for (int q = 0; q < H.n; ++q) c0[q] = hp.color[q] == 0;
        D.color_rows.assign(D.ncolors, 0);
        for (int q = 0; q < H.n; ++q) ++D.color_rows[hp.color[q]];
        if ((rc = upload(ctx, &D.c0, c0.data(), c0.size()))) return rc;
        D.all_active = true;
        for (uint8_t v : H.vc) D.all_active &= v >= 3;
        D.color_wmax.resize(D.ncolors);
        D.color_wmax_u.resize(D.ncolors);
        for (int k = 0; k < D.ncolors; ++k) {
          D.color_wmax[k] = D.L.max_width(D.color_slices[k], D.color_slices[k + 1]);
          D.color_wmax_u[k] = D.U.max_width(D.color_slices[k], D.color_slices[k + 1]);
        }
      }
      if (l > bl) {
        D.p_nnz = H.pp[H.n];
        D.pt_nnz = (int)hp.pt_idx.size();
        if ((rc = upload_sell(ctx, hp.P, D.P))) return rc;
        if ((rc = upload_sell(ctx, hp.PT, D.PT))) return rc;
      }
      }
      D.A.dms = D.L.dms = D.U.dms = D.dms;
      // vector types: the interpolation reads no VECSKIP bits of edge rows
      D.ntypes = ps.ntypes;
      const bool mat = l > bl && !ctx->host[l].pblk.empty();
      if (ps.ntypes > 1 && D.vt && !mat) {
        if ((rc = dalloc(ctx, &D.iskip, std::max(D.n, 1)))) return rc;
        launch_iskip(ctx->stream, D.n, D.vt, D.skip, D.iskip);
        HIPCHK(hipGetLastError());
      } else {
        if (ps.ntypes == 1) D.vt = nullptr;
        D.iskip = D.skip;
      }
      // scalar GS levels: the colour map of the whole-slab launches over A and U
      if (D.ncolors > 0 && D.ncolors <= kMaxColors && D.nb == 1 && D.A.nslices == D.U.nslices) {
        const int per = kBlock / kSlice, nblk = (D.A.nslices + per - 1) / per;
        for (DevSellBuf *S : {&D.A, &D.U}) {
          S->ncb = D.ncolors;
          for (int k = 0; k < D.ncolors; ++k) S->cb[k] = std::min(D.color_slices[k] / per, nblk);
          S->cb[D.ncolors] = nblk;
        }
      }
      // matrix bytes: b*b values (FP64, or FP32 for the FP32 copy) and one
      // column index per entry
      D.nt = (double)D.A.nent * (8.0 * D.nb * D.nb + 4.0) > ps.nt_bytes;
      D.nt32 = (double)D.A.nent * (4.0 * D.nb * D.nb + 4.0) > ps.nt_bytes;   // the FP32 copy
      // columns padded to 64 entries (512-byte aligned FP64 columns)
      D.vs = ((size_t)std::max(D.n, 1) * D.nb + 63) / 64 * 64;
      const size_t nv = D.vs * (size_t)ctx->nrhs_cap;
      if ((rc = dalloc(ctx, &D.c, nv)) || (rc = dalloc(ctx, &D.b, nv)) ||
          (rc = dalloc(ctx, &D.t, nv)) || (rc = dalloc(ctx, &D.c32, nv)) ||
          (rc = dalloc(ctx, &D.b32, nv)) || (rc = dalloc(ctx, &D.t32, nv)) ||
          (rc = dalloc(ctx, &D.u, nv)) || (rc = dalloc(ctx, &D.u32, nv)))
        return rc;
      // every level vector starts at zero: the zero-valued phantom entries of
      // uniform slices gather from them before a sweep has written them
      for (double *v : {D.c, D.b, D.t, D.u}) HIPCHK(hipMemset(v, 0, sizeof(double) * nv));
      for (float *v : {D.c32, D.b32, D.t32, D.u32}) HIPCHK(hipMemset(v, 0, sizeof(float) * nv));
    }
    // the diagonal blocks the smoothers solve on levels bl+1..level, refused
    // where SolveSmallBlock would refuse them in the first step (l_jac,
    // l_lgs: NUM_SMALL_DIAG); served groups fill theirs per member later
    if ((ctx->b == 2 || ctx->b == 4) && !ctx->grouped && level > bl) {
      int *bad = nullptr;
      std::vector<int> badh(level - bl, INT_MAX);
      if ((rc = upload(ctx, &bad, badh.data(), badh.size()))) return rc;
      for (int l = bl + 1; l <= level; ++l)
        launch_block_check(ctx->stream, ctx->dev[l].n, ctx->b, ctx->dev[l].diag64, ctx->dev[l].vclass,
                           bad + (l - bl - 1));
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(badh.data(), bad, sizeof(int) * badh.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      for (int l = bl + 1; l <= level; ++l) {
        const int r = badh[l - bl - 1];
        if (r == INT_MAX) continue;
        const DevLevel &D = ctx->dev[l];
        const int v = D.perm_h.empty() ? r : D.perm_h[r];
        char msg[160];
        std::snprintf(msg, sizeof msg, "singular %dx%d diagonal block (SolveSmallBlock) of vector %d on level %d",
                      ctx->b, ctx->b, v, l);
        free_device(ctx);
        return fail(ctx, UGMG_E_SINGULAR, msg);
      }
    }
    ctx->base = std::move(ps.base);
    if (!ctx->base.inv.empty() || ctx->base.n == 0) {   // factored on the host
      if ((rc = upload(ctx, &ctx->base_inv, ctx->base.inv.data(), ctx->base.inv.size()))) return rc;
      std::vector<double>().swap(ctx->base.inv);
    } else if ((rc = factor_base_device(ctx, ctx->base))) {
      return rc;
    }
    {   // The GEMVs read the inverse without its unit rows.  A Dirichlet row
        // of the base matrix (its one entry a diagonal 1) is a unit row of
        // L, of U and so of the inverse, exactly, since the LU has no
        // pivoting; its product with b is b's entry.  An FP32 copy serves
        // the FP32 cycle (kept whatever the precision, so that
        // ugmg_set_params may switch it).
      const HostLevel &H = ctx->host[bl];
      const int nb = ctx->dev[bl].nb, nsc = ctx->dev[bl].n * nb;
      if (nsc > 0 && ctx->base_inv) {
        std::vector<int> rowmap(nsc, -1), act, keep, drop;
        for (int v = 0; v < H.n; ++v)
          for (int q = 0; q < nb; ++q) {
            bool unit = H.rp[v] < H.rp[v + 1] && H.ci[H.rp[v]] == v;
            for (int m = H.rp[v]; m < H.rp[v + 1] && unit; ++m)
              for (int j = 0; j < nb; ++j) {
                const double want = (H.ci[m] == v && j == q) ? 1.0 : 0.0;
                if (H.a[((size_t)m * nb + q) * nb + j] != want) {
                  unit = false;
                  break;
                }
              }
            const int u = v * nb + q;
            if (!unit) {
              rowmap[u] = (int)act.size();
              act.push_back(u);
            }
            // the restriction writes 0 to this unknown: VECSKIP component of
            // a vector with VNCLASS >= 2 (StandardRestrictNodeVector)
            const bool zero = unit && ((H.skip[v] >> q) & 1u) && H.vnc[v] >= 2;
            (zero ? drop : keep).push_back(u);
          }
        const int nact = (int)act.size();
        ctx->base_nact = nact;
        ctx->base_ncyc = (int)keep.size();
        std::vector<int> cpos(nsc, -1);
        for (int j = 0; j < ctx->base_ncyc; ++j) cpos[keep[j]] = j;
        if ((rc = upload(ctx, &ctx->base_cpos, cpos.data(), cpos.size()))) return rc;
        keep.insert(keep.end(), drop.begin(), drop.end());
        const size_t ne = std::max<size_t>(1, (size_t)nact * nsc);
        if (act.empty()) act.push_back(0);   // (nothing to copy; keeps the upload non-empty)
        const int mb = ctx->members;
        if ((rc = upload(ctx, &ctx->base_rowmap, rowmap.data(), rowmap.size())) ||
            (rc = upload(ctx, &ctx->base_colperm, keep.data(), keep.size())) ||
            (rc = upload(ctx, &ctx->base_act, act.data(), act.size())) ||
            (rc = dalloc(ctx, &ctx->base_cinv, ne * mb)) || (rc = dalloc(ctx, &ctx->base_cinv32, ne * mb)))
          return rc;
        for (int k = 0; k < mb; ++k)
          launch_base_compact(ctx->stream, nact, nsc, ctx->base_act, ctx->base_colperm, ctx->base_inv,
                              ctx->base_cinv + k * ne, ctx->base_cinv32 + k * ne);
        ctx->base_ims = ctx->grouped ? ne : 0;
        if (ctx->grouped && (rc = dalloc(ctx, &ctx->base_stage, (size_t)nsc * nsc))) return rc;
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
      }
    }
    // fused coarse levels (opt-in): the base level and every level above it
    // up to fuse_rows vectors (and at most kMaxColors colours)
    ctx->fuse_top = -1;
    if (ctx->grouped) {
      size_t mx = 1;
      for (int l = bl; l <= level; ++l) mx = std::max(mx, ctx->host[l].a.size());
      if ((rc = dalloc(ctx, &ctx->stage, mx))) return rc;
    }
    if (ps.fuse_rows > 0 && ctx->dev[bl].n <= 2048 && ctx->b == 1 && !ctx->grouped && ps.ntypes == 1 &&
        !split_smoothers(ctx->p) && ctx->p.base_iter == UGMG_BASE_EX) {
      ctx->fuse_top = bl;
      for (int l = bl + 1; l <= level; ++l) {
        if (ctx->dev[l].n > ps.fuse_rows || ctx->dev[l].ncolors > kMaxColors) break;
        ctx->fuse_top = l;
      }
    }
    ctx->base_b_dead = true;
    for (uint8_t v : ctx->host[bl].vnc) ctx->base_b_dead &= v >= 2;
    std::vector<CLevel<double>> c64(std::max(ctx->fuse_top + 1, 1));
    std::vector<CLevel<float>> c32(std::max(ctx->fuse_top + 1, 1));
    for (int l = bl; l <= ctx->fuse_top; ++l) {
      const DevLevel &D = ctx->dev[l];
      auto fill = [&](auto &C, auto tag) {
        using T = decltype(tag);
        std::memset((void *)&C, 0, sizeof(C));
        C.n = D.n;
        C.ncolors = D.ncolors;
        for (int k = 0; k < (int)D.color_slices.size() && k <= kMaxColors; ++k)
          C.color_slices[k] = D.color_slices[k];
        C.A = D.A.d<T>();
        C.L = D.L.d<T>();
        C.U = D.U.d<T>();
        C.P = D.P.d<T>();
        C.PT = D.PT.d<T>();
        C.diag = D.diag<T>();
        C.vclass = D.vclass;
        C.vnclass = D.vnclass;
        C.skip = D.skip;
        C.c = D.cv<T>();
        C.b = D.bv<T>();
        C.t = D.tv<T>();
      };
      fill(c64[l], double());
      fill(c32[l], float());
    }
    if ((rc = upload(ctx, &ctx->clev64, c64.data(), c64.size()))) return rc;
    if ((rc = upload(ctx, &ctx->clev32, c32.data(), c32.size()))) return rc;
    ctx->part_cap = std::max(1024, (max_slices + 3) / 4 + 1);
    // partial-sum columns: rhs x norm entry (component, or type x component)
    const int nc = ctx->nrhs_cap * ctx->b * ps.ntypes;
    if ((rc = dalloc(ctx, &ctx->part, (size_t)ctx->part_cap * nc))) return rc;
    if ((rc = dalloc(ctx, &ctx->red, std::max(8, nc)))) return rc;
    // restriction partials into the base level: one per block of 4 slices
    const int base_parts = std::max(1024, (ctx->dev[std::min(bl + 1, level)].PT.nslices + 3) / 4 + 1);
    if ((rc = dalloc(ctx, &ctx->base_part, (size_t)base_parts * nc))) return rc;
    if ((rc = dalloc(ctx, &ctx->base_first, std::max(8, nc)))) return rc;
  } catch (const std::bad_alloc &) {
    free_device(ctx);
    return fail(ctx, UGMG_E_NOMEM, "host allocation during preprocess");
  }
  ctx->top = level;
  ctx->ntypes = ps.ntypes;
  ctx->nc = ps.ntypes * ctx->b;
  // FULLREFINELEVEL (SetSurfaceClasses, gm/algebra.cc:2885-2916): the lowest
  // level with a FINE_GRID_DOF vector, the top level if there is none
  ctx->frl = host_fullrefinelevel(ctx, level);
  ctx->prepared = true;
  return UGMG_OK;
}

// ================================================================== C ABI

extern "C" {

void ugmg_params_default(ugmg_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->nu1 = p->nu2 = p->gamma = 1;
  p->baselevel = 0;
  p->smoother = UGMG_SMOOTH_JAC;
  for (int i = 0; i < UGMG_MAXB; ++i) {
    p->damp[i] = p->lmgc_damp[i] = p->base_damp[i] = 1.0;
    p->base_abslimit[i] = 1e-10;
    p->base_red[i] = 1e-12;
  }
  p->base_maxiter = 1;
  p->ex_reorder = 1;
  p->precision = UGMG_FP64;
  p->inner_cycles = 1;
  p->use_graph = 0;
  p->ex_float = 0;
  p->inner_adapt = 1;
  p->base_inv32 = 0;
  p->post_set = 0;
  p->post_smoother = UGMG_SMOOTH_JAC;
  for (int i = 0; i < UGMG_MAXB; ++i) p->post_damp[i] = 1.0;
}

int ugmg_create(int device, ugmg_ctx **out) {
  if (!out) return UGMG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return UGMG_E_NODEV;
  if (device < 0 || device >= ndev) return UGMG_E_ARG;
  ugmg_ctx *ctx = new (std::nothrow) ugmg_ctx;
  if (!ctx) return UGMG_E_NOMEM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&ctx->red_host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    delete ctx;
    return UGMG_E_HIP;
  }
  ctx->own_stream = true;
  ugmg_params_default(&ctx->p);
  resolve_scratch(device);
  *out = ctx;
  return UGMG_OK;
}

int ugmg_destroy(ugmg_ctx *ctx) {
  if (!ctx) return UGMG_E_ARG;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_device(ctx);
  for (int l = 0; l < (int)ctx->dcsr.size(); ++l) drop_dcsr(ctx, l);
  if (ctx->red_host) hipHostFree(ctx->red_host);
  if (ctx->surf_sq_host) hipHostFree(ctx->surf_sq_host);
  if (ctx->state_host) hipHostFree(ctx->state_host);
  if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return UGMG_OK;
}

const char *ugmg_last_error(const ugmg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ugmg_set_stream(ugmg_ctx *ctx, void *stream) {
  if (!ctx) return UGMG_E_ARG;
  drop_graph(ctx);
  if (ctx->own_stream && ctx->stream) {
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
    ctx->stream = nullptr;
  }
  if (stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  return UGMG_OK;
}

int ugmg_sync(ugmg_ctx *ctx) {
  if (!ctx) return UGMG_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return UGMG_OK;
}

int ugmg_set_nlevels(ugmg_ctx *ctx, int nlevels, int b) {
  if (!ctx || nlevels < 1 || nlevels > 32) return ctx ? fail(ctx, UGMG_E_ARG, "nlevels must be 1..32") : UGMG_E_ARG;
  if (b < 1 || b > UGMG_MAXB)
    return fail(ctx, UGMG_E_ARG, "block size must be 1 (scalar) .. 4 (b x b blocks)");
  free_device(ctx);
  for (int l = 0; l < (int)ctx->dcsr.size(); ++l) drop_dcsr(ctx, l);
  ctx->dcsr.assign(nlevels, ugmg_ctx::DevCsr());
  ctx->nlevels = nlevels;
  ctx->b = b;
  ctx->host.assign(nlevels, HostLevel());
  ctx->level_set.assign(nlevels, 0);
  ctx->transfer_set.assign(nlevels, 0);
  return UGMG_OK;
}

int ugmg_set_level(ugmg_ctx *ctx, int level, int n, const int *row_ptr, const int *col_idx,
                   const double *vals, const uint32_t *skip, const uint8_t *vclass,
                   const uint8_t *vnclass) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 0 || level >= ctx->nlevels) return fail(ctx, UGMG_E_ARG, "level out of range");
  if (n < 0 || (n > 0 && (!row_ptr || !col_idx || !vals || !skip || !vclass || !vnclass)))
    return fail(ctx, UGMG_E_ARG, "null array");
  drop_dcsr(ctx, level);
  try {
    HostLevel &L = ctx->host[level];
    L.n = n;
    L.rp.assign(row_ptr, row_ptr + n + 1);
    if (L.rp[0] != 0) return fail(ctx, UGMG_E_ARG, "row_ptr[0] != 0");
    for (int i = 0; i < n; ++i)
      if (L.rp[i + 1] <= L.rp[i]) return fail(ctx, UGMG_E_ARG, "every row needs its diagonal entry");
    L.nnzb = L.rp[n];
    L.ci.assign(col_idx, col_idx + L.nnzb);
    for (int i = 0; i < n; ++i) {
      if (L.ci[L.rp[i]] != i) return fail(ctx, UGMG_E_ARG, "row does not start with its diagonal");
      for (int m = L.rp[i]; m < L.rp[i + 1]; ++m)
        if (L.ci[m] < 0 || L.ci[m] >= n) return fail(ctx, UGMG_E_ARG, "column index out of range");
    }
    L.b = ctx->b;
    L.a.assign(vals, vals + (size_t)L.nnzb * ctx->b * ctx->b);
    L.skip.assign(skip, skip + n);
    L.vc.assign(vclass, vclass + n);
    L.vnc.assign(vnclass, vnclass + n);
    L.vt.clear();   // node vectors until ugmg_set_vtypes says otherwise
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  ctx->level_set[level] = 1;
  free_device(ctx);
  return UGMG_OK;
}

int ugmg_assemble_level(ugmg_ctx *ctx, int level, const ugmg_elements *el, double *rhs) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 0 || level >= ctx->nlevels) return fail(ctx, UGMG_E_ARG, "level out of range");
  if (!el || el->n < 0 || el->nelem < 0 || el->nc < 1 || el->nc > 64)
    return fail(ctx, UGMG_E_ARG, "bad element description");
  if (el->nelem > 0 && (!el->elem || (!el->emat && (!el->kref || !el->scale))))
    return fail(ctx, UGMG_E_ARG, "null element array");
  if ((long long)el->nelem * el->nc > 2000000000LL)
    return fail(ctx, UGMG_E_ARG, "more than 2e9 element corners");
  HIPCHK(hipSetDevice(ctx->device));
  AsmInput in{el->n, ctx->b, el->nelem, el->nc, el->elem, el->emat, el->kref, el->scale, el->evec,
              el->skip, el->dirichlet};
  std::string err;
  drop_dcsr(ctx, level);
  try {
    HostLevel L;
    AsmKeep kept;
    const int rc = assemble_device(ctx->stream, in, rhs, L.rp, L.ci, L.a, err, &kept);
    if (rc) return fail(ctx, rc, "ugmg_assemble_level: " + err);
    ctx->dcsr[level] = ugmg_ctx::DevCsr{kept.rp, kept.ci, kept.vals};
    const int n = el->n;
    L.n = n;
    L.b = ctx->b;
    L.nnzb = L.rp[n];
    L.skip.assign(n, 0u);
    L.vc.assign(n, 3);
    L.vnc.assign(n, 0);
    if (n > 0) {
      if (el->skip) HIPCHK(hipMemcpy(L.skip.data(), el->skip, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
      if (el->vclass) HIPCHK(hipMemcpy(L.vc.data(), el->vclass, n, hipMemcpyDeviceToHost));
      if (el->vnclass) HIPCHK(hipMemcpy(L.vnc.data(), el->vnclass, n, hipMemcpyDeviceToHost));
    }
    HostLevel &H = ctx->host[level];
    const bool keep_p = H.has_p && H.n == n;   // a transfer set for this level stays
    if (keep_p) {
      L.has_p = true;
      L.pp.swap(H.pp);
      L.pi.swap(H.pi);
      L.pw.swap(H.pw);
    } else {
      ctx->transfer_set[level] = 0;
    }
    H = std::move(L);
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  ctx->level_set[level] = 1;
  free_device(ctx);
  return UGMG_OK;
}

// AssembleDirichletBoundary (np/udm/disctools.cc:1837-1875) on the level's
// stored matrix and the caller's host arrays: for every Dirichlet component
// j of vector v (VECSKIP bit j), b_v,j = x_v,j, row j of the diagonal block
// the unit row, row j of every off-diagonal block zero (the columns stay)
int ugmg_assemble_dirichlet(ugmg_ctx *ctx, int level, const double *x, double *b) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 0 || level >= ctx->nlevels || !ctx->level_set[level])
    return fail(ctx, UGMG_E_ARG, "level not set");
  HostLevel &L = ctx->host[level];
  if (L.n > 0 && (!x || !b)) return fail(ctx, UGMG_E_ARG, "null array");
  const int nb = ctx->b, bb = nb * nb;
  bool changed = false;
  for (int v = 0; v < L.n; ++v) {
    const uint32_t sk = L.skip[v];
    if (!sk) continue;
    for (int j = 0; j < nb; ++j) {
      if (!((sk >> j) & 1u)) continue;
      b[(size_t)v * nb + j] = x[(size_t)v * nb + j];
      for (int m = L.rp[v]; m < L.rp[v + 1]; ++m) {
        double *blk = L.a.data() + (size_t)m * bb;
        for (int i = 0; i < nb; ++i) {
          const double want = (m == L.rp[v] && i == j) ? 1.0 : 0.0;
          if (blk[j * nb + i] != want) changed = true;
          blk[j * nb + i] = want;
        }
      }
    }
  }
  if (changed) {   // the layouts (and a device-assembled copy) hold the old rows
    drop_dcsr(ctx, level);
    free_device(ctx);
  }
  return UGMG_OK;
}

// ClearDirichletValues (np/udm/disctools.cc:1801-1817): x_v,j = 0 where bit
// j of VECSKIP is set; x a list-order device vector of a prepared level
int ugmg_clear_dirichlet(ugmg_ctx *ctx, int level, double *x) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!x) return fail(ctx, UGMG_E_ARG, "null array");
  const DevLevel &d = ctx->dev[level];
  launch_clear_dirichlet(ctx->stream, d.n, d.nb, d.perm, d.skip, x);
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_get_level(ugmg_ctx *ctx, int level, int *n, int *nnzb, int *row_ptr, int *col_idx,
                   double *vals, uint32_t *skip) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 0 || level >= ctx->nlevels || !ctx->level_set[level])
    return fail(ctx, UGMG_E_ARG, "level not set");
  const HostLevel &L = ctx->host[level];
  if (n) *n = L.n;
  if (nnzb) *nnzb = L.nnzb;
  if (row_ptr) std::copy(L.rp.begin(), L.rp.end(), row_ptr);
  if (col_idx) std::copy(L.ci.begin(), L.ci.end(), col_idx);
  if (vals) std::copy(L.a.begin(), L.a.end(), vals);
  if (skip) std::copy(L.skip.begin(), L.skip.end(), skip);
  return UGMG_OK;
}

int ugmg_set_transfer(ugmg_ctx *ctx, int level, const int *p_ptr, const int *p_idx,
                      const double *p_w) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 1 || level >= ctx->nlevels) return fail(ctx, UGMG_E_ARG, "transfer level out of range");
  if (!ctx->level_set[level] || !ctx->level_set[level - 1])
    return fail(ctx, UGMG_E_STATE, "set both levels before their transfer");
  HostLevel &L = ctx->host[level];
  const int nc = ctx->host[level - 1].n;
  if (!p_ptr || (L.n > 0 && (!p_idx || !p_w))) return fail(ctx, UGMG_E_ARG, "null array");
  try {
    L.pp.assign(p_ptr, p_ptr + L.n + 1);
    if (L.pp[0] != 0) return fail(ctx, UGMG_E_ARG, "p_ptr[0] != 0");
    for (int i = 0; i < L.n; ++i)
      if (L.pp[i + 1] < L.pp[i]) return fail(ctx, UGMG_E_ARG, "p_ptr not monotone");
    const int nnz = L.pp[L.n];
    L.pi.assign(p_idx, p_idx + nnz);
    for (int q = 0; q < nnz; ++q)
      if (L.pi[q] < 0 || L.pi[q] >= nc) return fail(ctx, UGMG_E_ARG, "coarse index out of range");
    L.pw.assign(p_w, p_w + nnz);
    L.pblk.clear();
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  L.has_p = true;
  ctx->transfer_set[level] = 1;
  free_device(ctx);
  return UGMG_OK;
}

int ugmg_set_transfer_matrix(ugmg_ctx *ctx, int level, const int *p_ptr, const int *p_idx,
                             const double *p_blocks) {
  if (!ctx) return UGMG_E_ARG;
  if (level >= 1 && level < ctx->nlevels && ctx->host[level].n > 0 && !p_blocks)
    return fail(ctx, UGMG_E_ARG, "null array");
  // the pattern as ugmg_set_transfer checks it, with unit weights
  std::vector<double> ones;
  try {
    if (level >= 1 && level < ctx->nlevels && p_ptr && ctx->host[level].n >= 0)
      ones.assign(std::max(0, p_ptr[ctx->host[level].n]), 1.0);
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  int rc = ugmg_set_transfer(ctx, level, p_ptr, p_idx, ones.data());
  if (rc) return rc;
  HostLevel &L = ctx->host[level];
  const size_t bb = (size_t)ctx->b * ctx->b;
  try {
    L.pblk.assign(p_blocks, p_blocks + L.pi.size() * bb);
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  return UGMG_OK;
}

int ugmg_set_vtypes(ugmg_ctx *ctx, int level, const uint8_t *vtype) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 0 || level >= ctx->nlevels) return fail(ctx, UGMG_E_ARG, "level out of range");
  if (!ctx->level_set[level]) return fail(ctx, UGMG_E_STATE, "set the level before its vector types");
  HostLevel &L = ctx->host[level];
  try {
    if (!vtype) {
      L.vt.clear();
    } else {
      for (int v = 0; v < L.n; ++v)
        if (vtype[v] > 1) return fail(ctx, UGMG_E_ARG, "vector type must be 0 (node) or 1 (edge)");
      L.vt.assign(vtype, vtype + L.n);
    }
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  free_device(ctx);
  return UGMG_OK;
}

int ugmg_set_ident(ugmg_ctx *ctx, const int *ident, int n) {
  if (!ctx) return UGMG_E_ARG;
  if (!ident || n <= 0) {
    ctx->nident = 0;
    drop_graph(ctx);
    return UGMG_OK;
  }
  if (n > UGMG_MAXB || n > 16) return fail(ctx, UGMG_E_ARG, "more VD_IDENT entries than UGMG_MAXB");
  for (int i = 0; i < n; ++i)
    if (ident[i] < 0 || ident[i] >= n || ident[ident[i]] != ident[i])
      return fail(ctx, UGMG_E_ARG, "VD_IDENT must map each entry to a group leader that maps to itself");
  for (int i = 0; i < n; ++i) ctx->ident[i] = ident[i];
  ctx->nident = n;
  drop_graph(ctx);   // the base solve's stop test is captured with the grouping
  return UGMG_OK;
}

int ugmg_set_params(ugmg_ctx *ctx, const ugmg_params *p) {
  if (!ctx || !p) return UGMG_E_ARG;
  if (p->nu1 < 0 || p->nu2 < 0 || p->gamma < 0 || p->base_maxiter < 0 ||
      (p->smoother != UGMG_SMOOTH_JAC && p->smoother != UGMG_SMOOTH_GS) ||
      (p->post_set && p->post_smoother != UGMG_SMOOTH_JAC && p->post_smoother != UGMG_SMOOTH_GS) ||
      (p->base_iter != UGMG_BASE_EX && p->base_iter != UGMG_BASE_JAC) ||
      (p->precision != UGMG_FP64 && p->precision != UGMG_MIXED) || p->inner_cycles < 0)
    return fail(ctx, UGMG_E_ARG, "invalid parameter");
  // the damping factors may differ per component ($damp, lmgc $damp, ex
  // $damp: dscalx per component); the base solve's stop limits may not
  if (ctx->b > 1)
    for (int i = 1; i < ctx->b; ++i)
      if (p->base_abslimit[i] != p->base_abslimit[0] || p->base_red[i] != p->base_red[0])
        return fail(ctx, UGMG_E_ARG, "block hierarchies take one base $abslimit / $red for all components");
  const bool relayout = !ctx->params_set || p->smoother != ctx->p.smoother ||
                        p->post_set != ctx->p.post_set || p->post_smoother != ctx->p.post_smoother ||
                        p->baselevel != ctx->p.baselevel || p->ex_reorder != ctx->p.ex_reorder ||
                        p->ex_float != ctx->p.ex_float;
  ctx->p = *p;
  ctx->params_set = true;
  drop_graph(ctx);
  if (relayout && ctx->prepared) free_device(ctx);
  return UGMG_OK;
}

int ugmg_preprocess(ugmg_ctx *ctx, int level) {
  if (!ctx) return UGMG_E_ARG;
  PrepState ps;
  int rc = prep_host(ctx, level, ps);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  return prep_device(ctx, ps);
}

int ugmg_postprocess(ugmg_ctx *ctx) {
  if (!ctx) return UGMG_E_ARG;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_device(ctx);
  return UGMG_OK;
}

int ugmg_block(const ugmg_ctx *ctx) { return ctx && ctx->nlevels > 0 ? ctx->b : UGMG_E_STATE; }

int ugmg_norm_entries(const ugmg_ctx *ctx) {
  if (!ctx) return UGMG_E_ARG;
  return ctx->prepared ? ctx->nc : UGMG_E_STATE;
}

int ugmg_level_size(const ugmg_ctx *ctx, int level, int *n, int *nnzb) {
  if (!ctx || level < 0 || level >= ctx->nlevels || !ctx->level_set[level]) return UGMG_E_ARG;
  if (n) *n = ctx->host[level].n;
  if (nnzb) *nnzb = ctx->host[level].nnzb;
  return UGMG_OK;
}

int ugmg_num_colors(const ugmg_ctx *ctx, int level) {
  if (!ctx || !ctx->prepared || level < 0 || level > ctx->top) return UGMG_E_ARG;
  return ctx->dev[level].ncolors;
}

int ugmg_get_colors(ugmg_ctx *ctx, int level, int *colors) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  const DevLevel &d = ctx->dev[level];
  if (d.color.empty()) return fail(ctx, UGMG_E_STATE, "level has no colouring (not a GS level)");
  std::copy(d.color.begin(), d.color.end(), colors);
  return UGMG_OK;
}

int ugmg_get_pattern(ugmg_ctx *ctx, int level, int *row_ptr, int *col_idx, double *vals) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  const DevLevel &d = ctx->dev[level];
  const int ns = d.A.nslices;
  const int bb = d.nb * d.nb;
  std::vector<int> soff(ns + 1), rows((size_t)ns * kSlice), col(d.A.nent), sdoff(std::max(ns, 1));
  std::vector<double> val(d.A.nent * bb);
  HIPCHK(hipMemcpy(soff.data(), d.A.soff, soff.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rows.data(), d.A.rows, rows.size() * sizeof(int), hipMemcpyDeviceToHost));
  if (ns) HIPCHK(hipMemcpy(sdoff.data(), d.A.sdoff, ns * sizeof(int), hipMemcpyDeviceToHost));
  if (d.A.nent) {
    HIPCHK(hipMemcpy(col.data(), d.A.col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(val.data(), d.A.v64, val.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  // explicit slices: the host copy's row lengths separate entries from slice
  // padding; uniform slices: phantom entries carry column -1.  Device rows
  // and columns map back to list indices through the device order.
  const HostLevel &H = ctx->host[level];
  const bool pm = !d.perm_h.empty();
  row_ptr[0] = 0;
  for (int s = 0; s < ns; ++s) {
    const int w = (soff[s + 1] - soff[s]) / kSlice;
    for (int lane = 0; lane < kSlice; ++lane) {
      const int rd = rows[(size_t)s * kSlice + lane];
      if (rd < 0) continue;
      const int r = pm ? d.perm_h[rd] : rd;
      const int len = H.rp[r + 1] - H.rp[r];
      int q = 0;
      for (int k = 0; k < w && q < len; ++k) {
        const size_t e = (size_t)soff[s] + (size_t)k * kSlice + lane;
        if (sdoff[s] >= 0 && col[e] < 0) continue;
        col_idx[H.rp[r] + q] = pm ? d.perm_h[col[e]] : col[e];
        if (vals)
          for (int k2 = 0; k2 < bb; ++k2)
            vals[(size_t)(H.rp[r] + q) * bb + k2] = val[(size_t)k2 * d.A.nent + e];
        ++q;
      }
      row_ptr[r + 1] = q;
    }
  }
  for (int i = 0; i < d.n; ++i) row_ptr[i + 1] += row_ptr[i];
  return UGMG_OK;
}

// On a level kept in device order the single-level entry points move their
// list-order arguments through the level's own vectors (c, b, t), which are
// scratch outside a cycle.
int ugmg_dmatmul_minus(ugmg_ctx *ctx, int level, double *x, const double *y) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  DevLevel &d = ctx->dev[level];
  if (!d.perm) {
    launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, y, x, nullptr, 0, d.nt, B1(d));
  } else {
    to_dev<double, double>(ctx, level, y, d.vs, d.t);
    to_dev<double, double>(ctx, level, x, d.vs, d.b);
    launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, d.t, d.b, nullptr, 0, d.nt, B1(d));
    to_user<double, double>(ctx, level, d.b, x, d.vs);
  }
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_dmatmul(ugmg_ctx *ctx, int level, double *x, const double *y) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  DevLevel &d = ctx->dev[level];
  if (!d.perm) {
    launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, y, x, nullptr, 1, d.nt, B1(d));
  } else {
    to_dev<double, double>(ctx, level, y, d.vs, d.t);
    launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, d.t, d.b, nullptr, 1, d.nt, B1(d));
    to_user<double, double>(ctx, level, d.b, x, d.vs);
  }
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_smatmul_minus(ugmg_ctx *ctx, int level, float *x, const float *y) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  DevLevel &d = ctx->dev[level];
  if (!d.perm) {
    launch_spmv<float, float, float>(ctx->stream, d.A.d32(), 0, d.A.nslices, y, x, nullptr, 0, d.nt32, B1(d));
  } else {
    to_dev<float, float>(ctx, level, y, d.vs, d.t32);
    to_dev<float, float>(ctx, level, x, d.vs, d.b32);
    launch_spmv<float, float, float>(ctx->stream, d.A.d32(), 0, d.A.nslices, d.t32, d.b32, nullptr, 0, d.nt32, B1(d));
    to_user<float, float>(ctx, level, d.b32, x, d.vs);
  }
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

// per-type sums read each row's VTYPE, kept in the level's device order: a
// list-order argument of such a level moves to that order first (into the
// level's scratch vector `tmp`)
const double *typed_order(ugmg_ctx *ctx, int level, const double *x, double *tmp) {
  const DevLevel &d = ctx->dev[level];
  if (!d.vt || !d.perm) return x;
  to_dev<double, double>(ctx, level, x, d.vs, tmp);
  return tmp;
}

int ugmg_dnrm2x(ugmg_ctx *ctx, int level, const double *x, double *out) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!out) return fail(ctx, UGMG_E_ARG, "null output");
  DevLevel &d = ctx->dev[level];
  x = typed_order(ctx, level, x, d.t);
  launch_dot<double>(ctx->stream, d.n, x, nullptr, nullptr, ctx->part, ctx->red, 1, B1(d));
  HIPCHK(hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double) * d.nc(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < d.nc(); ++i) out[i] = ctx->red_host[i];
  return UGMG_OK;
}

// ugmg_dnrm2x with the partial sums and the norms in the static scratch
// instead of the context's reduction buffers: the same kernels over the
// same partition, so the result is bit-identical; no allocation
int ugmg_dnrm2x_noalloc(ugmg_ctx *ctx, int level, const double *x, double *out) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!x || !out) return fail(ctx, UGMG_E_ARG, "null argument");
  const int dev = ctx->device;
  if (dev < 0 || dev >= kScratchMaxDevices || !g_scratch_done[dev])
    return fail(ctx, UGMG_E_STATE, "no scratch on this device");
  if (!g_scratch_ptr[dev])
    return fail(ctx, UGMG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(g_scratch_err[dev]));
  DevLevel &d = ctx->dev[level];
  const long long need = (long long)reduce_parts(d.n) * d.nc();
  if (need > kScratchParts || d.nc() > UGMG_MAXB)
    return fail(ctx, UGMG_E_ARG, "ugmg_dnrm2x_noalloc: " + std::to_string(need) +
                                     " partial sums exceed the scratch of " +
                                     std::to_string(kScratchParts));
  double *part = g_scratch_ptr[dev], *norms = g_scratch_ptr[dev] + kScratchParts;
  x = typed_order(ctx, level, x, d.t);
  launch_dot<double>(ctx->stream, d.n, x, nullptr, nullptr, part, norms, 1, B1(d));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ctx->red_host, norms, sizeof(double) * d.nc(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < d.nc(); ++i) out[i] = ctx->red_host[i];
  return UGMG_OK;
}

int ugmg_ddot(ugmg_ctx *ctx, int level, const double *x, const double *y, double *out) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!out || !y) return fail(ctx, UGMG_E_ARG, "null argument");
  DevLevel &d = ctx->dev[level];
  x = typed_order(ctx, level, x, d.t);
  y = typed_order(ctx, level, y, d.u);
  launch_dot<double>(ctx->stream, d.n, x, y, nullptr, ctx->part, ctx->red, 0, B1(d));
  HIPCHK(hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double) * d.nc(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < d.nc(); ++i) out[i] = ctx->red_host[i];
  return UGMG_OK;
}

int ugmg_jacobi_step(ugmg_ctx *ctx, int level, double *x, double *b) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  DevLevel &d = ctx->dev[level];
  double *xd = d.perm ? d.t : x, *bd = d.perm ? d.b : b;
  if (d.perm) to_dev<double, double>(ctx, level, b, d.vs, bd);
  if (per_comp(ctx, ctx->p.damp)) {
    launch_jac<double>(ctx->stream, d.n, bd, d.diag64, d.vclass, 1.0, xd, B1(d));
    launch_scale_comp<double>(ctx->stream, d.n, d.nb, ctx->p.damp, xd, nullptr, 0, B1(d));
  } else {
    launch_jac<double>(ctx->stream, d.n, bd, d.diag64, d.vclass, ctx->p.damp[0], xd, B1(d));
  }
  launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, xd, bd, nullptr, 0, d.nt, B1(d));
  if (d.perm) {
    to_user<double, double>(ctx, level, xd, x, d.vs);
    to_user<double, double>(ctx, level, bd, b, d.vs);
  }
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_gs_step(ugmg_ctx *ctx, int level, double *x, double *b) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  DevLevel &d = ctx->dev[level];
  if (d.ncolors == 0) return fail(ctx, UGMG_E_STATE, "level was not preprocessed for GS");
  double *xd = d.perm ? d.c : x, *bd = d.perm ? d.b : b;
  if (d.perm) to_dev<double, double>(ctx, level, b, d.vs, bd);
  const int saved = ctx->p.smoother;
  ctx->p.smoother = UGMG_SMOOTH_GS;
  smooth_step<double>(ctx, level, d.t, bd, nullptr, -1, xd);
  ctx->p.smoother = saved;
  if (d.perm) {
    to_user<double, double>(ctx, level, xd, x, d.vs);
    to_user<double, double>(ctx, level, bd, b, d.vs);
  }
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_restrict(ugmg_ctx *ctx, int level, double *to_c, const double *from_f,
                  const double *damp) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (level <= ctx->bl) return fail(ctx, UGMG_E_ARG, "restrict needs a level above the base level");
  DevLevel &d = ctx->dev[level], &dc = ctx->dev[level - 1];
  const double *fd = from_f;
  double *cd = to_c;
  if (d.perm) {
    to_dev<double, double>(ctx, level, from_f, d.vs, d.b);
    fd = d.b;
  }
  if (dc.perm) {
    to_dev<double, double>(ctx, level - 1, to_c, dc.vs, dc.b);
    cd = dc.b;
  }
  if (damp && per_comp_entries(ctx, damp)) {
    if (ctx->ntypes > 1) return fail(ctx, UGMG_E_ARG, "several vector types take one damping factor");
    if (d.PT.d64().bb > 1) {
      // RestrictByMatrix: the coarse rows of VNCLASS >= 2 scaled after the sum
      launch_restrict<double>(ctx->stream, d.PT.d64(), dc.skip, dc.vnclass, dc.vclass, fd, 1.0, cd,
                              nullptr, d.nt, 1, d.vs, dc.vs, d.nb);
      launch_scale_comp<double>(ctx->stream, dc.n, dc.nb, damp, cd, nullptr, 0, B1(dc), dc.vnclass, 2);
    } else {
      // the standard restriction damps each fine term: scale b_f first
      if (fd != d.b) HIPCHK(hipMemcpyAsync(d.t, fd, sizeof(double) * d.n * d.nb, hipMemcpyDeviceToDevice,
                                           ctx->stream));
      double *sf = fd != d.b ? d.t : d.b;
      launch_scale_comp<double>(ctx->stream, d.n, d.nb, damp, sf, nullptr, 0, B1(d));
      launch_restrict<double>(ctx->stream, d.PT.d64(), dc.skip, dc.vnclass, dc.vclass, sf, 1.0, cd,
                              nullptr, d.nt, 1, d.vs, dc.vs, d.nb);
    }
  } else {
    launch_restrict<double>(ctx->stream, d.PT.d64(), dc.skip, dc.vnclass, dc.vclass, fd,
                            damp ? damp[0] : 1.0, cd, nullptr, d.nt, 1, d.vs, dc.vs, d.nb);
  }
  if (dc.perm) to_user<double, double>(ctx, level - 1, cd, to_c, dc.vs);
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_interpolate(ugmg_ctx *ctx, int level, double *to_f, const double *from_c,
                     const double *damp) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (level <= ctx->bl) return fail(ctx, UGMG_E_ARG, "interpolate needs a level above the base level");
  DevLevel &d = ctx->dev[level], &dc = ctx->dev[level - 1];
  const double *cd = from_c;
  double *fd = d.perm ? d.t : to_f;
  if (dc.perm) {
    to_dev<double, double>(ctx, level - 1, from_c, dc.vs, dc.c);
    cd = dc.c;
  }
  if (damp && per_comp_entries(ctx, damp)) {   // dscalx after the sum
    if (ctx->ntypes > 1) return fail(ctx, UGMG_E_ARG, "several vector types take one damping factor");
    launch_interp<double>(ctx->stream, d.P.d64(), d.iskip, cd, 1.0, fd, nullptr, d.nt, 1, d.vs, dc.vs, d.nb);
    launch_scale_comp<double>(ctx->stream, d.n, d.nb, damp, fd, nullptr, 0, B1(d));
  } else {
    launch_interp<double>(ctx->stream, d.P.d64(), d.iskip, cd, damp ? damp[0] : 1.0, fd,
                          nullptr, d.nt, 1, d.vs, dc.vs, d.nb);
  }
  if (d.perm) to_user<double, double>(ctx, level, fd, to_f, d.vs);
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_base_solve(ugmg_ctx *ctx, double *c, double *b) {
  int rc = check_ready(ctx, ctx ? ctx->bl : 0);
  if (rc) return rc;
  if (base_singular(ctx)) return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  base_solve<double>(ctx, c, b, 0, false, true);
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_iter(ugmg_ctx *ctx, int level, double *c, double *b) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (base_singular(ctx)) return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  DevLevel &d = ctx->dev[level];
  if (ctx->p.precision == UGMG_MIXED) {
    to_dev<float, double>(ctx, level, b, d.vs, d.b32);
    const int k = std::max(1, ctx->p.inner_cycles);
    for (int j = 0; j < k; ++j) lmgc<float>(ctx, level, d.c32, d.b32, j == 0 ? 0 : 1);
    if (!d.perm) {
      launch_convert<double, float>(ctx->stream, d.n * d.nb, d.c32, c, 1, d.vs, d.vs);
      launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, c, b, nullptr, 0, d.nt, B1(d));
    } else {
      launch_convert<double, float>(ctx->stream, d.n * d.nb, d.c32, d.c, 1, d.vs, d.vs);
      to_dev<double, double>(ctx, level, b, d.vs, d.b);
      launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, d.c, d.b, nullptr, 0, d.nt, B1(d));
      to_user<double, double>(ctx, level, d.c, c, d.vs);
      to_user<double, double>(ctx, level, d.b, b, d.vs);
    }
  } else if (!d.perm) {
    lmgc<double>(ctx, level, c, b, 0);
  } else {
    to_dev<double, double>(ctx, level, b, d.vs, d.b);
    lmgc<double>(ctx, level, d.c, d.b, 0);
    to_user<double, double>(ctx, level, d.c, c, d.vs);
    to_user<double, double>(ctx, level, d.b, b, d.vs);
  }
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_iter_host(ugmg_ctx *ctx, int level, double *c, double *b) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  const size_t bytes = sizeof(double) * ctx->dev[level].n * ctx->b;
  double *dc = nullptr, *db = nullptr;
  HIPCHK(hipMalloc(&dc, bytes));
  if (hipMalloc(&db, bytes) != hipSuccess) {
    hipFree(dc);
    return fail(ctx, UGMG_E_NOMEM, "hipMalloc");
  }
  rc = UGMG_OK;
  if (hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(ctx, UGMG_E_HIP, "copy in");
  if (!rc) rc = ugmg_iter(ctx, level, dc, db);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, UGMG_E_HIP, "sync");
  if (!rc && (hipMemcpy(c, dc, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(b, db, bytes, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(ctx, UGMG_E_HIP, "copy out");
  hipFree(dc);
  hipFree(db);
  return rc;
}

int ugmg_solve(ugmg_ctx *ctx, int level, double *x, double *b, const double *abslimit,
               const double *reduction, int maxiter, int do_defect, ugmg_lresult *res,
               ugmg_defect_cb cb, void *user) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!x || !b || !abslimit || !reduction || !res || maxiter < 0)
    return fail(ctx, UGMG_E_ARG, "null argument");
  std::memset(res, 0, sizeof(*res));
  if (base_singular(ctx)) {
    res->error_code = UGMG_E_SINGULAR;
    return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  }
  DevLevel &d = ctx->dev[level];
  ctx->stat_cycles = 0;
  ctx->stat_bytes = 0.0;
  const double vb = vbytes(d, 8);
  // a level in device order iterates on its own c (x) and b, moved in and
  // out once per solve
  double *const xu = x, *const bu = b;
  if (d.perm) {
    to_dev<double, double>(ctx, level, xu, d.vs, d.c);
    to_dev<double, double>(ctx, level, bu, d.vs, d.b);
    x = d.c;
    b = d.b;
    ctx->stat_bytes += 2.0 * (2.0 * vb + 4.0 * d.n);
  }
  auto move_out = [&]() -> int {
    if (!d.perm) return 0;
    to_user<double, double>(ctx, level, x, xu, d.vs);
    to_user<double, double>(ctx, level, b, bu, d.vs);
    ctx->stat_bytes += 2.0 * (2.0 * vb + 4.0 * d.n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
  };
  if (do_defect) {  // LinearDefect: b -= A x
    launch_spmv<double, double, double>(ctx->stream, d.A.d64(), 0, d.A.nslices, x, b, nullptr, 0, d.nt, B1(d));
    ctx->stat_bytes += op_bytes(ctx, level, Op::Spmv, 8);
  }
  ctx->stat_bytes += vb + d.n;   // the first defect norm
  const int nc = d.nc();
  if (ctx->nident > 0 && ctx->nident != nc)
    return fail(ctx, UGMG_E_ARG, "ugmg_set_ident gave " + std::to_string(ctx->nident) + " entries, the norms have " +
                                     std::to_string(nc));
  // sc_cmp (npscan.cc:1027-1057): every component (or VD_IDENT group) below
  // its limit
  auto sc_cmp = [ctx, nc](const double *x, const double *y) { return sc_cmp_vd(ctx, nc, x, y); };
  double first[UGMG_MAXB] = {}, target[UGMG_MAXB], last[UGMG_MAXB];
  if ((rc = surface_norm(ctx, level, b, first))) return rc;
  for (int i = 0; i < nc; ++i) {
    res->first_defect[i] = res->last_defect[i] = first[i];
    target[i] = first[i] * reduction[i];  // sc_mul_check
    if (target[i] == 0.0) target[i] = reduction[i];
  }
  if (cb) cb(user, 0, nc, first);
  int conv = sc_cmp(first, abslimit);
  if (ctx->p.precision == UGMG_MIXED) {
    launch_convert<float, double>(ctx->stream, d.n * d.nb, b, d.b32, 1, d.vs, d.vs);
    ctx->stat_bytes += vbytes(d, 12);
  }
  // mixed steps: inner_cycles FP32 cycles, fewer when the per-cycle rate of
  // the previous step (the slowest component's) predicts that fewer bring
  // every component below 0.9 of the target (or of abslimit)
  const int kin = std::max(1, ctx->p.inner_cycles);
  const bool adapt = ctx->p.precision == UGMG_MIXED && ctx->p.inner_adapt && kin > 1;
  double prev[UGMG_MAXB] = {};
  int prev_cycles = 0;
  for (int k = 0; k < nc; ++k) last[k] = first[k];
  auto cycles_for = [&](const double *lim) {
    int need = 1;
    for (int q = 0; q < nc; ++q) {
      const double goal = 0.9 * std::fabs(lim[q]);
      if (last[q] <= goal) continue;
      if (!(prev[q] > 0.0) || !(last[q] > 0.0) || !(goal > 0.0)) return kin;
      const double rho = std::pow(last[q] / prev[q], 1.0 / prev_cycles);
      if (!(rho < 1.0) || !(rho > 0.0)) return kin;
      const double m = std::ceil(std::log(goal / last[q]) / std::log(rho));
      if (!(m < kin)) return kin;
      need = std::max(need, (int)m);
    }
    return need;
  };
  int i;
  for (i = 0; i < maxiter; ++i) {
    if (conv) break;
    int cycles = kin;
    if (adapt && prev_cycles > 0) cycles = std::min(cycles_for(target), cycles_for(abslimit));
    for (int k = 0; k < nc; ++k) prev[k] = last[k];
    prev_cycles = cycles;
    if ((rc = run_iteration(ctx, level, x, b, cycles))) {
      res->error_code = rc;
      move_out();
      return rc;
    }
    const int ncyc = ctx->p.precision == UGMG_MIXED ? std::max(1, cycles) : 1;
    ctx->stat_cycles += ncyc;
    ctx->stat_bytes += iteration_bytes(ctx, level, ncyc);
    for (int k = 0; k < nc; ++k) res->last_defect[k] = last[k] = ctx->red_host[k];
    if (cb) cb(user, i + 1, nc, last);
    if (sc_cmp(last, abslimit) || sc_cmp(last, target)) {
      conv = 1;
      res->number_of_linear_iterations = i + 1;
      break;
    }
  }
  if (!conv) res->number_of_linear_iterations = i;
  res->converged = conv;
  return move_out();
}

int ugmg_solve_stats(const ugmg_ctx *ctx, long long *cycles, double *bytes) {
  if (!ctx) return UGMG_E_ARG;
  if (cycles) *cycles = ctx->stat_cycles;
  if (bytes) *bytes = ctx->stat_bytes;
  return UGMG_OK;
}

int ugmg_base_level(const ugmg_ctx *ctx) {
  if (!ctx) return UGMG_E_ARG;
  return ctx->prepared ? ctx->bl : UGMG_E_STATE;
}

int ugmg_fullrefinelevel(const ugmg_ctx *ctx) {
  if (!ctx) return UGMG_E_ARG;
  return ctx->prepared ? ctx->frl : UGMG_E_STATE;
}

// LinearSolver (ls.cc:637-749) on a locally refined hierarchy.  The surface
// runs from FULLREFINELEVEL: LinearDefect and LinearResiduum take the
// FINE_GRID_DOF rows of the levels below `level` and the NEW_DEFECT rows of
// `level` (the T_SURFACE_MODE loops of matloop.ct / vecloop.ct, whatever
// their from-level); the coarse levels' b carry the defect recurrence from
// iteration to iteration (lmgc zeroes only what the restriction covers);
// LSUpdate adds each level's correction into x on baselevel..level.
//
// Mixed form: an iteration runs `cycles` FP32 cycles on the FP32 copies of
// every level's b, sums each coarse level's FP32 corrections, then per
// level in FP64: b -= A dx, x += dx, b32 = b (the refinement update of the
// plain solve, on every level of the recurrence), with the surface norms.

namespace {

struct SurfaceRun {
  int level = 0, bl = 0, frl = 0;
  std::vector<double *> X, B;   // device-order x and FP64 b per level
};

inline const uint8_t *surface_cls(const ugmg_ctx *ctx, const SurfaceRun &s, int l) {
  return l < s.level ? ctx->dev[l].fgd : ctx->dev[l].vclass;
}

// one iteration, enqueued on the stream; leaves each surface level's
// per-component squared sums in surf_sq[l * UGMG_MAXB ..]
void enqueue_surface_iteration(ugmg_ctx *ctx, const SurfaceRun &s, int cycles) {
  hipStream_t st = ctx->stream;
  const int level = s.level;
  if (ctx->p.precision == UGMG_MIXED) {
    DevLevel &dt = ctx->dev[level];
    const int k = std::max(1, cycles);
    for (int j = 0; j < k; ++j) {
      lmgc<float>(ctx, level, dt.c32, dt.b32, j == 0 ? 0 : 1);
      for (int l = s.bl; l < level; ++l) {
        DevLevel &d = ctx->dev[l];
        if (j == 0) launch_convert<float, float>(st, d.n * d.nb, d.c32, d.acc32, 1, d.vs, d.vs);
        else launch_add<float>(st, d.n, d.acc32, d.c32, B1(d));
      }
    }
    for (int l = s.bl; l <= level; ++l) {
      DevLevel &d = ctx->dev[l];
      int np = 0;
      launch_ir_update(st, d.A.d64(), l == level ? d.c32 : d.acc32, s.B[l], s.X[l], d.b32,
                       l >= s.frl ? surface_cls(ctx, s, l) : d.vclass, ctx->part, &np, d.nt, B1(d));
      if (l >= s.frl) launch_reduce_final(st, np, ctx->part, ctx->surf_sq + (size_t)l * UGMG_MAXB, 0, d.nb);
    }
    return;
  }
  // c accumulates straight into the top level's x; the coarse levels'
  // corrections stay in their c and are added after the cycle (LSUpdate)
  lmgc<double>(ctx, level, s.X[level], s.B[level], 1);
  for (int l = s.bl; l < level; ++l) launch_add<double>(st, ctx->dev[l].n, s.X[l], ctx->dev[l].c, B1(ctx->dev[l]));
  for (int l = s.frl; l <= level; ++l) {
    DevLevel &d = ctx->dev[l];
    launch_dot<double>(st, d.n, s.B[l], nullptr, surface_cls(ctx, s, l), ctx->part,
                       ctx->surf_sq + (size_t)l * UGMG_MAXB, 0, B1(d));
  }
}

// squared sums -> per-component norms, summed over the surface levels in
// level order
void surface_norms(const ugmg_ctx *ctx, const SurfaceRun &s, const double *sq, double *out) {
  const int nc = ctx->dev[s.level].nb;
  double acc[UGMG_MAXB] = {};
  for (int l = s.frl; l <= s.level; ++l)
    for (int i = 0; i < nc; ++i) acc[i] += sq[(size_t)l * UGMG_MAXB + i];
  for (int i = 0; i < nc; ++i) out[i] = std::sqrt(acc[i]);
}

// enqueue one iteration (from a cached graph when enabled) and the copy of
// the squared sums to the pinned surf_sq_host
int launch_surface_iteration(ugmg_ctx *ctx, const SurfaceRun &s, int cycles) {
  const bool mixed = ctx->p.precision == UGMG_MIXED;
  cycles = mixed ? std::max(1, cycles) : 1;
  if (!can_capture(ctx) || cycles >= ugmg_ctx::kMaxGraphs) {
    enqueue_surface_iteration(ctx, s, cycles);
    HIPCHK(hipGetLastError());
  } else {
    const int kind = ctx->p.precision;
    if (ctx->sgkind != kind || ctx->sglevel != s.level || ctx->sgx != s.X[s.level] ||
        ctx->sgb != s.B[s.level]) {
      for (hipGraphExec_t &g : ctx->sgexec)
        if (g) {
          hipGraphExecDestroy(g);
          g = nullptr;
        }
      ctx->sgkind = kind;
      ctx->sglevel = s.level;
      ctx->sgx = s.X[s.level];
      ctx->sgb = s.B[s.level];
    }
    hipGraphExec_t &ge = ctx->sgexec[cycles];
    if (!ge) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      enqueue_surface_iteration(ctx, s, cycles);
      HIPCHK(hipStreamEndCapture(ctx->stream, &g));
      hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        ge = nullptr;
        return fail(ctx, UGMG_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    HIPCHK(hipGraphLaunch(ge, ctx->stream));
  }
  HIPCHK(hipMemcpyAsync(ctx->surf_sq_host, ctx->surf_sq, sizeof(double) * UGMG_MAXB * (s.level + 1),
                        hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

// algorithmic bytes of one surface iteration (the byte model of the plain
// solve, plus the per-level LSUpdate / refinement updates and norms)
double surface_iteration_bytes(const ugmg_ctx *ctx, const SurfaceRun &s, int cycles) {
  double by = 0.0;
  if (ctx->p.precision == UGMG_MIXED) {
    for (int j = 0; j < cycles; ++j) {
      by += lmgc_bytes(ctx, s.level, 4, j == 0 ? 0 : 1, false);
      for (int l = s.bl; l < s.level; ++l) by += vbytes(ctx->dev[l], 4) * (j == 0 ? 2.0 : 3.0);
    }
    for (int l = s.bl; l <= s.level; ++l) {
      const DevLevel &d = ctx->dev[l];
      by += d.A.bytes(8) + (double)d.n * d.nb * (4.0 + 4.0 + 16.0 + 16.0 + 4.0) + d.n;
    }
    return by;
  }
  by = lmgc_bytes(ctx, s.level, 8, 1, false);
  for (int l = s.bl; l < s.level; ++l) by += 3.0 * vbytes(ctx->dev[l], 8);
  for (int l = s.frl; l <= s.level; ++l) by += vbytes(ctx->dev[l], 8) + ctx->dev[l].n;
  return by;
}

// FP32 cycles of a mixed step: inner_cycles, fewer when the per-cycle rate
// of the previous step (the slowest component's) predicts that fewer bring
// every component below 0.9 of `lim`
int adapt_cycles(int kin, int nc, const double *last, const double *prev, int prev_cycles,
                 const double *lim) {
  int need = 1;
  for (int q = 0; q < nc; ++q) {
    const double goal = 0.9 * std::fabs(lim[q]);
    if (last[q] <= goal) continue;
    if (!(prev[q] > 0.0) || !(last[q] > 0.0) || !(goal > 0.0)) return kin;
    const double rho = std::pow(last[q] / prev[q], 1.0 / prev_cycles);
    if (!(rho < 1.0) || !(rho > 0.0)) return kin;
    const double m = std::ceil(std::log(goal / last[q]) / std::log(rho));
    if (!(m < kin)) return kin;
    need = std::max(need, (int)m);
  }
  return need;
}

}  // namespace

int ugmg_solve_surface(ugmg_ctx *ctx, int level, double *const *x, double *const *b,
                       const double *abslimit, const double *reduction, int maxiter, int do_defect,
                       ugmg_lresult *res, ugmg_defect_cb cb, void *user) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!x || !b || !abslimit || !reduction || !res || maxiter < 0)
    return fail(ctx, UGMG_E_ARG, "null argument");
  if (ctx->ntypes > 1) return fail(ctx, UGMG_E_ARG, "the surface solve takes one vector type");
  std::memset(res, 0, sizeof(*res));
  // the level's own NEW_DEFECT loop always runs, the FINE_GRID_DOF loop
  // from FULLREFINELEVEL up to the level below it
  SurfaceRun s;
  s.level = level;
  s.bl = ctx->bl;
  s.frl = std::min(ctx->frl, level);
  const int bl = s.bl;
  if (s.frl < bl) return fail(ctx, UGMG_E_ARG, "FULLREFINELEVEL lies below the base level");
  for (int l = bl; l <= level; ++l)
    if (!x[l] || !b[l]) return fail(ctx, UGMG_E_ARG, "x[l] and b[l] are needed on baselevel..level");
  if (base_singular(ctx)) {
    res->error_code = UGMG_E_SINGULAR;
    return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  }
  const bool mixed = ctx->p.precision == UGMG_MIXED;
  hipStream_t st = ctx->stream;
  // surface work arrays, allocated on first use and kept with the hierarchy
  if (!ctx->surf_sq && (rc = dalloc(ctx, &ctx->surf_sq, (size_t)UGMG_MAXB * (ctx->top + 1)))) return rc;
  if (ctx->surf_sq_host_cap < UGMG_MAXB * (ctx->top + 1)) {
    if (ctx->surf_sq_host) hipHostFree(ctx->surf_sq_host);
    ctx->surf_sq_host = nullptr;
    ctx->surf_sq_host_cap = 0;
    if (hipHostMalloc((void **)&ctx->surf_sq_host, sizeof(double) * UGMG_MAXB * (ctx->top + 1),
                      hipHostMallocDefault) != hipSuccess)
      return fail(ctx, UGMG_E_NOMEM, "hipHostMalloc");
    ctx->surf_sq_host_cap = UGMG_MAXB * (ctx->top + 1);
  }
  for (int l = bl; l <= level; ++l) {
    DevLevel &d = ctx->dev[l];
    if (!d.fgd) {
      if ((rc = dalloc(ctx, &d.fgd, (size_t)d.n))) return rc;
      launch_fgd_class(st, d.n, d.vclass, d.vnclass, d.fgd);
    }
    if (l < level && !d.xs && (rc = dalloc(ctx, &d.xs, d.vs))) return rc;
    if (mixed && l < level && !d.acc32 && (rc = dalloc(ctx, &d.acc32, d.vs))) return rc;
  }
  // device-order working copies: the top level iterates on its c (x) and b
  // when it is permuted, a coarse level on xs and its own b (the cycle's)
  s.X.assign(level + 1, nullptr);
  s.B.assign(level + 1, nullptr);
  for (int l = bl; l <= level; ++l) {
    DevLevel &d = ctx->dev[l];
    if (l == level && !d.perm) {
      s.X[l] = x[l];
      s.B[l] = b[l];
      continue;
    }
    s.X[l] = l == level ? d.c : d.xs;
    s.B[l] = d.b;
    to_dev<double, double>(ctx, l, x[l], d.vs, s.X[l]);
    to_dev<double, double>(ctx, l, b[l], d.vs, s.B[l]);
  }
  ctx->stat_cycles = 0;
  ctx->stat_bytes = 0.0;
  auto move_out = [&]() -> int {
    for (int l = bl; l <= level; ++l) {
      if (s.X[l] == x[l]) continue;
      to_user<double, double>(ctx, l, s.X[l], x[l], ctx->dev[l].vs);
      to_user<double, double>(ctx, l, s.B[l], b[l], ctx->dev[l].vs);
    }
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  };
  const int nc = ctx->dev[level].nb;
  if (do_defect)   // LinearDefect ON_SURFACE: b -= A x on the surface rows
    for (int l = s.frl; l <= level; ++l) {
      DevLevel &d = ctx->dev[l];
      launch_convert<double, double>(st, d.n * d.nb, s.B[l], d.t, 1, d.vs, d.vs);
      launch_spmv<double, double, double>(st, d.A.d64(), 0, d.A.nslices, s.X[l], d.t, nullptr, 0, d.nt, B1(d));
      launch_copy_rows(st, d.n, d.nb, surface_cls(ctx, s, l), d.t, s.B[l]);
      ctx->stat_bytes += op_bytes(ctx, l, Op::Spmv, 8) + 2.0 * vbytes(d, 8) + d.n;
    }
  // dnrm2x ON_SURFACE: per-level squared sums, added in level order
  for (int l = s.frl; l <= level; ++l) {
    DevLevel &d = ctx->dev[l];
    launch_dot<double>(st, d.n, s.B[l], nullptr, surface_cls(ctx, s, l), ctx->part,
                       ctx->surf_sq + (size_t)l * UGMG_MAXB, 0, B1(d));
    ctx->stat_bytes += vbytes(d, 8) + d.n;
  }
  HIPCHK(hipMemcpyAsync(ctx->surf_sq_host, ctx->surf_sq, sizeof(double) * UGMG_MAXB * (level + 1),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (ctx->nident > 0 && ctx->nident != nc)
    return fail(ctx, UGMG_E_ARG, "ugmg_set_ident gave " + std::to_string(ctx->nident) + " entries, the norms have " +
                                     std::to_string(nc));
  auto sc_cmp = [ctx, nc](const double *u, const double *v) { return sc_cmp_vd(ctx, nc, u, v); };
  double first[UGMG_MAXB] = {}, target[UGMG_MAXB], last[UGMG_MAXB];
  surface_norms(ctx, s, ctx->surf_sq_host, first);
  for (int i = 0; i < nc; ++i) {
    res->first_defect[i] = res->last_defect[i] = first[i];
    target[i] = first[i] * reduction[i];
    if (target[i] == 0.0) target[i] = reduction[i];
  }
  if (cb) cb(user, 0, nc, first);
  int conv = sc_cmp(first, abslimit);
  if (mixed)   // the FP32 copies of every level's b
    for (int l = bl; l <= level; ++l) {
      DevLevel &d = ctx->dev[l];
      launch_convert<float, double>(st, d.n * d.nb, s.B[l], d.b32, 1, d.vs, d.vs);
      ctx->stat_bytes += vbytes(d, 12);
    }
  const int kin = std::max(1, ctx->p.inner_cycles);
  const bool adapt = mixed && ctx->p.inner_adapt && kin > 1;
  double prev[UGMG_MAXB] = {};
  int prev_cycles = 0;
  for (int k = 0; k < nc; ++k) last[k] = first[k];
  int i;
  for (i = 0; i < maxiter; ++i) {
    if (conv) break;
    int cycles = mixed ? kin : 1;
    if (adapt && prev_cycles > 0)
      cycles = std::min(adapt_cycles(kin, nc, last, prev, prev_cycles, target),
                        adapt_cycles(kin, nc, last, prev, prev_cycles, abslimit));
    for (int k = 0; k < nc; ++k) prev[k] = last[k];
    prev_cycles = cycles;
    if ((rc = launch_surface_iteration(ctx, s, cycles)) ||
        (hipStreamSynchronize(st) != hipSuccess && (rc = fail(ctx, UGMG_E_HIP, "hipStreamSynchronize")))) {
      res->error_code = rc;
      move_out();
      return rc;
    }
    ctx->stat_cycles += cycles;
    ctx->stat_bytes += surface_iteration_bytes(ctx, s, cycles);
    surface_norms(ctx, s, ctx->surf_sq_host, last);
    for (int k = 0; k < nc; ++k) res->last_defect[k] = last[k];
    if (cb) cb(user, i + 1, nc, last);
    if (sc_cmp(last, abslimit) || sc_cmp(last, target)) {
      conv = 1;
      res->number_of_linear_iterations = i + 1;
      break;
    }
  }
  if (!conv) res->number_of_linear_iterations = i;
  res->converged = conv;
  return move_out();
}

int ugmg_solve_surface_host(ugmg_ctx *ctx, int level, double *const *x, double *const *b,
                            const double *abslimit, const double *reduction, int maxiter,
                            int do_defect, ugmg_lresult *res, ugmg_defect_cb cb, void *user) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (!x || !b) return fail(ctx, UGMG_E_ARG, "null argument");
  const int bl = ctx->bl;
  std::vector<double *> dx(level + 1, nullptr), db(level + 1, nullptr);
  auto release = [&]() {
    for (double *p : dx) hipFree(p);
    for (double *p : db) hipFree(p);
  };
  rc = UGMG_OK;
  for (int l = bl; l <= level && !rc; ++l) {
    if (!x[l] || !b[l]) {
      rc = fail(ctx, UGMG_E_ARG, "x[l] and b[l] are needed on baselevel..level");
      break;
    }
    const size_t bytes = sizeof(double) * ctx->dev[l].n * ctx->b;
    if (hipMalloc(&dx[l], bytes) != hipSuccess || hipMalloc(&db[l], bytes) != hipSuccess)
      rc = fail(ctx, UGMG_E_NOMEM, "hipMalloc");
    else if (hipMemcpy(dx[l], x[l], bytes, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(db[l], b[l], bytes, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(ctx, UGMG_E_HIP, "copy in");
  }
  if (!rc)
    rc = ugmg_solve_surface(ctx, level, dx.data(), db.data(), abslimit, reduction, maxiter, do_defect,
                            res, cb, user);
  for (int l = bl; l <= level && !rc; ++l) {
    const size_t bytes = sizeof(double) * ctx->dev[l].n * ctx->b;
    if (hipMemcpy(x[l], dx[l], bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b[l], db[l], bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(ctx, UGMG_E_HIP, "copy out");
  }
  release();
  return rc;
}

int ugmg_solve_host(ugmg_ctx *ctx, int level, double *x, double *b, const double *abslimit,
                    const double *reduction, int maxiter, int do_defect, ugmg_lresult *res,
                    ugmg_defect_cb cb, void *user) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  const size_t bytes = sizeof(double) * ctx->dev[level].n * ctx->b;
  double *dx = nullptr, *db = nullptr;
  HIPCHK(hipMalloc(&dx, bytes));
  if (hipMalloc(&db, bytes) != hipSuccess) {
    hipFree(dx);
    return fail(ctx, UGMG_E_NOMEM, "hipMalloc");
  }
  rc = UGMG_OK;
  if (hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(ctx, UGMG_E_HIP, "copy in");
  if (!rc) rc = ugmg_solve(ctx, level, dx, db, abslimit, reduction, maxiter, do_defect, res, cb, user);
  if (!rc && (hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(b, db, bytes, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(ctx, UGMG_E_HIP, "copy out");
  drop_graph(ctx);
  hipFree(dx);
  hipFree(db);
  return rc;
}

int ugmg_set_batch(ugmg_ctx *ctx, int nrhs) {
  if (!ctx) return UGMG_E_ARG;
  if (nrhs < 1 || nrhs > 65535) return fail(ctx, UGMG_E_ARG, "nrhs must be 1..65535");
  if (nrhs != ctx->nrhs_cap) {
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    free_device(ctx);
  }
  ctx->nrhs_cap = nrhs;
  return UGMG_OK;
}

int ugmg_iter_batch(ugmg_ctx *ctx, int level, int nrhs, double *c, double *b, int ld) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (nrhs < 1 || nrhs > ctx->nrhs_cap)
    return fail(ctx, UGMG_E_ARG, "nrhs exceeds the batch the hierarchy was laid out for (ugmg_set_batch)");
  if (!c || !b || ld < ctx->dev[level].n * ctx->b) return fail(ctx, UGMG_E_ARG, "null vector or ld < n*b");
  if (base_singular(ctx)) return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  DevLevel &d = ctx->dev[level];
  hipStream_t st = ctx->stream;
  ctx->nb = nrhs;
  const Batch bt = batch(ctx, level);
  if (ctx->p.precision == UGMG_MIXED) {
    to_dev<float, double>(ctx, level, b, ld, d.b32, nrhs);
    const int k = std::max(1, ctx->p.inner_cycles);
    for (int j = 0; j < k; ++j) lmgc<float>(ctx, level, d.c32, d.b32, j == 0 ? 0 : 1);
    launch_convert<double, float>(st, d.n * d.nb, d.c32, d.c, nrhs, d.vs, d.vs);
    to_dev<double, double>(ctx, level, b, ld, d.b, nrhs);
    launch_spmv<double, double, double>(st, d.A.d64(), 0, d.A.nslices, d.c, d.b, nullptr, 0, d.nt, bt);
  } else {
    to_dev<double, double>(ctx, level, b, ld, d.b, nrhs);
    lmgc<double>(ctx, level, d.c, d.b, 0);
  }
  to_user<double, double>(ctx, level, d.c, c, ld, nrhs);
  to_user<double, double>(ctx, level, d.b, b, ld, nrhs);
  ctx->nb = 1;
  HIPCHK(hipGetLastError());
  return UGMG_OK;
}

int ugmg_bcgs_solve(ugmg_ctx *ctx, int level, int nrhs, double *x, double *b, int ld,
                    const double *abslimit, const double *reduction, int maxiter, int restart,
                    int use_iter, int do_defect, ugmg_lresult *res, double *defects, int maxdef) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (nrhs < 1 || nrhs > ctx->nrhs_cap)
    return fail(ctx, UGMG_E_ARG, "nrhs exceeds the batch the hierarchy was laid out for (ugmg_set_batch)");
  if (!x || !b || !abslimit || !reduction || !res || maxiter < 0 || restart < 0 ||
      ld < ctx->dev[level].n || (defects && maxdef < 1))
    return fail(ctx, UGMG_E_ARG, "bad argument");
  if (ctx->b != 1 || ctx->ntypes > 1)
    return fail(ctx, UGMG_E_ARG, "bcgs runs on scalar (b = 1) hierarchies of one vector type");
  std::memset(res, 0, sizeof(*res) * nrhs);
  if (use_iter && base_singular(ctx)) {
    for (int j = 0; j < nrhs; ++j) res[j].error_code = UGMG_E_SINGULAR;
    return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  }
  const int md = defects ? maxdef : 1;
  if ((rc = bcgs_alloc(ctx, level, nrhs, md))) return rc;
  auto &w = ctx->bc;
  DevLevel &d = ctx->dev[level];
  hipStream_t st = ctx->stream;
  ctx->nb = nrhs;
  const Batch bt = batch(ctx, level);
  to_dev<double, double>(ctx, level, x, ld, w.x, nrhs);
  to_dev<double, double>(ctx, level, b, ld, w.b, nrhs);
  if (do_defect)  // BCGSDefect: b -= A x
    launch_spmv<double, double, double>(st, d.A.d64(), 0, d.A.nslices, w.x, w.b, nullptr, 0, d.nt, bt);
  // first defect (the solver reads LinearResiduum's last_defect)
  launch_bcgs_norm(st, d.n, w.b, d.vclass, w.partA, bt);
  launch_bcgs_scalar(st, BCGS_FIRST, bcgs_nparts(d.n), w.partA, w.state, nrhs, 1, abslimit[0],
                     reduction[0], defects ? w.defects : nullptr, md);
  rc = 0;
  int it = 0;
  auto any_active = [&](bool &any) -> int {
    HIPCHK(hipMemcpyAsync(ctx->state_host, w.state, sizeof(BcgsState) * nrhs, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    any = false;
    for (int j = 0; j < nrhs; ++j) any |= ctx->state_host[j].active != 0;
    return 0;
  };
  bool any = false;
  if ((rc = any_active(any))) {
    ctx->nb = 1;
    return rc;
  }
  for (it = 0; it < maxiter && any; ++it) {
    const int rs = it == 0 || (restart > 0 && it % restart == 0);
    if ((rc = launch_bcgs_iteration(ctx, level, use_iter != 0, rs, abslimit[0], reduction[0],
                                    defects ? md : 0)))
      break;
    if ((rc = any_active(any))) break;
  }
  ctx->nb = 1;
  if (rc) return rc;
  to_user<double, double>(ctx, level, w.x, x, ld, nrhs);
  to_user<double, double>(ctx, level, w.b, b, ld, nrhs);
  HIPCHK(hipGetLastError());
  if (defects)
    HIPCHK(hipMemcpyAsync(defects, w.defects, sizeof(double) * (size_t)md * nrhs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int j = 0; j < nrhs; ++j) {
    const BcgsState &S = ctx->state_host[j];
    res[j].converged = S.conv && !S.active;
    res[j].first_defect[0] = S.first;
    res[j].last_defect[0] = S.last;
    res[j].number_of_linear_iterations = S.nit;
  }
  return UGMG_OK;
}

int ugmg_bcgs_solve_host(ugmg_ctx *ctx, int level, int nrhs, double *x, double *b, int ld,
                         const double *abslimit, const double *reduction, int maxiter,
                         int restart, int use_iter, int do_defect, ugmg_lresult *res,
                         double *defects, int maxdef) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (nrhs < 1 || ld < ctx->dev[level].n * ctx->b || !x || !b) return fail(ctx, UGMG_E_ARG, "bad argument");
  const size_t bytes = sizeof(double) * (size_t)ld * nrhs;
  double *dx = nullptr, *db = nullptr;
  HIPCHK(hipMalloc(&dx, bytes));
  if (hipMalloc(&db, bytes) != hipSuccess) {
    hipFree(dx);
    return fail(ctx, UGMG_E_NOMEM, "hipMalloc");
  }
  rc = UGMG_OK;
  if (hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(ctx, UGMG_E_HIP, "copy in");
  if (!rc)
    rc = ugmg_bcgs_solve(ctx, level, nrhs, dx, db, ld, abslimit, reduction, maxiter, restart,
                         use_iter, do_defect, res, defects, maxdef);
  if (!rc && (hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(b, db, bytes, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(ctx, UGMG_E_HIP, "copy out");
  hipFree(dx);
  hipFree(db);
  return rc;
}

int ugmg_galerkin(ugmg_ctx *ctx, int level) {
  if (!ctx) return UGMG_E_ARG;
  if (level < 1 || level >= ctx->nlevels) return fail(ctx, UGMG_E_ARG, "level must be 1..nlevels-1");
  if (!ctx->level_set[level] || !ctx->level_set[level - 1] || !ctx->transfer_set[level])
    return fail(ctx, UGMG_E_STATE, "set both levels and the transfer first");
  HostLevel &F = ctx->host[level], &Cl = ctx->host[level - 1];
  if (!F.pblk.empty())
    return fail(ctx, UGMG_E_ARG, "the Galerkin product takes scalar interpolation weights (ugmg_set_transfer)");
  const int bb = ctx->b * ctx->b;
  // P^T over every fine vector (the coarse matrix takes all of them)
  std::vector<int> ptp(Cl.n + 1, 0), pti, fillp;
  std::vector<double> ptw;
  try {
    for (int q = 0; q < F.pp[F.n]; ++q) ++ptp[F.pi[q] + 1];
    for (int i = 0; i < Cl.n; ++i) ptp[i + 1] += ptp[i];
    pti.resize(ptp[Cl.n]);
    ptw.resize(ptp[Cl.n]);
    fillp.assign(ptp.begin(), ptp.end() - 1);
    for (int v = 0; v < F.n; ++v)
      for (int q = F.pp[v]; q < F.pp[v + 1]; ++q) {
        const int I = F.pi[q];
        pti[fillp[I]] = v;
        ptw[fillp[I]] = F.pw[q];
        ++fillp[I];
      }
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation");
  }
  HIPCHK(hipSetDevice(ctx->device));
  if (ctx->stream) HIPCHK(hipStreamSynchronize(ctx->stream));
  free_device(ctx);   // the coarse values change: layouts are rebuilt by ugmg_preprocess
  std::vector<void *> tmp;
  auto up = [&](const void *src, size_t bytes, void **dst) -> int {
    *dst = nullptr;
    if (hipMalloc(dst, bytes ? bytes : 1) != hipSuccess) return UGMG_E_NOMEM;
    tmp.push_back(*dst);
    if (bytes && src && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return UGMG_E_HIP;
    return 0;
  };
  void *d_ptp, *d_pti, *d_ptw, *d_frp, *d_fci, *d_fa, *d_pp, *d_pi, *d_pw, *d_crp, *d_cci, *d_ca;
  int rc = 0;
  const size_t nzc = (size_t)Cl.nnzb;
  if ((rc = up(ptp.data(), ptp.size() * 4, &d_ptp)) || (rc = up(pti.data(), pti.size() * 4, &d_pti)) ||
      (rc = up(ptw.data(), ptw.size() * 8, &d_ptw)) || (rc = up(F.rp.data(), F.rp.size() * 4, &d_frp)) ||
      (rc = up(F.ci.data(), F.ci.size() * 4, &d_fci)) || (rc = up(F.a.data(), F.a.size() * 8, &d_fa)) ||
      (rc = up(F.pp.data(), F.pp.size() * 4, &d_pp)) || (rc = up(F.pi.data(), F.pi.size() * 4, &d_pi)) ||
      (rc = up(F.pw.data(), F.pw.size() * 8, &d_pw)) || (rc = up(Cl.rp.data(), Cl.rp.size() * 4, &d_crp)) ||
      (rc = up(Cl.ci.data(), Cl.ci.size() * 4, &d_cci)) || (rc = up(nullptr, nzc * bb * 8, &d_ca))) {
    for (void *p : tmp) hipFree(p);
    return fail(ctx, rc, "device buffers for the Galerkin product");
  }
  hipStream_t st = ctx->stream;
  launch_galerkin(st, Cl.n, bb, (int *)d_ptp, (int *)d_pti, (double *)d_ptw, (int *)d_frp,
                  (int *)d_fci, (double *)d_fa, (int *)d_pp, (int *)d_pi, (double *)d_pw,
                  (int *)d_crp, (int *)d_cci, (double *)d_ca);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(Cl.a.data(), d_ca, nzc * bb * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  for (void *p : tmp) hipFree(p);
  if (e != hipSuccess) return fail(ctx, UGMG_E_HIP, std::string("Galerkin product: ") + hipGetErrorString(e));
  return UGMG_OK;
}

int ugmg_color(int n, const int *row_ptr, const int *col_idx, int *colors) {
  if (n < 0 || (n > 0 && (!row_ptr || !col_idx || !colors))) return UGMG_E_ARG;
  try {
    HostLevel L;
    L.n = n;
    L.rp.assign(row_ptr, row_ptr + n + 1);
    L.ci.assign(col_idx, col_idx + L.rp[n]);
    for (int c : L.ci)
      if (c < 0 || c >= n) return UGMG_E_ARG;
    std::vector<int> col;
    const int nc = greedy_color(L, col);
    std::copy(col.begin(), col.end(), colors);
    return nc;
  } catch (const std::bad_alloc &) {
    return UGMG_E_NOMEM;
  }
}

int ugmg_color_device(ugmg_ctx *ctx, int n, const int *row_ptr, const int *col_idx, int *colors,
                      int *rounds) {
  if (!ctx || n < 0 || (n > 0 && (!row_ptr || !col_idx || !colors))) return UGMG_E_ARG;
  if (n == 0) return 0;
  const int nnz = row_ptr[n];
  if (row_ptr[0] != 0 || nnz < 0) return fail(ctx, UGMG_E_ARG, "row_ptr");
  for (int i = 0; i < n; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) return fail(ctx, UGMG_E_ARG, "row_ptr not ascending");
  for (int m = 0; m < nnz; ++m)
    if (col_idx[m] < 0 || col_idx[m] >= n) return fail(ctx, UGMG_E_ARG, "column index out of range");
  HIPCHK(hipSetDevice(ctx->device));
  int *rp = nullptr, *ci = nullptr, *col = nullptr;
  auto release = [&] {
    (void)hipFree(rp);
    (void)hipFree(ci);
    (void)hipFree(col);
  };
  hipError_t e = hipMalloc((void **)&rp, sizeof(int) * ((size_t)n + 1));
  if (e == hipSuccess) e = hipMalloc((void **)&ci, sizeof(int) * std::max<size_t>(nnz, 1));
  if (e == hipSuccess) e = hipMalloc((void **)&col, sizeof(int) * (size_t)n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rp, row_ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && nnz)
    e = hipMemcpyAsync(ci, col_idx, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) {
    release();
    return fail(ctx, UGMG_E_HIP, std::string("colouring upload: ") + hipGetErrorString(e));
  }
  std::string err;
  int sw = 0;
  const int nc = color_device(ctx->stream, n, rp, ci, col, &sw, err);
  if (nc < 0) {
    release();
    return fail(ctx, nc, err);
  }
  e = hipMemcpyAsync(colors, col, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release();
  if (e != hipSuccess) return fail(ctx, UGMG_E_HIP, std::string("colouring: ") + hipGetErrorString(e));
  if (rounds) *rounds = sw;
  return nc;
}

int ugmg_set_setup(ugmg_ctx *ctx, int where) {
  if (!ctx || (where != UGMG_SETUP_HOST && where != UGMG_SETUP_DEVICE)) return UGMG_E_ARG;
  ctx->setup_device = where == UGMG_SETUP_DEVICE;
  return UGMG_OK;
}

int ugmg_layout_digest(ugmg_ctx *ctx, int level, unsigned long long *out, int nout) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const DevLevel &d = ctx->dev[level];
  std::vector<unsigned long long> dg;
  auto fnv = [](const void *p, size_t bytes) {
    const unsigned char *c = (const unsigned char *)p;
    unsigned long long h = 1469598103934665603ULL;
    for (size_t q = 0; q < bytes; ++q) h = (h ^ c[q]) * 1099511628211ULL;
    return h;
  };
  std::vector<unsigned char> buf;
  auto dev = [&](const void *p, size_t bytes) -> int {
    if (!p) {
      dg.push_back(0);
      return 0;
    }
    buf.resize(bytes);
    if (bytes) HIPCHK(hipMemcpy(buf.data(), p, bytes, hipMemcpyDeviceToHost));
    dg.push_back(fnv(buf.data(), bytes));
    return 0;
  };
  const int n = d.n, bb = d.nb * d.nb;
  const int mb = std::max(ctx->members, 1);
  for (const DevSellBuf *S : {&d.A, &d.L, &d.U, &d.P, &d.PT}) {
    const size_t ns = S->soff ? (size_t)S->nslices : 0;
    if ((rc = dev(S->soff, S->soff ? 4 * (ns + 1) : 0)) || (rc = dev(S->rows, 4 * ns * kSlice)) ||
        (rc = dev(S->col, 4 * S->nent)) || (rc = dev(S->sdoff, 4 * ns)) ||
        (rc = dev(S->dtab, 4 * S->ndtab)) || (rc = dev(S->v64, 8 * S->nval * mb)) ||
        (rc = dev(S->v32, 4 * S->nval * mb)))
      return rc;
  }
  if ((rc = dev(d.diag64, 8 * (size_t)n * bb * mb)) || (rc = dev(d.diag32, 4 * (size_t)n * bb * mb)) ||
      (rc = dev(d.vclass, n)) || (rc = dev(d.vnclass, n)) || (rc = dev(d.skip, 4 * (size_t)n)) ||
      (rc = dev(d.c0, n)) || (rc = dev(d.perm, 4 * (size_t)n)) || (rc = dev(d.vt, n)) ||
      (rc = dev(d.iskip, 4 * (size_t)n)))
    return rc;
  // host-side description
  std::vector<long long> h;
  auto put = [&](const auto &v) {
    h.push_back((long long)v.size());
    for (auto x : v) h.push_back((long long)x);
  };
  put(d.color);
  put(d.color_slices);
  put(d.color_rows);
  put(d.color_wmax);
  put(d.color_wmax_u);
  put(d.perm_h);
  for (long long x : {(long long)d.n, (long long)d.nnzb, (long long)d.ncolors, (long long)d.all_active,
                      (long long)d.p_nnz, (long long)d.pt_nnz, (long long)d.nt, (long long)d.nt32})
    h.push_back(x);
  for (const DevSellBuf *S : {&d.A, &d.L, &d.U, &d.P, &d.PT}) {
    for (long long x : {(long long)S->nslices, (long long)S->nent, (long long)S->nuniform,
                        (long long)S->nent_explicit, (long long)S->ndtab, (long long)S->bb,
                        (long long)S->wmax, (long long)S->nval, (long long)S->ncb})
      h.push_back(x);
    put(S->width_h);
    for (int k = 0; k <= S->ncb && k <= kMaxColors; ++k) h.push_back(S->cb[k]);
    dg.push_back(fnv(S->slot_pre.data(), S->slot_pre.size() * sizeof(double)));
    dg.push_back(fnv(S->expl_pre.data(), S->expl_pre.size() * sizeof(double)));
  }
  dg.push_back(fnv(h.data(), h.size() * sizeof(long long)));
  if (out)
    for (int q = 0; q < nout && q < (int)dg.size(); ++q) out[q] = dg[q];
  return (int)dg.size();
}

int ugmg_coarse_inverse(int n, const int *row_ptr, const int *col_idx, const double *vals,
                        int reorder, double *inv, int *order, int *bandwidth) {
  if (n < 0 || (n > 0 && (!row_ptr || !col_idx || !vals || !inv))) return UGMG_E_ARG;
  try {
    HostLevel L;
    L.n = n;
    L.rp.assign(row_ptr, row_ptr + n + 1);
    L.ci.assign(col_idx, col_idx + L.rp[n]);
    L.a.assign(vals, vals + L.rp[n]);
    for (int c : L.ci)
      if (c < 0 || c >= n) return UGMG_E_ARG;
    HostBase B;
    build_base(L, reorder, B);
    if (order) std::copy(B.order.begin(), B.order.end(), order);
    if (bandwidth) *bandwidth = B.bw;
    if (B.failed) return UGMG_E_SINGULAR;
    std::copy(B.inv.begin(), B.inv.end(), inv);
    return UGMG_OK;
  } catch (const std::bad_alloc &) {
    return UGMG_E_NOMEM;
  }
}

int ugmg_time_kernel(ugmg_ctx *ctx, int level, int kind, int reps, double *ms, double *bytes) {
  int rc = check_ready(ctx, level);
  if (rc) return rc;
  if (reps <= 0 || !ms || !bytes) return fail(ctx, UGMG_E_ARG, "bad argument");
  DevLevel &d = ctx->dev[level];
  if ((kind == 4 || kind == 5) && level <= ctx->bl) return fail(ctx, UGMG_E_ARG, "transfer needs level > base");
  if ((kind == 3 || kind == 7 || kind == 8) && d.ncolors == 0) return fail(ctx, UGMG_E_STATE, "not a GS level");
  if (kind == 8 && (!d.all_active || d.nb != 1))
    return fail(ctx, UGMG_E_STATE, "the fused GS step needs a scalar level whose rows are all active");
  hipStream_t st = ctx->stream;
  // deterministic nonzero vectors
  launch_fill<double>(st, d.n, d.t, 1.0, B1(d));
  launch_fill<double>(st, d.n, d.b, 0.5, B1(d));
  launch_fill<float>(st, d.n, d.t32, 1.0f, B1(d));
  launch_fill<float>(st, d.n, d.b32, 0.5f, B1(d));
  double by = 0.0;
  auto once = [&]() {
    switch (kind) {
      case 0: launch_spmv<double, double, double>(st, d.A.d64(), 0, d.A.nslices, d.t, d.b, nullptr, 0, d.nt, B1(d)); break;
      case 1: launch_spmv<float, float, float>(st, d.A.d32(), 0, d.A.nslices, d.t32, d.b32, nullptr, 0, d.nt32, B1(d)); break;
      case 2: smooth_step<double>(ctx, level, d.t, d.b, d.c, 1, nullptr); break;
      case 3: {
        const int s = ctx->p.smoother;
        ctx->p.smoother = UGMG_SMOOTH_GS;
        smooth_step<double>(ctx, level, d.t, d.b, d.c, 1, nullptr);
        ctx->p.smoother = s;
        break;
      }
      case 4: {
        DevLevel &dc = ctx->dev[level - 1];
        launch_restrict<double>(st, d.PT.d64(), dc.skip, dc.vnclass, dc.vclass, d.b, 1.0, dc.b,
                                nullptr, d.nt, 1, d.vs, dc.vs, d.nb);
        break;
      }
      case 5: {
        DevLevel &dc = ctx->dev[level - 1];
        launch_interp<double>(st, d.P.d64(), d.iskip, dc.c, 1.0, d.t, d.c, d.nt, 1, d.vs, dc.vs, d.nb);
        break;
      }
      case 6:
        if (ctx->p.precision == UGMG_MIXED) lmgc<float>(ctx, level, d.c32, d.b32, 0);
        else lmgc<double>(ctx, level, d.c, d.b, 0);
        break;
      case 7: {
        const int s = ctx->p.smoother;
        ctx->p.smoother = UGMG_SMOOTH_GS;
        smooth_step<float>(ctx, level, d.t32, d.b32, d.c32, 1, nullptr);
        ctx->p.smoother = s;
        break;
      }
      case 8:   // a GS step as a cycle runs its steps after the first
        for (int k = 0; k < d.ncolors; ++k)
          launch_gs_color_fused<double, double>(st, d.L.d64(), d.U.d64(), d.color_slices[k],
                                                d.color_slices[k + 1] - d.color_slices[k],
                                                d.color_wmax[k], d.color_wmax_u[k], d.b,
                                                d.diag<double>(), ctx->p.damp[0], d.t, d.c, 1,
                                                d.nt, B1(d), ctx->gs_split);
        break;
      default: break;
    }
  };
  switch (kind) {
    case 0: by = op_bytes(ctx, level, Op::Spmv, 8); break;
    case 1: by = op_bytes(ctx, level, Op::Spmv, 4); break;
    case 2: by = op_bytes(ctx, level, Op::Jac, 8) + op_bytes(ctx, level, Op::SpmvC, 8); break;
    case 3: by = gs_step_bytes(ctx, level, 8, false); break;
    case 7: by = gs_step_bytes(ctx, level, 4, false); break;
    case 8: by = gs_step_bytes(ctx, level, 8, true); break;
    case 4: by = op_bytes(ctx, level, Op::Restrict, 8); break;
    case 5: by = op_bytes(ctx, level, Op::Interp, 8); break;
    case 6: by = cycle_bytes(ctx, level, ctx->p.precision == UGMG_MIXED ? 4 : 8); break;
    default: by = 0.0;
  }
  if (kind < 0 || kind > 8) return fail(ctx, UGMG_E_ARG, "unknown kernel kind");
  once();  // warm
  HIPCHK(hipStreamSynchronize(st));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) once();
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float msf = 0.0f;
  HIPCHK(hipEventElapsedTime(&msf, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms = (double)msf / reps;
  *bytes = by;
  return UGMG_OK;
}

}  // extern "C"

// ================================================================= serving
//
// Continuous batching of independent solves over resident, batched
// hierarchies.  Requests with the same hierarchy (pattern, transfers,
// classes and Dirichlet bits on every level) and the same solver parameters
// share a group: one context whose levels hold `cap` members -- one matrix
// per column of the level vectors, one pattern -- whose host layout is built
// once (on kSrvWorkers threads) and which stays resident while the HBM
// budget allows.  A request joins a free slot of its group at an iteration
// boundary: its matrices are gathered into the slot's value planes (device
// maps from its CSR values to slab positions), its base level is factored,
// its rhs and a zero x enter the slot's column.  Every ls iteration of a
// group is one batched launch sequence over columns 0..w-1 (w = the highest
// occupied slot + 1), replayed from a graph captured per w, so the
// same-level kernels of all of a group's solves run as single launches;
// groups overlap on the device through their streams.  Each column runs
// exactly the kernels of a stand-alone ugmg_solve of its request on its own
// values, so every result is the same alone and in the mix.

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <system_error>
#include <thread>

namespace {

constexpr int kSrvWorkers = 4;
constexpr int kSrvMaxCap = 64;

struct SrvReq {
  int id = 0;
  int state = UGMG_REQ_QUEUED;
  double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
  int top = 0, n = 0;
  int group = -1, slot = -1;
  std::vector<std::vector<double>> vals;   // per level: list-order CSR values
  std::vector<double> rhs;
  double red = 1e-8, abslimit = 1e-300;
  int maxiter = 100;
  int it = 0;
  double target = 0.0;
  ugmg_lresult res{};
  std::vector<double> xout;
  std::string err;
};

struct SrvGroup {
  uint64_t key = 0;
  ugmg_ctx *ctx = nullptr;
  int top = 0, cap = 1;
  PrepState ps;
  bool host_ready = false, dev_ready = false, dead = false;
  int prep_rc = 0;
  std::string err;
  double est_bytes = 0.0, charged = 0.0;
  std::vector<int> slot;        // request id per slot (-1: free)
  bool busy = false;            // an iteration is in flight
  int width = 0;                // its columns
  hipEvent_t ev = nullptr;
  double *x = nullptr, *b = nullptr;   // top level, cap columns, device order
  double *norms = nullptr;             // pinned, cap
  double last_used = 0.0;
  int occupied() const {
    int k = 0;
    for (int r : slot) k += r >= 0;
    return k;
  }
};

// 64-bit hash of a byte range, eight bytes per step (a key for finding a
// request's group; equal keys are confirmed by same_shape)
uint64_t fnv(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = static_cast<const unsigned char *>(p);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, c + i, 8);
    h = (h ^ w) * 0x9E3779B97F4A7C15ULL;
    h ^= h >> 29;
  }
  for (; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
  return h;
}

// everything a group shares: the level patterns, transfers, classes,
// Dirichlet bits and the solver parameters
uint64_t group_key(int nlevels, const ugmg_level_desc *lv, const ugmg_params &p) {
  uint64_t h = fnv(1469598103934665603ULL, &nlevels, sizeof nlevels);
  h = fnv(h, &p, sizeof p);
  for (int l = 0; l < nlevels; ++l) {
    const ugmg_level_desc &L = lv[l];
    const int n = L.n, nnz = L.row_ptr[L.n];
    h = fnv(h, &n, sizeof n);
    h = fnv(h, L.row_ptr, sizeof(int) * (n + 1));
    h = fnv(h, L.col_idx, sizeof(int) * nnz);
    h = fnv(h, L.skip, sizeof(uint32_t) * n);
    h = fnv(h, L.vclass, n);
    h = fnv(h, L.vnclass, n);
    if (l > 0) {
      h = fnv(h, L.p_ptr, sizeof(int) * (n + 1));
      h = fnv(h, L.p_idx, sizeof(int) * L.p_ptr[n]);
      h = fnv(h, L.p_w, sizeof(double) * L.p_ptr[n]);
    }
  }
  return h;
}

// the same data compared in full against a group's host levels
bool same_shape(const ugmg_ctx *ctx, int nlevels, const ugmg_level_desc *lv) {
  if (ctx->nlevels != nlevels) return false;
  for (int l = 0; l < nlevels; ++l) {
    const HostLevel &H = ctx->host[l];
    const ugmg_level_desc &L = lv[l];
    if (H.n != L.n || H.nnzb != L.row_ptr[L.n]) return false;
    if (!std::equal(H.rp.begin(), H.rp.end(), L.row_ptr) ||
        !std::equal(H.ci.begin(), H.ci.end(), L.col_idx) ||
        !std::equal(H.skip.begin(), H.skip.end(), L.skip) ||
        !std::equal(H.vc.begin(), H.vc.end(), L.vclass) ||
        !std::equal(H.vnc.begin(), H.vnc.end(), L.vnclass))
      return false;
    if (l > 0 && (!H.has_p || !std::equal(H.pp.begin(), H.pp.end(), L.p_ptr) ||
                  !std::equal(H.pi.begin(), H.pi.end(), L.p_idx) ||
                  !std::equal(H.pw.begin(), H.pw.end(), L.p_w)))
      return false;
  }
  return true;
}

double est_device_bytes(const ugmg_ctx *ctx, int top, int bl) {
  double by = 0.0;
  const bool gs = uses_gs(ctx->p);
  for (int l = bl; l <= top; ++l) {
    const HostLevel &H = ctx->host[l];
    by += (double)H.nnzb * 16.0 * 1.1;                      // A: col + FP64 + FP32, padding
    if (gs && l > bl) by += (double)(H.nnzb - H.n) * 16.0 * 1.1;  // L + U
    if (l > bl) by += 2.0 * (double)H.pp[H.n] * 16.0 * 1.1;      // P and P^T
    by += (double)H.n * 70.0;                                 // vectors, diag, flags
  }
  by += (double)ctx->host[bl].n * ctx->host[bl].n * 8.0;     // dense inverse
  return by;
}

}  // namespace

// ------------------------------------------------------------ members
// Member m of a context with members: its matrices (per level, list-order
// CSR values of the context's pattern) into its value planes, diagonal
// blocks and base inverse (EXPreProcess's band LU of its base level on the
// host, as for a stand-alone context with a small base level).  The base
// level's unit rows must be the prototype's.  Synchronous.
static int member_set(ugmg_ctx *ctx, int m, const std::vector<std::vector<double>> &vals) {
  if (m < 0 || m >= ctx->members) return fail(ctx, UGMG_E_ARG, "member out of range");
  hipStream_t st = ctx->stream;
  const int bl = ctx->bl, top = ctx->top;
  for (int l = bl; l <= top; ++l) {
    DevLevel &D = ctx->dev[l];
    const HostLevel &H = ctx->host[l];
    if (vals[l].size() != H.a.size()) return fail(ctx, UGMG_E_ARG, "member values of the wrong size");
    HIPCHK(hipMemcpyAsync(ctx->stage, vals[l].data(), sizeof(double) * H.a.size(), hipMemcpyHostToDevice, st));
    launch_member_gather(st, D.A.nval, D.map_a, ctx->stage, D.A.v64 + m * D.A.nval, D.A.v32 + m * D.A.nval);
    launch_member_gather(st, D.dms, D.map_d, ctx->stage, D.diag64 + m * D.dms, D.diag32 + m * D.dms);
    if (D.ncolors > 0) {
      launch_member_gather(st, D.L.nval, D.map_l, ctx->stage, D.L.v64 + m * D.L.nval, D.L.v32 + m * D.L.nval);
      launch_member_gather(st, D.U.nval, D.map_u, ctx->stage, D.U.v64 + m * D.U.nval, D.U.v32 + m * D.U.nval);
    }
    // the copy must land before the next level's overwrites the staging
    HIPCHK(hipStreamSynchronize(st));
  }
  // the base level: the unit rows of the inverse are those of the prototype
  HostLevel Hb = ctx->host[bl];
  Hb.a.assign(vals[bl].begin(), vals[bl].end());
  const int nb = ctx->b, nsc = Hb.n * nb;
  std::vector<int> rowmap(nsc);
  HIPCHK(hipMemcpy(rowmap.data(), ctx->base_rowmap, sizeof(int) * nsc, hipMemcpyDeviceToHost));
  for (int v = 0; v < Hb.n; ++v)
    for (int q = 0; q < nb; ++q) {
      bool unit = Hb.rp[v] < Hb.rp[v + 1] && Hb.ci[Hb.rp[v]] == v;
      for (int e = Hb.rp[v]; e < Hb.rp[v + 1] && unit; ++e)
        for (int j = 0; j < nb; ++j)
          if (Hb.a[((size_t)e * nb + q) * nb + j] != ((Hb.ci[e] == v && j == q) ? 1.0 : 0.0)) {
            unit = false;
            break;
          }
      if (unit != (rowmap[v * nb + q] < 0))
        return fail(ctx, UGMG_E_ARG, "the base level's Dirichlet rows differ from the group's");
    }
  HostBase B;
  try {
    build_base(Hb, ctx->p.ex_reorder, B, ctx->p.ex_float != 0);
  } catch (const std::bad_alloc &) {
    return fail(ctx, UGMG_E_NOMEM, "host allocation (base level)");
  }
  if (B.failed) return fail(ctx, UGMG_E_SINGULAR, "zero pivot in the band LU of the base level");
  if (nsc > 0 && ctx->base_nact > 0) {
    HIPCHK(hipMemcpyAsync(ctx->base_stage, B.inv.data(), sizeof(double) * (size_t)nsc * nsc,
                          hipMemcpyHostToDevice, st));
    launch_base_compact(st, ctx->base_nact, nsc, ctx->base_act, ctx->base_colperm, ctx->base_stage,
                        ctx->base_cinv + m * ctx->base_ims, ctx->base_cinv32 + m * ctx->base_ims);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

namespace {

// one ls iteration of a group over columns 0..w-1: the FP64 cycle into x,
// then every column's defect norm into red[0..w)
void enqueue_group_iteration(ugmg_ctx *ctx, int top, double *x, double *b, int w) {
  ctx->nb = w;
  lmgc<double>(ctx, top, x, b, 1);
  DevLevel &d = ctx->dev[top];
  launch_dot<double>(ctx->stream, d.n, b, nullptr, d.vclass, ctx->part, ctx->red, 1,
                     Batch{w, d.vs, d.nb});
  ctx->nb = 1;
}

// from the graph captured for w columns (kept with the group)
int launch_group_iteration(ugmg_ctx *ctx, int top, double *x, double *b, int w, double *norms_host) {
  if (!can_capture(ctx) || w >= ugmg_ctx::kMaxGraphs) {
    enqueue_group_iteration(ctx, top, x, b, w);
    HIPCHK(hipGetLastError());
  } else {
    if (ctx->gkind != -2 || ctx->glevel != top || ctx->gx != x || ctx->gb != b) {
      drop_graph(ctx);
      ctx->gkind = -2;
      ctx->glevel = top;
      ctx->gx = x;
      ctx->gb = b;
    }
    hipGraphExec_t &ge = ctx->gexec[w];
    if (!ge) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      enqueue_group_iteration(ctx, top, x, b, w);
      HIPCHK(hipStreamEndCapture(ctx->stream, &g));
      hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        ge = nullptr;
        return fail(ctx, UGMG_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    HIPCHK(hipGraphLaunch(ge, ctx->stream));
  }
  HIPCHK(hipMemcpyAsync(norms_host, ctx->red, sizeof(double) * w, hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

}  // namespace

struct ugmg_server {
  int device = 0;
  double budget = 0.0, used = 0.0;
  int max_slots = 1;
  std::vector<std::unique_ptr<SrvReq>> reqs;
  std::deque<int> queue;                       // submitted, not admitted
  std::vector<std::unique_ptr<SrvGroup>> groups;
  int nactive = 0, ndone = 0;
  long long launches = 0, columns = 0;   // group iterations, solve iterations in them
  std::string err;
  std::chrono::steady_clock::time_point t0;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<SrvGroup *> prep_queue;
  bool stop = false;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void server_worker(ugmg_server *s) {
  for (;;) {
    SrvGroup *g = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->prep_queue.empty(); });
      if (s->stop) return;
      g = s->prep_queue.front();
      s->prep_queue.pop_front();
    }
    const int rc = prep_host(g->ctx, g->top, g->ps);
    std::lock_guard<std::mutex> lk(s->mu);
    g->prep_rc = rc;
    if (rc) g->err = g->ctx->err;
    g->host_ready = true;
  }
}

int srv_fail(ugmg_server *s, int code, const std::string &msg) {
  s->err = msg;
  return code;
}

// request done: its solution (slot column of the group, device order) to
// the host in list order; the slot is freed and its columns cleared
void srv_finish(ugmg_server *s, SrvReq *r, int state) {
  if (r->group >= 0 && r->slot >= 0) {
    SrvGroup *g = s->groups[r->group].get();
    ugmg_ctx *ctx = g->ctx;
    const DevLevel &d = ctx->dev[g->top];
    if (state == UGMG_REQ_DONE) {
      r->xout.resize(d.n);
      if (hipMemcpyAsync(r->xout.data(), g->x + (size_t)r->slot * d.vs, sizeof(double) * d.n,
                         hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        state = UGMG_REQ_FAILED;
        r->err = "copy of the solution failed";
      } else if (!d.perm_h.empty()) {   // device order -> list order
        std::vector<double> xl(d.n);
        for (int q = 0; q < d.n; ++q) xl[d.perm_h[q]] = r->xout[q];
        r->xout.swap(xl);
      }
    }
    launch_fill<double>(ctx->stream, d.n, g->x + (size_t)r->slot * d.vs, 0.0, B1(d));
    launch_fill<double>(ctx->stream, d.n, g->b + (size_t)r->slot * d.vs, 0.0, B1(d));
    g->slot[r->slot] = -1;
    g->last_used = s->now();
    --s->nactive;
    r->slot = -1;
  }
  std::vector<std::vector<double>>().swap(r->vals);
  std::vector<double>().swap(r->rhs);
  r->state = state;
  r->t_done = s->now();
  ++s->ndone;
}

void group_release(ugmg_server *s, SrvGroup *g) {
  if (g->ctx) {
    if (g->ctx->stream) hipStreamSynchronize(g->ctx->stream);
    ugmg_destroy(g->ctx);
    g->ctx = nullptr;
  }
  if (g->ev) hipEventDestroy(g->ev);
  g->ev = nullptr;
  if (g->norms) hipHostFree(g->norms);
  g->norms = nullptr;
  s->used -= g->charged;
  g->charged = 0.0;
  g->dead = true;
  g->dev_ready = false;
}

// device set-up of a group whose host layout is ready
int group_setup(ugmg_server *s, SrvGroup *g) {
  ugmg_ctx *ctx = g->ctx;
  int rc = prep_device(ctx, g->ps);
  g->ps = PrepState();
  if (rc) {
    g->err = ctx->err;
    return rc;
  }
  const DevLevel &d = ctx->dev[g->top];
  const size_t nv = d.vs * (size_t)g->cap;
  if ((rc = dalloc(ctx, &g->x, nv)) || (rc = dalloc(ctx, &g->b, nv))) {
    g->err = ctx->err;
    return rc;
  }
  launch_fill<double>(ctx->stream, d.n, g->x, 0.0, Batch{g->cap, d.vs, d.nb});
  launch_fill<double>(ctx->stream, d.n, g->b, 0.0, Batch{g->cap, d.vs, d.nb});
  if (hipEventCreateWithFlags(&g->ev, hipEventDisableTiming) != hipSuccess ||
      hipHostMalloc((void **)&g->norms, sizeof(double) * g->cap, hipHostMallocDefault) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    g->err = "event / pinned memory / first fill";
    return UGMG_E_HIP;
  }
  g->charged = ctx->dev_bytes;
  s->used += g->charged;
  g->dev_ready = true;
  g->last_used = s->now();
  return UGMG_OK;
}

// admit a request into free slot j of its (idle) group: its matrices, its
// rhs and a zero x into column j, every level's vectors of that column
// cleared, the first defect norm and stop test.  Returns 1 when the request
// is already finished.
int srv_admit(ugmg_server *s, SrvReq *r, SrvGroup *g, int j) {
  ugmg_ctx *ctx = g->ctx;
  int rc = member_set(ctx, j, r->vals);
  if (rc) {
    r->err = ctx->err;
    return rc;
  }
  std::vector<std::vector<double>>().swap(r->vals);
  hipStream_t st = ctx->stream;
  for (int l = ctx->bl; l <= g->top; ++l) {
    DevLevel &D = ctx->dev[l];
    for (double *v : {D.c, D.b, D.t, D.u}) launch_fill<double>(st, D.n, v + (size_t)j * D.vs, 0.0, B1(D));
  }
  DevLevel &d = ctx->dev[g->top];
  if (!d.perm_h.empty()) {   // list order -> device order
    std::vector<double> rp(d.n);
    for (int q = 0; q < d.n; ++q) rp[q] = r->rhs[d.perm_h[q]];
    r->rhs.swap(rp);
  }
  double *bj = g->b + (size_t)j * d.vs, *xj = g->x + (size_t)j * d.vs;
  HIPCHK(hipMemcpyAsync(bj, r->rhs.data(), sizeof(double) * d.n, hipMemcpyHostToDevice, st));
  launch_fill<double>(st, d.n, xj, 0.0, B1(d));
  launch_dot<double>(st, d.n, bj, nullptr, d.vclass, ctx->part, ctx->red, 1, B1(d));
  HIPCHK(hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  g->slot[j] = r->id;
  r->group = (int)(std::find_if(s->groups.begin(), s->groups.end(),
                                [g](const std::unique_ptr<SrvGroup> &q) { return q.get() == g; }) -
                   s->groups.begin());
  r->slot = j;
  ++s->nactive;
  r->state = UGMG_REQ_ACTIVE;
  r->t_admit = s->now();
  g->last_used = r->t_admit;
  // the first defect: sc_mul_check target, stop on $abslimit
  const double val = ctx->red_host[0];
  r->res.first_defect[0] = r->res.last_defect[0] = val;
  r->target = val * r->red;
  if (r->target == 0.0) r->target = r->red;
  if (sc_cmp1(val, r->abslimit) || r->maxiter <= 0) {
    r->res.converged = sc_cmp1(val, r->abslimit) ? 1 : 0;
    r->res.number_of_linear_iterations = 0;
    srv_finish(s, r, UGMG_REQ_DONE);
    return 1;
  }
  return 0;
}

// a group's completed iteration: every column's stop test
void srv_progress(ugmg_server *s, SrvGroup *g) {
  for (int j = 0; j < g->width; ++j) {
    const int id = g->slot[j];
    if (id < 0) continue;
    SrvReq *r = s->reqs[id].get();
    const double val = g->norms[j];
    ++r->it;
    r->res.last_defect[0] = val;
    const bool conv = sc_cmp1(val, r->abslimit) || sc_cmp1(val, r->target);
    if (conv || r->it >= r->maxiter) {
      r->res.converged = conv ? 1 : 0;
      r->res.number_of_linear_iterations = r->it;
      srv_finish(s, r, UGMG_REQ_DONE);
    }
  }
}

void group_fail_all(ugmg_server *s, SrvGroup *g, int code, const std::string &msg) {
  for (int j = 0; j < (int)g->slot.size(); ++j)
    if (g->slot[j] >= 0) {
      SrvReq *r = s->reqs[g->slot[j]].get();
      r->err = msg;
      r->res.error_code = code;
      r->slot = -1;
      g->slot[j] = -1;
      --s->nactive;
      srv_finish(s, r, UGMG_REQ_FAILED);
    }
}

}  // namespace

extern "C" {

int ugmg_server_create(int device, double hbm_budget_bytes, int max_slots, ugmg_server **out) {
  if (!out || max_slots < 1 || hbm_budget_bytes <= 0.0) return UGMG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return UGMG_E_NODEV;
  if (device < 0 || device >= ndev) return UGMG_E_ARG;
  ugmg_server *s = new (std::nothrow) ugmg_server;
  if (!s) return UGMG_E_NOMEM;
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->max_slots = max_slots;
  s->t0 = std::chrono::steady_clock::now();
  try {
    for (int k = 0; k < kSrvWorkers; ++k) s->workers.emplace_back(server_worker, s);
  } catch (const std::system_error &) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    for (auto &w : s->workers) w.join();
    delete s;
    return UGMG_E_NOMEM;
  }
  *out = s;
  return UGMG_OK;
}

int ugmg_server_destroy(ugmg_server *s) {
  if (!s) return UGMG_E_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto &w : s->workers)
    if (w.joinable()) w.join();
  for (auto &g : s->groups)
    if (g && !g->dead) group_release(s, g.get());
  delete s;
  return UGMG_OK;
}

const char *ugmg_server_last_error(const ugmg_server *s) { return s ? s->err.c_str() : "null server"; }

int ugmg_server_submit(ugmg_server *s, int nlevels, const ugmg_level_desc *levels,
                       const ugmg_params *p, const double *rhs, double red, double abslimit,
                       int maxiter, int *id) {
  if (!s || !levels || !p || !rhs || !id || nlevels < 1 || maxiter < 0) return UGMG_E_ARG;
  if (p->precision != UGMG_FP64) return srv_fail(s, UGMG_E_ARG, "the server runs FP64 solves");
  if (p->base_iter != UGMG_BASE_EX || split_smoothers(*p))
    return srv_fail(s, UGMG_E_ARG, "the server runs one smoother class and the ex base solve");
  for (int l = 0; l < nlevels; ++l) {
    const ugmg_level_desc &L = levels[l];
    if (L.n < 0 || !L.row_ptr || !L.col_idx || !L.vals || !L.skip || !L.vclass || !L.vnclass ||
        (l > 0 && (!L.p_ptr || !L.p_idx || !L.p_w)))
      return srv_fail(s, UGMG_E_ARG, "incomplete level description");
  }
  auto r = std::make_unique<SrvReq>();
  r->top = nlevels - 1;
  r->n = levels[r->top].n;
  r->rhs.assign(rhs, rhs + r->n);
  r->red = red;
  r->abslimit = abslimit;
  r->maxiter = maxiter;
  r->vals.resize(nlevels);
  for (int l = 0; l < nlevels; ++l)
    r->vals[l].assign(levels[l].vals, levels[l].vals + levels[l].row_ptr[levels[l].n]);
  // the request's group: the first live one of its shape, or a new one
  const uint64_t key = group_key(nlevels, levels, *p);
  int gi = -1;
  for (int k = 0; k < (int)s->groups.size() && gi < 0; ++k) {
    SrvGroup *g = s->groups[k].get();
    if (!g->dead && g->key == key && same_shape(g->ctx, nlevels, levels)) gi = k;
  }
  if (gi < 0) {
    auto g = std::make_unique<SrvGroup>();
    g->key = key;
    g->top = r->top;
    g->cap = std::min(s->max_slots, kSrvMaxCap);
    g->slot.assign(g->cap, -1);
    ugmg_ctx *ctx = nullptr;
    int rc = ugmg_create(s->device, &ctx);
    if (rc) return srv_fail(s, rc, "ugmg_create failed");
    g->ctx = ctx;
    auto bail = [&](int code) {
      s->err = ctx->err;
      ugmg_destroy(ctx);
      return code;
    };
    if ((rc = ugmg_set_nlevels(ctx, nlevels, 1))) return bail(rc);
    for (int l = 0; l < nlevels; ++l) {
      const ugmg_level_desc &L = levels[l];
      if ((rc = ugmg_set_level(ctx, l, L.n, L.row_ptr, L.col_idx, L.vals, L.skip, L.vclass, L.vnclass)))
        return bail(rc);
    }
    for (int l = 1; l < nlevels; ++l)
      if ((rc = ugmg_set_transfer(ctx, l, levels[l].p_ptr, levels[l].p_idx, levels[l].p_w)))
        return bail(rc);
    ugmg_params q = *p;
    if ((rc = ugmg_set_params(ctx, &q))) return bail(rc);
    ctx->nrhs_cap = g->cap;
    ctx->members = g->cap;
    ctx->grouped = true;
    // every member's matrices, vectors and base inverse; patterns once
    g->est_bytes = est_device_bytes(ctx, r->top, resolve_baselevel(ctx, r->top)) * g->cap;
    gi = (int)s->groups.size();
    std::lock_guard<std::mutex> lk(s->mu);
    s->prep_queue.push_back(g.get());
    s->groups.push_back(std::move(g));
    s->cv.notify_one();
  }
  r->group = gi;
  r->t_submit = s->now();
  r->id = (int)s->reqs.size();
  *id = r->id;
  s->queue.push_back(r->id);
  s->reqs.push_back(std::move(r));
  return UGMG_OK;
}

int ugmg_server_step(ugmg_server *s, int *nactive, int *nqueued, int *ndone) {
  if (!s) return UGMG_E_ARG;
  if (hipSetDevice(s->device) != hipSuccess) return srv_fail(s, UGMG_E_HIP, "hipSetDevice");
  // 1. completed iterations: every column's stop test
  for (auto &gp : s->groups) {
    SrvGroup *g = gp.get();
    if (g->dead || !g->busy) continue;
    const hipError_t q = hipEventQuery(g->ev);
    if (q == hipErrorNotReady) continue;
    g->busy = false;
    if (q != hipSuccess) {
      group_fail_all(s, g, UGMG_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
      continue;
    }
    srv_progress(s, g);
  }
  // 2. device set-up of groups whose layout is ready, within the budget
  //    (idle resident groups are released, least recently used first, to
  //    make room); a group that alone exceeds it fails its requests
  std::vector<SrvGroup *> ready;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &gp : s->groups)
      if (!gp->dead && gp->host_ready && !gp->dev_ready) ready.push_back(gp.get());
  }
  std::vector<char> failed_group(s->groups.size(), 0);
  for (SrvGroup *g : ready) {
    const int gi = (int)(std::find_if(s->groups.begin(), s->groups.end(),
                                      [g](const std::unique_ptr<SrvGroup> &q) { return q.get() == g; }) -
                         s->groups.begin());
    int rc = g->prep_rc;
    if (!rc && g->est_bytes > s->budget) {
      rc = UGMG_E_BUDGET;
      g->err = "request exceeds the HBM budget";
    }
    while (!rc && s->used + g->est_bytes > s->budget) {
      SrvGroup *lru = nullptr;
      for (auto &q : s->groups)
        if (!q->dead && q->dev_ready && !q->busy && q->occupied() == 0 &&
            (!lru || q->last_used < lru->last_used))
          lru = q.get();
      if (!lru) break;
      group_release(s, lru);
    }
    if (!rc && s->used + g->est_bytes > s->budget) continue;   // waits for memory
    if (!rc) rc = group_setup(s, g);
    if (rc) {
      failed_group[gi] = 1;
      for (size_t k = 0; k < s->queue.size();) {
        SrvReq *r = s->reqs[s->queue[k]].get();
        if (r->group == gi) {
          s->queue.erase(s->queue.begin() + k);
          r->err = g->err;
          r->res.error_code = rc;
          r->group = -1;
          srv_finish(s, r, UGMG_REQ_FAILED);
        } else {
          ++k;
        }
      }
      group_release(s, g);
    }
  }
  // 3. admission in submission order into free slots of idle groups, while
  //    slots allow (a request whose group is not ready yet holds back none)
  for (size_t k = 0; k < s->queue.size() && s->nactive < s->max_slots;) {
    SrvReq *r = s->reqs[s->queue[k]].get();
    SrvGroup *g = s->groups[r->group].get();
    if (g->dead) {   // released while the request waited: a new group
      s->queue.erase(s->queue.begin() + k);
      r->err = "the request's group was released";
      r->res.error_code = UGMG_E_STATE;
      r->group = -1;
      srv_finish(s, r, UGMG_REQ_FAILED);
      continue;
    }
    if (!g->dev_ready || g->busy) {
      ++k;
      continue;
    }
    int j = 0;
    while (j < g->cap && g->slot[j] >= 0) ++j;
    if (j == g->cap) {
      ++k;
      continue;
    }
    s->queue.erase(s->queue.begin() + k);
    const int rc = srv_admit(s, r, g, j);
    if (rc < 0) {
      r->res.error_code = rc;
      if (r->slot >= 0) srv_finish(s, r, UGMG_REQ_FAILED);
      else {
        r->group = -1;
        srv_finish(s, r, UGMG_REQ_FAILED);
      }
    }
  }
  // 4. the next iteration of every idle group with solves in it
  for (auto &gp : s->groups) {
    SrvGroup *g = gp.get();
    if (g->dead || !g->dev_ready || g->busy || g->occupied() == 0) continue;
    int w = 0;
    for (int j = 0; j < g->cap; ++j)
      if (g->slot[j] >= 0) w = j + 1;
    const int rc = launch_group_iteration(g->ctx, g->top, g->x, g->b, w, g->norms);
    if (rc || hipEventRecord(g->ev, g->ctx->stream) != hipSuccess) {
      group_fail_all(s, g, rc ? rc : UGMG_E_HIP, g->ctx->err);
      continue;
    }
    g->width = w;
    g->busy = true;
    ++s->launches;
    s->columns += g->occupied();
  }
  if (nactive) *nactive = s->nactive;
  if (nqueued) *nqueued = (int)s->queue.size();
  if (ndone) *ndone = s->ndone;
  return UGMG_OK;
}

int ugmg_server_status(ugmg_server *s, int id, int *state, ugmg_lresult *res, double *times) {
  if (!s || id < 0 || id >= (int)s->reqs.size()) return UGMG_E_ARG;
  const SrvReq *r = s->reqs[id].get();
  if (state) *state = r->state;
  if (res) *res = r->res;
  if (times) {
    times[0] = r->t_submit;
    times[1] = r->t_admit;
    times[2] = r->t_done;
  }
  return UGMG_OK;
}

int ugmg_server_stats(const ugmg_server *s, int *groups, long long *launches, long long *columns) {
  if (!s) return UGMG_E_ARG;
  if (groups) {
    int k = 0;
    for (const auto &g : s->groups) k += !g->dead && g->dev_ready;
    *groups = k;
  }
  if (launches) *launches = s->launches;
  if (columns) *columns = s->columns;
  return UGMG_OK;
}

int ugmg_server_result(ugmg_server *s, int id, double *x, int n) {
  if (!s || id < 0 || id >= (int)s->reqs.size() || !x) return UGMG_E_ARG;
  SrvReq *r = s->reqs[id].get();
  if (r->state != UGMG_REQ_DONE) return srv_fail(s, UGMG_E_STATE, "request not done: " + r->err);
  if ((int)r->xout.size() != n) return srv_fail(s, UGMG_E_ARG, "wrong solution length");
  std::copy(r->xout.begin(), r->xout.end(), x);
  return UGMG_OK;
}

}  // extern "C"
