// Device half of the layout build: the colouring, GS device order,
// permutation and SELL-64 slabs of ugmg_setup.cc computed on the GPU, value
// for value (same colouring, window, order, slices, shared offset lists,
// phantoms, padding and transposed transfer), so a level assembled or
// uploaded to the device is laid out without a host pass over its entries.
//
//  * Greedy first-fit colouring in list order (what replaces l_lgs's
//    VINDEX order, np/algebra/ugiter.cc:412-615): vector i takes the lowest
//    colour none of its lower neighbours (w < i, adjacent in the symmetrised
//    pattern) holds.  The device colours in rounds (Jones-Plassmann with the
//    priority -i): a vector joins the frontier once its last lower neighbour
//    is coloured, so the frontier is independent and each of its vectors
//    sees exactly the colours the sequential loop would; the rounds are the
//    longest descending path of the pattern (about x + 2y + 4z on a
//    lexicographic lattice), each O(frontier) work.
//  * Device order (gs_device_order): a stable radix sort of the vectors by
//    (window, colour); the window is chosen by the host builder's heuristic
//    (most common smallest column offset >= 2 over 512 sampled rows, then
//    the candidate with the most uniform slices over 256 sampled slices).
//  * SELL-64 slabs of ConvertMatrix rows (np/udm/disctools.cc:2061-2128):
//    per-slice width, uniform-offset test against the slice's first longest
//    row, offset lists shared by equal lists in order of first appearance
//    (FNV-1a hash, then comparison), phantoms and padding as on the host.
//  * P^T rows in UG's scatter order (np/algebra/transgrid.cc:117-215) by a
//    stable sort of the interpolation entries by coarse vector.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ugmg.h"
#include "ugmg_internal.h"

namespace ugmg {
namespace {

constexpr int kT = 256;
inline unsigned nblk(long long n, int t = kT) {
  return (unsigned)std::max<long long>(1, (n + t - 1) / t);
}

__device__ inline void astore(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// ---------------------------------------------------------------- colouring

// the symmetrised pattern's lower neighbours that row i does not list:
// rows w < i that list i
__global__ void k_tcount(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                         int *__restrict__ cnt) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  for (int m = rp[w]; m < rp[w + 1]; ++m) {
    const int i = ci[m];
    if (i > w) atomicAdd(&cnt[i], 1);
  }
}
__global__ void k_tfill(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                        const int *__restrict__ tptr, int *__restrict__ cur, int *__restrict__ tadj) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  for (int m = rp[w]; m < rp[w + 1]; ++m) {
    const int i = ci[m];
    if (i > w) tadj[tptr[i] + atomicAdd(&cur[i], 1)] = w;
  }
}

// rows w > i that list i (the symmetrised pattern's higher neighbours that
// row i does not list)
__global__ void k_ucount(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                         int *__restrict__ cnt) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  for (int m = rp[w]; m < rp[w + 1]; ++m) {
    const int i = ci[m];
    if (i < w) atomicAdd(&cnt[i], 1);
  }
}
__global__ void k_ufill(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                        const int *__restrict__ uptr, int *__restrict__ cur, int *__restrict__ uadj) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  for (int m = rp[w]; m < rp[w + 1]; ++m) {
    const int i = ci[m];
    if (i < w) uadj[uptr[i] + atomicAdd(&cur[i], 1)] = w;
  }
}

// lower neighbours still uncoloured (all of them at the start); the vectors
// without any form the first frontier
__global__ void k_color_init(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                             const int *__restrict__ tptr, int *__restrict__ indeg,
                             int *__restrict__ front, int *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int d = tptr[i + 1] - tptr[i];
  for (int m = rp[i]; m < rp[i + 1]; ++m) d += ci[m] < i ? 1 : 0;
  indeg[i] = d;
  if (d == 0) front[atomicAdd(cnt, 1)] = i;
}

// lowest colour none of i's lower neighbours holds (all of them coloured)
__device__ int lowest_free(int i, const int *__restrict__ rp, const int *__restrict__ ci,
                           const int *__restrict__ tptr, const int *__restrict__ tadj,
                           const int *__restrict__ color) {
  unsigned long long used = 0;
  int deg = 0;
  for (int m = rp[i]; m < rp[i + 1]; ++m) {
    const int w = ci[m];
    if (w >= i) continue;
    const int c = color[w];
    if (c < 64) used |= 1ull << c;
    ++deg;
  }
  for (int q = tptr[i]; q < tptr[i + 1]; ++q) {
    const int c = color[tadj[q]];
    if (c < 64) used |= 1ull << c;
    ++deg;
  }
  if (~used) return __builtin_ctzll(~used);
  // 0..63 all taken: some colour in 64..64+deg is free
  for (int c = 64; c < 64 + deg; ++c) {
    bool taken = false;
    for (int m = rp[i]; m < rp[i + 1] && !taken; ++m) {
      const int w = ci[m];
      taken = w < i && color[w] == c;
    }
    for (int q = tptr[i]; q < tptr[i + 1] && !taken; ++q) taken = color[tadj[q]] == c;
    if (!taken) return c;
  }
  return 64 + deg;
}

// One round of the greedy colouring in index priority (Jones-Plassmann with
// priority -i): every vector of the frontier has all its lower neighbours
// coloured, so it takes its greedy colour now; a higher neighbour whose last
// uncoloured lower neighbour this was joins the next frontier.  The frontier
// is independent (of two adjacent vectors the higher waits for the lower).
// One wavefront per frontier vector, its lanes over the vector's entries.
// Counters rotate over three slots: this round reads cnt[r % 3], appends to
// cnt[(r + 1) % 3] and clears cnt[(r + 2) % 3] for the next round; cnt[3]
// counts the rounds that coloured something.
__global__ __launch_bounds__(256) void k_color_round(
    int r, const int *__restrict__ rp, const int *__restrict__ ci, const int *__restrict__ tptr,
    const int *__restrict__ tadj, const int *__restrict__ uptr, const int *__restrict__ uadj,
    int *__restrict__ color, int *__restrict__ indeg, const int *__restrict__ fin,
    int *__restrict__ fout, int *__restrict__ cnt) {
  const int m = cnt[r % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[(r + 2) % 3] = 0;
    if (m > 0) cnt[3] += 1;
  }
  int *next = cnt + (r + 1) % 3;
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < m; j += nw) {   // wave-uniform
    const int v = fin[j];
    const int r0 = rp[v], nr = rp[v + 1] - r0, t0 = tptr[v], nt = tptr[v + 1] - t0;
    unsigned long long used = 0;
    for (int q = lane; q < nr + nt; q += 64) {
      const int w = q < nr ? ci[r0 + q] : tadj[t0 + q - nr];
      if (w < v) {
        const int c = color[w];
        if (c < 64) used |= 1ull << c;
      }
    }
    unsigned lo = (unsigned)used, hi = (unsigned)(used >> 32);
    for (int o = 32; o > 0; o >>= 1) {
      lo |= __shfl_xor(lo, o, 64);
      hi |= __shfl_xor(hi, o, 64);
    }
    used = ((unsigned long long)hi << 32) | lo;
    if (lane == 0) color[v] = ~used ? __builtin_ctzll(~used) : lowest_free(v, rp, ci, tptr, tadj, color);
    const int u0 = uptr[v], nu = uptr[v + 1] - u0;
    for (int q = lane; q < nr + nu; q += 64) {
      const int w = q < nr ? ci[r0 + q] : uadj[u0 + q - nr];
      if (w > v && atomicSub(&indeg[w], 1) == 1) fout[atomicAdd(next, 1)] = w;
    }
  }
}

// first position of every colour in the colour-sorted keys (every colour
// below the count occurs: a vector of colour c has neighbours of 0..c-1)
__global__ void k_color_bounds(int n, const unsigned long long *__restrict__ key,
                               int *__restrict__ first) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (j == 0 || key[j] != key[j - 1]) first[key[j]] = j;
}

// ------------------------------------------------------------- device order

// smallest column offset >= 2 of 512 rows spread over the level (0: none)
__global__ void k_win_probe(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                            int *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 512) return;
  const int r = (int)((long long)k * (n - 1) / 511);
  int s = 0;
  for (int m = rp[r]; m < rp[r + 1]; ++m) {
    const int d = ci[m] - r;
    if (d >= 2 && (s == 0 || d < s)) s = d;
  }
  out[k] = s;
}

__global__ void k_order_keys(int n, int win, int ncolors, const int *__restrict__ color,
                             unsigned long long *__restrict__ key, int *__restrict__ val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (unsigned long long)(i / win) * (unsigned long long)ncolors + (unsigned long long)color[i];
  val[i] = i;
}

// pos[perm[p]] = p, colour of device position p, and sort keys by colour
__global__ void k_positions(int n, const int *__restrict__ perm, const int *__restrict__ color,
                            int *__restrict__ pos, int *__restrict__ colp,
                            unsigned long long *__restrict__ key, int *__restrict__ val) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  pos[perm[p]] = p;
  const int c = color[perm[p]];
  colp[p] = c;
  key[p] = (unsigned long long)c;
  val[p] = p;
}

// choose_gs_window's test on sampled slices of the colour-grouped order
// `list` (device positions): every row's offsets an in-order subsequence of
// the slice's first longest row's.  One wavefront per sample.
__global__ void k_win_sample(int nsl, const int *__restrict__ rp, const int *__restrict__ ci,
                             const int *__restrict__ perm, const int *__restrict__ pos,
                             const int *__restrict__ list, int *__restrict__ uni) {
  const int k = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (k >= 256) return;
  const int s0 = (int)((long long)k * nsl / 256) * kSlice;
  const int p = list[s0 + lane], r = perm[p];
  const int len = rp[r + 1] - rp[r];
  const int lw = wave_max(len);
  const unsigned long long at = __ballot(len == lw);
  const int jl = __builtin_ctzll(at);
  const int longest = __shfl(p, jl, 64);
  const int rl = perm[longest];
  int q = rp[rl];
  const int qe = rp[rl + 1];
  bool ok = true;
  for (int m = rp[r]; m < rp[r + 1] && ok; ++m) {
    const int d = pos[ci[m]] - p;
    while (q < qe && pos[ci[q]] - longest != d) ++q;
    if (q == qe) ok = false;
    else ++q;
  }
  const bool all = __ballot(!ok) == 0;
  if (lane == 0 && all) atomicAdd(uni, 1);
}

// ----------------------------------------------------------------- permute

__global__ void k_perm_len(int n, const int *__restrict__ perm, const int *__restrict__ rp,
                           int *__restrict__ len) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > n) return;
  len[p] = p < n ? rp[perm[p] + 1] - rp[perm[p]] : 0;
}

__global__ void k_perm_rows(int n, int bb, const int *__restrict__ perm, const int *__restrict__ pos,
                            const int *__restrict__ rp, const int *__restrict__ ci,
                            const double *__restrict__ a, const uint32_t *__restrict__ skip,
                            const uint8_t *__restrict__ vc, const uint8_t *__restrict__ vnc,
                            const int *__restrict__ rp2, int *__restrict__ ci2, double *__restrict__ a2,
                            uint32_t *__restrict__ skip2, uint8_t *__restrict__ vc2,
                            uint8_t *__restrict__ vnc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int r = perm[p];
  int q = rp2[p];
  for (int m = rp[r]; m < rp[r + 1]; ++m, ++q) {
    ci2[q] = pos[ci[m]];
    for (int e = 0; e < bb; ++e) a2[(size_t)q * bb + e] = a[(size_t)m * bb + e];
  }
  skip2[p] = skip[r];
  vc2[p] = vc[r];
  vnc2[p] = vnc[r];
}

__global__ void k_perm_p(int n, const int *__restrict__ perm, const int *__restrict__ pp,
                         const int *__restrict__ pi, const double *__restrict__ pw,
                         const int *__restrict__ cpos, const int *__restrict__ pp2,
                         int *__restrict__ pi2, double *__restrict__ pw2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int r = perm[p];
  int q = pp2[p];
  for (int m = pp[r]; m < pp[r + 1]; ++m, ++q) {
    pi2[q] = cpos ? cpos[pi[m]] : pi[m];
    pw2[q] = pw[m];
  }
}

// --------------------------------------------------------------- level prep

__global__ void k_diag(int n, int bb, const int *__restrict__ rp, const double *__restrict__ a,
                       double *__restrict__ diag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * bb) return;
  const int i = (int)(t / bb), e = (int)(t % bb);
  diag[t] = a[(size_t)rp[i] * bb + e];
}

__global__ void k_c0_flags(int n, const int *__restrict__ colp, uint8_t *__restrict__ c0,
                           const uint8_t *__restrict__ vc, int *__restrict__ inactive) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < n;
  if (live) c0[p] = colp[p] == 0;
  if (__ballot(live && vc[p] < 3) != 0 && (threadIdx.x & 63) == 0) *inactive = 1;
}

// rows grouped by colour (list = device positions stably sorted by colour),
// each colour from slot 64 color_slices[c]
__global__ void k_group_rows(int n, const int *__restrict__ list, const int *__restrict__ colp,
                             const int *__restrict__ cstart, const int *__restrict__ cslice,
                             int *__restrict__ rows) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int p = list[j], c = colp[p];
  rows[(size_t)cslice[c] * kSlice + (j - cstart[c])] = p;
}

// ----------------------------------------------------------------- SELL-64

enum : int { kAll = 0, kLower = 1, kUpper = 2 };

struct SellIn {
  int n, ncols, bb, nr;
  const int *rp, *ci;
  const double *a;
  const int *rows;   // slot -> row, or null: rows 0..nr-1
  int part;
  const int *color;
  const uint8_t *vc;
  bool uniform;
};

__device__ inline bool keep_entry(const SellIn &S, int r, int m) {
  if (S.part == kAll) return true;
  const int w = S.ci[m];
  if (w == r) return false;
  const int cr = S.color[r], cw = S.color[w];
  if (S.part == kLower) return cw < cr && S.vc[w] >= 3;
  if (S.vc[r] < 3) return true;
  return cw > cr;
}

__device__ inline int slot_row(const SellIn &S, int slot) {
  if (slot >= S.nr) return -1;
  return S.rows ? S.rows[slot] : slot;
}

// per slice (one wavefront): the slot -> row map, width x 64 and the first
// longest row (-1 when the slice keeps no entry)
__global__ void k_sell_width(SellIn S, int ns, int *__restrict__ rows_out, int *__restrict__ w64,
                             int *__restrict__ longest) {
  const int s = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (s >= ns) return;
  const int slot = s * kSlice + lane;
  const int r = slot_row(S, slot);
  rows_out[slot] = r;
  int cnt = 0;
  if (r >= 0)
    for (int m = S.rp[r]; m < S.rp[r + 1]; ++m) cnt += keep_entry(S, r, m) ? 1 : 0;
  const int w = wave_max(cnt);
  const unsigned long long at = __ballot(cnt == w && w > 0);
  const int rl = at ? __shfl(r, __builtin_ctzll(at), 64) : -1;
  if (lane == 0) {
    w64[s] = w * kSlice;
    longest[s] = rl;
  }
}

// uniform test of each slice against its first longest row; a uniform
// slice's offset list goes to ulist at soff[s] / 64 with its FNV-1a hash
__global__ void k_sell_uniform(SellIn S, int ns, const int *__restrict__ rows_out,
                               const int *__restrict__ soff, const int *__restrict__ longest,
                               int *__restrict__ uflag, unsigned long long *__restrict__ hash,
                               int *__restrict__ ulist) {
  const int s = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (s >= ns) return;
  const int w = (soff[s + 1] - soff[s]) / kSlice;
  const int rl = longest[s];
  if (w == 0 || rows_out[(size_t)s * kSlice] < 0 || rl < 0) {
    if (lane == 0) uflag[s] = 0;
    return;
  }
  const int r = rows_out[(size_t)s * kSlice + lane];
  bool ok = true;
  if (r >= 0) {
    for (int m = S.rp[rl]; m < S.rp[rl + 1] && ok; ++m) {
      if (!keep_entry(S, rl, m)) continue;
      const int u = S.ci[m] - rl;
      if (r + u < 0 || r + u >= S.ncols) ok = false;
    }
    int q = S.rp[rl];
    const int qe = S.rp[rl + 1];
    for (int m = S.rp[r]; m < S.rp[r + 1] && ok; ++m) {
      if (!keep_entry(S, r, m)) continue;
      const int d = S.ci[m] - r;
      while (q < qe && (!keep_entry(S, rl, q) || S.ci[q] - rl != d)) ++q;
      if (q == qe) ok = false;
      else ++q;
    }
  }
  const bool all = __ballot(!ok) == 0;
  if (lane != 0) return;
  uflag[s] = all ? 1 : 0;
  if (!all) return;
  unsigned long long hv = 1469598103934665603ULL;
  int k = 0;
  int *out = ulist + soff[s] / kSlice;
  for (int m = S.rp[rl]; m < S.rp[rl + 1]; ++m) {
    if (!keep_entry(S, rl, m)) continue;
    const int u = S.ci[m] - rl;
    out[k++] = u;
    hv = (hv ^ (uint32_t)u) * 1099511628211ULL;
  }
  hash[s] = hv ^ (unsigned long long)k;
}

__device__ inline bool same_list(const int *__restrict__ soff, const int *__restrict__ ulist,
                                 int s, int t) {
  const int ws = soff[s + 1] - soff[s], wt = soff[t + 1] - soff[t];
  if (ws != wt) return false;
  const int *a = ulist + soff[s] / kSlice, *b = ulist + soff[t] / kSlice;
  for (int k = 0; k < ws / kSlice; ++k)
    if (a[k] != b[k]) return false;
  return true;
}

__global__ void k_gather_hash(int m, const int *__restrict__ sel,
                              const unsigned long long *__restrict__ hash,
                              unsigned long long *__restrict__ hk) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) hk[j] = hash[sel[j]];
}

// first index of each run of equal hashes (sorted), as a value to max-scan
__global__ void k_run_heads(int m, const unsigned long long *__restrict__ hs, int *__restrict__ head) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  head[j] = (j == 0 || hs[j] != hs[j - 1]) ? j : 0;
}

// representative (lowest slice with an equal list) of each uniform slice:
// the run's first slice when the lists match, else -1 (k_rep_rest)
__global__ void k_rep_head(int m, const int *__restrict__ ss, const int *__restrict__ head,
                           const int *__restrict__ soff, const int *__restrict__ ulist,
                           int *__restrict__ rep, int *__restrict__ rest) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int s = ss[j], h = ss[head[j]];
  const bool eq = same_list(soff, ulist, s, h);
  rep[s] = eq ? h : -1;
  if (!eq) astore(rest, 1);
}

// hash collisions between different lists (one thread; none in practice):
// the lowest slice of the run whose list equals this one's
__global__ void k_rep_rest(int m, const int *__restrict__ ss, const int *__restrict__ head,
                           const int *__restrict__ soff, const int *__restrict__ ulist,
                           int *__restrict__ rep) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int j = 0; j < m; ++j) {
    const int s = ss[j];
    if (rep[s] >= 0) continue;
    for (int t = head[j]; t <= j; ++t)
      if (same_list(soff, ulist, ss[t], s)) {
        rep[s] = ss[t];
        break;
      }
  }
}

__global__ void k_rep_len(int ns, const int *__restrict__ uflag, const int *__restrict__ rep,
                          const int *__restrict__ soff, int *__restrict__ len) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > ns) return;
  len[s] = (s < ns && uflag[s] && rep[s] == s) ? (soff[s + 1] - soff[s]) / kSlice : 0;
}

__global__ void k_sdoff(int ns, const int *__restrict__ uflag, const int *__restrict__ rep,
                        const int *__restrict__ doff, const int *__restrict__ soff,
                        const int *__restrict__ ulist, int *__restrict__ sdoff, int *__restrict__ dtab) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  if (!uflag[s]) {
    sdoff[s] = -1;
    return;
  }
  sdoff[s] = doff[rep[s]];
  if (rep[s] == s) {
    const int w = (soff[s + 1] - soff[s]) / kSlice;
    for (int k = 0; k < w; ++k) dtab[doff[s] + k] = ulist[soff[s] / kSlice + k];
  }
}

__global__ void k_sell_fill(SellIn S, int ns, size_t nent, const int *__restrict__ rows_out,
                            const int *__restrict__ soff, const int *__restrict__ sdoff,
                            const int *__restrict__ dtab, int *__restrict__ col,
                            double *__restrict__ val) {
  const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= (long long)ns * kSlice) return;
  const int s = (int)(slot / kSlice), lane = (int)(slot % kSlice);
  const int r = rows_out[slot];
  const int w = (soff[s + 1] - soff[s]) / kSlice;
  const size_t base = (size_t)soff[s] + lane;
  auto put = [&](int k, int c, int m) {
    const size_t e = base + (size_t)k * kSlice;
    col[e] = c;
    for (int q = 0; q < S.bb; ++q) val[(size_t)q * nent + e] = m < 0 ? 0.0 : S.a[(size_t)m * S.bb + q];
  };
  int k = 0;
  if (sdoff[s] >= 0) {
    const int *U = dtab + sdoff[s];
    if (r >= 0)
      for (int m = S.rp[r]; m < S.rp[r + 1]; ++m) {
        if (!keep_entry(S, r, m)) continue;
        const int d = S.ci[m] - r;
        while (U[k] != d) put(k++, -1, -1);   // phantoms before this entry
        put(k++, S.ci[m], m);
      }
    for (; k < w; ++k) put(k, -1, -1);
    return;
  }
  int last = 0;
  if (r >= 0)
    for (int m = S.rp[r]; m < S.rp[r + 1]; ++m) {
      if (!keep_entry(S, r, m)) continue;
      put(k++, S.ci[m], m);
      last = S.ci[m];
    }
  for (; k < w; ++k) put(k, last, -1);   // padding: value 0 against the last column
}

// ----------------------------------------------------------- transfer P^T

// the rows P^T takes: node rows of VCLASS >= 2 and every edge row (vt null:
// node rows only)
__global__ void k_entry_rows(int n, const int *__restrict__ pp, const uint8_t *__restrict__ vc,
                             const uint8_t *__restrict__ vt, int *__restrict__ erow,
                             int *__restrict__ flag) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const bool take = vc[v] >= 2 || (vt && vt[v] == 1);
  for (int q = pp[v]; q < pp[v + 1]; ++q) {
    erow[q] = v;
    flag[q] = take;
  }
}

__global__ void k_perm_u8(int n, const int *__restrict__ perm, const uint8_t *__restrict__ src,
                          uint8_t *__restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[p] = src[perm[p]];
}

__global__ void k_pt_keys(int m, const int *__restrict__ sel, const int *__restrict__ pi,
                          unsigned long long *__restrict__ key, int *__restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int c = pi[sel[j]];
  key[j] = (unsigned long long)c;
  atomicAdd(&cnt[c], 1);
}

__global__ void k_pt_fill(int m, const int *__restrict__ sorted, const int *__restrict__ erow,
                          const double *__restrict__ pw, int *__restrict__ idx, double *__restrict__ w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  idx[j] = erow[sorted[j]];
  w[j] = pw[sorted[j]];
}

__global__ void k_iota(int n, int *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// ------------------------------------------------------------ host helpers

// device buffers of one build step: temporaries freed on scope exit
struct Temps {
  std::vector<void *> ptrs;
  ~Temps() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t get(T **p, size_t n) {
    *p = nullptr;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
      ptrs.push_back(q);
      *p = (T *)q;
    }
    return e;
  }
};

struct Ctx {
  hipStream_t st;
  const DevKeep &keep;
  std::string &err;
  int fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? UGMG_E_NOMEM : UGMG_E_HIP;
  }
  template <typename T>
  int own(T **p, size_t n) {
    *p = nullptr;
    void *q = nullptr;
    const int rc = keep.alloc(keep.user, &q, std::max<size_t>(n, 1) * sizeof(T));
    if (rc) {
      err = "device allocation (layout)";
      return rc;
    }
    *p = (T *)q;
    return 0;
  }
};

#define TRY(call, what)                       \
  do {                                        \
    hipError_t e_ = (call);                   \
    if (e_ != hipSuccess) return C.fail(e_, what); \
  } while (0)
#define TRYRC(call)      \
  do {                   \
    int rc_ = (call);    \
    if (rc_) return rc_; \
  } while (0)

int bits_for(unsigned long long maxkey) {
  int b = 1;
  while (b < 64 && (maxkey >> b) != 0) ++b;
  return b;
}

// stable sort of (key, val) pairs by key (LSD radix: equal keys keep their
// input order)
int sort_pairs(Ctx &C, Temps &T, const unsigned long long *kin, unsigned long long *kout,
               const int *vin, int *vout, int m, int end_bit) {
  if (m == 0) return 0;
  size_t tb = 0;
  TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, m, 0, end_bit, C.st),
      "radix sort set-up");
  void *tmp = nullptr;
  TRY(T.get((char **)&tmp, tb), "hipMalloc (sort)");
  TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin, vout, m, 0, end_bit, C.st),
      "radix sort");
  return 0;
}

int excl_scan(Ctx &C, Temps &T, const int *in, int *out, int m) {
  size_t tb = 0;
  TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, m, C.st), "scan set-up");
  void *tmp = nullptr;
  TRY(T.get((char **)&tmp, tb), "hipMalloc (scan)");
  TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, m, C.st), "scan");
  return 0;
}

template <typename T>
int to_host(Ctx &C, T *dst, const T *src, size_t n) {
  if (n == 0) return 0;
  TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, C.st), "copy to host");
  TRY(hipStreamSynchronize(C.st), "synchronize");
  return 0;
}

// one slab (sell_from_csr on the device)
int build_sell_dev(Ctx &C, const SellIn &S, DevSellParts &out) {
  Temps T;
  const int ns = (S.nr + kSlice - 1) / kSlice;
  out.bb = S.bb;
  out.nrows = S.nr;
  out.nslices = ns;
  TRYRC(C.own(&out.rows, (size_t)ns * kSlice));
  TRYRC(C.own(&out.soff, (size_t)ns + 1));
  TRYRC(C.own(&out.sdoff, (size_t)ns));
  int *w64, *longest;
  TRY(T.get(&w64, (size_t)ns + 1), "hipMalloc (layout)");
  TRY(T.get(&longest, (size_t)ns), "hipMalloc (layout)");
  TRY(hipMemsetAsync(w64, 0, sizeof(int) * ((size_t)ns + 1), C.st), "memset");
  const unsigned wb = nblk((long long)ns * 64);
  if (ns > 0) k_sell_width<<<wb, kT, 0, C.st>>>(S, ns, out.rows, w64, longest);
  TRYRC(excl_scan(C, T, w64, out.soff, ns + 1));
  int nent = 0;
  TRYRC(to_host(C, &nent, out.soff + ns, 1));
  out.nent = (size_t)nent;
  int ndtab = 0;
  int *ulist = nullptr, *uflag = nullptr, *rep = nullptr;
  if (S.uniform && ns > 0) {
    unsigned long long *hash;
    TRY(T.get(&ulist, (size_t)nent / kSlice), "hipMalloc (layout)");
    TRY(T.get(&uflag, (size_t)ns), "hipMalloc (layout)");
    TRY(T.get(&rep, (size_t)ns), "hipMalloc (layout)");
    TRY(T.get(&hash, (size_t)ns), "hipMalloc (layout)");
    k_sell_uniform<<<wb, kT, 0, C.st>>>(S, ns, out.rows, out.soff, longest, uflag, hash, ulist);
    // uniform slices in slice order, then stably by hash
    int *iota, *sel, *nsel;
    TRY(T.get(&iota, (size_t)ns), "hipMalloc (layout)");
    TRY(T.get(&sel, (size_t)ns), "hipMalloc (layout)");
    TRY(T.get(&nsel, 1), "hipMalloc (layout)");
    k_iota<<<nblk(ns), kT, 0, C.st>>>(ns, iota);
    size_t tb = 0;
    TRY(hipcub::DeviceSelect::Flagged(nullptr, tb, iota, uflag, sel, nsel, ns, C.st), "select set-up");
    void *tmp;
    TRY(T.get((char **)&tmp, tb), "hipMalloc (select)");
    TRY(hipcub::DeviceSelect::Flagged(tmp, tb, iota, uflag, sel, nsel, ns, C.st), "select");
    int nu = 0;
    TRYRC(to_host(C, &nu, nsel, 1));
    if (nu > 0) {
      unsigned long long *hk, *hs;
      int *ss, *head, *rest;
      TRY(T.get(&hk, (size_t)nu), "hipMalloc (layout)");
      TRY(T.get(&hs, (size_t)nu), "hipMalloc (layout)");
      TRY(T.get(&ss, (size_t)nu), "hipMalloc (layout)");
      TRY(T.get(&head, (size_t)nu), "hipMalloc (layout)");
      TRY(T.get(&rest, 1), "hipMalloc (layout)");
      // keys: the hashes of the selected slices
      TRY(hipMemsetAsync(rest, 0, sizeof(int), C.st), "memset");
      k_gather_hash<<<nblk(nu), kT, 0, C.st>>>(nu, sel, hash, hk);
      TRYRC(sort_pairs(C, T, hk, hs, sel, ss, nu, 64));
      int *h0;
      TRY(T.get(&h0, (size_t)nu), "hipMalloc (layout)");
      k_run_heads<<<nblk(nu), kT, 0, C.st>>>(nu, hs, h0);
      {
        size_t tb2 = 0;
        TRY(hipcub::DeviceScan::InclusiveScan(nullptr, tb2, h0, head, hipcub::Max(), nu, C.st),
            "scan set-up");
        void *tmp2;
        TRY(T.get((char **)&tmp2, tb2), "hipMalloc (scan)");
        TRY(hipcub::DeviceScan::InclusiveScan(tmp2, tb2, h0, head, hipcub::Max(), nu, C.st), "scan");
      }
      k_rep_head<<<nblk(nu), kT, 0, C.st>>>(nu, ss, head, out.soff, ulist, rep, rest);
      int hrest = 0;
      TRYRC(to_host(C, &hrest, rest, 1));
      if (hrest) k_rep_rest<<<1, 1, 0, C.st>>>(nu, ss, head, out.soff, ulist, rep);
    }
    int *len, *doff;
    TRY(T.get(&len, (size_t)ns + 1), "hipMalloc (layout)");
    TRY(T.get(&doff, (size_t)ns + 1), "hipMalloc (layout)");
    k_rep_len<<<nblk(ns + 1), kT, 0, C.st>>>(ns, uflag, rep, out.soff, len);
    TRYRC(excl_scan(C, T, len, doff, ns + 1));
    TRYRC(to_host(C, &ndtab, doff + ns, 1));
    TRYRC(C.own(&out.dtab, (size_t)std::max(ndtab, 1)));
    TRY(hipMemsetAsync(out.dtab, 0, sizeof(int), C.st), "memset");
    k_sdoff<<<nblk(ns), kT, 0, C.st>>>(ns, uflag, rep, doff, out.soff, ulist, out.sdoff, out.dtab);
  } else {
    TRYRC(C.own(&out.dtab, 1));
    TRY(hipMemsetAsync(out.dtab, 0, sizeof(int), C.st), "memset");
    TRY(hipMemsetAsync(out.sdoff, 0xff, sizeof(int) * std::max(ns, 1), C.st), "memset");
  }
  out.ndtab = (size_t)std::max(ndtab, 1);
  TRYRC(C.own(&out.col, out.nent));
  TRYRC(C.own(&out.val, out.nent * S.bb));
  if (ns > 0)
    k_sell_fill<<<wb, kT, 0, C.st>>>(S, ns, out.nent, out.rows, out.soff, out.sdoff, out.dtab,
                                     out.col, out.val);
  TRY(hipGetLastError(), "layout kernels");
  out.soff_h.resize((size_t)ns + 1);
  out.sdoff_h.resize((size_t)ns);
  TRYRC(to_host(C, out.soff_h.data(), out.soff, (size_t)ns + 1));
  TRYRC(to_host(C, out.sdoff_h.data(), out.sdoff, (size_t)ns));
  return 0;
}

int sell_csr_dev(Ctx &C, int n, const int *rp, const int *ci, const double *a, DevSellParts &out) {
  SellIn S{n, 0, 1, n, rp, ci, a, nullptr, kAll, nullptr, nullptr, false};
  return build_sell_dev(C, S, out);
}

// greedy colouring (see the file comment); colors: n ints (device),
// *rounds: the frontier rounds it took
int color_dev(Ctx &C, int n, const int *rp, const int *ci, int *colors, int *rounds) {
  Temps T;
  *rounds = 0;
  if (n == 0) return 0;
  int *tcnt, *tptr, *tcur, *ucnt, *uptr, *ucur, *indeg, *f0, *f1, *cnt;
  TRY(T.get(&tcnt, (size_t)n + 1), "hipMalloc (colouring)");
  TRY(T.get(&tptr, (size_t)n + 1), "hipMalloc (colouring)");
  TRY(T.get(&tcur, (size_t)n), "hipMalloc (colouring)");
  TRY(T.get(&ucnt, (size_t)n + 1), "hipMalloc (colouring)");
  TRY(T.get(&uptr, (size_t)n + 1), "hipMalloc (colouring)");
  TRY(T.get(&ucur, (size_t)n), "hipMalloc (colouring)");
  TRY(T.get(&indeg, (size_t)n), "hipMalloc (colouring)");
  TRY(T.get(&f0, (size_t)n), "hipMalloc (colouring)");
  TRY(T.get(&f1, (size_t)n), "hipMalloc (colouring)");
  TRY(T.get(&cnt, 4), "hipMalloc (colouring)");
  for (int *z : {tcnt, ucnt}) TRY(hipMemsetAsync(z, 0, sizeof(int) * ((size_t)n + 1), C.st), "memset");
  for (int *z : {tcur, ucur}) TRY(hipMemsetAsync(z, 0, sizeof(int) * (size_t)n, C.st), "memset");
  TRY(hipMemsetAsync(cnt, 0, sizeof(int) * 4, C.st), "memset");
  k_tcount<<<nblk(n), kT, 0, C.st>>>(n, rp, ci, tcnt);
  k_ucount<<<nblk(n), kT, 0, C.st>>>(n, rp, ci, ucnt);
  TRYRC(excl_scan(C, T, tcnt, tptr, n + 1));
  TRYRC(excl_scan(C, T, ucnt, uptr, n + 1));
  int nt[2] = {0, 0};
  TRYRC(to_host(C, &nt[0], tptr + n, 1));
  TRYRC(to_host(C, &nt[1], uptr + n, 1));
  int *tadj, *uadj;
  TRY(T.get(&tadj, (size_t)nt[0]), "hipMalloc (colouring)");
  TRY(T.get(&uadj, (size_t)nt[1]), "hipMalloc (colouring)");
  k_tfill<<<nblk(n), kT, 0, C.st>>>(n, rp, ci, tptr, tcur, tadj);
  k_ufill<<<nblk(n), kT, 0, C.st>>>(n, rp, ci, uptr, ucur, uadj);
  k_color_init<<<nblk(n), kT, 0, C.st>>>(n, rp, ci, tptr, indeg, f0, cnt);
  TRY(hipGetLastError(), "colouring kernels");
  // rounds in batches; a round that finds an empty frontier ends the
  // colouring (every vector has been coloured: the lower-neighbour relation
  // is acyclic).  Each non-empty round colours at least one vector.
  const unsigned grid = std::min<unsigned>(nblk((long long)n * 64), 1024);   // a wavefront per vector
  constexpr int kBatch = 32;
  for (long long r = 0; r <= (long long)n + kBatch; r += kBatch) {
    for (int k = 0; k < kBatch; ++k) {
      const int q = (int)((r + k) % 6);   // round index mod 6: frontier buffer and counter slot
      k_color_round<<<grid, kT, 0, C.st>>>(q, rp, ci, tptr, tadj, uptr, uadj, colors, indeg,
                                           (q & 1) ? f1 : f0, (q & 1) ? f0 : f1, cnt);
    }
    int left = 0;
    TRYRC(to_host(C, &left, cnt + (int)((r + kBatch) % 3), 1));
    if (left == 0) {
      int hc[4];
      TRYRC(to_host(C, hc, cnt, 4));
      *rounds = hc[3];
      return 0;
    }
  }
  C.err = "colouring did not finish";
  return UGMG_E_HIP;
}

int max_color(Ctx &C, int n, const int *colors, int *ncol) {
  Temps T;
  *ncol = 0;
  if (n == 0) return 0;
  int *mx;
  TRY(T.get(&mx, 1), "hipMalloc");
  size_t tb = 0;
  TRY(hipcub::DeviceReduce::Max(nullptr, tb, colors, mx, n, C.st), "reduce set-up");
  void *tmp;
  TRY(T.get((char **)&tmp, tb), "hipMalloc (reduce)");
  TRY(hipcub::DeviceReduce::Max(tmp, tb, colors, mx, n, C.st), "reduce");
  int h = 0;
  TRYRC(to_host(C, &h, mx, 1));
  *ncol = h + 1;
  return 0;
}

// the device order for window `win` (0: whole level): perm, pos, colour per
// device position and the colour-grouped list of positions
int order_dev(Ctx &C, Temps &T, int n, int ncolors, int win, const int *color, int *perm, int *pos,
              int *colp, int *list, int *first = nullptr) {
  const int W = win > 0 ? win : std::max(n, 1);
  unsigned long long *k0, *k1;
  int *v0;
  TRY(T.get(&k0, (size_t)n), "hipMalloc (order)");
  TRY(T.get(&k1, (size_t)n), "hipMalloc (order)");
  TRY(T.get(&v0, (size_t)n), "hipMalloc (order)");
  k_order_keys<<<nblk(n), kT, 0, C.st>>>(n, W, ncolors, color, k0, v0);
  const unsigned long long mk =
      (unsigned long long)((n - 1) / W) * (unsigned long long)ncolors + (unsigned long long)(ncolors - 1);
  TRYRC(sort_pairs(C, T, k0, k1, v0, perm, n, bits_for(mk)));
  k_positions<<<nblk(n), kT, 0, C.st>>>(n, perm, color, pos, colp, k0, v0);
  TRYRC(sort_pairs(C, T, k0, k1, v0, list, n, bits_for((unsigned long long)std::max(ncolors - 1, 0))));
  if (first) k_color_bounds<<<nblk(n), kT, 0, C.st>>>(n, k1, first);
  return 0;
}

int choose_window_dev(Ctx &C, int n, const int *rp, const int *ci, const int *color, int ncolors,
                      int *win_out) {
  *win_out = 0;
  if (n < 4 * kSlice) return 0;
  Temps T;
  int *probe;
  TRY(T.get(&probe, 512), "hipMalloc (window)");
  k_win_probe<<<2, kT, 0, C.st>>>(n, rp, ci, probe);
  std::vector<int> hp(512);
  TRYRC(to_host(C, hp.data(), probe, 512));
  std::vector<std::pair<int, int>> hist;
  for (int s : hp) {
    if (s == 0) continue;
    bool f = false;
    for (auto &h : hist)
      if (h.first == s) {
        ++h.second;
        f = true;
      }
    if (!f) hist.emplace_back(s, 1);
  }
  int best = 0, bc = 0;
  for (auto &h : hist)
    if (h.second > bc) {
      bc = h.second;
      best = h.first;
    }
  std::vector<int> cand;
  if (best > 0) {
    cand.push_back(best);
    cand.push_back(best + 1);
  }
  cand.push_back(0);
  int *perm, *pos, *colp, *list, *uni;
  TRY(T.get(&perm, (size_t)n), "hipMalloc (window)");
  TRY(T.get(&pos, (size_t)n), "hipMalloc (window)");
  TRY(T.get(&colp, (size_t)n), "hipMalloc (window)");
  TRY(T.get(&list, (size_t)n), "hipMalloc (window)");
  TRY(T.get(&uni, 1), "hipMalloc (window)");
  const int nsl = n / kSlice;
  double best_score = -1.0;
  for (int W : cand) {
    Temps T2;
    TRYRC(order_dev(C, T2, n, ncolors, W, color, perm, pos, colp, list));
    int u = 0, tried = 0;
    if (nsl > 0) {
      TRY(hipMemsetAsync(uni, 0, sizeof(int), C.st), "memset");
      k_win_sample<<<256 * 64 / kT, kT, 0, C.st>>>(nsl, rp, ci, perm, pos, list, uni);
      TRYRC(to_host(C, &u, uni, 1));
      tried = 256;
    }
    const double score = tried ? (double)u / tried : 0.0;
    if (score > best_score + 1e-9) {
      best_score = score;
      *win_out = W;
    }
  }
  return 0;
}

template <typename T>
int up(Ctx &C, Temps &Tm, T **d, const T *h, size_t n) {
  TRY(Tm.get(d, n), "hipMalloc (level upload)");
  if (n) TRY(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, C.st), "level upload");
  return 0;
}
template <typename T>
int up_own(Ctx &C, T **d, const T *h, size_t n) {
  TRYRC(C.own(d, n));
  if (n) TRY(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, C.st), "level upload");
  return 0;
}

}  // namespace

int color_device(void *stream, int n, const int *rp, const int *ci, int *colors, int *rounds,
                 std::string &err) {
  DevKeep none;
  Ctx C{(hipStream_t)stream, none, err};
  int rc = color_dev(C, n, rp, ci, colors, rounds);
  if (rc) return rc;
  int nc = 0;
  rc = max_color(C, n, colors, &nc);
  return rc ? rc : nc;
}

int build_level_device(void *stream, const HostLevel &H, int coarse_n, const int *coarse_pos,
                       bool gs, int force_win, const DevKeep &keep, DevPrep &out,
                       std::string &err, const int *drp, const int *dci, const double *da) {
  Ctx C{(hipStream_t)stream, keep, err};
  Temps T;
  const int n = H.n, bb = H.b * H.b;
  out = DevPrep();
  out.n = n;
  out.nnzb = H.nnzb;
  // the level as it crosses the ABI (list order)
  const int *rp, *ci;
  const double *a;
  uint32_t *skip;
  uint8_t *vc, *vnc;
  if (drp && dci && da) {   // resident on the device (assembled there)
    rp = drp;
    ci = dci;
    a = da;
  } else {
    int *urp, *uci;
    double *ua;
    TRYRC(up(C, T, &urp, H.rp.data(), H.rp.size()));
    TRYRC(up(C, T, &uci, H.ci.data(), H.ci.size()));
    TRYRC(up(C, T, &ua, H.a.data(), H.a.size()));
    rp = urp;
    ci = uci;
    a = ua;
  }
  int *pp = nullptr, *pi = nullptr;
  double *pw = nullptr;
  const bool has_p = coarse_n > 0 && H.has_p;
  if (has_p) {
    TRYRC(up(C, T, &pp, H.pp.data(), H.pp.size()));
    TRYRC(up(C, T, &pi, H.pi.data(), H.pi.size()));
    TRYRC(up(C, T, &pw, H.p_weights().data(), H.p_weights().size()));
  }
  // level in device order (GS levels) or as given
  const int *Lrp = rp, *Lci = ci, *Lpp = pp, *Lpi = pi;
  const double *La = a, *Lpw = pw;
  int *colp = nullptr;
  if (!gs) {
    TRYRC(up_own(C, &out.skip, H.skip.data(), H.skip.size()));
    TRYRC(up_own(C, &out.vc, H.vc.data(), H.vc.size()));
    TRYRC(up_own(C, &out.vnc, H.vnc.data(), H.vnc.size()));
    if (!H.vt.empty()) TRYRC(up_own(C, &out.vt, H.vt.data(), H.vt.size()));
  } else {
    TRYRC(up(C, T, &skip, H.skip.data(), H.skip.size()));
    TRYRC(up(C, T, &vc, H.vc.data(), H.vc.size()));
    TRYRC(up(C, T, &vnc, H.vnc.data(), H.vnc.size()));
    int *color;
    TRY(T.get(&color, (size_t)n), "hipMalloc (colouring)");
    int rounds = 0;
    TRYRC(color_dev(C, n, rp, ci, color, &rounds));
    TRYRC(max_color(C, n, color, &out.ncolors));
    int win = force_win;
    if (win < 0) TRYRC(choose_window_dev(C, n, rp, ci, color, out.ncolors, &win));
    out.window = win;
    int *list, *first;
    TRYRC(C.own(&out.perm, (size_t)n));
    TRY(hipMalloc((void **)&out.pos, sizeof(int) * std::max(n, 1)), "hipMalloc (order)");
    TRY(T.get(&colp, (size_t)n), "hipMalloc (order)");
    TRY(T.get(&list, (size_t)n), "hipMalloc (order)");
    TRY(T.get(&first, (size_t)out.ncolors), "hipMalloc (order)");
    {
      Temps T2;
      TRYRC(order_dev(C, T2, n, out.ncolors, win, color, out.perm, out.pos, colp, list, first));
    }
    // permute_level
    int *len, *rp2, *ci2;
    double *a2;
    TRY(T.get(&len, (size_t)n + 1), "hipMalloc (permute)");
    TRY(T.get(&rp2, (size_t)n + 1), "hipMalloc (permute)");
    TRY(T.get(&ci2, H.ci.size()), "hipMalloc (permute)");
    TRY(T.get(&a2, H.a.size()), "hipMalloc (permute)");
    TRYRC(C.own(&out.skip, (size_t)n));
    TRYRC(C.own(&out.vc, (size_t)n));
    TRYRC(C.own(&out.vnc, (size_t)n));
    k_perm_len<<<nblk(n + 1), kT, 0, C.st>>>(n, out.perm, rp, len);
    TRYRC(excl_scan(C, T, len, rp2, n + 1));
    k_perm_rows<<<nblk(n), kT, 0, C.st>>>(n, bb, out.perm, out.pos, rp, ci, a, skip, vc, vnc, rp2, ci2,
                                          a2, out.skip, out.vc, out.vnc);
    if (!H.vt.empty()) {
      uint8_t *vt;
      TRYRC(up(C, T, &vt, H.vt.data(), H.vt.size()));
      TRYRC(C.own(&out.vt, (size_t)n));
      k_perm_u8<<<nblk(n), kT, 0, C.st>>>(n, out.perm, vt, out.vt);
    }
    Lrp = rp2;
    Lci = ci2;
    La = a2;
    if (has_p) {
      int *pp2, *pi2;
      double *pw2;
      TRY(T.get(&pp2, (size_t)n + 1), "hipMalloc (permute)");
      TRY(T.get(&pi2, H.pi.size()), "hipMalloc (permute)");
      TRY(T.get(&pw2, H.p_weights().size()), "hipMalloc (permute)");
      k_perm_len<<<nblk(n + 1), kT, 0, C.st>>>(n, out.perm, pp, len);
      TRYRC(excl_scan(C, T, len, pp2, n + 1));
      k_perm_p<<<nblk(n), kT, 0, C.st>>>(n, out.perm, pp, pi, pw, coarse_pos, pp2, pi2, pw2);
      Lpp = pp2;
      Lpi = pi2;
      Lpw = pw2;
    }
    TRY(hipGetLastError(), "permute kernels");
    // rows grouped by colour, each colour padded to whole slices
    int *inactive;
    TRY(T.get(&inactive, 1), "hipMalloc (groups)");
    TRYRC(C.own(&out.c0, (size_t)n));
    TRY(hipMemsetAsync(inactive, 0, sizeof(int), C.st), "memset");
    k_c0_flags<<<nblk(n), kT, 0, C.st>>>(n, colp, out.c0, out.vc, inactive);
    std::vector<int> fh(out.ncolors + 1, n);
    TRYRC(to_host(C, fh.data(), first, (size_t)out.ncolors));
    out.color_rows.resize(out.ncolors);
    for (int c = 0; c < out.ncolors; ++c) out.color_rows[c] = fh[c + 1] - fh[c];
    int hin = 0;
    TRYRC(to_host(C, &hin, inactive, 1));
    out.all_active = hin == 0;
    out.color_slices.assign(out.ncolors + 1, 0);
    std::vector<int> cstart(out.ncolors, 0);
    for (int c = 0; c < out.ncolors; ++c) {
      out.color_slices[c + 1] = out.color_slices[c] + (out.color_rows[c] + kSlice - 1) / kSlice;
      if (c > 0) cstart[c] = cstart[c - 1] + out.color_rows[c - 1];
    }
    const int nrows = out.color_slices[out.ncolors] * kSlice;
    int *rows, *dcs, *dcsl;
    TRY(T.get(&rows, (size_t)nrows), "hipMalloc (groups)");
    TRYRC(up(C, T, &dcs, cstart.data(), cstart.size()));
    TRYRC(up(C, T, &dcsl, out.color_slices.data(), out.color_slices.size()));
    TRY(hipMemsetAsync(rows, 0xff, sizeof(int) * std::max(nrows, 1), C.st), "memset");
    k_group_rows<<<nblk(n), kT, 0, C.st>>>(n, list, colp, dcs, dcsl, rows);
    TRY(hipGetLastError(), "group kernels");
    SellIn S{n, n, bb, nrows, Lrp, Lci, La, rows, kLower, colp, out.vc, true};
    TRYRC(build_sell_dev(C, S, out.L));
    S.part = kUpper;
    TRYRC(build_sell_dev(C, S, out.U));
    S.part = kAll;
    TRYRC(build_sell_dev(C, S, out.A));
    out.perm_h.resize(n);
    out.color_ug_h.resize(n);
    TRYRC(to_host(C, out.perm_h.data(), out.perm, (size_t)n));
    TRYRC(to_host(C, out.color_ug_h.data(), color, (size_t)n));
  }
  TRYRC(C.own(&out.diag, (size_t)n * bb));
  k_diag<<<nblk((long long)n * bb), kT, 0, C.st>>>(n, bb, Lrp, La, out.diag);
  if (!gs) {
    SellIn S{n, n, bb, n, Lrp, Lci, La, nullptr, kAll, nullptr, nullptr, true};
    TRYRC(build_sell_dev(C, S, out.A));
  }
  if (has_p) {
    // P^T rows: the entries of the fine vectors the restriction reads (node
    // vectors of VCLASS >= 2, edge vectors), stably by coarse vector
    const int pn = H.pp[n];
    out.p_nnz = pn;
    int *erow, *flag, *iota, *sel, *nsel, *cnt, *ptr, *sorted;
    unsigned long long *k0, *k1;
    TRY(T.get(&erow, (size_t)pn), "hipMalloc (transfer)");
    TRY(T.get(&flag, (size_t)pn), "hipMalloc (transfer)");
    TRY(T.get(&iota, (size_t)pn), "hipMalloc (transfer)");
    TRY(T.get(&sel, (size_t)pn), "hipMalloc (transfer)");
    TRY(T.get(&nsel, 1), "hipMalloc (transfer)");
    TRY(T.get(&cnt, (size_t)coarse_n + 1), "hipMalloc (transfer)");
    TRY(T.get(&ptr, (size_t)coarse_n + 1), "hipMalloc (transfer)");
    TRY(T.get(&sorted, (size_t)pn), "hipMalloc (transfer)");
    TRY(T.get(&k0, (size_t)pn), "hipMalloc (transfer)");
    TRY(T.get(&k1, (size_t)pn), "hipMalloc (transfer)");
    k_entry_rows<<<nblk(n), kT, 0, C.st>>>(n, Lpp, out.vc, out.vt, erow, flag);
    k_iota<<<nblk(pn), kT, 0, C.st>>>(pn, iota);
    int m = 0;
    if (pn > 0) {
      size_t tb = 0;
      TRY(hipcub::DeviceSelect::Flagged(nullptr, tb, iota, flag, sel, nsel, pn, C.st), "select set-up");
      void *tmp;
      TRY(T.get((char **)&tmp, tb), "hipMalloc (select)");
      TRY(hipcub::DeviceSelect::Flagged(tmp, tb, iota, flag, sel, nsel, pn, C.st), "select");
      TRYRC(to_host(C, &m, nsel, 1));
    }
    out.pt_nnz = m;
    TRY(hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)coarse_n + 1), C.st), "memset");
    k_pt_keys<<<nblk(m), kT, 0, C.st>>>(m, sel, Lpi, k0, cnt);
    TRYRC(sort_pairs(C, T, k0, k1, sel, sorted, m, bits_for((unsigned long long)std::max(coarse_n - 1, 0))));
    TRYRC(excl_scan(C, T, cnt, ptr, coarse_n + 1));
    int *pidx;
    double *pwt;
    TRY(T.get(&pidx, (size_t)m), "hipMalloc (transfer)");
    TRY(T.get(&pwt, (size_t)m), "hipMalloc (transfer)");
    k_pt_fill<<<nblk(m), kT, 0, C.st>>>(m, sorted, erow, Lpw, pidx, pwt);
    TRY(hipGetLastError(), "transfer kernels");
    TRYRC(sell_csr_dev(C, n, Lpp, Lpi, Lpw, out.P));
    TRYRC(sell_csr_dev(C, coarse_n, ptr, pidx, pwt, out.PT));
  }
  TRY(hipStreamSynchronize(C.st), "layout");
  return 0;
}

}  // namespace ugmg
