// Host half of LmgcPreProcess: device layouts, colouring, coarse LU.
//
//  * SELL-64 row slabs of the ConvertMatrix arrays (np/udm/disctools.cc:
//    2061-2128); within a row the entries keep their list order (diagonal
//    first), so every row sum runs in UG's order (matloop.ct:75-99).
//  * greedy first-fit colouring in vector-list order over the symmetrised
//    matrix pattern; colour-ordered forward substitution is l_lgs
//    (np/algebra/ugiter.cc:412-615) with VINDEX replaced by (colour, list
//    position).
//  * P^T rows for the gather form of StandardRestrictNodeVector
//    (np/algebra/transgrid.cc:117-215): only fine vectors with VCLASS >= 2
//    contribute, entries ordered by fine list position (UG's scatter order).
//  * EXPreProcess (np/procs/iter.cc:8685-8970): breadth-first band reorder,
//    band copy (EX_MAT, gm/evm.h:392), LU without pivoting
//    (gm/evm.cc:1400-1418); the factors are then applied to unit vectors to
//    form the dense inverse that the GPU multiplies with.
#include "ugmg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace ugmg {

// SELL-64 slabs of the rows `rows` (or 0..n-1) of a CSR matrix, keeping the
// entries for which keep(row, entry) holds, in CSR order.
// SELL-64 slabs of the rows `rows` (or 0..n-1) of a CSR matrix with ncols
// columns, keeping the entries for which keep(row, entry) holds, in CSR
// order.  With `uniform`, a slice whose rows all list their columns as an
// in-order subsequence of one common offset sequence (column - row) is stored
// as values only against that sequence.
template <class Keep>
static void sell_from_csr(int n, int ncols, const int *rp, const int *ci, const double *a,
                          int bb, const std::vector<int> *rows, Keep keep, bool uniform,
                          HostSell &out) {
  const int nr = rows ? (int)rows->size() : n;
  out.bb = bb;
  out.nrows = nr;
  out.identity_rows = rows == nullptr;
  out.nslices = (nr + kSlice - 1) / kSlice;
  const int ns = out.nslices;
  out.soff.assign(ns + 1, 0);
  out.rows.assign((size_t)ns * kSlice, -1);
  out.sdoff.assign(ns, -1);
  std::vector<int> width(ns, 0);
  std::vector<std::vector<int>> sdel(ns);   // common offsets of uniform slices
  std::vector<uint64_t> shash(ns, 0);       // FNV-1a of each list
#pragma omp parallel
  {
  std::vector<int> U;   // per-thread scratch: the candidate offset list of a slice
  U.reserve(64);
#pragma omp for schedule(static)
  for (int s = 0; s < ns; ++s) {
    int wmax = 0, longest = -1;
    for (int lane = 0; lane < kSlice; ++lane) {
      const int slot = s * kSlice + lane;
      if (slot >= nr) break;
      const int r = rows ? (*rows)[slot] : slot;
      out.rows[slot] = r;
      if (r < 0) continue;  // padding slot
      int cnt = 0;
      for (int m = rp[r]; m < rp[r + 1]; ++m) cnt += keep(r, m) ? 1 : 0;
      if (cnt > wmax) { wmax = cnt; longest = r; }
    }
    width[s] = wmax;
    if (!uniform || wmax == 0 || out.rows[(size_t)s * kSlice] < 0) continue;
    U.clear();
    for (int m = rp[longest]; m < rp[longest + 1]; ++m)
      if (keep(longest, m)) U.push_back(ci[m] - longest);
    bool ok = true;
    for (int lane = 0; lane < kSlice && ok; ++lane) {
      const int r = out.rows[(size_t)s * kSlice + lane];
      if (r < 0) continue;
      for (int u : U)
        if (r + u < 0 || r + u >= ncols) { ok = false; break; }
      size_t k = 0;
      for (int m = rp[r]; m < rp[r + 1] && ok; ++m) {
        if (!keep(r, m)) continue;
        const int d = ci[m] - r;
        while (k < U.size() && U[k] != d) ++k;
        if (k == U.size()) ok = false;
        else ++k;
      }
    }
    if (ok) {
      uint64_t hv = 1469598103934665603ULL;
      for (int u : U) hv = (hv ^ (uint32_t)u) * 1099511628211ULL;
      shash[s] = hv ^ U.size();
      sdel[s].assign(U.begin(), U.end());
    }
  }
  }
  for (int s = 0; s < ns; ++s) out.soff[s + 1] = out.soff[s] + width[s] * kSlice;
  // shared offset lists (most slices of a stencil matrix share one): equal
  // lists found by hash, confirmed by comparison
  out.dtab.clear();
  {
    // hash -> (dtab offset, length) of the lists stored so far
    std::unordered_map<uint64_t, std::vector<std::pair<int, int>>> seen;
    for (int s = 0; s < ns; ++s) {
      if (sdel[s].empty()) continue;
      const std::vector<int> &U = sdel[s];
      int off = -1;
      auto &cands = seen[shash[s]];
      for (const auto &o : cands)
        if (o.second == (int)U.size() && std::equal(U.begin(), U.end(), out.dtab.begin() + o.first)) {
          off = o.first;
          break;
        }
      if (off < 0) {
        off = (int)out.dtab.size();
        out.dtab.insert(out.dtab.end(), U.begin(), U.end());
        cands.emplace_back(off, (int)U.size());
      }
      out.sdoff[s] = off;
    }
    if (out.dtab.empty()) out.dtab.push_back(0);
  }
  const size_t nent = (size_t)out.soff[ns];
  out.col.resize(nent);
  out.val.resize(nent * bb);
  auto put = [&](size_t e, int m) {
    for (int q = 0; q < bb; ++q) out.val[(size_t)q * nent + e] = m < 0 ? 0.0 : a[(size_t)m * bb + q];
  };
#pragma omp parallel for schedule(static)
  for (int s = 0; s < ns; ++s) {
    const int w = width[s];
    for (int lane = 0; lane < kSlice; ++lane) {
      const int slot = s * kSlice + lane;
      const int r = slot < nr ? out.rows[slot] : -1;
      if (out.sdoff[s] >= 0) {
        const int *U = out.dtab.data() + out.sdoff[s];
        int k = 0;
        if (r >= 0)
          for (int m = rp[r]; m < rp[r + 1]; ++m) {
            if (!keep(r, m)) continue;
            const int d = ci[m] - r;
            while (U[k] != d) {  // phantom entries before this one
              const size_t e = (size_t)out.soff[s] + (size_t)k * kSlice + lane;
              out.col[e] = -1;
              put(e, -1);
              ++k;
            }
            const size_t e = (size_t)out.soff[s] + (size_t)k * kSlice + lane;
            out.col[e] = ci[m];
            put(e, m);
            ++k;
          }
        for (; k < w; ++k) {
          const size_t e = (size_t)out.soff[s] + (size_t)k * kSlice + lane;
          out.col[e] = -1;
          put(e, -1);
        }
        continue;
      }
      int k = 0;
      int last = 0;
      if (r >= 0) {
        for (int m = rp[r]; m < rp[r + 1]; ++m) {
          if (!keep(r, m)) continue;
          const size_t e = (size_t)out.soff[s] + (size_t)k * kSlice + lane;
          out.col[e] = ci[m];
          put(e, m);
          last = ci[m];
          ++k;
        }
      }
      // padding: value 0 against the row's last column (a line already loaded)
      for (; k < w; ++k) {
        const size_t e = (size_t)out.soff[s] + (size_t)k * kSlice + lane;
        out.col[e] = last;
        put(e, -1);
      }
    }
  }
}

void build_sell(const HostLevel &L, const std::vector<int> *rows, Part part,
                const std::vector<int> *color, HostSell &out) {
  auto keep = [&](int r, int m) {
    if (part == Part::All) return true;
    const int w = L.ci[m];
    if (w == r) return false;
    const int cr = (*color)[r], cw = (*color)[w];
    if (part == Part::Lower) return cw < cr && L.vc[w] >= 3;
    // Upper: higher colours for active rows, every off-diagonal entry for
    // rows that the smoother leaves at zero (VCLASS < 3)
    if (L.vc[r] < 3) return true;
    return cw > cr;
  };
  sell_from_csr(L.n, L.n, L.rp.data(), L.ci.data(), L.a.data(), L.b * L.b, rows, keep, true, out);
}

void build_sell_csr(int n, const int *rp, const int *ci, const double *a, HostSell &out, int bb) {
  sell_from_csr(n, 0, rp, ci, a, bb, nullptr, [](int, int) { return true; }, false, out);
}

void to_float(const double *in, size_t n, hvec<float> &out) {
  out.resize(n);
  const long long ne = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < ne; ++i) out[i] = (float)in[i];
}

int greedy_color(const HostLevel &L, std::vector<int> &color) {
  // First fit in list order over the symmetrised pattern.  When vector i is
  // coloured only its neighbours w < i carry a colour, so it looks at the
  // lower entries of its own row and at the rows w < i that list i (the
  // transposed lower part, gathered here).
  const int n = L.n;
  std::vector<int> tptr(n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < n; ++w)
    for (int m = L.rp[w]; m < L.rp[w + 1]; ++m) {
      const int i = L.ci[m];
      if (i > w) __atomic_fetch_add(&tptr[i + 1], 1, __ATOMIC_RELAXED);
    }
  for (int i = 0; i < n; ++i) tptr[i + 1] += tptr[i];
  hvec<int> tadj(tptr[n]);
  std::vector<int> cur(tptr.begin(), tptr.end() - 1);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < n; ++w)
    for (int m = L.rp[w]; m < L.rp[w + 1]; ++m) {
      const int i = L.ci[m];
      if (i > w) tadj[__atomic_fetch_add(&cur[i], 1, __ATOMIC_RELAXED)] = w;
    }
  color.assign(n, -1);
  int ncol = 0;
  std::vector<int> seen(64, -1);
  auto mark = [&](int c, int i) {
    if (c >= (int)seen.size()) seen.resize(c + 1, -1);
    seen[c] = i;
  };
  for (int i = 0; i < n; ++i) {
    for (int m = L.rp[i]; m < L.rp[i + 1]; ++m) {
      const int w = L.ci[m];
      if (w < i) mark(color[w], i);
    }
    for (int q = tptr[i]; q < tptr[i + 1]; ++q) mark(color[tadj[q]], i);
    int c = 0;
    while (c < (int)seen.size() && seen[c] == i) ++c;
    color[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void permute_level(const HostLevel &L, const std::vector<int> &perm, const std::vector<int> &pos,
                   const std::vector<int> *cpos, HostLevel &out) {
  const int n = L.n, bb = L.b * L.b;
  out.n = n;
  out.nnzb = L.nnzb;
  out.b = L.b;
  out.rp.assign(n + 1, 0);
  for (int p = 0; p < n; ++p) out.rp[p + 1] = out.rp[p] + (L.rp[perm[p] + 1] - L.rp[perm[p]]);
  out.ci.resize(L.ci.size());
  out.a.resize(L.a.size());
  out.skip.resize(n);
  out.vc.resize(n);
  out.vnc.resize(n);
  out.vt.resize(L.vt.empty() ? 0 : n);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    const int r = perm[p];
    if (!L.vt.empty()) out.vt[p] = L.vt[r];
    int q = out.rp[p];
    for (int m = L.rp[r]; m < L.rp[r + 1]; ++m, ++q) {
      out.ci[q] = pos[L.ci[m]];
      for (int e = 0; e < bb; ++e) out.a[(size_t)q * bb + e] = L.a[(size_t)m * bb + e];
    }
    out.skip[p] = L.skip[r];
    out.vc[p] = L.vc[r];
    out.vnc[p] = L.vnc[r];
  }
  out.has_p = L.has_p;
  if (L.has_p) {
    out.pp.assign(n + 1, 0);
    for (int p = 0; p < n; ++p) out.pp[p + 1] = out.pp[p] + (L.pp[perm[p] + 1] - L.pp[perm[p]]);
    out.pi.resize(L.pi.size());
    out.pw.resize(L.pw.size());
    out.pwf.resize(L.pwf.size());
    out.pblk.resize(L.pblk.size());
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) {
      const int r = perm[p];
      int q = out.pp[p];
      for (int m = L.pp[r]; m < L.pp[r + 1]; ++m, ++q) {
        out.pi[q] = cpos ? (*cpos)[L.pi[m]] : L.pi[m];
        out.pw[q] = L.pw[m];
        if (!L.pwf.empty()) out.pwf[q] = L.pwf[m];
        if (!L.pblk.empty())
          for (int e = 0; e < bb; ++e) out.pblk[(size_t)q * bb + e] = L.pblk[(size_t)m * bb + e];
      }
    }
  }
}

void gs_device_order(const std::vector<int> &color, int ncolors, int window,
                     std::vector<int> &perm) {
  const int n = (int)color.size();
  const int W = window > 0 ? window : std::max(n, 1);
  perm.resize(n);
  int p = 0;
  std::vector<int> cnt(ncolors + 1);
  for (int w0 = 0; w0 < n; w0 += W) {
    const int w1 = std::min(n, w0 + W);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = w0; i < w1; ++i) ++cnt[color[i] + 1];
    for (int c = 0; c < ncolors; ++c) cnt[c + 1] += cnt[c];
    for (int i = w0; i < w1; ++i) perm[p + cnt[color[i]]++] = i;
    p += w1 - w0;
  }
}

int choose_gs_window(const HostLevel &L, const std::vector<int> &color, int ncolors) {
  const int n = L.n;
  if (n < 4 * kSlice) return 0;
  // most common smallest column offset >= 2 over a sample of rows
  std::vector<int> cand;
  {
    std::vector<std::pair<int, int>> hist;
    for (int k = 0; k < 512; ++k) {
      const int r = (int)((long long)k * (n - 1) / 511);
      int s = 0;
      for (int m = L.rp[r]; m < L.rp[r + 1]; ++m) {
        const int d = L.ci[m] - r;
        if (d >= 2 && (s == 0 || d < s)) s = d;
      }
      if (s == 0) continue;
      bool f = false;
      for (auto &h : hist)
        if (h.first == s) { ++h.second; f = true; }
      if (!f) hist.emplace_back(s, 1);
    }
    int best = 0, bc = 0;
    for (auto &h : hist)
      if (h.second > bc) { bc = h.second; best = h.first; }
    if (best > 0) {
      cand.push_back(best);
      cand.push_back(best + 1);
    }
    cand.push_back(0);   // whole level
  }
  int win = 0;
  double best_score = -1.0;
  std::vector<int> perm, pos(n), list;
  for (int W : cand) {
    gs_device_order(color, ncolors, W, perm);
    for (int p = 0; p < n; ++p) pos[perm[p]] = p;
    // colour-grouped device positions (the GS slab order)
    list.clear();
    for (int c = 0; c < ncolors; ++c)
      for (int p = 0; p < n; ++p)
        if (color[perm[p]] == c) list.push_back(p);
    const int nsl = (int)list.size() / kSlice;
    int tried = 0, uni = 0;
    for (int k = 0; k < 256 && nsl > 0; ++k) {
      const int s0 = (int)((long long)k * nsl / 256) * kSlice;
      ++tried;
      // uniform: every row's offsets (device positions) an in-order
      // subsequence of the longest row's
      int longest = -1, lw = -1;
      for (int j = 0; j < kSlice; ++j) {
        const int r = perm[list[s0 + j]];
        const int w = L.rp[r + 1] - L.rp[r];
        if (w > lw) { lw = w; longest = list[s0 + j]; }
      }
      std::vector<int> U;
      const int rl = perm[longest];
      for (int m = L.rp[rl]; m < L.rp[rl + 1]; ++m) U.push_back(pos[L.ci[m]] - longest);
      bool ok = true;
      for (int j = 0; j < kSlice && ok; ++j) {
        const int p = list[s0 + j], r = perm[p];
        size_t q = 0;
        for (int m = L.rp[r]; m < L.rp[r + 1] && ok; ++m) {
          const int d = pos[L.ci[m]] - p;
          while (q < U.size() && U[q] != d) ++q;
          if (q == U.size()) ok = false;
          else ++q;
        }
      }
      uni += ok;
    }
    const double score = tried ? (double)uni / tried : 0.0;
    if (score > best_score + 1e-9) {
      best_score = score;
      win = W;
    }
  }
  return win;
}

void build_prep(const HostLevel &L, const HostLevel *coarse, const std::vector<int> *color,
                int ncolors, HostPrep &out) {
  const int bb = L.b * L.b;
  out.diag.resize((size_t)L.n * bb);
  for (int i = 0; i < L.n; ++i)
    for (int e = 0; e < bb; ++e) out.diag[(size_t)i * bb + e] = L.a[(size_t)L.rp[i] * bb + e];
  if (!color) build_sell(L, nullptr, Part::All, nullptr, out.A);
  if (color) {
    out.ncolors = ncolors;
    out.color = *color;
    // rows grouped by colour, list order inside a colour; each colour padded
    // to whole slices so one launch per colour covers whole wavefronts
    std::vector<int> rows;
    out.color_slices.assign(out.ncolors + 1, 0);
    std::vector<std::vector<int>> by(out.ncolors);
    for (int i = 0; i < L.n; ++i) by[out.color[i]].push_back(i);
    for (int c = 0; c < out.ncolors; ++c) {
      for (int r : by[c]) rows.push_back(r);
      while (rows.size() % kSlice) rows.push_back(-1);
      out.color_slices[c + 1] = (int)(rows.size() / kSlice);
    }
    build_sell(L, &rows, Part::Lower, &out.color, out.L);
    build_sell(L, &rows, Part::Upper, &out.color, out.U);
    // the whole matrix over the same colour-grouped rows: a slice then holds
    // one colour run, whose stencil offsets are uniform
    build_sell(L, &rows, Part::All, nullptr, out.A);
  }
  if (coarse && L.has_p) {
    const int nc = coarse->n;
    out.pt_ptr.assign(nc + 1, 0);
    for (int v = 0; v < L.n; ++v) {
      if (!L.restricts(v)) continue;
      for (int q = L.pp[v]; q < L.pp[v + 1]; ++q) ++out.pt_ptr[L.pi[q] + 1];
    }
    for (int i = 0; i < nc; ++i) out.pt_ptr[i + 1] += out.pt_ptr[i];
    out.pt_idx.resize(out.pt_ptr[nc]);
    out.pt_w.resize(out.pt_ptr[nc]);
    std::vector<int> fill(out.pt_ptr.begin(), out.pt_ptr.end() - 1);
    const std::vector<double> &W = L.p_weights();
    const int b = L.b, bb = b * b;
    const bool mat = !L.pblk.empty();
    hvec<double> ptb(mat ? out.pt_idx.size() * bb : 0);
    for (int v = 0; v < L.n; ++v) {
      if (!L.restricts(v)) continue;
      for (int q = L.pp[v]; q < L.pp[v + 1]; ++q) {
        int c = L.pi[q];
        out.pt_idx[fill[c]] = v;
        out.pt_w[fill[c]] = mat ? 0.0 : W[q];
        if (mat)   // restriction: M as it is (coarse rows, fine columns)
          for (int e = 0; e < bb; ++e) ptb[(size_t)fill[c] * bb + e] = L.pblk[(size_t)q * bb + e];
        ++fill[c];
      }
    }
    if (mat) {
      // interpolation: M^T per entry (fine rows, coarse columns)
      hvec<double> pb(L.pblk.size());
      const size_t ne = L.pblk.size() / bb;
      for (size_t q = 0; q < ne; ++q)
        for (int i = 0; i < b; ++i)
          for (int j = 0; j < b; ++j) pb[q * bb + i * b + j] = L.pblk[q * bb + j * b + i];
      build_sell_csr(L.n, L.pp.data(), L.pi.data(), pb.data(), out.P, bb);
      build_sell_csr(nc, out.pt_ptr.data(), out.pt_idx.data(), ptb.data(), out.PT, bb);
      return;
    }
    build_sell_csr(L.n, L.pp.data(), L.pi.data(), W.data(), out.P);
    build_sell_csr(nc, out.pt_ptr.data(), out.pt_idx.data(), out.pt_w.data(), out.PT);
  }
}

#define EXM(m, bw, i, j) ((m)[2 * (size_t)(bw) * (i) + (j)])

void build_base_band(const HostLevel &L, int reorder, HostBase &out, bool fmode) {
  const int nvec = L.n, B = L.b, bb = B * B;
  std::vector<int> order(nvec);
  for (int i = 0; i < nvec; ++i) order[i] = i;
  if (reorder && nvec > 0) {
    std::vector<int> fifo(nvec);
    std::vector<char> used(nvec, 0);
    int head = 0, tail = 0, v = 0;
    fifo[tail++] = 0;
    used[0] = 1;
    while (head < tail) {
      v = fifo[head++];
      for (int m = L.rp[v] + 1; m < L.rp[v + 1]; ++m) {
        int w = L.ci[m];
        if (!used[w]) { fifo[tail++] = w; used[w] = 1; }
      }
    }
    head = tail = 0;
    fifo[tail++] = v;
    used[v] = 0;
    int i = 0;
    while (head < tail) {
      int u = fifo[head++];
      order[i++] = u;
      for (int m = L.rp[u] + 1; m < L.rp[u + 1]; ++m) {
        int w = L.ci[m];
        if (used[w]) { fifo[tail++] = w; used[w] = 0; }
      }
    }
    if (i != nvec) {  // disconnected pattern: keep the list order
      for (int k = 0; k < nvec; ++k) order[k] = k;
    }
  }
  out.order = order;
  // VINDEX * b of every vector; band over scalar unknowns, bandwidth
  // max|i - j| (+ b - 1 for blocks)
  std::vector<int> index(nvec);
  for (int k = 0; k < nvec; ++k) index[order[k]] = k * B;
  int bw = 0;
  for (int v = 0; v < nvec; ++v)
    for (int m = L.rp[v] + 1; m < L.rp[v + 1]; ++m)
      bw = std::max(bw, std::abs(index[v] - index[L.ci[m]]));
  if (B > 1) bw += B - 1;
  const int n = nvec * B;
  out.n = n;
  out.bw = bw;
  out.pos.resize(n);
  for (int u = 0; u < n; ++u) out.pos[u] = index[u / B] + u % B;
  out.band.assign((size_t)n * (2 * (size_t)bw + 1) + 1, 0.0);
  std::vector<double> &band = out.band;
  for (int v = 0; v < nvec; ++v)
    for (int m = L.rp[v]; m < L.rp[v + 1]; ++m)
      for (int i = 0; i < B; ++i)
        for (int j = 0; j < B; ++j)
          EXM(band.data(), bw, index[v] + i, index[L.ci[m]] + j) = L.a[(size_t)m * bb + i * B + j];
  // ex $f: the band holds FLOAT entries (EXCopyMatrixFLOAT, iter.cc:8325)
  out.fmode = fmode;
  if (fmode)
    for (double &x : band) x = (double)(float)x;
  out.failed = false;
  out.inv.clear();
}

void factor_base_host(HostBase &out) {
  const int n = out.n, bw = out.bw;
  std::vector<double> &band = out.band;
  out.failed = false;
  if (out.fmode) {   // EXDecomposeMatrixFLOAT (gm/evm.cc:1355-1373): FLOAT arithmetic
    for (int i = 0; i < n - 1; ++i) {
      const float d = (float)EXM(band.data(), bw, i, i);
      if (d == 0.0f) { out.failed = true; break; }
      const int jmax = std::min(i + bw, n - 1);
      for (int j = i + 1; j <= jmax; ++j) {
        const float f = (float)EXM(band.data(), bw, j, i) / d;
        EXM(band.data(), bw, j, i) = f;
        for (int k = i + 1; k <= jmax; ++k) {
          const float p = f * (float)EXM(band.data(), bw, i, k);
          EXM(band.data(), bw, j, k) = (float)EXM(band.data(), bw, j, k) - p;
        }
      }
    }
  }
  for (int i = 0; i < (out.fmode ? 0 : n - 1); ++i) {
    double d = EXM(band.data(), bw, i, i);
    if (d == 0.0) { out.failed = true; break; }
    int jmax = std::min(i + bw, n - 1);
    for (int j = i + 1; j <= jmax; ++j) {
      double f = EXM(band.data(), bw, j, i) / d;
      EXM(band.data(), bw, j, i) = f;
      for (int k = i + 1; k <= jmax; ++k)
        EXM(band.data(), bw, j, k) -= f * EXM(band.data(), bw, i, k);
    }
  }
  out.inv.assign((size_t)n * n, 0.0);
  if (out.failed) {
    std::vector<double>().swap(out.band);
    return;
  }
  // column u of the inverse: EXApplyLUDOUBLE on the unit vector at band
  // position pos[u] (the forward substitution starts there)
#pragma omp parallel
  {
    std::vector<double> vec(n);
#pragma omp for schedule(dynamic, 16)
    for (int u = 0; u < n; ++u) {
      std::fill(vec.begin(), vec.end(), 0.0);
      const int p0 = out.pos[u];
      vec[p0] = 1.0;
      for (int i = p0 + 1; i < n; ++i)
        for (int k = std::max(i - bw, p0); k < i; ++k) vec[i] -= EXM(band.data(), bw, i, k) * vec[k];
      for (int i = n - 1; i >= 0; --i) {
        int kmax = std::min(i + bw, n - 1);
        for (int k = i + 1; k <= kmax; ++k) vec[i] -= EXM(band.data(), bw, i, k) * vec[k];
        vec[i] /= EXM(band.data(), bw, i, i);
      }
      for (int r = 0; r < n; ++r) out.inv[(size_t)r * n + u] = vec[out.pos[r]];
    }
  }
  std::vector<double>().swap(out.band);
}

void build_base(const HostLevel &L, int reorder, HostBase &out, bool fmode) {
  build_base_band(L, reorder, out, fmode);
  factor_base_host(out);
}

}  // namespace ugmg
