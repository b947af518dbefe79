// Synthetic UG-layout multigrid problems (host only).
//
// Plays the part of the UG application that assembles one matrix per level
// (UG has no discretisation of its own; the solve path reads whatever the
// application put into the MATRIX lists).  Conventions follow UG:
//   * vector list per level, one vector per node, listed lexicographically
//     (x fastest), flattened as in ConvertMatrix (np/udm/disctools.cc:
//     2061-2128): VINDEX = list position, each row starts with its diagonal
//     MATRIX (VSTART, gm/gm.h:2298) followed by the off-diagonal entries,
//     here in ascending column order;
//   * Dirichlet rows as AssembleDirichletBoundary (np/udm/disctools.cc:
//     1837-1875): row zeroed, diagonal 1, column kept, rhs = solution = 0;
//   * interpolation weights = GNs of the father element at the node's local
//     coordinate (gm/shapes.cc:748-819; corner order (0,0),(1,0),(1,1),(0,1)
//     and the hexahedron's bottom then top face), as used by
//     StandardIntCorNodeVector (np/algebra/transgrid.cc:235-336).
#include "../../include/ugmg_problems.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

namespace {

struct Level {
  int N = 0;        // nodes per direction
  int C = 0;        // cells per direction
  int n = 0;        // vectors
  std::vector<int> rp, ci;
  std::vector<double> v;
  std::vector<uint32_t> skip;
  std::vector<uint8_t> vc, vnc;
  std::vector<int> pp, pi;      // interpolation from level-1 (empty on 0)
  std::vector<double> pw;
};

inline uint64_t splitmix_next(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline double splitmix_u01(uint64_t &s) {
  return (double)(splitmix_next(s) >> 11) * 0x1.0p-53;
}

// UG corner order of the reference quadrilateral / hexahedron, as tensor bits
// (bit0 = x, bit1 = y, bit2 = z).
const int kQuadCorner[4] = {0, 1, 3, 2};
const int kHexCorner[8] = {0, 1, 3, 2, 4, 5, 7, 6};

}  // namespace

struct ugmg_problem {
  int dim = 0;
  int b = 1;
  double theta = 0.0;
  std::vector<Level> lv;
  std::vector<double> rhs;
  // unstructured tetrahedral meshes (ugmg_problem_elasticity_local only):
  // node coordinates, corners and kind (0 red, 1 green, 2 copy) per level
  std::vector<std::vector<double>> xyz;
  std::vector<std::vector<int>> tets;
  std::vector<std::vector<uint8_t>> tet_kind;
};

extern "C" void ugmg_splitmix_uniform(uint64_t seed, int64_t count, double a,
                                      double b, double *out) {
  uint64_t s = seed;
  for (int64_t i = 0; i < count; ++i) out[i] = a + (b - a) * splitmix_u01(s);
}

namespace {

// Q1 element stiffness of -div(grad) on a cube/square of side h, tensor
// local numbering: K[a][b] = sum_d s(a_d,b_d) prod_{e!=d} m(a_e,b_e) with the
// 1D stiffness s = (1/h)[[1,-1],[-1,1]] and mass m = (h/6)[[2,1],[1,2]].
void q1_stiffness(int dim, double h, double *K) {
  const int nc = 1 << dim;
  for (int a = 0; a < nc; ++a)
    for (int bb = 0; bb < nc; ++bb) {
      double sum = 0.0;
      for (int d = 0; d < dim; ++d) {
        int ad = (a >> d) & 1, bd = (bb >> d) & 1;
        double t = (ad == bd ? 1.0 : -1.0) / h;
        for (int e = 0; e < dim; ++e) {
          if (e == d) continue;
          int ae = (a >> e) & 1, be = (bb >> e) & 1;
          t *= (ae == be ? 2.0 : 1.0) * h / 6.0;
        }
        sum += t;
      }
      K[a * nc + bb] = sum;
    }
}

// Q[c][a][J]: coarse basis J (tensor numbering) at corner a of child c.
void child_interp(int dim, std::vector<double> &Q) {
  const int nc = 1 << dim;
  Q.assign((size_t)nc * nc * nc, 0.0);
  for (int c = 0; c < nc; ++c)
    for (int a = 0; a < nc; ++a)
      for (int J = 0; J < nc; ++J) {
        double w = 1.0;
        for (int d = 0; d < dim; ++d) {
          double p = 0.5 * (((c >> d) & 1) + ((a >> d) & 1));
          w *= ((J >> d) & 1) ? p : 1.0 - p;
        }
        Q[((size_t)c * nc + a) * nc + J] = w;
      }
}

// Kc += Q^T Kf Q  (nc x nc)
void galerkin_add(int nc, const double *Q, const double *Kf, double scale,
                  double *Kc) {
  double T[64];
  for (int a = 0; a < nc; ++a)
    for (int J = 0; J < nc; ++J) {
      double s = 0.0;
      for (int bb = 0; bb < nc; ++bb) s += Kf[a * nc + bb] * Q[bb * nc + J];
      T[a * nc + J] = s;
    }
  for (int I = 0; I < nc; ++I)
    for (int J = 0; J < nc; ++J) {
      double s = 0.0;
      for (int a = 0; a < nc; ++a) s += Q[a * nc + I] * T[a * nc + J];
      Kc[I * nc + J] += scale * s;
    }
}

inline int64_t node_id(int dim, int N, int x, int y, int z) {
  return dim == 2 ? (int64_t)y * N + x : ((int64_t)z * N + y) * N + x;
}

// Row-wise assembly of one level from element matrices.
// emat(e) returns a pointer to the element matrix (tensor numbering) of cell e
// or fills scratch for the finest level.
template <class EMat>
void assemble_level(int dim, Level &L, EMat emat) {
  const int N = L.N, C = L.C;
  const int64_t n = L.n;
  const int nsl = dim == 2 ? 9 : 27;
  // pattern: all stencil neighbours inside the grid
  L.rp.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    int x = (int)(i % N), y = (int)((i / N) % N), z = dim == 3 ? (int)(i / ((int64_t)N * N)) : 0;
    int cx = 1 + (x > 0) + (x < N - 1), cy = 1 + (y > 0) + (y < N - 1);
    int cz = dim == 3 ? 1 + (z > 0) + (z < N - 1) : 1;
    L.rp[i + 1] = cx * cy * cz;
  }
  for (int64_t i = 0; i < n; ++i) L.rp[i + 1] += L.rp[i];
  const int64_t nnz = L.rp[n];
  L.ci.assign(nnz, 0);
  L.v.assign(nnz, 0.0);
  L.skip.assign(n, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int x = (int)(i % N), y = (int)((i / N) % N), z = dim == 3 ? (int)(i / ((int64_t)N * N)) : 0;
    double acc[27];
    bool valid[27];
    for (int k = 0; k < nsl; ++k) { acc[k] = 0.0; valid[k] = false; }
    const int zlo = dim == 3 ? -1 : 0, zhi = dim == 3 ? 0 : 0;
    double scratch[64];
    // every element touching node i
    for (int ez = zlo; ez <= zhi; ++ez)
      for (int ey = -1; ey <= 0; ++ey)
        for (int ex = -1; ex <= 0; ++ex) {
          int exx = x + ex, eyy = y + ey, ezz = dim == 3 ? z + ez : 0;
          if (exx < 0 || exx >= C || eyy < 0 || eyy >= C) continue;
          if (dim == 3 && (ezz < 0 || ezz >= C)) continue;
          const int64_t e = dim == 2 ? (int64_t)eyy * C + exx
                                     : ((int64_t)ezz * C + eyy) * C + exx;
          const double *K = emat(e, scratch);
          const int nc = 1 << dim;
          int a = (-ex) | ((-ey) << 1) | (dim == 3 ? ((-ez) << 2) : 0);
          for (int bb = 0; bb < nc; ++bb) {
            int dx = (bb & 1) + ex, dy = ((bb >> 1) & 1) + ey;
            int dz = dim == 3 ? ((bb >> 2) & 1) + ez : 0;
            int k = dim == 2 ? (dy + 1) * 3 + (dx + 1)
                             : (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
            acc[k] += K[a * nc + bb];
          }
        }
    for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int xx = x + dx, yy = y + dy, zz = z + dz;
          if (xx < 0 || xx >= N || yy < 0 || yy >= N) continue;
          if (dim == 3 && (zz < 0 || zz >= N)) continue;
          int k = dim == 2 ? (dy + 1) * 3 + (dx + 1)
                           : (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
          valid[k] = true;
        }
    const int kdiag = dim == 2 ? 4 : 13;
    int64_t p = L.rp[i];
    L.ci[p] = (int)i;
    L.v[p] = acc[kdiag];
    ++p;
    for (int k = 0; k < nsl; ++k) {
      if (k == kdiag || !valid[k]) continue;
      int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = dim == 3 ? k / 9 - 1 : 0;
      L.ci[p] = (int)node_id(dim, N, x + dx, y + dy, z + dz);
      L.v[p] = acc[k];
      ++p;
    }
    const bool bnd = x == 0 || x == N - 1 || y == 0 || y == N - 1 ||
                     (dim == 3 && (z == 0 || z == N - 1));
    if (bnd) {
      // AssembleDirichletBoundary: zero the row, unit diagonal.
      L.skip[i] = 1u;
      for (int64_t q = L.rp[i]; q < L.rp[i + 1]; ++q) L.v[q] = 0.0;
      L.v[L.rp[i]] = 1.0;
    }
  }
  L.vc.assign(n, 3);
}

void build_transfer(int dim, const Level &coarse, Level &fine) {
  const int Nf = fine.N, Nc = coarse.N, Cc = coarse.C;
  const int64_t n = fine.n;
  const int nc = 1 << dim;
  const int *corner = dim == 2 ? kQuadCorner : kHexCorner;
  fine.pp.assign(n + 1, 0);
  std::vector<int> tmp_idx((size_t)n * nc);
  std::vector<double> tmp_w((size_t)n * nc);
  std::vector<int> cnt(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int c[3] = {(int)(i % Nf), (int)((i / Nf) % Nf),
                dim == 3 ? (int)(i / ((int64_t)Nf * Nf)) : 0};
    bool corner_node = true;
    for (int d = 0; d < dim; ++d) corner_node &= (c[d] % 2 == 0);
    int k = 0;
    if (corner_node) {
      tmp_idx[i * nc] = (int)node_id(dim, Nc, c[0] / 2, c[1] / 2, c[2] / 2);
      tmp_w[i * nc] = 1.0;
      k = 1;
    } else {
      int e[3] = {0, 0, 0};
      double xi[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) {
        if (c[d] % 2) { e[d] = (c[d] - 1) / 2; xi[d] = 0.5; }
        else {
          int X = c[d] / 2;
          e[d] = X < Cc ? X : Cc - 1;
          xi[d] = (double)(X - e[d]);
        }
      }
      for (int q = 0; q < nc; ++q) {
        int t = corner[q];
        double w = 1.0;
        for (int d = 0; d < dim; ++d) w *= ((t >> d) & 1) ? xi[d] : 1.0 - xi[d];
        if (w == 0.0) continue;
        tmp_idx[i * nc + k] = (int)node_id(dim, Nc, e[0] + (t & 1),
                                          e[1] + ((t >> 1) & 1),
                                          e[2] + ((t >> 2) & 1));
        tmp_w[i * nc + k] = w;
        ++k;
      }
    }
    cnt[i] = k;
  }
  for (int64_t i = 0; i < n; ++i) fine.pp[i + 1] = fine.pp[i] + cnt[i];
  fine.pi.resize(fine.pp[n]);
  fine.pw.resize(fine.pp[n]);
  for (int64_t i = 0; i < n; ++i)
    for (int q = 0; q < cnt[i]; ++q) {
      fine.pi[fine.pp[i] + q] = tmp_idx[i * nc + q];
      fine.pw[fine.pp[i] + q] = tmp_w[i * nc + q];
    }
}

}  // namespace

extern "C" int ugmg_problem_q1(int dim, int coarse_cells, int nref,
                               int kappa_mode, uint64_t seed,
                               ugmg_problem **out) {
  if (!out || (dim != 2 && dim != 3) || coarse_cells < 1 || nref < 0 ||
      nref > 12 || (kappa_mode != UGMG_KAPPA_ONE && kappa_mode != UGMG_KAPPA_JUMP))
    return -1;
  // int32 indices: keep nnz below 2^31
  {
    double Nf = (double)coarse_cells * std::ldexp(1.0, nref) + 1.0;
    double nnz = std::pow(3.0 * (Nf - 1) + 1, dim);
    if (nnz > 2.0e9) return -1;
  }
  ugmg_problem *p = new (std::nothrow) ugmg_problem;
  if (!p) return -4;
  p->dim = dim;
  p->b = 1;
  const int L = nref;
  p->lv.resize(L + 1);
  for (int l = 0; l <= L; ++l) {
    Level &lv = p->lv[l];
    lv.C = coarse_cells << l;
    lv.N = lv.C + 1;
    lv.n = dim == 2 ? lv.N * lv.N : lv.N * lv.N * lv.N;
  }
  const int nc = 1 << dim;
  const int64_t nel_f = dim == 2 ? (int64_t)p->lv[L].C * p->lv[L].C
                                 : (int64_t)p->lv[L].C * p->lv[L].C * p->lv[L].C;
  uint64_t s = seed;
  std::vector<double> kappa(nel_f, 1.0);
  if (kappa_mode == UGMG_KAPPA_JUMP)
    for (int64_t e = 0; e < nel_f; ++e)
      kappa[e] = std::pow(10.0, -2.0 + 4.0 * splitmix_u01(s));
  std::vector<double> f(p->lv[L].n);
  for (auto &fi : f) fi = -1.0 + 2.0 * splitmix_u01(s);

  // finest level: kappa_e * Kref(h)
  std::vector<double> Kref(nc * nc);
  q1_stiffness(dim, 1.0 / p->lv[L].C, Kref.data());
  assemble_level(dim, p->lv[L], [&](int64_t e, double *scratch) -> const double * {
    for (int q = 0; q < nc * nc; ++q) scratch[q] = kappa[e] * Kref[q];
    return scratch;
  });
  // coarser levels: element-local Galerkin products of the children
  std::vector<double> Q;
  child_interp(dim, Q);
  std::vector<double> emat_child, emat;
  for (int l = L - 1; l >= 0; --l) {
    const int Cc = p->lv[l].C, Cf = p->lv[l + 1].C;
    const int64_t nel = dim == 2 ? (int64_t)Cc * Cc : (int64_t)Cc * Cc * Cc;
    emat.assign((size_t)nel * nc * nc, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t E = 0; E < nel; ++E) {
      int X = (int)(E % Cc), Y = (int)((E / Cc) % Cc), Z = dim == 3 ? (int)(E / ((int64_t)Cc * Cc)) : 0;
      double Kf[64];
      for (int c = 0; c < nc; ++c) {
        int fx = 2 * X + (c & 1), fy = 2 * Y + ((c >> 1) & 1), fz = 2 * Z + ((c >> 2) & 1);
        int64_t fe = dim == 2 ? (int64_t)fy * Cf + fx : ((int64_t)fz * Cf + fy) * Cf + fx;
        const double *Kc;
        if (l == L - 1) {
          for (int q = 0; q < nc * nc; ++q) Kf[q] = kappa[fe] * Kref[q];
          Kc = Kf;
        } else {
          Kc = &emat_child[(size_t)fe * nc * nc];
        }
        galerkin_add(nc, &Q[(size_t)c * nc * nc], Kc, 1.0, &emat[(size_t)E * nc * nc]);
      }
    }
    assemble_level(dim, p->lv[l], [&](int64_t e, double *) -> const double * {
      return &emat[(size_t)e * nc * nc];
    });
    emat_child.swap(emat);
  }
  for (int l = 0; l <= L; ++l) p->lv[l].vnc.assign(p->lv[l].n, l < L ? 3 : 0);
  for (int l = 1; l <= L; ++l) build_transfer(dim, p->lv[l - 1], p->lv[l]);

  // rhs = M f (consistent Q1 mass, tensor product of (h/6)[1 4 1]); 0 on
  // Dirichlet vectors (rhs = sol = 0).
  {
    const Level &F = p->lv[L];
    const int N = F.N;
    const double h = 1.0 / F.C;
    const double m1[3] = {h / 6.0, 4.0 * h / 6.0, h / 6.0};
    p->rhs.assign(F.n, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F.n; ++i) {
      if (F.skip[i]) continue;
      int x = (int)(i % N), y = (int)((i / N) % N), z = dim == 3 ? (int)(i / ((int64_t)N * N)) : 0;
      double sum = 0.0;
      for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            double w = m1[dx + 1] * m1[dy + 1] * (dim == 3 ? m1[dz + 1] : 1.0);
            sum += w * f[node_id(dim, N, x + dx, y + dy, z + dz)];
          }
      p->rhs[i] = sum;
    }
  }
  *out = p;
  return 0;
}

// ----------------------------------------------------------------------
// P1 triangles, streamline-upwind Petrov-Galerkin convection-diffusion
//   -eps Laplace u + b . grad u = f,  b = (cos theta, sin theta)
// Unit square, 2x2 coarse squares each split by the (0,0)-(1,1) diagonal
// (8 triangles), uniform red refinement (every triangle into 4), so level l
// has (2*2^l+1)^2 nodes; every level is rediscretised with its own h.
// Element matrix: eps (grad phi_j, grad phi_i) + (b.grad phi_j, phi_i)
//               + tau (b.grad phi_j, b.grad phi_i),
// tau = h/(2|b|) (coth Pe - 1/Pe), Pe = |b| h / (2 eps);
// rhs: (f_h, phi_i) + tau (f_h, b.grad phi_i) with f_h the P1 interpolant.
namespace {

void p1_element(const double xy[3][2], double eps, const double bvec[2], double tau,
                double K[3][3], double M[3][3], double G[3]) {
  const double x0 = xy[0][0], y0 = xy[0][1], x1 = xy[1][0], y1 = xy[1][1], x2 = xy[2][0],
               y2 = xy[2][1];
  const double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
  const double area = 0.5 * std::fabs(det);
  double gr[3][2] = {{(y1 - y2) / det, (x2 - x1) / det},
                     {(y2 - y0) / det, (x0 - x2) / det},
                     {(y0 - y1) / det, (x1 - x0) / det}};
  for (int i = 0; i < 3; ++i) {
    const double bgi = bvec[0] * gr[i][0] + bvec[1] * gr[i][1];
    G[i] = tau * bgi * area;  // weight of mean(f) in the SUPG rhs term
    for (int j = 0; j < 3; ++j) {
      const double bgj = bvec[0] * gr[j][0] + bvec[1] * gr[j][1];
      K[i][j] = eps * area * (gr[i][0] * gr[j][0] + gr[i][1] * gr[j][1]) + area / 3.0 * bgj +
                tau * area * bgi * bgj;
      M[i][j] = area / 12.0 * (i == j ? 2.0 : 1.0);
    }
  }
}

void assemble_p1_level(Level &L, double eps, const double bvec[2], std::vector<double> *rhs,
                       const std::vector<double> *f) {
  const int N = L.N, C = L.C;
  const double h = 1.0 / C;
  const double pe = h / (2.0 * eps);
  const double tau = pe > 1e-8 ? h / 2.0 * (1.0 / std::tanh(pe) - 1.0 / pe) : 0.0;
  const int64_t n = (int64_t)N * N;
  // offsets (dx,dy) coupled by a triangle of this split, ascending column order
  const int offs[7][2] = {{-1, -1}, {0, -1}, {-1, 0}, {0, 0}, {1, 0}, {0, 1}, {1, 1}};
  std::vector<double> st((size_t)n * 7, 0.0);
  if (rhs) rhs->assign(n, 0.0);
  auto slot = [&](int dx, int dy) {
    for (int k = 0; k < 7; ++k)
      if (offs[k][0] == dx && offs[k][1] == dy) return k;
    return -1;
  };
  for (int cy = 0; cy < C; ++cy)
    for (int cx = 0; cx < C; ++cx)
      for (int t = 0; t < 2; ++t) {
        // lower triangle (a,b,c), upper (a,c,d); a = (cx,cy) ... d = (cx,cy+1)
        const int vx[2][3] = {{cx, cx + 1, cx + 1}, {cx, cx + 1, cx}};
        const int vy[2][3] = {{cy, cy, cy + 1}, {cy, cy + 1, cy + 1}};
        double xy[3][2];
        int64_t id[3];
        for (int q = 0; q < 3; ++q) {
          xy[q][0] = vx[t][q] * h;
          xy[q][1] = vy[t][q] * h;
          id[q] = (int64_t)vy[t][q] * N + vx[t][q];
        }
        double K[3][3], M[3][3], G[3];
        p1_element(xy, eps, bvec, tau, K, M, G);
        for (int i = 0; i < 3; ++i) {
          for (int j = 0; j < 3; ++j) {
            const int k = slot(vx[t][j] - vx[t][i], vy[t][j] - vy[t][i]);
            st[(size_t)id[i] * 7 + k] += K[i][j];
          }
          if (rhs) {
            double fm = 0.0;
            for (int j = 0; j < 3; ++j) {
              (*rhs)[id[i]] += M[i][j] * (*f)[id[j]];
              fm += (*f)[id[j]];
            }
            (*rhs)[id[i]] += G[i] * fm / 3.0;
          }
        }
      }
  L.rp.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % N), y = (int)(i / N);
    int cnt = 0;
    for (int k = 0; k < 7; ++k) {
      const int xx = x + offs[k][0], yy = y + offs[k][1];
      cnt += (xx >= 0 && xx < N && yy >= 0 && yy < N);
    }
    L.rp[i + 1] = L.rp[i] + cnt;
  }
  L.ci.resize(L.rp[n]);
  L.v.resize(L.rp[n]);
  L.skip.assign(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % N), y = (int)(i / N);
    int64_t q = L.rp[i];
    L.ci[q] = (int)i;
    L.v[q] = st[(size_t)i * 7 + 3];
    ++q;
    for (int k = 0; k < 7; ++k) {
      if (k == 3) continue;
      const int xx = x + offs[k][0], yy = y + offs[k][1];
      if (xx < 0 || xx >= N || yy < 0 || yy >= N) continue;
      L.ci[q] = yy * N + xx;
      L.v[q] = st[(size_t)i * 7 + k];
      ++q;
    }
    if (x == 0 || x == N - 1 || y == 0 || y == N - 1) {
      L.skip[i] = 1u;
      for (int64_t m = L.rp[i]; m < L.rp[i + 1]; ++m) L.v[m] = 0.0;
      L.v[L.rp[i]] = 1.0;
      if (rhs) (*rhs)[i] = 0.0;
    }
  }
  L.vc.assign(n, 3);
}

// P1 interpolation: corner nodes take their father, edge midpoints 1/2 of
// the edge's two end points (GNs of the father triangle, zero weight dropped)
void build_p1_transfer(const Level &coarse, Level &fine) {
  const int Nf = fine.N, Nc = coarse.N;
  const int64_t n = (int64_t)Nf * Nf;
  fine.pp.assign(n + 1, 0);
  fine.pi.clear();
  fine.pw.clear();
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % Nf), y = (int)(i / Nf);
    if (x % 2 == 0 && y % 2 == 0) {
      fine.pi.push_back((y / 2) * Nc + x / 2);
      fine.pw.push_back(1.0);
    } else {
      const int x0 = x / 2, y0 = y / 2;                 // lower end point
      const int x1 = (x + 1) / 2, y1 = (y + 1) / 2;     // upper end point
      fine.pi.push_back(y0 * Nc + x0);
      fine.pw.push_back(0.5);
      fine.pi.push_back(y1 * Nc + x1);
      fine.pw.push_back(0.5);
    }
    fine.pp[i + 1] = (int)fine.pi.size();
  }
}

}  // namespace

extern "C" int ugmg_problem_p1_cd(int nref, double eps, double theta, uint64_t seed,
                                  ugmg_problem **out) {
  if (!out || nref < 0 || nref > 11 || !(eps > 0.0)) return -1;
  ugmg_problem *p = new (std::nothrow) ugmg_problem;
  if (!p) return -4;
  p->dim = 2;
  p->b = 1;
  uint64_t s = seed;
  if (std::isnan(theta)) theta = 2.0 * M_PI * splitmix_u01(s);
  const double bvec[2] = {std::cos(theta), std::sin(theta)};
  const int L = nref;
  p->lv.resize(L + 1);
  for (int l = 0; l <= L; ++l) {
    p->lv[l].C = 2 << l;
    p->lv[l].N = p->lv[l].C + 1;
    p->lv[l].n = p->lv[l].N * p->lv[l].N;
  }
  std::vector<double> f(p->lv[L].n);
  for (auto &fi : f) fi = -1.0 + 2.0 * splitmix_u01(s);
  for (int l = 0; l <= L; ++l)
    assemble_p1_level(p->lv[l], eps, bvec, l == L ? &p->rhs : nullptr, l == L ? &f : nullptr);
  for (int l = 0; l <= L; ++l) p->lv[l].vnc.assign(p->lv[l].n, l < L ? 3 : 0);
  for (int l = 1; l <= L; ++l) build_p1_transfer(p->lv[l - 1], p->lv[l]);
  p->theta = theta;
  *out = p;
  return 0;
}

extern "C" double ugmg_problem_theta(const ugmg_problem *p) { return p ? p->theta : 0.0; }

namespace {

// ---------------------------------------------------------------- elasticity
// Kuhn (Freudenthal) split of the unit cube into the 6 tetrahedra
// 0 -> e_a -> e_a + e_b -> (1,1,1) over the permutations (a,b,c) of the axes;
// refining every cube of a level into 8 and splitting each the same way is a
// red refinement of the coarse tetrahedra, so the P1 spaces are nested and a
// fine node either is a coarse node or bisects a coarse edge e_S (S the set
// of its odd coordinates).
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// stencil offsets of a Kuhn lattice node: +-e_S for every nonempty S, and 0
void kuhn_offsets(std::vector<std::array<int, 3>> &offs, int N) {
  offs.clear();
  for (int sgn = -1; sgn <= 1; sgn += 2)
    for (int S = 1; S < 8; ++S) offs.push_back({sgn * (S & 1), sgn * ((S >> 1) & 1), sgn * ((S >> 2) & 1)});
  offs.push_back({0, 0, 0});
  // ascending column order for lexicographic ids (x fastest)
  std::sort(offs.begin(), offs.end(), [N](const std::array<int, 3> &a, const std::array<int, 3> &b) {
    return a[0] + (long long)N * (a[1] + (long long)N * a[2]) < b[0] + (long long)N * (b[1] + (long long)N * b[2]);
  });
}

// P1 linear elasticity on one tetrahedron: K[i][j] (3x3 block, row-major) =
// V (lambda d_a phi_i d_b phi_j + mu d_b phi_i d_a phi_j + mu delta_ab
// grad phi_i . grad phi_j), consistent mass V/20 (1 + delta_ij)
void p1_elasticity_element(const double X[4][3], double lam, double mu, double K[4][4][9],
                           double M[4][4]) {
  double J[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) J[r][c] = X[c + 1][r] - X[0][r];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double V = std::fabs(det) / 6.0;
  // inverse of J: rows are the gradients of the barycentric coordinates 1..3
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  double g[4][3];
  for (int d = 0; d < 3; ++d) {
    g[1][d] = Ji[0][d];
    g[2][d] = Ji[1][d];
    g[3][d] = Ji[2][d];
    g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const double gg = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
      for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 3; ++bb)
          K[i][j][a * 3 + bb] =
              V * (lam * g[i][a] * g[j][bb] + mu * g[i][bb] * g[j][a] + (a == bb ? mu * gg : 0.0));
      M[i][j] = V / 20.0 * (i == j ? 2.0 : 1.0);
    }
}

void assemble_elasticity_level(Level &L, double lam, double mu, std::vector<double> *rhs,
                               const std::vector<double> *f) {
  const int N = L.N, C = L.C;
  const double h = 1.0 / C;
  const int64_t n = (int64_t)N * N * N;
  std::vector<std::array<int, 3>> offs;
  kuhn_offsets(offs, N);
  const int no = (int)offs.size();   // 15
  auto slot = [&](int dx, int dy, int dz) {
    for (int k = 0; k < no; ++k)
      if (offs[k][0] == dx && offs[k][1] == dy && offs[k][2] == dz) return k;
    return -1;
  };
  std::vector<double> st((size_t)n * no * 9, 0.0);
  if (rhs) rhs->assign((size_t)n * 3, 0.0);
  for (int cz = 0; cz < C; ++cz)
    for (int cy = 0; cy < C; ++cy)
      for (int cx = 0; cx < C; ++cx)
        for (int t = 0; t < 6; ++t) {
          int vx[4][3] = {{0, 0, 0}};
          for (int q = 1; q < 4; ++q) {
            for (int d = 0; d < 3; ++d) vx[q][d] = vx[q - 1][d];
            vx[q][kPerm[t][q - 1]] = 1;
          }
          double X[4][3];
          int64_t id[4];
          int cc[4][3];
          for (int q = 0; q < 4; ++q) {
            cc[q][0] = cx + vx[q][0];
            cc[q][1] = cy + vx[q][1];
            cc[q][2] = cz + vx[q][2];
            for (int d = 0; d < 3; ++d) X[q][d] = cc[q][d] * h;
            id[q] = cc[q][0] + (int64_t)N * (cc[q][1] + (int64_t)N * cc[q][2]);
          }
          double K[4][4][9], M[4][4];
          p1_elasticity_element(X, lam, mu, K, M);
          for (int i = 0; i < 4; ++i) {
            for (int j = 0; j < 4; ++j) {
              const int k = slot(cc[j][0] - cc[i][0], cc[j][1] - cc[i][1], cc[j][2] - cc[i][2]);
              double *dst = &st[((size_t)id[i] * no + k) * 9];
              for (int e = 0; e < 9; ++e) dst[e] += K[i][j][e];
            }
            if (rhs)
              for (int j = 0; j < 4; ++j)
                for (int a = 0; a < 3; ++a) (*rhs)[(size_t)id[i] * 3 + a] += M[i][j] * (*f)[(size_t)id[j] * 3 + a];
          }
        }
  const int kd = slot(0, 0, 0);
  L.rp.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((int64_t)N * N));
    int cnt = 0;
    for (int k = 0; k < no; ++k) {
      const int xx = x + offs[k][0], yy = y + offs[k][1], zz = z + offs[k][2];
      cnt += (xx >= 0 && xx < N && yy >= 0 && yy < N && zz >= 0 && zz < N);
    }
    L.rp[i + 1] = L.rp[i] + cnt;
  }
  L.ci.resize(L.rp[n]);
  L.v.resize((size_t)L.rp[n] * 9);
  L.skip.assign(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((int64_t)N * N));
    int64_t q = L.rp[i];
    auto put = [&](int64_t col, int k) {
      L.ci[q] = (int)col;
      std::memcpy(&L.v[(size_t)q * 9], &st[((size_t)i * no + k) * 9], 9 * sizeof(double));
      ++q;
    };
    put(i, kd);
    for (int k = 0; k < no; ++k) {
      if (k == kd) continue;
      const int xx = x + offs[k][0], yy = y + offs[k][1], zz = z + offs[k][2];
      if (xx < 0 || xx >= N || yy < 0 || yy >= N || zz < 0 || zz >= N) continue;
      put(xx + (int64_t)N * (yy + (int64_t)N * zz), k);
    }
    if (x == 0) {   // clamped face: every component Dirichlet
      L.skip[i] = 7u;
      for (int64_t m = L.rp[i]; m < L.rp[i + 1]; ++m)
        for (int e = 0; e < 9; ++e) L.v[(size_t)m * 9 + e] = 0.0;
      for (int a = 0; a < 3; ++a) L.v[(size_t)L.rp[i] * 9 + a * 4] = 1.0;
      if (rhs)
        for (int a = 0; a < 3; ++a) (*rhs)[(size_t)i * 3 + a] = 0.0;
    }
  }
  L.vc.assign(n, 3);
}

// P1 interpolation on the Kuhn lattice: coarse nodes 1, bisecting nodes 1/2
// to the two end points of the coarse edge (lower end point first)
void build_kuhn_transfer(const Level &coarse, Level &fine) {
  const int Nf = fine.N, Nc = coarse.N;
  const int64_t n = (int64_t)Nf * Nf * Nf;
  fine.pp.assign(n + 1, 0);
  fine.pi.clear();
  fine.pw.clear();
  auto cid = [Nc](int x, int y, int z) { return x + Nc * (y + Nc * z); };
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % Nf), y = (int)((i / Nf) % Nf), z = (int)(i / ((int64_t)Nf * Nf));
    if (x % 2 == 0 && y % 2 == 0 && z % 2 == 0) {
      fine.pi.push_back(cid(x / 2, y / 2, z / 2));
      fine.pw.push_back(1.0);
    } else {
      fine.pi.push_back(cid(x / 2, y / 2, z / 2));
      fine.pw.push_back(0.5);
      fine.pi.push_back(cid((x + 1) / 2, (y + 1) / 2, (z + 1) / 2));
      fine.pw.push_back(0.5);
    }
    fine.pp[i + 1] = (int)fine.pi.size();
  }
}

}  // namespace

extern "C" int ugmg_problem_elasticity(int nref, double E, double nu, uint64_t seed,
                                       ugmg_problem **out) {
  if (!out || nref < 0 || nref > 8 || !(E > 0.0) || !(nu > -1.0 && nu < 0.5)) return -1;
  ugmg_problem *p = new (std::nothrow) ugmg_problem;
  if (!p) return -4;
  p->dim = 3;
  p->b = 3;
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
  const int L = nref;
  try {
    p->lv.resize(L + 1);
    for (int l = 0; l <= L; ++l) {
      p->lv[l].C = 1 << l;
      p->lv[l].N = p->lv[l].C + 1;
      p->lv[l].n = p->lv[l].N * p->lv[l].N * p->lv[l].N;
    }
    uint64_t s = seed;
    std::vector<double> f((size_t)p->lv[L].n * 3);
    for (auto &fi : f) fi = -1.0 + 2.0 * splitmix_u01(s);
    for (int l = 0; l <= L; ++l)
      assemble_elasticity_level(p->lv[l], lam, mu, l == L ? &p->rhs : nullptr, l == L ? &f : nullptr);
    for (int l = 0; l <= L; ++l) p->lv[l].vnc.assign(p->lv[l].n, l < L ? 3 : 0);
    for (int l = 1; l <= L; ++l) build_kuhn_transfer(p->lv[l - 1], p->lv[l]);
  } catch (const std::bad_alloc &) {
    delete p;
    return -4;
  }
  *out = p;
  return 0;
}

// ----------------------------------------------- locally refined elasticity
// P1 linear elasticity (the element matrices above) on a tetrahedral
// hierarchy refined the way UG refines tetrahedra in its default build,
// where tetrahedra have no complete rule set (TET_RULESET undefined:
// gm/rm.h:173, gm/refine.cc:160-166, gm/rm.cc:2764-2776):
//   * level 0: the unit cube split into the 6 Kuhn tetrahedra;
//   * refinement step s marks the regular (red) elements whose corners all
//     lie in [0, a]^3, a = 1 for the first nuni steps, then 1/2, 1/4, ...;
//   * a marked element is refined red into 8 (TET_RED, gm/rm.cc:536-575):
//     the four corner tetrahedra and the inner octahedron split along the
//     diagonal joining the midpoints of edges (0,2) and (1,3), which keeps
//     every child a Kuhn simplex of half the size, so a fully refined region
//     is the uniform Kuhn lattice of the next level;
//   * an unmarked element with refined edges is closed green
//     (BuildGreenClosure, gm/refine.cc:1850-2005; RefineElementGreen,
//     4493-5360): a centre node at the mean of its corners (CreateCenterNode,
//     gm/ugm.cc:1656-1770) and one son per triangle of each side, a side
//     with one refined edge cut in two, one with three in four, each son's
//     fourth corner the centre node; green elements are not refined again;
//   * every other element is copied (yellow).  UG copies only a
//     neighbourhood of the refined region; here every level keeps copies of
//     the whole cube (the far copies have class 0 and are off the surface).
// Vector classes: NCLASS 3 on the corners of the red and green elements of
// a level (refine.cc:6759-6790), NNCLASS 3 on the corners of the elements
// refined red or green (refine.cc:2890-2910); VCLASS / VNCLASS 3 on the
// vectors of elements whose corners all carry node class 3, then 2 on their
// matrix neighbours and 1 on the neighbours of those (SetSurfaceClasses,
// gm/algebra.cc:2838-2930; PropagateVectorClasses, 4063-4190); level 0
// VCLASS 3, top level VNCLASS 0.  Level l+1 lists the sons of level l's
// vectors in their order, then the edge midpoints (ascending edge), then
// the centre nodes (element order).  Interpolation (GNs of the father,
// gm/shapes.cc:748-819): sons 1, midpoints 1/2 to the edge's end points
// (lower first), centre nodes 1/4 to the father's corners in corner order.
// Clamped (all components Dirichlet) on x = 0; no load here (the surface
// solve takes one right-hand side per level).
namespace {

enum : uint8_t { kTetRed = 0, kTetGreen = 1, kTetCopy = 2 };

struct TetMesh {
  std::vector<double> X;            // 3 per node
  std::vector<int> T;               // 4 per element, UG corner order
  std::vector<uint8_t> kind;        // kTetRed / kTetGreen / kTetCopy
  int nodes() const { return (int)(X.size() / 3); }
  int64_t elems() const { return (int64_t)kind.size(); }
};

// UG's reference tetrahedron (gm/elements.cc:121-135): corners of each
// edge and of each side
const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
const int kTetSide[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};

inline uint64_t edge_key(int a, int b, int n) {
  return a < b ? (uint64_t)a * (uint64_t)n + (uint64_t)b : (uint64_t)b * (uint64_t)n + (uint64_t)a;
}

// refined[e] = 1 for elements of c refined red or green (NNCLASS seeds);
// fine gets its transfer rows (pp/pi/pw) and f its mesh
int refine_tets(const TetMesh &c, double a, TetMesh &f, std::vector<uint8_t> &refined, Level &fine) {
  const int n = c.nodes();
  const int64_t ne = c.elems();
  std::vector<uint8_t> marked(ne, 0);
  std::vector<uint64_t> ref;
  for (int64_t e = 0; e < ne; ++e) {
    if (c.kind[e] != kTetRed) continue;
    bool in = true;
    for (int q = 0; q < 4 && in; ++q)
      for (int d = 0; d < 3; ++d) in = in && c.X[(size_t)c.T[e * 4 + q] * 3 + d] <= a + 1e-12;
    if (!in) continue;
    marked[e] = 1;
    for (int k = 0; k < 6; ++k) ref.push_back(edge_key(c.T[e * 4 + kTetEdge[k][0]], c.T[e * 4 + kTetEdge[k][1]], n));
  }
  std::sort(ref.begin(), ref.end());
  ref.erase(std::unique(ref.begin(), ref.end()), ref.end());
  auto mid = [&](int u, int w) -> int {   // midpoint node of a refined edge, -1 if none
    const uint64_t k = edge_key(u, w, n);
    auto it = std::lower_bound(ref.begin(), ref.end(), k);
    return (it != ref.end() && *it == k) ? n + (int)(it - ref.begin()) : -1;
  };
  // green closure: unmarked elements with a refined edge
  std::vector<uint8_t> green(ne, 0);
  int64_t ngreen = 0;
  for (int64_t e = 0; e < ne; ++e) {
    if (marked[e]) continue;
    const int *t = &c.T[e * 4];
    int cnt = 0;
    for (int k = 0; k < 6; ++k) cnt += mid(t[kTetEdge[k][0]], t[kTetEdge[k][1]]) >= 0;
    if (!cnt) continue;
    if (c.kind[e] != kTetRed) return -1;   // only regular elements can be closed
    for (int s = 0; s < 4; ++s) {
      int m = 0;
      for (int j = 0; j < 3; ++j) m += mid(t[kTetSide[s][j]], t[kTetSide[s][(j + 1) % 3]]) >= 0;
      if (m == 2) return -1;               // not produced by refinement toward a corner
    }
    green[e] = 1;
    ++ngreen;
  }
  refined.assign(ne, 0);
  for (int64_t e = 0; e < ne; ++e) refined[e] = marked[e] | green[e];
  // nodes: sons, edge midpoints, centres
  const int nm = (int)ref.size();
  const int64_t nf = (int64_t)n + nm + ngreen;
  if (nf > 2000000000LL) return -1;
  f.X.resize((size_t)nf * 3);
  std::copy(c.X.begin(), c.X.end(), f.X.begin());
  fine.pp.assign(nf + 1, 0);
  fine.pi.clear();
  fine.pw.clear();
  for (int v = 0; v < n; ++v) {
    fine.pi.push_back(v);
    fine.pw.push_back(1.0);
    fine.pp[v + 1] = (int)fine.pi.size();
  }
  for (int k = 0; k < nm; ++k) {
    const int u = (int)(ref[k] / (uint64_t)n), w = (int)(ref[k] % (uint64_t)n);
    for (int d = 0; d < 3; ++d) f.X[(size_t)(n + k) * 3 + d] = 0.5 * (c.X[(size_t)u * 3 + d] + c.X[(size_t)w * 3 + d]);
    fine.pi.push_back(u);
    fine.pw.push_back(0.5);
    fine.pi.push_back(w);
    fine.pw.push_back(0.5);
    fine.pp[n + k + 1] = (int)fine.pi.size();
  }
  // sons in father element order
  f.T.clear();
  f.kind.clear();
  f.T.reserve((size_t)ne * 4 + (size_t)(ne - ngreen) * 28);
  int next_centre = n + nm;
  auto son = [&](int p, int q, int r, int s, uint8_t kind) {
    f.T.insert(f.T.end(), {p, q, r, s});
    f.kind.push_back(kind);
  };
  for (int64_t e = 0; e < ne; ++e) {
    const int *t = &c.T[e * 4];
    if (marked[e]) {
      const int x0 = t[0], x1 = t[1], x2 = t[2], x3 = t[3];
      const int m01 = mid(x0, x1), m02 = mid(x0, x2), m03 = mid(x0, x3);
      const int m12 = mid(x1, x2), m13 = mid(x1, x3), m23 = mid(x2, x3);
      son(x0, m01, m02, m03, kTetRed);
      son(m01, x1, m12, m13, kTetRed);
      son(m02, m12, x2, m23, kTetRed);
      son(m03, m13, m23, x3, kTetRed);
      son(m01, m02, m03, m13, kTetRed);
      son(m01, m02, m12, m13, kTetRed);
      son(m02, m03, m13, m23, kTetRed);
      son(m02, m12, m13, m23, kTetRed);
    } else if (green[e]) {
      const int ctr = next_centre++;
      for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int q = 0; q < 4; ++q) s += c.X[(size_t)t[q] * 3 + d];
        f.X[(size_t)ctr * 3 + d] = 0.25 * s;
      }
      for (int q = 0; q < 4; ++q) {
        fine.pi.push_back(t[q]);
        fine.pw.push_back(0.25);
      }
      fine.pp[ctr + 1] = (int)fine.pi.size();
      for (int s = 0; s < 4; ++s) {
        const int c0 = t[kTetSide[s][0]], c1 = t[kTetSide[s][1]], c2 = t[kTetSide[s][2]];
        const int cs[3] = {c0, c1, c2};
        const int m[3] = {mid(c0, c1), mid(c1, c2), mid(c2, c0)};
        const int k = (m[0] >= 0) + (m[1] >= 0) + (m[2] >= 0);
        // a side triangle (p, q, r) becomes the son with corners p, r, q and
        // the centre (the side's corners fill son corners 0, 2, 1)
        auto tri = [&](int p, int q, int r) { son(p, r, q, ctr, kTetGreen); };
        if (k == 0) {
          tri(c0, c1, c2);
        } else if (k == 1) {
          int j = 0;
          while (m[j] < 0) ++j;
          tri(m[j], cs[(j + 1) % 3], cs[(j + 2) % 3]);
          tri(m[j], cs[(j + 2) % 3], cs[j]);
        } else {
          tri(m[0], c1, m[1]);
          tri(m[0], m[2], c0);
          tri(m[1], c2, m[2]);
          tri(m[0], m[1], m[2]);
        }
      }
    } else {
      son(t[0], t[1], t[2], t[3], kTetCopy);
    }
  }
  fine.n = (int)nf;
  return 0;
}

// level matrix from the element list: pattern = vectors sharing an element
// (diagonal first, then ascending columns), element blocks added in element
// order, Dirichlet rows on x = 0
void assemble_tet_level(const TetMesh &m, double lam, double mu, Level &L) {
  const int n = m.nodes();
  const int64_t ne = m.elems();
  std::vector<uint64_t> keys;
  keys.reserve((size_t)ne * 16);
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const int r = m.T[e * 4 + i], c = m.T[e * 4 + j];
        keys.push_back((uint64_t)r * ((uint64_t)n + 1) + (r == c ? 0 : (uint64_t)c + 1));
      }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  L.n = n;
  L.rp.assign(n + 1, 0);
  L.ci.resize(keys.size());
  for (size_t q = 0; q < keys.size(); ++q) {
    const int r = (int)(keys[q] / ((uint64_t)n + 1));
    const uint64_t code = keys[q] % ((uint64_t)n + 1);
    L.ci[q] = code == 0 ? r : (int)(code - 1);
    ++L.rp[r + 1];
  }
  for (int r = 0; r < n; ++r) L.rp[r + 1] += L.rp[r];
  L.v.assign(keys.size() * 9, 0.0);
  auto slot = [&](int r, int c) -> size_t {
    if (r == c) return (size_t)L.rp[r];
    const int *b = &L.ci[L.rp[r] + 1], *en = &L.ci[L.rp[r + 1]];
    return (size_t)(std::lower_bound(b, en, c) - &L.ci[0]);
  };
  for (int64_t e = 0; e < ne; ++e) {
    double X[4][3];
    for (int q = 0; q < 4; ++q)
      for (int d = 0; d < 3; ++d) X[q][d] = m.X[(size_t)m.T[e * 4 + q] * 3 + d];
    double K[4][4][9], M[4][4];
    p1_elasticity_element(X, lam, mu, K, M);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double *dst = &L.v[slot(m.T[e * 4 + i], m.T[e * 4 + j]) * 9];
        for (int q = 0; q < 9; ++q) dst[q] += K[i][j][q];
      }
  }
  L.skip.assign(n, 0u);
  for (int r = 0; r < n; ++r) {
    if (m.X[(size_t)r * 3] != 0.0) continue;
    L.skip[r] = 7u;
    for (int64_t q = L.rp[r]; q < L.rp[r + 1]; ++q)
      for (int k = 0; k < 9; ++k) L.v[(size_t)q * 9 + k] = 0.0;
    for (int a = 0; a < 3; ++a) L.v[(size_t)L.rp[r] * 9 + a * 4] = 1.0;
  }
}

// class 3 on the vectors of elements whose corners all carry node class 3,
// then 2 on matrix neighbours of class 3 and 1 on neighbours of class 2
std::vector<uint8_t> tet_classes(const TetMesh &m, const std::vector<uint8_t> &node3, const Level &L) {
  const int n = L.n;
  std::vector<uint8_t> cls(n, 0);
  for (int64_t e = 0; e < m.elems(); ++e) {
    const int *t = &m.T[e * 4];
    if (node3[t[0]] && node3[t[1]] && node3[t[2]] && node3[t[3]])
      for (int q = 0; q < 4; ++q) cls[t[q]] = 3;
  }
  for (int k = 3; k >= 2; --k) {
    std::vector<uint8_t> hit(n, 0);
    for (int r = 0; r < n; ++r) {
      if (cls[r] != k) continue;
      for (int q = L.rp[r] + 1; q < L.rp[r + 1]; ++q) hit[L.ci[q]] = 1;
    }
    for (int r = 0; r < n; ++r)
      if (hit[r] && cls[r] < k) cls[r] = (uint8_t)(k - 1);
  }
  return cls;
}

}  // namespace

extern "C" int ugmg_problem_elasticity_local(int nuni, int depth, double E, double nu,
                                             ugmg_problem **out) {
  if (!out || nuni < 0 || depth < 0 || nuni + depth > 9 || !(E > 0.0) || !(nu > -1.0 && nu < 0.5))
    return -1;
  ugmg_problem *p = new (std::nothrow) ugmg_problem;
  if (!p) return -4;
  p->dim = 3;
  p->b = 3;
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
  const int L = nuni + depth;
  try {
    std::vector<TetMesh> mesh(L + 1);
    TetMesh &m0 = mesh[0];
    for (int v = 0; v < 8; ++v) m0.X.insert(m0.X.end(), {(double)(v & 1), (double)((v >> 1) & 1), (double)((v >> 2) & 1)});
    for (int t = 0; t < 6; ++t) {
      int id = 0, path[4] = {0, 0, 0, 0};
      for (int q = 1; q < 4; ++q) path[q] = (id |= 1 << kPerm[t][q - 1]);
      m0.T.insert(m0.T.end(), {path[0], path[1], path[2], path[3]});
      m0.kind.push_back(kTetRed);
    }
    p->lv.resize(L + 1);
    std::vector<std::vector<uint8_t>> refined(L);
    for (int s = 0; s < L; ++s) {
      const double a = s < nuni ? 1.0 : std::ldexp(1.0, -(s - nuni + 1));
      if (refine_tets(mesh[s], a, mesh[s + 1], refined[s], p->lv[s + 1])) {
        delete p;
        return -1;
      }
    }
    for (int l = 0; l <= L; ++l) {
      Level &lv = p->lv[l];
      assemble_tet_level(mesh[l], lam, mu, lv);
      lv.N = 0;
      lv.C = 0;
      if (l == 0) {
        lv.vc.assign(lv.n, 3);
      } else {
        std::vector<uint8_t> nc3(lv.n, 0);
        for (int64_t e = 0; e < mesh[l].elems(); ++e)
          if (mesh[l].kind[e] != kTetCopy)
            for (int q = 0; q < 4; ++q) nc3[mesh[l].T[e * 4 + q]] = 1;
        lv.vc = tet_classes(mesh[l], nc3, lv);
      }
      if (l == L) {
        lv.vnc.assign(lv.n, 0);
      } else {
        std::vector<uint8_t> nn3(lv.n, 0);
        for (int64_t e = 0; e < mesh[l].elems(); ++e)
          if (refined[l][e])
            for (int q = 0; q < 4; ++q) nn3[mesh[l].T[e * 4 + q]] = 1;
        lv.vnc = tet_classes(mesh[l], nn3, lv);
      }
    }
    p->rhs.assign((size_t)p->lv[L].n * 3, 0.0);
    p->xyz.resize(L + 1);
    p->tets.resize(L + 1);
    p->tet_kind.resize(L + 1);
    for (int l = 0; l <= L; ++l) {
      p->xyz[l].swap(mesh[l].X);
      p->tets[l].swap(mesh[l].T);
      p->tet_kind[l].swap(mesh[l].kind);
    }
  } catch (const std::bad_alloc &) {
    delete p;
    return -4;
  }
  *out = p;
  return 0;
}

extern "C" int ugmg_problem_mesh(const ugmg_problem *p, int level, int *nnodes, const double **xyz,
                                 int *nelem, const int **tets, const uint8_t **kind) {
  if (!p || level < 0 || level >= (int)p->xyz.size()) return -1;
  if (nnodes) *nnodes = (int)(p->xyz[level].size() / 3);
  if (xyz) *xyz = p->xyz[level].data();
  if (nelem) *nelem = (int)p->tet_kind[level].size();
  if (tets) *tets = p->tets[level].data();
  if (kind) *kind = p->tet_kind[level].data();
  return 0;
}

// ------------------------------------------------------------ element data
// The finest level as its element loop sees it (the data an application
// hands to AddVlistMValues, np/udm/disctools.cc:841): element e's corners
// and its (nc b) x (nc b) matrix, computed by the same code, from the same
// SplitMix64 draws, as the generators above, so that summing the element
// matrices in element order reproduces their finest levels bit for bit.

namespace {

ugmg_element_data *new_element_data(int n, int b, int64_t nelem, int nc, bool full) {
  ugmg_element_data *d = new ugmg_element_data();
  d->n = n;
  d->b = b;
  d->nelem = (int)nelem;
  d->nc = nc;
  const size_t m = (size_t)nc * b;
  d->elem = new int[(size_t)nelem * nc];
  d->skip = new uint32_t[n]();
  if (full) {
    d->emat = new double[(size_t)nelem * m * m];
  } else {
    d->scale = new double[nelem];
    d->kref = new double[m * m];
  }
  return d;
}

}  // namespace

extern "C" void ugmg_element_data_free(ugmg_element_data *d) {
  if (!d) return;
  delete[] d->elem;
  delete[] d->skip;
  delete[] d->emat;
  delete[] d->scale;
  delete[] d->kref;
  delete d;
}

extern "C" int ugmg_problem_q1_elements(int dim, int coarse_cells, int nref, int kappa_mode,
                                        uint64_t seed, ugmg_element_data **out) {
  if (!out || (dim != 2 && dim != 3) || coarse_cells < 1 || nref < 0 || nref > 12 ||
      (kappa_mode != UGMG_KAPPA_ONE && kappa_mode != UGMG_KAPPA_JUMP))
    return -1;
  const int C = coarse_cells << nref, N = C + 1, nc = 1 << dim;
  const int64_t nel = dim == 2 ? (int64_t)C * C : (int64_t)C * C * C;
  const int64_t n = dim == 2 ? (int64_t)N * N : (int64_t)N * N * N;
  if (nel * nc > 2000000000LL) return -1;
  ugmg_element_data *d;
  try {
    d = new_element_data((int)n, 1, nel, nc, false);
  } catch (const std::bad_alloc &) {
    return -4;
  }
  uint64_t s = seed;
  for (int64_t e = 0; e < nel; ++e)
    d->scale[e] = kappa_mode == UGMG_KAPPA_JUMP ? std::pow(10.0, -2.0 + 4.0 * splitmix_u01(s)) : 1.0;
  q1_stiffness(dim, 1.0 / C, d->kref);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nel; ++e) {
    const int x = (int)(e % C), y = (int)((e / C) % C), z = dim == 3 ? (int)(e / ((int64_t)C * C)) : 0;
    for (int a = 0; a < nc; ++a)
      d->elem[e * nc + a] = (int)node_id(dim, N, x + (a & 1), y + ((a >> 1) & 1), z + ((a >> 2) & 1));
  }
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % N), y = (int)((i / N) % N), z = dim == 3 ? (int)(i / ((int64_t)N * N)) : 0;
    d->skip[i] = (x == 0 || x == N - 1 || y == 0 || y == N - 1 || (dim == 3 && (z == 0 || z == N - 1))) ? 1u : 0u;
  }
  *out = d;
  return 0;
}

extern "C" int ugmg_problem_p1_cd_elements(int nref, double eps, double theta, uint64_t seed,
                                           ugmg_element_data **out) {
  if (!out || nref < 0 || nref > 11 || !(eps > 0.0)) return -1;
  uint64_t s = seed;
  if (std::isnan(theta)) theta = 2.0 * M_PI * splitmix_u01(s);
  const double bvec[2] = {std::cos(theta), std::sin(theta)};
  const int C = 2 << nref, N = C + 1;
  const double h = 1.0 / C;
  const double pe = h / (2.0 * eps);
  const double tau = pe > 1e-8 ? h / 2.0 * (1.0 / std::tanh(pe) - 1.0 / pe) : 0.0;
  ugmg_element_data *d;
  try {
    d = new_element_data(N * N, 1, (int64_t)C * C * 2, 3, true);
  } catch (const std::bad_alloc &) {
    return -4;
  }
  for (int cy = 0; cy < C; ++cy)
    for (int cx = 0; cx < C; ++cx)
      for (int t = 0; t < 2; ++t) {
        const int64_t e = ((int64_t)cy * C + cx) * 2 + t;
        const int vx[2][3] = {{cx, cx + 1, cx + 1}, {cx, cx + 1, cx}};
        const int vy[2][3] = {{cy, cy, cy + 1}, {cy, cy + 1, cy + 1}};
        double xy[3][2];
        for (int q = 0; q < 3; ++q) {
          xy[q][0] = vx[t][q] * h;
          xy[q][1] = vy[t][q] * h;
          d->elem[e * 3 + q] = vy[t][q] * N + vx[t][q];
        }
        double K[3][3], M[3][3], G[3];
        p1_element(xy, eps, bvec, tau, K, M, G);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) d->emat[e * 9 + i * 3 + j] = K[i][j];
      }
  for (int i = 0; i < N * N; ++i) {
    const int x = i % N, y = i / N;
    d->skip[i] = (x == 0 || x == N - 1 || y == 0 || y == N - 1) ? 1u : 0u;
  }
  *out = d;
  return 0;
}

extern "C" int ugmg_problem_elasticity_elements(int nref, double E, double nu, uint64_t seed,
                                                ugmg_element_data **out) {
  (void)seed;   // the element matrices do not depend on the body force
  if (!out || nref < 0 || nref > 8 || !(E > 0.0) || !(nu > -1.0 && nu < 0.5)) return -1;
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
  const int C = 1 << nref, N = C + 1;
  const double h = 1.0 / C;
  const int64_t nel = (int64_t)C * C * C * 6;
  ugmg_element_data *d;
  try {
    d = new_element_data(N * N * N, 3, nel, 4, true);
  } catch (const std::bad_alloc &) {
    return -4;
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)C * C * C; ++c) {
    const int cx = (int)(c % C), cy = (int)((c / C) % C), cz = (int)(c / ((int64_t)C * C));
    for (int t = 0; t < 6; ++t) {
      const int64_t e = c * 6 + t;
      int vx[4][3] = {{0, 0, 0}};
      for (int q = 1; q < 4; ++q) {
        for (int dd = 0; dd < 3; ++dd) vx[q][dd] = vx[q - 1][dd];
        vx[q][kPerm[t][q - 1]] = 1;
      }
      double X[4][3];
      for (int q = 0; q < 4; ++q) {
        const int cc[3] = {cx + vx[q][0], cy + vx[q][1], cz + vx[q][2]};
        for (int dd = 0; dd < 3; ++dd) X[q][dd] = cc[dd] * h;
        d->elem[e * 4 + q] = cc[0] + N * (cc[1] + N * cc[2]);
      }
      double K[4][4][9], M[4][4];
      p1_elasticity_element(X, lam, mu, K, M);
      double *em = d->emat + (size_t)e * 144;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) em[(i * 3 + a) * 12 + j * 3 + bb] = K[i][j][a * 3 + bb];
    }
  }
  for (int i = 0; i < N * N * N; ++i) d->skip[i] = (i % N == 0) ? 7u : 0u;
  *out = d;
  return 0;
}

extern "C" void ugmg_problem_free(ugmg_problem *p) { delete p; }

extern "C" int ugmg_problem_nlevels(const ugmg_problem *p) {
  return p ? (int)p->lv.size() : -1;
}
extern "C" int ugmg_problem_block(const ugmg_problem *p) { return p ? p->b : -1; }

extern "C" int ugmg_problem_level(const ugmg_problem *p, int level, int *n,
                                  int *nnzb, const int **row_ptr,
                                  const int **col_idx, const double **vals,
                                  const uint32_t **skip, const uint8_t **vclass,
                                  const uint8_t **vnclass) {
  if (!p || level < 0 || level >= (int)p->lv.size()) return -1;
  const Level &L = p->lv[level];
  if (n) *n = L.n;
  if (nnzb) *nnzb = L.rp[L.n];
  if (row_ptr) *row_ptr = L.rp.data();
  if (col_idx) *col_idx = L.ci.data();
  if (vals) *vals = L.v.data();
  if (skip) *skip = L.skip.data();
  if (vclass) *vclass = L.vc.data();
  if (vnclass) *vnclass = L.vnc.data();
  return 0;
}

extern "C" int ugmg_problem_transfer(const ugmg_problem *p, int level, int *nnz,
                                     const int **p_ptr, const int **p_idx,
                                     const double **p_w) {
  if (!p || level < 1 || level >= (int)p->lv.size()) return -1;
  const Level &L = p->lv[level];
  if (nnz) *nnz = L.pp[L.n];
  if (p_ptr) *p_ptr = L.pp.data();
  if (p_idx) *p_idx = L.pi.data();
  if (p_w) *p_w = L.pw.data();
  return 0;
}

extern "C" int ugmg_problem_rhs(const ugmg_problem *p, const double **b) {
  if (!p || !b) return -1;
  *b = p->rhs.data();
  return 0;
}

extern "C" int ugmg_problem_grid(const ugmg_problem *p, int level, int *dim,
                                 int *nodes_per_dir) {
  if (!p || level < 0 || level >= (int)p->lv.size()) return -1;
  if (dim) *dim = p->dim;
  if (nodes_per_dir) *nodes_per_dir = p->lv[level].N;
  return 0;
}
