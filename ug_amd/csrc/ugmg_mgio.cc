// UG's multigrid files: the sequential reader of gm/mgio.cc's sections and
// the rebuild of the levels as LoadMultiGrid / InsertLocalTree
// (gm/ugio.cc:2299-2625) performs it, with the node-vector interpolation of
// StandardIntCorNodeVector (np/algebra/transgrid.cc:235-336).  Host only;
// the layout is described in include/ugmg_mgio.h.
#include "../../include/ugmg_mgio.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ugmg_bio.h"

namespace {

using ugmg_bio::Bio;
using ugmg_bio::File;

constexpr const char *kTitle = "####.sparse.mg.storage.format.####";   // mgio.cc:64
constexpr int kTags = 8;                                                // MGIO_TAGS
constexpr int kMaxCorners = 8, kMaxEdges = 12, kMaxSides = 6, kMaxSons = 30;

// MGIO_ECTRL fields of a refinement record's control word (mgio.cc:73-98)
inline int ectrl(unsigned c, int shift, int len) { return (int)((c >> shift) & ((1u << len) - 1u)); }

struct GeElem {   // MGIO_GE_ELEMENT
  int tag = 0, nc = 0, ne = 0, ns = 0;
  int ce[kMaxEdges][2] = {};
  int cs[kMaxSides][4] = {};
};
struct Son {
  int tag = 0;
  int corners[kMaxCorners] = {};
  int nb[kMaxSides] = {};
  int path = 0;
};
struct Rule {   // MGIO_RR_RULE
  int rclass = 0, nsons = 0;
  std::vector<int> pattern;
  std::vector<Son> sons;
};

struct Node {
  int id = -1, level = 0;
  double x[3] = {0, 0, 0};
  int father_node = -1;             // a corner's son: the father node (global index)
  int father_elem = -1;             // other new nodes: the element that created them
  std::vector<int> wc;              // ... its corners (element corner numbers) with
  std::vector<double> ww;           // ... nonzero GNs weights, corner order
};
struct Elem {
  int tag = 0, level = 0;
  std::vector<int> corners;   // global node indices
};

}  // namespace

struct ugmg_mg {
  ugmg_mg_info info{};
  int dim = 3;
  std::vector<GeElem> ge;
  std::vector<Rule> rules;
  std::vector<Node> nodes;
  std::vector<Elem> elems;
  std::map<int, int> by_id;   // node ID -> global node index
  // per level, filled after the rebuild
  std::vector<std::vector<int>> lev_nodes;   // global node indices in ID order
  std::vector<std::vector<int>> lev_elems;
  std::vector<int> node_pos;                 // global node index -> index on its level
};

namespace {

struct Reader {
  ugmg_mg &mg;
  Bio io;
  std::string err;
  int max_new = 19;   // MGIO_MAX_NEW_CORNERS: 5 in 2D, 19 in 3D

  bool fail(const std::string &m) {
    if (err.empty()) err = m;
    return false;
  }
  int center_index(int tag) const {   // CENTER_NODE_INDEX (rm.cc:3778-4050)
    if (mg.dim == 2) return 4;
    return tag == 4 ? 10 : 18;
  }
  const GeElem *gel(int tag) {
    if (tag < 0 || tag >= kTags || mg.ge[tag].nc <= 0) return nullptr;
    return &mg.ge[tag];
  }

  bool head() {
    char title[256];
    if (!io.read_string(title, sizeof title) || std::strcmp(title, kTitle) != 0)
      return fail("not a multigrid file (title line)");
    int mode;
    if (!io.read_ints(1, &mode) || !ugmg_bio::valid_mode(mode)) return fail("bad mode");
    io.mode = mode;
    ugmg_mg_info &in = mg.info;
    in.mode = mode;
    if (!io.read_string(in.version, sizeof in.version)) return fail("version");
    if (std::strcmp(in.version, "UG_IO_2.2") == 0) std::strcpy(in.version, "UG_IO_2.3");
    if (std::strncmp(in.version, "UG_IO_2.", 8) != 0) return fail(std::string("version ") + in.version);
    if (!io.read_string(in.ident, sizeof in.ident) || !io.read_string(in.domain, sizeof in.domain) ||
        !io.read_string(in.multigrid, sizeof in.multigrid) || !io.read_string(in.format, sizeof in.format))
      return fail("names");
    int v[11];
    if (!io.read_ints(11, v)) return fail("general counts");
    in.dim = v[0];
    in.magic_cookie = v[1];
    in.heapsize = v[2];
    in.nlevels = v[3];
    in.nnodes = v[4];
    in.npoints = v[5];
    in.nelements = v[6];
    if (v[9] != 1) return fail("parallel multigrid files (nparfiles > 1) are not read");
    if (v[10] != 0) return fail("files with MGIO debug extensions are not read");
    if (in.dim != 2 && in.dim != 3) return fail("dimension");
    if (in.nlevels < 1 || in.nlevels > 32) return fail("level count");
    mg.dim = in.dim;
    max_new = in.dim == 2 ? 5 : 19;
    return true;
  }

  bool elements_and_rules() {
    int n;
    if (!io.read_ints(1, &n)) return fail("element types");
    mg.ge.assign(kTags, GeElem());
    for (int i = 0; i < kTags; ++i) {   // Read_GE_Elements(TAGS, ..)
      int v[4];
      if (!io.read_ints(4, v)) return fail("element type");
      GeElem &g = mg.ge[i];
      g.tag = v[0];
      g.nc = v[1];
      g.ne = v[2];
      g.ns = v[3];
      if (g.nc < 0 || g.nc > kMaxCorners || g.ne < 0 || g.ne > kMaxEdges || g.ns < 0 || g.ns > kMaxSides)
        return fail("element type sizes");
      if (g.ne > 0 || g.ns > 0) {
        std::vector<int> w(2 * g.ne + 4 * g.ns);
        if (!io.read_ints((int)w.size(), w.data())) return fail("element type tables");
        int s = 0;
        for (int j = 0; j < g.ne; ++j) {
          g.ce[j][0] = w[s++];
          g.ce[j][1] = w[s++];
        }
        for (int j = 0; j < g.ns; ++j)
          for (int k = 0; k < 4; ++k) g.cs[j][k] = w[s++];
      }
    }
    int rg[1 + kTags];
    if (!io.read_ints(1 + kTags, rg)) return fail("rule counts");
    mg.info.nrules = rg[0];
    if (rg[0] < 0 || rg[0] > (1 << 18)) return fail("rule count");
    mg.rules.assign(rg[0], Rule());
    for (Rule &r : mg.rules) {   // Read_RR_Rules
      int h[2];
      if (!io.read_ints(2, h)) return fail("rule head");
      r.rclass = h[0];
      r.nsons = h[1];
      if (r.nsons < 0 || r.nsons > kMaxSons) return fail("rule sons");
      const int m = 3 * max_new + r.nsons * (1 + kMaxCorners + kMaxSides + 1);
      std::vector<int> w(m);
      if (!io.read_ints(m, w.data())) return fail("rule body");
      int s = 0;
      r.pattern.assign(w.begin(), w.begin() + max_new);
      s = 3 * max_new;   // pattern, then sonandnode pairs
      r.sons.resize(r.nsons);
      for (Son &so : r.sons) {
        so.tag = w[s++];
        for (int k = 0; k < kMaxCorners; ++k) so.corners[k] = w[s++];
        for (int k = 0; k < kMaxSides; ++k) so.nb[k] = w[s++];
        so.path = w[s++];
      }
    }
    return true;
  }

  // level-0 points and elements; the records of every refined element
  bool coarse_grid_and_tree() {
    int cg[6];
    if (!io.read_ints(6, cg)) return fail("coarse grid counts");
    const int np = cg[0], ne = cg[3];
    if (np < 0 || ne < 0 || cg[1] + cg[2] != np) return fail("coarse grid counts");
    mg.info.ncoarse_points = np;
    mg.info.ncoarse_bnd_points = cg[1];
    mg.info.ncoarse_elements = ne;
    const int d = mg.dim;
    std::vector<double> pos((size_t)np * d);
    for (int i = 0; i < np; ++i)
      if (!io.read_doubles(d, pos.data() + (size_t)i * d)) return fail("coarse points");
    mg.nodes.resize(np);
    for (int i = 0; i < np; ++i) {   // level-0 node IDs are vertex IDs
      Node &nd = mg.nodes[i];
      nd.id = i;
      nd.level = 0;
      for (int k = 0; k < d; ++k) nd.x[k] = pos[(size_t)i * d + k];
      mg.by_id[i] = i;
    }
    std::vector<int> nref(ne);
    for (int e = 0; e < ne; ++e) {   // Read_CG_Elements
      int tag;
      if (!io.read_ints(1, &tag)) return fail("coarse element");
      const GeElem *g = gel(tag);
      if (!g) return fail("coarse element type");
      std::vector<int> w(g->nc + g->ns + 3);
      if (!io.read_ints((int)w.size(), w.data())) return fail("coarse element body");
      nref[e] = w[0];
      Elem el;
      el.tag = tag;
      el.level = 0;
      for (int j = 0; j < g->nc; ++j) {
        const int id = w[1 + j];
        if (id < 0 || id >= np) return fail("coarse corner ID");
        el.corners.push_back(id);
      }
      mg.elems.push_back(el);
    }
    // the domain's boundary points: a section of counted length
    if (!io.skip_jump()) return fail("boundary section");
    for (int e = 0; e < ne; ++e)
      if (nref[e] > 0 && !insert_tree(e)) return false;
    return true;
  }

  // a new node of the level above: its ID, position and father data
  int new_node(int id, int level, const double *x, int father_node, int father_elem,
               std::vector<int> wc, std::vector<double> ww) {
    auto it = mg.by_id.find(id);
    if (it != mg.by_id.end()) return it->second;
    Node nd;
    nd.id = id;
    nd.level = level;
    for (int k = 0; k < 3; ++k) nd.x[k] = x[k];
    nd.father_node = father_node;
    nd.father_elem = father_elem;
    nd.wc = std::move(wc);
    nd.ww = std::move(ww);
    mg.nodes.push_back(nd);
    const int g = (int)mg.nodes.size() - 1;
    mg.by_id[id] = g;
    return g;
  }

  // InsertLocalTree (gm/ugio.cc:2299-2625), sequential part
  bool insert_tree(int ei) {
    int h[2];
    if (!io.read_ints(2, h)) return fail("refinement record");
    const unsigned ctrl = (unsigned)h[0];
    const int sonref = h[1];
    const int refrule = ectrl(ctrl, 10, 18) - 1;
    if (refrule < 0) return fail("refinement record without a rule");
    if (refrule >= (int)mg.rules.size()) return fail("refinement rule out of range");
    const int nnew = ectrl(ctrl, 0, 5), nmoved = ectrl(ctrl, 5, 5);
    std::vector<int> ids(nnew + nmoved);
    if (!ids.empty() && !io.read_ints((int)ids.size(), ids.data())) return fail("new corner IDs");
    std::vector<double> mv((size_t)nmoved * mg.dim);
    if (nmoved > 0 && !io.read_doubles((long long)mv.size(), mv.data())) return fail("moved corners");
    const Rule &rule = mg.rules[refrule];
    const Elem fe = mg.elems[ei];   // copy: elems grows below
    const GeElem *g = gel(fe.tag);
    if (!g) return fail("element type");
    const int up = fe.level + 1;
    if (up >= mg.info.nlevels) return fail("refinement above the top level");
    const int d = mg.dim;
    auto pos = [&](int gidx) { return mg.nodes[gidx].x; };
    // moved corners: positions by ID
    auto moved = [&](int id, double *x) {
      for (int m = 0; m < nmoved; ++m)
        if (ids[nnew + m] == id) {
          for (int k = 0; k < d; ++k) x[k] = mv[(size_t)m * d + k];
          return true;
        }
      return false;
    };
    // NodeList: corners, edges, (3D) sides, centre
    std::vector<int> nl(g->nc + max_new, -1);
    int r = 0;
    auto take = [&]() -> int {
      if (r >= nnew) return -1;
      return ids[r++];
    };
    for (int i = 0; i < g->nc; ++i) {
      const int id = take();
      if (id < 0) return fail("too few new corner IDs");
      const int fn = fe.corners[i];
      double x[3];
      for (int k = 0; k < 3; ++k) x[k] = pos(fn)[k];
      moved(id, x);
      nl[i] = new_node(id, up, x, fn, -1, {}, {});
    }
    // weights of a point given as an average of some of the element's
    // corners (GNs at that local point for the element types read here)
    auto avg_node = [&](int id, const std::vector<int> &cs) -> int {
      double x[3] = {0, 0, 0};
      for (int c : cs)
        for (int k = 0; k < 3; ++k) x[k] += pos(fe.corners[c])[k] / (double)cs.size();
      std::vector<int> wc(cs);
      std::sort(wc.begin(), wc.end());
      std::vector<double> ww(wc.size(), 1.0 / (double)cs.size());
      if (moved(id, x)) {
        if (!(fe.tag == 3 && d == 2) && !(fe.tag == 4 && d == 3)) {
          fail("moved corners of non-simplex elements are not read");
          return -1;
        }
        bary(fe, x, wc, ww);
      }
      return new_node(id, up, x, -1, ei, wc, ww);
    };
    for (int e = 0; e < g->ne; ++e) {
      if (rule.pattern[e] != 1) continue;
      const int id = take();
      if (id < 0) return fail("too few new corner IDs");
      const int n = avg_node(id, {g->ce[e][0], g->ce[e][1]});
      if (n < 0) return false;
      nl[g->nc + e] = n;
    }
    if (d == 3)
      for (int s = 0; s < g->ns; ++s) {
        if (rule.pattern[g->ne + s] != 1) continue;
        const int id = take();
        if (id < 0) return fail("too few new corner IDs");
        std::vector<int> cs;
        for (int k = 0; k < 4; ++k)
          if (g->cs[s][k] >= 0) cs.push_back(g->cs[s][k]);
        const int n = avg_node(id, cs);
        if (n < 0) return false;
        nl[g->nc + g->ne + s] = n;
      }
    const int ci = center_index(fe.tag);
    if (ci < max_new && rule.pattern[ci] == 1) {
      if (fe.tag == 5) return fail("centre nodes of pyramids are not read");
      const int id = take();
      if (id < 0) return fail("too few new corner IDs");
      std::vector<int> cs(g->nc);
      for (int c = 0; c < g->nc; ++c) cs[c] = c;
      const int n = avg_node(id, cs);
      if (n < 0) return false;
      nl[g->nc + ci] = n;
    }
    if (r != nnew) return fail("new corner IDs left over");
    // the sons
    std::vector<int> sons(rule.nsons);
    for (int s = 0; s < rule.nsons; ++s) {
      const Son &so = rule.sons[s];
      const GeElem *sg = gel(so.tag);
      if (!sg) return fail("son element type");
      Elem el;
      el.tag = so.tag;
      el.level = up;
      for (int j = 0; j < sg->nc; ++j) {
        const int k = so.corners[j];
        if (k < 0 || k >= (int)nl.size() || nl[k] < 0) return fail("son corner not created by the rule");
        el.corners.push_back(nl[k]);
      }
      mg.elems.push_back(el);
      sons[s] = (int)mg.elems.size() - 1;
    }
    for (int s = 0; s < rule.nsons; ++s)
      if ((sonref >> s) & 1)
        if (!insert_tree(sons[s])) return false;
    return true;
  }

  // barycentric weights of x in a simplex (corner order, zeros dropped)
  void bary(const Elem &fe, const double *x, std::vector<int> &wc, std::vector<double> &ww) {
    const int d = mg.dim, n = d + 1;
    double A[3][3] = {}, b[3] = {};
    const double *x0 = mg.nodes[fe.corners[0]].x;
    for (int i = 0; i < d; ++i) {
      for (int j = 0; j < d; ++j) A[i][j] = mg.nodes[fe.corners[j + 1]].x[i] - x0[i];
      b[i] = x[i] - x0[i];
    }
    double lam[3] = {0, 0, 0};
    if (d == 2) {
      const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
      lam[0] = (b[0] * A[1][1] - A[0][1] * b[1]) / det;
      lam[1] = (A[0][0] * b[1] - b[0] * A[1][0]) / det;
    } else {
      const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                         A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
      for (int c = 0; c < 3; ++c) {
        double M[3][3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) M[i][j] = j == c ? b[i] : A[i][j];
        lam[c] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                  M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                  M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) /
                 det;
      }
    }
    wc.clear();
    ww.clear();
    double l0 = 1.0;
    for (int c = 0; c < d; ++c) l0 -= lam[c];
    const double w[4] = {l0, lam[0], lam[1], lam[2]};
    for (int c = 0; c < n; ++c)
      if (w[c] != 0.0) {
        wc.push_back(c);
        ww.push_back(w[c]);
      }
  }
};

}  // namespace

extern "C" {

int ugmg_mg_read(const char *path, ugmg_mg **out, char *err, int errlen) {
  auto say = [&](const std::string &m) {
    if (err && errlen > 0) {
      std::strncpy(err, m.c_str(), (size_t)errlen - 1);
      err[errlen - 1] = '\0';
    }
  };
  if (!path || !out) {
    say("null argument");
    return -1;
  }
  *out = nullptr;
  File f(std::fopen(path, "rb"));
  if (!f) {
    say(std::string("cannot open ") + path);
    return -2;
  }
  ugmg_mg *mg = new ugmg_mg();
  Reader rd{*mg, Bio{f.get(), ugmg_bio::ASCII}, {}};
  bool ok = false;
  try {
    ok = rd.head() && rd.elements_and_rules() && rd.coarse_grid_and_tree();
  } catch (const std::exception &e) {
    rd.err = std::string("allocation: ") + e.what();
  }
  if (ok) {   // levels in node-ID order
    const int nl = mg->info.nlevels;
    mg->lev_nodes.assign(nl, {});
    mg->lev_elems.assign(nl, {});
    for (int g = 0; g < (int)mg->nodes.size(); ++g) mg->lev_nodes[mg->nodes[g].level].push_back(g);
    for (auto &v : mg->lev_nodes)
      std::sort(v.begin(), v.end(), [&](int a, int b) { return mg->nodes[a].id < mg->nodes[b].id; });
    mg->node_pos.assign(mg->nodes.size(), -1);
    for (auto &v : mg->lev_nodes)
      for (int q = 0; q < (int)v.size(); ++q) mg->node_pos[v[q]] = q;
    for (int e = 0; e < (int)mg->elems.size(); ++e) mg->lev_elems[mg->elems[e].level].push_back(e);
  }
  if (!ok) {
    say(rd.err.empty() ? "read error" : rd.err);
    delete mg;
    return -1;
  }
  *out = mg;
  return 0;
}

void ugmg_mg_free(ugmg_mg *mg) { delete mg; }

int ugmg_mg_get_info(const ugmg_mg *mg, ugmg_mg_info *info) {
  if (!mg || !info) return -1;
  *info = mg->info;
  return 0;
}

int ugmg_mg_level(const ugmg_mg *mg, int level, int *nnodes, int *nelem, int *ncorners, int *np,
                  int *ids, double *xyz, int *elem_tag, int *elem_ptr, int *elem_corners, int *p_ptr,
                  int *p_idx, double *p_w) {
  if (!mg || level < 0 || level >= (int)mg->lev_nodes.size()) return -1;
  const auto &ln = mg->lev_nodes[level];
  const auto &le = mg->lev_elems[level];
  int nc = 0;
  for (int e : le) nc += (int)mg->elems[e].corners.size();
  int npp = 0;
  if (level > 0)
    for (int g : ln) npp += mg->nodes[g].father_node >= 0 ? 1 : (int)mg->nodes[g].wc.size();
  if (nnodes) *nnodes = (int)ln.size();
  if (nelem) *nelem = (int)le.size();
  if (ncorners) *ncorners = nc;
  if (np) *np = npp;
  const int d = mg->dim;
  for (int q = 0; q < (int)ln.size(); ++q) {
    const Node &nd = mg->nodes[ln[q]];
    if (ids) ids[q] = nd.id;
    if (xyz)
      for (int k = 0; k < d; ++k) xyz[(size_t)q * d + k] = nd.x[k];
  }
  if (elem_ptr) elem_ptr[0] = 0;
  int c = 0;
  for (int k = 0; k < (int)le.size(); ++k) {
    const Elem &el = mg->elems[le[k]];
    if (elem_tag) elem_tag[k] = el.tag;
    for (int g : el.corners) {
      if (elem_corners) elem_corners[c] = mg->node_pos[g];
      ++c;
    }
    if (elem_ptr) elem_ptr[k + 1] = c;
  }
  if (level == 0) {
    if (p_ptr)
      for (int q = 0; q <= (int)ln.size(); ++q) p_ptr[q] = 0;
    return 0;
  }
  int m = 0;
  if (p_ptr) p_ptr[0] = 0;
  for (int q = 0; q < (int)ln.size(); ++q) {
    const Node &nd = mg->nodes[ln[q]];
    if (nd.father_node >= 0) {
      if (p_idx) p_idx[m] = mg->node_pos[nd.father_node];
      if (p_w) p_w[m] = 1.0;
      ++m;
    } else {
      const Elem &fe = mg->elems[nd.father_elem];
      for (size_t k = 0; k < nd.wc.size(); ++k) {
        if (p_idx) p_idx[m] = mg->node_pos[fe.corners[nd.wc[k]]];
        if (p_w) p_w[m] = nd.ww[k];
        ++m;
      }
    }
    if (p_ptr) p_ptr[q + 1] = m;
  }
  return 0;
}

}  // extern "C"
