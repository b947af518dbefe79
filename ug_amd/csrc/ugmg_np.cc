// Numproc layer over the C ABI: npcreate / npinit / npexecute for the
// solve path (include/ugmg_np.h).  Option parsing follows UG's helpers:
//   ReadArgvINT / ReadArgvDOUBLE  low/scan.cc:103-160  (first character of
//                                 the argument, then "%s %d" / "%s %lf")
//   ReadArgvOption                low/scan.cc:316-333  (flag, or flag k)
//   sc_read                       np/udm/npscan.cc:1211-1280 over
//                                 ReadVecTypeDOUBLEs (npscan.cc:783-855): one
//                                 value for every entry, or values per vector
//                                 type ("n0.5:0.6|k0.7:0.8"); absent -> 2
// and each class reads the options its UG Init function reads, with the
// same defaults and the same returned state.
#include "../../include/ugmg_np.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

namespace {

constexpr double kAbsLimit = 1e-10;   // ABS_LIMIT, np/procs/ls.cc:66

using Argv = std::vector<std::string>;

// the command line split at '$' (UG's option separator), trimmed
Argv split_argv(const char *s) {
  Argv av;
  if (!s) return av;
  std::string cur;
  auto flush = [&]() {
    size_t a = cur.find_first_not_of(" \t\n"), b = cur.find_last_not_of(" \t\n");
    if (a != std::string::npos) av.push_back(cur.substr(a, b - a + 1));
    cur.clear();
  };
  for (const char *p = s; *p; ++p) {
    if (*p == '$') flush();
    else cur.push_back(*p);
  }
  flush();
  return av;
}

int read_int(const Argv &av, const char *name, int *out) {
  for (const auto &a : av) {
    if (a.empty() || a[0] != name[0]) continue;
    char opt[128];
    int v;
    if (std::sscanf(a.c_str(), "%127s %d", opt, &v) != 2) continue;
    if (std::strcmp(opt, name) == 0) {
      *out = v;
      return 0;
    }
  }
  return 1;
}

int read_option(const Argv &av, const char *name) {
  for (const auto &a : av) {
    if (a.empty() || a[0] != name[0]) continue;
    char opt[128];
    int v;
    if (std::sscanf(a.c_str(), "%127s %d", opt, &v) == 2 && std::strcmp(opt, name) == 0) return v;
    if (a == name) return 1;
  }
  return 0;
}

// "name value": the value string (the rest of the argument)
bool read_string(const Argv &av, const char *name, std::string &out) {
  const size_t ln = std::strlen(name);
  for (const auto &a : av)
    if (a.size() > ln && a.compare(0, ln, name) == 0 && (a[ln] == ' ' || a[ln] == '\t')) {
      size_t p = a.find_first_not_of(" \t", ln);
      out = p == std::string::npos ? "" : a.substr(p);
      return true;
    }
  return false;
}

// Values a VEC_SCALAR option gave per vector type (node 'n', edge 'k', the
// default type names of formats.cc:3062-3063); -1: one value for all entries
struct ScCount {
  int n[2] = {-1, -1};
  bool typed() const { return n[0] >= 0; }
};

// sc_read: 0 read, 2 absent, 3 unreadable, 4 malformed.  The option's value
// is either one number (every entry) or tokens separated by '|', each a type
// letter followed by that type's component values separated by blanks or
// ':'; x receives them type-major (the VD_OFFSET order of the norms)
int sc_read(const Argv &av, const char *name, double x[UGMG_MAXB], ScCount *cnt = nullptr) {
  std::string v;
  if (!read_string(av, name, v)) return 2;
  std::vector<double> per[2];
  bool typed = false, untyped = false;
  size_t p0 = 0;
  while (p0 <= v.size()) {
    size_t p1 = v.find('|', p0);
    if (p1 == std::string::npos) p1 = v.size();
    std::string tok = v.substr(p0, p1 - p0);
    p0 = p1 + 1;
    const size_t f = tok.find_first_not_of(" \t\n");
    const char c = f == std::string::npos ? '\0' : tok[f];
    const int t = c == 'n' ? 0 : c == 'k' ? 1 : -1;
    if (t < 0) {
      untyped = true;
      continue;
    }
    typed = true;
    if (f + 1 < tok.size() && std::isalpha((unsigned char)tok[f + 1])) return 4;   // two-letter type
    size_t q = f + 1;
    while (q < tok.size()) {
      const size_t a = tok.find_first_not_of(" \t:", q);
      if (a == std::string::npos) break;
      const size_t z = tok.find_first_of(" \t:", a);
      const std::string num = tok.substr(a, z == std::string::npos ? std::string::npos : z - a);
      double d;
      char extra;
      if (std::sscanf(num.c_str(), "%lf%c", &d, &extra) != 1) return 3;
      if (per[t].size() >= (size_t)UGMG_MAXB) return 4;
      per[t].push_back(d);
      q = z == std::string::npos ? tok.size() : z;
    }
  }
  if (typed && untyped) return 4;   // type letter missing on one of the tokens
  if (!typed) {                      // one value for every entry
    double d;
    if (std::sscanf(v.c_str(), "%lf", &d) != 1) return 3;
    const size_t z = v.find_first_of(" \t:");
    if (z != std::string::npos && v.find_first_not_of(" \t:", z) != std::string::npos) return 4;
    for (int i = 0; i < UGMG_MAXB; ++i) x[i] = d;
    if (cnt) *cnt = ScCount{};
    return 0;
  }
  if (per[0].size() + per[1].size() > (size_t)UGMG_MAXB) return 4;
  int k = 0;
  for (int t = 0; t < 2; ++t)
    for (double d : per[t]) x[k++] = d;
  if (cnt) {
    cnt->n[0] = (int)per[0].size();
    cnt->n[1] = (int)per[1].size();
  }
  return 0;
}

struct NumProc {
  std::string cls;
  int state = UGMG_NP_NOT_INIT;
  Argv av;
  // ls / bcgs
  int m = 0, restart = 0, nrhs = 1;
  double red[UGMG_MAXB], abslimit[UGMG_MAXB], weight[UGMG_MAXB];
  std::string I, x, b, A, c, r;
  std::vector<ugmg_lresult> res;
  bool pending_defect = false;
  int display = 0;   // PCR_NO_DISPLAY 0, PCR_RED_DISPLAY 1, PCR_FULL_DISPLAY 2
  ugmg_params params{};
  // lmgc
  std::string T, pre, post, base;
  int gamma = 1, nu1 = 1, nu2 = 1, bl = 0, precision = UGMG_FP64, inner = 1, graph = 1, adapt = 1;
  // smoothers
  double damp[UGMG_MAXB];
  int reorder = 1, fmode = 0;
  ScCount damp_n, abslimit_n, red_n;   // per-type counts of the typed forms

  NumProc() {
    for (int i = 0; i < UGMG_MAXB; ++i) {
      red[i] = 0.0;
      abslimit[i] = kAbsLimit;
      weight[i] = damp[i] = 1.0;
    }
  }
};

bool is_iter(const std::string &c) { return c == "lmgc" || c == "jac" || c == "gs" || c == "ex"; }
bool is_ls(const std::string &c) { return c == "ls" || c == "bcgs"; }

}  // namespace

struct ugmg_np_env {
  ugmg_ctx *ctx = nullptr;
  int level = -1;
  std::map<std::string, NumProc> np;
  std::map<std::string, std::pair<double *, int>> vec;
  // the same vectors on the coarser levels (locally refined hierarchies)
  std::map<std::string, std::map<int, double *>> lvec;
  std::string err;
  std::string out, taken;   // PCR display text: pending, and the last one handed out
};

namespace {

int fail(ugmg_np_env *e, int code, const std::string &msg) {
  e->err = msg;
  return code;
}

NumProc *find(ugmg_np_env *e, const std::string &name) {
  auto it = e->np.find(name);
  return it == e->np.end() ? nullptr : &it->second;
}

// NPIterInit (iter.cc:674-684): $A $c $r
int iter_state(const Argv &av, NumProc &o) {
  read_string(av, "A", o.A);
  read_string(av, "c", o.c);
  read_string(av, "r", o.r);
  return (o.A.empty() || o.c.empty() || o.r.empty()) ? UGMG_NP_ACTIVE : UGMG_NP_EXECUTABLE;
}

// ReadArgvDisplay (np/udm/npscan.cc:654-672): "$display no|red|full"
int read_display(const Argv &av) {
  for (const auto &a : av) {
    if (a.compare(0, 7, "display") != 0) continue;
    char v[64];
    if (std::sscanf(a.c_str(), "display %63s", v) != 1) continue;
    if (!std::strcmp(v, "no")) return 0;
    if (!std::strcmp(v, "red")) return 1;
    if (!std::strcmp(v, "full")) return 2;
  }
  return 0;
}

// NPLinearSolverInit (ls.cc:382-407)
int ls_state(ugmg_np_env *e, const Argv &av, NumProc &o) {
  o.display = read_display(av);
  read_string(av, "A", o.A);
  read_string(av, "x", o.x);
  read_string(av, "b", o.b);
  o.abslimit_n = o.red_n = ScCount{};
  if (sc_read(av, "abslimit", o.abslimit, &o.abslimit_n))
    for (int i = 0; i < UGMG_MAXB; ++i) o.abslimit[i] = kAbsLimit;
  if (sc_read(av, "red", o.red, &o.red_n)) return UGMG_NP_ACTIVE;
  if (o.x.empty() || o.b.empty() || o.A.empty() || !e->vec.count(o.x) || !e->vec.count(o.b))
    return UGMG_NP_ACTIVE;
  return UGMG_NP_EXECUTABLE;
}

// The display of one solve: PreparePCR / DoPCR (np/udm/pcr.cc:263-606) with
// UG's formats; component names from DEFAULT_NAMES (np/udm/udm.h:77).
class Pcr {
 public:
  Pcr(std::string &out, int mode, int ncomp) : out_(out), mode_(mode), nc_(std::min(ncomp, UGMG_MAXB)) {}
  // PreparePCR: the head line, the solver's name centred in 50 '*'
  // (CenterInPattern, low/misc.cc:167-190)
  void prepare(const std::string &name) {
    if (!mode_) return;
    out_ += "\n";
    out_ += center(name);
  }
  // DoPCR with PCR_CRATE (sd = false) or PCR_CRATE_SD
  void crate(const double *d, bool sd) {
    const double s = norm(d);
    if (nb_ == 0) {
      for (int j = 0; j < nc_; ++j) init_[j] = d[j];
      init_norm_ = s;
      if (mode_ == 2) {
        put(" %-3d  %c: %-12.7e   %-12.7s\n", nb_, kNames[0], d[0], "---");
        for (int i = 1; i < nc_; ++i) put("      %c: %-12.7e   %-12.7s\n", kNames[i], d[i], "---");
        if (nc_ > 1 && sd) put("   norm: %-12.7e   %-12.7s\n", s, "---");
        if (nc_ > 1) out_ += "\n";
      }
    } else if (mode_ == 2) {
      if (old_[0] != 0.0) put(" %-3d  %c: %-12.7e   %-12.7e\n", nb_, kNames[0], d[0], d[0] / old_[0]);
      else put(" %-3d  %c: %-12.7e   %-12.7s\n", nb_, kNames[0], d[0], "NaN");
      for (int i = 1; i < nc_; ++i) {
        if (old_[i] != 0.0) put("      %c: %-12.7e   %-12.7e\n", kNames[i], d[i], d[i] / old_[i]);
        else put("      %c: %-12.7e   %-12.7s\n", kNames[i], d[i], "NaN");
      }
      if (nc_ > 1 && sd) put("   norm: %-12.7e   %-12.7e\n", s, s / old_norm_);
      if (nc_ > 1) out_ += "\n";
    }
    for (int j = 0; j < nc_; ++j) old_[j] = d[j];
    old_norm_ = s;
    ++nb_;
  }
  // DoPCR with PCR_AVERAGE (sd = false) or PCR_AVERAGE_SD
  void average(const double *d, bool sd) {
    if (nb_ < 2 || !mode_) return;
    const double s = norm(d), e = 1.0 / (nb_ - 1);
    if (mode_ == 2) out_ += "\n";
    if (init_[0] != 0.0)
      put(" %-3d avg:  %c: %-12.7e   %-12.7e   %-12.7e\n", nb_ - 1, kNames[0], init_[0], d[0],
          std::pow(d[0] / init_[0], e));
    else
      put(" %-3d avg:  %c: %-12.7e   %-12.7e   %-12.7s\n", nb_ - 1, kNames[0], init_[0], d[0], "NaN");
    for (int i = 1; i < nc_; ++i) {
      if (init_[i] != 0.0)
        put("           %c: %-12.7e   %-12.7e   %-12.7e\n", kNames[i], init_[i], d[i],
            std::pow(d[i] / init_[i], e));
      else
        put("           %c: %-12.7e   %-12.7e   %-12.7s\n", kNames[i], init_[i], d[i], "NaN");
    }
    if (nc_ > 1 && sd)
      put("        norm: %-12.7e   %-12.7e   %-12.7e\n", init_norm_, s, std::pow(s / init_norm_, e));
    out_ += "\n";
  }
  // the solver's closing line ("LS  : " / "BCGS: "), times in seconds of
  // wall clock (UG's CSTART/CSTOP read the process clock)
  void times(const char *who, int level, int nit, double t) {
    if (!mode_) return;
    if (nit != 0) put("%s: L=%2d N=%2d TSOLVE=%10.4g TIT=%10.4g\n", who, level, nit, t, t / nit);
    else put("%s: L=%2d N=%2d TSOLVE=%10.4g\n", who, level, nit, t);
  }

 private:
  static constexpr const char *kNames = "uvwzpabcdefghijklmnoPQRSTUVWXYZ123456789";
  double norm(const double *d) const {
    double s = 0.0;
    for (int j = 0; j < nc_; ++j) s += d[j] * d[j];
    return std::sqrt(s);
  }
  static std::string center(const std::string &text) {
    constexpr int kWidth = 50;   // DISPLAY_WIDTH (np/np.h:98)
    const std::string t = (int)text.size() > kWidth ? std::string(" text too long ") : text;
    const int len = (int)t.size(), begin = (kWidth - len) / 2, end = begin + len;
    std::string s(kWidth, '*');
    int i = std::max(begin - 1, 0);
    s[i] = ' ';
    for (i = begin; i < end; ++i) s[i] = t[i - begin];
    if (i < kWidth) s[i] = ' ';
    return s + "\n";
  }
  void put(const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    out_ += buf;
  }
  std::string &out_;
  int mode_, nc_, nb_ = 0;
  double init_[UGMG_MAXB] = {}, old_[UGMG_MAXB] = {}, init_norm_ = 0.0, old_norm_ = 0.0;
};

void pcr_defect_cb(void *user, int, int, const double *d) { static_cast<Pcr *>(user)->crate(d, true); }

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int init_object(ugmg_np_env *e, NumProc &o) {
  const Argv &av = o.av;
  if (o.cls == "ls") {   // LinearSolverInit (ls.cc:771-797)
    if (read_int(av, "m", &o.m)) return fail(e, UGMG_NP_NOT_ACTIVE, "ls: $m missing");
    if (!read_string(av, "I", o.I) || !find(e, o.I) || !is_iter(find(e, o.I)->cls))
      return fail(e, UGMG_NP_NOT_ACTIVE, "ls: $I does not name an iteration");
    return ls_state(e, av, o);
  }
  if (o.cls == "bcgs") {   // BCGSInit (ls.cc:1750-1800)
    if (sc_read(av, "weight", o.weight))
      for (int i = 0; i < UGMG_MAXB; ++i) o.weight[i] = 1.0;
    for (int i = 0; i < UGMG_MAXB; ++i) o.weight[i] *= o.weight[i];
    if (read_int(av, "m", &o.m)) return fail(e, UGMG_NP_NOT_ACTIVE, "bcgs: $m missing");
    if (read_int(av, "R", &o.restart)) o.restart = 0;
    if (o.restart < 0) return fail(e, UGMG_NP_NOT_ACTIVE, "bcgs: $R < 0");
    for (int i = 0; i < UGMG_MAXB; ++i)
      if (o.weight[i] != 1.0) return fail(e, UGMG_NP_NOT_ACTIVE, "bcgs: only unit $weight runs on the GPU");
    o.I.clear();
    if (read_string(av, "I", o.I) && (!find(e, o.I) || find(e, o.I)->cls != "lmgc"))
      return fail(e, UGMG_NP_NOT_ACTIVE, "bcgs: $I must name an lmgc object");
    if (read_int(av, "nrhs", &o.nrhs)) o.nrhs = 1;
    if (o.nrhs < 1) return fail(e, UGMG_NP_NOT_ACTIVE, "bcgs: $nrhs < 1");
    return ls_state(e, av, o);
  }
  if (o.cls == "lmgc") {   // LmgcInit (iter.cc:7613-7656)
    read_string(av, "T", o.T);
    o.pre.clear();
    for (const auto &a : av)
      if (!a.empty() && a[0] == 'S') {
        char p1[128], p2[128], p3[128];
        if (std::sscanf(a.c_str(), "S %127s %127s %127s", p1, p2, p3) != 3) continue;
        o.pre = p1;
        o.post = p2;
        o.base = p3;
        break;
      }
    if (read_int(av, "g", &o.gamma)) o.gamma = 1;
    if (read_int(av, "n1", &o.nu1)) o.nu1 = 1;
    if (read_int(av, "n2", &o.nu2)) o.nu2 = 1;
    if (read_int(av, "b", &o.bl)) o.bl = 0;
    NumProc *t = find(e, o.T), *pr = find(e, o.pre), *po = find(e, o.post), *ba = find(e, o.base);
    if (!t || t->cls != "transfer") return fail(e, UGMG_NP_NOT_ACTIVE, "lmgc: $T does not name a transfer");
    if (!pr || !po || !is_iter(pr->cls) || !is_iter(po->cls))
      return fail(e, UGMG_NP_NOT_ACTIVE, "lmgc: $S pre/post do not name iterations");
    if (!ba || ba->cls != "ls") return fail(e, UGMG_NP_NOT_ACTIVE, "lmgc: $S base does not name an ls");
    o.damp_n = ScCount{};
    if (sc_read(av, "damp", o.damp, &o.damp_n))
      for (int i = 0; i < UGMG_MAXB; ++i) o.damp[i] = 1.0;
    std::string prec;
    o.precision = UGMG_FP64;
    if (read_string(av, "precision", prec)) {
      if (prec == "mixed") o.precision = UGMG_MIXED;
      else if (prec != "fp64") return fail(e, UGMG_NP_NOT_ACTIVE, "lmgc: $precision is fp64 or mixed");
    }
    if (read_int(av, "inner", &o.inner)) o.inner = 1;
    if (read_int(av, "adapt", &o.adapt)) o.adapt = 1;
    if (read_int(av, "graph", &o.graph)) o.graph = 1;
    return iter_state(av, o);
  }
  if (o.cls == "jac" || o.cls == "gs") {   // SmootherInit (iter.cc:763-790)
    for (int i = 0; i < UGMG_MAXB; ++i) o.damp[i] = 1.0;
    o.damp_n = ScCount{};
    if (sc_read(av, "damp", o.damp, &o.damp_n) > 2) return fail(e, UGMG_NP_NOT_ACTIVE, "jac/gs: bad $damp");
    return iter_state(av, o);
  }
  if (o.cls == "ex") {   // EXInit (iter.cc:8248-8260)
    if (read_int(av, "o", &o.reorder)) o.reorder = 1;
    o.fmode = read_option(av, "f");
    for (int i = 0; i < UGMG_MAXB; ++i) o.damp[i] = 1.0;
    o.damp_n = ScCount{};
    if (sc_read(av, "damp", o.damp, &o.damp_n) > 2) return fail(e, UGMG_NP_NOT_ACTIVE, "ex: bad $damp");
    return iter_state(av, o);
  }
  if (o.cls == "transfer") return UGMG_NP_EXECUTABLE;
  return fail(e, UGMG_E_ARG, "unknown class " + o.cls);
}

// A typed $damp as the damping per component: each type it names carries
// the block's b values (sc_read's check against the descriptor,
// npscan.cc:1258-1277); the device smoothers take one factor per component
// for every vector type, so node and edge values must agree
// (all typed ones of a chain the same count, as they describe one vector)
int resolve_damp(ugmg_np_env *e, const NumProc &o, const char *what, double out[UGMG_MAXB], int &count) {
  for (int i = 0; i < UGMG_MAXB; ++i) out[i] = o.damp[i];
  if (!o.damp_n.typed()) return 0;
  int b = e->ctx ? ugmg_block(e->ctx) : -1;
  if (b <= 0) b = count;
  const int nt = o.damp_n.n[0] > 0 && o.damp_n.n[1] > 0 ? 2 : 1;
  const int per = o.damp_n.n[0] > 0 ? o.damp_n.n[0] : o.damp_n.n[1];
  if ((nt == 2 && o.damp_n.n[0] != o.damp_n.n[1]) || (b > 0 && per != b))
    return fail(e, UGMG_E_ARG, std::string(what) + ": $damp gives a count per type other than the block's");
  for (int j = 0; j < per; ++j) {
    if (nt == 2 && o.damp[per + j] != o.damp[j])
      return fail(e, UGMG_E_ARG, std::string(what) + ": node and edge $damp differ (one factor per component)");
    out[j] = o.damp[j];
  }
  for (int j = per; j < UGMG_MAXB; ++j) out[j] = o.damp[per - 1];
  count = per;
  return 0;
}

// the ugmg_params a solver object's $I chain describes
int build_params(ugmg_np_env *e, NumProc &o, ugmg_params &p, bool &use_iter) {
  ugmg_params_default(&p);
  use_iter = !o.I.empty();
  if (!use_iter) return 0;
  NumProc *mg = find(e, o.I);
  if (!mg || mg->cls != "lmgc" || mg->state < UGMG_NP_ACTIVE)
    return fail(e, UGMG_E_STATE, "$I must be an initialised lmgc");
  NumProc *pr = find(e, mg->pre), *po = find(e, mg->post), *ba = find(e, mg->base);
  if (!pr || !po || !ba) return fail(e, UGMG_E_STATE, "lmgc parts missing");
  auto smoother_cls = [](const std::string &c) { return c == "jac" || c == "gs"; };
  if (!smoother_cls(pr->cls) || !smoother_cls(po->cls))
    return fail(e, UGMG_E_ARG, "pre- and post-smoother must be jac or gs objects");
  double dpre[UGMG_MAXB], dpost[UGMG_MAXB], dmg[UGMG_MAXB], dex[UGMG_MAXB];
  int rc, count = -1;
  if ((rc = resolve_damp(e, *pr, "pre-smoother", dpre, count)) ||
      (rc = resolve_damp(e, *po, "post-smoother", dpost, count)) ||
      (rc = resolve_damp(e, *mg, "lmgc", dmg, count)))
    return rc;
  // the base ls iterates with ex (the dense inverse) or jac (LinearSolver
  // over Smoother on the base level)
  NumProc *ex = find(e, ba->I);
  if (!ex || (ex->cls != "ex" && ex->cls != "jac"))
    return fail(e, UGMG_E_ARG, "the base ls must iterate with ex or jac");
  if ((rc = resolve_damp(e, *ex, ex->cls.c_str(), dex, count))) return rc;
  p.base_iter = ex->cls == "jac" ? UGMG_BASE_JAC : UGMG_BASE_EX;
  if (ba->abslimit_n.typed() || ba->red_n.typed())
    return fail(e, UGMG_E_ARG, "the base ls takes one $abslimit / $red for every entry");
  p.nu1 = mg->nu1;
  p.nu2 = mg->nu2;
  p.gamma = mg->gamma;
  p.baselevel = mg->bl;
  p.smoother = pr->cls == "jac" ? UGMG_SMOOTH_JAC : UGMG_SMOOTH_GS;
  // Lmgc runs $S's pre object nu1 times and its post object nu2 times
  // (np/procs/iter.cc:7812, 7922): a post object of another class or $damp is kept
  p.post_smoother = po->cls == "jac" ? UGMG_SMOOTH_JAC : UGMG_SMOOTH_GS;
  p.post_set = p.post_smoother != p.smoother;
  for (int i = 0; i < UGMG_MAXB; ++i) {
    p.post_damp[i] = dpost[i];
    if (dpost[i] != dpre[i]) p.post_set = 1;
  }
  for (int i = 0; i < UGMG_MAXB; ++i) {
    p.damp[i] = dpre[i];
    p.lmgc_damp[i] = dmg[i];
    p.base_damp[i] = dex[i];
    p.base_abslimit[i] = ba->abslimit[i];
    p.base_red[i] = ba->red[i];
  }
  p.base_maxiter = ba->m;
  p.ex_reorder = ex->reorder;
  p.ex_float = ex->fmode;
  p.precision = mg->precision;
  p.inner_cycles = mg->inner;
  p.inner_adapt = mg->adapt;
  p.use_graph = mg->graph;
  return 0;
}

// typed $abslimit / $red against the prepared hierarchy: b values for each
// of its vector types and none for the others (npscan.cc:1258-1277)
int check_limits(ugmg_np_env *e, const NumProc &o) {
  const int nc = ugmg_norm_entries(e->ctx), b = ugmg_block(e->ctx);
  if (nc <= 0 || b <= 0) return 0;
  const int nt = nc / b;
  for (const ScCount *c : {&o.abslimit_n, &o.red_n})
    if (c->typed())
      for (int t = 0; t < 2; ++t)
        if (c->n[t] != (t < nt ? b : 0))
          return fail(e, UGMG_E_ARG, "number of $abslimit / $red values per type does not match the vectors");
  return 0;
}

int level_n(ugmg_np_env *e, int level, int *n) {
  int nn = 0;
  int rc = ugmg_level_size(e->ctx, level, &nn, nullptr);
  if (rc) return fail(e, rc, "level not set");
  *n = nn;
  return 0;
}

}  // namespace

extern "C" {

int ugmg_np_env_create(ugmg_ctx *ctx, ugmg_np_env **out) {
  if (!out) return UGMG_E_ARG;
  *out = new (std::nothrow) ugmg_np_env;
  if (!*out) return UGMG_E_NOMEM;
  (*out)->ctx = ctx;
  return UGMG_OK;
}

int ugmg_np_env_destroy(ugmg_np_env *env) {
  delete env;
  return UGMG_OK;
}

const char *ugmg_np_last_error(const ugmg_np_env *env) { return env ? env->err.c_str() : "null environment"; }

int ugmg_np_set_level(ugmg_np_env *env, int level) {
  if (!env || level < 0) return UGMG_E_ARG;
  env->level = level;
  return UGMG_OK;
}

int ugmg_np_set_vector(ugmg_np_env *env, const char *name, double *data, int ncols) {
  if (!env || !name || !data || ncols < 1) return UGMG_E_ARG;
  env->vec[name] = {data, ncols};
  return UGMG_OK;
}

int ugmg_np_set_level_vector(ugmg_np_env *env, const char *name, int level, double *data) {
  if (!env || !name || !data || level < 0) return UGMG_E_ARG;
  env->lvec[name][level] = data;
  return UGMG_OK;
}

int ugmg_np_create(ugmg_np_env *env, const char *name, const char *cls) {
  if (!env || !name || !cls) return UGMG_E_ARG;
  const std::string c = cls;
  if (!is_iter(c) && !is_ls(c) && c != "transfer") return fail(env, UGMG_E_ARG, "unknown class " + c);
  if (env->np.count(name)) return fail(env, UGMG_E_ARG, std::string("numproc ") + name + " exists");
  NumProc o;
  o.cls = c;
  env->np[name] = o;
  return UGMG_OK;
}

int ugmg_np_init(ugmg_np_env *env, const char *name, const char *options) {
  if (!env || !name) return UGMG_E_ARG;
  NumProc *o = find(env, name);
  if (!o) return fail(env, UGMG_E_ARG, std::string("no numproc ") + name);
  o->av = split_argv(options);
  env->err.clear();
  const int st = init_object(env, *o);
  if (st < 0) return st;
  o->state = st;
  return st;
}

int ugmg_np_execute(ugmg_np_env *env, const char *name, const char *flags) {
  if (!env || !name) return UGMG_E_ARG;
  NumProc *o = find(env, name);
  if (!o) return fail(env, UGMG_E_ARG, std::string("no numproc ") + name);
  if (!env->ctx) return fail(env, UGMG_E_STATE, "no context (environment created without a GPU)");
  const Argv av = split_argv(flags);
  int level = env->level;
  if (level < 0) {
    level = 0;
    while (ugmg_level_size(env->ctx, level + 1, nullptr, nullptr) == UGMG_OK) ++level;
  }
  int n = 0, rc;
  if ((rc = level_n(env, level, &n))) return rc;
  if (is_iter(o->cls)) {   // NPIterExecute: $i PreProcess, $s Iter (c, r), $p PostProcess
    if (o->cls != "lmgc") return fail(env, UGMG_E_ARG, "only lmgc objects run on their own");
    NumProc fake;
    fake.I = name;
    ugmg_params p;
    bool use_iter;
    if (read_option(av, "i")) {
      if ((rc = build_params(env, fake, p, use_iter))) return rc;
      if ((rc = ugmg_set_params(env->ctx, &p)) || (rc = ugmg_preprocess(env->ctx, level)))
        return fail(env, rc, ugmg_last_error(env->ctx));
    }
    if (read_option(av, "s")) {
      if (o->state != UGMG_NP_EXECUTABLE || !env->vec.count(o->c) || !env->vec.count(o->r))
        return fail(env, UGMG_E_STATE, "lmgc: $c and $r must name registered vectors");
      if ((rc = ugmg_iter_host(env->ctx, level, env->vec[o->c].first, env->vec[o->r].first)))
        return fail(env, rc, ugmg_last_error(env->ctx));
    }
    if (read_option(av, "p")) ugmg_postprocess(env->ctx);
    return UGMG_OK;
  }
  if (!is_ls(o->cls)) return fail(env, UGMG_E_ARG, "transfer objects are not executed");
  if (o->state != UGMG_NP_EXECUTABLE) return fail(env, UGMG_E_STATE, std::string(name) + " is not executable");
  // NPLinearSolverExecute (ls.cc:436-535): $i, $d, $r, $s, $p in this order
  double *x = env->vec[o->x].first, *b = env->vec[o->b].first;
  {
    ugmg_params p;
    bool ui;
    if (read_option(av, "i")) {
      if ((rc = build_params(env, *o, p, ui))) return rc;
      o->params = p;
      if (o->cls == "bcgs" && (rc = ugmg_set_batch(env->ctx, o->nrhs)))
        return fail(env, rc, ugmg_last_error(env->ctx));
      if ((rc = ugmg_set_params(env->ctx, &p)) || (rc = ugmg_preprocess(env->ctx, level)))
        return fail(env, rc, ugmg_last_error(env->ctx));
    }
  }
  if ((rc = check_limits(env, *o))) return rc;
  const int ncols = o->cls == "bcgs" ? o->nrhs : 1;
  o->res.assign(ncols, ugmg_lresult{});
  // a locally refined hierarchy (FULLREFINELEVEL below the level): LinearSolver
  // runs ON_SURFACE over x and b of every level from the base level up
  const int frl = ugmg_fullrefinelevel(env->ctx);
  const bool surface = frl >= 0 && frl < level;
  std::vector<double *> XL(level + 1, nullptr), BL(level + 1, nullptr);
  if (surface) {
    if (o->cls != "ls") return fail(env, UGMG_E_ARG, "bcgs runs on uniformly refined hierarchies only");
    const int bl = ugmg_base_level(env->ctx);
    if (bl < 0) return fail(env, UGMG_E_STATE, "the hierarchy is not preprocessed ($i)");
    for (int l = bl; l < level; ++l) {
      auto fx = env->lvec[o->x].find(l), fb = env->lvec[o->b].find(l);
      if (fx == env->lvec[o->x].end() || fb == env->lvec[o->b].end())
        return fail(env, UGMG_E_STATE, "locally refined hierarchy: $x and $b need their coarser levels "
                                       "(ugmg_np_set_level_vector) from the base level up");
      XL[l] = fx->second;
      BL[l] = fb->second;
    }
    XL[level] = x;
    BL[level] = b;
  }
  auto ls_solve = [&](int maxit, ugmg_lresult *r, ugmg_defect_cb cb, void *user) {
    const int dd = o->pending_defect ? 1 : 0;
    return surface ? ugmg_solve_surface_host(env->ctx, level, XL.data(), BL.data(), o->abslimit, o->red,
                                             maxit, dd, r, cb, user)
                   : ugmg_solve_host(env->ctx, level, x, b, o->abslimit, o->red, maxit, dd, r, cb, user);
  };
  if (read_option(av, "d")) o->pending_defect = true;   // formed by the next Residuum/Solver
  const bool want_r = read_option(av, "r") != 0, want_s = read_option(av, "s") != 0;
  if (want_r && !want_s) {
    // LinearResiduum alone: LinearDefect (if pending) and the surface norm,
    // through a zero-iteration solve
    ugmg_lresult r;
    if ((rc = ls_solve(0, &r, nullptr, nullptr))) return fail(env, rc, ugmg_last_error(env->ctx));
    o->pending_defect = false;
    o->res[0] = r;
  }
  if (want_s) {
    // VD_NCOMP: the entries of the norms (components, times the vector
    // types of a node + edge hierarchy)
    const int ne = ugmg_norm_entries(env->ctx);
    const int nb = ne > 0 ? ne : std::max(1, ugmg_block(env->ctx));
    const auto t0 = std::chrono::steady_clock::now();
    if (o->cls == "ls") {
      // LinearSolver's display (ls.cc:666-744): DoPCR(PCR_CRATE_SD) on the
      // first and every iteration's defect, then PCR_AVERAGE and "LS  :"
      Pcr pcr(env->out, o->display, nb);
      pcr.prepare(name);
      if ((rc = ls_solve(o->m, &o->res[0], o->display ? pcr_defect_cb : nullptr, &pcr)))
        return fail(env, rc, ugmg_last_error(env->ctx));
      pcr.average(o->res[0].last_defect, false);
      pcr.times("LS  ", level, o->res[0].number_of_linear_iterations, seconds_since(t0));
    } else {
      // BCGSSolver's display (ls.cc:1896-2060): DoPCR(PCR_CRATE) on the first
      // defect and every LinearResiduum, then PCR_AVERAGE and "BCGS:", per
      // right-hand side (scalar hierarchies)
      const int md = 2 * std::max(o->m, 0) + 1;
      std::vector<double> dfs(o->display ? (size_t)ncols * md : 0);
      if ((rc = ugmg_bcgs_solve_host(env->ctx, level, ncols, x, b, n, o->abslimit, o->red, o->m,
                                     o->restart, o->I.empty() ? 0 : 1, o->pending_defect ? 1 : 0,
                                     o->res.data(), o->display ? dfs.data() : nullptr, md)))
        return fail(env, rc, ugmg_last_error(env->ctx));
      const double t = seconds_since(t0);
      for (int j = 0; j < ncols && o->display; ++j) {
        Pcr pcr(env->out, o->display, 1);
        pcr.prepare(name);
        const ugmg_lresult &r = o->res[j];
        const int nd = std::min(md, 1 + r.number_of_linear_iterations);
        for (int k = 0; k < nd; ++k) pcr.crate(&dfs[(size_t)j * md + k], false);
        pcr.average(r.last_defect, false);
        pcr.times("BCGS", level, r.number_of_linear_iterations, t / ncols);
      }
    }
    o->pending_defect = false;
  }
  if (read_option(av, "p")) ugmg_postprocess(env->ctx);
  return UGMG_OK;
}

const char *ugmg_np_output(ugmg_np_env *env) {
  if (!env) return "";
  env->taken.swap(env->out);
  env->out.clear();
  return env->taken.c_str();
}

int ugmg_np_result(const ugmg_np_env *env, const char *name, int col, ugmg_lresult *res) {
  if (!env || !name || !res) return UGMG_E_ARG;
  auto it = env->np.find(name);
  if (it == env->np.end() || col < 0 || col >= (int)it->second.res.size()) return UGMG_E_ARG;
  *res = it->second.res[col];
  return UGMG_OK;
}

int ugmg_np_params(const ugmg_np_env *env, const char *name, ugmg_params *p) {
  if (!env || !name || !p) return UGMG_E_ARG;
  auto it = env->np.find(name);
  if (it == env->np.end()) return UGMG_E_ARG;
  bool ui;
  NumProc o = it->second;
  return build_params(const_cast<ugmg_np_env *>(env), o, *p, ui);
}

}  // extern "C"
