// Deck reading and table construction for the walker engine (host, FP64).
//
// Input semantics follow the reference's readers:
//   tokens, comments, INCLUDE       src/program_utils/qmc_io.cpp:92-352
//   SYSTEM { MOLECULE }             src/system/Molecular_system.cpp:53-137, Particle_set.cpp:53-89
//   PSEUDO (RGAUSSIAN OLDQMC)       src/system/Pseudopotential.cpp:874-1033,
//                                   src/basis_function/Rgaussian_function.cpp:24-112
//   AOSPLINE GAMESS                 src/basis_function/Cubic_spline_build.cpp:140-655
//   spline fit / cutoff / pad       src/basis_function/Spline_fitter.cpp:153-289
//   CUTOFF_MO + ORBFILE             src/orbitals/MO_matrix_cutoff.h:141-247, MO_matrix.h:173-251
//   SLATER STATES / DETWT           src/wavefunction/Slat_wf_data.cpp:34-303
//   JASTROW2 groups                 src/wavefunction/Jastrow2_wf.cpp:29-160,
//                                   Jastrow2_one.cpp:24-100, Jastrow2_two.cpp:20-240
//   CUTOFF_CUSP / POLYPADE          src/basis_function/Cutoff_cusp.cpp:27-60,
//                                   Poly_pade_function.cpp:27-90
//   12/6/4-point quadratures        src/system/gesqua.cpp:32-120
//   METHOD defaults                 src/method/Vmc_method.cpp:33-132, Dmc_method.cpp:33-160
#include "qw_setup.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>

namespace qw {
namespace {

using Words = std::vector<std::string>;
const double kPi = 3.14159265358979323846;

struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] void bad(const std::string& m) { throw std::runtime_error(m); }
[[noreturn]] void unsupported(const std::string& m) { throw Unsupported(m); }

bool same(const std::string& a, const char* b) {
  if (a.size() != strlen(b)) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (std::toupper((unsigned char)a[i]) != std::toupper((unsigned char)b[i])) return false;
  return true;
}

std::string dir_of(const std::string& p) {
  auto k = p.rfind('/');
  return k == std::string::npos ? "." : p.substr(0, k);
}
std::string resolve(const std::string& dir, const std::string& f) { return f.size() && f[0] == '/' ? f : dir + "/" + f; }

void lex(const std::string& path, Words& out, int depth = 0) {
  if (depth > 16) bad("INCLUDE nesting too deep at " + path);
  std::ifstream in(path);
  if (!in) bad("cannot open " + path);
  Words w;
  std::string line;
  while (std::getline(in, line)) {
    auto hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    size_t i = 0;
    while (i < line.size()) {
      char ch = line[i];
      if (std::isspace((unsigned char)ch)) {
        i++;
      } else if (ch == '{' || ch == '}') {
        w.emplace_back(1, ch);
        i++;
      } else {
        size_t j = i;
        while (j < line.size() && !std::isspace((unsigned char)line[j]) && line[j] != '{' && line[j] != '}') j++;
        w.push_back(line.substr(i, j - i));
        i = j;
      }
    }
  }
  for (size_t i = 0; i < w.size(); i++) {
    if (same(w[i], "INCLUDE")) {
      if (i + 1 >= w.size()) bad("INCLUDE without a file name in " + path);
      lex(resolve(dir_of(path), w[i + 1]), out, depth + 1);
      i++;
    } else {
      out.push_back(w[i]);
    }
  }
}

// Scans words at brace depth 0 from `pos`; returns the index of `name` or -1.
long find0(const Words& w, size_t pos, const char* name) {
  int depth = 0;
  for (size_t i = pos; i < w.size(); i++) {
    if (w[i] == "{") depth++;
    else if (w[i] == "}") depth--;
    if (depth < 0) return -1;
    if (depth == 0 && same(w[i], name)) return (long)i;
  }
  return -1;
}
bool section(const Words& w, size_t& pos, Words& out, const char* name) {
  out.clear();
  long k = find0(w, pos, name);
  if (k < 0) return false;
  size_t i = (size_t)k + 1;
  if (i >= w.size() || w[i] != "{") bad(std::string(name) + " must be followed by {");
  int depth = 0;
  size_t b = ++i;
  for (; i < w.size(); i++) {
    if (w[i] == "{") depth++;
    if (w[i] == "}" && --depth < 0) break;
  }
  if (i >= w.size()) bad(std::string("unterminated section ") + name);
  out.assign(w.begin() + b, w.begin() + i);
  pos = i + 1;
  return true;
}
bool section0(const Words& w, Words& out, const char* name) {
  size_t p = 0;
  return section(w, p, out, name);
}
bool value(const Words& w, std::string& v, const char* name) {
  long k = find0(w, 0, name);
  if (k < 0) return false;
  if ((size_t)k + 1 >= w.size()) bad(std::string("value expected after ") + name);
  v = w[k + 1];
  return true;
}
bool dvalue(const Words& w, double& v, const char* name) {
  std::string s;
  if (!value(w, s, name)) return false;
  v = std::atof(s.c_str());
  return true;
}
bool ivalue(const Words& w, int& v, const char* name) {
  std::string s;
  if (!value(w, s, name)) return false;
  v = std::atoi(s.c_str());
  return true;
}
bool keyword(const Words& w, const char* name) { return find0(w, 0, name) >= 0; }

// ------------------------------------------------------------------ AOSPLINE
struct SplineFit {
  double spacing, threshold;
  std::vector<double> coef;  // [n][4]
  size_t n() const { return coef.size() / 4; }
  double eval(double r) const {
    size_t i = (size_t)(r * (1.0 / spacing));  // same interval rule as the device lookup
    double h = r - i * spacing;
    const double* c = &coef[4 * i];
    return c[0] + h * (c[1] + h * (c[2] + h * c[3]));
  }
};

// Clamped cubic spline through (x_j, y_j) on a uniform grid (tridiagonal solve for the
// second derivatives, then per-interval polynomial coefficients).
SplineFit fit_spline(const std::vector<double>& x, const std::vector<double>& y, double d0, double dn) {
  size_t n = x.size();
  std::vector<double> m(n), u(n);
  m[0] = -0.5;
  u[0] = (3. / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - d0);
  for (size_t i = 1; i + 1 < n; i++) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * m[i - 1] + 2.;
    m[i] = (sig - 1.) / p;
    double dd = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6. * dd / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  double qn = .5;
  double un = (3. / (x[n - 1] - x[n - 2])) * (dn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
  m[n - 1] = (un - qn * u[n - 2]) / (qn * m[n - 2] + 1.);
  for (size_t k = n - 1; k-- > 0;) m[k] = m[k] * m[k + 1] + u[k];
  SplineFit s;
  s.spacing = x[1] - x[0];
  s.threshold = x[n - 1];
  s.coef.assign(4 * n, 0.0);
  for (size_t i = 0; i + 1 < n; i++) {
    double h = x[i + 1] - x[i];
    s.coef[4 * i] = y[i];
    s.coef[4 * i + 1] = (y[i + 1] - y[i]) / h - h * (2. * m[i] + m[i + 1]) / 6.;
    s.coef[4 * i + 2] = m[i] / 2.;
    s.coef[4 * i + 3] = (m[i + 1] - m[i]) / (6. * h);
  }
  s.coef[4 * (n - 1)] = y[n - 1];
  return s;
}

// Largest radius, on a 0.5 bohr scan from the end of the table, where |f| exceeds
// 1e-10 * min(max|f_grid|, 1), plus one scan step.
double spline_cutoff(const SplineFit& s) {
  double mx = 0;
  for (size_t i = 0; i < s.n(); i++) mx = std::max(mx, std::fabs(s.coef[4 * i]));
  double thr = 1e-10 * std::min(mx, 1.0);
  for (double r = s.threshold - 0.5; r >= 0; r -= 0.5)
    if (std::fabs(s.eval(r)) > thr) return r + 0.5;
  return 0;
}

int sym_of(const std::string& s) {
  if (same(s, "S")) return SK_S;
  if (same(s, "P")) return SK_P;
  if (same(s, "5D")) return SK_5D;
  if (same(s, "D") || same(s, "6D")) return SK_6D;
  if (same(s, "F") || same(s, "10F")) return SK_10F;
  if (same(s, "7F")) return SK_7F;
  if (same(s, "7F_crystal")) return SK_7FC;
  if (same(s, "G") || same(s, "15G")) return SK_15G;
  if (same(s, "9G")) return SK_9G;
  if (same(s, "9G_pyscf")) return SK_9GP;
  unsupported("AOSPLINE symmetry " + s + " is not supported");
}
int sym_nf(int s) {
  static const int nf[] = {1, 3, 5, 6, 10, 7, 7, 9, 9, 15};
  return nf[s];
}
int sym_l(int s) {
  static const int l[] = {0, 1, 2, 2, 3, 3, 3, 4, 4, 4};
  return l[s];
}

struct BasisObj {
  std::string label;
  std::vector<int> sym;
  std::vector<SplineFit> spl;
  std::vector<double> cut;  // per spline
  double thr = 0;
  int nfunc = 0;
};

BasisObj read_aospline(const Words& w) {
  BasisObj b;
  b.label = w[0];
  double dx = 0.02;
  int npts = 2500;
  std::string norm = "GAMESSNORM";
  value(w, norm, "NORMTYPE");
  bool renorm = !keyword(w, "NORENORMALIZE");
  double reqcut = -1;
  dvalue(w, reqcut, "CUTOFF");
  if (keyword(w, "CUSP") || keyword(w, "CUSP_MATCHING")) unsupported("AOSPLINE CUSP is not supported");
  bool zero_d = keyword(w, "ZERO_DERIVATIVE");
  double sp;
  if (dvalue(w, sp, "SPACING")) {
    dx = sp;
    npts = (int)(50.0 / sp);
  }
  Words g;
  if (!section0(w, g, "GAMESS")) unsupported("AOSPLINE without a GAMESS section is not supported");
  std::vector<double> x(npts), y(npts);
  for (int j = 0; j < npts; j++) x[j] = j * dx;
  for (size_t p = 0; p < g.size();) {
    int s = sym_of(g[p]);
    if (p + 1 >= g.size()) bad("GAMESS section ends early");
    int ng = std::atoi(g[p + 1].c_str());
    p += 2;
    if (p + 3 * (size_t)ng > g.size()) bad("GAMESS section count is wrong");
    std::fill(y.begin(), y.end(), 0.0);
    for (int k = 0; k < ng; k++, p += 3) {
      double a = std::atof(g[p + 1].c_str()), c = std::atof(g[p + 2].c_str()), nrm;
      if (norm == "GAMESSNORM") {
        // sqrt(2 (4a)^(l+1) sqrt(2a/pi) / (2l+1)!!)
        double f4 = 4. * a, fac = std::sqrt(2. * a / kPi);
        int l = sym_l(s);
        double pw = f4, dfact = 1;
        for (int q = 0; q < l; q++) pw *= f4, dfact *= (2 * q + 3);
        if (l == 0) nrm = std::sqrt(2. * f4 * fac);
        else if (l == 1) nrm = std::sqrt(2. * f4 * f4 * fac / 3.);
        else if (l == 2) nrm = std::sqrt(2. * f4 * f4 * f4 * fac / 15.);
        else if (l == 3) nrm = std::sqrt(2. * (f4 * f4) * (f4 * f4) * fac / 105.);
        else nrm = std::sqrt(2. * (f4 * f4) * (f4 * f4) * f4 * fac / 945.);
        (void)pw;
        (void)dfact;
      } else if (norm == "CRYSTAL") {
        int l = sym_l(s);
        nrm = l == 0 ? 1.0 : 1.0 / std::sqrt(2.0 * l + 1.0);
      } else if (norm == "NONE") {
        nrm = 1;
      } else {
        bad("unknown NORMTYPE " + norm);
      }
      for (int j = 0; j < npts; j++) y[j] += nrm * c * std::exp(-a * x[j] * x[j]);
    }
    if (renorm) {
      int pexp = 2 * (sym_l(s) + 1);
      double sum = 0;
      for (int j = 0; j < npts; j++) sum += dx * std::pow(x[j], pexp) * y[j] * y[j];
      if (std::fabs(sum - 1) > 1e-5) {
        double r = std::sqrt(sum);
        for (double& v : y) v = v / r;
      }
    }
    if (reqcut > 0) {
      const double smooth = 1.2, hi = reqcut - 1e-6, lo = hi - smooth;
      for (int j = 0; j < npts; j++)
        if (x[j] > lo) {
          if (x[j] > hi) y[j] = 0;
          else {
            double z = (x[j] - lo) / smooth;
            y[j] *= (1 - z * z * z * (6 * z * z - 15 * z + 10));
          }
        }
    }
    double d0 = zero_d ? 0.0 : (y[1] - y[0]) / dx;
    double dn = (y[npts - 1] - y[npts - 2]) / dx;
    b.sym.push_back(s);
    b.spl.push_back(fit_spline(x, y, d0, dn));
    b.nfunc += sym_nf(s);
  }
  for (auto& s : b.spl) {
    double c = spline_cutoff(s);
    b.cut.push_back(c);
    b.thr = std::max(b.thr, c);
  }
  // pad every spline with zero intervals out to the object cutoff
  for (auto& s : b.spl)
    if (s.threshold < b.thr) {
      int extra = (int)((b.thr - s.threshold) / s.spacing) + 1;
      s.coef.resize(s.coef.size() + 4 * extra, 0.0);
      s.threshold = b.thr;
    }
  return b;
}

// ------------------------------------------------------------------ Jastrow objects
JObj read_jobj(const Words& w) {
  JObj o{};
  if (w.size() < 2) bad("short Jastrow basis section");
  if (same(w[1], "CUTOFF_CUSP")) {
    o.kind = 0;
    o.nf = 1;
    if (!dvalue(w, o.cusp, "CUSP") || !dvalue(w, o.gamma, "GAMMA") || !dvalue(w, o.rcut, "CUTOFF"))
      bad("CUTOFF_CUSP needs CUSP, GAMMA and CUTOFF");
    o.rcutinv = 1.0 / o.rcut;
    o.pade0 = 1. / (3. + o.gamma);
  } else if (same(w[1], "POLYPADE")) {
    o.kind = 1;
    double b0;
    if (!dvalue(w, o.rcut, "RCUT") || !dvalue(w, b0, "BETA0") || !ivalue(w, o.nf, "NFUNC"))
      bad("POLYPADE needs RCUT, BETA0 and NFUNC");
    if (o.nf < 1 || o.nf > kMaxJF) unsupported("POLYPADE NFUNC above 8");
    o.rcutinv = 1.0 / o.rcut;
    o.beta[0] = b0;
    double b1 = std::log(b0 + 1.00001);
    for (int i = 1; i < o.nf; i++) o.beta[i] = std::exp(b1 + 1.6 * i) - 1;
  } else {
    unsupported("Jastrow basis " + w[1] + " is not supported");
  }
  return o;
}
bool same_shape(const JObj& a, const JObj& b) {
  if (a.kind != b.kind || a.nf != b.nf || a.rcut != b.rcut) return false;
  if (a.kind == 0) return a.cusp == b.cusp && a.gamma == b.gamma;
  for (int i = 0; i < a.nf; i++)
    if (a.beta[i] != b.beta[i]) return false;
  return true;
}
void add_merged(std::vector<JObj>& list, const JObj& o) {
  for (auto& x : list)
    if (same_shape(x, o)) {
      for (int i = 0; i < o.nf; i++) x.cl[i] += o.cl[i], x.cu[i] += o.cu[i];
      return;
    }
  list.push_back(o);
}

void gesqua(int& nq, PPType& p) {
  if (nq < 4) bad("AIP must be >= 4");
  for (int i = 0; i < kMaxAIP; i++) p.q[i][0] = p.q[i][1] = p.q[i][2] = p.w[i] = 0;
  if (nq < 6) {
    nq = 4;
    double q = 1. / std::sqrt(3.);
    double s[4][3] = {{q, q, q}, {q, -q, -q}, {-q, q, -q}, {-q, -q, q}};
    for (int i = 0; i < 4; i++) {
      for (int d = 0; d < 3; d++) p.q[i][d] = s[i][d];
      p.w[i] = 0.25;
    }
  } else if (nq < 12) {
    nq = 6;
    for (int i = 0; i < 6; i++) {
      p.q[i][i / 2] = (i % 2) ? -1 : 1;
      p.w[i] = 1. / 6.;
    }
  } else if (nq < 18) {
    nq = 12;
    double fi0 = kPi / 5., ct = 1. / std::sqrt(5.0), st = 2. / std::sqrt(5.0);
    p.q[0][2] = 1;
    p.q[1][2] = -1;
    for (int i = 0; i < 5; i++) {
      double k2 = 2 * i;
      p.q[i + 2][0] = st * std::cos(k2 * fi0);
      p.q[i + 2][1] = st * std::sin(k2 * fi0);
      p.q[i + 2][2] = ct;
      p.q[i + 7][0] = st * std::cos((k2 + 1) * fi0);
      p.q[i + 7][1] = st * std::sin((k2 + 1) * fi0);
      p.q[i + 7][2] = -ct;
    }
    for (int i = 0; i < 12; i++) p.w[i] = 1. / 12.;
  } else {
    unsupported("AIP above 12 is not supported");
  }
}

double pp_channel(const PPType& p, int l, double r) {
  double v = 0;
  for (int j = 0; j < p.nterm[l]; j++) v += p.c[l][j] * std::pow(r, p.n[l][j]) * std::exp(-std::min(p.a[l][j] * r * r, 60.0));
  return v;
}

}  // namespace

// ------------------------------------------------------------------ public
SysTables read_deck(const std::string& path, bool* unsup) {
  if (unsup) *unsup = false;
  try {
    Words all;
    lex(path, all);
    SysTables T;
    std::string dir = dir_of(path);
    Words sys, tmp;
    if (!section0(all, sys, "SYSTEM")) bad("missing SYSTEM section");
    if (!keyword(sys, "MOLECULE")) unsupported("only SYSTEM { MOLECULE } is supported");
    if (section0(sys, tmp, "ELECTRIC_FIELD")) unsupported("ELECTRIC_FIELD is not supported");
    if (section0(sys, tmp, "BOUNDING_BOX")) unsupported("BOUNDING_BOX is not supported");
    Words ns;
    if (!section0(sys, ns, "NSPIN") || ns.size() < 2) bad("need NSPIN { nup ndown }");
    T.nup = std::atoi(ns[0].c_str());
    T.ndn = std::atoi(ns[1].c_str());
    std::vector<std::string> labels;
    size_t pos = 0;
    Words at;
    while (section(sys, pos, at, "ATOM")) {
      if (at.size() < 2) bad("short ATOM section");
      labels.push_back(at[0]);
      T.atom_Z.push_back(std::atof(at[1].c_str()));
      long k = find0(at, 0, "COOR");
      if (k < 0 || (size_t)k + 3 >= at.size()) bad("ATOM needs COOR x y z");
      for (int d = 0; d < 3; d++) T.atom_pos.push_back(std::atof(at[k + 1 + d].c_str()));
    }
    T.na = (int)labels.size();
    if (!T.na) bad("no ATOM sections");
    for (int i = 0; i < T.na; i++)
      for (int j = 0; j < i; j++) {
        double r2 = 0;
        for (int d = 0; d < 3; d++) {
          double x = T.atom_pos[3 * i + d] - T.atom_pos[3 * j + d];
          r2 += x * x;
        }
        T.vii += T.atom_Z[i] * T.atom_Z[j] / std::sqrt(r2);
      }

    // ---- pseudopotentials
    T.atom_pp.assign(T.na, -1);
    pos = 0;
    Words ps;
    while (section(all, pos, ps, "PSEUDO")) {
      PPType p{};
      if (keyword(ps, "SPIN_DEP")) unsupported("SPIN_DEP pseudopotentials are not supported");
      Words b, oq;
      if (!section0(ps, b, "BASIS")) bad("PSEUDO needs BASIS");
      if (b.size() < 2 || !same(b[1], "RGAUSSIAN")) unsupported("pseudopotential radial form must be RGAUSSIAN");
      if (!section0(b, oq, "OLDQMC")) unsupported("RGAUSSIAN needs OLDQMC");
      size_t it = 1;
      p.L = std::atoi(oq.at(it++).c_str());
      if (p.L < 1 || p.L > kMaxPPL) unsupported("pseudopotential channel count out of range");
      for (int l = 0; l < p.L; l++) {
        p.nterm[l] = std::atoi(oq.at(it++).c_str());
        if (p.nterm[l] > kMaxPPT) unsupported("too many Gaussian terms in a pseudopotential channel");
      }
      for (int l = 0; l < p.L; l++)
        for (int j = 0; j < p.nterm[l]; j++) {
          p.n[l][j] = std::atof(oq.at(it++).c_str()) - 2;
          p.a[l][j] = std::atof(oq.at(it++).c_str());
          p.c[l][j] = std::atof(oq.at(it++).c_str());
        }
      p.aip = 6;
      ivalue(ps, p.aip, "AIP");
      p.addzeff = keyword(ps, "ADD_ZEFF");
      gesqua(p.aip, p);
      int id = (int)T.pptypes.size();
      bool used = false;
      for (int a = 0; a < T.na; a++)
        if (labels[a] == ps[0]) {
          if (T.atom_pp[a] >= 0) bad("two PSEUDO sections for " + ps[0]);
          T.atom_pp[a] = id;
          used = true;
          // static cutoff: scan inward from 20 bohr in 0.05 steps
          double cut = 0;
          for (int l = 0; l < p.L - 1; l++) {
            double cl = 20.0;
            for (double r = 20.0; r > 0; r -= .05)
              if (std::fabs(pp_channel(p, l, r)) > 1e-5) {
                cl = r - .05;
                break;
              }
            cut = std::max(cut, cl);
          }
          p.cutoff = cut;
        }
      if (used) T.pptypes.push_back(p);
    }

    // ---- trial function
    Words tf, wf1, wf2;
    if (!section0(all, tf, "TRIALFUNC")) bad("missing TRIALFUNC");
    if (keyword(tf, "SLATER-JASTROW")) {
      if (!section0(tf, wf1, "WF1") || !section0(tf, wf2, "WF2")) bad("SLATER-JASTROW needs WF1 and WF2");
    } else if (keyword(tf, "SLATER")) {
      wf1 = tf;
    } else {
      unsupported("TRIALFUNC must be SLATER-JASTROW or SLATER");
    }
    if (!keyword(wf1, "SLATER")) unsupported("WF1 must be a SLATER determinant");
    Words dw, st, orb;
    if (keyword(wf1, "OPTIMIZE_MO")) unsupported("OPTIMIZE_MO is not supported");
    // determinant weights: DETWT, or CSF sections { c w_1 w_2 ... } giving c * w_j to the next
    // determinants in STATES order (src/wavefunction/Slat_wf_data.cpp:69-118)
    T.detwt.clear();
    T.csf_c.clear();
    T.csf_w.clear();
    T.csf_d0.assign(1, 0);
    {
      size_t cp = 0;
      Words cs;
      int ncsf = 0;
      while (section(wf1, cp, cs, "CSF")) {
        if (cs.size() < 2) bad("a CSF needs a coefficient and at least one determinant weight");
        const double c = std::atof(cs[0].c_str());
        T.csf_c.push_back(c);
        for (size_t j = 1; j < cs.size(); j++) {
          T.csf_w.push_back(std::atof(cs[j].c_str()));
          T.detwt.push_back(c * T.csf_w.back());
        }
        T.csf_d0.push_back((int)T.detwt.size());
        ncsf++;
      }
      if (ncsf) {
        if (section0(wf1, dw, "DETWT")) bad("CSF is incompatible with DETWT");
      } else {
        if (!section0(wf1, dw, "DETWT") || dw.empty()) bad("need DETWT");
        for (auto& x : dw) {
          T.detwt.push_back(std::atof(x.c_str()));
          T.csf_c.push_back(T.detwt.back());
          T.csf_w.push_back(1.0);
          T.csf_d0.push_back((int)T.detwt.size());
        }
      }
    }
    T.optimize_det = keyword(wf1, "OPTIMIZE_DET");
    T.ndet = (int)T.detwt.size();
    if (T.ndet > kMaxDet) unsupported("at most " + std::to_string(kMaxDet) + " determinants are supported");
    if (T.detwt[0] == 0.0) bad("the first determinant must have a nonzero weight");
    if (!section0(wf1, st, "STATES")) bad("need STATES");
    if ((int)st.size() != T.ndet * T.ne()) bad("STATES must list one orbital per electron for every determinant");
    if (!section0(wf1, orb, "ORBITALS")) unsupported("only ORBITALS (real MOs) are supported");
    // CUTOFF_MO, or BLAS_MO: the same linear combination, which the reference evaluates densely
    // and without the functions whose value is below 1e-8 at the point
    // (src/orbitals/MO_matrix_blas.h:336-491); here both are evaluated exactly
    if (!keyword(orb, "CUTOFF_MO") && !keyword(orb, "BLAS_MO"))
      unsupported("only CUTOFF_MO and BLAS_MO orbitals are supported");
    // basis centers: the atoms (USEATOMS, or USEGLOBAL of a molecule), or the centers of a file
    // (READ: a count, then a label and x y z per center; src/orbitals/Center_set.cpp:51-90,
    // 227-244).  File centers carry the basis functions of their label's BASIS sections and
    // nothing else: they join the atom tables as centers of charge zero, with no
    // pseudopotential and no electron-ion Jastrow, after the real atoms.
    const int nreal = T.na;
    std::vector<int> cen_atom;         // atom-table index of each basis center
    std::vector<std::string> cen_label;
    {
      Words cw;
      std::string cfile;
      if (section0(orb, cw, "CENTERS") && value(cw, cfile, "READ")) {
        std::ifstream cin(resolve(dir, cfile));
        if (!cin) bad("cannot open centers file " + cfile);
        int nc = 0;
        if (!(cin >> nc) || nc <= 0) bad("centers file " + cfile + " needs a positive center count");
        for (int c = 0; c < nc; c++) {
          std::string lab;
          double x[3];
          if (!(cin >> lab >> x[0] >> x[1] >> x[2])) bad("centers file " + cfile + " ends early");
          for (int d = 0; d < 3; d++) T.atom_pos.push_back(x[d]);
          T.atom_Z.push_back(0.0);
          T.atom_pp.push_back(-1);
          labels.push_back("");  // matches no PSEUDO or JASTROW label
          cen_atom.push_back(T.na + c);
          cen_label.push_back(lab);
        }
        T.na += nc;
      } else {
        if (!cw.empty() && !keyword(cw, "USEATOMS") && !keyword(cw, "USEGLOBAL"))
          bad("CENTERS needs USEATOMS, USEGLOBAL or READ");
        for (int a = 0; a < T.na; a++) {
          cen_atom.push_back(a);
          cen_label.push_back(labels[a]);
        }
      }
    }
    const int ncen = (int)cen_atom.size();
    int nmo_tot = 0;
    if (!ivalue(orb, nmo_tot, "NMO")) bad("need NMO");
    double magnify = 1;
    dvalue(orb, magnify, "MAGNIFY");
    std::vector<BasisObj> objs;
    pos = 0;
    Words bs;
    while (section(orb, pos, bs, "BASIS")) {
      if (bs.size() < 2 || !same(bs[1], "AOSPLINE")) unsupported("orbital basis must be AOSPLINE");
      objs.push_back(read_aospline(bs));
    }
    // spline tables (one per spline of each object) and shells placed on atoms
    std::vector<int> obj_tab0;
    for (auto& o : objs) {
      obj_tab0.push_back(T.ntab);
      T.ntab += (int)o.spl.size();
      for (auto& s : o.spl) T.nint = std::max(T.nint, (int)s.n());
    }
    T.spline.assign((size_t)T.ntab * T.nint * 4, 0.0);
    for (size_t oi = 0; oi < objs.size(); oi++)
      for (size_t s = 0; s < objs[oi].spl.size(); s++) {
        const SplineFit& f = objs[oi].spl[s];
        std::copy(f.coef.begin(), f.coef.end(), T.spline.begin() + (size_t)(obj_tab0[oi] + s) * T.nint * 4);
        T.tab_spacing.push_back(f.spacing);
        T.tab_invspacing.push_back(1.0 / f.spacing);
      }
    std::vector<int> maxf_center(ncen, 0), center_f0(ncen, 0);
    for (int c = 0; c < ncen; c++) {
      const int a = cen_atom[c];
      int fa = 0;
      center_f0[c] = T.nb;
      for (size_t oi = 0; oi < objs.size(); oi++) {
        if (objs[oi].label != cen_label[c]) continue;
        for (size_t s = 0; s < objs[oi].spl.size(); s++) {
          ShellDesc d{};
          d.table = obj_tab0[oi] + (int)s;
          d.sym = objs[oi].sym[s];
          d.f0 = T.nb;
          d.nf = sym_nf(d.sym);
          d.atom = a;
          d.rcut = objs[oi].cut[s];
          d.objthr = objs[oi].thr;
          T.shells.push_back(d);
          T.nb += d.nf;
          fa += d.nf;
        }
      }
      maxf_center[c] = fa;
    }
    int maxbasis = *std::max_element(maxf_center.begin(), maxf_center.end());
    // orbital file: "mo basis center label" lines, COEFFICIENTS, values
    std::string of;
    if (!value(orb, of, "ORBFILE")) bad("need ORBFILE");
    std::ifstream oin(resolve(dir, of));
    if (!oin) bad("cannot open orbital file " + of);
    struct Ent { int mo, f, c, l; };
    std::vector<Ent> ents;
    std::string tok;
    int maxlabel = 0, nmo_read = 0;
    while (true) {
      if (!(oin >> tok)) bad("orbital file has no COEFFICIENTS");
      if (tok == "COEFFICIENTS") break;
      std::string b1, c1, l1;
      oin >> b1 >> c1 >> l1;
      int m = std::atoi(tok.c_str()) - 1;
      if (m > nmo_tot) continue;
      Ent e{m, std::atoi(b1.c_str()) - 1, std::atoi(c1.c_str()) - 1, std::atoi(l1.c_str()) - 1};
      if (e.f < 0 || e.f >= maxbasis) bad("orbital file basis index out of range");
      if (e.c < 0 || e.c >= ncen) bad("orbital file center index out of range");
      ents.push_back(e);
      maxlabel = std::max(maxlabel, e.l + 1);
      nmo_read = std::max(nmo_read, m + 1);
    }
    std::vector<double> coef(maxlabel);
    for (auto& v : coef)
      if (!(oin >> v)) bad("orbital file ends before all coefficients");
    // the last line naming a (mo, center, function) wins; tiny coefficients are dropped after
    std::vector<std::vector<double>> moc(nmo_tot, std::vector<double>(T.nb, 0.0));
    for (auto& e : ents) {
      if (e.mo >= nmo_tot || e.f >= maxf_center[e.c]) continue;
      moc[e.mo][center_f0[e.c] + e.f] = coef[e.l];
    }
    for (auto& row : moc)
      for (double& v : row) v = std::fabs(v) > 1e-12 ? magnify * v : 0.0;
    // per spin: the MOs any determinant occupies, in order of first appearance, and each
    // determinant's occupation as indices into that list (src/wavefunction/Slat_wf_data.cpp:192-263)
    for (int s = 0; s < 2; s++) {
      const int n = s ? T.ndn : T.nup;
      std::vector<int> list;
      T.occ[s].assign((size_t)T.ndet * n, 0);
      for (int d = 0; d < T.ndet; d++)
        for (int i = 0; i < n; i++) {
          const int mo = std::atoi(st[(size_t)d * T.ne() + (s ? T.nup : 0) + i].c_str()) - 1;
          if (mo < 0 || mo >= nmo_tot) bad("STATES index out of range");
          int place = -1;
          for (size_t q = 0; q < list.size(); q++)
            if (list[q] == mo) place = (int)q;
          if (place < 0) {
            place = (int)list.size();
            list.push_back(mo);
          }
          for (int j = 0; j < i; j++)
            if (T.occ[s][(size_t)d * n + j] == place) bad("repeated orbital in one spin channel of a determinant");
          T.occ[s][(size_t)d * n + i] = place;
        }
      const int nm = (int)list.size();
      if (T.ndet > 1 && nm > 64) unsupported("several determinants with more than 64 orbitals of one spin are not supported");
      T.nmo[s] = nm;
      T.C[s].assign((size_t)T.nb * nm, 0.0);
      for (int i = 0; i < nm; i++)
        for (int f = 0; f < T.nb; f++) T.C[s][(size_t)f * nm + i] = moc[list[i]][f];
    }

    // ---- Jastrow
    T.atom_ei0.assign(T.na + 1, 0);
    T.atom_label = labels;
    std::vector<std::vector<JObj>> ei_atom(T.na);
    if (!wf2.empty()) {
      if (!keyword(wf2, "JASTROW2")) unsupported("WF2 must be JASTROW2");
      pos = 0;
      Words g;
      while (section(wf2, pos, g, "GROUP")) {
        if (keyword(g, "THREEBODY") || keyword(g, "THREEBODY_SPIN")) unsupported("three-body Jastrow is not supported");
        std::vector<JObj> eib, eeb;
        std::vector<std::string> eib_label;
        size_t p2 = 0;
        Words b;
        while (section(g, p2, b, "EIBASIS")) eib.push_back(read_jobj(b)), eib_label.push_back(b[0]);
        p2 = 0;
        while (section(g, p2, b, "EEBASIS")) eeb.push_back(read_jobj(b));
        Words ob;
        if (section0(g, ob, "ONEBODY")) {
          double scale = 1;
          dvalue(ob, scale, "SCALE");
          const bool frozen = keyword(ob, "FREEZE");
          std::map<std::string, std::vector<double>> cf;
          size_t p3 = 0;
          Words cs;
          while (section(ob, p3, cs, "COEFFICIENTS")) {
            if (cs.empty()) bad("empty COEFFICIENTS");
            std::vector<double> c;
            for (size_t i = 1; i < cs.size(); i++) c.push_back(std::atof(cs[i].c_str()) * scale);
            cf[cs[0]] = c;
            // parameters of this section: coefficient k multiplies the k-th function of the
            // atom type's electron-ion basis objects, concatenated in input order
            for (size_t k = 0; k < c.size() && !frozen; k++) {
              size_t kk = k;
              for (size_t q = 0; q < eib.size(); q++) {
                if (eib_label[q] != cs[0]) continue;
                if (kk < (size_t)eib[q].nf) {
                  JParm jp{1, 0, cs[0], c[k], eib[q]};
                  for (int i = 0; i < kMaxJF; i++) jp.unit.cl[i] = jp.unit.cu[i] = i == (int)kk ? 1.0 : 0.0;
                  T.jparms.push_back(jp);
                  break;
                }
                kk -= eib[q].nf;
              }
            }
          }
          for (int a = 0; a < nreal; a++) {
            auto it = cf.find(labels[a]);
            if (it == cf.end()) bad("ONEBODY has no coefficients for atom " + labels[a]);
            int k = 0;
            for (size_t q = 0; q < eib.size(); q++) {
              if (eib_label[q] != labels[a]) continue;
              JObj o = eib[q];
              for (int i = 0; i < o.nf; i++, k++) o.cl[i] = o.cu[i] = k < (int)it->second.size() ? it->second[k] : 0.0;
              add_merged(ei_atom[a], o);
            }
            if ((int)it->second.size() > k) bad("more one-body coefficients than electron-ion functions");
          }
        }
        Words tb, l1, l2;
        bool two = section0(g, tb, "TWOBODY"), twos = false;
        std::vector<double> cl, cu;
        if (two) {
          Words cs;
          if (!section0(tb, cs, "COEFFICIENTS")) bad("TWOBODY needs COEFFICIENTS");
          for (auto& x : cs) cl.push_back(std::atof(x.c_str()));
          cu = cl;
        }
        if (section0(g, tb, "TWOBODY_SPIN")) {
          if (two) bad("only one of TWOBODY and TWOBODY_SPIN per group");
          twos = true;
          if (!section0(tb, l1, "LIKE_COEFFICIENTS") || !section0(tb, l2, "UNLIKE_COEFFICIENTS"))
            bad("TWOBODY_SPIN needs LIKE_COEFFICIENTS and UNLIKE_COEFFICIENTS");
          if (l1.size() != l2.size()) bad("like and unlike coefficient counts differ");
          for (auto& x : l1) cl.push_back(std::atof(x.c_str()));
          for (auto& x : l2) cu.push_back(std::atof(x.c_str()));
        }
        if (two || twos) {
          int k = 0;
          int nfe = 0;
          for (auto& o : eeb) nfe += o.nf;
          if ((int)cl.size() > nfe) bad("not enough EE basis functions for the two-body coefficients");
          if (!keyword(tb, "FREEZE")) {
            // TWOBODY: one parameter per coefficient; TWOBODY_SPIN: the like ones, then unlike
            for (int sp = 0; sp < (twos ? 2 : 1); sp++)
              for (size_t c = 0; c < cl.size(); c++) {
                size_t kk = c;
                for (auto& o : eeb) {
                  if (kk < (size_t)o.nf) {
                    JParm jp{2, twos ? 1 + sp : 0, std::string(), sp ? cu[c] : cl[c], o};
                    for (int i = 0; i < kMaxJF; i++) {
                      const double one = i == (int)kk ? 1.0 : 0.0;
                      jp.unit.cl[i] = (!twos || sp == 0) ? one : 0.0;
                      jp.unit.cu[i] = (!twos || sp == 1) ? one : 0.0;
                    }
                    T.jparms.push_back(jp);
                    break;
                  }
                  kk -= o.nf;
                }
              }
          }
          for (auto o : eeb) {
            for (int i = 0; i < o.nf; i++, k++) {
              o.cl[i] = k < (int)cl.size() ? cl[k] : 0.0;
              o.cu[i] = k < (int)cu.size() ? cu[k] : 0.0;
            }
            add_merged(T.ee, o);
          }
        }
      }
    }
    for (int a = 0; a < T.na; a++) {
      T.atom_ei0[a] = (int)T.ei.size();
      for (auto& o : ei_atom[a]) T.ei.push_back(o);
    }
    T.atom_ei0[T.na] = (int)T.ei.size();
    return T;
  } catch (Unsupported& u) {
    if (unsup) *unsup = true;
    throw std::runtime_error(u.what());
  }
}

// One METHOD section (src/method/Vmc_method.cpp:33-80, src/method/Dmc_method.cpp:33-104).
MethodDesc method_of(const Words& m) {
  MethodDesc d;
  if (m.empty()) throw std::runtime_error("empty METHOD section");
  if (same(m[0], "VMC")) {
    d.method = 0;
    d.nblock = 100;
    d.nstep = 100;
    d.nconfig = 1;
    d.ndecorr = 2;
    d.timestep = -1;  // the reference tunes it when absent
    d.label = "vmc";
    dvalue(m, d.timestep, "TIMESTEP");
    ivalue(m, d.ndecorr, "NDECORR");
  } else if (same(m[0], "DMC")) {
    d.method = 1;
    d.label = "dmc";
    if (!dvalue(m, d.timestep, "TIMESTEP")) throw std::runtime_error("DMC needs TIMESTEP");
    d.nblock = 100;
    int target = 2048;
    ivalue(m, target, "TARGET_CONFIG");
    d.nconfig = std::max(target, 1);
    d.nstep = std::max(int(1.0 / d.timestep + 0.5), 1);
    dvalue(m, d.max_weight, "MAX_POSS_WEIGHT");
    if (ivalue(m, d.feedback_interval, "FEEDBACK_INTERVAL") && d.feedback_interval < 1)
      throw std::runtime_error("FEEDBACK_INTERVAL must be greater than or equal to 1");
    dvalue(m, d.feedback, "FEEDBACK");
    dvalue(m, d.start_feedback, "START_FEEDBACK");
    dvalue(m, d.branch_start_cutoff, "BRANCH_START_CUTOFF");
    // TMOVES takes precedence when both are given (src/method/Dmc_method.cpp:70-73,869)
    d.tmoves = keyword(m, "TMOVES") ? 1 : keyword(m, "TMOVESSC") ? 2 : 0;
    if (keyword(m, "PURE_DMC") || keyword(m, "ALL_ELECTRON"))
      throw std::runtime_error("DMC options PURE_DMC / ALL_ELECTRON are not supported");
  } else {
    throw std::runtime_error("METHOD must be VMC or DMC, found " + m[0]);
  }
  ivalue(m, d.nblock, "NBLOCK");
  ivalue(m, d.nstep, "NSTEP");
  ivalue(m, d.nconfig, "NCONFIG");
  value(m, d.label, "LABEL");
  return d;
}

// Every METHOD section of a deck, in input order (the reference runs them one after another,
// src/program_utils/ooqmc.cpp:148-170).
std::vector<MethodDesc> read_methods(const std::string& path) {
  Words all, m;
  lex(path, all);
  std::vector<MethodDesc> out;
  size_t pos = 0;
  while (section(all, pos, m, "METHOD")) {
    out.push_back(method_of(m));
    MethodDesc& d = out.back();
    d.readconfig = d.storeconfig = path + ".config";
    value(m, d.readconfig, "READCONFIG");
    value(m, d.storeconfig, "STORECONFIG");
  }
  if (out.empty()) throw std::runtime_error("no METHOD section in " + path);
  return out;
}

MethodDesc read_method(const std::string& path) { return read_methods(path)[0]; }

}  // namespace qw
