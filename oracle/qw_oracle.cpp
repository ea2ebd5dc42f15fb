// Test infrastructure only: the oracle (see qw_oracle.h).
//
// FP64 CPU restatement of the QWalk VMC/DMC hot path.  Every function names the reference
// file:line whose algorithm it restates; the code itself is written for this project.
// Built with g++ -O2 (no fast-math) by oracle/Makefile.

#include "qw_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

using Tok = std::vector<std::string>;
struct OErr : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] void fail(const std::string& m) { throw OErr(m); }

const double PI = 3.14159265358979323846;

bool ieq(const std::string& a, const char* b) {
  size_t n = strlen(b);
  if (a.size() != n) return false;
  for (size_t i = 0; i < n; i++)
    if (toupper((unsigned char)a[i]) != toupper((unsigned char)b[i])) return false;
  return true;
}

// ---------------------------------------------------------------- input text
// Tokeniser: whitespace words, '#' comments to end of line, braces as separate words,
// INCLUDE expanded recursively (src/program_utils/qmc_io.cpp:280-352).  Paths are resolved
// against the including file's directory.
std::string dirname_of(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
std::string join_path(const std::string& dir, const std::string& f) {
  if (!f.empty() && f[0] == '/') return f;
  return dir + "/" + f;
}

void parse_file(const std::string& path, Tok& out) {
  std::ifstream in(path);
  if (!in) fail("cannot open " + path);
  Tok raw;
  std::string line;
  while (std::getline(in, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line.erase(h);
    std::string cur;
    for (char ch : line) {
      if (isspace((unsigned char)ch)) {
        if (!cur.empty()) raw.push_back(cur), cur.clear();
      } else if (ch == '{' || ch == '}') {
        if (!cur.empty()) raw.push_back(cur), cur.clear();
        raw.push_back(std::string(1, ch));
      } else {
        cur.push_back(ch);
      }
    }
    if (!cur.empty()) raw.push_back(cur);
  }
  std::string dir = dirname_of(path);
  for (size_t i = 0; i < raw.size(); i++) {
    if (ieq(raw[i], "INCLUDE")) {
      if (i + 1 >= raw.size()) fail("INCLUDE without a file");
      parse_file(join_path(dir, raw[i + 1]), out);
      i++;
    } else {
      out.push_back(raw[i]);
    }
  }
}

// Level-0 section / keyword search (src/program_utils/qmc_io.cpp:92-120,168-269;
// src/program_utils/qmc_io.h:74-110).
bool readsection(const Tok& in, size_t& pos, Tok& sec, const char* name) {
  sec.clear();
  int level = 0;
  for (; pos < in.size(); pos++) {
    if (in[pos] == "{") level++;
    else if (in[pos] == "}") level--;
    if (level < 0) {
      pos++;
      return false;
    }
    if (level == 0 && ieq(in[pos], name)) {
      pos++;
      if (pos >= in.size() || in[pos] != "{") fail(std::string("section ") + name + " needs {");
      pos++;
      size_t begin = pos;
      int n = 0;
      for (; pos < in.size(); pos++) {
        if (in[pos] == "{") n++;
        if (in[pos] == "}") n--;
        if (n < 0) break;
      }
      sec.assign(in.begin() + begin, in.begin() + pos);
      pos++;
      return true;
    }
  }
  return false;
}
bool readsection0(const Tok& in, Tok& sec, const char* name) {
  size_t p = 0;
  return readsection(in, p, sec, name);
}
bool readvalue(const Tok& in, std::string& v, const char* name) {
  int level = 0;
  for (size_t pos = 0; pos < in.size(); pos++) {
    if (in[pos] == "{") level++;
    else if (in[pos] == "}") level--;
    if (level < 0) return false;
    if (level == 0 && ieq(in[pos], name)) {
      if (pos + 1 >= in.size()) fail(std::string("value expected after ") + name);
      v = in[pos + 1];
      return true;
    }
  }
  return false;
}
bool readdouble(const Tok& in, double& v, const char* name) {
  std::string s;
  if (!readvalue(in, s, name)) return false;
  v = atof(s.c_str());
  return true;
}
bool haskeyword(const Tok& in, const char* name) {
  int level = 0;
  for (size_t pos = 0; pos < in.size(); pos++) {
    if (in[pos] == "{") level++;
    else if (in[pos] == "}") level--;
    if (level < 0) return false;
    if (level == 0 && ieq(in[pos], name)) return true;
  }
  return false;
}

// ---------------------------------------------------------------- cubic spline
// Clamped cubic spline on a uniform grid, coefficients per interval
// (src/basis_function/Spline_fitter.cpp:208-289, Numerical Recipes 'spline').
struct Spline {
  double spacing = 0, invspacing = 0, threshold = 0;
  std::vector<double> c;  // [n][4]
  int n() const { return (int)c.size() / 4; }

  void fit(const std::vector<double>& x, const std::vector<double>& y, double yp1, double ypn) {
    int n = (int)x.size();
    c.assign(4 * n, 0.0);
    spacing = x[1] - x[0];
    threshold = x[n - 1];
    invspacing = 1.0 / spacing;
    std::vector<double> y2(n), u(n);
    if (yp1 > .99e30) {
      y2[0] = 0;
      u[0] = 0;
    } else {
      y2[0] = -.5;
      u[0] = (3. / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    }
    for (int i = 1; i < n - 1; i++) {
      double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
      double p = sig * y2[i - 1] + 2.;
      y2[i] = (sig - 1.) / p;
      u[i] = (6. * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                  (x[i + 1] - x[i - 1]) -
              sig * u[i - 1]) /
             p;
    }
    double qn, un;
    if (ypn > .99e30) {
      qn = 0;
      un = 0;
    } else {
      qn = .5;
      un = (3. / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    }
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
    for (int i = 0; i < n - 1; i++) {
      double h = x[i + 1] - x[i];
      c[4 * i + 0] = y[i];
      c[4 * i + 1] = (y[i + 1] - y[i]) / h - h * (2. * y2[i] + y2[i + 1]) / 6.;
      c[4 * i + 2] = y2[i] / 2.;
      c[4 * i + 3] = (y2[i + 1] - y2[i]) / (6. * h);
    }
    c[4 * (n - 1) + 0] = y[n - 1];
  }

  double value(double r, int i) const {
    double h = r - i * spacing;
    const double* q = &c[4 * i];
    return q[0] + h * (q[1] + h * (q[2] + h * q[3]));
  }
  // f, f'/r, f'' (src/basis_function/Spline_fitter.h:65-78)
  void ders(double r, int i, double& f, double& fd, double& f2) const {
    double h = r - i * spacing;
    const double* q = &c[4 * i];
    f = q[0] + h * (q[1] + h * (q[2] + h * q[3]));
    fd = (q[1] + h * (2 * q[2] + h * (3 * q[3]))) / r;
    f2 = 2 * q[2] + 6 * h * q[3];
  }
  // src/basis_function/Spline_fitter.cpp:153-182
  double find_cutoff() const {
    const double step = 0.5;
    double thr = 1e-10;
    double mx = 0;
    for (int i = 0; i < n(); i++) mx = std::max(mx, fabs(c[4 * i]));
    mx = std::min(mx, 1.0);
    thr *= mx;
    for (double x = threshold - step; x >= 0; x -= step) {
      int iv = int(x * invspacing);
      if (fabs(value(x, iv)) > thr) return x + step;
    }
    return 0;
  }
  // src/basis_function/Spline_fitter.cpp:186-204
  void pad(double thresh) {
    if (threshold >= thresh) return;
    int nnew = int((thresh - threshold) * invspacing) + 1;
    c.resize(c.size() + 4 * nnew, 0.0);
    threshold = thresh;
  }
};

// Angular symmetries (src/basis_function/Cubic_spline_build.cpp:26-120,664-800).
enum Sym { SYM_S, SYM_P, SYM_5D, SYM_6D, SYM_10F, SYM_7F, SYM_7FC, SYM_9G, SYM_9GP, SYM_15G };
enum ISym {
  I_S, I_PX, I_PY, I_PZ,
  I_DXX, I_DYY, I_DZZ, I_DXY, I_DXZ, I_DYZ, I_DX2Y2, I_DZ2R2,
  I_FXXX, I_FYYY, I_FZZZ, I_FXXY, I_FXXZ, I_FYYX, I_FYYZ, I_FZZX, I_FZZY, I_FXYZ
};
constexpr int kPoly = 1000;  // individual-function codes >= kPoly: ang_polys()[code - kPoly]

int sym_lvalue(Sym s) {
  switch (s) {
    case SYM_S: return 0;
    case SYM_P: return 1;
    case SYM_5D:
    case SYM_6D: return 2;
    case SYM_10F:
    case SYM_7F:
    case SYM_7FC: return 3;
    default: return 4;
  }
}
int sym_nfunc(Sym s) {
  switch (s) {
    case SYM_S: return 1;
    case SYM_P: return 3;
    case SYM_5D: return 5;
    case SYM_6D: return 6;
    case SYM_10F: return 10;
    case SYM_7F:
    case SYM_7FC: return 7;
    case SYM_9G:
    case SYM_9GP: return 9;
    case SYM_15G: return 15;
  }
  return 0;
}

// 7F and 9G real solid harmonics as the reference writes them (unnormalised) and the 15G
// Cartesian monomials, as polynomials {coefficient, power of x, y, z}
// (src/basis_function/Cubic_spline_calc.cpp:646-720,720-1070; orders
// src/basis_function/Cubic_spline_build.cpp:702-776).  Indexed by function: F0 Fm3 Fp3mod
// Fp2 Fxyz Fm1 Fp1 (0-6), G0..G8 (7-15), xxxx yyyy zzzz xxxy xxxz yyyx yyyz zzzx zzzy xxyy
// xxzz yyzz xxyz yyxz zzxy (16-30).
struct PolyTerm {
  int c, a, b, d;
};
const std::vector<std::vector<PolyTerm>>& ang_polys() {
  static const std::vector<std::vector<PolyTerm>> P = {
      {{2, 0, 0, 3}, {-3, 2, 0, 1}, {-3, 0, 2, 1}},   // z(2z^2 - 3x^2 - 3y^2)
      {{3, 2, 1, 0}, {-1, 0, 3, 0}},                  // y(3x^2 - y^2)
      {{1, 3, 0, 0}, {-3, 1, 2, 0}},                  // x(x^2 - 3y^2)
      {{1, 2, 0, 1}, {-1, 0, 2, 1}},                  // z(x^2 - y^2)
      {{1, 1, 1, 1}},                                 // xyz
      {{4, 0, 1, 2}, {-1, 2, 1, 0}, {-1, 0, 3, 0}},   // y(4z^2 - x^2 - y^2)
      {{4, 1, 0, 2}, {-1, 3, 0, 0}, {-1, 1, 2, 0}},   // x(4z^2 - x^2 - y^2)
      {{3, 4, 0, 0}, {3, 0, 4, 0}, {8, 0, 0, 4}, {6, 2, 2, 0}, {-24, 2, 0, 2}, {-24, 0, 2, 2}},  // 35z^4-30z^2r^2+3r^4
      {{4, 1, 0, 3}, {-3, 3, 0, 1}, {-3, 1, 2, 1}},   // xz(7z^2 - 3r^2)
      {{4, 0, 1, 3}, {-3, 2, 1, 1}, {-3, 0, 3, 1}},   // yz(7z^2 - 3r^2)
      {{-1, 4, 0, 0}, {1, 0, 4, 0}, {6, 2, 0, 2}, {-6, 0, 2, 2}},  // (x^2 - y^2)(7z^2 - r^2)
      {{6, 1, 1, 2}, {-1, 3, 1, 0}, {-1, 1, 3, 0}},   // xy(7z^2 - r^2)
      {{1, 3, 0, 1}, {-3, 1, 2, 1}},                  // xz(x^2 - 3y^2)
      {{3, 2, 1, 1}, {-1, 0, 3, 1}},                  // yz(3x^2 - y^2)
      {{1, 4, 0, 0}, {-6, 2, 2, 0}, {1, 0, 4, 0}},    // x^4 - 6x^2y^2 + y^4
      {{1, 3, 1, 0}, {-1, 1, 3, 0}},                  // xy(x^2 - y^2)
      {{1, 4, 0, 0}}, {{1, 0, 4, 0}}, {{1, 0, 0, 4}}, {{1, 3, 1, 0}}, {{1, 3, 0, 1}}, {{1, 1, 3, 0}},
      {{1, 0, 3, 1}}, {{1, 1, 0, 3}}, {{1, 0, 1, 3}}, {{1, 2, 2, 0}}, {{1, 2, 0, 2}}, {{1, 0, 2, 2}},
      {{1, 2, 1, 1}}, {{1, 1, 2, 1}}, {{1, 1, 1, 2}}};
  return P;
}
inline double ipow(double v, int k) {
  double r = 1;
  for (int i = 0; i < k; i++) r *= v;
  return r;
}
Sym sym_lookup(const std::string& s) {
  if (ieq(s, "S")) return SYM_S;
  if (ieq(s, "P")) return SYM_P;
  if (ieq(s, "D") || ieq(s, "6D")) return SYM_6D;
  if (ieq(s, "5D")) return SYM_5D;
  if (ieq(s, "F") || ieq(s, "10F")) return SYM_10F;
  if (ieq(s, "7F")) return SYM_7F;
  if (ieq(s, "7F_crystal")) return SYM_7FC;
  if (ieq(s, "G") || ieq(s, "15G")) return SYM_15G;
  if (ieq(s, "9G")) return SYM_9G;
  if (ieq(s, "9G_pyscf")) return SYM_9GP;
  fail("unsupported AOSPLINE symmetry " + s);
}

// AOSPLINE basis object (src/basis_function/Cubic_spline_build.cpp:140-212,328-382,396-655;
// evaluation src/basis_function/Cubic_spline_calc.cpp:25-700).
struct AOSpline {
  std::string label;
  std::vector<Sym> sym;
  std::vector<Spline> spl;
  std::vector<int> indiv;   // per function
  std::vector<double> rcut; // per function
  double threshold = 0;
  int nfunc = 0;

  void read(const Tok& w) {
    label = w[0];
    double spacing = 0.02;
    int npts = 2500;
    std::string norm = "GAMESSNORM";
    readvalue(w, norm, "NORMTYPE");
    bool renorm = !haskeyword(w, "NORENORMALIZE");
    double reqcut = -1;
    readdouble(w, reqcut, "CUTOFF");
    if (haskeyword(w, "CUSP") || haskeyword(w, "CUSP_MATCHING")) fail("AOSPLINE CUSP not supported");
    bool zero_der = haskeyword(w, "ZERO_DERIVATIVE");
    double cs;
    if (readdouble(w, cs, "SPACING")) {
      spacing = cs;
      npts = (int)(50.0 / cs);
    }
    Tok g;
    if (!readsection0(w, g, "GAMESS")) fail("AOSPLINE needs a GAMESS section");
    std::vector<std::vector<double>> ex, co;
    for (size_t p = 0; p < g.size(); p++) {
      Sym s = sym_lookup(g[p]);
      int ng = atoi(g[++p].c_str());
      if (p + ng * 3 >= g.size()) fail("GAMESS section count is wrong");
      std::vector<double> e(ng), c(ng);
      for (int i = 0; i < ng; i++) {
        p++;
        e[i] = atof(g[++p].c_str());
        c[i] = atof(g[++p].c_str());
      }
      sym.push_back(s);
      ex.push_back(e);
      co.push_back(c);
    }
    int ns = (int)sym.size();
    std::vector<double> x(npts), y(npts);
    spl.resize(ns);
    for (int f = 0; f < ns; f++) {
      for (int j = 0; j < npts; j++) x[j] = j * spacing;
      std::fill(y.begin(), y.end(), 0.0);
      for (size_t i = 0; i < ex[f].size(); i++) {
        double a = ex[f][i], nrm;
        if (norm == "GAMESSNORM") {
          double fac = sqrt(2. * a / PI), feg = 4. * a, feg2 = feg * feg;
          switch (sym[f]) {
            case SYM_S: nrm = sqrt(2. * feg * fac); break;
            case SYM_P: nrm = sqrt(2. * feg2 * fac / 3.); break;
            case SYM_5D:
            case SYM_6D: nrm = sqrt(2. * feg * feg2 * fac / 15.); break;
            case SYM_9G:
            case SYM_9GP:
            case SYM_15G: nrm = sqrt(2. * feg2 * feg2 * feg * fac / 945.); break;
            default: nrm = sqrt(2. * feg2 * feg2 * fac / 105.); break;
          }
        } else if (norm == "CRYSTAL") {
          int l = sym_lvalue(sym[f]);
          nrm = l == 0 ? 1.0 : 1.0 / sqrt(2.0 * l + 1.0);
        } else if (norm == "NONE") {
          nrm = 1;
        } else {
          fail("unknown NORMTYPE " + norm);
        }
        for (int j = 0; j < npts; j++) y[j] += nrm * co[f][i] * exp(-a * x[j] * x[j]);
      }
      if (renorm) {
        int p = sym_lvalue(sym[f]) + 1;
        double sum = 0;
        for (int j = 0; j < npts; j++) sum += spacing * pow(x[j], 2 * p) * y[j] * y[j];
        if (fabs(sum - 1) > 1e-5)
          for (int j = 0; j < npts; j++) y[j] = y[j] / sqrt(sum);
      }
      if (reqcut > 0) {
        const double smooth = 1.2, cutmax = reqcut - 1e-6, cutmin = cutmax - smooth;
        for (int j = 0; j < npts; j++) {
          if (x[j] > cutmin) {
            if (x[j] > cutmax) y[j] = 0;
            else {
              double zz = (x[j] - cutmin) / smooth;
              y[j] *= (1 - zz * zz * zz * (6 * zz * zz - 15 * zz + 10));
            }
          }
        }
      }
      double yp1 = zero_der ? 0.0 : (y[1] - y[0]) / spacing;
      double ypn = (y[npts - 1] - y[npts - 2]) / spacing;
      spl[f].fit(x, y, yp1, ypn);
    }
    // individual symmetries (src/basis_function/Cubic_spline_build.cpp:664-800)
    for (int f = 0; f < ns; f++) {
      switch (sym[f]) {
        case SYM_S: indiv.push_back(I_S); break;
        case SYM_P: indiv.insert(indiv.end(), {I_PX, I_PY, I_PZ}); break;
        case SYM_5D: indiv.insert(indiv.end(), {I_DZ2R2, I_DXZ, I_DYZ, I_DX2Y2, I_DXY}); break;
        case SYM_6D: indiv.insert(indiv.end(), {I_DXX, I_DYY, I_DZZ, I_DXY, I_DXZ, I_DYZ}); break;
        case SYM_10F:
          indiv.insert(indiv.end(), {I_FXXX, I_FYYY, I_FZZZ, I_FXXY, I_FXXZ, I_FYYX, I_FYYZ,
                                     I_FZZX, I_FZZY, I_FXYZ});
          break;
        // polynomial functions: kPoly + index into ang_polys()
        case SYM_7F:
          for (int q : {0, 1, 2, 3, 4, 5, 6}) indiv.push_back(kPoly + q);
          break;
        case SYM_7FC:
          for (int q : {0, 6, 5, 3, 4, 2, 1}) indiv.push_back(kPoly + q);
          break;
        case SYM_9G:
          for (int q = 7; q <= 15; q++) indiv.push_back(kPoly + q);
          break;
        case SYM_9GP:
          for (int q : {15, 13, 11, 9, 7, 8, 10, 12, 14}) indiv.push_back(kPoly + q);
          break;
        case SYM_15G:
          for (int q = 16; q <= 30; q++) indiv.push_back(kPoly + q);
          break;
      }
    }
    nfunc = (int)indiv.size();
    // cutoffs (src/basis_function/Cubic_spline_build.cpp:328-382)
    threshold = 0;
    for (int f = 0; f < ns; f++) {
      double cut = spl[f].find_cutoff();
      for (int k = 0; k < sym_nfunc(sym[f]); k++) rcut.push_back(cut);
      threshold = std::max(threshold, cut);
    }
    for (int f = 0; f < ns; f++) spl[f].pad(threshold);
  }

  // values+gradient+laplacian, r = [r, r^2, x, y, z]; out[nfunc][5]
  void calc_lap(const double* r, double* out) const {
    if (r[0] >= threshold) {
      std::fill(out, out + 5 * nfunc, 0.0);
      return;
    }
    int tot = 0;
    for (size_t i = 0; i < spl.size(); i++) {
      int iv = int(r[0] * spl[i].invspacing);
      double f, fd, f2;
      spl[i].ders(r[0], iv, f, fd, f2);
      int end = tot + sym_nfunc(sym[i]);
      for (; tot < end; tot++) {
        double* o = out + 5 * tot;
        double x = r[2], y = r[3], z = r[4], v, h, v2, v3, rdp;
        if (indiv[tot] >= kPoly) {
          // psi = P f: grad = f grad P + (f'/r) P r_vec, lap = f lap P + 2 (f'/r) r.grad P + P (f'' + 2 f'/r)
          const std::vector<PolyTerm>& T = ang_polys()[indiv[tot] - kPoly];
          double P = 0, G[3] = {0, 0, 0}, L = 0, rg = 0;
          for (const PolyTerm& t : T) {
            const int e[3] = {t.a, t.b, t.d};
            const double c[3] = {x, y, z};
            P += t.c * ipow(x, t.a) * ipow(y, t.b) * ipow(z, t.d);
            for (int k = 0; k < 3; k++) {
              if (e[k] == 0) continue;
              double term = t.c * e[k];
              for (int q = 0; q < 3; q++) term *= ipow(c[q], q == k ? e[q] - 1 : e[q]);
              G[k] += term;
              if (e[k] > 1) {
                double t2 = t.c * e[k] * (e[k] - 1);
                for (int q = 0; q < 3; q++) t2 *= ipow(c[q], q == k ? e[q] - 2 : e[q]);
                L += t2;
              }
            }
          }
          rg = x * G[0] + y * G[1] + z * G[2];
          o[0] = P * f;
          o[1] = f * G[0] + fd * P * x;
          o[2] = f * G[1] + fd * P * y;
          o[3] = f * G[2] + fd * P * z;
          o[4] = f * L + 2 * fd * rg + P * (f2 + 2 * fd);
          continue;
        }
        switch (indiv[tot]) {
          case I_S:
            o[0] = f; o[1] = fd * x; o[2] = fd * y; o[3] = fd * z; o[4] = f2 + 2. * fd;
            break;
          case I_PX:
            o[0] = f * x; rdp = fd * x;
            o[1] = rdp * x + f; o[2] = rdp * y; o[3] = rdp * z; o[4] = f2 * x + 4. * rdp;
            break;
          case I_PY:
            o[0] = f * y; rdp = fd * y;
            o[1] = rdp * x; o[2] = rdp * y + f; o[3] = rdp * z; o[4] = f2 * y + 4. * rdp;
            break;
          case I_PZ:
            o[0] = f * z; rdp = fd * z;
            o[1] = rdp * x; o[2] = rdp * y; o[3] = rdp * z + f; o[4] = f2 * z + 4. * rdp;
            break;
          case I_DXX:
            v = x * x; o[0] = v * f; h = v * fd;
            o[1] = h * x + f * 2. * x; o[2] = h * y; o[3] = h * z; o[4] = v * f2 + 6. * h + 2. * f;
            break;
          case I_DYY:
            v = y * y; o[0] = v * f; h = v * fd;
            o[1] = h * x; o[2] = h * y + f * 2. * y; o[3] = h * z; o[4] = v * f2 + 6. * h + 2. * f;
            break;
          case I_DZZ:
            v = z * z; o[0] = v * f; h = v * fd;
            o[1] = h * x; o[2] = h * y; o[3] = h * z + f * 2. * z; o[4] = v * f2 + 6. * h + 2. * f;
            break;
          case I_DXY:
            v = x * y; o[0] = v * f; h = v * fd;
            o[1] = h * x + f * y; o[2] = h * y + f * x; o[3] = h * z; o[4] = v * f2 + 6. * h;
            break;
          case I_DXZ:
            v = x * z; o[0] = v * f; h = v * fd;
            o[1] = h * x + f * z; o[2] = h * y; o[3] = h * z + f * x; o[4] = v * f2 + 6. * h;
            break;
          case I_DYZ:
            v = y * z; o[0] = v * f; h = v * fd;
            o[1] = h * x; o[2] = h * y + f * z; o[3] = h * z + f * y; o[4] = v * f2 + 6. * h;
            break;
          case I_DX2Y2:
            v = x * x - y * y; o[0] = v * f; h = v * fd;
            o[1] = h * x + 2. * f * x; o[2] = h * y - 2. * f * y; o[3] = h * z; o[4] = v * f2 + 6. * h;
            break;
          case I_DZ2R2:
            v = 2. * z * z - x * x - y * y; o[0] = v * f; h = v * fd;
            o[1] = h * x - 2 * f * x; o[2] = h * y - 2 * f * y; o[3] = h * z + 4 * f * z; o[4] = v * f2 + 6 * h;
            break;
          case I_FXXX:
            v2 = x * x; v3 = v2 * x; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x + f * 3. * v2; o[2] = h * y; o[3] = h * z; o[4] = v3 * f2 + 8. * h + 6. * f * x;
            break;
          case I_FYYY:
            v2 = y * y; v3 = v2 * y; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x; o[2] = h * y + 3. * f * v2; o[3] = h * z; o[4] = v3 * f2 + 8. * h + 6. * f * y;
            break;
          case I_FZZZ:
            v2 = z * z; v3 = v2 * z; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x; o[2] = h * y; o[3] = h * z + 3. * f * v2; o[4] = v3 * f2 + 8. * h + 6. * f * z;
            break;
          case I_FXXY:
            v2 = x * x; v3 = v2 * y; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x + 2. * f * x * y; o[2] = h * y + f * v2; o[3] = h * z; o[4] = v3 * f2 + 8. * h + 2. * f * y;
            break;
          case I_FXXZ:
            v2 = x * x; v3 = v2 * z; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x + 2. * f * x * z; o[2] = h * y; o[3] = h * z + f * v2; o[4] = v3 * f2 + 8. * h + 2. * f * z;
            break;
          case I_FYYX:
            v2 = y * y; v3 = v2 * x; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x + v2 * f; o[2] = h * y + 2. * f * x * y; o[3] = h * z; o[4] = v3 * f2 + 8. * h + 2. * f * x;
            break;
          case I_FYYZ:
            v2 = y * y; v3 = v2 * z; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x; o[2] = h * y + 2. * f * y * z; o[3] = h * z + f * v2; o[4] = v3 * f2 + 8. * h + 2. * f * z;
            break;
          case I_FZZX:
            v2 = z * z; v3 = x * v2; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x + f * v2; o[2] = h * y; o[3] = h * z + 2. * f * x * z; o[4] = v3 * f2 + 8. * h + 2. * f * x;
            break;
          case I_FZZY:
            v2 = z * z; v3 = v2 * y; o[0] = v3 * f; h = v3 * fd;
            o[1] = h * x; o[2] = h * y + f * v2; o[3] = h * z + 2. * f * y * z; o[4] = v3 * f2 + 8. * h + 2. * f * y;
            break;
          case I_FXYZ:
            v3 = x * y * z; o[0] = v3 * f; v2 = f * z; h = v3 * fd;
            o[1] = h * x + v2 * y; o[2] = h * y + v2 * x; o[3] = h * z + f * x * y; o[4] = v3 * f2 + 8. * h;
            break;
        }
      }
    }
  }
};

// ---------------------------------------------------------------- Jastrow radial functions
// CUTOFF_CUSP (src/basis_function/Cutoff_cusp.cpp:27-154) and POLYPADE
// (src/basis_function/Poly_pade_function.cpp:27-237).
struct JBasis {
  int kind = 0;  // 0 cusp, 1 polypade
  std::string label;
  double cusp = 0, gamma = 0, rcut = 0, rcutinv = 0, pade0 = 0;
  std::vector<double> beta;
  int nfunc() const { return kind == 0 ? 1 : (int)beta.size(); }
  double cutoff() const { return rcut; }

  void read(const Tok& w) {
    label = w[0];
    if (w.size() < 2) fail("basis section too short");
    if (ieq(w[1], "CUTOFF_CUSP")) {
      kind = 0;
      if (!readdouble(w, cusp, "CUSP")) fail("CUTOFF_CUSP needs CUSP");
      if (!readdouble(w, gamma, "GAMMA")) fail("CUTOFF_CUSP needs GAMMA");
      if (!readdouble(w, rcut, "CUTOFF")) fail("CUTOFF_CUSP needs CUTOFF");
      rcutinv = 1.0 / rcut;
      pade0 = 1. / (3. + gamma);
    } else if (ieq(w[1], "POLYPADE")) {
      kind = 1;
      if (!readdouble(w, rcut, "RCUT")) fail("POLYPADE needs RCUT");
      double beta0;
      std::string nf;
      if (!readdouble(w, beta0, "BETA0") || !readvalue(w, nf, "NFUNC")) fail("POLYPADE needs BETA0 and NFUNC");
      int nmax = atoi(nf.c_str());
      beta.resize(nmax);
      beta[0] = beta0;
      double beta1 = log(beta0 + 1.00001);
      for (int i = 1; i < nmax; i++) beta[i] = exp(beta1 + 1.6 * i) - 1;
    } else {
      fail("unsupported Jastrow basis " + w[1]);
    }
  }
  // out[nfunc][5]
  void calc_lap(const double* r, double* out) const {
    if (kind == 0) {
      if (r[0] < rcut) {
        double zz = r[0] * rcutinv, zz2 = zz * zz, pp = zz - zz2 + zz * zz2 / 3;
        double pade = 1. / (1 + gamma * pp);
        out[0] = cusp * rcut * (pp * pade - pade0);
        double pade2 = pade * pade, ppd = 1. - 2. * zz + zz2, ppdd = -2. + 2. * zz;
        double dadr = ppd * pade2 / r[0];
        double dadr2 = ppdd * pade2 * rcutinv - 2. * gamma * ppd * ppd * pade2 * pade * rcutinv + 2. * dadr;
        for (int i = 1; i < 4; i++) out[i] = cusp * dadr * r[i + 1];
        out[4] = cusp * dadr2;
      } else {
        for (int i = 0; i < 5; i++) out[i] = 0;
      }
      return;
    }
    int nmax = (int)beta.size();
    if (r[0] >= rcut) {
      std::fill(out, out + 5 * nmax, 0.0);
      return;
    }
    double zz = r[0] / rcut, zz2 = zz * zz, zpp = zz2 * (6. - 8. * zz + 3. * zz2), zpp1 = 1 - zpp;
    double zppd = 12. * zz * (1. - 2. * zz + zz2), zppdd2 = 6. - 24. * zz + 18 * zz2;
    for (int i = 0; i < nmax; i++) {
      double zpade = 1. / (1 + beta[i] * zpp), zpade2 = zpade * zpade;
      double crs = zpp1 * zpade;
      double crsd = -zppd * zpade2 * (beta[i] + 1) / (r[0] * rcut);
      double crsdd2 = -zpade2 * (zppdd2 - beta[i] * zppd * zppd * zpade) * (beta[i] + 1) / (rcut * rcut);
      double* o = out + 5 * i;
      o[0] = crs;
      for (int d = 1; d < 4; d++) o[d] = crsd * r[d + 1];
      o[4] = 2 * (crsdd2 + crsd);
    }
  }
};

struct JGroup {
  // one-body (src/wavefunction/Jastrow2_one.cpp:24-194)
  bool has_one = false, has_two = false, diffspin = false;
  std::vector<JBasis> eib, eeb;
  std::vector<std::vector<int>> atom2basis;   // per atom: eib indices
  std::vector<std::vector<double>> onecoef;   // per atom (resolved from label)
  std::vector<double> twocoef;                // TWOBODY
  std::vector<double> like, unlike;           // TWOBODY_SPIN
  double eecut = 0;                           // per EE basis cutoff handled per object
  // parameter bookkeeping (src/wavefunction/Jastrow2_one.cpp:25-100,250-258,
  // Jastrow2_two.h:96,148): FREEZE flags and the one-body COEFFICIENTS sections in input order
  bool one_frozen = false, two_frozen = false;
  std::vector<std::string> sec_label;
  std::vector<std::vector<double>> sec_coef;
};

struct Atom {
  std::string label;
  double Z;
  double r[3];
};

struct PPAtom {
  // RGAUSSIAN OLDQMC (src/basis_function/Rgaussian_function.cpp:24-202)
  std::vector<std::vector<double>> n, a, c;  // per channel
  int aip = 6;
  bool addzeff = false;
  double cutoff = 0;
  std::vector<double> qx, qy, qz, qw;        // unrotated quadrature
};

// ---------------------------------------------------------------- the system
struct System {
  std::vector<Atom> atoms;
  int nup = 0, ndn = 0;
  // basis / MOs
  std::vector<AOSpline> aos;                 // BASIS objects
  std::vector<std::vector<int>> center_basis;  // per atom: AOSpline indices
  int totbasis = 0;
  std::vector<double> fcut;                  // per function
  std::vector<int> fobj;                     // basis object of each function (for obj cutoff)
  std::vector<std::vector<double>> C;        // per spin [totbasis][nmo_s] dense
  std::vector<int> nmo;                      // per spin list length
  // determinants (src/wavefunction/Slat_wf_data.cpp:34-263): weights and, per spin, each
  // determinant's occupied orbitals as indices into that spin's MO list
  int ndet = 1;
  std::vector<double> detwt;
  std::vector<std::vector<int>> occ[2];      // [spin][det][n]
  // CSF structure: coefficient and determinant weights of each CSF, over consecutive
  // determinants (DETWT input: one CSF {detwt, 1} per determinant); OPTIMIZE_DET makes the
  // coefficients of CSFs 1.. parameters (src/wavefunction/Slat_wf_data.cpp:69-118,574-640)
  std::vector<double> csf_c;
  std::vector<std::vector<double>> csf_w;
  bool optimize_det = false;
  // Jastrow
  std::vector<JGroup> groups;
  // PP
  std::vector<int> pp_index;                 // per atom: -1 or index into pp
  std::vector<PPAtom> pp;

  int ne() const { return nup + ndn; }
  int nspin(int s) const { return s == 0 ? nup : ndn; }
  int spin_of(int e) const { return e < nup ? 0 : 1; }
};

// Rgaussian value: sum c r^n exp(-min(a r^2, 60)) with n already decremented by two
// (src/basis_function/Rgaussian_function.cpp:180-202).
void rgauss_val(const PPAtom& p, const double* r, double* vl) {
  for (size_t l = 0; l < p.n.size(); l++) {
    double v = 0;
    for (size_t j = 0; j < p.n[l].size(); j++) {
      double ex = std::min(p.a[l][j] * r[1], 60.0);
      v += p.c[l][j] * pow(r[0], p.n[l][j]) * exp(-ex);
    }
    vl[l] = v;
  }
}
// getRadial with optional ADD_ZEFF on the local channel (src/system/Pseudopotential.cpp:305-318).
void pp_radial(const PPAtom& p, double Z, const double* r, double* vl) {
  rgauss_val(p, r, vl);
  if (p.addzeff) {
    int l = (int)p.n.size() - 1;
    // cutoff of an RGAUSSIAN basis is 1e99 (src/basis_function/Rgaussian_function.cpp:33)
    vl[l] += Z / r[0];
  }
}

// 12/6/4-point spherical quadratures (src/system/gesqua.cpp:32-120).
void gesqua(int& nq, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z,
            std::vector<double>& w) {
  if (nq < 4) fail("AIP must be >= 4");
  if (nq < 6) {
    nq = 4;
    double q = 1. / sqrt(3.);
    x = {q, q, -q, -q};
    y = {q, -q, q, -q};
    z = {q, -q, -q, q};
    w.assign(4, 0.25);
  } else if (nq < 12) {
    nq = 6;
    x = {1, -1, 0, 0, 0, 0};
    y = {0, 0, 1, -1, 0, 0};
    z = {0, 0, 0, 0, 1, -1};
    w.assign(6, 1. / 6.);
  } else if (nq < 18) {
    nq = 12;
    w.assign(12, 1. / 12.);
    x.assign(12, 0);
    y.assign(12, 0);
    z.assign(12, 0);
    double fi0 = PI / 5., cstha = 1. / sqrt(5.0), sntha = 2. / sqrt(5.0);
    z[0] = 1;
    z[1] = -1;
    for (int i = 0; i < 5; i++) {
      double rk2 = 2 * i;
      x[i + 2] = sntha * cos(rk2 * fi0);
      y[i + 2] = sntha * sin(rk2 * fi0);
      z[i + 2] = cstha;
      x[i + 7] = sntha * cos((rk2 + 1) * fi0);
      y[i + 7] = sntha * sin((rk2 + 1) * fi0);
      z[i + 7] = -cstha;
    }
  } else {
    fail("AIP >= 18 not supported");
  }
}

void read_system(System& S, const Tok& words) {
  Tok sys;
  if (!readsection0(words, sys, "SYSTEM")) fail("missing SYSTEM section");
  if (!haskeyword(sys, "MOLECULE")) fail("only SYSTEM { MOLECULE } is supported");
  Tok ns;
  if (!readsection0(sys, ns, "NSPIN") || ns.size() < 2) fail("need NSPIN");
  S.nup = atoi(ns[0].c_str());
  S.ndn = atoi(ns[1].c_str());
  size_t pos = 0;
  Tok at;
  while (readsection(sys, pos, at, "ATOM")) {
    Atom a;
    a.label = at[0];
    a.Z = atof(at[1].c_str());
    size_t k = 0;
    for (; k < at.size(); k++)
      if (ieq(at[k], "COOR")) break;
    if (k + 3 >= at.size()) fail("ATOM needs COOR x y z");
    for (int d = 0; d < 3; d++) a.r[d] = atof(at[k + 1 + d].c_str());
    S.atoms.push_back(a);
  }
  if (S.atoms.empty()) fail("no ATOM sections");
  Tok ef;
  if (readsection0(sys, ef, "ELECTRIC_FIELD")) fail("ELECTRIC_FIELD not supported");

  // PSEUDO (src/system/Pseudopotential.cpp:874-1033)
  S.pp_index.assign(S.atoms.size(), -1);
  pos = 0;
  Tok ps;
  std::vector<Tok> pstexts;
  while (readsection(words, pos, ps, "PSEUDO")) pstexts.push_back(ps);
  for (auto& t : pstexts) {
    PPAtom p;
    Tok b;
    if (haskeyword(t, "SPIN_DEP")) fail("SPIN_DEP pseudopotentials not supported");
    if (!readsection0(t, b, "BASIS")) fail("PSEUDO needs BASIS");
    if (b.size() < 2 || !ieq(b[1], "RGAUSSIAN")) fail("PSEUDO BASIS must be RGAUSSIAN");
    Tok oq;
    if (!readsection0(b, oq, "OLDQMC")) fail("RGAUSSIAN needs OLDQMC");
    size_t it = 0;
    it++;  // zeff (unused by the radial function)
    int nl = atoi(oq[it++].c_str());
    std::vector<int> nper(nl);
    for (int l = 0; l < nl; l++) nper[l] = atoi(oq[it++].c_str());
    p.n.resize(nl);
    p.a.resize(nl);
    p.c.resize(nl);
    for (int l = 0; l < nl; l++)
      for (int j = 0; j < nper[l]; j++) {
        p.n[l].push_back(atof(oq[it++].c_str()) - 2);
        p.a[l].push_back(atof(oq[it++].c_str()));
        p.c[l].push_back(atof(oq[it++].c_str()));
      }
    std::string v;
    if (readvalue(t, v, "AIP")) p.aip = atoi(v.c_str());
    p.addzeff = haskeyword(t, "ADD_ZEFF");
    int aip = p.aip;
    gesqua(aip, p.qx, p.qy, p.qz, p.qw);
    p.aip = aip;
    int idx = (int)S.pp.size();
    bool used = false;
    for (size_t a = 0; a < S.atoms.size(); a++)
      if (S.atoms[a].label == t[0]) {
        if (S.pp_index[a] >= 0) fail("two PSEUDO sections for " + t[0]);
        S.pp_index[a] = idx;
        used = true;
      }
    if (used) S.pp.push_back(p);
  }
  // static cutoff per PP atom type (src/system/Pseudopotential.cpp:984-1031)
  for (size_t a = 0; a < S.atoms.size(); a++) {
    int k = S.pp_index[a];
    if (k < 0) continue;
    PPAtom& p = S.pp[k];
    int L = (int)p.n.size();
    std::vector<double> cutL(L, 20.0), vl(L);
    std::vector<int> found(L, 0);
    for (double r = 20.0; r > 0; r -= .05) {
      double tr[5] = {r, r * r, 0, 0, r};
      pp_radial(p, S.atoms[a].Z, tr, vl.data());
      for (int l = 0; l < L - 1; l++)
        if (fabs(vl[l]) > 1e-5 && !found[l]) {
          cutL[l] = r - .05;
          found[l] = 1;
        }
    }
    double cut = 0;
    for (int l = 0; l < L - 1; l++) cut = std::max(cut, cutL[l]);
    p.cutoff = cut;
  }
}

void read_slater(System& S, const Tok& wf1, const std::string& dir) {
  if (!haskeyword(wf1, "SLATER")) fail("WF1 must be SLATER");
  Tok detwt, states, orbs;
  // DETWT, or CSF { c w_1 w_2 ... } sections: determinant weights c * w_j in STATES order
  // (src/wavefunction/Slat_wf_data.cpp:69-118)
  S.detwt.clear();
  size_t cpos = 0;
  Tok csf;
  int ncsf = 0;
  S.csf_c.clear();
  S.csf_w.clear();
  while (readsection(wf1, cpos, csf, "CSF")) {
    if (csf.size() < 2) fail("wrong number of elements in a CSF");
    S.csf_c.push_back(atof(csf[0].c_str()));
    S.csf_w.emplace_back();
    for (size_t j = 1; j < csf.size(); j++) {
      S.csf_w.back().push_back(atof(csf[j].c_str()));
      S.detwt.push_back(atof(csf[0].c_str()) * atof(csf[j].c_str()));
    }
    ncsf++;
  }
  if (ncsf) {
    if (readsection0(wf1, detwt, "DETWT")) fail("CSF is incompatible with DETWT");
  } else {
    if (!readsection0(wf1, detwt, "DETWT") || detwt.empty()) fail("need DETWT");
    for (auto& x : detwt) {
      S.detwt.push_back(atof(x.c_str()));
      S.csf_c.push_back(S.detwt.back());
      S.csf_w.push_back({1.0});
    }
  }
  S.ndet = (int)S.detwt.size();
  S.optimize_det = haskeyword(wf1, "OPTIMIZE_DET");
  if (!readsection0(wf1, states, "STATES")) fail("need STATES");
  if ((int)states.size() != S.ndet * S.ne()) fail("STATES count must be ndet x the number of electrons");
  if (!readsection0(wf1, orbs, "ORBITALS")) fail("need ORBITALS");
  if (!haskeyword(orbs, "CUTOFF_MO")) fail("only CUTOFF_MO orbitals are supported");
  {
    Tok cw;
    if (readsection0(orbs, cw, "CENTERS") && haskeyword(cw, "READ")) fail("CENTERS { READ } is not supported");
  }
  std::string v;
  if (!readvalue(orbs, v, "NMO")) fail("need NMO");
  int nmo_total = atoi(v.c_str());
  double magnify = 1;
  readdouble(orbs, magnify, "MAGNIFY");
  size_t pos = 0;
  Tok b;
  while (readsection(orbs, pos, b, "BASIS")) {
    if (b.size() < 2 || !ieq(b[1], "AOSPLINE")) fail("only AOSPLINE basis is supported");
    AOSpline ao;
    ao.read(b);
    S.aos.push_back(ao);
  }
  // centers = atoms (USEATOMS; USEGLOBAL is the same set for a molecule)
  S.center_basis.assign(S.atoms.size(), {});
  int maxbasis = 0;
  for (size_t a = 0; a < S.atoms.size(); a++) {
    int nb = 0;
    for (size_t j = 0; j < S.aos.size(); j++)
      if (S.aos[j].label == S.atoms[a].label) {
        S.center_basis[a].push_back((int)j);
        nb += S.aos[j].nfunc;
      }
    maxbasis = std::max(maxbasis, nb);
    S.totbasis += nb;
  }
  for (size_t a = 0; a < S.atoms.size(); a++)
    for (int j : S.center_basis[a])
      for (int f = 0; f < S.aos[j].nfunc; f++) {
        S.fcut.push_back(S.aos[j].rcut[f]);
        S.fobj.push_back(j);
      }
  // ORBFILE (src/orbitals/MO_matrix.h:173-251)
  if (!readvalue(orbs, v, "ORBFILE")) fail("need ORBFILE");
  std::ifstream orb(join_path(dir, v));
  if (!orb) fail("cannot open orbital file " + v);
  std::vector<int> mo, bas, cen, lab;
  std::string w;
  while (true) {
    if (!(orb >> w)) fail("no COEFFICIENTS in orbital file");
    if (w == "COEFFICIENTS") break;
    int m = atoi(w.c_str()) - 1;
    std::string b1, c1, l1;
    orb >> b1 >> c1 >> l1;
    if (m > nmo_total) continue;
    mo.push_back(m);
    bas.push_back(atoi(b1.c_str()) - 1);
    cen.push_back(atoi(c1.c_str()) - 1);
    lab.push_back(atoi(l1.c_str()) - 1);
    if (bas.back() >= maxbasis || bas.back() < 0) fail("basis index out of range in orbital file");
  }
  int nmo_read = *std::max_element(mo.begin(), mo.end()) + 1;
  int maxlabel = *std::max_element(lab.begin(), lab.end()) + 1;
  std::vector<double> coeff(maxlabel);
  for (int i = 0; i < maxlabel; i++)
    if (!(orb >> coeff[i])) fail("orbital file ended early");
  int nat = (int)S.atoms.size();
  std::vector<int> cmat((size_t)nmo_read * nat * maxbasis, -1);
  for (size_t i = 0; i < mo.size(); i++) cmat[((size_t)mo[i] * nat + cen[i]) * maxbasis + bas[i]] = lab[i];
  // coefficient lists per MO with the 1e-12 drop (src/orbitals/MO_matrix_cutoff.h:200-245)
  std::vector<std::vector<double>> full(nmo_total, std::vector<double>(S.totbasis, 0.0));
  int tot = 0;
  for (int ion = 0; ion < nat; ion++) {
    int f = 0;
    for (int j : S.center_basis[ion])
      for (int i = 0; i < S.aos[j].nfunc; i++) {
        for (int m = 0; m < nmo_total; m++) {
          double t = 0;
          if (m < nmo_read) {
            int l = cmat[((size_t)m * nat + ion) * maxbasis + f];
            if (l != -1) t = coeff[l];
          }
          if (fabs(t) > 1e-12) full[m][tot] = magnify * t;
        }
        f++;
        tot++;
      }
  }
  // per-spin MO lists: every orbital a determinant of that spin occupies, in order of first
  // appearance, and each determinant's occupation as indices into the list
  // (src/wavefunction/Slat_wf_data.cpp:192-263, MO_matrix_cutoff::buildLists)
  S.C.assign(2, {});
  S.nmo.assign(2, 0);
  for (int s = 0; s < 2; s++) {
    int n = S.nspin(s);
    std::vector<int> list;
    S.occ[s].assign(S.ndet, std::vector<int>(n));
    for (int d = 0; d < S.ndet; d++) {
      std::vector<int> orig(n);
      for (int i = 0; i < n; i++) {
        orig[i] = atoi(states[(size_t)d * S.ne() + (s == 0 ? 0 : S.nup) + i].c_str()) - 1;
        if (orig[i] < 0 || orig[i] >= nmo_total) fail("STATES index out of range");
        for (int j = 0; j < i; j++)
          if (orig[i] == orig[j]) fail("repeated orbital in one spin channel");
        int place = -1;
        for (size_t q = 0; q < list.size(); q++)
          if (list[q] == orig[i]) place = (int)q;
        if (place < 0) {
          place = (int)list.size();
          list.push_back(orig[i]);
        }
        S.occ[s][d][i] = place;
      }
    }
    int nm = (int)list.size();
    S.nmo[s] = nm;
    S.C[s].assign((size_t)S.totbasis * nm, 0.0);
    for (int f = 0; f < S.totbasis; f++)
      for (int i = 0; i < nm; i++) S.C[s][(size_t)f * nm + i] = full[list[i]][f];
  }
}

// Jastrow2 groups (src/wavefunction/Jastrow2_wf.cpp:29-160,630-660).
void read_jastrow(System& S, const Tok& wf2) {
  if (!haskeyword(wf2, "JASTROW2")) fail("WF2 must be JASTROW2");
  size_t pos = 0;
  Tok g;
  while (readsection(wf2, pos, g, "GROUP")) {
    JGroup G;
    if (haskeyword(g, "THREEBODY") || haskeyword(g, "THREEBODY_SPIN")) fail("three-body Jastrow not supported");
    size_t p2 = 0;
    Tok b;
    while (readsection(g, p2, b, "EIBASIS")) {
      JBasis jb;
      jb.read(b);
      G.eib.push_back(jb);
    }
    p2 = 0;
    while (readsection(g, p2, b, "EEBASIS")) {
      JBasis jb;
      jb.read(b);
      G.eeb.push_back(jb);
    }
    int nat = (int)S.atoms.size();
    G.atom2basis.assign(nat, {});
    for (size_t k = 0; k < G.eib.size(); k++) {
      bool found = false;
      for (int a = 0; a < nat; a++)
        if (G.eib[k].label == S.atoms[a].label) G.atom2basis[a].push_back((int)k), found = true;
      if (!found) fail("no atom for Jastrow basis label " + G.eib[k].label);
    }
    Tok ob;
    if (readsection0(g, ob, "ONEBODY")) {
      G.has_one = true;
      G.onecoef.assign(nat, {});
      double scale = 1;
      bool has_scale = readdouble(ob, scale, "SCALE");
      size_t p3 = 0;
      Tok cs;
      std::vector<bool> got(nat, false);
      G.one_frozen = haskeyword(ob, "FREEZE");
      while (readsection(ob, p3, cs, "COEFFICIENTS")) {
        std::vector<double> c;
        for (size_t i = 1; i < cs.size(); i++) c.push_back(atof(cs[i].c_str()) * (has_scale ? scale : 1.0));
        G.sec_label.push_back(cs[0]);
        G.sec_coef.push_back(c);
        for (int a = 0; a < nat; a++)
          if (S.atoms[a].label == cs[0]) G.onecoef[a] = c, got[a] = true;
      }
      for (int a = 0; a < nat; a++) {
        if (!got[a]) fail("ONEBODY has no coefficients for atom " + S.atoms[a].label);
        int nb = 0;
        for (int k : G.atom2basis[a]) nb += G.eib[k].nfunc();
        if ((int)G.onecoef[a].size() > nb) fail("more one-body coefficients than EI functions");
      }
    }
    Tok tb;
    if (readsection0(g, tb, "TWOBODY")) {
      G.has_two = true;
      Tok cs;
      if (!readsection0(tb, cs, "COEFFICIENTS")) fail("TWOBODY needs COEFFICIENTS");
      for (auto& x : cs) G.twocoef.push_back(atof(x.c_str()));
      G.two_frozen = haskeyword(tb, "FREEZE");
    }
    if (readsection0(g, tb, "TWOBODY_SPIN")) {
      if (G.has_two) fail("only one of TWOBODY or TWOBODY_SPIN");
      G.has_two = true;
      G.diffspin = true;
      Tok l1, l2;
      if (!readsection0(tb, l1, "LIKE_COEFFICIENTS") || !readsection0(tb, l2, "UNLIKE_COEFFICIENTS"))
        fail("TWOBODY_SPIN needs LIKE_/UNLIKE_COEFFICIENTS");
      if (l1.size() != l2.size()) fail("like/unlike coefficient counts differ");
      for (auto& x : l1) G.like.push_back(atof(x.c_str()));
      for (auto& x : l2) G.unlike.push_back(atof(x.c_str()));
      G.two_frozen = haskeyword(tb, "FREEZE");
    }
    if (G.has_two) {
      int nb = 0;
      for (auto& e : G.eeb) nb += e.nfunc();
      int np = G.diffspin ? (int)G.like.size() : (int)G.twocoef.size();
      if (np > nb) fail("not enough EE basis functions for the two-body coefficients");
    }
    S.groups.push_back(G);
  }
}

// ---------------------------------------------------------------- evaluation pieces
inline void dist5(const double* a, const double* b, double* R) {
  R[1] = 0;
  for (int d = 0; d < 3; d++) {
    R[d + 2] = a[d] - b[d];
    R[1] += R[d + 2] * R[d + 2];
  }
  R[0] = sqrt(R[1]);
}

// MO values of spin s at point p, out[5][nmo] (src/orbitals/MO_matrix_cutoff.h:480-560).
void mo_lap(const System& S, int s, const double* p, double* out, std::vector<double>& tmp) {
  int n = S.nmo[s];
  std::fill(out, out + 5 * n, 0.0);
  int tot = 0;
  const std::vector<double>& C = S.C[s];
  for (size_t a = 0; a < S.atoms.size(); a++) {
    double R[5];
    dist5(p, S.atoms[a].r, R);
    for (int j : S.center_basis[a]) {
      const AOSpline& ao = S.aos[j];
      if (R[0] < ao.threshold) {
        tmp.resize(5 * ao.nfunc);
        ao.calc_lap(R, tmp.data());
        for (int i = 0; i < ao.nfunc; i++) {
          if (R[0] < S.fcut[tot]) {
            const double* crow = &C[(size_t)tot * n];
            const double* sv = &tmp[5 * i];
            for (int m = 0; m < n; m++) {
              double c = crow[m];
              if (c == 0.0) continue;
              for (int d = 0; d < 5; d++) out[d * n + m] += c * sv[d];
            }
          }
          tot++;
        }
      } else {
        tot += ao.nfunc;
      }
    }
  }
}

// Value-only MO evaluation (src/orbitals/MO_matrix_cutoff.h:413-476).
void mo_val(const System& S, int s, const double* p, double* out, std::vector<double>& tmp) {
  int n = S.nmo[s];
  std::fill(out, out + n, 0.0);
  int tot = 0;
  const std::vector<double>& C = S.C[s];
  for (size_t a = 0; a < S.atoms.size(); a++) {
    double R[5];
    dist5(p, S.atoms[a].r, R);
    for (int j : S.center_basis[a]) {
      const AOSpline& ao = S.aos[j];
      if (ao.threshold > R[0]) {
        tmp.resize(5 * ao.nfunc);
        ao.calc_lap(R, tmp.data());
        for (int i = 0; i < ao.nfunc; i++) {
          if (R[0] < S.fcut[tot]) {
            const double* crow = &C[(size_t)tot * n];
            for (int m = 0; m < n; m++)
              if (crow[m] != 0.0) out[m] += crow[m] * tmp[5 * i];
          }
          tot++;
        }
      } else {
        tot += ao.nfunc;
      }
    }
  }
}

// LU with partial pivoting and implicit row scaling, Crout order
// (src/program_utils/MatrixAlgebrac.cpp:108-190).  a is n x n row-major, overwritten.
int ludcmp(std::vector<double>& a, int n, std::vector<int>& indx) {
  std::vector<double> vv(n);
  for (int i = 0; i < n; i++) {
    double big = 0;
    for (int j = 0; j < n; j++) big = std::max(big, fabs(a[i * n + j]));
    if (big == 0.0) return 0;
    vv[i] = 1.0 / big;
  }
  for (int j = 0; j < n; j++) {
    int imax = j;
    for (int i = 0; i < j; i++) {
      double sum = a[i * n + j];
      for (int k = 0; k < i; k++) sum -= a[i * n + k] * a[k * n + j];
      a[i * n + j] = sum;
    }
    double big = 0.0;
    for (int i = j; i < n; i++) {
      double sum = a[i * n + j];
      for (int k = 0; k < j; k++) sum -= a[i * n + k] * a[k * n + j];
      a[i * n + j] = sum;
      double dum = vv[i] * fabs(sum);
      if (dum >= big) {
        big = dum;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 0; k < n; k++) std::swap(a[imax * n + k], a[j * n + k]);
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (a[j * n + j] == 0.0) return 0;
    if (j != n - 1) {
      double dum = 1.0 / a[j * n + j];
      for (int i = j + 1; i < n; i++) a[i * n + j] *= dum;
    }
  }
  return 1;
}
// src/program_utils/MatrixAlgebrac.cpp:192-224
void lubksb(const std::vector<double>& a, int n, const std::vector<int>& indx, double* b) {
  int ii = -1;
  for (int i = 0; i < n; i++) {
    int ip = indx[i];
    double sum = b[ip];
    b[ip] = b[i];
    if (ii >= 0)
      for (int j = ii; j < i; j++) sum -= a[i * n + j] * b[j];
    else if (sum != 0.0)
      ii = i;
    b[i] = sum;
  }
  for (int i = n - 1; i >= 0; i--) {
    double sum = b[i];
    for (int j = i + 1; j < n; j++) sum -= a[i * n + j] * b[j];
    b[i] = sum / a[i * n + i];
  }
}
// Transposed inverse and log-determinant (src/program_utils/MatrixAlgebrac.cpp:316-421).
// Returns 0 for a singular matrix (log value -1e201, sign +1 as the reference's log_value(0)).
int transpose_inverse(const double* a, double* a1, int n, double& logdet, double& sign) {
  std::vector<double> t(a, a + n * n);
  std::vector<int> indx(n);
  for (int i = 0; i < n * n; i++) a1[i] = 0;
  for (int i = 0; i < n; i++) a1[i * n + i] = 1;
  if (!ludcmp(t, n, indx)) {
    logdet = -1e201;
    sign = 1;
    return 0;
  }
  for (int j = 0; j < n; j++) lubksb(t, n, indx, a1 + j * n);
  logdet = 0;
  sign = 1;
  for (int i = 0; i < n; i++) {
    if (indx[i] != i) sign = -sign;
    logdet += log(fabs(t[i * n + i]));
    if (t[i * n + i] < 0) sign = -sign;
  }
  return 1;
}
// Sherman-Morrison for a replaced row of M, on the transposed inverse
// (src/program_utils/MatrixAlgebrac.cpp:766-838).  Returns det_old/det_new.
double inverse_update_row(double* a1, const double* v, int l, int n) {
  std::vector<double> tl(n), prod(n);
  double f = 0;
  for (int i = 0; i < n; i++) f += a1[l * n + i] * v[i];
  f = -1.0 / f;
  for (int j = 0; j < n; j++) {
    tl[j] = a1[l * n + j];
    double p = 0;
    for (int i = 0; i < n; i++) p += a1[j * n + i] * v[i];
    prod[j] = p * f;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) a1[i * n + j] += tl[j] * prod[i];
  f = -f;
  for (int j = 0; j < n; j++) a1[l * n + j] = f * tl[j];
  return f;
}

// ---------------------------------------------------------------- Jastrow evaluation
// One-body and two-body terms touching electron e at position p (others at pos).
// Returns the sum of pair/one-body values involving e; grad (3) and lap accumulate
// d/dr_e (src/wavefunction/Jastrow2_wf.cpp:229-330,988-1157; Jastrow2_one.cpp:137-194;
// Jastrow2_two.cpp:83-133,294-348).
struct JScratch {
  std::vector<double> b;
};
double jastrow_e(const System& S, int e, const double* p, const double* pos, double* grad, double* lap,
                 JScratch& js) {
  double u = 0, g[3] = {0, 0, 0}, l = 0;
  int ne = S.ne();
  for (const JGroup& G : S.groups) {
    if (G.has_one) {
      for (size_t a = 0; a < S.atoms.size(); a++) {
        double R[5];
        dist5(p, S.atoms[a].r, R);
        const std::vector<double>& cf = G.onecoef[a];
        int k = 0;
        for (int bi : G.atom2basis[a]) {
          const JBasis& jb = G.eib[bi];
          js.b.resize(5 * jb.nfunc());
          jb.calc_lap(R, js.b.data());
          for (int i = 0; i < jb.nfunc(); i++, k++) {
            if (k >= (int)cf.size()) continue;
            u += cf[k] * js.b[5 * i];
            for (int d = 0; d < 3; d++) g[d] += cf[k] * js.b[5 * i + 1 + d];
            l += cf[k] * js.b[5 * i + 4];
          }
        }
      }
    }
    if (G.has_two) {
      for (int j = 0; j < ne; j++) {
        if (j == e) continue;
        double R[5];
        dist5(p, pos + 3 * j, R);
        int s = (j < S.nup) != (e < S.nup);
        const std::vector<double>& cf = G.diffspin ? (s ? G.unlike : G.like) : G.twocoef;
        int k = 0;
        for (const JBasis& jb : G.eeb) {
          js.b.resize(5 * jb.nfunc());
          if (R[0] < jb.cutoff()) jb.calc_lap(R, js.b.data());
          else std::fill(js.b.begin(), js.b.end(), 0.0);
          for (int i = 0; i < jb.nfunc(); i++, k++) {
            if (k >= (int)cf.size()) continue;
            u += cf[k] * js.b[5 * i];
            for (int d = 0; d < 3; d++) g[d] += cf[k] * js.b[5 * i + 1 + d];
            l += cf[k] * js.b[5 * i + 4];
          }
        }
      }
    }
  }
  if (grad)
    for (int d = 0; d < 3; d++) grad[d] = g[d];
  if (lap) *lap = l;
  return u;
}

// Total Jastrow U = sum_i onebody(i) + sum_{i<j} twobody(i,j).
double jastrow_total(const System& S, const double* pos, JScratch& js) {
  int ne = S.ne();
  double u1 = 0, u2 = 0;
  for (int e = 0; e < ne; e++) {
    // one-body part of e, and pairs (e,j) with j>e
    for (const JGroup& G : S.groups) {
      if (G.has_one) {
        for (size_t a = 0; a < S.atoms.size(); a++) {
          double R[5];
          dist5(pos + 3 * e, S.atoms[a].r, R);
          const std::vector<double>& cf = G.onecoef[a];
          int k = 0;
          for (int bi : G.atom2basis[a]) {
            const JBasis& jb = G.eib[bi];
            js.b.resize(5 * jb.nfunc());
            jb.calc_lap(R, js.b.data());
            for (int i = 0; i < jb.nfunc(); i++, k++)
              if (k < (int)cf.size()) u1 += cf[k] * js.b[5 * i];
          }
        }
      }
      if (G.has_two) {
        for (int j = e + 1; j < ne; j++) {
          double R[5];
          dist5(pos + 3 * e, pos + 3 * j, R);
          int s = (j < S.nup) != (e < S.nup);
          const std::vector<double>& cf = G.diffspin ? (s ? G.unlike : G.like) : G.twocoef;
          int k = 0;
          for (const JBasis& jb : G.eeb) {
            js.b.resize(5 * jb.nfunc());
            if (R[0] < jb.cutoff()) jb.calc_lap(R, js.b.data());
            else std::fill(js.b.begin(), js.b.end(), 0.0);
            for (int i = 0; i < jb.nfunc(); i++, k++)
              if (k < (int)cf.size()) u2 += cf[k] * js.b[5 * i];
          }
        }
      }
    }
  }
  return u1 + u2;
}

// Coulomb local potential (src/system/Molecular_system.cpp:209-284).
double coulomb(const System& S, const double* pos) {
  int ne = S.ne(), na = (int)S.atoms.size();
  double ei = 0;
  for (int e = 0; e < ne; e++)
    for (int a = 0; a < na; a++) {
      double R[5];
      dist5(pos + 3 * e, S.atoms[a].r, R);
      ei += S.atoms[a].Z / R[0];
    }
  double pot = -ei;
  double ii = 0;
  for (int i = 0; i < na; i++)
    for (int j = 0; j < i; j++) {
      double R[5];
      dist5(S.atoms[i].r, S.atoms[j].r, R);
      ii += S.atoms[i].Z * S.atoms[j].Z / R[0];
    }
  pot += ii;
  double ee = 0;
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < i; j++) {
      double R[5];
      dist5(pos + 3 * j, pos + 3 * i, R);
      ee += 1 / R[0];
    }
  return pot + ee;
}

double legendre(double x, int n) {
  switch (n) {
    case 0: return 1;
    case 1: return x;
    case 2: return .5 * (3 * x * x - 1);
    case 3: return .5 * (5 * x * x * x - 3 * x);
    case 4: return 0.125 * (35 * x * x * x * x - 30 * x * x + 3);
    case 5: return 0.125 * (63 * x * x * x * x * x - 70 * x * x * x + 15 * x);
  }
  fail("legendre order too high");
}

// Walker state for evaluation and moves.
struct Walker {
  std::vector<double> pos;              // ne*3
  std::vector<double> mo;               // [ne][5][nmo_s]
  std::vector<std::vector<double>> inv[2];  // [det][n][n] transposed inverse per determinant
  std::vector<double> logdet[2], sign[2];   // [det] log|det| and sign per spin
};

void walker_init(const System& S, Walker& w) {
  w.mo.assign((size_t)S.ne() * 5 * std::max(S.nmo[0], S.nmo[1]), 0.0);
  for (int s = 0; s < 2; s++) {
    w.inv[s].assign(S.ndet, std::vector<double>((size_t)S.nspin(s) * S.nspin(s), 0.0));
    w.logdet[s].assign(S.ndet, 0.0);
    w.sign[s].assign(S.ndet, 1.0);
  }
}

// Relative determinant weights b_d = sign_d sign_0 exp(l_d - l_0) of the terms
// c_d D_up,d D_dn,d (log l_d, sign sign_d), pivoting on the first term as the reference's
// sum of log values does (src/program_utils/MatrixAlgebra.h:92-99); the total is
// sign_0 exp(l_0) sum_d b_d.  One determinant: b_0 = 1 exactly.
void det_terms(const System& S, const Walker& w, std::vector<double>& b, double& l0, double& s0) {
  b.resize(S.ndet);
  auto term = [&](int d, double& l, double& sg) {
    l = log(fabs(S.detwt[d])) + w.logdet[0][d] + w.logdet[1][d];
    sg = (S.detwt[d] < 0 ? -1.0 : 1.0) * w.sign[0][d] * w.sign[1][d];
  };
  term(0, l0, s0);
  b[0] = 1.0;
  for (int d = 1; d < S.ndet; d++) {
    double l, sg;
    term(d, l, sg);
    b[d] = sg * s0 * exp(l - l0);
  }
}
inline double* morow(const System& S, Walker& w, int e) {
  return &w.mo[(size_t)e * 5 * std::max(S.nmo[0], S.nmo[1])];
}

// Full recompute: MO rows for all electrons, LU per spin (src/wavefunction/Slat_wf.h:948-1016).
void walker_refresh(const System& S, Walker& w, std::vector<double>& tmp) {
  for (int e = 0; e < S.ne(); e++) mo_lap(S, S.spin_of(e), &w.pos[3 * e], morow(S, w, e), tmp);
  for (int s = 0; s < 2; s++) {
    int n = S.nspin(s);
    for (int d = 0; d < S.ndet; d++) {
      if (n == 0) {
        w.logdet[s][d] = 0;
        w.sign[s][d] = 1;
        continue;
      }
      const std::vector<int>& oc = S.occ[s][d];
      std::vector<double> m((size_t)n * n);
      for (int k = 0; k < n; k++) {
        double* r = morow(S, w, s * S.nup + k);
        for (int i = 0; i < n; i++) m[k * n + i] = r[oc[i]];  // value row, component 0
      }
      transpose_inverse(m.data(), w.inv[s][d].data(), n, w.logdet[s][d], w.sign[s][d]);
    }
  }
}

// sum_d b_d sum_j inv_d(k, j) v[occ_d(j)] / sum_d b_d for the row k of spin s: the
// determinant part of Psi(r')/Psi (v = MO values at r') or of grad Psi / Psi and
// lap Psi / Psi (v = MO derivatives at the electron) (src/wavefunction/Slat_wf.h:1021-1139).
double det_contract(const System& S, const Walker& w, int s, int k, const double* v) {
  const int n = S.nspin(s);
  std::vector<double> b;
  double l0, s0;
  det_terms(S, w, b, l0, s0);
  double num = 0, den = 0;
  for (int d = 0; d < S.ndet; d++) {
    const double* iv = &w.inv[s][d][(size_t)k * n];
    const std::vector<int>& oc = S.occ[s][d];
    double t = 0;
    for (int j = 0; j < n; j++) t += v[oc[j]] * iv[j];
    num += b[d] * t;
    den += b[d];
  }
  return num / den;
}

// det part of grad/lap for electron e from the cache (src/wavefunction/Slat_wf.h:1021-1139)
void det_gradlap(const System& S, Walker& w, int e, double* g, double* l) {
  int s = S.spin_of(e), k = e - s * S.nup, nm = S.nmo[s];
  const double* r = morow(S, w, e);
  double t[5] = {0, 0, 0, 0, 0};
  for (int d = 1; d < 5; d++) t[d] = det_contract(S, w, s, k, r + d * nm);
  for (int d = 0; d < 3; d++) g[d] = t[d + 1];
  *l = t[4];
}

struct EvalOut {
  double logpsi, sign, kin, vloc, vnl;
  std::vector<double> grad, lap;
};

// A T-move candidate (src/system/Pseudopotential.h:60-65): electron e displaced by disp onto
// a quadrature point, with its negative nonlocal matrix element vxx.
struct Tmove {
  int e;
  double vxx;
  double disp[3];
};

// Nonlocal pseudopotential for one walker (src/system/Pseudopotential.cpp:596-815).
// qrot: rotated quadrature points per PP atom type; rnd: uniforms [natoms][ne] or NULL.
// With tmoves set (Tmoves::negative_tmove, src/system/Pseudopotential.cpp:752-763), terms with
// vxx < 0 are left out of the sum and appended to *tmoves in the order they are met (atom,
// electron, quadrature point).
double pp_nonlocal(const System& S, Walker& w, const double* rot, int stochastic, const double* rnd,
                   std::vector<double>& tmp, JScratch& js, std::vector<Tmove>* tmoves = nullptr) {
  int ne = S.ne(), na = (int)S.atoms.size();
  double tot = 0;
  std::vector<double> phi;
  for (int a = 0; a < na; a++) {
    int pk = S.pp_index[a];
    if (pk < 0) continue;
    const PPAtom& P = S.pp[pk];
    int L = (int)P.n.size();
    std::vector<double> vl(L), qx(P.aip), qy(P.aip), qz(P.aip);
    for (int i = 0; i < P.aip; i++) {
      double o[3] = {P.qx[i], P.qy[i], P.qz[i]};
      // rotateQuadrature with axes x,y,z = rows of rot (src/system/Pseudopotential.cpp:828-864)
      if (rot) {
        qx[i] = o[0] * rot[0] + o[1] * rot[3] + o[2] * rot[6];
        qy[i] = o[0] * rot[1] + o[1] * rot[4] + o[2] * rot[7];
        qz[i] = o[0] * rot[2] + o[1] * rot[5] + o[2] * rot[8];
      } else {
        qx[i] = o[0];
        qy[i] = o[1];
        qz[i] = o[2];
      }
    }
    for (int e = 0; e < ne; e++) {
      double old[5];
      dist5(&w.pos[3 * e], S.atoms[a].r, old);
      pp_radial(P, S.atoms[a].Z, old, vl.data());
      bool accept;
      if (!stochastic) {
        accept = old[0] < P.cutoff;
      } else {
        double strength = 0;
        for (int l = 0; l < L - 1; l++) strength += 10.0 * (2 * l + 1) * fabs(vl[l]);
        strength = std::min(1.0, strength);
        double rr = rnd[a * ne + e];
        if (fabs(strength) > 0)
          for (int l = 0; l < L - 1; l++) vl[l] /= strength;
        accept = strength > rr;
      }
      double nonlocal = 0;
      if (accept) {
        int s = S.spin_of(e), n = S.nspin(s), k = e - s * S.nup;
        double jg[3];
        double u_old = jastrow_e(S, e, &w.pos[3 * e], w.pos.data(), jg, nullptr, js);
        phi.resize(S.nmo[s]);
        for (int i = 0; i < P.aip; i++) {
          double q[3] = {qx[i], qy[i], qz[i]}, np[3], nd[5];
          for (int d = 0; d < 3; d++) np[d] = w.pos[3 * e + d] + (q[d] * old[0] - old[d + 2]);
          dist5(np, S.atoms[a].r, nd);
          double rdr = 0;
          for (int d = 0; d < 3; d++) rdr += nd[d + 2] * old[d + 2];
          rdr /= (nd[0] * old[0]);
          mo_val(S, s, np, phi.data(), tmp);
          const double R = det_contract(S, w, s, k, phi.data());
          double u_new = jastrow_e(S, e, np, w.pos.data(), nullptr, nullptr, js);
          double ratio = R * exp(u_new - u_old) * P.qw[i];
          double ts = 0;
          for (int l = 0; l < L - 1; l++) ts += (2 * l + 1) * vl[l] * legendre(rdr, l);
          const double vxx = ts * ratio;
          if (tmoves && vxx < 0.0) {
            Tmove t{e, vxx, {q[0] * old[0] - old[2], q[1] * old[0] - old[3], q[2] * old[0] - old[4]}};
            tmoves->push_back(t);
          } else {
            nonlocal += vxx;
          }
        }
      }
      tot += vl[L - 1] + nonlocal;
    }
  }
  return tot;
}

// Kinetic energy from the caches plus a fresh O(N^2) Jastrow
// (src/system/System.cpp:53-140, src/wavefunction/Slat_Jastrow.cpp:314-358).
double kinetic(const System& S, Walker& w, double* grad, double* lap, JScratch& js) {
  double ke = 0;
  for (int e = 0; e < S.ne(); e++) {
    double gd[3], ld, gj[3], lj;
    det_gradlap(S, w, e, gd, &ld);
    jastrow_e(S, e, &w.pos[3 * e], w.pos.data(), gj, &lj, js);
    double dot = 0, jj = 0;
    for (int d = 0; d < 3; d++) {
      dot += gd[d] * gj[d];
      jj += gj[d] * gj[d];
    }
    double L = ld + (lj + jj) + 2 * dot;
    if (grad)
      for (int d = 0; d < 3; d++) grad[3 * e + d] = gd[d] + gj[d];
    if (lap) lap[e] = L;
    ke += -0.5 * L;
  }
  return ke;
}

// ---------------------------------------------------------------- Philox4x32-10
// Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11.
inline void philox(const unsigned* ctr, const unsigned* key, unsigned* out) {
  unsigned c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
    unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
    unsigned h0 = (unsigned)(p0 >> 32), l0 = (unsigned)p0;
    unsigned h1 = (unsigned)(p1 >> 32), l1 = (unsigned)p1;
    unsigned n0 = h1 ^ c1 ^ k0, n1 = l1, n2 = h0 ^ c3 ^ k1, n3 = l0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
// RNG streams of the walker engine: key = seed, counter = (walker, step, index, tag).
// The sweep number within a step goes into bits 8.. of the tag.
enum { TAG_GAUSS1 = 0, TAG_ACC1 = 1, TAG_GAUSS2 = 2, TAG_ACC2 = 3, TAG_PP = 4, TAG_ROT = 5, TAG_BRANCH = 6, TAG_TMOVE = 7 };
inline double u53(unsigned a, unsigned b) {
  return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0);
}
inline double u32o(unsigned a) { return (a + 0.5) * (1.0 / 4294967296.0); }
void rng_block(unsigned long long seed, unsigned walker, unsigned step, unsigned idx, unsigned tag, unsigned* out) {
  unsigned ctr[4] = {walker, step, idx, tag}, key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
  philox(ctr, key, out);
}
void gauss3(unsigned long long seed, unsigned walker, unsigned step, unsigned idx, unsigned tag, double* g) {
  unsigned o[4];
  rng_block(seed, walker, step, idx, tag, o);
  double r0 = sqrt(-2.0 * log(u32o(o[0]))), t0 = 2.0 * PI * u32o(o[1]);
  double r1 = sqrt(-2.0 * log(u32o(o[2]))), t1 = 2.0 * PI * u32o(o[3]);
  g[0] = r0 * cos(t0);
  g[1] = r0 * sin(t0);
  g[2] = r1 * cos(t1);
}
double uniform(unsigned long long seed, unsigned walker, unsigned step, unsigned idx, unsigned tag) {
  unsigned o[4];
  rng_block(seed, walker, step, idx, tag, o);
  return u53(o[0], o[1]);
}
// Uniform random rotation from a unit quaternion (Shoemake, Graphics Gems III, 1992).
void random_rotation(unsigned long long seed, unsigned walker, unsigned step, double* rot) {
  unsigned o[4];
  rng_block(seed, walker, step, 0, TAG_ROT, o);
  double u1 = u53(o[0], o[1]), u2 = u32o(o[2]), u3 = u32o(o[3]);
  double a = sqrt(1 - u1), b = sqrt(u1);
  double qw = a * sin(2 * PI * u2), qx = a * cos(2 * PI * u2), qy = b * sin(2 * PI * u3), qz = b * cos(2 * PI * u3);
  rot[0] = 1 - 2 * (qy * qy + qz * qz);
  rot[1] = 2 * (qx * qy - qz * qw);
  rot[2] = 2 * (qx * qz + qy * qw);
  rot[3] = 2 * (qx * qy + qz * qw);
  rot[4] = 1 - 2 * (qx * qx + qz * qz);
  rot[5] = 2 * (qy * qz - qx * qw);
  rot[6] = 2 * (qx * qz - qy * qw);
  rot[7] = 2 * (qy * qz + qx * qw);
  rot[8] = 1 - 2 * (qx * qx + qy * qy);
}

// Cyrus drift limiter (src/method/Split_sample.cpp:40-83).
void lim_drift(const double* v, double tau, double* out) {
  double d2 = 0;
  for (int d = 0; d < 3; d++) d2 += v[d] * v[d];
  d2 *= 0.25;
  double da = d2 > 1e-8 ? (sqrt(1. + 2 * tau * d2) - 1) / d2 : tau;
  for (int d = 0; d < 3; d++) out[d] = v[d] * da;
}
// log Green's function exponent from a to b (src/method/Split_sample.cpp:91-110).
double trans(const double* ra, const double* drifta, const double* rb, double tau) {
  double dr[3];
  lim_drift(drifta, tau, dr);
  double p = 0;
  for (int d = 0; d < 3; d++) {
    double t = rb[d] - ra[d] - dr[d];
    p -= t * t;
  }
  return p / (2.0 * tau);
}

struct Point {
  double r[3];
  double drift[3];
  double lnpsi;  // log|Psi| relative to the start point
  double sgn;    // sign of Psi relative to the start point
  double R;      // determinant ratio
  std::vector<double> phi;  // [5][nmo]
};

// Evaluate trial point: MO row, det ratio, det gradient, Jastrow of e.
void eval_point(const System& S, Walker& w, int e, Point& P, double u_old, std::vector<double>& tmp,
                JScratch& js) {
  int s = S.spin_of(e), n = S.nspin(s), k = e - s * S.nup, nm = S.nmo[s];
  P.phi.resize(5 * nm);
  mo_lap(S, s, P.r, P.phi.data(), tmp);
  double t[4];
  for (int d = 0; d < 4; d++) t[d] = det_contract(S, w, s, k, P.phi.data() + d * nm);
  (void)n;
  P.R = t[0];
  double gj[3];
  double u = jastrow_e(S, e, P.r, w.pos.data(), gj, nullptr, js);
  for (int d = 0; d < 3; d++) P.drift[d] = t[d + 1] / t[0] + gj[d];
  P.lnpsi = log(fabs(t[0])) + (u - u_old);
  P.sgn = t[0] < 0 ? -1 : 1;
}

// Accept a move: Sherman-Morrison, cache row, position, log-det.
void accept_point(const System& S, Walker& w, int e, const Point& P) {
  int s = S.spin_of(e), n = S.nspin(s), k = e - s * S.nup, nm = S.nmo[s];
  std::vector<double> v(n);
  for (int d = 0; d < S.ndet; d++) {
    for (int j = 0; j < n; j++) v[j] = P.phi[S.occ[s][d][j]];
    double f = inverse_update_row(w.inv[s][d].data(), v.data(), k, n);
    double ratio = 1. / f;
    w.logdet[s][d] += log(fabs(ratio));
    if (ratio < 0) w.sign[s][d] = -w.sign[s][d];
  }
  double* row = morow(S, w, e);
  for (int d = 0; d < 5; d++)
    for (int j = 0; j < nm; j++) row[d * nm + j] = P.phi[d * nm + j];
  for (int d = 0; d < 3; d++) w.pos[3 * e + d] = P.r[d];
}

// One electron move with the Split sampler (src/method/Split_sample.cpp:302-540):
// depth 1 with node restriction for DMC, depth 2 with timesteps tau, tau/2 for VMC.
int move_electron(const System& S, Walker& w, int e, double tau, int depth, int restrict_nodes,
                  unsigned long long seed, unsigned wid, unsigned step, unsigned sweep,
                  std::vector<double>& tmp, JScratch& js, Point* pts) {
  Point& P0 = pts[0];
  for (int d = 0; d < 3; d++) P0.r[d] = w.pos[3 * e + d];
  double gd[3], ld, gj[3];
  det_gradlap(S, w, e, gd, &ld);
  double u_old = jastrow_e(S, e, P0.r, w.pos.data(), gj, nullptr, js);
  for (int d = 0; d < 3; d++) P0.drift[d] = gd[d] + gj[d];
  P0.lnpsi = 0;
  P0.sgn = 1;
  double taus[3] = {tau, tau, tau / 2.0};
  auto ratio_sq = [&](const Point& a, const Point& b, bool& neg) {
    // (Psi_b/Psi_a)^2 and the sign of the ratio
    double r = exp(b.lnpsi - a.lnpsi);
    neg = (a.sgn * b.sgn) < 0;
    return r * r;
  };
  auto acc1 = [&](const Point& x, const Point& y, double t) {
    bool neg;
    double r2 = ratio_sq(x, y, neg);
    if (restrict_nodes && neg) return 0.0;
    double a = r2 * exp(trans(y.r, y.drift, x.r, t) - trans(x.r, x.drift, y.r, t));
    return std::min(1.0, a);
  };
  for (int dep = 1; dep <= depth; dep++) {
    double g[3], dr[3];
    gauss3(seed, wid, step, (unsigned)e, (dep == 1 ? TAG_GAUSS1 : TAG_GAUSS2) | (sweep << 8), g);
    lim_drift(P0.drift, taus[dep], dr);
    Point& Pn = pts[dep];
    for (int d = 0; d < 3; d++) Pn.r[d] = P0.r[d] + (g[d] * sqrt(taus[dep]) + dr[d]);
    eval_point(S, w, e, Pn, u_old, tmp, js);
    double acc;
    if (dep == 1) {
      acc = acc1(P0, Pn, taus[1]);
    } else {
      // delayed rejection (src/method/Split_sample.cpp:302-373) for x=0, y=2
      bool neg;
      double r2 = ratio_sq(P0, Pn, neg);
      if (restrict_nodes && neg) {
        acc = 0;
      } else {
        double pt = (trans(Pn.r, Pn.drift, pts[1].r, taus[1]) - trans(P0.r, P0.drift, pts[1].r, taus[1])) +
                    (trans(Pn.r, Pn.drift, P0.r, taus[2]) - trans(P0.r, P0.drift, Pn.r, taus[2]));
        double num = 1 - acc1(Pn, pts[1], taus[1]);
        double den = 1 - acc1(P0, pts[1], taus[1]);
        if (fabs(den) < 1e-10) acc = fabs(num) < 1e-10 ? 0 : 1;
        else acc = std::min(1.0, r2 * exp(pt) * num / den);
      }
    }
    double u = uniform(seed, wid, step, (unsigned)e, (dep == 1 ? TAG_ACC1 : TAG_ACC2) | (sweep << 8));
    if (acc + u > 1.0) {
      accept_point(S, w, e, Pn);
      return dep;
    }
  }
  return 0;
}

struct Energy {
  double ke, vloc, vnl;
  double total() const { return ke + vloc + vnl; }
};

// With tmoves set, the nonlocal part is evaluated as Pseudopotential::calcNonlocTmove does
// (negative terms collected into *tmoves) and then restored as Dmc_method::doTmove does,
// nonlocal -= (sum - 1) / tau with sum = 1 - tau * sum(vxx) (src/method/Dmc_method.cpp:850-866).
Energy local_energy(const System& S, Walker& w, unsigned long long seed, unsigned wid, unsigned step,
                    int pp_stochastic, std::vector<double>& tmp, JScratch& js, std::vector<Tmove>* tmoves = nullptr,
                    double tau = 0.0) {
  Energy E;
  E.ke = kinetic(S, w, nullptr, nullptr, js);
  E.vloc = coulomb(S, w.pos.data());
  E.vnl = 0;
  if (!S.pp.empty()) {
    double rot[9];
    random_rotation(seed, wid, step, rot);
    std::vector<double> rnd;
    if (pp_stochastic) {
      int na = (int)S.atoms.size(), ne = S.ne();
      rnd.resize((size_t)na * ne);
      for (int a = 0; a < na; a++)
        for (int e = 0; e < ne; e++) rnd[a * ne + e] = uniform(seed, wid, step, (unsigned)(a * ne + e), TAG_PP);
    }
    E.vnl = pp_nonlocal(S, w, rot, pp_stochastic, pp_stochastic ? rnd.data() : nullptr, tmp, js, tmoves);
    if (tmoves) {
      double sum = 1;
      for (const Tmove& m : *tmoves) sum -= tau * m.vxx;
      E.vnl -= (sum - 1) / tau;
    }
  }
  return E;
}

// Casula T-moves after the local energy (src/method/Dmc_method.cpp:867-905).  sc = 0 (TMOVES):
// one draw scaled by sum = 1 - tau * sum(vxx) chooses at most one candidate of the whole list,
// in list order; sc = 1 (TMOVESSC): the same per electron over its own candidates.  A chosen
// electron is translated by the candidate's displacement and the walker's inverse, MO cache and
// log|det| are updated as for an accepted move.  Uniforms: TAG_TMOVE, index 0 (TMOVES) or the
// electron (TMOVESSC).
int do_tmove(const System& S, Walker& w, const std::vector<Tmove>& tm, double tau, int sc, unsigned long long seed,
             unsigned wid, unsigned step, std::vector<double>& tmp, JScratch& js, Point& P) {
  auto move = [&](const Tmove& m) {
    for (int d = 0; d < 3; d++) P.r[d] = w.pos[3 * m.e + d] + m.disp[d];
    eval_point(S, w, m.e, P, 0.0, tmp, js);
    accept_point(S, w, m.e, P);
  };
  int moved = 0;
  if (!sc) {
    double sum = 1;
    for (const Tmove& m : tm) sum -= tau * m.vxx;
    const double rand = uniform(seed, wid, step, 0, TAG_TMOVE) * sum;
    sum = 1;
    if (rand > sum) {
      for (const Tmove& m : tm) {
        sum -= tau * m.vxx;
        if (rand < sum) {
          move(m);
          moved++;
          break;
        }
      }
    }
    return moved;
  }
  for (int e = 0; e < S.ne(); e++) {
    double sum_e = 1;
    for (const Tmove& m : tm)
      if (m.e == e) sum_e -= tau * m.vxx;
    const double rand = uniform(seed, wid, step, (unsigned)e, TAG_TMOVE) * sum_e;
    double sel = 1;
    if (rand > sel) {
      for (const Tmove& m : tm) {
        if (m.e != e) continue;
        sel -= tau * m.vxx;
        if (rand < sel) {
          move(m);
          moved++;
          break;
        }
      }
    }
  }
  return moved;
}

// DMC weight factor (src/method/Dmc_method.cpp:801-824).
double dmc_weight(double teff, double etr, double enew, double eold, double cut_start, double cut_stop) {
  double teffac = teff / 2.0;
  double fbet = std::max(etr - enew, etr - eold);
  if (fbet > cut_stop) teffac = 0;
  else if (fbet > cut_start) teffac = teffac * (1. - (fbet - cut_start) / (cut_stop - cut_start));
  return exp(teffac * (etr * 2 - enew - eold));
}

template <class F>
void parallel_for(int n, int nthreads, F f) {
  if (nthreads <= 1 || n < 2) {
    for (int i = 0; i < n; i++) f(i, 0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += nthreads) f(i, t);
    });
  for (auto& x : th) x.join();
}

}  // namespace

// ====================================================================== C API
struct orc_system {
  System S;
};

extern "C" orc_system* orc_create(const char* deck_path, char* err, int errlen) {
  try {
    Tok words;
    parse_file(deck_path, words);
    orc_system* o = new orc_system;
    read_system(o->S, words);
    Tok tf;
    if (!readsection0(words, tf, "TRIALFUNC")) fail("missing TRIALFUNC");
    Tok wf1, wf2;
    std::string dir = dirname_of(deck_path);
    if (haskeyword(tf, "SLATER-JASTROW")) {
      if (!readsection0(tf, wf1, "WF1") || !readsection0(tf, wf2, "WF2")) fail("SLATER-JASTROW needs WF1 and WF2");
      read_slater(o->S, wf1, dir);
      read_jastrow(o->S, wf2);
    } else if (haskeyword(tf, "SLATER")) {
      read_slater(o->S, tf, dir);
    } else {
      fail("TRIALFUNC must be SLATER-JASTROW or SLATER");
    }
    return o;
  } catch (std::exception& ex) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", ex.what());
    return nullptr;
  }
}

extern "C" void orc_destroy(orc_system* s) { delete s; }

extern "C" void orc_info(const orc_system* s, int* out) {
  out[0] = s->S.nup;
  out[1] = s->S.ndn;
  out[2] = (int)s->S.atoms.size();
  out[3] = s->S.totbasis;
  out[4] = s->S.pp.empty() ? 0 : 1;
}

// ---------------------------------------------------------------- parameter derivatives
// Linear Jastrow parameters and d ln Psi / dp at one configuration: Wavefunction::getParmDeriv
// of SLATER-JASTROW with one determinant (no determinant or orbital parameters) and JASTROW2
// (src/wavefunction/Slat_Jastrow.cpp:171-193, Slat_wf.h:555-575,680-686,
// Jastrow2_wf.cpp:1486-1553 and 1900-1945 (create_parm_deriv), Jastrow2_one.cpp:214-248,
// Jastrow2_two.cpp:153-180,407-440, Wavefunction.cpp:27-113 (extend_parm_deriv)).
// A piece is a group's one-body or two-body term with parameters (FREEZE: none, and the piece
// contributes nothing, not even to val_gradient).  For a piece with basis sums func(p,e,0..4)
// (value, gradient, laplacian; two-body values counted once per pair, gradients
// antisymmetric, laplacians on both electrons):
//   gradient(p) = sum_e func(p,e,0);  gd(p,e,0..2) = func(p,e,1..3);
//   vg_piece(e) = sum_p c_p func(p,e,1..3);  gd(p,e,3) = func(p,e,4) + 2 gd(p,e,0..2).vg_piece(e).
// Appending a piece adds to each earlier parameter 2 gd.vg_new and to each new one
// 2 gd.vg_before, and vg accumulates; the Slater determinant (no parameters) finally adds
// 2 gd(p,e,0..2).grad ln D(e) and its gradient to vg.
struct PDeriv {
  std::vector<double> g;   // [np]
  std::vector<double> gd;  // [np][ne][4]
  std::vector<double> vg;  // [ne][3]
  std::vector<double> parms;
};
void pd_append(PDeriv& acc, const std::vector<double>& func, const std::vector<double>& coef, int ne) {
  const int np = (int)coef.size(), np0 = (int)acc.g.size();
  std::vector<double> vgp((size_t)ne * 3, 0.0);
  for (int p = 0; p < np; p++)
    for (int e = 0; e < ne; e++)
      for (int d = 0; d < 3; d++) vgp[3 * e + d] += coef[p] * func[((size_t)p * ne + e) * 5 + 1 + d];
  // earlier parameters: cross term with the new piece's gradient
  for (int p = 0; p < np0; p++)
    for (int e = 0; e < ne; e++) {
      double dot = 0;
      for (int d = 0; d < 3; d++) dot += acc.gd[((size_t)p * ne + e) * 4 + d] * vgp[3 * e + d];
      acc.gd[((size_t)p * ne + e) * 4 + 3] += 2 * dot;
    }
  for (int p = 0; p < np; p++) {
    double g = 0;
    for (int e = 0; e < ne; e++) g += func[((size_t)p * ne + e) * 5];
    acc.g.push_back(g);
    acc.parms.push_back(coef[p]);
    for (int e = 0; e < ne; e++) {
      const double* f = &func[((size_t)p * ne + e) * 5];
      double own = 0, before = 0;
      for (int d = 0; d < 3; d++) own += vgp[3 * e + d] * f[1 + d], before += f[1 + d] * acc.vg[3 * e + d];
      double l = f[4] + 2 * own;
      l += 2 * before;
      for (int d = 0; d < 3; d++) acc.gd.push_back(f[1 + d]);
      acc.gd.push_back(l);
    }
  }
  for (int i = 0; i < 3 * ne; i++) acc.vg[i] += vgp[i];
}

void parm_deriv(const System& S, Walker& w, PDeriv& out, JScratch& js) {
  const int ne = S.ne(), nat = (int)S.atoms.size();
  out.g.clear();
  out.gd.clear();
  out.parms.clear();
  out.vg.assign((size_t)ne * 3, 0.0);
  std::vector<double> func;
  for (const JGroup& G : S.groups) {
    if (G.has_one && !G.one_frozen) {
      // parameters: COEFFICIENTS sections in order, coefficient k of the atom type's
      // concatenated EI basis
      std::vector<double> coef;
      std::vector<std::pair<int, int>> pk;  // (section, k)
      for (size_t sc = 0; sc < G.sec_label.size(); sc++)
        for (size_t k = 0; k < G.sec_coef[sc].size(); k++) coef.push_back(G.sec_coef[sc][k]), pk.push_back({(int)sc, (int)k});
      func.assign(coef.size() * ne * 5, 0.0);
      for (size_t p = 0; p < pk.size(); p++)
        for (int e = 0; e < ne; e++)
          for (int a = 0; a < nat; a++) {
            if (S.atoms[a].label != G.sec_label[pk[p].first]) continue;
            double R[5];
            dist5(&w.pos[3 * e], S.atoms[a].r, R);
            int k = 0;
            for (int bi : G.atom2basis[a]) {
              const JBasis& jb = G.eib[bi];
              js.b.resize(5 * jb.nfunc());
              jb.calc_lap(R, js.b.data());
              for (int i = 0; i < jb.nfunc(); i++, k++)
                if (k == pk[p].second)
                  for (int d = 0; d < 5; d++) func[((size_t)p * ne + e) * 5 + d] += js.b[5 * i + d];
            }
          }
      if (!coef.empty()) pd_append(out, func, coef, ne);
    }
    if (G.has_two && !G.two_frozen) {
      const int nb = G.diffspin ? (int)G.like.size() : (int)G.twocoef.size();
      std::vector<double> coef = G.diffspin ? G.like : G.twocoef;
      if (G.diffspin) coef.insert(coef.end(), G.unlike.begin(), G.unlike.end());
      func.assign(coef.size() * ne * 5, 0.0);
      for (int i = 0; i < ne; i++)
        for (int j = i + 1; j < ne; j++) {
          double R[5];
          dist5(&w.pos[3 * i], &w.pos[3 * j], R);
          const int shift = G.diffspin && ((j < S.nup) != (i < S.nup)) ? nb : 0;
          int k = 0;
          for (const JBasis& jb : G.eeb) {
            js.b.resize(5 * jb.nfunc());
            if (R[0] < jb.cutoff()) jb.calc_lap(R, js.b.data());
            else std::fill(js.b.begin(), js.b.end(), 0.0);
            for (int q = 0; q < jb.nfunc(); q++, k++) {
              if (k >= nb) continue;
              double* fi = &func[((size_t)(k + shift) * ne + i) * 5];
              double* fj = &func[((size_t)(k + shift) * ne + j) * 5];
              fi[0] += js.b[5 * q];
              for (int d = 1; d < 4; d++) fi[d] += js.b[5 * q + d], fj[d] -= js.b[5 * q + d];
              fi[4] += js.b[5 * q + 4];
              fj[4] += js.b[5 * q + 4];
            }
          }
        }
      if (!coef.empty()) pd_append(out, func, coef, ne);
    }
  }
  // the determinants: gradient of ln Psi_S per electron, and with OPTIMIZE_DET the CSF
  // coefficients 1.. as parameters (src/wavefunction/Slat_wf.h:555-686): for CSF p with
  // determinants j and weights w_j,
  //   gradient(p) = sum_j w_j D_j / Psi_S,
  //   gradderiv(p, e, c) = sum_j w_j D_j (d_c D_j / D_j - d_c Psi_S / Psi_S) / Psi_S,
  // d_c = d/dx, d/dy, d/dz and the laplacian of electron e
  std::vector<double> gD((size_t)ne * 3), lD(ne);
  for (int e = 0; e < ne; e++) det_gradlap(S, w, e, &gD[3 * e], &lD[e]);
  PDeriv sl;
  if (S.optimize_det && S.csf_c.size() > 1) {
    std::vector<double> b;
    double l0, s0, den = 0;
    det_terms(S, w, b, l0, s0);
    for (double x : b) den += x;
    // e_d = D_d / (c_0 D_0) without the weight c_d
    std::vector<double> ed(S.ndet);
    for (int d = 0; d < S.ndet; d++) {
      const double l = w.logdet[0][d] + w.logdet[1][d], sg = w.sign[0][d] * w.sign[1][d];
      ed[d] = sg * s0 * exp(l - l0);
    }
    const int ncsf = (int)S.csf_c.size();
    sl.g.assign(ncsf - 1, 0.0);
    sl.gd.assign((size_t)(ncsf - 1) * ne * 4, 0.0);
    int det = (int)S.csf_w[0].size();
    for (int c = 1; c < ncsf; c++)
      for (double wj : S.csf_w[c]) {
        sl.g[c - 1] += wj * ed[det] / den;
        for (int e = 0; e < ne; e++) {
          const int s = S.spin_of(e), n = S.nspin(s), k = e - s * S.nup, nm = S.nmo[s];
          const double* r = morow(S, w, e);
          const double* iv = &w.inv[s][det][(size_t)k * n];
          for (int q = 0; q < 4; q++) {
            double t = 0;
            for (int j = 0; j < n; j++) t += r[(q + 1) * nm + S.occ[s][det][j]] * iv[j];
            const double tot = q < 3 ? gD[3 * e + q] : lD[e];
            sl.gd[((size_t)(c - 1) * ne + e) * 4 + q] += wj * ed[det] * (t - tot) / den;
          }
        }
        det++;
      }
    for (int c = 1; c < ncsf; c++) sl.parms.push_back(S.csf_c[c]);
  }
  const int np = (int)out.g.size(), ns = (int)sl.g.size();
  // extend_parm_deriv(Slater, Jastrow) (src/wavefunction/Wavefunction.cpp:27-113)
  for (int p = 0; p < np; p++)
    for (int e = 0; e < ne; e++) {
      double dot = 0;
      for (int d = 0; d < 3; d++) dot += out.gd[((size_t)p * ne + e) * 4 + d] * gD[3 * e + d];
      out.gd[((size_t)p * ne + e) * 4 + 3] += 2 * dot;
    }
  for (int p = 0; p < ns && np > 0; p++)
    for (int e = 0; e < ne; e++) {
      double dot = 0;
      for (int d = 0; d < 3; d++) dot += sl.gd[((size_t)p * ne + e) * 4 + d] * out.vg[3 * e + d];
      sl.gd[((size_t)p * ne + e) * 4 + 3] += 2 * dot;
    }
  if (np == 0) {
    out.vg = gD;
  } else {
    for (int i = 0; i < 3 * ne; i++) out.vg[i] += gD[i];
  }
  if (ns) {
    out.g.insert(out.g.begin(), sl.g.begin(), sl.g.end());
    out.gd.insert(out.gd.begin(), sl.gd.begin(), sl.gd.end());
    out.parms.insert(out.parms.begin(), sl.parms.begin(), sl.parms.end());
  }
}

extern "C" int orc_eval(orc_system* o, int K, const double* pos, const double* rot, int pp_mode,
                        const double* pp_rand, double* logpsi, double* sign, double* grad, double* lap,
                        double* eparts) {
  const System& S = o->S;
  int ne = S.ne(), na = (int)S.atoms.size();
  std::vector<double> tmp;
  JScratch js;
  for (int k = 0; k < K; k++) {
    Walker w;
    walker_init(S, w);
    w.pos.assign(pos + (size_t)k * ne * 3, pos + (size_t)(k + 1) * ne * 3);
    walker_refresh(S, w, tmp);
    double U = jastrow_total(S, w.pos.data(), js);
    std::vector<double> b;
    double l0, s0, sb = 0;
    det_terms(S, w, b, l0, s0);
    for (double x : b) sb += x;
    logpsi[k] = l0 + log(fabs(sb)) + U;
    sign[k] = s0 * (sb < 0 ? -1.0 : 1.0);
    double ke = kinetic(S, w, grad + (size_t)k * ne * 3, lap + (size_t)k * ne, js);
    eparts[3 * k + 0] = ke;
    eparts[3 * k + 1] = coulomb(S, w.pos.data());
    eparts[3 * k + 2] = S.pp.empty() ? 0.0
                                     : pp_nonlocal(S, w, rot, pp_mode, pp_rand ? pp_rand + (size_t)k * na * ne : nullptr,
                                                   tmp, js);
  }
  return 0;
}

// T-move candidates of fixed configurations (deterministic or stochastic PP as in orc_eval).
// Returns the number of candidates of configuration k written to e_out[cap], vxx_out[cap],
// disp_out[cap][3] (-1 if cap is too small); vnl_out = the nonlocal energy without them.
extern "C" int orc_tmoves(orc_system* o, const double* pos, const double* rot, int pp_mode, const double* pp_rand,
                          int cap, int* e_out, double* vxx_out, double* disp_out, double* vnl_out) {
  const System& S = o->S;
  int ne = S.ne();
  std::vector<double> tmp;
  JScratch js;
  Walker w;
  walker_init(S, w);
  w.pos.assign(pos, pos + (size_t)ne * 3);
  walker_refresh(S, w, tmp);
  std::vector<Tmove> tl;
  *vnl_out = S.pp.empty() ? 0.0 : pp_nonlocal(S, w, rot, pp_mode, pp_rand, tmp, js, &tl);
  if ((int)tl.size() > cap) return -1;
  for (size_t t = 0; t < tl.size(); t++) {
    e_out[t] = tl[t].e;
    vxx_out[t] = tl[t].vxx;
    for (int d = 0; d < 3; d++) disp_out[3 * t + d] = tl[t].disp[d];
  }
  return (int)tl.size();
}

// T-move step of one walker at a fixed configuration (tests of the selection rule): the
// candidates at pos, then do_tmove with the given seed / walker / step counters.  pos is
// updated in place; returns the number of electrons moved.
extern "C" int orc_tmove_step(orc_system* o, double* pos, const double* rot, int pp_mode, const double* pp_rand,
                              double tau, int sc, unsigned long long seed, unsigned wid, unsigned step) {
  const System& S = o->S;
  int ne = S.ne();
  std::vector<double> tmp;
  JScratch js;
  Walker w;
  walker_init(S, w);
  w.pos.assign(pos, pos + (size_t)ne * 3);
  walker_refresh(S, w, tmp);
  std::vector<Tmove> tl;
  if (!S.pp.empty()) pp_nonlocal(S, w, rot, pp_mode, pp_rand, tmp, js, &tl);
  Point P;
  int moved = do_tmove(S, w, tl, tau, sc, seed, wid, step, tmp, js, P);
  std::copy(w.pos.begin(), w.pos.end(), pos);
  return moved;
}

// Number of linear Jastrow parameters, their values (may be NULL).
extern "C" int orc_nparm(orc_system* o, double* values) {
  const System& S = o->S;
  Walker w;
  walker_init(S, w);
  w.pos.assign((size_t)S.ne() * 3, 0.0);
  for (int e = 0; e < S.ne(); e++) w.pos[3 * e] = 100.0 + 7.0 * e;  // far apart: only the count matters
  std::vector<double> tmp;
  walker_refresh(S, w, tmp);
  PDeriv pd;
  JScratch js;
  parm_deriv(S, w, pd, js);
  if (values)
    for (size_t p = 0; p < pd.parms.size(); p++) values[p] = pd.parms[p];
  return (int)pd.g.size();
}

// Sets the linear Jastrow parameters (same order and count as orc_nparm) — for finite
// differences in the tests.  Returns the number set.
extern "C" int orc_set_parms(orc_system* o, const double* values) {
  System& S = o->S;
  int q = 0;
  if (S.optimize_det && S.csf_c.size() > 1) {
    for (size_t c = 1; c < S.csf_c.size(); c++) S.csf_c[c] = values[q++];
    int det = 0;
    for (size_t c = 0; c < S.csf_c.size(); c++)
      for (double wj : S.csf_w[c]) S.detwt[det++] = S.csf_c[c] * wj;
  }
  for (JGroup& G : S.groups) {
    if (G.has_one && !G.one_frozen)
      for (size_t sc = 0; sc < G.sec_label.size(); sc++)
        for (size_t k = 0; k < G.sec_coef[sc].size(); k++) {
          G.sec_coef[sc][k] = values[q++];
          for (size_t a = 0; a < S.atoms.size(); a++)
            if (S.atoms[a].label == G.sec_label[sc]) G.onecoef[a][k] = G.sec_coef[sc][k];
        }
    if (G.has_two && !G.two_frozen) {
      if (G.diffspin) {
        for (auto& c : G.like) c = values[q++];
        for (auto& c : G.unlike) c = values[q++];
      } else {
        for (auto& c : G.twocoef) c = values[q++];
      }
    }
  }
  return q;
}

// Parameter derivatives at K configurations: grad[K][np], gd[K][np][ne][4], vg[K][ne][3].
extern "C" int orc_parm_deriv(orc_system* o, int K, const double* pos, double* grad, double* gd, double* vg) {
  const System& S = o->S;
  const int ne = S.ne();
  std::vector<double> tmp;
  JScratch js;
  int np = -1;
  for (int k = 0; k < K; k++) {
    Walker w;
    walker_init(S, w);
    w.pos.assign(pos + (size_t)k * ne * 3, pos + (size_t)(k + 1) * ne * 3);
    walker_refresh(S, w, tmp);
    PDeriv pd;
    parm_deriv(S, w, pd, js);
    np = (int)pd.g.size();
    std::copy(pd.g.begin(), pd.g.end(), grad + (size_t)k * np);
    std::copy(pd.gd.begin(), pd.gd.end(), gd + (size_t)k * np * ne * 4);
    std::copy(pd.vg.begin(), pd.vg.end(), vg + (size_t)k * ne * 3);
  }
  return np;
}

extern "C" int orc_mo_eval(orc_system* o, int spin, int n, const double* pts, double* out) {
  std::vector<double> tmp;
  int nm = o->S.nmo[spin];
  for (int i = 0; i < n; i++) mo_lap(o->S, spin, pts + 3 * i, out + (size_t)i * 5 * nm, tmp);
  return nm;
}

extern "C" int orc_basis_eval(orc_system* o, int n, const double* pts, double* out) {
  const System& S = o->S;
  std::vector<double> tmp;
  for (int p = 0; p < n; p++) {
    double* op = out + (size_t)p * S.totbasis * 5;
    int tot = 0;
    for (size_t a = 0; a < S.atoms.size(); a++) {
      double R[5];
      dist5(pts + 3 * p, S.atoms[a].r, R);
      for (int j : S.center_basis[a]) {
        const AOSpline& ao = S.aos[j];
        tmp.resize(5 * ao.nfunc);
        ao.calc_lap(R, tmp.data());
        for (int i = 0; i < ao.nfunc; i++, tot++)
          for (int d = 0; d < 5; d++) op[5 * tot + d] = (R[0] < S.fcut[tot]) ? tmp[5 * i + d] : 0.0;
      }
    }
  }
  return S.totbasis;
}

extern "C" int orc_spline(orc_system* o, int shell, double* coeff, double* meta) {
  int k = 0;
  for (auto& ao : o->S.aos)
    for (size_t i = 0; i < ao.spl.size(); i++, k++)
      if (k == shell) {
        const Spline& sp = ao.spl[i];
        if (coeff) std::copy(sp.c.begin(), sp.c.end(), coeff);
        if (meta) {
          meta[0] = sp.spacing;
          meta[1] = sp.threshold;
          int f = 0;
          for (size_t q = 0; q < i; q++) f += sym_nfunc(ao.sym[q]);
          meta[2] = ao.rcut[f];
        }
        return sp.n();
      }
  return -1;
}

extern "C" int orc_mo_coeff(orc_system* o, int spin, double* C, double* rcut) {
  const System& S = o->S;
  if (C) std::copy(S.C[spin].begin(), S.C[spin].end(), C);
  if (rcut) std::copy(S.fcut.begin(), S.fcut.end(), rcut);
  return S.nmo[spin];
}

extern "C" int orc_transpose_inverse(int n, const double* a, double* a1, double* logdet, double* sign) {
  return transpose_inverse(a, a1, n, *logdet, *sign);
}

extern "C" double orc_inverse_update_row(int n, double* a1, const double* newrow, int row) {
  return inverse_update_row(a1, newrow, row, n);
}

extern "C" void orc_philox(const unsigned* ctr, const unsigned* key, unsigned* out) { philox(ctr, key, out); }

// ---------------------------------------------------------------------- walker engine
// VMC: src/method/Vmc_method.cpp:351-558.  DMC: src/method/Dmc_method.cpp:328-600,776-1156.
extern "C" double orc_run(orc_system* o, const orc_run_opts* opt, double* pos, double* weights, double* eblock,
                          double* state) {
  const System& S = o->S;
  int nw = opt->nwalkers, ne = S.ne();
  int nth = std::max(1, opt->nthreads);
  std::vector<Walker> W(nw);
  std::vector<std::vector<double>> tmps(nth);
  std::vector<JScratch> jss(nth);
  std::vector<std::vector<Point>> ptss(nth, std::vector<Point>(3));
  for (int i = 0; i < nw; i++) {
    walker_init(S, W[i]);
    W[i].pos.assign(pos + (size_t)i * ne * 3, pos + (size_t)(i + 1) * ne * 3);
  }
  std::vector<double> eold(nw), enew(nw), w(nw, 1.0), wk(nw), ekin(nw), evl(nw), evn(nw);
  std::vector<int> acc(nw);
  for (int i = 0; i < nw; i++) w[i] = weights ? weights[i] : 1.0;
  bool dmc = opt->method == 1;
  unsigned step = opt->step0;
  double tau = opt->timestep, teff = tau;
  double etrial = 0, eref = 0, cut_start = 1e300, cut_stop = 1e300;
  long long accepted = 0, tried = 0;

  // initial local energies (restorecheckpoint + find_cutoffs, src/method/Dmc_method.cpp:262-300,733-773)
  parallel_for(nw, nth, [&](int i, int t) {
    walker_refresh(S, W[i], tmps[t]);
    Energy E = local_energy(S, W[i], opt->seed, (unsigned)i, step, opt->pp_stochastic, tmps[t], jss[t]);
    eold[i] = E.total();
  });
  if (dmc) {
    if (state && state[3] > 0) {
      etrial = state[0];
      eref = state[1];
      teff = state[2];
      cut_start = opt->branch_start_cutoff * state[3];
      cut_stop = cut_start * 1.5;
    } else {
      double ea = 0, tw = 0;
      for (int i = 0; i < nw; i++) ea += eold[i] * w[i], tw += w[i];
      ea /= tw;
      double v = 0;
      for (int i = 0; i < nw; i++) v += (eold[i] - ea) * (eold[i] - ea);
      double sig = sqrt(v / nw);
      eref = ea;
      cut_start = opt->branch_start_cutoff * sig;
      cut_stop = cut_start * 1.5;
      double totw = 0;
      for (int i = 0; i < nw; i++) totw += w[i];
      etrial = eref - opt->feedback * log(totw / nw);
      if (state) state[3] = sig;
    }
  }
  step++;
  int fbi = std::max(1, opt->feedback_interval);
  for (int b = 0; b < opt->nblock; b++) {
    double sw = 0, swe = 0, swk = 0, swl = 0, swn = 0;
    if (!dmc)
      parallel_for(nw, nth, [&](int i, int t) { walker_refresh(S, W[i], tmps[t]); });
    for (int st = 0; st < opt->nstep;) {
      int np = dmc ? std::min(fbi, opt->nstep - st) : 1;
      long long acc_chunk = 0;
      for (int p = 0; p < np; p++) {
        unsigned cur = step;
        parallel_for(nw, nth, [&](int i, int t) {
          int a = 0;
          int nsweep = dmc ? 1 : std::max(1, opt->ndecorr);
          for (int sw2 = 0; sw2 < nsweep; sw2++)
            for (int e = 0; e < ne; e++)
              a += move_electron(S, W[i], e, tau, dmc ? 1 : 2, dmc ? 1 : 0, opt->seed, (unsigned)i, cur,
                                 (unsigned)sw2, tmps[t], jss[t], ptss[t].data()) > 0;
          acc[i] = a;
          const bool tmv = dmc && opt->tmoves && !S.pp.empty();
          std::vector<Tmove> tl;
          Energy E = local_energy(S, W[i], opt->seed, (unsigned)i, cur, opt->pp_stochastic, tmps[t], jss[t],
                                  tmv ? &tl : nullptr, tau);
          enew[i] = E.total();
          ekin[i] = E.ke;
          evl[i] = E.vloc;
          evn[i] = E.vnl;
          if (tmv) do_tmove(S, W[i], tl, tau, opt->tmoves == 2, opt->seed, (unsigned)i, cur, tmps[t], jss[t], ptss[t][0]);
        });
        for (int i = 0; i < nw; i++) {
          acc_chunk += acc[i];
          if (dmc) {
            w[i] *= dmc_weight(teff, etrial, enew[i], eold[i], cut_start, cut_stop);
            if (w[i] > opt->max_weight) w[i] = opt->max_weight;
          }
          eold[i] = enew[i];
          sw += w[i];
          swe += w[i] * enew[i];
          swk += w[i] * ekin[i];
          swl += w[i] * evl[i];
          swn += w[i] * evn[i];
        }
        step++;
      }
      accepted += acc_chunk;
      tried += (long long)nw * ne * np * (dmc ? 1 : std::max(1, opt->ndecorr));
      st += np;
      if (dmc) {
        teff = tau * (double)acc_chunk / ((double)nw * ne * np);
        double totw = 0;
        for (int i = 0; i < nw; i++) totw += w[i];
        etrial = eref - opt->feedback * log(totw / nw);
        // fixed-population pair branching (src/method/Dmc_method.cpp:961-995)
        std::vector<std::pair<double, int>> srt(nw);
        for (int i = 0; i < nw; i++) srt[i] = {w[i], i};
        std::stable_sort(srt.begin(), srt.end(),
                         [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
        int small = 0;
        unsigned bev = step;
        for (int i = nw - 1, pair = 0; i > 1 && 2 * pair + 1 < nw; i--, pair++) {
          if (srt[i].first < 1.8) break;
          int big = srt[i].second, sm = srt[small].second;
          double w1 = w[big] / (w[big] + w[sm]);
          double u = uniform(opt->seed, (unsigned)pair, bev, 0, TAG_BRANCH);
          double nwt = (w[big] + w[sm]) / 2.0;
          if (w1 + u >= 1.0) {
            W[sm] = W[big];
            eold[sm] = eold[big];
          } else {
            W[big] = W[sm];
            eold[big] = eold[sm];
          }
          w[big] = w[sm] = nwt;
          small++;
        }
      }
    }
    if (eblock) {
      eblock[5 * b + 0] = swe / sw;
      eblock[5 * b + 1] = swk / sw;
      eblock[5 * b + 2] = swl / sw;
      eblock[5 * b + 3] = swn / sw;
      eblock[5 * b + 4] = sw;
    }
    if (dmc) {
      // E_ref <- average over the blocks so far (src/method/Dmc_method.cpp:501-504)
      double s = 0;
      for (int q = 0; q <= b; q++) s += eblock ? eblock[5 * q] : swe / sw;
      eref = s / (b + 1);
      double totw = 0;
      for (int i = 0; i < nw; i++) totw += w[i];
      etrial = eref - opt->feedback * log(totw / nw);
    }
  }
  for (int i = 0; i < nw; i++) {
    std::copy(W[i].pos.begin(), W[i].pos.end(), pos + (size_t)i * ne * 3);
    if (weights) weights[i] = w[i];
  }
  if (state) {
    state[0] = etrial;
    state[1] = eref;
    state[2] = teff;
  }
  return tried ? (double)accepted / tried : 0.0;
}
