// Device-side data layout and helper functions shared by the walker-engine kernels.
// One 64-lane wavefront works on one walker; lanes map to molecular orbitals in the LCAO,
// to partner electrons / atoms in the Jastrow pair sums, and to spline shells in the basis.
#pragma once
#include <hip/hip_runtime.h>

#include "qw_setup.hpp"

namespace qw {

constexpr int kWave = 64;

// Read-only tables read at wave-uniform addresses go through the constant address space so
// the compiler can use scalar loads and hoist them out of loops; the tables are never
// written while a kernel runs.
#define QW_CONST __attribute__((address_space(4)))
template <typename T>
__device__ inline const QW_CONST T* cptr(const T* p) {
  return (const QW_CONST T*)p;
}

// Compact shell record for LDS staging: rc = min(shell cutoff, object cutoff);
// meta = atom | sym << 10 | nf << 14 | table << 18.
struct ShellLds {
  double rc;
  int meta;
  int f0;
};
__device__ inline int shl_atom(int m) { return m & 1023; }
__device__ inline int shl_sym(int m) { return (m >> 10) & 15; }
__device__ inline int shl_nf(int m) { return (m >> 14) & 15; }
__device__ inline int shl_table(int m) { return m >> 18; }

// Device form of a Jastrow radial object (CUTOFF_CUSP kind 0, POLYPADE kind 1) with the
// constants of its evaluation folded in: R = double (FP64 path) or float (mixed path).
//   kind 0: ccl/ccu = c*cusp (like/unlike), scale = rcut, gamma, pade0
//   kind 1: beta[i], cl/cu[i] = c_i, cbl/cbu[i] = c_i*(beta_i + 1)
template <typename R>
struct JRad {
  int kind, nf;
  R rcut, rcinv, rcinv2, gamma, pade0;
  R beta[kMaxJF];
  R cl[kMaxJF], cu[kMaxJF];
  R cbl[kMaxJF], cbu[kMaxJF];
  int nz[kMaxJF];  // coefficient i is non-zero for like or unlike pairs
};

// Read-only system tables in device memory.
struct DevSys {
  int nup, ndn, ne, na, nb, nsh, ntab, nint;
  int nee, nchunk, has_pp, nsp;  // nsp = max(nup, ndn): row length of the MO cache
  const double* atom_pos;
  const double* atom_Z;
  const ShellDesc* shells;
  const int* chunk_sh0;   // [nchunk+1] shell ranges whose functions fit one LDS tile
  const double* spline;   // [ntab][nint][4]
  const double* tab_inv;  // [ntab] 1/spacing (interval lookup)
  const double* tab_sp;   // [ntab] spacing (height within the interval)
  const double* C[2];     // [nb][nspin]
  const float* Cf[2];     // FP32 copies
  const JRad<double>* ee;   // [nee] electron-electron objects
  const JRad<float>* eef;
  // electron-ion Jastrow by atom type (atoms with identical object lists): atoms of type t
  // are eit_a0[t] .. eit_a0[t+1]-1 in type order (positions eit_pos, charges eit_Z), their
  // objects eit[eit_o0[t] .. eit_o0[t+1]-1]
  int n_eit;
  const JRad<double>* eit;
  const JRad<float>* eitf;
  const int* eit_a0;
  const int* eit_o0;
  const double* eit_pos;  // [na][3] in type order
  const double* eit_Z;    // [na] in type order
  const ShellLds* shl;    // [nsh] compact shell records
  const double* tabs;     // [ntab][2] 1/spacing, spacing
  const PPType* pp;
  const int* atom_pp;
  double vii;
  int chunk_nf;           // max functions per chunk (LDS tile width)
  // several determinants (ndet > 1): weights, per-spin MO list lengths (the MO cache row
  // holds nsp = max(nmo) values), occupations [2][ndet][nsig] as indices into the lists
  int ndet, nsig;         // nsig = max(nup, ndn): row stride of the determinant inverses
  int nmo[2];
  const double* detwt;    // [ndet]
  const double* logc;     // [ndet] log|detwt|
  const int* occ;         // [2][ndet][nsig]
};

// Per-walker state in device memory (walker-major).
struct DevWalkers {
  int nw;
  double* pos;     // [nw][ne][3]
  double* inv;     // [nw][2][nsp][nsp]  transposed inverse per spin (FP64 path)
  float* invf;     // [nw][2][nsp][nsp]  FP32 path
  double* mo;      // [nw][ne][5][nsp]   value, gradient, laplacian of the occupied MOs
  float* mof;      // FP32 path
  double* logdet;  // [nw][2]
  double* sgn;     // [nw][2]
  double* weight;  // [nw]
  double* eold;    // [nw] local energy of the previous step
  double* elast;   // [nw][4] last kinetic, Coulomb, nonlocal, total
  int* acc;        // [nw] accepted moves since the last reset
  double* bsum;    // [nw][16] block sums (see kBsum)
  double* dgl;     // [nw][ne][4] determinant gradient / laplacian ratios from the last sweep
  double* invd;    // several determinants: [nw][2][ndet][nsig][nsig] transposed inverses
  double* ldd;     // [nw][2][ndet] log|det|
  double* sgd;     // [nw][2][ndet] sign
};

// ---------------------------------------------------------------- wave reductions
// ---------------------------------------------------------------- wave reductions
// Called with all 64 lanes active.  Within each row of 16 lanes a DPP butterfly (quad_perm
// xor 1, quad_perm xor 2, row_half_mirror, row_mirror) leaves the row total in every lane; the
// four row totals are read with readlane and added in a fixed order, so every lane returns
// the bitwise-same value.
template <int CTRL>
__device__ inline int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ inline float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_i<CTRL>((int)(unsigned)b), hi = (unsigned)dpp_i<CTRL>((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline float readlane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

__device__ inline double wsum(double v) {
  v += dpp_d<kDppXor1>(v);
  v += dpp_d<kDppXor2>(v);
  v += dpp_d<kDppHalfMirror>(v);
  v += dpp_d<kDppMirror>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ inline float wsumf(float v) {
  v += dpp_f<kDppXor1>(v);
  v += dpp_f<kDppXor2>(v);
  v += dpp_f<kDppHalfMirror>(v);
  v += dpp_f<kDppMirror>(v);
  return (readlane_f(v, 0) + readlane_f(v, 16)) + (readlane_f(v, 32) + readlane_f(v, 48));
}
__device__ inline int wave_isum(int v) {
  v += dpp_i<kDppXor1>(v);
  v += dpp_i<kDppXor2>(v);
  v += dpp_i<kDppHalfMirror>(v);
  v += dpp_i<kDppMirror>(v);
  return (__builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16)) +
         (__builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48));
}
__device__ inline double wave_maxd(double v) {
  v = fmax(v, dpp_d<kDppXor1>(v));
  v = fmax(v, dpp_d<kDppXor2>(v));
  v = fmax(v, dpp_d<kDppHalfMirror>(v));
  v = fmax(v, dpp_d<kDppMirror>(v));
  return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }
// Orders the LDS accesses of one wavefront around this point (the lanes of a wave run in
// lockstep and the LDS serves a wave's requests in issue order, so the compiler fence is all
// that is needed when one wave owns the data).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Barrier between the LDS producer and consumer of one walker: the block barrier when a block
// is one walker (WPB = 1), the wavefront fence when a block holds WPB walkers, one per wave.
template <int WPB>
__device__ inline void walker_sync() {
  if constexpr (WPB == 1)
    __syncthreads();
  else
    wave_sync();
}
// Lane holding the largest v (the highest such lane on ties); -1 when no lane compares
// equal to the maximum (NaN inputs).
__device__ inline int wave_argmax_last(double v, int lane) {
  const double m = wave_maxd(v);
  const unsigned long long b = __ballot(v == m);
  return b ? 63 - __clzll((long long)b) : -1;
}
// Value of v in lane l (l uniform).
__device__ inline double wave_bcast(double v, int l) { return readlane_d(v, l); }

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
__device__ inline void philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1,
                              unsigned out[4]) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    unsigned h0 = __umulhi(0xD2511F53u, c0), l0 = 0xD2511F53u * c0;
    unsigned h1 = __umulhi(0xCD9E8D57u, c2), l1 = 0xCD9E8D57u * c2;
    unsigned n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
    c0 = n0;
    c1 = l1;
    c2 = n2;
    c3 = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
// Stream tags (counter word 3); the sweep number of a step sits in bits 8 and up.
enum : unsigned { TAG_GAUSS1 = 0, TAG_ACC1 = 1, TAG_GAUSS2 = 2, TAG_ACC2 = 3, TAG_PP = 4, TAG_ROT = 5, TAG_BRANCH = 6,
                  TAG_TMOVE = 7 };
__device__ inline double u53(unsigned a, unsigned b) {
  return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0);
}
__device__ inline double u32o(unsigned a) { return (a + 0.5) * (1.0 / 4294967296.0); }
__device__ inline void gauss3(unsigned long long seed, unsigned w, unsigned step, unsigned idx, unsigned tag, double g[3]) {
  unsigned o[4];
  philox(w, step, idx, tag, (unsigned)seed, (unsigned)(seed >> 32), o);
  const double twopi = 6.283185307179586476925;
  double r0 = sqrt(-2.0 * log(u32o(o[0]))), t0 = twopi * u32o(o[1]);
  double r1 = sqrt(-2.0 * log(u32o(o[2]))), t1 = twopi * u32o(o[3]);
  double s0, c0, s1, c1;
  sincos(t0, &s0, &c0);
  sincos(t1, &s1, &c1);
  g[0] = r0 * c0;
  g[1] = r0 * s0;
  g[2] = r1 * c1;
}
__device__ inline double uniform(unsigned long long seed, unsigned w, unsigned step, unsigned idx, unsigned tag) {
  unsigned o[4];
  philox(w, step, idx, tag, (unsigned)seed, (unsigned)(seed >> 32), o);
  return u53(o[0], o[1]);
}
// Uniform random rotation from a unit quaternion (Shoemake, Graphics Gems III, 1992);
// rows of rot are the rotated x, y, z axes.
__device__ inline void random_rotation(unsigned long long seed, unsigned w, unsigned step, double rot[9]) {
  unsigned o[4];
  philox(w, step, 0, TAG_ROT, (unsigned)seed, (unsigned)(seed >> 32), o);
  const double twopi = 6.283185307179586476925;
  double u1 = u53(o[0], o[1]), u2 = u32o(o[2]), u3 = u32o(o[3]);
  double a = sqrt(1 - u1), b = sqrt(u1);
  double s2, c2, s3, c3;
  sincos(twopi * u2, &s2, &c2);
  sincos(twopi * u3, &s3, &c3);
  double qw = a * s2, qx = a * c2, qy = b * s3, qz = b * c3;
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

// Cyrus drift limiter, a = 0.25 (src/method/Split_sample.cpp:58-78).
__device__ inline void lim_drift(const double v[3], double tau, double out[3]) {
  double d2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * 0.25;
  double da = d2 > 1e-8 ? (sqrt(1. + 2 * tau * d2) - 1) / d2 : tau;
  out[0] = v[0] * da;
  out[1] = v[1] * da;
  out[2] = v[2] * da;
}
// Exponent of the drift-diffusion Green's function from a to b (src/method/Split_sample.cpp:91-110).
__device__ inline double trans(const double ra[3], const double da[3], const double rb[3], double tau) {
  double dr[3];
  lim_drift(da, tau, dr);
  double p = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double t = rb[d] - ra[d] - dr[d];
    p -= t * t;
  }
  return p / (2.0 * tau);
}

// ---------------------------------------------------------------- basis
// ---------------------------------------------------------------- F and G shells
// 7F and 9G are the reference's real solid harmonics (unnormalised, its function order), 15G
// the Cartesian monomials, all as monomial lists c x^a y^b z^c (the reference's polynomials
// expanded; src/basis_function/Cubic_spline_calc.cpp:646-720, 720-1070, function orders
// src/basis_function/Cubic_spline_build.cpp:702-776).
struct AngMono {
  signed char c, ex, ey, ez;
};
__device__ inline double pw4(double v, int k) {
  const double v2 = v * v;
  return k == 0 ? 1.0 : k == 1 ? v : k == 2 ? v2 : k == 3 ? v2 * v : v2 * v2;
}
// Monomials of function id (0-6 F0 Fm3 Fp3mod Fp2 Fxyz Fm1 Fp1, 7-15 G0..G8, 16-30 the
// Cartesian 15G in the order xxxx yyyy zzzz xxxy xxxz yyyx yyyz zzzx zzzy xxyy xxzz yyzz
// xxyz yyxz zzxy); returns their count, *first the index of the first.
__device__ inline int ang_monos(int id, const AngMono** first) {
  static constexpr AngMono m[] = {
      {2, 0, 0, 3},   {-3, 2, 0, 1},  {-3, 0, 2, 1},                                  // F0
      {3, 2, 1, 0},   {-1, 0, 3, 0},                                                  // Fm3
      {1, 3, 0, 0},   {-3, 1, 2, 0},                                                  // Fp3mod
      {1, 2, 0, 1},   {-1, 0, 2, 1},                                                  // Fp2
      {1, 1, 1, 1},                                                                   // Fxyz
      {4, 0, 1, 2},   {-1, 2, 1, 0},  {-1, 0, 3, 0},                                  // Fm1
      {4, 1, 0, 2},   {-1, 3, 0, 0},  {-1, 1, 2, 0},                                  // Fp1
      {3, 4, 0, 0},   {3, 0, 4, 0},   {8, 0, 0, 4}, {6, 2, 2, 0}, {-24, 2, 0, 2}, {-24, 0, 2, 2},  // G0
      {4, 1, 0, 3},   {-3, 3, 0, 1},  {-3, 1, 2, 1},                                  // G1
      {4, 0, 1, 3},   {-3, 2, 1, 1},  {-3, 0, 3, 1},                                  // G2
      {-1, 4, 0, 0},  {1, 0, 4, 0},   {6, 2, 0, 2}, {-6, 0, 2, 2},                    // G3
      {6, 1, 1, 2},   {-1, 3, 1, 0},  {-1, 1, 3, 0},                                  // G4
      {1, 3, 0, 1},   {-3, 1, 2, 1},                                                  // G5
      {3, 2, 1, 1},   {-1, 0, 3, 1},                                                  // G6
      {1, 4, 0, 0},   {-6, 2, 2, 0},  {1, 0, 4, 0},                                   // G7
      {1, 3, 1, 0},   {-1, 1, 3, 0},                                                  // G8
      {1, 4, 0, 0},   {1, 0, 4, 0},   {1, 0, 0, 4}, {1, 3, 1, 0}, {1, 3, 0, 1}, {1, 1, 3, 0},
      {1, 0, 3, 1},   {1, 1, 0, 3},   {1, 0, 1, 3}, {1, 2, 2, 0}, {1, 2, 0, 2}, {1, 0, 2, 2},
      {1, 2, 1, 1},   {1, 1, 2, 1},   {1, 1, 1, 2}};                                  // 15G
  static constexpr unsigned char start[32] = {0,  3,  5,  7,  9,  10, 13, 16, 22, 25, 28, 32, 35, 37, 39, 42, 44,
                                              45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59};
  *first = m + start[id];
  return start[id + 1] - start[id];
}
// Function id of the i-th function of a 7F, 7F_crystal, 9G, 9G_pyscf or 15G shell.
__device__ inline int ang_fn(int sym, int i) {
  static constexpr unsigned char f7c[7] = {0, 6, 5, 3, 4, 2, 1};
  static constexpr unsigned char g9p[9] = {15, 13, 11, 9, 7, 8, 10, 12, 14};
  switch (sym) {
    case SK_7F: return i;
    case SK_7FC: return f7c[i];
    case SK_9G: return 7 + i;
    case SK_9GP: return g9p[i];
    default: return 16 + i;  // SK_15G
  }
}
// psi = P f(r): grad psi = f grad P + (f'/r) P r_vec; lap psi = f lap P + P (f'' + (2 + 2l) f'/r)
// for P homogeneous of degree l (r_vec . grad P = l P).
// Kept out of line: F and G shells are rare, and inlining would add their registers to every
// sweep kernel.
template <typename T>
__device__ __attribute__((noinline)) void shell_poly_lap(int sym, double f, double fd, double f2, double x, double y, double z, T* out,
                                      int ld) {
  const int nf = (sym == SK_7F || sym == SK_7FC) ? 7 : (sym == SK_15G ? 15 : 9);
  const double lr = (sym == SK_7F || sym == SK_7FC) ? 8.0 : 10.0;
  for (int i = 0; i < nf; i++) {
    const AngMono* m;
    const int nm = ang_monos(ang_fn(sym, i), &m);
    double P = 0, Px = 0, Py = 0, Pz = 0, L = 0;
    for (int k = 0; k < nm; k++) {
      const double c = m[k].c;
      const int a = m[k].ex, b = m[k].ey, d = m[k].ez;
      const double xa = pw4(x, a), yb = pw4(y, b), zd = pw4(z, d);
      P += c * xa * yb * zd;
      if (a) Px += c * a * pw4(x, a - 1) * yb * zd;
      if (b) Py += c * b * xa * pw4(y, b - 1) * zd;
      if (d) Pz += c * d * xa * yb * pw4(z, d - 1);
      if (a > 1) L += c * a * (a - 1) * pw4(x, a - 2) * yb * zd;
      if (b > 1) L += c * b * (b - 1) * xa * pw4(y, b - 2) * zd;
      if (d > 1) L += c * d * (d - 1) * xa * yb * pw4(z, d - 2);
    }
    const double h = P * fd;
    out[i] = (T)(P * f);
    out[ld + i] = (T)(f * Px + h * x);
    out[2 * ld + i] = (T)(f * Py + h * y);
    out[3 * ld + i] = (T)(f * Pz + h * z);
    out[4 * ld + i] = (T)(f * L + P * f2 + lr * h);
  }
}
template <typename T>
__device__ __attribute__((noinline)) void shell_poly_val(int sym, double f, double x, double y, double z, T* out) {
  const int nf = (sym == SK_7F || sym == SK_7FC) ? 7 : (sym == SK_15G ? 15 : 9);
  for (int i = 0; i < nf; i++) {
    const AngMono* m;
    const int nm = ang_monos(ang_fn(sym, i), &m);
    double P = 0;
    for (int k = 0; k < nm; k++) P += m[k].c * pw4(x, m[k].ex) * pw4(y, m[k].ey) * pw4(z, m[k].ez);
    out[i] = (T)(P * f);
  }
}

// Values of one spline shell at distance vector (x,y,z), r: writes nf functions x 5
// components to out[c*ld + i] (S, P, 5D, 6D, 10F; src/basis_function/Cubic_spline_calc.cpp:353-700).
template <typename T>
__device__ inline void shell_lap(int sym, double f, double fd, double f2, double x, double y, double z, T* out, int ld) {
  auto put = [&](int i, double v0, double v1, double v2, double v3, double v4) {
    out[i] = (T)v0;
    out[ld + i] = (T)v1;
    out[2 * ld + i] = (T)v2;
    out[3 * ld + i] = (T)v3;
    out[4 * ld + i] = (T)v4;
  };
  double v, h, v2, v3;
  switch (sym) {
    case SK_S:
      put(0, f, fd * x, fd * y, fd * z, f2 + 2. * fd);
      break;
    case SK_P: {
      double rx = fd * x, ry = fd * y, rz = fd * z;
      put(0, f * x, rx * x + f, rx * y, rx * z, f2 * x + 4. * rx);
      put(1, f * y, ry * x, ry * y + f, ry * z, f2 * y + 4. * ry);
      put(2, f * z, rz * x, rz * y, rz * z + f, f2 * z + 4. * rz);
      break;
    }
    case SK_5D:
      v = 2. * z * z - x * x - y * y; h = v * fd;
      put(0, v * f, h * x - 2 * f * x, h * y - 2 * f * y, h * z + 4 * f * z, v * f2 + 6 * h);
      v = x * z; h = v * fd;
      put(1, v * f, h * x + f * z, h * y, h * z + f * x, v * f2 + 6. * h);
      v = y * z; h = v * fd;
      put(2, v * f, h * x, h * y + f * z, h * z + f * y, v * f2 + 6. * h);
      v = x * x - y * y; h = v * fd;
      put(3, v * f, h * x + 2. * f * x, h * y - 2. * f * y, h * z, v * f2 + 6. * h);
      v = x * y; h = v * fd;
      put(4, v * f, h * x + f * y, h * y + f * x, h * z, v * f2 + 6. * h);
      break;
    case SK_6D:
      v = x * x; h = v * fd;
      put(0, v * f, h * x + f * 2. * x, h * y, h * z, v * f2 + 6. * h + 2. * f);
      v = y * y; h = v * fd;
      put(1, v * f, h * x, h * y + f * 2. * y, h * z, v * f2 + 6. * h + 2. * f);
      v = z * z; h = v * fd;
      put(2, v * f, h * x, h * y, h * z + f * 2. * z, v * f2 + 6. * h + 2. * f);
      v = x * y; h = v * fd;
      put(3, v * f, h * x + f * y, h * y + f * x, h * z, v * f2 + 6. * h);
      v = x * z; h = v * fd;
      put(4, v * f, h * x + f * z, h * y, h * z + f * x, v * f2 + 6. * h);
      v = y * z; h = v * fd;
      put(5, v * f, h * x, h * y + f * z, h * z + f * y, v * f2 + 6. * h);
      break;
    case SK_10F:
      v2 = x * x; v3 = v2 * x; h = v3 * fd;
      put(0, v3 * f, h * x + f * 3. * v2, h * y, h * z, v3 * f2 + 8. * h + 6. * f * x);
      v2 = y * y; v3 = v2 * y; h = v3 * fd;
      put(1, v3 * f, h * x, h * y + 3. * f * v2, h * z, v3 * f2 + 8. * h + 6. * f * y);
      v2 = z * z; v3 = v2 * z; h = v3 * fd;
      put(2, v3 * f, h * x, h * y, h * z + 3. * f * v2, v3 * f2 + 8. * h + 6. * f * z);
      v2 = x * x; v3 = v2 * y; h = v3 * fd;
      put(3, v3 * f, h * x + 2. * f * x * y, h * y + f * v2, h * z, v3 * f2 + 8. * h + 2. * f * y);
      v2 = x * x; v3 = v2 * z; h = v3 * fd;
      put(4, v3 * f, h * x + 2. * f * x * z, h * y, h * z + f * v2, v3 * f2 + 8. * h + 2. * f * z);
      v2 = y * y; v3 = v2 * x; h = v3 * fd;
      put(5, v3 * f, h * x + v2 * f, h * y + 2. * f * x * y, h * z, v3 * f2 + 8. * h + 2. * f * x);
      v2 = y * y; v3 = v2 * z; h = v3 * fd;
      put(6, v3 * f, h * x, h * y + 2. * f * y * z, h * z + f * v2, v3 * f2 + 8. * h + 2. * f * z);
      v2 = z * z; v3 = x * v2; h = v3 * fd;
      put(7, v3 * f, h * x + f * v2, h * y, h * z + 2. * f * x * z, v3 * f2 + 8. * h + 2. * f * x);
      v2 = z * z; v3 = v2 * y; h = v3 * fd;
      put(8, v3 * f, h * x, h * y + f * v2, h * z + 2. * f * y * z, v3 * f2 + 8. * h + 2. * f * y);
      v3 = x * y * z; v2 = f * z; h = v3 * fd;
      put(9, v3 * f, h * x + v2 * y, h * y + v2 * x, h * z + f * x * y, v3 * f2 + 8. * h);
      break;
    default:  // 7F, 7F_crystal, 9G, 9G_pyscf, 15G
      shell_poly_lap<T>(sym, f, fd, f2, x, y, z, out, ld);
      break;
  }
}

// Value-only variant for the pseudopotential quadrature (src/basis_function/Cubic_spline_calc.cpp:25-140).
template <typename T>
__device__ inline void shell_val(int sym, double f, double x, double y, double z, T* out) {
  switch (sym) {
    case SK_S: out[0] = (T)f; break;
    case SK_P: out[0] = (T)(f * x); out[1] = (T)(f * y); out[2] = (T)(f * z); break;
    case SK_5D:
      out[0] = (T)((2. * z * z - x * x - y * y) * f);
      out[1] = (T)((x * z) * f);
      out[2] = (T)((y * z) * f);
      out[3] = (T)((x * x - y * y) * f);
      out[4] = (T)((x * y) * f);
      break;
    case SK_6D:
      out[0] = (T)((x * x) * f); out[1] = (T)((y * y) * f); out[2] = (T)((z * z) * f);
      out[3] = (T)((x * y) * f); out[4] = (T)((x * z) * f); out[5] = (T)((y * z) * f);
      break;
    case SK_10F:
      out[0] = (T)((x * x * x) * f); out[1] = (T)((y * y * y) * f); out[2] = (T)((z * z * z) * f);
      out[3] = (T)((x * x * y) * f); out[4] = (T)((x * x * z) * f); out[5] = (T)((y * y * x) * f);
      out[6] = (T)((y * y * z) * f); out[7] = (T)((z * z * x) * f); out[8] = (T)((z * z * y) * f);
      out[9] = (T)((x * y * z) * f);
      break;
    default:
      shell_poly_val<T>(sym, f, x, y, z, out);
      break;
  }
}

// Basis values of the shells [sh0, sh1) at point p, written to chi[c*ld + (f - f0)]
// (c = value, d/dx, d/dy, d/dz, laplacian), zero beyond each shell's cutoff.
// Lanes take shells round-robin.  Table indices are clamped to the table.
template <typename T>
__device__ inline void basis_lap(const DevSys& S, const double p[3], int sh0, int sh1, int f0, T* chi, int ld, int lane) {
  for (int sh = sh0 + lane; sh < sh1; sh += kWave) {
    const ShellDesc d = S.shells[sh];
    const double* ap = S.atom_pos + 3 * d.atom;
    double x = p[0] - ap[0], y = p[1] - ap[1], z = p[2] - ap[2];
    double r2 = x * x + y * y + z * z, r = sqrt(r2);
    T* o = chi + (d.f0 - f0);
    if (r < d.rcut && r < d.objthr) {
      int iv = (int)(r * S.tab_inv[d.table]);
      iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
      const double* c = S.spline + ((size_t)d.table * S.nint + iv) * 4;
      double h = r - iv * S.tab_sp[d.table];
      double f = c[0] + h * (c[1] + h * (c[2] + h * c[3]));
      double fd = (c[1] + h * (2 * c[2] + h * (3 * c[3]))) / r;
      double f2 = 2 * c[2] + 6 * h * c[3];
      shell_lap<T>(d.sym, f, fd, f2, x, y, z, o, ld);
    } else {
      for (int i = 0; i < d.nf; i++)
        for (int c5 = 0; c5 < 5; c5++) o[c5 * ld + i] = (T)0;
    }
  }
}

template <typename T>
__device__ inline void basis_val(const DevSys& S, const double p[3], int sh0, int sh1, int f0, T* chi, int lane) {
  for (int sh = sh0 + lane; sh < sh1; sh += kWave) {
    const ShellDesc d = S.shells[sh];
    const double* ap = S.atom_pos + 3 * d.atom;
    double x = p[0] - ap[0], y = p[1] - ap[1], z = p[2] - ap[2];
    double r = sqrt(x * x + y * y + z * z);
    T* o = chi + (d.f0 - f0);
    if (r < d.rcut && r < d.objthr) {
      int iv = (int)(r * S.tab_inv[d.table]);
      iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
      const double* c = S.spline + ((size_t)d.table * S.nint + iv) * 4;
      double h = r - iv * S.tab_sp[d.table];
      double f = c[0] + h * (c[1] + h * (c[2] + h * c[3]));
      shell_val<T>(d.sym, f, x, y, z, o);
    } else {
      for (int i = 0; i < d.nf; i++) o[i] = (T)0;
    }
  }
}

// Exclusive prefix sum across the wave of small counts (0 <= v < 16, the function count of one
// shell) from four ballots; *total receives the wave sum.
__device__ inline int wave_excl_scan(int v, int lane, int* total) {
  int ex = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const unsigned long long m = __ballot((v >> b) & 1);
    ex += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) << b;
    tot += __popcll(m) << b;
  }
  *total = tot;
  return ex;
}

// Sparse basis evaluation: only shells inside their cutoff contribute (the functions the
// reference's CUTOFF_MO loop visits, src/orbitals/MO_matrix_cutoff.h:516-530).  Active
// functions of shells [sh0, sh1) are packed in increasing function order:
// chi[c*ld + i] (i < returned count) and flist[i] = global function index.
template <typename T>
__device__ inline int basis_lap_sparse(const DevSys& S, const double p[3], int sh0, int sh1, T* chi, int ld,
                                       int* flist, int lane) {
  int base = 0;
  for (int s0 = sh0; s0 < sh1; s0 += kWave) {
    const int sh = s0 + lane;
    ShellDesc d;
    double x = 0, y = 0, z = 0, r = 0;
    bool act = false;
    if (sh < sh1) {
      d = S.shells[sh];
      const double* ap = S.atom_pos + 3 * d.atom;
      x = p[0] - ap[0];
      y = p[1] - ap[1];
      z = p[2] - ap[2];
      r = sqrt(x * x + y * y + z * z);
      act = r < d.rcut && r < d.objthr;
    }
    int tot;
    const int at = base + wave_excl_scan(act ? d.nf : 0, lane, &tot);
    if (act) {
      int iv = (int)(r * S.tab_inv[d.table]);
      iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
      const double* c = S.spline + ((size_t)d.table * S.nint + iv) * 4;
      double h = r - iv * S.tab_sp[d.table];
      double f = c[0] + h * (c[1] + h * (c[2] + h * c[3]));
      double fd = (c[1] + h * (2 * c[2] + h * (3 * c[3]))) / r;
      double f2 = 2 * c[2] + 6 * h * c[3];
      shell_lap<T>(d.sym, f, fd, f2, x, y, z, chi + at, ld);
      for (int i = 0; i < d.nf; i++) flist[at + i] = d.f0 + i;
    }
    base += tot;
  }
  return base;
}

template <typename T>
__device__ inline int basis_val_sparse(const DevSys& S, const double p[3], int sh0, int sh1, T* chi, int* flist,
                                       int lane) {
  int base = 0;
  for (int s0 = sh0; s0 < sh1; s0 += kWave) {
    const int sh = s0 + lane;
    ShellDesc d;
    double x = 0, y = 0, z = 0, r = 0;
    bool act = false;
    if (sh < sh1) {
      d = S.shells[sh];
      const double* ap = S.atom_pos + 3 * d.atom;
      x = p[0] - ap[0];
      y = p[1] - ap[1];
      z = p[2] - ap[2];
      r = sqrt(x * x + y * y + z * z);
      act = r < d.rcut && r < d.objthr;
    }
    int tot;
    const int at = base + wave_excl_scan(act ? d.nf : 0, lane, &tot);
    if (act) {
      int iv = (int)(r * S.tab_inv[d.table]);
      iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
      const double* c = S.spline + ((size_t)d.table * S.nint + iv) * 4;
      double h = r - iv * S.tab_sp[d.table];
      double f = c[0] + h * (c[1] + h * (c[2] + h * c[3]));
      shell_val<T>(d.sym, f, x, y, z, chi + at);
      for (int i = 0; i < d.nf; i++) flist[at + i] = d.f0 + i;
    }
    base += tot;
  }
  return base;
}

// ---------------------------------------------------------------- Jastrow radial objects
// Reciprocal and square root from the hardware estimates refined by Newton-Raphson steps
// (two for FP64: relative error near the double epsilon; the FP32 estimate is used as is).
__device__ inline double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ inline float rcp_nr(float x) { return __builtin_amdgcn_rcpf(x); }
// r = |d| and 1/r from d2 = |d|^2 (r = 0, 1/r = inf for coincident points, as sqrt and a
// division give).
__device__ inline void dist_inv(double d2, double& r, double& ri) {
  double y = __builtin_amdgcn_rsq(d2);
  double h = 0.5 * d2;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  const bool z = d2 == 0.0;
  ri = z ? __builtin_inf() : y;
  r = z ? 0.0 : d2 * y;
}
__device__ inline void dist_inv(float d2, float& r, float& ri) {
  const float y = __builtin_amdgcn_rsqf(d2);
  const bool z = d2 == 0.f;
  ri = z ? __builtin_inff() : y;
  r = z ? 0.f : d2 * y;
}

// CUTOFF_CUSP (src/basis_function/Cutoff_cusp.cpp:122-154) and POLYPADE
// (src/basis_function/Poly_pade_function.cpp:183-237), contracted with the like-spin /
// one-body coefficients (cl) or the unlike-spin ones (cu), chosen per lane by uf = 0 / 1
// (c = uf*cu + (1-uf)*cl, exact).  The POLYPADE sum is regrouped as
//   u = zpp1 S0,  gfac = -zppd/(r rcut) S1,  d2 = -(zppdd2 S1 - zppd^2 S2)/rcut^2,
//   lap = 2 (d2 + gfac),  S0 = sum c_i p_i, S1 = sum c_i (b_i+1) p_i^2, S2 = sum c_i (b_i+1) b_i p_i^3,
// p_i = 1/(1 + b_i zpp): the reference's per-function terms summed in another order.
// The evaluators are branch-free (the cutoff masks the result) and specialised on the
// object's function count, so every parameter load is unconditional and loop-invariant in
// the pair loops that call them: the compiler keeps the parameters in scalar registers.
template <typename R, bool LAP>
struct CuspEval {
  const QW_CONST JRad<R>& o;
  __device__ inline void operator()(R uf, R r, R ri, R& u, R& gfac, R& lap) const {
    const R zz = r * o.rcinv, zz2 = zz * zz;
    const R pp = zz - zz2 + zz * zz2 * (R)(1.0 / 3.0);
    const R pade = rcp_nr((R)1 + o.gamma * pp), pade2 = pade * pade;
    const R ppd = (R)1 - (R)2 * zz + zz2;
    const R dadr = ppd * pade2 * ri;
    const R c = fma(uf, o.cu[0], ((R)1 - uf) * o.cl[0]);
    const bool in = r < o.rcut;
    u += in ? c * (o.rcut * (pp * pade - o.pade0)) : (R)0;
    gfac += in ? c * dadr : (R)0;
    if (LAP) {
      const R ppdd = (R)-2 + (R)2 * zz;
      const R dadr2 = ppdd * pade2 * o.rcinv - (R)2 * o.gamma * ppd * ppd * pade2 * pade * o.rcinv + (R)2 * dadr;
      lap += in ? c * dadr2 : (R)0;
    }
  }
};
template <typename R, bool LAP, int NF>
struct PolyEval {
  const QW_CONST JRad<R>& o;
  __device__ inline void operator()(R uf, R r, R ri, R& u, R& gfac, R& lap) const {
    const R zz = r * o.rcinv, zz2 = zz * zz;
    const R zpp = zz2 * ((R)6 - (R)8 * zz + (R)3 * zz2), zpp1 = (R)1 - zpp;
    const R zppd = (R)12 * zz * ((R)1 - (R)2 * zz + zz2);
    const R omu = (R)1 - uf;
    R s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < NF; i++) {
      const R b = o.beta[i];
      const R pz = rcp_nr((R)1 + b * zpp);
      const R c = fma(uf, o.cu[i], omu * o.cl[i]), cb = fma(uf, o.cbu[i], omu * o.cbl[i]);
      s0 += c * pz;
      const R t = cb * pz * pz;
      s1 += t;
      if (LAP) s2 += t * b * pz;
    }
    const bool in = r < o.rcut;
    const R g = -zppd * ri * o.rcinv * s1;
    u += in ? zpp1 * s0 : (R)0;
    gfac += in ? g : (R)0;
    if (LAP) {
      const R zppdd2 = (R)6 - (R)24 * zz + (R)18 * zz2;
      const R d2 = -o.rcinv2 * (zppdd2 * s1 - zppd * zppd * s2);
      lap += in ? (R)2 * (d2 + g) : (R)0;
    }
  }
};

// Runs f(evaluator) with the evaluator specialised for object o (kind and function count).
template <typename R, bool LAP, typename F>
__device__ inline void with_jrad(const QW_CONST JRad<R>& o, F&& f) {
  if (o.kind == 0) {
    f(CuspEval<R, LAP>{o});
    return;
  }
  switch (o.nf) {
    case 1: f(PolyEval<R, LAP, 1>{o}); break;
    case 2: f(PolyEval<R, LAP, 2>{o}); break;
    case 3: f(PolyEval<R, LAP, 3>{o}); break;
    case 4: f(PolyEval<R, LAP, 4>{o}); break;
    case 5: f(PolyEval<R, LAP, 5>{o}); break;
    case 6: f(PolyEval<R, LAP, 6>{o}); break;
    case 7: f(PolyEval<R, LAP, 7>{o}); break;
    default: f(PolyEval<R, LAP, kMaxJF>{o}); break;
  }
}

// Jastrow terms involving electron e placed at p (others at pos): value, gradient and,
// when lap != nullptr, laplacian with respect to r_e.  epos: atom positions in e-i type
// order with stride estride (3 in device memory, 4 in the LDS copy).  Wave-cooperative; the
// result is in all lanes.  R: arithmetic type (double, or float on the mixed path with
// distances from FP64 positions).
template <typename R, bool LAP>
__device__ inline R jastrow_e_t(const DevSys& S, const JRad<R>* ee, const JRad<R>* eit, int e, const double p[3],
                                const double* pos, R g[3], R* lap, int lane, const double* epos, int estride) {
  R u = 0, gx = 0, gy = 0, gz = 0, l = 0;
  // object loops outermost: each object's parameters are loop-invariant over the pairs
  const QW_CONST int* ea0 = cptr(S.eit_a0);
  const QW_CONST int* eo0 = cptr(S.eit_o0);
  for (int t = 0; t < S.n_eit; t++) {
    const int a0 = ea0[t], a1 = ea0[t + 1], o1 = eo0[t + 1];
    for (int o = eo0[t]; o < o1; o++) {
      with_jrad<R, LAP>(cptr(eit)[o], [&](const auto& ev) {
        for (int k = a0 + lane; k < a1; k += kWave) {
          const double* ap = epos + estride * k;
          const R x = (R)(p[0] - ap[0]), y = (R)(p[1] - ap[1]), z = (R)(p[2] - ap[2]);
          R r, ri;
          dist_inv(x * x + y * y + z * z, r, ri);
          R gf = 0;
          ev((R)0, r, ri, u, gf, l);
          gx += gf * x;
          gy += gf * y;
          gz += gf * z;
        }
      });
    }
  }
  const bool eup = e < S.nup;
  for (int o = 0; o < S.nee; o++) {
    with_jrad<R, LAP>(cptr(ee)[o], [&](const auto& ev) {
      for (int j = lane; j < S.ne; j += kWave) {
        if (j == e) continue;
        const R x = (R)(p[0] - pos[3 * j]), y = (R)(p[1] - pos[3 * j + 1]), z = (R)(p[2] - pos[3 * j + 2]);
        R r, ri;
        dist_inv(x * x + y * y + z * z, r, ri);
        const R uf = ((j < S.nup) != eup) ? (R)1 : (R)0;
        R gf = 0;
        ev(uf, r, ri, u, gf, l);
        gx += gf * x;
        gy += gf * y;
        gz += gf * z;
      }
    });
  }
  if constexpr (sizeof(R) == sizeof(double)) {
    u = wsum(u);
    g[0] = wsum(gx);
    g[1] = wsum(gy);
    g[2] = wsum(gz);
    if (LAP) *lap = wsum(l);
  } else {
    u = wsumf(u);
    g[0] = wsumf(gx);
    g[1] = wsumf(gy);
    g[2] = wsumf(gz);
    if (LAP) *lap = wsumf(l);
  }
  return u;
}
__device__ inline double jastrow_e(const DevSys& S, int e, const double p[3], const double* pos, double g[3],
                                   double* lap, int lane, const double* epos, int estride) {
  if (lap) return jastrow_e_t<double, true>(S, S.ee, S.eit, e, p, pos, g, lap, lane, epos, estride);
  return jastrow_e_t<double, false>(S, S.ee, S.eit, e, p, pos, g, nullptr, lane, epos, estride);
}
__device__ inline double jastrow_e(const DevSys& S, int e, const double p[3], const double* pos, double g[3],
                                   double* lap, int lane) {
  return jastrow_e(S, e, p, pos, g, lap, lane, S.eit_pos, 3);
}

// Value-only Jastrow terms of electron e at p (pseudopotential quadrature).
__device__ inline double jastrow_e_val(const DevSys& S, int e, const double p[3], const double* pos, int lane) {
  double g[3];
  return jastrow_e_t<double, false>(S, S.ee, S.eit, e, p, pos, g, nullptr, lane, S.eit_pos, 3);
}

// The same value for electron e at each of the points p[0..nq-1] (NQ >= nq), in one pass over
// the objects and pairs: each point's terms are added in the order jastrow_e_val adds them, so
// u[i] equals jastrow_e_val at p[i] bit for bit.
template <int NQ>
__device__ inline void jastrow_e_val_multi(const DevSys& S, int e, const double (*p)[3], int nq, const double* pos,
                                           double u[NQ], int lane) {
  double uq[NQ];
#pragma unroll
  for (int i = 0; i < NQ; i++) uq[i] = 0;
  const QW_CONST int* ea0 = cptr(S.eit_a0);
  const QW_CONST int* eo0 = cptr(S.eit_o0);
  for (int t = 0; t < S.n_eit; t++) {
    const int a0 = ea0[t], a1 = ea0[t + 1], o1 = eo0[t + 1];
    for (int o = eo0[t]; o < o1; o++) {
      with_jrad<double, false>(cptr(S.eit)[o], [&](const auto& ev) {
        for (int k = a0 + lane; k < a1; k += kWave) {
          const double* ap = S.eit_pos + 3 * k;
#pragma unroll
          for (int i = 0; i < NQ; i++) {
            if (i >= nq) break;
            const double x = p[i][0] - ap[0], y = p[i][1] - ap[1], z = p[i][2] - ap[2];
            double r, ri;
            dist_inv(x * x + y * y + z * z, r, ri);
            double gf = 0, l = 0;
            ev(0.0, r, ri, uq[i], gf, l);
          }
        }
      });
    }
  }
  const bool eup = e < S.nup;
  for (int o = 0; o < S.nee; o++) {
    with_jrad<double, false>(cptr(S.ee)[o], [&](const auto& ev) {
      for (int j = lane; j < S.ne; j += kWave) {
        if (j == e) continue;
        const double uf = ((j < S.nup) != eup) ? 1.0 : 0.0;
#pragma unroll
        for (int i = 0; i < NQ; i++) {
          if (i >= nq) break;
          const double x = p[i][0] - pos[3 * j], y = p[i][1] - pos[3 * j + 1], z = p[i][2] - pos[3 * j + 2];
          double r, ri;
          dist_inv(x * x + y * y + z * z, r, ri);
          double gf = 0, l = 0;
          ev(uf, r, ri, uq[i], gf, l);
        }
      }
    });
  }
#pragma unroll
  for (int i = 0; i < NQ; i++) u[i] = i < nq ? wsum(uq[i]) : 0.0;
}

// Semilocal pseudopotential channel values v_l(r) (src/basis_function/Rgaussian_function.cpp:180-202,
// src/system/Pseudopotential.cpp:305-318).
__device__ inline void pp_radial(const QW_CONST PPType& P, double Z, double r, double* vl) {
  for (int l = 0; l < P.L; l++) {
    double v = 0;
    for (int j = 0; j < P.nterm[l]; j++) {
      double ex = P.a[l][j] * (r * r);
      ex = ex < 60.0 ? ex : 60.0;
      v += P.c[l][j] * pow(r, P.n[l][j]) * exp(-ex);
    }
    vl[l] = v;
  }
  if (P.addzeff) vl[P.L - 1] += Z / r;
}

__device__ inline double legendre(double x, int n) {
  switch (n) {
    case 0: return 1;
    case 1: return x;
    case 2: return .5 * (3 * x * x - 1);
    case 3: return .5 * (5 * x * x * x - 3 * x);
    case 4: return 0.125 * (35 * x * x * x * x - 30 * x * x + 3);
    default: return 0.125 * (63 * x * x * x * x * x - 70 * x * x * x + 15 * x);
  }
}

}  // namespace qw
