// Mixed-precision path.
//
//   inverse of each spin channel and the MO cache in FP32; the inverse is held in VGPRs for a
//   sweep and updated by FP32 Sherman-Morrison;
//   LCAO either FP32 on the VALU (lane = MO) or on MFMA (v_mfma_f32_16x16x32_bf16) with the
//   operands split into BF16 terms and FP32 accumulation: x = x0 + x1 (+ x2), products kept
//   down to the chosen order (3 products: "bf16x3", 6 products: "bf16x6"; Henry, Tang and
//   Heinecke, "Leveraging the bfloat16 artificial intelligence datatype for higher-precision
//   computations", ARITH 2019; error-compensated splitting as in Ootomo and Yokota, 2022);
//   basis splines and Jastrow pair terms in FP32 (distances from FP64 positions);
//   FP64: positions, the periodic LU recompute of each inverse (every `refresh` moves),
//   log|det| accumulation, local-energy sums, DMC weights and branching.
#pragma once
#include "qw_kernels.hpp"

struct qw_engine;

namespace qw {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum LcaoMode { LCAO_FP32 = 0, LCAO_BF16X3 = 1, LCAO_BF16X6 = 2 };

// Device tables of the mixed path.
struct MixedTables {
  int ready;
  int ldchi;                  // basis functions per LDS row (nb rounded up to 32)
  int nkstep;                 // ceil(nb / 32)
  int ntile[2];               // ceil(nspin / 16)
  const float* spline;        // [ntab][nint][4] FP32
  const uint4* Cfrag[2][3];   // per spin, split term: [kstep][tile][64 lanes] x 8 bf16 (16 bytes)
  const float* Cf[2];         // dense FP32 [nb][nspin]
};

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }

// Split x into up to three BF16 terms with x ~= t0 + t1 + t2.
__device__ inline void split3(float x, __bf16& t0, __bf16& t1, __bf16& t2) {
  t0 = to_bf16(x);
  float r = x - (float)t0;
  t1 = to_bf16(r);
  r = r - (float)t1;
  t2 = to_bf16(r);
}

// FP32 basis values of shells [sh0, sh1) at point p (FP64), written as FP32 to chi[c*ld + f].
__device__ inline void basis_lap_f32(const DevSys& S, const MixedTables& M, const double p[3], int sh0, int sh1,
                                     int f0, float* chi, int ld, int lane) {
  for (int sh = sh0 + lane; sh < sh1; sh += kWave) {
    const ShellDesc d = S.shells[sh];
    const double* ap = S.atom_pos + 3 * d.atom;
    float x = (float)(p[0] - ap[0]), y = (float)(p[1] - ap[1]), z = (float)(p[2] - ap[2]);
    float r = sqrtf(x * x + y * y + z * z);
    float* o = chi + (d.f0 - f0);
    if (r < (float)d.rcut && r < (float)d.objthr) {
      float inv = (float)S.tab_inv[d.table];
      int iv = (int)(r * inv);
      iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
      const float4 c = *(const float4*)(M.spline + ((size_t)d.table * S.nint + iv) * 4);
      float h = r - iv * (float)S.tab_sp[d.table];
      float f = c.x + h * (c.y + h * (c.z + h * c.w));
      float fd = (c.y + h * (2.f * c.z + h * (3.f * c.w))) / r;
      float f2 = 2.f * c.z + 6.f * h * c.w;
      shell_lap<float>(d.sym, f, fd, f2, x, y, z, o, ld);
    } else {
      for (int i = 0; i < d.nf; i++)
        for (int c5 = 0; c5 < 5; c5++) o[c5 * ld + i] = 0.f;
    }
  }
}

// Sparse FP32 basis evaluation: active functions packed in increasing order (see
// basis_lap_sparse); returns their count.
__device__ inline int basis_lap_sparse_f32(const DevSys& S, const MixedTables& M, const double p[3], float* chi,
                                           int ld, int* flist, int lane) {
  int base = 0;
  for (int s0 = 0; s0 < S.nsh; s0 += kWave) {
    const int sh = s0 + lane;
    ShellDesc d;
    float x = 0, y = 0, z = 0, r = 0;
    bool act = false;
    if (sh < S.nsh) {
      d = S.shells[sh];
      const double* ap = S.atom_pos + 3 * d.atom;
      x = (float)(p[0] - ap[0]);
      y = (float)(p[1] - ap[1]);
      z = (float)(p[2] - ap[2]);
      r = sqrtf(x * x + y * y + z * z);
      act = r < (float)d.rcut && r < (float)d.objthr;
    }
    int tot;
    const int at = base + wave_excl_scan(act ? d.nf : 0, lane, &tot);
    if (act) {
      float inv = (float)S.tab_inv[d.table];
      int iv = (int)(r * inv);
      iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
      const float4 c = *(const float4*)(M.spline + ((size_t)d.table * S.nint + iv) * 4);
      float h = r - iv * (float)S.tab_sp[d.table];
      float f = c.x + h * (c.y + h * (c.z + h * c.w));
      float fd = (c.y + h * (2.f * c.z + h * (3.f * c.w))) / r;
      float f2 = 2.f * c.z + 6.f * h * c.w;
      shell_lap<float>(d.sym, f, fd, f2, x, y, z, chi + at, ld);
      for (int i = 0; i < d.nf; i++) flist[at + i] = d.f0 + i;
    }
    base += tot;
  }
  return base;
}

// FP32 Jastrow value and gradient of electron e at p (distances from FP64 positions);
// epos: atom positions in e-i type order with stride estride.
__device__ inline float jastrow_e_f(const DevSys& S, int e, const double p[3], const double* pos, float g[3], int lane,
                                    const double* epos, int estride) {
  return jastrow_e_t<float, false>(S, S.eef, S.eitf, e, p, pos, g, nullptr, lane, epos, estride);
}
__device__ inline float jastrow_e_f(const DevSys& S, int e, const double p[3], const double* pos, float g[3], int lane) {
  return jastrow_e_f(S, e, p, pos, g, lane, S.eit_pos, 3);
}

// The same FP32 Jastrow value and gradient for a sweep that keeps the electron positions in
// registers: lane l holds electrons l and l + 64 (pe[0], pe[1]; at most 128 electrons).  Each
// partner's distance is formed once and every object is evaluated on it (objects innermost),
// so the pair loop reads no LDS.  Same terms as jastrow_e_t; the per-lane sums run in another
// order.
__device__ inline float jastrow_e_freg(const DevSys& S, int e, const double p[3], const double (*pe)[3], float g[3],
                                       int lane, const double* epos, int estride) {
  float u = 0, gx = 0, gy = 0, gz = 0, l = 0;
  const QW_CONST int* ea0 = cptr(S.eit_a0);
  const QW_CONST int* eo0 = cptr(S.eit_o0);
  const QW_CONST JRad<float>* eit = cptr(S.eitf);
  const QW_CONST JRad<float>* ee = cptr(S.eef);
  for (int t = 0; t < S.n_eit; t++) {
    const int a0 = ea0[t], a1 = ea0[t + 1], o0 = eo0[t], o1 = eo0[t + 1];
    for (int k = a0 + lane; k < a1; k += kWave) {
      const double* ap = epos + estride * k;
      const float x = (float)(p[0] - ap[0]), y = (float)(p[1] - ap[1]), z = (float)(p[2] - ap[2]);
      float r, ri;
      dist_inv(x * x + y * y + z * z, r, ri);
      float gf = 0;
      for (int o = o0; o < o1; o++) with_jrad<float, false>(eit[o], [&](const auto& ev) { ev(0.f, r, ri, u, gf, l); });
      gx += gf * x;
      gy += gf * y;
      gz += gf * z;
    }
  }
  const bool eup = e < S.nup;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int j = lane + h * kWave;
    if (j < S.ne && j != e) {
      const float x = (float)(p[0] - pe[h][0]), y = (float)(p[1] - pe[h][1]), z = (float)(p[2] - pe[h][2]);
      float r, ri;
      dist_inv(x * x + y * y + z * z, r, ri);
      const float uf = ((j < S.nup) != eup) ? 1.f : 0.f;
      float gf = 0;
      for (int o = 0; o < S.nee; o++) with_jrad<float, false>(ee[o], [&](const auto& ev) { ev(uf, r, ri, u, gf, l); });
      gx += gf * x;
      gy += gf * y;
      gz += gf * z;
    }
  }
  u = wsumf(u);
  g[0] = wsumf(gx);
  g[1] = wsumf(gy);
  g[2] = wsumf(gz);
  return u;
}

// LCAO of one point in the mixed path.  On return lane j < n holds phi[c] (c = value,
// gradient, laplacian) of MO j.  chi: LDS tile 5 x ld floats (FP32 basis values);
// phis: LDS 5 x 64 floats used to move MFMA accumulators to lane = MO.
template <int MODE>
__device__ inline void lcao_lap_mx(const DevSys& S, const MixedTables& M, int s, const double p[3], float* chi,
                                   float* phis, int* flist, float phi[5], int lane) {
  const int n = s ? S.ndn : S.nup, ld = M.ldchi;
  for (int c = 0; c < 5; c++) phi[c] = 0.f;
  if (MODE == LCAO_FP32) {
    {
      __syncthreads();
      const int nact = basis_lap_sparse_f32(S, M, p, chi, ld, flist, lane);
      __syncthreads();
      if (lane < n) {
        const float* Cp = M.Cf[s] + lane;
        float a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
#pragma unroll 8
        for (int b = 0; b < nact; b++) {
          float c = Cp[(size_t)flist[b] * n];
          a0 += c * chi[b];
          a1 += c * chi[ld + b];
          a2 += c * chi[2 * ld + b];
          a3 += c * chi[3 * ld + b];
          a4 += c * chi[4 * ld + b];
        }
        phi[0] += a0;
        phi[1] += a1;
        phi[2] += a2;
        phi[3] += a3;
        phi[4] += a4;
      }
    }
    return;
  }
  // MFMA: D[c][j] += sum_k A[c][k] B[k][j], A = basis values (rows = 5 components, padded to
  // 16), B = coefficient fragments pre-packed per (kstep, tile, lane).  The LDS tile holds
  // all basis functions (ldchi >= nb).
  __syncthreads();
  basis_lap_f32(S, M, p, 0, S.nsh, 0, chi, ld, lane);
  __syncthreads();
  const int row = lane & 15, kq = (lane >> 4) * 8;
  const int ntile = M.ntile[s];
  f32x4 acc[4];
  for (int t = 0; t < 4; t++) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < M.nkstep; ks++) {
    bf16x8 a0, a1, a2;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      int b = ks * 32 + kq + i;
      float x = (row < 5 && b < S.nb) ? chi[row * ld + b] : 0.f;
      __bf16 t0, t1, t2;
      split3(x, t0, t1, t2);
      a0[i] = t0;
      a1[i] = t1;
      a2[i] = t2;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (t < ntile) {
        const size_t fi = ((size_t)ks * ntile + t) * 64 + lane;
        uint4 q0 = M.Cfrag[s][0][fi], q1 = M.Cfrag[s][1][fi];
        bf16x8 b0 = *reinterpret_cast<bf16x8*>(&q0), b1 = *reinterpret_cast<bf16x8*>(&q1);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[t], 0, 0, 0);
        if (MODE == LCAO_BF16X6) {
          uint4 q2 = M.Cfrag[s][2][fi];
          bf16x8 b2 = *reinterpret_cast<bf16x8*>(&q2);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b2, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b0, acc[t], 0, 0, 0);
        }
      }
    }
  }
  // accumulator layout: lane holds rows 4*(lane/16)+i (i<4), column lane%16 of each tile
  __syncthreads();
  const int rbase = (lane >> 4) * 4, col = lane & 15;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    if (t < ntile) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int r = rbase + i;
        if (r < 5) phis[r * 64 + t * 16 + col] = acc[t][i];
      }
    }
  }
  __syncthreads();
  if (lane < n)
    for (int c = 0; c < 5; c++) phi[c] = phis[c * 64 + lane];
}

// ---------------------------------------------------------------- k_mo_rows_mx
template <int MODE>
__global__ __launch_bounds__(64) void k_mo_rows_mx(DevSys S, MixedTables M, DevWalkers W, int nw) {
  extern __shared__ float ldsf[];
  const int lane = threadIdx.x;
  const long blk = blockIdx.x;
  const int w = (int)(blk / S.ne), e = (int)(blk % S.ne);
  if (w >= nw) return;
  const int s = e < S.nup ? 0 : 1, n = s ? S.ndn : S.nup;
  double p[3];
  for (int d = 0; d < 3; d++) p[d] = W.pos[((size_t)w * S.ne + e) * 3 + d];
  float phi[5];
  lcao_lap_mx<MODE>(S, M, s, p, ldsf, ldsf + 5 * M.ldchi, (int*)(ldsf + 5 * M.ldchi + 5 * 64), phi, lane);
  if (lane < n) {
    float* row = W.mof + ((size_t)w * S.ne + e) * 5 * S.nsp;
    for (int c = 0; c < 5; c++) row[c * S.nsp + lane] = phi[c];
  }
}

// ---------------------------------------------------------------- k_sweep_mx
// Basis functions per LDS tile batch in the FP32 sweep (one walker per block), and with the
// coefficients staged in LDS (several walkers per block, where LDS is shared by all of them).
constexpr int kCapF32 = 256;
constexpr int kCapCL = 64;
// Walkers per block of the LDS-staged variant: two waves per SIMD of one compute unit.
constexpr int kWpbCL = 8;
constexpr size_t kLdsMax = 163840;  // LDS per workgroup on gfx950

__host__ __device__ inline size_t lds_round16(size_t b) { return (b + 15) & ~(size_t)15; }

// Per-walker LDS of k_sweep_mx: positions | (several walkers per block: active-shell list) |
// basis tile (FP32 sparse: 5 x cap floats + cap row offsets; MFMA: 5 x ldchi floats + 5 x 64
// accumulator exchange + ldchi ints) | inverse row and value row broadcast (2 x NS floats).
__host__ __device__ inline size_t sweep_mx_walker_lds(int ne, int nsh, int ns, int mode, int ldchi, int wpb,
                                                      int depth) {
  const int cap = wpb > 1 ? kCapCL : kCapF32;
  const size_t tile = mode == LCAO_FP32 ? (size_t)5 * cap * 4 + (size_t)cap * 4
                                        : (size_t)(5 * ldchi + 5 * 64) * 4 + (size_t)ldchi * 4;
  return lds_round16((size_t)3 * ne * sizeof(double)) + (wpb > 1 ? lds_round16((size_t)nsh * 4) : 0) +
         sweep_rand_bytes(ns, depth) + tile + (size_t)2 * ns * 4 + 16;
}
// LDS of k_sweep_mx: (several walkers per block: the spin's coefficients, nb x nsp floats) |
// system tables | per-walker regions.
__host__ __device__ inline size_t sweep_mx_lds_bytes(int ne, int na, int nsh, int ntab, int ns, int mode, int ldchi,
                                                     int depth, int nb = 0, int nsp = 0, int wpb = 1) {
  return (wpb > 1 ? lds_round16((size_t)nb * nsp * 4) : 0) + lds_round16(sys_lds_bytes(na, nsh, ntab)) +
         (size_t)wpb * sweep_mx_walker_lds(ne, nsh, ns, mode, ldchi, wpb, depth);
}

// Moves of electrons [k0, k1) of spin s for WPB walkers per block (one wavefront each); all
// FP32 except positions, drift/proposal, acceptance and log|det| (FP64).  With WPB > 1 the
// block first stages the spin's MO coefficients in LDS, so the LCAO of every move reads LDS
// instead of L2; the walkers then run independently (wavefront-level barriers only).  The
// arithmetic is the same for every WPB and U.
// The sweep of walker w by wavefront `wid` of a block of WPB wavefronts (lds0: the block's LDS).
// STAGE: the block shares system tables and the LDS-staged coefficients (the layout above);
// otherwise each wavefront has a whole region of its own (sweep_mx_lds_bytes with one walker)
// and runs exactly as a one-walker block does, with wavefront-level barriers.
template <int NS, int DEPTH, int MODE, int U, int WPB, bool STAGE>
__device__ __forceinline__ void sweep_mx_body(const DevSys& S, const MixedTables& M, const DevWalkers& W,
                                              const SweepArgs& A, int w, int lane, int wid, char* lds0) {
  static_assert(WPB == 1 || MODE == LCAO_FP32, "several walkers per block: FP32 LCAO only");
  static_assert(!STAGE || WPB > 1, "staged coefficients are shared by several walkers");
  const int ne = S.ne, s = A.s, n = s ? S.ndn : S.nup, nsp = S.nsp, off = s ? S.nup : 0;
  char* base = lds0;
  float* Cl = (float*)base;
  if constexpr (STAGE) base += lds_round16((size_t)S.nb * nsp * 4);
  if constexpr (!STAGE && WPB > 1)
    base += (size_t)wid * sweep_mx_lds_bytes(ne, S.na, S.nsh, S.ntab, NS, MODE, M.ldchi, DEPTH);
  SysLds V;
  base = sys_lds_carve(S, base, V);
  if constexpr (STAGE) {
    // stage the coefficients (16-byte chunks while aligned) and the system tables
    const int tid = (int)threadIdx.x, nt = 64 * WPB, nc = S.nb * n;
    const float4* src4 = reinterpret_cast<const float4*>(M.Cf[s]);
    float4* dst4 = reinterpret_cast<float4*>(Cl);
    for (int i = tid; i < nc / 4; i += nt) dst4[i] = src4[i];
    for (int i = (nc / 4) * 4 + tid; i < nc; i += nt) Cl[i] = M.Cf[s][i];
    stage_sys(S, V, tid, nt);
    base += (size_t)wid * sweep_mx_walker_lds(ne, S.nsh, NS, MODE, M.ldchi, WPB, DEPTH);
    __syncthreads();
  } else {
    stage_sys(S, V, lane);
  }
  if (w >= W.nw) return;
  double* pos = (double*)base;
  char* q = base + lds_round16((size_t)3 * ne * sizeof(double));
  if constexpr (STAGE) {
    V.actl = (int*)q;
    q += lds_round16((size_t)S.nsh * 4);
  }
  constexpr int cap = STAGE ? kCapCL : kCapF32;
  double* rnd = (double*)q;
  q += sweep_rand_bytes(NS, DEPTH);
  float* chi = (float*)q;
  float* phis = chi + 5 * (MODE == LCAO_FP32 ? cap : M.ldchi);  // MFMA modes only
  int* flist = (int*)(MODE == LCAO_FP32 ? phis : phis + 5 * 64);
  float* rowb = (float*)(flist + (MODE == LCAO_FP32 ? cap : M.ldchi));
  float* phib = rowb + NS;
  for (int i = lane; i < 3 * ne; i += kWave) pos[i] = W.pos[(size_t)w * ne * 3 + i];
  sweep_randoms<DEPTH>(A, (unsigned)w, off, rnd, lane);
  // electron positions of this lane's Jastrow partners, kept in registers for the sweep
  double pe[2][3];
#pragma unroll
  for (int h = 0; h < 2; h++)
    for (int d = 0; d < 3; d++) {
      const int j = lane + h * kWave;
      pe[h][d] = j < ne ? W.pos[((size_t)w * ne + j) * 3 + d] : 0.0;
    }
  float inv[NS];
  const float* ig = W.invf + ((size_t)w * 2 + s) * nsp * nsp;
  inv_row_load<NS>(ig, lane, n, nsp, inv);
  double logdet = W.logdet[2 * w + s], sgn = W.sgn[2 * w + s];
  int nacc = 0;
  double dprod = 1.0;
  float* mog = W.mof + (size_t)w * ne * 5 * nsp;
  const double tau = A.tau, stau = sqrt(tau);
  float m1 = 0, m2 = 0, m3 = 0;
  if (lane < n && A.k0 < A.k1) {
    const size_t e0 = off + A.k0;
    m1 = mog[(e0 * 5 + 1) * nsp + lane];
    m2 = mog[(e0 * 5 + 2) * nsp + lane];
    m3 = mog[(e0 * 5 + 3) * nsp + lane];
  }
  walker_sync<WPB>();
  for (int k = A.k0; k < A.k1; k++) {
    const int e = off + k;
    double r0[3] = {pos[3 * e], pos[3 * e + 1], pos[3 * e + 2]};
    if (lane == k) {
#pragma unroll
      for (int j = 0; j < NS; j++) rowb[j] = inv[j];
    }
    walker_sync<WPB>();
    const float rk = lane < n ? rowb[lane] : 0.f;
    float gj0[3];
    const float u0 = jastrow_e_freg(S, e, r0, pe, gj0, lane, V.epos, 4);
    double d0[3] = {(double)(wsumf(rk * m1) + gj0[0]), (double)(wsumf(rk * m2) + gj0[1]), (double)(wsumf(rk * m3) + gj0[2])};
    float phi[5];
    double r1[3], d1[3], ln1 = 0;
    float R = 0;
    bool accepted = false;
    double acc_first = 0;
    for (int dep = 1; dep <= DEPTH && !accepted; dep++) {
      const double t = dep == 1 ? tau : tau * 0.5, st = dep == 1 ? stau : sqrt(tau * 0.5);
      const double* rq = rnd + ((size_t)(k - A.k0) * DEPTH + (dep - 1)) * 4;
      double dr[3], rn[3], dn[3];
      lim_drift(d0, t, dr);
      for (int d = 0; d < 3; d++) rn[d] = r0[d] + (rq[d] * st + dr[d]);
      if constexpr (MODE == LCAO_FP32 && STAGE)
        mo_lap_point<float, float, U, WPB>(S, V, M.spline, (const float*)Cl, n, rn, chi, flist, cap, phi, lane);
      else if constexpr (MODE == LCAO_FP32)
        mo_lap_point<float, float, U, WPB>(S, V, M.spline, M.Cf[s], n, rn, chi, flist, cap, phi, lane);
      else
        lcao_lap_mx<MODE>(S, M, s, rn, chi, phis, flist, phi, lane);
      float t0 = wsumf(rk * phi[0]);
      float t1 = wsumf(rk * phi[1]), t2 = wsumf(rk * phi[2]), t3 = wsumf(rk * phi[3]);
      float gj[3];
      float un = jastrow_e_freg(S, e, rn, pe, gj, lane, V.epos, 4);
      float it0 = 1.f / t0;
      dn[0] = (double)(t1 * it0 + gj[0]);
      dn[1] = (double)(t2 * it0 + gj[1]);
      dn[2] = (double)(t3 * it0 + gj[2]);
      const bool neg = t0 < 0;
      double acc, lnn = 0;
      if constexpr (DEPTH == 1) {
        // |ratio|^2 exp(dtrans) as t0^2 exp(2 du + dtrans): one exponential, no logarithm
        const double td = (double)t0;
        acc = (A.restrict_nodes && neg)
                  ? 0.0
                  : fmin(1.0, td * td * exp(2.0 * (double)(un - u0) + (trans(rn, dn, r0, t) - trans(r0, d0, rn, t))));
      } else if (dep == 1) {
        lnn = log(fabs((double)t0)) + (double)(un - u0);
        double rr = exp(lnn);
        acc = (A.restrict_nodes && neg) ? 0.0 : fmin(1.0, rr * rr * exp(trans(rn, dn, r0, t) - trans(r0, d0, rn, t)));
        acc_first = acc;
        for (int d = 0; d < 3; d++) r1[d] = rn[d], d1[d] = dn[d];
        ln1 = lnn;
      } else {
        lnn = log(fabs((double)t0)) + (double)(un - u0);
        double rr = exp(lnn);
        double pt = (trans(rn, dn, r1, tau) - trans(r0, d0, r1, tau)) + (trans(rn, dn, r0, t) - trans(r0, d0, rn, t));
        double r21 = exp(ln1 - lnn);
        double a21 = fmin(1.0, r21 * r21 * exp(trans(r1, d1, rn, tau) - trans(rn, dn, r1, tau)));
        double num = 1 - a21, den = 1 - acc_first;
        if (A.restrict_nodes && neg) acc = 0;
        else if (fabs(den) < 1e-10) acc = fabs(num) < 1e-10 ? 0.0 : 1.0;
        else acc = fmin(1.0, rr * rr * exp(pt) * num / den);
      }
      if (acc + rq[3] > 1.0) {
        accepted = true;
        R = t0;
        for (int d = 0; d < 3; d++) r1[d] = rn[d];
      }
    }
    if (accepted) {
      if (lane < NS) phib[lane] = lane < n ? phi[0] : 0.f;
      walker_sync<WPB>();
      if (lane < n) {
        // four independent partial sums: the dot product is not one serial FMA chain
        float tq[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NS; j++) tq[j & 3] += inv[j] * phib[j];
        const float tk = (tq[0] + tq[1]) + (tq[2] + tq[3]);
        if (lane == k) {
          const float f = 1.f / R;
#pragma unroll
          for (int j = 0; j < NS; j++) inv[j] = rowb[j] * f;
        } else {
          const float pk = tk * (-1.f / R);
#pragma unroll
          for (int j = 0; j < NS; j++) inv[j] += rowb[j] * pk;
        }
        for (int c = 0; c < 5; c++) mog[((size_t)e * 5 + c) * nsp + lane] = phi[c];
      }
      // accepted ratios multiply into dprod, folded into log|det| before it could overflow
      dprod *= (double)R;
      if (!(fabs(dprod) < 0x1p+512 && fabs(dprod) > 0x1p-512)) {
        logdet += log(fabs(dprod));
        dprod = 1.0;
      }
      if (R < 0) sgn = -sgn;
      nacc++;
      if (lane < 3) pos[3 * e + lane] = r1[lane];
      if (lane == (e & (kWave - 1))) {
        if (e < kWave)
          for (int d = 0; d < 3; d++) pe[0][d] = r1[d];
        else
          for (int d = 0; d < 3; d++) pe[1][d] = r1[d];
      }
    }
    if (lane < n && k + 1 < A.k1) {
      const size_t e1 = e + 1;
      m1 = mog[(e1 * 5 + 1) * nsp + lane];
      m2 = mog[(e1 * 5 + 2) * nsp + lane];
      m3 = mog[(e1 * 5 + 3) * nsp + lane];
    }
    walker_sync<WPB>();
  }
  float* igw = W.invf + ((size_t)w * 2 + s) * nsp * nsp;
  inv_row_store<NS>(igw, lane, n, nsp, inv);
  if (A.dets) write_det_parts<NS>(S, W, w, s, inv, (const float*)mog, lane);
  for (int i = lane; i < 3 * ne; i += kWave) W.pos[(size_t)w * ne * 3 + i] = pos[i];
  logdet += log(fabs(dprod));
  if (lane == 0) {
    W.logdet[2 * w + s] = logdet;
    W.sgn[2 * w + s] = sgn;
    W.acc[w] += nacc;
  }
}

template <int NS, int DEPTH, int MODE, int U, int WPB>
__global__ __launch_bounds__(64 * WPB, WPB == 1 ? 2 : 1) void k_sweep_mx(DevSys S, MixedTables M, DevWalkers W,
                                                                          SweepArgs A) {
  extern __shared__ double ldsd[];
  const int lane = (int)threadIdx.x & (kWave - 1), wid = WPB == 1 ? 0 : (int)(threadIdx.x / kWave);
  sweep_mx_body<NS, DEPTH, MODE, U, WPB, (WPB > 1)>(S, M, W, A, (int)blockIdx.x * WPB + wid, lane, wid,
                                                    (char*)ldsd);
}

}  // namespace qw

void mixed_setup(qw_engine& e);
void mixed_refresh_all(qw_engine& e);
void mixed_sweep(qw_engine& e, const qw::SweepArgs& A, int depth);
void mixed_local(qw_engine& e, const qw::LocalArgs& A);
