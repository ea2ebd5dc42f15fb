// Walker kernels for more than 64 electrons of one spin (up to 256).
//
//   k_sweep_w<TI, R, DEPTH, NR>  moves of one spin channel by one 64-lane wavefront per walker
//                       (src/method/Split_sample.cpp:302-540,
//                       src/program_utils/MatrixAlgebrac.cpp:766-838): lane l owns rows and MOs
//                       l, l+64, ... (NR of them); the inverse stays in device memory (too
//                       large for one wavefront's registers) and each Sherman-Morrison update
//                       streams the lane's rows; TI = inverse / MO cache element type, R =
//                       Jastrow arithmetic
//   k_mo_rows_big<T>    basis + LCAO of every electron (src/orbitals/MO_matrix_cutoff.h:480-560)
//   k_inv_grid<G>       FP64 pivoted Gauss-Jordan over a G x G thread grid (up to 128 per spin)
//   k_inv_panel         the same Gauss-Jordan for 129 to 256 per spin, panel by panel with
//                       the working matrix in device memory (src/program_utils/MatrixAlgebrac.cpp:
//                       108-224,316-421)
#pragma once
#include "qw_kernels.hpp"

namespace qw {

constexpr int kBigMaxSpin = 256;
// Electrons of a spin up to which the inverse recompute is k_inv_grid<16> (the whole matrix in
// registers); above, k_inv_panel.
constexpr int kGridMaxSpin = 128;

// Spline table element type that goes with an inverse / MO cache element type.
template <typename TI>
struct SplineOf {
  using type = double;
};
template <>
struct SplineOf<float> {
  using type = float;
};
template <typename TI>
using TS_SPLINE_T = typename SplineOf<TI>::type;

// Jastrow objects of the arithmetic type R.
template <typename R>
__device__ inline const JRad<R>* jee(const DevSys& S);
template <>
__device__ inline const JRad<double>* jee<double>(const DevSys& S) { return S.ee; }
template <>
__device__ inline const JRad<float>* jee<float>(const DevSys& S) { return S.eef; }
template <typename R>
__device__ inline const JRad<R>* jei(const DevSys& S);
template <>
__device__ inline const JRad<double>* jei<double>(const DevSys& S) { return S.eit; }
template <>
__device__ inline const JRad<float>* jei<float>(const DevSys& S) { return S.eitf; }

// Pointers of one precision for the big kernels.
template <typename TI>
struct BigPtrs {
  TI* inv;           // [nw][2][nsp][nsp]
  TI* mo;            // [nw][ne][5][nsp]
  const TI* C[2];    // [nb][nspin] coefficients
  const void* spline;  // double or float table, matching TI
};

// ---------------------------------------------------------------- one wavefront per walker
// The same moves with one 64-lane wavefront per walker: lane l owns rows and MOs l, l+64, ...
// (NR of them, nsigma <= 64 NR), so the whole move runs without block barriers and a compute
// unit holds several walkers per SIMD.  The inverse stays in device memory (row-major, as the
// recompute writes it); row k and the accepted point's MO values go through LDS.

// LDS of one walker: positions | system tables | basis tile | inverse row k and new MO values.
__host__ __device__ inline size_t sweep_w_lds_bytes(int ne, int na, int nsh, int ntab, int nsp, int cap, int tsize) {
  return ((((size_t)3 * ne * sizeof(double) + 15) & ~(size_t)15) + ((sys_lds_bytes(na, nsh, ntab) + 15) & ~(size_t)15) +
          (size_t)5 * cap * tsize + (size_t)cap * sizeof(int) + (size_t)2 * nsp * tsize + 15) &
         ~(size_t)15;
}

// The sweep of walker w by one wavefront (base: its LDS region of sweep_w_lds_bytes); WPB > 1:
// other wavefronts of the block run other walkers (wavefront-level barriers only).
template <typename TI, typename R, int DEPTH, int NR, int WPB>
__device__ __forceinline__ void sweep_w_body(const DevSys& S, const DevWalkers& W, const BigPtrs<TI>& P,
                                             const SweepArgs& A, int cap, int w, int lane, char* base) {
  if (w >= W.nw) return;
  const int ne = S.ne, s = A.s, n = s ? S.ndn : S.nup, nsp = S.nsp, off = s ? S.nup : 0;
  // rows of the inverse and of the MO cache start 16-byte aligned when nsp is a multiple of
  // the values per 16 bytes
  constexpr int per = 16 / sizeof(TI);
  typedef TI tvr __attribute__((ext_vector_type(per)));
  const bool vrow = nsp % per == 0;
  double* pos = (double*)base;
  SysLds V;
  char* q = sys_lds_carve(S, base + (((size_t)3 * ne * sizeof(double) + 15) & ~(size_t)15), V);
  TI* chi = (TI*)q;
  int* roff = (int*)(chi + 5 * cap);
  TI* rowk = (TI*)(roff + cap);
  TI* phi0 = rowk + nsp;
  for (int i = lane; i < 3 * ne; i += kWave) pos[i] = W.pos[(size_t)w * ne * 3 + i];
  stage_sys(S, V, lane);
  TI* ig = P.inv + ((size_t)w * 2 + s) * nsp * nsp;
  TI* mog = P.mo + (size_t)w * ne * 5 * nsp;
  const TS_SPLINE_T<TI>* spline = (const TS_SPLINE_T<TI>*)P.spline;
  double logdet = W.logdet[2 * w + s], sgn = W.sgn[2 * w + s];
  int nacc = 0;
  const double tau = A.tau, stau = sqrt(tau);
  walker_sync<WPB>();
  for (int k = A.k0; k < A.k1; k++) {
    const int e = off + k;
    const double r0[3] = {pos[3 * e], pos[3 * e + 1], pos[3 * e + 2]};
    // row k of the inverse: lane l holds its entries l + 64 r
    double rk[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int j = lane + kWave * r;
      const TI v = j < n ? ig[(size_t)k * nsp + j] : (TI)0;
      rk[r] = (double)v;
      if (j < nsp) rowk[j] = v;
    }
    double gd[3] = {0, 0, 0};
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int j = lane + kWave * r;
      if (j < n)
        for (int c = 0; c < 3; c++) gd[c] += rk[r] * (double)mog[((size_t)e * 5 + 1 + c) * nsp + j];
    }
    for (int c = 0; c < 3; c++) gd[c] = wsum(gd[c]);
    R g0[3];
    const double u0 = (double)jastrow_e_t<R, false>(S, jee<R>(S), jei<R>(S), e, r0, pos, g0, nullptr, lane, V.epos, 4);
    const double d0[3] = {gd[0] + (double)g0[0], gd[1] + (double)g0[1], gd[2] + (double)g0[2]};
    TI acc[NR][5];
    double r1[3], d1[3], ln1 = 0, Rt = 0;
    bool accepted = false;
    double acc_first = 0;
    for (int dep = 1; dep <= DEPTH && !accepted; dep++) {
      const double tt = dep == 1 ? tau : tau * 0.5, st = dep == 1 ? stau : sqrt(tau * 0.5);
      double g[3], dr[3], rn[3], dn[3];
      gauss3(A.seed, (unsigned)w, step_of(A.step, A.stepp), (unsigned)e, (dep == 1 ? TAG_GAUSS1 : TAG_GAUSS2) | (A.sweep << 8), g);
      lim_drift(d0, tt, dr);
      for (int d = 0; d < 3; d++) rn[d] = r0[d] + (g[d] * st + dr[d]);
#pragma unroll
      for (int r = 0; r < NR; r++)
        for (int c = 0; c < 5; c++) acc[r][c] = (TI)0;
      const int nact = basis_active(S, V, rn, lane);
      walker_sync<WPB>();
      for (int i0 = 0; i0 < nact;) {
        int nfun;
        const int take =
            basis_batch_lap<TI, TS_SPLINE_T<TI>>(S, V, spline, rn, i0, nact, n, chi, cap, roff, cap, &nfun, lane);
        walker_sync<WPB>();
#pragma unroll
        for (int r = 0; r < NR; r++)
          if (lane + kWave * r < n) lcao_batch<TI, 4>(P.C[s] + kWave * r, roff, chi, cap, nfun, acc[r], lane);
        walker_sync<WPB>();
        i0 += take;
      }
      double tv[4] = {0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < NR; r++)
        for (int c = 0; c < 4; c++) tv[c] += rk[r] * (double)acc[r][c];
      for (int c = 0; c < 4; c++) tv[c] = wsum(tv[c]);
      const double t0 = tv[0];
      R g2[3];
      const double un = (double)jastrow_e_t<R, false>(S, jee<R>(S), jei<R>(S), e, rn, pos, g2, nullptr, lane, V.epos, 4);
      dn[0] = tv[1] / t0 + (double)g2[0];
      dn[1] = tv[2] / t0 + (double)g2[1];
      dn[2] = tv[3] / t0 + (double)g2[2];
      const double lnn = log(fabs(t0)) + (un - u0);
      const bool neg = t0 < 0;
      double acc_p;
      if (dep == 1) {
        const double rr = exp(lnn);
        acc_p = (A.restrict_nodes && neg) ? 0.0
                                          : fmin(1.0, rr * rr * exp(trans(rn, dn, r0, tt) - trans(r0, d0, rn, tt)));
        acc_first = acc_p;
        for (int d = 0; d < 3; d++) r1[d] = rn[d], d1[d] = dn[d];
        ln1 = lnn;
      } else {
        const double rr = exp(lnn);
        const double pt =
            (trans(rn, dn, r1, tau) - trans(r0, d0, r1, tau)) + (trans(rn, dn, r0, tt) - trans(r0, d0, rn, tt));
        const double r21 = exp(ln1 - lnn);
        const double a21 = fmin(1.0, r21 * r21 * exp(trans(r1, d1, rn, tau) - trans(rn, dn, r1, tau)));
        const double num = 1 - a21, den = 1 - acc_first;
        if (A.restrict_nodes && neg) acc_p = 0;
        else if (fabs(den) < 1e-10) acc_p = fabs(num) < 1e-10 ? 0.0 : 1.0;
        else acc_p = fmin(1.0, rr * rr * exp(pt) * num / den);
      }
      const double uu =
          uniform(A.seed, (unsigned)w, step_of(A.step, A.stepp), (unsigned)e, (dep == 1 ? TAG_ACC1 : TAG_ACC2) | (A.sweep << 8));
      if (acc_p + uu > 1.0) {
        accepted = true;
        Rt = t0;
        for (int d = 0; d < 3; d++) r1[d] = rn[d];
      }
    }
    if (accepted) {
      // Sherman-Morrison on the transposed inverse in device memory: lane l owns rows l + 64 r
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const int j = lane + kWave * r;
        if (j < nsp) phi0[j] = j < n ? acc[r][0] : (TI)0;
      }
      walker_sync<WPB>();
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const int i = lane + kWave * r;
        if (i >= n) continue;
        TI* row = ig + (size_t)i * nsp;
        // 16-byte accesses to the lane's row where the rows start 16-byte aligned (the same
        // arithmetic in the same order as the one-value loop that finishes the row)
        const int jv = vrow ? n / per * per : 0;
        if (i == k) {
          const double f = 1.0 / Rt;
          for (int j = 0; j < jv; j += per) {
            tvr u;
#pragma unroll
            for (int q = 0; q < per; q++) u[q] = (TI)((double)rowk[j + q] * f);
            *(tvr*)(row + j) = u;
          }
          for (int j = jv; j < n; j++) row[j] = (TI)((double)rowk[j] * f);
        } else {
          double tk = 0;
          for (int j = 0; j < jv; j += per) {
            const tvr u = *(const tvr*)(row + j);
#pragma unroll
            for (int q = 0; q < per; q++) tk += (double)u[q] * (double)phi0[j + q];
          }
          for (int j = jv; j < n; j++) tk += (double)row[j] * (double)phi0[j];
          const double pk = tk * (-1.0 / Rt);
          for (int j = 0; j < jv; j += per) {
            tvr u = *(const tvr*)(row + j);
#pragma unroll
            for (int q = 0; q < per; q++) u[q] = (TI)((double)u[q] + (double)rowk[j + q] * pk);
            *(tvr*)(row + j) = u;
          }
          for (int j = jv; j < n; j++) row[j] = (TI)((double)row[j] + (double)rowk[j] * pk);
        }
        for (int c = 0; c < 5; c++) mog[((size_t)e * 5 + c) * nsp + i] = acc[r][c];
      }
      logdet += log(fabs(Rt));
      if (Rt < 0) sgn = -sgn;
      nacc++;
      if (lane < 3) pos[3 * e + lane] = r1[lane];
    }
    // the device-memory rows written above are read by other lanes (row k) on the next move
    __threadfence_block();
    walker_sync<WPB>();
  }
  for (int i = lane; i < 3 * ne; i += kWave) W.pos[(size_t)w * ne * 3 + i] = pos[i];
  if (lane == 0) {
    W.logdet[2 * w + s] = logdet;
    W.sgn[2 * w + s] = sgn;
    W.acc[w] += nacc;
  }
  if (A.dets) {
    // determinant parts of the local energy for the lane's electrons
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int t = lane + kWave * r;
      if (t >= n) continue;
      const TI* row = ig + (size_t)t * nsp;
      const TI* mr = mog + (size_t)(off + t) * 5 * nsp;
      double dv[4] = {0, 0, 0, 0};
      const int jv = vrow ? n / per * per : 0;
      for (int j = 0; j < jv; j += per) {
        const tvr u = *(const tvr*)(row + j);
        tvr m[4];
#pragma unroll
        for (int c = 0; c < 4; c++) m[c] = *(const tvr*)(mr + (size_t)(c + 1) * nsp + j);
#pragma unroll
        for (int q = 0; q < per; q++) {
          const double iv = (double)u[q];
#pragma unroll
          for (int c = 0; c < 4; c++) dv[c] += (double)m[c][q] * iv;
        }
      }
      for (int j = jv; j < n; j++) {
        const double iv = (double)row[j];
        for (int c = 0; c < 4; c++) dv[c] += (double)mr[(size_t)(c + 1) * nsp + j] * iv;
      }
      double* o = W.dgl + ((size_t)w * ne + off + t) * 4;
      for (int c = 0; c < 4; c++) o[c] = dv[c];
    }
  }
}

template <typename TI, typename R, int DEPTH, int NR>
__global__ __launch_bounds__(64) void k_sweep_w(DevSys S, DevWalkers W, BigPtrs<TI> P, SweepArgs A, int cap) {
  extern __shared__ double lds[];
  sweep_w_body<TI, R, DEPTH, NR, 1>(S, W, P, A, cap, (int)blockIdx.x, (int)threadIdx.x, (char*)lds);
}

// Basis + LCAO of electron e of walker w (one 256-thread block per electron): wavefront 0
// evaluates the active functions chunk by chunk, every thread accumulates its MO.
template <typename T>
__global__ __launch_bounds__(256) void k_mo_rows_big(DevSys S, const double* wpos, const T* C0, const T* C1,
                                                     T* mo_out, int nw) {
  extern __shared__ double lds[];
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const long blk = blockIdx.x;
  const int w = (int)(blk / S.ne), e = (int)(blk % S.ne);
  if (w >= nw) return;
  const int s = e < S.nup ? 0 : 1, n = s ? S.ndn : S.nup, ld = S.chunk_nf, nsp = S.nsp;
  const T* C = s ? C1 : C0;
  T* chi = (T*)lds;
  int* flist = (int*)(chi + 5 * ld);
  int* nb = flist + ld;
  double p[3];
  for (int d = 0; d < 3; d++) p[d] = wpos[((size_t)w * S.ne + e) * 3 + d];
  T acc[5] = {0, 0, 0, 0, 0};
  for (int ch = 0; ch < S.nchunk; ch++) {
    const int sh0 = S.chunk_sh0[ch], sh1 = S.chunk_sh0[ch + 1];
    __syncthreads();
    if (wv == 0) {
      const int na_ = basis_lap_sparse<T>(S, p, sh0, sh1, chi, ld, flist, lane);
      if (lane == 0) *nb = na_;
    }
    __syncthreads();
    const int nact = *nb;
    if (t < n) {
      const T* Cp = C + t;
      for (int b = 0; b < nact; b++) {
        const T c = Cp[(size_t)flist[b] * n];
        acc[0] += c * chi[b];
        acc[1] += c * chi[ld + b];
        acc[2] += c * chi[2 * ld + b];
        acc[3] += c * chi[3 * ld + b];
        acc[4] += c * chi[4 * ld + b];
      }
    }
  }
  if (t < n) {
    T* row = mo_out + ((size_t)w * S.ne + e) * 5 * nsp;
    for (int c = 0; c < 5; c++) row[(size_t)c * nsp + t] = acc[c];
  }
}

// ---------------------------------------------------------------- k_inv_grid
// The pivoted Gauss-Jordan inversion of k_inv_blk for up to 8G electrons of a spin
// (src/program_utils/MatrixAlgebrac.cpp:108-224,316-421: the pivot sequence of the reference's
// Crout ludcmp), with the matrix spread over a G x G grid of threads, one walker-spin per block:
// thread (bi, bj) = (t / G, t % G) holds rows 8bi..8bi+7 and columns 8bj..8bj+7 as one
// register vector per column, as a lane of k_inv_blk does.  A step: the threads of column
// block J propose their best unpivoted row (largest vv[i]*|a[i][j]|, last row on ties), each
// wavefront keeps its highest-lane winner, the wavefronts' winners meet in LDS (the highest row
// wins ties), then the pivot row and the multipliers go through LDS and every thread updates
// its block with one FMA per element.  Every element sees the operations of k_inv_blk in the
// same order, so the two agree bit for bit on matrices both handle; G = 16 (256 threads)
// inverts up to 128 electrons of a spin in registers.
// The pivot row and the multipliers are stored interleaved (entry 8i + b at index G*b + i), so
// the eight values a thread reads in a step are spread over the LDS banks.
template <int G>
struct InvGridLds {
  double prowL[2][8 * G];
  double gL[2][8 * G];
  double pvs[8 * G];
  int pl[8 * G];
  int qinv[8 * G];
  double wbest[2][G * G / 64], wraw[2][G * G / 64];
  int wrow[2][G * G / 64];
  double redd[G * G / 64];
  int redi[G * G / 64];
};
// The inversion of walker w's spin s by a G*G-thread block (t = thread).
template <int G, typename TIN, typename TI>
__device__ __forceinline__ void k_inv_grid_body(const DevSys& S, const TIN* mo, TI* inv_out, double* logdet_out,
                                                double* sgn_out, int w, int s, int t, InvGridLds<G>& Ls) {
  static_assert(G == 8 || G == 16, "a row block's G threads share one 16-lane DPP row");
  constexpr int NT = G * G, NWV = NT / 64, N8 = 8 * G;
  typedef double dv8 __attribute__((ext_vector_type(8)));
  auto& prowL = Ls.prowL;
  auto& gL = Ls.gL;
  double* pvs = Ls.pvs;
  int* pl = Ls.pl;
  int* qinv = Ls.qinv;
  auto& wbest = Ls.wbest;
  auto& wraw = Ls.wraw;
  auto& wrow = Ls.wrow;
  double* redd = Ls.redd;
  int* redi = Ls.redi;
  const int lane = t & 63, wv = t >> 6, bi = t / G, bj = t % G;
  const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0, nsp = S.nsp;
  TI* out = inv_out + ((size_t)w * 2 + s) * nsp * nsp;
  if (n == 0) {
    if (t == 0) {
      logdet_out[2 * w + s] = 0;
      sgn_out[2 * w + s] = 1;
    }
    return;
  }
  dv8 a[8];
  double vv[8];
  bool zero_row = false;
  int pivm = 0;  // bit r: row 8bi + r already a pivot (rows >= n count as such)
  // a thread's 8 entries of a row are contiguous: 16-byte loads when every MO row starts on a
  // 16-byte boundary (entries past n are read, then replaced)
  const bool vec16 = sizeof(TIN) == 4 ? (nsp & 3) == 0 : (nsp & 1) == 0;
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const int row = 8 * bi + r;
    const TIN* rp = mo + ((size_t)w * S.ne + off + (row < n ? row : 0)) * 5 * nsp + 8 * bj;
    TIN vin[8];
    if (vec16) {
      constexpr int per = 16 / sizeof(TIN);
      typedef TIN tv16 __attribute__((ext_vector_type(per)));
#pragma unroll
      for (int h = 0; h < 8 / per; h++) {
        const tv16 u = *(const tv16*)(rp + per * h);
#pragma unroll
        for (int i = 0; i < per; i++) vin[per * h + i] = u[i];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 8; c++) vin[c] = 8 * bj + c < n ? rp[c] : (TIN)0;
    }
    double mx = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int col = 8 * bj + c;
      const double v = (row < n && col < n) ? (double)vin[c] : (row == col ? 1.0 : 0.0);
      a[c][r] = v;
      mx = fmax(mx, fabs(v));
    }
    // row maximum over the G threads of the row block (a contiguous group of G lanes)
    mx = fmax(mx, dpp_d<kDppXor1>(mx));
    mx = fmax(mx, dpp_d<kDppXor2>(mx));
    mx = fmax(mx, dpp_d<kDppHalfMirror>(mx));
    if constexpr (G == 16) mx = fmax(mx, dpp_d<kDppMirror>(mx));
    vv[r] = mx == 0.0 ? 0.0 : 1.0 / mx;
    zero_row |= row < n && mx == 0.0;
    if (row >= n) pivm |= 1 << r;
  }
  if (lane == 0) redi[wv] = wave_any(zero_row) ? 1 : 0;
  __syncthreads();
  bool singular = false;
#pragma unroll
  for (int k = 0; k < NWV; k++) singular |= redi[k] != 0;
  bool done = singular;
  for (int J = 0; J < G && !done; J++) {
    const bool mine = bj == J;  // this thread holds column 8J + c of its rows
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int j = 8 * J + c;
      if (done || j >= n) {
        done = true;
        break;
      }
      double best = -1.0, braw = 0.0;
      int br = 0;
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const double x = a[c][r];
        const double cd = (mine && !((pivm >> r) & 1)) ? vv[r] * fabs(x) : -1.0;
        if (cd >= best) {
          best = cd;
          br = r;
          braw = x;
        }
      }
      // this wavefront's winner (highest lane with the maximum), then the wavefronts'
      const int L = wave_argmax_last(best, lane);
      const int Lc = L < 0 ? 0 : L;
      const double bb = wave_bcast(best, Lc), bw = wave_bcast(braw, Lc);
      const int brr = __builtin_amdgcn_readlane(br, Lc);
      if (lane == 0) {
        wbest[j & 1][wv] = L < 0 ? __builtin_nan("") : bb;
        wraw[j & 1][wv] = L < 0 ? 0.0 : bw;
        wrow[j & 1][wv] = L < 0 ? -1 : 8 * ((64 * wv + L) / G) + brr;
      }
      __syncthreads();
      double gb = -2.0;
      int gw = -1;
#pragma unroll
      for (int k = 0; k < NWV; k++) {
        const double b = wbest[j & 1][k];
        if (b >= gb) {
          gb = b;
          gw = k;
        }
      }
      const double pv = gw < 0 ? 0.0 : wraw[j & 1][gw];
      if (gw < 0 || wrow[j & 1][gw] < 0 || !(gb >= 0.0) || pv == 0.0) {
        singular = done = true;
        break;
      }
      const int p = wrow[j & 1][gw];
      const int pb = p >> 3, pr = p & 7;
      if (t == 0) {
        pl[j] = p;
        pvs[j] = pv;
      }
      double* prow = prowL[j & 1];
      double* gw8 = gL[j & 1];
      if (bi == pb) {
#pragma unroll
        for (int c2 = 0; c2 < 8; c2++) prow[G * c2 + bj] = a[c2][pr];
      }
      const double rpv = 1.0 / pv;
      const bool pblk = bi == pb;  // row 8bi + r is the pivot row p = 8pb + pr
      if (mine) {
#pragma unroll
        for (int r = 0; r < 8; r++) gw8[G * r + bi] = (pblk && r == pr) ? 0.0 : a[c][r] * rpv;
      }
      __syncthreads();
      double pr8[8], g8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        pr8[k] = prow[G * k + bj];
        g8[k] = gw8[G * k + bi];
      }
#pragma unroll
      for (int c2 = 0; c2 < 8; c2++)
#pragma unroll
        for (int r = 0; r < 8; r++) a[c2][r] = fma(-g8[r], pr8[c2], a[c2][r]);
      if (mine) {
#pragma unroll
        for (int r = 0; r < 8; r++) a[c][r] = (pblk && r == pr) ? 1.0 : -g8[r];
      }
      if (pblk) pivm |= 1 << pr;
    }
  }
  if (singular) {
    if (t == 0) {
      logdet_out[2 * w + s] = -1e201;
      sgn_out[2 * w + s] = 1;
    }
    for (int k = t; k < n * nsp; k += NT) out[k] = (TI)0;
    return;
  }
  __syncthreads();
  for (int j = t; j < n; j += NT) qinv[pl[j]] = j;
  // log|det| (wavefront sums added in wavefront order), negative pivots and the inversion
  // count of the pivot sequence
  double lg = 0;
  int neg = 0, inv_cnt = 0;
  for (int j = t; j < n; j += NT) {
    lg += log(fabs(pvs[j]));
    neg += pvs[j] < 0;
    const int pj = pl[j];
    for (int m = j + 1; m < n; m++) inv_cnt += pl[m] < pj;
  }
  lg = wsum(lg);
  const int par = wave_isum(neg + inv_cnt);
  if (lane == 0) {
    redd[wv] = lg;
    redi[wv] = par;
  }
  __syncthreads();
  if (t == 0) {
    double lt = redd[0];
    int pt = redi[0];
    for (int k = 1; k < NWV; k++) {
      lt += redd[k];
      pt += redi[k];
    }
    logdet_out[2 * w + s] = lt;
    sgn_out[2 * w + s] = (pt & 1) ? -1.0 : 1.0;
  }
  // row 8bi+r was the pivot of step q: it holds row q of the inverse of the row-permuted
  // matrix, scaled by 1/pivot_q; its column m goes to out[p_m][q]
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const int row = 8 * bi + r;
    if (row < n) {
      const int q = qinv[row];
      const double sc = 1.0 / pvs[q];
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int m = 8 * bj + c;
        if (m < n) out[(size_t)pl[m] * nsp + q] = (TI)(a[c][r] * sc);
      }
    }
  }
}

template <int G, typename TIN, typename TI>
__global__ __launch_bounds__(G * G) void k_inv_grid(DevSys S, const TIN* mo, TI* inv_out, double* logdet_out,
                                                  double* sgn_out, int nw, int only_spin) {
  __shared__ InvGridLds<G> L;
  const int w = only_spin < 0 ? (int)(blockIdx.x >> 1) : (int)blockIdx.x;
  const int s = only_spin < 0 ? (int)(blockIdx.x & 1) : only_spin;
  if (w >= nw) return;
  k_inv_grid_body<G>(S, mo, inv_out, logdet_out, sgn_out, w, s, (int)threadIdx.x, L);
}

// ---------------------------------------------------------------- k_inv_panel
// FP64 pivoted Gauss-Jordan for 129..256 electrons of a spin, one 256-thread block per matrix,
// the working matrix W (n x n, row-major) in device memory and the current panel of kPanel
// columns in LDS.  The elimination is that of k_inv_blk / k_inv_grid (rows left in place,
// pivot = the unpivoted row with the largest vv[i]*|a[i][j]|, the last row on ties, multipliers
// a[i][j] * (1/pivot), column j replaced by -multiplier or 1), grouped by panels:
//   1. the panel's columns are eliminated in LDS (thread = row), recording the multipliers
//      g^(k) (k < kPanel) and pivot rows p_k;
//   2. every other column c (thread = column) takes the panel's steps at once: the pivot-row
//      values the steps used follow from the values before the panel by a triangular solve,
//      r_k = W[p_k][c] - sum_(m<k) g^(m)[p_k] r_m, and W[:,c] -= sum_k g^(k) r_k,
// so W is read and written once per panel instead of once per column
// (src/program_utils/MatrixAlgebrac.cpp:108-224,316-421 for the pivot rule).  The log|det|,
// sign and the transposed inverse follow from the pivots as in k_inv_blk.
constexpr int kPanel = 16;
constexpr int kRowsPanel = 16;  // rows of W in flight per thread in the whole-matrix passes
// Rows of the panel and multiplier tiles are padded to kPanel + 1 doubles: a thread reads and
// writes its own row, and with a 16-double stride every lane of a wavefront would hit the
// same LDS banks (a 64-way conflict on each such access).
struct InvPanelLds {
  double pan[kBigMaxSpin][kPanel + 1];  // the panel's columns
  double g[kBigMaxSpin][kPanel + 1];    // multipliers of the panel's steps
  double vv[kBigMaxSpin];           // 1 / max |row|
  double pvs[kBigMaxSpin];
  int pl[kBigMaxSpin];
  int qinv[kBigMaxSpin];
  int used[kBigMaxSpin];
  double wbest[4];
  int wrow[4];
  double redd[4];
  int redi[4];
};
template <typename TIN, typename TI>
__global__ __launch_bounds__(256) void k_inv_panel(DevSys S, const TIN* mo, double* work, TI* inv_out,
                                                   double* logdet_out, double* sgn_out, int nw, int s) {
  __shared__ InvPanelLds L;
  const int w = (int)blockIdx.x, t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
  if (w >= nw) return;
  const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0, nsp = S.nsp;
  TI* out = inv_out + ((size_t)w * 2 + s) * nsp * nsp;
  double* W = work + (size_t)w * n * n;
  if (n == 0) {
    if (t == 0) {
      logdet_out[2 * w + s] = 0;
      sgn_out[2 * w + s] = 1;
    }
    return;
  }
  // W = A (A[i][c] = phi_c(r_i), the MO value rows); row scales; thread t = column t.  W is
  // read and written kRowsPanel rows at a time, every load of a group issued before its
  // stores, so the device-memory latency is paid once per group and not once per row.
  for (int i0 = 0; i0 < n; i0 += kRowsPanel) {
    double v[kRowsPanel];
#pragma unroll
    for (int u = 0; u < kRowsPanel; u++)
      v[u] = (t < n && i0 + u < n) ? (double)mo[((size_t)w * S.ne + off + i0 + u) * 5 * nsp + t] : 0.0;
#pragma unroll
    for (int u = 0; u < kRowsPanel; u++)
      if (t < n && i0 + u < n) W[(size_t)(i0 + u) * n + t] = v[u];
  }
  bool zero_row = false;
  // a thread's own row is contiguous: 16-byte loads when every row starts 16-byte aligned
  constexpr int per = 16 / sizeof(TIN);
  const bool vec16 = (nsp * 5) % per == 0;
  if (t < n) {
    const TIN* rp = mo + ((size_t)w * S.ne + off + t) * 5 * nsp;
    double mx = 0;
    int c = 0;
    if (vec16) {
      typedef TIN tv16 __attribute__((ext_vector_type(per)));
      for (; c + per <= n; c += per) {
        const tv16 u = *(const tv16*)(rp + c);
#pragma unroll
        for (int i = 0; i < per; i++) mx = fmax(mx, fabs((double)u[i]));
      }
    }
    for (; c < n; c++) mx = fmax(mx, fabs((double)rp[c]));
    L.vv[t] = mx == 0.0 ? 0.0 : 1.0 / mx;
    zero_row = mx == 0.0;
  }
  L.used[t] = t >= n;
  if (lane == 0) L.redi[wv] = wave_any(zero_row) ? 1 : 0;
  __syncthreads();
  bool singular = (L.redi[0] | L.redi[1] | L.redi[2] | L.redi[3]) != 0;
  for (int j0 = 0; j0 < n && !singular; j0 += kPanel) {
    const int jb = n - j0 < kPanel ? n - j0 : kPanel;
    // 1. the panel into LDS (16 consecutive doubles of a row per 16 threads)
    for (int x = t; x < n * kPanel; x += 256) {
      const int i = x / kPanel, k = x % kPanel;
      L.pan[i][k] = k < jb ? W[(size_t)i * n + j0 + k] : 0.0;
    }
    __syncthreads();
    for (int kk = 0; kk < jb; kk++) {
      const int j = j0 + kk;
      // pivot: thread = row; per-wavefront winner, then the highest row among the best
      double cd = -1.0;
      if (t < n && !L.used[t]) {
        cd = L.vv[t] * fabs(L.pan[t][kk]);
        if (!(cd >= 0.0)) cd = -1.0;
      }
      const int Lw = wave_argmax_last(cd, lane);
      if (lane == 0) {
        L.wbest[wv] = Lw < 0 ? -2.0 : wave_bcast(cd, Lw);
        L.wrow[wv] = Lw < 0 ? -1 : 64 * wv + Lw;
      }
      __syncthreads();
      double gb = -2.0;
      int p = -1;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (L.wbest[q] >= gb) {
          gb = L.wbest[q];
          p = L.wrow[q];
        }
      const double pv = p < 0 ? 0.0 : L.pan[p][kk];
      if (p < 0 || !(gb >= 0.0) || pv == 0.0) {
        singular = true;
        break;
      }
      // every thread holds the pivot before the pivot row's column kk is rewritten
      __syncthreads();
      if (t == 0) {
        L.pl[j] = p;
        L.pvs[j] = pv;
      }
      const double rpv = 1.0 / pv;
      double g = 0.0;
      if (t < n && t != p) {
        g = L.pan[t][kk] * rpv;
        for (int c = 0; c < jb; c++)
          if (c != kk) L.pan[t][c] = fma(-g, L.pan[p][c], L.pan[t][c]);
      }
      if (t < n) {
        L.g[t][kk] = g;
        L.pan[t][kk] = t == p ? 1.0 : -g;
      }
      if (t == p) L.used[t] = 1;
      __syncthreads();
    }
    if (singular) break;
    // the panel back to W
    for (int x = t; x < n * kPanel; x += 256) {
      const int i = x / kPanel, k = x % kPanel;
      if (k < jb) W[(size_t)i * n + j0 + k] = L.pan[i][k];
    }
    // 2. every other column: thread = column
    if (t < n && (t < j0 || t >= j0 + jb)) {
      double r[kPanel];
#pragma unroll
      for (int k = 0; k < kPanel; k++) {
        r[k] = 0.0;
        if (k < jb) {
          const int pk = L.pl[j0 + k];
          double v = W[(size_t)pk * n + t];
#pragma unroll
          for (int m = 0; m < kPanel; m++)
            if (m < k) v = fma(-L.g[pk][m], r[m], v);
          r[k] = v;
        }
      }
      for (int i0 = 0; i0 < n; i0 += kRowsPanel) {
        double v[kRowsPanel];
#pragma unroll
        for (int u = 0; u < kRowsPanel; u++) v[u] = i0 + u < n ? W[(size_t)(i0 + u) * n + t] : 0.0;
#pragma unroll
        for (int u = 0; u < kRowsPanel; u++) {
          const int i = i0 + u < n ? i0 + u : n - 1;
#pragma unroll
          for (int k = 0; k < kPanel; k++)
            if (k < jb) v[u] = fma(-L.g[i][k], r[k], v[u]);
        }
#pragma unroll
        for (int u = 0; u < kRowsPanel; u++)
          if (i0 + u < n) W[(size_t)(i0 + u) * n + t] = v[u];
      }
    }
    __threadfence_block();
    __syncthreads();
  }
  if (singular) {
    if (t == 0) {
      logdet_out[2 * w + s] = -1e201;
      sgn_out[2 * w + s] = 1;
    }
    for (int k = t; k < n * nsp; k += 256) out[k] = (TI)0;
    return;
  }
  for (int j = t; j < n; j += 256) L.qinv[L.pl[j]] = j;
  // log|det|, negative pivots and the inversion count of the pivot sequence
  double lg = 0;
  int neg = 0, inv_cnt = 0;
  if (t < n) {
    lg = log(fabs(L.pvs[t]));
    neg = L.pvs[t] < 0;
    const int pj = L.pl[t];
    for (int m = t + 1; m < n; m++) inv_cnt += L.pl[m] < pj;
  }
  lg = wsum(lg);
  const int par = wave_isum(neg + inv_cnt);
  if (lane == 0) {
    L.redd[wv] = lg;
    L.redi[wv] = par;
  }
  __syncthreads();
  if (t == 0) {
    logdet_out[2 * w + s] = ((L.redd[0] + L.redd[1]) + L.redd[2]) + L.redd[3];
    sgn_out[2 * w + s] = ((L.redi[0] + L.redi[1] + L.redi[2] + L.redi[3]) & 1) ? -1.0 : 1.0;
  }
  // row p_q holds row q of the inverse of the row-permuted matrix scaled by 1/pivot_q: its
  // column m goes to out[p_m][q]; thread = m, reading row p_q contiguously
  // thread m = t writes its output row p_m, kRowsPanel consecutive entries at a time (16-byte
  // stores when the output rows start 16-byte aligned)
  const int pt = t < n ? L.pl[t] : 0;
  constexpr int po = 16 / sizeof(TI);
  const bool ovec = nsp % po == 0;
  for (int q0 = 0; q0 < n; q0 += kRowsPanel) {
    TI v[kRowsPanel];
#pragma unroll
    for (int u = 0; u < kRowsPanel; u++) {
      const int q = q0 + u;
      v[u] = (t < n && q < n) ? (TI)(W[(size_t)L.pl[q] * n + t] * (1.0 / L.pvs[q])) : (TI)0;
    }
    if (t >= n) continue;
    TI* orow = out + (size_t)pt * nsp + q0;
    if (ovec && q0 + kRowsPanel <= n) {
      typedef TI ov16 __attribute__((ext_vector_type(po)));
#pragma unroll
      for (int h = 0; h < kRowsPanel / po; h++) {
        ov16 x;
#pragma unroll
        for (int i = 0; i < po; i++) x[i] = v[po * h + i];
        *(ov16*)(orow + po * h) = x;
      }
    } else {
#pragma unroll
      for (int u = 0; u < kRowsPanel; u++)
        if (q0 + u < n) orow[u] = v[u];
    }
  }
}


}  // namespace qw
