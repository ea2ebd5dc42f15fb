// Walker-engine kernels (FP64 path).  One 64-lane wavefront = one walker per block.
//
//   k_mo_rows    basis + LCAO for every electron of every walker (walker refresh, fixed-config
//                evaluation); src/orbitals/MO_matrix_cutoff.h:480-560
//   k_inv<NS>    pivoted inversion, log|det|, sign and transposed inverse per walker and spin;
//                src/program_utils/MatrixAlgebrac.cpp:108-224,316-421
//   k_sweep<NS,D> single-electron moves of one spin channel: drift, proposal, ratio, Jastrow,
//                acceptance, Sherman-Morrison update with the inverse held in registers;
//                src/method/Split_sample.cpp:302-540, src/program_utils/MatrixAlgebrac.cpp:766-838
//   k_local      local energy (kinetic, Coulomb, nonlocal pseudopotential), DMC weight, block
//                sums; src/system/System.cpp:53-140, Molecular_system.cpp:209-284,
//                Pseudopotential.cpp:596-815, src/method/Dmc_method.cpp:801-824
//   k_branch     trial energy, effective timestep, fixed-population pair branching;
//                src/method/Dmc_method.cpp:776-797,961-995
//   k_copy       walker state copies decided by k_branch
#pragma once
#include "qw_move.hpp"

namespace qw {

// Per-walker block accumulators: w, wE, wKE, wV, wNL, steps, E, E^2, KE, KE^2, V, V^2, NL,
// NL^2, w^2, (spare).
constexpr int kBsum = 16;

// Scalars of the DMC control loop kept on the device (read by k_local, written by k_branch).
struct DmcCtl {
  double etrial, teff, cut_start, cut_stop, max_weight, eref, feedback, tau;
  double sumw;     // last weight sum
  int npairs;
  int pad;
};

// ---------------------------------------------------------------- LCAO of one point
// Lane j < n accumulates MO j: phi[c] = sum_b C[b][j] chi[c][b] over all basis chunks.
// chi is an LDS tile of 5 x S.chunk_nf doubles.  Block-wide barriers separate the tile
// producer (basis_lap) from the consumer.
// flist: LDS tile of S.chunk_nf ints (indices of the active functions).
__device__ inline void lcao_lap_n(const DevSys& S, int s, int n, const double p[3], double* chi, int* flist,
                                  double phi[5], int lane) {
  const int ld = S.chunk_nf;
  for (int c = 0; c < 5; c++) phi[c] = 0;
  for (int ch = 0; ch < S.nchunk; ch++) {
    int sh0 = S.chunk_sh0[ch], sh1 = S.chunk_sh0[ch + 1];
    __syncthreads();
    const int nact = basis_lap_sparse<double>(S, p, sh0, sh1, chi, ld, flist, lane);
    __syncthreads();
    if (lane < n) {
      const double* Cp = S.C[s] + lane;
      double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
#pragma unroll 8
      for (int b = 0; b < nact; b++) {
        double c = Cp[(size_t)flist[b] * n];
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
}

__device__ inline void lcao_lap(const DevSys& S, int s, const double p[3], double* chi, int* flist, double phi[5],
                                int lane) {
  lcao_lap_n(S, s, s ? S.ndn : S.nup, p, chi, flist, phi, lane);
}

// n: MOs per row of C[s] (the spin's electrons, or its MO list with several determinants).
__device__ inline double lcao_val_n(const DevSys& S, int s, int n, const double p[3], double* chi, int* flist,
                                    int lane) {
  double phi = 0;
  for (int ch = 0; ch < S.nchunk; ch++) {
    int sh0 = S.chunk_sh0[ch], sh1 = S.chunk_sh0[ch + 1];
    __syncthreads();
    const int nact = basis_val_sparse<double>(S, p, sh0, sh1, chi, flist, lane);
    __syncthreads();
    if (lane < n) {
      const double* Cp = S.C[s] + lane;
      double a = 0;
#pragma unroll 8
      for (int b = 0; b < nact; b++) a += Cp[(size_t)flist[b] * n] * chi[b];
      phi += a;
    }
  }
  return phi;
}
__device__ inline double lcao_val(const DevSys& S, int s, const double p[3], double* chi, int* flist, int lane) {
  return lcao_val_n(S, s, s ? S.ndn : S.nup, p, chi, flist, lane);
}

// Basis values at the points p[0..nq-1] (NQ >= nq) over the union of their active shells of
// shells sh0..sh1-1: chi[i*ld + b] for point i and packed function b (exactly zero where the
// shell is outside its cutoff for that point), flist[b] the function index.  Returns the
// number of packed functions.
template <int NQ>
__device__ inline int basis_val_multi(const DevSys& S, const double (*p)[3], int nq, int sh0, int sh1, double* chi,
                                      int ld, int* flist, int lane) {
  int base = 0;
  for (int s0 = sh0; s0 < sh1; s0 += kWave) {
    const int sh = s0 + lane;
    ShellDesc d;
    double x[NQ], y[NQ], z[NQ], r[NQ];
    bool act[NQ];
    bool any = false;
    if (sh < sh1) {
      d = S.shells[sh];
      const double* ap = S.atom_pos + 3 * d.atom;
#pragma unroll
      for (int i = 0; i < NQ; i++) {
        act[i] = false;
        if (i >= nq) continue;
        x[i] = p[i][0] - ap[0];
        y[i] = p[i][1] - ap[1];
        z[i] = p[i][2] - ap[2];
        r[i] = sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
        act[i] = r[i] < d.rcut && r[i] < d.objthr;
        any |= act[i];
      }
    }
    int tot;
    const int at = base + wave_excl_scan(any ? d.nf : 0, lane, &tot);
    if (any) {
#pragma unroll
      for (int i = 0; i < NQ; i++) {
        if (i >= nq) break;
        double* cq = chi + i * ld + at;
        if (act[i]) {
          int iv = (int)(r[i] * S.tab_inv[d.table]);
          iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
          const double* c = S.spline + ((size_t)d.table * S.nint + iv) * 4;
          double h = r[i] - iv * S.tab_sp[d.table];
          double f = c[0] + h * (c[1] + h * (c[2] + h * c[3]));
          shell_val<double>(d.sym, f, x[i], y[i], z[i], cq);
        } else {
          for (int k = 0; k < d.nf; k++) cq[k] = 0.0;
        }
      }
      for (int k = 0; k < d.nf; k++) flist[at + k] = d.f0 + k;
    }
    base += tot;
  }
  return base;
}

// MO j (lane j < n) at the points p[0..nq-1]: phi[i].  A function outside its cutoff for a
// point adds an exact zero there, so phi[i] equals lcao_val at p[i] bit for bit; each
// coefficient load serves every point.
template <int NQ>
__device__ inline void lcao_val_multi(const DevSys& S, int s, const double (*p)[3], int nq, double* chi, int* flist,
                                      double phi[NQ], int lane) {
  const int n = s ? S.ndn : S.nup, ld = S.chunk_nf;
#pragma unroll
  for (int i = 0; i < NQ; i++) phi[i] = 0;
  for (int ch = 0; ch < S.nchunk; ch++) {
    int sh0 = S.chunk_sh0[ch], sh1 = S.chunk_sh0[ch + 1];
    __syncthreads();
    const int nact = basis_val_multi<NQ>(S, p, nq, sh0, sh1, chi, ld, flist, lane);
    __syncthreads();
    if (lane < n) {
      const double* Cp = S.C[s] + lane;
      double a[NQ];
#pragma unroll
      for (int i = 0; i < NQ; i++) a[i] = 0;
#pragma unroll 4
      for (int b = 0; b < nact; b++) {
        const double c = Cp[(size_t)flist[b] * n];
#pragma unroll
        for (int i = 0; i < NQ; i++) a[i] += c * chi[i * ld + b];
      }
#pragma unroll
      for (int i = 0; i < NQ; i++) phi[i] += a[i];
    }
  }
}

// Quadrature points of a nonlocal pseudopotential term evaluated together in k_local.
constexpr int kPPBatch = 4;

// ---------------------------------------------------------------- k_mo_rows
static __global__ __launch_bounds__(64) void k_mo_rows(DevSys S, DevWalkers W, int nw) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const long blk = blockIdx.x;
  const int w = (int)(blk / S.ne), e = (int)(blk % S.ne);
  if (w >= nw) return;
  const int s = e < S.nup ? 0 : 1, n = S.nmo[s];
  double p[3];
  for (int d = 0; d < 3; d++) p[d] = W.pos[((size_t)w * S.ne + e) * 3 + d];
  double phi[5];
  lcao_lap_n(S, s, n, p, lds, (int*)(lds + 5 * S.chunk_nf), phi, lane);
  if (lane < n) {
    double* row = W.mo + ((size_t)w * S.ne + e) * 5 * S.nsp;
    for (int c = 0; c < 5; c++) row[c * S.nsp + lane] = phi[c];
  }
}

// ---------------------------------------------------------------- k_inv
// Pivoted Gauss-Jordan inversion with the rows in registers: lane i holds row i of the
// Slater matrix A[i][j] = phi_j(r_i) of spin s (TIN: FP64 or FP32 MO cache; arithmetic FP64).
// The pivot of column j is the unpivoted row with the largest vv[i]*|a[i][j]| (vv = 1/max|row|,
// last row on ties): the pivot sequence of the reference's Crout ludcmp, whose candidates are
// the same Schur-complement entries (src/program_utils/MatrixAlgebrac.cpp:108-198).  Rows are
// never exchanged.  With A' the row-permuted matrix (A'[j] = A[p_j]), lane p_j ends holding
// row j of inv(A') in place; the transposed inverse out[k][i] = inv(A)[i][k] is then
// inv(A')[i][q_k] with q = p^-1, so lane p_i stores its column m at out[p_m][i].
// The row is held as NB vectors of VW doubles.  The step loop runs over the vectors (unrolled)
// and inside each over its columns, so column j is read and written with one register-relative
// move in a vector known at compile time.  A pivot row is left unscaled in its lane with a
// per-lane scale (1/pivot), so each elimination step is one FMA per element.
// log|det| = sum log|pivot|; the sign comes from the pivots and the permutation parity; an
// exactly zero pivot flags the matrix singular as the reference does (log|det| = -1e201,
// inverse zero).
// LDS of one wavefront's inversion (k_inv_rows_body).
template <int NS>
struct InvRowsLds {
  double prow2[2][NS];
  double pvs[NS];
  int pl[NS];
};
// The inversion of walker w's spin s by one wavefront (WPB wavefronts per block, each on its
// own matrix with its own LDS L; wavefront-level barriers when WPB > 1).
template <int NS, int WPB, typename TIN, typename TI>
__device__ inline void k_inv_rows_body(const DevSys& S, const TIN* mo, TI* inv_out, double* logdet_out,
                                       double* sgn_out, int w, int s, int lane, InvRowsLds<NS>& L) {
  double(&prow2)[2][NS] = L.prow2;
  double* pvs = L.pvs;
  int* pl = L.pl;
  const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0, nsp = S.nsp;
  TI* out = inv_out + ((size_t)w * 2 + s) * nsp * nsp;
  if (n == 0) {
    if (lane == 0) {
      logdet_out[2 * w + s] = 0;
      sgn_out[2 * w + s] = 1;
    }
    return;
  }
  const bool row = lane < n;
  constexpr int VW = NS < 16 ? NS : 16, NB = NS / VW;
  typedef double dvec __attribute__((ext_vector_type(VW)));
  dvec a[NB];
  const TIN* rp = mo + ((size_t)w * S.ne + off + (row ? lane : 0)) * 5 * nsp;
  double big = 0;
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int k = 0; k < VW; k++) {
      const int m = b * VW + k;
      a[b][k] = (row && m < n) ? (double)rp[m] : 0.0;
      big = fmax(big, fabs(a[b][k]));
    }
  bool singular = wave_any(row && big == 0.0);
  const double vv = big == 0.0 ? 0.0 : 1.0 / big;
  bool pivoted = !row;
  double sc = 1.0;
  int qi = 0;
  bool done = singular;
#pragma unroll
  for (int jb = 0; jb < NB; jb++) {
    for (int jo = 0; jo < VW && !done; jo++) {
      const int j = jb * VW + jo;
      if (j >= n) {
        done = true;
        break;
      }
      const double aj = a[jb][jo];
      const double cand = pivoted ? -1.0 : vv * fabs(aj);
      const int p = wave_argmax_last(cand, lane);
      // no candidate (NaN entries) or a zero pivot: singular
      const double pv = p < 0 ? 0.0 : wave_bcast(aj, p);
      if (p < 0 || !(wave_bcast(cand, p) >= 0.0) || pv == 0.0) {
        singular = done = true;
        break;
      }
      if (lane == 0) {
        pl[j] = p;
        pvs[j] = pv;
      }
      double* prow = prow2[j & 1];
      const bool me = lane == p;
      if (me) {
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
          for (int k = 0; k < VW; k++) prow[b * VW + k] = a[b][k];
      }
      walker_sync<WPB>();
      const double rpv = 1.0 / pv;
      const double g = me ? 0.0 : aj * rpv;
#pragma unroll
      for (int b = 0; b < NB; b++)
#pragma unroll
        for (int k = 0; k < VW; k++) a[b][k] = fma(-g, prow[b * VW + k], a[b][k]);
      a[jb][jo] = me ? 1.0 : -g;
      if (me) {
        sc = rpv;
        qi = j;
        pivoted = true;
      }
    }
  }
  if (singular) {
    if (lane == 0) {
      logdet_out[2 * w + s] = -1e201;
      sgn_out[2 * w + s] = 1;
    }
    for (int k = lane; k < n * nsp; k += kWave) out[k] = (TI)0;
    return;
  }
  walker_sync<WPB>();
  // log|det|, negative pivots and the inversion count of the pivot sequence
  double lg = 0;
  int neg = 0, inv_cnt = 0;
  for (int j = lane; j < n; j += kWave) {
    lg += log(fabs(pvs[j]));
    neg += pvs[j] < 0;
    const int pj = pl[j];
    for (int m = j + 1; m < n; m++) inv_cnt += pl[m] < pj;
  }
  lg = wsum(lg);
  const int par = wave_isum(neg + inv_cnt);
  if (lane == 0) {
    logdet_out[2 * w + s] = lg;
    sgn_out[2 * w + s] = (par & 1) ? -1.0 : 1.0;
  }
  if (row) {
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int k = 0; k < VW; k++) {
        const int m = b * VW + k;
        if (m < n) out[(size_t)pl[m] * nsp + qi] = (TI)(a[b][k] * sc);
      }
  }
}

template <int NS, typename TIN, typename TI>
__global__ __launch_bounds__(64) void k_inv(DevSys S, const TIN* mo, TI* inv_out, double* logdet_out,
                                            double* sgn_out, int nw, int only_spin) {
  __shared__ InvRowsLds<NS> L;
  const int w = only_spin < 0 ? (int)(blockIdx.x >> 1) : (int)blockIdx.x;
  const int s = only_spin < 0 ? (int)(blockIdx.x & 1) : only_spin;
  if (w >= nw) return;
  k_inv_rows_body<NS, 1>(S, mo, inv_out, logdet_out, sgn_out, w, s, (int)threadIdx.x, L);
}

// ---------------------------------------------------------------- k_inv_blk
// The same inversion for 33..64 electrons of a spin with the matrix spread over the wave as
// 8x8 blocks: lane (bi, bj) = (lane / 8, lane % 8) holds rows 8bi..8bi+7 and columns
// 8bj..8bj+7, column c of its block as one register vector a[c][r].  A step exchanges the
// pivot column, the pivot row and the multipliers through LDS (8 values per lane each way),
// and every register index except the pivot row's is a compile-time constant.  The pivot
// search runs with lane l standing for row l: the eight lanes holding column j write it to
// LDS, every lane reads its row's entry and scale and the wave takes the largest candidate,
// so each lane tests one candidate instead of eight, and lane l's multiplier is its own
// entry times 1/pivot.  Each element sees the operations of k_inv in the same order (same
// pivots, multipliers a[i][j] * (1/pivot), one FMA per step, column j replaced by
// -multiplier or 1), so both kernels give bitwise-equal results.  Rows and columns >= n hold
// the identity and are never pivots before step n.
// The exchanged column, pivot row and multipliers are stored interleaved (entry 8i + b of a
// row or column at index 8b + i), so the eight values a lane reads or writes in one step are
// spread over the LDS banks instead of sharing two of them (12 % faster than row order).
struct InvBlkLds {
  double prowL[2][64];
  double gL[2][64];
  double colL[2][64];
  double pvs[64];
  int pl[64];
  int qinv[64];
};
template <int WPB, typename TIN, typename TI>
__device__ inline void k_inv_blk_body(const DevSys& S, const TIN* mo, TI* inv_out, double* logdet_out,
                                      double* sgn_out, int w, int s, int lane, InvBlkLds& L) {
  typedef double dv8 __attribute__((ext_vector_type(8)));
  double(&prowL)[2][64] = L.prowL;
  double(&gL)[2][64] = L.gL;
  double(&colL)[2][64] = L.colL;
  double* pvs = L.pvs;
  int* pl = L.pl;
  int* qinv = L.qinv;
  const int bi = lane >> 3, bj = lane & 7;
  const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0, nsp = S.nsp;
  TI* out = inv_out + ((size_t)w * 2 + s) * nsp * nsp;
  if (n == 0) {
    if (lane == 0) {
      logdet_out[2 * w + s] = 0;
      sgn_out[2 * w + s] = 1;
    }
    return;
  }
  dv8 a[8];
  double vv[8];
  bool zero_row = false;
  // a lane's 8 entries of a row are contiguous: 16-byte loads when every MO row starts on a
  // 16-byte boundary (row length 5 * nsp values; entries past n are read, then replaced)
  const bool vec16 = sizeof(TIN) == 4 ? (nsp & 3) == 0 : (nsp & 1) == 0;
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const int row = 8 * bi + r;
    const TIN* rp = mo + ((size_t)w * S.ne + off + (row < n ? row : 0)) * 5 * nsp + 8 * bj;
    TIN vin[8];
    if (vec16) {
      typedef TIN tv16 __attribute__((ext_vector_type(16 / sizeof(TIN))));
      constexpr int per = 16 / sizeof(TIN);
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
    // row maximum over the 8 lanes of the row block (lanes 8bi .. 8bi+7)
    mx = fmax(mx, dpp_d<kDppXor1>(mx));
    mx = fmax(mx, dpp_d<kDppXor2>(mx));
    mx = fmax(mx, dpp_d<kDppHalfMirror>(mx));
    vv[r] = mx == 0.0 ? 0.0 : 1.0 / mx;
    zero_row |= row < n && mx == 0.0;
  }
  // lane l = row l (row block bi, r = bj): its implicit scale, and whether it was a pivot
  double vvl = vv[0];
#pragma unroll
  for (int r = 1; r < 8; r++)
    if (bj == r) vvl = vv[r];
  bool used = lane >= n;
  bool singular = wave_any(zero_row);
  bool done = singular;
  for (int J = 0; J < 8 && !done; J++) {
    const bool mine = bj == J;  // this lane holds column 8J + c of its rows
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int j = 8 * J + c;
      if (done || j >= n) {
        done = true;
        break;
      }
      double* col = colL[j & 1];
      if (mine) {
#pragma unroll
        for (int r = 0; r < 8; r++) col[8 * r + bi] = a[c][r];
      }
      walker_sync<WPB>();
      const double x = col[8 * bj + bi];
      // candidate of row `lane` (largest vv * |a|, the last row on ties); NaN entries are none
      double cd = used ? -1.0 : vvl * fabs(x);
      if (!(cd >= 0.0)) cd = -1.0;
      const int p = wave_argmax_last(cd, lane);
      const double pv = p < 0 ? 0.0 : wave_bcast(x, p);
      if (p < 0 || !(wave_bcast(cd, p) >= 0.0) || pv == 0.0) {
        singular = done = true;
        break;
      }
      const int pb = p >> 3, pr = p & 7;
      if (lane == 0) {
        pl[j] = p;
        pvs[j] = pv;
      }
      double* prow = prowL[j & 1];
      double* gw = gL[j & 1];
      if (bi == pb) {
#pragma unroll
        for (int c2 = 0; c2 < 8; c2++) prow[8 * c2 + bj] = a[c2][pr];
      }
      const double rpv = 1.0 / pv;
      gw[8 * bj + bi] = lane == p ? 0.0 : x * rpv;
      walker_sync<WPB>();
      double pr8[8], g8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        pr8[k] = prow[8 * k + bj];
        g8[k] = gw[8 * k + bi];
      }
#pragma unroll
      for (int c2 = 0; c2 < 8; c2++)
#pragma unroll
        for (int r = 0; r < 8; r++) a[c2][r] = fma(-g8[r], pr8[c2], a[c2][r]);
      if (mine) {
        // row 8bi + r is the pivot row p = 8pb + pr (pr wave-uniform, so no per-row lane constants)
        const bool pblk = bi == pb;
#pragma unroll
        for (int r = 0; r < 8; r++) a[c][r] = (pblk && r == pr) ? 1.0 : -g8[r];
      }
      used |= lane == p;
    }
  }
  if (singular) {
    if (lane == 0) {
      logdet_out[2 * w + s] = -1e201;
      sgn_out[2 * w + s] = 1;
    }
    for (int k = lane; k < n * nsp; k += kWave) out[k] = (TI)0;
    return;
  }
  walker_sync<WPB>();
  for (int j = lane; j < n; j += kWave) qinv[pl[j]] = j;
  // log|det|, negative pivots and the inversion count of the pivot sequence
  double lg = 0;
  int neg = 0, inv_cnt = 0;
  for (int j = lane; j < n; j += kWave) {
    lg += log(fabs(pvs[j]));
    neg += pvs[j] < 0;
    const int pj = pl[j];
    for (int m = j + 1; m < n; m++) inv_cnt += pl[m] < pj;
  }
  lg = wsum(lg);
  const int par = wave_isum(neg + inv_cnt);
  if (lane == 0) {
    logdet_out[2 * w + s] = lg;
    sgn_out[2 * w + s] = (par & 1) ? -1.0 : 1.0;
  }
  walker_sync<WPB>();
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

// Three waves per SIMD (at most 168 VGPRs, a few spilled): 6 % faster than two.
template <typename TIN, typename TI>
__global__ __launch_bounds__(64, 3) void k_inv_blk(DevSys S, const TIN* mo, TI* inv_out, double* logdet_out,
                                                double* sgn_out, int nw, int only_spin) {
  __shared__ InvBlkLds L;
  const int w = only_spin < 0 ? (int)(blockIdx.x >> 1) : (int)blockIdx.x;
  const int s = only_spin < 0 ? (int)(blockIdx.x & 1) : only_spin;
  if (w >= nw) return;
  k_inv_blk_body<1>(S, mo, inv_out, logdet_out, sgn_out, w, s, (int)threadIdx.x, L);
}

// ---------------------------------------------------------------- k_sweep
struct SweepArgs {
  int s, k0, k1;
  double tau;
  unsigned long long seed;
  unsigned step, sweep;
  int restrict_nodes;
  int dets;  // write the determinant gradient / laplacian ratios of spin s to W.dgl at the end
  const unsigned* stepp;  // device step counter (replaces `step` when set; graph replays)
};
__device__ inline unsigned step_of(unsigned step, const unsigned* stepp) { return stepp ? *stepp : step; }

// Random numbers of the moves of one sweep launch, drawn one move per lane: lane l takes the
// electron off+k0+l and draws its proposal Gaussians and acceptance uniform for each stage from
// the same counter-based streams a per-move draw uses, so the wave evaluates the draws of up
// to 64 moves at once instead of repeating each on every lane.  rnd: LDS [move][DEPTH][4]
// (Gaussians x, y, z, uniform).
template <int DEPTH>
__device__ inline void sweep_randoms(const SweepArgs& A, unsigned w, int off, double* rnd, int lane) {
  const unsigned st = step_of(A.step, A.stepp);
  for (int k = A.k0 + lane; k < A.k1; k += kWave) {
    const unsigned e = (unsigned)(off + k);
#pragma unroll
    for (int d = 0; d < DEPTH; d++) {
      double g[3];
      gauss3(A.seed, w, st, e, (d == 0 ? TAG_GAUSS1 : TAG_GAUSS2) | (A.sweep << 8), g);
      double* r = rnd + ((size_t)(k - A.k0) * DEPTH + d) * 4;
      r[0] = g[0];
      r[1] = g[1];
      r[2] = g[2];
      r[3] = uniform(A.seed, w, st, e, (d == 0 ? TAG_ACC1 : TAG_ACC2) | (A.sweep << 8));
    }
  }
}
__host__ __device__ inline size_t sweep_rand_bytes(int ns, int depth) { return (size_t)ns * depth * 4 * sizeof(double); }

// Advances the device step counter (one thread).
static __global__ void k_step_advance(unsigned* stepp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *stepp += 1;
}

// Determinant parts of the local energy for lane k's electron of spin s from the inverse row
// in registers: dgl[c] = sum_j M_c[e_k][j] inv(k, j), c = d/dx, d/dy, d/dz, laplacian
// (src/wavefunction/Slat_wf.h:1021-1080, getDetLap).  inv[j] is zero for j >= n, so the row index is
// clamped instead of branched on.
// Row `lane` of a spin's transposed inverse into registers (entries >= n zero) and back: with
// full rows (n = NS) that start 16-byte aligned, 16-byte accesses instead of one value each.
template <int NS, typename T>
__device__ inline void inv_row_load(const T* ig, int lane, int n, int nsp, T (&inv)[NS]) {
  constexpr int per = 16 / sizeof(T);
  if (n == NS && nsp % per == 0) {
    typedef T tv16 __attribute__((ext_vector_type(per)));
    const T* r = ig + (size_t)(lane < n ? lane : 0) * nsp;
#pragma unroll
    for (int j0 = 0; j0 < NS; j0 += per) {
      const tv16 u = lane < n ? *(const tv16*)(r + j0) : tv16(0);
#pragma unroll
      for (int i = 0; i < per; i++) inv[j0 + i] = u[i];
    }
  } else {
#pragma unroll
    for (int j = 0; j < NS; j++) inv[j] = (lane < n && j < n) ? ig[(size_t)lane * nsp + j] : (T)0;
  }
}
template <int NS, typename T>
__device__ inline void inv_row_store(T* ig, int lane, int n, int nsp, const T (&inv)[NS]) {
  if (lane >= n) return;
  constexpr int per = 16 / sizeof(T);
  if (n == NS && nsp % per == 0) {
    typedef T tv16 __attribute__((ext_vector_type(per)));
    T* r = ig + (size_t)lane * nsp;
#pragma unroll
    for (int j0 = 0; j0 < NS; j0 += per) {
      tv16 u;
#pragma unroll
      for (int i = 0; i < per; i++) u[i] = inv[j0 + i];
      *(tv16*)(r + j0) = u;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NS; j++)
      if (j < n) ig[(size_t)lane * nsp + j] = inv[j];
  }
}

template <int NS, typename TI, typename TM>
__device__ inline void write_det_parts(const DevSys& S, const DevWalkers& W, int w, int s, const TI* inv, const TM* mog,
                                       int lane) {
  const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0, nsp = S.nsp;
  if (lane >= n) return;
  const int e = off + lane;
  const TM* mr = mog + (size_t)e * 5 * nsp;
  double t[4] = {0, 0, 0, 0};
  // a lane reads its own electron's rows: 16-byte loads when the rows are full and aligned
  constexpr int per = 16 / sizeof(TM);
  if (n == NS && nsp % per == 0) {
    typedef TM tv16 __attribute__((ext_vector_type(per)));
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const TM* row = mr + (size_t)(c + 1) * nsp;
#pragma unroll
      for (int j0 = 0; j0 < NS; j0 += per) {
        const tv16 u = *(const tv16*)(row + j0);
#pragma unroll
        for (int i = 0; i < per; i++) t[c] += (double)u[i] * (double)inv[j0 + i];
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const TM* row = mr + (size_t)(c + 1) * nsp;
#pragma unroll
      for (int j = 0; j < NS; j++) t[c] += (double)row[min(j, n - 1)] * (double)inv[j];
    }
  }
  double* o = W.dgl + ((size_t)w * S.ne + e) * 4;
  o[0] = t[0];
  o[1] = t[1];
  o[2] = t[2];
  o[3] = t[3];
}

// Basis functions per LDS tile batch in the FP64 sweep.
constexpr int kCapF64 = 128;

__host__ __device__ inline size_t sweep_lds_bytes(int ne, int na, int nsh, int ntab, int ns, int cap, int tsize,
                                                  int depth) {
  return (((size_t)3 * ne * sizeof(double) + 15) & ~(size_t)15) + ((sys_lds_bytes(na, nsh, ntab) + 15) & ~(size_t)15) +
         sweep_rand_bytes(ns, depth) + (size_t)5 * cap * tsize + (size_t)cap * sizeof(int) + (size_t)2 * ns * tsize + 16;
}

// U: coefficient loads in flight in the LCAO (chosen by the engine's online tuner; every U
// gives the same arithmetic in the same order).
// The sweep of walker w by wavefront `wid` of a block of WPB wavefronts, each with a whole
// region of sweep_lds_bytes of its own (lds0: the block's LDS) and wavefront-level barriers
// when WPB > 1: the same arithmetic as a one-walker block.
template <int NS, int DEPTH, int U, int WPB>
__device__ __forceinline__ void sweep_fp64_body(const DevSys& S, const DevWalkers& W, const SweepArgs& A, int w,
                                                int lane, int wid, char* lds0) {
  if (w >= W.nw) return;
  const int ne = S.ne, s = A.s, n = s ? S.ndn : S.nup, nsp = S.nsp, off = s ? S.nup : 0;
  char* base = lds0;
  if constexpr (WPB > 1)
    base += (size_t)wid * ((sweep_lds_bytes(ne, S.na, S.nsh, S.ntab, NS, kCapF64, 8, DEPTH) + 15) & ~(size_t)15);
  double* pos = (double*)base;
  SysLds V;
  char* q = sys_lds_carve(S, base + (((size_t)3 * ne * sizeof(double) + 15) & ~(size_t)15), V);
  double* rnd = (double*)q;
  double* chi = rnd + sweep_rand_bytes(NS, DEPTH) / sizeof(double);
  int* roff = (int*)(chi + 5 * kCapF64);
  double* rowb = (double*)(roff + kCapF64);
  double* phib = rowb + NS;
  for (int i = lane; i < 3 * ne; i += kWave) pos[i] = W.pos[(size_t)w * ne * 3 + i];
  stage_sys(S, V, lane);
  sweep_randoms<DEPTH>(A, (unsigned)w, off, rnd, lane);
  double inv[NS];
  const double* ig = W.inv + ((size_t)w * 2 + s) * nsp * nsp;
  inv_row_load<NS>(ig, lane, n, nsp, inv);
  double logdet = W.logdet[2 * w + s], sgn = W.sgn[2 * w + s];
  int nacc = 0;
  double dprod = 1.0;
  double* mog = W.mo + (size_t)w * ne * 5 * nsp;
  const double tau = A.tau, stau = sqrt(tau);
  // MO gradient row of the electron about to move, loaded one move ahead
  double m1 = 0, m2 = 0, m3 = 0;
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
    const double rk = lane < n ? rowb[lane] : 0.0;
    // drift at the old position: determinant part from the cached MO row, Jastrow part fresh
    double d0[3], gj0[3];
    const double gdx = wsum(rk * m1), gdy = wsum(rk * m2), gdz = wsum(rk * m3);
    const double u0 = jastrow_e(S, e, r0, pos, gj0, nullptr, lane, V.epos, 4);
    d0[0] = gdx + gj0[0];
    d0[1] = gdy + gj0[1];
    d0[2] = gdz + gj0[2];
    double phi[5];
    double r1[3], d1[3], ln1 = 0, R = 0;
    bool accepted = false;
    double acc_first = 0;
    for (int dep = 1; dep <= DEPTH && !accepted; dep++) {
      const double t = dep == 1 ? tau : tau * 0.5, st = dep == 1 ? stau : sqrt(tau * 0.5);
      const double* rq = rnd + ((size_t)(k - A.k0) * DEPTH + (dep - 1)) * 4;
      double dr[3], rn[3], dn[3];
      lim_drift(d0, t, dr);
      for (int d = 0; d < 3; d++) rn[d] = r0[d] + (rq[d] * st + dr[d]);
      mo_lap_point<double, double, U, WPB>(S, V, S.spline, S.C[s], n, rn, chi, roff, kCapF64, phi, lane);
      double t0 = wsum(rk * phi[0]);
      double t1 = wsum(rk * phi[1]), t2 = wsum(rk * phi[2]), t3 = wsum(rk * phi[3]);
      double gj[3];
      double un = jastrow_e(S, e, rn, pos, gj, nullptr, lane, V.epos, 4);
      dn[0] = t1 / t0 + gj[0];
      dn[1] = t2 / t0 + gj[1];
      dn[2] = t3 / t0 + gj[2];
      const bool neg = t0 < 0;
      double acc, lnn = 0;
      if constexpr (DEPTH == 1) {
        // |ratio|^2 exp(dtrans) as t0^2 exp(2 du + dtrans): one exponential, no logarithm
        acc = (A.restrict_nodes && neg)
                  ? 0.0
                  : fmin(1.0, t0 * t0 * exp(2.0 * (un - u0) + (trans(rn, dn, r0, t) - trans(r0, d0, rn, t))));
      } else if (dep == 1) {
        lnn = log(fabs(t0)) + (un - u0);
        double rr = exp(lnn);
        acc = (A.restrict_nodes && neg) ? 0.0
                                        : fmin(1.0, rr * rr * exp(trans(rn, dn, r0, t) - trans(r0, d0, rn, t)));
        acc_first = acc;
        for (int d = 0; d < 3; d++) r1[d] = rn[d], d1[d] = dn[d];
        ln1 = lnn;
      } else {
        // delayed rejection from the origin to the second point, through the rejected first
        // point (src/method/Split_sample.cpp:302-373)
        lnn = log(fabs(t0)) + (un - u0);
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
      // Sherman-Morrison on the transposed inverse for the replaced row k
      if (lane < NS) phib[lane] = lane < n ? phi[0] : 0.0;
      walker_sync<WPB>();
      if (lane < n) {
        // four independent partial sums: the dot product is not one serial FMA chain
        double tq[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NS; j++) tq[j & 3] += inv[j] * phib[j];
        const double tk = (tq[0] + tq[1]) + (tq[2] + tq[3]);
        if (lane == k) {
          const double f = 1.0 / R;
#pragma unroll
          for (int j = 0; j < NS; j++) inv[j] = rowb[j] * f;
        } else {
          const double pk = tk * (-1.0 / R);
#pragma unroll
          for (int j = 0; j < NS; j++) inv[j] += rowb[j] * pk;
        }
        for (int c = 0; c < 5; c++) mog[((size_t)e * 5 + c) * nsp + lane] = phi[c];
      }
      // accepted ratios multiply into dprod, folded into log|det| before it could overflow
      dprod *= R;
      if (!(fabs(dprod) < 0x1p+512 && fabs(dprod) > 0x1p-512)) {
        logdet += log(fabs(dprod));
        dprod = 1.0;
      }
      if (R < 0) sgn = -sgn;
      nacc++;
      if (lane < 3) pos[3 * e + lane] = r1[lane];
    }
    if (lane < n && k + 1 < A.k1) {
      const size_t e1 = e + 1;
      m1 = mog[(e1 * 5 + 1) * nsp + lane];
      m2 = mog[(e1 * 5 + 2) * nsp + lane];
      m3 = mog[(e1 * 5 + 3) * nsp + lane];
    }
    walker_sync<WPB>();
  }
  double* igw = W.inv + ((size_t)w * 2 + s) * nsp * nsp;
  inv_row_store<NS>(igw, lane, n, nsp, inv);
  if (A.dets) write_det_parts<NS>(S, W, w, s, inv, (const double*)mog, lane);
  for (int i = lane; i < 3 * ne; i += kWave) W.pos[(size_t)w * ne * 3 + i] = pos[i];
  logdet += log(fabs(dprod));
  if (lane == 0) {
    W.logdet[2 * w + s] = logdet;
    W.sgn[2 * w + s] = sgn;
    W.acc[w] += nacc;
  }
}

template <int NS, int DEPTH, int U>
__global__ __launch_bounds__(64, 2) void k_sweep(DevSys S, DevWalkers W, SweepArgs A) {
  extern __shared__ double lds[];
  sweep_fp64_body<NS, DEPTH, U, 1>(S, W, A, (int)blockIdx.x, (int)threadIdx.x, 0, (char*)lds);
}

// ---------------------------------------------------------------- k_local
struct LocalArgs {
  int mode;              // 0 = evaluate (write outputs), 1 = VMC sums, 2 = DMC weight + sums, 3 = init (elast only)
  int use_dgl;           // determinant parts from W.dgl (written by the last sweep) instead of the inverse
  const unsigned* stepp; // device step counter (replaces `step` when set)
  int pp_mode;           // 0 deterministic, 1 stochastic
  const double* rot;     // [9] fixed rotation (mode 0) or null
  const double* pp_rand; // [nw][na][ne] uniforms (mode 0, stochastic) or null
  unsigned long long seed;
  unsigned step;
  double* out_logpsi;    // mode 0 outputs
  double* out_sign;
  double* out_grad;
  double* out_lap;
  double* out_eparts;
  DmcCtl* ctl;
  double* tneg;          // T-moves (mode 2): [nw][na][ne] sums of -vxx over the negative terms, or null
};

// TM: element type of the MO cache and inverse (double on the FP64 path, float on the mixed
// path); the local-energy arithmetic and sums are FP64 on both.
template <typename TM>
__device__ inline const TM* mo_cache(const DevWalkers& W);
template <>
__device__ inline const double* mo_cache<double>(const DevWalkers& W) { return W.mo; }
template <>
__device__ inline const float* mo_cache<float>(const DevWalkers& W) { return W.mof; }
template <typename TM>
__device__ inline const TM* inv_store(const DevWalkers& W);
template <>
__device__ inline const double* inv_store<double>(const DevWalkers& W) { return W.inv; }
template <>
__device__ inline const float* inv_store<float>(const DevWalkers& W) { return W.invf; }

// Jastrow of electron i for the local energy: value (pairs i<j and its one-body terms),
// gradient and laplacian with respect to r_i, in arithmetic R (distances from FP64 positions).
// Object loops outermost so each object's parameters stay in scalar registers.
template <typename R>
__device__ inline void jastrow_local(const DevSys& S, const JRad<R>* ee, const JRad<R>* eit, const double* pos, int i,
                                     R& u, R& gx, R& gy, R& gz, R& l) {
  const double* pi = pos + 3 * i;
  const QW_CONST int* ea0 = cptr(S.eit_a0);
  const QW_CONST int* eo0 = cptr(S.eit_o0);
  const QW_CONST double* epos = cptr(S.eit_pos);
  for (int t = 0; t < S.n_eit; t++) {
    const int a0 = ea0[t], a1 = ea0[t + 1], o1 = eo0[t + 1];
    for (int o = eo0[t]; o < o1; o++) {
      with_jrad<R, true>(cptr(eit)[o], [&](const auto& ev) {
        for (int k = a0; k < a1; k++) {
          const QW_CONST double* ap = epos + 3 * k;
          const R x = (R)(pi[0] - ap[0]), y = (R)(pi[1] - ap[1]), z = (R)(pi[2] - ap[2]);
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
  const bool iup = i < S.nup;
  for (int o = 0; o < S.nee; o++) {
    with_jrad<R, true>(cptr(ee)[o], [&](const auto& ev) {
      for (int j = 0; j < S.ne; j++) {
        if (j == i) continue;
        const R x = (R)(pi[0] - pos[3 * j]), y = (R)(pi[1] - pos[3 * j + 1]), z = (R)(pi[2] - pos[3 * j + 2]);
        R r, ri;
        dist_inv(x * x + y * y + z * z, r, ri);
        const R uf = ((j < S.nup) != iup) ? (R)1 : (R)0;
        R gf = 0, upair = 0;
        ev(uf, r, ri, upair, gf, l);
        if (j > i) u += upair;
        gx += gf * x;
        gy += gf * y;
        gz += gf * z;
      }
    });
  }
}

// ---------------------------------------------------------------- several determinants
// Weights b_d = sign_d sign_0 exp(l_d - l_0) of the terms c_d D_up,d D_dn,d (log l_d, sign
// sign_d), pivoting on the first term as the reference's sum of log values does
// (src/program_utils/MatrixAlgebra.h:92-99); determinant parts as weighted sums over the
// determinants of each one's inverse row times its orbitals (src/wavefunction/Slat_wf.h:1021-1139).
constexpr int kMdMaxSpin = 32;

// b[d] for the walker's per-determinant log|det| and sign ([2][ndet] each, in LDS); returns
// sum_d b_d in every lane and the pivot term's log l0 and sign s0.
__device__ inline double md_weights(const DevSys& S, const double* ld, const double* sg, double* b, double& l0,
                                    double& s0, int lane) {
  const int nd = S.ndet;
  const QW_CONST double* lc = cptr(S.logc);
  const QW_CONST double* cw = cptr(S.detwt);
  l0 = lc[0] + ld[0] + ld[nd];
  s0 = (cw[0] < 0 ? -1.0 : 1.0) * sg[0] * sg[nd];
  double part = 0;
  for (int d = lane; d < nd; d += kWave) {
    const double l = lc[d] + ld[d] + ld[nd + d];
    const double sd = (cw[d] < 0 ? -1.0 : 1.0) * sg[d] * sg[nd + d];
    const double v = d == 0 ? 1.0 : sd * s0 * exp(l - l0);
    b[d] = v;
    part += v;
  }
  return wsum(part);
}

// out[c] = sum_d b_d sum_j inv_d(k, j) v[c][occ_d(j)] / den for NC components (v: LDS rows of
// length ldv indexed by the spin's MO list), lanes over determinants.
template <int NC>
__device__ inline void md_contract(const DevSys& S, const double* invw, int s, int k, const double* v, int ldv,
                                   const double* b, double den, double out[NC], int lane) {
  const int nd = S.ndet, n = s ? S.ndn : S.nup, ns = S.nsig;
  const QW_CONST int* oc = cptr(S.occ) + (size_t)s * nd * ns;
  double part[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) part[c] = 0;
  for (int d = lane; d < nd; d += kWave) {
    const double* row = invw + (((size_t)s * nd + d) * ns + k) * ns;
    const QW_CONST int* o = oc + (size_t)d * ns;
    double t[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) t[c] = 0;
    for (int j = 0; j < n; j++) {
      const double r = row[j];
      const int m = o[j];
#pragma unroll
      for (int c = 0; c < NC; c++) t[c] += v[c * ldv + m] * r;
    }
#pragma unroll
    for (int c = 0; c < NC; c++) part[c] += b[d] * t[c];
  }
#pragma unroll
  for (int c = 0; c < NC; c++) out[c] = wsum(part[c]) / den;
}

// MD: several determinants (FP64): determinant parts and quadrature ratios are weighted sums
// over the determinants (md_contract), from their inverses in W.invd.
// PP = false compiles the nonlocal pseudopotential out (systems without one): the kernel then
// needs far fewer registers and runs more waves per SIMD; the arithmetic is the same.
// The local energy of walker w by one 64-thread block (lds: the block's LDS).
template <typename TM, bool MD = false, bool PP = true>
__device__ __forceinline__ void local_body(const DevSys& S, const DevWalkers& W, const LocalArgs& A, int w, int lane,
                                           double* lds) {
  if (w >= W.nw) return;
  const int ne = S.ne, nsp = S.nsp;
  double* pos = lds;
  double* gD = pos + 3 * ne;
  double* lD = gD + 3 * ne;
  double* chi = lD + ne;
  int* flist = (int*)(chi + 5 * S.chunk_nf);
  // MD: determinant weights, their log/sign, and MO values at a quadrature point
  double* bw = (double*)(flist + ((S.chunk_nf + 1) & ~1));
  double* mdl = bw + (MD ? S.ndet : 0);
  double* phq = mdl + (MD ? 4 * S.ndet : 0);
  for (int i = lane; i < 3 * ne; i += kWave) pos[i] = W.pos[(size_t)w * ne * 3 + i];
  const TM* mog = mo_cache<TM>(W) + (size_t)w * ne * 5 * nsp;
  double md_den = 1, md_l0 = 0, md_s0 = 1;
  const double* md_inv = nullptr;
  if constexpr (MD) {
    for (int i = lane; i < 2 * S.ndet; i += kWave) {
      mdl[i] = W.ldd[(size_t)w * 2 * S.ndet + i];
      mdl[2 * S.ndet + i] = W.sgd[(size_t)w * 2 * S.ndet + i];
    }
    __syncthreads();
    md_den = md_weights(S, mdl, mdl + 2 * S.ndet, bw, md_l0, md_s0, lane);
    md_inv = W.invd + (size_t)w * 2 * S.ndet * S.nsig * S.nsig;
    __syncthreads();
    for (int e = 0; e < ne; e++) {
      const int s = e < S.nup ? 0 : 1, k = e - (s ? S.nup : 0);
      double t[4];
      md_contract<4>(S, md_inv, s, k, (const double*)mog + ((size_t)e * 5 + 1) * nsp, nsp, bw, md_den, t, lane);
      if (lane == 0) {
        gD[3 * e] = t[0];
        gD[3 * e + 1] = t[1];
        gD[3 * e + 2] = t[2];
        lD[e] = t[3];
      }
    }
  }
  // determinant gradient / laplacian per electron: lane k owns row k of its spin's inverse
  if (A.use_dgl && !MD) {
    for (int e = lane; e < ne; e += kWave) {
      const double* d = W.dgl + ((size_t)w * ne + e) * 4;
      gD[3 * e] = d[0];
      gD[3 * e + 1] = d[1];
      gD[3 * e + 2] = d[2];
      lD[e] = d[3];
    }
  }
  // otherwise from the inverse and the MO cache: electron by electron, lanes over the MOs, so
  // every row is read coalesced
  for (int s = 0; s < 2 && !A.use_dgl && !MD; s++) {
    const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0;
    const TM* ig = inv_store<TM>(W) + ((size_t)w * 2 + s) * nsp * nsp;
    for (int k = 0; k < n; k++) {
      const int e = off + k;
      double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
      for (int j = lane; j < n; j += kWave) {
        const double iv = ig[(size_t)k * nsp + j];
        t1 += mog[((size_t)e * 5 + 1) * nsp + j] * iv;
        t2 += mog[((size_t)e * 5 + 2) * nsp + j] * iv;
        t3 += mog[((size_t)e * 5 + 3) * nsp + j] * iv;
        t4 += mog[((size_t)e * 5 + 4) * nsp + j] * iv;
      }
      t1 = wsum(t1);
      t2 = wsum(t2);
      t3 = wsum(t3);
      t4 = wsum(t4);
      if (lane == 0) {
        gD[3 * e] = t1;
        gD[3 * e + 1] = t2;
        gD[3 * e + 2] = t3;
        lD[e] = t4;
      }
    }
  }
  __syncthreads();
  // Jastrow per electron (lanes over electrons), kinetic energy, Coulomb
  double ke = 0, utot = 0, vei = 0, vee = 0;
  for (int i = lane; i < ne; i += kWave) {
    const double* pi = pos + 3 * i;
    double gx = 0, gy = 0, gz = 0, l = 0, u = 0;
    // Coulomb (FP64)
    const QW_CONST double* epos = cptr(S.eit_pos);
    for (int k = 0; k < S.na; k++) {
      const QW_CONST double* ap = epos + 3 * k;
      double x = pi[0] - ap[0], y = pi[1] - ap[1], z = pi[2] - ap[2];
      double r, ri;
      dist_inv(x * x + y * y + z * z, r, ri);
      vei += cptr(S.eit_Z)[k] * ri;
    }
    for (int j = 0; j < i; j++) {
      double x = pi[0] - pos[3 * j], y = pi[1] - pos[3 * j + 1], z = pi[2] - pos[3 * j + 2];
      double r, ri;
      dist_inv(x * x + y * y + z * z, r, ri);
      vee += ri;
    }
    // Jastrow: FP64 on the FP64 path, FP32 pair terms on the mixed path
    if constexpr (sizeof(TM) == sizeof(double)) {
      jastrow_local<double>(S, S.ee, S.eit, pos, i, u, gx, gy, gz, l);
    } else {
      float uf = 0, gxf = 0, gyf = 0, gzf = 0, lf = 0;
      jastrow_local<float>(S, S.eef, S.eitf, pos, i, uf, gxf, gyf, gzf, lf);
      u = uf;
      gx = gxf;
      gy = gyf;
      gz = gzf;
      l = lf;
    }
    utot += u;
    double dot = gD[3 * i] * gx + gD[3 * i + 1] * gy + gD[3 * i + 2] * gz;
    double lap = lD[i] + (l + (gx * gx + gy * gy + gz * gz)) + 2 * dot;
    ke += -0.5 * lap;
    if (A.mode == 0) {
      double* og = A.out_grad + ((size_t)w * ne + i) * 3;
      og[0] = gD[3 * i] + gx;
      og[1] = gD[3 * i + 1] + gy;
      og[2] = gD[3 * i + 2] + gz;
      A.out_lap[(size_t)w * ne + i] = lap;
    }
  }
  ke = wsum(ke);
  utot = wsum(utot);
  double vloc = -wsum(vei) + S.vii + wsum(vee);

  // nonlocal pseudopotential (src/system/Pseudopotential.cpp:596-815)
  double vnl = 0;
  if (PP && S.has_pp) {
    double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (A.mode == 0) {
      if (A.rot)
        for (int i = 0; i < 9; i++) rot[i] = A.rot[i];
    } else {
      random_rotation(A.seed, (unsigned)w, step_of(A.step, A.stepp), rot);
    }
    for (int a = 0; a < S.na; a++) {
      const int pk = cptr(S.atom_pp)[a];
      if (pk < 0) continue;
      const QW_CONST PPType& P = cptr(S.pp)[pk];
      const QW_CONST double* ap = cptr(S.atom_pos) + 3 * a;
      for (int e0 = 0; e0 < ne; e0 += kWave) {
        const int e = e0 + lane;
        bool want = false;
        double vl[kMaxPPL];
        if (e < ne) {
          double x = pos[3 * e] - ap[0], y = pos[3 * e + 1] - ap[1], z = pos[3 * e + 2] - ap[2];
          double r = sqrt(x * x + y * y + z * z);
          pp_radial(P, cptr(S.atom_Z)[a], r, vl);
          if (A.pp_mode == 0) {
            want = r < P.cutoff;
          } else {
            double strength = 0;
            for (int l = 0; l < P.L - 1; l++) strength += 10.0 * (2 * l + 1) * fabs(vl[l]);
            strength = fmin(1.0, strength);
            double rr = A.mode == 0 ? A.pp_rand[((size_t)w * S.na + a) * ne + e]
                                    : uniform(A.seed, (unsigned)w, step_of(A.step, A.stepp), (unsigned)(a * ne + e), TAG_PP);
            want = strength > rr;
          }
          vnl += vl[P.L - 1];  // local channel for every electron
          if (A.tneg && !want) A.tneg[((size_t)w * S.na + a) * ne + e] = 0.0;
        }
        unsigned long long mask = __ballot(want);
        while (mask) {
          const int b = __ffsll((long long)mask) - 1;
          mask &= mask - 1;
          const int ee = e0 + b;
          const int s = ee < S.nup ? 0 : 1, k = ee - (s ? S.nup : 0), n = s ? S.ndn : S.nup;
          const double* re = pos + 3 * ee;
          double od[3] = {re[0] - ap[0], re[1] - ap[1], re[2] - ap[2]};
          double orr = sqrt(od[0] * od[0] + od[1] * od[1] + od[2] * od[2]);
          double v[kMaxPPL];
          pp_radial(P, cptr(S.atom_Z)[a], orr, v);
          if (A.pp_mode == 1) {
            double strength = 0;
            for (int l = 0; l < P.L - 1; l++) strength += 10.0 * (2 * l + 1) * fabs(v[l]);
            strength = fmin(1.0, strength);
            if (fabs(strength) > 0)
              for (int l = 0; l < P.L - 1; l++) v[l] /= strength;
          }
          double gtmp[3];
          const double u_old = jastrow_e_val(S, ee, re, pos, lane);
          (void)gtmp;
          const double ivk =
              (!MD && lane < n) ? (double)inv_store<TM>(W)[(((size_t)w * 2 + s) * nsp + k) * nsp + lane] : 0.0;
          double nl = 0, nneg = 0;
          // the quadrature points kPPBatch at a time: one pass of the basis, the LCAO
          // (each coefficient load serving every point) and the Jastrow pairs per batch,
          // each point's value bitwise that of a pass of its own
          for (int q0 = 0; q0 < P.aip; q0 += kPPBatch) {
            const int nq = P.aip - q0 < kPPBatch ? P.aip - q0 : kPPBatch;
            double np_[kPPBatch][3], cs[kPPBatch];
#pragma unroll
            for (int i = 0; i < kPPBatch; i++) {
              const int q = i < nq ? q0 + i : q0;
              double qx = P.q[q][0] * rot[0] + P.q[q][1] * rot[3] + P.q[q][2] * rot[6];
              double qy = P.q[q][0] * rot[1] + P.q[q][1] * rot[4] + P.q[q][2] * rot[7];
              double qz = P.q[q][0] * rot[2] + P.q[q][1] * rot[5] + P.q[q][2] * rot[8];
              np_[i][0] = re[0] + (qx * orr - od[0]);
              np_[i][1] = re[1] + (qy * orr - od[1]);
              np_[i][2] = re[2] + (qz * orr - od[2]);
              double nd[3] = {np_[i][0] - ap[0], np_[i][1] - ap[1], np_[i][2] - ap[2]};
              double nr = sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
              cs[i] = (nd[0] * od[0] + nd[1] * od[1] + nd[2] * od[2]) / (nr * orr);
            }
            double Rq[kPPBatch], uq[kPPBatch];
            if constexpr (MD) {
              for (int i = 0; i < nq; i++) {
                const double phm = lcao_val_n(S, s, S.nmo[s], np_[i], chi, flist, lane);
                __syncthreads();
                if (lane < S.nmo[s]) phq[lane] = phm;
                __syncthreads();
                md_contract<1>(S, md_inv, s, k, phq, nsp, bw, md_den, &Rq[i], lane);
              }
            } else {
              double phi[kPPBatch];
              lcao_val_multi<kPPBatch>(S, s, np_, nq, chi, flist, phi, lane);
#pragma unroll
              for (int i = 0; i < kPPBatch; i++) Rq[i] = i < nq ? wsum(ivk * phi[i]) : 0.0;
            }
            jastrow_e_val_multi<kPPBatch>(S, ee, np_, nq, pos, uq, lane);
            for (int i = 0; i < nq; i++) {
              double ratio = Rq[i] * exp(uq[i] - u_old) * P.w[q0 + i];
              double ts = 0;
              for (int l = 0; l < P.L - 1; l++) ts += (2 * l + 1) * v[l] * legendre(cs[i], l);
              const double vxx = ts * ratio;
              nl += vxx;
              if (vxx < 0.0) nneg -= vxx;
            }
          }
          if (lane == 0) {
            vnl += nl;
            if (A.tneg) A.tneg[((size_t)w * S.na + a) * ne + ee] = nneg;
          }
        }
      }
    }
    vnl = wsum(vnl);
  }
  const double etot = ke + vloc + vnl;
  if (lane == 0) {
    double* el = W.elast + 4 * (size_t)w;
    el[0] = ke;
    el[1] = vloc;
    el[2] = vnl;
    el[3] = etot;
    if (A.mode == 0) {
      if constexpr (MD) {
        A.out_logpsi[w] = md_l0 + log(fabs(md_den)) + utot;
        A.out_sign[w] = md_s0 * (md_den < 0 ? -1.0 : 1.0);
      } else {
        A.out_logpsi[w] = W.logdet[2 * w] + W.logdet[2 * w + 1] + utot;
        A.out_sign[w] = W.sgn[2 * w] * W.sgn[2 * w + 1];
      }
      A.out_eparts[3 * w] = ke;
      A.out_eparts[3 * w + 1] = vloc;
      A.out_eparts[3 * w + 2] = vnl;
    } else if (A.mode == 3) {
      W.eold[w] = etot;
    } else {
      double wt = W.weight[w];
      if (A.mode == 2) {
        const DmcCtl& C = *A.ctl;
        double teffac = C.teff / 2.0;
        double eo = W.eold[w];
        double fbet = fmax(C.etrial - etot, C.etrial - eo);
        if (fbet > C.cut_stop) teffac = 0;
        else if (fbet > C.cut_start) teffac = teffac * (1. - (fbet - C.cut_start) / (C.cut_stop - C.cut_start));
        wt *= exp(teffac * (C.etrial * 2 - etot - eo));
        if (wt > C.max_weight) wt = C.max_weight;
        W.weight[w] = wt;
      }
      W.eold[w] = etot;
      double* bs = W.bsum + kBsum * (size_t)w;
      bs[0] += wt;
      bs[1] += wt * etot;
      bs[2] += wt * ke;
      bs[3] += wt * vloc;
      bs[4] += wt * vnl;
      bs[5] += 1;
      // unweighted first and second moments (per-point variances of the block log)
      bs[6] += etot;
      bs[7] += etot * etot;
      bs[8] += ke;
      bs[9] += ke * ke;
      bs[10] += vloc;
      bs[11] += vloc * vloc;
      bs[12] += vnl;
      bs[13] += vnl * vnl;
      bs[14] += wt * wt;
    }
  }
}

template <typename TM, bool MD = false, bool PP = true>
__global__ __launch_bounds__(64) void k_local(DevSys S, DevWalkers W, LocalArgs A) {
  extern __shared__ double lds[];
  local_body<TM, MD, PP>(S, W, A, (int)blockIdx.x, (int)threadIdx.x, lds);
}

// ---------------------------------------------------------------- k_tmove
struct TmoveArgs {
  int sc;                // 0: one draw per walker (TMOVES), 1: one draw per electron (TMOVESSC)
  int pp_mode;           // as in the k_local launch that wrote tneg
  const double* tneg;    // [nw][na][ne] from k_local
  unsigned long long seed;
  const unsigned* stepp; // device step counter (the step k_local used)
  const DmcCtl* ctl;     // tau
};

// The quadrature point of atom a that the draw t selects for electron e, scanning the atom's
// negative terms from the cumulative value c (the order and terms of k_local's nonlocal loop);
// writes the new position.  All lanes take part (LCAO over the MOs, Jastrow over electrons).
// MD: several determinants, ratios as weighted sums over the determinants (weights bw, sum
// den, the walker's inverses md_inv; phq: LDS row for the MO values at a quadrature point).
template <typename TM, bool MD>
__device__ inline void tmove_target(const DevSys& S, const DevWalkers& W, int w, const double* pos, int e, int a,
                                    double t, double c, double tau, const double rot[9], int pp_mode, double* chi,
                                    int* flist, double np_out[3], const double* bw, double den,
                                    const double* md_inv, double* phq, int lane) {
  const int nsp = S.nsp, s = e < S.nup ? 0 : 1, k = e - (s ? S.nup : 0), n = s ? S.ndn : S.nup;
  const QW_CONST PPType& P = cptr(S.pp)[cptr(S.atom_pp)[a]];
  const QW_CONST double* ap = cptr(S.atom_pos) + 3 * a;
  const double* re = pos + 3 * e;
  const double od[3] = {re[0] - ap[0], re[1] - ap[1], re[2] - ap[2]};
  const double orr = sqrt(od[0] * od[0] + od[1] * od[1] + od[2] * od[2]);
  double v[kMaxPPL];
  pp_radial(P, cptr(S.atom_Z)[a], orr, v);
  if (pp_mode == 1) {
    double strength = 0;
    for (int l = 0; l < P.L - 1; l++) strength += 10.0 * (2 * l + 1) * fabs(v[l]);
    strength = fmin(1.0, strength);
    if (fabs(strength) > 0)
      for (int l = 0; l < P.L - 1; l++) v[l] /= strength;
  }
  const double u_old = jastrow_e_val(S, e, re, pos, lane);
  const double ivk =
      (!MD && lane < n) ? (double)inv_store<TM>(W)[(((size_t)w * 2 + s) * nsp + k) * nsp + lane] : 0.0;
  int pick = -1;
  for (int q = 0; q < P.aip && pick < 0; q++) {
    const double qx = P.q[q][0] * rot[0] + P.q[q][1] * rot[3] + P.q[q][2] * rot[6];
    const double qy = P.q[q][0] * rot[1] + P.q[q][1] * rot[4] + P.q[q][2] * rot[7];
    const double qz = P.q[q][0] * rot[2] + P.q[q][1] * rot[5] + P.q[q][2] * rot[8];
    const double np_[3] = {re[0] + (qx * orr - od[0]), re[1] + (qy * orr - od[1]), re[2] + (qz * orr - od[2])};
    const double nd[3] = {np_[0] - ap[0], np_[1] - ap[1], np_[2] - ap[2]};
    const double nr = sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
    const double cs = (nd[0] * od[0] + nd[1] * od[1] + nd[2] * od[2]) / (nr * orr);
    double Rq;
    if constexpr (MD) {
      const double phm = lcao_val_n(S, s, S.nmo[s], np_, chi, flist, lane);
      __syncthreads();
      if (lane < S.nmo[s]) phq[lane] = phm;
      __syncthreads();
      md_contract<1>(S, md_inv, s, k, phq, nsp, bw, den, &Rq, lane);
    } else {
      const double phi = lcao_val(S, s, np_, chi, flist, lane);
      Rq = wsum(ivk * phi);
    }
    const double uq = jastrow_e_val(S, e, np_, pos, lane);
    const double ratio = Rq * exp(uq - u_old) * P.w[q];
    double ts = 0;
    for (int l = 0; l < P.L - 1; l++) ts += (2 * l + 1) * v[l] * legendre(cs, l);
    const double vxx = ts * ratio;
    if (vxx < 0.0) {
      c += tau * (-vxx);
      // the last negative term stands in when rounding leaves t just above the running sum
      for (int d = 0; d < 3; d++) np_out[d] = np_[d];
      if (t < c) pick = q;
    }
  }
}

// Casula T-moves after the local energy of a DMC step (src/method/Dmc_method.cpp:850-906,
// src/system/Pseudopotential.cpp:371-389,752-763), one wavefront per walker.  From k_local's
// per-(atom, electron) sums S_ae of -vxx over the negative nonlocal terms:
//   TMOVES    one uniform x (1 + tau sum S_ae) picks at most one term of the whole list, in
//             k_local's order (atom, electron, quadrature point);
//   TMOVESSC  per electron, one uniform x (1 + tau sum_a S_ae) over that electron's terms.
// The chosen terms are found again at the unmoved configuration (the atom from the sums, the
// quadrature point by recomputing that atom's terms), then applied in electron order as
// accepted moves: Sherman-Morrison on the inverse, the MO cache row, log|det| and sign.
template <typename TM, bool MD = false>
__global__ __launch_bounds__(64) void k_tmove(DevSys S, DevWalkers W, TmoveArgs A) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x, w = blockIdx.x;
  if (w >= W.nw) return;
  const int ne = S.ne, na = S.na, nsp = S.nsp;
  double* pos = lds;
  double* chi = pos + 3 * ne;
  double* newp = chi + 5 * S.chunk_nf;
  double* selt = newp + 3 * ne;
  double* selc = selt + ne;
  double* phib = selc + ne;
  double* rowk = phib + kWave;
  double* xs = rowk + kWave;
  int* sela = (int*)(xs + ne);
  int* flist = sela + ne;
  // MD: determinant logs/signs [2][2][nd], weights, quadrature / trial MO rows, rows k
  double* mdl = (double*)(flist + ((S.chunk_nf + 1) & ~1));
  double* bw = mdl + (MD ? 4 * S.ndet : 0);
  double* phq = bw + (MD ? S.ndet : 0);
  double* phn = phq + (MD ? nsp : 0);
  double* rwk = phn + (MD ? 5 * nsp : 0);
  for (int i = lane; i < 3 * ne; i += kWave) pos[i] = W.pos[(size_t)w * ne * 3 + i];
  const unsigned step = *A.stepp;
  const double tau = A.ctl->tau;
  const double* T = A.tneg + (size_t)w * na * ne;
  auto has_pp = [&](int a) { return cptr(S.atom_pp)[a] >= 0; };
  for (int e = lane; e < ne; e += kWave) sela[e] = -1;
  __syncthreads();
  if (A.sc) {
    for (int e = lane; e < ne; e += kWave) {
      double se = 0;
      for (int a = 0; a < na; a++)
        if (has_pp(a)) se += tau * T[(size_t)a * ne + e];
      const double t = uniform(A.seed, (unsigned)w, step, (unsigned)e, TAG_TMOVE) * (1.0 + se);
      if (t > 1.0) {
        double c = 1.0;
        for (int a = 0; a < na; a++) {
          const double x = has_pp(a) ? tau * T[(size_t)a * ne + e] : 0.0;
          if (x > 0) {
            sela[e] = a;
            selc[e] = c;
            if (t < c + x) break;
          }
          c += x;
        }
        selt[e] = t;
      }
    }
  } else {
    double part = 0;
    for (int a = 0; a < na; a++)
      if (has_pp(a))
        for (int e = lane; e < ne; e += kWave) part += tau * T[(size_t)a * ne + e];
    const double stot = wsum(part);
    const double t = uniform(A.seed, (unsigned)w, step, 0u, TAG_TMOVE) * (1.0 + stot);
    if (t > 1.0) {
      // the atom, from per-atom totals in order; then the electron, scanning that atom's sums
      double c = 1.0;
      int as = -1;
      double cs = 1.0;
      for (int a = 0; a < na; a++) {
        if (!has_pp(a)) continue;
        double xa = 0;
        for (int e = lane; e < ne; e += kWave) xa += tau * T[(size_t)a * ne + e];
        xa = wsum(xa);
        if (xa > 0) {
          as = a;
          cs = c;
          if (t < c + xa) break;
        }
        c += xa;
      }
      if (as >= 0) {
        __syncthreads();
        for (int e = lane; e < ne; e += kWave) xs[e] = tau * T[(size_t)as * ne + e];
        __syncthreads();
        if (lane == 0) {
          int es = -1;
          double ce = cs, c2 = cs;
          for (int i = 0; i < ne; i++) {
            const double x = xs[i];
            if (x > 0) {
              es = i;  // the last candidate stands in when rounding leaves t above the sum
              ce = c2;
              if (t < c2 + x) break;
            }
            c2 += x;
          }
          if (es >= 0) {
            sela[es] = as;
            selc[es] = ce;
            selt[es] = t;
          }
        }
      }
    }
  }
  __syncthreads();
  double rot[9];
  random_rotation(A.seed, (unsigned)w, step, rot);
  double md_den = 1, md_l0, md_s0;
  double* md_inv = nullptr;
  const int nd = S.ndet, nsg = S.nsig;
  if constexpr (MD) {
    for (int i = lane; i < 2 * nd; i += kWave) {
      mdl[i] = W.ldd[(size_t)w * 2 * nd + i];
      mdl[2 * nd + i] = W.sgd[(size_t)w * 2 * nd + i];
    }
    __syncthreads();
    md_den = md_weights(S, mdl, mdl + 2 * nd, bw, md_l0, md_s0, lane);
    md_inv = W.invd + (size_t)w * 2 * nd * nsg * nsg;
    __syncthreads();
  }
  // targets at the unmoved configuration
  int* flst = flist;
  for (int e = 0; e < ne; e++) {
    if (sela[e] < 0) continue;
    double np_[3];
    tmove_target<TM, MD>(S, W, w, pos, e, sela[e], selt[e], selc[e], tau, rot, A.pp_mode, chi, flst, np_, bw, md_den,
                         md_inv, phq, lane);
    if (lane == 0)
      for (int d = 0; d < 3; d++) newp[3 * e + d] = np_[d];
  }
  __syncthreads();
  // apply in electron order
  TM* invw = (sizeof(TM) == sizeof(double) ? (TM*)W.inv : (TM*)W.invf) + (size_t)w * 2 * nsp * nsp;
  TM* mow = (sizeof(TM) == sizeof(double) ? (TM*)W.mo : (TM*)W.mof) + (size_t)w * ne * 5 * nsp;
  for (int e = 0; e < ne; e++) {
    if (sela[e] < 0) continue;
    const int s = e < S.nup ? 0 : 1, k = e - (s ? S.nup : 0), n = s ? S.ndn : S.nup;
    const double np_[3] = {newp[3 * e], newp[3 * e + 1], newp[3 * e + 2]};
    double phi[5];
    if constexpr (MD) {
      // every determinant's inverse by Sherman-Morrison, as an accepted move of k_sweep_md
      const int nm = S.nmo[s];
      lcao_lap_n(S, s, nm, np_, chi, flist, phi, lane);
      __syncthreads();
      if (lane < nm)
        for (int c = 0; c < 5; c++) phn[c * nsp + lane] = phi[c];
      const QW_CONST int* oc = cptr(S.occ) + (size_t)s * nd * nsg;
      for (int i = lane; i < nd * n; i += kWave) {
        const int d = i / n, j = i % n;
        rwk[d * nsg + j] = md_inv[(((size_t)s * nd + d) * nsg + k) * nsg + j];
      }
      __syncthreads();
      for (int i = lane; i < nd * n; i += kWave) {
        const int d = i / n, r = i % n;
        const QW_CONST int* o = oc + (size_t)d * nsg;
        const double* rk = rwk + d * nsg;
        double* row = md_inv + (((size_t)s * nd + d) * nsg + r) * nsg;
        double Rd = 0, tr = 0;
        for (int j = 0; j < n; j++) {
          const double vj = phn[o[j]];
          Rd += rk[j] * vj;
          tr += row[j] * vj;
        }
        if (r == k) {
          const double f = 1.0 / Rd;
          for (int j = 0; j < n; j++) row[j] = rk[j] * f;
          W.ldd[(size_t)w * 2 * nd + s * nd + d] += log(fabs(Rd));
          if (Rd < 0) W.sgd[(size_t)w * 2 * nd + s * nd + d] = -W.sgd[(size_t)w * 2 * nd + s * nd + d];
        } else {
          const double pk = -tr / Rd;
          for (int j = 0; j < n; j++) row[j] += rk[j] * pk;
        }
      }
      if (lane < nm)
        for (int c = 0; c < 5; c++) mow[((size_t)e * 5 + c) * nsp + lane] = (TM)phn[c * nsp + lane];
      if (lane < 3) pos[3 * e + lane] = np_[lane];
      __syncthreads();
      continue;
    }
    lcao_lap(S, s, np_, chi, flist, phi, lane);
    TM* ig = invw + (size_t)s * nsp * nsp;
    const double rk = lane < n ? (double)ig[(size_t)k * nsp + lane] : 0.0;
    const double R = wsum(rk * (lane < n ? phi[0] : 0.0));
    phib[lane] = lane < n ? phi[0] : 0.0;
    rowk[lane] = rk;
    __syncthreads();
    if (lane < n) {
      TM* row = ig + (size_t)lane * nsp;
      double tk = 0;
      for (int j = 0; j < n; j++) tk += (double)row[j] * phib[j];
      if (lane == k) {
        for (int j = 0; j < n; j++) row[j] = (TM)(rowk[j] / R);
      } else {
        const double pk = -tk / R;
        for (int j = 0; j < n; j++) row[j] = (TM)((double)row[j] + rowk[j] * pk);
      }
      for (int c = 0; c < 5; c++) mow[((size_t)e * 5 + c) * nsp + lane] = (TM)phi[c];
    }
    if (lane == 0) {
      W.logdet[2 * w + s] += log(fabs(R));
      if (R < 0) W.sgn[2 * w + s] = -W.sgn[2 * w + s];
    }
    if (lane < 3) pos[3 * e + lane] = np_[lane];
    __syncthreads();
  }
  for (int i = lane; i < 3 * ne; i += kWave) W.pos[(size_t)w * ne * 3 + i] = pos[i];
}

// Dynamic LDS of k_tmove.
// With several determinants add (ndet, nsig, nsp): their logs, weights and row copies.
inline size_t tmove_lds_bytes(int ne, int chunk_nf, int ndet = 0, int nsig = 0, int nsp = 0) {
  const size_t md = ndet ? (size_t)5 * ndet + 6 * (size_t)nsp + (size_t)ndet * nsig + 1 : 0;
  return ((size_t)3 * ne + 5 * chunk_nf + 3 * ne + 3 * ne + 2 * kWave + md) * sizeof(double) +
         ((size_t)ne + chunk_nf + 1) * sizeof(int) + 16;
}

// ---------------------------------------------------------------- k_branch
// Single block.  Trial energy and effective timestep from the population, then the
// fixed-population pair comb: sort by weight (ties by walker index), pair the p-th largest
// weight >= 1.8 with the p-th smallest, keep one of the two with probability
// w_big / (w_big + w_small), both slots get the mean weight.
// Populations of up to kBranchLdsMax walkers are sorted in LDS (the same network on the same
// keys as the device-memory sort used above that size, whose every stage waits on L2).
constexpr int kBranchLdsMax = 8192;
struct BranchLds {
  double red[1024];
  long long redi[1024];
  int npair;
  unsigned long long skey[kBranchLdsMax];
  int sval[kBranchLdsMax];
};
// Bitonic sort of the walkers by weight (ties by walker index) into keys/vals[0..P): P a power
// of two >= nw, padded with the largest key.  Called with LDS or device-memory arrays.
__device__ __forceinline__ void branch_sort(const DevWalkers& W, unsigned long long* keys, int* vals, int nw, int P,
                                            int t) {
  for (int i = t; i < P; i += 1024) {
    if (i < nw) {
      keys[i] = (unsigned long long)__double_as_longlong(W.weight[i]);
      vals[i] = i;
    } else {
      keys[i] = ~0ull;
      vals[i] = i;
    }
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P; i += 1024) {
        int l = i ^ j;
        if (l > i) {
          bool up = (i & k) == 0;
          unsigned long long ki = keys[i], kl = keys[l];
          int vi = vals[i], vl = vals[l];
          bool gt = ki > kl || (ki == kl && vi > vl);
          if (gt == up) {
            keys[i] = kl;
            keys[l] = ki;
            vals[i] = vl;
            vals[l] = vi;
          }
        }
      }
      __syncthreads();
    }
  }
}
__device__ __forceinline__ void branch_body(const DevWalkers& W, DmcCtl* ctl, unsigned long long* keys, int* vals,
                                            int* pair_src, int* pair_dst, int nmoves, int do_branch,
                                            unsigned long long seed, unsigned event_host, const unsigned* eventp,
                                            int t, BranchLds& L) {
  const unsigned event = eventp ? *eventp : event_host;
  double* red = L.red;
  long long* redi = L.redi;
  int& npair = L.npair;
  const int nw = W.nw;
  double sw = 0;
  long long sa = 0;
  for (int i = t; i < nw; i += 1024) {
    sw += W.weight[i];
    sa += W.acc[i];
  }
  red[t] = sw;
  redi[t] = sa;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) {
      red[t] += red[t + o];
      redi[t] += redi[t + o];
    }
    __syncthreads();
  }
  DmcCtl& C = *ctl;
  if (t == 0) {
    if (nmoves > 0) C.teff = C.tau * (double)redi[0] / (double)nmoves;
    C.sumw = red[0];
    C.etrial = C.eref - C.feedback * log(red[0] / nw);
    C.npairs = 0;
  }
  __syncthreads();
  for (int i = t; i < nw; i += 1024) W.acc[i] = 0;
  if (!do_branch) return;
  int P = 1;
  while (P < nw) P <<= 1;
  const bool in_lds = P <= kBranchLdsMax;
  if (in_lds) {
    branch_sort(W, L.skey, L.sval, nw, P, t);
    vals = L.sval;
  } else {
    branch_sort(W, keys, vals, nw, P, t);
  }
  // pairs: p-th from the top with p-th from the bottom while the top weight >= 1.8
  if (t == 0) {
    int np = 0;
    for (int i = nw - 1; i > 1 && 2 * np + 1 < nw; i--, np++)
      if (W.weight[vals[i]] < 1.8) break;
    npair = np;
    C.npairs = np;
  }
  __syncthreads();
  for (int p = t; p < npair; p += 1024) {
    int big = vals[nw - 1 - p], sm = vals[p];
    double wb = W.weight[big], ws = W.weight[sm];
    double w1 = wb / (wb + ws);
    double u = uniform(seed, (unsigned)p, event, 0, TAG_BRANCH);
    double nwt = (wb + ws) / 2.0;
    if (w1 + u >= 1.0) {
      pair_src[p] = big;
      pair_dst[p] = sm;
    } else {
      pair_src[p] = sm;
      pair_dst[p] = big;
    }
    W.weight[big] = nwt;
    W.weight[sm] = nwt;
  }
}

static __global__ __launch_bounds__(1024) void k_branch(DevWalkers W, DmcCtl* ctl, unsigned long long* keys, int* vals,
                                                 int* pair_src, int* pair_dst, int nmoves, int do_branch,
                                                 unsigned long long seed, unsigned event_host,
                                                 const unsigned* eventp) {
  __shared__ BranchLds L;
  branch_body(W, ctl, keys, vals, pair_src, pair_dst, nmoves, do_branch, seed, event_host, eventp, (int)threadIdx.x, L);
}

// Copies the state of walker pair_src[p] into walker pair_dst[p] for p < ctl->npairs.
__device__ __forceinline__ void copy_body(const DevSys& S, const DevWalkers& W, const DmcCtl* ctl,
                                          const int* pair_src, const int* pair_dst, int p, int t) {
  if (p >= ctl->npairs) return;
  const int src = pair_src[p], dst = pair_dst[p];
  const int ne = S.ne, nsp = S.nsp;
  const size_t npos = (size_t)ne * 3, ninv = (size_t)2 * nsp * nsp, nmo = (size_t)ne * 5 * nsp;
  for (size_t i = t; i < npos; i += 256) W.pos[dst * npos + i] = W.pos[src * npos + i];
  if (W.inv)
    for (size_t i = t; i < ninv; i += 256) W.inv[dst * ninv + i] = W.inv[src * ninv + i];
  if (W.invf)
    for (size_t i = t; i < ninv; i += 256) W.invf[dst * ninv + i] = W.invf[src * ninv + i];
  if (W.mo)
    for (size_t i = t; i < nmo; i += 256) W.mo[dst * nmo + i] = W.mo[src * nmo + i];
  if (W.mof)
    for (size_t i = t; i < nmo; i += 256) W.mof[dst * nmo + i] = W.mof[src * nmo + i];
  if (W.invd) {
    const size_t nd = (size_t)2 * S.ndet * S.nsig * S.nsig, nl = (size_t)2 * S.ndet;
    for (size_t i = t; i < nd; i += 256) W.invd[dst * nd + i] = W.invd[src * nd + i];
    for (size_t i = t; i < nl; i += 256) {
      W.ldd[dst * nl + i] = W.ldd[src * nl + i];
      W.sgd[dst * nl + i] = W.sgd[src * nl + i];
    }
  }
  if (t < 2) {
    W.logdet[2 * dst + t] = W.logdet[2 * src + t];
    W.sgn[2 * dst + t] = W.sgn[2 * src + t];
  }
  if (t < 4) W.elast[4 * (size_t)dst + t] = W.elast[4 * (size_t)src + t];
  if (t == 0) W.eold[dst] = W.eold[src];
}

static __global__ __launch_bounds__(256) void k_copy(DevSys S, DevWalkers W, const DmcCtl* ctl, const int* pair_src,
                                              const int* pair_dst) {
  copy_body(S, W, ctl, pair_src, pair_dst, (int)blockIdx.x, (int)threadIdx.x);
}

// Block sums over walkers: out[0..kBsum-1] = sums of the bsum columns, out[kBsum] = current
// weight sum.
static __global__ __launch_bounds__(256) void k_reduce(DevWalkers W, double* out, int reset) {
  constexpr int NC = kBsum + 1;
  __shared__ double red[NC][256];
  const int t = threadIdx.x;
  double acc[NC];
  for (int c = 0; c < NC; c++) acc[c] = 0;
  for (int i = t; i < W.nw; i += blockDim.x) {
    for (int c = 0; c < kBsum; c++) acc[c] += W.bsum[kBsum * (size_t)i + c];
    acc[kBsum] += W.weight[i];
    if (reset)
      for (int c = 0; c < kBsum; c++) W.bsum[kBsum * (size_t)i + c] = 0;
  }
  for (int c = 0; c < NC; c++) red[c][t] = acc[c];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int c = 0; c < NC; c++) red[c][t] += red[c][t + o];
    __syncthreads();
  }
  if (t < NC) out[t] = red[t][0];
}

}  // namespace qw
