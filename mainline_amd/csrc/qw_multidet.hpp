// Several determinants, FP64 (Psi = sum_d c_d D_up,d D_dn,d times the Jastrow factor):
//
//   k_inv_md       LU-free Gauss-Jordan with partial pivoting of every determinant's Slater
//                  matrix (one block per walker, spin and determinant), from the MO cache,
//                  as the transposed inverse, log|det| and sign (src/wavefunction/Slat_wf.h:948-1016,
//                  src/program_utils/MatrixAlgebrac.cpp:316-421)
//   k_sweep_md     moves of one spin channel: each determinant keeps its own inverse and is
//                  updated by Sherman-Morrison (the reference's updates without excitation
//                  tables, src/wavefunction/Slat_wf.h:706-800,858-903); ratios, drift and
//                  acceptance as the one-determinant sweep with the determinant parts replaced by
//                  weighted sums over determinants (src/wavefunction/Slat_wf.h:905-935,1021-1139)
//
// The local energy is k_local<double, true>.  Weights: b_d = sign_d sign_0 exp(l_d - l_0) of
// the terms c_d D_up,d D_dn,d (log l_d, sign sign_d), pivoting on the first term as the
// reference's sum of log values does (src/program_utils/MatrixAlgebra.h:92-99).  Up to 32
// electrons per spin and 64 orbitals per spin list.
#pragma once

#include "qw_kernels.hpp"

namespace qw {

// Gauss-Jordan on [A | I] in LDS, lanes over rows; A(r, c) = phi_{occ_d(c)}(r_r) of spin s.
static __global__ __launch_bounds__(64) void k_inv_md(DevSys S, DevWalkers W, int nw) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x, nd = S.ndet, ns = S.nsig;
  const int blk = (int)blockIdx.x, d = blk % nd, s = (blk / nd) & 1, w = blk / (2 * nd);
  if (w >= nw) return;
  const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0, nsp = S.nsp, n2 = 2 * n;
  double* a = lds;  // [n][2n]
  double* out = W.invd + (((size_t)w * 2 + s) * nd + d) * ns * ns;
  double* ldo = W.ldd + ((size_t)w * 2 + s) * nd + d;
  double* sgo = W.sgd + ((size_t)w * 2 + s) * nd + d;
  if (n == 0) {
    if (lane == 0) {
      *ldo = 0;
      *sgo = 1;
    }
    return;
  }
  const QW_CONST int* o = cptr(S.occ) + ((size_t)s * nd + d) * ns;
  const double* mo = W.mo + (size_t)w * S.ne * 5 * nsp;
  if (lane < n)
    for (int c = 0; c < n2; c++)
      a[lane * n2 + c] = c < n ? mo[(size_t)(off + lane) * 5 * nsp + o[c]] : (c - n == lane ? 1.0 : 0.0);
  __syncthreads();
  double logdet = 0, sgn = 1;
  bool singular = false;
  for (int c = 0; c < n && !singular; c++) {
    // pivot: the largest |a(r, c)| over r >= c, lowest row on ties
    double v = (lane >= c && lane < n) ? fabs(a[lane * n2 + c]) : -1.0;
    int p = lane;
    for (int sh = 1; sh < kWave; sh <<= 1) {
      const double v2 = __shfl_xor(v, sh);
      const int p2 = __shfl_xor(p, sh);
      if (v2 > v || (v2 == v && p2 < p)) {
        v = v2;
        p = p2;
      }
    }
    const double piv = a[p * n2 + c];
    if (piv == 0.0) {
      singular = true;
      break;
    }
    __syncthreads();
    if (p != c) {
      for (int q = lane; q < n2; q += kWave) {
        const double t = a[c * n2 + q];
        a[c * n2 + q] = a[p * n2 + q];
        a[p * n2 + q] = t;
      }
      sgn = -sgn;
    }
    logdet += log(fabs(piv));
    if (piv < 0) sgn = -sgn;
    __syncthreads();
    const double ip = 1.0 / piv;
    for (int q = lane; q < n2; q += kWave) a[c * n2 + q] *= ip;
    __syncthreads();
    if (lane < n && lane != c) {
      const double f = a[lane * n2 + c];
      for (int q = 0; q < n2; q++) a[lane * n2 + q] -= f * a[c * n2 + q];
    }
    __syncthreads();
  }
  // transposed inverse: out(k, j) = A^-1(j, k)
  for (int i = lane; i < n * n; i += kWave) {
    const int k = i / n, j = i % n;
    out[(size_t)k * ns + j] = singular ? 0.0 : a[j * n2 + n + k];
  }
  if (lane == 0) {
    *ldo = singular ? -1e201 : logdet;
    *sgo = singular ? 1.0 : sgn;
  }
}

inline size_t inv_md_lds_bytes(int nsig) { return (size_t)2 * nsig * nsig * sizeof(double) + 16; }

// LDS of k_sweep_md.
inline size_t sweep_md_lds_bytes(int ne, int ndet, int nsig, int nsp, int chunk_nf, int depth) {
  return ((size_t)3 * ne + 4 * (size_t)ndet + ndet + 5 * (size_t)nsp + (size_t)ndet * nsig + 5 * (size_t)chunk_nf) *
             sizeof(double) +
         (size_t)chunk_nf * sizeof(int) + sweep_rand_bytes(kMdMaxSpin, depth) + 64;
}

template <int DEPTH>
__global__ __launch_bounds__(64) void k_sweep_md(DevSys S, DevWalkers W, SweepArgs A) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x, w = blockIdx.x;
  if (w >= W.nw) return;
  const int ne = S.ne, s = A.s, n = s ? S.ndn : S.nup, nsp = S.nsp, off = s ? S.nup : 0;
  const int nd = S.ndet, ns = S.nsig, nm = S.nmo[s];
  double* pos = lds;
  double* ld = pos + 3 * ne;  // [2][nd]
  double* sg = ld + 2 * nd;   // [2][nd]
  double* b = sg + 2 * nd;    // [nd]
  double* phin = b + nd;      // [5][nsp] MO values at the trial point
  double* rowk = phin + 5 * nsp;  // [nd][ns] row k of every inverse before an update
  double* rnd = rowk + (size_t)nd * ns;
  double* chi = rnd + sweep_rand_bytes(kMdMaxSpin, DEPTH) / sizeof(double);
  int* flist = (int*)(chi + 5 * S.chunk_nf);
  for (int i = lane; i < 3 * ne; i += kWave) pos[i] = W.pos[(size_t)w * ne * 3 + i];
  for (int i = lane; i < 2 * nd; i += kWave) {
    ld[i] = W.ldd[(size_t)w * 2 * nd + i];
    sg[i] = W.sgd[(size_t)w * 2 * nd + i];
  }
  sweep_randoms<DEPTH>(A, (unsigned)w, off, rnd, lane);
  double* invw = W.invd + (size_t)w * 2 * nd * ns * ns;
  const QW_CONST int* oc = cptr(S.occ) + (size_t)s * nd * ns;
  double* mog = W.mo + (size_t)w * ne * 5 * nsp;
  const double tau = A.tau, stau = sqrt(tau);
  int nacc = 0;
  __syncthreads();
  for (int k = A.k0; k < A.k1; k++) {
    const int e = off + k;
    double r0[3] = {pos[3 * e], pos[3 * e + 1], pos[3 * e + 2]};
    double l0, s0;
    const double den = md_weights(S, ld, sg, b, l0, s0, lane);
    __syncthreads();
    // drift at the old position: determinant part from the cached MO row, Jastrow part fresh
    double gd[3], gj0[3], d0[3];
    md_contract<3>(S, invw, s, k, mog + ((size_t)e * 5 + 1) * nsp, nsp, b, den, gd, lane);
    const double u0 = jastrow_e(S, e, r0, pos, gj0, nullptr, lane);
    for (int d = 0; d < 3; d++) d0[d] = gd[d] + gj0[d];
    double phi[5];
    double r1[3], d1[3], ln1 = 0;
    bool accepted = false;
    double acc_first = 0;
    for (int dep = 1; dep <= DEPTH && !accepted; dep++) {
      const double t = dep == 1 ? tau : tau * 0.5, st = dep == 1 ? stau : sqrt(tau * 0.5);
      const double* rq = rnd + ((size_t)(k - A.k0) * DEPTH + (dep - 1)) * 4;
      double dr[3], rn[3], dn[3];
      lim_drift(d0, t, dr);
      for (int d = 0; d < 3; d++) rn[d] = r0[d] + (rq[d] * st + dr[d]);
      lcao_lap_n(S, s, nm, rn, chi, flist, phi, lane);
      __syncthreads();
      if (lane < nm)
        for (int c = 0; c < 5; c++) phin[c * nsp + lane] = phi[c];
      __syncthreads();
      double tt[4];
      md_contract<4>(S, invw, s, k, phin, nsp, b, den, tt, lane);
      const double t0 = tt[0];
      double gj[3];
      const double un = jastrow_e(S, e, rn, pos, gj, nullptr, lane);
      dn[0] = tt[1] / t0 + gj[0];
      dn[1] = tt[2] / t0 + gj[1];
      dn[2] = tt[3] / t0 + gj[2];
      const bool neg = t0 < 0;
      double acc, lnn = 0;
      if constexpr (DEPTH == 1) {
        acc = (A.restrict_nodes && neg)
                  ? 0.0
                  : fmin(1.0, t0 * t0 * exp(2.0 * (un - u0) + (trans(rn, dn, r0, t) - trans(r0, d0, rn, t))));
      } else if (dep == 1) {
        lnn = log(fabs(t0)) + (un - u0);
        const double rr = exp(lnn);
        acc = (A.restrict_nodes && neg) ? 0.0
                                        : fmin(1.0, rr * rr * exp(trans(rn, dn, r0, t) - trans(r0, d0, rn, t)));
        acc_first = acc;
        for (int d = 0; d < 3; d++) r1[d] = rn[d], d1[d] = dn[d];
        ln1 = lnn;
      } else {
        lnn = log(fabs(t0)) + (un - u0);
        const double rr = exp(lnn);
        const double pt =
            (trans(rn, dn, r1, tau) - trans(r0, d0, r1, tau)) + (trans(rn, dn, r0, t) - trans(r0, d0, rn, t));
        const double r21 = exp(ln1 - lnn);
        const double a21 = fmin(1.0, r21 * r21 * exp(trans(r1, d1, rn, tau) - trans(rn, dn, r1, tau)));
        const double num = 1 - a21, dd = 1 - acc_first;
        if (A.restrict_nodes && neg) acc = 0;
        else if (fabs(dd) < 1e-10) acc = fabs(num) < 1e-10 ? 0.0 : 1.0;
        else acc = fmin(1.0, rr * rr * exp(pt) * num / dd);
      }
      if (acc + rq[3] > 1.0) {
        accepted = true;
        for (int d = 0; d < 3; d++) r1[d] = rn[d];
      }
    }
    if (accepted) {
      // every determinant: Sherman-Morrison for the replaced row k with its own orbitals
      for (int i = lane; i < nd * n; i += kWave) {
        const int d = i / n, j = i % n;
        rowk[d * ns + j] = invw[(((size_t)s * nd + d) * ns + k) * ns + j];
      }
      __syncthreads();
      for (int i = lane; i < nd * n; i += kWave) {
        const int d = i / n, r = i % n;
        const QW_CONST int* o = oc + (size_t)d * ns;
        const double* rk = rowk + d * ns;
        double Rd = 0, tr = 0;
        double* row = invw + (((size_t)s * nd + d) * ns + r) * ns;
        for (int j = 0; j < n; j++) {
          const double v = phin[o[j]];
          Rd += rk[j] * v;
          tr += row[j] * v;
        }
        if (r == k) {
          const double f = 1.0 / Rd;
          for (int j = 0; j < n; j++) row[j] = rk[j] * f;
          ld[s * nd + d] += log(fabs(Rd));
          if (Rd < 0) sg[s * nd + d] = -sg[s * nd + d];
        } else {
          const double pk = -tr / Rd;
          for (int j = 0; j < n; j++) row[j] += rk[j] * pk;
        }
      }
      if (lane < nm)
        for (int c = 0; c < 5; c++) mog[((size_t)e * 5 + c) * nsp + lane] = phin[c * nsp + lane];
      nacc++;
      if (lane < 3) pos[3 * e + lane] = r1[lane];
    }
    __syncthreads();
  }
  for (int i = lane; i < 3 * ne; i += kWave) W.pos[(size_t)w * ne * 3 + i] = pos[i];
  for (int i = lane; i < 2 * nd; i += kWave) {
    W.ldd[(size_t)w * 2 * nd + i] = ld[i];
    W.sgd[(size_t)w * 2 * nd + i] = sg[i];
  }
  if (lane == 0) W.acc[w] += nacc;
}

// OPTIMIZE_DET: derivatives with respect to the coefficients of CSFs 1.. (CSF c covers
// determinants csf_d0[c] .. csf_d0[c+1]-1 with weights csf_w), one wavefront per configuration
// (src/wavefunction/Slat_wf.h:555-686, the cross term of src/wavefunction/Wavefunction.cpp:27-113):
//   grad(p) = sum_j w_j D_j / Psi_S,
//   gd(p, e, c) = sum_j w_j D_j (d_c D_j / D_j - d_c Psi_S / Psi_S) / Psi_S
//                 (+ 2 gd(p, e, 0:3) . vj(e) on the laplacian when the Jastrow has parameters),
// D_j = D_up,j D_dn,j, Psi_S = sum_d c_d D_d, d_c = d/dx, d/dy, d/dz, laplacian of electron e;
// vj: gradient of the parameterised Jastrow pieces (from k_parm_deriv) or null.
static __global__ __launch_bounds__(64) void k_parm_det(DevSys S, DevWalkers W, const int* csf_d0, const double* csf_w,
                                                 int ncsf, int nw, const double* vj, double* grad, double* gd) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x, w = blockIdx.x;
  if (w >= nw) return;
  const int ne = S.ne, nd = S.ndet, ns = S.nsig, nsp = S.nsp, np = ncsf - 1;
  double* mdl = lds;          // [4][nd]
  double* bw = mdl + 4 * nd;  // [nd]
  double* ed = bw + nd;       // [nd] D_d / (c_0 D_0)
  double* tot = ed + nd;      // [ne][4] determinant parts of Psi_S
  for (int i = lane; i < 2 * nd; i += kWave) {
    mdl[i] = W.ldd[(size_t)w * 2 * nd + i];
    mdl[2 * nd + i] = W.sgd[(size_t)w * 2 * nd + i];
  }
  __syncthreads();
  double l0, s0;
  const double den = md_weights(S, mdl, mdl + 2 * nd, bw, l0, s0, lane);
  for (int d = lane; d < nd; d += kWave)
    ed[d] = mdl[2 * nd + d] * mdl[3 * nd + d] * s0 * exp((mdl[d] + mdl[nd + d]) - l0);
  const double* minv = W.invd + (size_t)w * 2 * nd * ns * ns;
  const double* mog = W.mo + (size_t)w * ne * 5 * nsp;
  __syncthreads();
  for (int e = 0; e < ne; e++) {
    const int s = e < S.nup ? 0 : 1, k = e - (s ? S.nup : 0);
    double t[4];
    md_contract<4>(S, minv, s, k, mog + ((size_t)e * 5 + 1) * nsp, nsp, bw, den, t, lane);
    if (lane == 0)
      for (int c = 0; c < 4; c++) tot[4 * e + c] = t[c];
  }
  __syncthreads();
  for (int p = 0; p < np; p++) {
    const int d0 = csf_d0[p + 1], d1 = csf_d0[p + 2];
    if (lane == 0) {
      double g = 0;
      for (int j = d0; j < d1; j++) g += csf_w[j] * ed[j] / den;
      grad[(size_t)w * np + p] = g;
    }
    for (int e = lane; e < ne; e += kWave) {
      const int s = e < S.nup ? 0 : 1, k = e - (s ? S.nup : 0), n = s ? S.ndn : S.nup;
      double acc[4] = {0, 0, 0, 0};
      for (int j = d0; j < d1; j++) {
        const double* row = minv + (((size_t)s * nd + j) * ns + k) * ns;
        const QW_CONST int* o = cptr(S.occ) + ((size_t)s * nd + j) * ns;
        double G[4] = {0, 0, 0, 0};
        for (int q = 0; q < n; q++)
          for (int c = 0; c < 4; c++) G[c] += mog[((size_t)e * 5 + 1 + c) * nsp + o[q]] * row[q];
        for (int c = 0; c < 4; c++) acc[c] += csf_w[j] * ed[j] * (G[c] - tot[4 * e + c]) / den;
      }
      if (vj) acc[3] += 2 * (acc[0] * vj[((size_t)w * ne + e) * 3] + acc[1] * vj[((size_t)w * ne + e) * 3 + 1] +
                             acc[2] * vj[((size_t)w * ne + e) * 3 + 2]);
      double* out = gd + (((size_t)w * np + p) * ne + e) * 4;
      for (int c = 0; c < 4; c++) out[c] = acc[c];
    }
  }
}

inline size_t parm_det_lds_bytes(int ne, int ndet) { return ((size_t)6 * ndet + 4 * (size_t)ne) * sizeof(double) + 16; }

}  // namespace qw
