// Linear-parameter derivatives of ln Psi at fixed configurations: Wavefunction::getParmDeriv
// of SLATER-JASTROW with one determinant and JASTROW2 (src/wavefunction/Slat_Jastrow.cpp:171-193,
// Slat_wf.h:555-575,680-686, Jastrow2_wf.cpp:1486-1553,1900-1945, Jastrow2_one.cpp:214-248,
// Jastrow2_two.cpp:153-180,407-440, Wavefunction.cpp:27-113).
//
// Each parameter p multiplies one Jastrow basis function: an electron-ion function summed
// over the atoms of one type, or an electron-electron function over all pairs, like-spin
// pairs or unlike-spin pairs.  With the basis sums f_p(e) (value, gradient and laplacian with
// respect to r_e):
//   gradient[p]          = sum over atoms/pairs of the function values (pairs counted once)
//   gradderiv[p][e][0:3] = grad_e f_p(e)
//   gradderiv[p][e][3]   = lap_e f_p(e) + 2 grad_e f_p(e) . val_gradient[e]
//   val_gradient[e]      = grad_e ln D + sum_p c_p grad_e f_p(e)
// val_gradient is the gradient of the determinant and of the Jastrow pieces that have
// parameters: pieces marked FREEZE take no part (the reference adds a piece's gradient only
// when it has parameters).  The hessian is gradient gradient^T (Jastrow2_wf.cpp:1546-1551).
#pragma once
#include "qw_kernels.hpp"

namespace qw {

struct JParmDev {
  int body;    // 1: electron-ion, 2: electron-electron
  int spin;    // electron-electron pairs: 0 all, 1 like spins, 2 unlike spins
  int a0, a1;  // electron-ion: range in the atom index list
  double coef;
};

// One block (wavefront) per configuration; outputs grad[nw][np], gd[nw][np][ne][4],
// vg[nw][ne][3].  LDS: positions, determinant gradients, Jastrow gradients (3 x 3 x ne doubles).
// MD: several determinants (the determinant gradient is their weighted sum); vjo, when set,
// receives the gradient of the parameterised Jastrow pieces alone ([nw][ne][3]).
template <typename TM, bool MD = false>
__global__ __launch_bounds__(64) void k_parm_deriv(DevSys S, DevWalkers W, const JParmDev* P, const JRad<double>* U,
                                                   const int* patoms, int np, int nw, double* grad, double* gd,
                                                   double* vg, double* vjo) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x, w = blockIdx.x;
  if (w >= nw) return;
  const int ne = S.ne, nsp = S.nsp;
  double* pos = lds;
  double* gD = pos + 3 * ne;
  double* vj = gD + 3 * ne;
  for (int i = lane; i < 3 * ne; i += kWave) {
    pos[i] = W.pos[(size_t)w * ne * 3 + i];
    vj[i] = 0.0;
  }
  // determinant gradient per electron from the inverse and the MO cache: electron by
  // electron, lanes over the MOs
  const TM* mog = mo_cache<TM>(W) + (size_t)w * ne * 5 * nsp;
  if constexpr (MD) {
    double* mdl = vj + 3 * ne;  // [4][ndet] logs and signs, then [ndet] weights
    double* bw = mdl + 4 * S.ndet;
    for (int i = lane; i < 2 * S.ndet; i += kWave) {
      mdl[i] = W.ldd[(size_t)w * 2 * S.ndet + i];
      mdl[2 * S.ndet + i] = W.sgd[(size_t)w * 2 * S.ndet + i];
    }
    __syncthreads();
    double l0, s0;
    const double den = md_weights(S, mdl, mdl + 2 * S.ndet, bw, l0, s0, lane);
    const double* minv = W.invd + (size_t)w * 2 * S.ndet * S.nsig * S.nsig;
    __syncthreads();
    for (int e = 0; e < ne; e++) {
      const int s = e < S.nup ? 0 : 1, k = e - (s ? S.nup : 0);
      double t[3];
      md_contract<3>(S, minv, s, k, (const double*)mog + ((size_t)e * 5 + 1) * nsp, nsp, bw, den, t, lane);
      if (lane == 0)
        for (int d = 0; d < 3; d++) gD[3 * e + d] = t[d];
    }
  }
  for (int s = 0; s < 2 && !MD; s++) {
    const int n = s ? S.ndn : S.nup, off = s ? S.nup : 0;
    const TM* ig = inv_store<TM>(W) + ((size_t)w * 2 + s) * nsp * nsp;
    for (int k = 0; k < n; k++) {
      const int e = off + k;
      double t[3] = {0, 0, 0};
      for (int j = lane; j < n; j += kWave) {
        const double iv = ig[(size_t)k * nsp + j];
        for (int d = 0; d < 3; d++) t[d] += (double)mog[((size_t)e * 5 + 1 + d) * nsp + j] * iv;
      }
      for (int d = 0; d < 3; d++) t[d] = wsum(t[d]);
      if (lane == 0)
        for (int d = 0; d < 3; d++) gD[3 * e + d] = t[d];
    }
  }
  __syncthreads();
  for (int p = 0; p < np; p++) {
    const JParmDev d = P[p];
    double gsum = 0;
    for (int e = lane; e < ne; e += kWave) {
      const double* pe = pos + 3 * e;
      double f0 = 0, gx = 0, gy = 0, gz = 0, l = 0;
      if (d.body == 1) {
        with_jrad<double, true>(cptr(U)[p], [&](const auto& ev) {
          for (int q = d.a0; q < d.a1; q++) {
            const int a = patoms[q];
            const double x = pe[0] - S.atom_pos[3 * a], y = pe[1] - S.atom_pos[3 * a + 1],
                         z = pe[2] - S.atom_pos[3 * a + 2];
            double r, ri, gf = 0;
            dist_inv(x * x + y * y + z * z, r, ri);
            ev(0.0, r, ri, f0, gf, l);
            gx += gf * x;
            gy += gf * y;
            gz += gf * z;
          }
        });
      } else {
        const bool eup = e < S.nup;
        with_jrad<double, true>(cptr(U)[p], [&](const auto& ev) {
          for (int j = 0; j < ne; j++) {
            const bool unlike = (j < S.nup) != eup;
            if (j == e || (d.spin == 1 && unlike) || (d.spin == 2 && !unlike)) continue;
            const double x = pe[0] - pos[3 * j], y = pe[1] - pos[3 * j + 1], z = pe[2] - pos[3 * j + 2];
            double r, ri, gf = 0, u = 0;
            dist_inv(x * x + y * y + z * z, r, ri);
            ev(unlike ? 1.0 : 0.0, r, ri, u, gf, l);
            if (j > e) f0 += u;  // each pair's value once
            gx += gf * x;
            gy += gf * y;
            gz += gf * z;
          }
        });
      }
      gsum += f0;
      double* o = gd + (((size_t)w * np + p) * ne + e) * 4;
      o[0] = gx;
      o[1] = gy;
      o[2] = gz;
      o[3] = l;
      vj[3 * e] += d.coef * gx;
      vj[3 * e + 1] += d.coef * gy;
      vj[3 * e + 2] += d.coef * gz;
    }
    gsum = wsum(gsum);
    if (lane == 0) grad[(size_t)w * np + p] = gsum;
  }
  // cross terms of the laplacian derivative with the gradient of the parameterised Jastrow
  // pieces and of the determinant
  for (int e = lane; e < ne; e += kWave) {
    const double vx = vj[3 * e] + gD[3 * e], vy = vj[3 * e + 1] + gD[3 * e + 1], vz = vj[3 * e + 2] + gD[3 * e + 2];
    for (int p = 0; p < np; p++) {
      double* o = gd + (((size_t)w * np + p) * ne + e) * 4;
      o[3] += 2 * (o[0] * vx + o[1] * vy + o[2] * vz);
    }
    double* v = vg + ((size_t)w * ne + e) * 3;
    v[0] = vx;
    v[1] = vy;
    v[2] = vz;
    if (vjo)
      for (int d = 0; d < 3; d++) vjo[((size_t)w * ne + e) * 3 + d] = vj[3 * e + d];
  }
}

}  // namespace qw
