// Per-move building blocks of the sweep kernels: system tables staged in LDS, the two-pass
// sparse basis evaluation and the LCAO accumulation over active functions.
//
//   pass 1  lane = shell: distance to its atom from the LDS copies, activity r < cutoff
//           (the functions the reference's CUTOFF_MO loop visits,
//           src/orbitals/MO_matrix_cutoff.h:516-530); active shells are compacted in shell
//           order into an LDS list.  No global memory is touched.
//   pass 2  lane = active shell: one spline-row load per lane (all lanes in flight at once),
//           values / gradients / laplacians of its functions written packed into the basis
//           tile, with the row offset f*n of each function's MO coefficients.  A batch holds
//           at most `cap` functions; the caller contracts a batch, then evaluates the next.
#pragma once
#include "qw_device.hpp"

namespace qw {

// LDS views of the system tables (filled once per kernel by stage_sys).
struct SysLds {
  double* apos;    // [na][4] atom positions (shell order)
  double* epos;    // [na][4] atom positions in e-i Jastrow type order
  ShellLds* sh;    // [nsh]
  double* tabs;    // [ntab][2] 1/spacing, spacing
  int* actl;       // [nsh] active shells of the current point
};

__host__ __device__ inline size_t sys_lds_bytes(int na, int nsh, int ntab) {
  return (size_t)8 * na * sizeof(double) + (size_t)nsh * sizeof(ShellLds) + (size_t)2 * ntab * sizeof(double) +
         (size_t)nsh * sizeof(int);
}

// Carves the views from `base` (16-byte aligned) and returns the first 16-byte aligned byte
// past them.  Only pointer arithmetic on the LDS base is used, so the compiler keeps LDS
// (ds_*) instructions for every access.
__device__ inline char* sys_lds_carve(const DevSys& S, char* base, SysLds& V) {
  V.apos = (double*)base;
  V.epos = V.apos + 4 * S.na;
  V.sh = (ShellLds*)(V.epos + 4 * S.na);
  V.tabs = (double*)(V.sh + S.nsh);
  V.actl = (int*)(V.tabs + 2 * S.ntab);
  return base + ((sys_lds_bytes(S.na, S.nsh, S.ntab) + 15) & ~(size_t)15);
}

__device__ inline void stage_sys(const DevSys& S, const SysLds& V, int lane, int stride = kWave) {
  for (int a = lane; a < S.na; a += stride) {
    for (int d = 0; d < 3; d++) {
      V.apos[4 * a + d] = S.atom_pos[3 * a + d];
      V.epos[4 * a + d] = S.eit_pos[3 * a + d];
    }
    V.apos[4 * a + 3] = 0.0;
    V.epos[4 * a + 3] = 0.0;
  }
  for (int i = lane; i < S.nsh; i += stride) V.sh[i] = S.shl[i];
  for (int i = lane; i < 2 * S.ntab; i += stride) V.tabs[i] = S.tabs[i];
}

// Pass 1: active shells of point p, compacted into V.actl; returns their count.
__device__ inline int basis_active(const DevSys& S, const SysLds& V, const double p[3], int lane) {
  int nact = 0;
  for (int s0 = 0; s0 < S.nsh; s0 += kWave) {
    const int sh = s0 + lane;
    bool act = false;
    if (sh < S.nsh) {
      const ShellLds d = V.sh[sh];
      const double* ap = V.apos + 4 * shl_atom(d.meta);
      const double x = p[0] - ap[0], y = p[1] - ap[1], z = p[2] - ap[2];
      act = sqrt(x * x + y * y + z * z) < d.rc;
    }
    const unsigned long long m = __ballot(act);
    if (act) V.actl[nact + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = sh;
    nact += __popcll(m);
  }
  return nact;
}

// Pass 2, one batch: active shells i0, i0+1, ... (lane i0+l) while their functions fit in `cap`.
// chi[c*ld + b] (c = value, d/dx, d/dy, d/dz, laplacian) and roff[b] = f*n for the packed
// functions b < *nfun.  TS: spline element type (double or float); T: tile element type.
// Returns the number of shells consumed.
template <typename T, typename TS>
__device__ inline int basis_batch_lap(const DevSys& S, const SysLds& V, const TS* spline, const double p[3], int i0,
                                      int nact, int n, T* chi, int ld, int* roff, int cap, int* nfun, int lane) {
  const int i = i0 + lane;
  ShellLds d;
  int nf = 0;
  if (i < nact) {
    d = V.sh[V.actl[i]];
    nf = shl_nf(d.meta);
  }
  int tot;
  const int at = wave_excl_scan(nf, lane, &tot);
  const bool fit = nf > 0 && at + nf <= cap;
  const unsigned long long fm = __ballot(fit);
  const int take = __popcll(fm);
  *nfun = take ? __builtin_amdgcn_readlane(at + nf, take - 1) : 0;
  if (fit) {
    const int atom = shl_atom(d.meta), tab = shl_table(d.meta), sym = shl_sym(d.meta);
    const double* ap = V.apos + 4 * atom;
    const double x = p[0] - ap[0], y = p[1] - ap[1], z = p[2] - ap[2];
    const double r = sqrt(x * x + y * y + z * z);
    int iv = (int)(r * V.tabs[2 * tab]);
    iv = iv < 0 ? 0 : (iv >= S.nint ? S.nint - 1 : iv);
    const TS* c = spline + ((size_t)tab * S.nint + iv) * 4;
    const TS c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    if constexpr (sizeof(T) == sizeof(double)) {
      const double h = r - iv * V.tabs[2 * tab + 1];
      const double f = c0 + h * (c1 + h * (c2 + h * c3));
      const double fd = (c1 + h * (2 * c2 + h * (3 * c3))) / r;
      const double f2 = 2 * c2 + 6 * h * c3;
      shell_lap<T>(sym, f, fd, f2, x, y, z, chi + at, ld);
    } else {
      const float rf = (float)r, xf = (float)x, yf = (float)y, zf = (float)z;
      const float h = rf - iv * (float)V.tabs[2 * tab + 1];
      const float f = c0 + h * (c1 + h * (c2 + h * c3));
      const float fd = (c1 + h * (2.f * c2 + h * (3.f * c3))) / rf;
      const float f2 = 2.f * c2 + 6.f * h * c3;
      shell_lap<float>(sym, f, fd, f2, xf, yf, zf, (float*)(chi + at), ld);
    }
    for (int q = 0; q < nf; q++) roff[at + q] = (d.f0 + q) * n * (int)sizeof(T);
  }
  return take;
}

// acc[c] += sum_b C[f_b][lane] * chi[c*ld + b] for b < nfun (lane = MO j < n), with U
// coefficient loads in flight.  roff[b] is the byte offset of function f_b's coefficient row,
// so each load is a uniform base plus a 32-bit per-lane offset (one integer add).
template <typename T, int U>
__device__ inline void lcao_batch(const T* __restrict__ C, const int* roff, const T* chi, int ld, int nfun, T acc[5],
                                  int lane) {
  const char* Cb = reinterpret_cast<const char*>(C);
  const unsigned lb = (unsigned)lane * (unsigned)sizeof(T);
  int b = 0;
  for (; b + U <= nfun; b += U) {
    T cv[U];
#pragma unroll
    for (int q = 0; q < U; q++) cv[q] = *reinterpret_cast<const T*>(Cb + ((unsigned)roff[b + q] + lb));
    __builtin_amdgcn_sched_barrier(0);  // all U loads issue before the first use
#pragma unroll
    for (int q = 0; q < U; q++) {
      acc[0] += cv[q] * chi[b + q];
      acc[1] += cv[q] * chi[ld + b + q];
      acc[2] += cv[q] * chi[2 * ld + b + q];
      acc[3] += cv[q] * chi[3 * ld + b + q];
      acc[4] += cv[q] * chi[4 * ld + b + q];
    }
  }
  for (; b < nfun; b++) {
    const T cv = *reinterpret_cast<const T*>(Cb + ((unsigned)roff[b] + lb));
    acc[0] += cv * chi[b];
    acc[1] += cv * chi[ld + b];
    acc[2] += cv * chi[2 * ld + b];
    acc[3] += cv * chi[3 * ld + b];
    acc[4] += cv * chi[4 * ld + b];
  }
}

// Value, gradient and laplacian of the n MOs of spin s at p (lane j < n holds MO j):
// pass 1, then batches of pass 2 + contraction.  Barriers separate the LDS producers and
// consumers of the walker (WPB walkers per block, one per wave).  C: coefficients in global
// memory or staged in LDS.
template <typename T, typename TS, int U, int WPB = 1>
__device__ inline void mo_lap_point(const DevSys& S, const SysLds& V, const TS* spline, const T* C, int n,
                                    const double p[3], T* chi, int* roff, int cap, T phi[5], int lane) {
  for (int c = 0; c < 5; c++) phi[c] = (T)0;
  const int nact = basis_active(S, V, p, lane);
  walker_sync<WPB>();
  for (int i0 = 0; i0 < nact;) {
    int nfun;
    const int take = basis_batch_lap<T, TS>(S, V, spline, p, i0, nact, n, chi, cap, roff, cap, &nfun, lane);
    walker_sync<WPB>();
    if (lane < n) lcao_batch<T, U>(C, roff, chi, cap, nfun, phi, lane);
    walker_sync<WPB>();
    i0 += take;
  }
}

}  // namespace qw
