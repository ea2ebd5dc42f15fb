// Host side of the mixed-precision path (included by qw_engine.hip after qw_engine).
#pragma once
#include <cstdint>

namespace {

// Round-to-nearest-even BF16 of a float (returned as the raw 16-bit pattern).
uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
float bf16_val(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

int lcao_mode(int precision) {
  return precision == QW_PREC_MIXED_FP32 ? LCAO_FP32 : (precision == QW_PREC_MIXED_BF16X6 ? LCAO_BF16X6 : LCAO_BF16X3);
}

}  // namespace

void mixed_setup(qw_engine& e) {
  const SysTables& T = e.T;
  MixedTables& M = e.MT;
  M.ldchi = (T.nb + 31) / 32 * 32;
  M.nkstep = M.ldchi / 32;
  std::vector<float> spf(T.spline.begin(), T.spline.end());
  M.spline = e.upload(spf);
  for (int s = 0; s < 2; s++) {
    const int n = s ? T.ndn : T.nup;
    M.ntile[s] = (n + 15) / 16;
    std::vector<float> cf(T.C[s].begin(), T.C[s].end());
    M.Cf[s] = e.upload(cf);
    // B fragments of v_mfma_f32_16x16x32_bf16: lane l holds rows k = 8*(l/16)+i, column l%16
    const size_t nfrag = (size_t)M.nkstep * std::max(M.ntile[s], 1) * 64;
    std::vector<uint16_t> part[3];
    for (auto& v : part) v.assign(nfrag * 8, 0);
    for (int ks = 0; ks < M.nkstep; ks++)
      for (int t = 0; t < M.ntile[s]; t++)
        for (int l = 0; l < 64; l++) {
          const int j = t * 16 + (l & 15);
          for (int i = 0; i < 8; i++) {
            const int b = ks * 32 + (l >> 4) * 8 + i;
            double c = (b < T.nb && j < n) ? T.C[s][(size_t)b * n + j] : 0.0;
            const size_t at = (((size_t)ks * M.ntile[s] + t) * 64 + l) * 8 + i;
            for (int q = 0; q < 3; q++) {
              uint16_t bb = bf16_bits((float)c);
              part[q][at] = bb;
              c -= (double)bf16_val(bb);
            }
          }
        }
    for (int q = 0; q < 3; q++) {
      std::vector<uint4> packed(nfrag);
      std::memcpy(packed.data(), part[q].data(), nfrag * 16);
      M.Cfrag[s][q] = e.upload(packed);
    }
  }
  M.ready = 1;
  // the LDS-staged sweep variant is a tuning candidate where its LDS fits one workgroup
  int maxlds = 0;
  if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, e.device) != hipSuccess) maxlds = 0;
  const size_t need = sweep_mx_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, LCAO_FP32, M.ldchi, 2, e.S.nb,
                                         e.S.nsp, kWpbCL);
  e.cl_ok = e.precision == QW_PREC_MIXED_FP32 && !e.big && need <= std::min<size_t>(kLdsMax, (size_t)maxlds);
}

namespace {
// LDS: FP32 basis tile (5 x ldchi), MFMA accumulator exchange (5 x 64), active-function list (ldchi ints)
size_t lds_mx_rows(const qw_engine& e) {
  return (5 * (size_t)e.MT.ldchi + 5 * 64) * sizeof(float) + (size_t)e.MT.ldchi * sizeof(int);
}
size_t lds_mx_sweep(const qw_engine& e, int mode, int depth) {
  return sweep_mx_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, mode, e.MT.ldchi, depth);
}

// FP64 recompute of the FP32 inverses of spin `only_spin` (or both when -1) from the current
// FP32 MO value rows: log|det| and sign in FP64.
void mixed_lu(qw_engine& e, int only_spin) { e.run_inv<float, float>(e.W.mof, e.W.invf, only_spin); }

template <int MODE>
void mixed_rows_t(qw_engine& e) {
  e.launch("k_mo_rows", [&] {
    hipLaunchKernelGGL(k_mo_rows_mx<MODE>, dim3((unsigned)((size_t)e.nw * e.S.ne)), dim3(64), lds_mx_rows(e), e.stream,
                       e.S, e.MT, e.W, e.nw);
  });
}

// LDS of the variant with the coefficients staged in LDS (kWpbCL walkers per block).
size_t lds_mx_cl(const qw_engine& e, int depth) {
  return sweep_mx_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, LCAO_FP32, e.MT.ldchi, depth, e.S.nb, e.S.nsp,
                            kWpbCL);
}

template <int DEPTH, int MODE, int U, int WPB>
void mixed_sweep_u(qw_engine& e, const SweepArgs& A) {
  const size_t lds = WPB > 1 ? lds_mx_cl(e, DEPTH) : lds_mx_sweep(e, MODE, DEPTH);
  const unsigned grid = (unsigned)((e.nw + WPB - 1) / WPB);
  auto go = [&](auto kern) {
    e.tuned_launch([&] { hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WPB), lds, e.stream, e.S, e.MT, e.W, A); });
  };
  switch (e.ns) {
    case 4: go(k_sweep_mx<4, DEPTH, MODE, U, WPB>); break;
    case 8: go(k_sweep_mx<8, DEPTH, MODE, U, WPB>); break;
    case 16: go(k_sweep_mx<16, DEPTH, MODE, U, WPB>); break;
    case 32: go(k_sweep_mx<32, DEPTH, MODE, U, WPB>); break;
    default: go(k_sweep_mx<64, DEPTH, MODE, U, WPB>); break;
  }
}
template <int DEPTH, int MODE>
void mixed_sweep_t(qw_engine& e, const SweepArgs& A) {
  if constexpr (MODE != LCAO_FP32) {  // the MFMA modes read the dense coefficient fragments (no U)
    mixed_sweep_u<DEPTH, MODE, 8, 1>(e, A);
  } else {
    // launch choice: 100 * (coefficients staged in LDS) + U
    switch (e.pick_u()) {
      case 4: mixed_sweep_u<DEPTH, MODE, 4, 1>(e, A); break;
      case 16: mixed_sweep_u<DEPTH, MODE, 16, 1>(e, A); break;
      case 104: mixed_sweep_u<DEPTH, MODE, 4, kWpbCL>(e, A); break;
      case 108: mixed_sweep_u<DEPTH, MODE, 8, kWpbCL>(e, A); break;
      case 116: mixed_sweep_u<DEPTH, MODE, 16, kWpbCL>(e, A); break;
      default: mixed_sweep_u<DEPTH, MODE, 8, 1>(e, A); break;
    }
  }
}
}  // namespace

void mixed_refresh_all(qw_engine& e) {
  switch (lcao_mode(e.precision)) {
    case LCAO_FP32: mixed_rows_t<LCAO_FP32>(e); break;
    case LCAO_BF16X6: mixed_rows_t<LCAO_BF16X6>(e); break;
    default: mixed_rows_t<LCAO_BF16X3>(e); break;
  }
  mixed_lu(e, -1);
}

// Moves of electrons [k0, k1) of spin s, with an FP64 LU recompute of that spin's inverse
// after every `refresh` moves.
void mixed_sweep(qw_engine& e, const SweepArgs& A0, int depth) {
  const int mode = lcao_mode(e.precision);
  for (int k0 = A0.k0; k0 < A0.k1; k0 += e.refresh) {
    SweepArgs A = A0;
    A.k0 = k0;
    A.k1 = std::min(A0.k1, k0 + e.refresh);
    A.dets = A0.dets && A.k1 == A0.k1;
    if (depth == 1) {
      if (mode == LCAO_FP32) mixed_sweep_t<1, LCAO_FP32>(e, A);
      else if (mode == LCAO_BF16X6) mixed_sweep_t<1, LCAO_BF16X6>(e, A);
      else mixed_sweep_t<1, LCAO_BF16X3>(e, A);
    } else {
      if (mode == LCAO_FP32) mixed_sweep_t<2, LCAO_FP32>(e, A);
      else if (mode == LCAO_BF16X6) mixed_sweep_t<2, LCAO_BF16X6>(e, A);
      else mixed_sweep_t<2, LCAO_BF16X3>(e, A);
    }
    mixed_lu(e, A.s);
  }
}

void mixed_local(qw_engine& e, const LocalArgs& A) {
  e.launch("k_local", [&] {
    if (e.S.has_pp) hipLaunchKernelGGL(k_local<float>, dim3(e.nw), dim3(64), e.lds_local(), e.stream, e.S, e.W, A);
    else hipLaunchKernelGGL((k_local<float, false, false>), dim3(e.nw), dim3(64), e.lds_local_nopp(), e.stream, e.S, e.W, A);
  });
}
