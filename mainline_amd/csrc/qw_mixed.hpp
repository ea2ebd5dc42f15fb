// Mixed-precision path: FP32 inverse and MO cache, BF16 split-precision or FP32 LCAO,
// FP64 LU refresh, log-determinant and sums.
#pragma once
#include "qw_kernels.hpp"

struct qw_engine;

namespace qw {
struct MixedTables {
  int ready;
};
}  // namespace qw

void mixed_setup(qw_engine& e);
void mixed_refresh_all(qw_engine& e);
void mixed_sweep(qw_engine& e, const qw::SweepArgs& A, int depth);
void mixed_local(qw_engine& e, const qw::LocalArgs& A);
