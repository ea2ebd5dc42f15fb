// Host side of the mixed-precision path (included by qw_engine.hip after qw_engine).
#pragma once

void mixed_setup(qw_engine& e) { throw std::runtime_error("mixed precision is not available in this build"); }
void mixed_refresh_all(qw_engine& e) { throw std::runtime_error("mixed precision is not available in this build"); }
void mixed_sweep(qw_engine& e, const qw::SweepArgs& A, int depth) {
  throw std::runtime_error("mixed precision is not available in this build");
}
void mixed_local(qw_engine& e, const qw::LocalArgs& A) {
  throw std::runtime_error("mixed precision is not available in this build");
}
