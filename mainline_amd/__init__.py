"""MI355X walker engine for QWalk's VMC/DMC hot path (C++ host + HIP kernels behind a C ABI)."""
