// Several systems in shared launches: one DMC step of every engine of a batch is one sweep
// launch per spin channel (per FP64 recompute interval on the mixed path), one inverse
// recompute launch after each, one local-energy launch, and every feedback interval one
// branching and one copy launch, all covering every system, recorded once into a HIP graph
// and replayed.  A launch's blocks read their system from a table of slots (tables, walkers,
// random-stream key and step counter of one engine); each block runs the same device code as
// the engine's own launches, so a system's walkers, weights and block energies in a batch
// are bitwise those of the same engine run alone.  No reference counterpart: QWalk runs one
// system per process (src/program_utils/ooqmc.cpp:148-170, src/method/Dmc_method.cpp:328-567).
#include "qw_engine.hpp"

namespace qw {

// One system of a batched launch.
struct BatchSlot {
  DevSys S;
  MixedTables M;
  DevWalkers W;
  BigPtrs<float> Pf;
  BigPtrs<double> Pd;
  DmcCtl* ctl;
  unsigned long long* keys;
  int* vals;
  int* psrc;
  int* pdst;
  unsigned* stepp;
  unsigned long long seed;
  int n[2];     // electrons of each spin
  int ns;       // register tile of the inverse (4..64); 0 = block-per-walker path
  int u;        // LCAO coefficient loads in flight (per size bucket, tuned)
  int inv_rows; // 33..64 electrons of a spin: row-per-lane inversion instead of 8x8 blocks
  int cap;      // basis tile of the block-per-walker sweep
};

constexpr int kBatchWpb = 4;  // walkers (wavefronts) per 256-thread block for <= 64 per spin

// A field of the slot table (constant address space) as the generic reference the kernel
// bodies take.
template <typename T>
__device__ inline const T& gref(const QW_CONST T& r) {
  return *(const T*)&r;
}

__device__ inline int batch_blocks(const QW_CONST BatchSlot& B) {
  return B.ns ? (B.W.nw + kBatchWpb - 1) / kBatchWpb : B.W.nw;
}

// The slot of block `blk` among the slots with moves k0.. of spin s, and the slot's first
// block (*b0).  Uniform over the block; nslot is small.
__device__ inline int batch_find(const QW_CONST BatchSlot* sl, int nslot, int s, int k0, int blk, int* b0) {
  int acc = 0;
  for (int i = 0; i < nslot; i++) {
    if (sl[i].n[s] <= k0) continue;
    const int nb = batch_blocks(sl[i]);
    if (blk < acc + nb) {
      *b0 = acc;
      return i;
    }
    acc += nb;
  }
  return -1;
}

// Moves k0 .. min(n_s, k0 + R) of spin s of every slot with more than k0 electrons of that
// spin.  PREC: 0 FP64, 1 mixed (FP32 LCAO and inverse).
template <int PREC>
__global__ __launch_bounds__(256, 2) void k_sweep_batch(const BatchSlot* __restrict__ slots, int nslot, int s, int k0,
                                                        int R, double tau) {
  extern __shared__ double ldsb[];
  const QW_CONST BatchSlot* sl = cptr(slots);
  int b0 = 0;
  const int i = batch_find(sl, nslot, s, k0, (int)blockIdx.x, &b0);
  if (i < 0) return;
  const QW_CONST BatchSlot& B = sl[i];
  const int n = B.n[s];
  SweepArgs A{s, k0, min(n, k0 + R), tau, B.seed, 0u, 0u, 1, 0, B.stepp};
  A.dets = A.k1 == n;
  const int t = (int)threadIdx.x, blk = (int)blockIdx.x - b0;
  char* lds = (char*)ldsb;
  if (B.ns == 0) {
    if constexpr (PREC == 0)
      sweep_big_body<double, double, 1>(gref(B.S), gref(B.W), gref(B.Pd), A, B.cap, blk, t, lds);
    else
      sweep_big_body<float, float, 1>(gref(B.S), gref(B.W), gref(B.Pf), A, B.cap, blk, t, lds);
    return;
  }
  const int lane = t & 63, wid = t >> 6, w = blk * kBatchWpb + wid;
#define QW_MX(NS, U) \
  sweep_mx_body<NS, 1, LCAO_FP32, U, kBatchWpb, false>(gref(B.S), gref(B.M), gref(B.W), A, w, lane, wid, lds)
#define QW_FP(NS, U) sweep_fp64_body<NS, 1, U, kBatchWpb>(gref(B.S), gref(B.W), A, w, lane, wid, lds)
#define QW_CASES(X, U0, U1, U2)                 \
  switch (B.ns * 100 + B.u) {                   \
    case 400 + U0: X(4, U0); break;             \
    case 400 + U1: X(4, U1); break;             \
    case 400 + U2: X(4, U2); break;             \
    case 800 + U0: X(8, U0); break;             \
    case 800 + U1: X(8, U1); break;             \
    case 800 + U2: X(8, U2); break;             \
    case 1600 + U0: X(16, U0); break;           \
    case 1600 + U1: X(16, U1); break;           \
    case 1600 + U2: X(16, U2); break;           \
    case 3200 + U0: X(32, U0); break;           \
    case 3200 + U1: X(32, U1); break;           \
    case 3200 + U2: X(32, U2); break;           \
    case 6400 + U0: X(64, U0); break;           \
    case 6400 + U1: X(64, U1); break;           \
    default: X(64, U2); break;                  \
  }
  if constexpr (PREC == 0) {
    QW_CASES(QW_FP, 2, 4, 8)
  } else {
    QW_CASES(QW_MX, 4, 8, 16)
  }
#undef QW_CASES
#undef QW_FP
#undef QW_MX
}

// FP64 recompute of spin s's inverse for the slots with moves k0.. of spin s (after their
// FP32 updates): four walkers per block with a wavefront each up to 64 electrons of a spin,
// one walker per block (k_inv_grid's 16 x 16 grid) above.
union InvBatchLds {
  InvRowsLds<4> r4[kBatchWpb];
  InvRowsLds<8> r8[kBatchWpb];
  InvRowsLds<16> r16[kBatchWpb];
  InvRowsLds<32> r32[kBatchWpb];
  InvRowsLds<64> r64[kBatchWpb];
  InvBlkLds blk[kBatchWpb];
  InvGridLds<16> grid;
};
__global__ __launch_bounds__(256) void k_inv_batch(const BatchSlot* __restrict__ slots, int nslot, int s, int k0) {
  __shared__ InvBatchLds L;
  const QW_CONST BatchSlot* sl = cptr(slots);
  int b0 = 0;
  const int i = batch_find(sl, nslot, s, k0, (int)blockIdx.x, &b0);
  if (i < 0) return;
  const QW_CONST BatchSlot& B = sl[i];
  const int t = (int)threadIdx.x, blk = (int)blockIdx.x - b0;
  const float* mo = B.W.mof;
  float* inv = B.W.invf;
  if (B.ns == 0) {
    k_inv_grid_body<16>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, blk, s, t, L.grid);
    return;
  }
  const int lane = t & 63, wid = t >> 6, w = blk * kBatchWpb + wid;
  if (w >= B.W.nw) return;
  switch (B.ns) {
    case 4: k_inv_rows_body<4, kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.r4[wid]); break;
    case 8: k_inv_rows_body<8, kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.r8[wid]); break;
    case 16: k_inv_rows_body<16, kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.r16[wid]); break;
    case 32: k_inv_rows_body<32, kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.r32[wid]); break;
    default:
      if (B.inv_rows)
        k_inv_rows_body<64, kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.r64[wid]);
      else
        k_inv_blk_body<kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.blk[wid]);
      break;
  }
}

// Local energy and DMC weight of every walker of every slot (one 64-thread block each).
template <int PREC>
__global__ __launch_bounds__(64) void k_local_batch(const BatchSlot* __restrict__ slots, int nslot, int pp_mode) {
  extern __shared__ double ldsl[];
  const QW_CONST BatchSlot* sl = cptr(slots);
  int i = 0, b0 = 0;
  while (i + 1 < nslot && (int)blockIdx.x >= b0 + sl[i].W.nw) b0 += sl[i++].W.nw;
  const QW_CONST BatchSlot& B = sl[i];
  LocalArgs A{};
  A.mode = 2;
  A.use_dgl = 1;
  A.stepp = B.stepp;
  A.pp_mode = pp_mode;
  A.seed = B.seed;
  A.ctl = B.ctl;
  if constexpr (PREC == 0)
    local_body<double>(gref(B.S), gref(B.W), A, (int)blockIdx.x - b0, (int)threadIdx.x, ldsl);
  else
    local_body<float>(gref(B.S), gref(B.W), A, (int)blockIdx.x - b0, (int)threadIdx.x, ldsl);
}

// Trial energy, effective timestep and the pair comb of each slot (one block each).
__global__ __launch_bounds__(1024) void k_branch_batch(const BatchSlot* __restrict__ slots, int nslot, int np) {
  __shared__ BranchLds L;
  const QW_CONST BatchSlot& B = cptr(slots)[blockIdx.x];
  branch_body(gref(B.W), B.ctl, B.keys, B.vals, B.psrc, B.pdst, B.W.nw * B.S.ne * np, 1, B.seed, 0u, B.stepp,
              (int)threadIdx.x, L);
}

// The branching copies of every slot (one block per pair slot of each).
__global__ __launch_bounds__(256) void k_copy_batch(const BatchSlot* __restrict__ slots, int nslot) {
  const QW_CONST BatchSlot* sl = cptr(slots);
  int i = 0, b0 = 0;
  while (i + 1 < nslot && (int)blockIdx.x >= b0 + sl[i].W.nw / 2 + 1) b0 += sl[i++].W.nw / 2 + 1;
  const QW_CONST BatchSlot& B = sl[i];
  copy_body(gref(B.S), gref(B.W), B.ctl, B.psrc, B.pdst, (int)blockIdx.x - b0, (int)threadIdx.x);
}

// Every slot's device step counter advances by one.
__global__ void k_step_batch(const BatchSlot* __restrict__ slots, int nslot) {
  const int i = (int)threadIdx.x;
  if (i < nslot) *slots[i].stepp += 1;
}

}  // namespace qw

// ====================================================================== the batch
struct qw_batch {
  std::string err;
  std::vector<qw_engine*> eng;
  int device = 0, precision = 0, refresh = kRefreshDefault;
  hipStream_t stream = nullptr;
  std::vector<BatchSlot> slots;
  BatchSlot* d_slots = nullptr;
  hipGraphExec_t gx = nullptr;
  int g_np = 0;
  double g_tau = 0;
  int g_ppm = -1;
  bool graphs = true;
  // online tuning per size bucket (slots of one register tile): LCAO loads in flight, and for
  // 33..64 electrons of a spin the inversion layout; every choice gives the same results
  std::vector<int> buckets;            // register tiles present, in order
  int tune_bucket = 0, tune_cand = 0;  // position of the tuning walk
  bool tuned = false;
  std::vector<double> tune_best;
  std::vector<hipEvent_t> evpool;
  struct Timed {
    hipEvent_t a, b;
    int kind;  // 0 sweep, 1 inverse, 2 local, 3 branch + copy
  };
  std::vector<Timed> timed;
  double ktime[4] = {0, 0, 0, 0};
  int kcount[4] = {0, 0, 0, 0};
  size_t lds_sweep = 0, lds_local = 0;

  ~qw_batch() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (gx) (void)hipGraphExecDestroy(gx);
    for (auto& t : timed) evpool.push_back(t.a), evpool.push_back(t.b);
    for (auto e : evpool) (void)hipEventDestroy(e);
    if (d_slots) {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      (void)hipFree(d_slots);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
  std::vector<int> ucands() const { return precision == QW_PREC_FP64 ? std::vector<int>{2, 4, 8} : std::vector<int>{4, 8, 16}; }
  hipEvent_t ev() {
    if (evpool.empty()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = evpool.back();
    evpool.pop_back();
    return e;
  }
  bool capturing = false;
  template <typename F>
  void launch(int kind, F f) {
    if (capturing) {
      f();
      HIPCHK(hipGetLastError());
      return;
    }
    Timed t{ev(), ev(), kind};
    HIPCHK(hipEventRecord(t.a, stream));
    f();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(t.b, stream));
    timed.push_back(t);
  }
  void drain() {
    HIPCHK(hipStreamSynchronize(stream));
    for (auto& t : timed) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
      ktime[t.kind] += ms;
      kcount[t.kind] += 1;
      evpool.push_back(t.a);
      evpool.push_back(t.b);
    }
    timed.clear();
  }

  // slot table from the engines' current walkers (walker counts change with the engines)
  void build_slots() {
    slots.clear();
    lds_sweep = lds_local = 0;
    // largest systems first: their blocks, the longest-running, are dispatched first and the
    // small systems' blocks fill in behind them (longest-processing-time-first order)
    std::vector<size_t> order(eng.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return eng[a]->S.ne > eng[b]->S.ne; });
    for (size_t k : order) {
      qw_engine& e = *eng[k];
      BatchSlot B{};
      B.S = e.S;
      B.M = e.MT;
      B.W = e.W;
      B.Pf = BigPtrs<float>{e.W.invf, e.W.mof, {e.MT.Cf[0], e.MT.Cf[1]}, e.MT.spline};
      B.Pd = BigPtrs<double>{e.W.inv, e.W.mo, {e.S.C[0], e.S.C[1]}, e.S.spline};
      B.ctl = e.dctl;
      B.keys = e.keys;
      B.vals = e.vals;
      B.psrc = e.psrc;
      B.pdst = e.pdst;
      B.stepp = e.dstep;
      B.seed = e.seed;
      B.n[0] = e.T.nup;
      B.n[1] = e.T.ndn;
      B.ns = e.big ? 0 : e.ns;
      B.cap = precision == QW_PREC_FP64 ? kCapF64 : kCapF32;
      B.u = precision == QW_PREC_FP64 ? 4 : 8;
      B.inv_rows = e.inv_rows ? 1 : 0;
      if (B.ns) {
        const size_t per = precision == QW_PREC_FP64
                               ? ((sweep_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, kCapF64, 8, 1) + 15) & ~(size_t)15)
                               : sweep_mx_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, LCAO_FP32, e.MT.ldchi, 1);
        lds_sweep = std::max(lds_sweep, (size_t)kBatchWpb * per);
      } else {
        lds_sweep = std::max(lds_sweep, sweep_big_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.S.nsp, B.cap,
                                                            precision == QW_PREC_FP64 ? 8 : 4));
      }
      lds_local = std::max(lds_local, e.lds_local());
      slots.push_back(B);
    }
    // the tuned choices of each bucket
    for (auto& B : slots) {
      const int bi = (int)(std::find(buckets.begin(), buckets.end(), B.ns) - buckets.begin());
      if (bi < (int)buckets.size() && B.ns) B.u = bucket_u[bi];
    }
    if (!d_slots) {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      HIPCHK(hipMalloc(&d_slots, std::max<size_t>(eng.size(), 1) * sizeof(BatchSlot)));
    }
    HIPCHK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, stream));
  }
  std::vector<int> bucket_u;

  int nslot() const { return (int)slots.size(); }
  int max_n(int s) const {
    int m = 0;
    for (auto& B : slots) m = std::max(m, B.n[s]);
    return m;
  }
  unsigned sweep_grid(int s, int k0) const {
    unsigned g = 0;
    for (auto& B : slots)
      if (B.n[s] > k0) g += B.ns ? (unsigned)((B.W.nw + kBatchWpb - 1) / kBatchWpb) : (unsigned)B.W.nw;
    return g;
  }
  void sweep(int s, int k0, int R, double tau) {
    const unsigned g = sweep_grid(s, k0);
    if (!g) return;
    launch(0, [&] {
      if (precision == QW_PREC_FP64)
        hipLaunchKernelGGL(k_sweep_batch<0>, dim3(g), dim3(256), lds_sweep, stream, d_slots, nslot(), s, k0, R, tau);
      else
        hipLaunchKernelGGL(k_sweep_batch<1>, dim3(g), dim3(256), lds_sweep, stream, d_slots, nslot(), s, k0, R, tau);
    });
  }
  void recompute(int s, int k0) {
    const unsigned g = sweep_grid(s, k0);
    if (!g) return;
    launch(1, [&] { hipLaunchKernelGGL(k_inv_batch, dim3(g), dim3(256), 0, stream, d_slots, nslot(), s, k0); });
  }
  // one DMC feedback interval of every slot: the launches of qw_engine::dmc_interval
  void interval(int np, double tau, int pp_mode) {
    unsigned nwt = 0, ncp = 0;
    for (auto& B : slots) {
      nwt += (unsigned)B.W.nw;
      ncp += (unsigned)(B.W.nw / 2 + 1);
    }
    for (int p = 0; p < np; p++) {
      for (int s = 0; s < 2; s++) {
        const int nmax = max_n(s);
        if (precision == QW_PREC_FP64) {
          sweep(s, 0, nmax, tau);
        } else {
          for (int k0 = 0; k0 < nmax; k0 += refresh) {
            sweep(s, k0, refresh, tau);
            recompute(s, k0);
          }
        }
      }
      launch(2, [&] {
        if (precision == QW_PREC_FP64)
          hipLaunchKernelGGL(k_local_batch<0>, dim3(nwt), dim3(64), lds_local, stream, d_slots, nslot(), pp_mode);
        else
          hipLaunchKernelGGL(k_local_batch<1>, dim3(nwt), dim3(64), lds_local, stream, d_slots, nslot(), pp_mode);
        hipLaunchKernelGGL(k_step_batch, dim3(1), dim3(64), 0, stream, d_slots, nslot());
      });
    }
    launch(3, [&] {
      hipLaunchKernelGGL(k_branch_batch, dim3(nslot()), dim3(1024), 0, stream, d_slots, nslot(), np);
      hipLaunchKernelGGL(k_copy_batch, dim3(ncp), dim3(256), 0, stream, d_slots, nslot());
    });
  }
  void interval_graph(int np, double tau, int pp_mode) {
    if (!gx || g_np != np || g_tau != tau || g_ppm != pp_mode) {
      if (gx) {
        HIPCHK(hipGraphExecDestroy(gx));
        gx = nullptr;
      }
      hipGraph_t g = nullptr;
      std::lock_guard<std::mutex> lk(g_devmem_mu);
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      capturing = true;
      try {
        interval(np, tau, pp_mode);
      } catch (...) {
        capturing = false;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      capturing = false;
      HIPCHK(hipStreamEndCapture(stream, &g));
      HIPCHK(hipGraphInstantiate(&gx, g, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(g));
      g_np = np;
      g_tau = tau;
      g_ppm = pp_mode;
    }
    HIPCHK(hipGraphLaunch(gx, stream));
  }
  void drop_graph() {
    if (gx) (void)hipGraphExecDestroy(gx);
    gx = nullptr;
  }

  // Online tuning, one size bucket at a time: each candidate of the bucket runs one full DMC
  // step's first sweep launch (spin 0, moves 0..), timed with events; the fastest is kept.
  // Runs eagerly during the first intervals of the first run.
  int tune_pending() const { return !tuned; }
  void tune_step_choose() {
    if (tuned) return;
    const auto c = ucands();
    if (tune_bucket >= (int)buckets.size()) {
      tuned = true;
      return;
    }
    bucket_u[tune_bucket] = c[tune_cand];
    for (auto& B : slots)
      if (B.ns == buckets[tune_bucket]) B.u = c[tune_cand];
    HIPCHK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, stream));
  }
  // after an interval of np steps run eagerly while tuning: the sweep launches of spin 0 at
  // moves 0.. are the first sweep launch of each step
  void tune_step_record(int np) {
    if (tuned) return;
    drain_tuning(np);
  }
  std::vector<std::pair<int, double>> tune_hist;  // (candidate index, ms) of the bucket
  void drain_tuning(int np) {
    HIPCHK(hipStreamSynchronize(stream));
    // the first sweep launch of each step is every (2 * segments_0 + ...)th; find them by order
    const int seg0 = precision == QW_PREC_FP64 ? 1 : (max_n(0) + refresh - 1) / refresh;
    const int seg1 = precision == QW_PREC_FP64 ? 1 : (max_n(1) + refresh - 1) / refresh;
    const int per_step = seg0 + seg1;
    int k = 0;
    double best = 1e300;
    for (auto& t : timed)
      if (t.kind == 0) {
        if (k % per_step == 0) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
          best = std::min(best, (double)ms);
        }
        k++;
      }
    (void)np;
    drain();
    tune_hist.push_back({tune_cand, best});
    const auto c = ucands();
    if (++tune_cand >= (int)c.size()) {
      int bi = 0;
      for (size_t q = 1; q < tune_hist.size(); q++)
        if (tune_hist[q].second < tune_hist[bi].second) bi = (int)q;
      bucket_u[tune_bucket] = c[tune_hist[bi].first];
      for (auto& B : slots)
        if (B.ns == buckets[tune_bucket]) B.u = bucket_u[tune_bucket];
      HIPCHK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, stream));
      tune_hist.clear();
      tune_cand = 0;
      if (++tune_bucket >= (int)buckets.size()) tuned = true;
    }
  }
};

namespace {
template <typename F>
qw_status bguard(qw_batch* b, F f) {
  try {
    HIPCHK(hipSetDevice(b->device));
    return f();
  } catch (DevErr& x) {
    b->err = x.what();
    return QW_ERR_DEVICE;
  } catch (Unsupported& x) {
    b->err = x.what();
    return QW_ERR_UNSUPPORTED;
  } catch (std::bad_alloc&) {
    b->err = "host allocation failed";
    return QW_ERR_NOMEM;
  } catch (std::exception& x) {
    b->err = x.what();
    return QW_ERR_INPUT;
  }
}
thread_local std::string g_batch_error;
}  // namespace

extern "C" qw_status qw_batch_create(int n, qw_engine* const* engines, qw_batch** out) {
  if (!out || n <= 0 || !engines) return QW_ERR_ARG;
  *out = nullptr;
  if (n > 64) {  // k_step_batch advances the step counters with one 64-thread block
    g_batch_error = "at most 64 engines per batch";
    return QW_ERR_ARG;
  }
  std::unique_ptr<qw_batch> b(new qw_batch);
  for (int i = 0; i < n; i++) {
    qw_engine* e = engines[i];
    if (!e) return QW_ERR_ARG;
    for (int j = 0; j < i; j++)
      if (engines[j] == e) {
        g_batch_error = "an engine appears twice in the batch";
        return QW_ERR_ARG;
      }
    if (i == 0) {
      b->device = e->device;
      b->precision = e->precision;
      b->refresh = e->refresh;
    }
    if (e->device != b->device || e->precision != b->precision || e->refresh != b->refresh) {
      g_batch_error = "the engines of a batch share one device, precision and recompute interval";
      return QW_ERR_ARG;
    }
    if (b->precision != QW_PREC_FP64 && b->precision != QW_PREC_MIXED_FP32) {
      g_batch_error = "batches run in fp64 or mixed_fp32";
      return QW_ERR_UNSUPPORTED;
    }
    if (e->md) {
      g_batch_error = "several determinants are not batched";
      return QW_ERR_UNSUPPORTED;
    }
    if (e->big && std::max(e->T.nup, e->T.ndn) > kGridMaxSpin && b->precision != QW_PREC_FP64) {
      g_batch_error = "mixed-precision batches hold systems of up to 128 electrons of a spin";
      return QW_ERR_UNSUPPORTED;
    }
    b->eng.push_back(e);
  }
  qw_status st = bguard(b.get(), [&] {
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (qw_engine* e : b->eng) {
      const int key = e->big ? 0 : e->ns;
      if (key && std::find(b->buckets.begin(), b->buckets.end(), key) == b->buckets.end()) b->buckets.push_back(key);
    }
    std::sort(b->buckets.begin(), b->buckets.end());
    b->bucket_u.assign(b->buckets.size(), b->precision == QW_PREC_FP64 ? 4 : 8);
    if (const char* t = std::getenv("QW_TUNE")) b->tuned = t[0] == '0';
    if (const char* g = std::getenv("QW_GRAPHS")) b->graphs = g[0] != '0';
    return QW_OK;
  });
  if (st != QW_OK) {
    g_batch_error = b->err;
    return st;
  }
  *out = b.release();
  return QW_OK;
}

extern "C" void qw_batch_destroy(qw_batch* b) { delete b; }

extern "C" const char* qw_batch_error(const qw_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

extern "C" qw_status qw_batch_run(qw_batch* b, const qw_method_desc* m, double* eblock, double* stats) {
  if (!b || !m) return QW_ERR_ARG;
  if (m->method != 1 || m->timestep <= 0 || m->nblock < 0 || m->nstep < 1) {
    b->err = "qw_batch_run runs DMC with timestep > 0, nblock >= 0, nstep >= 1";
    return QW_ERR_ARG;
  }
  for (qw_engine* e : b->eng)
    if (e->nw <= 0) {
      b->err = "every engine of the batch needs walkers (qw_walkers_set)";
      return QW_ERR_ARG;
    }
  return bguard(b, [&] {
    const int ne_sys = (int)b->eng.size();
    std::vector<qw_engine::RunState> R(ne_sys);
    for (int i = 0; i < ne_sys; i++) {
      qw_engine& e = *b->eng[i];
      e.run_begin(*m, R[i]);
      if (e.tm_mode) throw Unsupported("T-moves are not batched");
      e.set_device_step();
      HIPCHK(hipStreamSynchronize(e.stream));
    }
    // walker buffers and counts may have changed since the last run: new slot table, and the
    // recorded interval (grid sizes and slot contents) is recorded again
    b->drop_graph();
    b->build_slots();
    for (int k = 0; k < 4; k++) b->ktime[k] = 0, b->kcount[k] = 0;
    const int pp_mode = m->pp_stochastic ? 1 : 0;
    const int fbi = std::max(1, m->feedback_interval);
    auto t0 = std::chrono::steady_clock::now();
    long long moves = 0;
    long long per_step = 0;
    for (qw_engine* e : b->eng) per_step += (long long)e->nw * e->S.ne;
    for (int blk = 0; blk < m->nblock; blk++) {
      for (int st = 0; st < m->nstep;) {
        const int np = std::min(fbi, m->nstep - st);
        if (!b->tuned) {
          b->tune_step_choose();
          if (!b->tuned) {
            b->interval(np, m->timestep, pp_mode);
            b->tune_step_record(np);
          } else if (b->graphs) {
            b->interval_graph(np, m->timestep, pp_mode);
          } else {
            b->interval(np, m->timestep, pp_mode);
          }
        } else if (b->graphs) {
          b->interval_graph(np, m->timestep, pp_mode);
        } else {
          b->interval(np, m->timestep, pp_mode);
        }
        for (int i = 0; i < ne_sys; i++) b->eng[i]->run_advance(R[i], np);
        moves += per_step * np;
        st += np;
      }
      HIPCHK(hipStreamSynchronize(b->stream));
      for (int i = 0; i < ne_sys; i++)
        b->eng[i]->run_block_end(*m, R[i], blk, eblock ? eblock + (size_t)i * m->nblock * QW_BLOCK_FIELDS : nullptr);
      for (qw_engine* e : b->eng) HIPCHK(hipStreamSynchronize(e->stream));
    }
    b->drain();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < ne_sys; i++) {
      double st4[4];
      b->eng[i]->run_finish(*m, R[i], st4);
      if (stats) stats[4 + i] = st4[0];
    }
    if (stats) {
      stats[0] = (double)moves;
      stats[1] = secs;
      stats[2] = secs > 0 ? moves / secs : 0.0;
      stats[3] = b->tuned ? 1.0 : 0.0;
    }
    return QW_OK;
  });
}

extern "C" qw_status qw_batch_kernel_times(const qw_batch* b, double* avg_ms4, int* count4, int* tuning) {
  if (!b) return QW_ERR_ARG;
  for (int k = 0; k < 4; k++) {
    if (avg_ms4) avg_ms4[k] = b->kcount[k] ? b->ktime[k] / b->kcount[k] : 0.0;
    if (count4) count4[k] = b->kcount[k];
  }
  if (tuning)
    for (size_t i = 0; i < b->buckets.size(); i++) {
      tuning[2 * i] = b->buckets[i];
      tuning[2 * i + 1] = b->bucket_u[i];
    }
  return QW_OK;
}
