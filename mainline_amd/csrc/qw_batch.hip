// Several systems in shared launches: one DMC step of every engine of a batch is one sweep
// launch per spin channel (per FP64 recompute interval on the mixed path), one inverse
// recompute launch after each, one local-energy launch, and every feedback interval one
// branching and one copy launch, all covering every system, recorded once into a HIP graph
// and replayed.  A launch's blocks read their system from a table of slots (tables, walkers,
// random-stream key and step counter of one engine); each block runs the same device code as
// the engine's own launches, so a system's walkers, weights and block energies in a batch
// are bitwise those of the same engine run alone.  No reference counterpart: QWalk runs one
// system per process (src/program_utils/ooqmc.cpp:148-170, src/method/Dmc_method.cpp:328-567).
#include <numeric>

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
  // the slot's method: DMC (depth 1, node-crossing moves rejected, weights and branching)
  // or VMC (depth 2 with delayed rejection, ndecorr sweeps per step)
  int depth, ndecorr, lmode, pp_mode;
  int branch_np;  // branching after this step: the feedback interval's steps (0 = none)
  double tau;
  // block-per-... layout: walkers (wavefronts) per 256-thread block, the LDS bytes of each,
  // rows per lane of the one-wavefront sweep above 64 electrons of a spin (2 or 4)
  int wpb, lds_w, nr;
  // the slot's FP64 recompute interval (mixed path): a multiple of the batch's launch interval
  // R, or at least n (once per sweep); the recompute launch after moves [k0, k0 + R) covers
  // the slot when k0 + R is a multiple of rf or ends the sweep
  int rf;
};

constexpr int kBatchWpb = 4;  // walkers (wavefronts) per 256-thread block for <= 64 per spin

// A field of the slot table (constant address space) as the generic reference the kernel
// bodies take.
template <typename T>
__device__ inline const T& gref(const QW_CONST T& r) {
  return *(const T*)&r;
}

// Blocks of a slot in a sweep launch (wpb walkers each), and in a recompute launch (a
// wavefront per walker up to 64 electrons of a spin, a whole block per walker above).
__device__ inline int batch_blocks(const QW_CONST BatchSlot& B) { return (B.W.nw + B.wpb - 1) / B.wpb; }
__device__ inline int batch_blocks_inv(const QW_CONST BatchSlot& B) {
  return B.ns ? (B.W.nw + kBatchWpb - 1) / kBatchWpb : B.W.nw;
}

// Whether slot B recomputes spin s after the moves [k0, k0 + R) (host and device).
__host__ __device__ inline bool batch_recomputes(int rf, int n, int k0, int R) {
  return (k0 + R) % rf == 0 || k0 + R >= n;
}

// The slot of block `blk` among the slots of method depth `depth` with a sweep `sw` and moves
// k0.. of spin s (INV: and a recompute after the R moves from k0), and the slot's first block
// (*b0).  Uniform over the block; nslot is small.
template <bool INV = false>
__device__ inline int batch_find(const QW_CONST BatchSlot* sl, int nslot, int s, int k0, int depth, int sw, int blk,
                                 int* b0, int R = 0) {
  int acc = 0;
  for (int i = 0; i < nslot; i++) {
    if (sl[i].n[s] <= k0 || sl[i].depth != depth || sw >= sl[i].ndecorr) continue;
    if (INV && !batch_recomputes(sl[i].rf, sl[i].n[s], k0, R)) continue;
    const int nb = INV ? batch_blocks_inv(sl[i]) : batch_blocks(sl[i]);
    if (blk < acc + nb) {
      *b0 = acc;
      return i;
    }
    acc += nb;
  }
  return -1;
}

// Moves k0 .. min(n_s, k0 + R) of spin s, in sweep `sw` of the step, of every slot of method
// depth DEPTH (1 DMC, 2 VMC) with more than k0 electrons of that spin.  PREC: 0 FP64, 1 mixed
// (FP32 LCAO and inverse).
template <int PREC, int DEPTH>
__global__ __launch_bounds__(256, 2) void k_sweep_batch(const BatchSlot* __restrict__ slots, int nslot, int s, int k0,
                                                        int R, int sw) {
  extern __shared__ double ldsb[];
  const QW_CONST BatchSlot* sl = cptr(slots);
  int b0 = 0;
  const int i = batch_find(sl, nslot, s, k0, DEPTH, sw, (int)blockIdx.x, &b0);
  if (i < 0) return;
  const QW_CONST BatchSlot& B = sl[i];
  const int n = B.n[s];
  SweepArgs A{s, k0, min(n, k0 + R), B.tau, B.seed, 0u, (unsigned)sw, DEPTH == 1 ? 1 : 0, 0, B.stepp};
  A.dets = A.k1 == n && sw == B.ndecorr - 1;
  const int t = (int)threadIdx.x, blk = (int)blockIdx.x - b0;
  char* lds = (char*)ldsb;
  if (B.ns == 0) {
    // more than 64 electrons of a spin: B.wpb walkers per block, a wavefront each
    const int lane = t & 63, wid = t >> 6;
    if (wid >= B.wpb) return;
    const int w = blk * B.wpb + wid;
    char* base = lds + (size_t)wid * B.lds_w;
    if constexpr (PREC == 0) {
      if (B.nr == 2)
        sweep_w_body<double, double, DEPTH, 2, kBatchWpb>(gref(B.S), gref(B.W), gref(B.Pd), A, B.cap, w, lane, base);
      else
        sweep_w_body<double, double, DEPTH, 4, kBatchWpb>(gref(B.S), gref(B.W), gref(B.Pd), A, B.cap, w, lane, base);
    } else {
      if (B.nr == 2)
        sweep_w_body<float, float, DEPTH, 2, kBatchWpb>(gref(B.S), gref(B.W), gref(B.Pf), A, B.cap, w, lane, base);
      else
        sweep_w_body<float, float, DEPTH, 4, kBatchWpb>(gref(B.S), gref(B.W), gref(B.Pf), A, B.cap, w, lane, base);
    }
    return;
  }
  const int lane = t & 63, wid = t >> 6, w = blk * kBatchWpb + wid;
#define QW_MX(NS, U) \
  sweep_mx_body<NS, DEPTH, LCAO_FP32, U, kBatchWpb, false>(gref(B.S), gref(B.M), gref(B.W), A, w, lane, wid, lds)
#define QW_FP(NS, U) sweep_fp64_body<NS, DEPTH, U, kBatchWpb>(gref(B.S), gref(B.W), A, w, lane, wid, lds)
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
  InvBlkLds blk[kBatchWpb];
  InvGridLds<16> grid;
};
__global__ __launch_bounds__(256, 2) void k_inv_batch(const BatchSlot* __restrict__ slots, int nslot, int s, int k0,
                                                   int depth, int sw, int R) {
  __shared__ InvBatchLds L;
  const QW_CONST BatchSlot* sl = cptr(slots);
  int b0 = 0;
  const int i = batch_find<true>(sl, nslot, s, k0, depth, sw, (int)blockIdx.x, &b0, R);
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
      // 33..64 electrons of a spin: the 8x8-block layout (bitwise equal to the row layout an
      // engine may be set to, so one body serves both and the kernel keeps two waves a SIMD)
      k_inv_blk_body<kBatchWpb>(gref(B.S), mo, inv, B.W.logdet, B.W.sgn, w, s, lane, L.blk[wid]);
      break;
  }
}

// Local energy and DMC weight of every walker of every slot (one 64-thread block each).
// PP = false when no slot's system has a pseudopotential (the leaner kernel).
template <int PREC, bool PP>
__global__ __launch_bounds__(64) void k_local_batch(const BatchSlot* __restrict__ slots, int nslot) {
  extern __shared__ double ldsl[];
  const QW_CONST BatchSlot* sl = cptr(slots);
  int i = 0, b0 = 0;
  while (i + 1 < nslot && (int)blockIdx.x >= b0 + sl[i].W.nw) b0 += sl[i++].W.nw;
  const QW_CONST BatchSlot& B = sl[i];
  LocalArgs A{};
  A.mode = B.lmode;
  A.use_dgl = 1;
  A.stepp = B.stepp;
  A.pp_mode = B.pp_mode;
  A.seed = B.seed;
  A.ctl = B.ctl;
  if constexpr (PREC == 0)
    local_body<double, false, PP>(gref(B.S), gref(B.W), A, (int)blockIdx.x - b0, (int)threadIdx.x, ldsl);
  else
    local_body<float, false, PP>(gref(B.S), gref(B.W), A, (int)blockIdx.x - b0, (int)threadIdx.x, ldsl);
}

// Trial energy, effective timestep and the pair comb of each branching slot (one block each):
// every slot over np steps, or with np = 0 the slots whose branch_np is set (over that many).
__global__ __launch_bounds__(1024) void k_branch_batch(const BatchSlot* __restrict__ slots, int nslot, int np) {
  __shared__ BranchLds L;
  const QW_CONST BatchSlot& B = cptr(slots)[blockIdx.x];
  const int m = np ? np : B.branch_np;
  if (m <= 0) return;
  branch_body(gref(B.W), B.ctl, B.keys, B.vals, B.psrc, B.pdst, B.W.nw * B.S.ne * m, 1, B.seed, 0u, B.stepp,
              (int)threadIdx.x, L);
}

// The branching copies of the slots that branched (one block per pair slot of each).
__global__ __launch_bounds__(256) void k_copy_batch(const BatchSlot* __restrict__ slots, int nslot, int np) {
  const QW_CONST BatchSlot* sl = cptr(slots);
  int i = 0, b0 = 0;
  while (i + 1 < nslot && (int)blockIdx.x >= b0 + sl[i].W.nw / 2 + 1) b0 += sl[i++].W.nw / 2 + 1;
  const QW_CONST BatchSlot& B = sl[i];
  if (!np && B.branch_np <= 0) return;
  copy_body(gref(B.S), gref(B.W), B.ctl, B.psrc, B.pdst, (int)blockIdx.x - b0, (int)threadIdx.x);
}

// Every slot's device step counter advances by one.
__global__ void k_step_batch(const BatchSlot* __restrict__ slots, int nslot) {
  const int i = (int)threadIdx.x;
  if (i < nslot) *slots[i].stepp += 1;
}

}  // namespace qw

// ====================================================================== the batch
// A member: one engine with the runs it still has to do (an optional VMC equilibration, then
// its method), advanced one step per batch step in step mode, or a whole DMC run at a time in
// qw_batch_run.
struct BatchMember {
  qw_engine* e = nullptr;
  int id = 0;
  std::vector<qw_method_desc> runs;
  int run = 0;            // the run being done
  bool begun = false;     // run_begin done for it
  qw_engine::RunState R;
  int b = 0, st = 0;      // block, step within the block
  int ip = 0, np = 0;     // DMC: steps done in the feedback interval, its length
  std::vector<double> eblock;  // block records of the current run
  double stats[4] = {0, 0, 0, 0};
  bool done = false;
  const qw_method_desc& m() const { return runs[run]; }
  bool dmc() const { return m().method == 1; }
};

struct qw_batch {
  std::string err;
  int device = 0, precision = 0, refresh = kRefreshDefault;
  hipStream_t stream = nullptr;
  std::vector<BatchMember> mem;  // in join order
  int next_id = 1;
  std::vector<BatchSlot> slots;  // largest systems first
  BatchSlot* d_slots = nullptr;
  int slot_cap = 0;
  hipGraphExec_t gx = nullptr;
  int g_np = 0;
  bool graphs = true;
  // online tuning per size bucket (slots of one register tile): LCAO loads in flight; every
  // choice gives the same results
  std::vector<int> buckets;  // register tiles seen, sorted
  std::vector<int> bucket_u;
  int tune_bucket = 0, tune_cand = 0;
  bool tuned = false;
  std::vector<std::pair<int, double>> tune_hist;
  std::vector<hipEvent_t> evpool;
  struct Timed {
    hipEvent_t a, b;
    int kind;  // 0 sweep, 1 inverse, 2 local, 3 branch + copy
  };
  std::vector<Timed> timed;
  double ktime[4] = {0, 0, 0, 0};
  int kcount[4] = {0, 0, 0, 0};
  size_t lds_sweep = 0, lds_local = 0;
  bool any_pp = false;  // some slot's system has a pseudopotential
  bool g_pp = false;    // any_pp of the recorded graph
  bool capturing = false;
  bool step_mode = false;  // made by qw_batch_open: members join and leave between steps

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
  std::vector<int> ucands() const {
    return precision == QW_PREC_FP64 ? std::vector<int>{2, 4, 8} : std::vector<int>{4, 8, 16};
  }
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

  // Whether an engine can be a member (same device and precision, and a recompute interval
  // that is a multiple of the batch's launch interval or a whole sweep).
  std::string admissible(const qw_engine* e) const {
    const bool rf_ok = precision == QW_PREC_FP64 || e->refresh % refresh == 0 ||
                       e->refresh >= std::max(e->T.nup, e->T.ndn);
    if (e->device != device || e->precision != precision || !rf_ok)
      return "the engines of a batch share one device and precision, and their recompute intervals are multiples "
             "of the batch's";
    if (precision != QW_PREC_FP64 && precision != QW_PREC_MIXED_FP32) return "batches run in fp64 or mixed_fp32";
    if (e->md) return "several determinants are not batched";
    if (e->big && std::max(e->T.nup, e->T.ndn) > kGridMaxSpin && precision != QW_PREC_FP64)
      return "mixed-precision batches hold systems of up to 128 electrons of a spin";
    for (const auto& m : mem)
      if (m.e == e) return "an engine appears twice in the batch";
    return std::string();
  }
  void add_bucket(const qw_engine* e) {
    const int key = e->big ? 0 : e->ns;
    if (!key || std::find(buckets.begin(), buckets.end(), key) != buckets.end()) return;
    buckets.push_back(key);
    std::sort(buckets.begin(), buckets.end());
    bucket_u.assign(buckets.size(), precision == QW_PREC_FP64 ? 4 : 8);
    tuned = false;  // a new size bucket is tuned from the start
    tune_bucket = tune_cand = 0;
    tune_hist.clear();
  }
  int bucket_of(int ns) const {
    return (int)(std::find(buckets.begin(), buckets.end(), ns) - buckets.begin());
  }

  // Slot table of the members (largest systems first: their blocks, the longest-running, are
  // dispatched first and the small systems' blocks fill in behind them).  branch: DMC
  // members branch after this step (step mode).
  void build_slots(bool step_mode) {
    slots.clear();
    lds_sweep = lds_local = 0;
    any_pp = false;
    std::vector<size_t> order(mem.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return mem[a].e->S.ne > mem[b].e->S.ne; });
    for (size_t k : order) {
      const BatchMember& mb = mem[k];
      if (mb.done) continue;
      qw_engine& e = *mb.e;
      const qw_method_desc& m = mb.m();
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
      const int bi = bucket_of(B.ns);
      B.u = (B.ns && bi < (int)buckets.size()) ? bucket_u[bi] : (precision == QW_PREC_FP64 ? 4 : 8);
      B.inv_rows = e.inv_rows ? 1 : 0;
      B.depth = m.method == 1 ? 1 : 2;
      B.ndecorr = m.method == 1 ? 1 : std::max(1, m.ndecorr);
      B.lmode = m.method == 1 ? 2 : 1;
      B.pp_mode = m.pp_stochastic ? 1 : 0;
      B.tau = m.timestep;
      B.branch_np = (step_mode && m.method == 1 && mb.ip + 1 == mb.np) ? mb.np : 0;
      B.rf = std::max(e.refresh, refresh);
      if (B.ns) {
        const size_t per =
            precision == QW_PREC_FP64
                ? ((sweep_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, kCapF64, 8, B.depth) + 15) & ~(size_t)15)
                : sweep_mx_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.ns, LCAO_FP32, e.MT.ldchi, B.depth);
        B.wpb = kBatchWpb;
        lds_sweep = std::max(lds_sweep, (size_t)kBatchWpb * per);
      } else {
        const size_t per = sweep_w_lds_bytes(e.S.ne, e.S.na, e.S.nsh, e.S.ntab, e.S.nsp, B.cap,
                                             precision == QW_PREC_FP64 ? 8 : 4);
        // walkers per block while their LDS leaves room for two blocks per compute unit
        B.wpb = (int)std::max<size_t>(1, std::min<size_t>(kBatchWpb, (80u << 10) / per));
        B.lds_w = (int)per;
        B.nr = std::max(e.T.nup, e.T.ndn) <= 2 * kWave ? 2 : 4;
        lds_sweep = std::max(lds_sweep, (size_t)B.wpb * per);
      }
      any_pp |= e.S.has_pp != 0;
      lds_local = std::max(lds_local, e.S.has_pp ? e.lds_local() : e.lds_local_nopp());
      slots.push_back(B);
    }
    if ((int)slots.size() > slot_cap) {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      if (d_slots) HIPCHK(hipFree(d_slots));
      d_slots = nullptr;
      slot_cap = std::max<int>(64, (int)slots.size());
      HIPCHK(hipMalloc(&d_slots, (size_t)slot_cap * sizeof(BatchSlot)));
    }
    if (!slots.empty())
      HIPCHK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, stream));
  }

  int nslot() const { return (int)slots.size(); }
  int max_n(int s, int depth, int sw) const {
    int m = 0;
    for (auto& B : slots)
      if (B.depth == depth && sw < B.ndecorr) m = std::max(m, B.n[s]);
    return m;
  }
  unsigned sweep_grid(int s, int k0, int depth, int sw) const {
    unsigned g = 0;
    for (auto& B : slots)
      if (B.n[s] > k0 && B.depth == depth && sw < B.ndecorr) g += (unsigned)((B.W.nw + B.wpb - 1) / B.wpb);
    return g;
  }
  template <int DEPTH>
  void sweep(int s, int k0, int R, int sw) {
    const unsigned g = sweep_grid(s, k0, DEPTH, sw);
    if (!g) return;
    launch(0, [&] {
      if (precision == QW_PREC_FP64)
        hipLaunchKernelGGL((k_sweep_batch<0, DEPTH>), dim3(g), dim3(256), lds_sweep, stream, d_slots, nslot(), s, k0,
                           R, sw);
      else
        hipLaunchKernelGGL((k_sweep_batch<1, DEPTH>), dim3(g), dim3(256), lds_sweep, stream, d_slots, nslot(), s, k0,
                           R, sw);
    });
  }
  void recompute(int s, int k0, int depth, int sw) {
    unsigned g = 0;
    for (auto& B : slots)
      if (B.n[s] > k0 && B.depth == depth && sw < B.ndecorr && batch_recomputes(B.rf, B.n[s], k0, refresh))
        g += B.ns ? (unsigned)((B.W.nw + kBatchWpb - 1) / kBatchWpb) : (unsigned)B.W.nw;
    if (!g) return;
    launch(1, [&] {
      hipLaunchKernelGGL(k_inv_batch, dim3(g), dim3(256), 0, stream, d_slots, nslot(), s, k0, depth, sw, refresh);
    });
  }
  // the moves of one step of every slot of method depth DEPTH: its sweeps of both spin
  // channels, on the mixed path in recompute intervals each followed by the recompute
  template <int DEPTH>
  void moves() {
    int nsw = 0;
    for (auto& B : slots)
      if (B.depth == DEPTH) nsw = std::max(nsw, B.ndecorr);
    for (int sw = 0; sw < nsw; sw++)
      for (int s = 0; s < 2; s++) {
        const int nmax = max_n(s, DEPTH, sw);
        if (nmax == 0) continue;
        if (precision == QW_PREC_FP64) {
          sweep<DEPTH>(s, 0, nmax, sw);
        } else {
          for (int k0 = 0; k0 < nmax; k0 += refresh) {
            sweep<DEPTH>(s, k0, refresh, sw);
            recompute(s, k0, DEPTH, sw);
          }
        }
      }
  }
  void local_and_step() {
    unsigned nwt = 0;
    for (auto& B : slots) nwt += (unsigned)B.W.nw;
    launch(2, [&] {
      if (precision == QW_PREC_FP64) {
        if (any_pp) hipLaunchKernelGGL((k_local_batch<0, true>), dim3(nwt), dim3(64), lds_local, stream, d_slots, nslot());
        else hipLaunchKernelGGL((k_local_batch<0, false>), dim3(nwt), dim3(64), lds_local, stream, d_slots, nslot());
      } else {
        if (any_pp) hipLaunchKernelGGL((k_local_batch<1, true>), dim3(nwt), dim3(64), lds_local, stream, d_slots, nslot());
        else hipLaunchKernelGGL((k_local_batch<1, false>), dim3(nwt), dim3(64), lds_local, stream, d_slots, nslot());
      }
      hipLaunchKernelGGL(k_step_batch, dim3(1), dim3(64), 0, stream, d_slots, nslot());
    });
  }
  // branching: every slot over np steps, or (np = 0) the slots whose branch_np is set
  void branch(int np) {
    unsigned ncp = 0;
    for (auto& B : slots) ncp += (unsigned)(B.W.nw / 2 + 1);
    launch(3, [&] {
      hipLaunchKernelGGL(k_branch_batch, dim3(nslot()), dim3(1024), 0, stream, d_slots, nslot(), np);
      hipLaunchKernelGGL(k_copy_batch, dim3(ncp), dim3(256), 0, stream, d_slots, nslot(), np);
    });
  }
  // one DMC feedback interval of every slot (qw_batch_run): the launches of
  // qw_engine::dmc_interval
  void interval(int np) {
    for (int p = 0; p < np; p++) {
      moves<1>();
      local_and_step();
    }
    branch(np);
  }
  void interval_graph(int np) {
    if (!gx || g_np != np || g_pp != any_pp) {
      if (gx) {
        HIPCHK(hipGraphExecDestroy(gx));
        gx = nullptr;
      }
      hipGraph_t g = nullptr;
      std::lock_guard<std::mutex> lk(g_devmem_mu);
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      capturing = true;
      try {
        interval(np);
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
      g_pp = any_pp;
    }
    HIPCHK(hipGraphLaunch(gx, stream));
  }
  void drop_graph() {
    if (gx) (void)hipGraphExecDestroy(gx);
    gx = nullptr;
  }

  // Online tuning, one size bucket at a time: a candidate is set for the bucket, one step (or
  // interval) runs with every launch timed, and the sum of its sweep launches is the
  // candidate's time; after every candidate of a bucket the fastest is kept.
  void tune_choose() {
    if (tuned) return;
    if (tune_bucket >= (int)buckets.size()) {
      tuned = true;
      return;
    }
    bucket_u[tune_bucket] = ucands()[tune_cand];
    for (auto& B : slots)
      if (B.ns == buckets[tune_bucket]) B.u = bucket_u[tune_bucket];
    if (!slots.empty())
      HIPCHK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, stream));
  }
  // nsteps: steps the timed launches cover (the candidate's time is per step)
  void tune_record(int nsteps) {
    if (tuned) return;
    HIPCHK(hipStreamSynchronize(stream));
    double t = 0;
    for (auto& x : timed)
      if (x.kind == 0) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, x.a, x.b));
        t += ms;
      }
    drain();
    tune_hist.push_back({tune_cand, t / std::max(1, nsteps)});
    const auto c = ucands();
    if (++tune_cand >= (int)c.size()) {
      int bi = 0;
      for (size_t q = 1; q < tune_hist.size(); q++)
        if (tune_hist[q].second < tune_hist[bi].second) bi = (int)q;
      bucket_u[tune_bucket] = c[tune_hist[bi].first];
      for (auto& B : slots)
        if (B.ns == buckets[tune_bucket]) B.u = bucket_u[tune_bucket];
      if (!slots.empty())
        HIPCHK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, stream));
      tune_hist.clear();
      tune_cand = 0;
      if (++tune_bucket >= (int)buckets.size()) tuned = true;
    }
  }

  // ---- step mode: every member advances one step (DMC: one step of its feedback interval;
  // VMC: one step of ndecorr sweeps), in shared launches per method
  void step_once(std::vector<int>& finished) {
    // engine-side work of the step's start, on each engine's stream: the run's start, the
    // all-electron recompute at a VMC block's start, the device step counter
    for (auto& mb : mem) {
      if (mb.done) continue;
      qw_engine& e = *mb.e;
      if (!mb.begun) {
        e.run_begin(mb.m(), mb.R);
        if (e.tm_mode) throw Unsupported("T-moves are not batched");
        mb.begun = true;
        mb.b = mb.st = mb.ip = 0;
        mb.eblock.assign((size_t)std::max(mb.m().nblock, 1) * QW_BLOCK_FIELDS, 0.0);
      }
      if (!mb.dmc() && mb.st == 0) e.refresh_all();
      if (mb.dmc() && mb.ip == 0) mb.np = std::min(mb.R.fbi, mb.m().nstep - mb.st);
      if (!mb.dmc() || mb.ip == 0) e.set_device_step();
    }
    for (auto& mb : mem)
      if (!mb.done) HIPCHK(hipStreamSynchronize(mb.e->stream));
    build_slots(true);
    if (slots.empty()) return;
    tune_choose();
    bool any_dmc = false, any_vmc = false, any_branch = false;
    for (auto& B : slots) {
      (B.depth == 1 ? any_dmc : any_vmc) = true;
      any_branch |= B.branch_np > 0;
    }
    if (any_dmc) moves<1>();
    if (any_vmc) moves<2>();
    local_and_step();
    if (any_branch) branch(0);
    tune_record(1);
    HIPCHK(hipStreamSynchronize(stream));
    drain();
    // bookkeeping of the members, and their block ends on their own streams
    for (auto& mb : mem) {
      if (mb.done) continue;
      qw_engine& e = *mb.e;
      const qw_method_desc& m = mb.m();
      mb.st++;
      if (mb.dmc()) {
        if (++mb.ip == mb.np) {
          e.run_advance(mb.R, mb.np);
          mb.ip = 0;
        }
      } else {
        mb.R.moves += (long long)e.nw * e.S.ne * (std::max(1, m.ndecorr) - 1);
        e.run_advance(mb.R, 1);
      }
      if (mb.st == m.nstep) {
        e.run_block_end(m, mb.R, mb.b, mb.eblock.data());
        mb.b++;
        mb.st = mb.ip = 0;
        if (mb.b == m.nblock) {
          e.run_finish(m, mb.R, mb.stats);
          mb.begun = false;
          if (++mb.run == (int)mb.runs.size()) {
            mb.done = true;
            finished.push_back(mb.id);
          }
        }
      }
    }
    for (auto& mb : mem) HIPCHK(hipStreamSynchronize(mb.e->stream));
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

qw_status batch_new(int device, int precision, int refresh, qw_batch** out) {
  std::unique_ptr<qw_batch> b(new qw_batch);
  b->device = device;
  b->precision = precision;
  b->refresh = refresh > 0 ? refresh : kRefreshDefault;
  qw_status st = bguard(b.get(), [&] {
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
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
}  // namespace

extern "C" qw_status qw_batch_create(int n, qw_engine* const* engines, qw_batch** out) {
  if (!out || n <= 0 || !engines) return QW_ERR_ARG;
  *out = nullptr;
  if (n > 64) {  // k_step_batch advances the step counters with one 64-thread block
    g_batch_error = "at most 64 engines per batch";
    return QW_ERR_ARG;
  }
  for (int i = 0; i < n; i++)
    if (!engines[i]) return QW_ERR_ARG;
  // launch interval: the largest common divisor of the members' recompute intervals (a
  // member that recomputes once per sweep does not constrain it)
  int R = 0, nmax = 1;
  for (int i = 0; i < n; i++) {
    const int ns = std::max(engines[i]->T.nup, engines[i]->T.ndn);
    nmax = std::max(nmax, ns);
    if (engines[i]->refresh < ns) R = std::gcd(R, engines[i]->refresh);
  }
  if (R == 0) R = nmax;
  qw_batch* b = nullptr;
  qw_status st = batch_new(engines[0]->device, engines[0]->precision, R, &b);
  if (st != QW_OK) return st;
  for (int i = 0; i < n; i++) {
    const std::string why = b->admissible(engines[i]);
    if (!why.empty()) {
      g_batch_error = why;
      const bool prec_ok = b->precision == QW_PREC_FP64 || b->precision == QW_PREC_MIXED_FP32;
      delete b;
      return prec_ok ? QW_ERR_ARG : QW_ERR_UNSUPPORTED;
    }
    BatchMember mb;
    mb.e = engines[i];
    mb.id = b->next_id++;
    b->mem.push_back(mb);
    b->add_bucket(engines[i]);
  }
  *out = b;
  return QW_OK;
}

extern "C" qw_status qw_batch_open(int device, int precision, int refresh_interval, qw_batch** out) {
  if (!out) return QW_ERR_ARG;
  *out = nullptr;
  if (precision != QW_PREC_FP64 && precision != QW_PREC_MIXED_FP32) {
    g_batch_error = "batches run in fp64 or mixed_fp32";
    return QW_ERR_UNSUPPORTED;
  }
  qw_status st = batch_new(device, precision, refresh_interval, out);
  if (st == QW_OK) (*out)->step_mode = true;
  return st;
}

extern "C" qw_status qw_batch_join(qw_batch* b, qw_engine* e, int nruns, const qw_method_desc* runs, int* member) {
  if (!b || !e || nruns <= 0 || !runs || !member) return QW_ERR_ARG;
  if (!b->step_mode) {
    b->err = "engines join a batch made by qw_batch_open";
    return QW_ERR_ARG;
  }
  const std::string why = b->admissible(e);
  if (!why.empty()) {
    b->err = why;
    return e->device == b->device && e->precision == b->precision ? QW_ERR_UNSUPPORTED : QW_ERR_ARG;
  }
  if (e->nw <= 0) {
    b->err = "an engine joins a batch with its walkers set";
    return QW_ERR_ARG;
  }
  int live = 0;
  for (auto& m : b->mem) live += !m.done;
  if (live >= 64) {
    b->err = "at most 64 engines per batch";
    return QW_ERR_ARG;
  }
  for (int r = 0; r < nruns; r++) {
    if (runs[r].timestep <= 0 || runs[r].nblock < 1 || runs[r].nstep < 1) {
      b->err = "every run needs timestep > 0, nblock >= 1, nstep >= 1";
      return QW_ERR_ARG;
    }
    if (runs[r].method == 1 && runs[r].tmoves && e->S.has_pp) {
      b->err = "T-moves are not batched";
      return QW_ERR_UNSUPPORTED;
    }
  }
  BatchMember mb;
  mb.e = e;
  mb.id = b->next_id++;
  mb.runs.assign(runs, runs + nruns);
  b->mem.push_back(mb);
  b->add_bucket(e);
  *member = mb.id;
  return QW_OK;
}

extern "C" qw_status qw_batch_step(qw_batch* b, int* finished, int* nfinished) {
  if (!b || !b->step_mode) return QW_ERR_ARG;
  return bguard(b, [&] {
    std::vector<int> fin;
    b->step_once(fin);
    if (nfinished) *nfinished = (int)fin.size();
    if (finished)
      for (size_t i = 0; i < fin.size(); i++) finished[i] = fin[i];
    return QW_OK;
  });
}

extern "C" qw_status qw_batch_leave(qw_batch* b, int member, double* eblock, double* stats) {
  if (!b) return QW_ERR_ARG;
  for (size_t i = 0; i < b->mem.size(); i++) {
    BatchMember& mb = b->mem[i];
    if (mb.id != member) continue;
    if (!mb.done) {
      b->err = "a member leaves the batch when its runs are done";
      return QW_ERR_ARG;
    }
    if (eblock) std::copy(mb.eblock.begin(), mb.eblock.end(), eblock);
    if (stats) std::copy(mb.stats, mb.stats + 4, stats);
    b->mem.erase(b->mem.begin() + i);
    return QW_OK;
  }
  b->err = "no such member";
  return QW_ERR_ARG;
}

extern "C" void qw_batch_destroy(qw_batch* b) { delete b; }

extern "C" const char* qw_batch_error(const qw_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

extern "C" qw_status qw_batch_run(qw_batch* b, const qw_method_desc* m, double* eblock, double* stats) {
  if (!b || !m) return QW_ERR_ARG;
  if (m->method != 1 || m->timestep <= 0 || m->nblock < 0 || m->nstep < 1) {
    b->err = "qw_batch_run runs DMC with timestep > 0, nblock >= 0, nstep >= 1";
    return QW_ERR_ARG;
  }
  if (b->step_mode) {
    b->err = "qw_batch_run runs a batch made by qw_batch_create";
    return QW_ERR_ARG;
  }
  for (auto& mb : b->mem)
    if (mb.e->nw <= 0) {
      b->err = "every engine of the batch needs walkers (qw_walkers_set)";
      return QW_ERR_ARG;
    }
  return bguard(b, [&] {
    const int nsys = (int)b->mem.size();
    for (auto& mb : b->mem) {
      mb.runs.assign(1, *m);
      mb.run = 0;
      mb.done = false;
      qw_engine& e = *mb.e;
      e.run_begin(*m, mb.R);
      if (e.tm_mode) throw Unsupported("T-moves are not batched");
      e.set_device_step();
      HIPCHK(hipStreamSynchronize(e.stream));
    }
    // walker buffers and counts may have changed since the last run: new slot table, and the
    // recorded interval (grid sizes and slot contents) is recorded again
    b->drop_graph();
    b->build_slots(false);
    for (int k = 0; k < 4; k++) b->ktime[k] = 0, b->kcount[k] = 0;
    const int fbi = std::max(1, m->feedback_interval);
    auto t0 = std::chrono::steady_clock::now();
    long long moves = 0, per_step = 0;
    for (auto& mb : b->mem) per_step += (long long)mb.e->nw * mb.e->S.ne;
    for (int blk = 0; blk < m->nblock; blk++) {
      for (int st = 0; st < m->nstep;) {
        const int np = std::min(fbi, m->nstep - st);
        if (!b->tuned) {
          b->tune_choose();
          b->interval(np);
          b->tune_record(np);
        } else if (b->graphs) {
          b->interval_graph(np);
        } else {
          b->interval(np);
        }
        for (auto& mb : b->mem) mb.e->run_advance(mb.R, np);
        moves += per_step * np;
        st += np;
      }
      HIPCHK(hipStreamSynchronize(b->stream));
      for (int i = 0; i < nsys; i++)
        b->mem[i].e->run_block_end(*m, b->mem[i].R, blk,
                                   eblock ? eblock + (size_t)i * m->nblock * QW_BLOCK_FIELDS : nullptr);
      for (auto& mb : b->mem) HIPCHK(hipStreamSynchronize(mb.e->stream));
    }
    b->drain();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < nsys; i++) {
      double st4[4];
      b->mem[i].e->run_finish(*m, b->mem[i].R, st4);
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
    for (size_t i = 0; i < b->buckets.size() && i < 5; i++) {
      tuning[2 * i] = b->buckets[i];
      tuning[2 * i + 1] = b->bucket_u[i];
    }
  return QW_OK;
}
