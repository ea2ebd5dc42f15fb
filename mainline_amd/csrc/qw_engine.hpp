// Engine state shared by the library's translation units (qw_engine.hip: one engine and
// the C ABI; qw_batch.hip: launches covering several engines).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>


#include "../../include/qwalk_amd.h"
#include "qw_big.hpp"
#include "qw_kernels.hpp"
#include "qw_mixed.hpp"
#include "qw_multidet.hpp"
#include "qw_parm.hpp"
#include "qw_setup.hpp"

using namespace qw;

// Device allocation and release, and graph recording, are serialised across the engines of a
// process: an allocation or free (which may synchronise the device) issued by one host
// thread while another records a graph would invalidate that recording.
extern std::mutex g_devmem_mu;

struct DevErr : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// Valid input outside what this engine supports (QW_ERR_UNSUPPORTED).
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};
#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) throw DevErr(std::string(#x) + ": " + hipGetErrorString(e_));            \
  } while (0)

constexpr int kChunkNF = 640;  // basis functions per LDS tile
constexpr int kRefreshDefault = 32;
// Default FP64 recompute interval of the mixed path for n electrons of a spin: every 32 moves
// up to 64; above, half a sweep rounded up to a multiple of 32.  The recompute costs O(n^3) a
// spin and an update O(n^2) a move, so a fixed 32 would make the recompute the larger part
// (eight recomputes per sweep at 256 a spin).
inline int default_refresh(int n) { return n <= 64 ? kRefreshDefault : (n / 2 + 31) / 32 * 32; }

inline int pick_ns(int n) {
  int ns = 4;
  while (ns < n) ns <<= 1;
  return ns;
}

// Device forms of the Jastrow radial objects (constants folded, see JRad).
template <typename R>
std::vector<JRad<R>> jrad_table(const std::vector<JObj>& objs) {
  std::vector<JRad<R>> out;
  for (const JObj& o : objs) {
    JRad<R> d{};
    d.kind = o.kind;
    d.nf = o.nf;
    d.rcut = (R)o.rcut;
    d.rcinv = (R)o.rcutinv;
    d.rcinv2 = (R)(o.rcutinv * o.rcutinv);
    d.gamma = (R)o.gamma;
    d.pade0 = (R)o.pade0;
    for (int i = 0; i < kMaxJF; i++) {
      d.beta[i] = (R)o.beta[i];
      if (o.kind == 0) {
        d.cl[i] = (R)(o.cl[i] * o.cusp);
        d.cu[i] = (R)(o.cu[i] * o.cusp);
      } else {
        d.cl[i] = (R)o.cl[i];
        d.cu[i] = (R)o.cu[i];
      }
      d.cbl[i] = (R)(o.cl[i] * (o.beta[i] + 1));
      d.cbu[i] = (R)(o.cu[i] * (o.beta[i] + 1));
      d.nz[i] = (o.cl[i] != 0.0 || o.cu[i] != 0.0) ? 1 : 0;
    }
    out.push_back(d);
  }
  if (out.empty()) out.push_back(JRad<R>{});
  return out;
}

struct qw_engine {
  std::string err;
  // mixed precision: FP64 recompute of a spin's inverse after this many moves (FP32 rank-1
  // updates in between); by default every 32 moves
  int device = 0, precision = 0, refresh = kRefreshDefault;
  unsigned long long seed = 1;
  SysTables T;
  DevSys S{};
  DevWalkers W{};
  int cap = 0, nw = 0, ns = 4;
  hipStream_t stream = nullptr;
  unsigned step = 0;
  std::vector<void*> sys_allocs, walker_allocs;
  DmcCtl* dctl = nullptr;
  unsigned long long* keys = nullptr;
  int *vals = nullptr, *psrc = nullptr, *pdst = nullptr;
  double* dred = nullptr;
  int sort_cap = 0;
  bool mixed_ready = false;
  bool dmc_ready = false;  // DMC control state valid for the resident walkers
  DmcCtl ctl_host{};
  MixedTables MT{};
  // per-kernel timing with events on the engine stream
  bool timing = true;
  bool capturing = false;  // launches are being recorded into a graph (no timing events)
  int graph_mode = -1;     // QW_GRAPHS: 0 never, 1 always, otherwise for small step work
  unsigned* dstep = nullptr;  // device step counter read by the kernels
  hipGraphExec_t gx = nullptr;
  int g_np = 0, g_nw = 0, g_tm = 0, g_ppm = 0;
  double g_tau = 0;
  struct Ev {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Ev> evs;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> knames;
  std::vector<double> ktime;
  std::vector<int> kcount;

  // more than 64 electrons of one spin: one wavefront per walker, inverse in device memory
  bool big = false;
  bool md = false;  // several determinants (FP64, k_inv_md / k_sweep_md / k_local<double, true>)
  double* luw = nullptr;  // k_inv_panel working matrices (more than 128 of a spin)
  int lucap = 0;
  // T-moves: per-(atom, electron) negative nonlocal sums of the last local energy
  double* tneg = nullptr;
  int tnegcap = 0;
  int tm_mode = 0;  // 0 off, 1 TMOVES, 2 TMOVESSC (for the current run)
  void* scratch = nullptr;  // this device's copy of the allocation-free evaluation's array

  ~qw_engine() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (gx) (void)hipGraphExecDestroy(gx);
    {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      for (void* p : sys_allocs) (void)hipFree(p);
      for (void* p : walker_allocs) (void)hipFree(p);
    }
    for (auto& e : evs) pool.push_back(e.a), pool.push_back(e.b);
    for (auto ev : pool) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  T* upload(const std::vector<T>& v) {
    void* p = nullptr;
    size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      HIPCHK(hipMalloc(&p, n));
    }
    sys_allocs.push_back(p);
    if (!v.empty()) {
      HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    return (T*)p;
  }
  void drop_graph() {
    if (gx) (void)hipGraphExecDestroy(gx);
    gx = nullptr;
    g_np = g_nw = g_tm = g_ppm = 0;
    g_tau = 0;
  }
  template <typename T>
  T* walloc(size_t n) {
    // never inside a recorded interval: the allocation and its clearing memset would be
    // replayed against memory this call does not own
    if (capturing) throw DevErr("walker allocation while recording a graph");
    void* p = nullptr;
    {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    }
    walker_allocs.push_back(p);
    HIPCHK(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream));
    return (T*)p;
  }

  int kind_of(const char* name) {
    for (size_t i = 0; i < knames.size(); i++)
      if (knames[i] == name) return (int)i;
    knames.push_back(name);
    ktime.push_back(0);
    kcount.push_back(0);
    return (int)knames.size() - 1;
  }
  hipEvent_t ev() {
    if (pool.empty()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // Launch wrapper: records start/stop events around the launch on the engine stream.
  bool quiet = false;  // launches issue no timing events (the allocation-free entry point)
  template <typename F>
  void launch(const char* name, F f) {
    if (!timing || capturing || quiet) {
      f();
      HIPCHK(hipGetLastError());
      return;
    }
    Ev e{kind_of(name), ev(), ev()};
    HIPCHK(hipEventRecord(e.a, stream));
    f();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e.b, stream));
    evs.push_back(e);
    if (evs.size() > 4096) drain_events();
  }
  void drain_events() {
    HIPCHK(hipStreamSynchronize(stream));
    for (auto& e : evs) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
      ktime[e.kind] += ms;
      kcount[e.kind] += 1;
      pool.push_back(e.a);
      pool.push_back(e.b);
    }
    evs.clear();
  }
  void reset_times() {
    drain_events();
    std::fill(ktime.begin(), ktime.end(), 0.0);
    std::fill(kcount.begin(), kcount.end(), 0);
  }

  void build_device_system() {
    S.nup = T.nup;
    S.ndn = T.ndn;
    S.ne = T.ne();
    S.na = T.na;
    S.nb = T.nb;
    S.nsh = (int)T.shells.size();
    S.ntab = T.ntab;
    S.nint = T.nint;
    S.nee = (int)T.ee.size();
    S.has_pp = T.pptypes.empty() ? 0 : 1;
    // MO cache row: the spin's MO list (its electrons' orbitals with one determinant)
    S.nsp = std::max(T.nmo[0], T.nmo[1]);
    S.nsig = std::max(T.nup, T.ndn);
    S.ndet = T.ndet;
    S.nmo[0] = T.nmo[0];
    S.nmo[1] = T.nmo[1];
    S.vii = T.vii;
    // shell chunks whose functions fit one LDS tile
    std::vector<int> ch{0};
    int acc = 0, maxnf = 0;
    for (size_t i = 0; i < T.shells.size(); i++) {
      if (acc + T.shells[i].nf > kChunkNF) {
        ch.push_back((int)i);
        maxnf = std::max(maxnf, acc);
        acc = 0;
      }
      acc += T.shells[i].nf;
    }
    maxnf = std::max(maxnf, acc);
    ch.push_back((int)T.shells.size());
    S.nchunk = (int)ch.size() - 1;
    S.chunk_nf = maxnf;
    S.atom_pos = upload(T.atom_pos);
    S.atom_Z = upload(T.atom_Z);
    S.shells = upload(T.shells);
    S.chunk_sh0 = upload(ch);
    S.spline = upload(T.spline);
    S.tab_inv = upload(T.tab_invspacing);
    S.tab_sp = upload(T.tab_spacing);
    for (int s = 0; s < 2; s++) {
      S.C[s] = upload(T.C[s]);
      std::vector<float> cf(T.C[s].begin(), T.C[s].end());
      S.Cf[s] = upload(cf);
    }
    {
      std::vector<double> lc(T.ndet);
      for (int d = 0; d < T.ndet; d++) lc[d] = std::log(std::fabs(T.detwt[d]));
      std::vector<int> oc((size_t)2 * T.ndet * S.nsig, 0);
      for (int s = 0; s < 2; s++) {
        const int n = s ? T.ndn : T.nup;
        for (int d = 0; d < T.ndet; d++)
          for (int i = 0; i < n; i++) oc[((size_t)s * T.ndet + d) * S.nsig + i] = T.occ[s][(size_t)d * n + i];
      }
      S.detwt = upload(T.detwt);
      S.logc = upload(lc);
      S.occ = upload(oc);
    }
    // linear Jastrow parameters: one unit-coefficient object each, atom lists for one-body ones
    {
      std::vector<JParmDev> pd;
      std::vector<JObj> units;
      std::vector<int> patoms;
      for (const JParm& jp : T.jparms) {
        JParmDev d{jp.body, jp.spin, (int)patoms.size(), 0, jp.value};
        if (jp.body == 1)
          for (int a = 0; a < T.na; a++)
            if (T.atom_label[a] == jp.label) patoms.push_back(a);
        d.a1 = (int)patoms.size();
        pd.push_back(d);
        units.push_back(jp.unit);
        parm_values.push_back(jp.value);
      }
      npar = (int)pd.size();
      d_parm = upload(pd);
      d_punit = upload(jrad_table<double>(units));
      d_patoms = upload(patoms);
      // OPTIMIZE_DET: the coefficients of CSFs 1.. come first (Slat_Jastrow order)
      if (T.optimize_det && T.ndet > 1 && T.csf_c.size() > 1) {
        nslp = (int)T.csf_c.size() - 1;
        parm_values.insert(parm_values.begin(), T.csf_c.begin() + 1, T.csf_c.end());
        d_csf_d0 = upload(T.csf_d0);
        d_csf_w = upload(T.csf_w);
      }
    }
    S.ee = upload(jrad_table<double>(T.ee));
    S.eef = upload(jrad_table<float>(T.ee));
    {
      // electron-ion Jastrow objects grouped by atom type: atoms whose object lists are equal
      std::vector<std::vector<int>> members;
      std::vector<int> rep;
      auto same = [&](int a, int b) {
        const int na0 = T.atom_ei0[a], nb0 = T.atom_ei0[b];
        const int n1 = T.atom_ei0[a + 1] - na0, n2 = T.atom_ei0[b + 1] - nb0;
        return n1 == n2 && (n1 == 0 || std::memcmp(&T.ei[na0], &T.ei[nb0], sizeof(JObj) * n1) == 0);
      };
      for (int a = 0; a < T.na; a++) {
        size_t t = 0;
        while (t < rep.size() && !same(rep[t], a)) t++;
        if (t == rep.size()) {
          rep.push_back(a);
          members.emplace_back();
        }
        members[t].push_back(a);
      }
      std::vector<JObj> eit;
      std::vector<int> a0{0}, o0{0};
      std::vector<double> epos, eZ;
      for (size_t t = 0; t < rep.size(); t++) {
        for (int o = T.atom_ei0[rep[t]]; o < T.atom_ei0[rep[t] + 1]; o++) eit.push_back(T.ei[o]);
        o0.push_back((int)eit.size());
        for (int a : members[t]) {
          for (int d = 0; d < 3; d++) epos.push_back(T.atom_pos[3 * a + d]);
          eZ.push_back(T.atom_Z[a]);
        }
        a0.push_back(a0.back() + (int)members[t].size());
      }
      if (eit.empty()) eit.push_back(JObj{});
      S.n_eit = (int)rep.size();
      S.eit = upload(jrad_table<double>(eit));
      S.eitf = upload(jrad_table<float>(eit));
      S.eit_a0 = upload(a0);
      S.eit_o0 = upload(o0);
      S.eit_pos = upload(epos);
      S.eit_Z = upload(eZ);
    }
    {
      if (T.na > 1023 || T.ntab >= (1 << 13)) throw std::runtime_error("too many atoms or spline tables");
      std::vector<ShellLds> shl(T.shells.size());
      for (size_t i = 0; i < T.shells.size(); i++) {
        const ShellDesc& d = T.shells[i];
        shl[i].rc = std::min(d.rcut, d.objthr);
        shl[i].meta = d.atom | (d.sym << 10) | (d.nf << 14) | (d.table << 18);
        shl[i].f0 = d.f0;
      }
      S.shl = upload(shl);
      std::vector<double> tabs;
      for (int t = 0; t < T.ntab; t++) {
        tabs.push_back(T.tab_invspacing[t]);
        tabs.push_back(T.tab_spacing[t]);
      }
      if (tabs.empty()) tabs.assign(2, 0.0);
      S.tabs = upload(tabs);
    }
    S.pp = upload(T.pptypes);
    S.atom_pp = upload(T.atom_pp);
    dctl = upload(std::vector<DmcCtl>(1));
    dstep = upload(std::vector<unsigned>(1));
    dred = upload(std::vector<double>(kBsum + 1));
  }

  // device bytes of ensure_walkers(n) (keep in step with it)
  unsigned long long walker_bytes(int n) const {
    const unsigned long long ne = S.ne, nsp = S.nsp, N = n;
    const unsigned long long es = precision == QW_PREC_FP64 ? 8 : 4;
    unsigned long long b = N * ne * 3 * 8 + N * 2 * nsp * nsp * es + N * ne * 5 * nsp * es;
    b += N * (2 * 8 + 2 * 8 + 8 + 8 + 4 * 8 + 4 + kBsum * 8) + N * ne * 4 * 8;
    unsigned long long P = 1;
    while (P < N) P <<= 1;
    b += P * (8 + 4) + 2 * (N / 2 + 1) * 4;
    if (big) {
      const unsigned long long nm = std::max(T.nup, T.ndn);
      if (nm > kGridMaxSpin) b += N * nm * nm * 8;  // k_inv_panel working matrices
    }
    if (md) b += N * 2 * S.ndet * (S.nsig * S.nsig + 2) * 8;
    return b;
  }

  // Walker arrays of K configurations carved from `base` (nullptr: only count), each 256-byte
  // aligned, in the layout ensure_walkers allocates; returns the bytes used.  The branching
  // buffers and T-move sums are not carved (fixed-configuration evaluation uses neither).
  size_t carve_walkers(char* base, int K, DevWalkers& V) const {
    size_t at = 0;
    auto take = [&](size_t bytes) -> void* {
      void* p = base ? base + at : nullptr;
      at += (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
      return p;
    };
    const size_t ne = S.ne, nsp = S.nsp, n = (size_t)K;
    V = DevWalkers{};
    V.nw = K;
    V.pos = (double*)take(n * ne * 3 * 8);
    if (precision == QW_PREC_FP64) {
      V.inv = (double*)take(n * 2 * nsp * nsp * 8);
      V.mo = (double*)take(n * ne * 5 * nsp * 8);
    } else {
      V.invf = (float*)take(n * 2 * nsp * nsp * 4);
      V.mof = (float*)take(n * ne * 5 * nsp * 4);
    }
    V.logdet = (double*)take(n * 2 * 8);
    V.sgn = (double*)take(n * 2 * 8);
    V.weight = (double*)take(n * 8);
    V.eold = (double*)take(n * 8);
    V.elast = (double*)take(n * 4 * 8);
    V.acc = (int*)take(n * 4);
    V.bsum = (double*)take(n * kBsum * 8);
    V.dgl = (double*)take(n * ne * 4 * 8);
    if (md) {
      V.invd = (double*)take(n * 2 * S.ndet * S.nsig * S.nsig * 8);
      V.ldd = (double*)take(n * 2 * S.ndet * 8);
      V.sgd = (double*)take(n * 2 * S.ndet * 8);
    }
    return at;
  }

  void ensure_walkers(int n) {
    if (n <= cap) {
      W.nw = n;
      nw = n;
      return;
    }
    HIPCHK(hipStreamSynchronize(stream));
    // a recorded DMC interval holds the old walker pointers: it is recorded again on next use
    drop_graph();
    {
      std::lock_guard<std::mutex> g(g_devmem_mu);
      for (void* p : walker_allocs) (void)hipFree(p);
    }
    walker_allocs.clear();
    luw = nullptr;  // the inversion and T-move buffers are walker allocations too
    lucap = 0;
    tneg = nullptr;
    tnegcap = 0;
    const size_t ne = S.ne, nsp = S.nsp;
    W.pos = walloc<double>((size_t)n * ne * 3);
    if (precision == QW_PREC_FP64) {
      W.inv = walloc<double>((size_t)n * 2 * nsp * nsp);
      W.mo = walloc<double>((size_t)n * ne * 5 * nsp);
      W.invf = nullptr;
      W.mof = nullptr;
    } else {
      // FP32 caches only: the FP64 inverse recompute reads the FP32 MO values directly
      W.mo = nullptr;
      W.inv = nullptr;
      W.invf = walloc<float>((size_t)n * 2 * nsp * nsp);
      W.mof = walloc<float>((size_t)n * ne * 5 * nsp);
    }
    W.logdet = walloc<double>((size_t)n * 2);
    W.sgn = walloc<double>((size_t)n * 2);
    W.weight = walloc<double>(n);
    W.eold = walloc<double>(n);
    W.elast = walloc<double>((size_t)n * 4);
    W.acc = walloc<int>(n);
    W.bsum = walloc<double>((size_t)n * kBsum);
    W.dgl = walloc<double>((size_t)n * ne * 4);
    if (md) {
      W.invd = walloc<double>((size_t)n * 2 * S.ndet * S.nsig * S.nsig);
      W.ldd = walloc<double>((size_t)n * 2 * S.ndet);
      W.sgd = walloc<double>((size_t)n * 2 * S.ndet);
    } else {
      W.invd = W.ldd = W.sgd = nullptr;
    }
    int P = 1;
    while (P < n) P <<= 1;
    keys = walloc<unsigned long long>(P);
    vals = walloc<int>(P);
    psrc = walloc<int>(n / 2 + 1);
    pdst = walloc<int>(n / 2 + 1);
    cap = n;
    W.nw = n;
    nw = n;
  }

  // ---- kernels
  // LDS: tiles of basis values (5 x chunk_nf doubles) and active-function indices (chunk_nf ints)
  size_t lds_rows() const { return 5 * (size_t)S.chunk_nf * sizeof(double) + (size_t)S.chunk_nf * sizeof(int); }
  size_t lds_sweep(int depth) const {
    return sweep_lds_bytes(S.ne, S.na, S.nsh, S.ntab, ns, kCapF64, sizeof(double), depth);
  }
  size_t lds_local() const {
    return (7 * (size_t)S.ne + 5 * (size_t)S.chunk_nf) * sizeof(double) + (size_t)S.chunk_nf * sizeof(int);
  }
  // without a pseudopotential: positions, determinant gradients and laplacians only
  size_t lds_local_nopp() const { return 7 * (size_t)S.ne * sizeof(double); }

  void run_mo_rows() {
    launch("k_mo_rows", [&] {
      hipLaunchKernelGGL(k_mo_rows, dim3((unsigned)((size_t)nw * S.ne)), dim3(64), lds_rows(), stream, S, W, nw);
    });
  }
  // inverse, log|det| and sign of spin `only_spin` (both spins when -1) from the MO value rows
  template <typename TIN, typename TI>
  void run_inv(const TIN* mo, TI* out, int only_spin) {
    auto go = [&](auto kern) {
      launch("k_inv", [&] {
        hipLaunchKernelGGL(kern, dim3(only_spin < 0 ? 2 * nw : nw), dim3(64), 0, stream, S, mo, out, W.logdet, W.sgn,
                           nw, only_spin);
      });
    };
    switch (ns) {
      case 4: go(k_inv<4, TIN, TI>); break;
      case 8: go(k_inv<8, TIN, TI>); break;
      case 16: go(k_inv<16, TIN, TI>); break;
      case 32: go(k_inv<32, TIN, TI>); break;
      default:
        if (inv_rows) go(k_inv<64, TIN, TI>);
        else go(k_inv_blk<TIN, TI>);
        break;
    }
  }
  bool inv_rows = false;  // 33..64 per spin: row-per-lane inversion instead of 8x8 blocks (QW_INV=rows)
  // linear Jastrow parameters (getParmDeriv)
  int npar = 0;   // Jastrow parameters
  int nslp = 0;   // CSF coefficients (OPTIMIZE_DET), before the Jastrow parameters
  std::vector<double> parm_values;
  int* d_csf_d0 = nullptr;
  double* d_csf_w = nullptr;
  JParmDev* d_parm = nullptr;
  JRad<double>* d_punit = nullptr;
  int* d_patoms = nullptr;
  void refresh_all() {
    if (big) {
      big_refresh_all();
      return;
    }
    if (md) {
      run_mo_rows();
      launch("k_inv", [&] {
        hipLaunchKernelGGL(k_inv_md, dim3((unsigned)((size_t)nw * 2 * S.ndet)), dim3(64), inv_md_lds_bytes(S.nsig),
                           stream, S, W, nw);
      });
      return;
    }
    if (precision == QW_PREC_FP64) {
      run_mo_rows();
      run_inv<double, double>(W.mo, W.inv, -1);
    } else {
      mixed_refresh_all(*this);
    }
  }
  void run_sweep(int s, int k0, int k1, double tau, unsigned sweep, int depth, int restrict_nodes, int dets) {
    SweepArgs A{s, k0, k1, tau, seed, step, sweep, restrict_nodes, dets, dstep};
    if (big) {
      big_sweep(A, depth);
      return;
    }
    if (md) {
      const size_t lds = sweep_md_lds_bytes(S.ne, S.ndet, S.nsig, S.nsp, S.chunk_nf, depth);
      launch("k_sweep", [&] {
        if (depth == 1) hipLaunchKernelGGL(k_sweep_md<1>, dim3(nw), dim3(64), lds, stream, S, W, A);
        else hipLaunchKernelGGL(k_sweep_md<2>, dim3(nw), dim3(64), lds, stream, S, W, A);
      });
      return;
    }
    if (precision != QW_PREC_FP64) {
      mixed_sweep(*this, A, depth);
      return;
    }
    const int u = pick_u();
    if (depth == 1) {
      if (u == 2) sweep_fp64<1, 2>(A);
      else if (u == 8) sweep_fp64<1, 8>(A);
      else sweep_fp64<1, 4>(A);
    } else {
      if (u == 2) sweep_fp64<2, 2>(A);
      else if (u == 8) sweep_fp64<2, 8>(A);
      else sweep_fp64<2, 4>(A);
    }
  }
  template <int DEPTH, int U>
  void sweep_fp64(const SweepArgs& A) {
    auto go = [&](auto kern) {
      tuned_launch([&] { hipLaunchKernelGGL(kern, dim3(nw), dim3(64), lds_sweep(DEPTH), stream, S, W, A); });
    };
    switch (ns) {
      case 4: go(k_sweep<4, DEPTH, U>); break;
      case 8: go(k_sweep<8, DEPTH, U>); break;
      case 16: go(k_sweep<16, DEPTH, U>); break;
      case 32: go(k_sweep<32, DEPTH, U>); break;
      default: go(k_sweep<64, DEPTH, U>); break;
    }
  }

  // ---- online tuning of the sweep launch: LCAO coefficient loads in flight (U) and, on the
  // FP32 mixed path, whether a block stages the coefficients in LDS for several walkers
  // (choice 100 + U).  The candidates do the same arithmetic in the same order, so the first
  // sweep launches of an engine simply run with each candidate in turn (kTuneReps times),
  // timed with events; the fastest is kept for the engine's lifetime.  QW_TUNE=0 keeps the
  // default; QW_SWEEP=<choice> forces one.
  static constexpr int kTuneReps = 2;
  bool tune_enabled = true;
  int tune_u = 0;      // chosen launch, 0 while measuring
  int tune_n = 0;      // candidate launches issued
  int tune_cur = 0;    // candidate of the launch being issued
  bool cl_ok = false;  // LDS-staged coefficients fit this device and system
  std::vector<int> tune_cands;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> tune_ev;
  int u_default() const { return precision == QW_PREC_FP64 ? 4 : 8; }
  const std::vector<int>& cands() {
    if (tune_cands.empty()) {
      if (precision == QW_PREC_FP64) tune_cands = {2, 4, 8};
      else tune_cands = {4, 8, 16};
      if (cl_ok) tune_cands.insert(tune_cands.end(), {104, 108, 116});
    }
    return tune_cands;
  }
  int pick_u() {
    if (tune_u) return tune_u;
    if (!tune_enabled || capturing || big || (precision != QW_PREC_FP64 && precision != QW_PREC_MIXED_FP32))
      return u_default();
    tune_cur = cands()[tune_n % cands().size()];
    return tune_cur;
  }
  // Launches a sweep; while the launch choice is being tuned the launch is bracketed by events.
  template <typename F>
  void tuned_launch(F f) {
    const bool measuring = !tune_u && tune_enabled && !capturing && tune_cur != 0;
    if (!measuring) {
      launch("k_sweep", f);
      return;
    }
    hipEvent_t a = ev(), b = ev();
    HIPCHK(hipEventRecord(a, stream));
    launch("k_sweep", f);
    HIPCHK(hipEventRecord(b, stream));
    tune_ev.push_back({tune_cur, {a, b}});
    tune_cur = 0;
    const size_t nc = cands().size();
    if (++tune_n < (int)(nc * kTuneReps)) return;
    HIPCHK(hipEventSynchronize(b));
    double best = 1e300;
    for (size_t i = 0; i < nc; i++) {
      const int u = tune_cands[i];
      double t = 1e300;
      for (auto& x : tune_ev)
        if (x.first == u) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, x.second.first, x.second.second));
          t = std::min(t, (double)ms);
        }
      if (t < best) {
        best = t;
        tune_u = u;
      }
    }
    for (auto& x : tune_ev) pool.push_back(x.second.first), pool.push_back(x.second.second);
    tune_ev.clear();
  }
  void ensure_tneg() {
    if (tneg && tnegcap >= nw) return;
    drop_graph();
    tneg = walloc<double>((size_t)nw * S.na * S.ne);
    tnegcap = nw;
  }
  // Casula T-moves after the local energy of a DMC step (k_local wrote A.tneg)
  void run_tmove(const LocalArgs& A) {
    TmoveArgs t{tm_mode == 2 ? 1 : 0, A.pp_mode, A.tneg, A.seed, dstep, dctl};
    const size_t lds = md ? tmove_lds_bytes(S.ne, S.chunk_nf, S.ndet, S.nsig, S.nsp) : tmove_lds_bytes(S.ne, S.chunk_nf);
    launch("k_tmove", [&] {
      if (md)
        hipLaunchKernelGGL((k_tmove<double, true>), dim3(nw), dim3(64), lds, stream, S, W, t);
      else if (precision == QW_PREC_FP64)
        hipLaunchKernelGGL(k_tmove<double>, dim3(nw), dim3(64), lds, stream, S, W, t);
      else
        hipLaunchKernelGGL(k_tmove<float>, dim3(nw), dim3(64), lds, stream, S, W, t);
    });
  }
  void run_local(const LocalArgs& A) {
    if (md) {
      const size_t lds = lds_local() + ((size_t)5 * S.ndet + S.nsp + 2) * sizeof(double);
      launch("k_local", [&] { hipLaunchKernelGGL((k_local<double, true>), dim3(nw), dim3(64), lds, stream, S, W, A); });
      return;
    }
    if (precision != QW_PREC_FP64) {
      mixed_local(*this, A);
      return;
    }
    launch("k_local", [&] {
      if (S.has_pp) hipLaunchKernelGGL(k_local<double>, dim3(nw), dim3(64), lds_local(), stream, S, W, A);
      else hipLaunchKernelGGL((k_local<double, false, false>), dim3(nw), dim3(64), lds_local_nopp(), stream, S, W, A);
    });
  }
  // ---- more than 64 electrons per spin
  // working matrices of k_inv_panel (more than 128 electrons of a spin)
  void lu_buffers() {
    const int n = std::max(T.nup, T.ndn);
    if (lucap >= nw && luw) return;
    luw = walloc<double>((size_t)nw * n * n);
    lucap = nw;
  }
  // FP64 inverse, log|det| and sign of spin s from the MO value rows (FP64 or FP32 cache)
  void lu_refresh(int s) {
    const int n = s ? T.ndn : T.nup;
    if (n == 0) return;
    if (std::max(T.nup, T.ndn) <= kGridMaxSpin) {
      // up to 128 electrons of a spin: the register-grid Gauss-Jordan (256 threads per matrix)
      launch("k_inv", [&] {
        if (precision == QW_PREC_FP64)
          hipLaunchKernelGGL((k_inv_grid<16, double, double>), dim3(nw), dim3(256), 0, stream, S,
                             (const double*)W.mo, W.inv, W.logdet, W.sgn, nw, s);
        else
          hipLaunchKernelGGL((k_inv_grid<16, float, float>), dim3(nw), dim3(256), 0, stream, S,
                             (const float*)W.mof, W.invf, W.logdet, W.sgn, nw, s);
      });
      return;
    }
    lu_buffers();
    launch("k_inv", [&] {
      if (precision == QW_PREC_FP64)
        hipLaunchKernelGGL((k_inv_panel<double, double>), dim3(nw), dim3(256), 0, stream, S, (const double*)W.mo,
                           luw, W.inv, W.logdet, W.sgn, nw, s);
      else
        hipLaunchKernelGGL((k_inv_panel<float, float>), dim3(nw), dim3(256), 0, stream, S, (const float*)W.mof, luw,
                           W.invf, W.logdet, W.sgn, nw, s);
    });
  }
  void big_refresh_all() {
    const size_t lds = 5 * (size_t)S.chunk_nf * sizeof(double) + (size_t)S.chunk_nf * sizeof(int) + 16;
    launch("k_mo_rows", [&] {
      if (precision == QW_PREC_FP64)
        hipLaunchKernelGGL(k_mo_rows_big<double>, dim3((unsigned)((size_t)nw * S.ne)), dim3(256), lds, stream, S,
                           (const double*)W.pos, S.C[0], S.C[1], W.mo, nw);
      else
        hipLaunchKernelGGL(k_mo_rows_big<float>, dim3((unsigned)((size_t)nw * S.ne)), dim3(256), lds, stream, S,
                           (const double*)W.pos, MT.Cf[0], MT.Cf[1], W.mof, nw);
    });
    lu_refresh(0);
    lu_refresh(1);
  }
  template <typename TI, typename R>
  void big_sweep_t(const SweepArgs& A, int depth, const BigPtrs<TI>& P, int cap) {
    // one wavefront per walker, lanes over two (<= 128 per spin) or four rows
    const size_t lds = sweep_w_lds_bytes(S.ne, S.na, S.nsh, S.ntab, S.nsp, cap, sizeof(TI));
    const bool nr2 = std::max(T.nup, T.ndn) <= 2 * kWave;
    launch("k_sweep", [&] {
      if (depth == 1) {
        if (nr2) hipLaunchKernelGGL((k_sweep_w<TI, R, 1, 2>), dim3(nw), dim3(kWave), lds, stream, S, W, P, A, cap);
        else hipLaunchKernelGGL((k_sweep_w<TI, R, 1, 4>), dim3(nw), dim3(kWave), lds, stream, S, W, P, A, cap);
      } else {
        if (nr2) hipLaunchKernelGGL((k_sweep_w<TI, R, 2, 2>), dim3(nw), dim3(kWave), lds, stream, S, W, P, A, cap);
        else hipLaunchKernelGGL((k_sweep_w<TI, R, 2, 4>), dim3(nw), dim3(kWave), lds, stream, S, W, P, A, cap);
      }
    });
  }
  void big_sweep(const SweepArgs& A0, int depth) {
    if (precision == QW_PREC_FP64) {
      BigPtrs<double> P{W.inv, W.mo, {S.C[0], S.C[1]}, S.spline};
      big_sweep_t<double, double>(A0, depth, P, kCapF64);
      return;
    }
    BigPtrs<float> P{W.invf, W.mof, {MT.Cf[0], MT.Cf[1]}, MT.spline};
    for (int k0 = A0.k0; k0 < A0.k1; k0 += refresh) {
      SweepArgs A = A0;
      A.k0 = k0;
      A.k1 = std::min(A0.k1, k0 + refresh);
      A.dets = A0.dets && A.k1 == A0.k1;
      big_sweep_t<float, float>(A, depth, P, kCapF32);
      lu_refresh(A.s);
    }
  }

  void set_device_step() {
    const unsigned v = step;
    HIPCHK(hipMemcpyAsync(dstep, &v, sizeof(unsigned), hipMemcpyHostToDevice, stream));
  }
  // One DMC feedback interval on the stream: np steps (both spin sweeps, local energy and
  // weights, step counter) then the trial-energy update and branching.  The kernels read the
  // step from the device counter, so the same launches can be recorded once and replayed.
  void dmc_interval(int np, double tau, LocalArgs A) {
    const int ne = S.ne;
    for (int p = 0; p < np; p++) {
      for (int s = 0; s < 2; s++) {
        const int n = s ? T.ndn : T.nup;
        if (n == 0) continue;
        run_sweep(s, 0, n, tau, 0u, 1, 1, 1);
      }
      A.mode = 2;
      A.use_dgl = 1;
      A.stepp = dstep;
      A.tneg = tm_mode ? tneg : nullptr;
      run_local(A);
      if (tm_mode) run_tmove(A);
      launch("k_step", [&] { hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, stream, dstep); });
    }
    launch("k_branch", [&] {
      hipLaunchKernelGGL(k_branch, dim3(1), dim3(1024), 0, stream, W, dctl, keys, vals, psrc, pdst, nw * ne * np, 1,
                         seed, 0u, (const unsigned*)dstep);
    });
    launch("k_copy", [&] {
      hipLaunchKernelGGL(k_copy, dim3(nw / 2 + 1), dim3(256), 0, stream, S, W, (const DmcCtl*)dctl,
                         (const int*)psrc, (const int*)pdst);
    });
  }
  bool use_graphs() const {
    if (big || graph_mode == 0) return false;
    return graph_mode == 1 || (long long)nw * S.ne <= (1ll << 17);
  }
  // dmc_interval through a recorded graph (recorded again when np, tau or nw change)
  void dmc_interval_graph(int np, double tau, const LocalArgs& A) {
    if (!gx || g_np != np || g_tau != tau || g_nw != nw || g_tm != tm_mode || g_ppm != A.pp_mode) {
      if (gx) {
        HIPCHK(hipGraphExecDestroy(gx));
        gx = nullptr;
      }
      hipGraph_t g = nullptr;
      std::lock_guard<std::mutex> lk(g_devmem_mu);
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      capturing = true;
      try {
        dmc_interval(np, tau, A);
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
      g_nw = nw;
      g_tm = tm_mode;
      g_ppm = A.pp_mode;
    }
    HIPCHK(hipGraphLaunch(gx, stream));
  }

  // ---- a VMC / DMC run in phases (qw_run; the batch drives several engines through the same
  // phases with shared launches for the steps)
  struct RunState {
    bool dmc = false;
    int fbi = 1;
    LocalArgs A{};
    DmcCtl ctl{};
    std::vector<double> bavg;
    long long moves = 0;
    std::chrono::steady_clock::time_point t0;
  };
  void run_begin(const qw_method_desc& m, RunState& R) {
    reset_times();
    R = RunState{};  // a run starts afresh (block averages, move count)
    R.dmc = m.method == 1;
    R.t0 = std::chrono::steady_clock::now();
    R.A = LocalArgs{};
    R.A.pp_mode = m.pp_stochastic ? 1 : 0;
    R.A.seed = seed;
    R.A.ctl = dctl;
    tm_mode = R.dmc && S.has_pp ? (m.tmoves == 1 ? 1 : m.tmoves == 2 ? 2 : 0) : 0;
    if (tm_mode) ensure_tneg();
    HIPCHK(hipMemsetAsync(W.bsum, 0, (size_t)nw * kBsum * sizeof(double), stream));
    HIPCHK(hipMemsetAsync(W.acc, 0, (size_t)nw * sizeof(int), stream));
    DmcCtl& ctl = R.ctl;
    if (R.dmc && dmc_ready) {
      // continuing a DMC run on the same population: keep E_ref, E_T, cutoffs, tau_eff
      ctl = ctl_host;
      ctl.tau = m.timestep;
      ctl.feedback = m.feedback;
      ctl.max_weight = m.max_weight;
      HIPCHK(hipMemcpyAsync(dctl, &ctl, sizeof(DmcCtl), hipMemcpyHostToDevice, stream));
    } else if (R.dmc) {
      // initial local energies, reference energy and branching cutoffs
      R.A.mode = 3;
      R.A.step = step;
      run_local(R.A);
      std::vector<double> el((size_t)nw * 4), w(nw);
      HIPCHK(hipMemcpyAsync(el.data(), W.elast, el.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(w.data(), W.weight, nw * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      double ea = 0, tw = 0;
      for (int i = 0; i < nw; i++) ea += el[4 * i + 3] * w[i], tw += w[i];
      ea /= tw;
      double v = 0;
      for (int i = 0; i < nw; i++) v += (el[4 * i + 3] - ea) * (el[4 * i + 3] - ea);
      double sig = std::sqrt(v / nw);
      ctl.eref = ea;
      ctl.cut_start = m.branch_start_cutoff * sig;
      ctl.cut_stop = ctl.cut_start * 1.5;
      ctl.feedback = m.feedback;
      ctl.max_weight = m.max_weight;
      ctl.tau = m.timestep;
      ctl.teff = m.timestep;
      // first trial energy with START_FEEDBACK (src/method/Dmc_method.cpp:771,780-787)
      ctl.etrial = ea - (m.start_feedback > 0 ? m.start_feedback : m.feedback) * std::log(tw / nw);
      HIPCHK(hipMemcpyAsync(dctl, &ctl, sizeof(DmcCtl), hipMemcpyHostToDevice, stream));
    }
    step++;
    R.fbi = std::max(1, m.feedback_interval);
  }
  // one VMC step: ndecorr sweeps of both spin channels, then the local energy
  void vmc_step(const qw_method_desc& m, RunState& R) {
    const int nsw = std::max(1, m.ndecorr);
    for (int sw = 0; sw < nsw; sw++) {
      for (int s = 0; s < 2; s++) {
        int n = s ? T.ndn : T.nup;
        if (n == 0) continue;
        run_sweep(s, 0, n, m.timestep, (unsigned)sw, 2, 0, sw == nsw - 1);
      }
    }
    R.moves += (long long)nw * S.ne * (nsw - 1);  // run_advance counts one sweep per step
    R.A.mode = 1;
    R.A.use_dgl = 1;
    R.A.stepp = dstep;
    run_local(R.A);
  }
  // after np steps (a DMC interval, or one VMC step)
  void run_advance(RunState& R, int np) {
    R.moves += (long long)nw * S.ne * np;
    step += R.dmc ? np : 1;
  }
  void run_block_end(const qw_method_desc& m, RunState& R, int b, double* eblock) {
    double red[kBsum + 1];
    reduce(red, 1);
    double sw = red[0], np_ = red[5];
    if (eblock) {
      double* eb = eblock + (size_t)QW_BLOCK_FIELDS * b;
      auto var = [&](int c) { return red[c + 1] / np_ - (red[c] / np_) * (red[c] / np_); };
      eb[0] = red[1] / sw;
      eb[1] = red[2] / sw;
      eb[2] = red[3] / sw;
      eb[3] = red[4] / sw;
      eb[4] = sw;
      eb[5] = np_;
      eb[6] = var(6);
      eb[7] = var(8);
      eb[8] = var(10);
      eb[9] = var(12);
      eb[10] = sw / np_;
      eb[11] = red[14] / np_ - (sw / np_) * (sw / np_);
    }
    R.bavg.push_back(red[1] / sw);
    if (R.dmc) {
      // E_ref <- average over the blocks so far, then the trial energy from the population
      double s = 0;
      for (double x : R.bavg) s += x;
      DmcCtl& ctl = R.ctl;
      HIPCHK(hipMemcpyAsync(&ctl, dctl, sizeof(DmcCtl), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      ctl.eref = s / R.bavg.size();
      ctl.etrial = ctl.eref - m.feedback * std::log(red[kBsum] / nw);
      HIPCHK(hipMemcpyAsync(dctl, &ctl, sizeof(DmcCtl), hipMemcpyHostToDevice, stream));
    }
  }
  void run_finish(const qw_method_desc& m, RunState& R, double* stats) {
    DmcCtl& ctl = R.ctl;
    if (R.dmc) {
      HIPCHK(hipMemcpyAsync(&ctl, dctl, sizeof(DmcCtl), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      ctl_host = ctl;
      dmc_ready = true;
    }
    // acceptance: sum of per-walker accepted counters (DMC resets them at each branch)
    std::vector<int> acc(nw);
    HIPCHK(hipMemcpyAsync(acc.data(), W.acc, nw * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double accepted_total = 0;
    for (int a : acc) accepted_total += a;
    drain_events();
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - R.t0).count();
    if (stats) {
      stats[0] = R.dmc ? ctl.teff / m.timestep : (R.moves ? accepted_total / (double)R.moves : 0.0);
      stats[1] = (double)R.moves;
      stats[2] = secs;
      stats[3] = secs > 0 ? R.moves / secs : 0.0;
    }
  }

  void reduce(double* host, int reset) {
    launch("k_reduce", [&] { hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, stream, W, dred, reset); });
    HIPCHK(hipMemcpyAsync(host, dred, (kBsum + 1) * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
};

