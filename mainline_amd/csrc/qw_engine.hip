// Walker engine: device memory, kernel launches, VMC/DMC drivers and the C ABI.
#include "qw_engine.hpp"

std::mutex g_devmem_mu;

namespace {
thread_local std::string g_create_error;
}  // namespace

// Per-configuration state of the allocation-free evaluation (qw_eval_fixed_static): a
// file-scope device array, one copy on every GPU the library runs on, held until the process
// ends.  Host code never passes the array's name to a kernel or a copy: it uses the device
// address that hipGetSymbolAddress returns for the current device (scratch_address).
constexpr size_t kScratchBytes = (size_t)QW_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device ordinal -> address of that device's scratch

// Address of `scratch` on device `dev`, which must be the current device (looked up once per
// device).
void* scratch_address(int dev) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (r != hipSuccess || !p)
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
  g_scratch_ptr[dev] = p;
  return p;
}

// Clears n16 16-byte words (a fill kernel rather than a memset node, so the call can be
// recorded into a graph).
__global__ __launch_bounds__(256) void k_fill_zero(uint4* p, size_t n16) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}
}  // namespace

#include "qw_mixed_host.hpp"

// ====================================================================== C ABI
namespace {
template <typename F>
qw_status guarded(qw_engine* e, F f) {
  try {
    if (e) HIPCHK(hipSetDevice(e->device));
    return f();
  } catch (DevErr& x) {
    if (e) e->err = x.what();
    return QW_ERR_DEVICE;
  } catch (Unsupported& x) {
    if (e) e->err = x.what();
    return QW_ERR_UNSUPPORTED;
  } catch (std::bad_alloc&) {
    if (e) e->err = "host allocation failed";
    return QW_ERR_NOMEM;
  } catch (std::exception& x) {
    if (e) e->err = x.what();
    return QW_ERR_INPUT;
  }
}
}  // namespace

namespace {
// An engine for parsed deck tables: the options, the size and feature checks, then the
// device tables on opts->device.
qw_status make_engine(SysTables T, const qw_engine_opts* opts, qw_engine** out) {
  std::unique_ptr<qw_engine> e(new qw_engine);
  if (opts) {
    e->precision = opts->precision;
    e->device = opts->device;
    e->seed = opts->seed;
    if (opts->refresh_interval > 0) e->refresh = opts->refresh_interval;
  }
  if (e->precision < 0 || e->precision > 3) {
    g_create_error = "unknown precision";
    return QW_ERR_ARG;
  }
  e->T = std::move(T);
  const int nmax = std::max(e->T.nup, e->T.ndn);
  if (!opts || opts->refresh_interval <= 0) e->refresh = default_refresh(nmax);
  if (nmax > kBigMaxSpin) {
    g_create_error = "more than 256 electrons of one spin are not supported by this engine build";
    return QW_ERR_UNSUPPORTED;
  }
  if (nmax > 64) {
    // one wavefront per walker with the inverse in device memory: no pseudopotential
    // quadrature, no MFMA LCAO
    if (!e->T.pptypes.empty()) {
      g_create_error = "pseudopotentials with more than 64 electrons of one spin are not supported";
      return QW_ERR_UNSUPPORTED;
    }
    if (e->precision == QW_PREC_MIXED_BF16 || e->precision == QW_PREC_MIXED_BF16X6) {
      g_create_error = "the BF16 LCAO modes support up to 64 electrons of one spin";
      return QW_ERR_UNSUPPORTED;
    }
    e->big = true;
  }
  if (e->T.ndet > 1) {
    if (e->precision != QW_PREC_FP64) {
      g_create_error = "several determinants are supported with the fp64 precision only";
      return QW_ERR_UNSUPPORTED;
    }
    if (nmax > kMdMaxSpin) {
      g_create_error = "several determinants are supported up to 32 electrons of one spin";
      return QW_ERR_UNSUPPORTED;
    }
    e->md = true;
  }
  try {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (e->device < 0 || e->device >= ndev) throw DevErr("no HIP device " + std::to_string(e->device));
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->ns = pick_ns(std::max(e->T.nup, e->T.ndn));
    e->build_device_system();
    if (e->precision != QW_PREC_FP64) mixed_setup(*e);
    e->scratch = scratch_address(e->device);
  } catch (std::exception& x) {
    g_create_error = x.what();
    return QW_ERR_DEVICE;
  }
  if (const char* g = std::getenv("QW_GRAPHS")) e->graph_mode = (g[0] == '0') ? 0 : (g[0] == '1') ? 1 : -1;
  if (const char* t = std::getenv("QW_TUNE")) e->tune_enabled = t[0] != '0';
  if (const char* v = std::getenv("QW_INV")) e->inv_rows = std::strcmp(v, "rows") == 0;
  if (const char* c = std::getenv("QW_SWEEP"))
    for (int x : e->cands())
      if (x == std::atoi(c)) e->tune_u = x;
  *out = e.release();
  return QW_OK;
}
}  // namespace

extern "C" qw_status qw_engine_create(const char* deck_path, const qw_engine_opts* opts, qw_engine** out) {
  if (!out || !deck_path) {
    g_create_error = "null argument";
    return QW_ERR_ARG;
  }
  *out = nullptr;
  if (opts && (opts->precision < 0 || opts->precision > 3)) {
    g_create_error = "unknown precision";
    return QW_ERR_ARG;
  }
  bool unsup = false;
  SysTables T;
  try {
    T = read_deck(deck_path, &unsup);
  } catch (std::exception& x) {
    g_create_error = x.what();
    return unsup ? QW_ERR_UNSUPPORTED : QW_ERR_INPUT;
  }
  return make_engine(std::move(T), opts, out);
}

extern "C" qw_status qw_engine_clone(const qw_engine* src, const qw_engine_opts* opts, qw_engine** out) {
  if (!out || !src) {
    g_create_error = "null argument";
    return QW_ERR_ARG;
  }
  *out = nullptr;
  return make_engine(src->T, opts, out);
}

extern "C" void qw_engine_destroy(qw_engine* e) { delete e; }

extern "C" qw_status qw_engine_tuning(const qw_engine* e, int* out4) {
  if (!e || !out4) return QW_ERR_ARG;
  out4[0] = e->tune_u;
  out4[1] = e->use_graphs() ? 1 : 0;
  out4[2] = e->big ? 1 : 0;
  out4[3] = e->ns;
  return QW_OK;
}

extern "C" qw_status qw_engine_walker_bytes(const qw_engine* e, int nw, unsigned long long* bytes) {
  if (!e || !bytes || nw < 0) return QW_ERR_ARG;
  *bytes = e->walker_bytes(nw);
  return QW_OK;
}

extern "C" const char* qw_last_error(const qw_engine* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

extern "C" qw_status qw_engine_info(const qw_engine* e, int* out6) {
  if (!e || !out6) return QW_ERR_ARG;
  out6[0] = e->T.nup;
  out6[1] = e->T.ndn;
  out6[2] = e->T.na;
  out6[3] = e->T.nb;
  out6[4] = e->T.pptypes.empty() ? 0 : 1;
  out6[5] = e->cap;
  return QW_OK;
}

extern "C" qw_status qw_walkers_set(qw_engine* e, int n, const double* pos, const double* weights) {
  if (!e || !pos || n <= 0) return QW_ERR_ARG;
  return guarded(e, [&] {
    e->ensure_walkers(n);
    const size_t ne = e->S.ne;
    HIPCHK(hipMemcpyAsync(e->W.pos, pos, (size_t)n * ne * 3 * sizeof(double), hipMemcpyHostToDevice, e->stream));
    std::vector<double> w(n, 1.0);
    if (weights) std::copy(weights, weights + n, w.begin());
    HIPCHK(hipMemcpyAsync(e->W.weight, w.data(), n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemsetAsync(e->W.bsum, 0, (size_t)n * kBsum * sizeof(double), e->stream));
    HIPCHK(hipMemsetAsync(e->W.acc, 0, (size_t)n * sizeof(int), e->stream));
    e->refresh_all();
    HIPCHK(hipStreamSynchronize(e->stream));
    e->dmc_ready = false;
    return QW_OK;
  });
}

namespace {
// Host Philox4x32-10 (same streams as the device helper) for walker initialisation.
void philox_host(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1, unsigned out[4]) {
  for (int r = 0; r < 10; r++) {
    unsigned long long p0 = 0xD2511F53ull * c0, p1 = 0xCD9E8D57ull * c2;
    unsigned n0 = (unsigned)(p1 >> 32) ^ c1 ^ k0, n2 = (unsigned)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (unsigned)p1;
    c2 = n2;
    c3 = (unsigned)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
}  // namespace

extern "C" qw_status qw_walkers_random(const qw_engine* e, int nw, unsigned long long seed, double inirange,
                                       double* pos) {
  if (!e || nw <= 0 || !pos || inirange <= 0) return QW_ERR_ARG;
  const SysTables& T = e->T;
  const int ne = T.ne(), na = T.na;
  double cen[3] = {0, 0, 0};
  for (int a = 0; a < na; a++)
    for (int d = 0; d < 3; d++) cen[d] += T.atom_pos[3 * a + d] / na;
  for (int w = 0; w < nw; w++) {
    double* p = pos + (size_t)w * ne * 3;
    auto u3 = [&](int el, double* out) {
      unsigned o[4];
      philox_host((unsigned)w, (unsigned)el, 0u, 7u, (unsigned)seed, (unsigned)(seed >> 32), o);
      for (int d = 0; d < 3; d++) out[d] = (o[d] + 0.5) * (1.0 / 4294967296.0);
    };
    int el = 0;
    for (int spin = 0; spin < 2; spin++)
      for (int a = 0; a < na; a++) {
        const int zi = (int)T.atom_Z[a];
        const int nplace = (int)(T.atom_Z[a] / 2) + ((1 - spin) * zi) % 2;
        for (int k = 0; k < nplace; k++, el++) {
          if (el >= ne) continue;
          double u[3];
          u3(el, u);
          for (int d = 0; d < 3; d++) p[3 * el + d] = T.atom_pos[3 * a + d] + 2 * inirange * (u[d] - .5);
        }
      }
    for (; el < ne; el++) {
      double u[3];
      u3(el, u);
      for (int d = 0; d < 3; d++) p[3 * el + d] = 4 * inirange * (u[d] - .5) + cen[d];
    }
  }
  return QW_OK;
}

extern "C" qw_status qw_walkers_get(qw_engine* e, int n, double* pos, double* weights) {
  if (!e || n <= 0 || n > e->nw) return QW_ERR_ARG;
  return guarded(e, [&] {
    const size_t ne = e->S.ne;
    if (pos) HIPCHK(hipMemcpyAsync(pos, e->W.pos, (size_t)n * ne * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (weights) HIPCHK(hipMemcpyAsync(weights, e->W.weight, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return QW_OK;
  });
}

extern "C" qw_status qw_walkers_state(qw_engine* e, int n, double* logdet, double* sign, double* elast, double* dgl) {
  if (!e || n <= 0 || n > e->nw) return QW_ERR_ARG;
  return guarded(e, [&] {
    if (logdet) HIPCHK(hipMemcpyAsync(logdet, e->W.logdet, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (sign) HIPCHK(hipMemcpyAsync(sign, e->W.sgn, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (elast) HIPCHK(hipMemcpyAsync(elast, e->W.elast, (size_t)n * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (dgl)
      HIPCHK(hipMemcpyAsync(dgl, e->W.dgl, (size_t)n * e->S.ne * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return QW_OK;
  });
}

extern "C" qw_status qw_parm_count(const qw_engine* e, int* np, double* values) {
  if (!e || !np) return QW_ERR_ARG;
  *np = e->nslp + e->npar;
  if (values)
    for (int p = 0; p < *np; p++) values[p] = e->parm_values[p];
  return QW_OK;
}

extern "C" qw_status qw_parm_deriv(qw_engine* e, int K, const double* pos, double* gradient, double* hessian,
                                   double* gradderiv, double* val_gradient) {
  if (!e || K <= 0 || !pos || !gradient || !gradderiv || !val_gradient) return QW_ERR_ARG;
  return guarded(e, [&] {
    e->ensure_walkers(K);
    e->dmc_ready = false;
    const size_t ne = e->S.ne, np = (size_t)e->npar, nps = (size_t)e->nslp, nt = np + nps;
    HIPCHK(hipMemcpyAsync(e->W.pos, pos, (size_t)K * ne * 3 * sizeof(double), hipMemcpyHostToDevice, e->stream));
    e->refresh_all();
    std::vector<void*> tmp;
    auto dalloc = [&](size_t bytes) {
      void* p = nullptr;
      {
        std::lock_guard<std::mutex> g(g_devmem_mu);
        HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
      }
      tmp.push_back(p);
      return (double*)p;
    };
    struct Free {
      std::vector<void*>& v;
      hipStream_t st;
      ~Free() {
        (void)hipStreamSynchronize(st);
        std::lock_guard<std::mutex> g(g_devmem_mu);
        for (void* p : v) (void)hipFree(p);
      }
    } fr{tmp, e->stream};
    double* d_g = dalloc((size_t)K * np * sizeof(double));
    double* d_gd = dalloc((size_t)K * np * ne * 4 * sizeof(double));
    double* d_vg = dalloc((size_t)K * ne * 3 * sizeof(double));
    double* d_vj = nps && np ? dalloc((size_t)K * ne * 3 * sizeof(double)) : nullptr;
    const size_t lds = (size_t)9 * ne * sizeof(double) + (e->md ? (size_t)5 * e->S.ndet * sizeof(double) : 0);
    e->launch("k_parm_deriv", [&] {
      if (e->md)
        hipLaunchKernelGGL((k_parm_deriv<double, true>), dim3(K), dim3(64), lds, e->stream, e->S, e->W,
                           (const JParmDev*)e->d_parm, (const JRad<double>*)e->d_punit, (const int*)e->d_patoms,
                           (int)np, K, d_g, d_gd, d_vg, d_vj);
      else if (e->precision == QW_PREC_FP64)
        hipLaunchKernelGGL(k_parm_deriv<double>, dim3(K), dim3(64), lds, e->stream, e->S, e->W,
                           (const JParmDev*)e->d_parm, (const JRad<double>*)e->d_punit, (const int*)e->d_patoms,
                           (int)np, K, d_g, d_gd, d_vg, (double*)nullptr);
      else
        hipLaunchKernelGGL(k_parm_deriv<float>, dim3(K), dim3(64), lds, e->stream, e->S, e->W,
                           (const JParmDev*)e->d_parm, (const JRad<double>*)e->d_punit, (const int*)e->d_patoms,
                           (int)np, K, d_g, d_gd, d_vg, (double*)nullptr);
    });
    double* d_gs = nps ? dalloc((size_t)K * nps * sizeof(double)) : nullptr;
    double* d_gds = nps ? dalloc((size_t)K * nps * ne * 4 * sizeof(double)) : nullptr;
    if (nps) {
      e->launch("k_parm_det", [&] {
        hipLaunchKernelGGL(k_parm_det, dim3(K), dim3(64), parm_det_lds_bytes((int)ne, e->S.ndet), e->stream, e->S,
                           e->W, (const int*)e->d_csf_d0, (const double*)e->d_csf_w, (int)nps + 1, K,
                           (const double*)d_vj, d_gs, d_gds);
      });
    }
    std::vector<double> gj((size_t)K * np), gdj((size_t)K * np * ne * 4), gs((size_t)K * nps),
        gds((size_t)K * nps * ne * 4);
    if (np) {
      HIPCHK(hipMemcpyAsync(gj.data(), d_g, gj.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipMemcpyAsync(gdj.data(), d_gd, gdj.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    if (nps) {
      HIPCHK(hipMemcpyAsync(gs.data(), d_gs, gs.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipMemcpyAsync(gds.data(), d_gds, gds.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipMemcpyAsync(val_gradient, d_vg, (size_t)K * ne * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    // the caller's order: CSF coefficients, then Jastrow parameters
    const size_t row = ne * 4;
    for (int k = 0; k < K; k++) {
      for (size_t p = 0; p < nps; p++) {
        gradient[(size_t)k * nt + p] = gs[(size_t)k * nps + p];
        std::copy(&gds[((size_t)k * nps + p) * row], &gds[((size_t)k * nps + p) * row] + row,
                  gradderiv + ((size_t)k * nt + p) * row);
      }
      for (size_t p = 0; p < np; p++) {
        gradient[(size_t)k * nt + nps + p] = gj[(size_t)k * np + p];
        std::copy(&gdj[((size_t)k * np + p) * row], &gdj[((size_t)k * np + p) * row] + row,
                  gradderiv + ((size_t)k * nt + nps + p) * row);
      }
    }
    // hessian: g_i g_j, except between two CSF coefficients (zero), as the reference's
    // Slater determinant leaves its block zero and extend_parm_deriv fills the rest
    if (hessian)
      for (int k = 0; k < K; k++)
        for (size_t i = 0; i < nt; i++)
          for (size_t j = 0; j < nt; j++)
            hessian[((size_t)k * nt + i) * nt + j] =
                (i < nps && j < nps) ? 0.0 : gradient[(size_t)k * nt + i] * gradient[(size_t)k * nt + j];
    return QW_OK;
  });
}

extern "C" qw_status qw_eval_fixed(qw_engine* e, int K, const double* pos, const double* rot, int pp_mode,
                                   const double* pp_rand, double* logpsi, double* sign, double* grad, double* lap,
                                   double* eparts) {
  if (!e || K <= 0 || !pos || !logpsi || !sign || !grad || !lap || !eparts) return QW_ERR_ARG;
  if (pp_mode == 1 && !pp_rand && e->S.has_pp) return QW_ERR_ARG;
  return guarded(e, [&] {
    e->ensure_walkers(K);
    e->dmc_ready = false;
    const size_t ne = e->S.ne, na = e->S.na;
    HIPCHK(hipMemcpyAsync(e->W.pos, pos, (size_t)K * ne * 3 * sizeof(double), hipMemcpyHostToDevice, e->stream));
    e->refresh_all();
    std::vector<void*> tmp;
    auto dalloc = [&](size_t bytes) {
      void* p = nullptr;
      {
        std::lock_guard<std::mutex> g(g_devmem_mu);
        HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
      }
      tmp.push_back(p);
      return p;
    };
    struct Free {
      std::vector<void*>& v;
      hipStream_t st;
      ~Free() {
        (void)hipStreamSynchronize(st);
        std::lock_guard<std::mutex> g(g_devmem_mu);
        for (void* p : v) (void)hipFree(p);
      }
    } fr{tmp, e->stream};
    double* d_rot = nullptr;
    double* d_rand = nullptr;
    if (rot) {
      d_rot = (double*)dalloc(9 * sizeof(double));
      HIPCHK(hipMemcpyAsync(d_rot, rot, 9 * sizeof(double), hipMemcpyHostToDevice, e->stream));
    }
    if (pp_mode == 1 && pp_rand) {
      d_rand = (double*)dalloc((size_t)K * na * ne * sizeof(double));
      HIPCHK(hipMemcpyAsync(d_rand, pp_rand, (size_t)K * na * ne * sizeof(double), hipMemcpyHostToDevice, e->stream));
    }
    double* d_lp = (double*)dalloc(K * sizeof(double));
    double* d_sg = (double*)dalloc(K * sizeof(double));
    double* d_gr = (double*)dalloc((size_t)K * ne * 3 * sizeof(double));
    double* d_la = (double*)dalloc((size_t)K * ne * sizeof(double));
    double* d_ep = (double*)dalloc((size_t)K * 3 * sizeof(double));
    LocalArgs A{};
    A.mode = 0;
    A.pp_mode = pp_mode;
    A.rot = d_rot;
    A.pp_rand = d_rand;
    A.seed = e->seed;
    A.step = e->step;
    A.out_logpsi = d_lp;
    A.out_sign = d_sg;
    A.out_grad = d_gr;
    A.out_lap = d_la;
    A.out_eparts = d_ep;
    A.ctl = e->dctl;
    e->run_local(A);
    HIPCHK(hipMemcpyAsync(logpsi, d_lp, K * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(sign, d_sg, K * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(grad, d_gr, (size_t)K * ne * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(lap, d_la, (size_t)K * ne * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(eparts, d_ep, (size_t)K * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return QW_OK;
  });
}

extern "C" qw_status qw_eval_fixed_static_max(const qw_engine* e, int* kmax) {
  if (!e || !kmax) return QW_ERR_ARG;
  DevWalkers V;
  const size_t per = e->carve_walkers(nullptr, 1, V);
  *kmax = (std::max(e->T.nup, e->T.ndn) > kGridMaxSpin || per > kScratchBytes) ? 0 : (int)std::min<size_t>(kScratchBytes / per, 1 << 30);
  // the per-array 256-byte rounding makes K configurations need at most K x (one's bytes)
  while (*kmax > 0 && e->carve_walkers(nullptr, *kmax, V) > kScratchBytes) (*kmax)--;
  return QW_OK;
}

extern "C" qw_status qw_eval_fixed_static(qw_engine* e, int K, const double* pos, const double* rot, int pp_mode,
                                          const double* pp_rand, double* logpsi, double* sign, double* grad,
                                          double* lap, double* eparts, void* stream) {
  if (!e || K <= 0 || !pos || !logpsi || !sign || !grad || !lap || !eparts) return QW_ERR_ARG;
  if (pp_mode == 1 && !pp_rand && e->S.has_pp) return QW_ERR_ARG;
  if (std::max(e->T.nup, e->T.ndn) > kGridMaxSpin) {
    e->err = "qw_eval_fixed_static: more than 128 electrons of one spin need working matrices outside the fixed array";
    return QW_ERR_UNSUPPORTED;
  }
  DevWalkers V;
  const size_t need = e->carve_walkers(nullptr, K, V);
  if (need > kScratchBytes) {
    int kmax = 0;
    (void)qw_eval_fixed_static_max(e, &kmax);
    e->err = "qw_eval_fixed_static: " + std::to_string(K) + " configurations need " + std::to_string(need) +
             " bytes of scratch; at most " + std::to_string(kmax) + " fit in " + std::to_string(kScratchBytes);
    return QW_ERR_ARG;
  }
  return guarded(e, [&] {
    // the engine's walker view, population, stream and timing are swapped for this call and
    // restored on every exit path
    struct Swap {
      qw_engine* e;
      DevWalkers W;
      int nw;
      hipStream_t st;
      bool quiet;
      ~Swap() {
        e->W = W;
        e->nw = nw;
        e->stream = st;
        e->quiet = quiet;
      }
    } sw{e, e->W, e->nw, e->stream, e->quiet};
    char* base = (char*)e->scratch;
    e->carve_walkers(base, K, V);
    if (stream) e->stream = (hipStream_t)stream;
    e->quiet = true;
    e->W = V;
    e->nw = K;
    e->launch("k_fill", [&] {
      hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<size_t>((need / 16 + 255) / 256, 4096)), dim3(256), 0,
                         e->stream, (uint4*)base, need / 16);
    });
    const size_t ne = e->S.ne;
    HIPCHK(hipMemcpyAsync(V.pos, pos, (size_t)K * ne * 3 * sizeof(double), hipMemcpyDefault, e->stream));
    e->refresh_all();
    LocalArgs A{};
    A.mode = 0;
    A.pp_mode = pp_mode;
    A.rot = rot;
    A.pp_rand = pp_mode == 1 ? pp_rand : nullptr;
    A.seed = e->seed;
    A.step = e->step;
    A.out_logpsi = logpsi;
    A.out_sign = sign;
    A.out_grad = grad;
    A.out_lap = lap;
    A.out_eparts = eparts;
    A.ctl = e->dctl;
    e->run_local(A);
    return QW_OK;
  });
}

namespace {
void method_to_abi(const MethodDesc& d, qw_method_desc* out) {
  out->method = d.method;
  out->nconfig = d.nconfig;
  out->nblock = d.nblock;
  out->nstep = d.nstep;
  out->ndecorr = d.ndecorr;
  out->timestep = d.timestep;
  out->feedback_interval = d.feedback_interval;
  out->feedback = d.feedback;
  out->max_weight = d.max_weight;
  out->branch_start_cutoff = d.branch_start_cutoff;
  out->pp_stochastic = d.pp_stochastic;
  out->tmoves = d.tmoves;
  out->start_feedback = d.start_feedback;
}
}  // namespace

extern "C" qw_status qw_method_read(const char* deck_path, qw_method_desc* out) {
  return qw_method_read_at(deck_path, 0, out, nullptr, 0);
}

extern "C" qw_status qw_method_count(const char* deck_path, int* n) {
  if (!deck_path || !n) return QW_ERR_ARG;
  try {
    *n = (int)read_methods(deck_path).size();
    return QW_OK;
  } catch (std::exception& x) {
    g_create_error = x.what();
    return QW_ERR_INPUT;
  }
}

extern "C" qw_status qw_method_read_at(const char* deck_path, int index, qw_method_desc* out, char* label,
                                       int label_len) {
  if (!deck_path || !out || index < 0) return QW_ERR_ARG;
  try {
    std::vector<MethodDesc> all = read_methods(deck_path);
    if (index >= (int)all.size()) {
      g_create_error = "no METHOD section " + std::to_string(index);
      return QW_ERR_ARG;
    }
    method_to_abi(all[index], out);
    if (label && label_len > 0) snprintf(label, (size_t)label_len, "%s", all[index].label.c_str());
    return QW_OK;
  } catch (std::exception& x) {
    g_create_error = x.what();
    return QW_ERR_INPUT;
  }
}

// VMC: src/method/Vmc_method.cpp:351-558.  DMC: src/method/Dmc_method.cpp:328-600,776-1156.
extern "C" qw_status qw_run(qw_engine* e, const qw_method_desc* m, double* eblock, double* stats) {
  if (!e || !m || e->nw <= 0) return QW_ERR_ARG;
  if (m->timestep <= 0 || m->nblock < 0 || m->nstep < 1) {
    e->err = "qw_run needs timestep > 0, nblock >= 0, nstep >= 1";
    return QW_ERR_ARG;
  }
  return guarded(e, [&] {
    qw_engine::RunState R;
    e->run_begin(*m, R);
    for (int b = 0; b < m->nblock; b++) {
      if (!R.dmc) e->refresh_all();
      for (int st = 0; st < m->nstep;) {
        const int np = R.dmc ? std::min(R.fbi, m->nstep - st) : 1;
        e->set_device_step();
        if (R.dmc) {
          if (e->use_graphs()) e->dmc_interval_graph(np, m->timestep, R.A);
          else e->dmc_interval(np, m->timestep, R.A);
        } else {
          e->vmc_step(*m, R);
        }
        e->run_advance(R, np);
        st += np;
      }
      e->run_block_end(*m, R, b, eblock);
    }
    e->run_finish(*m, R, stats);
    return QW_OK;
  });
}

extern "C" qw_status qw_kernel_times(const qw_engine* e, char* buf, int buflen, double* avg_ms, int* count, int n,
                                     int* nk) {
  if (!e) return QW_ERR_ARG;
  // fold in the launches still pending on the stream (reading the totals does not change
  // what the engine computes)
  try {
    const_cast<qw_engine*>(e)->drain_events();
  } catch (std::exception&) {
    return QW_ERR_DEVICE;
  }
  std::string names;
  int k = (int)e->knames.size();
  for (int i = 0; i < k; i++) {
    if (i) names += ",";
    names += e->knames[i];
    if (i < n) {
      if (avg_ms) avg_ms[i] = e->kcount[i] ? e->ktime[i] / e->kcount[i] : 0.0;
      if (count) count[i] = e->kcount[i];
    }
  }
  if (buf && buflen > 0) snprintf(buf, buflen, "%s", names.c_str());
  if (nk) *nk = k;
  return QW_OK;
}
