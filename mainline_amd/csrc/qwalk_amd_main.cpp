// qwalk_amd: runs the METHOD { VMC | DMC } sections of a QWalk deck on the MI355X walker engine,
// through the C ABI only, and writes QWalk's own output files so the reference's tooling
// (gosling, tests/qwtest.py) reads them unchanged:
//   <deck>.log     block records (src/properties/Properties.cpp:578-641,
//                  src/properties/Properties_block.cpp:73-138)
//   <deck>.json    one record per block, each the one-block final average followed by <RS>
//                  (src/method/Dmc_method.cpp:541-549, src/method/Vmc_method.cpp:526-534,
//                  src/properties/Properties_average.cpp:449-520)
//   STORECONFIG    walker positions (+ DMC weight and sign), <deck>.config by default
//                  (qw_configs_write: src/program_utils/qmc_io.h:371-423)
// Every METHOD section runs in input order (src/program_utils/ooqmc.cpp:148-170); each takes
// its walkers from its READCONFIG file (<deck>.config by default, so the previous method's
// STORECONFIG) when it exists, otherwise from the reference's random guess followed by VMC
// equilibration (src/method/Generate_sample.cpp:24-71, src/method/Vmc_method.cpp:362,
// src/method/Dmc_method.cpp:347,733-759);
// a VMC section without TIMESTEP gets a timestep tuned toward acceptance 0.7.
//
// usage: qwalk_amd <deck> [--precision fp64|mixed_fp32|mixed_bf16|mixed_bf16x6] [--device N]
//                         [--seed S] [--walkers N]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/qwalk_amd.h"

namespace {

int fail(const char* what, const qw_engine* e) {
  fprintf(stderr, "qwalk_amd: %s: %s\n", what, qw_last_error(e));
  return 1;
}

void write_log(const std::string& path, const char* label, double tau, bool dmc, const std::vector<double>& eb,
               int nblock) {
  std::ofstream lo(path, std::ios::app);
  lo.precision(16);
  lo << "#-------------------------------------------------\n";
  lo << "#" << (dmc ? "DMC" : "VMC") << " run: timestep " << tau << "\n";
  lo << "#-------------------------------------------------\n\n\n";
  // the run's header record, with no average generators (src/properties/Properties.cpp:1020-1034)
  lo << "init { \n   label " << label << "\n}\n";
  for (int b = 0; b < nblock; b++) {
    const double* r = &eb[(size_t)b * QW_BLOCK_FIELDS];
    auto sec = [&](const char* name, double v) { lo << "   " << name << " { " << v << "   } \n"; };
    lo << "block { \n";
    lo << "   label " << label << "\n";
    lo << "   totweight " << r[5] << "\n";
    lo << "   nautocorr 0\n";
    sec("total_energy", r[0]);
    sec("total_energyvar", r[6]);
    sec("weight", r[10]);
    sec("weightvar", r[11]);
    sec("kinetic_energy", r[1]);
    sec("kinetic_energyvar", r[7]);
    sec("potential_energy", r[2]);
    sec("potential_energyvar", r[8]);
    sec("nonlocal_energy", r[3]);
    sec("nonlocal_energyvar", r[9]);
    lo << "   autocorr_energy { } \n";
    lo << "} \n\n";
  }
}

// Per block: the reduction of that one block alone, so value = the block average, error = 0,
// sigma = the square root of the block variance, and the independent-point estimate
// sigma² / error² (infinite, with the reference's warning, whenever the variance is nonzero).
void write_json(const std::string& path, const std::vector<double>& eb, int nblock) {
  std::ofstream js(path, std::ios::app);  // default stream precision, as the reference's
  static const char* names[] = {"total_energy", "kinetic", "potential", "nonlocal", "weight"};
  static const int avg_at[] = {0, 1, 2, 3, 10}, var_at[] = {6, 7, 8, 9, 11};
  for (int b = 0; b < nblock; b++) {
    const double* r = &eb[(size_t)b * QW_BLOCK_FIELDS];
    const double err = 0.0;
    js << "{\n\"warmup blocks\":0,\n\"properties\":{\n";
    for (int i = 0; i < 5; i++) {
      js << "\"" << names[i] << "\":{\n";
      js << "\"value\":[" << r[avg_at[i]] << "],\n";
      js << "\"error\":[" << std::sqrt(err) << "],\n";
      js << "\"sigma\":[" << std::sqrt(r[var_at[i]]) << "]\n";
      js << "}" << (i != 4 ? "," : "") << "\n";
    }
    js << "},\n";
    const double indep = std::sqrt(r[6]) / std::sqrt(err), totpoints = indep * indep;
    js << "\"independent points\":" << totpoints << ",\n";
    js << "\"commments\": [\n";
    if (totpoints > r[5])
      js << "\"Warning: The estimated number of independent points is _greater_ that the points in the log "
            "file. Error bars are probably unreliable! \"\n";
    js << "]\n}\n<RS>\n";
  }
}

// Mean acceptance of a short VMC segment at the given timestep (the walkers move on).
bool vmc_acceptance(qw_engine* e, qw_method_desc vmc, double tau, int nstep, double* acc) {
  vmc.method = 0;
  vmc.nblock = 1;
  vmc.nstep = nstep;
  vmc.timestep = tau;
  std::vector<double> tmp(QW_BLOCK_FIELDS);
  double st[4];
  if (qw_run(e, &vmc, tmp.data(), st) != QW_OK) return false;
  *acc = st[0];
  return true;
}

// VMC without TIMESTEP: the reference nudges the timestep by 0.01 after every move of its
// warm-up chain, up while the move's acceptance exceeds 0.7 and down otherwise
// (src/method/Generate_sample.cpp:40-53, src/method/Vmc_method.cpp:364-369).  With a whole
// population the same target is reached by measuring the acceptance of short segments and
// solving for the timestep whose mean acceptance is 0.7 (secant steps on log tau).
bool tune_timestep(qw_engine* e, const qw_method_desc& vmc, int nw, int ne, double* tau_out) {
  const double target = 0.7;
  // segments of at least ~2e4 moves, so a segment's acceptance is known to about 0.005
  const int nstep = std::max(4, std::min(2000, (int)std::ceil(2e4 / ((double)nw * ne))));
  double t0 = 1.0, a0 = 0;
  if (!vmc_acceptance(e, vmc, t0, nstep, &a0)) return false;
  double t1 = a0 > target ? 2.0 : 0.5, a1 = 0;
  for (int it = 0; it < 12; it++) {
    if (!vmc_acceptance(e, vmc, t1, nstep, &a1)) return false;
    if (std::fabs(a1 - target) < 0.01) break;
    double l0 = std::log(t0), l1 = std::log(t1), l2;
    if (std::fabs(a1 - a0) > 1e-6)
      l2 = l1 + (target - a1) * (l1 - l0) / (a1 - a0);
    else
      l2 = l1 + (a1 > target ? 0.5 : -0.5);
    l2 = std::max(l1 - 1.0, std::min(l1 + 1.0, l2));  // at most a factor e per step
    t0 = t1;
    a0 = a1;
    t1 = std::exp(l2);
  }
  *tau_out = t1;
  return true;
}

bool parse_precision(const std::string& v, int* p) {
  if (v == "fp64") *p = QW_PREC_FP64;
  else if (v == "mixed_fp32") *p = QW_PREC_MIXED_FP32;
  else if (v == "mixed_bf16") *p = QW_PREC_MIXED_BF16;
  else if (v == "mixed_bf16x6") *p = QW_PREC_MIXED_BF16X6;
  else return false;
  return true;
}

const char* kUsage =
    "usage: qwalk_amd <deck> [--precision fp64|mixed_fp32|mixed_bf16|mixed_bf16x6] [--device N] [--seed S] "
    "[--walkers N]\n";

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "%s", kUsage);
    return 2;
  }
  std::string deck = argv[1];
  qw_engine_opts opts{QW_PREC_FP64, 0, 20261202ull, 0};
  int walkers_override = 0;
  for (int i = 2; i < argc; i += 2) {
    std::string k = argv[i];
    if (i + 1 >= argc) {
      fprintf(stderr, "qwalk_amd: %s needs a value\n%s", k.c_str(), kUsage);
      return 2;
    }
    std::string v = argv[i + 1];
    if (k == "--precision") {
      if (!parse_precision(v, &opts.precision)) {
        fprintf(stderr, "qwalk_amd: unknown precision '%s'\n%s", v.c_str(), kUsage);
        return 2;
      }
    } else if (k == "--device") {
      opts.device = atoi(v.c_str());
    } else if (k == "--seed") {
      opts.seed = strtoull(v.c_str(), nullptr, 10);
    } else if (k == "--walkers") {
      walkers_override = atoi(v.c_str());
    } else {
      fprintf(stderr, "qwalk_amd: unknown option '%s'\n%s", k.c_str(), kUsage);
      return 2;
    }
  }
  int nmethod = 0;
  if (qw_method_count(deck.c_str(), &nmethod) != QW_OK) return fail("METHOD", nullptr);
  qw_engine* e = nullptr;
  if (qw_engine_create(deck.c_str(), &opts, &e) != QW_OK) return fail("engine", nullptr);
  int info[6];
  qw_engine_info(e, info);
  const int ne = info[0] + info[1];
  std::vector<double> pos, w;
  for (int mi = 0; mi < nmethod; mi++) {
    qw_method_desc m;
    char label[64];
    if (qw_method_read_at(deck.c_str(), mi, &m, label, sizeof label) != QW_OK) return fail("METHOD", nullptr);
    const bool dmc = m.method == 1;
    if (walkers_override > 0) m.nconfig = walkers_override;
    char readcfg[4096], storecfg[4096];
    if (qw_method_files(deck.c_str(), mi, readcfg, storecfg, sizeof readcfg) != QW_OK) {
      fprintf(stderr, "qwalk_amd: METHOD: %s\n", qw_configs_error());
      return 1;
    }
    // walkers, as the reference takes them: every method reads its READCONFIG file (the
    // previous method's STORECONFIG by default) when that exists.  DMC keeps the first NCONFIG
    // walkers with their weights and stops when the file holds fewer
    // (src/method/Dmc_method.cpp:733-759); VMC keeps the first NCONFIG and fills the rest with
    // the random guess (src/method/Vmc_method.cpp:270-289).  Without a file both start from the
    // random guess; VMC always runs its warm-up first (src/method/Vmc_method.cpp:362-364).
    const int nwn = m.nconfig;
    pos.assign((size_t)nwn * ne * 3, 0.0);
    w.assign(nwn, 1.0);
    int nr = 0;
    if (FILE* f = fopen(readcfg, "r")) {
      fclose(f);
      if (qw_configs_read(readcfg, ne, nwn, &nr, pos.data(), w.data(), nullptr) != QW_OK) {
        fprintf(stderr, "qwalk_amd: %s\n", qw_configs_error());
        return 1;
      }
      if (dmc && nr < nwn) {
        fprintf(stderr, "qwalk_amd: not enough configurations in %s (%d, DMC needs %d)\n", readcfg, nr, nwn);
        return 1;
      }
      printf("read %d walkers from %s\n", std::min(nr, nwn), readcfg);
      if (!dmc) std::fill(w.begin(), w.end(), 1.0);
    }
    const bool fresh = nr == 0;
    if (nr < nwn &&
        qw_walkers_random(e, nwn - nr, opts.seed + (unsigned long long)mi, 3.0, pos.data() + (size_t)nr * ne * 3) !=
            QW_OK)
      return fail("walkers", e);
    const int nw = nwn;
    if (qw_walkers_set(e, nw, pos.data(), w.data()) != QW_OK) return fail("walkers", e);
    const bool auto_tau = m.timestep <= 0;
    if (fresh || !dmc) {
      // warm-up from the random guess, and before every VMC method: VMC with depth-2 split
      // moves (the reference's warm-up starts from timestep 1 when it tunes it)
      qw_method_desc eq = m;
      eq.method = 0;
      eq.nblock = 1;
      eq.nstep = 50;
      eq.ndecorr = 2;
      eq.timestep = auto_tau ? 1.0 : dmc ? std::max(m.timestep, 0.05) : m.timestep;
      std::vector<double> tmp(QW_BLOCK_FIELDS);
      if (qw_run(e, &eq, tmp.data(), nullptr) != QW_OK) return fail("equilibration", e);
    }
    if (auto_tau && !tune_timestep(e, m, nw, ne, &m.timestep)) return fail("timestep", e);
    std::vector<double> eb((size_t)std::max(m.nblock, 1) * QW_BLOCK_FIELDS);
    double stats[4];
    if (qw_run(e, &m, eb.data(), stats) != QW_OK) return fail("run", e);
    write_log(deck + ".log", label, m.timestep, dmc, eb, m.nblock);
    write_json(deck + ".json", eb, m.nblock);
    if (qw_walkers_get(e, nw, pos.data(), w.data()) != QW_OK) return fail("walkers", e);
    if (qw_configs_write(storecfg, nw, ne, pos.data(), dmc ? w.data() : nullptr, nullptr) != QW_OK) {
      fprintf(stderr, "qwalk_amd: %s\n", qw_configs_error());
      return 1;
    }
    double s = 0, s2 = 0;
    for (int b = 0; b < m.nblock; b++) s += eb[(size_t)b * QW_BLOCK_FIELDS];
    s /= std::max(m.nblock, 1);
    for (int b = 0; b < m.nblock; b++) s2 += std::pow(eb[(size_t)b * QW_BLOCK_FIELDS] - s, 2);
    double err = m.nblock > 1 ? std::sqrt(s2 / (m.nblock - 1) / m.nblock) : 0.0;
    printf("%s (%s): %d walkers, %d blocks x %d steps, timestep %g%s\n", dmc ? "DMC" : "VMC", label, nw, m.nblock,
           m.nstep, m.timestep, auto_tau ? " (tuned)" : "");
    printf("total_energy %.10f +/- %.10f\n", s, err);
    printf("acceptance %.4f  electron-moves/s %.4e  wall %.2f s\n", stats[0], stats[3], stats[2]);
  }
  qw_engine_destroy(e);
  return 0;
}
