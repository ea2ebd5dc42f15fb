// Walker files (<runid>.config): QWalk's text checkpoint of a walker population, read and
// written in the reference's own layout so either program restarts from the other's file.
//
//   write_configurations (src/program_utils/qmc_io.h:371-423): the old file is renamed to
//   <file>.backup, then per walker " walker { \n", the walker's own record, " } \n", at
//   stream precision 15.  A walker record is Config_save_point::write
//   (src/system/Sample_point.cpp:140-148): "nElec N ndim 3", then one line "x y z " per
//   electron; a DMC walker (Dmc_point::write, src/method/Dmc_method.cpp:1228-1250) follows it
//   with "DMC { ", "weight w", "sign s", "}".
//
//   read_configurations (src/program_utils/qmc_io.h:427-503): every "walker" token opens a
//   record; Config_save_point::read (src/system/Sample_point.cpp:120-139) skips to "nElec",
//   reads the count, "ndim" and its value, then three coordinates per electron (only three are
//   read whatever ndim says); Dmc_point::read (src/method/Dmc_method.cpp:1252-1265) then looks
//   at the next token and, when it is "DMC", reads "{", the weight and the sign by position.
//   A walker without a DMC record has weight 1.
//
// Host code only; no device calls.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <strings.h>
#include <vector>

#include "../../include/qwalk_amd.h"
#include "qw_setup.hpp"

namespace {

thread_local std::string g_config_error;

qw_status config_fail(qw_status s, const std::string& why) {
  g_config_error = why;
  return s;
}

bool caseless_eq(const std::string& a, const char* b) { return strcasecmp(a.c_str(), b) == 0; }

}  // namespace

extern "C" const char* qw_configs_error(void) { return g_config_error.c_str(); }

extern "C" qw_status qw_configs_read(const char* path, int ne, int max_nw, int* nw, double* pos, double* weights,
                                     int* signs) {
  if (!path || ne < 1 || !nw || max_nw < 0) return config_fail(QW_ERR_ARG, "qw_configs_read: bad arguments");
  std::ifstream in(path);
  if (!in) return config_fail(QW_ERR_INPUT, std::string("cannot open ") + path);
  std::string tok;
  int n = 0;
  while (in >> tok) {
    if (!caseless_eq(tok, "walker")) continue;
    if (!(in >> tok) || tok != "{") return config_fail(QW_ERR_INPUT, "expected { after walker in " + std::string(path));
    // Config_save_point::read
    while (in >> tok && tok != "nElec") {
    }
    if (tok != "nElec") return config_fail(QW_ERR_INPUT, "no nElec in walker " + std::to_string(n));
    int nel = 0, ndim = 0;
    std::string dimword;
    if (!(in >> nel >> dimword >> ndim) || dimword != "ndim")
      return config_fail(QW_ERR_INPUT, "bad nElec/ndim header in walker " + std::to_string(n));
    if (nel != ne)
      return config_fail(QW_ERR_INPUT, "walker " + std::to_string(n) + " has " + std::to_string(nel) +
                                           " electrons, the system " + std::to_string(ne));
    const bool keep = pos && n < max_nw;
    for (int k = 0; k < 3 * ne; k++) {
      double x;
      if (!(in >> x)) return config_fail(QW_ERR_INPUT, "short coordinates in walker " + std::to_string(n));
      if (keep) pos[(size_t)n * ne * 3 + k] = x;
    }
    // Dmc_point::read
    double w = 1.0;
    int sg = 1;
    std::streampos at = in.tellg();
    std::string next;
    if (in >> next && caseless_eq(next, "DMC")) {
      std::string brace, key1, key2;
      if (!(in >> brace >> key1 >> w >> key2 >> sg))
        return config_fail(QW_ERR_INPUT, "bad DMC record in walker " + std::to_string(n));
    } else {
      in.clear();
      in.seekg(at);
    }
    if (n < max_nw) {
      if (weights) weights[n] = w;
      if (signs) signs[n] = sg;
    }
    n++;
  }
  *nw = n;
  if (n == 0) return config_fail(QW_ERR_INPUT, std::string("no walkers in ") + path);
  return QW_OK;
}

extern "C" qw_status qw_configs_write(const char* path, int nw, int ne, const double* pos, const double* weights,
                                      const int* signs) {
  if (!path || nw < 1 || ne < 1 || !pos) return config_fail(QW_ERR_ARG, "qw_configs_write: bad arguments");
  const std::string file = path, backup = file + ".backup";
  std::rename(file.c_str(), backup.c_str());  // as the reference: no file yet is not an error
  std::ostringstream os;
  os.precision(15);
  for (int i = 0; i < nw; i++) {
    os << " walker { \n";
    os << "nElec " << ne << " ndim " << 3 << "\n";
    for (int e = 0; e < ne; e++) {
      const double* p = &pos[((size_t)i * ne + e) * 3];
      os << p[0] << " " << p[1] << " " << p[2] << " " << "\n";
    }
    if (weights) {
      os << "DMC { \n";
      os << "weight " << weights[i] << "\n";
      os << "sign " << (signs ? signs[i] : 1) << "\n";
      os << "}\n";
    }
    os << " } \n";
  }
  std::ofstream out(file);
  if (!out) return config_fail(QW_ERR_INPUT, "cannot write " + file + ": " + strerror(errno));
  out << os.str();
  out.close();
  if (!out) return config_fail(QW_ERR_INPUT, "write failed: " + file);
  return QW_OK;
}

extern "C" qw_status qw_method_files(const char* deck_path, int index, char* readconfig, char* storeconfig,
                                     int len) {
  if (!deck_path || index < 0 || len < 1) return config_fail(QW_ERR_ARG, "qw_method_files: bad arguments");
  try {
    std::vector<qw::MethodDesc> all = qw::read_methods(deck_path);
    if (index >= (int)all.size()) return config_fail(QW_ERR_ARG, "no METHOD section " + std::to_string(index));
    const qw::MethodDesc& m = all[index];
    if ((int)m.readconfig.size() >= len || (int)m.storeconfig.size() >= len)
      return config_fail(QW_ERR_ARG, "qw_method_files: file name longer than the buffer");
    if (readconfig) snprintf(readconfig, (size_t)len, "%s", m.readconfig.c_str());
    if (storeconfig) snprintf(storeconfig, (size_t)len, "%s", m.storeconfig.c_str());
    return QW_OK;
  } catch (std::exception& x) {
    return config_fail(QW_ERR_INPUT, x.what());
  }
}
