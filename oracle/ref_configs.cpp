// Test infrastructure only (never linked into the product).
//
// Our own driver around the reference's walker-file routines, compiled with the reference
// sources by oracle/Makefile.ref: read_configurations / write_configurations
// (src/program_utils/qmc_io.h:371-503) over Dmc_point (src/method/Dmc_method.cpp:1228-1265,
// a Config_save_point plus DMC weight and sign) or Config_save_point
// (src/system/Sample_point.cpp:120-148).
//
// Usage:
//   qw_configs read <file> <Ne>
//       reads <file> with read_configurations<Dmc_point> and prints
//         nwalkers N
//         walker i weight w sign s        (then Ne lines "x y z", %.17g)
//   qw_configs write <file> <dmc 0|1>     (walkers on stdin: "N Ne", then per walker its
//       weight and Ne*3 coordinates)
//       writes <file> with write_configurations<Dmc_point> (dmc = 1) or
//       write_configurations<Config_save_point> (dmc = 0), coordinates handed to the
//       reference's Config_save_point::read as "nElec Ne ndim 3 x y z ..." text.

#include "Qmc_std.h"
#include "qmc_io.h"
#include "Sample_point.h"
#include "Dmc_method.h"
#include <cstdio>
#include <sstream>

static Config_save_point config_of(int ne, const double* xyz) {
  std::stringstream ss;
  ss.precision(17);
  ss << "nElec " << ne << " ndim 3\n";
  for (int k = 0; k < 3 * ne; k++) ss << xyz[k] << " ";
  Config_save_point c;
  c.read(ss);
  return c;
}

int main(int argc, char** argv) {
  mpi_info.nprocs = 1;
  mpi_info.node = 0;
  if (argc == 4 && string(argv[1]) == "read") {
    string file = argv[2];
    int ne = atoi(argv[3]);
    Array1<Dmc_point> pts;
    read_configurations(file, pts);
    printf("nwalkers %d\n", pts.GetDim(0));
    Array1<doublevar> r(3);
    for (int i = 0; i < pts.GetDim(0); i++) {
      printf("walker %d weight %.17g sign %d\n", i, pts(i).weight, pts(i).sign);
      for (int e = 0; e < ne; e++) {
        pts(i).config_pos.getPos(e, r);
        printf("%.17g %.17g %.17g\n", r(0), r(1), r(2));
      }
    }
    return 0;
  }
  if (argc == 4 && string(argv[1]) == "write") {
    string file = argv[2];
    int dmc = atoi(argv[3]);
    int nw = 0, ne = 0;
    if (scanf("%d %d", &nw, &ne) != 2 || nw < 1 || ne < 1) {
      fprintf(stderr, "expected N Ne on stdin\n");
      return 2;
    }
    vector<double> w(nw), xyz((size_t)nw * ne * 3);
    for (int i = 0; i < nw; i++) {
      if (scanf("%lf", &w[i]) != 1) return 2;
      for (int k = 0; k < 3 * ne; k++)
        if (scanf("%lf", &xyz[(size_t)i * ne * 3 + k]) != 1) return 2;
    }
    if (dmc) {
      Array1<Dmc_point> pts(nw);
      for (int i = 0; i < nw; i++) {
        pts(i).config_pos = config_of(ne, &xyz[(size_t)i * ne * 3]);
        pts(i).weight = w[i];
        pts(i).sign = 1;
      }
      write_configurations(file, pts);
    } else {
      Array1<Config_save_point> pts(nw);
      for (int i = 0; i < nw; i++) pts(i) = config_of(ne, &xyz[(size_t)i * ne * 3]);
      write_configurations(file, pts);
    }
    return 0;
  }
  fprintf(stderr, "usage: qw_configs read <file> <Ne> | qw_configs write <file> <dmc> < walkers\n");
  return 2;
}
