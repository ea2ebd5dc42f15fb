// Test helper: writes the qwalk_amd driver's block log and per-block JSON for two made-up
// blocks (no GPU needed), so that their layout can be compared with the reference program's
// own files.  Built by tests/test_output_format.py.
#define main driver_main
#include "../mainline_amd/csrc/qwalk_amd_main.cpp"
#undef main

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::vector<double> eb(2 * QW_BLOCK_FIELDS, 0.0);
  for (int b = 0; b < 2; b++) {
    double* r = &eb[(size_t)b * QW_BLOCK_FIELDS];
    r[0] = -0.50208566 + 0.001 * b;  // E, kinetic, Coulomb, nonlocal
    r[1] = 0.3265;
    r[2] = -0.8286;
    r[5] = 80;  // points
    r[6] = 0.0019923765;  // variances
    r[7] = 0.15;
    r[8] = 0.138;
    r[10] = 0.9990838;  // mean weight and its variance
    r[11] = 3.8125e-06;
  }
  const std::string base = argv[1];
  write_log(base + ".log", "dmc", 0.02, true, eb, 2);
  write_json(base + ".json", eb, 2);
  return 0;
}
