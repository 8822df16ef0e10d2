// sanitize_driver.cpp -- runs the oracle on case descriptions written by the test suite
// (raw MhMesh, MhEnvironment[m], MhOptions records, as laid out by mahyco_types.h) under
// AddressSanitizer and UndefinedBehaviorSanitizer.  TEST INFRASTRUCTURE ONLY: built by
// oracle/sanitize.mk for tests/test_oracle_sanitizers.py; no product path uses it.
//
//   sanitize_driver <case file> <cycles>   -> prints "iteration=<n> status=<s> time=<t>"
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mahyco_oracle.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s <case file> <cycles>\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  MhMesh mesh;
  int m = 0;
  MhOptions opt;
  bool ok = fread(&mesh, sizeof mesh, 1, f) == 1 && fread(&m, sizeof m, 1, f) == 1 && m >= 1 && m <= MH_MAX_ENV;
  std::vector<MhEnvironment> envs(ok ? m : 0);
  ok = ok && fread(envs.data(), sizeof(MhEnvironment), m, f) == (size_t)m && fread(&opt, sizeof opt, 1, f) == 1;
  fclose(f);
  if (!ok) return 2;
  oracle_set_threads(1);
  OracleCase* h = oracle_create(&mesh, envs.data(), m, &opt);
  if (!h) return 3;
  const int st = oracle_run(h, atoi(argv[2]));
  std::vector<double> rho(oracle_nb_cells(h));
  oracle_get_field(h, MH_F_DENSITY, -1, 0, rho.data());
  double sum = 0.;
  for (double r : rho) sum += r;
  printf("iteration=%d status=%d time=%.17g sum_rho=%.17g\n", oracle_iteration(h), st, oracle_time(h), sum);
  oracle_destroy(h);
  return 0;
}
