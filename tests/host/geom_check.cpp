// Host check of the corner-vector forms in mh_geom.h: cqs_hex_comp<C> and cqs_hex_comp_fm<C>
// must give cqs_hex's component C bit for bit.  Reads hexahedra (24 doubles each, as
// hexadecimal bit patterns) on stdin, prints the number of mismatches.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include "../../mahyco_amd/csrc/mh_geom.h"

static uint64_t bits(double v) { uint64_t u; std::memcpy(&u, &v, 8); return u; }
static double from_bits(uint64_t u) { double v; std::memcpy(&v, &u, 8); return v; }

int main() {
  long cells = 0, bad = 0;
  for (;;) {
    uint64_t raw[24];
    int got = 0;
    for (int k = 0; k < 24; ++k) got += std::scanf("%" SCNx64, &raw[k]);
    if (got != 24) break;
    D3 x[8];
    for (int l = 0; l < 8; ++l) x[l] = d3(from_bits(raw[3 * l]), from_bits(raw[3 * l + 1]), from_bits(raw[3 * l + 2]));
    D3 q[8];
    cqs_hex(x, q);
    double qc[3][8], qf[3][8];
    cqs_hex_comp<0>(x, qc[0]); cqs_hex_comp<1>(x, qc[1]); cqs_hex_comp<2>(x, qc[2]);
    cqs_hex_comp_fm<0>(x, qf[0]); cqs_hex_comp_fm<1>(x, qf[1]); cqs_hex_comp_fm<2>(x, qf[2]);
    for (int l = 0; l < 8; ++l) {
      const double ref[3] = {q[l].x, q[l].y, q[l].z};
      for (int c = 0; c < 3; ++c)
        if (bits(ref[c]) != bits(qc[c][l]) || bits(ref[c]) != bits(qf[c][l])) ++bad;
    }
    ++cells;
  }
  std::printf("%ld %ld\n", cells, bad);
  return 0;
}
