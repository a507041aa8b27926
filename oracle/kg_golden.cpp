// Writes the bead positions of the reference's Kremer-Grest builder (core/src/kremer_grest.cpp:15-83) for given
// options as raw little-endian doubles (3 per bead) on stdout, to pin the draw order of the oracle's restatement.
// Test infrastructure only: built by oracle/Makefile from the reference sources where they lie.
//   kg_golden chains beads density k_theta bond seed > positions.bin
#include <cstdio>
#include <cstdlib>

#include "caps/kremer_grest.hpp"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: kg_golden chains beads density k_theta bond seed\n");
    return 2;
  }
  caps::KgOptions o;
  o.chains = std::atoi(argv[1]);
  o.beads = std::atoi(argv[2]);
  o.density = std::atof(argv[3]);
  o.k_theta = std::atof(argv[4]);
  o.bond = std::atof(argv[5]);
  o.seed = std::strtoull(argv[6], nullptr, 10);
  const caps::System s = caps::kremer_grest(o, nullptr);
  for (const auto& a : s.atoms) std::fwrite(a.pos.data(), sizeof(double), 3, stdout);
  return 0;
}
