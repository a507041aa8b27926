// The host-side owner of a capsgpu_system: the arrays it points into, published as plain pointers for the C ABI, and
// freed by capsgpu_system_free whichever entry point (a synthetic builder, the LAMMPS reader) made it.
#pragma once

#include <cstdint>
#include <set>
#include <vector>

#include "../../include/capsgpu_build.h"

namespace capsgpu {

struct Sys {
  capsgpu_system s{};
  std::vector<double> x, v, charge, mass, eps, sigma, bond_param, angle_param, fene_param;
  std::vector<int32_t> type, bond_atoms, angle_atoms, fene_atoms, excl_offset, excl_index;
  void publish() {
    s.x = x.data(), s.type = type.data(), s.charge = charge.data(), s.mass = mass.data(), s.eps = eps.data(), s.sigma = sigma.data();
    s.n_bonds = int32_t(bond_atoms.size() / 2), s.bond_atoms = bond_atoms.data(), s.bond_param = bond_param.data();
    s.n_angles = int32_t(angle_atoms.size() / 3), s.angle_atoms = angle_atoms.data(), s.angle_param = angle_param.data();
    s.n_fene = int32_t(fene_atoms.size() / 2), s.fene_atoms = fene_atoms.data(), s.fene_param = fene_param.data();
    s.excl_offset = excl_offset.data(), s.excl_index = excl_index.data();
    s.v = v.empty() ? nullptr : v.data();
    s.opaque = this;
  }
};

// 1-2 (and 1-3 when asked) partners from the harmonic and FENE bonds, sorted, as CSR
inline void exclusions(Sys& S, int n, bool with13) {
  std::vector<std::vector<int32_t>> nb(static_cast<size_t>(n));
  auto link = [&](const std::vector<int32_t>& pairs) {
    for (size_t b = 0; b + 1 < pairs.size(); b += 2) {
      const int32_t i = pairs[b], j = pairs[b + 1];
      nb[size_t(i)].push_back(j), nb[size_t(j)].push_back(i);
    }
  };
  link(S.bond_atoms);
  link(S.fene_atoms);
  S.excl_offset.assign(size_t(n) + 1, 0);
  S.excl_index.clear();
  for (int i = 0; i < n; ++i) {
    std::set<int32_t> p;
    for (int32_t j : nb[size_t(i)]) {
      p.insert(j);
      if (with13)
        for (int32_t k : nb[size_t(j)])
          if (k != i) p.insert(k);
    }
    p.erase(i);
    S.excl_index.insert(S.excl_index.end(), p.begin(), p.end());
    S.excl_offset[size_t(i) + 1] = int32_t(S.excl_index.size());
  }
  if (S.excl_index.empty()) S.excl_index.push_back(0);   // a valid pointer for an empty list
}

}  // namespace capsgpu
