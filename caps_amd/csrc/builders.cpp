// Synthetic systems for the parity cases and the benchmark, with their flattened force fields (capsgpu_ff).
//
// capsgpu_build_kg: the Kremer–Grest melt of core/src/kremer_grest.cpp:15-124 — the same chains from the same seed
//   (std::mt19937_64 with CAPS's fixed uniform and normal distributions), FENE K = 30, R0 = 1.5 with WCA on every bond,
//   one type, ε = σ = 1, r_c = 2^(1/6), LJ shifted to zero there, 1-2 exclusions.
// capsgpu_build_melt: an all-atom polyethylene-like melt (no counterpart in the reference; the builders there are out of
//   this path): 128-atom chains CH3-(CH2)40-CH3 grown as self-avoiding walks of carbons (C-C 1.53 Å, 112°, torsions
//   near trans/gauche, the farthest of several trials), hydrogens at tetrahedral positions, six types with σ in
//   2.5-3.9 Å and ε in 0.03-0.15 kcal/mol (geometric mixing), harmonic bonds and angles, 1-2 and 1-3 exclusions,
//   charges drawn N(0, 0.25 e) and neutralised per chain, in a cubic box at the given density.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/capsgpu.h"
#include "../../include/capsgpu_build.h"
#include "host_rng.hpp"
#include "system.hpp"

namespace {
using capsgpu::exclusions;
using capsgpu::Sys;
thread_local std::string g_err;

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 unit(V3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }


// a cubic grid of the placed backbone carbons, for the self-avoidance test
struct Grid {
  double L, h;
  int m;
  std::vector<std::vector<V3>> cell;
  Grid(double L_, double h_) : L(L_) {
    m = std::max(1, int(L / h_));
    h = L / m;
    cell.resize(size_t(m) * m * m);
  }
  int idx(double v) const {
    int k = int(std::floor(v / h)) % m;
    return k < 0 ? k + m : k;
  }
  void add(V3 p) { cell[(size_t(idx(p.x)) * m + idx(p.y)) * m + idx(p.z)].push_back(p); }
  // the smallest distance from p to a stored point (minimum image), looking one cell around
  double closest(V3 p) const {
    double best = 1e30;
    const int ix = idx(p.x), iy = idx(p.y), iz = idx(p.z);
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          const auto& v = cell[(size_t((ix + a + m) % m) * m + (iy + b + m) % m) * m + (iz + c + m) % m];
          for (const V3& q : v) {
            V3 d = q - p;
            d.x -= L * std::round(d.x / L), d.y -= L * std::round(d.y / L), d.z -= L * std::round(d.z / L);
            best = std::min(best, dot(d, d));
          }
        }
    return std::sqrt(best);
  }
};

constexpr double kDeg = 3.14159265358979323846 / 180.0;

Sys* build_melt(int chains, uint64_t seed, double density) {
  const int per = 128, ncarb = 42;
  const int n = chains * per;
  auto* S = new Sys;
  const double mC = 12.011, mH = 1.008;
  const double mtot = chains * (ncarb * mC + (per - ncarb) * mH);
  const double L = std::cbrt(mtot / (density * 0.602214076));
  S->s.n_atoms = n;
  S->s.n_types = 6;
  double cell[12] = {0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L};
  std::memcpy(S->s.cell, cell, sizeof cell);
  // types: 0 C of CH3, 1 C next to CH3, 2 C inside, 3 H of CH3, 4 H next to CH3, 5 H inside
  const double sg[6] = {3.90, 3.70, 3.50, 2.60, 2.55, 2.50}, ep[6] = {0.150, 0.100, 0.066, 0.035, 0.032, 0.030};
  S->eps.resize(36), S->sigma.resize(36);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) S->eps[size_t(6 * a + b)] = std::sqrt(ep[a] * ep[b]), S->sigma[size_t(6 * a + b)] = std::sqrt(sg[a] * sg[b]);
  std::mt19937_64 eng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> G(0.0, 1.0);
  Grid grid(L, 3.0);
  const double bcc = 1.53, bch = 1.09, th_ccc = 112.0 * kDeg;
  S->x.resize(3 * size_t(n));
  S->type.resize(size_t(n));
  S->charge.resize(size_t(n));
  S->mass.resize(size_t(n));
  for (int ch = 0; ch < chains; ++ch) {
    std::vector<V3> C(static_cast<size_t>(ncarb));
    // the first carbon where it is farthest from those placed, of a few random tries
    double bestd = -1;
    for (int t = 0; t < 16; ++t) {
      const V3 p{U(eng) * L, U(eng) * L, U(eng) * L};
      const double d = grid.closest(p);
      if (d > bestd) bestd = d, C[0] = p;
    }
    for (int k = 1; k < ncarb; ++k) {
      V3 best{};
      double bd = -1;
      for (int t = 0; t < 24; ++t) {
        V3 dir;
        if (k == 1) {
          dir = unit(V3{G(eng), G(eng), G(eng)});
        } else {
          // bond angle θ at C[k-1], torsion φ about C[k-2]-C[k-1] (from C[k-3] when there is one)
          const V3 b1 = unit(C[size_t(k - 1)] - C[size_t(k - 2)]);
          V3 ref = k >= 3 ? C[size_t(k - 3)] - C[size_t(k - 2)] : V3{G(eng), G(eng), G(eng)};
          V3 nrm = cross(ref, b1);
          if (dot(nrm, nrm) < 1e-12) nrm = cross(V3{1, 0.3, 0.1}, b1);
          nrm = unit(nrm);
          const V3 m = cross(nrm, b1);
          const double r = U(eng);
          const double phi0 = r < 0.6 ? 180.0 : r < 0.8 ? 60.0 : -60.0;
          const double phi = (phi0 + 20.0 * (U(eng) - 0.5)) * kDeg;
          const double th = 3.14159265358979323846 - th_ccc;
          dir = b1 * std::cos(th) + (m * std::cos(phi) + nrm * std::sin(phi)) * std::sin(th);
        }
        const V3 p = C[size_t(k - 1)] + dir * bcc;
        double d = grid.closest(p);
        // not against its own chain's last carbons
        for (int q = 0; q < k - 2; ++q) {
          V3 dd = C[size_t(q)] - p;
          dd.x -= L * std::round(dd.x / L), dd.y -= L * std::round(dd.y / L), dd.z -= L * std::round(dd.z / L);
          d = std::min(d, std::sqrt(dot(dd, dd)));
        }
        if (d > bd) bd = d, best = p;
        if (d > 3.2) break;
      }
      C[size_t(k)] = best;
    }
    for (const V3& p : C) grid.add(p);
    // atoms of the chain in order: each carbon followed by its hydrogens
    const int base = ch * per;
    int at = base;
    std::vector<int> cidx(static_cast<size_t>(ncarb));
    std::vector<double> q(static_cast<size_t>(per));
    auto put = [&](V3 p, int t, double m) {
      S->x[3 * size_t(at)] = p.x, S->x[3 * size_t(at) + 1] = p.y, S->x[3 * size_t(at) + 2] = p.z;
      S->type[size_t(at)] = t;
      S->mass[size_t(at)] = m;
      return at++;
    };
    for (int k = 0; k < ncarb; ++k) {
      const bool end = k == 0 || k == ncarb - 1, near = k == 1 || k == ncarb - 2;
      cidx[size_t(k)] = put(C[size_t(k)], end ? 0 : near ? 1 : 2, mC);
      const int ht = end ? 3 : near ? 4 : 5;
      const V3 c = C[size_t(k)];
      if (!end) {
        const V3 u = unit(C[size_t(k - 1)] - c), w = unit(C[size_t(k + 1)] - c);
        const V3 bis = unit((u + w) * -1.0), nrm = unit(cross(u, w));
        const double a = 0.5 * 107.8 * kDeg;
        const int h1 = put(c + (bis * std::cos(a) + nrm * std::sin(a)) * bch, ht, mH);
        const int h2 = put(c + (bis * std::cos(a) - nrm * std::sin(a)) * bch, ht, mH);
        S->bond_atoms.insert(S->bond_atoms.end(), {cidx[size_t(k)], h1, cidx[size_t(k)], h2});
      } else {
        // three hydrogens around the axis away from the neighbouring carbon
        const V3 nb = C[size_t(k == 0 ? 1 : ncarb - 2)];
        const V3 ax = unit(c - nb);
        V3 p0 = cross(ax, V3{0.3, 1, 0.2});
        p0 = unit(p0);
        const V3 p1 = cross(ax, p0);
        const double tilt = (180.0 - 110.7) * kDeg;
        for (int j = 0; j < 3; ++j) {
          const double az = (120.0 * j) * kDeg;
          const V3 d = ax * std::cos(tilt) + (p0 * std::cos(az) + p1 * std::sin(az)) * std::sin(tilt);
          const int h = put(c + d * bch, ht, mH);
          S->bond_atoms.insert(S->bond_atoms.end(), {cidx[size_t(k)], h});
        }
      }
    }
    for (int k = 0; k + 1 < ncarb; ++k) S->bond_atoms.insert(S->bond_atoms.end(), {cidx[size_t(k)], cidx[size_t(k + 1)]});
    // charges N(0, 0.25 e), neutral chain
    double sum = 0;
    for (int a = 0; a < per; ++a) q[size_t(a)] = 0.25 * G(eng), sum += q[size_t(a)];
    for (int a = 0; a < per; ++a) S->charge[size_t(base + a)] = q[size_t(a)] - sum / per;
  }
  // bond parameters; angles from the bond graph: every pair of bonds at an atom
  std::vector<std::vector<int32_t>> nb(static_cast<size_t>(n));
  for (size_t b = 0; b < S->bond_atoms.size(); b += 2) {
    const int32_t i = S->bond_atoms[b], j = S->bond_atoms[b + 1];
    const bool cc = S->type[size_t(i)] <= 2 && S->type[size_t(j)] <= 2;
    S->bond_param.insert(S->bond_param.end(), {cc ? 268.0 : 340.0, cc ? 1.529 : 1.090});
    nb[size_t(i)].push_back(j), nb[size_t(j)].push_back(i);
  }
  for (int j = 0; j < n; ++j)
    for (size_t a = 0; a < nb[size_t(j)].size(); ++a)
      for (size_t b = a + 1; b < nb[size_t(j)].size(); ++b) {
        const int32_t i = nb[size_t(j)][a], k = nb[size_t(j)][b];
        const bool hi = S->type[size_t(i)] >= 3, hk = S->type[size_t(k)] >= 3;
        S->angle_atoms.insert(S->angle_atoms.end(), {i, j, k});
        if (!hi && !hk) S->angle_param.insert(S->angle_param.end(), {58.35, 112.7 * kDeg});
        else if (hi && hk) S->angle_param.insert(S->angle_param.end(), {33.0, 107.8 * kDeg});
        else S->angle_param.insert(S->angle_param.end(), {37.5, 110.7 * kDeg});
      }
  exclusions(*S, n, true);
  S->publish();
  return S;
}

// kremer_grest (kremer_grest.cpp:15-83) and kremer_grest_forcefield (:85-124), reduced units read as Å, kcal/mol, g/mol
Sys* build_kg(int chains, int beads, double density, double k_theta, double bond, uint64_t seed) {
  if (chains < 1 || beads < 2) throw std::invalid_argument("at least one chain of two beads");
  if (density <= 0.05 || density > 1.2) throw std::invalid_argument("reduced density ρσ³ between 0.05 and 1.2");
  const double L = std::cbrt(double(chains) * beads / density);
  auto* S = new Sys;
  const int n = chains * beads;
  S->s.n_atoms = n;
  S->s.n_types = 1;
  double cell[12] = {0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L};
  std::memcpy(S->s.cell, cell, sizeof cell);
  capsgpu::HostRng r(seed);
  auto unitv = [&] {
    V3 v;
    v.x = r.normal(), v.y = r.normal(), v.z = r.normal();
    return v * (1 / std::sqrt(dot(v, v)));
  };
  for (int c = 0; c < chains; ++c) {
    V3 p;
    p.x = r.uniform01() * L, p.y = r.uniform01() * L, p.z = r.uniform01() * L;
    V3 prev{0, 0, 0};
    for (int k = 0; k < beads; ++k) {
      if (k > 0) {
        V3 d;
        for (int tries = 0;; ++tries) {
          d = unitv();
          if (k == 1) break;
          const double cosb = dot(d, prev);
          if (cosb < -0.5) continue;
          if (k_theta <= 0 || r.uniform01() < std::exp(-k_theta * (1 - cosb)) || tries > 1000) break;
        }
        p = p + d * bond;
        prev = d;
      }
      S->x.insert(S->x.end(), {p.x, p.y, p.z});
      if (k > 0) {
        const int32_t a = int32_t(S->x.size() / 3) - 2;
        S->fene_atoms.insert(S->fene_atoms.end(), {a, a + 1});
        S->fene_param.insert(S->fene_param.end(), {30.0, 1.5, 1.0, 1.0});
      }
    }
  }
  S->type.assign(size_t(n), 0);
  S->charge.assign(size_t(n), 0.0);
  S->mass.assign(size_t(n), 1.0);
  S->eps = {1.0};
  S->sigma = {1.0};
  exclusions(*S, n, false);
  S->publish();
  return S;
}
}  // namespace

extern "C" {
const char* capsgpu_build_error(void) { return g_err.c_str(); }

int32_t capsgpu_build_melt(int32_t chains, uint64_t seed, double density, capsgpu_system** out) {
  try {
    if (chains < 1 || density <= 0) throw std::invalid_argument("chains and density must be positive");
    *out = &build_melt(chains, seed, density)->s;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int32_t capsgpu_build_kg(int32_t chains, int32_t beads, double density, double k_theta, double bond, uint64_t seed, capsgpu_system** out) {
  try {
    *out = &build_kg(chains, beads, density, k_theta, bond, seed)->s;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void capsgpu_system_free(capsgpu_system* s) {
  if (s) delete static_cast<Sys*>(s->opaque);
}
}
