// Internal: the host-side random numbers of a dynamics run, the same stream as CAPS's (core/include/caps/rng.hpp:26-147,
// core/src/dynamics.cpp:207-256, 466-477): std::mt19937_64 (exactly specified by the C++ standard), the top 53 bits as a
// uniform double, Marsaglia's polar normal with its spare kept, and Marsaglia–Tsang's gamma with a normal generator of
// its own. Initial velocities and Bussi's draws come from here, in the reference's order; Langevin noise is drawn on
// the device (counter-based) instead.
#pragma once

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace capsgpu {

struct PolarNormal {
  bool have = false;
  double spare = 0;
  template <class U>
  double draw(U&& uniform) {
    if (have) {
      have = false;
      return spare;
    }
    double u, v, q;
    do {
      u = 2 * uniform() - 1, v = 2 * uniform() - 1, q = u * u + v * v;
    } while (q >= 1 || q == 0);
    const double f = std::sqrt(-2 * std::log(q) / q);
    spare = v * f, have = true;
    return u * f;
  }
};

struct HostRng {
  std::mt19937_64 eng;
  PolarNormal gauss;
  explicit HostRng(uint64_t seed) : eng(seed) {}
  double uniform01() { return double(eng() >> 11) * 0x1.0p-53; }
  double normal() {
    return gauss.draw([this] { return uniform01(); });
  }
  // shape k, scale θ; a fresh distribution per call (as the reference constructs one per Bussi step)
  double gamma(double k, double th) {
    PolarNormal own;
    return gamma_with(k, th, own);
  }

 private:
  double gamma_with(double k, double th, PolarNormal& own) {
    if (k <= 0) return 0;
    if (k < 1) {
      double u;
      do u = uniform01(); while (u == 0);
      PolarNormal up;
      return gamma_with(k + 1, th, up) * std::pow(u, 1.0 / k);
    }
    const double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9 * d);
    for (;;) {
      double x, v;
      do {
        x = own.draw([this] { return uniform01(); });
        v = 1 + c * x;
      } while (v <= 0);
      v = v * v * v;
      double u;
      do u = uniform01(); while (u == 0);
      if (u < 1 - 0.0331 * x * x * x * x || std::log(u) < 0.5 * x * x + d * (1 - v + std::log(v))) return d * v * th;
    }
  }
};

// Maxwell–Boltzmann velocities at T without centre-of-mass motion, rescaled to T exactly (dynamics.cpp:236-256)
inline void initial_velocities(HostRng& r, const std::vector<double>& m, double temperature, std::vector<double>& v) {
  constexpr double kB = 0.0019872041, kAcc = 4.184e-4;
  const size_t n = m.size();
  v.assign(3 * n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const double sd = std::sqrt(kB * temperature * kAcc / m[i]);
    for (int k = 0; k < 3; ++k) v[3 * i + k] = sd * r.normal();
  }
  double p[3] = {0, 0, 0}, mfree = 0;
  for (size_t i = 0; i < n; ++i) {
    mfree += m[i];
    for (int k = 0; k < 3; ++k) p[k] += m[i] * v[3 * i + k];
  }
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) v[3 * i + k] = v[3 * i + k] - p[k] / mfree;
  const double ndof = 3.0 * double(n) - 3.0;
  double ke = 0;
  for (size_t i = 0; i < n; ++i) ke += m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  ke = 0.5 * ke / kAcc;
  const double t0 = 2 * ke / (ndof * kB);
  if (t0 > 0)
    for (auto& q : v) q *= std::sqrt(temperature / t0);
}

}  // namespace capsgpu
