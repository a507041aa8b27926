/* Oracle: random numbers, integrator pieces, the Kremer–Grest builder and the learned-RESPA MLP in FP64 —
 * TEST INFRASTRUCTURE ONLY (see caps_oracle.h).
 *
 * orc_rng_*: std::mt19937_64 (the engine the C++ standard specifies exactly: Matsumoto & Nishimura's MT19937-64) and
 *   the fixed distributions of core/include/caps/rng.hpp:26-147 (top-53-bit uniform, Marsaglia polar normal with the
 *   spare kept, Marsaglia–Tsang gamma). Pinned by the standard's known answer (the 10000th output of the default seed)
 *   and by moments of the distributions.
 * orc_kinetic_energy, orc_initial_velocities, orc_bussi_alpha: core/src/dynamics.cpp:65-69, 236-256, 466-477.
 * orc_kremer_grest: core/src/kremer_grest.cpp:15-83. Pinned bit for bit against the reference builder's positions
 *   (tests/golden/kg.npz) and by its own invariants (tests/test_provenance.cpp:315-345): every bond 0.97, no fold past
 *   120°, box ∛(MN/ρ).
 * orc_mlp_*: the 96-128-128-3 SiLU network of the learned r-RESPA (no reference counterpart), pinned by central finite
 *   differences of its loss.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "caps_oracle.h"

static const double kB = 0.0019872041, kAcc = 4.184e-4;

#define MT_N 312
#define MT_M 156
void orc_rng_seed(orc_rng* r, uint64_t seed) {
  r->mt[0] = seed;
  for (int i = 1; i < MT_N; ++i) r->mt[i] = 6364136223846793005ULL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 62)) + (uint64_t)i;
  r->mti = MT_N;
  r->have_spare = 0;
  r->spare = 0;
}

uint64_t orc_rng_next(orc_rng* r) {
  static const uint64_t mag[2] = {0ULL, 0xB5026F5AA96619E9ULL};
  const uint64_t um = 0xFFFFFFFF80000000ULL, lm = 0x7FFFFFFFULL;
  if (r->mti >= MT_N) {
    int i;
    for (i = 0; i < MT_N - MT_M; ++i) {
      const uint64_t y = (r->mt[i] & um) | (r->mt[i + 1] & lm);
      r->mt[i] = r->mt[i + MT_M] ^ (y >> 1) ^ mag[y & 1];
    }
    for (; i < MT_N - 1; ++i) {
      const uint64_t y = (r->mt[i] & um) | (r->mt[i + 1] & lm);
      r->mt[i] = r->mt[i + (MT_M - MT_N)] ^ (y >> 1) ^ mag[y & 1];
    }
    const uint64_t y = (r->mt[MT_N - 1] & um) | (r->mt[0] & lm);
    r->mt[MT_N - 1] = r->mt[MT_M - 1] ^ (y >> 1) ^ mag[y & 1];
    r->mti = 0;
  }
  uint64_t x = r->mt[r->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= x >> 43;
  return x;
}

double orc_rng_uniform01(orc_rng* r) { return (double)(orc_rng_next(r) >> 11) * 0x1.0p-53; }

static double polar(orc_rng* r, int32_t* have, double* spare) {
  if (*have) { *have = 0; return *spare; }
  double u, v, q;
  do {
    u = 2 * orc_rng_uniform01(r) - 1, v = 2 * orc_rng_uniform01(r) - 1, q = u * u + v * v;
  } while (q >= 1 || q == 0);
  const double f = sqrt(-2 * log(q) / q);
  *spare = v * f, *have = 1;
  return u * f;
}

double orc_rng_normal(orc_rng* r) { return polar(r, &r->have_spare, &r->spare); }

double orc_rng_gamma(orc_rng* r, double shape, double scale, int32_t* g_have, double* g_spare) {
  if (shape <= 0) return 0;
  if (shape < 1) {
    /* Γ(k) = Γ(k + 1) U^(1/k): the boosted distribution has its own normal generator */
    int32_t h2 = 0;
    double s2 = 0;
    double u;
    do u = orc_rng_uniform01(r); while (u == 0);
    /* u is drawn before the boosted gamma (rng.hpp:124-129) */
    return orc_rng_gamma(r, shape + 1, scale, &h2, &s2) * pow(u, 1.0 / shape);
  }
  const double d = shape - 1.0 / 3.0, c = 1.0 / sqrt(9 * d);
  for (;;) {
    double x, v;
    do {
      x = polar(r, g_have, g_spare);
      v = 1 + c * x;
    } while (v <= 0);
    v = v * v * v;
    double u;
    do u = orc_rng_uniform01(r); while (u == 0);
    if (u < 1 - 0.0331 * x * x * x * x || log(u) < 0.5 * x * x + d * (1 - v + log(v))) return d * v * scale;
  }
}

double orc_kinetic_energy(const double* v, const double* m, int32_t n) {
  double k = 0;
  for (int32_t i = 0; i < n; ++i) k += m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  return 0.5 * k / kAcc;
}

void orc_initial_velocities(orc_rng* r, const double* m, int32_t n, double temperature, double* v) {
  for (int32_t i = 0; i < n; ++i) {
    const double sd = sqrt(kB * temperature * kAcc / m[i]);
    for (int k = 0; k < 3; ++k) v[3 * i + k] = sd * orc_rng_normal(r);
  }
  double p[3] = {0, 0, 0}, mfree = 0;
  for (int32_t i = 0; i < n; ++i) {
    mfree += m[i];
    for (int k = 0; k < 3; ++k) p[k] += m[i] * v[3 * i + k];
  }
  for (int32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) v[3 * i + k] = v[3 * i + k] - p[k] / mfree;
  const double ndof = 3.0 * n - 3.0;
  const double t0 = 2 * orc_kinetic_energy(v, m, n) / (ndof * kB);
  if (t0 > 0)
    for (int32_t k = 0; k < 3 * n; ++k) v[k] *= sqrt(temperature / t0);
}

double orc_bussi_alpha(double k, double kt_target, double ndof, double c, double r1, double sum) {
  const double knew = k + (1 - c) * (kt_target * (r1 * r1 + sum) / ndof - k) + 2 * r1 * sqrt(c * (1 - c) * kt_target * k / ndof);
  return sqrt((knew > 0 ? knew : 0.0) / k);
}

double orc_kremer_grest(int32_t chains, int32_t beads, double density, double k_theta, double bond, uint64_t seed, double* x) {
  const double L = cbrt((double)chains * beads / density);
  orc_rng r;
  orc_rng_seed(&r, seed);
  size_t at = 0;
  for (int32_t c = 0; c < chains; ++c) {
    double p[3];
    for (int k = 0; k < 3; ++k) p[k] = orc_rng_uniform01(&r) * L;
    double prev[3] = {0, 0, 0};
    for (int32_t kk = 0; kk < beads; ++kk) {
      if (kk > 0) {
        double d[3];
        for (int tries = 0;; ++tries) {
          for (int k = 0; k < 3; ++k) d[k] = orc_rng_normal(&r);
          const double nn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          for (int k = 0; k < 3; ++k) d[k] *= 1 / nn;
          if (kk == 1) break;
          const double cosb = d[0] * prev[0] + d[1] * prev[1] + d[2] * prev[2];
          if (cosb < -0.5) continue;
          if (k_theta <= 0 || orc_rng_uniform01(&r) < exp(-k_theta * (1 - cosb)) || tries > 1000) break;
        }
        for (int k = 0; k < 3; ++k) p[k] = p[k] + d[k] * bond, prev[k] = d[k];
      }
      for (int k = 0; k < 3; ++k) x[3 * at + k] = p[k];
      ++at;
    }
  }
  return L;
}

/* ---- the MLP ------------------------------------------------------------------------------------------------------ */
static double silu(double z) { return z / (1 + exp(-z)); }
static double dsilu(double z) {
  const double s = 1 / (1 + exp(-z));
  return s * (1 + z * (1 - s));
}

void orc_mlp_forward(const double* w1, const double* b1, const double* w2, const double* b2, const double* w3,
                     const double* b3, const double* x, int32_t n, double* y) {
  double h1[128], h2[128];
  for (int32_t s = 0; s < n; ++s) {
    const double* xs = x + (size_t)s * 96;
    for (int o = 0; o < 128; ++o) {
      double z = b1[o];
      for (int i = 0; i < 96; ++i) z += w1[o * 96 + i] * xs[i];
      h1[o] = silu(z);
    }
    for (int o = 0; o < 128; ++o) {
      double z = b2[o];
      for (int i = 0; i < 128; ++i) z += w2[o * 128 + i] * h1[i];
      h2[o] = silu(z);
    }
    for (int o = 0; o < 3; ++o) {
      double z = b3[o];
      for (int i = 0; i < 128; ++i) z += w3[o * 128 + i] * h2[i];
      y[(size_t)s * 3 + o] = z;
    }
  }
}

double orc_mlp_backward(const double* w1, const double* b1, const double* w2, const double* b2, const double* w3,
                        const double* b3, const double* x, const double* t, int32_t n, double* gw1, double* gb1,
                        double* gw2, double* gb2, double* gw3, double* gb3) {
  double z1[128], h1[128], z2[128], h2[128], y[3], d3[3], d2[128], d1[128];
  double loss = 0;
  for (int32_t s = 0; s < n; ++s) {
    const double* xs = x + (size_t)s * 96;
    for (int o = 0; o < 128; ++o) {
      double z = b1[o];
      for (int i = 0; i < 96; ++i) z += w1[o * 96 + i] * xs[i];
      z1[o] = z, h1[o] = silu(z);
    }
    for (int o = 0; o < 128; ++o) {
      double z = b2[o];
      for (int i = 0; i < 128; ++i) z += w2[o * 128 + i] * h1[i];
      z2[o] = z, h2[o] = silu(z);
    }
    for (int o = 0; o < 3; ++o) {
      double z = b3[o];
      for (int i = 0; i < 128; ++i) z += w3[o * 128 + i] * h2[i];
      y[o] = z;
      const double r = z - t[(size_t)s * 3 + o];
      loss += 0.5 * r * r / n;
      d3[o] = r / n;
    }
    for (int o = 0; o < 3; ++o) {
      gb3[o] += d3[o];
      for (int i = 0; i < 128; ++i) gw3[o * 128 + i] += d3[o] * h2[i];
    }
    for (int i = 0; i < 128; ++i) {
      double g = 0;
      for (int o = 0; o < 3; ++o) g += w3[o * 128 + i] * d3[o];
      d2[i] = g * dsilu(z2[i]);
    }
    for (int o = 0; o < 128; ++o) {
      gb2[o] += d2[o];
      for (int i = 0; i < 128; ++i) gw2[o * 128 + i] += d2[o] * h1[i];
    }
    for (int i = 0; i < 128; ++i) {
      double g = 0;
      for (int o = 0; o < 128; ++o) g += w2[o * 128 + i] * d2[o];
      d1[i] = g * dsilu(z1[i]);
    }
    for (int o = 0; o < 128; ++o) {
      gb1[o] += d1[o];
      for (int i = 0; i < 96; ++i) gw1[o * 96 + i] += d1[o] * xs[i];
    }
  }
  return loss;
}

/* Learned r-RESPA descriptors (no reference counterpart): per atom i, over the atoms j within 8 Å (nearest image; the
 * cell is at least 24 Å wide), 24 hat functions h_k of r on centres k·8/23 (k = 0 … 23) times the cut-off
 * ½(1 + cos πr/8) and q_j: out[i][k] = Σ q_j f h_k, out[i][24 + 24c + k] = Σ q_j f h_k d_c / r (c = x, y, z). */
void orc_descriptors(const double* x, const double* q, int32_t n, const double cell[12], double* out) {
  const double *a = cell + 3, *b = cell + 6, *c = cell + 9;
  double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
  double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double v = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
  const double dlt = 8.0 / 23.0;
  for (int i = 0; i < n; ++i) {
    double* o = out + (size_t)i * 96;
    for (int k = 0; k < 96; ++k) o[k] = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      const double f[3] = {(d[0] * bc[0] + d[1] * bc[1] + d[2] * bc[2]) / v, (d[0] * ca[0] + d[1] * ca[1] + d[2] * ca[2]) / v,
                           (d[0] * ab[0] + d[1] * ab[1] + d[2] * ab[2]) / v};
      const double m[3] = {floor(f[0] + 0.5), floor(f[1] + 0.5), floor(f[2] + 0.5)};
      for (int k = 0; k < 3; ++k) d[k] -= m[0] * a[k] + m[1] * b[k] + m[2] * c[k];
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 >= 64.0 || r2 < 1e-6) continue;
      const double r = sqrt(r2), fc = 0.5 * (1 + cos(3.14159265358979323846 * r / 8.0));
      for (int k = 0; k < 24; ++k) {
        const double h = fmax(0.0, 1.0 - fabs(r - k * dlt) / dlt) * fc * q[j];
        if (h == 0) continue;
        o[k] += h;
        for (int cc = 0; cc < 3; ++cc) o[24 + 24 * cc + k] += h * d[cc] / r;
      }
    }
  }
}
