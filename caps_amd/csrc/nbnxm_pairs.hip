// The non-bonded cluster-pair kernel for gfx950: Lennard-Jones 12-6 or 9-6 with per-type factors, Ewald real space
// (or damped shifted force) by the rational fits of core/src/nbnxm_fit.hpp:8-16 (κ r_c ≤ 4) or Abramowitz & Stegun erfc
// beyond them, the excluded pairs' corrections in their lanes — the pair terms of core/src/nbnxm.cpp:408-567.
//
// Mapping: one 64-wide wavefront per i cluster of 4 atoms; the wave walks the cluster's pruned FULL list 16 entries at a
// time, one j atom per lane, and each lane takes its j atom against the 4 i atoms (i data is wave-uniform). Every
// cluster pair is in the full list from both sides, so a wave writes only its own i forces: no atomics, and the sums
// run in a fixed order (bit-reproducible). Energies and the virial are computed only when asked (energy steps).
// Virial: Σ d⊗f over pairs = Σ_slots x⊗F + ½ Σ_full S⊗f_(j←i), S = qa·a + qb·b + qc·c, kept as three vectors per
// lattice coefficient (9 floats per lane); the slot part is added in double by k_slot_virial.
#include "context.hpp"
#include "prof.hpp"

namespace capsgpu {
namespace {

// Ewald real space: erf(κr)/r = κ gV(u), 2κ³ dgV/du = κ³ gF(u), u = (κr)², as P(t)/Q(t), t = u/16 (nbnxm_fit.hpp:8-12)
__constant__ float cGvP[7] = {1.1283791379e+00f, 3.1401995010e+00f, 1.2811100603e+01f, 1.4467038445e+01f, 8.9639355099e+00f,
                              -2.0071435048e+00f, 3.2804742414e-01f};
__constant__ float cGvQ[5] = {1.0f, 8.1162471304e+00f, 2.9041673956e+01f, 5.7404020415e+01f, 5.9764272617e+01f};
__constant__ float cGfP[7] = {-7.5225272615e-01f, 1.2926832369e+00f, -5.1479940366e+00f, 5.0524604436e+00f, -4.3528567753e+00f,
                              2.0075107894e+00f, -3.9860971314e-01f};
__constant__ float cGfQ[5] = {1.0f, 7.8815427017e+00f, 2.7653014088e+01f, 5.3807163983e+01f, 5.6797690753e+01f};
constexpr float kErfcP = 0.3275911f;
constexpr float kA1 = 0.254829592f, kA2 = -0.284496736f, kA3 = 1.421413741f, kA4 = -1.453152027f, kA5 = 1.061405429f;

__device__ inline float horner7(const float* c, float t) {
  float p = c[6];
#pragma unroll
  for (int k = 5; k >= 0; --k) p = __builtin_fmaf(p, t, c[k]);
  return p;
}
__device__ inline float horner5(const float* c, float t) {
  float p = c[4];
#pragma unroll
  for (int k = 3; k >= 0; --k) p = __builtin_fmaf(p, t, c[k]);
  return p;
}
__device__ inline float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

struct PairArgs {
  int64_t ncl;
  const int64_t* pfirst;
  const uint2* pent;
  const uint16_t* pent14;
  const float4* xq;
  const float2* pp;
  const double* org;
  const double* qvec;
  const int8_t* qint8;
  float4* fnb;
  double* epart;
  int zero_code;
  float rc2, rc, rcn, rc6, shift;
  float kappa, kap3, ku, kc, e0, f0, kcoul, f14, epk, kap2;
};

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

template <int kLj, int kCoul, bool kE>
__global__ void __launch_bounds__(256) k_pairs(PairArgs P) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ci = xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (ci >= P.ncl) return;
  const int l = lane & 3, eo = lane >> 2;
  float4 xi[4];
  float2 pi[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    xi[a] = P.xq[4 * ci + a];
    pi[a] = P.pp[4 * ci + a];
    xi[a].w *= P.kcoul;
  }
  const double oi0 = P.org[3 * ci], oi1 = P.org[3 * ci + 1], oi2 = P.org[3 * ci + 2];
  float fx[4] = {0, 0, 0, 0}, fy[4] = {0, 0, 0, 0}, fz[4] = {0, 0, 0, 0};
  float ev = 0, ec = 0;
  float va[3] = {0, 0, 0}, vb[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
  const float twelve = kLj == 2 ? 9.0f : 12.0f;
  const int64_t end = P.pfirst[ci + 1];
  for (int64_t e0 = P.pfirst[ci]; e0 < end; e0 += 16) {
    const int64_t e = e0 + eo;
    const bool valid = e < end;
    const uint2 en = valid ? P.pent[e] : make_uint2(uint32_t(ci) | (uint32_t(P.zero_code) << 24), 0u);
    const uint32_t cj = en.x & 0xffffffu, code = en.x >> 24;
    const unsigned mask = en.y & 0xffffu, ex = en.y >> 16;
    // the shift: image plus the difference of the cluster origins, rounded once to float as the reference does
    const double* sq = P.qvec + 3 * code;
    const double* oj = P.org + 3 * size_t(cj);
    const float sx = float(sq[0] + oj[0] - oi0), sy = float(sq[1] + oj[1] - oi1), sz = float(sq[2] + oj[2] - oi2);
    float4 xj = P.xq[4 * size_t(cj) + l];
    const float2 pj = P.pp[4 * size_t(cj) + l];
    xj.x += sx, xj.y += sy, xj.z += sz;
    const bool any_ex = __ballot(ex != 0) != 0;
    unsigned ex14 = 0;
    if (any_ex && ex) ex14 = P.pent14[e];
    float fjx = 0, fjy = 0, fjz = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const unsigned bit = (mask >> (4 * a + l)) & 1u, xb = (ex >> (4 * a + l)) & 1u;
      const float dx = xj.x - xi[a].x, dy = xj.y - xi[a].y, dz = xj.z - xi[a].z;
      float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
      const bool inr = r2 < P.rc2;
      const bool m = bit && inr, mall = (bit | xb) && inr;
      r2 = mall ? r2 : 1.0f;
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float r2i = rinv * rinv, r6i = r2i * r2i * r2i;
      const float c12 = pi[a].x * pj.x, c6 = pi[a].y * pj.y;
      const float rn = kLj == 2 ? r6i * r2i * rinv : r6i * r6i;
      const float Aa = c12 * rn, Bb = c6 * r6i;
      float fr = __builtin_fmaf(-6.0f, Bb, twelve * Aa) * r2i;
      fr = m ? fr : 0.0f;
      if (kE) {
        float vlj = Aa - Bb;
        vlj -= P.shift * __builtin_fmaf(-c6, P.rc6, c12 * P.rcn);
        ev += m ? vlj : 0.0f;
      }
      // the weight of the bare 1/r: 1 for a pair that interacts, f14 for an excluded 1-4 pair, 0 for the others
      float wr = rinv;
      if (any_ex) wr = (m ? 1.0f : (((ex14 >> (4 * a + l)) & 1u) ? P.f14 : 0.0f)) * rinv;
      if (kCoul == 1 || kCoul == 2) {
        const float qq = xi[a].w * xj.w;
        const float t = P.ku * r2;
        const float Fc = __builtin_fmaf(P.kap3, fast_div(horner7(cGfP, t), horner5(cGfQ, t)), wr * r2i);
        if (kCoul == 1) {
          fr = __builtin_fmaf(qq, Fc, fr);
          if (kE) ec += mall ? qq * __builtin_fmaf(-P.kappa, fast_div(horner7(cGvP, t), horner5(cGvQ, t)), wr) : 0.0f;
        } else {
          fr = __builtin_fmaf(qq, __builtin_fmaf(-P.f0, rinv, Fc), fr);
          if (kE) {
            const float gv = __builtin_fmaf(-P.kappa, fast_div(horner7(cGvP, t), horner5(cGvQ, t)), wr) - P.e0;
            ec += mall ? qq * __builtin_fmaf(P.f0, r2 * rinv - P.rc, gv) : 0.0f;
          }
        }
      }
      if (kCoul == 3 || kCoul == 4) {
        const float qq = xi[a].w * xj.w;
        const float r = r2 * rinv;
        const float exv = __expf(P.kap2 * r2);
        const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(P.epk, r, 1.0f));
        const float poly = t * __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, kA5, kA4), kA3), kA2), kA1);
        const float E0 = poly * exv * rinv;
        const float Fc0 = __builtin_fmaf(P.kc, exv, E0) * r2i;
        const float E = E0 - (rinv - wr), Fc = Fc0 - (rinv - wr) * r2i;
        if (kCoul == 3) {
          if (kE) ec += mall ? qq * E : 0.0f;
          fr = __builtin_fmaf(qq, Fc, fr);
        } else {
          if (kE) ec += mall ? qq * __builtin_fmaf(P.f0, r - P.rc, E - P.e0) : 0.0f;
          fr = __builtin_fmaf(qq, __builtin_fmaf(-P.f0, rinv, Fc), fr);
        }
      }
      fr = mall ? fr : 0.0f;
      const float gx = fr * dx, gy = fr * dy, gz = fr * dz;   // force on j from i atom a
      fx[a] -= gx, fy[a] -= gy, fz[a] -= gz;
      if (kE) fjx += gx, fjy += gy, fjz += gz;
    }
    if (kE && code != unsigned(P.zero_code)) {
      const float qa = float(P.qint8[4 * code]), qb = float(P.qint8[4 * code + 1]), qc = float(P.qint8[4 * code + 2]);
      va[0] += qa * fjx, va[1] += qa * fjy, va[2] += qa * fjz;
      vb[0] += qb * fjx, vb[1] += qb * fjy, vb[2] += qb * fjz;
      vc[0] += qc * fjx, vc[1] += qc * fjy, vc[2] += qc * fjz;
    }
  }
  // the i forces: every lane holds part of each i atom's sum
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const float sx = wave_sum(fx[a]), sy = wave_sum(fy[a]), sz = wave_sum(fz[a]);
    if (lane == a) P.fnb[4 * ci + a] = make_float4(sx, sy, sz, 0.0f);
  }
  if (kE) {
    const float s[11] = {wave_sum(ev), wave_sum(ec), wave_sum(va[0]), wave_sum(va[1]), wave_sum(va[2]), wave_sum(vb[0]),
                         wave_sum(vb[1]),   wave_sum(vb[2]), wave_sum(vc[0]), wave_sum(vc[1]), wave_sum(vc[2])};
    if (lane < 11) P.epart[11 * ci + lane] = double(s[lane]);
  }
}

template <int kLj, int kCoul>
void launch_kE(const PairArgs& P, bool want, hipStream_t s, unsigned grid) {
  if (want) hipLaunchKernelGGL((k_pairs<kLj, kCoul, true>), dim3(grid), dim3(256), 0, s, P);
  else hipLaunchKernelGGL((k_pairs<kLj, kCoul, false>), dim3(grid), dim3(256), 0, s, P);
}
template <int kLj>
void launch_coul(const PairArgs& P, int coul, bool want, hipStream_t s, unsigned grid) {
  switch (coul) {
    case 0: launch_kE<kLj, 0>(P, want, s, grid); break;
    case 1: launch_kE<kLj, 1>(P, want, s, grid); break;
    case 2: launch_kE<kLj, 2>(P, want, s, grid); break;
    case 3: launch_kE<kLj, 3>(P, want, s, grid); break;
    default: launch_kE<kLj, 4>(P, want, s, grid); break;
  }
}

}  // namespace

void pair_forces(capsgpu_ctx& c, bool want) {
  DevList& L = c.L;
  const int64_t ncl = L.g.nclusters;
  if (!c.comb) throw Error(-1, "the GPU pair kernel takes Lennard-Jones coefficients that are products of per-type factors (geometric mixing)");
  if (ncl >= (int64_t(1) << 24)) throw Error(-1, "more than 2^24 clusters");
  const double rc = c.opt.cutoff, rc2 = rc * rc;
  PairArgs P{};
  P.ncl = ncl;
  P.pfirst = L.pfirst.p;
  P.pent = L.pent.p;
  P.pent14 = L.pent14.p;
  P.xq = L.xq.p;
  P.pp = L.pp.p;
  P.org = L.org.p;
  P.qvec = L.qvec.p;
  P.qint8 = L.qint8.p;
  P.fnb = L.fnb.p;
  P.epart = L.epart.p;
  P.zero_code = L.g.zero_code;
  P.rc2 = float(rc2);
  P.rc = float(rc);
  P.rcn = float(c.lj96 ? std::pow(rc, -9.0) : std::pow(rc, -12.0));
  P.rc6 = float(std::pow(rc, -6.0));
  P.shift = c.opt.tail ? 0.0f : 1.0f;
  // Coulomb: PME → Ewald real space, else DSF (field.cpp:1345-1350, 1819-1838)
  int coul = 0;
  double kappa = 0, e0 = 0, f0 = 0;
  if (c.opt.coulomb) {
    const bool pme = c.opt.electrostatics == 1;
    kappa = pme ? c.pme.beta : c.opt.dsf_alpha;
    if (!pme) {
      const double a = c.opt.dsf_alpha, P0 = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027,
                   A5 = 1.061405429;
      const double x = a * rc, ex2 = std::exp(-x * x), t = 1.0 / (1.0 + P0 * x);
      const double erfc_rc = t * (A1 + t * (A2 + t * (A3 + t * (A4 + t * A5)))) * ex2;
      e0 = erfc_rc / rc;
      f0 = erfc_rc / (rc * rc) + 2 * a / std::sqrt(kPi) * std::exp(-a * a * rc2) / rc;
    }
    const bool fit = float(kappa) * float(kappa) * float(rc2) <= 16.0f;
    coul = pme ? (fit ? 1 : 3) : (fit ? 2 : 4);
  }
  P.kappa = float(kappa);
  P.kap3 = float(kappa * kappa * kappa);
  P.ku = float(kappa * kappa / 16.0);
  P.kc = float(2 * kappa / std::sqrt(kPi));
  P.e0 = float(e0);
  P.f0 = float(f0);
  P.kcoul = float(kCoulomb);
  P.f14 = float(c.f14);
  P.epk = float(0.3275911 * kappa);
  P.kap2 = float(-kappa * kappa);
  if (want) L.epart.resize(11 * size_t(ncl));
  P.epart = L.epart.p;
  const unsigned grid = unsigned((ncl + 3) / 4);
  if (grid == 0) return;
  kt_begin(c, "k_pairs");
  if (c.lj96) launch_coul<2>(P, coul, want, c.stream, grid);
  else launch_coul<1>(P, coul, want, c.stream, grid);
  LAUNCHCHECK();
  kt_end(c);
}

}  // namespace capsgpu
