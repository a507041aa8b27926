// Bonded terms, force summation, integrator steps and deterministic reductions on gfx950, all in FP64.
//
//  k_bonded     one thread per term: the term's force on each member atom into that member's slot, energy and
//               virial into per-block sums. Harmonic bond and angle (core/src/field.cpp:1980-2003), FENE + WCA
//               (:2150-2180).
//  k_sum_forces per atom: the pair, reciprocal and bonded forces; its bonded slots (contiguous, in the order of its
//               terms) added first (no atomics, a fixed order)
//  k_bonded_sum the two above in one launch by ranges of consecutive atoms, the slots kept in LDS (the same sums)
//  k_sum_forces the total force per atom: pair forces (by slot, float) + reciprocal (float) + bonded (double)
//  k_vv_open    velocity Verlet's opening half kick and drift (core/src/dynamics.cpp:405-431, 657-666), Bussi's
//               scaling of the previous step folded in
//  k_baoab_open Langevin BAOAB's B, A, O, A (dynamics.cpp:613-621), O with counter-based normal noise
//  k_kick       the closing half kick, with kinetic-energy partial sums on thermo steps
//  k_colsum     column sums of per-block partials in a fixed order (bit-reproducible energies and virials)
#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "pme_atoms.hpp"
#include "prof.hpp"
#include "step.hpp"

namespace capsgpu {
namespace {

struct CellArgs {
  double a[3], b[3], c[3], hi[3][3];
  double il[3];   // orthorhombic: 1 / the edge lengths
  int ortho;
};

// kOrtho: 1 the cell is orthorhombic, 0 it is not, −1 C.ortho says (a kernel for one kind of cell keeps fewer registers)
template <int kOrtho = -1>
__device__ inline void min_image(const CellArgs& C, double* d) {
  if (kOrtho == 1 || (kOrtho == -1 && C.ortho)) {
    const double L[3] = {C.a[0], C.b[1], C.c[2]};
    for (int k = 0; k < 3; ++k) d[k] -= L[k] * round(d[k] * C.il[k]);
    return;
  }
  // Cell::minimum_image (core/src/system.cpp:27-45)
  double f[3];
  for (int k = 0; k < 3; ++k) f[k] = C.hi[k][0] * d[0] + C.hi[k][1] * d[1] + C.hi[k][2] * d[2];
  for (int k = 0; k < 3; ++k) f[k] -= round(f[k]);
  double r[3];
  for (int k = 0; k < 3; ++k) r[k] = C.a[k] * f[0] + C.b[k] * f[1] + C.c[k] * f[2];
  double best = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], bt[3] = {r[0], r[1], r[2]};
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        double t[3];
        for (int m = 0; m < 3; ++m) t[m] = r[m] + C.a[m] * i + C.b[m] * j + C.c[m] * k;
        const double q = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        if (q < best - 1e-12) best = q, bt[0] = t[0], bt[1] = t[1], bt[2] = t[2];
      }
  d[0] = bt[0], d[1] = bt[1], d[2] = bt[2];
}

__device__ inline void vir_add(double* A, const double* d, const double* f) {
  A[2] += d[0] * f[0] + d[1] * f[1] + d[2] * f[2];
  A[3] += d[0] * f[0], A[4] += d[1] * f[1], A[5] += d[2] * f[2];
  A[6] += 0.5 * (d[0] * f[1] + d[1] * f[0]), A[7] += 0.5 * (d[0] * f[2] + d[2] * f[0]), A[8] += 0.5 * (d[1] * f[2] + d[2] * f[1]);
}

constexpr int kEB = 9;   // bond, angle, virial, w[6]

// term g's member atoms (two for a bond or FENE bond, three for an angle; the third −1 for a bond)
__device__ __forceinline__ int3 term_members(int g, int nb, int na, int nf, const int32_t* __restrict__ bij, const int32_t* __restrict__ aijk,
                                             const int32_t* __restrict__ fij) {
  if (g < nb) return make_int3(bij[2 * g], bij[2 * g + 1], -1);
  if (g < nb + na) return make_int3(aijk[3 * (g - nb)], aijk[3 * (g - nb) + 1], aijk[3 * (g - nb) + 2]);
  if (g < nb + na + nf) return make_int3(fij[2 * (g - nb - na)], fij[2 * (g - nb - na) + 1], -1);
  return make_int3(-1, -1, -1);
}

// The forces of term g (bonds [0, nb), then angles, then FENE bonds) with members m (term_members) on those members:
// put(place, fx, fy, fz) for each member, place being the member's place in the per-atom order of slots (bond t: member
// slots 2t, 2t + 1; angle t: 2nb + 3t + 0…2; FENE t: 2nb + 3na + 2t + 0…1); with kE and own, the term's energy and
// virial into A. No contraction of a·b + c into FMAs here: the energy variant's extra arithmetic must not change how
// the forces are rounded (forces are the same whether or not energies are asked for, and wherever the term is
// evaluated).
template <bool kE, int kOrtho, class Put>
__device__ __forceinline__ void term_forces(int g, int3 mem, bool own, int nb, int na, int nf, const CellArgs& C, const double* __restrict__ x,
                                            const double* __restrict__ bkr, const double* __restrict__ akt, const double* __restrict__ fp,
                                            const int32_t* __restrict__ spos, Put&& put, double* A, uint32_t* __restrict__ flags) {
#pragma clang fp contract(off)
  if (g < nb || (g >= nb + na && g < nb + na + nf)) {
    const bool fene = g >= nb;
    const int t = fene ? g - nb - na : g;
    const int32_t a = mem.x, b = mem.y;
    // the members' places, requested with the atoms (they depend on the term only)
    const size_t s0 = fene ? 2 * size_t(nb) + 3 * size_t(na) + 2 * size_t(t) : 2 * size_t(t);
    const int32_t p0 = spos[s0], p1 = spos[s0 + 1];
    double d[3] = {x[3 * b] - x[3 * a], x[3 * b + 1] - x[3 * a + 1], x[3 * b + 2] - x[3 * a + 2]};
    min_image<kOrtho>(C, d);
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double en = 0, dEdr = 0;
    bool ok = true;
    if (!fene) {
      const double k = bkr[2 * t], dr = r - bkr[2 * t + 1];
      en = k * dr * dr;
      dEdr = 2 * k * dr;
    } else {
      const double K = fp[4 * t], R0 = fp[4 * t + 1], eps = fp[4 * t + 2], sg = fp[4 * t + 3], q = r / R0;
      if (q >= 1) {
        atomicOr(&flags[2], 1u);
        ok = false;
      } else {
        en = -0.5 * K * R0 * R0 * log(1 - q * q);
        dEdr = K * r / (1 - q * q);
        if (eps > 0 && r < 1.122462048309373 * sg) {
          const double s2 = (sg / r) * (sg / r), s6 = s2 * s2 * s2;
          en += 4 * eps * (s6 * s6 - s6) + eps;
          dEdr += -24 * eps * (2 * s6 * s6 - s6) / r;
        }
      }
    }
    double fj[3] = {0, 0, 0};
    if (ok && r > 1e-12)
      for (int m = 0; m < 3; ++m) fj[m] = d[m] * (-dEdr / r);
    put(p0, -fj[0], -fj[1], -fj[2]);
    put(p1, fj[0], fj[1], fj[2]);
    if (kE && own && ok) {
      A[0] += en;
      vir_add(A, d, fj);
    }
    return;
  }
  if (g >= nb && g < nb + na) {
    const int t = g - nb;
    const int32_t ia = mem.x, ja = mem.y, ka = mem.z;
    const size_t s0 = 2 * size_t(nb) + 3 * size_t(t);
    const int32_t p0 = spos[s0], p1 = spos[s0 + 1], p2 = spos[s0 + 2];
    double u[3] = {x[3 * ia] - x[3 * ja], x[3 * ia + 1] - x[3 * ja + 1], x[3 * ia + 2] - x[3 * ja + 2]};
    double v[3] = {x[3 * ka] - x[3 * ja], x[3 * ka + 1] - x[3 * ja + 1], x[3 * ka + 2] - x[3 * ja + 2]};
    min_image<kOrtho>(C, u);
    min_image<kOrtho>(C, v);
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double ilu = 1 / lu, ilv = 1 / lv, iuv = ilu * ilv;
    double cs = (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) * iuv;
    cs = cs < -1 ? -1 : cs > 1 ? 1 : cs;
    const double th = acos(cs), sn = fmax(sqrt(1 - cs * cs), 1e-8);
    const double dth = th - akt[2 * t + 1], kt = akt[2 * t];
    const double gg = 2 * kt * dth / sn, cu = cs * ilu * ilu, cv = cs * ilv * ilv;
    double fi[3], fk[3];
    for (int m = 0; m < 3; ++m) {
      fi[m] = (v[m] * iuv - u[m] * cu) * gg;
      fk[m] = (u[m] * iuv - v[m] * cv) * gg;
    }
    put(p0, fi[0], fi[1], fi[2]);
    put(p1, -(fi[0] + fk[0]), -(fi[1] + fk[1]), -(fi[2] + fk[2]));
    put(p2, fk[0], fk[1], fk[2]);
    if (kE && own) {
      A[1] += kt * dth * dth;
      vir_add(A, u, fi);
      vir_add(A, v, fk);
    }
  }
}

// the sums of kEB partials over a 256-thread workgroup (red holds kEB·256 doubles), written to out by the first threads
__device__ __forceinline__ void block_sums(const double* A, double* red, double* __restrict__ out) {
  for (int q = 0; q < kEB; ++q) red[q * 256 + threadIdx.x] = A[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < kEB; ++q) red[q * 256 + threadIdx.x] += red[q * 256 + threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < kEB) out[threadIdx.x] = red[threadIdx.x * 256];
}

// one thread per term: the term's force on each member atom into that member's slot, at the member's place spos[member]
// in the per-atom order of slots, energies and virial into per-block sums. Every term is computed once; k_sum_forces
// adds each atom's slots, which lie together, in a fixed order. (Used when the atoms cannot be cut into ranges whose
// slots fit a workgroup's LDS; k_bonded_sum otherwise.)
template <bool kE>
__device__ __forceinline__ void b_bonded(int bx, int gx, int nb, int na, int nf, CellArgs C, const double* __restrict__ x,
                                                const int32_t* __restrict__ bij, const double* __restrict__ bkr,
                                                const int32_t* __restrict__ aijk, const double* __restrict__ akt,
                                                const int32_t* __restrict__ fij, const double* __restrict__ fp, const int32_t* __restrict__ spos,
                                                double* __restrict__ slot, double* __restrict__ epart, uint32_t* __restrict__ flags) {
  const int g = bx * blockDim.x + threadIdx.x;
  double A[kEB] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  term_forces<kE, -1>(g, term_members(g, nb, na, nf, bij, aijk, fij), true, nb, na, nf, C, x, bkr, akt, fp, spos,
                  [&](int32_t p, double f0, double f1, double f2) {
                    double* o = slot + 3 * size_t(p);
                    o[0] = f0, o[1] = f1, o[2] = f2;
                  },
                  A, flags);
  if (kE) {
    __shared__ double red[kEB * 256];
    block_sums(A, red, epart + kEB * size_t(bx));
  }
}
template <bool kE>
__global__ void __launch_bounds__(256) k_bonded(int nb, int na, int nf, CellArgs C, const double* __restrict__ x,
                                                const int32_t* __restrict__ bij, const double* __restrict__ bkr,
                                                const int32_t* __restrict__ aijk, const double* __restrict__ akt,
                                                const int32_t* __restrict__ fij, const double* __restrict__ fp, const int32_t* __restrict__ spos,
                                                double* __restrict__ slot, double* __restrict__ epart, uint32_t* __restrict__ flags) {
  b_bonded<kE>(blockIdx.x, gridDim.x, nb, na, nf, C, x, bij, bkr, aijk, akt, fij, fp, spos, slot, epart, flags);
}

// per atom: the pair forces (by slot of the cluster list), the reciprocal forces and the bonded forces — the atom's
// bonded slots aoff[i] … aoff[i+1] (contiguous) summed first, in the order of its terms (aoff null: no bonded part)
__device__ __forceinline__ void b_sum_forces(int bx, int gx, int n, const uint32_t* __restrict__ iperm, const float4* __restrict__ fnb,
                                                    const float* __restrict__ frec, const int32_t* __restrict__ aoff,
                                                    const double* __restrict__ slot, double* __restrict__ f) {
  const int i = bx * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // every load first (the pair and reciprocal forces, then the slots four at a time), the sums in the fixed order
  const float4 p = fnb ? fnb[iperm[i]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float pr[3] = {0.0f, 0.0f, 0.0f};
  if (frec)
    for (int m = 0; m < 3; ++m) pr[m] = frec[3 * i + m];
  double Fb[3] = {0, 0, 0};
  if (aoff) {
    const int32_t e0 = aoff[i], e1 = aoff[i + 1];
    int32_t e = e0;
    for (; e + 4 <= e1; e += 4) {
      double v[12];
#pragma unroll
      for (int u = 0; u < 12; ++u) v[u] = slot[3 * size_t(e) + u];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        for (int m = 0; m < 3; ++m) Fb[m] += v[3 * u + m];
    }
    for (; e < e1; ++e)
      for (int m = 0; m < 3; ++m) Fb[m] += slot[3 * size_t(e) + m];
  }
  double F[3] = {0, 0, 0};
  if (fnb) F[0] += double(p.x), F[1] += double(p.y), F[2] += double(p.z);
  if (frec)
    for (int m = 0; m < 3; ++m) F[m] += double(pr[m]);
  if (aoff)
    for (int m = 0; m < 3; ++m) F[m] += Fb[m];
  for (int m = 0; m < 3; ++m) f[3 * i + m] = F[m];
}
__global__ void __launch_bounds__(256) k_sum_forces(int n, const uint32_t* __restrict__ iperm, const float4* __restrict__ fnb,
                                                    const float* __restrict__ frec, const int32_t* __restrict__ aoff,
                                                    const double* __restrict__ slot, double* __restrict__ f) {
  b_sum_forces(blockIdx.x, gridDim.x, n, iperm, fnb, frec, aoff, slot, f);
}

// The bonded terms and the force sum in one launch, by ranges of consecutive atoms (cut on the host so that a range's
// slots fit in LDS): a workgroup evaluates every term with a member in its range (a term reaching two ranges is
// evaluated in both, with the same arithmetic and so the same values; its energy is counted by the range of its first
// member), keeps the forces on its own atoms in LDS at their places in the per-atom order of slots, then adds each
// atom's slots in that order and writes the atom's total force: the sums of k_bonded + k_sum_forces, bit for bit,
// without the slots' round trip through device memory.
// (one wave per range of 64 atoms measured slower: 18.0 -> 22.4 us on the 262,144-atom melt)
constexpr int kRangeAtoms = 256, kRangeSlots = 1024, kRangeThreads = 256;   // 25 KB of slots per workgroup
// slot q's three doubles in LDS: 3q, with one double skipped after every eighth slot, so neighbouring atoms' runs of
// slots (a few slots apart) start in different banks
__device__ __forceinline__ int slot_at(int q) { return 3 * q + (q >> 3); }
struct BondSumArgs {
  int nb, na, nf;
  CellArgs C;
  const double* x;
  const int32_t *bij, *aijk, *fij, *spos, *aoff;
  const double *bkr, *akt, *fp;
  const int32_t *rstart, *roff;   // range r: atoms rstart[r] … rstart[r+1] − 1, terms rterm[roff[r]] …
  const int4* rterm;              // (term | top bit when its energy is counted here, its members as term_members)
  const uint32_t* iperm;          // the pair forces by slot of the cluster list (null: none)
  const float4* fnb;
  const float* frec;              // the reciprocal forces (null: none)
  double* f;
  double* epart;                  // kEB per range
  uint32_t* flags;
  int bonded;                     // 0: no bonded part in this evaluation (the force sum alone)
};

template <bool kE, int kOrtho>
__device__ __forceinline__ void b_bonded_sum(int r, const BondSumArgs& a) {
  __shared__ double sl[3 * kRangeSlots + kRangeSlots / 8];   // the range's slots, by place − its first place; then the energy sums
  const int r0 = a.rstart[r], r1 = a.rstart[r + 1];
  const int s0 = a.bonded ? a.aoff[r0] : 0, ns = a.bonded ? a.aoff[r1] - s0 : 0;
  double A[kEB] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (a.bonded) {
    // each term's record (term and members) requested one term ahead, so its load passes under the previous term's work
    const int k1 = a.roff[r + 1];
    int k = a.roff[r] + int(threadIdx.x);
    int4 rec = k < k1 ? a.rterm[k] : make_int4(-1, -1, -1, -1);
#pragma unroll 1
    for (; k < k1; k += blockDim.x) {
      const int4 cur = rec;
      rec = k + int(blockDim.x) < k1 ? a.rterm[k + blockDim.x] : make_int4(-1, -1, -1, -1);
      const uint32_t code = uint32_t(cur.x);
      term_forces<kE, kOrtho>(int(code & 0x7fffffffu), make_int3(cur.y, cur.z, cur.w), (code >> 31) != 0, a.nb, a.na, a.nf, a.C, a.x,
                      a.bkr, a.akt, a.fp, a.spos,
                      [&](int32_t p, double f0, double f1, double f2) {   // the range's own members only
                        const int q = p - s0;
                        if (q >= 0 && q < ns) sl[slot_at(q)] = f0, sl[slot_at(q) + 1] = f1, sl[slot_at(q) + 2] = f2;
                      },
                      A, a.flags);
    }
  }
  __syncthreads();
  // per atom, as k_sum_forces: the pair and reciprocal forces, the bonded slots in order
  for (int i = r0 + int(threadIdx.x); i < r1; i += blockDim.x) {
    const float4 pq = a.fnb ? a.fnb[a.iperm[i]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float pr[3] = {0.0f, 0.0f, 0.0f};
    if (a.frec)
      for (int m = 0; m < 3; ++m) pr[m] = a.frec[3 * i + m];
    double Fb[3] = {0, 0, 0};
    if (a.bonded)
      for (int e = a.aoff[i] - s0, e1 = a.aoff[i + 1] - s0; e < e1; ++e)
        for (int m = 0; m < 3; ++m) Fb[m] += sl[slot_at(e) + m];
    double F[3] = {0, 0, 0};
    if (a.fnb) F[0] += double(pq.x), F[1] += double(pq.y), F[2] += double(pq.z);
    if (a.frec)
      for (int m = 0; m < 3; ++m) F[m] += double(pr[m]);
    if (a.bonded)
      for (int m = 0; m < 3; ++m) F[m] += Fb[m];
    for (int m = 0; m < 3; ++m) a.f[3 * size_t(i) + m] = F[m];
  }
  if (kE) {
    static_assert(3 * kRangeSlots >= kEB * kRangeThreads && kRangeThreads == 256, "the energy sums reuse the slots' space");
    __syncthreads();
    block_sums(A, sl, a.epart + kEB * size_t(r));
  }
}
template <bool kE, int kOrtho>
__global__ void __launch_bounds__(kRangeThreads) k_bonded_sum(BondSumArgs a) {
  b_bonded_sum<kE, kOrtho>(int(blockIdx.x), a);
}
// the force-only variant for orthorhombic cells (every step of most runs): held to 80 registers, so six workgroups
// share a CU as their LDS allows, without spills
__global__ void __launch_bounds__(kRangeThreads, 6) k_bonded_sum_ortho(BondSumArgs a) {
  b_bonded_sum<false, 1>(int(blockIdx.x), a);
}

// Σ_slots x ⊗ F of the pair forces, x the listed positions X + origin (core/src/nbnxm.cpp:585-592)
__global__ void __launch_bounds__(256) k_slot_virial(int64_t ns, const float4* __restrict__ xq, const double* __restrict__ org,
                                                     const float4* __restrict__ fnb, double* __restrict__ part) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < ns) {
    const float4 p = xq[s], F = fnb[s];
    const double* o = org + 3 * (s / 4);
    const double r[3] = {double(p.x) + o[0], double(p.y) + o[1], double(p.z) + o[2]};
    const double f[3] = {double(F.x), double(F.y), double(F.z)};
    for (int u = 0; u < 3; ++u)
      for (int w = 0; w < 3; ++w) W[3 * u + w] = r[u] * f[w];
  }
  __shared__ double red[9][256];
  for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = W[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 9; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 9) part[9 * size_t(blockIdx.x) + threadIdx.x] = red[threadIdx.x][0];
}

// column sums of rows × cols partials, each column by one workgroup in a fixed order
__global__ void __launch_bounds__(256) k_colsum(const double* __restrict__ part, int64_t rows, int cols, double* __restrict__ out) {
  const int c = blockIdx.x;
  double s = 0;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += part[r * cols + c];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// largest squared displacement against the positions of the last build (flag 0) and the last prune (flag 1), in
// float-ordered bits (non-negative floats compare as their bit patterns)
// flags[0] / flags[1]: some atom moved more than √lim_b / √lim_p since the list's build / prune (no check when the
// limit is negative: no list); flags[3]: a position is not a number. One atomic per wave that finds something, and
// none once the flag reads as set.
__device__ inline void move_flags(bool valid, const double* x, const double* x0, const double* x0p, int i, double lim_b, double lim_p,
                                  uint32_t* flags) {
  bool mb = false, mp = false, bad = false;
  if (valid) {
    if (lim_b >= 0.0) {
      double d0 = 0, d1 = 0;
      for (int m = 0; m < 3; ++m) {
        const double a = x[3 * i + m] - x0[3 * i + m], b = x[3 * i + m] - x0p[3 * i + m];
        d0 += a * a, d1 += b * b;
      }
      mb = d0 > lim_b, mp = d1 > lim_p;
      bad = !isfinite(d0);
    } else {
      bad = !isfinite(x[3 * i] + x[3 * i + 1] + x[3 * i + 2]);
    }
  }
  const uint64_t bm = __ballot(mb), pm = __ballot(mp), nm = __ballot(bad);
  if ((threadIdx.x & 63) == 0) {
    if (bm && !flags[0]) atomicOr(&flags[0], 1u);
    if (pm && !flags[1]) atomicOr(&flags[1], 1u);
    if (nm && !flags[3]) atomicOr(&flags[3], 1u);
  }
}

__global__ void k_check_moves(int n, const double* __restrict__ x, const double* __restrict__ x0, const double* __restrict__ x0p,
                              double lim_b, double lim_p, uint32_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  move_flags(i < n, x, x0, x0p, i, lim_b, lim_p, flags);
}

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC 2011): four 32-bit words from a counter and a key
__device__ inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c[0], p1 = uint64_t(0xCD9E8D57u) * c[2];
    const uint32_t h0 = uint32_t(p0 >> 32), l0 = uint32_t(p0), h1 = uint32_t(p1 >> 32), l1 = uint32_t(p1);
    c[0] = h1 ^ c[1] ^ k0, c[1] = l1, c[2] = h0 ^ c[3] ^ k1, c[3] = l0;
    k0 += 0x9E3779B9u, k1 += 0xBB67AE85u;
  }
}

// three standard normals for (atom, step) by Box–Muller on Philox uniforms, in FP32 with the hardware's log2, sine and
// cosine (argument in turns) and square root: u = (top 23 bits + ½) / 2^23 is exact in FP32 and lies in (0, 1), so every
// log is finite; the normals are good to about 1e-6, far below what a thermostat's noise needs
__device__ inline void normals3(uint64_t seed, uint64_t step, uint32_t atom, double* g) {
  uint32_t c[4] = {atom, uint32_t(step), uint32_t(step >> 32), 0x43415053u};
  philox(c, uint32_t(seed), uint32_t(seed >> 32));
  float u[4];
  for (int k = 0; k < 4; ++k) u[k] = (float(c[k] >> 9) + 0.5f) * 1.1920928955078125e-7f;
  const float ln2x2 = -1.3862943611198906f;   // −2 ln 2: −2 ln u = −2 ln 2 · log2 u
  const float r0 = __builtin_amdgcn_sqrtf(ln2x2 * __builtin_amdgcn_logf(u[0]));
  const float r1 = __builtin_amdgcn_sqrtf(ln2x2 * __builtin_amdgcn_logf(u[2]));
  g[0] = double(r0 * __builtin_amdgcn_cosf(u[1]));
  g[1] = double(r0 * __builtin_amdgcn_sinf(u[1]));
  g[2] = double(r1 * __builtin_amdgcn_cosf(u[3]));
}

// the workgroup's sums of v into its row of part (kAdd: added to prior, what the row held, read by the caller when it
// started so the load's latency passes under its work): each wave's sums by a butterfly, then the waves' in order
template <int kParts, bool kAdd = false>
__device__ inline void block_sum_store(double (&v)[kParts], double* part, int row, double prior = 0.0) {
  __shared__ double red[kParts][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = int(blockDim.x + 63) >> 6;
  for (int q = 0; q < kParts; ++q) {
    double s = v[q];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < kParts) {
    double t = red[threadIdx.x][0];
    for (int w = 1; w < nw; ++w) t += red[threadIdx.x][w];
    part[kParts * size_t(row) + threadIdx.x] = kAdd ? prior + t : t;
  }
}

// atom i's velocity with the put-off half kicks applied, in order: the sums of as many k_kick launches
// The integrator's arithmetic is written out (fma where a product is added, contraction off otherwise), so a kick
// gives the same bits in whichever launch it runs: alone, put off into the next launch, or in a grouped launch.
__device__ __forceinline__ void kicked_v(const Kicks& K, const double* __restrict__ minv, const double* __restrict__ v, int i, double* w) {
#pragma clang fp contract(off)
  for (int m = 0; m < 3; ++m) w[m] = v[3 * i + m];
#pragma unroll
  for (int j = 0; j < Kicks::kMax; ++j)
    if (j < K.n) {
      const double a = K.hk[j] * minv[i];
      for (int m = 0; m < 3; ++m) w[m] = fma(a, K.f[j][3 * i + m], w[m]);
    }
}

// the put-off half kicks, then v ← α v (Bussi, the previous step) and v += h·kAcc/m F; x += dt v; with sp.wq set, the
// per-atom half of the next PME evaluation's bookkeeping on the new positions (its splines launch is then left out)
__global__ void __launch_bounds__(256) k_vv_open(int n, double hk, double dt, const double* __restrict__ minv, const double* __restrict__ alpha,
                                                 const double* __restrict__ f, double* __restrict__ x, double* __restrict__ v,
                                                 const double* __restrict__ x0, const double* __restrict__ x0p, double lim_b, double lim_p,
                                                 uint32_t* __restrict__ flags, Kicks pre, AtomSplines sp) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double xn[3] = {0, 0, 0};
  if (i < n) {
    const double al = alpha ? *alpha : 1.0, a = hk * minv[i];
    double w[3];
    kicked_v(pre, minv, v, i, w);
    for (int m = 0; m < 3; ++m) {
      const double vv = fma(a, f[3 * i + m], al * w[m]);
      v[3 * i + m] = vv;
      xn[m] = fma(dt, vv, x[3 * i + m]);
      x[3 * i + m] = xn[m];
    }
  }
  move_flags(i < n, x, x0, x0p, i, lim_b, lim_p, flags);
  if (sp.wq) spline_atom(sp, i < n, i, xn);
}

// the put-off half kicks, then B (half kick), A (half drift), O, A; partials: kinetic energy before and after O (the
// bath's share)
__device__ __forceinline__ void b_baoab_open(int bx, int gx, int n, double hk, double hdt, double c1, double c2, const double* __restrict__ minv,
                                                    const double* __restrict__ sdv, uint64_t seed, const uint64_t* __restrict__ stepc,
                                                    uint64_t step_add, const double* __restrict__ f,
                                                    double* __restrict__ x, double* __restrict__ v, const double* __restrict__ x0,
                                                    const double* __restrict__ x0p, double lim_b, double lim_p, uint32_t* __restrict__ flags,
                                                    double* __restrict__ part, const Kicks& pre, const AtomSplines* sp = nullptr) {
#pragma clang fp contract(off)
  const int i = bx * blockDim.x + threadIdx.x;
  const uint64_t step = (stepc ? *stepc : 0) + step_add;
  const double prior = threadIdx.x == 0 ? part[bx] : 0.0;   // the row the bath's share is added to
  double ke[1] = {0}, xn[3] = {0, 0, 0};
  if (i < n) {
    const double a = hk * minv[i], mi = 1.0 / minv[i], cs = c2 * sdv[i];
    double g[3], w[3];
    normals3(seed, step, uint32_t(i), g);
    kicked_v(pre, minv, v, i, w);
    double k0 = 0, k1 = 0;
    for (int m = 0; m < 3; ++m) {
      double vv = fma(a, f[3 * i + m], w[m]);
      double xx = fma(hdt, vv, x[3 * i + m]);
      k0 = fma(vv, vv, k0);
      vv = fma(c1, vv, cs * g[m]);
      k1 = fma(vv, vv, k1);
      xx = fma(hdt, vv, xx);
      v[3 * i + m] = vv;
      x[3 * i + m] = xx;
      xn[m] = xx;
    }
    ke[0] = 0.5 * mi * (k1 - k0) / kAcc;
  }
  move_flags(i < n, x, x0, x0p, i, lim_b, lim_p, flags);
  if (sp && sp->wq) spline_atom(*sp, i < n, i, xn);
  block_sum_store<1, true>(ke, part, bx, prior);
}
__global__ void __launch_bounds__(256) k_baoab_open(int n, double hk, double hdt, double c1, double c2, const double* __restrict__ minv,
                                                    const double* __restrict__ sdv, uint64_t seed, const uint64_t* __restrict__ stepc,
                                                    uint64_t step_add, const double* __restrict__ f,
                                                    double* __restrict__ x, double* __restrict__ v, const double* __restrict__ x0,
                                                    const double* __restrict__ x0p, double lim_b, double lim_p, uint32_t* __restrict__ flags,
                                                    double* __restrict__ part, Kicks pre, AtomSplines sp) {
  b_baoab_open(blockIdx.x, gridDim.x, n, hk, hdt, c1, c2, minv, sdv, seed, stepc, step_add, f, x, v, x0, x0p, lim_b, lim_p, flags, part, pre,
               &sp);
}

// the put-off half kicks and the closing half kick; kinetic energy partials (Σ ½ m v², and the tensor m v_a v_b) when
// asked
template <bool kKE>
__device__ __forceinline__ void b_kick(int bx, int gx, int n, double hk, const double* __restrict__ minv, const double* __restrict__ f,
                                              double* __restrict__ v, double* __restrict__ part, const Kicks& pre) {
#pragma clang fp contract(off)
  const int i = bx * blockDim.x + threadIdx.x;
  double k[7] = {0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const double a = hk * minv[i];
    double vv[3], w[3];
    kicked_v(pre, minv, v, i, w);
    for (int m = 0; m < 3; ++m) vv[m] = fma(a, f[3 * i + m], w[m]), v[3 * i + m] = vv[m];
    if (kKE) {
      const double mi = 1.0 / minv[i];
      k[0] = 0.5 * mi * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]) / kAcc;
      k[1] = mi * vv[0] * vv[0], k[2] = mi * vv[1] * vv[1], k[3] = mi * vv[2] * vv[2];
      k[4] = mi * vv[0] * vv[1], k[5] = mi * vv[0] * vv[2], k[6] = mi * vv[1] * vv[2];
    }
  }
  if (kKE) block_sum_store<7>(k, part, bx);
}
template <bool kKE>
__global__ void __launch_bounds__(256) k_kick(int n, double hk, const double* __restrict__ minv, const double* __restrict__ f,
                                              double* __restrict__ v, double* __restrict__ part, Kicks pre) {
  b_kick<kKE>(blockIdx.x, gridDim.x, n, hk, minv, f, v, part, pre);
}

// Bussi's α from the reduced kinetic energy (dynamics.cpp:466-477); r1 and Σ R² drawn on the host in the reference's
// order. bath −= K_new − K.
__global__ void k_bussi(const double* __restrict__ part, int64_t rows, double kt_target, double ndof, double c, double r1, double sum,
                        double* __restrict__ alpha, double* __restrict__ bath) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += part[7 * r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double k = red[0];
    double al = 1.0;
    if (k > 0) {
      const double knew = k + (1 - c) * (kt_target * (r1 * r1 + sum) / ndof - k) + 2 * r1 * sqrt(c * (1 - c) * kt_target * k / ndof);
      al = sqrt(fmax(knew, 0.0) / k);
      *bath -= al * al * k - k;
    }
    *alpha = al;
  }
}

// the barostat's isotropic scaling: positions about the cell origin by mu, velocities by vf (dynamics.cpp:558-564, 716-719)
__global__ void k_scale_xv(int n, double o0, double o1, double o2, double mu, double vf, double* __restrict__ x, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double o[3] = {o0, o1, o2};
  for (int k = 0; k < 3; ++k) {
    x[3 * i + k] = o[k] + mu * (x[3 * i + k] - o[k]);
    if (vf != 1.0) v[3 * i + k] *= vf;
  }
}

__global__ void k_scale_v(int n, const double* __restrict__ alpha, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * n) v[i] *= *alpha;
}

// acc −= Σ part, and the rows cleared for the next accumulation (each thread clears the rows it read)
__global__ void k_add_scalar(double* __restrict__ part, int64_t rows, double* __restrict__ acc) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += part[r], part[r] = 0.0;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc -= red[0];
}

// the step counter of captured dynamics: the steps a replay took
__device__ __forceinline__ void b_step_advance(int bx, int gx, uint64_t* __restrict__ stepc, uint64_t by) {
  if (bx == 0 && threadIdx.x == 0) *stepc += by;
}
__global__ void k_step_advance(uint64_t* __restrict__ stepc, uint64_t by) {
  b_step_advance(blockIdx.x, gridDim.x, stepc, by);
}

CellArgs cell_args(const capsgpu_ctx& c) {
  CellArgs C{};
  for (int k = 0; k < 3; ++k) C.a[k] = c.cell[3 + k], C.b[k] = c.cell[6 + k], C.c[k] = c.cell[9 + k];
  C.ortho = C.a[1] == 0 && C.a[2] == 0 && C.b[0] == 0 && C.b[2] == 0 && C.c[0] == 0 && C.c[1] == 0;
  C.il[0] = 1 / C.a[0], C.il[1] = 1 / C.b[1], C.il[2] = 1 / C.c[2];
  const double* a = C.a;
  const double* b = C.b;
  const double* cc = C.c;
  const double bc[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
  const double ca[3] = {cc[1] * a[2] - cc[2] * a[1], cc[2] * a[0] - cc[0] * a[2], cc[0] * a[1] - cc[1] * a[0]};
  const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double v = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
  for (int k = 0; k < 3; ++k) C.hi[0][k] = bc[k] / v, C.hi[1][k] = ca[k] / v, C.hi[2][k] = ab[k] / v;
  return C;
}

}  // namespace

void bonded_setup(capsgpu_ctx& c, const capsgpu_ff& ff) {
  Bonded& B = c.bonded;
  hipStream_t s = c.stream;
  B.nb = ff.n_bonds, B.na = ff.n_angles, B.nf = ff.n_fene;
  const int n = c.n;
  std::vector<std::vector<int32_t>> per(static_cast<size_t>(n));
  auto chk = [&](int32_t a) {
    if (a < 0 || a >= n) throw Error(-1, "a bonded term names an atom out of range");
  };
  // per atom: its slots (bond t, member k: 2t + k; angle t: 2nb + 3t + k; FENE t: 2nb + 3na + 2t + k), terms in order
  for (int t = 0; t < B.nb; ++t)
    for (int k = 0; k < 2; ++k) chk(ff.bond_atoms[2 * t + k]), per[size_t(ff.bond_atoms[2 * t + k])].push_back(2 * t + k);
  for (int t = 0; t < B.na; ++t)
    for (int k = 0; k < 3; ++k) chk(ff.angle_atoms[3 * t + k]), per[size_t(ff.angle_atoms[3 * t + k])].push_back(2 * B.nb + 3 * t + k);
  for (int t = 0; t < B.nf; ++t)
    for (int k = 0; k < 2; ++k)
      chk(ff.fene_atoms[2 * t + k]), per[size_t(ff.fene_atoms[2 * t + k])].push_back(2 * B.nb + 3 * B.na + 2 * t + k);
  std::vector<int32_t> off(size_t(n) + 1, 0), idx;
  for (int i = 0; i < n; ++i) {
    off[size_t(i) + 1] = off[size_t(i)] + int32_t(per[size_t(i)].size());
    idx.insert(idx.end(), per[size_t(i)].begin(), per[size_t(i)].end());
  }
  if (idx.empty()) idx.push_back(0);
  // each member slot's place in that per-atom order (k_bonded writes there; k_sum_forces reads an atom's run)
  std::vector<int32_t> spos(2 * size_t(B.nb) + 3 * size_t(B.na) + 2 * size_t(B.nf) + 1, 0);
  for (int32_t e = 0; e < off[size_t(n)]; ++e) spos[size_t(idx[size_t(e)])] = e;
  B.aoff.upload(off, s);
  B.aslot.upload(idx, s);
  B.spos.upload(spos, s);
  // ranges of consecutive atoms: at most kRangeAtoms atoms and kRangeSlots slots each (none when one atom alone has more
  // slots: then the two-launch path). Each range lists the terms with a member in it, in term order.
  const int64_t nterms = int64_t(B.nb) + B.na + B.nf;
  std::vector<int32_t> rstart(1, 0), rof(size_t(n), 0);
  bool ranged = n > 0 && nterms < (int64_t(1) << 31);
  for (int i = 0, cnt = 0, sl = 0; i < n && ranged; ++i) {
    const int d = off[size_t(i) + 1] - off[size_t(i)];
    if (d > kRangeSlots) ranged = false;
    if (cnt == kRangeAtoms || sl + d > kRangeSlots) rstart.push_back(i), cnt = 0, sl = 0;
    ++cnt, sl += d;
    rof[size_t(i)] = int32_t(rstart.size()) - 1;
  }
  if (ranged) {
    rstart.push_back(n);
    const int nr = int(rstart.size()) - 1;
    std::vector<int32_t> roff(size_t(nr) + 1, 0);
    auto members = [&](int64_t g, int32_t* m) {
      if (g < B.nb) return m[0] = ff.bond_atoms[2 * g], m[1] = ff.bond_atoms[2 * g + 1], 2;
      if (g < B.nb + B.na) {
        const int64_t t = g - B.nb;
        return m[0] = ff.angle_atoms[3 * t], m[1] = ff.angle_atoms[3 * t + 1], m[2] = ff.angle_atoms[3 * t + 2], 3;
      }
      const int64_t t = g - B.nb - B.na;
      return m[0] = ff.fene_atoms[2 * t], m[1] = ff.fene_atoms[2 * t + 1], 2;
    };
    // each distinct range of a term's members, the first member's range first
    auto ranges_of = [&](int64_t g, int32_t* rr) {
      int32_t m[3];
      const int k = members(g, m);
      int nd = 0;
      for (int j = 0; j < k; ++j) {
        const int32_t q = rof[size_t(m[j])];
        bool seen = false;
        for (int u = 0; u < nd; ++u) seen = seen || rr[u] == q;
        if (!seen) rr[nd++] = q;
      }
      return nd;
    };
    for (int64_t g = 0; g < nterms; ++g) {
      int32_t rr[3];
      const int nd = ranges_of(g, rr);
      for (int u = 0; u < nd; ++u) ++roff[size_t(rr[u]) + 1];
    }
    for (int q = 0; q < nr; ++q) roff[size_t(q) + 1] += roff[size_t(q)];
    std::vector<int4> rterm(size_t(std::max(roff[size_t(nr)], 1)), make_int4(-1, -1, -1, -1));
    std::vector<int32_t> fill(roff.begin(), roff.end() - 1);
    for (int64_t g = 0; g < nterms; ++g) {
      int32_t rr[3], m[3] = {-1, -1, -1};
      const int nd = ranges_of(g, rr);
      members(g, m);
      for (int u = 0; u < nd; ++u)
        rterm[size_t(fill[size_t(rr[u])]++)] = make_int4(int32_t(uint32_t(g) | (u == 0 ? 0x80000000u : 0u)), m[0], m[1], m[2]);
    }
    B.rstart.upload(rstart, s);
    B.roff.upload(roff, s);
    B.rterm.upload(rterm, s);
    B.nranges = nr;
  }
  B.ranged = ranged;
  B.slot.resize(3 * (2 * size_t(B.nb) + 3 * size_t(B.na) + 2 * size_t(B.nf)) + 3);
  auto up_i = [&](DevBuf<int32_t>& d, const int32_t* p, size_t m) {
    if (m) d.upload(p, m, s);
    else d.resize(1);
  };
  auto up_d = [&](DevBuf<double>& d, const double* p, size_t m) {
    if (m) d.upload(p, m, s);
    else d.resize(1);
  };
  up_i(B.bij, ff.bond_atoms, 2 * size_t(B.nb));
  up_d(B.bkr, ff.bond_param, 2 * size_t(B.nb));
  up_i(B.aijk, ff.angle_atoms, 3 * size_t(B.na));
  up_d(B.akt, ff.angle_param, 2 * size_t(B.na));
  up_i(B.fij, ff.fene_atoms, 2 * size_t(B.nf));
  up_d(B.fp, ff.fene_param, 4 * size_t(B.nf));
}

void bonded_forces(capsgpu_ctx& c, bool want) {
  Bonded& B = c.bonded;
  const int nt = B.nb + B.na + B.nf;
  const CellArgs C = cell_args(c);
  const unsigned tb = blocks_for(size_t(std::max(nt, 1)), 256);
  B.eblocks = int(tb);
  kt_begin(c, "k_bonded");
  if (want) {
    B.epart.resize(kEB * size_t(tb));
    hipLaunchKernelGGL(k_bonded<true>, dim3(tb), dim3(256), 0, c.stream, B.nb, B.na, B.nf, C, c.x.p, B.bij.p, B.bkr.p, B.aijk.p, B.akt.p,
                       B.fij.p, B.fp.p, B.spos.p, B.slot.p, B.epart.p, c.flags.p);
  } else {
    hipLaunchKernelGGL(k_bonded<false>, dim3(tb), dim3(256), 0, c.stream, B.nb, B.na, B.nf, C, c.x.p, B.bij.p, B.bkr.p, B.aijk.p, B.akt.p,
                       B.fij.p, B.fp.p, B.spos.p, B.slot.p, nullptr, c.flags.p);
  }
  LAUNCHCHECK();
  kt_end(c);
}

void bonded_and_sum(capsgpu_ctx& c, bool want, bool nb, bool rec, bool bond, double* fdst) {
  Bonded& B = c.bonded;
  if (!B.ranged) {
    if (bond) bonded_forces(c, want);
    sum_forces(c, nb, rec, bond, fdst);
    return;
  }
  BondSumArgs a{};
  a.nb = B.nb, a.na = B.na, a.nf = B.nf;
  a.C = cell_args(c);
  a.x = c.x.p;
  a.bij = B.bij.p, a.aijk = B.aijk.p, a.fij = B.fij.p, a.spos = B.spos.p, a.aoff = B.aoff.p;
  a.bkr = B.bkr.p, a.akt = B.akt.p, a.fp = B.fp.p;
  a.rstart = B.rstart.p, a.roff = B.roff.p, a.rterm = B.rterm.p;
  a.iperm = nb ? c.L.iperm.p : nullptr;
  a.fnb = nb ? c.L.fnb.p : nullptr;
  a.frec = rec ? c.pme.fpme.p : nullptr;
  a.f = fdst;
  a.flags = c.flags.p;
  a.bonded = bond ? 1 : 0;
  const bool ke = want && bond;
  if (ke) {
    B.eblocks = B.nranges;
    B.epart.resize(kEB * size_t(B.nranges));
    a.epart = B.epart.p;
  }
  kt_begin(c, "k_bonded_sum");
  const dim3 grid(unsigned(B.nranges)), block(kRangeThreads);
  if (a.C.ortho) {
    if (ke) hipLaunchKernelGGL((k_bonded_sum<true, 1>), grid, block, 0, c.stream, a);
    else hipLaunchKernelGGL(k_bonded_sum_ortho, grid, block, 0, c.stream, a);
  } else {
    if (ke) hipLaunchKernelGGL((k_bonded_sum<true, 0>), grid, block, 0, c.stream, a);
    else hipLaunchKernelGGL((k_bonded_sum<false, 0>), grid, block, 0, c.stream, a);
  }
  LAUNCHCHECK();
  kt_end(c);
}

void sum_forces(capsgpu_ctx& c, bool nb, bool rec, bool bond, double* fdst) {
  kt_begin(c, "k_sum_forces");
  hipLaunchKernelGGL(k_sum_forces, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, c.L.iperm.p, nb ? c.L.fnb.p : nullptr,
                     rec ? c.pme.fpme.p : nullptr, bond ? c.bonded.aoff.p : nullptr, c.bonded.slot.p, fdst);
  LAUNCHCHECK();
  kt_end(c);
}

void slot_virial(capsgpu_ctx& c, double* part) {
  const int64_t ns = c.L.g.ns;
  kt_begin(c, "k_slot_virial");
  hipLaunchKernelGGL(k_slot_virial, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, c.L.xq.p, c.L.org.p, c.L.fnb.p, part);
  LAUNCHCHECK();
  kt_end(c);
}

void colsum(capsgpu_ctx& c, const double* part, int64_t rows, int cols, double* out) {
  kt_begin(c, "k_colsum");
  hipLaunchKernelGGL(k_colsum, dim3(unsigned(cols)), dim3(256), 0, c.stream, part, rows, cols, out);
  LAUNCHCHECK();
  kt_end(c);
}

void check_moves(capsgpu_ctx& c, double lim_build, double lim_prune) {
  kt_begin(c, "k_check_moves");
  hipLaunchKernelGGL(k_check_moves, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, c.x.p, c.x0.p, c.x0p.p, lim_build, lim_prune,
                     c.flags.p);
  LAUNCHCHECK();
  kt_end(c);
}

void vv_open(capsgpu_ctx& c, double hk, double dt, const double* alpha, double lim_b, double lim_p, const double* f, const Kicks& pre,
             const AtomSplines* sp) {
  AtomSplines none{};
  kt_begin(c, "k_vv_open");
  hipLaunchKernelGGL(k_vv_open, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, hk, dt, c.d_minv.p, alpha, f ? f : c.f.p, c.x.p,
                     c.v.p, c.x0.p, c.x0p.p, lim_b, lim_p, c.flags.p, pre, sp ? *sp : none);
  LAUNCHCHECK();
  kt_end(c);
}

void baoab_open(capsgpu_ctx& c, double hk, double hdt, double c1, double c2, const double* sdv, uint64_t seed, const uint64_t* stepc,
                uint64_t step_add, double lim_b, double lim_p, double* part, const Kicks& pre, const AtomSplines* sp) {
  AtomSplines none{};
  kt_begin(c, "k_baoab_open");
  hipLaunchKernelGGL(k_baoab_open, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, hk, hdt, c1, c2, c.d_minv.p, sdv, seed, stepc,
                     step_add, c.f.p, c.x.p, c.v.p, c.x0.p, c.x0p.p, lim_b, lim_p, c.flags.p, part, pre, sp ? *sp : none);
  LAUNCHCHECK();
  kt_end(c);
}

void kick(capsgpu_ctx& c, double hk, bool ke, double* part, const double* f, const Kicks& pre) {
  if (!f) f = c.f.p;
  kt_begin(c, "k_kick");
  if (ke)
    hipLaunchKernelGGL(k_kick<true>, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, hk, c.d_minv.p, f, c.v.p, part, pre);
  else
    hipLaunchKernelGGL(k_kick<false>, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, hk, c.d_minv.p, f, c.v.p, part, pre);
  LAUNCHCHECK();
  kt_end(c);
}

void bussi(capsgpu_ctx& c, const double* part, double kt_target, double ndof, double cc, double r1, double sum, double* alpha, double* bath) {
  kt_begin(c, "k_bussi");
  hipLaunchKernelGGL(k_bussi, dim3(1), dim3(256), 0, c.stream, part, int64_t(blocks_for(c.n, 256)), kt_target, ndof, cc, r1, sum, alpha,
                     bath);
  LAUNCHCHECK();
  kt_end(c);
}

void scale_v(capsgpu_ctx& c, const double* alpha) {
  kt_begin(c, "k_scale_v");
  hipLaunchKernelGGL(k_scale_v, dim3(blocks_for(3 * size_t(c.n), 256)), dim3(256), 0, c.stream, c.n, alpha, c.v.p);
  LAUNCHCHECK();
  kt_end(c);
}

void scale_xv(capsgpu_ctx& c, const double origin[3], double mu, double vf) {
  kt_begin(c, "k_scale_xv");
  hipLaunchKernelGGL(k_scale_xv, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, origin[0], origin[1], origin[2], mu, vf, c.x.p,
                     c.v.p);
  LAUNCHCHECK();
  kt_end(c);
}

void step_advance(capsgpu_ctx& c, uint64_t* stepc, uint64_t by) {
  kt_begin(c, "k_step_advance");
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, c.stream, stepc, by);
  LAUNCHCHECK();
  kt_end(c);
}

void bath_add(capsgpu_ctx& c, double* part, double* bath) {
  kt_begin(c, "k_add_scalar");
  hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(256), 0, c.stream, part, int64_t(blocks_for(c.n, 256)), bath);
  LAUNCHCHECK();
  kt_end(c);
}


// ---- many contexts in one launch per kernel (the replica batch) ----------------------------------------------------
// As the grouped PME kernels: a workgroup finds its replica in a prefix array of workgroup counts and runs the
// single-context body with that replica's arguments, so each replica's arithmetic (and result) is unchanged.
namespace {
struct StepRep {
  int n;
  double hk, hdt, c1, c2;
  uint64_t seed;
  uint64_t* stepc;
  const double *minv, *sdv;
  double *x, *v, *f;
  const double *x0, *x0p;
  uint32_t* flags;
  double *bpart, *kpart;
  int nb, na, nf;
  CellArgs C;
  const int32_t *bij, *aijk, *fij, *spos, *aoff;
  const double *bkr, *akt, *fp;
  double* slot;
  const float* frec;
  const int32_t *rstart, *roff;
  const int4* rterm;
  int bonded;
  AtomSplines sp;   // the replica's per-atom PME bookkeeping, done by the opening launch (sp.wq null: none)
};

__device__ inline int find_rep(const int* __restrict__ off, int nr, int b) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) kg_baoab_open(const StepRep* __restrict__ R, const int* __restrict__ off, int nr, uint64_t step_add) {
  const int r = find_rep(off, nr, blockIdx.x);
  const StepRep& a = R[r];
  b_baoab_open(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.n, a.hk, a.hdt, a.c1, a.c2, a.minv, a.sdv, a.seed, a.stepc, step_add, a.f,
               a.x, a.v, a.x0, a.x0p, -1.0, -1.0, a.flags, a.bpart, Kicks{}, &a.sp);
}

__global__ void __launch_bounds__(256) kg_bonded(const StepRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const StepRep& a = R[r];
  b_bonded<false>(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.nb, a.na, a.nf, a.C, a.x, a.bij, a.bkr, a.aijk, a.akt, a.fij, a.fp,
                  a.spos, a.slot, nullptr, a.flags);
}

__global__ void __launch_bounds__(256) kg_sum_forces(const StepRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const StepRep& a = R[r];
  b_sum_forces(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.n, nullptr, nullptr, a.frec, a.aoff, a.slot, a.f);
}

__global__ void __launch_bounds__(kRangeThreads) kg_bonded_sum(const StepRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const StepRep& a = R[r];
  BondSumArgs b{};
  b.nb = a.nb, b.na = a.na, b.nf = a.nf, b.C = a.C, b.x = a.x;
  b.bij = a.bij, b.aijk = a.aijk, b.fij = a.fij, b.spos = a.spos, b.aoff = a.aoff;
  b.bkr = a.bkr, b.akt = a.akt, b.fp = a.fp;
  b.rstart = a.rstart, b.roff = a.roff, b.rterm = a.rterm;
  b.frec = a.frec, b.f = a.f, b.flags = a.flags, b.bonded = a.bonded;
  b_bonded_sum<false, -1>(int(blockIdx.x) - off[r], b);
}

__global__ void __launch_bounds__(256) kg_kick(const StepRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const StepRep& a = R[r];
  b_kick<false>(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.n, a.hk, a.minv, a.f, a.v, a.kpart, Kicks{});
}

__global__ void kg_step_advance(const StepRep* __restrict__ R, int nr, uint64_t by) {
  const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < nr) *R[r].stepc += by;
}
}  // namespace

struct StepGroup {
  int nr = 0;
  DevBuf<unsigned char> reps;             // StepRep[nr]
  DevBuf<int32_t> off_atoms, off_terms, off_ranges;   // workgroup prefixes
  int h_atoms = 0, h_terms = 0, h_ranges = 0;
  bool ranged = true;                                 // every replica's atoms in ranges: one bonded-and-sum launch
};

StepGroup* step_group_create(const std::vector<StepGroupInput>& in, hipStream_t s) {
  auto g = std::make_unique<StepGroup>();
  g->nr = int(in.size());
  if (g->nr < 1) throw Error(-1, "a step group needs a replica");
  std::vector<StepRep> R(in.size());
  std::vector<int32_t> oa(1, 0), ot(1, 0), orr(1, 0);
  for (size_t i = 0; i < in.size(); ++i) {
    capsgpu_ctx& c = *in[i].c;
    Bonded& B = c.bonded;
    StepRep& a = R[i];
    a.n = c.n;
    a.hk = in[i].hk, a.hdt = in[i].hdt, a.c1 = in[i].c1, a.c2 = in[i].c2, a.seed = in[i].seed, a.stepc = in[i].stepc;
    a.minv = c.d_minv.p, a.sdv = in[i].sdv, a.x = c.x.p, a.v = c.v.p, a.f = c.f.p, a.x0 = c.x0.p, a.x0p = c.x0p.p;
    a.flags = c.flags.p, a.bpart = in[i].bpart, a.kpart = in[i].kpart;
    a.nb = in[i].bonded ? B.nb : 0, a.na = in[i].bonded ? B.na : 0, a.nf = in[i].bonded ? B.nf : 0;
    a.C = cell_args(c);
    a.bij = B.bij.p, a.aijk = B.aijk.p, a.fij = B.fij.p, a.spos = B.spos.p, a.aoff = in[i].bonded ? B.aoff.p : nullptr;
    a.bkr = B.bkr.p, a.akt = B.akt.p, a.fp = B.fp.p, a.slot = B.slot.p;
    a.frec = in[i].recip ? c.pme.fpme.p : nullptr;
    a.rstart = B.rstart.p, a.roff = B.roff.p, a.rterm = B.rterm.p, a.bonded = in[i].bonded ? 1 : 0;
    a.sp = AtomSplines{};
    if (in[i].recip) a.sp = pme_atom_splines(c);
    g->ranged = g->ranged && B.ranged;
    orr.push_back(orr.back() + B.nranges);
    oa.push_back(oa.back() + int(blocks_for(size_t(c.n), 256)));
    ot.push_back(ot.back() + int(blocks_for(size_t(std::max(a.nb + a.na + a.nf, 1)), 256)));
  }
  g->reps.upload(reinterpret_cast<const unsigned char*>(R.data()), R.size() * sizeof(StepRep), s);
  g->off_atoms.upload(oa, s), g->off_terms.upload(ot, s), g->off_ranges.upload(orr, s);
  g->h_atoms = oa.back(), g->h_terms = ot.back(), g->h_ranges = orr.back();
  return g.release();
}

void step_group_destroy(StepGroup* g) { delete g; }

void step_group_open(StepGroup& g, hipStream_t s, uint64_t step_add) {
  hipLaunchKernelGGL(kg_baoab_open, dim3(unsigned(g.h_atoms)), dim3(256), 0, s, reinterpret_cast<const StepRep*>(g.reps.p), g.off_atoms.p,
                     g.nr, step_add);
  LAUNCHCHECK();
}

void step_group_close(StepGroup& g, hipStream_t s) {
  const StepRep* R = reinterpret_cast<const StepRep*>(g.reps.p);
  if (g.ranged) {
    hipLaunchKernelGGL(kg_bonded_sum, dim3(unsigned(g.h_ranges)), dim3(kRangeThreads), 0, s, R, g.off_ranges.p, g.nr);
    LAUNCHCHECK();
  } else {
    hipLaunchKernelGGL(kg_bonded, dim3(unsigned(g.h_terms)), dim3(256), 0, s, R, g.off_terms.p, g.nr);
    LAUNCHCHECK();
    hipLaunchKernelGGL(kg_sum_forces, dim3(unsigned(g.h_atoms)), dim3(256), 0, s, R, g.off_atoms.p, g.nr);
    LAUNCHCHECK();
  }
  hipLaunchKernelGGL(kg_kick, dim3(unsigned(g.h_atoms)), dim3(256), 0, s, R, g.off_atoms.p, g.nr);
  LAUNCHCHECK();
}

void step_group_advance(StepGroup& g, hipStream_t s, uint64_t by) {
  hipLaunchKernelGGL(kg_step_advance, dim3(unsigned((g.nr + 63) / 64)), dim3(64), 0, s, reinterpret_cast<const StepRep*>(g.reps.p), g.nr, by);
  LAUNCHCHECK();
}

}  // namespace capsgpu
