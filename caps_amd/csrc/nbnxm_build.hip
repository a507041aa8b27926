// The cluster pair list on the GPU, bit-identical to CAPS's CPU build (core/src/nbnxm.cpp:22-326).
//
// Built with -ffp-contract=off: every FP64 binning and bounding-box operation and every FP32 mask operation rounds as
// the reference's (x86-64, SSE2 lanes with an unfused multiply and add, core/src/simd4.hpp:79-80).
// The order of the list does not come from the order threads run in:
//   * the column sort is a rank sort on the total order (f_c, atom index) (nbnxm.cpp:70-71);
//   * the candidate search runs one thread per (i cluster, column window), counts first, then writes at exclusive
//     offsets, so the entries land in the reference's loop order du -> dv -> qc -> cj (nbnxm.cpp:177-197);
//   * the mirrored entries of the full list are sorted by source entry per cluster.
#include <algorithm>
#include <cmath>

#include "context.hpp"

namespace capsgpu {
namespace {

struct D3 { double v[3]; };

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline double clampd(double x, double lo, double hi) { return x < lo ? lo : x > hi ? hi : x; }

struct GeomArgs {
  double o[3], a[3], b[3], c[3];
  double bc[3], ca[3], ab[3], vs;
  int n, nu, nv, ncols, ru, rv, rq, nwin;
  int qlo[3], span[3];
  double rl2, dw;
  float rl2f;
};

GeomArgs geom_args(const ListGeom& g) {
  GeomArgs A;
  for (int k = 0; k < 3; ++k) {
    A.o[k] = g.cell[k], A.a[k] = g.cell[3 + k], A.b[k] = g.cell[6 + k], A.c[k] = g.cell[9 + k];
    A.bc[k] = g.bc[k], A.ca[k] = g.ca[k], A.ab[k] = g.ab[k];
    A.qlo[k] = g.qlo[k], A.span[k] = g.span[k];
  }
  A.vs = g.vsigned;
  A.n = g.n, A.nu = g.nu, A.nv = g.nv, A.ncols = g.ncols, A.ru = g.ru, A.rv = g.rv, A.rq = g.rq, A.nwin = g.nwin;
  A.rl2 = g.rl2, A.dw = g.dw, A.rl2f = g.rl2f;
  return A;
}

// Cell::to_fractional (core/src/system.cpp:18-23) with the cross products precomputed the same way
__device__ inline void to_frac(const GeomArgs& A, const double r[3], double f[3]) {
  const double d[3] = {r[0] - A.o[0], r[1] - A.o[1], r[2] - A.o[2]};
  f[0] = dot3(d, A.bc) / A.vs;
  f[1] = dot3(d, A.ca) / A.vs;
  f[2] = dot3(d, A.ab) / A.vs;
}

__global__ void k_bin(GeomArgs A, const double* __restrict__ x, double* __restrict__ fr, uint32_t* __restrict__ col,
                      int32_t* __restrict__ colcnt, uint32_t* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double r[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  double f[3];
  to_frac(A, r, f);
  for (int k = 0; k < 3; ++k) {
    // a position that is not a number cannot be binned: flag it and keep the index in range
    if (!isfinite(f[k])) {
      atomicOr(bad, 1u);
      f[k] = 0.0;
    }
    f[k] = clampd(f[k] - floor(f[k]), 0.0, 1.0 - 1e-12);
  }
  const double fu = clampd(f[0], 0.0, 1.0 - 1e-12), fv = clampd(f[1], 0.0, 1.0 - 1e-12);
  const int cu = min(A.nu - 1, int(fu * A.nu)), cv = min(A.nv - 1, int(fv * A.nv));
  const uint32_t c = uint32_t(cu * A.nv + cv);
  col[i] = c;
  for (int k = 0; k < 3; ++k) fr[3 * i + k] = f[k];
  atomicAdd(&colcnt[c], 1);
}

__global__ void k_bucket(int n, const uint32_t* __restrict__ col, const int64_t* __restrict__ cstart, int32_t* __restrict__ bfill,
                         uint32_t* __restrict__ bucket) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = col[i];
  bucket[cstart[c] + atomicAdd(&bfill[c], 1)] = uint32_t(i);
}

// rank of atom i in its column under (f_c, index)
__global__ void k_rank(int n, const uint32_t* __restrict__ col, const int64_t* __restrict__ cstart, const uint32_t* __restrict__ bucket,
                       const double* __restrict__ fr, uint32_t* __restrict__ rank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = col[i];
  const double fi = fr[3 * i + 2];
  uint32_t r = 0;
  for (int64_t k = cstart[c]; k < cstart[c + 1]; ++k) {
    const uint32_t j = bucket[k];
    const double fj = fr[3 * j + 2];
    r += (fj < fi || (fj == fi && j < uint32_t(i))) ? 1u : 0u;
  }
  rank[i] = r;
}

__global__ void k_col_clusters(int ncols, const int32_t* __restrict__ colcnt, uint32_t* __restrict__ ccls) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncols) ccls[c] = uint32_t((colcnt[c] + 3) / 4);
}

__global__ void k_perm(int n, const uint32_t* __restrict__ col, const uint32_t* __restrict__ rank, const int64_t* __restrict__ ccl,
                       uint32_t* __restrict__ perm, uint32_t* __restrict__ iperm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = 4 * ccl[col[i]] + rank[i];
  perm[s] = uint32_t(i);
  iperm[i] = uint32_t(s);
}

__global__ void k_ccol(int ncols, const int64_t* __restrict__ ccl, uint32_t* __restrict__ ccol) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  for (int64_t k = ccl[c]; k < ccl[c + 1]; ++k) ccol[k] = uint32_t(c);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// wrapped Cartesian positions by slot, padding on the cluster's first atom (nbnxm.cpp:88-93)
__global__ void k_slots(GeomArgs A, int64_t ns, const uint32_t* __restrict__ perm, const double* __restrict__ fr,
                        double* __restrict__ cw, double* __restrict__ frs) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const uint32_t a = perm[s] != kPad ? perm[s] : perm[s & ~int64_t(3)];
  const double f0 = fr[3 * a], f1 = fr[3 * a + 1], f2 = fr[3 * a + 2];
  for (int k = 0; k < 3; ++k) {
    frs[3 * s + k] = fr[3 * a + k];
    cw[3 * s + k] = ((A.o[k] + A.a[k] * f0) + A.b[k] * f1) + A.c[k] * f2;
  }
}

// bounding boxes, c ranges, origins, single-precision offsets and real lanes per cluster (nbnxm.cpp:94-117)
__global__ void k_clusters(int64_t ncl, const uint32_t* __restrict__ perm, const double* __restrict__ cw, const double* __restrict__ frs,
                           double* __restrict__ blo, double* __restrict__ bhi, double* __restrict__ wlo, double* __restrict__ whi,
                           double* __restrict__ org, float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz,
                           uint8_t* __restrict__ real) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) lo[k] = hi[k] = cw[3 * (4 * c) + k];
  double wl = frs[3 * (4 * c) + 2], wh = wl;
  for (int l = 1; l < 4; ++l)
    for (int k = 0; k < 3; ++k) {
      const double v = cw[3 * (4 * c + l) + k];
      lo[k] = v < lo[k] ? v : lo[k];
      hi[k] = hi[k] < v ? v : hi[k];
    }
  for (int l = 1; l < 4; ++l) {
    const double v = frs[3 * (4 * c + l) + 2];
    wl = v < wl ? v : wl;
    wh = wh < v ? v : wh;
  }
  double o[3];
  for (int k = 0; k < 3; ++k) {
    blo[3 * c + k] = lo[k], bhi[3 * c + k] = hi[k];
    o[k] = 0.5 * (lo[k] + hi[k]);
    org[3 * c + k] = o[k];
  }
  wlo[c] = wl, whi[c] = wh;
  uint8_t rm = 0;
  for (int l = 0; l < 4; ++l) {
    const int64_t s = 4 * c + l;
    bx[s] = float(cw[3 * s] - o[0]), by[s] = float(cw[3 * s + 1] - o[1]), bz[s] = float(cw[3 * s + 2] - o[2]);
    if (perm[s] != kPad) rm |= uint8_t(1u << l);
  }
  real[c] = rm;
}

__global__ void k_excl_count(int64_t ncl, const uint32_t* __restrict__ perm, const int32_t* __restrict__ eoff, uint32_t* __restrict__ cnt) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  uint32_t t = 0;
  for (int a = 0; a < 4; ++a) {
    const uint32_t ai = perm[4 * c + a];
    if (ai != kPad) t += uint32_t(eoff[ai + 1] - eoff[ai]);
  }
  cnt[c] = t;
}

__device__ inline bool csr_has(const int32_t* off, const int32_t* idx, uint32_t i, uint32_t j) {
  int32_t lo = off[i], hi = off[i + 1];
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (uint32_t(idx[mid]) < j) lo = mid + 1;
    else hi = mid;
  }
  return lo < off[i + 1] && uint32_t(idx[lo]) == j;
}

// per i cluster: (j cluster, bit, 1-4 bit) of its excluded pairs, sorted by (cj, bit) and merged per cj (nbnxm.cpp:118-145)
__global__ void k_excl_fill(int64_t ncl, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ iperm,
                            const int32_t* __restrict__ eoff, const int32_t* __restrict__ eidx, const int32_t* __restrict__ p14off,
                            const int32_t* __restrict__ p14idx, int keep, const int64_t* __restrict__ xoff, uint32_t* __restrict__ xcj,
                            uint16_t* __restrict__ xbits, uint16_t* __restrict__ x14, int32_t* __restrict__ xcnt) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  const int64_t b = xoff[c];
  int m = 0;
  for (int a = 0; a < 4; ++a) {
    const uint32_t ai = perm[4 * c + a];
    if (ai == kPad) continue;
    for (int32_t e = eoff[ai]; e < eoff[ai + 1]; ++e) {
      const uint32_t aj = uint32_t(eidx[e]), sj = iperm[aj];
      const uint16_t bit = uint16_t(1u << (4 * a + int(sj % 4)));
      const bool k14 = keep && p14off && csr_has(p14off, p14idx, ai, aj);
      // insertion by (cj, bit)
      const uint32_t cj = sj / 4;
      int p = m;
      while (p > 0 && (xcj[b + p - 1] > cj || (xcj[b + p - 1] == cj && xbits[b + p - 1] > bit))) {
        xcj[b + p] = xcj[b + p - 1], xbits[b + p] = xbits[b + p - 1], x14[b + p] = x14[b + p - 1];
        --p;
      }
      xcj[b + p] = cj, xbits[b + p] = bit, x14[b + p] = k14 ? bit : 0;
      ++m;
    }
  }
  int w = 0;
  for (int r = 0; r < m; ++r) {
    if (w > 0 && xcj[b + w - 1] == xcj[b + r]) {
      xbits[b + w - 1] |= xbits[b + r], x14[b + w - 1] |= x14[b + r];
    } else {
      xcj[b + w] = xcj[b + r], xbits[b + w] = xbits[b + r], x14[b + w] = x14[b + r];
      ++w;
    }
  }
  xcnt[c] = w;
}

struct SearchArgs {
  GeomArgs A;
  int64_t ncl;
  const uint32_t* ccol;
  const int64_t* ccl;
  const double *blo, *bhi, *wlo, *whi, *org, *cw, *frs;
  const float *bx, *by, *bz;
  const uint8_t* real;
  const uint32_t* perm;
  const int64_t* xoff;
  const int32_t* xcnt;
  const uint32_t* xcj;
  const uint16_t *xbits, *x14;
  int keep;
};

// One (i cluster, column window) of the candidate search (nbnxm.cpp:174-258): count mode or write mode.
template <bool kWrite>
__global__ void k_search(SearchArgs P, uint32_t* __restrict__ wcount, const int64_t* __restrict__ woff, uint32_t* __restrict__ o_cj,
                         uint32_t* __restrict__ o_ci, uint16_t* __restrict__ o_code, uint16_t* __restrict__ o_mask,
                         uint16_t* __restrict__ o_ex, uint16_t* __restrict__ o_ex14) {
  const GeomArgs& A = P.A;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= P.ncl * A.nwin) return;
  const int64_t ci = t / A.nwin;
  const int wv = int(t % A.nwin);
  const int du = wv / (2 * A.rv + 1) - A.ru, dvv = wv % (2 * A.rv + 1) - A.rv;
  const int cu = int(P.ccol[ci]) / A.nv, cv = int(P.ccol[ci]) % A.nv;
  int U = cu + du;
  const int qa = int(floor(double(U) / A.nu));
  U -= qa * A.nu;
  int V = cv + dvv;
  const int qb = int(floor(double(V) / A.nv));
  V -= qb * A.nv;
  const size_t cc = size_t(U) * A.nv + V;
  const double wlo_i = P.wlo[ci], whi_i = P.whi[ci];
  int qc0 = int(floor(wlo_i - A.dw)), qc1 = int(floor(whi_i + A.dw));
  qc0 = qc0 < -A.rq ? -A.rq : qc0;
  qc1 = qc1 > A.rq ? A.rq : qc1;
  const double* bhi_i = P.bhi + 3 * ci;
  const double* blo_i = P.blo + 3 * ci;
  const double* oi = P.org + 3 * ci;
  const uint8_t real_i = P.real[ci];
  uint32_t count = 0;
  int64_t out = kWrite ? woff[t] : 0;
  const int64_t c_end = P.ccl[cc + 1];
  for (int qc = qc0; qc <= qc1; ++qc) {
    double S[3];
    for (int k = 0; k < 3; ++k) S[k] = (A.a[k] * double(qa) + A.b[k] * double(qb)) + A.c[k] * double(qc);
    const bool zero = qa == 0 && qb == 0 && qc == 0;
    const bool positive = qa > 0 || (qa == 0 && (qb > 0 || (qb == 0 && qc > 0)));
    const uint16_t code = uint16_t(((qa - A.qlo[0]) * A.span[1] + (qb - A.qlo[1])) * A.span[2] + (qc - A.qlo[2]));
    const double lo = wlo_i - A.dw - qc, hi = whi_i + A.dw - qc;
    int64_t b0 = P.ccl[cc], b1 = c_end;
    while (b0 < b1) {
      const int64_t mid = b0 + (b1 - b0) / 2;
      if (P.whi[mid] < lo) b0 = mid + 1;
      else b1 = mid;
    }
    for (int64_t cj = b0; cj < c_end && P.wlo[cj] <= hi; ++cj) {
      if (cj != ci) {
        if ((cj > ci) != (((ci + cj) & 1) == 0)) continue;
      } else if (!zero && !positive) {
        continue;
      }
      double d2 = 0;
      for (int k = 0; k < 3; ++k) {
        const double g1 = P.blo[3 * cj + k] + S[k] - bhi_i[k], g2 = blo_i[k] - (P.bhi[3 * cj + k] + S[k]);
        const double g = g1 > 0 ? g1 : g2 > 0 ? g2 : 0.0;
        d2 += g * g;
      }
      if (d2 >= A.rl2) continue;
      const double* oj = P.org + 3 * cj;
      const float sx = float(S[0] + oj[0] - oi[0]), sy = float(S[1] + oj[1] - oi[1]), sz = float(S[2] + oj[2] - oi[2]);
      float xj[4], yj[4], zj[4];
      for (int l = 0; l < 4; ++l) xj[l] = P.bx[4 * cj + l] + sx, yj[l] = P.by[4 * cj + l] + sy, zj[l] = P.bz[4 * cj + l] + sz;
      const uint8_t real_j = P.real[cj];
      unsigned mask = 0;
      for (int a = 0; a < 4; ++a) {
        if (!((real_i >> a) & 1u)) continue;
        const int64_t sa = 4 * ci + a;
        const float xa = P.bx[sa], ya = P.by[sa], za = P.bz[sa];
        unsigned row = 0;
        for (int l = 0; l < 4; ++l) {
          const float dx = xj[l] - xa, dy = yj[l] - ya, dz = zj[l] - za;
          const float zz = dz * dz;
          const float yy = dy * dy + zz;
          const float q = dx * dx + yy;
          row |= (q < A.rl2f ? 1u : 0u) << l;
        }
        row &= real_j;
        if (cj == ci && zero) row &= (0xFu << (a + 1)) & 0xFu;
        mask |= row << (4 * a);
      }
      unsigned exm = 0, e14 = 0;
      if (mask) {
        int64_t l0 = P.xoff[ci], l1 = l0 + P.xcnt[ci];
        const int64_t lend = l1;
        while (l0 < l1) {
          const int64_t mid = (l0 + l1) >> 1;
          if (P.xcj[mid] < uint32_t(cj)) l0 = mid + 1;
          else l1 = mid;
        }
        if (l0 < lend && P.xcj[l0] == uint32_t(cj) && (P.xbits[l0] & mask)) {
          const unsigned exb = P.xbits[l0] & mask, k14 = P.x14[l0];
          for (int b = 0; b < 16; ++b) {
            if (!((exb >> b) & 1u)) continue;
            const int64_t sa = 4 * ci + b / 4, sb = 4 * cj + b % 4;
            if (P.perm[sa] == P.perm[sb]) continue;
            double d[3], m[3];
            for (int k = 0; k < 3; ++k) {
              d[k] = (P.cw[3 * sb + k] + S[k]) - P.cw[3 * sa + k];
              m[k] = P.frs[3 * sb + k] - P.frs[3 * sa + k];
              m[k] -= round(m[k]);
            }
            double dm[3];
            for (int k = 0; k < 3; ++k) dm[k] = (A.a[k] * m[0] + A.b[k] * m[1]) + A.c[k] * m[2];
            if (dot3(d, d) <= dot3(dm, dm) + 1e-6) {
              mask &= ~(1u << b);
              if (P.keep) exm |= 1u << b, e14 |= k14 & (1u << b);
            }
          }
        }
      }
      if (mask | exm) {
        if (kWrite) {
          o_cj[out] = uint32_t(cj), o_ci[out] = uint32_t(ci), o_code[out] = code, o_mask[out] = uint16_t(mask),
          o_ex[out] = uint16_t(exm), o_ex14[out] = uint16_t(e14);
          ++out;
        }
        ++count;
      }
    }
  }
  if (!kWrite) wcount[t] = count;
}

__global__ void k_first(int64_t ncl, int nwin, const int64_t* __restrict__ woff, int64_t* __restrict__ first) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c <= ncl) first[c] = woff[c * nwin];
}

// offsets to the binned positions as Cartesian vectors and integer lattice coefficients (nbnxm.cpp:267-277)
__global__ void k_offsets(GeomArgs A, int64_t ns, const uint32_t* __restrict__ perm, const double* __restrict__ cw,
                          const double* __restrict__ x, double* __restrict__ off, int32_t* __restrict__ ioff) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  if (perm[s] == kPad) {
    for (int k = 0; k < 3; ++k) off[3 * s + k] = 0.0, ioff[3 * s + k] = 0;
    return;
  }
  const size_t a = perm[s];
  double o[3];
  for (int k = 0; k < 3; ++k) o[k] = cw[3 * s + k] - x[3 * a + k];
  const double r[3] = {A.o[0] + o[0], A.o[1] + o[1], A.o[2] + o[2]};
  double f[3];
  to_frac(A, r, f);
  for (int k = 0; k < 3; ++k) off[3 * s + k] = o[k], ioff[3 * s + k] = int32_t(llround(f[k]));
}

// per slot: charge and LJ factors (padding: none) (nbnxm.cpp:284-294)
__global__ void k_slot_params(int64_t ns, const uint32_t* __restrict__ perm, const int32_t* __restrict__ type, const double* __restrict__ q,
                              const float* __restrict__ f12, const float* __restrict__ f6, float4* __restrict__ xq, float2* __restrict__ pp) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  if (perm[s] == kPad) {
    xq[s].w = 0.0f;
    pp[s] = make_float2(0.0f, 0.0f);
    return;
  }
  const uint32_t a = perm[s];
  xq[s].w = float(q[a]);
  const int t = type[a];
  pp[s] = make_float2(f12 ? f12[t] : 0.0f, f6 ? f6[t] : 0.0f);
}

// ---- the full list -------------------------------------------------------------------------------------------------
__global__ void k_full_count(int64_t ne, int zero_code, const uint32_t* __restrict__ e_cj, const uint32_t* __restrict__ e_ci,
                             const uint16_t* __restrict__ e_code, uint32_t* __restrict__ own, uint32_t* __restrict__ mir) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  atomicAdd(&own[e_ci[e]], 1u);
  if (e_cj[e] != e_ci[e] || e_code[e] != zero_code) atomicAdd(&mir[e_cj[e]], 1u);
}

__global__ void k_full_sum(int64_t ncl, const uint32_t* __restrict__ own, const uint32_t* __restrict__ mir, uint32_t* __restrict__ cnt) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < ncl) cnt[c] = own[c] + mir[c];
}

__global__ void k_full_fill(int64_t ne, int zero_code, const uint32_t* __restrict__ e_cj, const uint32_t* __restrict__ e_ci,
                            const uint16_t* __restrict__ e_code, const int64_t* __restrict__ first, const int64_t* __restrict__ fstart,
                            const uint32_t* __restrict__ own, uint32_t* __restrict__ mfill, uint32_t* __restrict__ fsrc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const uint32_t ci = e_ci[e], cj = e_cj[e];
  fsrc[fstart[ci] + (e - first[ci])] = uint32_t(e);
  if (cj != ci || e_code[e] != zero_code) fsrc[fstart[cj] + own[cj] + atomicAdd(&mfill[cj], 1u)] = uint32_t(e) | 0x80000000u;
}

// the mirrored segment of each cluster sorted by source entry: one workgroup per cluster, ranks counted in LDS
constexpr int kSortCap = 2048;
__global__ void __launch_bounds__(256) k_full_sort(int64_t ncl, const int64_t* __restrict__ fstart, const uint32_t* __restrict__ own,
                                                   uint32_t* __restrict__ fsrc) {
  __shared__ uint32_t keys[kSortCap];
  const int64_t c = blockIdx.x;
  if (c >= ncl) return;
  const int64_t b = fstart[c] + own[c];
  const int m = int(fstart[c + 1] - b);
  if (m <= 1) return;
  if (m <= kSortCap) {
    for (int k = threadIdx.x; k < m; k += blockDim.x) keys[k] = fsrc[b + k];
    __syncthreads();
    uint32_t mine[kSortCap / 256];
    int rk[kSortCap / 256];
    int cnt = 0;
    for (int k = threadIdx.x; k < m; k += blockDim.x, ++cnt) {
      const uint32_t v = keys[k];
      int r = 0;
      for (int u = 0; u < m; ++u) r += keys[u] < v ? 1 : 0;
      mine[cnt] = v, rk[cnt] = r;
    }
    __syncthreads();
    for (int q = 0; q < cnt; ++q) fsrc[b + rk[q]] = mine[q];
  } else if (threadIdx.x == 0) {
    // a very long segment: insertion sort in place (keys are distinct)
    for (int k = 1; k < m; ++k) {
      const uint32_t v = fsrc[b + k];
      int p = k;
      while (p > 0 && fsrc[b + p - 1] > v) fsrc[b + p] = fsrc[b + p - 1], --p;
      fsrc[b + p] = v;
    }
  }
}

}  // namespace

void list_build(capsgpu_ctx& c) {
  DevList& L = c.L;
  ListGeom& g = L.g;
  hipStream_t s = c.stream;
  const int n = c.n;
  // geometry on the host, as nbnxm.cpp:40-47 and 146-164
  g.n = n;
  for (int k = 0; k < 12; ++k) g.cell[k] = c.cell[k];
  const double *a = c.cell + 3, *b = c.cell + 6, *cc = c.cell + 9;
  auto cross = [](const double* p, const double* q, double* r) {
    r[0] = p[1] * q[2] - p[2] * q[1], r[1] = p[2] * q[0] - p[0] * q[2], r[2] = p[0] * q[1] - p[1] * q[0];
  };
  auto dot = [](const double* p, const double* q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
  cross(b, cc, g.bc);
  cross(cc, a, g.ca);
  cross(a, b, g.ab);
  g.vsigned = dot(a, g.bc);
  const double v = std::fabs(g.vsigned);
  if (!(v > 1e-9)) throw Error(-1, "the pair list needs a periodic cell");
  const double w[3] = {v / std::sqrt(dot(g.bc, g.bc)), v / std::sqrt(dot(g.ca, g.ca)), v / std::sqrt(dot(g.ab, g.ab))};
  const double side = std::cbrt(4.0 * v / std::max<double>(1.0, double(n)));
  g.nu = std::clamp(int(w[0] / side), 1, 1024);
  g.nv = std::clamp(int(w[1] / side), 1, 1024);
  g.ncols = g.nu * g.nv;
  g.rl = c.opt.cutoff + c.opt.skin;
  const double rl = g.rl;
  g.ru = int(std::floor(rl * g.nu / w[0])) + 1;
  g.rv = int(std::floor(rl * g.nv / w[1])) + 1;
  g.rq = int(std::ceil(rl / w[2])) + 1;
  g.dw = rl / w[2];
  g.qlo[0] = int(std::floor(double(-g.ru) / g.nu));
  g.qlo[1] = int(std::floor(double(-g.rv) / g.nv));
  g.qlo[2] = -g.rq;
  g.span[0] = int(std::floor(double(g.nu - 1 + g.ru) / g.nu)) - g.qlo[0] + 1;
  g.span[1] = int(std::floor(double(g.nv - 1 + g.rv) / g.nv)) - g.qlo[1] + 1;
  g.span[2] = 2 * g.rq + 1;
  if (size_t(g.span[0]) * g.span[1] * g.span[2] > 256)
    throw Error(-1, "the cell is too small for the cut-off: more than 256 periodic images in range");
  g.ncodes = g.span[0] * g.span[1] * g.span[2];
  g.nwin = (2 * g.ru + 1) * (2 * g.rv + 1);
  g.rl2 = rl * rl;
  g.rl2f = float(g.rl2 * (1 + 1e-5));
  std::vector<int32_t> qint;
  for (int ia = 0; ia < g.span[0]; ++ia)
    for (int ib = 0; ib < g.span[1]; ++ib)
      for (int ic = 0; ic < g.span[2]; ++ic) qint.insert(qint.end(), {g.qlo[0] + ia, g.qlo[1] + ib, g.qlo[2] + ic});
  std::vector<double> qvec(qint.size());
  std::vector<int32_t> mirror(size_t(g.ncodes), -1);
  for (int q = 0; q < g.ncodes; ++q) {
    for (int k = 0; k < 3; ++k)
      qvec[3 * q + k] = (a[k] * qint[3 * q] + b[k] * qint[3 * q + 1]) + cc[k] * qint[3 * q + 2];
    if (qint[3 * q] == 0 && qint[3 * q + 1] == 0 && qint[3 * q + 2] == 0) g.zero_code = q;
    const int ma = -qint[3 * q] - g.qlo[0], mb = -qint[3 * q + 1] - g.qlo[1], mc = -qint[3 * q + 2] - g.qlo[2];
    if (ma >= 0 && ma < g.span[0] && mb >= 0 && mb < g.span[1] && mc >= 0 && mc < g.span[2]) mirror[size_t(q)] = (ma * g.span[1] + mb) * g.span[2] + mc;
  }
  for (int q = 0; q < g.ncodes; ++q)
    if (mirror[size_t(q)] < 0) throw Error(-1, "the cell is too small for the cut-off: an image has no opposite in range");
  L.qint.upload(qint, s);
  L.qvec.upload(qvec, s);
  std::vector<int8_t> q8(4 * size_t(g.ncodes), 0);
  for (int q = 0; q < g.ncodes; ++q)
    for (int k = 0; k < 3; ++k) q8[4 * size_t(q) + k] = int8_t(qint[3 * size_t(q) + k]);
  L.qint8.upload(q8, s);
  L.mirror.upload(mirror, s);

  const GeomArgs A = geom_args(g);
  const unsigned bs = 256;
  // binning, column sort
  L.fr.resize(3 * size_t(n));
  L.col.resize(size_t(n));
  L.colcnt.resize(size_t(g.ncols) + 1);
  L.colcnt.zero(s);
  hipLaunchKernelGGL(k_bin, dim3(blocks_for(n, bs)), dim3(bs), 0, s, A, c.x.p, L.fr.p, L.col.p, L.colcnt.p, c.flags.p + 3);
  LAUNCHCHECK();
  L.cstart.resize(size_t(g.ncols) + 1);
  c.scan.exclusive(L.colcnt.p, L.cstart.p, size_t(g.ncols), s);
  L.bfill.resize(size_t(g.ncols));
  L.bfill.zero(s);
  L.bucket.resize(size_t(n));
  hipLaunchKernelGGL(k_bucket, dim3(blocks_for(n, bs)), dim3(bs), 0, s, n, L.col.p, L.cstart.p, L.bfill.p, L.bucket.p);
  LAUNCHCHECK();
  L.rank.resize(size_t(n));
  hipLaunchKernelGGL(k_rank, dim3(blocks_for(n, bs)), dim3(bs), 0, s, n, L.col.p, L.cstart.p, L.bucket.p, L.fr.p, L.rank.p);
  LAUNCHCHECK();
  L.ccls.resize(size_t(g.ncols));
  hipLaunchKernelGGL(k_col_clusters, dim3(blocks_for(g.ncols, bs)), dim3(bs), 0, s, g.ncols, L.colcnt.p, L.ccls.p);
  LAUNCHCHECK();
  L.ccl.resize(size_t(g.ncols) + 1);
  c.scan.exclusive(L.ccls.p, L.ccl.p, size_t(g.ncols), s);
  int64_t ncl = 0;
  HIPCHECK(hipMemcpyAsync(&ncl, L.ccl.p + g.ncols, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  g.nclusters = ncl;
  g.ns = 4 * ncl;
  const int64_t ns = g.ns;
  L.perm.resize(size_t(ns));
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ns, bs)), dim3(bs), 0, s, L.perm.p, size_t(ns), kPad);
  LAUNCHCHECK();
  L.iperm.resize(size_t(n));
  hipLaunchKernelGGL(k_perm, dim3(blocks_for(n, bs)), dim3(bs), 0, s, n, L.col.p, L.rank.p, L.ccl.p, L.perm.p, L.iperm.p);
  LAUNCHCHECK();
  L.ccol.resize(size_t(ncl));
  hipLaunchKernelGGL(k_ccol, dim3(blocks_for(g.ncols, bs)), dim3(bs), 0, s, g.ncols, L.ccl.p, L.ccol.p);
  LAUNCHCHECK();
  // slots and clusters
  L.cw.resize(3 * size_t(ns));
  L.frs.resize(3 * size_t(ns));
  hipLaunchKernelGGL(k_slots, dim3(blocks_for(ns, bs)), dim3(bs), 0, s, A, ns, L.perm.p, L.fr.p, L.cw.p, L.frs.p);
  LAUNCHCHECK();
  L.blo.resize(3 * size_t(ncl));
  L.bhi.resize(3 * size_t(ncl));
  L.org.resize(3 * size_t(ncl));
  L.wlo.resize(size_t(ncl));
  L.whi.resize(size_t(ncl));
  L.bx.resize(size_t(ns));
  L.by.resize(size_t(ns));
  L.bz.resize(size_t(ns));
  L.real.resize(size_t(ncl));
  hipLaunchKernelGGL(k_clusters, dim3(blocks_for(ncl, bs)), dim3(bs), 0, s, ncl, L.perm.p, L.cw.p, L.frs.p, L.blo.p, L.bhi.p, L.wlo.p,
                     L.whi.p, L.org.p, L.bx.p, L.by.p, L.bz.p, L.real.p);
  LAUNCHCHECK();
  // exclusions by cluster pair
  L.xcount.resize(size_t(ncl));
  hipLaunchKernelGGL(k_excl_count, dim3(blocks_for(ncl, bs)), dim3(bs), 0, s, ncl, L.perm.p, c.d_excl_off.p, L.xcount.p);
  LAUNCHCHECK();
  L.xoff.resize(size_t(ncl) + 1);
  c.scan.exclusive(L.xcount.p, L.xoff.p, size_t(ncl), s);
  int64_t nx = 0;
  HIPCHECK(hipMemcpyAsync(&nx, L.xoff.p + ncl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  L.xcj.resize(size_t(std::max<int64_t>(nx, 1)));
  L.xbits.resize(size_t(std::max<int64_t>(nx, 1)));
  L.x14.resize(size_t(std::max<int64_t>(nx, 1)));
  L.xcnt.resize(size_t(ncl));
  // the excluded lanes stay in the entries (the kernel's Ewald / DSF corrections), as the evaluator builds its list
  // whenever every 1-4 pair keeps the same Coulomb fraction (core/src/field.cpp:1293-1296, 878-889)
  const int keep = 1;
  hipLaunchKernelGGL(k_excl_fill, dim3(blocks_for(ncl, bs)), dim3(bs), 0, s, ncl, L.perm.p, L.iperm.p, c.d_excl_off.p, c.d_excl_idx.p,
                     c.p14_off.empty() ? nullptr : c.d_p14_off.p, c.p14_off.empty() ? nullptr : c.d_p14_idx.p, keep, L.xoff.p, L.xcj.p,
                     L.xbits.p, L.x14.p, L.xcnt.p);
  LAUNCHCHECK();
  // the candidate search: count, scan, write
  SearchArgs P{A, ncl, L.ccol.p, L.ccl.p, L.blo.p, L.bhi.p, L.wlo.p, L.whi.p, L.org.p, L.cw.p, L.frs.p, L.bx.p, L.by.p, L.bz.p, L.real.p,
               L.perm.p, L.xoff.p, L.xcnt.p, L.xcj.p, L.xbits.p, L.x14.p, keep};
  const size_t nw = size_t(ncl) * size_t(g.nwin);
  L.wcount.resize(nw);
  hipLaunchKernelGGL(k_search<false>, dim3(blocks_for(nw, bs)), dim3(bs), 0, s, P, L.wcount.p, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr);
  LAUNCHCHECK();
  L.woff.resize(nw + 1);
  c.scan.exclusive(L.wcount.p, L.woff.p, nw, s);
  int64_t ne = 0;
  HIPCHECK(hipMemcpyAsync(&ne, L.woff.p + nw, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  g.nentries = ne;
  if (ne >= int64_t(0x7fffffff)) throw Error(-1, "the pair list has more than 2^31 entries");
  const size_t ne1 = size_t(std::max<int64_t>(ne, 1));
  L.e_cj.resize(ne1);
  L.e_ci.resize(ne1);
  L.e_code.resize(ne1);
  L.e_mask.resize(ne1);
  L.e_ex.resize(ne1);
  L.e_ex14.resize(ne1);
  hipLaunchKernelGGL(k_search<true>, dim3(blocks_for(nw, bs)), dim3(bs), 0, s, P, nullptr, L.woff.p, L.e_cj.p, L.e_ci.p, L.e_code.p,
                     L.e_mask.p, L.e_ex.p, L.e_ex14.p);
  LAUNCHCHECK();
  L.first.resize(size_t(ncl) + 1);
  hipLaunchKernelGGL(k_first, dim3(blocks_for(ncl + 1, bs)), dim3(bs), 0, s, ncl, g.nwin, L.woff.p, L.first.p);
  LAUNCHCHECK();
  // offsets, per-slot parameters
  L.off.resize(3 * size_t(ns));
  L.ioff.resize(3 * size_t(ns));
  hipLaunchKernelGGL(k_offsets, dim3(blocks_for(ns, bs)), dim3(bs), 0, s, A, ns, L.perm.p, L.cw.p, c.x.p, L.off.p, L.ioff.p);
  LAUNCHCHECK();
  L.xq.resize(size_t(ns));
  L.pp.resize(size_t(ns));
  hipLaunchKernelGGL(k_slot_params, dim3(blocks_for(ns, bs)), dim3(bs), 0, s, ns, L.perm.p, c.d_type.p, c.d_charge.p,
                     c.comb ? c.d_f12.p : nullptr, c.comb ? c.d_f6.p : nullptr, L.xq.p, L.pp.p);
  LAUNCHCHECK();
  // the full outer list
  L.own.resize(size_t(ncl));
  L.mir.resize(size_t(ncl));
  L.mfill.resize(size_t(ncl));
  L.own.zero(s);
  L.mir.zero(s);
  L.mfill.zero(s);
  if (ne > 0) {
    hipLaunchKernelGGL(k_full_count, dim3(blocks_for(ne, bs)), dim3(bs), 0, s, ne, g.zero_code, L.e_cj.p, L.e_ci.p, L.e_code.p, L.own.p,
                       L.mir.p);
    LAUNCHCHECK();
  }
  L.fcount.resize(size_t(ncl));
  hipLaunchKernelGGL(k_full_sum, dim3(blocks_for(ncl, bs)), dim3(bs), 0, s, ncl, L.own.p, L.mir.p, L.fcount.p);
  LAUNCHCHECK();
  L.fstart.resize(size_t(ncl) + 1);
  c.scan.exclusive(L.fcount.p, L.fstart.p, size_t(ncl), s);
  int64_t nfull = 0;
  HIPCHECK(hipMemcpyAsync(&nfull, L.fstart.p + ncl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  L.fsrc.resize(size_t(std::max<int64_t>(nfull, 1)));
  if (ne > 0) {
    hipLaunchKernelGGL(k_full_fill, dim3(blocks_for(ne, bs)), dim3(bs), 0, s, ne, g.zero_code, L.e_cj.p, L.e_ci.p, L.e_code.p, L.first.p,
                       L.fstart.p, L.own.p, L.mfill.p, L.fsrc.p);
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_full_sort, dim3(unsigned(ncl)), dim3(256), 0, s, ncl, L.fstart.p, L.own.p, L.fsrc.p);
    LAUNCHCHECK();
  }
  L.fnb.resize(3 * size_t(ns));
  L.built = true;
}

}  // namespace capsgpu
