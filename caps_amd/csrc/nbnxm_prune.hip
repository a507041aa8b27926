// Per-step slot coordinates (core/src/nbnxm.cpp:378-390) and dynamic pruning (core/src/nbnxm.cpp:328-363) on the GPU,
// bit-identical to the reference (built with -ffp-contract=off), then the pruned full list the pair kernel walks:
// the outer full list filtered to the entries that keep a lane, in order (a stable compaction: a scan of keep flags).
#include "context.hpp"

namespace capsgpu {
namespace {

// X = float(x + off − org) per slot, padding on the cluster's first atom; the charge stays in .w
__global__ void k_positions(int64_t ns, const uint32_t* __restrict__ perm, const double* __restrict__ x, const double* __restrict__ off,
                            const double* __restrict__ org, float4* __restrict__ xq) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int64_t r = perm[s] != kPad ? s : (s & ~int64_t(3));
  const size_t a = 3 * size_t(perm[r]);
  const double* o = org + 3 * (s / 4);
  float4 v = xq[s];
  v.x = float(x[a] + off[3 * r] - o[0]);
  v.y = float(x[a + 1] + off[3 * r + 1] - o[1]);
  v.z = float(x[a + 2] + off[3 * r + 2] - o[2]);
  xq[s] = v;
}

// the lanes of each outer entry within `radius` at the current slot coordinates
__global__ void k_prune(int64_t ne, float r2, const uint32_t* __restrict__ e_cj, const uint32_t* __restrict__ e_ci,
                        const uint16_t* __restrict__ e_code, const uint16_t* __restrict__ e_mask, const uint16_t* __restrict__ e_ex,
                        const uint16_t* __restrict__ e_ex14, const double* __restrict__ qvec, const double* __restrict__ org,
                        const float4* __restrict__ xq, uint16_t* __restrict__ pmask, uint16_t* __restrict__ pex,
                        uint16_t* __restrict__ pex14) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const size_t ci = e_ci[e], cj = e_cj[e];
  const double* sq = qvec + 3 * size_t(e_code[e]);
  const double* oi = org + 3 * ci;
  const double* oj = org + 3 * cj;
  const float sx = float(sq[0] + oj[0] - oi[0]), sy = float(sq[1] + oj[1] - oi[1]), sz = float(sq[2] + oj[2] - oi[2]);
  const unsigned both = unsigned(e_mask[e]) | unsigned(e_ex[e]);
  float xj[4], yj[4], zj[4];
  for (int l = 0; l < 4; ++l) {
    const float4 p = xq[4 * cj + l];
    xj[l] = p.x + sx, yj[l] = p.y + sy, zj[l] = p.z + sz;
  }
  unsigned in = 0;
  for (int a = 0; a < 4; ++a) {
    const unsigned rowbits = (both >> (4 * a)) & 15u;
    if (!rowbits) continue;
    const float4 pi = xq[4 * ci + a];
    unsigned row = 0;
    for (int l = 0; l < 4; ++l) {
      const float dx = xj[l] - pi.x, dy = yj[l] - pi.y, dz = zj[l] - pi.z;
      const float zz = dz * dz;
      const float yy = dy * dy + zz;
      const float d2 = dx * dx + yy;
      row |= (d2 < r2 ? 1u : 0u) << l;
    }
    in |= (row & rowbits) << (4 * a);
  }
  pmask[e] = uint16_t(e_mask[e] & in);
  pex[e] = uint16_t(e_ex[e] & in);
  pex14[e] = uint16_t(e_ex14[e] & in);
}

// bit 4a + l -> bit 4l + a
__device__ inline unsigned transpose16(unsigned m) {
  unsigned t = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int l = 0; l < 4; ++l) t |= ((m >> (4 * a + l)) & 1u) << (4 * l + a);
  return t;
}

struct FullEntry {
  uint32_t cj;
  unsigned code, mask, ex, ex14;
};

// one entry of the full list from its source half entry: as is, mirrored (the other side: transposed lanes, opposite
// image), or a cluster with itself at the zero image made symmetric
__device__ inline FullEntry full_entry(uint32_t src, int zero_code, const uint32_t* e_cj, const uint32_t* e_ci, const uint16_t* e_code,
                                       const uint16_t* pm, const uint16_t* pe, const uint16_t* p14, const int32_t* mirror) {
  const uint32_t e = src & 0x7fffffffu;
  FullEntry f;
  const unsigned m = pm[e], x = pe[e], x14 = p14[e];
  if (src & 0x80000000u) {
    f.cj = e_ci[e], f.code = unsigned(mirror[e_code[e]]);
    f.mask = transpose16(m), f.ex = transpose16(x), f.ex14 = transpose16(x14);
  } else {
    f.cj = e_cj[e], f.code = e_code[e];
    if (e_cj[e] == e_ci[e] && e_code[e] == zero_code) f.mask = m | transpose16(m), f.ex = x | transpose16(x), f.ex14 = x14 | transpose16(x14);
    else f.mask = m, f.ex = x, f.ex14 = x14;
  }
  return f;
}

__global__ void k_full_flags(int64_t nf, int zero_code, const uint32_t* __restrict__ fsrc, const uint32_t* __restrict__ e_cj,
                             const uint32_t* __restrict__ e_ci, const uint16_t* __restrict__ e_code, const uint16_t* __restrict__ pm,
                             const uint16_t* __restrict__ pe, const uint16_t* __restrict__ p14, const int32_t* __restrict__ mirror,
                             uint32_t* __restrict__ flag) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const uint32_t e = fsrc[k] & 0x7fffffffu;
  flag[k] = (pm[e] | pe[e]) ? 1u : 0u;
}

__global__ void k_full_write(int64_t nf, int zero_code, const uint32_t* __restrict__ fsrc, const uint32_t* __restrict__ e_cj,
                             const uint32_t* __restrict__ e_ci, const uint16_t* __restrict__ e_code, const uint16_t* __restrict__ pm,
                             const uint16_t* __restrict__ pe, const uint16_t* __restrict__ p14, const int32_t* __restrict__ mirror,
                             const uint32_t* __restrict__ flag, const int64_t* __restrict__ pos, uint2* __restrict__ out,
                             uint16_t* __restrict__ out14) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nf || !flag[k]) return;
  const FullEntry f = full_entry(fsrc[k], zero_code, e_cj, e_ci, e_code, pm, pe, p14, mirror);
  const int64_t o = pos[k];
  out[o] = make_uint2(f.cj | (f.code << 24), f.mask | (f.ex << 16));
  out14[o] = uint16_t(f.ex14);
}

__global__ void k_pfirst(int64_t ncl, const int64_t* __restrict__ fstart, const int64_t* __restrict__ pos, int64_t* __restrict__ pfirst) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c <= ncl) pfirst[c] = pos[fstart[c]];
}

}  // namespace

void list_positions(capsgpu_ctx& c) {
  DevList& L = c.L;
  const int64_t ns = L.g.ns;
  hipLaunchKernelGGL(k_positions, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, L.perm.p, c.x.p, L.off.p, L.org.p, L.xq.p);
  LAUNCHCHECK();
}

void list_prune(capsgpu_ctx& c, double radius) {
  DevList& L = c.L;
  hipStream_t s = c.stream;
  const int64_t ne = L.g.nentries, ncl = L.g.nclusters;
  const size_t ne1 = size_t(std::max<int64_t>(ne, 1));
  L.pmask.resize(ne1);
  L.pex.resize(ne1);
  L.pex14.resize(ne1);
  if (ne > 0) {
    hipLaunchKernelGGL(k_prune, dim3(blocks_for(ne, 256)), dim3(256), 0, s, ne, float(radius * radius), L.e_cj.p, L.e_ci.p, L.e_code.p,
                       L.e_mask.p, L.e_ex.p, L.e_ex14.p, L.qvec.p, L.org.p, L.xq.p, L.pmask.p, L.pex.p, L.pex14.p);
    LAUNCHCHECK();
  }
  int64_t nf = 0;
  HIPCHECK(hipMemcpyAsync(&nf, L.fstart.p + ncl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  L.pflag.resize(size_t(std::max<int64_t>(nf, 1)));
  L.ppos.resize(size_t(nf) + 1);
  if (nf > 0) {
    hipLaunchKernelGGL(k_full_flags, dim3(blocks_for(nf, 256)), dim3(256), 0, s, nf, L.g.zero_code, L.fsrc.p, L.e_cj.p, L.e_ci.p,
                       L.e_code.p, L.pmask.p, L.pex.p, L.pex14.p, L.mirror.p, L.pflag.p);
    LAUNCHCHECK();
  }
  c.scan.exclusive(L.pflag.p, L.ppos.p, size_t(nf), s);
  int64_t np = 0;
  HIPCHECK(hipMemcpyAsync(&np, L.ppos.p + nf, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  L.npruned = np;
  L.pent.resize(size_t(std::max<int64_t>(np, 1)));
  L.pent14.resize(size_t(std::max<int64_t>(np, 1)));
  if (nf > 0) {
    hipLaunchKernelGGL(k_full_write, dim3(blocks_for(nf, 256)), dim3(256), 0, s, nf, L.g.zero_code, L.fsrc.p, L.e_cj.p, L.e_ci.p,
                       L.e_code.p, L.pmask.p, L.pex.p, L.pex14.p, L.mirror.p, L.pflag.p, L.ppos.p, L.pent.p, L.pent14.p);
    LAUNCHCHECK();
  }
  L.pfirst.resize(size_t(ncl) + 1);
  hipLaunchKernelGGL(k_pfirst, dim3(blocks_for(ncl + 1, 256)), dim3(256), 0, s, ncl, L.fstart.p, L.ppos.p, L.pfirst.p);
  LAUNCHCHECK();
}

}  // namespace capsgpu
