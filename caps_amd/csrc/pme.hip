// Smooth PME reciprocal space on gfx950 in FP32 (the mesh), with FP64 energy and virial sums — the pipeline of
// core/src/kspace.cpp:80-175 and core/fortran/caps_kspace.f90 (splines :776-796, spread :800-822, r2c/c2r lines
// :392-461, FFT lines :348-387, influence function :518-564 / :880-928, gather :931-961).
//
//  k_splines      per atom: fractional grid coordinate, first grid point, B-spline weights and derivatives; home tile
//  k_tile_*       atoms binned by the 4x4x4 grid tile of their first point (sorted by atom in each tile)
//  k_spread_blk   one 256-thread workgroup per block of 2×2×2 tiles (every K a multiple of 4): each atom of the home
//                 tiles that reach the block staged once (splines, weight windows) in LDS, then each wave sums two
//                 tiles, a lane taking a row of 4 points over a quarter of the atoms (every point written once, no
//                 atomics, a fixed order)
//  k_spread       the same per single tile (64 lanes) for the other grids
//  k_fft_x_r2c    two real lines per complex transform along a (as the reference), the half spectrum written
//  k_fft_y        complex lines along b, 16 lines of consecutive a-frequencies per workgroup (coalesced)
//  k_fft_xy       a whole c-plane per workgroup when it fits in LDS: the two passes above (or k_fft_y and
//                 k_fft_x_c2r for the inverse) with the plane kept in LDS between them, one read and one write of it
//  k_fft_z_conv   lines along c: forward transform, × influence function (energy and virial on energy steps),
//                 inverse transform — the convolution without a round trip through HBM
//  k_fft_x_c2r    back to the real grid, two lines per complex transform
//  k_gather       per tile: its mesh box in LDS, p lanes per atom: F = −q Σ_d K_d h⁻¹_d ∂_d over its p³ points
// The FFT is a mixed-radix (2, 3, 4, 5, 7, 8) Stockham transform in LDS, X(m) = Σ x(k) exp(sign 2πi mk/n),
// forward sign +1 as the reference; twiddles from a table computed in double on the host.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "context.hpp"
#include "pme_atoms.hpp"
#include "prof.hpp"

namespace capsgpu {
namespace {

constexpr int kMaxOrder = 8;
constexpr int kMaxFft = 512;

__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// B-spline weights θ[j] = M_p(w + p − 1 − j) … in the reference's order (caps_kspace.f90:568-598), FP32; the order
// P a compile-time constant, so every loop unrolls and the arrays stay in registers
template <int P>
__device__ inline void bspline(float w, float* th, float* dth) {
#pragma unroll
  for (int j = 0; j < P; ++j) th[j] = 0.0f;
  th[1] = w;
  th[0] = 1.0f - w;
#pragma unroll
  for (int k = 3; k <= P - 1; ++k) {
    const float div = 1.0f / float(k - 1);
    th[k - 1] = div * w * th[k - 2];
#pragma unroll
    for (int j = 1; j <= k - 2; ++j) th[k - j - 1] = div * ((w + float(j)) * th[k - j - 2] + (float(k - j) - w) * th[k - j - 1]);
    th[0] = div * (1.0f - w) * th[0];
  }
  dth[0] = -th[0];
#pragma unroll
  for (int j = 2; j <= P; ++j) dth[j - 1] = th[j - 2] - th[j - 1];
  const float div = 1.0f / float(P - 1);
  th[P - 1] = div * w * th[P - 2];
#pragma unroll
  for (int j = 1; j <= P - 2; ++j) th[P - j - 1] = div * ((w + float(j)) * th[P - j - 2] + (float(P - j) - w) * th[P - j - 1]);
  th[0] = div * (1.0f - w) * th[0];
}

struct GridArgs {
  int k[3], nt[3], p, h1, hs, ntm;   // hs: half_stride(K1); ntm: the largest of nt (the home table's stride)
  int lx;                        // the transforms' tile shape: pairs of real lines per workgroup along a,
  int ly, ly_log2;               // and lines per workgroup along b and c (powers of two)
  double hi[3][3];
};

// the arrays of the per-atom half of the bookkeeping for the mesh of G
__device__ __host__ inline AtomSplines atom_splines(const GridArgs& G, const double* q, float4* wq, int4* base, uint32_t* tcount,
                                                    uint32_t* bucket, int cap, uint32_t* ovf, uint32_t* cnt) {
  AtomSplines A;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) A.hi[r][c] = G.hi[r][c];
    A.k[r] = G.k[r], A.nt[r] = G.nt[r];
  }
  A.p = G.p, A.cap = cap, A.q = q, A.wq = wq, A.base = base, A.tcount = tcount, A.bucket = bucket, A.ovf = ovf, A.cnt = cnt;
  return A;
}
// Per atom: its splines' offsets w, charge, first mesh point per axis and home tile (the 4x4x4 tile of that point), and
// a place in the home tile's bucket of cap slots (spline_atom). An atom past its bucket's cap goes to the overflow list
// (ovf, its length in cnt[0]): correct for any filling, slow only where a tile holds more than cap atoms, which the
// cap's margin over the mean makes a rarity in a liquid or solid.
__device__ __forceinline__ void b_splines(int bx, int gx, GridArgs G, int n, const double* __restrict__ x, const double* __restrict__ q,
                                          float4* __restrict__ wq, int4* __restrict__ base, uint32_t* __restrict__ tcount,
                                          uint32_t* __restrict__ bucket, int cap, uint32_t* __restrict__ ovf, uint32_t* __restrict__ cnt) {
  const int i = bx * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const double xi[3] = {valid ? x[3 * i] : 0.0, valid ? x[3 * i + 1] : 0.0, valid ? x[3 * i + 2] : 0.0};
  spline_atom(atom_splines(G, q, wq, base, tcount, bucket, cap, ovf, cnt), valid, i, xi);
}
__global__ void k_splines(GridArgs G, int n, const double* __restrict__ x, const double* __restrict__ q, float4* __restrict__ wq,
                          int4* __restrict__ base, uint32_t* __restrict__ tcount, uint32_t* __restrict__ bucket, int cap,
                          uint32_t* __restrict__ ovf, uint32_t* __restrict__ cnt) {
  b_splines(blockIdx.x, gridDim.x, G, n, x, q, wq, base, tcount, bucket, cap, ovf, cnt);
}

// The atoms of each tile in index order: an atom's place is the tile's first place (the scan of the counts) plus its
// rank among the tile's atoms — those in its bucket (filled unordered) and, for a tile past its cap, those in the
// overflow list; its spline offsets, charge and base point are copied to that place too, so the spread and the gather
// read a tile's atoms contiguously. tstart: the tiles' first places, or with sums (the scan's block totals, nsums of
// them) the offsets within each scan block of kScanTile tiles: then the block offsets are added here, and the tiles'
// first places written to tout (0 … ntiles) for the kernels after this one. The tile counts (read by the scan, done)
// are cleared here for the next evaluation; the overflow count, read here, by the spread that follows.
__device__ __forceinline__ void b_tile_rank(int bx, int gx, int n, int ntiles, const float4* __restrict__ wq, const int4* __restrict__ base,
                                            const int64_t* __restrict__ tstart, const uint32_t* __restrict__ bucket, int cap,
                                            const uint32_t* __restrict__ ovf, uint32_t* __restrict__ cnt, uint32_t* __restrict__ tatoms,
                                            float4* __restrict__ wsort, int4* __restrict__ bsort, uint32_t* __restrict__ tcount,
                                            const int64_t* __restrict__ sums = nullptr, int nsums = 0, int64_t* __restrict__ tout = nullptr) {
  const int i = bx * blockDim.x + threadIdx.x;
  // the atom's data requested first, so its latency passes under the block offsets' sums
  int4 bi = make_int4(0, 0, 0, 0);
  float4 wi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < n) bi = base[i], wi = wq[i];
  __shared__ int64_t boff[64];
  if (sums) {
    block_offsets(sums, nsums, boff);
    for (int u = i; u <= ntiles; u += gx * blockDim.x) tout[u] = tstart[u] + boff[u / kScanTile];
  }
  for (int t = i; t < ntiles; t += gx * blockDim.x) tcount[t] = 0u;
  const uint32_t novf = cnt[0];
  if (i < n) {
    const int t = bi.w;
    const uint32_t* bk = bucket + size_t(t) * cap;
    // the bucket's first 16 slots requested with the tile's places, before its count is known (a bucket holds at least
    // 32 slots, a multiple of 8, so these loads stay inside it and are 16-byte aligned): most tiles hold fewer atoms
    uint4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = reinterpret_cast<const uint4*>(bk)[u];
    int64_t b = tstart[t], e = tstart[t + 1];
    if (sums) b += boff[t / kScanTile], e += boff[(t + 1) / kScanTile];
    const int64_t m = e - b;
    const int nb = int(m < cap ? m : cap);
    int64_t r = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      r += (4 * u + 0 < nb && q[u].x < uint32_t(i)) ? 1 : 0;
      r += (4 * u + 1 < nb && q[u].y < uint32_t(i)) ? 1 : 0;
      r += (4 * u + 2 < nb && q[u].z < uint32_t(i)) ? 1 : 0;
      r += (4 * u + 3 < nb && q[u].w < uint32_t(i)) ? 1 : 0;
    }
    int k = 16;
    for (; k + 8 <= nb; k += 8) {   // eight loads in flight
      uint32_t t8[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t8[u] = bk[k + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) r += t8[u] < uint32_t(i) ? 1 : 0;
    }
    for (; k < nb; ++k) r += bk[k] < uint32_t(i) ? 1 : 0;
    if (m > cap)
      for (uint32_t o = 0; o < novf; ++o) {
        const uint32_t j = ovf[o];
        r += (base[j].w == t && j < uint32_t(i)) ? 1 : 0;
      }
    tatoms[b + r] = uint32_t(i);
    wsort[b + r] = wi;
    bsort[b + r] = bi;
  }
}
__global__ void k_tile_rank(int n, int ntiles, const float4* __restrict__ wq, const int4* __restrict__ base,
                            const int64_t* __restrict__ tstart, const uint32_t* __restrict__ bucket, int cap, const uint32_t* __restrict__ ovf,
                            uint32_t* __restrict__ cnt, uint32_t* __restrict__ tatoms, float4* __restrict__ wsort, int4* __restrict__ bsort,
                            uint32_t* __restrict__ tcount, const int64_t* __restrict__ sums, int nsums, int64_t* __restrict__ tout) {
  b_tile_rank(blockIdx.x, gridDim.x, n, ntiles, wq, base, tstart, bucket, cap, ovf, cnt, tatoms, wsort, bsort, tcount, sums, nsums, tout);
}

// one 64-lane workgroup per 4x4x4 tile. The atoms whose first point lies in a tile that can reach this one (the home
// table lists those tiles per axis) are staged in LDS 64 at a time. With r_d = 4 t_d − b_d the offset of an atom's
// first point from the tile's corner (folded into [−3, p]), a point (i, j, k) of the tile sees θa[i + r_a],
// θb[j + r_b], θc[k + r_c]. Staging (one atom per lane): the four θa[i + r_a] of the tile's points along a and the
// sixteen row weights θb[j + r_b]·(q θc[k + r_c]) of the tile's rows (j, k), zero outside the support; each window of
// four is read from the lane's own zero-padded column of LDS scratch at offset r + 3, so no selects (column-major:
// element e of lane L at 64 e + L, every access of the wave one word per bank). Lane = (row (j, k), quarter of the
// atoms): per atom one row weight, one float4 of θa and four FMAs; the quarters are then added in a fixed order
// (every point written once, no atomics). An atom's row weights take 20 floats (16 and a pad: its float4 stores meet
// no bank twice). Static LDS 6 KB (eight workgroups per SIMD).
constexpr int kRowStride = 20;
constexpr int kSpreadChunk = 64, kMaxHome = 64;   // at most 4 home tiles per axis reach a tile (p <= 8, partial last tile)
template <int P>
__device__ __forceinline__ void b_spread(int tile, GridArgs G, const int* __restrict__ home, const int64_t* __restrict__ tstart,
                                         const float4* __restrict__ wsort, const int4* __restrict__ bsort, float* __restrict__ grid) {
  static_assert(P + 6 < 16, "a window's last read must stay inside the lane's 16 scratch elements");
  __shared__ float4 s_row[kSpreadChunk * kRowStride / 4];   // per atom: row weights (row j + 4k); first the scratch
  __shared__ float4 s_wx[kSpreadChunk];                     // per atom: θa[r_a + i] for the tile's four points i along a
  const int lane = threadIdx.x;
  const int tx = tile % G.nt[0], ty = (tile / G.nt[0]) % G.nt[1], tz = tile / (G.nt[0] * G.nt[1]);
  // the candidate tiles in a fixed order (c, b, a) from the home table, their atom counts in parallel, a running sum
  const int* hx = home + (0 * G.ntm + tx) * 5;
  const int* hy = home + (1 * G.ntm + ty) * 5;
  const int* hz = home + (2 * G.ntm + tz) * 5;
  const int mx = hx[0], my = hy[0], mz = hz[0], nh = mx * my * mz;
  // lane k < nh: home tile k's first atom (hb) and atom count; the counts' running sum gives each tile's first candidate
  // number (ex); the lanes' values are read back with readlane (the home tiles are few: at most 27 for p = 8)
  int cnt = 0, hb = 0;   // atom counts fit 32 bits (the library's atom count is an int)
  if (lane < nh) {
    // lane = ix + mx (iy + my iz) with lane < 64 and mx, my <= 4: a quotient q = ⌊lane / d⌋ is ⌊lane · rcp(d) + 1e-3⌋
    // exactly (the product's error is below 1e-5 and a non-integer quotient's fraction at most 15/16)
    const int mxy = mx * my;
    const int iz = int(float(lane) * __builtin_amdgcn_rcpf(float(mxy)) + 1e-3f), rem = lane - iz * mxy;
    const int iy = int(float(rem) * __builtin_amdgcn_rcpf(float(mx)) + 1e-3f), ix = rem - iy * mx;
    const int h = (hz[1 + iz] * G.nt[1] + hy[1 + iy]) * G.nt[0] + hx[1 + ix];
    hb = int(tstart[h]);
    cnt = int(tstart[h + 1] - tstart[h]);
  }
  const int own = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(cnt, d, 64);
    if (lane >= d) cnt += o;
  }
  const int ex = cnt - own;
  const int m = __builtin_amdgcn_readlane(cnt, nh - 1);
  // lane = (row (j, k) of the tile, quarter of the atoms); row j + 4 k = lane & 15
  const int gy = lane & 3, gz = (lane >> 2) & 3, quarter = lane >> 4;
  float* scr = reinterpret_cast<float*>(s_row) + lane;   // scratch element e of this lane: scr[64 e], e < 16
  const float* srow = reinterpret_cast<const float*>(s_row) + (lane & 15);
  float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
  for (int c0 = 0; c0 < m; c0 += kSpreadChunk) {
    const int f = c0 + lane;
    // candidate f lies in the last home tile whose first candidate number is ≤ f (empty tiles are passed over)
    int at = 0;
    for (int k = 0; k < nh; ++k) {
      const int exk = __builtin_amdgcn_readlane(ex, k), hbk = __builtin_amdgcn_readlane(hb, k);
      at = f >= exk ? hbk + (f - exk) : at;
    }
    // the scratch's zero padding around θ (elements [0, 3) and [3 + p, p + 7)); the lane's writes and reads below
    // touch its own column only, and its LDS accesses complete in order, so each read sees the write before it
#pragma unroll
    for (int e = 0; e < 3; ++e) scr[64 * e] = 0.0f;
#pragma unroll
    for (int e = 3 + P; e < P + 7; ++e) scr[64 * e] = 0.0f;
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 rw[4] = {z4, z4, z4, z4}, wxv = z4;   // a slot past the candidates adds exact zeros
    if (f < m) {
      const int4 b = bsort[at];
      const float4 wv = wsort[at];
      // r = 4t − b lies in (−K, K + 3): to [0, K) and then into [−3, K − 4], and at most p (beyond, no point is reached)
      const int bb[3] = {b.x, b.y, b.z}, tt[3] = {tx, ty, tz};
      int r[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int v = 4 * tt[d] - bb[d];
        v += v < 0 ? G.k[d] : 0;
        v -= v >= G.k[d] ? G.k[d] : 0;
        v -= v > G.k[d] - 4 ? G.k[d] : 0;
        r[d] = v < P ? v : P;
      }
      float ta[P], tb[P], tc[P], dd[P];
      bspline<P>(wv.x, ta, dd);
      bspline<P>(wv.y, tb, dd);
      bspline<P>(wv.z, tc, dd);
      // the windows: θ at elements [3, 3 + p) between zeros, four read from element r + 3 on (r + 6 <= p + 6 < 16)
#pragma unroll
      for (int d = 0; d < P; ++d) scr[64 * (3 + d)] = ta[d];
      const float* wa = scr + 64 * (r[0] + 3);
      wxv = make_float4(wa[0], wa[64], wa[128], wa[192]);
#pragma unroll
      for (int d = 0; d < P; ++d) scr[64 * (3 + d)] = tb[d];
      const float* wb = scr + 64 * (r[1] + 3);
      const float b0 = wb[0], b1 = wb[64], b2 = wb[128], b3 = wb[192];
#pragma unroll
      for (int d = 0; d < P; ++d) scr[64 * (3 + d)] = wv.w * tc[d];
      const float* wc = scr + 64 * (r[2] + 3);
      const float cw[4] = {wc[0], wc[64], wc[128], wc[192]};
#pragma unroll
      for (int k = 0; k < 4; ++k) rw[k] = make_float4(b0 * cw[k], b1 * cw[k], b2 * cw[k], b3 * cw[k]);
    }
    // every lane's scratch reads are done before the row weights overwrite the scratch
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) s_row[lane * (kRowStride / 4) + k] = rw[k];
    s_wx[lane] = wxv;
    __syncthreads();
    const int n_here = m - c0 < kSpreadChunk ? m - c0 : kSpreadChunk;
#pragma unroll 4
    for (int j0 = 0; j0 < n_here; j0 += 4) {
      const int j = j0 + quarter;   // j < 64; slots at or past n_here hold zeros
      const float v = srow[kRowStride * j];
      const float4 w = s_wx[j];
      acc0 += v * w.x, acc1 += v * w.y, acc2 += v * w.z, acc3 += v * w.w;
    }
    __syncthreads();
  }
  // the quarters' sums by point (i + 4 j + 16 k), in the staging area, then each lane adds its point's four in order
  float* s_part = reinterpret_cast<float*>(s_row);   // [4][64]
  const int row = 4 * gy + 16 * gz;
  s_part[64 * quarter + row] = acc0, s_part[64 * quarter + row + 1] = acc1, s_part[64 * quarter + row + 2] = acc2,
  s_part[64 * quarter + row + 3] = acc3;
  __syncthreads();
  const float sum = ((s_part[lane] + s_part[64 + lane]) + s_part[128 + lane]) + s_part[192 + lane];
  const int gx = 4 * tx + (lane & 3), gy_ = 4 * ty + ((lane >> 2) & 3), gz_ = 4 * tz + (lane >> 4);
  if (gx < G.k[0] && gy_ < G.k[1] && gz_ < G.k[2]) grid[(size_t(gz_) * G.k[1] + gy_) * G.k[0] + gx] = sum;
}
template <int P>
// (the spread launches also clear the bucket overflow count, ovfc, for the next evaluation: k_tile_rank, its last
// reader, is done)
__global__ void __launch_bounds__(64) k_spread(GridArgs G, const int* __restrict__ home, const int64_t* __restrict__ tstart,
                                               const float4* __restrict__ wsort, const int4* __restrict__ bsort, float* __restrict__ grid,
                                               uint32_t* __restrict__ ovfc) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ovfc = 0u;
  const int ntiles = G.nt[0] * G.nt[1] * G.nt[2];
  const int tile = int(xcd_block(blockIdx.x, ntiles));   // neighbouring tiles on one XCD (shared atoms and mesh lines)
  if (tile >= ntiles) return;
  b_spread<P>(tile, G, home, tstart, wsort, bsort, grid);
}

// ---- the spread by blocks of 2×2×2 tiles (every K a multiple of 4) ---------------------------------------------------
// k_spread stages each atom once per tile it reaches (eight tiles at order 5), its splines recomputed each time; here a
// workgroup of four waves takes a block of eight tiles and stages each atom of the home tiles that reach the block
// once: its splines (the B-spline pieces as polynomials in w, by Horner), and per axis the NTP windows of four weights
// it has on the tiles past its home tile (θa, θb, q·θc). A window is cut out by where the weights are written, not by
// selects: each staging thread owns one column of the LDS rows (row e of column f holds element e of staged atom f),
// writes its P weights into zeroed rows s … s + P − 1 (s its first point's place in its tile) and, for θa, reads its two
// float4 windows back. Every read of the accumulation meets distinct banks or one broadcast. Each wave then sums two
// tiles, lane = (rows (j, k) and (j + 2, k) of the tile, eighth of the candidates): per candidate two row weights θb·qθc
// and one float4 of θa from LDS, eight FMAs, in wave-uniform rounds of eight candidates over each run of home tiles. A
// tile's candidates are walked home tile by home tile (each with fixed window offsets), in (c, b, a) order; the eighths
// are added in a fixed order and every mesh point is written once.
constexpr int kBlkThreads = 256, kBlkStage = 256, kBlkPitch = kBlkStage + 8;   // atoms staged per chunk: at order 5 the
                                                                                 // block's 27 home tiles hold ~216

// The weights θ[j](w) of order P as polynomials, θ[j] = Σ_m c[j][m] w^m: bspline()'s recursion carried out on the
// coefficients in double, then rounded to float.
template <int P>
struct BsplinePoly {
  float c[P][P];
};
template <int P>
constexpr BsplinePoly<P> bspline_poly() {
  double t[P][P] = {};
  t[0][0] = 1.0, t[0][1] = -1.0, t[1][1] = 1.0;   // θ[0] = 1 − w, θ[1] = w
  for (int k = 3; k <= P; ++k) {
    const double div = 1.0 / double(k - 1);
    for (int m = P - 1; m >= 1; --m) t[k - 1][m] = div * t[k - 2][m - 1];   // θ[k−1] = w θ[k−2] / (k−1)
    t[k - 1][0] = 0.0;
    for (int j = 1; j <= k - 2; ++j) {   // θ[k−j−1] = ((w + j) θ[k−j−2] + (k − j − w) θ[k−j−1]) / (k−1)
      double r[P] = {};
      for (int m = 0; m < P; ++m) {
        r[m] += double(j) * t[k - j - 2][m] + double(k - j) * t[k - j - 1][m];
        if (m + 1 < P) r[m + 1] += t[k - j - 2][m] - t[k - j - 1][m];
      }
      for (int m = 0; m < P; ++m) t[k - j - 1][m] = div * r[m];
    }
    double r[P] = {};   // θ[0] = (1 − w) θ[0] / (k−1)
    for (int m = 0; m < P; ++m) {
      r[m] += t[0][m];
      if (m + 1 < P) r[m + 1] -= t[0][m];
    }
    for (int m = 0; m < P; ++m) t[0][m] = div * r[m];
  }
  BsplinePoly<P> out{};
  for (int j = 0; j < P; ++j)
    for (int m = 0; m < P; ++m) out.c[j][m] = float(t[j][m]);
  return out;
}

// θ[j](w), j < P, by Horner in FP32: P − 1 FMAs each, within a few 1e-8 of the recursion's values for w in [0, 1)
template <int P>
__device__ __forceinline__ void bspline_horner(float w, float* th) {
  constexpr BsplinePoly<P> C = bspline_poly<P>();
#pragma unroll
  for (int j = 0; j < P; ++j) {
    float p = C.c[j][P - 1];
#pragma unroll
    for (int m = P - 2; m >= 0; --m) p = __builtin_fmaf(p, w, C.c[j][m]);
    th[j] = p;
  }
}

// the block spread's workgroups: blocks of 2x2x2 tiles, the last along an axis with an odd tile count one tile deep
__host__ __device__ inline int spread_block_count(const GridArgs& G) {
  return ((G.nt[0] + 1) / 2) * ((G.nt[1] + 1) / 2) * ((G.nt[2] + 1) / 2);
}

template <int P>
__device__ __forceinline__ void b_spread_blk(int blk, const GridArgs& G, const int64_t* __restrict__ tstart, const float4* __restrict__ wsort,
                                             const int4* __restrict__ bsort, float* __restrict__ grid) {
  constexpr int NTP = (P + 6) / 4, NA = NTP + 1, NH = NA * NA * NA;   // home tiles per axis that reach the block, in all
  constexpr int NR = 8 * NTP;                                          // LDS rows of a staged atom: θb's, then q·θc's
  static_assert(NH <= 64 && NH < 256, "a lane per home tile, its index in a byte");
  static_assert(4 * NTP + 8 <= NR && P + 7 <= NR && P + 2 < 4 * NTP, "θa's rows and the zero windows either side fit the rows");
  static_assert(kBlkThreads == kBlkStage, "one staged atom per thread");
  __shared__ float4 sWa[2 * kBlkStage];   // θa of staged atom f on the block's tile column dx at dx·stage + f
  __shared__ float sW[NR * kBlkPitch];    // element e = 4o + j of the θb windows at row e (e·pitch + f), q·θc's at 4 NTP + e
  __shared__ int s_ex[NH + 1], s_hb[NH];  // home tile h's first staged place and first atom
  __shared__ uint8_t s_home[kBlkStage];   // the home tile of each staged place of the first chunk
  float* sWb = sW;
  float* sWc = sW + 4 * NTP * kBlkPitch;
  static_assert(sizeof(sW) >= 8 * 8 * 64 * sizeof(float), "the tiles' partial sums reuse the rows' space");
  float(*s_part)[8][64] = reinterpret_cast<float(*)[8][64]>(sW);   // per tile: the eighths' sums by point (after the chunks)
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  // blocks of 2x2x2 tiles; with an odd tile count along an axis the last block along it has one tile (its second
  // column's sums are not written)
  const int nbx = (G.nt[0] + 1) / 2, nby = (G.nt[1] + 1) / 2;
  const int bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
  const int tx0 = 2 * bx, ty0 = 2 * by, tz0 = 2 * bz;
  // home tile h = ix + NA (iy + NA iz): tiles t0 − NTP + 1 + i along each axis (periodic)
  if (wave == 0) {
    int cnt = 0, hb = 0;
    if (lane < NH) {
      const int ix = lane % NA, iy = (lane / NA) % NA, iz = lane / (NA * NA);
      int hx = tx0 - NTP + 1 + ix, hy = ty0 - NTP + 1 + iy, hz = tz0 - NTP + 1 + iz;
      hx += hx < 0 ? G.nt[0] : 0, hx -= hx >= G.nt[0] ? G.nt[0] : 0;
      hy += hy < 0 ? G.nt[1] : 0, hy -= hy >= G.nt[1] ? G.nt[1] : 0;
      hz += hz < 0 ? G.nt[2] : 0, hz -= hz >= G.nt[2] ? G.nt[2] : 0;
      const int h = (hz * G.nt[1] + hy) * G.nt[0] + hx;
      hb = int(tstart[h]);
      cnt = int(tstart[h + 1] - tstart[h]);
    }
    const int own = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(cnt, d, 64);
      if (lane >= d) cnt += o;
    }
    if (lane < NH) {
      s_ex[lane] = cnt - own, s_hb[lane] = hb;
      const int q1 = cnt < kBlkStage ? cnt : kBlkStage;
      for (int q = cnt - own; q < q1; ++q) s_home[q] = uint8_t(lane);
    }
    if (lane == NH - 1) s_ex[NH] = cnt;
  }
  __syncthreads();
  const int m = __builtin_amdgcn_readfirstlane(s_ex[NH]);
  // this wave's two tiles: dx = 0, 1 at (dy, dz) = (wave & 1, wave >> 1); lane = (rows j, j + 2 at k, eighth), j < 2.
  // With the rows' pitch ≡ 8 (mod 32 banks), the eight eighths' reads of rows j, j + 2 or k meet distinct banks.
  const int dy = wave & 1, dz = wave >> 1;
  const int gy = lane & 1, gz = (lane >> 1) & 3, eighth = lane >> 3;
  float acc[2][2][4] = {};   // [dx][row j, j + 2][i]
  for (int c0 = 0; c0 < m; c0 += kBlkStage) {
    const int f = c0 + tid;
    if (f < m) {
      int h;
      if (c0 == 0) {
        h = s_home[tid];
      } else {   // the last h with s_ex[h] <= f (empty tiles are passed over)
        int lo = 0, hi = NH - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_ex[mid] <= f) lo = mid;
          else hi = mid - 1;
        }
        h = lo;
      }
      const int at = s_hb[h] + (f - s_ex[h]), ix = h % NA;
      const float4 w = wsort[at];
      const int4 b = bsort[at];
      float ta[P], tb[P], tc[P];
      bspline_horner<P>(w.x, ta);
      bspline_horner<P>(w.y, tb);
      bspline_horner<P>(w.z, tc);
      const int sx = b.x & 3, sy = b.y & 3, sz = b.z & 3;
      float* col = sW + tid;
#pragma unroll
      for (int e = 0; e < NR; ++e) col[e * kBlkPitch] = 0.0f;
      // θa at rows 4 + sx + j: window o (elements 4o … 4o + 3) at rows 4 + 4o …, zero windows at o = −1 and o = NTP.
      // The block's tile column dx takes window dx + NTP − 1 − ix, at rows 4 (dx + NTP − ix) … .
      float* ca = col + (4 + sx) * kBlkPitch;
#pragma unroll
      for (int j = 0; j < P; ++j) ca[j * kBlkPitch] = ta[j];
      const float* r0 = col + 4 * (NTP - ix) * kBlkPitch;
      const float* r1 = r0 + 4 * kBlkPitch;
      sWa[tid] = make_float4(r0[0], r0[kBlkPitch], r0[2 * kBlkPitch], r0[3 * kBlkPitch]);
      sWa[kBlkStage + tid] = make_float4(r1[0], r1[kBlkPitch], r1[2 * kBlkPitch], r1[3 * kBlkPitch]);
      // θa's rows zeroed again, then θb at rows sy + j and q·θc at rows 4 NTP + sz + j
#pragma unroll
      for (int e = 4; e < P + 7; ++e) col[e * kBlkPitch] = 0.0f;
      float* cb = col + sy * kBlkPitch;
      float* cc = col + (4 * NTP + sz) * kBlkPitch;
#pragma unroll
      for (int j = 0; j < P; ++j) cb[j * kBlkPitch] = tb[j], cc[j * kBlkPitch] = w.w * tc[j];
    }
    __syncthreads();
    const int c1 = min(m, c0 + kBlkStage);
    // per home row (iy, iz): its NA home tiles along a, consecutive in the staged order, for both of the wave's tiles at
    // once — each candidate's θa staged for either tile column, zero for the column it does not reach (the home tiles
    // dx … dx + NTP − 1 reach column dx)
#pragma unroll
    for (int iz = dz; iz < dz + NTP; ++iz) {
#pragma unroll
      for (int iy = dy; iy < dy + NTP; ++iy) {
        const int h = NA * (iy + NA * iz);
        const int a = __builtin_amdgcn_readfirstlane(max(s_ex[h], c0));
        const int n = __builtin_amdgcn_readfirstlane(min(s_ex[h + NA], c1)) - a;
        if (n <= 0) continue;
        const int oy = dy + NTP - 1 - iy, oz = dz + NTP - 1 - iz;
        const float* pb = sWb + (4 * oy + gy) * kBlkPitch + (a - c0) + eighth;
        const float* pc = sWc + (4 * oz + gz) * kBlkPitch + (a - c0) + eighth;
        const float4* pa = sWa + (a - c0) + eighth;
        auto add = [&](int g, float z) {
          const float c = pc[g] * z, v0 = pb[g] * c, v1 = pb[g + 2 * kBlkPitch] * c;
          const float4 w0 = pa[g], w1 = pa[kBlkStage + g];
          acc[0][0][0] += v0 * w0.x, acc[0][0][1] += v0 * w0.y, acc[0][0][2] += v0 * w0.z, acc[0][0][3] += v0 * w0.w;
          acc[0][1][0] += v1 * w0.x, acc[0][1][1] += v1 * w0.y, acc[0][1][2] += v1 * w0.z, acc[0][1][3] += v1 * w0.w;
          acc[1][0][0] += v0 * w1.x, acc[1][0][1] += v0 * w1.y, acc[1][0][2] += v0 * w1.z, acc[1][0][3] += v0 * w1.w;
          acc[1][1][0] += v1 * w1.x, acc[1][1][1] += v1 * w1.y, acc[1][1][2] += v1 * w1.z, acc[1][1][3] += v1 * w1.w;
        };
        // rounds of eight candidates, one per eighth: every lane has one in each round but the last
        const int T = (n + 7) >> 3;
        int k = 0;
        for (; k + 2 < T; k += 2) add(8 * k, 1.0f), add(8 * k + 8, 1.0f);
        if (k + 1 < T) add(8 * k, 1.0f), ++k;
        // the last round: a lane past the run takes the run's last candidate with a zero weight (finite products of
        // zero, so its sums are unchanged)
        const bool in = 8 * k + eighth < n;
        add(in ? 8 * k : n - 1 - eighth, in ? 1.0f : 0.0f);
      }
    }
    __syncthreads();
  }
  // the eighths' sums by point (i + 4 j + 16 k) of each tile, then each lane adds its point's eight in order
  const int prow = 4 * gy + 16 * gz;
#pragma unroll
  for (int dx = 0; dx < 2; ++dx) {
    const int t = dx + 2 * dy + 4 * dz;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int i = 0; i < 4; ++i) s_part[t][eighth][prow + 8 * r + i] = acc[dx][r][i];
  }
  __syncthreads();
#pragma unroll
  for (int dx = 0; dx < 2; ++dx) {
    const int t = dx + 2 * dy + 4 * dz;
    float sum = s_part[t][0][lane];
#pragma unroll
    for (int e = 1; e < 8; ++e) sum += s_part[t][e][lane];
    const int gx = 4 * (tx0 + dx) + (lane & 3), gy_ = 4 * (ty0 + dy) + ((lane >> 2) & 3), gz_ = 4 * (tz0 + dz) + (lane >> 4);
    if (tx0 + dx < G.nt[0] && ty0 + dy < G.nt[1] && tz0 + dz < G.nt[2]) grid[(size_t(gz_) * G.k[1] + gy_) * G.k[0] + gx] = sum;
  }
}
template <int P>
__global__ void __launch_bounds__(kBlkThreads) k_spread_blk(GridArgs G, const int64_t* __restrict__ tstart, const float4* __restrict__ wsort,
                                                            const int4* __restrict__ bsort, float* __restrict__ grid, uint32_t* __restrict__ ovfc) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ovfc = 0u;
  const int nblk = spread_block_count(G);
  b_spread_blk<P>(int(xcd_block(blockIdx.x, nblk)), G, tstart, wsort, bsort, grid);
}

// ---- FFT ------------------------------------------------------------------------------------------------------------
struct Fft1 {
  int n, nr;
  uint64_t rad;   // the stages' radices, 4 bits each from the lowest (no array: a plan copied into a kernel body stays in
                  // registers rather than private memory)
  int tw;         // offset of this axis's twiddles in the table
};
struct FftPlan {
  Fft1 ax[3];
};

Fft1 make_fft(int n, int twoff) {
  Fft1 f{};
  f.n = n, f.tw = twoff;
  int m = n;
  for (const int r : {8, 4, 2, 3, 5, 7}) {
    while (m % r == 0) {
      if (r == 8 && m % 8 != 0) break;
      if (f.nr == 12) throw Error(-1, "PME grid sizes need at most 12 transform stages");
      f.rad |= uint64_t(r) << (4 * f.nr++);
      m /= r;
    }
  }
  if (m != 1) throw Error(-1, "PME grid sizes must be products of 2, 3, 5 and 7");
  return f;
}

// R-point DFTs, o[u] = Σ_q v[q] exp(sign 2πi qu/R), in place; radix 4 and 8 with their exact ±i and (±1 ± i)/√2
// factors, the odd radices with cos/sin(2πm/R) as literals
__device__ inline float2 mul_si(float2 a, int sign) { return sign > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x); }   // × sign·i
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline void dft4(float2& v0, float2& v1, float2& v2, float2& v3, int sign) {
  const float2 a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = mul_si(csub(v1, v3), sign);
  v0 = cadd(a, c), v2 = csub(a, c), v1 = cadd(b, d), v3 = csub(b, d);
}
// × exp(sign 2πi m/16) for a compile-time m: cos and sin of π/8 steps as literals, the quarter turns exact
template <int M>
__device__ inline float2 tw16(float2 a, int sign) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, h = 0.70710678118654752f;
  constexpr int m = M % 16;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (m == 4) {
    return mul_si(a, sign);
  } else {
    constexpr float C[16] = {1.0f, c1, h, s1, 0.0f, -s1, -h, -c1, -1.0f, -c1, -h, -s1, 0.0f, s1, h, c1};
    constexpr float S[16] = {0.0f, s1, h, c1, 1.0f, c1, h, s1, 0.0f, -s1, -h, -c1, -1.0f, -c1, -h, -s1};
    const float c = C[m], sn = sign * S[m];
    return make_float2(a.x * c - a.y * sn, a.x * sn + a.y * c);
  }
}

template <int R>
__device__ inline void dft(float2* v, int sign) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b), v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4(v[0], v[1], v[2], v[3], sign);
  } else if constexpr (R == 16) {
    // n = n1 + 4 n2, k = k2 + 4 k1: four 4-point transforms over n2, × exp(sign 2πi n1 k2/16), four over n1
    float2 y[16];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
      float2 a0 = v[n1], a1 = v[n1 + 4], a2 = v[n1 + 8], a3 = v[n1 + 12];
      dft4(a0, a1, a2, a3, sign);
      y[4 * n1] = a0, y[4 * n1 + 1] = a1, y[4 * n1 + 2] = a2, y[4 * n1 + 3] = a3;   // y[4 n1 + k2]
    }
    y[5] = tw16<1>(y[5], sign), y[6] = tw16<2>(y[6], sign), y[7] = tw16<3>(y[7], sign);
    y[9] = tw16<2>(y[9], sign), y[10] = tw16<4>(y[10], sign), y[11] = tw16<6>(y[11], sign);
    y[13] = tw16<3>(y[13], sign), y[14] = tw16<6>(y[14], sign), y[15] = tw16<9>(y[15], sign);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      float2 b0 = y[k2], b1 = y[4 + k2], b2 = y[8 + k2], b3 = y[12 + k2];
      dft4(b0, b1, b2, b3, sign);
      v[k2] = b0, v[k2 + 4] = b1, v[k2 + 8] = b2, v[k2 + 12] = b3;
    }
  } else if constexpr (R == 8) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sign);
    dft4(o0, o1, o2, o3, sign);
    const float h = 0.70710678118654752f;   // √2 / 2
    // o[u] × exp(sign 2πi u/8), u = 1, 2, 3
    const float2 t1 = make_float2(h * (o1.x - sign * o1.y), h * (o1.y + sign * o1.x));
    const float2 t2 = mul_si(o2, sign);
    const float2 t3 = make_float2(-h * (o3.x + sign * o3.y), h * (sign * o3.x - o3.y));
    v[0] = cadd(e0, o0), v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1), v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2), v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3), v[7] = csub(e3, t3);
  } else {
    static_assert(R == 3 || R == 5 || R == 7, "radix");
    // R = 3, 5, 7: cos and sin of 2πm/R
    constexpr float C3[3] = {1.0f, -0.5f, -0.5f}, S3[3] = {0.0f, 0.86602540378443865f, -0.86602540378443865f};
    constexpr float C5[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f, -0.80901699437494742f, 0.30901699437494742f};
    constexpr float S5[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f, -0.58778525229247313f, -0.95105651629515357f};
    constexpr float C7[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f, -0.90096886790241913f,
                             -0.22252093395631440f, 0.62348980185873353f};
    constexpr float S7[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f, -0.43388373911755812f,
                             -0.97492791218182361f, -0.78183148246802981f};
    const float* Cs = R == 3 ? C3 : R == 5 ? C5 : C7;
    const float* Ss = R == 3 ? S3 : R == 5 ? S5 : S7;
    float2 o[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      float2 acc = v[0];
#pragma unroll
      for (int q = 1; q < R; ++q) {
        const int m = (q * u) % R;
        const float c = Cs[m], sn = sign * Ss[m];
        acc.x += v[q].x * c - v[q].y * sn;
        acc.y += v[q].x * sn + v[q].y * c;
      }
      o[u] = acc;
    }
#pragma unroll
    for (int u = 0; u < R; ++u) v[u] = o[u];
  }
}

// LDS layout of the lines being transformed: element j of a line at j + ⌊j/16⌋, lines n + ⌊n/16⌋ + 1 apart. The
// radix-strided writes of the early Stockham stages and the column-wise loads and stores of the b and c passes then
// fall in different banks (with plain n-apart lines of a power of two, sixteen lanes would hit one bank).
__host__ __device__ inline int lds_at(int j) { return j + (j >> 4); }
__host__ __device__ inline int line_stride(int n) { return n + (n >> 4) + 1; }
// The c-line pass's blocks of LY lines along a (of h1): a tail of at most LY / 4 lines joins the last block instead of
// taking a workgroup of its own (65 = 4 x 16 + 1 at K1 = 128) where the larger block's LDS stays within 64 KiB, so a
// block holds at most zline_cap() lines.
__host__ __device__ inline int zline_cap(int LY, int K3) {
  const int m = LY + LY / 4;
  return (2 * m * line_stride(K3) + K3) * int(sizeof(float2)) <= 65536 ? m : LY;
}
__host__ __device__ inline int zline_blocks(int h1, int LY, int K3) {
  const int full = h1 / LY, rem = h1 - full * LY;
  return full > 0 && rem > 0 && LY + rem <= zline_cap(LY, K3) ? full : full + (rem > 0);
}

// one Stockham stage of radix R over nl lines of length n (ls apart): butterfly i of a line takes
// x_q = src[i + q n/R] × w^(q k), w = exp(sign 2πi / (pR)), k = i mod p, and writes its DFT to dst[(i − k) R + k + q p];
// twl holds exp(2πi m/n). The butterfly is one function for the stages with run-time sizes and the compile-time ones
// below, so both compute a line with the same instructions.
template <int R, int Sign>
__device__ inline void butterfly(const float2* s, float2* d, int i, int k, int nb, int p, int step, const float2* twl) {
  float2 v[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const float2 x = s[lds_at(i + q * nb)];
    if (q == 0) {
      v[q] = x;
    } else {
      float2 w = twl[q * k * step];   // q k step < n
      if (Sign < 0) w.y = -w.y;
      v[q] = cmul(x, w);
    }
  }
  dft<R>(v, Sign);
  const int j = (i - k) * R + k;
#pragma unroll
  for (int q = 0; q < R; ++q) d[lds_at(j + q * p)] = v[q];
}

template <int R, int Sign>
__device__ inline void fft_stage(const float2* src, float2* dst, int nl, int n, int ls, int p, const float2* twl) {
  const int nb = n / R, step = n / (p * R);
  for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
    const int line = t / nb, i = t - line * nb;
    butterfly<R, Sign>(src + line * ls, dst + line * ls, i, i % p, nb, p, step, twl);
  }
}

// power-of-two lengths with every size a compile-time constant (the index arithmetic shifts and masks) and radix 16
// while it divides (then 8, 4, 2): 128 points in two stages. Every kernel transforms these lengths here, so a line's
// result is the same wherever it is transformed.
template <int R, int N, int P, int Sign>
__device__ inline void fft_stage_n(const float2* src, float2* dst, int nl, const float2* twl) {
  constexpr int nb = N / R, step = N / (P * R), LS = N + (N >> 4) + 1;
  for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
    const int line = t / nb, i = t % nb;
    butterfly<R, Sign>(src + line * LS, dst + line * LS, i, i % P, nb, P, step, twl);
  }
}
template <int N, int Sign, int P>
__device__ inline float2* stages_n(float2* src, float2* dst, int nl, const float2* twl) {
  if constexpr (P >= N) {
    return src;
  } else {
    constexpr int M = N / P, R = M % 16 == 0 ? 16 : M % 8 == 0 ? 8 : M % 4 == 0 ? 4 : 2;
    fft_stage_n<R, N, P, Sign>(src, dst, nl, twl);
    __syncthreads();
    return stages_n<N, Sign, P * R>(dst, src, nl, twl);
  }
}

// the axis's twiddles exp(2πi m/n): fetched into registers first (two per thread of 256; n ≤ 512) so their loads are
// in flight with the data's, then stored to LDS (the caller synchronises before the transform)
struct TwFetch {
  float2 v[2];
};
static_assert(kMaxFft <= 2 * 256, "two twiddles per thread of a 256-thread transform");
__device__ inline TwFetch fetch_twiddles(const float2* __restrict__ tw, const Fft1& F) {
  TwFetch r;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int t = threadIdx.x + u * 256;
    r.v[u] = t < F.n ? tw[F.tw + t] : make_float2(0.0f, 0.0f);
  }
  return r;
}
__device__ inline void store_twiddles(float2* twl, const TwFetch& r, const Fft1& F) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int t = threadIdx.x + u * 256;
    if (t < F.n) twl[t] = r.v[u];
  }
}

// nl lines of length n in A (laid out as lds_at and line_stride say), transformed by the whole workgroup; returns the
// buffer with the result
// (the sign a compile-time constant in every kernel, so each butterfly compiles to the same instructions wherever it runs
// and every tile shape transforms a line bit-identically)
template <int Sign>
__device__ float2* fft_lines(float2* A, float2* B, int nl, const Fft1& F, const float2* twl) {
  switch (F.n) {
    case 32: return stages_n<32, Sign, 1>(A, B, nl, twl);
    case 64: return stages_n<64, Sign, 1>(A, B, nl, twl);
    case 128: return stages_n<128, Sign, 1>(A, B, nl, twl);
    case 256: return stages_n<256, Sign, 1>(A, B, nl, twl);
    case 512: return stages_n<512, Sign, 1>(A, B, nl, twl);
    default: break;
  }
  float2* src = A;
  float2* dst = B;
  const int n = F.n, ls = line_stride(n);
  int p = 1;
  for (int s = 0; s < F.nr; ++s) {
    const int R = int((F.rad >> (4 * s)) & 15u);
    switch (R) {
      case 2: fft_stage<2, Sign>(src, dst, nl, n, ls, p, twl); break;
      case 3: fft_stage<3, Sign>(src, dst, nl, n, ls, p, twl); break;
      case 4: fft_stage<4, Sign>(src, dst, nl, n, ls, p, twl); break;
      case 5: fft_stage<5, Sign>(src, dst, nl, n, ls, p, twl); break;
      case 7: fft_stage<7, Sign>(src, dst, nl, n, ls, p, twl); break;
      default: fft_stage<8, Sign>(src, dst, nl, n, ls, p, twl); break;
    }
    __syncthreads();
    float2* tmp = src;
    src = dst, dst = tmp;
    p *= R;
  }
  return src;
}

// the transform kernels run 256 threads; their global loads are issued kFftBatch at a time into registers before the
// LDS stores, so a thread waits once per batch rather than once per element
constexpr int kFftThreads = 256, kFftBatch = 8;

// real lines along a (index j + K2·k, pairs 2l, 2l+1) -> the half spectrum, as caps_f_r2c_lines (f90:392-423)
__device__ __forceinline__ void b_fft_x_r2c(int bx, int gx, GridArgs G, FftPlan FP, const float2* __restrict__ tw, const float* __restrict__ grid,
                                                   float2* __restrict__ half) {
  extern __shared__ float2 lds[];
  const int K1 = G.k[0], h1 = G.h1, hs = G.hs, nl = G.k[1] * G.k[2], npairs = (nl + 1) / 2, ls = line_stride(K1);
  const int LP = G.lx;
  float2* A = lds;
  float2* B = lds + LP * ls;
  float2* twl = B + LP * ls;
  const TwFetch twf = fetch_twiddles(tw, FP.ax[0]);
  const int pb = bx * LP;
  for (int t0 = 0; t0 < LP * K1; t0 += kFftBatch * kFftThreads) {
    float2 r[kFftBatch];
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, pl = t / K1, m = t - pl * K1, pr = pb + pl;
      r[u] = make_float2(0.0f, 0.0f);
      if (t < LP * K1 && pr < npairs) {
        const int la = 2 * pr, lb = min(2 * pr + 1, nl - 1);
        r[u].x = grid[size_t(la) * K1 + m];
        r[u].y = lb != la ? grid[size_t(lb) * K1 + m] : 0.0f;
      }
    }
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, pl = t / K1, m = t - pl * K1;
      if (t < LP * K1) A[pl * ls + lds_at(m)] = r[u];
    }
  }
  store_twiddles(twl, twf, FP.ax[0]);
  __syncthreads();
  const float2* Z = fft_lines<+1>(A, B, LP, FP.ax[0], twl);
  for (int t = threadIdx.x; t < LP * h1; t += blockDim.x) {
    const int pl = t / h1, m = t - pl * h1, pr = pb + pl;
    if (pr >= npairs) continue;
    const int la = 2 * pr, lb = min(2 * pr + 1, nl - 1);
    const float2 zm = Z[pl * ls + lds_at(m)], zc0 = Z[pl * ls + lds_at((K1 - m) % K1)];
    const float2 zc = make_float2(zc0.x, -zc0.y);
    half[size_t(la) * hs + m] = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y + zc.y));
    if (lb != la) {
      const float2 d = make_float2(zm.x - zc.x, zm.y - zc.y);
      half[size_t(lb) * hs + m] = make_float2(0.5f * d.y, -0.5f * d.x);
    }
  }
}
__global__ void __launch_bounds__(256) k_fft_x_r2c(GridArgs G, FftPlan FP, const float2* __restrict__ tw, const float* __restrict__ grid,
                                                   float2* __restrict__ half) {
  b_fft_x_r2c(blockIdx.x, gridDim.x, G, FP, tw, grid, half);
}

// the inverse: half spectra of two real lines rebuilt whole, Z = A + iB, one inverse transform (f90:427-461)
__device__ __forceinline__ void b_fft_x_c2r(int bx, int gx, GridArgs G, FftPlan FP, const float2* __restrict__ tw, const float2* __restrict__ half,
                                                   float* __restrict__ grid) {
  extern __shared__ float2 lds[];
  const int K1 = G.k[0], h1 = G.h1, hs = G.hs, nl = G.k[1] * G.k[2], npairs = (nl + 1) / 2, ls = line_stride(K1);
  const int LP = G.lx;
  float2* A = lds;
  float2* B = lds + LP * ls;
  float2* twl = B + LP * ls;
  const TwFetch twf = fetch_twiddles(tw, FP.ax[0]);
  const int pb = bx * LP;
  for (int t0 = 0; t0 < LP * K1; t0 += kFftBatch * kFftThreads) {
    float2 ra[kFftBatch], rb[kFftBatch];
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, pl = t / K1, m = t - pl * K1, pr = pb + pl;
      ra[u] = rb[u] = make_float2(0.0f, 0.0f);
      if (t < LP * K1 && pr < npairs) {
        const int la = 2 * pr, lb = min(2 * pr + 1, nl - 1);
        const int mh = m > K1 / 2 ? K1 - m : m;
        ra[u] = half[size_t(la) * hs + mh];
        if (lb != la) rb[u] = half[size_t(lb) * hs + mh];
      }
    }
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, pl = t / K1, m = t - pl * K1;
      if (t < LP * K1) {
        float2 a = ra[u], b = rb[u];
        if (m > K1 / 2) a.y = -a.y, b.y = -b.y;
        A[pl * ls + lds_at(m)] = make_float2(a.x - b.y, a.y + b.x);
      }
    }
  }
  store_twiddles(twl, twf, FP.ax[0]);
  __syncthreads();
  const float2* Z = fft_lines<-1>(A, B, LP, FP.ax[0], twl);
  for (int t = threadIdx.x; t < LP * K1; t += blockDim.x) {
    const int pl = t / K1, m = t - pl * K1, pr = pb + pl;
    if (pr >= npairs) continue;
    const int la = 2 * pr, lb = min(2 * pr + 1, nl - 1);
    const float2 z = Z[pl * ls + lds_at(m)];
    grid[size_t(la) * K1 + m] = z.x;
    if (lb != la) grid[size_t(lb) * K1 + m] = z.y;
  }
}
__global__ void __launch_bounds__(256) k_fft_x_c2r(GridArgs G, FftPlan FP, const float2* __restrict__ tw, const float2* __restrict__ half,
                                                   float* __restrict__ grid) {
  b_fft_x_c2r(blockIdx.x, gridDim.x, G, FP, tw, half, grid);
}

// lines along b of the half spectrum: element (m, j, k) at m + hs (j + K2 k)
template <int Sign>
__device__ __forceinline__ void b_fft_y(int bx, int gx, GridArgs G, FftPlan FP, const float2* __restrict__ tw, float2* __restrict__ half) {
  extern __shared__ float2 lds[];
  const int h1 = G.h1, K2 = G.k[1];
  const int LY = G.ly, lg = G.ly_log2;   // lines (consecutive a-frequencies) per workgroup
  const int mblocks = (h1 + LY - 1) >> lg;
  const int kz = bx / mblocks, m0 = (bx % mblocks) << lg;
  const int nl = min(LY, h1 - m0), ls = line_stride(K2);
  float2* A = lds;
  float2* B = lds + LY * ls;
  float2* twl = B + LY * ls;
  const TwFetch twf = fetch_twiddles(tw, FP.ax[1]);
  for (int t0 = 0; t0 < LY * K2; t0 += kFftBatch * kFftThreads) {
    float2 r[kFftBatch];
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, j = t >> lg, li = t & (LY - 1);
      r[u] = t < LY * K2 && li < nl ? half[size_t(m0 + li) + size_t(G.hs) * (j + size_t(K2) * kz)] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, j = t >> lg, li = t & (LY - 1);
      if (t < LY * K2 && li < nl) A[li * ls + lds_at(j)] = r[u];
    }
  }
  store_twiddles(twl, twf, FP.ax[1]);
  __syncthreads();
  const float2* Z = fft_lines<Sign>(A, B, nl, FP.ax[1], twl);
  for (int t = threadIdx.x; t < LY * K2; t += blockDim.x) {
    const int j = t >> lg, li = t & (LY - 1);
    if (li < nl) half[size_t(m0 + li) + size_t(G.hs) * (j + size_t(K2) * kz)] = Z[li * ls + lds_at(j)];
  }
}
template <int Sign>
__global__ void __launch_bounds__(256) k_fft_y(GridArgs G, FftPlan FP, const float2* __restrict__ tw, float2* __restrict__ half) {
  b_fft_y<Sign>(blockIdx.x, gridDim.x, G, FP, tw, half);
}

// ---- a c-plane per workgroup: the transforms along a and b with the plane in LDS --------------------------------------
// The plane k's K2 real lines along a, paired (2l, 2l + 1) into one complex transform each as in k_fft_x_r2c, split into
// their half spectra, then the h1 lines along b as in k_fft_y: every line's arithmetic is those kernels', only the half
// spectrum stays in LDS between the passes (one read of the plane and one write instead of two each). The inverse runs
// the same steps backwards (k_fft_y(−1), then k_fft_x_c2r). Used when K2 is even (the pairs stay within a plane) and the
// plane fits: LDS holds two regions of plane_region() elements (the Stockham stages' source and destination), then the
// twiddles of a and b. With fewer planes than CUs, each plane goes to two workgroups (part 0 and 1 of parts = 2): the
// forward transform's both run every line along a and each its half of the lines along b, the inverse's both every line
// along b and each its half of the row pairs along a — three quarters of a plane's work each, on twice the CUs.
constexpr int kPlaneThreads = 1024, kPlaneBatch = 8;
__host__ __device__ inline int plane_region(int K1, int K2) {
  const int h1 = K1 / 2 + 1, a = ((K2 + 1) / 2) * line_stride(K1), b = h1 * line_stride(K2);
  return a > b ? a : b;
}

__device__ inline void plane_twiddles(const float2* __restrict__ tw, const FftPlan& FP, float2* twa, float2* twb) {
  for (int t = threadIdx.x; t < FP.ax[0].n; t += blockDim.x) twa[t] = tw[FP.ax[0].tw + t];
  for (int t = threadIdx.x; t < FP.ax[1].n; t += blockDim.x) twb[t] = tw[FP.ax[1].tw + t];
}

// KA, KB: the plane's sizes as compile-time constants (0: the run-time G.k), so the common grids index with shifts
template <int KA, int KB>
__device__ __forceinline__ void b_fft_xy_fwd(int kz, int part, int parts, GridArgs G, FftPlan FP, const float2* __restrict__ tw,
                                             const float* __restrict__ grid, float2* __restrict__ half) {
  extern __shared__ float2 lds[];
  const int K1 = KA ? KA : G.k[0], K2 = KB ? KB : G.k[1], h1 = K1 / 2 + 1, hs = half_stride(K1), ls1 = line_stride(K1), ls2 = line_stride(K2),
            npair = K2 / 2;
  const int S = plane_region(K1, K2);
  float2* A = lds;
  float2* B = lds + S;
  float2* twa = lds + 2 * S;
  float2* twb = twa + K1;
  // the plane's rows, row r into component r & 1 of pair line r >> 1; loads batched ahead of the LDS stores
  const float* src = grid + size_t(kz) * K2 * K1;
  float* Af = reinterpret_cast<float*>(A);
  const int total = K2 * K1;
  for (int t0 = 0; t0 < total; t0 += kPlaneBatch * kPlaneThreads) {
    float r[kPlaneBatch];
#pragma unroll
    for (int u = 0; u < kPlaneBatch; ++u) {
      const int t = t0 + u * kPlaneThreads + threadIdx.x;
      r[u] = t < total ? src[t] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kPlaneBatch; ++u) {
      const int t = t0 + u * kPlaneThreads + threadIdx.x;
      if (t < total) {
        const int row = t / K1, x = t - row * K1;
        Af[2 * ((row >> 1) * ls1 + lds_at(x)) + (row & 1)] = r[u];
      }
    }
  }
  plane_twiddles(tw, FP, twa, twb);
  __syncthreads();
  float2* Z = fft_lines<+1>(A, B, npair, FP.ax[0], twa);
  float2* Y = Z == A ? B : A;
  // this part's lines along b: a-frequencies m0 … m1 − 1
  const int hp = (h1 + parts - 1) / parts, m0 = part * hp, m1 = min(h1, m0 + hp), nm = m1 - m0;
  // the pairs' half spectra as the b-lines' elements: line m (a-frequency), element j (row)
  for (int t = threadIdx.x; t < npair * nm; t += blockDim.x) {
    const int pl = t / nm, m = m0 + (t - pl * nm);
    const float2 zm = Z[pl * ls1 + lds_at(m)], zc0 = Z[pl * ls1 + lds_at((K1 - m) % K1)];
    const float2 zc = make_float2(zc0.x, -zc0.y);
    const float2 d = make_float2(zm.x - zc.x, zm.y - zc.y);
    Y[m * ls2 + lds_at(2 * pl)] = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y + zc.y));
    Y[m * ls2 + lds_at(2 * pl + 1)] = make_float2(0.5f * d.y, -0.5f * d.x);
  }
  __syncthreads();
  const float2* W = fft_lines<+1>(Y + m0 * ls2, Z + m0 * ls2, nm, FP.ax[1], twb);   // line m at W + (m − m0) ls2
  float2* dst = half + size_t(kz) * K2 * hs;
  for (int t = threadIdx.x; t < K2 * nm; t += blockDim.x) {
    const int j = t / nm, m = t - j * nm;
    dst[size_t(j) * hs + m0 + m] = W[m * ls2 + lds_at(j)];
  }
}

template <int KA, int KB>
__device__ __forceinline__ void b_fft_xy_inv(int kz, int part, int parts, GridArgs G, FftPlan FP, const float2* __restrict__ tw,
                                             const float2* __restrict__ half, float* __restrict__ grid) {
  extern __shared__ float2 lds[];
  const int K1 = KA ? KA : G.k[0], K2 = KB ? KB : G.k[1], h1 = K1 / 2 + 1, hs = half_stride(K1), ls1 = line_stride(K1), ls2 = line_stride(K2),
            npair = K2 / 2;
  const int S = plane_region(K1, K2);
  float2* A = lds;
  float2* B = lds + S;
  float2* twa = lds + 2 * S;
  float2* twb = twa + K1;
  const float2* src = half + size_t(kz) * K2 * hs;
  const int total = K2 * h1;
  for (int t0 = 0; t0 < total; t0 += kPlaneBatch * kPlaneThreads) {
    float2 r[kPlaneBatch];
#pragma unroll
    for (int u = 0; u < kPlaneBatch; ++u) {
      const int t = t0 + u * kPlaneThreads + threadIdx.x, j = t / h1, m = t - j * h1;
      r[u] = t < total ? src[size_t(j) * hs + m] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < kPlaneBatch; ++u) {
      const int t = t0 + u * kPlaneThreads + threadIdx.x;
      if (t < total) {
        const int j = t / h1, m = t - j * h1;
        A[m * ls2 + lds_at(j)] = r[u];
      }
    }
  }
  plane_twiddles(tw, FP, twa, twb);
  __syncthreads();
  float2* Y = fft_lines<-1>(A, B, h1, FP.ax[1], twb);
  float2* X = Y == A ? B : A;
  // this part's row pairs: p0 … p1 − 1 (rows 2 p0 … 2 p1 − 1)
  const int pp = (npair + parts - 1) / parts, p0 = part * pp, p1 = min(npair, p0 + pp), np = p1 - p0;
  // two rows' half spectra rebuilt whole into one complex line, Z = A + iB
  for (int t = threadIdx.x; t < np * K1; t += blockDim.x) {
    const int pl = p0 + t / K1, x = t - (pl - p0) * K1;
    const int m = x > K1 / 2 ? K1 - x : x;
    float2 a = Y[m * ls2 + lds_at(2 * pl)], b = Y[m * ls2 + lds_at(2 * pl + 1)];
    if (x > K1 / 2) a.y = -a.y, b.y = -b.y;
    X[pl * ls1 + lds_at(x)] = make_float2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  const float2* Z = fft_lines<-1>(X + p0 * ls1, Y + p0 * ls1, np, FP.ax[0], twa);   // pair pl at Z + (pl − p0) ls1
  const float* Zf = reinterpret_cast<const float*>(Z);
  float* dst = grid + size_t(kz) * K2 * K1 + size_t(2 * p0) * K1;
  for (int t = threadIdx.x; t < 2 * np * K1; t += blockDim.x) {
    const int row = t / K1, x = t - row * K1;
    dst[t] = Zf[2 * ((row >> 1) * ls1 + lds_at(x)) + (row & 1)];
  }
}

__device__ __forceinline__ void b_fft_xy(int kz, int part, int parts, const GridArgs& G, const FftPlan& FP, const float2* __restrict__ tw,
                                         float* __restrict__ grid, float2* __restrict__ half, int forward) {
  const int K1 = G.k[0], K2 = G.k[1];
#define CAPSGPU_PLANE(A, B)                                                   \
  if (K1 == A && K2 == B) {                                                   \
    if (forward) b_fft_xy_fwd<A, B>(kz, part, parts, G, FP, tw, grid, half);  \
    else b_fft_xy_inv<A, B>(kz, part, parts, G, FP, tw, half, grid);          \
    return;                                                                   \
  }
  CAPSGPU_PLANE(64, 64)
  CAPSGPU_PLANE(128, 128)
  CAPSGPU_PLANE(256, 256)
#undef CAPSGPU_PLANE
  if (forward) b_fft_xy_fwd<0, 0>(kz, part, parts, G, FP, tw, grid, half);
  else b_fft_xy_inv<0, 0>(kz, part, parts, G, FP, tw, half, grid);
}

// parts workgroups per plane (1 or 2); with two and a multiple of eight planes, a plane's two workgroups are b and b + 8,
// which run on one XCD (workgroups go round-robin over the eight) and share its L2
__global__ void __launch_bounds__(kPlaneThreads) k_fft_xy(GridArgs G, FftPlan FP, const float2* __restrict__ tw, float* __restrict__ grid,
                                                         float2* __restrict__ half, int forward, int parts) {
  const int b = blockIdx.x;
  if (parts == 2 && G.k[2] % 8 == 0) b_fft_xy((b >> 4) * 8 + (b & 7), (b >> 3) & 1, 2, G, FP, tw, grid, half, forward);
  else b_fft_xy(b / parts, b % parts, parts, G, FP, tw, grid, half, forward);
}

// the influence function and its virial weights at half-spectrum index (a, b, c), in double (f90:518-564)
__device__ inline void green_at(const GridArgs& G, int a, int b, int c, double fac, double kcoul, double vol, const double* bm,
                                double* g, double* w) {
  const int idx[3] = {a, b, c};
  int m[3], nyq[3], nn = 0;
  for (int d = 0; d < 3; ++d) {
    m[d] = idx[d];
    if (m[d] >= (G.k[d] + 1) / 2) m[d] -= G.k[d];
    nyq[d] = (G.k[d] % 2 == 0) && 2 * idx[d] == G.k[d];
    nn += nyq[d];
  }
  const double bmod = bm[a] * bm[G.k[0] + b] * bm[G.k[0] + G.k[1] + c];
  const int ns = 1 << nn;
  double gs = 0, ws[6] = {0, 0, 0, 0, 0, 0};
  for (int s = 0; s < ns; ++s) {
    int sel = 0;
    double mv[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      double sg = 1;
      if (nyq[d]) {
        if ((s >> sel) & 1) sg = -1;
        ++sel;
      }
      for (int q = 0; q < 3; ++q) mv[q] += sg * m[d] * G.hi[d][q];
    }
    const double m2 = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
    const double e = kcoul * exp(-fac * m2) / (kPi * vol * m2 * bmod);
    const double ss = 2 * (1 + fac * m2) / m2;
    gs += e;
    ws[0] += e * (1 - ss * mv[0] * mv[0]);
    ws[1] += e * (1 - ss * mv[1] * mv[1]);
    ws[2] += e * (1 - ss * mv[2] * mv[2]);
    ws[3] -= e * ss * mv[0] * mv[1];
    ws[4] -= e * ss * mv[0] * mv[2];
    ws[5] -= e * ss * mv[1] * mv[2];
  }
  *g = gs / ns;
  if (w)
    for (int q = 0; q < 6; ++q) w[q] = ws[q] / ns;
}

__global__ void k_green(GridArgs G, double fac, double kcoul, double vol, const double* __restrict__ bm, float* __restrict__ gtab) {
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = size_t(G.h1) * G.k[1] * G.k[2];
  if (t >= total) return;
  const int a = int(t % G.h1), b = int((t / G.h1) % G.k[1]), c = int(t / (size_t(G.h1) * G.k[1]));
  double g = 0;
  if (a || b || c) green_at(G, a, b, c, fac, kcoul, vol, bm, &g, nullptr);
  gtab[size_t(a) + size_t(G.hs) * (b + size_t(G.k[1]) * c)] = float(g);
}

// lines along c: forward transform, × G (energy and virial when kE), inverse transform
template <bool kE>
__device__ __forceinline__ void b_fft_z_conv(int bx, int gx, GridArgs G, FftPlan FP, const float2* __restrict__ tw, float2* __restrict__ half,
                                                    const float* __restrict__ gtab, double fac, double kcoul, double vol,
                                                    const double* __restrict__ bm, double* __restrict__ epart) {
  extern __shared__ float2 lds[];
  const int h1 = G.h1, K2 = G.k[1], K3 = G.k[2];
  const int LY = G.ly, lg = G.ly_log2;
  const int mblocks = zline_blocks(h1, LY, K3);
  const int jy = bx / mblocks, mb = bx % mblocks, m0 = mb << lg;
  // the block's lines: LY, the last block's the rest (a short tail joins it); those past LY go through side loops
  const int nl = mb == mblocks - 1 ? h1 - m0 : LY, nmain = min(nl, LY), nx = nl - nmain;
  const int ls = line_stride(K3);
  float2* A = lds;
  float2* B = lds + zline_cap(LY, K3) * ls;
  float2* twl = B + zline_cap(LY, K3) * ls;
  const TwFetch twf = fetch_twiddles(tw, FP.ax[2]);
  const size_t zs = size_t(G.hs) * K2;
  for (int t0 = 0; t0 < LY * K3; t0 += kFftBatch * kFftThreads) {
    float2 r[kFftBatch];
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, k = t >> lg, li = t & (LY - 1);
      r[u] = t < LY * K3 && li < nmain ? half[size_t(m0 + li) + size_t(G.hs) * jy + zs * k] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, k = t >> lg, li = t & (LY - 1);
      if (t < LY * K3 && li < nmain) A[li * ls + lds_at(k)] = r[u];
    }
  }
  for (int t = threadIdx.x; t < nx * K3; t += blockDim.x) {
    const int k = t / nx, li = LY + (t - k * nx);
    A[li * ls + lds_at(k)] = half[size_t(m0 + li) + size_t(G.hs) * jy + zs * k];
  }
  // the influence function, element by element as the lines were read (consecutive lanes, consecutive a-frequencies:
  // whole cache lines), the first batch requested now so its latency passes under the transform
  float gpre[kFftBatch];
#pragma unroll
  for (int u = 0; u < kFftBatch; ++u) {
    const int t = u * kFftThreads + threadIdx.x, k = t >> lg, li = t & (LY - 1);
    gpre[u] = t < LY * K3 && li < nmain ? gtab[size_t(m0 + li) + size_t(G.hs) * jy + zs * k] : 0.0f;
  }
  store_twiddles(twl, twf, FP.ax[2]);
  __syncthreads();
  float2* Z = fft_lines<+1>(A, B, nl, FP.ax[2], twl);
  double e = 0, w[6] = {0, 0, 0, 0, 0, 0};
  // element (line li, point k) × G, with its energy and virial terms when kE
  auto convolve = [&](int li, int k, float g) {
    const int a = m0 + li;
    float2 v = Z[li * ls + lds_at(k)];
    if (kE && (a || jy || k)) {
      double em = 0.5 * (double(v.x) * v.x + double(v.y) * v.y);
      if (a > 0 && 2 * a != G.k[0]) em *= 2;
      double gd, wv[6];
      green_at(G, a, jy, k, fac, kcoul, vol, bm, &gd, wv);
      e += em * double(g);
      for (int q = 0; q < 6; ++q) w[q] += em * wv[q];
    }
    v.x *= g, v.y *= g;
    Z[li * ls + lds_at(k)] = v;
  };
  for (int t0 = 0; t0 < LY * K3; t0 += kFftBatch * kFftThreads) {
    float gb[kFftBatch];
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, k = t >> lg, li = t & (LY - 1);
      gb[u] = t0 == 0 ? gpre[u] : t < LY * K3 && li < nmain ? gtab[size_t(m0 + li) + size_t(G.hs) * jy + zs * k] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kFftBatch; ++u) {
      const int t = t0 + u * kFftThreads + threadIdx.x, k = t >> lg, li = t & (LY - 1);
      if (t >= LY * K3 || li >= nmain) continue;
      convolve(li, k, gb[u]);
    }
  }
  for (int t = threadIdx.x; t < nx * K3; t += blockDim.x) {
    const int k = t / nx, li = LY + (t - k * nx);
    convolve(li, k, gtab[size_t(m0 + li) + size_t(G.hs) * jy + zs * k]);
  }
  __syncthreads();
  float2* other = Z == A ? B : A;
  const float2* R = fft_lines<-1>(Z, other, nl, FP.ax[2], twl);
  for (int t = threadIdx.x; t < LY * K3; t += blockDim.x) {
    const int k = t >> lg, li = t & (LY - 1);
    if (li < nmain) half[size_t(m0 + li) + size_t(G.hs) * jy + zs * k] = R[li * ls + lds_at(k)];
  }
  for (int t = threadIdx.x; t < nx * K3; t += blockDim.x) {
    const int k = t / nx, li = LY + (t - k * nx);
    half[size_t(m0 + li) + size_t(G.hs) * jy + zs * k] = R[li * ls + lds_at(k)];
  }
  if (kE) {
    // the workgroup's partial sums in a fixed order
    __shared__ double red[7][256];
    red[0][threadIdx.x] = e;
    for (int q = 0; q < 6; ++q) red[1 + q][threadIdx.x] = w[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        for (int q = 0; q < 7; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x < 7) epart[7 * size_t(bx) + threadIdx.x] = red[threadIdx.x][0];
  }
}
template <bool kE>
__global__ void __launch_bounds__(256) k_fft_z_conv(GridArgs G, FftPlan FP, const float2* __restrict__ tw, float2* __restrict__ half,
                                                    const float* __restrict__ gtab, double fac, double kcoul, double vol,
                                                    const double* __restrict__ bm, double* __restrict__ epart) {
  b_fft_z_conv<kE>(blockIdx.x, gridDim.x, G, FP, tw, half, gtab, fac, kcoul, vol, bm, epart);
}

// one 64-lane workgroup per 4x4x4 tile: the tile's box of (p + 3)³ mesh points in LDS (loaded row by row, wrapped),
// then its atoms ⌊64/p⌋ at a time, p lanes per atom: lane j of an atom takes its p rows (dy = j, dz = 0 … p − 1), so
// every lane has whole rows and keeps its b-weights; the p lanes' sums are added in order by the first and it writes
// the atom's force. Dynamic LDS: the box (p + 3)² rows of p + 4 floats, then per atom slot 6p weights.
template <int P>
__device__ __forceinline__ void b_gather(int tile, GridArgs G, const int64_t* __restrict__ tstart, const uint32_t* __restrict__ tatoms,
                                               const float4* __restrict__ wsort, const int4* __restrict__ bsort,
                                               const float* __restrict__ grid, float* __restrict__ f) {
  extern __shared__ float2 lds[];
  constexpr int APG = 64 / P;                            // atoms per group
  const int lane = threadIdx.x;
  const int64_t a0 = tstart[tile], na = tstart[tile + 1] - a0;
  if (na == 0) return;
  constexpr int W = P + 3, WX = W + 1;                   // rows padded to W + 1 floats (fewer bank conflicts)
  float* box = reinterpret_cast<float*>(lds);            // [W][W][WX] (c, b, a)
  float* s_w = box + W * W * WX;                         // [APG][6 P]: θa θb θc dθa dθb dθc
  const int tx = tile % G.nt[0], ty = (tile / G.nt[0]) % G.nt[1], tz = tile / (G.nt[0] * G.nt[1]);
  const int slot = lane / P, sub = lane - slot * P;
  const bool in_group = slot < APG;
  // the first group's atoms are requested before the box, so the two loads' latencies overlap
  int64_t at = a0 + (in_group && slot < na ? slot : 0);
  uint32_t i_next = tatoms[at];
  float4 w_next = wsort[at];
  int4 b_next = bsort[at];
  // the box, element e = x + W (y + W z) by lane: a load instruction reads 64 / W whole rows (few cache lines), all
  // loads in flight before the LDS stores
  constexpr int kBox = W * W * W, kIter = (kBox + 63) / 64;
  float bv[kIter];
#pragma unroll
  for (int it = 0; it < kIter; ++it) {
    const int e = lane + 64 * it;
    if (e < kBox) {
      const int xx = e % W, r = e / W;
      int yy = 4 * ty + r % W, zz = 4 * tz + r / W, gx = 4 * tx + xx;
      yy -= yy >= G.k[1] ? G.k[1] : 0;
      yy -= yy >= G.k[1] ? G.k[1] : 0;
      zz -= zz >= G.k[2] ? G.k[2] : 0;
      zz -= zz >= G.k[2] ? G.k[2] : 0;
      gx -= gx >= G.k[0] ? G.k[0] : 0;
      gx -= gx >= G.k[0] ? G.k[0] : 0;
      // the mesh index fits 32 bits and each product's factors 24 bits (K <= 512 per axis)
      bv[it] = grid[__umul24(__umul24(unsigned(zz), unsigned(G.k[1])) + unsigned(yy), unsigned(G.k[0])) + unsigned(gx)];
    }
  }
#pragma unroll
  for (int it = 0; it < kIter; ++it) {
    const int e = lane + 64 * it;
    if (e < kBox) box[(e / W) * WX + e % W] = bv[it];
  }
  for (int64_t g0 = 0; g0 < na; g0 += APG) {
    const bool has = in_group && g0 + slot < na;
    const uint32_t i = i_next;
    const float4 w = w_next;
    const int4 b = b_next;
    if (g0 + APG < na) {
      const int64_t ia = g0 + APG + slot;
      at = a0 + (in_group && ia < na ? ia : 0);
      i_next = tatoms[at], w_next = wsort[at], b_next = bsort[at];
    }
    float* ws = s_w + (in_group ? slot : 0) * 6 * P;
    if (in_group && sub < 3) {
      // lane sub of the atom's p: the splines of axis sub
      float th[P], dth[P];
      bspline<P>(sub == 0 ? w.x : sub == 1 ? w.y : w.z, th, dth);
#pragma unroll
      for (int j = 0; j < P; ++j) ws[sub * P + j] = th[j], ws[(3 + sub) * P + j] = dth[j];
    }
    __syncthreads();
    float g0s = 0.0f, g1s = 0.0f, g2s = 0.0f;
    if (has) {
      // the atom's first point relative to the tile's corner (its home tile: 0 … 3 per axis); this lane's rows dy = sub
      const int ox = b.x - 4 * tx, oy = b.y - 4 * ty, oz = b.z - 4 * tz;
      float ta[P], da[P];
#pragma unroll
      for (int j = 0; j < P; ++j) ta[j] = ws[j], da[j] = ws[3 * P + j];
      const float tyv = ws[P + sub], dyv = ws[4 * P + sub];
      const float* row = box + (oz * W + (oy + sub)) * WX + ox;
#pragma unroll
      for (int dz = 0; dz < P; ++dz) {
        float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int a = 0; a < P; ++a) {
          const float v = row[dz * W * WX + a];
          s0 += da[a] * v;
          s1 += ta[a] * v;
        }
        const float tzv = ws[2 * P + dz], dzv = ws[5 * P + dz];
        g0s += s0 * tyv * tzv;
        g1s += s1 * dyv * tzv;
        g2s += s1 * tyv * dzv;
      }
    }
    // the atom's p lanes added in order by its first lane (every lane takes part in the shuffles)
    const int base = (in_group ? slot : 0) * P;
    float t0 = g0s, t1 = g1s, t2 = g2s;
#pragma unroll
    for (int k = 1; k < P; ++k) {
      const float u0 = __shfl(g0s, base + k, 64), u1 = __shfl(g1s, base + k, 64), u2 = __shfl(g2s, base + k, 64);
      t0 += u0, t1 += u1, t2 += u2;
    }
    if (has && sub == 0)
      for (int m = 0; m < 3; ++m) {
        const double grad = G.k[0] * double(t0) * G.hi[0][m] + G.k[1] * double(t1) * G.hi[1][m] + G.k[2] * double(t2) * G.hi[2][m];
        f[3 * size_t(i) + m] = float(-double(w.w) * grad);
      }
    __syncthreads();
  }
}
// The gather by runs of the tile order: one 64-lane workgroup per 64 consecutive atoms of the tile-sorted arrays, one
// atom per lane (every lane busy, where the gather by tiles keeps p lanes per atom and a tile's ~8 atoms fill 40 of 64).
// A pass takes the lanes whose home tiles share one (b, c) row of tiles and span at most kSeqTiles tiles along a; it
// loads the box of mesh points those atoms reach — 4 (tiles + 2) points along a by p + 3 along b and c, as float4s
// (rows of the mesh are 16-byte aligned: K1 a multiple of 4) — into LDS, then each lane computes its atom's force with
// the per-tile gather's arithmetic in its order: the same splines, for each row dy the sums over dz and a as that
// kernel's lane dy computes them, the rows' sums added in order dy = 0 … p − 1 as its first lane adds them. The forces
// are bit-identical to the per-tile gather's. Dynamic LDS: the largest box, 4 (kSeqTiles + 2) (p + 3)² floats.
constexpr int kSeqTiles = 12;
template <int P>
__device__ __forceinline__ void b_gather_seq(int wg, int n, GridArgs G, const uint32_t* __restrict__ tatoms,
                                             const float4* __restrict__ wsort, const int4* __restrict__ bsort,
                                             const float* __restrict__ grid, float* __restrict__ f) {
  extern __shared__ float2 lds[];
  constexpr int NY = P + 3;                               // box points along b and c
  constexpr int NX4 = kSeqTiles + 2, RX = 4 * NX4;        // a box row: float4s, floats
  constexpr int PS = NY * RX + 4;                         // a box plane along c, floats
  static_assert(NX4 <= 16, "a box row's float4s are loaded by the 16 lanes of a row of four");
  float* box = reinterpret_cast<float*>(lds);             // [c][b][a]: planes of PS floats, rows of RX
  float4* box4 = reinterpret_cast<float4*>(lds);
  const int lane = threadIdx.x, c4 = lane & 15, r4 = lane >> 4;
  const int64_t s = int64_t(wg) * 64 + lane;
  const bool valid = s < n;
  const int64_t sl = valid ? s : int64_t(n) - 1;
  const uint32_t i = tatoms[sl];
  const float4 w = wsort[sl];
  const int4 b = bsort[sl];
  const int tx = b.x >> 2, ty = b.y >> 2, tz = b.z >> 2;
  const int key = tz * G.nt[1] + ty;                     // the atom's (b, c) row of tiles
  bool done = !valid;
  float th[3][P], dth[3][P];
  bspline<P>(w.x, th[0], dth[0]);
  bspline<P>(w.y, th[1], dth[1]);
  bspline<P>(w.z, th[2], dth[2]);
  for (;;) {
    const uint64_t left = __ballot(!done);
    if (!left) break;
    // this pass: the lanes of the first remaining lane's row, up to kSeqTiles tiles along a (the sorted order keeps a
    // row's atoms in consecutive lanes, their tiles along a non-decreasing)
    const int first = __builtin_ffsll((long long)left) - 1;
    const int key0 = __builtin_amdgcn_readlane(key, first), tx0 = __builtin_amdgcn_readlane(tx, first);
    const int ty0 = __builtin_amdgcn_readlane(ty, first), tz0 = __builtin_amdgcn_readlane(tz, first);
    const bool active = !done && key == key0 && tx < tx0 + kSeqTiles;
    const uint64_t am = __ballot(active);
    const int last = 63 - __builtin_clzll(am);
    const int ntx = __builtin_amdgcn_readlane(tx, last) - tx0 + 1;
    // the box: rows of NX4 float4s (the widest box's), ntx + 2 of them loaded (4 ntx + 8 >= 4 ntx + p - 1 points); lane
    // = (float4 c4 of a row, row of four), so each row's column and mesh index along a are the lane's own. Planes along c
    // are PS floats apart (one float4 more than their rows), so atoms at different heights in a tile read other banks.
    if (c4 < ntx + 2) {
#pragma unroll 4
      for (int r = r4; r < NY * NY; r += 4) {
        const int yy = r % NY, zz = r / NY;
        int gy = 4 * ty0 + yy, gz = 4 * tz0 + zz;
        gy -= gy >= G.k[1] ? G.k[1] : 0;
        gy -= gy >= G.k[1] ? G.k[1] : 0;
        gz -= gz >= G.k[2] ? G.k[2] : 0;
        gz -= gz >= G.k[2] ? G.k[2] : 0;
        int gx = 4 * (tx0 + c4);
        gx -= gx >= G.k[0] ? G.k[0] : 0;
        gx -= gx >= G.k[0] ? G.k[0] : 0;
        box4[zz * (PS / 4) + yy * NX4 + c4] = *reinterpret_cast<const float4*>(grid + (size_t(gz) * G.k[1] + gy) * G.k[0] + gx);
      }
    }
    __syncthreads();
    if (active) {
      // the atom's first point relative to the box's corner: 0 … 4 ntx − 1 along a, 0 … 3 along b and c
      const int ox = b.x - 4 * tx0, oy = b.y - 4 * ty0, oz = b.z - 4 * tz0;
      float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
#pragma unroll
      for (int sub = 0; sub < P; ++sub) {
        const float tyv = th[1][sub], dyv = dth[1][sub];
        const float* row = box + oz * PS + (oy + sub) * RX + ox;
        float g0s = 0.0f, g1s = 0.0f, g2s = 0.0f;
#pragma unroll
        for (int dz = 0; dz < P; ++dz) {
          float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
          for (int a = 0; a < P; ++a) {
            const float v = row[dz * PS + a];
            s0 += dth[0][a] * v;
            s1 += th[0][a] * v;
          }
          const float tzv = th[2][dz], dzv = dth[2][dz];
          g0s += s0 * tyv * tzv;
          g1s += s1 * dyv * tzv;
          g2s += s1 * tyv * dzv;
        }
        if (sub == 0) t0 = g0s, t1 = g1s, t2 = g2s;
        else t0 += g0s, t1 += g1s, t2 += g2s;
      }
      for (int m = 0; m < 3; ++m) {
        const double grad = G.k[0] * double(t0) * G.hi[0][m] + G.k[1] * double(t1) * G.hi[1][m] + G.k[2] * double(t2) * G.hi[2][m];
        f[3 * size_t(i) + m] = float(-double(w.w) * grad);
      }
    }
    done = done || active;
    __syncthreads();
  }
}
template <int P>
__global__ void __launch_bounds__(64) k_gather_seq(GridArgs G, int n, const uint32_t* __restrict__ tatoms, const float4* __restrict__ wsort,
                                                   const int4* __restrict__ bsort, const float* __restrict__ grid, float* __restrict__ f) {
  const int nwg = (n + 63) / 64;
  b_gather_seq<P>(int(xcd_block(blockIdx.x, nwg)), n, G, tatoms, wsort, bsort, grid, f);
}
// its dynamic LDS: the largest box
size_t gather_seq_lds(int p) { return size_t(p + 3) * (size_t(4) * (kSeqTiles + 2) * (p + 3) + 4) * sizeof(float); }

template <int P>
__global__ void __launch_bounds__(64) k_gather(GridArgs G, const int64_t* __restrict__ tstart, const uint32_t* __restrict__ tatoms,
                                               const float4* __restrict__ wsort, const int4* __restrict__ bsort,
                                               const float* __restrict__ grid, float* __restrict__ f) {
  const int ntiles = G.nt[0] * G.nt[1] * G.nt[2];
  const int tile = int(xcd_block(blockIdx.x, ntiles));   // neighbouring tiles on one XCD (shared atoms and mesh lines)
  if (tile >= ntiles) return;
  b_gather<P>(tile, G, tstart, tatoms, wsort, bsort, grid, f);
}

// the spread by blocks: every K a multiple of 4 (whole tiles; with an odd tile count the last block along that axis is
// one tile deep) and more tiles per axis than reach a block's tile (so its home tiles are distinct)
bool spread_blocks(const GridArgs& G) {
  const int ntp = (G.p + 6) / 4;
  for (int d = 0; d < 3; ++d)
    if (G.k[d] % 4 != 0 || G.nt[d] < ntp + 1) return false;
  return true;
}

GridArgs grid_args(const Pme& P) {
  GridArgs G{};
  for (int d = 0; d < 3; ++d) G.k[d] = P.k[d], G.nt[d] = (P.k[d] + 3) / 4;
  G.ntm = std::max(G.nt[0], std::max(G.nt[1], G.nt[2]));
  G.p = P.order;
  G.h1 = P.h1;
  G.hs = P.hs;
  G.lx = P.lines_x > 0 ? P.lines_x : P.plane_ok ? 0 : 8;   // 0: the plane transform
  G.ly = P.lines_yz;
  G.ly_log2 = 0;
  while ((1 << G.ly_log2) < G.ly) ++G.ly_log2;
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) G.hi[r][q] = P.hi[r][q];
  return G;
}

// B-spline weights at the integers and |b(m)|² with the fix at zeros (f90:601-620)
void moduli(int k, int p, double* out) {
  double th[16] = {0};
  th[1] = 0, th[0] = 1;
  // bspline(0, p) in double
  for (int j = 0; j < p; ++j) th[j] = 0;
  th[1] = 0.0;
  th[0] = 1.0;
  for (int kk = 3; kk <= p - 1; ++kk) {
    const double div = 1.0 / (kk - 1);
    th[kk - 1] = div * 0.0 * th[kk - 2];
    for (int j = 1; j <= kk - 2; ++j) th[kk - j - 1] = div * ((0.0 + j) * th[kk - j - 2] + (kk - j - 0.0) * th[kk - j - 1]);
    th[0] = div * (1 - 0.0) * th[0];
  }
  {
    const double div = 1.0 / (p - 1);
    th[p - 1] = div * 0.0 * th[p - 2];
    for (int j = 1; j <= p - 2; ++j) th[p - j - 1] = div * ((0.0 + j) * th[p - j - 2] + (p - j - 0.0) * th[p - j - 1]);
    th[0] = div * (1 - 0.0) * th[0];
  }
  for (int m = 0; m < k; ++m) {
    double sr = 0, si = 0;
    for (int j = 0; j <= p - 2; ++j) {
      const double arg = 2 * kPi * m * j / double(k);
      sr += th[j] * std::cos(arg);
      si += th[j] * std::sin(arg);
    }
    out[m] = sr * sr + si * si;
  }
  for (int m = 0; m < k; ++m)
    if (out[m] < 1e-7) out[m] = 0.5 * (out[(m - 1 + k) % k] + out[(m + 1) % k]);
}

FftPlan plan_of(const Pme& P) {
  FftPlan F{};
  F.ax[0] = make_fft(P.k[0], 0);
  F.ax[1] = make_fft(P.k[1], P.k[0]);
  F.ax[2] = make_fft(P.k[2], P.k[0] + P.k[1]);
  return F;
}

}  // namespace

// the smallest 2-3-5-7 size of at least n points (kspace.cpp:45-52, f90:23-36)
static int good_fft_size(int n) {
  int m = std::max(n, 1);
  for (;;) {
    int k = m;
    for (int f : {2, 3, 5, 7})
      while (k % f == 0) k /= f;
    if (k == 1) return m;
    ++m;
  }
}

// the plane transform's dynamic LDS, and whether a grid's planes take it on this device: K2 even and the LDS within the
// device's limit per workgroup (the kernel's limit raised to it once)
size_t plane_lds(int K1, int K2) { return (2 * size_t(plane_region(K1, K2)) + size_t(K1) + size_t(K2)) * sizeof(float2); }

bool plane_fits(int device, int K1, int K2) {
  if (K2 % 2) return false;
  static std::mutex mu;
  static int limit[64];
  static bool known[64];
  if (device < 0 || device >= 64) return false;
  std::lock_guard<std::mutex> lk(mu);
  if (!known[device]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) v = 0;
    if (v > 0 && hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_xy), hipFuncAttributeMaxDynamicSharedMemorySize, v) != hipSuccess)
      v = 65536;
    (void)hipGetLastError();
    limit[device] = v, known[device] = true;
  }
  return plane_lds(K1, K2) <= size_t(limit[device]);
}

// the current device's compute units (cached per device)
int device_cus() {
  static std::mutex mu;
  static int cus[64];
  int device = 0;
  HIPCHECK(hipGetDevice(&device));
  if (device < 0 || device >= 64) return 0;
  std::lock_guard<std::mutex> lk(mu);
  if (!cus[device] && hipDeviceGetAttribute(&cus[device], hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus[device] = 0;
  return cus[device];
}

void pme_grid_dims(const capsgpu_ctx& c, const double cell[12], int k[3]) {
  for (int d = 0; d < 3; ++d) {
    const double* e = cell + 3 + 3 * d;
    const double len = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    k[d] = good_fft_size(std::max(c.opt.pme_order + 1, int(std::ceil(len / c.opt.pme_spacing))));
  }
}

// the grid, its tables and the influence function for the current cell (kept while the cell stays the same)
void pme_setup(capsgpu_ctx& c) {
  Pme& P = c.pme;
  const double* cell = c.cell;
  bool same = P.green_ok;
  for (int q = 0; q < 12 && same; ++q) same = P.cell[q] == cell[q];
  if (same) return;
  // h = cell rows as columns; h⁻¹ rows the reciprocal vectors (f90:484-500)
  double h[3][3];
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) h[r][q] = cell[3 + 3 * q + r];
  const double v = h[0][0] * (h[1][1] * h[2][2] - h[2][1] * h[1][2]) - h[0][1] * (h[1][0] * h[2][2] - h[2][0] * h[1][2]) +
                   h[0][2] * (h[1][0] * h[2][1] - h[2][0] * h[1][1]);
  P.hi[0][0] = (h[1][1] * h[2][2] - h[1][2] * h[2][1]) / v;
  P.hi[0][1] = (h[0][2] * h[2][1] - h[0][1] * h[2][2]) / v;
  P.hi[0][2] = (h[0][1] * h[1][2] - h[0][2] * h[1][1]) / v;
  P.hi[1][0] = (h[1][2] * h[2][0] - h[1][0] * h[2][2]) / v;
  P.hi[1][1] = (h[0][0] * h[2][2] - h[0][2] * h[2][0]) / v;
  P.hi[1][2] = (h[0][2] * h[1][0] - h[0][0] * h[1][2]) / v;
  P.hi[2][0] = (h[1][0] * h[2][1] - h[1][1] * h[2][0]) / v;
  P.hi[2][1] = (h[0][1] * h[2][0] - h[0][0] * h[2][1]) / v;
  P.hi[2][2] = (h[0][0] * h[1][1] - h[0][1] * h[1][0]) / v;
  P.vol = std::fabs(v);
  // grid: the smallest 2-3-5-7 sizes with spacing at most pme_spacing
  pme_grid_dims(c, cell, P.k);
  for (int d = 0; d < 3; ++d)
    if (P.k[d] > kMaxFft) throw Error(-1, "PME grid larger than 512 points along an axis");
  P.order = c.opt.pme_order;
  if (P.order < 3 || P.order > kMaxOrder) throw Error(-1, "PME order must be 3 … 8");
  P.h1 = P.k[0] / 2 + 1;
  P.hs = half_stride(P.k[0]);
  hipStream_t s = c.stream;
  std::vector<double> bm(size_t(P.k[0] + P.k[1] + P.k[2]));
  moduli(P.k[0], P.order, bm.data());
  moduli(P.k[1], P.order, bm.data() + P.k[0]);
  moduli(P.k[2], P.order, bm.data() + P.k[0] + P.k[1]);
  P.bmod.upload(bm, s);
  std::vector<float2> tw;
  for (int d = 0; d < 3; ++d)
    for (int m = 0; m < P.k[d]; ++m) {
      const double ang = 2 * kPi * m / P.k[d];
      tw.push_back(make_float2(float(std::cos(ang)), float(std::sin(ang))));
    }
  P.tw.upload(tw, s);
  const size_t nh = size_t(P.hs) * P.k[1] * P.k[2];
  P.gtab.resize(nh);
  const GridArgs G = grid_args(P);
  const double fac = kPi * kPi / (P.beta * P.beta);
  kt_begin(c, "k_green");
  hipLaunchKernelGGL(k_green, dim3(blocks_for(size_t(P.h1) * P.k[1] * P.k[2], 256)), dim3(256), 0, s, G, fac, kCoulomb, P.vol, P.bmod.p,
                     P.gtab.p);
  LAUNCHCHECK();
  kt_end(c);
  // per axis and tile T: the distinct home tiles of the first points that reach points 4T … 4T + 3 (count, then up to
  // four tiles in the order of those first points)
  std::vector<int32_t> home(size_t(3) * G.ntm * 5, 0);
  for (int d = 0; d < 3; ++d)
    for (int T = 0; T < G.nt[d]; ++T) {
      int32_t* h = home.data() + (size_t(d) * G.ntm + T) * 5;
      for (int bb = 4 * T - P.order + 1; bb <= 4 * T + 3; ++bb) {
        const int t = (((bb % P.k[d]) + P.k[d]) % P.k[d]) >> 2;
        bool seen = false;
        for (int q = 0; q < h[0]; ++q) seen |= h[1 + q] == t;
        if (!seen) {
          if (h[0] == 4) throw Error(-1, "PME: more than four home tiles along an axis");
          h[1 + h[0]++] = t;
        }
      }
    }
  P.home.upload(home, s);
  P.grid.resize(size_t(P.k[0]) * P.k[1] * P.k[2]);
  P.half.resize(nh);
  P.plane_ok = plane_fits(c.device, P.k[0], P.k[1]);
  for (int q = 0; q < 12; ++q) P.cell[q] = cell[q];
  P.green_ok = true;
  P.counts_clear = false;   // a new grid: its tile counters are cleared before the next spread
  P.presplined = false;
}

namespace {
// the transforms' launches with the context's tile shape (G.lx line pairs, G.ly lines per workgroup)
void launch_fft_x(capsgpu_ctx& c, const GridArgs& G, const FftPlan& F, bool forward) {
  Pme& P = c.pme;
  const int npairs = (P.k[1] * P.k[2] + 1) / 2;
  const size_t lds_x = (2 * size_t(G.lx) * size_t(line_stride(P.k[0])) + P.k[0]) * sizeof(float2);
  const dim3 grid(unsigned((npairs + G.lx - 1) / G.lx));
  if (forward) {
    kt_begin(c, "k_fft_x_r2c");
    hipLaunchKernelGGL(k_fft_x_r2c, grid, dim3(256), lds_x, c.stream, G, F, P.tw.p, P.grid.p, P.half.p);
  } else {
    kt_begin(c, "k_fft_x_c2r");
    hipLaunchKernelGGL(k_fft_x_c2r, grid, dim3(256), lds_x, c.stream, G, F, P.tw.p, P.half.p, P.grid.p);
  }
  LAUNCHCHECK();
  kt_end(c);
}

void launch_fft_y(capsgpu_ctx& c, const GridArgs& G, const FftPlan& F, int sign) {
  Pme& P = c.pme;
  const int mblocks = (P.h1 + G.ly - 1) / G.ly;
  const size_t lds_y = (2 * size_t(G.ly) * size_t(line_stride(P.k[1])) + P.k[1]) * sizeof(float2);
  kt_begin(c, "k_fft_y");
  if (sign > 0) hipLaunchKernelGGL(k_fft_y<+1>, dim3(unsigned(mblocks * P.k[2])), dim3(256), lds_y, c.stream, G, F, P.tw.p, P.half.p);
  else hipLaunchKernelGGL(k_fft_y<-1>, dim3(unsigned(mblocks * P.k[2])), dim3(256), lds_y, c.stream, G, F, P.tw.p, P.half.p);
  LAUNCHCHECK();
  kt_end(c);
}

void launch_fft_z_conv(capsgpu_ctx& c, const GridArgs& G, const FftPlan& F, bool want) {
  Pme& P = c.pme;
  const int mblocks = zline_blocks(P.h1, G.ly, P.k[2]);
  const size_t lds_z = (2 * size_t(zline_cap(G.ly, P.k[2])) * size_t(line_stride(P.k[2])) + P.k[2]) * sizeof(float2);
  const unsigned zb = unsigned(mblocks * P.k[1]);
  const double fac = kPi * kPi / (P.beta * P.beta);
  kt_begin(c, "k_fft_z_conv");
  if (want) {
    P.epart.resize(7 * size_t(zb));
    P.erows = int64_t(zb);
    hipLaunchKernelGGL(k_fft_z_conv<true>, dim3(zb), dim3(256), lds_z, c.stream, G, F, P.tw.p, P.half.p, P.gtab.p, fac, kCoulomb, P.vol,
                       P.bmod.p, P.epart.p);
  } else {
    hipLaunchKernelGGL(k_fft_z_conv<false>, dim3(zb), dim3(256), lds_z, c.stream, G, F, P.tw.p, P.half.p, P.gtab.p, fac, kCoulomb,
                       P.vol, P.bmod.p, nullptr);
  }
  LAUNCHCHECK();
  kt_end(c);
}

void launch_fft_xy(capsgpu_ctx& c, const GridArgs& G, const FftPlan& F, bool forward) {
  Pme& P = c.pme;
  kt_begin(c, forward ? "k_fft_xy_fwd" : "k_fft_xy_inv");
  // two workgroups per plane when the planes alone would leave half the CUs idle
  const int parts = 2 * P.k[2] <= device_cus() ? 2 : 1;
  hipLaunchKernelGGL(k_fft_xy, dim3(unsigned(parts * P.k[2])), dim3(kPlaneThreads), plane_lds(P.k[0], P.k[1]), c.stream, G, F, P.tw.p,
                     P.grid.p, P.half.p, forward ? 1 : 0, parts);
  LAUNCHCHECK();
  kt_end(c);
}

// the passes along a and b: one plane kernel (G.lx = 0) or the a-lines and b-lines kernels
void fft_ab(capsgpu_ctx& c, const GridArgs& G, const FftPlan& F, bool forward) {
  if (G.lx == 0) {
    launch_fft_xy(c, G, F, forward);
  } else if (forward) {
    launch_fft_x(c, G, F, true);
    launch_fft_y(c, G, F, +1);
  } else {
    launch_fft_y(c, G, F, -1);
    launch_fft_x(c, G, F, false);
  }
}

// mesh → half spectrum → × influence function → mesh, in place on P.grid
void fft_forward_conv_inverse(capsgpu_ctx& c, const GridArgs& G, const FftPlan& F, bool want) {
  fft_ab(c, G, F, true);
  launch_fft_z_conv(c, G, F, want);
  fft_ab(c, G, F, false);
}

// the gather by runs of 64 sorted atoms where the context's shape asks for it (0 tiles, 1 runs) and the mesh rows are
// 16-byte aligned (K1 a multiple of 4)
bool gather_by_runs(const Pme& P, const GridArgs& G) { return P.gather_shape == 1 && G.k[0] % 4 == 0; }

// the gather of the last spread's atoms from P.grid into P.fpme (the arrays of the last evaluation)
void launch_gather(capsgpu_ctx& c, const GridArgs& G, bool seq) {
  Pme& P = c.pme;
  const float4* wsort = reinterpret_cast<const float4*>(P.wsort.p);
  const int4* bsort = reinterpret_cast<const int4*>(P.bsort.p);
  const int ntiles = G.nt[0] * G.nt[1] * G.nt[2];
  const size_t lds_gather = (size_t(G.p + 3) * (G.p + 3) * (G.p + 4) + 64 * 6) * sizeof(float);   // box + ⌊64/p⌋ slots of 6p
  kt_begin(c, "k_gather");
  switch (G.p) {
#define CAPSGPU_GATHER(Q)                                                                                                      \
  case Q:                                                                                                                      \
    if (seq)                                                                                                                   \
      hipLaunchKernelGGL(k_gather_seq<Q>, dim3(unsigned((c.n + 63) / 64)), dim3(64), gather_seq_lds(Q), c.stream, G, c.n,       \
                         P.tatoms.p, wsort, bsort, P.grid.p, P.fpme.p);                                                        \
    else                                                                                                                       \
      hipLaunchKernelGGL(k_gather<Q>, dim3(unsigned(ntiles)), dim3(64), lds_gather, c.stream, G, P.tstart.p, P.tatoms.p, wsort, \
                         bsort, P.grid.p, P.fpme.p);                                                                           \
    break;
    CAPSGPU_GATHER(3) CAPSGPU_GATHER(4) CAPSGPU_GATHER(5) CAPSGPU_GATHER(6) CAPSGPU_GATHER(7) CAPSGPU_GATHER(8)
#undef CAPSGPU_GATHER
  }
  LAUNCHCHECK();
  kt_end(c);
}
}  // namespace

// the gather's workgroup shape and the transforms' tile shape of this context, picked by timing on its own mesh: the
// gather by tiles or by runs of sorted atoms, and every pair of the passes along a and b
// (the plane kernel when the planes fit, lines_x 0, or the a-lines with 2, 4, 8 or 16 line pairs per workgroup) and the
// lines along b and c (4, 8 or 16 per workgroup), the whole transform with the convolution, each the best of three runs
// timed with HIP events on the context's stream. The passes run in place on the
// mesh of the last evaluation, which the next evaluation overwrites; every shape transforms each line with the same
// arithmetic, so the choice changes the time and not the forces.
void pme_tune(capsgpu_ctx& c, int out[2]) {
  Pme& P = c.pme;
  if (!P.green_ok) throw Error(-1, "tuning the mesh transforms needs one evaluation first");
  const FftPlan F = plan_of(P);
  const bool prof = c.profile;
  c.profile = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto best_of = [&](auto&& run) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      HIPCHECK(hipEventRecord(e0, c.stream));
      run();
      HIPCHECK(hipEventRecord(e1, c.stream));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    return best;
  };
  try {
    // the gather's workgroups (one per tile, or one per run of 64 sorted atoms where the mesh rows allow it), each timed
    // alone on the last evaluation's sorted atoms and mesh (before the transforms below overwrite the mesh); the same
    // forces either way
    // (the runs are kept unless the tiles are clearly faster: three launches a sample, the best of three samples)
    const GridArgs G0 = grid_args(P);
    if (G0.k[0] % 4 == 0 && P.fpme.n >= 3 * size_t(c.n)) {
      float tg[2];
      for (int shape = 0; shape < 2; ++shape)
        tg[shape] = best_of([&] {
          for (int rep = 0; rep < 3; ++rep) launch_gather(c, G0, shape == 1);
        });
      P.gather_shape = tg[0] < 0.95f * tg[1] ? 0 : 1;
    }
    // every pair of shapes over the whole transform (the b pass belongs to the plane kernel or to the line tiles, so the
    // two choices are not independent)
    float tbest = 1e30f;
    int pick_x = 8, pick_yz = P.lines_yz;
    for (int lx : {0, 2, 4, 8, 16}) {
      if (lx == 0 && !P.plane_ok) continue;
      for (int ly : {4, 8, 16}) {
        P.lines_x = lx, P.lines_yz = ly;
        const GridArgs G = grid_args(P);
        const float t = best_of([&] {   // four transforms a sample: one alone is too short to rank the shapes reliably
          for (int rep = 0; rep < 4; ++rep) fft_forward_conv_inverse(c, G, F, false);
        });
        if (t < tbest) tbest = t, pick_x = lx, pick_yz = ly;
      }
    }
    P.lines_x = pick_x, P.lines_yz = pick_yz;
  } catch (...) {
    c.profile = prof;
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    throw;
  }
  c.profile = prof;
  (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
  out[0] = P.lines_x, out[1] = P.lines_yz;
}

bool pme_presplines(capsgpu_ctx& c, AtomSplines& out) {
  Pme& P = c.pme;
  pme_setup(c);
  if (!P.counts_clear) return false;
  const GridArgs G = grid_args(P);
  const int n = c.n, ntiles = G.nt[0] * G.nt[1] * G.nt[2];
  P.theta.resize(4 * size_t(n));
  P.base.resize(4 * size_t(n));
  P.tcount.resize(size_t(ntiles));
  P.tfill.resize(2);
  P.cap = tile_capacity(n, ntiles);
  P.tile.resize(size_t(ntiles) * size_t(P.cap));
  P.ovf.resize(size_t(n));
  out = atom_splines(G, c.d_charge.p, reinterpret_cast<float4*>(P.theta.p), reinterpret_cast<int4*>(P.base.p), P.tcount.p, P.tile.p, P.cap,
                     P.ovf.p, P.tfill.p);
  P.counts_clear = false;
  P.presplined = true;
  return true;
}

void pme_drop_presplines(capsgpu_ctx& c) { c.pme.presplined = false, c.pme.counts_clear = false; }

void pme_forces(capsgpu_ctx& c, bool want) {
  Pme& P = c.pme;
  pme_setup(c);
  hipStream_t s = c.stream;
  const int n = c.n;
  const GridArgs G = grid_args(P);
  const FftPlan F = plan_of(P);
  const int ntiles = G.nt[0] * G.nt[1] * G.nt[2];
  P.theta.resize(4 * size_t(n));
  P.base.resize(4 * size_t(n));
  float4* wq = reinterpret_cast<float4*>(P.theta.p);
  int4* base = reinterpret_cast<int4*>(P.base.p);
  if (P.presplined) {
    // the integrator's launch did the per-atom part on these positions (pme_presplines sized the arrays)
    P.presplined = false;
  } else {
    P.tcount.resize(size_t(ntiles));
    P.tfill.resize(2);
    P.cap = tile_capacity(n, ntiles);
    P.tile.resize(size_t(ntiles) * size_t(P.cap));
    P.ovf.resize(size_t(n));
    // the tile counts and the overflow count start at zero: cleared here after a new grid (or an evaluation that did
    // not finish), else left clear by the previous evaluation's k_tile_rank and spread
    if (!P.counts_clear) {
      P.tcount.zero(s);
      P.tfill.zero(s);
    }
    P.counts_clear = false;
    kt_begin(c, "k_splines");
    hipLaunchKernelGGL(k_splines, dim3(blocks_for(n, 256)), dim3(256), 0, s, G, n, c.x.p, c.d_charge.p, wq, base, P.tcount.p, P.tile.p,
                       P.cap, P.ovf.p, P.tfill.p);
    LAUNCHCHECK();
    kt_end(c);
  }
  P.tstart.resize(size_t(ntiles) + 1);
  // the count scan: its blocks' offsets added by k_tile_rank when there are few blocks (one launch less), else whole
  P.tlocal.resize(size_t(ntiles) + 1);
  kt_begin(c, "k_scan_tiles");
  const int nsums = c.scan.blocks_only(P.tcount.p, P.tlocal.p, size_t(ntiles), s);
  if (!nsums) c.scan.exclusive(P.tcount.p, P.tstart.p, size_t(ntiles), s);
  kt_end(c);
  P.tatoms.resize(size_t(n));
  P.wsort.resize(4 * size_t(n));
  P.bsort.resize(4 * size_t(n));
  float4* wsort = reinterpret_cast<float4*>(P.wsort.p);
  int4* bsort = reinterpret_cast<int4*>(P.bsort.p);
  kt_begin(c, "k_tile_rank");
  hipLaunchKernelGGL(k_tile_rank, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ntiles, wq, base, nsums ? P.tlocal.p : P.tstart.p, P.tile.p,
                     P.cap, P.ovf.p, P.tfill.p, P.tatoms.p, wsort, bsort, P.tcount.p, nsums ? c.scan.tsum[0].p : nullptr, nsums,
                     nsums ? P.tstart.p : nullptr);
  LAUNCHCHECK();
  P.counts_clear = true;
  kt_end(c);
  kt_begin(c, "k_spread");
  const bool blocks = spread_blocks(G);
  const unsigned nblk = unsigned(spread_block_count(G));
  switch (G.p) {
#define CAPSGPU_SPREAD(Q)                                                                                                          \
  case Q:                                                                                                                          \
    if (blocks)                                                                                                                    \
      hipLaunchKernelGGL(k_spread_blk<Q>, dim3(nblk), dim3(kBlkThreads), 0, s, G, P.tstart.p, wsort, bsort, P.grid.p, P.tfill.p);   \
    else                                                                                                                           \
      hipLaunchKernelGGL(k_spread<Q>, dim3(unsigned(ntiles)), dim3(64), 0, s, G, P.home.p, P.tstart.p, wsort, bsort, P.grid.p,      \
                         P.tfill.p);                                                                                               \
    break;
    CAPSGPU_SPREAD(3) CAPSGPU_SPREAD(4) CAPSGPU_SPREAD(5) CAPSGPU_SPREAD(6) CAPSGPU_SPREAD(7) CAPSGPU_SPREAD(8)
#undef CAPSGPU_SPREAD
  }
  LAUNCHCHECK();
  kt_end(c);
  fft_forward_conv_inverse(c, G, F, want);
  P.fpme.resize(3 * size_t(n));
  launch_gather(c, G, gather_by_runs(P, G));
}


// ---- many contexts in one launch per kernel (the replica batch) ----------------------------------------------------
// Each grouped kernel covers every replica's workgroups: a workgroup finds its replica in a prefix array of workgroup
// counts (binary search, uniform across the workgroup) and runs that replica's body with its own arguments from a
// table in device memory; the tile kernels map the whole range of tiles to XCDs first, so a replica's neighbouring
// tiles stay together. The arithmetic per replica is the single-context kernels', so results are bit-identical.
namespace {
struct PmeRep {
  GridArgs G;
  FftPlan F;
  int n, ntiles;
  const double* x;
  const double* q;
  float4* wq;
  int4* base;
  uint32_t *tcount, *tfill, *tile, *tatoms, *ovf;   // tile: the buckets (cap slots per tile); tfill: the overflow counts
  int cap;
  int64_t* tstart;
  float4* wsort;
  int4* bsort;
  const int* home;
  float* grid;
  float2* half;
  const float2* tw;
  const float* gtab;
  const double* bmod;
  double fac, vol;
  float* fpme;
};

// the largest r with off[r] ≤ b (off[0] = 0, off[nr] = the launch's workgroups)
__device__ inline int find_rep(const int* __restrict__ off, int nr, int b) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) kg_splines(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const PmeRep& a = R[r];
  b_splines(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.G, a.n, a.x, a.q, a.wq, a.base, a.tcount, a.tile, a.cap, a.ovf, a.tfill);
}

// one workgroup of 1024 threads per replica: its tile counts' exclusive sums (tstart[ntiles] = the total); thread t
// takes a run of consecutive tiles, the runs' sums are scanned across the workgroup's 16 waves
__global__ void __launch_bounds__(1024) kg_tile_scan(const PmeRep* __restrict__ R) {
  __shared__ int64_t wsum[16];
  const PmeRep& a = R[blockIdx.x];
  const int n = a.ntiles, per = (n + 1 + 1023) / 1024, b0 = int(threadIdx.x) * per;
  int64_t s = 0;
  for (int k = 0; k < per; ++k) s += b0 + k < n ? int64_t(a.tcount[b0 + k]) : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t v = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  int64_t run = v - s;
  for (int w = 0; w < wave; ++w) run += wsum[w];
  for (int k = 0; k < per && b0 + k <= n; ++k) {
    a.tstart[b0 + k] = run;
    run += b0 + k < n ? int64_t(a.tcount[b0 + k]) : 0;
  }
}

__global__ void __launch_bounds__(256) kg_tile_rank(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const PmeRep& a = R[r];
  b_tile_rank(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.n, a.ntiles, a.wq, a.base, a.tstart, a.tile, a.cap, a.ovf, a.tfill, a.tatoms,
              a.wsort, a.bsort, a.tcount);
}

template <int P>
__global__ void __launch_bounds__(64) kg_spread(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int g = int(xcd_block(blockIdx.x, off[nr]));
  const int r = find_rep(off, nr, g);
  const PmeRep& a = R[r];
  if (g == off[r] && threadIdx.x == 0) *a.tfill = 0u;   // the replica's overflow count, for its next evaluation
  b_spread<P>(g - off[r], a.G, a.home, a.tstart, a.wsort, a.bsort, a.grid);
}

template <int P>
__global__ void __launch_bounds__(kBlkThreads) kg_spread_blk(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int g = int(xcd_block(blockIdx.x, off[nr]));
  const int r = find_rep(off, nr, g);
  const PmeRep& a = R[r];
  if (g == off[r] && threadIdx.x == 0) *a.tfill = 0u;   // the replica's overflow count, for its next evaluation
  b_spread_blk<P>(g - off[r], a.G, a.tstart, a.wsort, a.bsort, a.grid);
}

template <int P>
__global__ void __launch_bounds__(64) kg_gather(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int g = int(xcd_block(blockIdx.x, off[nr]));
  const int r = find_rep(off, nr, g);
  const PmeRep& a = R[r];
  b_gather<P>(g - off[r], a.G, a.tstart, a.tatoms, a.wsort, a.bsort, a.grid, a.fpme);
}
template <int P>
__global__ void __launch_bounds__(64) kg_gather_seq(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int g = int(xcd_block(blockIdx.x, off[nr]));
  const int r = find_rep(off, nr, g);
  const PmeRep& a = R[r];
  b_gather_seq<P>(g - off[r], a.n, a.G, a.tatoms, a.wsort, a.bsort, a.grid, a.fpme);
}

__global__ void __launch_bounds__(256) kg_fft_x(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr, int forward) {
  const int r = find_rep(off, nr, blockIdx.x);
  const PmeRep& a = R[r];
  if (forward) b_fft_x_r2c(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.G, a.F, a.tw, a.grid, a.half);
  else b_fft_x_c2r(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.G, a.F, a.tw, a.half, a.grid);
}

__global__ void __launch_bounds__(kPlaneThreads) kg_fft_xy(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr, int forward) {
  const int r = find_rep(off, nr, blockIdx.x);
  const PmeRep& a = R[r];
  b_fft_xy(int(blockIdx.x) - off[r], 0, 1, a.G, a.F, a.tw, a.grid, a.half, forward);
}

template <int Sign>
__global__ void __launch_bounds__(256) kg_fft_y(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const PmeRep& a = R[r];
  b_fft_y<Sign>(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.G, a.F, a.tw, a.half);
}

__global__ void __launch_bounds__(256) kg_fft_z_conv(const PmeRep* __restrict__ R, const int* __restrict__ off, int nr) {
  const int r = find_rep(off, nr, blockIdx.x);
  const PmeRep& a = R[r];
  b_fft_z_conv<false>(int(blockIdx.x) - off[r], off[r + 1] - off[r], a.G, a.F, a.tw, a.half, a.gtab, a.fac, kCoulomb, a.vol, a.bmod,
                      nullptr);
}
}  // namespace

struct PmeGroup {
  int nr = 0, order = 0;
  DevBuf<unsigned char> reps;                         // PmeRep[nr]
  DevBuf<int32_t> off_atoms, off_tiles, off_x, off_yz, off_z, off_xy, off_sp, off_blk, off_gt, off_gs;   // workgroup prefixes per launch
  int h_atoms = 0, h_tiles = 0, h_x = 0, h_yz = 0, h_z = 0, h_xy = 0, h_sp = 0, h_blk = 0, h_gt = 0, h_gs = 0;  // the launches' workgroups
  size_t lds_x = 0, lds_y = 0, lds_z = 0, lds_xy = 0;
};

PmeGroup* pme_group_create(const std::vector<capsgpu_ctx*>& cs, hipStream_t s) {
  auto g = std::make_unique<PmeGroup>();
  g->nr = int(cs.size());
  if (g->nr < 1) throw Error(-1, "a PME group needs a replica");
  std::vector<PmeRep> R(cs.size());
  std::vector<int32_t> oa(1, 0), ot(1, 0), ox(1, 0), oy(1, 0), oz(1, 0), oxy(1, 0), osp(1, 0), oblk(1, 0), ogt(1, 0), ogs(1, 0);
  for (size_t i = 0; i < cs.size(); ++i) {
    capsgpu_ctx& c = *cs[i];
    Pme& P = c.pme;
    if (!P.green_ok || !P.counts_clear) throw Error(-1, "a PME group needs each replica evaluated once");
    if (i == 0) g->order = P.order;
    if (P.order != g->order) throw Error(-1, "a PME group needs one spline order");
    const GridArgs G = grid_args(P);
    PmeRep& a = R[i];
    a.G = G;
    a.F = plan_of(P);
    a.n = c.n;
    a.ntiles = G.nt[0] * G.nt[1] * G.nt[2];
    a.x = c.x.p, a.q = c.d_charge.p;
    a.wq = reinterpret_cast<float4*>(P.theta.p), a.base = reinterpret_cast<int4*>(P.base.p);
    a.tcount = P.tcount.p, a.tfill = P.tfill.p, a.tile = P.tile.p, a.tatoms = P.tatoms.p, a.tstart = P.tstart.p;
    a.ovf = P.ovf.p, a.cap = P.cap;
    a.wsort = reinterpret_cast<float4*>(P.wsort.p), a.bsort = reinterpret_cast<int4*>(P.bsort.p);
    a.home = P.home.p, a.grid = P.grid.p, a.half = P.half.p, a.tw = P.tw.p, a.gtab = P.gtab.p, a.bmod = P.bmod.p;
    a.fac = kPi * kPi / (P.beta * P.beta), a.vol = P.vol, a.fpme = P.fpme.p;
    // the spread each replica takes alone: by blocks of 2×2×2 tiles or by tiles
    const bool blocks = spread_blocks(G);
    osp.push_back(osp.back() + (blocks ? 0 : a.ntiles));
    oblk.push_back(oblk.back() + (blocks ? spread_block_count(G) : 0));
    // the gather each replica takes alone: by runs of 64 sorted atoms or by tiles (the same forces)
    const bool runs = gather_by_runs(P, G);
    ogs.push_back(ogs.back() + (runs ? (c.n + 63) / 64 : 0));
    ogt.push_back(ogt.back() + (runs ? 0 : a.ntiles));
    // a replica on the plane transform (G.lx = 0) has workgroups in the plane launches only, the others in the a- and
    // b-line launches only
    const int npairs = (P.k[1] * P.k[2] + 1) / 2, mblocks = (P.h1 + G.ly - 1) / G.ly;
    const bool plane = G.lx == 0;
    oa.push_back(oa.back() + int(blocks_for(size_t(c.n), 256)));
    ot.push_back(ot.back() + a.ntiles);
    ox.push_back(ox.back() + (plane ? 0 : (npairs + G.lx - 1) / G.lx));
    oy.push_back(oy.back() + (plane ? 0 : mblocks * P.k[2]));
    oz.push_back(oz.back() + zline_blocks(P.h1, G.ly, P.k[2]) * P.k[1]);
    oxy.push_back(oxy.back() + (plane ? P.k[2] : 0));
    if (plane) g->lds_xy = std::max(g->lds_xy, plane_lds(P.k[0], P.k[1]));
    else g->lds_x = std::max(g->lds_x, (2 * size_t(G.lx) * size_t(line_stride(P.k[0])) + P.k[0]) * sizeof(float2));
    g->lds_y = std::max(g->lds_y, (2 * size_t(G.ly) * size_t(line_stride(P.k[1])) + P.k[1]) * sizeof(float2));
    g->lds_z = std::max(g->lds_z, (2 * size_t(zline_cap(G.ly, P.k[2])) * size_t(line_stride(P.k[2])) + P.k[2]) * sizeof(float2));
  }
  g->reps.upload(reinterpret_cast<const unsigned char*>(R.data()), R.size() * sizeof(PmeRep), s);
  g->off_atoms.upload(oa, s), g->off_tiles.upload(ot, s), g->off_x.upload(ox, s), g->off_yz.upload(oy, s), g->off_z.upload(oz, s);
  g->off_xy.upload(oxy, s), g->off_sp.upload(osp, s), g->off_blk.upload(oblk, s);
  g->h_sp = osp.back(), g->h_blk = oblk.back();
  g->off_gt.upload(ogt, s), g->off_gs.upload(ogs, s);
  g->h_gt = ogt.back(), g->h_gs = ogs.back();
  if (g->lds_xy > 65536)
    HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&kg_fft_xy), hipFuncAttributeMaxDynamicSharedMemorySize, int(g->lds_xy)));
  g->h_atoms = oa.back(), g->h_tiles = ot.back(), g->h_x = ox.back(), g->h_yz = oy.back(), g->h_z = oz.back(), g->h_xy = oxy.back();
  return g.release();
}

void pme_group_destroy(PmeGroup* g) { delete g; }

// the reciprocal forces of every replica into its pme.fpme: for the whole group seven launches when every replica is on
// the plane transform and shares one spread and one gather shape; one more for each of those two kernels whose
// replicas take both shapes; the line transforms' four when some replicas take them (replacing the plane kernel's two
// when all do); one more for the splines when the integrator's launch has not placed the atoms
AtomSplines pme_atom_splines(capsgpu_ctx& c) {
  Pme& P = c.pme;
  return atom_splines(grid_args(P), c.d_charge.p, reinterpret_cast<float4*>(P.theta.p), reinterpret_cast<int4*>(P.base.p), P.tcount.p,
                      P.tile.p, P.cap, P.ovf.p, P.tfill.p);
}

void pme_group_forces(PmeGroup& g, hipStream_t s, bool splined) {
  const PmeRep* R = reinterpret_cast<const PmeRep*>(g.reps.p);
  const int nr = g.nr;
  const int na = g.h_atoms, nt = g.h_tiles, nx = g.h_x, ny = g.h_yz, nz = g.h_z;
  if (!splined) {
    hipLaunchKernelGGL(kg_splines, dim3(unsigned(na)), dim3(256), 0, s, R, g.off_atoms.p, nr);
    LAUNCHCHECK();
  }
  hipLaunchKernelGGL(kg_tile_scan, dim3(unsigned(nr)), dim3(1024), 0, s, R);
  LAUNCHCHECK();
  hipLaunchKernelGGL(kg_tile_rank, dim3(unsigned(na)), dim3(256), 0, s, R, g.off_atoms.p, nr);
  LAUNCHCHECK();
  const size_t lds_gather = (size_t(g.order + 3) * (g.order + 3) * (g.order + 4) + 64 * 6) * sizeof(float);
  // the spread: replicas by blocks and the others by tiles (a launch with no workgroups is skipped)
  switch (g.order) {
#define CAPSGPU_GSPREAD(Q)                                                                                                     \
  case Q:                                                                                                                      \
    if (g.h_blk) hipLaunchKernelGGL(kg_spread_blk<Q>, dim3(unsigned(g.h_blk)), dim3(kBlkThreads), 0, s, R, g.off_blk.p, nr);  \
    if (g.h_sp) hipLaunchKernelGGL(kg_spread<Q>, dim3(unsigned(g.h_sp)), dim3(64), 0, s, R, g.off_sp.p, nr);                   \
    break;
    CAPSGPU_GSPREAD(3) CAPSGPU_GSPREAD(4) CAPSGPU_GSPREAD(5) CAPSGPU_GSPREAD(6) CAPSGPU_GSPREAD(7) CAPSGPU_GSPREAD(8)
#undef CAPSGPU_GSPREAD
  }
  LAUNCHCHECK();
  // the passes along a and b: the plane kernel for the replicas on it, the line kernels for the others (a launch with no
  // workgroups is skipped)
  const int nxy = g.h_xy;
  if (nxy) {
    hipLaunchKernelGGL(kg_fft_xy, dim3(unsigned(nxy)), dim3(kPlaneThreads), g.lds_xy, s, R, g.off_xy.p, nr, 1);
    LAUNCHCHECK();
  }
  if (nx) {
    hipLaunchKernelGGL(kg_fft_x, dim3(unsigned(nx)), dim3(256), g.lds_x, s, R, g.off_x.p, nr, 1);
    LAUNCHCHECK();
    hipLaunchKernelGGL(kg_fft_y<+1>, dim3(unsigned(ny)), dim3(256), g.lds_y, s, R, g.off_yz.p, nr);
    LAUNCHCHECK();
  }
  hipLaunchKernelGGL(kg_fft_z_conv, dim3(unsigned(nz)), dim3(256), g.lds_z, s, R, g.off_z.p, nr);
  LAUNCHCHECK();
  if (nx) {
    hipLaunchKernelGGL(kg_fft_y<-1>, dim3(unsigned(ny)), dim3(256), g.lds_y, s, R, g.off_yz.p, nr);
    LAUNCHCHECK();
    hipLaunchKernelGGL(kg_fft_x, dim3(unsigned(nx)), dim3(256), g.lds_x, s, R, g.off_x.p, nr, 0);
    LAUNCHCHECK();
  }
  if (nxy) {
    hipLaunchKernelGGL(kg_fft_xy, dim3(unsigned(nxy)), dim3(kPlaneThreads), g.lds_xy, s, R, g.off_xy.p, nr, 0);
    LAUNCHCHECK();
  }
  switch (g.order) {
#define CAPSGPU_GGATHER(Q)                                                                                                      \
  case Q:                                                                                                                       \
    if (g.h_gs) hipLaunchKernelGGL(kg_gather_seq<Q>, dim3(unsigned(g.h_gs)), dim3(64), gather_seq_lds(Q), s, R, g.off_gs.p, nr);  \
    if (g.h_gt) hipLaunchKernelGGL(kg_gather<Q>, dim3(unsigned(g.h_gt)), dim3(64), lds_gather, s, R, g.off_gt.p, nr);              \
    break;
    CAPSGPU_GGATHER(3) CAPSGPU_GGATHER(4) CAPSGPU_GGATHER(5) CAPSGPU_GGATHER(6) CAPSGPU_GGATHER(7) CAPSGPU_GGATHER(8)
#undef CAPSGPU_GGATHER
  }
  LAUNCHCHECK();
}

}  // namespace capsgpu
