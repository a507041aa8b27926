// Internal: device-wide exclusive prefix sums (uint32 or int32 counts -> int64 offsets), deterministic.
// Blocks of 2048 items scanned in LDS; up to 64 blocks, each block then adds the totals of the blocks before it (a wave
// sums them); beyond, the block totals are scanned recursively (three levels at most) and the offsets added back.
#pragma once

#include "common.hpp"

namespace capsgpu {

constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

// inclusive scan of one value per lane across a 256-thread block; returns the block total in *total
__device__ inline int64_t block_scan_incl(int64_t v, int64_t* lds, int64_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  int64_t add = 0, tot = 0;
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wave) add += lds[w];
    tot += lds[w];
  }
  *total = tot;
  __syncthreads();
  return v + add;
}

template <class T>
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const T* __restrict__ in, int64_t* __restrict__ out, size_t n,
                                                             int64_t* __restrict__ tile_sums) {
  __shared__ int64_t lds[kScanThreads / 64];
  const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
  int64_t v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? int64_t(in[base + k]) : 0;
    s += v[k];
  }
  int64_t total;
  const int64_t incl = block_scan_incl(s, lds, &total);
  int64_t run = incl - s;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k <= n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

static __global__ void __launch_bounds__(kScanThreads) k_scan_add(int64_t* __restrict__ out, size_t n, const int64_t* __restrict__ tile_offs) {
  const size_t i = size_t(blockIdx.x) * kScanTile + threadIdx.x;
  const int64_t a = tile_offs[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const size_t j = i + size_t(k) * kScanThreads;
    if (j < n) out[j] += a;
  }
}

// the offsets of at most 64 blocks summed by each block's first wave from the block totals, then added
// the exclusive offsets of at most 64 block totals, into boff[0 … 63] (LDS), by the workgroup's first wave; a barrier
// follows, so every thread of the workgroup must call it
__device__ inline void block_offsets(const int64_t* __restrict__ sums, int nblocks, int64_t* boff) {
  if (threadIdx.x < 64) {
    const int64_t v = int(threadIdx.x) < nblocks ? sums[threadIdx.x] : 0;
    int64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int64_t o = __shfl_up(inc, d, 64);
      if (int(threadIdx.x) >= d) inc += o;
    }
    boff[threadIdx.x] = inc - v;
  }
  __syncthreads();
}

static __global__ void __launch_bounds__(kScanThreads) k_scan_add_small(int64_t* __restrict__ out, size_t n, const int64_t* __restrict__ sums) {
  __shared__ int64_t off;
  if (threadIdx.x < 64) {
    int64_t v = int(threadIdx.x) < int(blockIdx.x) ? sums[threadIdx.x] : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (threadIdx.x == 0) off = v;
  }
  __syncthreads();
  const int64_t a = off;
  const size_t i = size_t(blockIdx.x) * kScanTile + threadIdx.x;
  for (int k = 0; k < kScanItems; ++k) {
    const size_t j = i + size_t(k) * kScanThreads;
    if (j < n) out[j] += a;
  }
}

// out[i] = Σ_{k<i} in[k] for i ≤ n (out has n + 1 entries: out[n] is the total)
struct Scanner {
  DevBuf<int64_t> tsum[4], toff[4];
  template <class T>
  void exclusive(const T* in, int64_t* out, size_t n, hipStream_t s) {
    run(in, out, n, 0, s);
  }
  // The blocks' scans alone, for a consumer that adds the block offsets itself: when n + 1 items take at most 64 blocks
  // of kScanTile, out[i] = Σ of in over i's block before i, tsum[0][b] = block b's total, and the block count is
  // returned; otherwise 0, with nothing launched.
  template <class T>
  int blocks_only(const T* in, int64_t* out, size_t n, hipStream_t s) {
    const size_t tiles = (n + 1 + kScanTile - 1) / kScanTile;
    if (tiles > 64) return 0;
    tsum[0].resize(tiles + 1);
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3(unsigned(tiles)), dim3(kScanThreads), 0, s, in, out, n, tsum[0].p);
    LAUNCHCHECK();
    return int(tiles);
  }

 private:
  template <class T>
  void run(const T* in, int64_t* out, size_t n, int depth, hipStream_t s) {
    // out must hold n + 1 values: the extra one gets the total
    const size_t tiles = (n + 1 + kScanTile - 1) / kScanTile;
    if (depth >= 3) throw Error(-1, "scan: too many elements");
    DevBuf<int64_t>& sums = tsum[depth];
    sums.resize(tiles + 1);
    // scan n + 1 items, the last being 0, so out[n] is the total
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3(unsigned(tiles)), dim3(kScanThreads), 0, s, in, out, n, sums.p);
    LAUNCHCHECK();
    if (tiles > 1 && tiles <= 64) {
      hipLaunchKernelGGL(k_scan_add_small, dim3(unsigned(tiles)), dim3(kScanThreads), 0, s, out, n + 1, sums.p);
      LAUNCHCHECK();
    } else if (tiles > 1) {
      DevBuf<int64_t>& offs = toff[depth];
      offs.resize(tiles + 1);
      run<int64_t>(sums.p, offs.p, tiles, depth + 1, s);
      hipLaunchKernelGGL(k_scan_add, dim3(unsigned(tiles)), dim3(kScanThreads), 0, s, out, n + 1, offs.p);
      LAUNCHCHECK();
    }
  }
};

}  // namespace capsgpu
