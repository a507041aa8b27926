// Internal: the per-atom half of the PME tile bookkeeping — each atom's B-spline offsets, charge, first mesh point, home
// tile and place in that tile's bucket — as device code shared by the PME splines launch (positions read from memory)
// and the integrator's opening launch (positions it has just written, so the evaluation that follows skips a launch).
#pragma once

#include "common.hpp"

namespace capsgpu {

// the mesh and the arrays the per-atom part writes (wq == nullptr: nothing to do)
struct AtomSplines {
  double hi[3][3];                      // rows of the cell's inverse (fractional coordinate d = hi[d] · x)
  int k[3], nt[3], p, cap;              // mesh points and 4x4x4 tiles per axis, spline order, bucket slots
  const double* q;                      // charges
  float4* wq;                           // per atom: w per axis (the fractional offset of the first point), charge
  int4* base;                           // per atom: first mesh point per axis, home tile
  uint32_t *tcount, *bucket, *ovf, *cnt;   // tile counts, buckets (cap per tile), overflow list and its count
};

// Runs of a wave's lanes: a valid lane whose key equals the previous lane's continues that lane's run. Returns the lane's
// place in its run and its run's first lane; the first lane also gets the run's length. Consecutive atoms (neighbours
// along a chain) mostly share a tile, so one atomic per run replaces one per atom. Every lane of the wave must call it.
__device__ inline int wave_run(bool valid, uint32_t key, int* head, int* len) {
  const int lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(key, 1, 64);
  const uint64_t vmask = __ballot(valid);
  const bool prev_valid = lane > 0 && ((vmask >> (lane - 1)) & 1ull);
  const bool first = valid && !(prev_valid && prev == key);
  const uint64_t heads = __ballot(first);
  const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;   // lanes 0 … lane
  const uint64_t below = heads & upto;
  *head = below ? 63 - __builtin_clzll(below) : lane;
  const uint64_t ends = (heads | ~vmask) & ~upto;                       // run boundaries past this lane
  *len = (ends ? __builtin_ffsll((long long)ends) - 1 : 64) - lane;
  return lane - *head;
}

// Atom i (valid) at x: fractional coordinate s = hi · x, u = (s − ⌊s⌋) K, the first point ⌊u⌋ − p + 1 (mod K), w = u − ⌊u⌋,
// the home tile of the first point; then a place in the home tile's bucket, one atomic per run of lanes in the same tile
// (the run's lanes take consecutive slots from its result), or past the bucket's cap a place in the overflow list.
// Every lane of the wave must call it.
__device__ inline void spline_atom(const AtomSplines& A, bool valid, int i, const double x[3]) {
  int tile = 0;
  if (valid) {
    float w[3];
    int b[3];
    for (int d = 0; d < 3; ++d) {
      const double s = A.hi[d][0] * x[0] + A.hi[d][1] * x[1] + A.hi[d][2] * x[2];
      const double u = (s - floor(s)) * A.k[d];
      const double fl = floor(u);
      w[d] = float(u - fl);
      int t = (int(fl) - A.p + 1) % A.k[d];
      if (t < 0) t += A.k[d];
      b[d] = t;
    }
    tile = ((b[2] >> 2) * A.nt[1] + (b[1] >> 2)) * A.nt[0] + (b[0] >> 2);
    A.wq[i] = make_float4(w[0], w[1], w[2], float(A.q[i]));
    A.base[i] = make_int4(b[0], b[1], b[2], tile);
  }
  int head, len;
  const int pos = wave_run(valid, uint32_t(tile), &head, &len);
  uint32_t at = 0;
  if (valid && pos == 0) at = atomicAdd(&A.tcount[tile], uint32_t(len));
  at = __shfl(at, head, 64);
  if (valid) {
    const uint32_t slot = at + uint32_t(pos);
    if (slot < uint32_t(A.cap)) A.bucket[size_t(tile) * A.cap + slot] = uint32_t(i);
    else A.ovf[atomicAdd(&A.cnt[0], 1u)] = uint32_t(i);
  }
}

}  // namespace capsgpu
