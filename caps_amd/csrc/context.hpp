// Internal: the device state of one capsgpu context and the host launchers of each kernel family.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/capsgpu.h"
#include "common.hpp"
#include "scan.hpp"

namespace capsgpu {

// the PME half spectrum's rows along a: K1/2 + 1 elements at a stride rounded up to 16 (128 bytes), so runs of
// consecutive a-frequencies start on cache lines; the padding is never read or written
__host__ __device__ constexpr int half_stride(int K1) { return (K1 / 2 + 1 + 15) & ~15; }
struct Learned;

// Geometry of one cluster-list build, fixed on the host the way core/src/nbnxm.cpp:40-47 and 146-164 fix it
struct ListGeom {
  int n = 0, nu = 1, nv = 1, ncols = 1;
  int ru = 0, rv = 0, rq = 0, nwin = 0;
  int qlo[3] = {0, 0, 0}, span[3] = {1, 1, 1};
  int ncodes = 0, zero_code = 0;
  double rl = 0, rl2 = 0, dw = 0;
  float rl2f = 0;
  double cell[12] = {};       // origin, a, b, c
  double bc[3] = {}, ca[3] = {}, ab[3] = {}, vsigned = 0;   // cross products and a·(b×c) for to_fractional
  int64_t nclusters = 0, ns = 0, nentries = 0;
};

// The cluster list on the device: the half list exactly as nbnxm::List holds it, plus the derived full list the pair
// kernel walks (each cluster pair from both sides, so no kernel writes another cluster's forces)
struct DevList {
  ListGeom g;
  bool built = false;
  // binning
  DevBuf<double> fr;                    // 3 per atom: wrapped fractional coordinates
  DevBuf<uint32_t> col, bucket, rank;
  DevBuf<int32_t> colcnt, bfill;
  DevBuf<int64_t> cstart, ccl;
  DevBuf<uint32_t> ccls;                // clusters per column
  DevBuf<uint32_t> perm, iperm, ccol;
  DevBuf<double> cw, frs;               // 3 per slot
  DevBuf<double> blo, bhi, org;         // 3 per cluster
  DevBuf<double> wlo, whi;
  DevBuf<float> bx, by, bz;
  DevBuf<uint8_t> real;
  // excluded pairs by cluster pair
  DevBuf<uint32_t> xcount;
  DevBuf<int64_t> xoff;
  DevBuf<int32_t> xcnt;
  DevBuf<uint32_t> xcj;
  DevBuf<uint16_t> xbits, x14;
  // the half list
  DevBuf<uint32_t> wcount;              // entries per (i cluster, column window)
  DevBuf<int64_t> woff;
  DevBuf<int64_t> first;
  DevBuf<uint32_t> e_cj, e_ci;
  DevBuf<uint16_t> e_code, e_mask, e_ex, e_ex14;
  DevBuf<double> off;                   // 3 per slot
  DevBuf<int32_t> ioff;
  DevBuf<int32_t> qint;
  DevBuf<double> qvec;
  DevBuf<int32_t> mirror;               // per image code: the code of the opposite image
  // the full outer list: per cluster its own entries (half-list order), then the mirrored ones (by source entry)
  DevBuf<uint32_t> own, mir, mfill;
  DevBuf<uint32_t> fcount;
  DevBuf<int64_t> fstart;
  DevBuf<uint32_t> fsrc;                // source half entry, bit 31 set for a mirrored one
  // pruning
  DevBuf<uint16_t> pmask, pex, pex14;   // per half entry
  DevBuf<uint32_t> pflag;               // per full entry
  DevBuf<int64_t> ppos;
  DevBuf<uint2> pent;                   // pruned full entries: {cj | code << 24, mask | ex << 16}
  DevBuf<uint16_t> pent14;
  DevBuf<int64_t> pfirst;
  int64_t npruned = 0;
  // per slot, every step / every build
  DevBuf<float4> xq;                    // X, Y, Z relative to the cluster origin, charge
  DevBuf<float2> pp;                    // LJ factors P12, P6
  DevBuf<float4> fnb;                   // per slot: pair forces
  DevBuf<double> epart;                 // per cluster: vdw, coulomb, Σ S⊗f by lattice coefficient (9)
  DevBuf<int8_t> qint8;                 // 4 per image code: its lattice coefficients
};

struct Pme {
  int k[3] = {0, 0, 0}, order = 5, h1 = 0, hs = 0;   // hs: the half spectrum's row stride (half_stride)
  double beta = 0, cell[12] = {}, hi[3][3] = {}, vol = 0;
  bool green_ok = false;
  DevBuf<float> grid;                   // real K1·K2·K3
  DevBuf<float2> half;                  // (K1/2 + 1)·K2·K3
  DevBuf<float> gtab;                   // influence function on the half spectrum
  DevBuf<double> bmod;                  // |b(m)|² per axis, K1 + K2 + K3
  DevBuf<float> theta;                  // per atom: the B-spline offsets w of the three axes and the charge (float4)
  DevBuf<int32_t> base;                 // per atom: the first grid index per axis and the home tile (int4)
  DevBuf<int32_t> home;                 // per axis and tile: the home tiles that reach it (count + 4)
  DevBuf<uint32_t> tile, tcount, tatoms;   // tile: each tile's bucket of cap slots (atoms filled unordered); tatoms: the
                                        // atoms in tile order, each tile's in index order
  DevBuf<uint32_t> ovf, tfill;          // the atoms past their bucket's cap; tfill: their count and a workgroup counter
  int cap = 32;                         // slots per bucket
  bool counts_clear = false;            // tcount and tfill are zero on the device (k_tile_rank clears them)
  bool presplined = false;              // the next evaluation's per-atom bookkeeping is done (by the integrator's launch)
  DevBuf<float> wsort;                  // theta in tatoms's order (float4 per atom)
  DevBuf<int32_t> bsort;                // base in tatoms's order (int4 per atom)
  DevBuf<int64_t> tstart;
  DevBuf<int64_t> tlocal;               // the tile counts' scan within each scan block (k_tile_rank adds the block offsets)
  DevBuf<float> fpme;                   // 3 per atom
  DevBuf<float2> tw;                    // twiddles of the three axes
  DevBuf<double> epart;                 // energy and virial partials
  int64_t erows = 0;                    // rows of epart written by the last evaluation with energies
  int lines_x = 0, lines_yz = 16;       // the transforms' tile shape (line pairs along a, lines along b and c); lines_x 0:
                                        // a c-plane per workgroup for a and b together when it fits in LDS (else 8 pairs)
  bool plane_ok = false;                // the grid's c-planes fit the plane transform (even K2, LDS)
  int gather_shape = 1;                 // the gather's workgroups: 0 one per tile, 1 one per run of 64 sorted atoms (K1 a
                                        // multiple of 4; else per tile); the same forces either way
};

// slots per tile of the PME buckets: the mean filling with a margin of eight standard deviations of a random filling
// and 16 more, at least 32, a multiple of 8
inline int tile_capacity(int64_t n, int64_t ntiles) {
  const double mean = double(n) / double(ntiles > 0 ? ntiles : 1);
  const int c = int(std::ceil(2.0 * mean + 8.0 * std::sqrt(mean) + 16.0));
  return std::max(32, (c + 7) / 8 * 8);
}

struct Bonded {
  int nb = 0, na = 0, nf = 0;
  DevBuf<int32_t> bij, aijk, fij;
  DevBuf<double> bkr, akt, fp;
  DevBuf<int32_t> aoff;                 // per atom: its slots (CSR)
  DevBuf<int32_t> aslot;                // a member's slot: bond t 2t + k, angle t 2nb + 3t + k, FENE t 2nb + 3na + 2t + k
  DevBuf<int32_t> spos;                 // per member slot: its place in the per-atom order (the CSR above)
  DevBuf<double> slot;                  // 3 per place: the term's force on that member, by atom
  DevBuf<double> epart;                 // per block of terms (or per range): bond, angle, virial, w[6]
  int eblocks = 0;                      // rows of epart
  // ranges of consecutive atoms whose slots fit a workgroup's LDS (the bonded terms and force sum in one launch)
  bool ranged = false;
  int nranges = 0;
  DevBuf<int32_t> rstart, roff;         // range r: atoms rstart[r] … rstart[r+1] − 1; its terms rterm[roff[r]] …
  DevBuf<int4> rterm;                   // term (bonds, angles, FENE bonds in one numbering; top bit: its energy counted
                                        // here) and its member atoms
};


}  // namespace capsgpu

struct capsgpu_ctx {
  int n = 0, nt = 0;
  int device = 0;                       // the HIP device current at creation: the stream's and every array's
  capsgpu_energy_opts opt{};
  hipStream_t stream = nullptr;
  // force field
  std::vector<int32_t> type;
  std::vector<double> charge, mass;
  std::vector<double> eps, sig6;        // per type pair (mixed)
  std::vector<double> f12, f6;          // per type, when the pair coefficients factorise (geometric mixing)
  bool comb = false, lj96 = false;
  std::vector<int32_t> excl_off, excl_idx, p14_off, p14_idx;
  double f14 = 0;
  std::vector<double> type_count;
  capsgpu::DevBuf<int32_t> d_type, d_excl_off, d_excl_idx, d_p14_off, d_p14_idx;
  capsgpu::DevBuf<double> d_charge, d_minv;
  capsgpu::DevBuf<float> d_f12, d_f6, d_c12, d_c6;
  // state
  double cell[12] = {};
  capsgpu::DevBuf<double> x, v, f;      // 3 per atom, atom order
  capsgpu::DevBuf<double> x0, x0p;      // positions at the last build / prune
  capsgpu::DevBuf<double> terms;        // 2 × 16 partial sums reduced on the device
  capsgpu::DevBuf<double> partials;
  capsgpu::DevBuf<uint32_t> flags;      // [0] rebuild, [1] prune, [2] error (FENE past R0)
  uint32_t* h_flags = nullptr;          // pinned
  double* h_terms = nullptr;            // pinned
  capsgpu::DevList L;
  capsgpu::Pme pme;
  capsgpu::Bonded bonded;
  capsgpu::Scanner scan;
  int builds = 0, prunes = 0;
  bool state_set = false, pruned_ok = false;
  double stage_ms[6] = {0, 0, 0, 0, 0, 0};
  bool timing = false;
  hipEvent_t ev[8] = {};
  // stage profiling of every evaluation (dynamics included): event sets recorded on the stream, summed when harvested
  bool profile = false;
  std::string kfilter;                  // when not empty: only launches of this kernel are timed, and no stages
  bool kt_open = false;                 // a launch timer is open (kt_begin recorded its first event)
  std::vector<std::array<hipEvent_t, 7>> pev;
  size_t pev_used = 0, pslot = 0;
  double prof_ms[6] = {0, 0, 0, 0, 0, 0};
  int64_t prof_n[6] = {0, 0, 0, 0, 0, 0};
  // per-kernel launch timing while profiling: event pairs around each launch, summed per kernel name when harvested
  struct KTime {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<KTime> kev;
  size_t kev_used = 0;
  std::vector<std::pair<std::string, std::pair<double, int64_t>>> ktot;
  // learned r-RESPA: the network and its state persist across dynamics runs of the context
  capsgpu::Learned* learned = nullptr;
  ~capsgpu_ctx();
};

namespace capsgpu {
// nbnxm_build.hip
void list_build(capsgpu_ctx& c);
// nbnxm_prune.hip
void list_positions(capsgpu_ctx& c);
void list_prune(capsgpu_ctx& c, double radius);
// nbnxm_pairs.hip: pair forces into L.fnb by slot; energies/virial partials into c.partials[0..] when want
void pair_forces(capsgpu_ctx& c, bool want);
// pme.hip: reciprocal forces into pme.fpme (atom order); energy and virial partials when want
void pme_setup(capsgpu_ctx& c);
// the mesh pme_setup picks for a cell (points per axis), without setting anything up
void pme_grid_dims(const capsgpu_ctx& c, const double cell[12], int k[3]);
void pme_forces(capsgpu_ctx& c, bool want);
// picks the transforms' tile shape by timing (out: line pairs along a, lines along b and c per workgroup)
void pme_tune(capsgpu_ctx& c, int out[2]);
struct AtomSplines;
// The per-atom half of the next evaluation's PME bookkeeping for a launch that moves the atoms to do on the positions it
// writes (the mesh set up for the current cell, the arrays sized): false when the tile counters are not clear (a new
// grid, an unfinished evaluation); else the next pme_forces leaves its splines launch out.
bool pme_presplines(capsgpu_ctx& c, AtomSplines& out);
// after a failure between such a launch and its evaluation: the next evaluation starts its bookkeeping afresh
void pme_drop_presplines(capsgpu_ctx& c);
// many contexts' reciprocal forces, one launch per kernel (each context evaluated once before; one spline order)
struct PmeGroup;
PmeGroup* pme_group_create(const std::vector<capsgpu_ctx*>& cs, hipStream_t s);
// splined: the step group's opening launch did the per-atom part of the bookkeeping (pme_atom_splines' arrays)
void pme_group_forces(PmeGroup& g, hipStream_t s, bool splined = false);
// the per-atom part's arguments for the context's current mesh and arrays (as the last evaluation sized them)
AtomSplines pme_atom_splines(capsgpu_ctx& c);
void pme_group_destroy(PmeGroup* g);
// bonded.hip
void bonded_setup(capsgpu_ctx& c, const capsgpu_ff& ff);
void bonded_forces(capsgpu_ctx& c, bool want);
// integrate.hip
void sum_forces(capsgpu_ctx& c, bool nb, bool rec, bool bond, double* fdst);
// the bonded terms (when bond) and the force sum into fdst: one launch by atom ranges, or k_bonded + k_sum_forces
void bonded_and_sum(capsgpu_ctx& c, bool want, bool nb, bool rec, bool bond, double* fdst);
void check_moves(capsgpu_ctx& c, double lim_build, double lim_prune);
// partial sums: slot k of c.partials; reduce_terms folds them (fixed order) into c.terms
enum Part { P_VDW = 0, P_COUL, P_W0, P_BOND = 8, P_ANGLE, P_BVIR, P_BW0, P_RECE = 16, P_RW0, NPART = 24 };
void zero_partials(capsgpu_ctx& c);
}  // namespace capsgpu
