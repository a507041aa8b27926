// The C boundary of capsgpu (include/capsgpu.h): exceptions become return codes and a thread-local message, as capi's
// guard() does (capi/src/caps_c.cpp:353, 373-383). Host orchestration of one force evaluation (the order of
// core/src/field.cpp:1279-2539's cluster path) and of the dynamics loop (core/src/dynamics.cpp:71-815).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <string>

#include "batch.hpp"
#include "context.hpp"
#include "host_rng.hpp"
#include "mlp.hpp"
#include "prof.hpp"
#include "step.hpp"

using namespace capsgpu;

// the allocation-free evaluation's device scratch (one copy per device, held until the process ends)
constexpr size_t kScratchBytes = size_t(96) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
thread_local std::string g_error;

template <class F>
int32_t guard(F&& f) {
  try {
    f();
    return CAPSGPU_OK;
  } catch (const Error& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return CAPSGPU_EINVAL;
  }
}

double erfc_as(double x) {   // erfc_exp (field.cpp:585-590): A&S 7.1.26
  const double P = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027, A5 = 1.061405429;
  const double t = 1.0 / (1.0 + P * x);
  return t * (A1 + t * (A2 + t * (A3 + t * (A4 + t * A5)))) * std::exp(-x * x);
}

// Ewald β with erfc(β r_c) = rtol by bisection (core/src/kspace.cpp:35-43)
double ewald_beta(double cutoff, double rtol) {
  double lo = 0, hi = 20.0 / cutoff;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    (std::erfc(mid * cutoff) > rtol ? lo : hi) = mid;
  }
  return 0.5 * (lo + hi);
}

double cell_volume(const double* c) {
  const double *a = c + 3, *b = c + 6, *cc = c + 9;
  return std::fabs(a[0] * (b[1] * cc[2] - b[2] * cc[1]) + a[1] * (b[2] * cc[0] - b[0] * cc[2]) + a[2] * (b[0] * cc[1] - b[1] * cc[0]));
}

bool pme_on(const capsgpu_ctx& c) { return c.opt.coulomb && c.opt.electrostatics == 1 && !c.opt.no_reciprocal; }
bool ewald_on(const capsgpu_ctx& c) { return c.opt.coulomb && c.opt.electrostatics == 1; }

// the energy terms from the reduced partial sums of one evaluation (field.cpp:2376-2537)
struct Sums {
  double pair[11], slot[9], rec[7], bond[9];
};

// nonb: the pair part (pairs, DSF self, tail); recip: PME's reciprocal part with the Ewald self and background terms
void assemble(const capsgpu_ctx& c, const Sums& S, bool nonb, bool recip, bool bonded, double* t) {
  for (int k = 0; k < CAPSGPU_NTERMS; ++k) t[k] = 0;
  double w[6] = {0, 0, 0, 0, 0, 0}, vir = 0;
  double q2 = 0, qt = 0;
  for (double q : c.charge) q2 += q * q, qt += q;
  const double rc = c.opt.cutoff, V = cell_volume(c.cell);
  if (recip) {
    t[CAPSGPU_T_COULOMB] += S.rec[0];
    vir += S.rec[1] + S.rec[2] + S.rec[3];
    for (int q = 0; q < 6; ++q) w[q] += S.rec[1 + q];
    const double beta = c.pme.beta;
    t[CAPSGPU_T_COULOMB] -= kCoulomb * beta / std::sqrt(kPi) * q2;
    if (std::fabs(qt) > 1e-10) {
      const double eb = -kCoulomb * kPi * qt * qt / (2 * V * beta * beta);
      t[CAPSGPU_T_COULOMB] += eb;
      vir += 3 * eb;
      for (int q = 0; q < 3; ++q) w[q] += eb;
    }
  }
  if (bonded) {
    t[CAPSGPU_T_BOND] = S.bond[0];
    t[CAPSGPU_T_ANGLE] = S.bond[1];
    vir += S.bond[2];
    for (int q = 0; q < 6; ++q) w[q] += S.bond[3 + q];
  }
  if (nonb) {
    // the full list counts every pair twice
    t[CAPSGPU_T_VDW] += 0.5 * S.pair[0];
    t[CAPSGPU_T_COULOMB] += 0.5 * S.pair[1];
    double W[9];
    for (int q = 0; q < 9; ++q) W[q] = S.slot[q];
    const double* A[3] = {c.cell + 3, c.cell + 6, c.cell + 9};
    for (int L = 0; L < 3; ++L)
      for (int u = 0; u < 3; ++u)
        for (int v = 0; v < 3; ++v) W[3 * u + v] += 0.5 * A[L][u] * S.pair[2 + 3 * L + v];
    vir += W[0] + W[4] + W[8];
    w[0] += W[0], w[1] += W[4], w[2] += W[8];
    w[3] += 0.5 * (W[1] + W[3]), w[4] += 0.5 * (W[2] + W[6]), w[5] += 0.5 * (W[5] + W[7]);
    if (c.opt.coulomb && !ewald_on(c)) {
      const double a = c.opt.dsf_alpha, erfc_rc = erfc_as(a * rc);
      const double f0 = erfc_rc / (rc * rc) + 2 * a / std::sqrt(kPi) * std::exp(-a * a * rc * rc) / rc;
      t[CAPSGPU_T_COULOMB] -= (erfc_rc / (2 * rc) + f0 * rc / 2 + a / std::sqrt(kPi)) * q2 * kCoulomb;
    }
    if (c.opt.tail) {
      const double rc2 = rc * rc, rc3 = rc2 * rc, rc9 = rc3 * rc3 * rc3;
      double et = 0, wt = 0;
      for (int a = 0; a < c.nt; ++a)
        for (int b = 0; b < c.nt; ++b) {
          const size_t tp = size_t(a) * c.nt + b;
          const double eps = c.eps[tp], s6 = c.sig6[tp], s12 = s6 * s6, nn = c.type_count[size_t(a)] * c.type_count[size_t(b)];
          if (c.lj96) {
            const double s3 = std::sqrt(s6), rc6 = rc3 * rc3;
            et += nn * 2 * kPi * eps * s6 * (s3 - 3 * rc3) / (3 * rc6);
            wt += nn * 3 * 2 * kPi * eps * s6 * (s3 - 2 * rc3) / rc6;
          } else {
            et += nn * 2 * kPi * 4 * eps * (s12 / (9 * rc9) - s6 / (3 * rc3));
            wt += nn * 2 * kPi * 4 * eps * (4 * s12 / (3 * rc9) - 2 * s6 / rc3);
          }
        }
      t[CAPSGPU_T_VDW] += et / V;
      vir += wt / V;
      t[CAPSGPU_T_TAIL_ENERGY] = et / V, t[CAPSGPU_T_TAIL_VIRIAL] = wt / V;
      for (int q = 0; q < 3; ++q) w[q] += wt / V / 3;
    }
  }
  t[CAPSGPU_T_VIRIAL] = vir;
  for (int q = 0; q < 6; ++q) t[CAPSGPU_T_WXX + q] = w[q];
}

void read_flags(capsgpu_ctx& c) {
  HIPCHECK(hipMemcpyAsync(c.h_flags, c.flags.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
}

// the summed stage times of the recorded evaluations (waits for the stream)
void prof_harvest(capsgpu_ctx& c) {
  kt_harvest(c);
  if (!c.pev_used) return;
  HIPCHECK(hipStreamSynchronize(c.stream));
  for (size_t q = 0; q < c.pev_used; ++q)
    for (int k = 0; k < 6; ++k) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, c.pev[q][size_t(k)], c.pev[q][size_t(k) + 1]));
      c.prof_ms[k] += ms;
      ++c.prof_n[k];
    }
  c.pev_used = 0;
}

// a set of events for the next profiled evaluation
void prof_begin(capsgpu_ctx& c) {
  if (!c.profile || c.timing || !c.kfilter.empty()) return;
  if (c.pev_used == c.pev.size()) {
    if (c.pev.size() >= 1024) {
      prof_harvest(c);
    } else {
      std::array<hipEvent_t, 7> e{};
      for (auto& x : e) HIPCHECK(hipEventCreate(&x));
      c.pev.push_back(e);
    }
  }
  c.pslot = c.pev_used++;
}

void ev_record(capsgpu_ctx& c, int k) {
  if (c.timing) HIPCHECK(hipEventRecord(c.ev[k], c.stream));
  else if (c.profile && c.kfilter.empty()) HIPCHECK(hipEventRecord(c.pev[c.pslot][size_t(k)], c.stream));
}

// forces at the device positions c.x into c.f; the energy terms into t when want. The list follows the reference's
// rules (field.cpp:1286-1312): rebuilt when an atom moved more than skin/2 since the build (flag 0), pruned at
// r_c + b (b = min(0.8, skin)) when one moved more than b/2 since the prune (flag 1) and after every build.
}  // namespace

namespace capsgpu {
// the calling thread's error message for capsgpu_last_error (entry points defined in other files)
void set_last_error(const std::string& m) { g_error = m; }

// parts: the terms to evaluate (-1: the options'; bit 8 leaves PME's reciprocal part out of the non-bonded terms),
// fdst: where the forces go (nullptr: c.f)
void evaluate(capsgpu_ctx& c, bool want, double* t, bool flags_known, int parts, double* fdst) {
  if (parts < 0) parts = c.opt.parts;
  if (!fdst) fdst = c.f.p;
  const bool nonb = parts & 2, bonded = parts & 1, recip = pme_on(c) && (parts & 6) && !(parts & 8);
  const double b = std::min(0.8, c.opt.skin);
  prof_begin(c);
  ev_record(c, 0);
#ifndef CAPSGPU_ENABLE_PAIR_PATH
  // held back: the first GPU runs of the cluster-pair path ended with a memory fault or hung queue whose cause is not
  // shown; until it is, the path raises instead of running (builds for CPU emulation define CAPSGPU_ENABLE_PAIR_PATH)
  if (nonb) throw Error(CAPSGPU_EDEVICE, "the cluster pair path (list build, prune, pair kernel) is held back in this build");
#endif
  if (nonb) {
    if (!flags_known && c.L.built) {
      c.flags.zero(c.stream);
      check_moves(c, 0.25 * c.opt.skin * c.opt.skin, 0.25 * b * b);
      read_flags(c);
    }
    if (c.h_flags[3]) throw Error(CAPSGPU_EFIELD, "positions are not finite numbers");
    const bool rebuild = !c.L.built || c.h_flags[0];
    if (rebuild) {
      list_build(c);
      read_flags(c);
      if (c.h_flags[3]) throw Error(CAPSGPU_EFIELD, "positions are not finite numbers");
      HIPCHECK(hipMemcpyAsync(c.x0.p, c.x.p, 3 * size_t(c.n) * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
      ++c.builds;
      c.pruned_ok = false;
    }
    ev_record(c, 1);
    list_positions(c);
    if (!c.pruned_ok || c.h_flags[1]) {
      list_prune(c, c.opt.cutoff + b);
      HIPCHECK(hipMemcpyAsync(c.x0p.p, c.x.p, 3 * size_t(c.n) * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
      ++c.prunes;
      c.pruned_ok = true;
    }
    ev_record(c, 2);
    pair_forces(c, want);
  } else {
    for (int k = 1; k <= 2; ++k) ev_record(c, k);
  }
  ev_record(c, 3);
  if (recip) pme_forces(c, want);
  ev_record(c, 4);
  bonded_and_sum(c, want, nonb, recip, bonded, fdst);
  ev_record(c, 5);
  if (want) {
    Sums S{};
    double* out = c.terms.p;
    if (nonb) {
      colsum(c, c.L.epart.p, c.L.g.nclusters, 11, out);
      c.partials.resize(9 * size_t(blocks_for(c.L.g.ns, 256)));
      slot_virial(c, c.partials.p);
      colsum(c, c.partials.p, blocks_for(c.L.g.ns, 256), 9, out + 11);
    }
    if (recip) {
      colsum(c, c.pme.epart.p, c.pme.erows, 7, out + 20);
    }
    if (bonded) colsum(c, c.bonded.epart.p, c.bonded.eblocks, 9, out + 27);
    HIPCHECK(hipMemcpyAsync(c.h_terms, out, 36 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    std::memcpy(S.pair, c.h_terms, sizeof S.pair);
    std::memcpy(S.slot, c.h_terms + 11, sizeof S.slot);
    if (recip) std::memcpy(S.rec, c.h_terms + 20, sizeof S.rec);
    if (bonded) std::memcpy(S.bond, c.h_terms + 27, sizeof S.bond);
    assemble(c, S, nonb, recip, bonded, t);
  }
  ev_record(c, 6);
}
}  // namespace capsgpu

namespace {
void evaluate(capsgpu_ctx& c, bool want, double* t, bool flags_known) { capsgpu::evaluate(c, want, t, flags_known, -1, nullptr); }

void upload_cell(capsgpu_ctx& c, const double* cell) {
  bool same = true;
  for (int k = 0; k < 12; ++k) same = same && c.cell[k] == cell[k];
  if (!same) {
    for (int k = 0; k < 12; ++k) c.cell[k] = cell[k];
    c.L.built = false;   // a new cell: the list is built again (the reference may rescale instead, field.cpp:1298-1302)
  }
  if (!(cell_volume(cell) > 1e-9)) throw Error(CAPSGPU_EINVAL, "the cell must have a volume (periodic along a, b and c)");
}

// Evaluator::compute on the device state: forces and terms at the host positions x (field.cpp:1249-2539)
void compute_body(capsgpu_ctx* c, const double* x, const double cell[12], double* f, double* terms, int32_t want) {
  upload_cell(*c, cell);
  for (size_t k = 0; k < 3 * size_t(c->n); ++k)
    if (!std::isfinite(x[k])) throw Error(CAPSGPU_EFIELD, "positions are not finite numbers");
  c->x.upload(x, 3 * size_t(c->n), c->stream);
  c->state_set = true;
  c->timing = true;
  c->flags.zero(c->stream);
  std::memset(c->h_flags, 0, 4 * sizeof(uint32_t));
  double t[CAPSGPU_NTERMS];
  evaluate(*c, want != 0 || terms != nullptr, t, false);
  read_flags(*c);
  if (c->h_flags[2]) {
    c->flags.zero(c->stream);
    throw Error(CAPSGPU_EFIELD, "a FENE bond is stretched to R0 or past; lower the time step or relax the structure first");
  }
  if (f) c->f.download(f, 3 * size_t(c->n), c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  float ms[6];
  for (int k = 0; k < 6; ++k) HIPCHECK(hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]));
  for (int k = 0; k < 6; ++k) c->stage_ms[k] = ms[k];
  c->timing = false;
  if (terms) std::memcpy(terms, t, sizeof t);
}

// The allocation-free evaluation's device memory: one block per device, holding the per-evaluation scratch of the
// bonded terms and PME's reciprocal part (mesh, spectrum, influence function, spline and tile arrays, member slots,
// partial sums). Its device address is looked up once per device; calls on one device share it, so they are serialised
// here, and calls from elsewhere that overlap one of them on the same device race.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// the scratch block's address on device dev (that device made current for the lookup, the caller's restored)
unsigned char* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw Error(CAPSGPU_EDEVICE, "device index out of range");
  if (!g_scratch[dev]) {
    int cur = 0;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != dev) HIPCHECK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHECK(hipSetDevice(cur));
    HIPCHECK(e);
    if (!p) throw Error(CAPSGPU_EDEVICE, "no device address for the scratch block");
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch[dev];
}

// For one evaluation: the context's per-evaluation arrays handed to the arena (their sizes kept, their contents not:
// the cached PME tables are rebuilt and the tile counters cleared), and given back afterwards, the PME set-up then
// redone for the owned arrays too.
class ScratchBinding {
 public:
  ScratchBinding(capsgpu_ctx& c, Arena& a) : c_(c) {
    Pme& P = c.pme;
    bind(P.grid, a), bind(P.half, a), bind(P.gtab, a), bind(P.bmod, a), bind(P.theta, a), bind(P.base, a), bind(P.home, a);
    bind(P.tile, a), bind(P.tcount, a), bind(P.tfill, a), bind(P.ovf, a), bind(P.tatoms, a), bind(P.wsort, a), bind(P.bsort, a);
    bind(P.tstart, a), bind(P.tlocal, a), bind(P.fpme, a), bind(P.tw, a), bind(P.epart, a);
    bind(c.bonded.slot, a), bind(c.bonded.epart, a), bind(c.partials, a);
    for (int d = 0; d < 4; ++d) bind(c.scan.tsum[d], a), bind(c.scan.toff[d], a);
    P.green_ok = false, P.counts_clear = false;
  }
  ~ScratchBinding() {
    for (auto it = undo_.rbegin(); it != undo_.rend(); ++it) (*it)();
    c_.pme.green_ok = false, c_.pme.counts_clear = false;
  }
  ScratchBinding(const ScratchBinding&) = delete;
  ScratchBinding& operator=(const ScratchBinding&) = delete;

 private:
  template <class T>
  void bind(DevBuf<T>& d, Arena& a) {
    auto owned = std::make_shared<DevBuf<T>>();
    owned->swap(d);
    d.arena = &a;
    d.resize(owned->n);
    undo_.push_back([&d, owned] { d.swap(*owned); });
  }
  capsgpu_ctx& c_;
  std::vector<std::function<void()>> undo_;
};

}  // namespace

capsgpu_ctx::~capsgpu_ctx() {
  if (h_flags) (void)hipHostFree(h_flags);
  if (h_terms) (void)hipHostFree(h_terms);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& set : pev)
    for (auto& e : set) (void)hipEventDestroy(e);
  for (auto& k : kev) (void)hipEventDestroy(k.a), (void)hipEventDestroy(k.b);
  learned_free(learned);
  if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

const char* capsgpu_last_error(void) { return g_error.c_str(); }

int32_t capsgpu_create(const capsgpu_ff* ff, const capsgpu_energy_opts* o, capsgpu_ctx** out) {
  return guard([&] {
    if (!ff || !o || !out) throw Error(CAPSGPU_EINVAL, "null argument");
    *out = nullptr;
    const int n = ff->n_atoms, nt = ff->n_types;
    if (n < 1 || nt < 1) throw Error(CAPSGPU_EINVAL, "the force field needs atoms and types");
    if (!ff->type || !ff->charge || !ff->mass || !ff->eps || !ff->sigma || !ff->excl_offset)
      throw Error(CAPSGPU_EINVAL, "the force field lacks types, charges, masses, LJ tables or exclusions");
    if (o->cutoff <= 0 || o->skin < 0) throw Error(CAPSGPU_EINVAL, "the cut-off must be positive and the skin not negative");
    if (o->parts < 1 || o->parts > 7) throw Error(CAPSGPU_EINVAL, "parts must be 1 … 7 (1 bonded, 2 non-bonded, 4 reciprocal space)");
    if ((o->parts & 4) && !(o->coulomb && o->electrostatics == 1 && !o->no_reciprocal))
      throw Error(CAPSGPU_EINVAL, "the reciprocal-space part needs PME electrostatics");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error(CAPSGPU_EDEVICE, "no GPU visible to HIP");
    auto c = std::make_unique<capsgpu_ctx>();
    c->n = n, c->nt = nt, c->opt = *o;
    HIPCHECK(hipGetDevice(&c->device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_flags), 4 * sizeof(uint32_t), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_terms), 64 * sizeof(double), hipHostMallocDefault));
    std::memset(c->h_flags, 0, 4 * sizeof(uint32_t));
    c->type.assign(ff->type, ff->type + n);
    for (int32_t t : c->type)
      if (t < 0 || t >= nt) throw Error(CAPSGPU_EINVAL, "an atom type is out of range");
    c->charge.assign(ff->charge, ff->charge + n);
    c->mass.assign(ff->mass, ff->mass + n);
    for (double m : c->mass)
      if (!(m > 0)) throw Error(CAPSGPU_EINVAL, "every atom needs a positive mass");
    c->lj96 = ff->lj96 != 0;
    c->eps.resize(size_t(nt) * nt);
    c->sig6.resize(size_t(nt) * nt);
    std::vector<double> c12(size_t(nt) * nt), c6(size_t(nt) * nt);
    for (size_t tp = 0; tp < c12.size(); ++tp) {
      c->eps[tp] = ff->eps[tp];
      c->sig6[tp] = std::pow(ff->sigma[tp], 6);
      // Evaluator::cluster_coefficients (field.cpp:853-890): 12-6 4εσ¹², 4εσ⁶; 9-6 2εσ⁹, 3εσ⁶; factors when products
      const double eps = c->eps[tp] > 0 ? c->eps[tp] : 0.0, s6 = c->sig6[tp];
      c12[tp] = c->lj96 ? 2 * eps * s6 * std::sqrt(s6) : 4 * eps * s6 * s6;
      c6[tp] = c->lj96 ? 3 * eps * s6 : 4 * eps * s6;
    }
    c->f12.assign(size_t(nt), 0.0);
    c->f6.assign(size_t(nt), 0.0);
    bool comb = true;
    for (int t = 0; t < nt; ++t) c->f12[size_t(t)] = std::sqrt(c12[size_t(t) * nt + t]), c->f6[size_t(t)] = std::sqrt(c6[size_t(t) * nt + t]);
    for (int a = 0; a < nt && comb; ++a)
      for (int b = 0; b < nt && comb; ++b) {
        const size_t tp = size_t(a) * nt + b;
        const double p12 = c->f12[size_t(a)] * c->f12[size_t(b)], p6 = c->f6[size_t(a)] * c->f6[size_t(b)];
        comb = std::fabs(p12 - c12[tp]) <= 1e-7 * std::max(std::fabs(c12[tp]), 1e-300) &&
               std::fabs(p6 - c6[tp]) <= 1e-7 * std::max(std::fabs(c6[tp]), 1e-300);
      }
    c->comb = comb;
    c->excl_off.assign(ff->excl_offset, ff->excl_offset + n + 1);
    c->excl_idx.assign(ff->excl_index, ff->excl_index + c->excl_off[size_t(n)]);
    for (int i = 0; i < n; ++i)
      for (int32_t e = c->excl_off[size_t(i)]; e < c->excl_off[size_t(i) + 1]; ++e) {
        if (c->excl_idx[size_t(e)] < 0 || c->excl_idx[size_t(e)] >= n) throw Error(CAPSGPU_EINVAL, "an excluded partner is out of range");
        if (e > c->excl_off[size_t(i)] && c->excl_idx[size_t(e)] <= c->excl_idx[size_t(e) - 1])
          throw Error(CAPSGPU_EINVAL, "the excluded partners of each atom must be sorted and distinct");
      }
    if (ff->p14_offset) {
      c->p14_off.assign(ff->p14_offset, ff->p14_offset + n + 1);
      c->p14_idx.assign(ff->p14_index, ff->p14_index + c->p14_off[size_t(n)]);
      c->f14 = ff->coul14;
    }
    c->type_count.assign(size_t(nt), 0.0);
    for (int32_t t : c->type) c->type_count[size_t(t)] += 1;
    hipStream_t s = c->stream;
    c->d_type.upload(c->type, s);
    c->d_charge.upload(c->charge, s);
    std::vector<double> minv(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) minv[size_t(i)] = 1.0 / c->mass[size_t(i)];
    c->d_minv.upload(minv, s);
    std::vector<float> f12f(c->f12.begin(), c->f12.end()), f6f(c->f6.begin(), c->f6.end());
    c->d_f12.upload(f12f, s);
    c->d_f6.upload(f6f, s);
    c->d_excl_off.upload(c->excl_off, s);
    if (c->excl_idx.empty()) c->d_excl_idx.resize(1);
    else c->d_excl_idx.upload(c->excl_idx, s);
    if (!c->p14_off.empty()) {
      c->d_p14_off.upload(c->p14_off, s);
      if (c->p14_idx.empty()) c->d_p14_idx.resize(1);
      else c->d_p14_idx.upload(c->p14_idx, s);
    }
    bonded_setup(*c, *ff);
    c->x.resize(3 * size_t(n));
    c->v.resize(3 * size_t(n));
    c->f.resize(3 * size_t(n));
    c->x0.resize(3 * size_t(n));
    c->x0p.resize(3 * size_t(n));
    c->v.zero(s);
    c->flags.resize(4);
    c->flags.zero(s);
    c->terms.resize(64);
    if (ewald_on(*c)) {
      if (o->pme_order < 3 || o->pme_order > 8) throw Error(CAPSGPU_EINVAL, "PME order must be 3 … 8");
      c->pme.beta = ewald_beta(o->cutoff, o->ewald_rtol);
    }
    HIPCHECK(hipStreamSynchronize(s));
    *out = c.release();
  });
}

int32_t capsgpu_destroy(capsgpu_ctx* ctx) {
  return guard([&] { delete ctx; });
}

int32_t capsgpu_set_state(capsgpu_ctx* c, const double* x, const double* v, const double cell[12]) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null context");
    if (cell) upload_cell(*c, cell);
    if (x) {
      c->x.upload(x, 3 * size_t(c->n), c->stream);
      c->state_set = true;
    }
    if (v) c->v.upload(v, 3 * size_t(c->n), c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t capsgpu_get_state(capsgpu_ctx* c, double* x, double* v, double* cell) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null context");
    if (x) c->x.download(x, 3 * size_t(c->n), c->stream);
    if (v) c->v.download(v, 3 * size_t(c->n), c->stream);
    if (cell) std::memcpy(cell, c->cell, 12 * sizeof(double));
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t capsgpu_get_forces(capsgpu_ctx* c, double* f) {
  return guard([&] {
    if (!c || !f) throw Error(CAPSGPU_EINVAL, "null argument");
    c->f.download(f, 3 * size_t(c->n), c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t capsgpu_compute(capsgpu_ctx* c, const double* x, const double cell[12], double* f, double* terms, int32_t want) {
  return guard([&] {
    if (!c || !x || !cell) throw Error(CAPSGPU_EINVAL, "null argument");
    compute_body(c, x, cell, f, terms, want);
  });
}

int32_t capsgpu_compute_noalloc(capsgpu_ctx* c, const double* x, const double cell[12], double* f, double* terms, int32_t want) {
  return guard([&] {
    if (!c || !x || !cell) throw Error(CAPSGPU_EINVAL, "null argument");
    if (c->opt.parts & 2)
      throw Error(CAPSGPU_EINVAL, "the allocation-free evaluation takes the bonded terms and PME's reciprocal part (parts 1, 4 or 5)");
    if (!(cell_volume(cell) > 1e-9)) throw Error(CAPSGPU_EINVAL, "the cell must have a volume (periodic along a, b and c)");
    // the scratch this evaluation needs at most, checked before any work: mesh, spectrum and influence function, the
    // per-atom spline, tile and force arrays, the bonded terms' member slots, the reductions
    const size_t n = size_t(c->n);
    size_t need = 3 * size_t(c->bonded.nb * 2 + c->bonded.na * 3 + c->bonded.nf * 2 + 1) * sizeof(double) + 64 * 1024;
    if (pme_on(*c)) {
      int k[3];
      pme_grid_dims(*c, cell, k);
      const size_t k3 = size_t(k[0]) * k[1] * k[2], nh = size_t(half_stride(k[0])) * k[1] * k[2];
      const size_t tiles = size_t((k[0] + 3) / 4) * ((k[1] + 3) / 4) * ((k[2] + 3) / 4);
      need += 4 * k3 + 12 * nh + 16 * size_t(k[0] + k[1] + k[2]) + (20 + 4 * size_t(tile_capacity(int64_t(n), int64_t(tiles)))) * tiles +
              60 * std::max({k[0], k[1], k[2]}) + 88 * n;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    unsigned char* base = scratch_on(c->device);
    if (need > kScratchBytes)
      throw Error(CAPSGPU_EINVAL, "the allocation-free evaluation holds " + std::to_string(kScratchBytes >> 20) + " MiB of device scratch; this "
                                  "system needs about " + std::to_string((need + (1u << 20) - 1) >> 20) + " MiB: use capsgpu_compute");
    Arena arena{base, kScratchBytes, 0};
    {
      ScratchBinding bind(*c, arena);
      compute_body(c, x, cell, f, terms, want);
    }
  });
}


int32_t capsgpu_list_builds(capsgpu_ctx* c, int32_t* builds) {
  return guard([&] {
    if (!c || !builds) throw Error(CAPSGPU_EINVAL, "null argument");
    *builds = c->builds;
  });
}

int32_t capsgpu_stage_times(capsgpu_ctx* c, double out[6]) {
  return guard([&] {
    if (!c || !out) throw Error(CAPSGPU_EINVAL, "null argument");
    for (int k = 0; k < 6; ++k) out[k] = c->stage_ms[k];
  });
}

int32_t capsgpu_profile(capsgpu_ctx* c, int32_t on) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null context");
    prof_harvest(*c);
    c->profile = on != 0;
    for (int k = 0; k < 6; ++k) c->prof_ms[k] = 0, c->prof_n[k] = 0;
    c->ktot.clear();
  });
}

int32_t capsgpu_profile_only(capsgpu_ctx* c, const char* kernel) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null context");
    prof_harvest(*c);
    c->kfilter = kernel ? kernel : "";
  });
}

int32_t capsgpu_profile_read(capsgpu_ctx* c, double ms[6], int64_t counts[6]) {
  return guard([&] {
    if (!c || !ms || !counts) throw Error(CAPSGPU_EINVAL, "null argument");
    prof_harvest(*c);
    for (int k = 0; k < 6; ++k) ms[k] = c->prof_ms[k], counts[k] = c->prof_n[k];
  });
}

int32_t capsgpu_profile_kernels(capsgpu_ctx* c, char* buf, int32_t cap) {
  return guard([&] {
    if (!c || !buf || cap < 1) throw Error(CAPSGPU_EINVAL, "null argument");
    prof_harvest(*c);
    std::string out;
    char line[160];
    for (const auto& k : c->ktot) {
      std::snprintf(line, sizeof line, "%s %.6f %lld\n", k.first.c_str(), k.second.first, static_cast<long long>(k.second.second));
      out += line;
    }
    std::snprintf(buf, size_t(cap), "%s", out.c_str());
  });
}

int32_t capsgpu_learned_forward(capsgpu_ctx* c, const float* params, const float* x, int32_t m, float* y) {
  return guard([&] {
    if (!c || !params || !x || !y || m < 1) throw Error(CAPSGPU_EINVAL, "null argument");
    learned_forward_host(*c, params, x, m, y);
  });
}

int32_t capsgpu_learned_gradient(capsgpu_ctx* c, const float* params, const float* x, const float* t, int32_t m, float* grad,
                                 double* loss) {
  return guard([&] {
    if (!c || !params || !x || !t || !grad || m < 1) throw Error(CAPSGPU_EINVAL, "null argument");
    const double l = learned_gradient_host(*c, params, x, t, m, grad);
    if (loss) *loss = l;
  });
}

int32_t capsgpu_learned_descriptors(capsgpu_ctx* c, float* raw) {
  return guard([&] {
    if (!c || !raw) throw Error(CAPSGPU_EINVAL, "null argument");
    if (!c->state_set) throw Error(CAPSGPU_EINVAL, "set the positions first (capsgpu_set_state)");
    learned_descriptors_host(*c, raw);
  });
}

int32_t capsgpu_list_sizes(capsgpu_ctx* c, int64_t out[8]) {
  return guard([&] {
    if (!c || !out) throw Error(CAPSGPU_EINVAL, "null argument");
    const ListGeom& g = c->L.g;
    out[0] = c->n, out[1] = g.nclusters, out[2] = g.ns, out[3] = g.nentries, out[4] = g.ncodes, out[5] = g.nu, out[6] = g.nv;
    out[7] = c->L.npruned;
  });
}

int32_t capsgpu_list_get(capsgpu_ctx* c, uint32_t* perm, uint32_t* iperm, int64_t* first, uint32_t* cj, uint16_t* code, uint16_t* mask,
                         uint16_t* ex, uint16_t* ex14, double* org, double* off, int32_t* ioff, int32_t* qint, double* qvec) {
  return guard([&] {
    if (!c || !c->L.built) throw Error(CAPSGPU_EINVAL, "no list has been built");
    DevList& L = c->L;
    hipStream_t s = c->stream;
    const size_t ns = size_t(L.g.ns), ne = size_t(L.g.nentries), ncl = size_t(L.g.nclusters), nc = size_t(L.g.ncodes);
    if (perm) L.perm.download(perm, ns, s);
    if (iperm) L.iperm.download(iperm, size_t(c->n), s);
    if (first) L.first.download(first, ncl + 1, s);
    if (cj) L.e_cj.download(cj, ne, s);
    if (code) L.e_code.download(code, ne, s);
    if (mask) L.e_mask.download(mask, ne, s);
    if (ex) L.e_ex.download(ex, ne, s);
    if (ex14) L.e_ex14.download(ex14, ne, s);
    if (org) L.org.download(org, 3 * ncl, s);
    if (off) L.off.download(off, 3 * ns, s);
    if (ioff) L.ioff.download(ioff, 3 * ns, s);
    if (qint) L.qint.download(qint, 3 * nc, s);
    if (qvec) L.qvec.download(qvec, 3 * nc, s);
    HIPCHECK(hipStreamSynchronize(s));
  });
}

int32_t capsgpu_list_pruned(capsgpu_ctx* c, float* X, float* Y, float* Z, uint16_t* pmask, uint16_t* pex, uint16_t* pex14) {
  return guard([&] {
    if (!c || !c->L.built) throw Error(CAPSGPU_EINVAL, "no list has been built");
    DevList& L = c->L;
    hipStream_t s = c->stream;
    const size_t ns = size_t(L.g.ns), ne = size_t(L.g.nentries);
    if (X || Y || Z) {
      std::vector<float4> xq = L.xq.to_host(s);
      for (size_t k = 0; k < ns; ++k) {
        if (X) X[k] = xq[k].x;
        if (Y) Y[k] = xq[k].y;
        if (Z) Z[k] = xq[k].z;
      }
    }
    if (pmask) L.pmask.download(pmask, ne, s);
    if (pex) L.pex.download(pex, ne, s);
    if (pex14) L.pex14.download(pex14, ne, s);
    HIPCHECK(hipStreamSynchronize(s));
  });
}

int32_t capsgpu_run_dynamics(capsgpu_ctx* c, const capsgpu_dyn_opts* o, capsgpu_progress_fn progress, void* user, capsgpu_thermo* rows,
                             int32_t cap, int32_t* nrows, capsgpu_dyn_report* rep) {
  return guard([&] {
    if (!c || !o) throw Error(CAPSGPU_EINVAL, "null argument");
    if (!c->state_set) throw Error(CAPSGPU_EINVAL, "set the positions first (capsgpu_set_state)");
    if (o->dt <= 0 || o->dt > 5) throw Error(CAPSGPU_EINVAL, "time step must be in (0, 5] fs");
    if (o->steps < 0) throw Error(CAPSGPU_EINVAL, "the number of steps must be 0 or more");
    if (o->respa < 1 || o->respa > 16) throw Error(CAPSGPU_EINVAL, "r-RESPA inner steps must be 1 … 16");
    if (o->respa > 1 && o->thermostat == 2) throw Error(CAPSGPU_EINVAL, "r-RESPA runs with the Bussi thermostat or none");
    if (o->thermostat < 0 || o->thermostat > 2) throw Error(CAPSGPU_EINVAL, "thermostat must be 0 (none), 1 (Bussi) or 2 (Langevin)");
    if (o->learned_k > 1 && (o->respa > 1)) throw Error(CAPSGPU_EINVAL, "learned r-RESPA replaces the bonded/non-bonded split");
    if (o->learned_k > 1 && !pme_on(*c)) throw Error(CAPSGPU_EINVAL, "learned r-RESPA needs PME");
    const int baro = o->barostat;
    if (baro < 0 || baro > 2) throw Error(CAPSGPU_EINVAL, "barostat must be 0 (none), 1 (Berendsen) or 2 (C-rescale)");
    if (baro && o->learned_k > 1) throw Error(CAPSGPU_EINVAL, "learned r-RESPA runs at a fixed cell: no barostat");
    if (baro && !(o->tau_p > 0 && o->compressibility > 0)) throw Error(CAPSGPU_EINVAL, "the barostat needs a positive tau_p and compressibility");
    const int bevery = o->barostat_every > 0 ? o->barostat_every : 10;
    if (nrows) *nrows = 0;
    const int n = c->n;
    hipStream_t s = c->stream;
    const double ndof = 3.0 * n - 3.0;
    if (ndof < 1) throw Error(CAPSGPU_EINVAL, "dynamics needs at least two atoms");
    HostRng rng(o->seed);
    if (o->new_velocities) {
      // Maxwell–Boltzmann from the run's own mt19937_64 and normal generator (dynamics.cpp:236-256)
      std::vector<double> v(3 * size_t(n));
      initial_velocities(rng, c->mass, o->temperature, v);
      c->v.upload(v, s);
    }
    const double dt = o->dt, hk = 0.5 * dt * kAcc, bskin = std::min(0.8, c->opt.skin);
    // move limits for the pair list's rebuild and prune checks; with no pair terms there is no list (negative: only the
    // positions' finiteness is checked)
    const bool listed = (c->opt.parts & 2) != 0;
    const double lim_b = listed ? 0.25 * c->opt.skin * c->opt.skin : -1.0, lim_p = listed ? 0.25 * bskin * bskin : -1.0;
    const double c1 = std::exp(-dt / std::max(o->tau_t, 1e-9)), c2 = std::sqrt(1 - c1 * c1);
    double kt_target = 0.5 * ndof * kBoltz * o->temperature;
    // a ramped thermostat target (DynamicsOptions::temperature_end, dynamics.cpp:342-345, 608-612): T(step) from the
    // starting temperature to temperature_end over the run; Langevin's noise scales with √(T(step)/T(0))
    const bool ramp = o->temperature_end > 0 && o->steps > 0;
    if (ramp && !(o->temperature > 0)) throw Error(CAPSGPU_EINVAL, "a temperature ramp needs a positive starting temperature");
    double t_now = o->temperature, c2_now = c2;
    auto target_t = [&](int64_t step) {
      return ramp ? o->temperature + (o->temperature_end - o->temperature) * double(step) / double(o->steps) : o->temperature;
    };
    DevBuf<double> sdv, kpart, bpart, scal;   // scal: [0] α, [1] bath
    if (o->thermostat == 2) {
      std::vector<double> sd(static_cast<size_t>(n));
      for (int i = 0; i < n; ++i) sd[size_t(i)] = std::sqrt(kBoltz * o->temperature * kAcc / c->mass[size_t(i)]);
      sdv.upload(sd, s);
    }
    const unsigned nb = blocks_for(n, 256);
    kpart.resize(7 * size_t(nb));
    bpart.resize(size_t(nb));
    bpart.zero(s);
    scal.resize(2);
    scal.zero(s);
    std::vector<double> one = {1.0, 0.0};
    scal.upload(one, s);
    const int every = std::max(1, o->thermo_every);
    double mtot = 0;
    for (double m : c->mass) mtot += m;
    double baro_bath = 0;   // the energy the stochastic cell rescaling took out of the kinetic energy (the conserved sum)
    int32_t nr = 0;
    int64_t learned_steps = 0, full_steps = 0;
    double last_rmse = 0;
    bool last_learned = false;
    Learned* LR = nullptr;
    if (o->learned_k > 1) LR = learned_begin(*c, *o);
    auto emit = [&](int64_t step, const double* t) {
      // the kinetic energy and its tensor after the closing kick (and Bussi's scaling)
      colsum(*c, kpart.p, nb, 7, c->terms.p + 40);
      HIPCHECK(hipMemcpyAsync(c->h_terms + 40, c->terms.p + 40, 7 * sizeof(double), hipMemcpyDeviceToHost, s));
      double hb[2];
      HIPCHECK(hipMemcpyAsync(hb, scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      capsgpu_thermo r{};
      r.step = step + o->step_offset;
      r.time_ps = double(step + o->step_offset) * dt * 1e-3;
      r.kinetic = c->h_terms[40];
      r.temperature = 2 * r.kinetic / (ndof * kBoltz);
      for (int k = 0; k < CAPSGPU_NTERMS; ++k) r.terms[k] = t[k];
      r.potential = t[CAPSGPU_T_BOND] + t[CAPSGPU_T_ANGLE] + t[CAPSGPU_T_DIHEDRAL] + t[CAPSGPU_T_IMPROPER] + t[CAPSGPU_T_VDW] + t[CAPSGPU_T_COULOMB];
      r.total = r.potential + r.kinetic;
      const double V = cell_volume(c->cell);
      r.pressure = (2 * r.kinetic + t[CAPSGPU_T_VIRIAL]) / (3 * V) * kAtm;
      r.conserved = r.total + hb[1] + baro_bath;
      r.volume = V;
      r.density = mtot / 6.02214076e23 / (V * 1e-24);
      for (int q = 0; q < 6; ++q) r.p[q] = (c->h_terms[41 + q] / kAcc + t[CAPSGPU_T_WXX + q]) / V * kAtm;
      r.learned = last_learned ? 1 : 0;
      r.learned_rmse = last_rmse;
      r.learned_rmse_stale = LR ? learned_rmse_stale(*LR) : 0.0;
      if (rows && nr < cap) rows[nr] = r;
      ++nr;
      if (progress && !progress(&r, user)) throw Error(CAPSGPU_ECANCEL, "dynamics cancelled");
    };
    // r-RESPA (dynamics.cpp:113-123, 643-656): the slow terms (all but the bonded ones) kick over the whole step, the
    // bonded ones over `respa` inner velocity-Verlet steps
    const int R = o->respa, slow_parts = c->opt.parts & ~1;
    DevBuf<double> fslow, ffast;
    if (R > 1) {
      if (!(c->opt.parts & 1) || !slow_parts) throw Error(CAPSGPU_EINVAL, "r-RESPA splits the bonded terms from the others: parts must hold both");
      fslow.resize(3 * size_t(n));
      ffast.resize(3 * size_t(n));
    }
    auto full_eval = [&](bool want, double* tt, bool flags_known) {
      if (R > 1) {
        double ts[CAPSGPU_NTERMS], tf[CAPSGPU_NTERMS];
        evaluate(*c, want, ts, flags_known, slow_parts, fslow.p);
        evaluate(*c, want, tf, true, 1, ffast.p);
        if (want)
          for (int k = 0; k < CAPSGPU_NTERMS; ++k) tt[k] = ts[k] + tf[k];
      } else {
        evaluate(*c, want, tt, flags_known);
      }
    };
    // without a pair list the flags record only non-finite positions and FENE bonds past R0: they gather over the run and
    // are read at thermo steps and at the end, so the steps in between need no host synchronisation
    auto check_flags = [&](int64_t step) {
      read_flags(*c);
      if (c->h_flags[3]) throw Error(CAPSGPU_EFIELD, "the simulation became unstable by step " + std::to_string(step));
      if (c->h_flags[2]) throw Error(CAPSGPU_EFIELD, "a FENE bond is stretched to R0 or past; lower the time step or relax the structure first");
    };
    double t[CAPSGPU_NTERMS];
    // the barostat after a step (dynamics.cpp:701-723): the pressure from the kinetic energy after the step and the
    // virial of the step's forces; the cell and the positions (and, for stochastic cell rescaling, the velocities)
    // scaled about the cell origin; then the forces at the new positions and cell
    auto barostat_update = [&](bool thermo) {
      kick(*c, 0.0, true, kpart.p);
      colsum(*c, kpart.p, nb, 7, c->terms.p + 48);
      double kin = 0;
      HIPCHECK(hipMemcpyAsync(&kin, c->terms.p + 48, sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      const double V = cell_volume(c->cell), hp = dt * bevery;
      const double p = (2 * kin + t[CAPSGPU_T_VIRIAL]) / (3 * V) * kAtm;
      double deps = -o->compressibility / o->tau_p * (o->pressure - p) * hp;
      if (baro == 2) {
        const double ktv = kBoltz * t_now / V * kAtm;
        deps += o->compressibility / o->tau_p * ktv * hp + std::sqrt(2 * ktv * o->compressibility * hp / o->tau_p) * rng.normal();
      }
      deps = std::clamp(deps, -0.01, 0.01);
      const double mu = std::exp(deps / 3);
      scale_xv(*c, c->cell, mu, baro == 2 ? 1.0 / mu : 1.0);
      if (baro == 2) baro_bath -= kin / (mu * mu) - kin;
      for (int k = 3; k < 12; ++k) c->cell[k] *= mu;
      c->L.built = false;
      full_eval(thermo, t, false);
      if (thermo) kick(*c, 0.0, true, kpart.p);
    };
    c->h_flags[0] = c->h_flags[1] = 0;
    if (!listed) c->flags.zero(s);
    full_eval(true, t, false);
    // step 0: the kinetic energy of the starting velocities
    kick(*c, 0.0, true, kpart.p);
    emit(0, t);
    HIPCHECK(hipStreamSynchronize(s));
    // checkpoints of a whole state only: every coordinate and velocity finite (dynamics.cpp:758-763)
    std::vector<double> cpx, cpv;
    auto checkpoint_due = [&](int64_t step) { return o->checkpoint && o->checkpoint_every > 0 && step % o->checkpoint_every == 0 && step < o->steps; };
    auto checkpoint = [&](int64_t step) {
      cpx.resize(3 * size_t(n)), cpv.resize(3 * size_t(n));
      c->x.download(cpx.data(), cpx.size(), s);
      c->v.download(cpv.data(), cpv.size(), s);
      HIPCHECK(hipStreamSynchronize(s));
      for (size_t k = 0; k < cpx.size(); ++k)
        if (!std::isfinite(cpx[k]) || !std::isfinite(cpv[k])) return;
      if (o->checkpoint(cpx.data(), cpv.data(), c->cell, step + o->step_offset, o->checkpoint_user) != 0)
        throw Error(CAPSGPU_EINVAL, "the checkpoint callback failed at step " + std::to_string(step + o->step_offset));
    };
    const auto t0 = std::chrono::steady_clock::now();
    bool alpha_pending = false;
    // a step's closing half kicks are put off to the next step's first launch when nothing reads the velocities in
    // between (no thermo row, barostat, checkpoint or Bussi scaling, not the last step): one read and write of v less
    Kicks pend;
    auto quiet = [&](int64_t step, bool thermo, bool bstep) {
      return o->thermostat != 1 && !thermo && !bstep && !checkpoint_due(step) && step < o->steps;
    };
    // the plain evaluation of a step runs PME's reciprocal part (evaluate's rule for the options' parts)
    const bool recip_in_step = pme_on(*c) && (c->opt.parts & 6) && !(c->opt.parts & 8);
    // a failure inside the loop leaves no half-done PME bookkeeping behind for the context's next evaluation
    struct DropPresplines {
      capsgpu_ctx* c;
      ~DropPresplines() {
        if (c->pme.presplined) pme_drop_presplines(*c);
      }
    } drop_presplines{c};
    for (int64_t step = 1; step <= o->steps; ++step) {
      const bool thermo = step % every == 0 || step == o->steps;
      if (ramp) {
        t_now = target_t(step);
        kt_target = 0.5 * ndof * kBoltz * t_now;
        c2_now = c2 * std::sqrt(t_now / o->temperature);
      }
      const bool bstep = baro && step % bevery == 0;
      if (listed) c->flags.zero(s);
      if (R > 1) {
        const double h = dt / R, hkh = 0.5 * h * kAcc;
        pend.put(hk, fslow.p);
        for (int j = 0; j < R; ++j) {
          vv_open(*c, hkh, h, nullptr, lim_b, lim_p, ffast.p, pend);
          pend = Kicks{};
          if (j + 1 < R) {
            evaluate(*c, false, t, true, 1, ffast.p);
          } else {
            if (listed) {
              read_flags(*c);
              if (c->h_flags[3]) throw Error(CAPSGPU_EFIELD, "the simulation became unstable at step " + std::to_string(step));
            }
            full_eval(thermo || bstep, t, true);
          }
          pend.put(hkh, ffast.p);
        }
        if (quiet(step, thermo, bstep)) {
          pend.put(hk, fslow.p);
        } else {
          kick(*c, hk, thermo || o->thermostat == 1, kpart.p, fslow.p, pend);
          pend = Kicks{};
        }
        if (o->thermostat == 1) {
          const double r1 = rng.normal();
          const double sum = ndof > 1 ? rng.gamma((ndof - 1) / 2, 2.0) : 0.0;
          bussi(*c, kpart.p, kt_target, ndof, std::exp(-dt / o->tau_t), r1, sum, scal.p, scal.p + 1);
          scale_v(*c, scal.p);
          if (thermo) kick(*c, 0.0, true, kpart.p);
        }
        if (bstep) barostat_update(thermo);
        if (thermo) {
          if (!listed) check_flags(step);
          emit(step, t);
        }
        if (checkpoint_due(step)) checkpoint(step);
        continue;
      }
      // the opening launch also does the per-atom half of the evaluation's PME bookkeeping on the positions it writes
      // (the evaluation below is the plain one with PME's reciprocal part)
      AtomSplines sp{};
      const AtomSplines* spp = !LR && recip_in_step && pme_presplines(*c, sp) ? &sp : nullptr;
      if (o->thermostat == 2) {
        baoab_open(*c, hk, 0.5 * dt, c1, c2_now, sdv.p, o->seed, nullptr, uint64_t(step + o->step_offset), lim_b, lim_p, bpart.p, pend, spp);
      } else {
        vv_open(*c, hk, dt, alpha_pending ? scal.p : nullptr, lim_b, lim_p, nullptr, pend, spp);
        alpha_pending = false;
      }
      pend = Kicks{};
      if (listed) {
        read_flags(*c);
        if (c->h_flags[3]) throw Error(CAPSGPU_EFIELD, "the simulation became unstable at step " + std::to_string(step));
      }
      if (LR) {
        const int r = learned_step(*c, *LR, step, thermo, t);
        last_learned = r == 1;
        if (r == 1) ++learned_steps;
        else ++full_steps;
        last_rmse = learned_rmse(*LR);
      } else {
        evaluate(*c, thermo || bstep, t, true);
      }
      if (!LR && quiet(step, thermo, bstep)) pend.put(hk, c->f.p);
      else kick(*c, hk, thermo || o->thermostat == 1, kpart.p);
      if (o->thermostat == 1) {
        const double r1 = rng.normal();
        const double sum = ndof > 1 ? rng.gamma((ndof - 1) / 2, 2.0) : 0.0;
        bussi(*c, kpart.p, kt_target, ndof, std::exp(-dt / o->tau_t), r1, sum, scal.p, scal.p + 1);
        if (thermo || bstep) scale_v(*c, scal.p);
        else alpha_pending = true;
      }
      if (bstep) barostat_update(thermo);
      if (thermo) {
        if (o->thermostat == 1 && !bstep) kick(*c, 0.0, true, kpart.p);
        if (o->thermostat == 2) bath_add(*c, bpart.p, scal.p + 1);
        if (!listed) check_flags(step);
        emit(step, t);
      }
      if (checkpoint_due(step)) {
        if (alpha_pending) scale_v(*c, scal.p), alpha_pending = false;   // the velocities of the step, whole
        checkpoint(step);
      }
    }
    if (alpha_pending) scale_v(*c, scal.p);
    read_flags(*c);
    if (c->h_flags[2]) throw Error(CAPSGPU_EFIELD, "a FENE bond is stretched to R0 or past; lower the time step or relax the structure first");
    HIPCHECK(hipStreamSynchronize(s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (LR) learned_end(*c, LR);
    if (nrows) *nrows = nr;
    if (rep) {
      rep->steps = o->steps;
      rep->seconds = secs;
      rep->ns_per_day = secs > 0 ? double(o->steps) * dt * 1e-6 / secs * 86400 : 0;
      rep->list_builds = c->builds;
      rep->prunes = c->prunes;
      rep->learned_steps = int32_t(learned_steps);
      rep->full_steps = int32_t(full_steps);
      rep->fallbacks = LR ? int32_t(learned_failures(*LR)) : 0;
      rep->judgments = LR ? int32_t(learned_judgments(*LR)) : 0;
    }
  });
}

int32_t capsgpu_tune(capsgpu_ctx* c, int32_t tiles[2]) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null argument");
    if (!pme_on(*c)) throw Error(CAPSGPU_EINVAL, "the tuned tile shapes are the PME transforms': the context has no PME");
    if (!c->state_set) throw Error(CAPSGPU_EINVAL, "set the positions first (capsgpu_set_state)");
    pme_forces(*c, false);   // the arrays the gather's candidates are timed on, for the current cell and grid
    int t[2];
    pme_tune(*c, t);
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (tiles) tiles[0] = t[0], tiles[1] = t[1];
  });
}

int32_t capsgpu_set_tiles(capsgpu_ctx* c, int32_t lines_x, int32_t lines_yz, int32_t tiles[2]) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null argument");
    auto pow2 = [](int32_t v) { return v >= 1 && v <= 16 && (v & (v - 1)) == 0; };
    if ((lines_x != 0 && !pow2(lines_x)) || !pow2(lines_yz))
      throw Error(CAPSGPU_EINVAL, "tile shapes are powers of two from 1 to 16 lines (lines_x 0: the plane transform)");
    c->pme.lines_x = lines_x, c->pme.lines_yz = lines_yz;
    if (tiles) tiles[0] = lines_x, tiles[1] = lines_yz;
  });
}

int32_t capsgpu_set_gather(capsgpu_ctx* c, int32_t shape, int32_t* in_use) {
  return guard([&] {
    if (!c) throw Error(CAPSGPU_EINVAL, "null argument");
    if (shape != 0 && shape != 1) throw Error(CAPSGPU_EINVAL, "the gather's shape is 0 (a workgroup per tile) or 1 (per run of 64 sorted atoms)");
    c->pme.gather_shape = shape;
    if (in_use) *in_use = shape;
  });
}

int32_t capsgpu_get_gather(capsgpu_ctx* c, int32_t* shape) {
  return guard([&] {
    if (!c || !shape) throw Error(CAPSGPU_EINVAL, "null argument");
    *shape = c->pme.gather_shape;
  });
}

int32_t capsgpu_batch_create(capsgpu_ctx* const* replicas, int32_t n, capsgpu_batch** out) {
  return guard([&] {
    if (!out) throw Error(CAPSGPU_EINVAL, "null argument");
    *out = nullptr;
    *out = batch_create(replicas, n);
  });
}

int32_t capsgpu_batch_run(capsgpu_batch* b, const capsgpu_batch_opts* o, const double* temperatures, double* temperature_out,
                          capsgpu_batch_report* rep) {
  return guard([&] {
    if (!b || !o) throw Error(CAPSGPU_EINVAL, "null argument");
    batch_run(*b, *o, temperatures, temperature_out, rep);
  });
}

int32_t capsgpu_batch_destroy(capsgpu_batch* b) {
  return guard([&] { batch_destroy(b); });
}

}  // extern "C"
