// C entry points over the reference's own nbnxm and k-space code, compiled from /root/reference by oracle/Makefile
// into oracle/_ref/libcapsref.so — TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline; never the product).
//
// The reference objects linked: core/src/nbnxm.cpp, nbnxm_avx2.cpp, system.cpp, elements.cpp, power.cpp, kspace.cpp,
// fortran_runtime.cpp and core/fortran/caps_kspace.f90 (the recipe of the reference's tests/nbnxm_check.cpp:2-4 plus
// k-space). core/src/field.cpp is left out: its link closure pulls in most of the core (JSON, perception, I/O). This
// file therefore drives the cluster path the way Evaluator::compute_placed does (core/src/field.cpp:1286-1312,
// 1819-1851, 2285-2375, 2442-2467): build, positions, prune at r_c + 0.8, the kernel per chunk, the reduction in
// chunk order, pair and shift virial, then pme_reciprocal with the pool.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <random>

#include "caps/kspace.hpp"
#include "caps/rng.hpp"
#include "kspace_pool.hpp"
#include "nbnxm.hpp"
#include "parallel.hpp"

using namespace caps;

extern "C" void caps_f_ewald(int n, const double* x, const double* q, const double* cell, const int* kmax, double beta,
                             double kcoul, double* energy, double* f, double* vir);

namespace {
Cell make_cell(const double c[12]) {
  Cell cell;
  for (int k = 0; k < 3; ++k) cell.origin[k] = c[k], cell.a[k] = c[3 + k], cell.b[k] = c[6 + k], cell.c[k] = c[9 + k];
  return cell;
}
std::vector<std::vector<uint32_t>> csr(int n, const int32_t* off, const int32_t* idx) {
  std::vector<std::vector<uint32_t>> v(static_cast<size_t>(n));
  if (!off) return v;
  for (int i = 0; i < n; ++i)
    for (int32_t e = off[i]; e < off[i + 1]; ++e) v[size_t(i)].push_back(uint32_t(idx[e]));
  return v;
}
thread_local std::string g_err;

struct RefEval {
  int n = 0;
  Cell cell;
  std::unique_ptr<ThreadPool> pool;
  nbnxm::List L;
  nbnxm::Params p;
  std::vector<double> x, q;
  std::vector<double> f12, f6;
  std::vector<int> type;
  std::vector<std::vector<uint32_t>> excl, p14;
  double rlist = 0, rc = 0, prune_r = 0;
  int pme = 0, order = 5, grid[3] = {0, 0, 0};
  double beta = 0;
  std::vector<std::vector<float>> tfc;
  std::vector<double> P;
};
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

double ref_ewald_beta(double rc, double rtol) { return ewald_beta(rc, rtol); }
int ref_good_size(int n) { return fft_good_size(n); }
void ref_pme_grid(const double c[12], double beta, double spacing, int order, int k[3]) {
  const PmeGrid g = pme_grid(make_cell(c), beta, spacing, order);
  for (int d = 0; d < 3; ++d) k[d] = g.k[d];
}
void ref_fft3d(double* re, double* im, int k1, int k2, int k3, int sign) {
  std::vector<std::complex<double>> v(size_t(k1) * k2 * k3);
  for (size_t i = 0; i < v.size(); ++i) v[i] = {re[i], im[i]};
  fft3d(v, k1, k2, k3, sign);
  for (size_t i = 0; i < v.size(); ++i) re[i] = v[i].real(), im[i] = v[i].imag();
}
// the staged, threaded reciprocal part the evaluator calls (kspace.cpp:80-175)
double ref_pme(const double* x, const double* q, int n, const double c[12], const int k[3], int order, double beta, int threads,
               double* f, double* vir) {
  ThreadPool pool(threads);
  std::vector<double> xv(x, x + 3 * size_t(n)), qv(q, q + n), fv(3 * size_t(n), 0.0);
  PmeGrid g;
  g.k[0] = k[0], g.k[1] = k[1], g.k[2] = k[2], g.order = order, g.beta = beta;
  const double e = pme_reciprocal(xv, qv, make_cell(c), g, fv, vir, pool);
  for (size_t i = 0; i < fv.size(); ++i) f[i] += fv[i];
  return e;
}
double ref_ewald_recip(const double* x, const double* q, int n, const double c[12], const int kmax[3], double beta, double* f,
                       double* vir) {
  double cell[9];
  for (int i = 0; i < 9; ++i) cell[i] = c[3 + i];
  double e = 0;
  caps_f_ewald(n, x, q, cell, kmax, beta, kCoulombConstant, &e, f, vir);
  return e;
}

// ---- the list -------------------------------------------------------------------------------------------------------
void* ref_list_build(const double* x, int n, const double c[12], double rlist, const int32_t* excl_off, const int32_t* excl_idx,
                     const int32_t* p14_off, const int32_t* p14_idx, int keep_excluded, int threads) {
  try {
    auto* R = new RefEval;
    R->n = n;
    R->cell = make_cell(c);
    R->pool = std::make_unique<ThreadPool>(threads);
    R->x.assign(x, x + 3 * size_t(n));
    R->excl = csr(n, excl_off, excl_idx);
    R->p14 = csr(n, p14_off, p14_idx);
    R->type.assign(size_t(n), 0);
    R->q.assign(size_t(n), 0.0);
    const std::vector<double> none;
    R->L.build(R->x, R->cell, rlist, R->excl, R->type, 1, {}, R->q, none, none, *R->pool, nullptr, nullptr,
               keep_excluded ? &R->p14 : nullptr);
    R->rlist = rlist;
    return R;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void ref_list_free(void* h) { delete static_cast<RefEval*>(h); }
void ref_list_sizes(void* h, int64_t out[5]) {
  const auto& L = static_cast<RefEval*>(h)->L;
  out[0] = int64_t(L.iperm.size()), out[1] = int64_t(L.nclusters), out[2] = int64_t(L.perm.size()), out[3] = int64_t(L.entries.size()),
  out[4] = int64_t(L.qint.size() / 3);
}
void ref_list_get(void* h, uint32_t* perm, uint32_t* iperm, int64_t* first, uint32_t* cj, uint16_t* code, uint16_t* mask, uint16_t* ex,
                  uint16_t* ex14, double* org, double* off, int32_t* ioff, int32_t* qint, double* qvec) {
  const auto& L = static_cast<RefEval*>(h)->L;
  std::memcpy(perm, L.perm.data(), 4 * L.perm.size());
  std::memcpy(iperm, L.iperm.data(), 4 * L.iperm.size());
  for (size_t i = 0; i < L.first.size(); ++i) first[i] = int64_t(L.first[i]);
  for (size_t e = 0; e < L.entries.size(); ++e)
    cj[e] = L.entries[e].cj, code[e] = L.entries[e].code, mask[e] = L.entries[e].mask, ex[e] = L.entries[e].ex, ex14[e] = L.entries[e].ex14;
  std::memcpy(org, L.org.data(), 8 * L.org.size());
  std::memcpy(off, L.off.data(), 8 * L.off.size());
  std::memcpy(ioff, L.ioff.data(), 4 * L.ioff.size());
  std::memcpy(qint, L.qint.data(), 4 * L.qint.size());
  std::memcpy(qvec, L.qvec.data(), 8 * L.qvec.size());
}
// positions at x, then prune at radius; returns the pruned entries concatenated over the chunks (i-cluster order)
int64_t ref_list_prune(void* h, const double* x, double radius, float* X, float* Y, float* Z, uint32_t* cj, uint16_t* code,
                       uint16_t* mask, uint16_t* ex, uint16_t* ex14, int64_t cap) {
  auto* R = static_cast<RefEval*>(h);
  std::vector<double> xv(x, x + 3 * size_t(R->n));
  R->L.positions(xv, *R->pool);
  R->L.prune(radius, *R->pool);
  std::memcpy(X, R->L.X.data(), 4 * R->L.X.size());
  std::memcpy(Y, R->L.Y.data(), 4 * R->L.Y.size());
  std::memcpy(Z, R->L.Z.data(), 4 * R->L.Z.size());
  int64_t k = 0;
  for (const auto& pe : R->L.pe)
    for (const auto& e : pe) {
      if (k < cap) cj[k] = e.cj, code[k] = e.code, mask[k] = e.mask, ex[k] = e.ex, ex14[k] = e.ex14;
      ++k;
    }
  return k;
}

// ---- the cluster path of the evaluator (build, positions, prune, kernel, reduction, PME) -----------------------------
// lj: 1 12-6, 2 9-6; coul: 0 none, 1 PME (Ewald real space), 2 DSF; f12/f6 per type (geometric factors); shift 0/1
void* ref_eval_create(const double* x, int n, const double c[12], const int32_t* type, int nt, const double* f12, const double* f6,
                      const double* q, int lj, int coul, double rc, double skin, double kappa, double dsf_e0, double dsf_f0,
                      int shift, const int32_t* excl_off, const int32_t* excl_idx, const int32_t* p14_off, const int32_t* p14_idx,
                      double f14, int pme_order, const int grid[3], int threads) {
  try {
    auto* R = new RefEval;
    R->n = n;
    R->cell = make_cell(c);
    R->pool = std::make_unique<ThreadPool>(threads);
    R->x.assign(x, x + 3 * size_t(n));
    R->q.assign(q, q + n);
    R->type.assign(type, type + n);
    R->f12.assign(f12, f12 + nt);
    R->f6.assign(f6, f6 + nt);
    R->excl = csr(n, excl_off, excl_idx);
    R->p14 = csr(n, p14_off, p14_idx);
    R->rc = rc;
    R->rlist = rc + skin;
    R->prune_r = rc + std::min(0.8, skin);
    R->L.build(R->x, R->cell, R->rlist, R->excl, R->type, size_t(nt), {}, R->q, R->f12, R->f6, *R->pool, nullptr, nullptr,
               coul ? &R->p14 : nullptr);
    nbnxm::Params& p = R->p;
    const double pw = lj == 2 ? 9 : 12;
    p.rc2 = float(rc * rc);
    p.rc = float(rc);
    p.lj = lj;
    p.comb = true;
    p.nt = size_t(nt);
    p.shift = shift ? 1.0f : 0.0f;
    p.rcn = float(std::pow(rc, -pw));
    p.rc6 = float(std::pow(rc, -6.0));
    p.coul = coul;
    p.kappa = float(kappa);
    p.kc = float(2 * kappa / std::sqrt(M_PI));
    p.e0 = float(dsf_e0);
    p.f0 = float(dsf_f0);
    p.kcoul = float(kCoulombConstant);
    p.f14 = float(f14);
    R->pme = coul == 1 && grid;
    if (R->pme) R->order = pme_order, R->grid[0] = grid[0], R->grid[1] = grid[1], R->grid[2] = grid[2], R->beta = kappa;
    return R;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void ref_eval_free(void* h) { delete static_cast<RefEval*>(h); }
// forces (3N, overwritten), out[0..1] = vdw, coulomb (pair part, + reciprocal when PME), w[6] (pair + reciprocal);
// prune: 1 to prune at r_c + 0.8 first
int ref_eval_compute(void* h, const double* x, int energies, int prune, double* f, double out[2], double w[6]) {
  auto* R = static_cast<RefEval*>(h);
  try {
    std::memcpy(R->x.data(), x, 8 * R->x.size());
    R->L.positions(R->x, *R->pool);
    if (prune) R->L.prune(R->prune_r, *R->pool);
    R->p.energies = energies != 0;
    const size_t ns = R->L.perm.size(), nch = R->L.tb.size();
    R->tfc.resize(nch);
    std::vector<nbnxm::Sums> sums(nch);
    std::atomic<size_t> next{0};
    R->pool->run(size_t(R->pool->size()), [&](int, size_t, size_t) {
      for (size_t t; (t = next.fetch_add(1)) < nch;) {
        R->tfc[t].assign(3 * ns, 0.0f);
        nbnxm::kernel(R->L, R->p, int(t), R->tfc[t].data(), R->tfc[t].data() + ns, R->tfc[t].data() + 2 * ns, sums[t]);
      }
    });
    R->P.assign(3 * ns, 0.0);
    R->pool->run_slices(ns, R->pool->slices(), [&](size_t, size_t b, size_t e) {
      for (size_t sl = b; sl < e; ++sl) {
        double px = 0, py = 0, pz = 0;
        for (size_t c = 0; c < nch; ++c) {
          const float* fc = R->tfc[c].data();
          px += double(fc[sl]), py += double(fc[ns + sl]), pz += double(fc[2 * ns + sl]);
        }
        R->P[3 * sl] = px, R->P[3 * sl + 1] = py, R->P[3 * sl + 2] = pz;
      }
    });
    double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    nbnxm::pair_virial(R->L, R->P.data(), 0, ns, W);
    std::vector<double> fsh;
    double ev = 0, ec = 0;
    for (auto& s : sums) {
      if (fsh.size() < s.fsh.size()) fsh.resize(s.fsh.size(), 0.0);
      for (size_t k = 0; k < s.fsh.size(); ++k) fsh[k] += s.fsh[k];
      ev += s.evdw, ec += s.ecoul;
    }
    nbnxm::shift_virial(R->L, fsh, W);
    for (int i = 0; i < R->n; ++i)
      for (int k = 0; k < 3; ++k) f[3 * i + k] = R->P[3 * size_t(R->L.iperm[size_t(i)]) + size_t(k)];
    w[0] = W[0], w[1] = W[4], w[2] = W[8], w[3] = 0.5 * (W[1] + W[3]), w[4] = 0.5 * (W[2] + W[6]), w[5] = 0.5 * (W[5] + W[7]);
    if (R->pme) {
      std::vector<double> fv(3 * size_t(R->n), 0.0);
      double vk[6];
      PmeGrid g;
      g.k[0] = R->grid[0], g.k[1] = R->grid[1], g.k[2] = R->grid[2], g.order = R->order, g.beta = R->beta;
      ec += pme_reciprocal(R->x, R->q, R->cell, g, fv, vk, *R->pool);
      for (size_t k = 0; k < fv.size(); ++k) f[k] += fv[k];
      for (int k = 0; k < 6; ++k) w[k] += vk[k];
    }
    out[0] = ev, out[1] = ec;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
int ref_eval_threads(void* h) { return static_cast<RefEval*>(h)->pool->size(); }

// ---- the reference's random numbers (core/include/caps/rng.hpp, header-only) -----------------------------------------
// n draws of caps::Normal<double>(0, 1) from std::mt19937_64(seed), one distribution object (its spare kept)
void ref_rng_normal(uint64_t seed, int n, double* out) {
  std::mt19937_64 g(seed);
  caps::Normal<double> N(0, 1);
  for (int i = 0; i < n; ++i) out[i] = N(g);
}
// the Bussi draw pattern of dynamics.cpp:466-477: per step r1 from the run's normal generator, then Σ R² from a fresh
// caps::Gamma((ndof − 1)/2, 2) on the same engine; out = (r1, sum) pairs
void ref_rng_bussi(uint64_t seed, double ndof, int steps, double* out) {
  std::mt19937_64 g(seed);
  caps::Normal<double> gauss(0, 1);
  for (int s = 0; s < steps; ++s) {
    const double r1 = gauss(g);
    caps::Gamma<double> chi((ndof - 1) / 2, 2.0);
    out[2 * s] = r1, out[2 * s + 1] = chi(g);
  }
}
// n draws of caps::UniformReal<double>(a, b)
void ref_rng_uniform(uint64_t seed, double a, double b, int n, double* out) {
  std::mt19937_64 g(seed);
  caps::UniformReal<double> U(a, b);
  for (int i = 0; i < n; ++i) out[i] = U(g);
}

}  // extern "C"
