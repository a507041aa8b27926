// Energy minimisation on the GPU: steepest descent, Polak–Ribière conjugate gradients, L-BFGS (m = 10, Liu & Nocedal
// 1989, two-loop recursion) and FIRE (Bitzek et al. 2006) — caps::minimise (core/src/relax.cpp:65-282) with every 3N
// vector (positions, forces, directions, the L-BFGS pairs, FIRE's velocities) in device memory and the forces and energy
// from the context's own evaluation. Only scalars (dot products, largest atomic force or move, energies) come back to
// the host, each from a fixed-order reduction, so a minimisation is reproducible.
#include <algorithm>
#include <cmath>
#include <deque>
#include <string>
#include <vector>

#include "context.hpp"

namespace capsgpu {
void evaluate(capsgpu_ctx& c, bool want, double* t, bool flags_known, int parts, double* fdst);
void set_last_error(const std::string& m);

namespace {

constexpr int kRedThreads = 256, kRedBlocks = 256;

// per-workgroup partial sums of a·b over n elements (grid-stride, each workgroup's lanes in a fixed order)
__global__ void __launch_bounds__(256) k_dot_part(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ part) {
  __shared__ double red[kRedThreads];
  double s = 0;
  for (size_t i = size_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += size_t(gridDim.x) * kRedThreads) s += a[i] * b[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// per-workgroup maxima of |v_i|² over the atoms (3 per atom)
__global__ void __launch_bounds__(256) k_max3_part(int natoms, const double* __restrict__ v, double* __restrict__ part) {
  __shared__ double red[kRedThreads];
  double m = 0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < natoms; i += gridDim.x * kRedThreads) {
    const double q = v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2];
    m = q > m ? q : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + st] ? red[threadIdx.x] : red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// the partials folded by one workgroup in a fixed order: sum (kMax false) or maximum
template <bool kMax>
__global__ void __launch_bounds__(256) k_fold(const double* __restrict__ part, int rows, double* __restrict__ out) {
  __shared__ double red[kRedThreads];
  double s = 0;
  for (int r = threadIdx.x; r < rows; r += kRedThreads) s = kMax ? (part[r] > s ? part[r] : s) : s + part[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = kMax ? (red[threadIdx.x] > red[threadIdx.x + st] ? red[threadIdx.x] : red[threadIdx.x + st])
                                                  : red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// out = α x + β y (out may alias x or y)
__global__ void k_axpby(size_t n, double al, const double* __restrict__ x, double be, const double* y, double* out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = al * x[i] + (y ? be * y[i] : 0.0);
}

// FIRE's mixing v = (1 − α) v + α |v| f / |f|, then v += dt f and the move d = dt v
__global__ void k_fire_mix(size_t n, double keep, double mix, double dt, const double* __restrict__ f, double* __restrict__ v,
                           double* __restrict__ d) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double vi = keep * v[i] + mix * f[i];
  vi += dt * f[i];
  v[i] = vi;
  d[i] = dt * vi;
}

// held atoms feel no force (and so never move)
__global__ void k_hold(int natoms, const uint8_t* __restrict__ fixed, double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < natoms && fixed[i]) f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = 0.0;
}

struct Min {
  capsgpu_ctx& c;
  size_t n3;
  DevBuf<double> part, scal;
  DevBuf<uint8_t> fixed;
  bool have_fixed = false;
  int nev = 0;
  explicit Min(capsgpu_ctx& cc) : c(cc), n3(3 * size_t(cc.n)) {
    part.resize(kRedBlocks);
    scal.resize(8);
  }
  double fetch(const double* dev) {
    double h = 0;
    HIPCHECK(hipMemcpyAsync(&h, dev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    return h;
  }
  double dot(const double* a, const double* b) {
    hipLaunchKernelGGL(k_dot_part, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, n3, a, b, part.p);
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_fold<false>, dim3(1), dim3(kRedThreads), 0, c.stream, part.p, kRedBlocks, scal.p);
    LAUNCHCHECK();
    return fetch(scal.p);
  }
  // the largest atomic magnitude of a 3N vector (max_force / max_disp, relax.cpp:47-61)
  double max3(const double* v) {
    hipLaunchKernelGGL(k_max3_part, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, c.n, v, part.p);
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_fold<true>, dim3(1), dim3(kRedThreads), 0, c.stream, part.p, kRedBlocks, scal.p);
    LAUNCHCHECK();
    return std::sqrt(fetch(scal.p));
  }
  void axpby(double a, const double* x, double b, const double* y, double* out) {
    hipLaunchKernelGGL(k_axpby, dim3(blocks_for(n3, 256)), dim3(256), 0, c.stream, n3, a, x, b, y, out);
    LAUNCHCHECK();
  }
  void copy(const double* x, double* out) { axpby(1.0, x, 0.0, nullptr, out); }
  // energy and forces at the positions p (the context's positions are p during the evaluation)
  double eval(DevBuf<double>& p, DevBuf<double>& f) {
    ++nev;
    c.x.swap(p);
    double t[CAPSGPU_NTERMS];
    struct Back {
      capsgpu_ctx& c;
      DevBuf<double>& p;
      ~Back() { c.x.swap(p); }
    } back{c, p};
    evaluate(c, true, t, false, -1, f.p);
    if (have_fixed) {
      hipLaunchKernelGGL(k_hold, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, fixed.p, f.p);
      LAUNCHCHECK();
    }
    return t[CAPSGPU_T_BOND] + t[CAPSGPU_T_ANGLE] + t[CAPSGPU_T_DIHEDRAL] + t[CAPSGPU_T_IMPROPER] + t[CAPSGPU_T_VDW] + t[CAPSGPU_T_COULOMB];
  }
};

}  // namespace

// caps::minimise on the context's device state, from the positions in X (3N, device); X holds the result
void minimise_device(capsgpu_ctx& c, DevBuf<double>& X, const capsgpu_relax_opts& o, const uint8_t* fixed_host, capsgpu_relax_stage& st) {
  Min M(c);
  const size_t n3 = M.n3;
  if (fixed_host) {
    M.fixed.upload(fixed_host, size_t(c.n), c.stream);
    M.have_fixed = true;
  }
  DevBuf<double> f, fn, xn, d;
  f.resize(n3), fn.resize(n3), xn.resize(n3), d.resize(n3);
  double e = M.eval(X, f);
  double fmax = M.max3(f.p);
  int it = 0;
  st.stopped_by = 0;   // iterations
  if (fmax < o.ftol) st.stopped_by = 1;
  if (o.method == 3) {
    // FIRE with the standard parameters, the move capped at 0.1 Å per atom (relax.cpp:133-160)
    const int nmin = 5;
    const double finc = 1.1, fdec = 0.5, astart = 0.1, falpha = 0.99, dtmax = 0.5, maxstep = 0.1;
    double dt = 0.05, alpha = astart;
    int npos = 0;
    DevBuf<double> v;
    v.resize(n3);
    v.zero(c.stream);
    for (; it < o.max_iterations && st.stopped_by != 1; ++it) {
      const double p = M.dot(f.p, v.p);
      double keep = 1.0, mix = 0.0;
      if (p > 0) {
        const double vn = std::sqrt(M.dot(v.p, v.p)), fnn = std::sqrt(M.dot(f.p, f.p));
        keep = 1 - alpha, mix = alpha * vn / std::max(fnn, 1e-300);
        if (++npos > nmin) dt = std::min(dt * finc, dtmax), alpha *= falpha;
      } else {
        v.zero(c.stream);
        dt *= fdec;
        alpha = astart;
        npos = 0;
      }
      hipLaunchKernelGGL(k_fire_mix, dim3(blocks_for(n3, 256)), dim3(256), 0, c.stream, n3, keep, mix, dt, f.p, v.p, d.p);
      LAUNCHCHECK();
      const double md = M.max3(d.p);
      const double sc = md > maxstep ? maxstep / md : 1.0;
      M.axpby(1.0, X.p, sc, d.p, X.p);
      e = M.eval(X, f);
      fmax = M.max3(f.p);
      if (fmax < o.ftol) st.stopped_by = 1;
    }
  } else {
    // line-search methods: backtracking (Armijo), the first trial's largest atomic move limited (relax.cpp:161-275)
    const int m = 10;
    std::deque<DevBuf<double>> S, Y;
    std::deque<double> rho;
    DevBuf<double> fprev, dprev, xq, fq, q;
    bool have_prev = false, cg_restarted = false;
    dprev.resize(n3);
    dprev.zero(c.stream);
    if (o.method == 1) xq.resize(n3), fq.resize(n3);
    if (o.method == 2) q.resize(n3);
    double step = 0.1;
    const double step_max = o.method == 0 ? 0.3 : 0.5;
    for (; it < o.max_iterations && st.stopped_by != 1; ++it) {
      if (o.method == 0) {
        M.copy(f.p, d.p);
      } else if (o.method == 1) {
        if (!have_prev) {
          M.copy(f.p, d.p);
        } else {
          const double den = M.dot(fprev.p, fprev.p);
          const double num = M.dot(f.p, f.p) - M.dot(f.p, fprev.p);
          const double beta = den > 0 ? std::max(0.0, num / den) : 0.0;
          M.axpby(1.0, f.p, beta, dprev.p, d.p);
        }
      } else {
        // the two-loop recursion on the gradient g = −f
        M.axpby(-1.0, f.p, 0.0, nullptr, q.p);
        std::vector<double> al(S.size());
        for (int j = int(S.size()) - 1; j >= 0; --j) {
          al[size_t(j)] = rho[size_t(j)] * M.dot(S[size_t(j)].p, q.p);
          M.axpby(1.0, q.p, -al[size_t(j)], Y[size_t(j)].p, q.p);
        }
        double gamma = 1.0 / 300.0;   // initial inverse-Hessian scale (Å² mol/kcal), about a stiff bond
        if (!S.empty()) gamma = M.dot(S.back().p, Y.back().p) / M.dot(Y.back().p, Y.back().p);
        M.axpby(gamma, q.p, 0.0, nullptr, q.p);
        for (size_t j = 0; j < S.size(); ++j) {
          const double b = rho[j] * M.dot(Y[j].p, q.p);
          M.axpby(1.0, q.p, al[j] - b, S[j].p, q.p);
        }
        M.axpby(-1.0, q.p, 0.0, nullptr, d.p);
      }
      double slope = M.dot(f.p, d.p);
      if (slope <= 0) {
        M.copy(f.p, d.p);
        slope = M.dot(f.p, f.p);
        S.clear(), Y.clear(), rho.clear();
      }
      const double md = M.max3(d.p);
      double a = (o.method == 2 && !S.empty()) ? std::min(1.0, step_max / md) : step / md;
      bool ok = false;
      double en = 0;
      for (int ls = 0; ls < 30; ++ls) {
        M.axpby(1.0, X.p, a, d.p, xn.p);
        en = M.eval(xn, fn);
        if (en <= e - 1e-4 * a * slope) {
          ok = true;
          break;
        }
        a *= 0.5;
      }
      // conjugate gradients: the minimum of the parabola through E(0), its slope and E(a) tried as well
      if (ok && o.method == 1) {
        const double curv = en - e + a * slope;
        if (curv > 0) {
          const double aq = 0.5 * slope * a * a / curv;
          if (aq > 0 && aq < 4 * a && std::fabs(aq - a) > 1e-3 * a && aq * md <= step_max) {
            M.axpby(1.0, X.p, aq, d.p, xq.p);
            const double eq = M.eval(xq, fq);
            if (eq < en) xn.swap(xq), fn.swap(fq), en = eq, a = aq;
          }
        }
      }
      if (!ok) {
        if (!S.empty() || have_prev) {   // start over from steepest descent once
          S.clear(), Y.clear(), rho.clear();
          have_prev = false;
          step = 0.02;
          continue;
        }
        st.stopped_by = 3;   // line search
        break;
      }
      const double moved = a * md;
      step = std::clamp(moved * 1.5, 1e-4, step_max);
      if (o.method == 2) {
        DevBuf<double> s, y;
        s.resize(n3), y.resize(n3);
        M.axpby(1.0, xn.p, -1.0, X.p, s.p);
        M.axpby(1.0, f.p, -1.0, fn.p, y.p);
        const double sy = M.dot(s.p, y.p);
        if (sy > 1e-10) {
          S.push_back(std::move(s)), Y.push_back(std::move(y)), rho.push_back(1.0 / sy);
          if (int(S.size()) > m) S.pop_front(), Y.pop_front(), rho.pop_front();
        }
      }
      if (o.method == 1) {
        fprev.resize(n3);
        M.copy(f.p, fprev.p);
        M.copy(d.p, dprev.p);
        have_prev = true;
      } else {
        have_prev = true;
      }
      const double de = e - en;
      X.swap(xn);
      f.swap(fn);
      e = en;
      fmax = M.max3(f.p);
      if (fmax < o.ftol) {
        st.stopped_by = 1;
      } else if (de >= 0 && de < o.etol * std::max(1.0, std::fabs(e))) {
        if (o.method == 1 && !cg_restarted) {
          cg_restarted = true;
          have_prev = false;
          step = 0.1;
          continue;
        }
        st.stopped_by = 2;   // energy
        ++it;
        break;
      } else {
        cg_restarted = false;
      }
    }
  }
  HIPCHECK(hipStreamSynchronize(c.stream));
  st.iterations = it;
  st.evaluations = M.nev;
  st.energy = e;
  st.fmax = fmax;
}

}  // namespace capsgpu

extern "C" int32_t capsgpu_minimise(capsgpu_ctx* c, double* x, const double cell[12], const capsgpu_relax_opts* o, const uint8_t* fixed,
                                    capsgpu_relax_stage* out) {
  using namespace capsgpu;
  try {
    if (!c || !x || !cell || !o || !out) throw Error(CAPSGPU_EINVAL, "null argument");
    if (o->method < 0 || o->method > 3) throw Error(CAPSGPU_EINVAL, "method must be 0 (steepest descent), 1 (CG), 2 (L-BFGS) or 3 (FIRE)");
    if (o->max_iterations < 0) throw Error(CAPSGPU_EINVAL, "max_iterations must be 0 or more");
    for (size_t k = 0; k < 3 * size_t(c->n); ++k)
      if (!std::isfinite(x[k])) throw Error(CAPSGPU_EFIELD, "positions are not finite numbers");
    const int32_t rc = capsgpu_set_state(c, x, nullptr, cell);
    if (rc != CAPSGPU_OK) return rc;
    DevBuf<double> X;
    X.resize(3 * size_t(c->n));
    HIPCHECK(hipMemcpyAsync(X.p, c->x.p, 3 * size_t(c->n) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    *out = capsgpu_relax_stage{};
    minimise_device(*c, X, *o, fixed, *out);
    // the result becomes the context's state and is returned
    HIPCHECK(hipMemcpyAsync(c->x.p, X.p, 3 * size_t(c->n) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    X.download(x, 3 * size_t(c->n), c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CAPSGPU_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return CAPSGPU_EINVAL;
  }
}
