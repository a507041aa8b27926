// Protocols on the GPU dynamics. The glass-transition scan: stepwise cooling (core/src/mechanics.cpp:1381-1487,
// caps::run_cooling) of one context through capsgpu_run_dynamics, and of many replicas at once through the replica batch;
// the Tg fits of mechanics.cpp:1218-1379 (a continuous two-line fit with the hinge free, or two lines through a glassy
// and a rubbery range), with their bootstrap errors drawn as the reference draws them (std::mt19937_64(12345),
// caps::UniformInt). Equilibration: chained NVT / NPT / NVE stages (core/src/equilibrate.cpp:188-262, caps::equilibrate)
// and the 21-step and annealing protocols (equilibrate.cpp:43-75).
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "batch.hpp"
#include "context.hpp"

namespace capsgpu {
void evaluate(capsgpu_ctx& c, bool want, double* t, bool flags_known, int parts, double* fdst);
void set_last_error(const std::string& m);
}

using namespace capsgpu;

namespace {

// caps::UniformInt<size_t>(a, b) (core/include/caps/rng.hpp:54-67): rejection of the top partial range, then x mod n
size_t uniform_int(std::mt19937_64& g, size_t a, size_t b) {
  const uint64_t range = uint64_t(b) - uint64_t(a);
  if (range == std::numeric_limits<uint64_t>::max()) return size_t(g());
  const uint64_t n = range + 1, limit = std::numeric_limits<uint64_t>::max() - std::numeric_limits<uint64_t>::max() % n;
  uint64_t x;
  do x = g(); while (x >= limit);
  return size_t(uint64_t(a) + x % n);
}

// y = c0 + c1 min(T − h, 0) + c2 max(T − h, 0) by least squares for a fixed hinge h (T sorted); returns the residual sum
double hinge_solve(const std::vector<double>& T, const std::vector<double>& y, double h, double c[3]) {
  const size_t n = T.size();
  double A[3][4] = {};
  for (size_t i = 0; i < n; ++i) {
    const double u[3] = {1, std::min(T[i] - h, 0.0), std::max(T[i] - h, 0.0)};
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b) A[a][b] += u[a] * u[b];
      A[a][3] += u[a] * y[i];
    }
  }
  // Gauss–Jordan with partial pivoting
  for (int k = 0; k < 3; ++k) {
    int p = k;
    for (int i = k + 1; i < 3; ++i)
      if (std::fabs(A[i][k]) > std::fabs(A[p][k])) p = i;
    for (int j = 0; j < 4; ++j) std::swap(A[k][j], A[p][j]);
    if (std::fabs(A[k][k]) < 1e-300) return std::numeric_limits<double>::max();
    for (int i = 0; i < 3; ++i)
      if (i != k) {
        const double m = A[i][k] / A[k][k];
        for (int j = 0; j < 4; ++j) A[i][j] -= m * A[k][j];
      }
  }
  for (int k = 0; k < 3; ++k) c[k] = A[k][3] / A[k][k];
  double rss = 0;
  for (size_t i = 0; i < n; ++i) {
    const double r = y[i] - (c[0] + c[1] * std::min(T[i] - h, 0.0) + c[2] * std::max(T[i] - h, 0.0));
    rss += r * r;
  }
  return rss;
}

// the hinge between the second and the second-to-last temperature: a scan of `scan` + 1 points, then 60 golden-section
// steps around the best one
double best_hinge(const std::vector<double>& T, const std::vector<double>& y, int scan) {
  const size_t n = T.size();
  const double lo = T[1], hi = T[n - 2];
  double c[3], bh = lo, br = std::numeric_limits<double>::max();
  for (int k = 0; k <= scan; ++k) {
    const double h = lo + (hi - lo) * k / scan, r = hinge_solve(T, y, h, c);
    if (r < br) br = r, bh = h;
  }
  double a = std::max(lo, bh - (hi - lo) / scan), b = std::min(hi, bh + (hi - lo) / scan);
  const double g = 0.5 * (std::sqrt(5.0) - 1);
  double x1 = b - g * (b - a), x2 = a + g * (b - a), f1 = hinge_solve(T, y, x1, c), f2 = hinge_solve(T, y, x2, c);
  for (int it = 0; it < 60; ++it) {
    if (f1 < f2) {
      b = x2, x2 = x1, f2 = f1, x1 = b - g * (b - a), f1 = hinge_solve(T, y, x1, c);
    } else {
      a = x1, x1 = x2, f1 = f2, x2 = a + g * (b - a), f2 = hinge_solve(T, y, x2, c);
    }
  }
  const double h = 0.5 * (a + b);
  return hinge_solve(T, y, h, c) <= br ? h : bh;
}

}  // namespace

namespace capsgpu {

capsgpu_tg_fit fit_bilinear(const std::vector<double>& T0, const std::vector<double>& y0) {
  capsgpu_tg_fit fit{};
  const size_t n = T0.size();
  if (n != y0.size() || n < 5) {
    fit.note = 1;   // too few temperatures
    return fit;
  }
  std::vector<size_t> ord(n);
  std::iota(ord.begin(), ord.end(), size_t(0));
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return T0[a] < T0[b]; });
  std::vector<double> T(n), Y(n);
  for (size_t i = 0; i < n; ++i) T[i] = T0[ord[i]], Y[i] = y0[ord[i]];
  const double h = best_hinge(T, Y, 400);
  double c[3];
  fit.rss = hinge_solve(T, Y, h, c);
  fit.tg = h;
  fit.value_at_tg = c[0];
  fit.slope_low = c[1], fit.slope_high = c[2];
  fit.alpha_low = c[1] / c[0], fit.alpha_high = c[2] / c[0];
  fit.ok = 1;
  // bootstrap of residuals: 200 resamples, each hinge found with a coarser scan
  std::vector<double> fitted(n), res(n);
  for (size_t i = 0; i < n; ++i) {
    fitted[i] = c[0] + c[1] * std::min(T[i] - h, 0.0) + c[2] * std::max(T[i] - h, 0.0);
    res[i] = Y[i] - fitted[i];
  }
  std::mt19937_64 rng(12345);
  std::vector<double> hs;
  for (int b = 0; b < 200; ++b) {
    std::vector<double> yb(n);
    for (size_t i = 0; i < n; ++i) yb[i] = fitted[i] + res[uniform_int(rng, 0, n - 1)];
    hs.push_back(best_hinge(T, yb, 100));
  }
  double m = 0, v = 0;
  for (double x : hs) m += x / double(hs.size());
  for (double x : hs) v += (x - m) * (x - m) / double(hs.size() - 1);
  fit.tg_err = std::sqrt(v);
  if (std::fabs(c[2] - c[1]) < 0.1 * std::max(std::fabs(c[1]), std::fabs(c[2]))) fit.note = 2;   // slopes within 10 %
  else if (h - T[0] < 1.5 * (T[1] - T[0]) || T[n - 1] - h < 1.5 * (T[n - 1] - T[n - 2])) fit.note = 3;   // hinge at an edge
  return fit;
}

capsgpu_tg_fit fit_two_ranges(const std::vector<double>& T, const std::vector<double>& y, double glassy_max, double rubbery_min) {
  capsgpu_tg_fit fit{};
  const size_t n = T.size();
  if (n != y.size() || n < 4) {
    fit.note = 1;
    return fit;
  }
  const auto mm = std::minmax_element(T.begin(), T.end());
  const double tmin = *mm.first, tmax = *mm.second;
  if (glassy_max <= 0) glassy_max = tmin + (tmax - tmin) / 3.0;
  if (rubbery_min <= 0) rubbery_min = tmax - (tmax - tmin) / 3.0;
  if (glassy_max >= rubbery_min) {
    fit.note = 4;   // the ranges overlap
    return fit;
  }
  std::vector<size_t> lo, hi;
  for (size_t i = 0; i < n; ++i) {
    if (T[i] <= glassy_max + 1e-9) lo.push_back(i);
    if (T[i] >= rubbery_min - 1e-9) hi.push_back(i);
  }
  if (lo.size() < 2 || hi.size() < 2) {
    fit.note = 5;   // a range with fewer than two temperatures
    return fit;
  }
  auto line = [&](const std::vector<size_t>& idx, const std::vector<double>& Y, double& a, double& b) {
    double st = 0, sy = 0, stt = 0, sty = 0;
    for (size_t i : idx) st += T[i], sy += Y[i], stt += T[i] * T[i], sty += T[i] * Y[i];
    const double m = double(idx.size()), den = m * stt - st * st;
    b = den != 0 ? (m * sty - st * sy) / den : 0;
    a = (sy - b * st) / m;
  };
  auto cross_at = [&](const std::vector<double>& Y, double& tg, double& bl, double& bh, double& al, double& ah) {
    line(lo, Y, al, bl);
    line(hi, Y, ah, bh);
    tg = bl != bh ? (ah - al) / (bl - bh) : 0;
  };
  double al, bl, ah, bh, tg;
  cross_at(y, tg, bl, bh, al, ah);
  if (bl == bh) {
    fit.note = 6;   // parallel lines
    return fit;
  }
  fit.ok = 1;
  fit.tg = tg;
  fit.slope_low = bl, fit.slope_high = bh;
  fit.value_at_tg = al + bl * tg;
  if (fit.value_at_tg != 0) fit.alpha_low = bl / fit.value_at_tg, fit.alpha_high = bh / fit.value_at_tg;
  std::vector<double> res(n, 0), fitted(n, 0);
  for (size_t i : lo) fitted[i] = al + bl * T[i], res[i] = y[i] - fitted[i];
  for (size_t i : hi) fitted[i] = ah + bh * T[i], res[i] = y[i] - fitted[i];
  for (size_t i : lo) fit.rss += res[i] * res[i];
  for (size_t i : hi) fit.rss += res[i] * res[i];
  std::vector<size_t> used(lo);
  used.insert(used.end(), hi.begin(), hi.end());
  std::mt19937_64 rng(12345);
  std::vector<double> tgs;
  for (int b = 0; b < 400; ++b) {
    std::vector<double> yb = y;
    for (size_t i : used) yb[i] = fitted[i] + res[used[uniform_int(rng, 0, used.size() - 1)]];
    double t2, b1, b2, a1, a2;
    cross_at(yb, t2, b1, b2, a1, a2);
    if (b1 != b2 && std::isfinite(t2)) tgs.push_back(t2);
  }
  if (tgs.size() > 10) {
    const double m = std::accumulate(tgs.begin(), tgs.end(), 0.0) / double(tgs.size());
    double v = 0;
    for (double x : tgs) v += (x - m) * (x - m);
    fit.tg_err = std::sqrt(v / double(tgs.size() - 1));
  }
  if (lo.size() < 3 || hi.size() < 3) {
    fit.tg_err = std::numeric_limits<double>::quiet_NaN();
    fit.note = 7;   // two points fix a line exactly: no error estimate
  }
  if (tg < glassy_max || tg > rubbery_min) {
    fit.ok = 0;
    fit.note = 8;   // the lines cross outside the gap between the ranges
  }
  return fit;
}

// the point of one hold from its thermo rows: density and potential averaged over the rows from average_from on, the
// density's standard error from five blocks (mechanics.cpp:1440-1464)
capsgpu_cooling_point hold_point(double T, const std::vector<double>& dens_all, const std::vector<double>& pot_all, double average_from) {
  std::vector<double> dens, pot;
  for (size_t i = 0; i < dens_all.size(); ++i)
    if (i >= size_t(average_from * double(dens_all.size()))) dens.push_back(dens_all[i]), pot.push_back(pot_all[i]);
  capsgpu_cooling_point p{};
  p.temperature = T;
  p.density = std::accumulate(dens.begin(), dens.end(), 0.0) / double(std::max<size_t>(1, dens.size()));
  p.potential = std::accumulate(pot.begin(), pot.end(), 0.0) / double(std::max<size_t>(1, pot.size()));
  const size_t nb = std::min<size_t>(5, dens.size());
  if (nb >= 2) {
    std::vector<double> bm;
    for (size_t b = 0; b < nb; ++b) {
      double sum = 0;
      size_t cnt = 0;
      for (size_t i = b * dens.size() / nb; i < (b + 1) * dens.size() / nb; ++i) sum += dens[i], ++cnt;
      bm.push_back(sum / double(std::max<size_t>(1, cnt)));
    }
    double v = 0;
    for (double x : bm) v += (x - p.density) * (x - p.density);
    p.density_err = std::sqrt(v / double(nb - 1) / double(nb));
  }
  p.specific_volume = p.density > 0 ? 1 / p.density : 0;
  return p;
}

std::vector<double> cooling_ladder(const capsgpu_cooling_opts& o) {
  std::vector<double> temps;
  const double dir = o.t_end < o.t_start ? -1 : 1;
  for (double T = o.t_start; dir * (T - o.t_end) <= 1e-9; T += dir * o.t_step) temps.push_back(T);
  return temps;
}

capsgpu_tg_fit cooling_fit(const capsgpu_cooling_opts& o, const std::vector<capsgpu_cooling_point>& pts, double atoms) {
  std::vector<double> T, v;
  for (const auto& p : pts) T.push_back(p.temperature), v.push_back(o.property == 1 ? p.potential / atoms : p.specific_volume);
  return o.fit == 1 ? fit_two_ranges(T, v, o.glassy_max, o.rubbery_min) : fit_bilinear(T, v);
}

void check_cooling(const capsgpu_cooling_opts& o) {
  if (!(o.t_step > 0) || !(o.ps_per_step > 0)) throw Error(CAPSGPU_EINVAL, "temperature step and hold time must be positive");
  if (!(o.dt > 0) || o.dt > 5) throw Error(CAPSGPU_EINVAL, "time step must be in (0, 5] fs");
  if (!(o.average_from >= 0 && o.average_from < 1)) throw Error(CAPSGPU_EINVAL, "average_from must be in [0, 1)");
  if (o.property < 0 || o.property > 1 || o.fit < 0 || o.fit > 1) throw Error(CAPSGPU_EINVAL, "property and fit are 0 or 1");
  if (!(o.t_start > 0) || !(o.t_end > 0)) throw Error(CAPSGPU_EINVAL, "temperatures must be positive");
}

}  // namespace capsgpu

namespace {
template <class F>
int32_t guard_cool(F&& f) {
  try {
    f();
    return CAPSGPU_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return CAPSGPU_EINVAL;
  }
}

void rethrow(int32_t rc) {
  if (rc != CAPSGPU_OK) throw Error(rc, capsgpu_last_error());
}
}  // namespace

extern "C" {

int32_t capsgpu_tg_fit_points(const double* T, const double* y, int32_t n, int32_t fit, double glassy_max, double rubbery_min,
                              capsgpu_tg_fit* out) {
  return guard_cool([&] {
    if (!T || !y || !out || n < 0) throw Error(CAPSGPU_EINVAL, "null argument");
    const std::vector<double> t(T, T + n), v(y, y + n);
    *out = fit == 1 ? fit_two_ranges(t, v, glassy_max, rubbery_min) : fit_bilinear(t, v);
  });
}

int32_t capsgpu_run_cooling(capsgpu_ctx* c, const capsgpu_cooling_opts* o, capsgpu_cooling_progress_fn progress, void* user,
                            capsgpu_cooling_point* points, int32_t cap, int32_t* npoints, capsgpu_tg_fit* fit) {
  return guard_cool([&] {
    if (!c || !o) throw Error(CAPSGPU_EINVAL, "null argument");
    check_cooling(*o);
    const std::vector<double> temps = cooling_ladder(*o);
    const int64_t steps = std::llround(o->ps_per_step * 1000 / o->dt);
    const double eq_ps = o->equilibrate_ps < 0 ? o->ps_per_step : o->equilibrate_ps;
    bool first_velocities = o->new_velocities != 0;
    int64_t offset = 0;
    struct Ctx {
      capsgpu_cooling_progress_fn fn;
      void* user;
      int index, total;
    } pc{progress, user, -1, int(temps.size())};
    auto relay = [](const capsgpu_thermo* r, void* u) -> int32_t {
      const Ctx* p = static_cast<const Ctx*>(u);
      return p->fn ? p->fn(r, p->index, p->total, p->user) : 1;
    };
    auto stage = [&](double T, int64_t nsteps, uint64_t seed, bool newv, int64_t every, std::vector<capsgpu_thermo>& rows) {
      capsgpu_dyn_opts d{};
      d.dt = o->dt, d.steps = nsteps, d.temperature = T, d.thermostat = 1, d.tau_t = o->tau_t, d.seed = seed;
      d.thermo_every = int32_t(std::max<int64_t>(1, every)), d.new_velocities = newv ? 1 : 0, d.respa = 1, d.step_offset = offset;
      d.learned_train = -1;
      d.barostat = o->barostat, d.pressure = o->pressure, d.tau_p = o->tau_p, d.compressibility = o->compressibility;
      d.barostat_every = o->barostat_every;
      d.temperature_end = -1;
      rows.assign(size_t(nsteps / std::max<int64_t>(1, every) + 2), capsgpu_thermo{});
      int32_t nr = 0;
      rethrow(capsgpu_run_dynamics(c, &d, relay, &pc, rows.data(), int32_t(rows.size()), &nr, nullptr));
      rows.resize(size_t(std::min<int32_t>(nr, int32_t(rows.size()))));
      offset += nsteps;
    };
    std::vector<capsgpu_thermo> rows;
    // an unsampled run at the starting temperature first (Bussi, seed + 1000)
    if (eq_ps > 0) {
      const int64_t eq = std::llround(eq_ps * 1000 / o->dt);
      stage(o->t_start, eq, o->seed + 1000, first_velocities, eq / 50, rows);
      first_velocities = false;
    }
    std::vector<capsgpu_cooling_point> pts;
    for (size_t k = 0; k < temps.size(); ++k) {
      pc.index = int(k);
      stage(temps[k], steps, o->seed + k, k == 0 && first_velocities, steps / 200, rows);
      std::vector<double> dens, pot;
      for (const auto& r : rows) dens.push_back(r.density), pot.push_back(r.potential);
      pts.push_back(hold_point(temps[k], dens, pot, o->average_from));
    }
    if (points)
      for (size_t k = 0; k < pts.size() && int32_t(k) < cap; ++k) points[k] = pts[k];
    if (npoints) *npoints = int32_t(pts.size());
    if (fit) *fit = cooling_fit(*o, pts, double(std::max(1, c->n)));
  });
}

int32_t capsgpu_batch_cooling(capsgpu_batch* b, const capsgpu_cooling_opts* o, int32_t samples, capsgpu_cooling_point* points,
                              int32_t cap, int32_t* npoints, capsgpu_tg_fit* fits) {
  return guard_cool([&] {
    if (!b || !o) throw Error(CAPSGPU_EINVAL, "null argument");
    check_cooling(*o);
    if (o->barostat != 0) throw Error(CAPSGPU_EINVAL, "the replica batch steps at constant volume: barostat must be 0");
    if (o->property != 1) throw Error(CAPSGPU_EINVAL, "at constant volume Tg is read from the potential energy (property 1)");
    if (samples < 1) throw Error(CAPSGPU_EINVAL, "at least one sample per hold");
    const int nrep = batch_size(*b);
    const std::vector<double> temps = cooling_ladder(*o);
    const int64_t steps = std::llround(o->ps_per_step * 1000 / o->dt);
    const int64_t skip = int64_t(std::floor(o->average_from * double(steps)));
    const int64_t chunk = std::max<int64_t>(1, (steps - skip) / samples);
    const double eq_ps = o->equilibrate_ps < 0 ? o->ps_per_step : o->equilibrate_ps;
    std::vector<double> Tr(static_cast<size_t>(nrep));
    uint64_t run = 0;   // every batch run draws from its own seeds: seed + 1000 + nrep·run + replica
    auto advance = [&](double T, int64_t nsteps, bool newv) {
      if (nsteps <= 0) return;
      capsgpu_batch_opts bo{};
      bo.dt = o->dt, bo.steps = nsteps, bo.tau_t = o->tau_t, bo.seed = o->seed + 1000 + uint64_t(nrep) * run++;
      bo.new_velocities = newv ? 1 : 0, bo.mode = 0, bo.branches = 1, bo.graph_steps = 1;
      std::fill(Tr.begin(), Tr.end(), T);
      batch_run(*b, bo, Tr.data(), nullptr, nullptr);
    };
    bool newv = o->new_velocities != 0;
    if (eq_ps > 0) advance(o->t_start, std::llround(eq_ps * 1000 / o->dt), newv), newv = false;
    std::vector<std::vector<capsgpu_cooling_point>> pts(static_cast<size_t>(nrep));
    std::vector<std::vector<double>> pot(static_cast<size_t>(nrep));
    for (size_t k = 0; k < temps.size(); ++k) {
      advance(temps[k], steps - chunk * samples, newv && k == 0);
      newv = false;
      for (auto& p : pot) p.clear();
      for (int q = 0; q < samples; ++q) {
        advance(temps[k], chunk, false);
        for (int r = 0; r < nrep; ++r) {
          capsgpu_ctx& c = *batch_replica(*b, r);
          double t[CAPSGPU_NTERMS];
          evaluate(c, true, t, false, -1, nullptr);
          pot[size_t(r)].push_back(t[CAPSGPU_T_BOND] + t[CAPSGPU_T_ANGLE] + t[CAPSGPU_T_DIHEDRAL] + t[CAPSGPU_T_IMPROPER] +
                                   t[CAPSGPU_T_VDW] + t[CAPSGPU_T_COULOMB]);
        }
      }
      for (int r = 0; r < nrep; ++r) {
        capsgpu_ctx& c = *batch_replica(*b, r);
        double mtot = 0;
        for (double m : c.mass) mtot += m;
        const double* e = c.cell + 3;
        const double V = std::fabs(e[0] * (e[4] * e[8] - e[5] * e[7]) - e[1] * (e[3] * e[8] - e[5] * e[6]) + e[2] * (e[3] * e[7] - e[4] * e[6]));
        const std::vector<double> dens(pot[size_t(r)].size(), mtot / 6.02214076e23 / (V * 1e-24));
        pts[size_t(r)].push_back(hold_point(temps[k], dens, pot[size_t(r)], 0.0));
      }
    }
    for (int r = 0; r < nrep; ++r) {
      if (points)
        for (size_t k = 0; k < pts[size_t(r)].size() && int32_t(k) < cap; ++k) points[size_t(r) * size_t(cap) + k] = pts[size_t(r)][k];
      if (fits) fits[r] = cooling_fit(*o, pts[size_t(r)], double(std::max(1, batch_replica(*b, r)->n)));
    }
    if (npoints) *npoints = int32_t(temps.size());
  });
}

int32_t capsgpu_protocol(const char* name, const capsgpu_protocol_params* p, capsgpu_eq_stage* out, int32_t cap, int32_t* n) {
  return guard_cool([&] {
    if (!name || !p || !n) throw Error(CAPSGPU_EINVAL, "null argument");
    if (!(p->time_scale > 0)) throw Error(CAPSGPU_EINVAL, "the time scale must be above 0");
    if (!(p->ramp_ps >= 0) || !(p->hold_ps >= 0)) throw Error(CAPSGPU_EINVAL, "stage lengths must be 0 ps or more");
    if (p->cycles < 0) throw Error(CAPSGPU_EINVAL, "the number of annealing cycles must be 0 or more");
    std::vector<capsgpu_eq_stage> st;
    auto add = [&](int ens, double ps, double t0, double t1, double pr) {
      capsgpu_eq_stage s{};
      s.ensemble = ens, s.ps = ps * p->time_scale, s.t_start = t0, s.t_end = t1, s.pressure = pr;
      st.push_back(s);
    };
    const std::string nm = name;
    if (nm == "larsen21" || nm == "21-step") {
      // Larsen, Lin, Hart and Colina, Macromolecules 44, 6944 (2011), Table 1
      const double tm = p->t_max, tf = p->t_final, pm = p->p_max;
      const double fr[6] = {0.02, 0.6, 1.0, 0.5, 0.1, 0.01};
      const double heat[6] = {50, 50, 50, 50, 5, 5}, cool[6] = {50, 100, 100, 100, 10, 10}, comp[6] = {50, 50, 50, 5, 5, 5};
      for (int k = 0; k < 6; ++k) {
        add(0, heat[k], tm, -1, 1.0);
        add(0, cool[k], tf, -1, 1.0);
        add(1, comp[k], tf, -1, fr[k] * pm);
      }
      // steps 19 and 20 (heat, cool) and the final NPT stage
      add(0, 5, tm, -1, 1.0);
      add(0, 10, tf, -1, 1.0);
      add(1, 800, tf, -1, p->p_final);
    } else if (nm == "annealing" || nm == "anneal") {
      add(1, p->hold_ps, p->t_low, -1, p->p_final);
      for (int c = 1; c <= p->cycles; ++c) {
        add(1, p->ramp_ps, p->t_low, p->t_high, p->p_final);
        add(1, p->hold_ps, p->t_high, -1, p->p_final);
        add(1, p->ramp_ps, p->t_high, p->t_low, p->p_final);
        add(1, p->hold_ps, p->t_low, -1, p->p_final);
      }
    } else {
      throw Error(CAPSGPU_EINVAL, "unknown protocol '" + nm + "' (larsen21, annealing)");
    }
    *n = int32_t(st.size());
    if (out)
      for (size_t k = 0; k < st.size() && int32_t(k) < cap; ++k) out[k] = st[k];
  });
}

int32_t capsgpu_equilibrate(capsgpu_ctx* c, const capsgpu_eq_stage* stages, int32_t nstages, const capsgpu_eq_opts* o,
                            capsgpu_eq_progress_fn progress, void* user, capsgpu_eq_summary* out) {
  return guard_cool([&] {
    if (!c || !stages || !o || nstages < 1) throw Error(CAPSGPU_EINVAL, "the protocol has no stages");
    if (!(o->dt > 0) || o->dt > 5) throw Error(CAPSGPU_EINVAL, "time step must be in (0, 5] fs");
    for (int k = 0; k < nstages; ++k) {
      if (stages[k].ensemble < 0 || stages[k].ensemble > 2) throw Error(CAPSGPU_EINVAL, "a stage's ensemble is 0 (NVT), 1 (NPT) or 2 (NVE)");
      if (!(stages[k].ps > 0)) throw Error(CAPSGPU_EINVAL, "every stage needs a positive duration");
    }
    int64_t offset = 0;
    int64_t last_step = -1;
    struct Ctx {
      capsgpu_eq_progress_fn fn;
      void* user;
      int index, total;
      int64_t* last;
      std::vector<capsgpu_thermo>* rows;
    };
    std::vector<capsgpu_thermo> kept;
    Ctx pc{progress, user, 0, nstages, &last_step, &kept};
    auto relay = [](const capsgpu_thermo* r, void* u) -> int32_t {
      Ctx* p = static_cast<Ctx*>(u);
      // the first row of a stage repeats the last row of the previous one
      if (*p->last == r->step) return 1;
      *p->last = r->step;
      p->rows->push_back(*r);
      return p->fn ? p->fn(r, p->index, p->total, p->user) : 1;
    };
    for (int k = 0; k < nstages; ++k) {
      const capsgpu_eq_stage& st = stages[k];
      capsgpu_dyn_opts d{};
      const int64_t steps = std::max<int64_t>(1, std::llround(st.ps * 1000.0 / o->dt));
      d.dt = o->dt, d.steps = steps, d.temperature = st.t_start, d.temperature_end = st.t_end;
      d.thermostat = st.ensemble == 2 ? 0 : (o->thermostat == 0 ? 1 : o->thermostat);
      d.tau_t = o->tau_t;
      d.barostat = st.ensemble == 1 ? (o->barostat == 0 ? 2 : o->barostat) : 0;
      d.pressure = st.pressure, d.tau_p = o->tau_p, d.compressibility = o->compressibility, d.barostat_every = o->barostat_every;
      d.seed = o->seed + 1000003ull * uint64_t(k);
      d.step_offset = offset;
      const double tps = o->thermo_ps > 0 ? o->thermo_ps : std::clamp(st.ps / 200.0, 10.0 * o->dt / 1000.0, 0.5);
      d.thermo_every = std::max(1, int(std::lround(tps * 1000.0 / o->dt)));
      d.new_velocities = 0, d.respa = 1, d.learned_train = -1;
      pc.index = k;
      const size_t first = kept.size();
      std::vector<capsgpu_thermo> rows(size_t(steps / d.thermo_every + 2));
      int32_t nr = 0;
      rethrow(capsgpu_run_dynamics(c, &d, relay, &pc, rows.data(), int32_t(rows.size()), &nr, nullptr));
      offset += steps;
      capsgpu_eq_summary sum{};
      sum.ps = double(steps) * o->dt / 1000.0;
      const size_t a = first + (kept.size() - first) / 2, m = kept.size() - a;
      for (size_t q = a; q < kept.size(); ++q) {
        sum.temperature += kept[q].temperature / double(m);
        sum.pressure += kept[q].pressure / double(m);
        sum.density += kept[q].density / double(m);
        sum.potential += kept[q].potential / double(m);
      }
      sum.density_end = kept.empty() ? 0 : kept.back().density;
      if (out) out[k] = sum;
    }
  });
}

}  // extern "C"
