// Replicas stepped together: many contexts (the melts of a cooling ladder, each with its own force field, cell, PME grid
// and temperature) take Langevin BAOAB steps as one unit of work. The reference runs such replicas one after another,
// each through run_dynamics (core/src/mechanics.cpp:1381-1470, the calls at :1418 and :1443).
//
// A replica's step without a pair list (BAOAB open, PME and bonded forces, closing kick) needs no host synchronisation
// and no memset, so a chunk of S steps of every replica is captured once into HIP graphs and replayed. The replicas are
// spread over B streams (longest first, each to the stream with the fewest atoms so far); each stream captures and
// replays its own graph, so the replicas of different streams run side by side on the GPU's queues. The step number
// the Langevin noise is drawn for comes from a per-replica counter in device memory (the graph's kernels read it, a
// last node advances it by S), so every replay draws fresh noise. The launch shape (B, S), or direct launches without a
// graph, is picked by timing each candidate for a few steps from a snapshot of the state that is restored afterwards.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "batch.hpp"
#include "context.hpp"
#include "host_rng.hpp"
#include "step.hpp"

namespace capsgpu {
void evaluate(capsgpu_ctx& c, bool want, double* t, bool flags_known, int parts, double* fdst);
}

namespace {
bool pme_on_ctx(const capsgpu_ctx& c) { return c.opt.coulomb && c.opt.electrostatics == 1 && !c.opt.no_reciprocal && c.pme.green_ok; }
}

using namespace capsgpu;

namespace {

// one replica's run: its Langevin constants and the device arrays the captured steps use
struct Rep {
  capsgpu_ctx* c = nullptr;
  double hk = 0, hdt = 0, c1 = 0, c2 = 0;
  uint64_t seed = 0;
  DevBuf<double> sdv;        // per atom: √(k_B T kAcc / m)
  DevBuf<double> bpart;      // per atom block: the bath's heat (accumulated, not reported by the batch)
  DevBuf<double> kpart;      // per atom block: kinetic energy partials
  DevBuf<double> ke;         // the reduced kinetic energy
  DevBuf<uint64_t> stepc;    // steps taken
  // the snapshot the tuner restores
  DevBuf<double> sx, sv, sf;
  DevBuf<uint64_t> sstep;
};

struct Shape {
  int mode = 0;              // 1 graphs per stream, 2 direct, 3 graphs of grouped launches (one stream)
  int branches = 1, steps = 1;
  bool operator==(const Shape& o) const { return mode == o.mode && branches == o.branches && steps == o.steps; }
};

}  // namespace

struct capsgpu_batch {
  std::vector<Rep> r;
  std::vector<hipStream_t> streams;           // one per branch (created on demand)
  // the graphs of the current shape, one per branch. They hold the device addresses of the replicas' arrays, which a new
  // cell (another PME grid) may reallocate, so every run captures afresh.
  std::vector<hipGraphExec_t> exec;
  Shape captured;
  // the tuned shape and the run constants it was timed with
  bool have_tuned = false;
  Shape tuned;
  std::vector<double> tuned_key;
  // the grouped launches of this run (null when the replicas cannot share launches)
  PmeGroup* pg = nullptr;
  StepGroup* sg = nullptr;
  ~capsgpu_batch();
};

namespace {

void drop_groups(capsgpu_batch& b) {
  if (b.pg) pme_group_destroy(b.pg);
  if (b.sg) step_group_destroy(b.sg);
  b.pg = nullptr, b.sg = nullptr;
}

// one Langevin step of every replica as grouped launches on stream s (the noise for step *stepc + k)
void grouped_step(capsgpu_batch& b, hipStream_t s, uint64_t k) {
  step_group_open(*b.sg, s, k);   // with the per-atom half of the PME bookkeeping on the new positions
  pme_group_forces(*b.pg, s, true);
  step_group_close(*b.sg, s);
}

void drop_graphs(capsgpu_batch& b) {
  for (auto& e : b.exec)
    if (e) (void)hipGraphExecDestroy(e);
  b.exec.clear();
  b.captured = Shape{};
}

void ensure_streams(capsgpu_batch& b, int m) {
  while (int(b.streams.size()) < m) {
    hipStream_t s = nullptr;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b.streams.push_back(s);
  }
}

// replicas by branch: longest first, each to the branch holding the fewest atoms so far
std::vector<std::vector<int>> assign(const capsgpu_batch& b, int branches) {
  std::vector<int> order(b.r.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return b.r[size_t(p)].c->n > b.r[size_t(q)].c->n; });
  std::vector<std::vector<int>> out(static_cast<size_t>(branches));
  std::vector<int64_t> load(static_cast<size_t>(branches), 0);
  for (int i : order) {
    const size_t j = size_t(std::min_element(load.begin(), load.end()) - load.begin());
    out[j].push_back(i);
    load[j] += b.r[size_t(i)].c->n;
  }
  return out;
}

// one BAOAB step of replica R on stream s, the noise drawn for step *stepc + k
void one_step(Rep& R, hipStream_t s, uint64_t k) {
  capsgpu_ctx& c = *R.c;
  const hipStream_t own = c.stream;
  c.stream = s;
  try {
    baoab_open(c, R.hk, R.hdt, R.c1, R.c2, R.sdv.p, R.seed, R.stepc.p, k, -1.0, -1.0, R.bpart.p);
    evaluate(c, false, nullptr, true, -1, nullptr);
    kick(c, R.hk, false, R.kpart.p);
  } catch (...) {
    c.stream = own;
    throw;
  }
  c.stream = own;
}

// the graphs of shape sh: branch j captures S steps of each of its replicas and the counters' advance; in the grouped
// shape one stream captures S steps of grouped launches and one advance of every counter
void capture(capsgpu_batch& b, const Shape& sh) {
  drop_graphs(b);
  if (sh.mode == 3) {
    ensure_streams(b, 1);
    b.exec.assign(1, nullptr);
    hipStream_t s = b.streams[0];
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      for (int k = 1; k <= sh.steps; ++k) grouped_step(b, s, uint64_t(k));
      step_group_advance(*b.sg, s, uint64_t(sh.steps));
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      drop_graphs(b);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&b.exec[0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      b.exec[0] = nullptr;
      drop_graphs(b);
      HIPCHECK(e);
    }
    b.captured = sh;
    return;
  }
  const auto groups = assign(b, sh.branches);
  ensure_streams(b, sh.branches);
  b.exec.assign(size_t(sh.branches), nullptr);
  for (int j = 0; j < sh.branches; ++j) {
    hipStream_t s = b.streams[size_t(j)];
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      for (int i : groups[size_t(j)]) {
        Rep& R = b.r[size_t(i)];
        for (int k = 1; k <= sh.steps; ++k) one_step(R, s, uint64_t(k));
        const hipStream_t own = R.c->stream;
        R.c->stream = s;
        step_advance(*R.c, R.stepc.p, uint64_t(sh.steps));
        R.c->stream = own;
      }
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      drop_graphs(b);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&b.exec[size_t(j)], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      b.exec[size_t(j)] = nullptr;
      drop_graphs(b);
      HIPCHECK(e);
    }
  }
  b.captured = sh;
}

// steps with shape sh from the current state; the host waits for the device at the end
void advance(capsgpu_batch& b, const Shape& sh, int64_t steps) {
  if (steps <= 0) return;
  HIPCHECK(hipDeviceSynchronize());   // the replicas' own streams are done before other streams touch their state
  if (sh.mode == 2) {
    // direct: per step, every replica's kernels on its own stream (the streams run side by side)
    for (int64_t q = 0; q < steps; ++q)
      for (Rep& R : b.r) {
        one_step(R, R.c->stream, 1);
        step_advance(*R.c, R.stepc.p, 1);
      }
  } else {
    const int64_t reps = steps / sh.steps, rest = steps % sh.steps;
    const int nb = sh.mode == 3 ? 1 : sh.branches;
    for (int64_t q = 0; q < reps; ++q)
      for (int j = 0; j < nb; ++j) HIPCHECK(hipGraphLaunch(b.exec[size_t(j)], b.streams[size_t(j)]));
    HIPCHECK(hipDeviceSynchronize());
    for (int64_t q = 0; q < rest; ++q) {
      if (sh.mode == 3) {
        grouped_step(b, b.streams[0], 1);
        step_group_advance(*b.sg, b.streams[0], 1);
        continue;
      }
      for (Rep& R : b.r) {
        one_step(R, R.c->stream, 1);
        step_advance(*R.c, R.stepc.p, 1);
      }
    }
  }
  HIPCHECK(hipDeviceSynchronize());
}

void snapshot(capsgpu_batch& b, bool restore) {
  for (Rep& R : b.r) {
    capsgpu_ctx& c = *R.c;
    const size_t m = 3 * size_t(c.n);
    if (!restore) R.sx.resize(m), R.sv.resize(m), R.sf.resize(m), R.sstep.resize(1);
    auto cp = [&](double* dst, const double* src) {
      HIPCHECK(hipMemcpyAsync(dst, src, m * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    };
    if (restore) {
      cp(c.x.p, R.sx.p), cp(c.v.p, R.sv.p), cp(c.f.p, R.sf.p);
      HIPCHECK(hipMemcpyAsync(R.stepc.p, R.sstep.p, sizeof(uint64_t), hipMemcpyDeviceToDevice, c.stream));
    } else {
      cp(R.sx.p, c.x.p), cp(R.sv.p, c.v.p), cp(R.sf.p, c.f.p);
      HIPCHECK(hipMemcpyAsync(R.sstep.p, R.stepc.p, sizeof(uint64_t), hipMemcpyDeviceToDevice, c.stream));
    }
  }
  HIPCHECK(hipDeviceSynchronize());
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// the fastest shape by timing: direct launches and graphs of B ∈ {1, 2, 4, …, 32} streams × S ∈ {4, 16, 64} steps (at
// most 1,024 replica steps in one graph), each timed twice over at least 32 steps after one untimed chunk (which also
// captures), the better sample kept, from the snapshot
// per step, the better of two samples over at least 32 steps after one untimed chunk (which also captures); the caller
// saves the state before and restores it after
double time_shape(capsgpu_batch& b, const Shape& sh) {
  const int64_t chunk = sh.mode == 2 ? 1 : sh.steps, timed = std::max<int64_t>(32, chunk);
  if (sh.mode == 1 || sh.mode == 3) capture(b, sh);
  advance(b, sh, chunk);
  double per = 1e300;
  for (int rep = 0; rep < 2; ++rep) {   // one slow replay does not decide the shape
    const auto t0 = std::chrono::steady_clock::now();
    advance(b, sh, timed);
    per = std::min(per, seconds_since(t0) / double(timed));
  }
  return per;
}

Shape tune(capsgpu_batch& b) {
  std::vector<Shape> cand;
  cand.push_back({2, 0, 1});
  const int n = int(b.r.size());
  for (int B : {1, 2, 4, 8, 16, 32})
    if (B <= n)
      for (int S : {4, 16, 64})
        if (int64_t(S) * ((n + B - 1) / B) <= 1024) cand.push_back({1, B, S});
  if (b.pg)
    for (int S : {4, 16, 64}) cand.push_back({3, 1, S});
  snapshot(b, false);
  Shape best;
  double best_t = 1e300;
  for (const Shape& sh : cand) {
    const double per = time_shape(b, sh);
    snapshot(b, true);
    if (per < best_t) best_t = per, best = sh;
  }
  return best;
}

}  // namespace

capsgpu_batch::~capsgpu_batch() {
  (void)hipDeviceSynchronize();
  drop_graphs(*this);
  drop_groups(*this);
  for (hipStream_t s : streams) (void)hipStreamDestroy(s);
}

namespace capsgpu {

capsgpu_batch* batch_create(capsgpu_ctx* const* replicas, int n) {
  if (!replicas || n < 1) throw Error(CAPSGPU_EINVAL, "a batch needs at least one replica");
  auto b = std::make_unique<capsgpu_batch>();
  b->r.resize(size_t(n));
  for (int i = 0; i < n; ++i) {
    if (!replicas[i]) throw Error(CAPSGPU_EINVAL, "a replica context is null");
    for (int k = 0; k < i; ++k)
      if (replicas[k] == replicas[i]) throw Error(CAPSGPU_EINVAL, "a context appears twice in the batch");
    b->r[size_t(i)].c = replicas[i];
  }
  return b.release();
}

void batch_destroy(capsgpu_batch* b) { delete b; }
int batch_size(const capsgpu_batch& b) { return int(b.r.size()); }
capsgpu_ctx* batch_replica(capsgpu_batch& b, int i) { return b.r.at(size_t(i)).c; }

void batch_run(capsgpu_batch& b, const capsgpu_batch_opts& o, const double* temperatures, double* temperature_out, capsgpu_batch_report* rep) {
  const int n = int(b.r.size());
  if (!temperatures) throw Error(CAPSGPU_EINVAL, "null temperatures");
  if (!(o.dt > 0) || o.dt > 5) throw Error(CAPSGPU_EINVAL, "time step must be in (0, 5] fs");
  if (o.steps < 0) throw Error(CAPSGPU_EINVAL, "the number of steps must be 0 or more");
  if (!(o.tau_t > 0)) throw Error(CAPSGPU_EINVAL, "the Langevin time constant must be positive");
  if (o.mode < 0 || o.mode > 3)
    throw Error(CAPSGPU_EINVAL, "mode must be 0 (tuned), 1 (graphs), 2 (direct launches) or 3 (grouped launches)");
  if (o.mode == 3 && (o.graph_steps < 1 || o.graph_steps > 256)) throw Error(CAPSGPU_EINVAL, "grouped launches need 1 … 256 steps per graph");
  if (o.mode == 1 && (o.branches < 1 || o.branches > n || o.graph_steps < 1 || o.graph_steps > 256))
    throw Error(CAPSGPU_EINVAL, "graphs need 1 … replicas branches and 1 … 256 steps per graph");
  for (int i = 0; i < n; ++i) {
    const capsgpu_ctx& c = *b.r[size_t(i)].c;
    if (!(temperatures[i] > 0) || !std::isfinite(temperatures[i])) throw Error(CAPSGPU_EINVAL, "every replica needs a positive temperature");
    if (!c.state_set) throw Error(CAPSGPU_EINVAL, "set the positions of every replica first (capsgpu_set_state)");
    if (c.opt.parts & 2) throw Error(CAPSGPU_EINVAL, "a batch steps replicas without the pair path (energy parts 1, 4 or 5)");
    if (c.n < 2) throw Error(CAPSGPU_EINVAL, "dynamics needs at least two atoms");
  }
  // per replica: the constants of run_dynamics's Langevin branch, velocities, counters and the first forces
  std::vector<double> key = {o.dt, o.tau_t, double(o.seed)};
  for (int i = 0; i < n; ++i) {
    Rep& R = b.r[size_t(i)];
    capsgpu_ctx& c = *R.c;
    hipStream_t s = c.stream;
    const double T = temperatures[i];
    key.push_back(T);
    R.hk = 0.5 * o.dt * kAcc, R.hdt = 0.5 * o.dt;
    R.c1 = std::exp(-o.dt / std::max(o.tau_t, 1e-9)), R.c2 = std::sqrt(1 - R.c1 * R.c1);
    R.seed = o.seed + uint64_t(i);
    std::vector<double> sd(size_t(c.n));
    for (int a = 0; a < c.n; ++a) sd[size_t(a)] = std::sqrt(kBoltz * T * kAcc / c.mass[size_t(a)]);
    R.sdv.upload(sd, s);
    if (o.new_velocities) {
      HostRng rng(R.seed);
      std::vector<double> v(3 * size_t(c.n));
      initial_velocities(rng, c.mass, T, v);
      c.v.upload(v, s);
    }
    const size_t nb = blocks_for(size_t(c.n), 256);
    R.bpart.resize(nb);
    R.bpart.zero(s);
    R.kpart.resize(7 * nb);
    R.ke.resize(7);
    R.stepc.resize(1);
    R.stepc.zero(s);
    c.flags.zero(s);
    c.profile = false;
    double t[CAPSGPU_NTERMS];
    evaluate(c, true, t, false, -1, nullptr);
  }
  HIPCHECK(hipDeviceSynchronize());
  // the launch shape
  // tuned mode, first run with these constants: each replica's transform tiles (on its mesh of the first evaluation,
  // which the next evaluation overwrites), before the group tables take them and the launch shape is timed
  const bool retune = o.mode == 0 && !(b.have_tuned && b.tuned_key == key);
  double tile_s = 0;
  // each replica's kernel shapes as they were set (the transforms' tiles, the gather's workgroups), kept to compare
  std::vector<std::array<int, 3>> set_shapes;
  auto swap_shapes = [&] {
    for (size_t i = 0; i < b.r.size(); ++i) {
      Pme& P = b.r[i].c->pme;
      std::array<int, 3> cur{P.lines_x, P.lines_yz, P.gather_shape};
      P.lines_x = set_shapes[i][0], P.lines_yz = set_shapes[i][1], P.gather_shape = set_shapes[i][2];
      set_shapes[i] = cur;
    }
  };
  if (retune) {
    const auto t0 = std::chrono::steady_clock::now();
    for (Rep& R : b.r) set_shapes.push_back({R.c->pme.lines_x, R.c->pme.lines_yz, R.c->pme.gather_shape});
    for (Rep& R : b.r)
      if (pme_on_ctx(*R.c)) {
        int t[2];
        pme_tune(*R.c, t);
      }
    HIPCHECK(hipDeviceSynchronize());
    tile_s = seconds_since(t0);
  }
  // grouped launches when every replica has PME's reciprocal part with one spline order
  auto build_groups = [&] {
    drop_graphs(b);
    drop_groups(b);
    bool ok = true;
    for (Rep& R : b.r) ok = ok && pme_on_ctx(*R.c) && R.c->pme.counts_clear && R.c->pme.order == b.r[0].c->pme.order;
    if (ok) {
      std::vector<capsgpu_ctx*> cs;
      std::vector<StepGroupInput> in;
      for (Rep& R : b.r) {
        capsgpu_ctx& c = *R.c;
        cs.push_back(&c);
        in.push_back({&c, R.hk, R.hdt, R.c1, R.c2, R.seed, R.stepc.p, R.sdv.p, R.bpart.p, R.kpart.p, (c.opt.parts & 1) != 0, true});
      }
      ensure_streams(b, 1);
      b.pg = pme_group_create(cs, b.streams[0]);
      b.sg = step_group_create(in, b.streams[0]);
      HIPCHECK(hipStreamSynchronize(b.streams[0]));
    } else if (o.mode == 3) {
      throw Error(CAPSGPU_EINVAL, "grouped launches need PME's reciprocal part with one spline order in every replica");
    }
  };
  build_groups();
  Shape sh;
  double tune_s = 0;
  if (o.mode == 2) {
    sh = {2, 0, 1};
  } else if (o.mode == 1) {
    sh = {1, o.branches, o.graph_steps};
  } else if (o.mode == 3) {
    sh = {3, 1, o.graph_steps};
  } else if (b.have_tuned && b.tuned_key == key) {
    sh = b.tuned;
  } else {
    const auto t0 = std::chrono::steady_clock::now();
    sh = tune(b);
    if (sh.mode == 3) {
      // the replicas' shapes were picked each alone; grouped launches may run faster with the shapes they had: the
      // grouped shape is timed with both and the faster kept
      snapshot(b, false);
      const double t_alone = time_shape(b, sh);
      snapshot(b, true);
      swap_shapes();
      build_groups();
      const double t_set = time_shape(b, sh);
      snapshot(b, true);
      if (t_alone < t_set) {
        swap_shapes();
        build_groups();
      }
    }
    tune_s = seconds_since(t0) + tile_s;
    b.have_tuned = true, b.tuned = sh, b.tuned_key = key;
  }
  if (sh.mode == 1 || sh.mode == 3) capture(b, sh);
  const auto t0 = std::chrono::steady_clock::now();
  advance(b, sh, o.steps);
  const double secs = seconds_since(t0);
  // the end: finiteness and FENE flags, and each replica's kinetic temperature
  int64_t atoms = 0;
  for (int i = 0; i < n; ++i) {
    Rep& R = b.r[size_t(i)];
    capsgpu_ctx& c = *R.c;
    atoms += c.n;
    HIPCHECK(hipMemcpyAsync(c.h_flags, c.flags.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    kick(c, 0.0, true, R.kpart.p);
    colsum(c, R.kpart.p, int64_t(blocks_for(size_t(c.n), 256)), 7, R.ke.p);
    double ke = 0;
    HIPCHECK(hipMemcpyAsync(&ke, R.ke.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    if (c.h_flags[3]) throw Error(CAPSGPU_EFIELD, "replica " + std::to_string(i) + " became unstable");
    if (c.h_flags[2]) throw Error(CAPSGPU_EFIELD, "replica " + std::to_string(i) + ": a FENE bond is stretched to R0 or past");
    if (temperature_out) temperature_out[i] = 2 * ke / ((3.0 * c.n - 3.0) * kBoltz);
  }
  if (rep) {
    rep->steps = o.steps;
    rep->seconds = secs;
    rep->atom_steps_per_s = secs > 0 ? double(atoms) * double(o.steps) / secs : 0;
    rep->ns_per_day = secs > 0 ? double(o.steps) * o.dt * 1e-6 / secs * 86400 : 0;
    rep->mode = sh.mode;
    rep->branches = sh.mode == 1 ? sh.branches : sh.mode == 3 ? 1 : n;
    rep->graph_steps = sh.mode == 1 || sh.mode == 3 ? sh.steps : 0;
    rep->tune_seconds = tune_s;
  }
}

}  // namespace capsgpu
