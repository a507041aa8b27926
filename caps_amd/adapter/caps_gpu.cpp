// The drop-in's reference side: caps::ForceField / EnergyOptions / Cell / System / DynamicsOptions into the capsgpu C ABI
// (include/capsgpu.h) and back. Compiled inside CAPS's core with core/include and this repository's include/ on the
// include path, linked with libcapsgpu.so.
#include "caps_gpu.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <stdexcept>
#include <string>

namespace caps {
namespace gpu {
namespace {

// capsgpu's error classes as the reference's exceptions (capi/src/caps_c.cpp:373-383 maps them the other way)
void check(int32_t rc) {
  if (rc == CAPSGPU_OK) return;
  const std::string m = capsgpu_last_error();
  if (rc == CAPSGPU_EFIELD) throw FieldError(m);
  if (rc == CAPSGPU_ECANCEL) throw DynamicsCancelled();
  if (rc == CAPSGPU_EINVAL) throw std::invalid_argument(m);
  throw std::runtime_error("GPU: " + m);
}

void refuse_if(bool has, const char* what) {
  if (has) throw std::invalid_argument(std::string(what) + " are outside the GPU path (harmonic bonds and angles, FENE, LJ 12-6/9-6, DSF or PME)");
}

// one thermo row of the C ABI as the reference's ThermoRow
ThermoRow thermo_row(const capsgpu_thermo& r, double target) {
  ThermoRow t;
  t.step = r.step, t.time_ps = r.time_ps, t.temperature = r.temperature;
  t.potential = r.potential, t.kinetic = r.kinetic, t.total = r.total, t.conserved = r.conserved, t.pressure = r.pressure;
  t.target_temperature = target;
  t.volume = r.volume, t.density = r.density;
  for (int q = 0; q < 6; ++q) t.p[q] = r.p[q];
  return t;
}

struct RunState {
  const DynamicsOptions* o;
  DynamicsReport* rep;
  std::vector<ThermoRow> rows;
  std::exception_ptr failure;
};

int32_t progress_trampoline(const capsgpu_thermo* row, void* user) {
  auto* st = static_cast<RunState*>(user);
  const DynamicsOptions& o = *st->o;
  const double frac = o.steps > 0 ? double(row->step - o.step_offset) / double(o.steps) : 0.0;
  const double target = o.temperature_end >= 0 ? o.temperature + (o.temperature_end - o.temperature) * frac : o.temperature;
  st->rows.push_back(thermo_row(*row, target));
  if (!o.progress) return 1;
  try {
    return o.progress(st->rows.back()) ? 1 : 0;
  } catch (...) {
    st->failure = std::current_exception();
    return 0;
  }
}

int32_t checkpoint_trampoline(const double* x, const double* v, const double c[12], int64_t step, void* user) {
  auto* st = static_cast<RunState*>(user);
  try {
    const size_t n3 = st->o->field ? 3 * st->o->field->mass.size() : 0;
    const std::vector<double> xs(x, x + n3), vs(v, v + n3);
    Cell cell;
    for (int k = 0; k < 3; ++k) cell.origin[k] = c[k], cell.a[k] = c[3 + k], cell.b[k] = c[6 + k], cell.c[k] = c[9 + k];
    st->o->checkpoint(xs, vs, cell, step);
    return 0;
  } catch (...) {
    st->failure = std::current_exception();
    return 1;
  }
}

}  // namespace

void flatten(const ForceField& ff, FlatForceField& out) {
  // the terms this path does not evaluate: refused, never dropped
  refuse_if(!ff.dihedrals.empty() || !ff.dihedrals2.empty() || !ff.cbt.empty(), "dihedral terms");
  refuse_if(!ff.impropers.empty() || !ff.impropers_harmonic.empty() || !ff.inversions.empty() || !ff.impropers2.empty(),
            "improper terms");
  refuse_if(!ff.angles_x.empty() || !ff.angles2.empty() || !ff.urey_bradley.empty() || !ff.bonds2.empty(), "class II and other angle forms");
  refuse_if(!ff.lj_pairs.empty() || !ff.pair_func.empty() || !ff.tables.empty() || !ff.lj14_types.empty() || ff.per_pair14(),
            "explicit, tabulated and per-pair 1-4 pair terms");
  refuse_if(ff.sw.on || ff.swap3.on() || ff.manybody.on() || ff.hbond.on() || !ff.vsites.empty(), "many-body terms, hydrogen bonds and virtual sites");
  refuse_if(ff.coul_gromacs || ff.coul_rf || ff.lj_fsw || ff.keep13 || !ff.excluded_type_pairs.empty(),
            "switched, reaction-field and type-pair-excluded interactions");
  refuse_if(!ff.pairs14.empty() && ff.lj14 != 0.0, "1-4 Lennard-Jones pairs");
  if (ff.pair_form != "lj12-6" && ff.pair_form != "lj9-6") refuse_if(true, "pair forms other than lj12-6 and lj9-6");
  if (ff.dielectric <= 0) throw FieldError("the relative permittivity must be positive");
  const size_t n = ff.mass.size();
  const int nt = int(ff.lj.size());
  out = FlatForceField{};
  out.type.assign(ff.type_index.begin(), ff.type_index.end());
  out.mass = ff.mass;
  out.charge = ff.charge;
  if (ff.dielectric != 1)
    for (double& q : out.charge) q /= std::sqrt(ff.dielectric);
  out.eps.resize(size_t(nt) * nt), out.sigma.resize(size_t(nt) * nt);
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const PairType p = mixed_pair(ff, a, b);
      out.eps[size_t(a) * nt + b] = p.eps, out.sigma[size_t(a) * nt + b] = p.sigma;
    }
  // exclusions (sorted per atom) and, of them, the 1-4 pairs whose Coulomb keeps coul14
  out.excl_offset.assign(n + 1, 0);
  std::vector<std::vector<int32_t>> p14(n);
  for (const auto& pr : ff.pairs14) p14[pr[0]].push_back(int32_t(pr[1])), p14[pr[1]].push_back(int32_t(pr[0]));
  out.p14_offset.assign(n + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    if (i < ff.excluded.size())
      for (uint32_t j : ff.excluded[i]) out.excl_index.push_back(int32_t(j));
    out.excl_offset[i + 1] = int32_t(out.excl_index.size());
    std::sort(p14[i].begin(), p14[i].end());
    out.p14_index.insert(out.p14_index.end(), p14[i].begin(), p14[i].end());
    out.p14_offset[i + 1] = int32_t(out.p14_index.size());
  }
  for (const auto& b : ff.bonds) {
    out.bond_atoms.push_back(int32_t(b.i)), out.bond_atoms.push_back(int32_t(b.j));
    out.bond_param.push_back(b.k), out.bond_param.push_back(b.r0);
  }
  for (const auto& a : ff.angles) {
    out.angle_atoms.insert(out.angle_atoms.end(), {int32_t(a.i), int32_t(a.j), int32_t(a.k)});
    out.angle_param.push_back(a.kt), out.angle_param.push_back(a.theta0);
  }
  for (const auto& b : ff.bonds_x) {
    if (b.form != 3) refuse_if(true, "bond forms other than harmonic and FENE (BondX form 3)");
    out.fene_atoms.push_back(int32_t(b.i)), out.fene_atoms.push_back(int32_t(b.j));
    out.fene_param.insert(out.fene_param.end(), {b.a, b.b, b.c, b.d});
  }
  // empty arrays still need a valid pointer
  for (auto* v : {&out.excl_index, &out.p14_index, &out.bond_atoms, &out.angle_atoms, &out.fene_atoms})
    if (v->empty()) v->push_back(0);
  for (auto* v : {&out.bond_param, &out.angle_param, &out.fene_param})
    if (v->empty()) v->push_back(0.0);
  capsgpu_ff& c = out.c;
  c.n_atoms = int32_t(n), c.n_types = nt;
  c.type = out.type.data(), c.charge = out.charge.data(), c.mass = out.mass.data();
  c.eps = out.eps.data(), c.sigma = out.sigma.data(), c.lj96 = ff.pair_form == "lj9-6" ? 1 : 0;
  c.excl_offset = out.excl_offset.data(), c.excl_index = out.excl_index.data();
  c.p14_offset = out.p14_offset.data(), c.p14_index = out.p14_index.data(), c.coul14 = ff.coul14;
  c.n_bonds = int32_t(ff.bonds.size()), c.bond_atoms = out.bond_atoms.data(), c.bond_param = out.bond_param.data();
  c.n_angles = int32_t(ff.angles.size()), c.angle_atoms = out.angle_atoms.data(), c.angle_param = out.angle_param.data();
  c.n_fene = int32_t(ff.bonds_x.size()), c.fene_atoms = out.fene_atoms.data(), c.fene_param = out.fene_param.data();
}

capsgpu_energy_opts energy_options(const ForceField& ff, const EnergyOptions& o) {
  refuse_if(o.slab > 0, "slab PME corrections");
  refuse_if(o.force_cap > 0, "LJ force caps (push-off)");
  refuse_if(o.alchemical != nullptr, "alchemical pairs");
  capsgpu_energy_opts e{};
  e.cutoff = ff.cutoff > 0 ? ff.cutoff : o.cutoff;   // the model's own cut-off (field.cpp:438)
  e.skin = o.skin;
  e.coulomb = o.coulomb ? 1 : 0;
  e.electrostatics = o.electrostatics == EnergyOptions::Electrostatics::PME ? 1 : 0;
  e.dsf_alpha = o.dsf_alpha;
  e.ewald_rtol = o.ewald_rtol;
  e.pme_spacing = o.pme_spacing;
  e.pme_order = o.pme_order;
  e.tail = (o.tail && !ff.lj_shift) ? 1 : 0;         // shifted at the cut-off without tail terms (field.cpp:439)
  e.parts = o.parts;
  e.no_reciprocal = 0;
  return e;
}

void cell12(const Cell& c, double out[12]) {
  for (int k = 0; k < 3; ++k) out[k] = c.origin[k], out[3 + k] = c.a[k], out[6 + k] = c.b[k], out[9 + k] = c.c[k];
}

Evaluator::Evaluator(const ForceField& ff, const EnergyOptions& o) {
  flatten(ff, ff_);
  const capsgpu_energy_opts e = energy_options(ff, o);
  check(capsgpu_create(&ff_.c, &e, &ctx_));
}

Evaluator::~Evaluator() { capsgpu_destroy(ctx_); }

EnergyTerms Evaluator::compute(const std::vector<double>& x, const Cell& cell, std::vector<double>& f) {
  double c[12];
  cell12(cell, c);
  f.assign(x.size(), 0.0);
  double t[CAPSGPU_NTERMS];
  check(capsgpu_compute(ctx_, x.data(), c, f.data(), t, energies_ ? 1 : 0));
  EnergyTerms e;
  e.bond = t[CAPSGPU_T_BOND], e.angle = t[CAPSGPU_T_ANGLE], e.dihedral = t[CAPSGPU_T_DIHEDRAL], e.improper = t[CAPSGPU_T_IMPROPER];
  e.vdw = t[CAPSGPU_T_VDW], e.coulomb = t[CAPSGPU_T_COULOMB], e.virial = t[CAPSGPU_T_VIRIAL];
  for (int q = 0; q < 6; ++q) e.w[q] = t[CAPSGPU_T_WXX + q];
  e.tail_energy = t[CAPSGPU_T_TAIL_ENERGY], e.tail_virial = t[CAPSGPU_T_TAIL_VIRIAL];
  return e;
}

int Evaluator::list_builds() const {
  int32_t b = 0;
  check(capsgpu_list_builds(ctx_, &b));
  return b;
}

void run_dynamics(System& s, const DynamicsOptions& o, DynamicsReport* report) {
  if (!o.field) throw std::invalid_argument("run_dynamics needs DynamicsOptions::field");
  refuse_if(o.thermostat == Thermostat::NoseHoover, "Nosé–Hoover thermostats");
  refuse_if(o.barostat == Barostat::MTK || o.anisotropic || o.full_shape || o.deform_axis >= 0, "MTK, anisotropic and deforming barostats");
  refuse_if(o.constraints != ConstraintMode::None || o.hmr > 0, "constraints and mass repartitioning");
  refuse_if(!o.fixed.empty() || !o.pull_group.empty() || !o.move_group.empty() || !o.restraints.empty(), "held, pulled and restrained atoms");
  refuse_if(o.efield[0] != 0 || o.efield[1] != 0 || o.efield[2] != 0 || o.shear_rate != 0, "external fields and shear");
  refuse_if(bool(o.topology) || bool(o.each_step) || bool(o.after_step), "per-step host callbacks (topology, each_step, after_step)");
  const size_t n = s.atoms.size();
  if (o.field->mass.size() != n) throw std::invalid_argument("the force field is for another number of atoms");
  Evaluator ev(*o.field, o.energy);
  std::vector<double> x(3 * n), v;
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) x[3 * i + k] = s.atoms[i].pos[k];
  const bool have_v = s.velocities.size() == n && !o.new_velocities;
  if (have_v) {
    v.resize(3 * n);
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] = s.velocities[i][k];
  }
  double c[12];
  cell12(s.cell, c);
  check(capsgpu_set_state(ev.context(), x.data(), have_v ? v.data() : nullptr, c));
  capsgpu_dyn_opts d{};
  d.dt = o.dt, d.steps = o.steps, d.temperature = o.temperature;
  d.thermostat = o.thermostat == Thermostat::Bussi ? 1 : o.thermostat == Thermostat::Langevin ? 2 : 0;
  d.tau_t = o.tau_t, d.seed = o.seed, d.thermo_every = o.thermo_every, d.new_velocities = have_v ? 0 : 1;
  d.respa = o.respa, d.step_offset = o.step_offset;
  d.barostat = o.barostat == Barostat::Berendsen ? 1 : o.barostat == Barostat::CRescale ? 2 : 0;
  d.pressure = o.pressure, d.tau_p = o.tau_p, d.compressibility = o.compressibility, d.barostat_every = o.barostat_every;
  d.temperature_end = o.temperature_end >= 0 ? o.temperature_end : -1.0;
  RunState st{&o, report, {}, nullptr};
  if (o.checkpoint && o.checkpoint_every > 0) {
    d.checkpoint_every = o.checkpoint_every;
    d.checkpoint = checkpoint_trampoline;
    d.checkpoint_user = &st;
  }
  capsgpu_dyn_report r{};
  const int32_t rc = capsgpu_run_dynamics(ev.context(), &d, progress_trampoline, &st, nullptr, 0, nullptr, &r);
  if (st.failure) std::rethrow_exception(st.failure);
  check(rc);
  v.resize(3 * n);
  check(capsgpu_get_state(ev.context(), x.data(), v.data(), c));
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) s.atoms[i].pos[k] = x[3 * i + k];
  s.velocities.resize(n);
  for (size_t i = 0; i < n; ++i) s.velocities[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  for (int k = 0; k < 3; ++k) s.cell.origin[k] = c[k], s.cell.a[k] = c[3 + k], s.cell.b[k] = c[6 + k], s.cell.c[k] = c[9 + k];
  s.unwrapped = true;
  if (report) {
    report->thermo = std::move(st.rows);
    report->steps = r.steps, report->seconds = r.seconds, report->ns_per_day = r.ns_per_day, report->list_builds = r.list_builds;
    report->threads = 0;
  }
}

}  // namespace gpu
}  // namespace caps
