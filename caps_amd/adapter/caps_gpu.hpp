// The reference side of the drop-in: CAPS's own types (caps::ForceField, caps::EnergyOptions, caps::Cell, caps::System,
// caps::DynamicsOptions) turned into the capsgpu C ABI and back. Compiled inside CAPS's core (core/include on the include
// path, libcapsgpu.so linked); nothing in it is used by the Python layer or the kernels.
//
//   caps::gpu::Evaluator    stands in for caps::Evaluator (core/include/caps/field.hpp:327-352): same constructor
//                           arguments, compute(x, cell, f) -> EnergyTerms, energies(on), list_builds(); the error classes
//                           rethrown as the reference's (FieldError, std::invalid_argument, DynamicsCancelled)
//   caps::gpu::run_dynamics stands in for caps::run_dynamics (core/include/caps/dynamics.hpp:180) for the options this
//                           path runs: thermostats None / Bussi / Langevin, barostats Berendsen / CRescale (isotropic),
//                           r-RESPA, temperature ramps, checkpoints, progress with cancellation
// Force-field terms or options outside the path are refused with std::invalid_argument naming them, never dropped.
#pragma once

#include <cstdint>
#include <vector>

#include "caps/dynamics.hpp"
#include "caps/field.hpp"
#include "caps/system.hpp"
#include "capsgpu.h"

namespace caps {
namespace gpu {

// The force field as the C ABI reads it: the arrays capsgpu_ff points into, owned here.
struct FlatForceField {
  std::vector<int32_t> type, excl_offset, excl_index, p14_offset, p14_index, bond_atoms, angle_atoms, fene_atoms;
  std::vector<double> charge, mass, eps, sigma, bond_param, angle_param, fene_param;
  capsgpu_ff c{};
};

// Flattens ff (every term this path evaluates; anything else raises) with the model's own settings applied as
// caps::Evaluator applies them: charges over √εr, the mixed ε, σ of every type pair (caps::mixed_pair), the exclusions
// and their 1-4 pairs.
void flatten(const ForceField& ff, FlatForceField& out);

// EnergyOptions as the C ABI reads them; raises for the options this path does not run (slab PME, force caps,
// alchemical pairs). lj_shift / the model's own cut-off come from the force field, as in caps::Evaluator.
capsgpu_energy_opts energy_options(const ForceField& ff, const EnergyOptions& o);

// The reference's cell as the C ABI's twelve numbers (origin, a, b, c).
void cell12(const Cell& c, double out[12]);

class Evaluator {
 public:
  Evaluator(const ForceField& ff, const EnergyOptions& o);
  ~Evaluator();
  Evaluator(const Evaluator&) = delete;
  Evaluator& operator=(const Evaluator&) = delete;
  // x: 3N positions (Å). f: 3N forces out (kcal/mol/Å). Returns the energy terms.
  EnergyTerms compute(const std::vector<double>& x, const Cell& cell, std::vector<double>& f);
  void energies(bool on) { energies_ = on; }
  int list_builds() const;
  capsgpu_ctx* context() const { return ctx_; }

 private:
  FlatForceField ff_;
  capsgpu_ctx* ctx_ = nullptr;
  bool energies_ = true;
};

// run_dynamics on the GPU: the state leaves the host once, the steps run on the device, the thermo rows, checkpoints and
// progress come back through callbacks, and s's positions, velocities and cell are updated at the end.
void run_dynamics(System& s, const DynamicsOptions& o, DynamicsReport* report = nullptr);

}  // namespace gpu
}  // namespace caps
