// Internal: launchers of the integrator and reduction kernels (step.hip).
#pragma once

#include "context.hpp"
#include "pme_atoms.hpp"

namespace capsgpu {
// Half kicks put off from the end of a step to the next launch that moves the atoms (the forces they use are unchanged
// until then): that launch applies them first, in order, with the arithmetic of separate kicks, so the velocities are
// the same to the last bit. Nothing may read the velocities while kicks are put off.
struct Kicks {
  static constexpr int kMax = 3;
  int n = 0;
  double hk[kMax] = {0, 0, 0};
  const double* f[kMax] = {nullptr, nullptr, nullptr};
  void put(double h, const double* force) {
    if (n >= kMax) throw Error(-1, "too many half kicks put off");
    hk[n] = h, f[n] = force, ++n;
  }
};
void slot_virial(capsgpu_ctx& c, double* part);
void colsum(capsgpu_ctx& c, const double* part, int64_t rows, int cols, double* out);
// f: the force to kick with (nullptr: the context's c.f); sp: the next PME evaluation's per-atom bookkeeping, done on the
// new positions here (from pme_presplines), or nullptr
void vv_open(capsgpu_ctx& c, double hk, double dt, const double* alpha, double lim_b, double lim_p, const double* f = nullptr,
             const Kicks& pre = Kicks{}, const AtomSplines* sp = nullptr);
// the Langevin step's noise is drawn for step (stepc ? *stepc : 0) + step_add; part: per-workgroup rows the bath's heat
// is added to (bath_add takes them)
void baoab_open(capsgpu_ctx& c, double hk, double hdt, double c1, double c2, const double* sdv, uint64_t seed, const uint64_t* stepc,
                uint64_t step_add, double lim_b, double lim_p, double* part, const Kicks& pre = Kicks{}, const AtomSplines* sp = nullptr);
// *stepc += by on the stream
void step_advance(capsgpu_ctx& c, uint64_t* stepc, uint64_t by);
void kick(capsgpu_ctx& c, double hk, bool ke, double* part, const double* f = nullptr, const Kicks& pre = Kicks{});
void bussi(capsgpu_ctx& c, const double* part, double kt_target, double ndof, double cc, double r1, double sum, double* alpha, double* bath);
void scale_v(capsgpu_ctx& c, const double* alpha);
// positions scaled by mu about origin, velocities by vf (the barostat's isotropic update)
void scale_xv(capsgpu_ctx& c, const double origin[3], double mu, double vf);
// *bath −= Σ part over the rows of c's atom blocks; the rows are cleared
void bath_add(capsgpu_ctx& c, double* part, double* bath);
// many contexts' Langevin step, one launch per kernel: per replica its constants and device arrays
struct StepGroupInput {
  capsgpu_ctx* c;
  double hk, hdt, c1, c2;
  uint64_t seed;
  uint64_t* stepc;
  const double* sdv;
  double *bpart, *kpart;
  bool bonded, recip;
};
struct StepGroup;
StepGroup* step_group_create(const std::vector<StepGroupInput>& in, hipStream_t s);
void step_group_destroy(StepGroup* g);
// BAOAB's opening B, A, O, A (the noise for step *stepc + step_add of each replica)
void step_group_open(StepGroup& g, hipStream_t s, uint64_t step_add);
// the bonded terms, each atom's force sum (reciprocal + bonded) and the closing kick
void step_group_close(StepGroup& g, hipStream_t s);
void step_group_advance(StepGroup& g, hipStream_t s, uint64_t by);
}  // namespace capsgpu
