// Internal: launchers of the integrator and reduction kernels (step.hip).
#pragma once

#include "context.hpp"

namespace capsgpu {
void slot_virial(capsgpu_ctx& c, double* part);
void colsum(capsgpu_ctx& c, const double* part, int64_t rows, int cols, double* out);
// f: the force to kick with (nullptr: the context's c.f)
void vv_open(capsgpu_ctx& c, double hk, double dt, const double* alpha, double lim_b, double lim_p, const double* f = nullptr);
void baoab_open(capsgpu_ctx& c, double hk, double hdt, double c1, double c2, const double* sdv, uint64_t seed, uint64_t step, double lim_b,
                double lim_p, double* part);
void kick(capsgpu_ctx& c, double hk, bool ke, double* part, const double* f = nullptr);
void bussi(capsgpu_ctx& c, const double* part, double kt_target, double ndof, double cc, double r1, double sum, double* alpha, double* bath);
void scale_v(capsgpu_ctx& c, const double* alpha);
void bath_add(capsgpu_ctx& c, const double* part, double* bath);
}  // namespace capsgpu
