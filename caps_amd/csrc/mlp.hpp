// Internal: learned r-RESPA (mlp.hip) — no counterpart in the reference.
#pragma once

#include "context.hpp"

namespace capsgpu {
struct Learned;
Learned* learned_begin(capsgpu_ctx& c, const capsgpu_dyn_opts& o);
// one step's forces: 1 predicted (MLP), 0 a full PME step, 2 a full step that switched to plain PME (fallback)
int learned_step(capsgpu_ctx& c, Learned& L, int64_t step, bool want, double* terms);
double learned_rmse(const Learned& L);
void learned_end(capsgpu_ctx& c, Learned* L);
}  // namespace capsgpu
