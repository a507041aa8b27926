// Internal: learned r-RESPA (mlp.hip) — no counterpart in the reference.
#pragma once

#include "context.hpp"

namespace capsgpu {
struct Learned;
Learned* learned_begin(capsgpu_ctx& c, const capsgpu_dyn_opts& o);
// one step's forces: 1 predicted (MLP), 0 a full PME step, 2 a full step that switched to plain PME (fallback)
int learned_step(capsgpu_ctx& c, Learned& L, int64_t step, bool want, double* terms);
double learned_rmse(const Learned& L);
double learned_rmse_stale(const Learned& L);
// this run's judgments and the judgments that missed the tolerance (each one a fallback to plain PME)
int64_t learned_judgments(const Learned& L);
int64_t learned_failures(const Learned& L);
void learned_end(capsgpu_ctx& c, Learned* L);
void learned_free(Learned* L);
// inspection for the parity checks (parameters in the layout w1 [128][96], b1, w2 [128][128], b2, w3 [3][128], b3)
void learned_forward_host(capsgpu_ctx& c, const float* params, const float* x, int m, float* y);
double learned_gradient_host(capsgpu_ctx& c, const float* params, const float* x, const float* t, int m, float* grad);
void learned_descriptors_host(capsgpu_ctx& c, float* raw);
}  // namespace capsgpu
