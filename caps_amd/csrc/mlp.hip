// Learned r-RESPA (placeholder until the MLP kernels land).
#include "mlp.hpp"

namespace capsgpu {
struct Learned {
  double rmse = 0;
};
Learned* learned_begin(capsgpu_ctx&, const capsgpu_dyn_opts&) { throw Error(-1, "learned r-RESPA is not available in this build"); }
int learned_step(capsgpu_ctx&, Learned&, int64_t, bool, double*) { return 0; }
double learned_rmse(const Learned& L) { return L.rmse; }
void learned_end(capsgpu_ctx&, Learned* L) { delete L; }
}  // namespace capsgpu
