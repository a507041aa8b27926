// Internal: replicas stepped together (batch.hip); the C entry points in api.hip wrap these.
#pragma once

#include "../../include/capsgpu.h"

namespace capsgpu {
capsgpu_batch* batch_create(capsgpu_ctx* const* replicas, int n);
void batch_destroy(capsgpu_batch* b);
void batch_run(capsgpu_batch& b, const capsgpu_batch_opts& o, const double* temperatures, double* temperature_out, capsgpu_batch_report* rep);
int batch_size(const capsgpu_batch& b);
capsgpu_ctx* batch_replica(capsgpu_batch& b, int i);
}  // namespace capsgpu
