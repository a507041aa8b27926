// Internal: per-kernel launch timing while a context is profiling (events on its stream around each launch).
#pragma once

#include "context.hpp"

namespace capsgpu {

// sums the recorded launches per kernel name (waits for the stream)
inline void kt_harvest(capsgpu_ctx& c) {
  if (!c.kev_used) return;
  HIPCHECK(hipStreamSynchronize(c.stream));
  for (size_t q = 0; q < c.kev_used; ++q) {
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, c.kev[q].a, c.kev[q].b));
    auto it = c.ktot.begin();
    for (; it != c.ktot.end(); ++it)
      if (it->first == c.kev[q].name) break;
    if (it == c.ktot.end()) {
      c.ktot.push_back({c.kev[q].name, {0.0, 0}});
      it = c.ktot.end() - 1;
    }
    it->second.first += ms;
    ++it->second.second;
  }
  c.kev_used = 0;
}

// before a launch: an event pair for it when profiling
inline void kt_begin(capsgpu_ctx& c, const char* name) {
  if (!c.profile || (!c.kfilter.empty() && c.kfilter != name)) return;
  if (c.kev_used == c.kev.size()) {
    if (c.kev.size() >= 8192) {
      kt_harvest(c);
    } else {
      capsgpu_ctx::KTime k{name, nullptr, nullptr};
      HIPCHECK(hipEventCreate(&k.a));
      HIPCHECK(hipEventCreate(&k.b));
      c.kev.push_back(k);
    }
  }
  c.kev[c.kev_used].name = name;
  HIPCHECK(hipEventRecord(c.kev[c.kev_used].a, c.stream));
  c.kt_open = true;
}

// after the launch
inline void kt_end(capsgpu_ctx& c) {
  if (!c.profile || !c.kt_open) return;
  HIPCHECK(hipEventRecord(c.kev[c.kev_used].b, c.stream));
  ++c.kev_used;
  c.kt_open = false;
}

}  // namespace capsgpu
