// Optional per-kernel timing: when enabled (fgs_profile_enable), every
// launch wrapped in FGS_TIMED records a pair of HIP events on the stream it
// runs on; fgs_profile_read aggregates the elapsed times per name. Meant for
// single-threaded measurement runs (bench.py); off by default, in which case
// FGS_TIMED costs one branch.
#pragma once
#include <hip/hip_runtime.h>

namespace fgs {

bool prof_on();
hipEvent_t prof_begin(hipStream_t s);
void prof_end(const char* name, hipEvent_t start, hipStream_t s);

}  // namespace fgs

#define FGS_TIMED(name, stream, launch)                          \
  do {                                                           \
    hipEvent_t fgs_ev0__ = nullptr;                              \
    if (::fgs::prof_on()) fgs_ev0__ = ::fgs::prof_begin(stream); \
    launch;                                                      \
    if (fgs_ev0__) ::fgs::prof_end(name, fgs_ev0__, stream);     \
  } while (0)
