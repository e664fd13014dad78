// The plan object behind em3d_plan_t and the internal apply entry shared by
// the translation units of libem3d_hip.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/em3d_hip.h"
#include "e3_pipeline.hpp"

namespace e3 {
constexpr int kMaxTimedLaunches = 8;
}

struct em3d_plan_s {
  e3::Plan p;
  void* tw_dev = nullptr;
  int device = 0;
  // per-pass timing (em3d_plan_set_timing): events around the launches of
  // each apply call and the accumulated elapsed milliseconds per launch slot
  bool timing = false;
  hipEvent_t ev[e3::kMaxTimedLaunches + 1] = {};
  double ms[e3::kMaxTimedLaunches] = {};
  long long calls = 0;
  int nlaunch = 0;
};


namespace e3 {

// y = A u (or A^H u) on the plan's stream; when skip is non-null every kernel
// returns at once while *skip is non-zero (device-side early exit).  k1 / k5:
// fused Krylov steps of the first / last z pass; skip1: P1's flag when given;
// passes: which of P1..P5 to run (PASSES_ALL, PASSES_FIRST, PASSES_REST);
// win: the input box of the forward passes (run_apply), validated here.
int apply_plan(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
               void* work, hipStream_t stream, bool adjoint, const char* name, const int* skip,
               const ZKry* k1 = nullptr, const ZKry* k5 = nullptr, const int* skip1 = nullptr,
               unsigned passes = PASSES_ALL, const int* win = nullptr);

}  // namespace e3
