// em3d_matvec_noalloc: em3d_matvec with its FFT workspace in a fixed,
// file-scope device array instead of a caller-provided buffer, for callers
// that cannot allocate device memory (code under stream capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The array `scratch` holds EM3D_NOALLOC_SCRATCH_BYTES per GPU for the life
// of the process.  Host code never passes the array's name to a kernel: its
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and kept per device; that pointer is the workspace
// handed to the same five pass kernels em3d_matvec runs.  The passes write
// every workspace element before reading it, so nothing needs clearing
// between calls.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/em3d_hip.h"
#include "e3_error.hpp"
#include "e3_plan.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)EM3D_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int fail(const std::string& msg) { return e3::set_error(msg); }

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail("em3d_matvec_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("em3d_matvec_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace

extern "C" int em3d_matvec_noalloc(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind,
                                   const void* u, void* y, hipStream_t stream) {
  if (!plan) return fail("em3d_matvec_noalloc: null plan");
  size_t work_bytes = 0;
  if (em3d_plan_sizes(plan, nullptr, &work_bytes, nullptr)) return 1;
  if (work_bytes > kScratchBytes)
    return fail("em3d_matvec_noalloc: the plan needs " + std::to_string(work_bytes) +
                " bytes of FFT workspace, above the fixed scratch of " + std::to_string(kScratchBytes) +
                " bytes; use em3d_matvec with a caller workspace");
  int device = -1;
  const hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return fail(std::string("em3d_matvec_noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (device != plan->device)
    return fail("em3d_matvec_noalloc: the plan belongs to device " + std::to_string(plan->device) +
                ", the current device is " + std::to_string(device));
  void* work = nullptr;
  if (scratch_address(device, &work)) return 1;
  return e3::apply_plan(plan, khat, chi, chi_kind, u, y, work, stream, false, "em3d_matvec_noalloc", nullptr);
}
