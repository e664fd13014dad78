// Fast pass instantiations, float precision.
#include "e3_fast_launch.hpp"

namespace e3 {

bool launch_fast_s_f32(const SArgs& a, hipStream_t st, hipError_t* err) {
  if (a.nt) return launch_fast_s_f32_nt(a, st, err);
  HipFastRun r{st};
  const bool ok = fast_s<float>(r, a);
  *err = r.err;
  return ok;
}

bool launch_fast_x_f32(const XArgs& a, hipStream_t st, hipError_t* err) {
  if (a.nt) return launch_fast_x_f32_nt(a, st, err);
  HipFastRun r{st};
  const bool ok = fast_x<float>(r, a);
  *err = r.err;
  return ok;
}

bool launch_fast_z_f32(const ZArgs& a, hipStream_t st, hipError_t* err) {
  if (a.nt) return launch_fast_z_f32_nt(a, st, err);
  HipFastRun r{st};
  const bool ok = fast_z<float>(r, a);
  *err = r.err;
  return ok;
}

}  // namespace e3
