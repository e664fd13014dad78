// Fast pass instantiations, double precision.
#include "e3_fast_launch.hpp"

namespace e3 {

bool launch_fast_s_f64(const SArgs& a, hipStream_t st, hipError_t* err) {
  if (a.nt) return launch_fast_s_f64_nt(a, st, err);
  HipFastRun r{st};
  const bool ok = fast_s<double>(r, a);
  *err = r.err;
  return ok;
}

bool launch_fast_x_f64(const XArgs& a, hipStream_t st, hipError_t* err) {
  if (a.nt) return launch_fast_x_f64_nt(a, st, err);
  HipFastRun r{st};
  const bool ok = fast_x<double>(r, a);
  *err = r.err;
  return ok;
}

bool launch_fast_z_f64(const ZArgs& a, hipStream_t st, hipError_t* err) {
  if (a.nt) return launch_fast_z_f64_nt(a, st, err);
  HipFastRun r{st};
  const bool ok = fast_z<double>(r, a);
  *err = r.err;
  return ok;
}

}  // namespace e3
