// Fast x, y and forward z pass instantiations, double precision, with non-temporal row
// loads and stores (plans whose working set exceeds the last-level caches).
// A translation unit of its own, compiled beside the plain instantiations.
#include "e3_fast_launch.hpp"

namespace e3 {

bool launch_fast_s_f64_nt(const SArgs& a, hipStream_t st, hipError_t* err) {
  HipFastRun r{st};
  const bool ok = fast_s<double, true>(r, a);
  *err = r.err;
  return ok;
}

bool launch_fast_x_f64_nt(const XArgs& a, hipStream_t st, hipError_t* err) {
  HipFastRun r{st};
  const bool ok = fast_x<double, true>(r, a);
  *err = r.err;
  return ok;
}

bool launch_fast_z_f64_nt(const ZArgs& a, hipStream_t st, hipError_t* err) {
  HipFastRun r{st};
  const bool ok = fast_z<double, true>(r, a);
  *err = r.err;
  return ok;
}

}  // namespace e3
