// Internal interface of the FNO preconditioner (e3_fno.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "em3d_hip.h"

namespace e3 {

constexpr int kFnoMaxLayers = 16;

// offsets (in floats) of the parameters and of the work buffer
struct FnoLayout {
  long long cells, modes3;
  long long ldx;   // row stride of the input features: residual then material channels
  long long w_in, w[kFnoMaxLayers], u[kFnoMaxLayers], d[kFnoMaxLayers], r[kFnoMaxLayers], w_out, n_params;
  long long act[kFnoMaxLayers + 1], zpre[kFnoMaxLayers], gate[kFnoMaxLayers], a1, a2, a3[kFnoMaxLayers], b3, spec,
      d0, d1, part, tw;
  long long work_floats;
};

std::string fno_check(const em3d_fno_config* c);
FnoLayout fno_layout(const em3d_fno_config& c, int samples);
int fno_precond_apply(const em3d_fno_model* m, int batch, int ncomp, int nx, int ny, int nz, int precision,
                      const void* r, void* out, const int* skip, hipStream_t stream);

}  // namespace e3
