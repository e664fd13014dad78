// Support of the contrast and the sub-grid it lies in.
//
// A = I - B chi only couples cells where chi != 0, so the operator restricted
// to a box that holds supp(chi) is the same operator on a smaller grid with
// the same cell size.  These kernels find that box and move fields between
// the full grid and the box: a gather (crop) of a full field into the box,
// and a scatter that writes the box values inside the box and either zeros
// or f - y outside it.  All of them are HBM-bound streaming kernels (one read
// and one write per element, coalesced along z).
#include <hip/hip_runtime.h>

#include <climits>
#include <string>

#include "../../include/em3d_hip.h"
#include "e3_common.hpp"
#include "e3_error.hpp"

using namespace e3;

namespace {

int sfail(const std::string& m) { return set_error(m); }

constexpr int kBoundThreads = 256;

__global__ void k_bounds_init(int nx, int ny, int nz, int* b) {
  if (threadIdx.x == 0) {
    b[0] = nx; b[1] = -1;
    b[2] = ny; b[3] = -1;
    b[4] = nz; b[5] = -1;
  }
}

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// b = (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi) of the cells where any entry of
// the packed contrast is non-zero.  One wavefront reduction per axis, then
// one atomic per wavefront and bound.
template <typename T>
__global__ __launch_bounds__(kBoundThreads) void k_bounds(int nx, int ny, int nz, const cpx<T>* chi, int nent,
                                                          int* b) {
  const long long m = (long long)nx * ny * nz;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {-1, -1, -1};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    bool nz_cell = false;
    for (int e = 0; e < nent; ++e) {
      const cpx<T> v = chi[(size_t)e * m + i];
      nz_cell |= (v.re != T(0)) || (v.im != T(0));
    }
    if (nz_cell) {
      const int z = (int)(i % nz);
      const long long xy = i / nz;
      const int y = (int)(xy % ny), x = (int)(xy / ny);
      lo[0] = min(lo[0], x); hi[0] = max(hi[0], x);
      lo[1] = min(lo[1], y); hi[1] = max(hi[1], y);
      lo[2] = min(lo[2], z); hi[2] = max(hi[2], z);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int l = wave_min(lo[a]), h = wave_max(hi[a]);
    if ((threadIdx.x & 63) == 0 && h >= 0) {
      atomicMin(&b[2 * a], l);
      atomicMax(&b[2 * a + 1], h);
    }
  }
}

// dst[p][i][j][k] = src[p][x0 + i][y0 + j][z0 + k] over planes p (batch x components)
template <typename T>
__global__ __launch_bounds__(256) void k_gather(long long nplanes, int nx, int ny, int nz, int bx, int by, int bz,
                                                int x0, int y0, int z0, const cpx<T>* src, cpx<T>* dst) {
  const long long mb = (long long)bx * by * bz;
  const long long n = nplanes * mb;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const long long p = t / mb;
    long long r = t - p * mb;
    const int k = (int)(r % bz);
    r /= bz;
    const int j = (int)(r % by), i = (int)(r / by);
    dst[t] = src[((p * nx + (x0 + i)) * ny + (y0 + j)) * (long long)nz + (z0 + k)];
  }
}

// out = box value inside the box; outside it f - y (f and y given) or 0.
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(long long nplanes, int nx, int ny, int nz, int bx, int by, int bz,
                                                 int x0, int y0, int z0, const cpx<T>* box, const cpx<T>* f,
                                                 const cpx<T>* y, cpx<T>* out) {
  const long long m = (long long)nx * ny * nz;
  const long long n = nplanes * m;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const long long p = t / m;
    long long r = t - p * m;
    const int z = (int)(r % nz);
    r /= nz;
    const int yy = (int)(r % ny), x = (int)(r / ny);
    const int i = x - x0, j = yy - y0, k = z - z0;
    cpx<T> v;
    if (i >= 0 && i < bx && j >= 0 && j < by && k >= 0 && k < bz)
      v = box[((p * bx + i) * by + j) * (long long)bz + k];
    else if (f)
      v = f[t] - y[t];
    else
      v = czero<T>();
    out[t] = v;
  }
}

unsigned grid_for(long long n) {
  const long long want = (n + 255) / 256;
  return (unsigned)(want < 1 ? 1 : (want > 65536 ? 65536 : want));
}

int check_box(const char* who, int nx, int ny, int nz, int bx, int by, int bz, int x0, int y0, int z0) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return sfail(std::string(who) + ": grid sizes must be positive");
  if (bx <= 0 || by <= 0 || bz <= 0) return sfail(std::string(who) + ": box sizes must be positive");
  if (x0 < 0 || y0 < 0 || z0 < 0 || x0 + bx > nx || y0 + by > ny || z0 + bz > nz)
    return sfail(std::string(who) + ": box does not lie inside the grid");
  return 0;
}

}  // namespace

extern "C" {

int em3d_support_bounds(int nx, int ny, int nz, int precision, const void* chi, int chi_kind, int* bounds,
                        hipStream_t stream) {
  if (!chi || !bounds) return sfail("em3d_support_bounds: null argument");
  if (nx <= 0 || ny <= 0 || nz <= 0) return sfail("em3d_support_bounds: grid sizes must be positive");
  int nent;
  switch (chi_kind) {
    case EM3D_CHI_ISO: nent = 1; break;
    case EM3D_CHI_DIAG: nent = 3; break;
    case EM3D_CHI_FULL: nent = 9; break;
    default: return sfail("em3d_support_bounds: chi_kind must be ISO, DIAG or FULL");
  }
  const long long m = (long long)nx * ny * nz;
  const long long want = (m + kBoundThreads - 1) / kBoundThreads;
  const unsigned nb = (unsigned)(want > 2048 ? 2048 : want);
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, stream, nx, ny, nz, bounds);
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_bounds<double>, dim3(nb), dim3(kBoundThreads), 0, stream, nx, ny, nz,
                       (const cpx<double>*)chi, nent, bounds);
  else
    hipLaunchKernelGGL(k_bounds<float>, dim3(nb), dim3(kBoundThreads), 0, stream, nx, ny, nz,
                       (const cpx<float>*)chi, nent, bounds);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(std::string("em3d_support_bounds: ") + hipGetErrorString(e));
  return 0;
}

int em3d_support_gather(int precision, long long nplanes, int nx, int ny, int nz, int bx, int by, int bz, int x0,
                        int y0, int z0, const void* src, void* dst, hipStream_t stream) {
  if (!src || !dst) return sfail("em3d_support_gather: null argument");
  if (nplanes < 0) return sfail("em3d_support_gather: negative plane count");
  if (int rc = check_box("em3d_support_gather", nx, ny, nz, bx, by, bz, x0, y0, z0)) return rc;
  const long long n = nplanes * bx * by * bz;
  if (n == 0) return 0;
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, nx, ny, nz, bx, by, bz,
                       x0, y0, z0, (const cpx<double>*)src, (cpx<double>*)dst);
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, nx, ny, nz, bx, by, bz,
                       x0, y0, z0, (const cpx<float>*)src, (cpx<float>*)dst);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(std::string("em3d_support_gather: ") + hipGetErrorString(e));
  return 0;
}

int em3d_support_scatter(int precision, long long nplanes, int nx, int ny, int nz, int bx, int by, int bz, int x0,
                         int y0, int z0, const void* box, const void* f, const void* y, void* out,
                         hipStream_t stream) {
  if (!box || !out) return sfail("em3d_support_scatter: null argument");
  if ((f == nullptr) != (y == nullptr)) return sfail("em3d_support_scatter: give both f and y, or neither");
  if (nplanes < 0) return sfail("em3d_support_scatter: negative plane count");
  if (int rc = check_box("em3d_support_scatter", nx, ny, nz, bx, by, bz, x0, y0, z0)) return rc;
  const long long n = nplanes * nx * ny * nz;
  if (n == 0) return 0;
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_scatter<double>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, nx, ny, nz, bx, by, bz,
                       x0, y0, z0, (const cpx<double>*)box, (const cpx<double>*)f, (const cpx<double>*)y,
                       (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_scatter<float>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, nx, ny, nz, bx, by, bz,
                       x0, y0, z0, (const cpx<float>*)box, (const cpx<float>*)f, (const cpx<float>*)y,
                       (cpx<float>*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(std::string("em3d_support_scatter: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
