// Device generation of right-hand sides and materials.
//
// Plane waves  f_b = A_b exp(i sign (k o_b . r + phi0_b))  for a batch of
// incident directions (flat_wave_vec, src/em3d/wave.py:9-37, one call per
// field in the reference), and piecewise-constant materials made of
// ellipsoids evaluated at cell centres (center_mask + sample_contrast_tensor,
// src/em3d/geometry.py:67-93, 233-252; eta_ellipsoid,
// src/em3d/acoustics/materials.py:62-83), later ellipsoids overwriting
// earlier ones.  The ellipsoid test follows the reference's arithmetic
// exactly (same operation order, no contraction into FMAs, same precision),
// so masks agree bit for bit, boundary cells included.
#include <hip/hip_runtime.h>

#include <string>

#include "e3_common.hpp"
#include "e3_error.hpp"

using namespace e3;

namespace {

constexpr int kT = 256;

template <typename R>
__global__ __launch_bounds__(kT) void k_plane_waves(int nx, int ny, int nz, const R* ax, const R* ay, const R* az,
                                                    int B, const double* prm, double k, int sign, cpx<R>* out) {
  const long long M = (long long)nx * ny * nz;
  const int b = blockIdx.y;
  const double* q = prm + (size_t)b * 10;   // o(3), amp re(3), amp im(3), phi0
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(i % nz);
    const long long t = i / nz;
    const int iy = (int)(t % ny), ix = (int)(t / ny);
    const double ph = sign * (k * (q[0] * (double)ax[ix] + q[1] * (double)ay[iy] + q[2] * (double)az[iz]) + q[9]);
    double sn, cs;
    sincos(ph, &sn, &cs);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double ar = q[3 + c], ai = q[6 + c];
      out[((size_t)b * 3 + c) * M + i] = mk((R)(ar * cs - ai * sn), (R)(ar * sn + ai * cs));
    }
  }
}

// metric of one ellipsoid at a cell centre, in the arithmetic type A
template <typename A>
__device__ __forceinline__ bool inside(A x, A y, A z, const double* e, A limit) {
#pragma clang fp contract(off)
  const A dx = (x - (A)e[0]) / (A)e[3];
  const A dy = (y - (A)e[1]) / (A)e[4];
  const A dz = (z - (A)e[2]) / (A)e[5];
  A v = dx * dx;
  v = v + dy * dy;
  v = v + dz * dz;
  return v <= limit;
}

// EM: chi = eps - I, eps = background or the last ellipsoid containing the
// cell; ellipsoid record: centre(3), radii(3), eps 3x3 complex (18, row
// major re/im pairs).  Test in FP64 with tolerance `tol` (geometry.py:62-64).
// kind: 1 ISO (Nx,Ny,Nz) from eps[0][0], 2 DIAG (3,...), 3 FULL (3,3,...).
template <typename R>
__global__ __launch_bounds__(kT) void k_em_contrast(int nx, int ny, int nz, const R* ax, const R* ay, const R* az,
                                                    int n_ell, const double* ell, const double* bg, double tol,
                                                    int kind, cpx<R>* chi) {
  const long long M = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(i % nz);
    const long long t = i / nz;
    const int iy = (int)(t % ny), ix = (int)(t / ny);
    const double x = (double)ax[ix], y = (double)ay[iy], z = (double)az[iz];
    const double* eps = bg;
    for (int e = 0; e < n_ell; ++e) {
      const double* r = ell + (size_t)e * 24;
      if (inside<double>(x, y, z, r, 1.0 + tol)) eps = r + 6;
    }
    if (kind == 1) {
      chi[i] = mk((R)(eps[0] - 1.0), (R)eps[1]);
    } else if (kind == 2) {
#pragma unroll
      for (int a = 0; a < 3; ++a) chi[(size_t)a * M + i] = mk((R)(eps[2 * (4 * a)] - 1.0), (R)eps[2 * (4 * a) + 1]);
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double re = eps[2 * (3 * a + c)] - (a == c ? 1.0 : 0.0);
          chi[(size_t)(3 * a + c) * M + i] = mk((R)re, (R)eps[2 * (3 * a + c) + 1]);
        }
    }
  }
}

// acoustic: eta = outside or the last inclusion containing the cell; record:
// centre(3), radii(3), eta re, eta im.  Test in the axes' own real type,
// without tolerance (acoustics/materials.py:62-83).
template <typename R>
__global__ __launch_bounds__(kT) void k_eta(int nx, int ny, int nz, const R* ax, const R* ay, const R* az, int n_inc,
                                            const double* inc, double out_re, double out_im, cpx<R>* eta) {
  const long long M = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(i % nz);
    const long long t = i / nz;
    const int iy = (int)(t % ny), ix = (int)(t / ny);
    double re = out_re, im = out_im;
    for (int e = 0; e < n_inc; ++e) {
      const double* r = inc + (size_t)e * 8;
      if (inside<R>(ax[ix], ay[iy], az[iz], r, (R)1)) { re = r[6]; im = r[7]; }
    }
    eta[i] = mk((R)re, (R)im);
  }
}

unsigned blocks_for(long long m) {
  const long long b = (m + kT - 1) / kT;
  return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

int done(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int em3d_plane_waves(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision, int batch,
                     const double* params, double k, int sign, void* out, hipStream_t stream) {
  if (!ax || !ay || !az || !params || !out) return set_error("em3d_plane_waves: null argument");
  if (nx < 1 || ny < 1 || nz < 1 || batch < 1) return set_error("em3d_plane_waves: sizes must be positive");
  if (sign != 1 && sign != -1) return set_error("em3d_plane_waves: sign must be 1 or -1");
  const long long M = (long long)nx * ny * nz;
  const dim3 g(blocks_for(M), batch);
  if (precision == 1)
    hipLaunchKernelGGL(k_plane_waves<double>, g, dim3(kT), 0, stream, nx, ny, nz, (const double*)ax,
                       (const double*)ay, (const double*)az, batch, params, k, sign, (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_plane_waves<float>, g, dim3(kT), 0, stream, nx, ny, nz, (const float*)ax, (const float*)ay,
                       (const float*)az, batch, params, k, sign, (cpx<float>*)out);
  return done("em3d_plane_waves");
}

int em3d_ellipsoid_contrast(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                            int n_ellipsoids, const double* ellipsoids, const double* background, double tol,
                            int chi_kind, void* chi, hipStream_t stream) {
  if (!ax || !ay || !az || !background || !chi || (n_ellipsoids > 0 && !ellipsoids))
    return set_error("em3d_ellipsoid_contrast: null argument");
  if (chi_kind < 1 || chi_kind > 3) return set_error("em3d_ellipsoid_contrast: chi_kind must be 1, 2 or 3");
  const long long M = (long long)nx * ny * nz;
  if (precision == 1)
    hipLaunchKernelGGL(k_em_contrast<double>, dim3(blocks_for(M)), dim3(kT), 0, stream, nx, ny, nz, (const double*)ax,
                       (const double*)ay, (const double*)az, n_ellipsoids, ellipsoids, background, tol, chi_kind,
                       (cpx<double>*)chi);
  else
    hipLaunchKernelGGL(k_em_contrast<float>, dim3(blocks_for(M)), dim3(kT), 0, stream, nx, ny, nz, (const float*)ax,
                       (const float*)ay, (const float*)az, n_ellipsoids, ellipsoids, background, tol, chi_kind,
                       (cpx<float>*)chi);
  return done("em3d_ellipsoid_contrast");
}

int em3d_eta_inclusions(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                        int n_inclusions, const double* inclusions, double outside_re, double outside_im, void* eta,
                        hipStream_t stream) {
  if (!ax || !ay || !az || !eta || (n_inclusions > 0 && !inclusions))
    return set_error("em3d_eta_inclusions: null argument");
  const long long M = (long long)nx * ny * nz;
  if (precision == 1)
    hipLaunchKernelGGL(k_eta<double>, dim3(blocks_for(M)), dim3(kT), 0, stream, nx, ny, nz, (const double*)ax,
                       (const double*)ay, (const double*)az, n_inclusions, inclusions, outside_re, outside_im,
                       (cpx<double>*)eta);
  else
    hipLaunchKernelGGL(k_eta<float>, dim3(blocks_for(M)), dim3(kT), 0, stream, nx, ny, nz, (const float*)ax,
                       (const float*)ay, (const float*)az, n_inclusions, inclusions, outside_re, outside_im,
                       (cpx<float>*)eta);
  return done("em3d_eta_inclusions");
}

}  // extern "C"
