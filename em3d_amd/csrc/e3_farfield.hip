// Far-field scatter integral of the polarisation current:
//   F_b(d_m) = dv sum_q (chi u_b)(r_q) exp(-i k0 d_m . r_q)
// for a batch of fields and many observation directions (EM: 3 components
// of chi u; acoustic: the scalar chi u).  The reference evaluates it as a
// dense (3 x cells) x (cells x directions) product of explicit phase
// matrices, batch by batch (src/em3d/farfield/_core.py:8-71); here every
// thread forms chi u once per cell and accumulates a tile of directions in
// FP64 with the phases computed on the fly, so the current is read once per
// direction tile and no phase matrix is stored.
#include <hip/hip_runtime.h>

#include <string>

#include "e3_common.hpp"
#include "e3_error.hpp"

using namespace e3;

namespace {

constexpr int kT = 256;
constexpr int kDT = 8;   // directions per tile

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ cpx<double> ld(const cpx<T>* p, long long i) {
  const cpx<T> v = p[i];
  return mk((double)v.re, (double)v.im);
}

// grid (blocks over cells, direction tiles, batch); part[b][m][blk][2 C]
template <typename T, int C>
__global__ __launch_bounds__(kT) void k_scatter(int nx, int ny, int nz, const T* ax, const T* ay, const T* az,
                                                const cpx<T>* chi, int chi_kind, const cpx<T>* u, int ndir,
                                                const double* dirs, double k0, double* part) {
  const long long M = (long long)nx * ny * nz;
  const int b = blockIdx.z, m0 = blockIdx.y * kDT;
  const cpx<T>* ub = u + (size_t)b * C * M;
  double acc[kDT][2 * C];
#pragma unroll
  for (int d = 0; d < kDT; ++d)
#pragma unroll
    for (int c = 0; c < 2 * C; ++c) acc[d][c] = 0.0;
  double dir[kDT][3];
#pragma unroll
  for (int d = 0; d < kDT; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) dir[d][a] = (m0 + d < ndir) ? dirs[(m0 + d) * 3 + a] : 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(i % nz);
    const long long t = i / nz;
    const int iy = (int)(t % ny), ix = (int)(t / ny);
    const double x = (double)ax[ix], y = (double)ay[iy], z = (double)az[iz];
    cpx<double> pol[C];
    if (C == 1) {
      pol[0] = ld(chi, i) * ld(ub, i);
    } else if (chi_kind == 1) {
      const cpx<double> c0 = ld(chi, i);
#pragma unroll
      for (int a = 0; a < C; ++a) pol[a] = c0 * ld(ub, (long long)a * M + i);
    } else if (chi_kind == 2) {
#pragma unroll
      for (int a = 0; a < C; ++a) pol[a] = ld(chi, (long long)a * M + i) * ld(ub, (long long)a * M + i);
    } else {
#pragma unroll
      for (int a = 0; a < C; ++a) {
        cpx<double> s = mk(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < C; ++c) s = s + ld(chi, (long long)(a * C + c) * M + i) * ld(ub, (long long)c * M + i);
        pol[a] = s;
      }
    }
#pragma unroll
    for (int d = 0; d < kDT; ++d) {
      double sn, cs;
      sincos(-k0 * (dir[d][0] * x + dir[d][1] * y + dir[d][2] * z), &sn, &cs);
#pragma unroll
      for (int a = 0; a < C; ++a) {
        acc[d][2 * a] += pol[a].re * cs - pol[a].im * sn;
        acc[d][2 * a + 1] += pol[a].re * sn + pol[a].im * cs;
      }
    }
  }
  __shared__ double sh[kT / 64][kDT * 2 * C];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < kDT; ++d)
#pragma unroll
    for (int c = 0; c < 2 * C; ++c) {
      const double v = wsum(acc[d][c]);
      if (lane == 0) sh[w][d * 2 * C + c] = v;
    }
  __syncthreads();
  for (int k = threadIdx.x; k < kDT * 2 * C; k += blockDim.x) {
    const int d = k / (2 * C), c = k % (2 * C);
    if (m0 + d >= ndir) continue;
    double s = 0.0;
    for (int q = 0; q < kT / 64; ++q) s += sh[q][k];
    part[(((size_t)b * ndir + m0 + d) * gridDim.x + blockIdx.x) * 2 * C + c] = s;
  }
}

// out[b][m][c] = dv sum over blocks (complex128)
__global__ void k_scatter_final(int B, int ndir, int C, int nb, const double* part, double dv, double* out) {
  const long long total = (long long)B * ndir * 2 * C;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % (2 * C));
    const long long bm = t / (2 * C);
    double s = 0.0;
    for (int k = 0; k < nb; ++k) s += part[((size_t)bm * nb + k) * 2 * C + c];
    out[t] = dv * s;
  }
}

constexpr int kScatterBlocks = 512;

}  // namespace

extern "C" {

size_t em3d_scatter_work_bytes(int batch, int ndir, int ncomp) {
  return (size_t)batch * ndir * kScatterBlocks * 2 * ncomp * sizeof(double);
}

int em3d_scatter_integral(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                          int batch, int ncomp, const void* chi, int chi_kind, const void* u, int ndir,
                          const double* dirs, double k0, double dv, double* out, void* work, hipStream_t stream) {
  if (!ax || !ay || !az || !chi || !u || !dirs || !out || !work) return set_error("em3d_scatter_integral: null argument");
  if (batch < 1 || ndir < 1 || nx < 1 || ny < 1 || nz < 1) return set_error("em3d_scatter_integral: bad sizes");
  if (ncomp != 1 && ncomp != 3) return set_error("em3d_scatter_integral: ncomp must be 1 or 3");
  if (ncomp == 3 && (chi_kind < 1 || chi_kind > 3)) return set_error("em3d_scatter_integral: chi_kind must be 1, 2, 3");
  if (ndir > 65535 * kDT || batch > 65535) return set_error("em3d_scatter_integral: too many directions or fields");
  const dim3 g(kScatterBlocks, (ndir + kDT - 1) / kDT, batch);
  double* part = (double*)work;
#define E3_SCATTER(TT, CC)                                                                                    \
  hipLaunchKernelGGL((k_scatter<TT, CC>), g, dim3(kT), 0, stream, nx, ny, nz, (const TT*)ax, (const TT*)ay,   \
                     (const TT*)az, (const cpx<TT>*)chi, chi_kind, (const cpx<TT>*)u, ndir, dirs, k0, part)
  if (precision == 1) {
    if (ncomp == 3) E3_SCATTER(double, 3);
    else E3_SCATTER(double, 1);
  } else {
    if (ncomp == 3) E3_SCATTER(float, 3);
    else E3_SCATTER(float, 1);
  }
#undef E3_SCATTER
  const long long total = (long long)batch * ndir * 2 * ncomp;
  hipLaunchKernelGGL(k_scatter_final, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, batch, ndir, ncomp,
                     kScatterBlocks, (const double*)part, dv, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_scatter_integral: ") + hipGetErrorString(e));
}

}  // extern "C"
