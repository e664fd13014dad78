// Far-field scatter integral of the polarisation current:
//   F_b(d_m) = dv sum_q (chi u_b)(r_q) exp(-i k0 d_m . r_q)
// for a batch of fields and many observation directions (EM: 3 components
// of chi u; acoustic: the scalar chi u).  The reference evaluates it as a
// dense (3 x cells) x (cells x directions) product of explicit phase
// matrices, batch by batch (src/em3d/farfield/_core.py:8-71).  Here the
// phase is separated over the axes of the grid,
//   exp(-i k0 d . r) = X_d(x) Y_d(y) Z_d(z),
// and so is the sum: F_d = sum_x X_d(x) sum_y Y_d(y) sum_z Z_d(z) p(x, y, z).
// A first kernel tabulates X, Y, Z for every direction (FP64 sincos over
// ndir (nx + ny + nz) entries); the z sum, one complex multiply-add per
// cell and direction, runs as an LDS-tiled product in FP64 registers with
// the y and x sums in its epilogue, and no phase matrix is ever stored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "e3_common.hpp"
#include "e3_error.hpp"

using namespace e3;

namespace {

constexpr int kT = 256;

template <typename T>
__device__ __forceinline__ cpx<double> ld(const cpx<T>* p, long long i) {
  const cpx<T> v = p[i];
  return mk((double)v.re, (double)v.im);
}

// tab[m][0 .. nx) = X_m, [nx .. nx + ny) = Y_m, [nx + ny .. nx + ny + nz) = Z_m
template <typename T>
__global__ __launch_bounds__(kT) void k_phase(int nx, int ny, int nz, const T* ax, const T* ay, const T* az,
                                              int ndir, const double* dirs, double k0, cpx<double>* tab) {
  const int S = nx + ny + nz;
  const long long total = (long long)ndir * S;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(t / S), r = (int)(t - (long long)m * S);
    const int a = r < nx ? 0 : (r < nx + ny ? 1 : 2);
    const double v = a == 0 ? (double)ax[r] : (a == 1 ? (double)ay[r - nx] : (double)az[r - nx - ny]);
    double sn, cs;
    sincos(-k0 * dirs[m * 3 + a] * v, &sn, &cs);
    tab[t] = mk(cs, sn);
  }
}

// The contrast current p_a = (chi u)_a of cell i in FP64, from any packed chi
// (C == 1: scalar; C == 3: kind 1 isotropic, 2 diagonal, 3 full 3 x 3).
template <typename T, int C>
__device__ __forceinline__ cpx<double> pol_at(const cpx<T>* chi, int chi_kind, const cpx<T>* ub, long long M,
                                              long long i, int a) {
  if (C == 1) return ld(chi, i) * ld(ub, i);
  if (chi_kind == 1) return ld(chi, i) * ld(ub, (long long)a * M + i);
  if (chi_kind == 2) return ld(chi, (long long)a * M + i) * ld(ub, (long long)a * M + i);
  cpx<double> s = mk(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < C; ++c) s = s + ld(chi, (long long)(a * C + c) * M + i) * ld(ub, (long long)c * M + i);
  return s;
}

// Tiles of the factored sum F_d = sum_x X_d(x) sum_y Y_d(y) sum_z Z_d(z) p(x, y, z).
// A block takes one x plane, one tile of kDT directions and one field.  The
// z sum is a complex product (y rows x z) . (z x directions) staged through
// LDS in chunks of kKZ, each thread holding a kFR x kFD tile of it in FP64
// registers; the y sum and the X factor follow in the block's epilogue.
constexpr int kFR = 4, kFD = 4;         // rows (y) x directions per thread
constexpr int kRT = 16 * kFR;           // y rows per tile
constexpr int kDT = 16 * kFD;           // directions per tile
#ifndef E3_FF_KZ
#define E3_FF_KZ 16                     // z per LDS chunk
#endif
#ifndef E3_FF_UNROLL
#define E3_FF_UNROLL 1                  // unroll of the z loop of a chunk
#endif
#ifndef E3_FF_PREFETCH
#define E3_FF_PREFETCH 0                // 1: next chunk's loads in flight during the products (more VGPRs)
#endif
#ifndef E3_FF_WAVES
#define E3_FF_WAVES 1                   // minimum waves per SIMD the register budget must allow
#endif
constexpr int kKZ = E3_FF_KZ;
constexpr int kKP = kKZ + 1;            // padded LDS row

// grid (nx, direction tiles, batch); part[b][m][a][x] (complex).  The block
// walks chunks ch = (component a, y tile, z chunk) in that order; a chunk's
// p and Z values are loaded into registers (one chunk ahead with
// E3_FF_PREFETCH) and stored to LDS between the two barriers that open it.
template <typename T, int C>
__global__ __launch_bounds__(kT, E3_FF_WAVES) void k_scatter(int nx, int ny, int nz, const cpx<T>* chi, int chi_kind,
                                                const cpx<T>* u, int ndir, const cpx<double>* tab,
                                                cpx<double>* part) {
  constexpr int NQ = kRT * kKZ / kT, NQZ = kDT * kKZ / kT;   // values per thread per chunk
  __shared__ cpx<double> sp[kRT * kKP];   // p[y][z] of the chunk; the y-sum reduction afterwards
  __shared__ cpx<double> sz[kDT * kKP];   // Z[d][z] of the chunk
  const long long M = (long long)nx * ny * nz;
  const int ix = blockIdx.x, m0 = blockIdx.y * kDT, b = blockIdx.z;
  const cpx<T>* ub = u + (size_t)b * C * M;
  const int S = nx + ny + nz;
  const int nd = ndir - m0 < kDT ? ndir - m0 : kDT;   // live directions of the tile
  const int tid = threadIdx.x, td = tid & 15, tr = tid >> 4;
  const long long xb = (long long)ix * ny * nz;
  const int nyt = (ny + kRT - 1) / kRT, nzt = (nz + kKZ - 1) / kKZ, nch = C * nyt * nzt;
  cpx<double> pv[NQ], zv[NQZ];
  auto fetch = [&](int ch) {
    const int a = ch / (nyt * nzt), rem = ch - a * nyt * nzt;
    const int y0 = (rem / nzt) * kRT, z0 = (rem % nzt) * kKZ;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + q * kT, y = y0 + e / kKZ, z = z0 + e % kKZ;
      pv[q] = (y < ny && z < nz) ? pol_at<T, C>(chi, chi_kind, ub, M, xb + (long long)y * nz + z, a) : mk(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < NQZ; ++q) {
      const int e = tid + q * kT, d = e / kKZ, z = z0 + e % kKZ;
      zv[q] = (d < nd && z < nz) ? tab[(size_t)(m0 + d) * S + nx + ny + z] : mk(0.0, 0.0);
    }
  };
  cpx<double> s[kFD], acc[kFR][kFD];
#pragma unroll
  for (int k = 0; k < kFD; ++k) {
    s[k] = mk(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < kFR; ++j) acc[j][k] = mk(0.0, 0.0);
  }
  if (E3_FF_PREFETCH) fetch(0);
  for (int ch = 0; ch < nch; ++ch) {
    const int a = ch / (nyt * nzt), rem = ch - a * nyt * nzt, yt = rem / nzt, zt = rem - yt * nzt;
    __syncthreads();
    if (!E3_FF_PREFETCH) fetch(ch);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + q * kT;
      sp[(e / kKZ) * kKP + e % kKZ] = pv[q];
    }
#pragma unroll
    for (int q = 0; q < NQZ; ++q) {
      const int e = tid + q * kT;
      sz[(e / kKZ) * kKP + e % kKZ] = zv[q];
    }
    __syncthreads();
    if (E3_FF_PREFETCH && ch + 1 < nch) fetch(ch + 1);
#pragma unroll E3_FF_UNROLL
    for (int kz = 0; kz < kKZ; ++kz) {
      cpx<double> p[kFR], w[kFD];
#pragma unroll
      for (int j = 0; j < kFR; ++j) p[j] = sp[(tr + 16 * j) * kKP + kz];
#pragma unroll
      for (int k = 0; k < kFD; ++k) w[k] = sz[(td + 16 * k) * kKP + kz];
#pragma unroll
      for (int j = 0; j < kFR; ++j)
#pragma unroll
        for (int k = 0; k < kFD; ++k) cmac_f<false>(acc[j][k], w[k], p[j]);
    }
    if (zt + 1 < nzt) continue;
    // end of a y tile: s_d += Y_d(y) acc(y, d)
#pragma unroll
    for (int j = 0; j < kFR; ++j) {
      const int y = yt * kRT + tr + 16 * j;
#pragma unroll
      for (int k = 0; k < kFD; ++k) {
        const int d = td + 16 * k;
        if (y < ny && d < nd) cmac_f<false>(s[k], tab[(size_t)(m0 + d) * S + nx + y], acc[j][k]);
        acc[j][k] = mk(0.0, 0.0);
      }
    }
    if (yt + 1 < nyt) continue;
    // end of component a: sum s over the 16 row groups, times X_d(x)
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kFD; ++k) {
      sp[tr * kDT + td + 16 * k] = s[k];
      s[k] = mk(0.0, 0.0);
    }
    __syncthreads();
    if (tid < nd) {
      cpx<double> v = mk(0.0, 0.0);
      for (int q = 0; q < kT / 16; ++q) v = v + sp[q * kDT + tid];
      part[(((size_t)b * ndir + m0 + tid) * C + a) * nx + ix] = tab[(size_t)(m0 + tid) * S + ix] * v;
    }
  }
}

// out[b][m][a] = dv sum_x part[b][m][a][x] (complex128)
__global__ void k_scatter_final(long long total, int nx, const cpx<double>* part, double dv, cpx<double>* out) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    cpx<double> s = mk(0.0, 0.0);
    for (int x = 0; x < nx; ++x) s = s + part[(size_t)t * nx + x];
    out[t] = dv * s;
  }
}

}  // namespace

extern "C" {

static size_t scatter_part_bytes(int nx, int batch, int ndir, int ncomp) {
  return ((size_t)batch * ndir * ncomp * nx * sizeof(cpx<double>) + 255) & ~(size_t)255;
}

size_t em3d_scatter_work_bytes(int nx, int ny, int nz, int batch, int ndir, int ncomp) {
  return scatter_part_bytes(nx, batch, ndir, ncomp) + (size_t)ndir * (nx + ny + nz) * sizeof(cpx<double>);
}

int em3d_scatter_integral(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                          int batch, int ncomp, const void* chi, int chi_kind, const void* u, int ndir,
                          const double* dirs, double k0, double dv, double* out, void* work, hipStream_t stream) {
  if (!ax || !ay || !az || !chi || !u || !dirs || !out || !work) return set_error("em3d_scatter_integral: null argument");
  if (batch < 1 || ndir < 1 || nx < 1 || ny < 1 || nz < 1) return set_error("em3d_scatter_integral: bad sizes");
  if (ncomp != 1 && ncomp != 3) return set_error("em3d_scatter_integral: ncomp must be 1 or 3");
  if (ncomp == 3 && (chi_kind < 1 || chi_kind > 3)) return set_error("em3d_scatter_integral: chi_kind must be 1, 2, 3");
  if (ndir > 65535 * kDT || batch > 65535) return set_error("em3d_scatter_integral: too many directions or fields");
  const dim3 g(nx, (ndir + kDT - 1) / kDT, batch);
  cpx<double>* part = (cpx<double>*)work;
  cpx<double>* tab = (cpx<double>*)((char*)work + scatter_part_bytes(nx, batch, ndir, ncomp));
  const long long ntab = (long long)ndir * (nx + ny + nz);
  const unsigned gtab = (unsigned)std::min<long long>((ntab + kT - 1) / kT, 4096);
  if (precision == 1)
    hipLaunchKernelGGL(k_phase<double>, dim3(gtab), dim3(kT), 0, stream, nx, ny, nz, (const double*)ax,
                       (const double*)ay, (const double*)az, ndir, dirs, k0, tab);
  else
    hipLaunchKernelGGL(k_phase<float>, dim3(gtab), dim3(kT), 0, stream, nx, ny, nz, (const float*)ax,
                       (const float*)ay, (const float*)az, ndir, dirs, k0, tab);
#define E3_SCATTER(TT, CC)                                                                                    \
  hipLaunchKernelGGL((k_scatter<TT, CC>), g, dim3(kT), 0, stream, nx, ny, nz, (const cpx<TT>*)chi, chi_kind,   \
                     (const cpx<TT>*)u, ndir, (const cpx<double>*)tab, part)
  if (precision == 1) {
    if (ncomp == 3) E3_SCATTER(double, 3);
    else E3_SCATTER(double, 1);
  } else {
    if (ncomp == 3) E3_SCATTER(float, 3);
    else E3_SCATTER(float, 1);
  }
#undef E3_SCATTER
  const long long total = (long long)batch * ndir * ncomp;
  const unsigned gfin = (unsigned)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_scatter_final, dim3(gfin), dim3(256), 0, stream, total, nx, (const cpx<double>*)part, dv,
                     (cpx<double>*)out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_scatter_integral: ") + hipGetErrorString(e));
}

}  // extern "C"
