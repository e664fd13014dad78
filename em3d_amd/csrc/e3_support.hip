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

// The box's own axes: box axis a is grid axis perm[a] (the sides of the box
// operator are n[perm[0]], n[perm[1]], n[perm[2]]); planes come in groups of
// ncc = 1 (scalar), 3 (vector components, which follow the axes) or 9 (a
// 3 x 3 tensor, both indices following the axes).
struct BoxMap {
  int N[3];      // grid
  int n[3];      // box sides along the grid axes
  int lo[3];     // box offset along the grid axes
  int perm[3];   // box axis a = grid axis perm[a]
  int ncc;
};

// grid plane of box plane q within its group
__device__ __forceinline__ int grid_plane(const BoxMap& m, int q) {
  if (m.ncc == 3) return m.perm[q];
  if (m.ncc == 9) return 3 * m.perm[q / 3] + m.perm[q % 3];
  return q;
}

// dst[p][d0][d1][d2] (box axes) = src[grid plane][lo + d on the grid axes]
template <typename T>
__global__ __launch_bounds__(256) void k_gather(long long nplanes, BoxMap m, const cpx<T>* src, cpx<T>* dst) {
  const int b0 = m.n[m.perm[0]], b1 = m.n[m.perm[1]], b2 = m.n[m.perm[2]];
  const long long mb = (long long)b0 * b1 * b2;
  const long long n = nplanes * mb;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const long long p = t / mb;
    long long r = t - p * mb;
    int d[3];
    d[2] = (int)(r % b2);
    r /= b2;
    d[1] = (int)(r % b1);
    d[0] = (int)(r / b1);
    int g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) g[m.perm[a]] = m.lo[m.perm[a]] + d[a];
    const long long grp = p / m.ncc;
    const long long sp = grp * m.ncc + grid_plane(m, (int)(p - grp * m.ncc));
    dst[t] = src[((sp * m.N[0] + g[0]) * m.N[1] + g[1]) * (long long)m.N[2] + g[2]];
  }
}

// out[p][grid] = the box value inside the box; outside it f - y (f and y given) or 0
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(long long nplanes, BoxMap m, const cpx<T>* box, const cpx<T>* f,
                                                 const cpx<T>* y, cpx<T>* out) {
  const long long mg = (long long)m.N[0] * m.N[1] * m.N[2];
  const int b1 = m.n[m.perm[1]], b2 = m.n[m.perm[2]];
  const long long n = nplanes * mg;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const long long p = t / mg;
    long long r = t - p * mg;
    int g[3];
    g[2] = (int)(r % m.N[2]);
    r /= m.N[2];
    g[1] = (int)(r % m.N[1]);
    g[0] = (int)(r / m.N[1]);
    int i[3];
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      i[a] = g[a] - m.lo[a];
      in = in && i[a] >= 0 && i[a] < m.n[a];
    }
    cpx<T> v;
    if (in) {
      // the box plane whose grid plane is p's, within p's group
      const long long grp = p / m.ncc;
      const int gp = (int)(p - grp * m.ncc);
      int q = gp;
      for (int c = 0; c < m.ncc; ++c)
        if (grid_plane(m, c) == gp) q = c;
      const long long bp = grp * m.ncc + q;
      v = box[((bp * m.n[m.perm[0]] + i[m.perm[0]]) * b1 + i[m.perm[1]]) * (long long)b2 + i[m.perm[2]]];
    } else if (f) {
      v = f[t] - y[t];
    } else {
      v = czero<T>();
    }
    out[t] = v;
  }
}

unsigned grid_for(long long n) {
  const long long want = (n + 255) / 256;
  return (unsigned)(want < 1 ? 1 : (want > 65536 ? 65536 : want));
}

int make_map(const char* who, long long nplanes, int ncc, const int* grid, const int* box, const int* lo,
             const int* perm, BoxMap* m) {
  if (!grid || !box || !lo || !perm) return sfail(std::string(who) + ": null argument");
  if (ncc != 1 && ncc != 3 && ncc != 9) return sfail(std::string(who) + ": ncc must be 1, 3 or 9");
  if (nplanes < 0 || nplanes % ncc) return sfail(std::string(who) + ": nplanes must be a multiple of ncc");
  bool seen[3] = {false, false, false};
  for (int a = 0; a < 3; ++a) {
    if (perm[a] < 0 || perm[a] > 2 || seen[perm[a]]) return sfail(std::string(who) + ": perm is not a permutation");
    seen[perm[a]] = true;
    if (grid[a] <= 0) return sfail(std::string(who) + ": grid sizes must be positive");
    if (box[a] <= 0) return sfail(std::string(who) + ": box sizes must be positive");
    if (lo[a] < 0 || lo[a] + box[a] > grid[a]) return sfail(std::string(who) + ": box does not lie inside the grid");
    m->N[a] = grid[a];
    m->n[a] = box[a];
    m->lo[a] = lo[a];
    m->perm[a] = perm[a];
  }
  m->ncc = ncc;
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

int em3d_support_gather(int precision, long long nplanes, int ncc, const int* grid, const int* box, const int* lo,
                        const int* perm, const void* src, void* dst, hipStream_t stream) {
  if (!src || !dst) return sfail("em3d_support_gather: null argument");
  BoxMap m;
  if (int rc = make_map("em3d_support_gather", nplanes, ncc, grid, box, lo, perm, &m)) return rc;
  const long long n = nplanes * m.n[0] * m.n[1] * m.n[2];
  if (n == 0) return 0;
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, m,
                       (const cpx<double>*)src, (cpx<double>*)dst);
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, m, (const cpx<float>*)src,
                       (cpx<float>*)dst);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(std::string("em3d_support_gather: ") + hipGetErrorString(e));
  return 0;
}

int em3d_support_scatter(int precision, long long nplanes, int ncc, const int* grid, const int* box, const int* lo,
                         const int* perm, const void* boxed, const void* f, const void* y, void* out,
                         hipStream_t stream) {
  if (!boxed || !out) return sfail("em3d_support_scatter: null argument");
  if ((f == nullptr) != (y == nullptr)) return sfail("em3d_support_scatter: give both f and y, or neither");
  BoxMap m;
  if (int rc = make_map("em3d_support_scatter", nplanes, ncc, grid, box, lo, perm, &m)) return rc;
  const long long n = nplanes * m.N[0] * m.N[1] * m.N[2];
  if (n == 0) return 0;
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_scatter<double>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, m,
                       (const cpx<double>*)boxed, (const cpx<double>*)f, (const cpx<double>*)y, (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_scatter<float>, dim3(grid_for(n)), dim3(256), 0, stream, nplanes, m,
                       (const cpx<float>*)boxed, (const cpx<float>*)f, (const cpx<float>*)y, (cpx<float>*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(std::string("em3d_support_scatter: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
