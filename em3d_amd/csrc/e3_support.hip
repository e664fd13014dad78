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
// 3 x 3 tensor, both indices following the axes).  The strides and plane
// maps are worked out on the host, so that the kernels index no array by a
// run-time axis number.
struct BoxMap {
  int N[3];           // grid
  int n[3];           // box sides along the grid axes
  int lo[3];          // box offset along the grid axes
  int perm[3];        // box axis a = grid axis perm[a]
  int ncc;
  int bn[3];          // box sides in box-axis order, bn[k] = n[perm[k]]
  long long mg, mb;   // cells of the grid and of the box
  long long bst[3];   // box-layout stride of one step along grid axis a
  long long gst[3];   // grid-layout stride of one step along box axis k
  long long lo_off;   // grid index of the box's first cell
  int gplane[9];      // grid plane of box plane q within its group
  int bplane[9];      // box plane of grid plane g within its group
};

// Threads of a row kernel: W along the contiguous axis, R rows per block.
struct RowShape {
  int W, R;
};
inline RowShape row_shape(int len) {
  RowShape r;
  r.W = len < 256 ? len : 256;
  r.R = 256 / r.W;
  return r;
}

// dst[p][d0][d1][d2] (box axes) = src[grid plane][lo + d on the grid axes]:
// one thread per d2 of a box row, blockIdx.y the plane
template <typename T>
__global__ __launch_bounds__(256) void k_gather(BoxMap m, const cpx<T>* src, cpx<T>* dst) {
  const long long p = blockIdx.y;
  const int row = (int)blockIdx.x * (int)blockDim.y + (int)threadIdx.y;
  if (row >= m.bn[0] * m.bn[1]) return;
  const int d0 = row / m.bn[1], d1 = row - d0 * m.bn[1];
  const long long grp = p / m.ncc;
  const long long sp = grp * m.ncc + m.gplane[p - grp * m.ncc];
  const cpx<T>* s = src + sp * m.mg + m.lo_off + d0 * m.gst[0] + d1 * m.gst[1];
  cpx<T>* d = dst + p * m.mb + (long long)row * m.bn[2];
  for (int d2 = threadIdx.x; d2 < m.bn[2]; d2 += blockDim.x) d[d2] = s[d2 * m.gst[2]];
}

// out[p][grid] = the box value inside the box; outside it f - y (f and y
// given) or 0: one thread per z of a grid row, blockIdx.y the plane
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(BoxMap m, const cpx<T>* box, const cpx<T>* f, const cpx<T>* y,
                                                 cpx<T>* out) {
  const long long p = blockIdx.y;
  const int row = (int)blockIdx.x * (int)blockDim.y + (int)threadIdx.y;
  if (row >= m.N[0] * m.N[1]) return;
  const int g0 = row / m.N[1], g1 = row - g0 * m.N[1];
  const int i0 = g0 - m.lo[0], i1 = g1 - m.lo[1];
  const bool rin = i0 >= 0 && i0 < m.n[0] && i1 >= 0 && i1 < m.n[1];
  const cpx<T>* bx = box;
  if (rin) {
    // the box plane whose grid plane is p's, within p's group
    const long long grp = p / m.ncc;
    const long long bp = grp * m.ncc + m.bplane[p - grp * m.ncc];
    bx = box + bp * m.mb + i0 * m.bst[0] + i1 * m.bst[1];
  }
  const long long base = p * m.mg + (long long)row * m.N[2];
  for (int g2 = threadIdx.x; g2 < m.N[2]; g2 += blockDim.x) {
    const int i2 = g2 - m.lo[2];
    cpx<T> v;
    if (rin && i2 >= 0 && i2 < m.n[2]) v = bx[i2 * m.bst[2]];
    else if (f) v = f[base + g2] - y[base + g2];
    else v = czero<T>();
    out[base + g2] = v;
  }
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
  if (nplanes > 65535) return sfail(std::string(who) + ": at most 65535 planes per call");
  if ((long long)grid[0] * grid[1] > INT_MAX) return sfail(std::string(who) + ": grid too large");
  long long gstride[3];   // grid-layout stride of grid axis a
  gstride[2] = 1;
  gstride[1] = grid[2];
  gstride[0] = (long long)grid[1] * grid[2];
  for (int k = 0; k < 3; ++k) m->bn[k] = box[perm[k]];
  m->mg = (long long)grid[0] * grid[1] * grid[2];
  m->mb = (long long)box[0] * box[1] * box[2];
  const long long bstride_k[3] = {(long long)m->bn[1] * m->bn[2], m->bn[2], 1};   // box axis k
  for (int k = 0; k < 3; ++k) {
    m->bst[perm[k]] = bstride_k[k];
    m->gst[k] = gstride[perm[k]];
  }
  m->lo_off = lo[0] * gstride[0] + lo[1] * gstride[1] + lo[2];
  for (int q = 0; q < 9; ++q) m->gplane[q] = m->bplane[q] = 0;
  for (int q = 0; q < ncc; ++q) {
    const int g = ncc == 3 ? perm[q] : ncc == 9 ? 3 * perm[q / 3] + perm[q % 3] : q;
    m->gplane[q] = g;
    m->bplane[g] = q;
  }
  return 0;
}

// launch shape of a row kernel over `rows` rows of length `len`, `planes` planes
void row_launch(int rows, int len, long long planes, dim3* grid, dim3* block) {
  const RowShape r = row_shape(len);
  *block = dim3(r.W, r.R);
  *grid = dim3((unsigned)((rows + r.R - 1) / r.R), (unsigned)planes);
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
  if (nplanes == 0) return 0;
  dim3 g, blk;
  row_launch(m.bn[0] * m.bn[1], m.bn[2], nplanes, &g, &blk);
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_gather<double>, g, blk, 0, stream, m, (const cpx<double>*)src, (cpx<double>*)dst);
  else
    hipLaunchKernelGGL(k_gather<float>, g, blk, 0, stream, m, (const cpx<float>*)src, (cpx<float>*)dst);
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
  if (nplanes == 0) return 0;
  dim3 g, blk;
  row_launch(m.N[0] * m.N[1], m.N[2], nplanes, &g, &blk);
  if (precision == EM3D_DOUBLE)
    hipLaunchKernelGGL(k_scatter<double>, g, blk, 0, stream, m, (const cpx<double>*)boxed, (const cpx<double>*)f,
                       (const cpx<double>*)y, (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_scatter<float>, g, blk, 0, stream, m, (const cpx<float>*)boxed, (const cpx<float>*)f,
                       (const cpx<float>*)y, (cpx<float>*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(std::string("em3d_support_scatter: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
