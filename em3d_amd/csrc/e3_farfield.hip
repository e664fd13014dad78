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
// When every direction has the same component along one axis (the plane
// cuts of rcs_plane and pattern_plane), that axis's phase is common to all
// directions: the sum over it is taken once (k_collapse, one pass over the
// field), and the same product then runs on the remaining 2-D field.
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

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// tab[m][slot 0: n0 entries | slot 1: n1 | slot 2: n2], entry exp(-i k0 d_m[c_s] v_s[j]) of slot s;
// a slot without coordinates (v_s null) holds 1.  Slots (x, y, z) give X_m, Y_m, Z_m.
template <typename T>
__global__ __launch_bounds__(kT) void k_phase(int n0, int n1, int n2, const T* v0, const T* v1, const T* v2,
                                              int c0, int c1, int c2, int ndir, const double* dirs, double k0,
                                              cpx<double>* tab) {
  const int S = n0 + n1 + n2;
  const long long total = (long long)ndir * S;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(t / S), r = (int)(t - (long long)m * S);
    const int s = r < n0 ? 0 : (r < n0 + n1 ? 1 : 2);
    const T* v = s == 0 ? v0 : (s == 1 ? v1 : v2);
    const int j = s == 0 ? r : (s == 1 ? r - n0 : r - n0 - n1);
    const int c = s == 0 ? c0 : (s == 1 ? c1 : c2);
    const double x = v ? (double)v[j] : 0.0;
    double sn, cs;
    sincos(-k0 * dirs[m * 3 + c] * x, &sn, &cs);
    tab[t] = mk(cs, sn);
  }
}

// The contrast current p_a = (chi u)_a of cell i in FP64, from any packed chi
// (C == 1: scalar; C == 3: kind 1 isotropic, 2 diagonal, 3 full 3 x 3), or
// ub itself for kind 0 (a current already formed, such as a collapsed field).
template <typename T, int C>
__device__ __forceinline__ cpx<double> pol_at(const cpx<T>* chi, int chi_kind, const cpx<T>* ub, long long M,
                                              long long i, int a) {
  if (chi_kind == 0) return ld(ub, (long long)a * M + i);
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
// LDS in chunks of KZ: on the FP64 MFMA, each wave holding a 32 x 32 block
// of it (E3_FF_MFMA), or as FP64 FMAs with each thread holding kFR x kFD.
// The y sum and the X factor follow in the block's epilogue.
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
#define E3_FF_WAVES 3                   // minimum waves per SIMD the register budget must allow
#endif
#ifndef E3_FF_MFMA
#define E3_FF_MFMA 1                    // z sum on v_mfma_f64_16x16x4f64 (0: FP64 FMA register tiles)
#endif
constexpr int kKZ = E3_FF_KZ;
static_assert(kKZ % 4 == 0 && kRT == 64 && kDT == 64 && kT == 256, "tile shape of the far-field product");
typedef double d4v __attribute__((ext_vector_type(4)));

// grid (nx, direction tiles, batch); part[b][m][a][x] (complex).  The block
// walks chunks ch = (component a, y tile, z chunk) in that order; a chunk's
// p and Z values are loaded into registers (one chunk ahead with
// E3_FF_PREFETCH) and stored to LDS between the two barriers that open it.
template <typename T, int C, int KZ = kKZ>
__global__ __launch_bounds__(kT, E3_FF_WAVES) void k_scatter(int nx, int ny, int nz, const cpx<T>* chi, int chi_kind,
                                                const cpx<T>* u, int ndir, const cpx<double>* tab,
                                                cpx<double>* part) {
  static_assert(KZ % 4 == 0 && (kRT * KZ) % kT == 0 && (kDT * KZ) % kT == 0, "z chunk of the product");
  constexpr int KP = KZ + 1;                                 // padded LDS row
  constexpr int NQ = kRT * KZ / kT, NQZ = kDT * KZ / kT;     // values per thread per chunk
  constexpr int SPN = kRT * KP > kT * kFD ? kRT * KP : kT * kFD;
  __shared__ cpx<double> sp[SPN];         // p[y][z] of the chunk; the y-sum reduction afterwards
  __shared__ cpx<double> sz[kDT * KP];    // Z[d][z] of the chunk
  const long long M = (long long)nx * ny * nz;
  const int ix = blockIdx.x, m0 = blockIdx.y * kDT, b = blockIdx.z;
  const cpx<T>* ub = u + (size_t)b * C * M;
  const int S = nx + ny + nz;
  const int nd = ndir - m0 < kDT ? ndir - m0 : kDT;   // live directions of the tile
  const int tid = threadIdx.x, td = tid & 15, tr = tid >> 4;
  const long long xb = (long long)ix * ny * nz;
  const int nyt = (ny + kRT - 1) / kRT, nzt = (nz + KZ - 1) / KZ, nch = C * nyt * nzt;
  cpx<double> pv[NQ], zv[NQZ];
  auto fetch = [&](int ch) {
    const int a = ch / (nyt * nzt), rem = ch - a * nyt * nzt;
    const int y0 = (rem / nzt) * kRT, z0 = (rem % nzt) * KZ;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {   // cells off the grid load cell 0 and store zero: no branch, one wait
      const int e = tid + q * kT, y = y0 + e / KZ, z = z0 + e % KZ;
      const bool ok = y < ny && z < nz;
      const cpx<double> v = pol_at<T, C>(chi, chi_kind, ub, M, ok ? xb + (long long)y * nz + z : 0, a);
      pv[q] = ok ? v : mk(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < NQZ; ++q) {
      const int e = tid + q * kT, d = e / KZ, z = z0 + e % KZ;
      const bool ok = d < nd && z < nz;
      const cpx<double> v = tab[ok ? (size_t)(m0 + d) * S + nx + ny + z : 0];
      zv[q] = ok ? v : mk(0.0, 0.0);
    }
  };
#if E3_FF_MFMA
  // wave w: rows 32 (w >> 1) .. +32 and directions 32 (w & 1) .. +32 of the tile, as 2 x 2 blocks of
  // 16 x 16; lane l holds D[4 i + l / 16][l % 16] of a block (i < 4; the f64 MFMA's layout, unlike the
  // 4 (l / 16) + i rows of the 32-bit ones), real and imaginary parts apart
  const int lane = tid & 63, wr = (tid >> 6) >> 1, wc = (tid >> 6) & 1, li = lane & 15, lk = lane >> 4;
  d4v dre[2][2], dim[2][2];
  cpx<double> s[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) dre[rb][cb] = dim[rb][cb] = d4v{0.0, 0.0, 0.0, 0.0};
  s[0] = s[1] = mk(0.0, 0.0);
#else
  cpx<double> s[kFD], acc[kFR][kFD];
#pragma unroll
  for (int k = 0; k < kFD; ++k) {
    s[k] = mk(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < kFR; ++j) acc[j][k] = mk(0.0, 0.0);
  }
#endif
  if (E3_FF_PREFETCH) fetch(0);
  for (int ch = 0; ch < nch; ++ch) {
    const int a = ch / (nyt * nzt), rem = ch - a * nyt * nzt, yt = rem / nzt, zt = rem - yt * nzt;
    __syncthreads();
    if (!E3_FF_PREFETCH) fetch(ch);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + q * kT;
      sp[(e / KZ) * KP + e % KZ] = pv[q];
    }
#pragma unroll
    for (int q = 0; q < NQZ; ++q) {
      const int e = tid + q * kT;
      sz[(e / KZ) * KP + e % KZ] = zv[q];
    }
    __syncthreads();
    if (E3_FF_PREFETCH && ch + 1 < nch) fetch(ch + 1);
#if E3_FF_MFMA
    // D += P Z over the chunk, k = 4 z per step: lane l brings P[row l % 16][z 4 ks + l / 16] and
    // Z[z 4 ks + l / 16][direction l % 16]; four real products per complex block
#pragma unroll
    for (int ks = 0; ks < KZ / 4; ++ks) {
      const int kz = 4 * ks + lk;
      cpx<double> av[2], bv[2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) av[rb] = sp[(wr * 32 + rb * 16 + li) * KP + kz];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) bv[cb] = sz[(wc * 32 + cb * 16 + li) * KP + kz];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          dre[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rb].re, bv[cb].re, dre[rb][cb], 0, 0, 0);
          dim[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rb].re, bv[cb].im, dim[rb][cb], 0, 0, 0);
        }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          dre[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[rb].im, bv[cb].im, dre[rb][cb], 0, 0, 0);
          dim[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rb].im, bv[cb].re, dim[rb][cb], 0, 0, 0);
        }
    }
    if (zt + 1 < nzt) continue;
    // end of a y tile: s_d += Y_d(y) D(y, d)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const int d = wc * 32 + cb * 16 + li;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int y = yt * kRT + wr * 32 + rb * 16 + 4 * i + lk;
          const bool ok = y < ny && d < nd;
          const cpx<double> yv = tab[ok ? (size_t)(m0 + d) * S + nx + y : 0];
          cmac_f<false>(s[cb], ok ? yv : mk(0.0, 0.0), mk(dre[rb][cb][i], dim[rb][cb][i]));
        }
        dre[rb][cb] = dim[rb][cb] = d4v{0.0, 0.0, 0.0, 0.0};
        E3_SCHED_FENCE();   // four loads in flight at a time, not sixteen (registers)
      }
    if (yt + 1 < nyt) continue;
    // end of component a: sum s over the 8 lane groups holding a direction, times X_d(x)
    __syncthreads();
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      sp[(wr * 4 + lk) * kDT + wc * 32 + cb * 16 + li] = s[cb];
      s[cb] = mk(0.0, 0.0);
    }
    __syncthreads();
    if (tid < nd) {
      cpx<double> v = mk(0.0, 0.0);
      for (int q = 0; q < 8; ++q) v = v + sp[q * kDT + tid];
      part[(((size_t)b * ndir + m0 + tid) * C + a) * nx + ix] = tab[(size_t)(m0 + tid) * S + ix] * v;
    }
#else
#pragma unroll E3_FF_UNROLL
    for (int kz = 0; kz < KZ; ++kz) {
      cpx<double> p[kFR], w[kFD];
#pragma unroll
      for (int j = 0; j < kFR; ++j) p[j] = sp[(tr + 16 * j) * KP + kz];
#pragma unroll
      for (int k = 0; k < kFD; ++k) w[k] = sz[(td + 16 * k) * KP + kz];
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
        const bool ok = y < ny && d < nd;
        const cpx<double> yv = tab[ok ? (size_t)(m0 + d) * S + nx + y : 0];
        cmac_f<false>(s[k], ok ? yv : mk(0.0, 0.0), acc[j][k]);
        acc[j][k] = mk(0.0, 0.0);
      }
      E3_SCHED_FENCE();   // four loads in flight at a time, not sixteen (registers)
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
#endif
  }
}

// out[b][m][a] = dv sum_x part[b][m][a][x] (complex128), one wave per output
__global__ __launch_bounds__(kT) void k_scatter_final(long long total, int nx, const cpx<double>* part, double dv,
                                                      cpx<double>* out) {
  const int lane = threadIdx.x & 63;
  for (long long t = (long long)blockIdx.x * (kT / 64) + (threadIdx.x >> 6); t < total;
       t += (long long)gridDim.x * (kT / 64)) {
    cpx<double> s = mk(0.0, 0.0);
    for (int x = lane; x < nx; x += 64) s = s + part[(size_t)t * nx + x];
    s.re = wsum(s.re);
    s.im = wsum(s.im);
    if (lane == 0) out[t] = dv * s;
  }
}

// P[b][a][ib][ig] = sum_ia phi(ia) p_a(cell ia sa + ib sb + ig sg): the sum over an axis whose phase
// phi is the same for every direction.  A shared axis that is contiguous (sa == 1) is summed by
// the 64 lanes of a wave per output; otherwise a thread per output walks it, neighbouring
// threads on neighbouring cells.
template <typename T, int C>
__global__ __launch_bounds__(kT) void k_collapse(long long M, int na, int nb, int ng, long long sa, long long sb,
                                                 long long sg, const cpx<T>* chi, int chi_kind, const cpx<T>* u,
                                                 const cpx<double>* phi, int batch, cpx<double>* P) {
  const long long per = (long long)nb * ng, nout = (long long)batch * per;
  if (sa == 1) {
    const int lane = threadIdx.x & 63;
    for (long long o = (long long)blockIdx.x * (kT / 64) + (threadIdx.x >> 6); o < nout;
         o += (long long)gridDim.x * (kT / 64)) {
      const int b = (int)(o / per);
      const long long r = o - (long long)b * per;
      const long long base = (r / ng) * sb + (r % ng) * sg;
      const cpx<T>* ub = u + (size_t)b * C * M;
      cpx<double> s[C];
#pragma unroll
      for (int a = 0; a < C; ++a) s[a] = mk(0.0, 0.0);
      for (int ia = lane; ia < na; ia += 64) {
        const cpx<double> ph = phi[ia];
#pragma unroll
        for (int a = 0; a < C; ++a) cmac_f<false>(s[a], ph, pol_at<T, C>(chi, chi_kind, ub, M, base + ia, a));
      }
#pragma unroll
      for (int a = 0; a < C; ++a) {
        s[a].re = wsum(s[a].re);
        s[a].im = wsum(s[a].im);
        if (lane == 0) P[((size_t)b * C + a) * per + r] = s[a];
      }
    }
  } else {
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < nout;
         o += (long long)gridDim.x * blockDim.x) {
      const int b = (int)(o / per);
      const long long r = o - (long long)b * per;
      const long long base = (r / ng) * sb + (r % ng) * sg;
      const cpx<T>* ub = u + (size_t)b * C * M;
      cpx<double> s[C];
#pragma unroll
      for (int a = 0; a < C; ++a) s[a] = mk(0.0, 0.0);
      for (int ia = 0; ia < na; ++ia) {
        const cpx<double> ph = phi[ia];
#pragma unroll
        for (int a = 0; a < C; ++a)
          cmac_f<false>(s[a], ph, pol_at<T, C>(chi, chi_kind, ub, M, base + (long long)ia * sa, a));
      }
#pragma unroll
      for (int a = 0; a < C; ++a) P[((size_t)b * C + a) * per + r] = s[a];
    }
  }
}

}  // namespace

namespace {

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// work = part | tab | phi | P, each sized for either path
struct ScatterWork {
  size_t part, tab, phi, P;
  ScatterWork(int nx, int ny, int nz, int batch, int ndir, int ncomp) {
    const size_t c = sizeof(cpx<double>);
    const int nmax = std::max(nx, std::max(ny, nz));
    const size_t pair = std::max((size_t)nx * ny, std::max((size_t)nx * nz, (size_t)ny * nz));
    part = al256((size_t)batch * ndir * ncomp * nmax * c);
    tab = al256((size_t)ndir * (nx + ny + nz + 1) * c);
    phi = al256((size_t)nmax * c);
    P = al256((size_t)batch * ncomp * pair * c);
  }
  size_t total() const { return part + tab + phi + P; }
};

template <typename T>
void launch_phase(hipStream_t st, int n0, int n1, int n2, const void* v0, const void* v1, const void* v2, int c0,
                  int c1, int c2, int ndir, const double* dirs, double k0, cpx<double>* tab) {
  const long long n = (long long)ndir * (n0 + n1 + n2);
  const unsigned g = (unsigned)std::min<long long>((n + kT - 1) / kT, 4096);
  hipLaunchKernelGGL(k_phase<T>, dim3(g), dim3(kT), 0, st, n0, n1, n2, (const T*)v0, (const T*)v1, (const T*)v2,
                     c0, c1, c2, ndir, dirs, k0, tab);
}

template <typename T, int C>
void launch_scatter(hipStream_t st, int nx, int ny, int nz, int batch, const void* chi, int chi_kind, const void* u,
                    int ndir, int shared_axis, const void* const ax[3], const double* dirs, double k0, double dv,
                    double* out, const ScatterWork& w, char* work) {
  cpx<double>* part = (cpx<double>*)work;
  cpx<double>* tab = (cpx<double>*)(work + w.part);
  cpx<double>* phi = (cpx<double>*)(work + w.part + w.tab);
  cpx<double>* P = (cpx<double>*)(work + w.part + w.tab + w.phi);
  const unsigned nt = (unsigned)((ndir + kDT - 1) / kDT);
  int nsum = nx;   // extent of the x sum left to k_scatter_final
  if (shared_axis < 0) {
    launch_phase<T>(st, nx, ny, nz, ax[0], ax[1], ax[2], 0, 1, 2, ndir, dirs, k0, tab);
    hipLaunchKernelGGL((k_scatter<T, C>), dim3(nx, nt, batch), dim3(kT), 0, st, nx, ny, nz, (const cpx<T>*)chi,
                       chi_kind, (const cpx<T>*)u, ndir, (const cpx<double>*)tab, part);
  } else {
    const int n[3] = {nx, ny, nz};
    const long long stride[3] = {(long long)ny * nz, nz, 1};
    const int al = shared_axis, be = al == 0 ? 1 : 0, ga = al == 2 ? 1 : 2;   // remaining axes, in memory order
    launch_phase<T>(st, n[al], 0, 0, ax[al], nullptr, nullptr, al, 0, 0, 1, dirs, k0, phi);
    launch_phase<T>(st, n[be], n[ga], 1, ax[be], ax[ga], nullptr, be, ga, 0, ndir, dirs, k0, tab);
    const long long nout = (long long)batch * n[be] * n[ga];
    const long long threads = stride[al] == 1 ? 64 * nout : nout;
    const unsigned g = (unsigned)std::min<long long>((threads + kT - 1) / kT, 16384);
    hipLaunchKernelGGL((k_collapse<T, C>), dim3(g), dim3(kT), 0, st, (long long)nx * ny * nz, n[al], n[be], n[ga],
                       stride[al], stride[be], stride[ga], (const cpx<T>*)chi, chi_kind, (const cpx<T>*)u,
                       (const cpx<double>*)phi, batch, P);
    hipLaunchKernelGGL((k_scatter<double, C, 4>), dim3(n[be], nt, batch), dim3(kT), 0, st, n[be], n[ga], 1,
                       (const cpx<double>*)nullptr, 0, (const cpx<double>*)P, ndir, (const cpx<double>*)tab, part);
    nsum = n[be];
  }
  const long long total = (long long)batch * ndir * C;
  const unsigned gf = (unsigned)std::min<long long>((total + kT / 64 - 1) / (kT / 64), 16384);
  hipLaunchKernelGGL(k_scatter_final, dim3(gf), dim3(kT), 0, st, total, nsum, (const cpx<double>*)part, dv,
                     (cpx<double>*)out);
}

}  // namespace

extern "C" {

size_t em3d_scatter_work_bytes(int nx, int ny, int nz, int batch, int ndir, int ncomp) {
  return ScatterWork(nx, ny, nz, batch, ndir, ncomp).total();
}

int em3d_scatter_integral(int nx, int ny, int nz, const void* ax, const void* ay, const void* az, int precision,
                          int batch, int ncomp, const void* chi, int chi_kind, const void* u, int ndir,
                          const double* dirs, int shared_axis, double k0, double dv, double* out, void* work,
                          hipStream_t stream) {
  if (!ax || !ay || !az || !chi || !u || !dirs || !out || !work) return set_error("em3d_scatter_integral: null argument");
  if (batch < 1 || ndir < 1 || nx < 1 || ny < 1 || nz < 1) return set_error("em3d_scatter_integral: bad sizes");
  if (ncomp != 1 && ncomp != 3) return set_error("em3d_scatter_integral: ncomp must be 1 or 3");
  if (chi_kind < 1 || chi_kind > 3 || (ncomp == 1 && chi_kind != 1))
    return set_error("em3d_scatter_integral: chi_kind must be 1 (scalar), or 2, 3 with ncomp 3");
  if (shared_axis < -1 || shared_axis > 2) return set_error("em3d_scatter_integral: shared_axis must be -1, 0, 1, 2");
  if (ndir > 65535 * kDT || batch > 65535) return set_error("em3d_scatter_integral: too many directions or fields");
  const ScatterWork w(nx, ny, nz, batch, ndir, ncomp);
  const void* const axes[3] = {ax, ay, az};
#define E3_SCATTER(TT, CC)                                                                                      \
  launch_scatter<TT, CC>(stream, nx, ny, nz, batch, chi, chi_kind, u, ndir, shared_axis, axes, dirs, k0, dv, out, \
                         w, (char*)work)
  if (precision == 1) {
    if (ncomp == 3) E3_SCATTER(double, 3);
    else E3_SCATTER(double, 1);
  } else {
    if (ncomp == 3) E3_SCATTER(float, 3);
    else E3_SCATTER(float, 1);
  }
#undef E3_SCATTER
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_scatter_integral: ") + hipGetErrorString(e));
}

}  // extern "C"
