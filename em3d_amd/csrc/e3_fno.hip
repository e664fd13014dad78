// Fourier neural operator (FNO) used as a learned right preconditioner of
// the Krylov solvers: forward, backward and Adam kernels, and the glue that
// maps a fine-grid residual to the network's coarse grid and back.
//
// No counterpart in the reference (its solvers are unpreconditioned,
// src/em3d/solvers/bicgstab.py:12-211).
//
// Network on a G^3 coarse grid, channels last (cell-major, channel-minor).
// It is linear in the residual and conditioned on the material through
// multiplicative gates, so that the preconditioner is a fixed linear map for
// a given operator (BiCGStab's recurrences assume one):
//   v0      = W_in x                                   (in -> H; x: residual channels)
//   g_l     = sigmoid(U_l m + d_l)                     (material channels -> H)
//   z_l     = W_l v_l + S_l v_l
//   v_{l+1} = g_l * z_l                                (l < layers)
//   y       = W_out v_L                                (H -> out)
//   S_l v   = Re( E+ (R_l . (E- v)) ) / G^3
// E- is the forward DFT restricted to the modes k in [-M/2, M/2) per axis
// (M^3 modes), E+ the matching inverse sum, R_l(k) a complex H x H matrix
// per mode.  Pointwise channel mixing (W_in, W_l, U_l, W_out and the
// transposed products of the backward pass) runs on bf16 MFMA with fp32
// accumulation; parameters are fp32 masters with a bf16 copy the forward
// pass reads.
//
// Preconditioner:  M r = r + P( y(R r, m) ): R averages f^3 fine cells into
// one coarse cell (scaled by s = |r| / sqrt(n), undone after), P is the
// trilinear interpolation back to the fine grid; m holds the coarse means of
// the diagonal of chi and the wave-number feature.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "e3_common.hpp"
#include "e3_error.hpp"
#include "e3_fno.hpp"

using namespace e3;

namespace {

constexpr int kT = 256;
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

unsigned grid1(long long n) {
  const long long b = (n + kT - 1) / kT;
  return (unsigned)(b < 1 ? 1 : (b > 1048576 ? 1048576 : b));
}

// ---- pointwise channel mixing on MFMA ---------------------------------------
// z[r][o] = sum_{i<K} A[r * lda + i] W[o][i] + bias[o] + add[r][o] for r < rows,
// W bf16 row-major [cout][K] (transpose: W given as [K][cout], i.e. the
// product with W^T).  Y = z (act 0), gate[r][o] * z (act 1) or sigmoid(z)
// (act 2); Z (optional) receives z.  One wave computes a 16-row tile for all
// output columns, K in steps of 16: v_mfma_f32_16x16x16_bf16 with lane l
// holding A[l%16][4(l/16)+i], B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16].
__global__ __launch_bounds__(kT) void k_mix(long long rows, int K, int lda, int cout, const float* A,
                                            const unsigned short* W, int transpose, const float* bias,
                                            const float* add, const float* gate, float* Z, float* Y, int act,
                                            const int* skip) {
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long r0 = ((long long)blockIdx.x * (kT / 64) + wave) * 16;
  if (r0 >= rows) return;
  const int lr = lane & 15, lk = (lane >> 4) * 4;
  for (int c0 = 0; c0 < cout; c0 += 16) {
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
      s4v a, b;
      const long long ra = r0 + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = k0 + lk + i;
        const float av = (ra < rows && k < K) ? A[ra * lda + k] : 0.f;
        a[i] = (short)to_bf16(av);
        const int c = c0 + lr;
        unsigned short bv = 0;
        if (c < cout && k < K) bv = transpose ? W[(size_t)k * cout + c] : W[(size_t)c * K + k];
        b[i] = (short)bv;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int c = c0 + lr;
    if (c < cout) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long long r = r0 + lk + i;
        if (r >= rows) continue;
        const long long o = r * cout + c;
        const float z = acc[i] + (bias ? bias[c] : 0.f) + (add ? add[o] : 0.f);
        if (Z) Z[o] = z;
        Y[o] = act == 1 ? gate[o] * z : (act == 2 ? 1.f / (1.f + expf(-z)) : z);
      }
    }
  }
}

// ---- truncated DFT along one axis --------------------------------------------
// in  [outer][n_in][inner], out [outer][n_out][inner] (complex as float2 or
// real).  dir 0: input index = position x < G, output index = mode m < M;
// dir 1: input = mode, output = position.  Term: exp(sign 2 pi i k x / G),
// k = m < M/2 ? m : m - M.  out_real: out = scale Re(sum), else scale sum.
__global__ __launch_bounds__(kT) void k_dft(long long outer, int n_in, int n_out, int inner, int G, int M, int dir,
                                            int sign, const float* in, int in_real, float* out, int out_real,
                                            float scale, const float2* tw, const int* skip) {
  if (skip && *skip) return;
  const long long total = outer * n_out * inner;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % inner);
    const long long q = t / inner;
    const int j = (int)(q % n_out);
    const long long o = q / n_out;
    float re = 0.f, im = 0.f;
    for (int i = 0; i < n_in; ++i) {
      const int m = dir == 0 ? j : i, x = dir == 0 ? i : j;
      const int k = m < M / 2 ? m : m - M;
      int e = (k * x) % G;
      if (e < 0) e += G;
      const float2 w = tw[e];                       // exp(-2 pi i e / G)
      const float wi = sign > 0 ? -w.y : w.y;
      const long long src = (o * n_in + i) * inner + c;
      if (in_real) {
        const float v = in[src];
        re += v * w.x;
        im += v * wi;
      } else {
        const float2 v = ((const float2*)in)[src];
        re += v.x * w.x - v.y * wi;
        im += v.x * wi + v.y * w.x;
      }
    }
    const long long dst = (o * n_out + j) * inner + c;
    if (out_real) out[dst] = scale * re;
    else ((float2*)out)[dst] = make_float2(scale * re, scale * im);
  }
}

// ---- spectral mixing per mode --------------------------------------------------
// forward: B[s][m][o] = sum_i R[m][o][i] A[s][m][i];  adjoint: B[s][m][i] = sum_o conj(R[m][o][i]) A[s][m][o]
// R complex bf16 pairs [modes][H][H][2]
__global__ __launch_bounds__(kT) void k_specmix(int S, int modes, int H, const float2* A, const unsigned short* R,
                                                float2* B, int adjoint, const int* skip) {
  if (skip && *skip) return;
  const long long total = (long long)S * modes * H;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int o = (int)(t % H);
    const long long sm = t / H;
    const int m = (int)(sm % modes);
    const float2* a = A + sm * H;
    float re = 0.f, im = 0.f;
    for (int i = 0; i < H; ++i) {
      const size_t ri = adjoint ? (((size_t)m * H + i) * H + o) : (((size_t)m * H + o) * H + i);
      const float wr = from_bf16(R[2 * ri]), wi = adjoint ? -from_bf16(R[2 * ri + 1]) : from_bf16(R[2 * ri + 1]);
      const float2 v = a[i];
      re += wr * v.x - wi * v.y;
      im += wr * v.y + wi * v.x;
    }
    B[sm * H + o] = make_float2(re, im);
  }
}

// gR[m][o][i] += sum_s gB[s][m][o] conj(A[s][m][i])  (fp32 pairs)
__global__ __launch_bounds__(kT) void k_specgrad(int S, int modes, int H, const float2* gB, const float2* A, float* gR) {
  const long long total = (long long)modes * H * H;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % H);
    const long long q = t / H;
    const int o = (int)(q % H);
    const int m = (int)(q / H);
    float re = 0.f, im = 0.f;
    for (int s = 0; s < S; ++s) {
      const float2 g = gB[((size_t)s * modes + m) * H + o], a = A[((size_t)s * modes + m) * H + i];
      re += g.x * a.x + g.y * a.y;
      im += g.y * a.x - g.x * a.y;
    }
    gR[2 * t] += re;
    gR[2 * t + 1] += im;
  }
}

// gW[o][i] += sum_r dZ[r][o] X[r * ldx + i], gb[o] += sum_r dZ[r][o] (gb
// optional): a block sums a chunk of rows staged in LDS, then adds to the
// gradient with atomics
constexpr int kRowChunk = 64;
__global__ __launch_bounds__(kT) void k_wgrad(long long rows, int K, int ldx, int cout, const float* dZ,
                                              const float* X, float* gW, float* gb) {
  __shared__ float sz[kRowChunk * 32];
  __shared__ float sx[kRowChunk * 32];
  const long long r0 = (long long)blockIdx.x * kRowChunk;
  const int nr = (int)((rows - r0) < kRowChunk ? (rows - r0) : kRowChunk);
  for (int t = threadIdx.x; t < nr * cout; t += blockDim.x) sz[t] = dZ[r0 * cout + t];
  for (int t = threadIdx.x; t < nr * K; t += blockDim.x) sx[t] = X[(r0 + t / K) * ldx + t % K];
  __syncthreads();
  const int ncol = gb ? K + 1 : K;
  for (int p = threadIdx.x; p < cout * ncol; p += blockDim.x) {
    const int o = p / ncol, i = p % ncol;
    float acc = 0.f;
    if (i < K)
      for (int r = 0; r < nr; ++r) acc += sz[r * cout + o] * sx[r * K + i];
    else
      for (int r = 0; r < nr; ++r) acc += sz[r * cout + o];
    if (i < K) atomicAdd(&gW[o * K + i], acc);
    else atomicAdd(&gb[o], acc);
  }
}

// gate backward: dz = d * g (in place in d), dgp = d * z * g (1 - g)
__global__ __launch_bounds__(kT) void k_gate_bwd(long long n, const float* g, const float* z, float* d, float* dgp) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const float gv = g[t], dv = d[t];
    dgp[t] = dv * z[t] * gv * (1.f - gv);
    d[t] = dv * gv;
  }
}

__global__ __launch_bounds__(kT) void k_add(long long n, float* y, const float* x) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    y[t] += x[t];
}

// squared error per sample: partial[s][blk], dy = 2 (y - t) / (|t_s|^2 S)
__global__ __launch_bounds__(kT) void k_mse(int S, long long per, const float* y, const float* t, const double* tn2,
                                            float* dy, double* part) {
  const int s = blockIdx.y;
  double acc = 0.0;
  const double w = 2.0 / (tn2[s] * S);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += (long long)gridDim.x * blockDim.x) {
    const long long k = (long long)s * per + i;
    const float d = y[k] - t[k];
    acc += (double)d * d;
    if (dy) dy[k] = (float)(w * d);
  }
  __shared__ double sh[kT];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kT) void k_sumsq(int S, long long per, const float* t, double* part) {
  const int s = blockIdx.y;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += (long long)gridDim.x * blockDim.x) {
    const float v = t[(long long)s * per + i];
    acc += (double)v * v;
  }
  __shared__ double sh[kT];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = sh[0];
}

// out[s] = sum of the sample's partials; total += out[s] / denom[s]
__global__ void k_rowsum(int S, int nb, const double* part, double* out, double* total, const double* denom) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double a = 0.0;
  for (int b = 0; b < nb; ++b) a += part[(size_t)s * nb + b];
  out[s] = a;
  if (total) atomicAdd(total, denom ? a / denom[s] : a);
}

__global__ __launch_bounds__(kT) void k_adam(long long n, float* p, unsigned short* pb, const float* g, float* m,
                                             float* v, float lr, float b1, float b2, float eps, float wd, float c1,
                                             float c2) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const float gt = g[t] + wd * p[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    const float q = p[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
    p[t] = q;
    pb[t] = to_bf16(q);
  }
}

__global__ __launch_bounds__(kT) void k_tobf16(long long n, const float* p, unsigned short* pb) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    pb[t] = to_bf16(p[t]);
}

// ---- fine <-> coarse glue (preconditioner) -------------------------------------
// per-RHS sum |r|^2 partials over (3 or 1) x N^3 complex values
template <typename T>
__global__ __launch_bounds__(kT) void k_norm2(long long n, const cpx<T>* r, double* part, const int* skip) {
  if (skip && *skip) return;
  const int b = blockIdx.y;
  const cpx<T>* x = r + (size_t)b * n;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc += (double)x[i].re * x[i].re + (double)x[i].im * x[i].im;
  __shared__ double sh[kT];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)b * gridDim.x + blockIdx.x] = sh[0];
}

// scale[b] = sqrt(sum / n) (0 -> 1): one block per right-hand side sums its
// nb <= kT partials in a fixed tree order
__global__ __launch_bounds__(kT) void k_scale(int nb, long long n, const double* part, double* scale,
                                              const int* skip) {
  if (skip && *skip) return;
  const int b = blockIdx.x;
  __shared__ double sh[kT];
  sh[threadIdx.x] = (int)threadIdx.x < nb ? part[(size_t)b * nb + threadIdx.x] : 0.0;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x) return;
  const double s = sqrt(sh[0] / (double)n);
  scale[b] = s > 0.0 && isfinite(s) ? s : 1.0;
}

// features of coarse cell (b, X, Y, Z) for the two real samples of RHS b
// (sample 2b: real parts, 2b + 1: imaginary parts): the mean of the field /
// scale over the f^3 fine cells (C channels), then the nchi material
// features.  Running one real network on both samples makes the
// preconditioner complex-linear.
template <typename T>
__global__ __launch_bounds__(kT) void k_restrict(int B, int C, int nx, int ny, int nz, int f, const cpx<T>* r,
                                                 const double* scale, const float* chi_feat, int nchi, float* feat,
                                                 const int* skip) {
  if (skip && *skip) return;
  const int gx = nx / f, gy = ny / f, gz = nz / f;
  const long long cells = (long long)gx * gy * gz;
  const long long total = (long long)B * cells;
  const long long M = (long long)nx * ny * nz;
  const int nin = C + nchi;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / cells);
    const long long cell = t - (long long)b * cells;
    const int Z = (int)(cell % gz);
    const int Y = (int)((cell / gz) % gy);
    const int X = (int)(cell / ((long long)gz * gy));
    const float inv = (float)(1.0 / (scale[b] * (double)f * f * f));
    float* ore = feat + ((size_t)(2 * b) * cells + cell) * nin;
    float* oim = feat + ((size_t)(2 * b + 1) * cells + cell) * nin;
    for (int c = 0; c < C; ++c) {
      float re = 0.f, im = 0.f;
      const cpx<T>* rc = r + ((size_t)b * C + c) * M;
      for (int dx = 0; dx < f; ++dx)
        for (int dy = 0; dy < f; ++dy)
          for (int dz = 0; dz < f; ++dz) {
            const cpx<T> v = rc[((long long)(X * f + dx) * ny + (Y * f + dy)) * nz + (Z * f + dz)];
            re += (float)v.re;
            im += (float)v.im;
          }
      ore[c] = re * inv;
      oim[c] = im * inv;
    }
    for (int k = 0; k < nchi; ++k) ore[C + k] = oim[C + k] = chi_feat[cell * nchi + k];
  }
}

// Trilinear prolongation P: the coarse value of cell X sits at fine
// position (X + 1/2) f - 1/2; a fine cell takes the trilinear blend of the
// 8 surrounding coarse cells (clamped at the faces, weights summing to 1).
struct Lerp {
  int i0, i1;
  float w0, w1;
};
__device__ __forceinline__ Lerp lerp_axis(int x, int f, int G) {
  const float xi = ((float)x + 0.5f) / (float)f - 0.5f;
  int i0 = (int)floorf(xi);
  float w1 = xi - (float)i0;
  int i1 = i0 + 1;
  if (i0 < 0) { i0 = 0; w1 = 0.f; }
  if (i1 > G - 1) { i1 = G - 1; }
  Lerp l;
  l.i0 = i0; l.i1 = i1; l.w0 = 1.f - w1; l.w1 = w1;
  return l;
}

// sum_corners w y[cell][c] for the real and imaginary samples (2b, 2b + 1)
__device__ __forceinline__ void blend(const float* y, int b, long long cells, int C, int c, int gy, int gz, Lerp lx,
                                      Lerp ly, Lerp lz, float* qre, float* qim) {
  float re = 0.f, im = 0.f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int X = (a & 4) ? lx.i1 : lx.i0, Y = (a & 2) ? ly.i1 : ly.i0, Z = (a & 1) ? lz.i1 : lz.i0;
    const float w = ((a & 4) ? lx.w1 : lx.w0) * ((a & 2) ? ly.w1 : ly.w0) * ((a & 1) ? lz.w1 : lz.w0);
    const long long cell = ((long long)X * gy + Y) * gz + Z;
    re += w * y[((size_t)(2 * b) * cells + cell) * C + c];
    im += w * y[((size_t)(2 * b + 1) * cells + cell) * C + c];
  }
  *qre = re;
  *qim = im;
}

// out = base + scale P(y) (base may be null), y [2B][cells][C] (sample 2b:
// real parts, 2b + 1: imaginary parts)
template <typename T>
__global__ __launch_bounds__(kT) void k_prolong(int B, int C, int nx, int ny, int nz, int f, const cpx<T>* base,
                                                const double* scale, const float* y, cpx<T>* out, const int* skip) {
  if (skip && *skip) return;
  const long long M = (long long)nx * ny * nz;
  const long long total = (long long)B * C * M;
  const int gx = nx / f, gy = ny / f, gz = nz / f;
  const long long cells = (long long)gx * gy * gz;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long i = t % M;
    const long long bc = t / M;
    const int c = (int)(bc % C), b = (int)(bc / C);
    const int z = (int)(i % nz), yy = (int)((i / nz) % ny), x = (int)(i / ((long long)nz * ny));
    float qre, qim;
    blend(y, b, cells, C, c, gy, gz, lerp_axis(x, f, gx), lerp_axis(yy, f, gy), lerp_axis(z, f, gz), &qre, &qim);
    const double s = scale[b];
    const cpx<T> v = base ? base[t] : mk((T)0, (T)0);
    out[t] = mk((T)((double)v.re + s * qre), (T)((double)v.im + s * qim));
  }
}

// adjoint of P: out[2b + p][cell][c] = (sum over fine cells x of the weight of
// `cell` at x times Re / Im g_b(x)) / scale[b]; a coarse cell's support spans
// its own f cells and half a cell's worth on each side
template <typename T>
__global__ __launch_bounds__(kT) void k_prolong_t(int B, int C, int nx, int ny, int nz, int f, const cpx<T>* g,
                                                  const double* scale, float* out) {
  const int gx = nx / f, gy = ny / f, gz = nz / f;
  const long long cells = (long long)gx * gy * gz;
  const long long M = (long long)nx * ny * nz;
  const long long total = (long long)B * cells * C;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t % C);
    const long long q = t / C;
    const long long cell = q % cells;
    const int b = (int)(q / cells);
    const int Z = (int)(cell % gz), Y = (int)((cell / gz) % gy), X = (int)(cell / ((long long)gz * gy));
    const cpx<T>* gc = g + ((size_t)b * C + c) * M;
    float re = 0.f, im = 0.f;
    for (int x = (X - 1) * f; x < (X + 2) * f; ++x) {
      if (x < 0 || x >= nx) continue;
      const Lerp lx = lerp_axis(x, f, gx);
      const float wx = (lx.i0 == X ? lx.w0 : 0.f) + (lx.i1 == X ? lx.w1 : 0.f);
      if (wx == 0.f) continue;
      for (int yy = (Y - 1) * f; yy < (Y + 2) * f; ++yy) {
        if (yy < 0 || yy >= ny) continue;
        const Lerp ly = lerp_axis(yy, f, gy);
        const float wy = (ly.i0 == Y ? ly.w0 : 0.f) + (ly.i1 == Y ? ly.w1 : 0.f);
        if (wy == 0.f) continue;
        for (int z = (Z - 1) * f; z < (Z + 2) * f; ++z) {
          if (z < 0 || z >= nz) continue;
          const Lerp lz = lerp_axis(z, f, gz);
          const float wz = (lz.i0 == Z ? lz.w0 : 0.f) + (lz.i1 == Z ? lz.w1 : 0.f);
          if (wz == 0.f) continue;
          const cpx<T> v = gc[((long long)x * ny + yy) * nz + z];
          const float w = wx * wy * wz;
          re += w * (float)v.re;
          im += w * (float)v.im;
        }
      }
    }
    const float inv = (float)(1.0 / scale[b]);
    out[((size_t)(2 * b) * cells + cell) * C + c] = re * inv;
    out[((size_t)(2 * b + 1) * cells + cell) * C + c] = im * inv;
  }
}

// chi features of coarse cells: mean of the diagonal of chi (re, im of each
// of the 3 diagonal entries; 1 entry for the acoustic / isotropic case)
template <typename T>
__global__ __launch_bounds__(kT) void k_chi_feat(int nx, int ny, int nz, int f, const cpx<T>* chi, int chi_kind,
                                                 int nchi, float extra, float* out) {
  const int gx = nx / f, gy = ny / f, gz = nz / f;
  const long long cells = (long long)gx * gy * gz;
  const long long M = (long long)nx * ny * nz;
  for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell < cells;
       cell += (long long)gridDim.x * blockDim.x) {
    const int Z = (int)(cell % gz), Y = (int)((cell / gz) % gy), X = (int)(cell / ((long long)gz * gy));
    const int nd = nchi / 2;
    for (int a = 0; a < nd; ++a) {
      float re = 0.f, im = 0.f;
      for (int dx = 0; dx < f; ++dx)
        for (int dy = 0; dy < f; ++dy)
          for (int dz = 0; dz < f; ++dz) {
            const long long i = ((long long)(X * f + dx) * ny + (Y * f + dy)) * nz + (Z * f + dz);
            cpx<T> v = mk((T)0, (T)0);
            if (chi) {
              if (chi_kind == 1) v = chi[i];
              else if (chi_kind == 2) v = chi[(size_t)(a < 3 ? a : 0) * M + i];
              else v = chi[(size_t)(4 * a) * M + i];
            }
            re += (float)v.re;
            im += (float)v.im;
          }
      const float inv = 1.f / (float)(f * f * f);
      out[cell * nchi + 2 * a] = re * inv;
      out[cell * nchi + 2 * a + 1] = im * inv;
    }
    if (nchi & 1) out[cell * nchi + nchi - 1] = extra;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace e3 {

std::string fno_check(const em3d_fno_config* c) {
  if (!c) return "null config";
  if (c->grid < 2 || c->channels < 1 || c->channels > 32 || c->modes < 2 || (c->modes & 1) || c->modes > c->grid ||
      c->layers < 0 || c->layers > kFnoMaxLayers || c->in_channels < 1 || c->in_channels > 32 ||
      c->out_channels < 1 || c->out_channels > 32 || c->material_channels < 0 || c->material_channels > 32)
    return "invalid FNO config (grid >= 2, 1 <= channels/in/out <= 32, material <= 32, even modes <= grid, "
           "layers <= 16)";
  return "";
}

FnoLayout fno_layout(const em3d_fno_config& c, int S) {
  FnoLayout L{};
  const long long G = c.grid, H = c.channels, M = c.modes, nm = c.material_channels;
  L.cells = G * G * G;
  L.modes3 = M * M * M;
  L.ldx = c.in_channels + c.material_channels;
  long long o = 0;
  L.w_in = o; o += H * c.in_channels;
  for (int l = 0; l < c.layers; ++l) {
    L.w[l] = o; o += H * H;
    L.u[l] = o; o += H * nm;
    L.d[l] = o; o += H;
    L.r[l] = o; o += L.modes3 * H * H * 2;
  }
  L.w_out = o; o += (long long)c.out_channels * H;
  L.n_params = o;
  // work (floats)
  long long w = 0;
  auto take = [&](long long n) { const long long at = w; w += (n + 63) / 64 * 64; return at; };
  const long long ch = L.cells * H;
  for (int l = 0; l <= c.layers; ++l) L.act[l] = take(S * ch);
  for (int l = 0; l < c.layers; ++l) L.zpre[l] = take(S * ch);
  for (int l = 0; l < c.layers; ++l) L.gate[l] = take(S * ch);
  L.a1 = take(S * G * G * M * H * 2);
  L.a2 = take(S * G * M * M * H * 2);
  for (int l = 0; l < c.layers; ++l) L.a3[l] = take(S * L.modes3 * H * 2);
  L.b3 = take(S * L.modes3 * H * 2);
  L.spec = take(S * ch);
  L.d0 = take(S * ch);
  L.d1 = take(S * ch);
  L.part = take(2 * (S * 1024 + 64));
  L.tw = take(2 * G);
  L.work_floats = w;
  return L;
}

namespace {

struct Ctx {
  const em3d_fno_config& c;
  const FnoLayout& L;
  float* w;
  int S;
  hipStream_t st;
  const int* skip;
  hipError_t err = hipSuccess;
  void note() {
    hipError_t e = hipGetLastError();
    if (err == hipSuccess) err = e;
  }
  float* at(long long off) const { return w + off; }
  const float2* tw() const { return (const float2*)(w + L.tw); }
};

void mix(Ctx& x, long long rows, int K, int lda, int cout, const float* A, const unsigned short* W, int transpose,
         const float* bias, const float* add, const float* gate, float* Z, float* Y, int act) {
  const unsigned nb = (unsigned)((rows + 63) / 64);
  hipLaunchKernelGGL(k_mix, dim3(nb), dim3(kT), 0, x.st, rows, K, lda, cout, A, W, transpose, bias, add, gate, Z, Y,
                     act, x.skip);
  x.note();
}

void wgrad(Ctx& x, long long rows, int K, int ldx, int cout, const float* dZ, const float* X, float* gW, float* gb) {
  const unsigned nb = (unsigned)((rows + kRowChunk - 1) / kRowChunk);
  hipLaunchKernelGGL(k_wgrad, dim3(nb), dim3(kT), 0, x.st, rows, K, ldx, cout, dZ, X, gW, gb);
  x.note();
}

void dft(Ctx& x, long long outer, int n_in, int n_out, int inner, int dir, int sign, const float* in, int in_real,
         float* out, int out_real, float scale) {
  const long long total = outer * n_out * inner;
  hipLaunchKernelGGL(k_dft, dim3(grid1(total)), dim3(kT), 0, x.st, outer, n_in, n_out, inner, x.c.grid, x.c.modes, dir,
                     sign, in, in_real, out, out_real, scale, x.tw(), x.skip);
  x.note();
}

// spectral convolution S v (v real [S][G^3][H]) -> out real, keeping the
// forward spectrum (E- v) in a3 for the backward pass
void spectral(Ctx& x, int l, const float* v, const unsigned short* R, float* out, float* a3) {
  const int G = x.c.grid, M = x.c.modes, H = x.c.channels, S = x.S;
  float* a1 = x.at(x.L.a1);
  float* a2 = x.at(x.L.a2);
  float* b3 = x.at(x.L.b3);
  // forward: z, y, x
  dft(x, (long long)S * G * G, G, M, H, 0, -1, v, 1, a1, 0, 1.f);
  dft(x, (long long)S * G, G, M, M * H, 0, -1, a1, 0, a2, 0, 1.f);
  dft(x, (long long)S, G, M, M * M * H, 0, -1, a2, 0, a3, 0, 1.f);
  hipLaunchKernelGGL(k_specmix, dim3(grid1((long long)S * x.L.modes3 * H)), dim3(kT), 0, x.st, S, (int)x.L.modes3, H,
                     (const float2*)a3, R, (float2*)b3, 0, x.skip);
  x.note();
  // inverse: x, y, z (real part, 1 / G^3)
  dft(x, (long long)S, M, G, M * M * H, 1, 1, b3, 0, a2, 0, 1.f);
  dft(x, (long long)S * G, M, G, M * H, 1, 1, a2, 0, a1, 0, 1.f);
  dft(x, (long long)S * G * G, M, G, H, 1, 1, a1, 0, out, 1, 1.f / ((float)G * G * G));
  (void)l;
}

}  // namespace

int fno_forward_impl(const em3d_fno_config& c, const FnoLayout& L, const unsigned short* pb, const float* params,
                     const float* xin, float* y, float* work, int S, hipStream_t st, const int* skip) {
  Ctx x{c, L, work, S, st, skip};
  const long long rows = (long long)S * L.cells;
  const int H = c.channels, ldx = (int)L.ldx, nm = c.material_channels;
  mix(x, rows, c.in_channels, ldx, H, xin, pb + L.w_in, 0, nullptr, nullptr, nullptr, nullptr, x.at(L.act[0]), 0);
  for (int l = 0; l < c.layers; ++l) {
    // gates from the material channels (zero channels: sigmoid(d_l))
    mix(x, rows, nm, ldx, H, xin + c.in_channels, pb + L.u[l], 0, params + L.d[l], nullptr, nullptr, nullptr,
        x.at(L.gate[l]), 2);
    spectral(x, l, x.at(L.act[l]), pb + L.r[l], x.at(L.spec), x.at(L.a3[l]));
    mix(x, rows, H, H, H, x.at(L.act[l]), pb + L.w[l], 0, nullptr, x.at(L.spec), x.at(L.gate[l]), x.at(L.zpre[l]),
        x.at(L.act[l + 1]), 1);
  }
  mix(x, rows, H, H, c.out_channels, x.at(L.act[c.layers]), pb + L.w_out, 0, nullptr, nullptr, nullptr, nullptr, y, 0);
  return x.err == hipSuccess ? 0 : set_error(std::string("em3d_fno_forward: ") + hipGetErrorString(x.err));
}

int fno_backward_impl(const em3d_fno_config& c, const FnoLayout& L, const unsigned short* pb, const float* xin,
                      const float* dy, float* grad, float* work, int S, hipStream_t st) {
  Ctx x{c, L, work, S, st, nullptr};
  const long long rows = (long long)S * L.cells;
  const int H = c.channels, G = c.grid, M = c.modes, ldx = (int)L.ldx, nm = c.material_channels;
  wgrad(x, rows, H, H, c.out_channels, dy, x.at(L.act[c.layers]), grad + L.w_out, nullptr);
  float* d = x.at(L.d0);   // dL/d act[l]
  float* e = x.at(L.d1);
  mix(x, rows, c.out_channels, c.out_channels, H, dy, pb + L.w_out, 1, nullptr, nullptr, nullptr, nullptr, d, 0);
  float* a1 = x.at(L.a1);
  float* a2 = x.at(L.a2);
  float* b3 = x.at(L.b3);
  float* dgp = x.at(L.spec);
  for (int l = c.layers - 1; l >= 0; --l) {
    // d <- dz = d * g, dgp = d * z * g (1 - g)
    hipLaunchKernelGGL(k_gate_bwd, dim3(grid1(rows * H)), dim3(kT), 0, st, rows * H, x.at(L.gate[l]),
                       x.at(L.zpre[l]), d, dgp);
    x.note();
    if (nm > 0) wgrad(x, rows, nm, ldx, H, dgp, xin + c.in_channels, grad + L.u[l], grad + L.d[l]);
    else wgrad(x, rows, 0, ldx, H, dgp, xin, grad + L.u[l], grad + L.d[l]);
    wgrad(x, rows, H, H, H, d, x.at(L.act[l]), grad + L.w[l], nullptr);
    // spectral adjoint: gB = E-(dz) / G^3, gR += gB conj(A), gA = R^H gB, e = Re(E+ gA)
    dft(x, (long long)S * G * G, G, M, H, 0, -1, d, 1, a1, 0, 1.f / ((float)G * G * G));
    dft(x, (long long)S * G, G, M, M * H, 0, -1, a1, 0, a2, 0, 1.f);
    dft(x, (long long)S, G, M, M * M * H, 0, -1, a2, 0, b3, 0, 1.f);
    hipLaunchKernelGGL(k_specgrad, dim3(grid1(L.modes3 * H * H)), dim3(kT), 0, st, S, (int)L.modes3, H,
                       (const float2*)b3, (const float2*)x.at(L.a3[l]), grad + L.r[l]);
    x.note();
    float* ga = x.at(L.a3[l]);   // the forward spectrum is no longer needed: reuse for R^H gB
    hipLaunchKernelGGL(k_specmix, dim3(grid1((long long)S * L.modes3 * H)), dim3(kT), 0, st, S, (int)L.modes3, H,
                       (const float2*)b3, pb + L.r[l], (float2*)ga, 1, (const int*)nullptr);
    x.note();
    dft(x, (long long)S, M, G, M * M * H, 1, 1, ga, 0, a2, 0, 1.f);
    dft(x, (long long)S * G, M, G, M * H, 1, 1, a2, 0, a1, 0, 1.f);
    dft(x, (long long)S * G * G, M, G, H, 1, 1, a1, 0, e, 1, 1.f);
    // e += W^T dz  (the pointwise path), then e is dL/d act[l]
    mix(x, rows, H, H, H, d, pb + L.w[l], 1, nullptr, e, nullptr, nullptr, e, 0);
    float* t = d;
    d = e;
    e = t;
  }
  wgrad(x, rows, c.in_channels, ldx, H, d, xin, grad + L.w_in, nullptr);
  return x.err == hipSuccess ? 0 : set_error(std::string("em3d_fno_backward: ") + hipGetErrorString(x.err));
}

int fno_twiddles(const em3d_fno_config& c, const FnoLayout& L, float* work, hipStream_t st) {
  std::vector<float> tw(2 * (size_t)c.grid);
  for (int t = 0; t < c.grid; ++t) {
    const double a = -2.0 * 3.14159265358979323846 * t / c.grid;
    tw[2 * t] = (float)std::cos(a);
    tw[2 * t + 1] = (float)std::sin(a);
  }
  hipError_t e = hipMemcpyAsync(work + L.tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno twiddles: ") + hipGetErrorString(e));
}

int fno_precond_apply(const em3d_fno_model* m, int B, int C, int nx, int ny, int nz, int prec, const void* r, void* out,
                      const int* skip, hipStream_t st) {
  const em3d_fno_config& c = m->config;
  const FnoLayout L = fno_layout(c, 2 * B);   // two real samples per right-hand side
  float* w = (float*)m->work;
  const long long n = (long long)C * nx * ny * nz;
  const unsigned nb = 256;
  double* part = (double*)(w + L.part);
  double* scale = part + (size_t)B * nb;
  if (prec == 1)
    hipLaunchKernelGGL(k_norm2<double>, dim3(nb, B), dim3(kT), 0, st, n, (const cpx<double>*)r, part, skip);
  else
    hipLaunchKernelGGL(k_norm2<float>, dim3(nb, B), dim3(kT), 0, st, n, (const cpx<float>*)r, part, skip);
  hipLaunchKernelGGL(k_scale, dim3(B), dim3(kT), 0, st, (int)nb, n, part, scale, skip);
  const long long cells = L.cells;
  float* feat = (float*)m->features;
  const int f = m->factor;
  if (prec == 1)
    hipLaunchKernelGGL(k_restrict<double>, dim3(grid1(B * cells)), dim3(kT), 0, st, B, C, nx, ny, nz, f,
                       (const cpx<double>*)r, scale, m->chi_features, m->chi_channels, feat, skip);
  else
    hipLaunchKernelGGL(k_restrict<float>, dim3(grid1(B * cells)), dim3(kT), 0, st, B, C, nx, ny, nz, f,
                       (const cpx<float>*)r, scale, m->chi_features, m->chi_channels, feat, skip);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(std::string("fno precond: ") + hipGetErrorString(e));
  float* y = feat + 2 * B * cells * L.ldx;
  if (fno_forward_impl(c, L, m->params_bf16, m->params, feat, y, w, 2 * B, st, skip)) return 1;
  if (prec == 1)
    hipLaunchKernelGGL(k_prolong<double>, dim3(grid1(B * n)), dim3(kT), 0, st, B, C, nx, ny, nz, f,
                       (const cpx<double>*)r, scale, y, (cpx<double>*)out, skip);
  else
    hipLaunchKernelGGL(k_prolong<float>, dim3(grid1(B * n)), dim3(kT), 0, st, B, C, nx, ny, nz, f,
                       (const cpx<float>*)r, scale, y, (cpx<float>*)out, skip);
  e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("fno precond: ") + hipGetErrorString(e));
}

}  // namespace e3

extern "C" {

long long em3d_fno_param_count(const em3d_fno_config* c) {
  if (!fno_check(c).empty()) return -1;
  return fno_layout(*c, 1).n_params;
}

size_t em3d_fno_work_bytes(const em3d_fno_config* c, int samples) {
  if (!fno_check(c).empty() || samples < 1) return 0;
  return (size_t)fno_layout(*c, samples).work_floats * sizeof(float);
}

int em3d_fno_init_work(const em3d_fno_config* c, int samples, void* work, hipStream_t stream) {
  const std::string why = fno_check(c);
  if (!why.empty()) return set_error("em3d_fno_init_work: " + why);
  if (!work || samples < 1) return set_error("em3d_fno_init_work: null work or samples < 1");
  return fno_twiddles(*c, fno_layout(*c, samples), (float*)work, stream);
}

int em3d_fno_forward(const em3d_fno_config* c, const float* params, const unsigned short* params_bf16,
                     const float* x, float* y, void* work, int samples, hipStream_t stream) {
  const std::string why = fno_check(c);
  if (!why.empty()) return set_error("em3d_fno_forward: " + why);
  if (!params || !params_bf16 || !x || !y || !work || samples < 1) return set_error("em3d_fno_forward: null argument");
  return fno_forward_impl(*c, fno_layout(*c, samples), params_bf16, params, x, y, (float*)work, samples, stream,
                          nullptr);
}

int em3d_fno_backward(const em3d_fno_config* c, const unsigned short* params_bf16, const float* x, const float* dy,
                      float* grad, void* work, int samples, hipStream_t stream) {
  const std::string why = fno_check(c);
  if (!why.empty()) return set_error("em3d_fno_backward: " + why);
  if (!params_bf16 || !x || !dy || !grad || !work || samples < 1) return set_error("em3d_fno_backward: null argument");
  return fno_backward_impl(*c, fno_layout(*c, samples), params_bf16, x, dy, grad, (float*)work, samples, stream);
}

int em3d_fno_loss(int samples, long long per_sample, const float* y, const float* target, float* dy, double* loss,
                  double* work, hipStream_t stream) {
  if (!y || !target || !loss || !work || samples < 1 || per_sample < 1) return set_error("em3d_fno_loss: bad argument");
  const unsigned nb = 128;
  double* part = work;
  double* tn2 = work + (size_t)samples * nb;
  hipLaunchKernelGGL(k_sumsq, dim3(nb, samples), dim3(kT), 0, stream, samples, per_sample, target, part);
  hipLaunchKernelGGL(k_rowsum, dim3((samples + 63) / 64), dim3(64), 0, stream, samples, (int)nb, part, tn2,
                     (double*)nullptr, (const double*)nullptr);
  hipLaunchKernelGGL(k_mse, dim3(nb, samples), dim3(kT), 0, stream, samples, per_sample, y, target, tn2, dy, part);
  double* err2 = tn2 + samples;
  hipLaunchKernelGGL(k_rowsum, dim3((samples + 63) / 64), dim3(64), 0, stream, samples, (int)nb, part, err2, loss,
                     (const double*)tn2);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_loss: ") + hipGetErrorString(e));
}

size_t em3d_fno_loss_work_bytes(int samples) { return (size_t)(samples * 128 + 2 * samples) * sizeof(double); }

int em3d_fno_adam(long long n, float* params, unsigned short* params_bf16, const float* grad, float* m, float* v,
                  double lr, double beta1, double beta2, double eps, double weight_decay, int step,
                  hipStream_t stream) {
  if (!params || !params_bf16 || !grad || !m || !v || n < 1 || step < 1) return set_error("em3d_fno_adam: bad argument");
  const float c1 = (float)(1.0 - std::pow(beta1, step)), c2 = (float)(1.0 - std::pow(beta2, step));
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(kT), 0, stream, n, params, params_bf16, grad, m, v, (float)lr,
                     (float)beta1, (float)beta2, (float)eps, (float)weight_decay, c1, c2);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_adam: ") + hipGetErrorString(e));
}

int em3d_fno_to_bf16(long long n, const float* params, unsigned short* params_bf16, hipStream_t stream) {
  if (!params || !params_bf16 || n < 1) return set_error("em3d_fno_to_bf16: bad argument");
  hipLaunchKernelGGL(k_tobf16, dim3(grid1(n)), dim3(kT), 0, stream, n, params, params_bf16);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_to_bf16: ") + hipGetErrorString(e));
}

int em3d_fno_chi_features(int nx, int ny, int nz, int factor, int precision, const void* chi, int chi_kind,
                          int chi_channels, double extra, float* out, hipStream_t stream) {
  if (!out || factor < 1 || nx % factor || ny % factor || nz % factor ||
      (chi_channels != 2 && chi_channels != 3 && chi_channels != 6 && chi_channels != 7))
    return set_error("em3d_fno_chi_features: bad argument (factor must divide the grid; 2, 3, 6 or 7 channels)");
  const long long cells = (long long)(nx / factor) * (ny / factor) * (nz / factor);
  if (precision == 1)
    hipLaunchKernelGGL(k_chi_feat<double>, dim3(grid1(cells)), dim3(kT), 0, stream, nx, ny, nz, factor,
                       (const cpx<double>*)chi, chi_kind, chi_channels, (float)extra, out);
  else
    hipLaunchKernelGGL(k_chi_feat<float>, dim3(grid1(cells)), dim3(kT), 0, stream, nx, ny, nz, factor,
                       (const cpx<float>*)chi, chi_kind, chi_channels, (float)extra, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_chi_features: ") + hipGetErrorString(e));
}

size_t em3d_fno_features_bytes(const em3d_fno_config* c, int samples) {
  if (!fno_check(c).empty() || samples < 1) return 0;
  const long long cells = (long long)c->grid * c->grid * c->grid;
  return (size_t)samples * cells * (c->in_channels + c->material_channels + c->out_channels) * sizeof(float);
}

size_t em3d_fno_scale_work_bytes(int batch) { return (size_t)batch * 256 * sizeof(double); }

int em3d_fno_scale(int batch, long long n, int precision, const void* field, double* scale, void* work,
                   hipStream_t stream) {
  if (!field || !scale || !work || batch < 1 || n < 1) return set_error("em3d_fno_scale: bad argument");
  const unsigned nb = 256;
  if (precision == 1)
    hipLaunchKernelGGL(k_norm2<double>, dim3(nb, batch), dim3(kT), 0, stream, n, (const cpx<double>*)field,
                       (double*)work, (const int*)nullptr);
  else
    hipLaunchKernelGGL(k_norm2<float>, dim3(nb, batch), dim3(kT), 0, stream, n, (const cpx<float>*)field,
                       (double*)work, (const int*)nullptr);
  hipLaunchKernelGGL(k_scale, dim3(batch), dim3(kT), 0, stream, (int)nb, n, (const double*)work, scale,
                     (const int*)nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_scale: ") + hipGetErrorString(e));
}

int em3d_fno_restrict(int batch, int ncomp, int nx, int ny, int nz, int factor, int precision, const void* field,
                      const double* scale, const float* chi_features, int chi_channels, float* out,
                      hipStream_t stream) {
  if (!field || !scale || !out || batch < 1 || factor < 1 || nx % factor || ny % factor || nz % factor ||
      (chi_channels > 0 && !chi_features))
    return set_error("em3d_fno_restrict: bad argument (factor must divide the grid)");
  const long long cells = (long long)(nx / factor) * (ny / factor) * (nz / factor);
  if (precision == 1)
    hipLaunchKernelGGL(k_restrict<double>, dim3(grid1(batch * cells)), dim3(kT), 0, stream, batch, ncomp, nx, ny, nz,
                       factor, (const cpx<double>*)field, scale, chi_features, chi_channels, out, (const int*)nullptr);
  else
    hipLaunchKernelGGL(k_restrict<float>, dim3(grid1(batch * cells)), dim3(kT), 0, stream, batch, ncomp, nx, ny, nz,
                       factor, (const cpx<float>*)field, scale, chi_features, chi_channels, out, (const int*)nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_restrict: ") + hipGetErrorString(e));
}

int em3d_fno_prolong(int batch, int ncomp, int nx, int ny, int nz, int factor, int precision, const void* base,
                     const double* scale, const float* y, void* out, hipStream_t stream) {
  if (!scale || !y || !out || batch < 1 || factor < 1 || nx % factor || ny % factor || nz % factor)
    return set_error("em3d_fno_prolong: bad argument (factor must divide the grid)");
  const long long n = (long long)batch * ncomp * nx * ny * nz;
  if (precision == 1)
    hipLaunchKernelGGL(k_prolong<double>, dim3(grid1(n)), dim3(kT), 0, stream, batch, ncomp, nx, ny, nz, factor,
                       (const cpx<double>*)base, scale, y, (cpx<double>*)out, (const int*)nullptr);
  else
    hipLaunchKernelGGL(k_prolong<float>, dim3(grid1(n)), dim3(kT), 0, stream, batch, ncomp, nx, ny, nz, factor,
                       (const cpx<float>*)base, scale, y, (cpx<float>*)out, (const int*)nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_prolong: ") + hipGetErrorString(e));
}

int em3d_fno_prolong_adjoint(int batch, int ncomp, int nx, int ny, int nz, int factor, int precision, const void* g,
                             const double* scale, float* out, hipStream_t stream) {
  if (!g || !scale || !out || batch < 1 || factor < 1 || nx % factor || ny % factor || nz % factor)
    return set_error("em3d_fno_prolong_adjoint: bad argument (factor must divide the grid)");
  const long long n = (long long)batch * ncomp * (nx / factor) * (ny / factor) * (nz / factor);
  if (precision == 1)
    hipLaunchKernelGGL(k_prolong_t<double>, dim3(grid1(n)), dim3(kT), 0, stream, batch, ncomp, nx, ny, nz, factor,
                       (const cpx<double>*)g, scale, out);
  else
    hipLaunchKernelGGL(k_prolong_t<float>, dim3(grid1(n)), dim3(kT), 0, stream, batch, ncomp, nx, ny, nz, factor,
                       (const cpx<float>*)g, scale, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_error(std::string("em3d_fno_prolong_adjoint: ") + hipGetErrorString(e));
}

int em3d_fno_precondition(const em3d_fno_model* model, int batch, int ncomp, int nx, int ny, int nz, int precision,
                          const void* r, void* out, hipStream_t stream) {
  if (!model || !r || !out || batch < 1) return set_error("em3d_fno_precondition: null argument");
  const std::string why = fno_check(&model->config);
  if (!why.empty()) return set_error("em3d_fno_precondition: " + why);
  const int f = model->factor;
  if (f < 1 || nx != model->config.grid * f || ny != model->config.grid * f || nz != model->config.grid * f)
    return set_error("em3d_fno_precondition: the grid must be (config.grid * factor)^3");
  if (model->config.in_channels != ncomp || model->config.material_channels != model->chi_channels ||
      model->config.out_channels != ncomp)
    return set_error("em3d_fno_precondition: in_channels and out_channels must be ncomp, material_channels "
                     "chi_channels");
  return fno_precond_apply(model, batch, ncomp, nx, ny, nz, precision, r, out, nullptr, stream);
}

}  // extern "C"
