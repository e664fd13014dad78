// HIP kernels and launch helpers of the fast (power-of-two) passes.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "e3_fast.hpp"

namespace e3 {

template <class K>
__global__ __launch_bounds__(K::NT) void k_fast_s(SArgs a, long long ncb) {
  if (a.skip && *a.skip) return;
  using T = typename K::Real;
  __shared__ cpx<T> lds[K::LDS];
  const long long bid = blockIdx.x;
  const long long q = bid / ncb;
  K k(a, bid - q * ncb, q, lds);
  typename K::St st;
  k.stage_a(threadIdx.x, st);
  __syncthreads();
  k.stage_b(threadIdx.x, st);
}

// Phases 0 .. NPH-1 of a block, each a compile-time instance, barriers between.
template <class K, int... I>
__device__ __forceinline__ void run_phases(K& k, int h, typename K::St& st, std::integer_sequence<int, I...>) {
  ((k.template phase<I>(threadIdx.x, h, st), (I + 1 < K::NPH ? __syncthreads() : (void)0)), ...);
}

template <class K, int C>
__global__ __launch_bounds__(K::NT, K::MIN_WAVES_PER_EU) void k_fast_x(XArgs a, long long ncb) {
  if (a.skip && *a.skip) return;
  using T = typename K::Real;
  __shared__ cpx<T> lds[K::LDS];
  const long long bid = blockIdx.x;
  const int b = (int)(bid / ncb);
  K k(a, bid - (long long)b * ncb, b, lds);
  typename K::St st;
  const int h0 = k.row_begin(), h1 = k.row_end();
  if (h0 >= h1) return;   // block-uniform: the row does not exist
  k.h_end = h1;
  if constexpr (K::PREFETCH) k.load_a(0, threadIdx.x, h0, st);
#pragma unroll 1
  for (int h = h0; h < h1; ++h) {
    if (h != h0) __syncthreads();
    run_phases(k, h, st, std::make_integer_sequence<int, K::NPH>{});
  }
}

#ifndef E3_ZF_SPLIT
#define E3_ZF_SPLIT 0   // 1: one block per (component, pencil group) in single precision too
#endif

// forward z blocks per (component, pencil group): FP64 (measured faster there), or all with E3_ZF_SPLIT
template <class K>
constexpr bool zf_split() {
  return E3_ZF_SPLIT || sizeof(typename K::Real) == 8;
}

// Block sum of the fused Krylov partials (every thread of the block calls
// it): thread 0 writes the block's record, block 0 the blocks per right-hand
// side.  The grid holds gridDim.x / batch consecutive blocks per right-hand side.
__device__ __forceinline__ void kry_block_store(const ZArgs& a, KAcc& acc) {
  __shared__ double red[16][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double v = acc.v[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[w][i] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const long long per = gridDim.x / a.batch;
    const long long b = blockIdx.x / per;
    double* o = a.k.part + ((size_t)b * a.k.part_stride + (blockIdx.x - b * per)) * 4;
    const int nw = (int)((blockDim.x + 63) >> 6);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
      for (int k = 0; k < nw; ++k) s += red[k][i];
      o[i] = s;
    }
    if (blockIdx.x == 0) *a.k.count = (int)per;
  }
}

// KRY: the forward pass with a fused Krylov prologue (ZArgs::k).  SPLIT (or
// zf_split): one block per (pencil group, component).
template <class K, bool KRY, bool SPLIT>
__global__ __launch_bounds__(K::NT) void k_fast_zf(ZArgs a, long long nblk) {
  if (a.skip && *a.skip) return;
  static_assert(!KRY || K::NT % 64 == 0, "fused reductions take whole wavefronts");
  using T = typename K::Real;
  __shared__ cpx<T> lds[K::LDS];
  KAcc acc{};
  if constexpr (SPLIT || zf_split<K>()) {
    const long long bid = blockIdx.x / a.C;
    const int c = (int)(blockIdx.x - bid * a.C);
    const int b = (int)(bid / nblk);
    K k(a, bid - (long long)b * nblk, b, lds);
    if constexpr (KRY) k.stage_a(c, threadIdx.x, acc);
    else k.stage_a(c, threadIdx.x);
    __syncthreads();
    k.stage_b(c, threadIdx.x);
  } else {
    const long long bid = blockIdx.x;
    const int b = (int)(bid / nblk);
    K k(a, bid - (long long)b * nblk, b, lds);
    for (int c = 0; c < a.C; ++c) {
      if constexpr (KRY) k.stage_a(c, threadIdx.x, acc);
      else k.stage_a(c, threadIdx.x);
      __syncthreads();
      k.stage_b(c, threadIdx.x);
      if (c + 1 < a.C) __syncthreads();
    }
  }
  if constexpr (KRY) {
    if (a.k.op == KRY_S) kry_block_store(a, acc);   // KRY_P has no sums
  }
}

// KRY: the inverse pass with fused Krylov sums in its epilogue (ZArgs::k).
// SPLIT (per-component epilogues only): one block per (pencil group, component).
template <class K, int C, bool KRY, bool SPLIT>
__global__ __launch_bounds__(K::NT) void k_fast_zi(ZArgs a, long long nblk) {
  if (a.skip && *a.skip) return;
  static_assert(!KRY || (K::NT % 64 == 0 && !K::HOLD), "fused sums: whole wavefronts, per-component epilogue");
  static_assert(!SPLIT || !K::HOLD, "a held epilogue needs every component in one block");
  using T = typename K::Real;
  __shared__ cpx<T> lds[K::LDS];
  typename K::St st;
  KAcc acc{};
  if constexpr (SPLIT) {
    const long long bid = blockIdx.x / C;
    const int c = (int)(blockIdx.x - bid * C);
    const int b = (int)(bid / nblk);
    K k(a, bid - (long long)b * nblk, b, lds);
    k.c_first = c;
    k.stage_a(c, threadIdx.x);
    __syncthreads();
    k.stage_b(c, threadIdx.x, st);
    if constexpr (KRY) {
      k.epilogue_one_kry(c, threadIdx.x, st, acc);
      kry_block_store(a, acc);
    } else {
      k.epilogue_one(c, threadIdx.x, st);
    }
    return;
  }
  const long long bid = blockIdx.x;
  const int b = (int)(bid / nblk);
  K k(a, bid - (long long)b * nblk, b, lds);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    k.stage_a(c, threadIdx.x);
    __syncthreads();
    k.stage_b(c, threadIdx.x, st);
    if constexpr (KRY) k.epilogue_one_kry(c, threadIdx.x, st, acc);
    else if constexpr (!K::HOLD) k.epilogue_one(c, threadIdx.x, st);
    if (c + 1 < C) __syncthreads();
  }
  if constexpr (K::HOLD) k.epilogue(threadIdx.x, st);
  if constexpr (KRY) kry_block_store(a, acc);
}

struct HipFastRun {
  hipStream_t st;
  hipError_t err = hipSuccess;
  template <class K>
  void s(const SArgs& a) {
    const long long ncb = (a.W + K::TC - 1) / K::TC;
    hipLaunchKernelGGL(k_fast_s<K>, dim3((unsigned)(ncb * a.Q)), dim3(K::NT), 0, st, a, ncb);
    err = hipGetLastError();
  }
  template <class K>
  void x(const XArgs& a) {
    const long long ncb = K::blocks(a);
    hipLaunchKernelGGL((k_fast_x<K, K::NCOMP>), dim3((unsigned)(ncb * a.batch)), dim3(K::NT), 0, st,
                       a, ncb);
    err = hipGetLastError();
  }
  template <class K, bool SPLIT = false>
  void zf(const ZArgs& a) {
    const long long nblk = (a.P + K::P - 1) / K::P;
    const long long split = (SPLIT || zf_split<K>()) ? a.C : 1;
    const dim3 grid((unsigned)(nblk * a.batch * split));
    if (a.k.op == KRY_P || a.k.op == KRY_S)
      hipLaunchKernelGGL((k_fast_zf<K, true, SPLIT>), grid, dim3(K::NT), 0, st, a, nblk);
    else
      hipLaunchKernelGGL((k_fast_zf<K, false, SPLIT>), grid, dim3(K::NT), 0, st, a, nblk);
    err = hipGetLastError();
  }
  template <class K, bool SPLIT_ = false>
  void zi(const ZArgs& a) {
    constexpr bool SPLIT = SPLIT_ && !K::HOLD;
    const long long nblk = (a.P + K::P - 1) / K::P;
    const dim3 grid((unsigned)(nblk * a.batch * (SPLIT ? K::NCOMP : 1)));
    if constexpr (!K::HOLD) {
      if (a.k.op == KRY_DOT || a.k.op == KRY_DOT2) {
        hipLaunchKernelGGL((k_fast_zi<K, K::NCOMP, true, SPLIT>), grid, dim3(K::NT), 0, st, a, nblk);
        err = hipGetLastError();
        return;
      }
    }
    hipLaunchKernelGGL((k_fast_zi<K, K::NCOMP, false, SPLIT>), grid, dim3(K::NT), 0, st, a, nblk);
    err = hipGetLastError();
  }
};

// Defined in e3_fast_f32.hip / e3_fast_f64.hip.
bool launch_fast_s_f32(const SArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_x_f32(const XArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_z_f32(const ZArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_s_f64(const SArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_x_f64(const XArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_z_f64(const ZArgs& a, hipStream_t st, hipError_t* err);
// Defined in e3_fast_f32_nt.hip / e3_fast_f64_nt.hip (non-temporal rows).
bool launch_fast_s_f32_nt(const SArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_x_f32_nt(const XArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_s_f64_nt(const SArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_x_f64_nt(const XArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_z_f32_nt(const ZArgs& a, hipStream_t st, hipError_t* err);
bool launch_fast_z_f64_nt(const ZArgs& a, hipStream_t st, hipError_t* err);

}  // namespace e3
