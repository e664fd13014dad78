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
  const int nrow = k.rows();
  k.h_end = nrow;
  if constexpr (K::PREFETCH) k.load_a(0, threadIdx.x, 0, st);
#pragma unroll 1
  for (int h = 0; h < nrow; ++h) {
    if (h) __syncthreads();
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

template <class K>
__global__ __launch_bounds__(K::NT) void k_fast_zf(ZArgs a, long long nblk) {
  if (a.skip && *a.skip) return;
  using T = typename K::Real;
  __shared__ cpx<T> lds[K::LDS];
  if constexpr (zf_split<K>()) {
    const long long bid = blockIdx.x / a.C;
    const int c = (int)(blockIdx.x - bid * a.C);
    const int b = (int)(bid / nblk);
    K k(a, bid - (long long)b * nblk, b, lds);
    k.stage_a(c, threadIdx.x);
    __syncthreads();
    k.stage_b(c, threadIdx.x);
  } else {
    const long long bid = blockIdx.x;
    const int b = (int)(bid / nblk);
    K k(a, bid - (long long)b * nblk, b, lds);
    for (int c = 0; c < a.C; ++c) {
      k.stage_a(c, threadIdx.x);
      __syncthreads();
      k.stage_b(c, threadIdx.x);
      if (c + 1 < a.C) __syncthreads();
    }
  }
}

template <class K, int C>
__global__ __launch_bounds__(K::NT) void k_fast_zi(ZArgs a, long long nblk) {
  if (a.skip && *a.skip) return;
  using T = typename K::Real;
  __shared__ cpx<T> lds[K::LDS];
  const long long bid = blockIdx.x;
  const int b = (int)(bid / nblk);
  K k(a, bid - (long long)b * nblk, b, lds);
  typename K::St st;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    k.stage_a(c, threadIdx.x);
    __syncthreads();
    k.stage_b(c, threadIdx.x, st);
    if constexpr (!K::HOLD) k.epilogue_one(c, threadIdx.x, st);
    if (c + 1 < C) __syncthreads();
  }
  if constexpr (K::HOLD) k.epilogue(threadIdx.x, st);
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
  template <class K>
  void zf(const ZArgs& a) {
    const long long nblk = (a.P + K::P - 1) / K::P;
    const long long split = zf_split<K>() ? a.C : 1;
    hipLaunchKernelGGL(k_fast_zf<K>, dim3((unsigned)(nblk * a.batch * split)), dim3(K::NT), 0, st, a, nblk);
    err = hipGetLastError();
  }
  template <class K>
  void zi(const ZArgs& a) {
    const long long nblk = (a.P + K::P - 1) / K::P;
    hipLaunchKernelGGL((k_fast_zi<K, K::NCOMP>), dim3((unsigned)(nblk * a.batch)), dim3(K::NT), 0, st,
                       a, nblk);
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
