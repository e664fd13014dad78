// Common device/host helpers for the em3d HIP library: complex arithmetic,
// radix plans for the Stockham FFT and the small radix-R DFTs.
//
// Every function here is usable from HIP device code and, through the
// E3_HD macro, from plain host C++ (the block-phase emulator used by the CPU
// tests compiles these headers with g++).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define E3_HD __host__ __device__ __forceinline__
#else
#define E3_HD inline
#endif

// Scheduling fence for the device compiler: instructions are not moved across
// it (bounds how many loads the scheduler hoists, and so the register peak).
// The host build of these headers (CPU tests) compiles it to nothing.
#if defined(__HIP_DEVICE_COMPILE__)
#define E3_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define E3_SCHED_FENCE() ((void)0)
#endif

// E3_OPAQUE(x): an empty volatile asm that takes and returns the integer x
// in a vector register.  Values computed from x after this point cannot be
// hoisted above it, in particular not out of a loop, so per-lane addresses
// are formed in the phase that uses them instead of being held in registers
// across the whole kernel.  No instruction is emitted; host build: nothing.
#if defined(__HIP_DEVICE_COMPILE__)
#define E3_OPAQUE(x) asm volatile("" : "+v"(x))
#else
#define E3_OPAQUE(x) ((void)0)
#endif

namespace e3 {

template <typename T>
struct cpx {
  T re, im;
};

template <typename T> E3_HD cpx<T> mk(T r, T i) { cpx<T> c; c.re = r; c.im = i; return c; }
template <typename T> E3_HD cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T> E3_HD cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T> E3_HD cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> E3_HD cpx<T> operator*(T s, cpx<T> a) { return mk(s * a.re, s * a.im); }
template <typename T> E3_HD cpx<T> operator*(cpx<T> a, T s) { return mk(a.re * s, a.im * s); }
template <typename T> E3_HD cpx<T> conj(cpx<T> a) { return mk(a.re, -a.im); }
template <typename T> E3_HD cpx<T> czero() { return mk(T(0), T(0)); }

// fused multiply-add of the real type
E3_HD float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
E3_HD double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
// k x and acc + k x with k conjugated when CJ, as 2 + 2 fused operations
template <bool CJ, typename T>
E3_HD cpx<T> cmul_f(cpx<T> k, cpx<T> x) {
  const T ki = CJ ? -k.im : k.im;
  return mk(fmaT(k.re, x.re, -(ki * x.im)), fmaT(k.re, x.im, ki * x.re));
}
template <bool CJ, typename T>
E3_HD void cmac_f(cpx<T>& acc, cpx<T> k, cpx<T> x) {
  const T ki = CJ ? -k.im : k.im;
  acc.re = fmaT(-ki, x.im, fmaT(k.re, x.re, acc.re));
  acc.im = fmaT(ki, x.re, fmaT(k.re, x.im, acc.im));
}
// conj(a) * b
template <typename T> E3_HD cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> E3_HD cpx<T> mul_mi(cpx<T> a) {
  return INV ? mk(-a.im, a.re) : mk(a.im, -a.re);
}

// Radix plan of one FFT length.  Lengths factor into {2,3,4,5,7,8,16}.
constexpr int kMaxStages = 12;
struct AxisPlan {
  int L;                      // transform length (the doubled axis, 2N)
  int nst;                    // number of Stockham stages
  int rad[kMaxStages];        // radix of each stage, first stage first
};

// Greedy factorisation: 16s and 8s first, then 4, 2, 3, 5, 7.
inline bool make_axis_plan(int L, AxisPlan* p, int max_radix = 8) {
  p->L = L;
  p->nst = 0;
  int r = L;
  const int order[] = {16, 8, 4, 2, 3, 5, 7};
  for (int f : order) {
    if (f > max_radix) continue;
    while (r % f == 0 && r > 1) {
      if (p->nst >= kMaxStages) return false;
      p->rad[p->nst++] = f;
      r /= f;
    }
  }
  return r == 1 && L >= 1;
}

// Small DFT of radix R on v[0..R) in place, twiddles W_R^{jk} from the
// length-L forward table tw (tw[t] = exp(-2 pi i t / L)), R | L.
template <typename T, int R, bool INV>
E3_HD void dft_small(cpx<T>* v, const cpx<T>* tw, int L) {
  if constexpr (R == 2) {
    cpx<T> a = v[0] + v[1], b = v[0] - v[1];
    v[0] = a; v[1] = b;
    return;
  }
  if constexpr (R == 4) {
    cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
    cpx<T> t3 = mul_mi<T, INV>(v[1] - v[3]);
    v[0] = t0 + t2; v[2] = t0 - t2; v[1] = t1 + t3; v[3] = t1 - t3;
    return;
  }
  if constexpr (R == 8) {
    // radix-8 as two radix-4 halves combined with W8^k
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<T, 4, INV>(e, tw, L);
    dft_small<T, 4, INV>(o, tw, L);
    const T h = T(0.70710678118654752440);
    // W8^1, W8^2, W8^3 applied to o[1], o[2], o[3]
    cpx<T> o1 = INV ? mk(h * (o[1].re - o[1].im), h * (o[1].re + o[1].im))
                    : mk(h * (o[1].re + o[1].im), h * (o[1].im - o[1].re));
    cpx<T> o2 = mul_mi<T, INV>(o[2]);
    cpx<T> o3 = INV ? mk(-h * (o[3].re + o[3].im), h * (o[3].re - o[3].im))
                    : mk(h * (o[3].im - o[3].re), -h * (o[3].re + o[3].im));
    v[0] = e[0] + o[0]; v[4] = e[0] - o[0];
    v[1] = e[1] + o1;   v[5] = e[1] - o1;
    v[2] = e[2] + o2;   v[6] = e[2] - o2;
    v[3] = e[3] + o3;   v[7] = e[3] - o3;
    return;
  }
  cpx<T> out[R];
  const int step = L / R;
  for (int k = 0; k < R; ++k) {
    cpx<T> acc = v[0];
    for (int j = 1; j < R; ++j) {
      cpx<T> w = tw[((j * k) % R) * step];
      if (INV) w = conj(w);
      acc = acc + v[j] * w;
    }
    out[k] = acc;
  }
  for (int k = 0; k < R; ++k) v[k] = out[k];
}

// Fold a doubled-grid frequency index into the octant [0, N] with the sign
// an odd spectrum takes there: F(L - k) = -F(k) for an odd sequence.
E3_HD int fold_index(int k, int N, int L, int* flipped) {
  if (k > N) { *flipped = 1; return L - k; }
  *flipped = 0;
  return k;
}

}  // namespace e3
