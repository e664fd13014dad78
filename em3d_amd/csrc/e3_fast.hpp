// Fast pass bodies for doubled axes of 2^k (16 ... 512), 3 * 2^k and 5 * 2^k.
//
// The z and strided passes run each transform of length L = R1 * R2 as a
// two-stage Stockham FFT held in registers: stage A (radix R1) reads straight
// from global memory, with the zero half of the doubled axis never loaded;
// one LDS transpose; stage B (radix R2) writes straight to global memory,
// with the discarded half never stored.  Lanes run along the contiguous
// dimension of each access (the column index of a strided pass, the element
// index of a z pass), so every global access of a wavefront is a run of
// 128-byte segments.
//
// The x pass uses up to three register stages (radices RA, RB, RC) with an
// LDS exchange between stages, keeps the forward spectrum of all components
// in registers, applies the 3x3 (or scalar) octant spectrum there and runs the
// inverse transform from the same registers: the forward last stage and the
// inverse first stage share one thread-to-butterfly mapping.
//
// As in e3_blocks.hpp, kernels are phase functions of (thread id, per-thread
// state) separated by block barriers, so the CPU tests can replay them.
#pragma once

#include <type_traits>

#include "e3_blocks.hpp"

namespace e3 {

// cos and sin of 2 pi k / 32, k = 0 .. 15
constexpr double kCos32[16] = {
    1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173, 0.19509032201612826785,
    0.0, -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
    -0.70710678118654752440, -0.83146961230254523708, -0.92387953251128675613, -0.98078528040323044913};
constexpr double kSin32[16] = {
    0.0, 0.19509032201612826785, 0.38268343236508977173, 0.55557023301960222474,
    0.70710678118654752440, 0.83146961230254523708, 0.92387953251128675613, 0.98078528040323044913,
    1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173, 0.19509032201612826785};

// b * exp(-/+ 2 pi i e / 32) (forward / inverse), 0 <= e < 16
template <typename T, bool INV>
E3_HD cpx<T> rot32(cpx<T> b, int e) {
  if (e == 0) return b;
  if (e == 8) return mul_mi<T, INV>(b);
  const T c = (T)kCos32[e];
  const T s = INV ? (T)kSin32[e] : -(T)kSin32[e];
  return mk(b.re * c - b.im * s, b.re * s + b.im * c);
}

template <int R>
constexpr int bit_reverse(int i) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

constexpr bool is_pow2(int r) { return r > 0 && (r & (r - 1)) == 0; }

// cos and sin of 2 pi m / R, evaluated at compile time: the angle is reduced
// exactly to the first quadrant with integer arithmetic, then summed as Taylor
// series (|x| <= pi/2, 30 terms: truncation far below double rounding).
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 30; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 30; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double kHalfPi = 1.57079632679489661923;

template <int R>
struct Roots {
  double c[R], s[R];
  constexpr Roots() : c(), s() {
    for (int m = 0; m < R; ++m) {
      const int q = (4 * m) / R;                   // quadrant of 2 pi m / R
      const double x = kHalfPi * (double)(4 * m - q * R) / (double)R;
      const double cr = taylor_cos(x), sr = taylor_sin(x);
      c[m] = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
      s[m] = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    }
  }
};
template <int R>
inline constexpr Roots<R> kRoots{};

// b * exp(-/+ 2 pi i e / R) (forward / inverse)
template <typename T, int R, bool INV>
E3_HD cpx<T> rotR(cpx<T> b, int e) {
  if (e == 0) return b;
  const T c = (T)kRoots<R>.c[e];
  const T s = INV ? (T)kRoots<R>.s[e] : -(T)kRoots<R>.s[e];
  return mk(b.re * c - b.im * s, b.re * s + b.im * c);
}

// In-place DFT of 3 or 5 points.
template <typename T, int P, bool INV>
E3_HD void dft_odd(cpx<T>* x) {
  if constexpr (P == 3) {
    const T h = (T)0.86602540378443864676;   // sin(2 pi / 3)
    const cpx<T> t1 = x[1] + x[2], t2 = x[1] - x[2];
    const cpx<T> r = x[0] - t1 * (T)0.5;
    const cpx<T> i = mul_mi<T, INV>(t2 * h);
    x[0] = x[0] + t1;
    x[1] = r + i;
    x[2] = r - i;
  } else {
    static_assert(P == 5, "odd radix must be 3 or 5");
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;   // cos(2 pi / 5), cos(4 pi / 5)
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;    // sin(2 pi / 5), sin(4 pi / 5)
    const cpx<T> a1 = x[1] + x[4], b1 = x[1] - x[4], a2 = x[2] + x[3], b2 = x[2] - x[3];
    const cpx<T> r1 = x[0] + a1 * c1 + a2 * c2, r2 = x[0] + a1 * c2 + a2 * c1;
    const cpx<T> i1 = mul_mi<T, INV>(b1 * s1 + b2 * s2), i2 = mul_mi<T, INV>(b1 * s2 - b2 * s1);
    x[0] = x[0] + a1 + a2;
    x[1] = r1 + i1;
    x[4] = r1 - i1;
    x[2] = r2 + i2;
    x[3] = r2 - i2;
  }
}

template <typename T, int R, bool INV>
E3_HD void fft_reg(cpx<T>* v);

// In-register DFT of size R = P Q, P in {3, 5}, Q a power of two: P-point
// DFTs over stride Q, twiddles exp(-/+ 2 pi i n2 k1 / R), Q-point FFTs over
// contiguous runs, then the index map k1 + P k2 <- Q k1 + k2.
template <typename T, int R, bool INV>
E3_HD void fft_mixed(cpx<T>* v) {
  constexpr int P = R % 3 == 0 ? 3 : 5, Q = R / P;
  static_assert(is_pow2(Q), "mixed radix: R must be 3 or 5 times a power of two");
#pragma unroll
  for (int n2 = 0; n2 < Q; ++n2) {
    cpx<T> x[P];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) x[n1] = v[Q * n1 + n2];
    dft_odd<T, P, INV>(x);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) v[Q * k1 + n2] = rotR<T, R, INV>(x[k1], (n2 * k1) % R);
  }
  if constexpr (Q > 1) {
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) fft_reg<T, Q, INV>(v + Q * k1);
  }
  cpx<T> t[R];
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) t[k1 + P * k2] = v[Q * k1 + k2];
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = t[i];
}

// In-register DFT of size R: iterative radix-2 DIT for powers of two <= 32,
// fft_mixed for 3 or 5 times a power of two.
template <typename T, int R, bool INV>
E3_HD void fft_reg(cpx<T>* v) {
  if constexpr (!is_pow2(R)) {
    fft_mixed<T, R, INV>(v);
    return;
  } else {
  static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16 || R == 32, "radix must be a power of two <= 32");
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = bit_reverse<R>(i);
    if (j > i) {
      const cpx<T> t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        const cpx<T> a = v[i + j];
        const cpx<T> b = rot32<T, INV>(v[i + j + half], j * (32 / len));
        v[i + j] = a + b;
        v[i + j + half] = a - b;
      }
    }
  }
  }
}

// Element at a block-uniform base plus a per-lane byte offset below 4 GiB:
// written this way the device compiler addresses it as scalar base + one
// 32-bit vector offset shared by all rows, instead of a 64-bit vector
// address per row.
template <typename P>
E3_HD P* at_offset(P* uniform_base, unsigned lane_bytes) {
  return reinterpret_cast<P*>(reinterpret_cast<typename std::conditional<std::is_const<P>::value, const char, char>::type*>(uniform_base) + lane_bytes);
}

#ifndef E3_X_LD_NT
#define E3_X_LD_NT 1   // with a plan's nt: x-pass row loads non-temporal
#endif
#ifndef E3_X_ST_NT
#define E3_X_ST_NT 1   // with a plan's nt: x-pass row stores non-temporal
#endif
#ifndef E3_X_K_NT
#define E3_X_K_NT 0    // with a plan's nt: octant loads of the x pass non-temporal
#endif
#ifndef E3_SF_LD_NT
#define E3_SF_LD_NT 1  // with a plan's nt: forward y-pass loads / stores non-temporal
#endif
#ifndef E3_SF_ST_NT
#define E3_SF_ST_NT 1
#endif
#ifndef E3_SI_LD_NT
#define E3_SI_LD_NT 1  // with a plan's nt: inverse y-pass loads / stores non-temporal
#endif
#ifndef E3_SI_ST_NT
#define E3_SI_ST_NT 1
#endif
#ifndef E3_ZF_NT_U
#define E3_ZF_NT_U 1    // with a plan's nt: field loads of the forward z pass non-temporal
#endif
#ifndef E3_ZF_NT_CHI
#define E3_ZF_NT_CHI 1  // with a plan's nt: contrast loads of the forward z pass non-temporal
#endif
#ifndef E3_Z_NT
#define E3_Z_NT 0    // 1: loads of the inverse z pass (field, contrast) non-temporal
#endif
#ifndef E3_Z_NT_ST
#define E3_Z_NT_ST 0 // 1: stores of the z passes non-temporal (measured slower)
#endif

// Element access, non-temporal when NT.  The x and y passes take NT when the
// x pass's working set exceeds what the last-level caches hold (the plan's
// nt flag): their streaming rows then do not displace what L2 keeps for
// reuse, such as the octant run a block of the x pass reads twice.  Smaller
// sets stay in the 256 MB Infinity Cache between passes, and plain accesses
// are faster there.
template <bool NT, typename T>
E3_HD cpx<T> ld_nt(const cpx<T>* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return mk(v.x, v.y);
  }
#endif
  return *p;
}
template <bool NT, typename T>
E3_HD void st_nt(cpx<T>* p, cpx<T> x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 v;
    v.x = x.re;
    v.y = x.im;
    __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
    return;
  }
#endif
  *p = x;
}

template <typename T, bool INV>
E3_HD cpx<T> twiddle(const cpx<T>* tw, int idx, cpx<T> x) {
  const cpx<T> w = tw[idx];
  return x * (INV ? conj(w) : w);
}

// ---------------------------------------------------------------------------
// Tile configuration per (precision, doubled length).  R1 = stage-A radix,
// TC = columns per block of a strided pass, P = pencils per block of a z pass.
// ---------------------------------------------------------------------------
template <typename T, int L>
struct FastCfg {
  static constexpr bool ok = false;
};
#define E3_FAST_CFG(TYPE, LEN, RAD1, COLS, PENS)        \
  template <>                                            \
  struct FastCfg<TYPE, LEN> {                            \
    static constexpr bool ok = true;                     \
    static constexpr int R1 = RAD1, TC = COLS, P = PENS; \
  };
E3_FAST_CFG(float, 16, 4, 64, 64)
E3_FAST_CFG(float, 32, 8, 64, 64)
E3_FAST_CFG(float, 64, 8, 32, 32)
E3_FAST_CFG(float, 128, 16, 32, 32)
E3_FAST_CFG(float, 256, 16, 16, 16)
E3_FAST_CFG(float, 512, 32, 16, 16)
E3_FAST_CFG(double, 16, 4, 32, 32)
E3_FAST_CFG(double, 32, 8, 32, 32)
E3_FAST_CFG(double, 64, 8, 16, 16)
E3_FAST_CFG(double, 128, 16, 16, 16)
E3_FAST_CFG(double, 256, 16, 8, 8)
E3_FAST_CFG(double, 512, 32, 8, 8)
// 3 and 5 times a power of two (mixed-radix stage A)
E3_FAST_CFG(float, 48, 12, 64, 64)
E3_FAST_CFG(float, 80, 20, 64, 64)
E3_FAST_CFG(float, 96, 24, 64, 64)
E3_FAST_CFG(float, 160, 40, 32, 32)
E3_FAST_CFG(float, 192, 24, 32, 32)
E3_FAST_CFG(float, 320, 40, 16, 16)
E3_FAST_CFG(float, 384, 48, 16, 16)
E3_FAST_CFG(double, 48, 12, 32, 32)
E3_FAST_CFG(double, 80, 20, 32, 32)
E3_FAST_CFG(double, 96, 24, 32, 32)
E3_FAST_CFG(double, 160, 40, 16, 16)
E3_FAST_CFG(double, 192, 24, 16, 16)
E3_FAST_CFG(double, 320, 40, 8, 8)
E3_FAST_CFG(double, 384, 48, 8, 8)
#undef E3_FAST_CFG

// ---------------------------------------------------------------------------
// Strided pass: (Q planes) x (L rows, live n_in) x (W columns), row stride W.
// PIN: only the first L/2 rows are live; POUT: only the first L/2 rows kept.
// ---------------------------------------------------------------------------
template <typename T, int L, int R1_, int TC_, bool INV, bool PIN, bool POUT, bool NTS = false>
struct FastS {
  using Real = T;
  static constexpr int R1 = R1_, R2 = L / R1_, TC = TC_, NT = TC * R2, TCP = TC + 1;
  static constexpr int NB = R1 / R2;          // stage-B butterflies per thread
  static constexpr int LDS = L * TCP;         // LDS elements
  static constexpr int MA = PIN ? R1 / 2 : R1;
  static constexpr int MB = POUT ? R2 / 2 : R2;
  static_assert(R1 % R2 == 0, "R2 must divide R1");
  struct St {
    cpx<T> v[R1];
  };
  const SArgs& a;
  long long q, w0;
  cpx<T>* lds;
  E3_HD FastS(const SArgs& args, long long colblk, long long plane, cpx<T>* l)
      : a(args), q(s_plane(args, plane)), w0(colblk * TC), lds(l) {}

  E3_HD void stage_a(int tid, St& s) const {
    const int col = tid % TC, j = tid / TC;
    const long long w = w0 + col;
    const bool ok = w < a.W;
    const cpx<T>* in = (const cpx<T>*)a.in + (size_t)q * a.in_qs;
    const unsigned lane = (unsigned)(((long long)j * a.W + w) * sizeof(cpx<T>));
    // input rows outside a forward pass's window [r0, r1) read as zero (one
    // loop: two code paths here leave the mixed-radix row values in scratch)
    constexpr bool NTL = NTS && (INV ? E3_SI_LD_NT : E3_SF_LD_NT);
    const int r0 = (!INV && a.wr1) ? a.wr0 : 0, r1 = (!INV && a.wr1) ? a.wr1 : L;
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int row = j + m * R2;
      const bool live = INV ? (m < MA && ok) : (m < MA && ok && row >= r0 && row < r1);
      s.v[m] = live ? ld_nt<NTL>(at_offset(in + (size_t)m * R2 * a.W, lane)) : czero<T>();
    }
    fft_reg<T, R1, INV>(s.v);
#pragma unroll
    for (int k = 0; k < R1; ++k) lds[(k * R2 + j) * TCP + col] = s.v[k];
  }
  E3_HD void stage_b(int tid, St& s) const {
    const int col = tid % TC, j = tid / TC;
    const long long w = w0 + col;
    const cpx<T>* tw = (const cpx<T>*)a.tw;
    cpx<T>* out = (cpx<T>*)a.out + (size_t)q * a.out_qs;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int jp = j + b * R2;
      cpx<T>* v = s.v + b * R2;
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const cpx<T> x = lds[(jp * R2 + r) * TCP + col];
        v[r] = r ? twiddle<T, INV>(tw, jp * r, x) : x;
      }
      fft_reg<T, R2, INV>(v);
      if (w < a.W) {
        const unsigned lane = (unsigned)(((long long)jp * a.W + w) * sizeof(cpx<T>));
#pragma unroll
        for (int m = 0; m < MB; ++m) st_nt<NTS && (INV ? E3_SI_ST_NT : E3_SF_ST_NT)>(at_offset(out + (size_t)m * R1 * a.W, lane), v[m]);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// x pass: forward (Nx = L/2 live rows), octant spectral multiply, inverse,
// Nx rows kept, in place on W2 viewed per (b, c) as (Nx, Ly, Lz).
//
// A block takes one octant row iy = p (0 <= p <= Ny) and one group of S
// octant columns iz (a TC = 2S wide tile of the mirror-interleaved z axis):
// the frequency rows ky = p and ky = Ly - p (one row for p = 0 and p = Ny),
// one after the other, and in each the columns kz and Lz - kz.  Those four
// frequency columns share one octant entry, so the block's part of the
// spectrum is one contiguous run of (Nx+1) K S entries (plus the Nyquist
// column for tile 0), copied once into LDS: every octant entry is fetched
// once per pass.
//
// The transform along x is a three-stage Stockham FFT, L = RA * RB * RC with
// RC = 8, so each thread keeps only 8 frequencies per component between the
// forward transform, the spectral product and the inverse transform (the
// forward last stage and the inverse first stage share one butterfly
// mapping).  Stages exchange through LDS, in place, with a barrier between
// the reads and the writes of a stage.
// ---------------------------------------------------------------------------
template <int L>
struct XRadix;
template <> struct XRadix<16> { static constexpr int A = 2, B = 1, C = 8; };
template <> struct XRadix<32> { static constexpr int A = 4, B = 1, C = 8; };
template <> struct XRadix<64> { static constexpr int A = 8, B = 1, C = 8; };
template <> struct XRadix<128> { static constexpr int A = 4, B = 4, C = 8; };
#ifndef E3_XR256
#define E3_XR256 8, 4, 8
#endif
template <> struct XRadix<256> {
  static constexpr int r[3] = {E3_XR256};
  static constexpr int A = r[0], B = r[1], C = r[2];
};
template <> struct XRadix<512> { static constexpr int A = 8, B = 8, C = 8; };
// 3 and 5 times a power of two: the odd factor sits in the last (register) stage
template <> struct XRadix<48> { static constexpr int A = 4, B = 1, C = 12; };
template <> struct XRadix<80> { static constexpr int A = 4, B = 1, C = 20; };
template <> struct XRadix<96> { static constexpr int A = 4, B = 2, C = 12; };
template <> struct XRadix<160> { static constexpr int A = 4, B = 2, C = 20; };
template <> struct XRadix<192> { static constexpr int A = 4, B = 4, C = 12; };
template <> struct XRadix<320> { static constexpr int A = 4, B = 4, C = 20; };
template <> struct XRadix<384> { static constexpr int A = 4, B = 4, C = 24; };

// an explicit radix plan (A, B, C) of the x pass
template <int A_, int B_, int C_>
struct XR {
  static constexpr int A = A_, B = B_, C = C_;
};

#ifndef E3_X_PINGPONG
#define E3_X_PINGPONG 1
#endif
// row-tile padding (elements) against LDS bank conflicts: one complex64
// element; none for complex128 (measured: the 256-point FP64 x pass runs
// 285 us unpadded against 297 us padded by one, 326 us by two)
#ifndef E3_X_PAD
#define E3_X_PAD 1
#endif
#ifndef E3_X_PAD_F64
#define E3_X_PAD_F64 0
#endif
#ifndef E3_X_PREFETCH
#define E3_X_PREFETCH 1
#endif
// spectral products scheduled together (their octant reads in flight at
// once): two for complex64; one for complex128, whose x pass runs two
// wavefronts per SIMD (measured at 128^3, same box: FP64 x pass 288 -> 277 us
// with one, 295 us with four; complex64: the default plan unchanged with one
// or four, the 16 x 1 x 16 plan 7 % slower with one)
#ifndef E3_X_MULGROUP
#define E3_X_MULGROUP 2
#endif
#ifndef E3_X_MULGROUP_F64
#define E3_X_MULGROUP_F64 1
#endif
template <typename T, int L, int C, int TC_ = 128 / (int)sizeof(cpx<T>), bool STAGE_ = true, int WAVES_ = 2,
          class RX = XRadix<L>, bool PAIR_ = false, bool NTX = false, bool RSPLIT_ = false>
struct FastX {
  using Real = T;
  static constexpr int NCOMP = C;
  // RSPLIT (3 components, small grids): the rows ky = p and Ly - p of a pair
  // go to two blocks instead of one after the other in one block, which
  // doubles the wavefronts of a launch that would not fill the chip
  static constexpr bool RSPLIT = RSPLIT_ && C == 3;
  // PAIR (scalar operator): the rows ky = p and Ly - p of a block run together
  // as the two "components" c = 0, 1 of one schedule; they share every octant
  // entry (the scalar kernel is even in ky), loaded once for both.
  static constexpr bool PAIR = PAIR_ && C == 1;
  static constexpr int CE = PAIR ? 2 : C;              // components of the phase schedule
  static constexpr int RA = RX::A, RB = RX::B, RC = RX::C;
  static constexpr int NTC = L / RC;                   // threads per column
  static constexpr int S = 64 / (int)sizeof(cpx<T>);   // octant group: 64 bytes of iz
  static constexpr int TC = TC_, TCP = TC + (sizeof(T) == 8 ? E3_X_PAD_F64 : E3_X_PAD);   // columns per block, TC divides 2S
  static constexpr int NT = TC * NTC;                  // threads of a block (one row at a time)
  static constexpr int BA = RC / RA, BB = RB > 1 ? RC / RB : 0;   // butterflies per thread
  static constexpr int K = C == 3 ? 6 : 1;
  static constexpr int NX1 = L / 2 + 1;
  static constexpr int SEG = L * TCP;                  // LDS elements of the row tile
  // the block's octant run: staged once in LDS (STAGE_) or read through the caches
  static constexpr int KGRP = TC / 2;                  // octant columns of the block
  static constexpr int KS_MAIN = NX1 * K * S, KS = NX1 * K * (S + 1);
  // two row tiles (ping-pong: every LDS stage writes the other tile) when a
  // block then still fits twice into a CU's 160 KB, or when it has 1024
  // threads (16 waves) and so runs alone on its CU anyway; else one tile, in place
  static constexpr size_t PP_BYTES = (size_t)(2 * SEG + L + (STAGE_ ? KS : 0)) * sizeof(cpx<T>);
  static constexpr bool PP = E3_X_PINGPONG && (PP_BYTES <= 80 * 1024 || (NT >= 1024 && PP_BYTES <= 160 * 1024));
  static constexpr int NSEG = PP ? 2 : 1;
  static constexpr bool STAGE = STAGE_ && (size_t)(NSEG * SEG + L + KS) * sizeof(cpx<T>) <= 160 * 1024;
  static constexpr int MIN_WAVES_PER_EU = WAVES_;
  static constexpr int LDS = NSEG * SEG + L + (STAGE ? KS : 0);
  static constexpr int kMulGroup = sizeof(T) == 8 ? E3_X_MULGROUP_F64 : E3_X_MULGROUP;
  static_assert(RA * (RB > 1 ? RB : 1) * RC == L, "radix plan");
  static_assert(RC % RA == 0 && (RB == 1 || RC % RB == 0), "stage radices must divide the last one");
  struct St {
    cpx<T> X[CE][RC];           // spectrum at kx = j + m (L / RC)
    cpx<T> mid[BB > 0 ? BB : 1][RB > 1 ? RB : 1];   // middle stage values in flight
    cpx<T> pre[BA][RA / 2];     // inputs of the next stage A, loaded one phase ahead
  };
  // loads of stage A issued one phase before the stage (ping-pong schedule)
  static constexpr bool PREFETCH = E3_X_PREFETCH && PP;
  const XArgs& a;
  int pr, ky0, ky1, pos0, tile, b;
  bool two;            // both rows ky0 = p and ky1 = Ly - p are live
  cpx<T>* lds;
  int hsel = 0;        // RSPLIT: the row of the pair this block runs
  E3_HD FastX(const XArgs& args, long long blk, int bidx, cpx<T>* l) : a(args), b(bidx), lds(l) {
    const int ntile = args.L[2] / TC;
    if (RSPLIT) {
      hsel = (int)(blk & 1);
      blk >>= 1;
    }
    pr = (int)(blk / ntile);
    tile = (int)(blk - (long long)pr * ntile);
    pos0 = tile * TC;
    ky0 = pr;
    ky1 = args.L[1] - pr;
    two = pr > 0 && pr < args.N[1];
  }
  static long long blocks(const XArgs& a) { return (long long)(a.N[1] + 1) * (a.L[2] / TC) * (RSPLIT ? 2 : 1); }

  struct Lane {
    int h, col, j, ky;
    long long w;
  };
  // lane of thread tid while the block works on row h (0: ky = p, 1: ky = Ly - p)
  E3_HD Lane lane(int tid, int h) const {
    Lane q;
    q.h = h;
    q.col = tid % TC;
    q.j = tid / TC;
    q.ky = h ? ky1 : ky0;
    q.w = (long long)q.ky * a.L[2] + pos0 + q.col;
    return q;
  }
  E3_HD cpx<T>* plane(int c) const { return (cpx<T>*)a.data + (size_t)b * a.bs + (size_t)c * a.cs; }
  E3_HD cpx<T>* seg(int buf) const { return lds + buf * SEG; }
  E3_HD const cpx<T>* twl() const { return lds + NSEG * SEG; }
  E3_HD cpx<T>* kst() const { return lds + NSEG * SEG + L; }
  E3_HD cpx<T>& cell(cpx<T>* t, int idx, int col) const { return t[idx * TCP + col]; }

  // ---- loads: twiddles, the block's octant run, all components' inputs ----
  E3_HD void stage_tables(int tid) const {
    for (int t = tid; t < L; t += NT) lds[NSEG * SEG + t] = ((const cpx<T>*)a.tw)[t];
    if (STAGE) {
      // LDS copy [ix][comp][S + 1]: slots 0 .. S-1 the tile's column group,
      // slot S the column iz = Nz (tile 0 only)
      const cpx<T>* row = krow();
      const cpx<T>* kg = row + (long long)((pos0 >> 1) / S) * NX1 * K * S;
      for (int t = tid; t < KS_MAIN; t += NT) kst()[(t / S) * (S + 1) + t % S] = kg[t];
      if (pos0 == 0) {
        const int nz = a.N[2];
        const cpx<T>* kn = row + (long long)(nz / S) * NX1 * K * S + nz % S;
        for (int t = tid; t < NX1 * K; t += NT) kst()[t * (S + 1) + S] = kn[(long long)t * S];
      }
    }
  }
  // sign of component c on this lane's column: the spectrum is applied as
  // diag(sx, sy, sz) K diag(sx, sy, sz) with K the octant entries; sy (row
  // Ly - p) and sz (odd positions >= 3, kz > Nz) do not depend on kx, so they
  // are applied to the rows of components 1 and 2 as they are loaded and
  // stored, and only sx remains in the product.
  E3_HD T comp_sign(int c, const Lane& q) const {
    if (C != 3 || c == 0) return T(1);
    if (c == 1) return q.h ? T(-1) : T(1);
    const int pos = pos0 + q.col;
    return ((pos & 1) && pos > 1) ? T(-1) : T(1);
  }
  // ---- forward stage A (Ns = 1): global rows < L/2 -> regs -> LDS ----
  // stage-A butterfly jj = q.j + bb NTC reads rows jj + m L/RA, live for m < RA/2
  E3_HD void fwd_a(int c, int tid, int h, St& s, int buf) const {
    Lane q = lane(tid, prow(c, h));
    E3_OPAQUE(q.j);   // row addresses are formed per phase, not held across the row
    cpx<T>* t = seg(buf);
    const cpx<T>* d = plane(pcomp(c));
    const bool lv = live(c);
#pragma unroll
    for (int bb = 0; bb < BA; ++bb) {
      const int jj = q.j + bb * NTC;
      const unsigned lane_b = (unsigned)(((long long)jj * a.W + q.w) * sizeof(cpx<T>));
      const T sg = comp_sign(c, q);
      cpx<T> v[RA];
      auto load = [&](auto win) {
#pragma unroll
        for (int m = 0; m < RA; ++m)
          v[m] = (m < RA / 2 && lv && (!decltype(win)::value || x_live(a, jj + m * (L / RA))))
                     ? ld_nt<NTX && E3_X_LD_NT>(at_offset(d + (size_t)m * (L / RA) * a.W, lane_b)) * sg
                     : czero<T>();
      };
      if (a.wx1) load(std::true_type{});   // uniform: a windowed forward transform
      else load(std::false_type{});
      fft_reg<T, RA, false>(v);
#pragma unroll
      for (int m = 0; m < RA; ++m) cell(t, jj * RA + m, q.col) = v[m];
    }
  }
  // stage A in two halves: the row loads of component c of row h into s.pre,
  // and (a phase later) the sign, the radix-RA butterflies and the LDS write
  E3_HD void load_a(int c, int tid, int h, St& s) const {
    Lane q = lane(tid, prow(c, h));
    E3_OPAQUE(q.j);
    const cpx<T>* d = plane(pcomp(c));
    const bool lv = live(c);
#pragma unroll
    for (int bb = 0; bb < BA; ++bb) {
      const int jj = q.j + bb * NTC;
      const unsigned lane_b = (unsigned)(((long long)jj * a.W + q.w) * sizeof(cpx<T>));
      auto load = [&](auto win) {
#pragma unroll
        for (int m = 0; m < RA / 2; ++m)
          s.pre[bb][m] = (lv && (!decltype(win)::value || x_live(a, jj + m * (L / RA))))
                             ? ld_nt<NTX && E3_X_LD_NT>(at_offset(d + (size_t)m * (L / RA) * a.W, lane_b))
                             : czero<T>();
      };
      if (a.wx1) load(std::true_type{});   // uniform: a windowed forward transform
      else load(std::false_type{});
    }
  }
  E3_HD void fwd_a_pre(int c, int tid, int h, St& s, int buf) const {
    const Lane q = lane(tid, prow(c, h));
    cpx<T>* t = seg(buf);
    const T sg = comp_sign(c, q);
#pragma unroll
    for (int bb = 0; bb < BA; ++bb) {
      const int jj = q.j + bb * NTC;
      cpx<T> v[RA];
#pragma unroll
      for (int m = 0; m < RA; ++m) v[m] = m < RA / 2 ? s.pre[bb][m] * sg : czero<T>();
      fft_reg<T, RA, false>(v);
#pragma unroll
      for (int m = 0; m < RA; ++m) cell(t, jj * RA + m, q.col) = v[m];
    }
  }
  // stage A of component c, its loads either issued a phase ago or now
  E3_HD void stage_a(int c, int tid, int h, St& s, int buf) const {
    if constexpr (PREFETCH) fwd_a_pre(c, tid, h, s, buf);
    else fwd_a(c, tid, h, s, buf);
  }
  // ---- middle stage (radix RB): read | barrier | write, in place ----
  E3_HD void mid_read(int tid, int h, St& s, int buf) const {
    if constexpr (RB > 1) {
      const Lane q = lane(tid, h);
      const cpx<T>* t = seg(buf);
#pragma unroll
      for (int bb = 0; bb < BB; ++bb) {
        const int jj = q.j + bb * NTC;
#pragma unroll
        for (int m = 0; m < RB; ++m) s.mid[bb][m] = t[(jj + m * (L / RB)) * TCP + q.col];
      }
    }
  }
  template <bool INV>
  E3_HD void mid_write(int tid, int h, St& s, int ns, int buf) const {
    if constexpr (RB > 1) {
      const Lane q = lane(tid, h);
      cpx<T>* t = seg(buf);
      const cpx<T>* tw = twl();
#pragma unroll
      for (int bb = 0; bb < BB; ++bb) {
        const int jj = q.j + bb * NTC;
        const int k = jj % ns;
        cpx<T>* v = s.mid[bb];
#pragma unroll
        for (int m = 1; m < RB; ++m) v[m] = twiddle<T, INV>(tw, k * m * (L / (ns * RB)), v[m]);
        fft_reg<T, RB, INV>(v);
#pragma unroll
        for (int m = 0; m < RB; ++m) t[((jj / ns) * ns * RB + k + m * ns) * TCP + q.col] = v[m];
      }
    }
  }
  // ---- phase schedule: NPH phases separated by block barriers ----
  // per row: forward per component, spectral product, inverse per component.
  // Rows run in a runtime loop (rows() of them): the phases of a row are
  // compile-time instances, the row index is not, so nothing is carried from
  // one row to the next in registers.
  static constexpr int MID = RB > 1 ? 1 : 0;
  static constexpr int PPC = 2 + 2 * MID;              // phases per component and direction (one tile)
  static constexpr int NPH = PP ? (MID ? 4 * CE + 1 : 2 * CE + 1) : 2 * CE * PPC;
  E3_HD int rows() const { return PAIR ? 1 : (two ? 2 : 1); }
  // the rows h in [row_begin, row_end) this block runs (empty: the second
  // row of a pair 0 or Ny, which does not exist)
  E3_HD int row_begin() const { return RSPLIT ? hsel : 0; }
  E3_HD int row_end() const { return RSPLIT ? (hsel == 0 || two ? hsel + 1 : hsel) : rows(); }
  // field component and row of schedule component c in row iteration h, and
  // whether that row exists (a pair 0 or Ny has one row)
  E3_HD int pcomp(int c) const { return PAIR ? 0 : c; }
  E3_HD int prow(int c, int h) const { return PAIR ? c : h; }
  E3_HD bool live(int c) const { return !PAIR || c == 0 || two; }
  int h_end = 1;       // rows() once the kernel starts the row loop (prefetch of the next row)
  template <int I>
  E3_HD void phase(int tid, int h, St& s) const {
    if constexpr (I == 0) {
      if (h == row_begin()) stage_tables(tid);
    }
    if constexpr (PP) phase_pp<I>(tid, h, s);
    else phase_one<I>(tid, h, s);
  }
  // Two tiles.  A stage that reads the tile written in the previous phase
  // writes the other one, so a middle stage is one phase and the last stage
  // of a component shares its phase with the first stage of the next one.
  // With a middle stage (MID): stage A and inverse stage C write tile 0, the
  // middle stage reads 0 and writes 1, stage C and inverse stage A read 1:
  //   A0 | M0 | C0 A1 | M1 | C1 A2 | M2 | C2 x IC0 | IM0 | IA0 IC1 | IM1 | IA1 IC2 | IM2 | IA2
  // Without: the k-th tile write of the row goes to tile k % 2:
  //   A0 | C0 A1 | C1 A2 | C2 x IC0 | IA0 IC1 | IA1 IC2 | IA2
  // With PREFETCH the loads of stage A of component c + 1 are issued in the
  // phase before the one that transforms them (the block barrier waits for
  // LDS operations only, so they stay in flight across it), and the last
  // phase of a row issues those of component 0 of the next row (the kernel
  // issues them for the block's first row).
  template <int I>
  E3_HD void phase_pp(int tid, int h, St& s) const {
    if constexpr (I == 0) {
      stage_a(0, tid, h, s, 0);
      if constexpr (PREFETCH && !MID && CE > 1) load_a(1, tid, h, s);
    } else if constexpr (MID) {
      constexpr bool fwd = I <= 2 * CE;
      constexpr int J = fwd ? I - 1 : I - 2 * CE - 1;
      constexpr int c = J / 2;
      if constexpr (J % 2 == 0) {
        if constexpr (PREFETCH && fwd && c + 1 < CE) load_a(c + 1, tid, h, s);
        mid_read(tid, h, s, 0);
        mid_write<!fwd>(tid, h, s, fwd ? RA : RC, 1);
      } else if constexpr (fwd) {
        fwd_c(c, tid, h, s, 1);
        if constexpr (c + 1 < CE) {
          stage_a(c + 1, tid, h, s, 0);
        } else {
          multiply(tid, h, s);
          inv_c(0, tid, h, s, 0);
        }
      } else {
        if constexpr (PREFETCH && c + 1 == CE) {
          if (h + 1 < h_end) load_a(0, tid, h + 1, s);
        }
        inv_a(c, tid, h, 1);
        if constexpr (c + 1 < CE) inv_c(c + 1, tid, h, s, 0);
      }
    } else if constexpr (I <= CE) {
      constexpr int c = I - 1;
      fwd_c(c, tid, h, s, c % 2);
      if constexpr (c + 1 < CE) {
        stage_a(c + 1, tid, h, s, (c + 1) % 2);             // consumes s.pre ...
        if constexpr (PREFETCH && c + 2 < CE) load_a(c + 2, tid, h, s);   // ... before it is refilled
      } else {
        multiply(tid, h, s);
        inv_c(0, tid, h, s, CE % 2);
      }
    } else {
      constexpr int c = I - CE - 1;
      if constexpr (PREFETCH && c + 1 == CE) {
        if (h + 1 < h_end) load_a(0, tid, h + 1, s);
      }
      inv_a(c, tid, h, (CE + c) % 2);
      if constexpr (c + 1 < CE) inv_c(c + 1, tid, h, s, (CE + c + 1) % 2);
    }
  }
  // One tile, stages in place (the middle stage reads, then writes after a barrier).
  template <int I>
  E3_HD void phase_one(int tid, int h, St& s) const {
    constexpr bool fwd = I < CE * PPC;
    constexpr int ii = fwd ? I : I - CE * PPC;
    constexpr int c = ii / PPC, r = ii - c * PPC;
    if constexpr (fwd) {
      if constexpr (r == 0) {
        fwd_a(c, tid, h, s, 0);
      } else if constexpr (MID && r == 1) {
        mid_read(tid, h, s, 0);
      } else if constexpr (MID && r == 2) {
        mid_write<false>(tid, h, s, RA, 0);
      } else {
        fwd_c(c, tid, h, s, 0);
      }
    } else {
      if constexpr (r == 0) {
        if constexpr (c == 0) multiply(tid, h, s);
        inv_c(c, tid, h, s, 0);
      } else if constexpr (MID && r == 1) {
        mid_read(tid, h, s, 0);
      } else if constexpr (MID && r == 2) {
        mid_write<true>(tid, h, s, RC, 0);
      } else {
        inv_a(c, tid, h, 0);
      }
    }
  }
  // ---- forward last stage (radix RC, Ns = L/RC): LDS -> X regs ----
  E3_HD void fwd_c(int c, int tid, int h, St& s, int buf) const {
    const Lane q = lane(tid, h);
    const cpx<T>* t = seg(buf);
    const cpx<T>* tw = twl();
    cpx<T>* v = s.X[c];
#pragma unroll
    for (int m = 0; m < RC; ++m) {
      const cpx<T> x = cell((cpx<T>*)t, q.j + m * NTC, q.col);
      v[m] = m ? twiddle<T, false>(tw, q.j * m, x) : x;
    }
    fft_reg<T, RC, false>(v);
  }
  // ---- spectral product at kx = j + m NTC ----
  // The block's octant row (|ky| = p): entry (ix, comp) of column iz at element
  // ((iz / S) NX1 + ix) K S + comp S + iz % S.
  E3_HD const cpx<T>* krow() const {
    const int Nn[3] = {a.N[0], a.N[1], a.N[2]};
    const KhatLayout lay = khat_layout(Nn, C, (int)sizeof(cpx<T>));
    return (const cpx<T>*)a.khat + (long long)pr * lay.row_elems();
  }
  static constexpr int KSTR = STAGE ? S + 1 : S;   // element stride between components
  E3_HD cpx<T> kload(const cpx<T>* base, int off) const {
    if constexpr (STAGE) return base[off];
    return ld_nt<NTX && E3_X_K_NT>(at_offset(base, (unsigned)off * (unsigned)sizeof(cpx<T>)));
  }
  // element of (ix = 0, comp = 0) of this lane's octant column
  E3_HD int kcol(const Lane& q) const {
    const int pos = pos0 + q.col;
    const int iz = pos == 1 ? a.N[2] : (pos >> 1);    // octant column (position 1 holds kz = Nz)
    if constexpr (STAGE) return pos == 1 ? S : iz % S;
    return (iz / S) * NX1 * K * S + iz % S;
  }
  E3_HD int koff(const Lane& q, int t0, int m) const {
    const int kx = q.j + m * NTC;
    const bool fx = (RC % 2 == 0) ? (m >= RC / 2) : (kx > L / 2);
    return t0 + (fx ? L - kx : kx) * K * KSTR;
  }
  E3_HD void multiply(int tid, int h, St& s) const {
    Lane q = lane(tid, h);
    E3_OPAQUE(q.j);   // the octant addresses are formed here, not held across the unit
    E3_OPAQUE(q.col);
    const cpx<T>* base = STAGE ? (const cpx<T>*)kst() : krow();
    product(q, base, kcol(q), a.adj ? T(-1) : T(1), s);
  }
  // kx = j + m NTC folds to ix = L - kx for m >= RC/2 (kx > L/2 there, except
  // kx = L/2 itself, where both folds give ix = L/2 and the odd entries xy, xz
  // are stored as zeros): the sign sx is a compile-time constant per m.
  // sj = -1 applies the adjoint (conjugated entries)
  E3_HD void product(const Lane& q, const cpx<T>* base, int t0, T sj, St& s) const {
#pragma unroll
    for (int m = 0; m < RC; ++m) {
      if (m % kMulGroup == 0) E3_SCHED_FENCE();        // bound the reads in flight
      const bool fx = (RC % 2 == 0) ? (m >= RC / 2) : (q.j + m * NTC > L / 2);
      const int off = koff(q, t0, m);
      if constexpr (C == 1) {
        const cpx<T> k0 = kload(base, off);
        const cpx<T> k = mk(k0.re, k0.im * sj);
        s.X[0][m] = cmul_f<false>(k, s.X[0][m]);
        if constexpr (PAIR) s.X[1][m] = cmul_f<false>(k, s.X[1][m]);
      } else {
        cpx<T> k[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          k[i] = kload(base, off + i * KSTR);
          k[i].im *= sj;
        }
        const T sx = fx ? T(-1) : T(1);
        // X1 = sy f1, X2 = sz f2 (signs applied at load); outputs 1, 2 carry
        // sy, sz, removed at store
        const cpx<T> f0 = s.X[0][m], g1 = s.X[1][m], g2 = s.X[2][m];
        const cpx<T> h0 = f0 * sx, h1 = g1 * sx, h2 = g2 * sx;
        cpx<T> o0 = cmul_f<false>(k[0], f0);
        cmac_f<false>(o0, k[3], h1);
        cmac_f<false>(o0, k[4], h2);
        cpx<T> o1 = cmul_f<false>(k[3], h0);
        cmac_f<false>(o1, k[1], g1);
        cmac_f<false>(o1, k[5], g2);
        cpx<T> o2 = cmul_f<false>(k[4], h0);
        cmac_f<false>(o2, k[5], g1);
        cmac_f<false>(o2, k[2], g2);
        s.X[0][m] = o0;
        s.X[1][m] = o1;
        s.X[2][m] = o2;
      }
    }
  }
  // ---- inverse first stage (radix RC, Ns = 1) from X regs -> LDS ----
  E3_HD void inv_c(int c, int tid, int h, St& s, int buf) const {
    const Lane q = lane(tid, h);
    cpx<T>* t = seg(buf);
    cpx<T>* v = s.X[c];
    fft_reg<T, RC, true>(v);
#pragma unroll
    for (int m = 0; m < RC; ++m) cell(t, q.j * RC + m, q.col) = v[m];
  }
  // ---- inverse last stage (radix RA, Ns = L/RA): LDS -> rows < Nx ----
  E3_HD void inv_a(int c, int tid, int h, int buf) const {
    Lane q = lane(tid, prow(c, h));
    E3_OPAQUE(q.j);
    const cpx<T>* t = seg(buf);
    const cpx<T>* tw = twl();
    cpx<T>* d = plane(pcomp(c));
    if (!live(c)) return;   // block-uniform: the row does not exist
#pragma unroll
    for (int bb = 0; bb < BA; ++bb) {
      const int jj = q.j + bb * NTC;
      cpx<T> v[RA];
#pragma unroll
      for (int m = 0; m < RA; ++m) {
        const cpx<T> x = cell((cpx<T>*)t, jj + m * (L / RA), q.col);
        v[m] = m ? twiddle<T, true>(tw, jj * m, x) : x;
      }
      fft_reg<T, RA, true>(v);
      const unsigned lane_b = (unsigned)(((long long)jj * a.W + q.w) * sizeof(cpx<T>));
      const T sg = comp_sign(c, q);
#pragma unroll
      for (int m = 0; m < RA / 2; ++m) st_nt<NTX && E3_X_ST_NT>(at_offset(d + (size_t)m * (L / RA) * a.W, lane_b), v[m] * sg);
    }
  }
};

// ---------------------------------------------------------------------------
// z pass, forward: pencils of the contiguous axis, components one after the
// other.  PIN: n_in = L/2 with the contrast prologue; else a full transform.
// ---------------------------------------------------------------------------
// NTZ: contrast loads non-temporal; NTU: field loads too (the scalar operator:
// a 3-component epilogue re-reads u, which plain loads leave in the caches)
template <typename T, bool NTZ = false, bool NTU = false>
E3_HD cpx<T> contrast_row(const ZArgs& a, const cpx<T>* in, int c, long long p, int n) {
  constexpr bool NU = NTU && E3_ZF_NT_U, NC = NTZ && E3_ZF_NT_CHI;
  const cpx<T> uc = ld_nt<NU>(&in[(size_t)c * a.in_cs + (size_t)p * a.in_ps + n]);
  if (a.pro != PRO_CHI || a.chi_kind == CHI_NONE) return uc;
  const cpx<T>* chi = (const cpx<T>*)a.chi;
  const long long cell = p * a.n_in + n;
  if (a.chi_kind == CHI_ISO) return ld_nt<NC>(&chi[cell]) * uc;
  if (a.chi_kind == CHI_DIAG) return ld_nt<NC>(&chi[(size_t)c * a.ncell + cell]) * uc;
  cpx<T> acc = czero<T>();
  for (int k = 0; k < 3; ++k)
    acc = acc + ld_nt<NC>(&chi[(size_t)(c * 3 + k) * a.ncell + cell]) *
                    ld_nt<NU>(&in[(size_t)k * a.in_cs + (size_t)p * a.in_ps + n]);
  return acc;
}

// Fused Krylov prologue of one thread's MA live elements of a pencil row
// (n = j + m R2): x = r + beta (p - omega v) (KRY_P; x = r on the first
// iteration) or x = r - alpha v (KRY_S, |x|^2 into acc), stored to w, then
// v[m] = chi[n] x (chi null: x).  __restrict__ lets every load go first.
template <typename T, int R1, int R2, int MA>
E3_HD void kry_row(const cpx<T>* __restrict__ r, const cpx<T>* __restrict__ vv, cpx<T>* __restrict__ w,
                   const cpx<T>* __restrict__ chi, int op, const KryCoef k, int j, cpx<T>* out, KAcc& acc) {
  cpx<T> rr[MA], vx[MA], pw[MA], cx[MA];
  const bool upd = op == KRY_P && !k.first;
#pragma unroll
  for (int m = 0; m < MA; ++m) {
    const int n = j + m * R2;
    rr[m] = r[n];
    vx[m] = (op == KRY_S || upd) ? vv[n] : czero<T>();
    pw[m] = upd ? w[n] : czero<T>();
    cx[m] = chi ? chi[n] : mk((T)1, (T)0);
  }
  const cpx<T> om = mk((T)k.omega[0], (T)k.omega[1]), be = mk((T)k.beta[0], (T)k.beta[1]);
  const cpx<T> al = mk((T)k.alpha[0], (T)k.alpha[1]);
#pragma unroll
  for (int m = 0; m < MA; ++m) {
    cpx<T> x;
    if (op == KRY_P) {
      if (upd) {
        const cpx<T> d = pw[m] - om * vx[m];
        x = rr[m] + be * d;
      } else {
        x = rr[m];
      }
    } else {
      x = rr[m] - al * vx[m];
      acc.v[0] += (double)x.re * x.re + (double)x.im * x.im;
    }
    w[j + m * R2] = x;
    out[m] = chi ? cx[m] * x : x;
  }
}

// Inverse z epilogue y = u - t of one thread's NB x MH outputs of a pencil
// row (z = j + bb R2 + m R1) with the fused Krylov sums: (r_hat, y) (KRY_DOT)
// or (y, y) and (y, u) (KRY_DOT2).  Loads first, then stores.
// chi (when given): y = u - conj(chi) t (the adjoint's epilogue, a contrast
// that does not mix components); op KRY_NONE: no sums.
template <typename T, int NB, int MH, int R1, int R2>
E3_HD void kry_epilogue_row(cpx<T>* __restrict__ out, const cpx<T>* __restrict__ src,
                            const cpx<T>* __restrict__ rh, const cpx<T>* __restrict__ chi, int op, int j,
                            const cpx<T> (*t)[MH], KAcc& acc) {
  cpx<T> x[NB][MH], h[NB][MH];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int m = 0; m < MH; ++m) {
      const int z = j + bb * R2 + m * R1;
      x[bb][m] = src[z];
      h[bb][m] = rh ? rh[z] : (chi ? chi[z] : czero<T>());
    }
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int m = 0; m < MH; ++m) {
      const int z = j + bb * R2 + m * R1;
      const cpx<T> y = x[bb][m] - (chi ? cmulc(h[bb][m], t[bb][m]) : t[bb][m]);
      out[z] = y;
      if (op == KRY_NONE) continue;
      if (op == KRY_DOT) {
        acc.v[0] += (double)h[bb][m].re * y.re + (double)h[bb][m].im * y.im;
        acc.v[1] += (double)h[bb][m].re * y.im - (double)h[bb][m].im * y.re;
      } else {
        acc.v[0] += (double)y.re * y.re + (double)y.im * y.im;
        acc.v[2] += (double)y.re * x[bb][m].re + (double)y.im * x[bb][m].im;
        acc.v[3] += (double)y.re * x[bb][m].im - (double)y.im * x[bb][m].re;
      }
    }
}

template <typename T, int L, int R1_, int P_, bool PIN, bool NTZ = false, bool NTU = false>
struct FastZF {
  using Real = T;
  static constexpr int R1 = R1_, R2 = L / R1_, P = P_, NT = P * R2, R2P = R2 + 1;
  static constexpr int PS = R1 * R2P;          // LDS elements per pencil
  static constexpr int NB = R1 / R2;
  static constexpr int LDS = P * PS;
  static constexpr int MA = PIN ? R1 / 2 : R1;
  struct St {
    int unused;
  };
  const ZArgs& a;
  long long p0;
  int b;
  cpx<T>* lds;
  E3_HD FastZF(const ZArgs& args, long long blk, int bidx, cpx<T>* l) : a(args), p0(blk * P), b(bidx), lds(l) {}

  // stage A in two halves, so a kernel can issue the loads of the next
  // component before transforming the current one
  E3_HD void load_a(int c, int tid, cpx<T>* v) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    const bool ok = p0 + pl < a.P;
    const long long p = z_pencil(a, p0 + pl);
    const cpx<T>* in = (const cpx<T>*)a.in + (size_t)b * a.in_bs;
    auto load = [&](auto win) {
#pragma unroll
      for (int m = 0; m < R1; ++m)
        v[m] = (m < MA && ok && (!decltype(win)::value || z_live(a, j + m * R2)))
                   ? contrast_row<T, NTZ, NTU>(a, in, c, p, j + m * R2)
                   : czero<T>();
    };
    if (a.wz1) load(std::true_type{});   // uniform: a windowed forward transform
    else load(std::false_type{});
  }
  // stage A input with a fused Krylov prologue (KRY_P / KRY_S, iso or diag
  // contrast): the update of each live element, written back, then chi.  The
  // pencil rows go through __restrict__ pointers and every load of the row is
  // issued before its first store (kry_row).
  E3_HD void load_a_kry(int c, int tid, cpx<T>* v, KAcc& acc) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);
    E3_OPAQUE(pl);
    const KryCoef k = a.k.coef[b];
#pragma unroll
    for (int m = 0; m < R1; ++m) v[m] = czero<T>();
    if (p0 + pl >= a.P || k.done) return;
    const long long p = z_pencil(a, p0 + pl);
    const size_t row = (size_t)b * a.in_bs + (size_t)c * a.in_cs + (size_t)p * a.in_ps;
    const cpx<T>* chi = nullptr;
    if (a.pro == PRO_CHI && a.chi_kind != CHI_NONE)
      chi = (const cpx<T>*)a.chi + (a.chi_kind == CHI_ISO ? 0 : (size_t)c * a.ncell) + (size_t)p * a.n_in;
    kry_row<T, R1, R2, MA>((const cpx<T>*)a.k.r + row, (const cpx<T>*)a.k.v + row, (cpx<T>*)a.k.w + row, chi,
                           a.k.op, k, j, v, acc);
  }
  E3_HD void fft_a(int tid, cpx<T>* v) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    fft_reg<T, R1, false>(v);
#pragma unroll
    for (int k = 0; k < R1; ++k) lds[pl * PS + k * R2P + j] = v[k];
  }
  E3_HD void stage_a(int c, int tid) const {
    cpx<T> v[R1];
    load_a(c, tid, v);
    fft_a(tid, v);
  }
  E3_HD void stage_a(int c, int tid, KAcc& acc) const {
    cpx<T> v[R1];
    if (a.k.op == KRY_P || a.k.op == KRY_S) load_a_kry(c, tid, v, acc);
    else load_a(c, tid, v);
    fft_a(tid, v);
  }
  E3_HD void stage_b(int c, int tid) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    const bool ok = p0 + pl < a.P;
    const long long p = z_pencil(a, p0 + pl);
    const cpx<T>* tw = (const cpx<T>*)a.tw;
    cpx<T>* out = (cpx<T>*)a.out + (size_t)b * a.out_bs + (size_t)c * a.out_cs + (size_t)p * a.out_ps;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const int jp = j + bb * R2;
      cpx<T> v[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const cpx<T> x = lds[pl * PS + jp * R2P + r];
        v[r] = r ? twiddle<T, false>(tw, jp * r, x) : x;
      }
      fft_reg<T, R2, false>(v);
      if (ok) {
#pragma unroll
        for (int m = 0; m < R2; ++m)
          st_nt<E3_Z_NT_ST>(&out[a.perm ? mirror_position(jp + m * R1, L / 2, L) : jp + m * R1], v[m]);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// z pass, inverse with the epilogue y = u - t or y = v - chi^H t.  HOLD
// (the adjoint with a full 3x3 contrast, whose chi^H mixes components): the
// first L/2 outputs of all C components stay in registers until one
// epilogue; otherwise each component's epilogue follows its transform.
// ---------------------------------------------------------------------------
template <typename T, int L, int R1_, int P_, int C, bool HOLD_ = true>
struct FastZI {
  using Real = T;
  static constexpr int NCOMP = C;
  static constexpr bool HOLD = HOLD_ && C > 1;
  static constexpr int R1 = R1_, R2 = L / R1_, P = P_, NT = P * R2, R2P = R2 + 1;
  static constexpr int PS = R1 * R2P;
  static constexpr int NB = R1 / R2;
  static constexpr int LDS = P * PS + L;       // pencil tiles + twiddle table
  static constexpr int MH = R2 / 2;
  struct St {
    cpx<T> y[HOLD ? C : 1][NB][MH];
  };
  const ZArgs& a;
  long long p0;
  int b;
  cpx<T>* lds;
  E3_HD FastZI(const ZArgs& args, long long blk, int bidx, cpx<T>* l) : a(args), p0(blk * P), b(bidx), lds(l) {}

  E3_HD void load_a(int c, int tid, cpx<T>* v) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    const long long p = p0 + pl;
    const bool ok = p < a.P;
    const cpx<T>* in = (const cpx<T>*)a.in + (size_t)b * a.in_bs + (size_t)c * a.in_cs + (size_t)p * a.in_ps;
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int n = j + m * R2;
      v[m] = ok ? ld_nt<E3_Z_NT>(&in[a.perm ? mirror_position(n, L / 2, L) : n]) : czero<T>();
    }
  }
  int c_first = 0;     // the block's first component (it stages the twiddle table)
  E3_HD void fft_a(int c, int tid, cpx<T>* v) const {
    if (c == c_first)
      for (int t = tid; t < L; t += NT) lds[P * PS + t] = ((const cpx<T>*)a.tw)[t];
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    fft_reg<T, R1, true>(v);
#pragma unroll
    for (int k = 0; k < R1; ++k) lds[pl * PS + k * R2P + j] = v[k];
  }
  E3_HD void stage_a(int c, int tid) const {
    cpx<T> v[R1];
    load_a(c, tid, v);
    fft_a(c, tid, v);
  }
  E3_HD void stage_b(int c, int tid, St& s) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    const cpx<T>* tw = lds + P * PS;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const int jp = j + bb * R2;
      cpx<T> v[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const cpx<T> x = lds[pl * PS + jp * R2P + r];
        v[r] = r ? twiddle<T, true>(tw, jp * r, x) : x;
      }
      fft_reg<T, R2, true>(v);
#pragma unroll
      for (int m = 0; m < MH; ++m) s.y[HOLD ? c : 0][bb][m] = v[m];
    }
  }
  // epilogue of component c alone (not HOLD): y_c = u_c - t_c or u_c - conj(chi_c) t_c
  E3_HD void epilogue_one(int c, int tid, St& s) const {
    KAcc acc{};
    epilogue_one(c, tid, s, acc);
  }
  // the epilogue of component c with fused Krylov sums: the field and r_hat
  // values of the thread are loaded before its first store
  E3_HD void epilogue_one_kry(int c, int tid, St& s, KAcc& acc) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);
    E3_OPAQUE(pl);
    const long long p = p0 + pl;
    if (p >= a.P) return;
    const size_t row = (size_t)b * a.out_bs + (size_t)c * a.out_cs + (size_t)p * a.out_ps;
    kry_epilogue_row<T, NB, MH, R1, R2>((cpx<T>*)a.out + row, (const cpx<T>*)a.u + row,
                                        a.k.op == KRY_DOT ? (const cpx<T>*)a.k.rh + row : nullptr, nullptr, a.k.op,
                                        j, s.y[0], acc);
  }
  E3_HD void epilogue_one(int c, int tid, St& s, KAcc& acc) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);
    E3_OPAQUE(pl);
    const long long p = p0 + pl;
    if (p >= a.P) return;
    const size_t row = (size_t)b * a.out_bs + (size_t)c * a.out_cs + (size_t)p * a.out_ps;
    const cpx<T>* chi = nullptr;
    if (a.epi == EPI_SUB_CHIH && a.chi_kind != CHI_NONE)
      chi = (const cpx<T>*)a.chi + (a.chi_kind == CHI_ISO ? 0 : (size_t)c * a.ncell) + (size_t)p * a.n_out;
    kry_epilogue_row<T, NB, MH, R1, R2>((cpx<T>*)a.out + row, (const cpx<T>*)a.u + row, nullptr, chi, KRY_NONE, j,
                                        s.y[0], acc);
  }
  E3_HD void epilogue(int tid, St& s) const {
    int j = tid % R2, pl = tid / R2;
    E3_OPAQUE(j);   // lane addresses are formed per stage, not held across components
    E3_OPAQUE(pl);
    const long long p = p0 + pl;
    if (p >= a.P) return;
    cpx<T>* out = (cpx<T>*)a.out + (size_t)b * a.out_bs;
    const cpx<T>* src = (const cpx<T>*)a.u + (size_t)b * a.out_bs;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
      for (int m = 0; m < MH; ++m) {
        if (m % 2 == 0) E3_SCHED_FENCE();   // bound the contrast and field reads in flight
        const int z = j + bb * R2 + m * R1;
        cpx<T> t[C];
#pragma unroll
        for (int c = 0; c < C; ++c) t[c] = s.y[c][bb][m];
        if (a.epi == EPI_SUB_CHIH) {
          cpx<T> wv[3];
          chi_apply<T>((const cpx<T>*)a.chi, a.chi_kind, a.ncell, p * a.n_out + z, C, t, wv, true);
#pragma unroll
          for (int c = 0; c < C; ++c) t[c] = wv[c];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const size_t o = (size_t)c * a.out_cs + (size_t)p * a.out_ps + z;
          st_nt<E3_Z_NT_ST>(&out[o], ld_nt<E3_Z_NT>(&src[o]) - t[c]);
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Dispatch: map a generic pass description onto a fast instantiation.  The
// Runner supplies  template <class K> void s(const SArgs&), x(const XArgs&),
// zf<K>(const ZArgs&), zi<K>(const ZArgs&).  Returns false when no fast
// kernel covers the pass (the caller then uses the generic kernels).
// ---------------------------------------------------------------------------
// Kernel variants the plan may choose per pass kind (the autotuner times
// them): 0 is the default configuration; alternatives exist for the
// single-precision lengths that dominate the measured workloads.
// The x and z passes have one more variant at the last index for every
// length and precision: the small-grid tile (small_variant, below).
inline int fast_variant_count(int kind, int L, int prec) {
  if (kind == 0) return ((prec == 0 && (L == 256 || L == 192)) || (prec == 1 && L == 256)) ? 3 : 1;   // strided passes
  int n = 1;
  if (prec == 0 && kind == 1) n = (L == 256) ? 4 : (L == 128 || L == 192) ? 3 : 1;   // x pass
  if (prec == 0 && kind == 2) n = (L == 256 || L == 192) ? 2 : 1;                    // z passes
  return n + 1;
}
inline int small_variant(int kind, int L, int prec) { return kind == 0 ? -2 : fast_variant_count(kind, L, prec) - 1; }

// strided-pass configuration of a variant: (R1, TC)
template <typename T, int L, int V>
struct SVar {
  static constexpr int R1 = FastCfg<T, L>::R1, TC = FastCfg<T, L>::TC;
};
template <> struct SVar<float, 256, 1> { static constexpr int R1 = 32, TC = 16; };
template <> struct SVar<float, 256, 2> { static constexpr int R1 = 16, TC = 32; };
template <> struct SVar<float, 192, 1> { static constexpr int R1 = 48, TC = 32; };
template <> struct SVar<float, 192, 2> { static constexpr int R1 = 24, TC = 16; };
// complex128 at 256 points: 16-column tiles (256-byte rows) by default;
// measured at 128^3, same box: y forward 113 -> 108 us, y inverse 121 -> 110 us
// against the 8-column tile of FastCfg (now variant 1)
template <> struct SVar<double, 256, 0> { static constexpr int R1 = 16, TC = 16; };
template <> struct SVar<double, 256, 1> { static constexpr int R1 = 16, TC = 8; };
template <> struct SVar<double, 256, 2> { static constexpr int R1 = 32, TC = 8; };

template <typename T, int L, int V, bool NT, class Run>
bool fast_s_var(Run& run, const SArgs& a) {
  using Cfg = SVar<T, L, V>;
  if (a.n_in * 2 == L && a.n_out == L && !a.inv) {
    run.template s<FastS<T, L, Cfg::R1, Cfg::TC, false, true, false, NT>>(a);
    return true;
  }
  if (a.n_in == L && a.n_out * 2 == L && a.inv) {
    run.template s<FastS<T, L, Cfg::R1, Cfg::TC, true, false, true, NT>>(a);
    return true;
  }
  if constexpr (sizeof(T) == 8) {
    if (a.n_in == L && a.n_out == L && !a.inv) {
      run.template s<FastS<T, L, Cfg::R1, Cfg::TC, false, false, false, NT>>(a);
      return true;
    }
  }
  return false;
}

template <typename T, int L, bool NT, class Run>
bool fast_s_len(Run& run, const SArgs& a) {
  if constexpr ((std::is_same<T, float>::value && (L == 256 || L == 192)) ||
                (std::is_same<T, double>::value && L == 256)) {
    if (a.variant == 1) return fast_s_var<T, L, 1, NT>(run, a);
    if (a.variant == 2) return fast_s_var<T, L, 2, NT>(run, a);
  }
  return fast_s_var<T, L, 0, NT>(run, a);
}

// x-pass tile of the 3-component operator: columns per block, octant run
// staged in LDS, minimum waves per SIMD (register budget)
#ifndef E3_X3_TC
#define E3_X3_TC (128 / (int)sizeof(cpx<T>))
#endif
#ifndef E3_X3_STAGE
#define E3_X3_STAGE false
#endif
#ifndef E3_X3_WAVES
#define E3_X3_WAVES 2
#endif

// RSPLIT: the small-grid variant, the rows of a ky pair in two blocks
template <typename T, int L, class RX, int TC, bool NT, class Run, bool RSPLIT = false>
bool fast_x3(Run& run, const XArgs& a) {
  // mixed-radix lengths hold more live values: one wave per SIMD keeps them
  // in registers (VGPR + AGPR) instead of scratch
  using X3 = FastX<T, L, 3, TC, E3_X3_STAGE, is_pow2(L) ? E3_X3_WAVES : 1, RX, false, NT, RSPLIT>;
  if (a.L[2] % X3::TC != 0 || X3::NT > 1024) return false;
  run.template x<X3>(a);
  return true;
}

// x-pass tile of the scalar operator: columns per block, octant run staged in LDS
#ifndef E3_X1_TC
#define E3_X1_TC (128 / (int)sizeof(cpx<T>))
#endif
#ifndef E3_X1_STAGE
#define E3_X1_STAGE true
#endif
#ifndef E3_X1_PAIR
#define E3_X1_PAIR 1   // both rows of a ky pair in one schedule
#endif

template <typename T, int L, bool NT, class Run>
bool fast_x_len(Run& run, const XArgs& a) {
  using X1 = FastX<T, L, 1, E3_X1_TC, E3_X1_STAGE, 2, XRadix<L>, E3_X1_PAIR, NT>;
  if (a.N[0] * 2 != L || !a.perm) return false;
  if (a.C == 3) {
    constexpr int TCd = E3_X3_TC;
    if (a.variant == small_variant(1, L, sizeof(T) == 8 ? 1 : 0))
      return fast_x3<T, L, XRadix<L>, TCd, NT, Run, true>(run, a);
    if constexpr (std::is_same<T, float>::value && L == 256) {
      if (a.variant == 1) return fast_x3<T, L, XR<4, 4, 16>, TCd, NT>(run, a);
      if (a.variant == 2) return fast_x3<T, L, XR<8, 4, 8>, 8, NT>(run, a);
      if (a.variant == 3) return fast_x3<T, L, XR<16, 1, 16>, TCd, NT>(run, a);
    }
    if constexpr (std::is_same<T, float>::value && L == 128) {
      if (a.variant == 1) return fast_x3<T, L, XR<8, 2, 8>, TCd, NT>(run, a);
      if (a.variant == 2) return fast_x3<T, L, XR<8, 1, 16>, TCd, NT>(run, a);
    }
    if constexpr (std::is_same<T, float>::value && L == 192) {
      if (a.variant == 1) return fast_x3<T, L, XR<4, 2, 24>, TCd, NT>(run, a);
      if (a.variant == 2) return fast_x3<T, L, XR<8, 1, 24>, TCd, NT>(run, a);
    }
    return fast_x3<T, L, XRadix<L>, TCd, NT>(run, a);
  }
  if (a.L[2] % X1::TC != 0 || X1::NT > 1024) return false;
  run.template x<X1>(a);
  return true;
}

// z-pass configuration of a variant: (R1, P)
template <typename T, int L, int V>
struct ZVar {
  static constexpr int R1 = FastCfg<T, L>::R1, P = FastCfg<T, L>::P;
};
template <> struct ZVar<float, 256, 1> { static constexpr int R1 = 32, P = 32; };
template <> struct ZVar<float, 192, 1> { static constexpr int R1 = 48, P = 32; };

// Small grids (a support box, a packed member: a few hundred thousand
// cells): a quarter of the pencils per block, at least one wavefront, and one
// component per block, so the launch still spreads over every CU.
template <class Cfg, int L>
struct ZSmall {
  static constexpr int R1 = Cfg::R1, R2 = L / Cfg::R1;
  static constexpr int P = (Cfg::P / 4) * R2 >= 64 ? Cfg::P / 4 : 64 / R2;
};

template <typename T, int L, class Cfg, bool SPLIT, bool NT, class Run>
bool fast_z_cfg(Run& run, const ZArgs& a) {
  if (a.C != 1 && a.C != 3) return false;
  if (a.n_in * 2 == L && a.n_out == L && a.inv == 0 && a.epi == EPI_STORE) {
    if (NT && a.C == 1)
      run.template zf<FastZF<T, L, Cfg::R1, Cfg::P, true, NT, NT>, SPLIT>(a);
    else
      run.template zf<FastZF<T, L, Cfg::R1, Cfg::P, true, NT, false>, SPLIT>(a);
    return true;
  }
  if constexpr (NT) return false;   // the non-temporal set holds the forward pass only
  if (a.n_in == L && a.n_out * 2 == L && a.inv == 1 && a.pro == PRO_COPY &&
      (a.epi == EPI_SUB || a.epi == EPI_SUB_CHIH)) {
    if (a.C == 3 && a.epi == EPI_SUB_CHIH && a.chi_kind == CHI_FULL)
      run.template zi<FastZI<T, L, Cfg::R1, Cfg::P, 3, true>, SPLIT>(a);
    else if (a.C == 3)
      run.template zi<FastZI<T, L, Cfg::R1, Cfg::P, 3, false>, SPLIT>(a);
    else
      run.template zi<FastZI<T, L, Cfg::R1, Cfg::P, 1, false>, SPLIT>(a);
    return true;
  }
  if constexpr (sizeof(T) == 8) {
    if (a.n_in == L && a.n_out == L && a.inv == 0 && a.pro == PRO_COPY && a.epi == EPI_STORE) {
      run.template zf<FastZF<T, L, Cfg::R1, Cfg::P, false>, SPLIT>(a);
      return true;
    }
  }
  return false;
}

template <typename T, int L, int V, bool NT, class Run>
bool fast_z_var(Run& run, const ZArgs& a) {
  return fast_z_cfg<T, L, ZVar<T, L, V>, false, NT>(run, a);
}

template <typename T, int L, bool NT, class Run>
bool fast_z_len(Run& run, const ZArgs& a) {
  if constexpr (std::is_same<T, float>::value && (L == 256 || L == 192)) {
    if (a.variant == 1) return fast_z_var<T, L, 1, NT>(run, a);
  }
  if (a.variant == small_variant(2, L, sizeof(T) == 8 ? 1 : 0))
    return fast_z_cfg<T, L, ZSmall<ZVar<T, L, 0>, L>, true, NT>(run, a);
  // From 256 points on, the plain forward pass of the default tiles runs the
  // small tiles too: its default block holds so large an LDS tile that only
  // two blocks fit a CU, and smaller blocks fit more wavefronts (measured,
  // same box: P1 of the acoustic 256^3 matvec 200 -> 168 us, of the FP64 EM
  // 128^3 one 92 -> 88 us).  The inverse pass keeps its default tiles (the
  // small ones are slower there: FP64 75 -> 79 us, acoustic 111 -> 127 us),
  // and so does the fused Krylov forward pass (the FP64 solve lost 0.5 %).
  if (a.variant == 0 && a.inv == 0 && a.k.op == KRY_NONE && L >= 256)
    return fast_z_cfg<T, L, ZSmall<ZVar<T, L, 0>, L>, true, NT>(run, a);
  return fast_z_var<T, L, 0, NT>(run, a);
}

// Launch sizes of the default tiles, for the plan's small-grid rule: pencils
// per z block and the register radix of the 3-component x pass per length.
inline int fast_z_pencils(int L, int prec) {
#define E3_ZP(LEN) \
  case LEN: return prec ? FastCfg<double, LEN>::P : FastCfg<float, LEN>::P;
  switch (L) {
    E3_ZP(16) E3_ZP(32) E3_ZP(48) E3_ZP(64) E3_ZP(80) E3_ZP(96) E3_ZP(128) E3_ZP(160) E3_ZP(192) E3_ZP(256)
    E3_ZP(320) E3_ZP(384) E3_ZP(512)
    default: return 0;
  }
#undef E3_ZP
}
inline int fast_x_rc(int L) {
#define E3_XC(LEN) \
  case LEN: return XRadix<LEN>::C;
  switch (L) {
    E3_XC(16) E3_XC(32) E3_XC(48) E3_XC(64) E3_XC(80) E3_XC(96) E3_XC(128) E3_XC(160) E3_XC(192) E3_XC(256)
    E3_XC(320) E3_XC(384) E3_XC(512)
    default: return 0;
  }
#undef E3_XC
}

// Default variants of a new plan (variant[1] x pass, variant[2] z passes):
// the small-grid tiles when the default tiles would launch fewer than 1024
// z blocks (one per component counted) or 2048 x-pass wavefronts, i.e. when a
// launch would leave CUs idle (support boxes, packed members, grids below
// about 96^3); 0 otherwise.
inline void small_grid_variants(const int N[3], const int L[3], int C, int batch, int prec, int variant[3]) {
  const int zp = fast_z_pencils(L[2], prec);
  if (zp > 0 && ((long long)N[0] * N[1] + zp - 1) / zp * batch * C < 1024) variant[2] = small_variant(2, L[2], prec);
  const int rc = fast_x_rc(L[0]), tc = prec ? 8 : 16;   // E3_X3_TC: 128 bytes of columns
  if (C == 3 && rc > 0 && L[2] % tc == 0) {
    const long long waves = (long long)(N[1] + 1) * (L[2] / tc) * ((tc * (L[0] / rc) + 63) / 64) * batch;
    if (waves < 2048) variant[1] = small_variant(1, L[0], prec);
  }
}

#define E3_FAST_SWITCH(FN, T, RUN, ARGS)        \
  switch (ARGS.plan.L) {                        \
    case 16: return FN<T, 16, NT>(RUN, ARGS);       \
    case 32: return FN<T, 32, NT>(RUN, ARGS);       \
    case 64: return FN<T, 64, NT>(RUN, ARGS);       \
    case 128: return FN<T, 128, NT>(RUN, ARGS);     \
    case 256: return FN<T, 256, NT>(RUN, ARGS);     \
    case 512: return FN<T, 512, NT>(RUN, ARGS);     \
    default: return false;                      \
  }

// strided and z passes also cover 3 and 5 times a power of two
#define E3_FAST_SWITCH_MIXED(FN, T, RUN, ARGS)  \
  switch (ARGS.plan.L) {                        \
    case 16: return FN<T, 16, NT>(RUN, ARGS);       \
    case 32: return FN<T, 32, NT>(RUN, ARGS);       \
    case 48: return FN<T, 48, NT>(RUN, ARGS);       \
    case 64: return FN<T, 64, NT>(RUN, ARGS);       \
    case 80: return FN<T, 80, NT>(RUN, ARGS);       \
    case 96: return FN<T, 96, NT>(RUN, ARGS);       \
    case 128: return FN<T, 128, NT>(RUN, ARGS);     \
    case 160: return FN<T, 160, NT>(RUN, ARGS);     \
    case 192: return FN<T, 192, NT>(RUN, ARGS);     \
    case 256: return FN<T, 256, NT>(RUN, ARGS);     \
    case 320: return FN<T, 320, NT>(RUN, ARGS);     \
    case 384: return FN<T, 384, NT>(RUN, ARGS);     \
    case 512: return FN<T, 512, NT>(RUN, ARGS);     \
    default: return false;                      \
  }

// NT: the non-temporal instantiations of the y, x and forward z passes (the args' nt)
template <typename T, bool NT = false, class Run>
bool fast_s(Run& run, const SArgs& a) {
  if (a.variant < 0) return false;
  E3_FAST_SWITCH_MIXED(fast_s_len, T, run, a)
}
template <typename T, bool NT = false, class Run>
bool fast_x(Run& run, const XArgs& a) {
  if (a.variant < 0) return false;
  E3_FAST_SWITCH_MIXED(fast_x_len, T, run, a)
}
template <typename T, bool NT = false, class Run>
bool fast_z(Run& run, const ZArgs& a) {
  if (a.variant < 0) return false;
  E3_FAST_SWITCH_MIXED(fast_z_len, T, run, a)
}
#undef E3_FAST_SWITCH
#undef E3_FAST_SWITCH_MIXED

}  // namespace e3
