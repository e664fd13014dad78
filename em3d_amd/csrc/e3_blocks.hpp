// Block-phase bodies of the generic (any 7-smooth length) FFT pass kernels.
//
// Each pass kernel is a sequence of phases separated by block barriers:
//   load (global -> LDS, prologue) | Stockham stages (LDS ping-pong) |
//   [spectral multiply] | [inverse stages] | store (LDS -> global, epilogue)
// A phase is a function of (block id, thread id); the HIP kernels in
// e3_kernels.hip run the phases with __syncthreads() between them.  Keeping
// the phases as plain functions lets the CPU tests run the very same index
// arithmetic thread by thread on the host.
//
// Layouts (C order, z fastest):
//   field   u, y : (B, C, Nx, Ny, Nz)
//   W1           : (B, C, Nx, Ny, Lz)      after the z pass
//   W2           : (B, C, Nx, Ly, Lz)      after the y pass (x still unpadded)
//   Khat octant  : (K, Nx+1, Ny+1, Nz+1)   K = 6 (xx,yy,zz,xy,xz,yz) or 1
//   chi          : iso (Nx,Ny,Nz) | diag (3,Nx,Ny,Nz) | full (3,3,Nx,Ny,Nz)
#pragma once

#include "e3_common.hpp"

namespace e3 {

enum ChiKind { CHI_NONE = 0, CHI_ISO = 1, CHI_DIAG = 2, CHI_FULL = 3 };
enum Prologue { PRO_COPY = 0, PRO_CHI = 1 };
enum Epilogue { EPI_STORE = 0, EPI_SUB = 1, EPI_SUB_CHIH = 2 };

// ---------------------------------------------------------------------------
// Krylov steps of the device BiCGStab fused into the z passes of its two
// matvecs (solvers/bicgstab.py:96-116, 145-175 of the reference):
//   KRY_P    forward z prologue of v = A p:  p = r + beta (p - omega v)
//            (p = r on the first iteration), written back, then chi p;
//   KRY_S    forward z prologue of t = A s:  s = r - alpha v, written, then
//            chi s, with the partial sums of |s|^2;
//   KRY_DOT  inverse z epilogue of v = A p:  partial sums of (r_hat, v);
//   KRY_DOT2 inverse z epilogue of t = A s:  partial sums of (t, t), (t, s).
// The coefficients come per right-hand side from the solver's decide kernels
// (KryCoef); a right-hand side that is done keeps its fields.  Partial sums
// go to part[(b * part_stride + block) * 4 + i], block counted within the
// right-hand side, and the number of blocks per right-hand side to *count.
// The prologues are elementwise per component, so they need a contrast that
// does not mix components (none, iso, diag).
// ---------------------------------------------------------------------------
enum KryOp { KRY_NONE = 0, KRY_P = 1, KRY_S = 2, KRY_DOT = 3, KRY_DOT2 = 4 };

struct KryCoef {
  double beta[2], omega[2], alpha[2];
  int first, done;
};

struct ZKry {
  int op;                  // KryOp
  const void* r;           // KRY_P, KRY_S
  const void* v;           // KRY_P, KRY_S
  void* w;                 // KRY_P, KRY_S: the updated p or s (the pass input)
  const void* rh;          // KRY_DOT
  const KryCoef* coef;     // per right-hand side
  double* part;
  long long part_stride;   // partial records per right-hand side
  int* count;              // blocks per right-hand side (written by block 0)
};

// per-thread partial sums of a fused reduction
struct KAcc {
  double v[4];
};

// ---------------------------------------------------------------------------
// Stockham stage over ncol pencils held in LDS (pencil p at base + p*ps).
// Butterfly j of radix R with Ns = product of the radices already applied:
//   v[m] = in[j + m L/R] * W_{Ns R}^{(j mod Ns) m};  v = DFT_R(v);
//   out[(j / Ns) Ns R + (j mod Ns) + m Ns] = v[m]
// ---------------------------------------------------------------------------
template <typename T, int R, bool INV>
E3_HD void stockham_stage(const cpx<T>* in, cpx<T>* out, int ncol, int ps, int L, int Ns,
                          const cpx<T>* tw, int tid, int nt) {
  const int LR = L / R;
  const int nb = ncol * LR;
  const int twstep = L / (Ns * R);
  for (int b = tid; b < nb; b += nt) {
    const int col = b / LR;
    const int j = b - col * LR;
    const cpx<T>* src = in + (size_t)col * ps;
    cpx<T> v[R];
    for (int m = 0; m < R; ++m) v[m] = src[j + m * LR];
    const int k = j % Ns;
    if (Ns > 1) {
      for (int m = 1; m < R; ++m) {
        cpx<T> w = tw[k * m * twstep];
        if (INV) w = conj(w);
        v[m] = v[m] * w;
      }
    }
    dft_small<T, R, INV>(v, tw, L);
    cpx<T>* dst = out + (size_t)col * ps + (j / Ns) * Ns * R + k;
    for (int m = 0; m < R; ++m) dst[m * Ns] = v[m];
  }
}

template <typename T, bool INV>
E3_HD void stage_dispatch(int R, const cpx<T>* in, cpx<T>* out, int ncol, int ps, int L, int Ns,
                          const cpx<T>* tw, int tid, int nt) {
  switch (R) {
    case 2: stockham_stage<T, 2, INV>(in, out, ncol, ps, L, Ns, tw, tid, nt); break;
    case 3: stockham_stage<T, 3, INV>(in, out, ncol, ps, L, Ns, tw, tid, nt); break;
    case 4: stockham_stage<T, 4, INV>(in, out, ncol, ps, L, Ns, tw, tid, nt); break;
    case 5: stockham_stage<T, 5, INV>(in, out, ncol, ps, L, Ns, tw, tid, nt); break;
    case 7: stockham_stage<T, 7, INV>(in, out, ncol, ps, L, Ns, tw, tid, nt); break;
    case 8: stockham_stage<T, 8, INV>(in, out, ncol, ps, L, Ns, tw, tid, nt); break;
    default: break;
  }
}

// Stage s of a plan: reads buffer (s & 1), writes buffer (s & 1) ^ 1.
template <typename T>
E3_HD void fft_stage(const AxisPlan& pl, int s, bool inv, cpx<T>* b0, cpx<T>* b1, int ncol, int ps,
                     const cpx<T>* tw, int tid, int nt) {
  int Ns = 1;
  for (int i = 0; i < s; ++i) Ns *= pl.rad[i];
  const cpx<T>* in = (s & 1) ? b1 : b0;
  cpx<T>* out = (s & 1) ? b0 : b1;
  if (inv) stage_dispatch<T, true>(pl.rad[s], in, out, ncol, ps, pl.L, Ns, tw, tid, nt);
  else stage_dispatch<T, false>(pl.rad[s], in, out, ncol, ps, pl.L, Ns, tw, tid, nt);
}

// Buffer that holds the result after all stages of a plan.
template <typename T>
E3_HD cpx<T>* fft_result(const AxisPlan& pl, cpx<T>* b0, cpx<T>* b1) {
  return (pl.nst & 1) ? b1 : b0;
}

// ---------------------------------------------------------------------------
// contrast helpers: w = chi u, w = chi^H t at one cell
// ---------------------------------------------------------------------------
template <typename T>
E3_HD void chi_apply(const cpx<T>* chi, int kind, long long ncell, long long cell, int C,
                     const cpx<T>* u, cpx<T>* w, bool herm) {
  if (kind == CHI_NONE) {
    for (int c = 0; c < C; ++c) w[c] = u[c];
    return;
  }
  if (kind == CHI_ISO) {
    cpx<T> x = chi[cell];
    if (herm) x = conj(x);
    for (int c = 0; c < C; ++c) w[c] = x * u[c];
    return;
  }
  if (kind == CHI_DIAG) {
    for (int c = 0; c < C; ++c) {
      cpx<T> x = chi[c * ncell + cell];
      if (herm) x = conj(x);
      w[c] = x * u[c];
    }
    return;
  }
  // full 3x3: w_a = sum_b chi_ab u_b, or (chi^H t)_a = sum_b conj(chi_ba) t_b
  for (int a = 0; a < 3; ++a) {
    cpx<T> acc = czero<T>();
    for (int b = 0; b < 3; ++b) {
      if (herm) acc = acc + cmulc(chi[(b * 3 + a) * ncell + cell], u[b]);
      else acc = acc + chi[(a * 3 + b) * ncell + cell] * u[b];
    }
    w[a] = acc;
  }
}

// ---------------------------------------------------------------------------
// Mirror-interleaved order of a doubled axis (length L = 2N) in the
// intermediates W1/W2: position 0 <- k = 0, 1 <- k = N, 2i <- k = i and
// 2i + 1 <- k = L - i (0 < i < N).  Frequencies k and L - k, which share one
// octant entry of the spectrum, become neighbours, so a 16-wide column tile of
// the x pass reads 8 distinct octant entries instead of 16.
// ---------------------------------------------------------------------------
E3_HD int mirror_position(int k, int N, int L) {
  if (k == 0) return 0;
  if (k == N) return 1;
  return k < N ? 2 * k : 2 * (L - k) + 1;
}
E3_HD int mirror_frequency(int pos, int N, int L) {
  if (pos == 0) return 0;
  if (pos == 1) return N;
  return (pos & 1) ? L - (pos >> 1) : (pos >> 1);
}

// ---------------------------------------------------------------------------
// Z pass: transform along the contiguous axis, C components per pencil.
// ---------------------------------------------------------------------------
struct ZArgs {
  const void* in;
  void* out;
  int C;
  long long P;                 // pencils per component
  long long in_ps, in_cs, in_bs;
  long long out_ps, out_cs, out_bs;
  int n_in, n_out;             // live input length, kept output length
  int T;                       // pencils per block
  int ps;                      // LDS pencil stride
  AxisPlan plan;
  const void* tw;
  int inv;
  int pro;                     // Prologue
  int epi;                     // Epilogue
  const void* chi;
  int chi_kind;
  long long ncell;
  const void* u;               // epilogue source, same layout as out
  int batch;
  int perm;                    // doubled axis in mirror-interleaved order (output of a
                               // forward pass, input of an inverse pass)
  int variant;         // fast-kernel variant chosen by the plan (-1: the generic kernels)
  int nt;              // fast kernels: rows read and written non-temporally
  const int* skip;             // device flag: when non-zero the whole pass is a no-op
  ZKry k;                      // fused Krylov step (k.op == KRY_NONE: none)
  // input window of a forward pass (wn1 == 0: none): logical pencil q of the
  // P = n0 n1 pencils is grid pencil wbase + (q / wn1) wny + q % wn1, and
  // input z outside [wz0, wz1) reads as zero
  int wn1, wny, wz0, wz1;
  long long wbase;
};

// grid pencil of logical pencil q, and whether input z lies in the window
E3_HD long long z_pencil(const ZArgs& a, long long q) {
  if (!a.wn1) return q;
  const int qi = (int)q, r = qi / a.wn1;
  return a.wbase + (long long)r * a.wny + (qi - r * a.wn1);
}
E3_HD bool z_live(const ZArgs& a, int z) { return !a.wz1 || (z >= a.wz0 && z < a.wz1); }

// The Krylov prologue value of component c at element o (offset within the
// right-hand side b), before the contrast: KRY_P p, KRY_S s (and |s|^2 into
// acc).  Same arithmetic, in T, as the stand-alone vector kernels.
template <typename T>
E3_HD cpx<T> kry_update(const ZArgs& a, int b, size_t o, KAcc& acc) {
  const KryCoef& k = a.k.coef[b];
  const size_t bo = (size_t)b * a.in_bs + o;
  const cpx<T> r = ((const cpx<T>*)a.k.r)[bo];
  const cpx<T> v = ((const cpx<T>*)a.k.v)[bo];
  cpx<T>* w = (cpx<T>*)a.k.w + bo;
  cpx<T> x;
  if (a.k.op == KRY_P) {
    if (k.first) {
      x = r;
    } else {
      const cpx<T> om = mk((T)k.omega[0], (T)k.omega[1]), be = mk((T)k.beta[0], (T)k.beta[1]);
      const cpx<T> d = *w - om * v;
      x = r + be * d;
    }
  } else {
    const cpx<T> al = mk((T)k.alpha[0], (T)k.alpha[1]);
    x = r - al * v;
    acc.v[0] += (double)x.re * x.re + (double)x.im * x.im;
  }
  *w = x;
  return x;
}

// Krylov epilogue sums of one stored output y (src: the matvec input at the
// same element, b-relative offset o)
template <typename T>
E3_HD void kry_epilogue(const ZArgs& a, int b, size_t o, cpx<T> y, cpx<T> src, KAcc& acc) {
  if (a.k.op == KRY_DOT) {
    const cpx<T> h = ((const cpx<T>*)a.k.rh)[(size_t)b * a.out_bs + o];
    acc.v[0] += (double)h.re * y.re + (double)h.im * y.im;
    acc.v[1] += (double)h.re * y.im - (double)h.im * y.re;
  } else if (a.k.op == KRY_DOT2) {
    acc.v[0] += (double)y.re * y.re + (double)y.im * y.im;
    acc.v[2] += (double)y.re * src.re + (double)y.im * src.im;
    acc.v[3] += (double)y.re * src.im - (double)y.im * src.re;
  }
}

template <typename T>
struct ZPass {
  const ZArgs& a;
  long long p0;
  int b;
  cpx<T>* b0;
  cpx<T>* b1;
  E3_HD ZPass(const ZArgs& args, long long blk, int bidx, cpx<T>* lds)
      : a(args), p0(blk * args.T), b(bidx), b0(lds), b1(lds + (size_t)args.C * args.T * args.ps) {}

  E3_HD void load(int tid, int nt, KAcc& acc) const {
    const cpx<T>* in = (const cpx<T>*)a.in + (size_t)b * a.in_bs;
    const int L = a.plan.L;
    const bool kry = (a.k.op == KRY_P || a.k.op == KRY_S) && !a.k.coef[b].done;
    for (int e = tid; e < a.T * L; e += nt) {
      const int t = e / L, z = e - t * L;
      const long long p = z_pencil(a, p0 + t);
      cpx<T> w[6];
      if (p0 + t < a.P && z < a.n_in && z_live(a, z)) {
        cpx<T> v[6];
        const int zi = (a.perm && a.inv) ? mirror_position(z, a.n_in / 2, a.n_in) : z;
        for (int c = 0; c < a.C; ++c) {
          const size_t o = c * a.in_cs + p * a.in_ps + zi;
          v[c] = kry ? kry_update<T>(a, b, o, acc) : in[o];
        }
        if (a.pro == PRO_CHI)
          chi_apply<T>((const cpx<T>*)a.chi, a.chi_kind, a.ncell, p * a.n_in + z, a.C, v, w, false);
        else
          for (int c = 0; c < a.C; ++c) w[c] = v[c];
      } else {
        for (int c = 0; c < a.C; ++c) w[c] = czero<T>();
      }
      for (int c = 0; c < a.C; ++c) b0[(size_t)(c * a.T + t) * a.ps + z] = w[c];
    }
  }
  E3_HD void stage(int s, int tid, int nt) const {
    fft_stage<T>(a.plan, s, a.inv != 0, b0, b1, a.C * a.T, a.ps, (const cpx<T>*)a.tw, tid, nt);
  }
  E3_HD void store(int tid, int nt, KAcc& acc) const {
    const cpx<T>* res = fft_result<T>(a.plan, b0, b1);
    cpx<T>* out = (cpx<T>*)a.out + (size_t)b * a.out_bs;
    const cpx<T>* src = (const cpx<T>*)a.u + (size_t)b * a.out_bs;
    for (int e = tid; e < a.T * a.n_out; e += nt) {
      const int t = e / a.n_out, z = e - t * a.n_out;
      if (p0 + t >= a.P) continue;
      const long long p = z_pencil(a, p0 + t);
      cpx<T> v[6];
      for (int c = 0; c < a.C; ++c) v[c] = res[(size_t)(c * a.T + t) * a.ps + z];
      if (a.epi == EPI_SUB_CHIH) {
        cpx<T> w[6];
        chi_apply<T>((const cpx<T>*)a.chi, a.chi_kind, a.ncell, p * a.n_out + z, a.C, v, w, true);
        for (int c = 0; c < a.C; ++c) v[c] = w[c];
      }
      const int zo = (a.perm && !a.inv) ? mirror_position(z, a.n_out / 2, a.n_out) : z;
      for (int c = 0; c < a.C; ++c) {
        const size_t o = c * a.out_cs + p * a.out_ps + zo;
        if (a.epi == EPI_STORE) {
          out[o] = v[c];
        } else {
          const cpx<T> x = src[o], y = x - v[c];
          out[o] = y;
          if (a.k.op == KRY_DOT || a.k.op == KRY_DOT2) kry_epilogue<T>(a, b, o, y, x, acc);
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// S pass: transform along a strided axis.  Data viewed as (Q, rows, W) with
// row stride W; a block takes T consecutive columns of one plane.
// ---------------------------------------------------------------------------
struct SArgs {
  const void* in;
  void* out;
  long long Q, W;
  long long in_qs, out_qs;
  int n_in, n_out;
  int T, ps;
  AxisPlan plan;
  const void* tw;
  int inv;
  int variant;         // fast-kernel variant chosen by the plan (-1: the generic kernels)
  int nt;              // fast kernels: rows read and written non-temporally
  const int* skip;             // device flag: when non-zero the whole pass is a no-op
  // input window of a forward pass (wn == 0: none): logical plane q of Q is
  // plane (q / wn) wnq + wlo + q % wn, and input rows outside [wr0, wr1)
  // read as zero
  int wn, wnq, wlo, wr0, wr1;
};

E3_HD long long s_plane(const SArgs& a, long long q) {
  if (!a.wn) return q;
  const long long r = q / a.wn;
  return r * a.wnq + a.wlo + (q - r * a.wn);
}
E3_HD bool s_live(const SArgs& a, int row) { return !a.wr1 || (row >= a.wr0 && row < a.wr1); }

template <typename T>
struct SPass {
  const SArgs& a;
  long long q, w0;
  cpx<T>* b0;
  cpx<T>* b1;
  E3_HD SPass(const SArgs& args, long long colblk, long long plane, cpx<T>* lds)
      : a(args), q(s_plane(args, plane)), w0(colblk * args.T), b0(lds), b1(lds + (size_t)args.T * args.ps) {}

  E3_HD void load(int tid, int nt) const {
    const cpx<T>* in = (const cpx<T>*)a.in + (size_t)q * a.in_qs;
    const int L = a.plan.L;
    for (int e = tid; e < a.T * L; e += nt) {
      const int row = e / a.T, col = e - row * a.T;
      const long long w = w0 + col;
      cpx<T> v = czero<T>();
      if (row < a.n_in && w < a.W && s_live(a, row)) v = in[(size_t)row * a.W + w];
      b0[(size_t)col * a.ps + row] = v;
    }
  }
  E3_HD void stage(int s, int tid, int nt) const {
    fft_stage<T>(a.plan, s, a.inv != 0, b0, b1, a.T, a.ps, (const cpx<T>*)a.tw, tid, nt);
  }
  E3_HD void store(int tid, int nt) const {
    const cpx<T>* res = fft_result<T>(a.plan, b0, b1);
    cpx<T>* out = (cpx<T>*)a.out + (size_t)q * a.out_qs;
    for (int e = tid; e < a.T * a.n_out; e += nt) {
      const int row = e / a.T, col = e - row * a.T;
      const long long w = w0 + col;
      if (w < a.W) out[(size_t)row * a.W + w] = res[(size_t)col * a.ps + row];
    }
  }
};

// ---------------------------------------------------------------------------
// X pass: forward transform along x (Nx live rows), multiply by the octant
// spectrum (3x3 symmetric or scalar), inverse transform, keep Nx rows.  In
// place on W2 viewed per (b, c) as (Nx, W = Ly*Lz).
// ---------------------------------------------------------------------------
struct XArgs {
  void* data;
  int C;
  long long W, cs, bs;
  int N[3], L[3];
  int T, ps;
  AxisPlan plan;
  const void* tw;
  const void* khat;
  int adj;
  int batch;
  int perm;            // z axis of W2 in mirror-interleaved order
  int variant;         // fast-kernel variant chosen by the plan (-1: the generic kernels)
  int nt;              // fast kernels: rows read and written non-temporally
  const int* skip;     // device flag: when non-zero the whole pass is a no-op
  int wx0, wx1;        // input window of the forward transform (wx1 == 0: none): rows outside read as zero
};

E3_HD bool x_live(const XArgs& a, int x) { return !a.wx1 || (x >= a.wx0 && x < a.wx1); }

// ---------------------------------------------------------------------------
// Octant spectrum layout.  Entries (comp, ix, iy, iz), 0 <= i* <= N*, are
// stored as
//   [iy][iz / S][ix][comp][iz % S]          for 0 <= iz <= Nz
// in groups of S = 64 bytes of iz (S = 8 complex64 or 4 complex128), so that
// the entries one x-pass block needs (one iy, S consecutive iz, all ix and
// components) are one contiguous run, and every column, iz = Nz included,
// has the same strides.  G = Nz / S + 1 groups per iy; slots past Nz in the
// last group are padding (never read).
// ---------------------------------------------------------------------------
struct KhatLayout {
  int nx1, ny1, nz, K, S, G;
  E3_HD long long row_elems() const { return (long long)G * nx1 * K * S; }
  E3_HD long long total() const { return (long long)ny1 * row_elems(); }
  E3_HD long long index(int comp, int ix, int iy, int iz) const {
    return ((((long long)iy * G + iz / S) * nx1 + ix) * K + comp) * S + iz % S;
  }
};

E3_HD KhatLayout khat_layout(const int N[3], int ncomp, int elem_bytes) {
  KhatLayout l;
  l.nx1 = N[0] + 1;
  l.ny1 = N[1] + 1;
  l.nz = N[2];
  l.K = ncomp == 3 ? 6 : 1;
  l.S = 64 / elem_bytes;
  l.G = N[2] / l.S + 1;
  return l;
}

template <typename T>
E3_HD void spectral_multiply(const cpx<T>* K, int C, int N[3], int L[3], int kx, int ky, int kz,
                             bool adj, cpx<T>* f) {
  int fx, fy, fz;
  const int ix = fold_index(kx, N[0], L[0], &fx);
  const int iy = fold_index(ky, N[1], L[1], &fy);
  const int iz = fold_index(kz, N[2], L[2], &fz);
  const KhatLayout lay = khat_layout(N, C, (int)sizeof(cpx<T>));
  if (C == 1) {
    cpx<T> k = K[lay.index(0, ix, iy, iz)];
    if (adj) k = conj(k);
    f[0] = k * f[0];
    return;
  }
  cpx<T> k[6];
  for (int i = 0; i < 6; ++i) {
    k[i] = K[lay.index(i, ix, iy, iz)];
    if (adj) k[i] = conj(k[i]);
  }
  const T sx = fx ? T(-1) : T(1), sy = fy ? T(-1) : T(1), sz = fz ? T(-1) : T(1);
  const cpx<T> kxy = (sx * sy) * k[3], kxz = (sx * sz) * k[4], kyz = (sy * sz) * k[5];
  const cpx<T> f0 = f[0], f1 = f[1], f2 = f[2];
  f[0] = k[0] * f0 + kxy * f1 + kxz * f2;
  f[1] = kxy * f0 + k[1] * f1 + kyz * f2;
  f[2] = kxz * f0 + kyz * f1 + k[2] * f2;
}

template <typename T>
struct XPass {
  const XArgs& a;
  long long w0;
  int b;
  cpx<T>* b0;
  cpx<T>* b1;
  E3_HD XPass(const XArgs& args, long long colblk, int bidx, cpx<T>* lds)
      : a(args), w0(colblk * args.T), b(bidx), b0(lds), b1(lds + (size_t)args.C * args.T * args.ps) {}

  E3_HD void load(int tid, int nt) const {
    const cpx<T>* d = (const cpx<T>*)a.data + (size_t)b * a.bs;
    const int L = a.plan.L, TL = a.T * L;
    for (int e = tid; e < a.C * TL; e += nt) {
      const int c = e / TL, r = e - c * TL;
      const int row = r / a.T, col = r - row * a.T;
      const long long w = w0 + col;
      cpx<T> v = czero<T>();
      if (row < a.N[0] && w < a.W && x_live(a, row)) v = d[(size_t)c * a.cs + (size_t)row * a.W + w];
      b0[(size_t)(c * a.T + col) * a.ps + row] = v;
    }
  }
  E3_HD void stage_fwd(int s, int tid, int nt) const {
    fft_stage<T>(a.plan, s, false, b0, b1, a.C * a.T, a.ps, (const cpx<T>*)a.tw, tid, nt);
  }
  // After the forward stages the spectrum sits in fft_result(b0,b1); the
  // multiply runs in place there and the inverse stages continue from it.
  E3_HD cpx<T>* spec() const { return fft_result<T>(a.plan, b0, b1); }
  E3_HD void multiply(int tid, int nt) const {
    cpx<T>* sp = spec();
    const int L = a.plan.L;
    for (int e = tid; e < a.T * L; e += nt) {
      const int col = e / L, kx = e - col * L;
      const long long w = w0 + col;
      if (w >= a.W) continue;
      const int ky = (int)(w / a.L[2]);
      int kz = (int)(w - (long long)ky * a.L[2]);
      if (a.perm) kz = mirror_frequency(kz, a.N[2], a.L[2]);
      cpx<T> f[3];
      for (int c = 0; c < a.C; ++c) f[c] = sp[(size_t)(c * a.T + col) * a.ps + kx];
      int N[3] = {a.N[0], a.N[1], a.N[2]}, Ld[3] = {a.L[0], a.L[1], a.L[2]};
      spectral_multiply<T>((const cpx<T>*)a.khat, a.C, N, Ld, kx, ky, kz, a.adj != 0, f);
      for (int c = 0; c < a.C; ++c) sp[(size_t)(c * a.T + col) * a.ps + kx] = f[c];
    }
  }
  // inverse stage s reads/writes starting from the spectrum buffer
  E3_HD void stage_inv(int s, int tid, int nt) const {
    cpx<T>* first = spec();
    cpx<T>* other = (first == b0) ? b1 : b0;
    fft_stage<T>(a.plan, s, true, first, other, a.C * a.T, a.ps, (const cpx<T>*)a.tw, tid, nt);
  }
  E3_HD void store(int tid, int nt) const {
    cpx<T>* first = spec();
    cpx<T>* other = (first == b0) ? b1 : b0;
    const cpx<T>* res = fft_result<T>(a.plan, first, other);
    cpx<T>* d = (cpx<T>*)a.data + (size_t)b * a.bs;
    const int TN = a.T * a.N[0];
    for (int e = tid; e < a.C * TN; e += nt) {
      const int c = e / TN, r = e - c * TN;
      const int row = r / a.T, col = r - row * a.T;
      const long long w = w0 + col;
      if (w < a.W) d[(size_t)c * a.cs + (size_t)row * a.W + w] = res[(size_t)(c * a.T + col) * a.ps + row];
    }
  }
};

// ---------------------------------------------------------------------------
// Kernel build (FP64): spatial kernel on the doubled grid with the index-N
// planes set to zero, then (after a full 3-D FFT) the octant extraction.
// ---------------------------------------------------------------------------
struct KBArgs {
  void* full;          // complex<double> (K, Lx, Ly, Lz)
  void* oct;           // complex<T> (K, Nx+1, Ny+1, Nz+1)
  int kind;            // 0 EM (K = 6), 1 acoustic (K = 1)
  double k0, h[3], dv;
  double self_re, self_im;   // acoustic self-cell value
  int N[3], L[3];
  double scale;        // 1 / (Lx Ly Lz)
};

E3_HD double grid_offset(int i, int N, double h) { return (i < N) ? i * h : (i - 2 * N) * h; }

E3_HD void kernel_point(const KBArgs& a, long long idx) {
  cpx<double>* out = (cpx<double>*)a.full;
  const long long D = (long long)a.L[0] * a.L[1] * a.L[2];
  const int iz = (int)(idx % a.L[2]);
  const int iy = (int)((idx / a.L[2]) % a.L[1]);
  const int ix = (int)(idx / ((long long)a.L[2] * a.L[1]));
  const int ncomp = a.kind == 0 ? 6 : 1;
  const bool nyq = (ix == a.N[0]) || (iy == a.N[1]) || (iz == a.N[2]);
  if (nyq) {
    for (int c = 0; c < ncomp; ++c) out[c * D + idx] = mk(0.0, 0.0);
    return;
  }
  const double s[3] = {grid_offset(ix, a.N[0], a.h[0]), grid_offset(iy, a.N[1], a.h[1]),
                       grid_offset(iz, a.N[2], a.h[2])};
  const double R = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  const double k = a.k0;
  if (R < 1e-15) {
    if (a.kind == 0) {
      for (int c = 0; c < 3; ++c) out[c * D + idx] = mk(-1.0 / 3.0, 0.0);
      for (int c = 3; c < 6; ++c) out[c * D + idx] = mk(0.0, 0.0);
    } else {
      out[idx] = mk(a.self_re, a.self_im);
    }
    return;
  }
  const double inv = 1.0 / R;
  const double g = 1.0 / (4.0 * 3.14159265358979323846 * R);
  const cpx<double> e = mk(cos(k * R) * g, sin(k * R) * g);   // exp(ikR)/(4 pi R)
  if (a.kind == 1) {
    out[idx] = (k * k * a.dv) * e;
    return;
  }
  const cpx<double> wg = a.dv * e;
  const cpx<double> c1 = mk(3.0 * inv * inv - k * k, -3.0 * k * inv);
  const cpx<double> c2 = mk(k * k - inv * inv, k * inv);
  const cpx<double> w1 = wg * c1, w2 = wg * c2;
  const double al[3] = {s[0] * inv, s[1] * inv, s[2] * inv};
  out[0 * D + idx] = (al[0] * al[0]) * w1 + w2;
  out[1 * D + idx] = (al[1] * al[1]) * w1 + w2;
  out[2 * D + idx] = (al[2] * al[2]) * w1 + w2;
  out[3 * D + idx] = (al[0] * al[1]) * w1;
  out[4 * D + idx] = (al[0] * al[2]) * w1;
  out[5 * D + idx] = (al[1] * al[2]) * w1;
}

// idx runs over the logical entries (comp, ix, iy, iz), iz fastest.
template <typename T>
E3_HD void octant_point(const KBArgs& a, long long idx) {
  const cpx<double>* full = (const cpx<double>*)a.full;
  cpx<T>* oct = (cpx<T>*)a.oct;
  const long long nz1 = a.N[2] + 1, ny1 = a.N[1] + 1, nx1 = a.N[0] + 1;
  const long long per = nx1 * ny1 * nz1;
  const int c = (int)(idx / per);
  const long long r = idx - c * per;
  const int iz = (int)(r % nz1);
  const int iy = (int)((r / nz1) % ny1);
  const int ix = (int)(r / (nz1 * ny1));
  const long long D = (long long)a.L[0] * a.L[1] * a.L[2];
  cpx<double> v = full[c * D + ((long long)ix * a.L[1] + iy) * a.L[2] + iz];
  // the off-diagonal components xy, xz, yz are odd along both their axes, so
  // their spectrum vanishes on those axes' index-N planes: stored as exact
  // zeros (the FFT leaves rounding residue there), which lets the x pass
  // take either sign for kx = Nx.
  if (a.kind == 0 && c >= 3) {
    const bool ox = c != 5 && ix == a.N[0], oy = c != 4 && iy == a.N[1], oz = c != 3 && iz == a.N[2];
    if (ox || oy || oz) v = mk(0.0, 0.0);
  }
  const KhatLayout lay = khat_layout(a.N, a.kind == 0 ? 3 : 1, (int)sizeof(cpx<T>));
  oct[lay.index(c, ix, iy, iz)] = mk((T)(v.re * a.scale), (T)(v.im * a.scale));
}

}  // namespace e3
