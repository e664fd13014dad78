// Plan construction and the pass sequences of the operator (host logic).
//
// The sequences are templated on a Launcher so the same host code drives the
// HIP kernels (e3_capi.hip) and the CPU block-phase emulator of the tests.
#pragma once

#include <cstring>
#include <string>

#include "e3_blocks.hpp"

namespace e3 {

enum Precision { PREC_SINGLE = 0, PREC_DOUBLE = 1 };
enum OperatorKind { KIND_EM = 0, KIND_ACOUSTIC = 1 };

constexpr size_t kLdsBudget = 64 * 1024;
constexpr size_t kLdsMax = 160 * 1024;
constexpr int kThreads = 256;

struct Plan {
  int N[3], L[3];
  int C;          // 3 (EM) or 1 (acoustic)
  int batch;      // right-hand sides per call
  int prec;       // Precision
  AxisPlan ax[3];
  const void* tw_f[3];   // forward twiddles exp(-2 pi i t / L), complex<float>
  const void* tw_d[3];   // same in complex<double>
  long long w1_elems, w2_elems;   // workspace in plan-precision complex elements
  long long khat_elems;           // octant spectrum storage elements (with group padding)
  long long khat_entries;         // logical octant entries K (Nx+1)(Ny+1)(Nz+1)
  long long build_elems;          // complex<double> scratch of the kernel build
  int variant[3];                 // kernel variant per pass kind (strided, x, z); -1: generic kernels
};

enum PassKind { PASS_S = 0, PASS_X = 1, PASS_Z = 2 };

inline size_t elem_bytes(int prec) { return prec == PREC_DOUBLE ? 16 : 8; }

inline int tile_count(size_t per_col_bytes, int cap) {
  size_t t = kLdsBudget / (per_col_bytes ? per_col_bytes : 1);
  if (t < 1) t = 1;
  if ((int)t > cap) t = cap;
  return (int)t;
}

// Fills *p; returns an empty string on success, else the reason.
inline std::string plan_init(Plan* p, int nx, int ny, int nz, int ncomp, int batch, int prec) {
  std::memset(p, 0, sizeof(*p));
  const int n[3] = {nx, ny, nz};
  if (ncomp != 1 && ncomp != 3) return "ncomp must be 1 (acoustic) or 3 (EM)";
  if (batch < 1) return "batch must be >= 1";
  if (prec != PREC_SINGLE && prec != PREC_DOUBLE) return "precision must be 0 (single) or 1 (double)";
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 1) return "grid sizes must be positive";
    if (n[a] > 4096) return "grid size above 4096 per axis is not supported";
    p->N[a] = n[a];
    p->L[a] = 2 * n[a];
    if (!make_axis_plan(p->L[a], &p->ax[a]))
      return "doubled axis length " + std::to_string(p->L[a]) +
             " has a prime factor above 7; choose grid sizes whose prime factors are 2, 3, 5, 7";
    const size_t need = 2 * (size_t)(p->L[a] + 1) * 16 * 3;
    if (need > kLdsMax) return "axis too long for one LDS tile";
  }
  p->C = ncomp;
  p->batch = batch;
  p->prec = prec;
  const long long B = batch, C = ncomp;
  p->w1_elems = B * C * n[0] * n[1] * (long long)p->L[2];
  p->w2_elems = B * C * n[0] * (long long)p->L[1] * p->L[2];
  const long long K = ncomp == 3 ? 6 : 1;
  p->khat_elems = khat_layout(p->N, ncomp, (int)elem_bytes(prec)).total();
  p->khat_entries = K * (long long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
  p->build_elems = K * (long long)p->L[0] * p->L[1] * p->L[2];
  return "";
}

inline size_t lds_bytes(int ncols, int ps, int prec) { return 2 * (size_t)ncols * ps * elem_bytes(prec); }

// ---------------------------------------------------------------------------
// Pass sequences
// ---------------------------------------------------------------------------
template <class Launcher>
void z_pass(Launcher& go, const Plan& p, int prec, const void* tw, const AxisPlan& ax, ZArgs z) {
  z.plan = ax;
  z.tw = tw;
  z.ps = ax.L + 1;
  const size_t per = lds_bytes(z.C, z.ps, prec);
  z.T = tile_count(per, 64);
  if (z.T > z.P) z.T = (int)z.P;
  const long long nblk = (z.P + z.T - 1) / z.T;
  go.z(z, nblk, z.batch, prec, lds_bytes(z.C * z.T, z.ps, prec));
}

template <class Launcher>
void s_pass(Launcher& go, int prec, const void* tw, const AxisPlan& ax, SArgs s) {
  s.plan = ax;
  s.tw = tw;
  s.ps = ax.L + 1;
  s.T = tile_count(lds_bytes(1, s.ps, prec), 16);
  if (s.T > s.W) s.T = (int)s.W;
  const long long ncb = (s.W + s.T - 1) / s.T;
  go.s(s, ncb, s.Q, prec, lds_bytes(s.T, s.ps, prec));
}

// Passes of run_apply (bit i: pass P(i+1)); a solver runs P1 and P2..P5 of
// one matvec as two calls when a decision falls between them.
enum : unsigned { PASSES_ALL = 0x1Fu, PASSES_FIRST = 0x01u, PASSES_REST = 0x1Eu };

// y = A u (adjoint = false) or y = A^H u (adjoint = true) for all batch RHS.
// k1 / k5: fused Krylov steps of the first / last z pass (device BiCGStab);
// skip: device flag of every pass run here (skip1, when given, of P1).
// win (lo[3], n[3]; null: none): the forward passes read only the cells of
// that box, i.e. the transformed field (chi u, or u for the adjoint) is taken
// as zero outside it.  P1 runs on the box's n0 n1 pencils, P2 on its n0 x
// planes and n1 rows, P3 on its n0 rows; P4 and P5 cover the whole grid.
template <class Launcher>
void run_apply(Launcher& go, const Plan& p, const void* khat, const void* chi, int chi_kind,
               const void* u, void* y, void* work, bool adjoint, const int* skip = nullptr,
               const ZKry* k1 = nullptr, const ZKry* k5 = nullptr, const int* skip1 = nullptr,
               unsigned passes = PASSES_ALL, const int* win = nullptr) {
  const int prec = p.prec;
  const size_t es = elem_bytes(prec);
  void* W1 = work;
  void* W2 = (char*)work + (size_t)p.w1_elems * es;
  const int Nx = p.N[0], Ny = p.N[1], Nz = p.N[2];
  const int Lx = p.L[0], Ly = p.L[1], Lz = p.L[2];
  const long long P = (long long)Nx * Ny, C = p.C;
  const void* twz = prec == PREC_DOUBLE ? p.tw_d[2] : p.tw_f[2];
  const void* twy = prec == PREC_DOUBLE ? p.tw_d[1] : p.tw_f[1];
  const void* twx = prec == PREC_DOUBLE ? p.tw_d[0] : p.tw_f[0];
  // the x pass's working set (W2) beyond 128 MiB: the forward z, y and x passes stream non-temporally
  const int nt = (double)p.batch * C * Nx * Ly * Lz * es > 128.0 * 1024 * 1024 ? 1 : 0;

  // P1: z forward, chi prologue (matvec only), zero half skipped
  ZArgs z;
  std::memset(&z, 0, sizeof(z));
  z.in = u; z.out = W1; z.C = p.C; z.P = P; z.batch = p.batch;
  z.in_ps = Nz; z.in_cs = P * Nz; z.in_bs = C * P * Nz;
  z.out_ps = Lz; z.out_cs = P * Lz; z.out_bs = C * P * Lz;
  z.n_in = Nz; z.n_out = Lz; z.inv = 0;
  z.pro = adjoint ? PRO_COPY : PRO_CHI; z.epi = EPI_STORE;
  z.chi = chi; z.chi_kind = chi_kind; z.ncell = P * Nz; z.perm = 1; z.skip = skip1 ? skip1 : skip;
  z.variant = p.variant[PASS_Z];
  z.nt = nt;
  if (k1) z.k = *k1;
  if (win) {
    z.P = (long long)win[3] * win[4];
    z.wn1 = win[4]; z.wny = Ny; z.wbase = (long long)win[0] * Ny + win[1];
    z.wz0 = win[2]; z.wz1 = win[2] + win[5];
  }
  if (passes & 0x01u) z_pass(go, p, prec, twz, p.ax[2], z);

  // P2: y forward (Ny live rows -> Ly)
  SArgs s;
  std::memset(&s, 0, sizeof(s));
  s.in = W1; s.out = W2; s.Q = (long long)p.batch * C * Nx; s.W = Lz;
  s.in_qs = (long long)Ny * Lz; s.out_qs = (long long)Ly * Lz;
  s.n_in = Ny; s.n_out = Ly; s.inv = 0; s.skip = skip; s.variant = p.variant[PASS_S]; s.nt = nt;
  if (win) {
    s.Q = (long long)p.batch * C * win[3];
    s.wn = win[3]; s.wnq = Nx; s.wlo = win[0];
    s.wr0 = win[1]; s.wr1 = win[1] + win[4];
  }
  if (passes & 0x02u) s_pass(go, prec, twy, p.ax[1], s);
  s.Q = (long long)p.batch * C * Nx;
  s.wn = s.wnq = s.wlo = s.wr0 = s.wr1 = 0;

  // P3: x forward, spectral multiply, x inverse (Nx rows kept), in place
  XArgs x;
  std::memset(&x, 0, sizeof(x));
  x.data = W2; x.C = p.C; x.W = (long long)Ly * Lz; x.cs = (long long)Nx * x.W; x.bs = C * x.cs;
  for (int a = 0; a < 3; ++a) { x.N[a] = p.N[a]; x.L[a] = p.L[a]; }
  x.plan = p.ax[0]; x.tw = twx; x.khat = khat; x.adj = adjoint ? 1 : 0; x.batch = p.batch; x.perm = 1; x.skip = skip;
  x.variant = p.variant[PASS_X];
  x.nt = nt;
  if (win) {
    x.wx0 = win[0]; x.wx1 = win[0] + win[3];
  }
  x.ps = Lx + 1;
  x.T = tile_count(lds_bytes(p.C, x.ps, prec), 16);
  if (x.T > x.W) x.T = (int)x.W;
  if (passes & 0x04u) go.x(x, (x.W + x.T - 1) / x.T, p.batch, prec, lds_bytes(p.C * x.T, x.ps, prec));

  // P4: y inverse (Ny rows kept)
  s.in = W2; s.out = W1; s.in_qs = (long long)Ly * Lz; s.out_qs = (long long)Ny * Lz;
  s.n_in = Ly; s.n_out = Ny; s.inv = 1;
  if (passes & 0x08u) s_pass(go, prec, twy, p.ax[1], s);

  // P5: z inverse, Nz kept, epilogue y = u - t or y = v - chi^H t
  std::memset(&z, 0, sizeof(z));
  z.in = W1; z.out = y; z.C = p.C; z.P = P; z.batch = p.batch;
  z.in_ps = Lz; z.in_cs = P * Lz; z.in_bs = C * P * Lz;
  z.out_ps = Nz; z.out_cs = P * Nz; z.out_bs = C * P * Nz;
  z.n_in = Lz; z.n_out = Nz; z.inv = 1;
  z.pro = PRO_COPY; z.epi = adjoint ? EPI_SUB_CHIH : EPI_SUB;
  z.chi = chi; z.chi_kind = chi_kind; z.ncell = P * Nz; z.u = u; z.perm = 1; z.skip = skip; z.variant = p.variant[PASS_Z];
  if (k5) z.k = *k5;
  if (passes & 0x10u) z_pass(go, p, prec, twz, p.ax[2], z);
}

// Octant spectrum of the operator kernel, built in FP64 then cast.
template <class Launcher>
void run_kernel_build(Launcher& go, const Plan& p, double k0, const double h[3], double self_re,
                      double self_im, void* khat, void* build_work) {
  KBArgs kb;
  std::memset(&kb, 0, sizeof(kb));
  kb.full = build_work; kb.oct = khat;
  kb.kind = p.C == 3 ? KIND_EM : KIND_ACOUSTIC;
  kb.k0 = k0;
  for (int a = 0; a < 3; ++a) { kb.h[a] = h[a]; kb.N[a] = p.N[a]; kb.L[a] = p.L[a]; }
  kb.dv = h[0] * h[1] * h[2];
  kb.self_re = self_re; kb.self_im = self_im;
  const long long D = (long long)p.L[0] * p.L[1] * p.L[2];
  kb.scale = 1.0 / (double)D;
  const int K = p.C == 3 ? 6 : 1;
  go.kspatial(kb, D);

  // full forward 3-D FFT in place, FP64
  ZArgs z;
  std::memset(&z, 0, sizeof(z));
  z.in = build_work; z.out = build_work; z.C = 1; z.P = (long long)p.L[0] * p.L[1]; z.batch = K;
  z.in_ps = z.out_ps = p.L[2]; z.in_cs = z.out_cs = 0; z.in_bs = z.out_bs = D;
  z.n_in = z.n_out = p.L[2]; z.inv = 0; z.pro = PRO_COPY; z.epi = EPI_STORE;
  z_pass(go, p, PREC_DOUBLE, p.tw_d[2], p.ax[2], z);

  SArgs s;
  std::memset(&s, 0, sizeof(s));
  s.in = s.out = build_work; s.Q = (long long)K * p.L[0]; s.W = p.L[2];
  s.in_qs = s.out_qs = (long long)p.L[1] * p.L[2]; s.n_in = s.n_out = p.L[1]; s.inv = 0;
  s_pass(go, PREC_DOUBLE, p.tw_d[1], p.ax[1], s);

  s.Q = K; s.W = (long long)p.L[1] * p.L[2]; s.in_qs = s.out_qs = D;
  s.n_in = s.n_out = p.L[0];
  s_pass(go, PREC_DOUBLE, p.tw_d[0], p.ax[0], s);

  go.koct(kb, p.khat_entries, p.prec);
}

}  // namespace e3
