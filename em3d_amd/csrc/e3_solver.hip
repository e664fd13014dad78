// Device-resident BiCGStab on the FFT operator of a plan, for every
// right-hand side of the plan's batch at once.
//
// Same recurrence, early exit on ||s||, statuses and residual-history
// semantics as the reference solver (src/em3d/solvers/bicgstab.py:12-211),
// run independently per right-hand side, with every scalar of the iteration
// kept on the device: the reductions end in "decide" kernels (one block per
// right-hand side) that compute rho, alpha, omega, the relative residuals and
// the status, and append to that right-hand side's residual history.  Once a
// right-hand side has a status its vector kernels return at once; once every
// right-hand side has one (or, for t = A s, has stopped at ||s||), group flags
// turn the five FFT passes of the batched matvec into no-ops too.  The host
// enqueues iterations in chunks and reads the states back once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <string>
#include <vector>

#include "e3_common.hpp"
#include "e3_error.hpp"
#include "e3_fno.hpp"
#include "e3_plan.hpp"

using namespace e3;

namespace {

constexpr int kRB = 1024;   // reduction blocks
constexpr int kRT = 256;    // reduction threads

enum Status {
  ST_RUNNING = -1,
  ST_CONVERGED = 0,
  ST_NONFINITE = 1,
  ST_MAX_ACTIONS = 2,
  ST_BREAKDOWN_RHO = 3,
  ST_BREAKDOWN_OMEGA = 4,
  ST_BREAKDOWN_ALPHA = 5,
  ST_BREAKDOWN_T = 6,
  ST_DIVERGENCE = 7,
  ST_MAX_ITER = 8,
};

struct BiState {
  // inputs
  double rhs_norm, tol, divg;
  long long max_act;
  int max_iter, use_divg, use_max_act;
  // iteration scalars
  cpx<double> rho, rho_prev, alpha, omega, beta;
  double s_rel;
  int it;          // completed updates
  long long mv;    // matvec count
  int status;
  int done;        // nothing more to do: every kernel returns
  int halt;        // stop decided at ||s|| or omega: finish with u += alpha p
  int skip_t;      // done || halt: the t matvec and its reductions are skipped
  int nhist;
  int hist_cap;
  double last_rel;  // most recent history entry
};

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block sum of NV doubles; result valid in thread 0
template <int NV>
__device__ __forceinline__ void bsum(double* v) {
  __shared__ double sh[NV][kRT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    v[i] = wsum(v[i]);
    if (lane == 0) sh[i][w] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double a = 0.0;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) a += sh[i][k];
      v[i] = a;
    }
  }
}

__device__ __forceinline__ double cabs_d(cpx<double> z) { return sqrt(z.re * z.re + z.im * z.im); }
__device__ __forceinline__ cpx<double> cdiv(cpx<double> a, cpx<double> b) {
  const double d = b.re * b.re + b.im * b.im;
  return mk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
template <typename T>
__device__ __forceinline__ cpx<T> cast(cpx<double> z) { return mk((T)z.re, (T)z.im); }

__device__ __forceinline__ void finish(BiState* st, int status) {
  st->status = status;
  st->done = 1;
  st->skip_t = 1;
}
__device__ __forceinline__ void halt(BiState* st, int status) {
  st->status = status;
  st->halt = 1;
  st->skip_t = 1;
}

// ---- vector kernels: grid (blocks, batch); blockIdx.y selects the RHS ------
// part[b][blk] = sum conj(x) y over the block's slice; skip_which 0: done, 1: skip_t
template <typename T>
__global__ __launch_bounds__(kRT) void k_dot(long long n, const cpx<T>* x, const cpx<T>* y, double* part,
                                             const BiState* st, int skip_which) {
  const int b = blockIdx.y;
  if (st && (skip_which ? st[b].skip_t : st[b].done)) return;
  x += (long long)b * n; y += (long long)b * n;
  double v[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const cpx<T> a = x[i], c = y[i];
    v[0] += (double)a.re * c.re + (double)a.im * c.im;
    v[1] += (double)a.re * c.im - (double)a.im * c.re;
  }
  bsum<2>(v);
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)b * kRB + blockIdx.x) * 4;
    o[0] = v[0]; o[1] = v[1];
  }
}

// part = (sum conj(t) t, sum conj(t) s)  (t t is complex like the reference's vdot)
template <typename T>
__global__ __launch_bounds__(kRT) void k_dot2(long long n, const cpx<T>* t, const cpx<T>* s, double* part,
                                              const BiState* st) {
  const int b = blockIdx.y;
  if (st[b].skip_t) return;
  t += (long long)b * n; s += (long long)b * n;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const cpx<T> a = t[i], c = s[i];
    v[0] += (double)a.re * a.re + (double)a.im * a.im;
    v[2] += (double)a.re * c.re + (double)a.im * c.im;
    v[3] += (double)a.re * c.im - (double)a.im * c.re;
  }
  bsum<4>(v);
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)b * kRB + blockIdx.x) * 4;
    for (int i = 0; i < 4; ++i) o[i] = v[i];
  }
}

// p = r (first iteration) or p = r + beta (p - omega v)
template <typename T>
__global__ __launch_bounds__(kRT) void k_p(long long n, cpx<T>* p, const cpx<T>* r, const cpx<T>* v, const BiState* st) {
  const int b = blockIdx.y;
  if (st[b].done) return;
  const long long off = (long long)b * n;
  p += off; r += off; v += off;
  const bool first = st[b].it == 0;
  const cpx<T> be = cast<T>(st[b].beta), om = cast<T>(st[b].omega);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if (first) {
      p[i] = r[i];
    } else {
      const cpx<T> w = p[i] - om * v[i];
      p[i] = r[i] + be * w;
    }
  }
}

// s = r - alpha v; part = sum |s|^2
template <typename T>
__global__ __launch_bounds__(kRT) void k_s(long long n, cpx<T>* s, const cpx<T>* r, const cpx<T>* v, const BiState* st,
                                           double* part) {
  const int b = blockIdx.y;
  if (st[b].done) return;
  const long long off = (long long)b * n;
  s += off; r += off; v += off;
  const cpx<T> al = cast<T>(st[b].alpha);
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const cpx<T> x = r[i] - al * v[i];
    s[i] = x;
    acc += (double)x.re * x.re + (double)x.im * x.im;
  }
  bsum<1>(&acc);
  if (threadIdx.x == 0) part[((size_t)b * kRB + blockIdx.x) * 4] = acc;
}

// u_alpha = u + alpha pu; then (no halt) u = u_alpha + omega su, r = s - omega t, part = sum |r|^2
// (pu, su: the search directions after the right preconditioner, p and s without one)
template <typename T>
__global__ __launch_bounds__(kRT) void k_ur(long long n, cpx<T>* u, const cpx<T>* pu, const cpx<T>* su,
                                            const cpx<T>* s, cpx<T>* r, const cpx<T>* t, const BiState* st,
                                            double* part) {
  const int b = blockIdx.y;
  if (st[b].done) return;
  const long long off = (long long)b * n;
  u += off; pu += off; su += off; s += off; r += off; t += off;
  const bool h = st[b].halt != 0;
  const cpx<T> al = cast<T>(st[b].alpha), om = cast<T>(st[b].omega);
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const cpx<T> ua = u[i] + al * pu[i];
    if (h) {
      u[i] = ua;
    } else {
      const cpx<T> si = s[i];
      u[i] = ua + om * su[i];
      const cpx<T> x = si - om * t[i];
      r[i] = x;
      acc += (double)x.re * x.re + (double)x.im * x.im;
    }
  }
  bsum<1>(&acc);
  if (threadIdx.x == 0) part[((size_t)b * kRB + blockIdx.x) * 4] = acc;
}

// Fused form of k_ur followed by the next iteration's (r_hat, r): u_alpha =
// u + alpha p; then (no halt) u = u_alpha + omega s, r = s - omega t, and the
// partials (r_hat, r) in entries 0-1 and |r|^2 in entry 2.
template <typename T>
__global__ __launch_bounds__(kRT) void k_ur_dot(long long n, cpx<T>* u, const cpx<T>* p, const cpx<T>* s, cpx<T>* r,
                                                const cpx<T>* t, const cpx<T>* rh, const BiState* st, double* part) {
  const int b = blockIdx.y;
  if (st[b].done) return;
  const long long off = (long long)b * n;
  u += off; p += off; s += off; r += off; t += off; rh += off;
  const bool h = st[b].halt != 0;
  const cpx<T> al = cast<T>(st[b].alpha), om = cast<T>(st[b].omega);
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const cpx<T> ua = u[i] + al * p[i];
    if (h) {
      u[i] = ua;
    } else {
      const cpx<T> si = s[i];
      u[i] = ua + om * si;
      const cpx<T> x = si - om * t[i];
      r[i] = x;
      const cpx<T> q = rh[i];
      acc[0] += (double)q.re * x.re + (double)q.im * x.im;
      acc[1] += (double)q.re * x.im - (double)q.im * x.re;
      acc[2] += (double)x.re * x.re + (double)x.im * x.im;
    }
  }
  bsum<3>(acc);
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)b * kRB + blockIdx.x) * 4;
    o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2];
  }
}

// ---- decide kernels: one block per RHS ----------------------------------------
struct Hist {
  double* h;
  long long* a;
  int cap;
};

// Where a decide kernel finds its partial sums: records of 4 doubles, `stride`
// records per RHS, the first np (or *count, when given) used, from entry off.
struct PartSrc {
  const double* part;
  long long stride;
  int np;
  const int* count;
  int off;
};

__host__ __device__ inline PartSrc vec_parts(const double* part, int np, int off = 0) {
  return PartSrc{part, kRB, np, nullptr, off};
}

template <int NV>
__device__ __forceinline__ void final_sum4(PartSrc ps, double* out) {
  const double* part = ps.part + (size_t)blockIdx.x * ps.stride * 4 + ps.off;
  const int np = ps.count ? *ps.count : ps.np;
  double v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0.0;
  // four records per step, all loaded before the first add: the loads are in
  // flight together, and the additions keep the record order
  const int bd = (int)blockDim.x;
  int k = threadIdx.x;
  for (; k + 3 * bd < np; k += 4 * bd) {
    double q[4][NV];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < NV; ++i) q[j][i] = part[(size_t)(k + j * bd) * 4 + i];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] += q[j][i];
  }
  for (; k < np; k += bd)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += part[(size_t)k * 4 + i];
  bsum<NV>(v);
#pragma unroll
  for (int i = 0; i < NV; ++i) out[i] = v[i];
}

// The coefficients the fused z passes read (KryCoef), written by thread 0 of
// a decide kernel when it leaves (kc null: the stand-alone vector kernels run).
struct Publish {
  const BiState* st;
  KryCoef* kc;
  __device__ ~Publish() {
    if (!kc) return;
    KryCoef c;
    c.beta[0] = st->beta.re; c.beta[1] = st->beta.im;
    c.omega[0] = st->omega.re; c.omega[1] = st->omega.im;
    c.alpha[0] = st->alpha.re; c.alpha[1] = st->alpha.im;
    c.first = st->it == 0 ? 1 : 0;
    c.done = st->done;
    *kc = c;
  }
};

__device__ __forceinline__ void record_b(BiState* st, Hist H, double v) {
  st->last_rel = v;
  if (st->nhist < st->hist_cap) {
    H.h[(size_t)blockIdx.x * H.cap + st->nhist] = v;
    H.a[(size_t)blockIdx.x * H.cap + st->nhist] = st->mv;
  }
  st->nhist += 1;
}

// rho = (r_hat, r) of the iteration about to start, and its beta
__device__ __forceinline__ void rho_step(BiState* st, double re, double im) {
  const cpx<double> rho = mk(re, im);
  st->rho = rho;
  if (cabs_d(rho) == 0.0) return finish(st, ST_BREAKDOWN_RHO);
  if (st->it > 0) {
    if (cabs_d(st->omega) == 0.0) return finish(st, ST_BREAKDOWN_OMEGA);
    st->beta = cdiv(rho, st->rho_prev) * cdiv(st->alpha, st->omega);
  }
  if (st->use_max_act && st->mv + 1 > st->max_act) return finish(st, ST_MAX_ACTIONS);
}

__global__ __launch_bounds__(kRT) void k_rho(PartSrc ps, BiState* sts, KryCoef* kcs) {
  BiState* st = sts + blockIdx.x;
  if (st->done) return;
  double v[2];
  final_sum4<2>(ps, v);
  if (threadIdx.x) return;
  Publish pub{st, kcs ? kcs + blockIdx.x : nullptr};
  rho_step(st, v[0], v[1]);
}

__global__ __launch_bounds__(kRT) void k_alpha(PartSrc ps, BiState* sts, KryCoef* kcs) {
  BiState* st = sts + blockIdx.x;
  if (st->done) return;
  double v[2];
  final_sum4<2>(ps, v);
  if (threadIdx.x) return;
  Publish pub{st, kcs ? kcs + blockIdx.x : nullptr};
  st->mv += 1;
  const cpx<double> den = mk(v[0], v[1]);
  if (cabs_d(den) == 0.0) return finish(st, ST_BREAKDOWN_ALPHA);
  st->alpha = cdiv(st->rho, den);
}

__global__ __launch_bounds__(kRT) void k_sfin(PartSrc ps, BiState* sts, Hist H) {
  BiState* st = sts + blockIdx.x;
  if (st->done) return;
  double v;
  final_sum4<1>(ps, &v);
  if (threadIdx.x) return;
  const double sr = sqrt(v) / st->rhs_norm;
  st->s_rel = sr;
  int stop = ST_RUNNING;
  if (!isfinite(sr)) stop = ST_NONFINITE;
  else if (st->use_divg && sr > st->divg) stop = ST_DIVERGENCE;
  else if (sr < st->tol) stop = ST_CONVERGED;
  else if (st->use_max_act && st->mv + 1 > st->max_act) stop = ST_MAX_ACTIONS;
  if (stop != ST_RUNNING) {
    record_b(st, H, sr);
    halt(st, stop);
  }
}

__global__ __launch_bounds__(kRT) void k_omega(PartSrc ps, BiState* sts, Hist H) {
  BiState* st = sts + blockIdx.x;
  if (st->skip_t) return;
  double v[4];
  final_sum4<4>(ps, v);
  if (threadIdx.x) return;
  st->mv += 1;
  const cpx<double> tt = mk(v[0], v[1]);
  if (cabs_d(tt) < 1e-30 * st->rhs_norm * st->rhs_norm) {
    record_b(st, H, st->s_rel);
    return halt(st, ST_BREAKDOWN_T);
  }
  st->omega = cdiv(mk(v[2], v[3]), tt);
  if (cabs_d(st->omega) == 0.0) {
    record_b(st, H, st->s_rel);
    return halt(st, ST_BREAKDOWN_OMEGA);
  }
}

// end of an iteration: |r|^2 of the updated residual, history, stop tests
__device__ __forceinline__ void rfin_step(BiState* st, Hist H, double rr) {
  st->it += 1;
  if (st->halt) {
    st->done = 1;
    return;
  }
  const double rel = sqrt(rr) / st->rhs_norm;
  record_b(st, H, rel);
  if (!isfinite(rel)) return finish(st, ST_NONFINITE);
  if (rel < st->tol) return finish(st, ST_CONVERGED);
  if (st->use_divg && rel > st->divg) return finish(st, ST_DIVERGENCE);
  st->rho_prev = st->rho;
  if (st->it >= st->max_iter) return finish(st, ST_MAX_ITER);
}

__global__ __launch_bounds__(kRT) void k_rfin(PartSrc ps, BiState* sts, Hist H, KryCoef* kcs) {
  BiState* st = sts + blockIdx.x;
  if (st->done) return;
  double v;
  final_sum4<1>(ps, &v);
  if (threadIdx.x) return;
  Publish pub{st, kcs ? kcs + blockIdx.x : nullptr};
  rfin_step(st, H, v);
}

// k_rfin, then (the solve going on) k_rho of the next iteration, in one
// launch: k_ur_dot's records hold (r_hat, r) and |r|^2 side by side
__global__ __launch_bounds__(kRT) void k_rfin_rho(PartSrc ps, BiState* sts, Hist H, KryCoef* kcs) {
  BiState* st = sts + blockIdx.x;
  if (st->done) return;
  double v[3];
  final_sum4<3>(ps, v);
  if (threadIdx.x) return;
  Publish pub{st, kcs ? kcs + blockIdx.x : nullptr};
  rfin_step(st, H, v[2]);
  if (!st->done) rho_step(st, v[0], v[1]);
}

// group flags of the batched matvecs: all_done = AND done, all_skip_t = AND skip_t
struct Group {
  int all_done;
  int all_skip_t;
};

__global__ __launch_bounds__(kRT) void k_group(const BiState* sts, int nb, Group* g) {
  int d = 1, s = 1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    d &= sts[b].done ? 1 : 0;
    s &= sts[b].skip_t ? 1 : 0;
  }
  __shared__ int sd[kRT], ss[kRT];
  sd[threadIdx.x] = d;
  ss[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)blockDim.x; ++k) { d &= sd[k]; s &= ss[k]; }
    g->all_done = d;
    g->all_skip_t = s;
  }
}

// first history entry of every RHS (relative residual of u = 0 and its action count)
// Start state of every RHS from the norm partials of f (r = f - A 0 = f, A 0
// counted as one action) and its first history entry.  One thread per RHS
// sums the nb partials in order, as a host loop would.
// one block per right-hand side: ||f_b||^2 from the norm kernel's partials
// (the same fixed-order block sum as the decide kernels), then the start state
__global__ __launch_bounds__(kRT) void k_init(PartSrc ps, em3d_bicgstab_params prm, BiState* sts, Hist H,
                                              KryCoef* kcs) {
  const int b = blockIdx.x;
  double ff[2];          // (f, f): |f|^2 and its imaginary part as summed
  final_sum4<2>(ps, ff);
  if (threadIdx.x) return;
  const double nn = ff[0];
  const double rn = sqrt(nn);
  BiState s = {};
  s.rhs_norm = rn;
  s.divg = prm.divergence_guard;
  s.use_divg = prm.divergence_guard >= 0.0 ? 1 : 0;
  s.max_act = prm.max_operator_actions;
  s.use_max_act = prm.max_operator_actions >= 0 ? 1 : 0;
  s.max_iter = prm.max_iter;
  s.rho_prev = mk(1.0, 0.0);
  s.alpha = mk(1.0, 0.0);
  s.omega = mk(1.0, 0.0);
  s.nhist = 1;
  s.hist_cap = H.cap;
  s.status = ST_RUNNING;
  double rel0;
  if (rn == 0.0) {
    rel0 = 0.0;
    s.mv = 0;
    s.status = ST_CONVERGED;
  } else {
    s.tol = fmax(prm.rtol, prm.atol / rn);
    rel0 = rn / rn;
    s.mv = 1;
    if (!isfinite(rel0)) s.status = ST_NONFINITE;
    else if (rel0 < s.tol) s.status = ST_CONVERGED;
    else if (prm.max_operator_actions >= 0 && s.mv >= prm.max_operator_actions) s.status = ST_MAX_ACTIONS;
    else if (prm.max_iter <= 0) s.status = ST_MAX_ITER;
  }
  s.last_rel = rel0;
  if (s.status != ST_RUNNING) {
    s.done = 1;
    s.skip_t = 1;
  } else if (kcs) {
    // the fused iterations start without k_rho: r_hat = r = f, so the first
    // rho is (f, f), the sum just taken
    rho_step(&s, ff[0], ff[1]);
  }
  sts[b] = s;
  H.h[(size_t)b * H.cap] = rel0;
  H.a[(size_t)b * H.cap] = s.mv;
  if (kcs) {
    Publish pub{&sts[b], kcs + b};   // publishes as it goes out of scope
  }
}

int sfail(const std::string& m) { return set_error(m); }

int chk(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return sfail(std::string("em3d_bicgstab: ") + what + ": " + hipGetErrorString(e));
}

long long field_elems(const em3d_plan_s* plan) {
  const Plan& p = plan->p;
  return (long long)p.C * p.N[0] * p.N[1] * p.N[2];
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Iterations to enqueue before the next state readback: for every running
// right-hand side, ceil(log(tol / rel) / log(rate)) with rate = rel^(1/it) its
// mean reduction per iteration (rel = 1 at it = 0); the largest of them,
// clamped to [1, chunk].  A right-hand side with no reduction yet asks for
// a full chunk.  The choice only moves the host checks: the iterations and
// their results are the same for any chunk length.
int next_chunk(const std::vector<BiState>& cur, int chunk) {
  int len = 1;
  for (const BiState& s : cur) {
    if (s.done) continue;
    const double rel = s.last_rel;
    if (!(s.it > 0 && rel > 0.0 && rel < 1.0 && s.tol > 0.0 && std::isfinite(rel))) return chunk;
    const double need = std::ceil(std::log(s.tol / rel) / (std::log(rel) / (double)s.it));
    if (!(need < (double)chunk)) return chunk;
    len = std::max(len, (int)need);
  }
  return len;
}

// solver scratch: r, r_hat, p, v, s, t (batch fields each) | partials | states | group | rel0 (unused)
// | fused-step coefficients | z-pass partials of P1 and P5 | their block counts
// | preconditioned p and s (preconditioned solves only)
struct Work {
  void* r;
  void* rh;
  void* p;
  void* v;
  void* s;
  void* t;
  double* part;
  BiState* st;
  Group* g;
  double* rel0;
  KryCoef* kc;
  double* zpart1;
  double* zpart5;
  int* zcount;
  long long zstride;
  void* ph;
  void* sh;
};

size_t part_bytes(int B) { return align256((size_t)B * kRB * 4 * sizeof(double)); }

// records per RHS of the z-pass partials: at most one block per pencil and component
long long zpart_stride(const Plan& p) { return (long long)p.C * p.N[0] * p.N[1]; }
size_t zpart_bytes(const Plan& p) { return align256((size_t)p.batch * zpart_stride(p) * 4 * sizeof(double)); }

size_t work_total(const Plan& p, size_t vb);

Work carve(const Plan& p, void* base, size_t vb) {
  const int B = p.batch;
  Work w;
  char* c = (char*)base;
  const size_t a = align256(vb);
  w.r = c; w.rh = c + a; w.p = c + 2 * a; w.v = c + 3 * a; w.s = c + 4 * a; w.t = c + 5 * a;
  w.ph = (char*)base + work_total(p, vb);
  w.sh = (char*)w.ph + a;
  c += 6 * a;
  w.part = (double*)c;
  c += part_bytes(B);
  w.st = (BiState*)c;
  c += align256((size_t)B * sizeof(BiState));
  w.g = (Group*)c;
  c += align256(sizeof(Group));
  w.rel0 = (double*)c;
  c += align256((size_t)B * sizeof(double));
  w.kc = (KryCoef*)c;
  c += align256((size_t)B * sizeof(KryCoef));
  w.zpart1 = (double*)c;
  c += zpart_bytes(p);
  w.zpart5 = (double*)c;
  c += zpart_bytes(p);
  w.zcount = (int*)c;
  w.zstride = zpart_stride(p);
  return w;
}

// The flags the passes of the matvecs test: for one right-hand side the
// flags of its own state (no group kernel is launched), for a batch the
// group's AND over the states (k_group after each decision that moves them).
struct Flags {
  const int* all_done;
  const int* all_skip_t;
  bool group;
};
Flags pass_flags(const Work& w, int B) {
  if (B == 1)
    return {(const int*)((const char*)w.st + offsetof(BiState, done)),
            (const int*)((const char*)w.st + offsetof(BiState, skip_t)), false};
  return {&w.g->all_done, &w.g->all_skip_t, true};
}

size_t work_total(const Plan& p, size_t vb) {
  const int B = p.batch;
  return 6 * align256(vb) + part_bytes(B) + align256((size_t)B * sizeof(BiState)) + align256(sizeof(Group)) +
         align256((size_t)B * sizeof(double)) + align256((size_t)B * sizeof(KryCoef)) + 2 * zpart_bytes(p) +
         align256(2 * sizeof(int));
}

// One BiCGStab iteration with the vector updates fused into the matvecs'
// z passes (ZKry): p = r + beta (p - omega v) in the first pass of v = A p and
// (r_hat, v) in its last; s = r - alpha v and |s|^2 in the first pass of
// t = A s, (t, t) and (t, s) in its last; then one field kernel for u and r
// with the next (r_hat, r) and |r|^2.  Same decisions as enqueue_iteration.
template <typename T>
void enqueue_iteration_fused(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, void* u, void* work,
                             const Work& w, long long n, int nb, int B, Hist H, hipStream_t stream, int* fail) {
  const dim3 gv(nb, B), gd(B);
  const Flags fl = pass_flags(w, B);
  const int* all_done = fl.all_done;
  const int* all_skip_t = fl.all_skip_t;
  const PartSrc z1{w.zpart1, w.zstride, 0, &w.zcount[0], 0};
  const PartSrc z5{w.zpart5, w.zstride, 0, &w.zcount[1], 0};
  ZKry kp{KRY_P, w.r, w.v, w.p, nullptr, w.kc, w.zpart1, w.zstride, &w.zcount[0]};
  ZKry kd{KRY_DOT, nullptr, nullptr, nullptr, w.rh, w.kc, w.zpart5, w.zstride, &w.zcount[1]};
  if (apply_plan(plan, khat, chi, chi_kind, w.p, w.v, work, stream, false, "em3d_bicgstab matvec", all_done, &kp,
                 &kd)) {
    *fail = 1;
    return;
  }
  hipLaunchKernelGGL(k_alpha, gd, dim3(kRT), 0, stream, z5, w.st, w.kc);
  ZKry ks{KRY_S, w.r, w.v, w.s, nullptr, w.kc, w.zpart1, w.zstride, &w.zcount[0]};
  ZKry kt{KRY_DOT2, nullptr, nullptr, nullptr, nullptr, w.kc, w.zpart5, w.zstride, &w.zcount[1]};
  if (apply_plan(plan, khat, chi, chi_kind, w.s, w.t, work, stream, false, "em3d_bicgstab matvec", all_skip_t, &ks,
                 &kt, all_done, PASSES_FIRST)) {
    *fail = 1;
    return;
  }
  hipLaunchKernelGGL(k_sfin, gd, dim3(kRT), 0, stream, z1, w.st, H);
  if (fl.group) hipLaunchKernelGGL(k_group, dim3(1), dim3(kRT), 0, stream, w.st, B, w.g);
  if (apply_plan(plan, khat, chi, chi_kind, w.s, w.t, work, stream, false, "em3d_bicgstab matvec", all_skip_t, &ks,
                 &kt, all_done, PASSES_REST)) {
    *fail = 1;
    return;
  }
  hipLaunchKernelGGL(k_omega, gd, dim3(kRT), 0, stream, z5, w.st, H);
  hipLaunchKernelGGL(k_ur_dot<T>, gv, dim3(kRT), 0, stream, n, (cpx<T>*)u, (const cpx<T>*)w.p, (const cpx<T>*)w.s,
                     (cpx<T>*)w.r, (const cpx<T>*)w.t, (const cpx<T>*)w.rh, w.st, w.part);
  hipLaunchKernelGGL(k_rfin_rho, gd, dim3(kRT), 0, stream, vec_parts(w.part, nb), w.st, H, w.kc);
  if (fl.group) hipLaunchKernelGGL(k_group, dim3(1), dim3(kRT), 0, stream, w.st, B, w.g);
}

template <typename T>
void enqueue_iteration(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, void* u, void* work,
                       const Work& w, long long n, int nb, int B, Hist H, const em3d_fno_model* pc,
                       hipStream_t stream, int* fail) {
  const dim3 gv(nb, B), gd(B);
  const Flags fl = pass_flags(w, B);
  const int* all_done = fl.all_done;
  const int* all_skip_t = fl.all_skip_t;
  const Plan& P = plan->p;
  void* pu = pc ? w.ph : w.p;   // search directions the matvecs and the update use
  void* su = pc ? w.sh : w.s;
  hipLaunchKernelGGL(k_dot<T>, gv, dim3(kRT), 0, stream, n, (const cpx<T>*)w.rh, (const cpx<T>*)w.r, w.part, w.st, 0);
  hipLaunchKernelGGL(k_rho, gd, dim3(kRT), 0, stream, vec_parts(w.part, nb), w.st, (KryCoef*)nullptr);
  hipLaunchKernelGGL(k_p<T>, gv, dim3(kRT), 0, stream, n, (cpx<T>*)w.p, (const cpx<T>*)w.r, (const cpx<T>*)w.v, w.st);
  if (fl.group) hipLaunchKernelGGL(k_group, dim3(1), dim3(kRT), 0, stream, w.st, B, w.g);
  if (pc && fno_precond_apply(pc, B, P.C, P.N[0], P.N[1], P.N[2], P.prec, w.p, w.ph, all_done, stream)) {
    *fail = 1;
    return;
  }
  if (apply_plan(plan, khat, chi, chi_kind, pu, w.v, work, stream, false, "em3d_bicgstab matvec", all_done)) {
    *fail = 1;
    return;
  }
  hipLaunchKernelGGL(k_dot<T>, gv, dim3(kRT), 0, stream, n, (const cpx<T>*)w.rh, (const cpx<T>*)w.v, w.part, w.st, 0);
  hipLaunchKernelGGL(k_alpha, gd, dim3(kRT), 0, stream, vec_parts(w.part, nb), w.st, (KryCoef*)nullptr);
  hipLaunchKernelGGL(k_s<T>, gv, dim3(kRT), 0, stream, n, (cpx<T>*)w.s, (const cpx<T>*)w.r, (const cpx<T>*)w.v, w.st,
                     w.part);
  hipLaunchKernelGGL(k_sfin, gd, dim3(kRT), 0, stream, vec_parts(w.part, nb), w.st, H);
  if (fl.group) hipLaunchKernelGGL(k_group, dim3(1), dim3(kRT), 0, stream, w.st, B, w.g);
  if (pc && fno_precond_apply(pc, B, P.C, P.N[0], P.N[1], P.N[2], P.prec, w.s, w.sh, all_skip_t, stream)) {
    *fail = 1;
    return;
  }
  if (apply_plan(plan, khat, chi, chi_kind, su, w.t, work, stream, false, "em3d_bicgstab matvec", all_skip_t)) {
    *fail = 1;
    return;
  }
  hipLaunchKernelGGL(k_dot2<T>, gv, dim3(kRT), 0, stream, n, (const cpx<T>*)w.t, (const cpx<T>*)w.s, w.part, w.st);
  hipLaunchKernelGGL(k_omega, gd, dim3(kRT), 0, stream, vec_parts(w.part, nb), w.st, H);
  hipLaunchKernelGGL(k_ur<T>, gv, dim3(kRT), 0, stream, n, (cpx<T>*)u, (const cpx<T>*)pu, (const cpx<T>*)su,
                     (const cpx<T>*)w.s, (cpx<T>*)w.r, (const cpx<T>*)w.t, w.st, w.part);
  hipLaunchKernelGGL(k_rfin, gd, dim3(kRT), 0, stream, vec_parts(w.part, nb), w.st, H, (KryCoef*)nullptr);
}

}  // namespace

extern "C" {

size_t em3d_bicgstab_work_bytes(em3d_plan_t plan) {
  if (!plan) return 0;
  const int B = plan->p.batch;
  return work_total(plan->p, (size_t)B * field_elems(plan) * elem_bytes(plan->p.prec));
}

size_t em3d_bicgstab_pc_work_bytes(em3d_plan_t plan) {
  if (!plan) return 0;
  const size_t vb = (size_t)plan->p.batch * field_elems(plan) * elem_bytes(plan->p.prec);
  return em3d_bicgstab_work_bytes(plan) + 2 * align256(vb);
}

}  // extern "C"

namespace {

int bicgstab_impl(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* rhs, void* u,
                  void* work, void* solver_work, const em3d_bicgstab_params* prm, double* history,
                  long long* action_counts, int history_cap, em3d_solver_result* res, const em3d_fno_model* pc,
                  hipStream_t stream) {
  if (!plan || !khat || !rhs || !u || !work || !solver_work || !prm || !res || !history || !action_counts)
    return sfail("em3d_bicgstab: null argument");
  if (history_cap < 1) return sfail("em3d_bicgstab: history_cap must be >= 1");
  const Plan& P = plan->p;
  const int prec = P.prec;
  const int B = P.batch;
  const long long n = field_elems(plan);
  const size_t es = elem_bytes(prec);
  const size_t vb = (size_t)B * n * es;
  Work w = carve(P, solver_work, vb);
  // fused vector steps: the matvec contrast must not mix components, and the
  // preconditioned form keeps its own directions
  const bool fused = prm->fused != 0 && !pc && chi_kind != CHI_FULL;
  const int nb = (int)std::min<long long>(kRB, (n + kRT - 1) / kRT);
  const Hist H{history, action_counts, history_cap};
  hipError_t e;
  for (int b = 0; b < B; ++b) res[b] = em3d_solver_result{};

  // ||f_b|| for every RHS (the states are not yet set: no skip)
  if (prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_dot<double>, dim3(nb, B), dim3(kRT), 0, stream, n, (const cpx<double>*)rhs,
                       (const cpx<double>*)rhs, w.part, (const BiState*)nullptr, 0);
  else
    hipLaunchKernelGGL(k_dot<float>, dim3(nb, B), dim3(kRT), 0, stream, n, (const cpx<float>*)rhs,
                       (const cpx<float>*)rhs, w.part, (const BiState*)nullptr, 0);
  if ((e = hipGetLastError()) != hipSuccess) return chk(e, "norm launch");
  // u = 0, r = r_hat = f, p = v = 0, start states and first history entries, all on the stream
  if ((e = hipMemsetAsync(u, 0, vb, stream)) != hipSuccess ||
      (e = hipMemcpyAsync(w.r, rhs, vb, hipMemcpyDeviceToDevice, stream)) != hipSuccess ||
      (e = hipMemcpyAsync(w.rh, rhs, vb, hipMemcpyDeviceToDevice, stream)) != hipSuccess ||
      (e = hipMemsetAsync(w.p, 0, vb, stream)) != hipSuccess || (e = hipMemsetAsync(w.v, 0, vb, stream)) != hipSuccess)
    return chk(e, "init");
  hipLaunchKernelGGL(k_init, dim3(B), dim3(kRT), 0, stream, vec_parts(w.part, nb), *prm, w.st, H,
                     fused ? w.kc : (KryCoef*)nullptr);
  if (fused && pass_flags(w, B).group) hipLaunchKernelGGL(k_group, dim3(1), dim3(kRT), 0, stream, w.st, B, w.g);
  if ((e = hipGetLastError()) != hipSuccess) return chk(e, "init launch");

  // chunks of iterations, the states read back after each (a solve that is
  // already finished at the start runs one chunk of no-op launches).  After
  // the first chunk, the next one is as long as the iterations the running
  // right-hand sides still need at their mean residual reduction per
  // iteration so far (at most `chunk`): iterations enqueued after a solve
  // has finished still launch every kernel, each of which exits at once.
  const int chunk = prm->chunk > 0 ? prm->chunk : 4;
  std::vector<BiState> cur((size_t)B);
  bool read = false;
  int len = chunk;
  for (int it0 = 0; it0 < prm->max_iter; it0 += len) {
    if (read) len = next_chunk(cur, chunk);
    for (int k = 0; k < len && it0 + k < prm->max_iter; ++k) {
      int fail = 0;
      if (fused && prec == PREC_DOUBLE)
        enqueue_iteration_fused<double>(plan, khat, chi, chi_kind, u, work, w, n, nb, B, H, stream, &fail);
      else if (fused)
        enqueue_iteration_fused<float>(plan, khat, chi, chi_kind, u, work, w, n, nb, B, H, stream, &fail);
      else if (prec == PREC_DOUBLE)
        enqueue_iteration<double>(plan, khat, chi, chi_kind, u, work, w, n, nb, B, H, pc, stream, &fail);
      else
        enqueue_iteration<float>(plan, khat, chi, chi_kind, u, work, w, n, nb, B, H, pc, stream, &fail);
      if (fail) return 1;
      if ((e = hipGetLastError()) != hipSuccess) return chk(e, "iteration launch");
    }
    if ((e = hipMemcpyAsync(cur.data(), w.st, (size_t)B * sizeof(BiState), hipMemcpyDeviceToHost, stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
      return chk(e, "state readback");
    read = true;
    bool all = true;
    for (int b = 0; b < B; ++b) all = all && cur[b].done;
    if (all) break;
  }
  if (!read) {
    if ((e = hipMemcpyAsync(cur.data(), w.st, (size_t)B * sizeof(BiState), hipMemcpyDeviceToHost, stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
      return chk(e, "state readback");
  }
  for (int b = 0; b < B; ++b) {
    const BiState& s = cur[b];
    res[b].status = s.done ? s.status : ST_MAX_ITER;   // k_rfin decides at max_iter; kept for safety
    res[b].iterations = s.it;
    res[b].matvec_count = (int)s.mv;
    res[b].history_len = s.nhist < history_cap ? s.nhist : history_cap;
    res[b].true_final_residual = s.last_rel;
  }
  return 0;
}

}  // namespace

extern "C" {

int em3d_bicgstab(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* rhs, void* u,
                  void* work, void* solver_work, const em3d_bicgstab_params* prm, double* history,
                  long long* action_counts, int history_cap, em3d_solver_result* res, hipStream_t stream) {
  return bicgstab_impl(plan, khat, chi, chi_kind, rhs, u, work, solver_work, prm, history, action_counts, history_cap,
                       res, nullptr, stream);
}

int em3d_bicgstab_pc(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* rhs, void* u,
                     void* work, void* solver_work, const em3d_bicgstab_params* prm, double* history,
                     long long* action_counts, int history_cap, em3d_solver_result* res,
                     const em3d_fno_model* precond, hipStream_t stream) {
  if (!precond) return sfail("em3d_bicgstab_pc: null preconditioner");
  const std::string why = fno_check(&precond->config);
  if (!why.empty()) return sfail("em3d_bicgstab_pc: " + why);
  const Plan& P = plan ? plan->p : Plan{};
  if (!plan || precond->factor < 1 || P.N[0] != precond->config.grid * precond->factor ||
      P.N[1] != precond->config.grid * precond->factor || P.N[2] != precond->config.grid * precond->factor)
    return sfail("em3d_bicgstab_pc: the plan grid must be (config.grid * factor)^3");
  if (precond->config.in_channels != P.C || precond->config.material_channels != precond->chi_channels ||
      precond->config.out_channels != P.C)
    return sfail("em3d_bicgstab_pc: in_channels and out_channels must be ncomp, material_channels chi_channels");
  return bicgstab_impl(plan, khat, chi, chi_kind, rhs, u, work, solver_work, prm, history, action_counts, history_cap,
                       res, precond, stream);
}

}  // extern "C"
