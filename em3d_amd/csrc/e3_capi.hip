// libem3d_hip: HIP kernels of the FFT volume-integral operator and the C-ABI.
//
// Entry points are declared (with the reference interface each one replaces)
// in include/em3d_hip.h.  Rules of the ABI: every function returns int status
// (0 = OK, else the message is in em3d_last_error()); no C++ exception crosses
// it; all data pointers are device pointers owned by the caller; the stream is
// the caller's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/em3d_hip.h"
#include "e3_error.hpp"
#include "e3_fast_launch.hpp"
#include "e3_plan.hpp"
#include "e3_pipeline.hpp"

using namespace e3;

namespace {

int fail(const std::string& msg) { return set_error(msg); }

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" __shared__ double2 e3_smem[];

// ---------------------------------------------------------------------------
// pass kernels: phases of e3_blocks.hpp separated by barriers
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_zpass(ZArgs a, long long nblk) {
  if (a.skip && *a.skip) return;
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(e3_smem);
  const long long bid = blockIdx.x;
  const int b = (int)(bid / nblk);
  ZPass<T> z(a, bid - (long long)b * nblk, b, lds);
  KAcc acc{};
  z.load(threadIdx.x, blockDim.x, acc);
  __syncthreads();
  for (int s = 0; s < a.plan.nst; ++s) {
    z.stage(s, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  z.store(threadIdx.x, blockDim.x, acc);
  if (a.k.op == KRY_S || a.k.op == KRY_DOT || a.k.op == KRY_DOT2) kry_block_store(a, acc);
}

template <typename T>
__global__ __launch_bounds__(256) void k_spass(SArgs a, long long ncb) {
  if (a.skip && *a.skip) return;
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(e3_smem);
  const long long bid = blockIdx.x;
  const long long q = bid / ncb;
  SPass<T> s(a, bid - q * ncb, q, lds);
  s.load(threadIdx.x, blockDim.x);
  __syncthreads();
  for (int i = 0; i < a.plan.nst; ++i) {
    s.stage(i, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  s.store(threadIdx.x, blockDim.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_xpass(XArgs a, long long ncb) {
  if (a.skip && *a.skip) return;
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(e3_smem);
  const long long bid = blockIdx.x;
  const int b = (int)(bid / ncb);
  XPass<T> x(a, bid - (long long)b * ncb, b, lds);
  x.load(threadIdx.x, blockDim.x);
  __syncthreads();
  for (int i = 0; i < a.plan.nst; ++i) {
    x.stage_fwd(i, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  x.multiply(threadIdx.x, blockDim.x);
  __syncthreads();
  for (int i = 0; i < a.plan.nst; ++i) {
    x.stage_inv(i, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  x.store(threadIdx.x, blockDim.x);
}

__global__ __launch_bounds__(256) void k_kspatial(KBArgs a, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) kernel_point(a, i);
}

template <typename T>
__global__ __launch_bounds__(256) void k_koct(KBArgs a, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) octant_point<T>(a, i);
}

// Lets a kernel take up to kLdsMax bytes of LDS, its static LDS included.  A
// refused attribute is not an error of the caller's later calls: the pending
// error is cleared here.
template <typename K>
void allow_big_lds(K kernel) {
  const void* f = reinterpret_cast<const void*>(kernel);
  hipFuncAttributes fa;
  size_t stat = 0;
  if (hipFuncGetAttributes(&fa, f) == hipSuccess) stat = fa.sharedSizeBytes;
  else (void)hipGetLastError();
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kLdsMax - stat)) != hipSuccess)
    (void)hipGetLastError();
}

struct HipLauncher {
  hipStream_t st;
  hipError_t err = hipSuccess;
  hipEvent_t* ev = nullptr;   // when set, ev[i] is recorded before launch i
  int launches = 0;
  void note(hipError_t e) {
    if (err == hipSuccess && e != hipSuccess) err = e;
  }
  void mark() {
    if (ev && launches < kMaxTimedLaunches) note(hipEventRecord(ev[launches], st));
    ++launches;
  }
  // the fast (power-of-two) kernels take the pass when they cover it
  bool fast_off = false;
  template <class F, class A>
  bool try_fast(F f32, F f64, const A& a, int prec) {
    if (fast_off) return false;
    hipError_t e = hipSuccess;
    const bool done = prec == PREC_DOUBLE ? f64(a, st, &e) : f32(a, st, &e);
    if (done) note(e);
    return done;
  }
  void z(const ZArgs& a, long long nblk, int batch, int prec, size_t lds) {
    mark();
    if (try_fast(&launch_fast_z_f32, &launch_fast_z_f64, a, prec)) return;
    const dim3 g((unsigned)(nblk * batch));
    if (prec == PREC_DOUBLE) hipLaunchKernelGGL(k_zpass<double>, g, dim3(kThreads), lds, st, a, nblk);
    else hipLaunchKernelGGL(k_zpass<float>, g, dim3(kThreads), lds, st, a, nblk);
    note(hipGetLastError());
  }
  void s(const SArgs& a, long long ncb, long long Q, int prec, size_t lds) {
    mark();
    if (try_fast(&launch_fast_s_f32, &launch_fast_s_f64, a, prec)) return;
    const dim3 g((unsigned)(ncb * Q));
    if (prec == PREC_DOUBLE) hipLaunchKernelGGL(k_spass<double>, g, dim3(kThreads), lds, st, a, ncb);
    else hipLaunchKernelGGL(k_spass<float>, g, dim3(kThreads), lds, st, a, ncb);
    note(hipGetLastError());
  }
  void x(const XArgs& a, long long ncb, int batch, int prec, size_t lds) {
    mark();
    if (try_fast(&launch_fast_x_f32, &launch_fast_x_f64, a, prec)) return;
    const dim3 g((unsigned)(ncb * batch));
    if (prec == PREC_DOUBLE) hipLaunchKernelGGL(k_xpass<double>, g, dim3(kThreads), lds, st, a, ncb);
    else hipLaunchKernelGGL(k_xpass<float>, g, dim3(kThreads), lds, st, a, ncb);
    note(hipGetLastError());
  }
  void kspatial(const KBArgs& a, long long n) {
    mark();
    hipLaunchKernelGGL(k_kspatial, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
    note(hipGetLastError());
  }
  void koct(const KBArgs& a, long long n, int prec) {
    mark();
    const dim3 g((unsigned)((n + 255) / 256));
    if (prec == PREC_DOUBLE) hipLaunchKernelGGL(k_koct<double>, g, dim3(256), 0, st, a, n);
    else hipLaunchKernelGGL(k_koct<float>, g, dim3(256), 0, st, a, n);
    note(hipGetLastError());
  }
};

void set_lds_limits_once() {
  static bool done = false;
  if (done) return;
  done = true;
  allow_big_lds(k_zpass<float>);
  allow_big_lds(k_zpass<double>);
  allow_big_lds(k_spass<float>);
  allow_big_lds(k_spass<double>);
  allow_big_lds(k_xpass<float>);
  allow_big_lds(k_xpass<double>);
}

}  // namespace

int e3::apply_plan(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                   void* work, hipStream_t stream, bool adjoint, const char* name, const int* skip,
                   const ZKry* k1, const ZKry* k5, const int* skip1, unsigned passes, const int* win) {
  if (!plan || !khat || !u || !y || !work) return fail(std::string(name) + ": null argument");
  if (chi_kind < CHI_NONE || chi_kind > CHI_FULL) return fail(std::string(name) + ": bad chi_kind");
  if (chi_kind != CHI_NONE && !chi) return fail(std::string(name) + ": null chi");
  if (plan->p.C == 1 && chi_kind > CHI_ISO) return fail(std::string(name) + ": acoustic contrast must be scalar");
  if (u == y) return fail(std::string(name) + ": output must not alias the input");
  HipLauncher go{stream};
  if (plan->timing) go.ev = plan->ev;
  if ((k1 && k1->op != KRY_NONE && (chi_kind == CHI_FULL || adjoint)) || (k5 && k5->op != KRY_NONE && adjoint))
    return fail(std::string(name) + ": fused Krylov steps need a matvec with a non-mixing contrast");
  if (win) {
    if (k1 && k1->op != KRY_NONE) return fail(std::string(name) + ": an input window takes no fused Krylov step");
    for (int a = 0; a < 3; ++a)
      if (win[a] < 0 || win[3 + a] < 1 || win[a] + win[3 + a] > plan->p.N[a])
        return fail(std::string(name) + ": the window must be a non-empty box inside the grid");
  }
  run_apply(go, plan->p, khat, chi, chi_kind, u, y, work, adjoint, skip, k1, k5, skip1, passes, win);
  if (plan->timing && go.launches <= kMaxTimedLaunches) {
    go.note(hipEventRecord(plan->ev[go.launches], stream));
    go.note(hipEventSynchronize(plan->ev[go.launches]));
    for (int i = 0; i < go.launches; ++i) {
      float t = 0.f;
      go.note(hipEventElapsedTime(&t, plan->ev[i], plan->ev[i + 1]));
      plan->ms[i] += t;
    }
    plan->nlaunch = go.launches;
    plan->calls += 1;
  }
  if (go.err != hipSuccess) return hip_check(go.err, name);
  return 0;
}

extern "C" {

const char* em3d_version(void) { return "em3d_hip 0.1.0 (gfx950)"; }

const char* em3d_last_error(void) { return last_error_store().c_str(); }

int em3d_plan_create(int nx, int ny, int nz, int ncomp, int batch, int precision, em3d_plan_t* out) {
  if (!out) return fail("em3d_plan_create: null output handle");
  *out = nullptr;
  em3d_plan_s* h = new (std::nothrow) em3d_plan_s();
  if (!h) return fail("em3d_plan_create: out of host memory");
  const std::string why = plan_init(&h->p, nx, ny, nz, ncomp, batch, precision);
  if (!why.empty()) {
    delete h;
    return fail("em3d_plan_create: " + why);
  }
  small_grid_variants(h->p.N, h->p.L, h->p.C, h->p.batch, h->p.prec, h->p.variant);
  // twiddle tables exp(-2 pi i t / L) for the three axes, float then double
  size_t nf = 0;
  for (int a = 0; a < 3; ++a) nf += (size_t)h->p.L[a];
  std::vector<float> tf(2 * nf);
  std::vector<double> td(2 * nf);
  size_t off = 0;
  for (int a = 0; a < 3; ++a) {
    const int L = h->p.L[a];
    for (int t = 0; t < L; ++t) {
      const double ang = -2.0 * M_PI * (double)t / (double)L;
      td[2 * (off + t)] = std::cos(ang);
      td[2 * (off + t) + 1] = std::sin(ang);
      tf[2 * (off + t)] = (float)std::cos(ang);
      tf[2 * (off + t) + 1] = (float)std::sin(ang);
    }
    off += (size_t)L;
  }
  const size_t bytes_f = tf.size() * sizeof(float), bytes_d = td.size() * sizeof(double);
  if (hip_check(hipGetDevice(&h->device), "hipGetDevice") ||
      hip_check(hipMalloc(&h->tw_dev, bytes_d + bytes_f), "hipMalloc(twiddles)")) {
    delete h;
    return 1;
  }
  char* base = (char*)h->tw_dev;
  if (hip_check(hipMemcpy(base, td.data(), bytes_d, hipMemcpyHostToDevice), "hipMemcpy(twiddles)") ||
      hip_check(hipMemcpy(base + bytes_d, tf.data(), bytes_f, hipMemcpyHostToDevice), "hipMemcpy(twiddles)")) {
    (void)hipFree(h->tw_dev);
    delete h;
    return 1;
  }
  off = 0;
  for (int a = 0; a < 3; ++a) {
    h->p.tw_d[a] = base + off * 16;
    h->p.tw_f[a] = base + bytes_d + off * 8;
    off += (size_t)h->p.L[a];
  }
  set_lds_limits_once();
  *out = h;
  return 0;
}

int em3d_plan_set_timing(em3d_plan_t plan, int enable) {
  if (!plan) return fail("em3d_plan_set_timing: null plan");
  if (enable && !plan->timing) {
    for (int i = 0; i <= kMaxTimedLaunches; ++i)
      if (!plan->ev[i] && hip_check(hipEventCreate(&plan->ev[i]), "hipEventCreate")) return 1;
  }
  plan->timing = enable != 0;
  for (int i = 0; i < kMaxTimedLaunches; ++i) plan->ms[i] = 0.0;
  plan->calls = 0;
  plan->nlaunch = 0;
  return 0;
}

int em3d_plan_timing(em3d_plan_t plan, double* ms_per_launch, int max_launches, int* nlaunch, long long* calls) {
  if (!plan || !ms_per_launch || !nlaunch || !calls) return fail("em3d_plan_timing: null argument");
  *nlaunch = plan->nlaunch;
  *calls = plan->calls;
  for (int i = 0; i < plan->nlaunch && i < max_launches; ++i)
    ms_per_launch[i] = plan->calls ? plan->ms[i] / (double)plan->calls : 0.0;
  return 0;
}

int em3d_plan_variant_count(em3d_plan_t plan, int kind) {
  if (!plan || kind < 0 || kind > 2) return 0;
  const int axis = kind == PASS_S ? 1 : (kind == PASS_X ? 0 : 2);
  return fast_variant_count(kind, plan->p.L[axis], plan->p.prec);
}

int em3d_plan_set_variants(em3d_plan_t plan, const int* variants) {
  if (!plan || !variants) return fail("em3d_plan_set_variants: null argument");
  for (int k = 0; k < 3; ++k)
    if (variants[k] < -1 || variants[k] >= em3d_plan_variant_count(plan, k))
      return fail("em3d_plan_set_variants: variant " + std::to_string(variants[k]) + " out of range for pass kind " +
                  std::to_string(k));
  for (int k = 0; k < 3; ++k) plan->p.variant[k] = variants[k];
  return 0;
}

int em3d_plan_get_variants(em3d_plan_t plan, int* variants) {
  if (!plan || !variants) return fail("em3d_plan_get_variants: null argument");
  for (int k = 0; k < 3; ++k) variants[k] = plan->p.variant[k];
  return 0;
}

int em3d_plan_autotune(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                       void* work, int reps, double* best_ms, hipStream_t stream) {
  if (!plan || !khat || !u || !y || !work) return fail("em3d_plan_autotune: null argument");
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  if (hip_check(hipEventCreate(&e0), "hipEventCreate") || hip_check(hipEventCreate(&e1), "hipEventCreate")) return 1;
  const bool timing = plan->timing;
  plan->timing = false;
  int rc = 0;
  double best_all = 0.0;
  auto measure = [&](double* ms) -> int {
    if (apply_plan(plan, khat, chi, chi_kind, u, y, work, stream, false, "em3d_plan_autotune", nullptr)) return 1;
    if (hip_check(hipEventRecord(e0, stream), "hipEventRecord")) return 1;
    for (int r = 0; r < reps; ++r)
      if (apply_plan(plan, khat, chi, chi_kind, u, y, work, stream, false, "em3d_plan_autotune", nullptr)) return 1;
    if (hip_check(hipEventRecord(e1, stream), "hipEventRecord") || hip_check(hipEventSynchronize(e1), "sync"))
      return 1;
    float t = 0.f;
    if (hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime")) return 1;
    *ms = (double)t / reps;
    return 0;
  };
  // coordinate search: the x pass first (it dominates), then the strided and z
  // passes.  The plan's current variant is timed first and kept unless another
  // is faster by more than 1 %, so run-to-run noise does not flip the choice.
  const int order[3] = {PASS_X, PASS_S, PASS_Z};
  for (int oi = 0; oi < 3 && !rc; ++oi) {
    const int k = order[oi];
    const int n = em3d_plan_variant_count(plan, k);
    const int incumbent = plan->p.variant[k];
    int best_v = incumbent;
    double best = -1.0;
    for (int i = -1; i <= n && !rc; ++i) {
      const int v = i < 0 ? incumbent : i - 1;     // incumbent, then -1 .. n-1
      if (i >= 0 && v == incumbent) continue;
      plan->p.variant[k] = v;
      double ms = 0.0;
      rc = measure(&ms);
      if (!rc && (best < 0.0 || ms < 0.99 * best)) {
        best = ms;
        best_v = v;
      }
    }
    plan->p.variant[k] = best_v;
    best_all = best;
  }
  plan->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc && best_ms) *best_ms = best_all;
  return rc;
}

// ---- packed groups of operators replayed from one captured hipGraph ----------
}  // extern "C"

struct em3d_group_s {
  struct Member {
    em3d_plan_t plan;
    const void* khat;
    const void* chi;
    int chi_kind;
    const void* u;
    void* y;
    void* work;
  };
  std::vector<Member> m;
  int device = 0;
  hipStream_t cap = nullptr;       // private capture stream (default priority)
  std::vector<hipStream_t> branch; // one stream per member: parallel branches of the graph
  std::vector<hipEvent_t> ev;      // fork event + one join event per member
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int em3d_group_destroy(em3d_group_t g);

int em3d_group_create(int n, const em3d_plan_t* plans, const void* const* khats, const void* const* chis,
                      const int* chi_kinds, const void* const* inputs, void* const* outputs, void* const* works,
                      em3d_group_t* group) {
  if (n < 1 || !plans || !khats || !chis || !chi_kinds || !inputs || !outputs || !works || !group)
    return fail("em3d_group_create: null argument or empty group");
  auto* g = new em3d_group_s();
  for (int i = 0; i < n; ++i) {
    if (!plans[i] || !khats[i] || !inputs[i] || !outputs[i] || !works[i]) {
      delete g;
      return fail("em3d_group_create: null member " + std::to_string(i));
    }
    g->m.push_back({plans[i], khats[i], chis[i], chi_kinds[i], inputs[i], outputs[i], works[i]});
  }
  int rc = hip_check(hipGetDevice(&g->device), "hipGetDevice");
  if (!rc) rc = hip_check(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  g->branch.assign(n, nullptr);
  g->ev.assign(n + 1, nullptr);
  for (int i = 0; i < n && !rc; ++i)
    rc = hip_check(hipStreamCreateWithFlags(&g->branch[i], hipStreamNonBlocking), "hipStreamCreateWithFlags");
  for (int i = 0; i <= n && !rc; ++i)
    rc = hip_check(hipEventCreateWithFlags(&g->ev[i], hipEventDisableTiming), "hipEventCreateWithFlags");
  if (rc) {
    em3d_group_destroy(g);
    return 1;
  }
  *group = g;
  return 0;
}

// y_i = A_i u_i for every member.  The first call captures the members'
// pass sequences (no memsets, no events) into one graph; every call then
// launches that graph on `stream`.  Inputs and outputs are the buffers given
// at creation.
int em3d_group_matvec(em3d_group_t g, hipStream_t stream) {
  if (!g) return fail("em3d_group_matvec: null group");
  if (!g->exec) {
    std::vector<bool> timing(g->m.size());
    for (size_t i = 0; i < g->m.size(); ++i) {
      timing[i] = g->m[i].plan->timing;
      g->m[i].plan->timing = false;
    }
    // fork: every member on its own branch stream, joined back into the capture stream
    int rc = hip_check(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!rc) rc = hip_check(hipEventRecord(g->ev[0], g->cap), "hipEventRecord(fork)");
    for (size_t i = 0; i < g->m.size() && !rc; ++i) {
      const auto& q = g->m[i];
      rc = hip_check(hipStreamWaitEvent(g->branch[i], g->ev[0], 0), "hipStreamWaitEvent(fork)");
      if (!rc)
        rc = apply_plan(q.plan, q.khat, q.chi, q.chi_kind, q.u, q.y, q.work, g->branch[i], false,
                        "em3d_group_matvec", nullptr);
      if (!rc) rc = hip_check(hipEventRecord(g->ev[i + 1], g->branch[i]), "hipEventRecord(join)");
      if (!rc) rc = hip_check(hipStreamWaitEvent(g->cap, g->ev[i + 1], 0), "hipStreamWaitEvent(join)");
    }
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(g->cap, &graph);
    for (size_t i = 0; i < g->m.size(); ++i) g->m[i].plan->timing = timing[i];
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return 1;
    }
    if (hip_check(e, "hipStreamEndCapture")) return 1;
    g->graph = graph;
    if (hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate")) return 1;
  }
  return hip_check(hipGraphLaunch(g->exec, stream), "hipGraphLaunch");
}

int em3d_group_destroy(em3d_group_t g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (hipEvent_t e : g->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t b : g->branch)
    if (b) (void)hipStreamDestroy(b);
  if (g->cap) (void)hipStreamDestroy(g->cap);
  delete g;
  return 0;
}

int em3d_plan_destroy(em3d_plan_t plan) {
  if (!plan) return 0;
  int rc = 0;
  for (int i = 0; i <= kMaxTimedLaunches; ++i)
    if (plan->ev[i]) (void)hipEventDestroy(plan->ev[i]);
  if (plan->tw_dev) rc = hip_check(hipFree(plan->tw_dev), "hipFree(twiddles)");
  delete plan;
  return rc;
}

int em3d_plan_sizes(em3d_plan_t plan, size_t* khat_bytes, size_t* work_bytes, size_t* build_bytes) {
  if (!plan) return fail("em3d_plan_sizes: null plan");
  const size_t es = elem_bytes(plan->p.prec);
  if (khat_bytes) *khat_bytes = (size_t)plan->p.khat_elems * es;
  if (work_bytes) *work_bytes = (size_t)(plan->p.w1_elems + plan->p.w2_elems) * es;
  if (build_bytes) *build_bytes = (size_t)plan->p.build_elems * 16;
  return 0;
}

int em3d_kernel_build(em3d_plan_t plan, double k0, double hx, double hy, double hz, void* khat,
                      void* build_work, hipStream_t stream) {
  if (!plan || !khat || !build_work) return fail("em3d_kernel_build: null argument");
  if (!(k0 > 0.0)) return fail("em3d_kernel_build: k0 must be positive");
  if (!(hx > 0.0 && hy > 0.0 && hz > 0.0)) return fail("em3d_kernel_build: cell sizes must be positive");
  const double h[3] = {hx, hy, hz};
  double sre = 0.0, sim = 0.0;
  if (plan->p.C == 1) {
    // equivalent-sphere self cell exp(ika)(1 - ika) - 1, a = (3 dv / 4 pi)^(1/3)
    const double dv = hx * hy * hz;
    const double a = std::cbrt(3.0 * dv / (4.0 * M_PI));
    const double ka = k0 * a;
    const double cr = std::cos(ka), ci = std::sin(ka);
    sre = cr + ci * ka - 1.0;
    sim = ci - cr * ka;
  }
  HipLauncher go{stream};
  run_kernel_build(go, plan->p, k0, h, sre, sim, khat, build_work);
  if (go.err != hipSuccess) return hip_check(go.err, "em3d_kernel_build launch");
  return 0;
}

int em3d_matvec(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                void* work, hipStream_t stream) {
  return e3::apply_plan(plan, khat, chi, chi_kind, u, y, work, stream, false, "em3d_matvec", nullptr);
}

int em3d_rmatvec(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* v, void* y,
                 void* work, hipStream_t stream) {
  return e3::apply_plan(plan, khat, chi, chi_kind, v, y, work, stream, true, "em3d_rmatvec", nullptr);
}

int em3d_matvec_window(em3d_plan_t plan, const void* khat, const void* chi, int chi_kind, const void* u, void* y,
                       void* work, const int* lo, const int* n, hipStream_t stream) {
  if (!lo || !n) return fail("em3d_matvec_window: null window");
  const int win[6] = {lo[0], lo[1], lo[2], n[0], n[1], n[2]};
  return e3::apply_plan(plan, khat, chi, chi_kind, u, y, work, stream, false, "em3d_matvec_window", nullptr,
                        nullptr, nullptr, nullptr, e3::PASSES_ALL, win);
}

}  // extern "C"
