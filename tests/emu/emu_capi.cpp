// Host emulator of the HIP pass kernels (TEST INFRASTRUCTURE).
//
// Compiles the product's block-phase headers with g++ and runs every block of
// every pass thread by thread, phase by phase, on host memory.  It exercises
// the exact index arithmetic, tiling, pruning and spectral folding of the GPU
// kernels on the CPU; the LDS tile is poisoned with NaN before each block so a
// read of an unwritten LDS element shows up in the results.
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "../../em3d_amd/csrc/e3_fast.hpp"
#include "../../em3d_amd/csrc/e3_pipeline.hpp"

using namespace e3;

namespace {

std::string g_err;

template <typename T>
void poison(std::vector<unsigned char>& mem) {
  cpx<T>* p = (cpx<T>*)mem.data();
  const size_t n = mem.size() / sizeof(cpx<T>);
  const T nan = std::numeric_limits<T>::quiet_NaN();
  for (size_t i = 0; i < n; ++i) p[i] = mk(nan, nan);
}

// Fused Krylov partials of one block: the threads' sums in order, as the
// kernels' block reduction writes them (record `blk` of right-hand side b).
void kry_store(const ZArgs& a, const std::vector<KAcc>& acc, int b, long long blk, long long per) {
  double* o = a.k.part + ((size_t)b * a.k.part_stride + blk) * 4;
  for (int i = 0; i < 4; ++i) {
    double s = 0.0;
    for (const KAcc& x : acc) s += x.v[i];
    o[i] = s;
  }
  if (b == 0 && blk == 0) *a.k.count = (int)per;
}

// Replays the fast (power-of-two) kernels: per-thread register state lives
// in an array indexed by thread id; barriers separate the phase loops.
long long g_fast_calls[4] = {0, 0, 0, 0};   // s, x, zf, zi

struct EmuFastRun {
  template <class K>
  static std::vector<unsigned char> tile() {
    std::vector<unsigned char> m((size_t)K::LDS * sizeof(cpx<typename K::Real>));
    poison<typename K::Real>(m);
    return m;
  }
  template <class K>
  void s(const SArgs& a) {
    using T = typename K::Real;
    ++g_fast_calls[0];
    const long long ncb = (a.W + K::TC - 1) / K::TC;
    for (long long bid = 0; bid < ncb * a.Q; ++bid) {
      auto mem = tile<K>();
      const long long q = bid / ncb;
      K k(a, bid - q * ncb, q, (cpx<T>*)mem.data());
      std::vector<typename K::St> st(K::NT);
      for (int t = 0; t < K::NT; ++t) k.stage_a(t, st[t]);
      for (int t = 0; t < K::NT; ++t) k.stage_b(t, st[t]);
    }
  }
  template <class K>
  void x(const XArgs& a) {
    using T = typename K::Real;
    ++g_fast_calls[1];
    const long long ncb = K::blocks(a);
    for (long long bid = 0; bid < ncb * a.batch; ++bid) {
      auto mem = tile<K>();
      const int b = (int)(bid / ncb);
      K k(a, bid - (long long)b * ncb, b, (cpx<T>*)mem.data());
      std::vector<typename K::St> st(K::NT);
      const int h0 = k.row_begin(), h1 = k.row_end();
      if (h0 >= h1) continue;
      k.h_end = h1;
      if constexpr (K::PREFETCH)
        for (int t = 0; t < K::NT; ++t) k.load_a(0, t, h0, st[t]);
      for (int h = h0; h < h1; ++h) phases(k, h, st, std::make_integer_sequence<int, K::NPH>{});
    }
  }
  template <class K, int... I>
  static void phases(K& k, int h, std::vector<typename K::St>& st, std::integer_sequence<int, I...>) {
    auto one = [&](auto ic) {
      for (int t = 0; t < K::NT; ++t) k.template phase<decltype(ic)::value>(t, h, st[t]);
    };
    (one(std::integral_constant<int, I>{}), ...);
  }
  // SPLIT: one block per (pencil group, component), as the GPU launch
  template <class K, bool SPLIT = false>
  void zf(const ZArgs& a) {
    using T = typename K::Real;
    ++g_fast_calls[2];
    const long long nblk = (a.P + K::P - 1) / K::P;
    const bool kry = a.k.op == KRY_P || a.k.op == KRY_S;
    const int per = SPLIT ? a.C : 1;
    for (long long bid = 0; bid < nblk * a.batch * per; ++bid) {
      auto mem = tile<K>();
      const long long pg = bid / per;
      const int b = (int)(pg / nblk);
      K k(a, pg - (long long)b * nblk, b, (cpx<T>*)mem.data());
      std::vector<KAcc> acc(K::NT, KAcc{});
      const int c0 = SPLIT ? (int)(bid % per) : 0, c1 = SPLIT ? c0 + 1 : a.C;
      for (int c = c0; c < c1; ++c) {
        for (int t = 0; t < K::NT; ++t) {
          if (kry) k.stage_a(c, t, acc[t]);
          else k.stage_a(c, t);
        }
        for (int t = 0; t < K::NT; ++t) k.stage_b(c, t);
      }
      if (a.k.op == KRY_S) kry_store(a, acc, b, bid - (long long)b * nblk * per, nblk * per);
    }
  }
  template <class K, bool SPLIT_ = false>
  void zi(const ZArgs& a) {
    using T = typename K::Real;
    ++g_fast_calls[3];
    constexpr bool SPLIT = SPLIT_ && !K::HOLD;
    if constexpr (SPLIT) {
      const long long nblk = (a.P + K::P - 1) / K::P;
      const int C = K::NCOMP;
      for (long long bid = 0; bid < nblk * a.batch * C; ++bid) {
        auto mem = tile<K>();
        const long long pg = bid / C;
        const int c = (int)(bid % C);
        const int b = (int)(pg / nblk);
        K k(a, pg - (long long)b * nblk, b, (cpx<T>*)mem.data());
        k.c_first = c;
        std::vector<typename K::St> st(K::NT);
        std::vector<KAcc> acc(K::NT, KAcc{});
        for (int t = 0; t < K::NT; ++t) k.stage_a(c, t);
        for (int t = 0; t < K::NT; ++t) k.stage_b(c, t, st[t]);
        for (int t = 0; t < K::NT; ++t) {
          if (a.k.op == KRY_DOT || a.k.op == KRY_DOT2) k.epilogue_one_kry(c, t, st[t], acc[t]);
          else k.epilogue_one(c, t, st[t]);
        }
        if (a.k.op == KRY_DOT || a.k.op == KRY_DOT2)
          kry_store(a, acc, b, bid - (long long)b * nblk * C, nblk * C);
      }
      return;
    }
    const long long nblk = (a.P + K::P - 1) / K::P;
    for (long long bid = 0; bid < nblk * a.batch; ++bid) {
      auto mem = tile<K>();
      const int b = (int)(bid / nblk);
      K k(a, bid - (long long)b * nblk, b, (cpx<T>*)mem.data());
      std::vector<typename K::St> st(K::NT);
      std::vector<KAcc> acc(K::NT, KAcc{});
      for (int c = 0; c < a.C; ++c) {
        for (int t = 0; t < K::NT; ++t) k.stage_a(c, t);
        for (int t = 0; t < K::NT; ++t) k.stage_b(c, t, st[t]);
        if constexpr (!K::HOLD) {
          for (int t = 0; t < K::NT; ++t) {
            if (a.k.op == KRY_DOT || a.k.op == KRY_DOT2) k.epilogue_one_kry(c, t, st[t], acc[t]);
            else k.epilogue_one(c, t, st[t]);
          }
        }
      }
      if constexpr (K::HOLD)
        for (int t = 0; t < K::NT; ++t) k.epilogue(t, st[t]);
      if (a.k.op == KRY_DOT || a.k.op == KRY_DOT2) kry_store(a, acc, b, bid - (long long)b * nblk, nblk);
    }
  }
};

int g_use_fast = 1;

struct EmuLauncher {
  int nt = kThreads;
  EmuFastRun fr;
  template <typename T>
  void z_t(const ZArgs& a, long long nblk, int batch, size_t lds) {
    std::vector<unsigned char> mem(lds);
    for (long long bid = 0; bid < nblk * batch; ++bid) {
      poison<T>(mem);
      const int b = (int)(bid / nblk);
      ZPass<T> z(a, bid - (long long)b * nblk, b, (cpx<T>*)mem.data());
      std::vector<KAcc> acc(nt, KAcc{});
      for (int t = 0; t < nt; ++t) z.load(t, nt, acc[t]);
      for (int s = 0; s < a.plan.nst; ++s)
        for (int t = 0; t < nt; ++t) z.stage(s, t, nt);
      for (int t = 0; t < nt; ++t) z.store(t, nt, acc[t]);
      if (a.k.op == KRY_S || a.k.op == KRY_DOT || a.k.op == KRY_DOT2)
        kry_store(a, acc, b, bid - (long long)b * nblk, nblk);
    }
  }
  void z(const ZArgs& a, long long nblk, int batch, int prec, size_t lds) {
    if (g_use_fast && (prec == PREC_DOUBLE ? fast_z<double>(fr, a) : fast_z<float>(fr, a))) return;
    if (prec == PREC_DOUBLE) z_t<double>(a, nblk, batch, lds);
    else z_t<float>(a, nblk, batch, lds);
  }
  template <typename T>
  void s_t(const SArgs& a, long long ncb, long long Q, size_t lds) {
    std::vector<unsigned char> mem(lds);
    for (long long bid = 0; bid < ncb * Q; ++bid) {
      poison<T>(mem);
      const long long q = bid / ncb;
      SPass<T> s(a, bid - q * ncb, q, (cpx<T>*)mem.data());
      for (int t = 0; t < nt; ++t) s.load(t, nt);
      for (int i = 0; i < a.plan.nst; ++i)
        for (int t = 0; t < nt; ++t) s.stage(i, t, nt);
      for (int t = 0; t < nt; ++t) s.store(t, nt);
    }
  }
  void s(const SArgs& a, long long ncb, long long Q, int prec, size_t lds) {
    if (g_use_fast && (prec == PREC_DOUBLE ? fast_s<double>(fr, a) : fast_s<float>(fr, a))) return;
    if (prec == PREC_DOUBLE) s_t<double>(a, ncb, Q, lds);
    else s_t<float>(a, ncb, Q, lds);
  }
  template <typename T>
  void x_t(const XArgs& a, long long ncb, int batch, size_t lds) {
    std::vector<unsigned char> mem(lds);
    for (long long bid = 0; bid < ncb * batch; ++bid) {
      poison<T>(mem);
      const int b = (int)(bid / ncb);
      XPass<T> x(a, bid - (long long)b * ncb, b, (cpx<T>*)mem.data());
      for (int t = 0; t < nt; ++t) x.load(t, nt);
      for (int i = 0; i < a.plan.nst; ++i)
        for (int t = 0; t < nt; ++t) x.stage_fwd(i, t, nt);
      for (int t = 0; t < nt; ++t) x.multiply(t, nt);
      for (int i = 0; i < a.plan.nst; ++i)
        for (int t = 0; t < nt; ++t) x.stage_inv(i, t, nt);
      for (int t = 0; t < nt; ++t) x.store(t, nt);
    }
  }
  void x(const XArgs& a, long long ncb, int batch, int prec, size_t lds) {
    if (g_use_fast && (prec == PREC_DOUBLE ? fast_x<double>(fr, a) : fast_x<float>(fr, a))) return;
    if (prec == PREC_DOUBLE) x_t<double>(a, ncb, batch, lds);
    else x_t<float>(a, ncb, batch, lds);
  }
  void kspatial(const KBArgs& a, long long n) {
    for (long long i = 0; i < n; ++i) kernel_point(a, i);
  }
  void koct(const KBArgs& a, long long n, int prec) {
    for (long long i = 0; i < n; ++i) {
      if (prec == PREC_DOUBLE) octant_point<double>(a, i);
      else octant_point<float>(a, i);
    }
  }
};

struct EmuPlan {
  Plan p;
  std::vector<double> twd;
  std::vector<float> twf;
};

}  // namespace

extern "C" {

const char* emu_last_error(void) { return g_err.c_str(); }

void emu_set_fast(int on) { g_use_fast = on; }

long long emu_fast_calls(int kind) { return g_fast_calls[kind]; }

void* emu_plan_create(int nx, int ny, int nz, int ncomp, int batch, int precision) {
  EmuPlan* h = new EmuPlan();
  g_err = plan_init(&h->p, nx, ny, nz, ncomp, batch, precision);
  if (!g_err.empty()) {
    delete h;
    return nullptr;
  }
  small_grid_variants(h->p.N, h->p.L, h->p.C, h->p.batch, h->p.prec, h->p.variant);   // as em3d_plan_create
  size_t nf = 0;
  for (int a = 0; a < 3; ++a) nf += (size_t)h->p.L[a];
  h->twd.resize(2 * nf);
  h->twf.resize(2 * nf);
  size_t off = 0;
  for (int a = 0; a < 3; ++a) {
    const int L = h->p.L[a];
    for (int t = 0; t < L; ++t) {
      const double ang = -2.0 * M_PI * (double)t / (double)L;
      h->twd[2 * (off + t)] = std::cos(ang);
      h->twd[2 * (off + t) + 1] = std::sin(ang);
      h->twf[2 * (off + t)] = (float)std::cos(ang);
      h->twf[2 * (off + t) + 1] = (float)std::sin(ang);
    }
    h->p.tw_d[a] = h->twd.data() + 2 * off;
    h->p.tw_f[a] = h->twf.data() + 2 * off;
    off += (size_t)L;
  }
  return h;
}

void emu_plan_destroy(void* h) { delete (EmuPlan*)h; }

int emu_plan_variant_count(void* h, int kind) {
  const Plan& p = ((EmuPlan*)h)->p;
  return fast_variant_count(kind, p.L[kind == PASS_S ? 1 : (kind == PASS_X ? 0 : 2)], p.prec);
}

void emu_plan_set_variant(void* h, int kind, int v) { ((EmuPlan*)h)->p.variant[kind] = v; }

void emu_plan_sizes(void* h, long long* khat_elems, long long* work_elems, long long* build_elems) {
  const Plan& p = ((EmuPlan*)h)->p;
  *khat_elems = p.khat_elems;
  *work_elems = p.w1_elems + p.w2_elems;
  *build_elems = p.build_elems;
}

void emu_kernel_build(void* h, double k0, double hx, double hy, double hz, void* khat, void* build_work) {
  const Plan& p = ((EmuPlan*)h)->p;
  const double hh[3] = {hx, hy, hz};
  double sre = 0.0, sim = 0.0;
  if (p.C == 1) {
    const double dv = hx * hy * hz;
    const double a = std::cbrt(3.0 * dv / (4.0 * M_PI));
    const double ka = k0 * a;
    sre = std::cos(ka) + std::sin(ka) * ka - 1.0;
    sim = std::sin(ka) - std::cos(ka) * ka;
  }
  EmuLauncher go;
  run_kernel_build(go, p, k0, hh, sre, sim, khat, build_work);
}

void emu_apply(void* h, const void* khat, const void* chi, int chi_kind, const void* u, void* y, void* work,
               int adjoint) {
  EmuLauncher go;
  run_apply(go, ((EmuPlan*)h)->p, khat, chi, chi_kind, u, y, work, adjoint != 0);
}

// y = A u with the forward passes restricted to the box [lo, lo + n) (win =
// lo[3], n[3]), as em3d_matvec_window
void emu_apply_window(void* h, const void* khat, const void* chi, int chi_kind, const void* u, void* y, void* work,
                      const int* win) {
  EmuLauncher go;
  run_apply(go, ((EmuPlan*)h)->p, khat, chi, chi_kind, u, y, work, false, nullptr, nullptr, nullptr, nullptr,
            PASSES_ALL, win);
}

// y = A u with the fused Krylov steps of the device BiCGStab: op1 (KRY_P /
// KRY_S, updating u in place from r and v) in the first z pass, op5 (KRY_DOT
// / KRY_DOT2) in the last; partial sums of P1 / P5 to part1 / part5.
void emu_apply_kry(void* h, const void* khat, const void* chi, int chi_kind, void* u, void* y, void* work,
                   int op1, const void* r, const void* v, int op5, const void* rh, const void* coef, double* part1,
                   double* part5, long long stride, int* counts) {
  ZKry k1{}, k5{};
  k1.op = op1; k1.r = r; k1.v = v; k1.w = u; k1.coef = (const KryCoef*)coef;
  k1.part = part1; k1.part_stride = stride; k1.count = &counts[0];
  k5.op = op5; k5.rh = rh; k5.coef = (const KryCoef*)coef;
  k5.part = part5; k5.part_stride = stride; k5.count = &counts[1];
  EmuLauncher go;
  run_apply(go, ((EmuPlan*)h)->p, khat, chi, chi_kind, u, y, work, false, nullptr, &k1, &k5);
}

}  // extern "C"
