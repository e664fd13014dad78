// Krylov vector kernels: conjugated dot products reduced in FP64 with
// wavefront shuffles, and fused three-term complex updates.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/em3d_hip.h"
#include "e3_common.hpp"
#include "e3_error.hpp"

using namespace e3;

namespace {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-level sum of (re, im); result valid in thread 0.
__device__ __forceinline__ void block_sum2(double& re, double& im) {
  __shared__ double sre[kRedThreads / 64], sim[kRedThreads / 64];
  re = wave_sum(re);
  im = wave_sum(im);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sre[w] = re; sim[w] = im; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) { a += sre[i]; b += sim[i]; }
    re = a;
    im = b;
  }
}

// partial[block] = sum conj(x_i) y_i over a grid-stride slice, FP64 accumulation
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_dotc_partial(long long n, const cpx<T>* x, const cpx<T>* y,
                                                              double2* partial) {
  double re = 0.0, im = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const cpx<T> a = x[i], b = y[i];
    re += (double)a.re * (double)b.re + (double)a.im * (double)b.im;
    im += (double)a.re * (double)b.im - (double)a.im * (double)b.re;
  }
  block_sum2(re, im);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(re, im);
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_final(int np, const double2* partial, double2* out) {
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) { re += partial[i].x; im += partial[i].y; }
  block_sum2(re, im);
  if (threadIdx.x == 0) *out = make_double2(re, im);
}

// out = a x + b y + c z (z may be null)
template <typename T>
__global__ __launch_bounds__(256) void k_axpbypcz(long long n, cpx<T> a, const cpx<T>* x, cpx<T> b,
                                                  const cpx<T>* y, cpx<T> c, const cpx<T>* z, cpx<T>* out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    cpx<T> v = a * x[i] + b * y[i];
    if (z) v = v + c * z[i];
    out[i] = v;
  }
}

// dst = beta dst + a src, converting between complex64 and complex128
template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert_axpy(long long n, const cpx<S>* src, cpx<double> a, cpx<D>* dst,
                                                      int accumulate) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const cpx<double> x = mk((double)src[i].re, (double)src[i].im);
    cpx<double> y = a * x;
    if (accumulate) y = y + mk((double)dst[i].re, (double)dst[i].im);
    dst[i] = mk((D)y.re, (D)y.im);
  }
}

int kfail(const std::string& m) { return set_error(m); }

}  // namespace

extern "C" {

size_t em3d_reduce_work_bytes(void) { return (size_t)kRedBlocks * sizeof(double2); }

int em3d_dotc(int precision, long long n, const void* x, const void* y, void* out, void* red_work,
              hipStream_t stream) {
  if (!x || !y || !out || !red_work) return kfail("em3d_dotc: null argument");
  if (n < 0) return kfail("em3d_dotc: negative length");
  long long want = (n + kRedThreads - 1) / kRedThreads;
  const int nb = (int)(want < 1 ? 1 : (want > kRedBlocks ? kRedBlocks : want));
  double2* part = (double2*)red_work;
  if (precision == 1)
    hipLaunchKernelGGL(k_dotc_partial<double>, dim3(nb), dim3(kRedThreads), 0, stream, n,
                       (const cpx<double>*)x, (const cpx<double>*)y, part);
  else
    hipLaunchKernelGGL(k_dotc_partial<float>, dim3(nb), dim3(kRedThreads), 0, stream, n,
                       (const cpx<float>*)x, (const cpx<float>*)y, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream, nb, part, (double2*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kfail(std::string("em3d_dotc: ") + hipGetErrorString(e));
  return 0;
}

int em3d_axpbypcz(int precision, long long n, double a_re, double a_im, const void* x, double b_re,
                  double b_im, const void* y, double c_re, double c_im, const void* z, void* out,
                  hipStream_t stream) {
  if (!x || !y || !out) return kfail("em3d_axpbypcz: null argument");
  if (n <= 0) return 0;
  long long want = (n + 255) / 256;
  const unsigned nb = (unsigned)(want > 65536 ? 65536 : want);
  if (precision == 1)
    hipLaunchKernelGGL(k_axpbypcz<double>, dim3(nb), dim3(256), 0, stream, n, mk(a_re, a_im),
                       (const cpx<double>*)x, mk(b_re, b_im), (const cpx<double>*)y, mk(c_re, c_im),
                       (const cpx<double>*)z, (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_axpbypcz<float>, dim3(nb), dim3(256), 0, stream, n, mk((float)a_re, (float)a_im),
                       (const cpx<float>*)x, mk((float)b_re, (float)b_im), (const cpx<float>*)y,
                       mk((float)c_re, (float)c_im), (const cpx<float>*)z, (cpx<float>*)out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kfail(std::string("em3d_axpbypcz: ") + hipGetErrorString(e));
  return 0;
}

int em3d_convert_axpy(long long n, int src_precision, const void* src, double a_re, double a_im, int dst_precision,
                      void* dst, int accumulate, hipStream_t stream) {
  if (!src || !dst) return kfail("em3d_convert_axpy: null argument");
  if (n <= 0) return 0;
  const long long want = (n + 255) / 256;
  const unsigned nb = (unsigned)(want > 65536 ? 65536 : want);
  const cpx<double> a = mk(a_re, a_im);
  if (src_precision == 1 && dst_precision == 1)
    hipLaunchKernelGGL((k_convert_axpy<double, double>), dim3(nb), dim3(256), 0, stream, n, (const cpx<double>*)src, a,
                       (cpx<double>*)dst, accumulate);
  else if (src_precision == 1)
    hipLaunchKernelGGL((k_convert_axpy<double, float>), dim3(nb), dim3(256), 0, stream, n, (const cpx<double>*)src, a,
                       (cpx<float>*)dst, accumulate);
  else if (dst_precision == 1)
    hipLaunchKernelGGL((k_convert_axpy<float, double>), dim3(nb), dim3(256), 0, stream, n, (const cpx<float>*)src, a,
                       (cpx<double>*)dst, accumulate);
  else
    hipLaunchKernelGGL((k_convert_axpy<float, float>), dim3(nb), dim3(256), 0, stream, n, (const cpx<float>*)src, a,
                       (cpx<float>*)dst, accumulate);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kfail(std::string("em3d_convert_axpy: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
