// Ceilings of the roofline figures, measured on the device the library runs
// on: the MFMA rate on register operands (no memory traffic) in FP64 and FP32
// with the 16x16x4 instructions the factorisation uses, and the bandwidth of
// a streaming copy through HBM.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "nelson_amd.h"
#include "tile_device.h"

using namespace nsn;

namespace {

constexpr int kChains = 8;  // independent accumulators per wave (covers the MFMA latency)

template <typename T>
__global__ __launch_bounds__(256) void k_mfma_rate(int iters, T* out) {
  typedef typename Mfma<T>::acc_t acc_t;
  acc_t acc[kChains];
#pragma unroll
  for (int c = 0; c < kChains; c++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[c][v] = (T)0;
  const T a = (T)(1.0 + 1e-3 * (threadIdx.x & 63)), b = (T)(1.0 - 1e-3 * (threadIdx.x & 15));
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int c = 0; c < kChains; c++) acc[c] = Mfma<T>::fma(a, b, acc[c]);
  }
  T s = (T)0;
#pragma unroll
  for (int c = 0; c < kChains; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  if (s == (T)-1) out[blockIdx.x] = s;  // never true; keeps the products alive
}

__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

#define CK(expr)                                                                                   \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) {                                                                        \
      rc = fail(NSN_EHIP, "nsn_measure_peaks: %s failed: %s", #expr, hipGetErrorString(e_));       \
      goto done;                                                                                   \
    }                                                                                              \
  } while (0)

}  // namespace

extern "C" int nsn_measure_peaks(int device, double* mfma_f64_tflops, double* mfma_f32_tflops, double* copy_gbps) {
  int rc = NSN_OK, ndev = 0, ncu = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void *a = nullptr, *b = nullptr, *o = nullptr;
  const long long bytes = 1LL << 30;
  const int iters = 4096;
  float ms = 0.f;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NSN_EHIP, "nsn_measure_peaks: no HIP device");
  if (device < 0 || device >= ndev) return fail(NSN_EINVAL, "nsn_measure_peaks: bad device %d", device);
  CK(hipSetDevice(device));
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipMalloc(&o, sizeof(double) * 65536));
  {
    // 8 workgroups of 4 waves per CU; every wave: iters x kChains MFMAs of 16 x 16 x 4 x 2 flop
    const int grid = 8 * ncu;
    const double flop = (double)grid * 4 * iters * kChains * (16.0 * 16 * 4 * 2);
    for (int rep = 0; rep < 2; rep++) {  // the first launch warms up
      CK(hipEventRecord(e0, st));
      hipLaunchKernelGGL(k_mfma_rate<double>, dim3(grid), dim3(256), 0, st, iters, (double*)o);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
    }
    if (mfma_f64_tflops) *mfma_f64_tflops = flop / (ms * 1e-3) / 1e12;
    for (int rep = 0; rep < 2; rep++) {
      CK(hipEventRecord(e0, st));
      hipLaunchKernelGGL(k_mfma_rate<float>, dim3(grid), dim3(256), 0, st, iters, (float*)o);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
    }
    if (mfma_f32_tflops) *mfma_f32_tflops = flop / (ms * 1e-3) / 1e12;
  }
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemsetAsync(a, 0, bytes, st));
  {
    const long long n = bytes / 16;
    float best = 1e30f;
    for (int rep = 0; rep < 4; rep++) {
      CK(hipEventRecord(e0, st));
      hipLaunchKernelGGL(k_stream_copy, dim3(16 * ncu), dim3(256), 0, st, (const float4*)a, (float4*)b, n);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    if (copy_gbps) *copy_gbps = 2.0 * (double)bytes / (best * 1e-3) / 1e9;  // read + write
  }
  CK(hipGetLastError());
done:
  if (st) (void)hipStreamSynchronize(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (o) (void)hipFree(o);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}
