/*
 * Sustained throughput of the current device, measured: the ceilings that
 * the engine's kernels reach in practice, reported beside the datasheet
 * peaks that the roofline fractions use.
 *   HBM read     streaming read of a 1 GiB buffer, 16 bytes per lane, grid
 *                stride (the access of the ACA panel passes, without their
 *                arithmetic)
 *   HBM copy     read + write of 512 MiB (bytes read plus bytes written)
 *   FP64 MFMA    v_mfma_f64_16x16x4_f64 back to back on four accumulators
 *   FP64 vector  v_fma_f64 on eight independent chains per lane
 *   FP32 MFMA    v_mfma_f32_16x16x4_f32 back to back on four accumulators
 * Each is the best of five timed launches (HIP events) after one untimed
 * launch, each launch long enough (10-30 ms) for the clocks to settle.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "peaks.h"

namespace hodlr {

namespace {

constexpr int TPB = 256;

__global__ void __launch_bounds__(TPB) read_kernel(const double2* __restrict__ p, size_t n2,
                                                   double* __restrict__ sink) {
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)TPB + threadIdx.x; i < n2; i += (size_t)gridDim.x * TPB) {
    const double2 v = p[i];
    s += v.x + v.y;
  }
  if (s == 1.2345678) sink[0] = s;  // never true for the zero-filled buffer
}

__global__ void __launch_bounds__(TPB) copy_kernel(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, size_t n2) {
  for (size_t i = blockIdx.x * (size_t)TPB + threadIdx.x; i < n2; i += (size_t)gridDim.x * TPB)
    dst[i] = src[i];
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(TPB) mfma64_kernel(int iters, double* __restrict__ sink) {
  v4d acc[4];
  for (int i = 0; i < 4; i++) acc[i] = (v4d){0.0, 0.0, 0.0, 0.0};
  const double a = 1e-3 * threadIdx.x, b = 1e-3 * blockIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double s = 0.0;
  for (int i = 0; i < 4; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  if (s == 1.2345678) sink[0] = s;
}

__global__ void __launch_bounds__(TPB) mfma32_kernel(int iters, double* __restrict__ sink) {
  v4f acc[4];
  for (int i = 0; i < 4; i++) acc[i] = (v4f){0.f, 0.f, 0.f, 0.f};
  const float a = 1e-3f * threadIdx.x, b = 1e-3f * blockIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
  }
  float s = 0.f;
  for (int i = 0; i < 4; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  if (s == 1.2345678f) sink[0] = s;
}

__global__ void __launch_bounds__(TPB) fma64_kernel(int iters, double* __restrict__ sink) {
  double c[8];
  for (int i = 0; i < 8; i++) c[i] = 1e-3 * (threadIdx.x + i);
  const double a = 0.999, b = 1e-3 * blockIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = fma(a, c[i], b);
  }
  double s = 0.0;
  for (int i = 0; i < 8; i++) s += c[i];
  if (s == 1.2345678) sink[0] = s;
}

/* best of five launches after one untimed, in seconds */
template <class F>
double best_time(hipStream_t s, hipEvent_t e0, hipEvent_t e1, F&& launch) {
  launch();
  HODLR_HIP_CHECK(hipGetLastError());
  HODLR_HIP_CHECK(hipStreamSynchronize(s));
  float best = 1e30f;
  for (int r = 0; r < 5; r++) {
    HODLR_HIP_CHECK(hipEventRecord(e0, s));
    launch();
    HODLR_HIP_CHECK(hipEventRecord(e1, s));
    HODLR_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HODLR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, ms);
  }
  return 1e-3 * best;
}

}  // namespace

DevicePeaks measure_device_peaks() {
  constexpr size_t BYTES = size_t(1) << 30;
  constexpr int WGS = 4096;
  DevicePeaks out{};
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double* buf = nullptr;
  double* sink = nullptr;
  auto cleanup = [&]() {
    if (s) (void)hipStreamSynchronize(s);
    if (buf) (void)hipFree(buf);
    if (sink) (void)hipFree(sink);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
  };
  try {
    HODLR_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HODLR_HIP_CHECK(hipEventCreate(&e0));
    HODLR_HIP_CHECK(hipEventCreate(&e1));
    HODLR_HIP_CHECK(hipMalloc(&buf, BYTES));
    HODLR_HIP_CHECK(hipMalloc(&sink, 64));
    HODLR_HIP_CHECK(hipMemsetAsync(buf, 0, BYTES, s));
    const size_t n2 = BYTES / sizeof(double2);
    const double2* src = reinterpret_cast<const double2*>(buf);
    double t = best_time(s, e0, e1, [&] { read_kernel<<<2048, TPB, 0, s>>>(src, n2, sink); });
    out.hbm_read_gbs = BYTES / t * 1e-9;
    t = best_time(s, e0, e1, [&] {
      copy_kernel<<<2048, TPB, 0, s>>>(src, reinterpret_cast<double2*>(buf) + n2 / 2, n2 / 2);
    });
    out.hbm_copy_gbs = BYTES / t * 1e-9;  // BYTES / 2 read + BYTES / 2 written
    const int it64 = 8192;
    t = best_time(s, e0, e1, [&] { mfma64_kernel<<<WGS, TPB, 0, s>>>(it64, sink); });
    out.fp64_mfma_tflops = (double)WGS * (TPB / 64) * it64 * 4 * 2.0 * 16 * 16 * 4 / t * 1e-12;
    const int itv = 65536;
    t = best_time(s, e0, e1, [&] { fma64_kernel<<<WGS, TPB, 0, s>>>(itv, sink); });
    out.fp64_valu_tflops = (double)WGS * TPB * itv * 8 * 2.0 / t * 1e-12;
    const int it32 = 16384;
    t = best_time(s, e0, e1, [&] { mfma32_kernel<<<WGS, TPB, 0, s>>>(it32, sink); });
    out.fp32_mfma_tflops = (double)WGS * (TPB / 64) * it32 * 4 * 2.0 * 16 * 16 * 4 / t * 1e-12;
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  return out;
}

}  // namespace hodlr
