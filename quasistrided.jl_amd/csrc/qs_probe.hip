// Measured peaks of the current device, beside the datasheet ones of the
// TargetProfile: what the FP64 MFMA units sustain (independent chains of
// v_mfma_f64_16x16x4_f64 on registers, enough waves per SIMD to hide their
// latency) and what a plain device-to-device copy moves.  The reference's
// target_profile() detects its host's facts (src/hardware/target.jl:181-191);
// on the GPU the rooflines are priced against both these and the datasheet.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "qs_gemm_impl.h"

namespace qs {

static constexpr int kProbeChains = 8;

__global__ __launch_bounds__(256) void probe_mfma_f64_kernel(double *out, int iters, double seed) {
  qs_d4 acc[kProbeChains];
#pragma unroll
  for (int c = 0; c < kProbeChains; ++c) acc[c] = qs_d4{0, 0, 0, 0};
  const double a = seed + 1e-3 * threadIdx.x, b = seed - 1e-3 * threadIdx.x;
  // Sixteen steps per trip (iters is a multiple of 16): the compiler moves
  // the accumulators between register files around the loop's back edge, and
  // with one step per trip those moves cost a third of the MFMA rate.
  for (int i = 0; i < iters; i += 16) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int c = 0; c < kProbeChains; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  double s = 0;
#pragma unroll
  for (int c = 0; c < kProbeChains; ++c) s += (acc[c][0] + acc[c][1]) + (acc[c][2] + acc[c][3]);
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

typedef double qs_d2 __attribute__((ext_vector_type(2)));

// 16 bytes per lane, four loads in flight per thread per trip.
__global__ __launch_bounds__(256) void probe_copy_kernel(const qs_d2 *__restrict__ src, qs_d2 *__restrict__ dst,
                                                         int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const qs_d2 v0 = src[i], v1 = src[i + stride], v2 = src[i + 2 * stride], v3 = src[i + 3 * stride];
    dst[i] = v0;
    dst[i + stride] = v1;
    dst[i + 2 * stride] = v2;
    dst[i + 3 * stride] = v3;
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace qs

using namespace qs;

extern "C" {

qs_status qs_target_measure(double *fp64_tflops, double *copy_gbs, void *stream) {
  if (!fp64_tflops || !copy_gbs) return fail(QS_ERR_ARGUMENT, "qs_target_measure: null argument");
  hipStream_t s = (hipStream_t)stream;
  const int cus = device_facts().cus;
  const int grid = 2 * cus, iters = 2000;
  const size_t copy_bytes = size_t(1) << 30;
  double *out = nullptr;
  void *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  qs_status st = QS_OK;
  auto bad = [&](hipError_t e, const char *what) {
    if (st == QS_OK) st = fail(QS_ERR_DEVICE, std::string("qs_target_measure: ") + what + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = dev_malloc(&out, sizeof(double) * (size_t)grid * 256)) != hipSuccess) bad(e, "buffer");
  if (st == QS_OK && (e = dev_malloc(&src, copy_bytes)) != hipSuccess) bad(e, "copy source");
  if (st == QS_OK && (e = dev_malloc(&dst, copy_bytes)) != hipSuccess) bad(e, "copy target");
  if (st == QS_OK && (e = hipEventCreate(&e0)) != hipSuccess) bad(e, "event");
  if (st == QS_OK && (e = hipEventCreate(&e1)) != hipSuccess) bad(e, "event");
  double best_mfma = 0, best_copy = 0;
  if (st == QS_OK && (e = hipMemsetAsync(src, 0, copy_bytes, s)) != hipSuccess) bad(e, "fill");
  for (int rep = 0; rep < 4 && st == QS_OK; ++rep) {
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(probe_mfma_f64_kernel, dim3(grid), dim3(256), 0, s, out, iters, 1.0 + rep);
    (void)hipEventRecord(e1, s);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) { bad(e, "FP64 MFMA probe"); break; }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double flops = (double)grid * 4 * iters * kProbeChains * 2.0 * 16 * 16 * 4;
    if (rep > 0 && ms > 0) best_mfma = std::max(best_mfma, flops / (ms * 1e-3) / 1e12);
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(probe_copy_kernel, dim3(4 * cus * 4), dim3(256), 0, s, (const qs_d2 *)src, (qs_d2 *)dst,
                       (int64_t)(copy_bytes / 16));
    (void)hipEventRecord(e1, s);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) { bad(e, "copy probe"); break; }
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (rep > 0 && ms > 0) best_copy = std::max(best_copy, 2.0 * (double)copy_bytes / (ms * 1e-3) / 1e9);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)dev_free(out);
  (void)dev_free(src);
  (void)dev_free(dst);
  *fp64_tflops = best_mfma;
  *copy_gbs = best_copy;
  return st;
}

}  // extern "C"
