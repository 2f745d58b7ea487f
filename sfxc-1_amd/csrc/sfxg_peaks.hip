// Achievable-peak microbenchmarks for the roofline fractions the benchmark
// reports beside the vendor peaks: a streaming copy (HBM read + write), a
// dense FP16 MFMA loop (v_mfma_f32_32x32x16_f16, FP32 accumulation) and an
// FP32 VALU loop (v_mul_f32 / v_add_f32, not packed: the issue rate of
// wave-level VALU instructions the correlator's FFTs and FP32 X-step run at;
// a gfx950 SIMD can issue two independent VALU instructions per quad-cycle).
// Built
// into its own library (libsfxg_peaks.so); not part of the correlator ABI.
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f16v_t __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                              int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Each wave keeps four independent accumulators so that back-to-back MFMAs
// do not wait on each other; the result is folded and stored so the loop is
// not removed.
__global__ __launch_bounds__(256) void k_mfma(int iters, float *out) {
  const int lane = threadIdx.x & 63;
  h8_t a, b;
  for (int k = 0; k < 8; k++) {
    a[k] = (_Float16)(0.001f * (lane + k));
    b[k] = (_Float16)(0.002f * (lane - k));
  }
  f16v_t acc[4];
  for (int q = 0; q < 4; q++)
    for (int k = 0; k < 16; k++) acc[q][k] = 0.0f;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[q], 0, 0, 0);
  }
  float s = 0.0f;
  for (int q = 0; q < 4; q++)
    for (int k = 0; k < 16; k++) s += acc[q][k];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Eight independent chains per thread of a multiply and an add each (16
// VALU instructions per iteration), so that a SIMD always has independent
// ready instructions; the sum is stored so the loop is not removed.
__global__ __launch_bounds__(256) void k_valu(int iters, float *out) {
#pragma clang fp contract(off)
  float a[8];
  const float b = 0.999f + 1e-7f * threadIdx.x, c = 1e-4f;
  for (int k = 0; k < 8; k++) a[k] = 0.001f * (threadIdx.x + k);
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = a[k] * b + c;
  }
  float s = 0.0f;
  for (int k = 0; k < 8; k++) s += a[k];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

#define PK_CHECK(call)                 \
  do {                                 \
    if ((call) != hipSuccess) return -2; \
  } while (0)

extern "C" {

// Copies n_bytes (a multiple of 16) from src to dst `iters` times on the
// default stream after one untimed pass; *ms = mean time per copy.
int sfxg_peak_copy(const void *src, void *dst, int64_t n_bytes, int iters, float *ms) {
  if (n_bytes <= 0 || n_bytes % 16 != 0 || iters <= 0 || ms == nullptr) return -1;
  const int64_t n = n_bytes / 16;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -2;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -2;
  const dim3 grid((unsigned)(cus * 8)), block(256);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -2;
  hipLaunchKernelGGL(k_copy, grid, block, 0, 0, (const float4 *)src, (float4 *)dst, n);
  PK_CHECK(hipEventRecord(e0, 0));
  for (int i = 0; i < iters; i++)
    hipLaunchKernelGGL(k_copy, grid, block, 0, 0, (const float4 *)src, (float4 *)dst, n);
  PK_CHECK(hipEventRecord(e1, 0));
  PK_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  PK_CHECK(hipEventElapsedTime(&t, e0, e1));
  PK_CHECK(hipEventDestroy(e0));
  PK_CHECK(hipEventDestroy(e1));
  if (hipGetLastError() != hipSuccess) return -3;
  *ms = t / iters;
  return 0;
}

// Runs the MFMA loop on `waves_per_cu` waves per CU; *tflops = dense FP16
// rate (2 * 32 * 32 * 16 flops per instruction).
int sfxg_peak_mfma_fp16(int iters, int waves_per_cu, double *tflops) {
  if (iters <= 0 || waves_per_cu <= 0 || waves_per_cu % 4 != 0 || tflops == nullptr) return -1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -2;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -2;
  const dim3 grid((unsigned)(cus * waves_per_cu / 4)), block(256);
  float *out = nullptr;
  if (hipMalloc(&out, (size_t)grid.x * 256 * sizeof(float)) != hipSuccess) return -2;
  hipEvent_t e0, e1;
  PK_CHECK(hipEventCreate(&e0));
  PK_CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_mfma, grid, block, 0, 0, iters / 10 + 1, out);
  PK_CHECK(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_mfma, grid, block, 0, 0, iters, out);
  PK_CHECK(hipEventRecord(e1, 0));
  PK_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  PK_CHECK(hipEventElapsedTime(&t, e0, e1));
  PK_CHECK(hipEventDestroy(e0));
  PK_CHECK(hipEventDestroy(e1));
  PK_CHECK(hipFree(out));
  if (hipGetLastError() != hipSuccess || t <= 0.0f) return -3;
  const double waves = (double)grid.x * 4;
  *tflops = waves * iters * 4 * 32768.0 / (t * 1e-3) / 1e12;
  return 0;
}

// Runs the multiply-add loop on `waves_per_cu` waves per CU; *insts_per_s =
// wave-level FP32 VALU instructions per second (16 per wave and iteration).
int sfxg_peak_valu(int iters, int waves_per_cu, double *insts_per_s) {
  if (iters <= 0 || waves_per_cu <= 0 || waves_per_cu % 4 != 0 || insts_per_s == nullptr) return -1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -2;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -2;
  const dim3 grid((unsigned)(cus * waves_per_cu / 4)), block(256);
  float *out = nullptr;
  if (hipMalloc(&out, (size_t)grid.x * 256 * sizeof(float)) != hipSuccess) return -2;
  hipEvent_t e0, e1;
  PK_CHECK(hipEventCreate(&e0));
  PK_CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_valu, grid, block, 0, 0, iters / 10 + 1, out);
  PK_CHECK(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_valu, grid, block, 0, 0, iters, out);
  PK_CHECK(hipEventRecord(e1, 0));
  PK_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  PK_CHECK(hipEventElapsedTime(&t, e0, e1));
  PK_CHECK(hipEventDestroy(e0));
  PK_CHECK(hipEventDestroy(e1));
  PK_CHECK(hipFree(out));
  if (hipGetLastError() != hipSuccess || t <= 0.0f) return -3;
  const double waves = (double)grid.x * 4;
  *insts_per_s = waves * iters * 16.0 / (t * 1e-3);
  return 0;
}

}  // extern "C"
