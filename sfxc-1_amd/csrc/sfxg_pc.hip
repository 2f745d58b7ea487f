// Extra phase centres of Correlation_core (correlation_core.cc:515-592):
// folding each sub-integration's partial sums into every phase centre.
// Autos are added to every centre as they are; crosses are added to the
// pointing centre as they are and to centre j as
//   amplitude * partial[ch] * exp(i (phi + ch * delta)),
// amplitude, phi and delta per (sub-integration, centre, baseline) from the
// host (the FP64 scalars of uvshift). The reference advances exp(i phi) by
// an FP64 recurrence; here each channel's angle is formed and rotated in
// FP64, which agrees to ~1e-13 before both round to float.
#include "sfxg_internal.h"

namespace sfxg {
namespace {

__global__ __launch_bounds__(256) void k_fold_combine_pc(const float2 *__restrict__ partial, int n_k,
                                                         int n_baselines, int n_ch, int M,
                                                         const UvShift *__restrict__ uv, int n_pc,
                                                         float2 *__restrict__ vis) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)n_baselines * n_ch;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = (int)(i / n_ch), ch = (int)(i - (int64_t)b * n_ch);
  // pointing centre: the sums in fold order
  float2 t = vis[i];
  for (int k = 0; k < n_k; k++) {
    const float2 p = partial[k * n + i];
    t.x = t.x + p.x;
    t.y = t.y + p.y;
  }
  vis[i] = t;
  for (int j = 1; j < n_pc; j++) {
    float2 u = vis[j * n + i];
    for (int k = 0; k < n_k; k++) {
      const float2 p = partial[k * n + i];
      if (b < M) {  // auto-correlation (baselines 0 .. M - 1)
        u.x = u.x + p.x;
        u.y = u.y + p.y;
      } else {
        const UvShift s = uv[((int64_t)k * (n_pc - 1) + (j - 1)) * n_baselines + b];
        double sn, cs;
        sincos(s.phi + (double)ch * s.delta, &sn, &cs);
        const float c = (float)cs, sf = (float)sn;
        const float ar = s.amp * p.x, ai = s.amp * p.y;
        u.x = u.x + (ar * c - ai * sf);
        u.y = u.y + (ar * sf + ai * c);
      }
    }
    vis[j * n + i] = u;
  }
}

}  // namespace

hipError_t launch_fold_combine_pc(const float2 *partial, int n_k, int n_baselines, int n_ch, int M,
                                  const UvShift *uv, int n_pc, float2 *vis, hipStream_t s) {
  if (n_k <= 0) return hipSuccess;
  const int64_t n = (int64_t)n_baselines * n_ch;
  hipLaunchKernelGGL(k_fold_combine_pc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, n_k,
                     n_baselines, n_ch, M, uv, n_pc, vis);
  return hipGetLastError();
}

}  // namespace sfxg
