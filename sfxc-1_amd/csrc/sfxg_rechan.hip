// Lag re-channelisation of integration_write when fft_size_correlation >
// number_channels (correlation_core.cc:417-440): per row of normalised
// visibilities (N_c + 1 bins), optional normalisation of every bin to unit
// magnitude, the mask, a 2 N_c-point c2r transform to lags, the fold of lags
// -nchan..nchan into 2 nchan points (the two lags at +-nchan averaged), the
// lag window, a 2 nchan-point r2c transform and division by 2 N_c.
//
// One workgroup per row. The row's lag series stays in LDS (N_c float2, at
// most 128 KB for N_c = 16384): both real transforms run as half-length
// complex Stockham transforms (radix 4, one radix-2 stage for odd powers)
// whose stages pass through registers, so one LDS buffer suffices.
#include "sfxg_internal.h"

namespace sfxg {
namespace {

constexpr int RC_MAX_THREADS = 1024;
constexpr int RC_Q4 = 4;  // radix-4 butterflies per thread and stage, at most
constexpr int RC_Q2 = 8;  // radix-2 butterflies (and fold pairs) per thread, at most

__device__ __forceinline__ float2 rc_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(SIGN 2 pi i e / Nt) from the table tw[e] = exp(-2 pi i e / Nt).
template <int SIGN>
__device__ __forceinline__ float2 rc_tw(const float2 *__restrict__ tw, int e) {
  float2 w = tw[e];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// In-place complex transform of n points (a power of two, n <= 16 blockDim)
// in buf, unnormalised, exp(SIGN 2 pi i jk / n). Stockham stages: butterfly j
// of a stage with ns-point sub-transforms reads buf[j + r n/R], twiddles by
// exp(SIGN 2 pi i (j mod ns) r / (ns R)) and writes buf[(j - j mod ns) R +
// j mod ns + r ns]. All reads of a stage land in registers before the first
// write. Ends with a barrier.
template <int SIGN>
__device__ void rc_fft(float2 *buf, int n, const float2 *__restrict__ tw, int Nt) {
  const int T = blockDim.x, tid = threadIdx.x;
  int ns = 1;
  while (ns < n) {
    if (n / ns >= 4) {
      const int nb = n >> 2, step = Nt / (4 * ns);
      float2 v[RC_Q4][4];
#pragma unroll
      for (int q = 0; q < RC_Q4; q++) {
        const int j = tid + q * T;
        if (j < nb) {
#pragma unroll
          for (int r = 0; r < 4; r++) v[q][r] = buf[j + r * nb];
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < RC_Q4; q++) {
        const int j = tid + q * T;
        if (j < nb) {
          const int k = j & (ns - 1);
#pragma unroll
          for (int r = 1; r < 4; r++) v[q][r] = rc_mul(v[q][r], rc_tw<SIGN>(tw, k * r * step));
          const float2 s0 = make_float2(v[q][0].x + v[q][2].x, v[q][0].y + v[q][2].y);
          const float2 s1 = make_float2(v[q][0].x - v[q][2].x, v[q][0].y - v[q][2].y);
          const float2 s2 = make_float2(v[q][1].x + v[q][3].x, v[q][1].y + v[q][3].y);
          const float2 s3 = make_float2(v[q][1].x - v[q][3].x, v[q][1].y - v[q][3].y);
          // s3 times exp(SIGN i pi / 2)
          const float2 r3 = SIGN < 0 ? make_float2(s3.y, -s3.x) : make_float2(-s3.y, s3.x);
          const int o = ((j - k) << 2) + k;
          buf[o] = make_float2(s0.x + s2.x, s0.y + s2.y);
          buf[o + ns] = make_float2(s1.x + r3.x, s1.y + r3.y);
          buf[o + 2 * ns] = make_float2(s0.x - s2.x, s0.y - s2.y);
          buf[o + 3 * ns] = make_float2(s1.x - r3.x, s1.y - r3.y);
        }
      }
      __syncthreads();
      ns <<= 2;
    } else {
      const int nb = n >> 1, step = Nt / (2 * ns);
      float2 v[RC_Q2][2];
#pragma unroll
      for (int q = 0; q < RC_Q2; q++) {
        const int j = tid + q * T;
        if (j < nb) {
          v[q][0] = buf[j];
          v[q][1] = buf[j + nb];
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < RC_Q2; q++) {
        const int j = tid + q * T;
        if (j < nb) {
          const int k = j & (ns - 1);
          const float2 b = rc_mul(v[q][1], rc_tw<SIGN>(tw, k * step));
          const int o = ((j - k) << 1) + k;
          buf[o] = make_float2(v[q][0].x + b.x, v[q][0].y + b.y);
          buf[o + ns] = make_float2(v[q][0].x - b.x, v[q][0].y - b.y);
        }
      }
      __syncthreads();
      ns <<= 1;
    }
  }
}

// Bin idx of the row as the c2r transform takes it: normalised to unit
// magnitude (when asked and nonzero), times the mask, and, for bins 0 and
// N_c, without its imaginary part (FFTW's c2r does not read it).
__device__ __forceinline__ float2 rc_bin(const float2 *__restrict__ x, int idx, int nc,
                                        const float *__restrict__ mask, int normalize) {
  float2 v = x[idx];
  if (normalize) {
    const float a = hypotf(v.x, v.y);
    if (a != 0.f) {
      v.x /= a;
      v.y /= a;
    }
  }
  const float m = mask[idx];
  v.x *= m;
  v.y *= m;
  if (idx == 0 || idx == nc) v.y = 0.f;
  return v;
}

// Lag m of the c2r output held as buf[n] = (lag 2n, lag 2n + 1).
__device__ __forceinline__ float rc_lag(const float2 *buf, int m) {
  const float2 v = buf[m >> 1];
  return (m & 1) ? v.y : v.x;
}

// Point i of the folded 2 nchan-point lag series (correlation_core.cc:427-432).
__device__ __forceinline__ float rc_fold(const float2 *buf, int i, int nc, int nchan) {
  if (i < nchan) return rc_lag(buf, i);
  if (i == nchan) return (rc_lag(buf, nchan) + rc_lag(buf, 2 * nc - nchan)) / 2;
  return rc_lag(buf, 2 * nc - 2 * nchan + i);
}

__global__ void __launch_bounds__(RC_MAX_THREADS)
    k_rechannel(const float2 *__restrict__ in, float2 *__restrict__ out, int nc, int nchan,
                const float *__restrict__ mask, const float *__restrict__ window, int normalize,
                const float2 *__restrict__ tw) {
  extern __shared__ float2 rbuf[];
  const int T = blockDim.x, tid = threadIdx.x, Nt = 2 * nc;
  const float2 *x = in + (size_t)blockIdx.x * (nc + 1);
  // c2r of 2 N_c points as an N_c-point backward transform of
  // Z[k] = (X[k] + conj X[N_c - k]) + i exp(+2 pi i k / 2N_c) (X[k] - conj X[N_c - k])
  for (int k = tid; k < nc; k += T) {
    const float2 a = rc_bin(x, k, nc, mask, normalize);
    const float2 b = rc_bin(x, nc - k, nc, mask, normalize);
    const float2 s = make_float2(a.x + b.x, a.y - b.y);
    const float2 d = make_float2(a.x - b.x, a.y + b.y);
    const float2 t = rc_mul(d, rc_tw<+1>(tw, k));
    rbuf[k] = make_float2(s.x - t.y, s.y + t.x);
  }
  __syncthreads();
  rc_fft<+1>(rbuf, nc, tw, Nt);
  // fold and window into nchan pairs (y[2m], y[2m + 1])
  float2 u[RC_Q2];
#pragma unroll
  for (int q = 0; q < RC_Q2; q++) {
    const int m = tid + q * T;
    if (m < nchan)
      u[q] = make_float2(rc_fold(rbuf, 2 * m, nc, nchan) * window[2 * m],
                         rc_fold(rbuf, 2 * m + 1, nc, nchan) * window[2 * m + 1]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RC_Q2; q++) {
    const int m = tid + q * T;
    if (m < nchan) rbuf[m] = u[q];
  }
  __syncthreads();
  rc_fft<-1>(rbuf, nchan, tw, Nt);
  // r2c split into nchan + 1 bins, divided by 2 N_c
  float2 *o = out + (size_t)blockIdx.x * (nchan + 1);
  const float inv = 1.0f / (float)Nt;
  const int wstep = Nt / (2 * nchan);
  for (int k = tid; k <= nchan; k += T) {
    const float2 zk = rbuf[k & (nchan - 1)], zn = rbuf[(nchan - k) & (nchan - 1)];
    const float er = 0.5f * (zk.x + zn.x), ei = 0.5f * (zk.y - zn.y);
    const float orr = 0.5f * (zk.y + zn.y), oi = -0.5f * (zk.x - zn.x);
    const float2 w = k < nchan ? tw[k * wstep] : make_float2(-1.f, 0.f);
    o[k] = make_float2((er + (orr * w.x - oi * w.y)) * inv, (ei + (orr * w.y + oi * w.x)) * inv);
  }
}

}  // namespace

int rechannel_threads(int nc) { return std::min(RC_MAX_THREADS, std::max(64, nc / 4)); }

bool rechannel_supported(int nc, int nchan) {
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  if (!pow2(nc) || !pow2(nchan) || nchan < 2 || nchan >= nc || nc > 16384) return false;
  const int T = rechannel_threads(nc);
  return nc <= 4 * RC_Q4 * T && nchan <= RC_Q2 * T;
}

hipError_t launch_rechannel(const float2 *in, float2 *out, int rows, int nc, int nchan, const float *mask,
                            const float *window, int normalize, const float2 *tw, hipStream_t s) {
  if (!rechannel_supported(nc, nchan) || rows <= 0) return hipErrorInvalidValue;
  const int T = rechannel_threads(nc);
  const size_t lds = (size_t)nc * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void *)k_rechannel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rechannel, dim3(rows), dim3(T), lds, s, in, out, nc, nchan, mask, window, normalize, tw);
  return hipGetLastError();
}

}  // namespace sfxg
