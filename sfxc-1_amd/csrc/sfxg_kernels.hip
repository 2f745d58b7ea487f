// Device kernels of the correlator-node data path for gfx950 (CDNA4).
//
//   k_unpack      Bit2float_worker output (bit2float_worker.cc:326-424) for verification
//   k_stats       byte histograms of bit_statistics (bit_statistics.h:36-41)
//   k_delay       unpack + integer shift fused with Delay_correction's
//                 fractional_bit_shift and fringe_stopping (delay_correction.cc:82-184)
//   k_channelise  window + r2c FFT of 2 N_c points (delay_correction.cc:57-75)
//   k_xstep       Correlation_core::integration_step + sub_integration folds
//                 (correlation_core.cc:241-273, 515-559)
//   k_xstep_pulsar Correlation_core_pulsar::integration_step + dedisperse_buffer
//                 (correlation_core_pulsar.cc:188-238)
//   k_auto_norms / k_normalise  integration_normalize (correlation_core.cc:275-314)
//   k_synthesize  device-side synthetic raw station data
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <vector>

#include "../../include/sfxg.h"
#include "sfxg_internal.h"

namespace sfxg {

namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (sign < 0) or +i (sign > 0)
template <int SIGN>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
template <int SIGN>
__device__ __forceinline__ float2 twiddle(const float2 *tw, int idx) {
  float2 w = tw[idx];
  return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

// DevStream holds generic pointers to global memory. Reaching them through
// these casts lets the compiler emit global loads and stores (scalar loads
// for wave-uniform addresses of read-only data) instead of flat ones, which
// also count against the LDS wait counter and would make every LDS wait
// drain the memory loads in flight.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *cmem(const T *p) {
  return (const __attribute__((address_space(4))) T *)p;
}

// float2 loads and stores in global memory through the same cast (HIP's
// float2 has no members for address-space-qualified objects).
typedef float vf2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gload2(const float2 *p) {
  const vf2_t v = *(const __attribute__((address_space(1))) vf2_t *)p;
  return make_float2(v.x, v.y);
}
// base (wave-uniform) + a 32-bit byte offset: the scalar-base form of the
// global load, one VGPR per address instead of two.
__device__ __forceinline__ float2 gload2(const float2 *base, uint32_t byte_off) {
  const auto *p = (const __attribute__((address_space(1))) char *)base + byte_off;
  const vf2_t v = *(const __attribute__((address_space(1))) vf2_t *)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gstore2(float2 *base, uint32_t byte_off, float2 v) {
  auto *p = (__attribute__((address_space(1))) char *)base + byte_off;
  *(__attribute__((address_space(1))) vf2_t *)p = vf2_t{v.x, v.y};
}

// Windowed point x * w, per component and never contracted into the
// transform's first additions: the reference forms the windowed series
// before transforming it.
__device__ __forceinline__ float2 wmul(float2 x, float2 w) {
#pragma clang fp contract(off)
  return make_float2(x.x * w.x, x.y * w.y);
}

// Spectrum stores. SP = 0 writes float2 bins. SP = SFXG_PREC_FP16 or
// SFXG_PREC_BF16 scales both parts by a power of two (exact) and rounds them
// (to nearest even) to the MFMA X-step's operand type, packed in 32 bits
// with the real part in the low half: the X-step then reads half the bytes
// and converts nothing. SP = SFXG_PREC_SPLIT_FP16 writes the split operands
// of the scaled value x: hi = fp16(x) and lo = fp16(x - hi), as two such
// words per bin (hi pair, then lo pair), 8 bytes like a float2.
template <int SP>
struct SpecOut {
  static constexpr uint32_t W = SP == SFXG_PREC_SPLIT_FP16 ? 2 : 1;  // 32-bit words per bin
  uint32_t *p;
  float scale;
  __device__ __forceinline__ SpecOut(float2 *base, int64_t off, float sc)
      : p(reinterpret_cast<uint32_t *>(base) + W * off), scale(sc) {}
  __device__ __forceinline__ void put(uint32_t k, float2 v) const {
    auto *q = (__attribute__((address_space(1))) char *)p + k * (4u * W);
    if constexpr (SP == SFXG_PREC_SPLIT_FP16) {
      const float xr = v.x * scale, xi = v.y * scale;
      const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
      const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi);
      const uint32_t h = (uint32_t)__builtin_bit_cast(uint16_t, hr) | ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
      const uint32_t l = (uint32_t)__builtin_bit_cast(uint16_t, lr) | ((uint32_t)__builtin_bit_cast(uint16_t, li) << 16);
      typedef uint32_t u2_t __attribute__((ext_vector_type(2)));
      *(__attribute__((address_space(1))) u2_t *)q = u2_t{h, l};
    } else {
      uint32_t re, im;
      if constexpr (SP == SFXG_PREC_BF16) {
        re = __builtin_bit_cast(uint16_t, (__bf16)(v.x * scale));
        im = __builtin_bit_cast(uint16_t, (__bf16)(v.y * scale));
      } else {
        static_assert(SP == SFXG_PREC_FP16, "FP16, BF16 or split FP16 spectra");
        re = __builtin_bit_cast(uint16_t, (_Float16)(v.x * scale));
        im = __builtin_bit_cast(uint16_t, (_Float16)(v.y * scale));
      }
      *(__attribute__((address_space(1))) uint32_t *)q = re | (im << 16);
    }
  }
};
template <>
struct SpecOut<0> {
  float2 *p;
  __device__ __forceinline__ SpecOut(float2 *base, int64_t off, float) : p(base + off) {}
  __device__ __forceinline__ void put(uint32_t k, float2 v) const { gstore2(p, k * (uint32_t)sizeof(float2), v); }
};

// 4-point DFT in place, exp(SIGN 2 pi i jk/4)
template <int SIGN>
__device__ __forceinline__ void radix4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
  const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
  const float2 s13 = cadd(a1, a3), d13 = mul_i<SIGN>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

// Piecewise-cubic delay model in FP64, same operation order as the host and
// the reference's spline evaluation; no contraction so the integer delay
// decisions agree bit for bit with the host.
__device__ double eval_model(const DevModel &m, const double *coef, int64_t ticks, double dflt) {
#pragma clang fp contract(off)
  if (coef == nullptr) return dflt;
  const double sec = (double)(ticks - m.scan_begin) / kClockRate;
  int64_t i = (int64_t)floor(sec / m.interval);
  if (i < 0) i = 0;
  if (i > m.n - 1) i = m.n - 1;
  const double u = sec - (double)i * m.interval;
  const double *c = coef + 4 * i;
  return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
}

__device__ __forceinline__ float decode_sample(const DevStream &s, int64_t psamp) {
  if (s.bits == 2) {
    const uint32_t b = gmem(s.payload)[psamp >> 2];
    return s.lut[(b >> ((psamp & 3) * 2)) & 3];
  }
  const uint32_t b = gmem(s.payload)[psamp >> 3];
  return s.lut[(b >> (psamp & 7)) & 1];
}

// Output sample o of a stream; r is a run index hint that only moves forward.
__device__ __forceinline__ float sample_at(const DevStream &s, int64_t o, int64_t &r) {
  const auto *runs = gmem(s.runs);
  while (r + 1 < s.n_runs && runs[r + 1].out_start <= o) r++;
  const int64_t src = runs[r].src;
  if (src < 0) return 0.0f;
  return decode_sample(s, src + (o - runs[r].out_start));
}

}  // namespace

// ---------------------------------------------------------------------------
// k_unpack: one thread per output sample (verification path)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_unpack(const DevStream *__restrict__ streams, int64_t n_samples) {
  const DevStream &s = streams[blockIdx.y];
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_samples || s.samples == nullptr) return;
  int64_t lo = 0, hi = s.n_runs - 1;
  while (lo < hi) {  // last run with out_start <= o
    const int64_t mid = (lo + hi + 1) >> 1;
    if (s.runs[mid].out_start <= o) lo = mid; else hi = mid - 1;
  }
  int64_t r = lo;
  s.samples[o] = sample_at(s, o, r);
}

// ---------------------------------------------------------------------------
// k_stats: histograms of the counted payload bytes, split by tsys phase.
// Each workgroup takes a contiguous span of counting ordinals of one stream,
// so that its LDS histogram is flushed to HBM once per span; each thread
// takes 16 consecutive ordinals at a time (the workgroup 4096). A thread
// finds its first byte's range and tsys phase once and advances both
// incrementally: ordinals are contiguous across ranges, and the phase flips
// at tsys_count0 and every tsys_half bytes after it. Where the 16 bytes lie
// in one range and one phase (nearly always), all 16 loads are issued before
// the first histogram update.
// ---------------------------------------------------------------------------
constexpr int STATS_RUN = 16;  // ordinals per thread and step

__global__ __launch_bounds__(256) void k_stats(const DevStream *__restrict__ streams, int64_t span) {
  __shared__ int hist[2][256];
  const DevStream &s = streams[blockIdx.y];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) (&hist[0][0])[i] = 0;
  __syncthreads();
  const int64_t n = s.n_counted;
  int64_t total = 0;
  if (n > 0) {
    const CountRange last = s.counted[n - 1];
    total = last.ordinal + last.n_bytes;
  }
  const int64_t g0 = (int64_t)blockIdx.x * span;
  const int64_t g1 = g0 + span < total ? g0 + span : total;
  const int64_t gt = g0 + (int64_t)threadIdx.x * STATS_RUN;
  if (gt < g1) {
    int64_t lo = 0, hi = n - 1;  // last range with ordinal <= gt
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (s.counted[mid].ordinal <= gt) lo = mid; else hi = mid - 1;
    }
    int64_t r = lo;
    int64_t next_ord = r + 1 < n ? s.counted[r + 1].ordinal : INT64_MAX;
    int64_t delta = s.counted[r].byte_start - s.counted[r].ordinal;
    // phase: tsys_on0 before tsys_count0, flipped at it and then every
    // tsys_half bytes (once only without a half period); 1-bit data has none
    int on = 1;
    int64_t flip = INT64_MAX;
    const int64_t c0 = s.tsys_count0, half = s.tsys_half;
    if (s.bits == 2) {
      on = s.tsys_on0;
      if (c0 > 0) {
        if (gt < c0) {
          flip = c0;
        } else if (half > 0) {
          const int64_t k2 = (gt - c0) / half;
          on = s.tsys_on0 ^ (int)((k2 + 1) & 1);
          flip = c0 + (k2 + 1) * half;
        } else {
          on = !s.tsys_on0;
        }
      }
    }
    const uint8_t *pay = s.payload;
    auto advance = [&](int64_t g) {
      while (g >= next_ord) {
        r++;
        delta = s.counted[r].byte_start - s.counted[r].ordinal;
        next_ord = r + 1 < n ? s.counted[r + 1].ordinal : INT64_MAX;
      }
      while (g >= flip) {
        on ^= 1;
        flip = half > 0 ? flip + half : INT64_MAX;
      }
    };
    for (int64_t g = gt; g < g1; g += 256 * STATS_RUN) {
      advance(g);
      const int64_t ge = g + STATS_RUN;
      if (ge <= g1 && ge <= next_ord && ge <= flip) {
        uint8_t b[STATS_RUN];
#pragma unroll
        for (int i = 0; i < STATS_RUN; i++) b[i] = pay[delta + g + i];
        int *h = hist[on ? 0 : 1];
#pragma unroll
        for (int i = 0; i < STATS_RUN; i++) atomicAdd(&h[b[i]], 1);
      } else {
        for (int64_t gi = g; gi < ge && gi < g1; gi++) {
          advance(gi);
          atomicAdd(&hist[on ? 0 : 1][pay[delta + gi]], 1);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    const int v = (&hist[0][0])[i];
    if (v) atomicAdd(&s.hist[i], v);
  }
}

// ---------------------------------------------------------------------------
// Delay correction of one N_d = 256 block by one wavefront
// ---------------------------------------------------------------------------
namespace {

// Orders this wave's LDS writes before its later LDS reads of other lanes'
// data. A wavefront's LDS operations execute in issue order, so keeping the
// compiler from reordering them across this point is enough; waves working
// on their own LDS regions never wait for each other.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave FFT256: one wavefront transforms 256 points, lane holding points
// lane + 64 r (r < 4) in registers, in four radix-4 Stockham stages of span
// ns = 1, 4, 16, 64. Each stage but the last exchanges its outputs through
// the wave's LDS scratch; the last stage's outputs lane + 64 r are already
// where the next step wants them and stay in registers.
//
// Stage twiddles come from LDS tables gathered from exp(-2 pi i j / 256):
// for span ns (4, 16, 64) and r = 1..3 the entry tw256_base(ns) +
// (r - 1) ns + k holds exp(-2 pi i k r / (4 ns)), so the lanes of a stage
// (k = lane mod ns) read consecutive or identical entries. The first stage's
// twiddles are all 1 and are not applied.
constexpr int TW256_ELEMS = 3 * (4 + 16 + 64);
constexpr int WAVE_FFT256_LDS = 320;  // float2 elements of one wave's scratch

__device__ __forceinline__ constexpr int tw256_base(int ns) { return ns == 4 ? 0 : ns == 16 ? 12 : 60; }

// Fills the per-stage twiddle tables from the 256-entry table tw_g.
__device__ __forceinline__ void fill_tw256(float2 *tw, const float2 *__restrict__ tw_g, int tid, int nthreads) {
  for (int i = tid; i < TW256_ELEMS; i += nthreads) {
    const int ns = i < 12 ? 4 : i < 60 ? 16 : 64;
    const int j = i - tw256_base(ns);
    tw[i] = tw_g[(j % ns) * (j / ns + 1) * (64 / ns)];
  }
}

// LDS position of element i written by the stage of span NS. The first two
// stages write with strides of 4 and 16 elements; padding by one element
// (NS = 1) or four (NS = 4) every 16 puts the 16 lanes of each 64-bit access
// on distinct bank pairs, while the next stage's reads of lane + 64 r stay
// consecutive within every 16 lanes. The third stage writes runs of 16 and
// needs none.
template <int NS>
__device__ __forceinline__ int pad256(int i) {
  return NS == 1 ? i + (i >> 4) : NS == 4 ? i + 4 * (i >> 4) : i;
}

template <int SIGN, int NS>
__device__ __forceinline__ void wstage256(float2 (&v)[4], float2 *buf, const float2 *tw, int lane) {
  const int k = lane % NS;
  if (NS > 1) {
    const float2 *t = tw + tw256_base(NS) + k;
    v[1] = cmul(v[1], twiddle<SIGN>(t, 0));
    v[2] = cmul(v[2], twiddle<SIGN>(t, NS));
    v[3] = cmul(v[3], twiddle<SIGN>(t, 2 * NS));
  }
  radix4<SIGN>(v[0], v[1], v[2], v[3]);
  if (NS == 64) return;
  const int idx = (lane / NS) * NS * 4 + k;
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 4; r++) buf[pad256<NS>(idx + r * NS)] = v[r];
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 4; r++) v[r] = buf[pad256<NS>(lane + 64 * r)];
}

// v[r] = data[lane + 64 r] on entry and X[lane + 64 r] on exit.
template <int SIGN>
__device__ __forceinline__ void fft256(float2 (&v)[4], float2 *buf, const float2 *tw, int lane) {
  wstage256<SIGN, 1>(v, buf, tw, lane);
  wstage256<SIGN, 4>(v, buf, tw, lane);
  wstage256<SIGN, 16>(v, buf, tw, lane);
  wstage256<SIGN, 64>(v, buf, tw, lane);
}

}  // namespace

// Per-block FP64 constants of Delay_correction (delay_correction.cc:40-47,
// 97-103, 139-160), one thread per N_d block, and the run index of the
// block's first sample: {constant and linear term of the
// fractional-delay ramp, fringe phase at the block start and its per-sample
// step} and the amplitude. Compiled without contraction in the host's
// operation order, so integer/fractional delay decisions match it bit for bit.
__global__ __launch_bounds__(256) void k_block_consts(const DevStream *__restrict__ streams, int64_t n_blocks,
                                                      int nd) {
#pragma clang fp contract(off)
  const DevStream &s = streams[blockIdx.y];
  const int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= n_blocks) return;
  {
    // the run holding the block's first sample: the payload sample behind
    // it where the whole block lies in that payload run, else ~run index
    const int64_t o = blk * nd;
    int64_t lo = 0, hi = s.n_runs - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (s.runs[mid].out_start <= o) lo = mid; else hi = mid - 1;
    }
    const Run &run = s.runs[lo];
    s.block_src[blk] = run.src >= 0 && run.out_start + run.count >= o + nd ? run.src + (o - run.out_start) : ~lo;
  }
  const int64_t current = s.start_ticks + blk * nd * s.tps;
  const double delay = eval_model(s.model, s.model.delay, current + s.fl_half, 0.0);
  const double dis = delay * s.fs;
  const int integer_delay = (int)floor(dis + .5);
  const double fractional_delay = dis - integer_delay;
  const double dfr = s.fs / nd;
  const double tmp1 = -2.0 * kPi * fractional_delay / s.fs;
  const double tmp2 = kPi * (integer_delay & (4 * s.oversamp - 1)) / (2 * s.oversamp);
  const double c_term = tmp2 - tmp1 * 0.5 * s.bandwidth;
  const double l_term = tmp1 * dfr;
  const double lo_phase = s.start_phase + s.LO_offset * ((double)(current - s.start_ticks) / kClockRate);
  const double *ph = s.model.has_phase ? s.model.phase : nullptr;
  const double p0 = s.center_freq * eval_model(s.model, s.model.delay, current, 0.0) + lo_phase +
                    eval_model(s.model, ph, current, 0.0) / (2 * kPi);
  const double floor_phi = floor(p0);
  const double phi = s.mult_factor_phi * (p0 - floor_phi);
  double p1 = s.center_freq * eval_model(s.model, s.model.delay, current + s.fl, 0.0) + lo_phase +
              s.fl_sec * s.LO_offset + eval_model(s.model, ph, current + s.fl, 0.0) / (2 * kPi);
  p1 = s.mult_factor_phi * (p1 - floor_phi);
  const double dphi = (p1 - phi) / nd;
  const double amp = s.model.has_amp ? eval_model(s.model, s.model.amp, current + s.fl_half, 1.0) : 1.0;
  BlockConst bc;
  bc.c_term = c_term;
  bc.l_term = l_term;
  bc.phi = phi;
  bc.dphi = dphi;
  bc.amp = (float)amp;
  double sn, cs;
  sincos(64.0 * l_term, &sn, &cs);
  bc.ramp_step[0] = (float)cs;
  bc.ramp_step[1] = (float)sn;
  sincos(64.0 * dphi, &sn, &cs);
  bc.fringe_step[0] = (float)cs;
  bc.fringe_step[1] = (float)sn;
  sincos(16.0 * l_term, &sn, &cs);
  bc.ramp_step16[0] = (float)cs;
  bc.ramp_step16[1] = (float)sn;
  sincos(16.0 * dphi, &sn, &cs);
  bc.fringe_step16[0] = (float)cs;
  bc.fringe_step16[1] = (float)sn;
  bc.pad = 0.0f;
  s.block_consts[blk] = bc;
}

// Decode table entry of a 2-bit code from registers.
__device__ __forceinline__ float lut4(uint32_t c, float l0, float l1, float l2, float l3) {
  const float lo = (c & 1) ? l1 : l0, hi = (c & 1) ? l3 : l2;
  return (c & 2) ? hi : lo;
}

// The 256 samples of block blk (wave-uniform), lane holding samples
// lane + 64 q (zeros for a block past the end). A block inside one payload
// run, the common case, takes its four payload bytes per lane with loads
// from one address (offsets 16 q bytes at 2 bit, 8 q at 1 bit) and decodes
// them from the table in registers; a block that crosses a run boundary or
// invalid data goes sample by sample.
__device__ __forceinline__ void load_block(const DevStream &s, int64_t blk, int64_t n_blocks, int lane,
                                           float (&x)[4]) {
  constexpr int nd = 256;
  if (blk >= n_blocks) {
#pragma unroll
    for (int q = 0; q < 4; q++) x[q] = 0.0f;
    return;
  }
  const int64_t o0 = blk * nd;
  const int64_t bs = cmem(s.block_src)[blk];
  if (bs >= 0) {
    const int64_t p0 = bs;
    const auto *pay = gmem(s.payload);
    uint32_t c[4];
    if (s.bits == 2) {
      const int u = (int)(p0 & 3) + lane;
      const auto *b = pay + (p0 >> 2) + (u >> 2);
#pragma unroll
      for (int q = 0; q < 4; q++) c[q] = b[16 * q];
      const int sh = (u & 3) * 2;
      const float l0 = s.lut[0], l1 = s.lut[1], l2 = s.lut[2], l3 = s.lut[3];
#pragma unroll
      for (int q = 0; q < 4; q++) x[q] = lut4((c[q] >> sh) & 3, l0, l1, l2, l3);
    } else {
      const int u = (int)(p0 & 7) + lane;
      const auto *b = pay + (p0 >> 3) + (u >> 3);
#pragma unroll
      for (int q = 0; q < 4; q++) c[q] = b[8 * q];
      const int sh = u & 7;
      const float l0 = s.lut[0], l1 = s.lut[1];
#pragma unroll
      for (int q = 0; q < 4; q++) x[q] = ((c[q] >> sh) & 1) ? l1 : l0;
    }
  } else {
    int64_t r = ~bs;
#pragma unroll
    for (int q = 0; q < 4; q++) x[q] = sample_at(s, o0 + lane + 64 * q, r);
  }
}

// Second half of fractional_bit_shift and fringe_stopping for one block
// whose spectrum X[k] (k = lane + 64 q) is in v: halve DC and Nyquist, drop
// negative frequencies, apply the ramp exp(-i (c + k l)) (one sincos per
// lane, then rotations by exp(-i 64 l)), transform back, fringe-rotate by
// phi + i dphi (the same sincos-and-rotate scheme) and write 256 samples.
template <class OutT>
__device__ __forceinline__ void finish_block(const DevStream &s, int64_t blk, float2 (&v)[4], float2 *buf,
                                             const float2 *tw, int lane, OutT *out) {
  constexpr int nd = 256;
  const auto *bcp = cmem(s.block_consts) + blk;
  struct {
    double c_term, l_term, phi, dphi;
    float amp, ramp_step[2], fringe_step[2];
  } bc;
  bc.c_term = bcp->c_term;
  bc.l_term = bcp->l_term;
  bc.phi = bcp->phi;
  bc.dphi = bcp->dphi;
  bc.amp = bcp->amp;
  bc.ramp_step[0] = bcp->ramp_step[0];
  bc.ramp_step[1] = bcp->ramp_step[1];
  bc.fringe_step[0] = bcp->fringe_step[0];
  bc.fringe_step[1] = bcp->fringe_step[1];
  {
    float sn, cs;
    sincosf((float)(bc.c_term + lane * bc.l_term), &sn, &cs);
    float2 e = make_float2(cs, -sn);
    const float2 step = make_float2(bc.ramp_step[0], -bc.ramp_step[1]);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = lane + 64 * q;
      if (k > nd / 2) {
        v[q] = make_float2(0.0f, 0.0f);
      } else {
        if (k == 0 || k == nd / 2) v[q] = make_float2(v[q].x * 0.5f, v[q].y * 0.5f);
        v[q] = cmul(e, v[q]);
      }
      e = cmul(e, step);
    }
  }
  fft256<+1>(v, buf, tw, lane);
  float sn, cs;
  sincosf((float)(bc.phi + lane * bc.dphi), &sn, &cs);
  float2 r = make_float2(cs, sn);
  const float2 rstep = make_float2(bc.fringe_step[0], bc.fringe_step[1]);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = lane + 64 * q;
    out[i] = bc.amp * (v[q].x * r.x + v[q].y * r.y);
    r = cmul(r, rstep);
  }
}

// Two N_d = 256 blocks (blk, blk + 1) of one stream by one wavefront
// (unpack + integer shift, fractional_bit_shift, fringe_stopping). The two
// real blocks share one complex forward FFT, z = a + i b, and are separated
// by A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i. `buf`
// is the wave's own LDS region; outb may be null when blk + 1 does not exist.
template <class OutT>
__device__ __forceinline__ void delay_block_pair(const DevStream &s, int64_t blk, int64_t n_blocks, float2 *buf,
                                                 const float2 *tw, int lane, OutT *outa, OutT *outb) {
  float xa[4], xb[4];
  load_block(s, blk, n_blocks, lane, xa);
  load_block(s, blk + 1, outb ? n_blocks : 0, lane, xb);
  float2 v[4];
#pragma unroll
  for (int q = 0; q < 4; q++) v[q] = make_float2(xa[q], xb[q]);
  fft256<-1>(v, buf, tw, lane);
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < 4; q++) buf[lane + 64 * q] = v[q];  // Z in natural order for Z[-k]
  wave_lds_sync();
  float2 va[4], vb[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int k = lane + 64 * q;
    const float2 zk = v[q], zn = buf[(256 - k) & 255];
    va[q] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
    vb[q] = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
  }
  finish_block(s, blk, va, buf, tw, lane, outa);
  if (outb) finish_block(s, blk + 1, vb, buf, tw, lane, outb);
}

// ---------------------------------------------------------------------------
// Register-resident Stockham FFT. T threads transform N = P * T complex
// points; thread t holds points t + i T (i < P) in registers. Each stage of
// radix R (R divides P) does P / R butterflies per thread straight from the
// registers, writes its outputs to LDS in Stockham order and reads the next
// stage's inputs back in the same t + i T layout, so a stage costs one LDS
// write and one read. The LDS buffer is padded by one element every P
// (lds_index), which keeps the stride-R writes of the early stages free of
// bank conflicts. Stages: radix P while it fits, then one radix N / P^k.
// ---------------------------------------------------------------------------
namespace {

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// Returns x unchanged but hides it from the optimiser, so that table loads
// indexed by it stay inside a loop instead of being hoisted out of it and
// holding registers across the loop body.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
// The same for a wave-uniform value, kept in a scalar register.
__device__ __forceinline__ int opaque_s(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Padded LDS position of FFT point i for a P-point-per-thread transform.
template <int P>
__device__ __forceinline__ int lds_index(int i) {
  return i + (i >> ilog2(P));
}

// LDS elements (float2) a padded N-point buffer needs.
template <int N, int P>
constexpr int fft_lds_size() {
  return N + N / P;
}

// R-point DFT in place, natural order in and out, exp(SIGN 2 pi i jk / R).
template <int R, int SIGN>
__device__ __forceinline__ void dft(float2 (&u)[R]);

template <>
__device__ __forceinline__ void dft<2, -1>(float2 (&u)[2]) {
  const float2 a = u[0];
  u[0] = cadd(a, u[1]);
  u[1] = csub(a, u[1]);
}
template <>
__device__ __forceinline__ void dft<4, -1>(float2 (&u)[4]) {
  radix4<-1>(u[0], u[1], u[2], u[3]);
}
// radix-2 layer, the W8^k twiddles, two 4-point DFTs for even and odd outputs
template <>
__device__ __forceinline__ void dft<8, -1>(float2 (&u)[8]) {
  constexpr float r = 0.70710678118654752f;
  float2 b[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    b[k] = cadd(u[k], u[k + 4]);
    c[k] = csub(u[k], u[k + 4]);
  }
  c[1] = cmul(c[1], make_float2(r, -r));
  c[2] = mul_i<-1>(c[2]);
  c[3] = cmul(c[3], make_float2(-r, -r));
  radix4<-1>(b[0], b[1], b[2], b[3]);
  radix4<-1>(c[0], c[1], c[2], c[3]);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    u[2 * m] = b[m];
    u[2 * m + 1] = c[m];
  }
}
// 4 x 4 decomposition: n = 4 n1 + n2, m = m1 + 4 m2; 4-point DFTs over n1,
// the W16^(n2 m1) twiddles, 4-point DFTs over n2.
template <>
__device__ __forceinline__ void dft<16, -1>(float2 (&u)[16]) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, r = 0.70710678118654752f;
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) radix4<-1>(u[n2], u[n2 + 4], u[n2 + 8], u[n2 + 12]);
  // u[n2 + 4 m1] *= W16^(n2 m1), W16 = exp(-2 pi i / 16)
  u[5] = cmul(u[5], make_float2(c1, -s1));
  u[9] = cmul(u[9], make_float2(r, -r));
  u[13] = cmul(u[13], make_float2(s1, -c1));
  u[6] = cmul(u[6], make_float2(r, -r));
  u[10] = mul_i<-1>(u[10]);
  u[14] = cmul(u[14], make_float2(-r, -r));
  u[7] = cmul(u[7], make_float2(s1, -c1));
  u[11] = cmul(u[11], make_float2(-r, -r));
  u[15] = cmul(u[15], make_float2(-c1, s1));
  float2 x[16];
#pragma unroll
  for (int m1 = 0; m1 < 4; m1++) {
    radix4<-1>(u[4 * m1], u[4 * m1 + 1], u[4 * m1 + 2], u[4 * m1 + 3]);
#pragma unroll
    for (int m2 = 0; m2 < 4; m2++) x[m1 + 4 * m2] = u[4 * m1 + m2];
  }
#pragma unroll
  for (int k = 0; k < 16; k++) u[k] = x[k];
}

// Twiddle sources: fft(m) = exp(-2 pi i m / N), split(k) = exp(-pi i k / N).
// TableTw reads full tables in global memory (N and N + 1 entries);
// LdsTw multiplies two short tables held in LDS, hi[j] = exp(-2 pi i 128 j /
// 2N) and lo[j] = exp(-2 pi i j / 2N) (j < 128), so that the FFT issues no
// global loads of its own.
struct TableTw {
  const float2 *__restrict__ n;
  const float2 *__restrict__ n2;
  __device__ __forceinline__ float2 fft(int m) const { return n[m]; }
  __device__ __forceinline__ float2 split(int k) const { return n2[k]; }
};
struct LdsTw {
  static constexpr bool kChain = false;
  const float2 *hi;
  const float2 *lo;
  // lo[0] = 1 exactly, so the product is exact where j is a multiple of 128;
  // it is skipped only where the compiler can prove that, and otherwise
  // always taken (no per-twiddle select)
  __device__ __forceinline__ float2 at(int j) const {
    const int l = j & 127;
    const float2 h = hi[j >> 7];
    if (__builtin_constant_p(l == 0) && l == 0) return h;
    return cmul(h, lo[l]);
  }
  __device__ __forceinline__ float2 fft(int m) const { return at(2 * m); }
  __device__ __forceinline__ float2 split(int k) const { return at(k); }
};
// The same tables, with the twiddles w^q (q = 2 .. R - 1) of a radix-R stage
// formed from w = W^(k step) by products of powers (at most four products
// deep): one table product per R points instead of R - 1. Their rounding
// error, a few units of 2^-24, is far below the FP16 / BF16 rounding of the
// spectra that take it.
struct LdsTwChain : LdsTw {
  static constexpr bool kChain = true;
};
// The twiddle source of a channeliser storing spectra of type SP: chained
// powers for packed FP16 / BF16, table products otherwise.
template <int SP>
__device__ __forceinline__ auto channeliser_tw(const float2 *hi, const float2 *lo) {
  std::conditional_t<SP == SFXG_PREC_FP16 || SP == SFXG_PREC_BF16, LdsTwChain, LdsTw> t;
  t.hi = hi;
  t.lo = lo;
  return t;
}

// w[q] = w[1]^q for q < R, from w[1]: squares for the powers of two, one
// product for the others (w[a + b] = w[a] w[b], a the largest power of two).
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 (&w)[R]) {
#pragma unroll
  for (int q = 2; q < R; q++) {
    const int a = 1 << (31 - __builtin_clz(q));
    w[q] = a == q ? cmul(w[q / 2], w[q / 2]) : cmul(w[a], w[q - a]);
  }
}

// One radix-R stage over the registers of all threads; ns = product of the
// earlier radices. Outputs go to LDS.
template <int N, int P, int R, class TW>
__device__ __forceinline__ void fft_stage(float2 (&v)[P], float2 *lds, const TW &tw, int ns, int tid) {
  constexpr int T = N / P, B = P / R;
  const int lns = __builtin_ctz(ns);
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int j = tid + b * T;
    const int k = j & (ns - 1);
    float2 u[R];
#pragma unroll
    for (int q = 0; q < R; q++) u[q] = v[b + q * B];
    if (ns > 1) {
      const int step = N / (ns * R);
      if constexpr (TW::kChain) {
        float2 w[R];
        w[1] = tw.fft(k * step);
        twiddle_powers<R>(w);
#pragma unroll
        for (int q = 1; q < R; q++) u[q] = cmul(u[q], w[q]);
      } else {
#pragma unroll
        for (int q = 1; q < R; q++) {
          u[q] = cmul(u[q], tw.fft(q * k * step));
          // at most four twiddles in flight, so their loads do not hold
          // registers needed by data the caller keeps live across the FFT
          if (q % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    dft<R, -1>(u);
    const int idx = ((j >> lns) << (lns + ilog2(R))) + k;
#pragma unroll
    for (int r = 0; r < R; r++) lds[lds_index<P>(idx + r * ns)] = u[r];
  }
}

// Forward FFT (exp(-2 pi i jk / N)) of the points in v, result in padded LDS
// (point k at lds_index<P>(k)). tid is threadIdx.x; callers that loop pass
// it through opaque() so that the per-thread index arithmetic is redone in
// each iteration instead of being hoisted into registers held across the
// loop. The caller guarantees that no thread still reads `lds` on entry; the
// function ends with a barrier.
template <int N, int P, class TW>
__device__ __forceinline__ void reg_fft(float2 (&v)[P], float2 *lds, const TW &tw, int tid) {
  constexpr int T = N / P;
  constexpr int NS = ilog2(N) / ilog2(P);               // full radix-P stages
  constexpr int F = 1 << (ilog2(N) - NS * ilog2(P));    // last radix, 1 if none
  static_assert(P * T == N && (P & (P - 1)) == 0 && P >= 2 && P <= 16, "bad FFT shape");
  static_assert(P % F == 0, "the last radix must divide the points per thread");
  int ns = 1;
#pragma unroll
  for (int st = 0; st < NS; st++) {
    fft_stage<N, P, P, TW>(v, lds, tw, ns, tid);
    ns *= P;
    __syncthreads();
    if (st + 1 < NS || F > 1) {
#pragma unroll
      for (int i = 0; i < P; i++) v[i] = lds[lds_index<P>(tid + i * T)];
      __syncthreads();
    }
  }
  if constexpr (F > 1) {
    fft_stage<N, P, F, TW>(v, lds, tw, ns, tid);
    __syncthreads();
  }
}

// Split of the N-point complex transform of 2N packed reals into bins
// k = 0..N of the real transform: X_k = E_k + W_2N^k O_k.
// Bins k and N - k come from the same two points: E_{N-k} = conj E_k,
// O_{N-k} = conj O_k and W_2N^(N-k) = -conj W_2N^k, so
// X_{N-k} = conj(E_k - W_2N^k O_k). Thread tid takes the pairs
// k = tid + j T (j < P / 2, so k < N / 2, with k = 0 pairing bins 0 and N)
// and thread 0 also bin N / 2: two LDS reads and one twiddle per two bins,
// and a fixed number of stores per thread, so that the compiler can wait for
// loads issued before them without waiting for the stores too.
template <int N, int P, int T, class TW, class OUT>
__device__ __forceinline__ void real_split(const float2 *lds, const TW &tw, const OUT &out, int tid) {
  static_assert(N == P * T && P % 2 == 0, "pairs of bins per thread");
  auto eo = [&](int k, float2 &e, float2 &o) {
    const float2 zk = lds[lds_index<P>(k)], zn = lds[lds_index<P>((N - k) & (N - 1))];
    e = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
    o = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
  };
#pragma unroll 1
  for (int j = 0; j < P / 2; j++) {
    const int k = tid + j * T;
    float2 e, o;
    eo(k, e, o);
    const float2 wo = cmul(tw.split(k), o);
    out.put((uint32_t)k, cadd(e, wo));
    out.put((uint32_t)(N - k), make_float2(e.x - wo.x, -(e.y - wo.y)));
    __builtin_amdgcn_sched_barrier(0);  // one pair's loads in flight at a time
  }
  if (tid == 0) {
    float2 e, o;
    eo(N / 2, e, o);
    out.put((uint32_t)(N / 2), cadd(e, cmul(tw.split(N / 2), o)));
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Delay correction with 16 points per lane (k_delay): a wavefront works on
// four pairs of N_d = 256 blocks at once, one pair per 16-lane group. A
// 256-point transform is two radix-16 stages, n = 16 n1 + n2, k = k1 + 16 k2:
//   X[k1 + 16 k2] = sum_n2 W256^(n2 k1) W16^(n2 k2) sum_n1 x[16 n1 + n2] W16^(n1 k1).
// Lane l of a group (l = n2) holds x[16 n1 + l] for n1 = 0..15; the first
// stage is a 16-point DFT in registers, times W256^(l k1); one transpose
// through the group's LDS region gives lane l (= k1) the Y[l][n2] of every
// n2, and the second stage leaves X[l + 16 k2] in register k2 of lane l.
// Against the four radix-4 stages of fft256 (4 points per lane, 3 LDS
// exchanges) this takes about 40 % fewer VALU operations per point.
// ---------------------------------------------------------------------------
namespace {

constexpr int G16_ROW = 17;              // padded row of a group's 16 x 16 exchange
constexpr int G16_LDS = 16 * G16_ROW;    // float2 elements of one group's region

// 16-point DFT in registers, natural order in and out, exp(SIGN 2 pi i jk / 16).
template <int SIGN>
__device__ __forceinline__ void dft16s(float2 (&u)[16]) {
  if constexpr (SIGN > 0) {
#pragma unroll
    for (int k = 0; k < 16; k++) u[k].y = -u[k].y;
  }
  dft<16, -1>(u);
  if constexpr (SIGN > 0) {
#pragma unroll
    for (int k = 0; k < 16; k++) u[k].y = -u[k].y;
  }
}

// v[n1] = x[16 n1 + l] on entry, X[l + 16 k2] in v[k2] on exit; tw holds
// exp(-2 pi i l k1 / 256) at 16 k1 + l (lane-major: the 16 lanes of a group
// read 16 consecutive entries, free of bank conflicts); buf is the group's
// LDS region.
template <int SIGN>
__device__ __forceinline__ void fft256_g16(float2 (&v)[16], float2 *buf, const float2 *tw, int l) {
  dft16s<SIGN>(v);
#pragma unroll
  for (int k1 = 1; k1 < 16; k1++) v[k1] = cmul(v[k1], twiddle<SIGN>(tw, 16 * k1 + l));
  wave_lds_sync();
#pragma unroll
  for (int k1 = 0; k1 < 16; k1++) buf[k1 * G16_ROW + l] = v[k1];
  wave_lds_sync();
#pragma unroll
  for (int n2 = 0; n2 < 16; n2++) v[n2] = buf[l * G16_ROW + n2];
  dft16s<SIGN>(v);
}

// Payload words of a block that lies in one payload run (bs >= 0, its
// first payload sample): the aligned 32-bit words that hold its 256 samples,
// 17 at most for 2-bit data and 9 for 1-bit; lane l of the group loads word
// l, and word 16 + l where the block reaches it (lane 0, 2-bit). Every word
// loaded holds at least one byte of the block, so it lies in a page of the
// payload's allocation. A block past the end or across a run boundary loads
// nothing here.
struct Words16 {
  uint32_t w0, w1;
};
// field offset of the block's first sample in its first word (samples of
// 2 or 1 bits, least significant first)
__device__ __forceinline__ int words16_f0(const DevStream &s, int64_t bs, uintptr_t &a) {
  const int lb = s.bits == 2 ? 2 : 3;  // log2 samples per byte
  a = reinterpret_cast<uintptr_t>(s.payload) + (uintptr_t)(bs >> lb);
  return (int)((a & 3) << lb) + (int)(bs & ((1 << lb) - 1));
}
__device__ __forceinline__ Words16 load_words16(const DevStream &s, int64_t bs, bool fast, int l) {
  Words16 r{0u, 0u};
  if (!fast) return r;
  uintptr_t a;
  const int f0 = words16_f0(s, bs, a);
  const int last = (f0 + 255) >> (s.bits == 2 ? 4 : 5);
  const auto *w = gmem(reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3));
  if (l <= last) r.w0 = w[l];
  if (l + 16 <= last) r.w1 = w[l + 16];
  return r;
}

// Samples 16 n1 + l (n1 < 16) of block blk: an in-run block decodes them
// from its payload words, staged in the group's LDS (wl), with the table in
// registers; a block that crosses a run boundary or invalid data goes sample
// by sample; a block past the end is zeros. bs: the block's block_src entry.
__device__ __forceinline__ void block16(const DevStream &s, int64_t blk, bool valid, int64_t bs, const uint32_t *wl,
                                        int l, float (&x)[16]) {
  constexpr int nd = 256;
  if (!valid) {
#pragma unroll
    for (int q = 0; q < 16; q++) x[q] = 0.0f;
    return;
  }
  if (bs >= 0) {
    uintptr_t a;
    const int f = words16_f0(s, bs, a) + l;
    if (s.bits == 2) {
      const int dq = f >> 4, sh = 2 * (f & 15);
      const float l0 = s.lut[0], l1 = s.lut[1], l2 = s.lut[2], l3 = s.lut[3];
#pragma unroll
      for (int q = 0; q < 16; q++) x[q] = lut4((wl[q + dq] >> sh) & 3, l0, l1, l2, l3);
    } else {
      const float l0 = s.lut[0], l1 = s.lut[1];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int i = f + 16 * q;
        x[q] = ((wl[i >> 5] >> (i & 31)) & 1) ? l1 : l0;
      }
    }
  } else {
    int64_t r = ~bs;
#pragma unroll
    for (int q = 0; q < 16; q++) x[q] = sample_at(s, blk * nd + 16 * q + l, r);
  }
}

// finish_block for a 16-lane group: v[k2] = X[k] of bin k = l + 16 k2;
// halve DC and Nyquist, drop negative frequencies, apply the ramp
// exp(-i (c + k l_term)) (sincos of bin l, then rotations by exp(-i 16
// l_term)), transform back, fringe-rotate (the same scheme) and write the
// 256 samples when `write` is set.
// The per-block constants finish_block16 takes, loaded by the caller before
// any store of the iteration (the vector memory counter is in order, so
// waiting for loads issued after stores would also wait for the stores).
// The FP64 phases of lane l are rounded to FP32 here, so that only FP32
// values stay live across the transforms.
struct BlockC16 {
  float ramp_arg, fringe_arg;  // c_term + l l_term, phi + l dphi
  float amp, ramp16c, ramp16s, fringe16c, fringe16s;
};
__device__ __forceinline__ BlockC16 block_c16(const DevStream &s, int64_t blk, int l) {
  const auto *bcp = cmem(s.block_consts) + blk;
  return BlockC16{(float)(bcp->c_term + l * bcp->l_term), (float)(bcp->phi + l * bcp->dphi), bcp->amp,
                  bcp->ramp_step16[0], bcp->ramp_step16[1], bcp->fringe_step16[0], bcp->fringe_step16[1]};
}

template <bool H>
__device__ __forceinline__ void finish_block16(const BlockC16 &bc, float2 (&v)[16], float2 *buf,
                                               const float2 *tw, int l, void *out, bool write) {
  constexpr int nd = 256;
  const float amp = bc.amp;
  // the FP16 series (H) takes the hardware sine and cosine (absolute error
  // near 1e-6 for these arguments of a few radians, far below its 2^-12
  // rounding); the FP32 series the full-precision sincosf
  auto sc = [](float x, float *sn, float *cs) {
    if constexpr (H) __sincosf(x, sn, cs);
    else sincosf(x, sn, cs);
  };
  {
    float sn, cs;
    sc(bc.ramp_arg, &sn, &cs);
    // H: the pair split's factor 1/2 rides on the ramp (exact)
    constexpr float rs = H ? 0.5f : 1.0f;
    float2 e = make_float2(rs * cs, -rs * sn);
    const float2 step = make_float2(bc.ramp16c, -bc.ramp16s);
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int k = l + 16 * q;
      if (k > nd / 2) {
        v[q] = make_float2(0.0f, 0.0f);
      } else {
        if (k == 0 || k == nd / 2) v[q] = make_float2(v[q].x * 0.5f, v[q].y * 0.5f);
        v[q] = cmul(e, v[q]);
      }
      e = cmul(e, step);
    }
  }
  fft256_g16<+1>(v, buf, tw, l);
  float sn, cs;
  sc(bc.fringe_arg, &sn, &cs);
  float2 r = make_float2(cs, sn);
  const float2 rstep = make_float2(bc.fringe16c, bc.fringe16s);
  if constexpr (H) r = make_float2(amp * cs, amp * sn);  // the amplitude rides on the phasor
  if constexpr (!H) {
    auto *o = gmem(static_cast<float *>(out));
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (write) o[l + 16 * q] = amp * (v[q].x * r.x + v[q].y * r.y);
      r = cmul(r, rstep);
    }
  } else {
    // FP16 samples (rounded to nearest even), whole 32-bit words per lane:
    // for the rows q0 = 2 j and q1 = 2 j + 1, an even lane l stores samples
    // 16 q0 + l and 16 q0 + l + 1, an odd lane 16 q1 + l - 1 and 16 q1 + l;
    // neighbouring lanes swap the other half (DPP quad_perm [1, 0, 3, 2])
    auto *o = gmem(static_cast<uint32_t *>(out));
    const bool odd = l & 1;
    const int wb = (l >> 1) + (odd ? 8 : 0);  // word 16 j + wb
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float y0 = v[2 * j].x * r.x + v[2 * j].y * r.y;
      r = cmul(r, rstep);
      const float y1 = v[2 * j + 1].x * r.x + v[2 * j + 1].y * r.y;
      r = cmul(r, rstep);
      const uint32_t h0 = __builtin_bit_cast(uint16_t, (_Float16)y0), h1 = __builtin_bit_cast(uint16_t, (_Float16)y1);
      const uint32_t recv = (uint32_t)__builtin_amdgcn_mov_dpp((int)(odd ? h0 : h1), 0xB1, 0xF, 0xF, false);
      const uint32_t w = odd ? (recv | (h1 << 16)) : (h0 | (recv << 16));
      if (write) o[16 * j + wb] = w;
    }
  }
}

}  // namespace

// k_delay: each wavefront takes four pairs of blocks at a time (grid-stride
// over the chunk's blocks), the real series to HBM (unfused path, used for
// N_c above the fused F-engine's LDS budget, before the MFMA channeliser and
// for verification) as FP32 or, H, FP16. The two real blocks of a pair
// share one complex forward transform, z = a + i b.
template <bool H>
__global__ __launch_bounds__(256, 3) void k_delay(const DevStream *__restrict__ streams, DevJob job,
                                               const float2 *__restrict__ tw_g) {
  __shared__ float2 lds[4][4][G16_LDS];
  __shared__ float2 tw[256];  // exp(-2 pi i l k1 / 256) at 16 k1 + l
  for (int i = threadIdx.x; i < 256; i += 256) tw[i] = tw_g[((i & 15) * (i >> 4)) & 255];
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int g = lane >> 4, l = lane & 15;
  float2 *buf = lds[wave][g];
  const DevStream &s = streams[blockIdx.y];
  const int64_t end = job.blk_end < job.n_blocks ? job.blk_end : job.n_blocks;
  // block_src of this group's pair two iterations ahead and its payload
  // words one iteration ahead, so that nothing at the top of an iteration
  // waits on a load of its own
  const int64_t step = (int64_t)gridDim.x * 32;
  auto src_of = [&](int64_t b) { return b < end ? cmem(s.block_src)[b] : (int64_t)0; };
  auto fast = [&](int64_t b, int64_t bs) { return b < end && bs >= 0; };
  int64_t blk = job.blk_base + ((int64_t)blockIdx.x * 4 + wave) * 8;
  int64_t bsa = src_of(blk + 2 * g), bsb = src_of(blk + 2 * g + 1);
  int64_t bsa_n = src_of(blk + step + 2 * g), bsb_n = src_of(blk + step + 2 * g + 1);
  Words16 wa = load_words16(s, bsa, fast(blk + 2 * g, bsa), l);
  Words16 wb = load_words16(s, bsb, fast(blk + 2 * g + 1, bsb), l);
  // complete before the loop as in it (before each pair's stores), so that
  // the loop top waits for nothing
  asm volatile("" ::"v"(wa.w0), "v"(wa.w1), "v"(wb.w0), "v"(wb.w1), "v"(bsa_n), "v"(bsb_n));
  uint32_t *ws = reinterpret_cast<uint32_t *>(buf);  // staged words: block a at 0, block b at 32
  for (; blk < end; blk += step) {
    const int64_t ba = blk + 2 * g;  // this group's pair (ba, ba + 1)
    const BlockC16 ca = block_c16(s, ba < end ? ba : 0, l), cb = block_c16(s, ba + 1 < end ? ba + 1 : 0, l);
    // this pair's words (loaded one iteration ago) through the group's LDS;
    // the next pair's loads are issued only after them, so that waiting for
    // these does not wait for those (the vector memory counter is in order)
    wave_lds_sync();  // the group's last reads of buf (previous pair) are done
    ws[l] = wa.w0;
    ws[32 + l] = wb.w0;
    if (l == 0) {
      ws[16] = wa.w1;
      ws[48] = wb.w1;
    }
    wave_lds_sync();
    float xa[16], xb[16];
    block16(s, ba, ba < end, bsa, ws, l, xa);
    block16(s, ba + 1, ba + 1 < end, bsb, ws + 32, l, xb);
    wa = load_words16(s, bsa_n, fast(ba + step, bsa_n), l);
    wb = load_words16(s, bsb_n, fast(ba + step + 1, bsb_n), l);
    bsa = bsa_n;
    bsb = bsb_n;
    bsa_n = src_of(ba + 2 * step);
    bsb_n = src_of(ba + 2 * step + 1);
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = make_float2(xa[q], xb[q]);
    fft256_g16<-1>(v, buf, tw, l);
    // Z[-k] of bin k = l + 16 q: lane (16 - l) % 16, register 15 - q
    // (l > 0) or (16 - q) % 16 (l = 0)
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < 16; q++) buf[l * G16_ROW + q] = v[q];
    wave_lds_sync();
    const int ln = (16 - l) & 15;
    float2 va[16], vb[16];
    constexpr float hf = H ? 1.0f : 0.5f;  // H: the 1/2 is applied with the ramp
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int rq = l ? 15 - q : (16 - q) & 15;
      const float2 zk = v[q], zn = buf[ln * G16_ROW + rq];
      va[q] = make_float2(hf * (zk.x + zn.x), hf * (zk.y - zn.y));
      vb[q] = make_float2(hf * (zk.y + zn.y), -hf * (zk.x - zn.x));
    }
    // the next pair's words and block_src complete here, before this pair's
    // series stores, so that no later wait for them also waits for stores
    asm volatile("" ::"v"(wa.w0), "v"(wa.w1), "v"(wb.w0), "v"(wb.w1), "v"(bsa_n), "v"(bsb_n));
    constexpr int64_t bb = H ? 512 : 1024;  // bytes per block
    char *base = reinterpret_cast<char *>(s.time_series) + (ba - job.blk_base) * bb;
    finish_block16<H>(ca, va, buf, tw, l, base, ba < end);
    finish_block16<H>(cb, vb, buf, tw, l, base + bb, ba + 1 < end);
  }
}

// ---------------------------------------------------------------------------
// k_channelise: one workgroup per (segment, stream), N / P threads. The 2N
// windowed reals of the segment are packed as N complex points, loaded from
// HBM straight into the FFT's registers, transformed and split into the
// N + 1 bins of the real transform.
// ---------------------------------------------------------------------------
template <int N, int P, int SP>
__global__ __launch_bounds__(N / P) void k_channelise(const DevStream *__restrict__ streams, DevJob job,
                                                      const float *__restrict__ window,
                                                      const float2 *__restrict__ tw_split) {
  constexpr int T = N / P, NHI = 2 * N / 128;
  __shared__ float2 lds[fft_lds_size<N, P>()];
  __shared__ float2 thi[NHI], tlo[128];
  for (int i = threadIdx.x; i < NHI + 128; i += T) {
    const float2 t = tw_split[i];
    if (i < NHI) thi[i] = t; else tlo[i - NHI] = t;
  }
  const DevStream &s = streams[blockIdx.y];
  const int64_t seg = blockIdx.x;
  const float2 *x = reinterpret_cast<const float2 *>(s.time_series + seg * job.seg_step);
  const float2 *w = reinterpret_cast<const float2 *>(window);
  const bool none = job.window == SFXG_WINDOW_NONE;
  float2 v[P];
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int m = threadIdx.x + i * T;
    float2 z;
    if (none && m >= N / 2) {
      z = make_float2(0.0f, 0.0f);
    } else {
      z = wmul(gload2(x + m), w[m]);
    }
    if (s.flip) z.y = -z.y;
    v[i] = z;
  }
  __syncthreads();  // twiddle tables
  const auto tw = channeliser_tw<SP>(thi, tlo);
  reg_fft<N, P>(v, lds, tw, threadIdx.x);
  real_split<N, P, T>(lds, tw, SpecOut<SP>(s.spectra, seg * job.spectra_stride, job.spec_scale), threadIdx.x);
}

// ---------------------------------------------------------------------------
// k_channelise_run: the channeliser for large N (one workgroup per CU by
// LDS), as persistent workgroups that each walk a run of consecutive
// segments of one stream. Overlapping segments share half their samples, so
// the second half of each segment stays in registers as the next one's first
// half and only N new reals are read per segment; they are prefetched one
// segment ahead. The FFT takes its twiddles from two short LDS tables, so
// the only global loads in flight during it are the prefetch.
// ---------------------------------------------------------------------------
template <int N, int P, int SP>
__global__ __launch_bounds__(N / P) void k_channelise_run(const DevStream *__restrict__ streams, DevJob job,
                                                          const float *__restrict__ window,
                                                          const float2 *__restrict__ tw_split, int segs_per_wg) {
  constexpr int T = N / P, H = P / 2, NHI = 2 * N / 128;
  __shared__ float2 lds[fft_lds_size<N, P>()];
  __shared__ float2 thi[NHI], tlo[128];
  const DevStream &s = streams[blockIdx.y];
  const int64_t sa = (int64_t)blockIdx.x * segs_per_wg;
  int64_t sb = sa + segs_per_wg;
  if (sb > job.n_segments) sb = job.n_segments;
  if (sa >= sb) return;  // uniform over the workgroup
  for (int i = threadIdx.x; i < NHI + 128; i += T) {
    const float2 t = tw_split[i];
    if (i < NHI) thi[i] = t; else tlo[i - NHI] = t;
  }
  const auto tw = channeliser_tw<SP>(thi, tlo);
  const bool none = job.window == SFXG_WINDOW_NONE;
  const float2 *w2 = reinterpret_cast<const float2 *>(window);
  // complex point m of segment seg is the pair of reals 2m, 2m + 1 at
  // seg * N reals; the segment's second half is the next one's first half.
  // A segment's windowed points are formed right after the previous
  // segment's FFT, before its spectrum stores (loads and stores share one
  // in-order counter), from the carried raw half and the prefetched new
  // reals; the loop top then waits for nothing.
  const float2 *ts = reinterpret_cast<const float2 *>(s.time_series);
  const bool flip = s.flip;
  auto win = [flip](float2 x, float2 w) {
    float2 z = wmul(x, w);
    if (flip) z.y = -z.y;
    return z;
  };
  auto off = [](int tid, int i) { return (uint32_t)(tid + i * T) * (uint32_t)sizeof(float2); };
  float2 lo_w[H], hi_w[H], raw[H];
  {
    const float2 *seg0 = ts + sa * (N / 2);
#pragma unroll
    for (int i = 0; i < H; i++) {
      lo_w[i] = win(gload2(seg0, off(threadIdx.x, i)), gload2(w2, off(threadIdx.x, i)));
      hi_w[i] = make_float2(0.0f, 0.0f);
      if (!none) {
        raw[i] = gload2(seg0 + N / 2, off(threadIdx.x, i));
        hi_w[i] = win(raw[i], gload2(w2, off(threadIdx.x, i + H)));
      }
    }
  }
  __syncthreads();  // twiddle tables
  for (int64_t seg = sa; seg < sb; seg++) {
    const int tid = opaque(threadIdx.x);
    float2 v[P];
#pragma unroll
    for (int i = 0; i < H; i++) {
      v[i] = lo_w[i];
      v[i + H] = hi_w[i];
    }
    // the reals the next segment adds (this segment's again after the last
    // one, so that the loads need no branch)
    const float2 *nseg = ts + (seg + 1 < sb ? seg + 1 : seg) * (N / 2) + (none ? 0 : N / 2);
    float2 nx[H];
#pragma unroll
    for (int i = 0; i < H; i++) nx[i] = gload2(nseg, off(tid, i));
    reg_fft<N, P>(v, lds, tw, tid);
    // branch-free, so that the window loads issue together (without a
    // window the second half is zero and the first takes the new reals)
    const int tid2 = opaque(threadIdx.x);
    float2 wlo[H], whi[H];
#pragma unroll
    for (int i = 0; i < H; i++) {
      wlo[i] = gload2(w2, off(tid2, i));
      whi[i] = gload2(w2, off(tid2, i + H));
    }
#pragma unroll
    for (int i = 0; i < H; i++) {
      lo_w[i] = win(none ? nx[i] : raw[i], wlo[i]);
      hi_w[i] = none ? make_float2(0.0f, 0.0f) : win(nx[i], whi[i]);
      raw[i] = nx[i];
    }
    real_split<N, P, T>(lds, tw, SpecOut<SP>(s.spectra, seg * job.spectra_stride, job.spec_scale), tid);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// k_chan_mfma: the channeliser of the mixed-precision X-steps (spectra
// stored as packed FP16 or BF16) for N_c = 1024 NT (NT = 1, 2, 4, 8, 16), its
// transform on MFMA. The 2 N_c windowed reals of a segment are N = N_c
// complex points z[n],
//   n = NT (32 m2 + m3) + n1,  q = k2 + 32 k3,  k = q + 1024 k1:
//   V[n1][q]      = sum_m3 W32^(m3 k3) W1024^(m3 k2) sum_m2 z[n] W32^(m2 k2)
//   Z[q + 1024 k1] = sum_n1 W_NT^(n1 k1) W_N^(n1 q) V[n1][q].
// Each of the NT residues n1 is a 32 x 32 tile (rows m3, columns m2); a
// pass holds 16 tiles, those of 16 / NT segments. Stage B1 (over m2) takes
// the tile as the A operand of v_mfma_f32_32x32x16_f16 (K = the real and
// imaginary parts of 32 m2) against the real form of the 32-point DFT
// matrix; its 32 x 32 FP32 result, times W1024^(m3 k2) and rounded to FP16,
// is the B operand of stage B2 (over m3) as it lies in the registers: a lane
// holds the result rows 8 g + 4 h + j (h = lane / 32), so B1 orders the K
// slots of its operands the same way (cm_kslot) and the two stages exchange
// nothing. The DFT matrix is symmetric, so one set of fragments serves as
// B1's B operand and B2's A operand. The last stage (NT points over n1,
// twiddles W_N^(n1 q)) and the split of the N-point transform into the N + 1
// bins of the real one run in FP32 registers: the thread of row q also takes
// row 1024 - q, which holds the bins N - k.
// Phases per pass: the windowed points into the tiles' LDS operands (16-byte
// chunks, XOR-swizzled); B1 and B2 for two tiles per wave, whose V (packed
// FP16 pairs) goes where those two tiles lay; then the last stage and the
// spectrum stores. 64 KiB of operands, the window table and per-thread
// constants: about 152 KiB of LDS at N_c = 16384 and up to 256 VGPRs, one
// workgroup per CU.
// The operands are FP16 (unit roundoff 2^-11) with FP32 accumulation: the
// series (FP16 from the delay stage, or FP32 rounded on load), the windowed
// points, the B1 results and V are rounded once each, which puts the
// transform's error near that of the FP16 spectra themselves, within the
// mixed-precision visibility bar.
// ---------------------------------------------------------------------------
namespace {

constexpr int CM_T = 512;
typedef _Float16 cm_h8 __attribute__((ext_vector_type(8)));
typedef float cm_f16 __attribute__((ext_vector_type(16)));

// Row of the B1 / B2 result that K slot 8 h + e of a 16-wide K slice holds.
__device__ __forceinline__ constexpr int cm_kslot(int h, int e) { return 4 * h + (e & 3) + 8 * (e >> 2); }

// LDS position (16-byte units) of chunk c (K slice s = c / 2, lane half
// h = c % 2) of row m3 of tile n1.
__device__ __forceinline__ int cm_chunk(int n1, int m3, int c) { return n1 * 256 + m3 * 8 + (c ^ ((m3 + n1) & 7)); }

// exp(-2 pi i x / n), evaluated in FP64 and rounded once
__device__ __forceinline__ float2 cm_w(int x, int n) {
  double sn, cs;
  sincospi(2.0 * (double)(x % n) / (double)n, &sn, &cs);
  return make_float2((float)cs, (float)-sn);
}

// cos(2 pi k / 32) for k = 0..8
__device__ __forceinline__ constexpr float cm_c32(int k) {
  constexpr float c[9] = {1.0f, 0.98078528040323044913f, 0.92387953251128675613f, 0.83146961230254523708f,
                          0.70710678118654752440f, 0.55557023301960222474f, 0.38268343236508977173f,
                          0.19509032201612826785f, 0.0f};
  return c[k];
}
// exp(-2 pi i k / 32) for k = 0..15 (compile-time k)
__device__ __forceinline__ constexpr float2 cm_w32(int k) {
  return make_float2(k <= 8 ? cm_c32(k) : -cm_c32(16 - k), -cm_c32(k <= 8 ? 8 - k : k - 8));
}

__device__ __forceinline__ uint32_t cm_pack(float re, float im) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)re) |
         ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)im) << 16);
}
__device__ __forceinline__ float2 cm_unpack(uint32_t w) {
  return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)),
                     (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)));
}

// Row q of segment sig of the pass from LDS: tile tau = NT sig + n1 lies in
// wave tau / 2's region, which holds the pairs (tiles 2 w, 2 w + 1) of every
// q, 8 bytes each.
template <int NT>
__device__ __forceinline__ void cm_read_rows(const uint2 *vlds, int sig, int q, float2 (&z)[NT]) {
  if constexpr (NT == 1) {
    const uint2 u = vlds[(sig >> 1) * 1024 + q];
    z[0] = cm_unpack(sig & 1 ? u.y : u.x);
  } else {
#pragma unroll
    for (int w = 0; w < NT / 2; w++) {
      const uint2 u = vlds[(sig * NT / 2 + w) * 1024 + q];
      z[2 * w] = cm_unpack(u.x);
      z[2 * w + 1] = cm_unpack(u.y);
    }
  }
}

// z[n1] *= w^n1 for n1 = 1..NT - 1 (running powers), then the NT-point DFT:
// z[k1] = Z[q + 1024 k1].
template <int NT>
__device__ __forceinline__ void cm_last_stage(float2 (&z)[NT], float2 w) {
  if constexpr (NT > 1) {
    float2 p = w;
#pragma unroll
    for (int i = 1; i < NT; i++) {
      z[i] = cmul(z[i], p);
      if (i < NT - 1) p = cmul(p, w);
    }
    dft<NT, -1>(z);
  }
}

// The last stage of rows q and 1024 - q together: the second row's twiddles
// are W_N^(n1 (1024 - q)) = W_NT^n1 conj(W_N^(n1 q)), so one running power
// serves both (conjugated for the second row), and the factor W_NT^n1 shifts
// the second row's NT-point result by one bin: on return z1[k1] holds
// Z[(1024 - q) + 1024 ((k1 + NT - 1) mod NT)].
template <int NT>
__device__ __forceinline__ void cm_last_stage2(float2 (&z0)[NT], float2 (&z1)[NT], float2 w) {
  if constexpr (NT > 1) {
    float2 p = w;
#pragma unroll
    for (int i = 1; i < NT; i++) {
      z0[i] = cmul(z0[i], p);
      z1[i] = cmul(z1[i], make_float2(p.x, -p.y));
      if (i < NT - 1) p = cmul(p, w);
    }
    dft<NT, -1>(z0);
    dft<NT, -1>(z1);
  }
}

// Real-transform bins from Z[k] and Z[N - k]: X[k] = E + W_2N^k O and
// X[N - k] = conj(E - W_2N^k O), E = (Z[k] + conj Z[N - k]) / 2,
// O = -i (Z[k] - conj Z[N - k]) / 2. The factor 1/2 and the spectra's
// power-of-two scale are already in Z (folded into the twiddles between the
// MFMA stages: exact, as scaling afterwards would be).
__device__ __forceinline__ void cm_split(float2 zk, float2 zn, float2 w, float2 &xk, float2 &xn) {
  const float2 e = make_float2(zk.x + zn.x, zk.y - zn.y);
  const float2 o = make_float2(zk.y + zn.y, zn.x - zk.x);
  const float2 wo = cmul(w, o);
  xk = cadd(e, wo);
  xn = make_float2(e.x - wo.x, -(e.y - wo.y));
}

// Bin k of a packed spectrum row: both parts of the (already scaled) value
// rounded to nearest even in the operand type, one conversion of the pair.
template <int SP>
__device__ __forceinline__ void cm_put(uint32_t *row, uint32_t k, float2 v) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  uint32_t wd;
  if constexpr (SP == SFXG_PREC_BF16) {
    typedef __bf16 b2v __attribute__((ext_vector_type(2)));
    wd = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){v.x, v.y}, b2v));
  } else {
    typedef _Float16 h2v __attribute__((ext_vector_type(2)));
    wd = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){v.x, v.y}, h2v));
  }
  *((__attribute__((address_space(1))) uint32_t *)row + k) = wd;
}

}  // namespace

// The window as a table of FP32 pairs: entry e < N / 2 holds the weights
// of reals 2 e and 2 e + 1 of the 2 N (NONE 1 on the first half, RECT 1 on
// the middle half, and the cosine family of theta_i = 2 pi i / (2 N - 1),
// delay_correction.cc:299-341, through s = sin(theta_i / 2): HANN
// 0.5 (1 - cos theta) = s^2, HAMMING 0.54 - 0.46 cos theta = 0.08 + 0.92 s^2,
// COS s), each evaluated in FP64 and rounded once. Every window but NONE is
// symmetric, w(2 N - 1 - i) = w(i), so point n >= N / 2 (reals 2 n, 2 n + 1)
// takes entry N - 1 - n with its halves swapped; NONE's second half is zero.
template <int N>
__device__ __forceinline__ void cm_window_table(float2 *wt, int window) {
  for (int e = threadIdx.x; e < N / 2; e += CM_T) {
    float w[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int i = 2 * e + b;
      double v = 1.0;
      if (window == SFXG_WINDOW_RECT) {
        v = i >= N / 2 ? 1.0 : 0.0;
      } else if (window != SFXG_WINDOW_NONE) {
        const double sn = sinpi((double)i / (double)(2 * N - 1));
        v = window == SFXG_WINDOW_HANN ? sn * sn : window == SFXG_WINDOW_HAMMING ? 0.08 + 0.92 * sn * sn : sn;
      }
      w[b] = (float)v;
    }
    wt[e] = make_float2(w[0], w[1]);
  }
}

// Phase 0 of k_chan_mfma: thread t = 16 m3 + tau turns its 32 points xp (FP16
// pairs; point i = 16 g + 8 hh + e at m2 = 16 g + cm_kslot(hh, e), point
// n = NT (32 m2 + m3) + tau % NT of its segment) into the FP16 K chunks of
// row m3 of tile tau: the FP32 product with the window pair (exact for the
// FP16 point) rounded once to FP16, the sideband flip as the sign of the
// imaginary half, then the real halves and the imaginary halves of 8 points
// regrouped into the two chunks.
template <int NT>
__device__ __forceinline__ void cm_phase0(uint4 *lds, const uint32_t (&xp)[32], const float2 *wt,
                                          uint32_t flipmask, bool zero_hi) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  typedef _Float16 h2v __attribute__((ext_vector_type(2)));
  constexpr int N = 1024 * NT;
  const int to = opaque(threadIdx.x);
  const int tau = to & 15, m3 = to >> 4, nb = NT * m3 + tau % NT;
  const float2 *wd = wt + nb, *wm = wt + (N - 1 - nb);  // direct and mirrored entries at m2 = 0
#pragma unroll
  for (int g = 0; g < 2; g++) {
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
      uint32_t pr[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int m2 = 16 * g + cm_kslot(hh, e);
        float2 w;
        if (g == 0) {
          w = wd[32 * NT * m2];
        } else {
          const float2 u = wm[-32 * NT * m2];
          w = make_float2(u.y, u.x);
        }
        const h2v x = __builtin_bit_cast(h2v, xp[16 * g + 8 * hh + e]);
        const h2v z = __builtin_convertvector((f2v){(float)x.x * w.x, (float)x.y * w.y}, h2v);
        pr[e] = g == 1 && zero_hi ? 0u : __builtin_bit_cast(uint32_t, z) ^ flipmask;
      }
      uint4 re, im;
      re.x = __builtin_amdgcn_perm(pr[1], pr[0], 0x05040100u);
      re.y = __builtin_amdgcn_perm(pr[3], pr[2], 0x05040100u);
      re.z = __builtin_amdgcn_perm(pr[5], pr[4], 0x05040100u);
      re.w = __builtin_amdgcn_perm(pr[7], pr[6], 0x05040100u);
      im.x = __builtin_amdgcn_perm(pr[1], pr[0], 0x07060302u);
      im.y = __builtin_amdgcn_perm(pr[3], pr[2], 0x07060302u);
      im.z = __builtin_amdgcn_perm(pr[5], pr[4], 0x07060302u);
      im.w = __builtin_amdgcn_perm(pr[7], pr[6], 0x07060302u);
      lds[cm_chunk(tau, m3, 2 * g + hh)] = re;
      lds[cm_chunk(tau, m3, 4 + 2 * g + hh)] = im;
    }
  }
}

// A workgroup walks a run of consecutive segments of one stream, a pass of
// 16 / NT segments at a time. With a window the second half of a segment is
// the first half of the next one, so a pass reads only the 16 new points of
// each thread; they are loaded before this pass's spectrum stores (the
// vector memory counter is in order, so waiting for them does not wait for
// the stores). Without a window a segment uses only its first half and
// nothing carries. One workgroup per CU.
template <int SP, int NT, bool H>
__global__ __launch_bounds__(CM_T, 2) void k_chan_mfma(const DevStream *__restrict__ streams, DevJob job,
                                                      int segs_per_wg) {
  static_assert(SP == SFXG_PREC_FP16 || SP == SFXG_PREC_BF16, "packed FP16 / BF16 spectra");
  static_assert(NT == 1 || NT == 2 || NT == 4 || NT == 8 || NT == 16, "1024 NT complex points, NT <= 16");
  constexpr int N = 1024 * NT;  // complex points of a segment (N_c)
  constexpr int SG = 16 / NT;   // segments per pass (16 tiles)
  __shared__ uint4 lds[4096];  // 64 KiB: the 16 tiles' operands, then V
  __shared__ float2 wt[N / 2];  // the window as FP32 pairs (cm_window_table)
  // per-thread constants (kept in LDS, not in registers held across the
  // segment loop): W_N^q0, W_2N^q0 and row 512's constants
  __shared__ float2 cth[3][CM_T];
  __shared__ float4 ctw[8][64];  // per lane: W1024^(m3 col) of registers 2 p, 2 p + 1
  const DevStream &s = streams[blockIdx.y];
  const int64_t sa = (int64_t)blockIdx.x * segs_per_wg;
  int64_t sb = sa + segs_per_wg;
  if (sb > job.n_segments) sb = job.n_segments;
  if (sa >= sb) return;  // uniform over the workgroup
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, h = lane >> 5, col = lane & 31;
  // thread t = 16 m3 + tau takes tile tau = NT sigma + n1 (segment sigma of
  // the pass, residue n1): points n = NT (32 m2 + m3) + n1
  cm_window_table<N>(wt, job.window);
  const bool zero_hi = job.window == SFXG_WINDOW_NONE;
  // DFT-32 fragments: lane holds F[16 g + cm_kslot(h, e)][col] (e = 0..7),
  // F[a][b] = W32^(a b): real parts, imaginary parts and their negatives
  cm_h8 fr[2], fi[2], nfi[2];
#pragma unroll
  for (int g = 0; g < 2; g++)
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const float2 w = cm_w((16 * g + cm_kslot(h, e)) * col, 32);
      fr[g][e] = (_Float16)w.x;
      fi[g][e] = (_Float16)w.y;
      nfi[g][e] = (_Float16)-w.y;
    }
  // W1024^(m3 col) of the B1 result in register r (m3 = 8 (r / 4) + 4 h + r % 4),
  // times hs = scale / 2: the split's factor 1/2 and the spectra's
  // power-of-two scale (exact), which also keeps the FP16 operand of B2 and
  // V near unit rms, far from the FP16 range's ends
  if (t < 64) {
    const float hs = 0.5f * job.spec_scale;
#pragma unroll
    for (int p = 0; p < 8; p++) {
      const float2 a = cm_w((8 * (p >> 1) + 4 * h + 2 * (p & 1)) * col, 1024);
      const float2 b = cm_w((8 * (p >> 1) + 4 * h + 2 * (p & 1) + 1) * col, 1024);
      ctw[p][t] = make_float4(hs * a.x, hs * a.y, hs * b.x, hs * b.y);
    }
  }
  // the last stage's rows: q0 = t and q1 = 1024 - t (twiddles from those of
  // q0, cm_last_stage2). Thread 0 takes row 0 twice, the second time as "row
  // 1024", whose twiddles W_N^(1024 n1) = W_NT^n1 shift the NT-point result
  // by one bin, so that the pairing below pairs row 0's bins k1 and NT - k1
  // and yields bin N. Row 512 (its bins k1 and NT - 1 - k1 pair with each
  // other) goes to lane 0 of wave 0 besides.
  {
    const int q0 = t;
    cth[0][t] = cm_w(q0, N);
    cth[1][t] = cm_w(q0, 2 * N);
    // row 512's constants: W_N^512 (last stage) and W_2N^512 (split)
    cth[2][t] = t == 0 ? cm_w(512, N) : t == 1 ? cm_w(512, 2 * N) : make_float2(0.0f, 0.0f);
  }
  const uint32_t flipmask = s.flip ? 0x80000000u : 0u;  // sign of the imaginary half
  // the real series as pairs of reals: float2 (FP32) or one 32-bit word of
  // two FP16 (H)
  const float2 *ts = reinterpret_cast<const float2 *>(s.time_series);
  const uint32_t *tsh = reinterpret_cast<const uint32_t *>(s.time_series);
  uint2 *vlds = reinterpret_cast<uint2 *>(lds);
  // point i = 16 g + 8 hh + e of the thread has m2 = 16 g + cm_kslot(hh, e);
  // xp[0..15]: the first half of its segment, xp[16..31]: the second, as
  // FP16 pairs (an FP32 series is rounded to FP16 as it is loaded)
  uint32_t xp[32], xn[16];
  // With a window the second half of a segment is the first half of the
  // next one, so a pass reads only the second halves of its segments (16
  // points per thread): for NT = 16 (one segment per pass) the thread
  // carries its second half into the next pass in registers; for NT < 16
  // the first half of segment sigma comes from the thread of segment
  // sigma - 1 (NT lanes lower in the same wave), and for sigma = 0 from the
  // previous pass's last segment (NT (SG - 1) lanes higher), by ds_bpermute.
  // Without a window every pass reads the first halves only (the second is
  // zero padding).
  const bool carry = !zero_hi;
  auto load_half = [&](uint32_t (&dst)[16], int64_t s0, int g) {
    const int to = opaque(threadIdx.x);
    const int tau = to & 15, m3 = to >> 4, sig = tau / NT, n1 = tau % NT;
    const bool ok = s0 + sig < sb;
    const int64_t p0 = (ok ? s0 + sig : s0) * (N / 2);  // the segment's first point
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t p = (uint32_t)(NT * (32 * (16 * g + cm_kslot(i >> 3, i & 7)) + m3) + n1);
      uint32_t v;
      if constexpr (H) {
        v = *((const __attribute__((address_space(1))) uint32_t *)(tsh + p0) + p);
      } else {
        typedef float f2v __attribute__((ext_vector_type(2)));
        typedef _Float16 h2v __attribute__((ext_vector_type(2)));
        const float2 x = gload2(ts + p0, p * (uint32_t)sizeof(float2));
        v = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){x.x, x.y}, h2v));
      }
      dst[i] = ok ? v : 0u;
    }
  };
  auto load_pass = [&](int64_t s0) {  // without carry
    uint32_t a[16];
    load_half(a, s0, 0);
#pragma unroll
    for (int i = 0; i < 16; i++) xp[i] = a[i];
    if (!zero_hi) {
      load_half(a, s0, 1);
#pragma unroll
      for (int i = 0; i < 16; i++) xp[16 + i] = a[i];
    }
  };
  // without a window the second half is zero padding: never loaded, and
  // zero (not left undefined: a non-finite value times the zero weight would
  // give NaN)
#pragma unroll
  for (int i = 0; i < 16; i++) xp[16 + i] = 0u;
  __syncthreads();  // the window table
  if (carry && NT == 16) {
    uint32_t a[16];
    load_half(a, sa, 0);
    load_half(xn, sa, 1);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      xp[i] = a[i];
      xp[16 + i] = xn[i];
    }
  } else {
    load_pass(sa);
  }
  // complete here, so that no wait at the loop top (where the previous
  // pass's stores are in flight) is needed for them
#pragma unroll
  for (int i = 0; i < 32; i++) asm volatile("" ::"v"(xp[i]));
  for (int64_t seg = sa; seg < sb; seg += SG) {
    // phase 0: windowed (and sideband-flipped) points as FP16 K chunks
    cm_phase0<NT>(lds, xp, wt, flipmask, zero_hi);
    // with carry this segment's second half is the next one's first
    if (carry && NT == 16) {
#pragma unroll
      for (int i = 0; i < 16; i++) xp[i] = xp[16 + i];
    }
    __syncthreads();
    // the next pass's points: they complete before this pass's spectrum
    // stores, so no wait for them also waits for stores
    const bool more = seg + SG < sb;
    if (more) {
      if (carry) load_half(xn, seg + SG, 1);
      else load_pass(seg + SG);
    }
    // wave w takes tiles 2 w and 2 w + 1 (K slices of their A operands: real
    // parts of m2 0..15, 16..31, then the imaginary parts) and leaves their V
    // in its own 8 KiB of LDS, which held just those two tiles: no other
    // wave reads it before the next barrier
    {
      const int lo = opaque(lane), colo = lo & 31, ho = lo >> 5;
      uint32_t v0[16];
#pragma unroll
      for (int tt = 0; tt < 2; tt++) {
        cm_h8 a[4];
#pragma unroll
        for (int sl = 0; sl < 4; sl++)
          a[sl] = *reinterpret_cast<const cm_h8 *>(&lds[cm_chunk(2 * wave + tt, colo, 2 * sl + ho)]);
        cm_f16 ur = {}, ui = {};
        ur = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0], fr[0], ur, 0, 0, 0);
        ur = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1], fr[1], ur, 0, 0, 0);
        ur = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[2], nfi[0], ur, 0, 0, 0);
        ur = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[3], nfi[1], ur, 0, 0, 0);
        ui = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0], fi[0], ui, 0, 0, 0);
        ui = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1], fi[1], ui, 0, 0, 0);
        ui = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[2], fr[0], ui, 0, 0, 0);
        ui = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[3], fr[1], ui, 0, 0, 0);
        // U[m3][col] W1024^(m3 col) as B2's B operand: K slice g2 holds
        // registers 8 g2 .. 8 g2 + 7 (rows m3 = 16 g2 + cm_kslot(h, e))
        cm_h8 br[2], bi[2];
#pragma unroll
        for (int p = 0; p < 8; p++) {
          const float4 w = ctw[p][lo];
#pragma unroll
          for (int q = 0; q < 2; q++) {
            const int r = 2 * p + q, g = r >> 2, j = r & 3;
            const float2 u = cmul(make_float2(ur[r], ui[r]), q ? make_float2(w.z, w.w) : make_float2(w.x, w.y));
            br[g >> 1][4 * (g & 1) + j] = (_Float16)u.x;
            bi[g >> 1][4 * (g & 1) + j] = (_Float16)u.y;
          }
        }
        cm_f16 vr = {}, vi = {};
        vr = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[0], br[0], vr, 0, 0, 0);
        vr = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[1], br[1], vr, 0, 0, 0);
        vr = __builtin_amdgcn_mfma_f32_32x32x16_f16(nfi[0], bi[0], vr, 0, 0, 0);
        vr = __builtin_amdgcn_mfma_f32_32x32x16_f16(nfi[1], bi[1], vr, 0, 0, 0);
        vi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fi[0], br[0], vi, 0, 0, 0);
        vi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fi[1], br[1], vi, 0, 0, 0);
        vi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[0], bi[0], vi, 0, 0, 0);
        vi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fr[1], bi[1], vi, 0, 0, 0);
        // V[q] with q = col + 32 k3, k3 = 8 g + 4 h + j in register 4 g + j
        if (tt == 0) {
#pragma unroll
          for (int r = 0; r < 16; r++) v0[r] = cm_pack(vr[r], vi[r]);
        } else {
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int q = colo + 32 * (8 * (r >> 2) + 4 * ho + (r & 3));
            vlds[wave * 1024 + q] = make_uint2(v0[r], cm_pack(vr[r], vi[r]));
          }
        }
      }
    }
    __syncthreads();
    // last stage and split: in each segment of the pass, rows q0 and q1
    // hold bins k and N - k
    {
      const int tq = opaque(threadIdx.x), r0 = tq, r1 = 1024 - tq;
      if (more) {
        if (carry && NT == 16) {
#pragma unroll
          for (int i = 0; i < 16; i++) {
            asm volatile("" ::"v"(xn[i]));  // the loads complete here
            xp[16 + i] = xn[i];
          }
        } else if (carry) {
          // first halves from the neighbouring segments' second halves: the
          // lane of segment SG - 1 hands on this pass's, the others the
          // next pass's
          const int lo = opaque(lane), sg = (lo & 15) / NT;
          const int src = 4 * (sg == 0 ? lo + NT * (SG - 1) : lo - NT);
          const bool last = sg == SG - 1;
#pragma unroll
          for (int i = 0; i < 16; i++) {
            asm volatile("" ::"v"(xn[i]));  // the loads complete here
            xp[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(last ? xp[16 + i] : xn[i]));
            xp[16 + i] = xn[i];
          }
        }
        if (!carry) {
#pragma unroll
          for (int i = 0; i < 32; i++) asm volatile("" ::"v"(xp[i]));
        }
      }
      const float2 wn0 = cth[0][tq], w0 = cth[1][tq];
#pragma unroll
      for (int sig = 0; sig < SG; sig++) {
        if (seg + sig >= sb) break;  // uniform
        float2 z0[NT], z1[NT];
        cm_read_rows<NT>(vlds, sig, r0, z0);
        cm_read_rows<NT>(vlds, sig, r1 & 1023, z1);
        cm_last_stage2<NT>(z0, z1, wn0);
        uint32_t *row = reinterpret_cast<uint32_t *>(s.spectra) + (seg + sig) * job.spectra_stride;
#pragma unroll
        for (int k1 = 0; k1 < NT; k1++) {
          float2 xk, xn;
          // Z[N - k] of bin k = q0 + 1024 k1: row q1, bin NT - 1 - k1, held
          // in z1[(NT - k1) mod NT]
          cm_split(z0[k1], z1[(NT - k1) & (NT - 1)], cmul(w0, cm_w32((16 / NT) * k1)), xk, xn);
          cm_put<SP>(row, (uint32_t)(r0 + 1024 * k1), xk);
          cm_put<SP>(row, (uint32_t)(r1 + 1024 * (NT - 1 - k1)), xn);
        }
        if (tq == 0) {
          // row 512: bins 512 + 1024 k1 and 512 + 1024 (NT - 1 - k1) pair
          cm_read_rows<NT>(vlds, sig, 512, z1);
          cm_last_stage<NT>(z1, cth[2][0]);
          const float2 w5 = cth[2][1];
#pragma unroll
          for (int k1 = 0; k1 < (NT + 1) / 2; k1++) {
            float2 xk, xn;
            cm_split(z1[k1], z1[NT - 1 - k1], cmul(w5, cm_w32((16 / NT) * k1)), xk, xn);
            cm_put<SP>(row, (uint32_t)(512 + 1024 * k1), xk);
            if (NT - 1 - k1 != k1) cm_put<SP>(row, (uint32_t)(512 + 1024 * (NT - 1 - k1)), xn);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Channeliser for N = 32768 complex points (2 N_c = 65536 reals), which do
// not fit in LDS: a four-step FFT through HBM scratch, N = 128 x 256,
//   X[k1 + 128 k2] = sum_n2 W256^(n2 k2) W_N^(n2 k1) sum_n1 x[256 n1 + n2] W128^(n1 k1).
// k_chan_cols: 128-point FFTs over n1 for 16 columns n2 per workgroup, times
// W_N^(n2 k1), into Y[k1][n2]; k_chan_rows: 256-point FFTs over n2 for 16 rows
// k1 into Z in natural order; k_chan_split: the N + 1 bins of the real
// transform. Work items are (stream, segment) pairs item0 + blockIdx.y; the
// scratch holds one N-point row per item of the launch.
// ---------------------------------------------------------------------------
constexpr int LG_N1 = 128, LG_N2 = 256, LG_N = LG_N1 * LG_N2;

struct StrideTw {  // exp(-2 pi i m / n) = tw_N[m * stride], stride = N / n
  static constexpr bool kChain = false;
  const float2 *__restrict__ tw;
  int stride;
  __device__ __forceinline__ float2 fft(int m) const { return tw[m * stride]; }
};

__global__ __launch_bounds__(256) void k_chan_cols(const DevStream *__restrict__ streams, DevJob job,
                                                   const float *__restrict__ window, const float2 *__restrict__ tw_n,
                                                   int64_t item0, float2 *__restrict__ Y) {
  constexpr int C = 16, P = 8, T1 = LG_N1 / P;  // 16 column FFTs of 16 threads
  constexpr int RS = fft_lds_size<LG_N1, P>() + 1;  // odd region stride
  __shared__ float2 lds[C * RS];
  const int64_t item = item0 + blockIdx.y;
  const int st = (int)(item / job.n_segments);
  const int64_t seg = item - (int64_t)st * job.n_segments;
  const DevStream &s = streams[st];
  const float2 *x = reinterpret_cast<const float2 *>(s.time_series + seg * job.seg_step);
  const float2 *w = reinterpret_cast<const float2 *>(window);
  const bool none = job.window == SFXG_WINDOW_NONE;
  const int c = threadIdx.x % C, tid = threadIdx.x / C;  // consecutive threads: consecutive columns
  const int n2 = blockIdx.x * C + c;
  float2 v[P];
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int m = LG_N2 * (tid + i * T1) + n2;
    float2 z = make_float2(0.0f, 0.0f);
    if (!(none && m >= LG_N / 2)) {
      z = wmul(gload2(x + m), w[m]);
    }
    if (s.flip) z.y = -z.y;
    v[i] = z;
  }
  float2 *reg = lds + c * RS;
  reg_fft<LG_N1, P>(v, reg, StrideTw{tw_n, LG_N / LG_N1}, tid);
  float2 *y = Y + (int64_t)blockIdx.y * LG_N;
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int k1 = tid + i * T1;
    y[k1 * LG_N2 + n2] = cmul(reg[lds_index<P>(k1)], tw_n[n2 * k1]);
  }
}

__global__ __launch_bounds__(256) void k_chan_rows(const float2 *__restrict__ Y, const float2 *__restrict__ tw_n,
                                                   float2 *__restrict__ Z) {
  constexpr int R = 16, P = 16, T1 = LG_N2 / P;  // 16 row FFTs of 16 threads
  constexpr int RS = fft_lds_size<LG_N2, P>() + 1;
  __shared__ float2 lds[R * RS];
  const int tid = threadIdx.x % T1, r = threadIdx.x / T1;  // consecutive threads: consecutive n2
  const int k1 = blockIdx.x * R + r;
  const float2 *y = Y + (int64_t)blockIdx.y * LG_N + k1 * LG_N2;
  float2 v[P];
#pragma unroll
  for (int i = 0; i < P; i++) v[i] = y[tid + i * T1];
  reg_fft<LG_N2, P>(v, lds + r * RS, StrideTw{tw_n, LG_N / LG_N2}, tid);
  // Z[k1 + 128 k2]: consecutive threads take consecutive k1 of the block
  float2 *z = Z + (int64_t)blockIdx.y * LG_N;
  for (int f = threadIdx.x; f < R * LG_N2; f += 256) {
    const int rr = f % R, k2 = f / R;
    z[blockIdx.x * R + rr + LG_N1 * k2] = lds[rr * RS + lds_index<P>(k2)];
  }
}

template <int SP>
__global__ __launch_bounds__(256) void k_chan_split(const DevStream *__restrict__ streams, DevJob job,
                                                    const float2 *__restrict__ Z, const float2 *__restrict__ tw_2n,
                                                    int64_t item0) {
  const int64_t item = item0 + blockIdx.x;
  const int st = (int)(item / job.n_segments);
  const int64_t seg = item - (int64_t)st * job.n_segments;
  const float2 *z = Z + (int64_t)blockIdx.x * LG_N;
  const SpecOut<SP> out(streams[st].spectra, seg * job.spectra_stride, job.spec_scale);
  for (int k = threadIdx.x; k <= LG_N; k += 256) {
    const float2 zk = z[k & (LG_N - 1)], zn = z[(LG_N - k) & (LG_N - 1)];
    const float2 e = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
    const float2 o = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
    out.put((uint32_t)k, cadd(e, cmul(tw_2n[k], o)));
  }
}

// ---------------------------------------------------------------------------
// k_fengine: Delay_correction and the channeliser fused. A workgroup walks a
// contiguous range of segments of one stream. The real series lives in an
// LDS ring of two N_c halves: each step computes the next N_c samples
// (N_c / 256 delay blocks, one per wave) into the older half, then windows
// the 2 N_c samples of the segment into the FFT buffer, transforms them and
// writes the N_c + 1 bins. The series never goes to HBM: per station-sample
// the kernel reads b/8 bytes of payload and writes the 8 (N_c+1)/N_c bytes of
// the spectrum. The FFT buffer doubles as the waves' scratch for the 256-point
// transforms.
// ---------------------------------------------------------------------------
template <int N, int T, int SP>
__global__ __launch_bounds__(T) void k_fengine(const DevStream *__restrict__ streams, DevJob job,
                                               const float *__restrict__ window, const float2 *__restrict__ tw_nd,
                                               const float2 *__restrict__ tw_split, int segs_per_wg) {
  static_assert(N >= 256 && (T / 64) * WAVE_FFT256_LDS <= fft_lds_size<N, N / T>(),
                "wave scratch must fit in the FFT buffer");
  constexpr int P = N / T, NHI = 2 * N / 128;
  __shared__ float ring[2 * N];
  __shared__ float2 fbuf[fft_lds_size<N, P>()];
  __shared__ float2 tw[TW256_ELEMS];
  __shared__ float2 thi[NHI], tlo[128];
  const DevStream &s = streams[blockIdx.y];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  constexpr int nwaves = T / 64;
  constexpr int BPH = N / 256;  // delay blocks per half
  const int64_t sa = (int64_t)blockIdx.x * segs_per_wg;
  int64_t sb = sa + segs_per_wg;
  if (sb > job.n_segments) sb = job.n_segments;
  if (sa >= sb) return;  // uniform over the workgroup
  fill_tw256(tw, tw_nd, threadIdx.x, T);
  for (int i = threadIdx.x; i < NHI + 128; i += T) {
    const float2 t = tw_split[i];
    if (i < NHI) thi[i] = t; else tlo[i - NHI] = t;
  }
  __syncthreads();
  const auto twl = channeliser_tw<SP>(thi, tlo);
  const bool none = job.window == SFXG_WINDOW_NONE;
  float2 *scratch = fbuf + wave * WAVE_FFT256_LDS;
  // blocks go in pairs (blk, blk + 1) with blk even, as in k_delay
  if (!none) {
    for (int b = 2 * wave; b < BPH; b += 2 * nwaves)
      delay_block_pair(s, (job.seg_base + sa) * BPH + b, job.n_blocks, scratch, tw, lane, ring + b * 256,
                       b + 1 < BPH ? ring + (b + 1) * 256 : nullptr);
    __syncthreads();
  }
  const float2 *w2 = reinterpret_cast<const float2 *>(window);
  for (int64_t seg = sa; seg < sb; seg++) {
    const int h_old = (int)((seg - sa) & 1);
    const int h_new = none ? 0 : 1 - h_old;
    const int64_t blk0 = (job.seg_base + seg + (none ? 0 : 1)) * BPH;
    for (int b = 2 * wave; b < BPH; b += 2 * nwaves)
      delay_block_pair(s, blk0 + b, job.n_blocks, scratch, tw, lane, ring + h_new * N + b * 256,
                       b + 1 < BPH ? ring + h_new * N + (b + 1) * 256 : nullptr);
    __syncthreads();
    // window the segment's 2 N real samples into the FFT's registers:
    // complex point m = tid + i T is the pair of reals 2m, 2m + 1, which lie
    // in the segment's first half exactly when i < P / 2 (N = P T). Without a
    // window the segment is the N reals of half 0 and N zeros. Branch-free,
    // with every window load issued first.
    const int tid = opaque(threadIdx.x);
    const float2 *ring2 = reinterpret_cast<const float2 *>(ring);
    float2 wv[P];
#pragma unroll
    for (int i = 0; i < P; i++) wv[i] = gload2(w2, (uint32_t)(tid + i * T) * (uint32_t)sizeof(float2));
    float2 v[P];
#pragma unroll
    for (int i = 0; i < P; i++) {
      const int m = tid + i * T;
      const int half = i < P / 2 ? h_old : 1 - h_old;
      const float2 x = ring2[none ? m : half * (N / 2) + (m & (N / 2 - 1))];
      float2 z = wmul(x, wv[i]);
      if (none && i >= P / 2) z = make_float2(0.0f, 0.0f);
      if (s.flip) z.y = -z.y;
      v[i] = z;
    }
    reg_fft<N, P>(v, fbuf, twl, tid);
    real_split<N, P, T>(fbuf, twl, SpecOut<SP>(s.spectra, seg * job.spectra_stride, job.spec_scale), tid);
    __syncthreads();
  }
}

namespace {

// Workgroup index of a one-dimensional grid after the XCD regrouping: the
// hardware deals workgroups to the 8 XCDs in turn, so workgroup i runs on
// XCD i % 8. Groups of G consecutive logical indices (workgroups that read the
// same spectra: channel blocks sharing 128-byte lines, or the stream-pair
// groups of one channel block) are given to one XCD, where they run
// together and its L2 fetches each line once. A bijection on [0, n): the last
// n % (8 G) workgroups keep their index.
__device__ __forceinline__ int xcd_grouped(int i, int n, int G) {
  const int full = n / (8 * G) * (8 * G);
  if (G <= 1 || i >= full) return i;
  const int xcd = i & 7, slot = i >> 3;
  return (slot / G) * (8 * G) + xcd * G + slot % G;
}

}  // namespace

// ---------------------------------------------------------------------------
// k_xstep_subint: plain FP32 X-step for one sub-integration per workgroup
// slice of the grid. Lane = channel (64-channel tile), each wave owns up to
// XBPW baselines. The segments of the tile are staged through LDS in groups
// of `sps` segments x M streams x 64 channels (coalesced 512 B rows), and
// every product reads its two operands from LDS.
//
// The reference accumulates every (baseline, channel) sequentially over the
// segments of a sub-integration, starting from zero, and adds the result to
// the phase-centre buffer at the fold (correlation_core.cc:241-273,515-559).
// Sub-integrations are therefore independent partial sums: each z-slice of
// the grid computes one into `partial`, and k_fold_combine adds them in fold
// order. Products and sums use the reference's operation order without
// contraction (std::complex<float> multiply, then +=).
// ---------------------------------------------------------------------------
constexpr int XBPW = 16;  // baselines per wave

__global__ __launch_bounds__(512) void k_xstep_subint(const float2 *__restrict__ spectra, int64_t stream_stride,
                                                      int64_t seg_stride, int64_t seg_base, int n_ch, int M,
                                                      const int32_t *__restrict__ fold_start,
                                                      const int32_t *__restrict__ bl_first,
                                                      const int32_t *__restrict__ bl_second, int n_baselines,
                                                      int sps, float2 *__restrict__ partial) {
#pragma clang fp contract(off)
  extern __shared__ float2 xlds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nwaves = blockDim.x >> 6;
  const int ch0 = blockIdx.x * 64;
  const int ch = ch0 + lane;
  const int k = blockIdx.z;
  const int64_t s_begin = fold_start[k], s_end = fold_start[k + 1];
  const int b0 = __builtin_amdgcn_readfirstlane((blockIdx.y * nwaves + wave) * XBPW);
  int nb = n_baselines - b0;
  nb = nb < 0 ? 0 : (nb > XBPW ? XBPW : nb);
  int oa[XBPW], ob[XBPW];
  float2 acc[XBPW];
#pragma unroll
  for (int q = 0; q < XBPW; q++) {
    const int b = q < nb ? b0 + q : 0;
    oa[q] = bl_first[b] * 64 + lane;
    ob[q] = bl_second[b] * 64 + lane;
    acc[q] = make_float2(0.0f, 0.0f);
  }
  const int rows_per_seg = M * 64;
  // stage = up to sps segments x M rows x 64 channels; thread t holds the
  // elements t, t + T, ... of the next stage in registers while the current
  // one is consumed from LDS
  constexpr int PF = 16;  // per-thread prefetch registers (float2)
  float2 pf[PF];
  auto fetch = [&](int64_t s0, int ns) {
    const int total = ns * rows_per_seg;
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int idx = threadIdx.x + u * blockDim.x;
      float2 v = make_float2(0.0f, 0.0f);
      if (idx < total) {
        const int c = idx & 63;
        const int r = idx >> 6;
        const int sl = r / M, st = r - sl * M;
        if (ch0 + c < n_ch) v = spectra[st * stream_stride + (s0 + sl - seg_base) * seg_stride + ch0 + c];
      }
      pf[u] = v;
    }
  };
  int64_t s0 = s_begin;
  int ns = (int)((s_end - s0) < sps ? (s_end - s0) : sps);
  if (s0 < s_end) fetch(s0, ns);
  while (s0 < s_end) {
    const int total = ns * rows_per_seg;
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int idx = threadIdx.x + u * blockDim.x;
      if (idx < total) xlds[idx] = pf[u];
    }
    __syncthreads();
    const int64_t s1 = s0 + ns;
    const int ns1 = (int)((s_end - s1) < sps ? (s_end - s1) : sps);
    if (s1 < s_end) fetch(s1, ns1);
    for (int sl = 0; sl < ns; sl++) {
      const float2 *row = xlds + sl * rows_per_seg;
#pragma unroll
      for (int q = 0; q < XBPW; q++) {
        if (q < nb) {
          const float2 x = row[oa[q]], y = row[ob[q]];
          const float pr = x.x * y.x + x.y * y.y;
          const float pi = x.y * y.x - x.x * y.y;
          acc[q].x = acc[q].x + pr;
          acc[q].y = acc[q].y + pi;
        }
      }
    }
    __syncthreads();
    s0 = s1;
    ns = ns1;
  }
  if (ch < n_ch) {
    float2 *o = partial + (int64_t)k * n_baselines * n_ch + ch;
#pragma unroll
    for (int q = 0; q < XBPW; q++)
      if (q < nb) o[(int64_t)(b0 + q) * n_ch] = acc[q];
  }
}

// ---------------------------------------------------------------------------
// k_xstep_tile: the plain FP32 X-step as register tiles of stream pairs. A
// wave owns a 4 x 4 tile (streams a0.. x b0..) of 64 channels (lane =
// channel): per segment it reads 8 spectra from LDS for 16 products, where a
// per-baseline scheme reads 2 per product. 1024-thread workgroups stage
// segments x all streams x 64 channels in LDS (one workgroup per CU). The
// ceil(tiles / 16) workgroups of one channel block and sub-integration read
// the same spectra: the one-dimensional grid puts them on one XCD
// (xcd_grouped), where they run together and share its L2. The tile
// table maps each of the 16 pairs to its baseline: b >= 0 stores the sum,
// ~b its conjugate (the baseline lists the pair the other way round: the
// product of the swapped pair is exactly the conjugate), XT_NONE nothing.
// Products and sums are the reference's, in its order, per baseline.
// ---------------------------------------------------------------------------
constexpr int XT = 4;          // streams per tile side
constexpr int XT_WAVES = 16;   // tiles per workgroup
constexpr int XT_NONE = INT32_MIN;
constexpr int XT_PF = 8;       // prefetch registers per thread (float2)

__global__ __launch_bounds__(64 * XT_WAVES) void k_xstep_tile(
    const float2 *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base, int n_ch,
    int n_run, int M, const int32_t *__restrict__ fold_start, const int32_t *__restrict__ tiles, int n_tiles,
    int n_baselines, int sps, int n_cb, int n_tg, float2 *__restrict__ partial) {
#pragma clang fp contract(off)
  extern __shared__ float2 xlds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // (tile group, channel block, sub-integration), tile groups innermost
  const int L = xcd_grouped(blockIdx.x, gridDim.x, n_tg);
  const int tg = L % n_tg, cb = (L / n_tg) % n_cb, k = L / n_tg / n_cb;
  const int ch0 = cb * 64;
  const int ch = ch0 + lane;
  const int64_t s_begin = fold_start[k], s_end = fold_start[k + 1];
  const int t = __builtin_amdgcn_readfirstlane(tg * XT_WAVES + wave);
  const bool active = t < n_tiles;
  const int a0 = active ? tiles[t * (2 + XT * XT)] : 0;
  const int b0 = active ? tiles[t * (2 + XT * XT) + 1] : 0;
  float2 acc[XT * XT];
#pragma unroll
  for (int q = 0; q < XT * XT; q++) acc[q] = make_float2(0.0f, 0.0f);
  const int rows_per_seg = M * 64;
  constexpr int PF = XT_PF;  // per-thread prefetch registers (float2)
  float2 pf[PF];
  auto fetch = [&](int64_t s0, int ns) {
    const int total = ns * rows_per_seg;
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int idx = threadIdx.x + u * (64 * XT_WAVES);
      float2 v = make_float2(0.0f, 0.0f);
      if (idx < total) {
        const int c = idx & 63;
        const int r = idx >> 6;
        const int sl = r / M, st = r - sl * M;
        if (ch0 + c < n_run) v = spectra[st * stream_stride + (s0 + sl - seg_base) * seg_stride + ch0 + c];
      }
      pf[u] = v;
    }
  };
  int64_t s0 = s_begin;
  int ns = (int)((s_end - s0) < sps ? (s_end - s0) : sps);
  if (s0 < s_end) fetch(s0, ns);
  while (s0 < s_end) {
    const int total = ns * rows_per_seg;
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int idx = threadIdx.x + u * (64 * XT_WAVES);
      if (idx < total) xlds[idx] = pf[u];
    }
    __syncthreads();
    const int64_t s1 = s0 + ns;
    const int ns1 = (int)((s_end - s1) < sps ? (s_end - s1) : sps);
    if (s1 < s_end) fetch(s1, ns1);
    if (active) {
#pragma unroll 1
      for (int sl = 0; sl < ns; sl++) {
        const float2 *row = xlds + sl * rows_per_seg + lane;
        float2 x[XT], y[XT];
#pragma unroll
        for (int i = 0; i < XT; i++) {
          x[i] = a0 + i < M ? row[(a0 + i) * 64] : make_float2(0.0f, 0.0f);
          y[i] = b0 + i < M ? row[(b0 + i) * 64] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int i = 0; i < XT; i++)
#pragma unroll
          for (int j = 0; j < XT; j++) {
            const float pr = x[i].x * y[j].x + x[i].y * y[j].y;
            const float pi = x[i].y * y[j].x - x[i].x * y[j].y;
            acc[XT * i + j].x = acc[XT * i + j].x + pr;
            acc[XT * i + j].y = acc[XT * i + j].y + pi;
          }
      }
    }
    __syncthreads();
    s0 = s1;
    ns = ns1;
  }
  if (active && ch < n_run) {
    float2 *o = partial + (int64_t)k * n_baselines * n_ch + ch;
#pragma unroll
    for (int q = 0; q < XT * XT; q++) {
      const int e = tiles[t * (2 + XT * XT) + 2 + q];
      if (e == XT_NONE) continue;
      if (e >= 0) o[(int64_t)e * n_ch] = acc[q];
      else o[(int64_t)(~e) * n_ch] = make_float2(acc[q].x, -acc[q].y);
    }
  }
}

// ---------------------------------------------------------------------------
// k_xstep_tile16: the plain FP32 X-step for many stream-pair tiles (large
// station counts). A workgroup of 12 waves holds every tile of a block of
// 16 channels at once: lane = (tile slot u, channel c), 4 slots of 16 lanes
// per wave, 48 slots of TPS tiles each (up to 144 tiles of 4 x 4 pairs, all
// of 64 streams). A stage of segments is staged as [segment][stream][16
// channels], 128-byte rows, so each spectrum line is fetched from HBM once
// and read from LDS by every tile that needs it. More tiles than one
// workgroup holds go to further tile groups on the same XCD. Products,
// sums and their order per (baseline, channel) are those of k_xstep_tile,
// so the results are the same bits.
// ---------------------------------------------------------------------------
constexpr int XQ_WAVES = 12;              // waves per workgroup (three per SIMD)
constexpr int XQ_SLOTS = 4 * XQ_WAVES;    // tile slots of 16 lanes
constexpr int XQ_PF = 3;                  // prefetch registers per thread (float4)

template <int TPS>
__global__ __launch_bounds__(64 * XQ_WAVES) void k_xstep_tile16(
    const float2 *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base, int n_ch,
    int n_run, int M, const int32_t *__restrict__ fold_start, const int32_t *__restrict__ tiles, int n_tiles,
    int n_baselines, int sps, int n_cb, int n_tg, float2 *__restrict__ partial) {
#pragma clang fp contract(off)
  extern __shared__ float4 xq[];  // [sps][M][8]: 16 channels of a stream-segment as 8 float4
  const float2 *xq2 = reinterpret_cast<const float2 *>(xq);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15;
  const int slot = wave * 4 + (lane >> 4);
  // (tile group, channel block, sub-integration), tile groups innermost
  const int L = xcd_grouped(blockIdx.x, gridDim.x, n_tg);
  const int tg = L % n_tg, cb = (L / n_tg) % n_cb, k = L / n_tg / n_cb;
  const int ch0 = cb * 16;
  const int64_t s_begin = fold_start[k], s_end = fold_start[k + 1];
  int a0[TPS], b0[TPS];
  bool act[TPS];
#pragma unroll
  for (int j = 0; j < TPS; j++) {
    const int t = (tg * TPS + j) * XQ_SLOTS + slot;
    act[j] = t < n_tiles;
    a0[j] = act[j] ? tiles[t * (2 + XT * XT)] : 0;
    b0[j] = act[j] ? tiles[t * (2 + XT * XT) + 1] : 0;
  }
  float2 acc[TPS][XT * XT];
#pragma unroll
  for (int j = 0; j < TPS; j++)
#pragma unroll
    for (int q = 0; q < XT * XT; q++) acc[j][q] = make_float2(0.0f, 0.0f);
  // stage loads: item idx = (segment, stream, float4 column): 16 bytes each;
  // the 16-channel block lies inside the row (rows hold a multiple of 64
  // bins), channels past n_run are read but never stored
  const int row_items = M * 8;
  float4 pf[XQ_PF];
  auto fetch = [&](int64_t s0, int ns) {
    const int total = ns * row_items;
#pragma unroll
    for (int v = 0; v < XQ_PF; v++) {
      const int idx = threadIdx.x + v * (64 * XQ_WAVES);
      float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (idx < total) {
        const int sl = idx / row_items, r = idx - sl * row_items;
        const int st = r >> 3, q = r & 7;
        x = *reinterpret_cast<const float4 *>(spectra + st * stream_stride + (s0 + sl - seg_base) * seg_stride +
                                              ch0 + 2 * q);
      }
      pf[v] = x;
    }
  };
  int64_t s0 = s_begin;
  int ns = (int)((s_end - s0) < sps ? (s_end - s0) : sps);
  if (s0 < s_end) fetch(s0, ns);
  while (s0 < s_end) {
    const int total = ns * row_items;
#pragma unroll
    for (int v = 0; v < XQ_PF; v++) {
      const int idx = threadIdx.x + v * (64 * XQ_WAVES);
      if (idx < total) xq[idx] = pf[v];
    }
    __syncthreads();
    const int64_t s1 = s0 + ns;
    const int ns1 = (int)((s_end - s1) < sps ? (s_end - s1) : sps);
    if (s1 < s_end) fetch(s1, ns1);
#pragma unroll 1
    for (int sl = 0; sl < ns; sl++) {
      const float2 *row = xq2 + sl * M * 16 + c;
#pragma unroll
      for (int j = 0; j < TPS; j++) {
        if (!act[j]) continue;
        float2 x[XT], y[XT];
#pragma unroll
        for (int i = 0; i < XT; i++) {
          x[i] = a0[j] + i < M ? row[(a0[j] + i) * 16] : make_float2(0.0f, 0.0f);
          y[i] = b0[j] + i < M ? row[(b0[j] + i) * 16] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int i = 0; i < XT; i++)
#pragma unroll
          for (int m = 0; m < XT; m++) {
            const float pr = x[i].x * y[m].x + x[i].y * y[m].y;
            const float pi = x[i].y * y[m].x - x[i].x * y[m].y;
            acc[j][XT * i + m].x = acc[j][XT * i + m].x + pr;
            acc[j][XT * i + m].y = acc[j][XT * i + m].y + pi;
          }
      }
    }
    __syncthreads();
    s0 = s1;
    ns = ns1;
  }
  const int ch = ch0 + c;
  if (ch >= n_run) return;
  float2 *o = partial + (int64_t)k * n_baselines * n_ch + ch;
#pragma unroll
  for (int j = 0; j < TPS; j++) {
    if (!act[j]) continue;
    const int t = (tg * TPS + j) * XQ_SLOTS + slot;
#pragma unroll
    for (int q = 0; q < XT * XT; q++) {
      const int e = tiles[t * (2 + XT * XT) + 2 + q];
      if (e == XT_NONE) continue;
      if (e >= 0) o[(int64_t)e * n_ch] = acc[j][q];
      else o[(int64_t)(~e) * n_ch] = make_float2(acc[j][q].x, -acc[j][q].y);
    }
  }
}

// The last bin (channel ch = N_c, the Nyquist bin) of every baseline and
// sub-integration for the FP32 tile X-step, whose workgroups map 64
// channels to the lanes: N_c + 1 bins would otherwise need one more block
// of workgroups for a single channel. Thread = baseline; the bin of `segs`
// segments of every stream is staged in LDS (one scattered load per thread
// and stream-segment), then each baseline adds its products in segment
// order with the tile kernel's expressions, so the sums are the same bits.
// partial[k * k_stride + b * b_stride].
__global__ __launch_bounds__(256) void k_xstep_last(const float2 *__restrict__ spectra, int64_t stream_stride,
                                                    int64_t seg_stride, int64_t seg_base, int ch, int M, int segs,
                                                    const int32_t *__restrict__ fold_start,
                                                    const int32_t *__restrict__ bl_first,
                                                    const int32_t *__restrict__ bl_second, int n_baselines,
                                                    float2 *__restrict__ partial, int64_t k_stride, int64_t b_stride) {
#pragma clang fp contract(off)
  extern __shared__ float2 col[];  // [segs][M]
  const int k = blockIdx.x;
  const int b = blockIdx.y * 256 + threadIdx.x;
  const int64_t s_begin = fold_start[k], s_end = fold_start[k + 1];
  const int a = b < n_baselines ? bl_first[b] : 0, c = b < n_baselines ? bl_second[b] : 0;
  float2 acc = make_float2(0.0f, 0.0f);
  for (int64_t s0 = s_begin; s0 < s_end; s0 += segs) {
    const int n = (int)(s_end - s0 < segs ? s_end - s0 : segs);
    for (int i = threadIdx.x; i < n * M; i += 256) {
      const int sl = i / M, st = i - sl * M;
      col[i] = spectra[st * stream_stride + (s0 + sl - seg_base) * seg_stride + ch];
    }
    __syncthreads();
    if (b < n_baselines) {
      for (int sl = 0; sl < n; sl++) {
        const float2 x = col[sl * M + a], y = col[sl * M + c];
        const float pr = x.x * y.x + x.y * y.y;
        const float pi = x.y * y.x - x.x * y.y;
        acc.x = acc.x + pr;
        acc.y = acc.y + pi;
      }
    }
    __syncthreads();
  }
  if (b < n_baselines) partial[k * k_stride + b * b_stride] = acc;
}

static hipError_t launch_xstep_last(const float2 *spectra, int64_t stream_stride, int64_t seg_stride,
                                    int64_t seg_base, int ch, int M, const int32_t *fold_start, int n_k,
                                    const int32_t *bl_first, const int32_t *bl_second, int n_baselines,
                                    float2 *partial, int64_t k_stride, int64_t b_stride, hipStream_t s) {
  // up to 64 segments per pass, within 64 KiB of LDS
  const int segs = std::max(1, std::min(64, 8192 / std::max(M, 1)));
  const size_t lds = (size_t)segs * M * sizeof(float2);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  dim3 grid((unsigned)n_k, (unsigned)((n_baselines + 255) / 256));
  hipLaunchKernelGGL(k_xstep_last, grid, dim3(256), lds, s, spectra, stream_stride, seg_stride, seg_base, ch, M,
                     segs, fold_start, bl_first, bl_second, n_baselines, partial, k_stride, b_stride);
  return hipGetLastError();
}

// vis += partial[k] for k = 0 .. n_k - 1, in fold order.
__global__ __launch_bounds__(256) void k_fold_combine(const float2 *__restrict__ partial, int n_k, int64_t n,
                                                      float2 *__restrict__ vis) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 t = vis[i];
  for (int k = 0; k < n_k; k++) {
    const float2 p = partial[k * n + i];
    t.x = t.x + p.x;
    t.y = t.y + p.y;
  }
  vis[i] = t;
}

// ---------------------------------------------------------------------------
// k_xstep_mfma: mixed-precision X-step. For every channel the
// correlation matrix of a sub-integration, V = sum_s x_s x_s^H over the M
// streams' spectra, is a Hermitian rank-k update. With Y = [Re X; Im X]
// (2M x k) it is the real Gram matrix G = Y Y^T:
//   V[a][b] = G[a][b] + G[M+a][M+b] + i (G[M+a][b] - G[a][M+b]).
// G is computed in 32 x 32 tiles (the upper triangle of the 2M x 2M matrix,
// padded to R = 32 RT rows) on v_mfma_f32_32x32x16 with FP32 accumulation.
// Operands are FP16, BF16, or FP16 split into hi + lo (three products,
// hi.hi + hi.lo + lo.hi). Spectra are scaled by a power of two before the
// conversion (exact; undone when the baselines are formed).
//
// Workgroup: 8 waves, CH channels, one sub-integration (grid z). Stages of
// KS segments are loaded with coalesced reads (consecutive threads take
// consecutive channels, each thread 8 segments), converted, and stored into
// LDS as Y[c][r][k] with 16-byte stores (a per-channel skew keeps them
// conflict-free). The loads of the next stage are issued into registers
// before the current stage's MFMAs and stay in flight across them (the
// barriers order LDS only), so HBM latency hides behind the matrix work.
// Each wave owns up to MF_MAXP (channel, tile) pairs and keeps their
// accumulators for the whole sub-integration. At its end the tiles of each
// channel go through LDS and the workgroup forms that channel's baseline
// visibilities into partial[z][ch][b] (channel-major, so the stores are
// coalesced); k_fold_combine_cm adds the sub-integrations in fold order into
// vis[b][ch].
// ---------------------------------------------------------------------------
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f16v_t __attribute__((ext_vector_type(16)));

constexpr int MF_KS = 16;          // segments per LDS stage (one K = 16 MFMA step)
constexpr int MF_RS = MF_KS + 8;   // row stride in halves (rows 16-byte aligned)
constexpr int MF_MAXP = 10;        // (channel, tile) pairs per wave
constexpr int MF_THREADS = 512;
constexpr int MF_WAVES = MF_THREADS / 64;
constexpr int MF_PF_WORDS = 16;    // registers per thread holding the next stage's loads

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it stay in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// (channel, tile) pair p of a workgroup: channel c and the tile (ti, tj),
// ti <= tj, of the upper triangle; t is the tile's index in row-major order
// of that triangle.
__device__ __forceinline__ void mf_pair(int p, int nT, int RT, int &c, int &t, int &ti, int &tj) {
  c = p / nT;
  t = p - c * nT;
  int r = t;
  ti = 0;
  while (r >= RT - ti) {
    r -= RT - ti;
    ti++;
  }
  tj = ti + r;
}

// Loads segments [s0, s0 + MF_KS) (zeros at and past s_end, and for channels
// past n_ch; spectra rows start at segment seg_base) of CH channels from ch0
// into the LDS planes as Y[c][r][k]:
// consecutive threads take consecutive channels, each thread 8 segments of
// one stream, converted to the MFMA operand type and stored 16 bytes at a time.
// HS: the spectra are already in the operand type (SpecOut), 32 bits per
// bin, and only their halves are regrouped.
template <int PREC, int KS = MF_KS, bool HS = false>
__device__ __forceinline__ void mf_load_stage(const void *__restrict__ spectra_v, int64_t stream_stride,
                                              int64_t seg_stride, int64_t seg_base, int n_ch, int M, int CH,
                                              int ch0, int64_t s0,
                                              int64_t s_end, float scale, int chan_stride, int plane_elems,
                                              _Float16 *ylds) {
  const int n_items = CH * (KS / 8) * M;
  for (int item = threadIdx.x; item < n_items; item += MF_THREADS) {
    const int c = item % CH;
    const int rest = item / CH;
    const int kg = rest % (KS / 8);
    const int st = rest / (KS / 8);
    const int o_re = c * chan_stride + st * (KS + 8) + kg * 8;
    const int o_im = c * chan_stride + (M + st) * (KS + 8) + kg * 8;
    if constexpr (HS) {
      static_assert(PREC != SFXG_PREC_SPLIT_FP16, "split FP16 reads FP32 spectra");
      const uint32_t *sp = static_cast<const uint32_t *>(spectra_v);
      uint32_t w[8];
#pragma unroll
      for (int kk = 0; kk < 8; kk++) {
        const int64_t seg = s0 + kg * 8 + kk;
        w[kk] = 0u;
        if (seg < s_end && ch0 + c < n_ch) w[kk] = sp[st * stream_stride + (seg - seg_base) * seg_stride + ch0 + c];
      }
      typedef uint32_t u4_t __attribute__((ext_vector_type(4)));
      u4_t r, i;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        r[e] = (w[2 * e] & 0xffffu) | (w[2 * e + 1] << 16);
        i[e] = (w[2 * e] >> 16) | (w[2 * e + 1] & 0xffff0000u);
      }
      *reinterpret_cast<u4_t *>(ylds + o_re) = r;
      *reinterpret_cast<u4_t *>(ylds + o_im) = i;
      continue;
    }
    const float2 *spectra = static_cast<const float2 *>(spectra_v);
    float re[8], im[8];
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      const int64_t seg = s0 + kg * 8 + kk;
      float2 v = make_float2(0.0f, 0.0f);
      if (seg < s_end && ch0 + c < n_ch) v = spectra[st * stream_stride + (seg - seg_base) * seg_stride + ch0 + c];
      re[kk] = v.x * scale;
      im[kk] = v.y * scale;
    }
    if (PREC == SFXG_PREC_BF16) {
      b8_t br, bi;
#pragma unroll
      for (int kk = 0; kk < 8; kk++) {
        br[kk] = (__bf16)re[kk];
        bi[kk] = (__bf16)im[kk];
      }
      *reinterpret_cast<b8_t *>(ylds + o_re) = br;
      *reinterpret_cast<b8_t *>(ylds + o_im) = bi;
    } else {
      h8_t hr, hi_;
#pragma unroll
      for (int kk = 0; kk < 8; kk++) {
        hr[kk] = (_Float16)re[kk];
        hi_[kk] = (_Float16)im[kk];
      }
      *reinterpret_cast<h8_t *>(ylds + o_re) = hr;
      *reinterpret_cast<h8_t *>(ylds + o_im) = hi_;
      if (PREC == SFXG_PREC_SPLIT_FP16) {
        h8_t lr, li;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          lr[kk] = (_Float16)(re[kk] - (float)hr[kk]);
          li[kk] = (_Float16)(im[kk] - (float)hi_[kk]);
        }
        *reinterpret_cast<h8_t *>(ylds + plane_elems + o_re) = lr;
        *reinterpret_cast<h8_t *>(ylds + plane_elems + o_im) = li;
      }
    }
  }
}

// acc += A B^T for one 32 x 32 tile and one stage; A rows at LDS offset oa,
// B rows at ob (this lane's 8 segments). keep: bit e set keeps segment
// 8 (lane / 32) + e of the A operand (the others count as zero).
template <int PREC>
__device__ __forceinline__ void mf_mma(f16v_t &acc, const _Float16 *ylds, int oa, int ob, int plane_elems,
                                       unsigned keep) {
  if (PREC == SFXG_PREC_BF16) {
    b8_t a = *reinterpret_cast<const b8_t *>(ylds + oa);
    const b8_t b = *reinterpret_cast<const b8_t *>(ylds + ob);
    if (keep != 0xffu) {
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (!((keep >> e) & 1u)) a[e] = (__bf16)0.0f;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  } else {
    h8_t a = *reinterpret_cast<const h8_t *>(ylds + oa);
    const h8_t b = *reinterpret_cast<const h8_t *>(ylds + ob);
    h8_t al;
    if (PREC == SFXG_PREC_SPLIT_FP16) al = *reinterpret_cast<const h8_t *>(ylds + plane_elems + oa);
    if (keep != 0xffu) {
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (!((keep >> e) & 1u)) {
          a[e] = (_Float16)0.0f;
          if (PREC == SFXG_PREC_SPLIT_FP16) al[e] = (_Float16)0.0f;
        }
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
    if (PREC == SFXG_PREC_SPLIT_FP16) {
      const h8_t bl = *reinterpret_cast<const h8_t *>(ylds + plane_elems + ob);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, b, acc, 0, 0, 0);
    }
  }
}

// 32-bit words one load item (8 segments of one channel and stream) takes:
// packed operand pairs, or FP32 real and imaginary parts.
template <bool HS>
constexpr int mf_item_words() { return HS ? 8 : 16; }

// SFXG_PREC_FP16, SFXG_PREC_BF16, SFXG_PREC_SPLIT_FP16; segments per stage;
// spectra in the operand type (HS) or FP32. The launcher guarantees
// CH * (KS / 8) * M <= MF_THREADS * MF_PF_WORDS / mf_item_words<HS>().
// One-dimensional grid of n_blk channel blocks x n_k sub-integrations,
// regrouped by XCD in groups of G channel blocks.
template <int PREC, int KS, bool HS>
__global__ __launch_bounds__(MF_THREADS) void k_xstep_mfma(
    const void *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base, int n_ch, int M,
    int RT, int CH, int n_blk, int G, const int32_t *__restrict__ fold_start, float scale,
    const int32_t *__restrict__ bl_first, const int32_t *__restrict__ bl_second, int n_baselines, float inv_scale2,
    float2 *__restrict__ partial) {
  extern __shared__ _Float16 ylds[];
  constexpr int planes = PREC == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
  constexpr int W = mf_item_words<HS>(), IPT = MF_PF_WORDS / W;
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = xcd_grouped(blockIdx.x, gridDim.x, G);
  const int ch0 = (L % n_blk) * CH;
  const int z = L / n_blk;
  if (ch0 >= n_ch) return;  // a block that pads n_blk to whole XCD groups
  const int64_t s_begin = fold_start[z], s_end = fold_start[z + 1];
  const int chan_stride = R * (KS + 8) + 8;  // halves, skewed by 16 bytes per channel
  const int plane_elems = CH * chan_stride;
  // this thread's load items j: channel c, segment group kg, stream st
  // (element offset from the stage's first row, which the launcher keeps
  // below 2^32; LDS offset of the real part's row times 64 plus the first
  // segment in the stage, KS for an item that loads nothing; negative for
  // an item past the end)
  const int n_items = CH * (KS / 8) * M;
  uint32_t goff[IPT];
  int lk[IPT];
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    const int item = threadIdx.x + j * MF_THREADS;
    const int c = item % CH, rest = item / CH;
    const int kg = rest % (KS / 8), st = rest / (KS / 8);
    goff[j] = (uint32_t)(st * stream_stride + (int64_t)kg * 8 * seg_stride + ch0 + c);
    lk[j] = item < n_items ? (c * chan_stride + st * (KS + 8) + kg * 8) * 64 + (ch0 + c < n_ch ? kg * 8 : KS) : -1;
  }
  for (int i = threadIdx.x; i < planes * plane_elems; i += MF_THREADS) ylds[i] = (_Float16)0.0f;
  f16v_t acc[MF_MAXP];
#pragma unroll
  for (int q = 0; q < MF_MAXP; q++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[q][e] = 0.0f;
  uint32_t pf[IPT][W];
  // global loads of the stage at s0 into pf (zeros past s_end and n_ch)
  auto fetch = [&](int64_t s0) {
    const int nv = (int)(s_end - s0 < KS ? s_end - s0 : KS);
    const int64_t row = (s0 - seg_base) * seg_stride;
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const int kg8 = lk[j] < 0 ? KS : lk[j] & 63;
#pragma unroll
      for (int kk = 0; kk < 8; kk++) {
        const bool ok = kg8 + kk < nv;
        const int64_t idx = row + goff[j] + kk * seg_stride;
        if constexpr (HS) {
          pf[j][kk] = ok ? gmem(static_cast<const uint32_t *>(spectra))[idx] : 0u;
        } else {
          const float2 v = ok ? gload2(static_cast<const float2 *>(spectra) + idx) : make_float2(0.0f, 0.0f);
          pf[j][2 * kk] = __float_as_uint(v.x);
          pf[j][2 * kk + 1] = __float_as_uint(v.y);
        }
      }
    }
  };
  // pf into the LDS planes, converted to the operand type
  auto put = [&]() {
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (lk[j] < 0) continue;
      const int o_re = lk[j] >> 6, o_im = o_re + M * (KS + 8);
      if constexpr (HS) {
        static_assert(PREC != SFXG_PREC_SPLIT_FP16, "split FP16 reads FP32 spectra");
        typedef uint32_t u4_t __attribute__((ext_vector_type(4)));
        u4_t r, i;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          r[e] = (pf[j][2 * e] & 0xffffu) | (pf[j][2 * e + 1] << 16);
          i[e] = (pf[j][2 * e] >> 16) | (pf[j][2 * e + 1] & 0xffff0000u);
        }
        *reinterpret_cast<u4_t *>(ylds + o_re) = r;
        *reinterpret_cast<u4_t *>(ylds + o_im) = i;
      } else {
        float re[8], im[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          re[kk] = __uint_as_float(pf[j][2 * kk]) * scale;
          im[kk] = __uint_as_float(pf[j][2 * kk + 1]) * scale;
        }
        if constexpr (PREC == SFXG_PREC_BF16) {
          b8_t br, bi;
#pragma unroll
          for (int kk = 0; kk < 8; kk++) {
            br[kk] = (__bf16)re[kk];
            bi[kk] = (__bf16)im[kk];
          }
          *reinterpret_cast<b8_t *>(ylds + o_re) = br;
          *reinterpret_cast<b8_t *>(ylds + o_im) = bi;
        } else {
          h8_t hr, hi_;
#pragma unroll
          for (int kk = 0; kk < 8; kk++) {
            hr[kk] = (_Float16)re[kk];
            hi_[kk] = (_Float16)im[kk];
          }
          *reinterpret_cast<h8_t *>(ylds + o_re) = hr;
          *reinterpret_cast<h8_t *>(ylds + o_im) = hi_;
          if constexpr (PREC == SFXG_PREC_SPLIT_FP16) {
            h8_t lr, li;
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
              lr[kk] = (_Float16)(re[kk] - (float)hr[kk]);
              li[kk] = (_Float16)(im[kk] - (float)hi_[kk]);
            }
            *reinterpret_cast<h8_t *>(ylds + plane_elems + o_re) = lr;
            *reinterpret_cast<h8_t *>(ylds + plane_elems + o_im) = li;
          }
        }
      }
    }
  };
  const int n_pairs = CH * nT;
  if (s_begin < s_end) fetch(s_begin);
  __syncthreads();  // LDS zeroed
  for (int64_t s0 = s_begin; s0 < s_end; s0 += KS) {
    put();
    lds_barrier();
    if (s0 + KS < s_end) fetch(s0 + KS);
    const int wv = opaque(wave);  // pair offsets recomputed per stage, not held in registers
#pragma unroll
    for (int q = 0; q < MF_MAXP; q++) {
      const int p = wv + MF_WAVES * q;
      if (p < n_pairs) {
        int c, t, ti, tj;
        mf_pair(p, nT, RT, c, t, ti, tj);
        const int ko = 8 * (lane >> 5);
        const int oa = c * chan_stride + (ti * 32 + (lane & 31)) * (KS + 8) + ko;
        const int ob = c * chan_stride + (tj * 32 + (lane & 31)) * (KS + 8) + ko;
#pragma unroll
        for (int kk = 0; kk < KS; kk += 16) mf_mma<PREC>(acc[q], ylds, oa + kk, ob + kk, plane_elems, 0xffu);
      }
    }
    lds_barrier();
  }
  // baselines, one channel at a time through an R x R FP32 buffer in LDS
  float *g = reinterpret_cast<float *>(ylds);
  for (int c = 0; c < CH; c++) {
#pragma unroll
    for (int q = 0; q < MF_MAXP; q++) {
      const int p = wave + MF_WAVES * q;
      if (p < n_pairs && p / nT == c) {
        int cc, t, ti, tj;
        mf_pair(p, nT, RT, cc, t, ti, tj);
        const int col = tj * 32 + (lane & 31);
        // D[8 i + 4 (lane / 32) + j][lane % 32] = acc[4 i + j]
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) g[(ti * 32 + 8 * i + 4 * (lane >> 5) + j) * R + col] = acc[q][4 * i + j];
      }
    }
    __syncthreads();
    if (ch0 + c < n_ch) {
      float2 *out = partial + ((int64_t)z * n_ch + ch0 + c) * n_baselines;
      for (int b = threadIdx.x; b < n_baselines; b += MF_THREADS) {
        const int a = bl_first[b], d = bl_second[b];
        auto G = [&](int r, int s2) { return r <= s2 ? g[r * R + s2] : g[s2 * R + r]; };
        const float re = (G(a, d) + G(M + a, M + d)) * inv_scale2;
        const float im = (G(M + a, d) - G(a, M + d)) * inv_scale2;
        out[b] = make_float2(re, im);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// k_xstep_mfma_dma: the FP16 / BF16 X-step on packed spectra with the
// spectra fetched by direct global-to-LDS loads (global_load_lds_dwordx4:
// 16 bytes, four channels, per lane), so that two stages are in flight while
// a third is regrouped and multiplied, at no register cost. Stages are 16
// segments. A stage's M x 16 rows (stream-major) land in a ring slot in
// 1 KiB blocks, one load instruction each (256 / CH rows of CH channels),
// padded to mfd_blk(CH) bytes so that the regrouping reads of a wave fall in
// distinct banks. Each wave issues exactly NPW block loads per stage (blocks
// past the stage's count go to a dummy block), which lets it wait for its
// own loads of a stage with a fixed vmcnt. The regrouping pass turns the raw
// rows into the operand layout of k_xstep_mfma (zeros past n_ch and past the
// sub-integration's end); the MFMA work, the XCD grouping and the epilogue
// are those of k_xstep_mfma.
// ---------------------------------------------------------------------------
constexpr int MFD_KS = 16;          // segments per stage
constexpr int MFD_SLOTS = 3;        // ring slots

// Bytes per raw block of rows of CW 32-bit words: 1 KiB of data, then 4 CW
// bytes of skew so that consecutive blocks start CW banks apart.
__host__ __device__ constexpr int mfd_blk(int CW) { return 1024 + (CW < 64 ? 4 * CW : 0); }
// Block loads per stage for M streams.
__host__ __device__ constexpr int mfd_blocks(int M, int CW) { return (M * MFD_KS * CW + 255) / 256; }

// LDS bytes of the DMA X-step with rows of CW words (CW / W channels of W
// operand planes each): ring, operand stage, dummy block; the epilogue's
// R x R FP32 buffer reuses the ring.
__host__ __device__ constexpr size_t mfd_lds_bytes(int npw, int CW, int R) {
  return (size_t)MFD_SLOTS * npw * MF_WAVES * mfd_blk(CW) + (size_t)CW * (R * (MFD_KS + 8) + 8) * 2 + 1024;
}

// Workgroup barrier over LDS, with the waits written out: the compiler's
// own LDS fence would also wait for every direct-to-LDS load in flight.
__device__ __forceinline__ void dma_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// PREC: FP16 or BF16 on packed pairs (one 32-bit word per bin), or split
// FP16 on hi and lo pairs (two words per bin, SpecOut). CW: 32-bit words
// per row of the workgroup's channels (CH = CW / W channels). Strides are in
// bins.
template <int PREC, int NPW, int CW>
__global__ __launch_bounds__(MF_THREADS) void k_xstep_mfma_dma(
    const uint32_t *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base, int n_ch,
    int M, int RT, int n_blk, int G, const int32_t *__restrict__ fold_start, const int32_t *__restrict__ bl_first,
    const int32_t *__restrict__ bl_second, int n_baselines, float inv_scale2, float2 *__restrict__ partial) {
  static_assert(PREC == SFXG_PREC_FP16 || PREC == SFXG_PREC_BF16 || PREC == SFXG_PREC_SPLIT_FP16,
                "packed operand pairs");
  static_assert(CW == 8 || CW == 16 || CW == 32 || CW == 64, "words per row");
  constexpr int W = PREC == SFXG_PREC_SPLIT_FP16 ? 2 : 1;  // words per bin, operand planes
  constexpr int CH = CW / W;                 // channels per workgroup
  constexpr int KS = MFD_KS;
  constexpr int CQ = CW / 4;                 // lanes per row
  constexpr int RPX = 64 / CQ;               // rows per block
  constexpr int BLK = mfd_blk(CW);
  constexpr int SLOT = NPW * MF_WAVES * BLK;
  extern __shared__ __attribute__((aligned(16))) char dlds[];
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int L = xcd_grouped(blockIdx.x, gridDim.x, G);
  const int ch0 = (L % n_blk) * CH;
  const int z = L / n_blk;
  if (ch0 >= n_ch) return;  // a block that pads n_blk to whole XCD groups
  const int64_t s_begin = fold_start[z], s_end = fold_start[z + 1];
  const int chan_stride = R * (KS + 8) + 8;  // halves
  char *ring = dlds;
  _Float16 *ylds = reinterpret_cast<_Float16 *>(dlds + MFD_SLOTS * SLOT);
  const int plane_elems = CH * chan_stride;     // halves per operand plane
  char *dummy = dlds + MFD_SLOTS * SLOT + W * plane_elems * 2;
  const int NI = mfd_blocks(M, CW);             // real blocks per stage
  const bool regular = NI * RPX == M * KS;      // every block inside the M streams
  const bool all_ch = ch0 + CH <= n_ch;
  // lane part of a block load: row r = lane / CQ of the block (stream
  // r / 16, segment r % 16 from the block's first row), channels
  // ch0 + 4 (lane % CQ) .. + 3, in bytes. Rows hold S >= n_blk CH bins, so
  // the four channels stay inside the row (bins past n_ch are masked later).
  const int lr = lane / CQ, lq = lane % CQ;
  const uint32_t lane_boff =
      (uint32_t)((((lr / KS) * stream_stride + (lr % KS) * seg_stride + ch0) * W + 4 * lq) * 4);
  // issue this wave's NPW block loads of the stage at s0 into ring slot sl:
  // block j = wave + 8 q holds rows j RPX .. j RPX + RPX - 1 of the stage;
  // a block past the stage's count reads the buffer's first KiB into the
  // dummy block. In a full stage of a regular shape, block q of the wave
  // starts RPX / 2 q streams after block 0, so its wave-uniform base is one
  // scalar add away; otherwise (the last stage, or an odd stream count with
  // 32-row blocks) every lane clamps its stream and row.
  auto dma = [](const void *src, char *dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                     (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
  };
  auto issue = [&](int64_t s0, int sl) {
    const int64_t row0 = s0 - seg_base;
    const int wv = opaque_s(wave);  // per-block bases recomputed per stage, not held in registers
    char *dst0 = ring + sl * SLOT;
    const char *dsrc = reinterpret_cast<const char *>(spectra) + lane * 16;
    if (regular && s0 + KS <= s_end) {
      const int rj = wv * RPX;
      const char *b0 = reinterpret_cast<const char *>(spectra + ((rj / KS) * stream_stride +
                                                                 (row0 + rj % KS) * seg_stride) * W);
#pragma unroll
      for (int q = 0; q < NPW; q++) {
        const int j = wv + MF_WAVES * q;
        const char *bq = b0 + (int64_t)q * (RPX / 2) * stream_stride * W * 4;
        const bool real = j < NI;
        dma(real ? bq + lane_boff : dsrc, real ? dst0 + j * BLK : dummy);
      }
    } else {
      const int64_t last = s_end - 1 - seg_base;
#pragma unroll
      for (int q = 0; q < NPW; q++) {
        const int j = wv + MF_WAVES * q;
        const int rho = j * RPX + lr;
        const int st = rho / KS < M ? rho / KS : M - 1;
        int64_t row = row0 + rho % KS;
        row = row < last ? row : last;
        const bool real = j < NI;
        dma(real ? spectra + (st * stream_stride + row * seg_stride + ch0) * W + 4 * lq : (const uint32_t *)dsrc,
            real ? dst0 + j * BLK : dummy);
      }
    }
  };
  // raw slot sl into the operand planes: item (c, group gm = 2 st + kg)
  // takes the 8 rows of its group, real halves to row st, imaginary to row
  // M + st; zeros past n_ch and past s_end (last stage only). Items are
  // ordered channel first, then block, so that a wave's reads fall in
  // distinct banks.
  constexpr int GPI = RPX >= 8 ? RPX / 8 : 1;        // groups per block
  const int NJ = (2 * M + GPI - 1) / GPI;            // blocks holding groups
  const int n_items = CH * NJ * GPI;
  auto regroup = [&](int64_t s0, int sl) {
    const int nv = (int)(s_end - s0 < KS ? s_end - s0 : KS);
    const bool full = all_ch && nv == KS;
    const char *slot = ring + sl * SLOT;
    for (int item = threadIdx.x; item < n_items; item += MF_THREADS) {
      const int c = item % CH, rest = item / CH;
      const int jj = rest % NJ, sub = rest / NJ;
      const int gm = jj * GPI + sub;
      if (gm >= 2 * M) continue;
      const char *src;
      if constexpr (RPX >= 8) src = slot + jj * BLK + sub * 8 * (CW * 4) + c * W * 4;
      else src = slot + gm * (8 / RPX) * BLK + c * W * 4;  // a group spans 8 / RPX blocks
      const int st = gm >> 1, kg = gm & 1;
      const int o_re = c * chan_stride + st * (KS + 8) + kg * 8;
#pragma unroll
      for (int pl = 0; pl < W; pl++) {  // hi (or the only) pair, then lo
        uint32_t w[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          if constexpr (RPX >= 8) w[kk] = *reinterpret_cast<const uint32_t *>(src + pl * 4 + kk * (CW * 4));
          else
            w[kk] = *reinterpret_cast<const uint32_t *>(src + pl * 4 + (kk / RPX) * BLK + (kk % RPX) * (CW * 4));
        }
        if (!full) {
          const bool cok = ch0 + c < n_ch;
#pragma unroll
          for (int kk = 0; kk < 8; kk++) w[kk] = cok && kg * 8 + kk < nv ? w[kk] : 0u;
        }
        typedef uint32_t u4_t __attribute__((ext_vector_type(4)));
        u4_t re, im;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          re[e] = __builtin_amdgcn_perm(w[2 * e + 1], w[2 * e], 0x05040100u);  // low halves
          im[e] = __builtin_amdgcn_perm(w[2 * e + 1], w[2 * e], 0x07060302u);  // high halves
        }
        *reinterpret_cast<u4_t *>(ylds + pl * plane_elems + o_re) = re;
        *reinterpret_cast<u4_t *>(ylds + pl * plane_elems + o_re + M * (KS + 8)) = im;
      }
    }
  };
  // padding rows (2M .. R) stay zero
  for (int i = threadIdx.x; i < W * plane_elems; i += MF_THREADS) ylds[i] = (_Float16)0.0f;
  f16v_t acc[MF_MAXP];
#pragma unroll
  for (int q = 0; q < MF_MAXP; q++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[q][e] = 0.0f;
  // the wave's (channel, tile) pairs: operand offsets, wave-uniform part
  const int n_pairs = CH * nT;
  const int lane_ab = (lane & 31) * (KS + 8) + 8 * (lane >> 5);
  int pa[MF_MAXP], pb[MF_MAXP];
#pragma unroll
  for (int q = 0; q < MF_MAXP; q++) {
    const int p = wave + MF_WAVES * q;
    pa[q] = pb[q] = -1;
    if (p < n_pairs) {
      int c, t, ti, tj;
      mf_pair(p, nT, RT, c, t, ti, tj);
      pa[q] = c * chan_stride + ti * 32 * (KS + 8);
      pb[q] = c * chan_stride + tj * 32 * (KS + 8);
    }
  }
  const int n_stages = (int)((s_end - s_begin + KS - 1) / KS);
  if (n_stages > 0) issue(s_begin, 0);
  if (n_stages > 1) issue(s_begin + KS, 1);
  for (int k = 0; k < n_stages; k++) {
    const int64_t s0 = s_begin + (int64_t)k * KS;
    // stage k + 2 into the slot stage k - 1 used (regrouped before the
    // previous iteration's second barrier)
    if (k + 2 < n_stages) {
      issue(s0 + 2 * KS, (k + 2) % MFD_SLOTS);
      wait_vmcnt<2 * NPW>();
    } else if (k + 1 < n_stages) {
      wait_vmcnt<NPW>();
    } else {
      wait_vmcnt<0>();
    }
    dma_barrier();  // every wave's blocks of stage k are in LDS; the previous MFMAs are done
    regroup(s0, k % MFD_SLOTS);
    dma_barrier();
#pragma unroll
    for (int q = 0; q < MF_MAXP; q++)
      if (pa[q] >= 0) mf_mma<PREC>(acc[q], ylds, pa[q] + lane_ab, pb[q] + lane_ab, plane_elems, 0xffu);
  }
  __syncthreads();
  // baselines, one channel at a time through an R x R FP32 buffer in LDS
  float *gb = reinterpret_cast<float *>(dlds);
  for (int c = 0; c < CH; c++) {
#pragma unroll
    for (int q = 0; q < MF_MAXP; q++) {
      const int p = wave + MF_WAVES * q;
      if (p < n_pairs && p / nT == c) {
        int cc, t, ti, tj;
        mf_pair(p, nT, RT, cc, t, ti, tj);
        const int col = tj * 32 + (lane & 31);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) gb[(ti * 32 + 8 * i + 4 * (lane >> 5) + j) * R + col] = acc[q][4 * i + j];
      }
    }
    __syncthreads();
    if (ch0 + c < n_ch) {
      float2 *out = partial + ((int64_t)z * n_ch + ch0 + c) * n_baselines;
      for (int b = threadIdx.x; b < n_baselines; b += MF_THREADS) {
        const int a = bl_first[b], d = bl_second[b];
        auto Gm = [&](int r, int s2) { return r <= s2 ? gb[r * R + s2] : gb[s2 * R + r]; };
        const float re = (Gm(a, d) + Gm(M + a, M + d)) * inv_scale2;
        const float im = (Gm(M + a, d) - Gm(a, M + d)) * inv_scale2;
        out[b] = make_float2(re, im);
      }
    }
    __syncthreads();
  }
}

// vis[b][ch] += partial[k][ch][b] for k = 0 .. n_k - 1 in order: the
// channel-major partials of the MFMA X-step, transposed through LDS in
// 64 x 64 tiles so that both sides are read and written coalesced.
__global__ __launch_bounds__(256) void k_fold_combine_cm(const float2 *__restrict__ partial, int n_k, int n_ch,
                                                         int n_baselines, float2 *__restrict__ vis) {
#pragma clang fp contract(off)
  __shared__ float2 tile[64][65];  // [baseline][channel]
  const int b0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int b = b0 + r, c = c0 + tx;
    tile[r][tx] = b < n_baselines && c < n_ch ? vis[(int64_t)b * n_ch + c] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int c = c0 + r, b = b0 + tx;
    if (b < n_baselines && c < n_ch) {
      float2 t = tile[tx][r];
      for (int k = 0; k < n_k; k++) {
        const float2 p = partial[((int64_t)k * n_ch + c) * n_baselines + b];
        t.x = t.x + p.x;
        t.y = t.y + p.y;
      }
      tile[tx][r] = t;
    }
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int b = b0 + r, c = c0 + tx;
    if (b < n_baselines && c < n_ch) vis[(int64_t)b * n_ch + c] = tile[r][tx];
  }
}

// ---------------------------------------------------------------------------
// k_xstep_pulsar: one thread per (baseline, channel); per segment its
// product goes to the bin its pulse phase selects (from the chunk's bin
// table, k_pulsar_bin_table). Runs of equal bins accumulate in registers;
// at a bin change the running sum is written back and the new bin's loaded,
// which keeps the reference's per-segment addition order.
// ---------------------------------------------------------------------------
constexpr int XWAVES = 4;  // baselines per workgroup of the pulsar kernel (one per wave)

// Pulse-phase bin of segment sgi in a channel with dispersion offset off_ch
// (get_phase and the gate test of Correlation_core_pulsar::integration_step,
// correlation_core_pulsar.h:40-44, correlation_core_pulsar.cc:188-208):
// 1 + floor((phase - gate_begin) (nbins - 1) / gate length) inside the gate,
// which may wrap through phase 1, else 0. FP64 without contraction, in the
// reference's operation order.
__device__ __forceinline__ int pulsar_bin(int64_t sgi, double off_ch, double start_phase, double phase_coef,
                                          double fft_duration, double gate_begin, double gate_end, int nbins) {
#pragma clang fp contract(off)
  const int64_t us_per_day = (int64_t)24 * 60 * 60 * 1000000;
  const double len = gate_end - gate_begin;
  const double obs = start_phase + phase_coef * fft_duration * (double)(int)sgi * 1440 / (double)us_per_day;
  double phase = obs - off_ch;
  phase = phase - floor(phase);
  if (phase >= gate_begin) return phase < gate_end ? (int)((phase - gate_begin) * (nbins - 1) / len) + 1 : 0;
  if (phase + 1 < gate_end) return (int)((phase + 1 - gate_begin) * (nbins - 1) / len) + 1;
  return 0;
}

// The bin of every (segment, channel) of a chunk into bins[segment][channel]
// (pulsar_bin, once per pair instead of once per baseline group).
__global__ __launch_bounds__(256) void k_pulsar_bin_table(int16_t *__restrict__ bins, int n_ch, int64_t seg_base,
                                                          int64_t n_segments, const double *__restrict__ offsets,
                                                          double start_phase, double phase_coef, double fft_duration,
                                                          double gate_begin, double gate_end, int nbins) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_segments * n_ch) return;
  const int64_t sl = i / n_ch;
  const int ch = (int)(i - sl * n_ch);
  bins[i] = (int16_t)pulsar_bin(seg_base + sl, offsets[ch], start_phase, phase_coef, fft_duration, gate_begin,
                                gate_end, nbins);
}

__global__ __launch_bounds__(64 * XWAVES) void k_xstep_pulsar(
    const float2 *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int n_ch, int64_t seg_base,
    int64_t n_segments, const int32_t *__restrict__ bl_first, const int32_t *__restrict__ bl_second,
    int n_baselines, const int16_t *__restrict__ bins, int n_bg, float2 *__restrict__ acc_out) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // (baseline group, channel block), baseline groups innermost: the groups
  // of one channel block read the same spectra and run on one XCD
  const int L = xcd_grouped(blockIdx.x, gridDim.x, n_bg);
  const int ch = (L / n_bg) * 64 + lane;
  const int b = (L % n_bg) * XWAVES + wave;
  if (ch >= n_ch || b >= n_baselines) return;
  const float2 *pa = spectra + bl_first[b] * stream_stride + ch;
  const float2 *pb = spectra + bl_second[b] * stream_stride + ch;
  const int16_t *pbin = bins + ch;
  float2 acc = make_float2(0.0f, 0.0f);
  float2 *out = acc_out + (int64_t)b * n_ch + ch;
  const int64_t bin_stride = (int64_t)n_baselines * n_ch;
  // segment i + 1 and i + 2 (spectra and bin) are in flight while segment
  // i's product is added (the last segment's again past the end, so that
  // the loads need no branch)
  auto at = [&](int64_t i) { return i < n_segments ? i : n_segments - 1; };
  float2 x0 = pa[0], y0 = pb[0];
  int k0 = pbin[0];
  float2 x1 = pa[at(1) * seg_stride], y1 = pb[at(1) * seg_stride];
  int k1 = pbin[at(1) * n_ch];
  int cur = -1;
  for (int64_t i = 0; i < n_segments; i++) {
    const int64_t i2 = at(i + 2);
    const float2 x2 = pa[i2 * seg_stride], y2 = pb[i2 * seg_stride];
    const int k2 = pbin[i2 * n_ch];
    if (k0 != cur) {
      if (cur >= 0) out[cur * bin_stride] = acc;
      acc = out[k0 * bin_stride];
      cur = k0;
    }
    const float pr = x0.x * y0.x + x0.y * y0.y;
    const float pi = x0.y * y0.x - x0.x * y0.y;
    acc.x = acc.x + pr;
    acc.y = acc.y + pi;
    x0 = x1;
    y0 = y1;
    k0 = k1;
    x1 = x2;
    y1 = y2;
    k1 = k2;
  }
  if (cur >= 0) out[cur * bin_stride] = acc;
}

// ---------------------------------------------------------------------------
// k_xstep_pulsar_bin: the FP32 pulsar X-step by bin, on the register tiles
// of k_xstep_tile16 (16 channels per workgroup, lane = tile slot x channel).
// A workgroup takes one bin of one channel block and walks, in order, the
// 16-segment stages where the bin occurs (per-stage bin masks, compacted a
// window at a time); each lane adds the products of its channel's segments
// of that bin. Every (bin, baseline, channel) thus still adds its products
// one segment after the other, from the sum the previous chunks left, as
// the reference does, and its sum is read and written once per chunk
// instead of at every bin change. NW waves per workgroup, TPS tiles per slot.
// ---------------------------------------------------------------------------
template <int NW, int TPS>
__global__ __launch_bounds__(64 * NW) void k_xstep_pulsar_bin(
    const float2 *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int n_ch, int64_t n_segments,
    int M, const int32_t *__restrict__ tiles, int n_tiles, int n_baselines, const int16_t *__restrict__ bins,
    const uint32_t *__restrict__ masks, int W, int n_tg, float2 *__restrict__ acc_out) {
#pragma clang fp contract(off)
  constexpr int SLOTS = 4 * NW, NT = 64 * NW;
  extern __shared__ float4 xq[];                                            // [MF_KS][M][8]
  int *sb = reinterpret_cast<int *>(xq + (size_t)MF_KS * M * 8);           // [MF_KS][16]: bin per (segment, channel)
  int *list = sb + MF_KS * 16;                                             // [NT]
  int *wcnt = list + NT;                                                   // [NW]
  const float2 *xq2 = reinterpret_cast<const float2 *>(xq);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15;
  const int slot = wave * 4 + (lane >> 4);
  const int tg = blockIdx.x % n_tg, cb = blockIdx.x / n_tg, b = blockIdx.y;
  const int n_cb = gridDim.x / n_tg;
  const int ch0 = cb * 16, ch = ch0 + c;
  const bool own = ch < n_ch;
  int a0[TPS], b0[TPS];
  bool act[TPS];
#pragma unroll
  for (int j = 0; j < TPS; j++) {
    const int t = (tg * TPS + j) * SLOTS + slot;
    act[j] = t < n_tiles;
    a0[j] = act[j] ? tiles[t * (2 + XT * XT)] : 0;
    b0[j] = act[j] ? tiles[t * (2 + XT * XT) + 1] : 0;
  }
  float2 acc[TPS][XT * XT];
  // the running sums of the bin from acc_out (store = false) or to it
  auto sums = [&](bool store) {
    float2 *o = acc_out + (int64_t)b * n_baselines * n_ch + ch;
#pragma unroll
    for (int j = 0; j < TPS; j++) {
      const int t = (tg * TPS + j) * SLOTS + slot;
#pragma unroll
      for (int q = 0; q < XT * XT; q++) {
        const int e = act[j] && own ? tiles[t * (2 + XT * XT) + 2 + q] : XT_NONE;
        if (e == XT_NONE) {
          if (!store) acc[j][q] = make_float2(0.0f, 0.0f);
          continue;
        }
        float2 *p = o + (int64_t)(e >= 0 ? e : ~e) * n_ch;
        if (store) {
          *p = e >= 0 ? acc[j][q] : make_float2(acc[j][q].x, -acc[j][q].y);
        } else {
          const float2 v = *p;
          acc[j][q] = e >= 0 ? v : make_float2(v.x, -v.y);
        }
      }
    }
  };
  sums(false);
  const int64_t n_stages = (n_segments + MF_KS - 1) / MF_KS;
  const int row_items = M * 8;
  // stage item idx: segment sl, stream st, channels ch0 + 2 q, + 1 (16 bytes)
  auto item = [&](int64_t seg0, int idx) {
    const int sl = idx / row_items, r = idx - sl * row_items;
    const int st = r >> 3, q = r & 7;
    float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (seg0 + sl < n_segments)
      x = *reinterpret_cast<const float4 *>(spectra + st * stream_stride + (seg0 + sl) * seg_stride + ch0 + 2 * q);
    return x;
  };
  constexpr int PFN = 8;  // items per thread loaded one stage ahead
  float4 pf[PFN];
  int pbin = -1;
  auto fetch = [&](int64_t seg0) {
#pragma unroll
    for (int k = 0; k < PFN; k++) {
      const int idx = threadIdx.x + k * NT;
      pf[k] = idx < MF_KS * row_items ? item(seg0, idx) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const int sl = threadIdx.x >> 4, cc = threadIdx.x & 15;
    pbin = (int)threadIdx.x < MF_KS * 16 && seg0 + sl < n_segments && ch0 + cc < n_ch
               ? bins[(seg0 + sl) * n_ch + ch0 + cc] : -1;
  };
  for (int64_t w0 = 0; w0 < n_stages; w0 += NT) {
    const int64_t g = w0 + threadIdx.x;
    const bool has = g < n_stages && ((masks[(g * n_cb + cb) * W + (b >> 5)] >> (b & 31)) & 1u);
    const uint64_t bal = __ballot(has);
    __syncthreads();  // the previous window's list is no longer read
    if (lane == 0) wcnt[wave] = __popcll(bal);
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
      base += w < wave ? wcnt[w] : 0;
      total += wcnt[w];
    }
    if (has) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = (int)(g - w0);
    __syncthreads();
    if (total > 0) fetch((w0 + list[0]) * MF_KS);
    for (int e = 0; e < total; e++) {
      const int64_t seg0 = (w0 + list[e]) * MF_KS;
      // this stage's items (the first PFN per thread loaded one listed stage
      // ahead, the rest now) and bins into LDS, then the next stage's loads
#pragma unroll
      for (int k = 0; k < PFN; k++) {
        const int idx = threadIdx.x + k * NT;
        if (idx < MF_KS * row_items) xq[idx] = pf[k];
      }
      for (int idx = threadIdx.x + PFN * NT; idx < MF_KS * row_items; idx += NT) xq[idx] = item(seg0, idx);
      if ((int)threadIdx.x < MF_KS * 16) sb[threadIdx.x] = pbin;
      if (e + 1 < total) fetch((w0 + list[e + 1]) * MF_KS);
      __syncthreads();
#pragma unroll 1
      for (int sl = 0; sl < MF_KS; sl++) {
        const bool mine = sb[sl * 16 + c] == b;
        if (!__any(mine)) continue;
        const float2 *row = xq2 + sl * M * 16 + c;
#pragma unroll
        for (int j = 0; j < TPS; j++) {
          if (!act[j]) continue;
          float2 x[XT], y[XT];
#pragma unroll
          for (int i = 0; i < XT; i++) {
            x[i] = a0[j] + i < M ? row[(a0[j] + i) * 16] : make_float2(0.0f, 0.0f);
            y[i] = b0[j] + i < M ? row[(b0[j] + i) * 16] : make_float2(0.0f, 0.0f);
          }
#pragma unroll
          for (int i = 0; i < XT; i++)
#pragma unroll
            for (int m = 0; m < XT; m++) {
              const float pr = x[i].x * y[m].x + x[i].y * y[m].y;
              const float pi = x[i].y * y[m].x - x[i].x * y[m].y;
              const float2 s2 = make_float2(acc[j][XT * i + m].x + pr, acc[j][XT * i + m].y + pi);
              acc[j][XT * i + m] = mine ? s2 : acc[j][XT * i + m];
            }
        }
      }
      __syncthreads();
    }
  }
  sums(true);
}

// ---------------------------------------------------------------------------
// k_xstep_mfma_pulsar: the mixed-precision X-step with pulsar binning. Same
// stages and (channel, tile) pairs as k_xstep_mfma; the grid's z slices are
// ranges of segments. Each stage the workgroup evaluates the bin of its
// channels' segments (pulsar_bin, FP64) into LDS. A pair keeps the Gram tile
// of the bin its channel is in; where the bin changes inside a stage the
// stage is split into runs of equal bin (the other segments masked out of
// the A operand), and the finished tile is added into
// gram[bin][ch][tile][32][32] with atomic adds: every (bin, channel) gets
// contributions from several runs and z slices, in no fixed order.
// k_pulsar_gram_baselines_run then forms the baselines of every bin.
// ---------------------------------------------------------------------------
constexpr int MFP_MAXP = 4;  // (channel, tile) pairs per wave with pulsar binning
constexpr int MFP_RUNS = 20;  // run-table entries per channel: count + up to MF_KS runs, padded to 16 bytes

template <int PREC, bool HS>
__global__ __launch_bounds__(MF_THREADS) void k_xstep_mfma_pulsar(
    const void *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int n_ch, int M, int RT,
    int CH, int64_t seg_base, int64_t n_segments, int64_t segs_per_z, float scale, const double *__restrict__ offsets,
    double start_phase, double phase_coef, double fft_duration, double gate_begin, double gate_end, int nbins,
    float *__restrict__ gram) {
  extern __shared__ _Float16 ylds[];
  constexpr int planes = PREC == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ch0 = blockIdx.x * CH;
  const int64_t s_begin = seg_base + (int64_t)blockIdx.z * segs_per_z;
  int64_t s_end = s_begin + segs_per_z;
  if (s_end > seg_base + n_segments) s_end = seg_base + n_segments;
  if (s_begin >= s_end) return;  // uniform over the workgroup
  const int chan_stride = R * MF_RS + 8;
  const int plane_elems = CH * chan_stride;
  int16_t *bins = reinterpret_cast<int16_t *>(ylds + planes * plane_elems);  // [CH][MF_KS]
  // runs of equal bin per channel and stage: runs[c][0] = count, runs[c][1 + r]
  // = bin << 16 | mask of the run's segments (bit k: segment s0 + k); rows of
  // MFP_RUNS entries, 16-byte aligned
  int32_t *runs = reinterpret_cast<int32_t *>(bins + 64 * MF_KS);  // [CH][MFP_RUNS]
  for (int i = threadIdx.x; i < planes * plane_elems; i += MF_THREADS) ylds[i] = (_Float16)0.0f;
  f16v_t acc[MFP_MAXP];
  int cur[MFP_MAXP];
#pragma unroll
  for (int q = 0; q < MFP_MAXP; q++) {
    cur[q] = -1;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[q][e] = 0.0f;
  }
  __syncthreads();
  const int n_pairs = CH * nT;
  auto flush = [&](int q, int c, int t) {
    float *o = gram + (((int64_t)cur[q] * n_ch + ch0 + c) * nT + t) * 1024 + (lane & 31);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        unsafeAtomicAdd(o + (8 * i + 4 * (lane >> 5) + j) * 32, acc[q][4 * i + j]);
        acc[q][4 * i + j] = 0.0f;
      }
  };
  for (int64_t s0 = s_begin; s0 < s_end; s0 += MF_KS) {
    mf_load_stage<PREC, MF_KS, HS>(spectra, stream_stride, seg_stride, seg_base, n_ch, M, CH, ch0, s0, s_end, scale,
                        chan_stride, plane_elems, ylds);
    for (int i = threadIdx.x; i < CH * MF_KS; i += MF_THREADS) {
      const int c = i / MF_KS, k = i % MF_KS;
      const int64_t seg = s0 + k;
      bins[i] = (seg < s_end && ch0 + c < n_ch)
                    ? (int16_t)pulsar_bin(seg, offsets[ch0 + c], start_phase, phase_coef, fft_duration,
                                          gate_begin, gate_end, nbins)
                    : (int16_t)-1;
    }
    __syncthreads();
    // one thread per channel splits its stage into runs of equal bin, in
    // segment order (segments without a bin are in no run), so that every
    // (channel, tile) pair reads its runs instead of scanning the bins
    if (threadIdx.x < CH) {
      const int16_t *bc = bins + threadIdx.x * MF_KS;
      int32_t *rc = runs + threadIdx.x * MFP_RUNS;
      int bv[MF_KS];
#pragma unroll
      for (int k = 0; k < MF_KS; k++) bv[k] = bc[k];
      int n = 0, b = bv[0];
      unsigned m = 1u;
#pragma unroll
      for (int k = 1; k < MF_KS; k++) {
        if (bv[k] != b) {
          if (b >= 0) rc[1 + n++] = (int32_t)(((unsigned)b << 16) | m);
          b = bv[k];
          m = 0;
        }
        m |= 1u << k;
      }
      if (b >= 0) rc[1 + n++] = (int32_t)(((unsigned)b << 16) | m);
      rc[0] = n;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MFP_MAXP; q++) {
      const int p = wave + MF_WAVES * q;
      if (p < n_pairs) {
        int c, t, ti, tj;
        mf_pair(p, nT, RT, c, t, ti, tj);
        const int ko = 8 * (lane >> 5);
        const int oa = c * chan_stride + (ti * 32 + (lane & 31)) * MF_RS + ko;
        const int ob = c * chan_stride + (tj * 32 + (lane & 31)) * MF_RS + ko;
        const int32_t *rc = runs + c * MFP_RUNS;
        // the count and the first three runs in one read
        const int4 h = *reinterpret_cast<const int4 *>(rc);
        const int n = h.x;
        for (int r = 0; r < n; r++) {
          const int32_t e = r == 0 ? h.y : r == 1 ? h.z : r == 2 ? h.w : rc[1 + r];
          const int b = e >> 16;
          if (b != cur[q]) {
            if (cur[q] >= 0 && ch0 + c < n_ch) flush(q, c, t);
            cur[q] = b;
          }
          // this lane half's eight segments of the run
          mf_mma<PREC>(acc[q], ylds, oa, ob, plane_elems, ((unsigned)e >> ko) & 0xffu);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < MFP_MAXP; q++) {
    const int p = wave + MF_WAVES * q;
    if (p < n_pairs && cur[q] >= 0) {
      int c, t, ti, tj;
      mf_pair(p, nT, RT, c, t, ti, tj);
      if (ch0 + c < n_ch) flush(q, c, t);
    }
  }
}

// ---------------------------------------------------------------------------
// k_xstep_mfma_pulsar_bin: the mixed-precision X-step with pulsar binning by
// bin. A workgroup takes one bin of one channel block (the off-gate bin 0,
// which holds most segments, in MFB_Z0 parts of the chunk) and walks only
// the 16-segment stages where that bin occurs: the stages come from the
// per-stage bin masks (k_pulsar_stage_masks), compacted a window at a time,
// and each channel's segments of the bin from the bin table
// (k_pulsar_bin_table). The Gram tiles of its (channel, tile) pairs stay in
// registers for the whole chunk and are added into gram once at the end,
// instead of at every bin change.
// ---------------------------------------------------------------------------
constexpr int MFB_Z0 = 8;     // parts of the off-gate bin
constexpr int MFB_WIN = MF_THREADS;  // stages examined per pass, one per thread

// Bins present in each stage of each channel block: bit b of
// masks[(g n_cb + cb) W + b / 32]; one thread per (stage, block).
__global__ __launch_bounds__(256) void k_pulsar_stage_masks(const int16_t *__restrict__ bins, int64_t n_segments,
                                                            int n_ch, int CH, int n_cb, int64_t n_stages, int W,
                                                            uint32_t *__restrict__ masks) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_stages * n_cb) return;
  const int64_t g = i / n_cb;
  const int cb = (int)(i - g * n_cb);
  uint32_t *m = masks + i * W;
  for (int w = 0; w < W; w++) m[w] = 0u;
  int prev = -1;
  for (int k = 0; k < MF_KS; k++) {
    const int64_t seg = g * MF_KS + k;
    if (seg >= n_segments) break;
    for (int c = 0; c < CH && cb * CH + c < n_ch; c++) {
      const int b = bins[seg * n_ch + cb * CH + c];
      if (b != prev) {
        m[b >> 5] |= 1u << (b & 31);
        prev = b;
      }
    }
  }
}

template <int PREC, bool HS>
__global__ __launch_bounds__(MF_THREADS, 4) void k_xstep_mfma_pulsar_bin(
    const void *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride, int n_ch, int M, int RT, int CH,
    int64_t seg_base, int64_t n_segments, float scale, const int16_t *__restrict__ bins,
    const uint32_t *__restrict__ masks, int W, float *__restrict__ gram, int n_cb, int G) {
  extern __shared__ _Float16 ylds[];
  constexpr int planes = PREC == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // (channel block, bin) of this workgroup: G neighbouring channel blocks of
  // one bin, whose rows share 128-byte lines and which walk nearly the same
  // stages, run on one XCD together (its L2 fetches a line once); the block
  // count is padded to whole groups, padding workgroups return at once
  const int n_cbp = (n_cb + G - 1) / G * G;
  const int L = xcd_grouped(blockIdx.x, gridDim.x, G);
  const int cb = L % n_cbp, j = L / n_cbp;
  if (cb >= n_cb) return;
  const int b = j < MFB_Z0 ? 0 : j - MFB_Z0 + 1;
  const int part = j < MFB_Z0 ? j : 0, parts = j < MFB_Z0 ? MFB_Z0 : 1;
  const int64_t n_stages = (n_segments + MF_KS - 1) / MF_KS;
  const int64_t g_lo = n_stages * part / parts, g_hi = n_stages * (part + 1) / parts;
  const int ch0 = cb * CH;
  const int chan_stride = R * MF_RS + 8;
  const int plane_elems = CH * chan_stride;
  uint32_t *cmask = reinterpret_cast<uint32_t *>(ylds + planes * plane_elems);  // [64]: the bin's segments per channel
  int32_t *list = reinterpret_cast<int32_t *>(cmask + 64);                      // [MFB_WIN]
  int32_t *wcnt = list + MFB_WIN;                                               // [MF_WAVES]
  int16_t *sbin = reinterpret_cast<int16_t *>(wcnt + MF_WAVES);                 // [channel][16]: the stage's bins
  for (int i = threadIdx.x; i < planes * plane_elems; i += MF_THREADS) ylds[i] = (_Float16)0.0f;
  f16v_t acc[MFP_MAXP];
#pragma unroll
  for (int q = 0; q < MFP_MAXP; q++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[q][e] = 0.0f;
  const int n_pairs = CH * nT;
  const int ko = 8 * (lane >> 5);
  // packed spectra whose stage items fit one per thread: each thread's item
  // (channel ic, segment group ikg, stream ist: 8 words) and one bin of the
  // stage (segment bsl of channel bcc) are loaded one listed stage ahead,
  // while the current stage is multiplied; otherwise a stage is loaded when
  // it is reached
  const int n_items = CH * 2 * M;
  const bool pf = HS && n_items <= MF_THREADS && 16 * CH <= MF_THREADS;  // uniform
  const int ic = threadIdx.x % CH, ikg = (threadIdx.x / CH) & 1, ist = threadIdx.x / CH / 2;
  const int bcc = threadIdx.x >> 4, bsl = threadIdx.x & 15;
  const uint32_t *sp32 = static_cast<const uint32_t *>(spectra);
  uint32_t pw[8];
  int pb = -1;
  auto fetch = [&](int64_t g_rel) {
    const bool it = (int)threadIdx.x < n_items && ch0 + ic < n_ch;
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      const int64_t seg = g_rel * MF_KS + ikg * 8 + kk;
      pw[kk] = it && seg < n_segments ? sp32[ist * stream_stride + seg * seg_stride + ch0 + ic] : 0u;
    }
    const int64_t sb = g_rel * MF_KS + bsl;
    pb = (int)threadIdx.x < 16 * CH && sb < n_segments && ch0 + bcc < n_ch ? bins[sb * n_ch + ch0 + bcc] : -1;
  };
  auto put = [&]() {
    if ((int)threadIdx.x < n_items) {
      typedef uint32_t u4_t __attribute__((ext_vector_type(4)));
      u4_t r, i;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        r[e] = (pw[2 * e] & 0xffffu) | (pw[2 * e + 1] << 16);
        i[e] = (pw[2 * e] >> 16) | (pw[2 * e + 1] & 0xffff0000u);
      }
      *reinterpret_cast<u4_t *>(ylds + ic * chan_stride + ist * MF_RS + ikg * 8) = r;
      *reinterpret_cast<u4_t *>(ylds + ic * chan_stride + (M + ist) * MF_RS + ikg * 8) = i;
    }
    if ((int)threadIdx.x < 16 * CH) sbin[bcc * 16 + bsl] = (int16_t)pb;
  };
  for (int64_t w0 = g_lo; w0 < g_hi; w0 += MFB_WIN) {
    // this window's stages holding the bin, in order
    const int64_t g = w0 + threadIdx.x;
    const bool has = g < g_hi && ((masks[(g * n_cb + cb) * W + (b >> 5)] >> (b & 31)) & 1u);
    const uint64_t bal = __ballot(has);
    __syncthreads();  // the previous window's list is no longer read
    if (lane == 0) wcnt[wave] = __popcll(bal);
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < MF_WAVES; w++) {
      base += w < wave ? wcnt[w] : 0;
      total += wcnt[w];
    }
    if (has) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)(g - w0);
    __syncthreads();
    if (pf && total > 0) fetch(w0 + list[0]);
    for (int e = 0; e < total; e++) {
      const int64_t g_rel = w0 + list[e];
      if (pf) {
        put();
        if (e + 1 < total) fetch(w0 + list[e + 1]);
      } else {
        mf_load_stage<PREC, MF_KS, HS>(spectra, stream_stride, seg_stride, seg_base, n_ch, M, CH, ch0,
                                       seg_base + g_rel * MF_KS, seg_base + n_segments, scale, chan_stride,
                                       plane_elems, ylds);
        if ((int)threadIdx.x < CH) {
          uint32_t mk = 0u;
          const int ch = ch0 + threadIdx.x;
          if (ch < n_ch) {
#pragma unroll
            for (int k = 0; k < MF_KS; k++) {
              const int64_t seg = g_rel * MF_KS + k;
              if (seg < n_segments && bins[seg * n_ch + ch] == b) mk |= 1u << k;
            }
          }
          cmask[threadIdx.x] = mk;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < MFP_MAXP; q++) {
        const int p = wave + MF_WAVES * q;
        if (p < n_pairs) {
          int c, t, ti, tj;
          mf_pair(p, nT, RT, c, t, ti, tj);
          uint32_t mk = 0u;
          if (pf) {  // the channel's 16 bins of the stage (wave-uniform reads)
            typedef int16_t s8_t __attribute__((ext_vector_type(8)));
            const s8_t lo8 = *reinterpret_cast<const s8_t *>(sbin + c * 16);
            const s8_t hi8 = *reinterpret_cast<const s8_t *>(sbin + c * 16 + 8);
#pragma unroll
            for (int k = 0; k < 8; k++) {
              mk |= (lo8[k] == b ? 1u : 0u) << k;
              mk |= (hi8[k] == b ? 1u : 0u) << (k + 8);
            }
          } else {
            mk = cmask[c];
          }
          if (mk != 0u) {
            const int oa = c * chan_stride + (ti * 32 + (lane & 31)) * MF_RS + ko;
            const int ob = c * chan_stride + (tj * 32 + (lane & 31)) * MF_RS + ko;
            mf_mma<PREC>(acc[q], ylds, oa, ob, plane_elems, (mk >> ko) & 0xffu);
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < MFP_MAXP; q++) {
    const int p = wave + MF_WAVES * q;
    if (p < n_pairs) {
      int c, t, ti, tj;
      mf_pair(p, nT, RT, c, t, ti, tj);
      if (ch0 + c < n_ch) {
        float *o = gram + (((int64_t)b * n_ch + ch0 + c) * nT + t) * 1024 + (lane & 31);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int jj = 0; jj < 4; jj++) unsafeAtomicAdd(o + (8 * i + 4 * (lane >> 5) + jj) * 32, acc[q][4 * i + jj]);
      }
    }
  }
}

// Visibilities of every bin from the binned Gram tiles:
// vis[bin][b][ch] = (G[a][d] + G[M+a][M+d] + i (G[M+a][d] - G[a][M+d])) / scale^2.
// gram holds channels ch_base .. ch_base + n_run - 1; vis rows have n_ch channels.
__global__ __launch_bounds__(256) void k_pulsar_gram_baselines_run(const float *__restrict__ gram, int n_run,
                                                                   int ch_base, int n_ch, int M, int RT,
                                                                   const int32_t *__restrict__ bl_first,
                                                                   const int32_t *__restrict__ bl_second,
                                                                   int n_baselines, float inv_scale2,
                                                                   float2 *__restrict__ vis) {
  const int ch = blockIdx.x, bin = blockIdx.y;
  const int nT = RT * (RT + 1) / 2;
  const float *g = gram + ((int64_t)bin * n_run + ch) * nT * 1024;
  auto G = [&](int r, int c) {
    if (r > c) {
      const int tmp = r;
      r = c;
      c = tmp;
    }
    const int ti = r >> 5, tj = c >> 5;
    const int t = ti * RT - ti * (ti - 1) / 2 + (tj - ti);
    return g[t * 1024 + (r & 31) * 32 + (c & 31)];
  };
  for (int b = threadIdx.x; b < n_baselines; b += blockDim.x) {
    const int a = bl_first[b], d = bl_second[b];
    const float re = (G(a, d) + G(M + a, M + d)) * inv_scale2;
    const float im = (G(M + a, d) - G(a, M + d)) * inv_scale2;
    vis[((int64_t)bin * n_baselines + b) * n_ch + ch_base + ch] = make_float2(re, im);
  }
}

__global__ __launch_bounds__(256) void k_fill_zero(float4 *__restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(256) void k_fill_zero_u32(uint32_t *__restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// normalisation (correlation_core.cc:275-314)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_auto_norms(const float2 *__restrict__ vis, int n_ch, int fft_size,
                                                    int n_baselines, double *__restrict__ norms, int n_streams) {
  __shared__ double part[256];
  const int i = blockIdx.x, bin = blockIdx.y;
  const float2 *row = vis + ((int64_t)bin * n_baselines + i) * n_ch;
  double sum = 0.0;
  for (int j = threadIdx.x; j < n_ch; j += blockDim.x) sum += (double)row[j].x;
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double n = part[0] / fft_size;
    if (n < 1) n = 1;
    norms[bin * n_streams + i] = n;
  }
}

__global__ __launch_bounds__(256) void k_normalise(const float2 *__restrict__ vis, const double *__restrict__ norms,
                                                   const int32_t *__restrict__ bl_first,
                                                   const int32_t *__restrict__ bl_second,
                                                   const double *__restrict__ flag_factor, int n_streams,
                                                   int n_baselines, int n_ch, float2 *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int bin = blockIdx.y;
  if (idx >= (int64_t)n_baselines * n_ch) return;
  const int b = (int)(idx / n_ch);
  const int64_t k = (int64_t)bin * n_baselines * n_ch + idx;
  const float2 v = vis[k];
  const double *nm = norms + (int64_t)bin * n_streams;
  if (b < n_streams) {
    out[k] = make_float2((float)((double)v.x / nm[b]), 0.0f);
  } else {
    const float norm = (float)sqrt(flag_factor[b] * nm[bl_first[b]] * nm[bl_second[b]]);
    out[k] = make_float2(v.x / norm, v.y / norm);
  }
}

// ---------------------------------------------------------------------------
// synthetic raw data: counter-based Gaussian samples (SplitMix64 + Box-Muller)
// ---------------------------------------------------------------------------
namespace {
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ float gauss(uint64_t seed, int64_t k) {
  const uint64_t h = splitmix64(seed ^ splitmix64((uint64_t)k));
  const float u1 = ((float)(uint32_t)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = (float)(uint32_t)(h & 0xFFFFFFu) * (1.0f / 16777216.0f);
  return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}
}  // namespace

__global__ __launch_bounds__(256) void k_synthesize(uint8_t *__restrict__ dst, int64_t n_bytes, int bits,
                                                    int64_t first_sample, int64_t delay_samples,
                                                    uint64_t common_seed, uint64_t stream_seed, float rho,
                                                    float threshold, int64_t delay2_samples, uint64_t common2_seed,
                                                    float rho2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_bytes) return;
  const int spb = 8 / bits;
  const float a = sqrtf(rho), a2 = sqrtf(rho2), b = sqrtf(1.0f - rho - rho2);
  uint32_t byte = 0;
  for (int t = 0; t < spb; t++) {
    const int64_t k = first_sample + i * spb + t;
    float x = a * gauss(common_seed, k - delay_samples) + b * gauss(stream_seed, k);
    if (rho2 > 0.0f) x += a2 * gauss(common2_seed, k - delay2_samples);
    uint32_t c;
    if (bits == 2) c = x < -threshold ? 0u : (x < 0.0f ? 1u : (x < threshold ? 2u : 3u));
    else c = x >= 0.0f ? 1u : 0u;
    byte |= c << (bits * t);
  }
  dst[i] = (uint8_t)byte;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_unpack(const DevStream *d_streams, int n_streams, int64_t n_samples, hipStream_t s) {
  dim3 grid((unsigned)((n_samples + 255) / 256), n_streams);
  hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, s, d_streams, n_samples);
  return hipGetLastError();
}

hipError_t launch_stats(const DevStream *d_streams, int n_streams, int64_t max_counted_bytes, hipStream_t s) {
  if (max_counted_bytes <= 0 || n_streams <= 0) return hipSuccess;
  // about 4096 workgroups over all streams, each over at least 16 KiB (a
  // multiple of the 4096 ordinals a workgroup takes per step)
  int64_t span = max_counted_bytes * n_streams / 4096;
  span = std::max<int64_t>(16384, (span + 4095) / 4096 * 4096);
  dim3 grid((unsigned)((max_counted_bytes + span - 1) / span), n_streams);
  hipLaunchKernelGGL(k_stats, grid, dim3(256), 0, s, d_streams, span);
  return hipGetLastError();
}

hipError_t launch_block_consts(const DevStream *d_streams, const DevJob &job, hipStream_t s) {
  if (job.nd != 256) return hipErrorInvalidValue;
  if (job.n_blocks <= 0) return hipSuccess;
  dim3 gridc((unsigned)((job.n_blocks + 255) / 256), job.n_streams);
  hipLaunchKernelGGL(k_block_consts, gridc, dim3(256), 0, s, d_streams, job.n_blocks, job.nd);
  return hipGetLastError();
}

hipError_t launch_delay(const DevStream *d_streams, const DevJob &job, const float2 *tw_nd, hipStream_t s) {
  if (job.nd != 256) return hipErrorInvalidValue;
  const int64_t end = job.blk_end < job.n_blocks ? job.blk_end : job.n_blocks;
  if (end <= job.blk_base) return hipSuccess;
  // about 8192 workgroups over all streams; each wave then loops over four
  // pairs at a time (32 blocks per workgroup step)
  const int64_t wgs = (end - job.blk_base + 31) / 32;
  const int64_t per_stream = std::max<int64_t>(1, (8192 + job.n_streams - 1) / job.n_streams);
  dim3 grid((unsigned)std::min(wgs, per_stream), job.n_streams);
  if (job.ts_half) hipLaunchKernelGGL(k_delay<true>, grid, dim3(256), 0, s, d_streams, job, tw_nd);
  else hipLaunchKernelGGL(k_delay<false>, grid, dim3(256), 0, s, d_streams, job, tw_nd);
  return hipGetLastError();
}

template <int N, int P, int SP>
static hipError_t launch_chan_small(const DevStream *d_streams, const DevJob &job, const float *window,
                                    const float2 *tw_split, hipStream_t s) {
  dim3 grid((unsigned)job.n_segments, job.n_streams);
  hipLaunchKernelGGL((k_channelise<N, P, SP>), grid, dim3(N / P), 0, s, d_streams, job, window, tw_split);
  return hipGetLastError();
}

template <int N, int SP>
static hipError_t launch_chan_n(const DevStream *d_streams, const DevJob &job, const float *window,
                                const float2 *tw_n, const float2 *tw_2n, const float2 *tw_split, const Tune &t,
                                hipStream_t s) {
  if constexpr (N >= 8192) {
    // persistent runs of consecutive segments, about chan_wgs (2048) workgroups in all
    constexpr int P = 16;
    if (job.seg_step != N) return hipErrorInvalidValue;
    const int64_t wgs = t.chan_wgs > 0 ? t.chan_wgs : 2048;
    const int64_t per_stream = (wgs + job.n_streams - 1) / job.n_streams;
    int64_t segs = (job.n_segments + per_stream - 1) / per_stream;
    if (segs < 1) segs = 1;
    const unsigned gx = (unsigned)((job.n_segments + segs - 1) / segs);
    hipLaunchKernelGGL((k_channelise_run<N, P, SP>), dim3(gx, job.n_streams), dim3(N / P), 0, s, d_streams, job, window,
                       tw_split, (int)segs);
  } else {
    // the fused F-engine's schedule, so both paths give the same bits:
    // threads min(N / 4, 512) (PA points each) or half as many
    constexpr int PA = N / (N / 4 < 512 ? N / 4 : 512);
    const int T = t.fe_threads > 0 ? t.fe_threads : N / PA;
    if (T == N / PA) return launch_chan_small<N, PA, SP>(d_streams, job, window, tw_split, s);
    if constexpr (2 * PA <= 16 && N / (2 * PA) >= 64) {
      if (T == N / (2 * PA)) return launch_chan_small<N, 2 * PA, SP>(d_streams, job, window, tw_split, s);
    }
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// compute units of the current device (cached per device)
static int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n <= 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

// the MFMA channeliser (N_c = 1024, 2048, 4096, 8192, 16384, packed FP16 / BF16
// spectra): each workgroup a run of consecutive segments of one stream, the
// same number of workgroups per stream, at most chan_wgs in all. One
// workgroup fits a CU (its LDS), so by default chan_wgs is 2, 4 or 1 times
// the CU count, whichever leaves the fewest CUs idle in the last round.
template <int SP, int NT>
static hipError_t launch_chan_mfma_nt(const DevStream *d_streams, const DevJob &job, const Tune &t, hipStream_t s) {
  constexpr int SG = 16 / NT;  // segments per pass
  if (job.seg_step != 1024 * NT) return hipErrorInvalidValue;
  int64_t wgs = t.chan_wgs;
  if (wgs <= 0) {
    const int cus = device_cus();
    double best = -1.0;
    for (int k : {2, 4, 1}) {
      const int64_t g = std::max<int64_t>(1, (int64_t)k * cus / job.n_streams) * job.n_streams;
      const double eff = (double)g / (double)((g + cus - 1) / cus * cus);
      if (eff > best + 1e-9) {
        best = eff;
        wgs = (int64_t)k * cus;
      }
    }
  }
  const int64_t per_stream = std::max<int64_t>(1, wgs / job.n_streams);
  int64_t segs = (job.n_segments + per_stream - 1) / per_stream;
  segs = (segs + SG - 1) / SG * SG;  // whole passes
  if (segs < SG) segs = SG;
  const unsigned gx = (unsigned)((job.n_segments + segs - 1) / segs);
  if (job.ts_half)
    hipLaunchKernelGGL((k_chan_mfma<SP, NT, true>), dim3(gx, job.n_streams), dim3(CM_T), 0, s, d_streams, job, (int)segs);
  else
    hipLaunchKernelGGL((k_chan_mfma<SP, NT, false>), dim3(gx, job.n_streams), dim3(CM_T), 0, s, d_streams, job, (int)segs);
  return hipGetLastError();
}
template <int SP>
static hipError_t launch_chan_mfma(const DevStream *d_streams, const DevJob &job, const Tune &t, hipStream_t s) {
  switch (job.nc) {
    case 1024: return launch_chan_mfma_nt<SP, 1>(d_streams, job, t, s);
    case 2048: return launch_chan_mfma_nt<SP, 2>(d_streams, job, t, s);
    case 4096: return launch_chan_mfma_nt<SP, 4>(d_streams, job, t, s);
    case 8192: return launch_chan_mfma_nt<SP, 8>(d_streams, job, t, s);
    case 16384: return launch_chan_mfma_nt<SP, 16>(d_streams, job, t, s);
    default: return hipErrorInvalidValue;
  }
}

bool chan_mfma_supported(int nc, int spec_prec) {
  return (nc == 1024 || nc == 2048 || nc == 4096 || nc == 8192 || nc == 16384) &&
         (spec_prec == SFXG_PREC_FP16 || spec_prec == SFXG_PREC_BF16);
}

size_t channelise_scratch_elems(int nc) { return nc == LG_N ? (size_t)2 * 1024 * LG_N : 0; }

template <int SP>
static hipError_t launch_chan_large(const DevStream *d_streams, const DevJob &job, const float *window,
                                    const float2 *tw_n, const float2 *tw_2n, float2 *scratch, hipStream_t s) {
  // scratch: Y and Z for up to 1024 (stream, segment) items per launch
  constexpr int64_t kItems = 1024;
  if (job.seg_step != LG_N) return hipErrorInvalidValue;
  float2 *Y = scratch, *Z = scratch + kItems * LG_N;
  const int64_t total = (int64_t)job.n_streams * job.n_segments;
  for (int64_t i0 = 0; i0 < total; i0 += kItems) {
    const unsigned n = (unsigned)(total - i0 < kItems ? total - i0 : kItems);
    hipLaunchKernelGGL(k_chan_cols, dim3(LG_N2 / 16, n), dim3(256), 0, s, d_streams, job, window, tw_n, i0, Y);
    hipLaunchKernelGGL(k_chan_rows, dim3(LG_N1 / 16, n), dim3(256), 0, s, Y, tw_n, Z);
    hipLaunchKernelGGL(k_chan_split<SP>, dim3(n), dim3(256), 0, s, d_streams, job, Z, tw_2n, i0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <int SP>
static hipError_t launch_channelise_sp(const DevStream *d_streams, const DevJob &job, const float *window,
                             const float2 *tw_n, const float2 *tw_2n, const float2 *tw_split, float2 *scratch,
                             const Tune &t, hipStream_t s) {
  if (job.n_segments <= 0) return hipSuccess;
  if constexpr (SP == SFXG_PREC_FP16 || SP == SFXG_PREC_BF16) {
    if (chan_mfma_supported(job.nc, SP) && t.chan_fp32 != 1) return launch_chan_mfma<SP>(d_streams, job, t, s);
  }
  if (job.ts_half) return hipErrorInvalidValue;  // an FP16 series only feeds the MFMA channeliser
  if (job.nc == LG_N) return launch_chan_large<SP>(d_streams, job, window, tw_n, tw_2n, scratch, s);
  switch (job.nc) {
    case 256: return launch_chan_n<256, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    case 512: return launch_chan_n<512, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    case 1024: return launch_chan_n<1024, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    case 2048: return launch_chan_n<2048, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    case 4096: return launch_chan_n<4096, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    case 8192: return launch_chan_n<8192, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    case 16384: return launch_chan_n<16384, SP>(d_streams, job, window, tw_n, tw_2n, tw_split, t, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_channelise(const DevStream *d_streams, const DevJob &job, const float *window,
                             const float2 *tw_n, const float2 *tw_2n, const float2 *tw_split, float2 *scratch,
                             const Tune &t, hipStream_t s) {
  switch (job.spec_prec) {
    case 0: return launch_channelise_sp<0>(d_streams, job, window, tw_n, tw_2n, tw_split, scratch, t, s);
    case SFXG_PREC_FP16:
      return launch_channelise_sp<SFXG_PREC_FP16>(d_streams, job, window, tw_n, tw_2n, tw_split, scratch, t, s);
    case SFXG_PREC_BF16:
      return launch_channelise_sp<SFXG_PREC_BF16>(d_streams, job, window, tw_n, tw_2n, tw_split, scratch, t, s);
    case SFXG_PREC_SPLIT_FP16:
      return launch_channelise_sp<SFXG_PREC_SPLIT_FP16>(d_streams, job, window, tw_n, tw_2n, tw_split, scratch, t, s);
    default: return hipErrorInvalidValue;
  }
}

std::vector<int> fe_thread_candidates(int nc) {
  switch (nc) {
    case 256: return {64};
    case 512: return {128, 64};
    case 1024: return {256, 128};
    case 2048: return {512, 256};
    case 4096: return {512, 256};
    default: return {0};
  }
}
std::vector<int> chan_mfma_wgs_candidates() { return {512, 256, 1024}; }
std::vector<int> chan_wgs_candidates(int nc) {
  if (nc == 8192 || nc == 16384) return {2048, 1024, 4096};
  return {0};
}
// 0: register tiles (16-channel blocks for many tiles, else 64-channel blocks),
// 1: register tiles on 64-channel blocks, 2: on 16-channel blocks, else the
// per-baseline kernel's waves per workgroup
std::vector<int> xs_waves_candidates() { return {0, 1, 2, 8, 4}; }
std::vector<int> mf_channel_candidates(int precision, int M) {
  const int RT = (2 * M + 31) / 32;
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int planes = precision == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
  int CH = (MF_WAVES * MF_MAXP) / nT;
  if (CH > 64) CH = 64;
  while (CH > 1 && (size_t)planes * CH * (R * MF_RS + 8) * 2 > 150 * 1024) CH--;
  std::vector<int> c{CH};
  if (CH / 2 >= 1) c.push_back(CH / 2);
  if (CH / 4 >= 1) c.push_back(CH / 4);
  return c;
}
// 0: by bin (k_xstep_mfma_pulsar_bin), else the time-sliced kernel's workgroups
std::vector<int> mfp_wgs_candidates() { return {0, 1024, 512}; }

template <int N, int T, int SP>
static hipError_t launch_fe_n(const DevStream *d_streams, const DevJob &job, const float *window,
                              const float2 *tw_nd, const float2 *tw_split, hipStream_t s) {
  // about 1024 workgroups in flight: two per CU
  int per_stream = (1024 + job.n_streams - 1) / job.n_streams;
  int64_t segs = (job.n_segments + per_stream - 1) / per_stream;
  if (segs < 1) segs = 1;
  const unsigned gx = (unsigned)((job.n_segments + segs - 1) / segs);
  hipLaunchKernelGGL((k_fengine<N, T, SP>), dim3(gx, job.n_streams), dim3(T), 0, s, d_streams, job, window, tw_nd,
                     tw_split, (int)segs);
  return hipGetLastError();
}

bool fengine_supported(int nc) { return nc >= 256 && nc <= 4096; }

template <int SP>
static hipError_t launch_fengine_sp(const DevStream *d_streams, const DevJob &job, const float *window, const float2 *tw_nd,
                          const float2 *tw_split, const Tune &t, hipStream_t s) {
  if (job.n_segments <= 0) return hipSuccess;
  if (job.nd != 256) return hipErrorInvalidValue;
  const int T = t.fe_threads > 0 ? t.fe_threads : fe_thread_candidates(job.nc)[0];
  switch (job.nc * 1024 + T) {
    case 256 * 1024 + 64: return launch_fe_n<256, 64, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 512 * 1024 + 128: return launch_fe_n<512, 128, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 512 * 1024 + 64: return launch_fe_n<512, 64, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 1024 * 1024 + 256: return launch_fe_n<1024, 256, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 1024 * 1024 + 128: return launch_fe_n<1024, 128, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 2048 * 1024 + 512: return launch_fe_n<2048, 512, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 2048 * 1024 + 256: return launch_fe_n<2048, 256, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 4096 * 1024 + 512: return launch_fe_n<4096, 512, SP>(d_streams, job, window, tw_nd, tw_split, s);
    case 4096 * 1024 + 256: return launch_fe_n<4096, 256, SP>(d_streams, job, window, tw_nd, tw_split, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fengine(const DevStream *d_streams, const DevJob &job, const float *window, const float2 *tw_nd,
                          const float2 *tw_split, const Tune &t, hipStream_t s) {
  switch (job.spec_prec) {
    case 0: return launch_fengine_sp<0>(d_streams, job, window, tw_nd, tw_split, t, s);
    case SFXG_PREC_FP16: return launch_fengine_sp<SFXG_PREC_FP16>(d_streams, job, window, tw_nd, tw_split, t, s);
    case SFXG_PREC_BF16: return launch_fengine_sp<SFXG_PREC_BF16>(d_streams, job, window, tw_nd, tw_split, t, s);
    case SFXG_PREC_SPLIT_FP16:
      return launch_fengine_sp<SFXG_PREC_SPLIT_FP16>(d_streams, job, window, tw_nd, tw_split, t, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_xstep(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base,
                        int n_ch, int M, const int32_t *fold_start, int n_k, const int32_t *bl_first,
                        const int32_t *bl_second, int n_baselines, float2 *partial, float2 *vis, const Tune &t,
                        hipStream_t s) {
  if (n_k <= 0) return hipSuccess;
  const int max_w = t.xs_waves > 0 ? t.xs_waves : 8;
  int nw = (n_baselines + XBPW - 1) / XBPW;
  if (nw > max_w) nw = max_w;
  const int bl_tiles = (n_baselines + nw * XBPW - 1) / (nw * XBPW);
  // a stage must fit the threads' prefetch registers (16 float2 each)
  int sps = (16 * 64 * nw) / (M * 64);
  if (sps > 16) sps = 16;
  if (sps < 1) return hipErrorInvalidValue;  // more streams than one stage holds
  const size_t lds = (size_t)sps * M * 64 * sizeof(float2);
  if (lds > 65536) return hipErrorInvalidValue;
  dim3 grid((unsigned)((n_ch + 63) / 64), (unsigned)bl_tiles, (unsigned)n_k);
  hipLaunchKernelGGL(k_xstep_subint, grid, dim3(64 * nw), lds, s, spectra, stream_stride, seg_stride, seg_base, n_ch,
                     M, fold_start, bl_first, bl_second, n_baselines, sps, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (!vis) return hipSuccess;  // the caller folds the partials itself
  const int64_t n = (int64_t)n_baselines * n_ch;
  hipLaunchKernelGGL(k_fold_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, n_k, n, vis);
  return hipGetLastError();
}

void xstep_tile_table(int M, const std::vector<int32_t> &bl_first, const std::vector<int32_t> &bl_second,
                      std::vector<int32_t> &table) {
  const int nt = (M + XT - 1) / XT;
  std::vector<int32_t> idx((size_t)M * M, XT_NONE);  // (lo, hi) -> baseline (or ~b if listed as (hi, lo))
  for (size_t b = 0; b < bl_first.size(); b++) {
    const int f = bl_first[b], g = bl_second[b];
    if (f <= g) idx[(size_t)f * M + g] = (int32_t)b;
    else idx[(size_t)g * M + f] = ~(int32_t)b;
  }
  table.clear();
  for (int ta = 0; ta < nt; ta++)
    for (int tb = ta; tb < nt; tb++) {
      std::vector<int32_t> e(2 + XT * XT, XT_NONE);
      e[0] = ta * XT;
      e[1] = tb * XT;
      bool any = false;
      for (int i = 0; i < XT; i++)
        for (int j = 0; j < XT; j++) {
          const int a = ta * XT + i, c = tb * XT + j;
          if (a >= M || c >= M || a > c) continue;
          e[2 + XT * i + j] = idx[(size_t)a * M + c];
          any = any || e[2 + XT * i + j] != XT_NONE;
        }
      if (any) table.insert(table.end(), e.begin(), e.end());
    }
}

bool xstep_tile16_default(int n_tiles) { return n_tiles > 64; }


hipError_t launch_xstep_tile(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base,
                             int n_ch, int M, const int32_t *fold_start, int n_k, const int32_t *tiles, int n_tiles,
                             const int32_t *bl_first, const int32_t *bl_second, int n_baselines, float2 *partial,
                             float2 *vis, bool tile16, hipStream_t s) {
  if (n_k <= 0) return hipSuccess;
  if (tile16) {
    // a stage of sps segments fills the threads' prefetch registers
    int sps = (XQ_PF * 64 * XQ_WAVES) / (M * 8);
    if (sps > 8) sps = 8;
    if (sps < 1) return hipErrorInvalidValue;
    if (stream_stride % 2 || seg_stride % 64) return hipErrorInvalidValue;  // 16-byte loads inside the rows
    const size_t lds = (size_t)sps * M * 8 * sizeof(float4);
    const int TPS = n_tiles <= XQ_SLOTS ? 1 : n_tiles <= 2 * XQ_SLOTS ? 2 : 3;
    const int n_run = n_ch > 1 && (n_ch - 1) % 16 == 0 ? n_ch - 1 : n_ch;
    const int n_cb = (n_run + 15) / 16, n_tg = (n_tiles + TPS * XQ_SLOTS - 1) / (TPS * XQ_SLOTS);
    if ((int64_t)n_cb * n_tg * n_k >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    dim3 grid((unsigned)(n_cb * n_tg * n_k));
    hipError_t e = hipErrorInvalidValue;
#define SFXG_XQ(T)                                                                                                \
  do {                                                                                                            \
    e = hipFuncSetAttribute((const void *)k_xstep_tile16<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                                                \
    hipLaunchKernelGGL(k_xstep_tile16<T>, grid, dim3(64 * XQ_WAVES), lds, s, spectra, stream_stride, seg_stride,  \
                       seg_base, n_ch, n_run, M, fold_start, tiles, n_tiles, n_baselines, sps, n_cb, n_tg, partial); \
    e = hipGetLastError();                                                                                        \
  } while (0)
    if (TPS == 1) SFXG_XQ(1);
    else if (TPS == 2) SFXG_XQ(2);
    else SFXG_XQ(3);
#undef SFXG_XQ
    if (e != hipSuccess) return e;
    if (n_run < n_ch) {
      e = launch_xstep_last(spectra, stream_stride, seg_stride, seg_base, n_ch - 1, M, fold_start, n_k, bl_first,
                            bl_second, n_baselines, partial + (n_ch - 1), (int64_t)n_baselines * n_ch, n_ch, s);
      if (e != hipSuccess) return e;
    }
    if (!vis) return hipSuccess;
    const int64_t n = (int64_t)n_baselines * n_ch;
    hipLaunchKernelGGL(k_fold_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, n_k, n, vis);
    return hipGetLastError();
  }
  // a stage fills the threads' prefetch registers
  int sps = (XT_PF * 64 * XT_WAVES) / (M * 64);
  if (sps > 16) sps = 16;
  if (sps < 1) return hipErrorInvalidValue;  // more streams than one stage holds
  const size_t lds = (size_t)sps * M * 64 * sizeof(float2);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void *)k_xstep_tile, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // the last bin of N_c + 1 on its own, so that the tile grid covers whole
  // blocks of 64 channels
  const int n_run = n_ch > 1 && (n_ch - 1) % 64 == 0 ? n_ch - 1 : n_ch;
  const int n_cb = (n_run + 63) / 64, n_tg = (n_tiles + XT_WAVES - 1) / XT_WAVES;
  if ((int64_t)n_cb * n_tg * n_k >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  dim3 grid((unsigned)(n_cb * n_tg * n_k));
  hipLaunchKernelGGL(k_xstep_tile, grid, dim3(64 * XT_WAVES), lds, s, spectra, stream_stride, seg_stride, seg_base,
                     n_ch, n_run, M, fold_start, tiles, n_tiles, n_baselines, sps, n_cb, n_tg, partial);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n_run < n_ch) {
    e = launch_xstep_last(spectra, stream_stride, seg_stride, seg_base, n_ch - 1, M, fold_start, n_k, bl_first,
                          bl_second, n_baselines, partial + (n_ch - 1), (int64_t)n_baselines * n_ch, n_ch, s);
    if (e != hipSuccess) return e;
  }
  if (!vis) return hipSuccess;  // the caller folds the partials itself
  const int64_t n = (int64_t)n_baselines * n_ch;
  hipLaunchKernelGGL(k_fold_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, n_k, n, vis);
  return hipGetLastError();
}

template <int PREC, int KS, bool HS>
static hipError_t launch_mfma_ks(dim3 grid, size_t lds, hipStream_t s, const void *spectra, int64_t stream_stride,
                                 int64_t seg_stride, int64_t seg_base, int n_ch, int M, int RT, int CH, int n_blk,
                                 int G, const int32_t *fold_start, float scale, const int32_t *bl_first,
                                 const int32_t *bl_second, int n_baselines, float inv2, float2 *partial) {
  hipError_t e = hipFuncSetAttribute((const void *)k_xstep_mfma<PREC, KS, HS>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_xstep_mfma<PREC, KS, HS>), grid, dim3(MF_THREADS), lds, s, spectra, stream_stride, seg_stride,
                     seg_base, n_ch, M, RT, CH, n_blk, G, fold_start, scale, bl_first, bl_second, n_baselines, inv2,
                     partial);
  return hipGetLastError();
}

bool xstep_split_planes_fit(int M, int64_t stream_stride) {
  const int RT = (2 * M + 31) / 32, R = 32 * RT, nT = RT * (RT + 1) / 2;
  if (nT > MF_WAVES * MF_MAXP || (MF_WAVES * MF_MAXP) / nT < 4) return false;
  if (stream_stride * 2 * M >= ((int64_t)1 << 32) || stream_stride * 2 * 8 >= ((int64_t)1 << 32)) return false;
  const int n = (mfd_blocks(M, 8) + MF_WAVES - 1) / MF_WAVES;  // CW = 8: four channels
  const int npw = n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : n <= 8 ? 8 : 0;
  return npw > 0 && std::max(mfd_lds_bytes(npw, 8, R), (size_t)R * R * 4) <= 160 * 1024;
}

hipError_t launch_xstep_mfma(int precision, const void *spectra, bool half_spectra, int64_t stream_stride, int64_t seg_stride,
                             int64_t seg_base, int n_ch, int M, const int32_t *fold_start, int n_k,
                             const int32_t *bl_first, const int32_t *bl_second, int n_baselines, float scale,
                             float2 *partial, float2 *vis, const Tune &t, hipStream_t s) {
  if (n_k <= 0) return hipSuccess;
  const int RT = (2 * M + 31) / 32;
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int planes = precision == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
  if (nT > MF_WAVES * MF_MAXP) return hipErrorInvalidValue;
  // split FP16 on hi / lo planes (two words per bin) has only the
  // direct-to-LDS form
  const bool split_planes = half_spectra && precision == SFXG_PREC_SPLIT_FP16;
  const int W = split_planes ? 2 : 1;
  if (half_spectra && (t.mf_ks != 32 || split_planes) && stream_stride * W * M < ((int64_t)1 << 32)) {
    // direct-to-LDS form: CW = W CH words per row, a power of two from 8 to
    // 64, whose (channel, tile) pairs fit the waves, NPW = 1, 2, 4 or 8 block
    // loads per wave and stage, and the LDS budget
    const int max_pairs_ch = (MF_WAVES * MF_MAXP) / nT;
    auto npw_of = [&](int cw) {
      const int n = (mfd_blocks(M, cw) + MF_WAVES - 1) / MF_WAVES;
      return n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : n <= 8 ? 8 : 0;
    };
    auto lds_of = [&](int cw) { return std::max(mfd_lds_bytes(npw_of(cw), cw, R), (size_t)R * R * 4); };
    int CW = 64;
    const int cap = t.mf_channels > 0 ? std::max(t.mf_channels, 8 / W) : 64;
    while (CW >= 8 && (CW / W > max_pairs_ch || npw_of(CW) == 0 || lds_of(CW) > 160 * 1024 || CW / W > cap))
      CW /= 2;
    if (CW >= 8 && stream_stride * W * 8 < ((int64_t)1 << 32)) {
      const int npw = npw_of(CW);
      const size_t lds = lds_of(CW);
      const int CH = CW / W;
      // G channel blocks share a 128-byte line of a spectra row; n_blk is
      // padded to whole groups, so that no group straddles two
      // sub-integrations (and two XCDs fetch no line twice); padding blocks
      // return at once
      const int G = std::max(1, 128 / (CW * 4));
      const int n_blk = ((n_ch + CH - 1) / CH + G - 1) / G * G;
      dim3 grid((unsigned)(n_blk * n_k));
      const float inv2 = 1.0f / (scale * scale);
      const uint32_t *sp = static_cast<const uint32_t *>(spectra);
      hipError_t e = hipErrorInvalidValue;
#define SFXG_MFD3(P, NW, C)                                                                                        \
  do {                                                                                                             \
    e = hipFuncSetAttribute((const void *)k_xstep_mfma_dma<P, NW, C>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            (int)lds);                                                                             \
    if (e != hipSuccess) return e;                                                                                 \
    hipLaunchKernelGGL((k_xstep_mfma_dma<P, NW, C>), grid, dim3(MF_THREADS), lds, s, sp, stream_stride,           \
                       seg_stride, seg_base, n_ch, M, RT, n_blk, G, fold_start, bl_first, bl_second, n_baselines,  \
                       inv2, partial);                                                                             \
    e = hipGetLastError();                                                                                         \
  } while (0)
#define SFXG_MFD(P, NW)                     \
  do {                                      \
    if (CW == 8) SFXG_MFD3(P, NW, 8);       \
    else if (CW == 16) SFXG_MFD3(P, NW, 16); \
    else if (CW == 32) SFXG_MFD3(P, NW, 32); \
    else SFXG_MFD3(P, NW, 64);              \
  } while (0)
#define SFXG_MFDP(P)                          \
  do {                                        \
    if (npw == 1) SFXG_MFD(P, 1);             \
    else if (npw == 2) SFXG_MFD(P, 2);        \
    else if (npw == 4) SFXG_MFD(P, 4);        \
    else SFXG_MFD(P, 8);                      \
  } while (0)
      if (precision == SFXG_PREC_FP16) SFXG_MFDP(SFXG_PREC_FP16);
      else if (precision == SFXG_PREC_BF16) SFXG_MFDP(SFXG_PREC_BF16);
      else if (precision == SFXG_PREC_SPLIT_FP16) SFXG_MFDP(SFXG_PREC_SPLIT_FP16);
#undef SFXG_MFDP
#undef SFXG_MFD
#undef SFXG_MFD3
      if (e != hipSuccess) return e;
      dim3 gc((unsigned)((n_baselines + 63) / 64), (unsigned)((n_ch + 63) / 64));
      hipLaunchKernelGGL(k_fold_combine_cm, gc, dim3(256), 0, s, partial, n_k, n_ch, n_baselines, vis);
      return hipGetLastError();
    }
  }
  if (split_planes) return hipErrorInvalidValue;  // the caller keeps FP32 spectra where this form does not fit
  // channels per workgroup: all (channel, tile) pairs fit MF_WAVES waves x
  // MF_MAXP and the stage fits 150 KiB of LDS. Stages of 32 segments (two
  // MFMA K steps per barrier) where that keeps the channel count, else 16.
  const int ch_pairs = std::min(64, (MF_WAVES * MF_MAXP) / nT);
  auto lds_of = [&](int ch, int ks) {
    const size_t stage = (size_t)planes * ch * (R * (ks + 8) + 8) * 2;
    const size_t gbuf = (size_t)R * R * 4;
    return stage > gbuf ? stage : gbuf;
  };
  auto fit = [&](int ks) {
    int ch = ch_pairs;
    while (ch > 1 && lds_of(ch, ks) > 150 * 1024) ch--;
    return ch;
  };
  // and a stage's loads fit the prefetch registers
  const int words = half_spectra ? 8 : 16;
  auto fit_pf = [&](int ks) {
    int ch = fit(ks);
    while (ch > 1 && ch * (ks / 8) * M * words > MF_THREADS * MF_PF_WORDS) ch--;
    return ch;
  };
  int KS = t.mf_ks > 0 ? t.mf_ks : (fit_pf(32) >= fit_pf(16) ? 32 : 16);
  int CH = fit_pf(KS);
  if (t.mf_channels > 0 && CH > t.mf_channels) CH = t.mf_channels;
  if (CH < 1 || CH * (KS / 8) * M * words > MF_THREADS * MF_PF_WORDS) return hipErrorInvalidValue;
  if (stream_stride * M >= ((int64_t)1 << 32)) return hipErrorInvalidValue;  // 32-bit item offsets
  const size_t lds = lds_of(CH, KS);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  // FP32 spectra: the last bin of N_c + 1 on its own (in FP32), so that the
  // grid covers whole blocks of CH channels
  const int n_run = !half_spectra && n_ch > 1 && (n_ch - 1) % CH == 0 ? n_ch - 1 : n_ch;
  // channel blocks that share a 128-byte line of a spectra row; with packed
  // spectra (n_run = n_ch) n_blk is padded to whole groups, so that no group
  // straddles two sub-integrations, and padding blocks return at once
  const int G = std::max(1, 128 / (CH * (half_spectra ? 4 : 8)));
  int n_blk = (n_run + CH - 1) / CH;
  if (n_run == n_ch) n_blk = (n_blk + G - 1) / G * G;
  dim3 grid((unsigned)(n_blk * n_k));
  const float inv2 = 1.0f / (scale * scale);
  hipError_t e = hipErrorInvalidValue;
#define SFXG_MFMA_CASE(P, H)                                                                                     \
  e = KS == 32 ? launch_mfma_ks<P, 32, H>(grid, lds, s, spectra, stream_stride, seg_stride, seg_base, n_ch, M, RT, \
                                          CH, n_blk, G, fold_start, scale, bl_first, bl_second, n_baselines, inv2,  \
                                          partial)                                                                  \
               : launch_mfma_ks<P, 16, H>(grid, lds, s, spectra, stream_stride, seg_stride, seg_base, n_ch, M, RT, \
                                          CH, n_blk, G, fold_start, scale, bl_first, bl_second, n_baselines, inv2,  \
                                          partial);
  if (half_spectra) {
    if (precision == SFXG_PREC_FP16) SFXG_MFMA_CASE(SFXG_PREC_FP16, true)
    else if (precision == SFXG_PREC_BF16) SFXG_MFMA_CASE(SFXG_PREC_BF16, true)
    else return hipErrorInvalidValue;
  } else {
    if (precision == SFXG_PREC_FP16) SFXG_MFMA_CASE(SFXG_PREC_FP16, false)
    else if (precision == SFXG_PREC_BF16) SFXG_MFMA_CASE(SFXG_PREC_BF16, false)
    else if (precision == SFXG_PREC_SPLIT_FP16) SFXG_MFMA_CASE(SFXG_PREC_SPLIT_FP16, false)
    else return hipErrorInvalidValue;
  }
#undef SFXG_MFMA_CASE
  if (e != hipSuccess) return e;
  if (n_run < n_ch) {
    e = launch_xstep_last(static_cast<const float2 *>(spectra), stream_stride, seg_stride, seg_base, n_ch - 1, M,
                          fold_start, n_k, bl_first, bl_second, n_baselines,
                          partial + (int64_t)(n_ch - 1) * n_baselines, (int64_t)n_ch * n_baselines, 1, s);
    if (e != hipSuccess) return e;
  }
  dim3 gc((unsigned)((n_baselines + 63) / 64), (unsigned)((n_ch + 63) / 64));
  hipLaunchKernelGGL(k_fold_combine_cm, gc, dim3(256), 0, s, partial, n_k, n_ch, n_baselines, vis);
  return hipGetLastError();
}

hipError_t launch_xstep_pulsar(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int n_ch,
                               int64_t seg_base, int64_t n_segments, const int32_t *bl_first,
                               const int32_t *bl_second, int n_baselines, const double *offsets, double start_phase,
                               double phase_coef, double fft_duration, double gate_begin, double gate_end, int nbins,
                               int16_t *bins, uint32_t *masks, float2 *acc, const int32_t *tiles, int n_tiles, int M,
                               bool by_bin, hipStream_t s) {
  if (n_segments <= 0) return hipSuccess;
  const int64_t n_pairs = n_segments * n_ch;
  hipLaunchKernelGGL(k_pulsar_bin_table, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, s, bins, n_ch,
                     seg_base, n_segments, offsets, start_phase, phase_coef, fft_duration, gate_begin, gate_end, nbins);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (by_bin) {
    // stage bin masks of 16-channel blocks, then one workgroup per (channel
    // block, tile group, bin)
    const int W = (nbins + 31) / 32;
    const int n_cb = (n_ch + 15) / 16;
    const int64_t n_stages = (n_segments + MF_KS - 1) / MF_KS;
    hipLaunchKernelGGL(k_pulsar_stage_masks, dim3((unsigned)((n_stages * n_cb + 255) / 256)), dim3(256), 0, s, bins,
                       n_segments, n_ch, 16, n_cb, n_stages, W, masks);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (stream_stride % 2 || seg_stride % 64) return hipErrorInvalidValue;  // 16-byte loads inside the rows
    const int NW = n_tiles <= 24 ? 6 : 12;
    const int slots = 4 * NW;
    const int TPS = n_tiles <= slots ? 1 : n_tiles <= 2 * slots ? 2 : 3;
    const int n_tg = (n_tiles + TPS * slots - 1) / (TPS * slots);
    const size_t lds = (size_t)MF_KS * M * 8 * sizeof(float4) + (size_t)(MF_KS * 16 + 64 * NW + NW) * sizeof(int);
    if (lds > 160 * 1024 || (int64_t)n_cb * n_tg >= ((int64_t)1 << 31)) by_bin = false;  // per baseline below
  }
  if (by_bin) {
    const int W = (nbins + 31) / 32;
    const int n_cb = (n_ch + 15) / 16;
    const int NW = n_tiles <= 24 ? 6 : 12;
    const int slots = 4 * NW;
    const int TPS = n_tiles <= slots ? 1 : n_tiles <= 2 * slots ? 2 : 3;
    const int n_tg = (n_tiles + TPS * slots - 1) / (TPS * slots);
    const size_t lds = (size_t)MF_KS * M * 8 * sizeof(float4) + (size_t)(MF_KS * 16 + 64 * NW + NW) * sizeof(int);
    dim3 grid((unsigned)(n_cb * n_tg), (unsigned)nbins);
#define SFXG_XPB(NWV, T)                                                                                          \
  do {                                                                                                            \
    e = hipFuncSetAttribute((const void *)k_xstep_pulsar_bin<NWV, T>, hipFuncAttributeMaxDynamicSharedMemorySize,  \
                            (int)lds);                                                                            \
    if (e != hipSuccess) return e;                                                                                \
    hipLaunchKernelGGL((k_xstep_pulsar_bin<NWV, T>), grid, dim3(64 * NWV), lds, s, spectra, stream_stride,        \
                       seg_stride, n_ch, n_segments, M, tiles, n_tiles, n_baselines, bins, masks, W, n_tg, acc);  \
    e = hipGetLastError();                                                                                        \
  } while (0)
    if (NW == 6) SFXG_XPB(6, 1);
    else if (TPS == 1) SFXG_XPB(12, 1);
    else if (TPS == 2) SFXG_XPB(12, 2);
    else SFXG_XPB(12, 3);
#undef SFXG_XPB
    return e;
  }
  const int n_cb = (n_ch + 63) / 64, n_bg = (n_baselines + XWAVES - 1) / XWAVES;
  if ((int64_t)n_cb * n_bg >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  dim3 grid((unsigned)(n_cb * n_bg));
  hipLaunchKernelGGL(k_xstep_pulsar, grid, dim3(64 * XWAVES), 0, s, spectra, stream_stride, seg_stride, n_ch,
                     seg_base, n_segments, bl_first, bl_second, n_baselines, bins, n_bg, acc);
  return hipGetLastError();
}

// channels per workgroup of the MFMA pulsar kernel: every (channel, tile) pair
// fits MF_WAVES waves x maxp and the stage (plus `extra` bytes) fits 150 KiB of LDS
static int mf_channels(int precision, int M, int maxp, size_t extra, size_t *lds_out) {
  const int RT = (2 * M + 31) / 32;
  const int R = 32 * RT;
  const int nT = RT * (RT + 1) / 2;
  const int planes = precision == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
  if (nT > MF_WAVES * maxp) return 0;
  int CH = (MF_WAVES * maxp) / nT;
  if (CH > 64) CH = 64;
  auto lds_of = [&](int ch) { return (size_t)planes * ch * (R * MF_RS + 8) * 2 + extra; };
  while (CH > 1 && lds_of(CH) > 150 * 1024) CH--;
  if (lds_of(CH) > 150 * 1024) return 0;
  *lds_out = lds_of(CH);
  return CH;
}

hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (bytes % 4 != 0) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 4);
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)p, n);
  return hipGetLastError();
}

size_t pulsar_gram_floats(int M, int n_ch, int nbins) {
  const int RT = (2 * M + 31) / 32;
  return (size_t)nbins * n_ch * (RT * (RT + 1) / 2) * 1024;
}

template <int PREC, bool HS>
static hipError_t launch_mfma_pulsar_t(dim3 grid, size_t lds, hipStream_t s, const void *spectra,
                                       int64_t stream_stride, int64_t seg_stride, int n_ch, int M, int RT, int CH,
                                       int64_t seg_base, int64_t n_segments, int64_t segs_per_z, float scale,
                                       const double *offsets, double start_phase, double phase_coef,
                                       double fft_duration, double gate_begin, double gate_end, int nbins,
                                       float *gram) {
  hipError_t e = hipFuncSetAttribute((const void *)k_xstep_mfma_pulsar<PREC, HS>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_xstep_mfma_pulsar<PREC, HS>), grid, dim3(MF_THREADS), lds, s, spectra, stream_stride, seg_stride,
                     n_ch, M, RT, CH, seg_base, n_segments, segs_per_z, scale, offsets, start_phase, phase_coef,
                     fft_duration, gate_begin, gate_end, nbins, gram);
  return hipGetLastError();
}

// LDS beyond the stage planes of the by-bin kernel: channel masks, stage list, wave counts
constexpr size_t MFB_EXTRA = 64 * sizeof(uint32_t) + (MFB_WIN + MF_WAVES) * sizeof(int32_t) + 64 * 16 * sizeof(int16_t);

size_t pulsar_bin_mask_words(int precision, int M, int n_ch_run, int nbins, int64_t n_segments) {
  size_t lds = 0;
  const int CH = mf_channels(precision, M, MFP_MAXP, MFB_EXTRA, &lds);
  if (CH < 1) return 0;
  const int64_t n_cb = (n_ch_run + CH - 1) / CH, n_stages = (n_segments + MF_KS - 1) / MF_KS;
  return (size_t)n_stages * n_cb * ((nbins + 31) / 32);
}

template <int PREC, bool HS>
static hipError_t launch_mfma_pulsar_bin_t(dim3 grid, size_t lds, hipStream_t s, const void *spectra,
                                           int64_t stream_stride, int64_t seg_stride, int n_ch, int M, int RT, int CH,
                                           int64_t seg_base, int64_t n_segments, float scale, const int16_t *bins,
                                           const uint32_t *masks, int W, float *gram, int n_cb, int G) {
  hipError_t e = hipFuncSetAttribute((const void *)k_xstep_mfma_pulsar_bin<PREC, HS>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_xstep_mfma_pulsar_bin<PREC, HS>), grid, dim3(MF_THREADS), lds, s, spectra, stream_stride,
                     seg_stride, n_ch, M, RT, CH, seg_base, n_segments, scale, bins, masks, W, gram, n_cb, G);
  return hipGetLastError();
}

hipError_t pulsar_gram_zero(int M, int n_ch_run, int nbins, float *gram, hipStream_t s) {
  return launch_fill_zero(gram, pulsar_gram_floats(M, n_ch_run, nbins) * sizeof(float), s);
}

hipError_t launch_xstep_mfma_pulsar(int precision, const void *spectra, bool half_spectra, int64_t stream_stride,
                                    int64_t seg_stride,
                                    int ch_base, int n_ch_run, int M, int64_t seg_base, int64_t n_segments,
                                    float scale, const double *offsets, double start_phase, double phase_coef,
                                    double fft_duration, double gate_begin, double gate_end, int nbins, float *gram,
                                    int16_t *bins, uint32_t *masks, const Tune &t, hipStream_t s) {
  if (n_segments <= 0 || n_ch_run <= 0) return hipSuccess;
  if (t.mfp_wgs == 0) {
    // by bin: the bin table and the stages' bin masks of the chunk, then one
    // workgroup per (channel block, bin)
    size_t lds = 0;
    const int CH = mf_channels(precision, M, MFP_MAXP, MFB_EXTRA, &lds);
    if (CH < 1 || nbins < 2) return hipErrorInvalidValue;
    const int RT = (2 * M + 31) / 32;
    const int n_cb = (n_ch_run + CH - 1) / CH;
    const int64_t n_stages = (n_segments + MF_KS - 1) / MF_KS;
    const int W = (nbins + 31) / 32;
    const double *off = offsets + ch_base;
    const int64_t n_pairs = n_segments * n_ch_run;
    hipLaunchKernelGGL(k_pulsar_bin_table, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, s, bins, n_ch_run,
                       seg_base, n_segments, off, start_phase, phase_coef, fft_duration, gate_begin, gate_end, nbins);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pulsar_stage_masks, dim3((unsigned)((n_stages * n_cb + 255) / 256)), dim3(256), 0, s, bins,
                       n_segments, n_ch_run, CH, n_cb, n_stages, W, masks);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // groups of G = 8 channel blocks on one XCD (a group's rows span several
    // 128-byte lines, so few lines are shared with the next group)
    const int G = 8;
    const int n_cbp = (n_cb + G - 1) / G * G;
    dim3 grid((unsigned)(n_cbp * (nbins - 1 + MFB_Z0)));
    const void *sp = static_cast<const char *>(spectra) + (size_t)ch_base * (half_spectra ? 4 : 8);
#define SFXG_MFB(P, H)                                                                                            \
  return launch_mfma_pulsar_bin_t<P, H>(grid, lds, s, sp, stream_stride, seg_stride, n_ch_run, M, RT, CH, seg_base, \
                                        n_segments, scale, bins, masks, W, gram, n_cb, G)
    if (half_spectra) {
      if (precision == SFXG_PREC_FP16) SFXG_MFB(SFXG_PREC_FP16, true);
      if (precision == SFXG_PREC_BF16) SFXG_MFB(SFXG_PREC_BF16, true);
      return hipErrorInvalidValue;
    }
    if (precision == SFXG_PREC_FP16) SFXG_MFB(SFXG_PREC_FP16, false);
    if (precision == SFXG_PREC_BF16) SFXG_MFB(SFXG_PREC_BF16, false);
    if (precision == SFXG_PREC_SPLIT_FP16) SFXG_MFB(SFXG_PREC_SPLIT_FP16, false);
#undef SFXG_MFB
    return hipErrorInvalidValue;
  }
  size_t lds = 0;
  const int CH = mf_channels(precision, M, MFP_MAXP,
                            (size_t)64 * MF_KS * sizeof(int16_t) + (size_t)64 * MFP_RUNS * sizeof(int32_t), &lds);
  if (CH < 1) return hipErrorInvalidValue;
  const int RT = (2 * M + 31) / 32;
  // about mfp_wgs (1024) workgroups: channel blocks x segment ranges (multiples of a stage)
  const int64_t nblk = (n_ch_run + CH - 1) / CH;
  const int64_t wgs = t.mfp_wgs;
  int64_t nz = (wgs + nblk - 1) / nblk;
  int64_t per = (n_segments + nz - 1) / nz;
  per = (per + MF_KS - 1) / MF_KS * MF_KS;
  nz = (n_segments + per - 1) / per;
  dim3 grid((unsigned)nblk, 1, (unsigned)nz);
  const void *sp = static_cast<const char *>(spectra) + (size_t)ch_base * (half_spectra ? 4 : 8);
  const double *off = offsets + ch_base;
  if (half_spectra) {
    if (precision == SFXG_PREC_FP16)
      return launch_mfma_pulsar_t<SFXG_PREC_FP16, true>(grid, lds, s, sp, stream_stride, seg_stride, n_ch_run, M, RT,
                                                        CH, seg_base, n_segments, per, scale, off, start_phase,
                                                        phase_coef, fft_duration, gate_begin, gate_end, nbins, gram);
    if (precision == SFXG_PREC_BF16)
      return launch_mfma_pulsar_t<SFXG_PREC_BF16, true>(grid, lds, s, sp, stream_stride, seg_stride, n_ch_run, M, RT,
                                                        CH, seg_base, n_segments, per, scale, off, start_phase,
                                                        phase_coef, fft_duration, gate_begin, gate_end, nbins, gram);
    return hipErrorInvalidValue;
  }
  switch (precision) {
    case SFXG_PREC_FP16:
      return launch_mfma_pulsar_t<SFXG_PREC_FP16, false>(grid, lds, s, sp, stream_stride, seg_stride, n_ch_run, M, RT, CH,
                                                  seg_base, n_segments, per, scale, off, start_phase, phase_coef,
                                                  fft_duration, gate_begin, gate_end, nbins, gram);
    case SFXG_PREC_BF16:
      return launch_mfma_pulsar_t<SFXG_PREC_BF16, false>(grid, lds, s, sp, stream_stride, seg_stride, n_ch_run, M, RT, CH,
                                                  seg_base, n_segments, per, scale, off, start_phase, phase_coef,
                                                  fft_duration, gate_begin, gate_end, nbins, gram);
    case SFXG_PREC_SPLIT_FP16:
      return launch_mfma_pulsar_t<SFXG_PREC_SPLIT_FP16, false>(grid, lds, s, sp, stream_stride, seg_stride, n_ch_run, M,
                                                        RT, CH, seg_base, n_segments, per, scale, off, start_phase,
                                                        phase_coef, fft_duration, gate_begin, gate_end, nbins, gram);
    default: return hipErrorInvalidValue;
  }
}

hipError_t pulsar_gram_baselines(const float *gram, int n_ch, int ch_base, int n_ch_run, int M, int nbins,
                                 const int32_t *bl_first, const int32_t *bl_second, int n_baselines, float scale,
                                 float2 *vis, hipStream_t s) {
  if (n_ch_run <= 0) return hipSuccess;
  const int RT = (2 * M + 31) / 32;
  // vis rows keep the full channel stride; this run writes channels ch_base..
  hipLaunchKernelGGL(k_pulsar_gram_baselines_run, dim3((unsigned)n_ch_run, (unsigned)nbins), dim3(256), 0, s, gram,
                     n_ch_run, ch_base, n_ch, M, RT, bl_first, bl_second, n_baselines, 1.0f / (scale * scale), vis);
  return hipGetLastError();
}

hipError_t launch_auto_norms(const float2 *vis, int n_streams, int n_ch, int fft_size, int n_bins,
                             int n_baselines, double *norms, hipStream_t s) {
  dim3 grid(n_streams, n_bins);
  hipLaunchKernelGGL(k_auto_norms, grid, dim3(256), 0, s, vis, n_ch, fft_size, n_baselines, norms, n_streams);
  return hipGetLastError();
}

hipError_t launch_normalise(const float2 *vis, const double *norms, const int32_t *bl_first,
                            const int32_t *bl_second, const double *flag_factor, int n_streams, int n_baselines,
                            int n_ch, int n_bins, float2 *out, hipStream_t s) {
  const int64_t n = (int64_t)n_baselines * n_ch;
  dim3 grid((unsigned)((n + 255) / 256), n_bins);
  hipLaunchKernelGGL(k_normalise, grid, dim3(256), 0, s, vis, norms, bl_first, bl_second, flag_factor,
                     n_streams, n_baselines, n_ch, out);
  return hipGetLastError();
}

hipError_t launch_synthesize(void *dst, size_t n_bytes, int bits, int64_t first_sample, int64_t delay_samples,
                             uint64_t common_seed, uint64_t stream_seed, float rho, float threshold,
                             int64_t delay2_samples, uint64_t common2_seed, float rho2, hipStream_t s) {
  dim3 grid((unsigned)((n_bytes + 255) / 256));
  hipLaunchKernelGGL(k_synthesize, grid, dim3(256), 0, s, (uint8_t *)dst, (int64_t)n_bytes, bits, first_sample,
                     delay_samples, common_seed, stream_seed, rho, threshold, delay2_samples, common2_seed, rho2);
  return hipGetLastError();
}

}  // namespace sfxg
