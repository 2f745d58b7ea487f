/*
 * sfxc_oracle.c -- CPU restatement of the SFXC correlator-node data path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker: it is loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * shipped library. Built twice (OR_REAL=float -> *_f32, OR_REAL=double ->
 * *_f64) by oracle/Makefile.
 *
 * Functions, the reference code each follows, and what pins it
 * (tests/test_oracle_*.py):
 *   or_bit2float       bit2float_worker.cc:35-197,278-313,326-436
 *                      pinned by: exhaustive LUT closed form (256 bytes x
 *                      {1,2} bits), an independent numpy expansion of random
 *                      HEADER streams (skip/insert/invalid semantics), and
 *                      numpy byte histograms for the statistics.
 *   or_fft_c2c/rfft    FFTW3f conventions of sfxc_fft_float.cc:191-242
 *                      (unnormalised, forward exp(-2 pi i jk/N)); pinned by
 *                      brute-force FP64 DFT and numpy.fft.
 *   or_window          delay_correction.cc:299-341; pinned by closed forms.
 *   or_delay_correction delay_correction.cc:21-184; pinned by pure-tone closed
 *                      forms (fractional delay of a tone, fringe rotation of a
 *                      tone to DC) and a numpy FP64 restatement on random data.
 *   or_correlate       correlation_core.cc:20-73,241-273,515-559; pinned by a
 *                      FP64 numpy einsum over the same spectra.
 *   or_correlate_pc    the same with extra phase centres: sub_integration's
 *                      uv shift (correlation_core.cc:515-592); pinned by a
 *                      zero shift (copies of the pointing centre), a numpy
 *                      FP64 restatement of the rotation (direct exponentials,
 *                      not the recurrence) and, on the GPU path, by a second
 *                      source that only the shifted centre fringe-stops.
 *   or_pc_weights      create_weights (correlation_core.cc:712-759); pinned
 *                      by numpy.correlate of the window in FP64.
 *   or_eval_rate       gsl_spline_eval_deriv of the piecewise cubic; pinned
 *                      by finite differences.
 *   or_normalize       correlation_core.cc:275-314; pinned by closed form.
 *   or_find_invalid    correlation_core.cc:599-665; pinned by brute-force
 *                      interval arithmetic on small invalid lists.
 *   or_pulsar_*        correlation_core_pulsar.cc:46-238; pinned by the
 *                      linear-ephemeris closed form of the bin index.
 *   or_write_output    correlation_core.cc:316-477, output_header.h:70-148;
 *                      pinned by the struct formats of utils/print_corfile.py.
 *   or_irfft           FFTW3f c2r convention of sfxc_fft_float.cc:235-242
 *                      (unnormalised, imaginary parts of bins 0 and n/2
 *                      unused); pinned by numpy.fft.irfft and a brute-force
 *                      FP64 sum.
 *   or_lag_window      create_window (correlation_core.cc:668-710,761-809);
 *                      pinned by the rectangular closed form and, for the
 *                      COS / HAMMING / HANN forms, numpy.convolve of the
 *                      window function with itself.
 *   or_phased_correlate correlation_core_phased.cc:13-50,101-126; pinned by a
 *                      FP64 numpy sum over the same spectra.
 *   or_rechannelize    integration_write's fft_size != number_channels branch
 *                      (correlation_core.cc:417-440); pinned by closed forms:
 *                      a lag function inside +-nchan gives the spectrum
 *                      sampled every N_c / nchan bins, lags at +-nchan give
 *                      their mean times (-1)^k, normalisation of a pure delay
 *                      gives the window's value at that lag, mask before
 *                      normalisation is a no-op.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sfxc_oracle.h"

#ifndef OR_REAL
#define OR_REAL float
#define OR_SUFFIX _f32
#endif
#define OR_CAT(a, b) a##b
#define OR_XCAT(a, b) OR_CAT(a, b)
#define FN(name) OR_XCAT(name, OR_SUFFIX)

typedef OR_REAL real;
typedef struct {
  real re, im;
} cplx;

#define CLOCK_RATE 512000000.0 /* MAX_SAMPLE_RATE, utils.h:36 */

/* ------------------------------------------------------------------------ */
/* 1. Bit2float_worker                                                       */
/* ------------------------------------------------------------------------ */

/* bit2float_worker.cc:4-10 (reference) and the decode table of the synthetic
 * configurations ({-3.3359,-1,+1,+3.3359}, 1-bit +/-1). Code order is
 * offset binary {0:-hi, 1:-lo, 2:+lo, 3:+hi}, first sample in the LSBs. */
static const double lut2_tab[2][4] = {{-7, -2, 2, 7}, {-3.3359, -1, 1, 3.3359}};
static const double lut1_tab[2][2] = {{-5, 5}, {-1, 1}};

typedef struct {
  const or_b2f_params *p;
  const uint8_t *in;
  int64_t n_in;
  int64_t read;
  real lut[256][8];
  int32_t *counts_on, *counts_off;
  int tsys_count, tsys_on;
} b2f_state;

/* bit2float_worker.cc:326-424, with the circular buffer replaced by the
 * linear stream. Returns the sub-byte offset of the next sample. */
static int b2f_convert(b2f_state *s, real *output, int start, int nsamples) {
  const or_b2f_params *p = s->p;
  int64_t read = s->read;
  int iout = 0;
  int spb = 8 / p->bits_per_sample;
  int half = p->bits_per_sample == 2 ? (p->sample_rate / (2 * p->tsys_freq)) / 4 : 0;
  int samp_to_write = nsamples < spb - start ? nsamples : spb - start;
  uint8_t b = s->in[read];
  memcpy(&output[iout], &s->lut[b][start], samp_to_write * sizeof(real));
  if (p->bits_per_sample == 2) {
    if (s->tsys_on) s->counts_on[b]++; else s->counts_off[b]++;
    if (--s->tsys_count == 0) { s->tsys_count = half; s->tsys_on = !s->tsys_on; }
  } else {
    s->counts_on[b]++;
  }
  nsamples -= samp_to_write;
  iout += samp_to_write;
  if (samp_to_write + start == spb)
    read++;
  else {
    s->read = read;
    return start + samp_to_write;
  }
  int nbytes = nsamples / spb;
  for (int k = 0; k < nbytes; k++) {
    b = s->in[read];
    memcpy(&output[iout], &s->lut[b][0], spb * sizeof(real));
    if (p->bits_per_sample == 2) {
      if (s->tsys_on) s->counts_on[b]++; else s->counts_off[b]++;
      if (--s->tsys_count == 0) { s->tsys_count = half; s->tsys_on = !s->tsys_on; }
    } else {
      s->counts_on[b]++;
    }
    iout += spb;
    read++;
  }
  nsamples -= nbytes * spb;
  if (nsamples > 0) {
    memcpy(&output[iout], &s->lut[s->in[read]][0], nsamples * sizeof(real));
    s->read = read;
    return nsamples;
  }
  s->read = read;
  return 0;
}

/*
 * Runs one Bit2float_worker over a complete HEADER stream for one slice
 * (set_parameters :278-313, do_task :35-197 called until the slice is full).
 * out: n_ffts_per_integration * fft_size samples.
 * chunk_nfft: the nfft of every output buffer pushed (allocate_element :426-436).
 * result[0..3] = {n_chunks, n_invalid_entries, nInvalid, bytes_read}.
 * Returns 0, or -1 for a malformed stream, -2 for a stream that ends before
 * the slice is full, -3 for an output array overflow.
 */
int FN(or_bit2float)(const or_b2f_params *p, const uint8_t *in, int64_t n_in, real *out,
                     int32_t *chunk_nfft, int32_t max_chunks, int32_t *inv_start,
                     int32_t *inv_n, int32_t max_inv, int32_t *counts_on,
                     int32_t *counts_off, int32_t *result) {
  b2f_state s;
  memset(&s, 0, sizeof(s));
  s.p = p;
  s.in = in;
  s.n_in = n_in;
  s.counts_on = counts_on;
  s.counts_off = counts_off;
  for (int i = 0; i < 256; i++) {
    for (int j = 0; j < 4; j++) s.lut[i][j] = (real)lut2_tab[p->lut][(i >> (2 * j)) & 3];
    if (p->bits_per_sample == 1)
      for (int j = 0; j < 8; j++) s.lut[i][j] = (real)lut1_tab[p->lut][(i >> j) & 1];
  }
  memset(counts_on, 0, 256 * sizeof(int32_t));
  memset(counts_off, 0, 256 * sizeof(int32_t));
  if (p->bits_per_sample == 2) {
    int full = p->sample_rate / p->tsys_freq, half = p->sample_rate / (2 * p->tsys_freq);
    int tc = (int)floor(p->delay_in_samples + 0.5);
    tc %= full;
    while (tc < 0) tc += full;
    s.tsys_on = tc < half;
    tc %= half;
    tc = half - tc;
    tc /= 4;
    s.tsys_count = tc;
  }
  enum { IDLE, SEND_INVALID, SEND_DATA } state = IDLE;
  size_t bytes_left = 0, invalid_left = 0, sample_in_byte = 0;
  int cur_delay = -1; /* int8_t in the reference */
  int current_fft = 0, n_chunks = 0, n_inv = 0, n_invalid_total = 0;
  const int spb = 8 / p->bits_per_sample;
  while (current_fft < p->n_ffts_per_integration) {
    int nfft = p->nfft_max < p->n_ffts_per_integration - current_fft
                   ? p->nfft_max
                   : p->n_ffts_per_integration - current_fft;
    size_t out_size = (size_t)nfft * p->fft_size, out_index = 0;
    real *frame = out + (int64_t)current_fft * p->fft_size;
    while (out_index != out_size) {
      int64_t avail = s.n_in - s.read;
      if (state == IDLE) {
        if (avail < 3) return -2;
        uint8_t header = s.in[s.read++];
        if (header == OR_HEADER_DATA) {
          bytes_left = s.in[s.read] | (s.in[s.read + 1] << 8);
          s.read += 2;
          if (bytes_left == 0) return -1;
          state = SEND_DATA;
        } else if (header == OR_HEADER_DELAY) {
          int new_delay = (int8_t)s.in[s.read++];
          if (cur_delay < 0) {
            sample_in_byte = (size_t)new_delay;
          } else {
            int diff = new_delay - cur_delay;
            if (diff == -1 || diff > 1) {
              frame[out_index++] = 0;
            } else {
              if (sample_in_byte != 0) return -1;
              sample_in_byte = 1;
            }
          }
          cur_delay = (int8_t)new_delay;
        } else if (header == OR_HEADER_INVALID) {
          invalid_left = s.in[s.read] | (s.in[s.read + 1] << 8);
          s.read += 2;
          n_invalid_total += (int)invalid_left;
          if (n_inv >= max_inv) return -3;
          inv_start[n_inv] = current_fft * p->fft_size + (int)out_index;
          inv_n[n_inv] = (int)invalid_left;
          n_inv++;
          state = SEND_INVALID;
        } else {
          return -1;
        }
      } else if (state == SEND_INVALID) {
        size_t free_ = out_size - out_index;
        size_t to_write = free_ < invalid_left ? free_ : invalid_left;
        if (sample_in_byte > 0) {
          size_t drop = sample_in_byte < to_write ? sample_in_byte : to_write;
          sample_in_byte -= drop;
          to_write -= drop;
          invalid_left -= drop;
        }
        if (to_write > 0) {
          memset(&frame[out_index], 0, to_write * sizeof(real));
          invalid_left -= to_write;
          out_index += to_write;
        }
        if (invalid_left == 0) state = IDLE;
      } else { /* SEND_DATA */
        if (avail == 0) return -2;
        size_t free_ = out_size - out_index;
        size_t in_buf = (size_t)avail < bytes_left ? (size_t)avail : bytes_left;
        size_t samp = in_buf * spb - sample_in_byte;
        if (samp > free_) samp = free_;
        bytes_left -= (samp + sample_in_byte) / spb;
        sample_in_byte = (size_t)b2f_convert(&s, &frame[out_index], (int)sample_in_byte, (int)samp);
        out_index += samp;
        if (bytes_left == 0) {
          if (sample_in_byte != 0) return -1;
          state = IDLE;
        }
      }
    }
    if (n_chunks >= max_chunks) return -3;
    chunk_nfft[n_chunks++] = nfft;
    current_fft += nfft;
  }
  result[0] = n_chunks;
  result[1] = n_inv;
  result[2] = n_invalid_total;
  result[3] = (int32_t)s.read;
  return 0;
}

/* bit_statistics::get_statistics / get_tsys (bit_statistics.cc:33-91). */
void FN(or_bit_statistics)(int bits_per_sample, int64_t sample_rate, int64_t base_sample_rate,
                           const int32_t *on, const int32_t *off, int32_t n_invalid,
                           int32_t *levels /*5*/, int32_t *tsys /*4*/) {
  int st[5] = {0, 0, 0, 0, 0};
  int ton[4] = {0, 0, 0, 0}, toff[4] = {0, 0, 0, 0};
  if (bits_per_sample == 2) {
    for (int i = 0; i < 256; i++)
      for (int k = 0; k < 4; k++) {
        st[(i >> (2 * k)) & 3] += on[i];
        st[(i >> (2 * k)) & 3] += off[i];
        ton[(i >> (2 * k)) & 3] += on[i];
        toff[(i >> (2 * k)) & 3] += off[i];
      }
    tsys[0] = ton[1] + ton[2];
    tsys[1] = ton[0] + ton[3];
    tsys[2] = toff[1] + toff[2];
    tsys[3] = toff[0] + toff[3];
  } else {
    for (int i = 0; i < 256; i++)
      for (int j = 0; j < 8; j++) {
        st[(i >> j) & 1] += on[i];
        st[(i >> j) & 1] += off[i];
      }
    tsys[0] = tsys[1] = tsys[2] = tsys[3] = 0;
  }
  st[4] += n_invalid;
  for (int i = 0; i < 5; i++) levels[i] = (int32_t)((base_sample_rate * (int64_t)st[i]) / sample_rate);
}

/* ------------------------------------------------------------------------ */
/* 2. FFT (FFTW conventions: unnormalised, sign -1 forward, +1 backward)     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, sign;
  cplx *tw;   /* n/2 twiddles exp(sign 2 pi i k / n), computed in double */
  cplx *st;   /* per radix-4 stage of length m = n, n/4, ...: w^p, w^2p, w^3p (p < m/4), w = exp(sign 2 pi i / m) */
  cplx *work; /* the second buffer of the autosort passes */
} fft_plan;

/* Per-thread plan cache: the Python driver runs independent calls on
 * several threads. */
static _Thread_local fft_plan plans[64];
static _Thread_local int n_plans = 0;

static fft_plan *get_plan(int n, int sign) {
  for (int i = 0; i < n_plans; i++)
    if (plans[i].n == n && plans[i].sign == sign) return &plans[i];
  if (n_plans == 64) { /* evict everything; plans are cheap to rebuild */
    for (int i = 0; i < n_plans; i++) { free(plans[i].tw); free(plans[i].st); free(plans[i].work); }
    n_plans = 0;
  }
  fft_plan *pl = &plans[n_plans++];
  pl->n = n;
  pl->sign = sign;
  pl->tw = (cplx *)malloc(sizeof(cplx) * (n / 2 > 0 ? n / 2 : 1));
  for (int k = 0; k < n / 2; k++) {
    double a = sign * 2.0 * M_PI * k / n;
    pl->tw[k].re = (real)cos(a);
    pl->tw[k].im = (real)sin(a);
  }
  /* stage tables: 3 m/4 entries for each stage length m >= 4 */
  size_t tot = 0;
  for (int m = n; m >= 4; m /= 4) tot += 3 * (size_t)(m / 4);
  pl->st = (cplx *)malloc(sizeof(cplx) * (tot > 0 ? tot : 1));
  cplx *t = pl->st;
  for (int m = n; m >= 4; m /= 4) {
    for (int p = 0; p < m / 4; p++)
      for (int r = 1; r <= 3; r++) {
        double a = sign * 2.0 * M_PI * (double)r * p / m;
        t[3 * p + r - 1].re = (real)cos(a);
        t[3 * p + r - 1].im = (real)sin(a);
      }
    t += 3 * (m / 4);
  }
  pl->work = (cplx *)malloc(sizeof(cplx) * (n > 0 ? n : 1));
  return pl;
}

static inline cplx cmulr(cplx a, cplx b) {
  cplx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* Stockham autosort FFT, radix 4 with a last radix-2 pass when log2 n is odd
 * (no bit reversal): a pass over sub-transforms of length m at stride s
 * reads x[q + s (p + j m/4)] (j < 4) and writes the butterfly's outputs,
 * times w^(jp), to y[q + s (4 p + j)]; the next pass has length m/4 and
 * stride 4 s with the buffers swapped. out_in_x: whether the result must end
 * in x. */
static void fft_pass(cplx *x, cplx *y, int m, int s, int out_in_x, int sign, const cplx *st) {
  if (m == 1) {
    if (!out_in_x) memcpy(y, x, sizeof(cplx) * (size_t)s);
    return;
  }
  cplx *z = out_in_x ? x : y;
  if (m == 2) {
    for (int q = 0; q < s; q++) {
      cplx a = x[q], b = x[q + s];
      z[q].re = a.re + b.re; z[q].im = a.im + b.im;
      z[q + s].re = a.re - b.re; z[q + s].im = a.im - b.im;
    }
    return;
  }
  const int m4 = m / 4;
  for (int p = 0; p < m4; p++) {
    const cplx w1 = st[3 * p], w2 = st[3 * p + 1], w3 = st[3 * p + 2];
    for (int q = 0; q < s; q++) {
      const cplx a = x[q + s * p], b = x[q + s * (p + m4)], c = x[q + s * (p + 2 * m4)], d = x[q + s * (p + 3 * m4)];
      cplx apc = {a.re + c.re, a.im + c.im}, amc = {a.re - c.re, a.im - c.im};
      cplx bpd = {b.re + d.re, b.im + d.im};
      /* jbmd = -sign i (b - d): i (b - d) for the forward transform */
      cplx bmd = {b.re - d.re, b.im - d.im};
      cplx jbmd = sign < 0 ? (cplx){-bmd.im, bmd.re} : (cplx){bmd.im, -bmd.re};
      cplx y0 = {apc.re + bpd.re, apc.im + bpd.im};
      cplx y1 = {amc.re - jbmd.re, amc.im - jbmd.im};
      cplx y2 = {apc.re - bpd.re, apc.im - bpd.im};
      cplx y3 = {amc.re + jbmd.re, amc.im + jbmd.im};
      y[q + s * (4 * p)] = y0;
      y[q + s * (4 * p + 1)] = cmulr(y1, w1);
      y[q + s * (4 * p + 2)] = cmulr(y2, w2);
      y[q + s * (4 * p + 3)] = cmulr(y3, w3);
    }
  }
  fft_pass(y, x, m4, 4 * s, !out_in_x, sign, st + 3 * m4);
}

/* In-place complex FFT; n a power of two. */
static void fft_inplace(cplx *x, int n, int sign) {
  if (n <= 1) return;
  fft_plan *pl = get_plan(n, sign);
  fft_pass(x, pl->work, n, 1, 1, sign, pl->st);
}

void FN(or_fft_c2c)(real *data, int n, int sign) { fft_inplace((cplx *)data, n, sign); }

/* r2c of n real points -> n/2+1 complex (sfxc_fft_fftw_float::rfft). */
static void rfft(const real *in, cplx *out, int n, cplx *work) {
  int h = n / 2;
  for (int m = 0; m < h; m++) { work[m].re = in[2 * m]; work[m].im = in[2 * m + 1]; }
  fft_inplace(work, h, -1);
  fft_plan *pl = get_plan(n, -1); /* exp(-2 pi i k / n), k < n/2 */
  for (int k = 0; k <= h; k++) {
    cplx zk = work[k % h], zn = work[(h - k) % h];
    /* E = (Z[k] + conj Z[h-k]) / 2 ; O = (Z[k] - conj Z[h-k]) / (2i) */
    real er = (real)0.5 * (zk.re + zn.re), ei = (real)0.5 * (zk.im - zn.im);
    real orr = (real)0.5 * (zk.im + zn.im), oi = (real)-0.5 * (zk.re - zn.re);
    cplx w;
    if (k < h) w = pl->tw[k];
    else { w.re = -1; w.im = 0; }
    out[k].re = er + (orr * w.re - oi * w.im);
    out[k].im = ei + (orr * w.im + oi * w.re);
  }
}

void FN(or_rfft)(const real *in, real *out, int n) {
  cplx *work = (cplx *)malloc(sizeof(cplx) * (n / 2));
  rfft(in, (cplx *)out, n, work);
  free(work);
}

/* ------------------------------------------------------------------------ */
/* 3. Delay model and Delay_correction                                       */
/* ------------------------------------------------------------------------ */

static double eval_piece(const or_delay_model *m, const double *coef, int64_t ticks, double dflt) {
  if (!coef) return dflt;
  double sec = (ticks - m->scan_begin_ticks) / CLOCK_RATE;
  int i = (int)floor(sec / m->interval);
  if (i < 0) i = 0;
  if (i > m->n - 1) i = m->n - 1;
  double u = sec - i * m->interval;
  const double *c = coef + 4 * i;
  return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
}

double FN(or_eval_delay)(const or_delay_model *m, int64_t ticks) {
  return eval_piece(m, m->delay, ticks, 0.0);
}

/* Delay_table_akima::rate (delay_table_akima.cc:128-136): the derivative of
 * the piece at ticks, in seconds per second. */
double FN(or_eval_rate)(const or_delay_model *m, int64_t ticks) {
  double sec = (ticks - m->scan_begin_ticks) / CLOCK_RATE;
  int i = (int)floor(sec / m->interval);
  if (i < 0) i = 0;
  if (i > m->n - 1) i = m->n - 1;
  double u = sec - i * m->interval;
  const double *c = m->delay + 4 * i;
  return c[1] + u * (2 * c[2] + u * 3 * c[3]);
}

/* create_window (delay_correction.cc:299-341) */
void FN(or_window)(int window, int n, real *w) {
  for (int i = 0; i < n; i++) {
    double v;
    switch (window) {
      case OR_WINDOW_NONE: v = i < n / 2 ? 1 : 0; break;
      case OR_WINDOW_RECT: v = (i >= n / 4 && i < 3 * n / 4) ? 1 : 0; break;
      case OR_WINDOW_COS: v = sin(M_PI * i / (n - 1)); break;
      case OR_WINDOW_HAMMING: v = 0.54 - 0.46 * cos(2 * M_PI * i / (n - 1)); break;
      default: v = 0.5 * (1 - cos(2 * M_PI * i / (n - 1))); break;
    }
    w[i] = (real)v;
  }
}

/*
 * Delay_correction over one slice: consumes the bit2float output in the
 * chunks it was pushed (chunk_nfft, in N_d blocks) and writes every
 * correlation segment (stride complex values each) to spectra.
 * time_series (optional, may be NULL) receives the fringe-stopped real series.
 * seg_per_chunk receives the number of segments produced per chunk.
 * Returns the total number of segments.
 */
int64_t FN(or_delay_correction)(const or_dc_params *p, const or_delay_model *m, const real *samples,
                                const int32_t *chunk_nfft, int32_t n_chunks, real *spectra_out,
                                real *time_series, int32_t *seg_per_chunk) {
  const int nd = p->fft_size_delaycor;
  const int fs = p->station_sample_rate;
  const int64_t fft_cor = 2LL * p->fft_size_correlation;
  const int64_t fft_rot = (fft_cor * (int64_t)fs) / p->sample_rate;
  const int64_t tps = (int64_t)llround(CLOCK_RATE / fs); /* Time::set_sample_rate */
  const int64_t fl = (int64_t)llround((double)nd / (fs / 1000000) * (CLOCK_RATE / 1000000));
  const int64_t fl_half = (int64_t)llround(fl / 2.0);
  const double fl_sec = fl / CLOCK_RATE;
  const int oversamp = (int)round(fs / (2.0 * p->bandwidth));
  const int sb = p->sideband;
  const int stride = p->stride;

  int64_t n_total = (int64_t)p->n_ffts_per_integration * nd;
  real *tbuf = time_series ? time_series : (real *)malloc(sizeof(real) * (n_total > 0 ? n_total : 1));
  cplx *freq = (cplx *)malloc(sizeof(cplx) * nd);
  cplx *work = (cplx *)malloc(sizeof(cplx) * (fft_rot / 2 + 1));
  real *win = (real *)malloc(sizeof(real) * fft_rot);
  real *temp = (real *)malloc(sizeof(real) * fft_rot);
  int64_t tfft_len = (fft_cor > fft_rot ? fft_cor : fft_rot) / 2 + 4;
  cplx *temp_fft = (cplx *)calloc(tfft_len, sizeof(cplx));
  FN(or_window)(p->window, (int)fft_rot, win);

  /* output offsets (delay_correction.cc:226-241) */
  int64_t freq_st = p->channel_freq, delta;
  if (p->job_sideband != p->sideband) freq_st += sb * (int64_t)p->bandwidth;
  if (p->job_sideband == -1) delta = freq_st - p->job_channel_freq;
  else delta = p->job_channel_freq - freq_st;
  int64_t output_offset = 0, temp_fft_offset = 0;
  if (delta > 0) output_offset = delta * fft_cor / p->sample_rate;
  else temp_fft_offset = -delta * fft_cor / p->sample_rate;

  /* LO start phase (delay_correction.cc:251-255) */
  double dt = (p->start_ticks - p->experiment_start_ticks) / CLOCK_RATE;
  double LO = p->LO_offset;
  double start_phase = LO * (dt - trunc(dt)) + (LO - trunc(LO)) * trunc(dt);
  start_phase = start_phase - floor(start_phase);
  const double center_freq = (double)p->channel_freq + sb * p->bandwidth * 0.5 + LO;
  const double mult_factor_phi = -sb * 2.0 * M_PI;

  int64_t current = p->start_ticks;
  int64_t tbuf_start = 0, tbuf_end = 0, nseg = 0, blk = 0;
  for (int c = 0; c < n_chunks; c++) {
    int nbuffer = chunk_nfft[c];
    int64_t nfft_cor = (nbuffer * (int64_t)nd + tbuf_end - tbuf_start) / (fft_rot / 2);
    if (p->window != OR_WINDOW_NONE) nfft_cor -= 1;
    for (int buf = 0; buf < nbuffer; buf++, blk++) {
      const real *input = samples + blk * (int64_t)nd;
      /* fractional_bit_shift (delay_correction.cc:82-137) */
      double delay = eval_piece(m, m->delay, current + fl_half, 0.0);
      double dis = delay * fs;
      int integer_delay = (int)floor(dis + .5);
      double fractional_delay = dis - integer_delay;
      rfft(input, freq, nd, work);
      freq[0].re *= (real)0.5; freq[0].im *= (real)0.5;
      freq[nd / 2].re *= (real)0.5; freq[nd / 2].im *= (real)0.5;
      for (int i = nd / 2 + 1; i < nd; i++) { freq[i].re = 0; freq[i].im = 0; }
      const double dfr = (double)fs / nd;
      const double tmp1 = -2.0 * M_PI * fractional_delay / fs;
      const double tmp2 = M_PI * (integer_delay & (4 * oversamp - 1)) / (2 * oversamp);
      const double constant_term = tmp2 - tmp1 * 0.5 * p->bandwidth;
      const double linear_term = tmp1 * dfr;
      double temp_ = sin(linear_term / 2);
      double a = 2 * temp_ * temp_, b = sin(linear_term);
      double sin_phi = sin(constant_term), cos_phi = cos(constant_term);
      for (int i = 0; i < nd / 2 + 1; i++) {
        real er = (real)cos_phi, ei = (real)-sin_phi;
        real fr = freq[i].re, fi = freq[i].im;
        freq[i].re = er * fr - ei * fi;
        freq[i].im = er * fi + ei * fr;
        temp_ = sin_phi - (a * sin_phi - b * cos_phi);
        cos_phi = cos_phi - (a * cos_phi + b * sin_phi);
        sin_phi = temp_;
      }
      fft_inplace(freq, nd, +1);
      /* fringe_stopping (delay_correction.cc:139-184) */
      double lo_phase = start_phase + LO * ((current - p->start_ticks) / CLOCK_RATE);
      double phi = center_freq * eval_piece(m, m->delay, current, 0.0) + lo_phase +
                   eval_piece(m, m->phase, current, 0.0) / (2 * M_PI);
      double floor_phi = floor(phi);
      phi = mult_factor_phi * (phi - floor_phi);
      double phi_end = center_freq * eval_piece(m, m->delay, current + fl, 0.0) + lo_phase +
                       fl_sec * LO + eval_piece(m, m->phase, current + fl, 0.0) / (2 * M_PI);
      phi_end = mult_factor_phi * (phi_end - floor_phi);
      double delta_phi = (phi_end - phi) / nd;
      double amplitude = eval_piece(m, m->amp, current + fl_half, 1.0);
      temp_ = sin(delta_phi / 2);
      a = 2 * temp_ * temp_;
      b = sin(delta_phi);
      sin_phi = amplitude * sin(phi);
      cos_phi = amplitude * cos(phi);
      real *o = tbuf + tbuf_end;
      for (int i = 0; i < nd; i++) {
        o[i] = (real)(freq[i].re * cos_phi + freq[i].im * sin_phi);
        temp_ = sin_phi - (a * sin_phi - b * cos_phi);
        cos_phi = cos_phi - (a * cos_phi + b * sin_phi);
        sin_phi = temp_;
      }
      tbuf_end += nd;
      current += nd * tps;
    }
    int64_t nsamp_per_window = p->window == OR_WINDOW_NONE ? fft_rot / 2 : fft_rot;
    for (int64_t i = 0; i < nfft_cor; i++) {
      for (int64_t k = 0; k < nsamp_per_window; k++) temp[k] = tbuf[tbuf_start + k] * win[k];
      if (p->window == OR_WINDOW_NONE)
        for (int64_t k = nsamp_per_window; k < 2 * nsamp_per_window; k++) temp[k] = 0;
      tbuf_start += fft_rot / 2;
      if (p->job_sideband != p->sideband)
        for (int64_t k = 1; k < fft_rot; k += 2) temp[k] = -temp[k];
      rfft(temp, temp_fft + temp_fft_offset, (int)fft_rot, work);
      memcpy(spectra_out + 2 * (nseg * stride), temp_fft + output_offset,
             sizeof(cplx) * stride);
      nseg++;
    }
    if (seg_per_chunk) seg_per_chunk[c] = (int32_t)(nfft_cor > 0 ? nfft_cor : 0);
  }
  if (!time_series) free(tbuf);
  free(freq); free(work); free(win); free(temp); free(temp_fft);
  return nseg;
}

/* ------------------------------------------------------------------------ */
/* 4. Correlation_core                                                       */
/* ------------------------------------------------------------------------ */

/*
 * Correlation_core::do_task over one integration (correlation_core.cc:20-73),
 * with integration_step (:241-273) and sub_integration (:515-559) for one phase
 * centre. spectra[i] points at stream i's segments (stride complex each),
 * consumed in chunks of seg_per_chunk segments. vis receives
 * phase_centers[0] (n_baselines x (N_c + 1) complex, before normalisation).
 */
static void sub_integration(const or_core_params *p, const or_pc_params *pp, const cplx *acc, cplx *pc,
                            int64_t previous_fft, int64_t current_fft);

static void correlate(const or_core_params *p, const or_pc_params *pp, const real *const *spectra,
                      const int32_t *seg_per_chunk, int32_t n_chunks, real *vis) {
  const int nf = p->fft_size + 1, B = p->n_baselines, M = p->n_streams;
  const int n_pc = pp ? pp->n_phase_centers : 1;
  cplx *acc = (cplx *)calloc((size_t)B * nf, sizeof(cplx));
  cplx *pc = (cplx *)vis;
  memset(pc, 0, sizeof(cplx) * (size_t)n_pc * B * nf);
  int64_t previous_fft = 0;
  cplx *conj = (cplx *)malloc(sizeof(cplx) * nf);
  int64_t current_fft = 0, next_sub = 1, seg = 0;
  for (int c = 0; c < n_chunks; c++) {
    int nbuffer = seg_per_chunk[c];
    if (nbuffer <= 0) continue;
    for (int b = 0; b < B; b++) {
      const cplx *x = (const cplx *)spectra[p->bl_first[b]];
      const cplx *y = (const cplx *)spectra[p->bl_second[b]];
      cplx *o = acc + (size_t)b * nf;
      for (int k = 0; k < nbuffer; k++) {
        const cplx *xs = x + (seg + k) * (int64_t)p->stride;
        const cplx *ys = y + (seg + k) * (int64_t)p->stride;
        for (int j = 0; j < nf; j++) {
          real yr = ys[j].re, yi = -ys[j].im;
          real pr = xs[j].re * yr - xs[j].im * yi;
          real pi = xs[j].re * yi + xs[j].im * yr;
          o[j].re += pr;
          o[j].im += pi;
        }
      }
    }
    seg += nbuffer;
    current_fft += nbuffer;
    int fold = 0;
    if (current_fft == p->number_ffts_in_integration) fold = 1;
    else if (current_fft >= next_sub * p->number_ffts_in_sub_integration) { fold = 1; next_sub++; }
    if (fold) {
      if (n_pc > 1) {
        sub_integration(p, pp, acc, pc, previous_fft, current_fft);
      } else {
        for (size_t i = 0; i < (size_t)B * nf; i++) {
          pc[i].re += acc[i].re;
          pc[i].im += acc[i].im;
        }
      }
      memset(acc, 0, sizeof(cplx) * (size_t)B * nf);
      previous_fft = current_fft;
    }
  }
  (void)M;
  free(acc);
  free(conj);
}

void FN(or_correlate)(const or_core_params *p, const real *const *spectra, const int32_t *seg_per_chunk,
                      int32_t n_chunks, real *vis) {
  correlate(p, NULL, spectra, seg_per_chunk, n_chunks, vis);
}

/* or_correlate with pp->n_phase_centers phase centres: vis receives
 * phase_centers[0 .. n - 1], each n_baselines x (N_c + 1) complex. */
void FN(or_correlate_pc)(const or_core_params *p, const or_pc_params *pp, const real *const *spectra,
                         const int32_t *seg_per_chunk, int32_t n_chunks, real *vis) {
  correlate(p, pp, spectra, seg_per_chunk, n_chunks, vis);
}

/* uvshift (correlation_core.cc:561-592): out += amplitude * in * exp(i phi_k),
 * phi_k advanced by an FP64 sine/cosine recurrence. */
static void uvshift(const or_core_params *p, const or_pc_params *pp, const cplx *in, cplx *out, double ddelay1,
                    double ddelay2, double rate1, double rate2) {
  const int nf = p->fft_size + 1;
  const int sb = pp->sideband;
  const double base_freq = (double)pp->channel_freq;
  const double dfreq = pp->sample_rate / (2. * p->fft_size);
  real amplitude = 1;
  const int lag = abs((int)round((ddelay1 - ddelay2) * pp->sample_rate));
  if (lag < p->fft_size && pp->weights[lag] > 1e-4) amplitude = (real)(1. / pp->weights[lag]);
  double phi = base_freq * (ddelay1 * (1 - rate1) - ddelay2 * (1 - rate2));
  phi = 2 * M_PI * sb * (phi - floor(phi));
  const double delta = 2 * M_PI * dfreq * (ddelay1 * (1 - rate1) - ddelay2 * (1 - rate2));
  double temp = sin(delta / 2);
  const double a = 2 * temp * temp, b = sin(delta);
  double cos_phi = cos(phi), sin_phi = sin(phi);
  for (int i = 0; i < nf; i++) {
    const real ar = amplitude * in[i].re, ai = amplitude * in[i].im;
    const real c = (real)cos_phi, s = (real)sin_phi;
    out[i].re += ar * c - ai * s;
    out[i].im += ar * s + ai * c;
    temp = sin_phi - (a * sin_phi - b * cos_phi);
    cos_phi = cos_phi - (a * cos_phi + b * sin_phi);
    sin_phi = temp;
  }
}

/* sub_integration (correlation_core.cc:515-559) with several phase centres:
 * autos go to every centre, crosses to the pointing centre as they are and
 * to centre j uv-shifted by the delay differences at the sub-integration's
 * middle, tmid = start + tfft * (previous_fft + (current_fft - previous_fft) / 2). */
static void sub_integration(const or_core_params *p, const or_pc_params *pp, const cplx *acc, cplx *pc,
                            int64_t previous_fft, int64_t current_fft) {
  const int nf = p->fft_size + 1, B = p->n_baselines, M = p->n_streams, n_pc = pp->n_phase_centers;
  const size_t pcs = (size_t)B * nf;
  const int64_t tfft = (int64_t)p->fft_size * (int64_t)round(CLOCK_RATE / pp->sample_rate);
  const int64_t tmid =
      pp->start_ticks + (int64_t)round(tfft * (previous_fft + (current_fft - previous_fft) / 2.));
  for (int i = 0; i < M; i++)
    for (int j = 0; j < n_pc; j++)
      for (int k = 0; k < nf; k++) {
        pc[j * pcs + (size_t)i * nf + k].re += acc[(size_t)i * nf + k].re;
        pc[j * pcs + (size_t)i * nf + k].im += acc[(size_t)i * nf + k].im;
      }
  for (int i = M; i < B; i++) {
    const int s1 = p->bl_first[i], s2 = p->bl_second[i];
    for (int k = 0; k < nf; k++) {
      pc[(size_t)i * nf + k].re += acc[(size_t)i * nf + k].re;
      pc[(size_t)i * nf + k].im += acc[(size_t)i * nf + k].im;
    }
    for (int j = 1; j < n_pc; j++) {
      const double delay1 = FN(or_eval_delay)(&pp->models[s1], tmid);
      const double delay2 = FN(or_eval_delay)(&pp->models[s2], tmid);
      const double ddelay1 = FN(or_eval_delay)(&pp->models[(size_t)j * M + s1], tmid) - delay1;
      const double ddelay2 = FN(or_eval_delay)(&pp->models[(size_t)j * M + s2], tmid) - delay2;
      const double rate1 = FN(or_eval_rate)(&pp->models[s1], tmid);
      const double rate2 = FN(or_eval_rate)(&pp->models[s2], tmid);
      uvshift(p, pp, acc + (size_t)i * nf, pc + j * pcs + (size_t)i * nf, ddelay1, ddelay2, rate1, rate2);
    }
  }
}

/* create_weights (correlation_core.cc:712-759): the autocorrelation of the
 * window f(2n - 1, i), i < 2n, zero padded to 4n, through a 4n-point real
 * transform, |.|^2 and the inverse transform; weights[i] = r[i] / r[0] for
 * i < n. The window functions are those of create_window. Returns -1 for an
 * unknown window. */
static double lw_rect(int n, int i) { return (i >= n / 4 && i < (3 * n) / 4) ? 1.0 : 0.0; }
static double lw_cos(int n, int i);
static double lw_hamming(int n, int i);
static double lw_hann(int n, int i);
static void irfft(const cplx *in, real *out, int n, cplx *work);

int FN(or_pc_weights)(int window, int n, float *weights) {
  double (*f)(int, int);
  switch (window) {
    case OR_WINDOW_NONE:
    case OR_WINDOW_RECT: f = lw_rect; break;
    case OR_WINDOW_COS: f = lw_cos; break;
    case OR_WINDOW_HAMMING: f = lw_hamming; break;
    case OR_WINDOW_HANN: f = lw_hann; break;
    default: return -1;
  }
  real *tbuf = (real *)calloc(4 * (size_t)n, sizeof(real));
  cplx *fbuf = (cplx *)malloc(sizeof(cplx) * (2 * (size_t)n + 1));
  cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)n);
  for (int i = 0; i < 2 * n; i++) tbuf[i] = (real)f(2 * n - 1, i);
  rfft(tbuf, fbuf, 4 * n, work);
  for (int k = 0; k <= 2 * n; k++) {
    const real re = fbuf[k].re, im = fbuf[k].im;
    fbuf[k].re = re * re + im * im;  /* conj(F) F */
    fbuf[k].im = re * im - im * re;
  }
  irfft(fbuf, tbuf, 4 * n, work);
  for (int i = 0; i < n; i++) weights[i] = (float)(tbuf[i] / tbuf[0]);
  free(tbuf); free(fbuf); free(work);
  return 0;
}

/*
 * integration_normalize (correlation_core.cc:275-314). vis: phase-centre
 * accumulators (REAL complex); out: normalised complex<float>.
 * levels4[i] = levels[4] (invalid samples) of stream i; n_flagged_{first,second}
 * per baseline from or_find_invalid.
 */
void FN(or_normalize)(const or_core_params *p, const real *vis, const int32_t *levels4,
                      const int64_t *nfl_first, const int64_t *nfl_second, float *out) {
  const int nf = p->fft_size + 1, M = p->n_streams;
  const cplx *v = (const cplx *)vis;
  double *norms = (double *)calloc(M, sizeof(double));
  for (int i = 0; i < M; i++) {
    for (int j = 0; j < nf; j++) norms[i] += v[(size_t)i * nf + j].re;
    norms[i] /= p->fft_size;
    if (norms[i] < 1) norms[i] = 1;
    for (int j = 0; j < nf; j++) {
      out[2 * ((size_t)i * nf + j)] = (float)(real)(v[(size_t)i * nf + j].re / norms[i]);
      out[2 * ((size_t)i * nf + j) + 1] = 0.0f;
    }
  }
  const int64_t total_samples = (int64_t)p->number_ffts_in_integration * p->fft_size;
  for (int b = M; b < p->n_baselines; b++) {
    int s1 = p->bl_first[b], s2 = p->bl_second[b];
    int64_t nv1 = total_samples - levels4[s1], nv2 = total_samples - levels4[s2];
    double N1 = nv1 > 0 ? 1 - nfl_first[b] * 1. / nv1 : 1;
    double N2 = nv2 > 0 ? 1 - nfl_second[b] * 1. / nv2 : 1;
    double N = N1 * N2;
    if (N < 0.01) N = 1;
    real norm = (real)sqrt(N * norms[s1] * norms[s2]);
    for (int j = 0; j < nf; j++) {
      size_t k = (size_t)b * nf + j;
      out[2 * k] = (float)(v[k].re / norm);
      out[2 * k + 1] = (float)(v[k].im / norm);
    }
  }
  free(norms);
}

/* find_invalid (correlation_core.cc:599-665) for one baseline; scale = station
 * sample rate / job sample rate per side. Adds to nflagged[0..1]. */
void FN(or_find_invalid)(const int32_t *start0, const int32_t *n0, int32_t size0,
                         const int32_t *start1, const int32_t *n1, int32_t size1, double scale0,
                         double scale1, int64_t *nflagged_out) {
  const int32_t *st[2] = {start0, start1}, *nn[2] = {n0, n1};
  const int sz[2] = {size0, size1};
  const double scale[2] = {scale0, scale1};
  int index[2] = {0, 0};
  int nflagged[2] = {0, 0};
  int is[2] = {INT32_MAX, INT32_MAX}, ie[2] = {INT32_MAX, INT32_MAX};
  for (int i = 0; i < 2; i++) {
    if (sz[i] > 0) {
      is[i] = (int)floor(st[i][0] / scale[i]);
      ie[i] = is[i] + (int)floor(nn[i][index[i]] / scale[i]);
      index[i]++;
    }
  }
  int i = is[0] < is[1] ? 0 : 1;
  while (is[i] < INT32_MAX) {
    if (is[0] == is[1]) {
      if (ie[0] == ie[1]) {
        is[0] = ie[0] + 1;
        is[1] = ie[1] + 1;
      } else {
        int j = ie[0] < ie[1] ? 0 : 1;
        is[1 - j] = ie[j];
        is[j] = ie[j] + 1;
      }
    } else if (ie[i] < is[1 - i]) {
      nflagged[1 - i] += ie[i] - is[i];
      is[i] = ie[i] + 1;
    } else {
      nflagged[1 - i] += is[1 - i] - is[i];
      is[i] = is[1 - i];
    }
    for (int j = 0; j < 2; j++) {
      if (is[j] > ie[j]) {
        if (index[j] < sz[j]) {
          is[j] = (int)floor(st[j][index[j]] / scale[j]);
          ie[j] = is[j] + (int)floor(nn[j][index[j]] / scale[j]);
          index[j]++;
        } else {
          is[j] = INT32_MAX;
          ie[j] = INT32_MAX;
        }
      }
    }
    i = is[0] < is[1] ? 0 : 1;
  }
  nflagged_out[0] += nflagged[0];
  nflagged_out[1] += nflagged[1];
}

/* ------------------------------------------------------------------------ */
/* 5. Correlation_core_pulsar                                                */
/* ------------------------------------------------------------------------ */

/*
 * Polyco evaluation of Correlation_core_pulsar::set_parameters
 * (correlation_core_pulsar.cc:46-93): start phase, phase slope coefficient
 * and per-channel dispersion offsets. coef has n_coef entries.
 * out3 = {start_phase, coef[1] + 60 * ref_freq, fft_duration (us)}.
 */
void FN(or_pulsar_setup)(double start_mjd, int window, int fft_size, int sample_rate, int sideband,
                         int64_t channel_freq, double tmid, double DM, double ref_phase,
                         double ref_freq, double obs_freq, const double *coef, int n_coef,
                         double *offsets, double *out3) {
  const double us_per_day = (double)((int64_t)24 * 60 * 60 * 1000000);
  double fft_duration = ((double)fft_size * 1000000) / sample_rate;
  double DT;
  if (window == OR_WINDOW_NONE) DT = ((start_mjd - tmid) + fft_duration / (2 * us_per_day)) * 1440;
  else DT = ((start_mjd - tmid) + fft_duration / us_per_day) * 1440;
  int N = n_coef;
  double start_phase = 0;
  for (int i = N - 1; i > 0; i--) start_phase = (start_phase + coef[i]) * DT;
  start_phase += (ref_phase - floor(ref_phase)) + DT * 60 * ref_freq + coef[0];
  double base_freq = channel_freq * 1e-6;
  double dfreq = sideband * sample_rate * 1e-6 / (2 * fft_size);
  double inv_freq_obs2 = 1 / (obs_freq * obs_freq);
  double freq = 0;
  for (int i = N - 1; i > 1; i--) freq = (freq + i * coef[i]) * DT;
  freq = (freq + coef[1]) / 60 + ref_freq;
  for (int i = 0; i < fft_size + 1; i++)
    offsets[i] = 4148.808 * DM * (1 / pow(base_freq + i * dfreq, 2) - inv_freq_obs2) * freq;
  out3[0] = start_phase;
  out3[1] = coef[1] + 60 * ref_freq;
  out3[2] = fft_duration;
}

/* get_phase (correlation_core_pulsar.h:40-44) followed by the bin rule of
 * dedisperse_buffer (correlation_core_pulsar.cc:210-227). */
static void pulsar_bins(const or_pulsar_params *pp, int nf, int64_t current_fft, int *bins) {
  const int64_t us_per_day = (int64_t)24 * 60 * 60 * 1000000;
  double obs = pp->start_phase + pp->phase_coef * pp->fft_duration * current_fft * 1440 / us_per_day;
  double len = pp->gate_end - pp->gate_begin;
  for (int j = 0; j < nf; j++) {
    double phase = obs - pp->offsets[j];
    phase = phase - floor(phase);
    if (phase >= pp->gate_begin) {
      if (phase < pp->gate_end) bins[j] = (int)((phase - pp->gate_begin) * (pp->nbins - 1) / len) + 1;
      else bins[j] = 0;
    } else if (phase + 1 < pp->gate_end) {
      bins[j] = (int)((phase + 1 - pp->gate_begin) * (pp->nbins - 1) / len) + 1;
    } else {
      bins[j] = 0;
    }
  }
}

void FN(or_pulsar_bins)(const or_pulsar_params *pp, int nf, int64_t current_fft, int32_t *bins) {
  pulsar_bins(pp, nf, current_fft, bins);
}

/*
 * Correlation_core_pulsar::do_task over one integration
 * (correlation_core_pulsar.cc:101-146,188-238). acc: nbins x B x (N_c+1)
 * complex accumulators.
 */
void FN(or_pulsar_correlate)(const or_core_params *p, const or_pulsar_params *pp,
                             const real *const *spectra, int64_t n_segments, real *acc_out) {
  const int nf = p->fft_size + 1, B = p->n_baselines;
  cplx *acc = (cplx *)acc_out;
  memset(acc, 0, sizeof(cplx) * (size_t)pp->nbins * B * nf);
  int *bins = (int *)malloc(sizeof(int) * nf);
  cplx *ded = (cplx *)malloc(sizeof(cplx) * nf);
  for (int64_t s = 0; s < n_segments; s++) {
    pulsar_bins(pp, nf, s, bins);
    for (int b = 0; b < B; b++) {
      const cplx *x = (const cplx *)spectra[p->bl_first[b]] + s * p->stride;
      const cplx *y = (const cplx *)spectra[p->bl_second[b]] + s * p->stride;
      for (int j = 0; j < nf; j++) {
        real yr = y[j].re, yi = -y[j].im;
        ded[j].re = 0 + (x[j].re * yr - x[j].im * yi);
        ded[j].im = 0 + (x[j].re * yi + x[j].im * yr);
      }
      for (int j = 0; j < nf; j++) {
        cplx *o = acc + ((size_t)bins[j] * B + b) * nf + j;
        o->re += ded[j].re;
        o->im += ded[j].im;
      }
    }
  }
  free(bins);
  free(ded);
}

/* ------------------------------------------------------------------------ */
/* 6. Output layout                                                          */
/* ------------------------------------------------------------------------ */

/*
 * integration_write (correlation_core.cc:316-477) for one phase centre or
 * bin, fft_size == number_channels. Returns the number of bytes written.
 * Per stream i: station[i], pol[i] (0 R, 1 L), stats5[5*i..] = levels[0..4]
 * as the output header stores them (already mapped for 1-bit). uvw[3*i..] of
 * stream i. weight per baseline precomputed by the caller.
 */
int64_t FN(or_write_output)(int32_t sourcenr, int32_t integration_slice, int n_streams,
                            const int32_t *station, const int32_t *pol, int sideband_u,
                            int frequency_nr, const double *uvw, const int32_t *stats5,
                            int n_baselines, const int32_t *bl_first, const int32_t *bl_second,
                            const int32_t *weight, int nchan, const float *vis, uint8_t *dst) {
  uint8_t *o = dst;
  memcpy(o, &sourcenr, 4); o += 4;
  int nst = 0;
  int32_t seen[256];
  int32_t first_stream[256];
  for (int i = 0; i < n_streams; i++) {
    int found = 0;
    for (int k = 0; k < nst; k++) if (seen[k] == station[i]) found = 1;
    if (!found) { seen[nst] = station[i]; first_stream[nst] = i; nst++; }
  }
  int32_t ts[4] = {integration_slice, n_baselines, nst, n_streams};
  memcpy(o, ts, 16); o += 16;
  for (int k = 0; k < nst; k++) {
    int32_t hdr[2] = {seen[k], 0};
    memcpy(o, hdr, 8); o += 8;
    memcpy(o, &uvw[3 * first_stream[k]], 24); o += 24;
  }
  for (int i = 0; i < n_streams; i++) {
    uint8_t h[4] = {(uint8_t)station[i], (uint8_t)frequency_nr, (uint8_t)sideband_u, (uint8_t)pol[i]};
    memcpy(o, h, 4); o += 4;
    memcpy(o, &stats5[5 * i], 20); o += 20;
  }
  for (int b = 0; b < n_baselines; b++) {
    memcpy(o, &weight[b], 4); o += 4;
    o[0] = (uint8_t)station[bl_first[b]];
    o[1] = (uint8_t)station[bl_second[b]];
    o[2] = (uint8_t)((pol[bl_first[b]] & 1) | ((pol[bl_second[b]] & 1) << 1) |
                     ((sideband_u & 1) << 2) | ((frequency_nr & 31) << 3));
    o[3] = ' ';
    o += 4;
    memcpy(o, vis + 2 * (size_t)b * (nchan + 1), sizeof(float) * 2 * (nchan + 1));
    o += sizeof(float) * 2 * (nchan + 1);
  }
  return (int64_t)(o - dst);
}

/* ------------------------------------------------------------------------ */
/* 7. Lag re-channelisation (fft_size_correlation > number_channels)         */
/* ------------------------------------------------------------------------ */

/* c2r of n real points from n/2 + 1 complex values (sfxc_fft_fftw_float::irfft,
 * sfxc_fft_float.cc:235-242): FFTW's unnormalised backward transform of a
 * Hermitian half spectrum, which does not use the imaginary parts of bins 0
 * and n/2. Computed as an n/2-point complex transform of
 * Z[k] = (X[k] + conj X[n/2-k]) + i exp(+2 pi i k/n) (X[k] - conj X[n/2-k]),
 * whose output holds out[2m] + i out[2m+1]. */
static void irfft(const cplx *in, real *out, int n, cplx *work) {
  const int h = n / 2;
  fft_plan *pl = get_plan(n, -1); /* exp(-2 pi i k / n), k < n/2 */
  for (int k = 0; k < h; k++) {
    cplx a = in[k], b = in[h - k];
    if (k == 0) { a.im = 0; b.im = 0; }
    real sr = a.re + b.re, si = a.im - b.im;
    real dr = a.re - b.re, di = a.im + b.im;
    real wr = pl->tw[k].re, wi = -pl->tw[k].im;
    real tr = dr * wr - di * wi, ti = dr * wi + di * wr;
    work[k].re = sr - ti;
    work[k].im = si + tr;
  }
  fft_inplace(work, h, +1);
  for (int m = 0; m < h; m++) { out[2 * m] = work[m].re; out[2 * m + 1] = work[m].im; }
}

void FN(or_irfft)(const real *in, real *out, int n) {
  cplx *work = (cplx *)malloc(sizeof(cplx) * (n / 2));
  irfft((const cplx *)in, out, n, work);
  free(work);
}

/* The window functions and their self-convolution (correlation_core.cc:668-710). */
static double lw_cos(int n, int i) { return sin(M_PI * i / n); }
static double lw_hamming(int n, int i) { return 0.54 - 0.46 * cos(2 * M_PI * i / n); }
static double lw_hann(int n, int i) { return 0.5 * (1 - cos(2 * M_PI * i / n)); }

static double lw_convolve(double (*f)(int, int), int n, int i) {
  double sum = 0.0;
  i += n / 2;
  for (int j = 0; j <= n; j++) {
    if (i - j < 0 || i - j > n) continue;
    sum += f(n, j) * f(n, i - j);
  }
  return sum;
}

/* create_window without a user window (correlation_core.cc:761-809): the lag
 * window of 2 nchan points for the delay-correction window `window` applied
 * over 2 fft_size points. (m / n) is an integer quotient, as in the
 * reference. Returns -1 for an unknown window. */
int FN(or_lag_window)(int window, int nchan, int fft_size, float *w) {
  const int n = 2 * nchan, m = 2 * fft_size;
  double (*f)(int, int);
  switch (window) {
    case OR_WINDOW_NONE:
    case OR_WINDOW_RECT:
      w[0] = 1.0f;
      for (int i = 1; i < n / 2; i++) w[i] = w[n - i] = (float)((1. - 2. * i / n) / (1. - 2. * i / m));
      w[n / 2] = 0.0f;
      return 0;
    case OR_WINDOW_COS: f = lw_cos; break;
    case OR_WINDOW_HAMMING: f = lw_hamming; break;
    case OR_WINDOW_HANN: f = lw_hann; break;
    default: return -1;
  }
  w[0] = (float)((m / n) * lw_convolve(f, n, n / 2) / lw_convolve(f, m, m / 2));
  for (int i = 1; i < n / 2; i++)
    w[i] = w[n - i] = (float)((m / n) * lw_convolve(f, n, n / 2 + i) / lw_convolve(f, m, m / 2 + i));
  w[n / 2] = (float)((m / n) * lw_convolve(f, n, n) / lw_convolve(f, m, (n + m) / 2));
  return 0;
}

/* integration_write for fft_size != number_channels (correlation_core.cc:417-440),
 * rows of normalised visibilities: optional normalisation of every bin to
 * unit magnitude, the mask, a 2 N_c-point c2r transform to lags, the fold of
 * lags -nchan..nchan into 2 nchan points (lag +-nchan averaged), the lag
 * window, a 2 nchan-point r2c transform and division by 2 N_c.
 * in: rows x (N_c + 1) complex<float>; out: rows x (nchan + 1) complex<float>. */
void FN(or_rechannelize)(const or_rechan_params *p, int rows, const float *in, float *out) {
  const int nc = p->fft_size, nch = p->nchan;
  cplx *X = (cplx *)malloc(sizeof(cplx) * (nc + 1));
  cplx *work = (cplx *)malloc(sizeof(cplx) * nc);
  real *r = (real *)malloc(sizeof(real) * 2 * nc);
  real *y = (real *)malloc(sizeof(real) * 2 * nch);
  cplx *Y = (cplx *)malloc(sizeof(cplx) * (nch + 1));
  for (int row = 0; row < rows; row++) {
    const float *x = in + 2 * (size_t)row * (nc + 1);
    for (int j = 0; j <= nc; j++) {
      real re = x[2 * j], im = x[2 * j + 1];
      if (p->normalize) {
        real a = (real)hypot((double)re, (double)im);
        if (a != 0) { re /= a; im /= a; }
      }
      X[j].re = re * (real)p->mask[j];
      X[j].im = im * (real)p->mask[j];
    }
    irfft(X, r, 2 * nc, work);
    for (int i = 0; i < nch; i++) y[i] = r[i];
    y[nch] = (r[nch] + r[2 * nc - nch]) / 2;
    for (int j = 1; j < nch; j++) y[nch + j] = r[2 * nc - nch + j];
    for (int i = 0; i < 2 * nch; i++) y[i] *= (real)p->window[i];
    rfft(y, Y, 2 * nch, work);
    float *o = out + 2 * (size_t)row * (nch + 1);
    for (int j = 0; j <= nch; j++) {
      o[2 * j] = (float)(Y[j].re / (real)(2 * nc));
      o[2 * j + 1] = (float)(Y[j].im / (real)(2 * nc));
    }
  }
  free(X); free(work); free(r); free(y); free(Y);
}

/* ------------------------------------------------------------------------ */
/* 8. Phased-array core                                                      */
/* ------------------------------------------------------------------------ */

/* Correlation_core_phased (correlation_core_phased.cc:13-50,101-126): for
 * every segment in order, the spectra of all streams added, in stream order,
 * to the accumulator of its sub-integration (segment / number_ffts_in_sub_
 * integration). rows receives number_ffts_in_integration /
 * number_ffts_in_sub_integration + 1 accumulators of N_c + 1 complex values.
 * Returns the number of rows. */
int FN(or_phased_correlate)(const or_core_params *p, const real *const *spectra, int64_t nseg, real *rows) {
  const int nf = p->fft_size + 1, M = p->n_streams;
  const int nsub = p->number_ffts_in_sub_integration;
  const int nrows = p->number_ffts_in_integration / nsub + 1;
  cplx *acc = (cplx *)rows;
  memset(acc, 0, sizeof(cplx) * (size_t)nrows * nf);
  for (int64_t seg = 0; seg < nseg; seg++) {
    const int64_t sub = seg / nsub;
    if (sub >= nrows) break;
    cplx *o = acc + sub * nf;
    for (int i = 0; i < M; i++) {
      const cplx *x = (const cplx *)spectra[i] + seg * (int64_t)p->stride;
      for (int j = 0; j < nf; j++) {
        o[j].re += x[j].re;
        o[j].im += x[j].im;
      }
    }
  }
  return nrows;
}
