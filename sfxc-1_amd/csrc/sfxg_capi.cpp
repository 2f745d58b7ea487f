// C ABI of the correlator-node data path (include/sfxg.h): parameter
// handling, host-side parsing, device memory and kernel scheduling for one
// time slice at a time.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/sfxg.h"
#include "sfxg_internal.h"

using namespace sfxg;

namespace {

struct DBuf {
  void *p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) n = bytes;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

struct ModelCopy {
  int64_t scan_begin = 0;
  double interval = 1.0;
  int n = 0;
  std::vector<double> delay, phase, amp;
};

// Delay_table_akima::delay / rate (delay_table_akima.cc:114-136) on the
// piecewise cubic: value and derivative of the piece holding ticks.
static double model_eval(const ModelCopy &m, int64_t ticks, bool derivative) {
  const double sec = (double)(ticks - m.scan_begin) / kClockRate;
  int64_t i = (int64_t)floor(sec / m.interval);
  if (i < 0) i = 0;
  if (i > m.n - 1) i = m.n - 1;
  const double u = sec - (double)i * m.interval;
  const double *c = &m.delay[4 * i];
  if (derivative) return c[1] + u * (2 * c[2] + u * 3 * c[3]);
  return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
}

const double kLut2[2][4] = {{-7, -2, 2, 7}, {-3.3359, -1, 1, 3.3359}};  // bit2float_worker.cc:4-6
const double kLut1[2][2] = {{-5, 5}, {-1, 1}};                          // bit2float_worker.cc:8-10

int nr_ffts_per_integration_slice(int64_t integration_ticks, int data_rate, int fft_size) {
  // Control_parameters::nr_ffts_per_integration_slice (control_parameters.h:301-307),
  // called with (int)integration_time.get_time_usec()
  const int usec = (int)((double)integration_ticks / (kClockRate / 1000000));
  const int64_t t_int = (int64_t)llround(usec * (kClockRate / 1000000));
  const int64_t t_fft = (int64_t)llround(fft_size / (data_rate / 1000000.) * (kClockRate / 1000000));
  return (int)floor((double)t_int / (double)t_fft);
}

}  // namespace

// Process-wide table of tuned launch shapes, by slice shape.
std::mutex g_tune_mu;
std::map<std::string, Tune> g_tune;

// A run of whole sub-integrations processed together: segments [seg0, seg1)
// (folds fold0 .. fold1 - 1); the real series of blocks [blk0, blk1).
struct Chunk {
  int64_t seg0, seg1, blk0, blk1;
  int fold0, fold1;
};

// Launch parameters of a prepared slice's device work.
struct LaunchPlan {
  std::vector<Chunk> chunks;
  int64_t chunk_segs = 0;   // spectra rows per stream (largest chunk)
  int64_t ts_stride = 0;    // real-series floats per stream (largest chunk)
  bool fused = false, mixed = false;
  bool fused_ok = false;    // the fused F-engine can run this plan (FP32 series, not kept)
  bool key_fused = false;   // the shape's tuning key: N_c within the fused F-engine's range
  int k_group = 1, psr_run = 0;
  size_t vis_elems = 0;
  int64_t max_counted = 0;
  DevJob job{};
  float mixed_scale = 1.0f;
  double start_phase = 0.0, phase_coef = 0.0, fft_duration = 0.0;  // pulsar
  // phased array: per chunk, its first block / row in the tables and their counts
  std::vector<int32_t> ph_b0, ph_nb, ph_r0, ph_nr;
};

struct sfxg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  sfxg_options opt{SFXG_LUT_REFERENCE, SFXG_PREC_FP32, 0, 0, 0};
  bool fused = false;  // the last slice ran the fused F-engine
  float mixed_scale = 1.0f;
  std::string err;

  // slice parameters
  bool have_params = false;
  sfxg_correlation_parameters p{};
  std::vector<sfxg_station_parameters> st;
  std::vector<ModelCopy> models;
  std::vector<double> uvw;
  bool pulsar = false;
  sfxg_pulsar psr{};
  std::vector<double> psr_coef;

  // derived
  int M = 0, nd = 0, nc = 0, nchan = 0, B = 0, nbins = 1;
  int64_t n_ffts_per_integration = 0, n_blocks = 0, n_samples = 0, nseg = 0, S = 0;
  int n_integration = 0, n_sub = 0, nfft_max = 0;
  std::vector<int32_t> bl_first, bl_second;
  std::vector<int32_t> h_tiles;  // register-tile table of the FP32 X-step
  std::vector<StreamPlan> plans;
  std::vector<int32_t> seg_per_chunk, folds;
  std::vector<double> flag_factor;
  std::vector<int64_t> n_flagged_first;
  int64_t total_samples = 0;

  // device buffers
  DBuf payload, runs, block_run, counted, dstreams, time_series, spectra, vis, norm_out, hist, norms,
      flags, bl, fold_buf, partial, gram, block_consts, model_buf, tw_nd, tw_n, tw_2n, tw_split, window, offsets, samples,
      chan_scratch, tiles, psr_bins, psr_masks;
  // Mask_parameters (control_parameters.h:107-114) and the lag
  // re-channelisation tables built from them (correlation_core.cc:761-824)
  bool mask_normalize = false;
  std::vector<double> mask_param, window_param;
  uint64_t mask_version = 0;
  int rc_nc = -1, rc_nchan = -1, rc_window = -1;
  uint64_t rc_version = ~(uint64_t)0;
  bool rc_normalize = false;  // normalize as the prepared slice uses it
  // extra phase centres (sfxg_set_phase_centers)
  int n_pc = 1;
  bool have_pc = true;
  std::vector<ModelCopy> pc_models;  // (n_pc - 1) x M, centre-major
  std::vector<double> pc_uvw;        // n_pc x M x 3
  std::vector<int32_t> pc_source;    // n_pc
  std::vector<UvShift> h_uv;         // [fold][centre - 1][baseline] of the enqueued slice
  DBuf uv_tab;
  int pcw_window = -1, pcw_nc = -1;  // create_weights cache (window, N_c)
  std::vector<float> pcw_win;        // the window f(2 N_c - 1, i) as stored (float)
  std::vector<double> pcw_r;         // autocorrelation by lag, NaN until computed
  // raw VDIF ingest: the gathered payload of each stream, header words, pieces
  std::vector<DBuf> vdif_payload;
  DBuf vdif_words, vdif_pieces, vdif_dst;
  // phased array (Correlation_core_phased) as of sfxg_set_parameters
  bool phased = false;
  DBuf ph_blocks, ph_rows;
  std::vector<PhasedBlock> h_ph_blocks;  // host staging of the enqueued slice
  std::vector<PhasedRow> h_ph_rows;
  DBuf rc_tw, rc_mask, rc_window_buf, chan_out;
  // host staging of the enqueued slice (kept until the next slice)
  std::vector<double> h_model, h_offsets;
  std::vector<DevStream> h_streams;
  std::vector<int32_t> h_bl, h_fold_start;
  int tables_nc = -1;
  int tables_window = -1;
  double window_sumsq = 0.0;  // sum of the uploaded window's squares
  bool enqueued = false;
  bool have_results = false;  // the current slice ran (correlate or a batch launch)
  LaunchPlan lp;
  Tune tune;              // launch shapes of the current slice (from the tuning table)
  bool prepared = false;  // the device work of the current slice can be enqueued
  int batch_member = 0;   // number of batches holding this context
  std::vector<struct sfxg_batch *> batches;  // those batches
  uint64_t prep_gen = 0;  // counts sfxg_prepare calls (graph reuse check)
  hipEvent_t batch_done = nullptr;  // completion of a batch graph holding this context
  hipEvent_t prep_done = nullptr;   // the slice's uploads are complete
  std::vector<hipEvent_t> cev;       // 4 per chunk: F-engine start, delay end, channeliser end, X-step end
  int timed_chunks = 0;              // chunks of the slice whose events are pending
  hipEvent_t done = nullptr;
  hipEvent_t ev[SFXG_N_KERNELS + 1] = {};
  bool times_pending = false;  // events of an enqueued slice not yet read
  double kt_sum[SFXG_N_KERNELS] = {};
  int64_t kt_count = 0;

  int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return SFXG_OK;
    return fail(SFXG_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  }
};

#define HIPCHK(ctx, call)                         \
  do {                                            \
    int _rc = (ctx)->hip((call), #call);          \
    if (_rc != SFXG_OK) return _rc;               \
  } while (0)

// Adds the kernel durations of the last enqueued slice to the running sums
// (the caller has synchronised the stream).
static void collect_times(sfxg_ctx *ctx) {
  if (!ctx->times_pending) return;
  auto span = [](hipEvent_t a, hipEvent_t b) {
    float t = 0.0f;
    return hipEventElapsedTime(&t, a, b) == hipSuccess ? (double)t : 0.0;
  };
  ctx->kt_sum[SFXG_KERNEL_STATS] += span(ctx->ev[SFXG_KERNEL_STATS], ctx->ev[SFXG_KERNEL_DELAY]);
  ctx->kt_sum[SFXG_KERNEL_NORMALISE] += span(ctx->ev[SFXG_KERNEL_NORMALISE], ctx->ev[SFXG_N_KERNELS]);
  for (int c = 0; c < ctx->timed_chunks; c++) {
    hipEvent_t *e = &ctx->cev[4 * c];
    ctx->kt_sum[SFXG_KERNEL_DELAY] += span(e[0], e[1]);
    // the fused F-engine runs the channeliser inside the delay stage: no
    // channeliser launch to time
    if (!ctx->fused) ctx->kt_sum[SFXG_KERNEL_CHANNELISE] += span(e[1], e[2]);
    ctx->kt_sum[SFXG_KERNEL_XSTEP] += span(e[2], e[3]);
  }
  ctx->kt_count++;
  ctx->times_pending = false;
}

static void sync_slice(sfxg_ctx *ctx) {
  if (ctx->batch_done) {
    (void)hipEventSynchronize(ctx->batch_done);
    ctx->batch_done = nullptr;
  }
  if (!ctx->enqueued) return;
  (void)hipStreamSynchronize(ctx->stream);
  collect_times(ctx);
}

extern "C" {

int sfxg_create(int device, sfxg_ctx **out) {
  if (!out) return SFXG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return SFXG_ERR_DEVICE;
  std::unique_ptr<sfxg_ctx> ctx(new sfxg_ctx());
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) return SFXG_ERR_DEVICE;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return SFXG_ERR_DEVICE;
  if (hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming) != hipSuccess) return SFXG_ERR_DEVICE;
  if (hipEventCreateWithFlags(&ctx->prep_done, hipEventDisableTiming) != hipSuccess) return SFXG_ERR_DEVICE;
  for (hipEvent_t &e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return SFXG_ERR_DEVICE;
  *out = ctx.release();
  return SFXG_OK;
}

static void leave_batches(sfxg_ctx *ctx);

void sfxg_destroy(sfxg_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  leave_batches(ctx);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DBuf *bufs[] = {&ctx->payload, &ctx->runs, &ctx->block_run, &ctx->counted, &ctx->dstreams,
                  &ctx->time_series, &ctx->spectra, &ctx->vis, &ctx->norm_out, &ctx->hist,
                  &ctx->norms, &ctx->flags, &ctx->bl, &ctx->fold_buf, &ctx->partial, &ctx->gram, &ctx->block_consts, &ctx->model_buf, &ctx->tw_nd,
                  &ctx->tw_n, &ctx->tw_2n, &ctx->tw_split, &ctx->window, &ctx->offsets, &ctx->samples, &ctx->chan_scratch,
                  &ctx->tiles, &ctx->psr_bins, &ctx->psr_masks, &ctx->rc_tw, &ctx->rc_mask, &ctx->rc_window_buf, &ctx->chan_out,
                  &ctx->ph_blocks, &ctx->ph_rows, &ctx->uv_tab};
  for (DBuf *b : bufs) b->release();
  for (DBuf &b : ctx->vdif_payload) b.release();
  ctx->vdif_words.release();
  ctx->vdif_pieces.release();
  ctx->vdif_dst.release();
  sync_slice(ctx);
  if (ctx->done) (void)hipEventDestroy(ctx->done);
  if (ctx->prep_done) (void)hipEventDestroy(ctx->prep_done);
  for (hipEvent_t e : ctx->cev) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *sfxg_last_error(const sfxg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sfxg_set_options(sfxg_ctx *ctx, const sfxg_options *opt) {
  if (!ctx || !opt) return SFXG_ERR_ARG;
  if (opt->lut != SFXG_LUT_REFERENCE && opt->lut != SFXG_LUT_CONFIG)
    return ctx->fail(SFXG_ERR_ARG, "unknown lut %d", opt->lut);
  if (opt->precision < SFXG_PREC_FP32 || opt->precision > SFXG_PREC_SPLIT_FP16)
    return ctx->fail(SFXG_ERR_ARG, "unsupported precision %d", opt->precision);
  if (opt->chunk_mib < 0) return ctx->fail(SFXG_ERR_ARG, "negative chunk_mib");
  if (opt->phased_array != 0 && opt->phased_array != 1) return ctx->fail(SFXG_ERR_ARG, "phased_array is 0 or 1");
  ctx->opt = *opt;
  return SFXG_OK;
}

int sfxg_set_phase_centers(sfxg_ctx *ctx, int32_t n_phase_centers, const sfxg_delay_model *models,
                           const double *uvw, const int32_t *source_nr) {
  if (!ctx) return SFXG_ERR_ARG;
  if (!ctx->have_params) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_parameters first");
  if (n_phase_centers != ctx->n_pc || n_phase_centers < 2)
    return ctx->fail(SFXG_ERR_ARG, "%d phase centres given, the parameters have %d", n_phase_centers, ctx->n_pc);
  if (!models) return ctx->fail(SFXG_ERR_ARG, "no delay models");
  const int M = ctx->M, n = n_phase_centers;
  ctx->pc_models.assign((size_t)(n - 1) * M, ModelCopy());
  for (size_t k = 0; k < ctx->pc_models.size(); k++) {
    const sfxg_delay_model &src = models[k];
    if (src.n_pieces <= 0 || !src.delay)
      return ctx->fail(SFXG_ERR_ARG, "phase centre %d, stream %d: empty delay model", (int)(k / M) + 1, (int)(k % M));
    ModelCopy &m = ctx->pc_models[k];
    m.scan_begin = src.scan_begin;
    m.interval = src.interval;
    m.n = src.n_pieces;
    m.delay.assign(src.delay, src.delay + 4 * m.n);
  }
  ctx->pc_uvw.assign((size_t)n * M * 3, 0.0);
  if (uvw) std::copy(uvw, uvw + (size_t)n * M * 3, ctx->pc_uvw.begin());
  ctx->pc_source.assign(n, 0);
  if (source_nr) std::copy(source_nr, source_nr + n, ctx->pc_source.begin());
  ctx->have_pc = true;
  ctx->prepared = false;
  return SFXG_OK;
}

int sfxg_set_mask(sfxg_ctx *ctx, int32_t normalize, const double *mask, int32_t n_mask, const double *window,
                  int32_t n_window) {
  if (!ctx) return SFXG_ERR_ARG;
  if (n_mask < 0 || n_window < 0 || (n_mask > 0 && !mask) || (n_window > 0 && !window))
    return ctx->fail(SFXG_ERR_ARG, "mask or window size without values");
  ctx->mask_normalize = normalize != 0;
  ctx->mask_param.assign(mask, mask + n_mask);
  ctx->window_param.assign(window, window + n_window);
  ctx->mask_version++;
  ctx->prepared = false;
  return SFXG_OK;
}

static int sb_sign(char c) { return c == 'L' ? -1 : 1; }

int sfxg_set_parameters(sfxg_ctx *ctx, const sfxg_correlation_parameters *params,
                        const sfxg_delay_model *models, const double *uvw, const sfxg_pulsar *pulsar) {
  if (!ctx || !params || !models) return SFXG_ERR_ARG;
  sync_slice(ctx);
  ctx->enqueued = false;
  ctx->have_params = false;
  ctx->have_results = false;
  ctx->prepared = false;
  const sfxg_correlation_parameters &p = *params;
  if (p.n_streams <= 0 || !p.station_streams) return ctx->fail(SFXG_ERR_ARG, "no station streams");
  if (p.n_phase_centers < 1 || p.n_phase_centers > 256)
    return ctx->fail(SFXG_ERR_ARG, "n_phase_centers %d outside 1 .. 256", p.n_phase_centers);
  if (p.n_phase_centers > 1 && pulsar)  // control_parameters.cc:156-160
    return ctx->fail(SFXG_ERR_ARG, "pulsar binning cannot be used together with multiple phase centres");
  if (p.n_phase_centers > 1 && ctx->opt.phased_array)
    return ctx->fail(SFXG_ERR_ARG, "the phased-array core has one phase centre");
  if (p.fft_size_delaycor != 256) return ctx->fail(SFXG_ERR_ARG, "fft_size_delaycor must be 256");
  const int nc = p.fft_size_correlation;
  if (nc < 256 || nc > 32768 || (nc & (nc - 1)))
    return ctx->fail(SFXG_ERR_ARG, "fft_size_correlation must be a power of two in [256, 32768]");
  if (p.window < SFXG_WINDOW_NONE || p.window > SFXG_WINDOW_HANN)
    return ctx->fail(SFXG_ERR_ARG, "unknown window %d", p.window);
  const bool phased = ctx->opt.phased_array != 0;
  if (phased && pulsar)
    return ctx->fail(SFXG_ERR_ARG, "pulsar binning cannot be used in phased array mode");
  // Correlation_core_phased builds no lag window or mask (correlation_core_phased.cc:53-76)
  if (phased && p.number_channels != nc)
    return ctx->fail(SFXG_ERR_ARG, "phased array mode takes number_channels == fft_size_correlation");
  // integration_write re-channelises in the lag domain when nchan < N_c
  // (correlation_core.cc:408-440)
  if (p.number_channels != nc && !rechannel_supported(nc, p.number_channels))
    return ctx->fail(SFXG_ERR_ARG, "number_channels %d with fft_size_correlation %d: the lag re-channelisation "
                                   "takes a power of two number_channels below fft_size_correlation <= 16384",
                     p.number_channels, nc);
  for (int i = 0; i < p.n_streams; i++) {
    const sfxg_station_parameters &s = p.station_streams[i];
    if (s.sample_rate != p.sample_rate)
      return ctx->fail(SFXG_ERR_ARG, "stream %d: station sample rate differs from the job's", i);
    if (s.bits_per_sample != 1 && s.bits_per_sample != 2)
      return ctx->fail(SFXG_ERR_ARG, "stream %d: bits_per_sample must be 1 or 2", i);
    if (s.bandwidth <= 0) return ctx->fail(SFXG_ERR_ARG, "stream %d: bandwidth", i);
    // output offset must be zero (delay_correction.cc:226-241)
    int64_t freq = s.channel_freq;
    if (p.sideband != s.sideband) freq += sb_sign(s.sideband) * (int64_t)s.bandwidth;
    const int64_t delta = p.sideband == 'L' ? freq - p.channel_freq : p.channel_freq - freq;
    if (delta != 0) return ctx->fail(SFXG_ERR_ARG, "stream %d: frequency offset to the job band", i);
    if (models[i].n_pieces <= 0 || !models[i].delay)
      return ctx->fail(SFXG_ERR_ARG, "stream %d: empty delay model", i);
  }
  ctx->p = p;
  ctx->st.assign(p.station_streams, p.station_streams + p.n_streams);
  ctx->p.station_streams = ctx->st.data();
  ctx->models.assign(p.n_streams, ModelCopy());
  for (int i = 0; i < p.n_streams; i++) {
    ModelCopy &m = ctx->models[i];
    m.scan_begin = models[i].scan_begin;
    m.interval = models[i].interval;
    m.n = models[i].n_pieces;
    m.delay.assign(models[i].delay, models[i].delay + 4 * m.n);
    if (models[i].phase) m.phase.assign(models[i].phase, models[i].phase + 4 * m.n);
    if (models[i].amplitude) m.amp.assign(models[i].amplitude, models[i].amplitude + 4 * m.n);
  }
  ctx->uvw.assign(3 * p.n_streams, 0.0);
  if (uvw) std::copy(uvw, uvw + 3 * p.n_streams, ctx->uvw.begin());
  ctx->pulsar = pulsar != nullptr;
  if (pulsar) {
    if (pulsar->nbins <= 0 || pulsar->n_coef < 2 || !pulsar->coef)
      return ctx->fail(SFXG_ERR_ARG, "invalid pulsar parameters");
    // the MFMA pulsar X-step keeps bin indices in 16 bits (bin 0 off-gate)
    if (pulsar->nbins + 1 > 32767)
      return ctx->fail(SFXG_ERR_ARG, "pulsar with %d bins: at most 32766 on-pulse bins", pulsar->nbins);
    ctx->psr = *pulsar;
    ctx->psr_coef.assign(pulsar->coef, pulsar->coef + pulsar->n_coef);
    ctx->psr.coef = ctx->psr_coef.data();
  }

  // derived sizes
  ctx->M = p.n_streams;
  ctx->nd = p.fft_size_delaycor;
  ctx->nc = nc;
  ctx->nchan = p.number_channels;
  // one record per pulsar bin or per phase centre
  ctx->n_pc = p.n_phase_centers;
  ctx->have_pc = ctx->n_pc == 1;
  ctx->nbins = pulsar ? pulsar->nbins + 1 : ctx->n_pc;
  const int nr = nr_ffts_per_integration_slice(p.integration_time, p.sample_rate, nc);
  ctx->n_ffts_per_integration = (int64_t)(nc / ctx->nd) * nr;  // delay_correction.cc:244-249
  ctx->n_blocks = ctx->n_ffts_per_integration;
  ctx->n_samples = ctx->n_blocks * ctx->nd;
  ctx->n_integration = p.window != SFXG_WINDOW_NONE ? nr - 1 : nr;  // correlation_core.cc:135-142
  ctx->n_sub = nr_ffts_per_integration_slice(p.sub_integration_time, p.sample_rate, nc);
  ctx->nfft_max = std::max(kCorrelatorBufferSize / ctx->nd, std::max(nc / ctx->nd, 1));
  ctx->nseg = ctx->n_integration;
  ctx->S = ((int64_t)nc + 1 + 63) / 64 * 64;
  ctx->total_samples = (int64_t)ctx->n_integration * nc;
  if (ctx->n_integration <= 0) return ctx->fail(SFXG_ERR_ARG, "integration shorter than two FFTs");

  // baselines (correlation_core.cc:149-170); the phased core's rows are one
  // per sub-integration plus one, each written as stream 0's auto-correlation
  // (correlation_core_phased.cc:84-98)
  ctx->phased = phased;
  ctx->bl_first.clear();
  ctx->bl_second.clear();
  if (phased) {
    if (ctx->n_sub <= 0) return ctx->fail(SFXG_ERR_ARG, "sub-integration shorter than one FFT");
    const int rows = ctx->n_integration / ctx->n_sub + 1;
    ctx->bl_first.assign(rows, 0);
    ctx->bl_second.assign(rows, 0);
  }
  for (int i = 0; i < ctx->M && !phased; i++) {
    ctx->bl_first.push_back(i);
    ctx->bl_second.push_back(i);
  }
  for (int i = 0; i < ctx->M && !phased; i++) {
    for (int j = i + 1; j < ctx->M; j++) {
      const int s1 = ctx->st[i].station_number, s2 = ctx->st[j].station_number;
      if (s1 == s2) continue;
      if (p.reference_station >= 0 && p.reference_station != s1 && p.reference_station != s2) continue;
      if (s1 < s2) {
        ctx->bl_first.push_back(i);
        ctx->bl_second.push_back(j);
      } else {
        ctx->bl_first.push_back(j);
        ctx->bl_second.push_back(i);
      }
    }
  }
  ctx->B = (int)ctx->bl_first.size();
  xstep_tile_table(ctx->M, ctx->bl_first, ctx->bl_second, ctx->h_tiles);
  ctx->plans.assign(ctx->M, StreamPlan());
  ctx->have_params = true;
  return SFXG_OK;
}

static int push_common(sfxg_ctx *ctx, int32_t stream_idx, const uint8_t *bytes, size_t n, bool elided,
                       const void *dev_payload, size_t payload_bytes) {
  if (!ctx) return SFXG_ERR_ARG;
  if (!ctx->have_params) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_parameters first");
  if (stream_idx < 0 || stream_idx >= ctx->M) return ctx->fail(SFXG_ERR_ARG, "stream index %d", stream_idx);
  if (!bytes && n) return ctx->fail(SFXG_ERR_ARG, "null stream");
  sync_slice(ctx);
  ctx->enqueued = false;
  ctx->have_results = false;
  ctx->prepared = false;
  const sfxg_station_parameters &s = ctx->st[stream_idx];
  B2fParams bp;
  bp.bits_per_sample = s.bits_per_sample;
  bp.sample_rate = s.sample_rate;
  bp.base_sample_rate = ctx->p.sample_rate;
  bp.tsys_freq = s.tsys_freq;
  bp.fft_size = ctx->nd;
  bp.nfft_max = (int)(((int64_t)ctx->nfft_max * s.sample_rate) / ctx->p.sample_rate);
  bp.n_ffts_per_integration = (int)ctx->n_ffts_per_integration;
  {
    // Bit2float_worker::set_new_parameters: delay(start_time) * sample_rate into an int
    const ModelCopy &m = ctx->models[stream_idx];
    const double sec = (double)(ctx->p.start_time - m.scan_begin) / kClockRate;
    int64_t i = (int64_t)floor(sec / m.interval);
    if (i < 0) i = 0;
    if (i > m.n - 1) i = m.n - 1;
    const double u = sec - (double)i * m.interval;
    const double *c = &m.delay[4 * i];
    const double d = c[0] + u * (c[1] + u * (c[2] + u * c[3]));
    bp.delay_in_samples = (int)(d * s.sample_rate);
  }
  StreamPlan &plan = ctx->plans[stream_idx];
  std::string e;
  int rc = parse_header_stream(bytes, n, elided, bp, plan, e);
  if (rc != SFXG_OK) return ctx->fail(rc, "stream %d: %s", stream_idx, e.c_str());
  if (elided) {
    if (!dev_payload && plan.payload_bytes > 0) return ctx->fail(SFXG_ERR_ARG, "null device payload");
    if ((int64_t)payload_bytes < plan.payload_bytes)
      return ctx->fail(SFXG_ERR_ARG, "stream %d: device payload holds %zu bytes, the records announce %lld",
                       stream_idx, payload_bytes, (long long)plan.payload_bytes);
    plan.ext_payload = dev_payload;
  }
  return SFXG_OK;
}

int sfxg_push_stream(sfxg_ctx *ctx, int32_t stream_idx, const uint8_t *bytes, size_t n) {
  return push_common(ctx, stream_idx, bytes, n, false, nullptr, 0);
}

int sfxg_push_stream_elided(sfxg_ctx *ctx, int32_t stream_idx, const uint8_t *control, size_t n,
                            const void *device_payload, size_t payload_bytes) {
  return push_common(ctx, stream_idx, control, n, true, device_payload, payload_bytes);
}

// Input_node_tasklet::get_delays (input_node_tasklet.cc:235-262) for a
// monotonic delay model: the slice start's integer delay and the first
// sample instant of every later integer level, each found by bisection
// over the slice's samples.
static std::vector<std::pair<int64_t, int64_t>> input_node_delays(const ModelCopy &m, int64_t start,
                                                                  int64_t nsamples, int fs) {
  const int64_t tps = (int64_t)round(kClockRate / fs);
  auto n_at = [&](int64_t t) { return (int64_t)floor(model_eval(m, t, false) * fs + 0.5); };
  std::vector<std::pair<int64_t, int64_t>> out{{start, n_at(start)}};
  const int64_t n0 = out[0].second, n1 = n_at(start + nsamples * tps);
  if (n0 == n1) return out;
  const int64_t step = n1 > n0 ? 1 : -1;
  for (int64_t level = n0 + step; level != n1 + step; level += step) {
    int64_t lo = 0, hi = nsamples;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) / 2;
      if ((n_at(start + mid * tps) - level) * step >= 0) hi = mid;
      else lo = mid;
    }
    out.push_back({start + hi * tps, level});
  }
  return out;
}

// The frames' header words lifted on the device, then VDIF_reader's walk
// on the host: channel v->channel's frame blocks.
static int vdif_channel_blocks(sfxg_ctx *ctx, int stream_idx, const void *device_frames, int64_t n_frames,
                               const sfxg_vdif_stream *v, std::vector<VdifBlock> &blocks, int &frame_payload,
                               uint32_t *dev_words = nullptr) {
  std::vector<uint32_t> words((size_t)n_frames * 4);
  if (!dev_words) {
    HIPCHK(ctx, ctx->vdif_words.ensure(words.size() * sizeof(uint32_t)));
    dev_words = ctx->vdif_words.as<uint32_t>();
  }
  HIPCHK(ctx, launch_vdif_words(device_frames, n_frames, v->frame_bytes, dev_words, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(words.data(), dev_words, words.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int> threads(v->thread_ids, v->thread_ids + v->n_threads);
  std::string err;
  const int rc = vdif_blocks(words.data(), n_frames, v->frame_bytes, ctx->st[stream_idx].sample_rate, v->ref_mjd,
                             v->offset, threads, v->channel, blocks, frame_payload, err);
  if (rc != SFXG_OK) return ctx->fail(rc, "stream %d: %s", stream_idx, err.c_str());
  return SFXG_OK;
}

static int vdif_check(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                      const sfxg_vdif_stream *v) {
  if (!ctx->have_params) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_parameters first");
  if (stream_idx < 0 || stream_idx >= ctx->M) return ctx->fail(SFXG_ERR_ARG, "stream index %d", stream_idx);
  if (v->frame_bytes < 32 || v->frame_bytes % 8) return ctx->fail(SFXG_ERR_ARG, "frame_bytes %d", v->frame_bytes);
  if (v->n_threads < 0 || (v->n_threads > 0 && !v->thread_ids) || v->channel < 0 ||
      (v->n_threads > 0 && v->channel >= v->n_threads) || (v->n_threads == 0 && v->channel != 0))
    return ctx->fail(SFXG_ERR_ARG, "channel %d of %d threads", v->channel, v->n_threads);
  if ((int64_t)(n_bytes / (size_t)v->frame_bytes) <= 0) return ctx->fail(SFXG_ERR_ARG, "no whole VDIF frame");
  (void)device_frames;
  return SFXG_OK;
}

// do_phasecal over a stream's frames. scratch == nullptr: the device tables
// and sums go to the context's word table and to buffers allocated for the
// call; otherwise they go to `scratch` (scratch_bytes long), and nothing is
// allocated on the device.
static int vdif_phasecal_impl(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                              const sfxg_vdif_stream *v, int64_t integration_ticks, uint8_t *dst, size_t cap,
                              size_t *written, uint8_t *scratch, size_t scratch_bytes) {
  if (!ctx || !device_frames || !v || (!dst && cap)) return SFXG_ERR_ARG;
  int rc = vdif_check(ctx, stream_idx, device_frames, n_bytes, v);
  if (rc != SFXG_OK) return rc;
  if (integration_ticks <= 0) return ctx->fail(SFXG_ERR_ARG, "phase-cal integration time %lld", (long long)integration_ticks);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  const sfxg_station_parameters &st = ctx->st[stream_idx];
  const int64_t n_frames = (int64_t)(n_bytes / (size_t)v->frame_bytes);
  std::vector<VdifBlock> blocks;
  int frame_payload = 0;
  if (scratch && (size_t)n_frames * 16 > scratch_bytes)
    return ctx->fail(SFXG_ERR_ARG, "%lld frames: their header words need %lld bytes, the static scratch holds %zu",
                     (long long)n_frames, (long long)n_frames * 16, scratch_bytes);
  rc = vdif_channel_blocks(ctx, stream_idx, device_frames, n_frames, v, blocks, frame_payload,
                           reinterpret_cast<uint32_t *>(scratch));
  if (rc != SFXG_OK) return rc;
  // do_phasecal's walk over the blocks (input_node_data_writer.cc:232-306):
  // an integration starts at a block whose start time is a multiple of the
  // integration time; the fold position restarts there and advances over
  // invalid blocks without adding
  const int P = (int)(st.sample_rate / 10e3);
  const int spb = 8 / st.bits_per_sample;
  const int S = frame_payload * spb;
  if (P <= 0 || P % spb) return ctx->fail(SFXG_ERR_ARG, "phase-cal period of %d samples", P);
  std::vector<PcalFrame> table;
  std::vector<std::pair<int64_t, int64_t>> closed;  // (start, duration) of the finished integrations
  int integ = -1;
  int64_t count = 0, pc_time = 0;
  for (const VdifBlock &b : blocks) {
    if (b.time % integration_ticks == 0) {
      if (integ >= 0) closed.push_back({pc_time, b.time - pc_time});
      integ++;
      count = 0;
      pc_time = b.time;
    }
    if (integ < 0) continue;
    if (!b.invalid) table.push_back({b.payload_offset, (int32_t)(count % P), integ});
    count = count % P + S;
  }
  const size_t rec = 4 + 8 + 8 + 4 + 4 * (size_t)P;
  const size_t need = rec * closed.size();
  if (written) *written = need;
  if (cap < need) return ctx->fail(SFXG_ERR_ARG, "phase-cal buffer holds %zu bytes, %zu needed", cap, need);
  if (closed.empty()) return SFXG_OK;
  const size_t n_sums = (size_t)(integ + 1) * P;
  const size_t tab_bytes = std::max<size_t>(table.size(), 1) * sizeof(PcalFrame);
  const size_t sums_at = (tab_bytes + 255) / 256 * 256;  // the header words are no longer needed
  DBuf sums, tab;
  void *tab_p, *sums_p;
  hipError_t e = hipSuccess;
  if (scratch) {
    if (sums_at + n_sums * sizeof(int32_t) > scratch_bytes)
      return ctx->fail(SFXG_ERR_ARG, "%zu frames in %zu integrations of %d bins need %zu bytes, the static scratch "
                       "holds %zu", table.size(), (size_t)integ + 1, P, sums_at + n_sums * sizeof(int32_t),
                       scratch_bytes);
    tab_p = scratch;
    sums_p = scratch + sums_at;
  } else {
    e = sums.ensure(n_sums * sizeof(int32_t));
    if (e == hipSuccess) e = tab.ensure(tab_bytes);
    tab_p = tab.p;
    sums_p = sums.p;
  }
  if (e == hipSuccess) e = launch_fill_zero(sums_p, n_sums * sizeof(int32_t), ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tab_p, table.data(), table.size() * sizeof(PcalFrame), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = launch_phasecal(device_frames, static_cast<const PcalFrame *>(tab_p), (int)table.size(), S, P,
                        st.bits_per_sample, static_cast<int32_t *>(sums_p), ctx->stream);
  std::vector<int32_t> h(n_sums);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), sums_p, n_sums * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  sums.release();
  tab.release();
  HIPCHK(ctx, e);
  // the MPI_Pack layout of the messages: station, frequency number,
  // sideband, polarisation (uint8), start and duration (int64 ticks),
  // sample count (int32), the sums (int32)
  uint8_t *o = dst;
  for (size_t i = 0; i < closed.size(); i++) {
    const uint8_t hd[4] = {(uint8_t)st.station_number, (uint8_t)ctx->p.frequency_nr,
                           (uint8_t)(st.sideband == 'U' ? 1 : 0), (uint8_t)(st.polarisation == 'L' ? 1 : 0)};
    memcpy(o, hd, 4);
    memcpy(o + 4, &closed[i].first, 8);
    memcpy(o + 12, &closed[i].second, 8);
    const int32_t n = P;
    memcpy(o + 20, &n, 4);
    memcpy(o + 24, &h[i * P], 4 * (size_t)P);
    o += rec;
  }
  return SFXG_OK;
}

int sfxg_vdif_phasecal(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                       const sfxg_vdif_stream *v, int64_t integration_ticks, uint8_t *dst, size_t cap,
                       size_t *written) {
  return vdif_phasecal_impl(ctx, stream_idx, device_frames, n_bytes, v, integration_ticks, dst, cap, written,
                            nullptr, 0);
}

}  // extern "C"

// Device scratch of sfxg_vdif_phasecal_static: one copy per GPU, held until
// the process ends. Host code never passes `scratch` itself to a kernel or
// a copy (in host code the name is a host-side stand-in): it uses the device
// address hipGetSymbolAddress gives for the device that is current.
static __device__ __align__(256) uint8_t scratch[SFXG_PHASECAL_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};

// The device address of this device's `scratch`, looked up once per device
// with that device current (the caller has made it current).
int scratch_of(sfxg_ctx *ctx, uint8_t **out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return ctx->fail(SFXG_ERR_ARG, "device %d", ctx->device);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[ctx->device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return ctx->fail(SFXG_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    scratch_ptr[ctx->device] = ptr;
  }
  *out = static_cast<uint8_t *>(scratch_ptr[ctx->device]);
  return SFXG_OK;
}
}  // namespace

extern "C" {

int sfxg_vdif_phasecal_static(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                              const sfxg_vdif_stream *v, int64_t integration_ticks, uint8_t *dst, size_t cap,
                              size_t *written) {
  if (!ctx || !device_frames || !v || (!dst && cap)) return SFXG_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  uint8_t *buf = nullptr;
  const int rc = scratch_of(ctx, &buf);
  if (rc != SFXG_OK) return rc;
  return vdif_phasecal_impl(ctx, stream_idx, device_frames, n_bytes, v, integration_ticks, dst, cap, written, buf,
                            SFXG_PHASECAL_SCRATCH_BYTES);
}

int sfxg_push_vdif(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                   const sfxg_vdif_stream *v) {
  if (!ctx || !device_frames || !v) return SFXG_ERR_ARG;
  const int vrc = vdif_check(ctx, stream_idx, device_frames, n_bytes, v);
  if (vrc != SFXG_OK) return vrc;
  const int64_t n_frames = (int64_t)(n_bytes / (size_t)v->frame_bytes);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  const sfxg_station_parameters &st = ctx->st[stream_idx];
  std::vector<VdifBlock> blocks;
  int frame_payload = 0;
  std::string err;
  int rc = vdif_channel_blocks(ctx, stream_idx, device_frames, n_frames, v, blocks, frame_payload);
  if (rc != SFXG_OK) return rc;
  const int64_t nsamples = (int64_t)ctx->n_ffts_per_integration * ctx->nd;
  const auto delays = input_node_delays(ctx->models[stream_idx], ctx->p.start_time, nsamples, st.sample_rate);
  std::vector<uint8_t> control;
  std::vector<VdifPiece> pieces;
  rc = vdif_write_slice(blocks, frame_payload, st.bits_per_sample, st.sample_rate, ctx->p.start_time, nsamples,
                        delays, control, pieces, err);
  if (rc != SFXG_OK) return ctx->fail(rc, "stream %d: %s", stream_idx, err.c_str());
  // gather the DATA payload on the device
  std::vector<int64_t> dst_off(pieces.size());
  int64_t total = 0;
  for (size_t k = 0; k < pieces.size(); k++) {
    if (pieces[k].src < 0 || pieces[k].src + pieces[k].n > (int64_t)n_bytes)
      return ctx->fail(SFXG_ERR_STREAM, "stream %d: payload outside the frames", stream_idx);
    dst_off[k] = total;
    total += pieces[k].n;
  }
  if (ctx->vdif_payload.size() < (size_t)ctx->M) ctx->vdif_payload.resize(ctx->M);
  DBuf &pay = ctx->vdif_payload[stream_idx];
  HIPCHK(ctx, pay.ensure(std::max<int64_t>(total, 1)));
  HIPCHK(ctx, ctx->vdif_pieces.ensure(std::max<size_t>(pieces.size(), 1) * sizeof(VdifPiece)));
  HIPCHK(ctx, ctx->vdif_dst.ensure(std::max<size_t>(dst_off.size(), 1) * sizeof(int64_t)));
  HIPCHK(ctx, hipMemcpyAsync(ctx->vdif_pieces.p, pieces.data(), pieces.size() * sizeof(VdifPiece),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->vdif_dst.p, dst_off.data(), dst_off.size() * sizeof(int64_t),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, launch_vdif_gather(device_frames, ctx->vdif_pieces.as<VdifPiece>(), ctx->vdif_dst.as<int64_t>(),
                                 (int)pieces.size(), pay.p, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return push_common(ctx, stream_idx, control.data(), control.size(), true, pay.p, (size_t)total);
}

static void make_tables(sfxg_ctx *ctx, std::vector<float2> &tnd, std::vector<float2> &tn,
                        std::vector<float2> &t2n, std::vector<float2> &tsp, std::vector<float> &win) {
  const int nd = ctx->nd, nc = ctx->nc;
  tnd.resize(nd);
  for (int k = 0; k < nd; k++) {
    const double a = -2.0 * M_PI * k / nd;
    tnd[k] = make_float2((float)cos(a), (float)sin(a));
  }
  tn.resize(nc);
  for (int k = 0; k < nc; k++) {
    const double a = -2.0 * M_PI * k / nc;
    tn[k] = make_float2((float)cos(a), (float)sin(a));
  }
  t2n.resize(nc + 1);
  for (int k = 0; k <= nc; k++) {
    const double a = -M_PI * k / nc;
    t2n[k] = make_float2((float)cos(a), (float)sin(a));
  }
  // exp(-2 pi i j / 2N) = coarse[j >> 7] * fine[j & 127]
  const int ncoarse = 2 * nc / 128;
  tsp.resize(ncoarse + 128);
  for (int j = 0; j < ncoarse; j++) {
    const double a = -M_PI * (128.0 * j) / nc;
    tsp[j] = make_float2((float)cos(a), (float)sin(a));
  }
  for (int j = 0; j < 128; j++) {
    const double a = -M_PI * j / nc;
    tsp[ncoarse + j] = make_float2((float)cos(a), (float)sin(a));
  }
  // create_window (delay_correction.cc:299-341), n = fft_rot_size = 2 N_c
  const int n = 2 * nc;
  win.resize(n);
  for (int i = 0; i < n; i++) {
    double v;
    switch (ctx->p.window) {
      case SFXG_WINDOW_NONE: v = i < n / 2 ? 1 : 0; break;
      case SFXG_WINDOW_RECT: v = (i >= n / 4 && i < 3 * n / 4) ? 1 : 0; break;
      case SFXG_WINDOW_COS: v = sin(M_PI * i / (n - 1)); break;
      case SFXG_WINDOW_HAMMING: v = 0.54 - 0.46 * cos(2 * M_PI * i / (n - 1)); break;
      default: v = 0.5 * (1 - cos(2 * M_PI * i / (n - 1))); break;
    }
    win[i] = (float)v;
  }
}

}  // extern "C"

// Shape of a slice as the tuning table sees it.
static std::string tune_key(const sfxg_ctx *ctx, bool fused) {
  char buf[160];
  snprintf(buf, sizeof(buf), "nc=%d M=%d prec=%d pulsar=%d fused=%d", ctx->nc, ctx->M, ctx->opt.precision,
           ctx->pulsar ? 1 : 0, fused ? 1 : 0);
  return buf;
}

// Host part of sfxg_correlate: chunk and invalid-data bookkeeping, device
// buffers, tables and the slice's uploads (asynchronous, on the context's
// stream), and the launch parameters of the device work (ctx->lp).
// create_weights (correlation_core.cc:712-759) at one lag: the
// autocorrelation of the window f(2N_c - 1, i), i < 2N_c, as stored in FLOAT,
// over its value at lag 0. The reference takes the whole autocorrelation
// through FP32 transforms; here the lags a slice needs are summed in FP64.
static float pc_weight(sfxg_ctx *ctx, int lag) {
  const int n = ctx->nc, w = ctx->p.window;
  if (ctx->pcw_window != w || ctx->pcw_nc != n) {
    ctx->pcw_win.assign(2 * (size_t)n, 0.0f);
    const int L = 2 * n - 1;
    for (int i = 0; i < 2 * n; i++) {
      double v;
      if (w == SFXG_WINDOW_COS) v = sin(M_PI * i / L);
      else if (w == SFXG_WINDOW_HAMMING) v = 0.54 - 0.46 * cos(2 * M_PI * i / L);
      else if (w == SFXG_WINDOW_HANN) v = 0.5 * (1 - cos(2 * M_PI * i / L));
      else v = (i >= L / 4 && i < (3 * L) / 4) ? 1.0 : 0.0;  // NONE, RECT
      ctx->pcw_win[i] = (float)v;
    }
    ctx->pcw_r.assign(n, std::nan(""));
    ctx->pcw_window = w;
    ctx->pcw_nc = n;
  }
  auto r = [ctx, n](int l) {
    if (std::isnan(ctx->pcw_r[l])) {
      double sum = 0.0;
      for (int j = 0; j + l < 2 * n; j++) sum += (double)ctx->pcw_win[j] * ctx->pcw_win[j + l];
      ctx->pcw_r[l] = sum;
    }
    return ctx->pcw_r[l];
  };
  return (float)(r(lag) / r(0));
}

// The uv shift of every sub-integration, extra phase centre and cross
// baseline (Correlation_core::sub_integration and uvshift,
// correlation_core.cc:515-592) into ctx->h_uv, [fold][centre - 1][baseline].
static void uv_table(sfxg_ctx *ctx) {
  const int M = ctx->M, B = ctx->B, npc = ctx->n_pc, nc = ctx->nc;
  const int nf = (int)ctx->folds.size();
  ctx->h_uv.assign((size_t)nf * (npc - 1) * B, UvShift{0.0, 0.0, 1.0f, 0});
  // Time tfft(0., sample_rate); tfft.inc_samples(fft_size())
  const int64_t tfft = (int64_t)nc * (int64_t)(int)round(kClockRate / ctx->p.sample_rate);
  const int sb = ctx->p.sideband == 'L' ? -1 : 1;
  const double base_freq = (double)ctx->p.channel_freq;
  const double dfreq = ctx->p.sample_rate / (2. * nc);
  int64_t previous_fft = 0;
  for (int f = 0; f < nf; f++) {
    const int64_t current_fft = ctx->folds[f];
    const int64_t tmid =
        ctx->p.start_time + (int64_t)round(tfft * (previous_fft + (current_fft - previous_fft) / 2.));
    for (int b = M; b < B; b++) {
      const int s1 = ctx->bl_first[b], s2 = ctx->bl_second[b];
      const double delay1 = model_eval(ctx->models[s1], tmid, false);
      const double delay2 = model_eval(ctx->models[s2], tmid, false);
      const double rate1 = model_eval(ctx->models[s1], tmid, true);
      const double rate2 = model_eval(ctx->models[s2], tmid, true);
      for (int j = 1; j < npc; j++) {
        const double ddelay1 = model_eval(ctx->pc_models[(size_t)(j - 1) * M + s1], tmid, false) - delay1;
        const double ddelay2 = model_eval(ctx->pc_models[(size_t)(j - 1) * M + s2], tmid, false) - delay2;
        UvShift &u = ctx->h_uv[((size_t)f * (npc - 1) + (j - 1)) * B + b];
        const int lag = abs((int)round((ddelay1 - ddelay2) * ctx->p.sample_rate));
        u.amp = 1.0f;
        if (lag < nc) {
          const float w = pc_weight(ctx, lag);
          if (w > 1e-4) u.amp = (float)(1. / w);
        }
        const double x = ddelay1 * (1 - rate1) - ddelay2 * (1 - rate2);
        const double phi = base_freq * x;
        u.phi = 2 * M_PI * sb * (phi - floor(phi));
        u.delta = 2 * M_PI * dfreq * x;
      }
    }
    previous_fft = current_fft;
  }
}

static int prepare_slice(sfxg_ctx *ctx) {
  if (!ctx->have_params) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_parameters first");
  if ((ctx->opt.phased_array != 0) != ctx->phased)
    return ctx->fail(SFXG_ERR_STATE, "phased_array changed after sfxg_set_parameters: set the parameters again");
  if (!ctx->have_pc) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_phase_centers first (%d phase centres)", ctx->n_pc);
  if (ctx->n_pc > 1 && ctx->opt.precision != SFXG_PREC_FP32)
    return ctx->fail(SFXG_ERR_ARG, "multiple phase centres take SFXG_PREC_FP32");
  for (int i = 0; i < ctx->M; i++)
    if (!ctx->plans[i].ready) return ctx->fail(SFXG_ERR_STATE, "stream %d has not been pushed", i);
  for (int i = 1; i < ctx->M; i++)
    if (ctx->plans[i].chunks != ctx->plans[0].chunks)
      return ctx->fail(SFXG_ERR_STATE, "streams deliver different chunkings");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  const int M = ctx->M, nd = ctx->nd, nc = ctx->nc, B = ctx->B;
  hipStream_t s = ctx->stream;

  // chunk bookkeeping (delay_correction.cc:24-36, correlation_core.cc:42-72)
  const bool windowed = ctx->p.window != SFXG_WINDOW_NONE;
  ctx->seg_per_chunk = segments_per_chunk(ctx->plans[0].chunks, nd, 2LL * nc, windowed);
  int64_t total_seg = 0;
  for (int32_t v : ctx->seg_per_chunk) total_seg += v;
  if (total_seg != ctx->n_integration)
    return ctx->fail(SFXG_ERR_STATE, "segment count %lld != %d", (long long)total_seg, ctx->n_integration);
  ctx->folds = fold_points(ctx->seg_per_chunk, ctx->n_integration, ctx->n_sub);

  // invalid data bookkeeping (find_invalid :599-665, integration_normalize :295-313)
  std::vector<int32_t> levels4(M);
  for (int i = 0; i < M; i++)
    levels4[i] = (int32_t)(((int64_t)ctx->p.sample_rate * ctx->plans[i].n_invalid) / ctx->st[i].sample_rate);
  ctx->flag_factor.assign(B, 1.0);
  ctx->n_flagged_first.assign(B, 0);
  for (int b = M; b < B; b++) {
    const int a = ctx->bl_first[b], c = ctx->bl_second[b];
    int64_t nf[2] = {0, 0};
    find_invalid(ctx->plans[a].invalid, ctx->plans[c].invalid,
                 ctx->st[a].sample_rate * 1. / ctx->p.sample_rate, ctx->st[c].sample_rate * 1. / ctx->p.sample_rate,
                 nf);
    const int64_t nv1 = ctx->total_samples - levels4[a], nv2 = ctx->total_samples - levels4[c];
    const double N1 = nv1 > 0 ? 1 - nf[0] * 1. / nv1 : 1;
    const double N2 = nv2 > 0 ? 1 - nf[1] * 1. / nv2 : 1;
    double N = N1 * N2;
    if (N < 0.01) N = 1;
    ctx->flag_factor[b] = N;
    ctx->n_flagged_first[b] = nf[0];
  }

  // upload tables
  if (ctx->tables_nc != nc || ctx->tables_window != ctx->p.window) {
    std::vector<float2> tnd, tn, t2n, tsp;
    std::vector<float> win;
    make_tables(ctx, tnd, tn, t2n, tsp, win);
    HIPCHK(ctx, ctx->tw_nd.ensure(tnd.size() * sizeof(float2)));
    HIPCHK(ctx, ctx->tw_n.ensure(tn.size() * sizeof(float2)));
    HIPCHK(ctx, ctx->tw_2n.ensure(t2n.size() * sizeof(float2)));
    HIPCHK(ctx, ctx->tw_split.ensure(tsp.size() * sizeof(float2)));
    HIPCHK(ctx, ctx->window.ensure(win.size() * sizeof(float)));
    HIPCHK(ctx, hipMemcpy(ctx->tw_nd.p, tnd.data(), tnd.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->tw_n.p, tn.data(), tn.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->tw_2n.p, t2n.data(), t2n.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->tw_split.p, tsp.data(), tsp.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->window.p, win.data(), win.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->tables_nc = nc;
    ctx->tables_window = ctx->p.window;
    ctx->window_sumsq = 0.0;
    for (float w : win) ctx->window_sumsq += (double)w * w;
  }

  // lag re-channelisation tables (create_mask / create_window,
  // correlation_core.cc:761-824), rebuilt when the shape, the window or the
  // mask parameters change
  if (ctx->nchan != nc && (ctx->rc_nc != nc || ctx->rc_nchan != ctx->nchan || ctx->rc_window != ctx->p.window ||
                           ctx->rc_version != ctx->mask_version)) {
    std::vector<float> mask, win;
    if (!ctx->mask_param.empty()) {
      if ((int)ctx->mask_param.size() != nc + 1)
        return ctx->fail(SFXG_ERR_ARG, "the mask holds %zu values, fft_size_correlation + 1 = %d needed",
                         ctx->mask_param.size(), nc + 1);
      mask.assign(ctx->mask_param.begin(), ctx->mask_param.end());
    } else {
      mask.assign(nc + 1, 1.0f);
    }
    if (!ctx->window_param.empty()) {
      if ((int)ctx->window_param.size() != 2 * ctx->nchan)
        return ctx->fail(SFXG_ERR_ARG, "the lag window holds %zu values, 2 * number_channels = %d needed",
                         ctx->window_param.size(), 2 * ctx->nchan);
      win.assign(ctx->window_param.begin(), ctx->window_param.end());
    } else if (!lag_window(ctx->p.window, ctx->nchan, nc, win)) {
      return ctx->fail(SFXG_ERR_ARG, "unknown window %d", ctx->p.window);
    }
    std::vector<float2> tw(2 * (size_t)nc);
    for (int k = 0; k < 2 * nc; k++) {
      const double a = -M_PI * k / nc;
      tw[k] = make_float2((float)cos(a), (float)sin(a));
    }
    HIPCHK(ctx, ctx->rc_tw.ensure(tw.size() * sizeof(float2)));
    HIPCHK(ctx, ctx->rc_mask.ensure(mask.size() * sizeof(float)));
    HIPCHK(ctx, ctx->rc_window_buf.ensure(win.size() * sizeof(float)));
    HIPCHK(ctx, hipMemcpy(ctx->rc_tw.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->rc_mask.p, mask.data(), mask.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->rc_window_buf.p, win.data(), win.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->rc_nc = nc;
    ctx->rc_nchan = ctx->nchan;
    ctx->rc_window = ctx->p.window;
    ctx->rc_version = ctx->mask_version;
  }

  ctx->rc_normalize = ctx->mask_normalize;

  // per-stream tables, concatenated
  std::vector<int64_t> pay_off(M), run_off(M), cnt_off(M), blk_off(M), mod_off(M);
  int64_t pay_tot = 0, run_tot = 0, cnt_tot = 0, mod_tot = 0, max_counted = 0;
  for (int i = 0; i < M; i++) {
    const StreamPlan &pl = ctx->plans[i];
    pay_off[i] = pay_tot;
    if (!pl.ext_payload) pay_tot += ((int64_t)pl.payload.size() + 255) / 256 * 256;
    run_off[i] = run_tot;
    run_tot += (int64_t)pl.runs.size();
    cnt_off[i] = cnt_tot;
    cnt_tot += (int64_t)pl.counted.size();
    if (!pl.counted.empty())
      max_counted = std::max<int64_t>(max_counted, pl.counted.back().ordinal + pl.counted.back().n_bytes);
    blk_off[i] = (int64_t)i * ctx->n_blocks;
    mod_off[i] = mod_tot;
    const ModelCopy &m = ctx->models[i];
    mod_tot += (int64_t)(m.delay.size() + m.phase.size() + m.amp.size());
  }
  HIPCHK(ctx, ctx->payload.ensure((size_t)pay_tot));
  HIPCHK(ctx, ctx->runs.ensure((size_t)run_tot * sizeof(Run)));
  HIPCHK(ctx, ctx->counted.ensure((size_t)std::max<int64_t>(cnt_tot, 1) * sizeof(CountRange)));
  HIPCHK(ctx, ctx->block_run.ensure((size_t)M * ctx->n_blocks * sizeof(int64_t)));
  HIPCHK(ctx, ctx->block_consts.ensure((size_t)M * ctx->n_blocks * sizeof(BlockConst)));
  HIPCHK(ctx, ctx->model_buf.ensure((size_t)mod_tot * sizeof(double)));
  HIPCHK(ctx, ctx->dstreams.ensure((size_t)M * sizeof(DevStream)));
  // N_c <= 4096 without keep_time_series may take the fused F-engine when
  // the tuning table picks it (fe_fused = 1), unless packed FP16 / BF16
  // spectra take the MFMA channeliser; by default the delay stage and the
  // channeliser run as two kernels, which measured faster at 1024 and 4096
  // channels (the table's chan_fp32 = 1 keeps the FP32 transform)
  const bool key_fused = !ctx->opt.keep_time_series && fengine_supported(nc);
  {
    const std::string key = tune_key(ctx, key_fused);
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    ctx->tune = it != g_tune.end() ? it->second : Tune();
  }
  const int packed = ctx->opt.precision == SFXG_PREC_FP16 || ctx->opt.precision == SFXG_PREC_BF16
                         ? ctx->opt.precision : 0;
  const bool mfma_chan = chan_mfma_supported(nc, packed) && ctx->tune.chan_fp32 == 0;
  const bool fused_ok = key_fused && !mfma_chan;
  const bool fused = fused_ok && ctx->tune.fe_fused == 1;
  // the MFMA channeliser rounds its input points to FP16 anyway: the delay
  // stage hands it an FP16 series (half the bytes each way), unless the
  // series is kept for reading back
  const bool ts_half = mfma_chan && !ctx->opt.keep_time_series;
  const size_t ts_bytes = ts_half ? sizeof(uint16_t) : sizeof(float);
  // FP16 and BF16 X-steps take their operands straight from spectra the
  // channeliser stores in that type (4 bytes per bin instead of 8); split
  // FP16 and FP32 need the FP32 spectra
  int spec_prec = ctx->opt.precision == SFXG_PREC_FP16 || ctx->opt.precision == SFXG_PREC_BF16
                      ? ctx->opt.precision : 0;
  const size_t spec_bytes = spec_prec ? 4 : sizeof(float2);  // split FP16 planes take 8, as FP32
  // chunks of whole sub-integrations whose spectra fit the budget (one
  // sub-integration at least)
  LaunchPlan &lp = ctx->lp;
  {
    const int64_t bph = nc / nd;  // delay blocks per segment advance
    const size_t seg_bytes = (size_t)M * ctx->S * spec_bytes;
    const size_t chunk_bytes = (size_t)(ctx->opt.chunk_mib > 0 ? ctx->opt.chunk_mib : 8192) << 20;
    const int64_t budget = std::max<int64_t>(1, (int64_t)(chunk_bytes / seg_bytes));
    lp.chunks.clear();
    int f0 = 0;
    int64_t c0 = 0;
    const int nf = (int)ctx->folds.size();
    for (int f = 0; f < nf; f++) {
      if (ctx->folds[f] - c0 > budget && f > f0) {
        lp.chunks.push_back({c0, ctx->folds[f - 1], 0, 0, f0, f});
        c0 = ctx->folds[f - 1];
        f0 = f;
      }
    }
    lp.chunks.push_back({c0, nf ? (int64_t)ctx->folds[nf - 1] : ctx->nseg, 0, 0, f0, nf});
    lp.chunk_segs = 0;
    lp.ts_stride = 0;
    for (size_t c = 0; c < lp.chunks.size(); c++) {
      Chunk &ch = lp.chunks[c];
      ch.blk0 = ch.seg0 * bph;
      ch.blk1 = c + 1 == lp.chunks.size() ? ctx->n_blocks : std::min<int64_t>(ctx->n_blocks, (ch.seg1 + 1) * bph);
      lp.chunk_segs = std::max<int64_t>(lp.chunk_segs, ch.seg1 - ch.seg0);
      lp.ts_stride = std::max<int64_t>(lp.ts_stride, (ch.blk1 - ch.blk0) * nd);
    }
    while (ctx->cev.size() < 4 * lp.chunks.size()) {
      hipEvent_t e;
      HIPCHK(ctx, hipEventCreate(&e));
      ctx->cev.push_back(e);
    }
  }
  // split FP16 without pulsar binning: the channelisers store the hi and lo
  // operand pairs (8 bytes per bin, as FP32), which the direct-to-LDS X-step
  // reads without converting anything
  if (ctx->opt.precision == SFXG_PREC_SPLIT_FP16 && !ctx->pulsar &&
      xstep_split_planes_fit(M, lp.chunk_segs * ctx->S))
    spec_prec = SFXG_PREC_SPLIT_FP16;
  if (!fused) HIPCHK(ctx, ctx->time_series.ensure((size_t)M * lp.ts_stride * ts_bytes));
  if (channelise_scratch_elems(nc) > 0)
    HIPCHK(ctx, ctx->chan_scratch.ensure(channelise_scratch_elems(nc) * sizeof(float2)));
  HIPCHK(ctx, ctx->spectra.ensure((size_t)M * lp.chunk_segs * ctx->S * spec_bytes));
  const size_t vis_elems = (size_t)ctx->nbins * B * (nc + 1);
  HIPCHK(ctx, ctx->vis.ensure(vis_elems * sizeof(float2)));
  HIPCHK(ctx, ctx->norm_out.ensure(vis_elems * sizeof(float2)));
  if (ctx->nchan != nc) HIPCHK(ctx, ctx->chan_out.ensure((size_t)ctx->nbins * B * (ctx->nchan + 1) * sizeof(float2)));
  HIPCHK(ctx, ctx->hist.ensure((size_t)M * 512 * sizeof(int32_t)));
  HIPCHK(ctx, ctx->norms.ensure((size_t)ctx->nbins * M * sizeof(double)));
  HIPCHK(ctx, ctx->flags.ensure((size_t)B * sizeof(double)));
  HIPCHK(ctx, ctx->bl.ensure((size_t)2 * B * sizeof(int32_t)));
  HIPCHK(ctx, ctx->tiles.ensure(std::max<size_t>(ctx->h_tiles.size(), 1) * sizeof(int32_t)));
  HIPCHK(ctx, ctx->fold_buf.ensure((ctx->folds.size() + 1) * sizeof(int32_t)));
  // sub-integration partial sums, at most ~1 GiB at a time
  const size_t part_one = (size_t)B * (nc + 1) * sizeof(float2);
  const int k_group = (int)std::max<size_t>(1, std::min<size_t>(ctx->folds.size(), ((size_t)1 << 30) / part_one));
  const bool mixed = ctx->opt.precision != SFXG_PREC_FP32;
  if (!ctx->pulsar) HIPCHK(ctx, ctx->partial.ensure(part_one * k_group));
  // mixed precision with pulsar binning: binned Gram tiles for runs of
  // channels, at most ~8 GiB at a time
  int psr_run = 0;
  if (mixed && ctx->pulsar) {
    // all channels in one run up to 8 GiB; longer slices than one chunk need that
    const size_t per_ch = pulsar_gram_floats(M, 1, ctx->nbins) * sizeof(float);
    psr_run = (int)std::max<size_t>(1, std::min<size_t>((size_t)nc + 1, ((size_t)8 << 30) / per_ch));
    if (psr_run < nc + 1 && lp.chunks.size() > 1)
      return ctx->fail(SFXG_ERR_ARG, "pulsar binning with mixed precision: the binned Gram tiles of %d channels "
                                     "exceed 8 GiB for a slice longer than one chunk", nc + 1);
    HIPCHK(ctx, ctx->gram.ensure(per_ch * psr_run));
  }
  // pulsar binning: the bin of every (segment, channel) of a chunk, evaluated
  // once before the X-step; with mixed precision also the bins present in
  // each stage of each channel block
  if (ctx->pulsar) HIPCHK(ctx, ctx->psr_bins.ensure((size_t)lp.chunk_segs * (nc + 1) * sizeof(int16_t)));
  if (ctx->pulsar) {
    const size_t words =
        mixed ? pulsar_bin_mask_words(ctx->opt.precision, M, psr_run, ctx->nbins, lp.chunk_segs)
              : (size_t)((lp.chunk_segs + 15) / 16) * ((nc + 1 + 15) / 16) * ((ctx->nbins + 31) / 32);
    HIPCHK(ctx, ctx->psr_masks.ensure(std::max<size_t>(words, 1) * sizeof(uint32_t)));
  }


  // host staging stays alive in the context until the next slice, so the
  // copies and kernels below run asynchronously
  std::vector<double> &model_host = ctx->h_model;
  model_host.assign((size_t)mod_tot, 0.0);
  std::vector<DevStream> &ds = ctx->h_streams;
  ds.assign(M, DevStream());
  for (int i = 0; i < M; i++) {
    const StreamPlan &pl = ctx->plans[i];
    if (!pl.ext_payload && !pl.payload.empty())
      HIPCHK(ctx, hipMemcpyAsync(ctx->payload.as<uint8_t>() + pay_off[i], pl.payload.data(), pl.payload.size(),
                                 hipMemcpyHostToDevice, s));
    if (!pl.runs.empty())
      HIPCHK(ctx, hipMemcpyAsync(ctx->runs.as<Run>() + run_off[i], pl.runs.data(), pl.runs.size() * sizeof(Run),
                                 hipMemcpyHostToDevice, s));
    if (!pl.counted.empty())
      HIPCHK(ctx, hipMemcpyAsync(ctx->counted.as<CountRange>() + cnt_off[i], pl.counted.data(),
                                 pl.counted.size() * sizeof(CountRange), hipMemcpyHostToDevice, s));
    const ModelCopy &m = ctx->models[i];
    double *mh = model_host.data() + mod_off[i];
    std::copy(m.delay.begin(), m.delay.end(), mh);
    std::copy(m.phase.begin(), m.phase.end(), mh + m.delay.size());
    std::copy(m.amp.begin(), m.amp.end(), mh + m.delay.size() + m.phase.size());
    const double *md = ctx->model_buf.as<double>() + mod_off[i];

    const sfxg_station_parameters &st = ctx->st[i];
    DevStream &d = ds[i];
    memset(&d, 0, sizeof(d));
    d.payload = pl.ext_payload ? (const uint8_t *)pl.ext_payload : ctx->payload.as<uint8_t>() + pay_off[i];
    d.runs = ctx->runs.as<Run>() + run_off[i];
    d.block_src = ctx->block_run.as<int64_t>() + blk_off[i];  // filled by k_block_consts
    d.counted = ctx->counted.as<CountRange>() + cnt_off[i];
    d.n_runs = (int64_t)pl.runs.size();
    d.n_counted = (int64_t)pl.counted.size();
    d.bits = st.bits_per_sample;
    d.tsys_count0 = pl.tsys_count0;
    d.tsys_on0 = pl.tsys_on0;
    d.tsys_half = pl.tsys_half;
    for (int k = 0; k < 8; k++) d.lut[k] = 0.0f;
    if (st.bits_per_sample == 2)
      for (int k = 0; k < 4; k++) d.lut[k] = (float)kLut2[ctx->opt.lut][k];
    else
      for (int k = 0; k < 2; k++) d.lut[k] = (float)kLut1[ctx->opt.lut][k];
    d.model.scan_begin = m.scan_begin;
    d.model.interval = m.interval;
    d.model.n = m.n;
    d.model.has_phase = !m.phase.empty();
    d.model.has_amp = !m.amp.empty();
    d.model.delay = md;
    d.model.phase = m.phase.empty() ? nullptr : md + m.delay.size();
    d.model.amp = m.amp.empty() ? nullptr : md + m.delay.size() + m.phase.size();
    const int sb = sb_sign(st.sideband);
    const int fs = st.sample_rate;
    d.fs = fs;
    d.bandwidth = st.bandwidth;
    d.center_freq = (double)st.channel_freq + sb * st.bandwidth * 0.5 + st.LO_offset;
    d.mult_factor_phi = -sb * 2.0 * M_PI;
    d.LO_offset = st.LO_offset;
    {
      // start phase of the LO offset (delay_correction.cc:251-255)
      const double dt = (double)(ctx->p.start_time - ctx->p.experiment_start) / kClockRate;
      const double LO = st.LO_offset;
      double sp = LO * (dt - trunc(dt)) + (LO - trunc(LO)) * trunc(dt);
      d.start_phase = sp - floor(sp);
    }
    d.start_ticks = ctx->p.start_time;
    d.tps = (int64_t)llround(kClockRate / fs);
    d.fl = (int64_t)llround((double)nd / (fs / 1000000) * (kClockRate / 1000000));
    d.fl_half = (int64_t)llround(d.fl / 2.0);
    d.fl_sec = d.fl / kClockRate;
    d.oversamp = (int)round(fs / (2.0 * st.bandwidth));
    d.flip = ctx->p.sideband != st.sideband;
    d.time_series = fused ? nullptr
                          : reinterpret_cast<float *>(ctx->time_series.as<char>() + (size_t)i * lp.ts_stride * ts_bytes);
    d.spectra = reinterpret_cast<float2 *>(ctx->spectra.as<char>() + (size_t)i * lp.chunk_segs * ctx->S * spec_bytes);
    d.samples = nullptr;
    d.hist = ctx->hist.as<int32_t>() + 512 * i;
    d.block_consts = ctx->block_consts.as<BlockConst>() + blk_off[i];
  }
  if (mod_tot)
    HIPCHK(ctx, hipMemcpyAsync(ctx->model_buf.p, model_host.data(), model_host.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(ctx->dstreams.p, ds.data(), ds.size() * sizeof(DevStream), hipMemcpyHostToDevice, s));
  std::vector<int32_t> &blv = ctx->h_bl;
  blv.assign(2 * B, 0);
  std::copy(ctx->bl_first.begin(), ctx->bl_first.end(), blv.begin());
  std::copy(ctx->bl_second.begin(), ctx->bl_second.end(), blv.begin() + B);
  HIPCHK(ctx, hipMemcpyAsync(ctx->bl.p, blv.data(), blv.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (!ctx->h_tiles.empty())
    HIPCHK(ctx, hipMemcpyAsync(ctx->tiles.p, ctx->h_tiles.data(), ctx->h_tiles.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(ctx->flags.p, ctx->flag_factor.data(), B * sizeof(double), hipMemcpyHostToDevice, s));
  std::vector<int32_t> &fold_start = ctx->h_fold_start;
  fold_start.assign(1, 0);
  fold_start.insert(fold_start.end(), ctx->folds.begin(), ctx->folds.end());
  HIPCHK(ctx, hipMemcpyAsync(ctx->fold_buf.p, fold_start.data(), fold_start.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, s));
  // extra phase centres: the uv shift of every sub-integration
  if (ctx->n_pc > 1) {
    uv_table(ctx);
    HIPCHK(ctx, ctx->uv_tab.ensure(std::max<size_t>(ctx->h_uv.size(), 1) * sizeof(UvShift)));
    HIPCHK(ctx, hipMemcpyAsync(ctx->uv_tab.p, ctx->h_uv.data(), ctx->h_uv.size() * sizeof(UvShift),
                               hipMemcpyHostToDevice, s));
  }
  // phased array: blocks and rows of every chunk (correlation_core_phased.cc:116-126)
  lp.ph_b0.clear();
  lp.ph_nb.clear();
  lp.ph_r0.clear();
  lp.ph_nr.clear();
  if (ctx->phased) {
    if (ctx->opt.precision != SFXG_PREC_FP32)
      return ctx->fail(SFXG_ERR_ARG, "phased array mode takes SFXG_PREC_FP32");
    ctx->h_ph_blocks.clear();
    ctx->h_ph_rows.clear();
    size_t max_nb = 1;
    for (const Chunk &ch : lp.chunks) {
      const size_t b0 = ctx->h_ph_blocks.size(), r0 = ctx->h_ph_rows.size();
      phased_tables(ch.seg0, ch.seg1, ctx->n_sub, ctx->h_ph_blocks, ctx->h_ph_rows);
      lp.ph_b0.push_back((int32_t)b0);
      lp.ph_nb.push_back((int32_t)(ctx->h_ph_blocks.size() - b0));
      lp.ph_r0.push_back((int32_t)r0);
      lp.ph_nr.push_back((int32_t)(ctx->h_ph_rows.size() - r0));
      max_nb = std::max(max_nb, ctx->h_ph_blocks.size() - b0);
    }
    for (const PhasedRow &r : ctx->h_ph_rows)
      if (r.row < 0 || r.row >= B) return ctx->fail(SFXG_ERR_STATE, "phased row %d outside %d rows", r.row, B);
    HIPCHK(ctx, ctx->ph_blocks.ensure(std::max<size_t>(ctx->h_ph_blocks.size(), 1) * sizeof(PhasedBlock)));
    HIPCHK(ctx, ctx->ph_rows.ensure(std::max<size_t>(ctx->h_ph_rows.size(), 1) * sizeof(PhasedRow)));
    HIPCHK(ctx, ctx->partial.ensure(max_nb * (size_t)(nc + 1) * sizeof(float2)));
    HIPCHK(ctx, hipMemcpyAsync(ctx->ph_blocks.p, ctx->h_ph_blocks.data(),
                               ctx->h_ph_blocks.size() * sizeof(PhasedBlock), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->ph_rows.p, ctx->h_ph_rows.data(), ctx->h_ph_rows.size() * sizeof(PhasedRow),
                               hipMemcpyHostToDevice, s));
  }
  // launch parameters of the device work
  lp.fused = fused;
  lp.fused_ok = fused_ok;
  lp.key_fused = key_fused;
  lp.mixed = mixed;
  lp.k_group = k_group;
  lp.psr_run = psr_run;
  lp.vis_elems = vis_elems;
  lp.max_counted = max_counted;
  lp.job.n_streams = M;
  lp.job.nd = nd;
  lp.job.nc = nc;
  lp.job.window = ctx->p.window;
  lp.job.n_blocks = ctx->n_blocks;
  lp.job.n_samples = ctx->n_samples;
  lp.job.n_segments = ctx->nseg;  // per chunk in enqueue_slice
  lp.job.spectra_stride = ctx->S;
  lp.job.seg_step = nc;  // fft_rot_size / 2
  lp.job.spec_prec = spec_prec;
  lp.job.ts_half = ts_half ? 1 : 0;
  lp.job.spec_scale = 1.0f;
  lp.mixed_scale = 1.0f;
  if (mixed) {
    // power-of-two scale that brings the spectra's rms near 16: the delay
    // correction's unnormalised FFT pair gains N_d / 2, the channeliser
    // sqrt(sum w^2); the decode table bounds the sample rms
    const double wsum = ctx->window_sumsq;
    double lmax = 0.0;
    for (int i = 0; i < M; i++) {
      const int lut = ctx->opt.lut;
      lmax = std::max(lmax, ctx->st[i].bits_per_sample == 2 ? fabs(kLut2[lut][0]) : fabs(kLut1[lut][0]));
    }
    const double rms = (nd / 2.0) * lmax * sqrt(std::max(wsum, 1.0));
    lp.mixed_scale = (float)ldexp(1.0, (int)lround(log2(16.0 / rms)));
    ctx->mixed_scale = lp.mixed_scale;
    if (spec_prec) lp.job.spec_scale = lp.mixed_scale;
  }
  if (ctx->pulsar) {
    // Correlation_core_pulsar::set_parameters (correlation_core_pulsar.cc:46-93)
    const int64_t tpd = (int64_t)86400 * (int64_t)kClockRate;
    const int64_t t = ctx->p.start_time;
    const double start_mjd = 51544 + (double)(t / tpd) + (t % tpd) / kClockRate / 86400;
    const double us_per_day = (double)((int64_t)24 * 60 * 60 * 1000000);
    const double fft_duration = ((double)nc * 1000000) / ctx->p.sample_rate;
    const sfxg_pulsar &ps = ctx->psr;
    double DT;
    if (ctx->p.window == SFXG_WINDOW_NONE) DT = ((start_mjd - ps.tmid) + fft_duration / (2 * us_per_day)) * 1440;
    else DT = ((start_mjd - ps.tmid) + fft_duration / us_per_day) * 1440;
    const int N = ps.n_coef;
    double start_phase = 0;
    for (int i = N - 1; i > 0; i--) start_phase = (start_phase + ps.coef[i]) * DT;
    start_phase += (ps.ref_phase - floor(ps.ref_phase)) + DT * 60 * ps.ref_freq + ps.coef[0];
    const int sb = sb_sign(ctx->p.sideband);
    const double base_freq = ctx->p.channel_freq * 1e-6;
    const double dfreq = sb * ctx->p.sample_rate * 1e-6 / (2 * nc);
    const double inv_freq_obs2 = 1 / (ps.obs_freq * ps.obs_freq);
    double freq = 0;
    for (int i = N - 1; i > 1; i--) freq = (freq + i * ps.coef[i]) * DT;
    freq = (freq + ps.coef[1]) / 60 + ps.ref_freq;
    std::vector<double> &offs = ctx->h_offsets;
    offs.assign(nc + 1, 0.0);
    for (int i = 0; i < nc + 1; i++)
      offs[i] = 4148.808 * ps.DM * (1 / pow(base_freq + i * dfreq, 2) - inv_freq_obs2) * freq;
    HIPCHK(ctx, ctx->offsets.ensure(offs.size() * sizeof(double)));
    HIPCHK(ctx, hipMemcpyAsync(ctx->offsets.p, offs.data(), offs.size() * sizeof(double), hipMemcpyHostToDevice, s));
    lp.start_phase = start_phase;
    lp.phase_coef = ps.coef[1] + 60 * ps.ref_freq;
    lp.fft_duration = fft_duration;
  }
  HIPCHK(ctx, hipEventRecord(ctx->prep_done, s));
  ctx->prepared = true;
  ctx->prep_gen++;
  return SFXG_OK;
}

// F-engine of chunk c (block constants first, with the first chunk);
// timing: the delay/channeliser split event of the chunk.
static int enqueue_fengine(sfxg_ctx *ctx, size_t c, hipStream_t s, bool timing, const Tune &t) {
  const LaunchPlan &lp = ctx->lp;
  const Chunk &ch = lp.chunks[c];
  DevJob cj = lp.job;
  cj.seg_base = ch.seg0;
  cj.n_segments = ch.seg1 - ch.seg0;
  cj.blk_base = ch.blk0;
  cj.blk_end = ch.blk1;
  const DevStream *dstr = ctx->dstreams.as<DevStream>();
  if (c == 0) HIPCHK(ctx, launch_block_consts(dstr, cj, s));
  // a plan without the real series (lp.fused) runs the fused F-engine; one
  // with it runs whichever route t names (the autotuner times both)
  if (lp.fused || (lp.fused_ok && t.fe_fused == 1)) {
    HIPCHK(ctx, launch_fengine(dstr, cj, ctx->window.as<float>(), ctx->tw_nd.as<float2>(), ctx->tw_split.as<float2>(),
                               t, s));
    HIPCHK(ctx, timing ? hipEventRecord(ctx->cev[4 * c + 1], s) : hipSuccess);
  } else {
    HIPCHK(ctx, launch_delay(dstr, cj, ctx->tw_nd.as<float2>(), s));
    HIPCHK(ctx, timing ? hipEventRecord(ctx->cev[4 * c + 1], s) : hipSuccess);
    HIPCHK(ctx, launch_channelise(dstr, cj, ctx->window.as<float>(), ctx->tw_n.as<float2>(), ctx->tw_2n.as<float2>(),
                                  ctx->tw_split.as<float2>(), ctx->chan_scratch.as<float2>(), t, s));
  }
  return SFXG_OK;
}

// X-step of chunk c: the chunk's sub-integrations into vis (in fold order),
// or the pulsar bins of its segments.
static int enqueue_xstep(sfxg_ctx *ctx, size_t c, hipStream_t s, const Tune &t) {
  const LaunchPlan &lp = ctx->lp;
  const Chunk &ch = lp.chunks[c];
  const int M = ctx->M, nc = ctx->nc, B = ctx->B;
  const int32_t *blf = ctx->bl.as<int32_t>(), *bls = ctx->bl.as<int32_t>() + B;
  const int64_t sstride = lp.chunk_segs * ctx->S;  // spectra elements per stream
  const sfxg_pulsar &ps = ctx->psr;
  if (ctx->phased) {
    HIPCHK(ctx, launch_phased(ctx->spectra.as<float2>(), sstride, ctx->S, ch.seg0, nc + 1, M,
                              ctx->ph_blocks.as<PhasedBlock>() + lp.ph_b0[c], lp.ph_nb[c],
                              ctx->ph_rows.as<PhasedRow>() + lp.ph_r0[c], lp.ph_nr[c], ctx->partial.as<float2>(),
                              ctx->vis.as<float2>(), s));
  } else if (lp.mixed && !ctx->pulsar) {
    for (int g0 = ch.fold0; g0 < ch.fold1; g0 += lp.k_group) {
      const int ng = std::min(lp.k_group, ch.fold1 - g0);
      HIPCHK(ctx, launch_xstep_mfma(ctx->opt.precision, ctx->spectra.p, lp.job.spec_prec != 0, sstride, ctx->S, ch.seg0, nc + 1,
                                    M, ctx->fold_buf.as<int32_t>() + g0, ng, blf, bls, B, lp.mixed_scale,
                                    ctx->partial.as<float2>(), ctx->vis.as<float2>(), t, s));
    }
  } else if (!ctx->pulsar) {
    // with extra phase centres the partials are folded into every centre below
    float2 *fold_vis = ctx->n_pc > 1 ? nullptr : ctx->vis.as<float2>();
    for (int g0 = ch.fold0; g0 < ch.fold1; g0 += lp.k_group) {
      const int ng = std::min(lp.k_group, ch.fold1 - g0);
      const int n_tiles = (int)(ctx->h_tiles.size() / 18);
      if (t.xs_waves >= 0 && t.xs_waves <= 2)
        HIPCHK(ctx, launch_xstep_tile(ctx->spectra.as<float2>(), sstride, ctx->S, ch.seg0, nc + 1, M,
                                      ctx->fold_buf.as<int32_t>() + g0, ng, ctx->tiles.as<int32_t>(), n_tiles, blf,
                                      bls, B, ctx->partial.as<float2>(), fold_vis,
                                      t.xs_waves == 2 || (t.xs_waves == 0 && xstep_tile16_default(n_tiles)), s));
      else
        HIPCHK(ctx, launch_xstep(ctx->spectra.as<float2>(), sstride, ctx->S, ch.seg0, nc + 1, M,
                                 ctx->fold_buf.as<int32_t>() + g0, ng, blf, bls, B, ctx->partial.as<float2>(),
                                 fold_vis, t, s));
      if (ctx->n_pc > 1)
        HIPCHK(ctx, launch_fold_combine_pc(ctx->partial.as<float2>(), ng, B, nc + 1, M,
                                           ctx->uv_tab.as<UvShift>() + (size_t)g0 * (ctx->n_pc - 1) * B, ctx->n_pc,
                                           ctx->vis.as<float2>(), s));
    }
  } else if (lp.mixed) {
    // with several chunks the channel run covers all channels (checked in prepare)
    for (int c0 = 0; c0 < nc + 1; c0 += lp.psr_run) {
      const int n_run = std::min(lp.psr_run, nc + 1 - c0);
      if (c == 0 || lp.psr_run < nc + 1) HIPCHK(ctx, pulsar_gram_zero(M, n_run, ctx->nbins, ctx->gram.as<float>(), s));
      HIPCHK(ctx, launch_xstep_mfma_pulsar(ctx->opt.precision, ctx->spectra.p, lp.job.spec_prec != 0, sstride, ctx->S, c0, n_run,
                                           M, ch.seg0, ch.seg1 - ch.seg0, lp.mixed_scale, ctx->offsets.as<double>(),
                                           lp.start_phase, lp.phase_coef, lp.fft_duration, ps.gate_start,
                                           ps.gate_stop, ctx->nbins, ctx->gram.as<float>(),
                                           ctx->psr_bins.as<int16_t>(), ctx->psr_masks.as<uint32_t>(), t, s));
      if (c + 1 == lp.chunks.size() || lp.psr_run < nc + 1)
        HIPCHK(ctx, pulsar_gram_baselines(ctx->gram.as<float>(), nc + 1, c0, n_run, M, ctx->nbins, blf, bls, B,
                                          lp.mixed_scale, ctx->vis.as<float2>(), s));
    }
  } else {
    HIPCHK(ctx, launch_xstep_pulsar(ctx->spectra.as<float2>(), sstride, ctx->S, nc + 1, ch.seg0, ch.seg1 - ch.seg0,
                                    blf, bls, B, ctx->offsets.as<double>(), lp.start_phase, lp.phase_coef,
                                    lp.fft_duration, ps.gate_start, ps.gate_stop, ctx->nbins,
                                    ctx->psr_bins.as<int16_t>(), ctx->psr_masks.as<uint32_t>(), ctx->vis.as<float2>(),
                                    ctx->tiles.as<int32_t>(), (int)(ctx->h_tiles.size() / 18), M,
                                    t.xs_waves == 0 || t.xs_waves == 2, s));
  }
  return SFXG_OK;
}

// The device work of a prepared slice on stream s: no allocation, no host
// synchronisation and no memset nodes, so that it can be captured into a
// graph. timing: bracket the stages with the context's events.
static int enqueue_slice(sfxg_ctx *ctx, hipStream_t s, bool timing) {
  const LaunchPlan &lp = ctx->lp;
  const int M = ctx->M, nc = ctx->nc, B = ctx->B;
  auto mark = [&](int k) { return timing ? hipEventRecord(ctx->ev[k], s) : hipSuccess; };
  HIPCHK(ctx, launch_fill_zero(ctx->hist.p, (size_t)M * 512 * sizeof(int32_t), s));
  HIPCHK(ctx, launch_fill_zero(ctx->vis.p, lp.vis_elems * sizeof(float2), s));
  const DevStream *dstr = ctx->dstreams.as<DevStream>();
  HIPCHK(ctx, mark(SFXG_KERNEL_STATS));
  HIPCHK(ctx, launch_stats(dstr, M, lp.max_counted, s));
  HIPCHK(ctx, mark(SFXG_KERNEL_DELAY));
  ctx->fused = lp.fused;
  for (size_t c = 0; c < lp.chunks.size(); c++) {
    HIPCHK(ctx, timing ? hipEventRecord(ctx->cev[4 * c], s) : hipSuccess);
    const int rc = enqueue_fengine(ctx, c, s, timing, ctx->tune);
    if (rc != SFXG_OK) return rc;
    HIPCHK(ctx, timing ? hipEventRecord(ctx->cev[4 * c + 2], s) : hipSuccess);
    const int rc2 = enqueue_xstep(ctx, c, s, ctx->tune);
    if (rc2 != SFXG_OK) return rc2;
    HIPCHK(ctx, timing ? hipEventRecord(ctx->cev[4 * c + 3], s) : hipSuccess);
  }
  const int32_t *blf = ctx->bl.as<int32_t>(), *bls = ctx->bl.as<int32_t>() + B;
  HIPCHK(ctx, mark(SFXG_KERNEL_NORMALISE));
  if (ctx->phased) {  // Correlation_core_phased writes the sums as they are
    HIPCHK(ctx, mark(SFXG_N_KERNELS));
    return SFXG_OK;
  }
  HIPCHK(ctx, launch_auto_norms(ctx->vis.as<float2>(), M, nc + 1, nc, ctx->nbins, B, ctx->norms.as<double>(), s));
  HIPCHK(ctx, launch_normalise(ctx->vis.as<float2>(), ctx->norms.as<double>(), blf, bls, ctx->flags.as<double>(), M,
                               B, nc + 1, ctx->nbins, ctx->norm_out.as<float2>(), s));
  if (ctx->nchan != nc)
    HIPCHK(ctx, launch_rechannel(ctx->norm_out.as<float2>(), ctx->chan_out.as<float2>(), ctx->nbins * B, nc,
                                 ctx->nchan, ctx->rc_mask.as<float>(), ctx->rc_window_buf.as<float>(),
                                 ctx->rc_normalize ? 1 : 0, ctx->rc_tw.as<float2>(), s));
  HIPCHK(ctx, mark(SFXG_N_KERNELS));
  return SFXG_OK;
}

extern "C" {

int sfxg_prepare(sfxg_ctx *ctx) { return ctx ? prepare_slice(ctx) : SFXG_ERR_ARG; }

int sfxg_correlate(sfxg_ctx *ctx) {
  if (!ctx) return SFXG_ERR_ARG;
  if (ctx->batch_member)
    return ctx->fail(SFXG_ERR_STATE, "the context belongs to a batch: sfxg_prepare and sfxg_batch_launch");
  const int rc = prepare_slice(ctx);
  if (rc != SFXG_OK) return rc;
  const int rc2 = enqueue_slice(ctx, ctx->stream, true);
  if (rc2 != SFXG_OK) return rc2;
  ctx->timed_chunks = (int)ctx->lp.chunks.size();
  HIPCHK(ctx, hipEventRecord(ctx->done, ctx->stream));
  ctx->enqueued = true;
  ctx->have_results = true;
  ctx->times_pending = true;
  return SFXG_OK;
}

int sfxg_synchronize(sfxg_ctx *ctx) {
  if (!ctx) return SFXG_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (ctx->batch_done) {
    HIPCHK(ctx, hipEventSynchronize(ctx->batch_done));
    ctx->batch_done = nullptr;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return SFXG_OK;
}

static size_t output_size(const sfxg_ctx *ctx) {
  std::set<int> stations;
  for (const auto &s : ctx->st) stations.insert(s.station_number);
  const size_t one = 4 + 16 + 32 * stations.size() + 24 * (size_t)ctx->M +
                     (size_t)ctx->B * (8 + 8 * (size_t)(ctx->nchan + 1));
  return one * ctx->nbins;
}

int sfxg_output_size(sfxg_ctx *ctx, size_t *bytes) {
  if (!ctx || !bytes) return SFXG_ERR_ARG;
  if (!ctx->have_params) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_parameters first");
  *bytes = output_size(ctx);
  return SFXG_OK;
}

static int stats_of(sfxg_ctx *ctx, std::vector<int32_t> &levels, std::vector<int32_t> &tsys) {
  std::vector<int32_t> hist((size_t)ctx->M * 512);
  HIPCHK(ctx, hipMemcpy(hist.data(), ctx->hist.p, hist.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  levels.assign(5 * ctx->M, 0);
  tsys.assign(4 * ctx->M, 0);
  for (int i = 0; i < ctx->M; i++)
    bit_statistics(ctx->st[i].bits_per_sample, ctx->st[i].sample_rate, ctx->p.sample_rate, &hist[512 * i],
                   &hist[512 * i + 256], ctx->plans[i].n_invalid, &levels[5 * i], &tsys[4 * i]);
  return SFXG_OK;
}

int sfxg_get_statistics(sfxg_ctx *ctx, int32_t stream_idx, int32_t levels[5], int32_t tsys[4]) {
  if (!ctx || !levels || !tsys) return SFXG_ERR_ARG;
  if (!ctx->have_results) return ctx->fail(SFXG_ERR_STATE, "sfxg_correlate first");
  if (stream_idx < 0 || stream_idx >= ctx->M) return ctx->fail(SFXG_ERR_ARG, "stream index %d", stream_idx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  std::vector<int32_t> lv, ts;
  int rc = stats_of(ctx, lv, ts);
  if (rc) return rc;
  std::copy(&lv[5 * stream_idx], &lv[5 * stream_idx + 5], levels);
  std::copy(&ts[4 * stream_idx], &ts[4 * stream_idx + 4], tsys);
  return SFXG_OK;
}

int sfxg_number_of_baselines(sfxg_ctx *ctx, int32_t *n) {
  if (!ctx || !n) return SFXG_ERR_ARG;
  if (!ctx->have_params) return ctx->fail(SFXG_ERR_STATE, "sfxg_set_parameters first");
  *n = ctx->B;
  return SFXG_OK;
}

// The rows integration_write stores: the phased core's sums, the lag
// re-channelised rows, or the normalised visibilities.
static const void *written_rows(const sfxg_ctx *ctx) {
  if (ctx->phased) return ctx->vis.p;
  return ctx->nchan != ctx->nc ? ctx->chan_out.p : ctx->norm_out.p;
}

// integration_write (correlation_core.cc:316-477)
int sfxg_fetch_output(sfxg_ctx *ctx, uint8_t *dst, size_t cap, size_t *written) {
  if (!ctx || !dst) return SFXG_ERR_ARG;
  if (!ctx->have_results) return ctx->fail(SFXG_ERR_STATE, "sfxg_correlate first");
  const size_t need = output_size(ctx);
  if (cap < need) return ctx->fail(SFXG_ERR_ARG, "output buffer holds %zu bytes, %zu needed", cap, need);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  const int M = ctx->M, B = ctx->B, nch = ctx->nchan + 1;
  std::vector<float2> vis((size_t)ctx->nbins * B * nch);
  HIPCHK(ctx, hipMemcpy(vis.data(), written_rows(ctx), vis.size() * sizeof(float2), hipMemcpyDeviceToHost));
  std::vector<int32_t> lv, ts;
  int rc = stats_of(ctx, lv, ts);
  if (rc) return rc;
  // unique stations in stream order and the first stream of each
  std::vector<int> stations, first_stream;
  for (int i = 0; i < M; i++) {
    if (std::find(stations.begin(), stations.end(), ctx->st[i].station_number) == stations.end()) {
      stations.push_back(ctx->st[i].station_number);
      first_stream.push_back(i);
    }
  }
  uint8_t *o = dst;
  auto put = [&o](const void *src, size_t n) {
    memcpy(o, src, n);
    o += n;
  };
  for (int bin = 0; bin < ctx->nbins; bin++) {
    int32_t sourcenr = ctx->pulsar ? bin : ctx->n_pc > 1 ? ctx->pc_source[bin] : (ctx->p.pulsar_binning ? 1 : 0);
    put(&sourcenr, 4);
    const int32_t ts_hdr[4] = {ctx->p.integration_nr, B, (int32_t)stations.size(), M};
    put(ts_hdr, 16);
    for (size_t k = 0; k < stations.size(); k++) {
      const int32_t h[2] = {stations[k], 0};
      put(h, 8);
      put(ctx->n_pc > 1 ? &ctx->pc_uvw[((size_t)bin * M + first_stream[k]) * 3] : &ctx->uvw[3 * first_stream[k]], 24);
    }
    for (int i = 0; i < M; i++) {
      const sfxg_station_parameters &st = ctx->st[i];
      const uint8_t h[4] = {(uint8_t)st.station_number, (uint8_t)ctx->p.frequency_nr,
                            (uint8_t)(ctx->p.sideband == 'L' ? 0 : 1), (uint8_t)(st.polarisation == 'R' ? 0 : 1)};
      put(h, 4);
      int32_t l5[5];
      const int32_t *L = &lv[5 * i];
      if (st.bits_per_sample == 2) {
        l5[0] = L[0]; l5[1] = L[1]; l5[2] = L[2]; l5[3] = L[3];
      } else {
        l5[0] = 0; l5[1] = L[0]; l5[2] = L[1]; l5[3] = 0;
      }
      l5[4] = L[4];
      put(l5, 20);
    }
    for (int b = 0; b < B; b++) {
      const int a = ctx->bl_first[b], c = ctx->bl_second[b];
      int64_t w;
      if (a == c) w = std::max<int64_t>(ctx->total_samples - lv[5 * a + 4], 0);
      else w = std::max<int64_t>(ctx->total_samples - lv[5 * a + 4] - ctx->n_flagged_first[b], 0);
      const int32_t w32 = (int32_t)w;
      put(&w32, 4);
      uint8_t h[4];
      h[0] = (uint8_t)ctx->st[a].station_number;
      h[1] = (uint8_t)ctx->st[c].station_number;
      const int p1 = ctx->st[a].polarisation == 'R' ? 0 : 1, p2 = ctx->st[c].polarisation == 'R' ? 0 : 1;
      const int sbu = ctx->p.sideband == 'U' ? 1 : 0;
      h[2] = (uint8_t)((p1 & 1) | ((p2 & 1) << 1) | (sbu << 2) | ((ctx->p.frequency_nr & 31) << 3));
      h[3] = ' ';
      put(h, 4);
      put(&vis[((size_t)bin * B + b) * nch], sizeof(float2) * nch);
    }
  }
  if (written) *written = (size_t)(o - dst);
  return SFXG_OK;
}

int sfxg_fetch_tsys(sfxg_ctx *ctx, uint8_t *dst, size_t cap, size_t *written) {
  if (!ctx || !dst) return SFXG_ERR_ARG;
  if (!ctx->have_results) return ctx->fail(SFXG_ERR_STATE, "sfxg_correlate first");
  const size_t rec = 4 + 8 + 4 * 8, need = rec * (size_t)ctx->M;
  if (cap < need) return ctx->fail(SFXG_ERR_ARG, "tsys buffer holds %zu bytes, %zu needed", cap, need);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  std::vector<int32_t> lv, ts;
  const int rc = stats_of(ctx, lv, ts);
  if (rc) return rc;
  uint8_t *o = dst;
  for (int i = 0; i < ctx->M; i++) {
    const sfxg_station_parameters &st = ctx->st[i];
    const uint8_t h[4] = {(uint8_t)st.station_number, (uint8_t)ctx->p.frequency_nr,
                          (uint8_t)(ctx->p.sideband == 'L' ? 0 : 1), (uint8_t)(st.polarisation == 'R' ? 0 : 1)};
    memcpy(o, h, 4);
    const uint64_t ticks = (uint64_t)ctx->p.start_time;
    memcpy(o + 4, &ticks, 8);
    // tsys_on_lo = tsys[0], tsys_on_hi = tsys[1], tsys_off_lo = tsys[2], tsys_off_hi = tsys[3]
    for (int k = 0; k < 4; k++) {
      const int64_t v = ts[4 * i + k];
      memcpy(o + 12 + 8 * k, &v, 8);
    }
    o += rec;
  }
  if (written) *written = need;
  return SFXG_OK;
}

int sfxg_fetch_stage(sfxg_ctx *ctx, int32_t stage, int32_t stream_idx, void *dst, size_t cap, size_t *written) {
  if (!ctx || !dst) return SFXG_ERR_ARG;
  if (!ctx->have_results) return ctx->fail(SFXG_ERR_STATE, "sfxg_correlate first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  const bool per_stream = stage == SFXG_STAGE_SAMPLES || stage == SFXG_STAGE_TIME_SERIES || stage == SFXG_STAGE_SPECTRA;
  if (per_stream && (stream_idx < 0 || stream_idx >= ctx->M))
    return ctx->fail(SFXG_ERR_ARG, "stream index %d", stream_idx);
  size_t need = 0;
  switch (stage) {
    case SFXG_STAGE_SAMPLES: {
      need = (size_t)ctx->n_samples * sizeof(float);
      if (cap < need) return ctx->fail(SFXG_ERR_ARG, "buffer too small");
      HIPCHK(ctx, ctx->samples.ensure(need));
      DevStream d;
      HIPCHK(ctx, hipMemcpy(&d, ctx->dstreams.as<DevStream>() + stream_idx, sizeof(d), hipMemcpyDeviceToHost));
      d.samples = ctx->samples.as<float>();
      DBuf one;
      HIPCHK(ctx, one.ensure(sizeof(d)));
      HIPCHK(ctx, hipMemcpy(one.p, &d, sizeof(d), hipMemcpyHostToDevice));
      hipError_t e = launch_unpack(one.as<DevStream>(), 1, ctx->n_samples, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e == hipSuccess) e = hipMemcpy(dst, ctx->samples.p, need, hipMemcpyDeviceToHost);
      one.release();
      HIPCHK(ctx, e);
      break;
    }
    case SFXG_STAGE_TIME_SERIES:
      if (ctx->fused)
        return ctx->fail(SFXG_ERR_STATE, "the real series stayed in LDS: set keep_time_series in sfxg_options");
      if (ctx->lp.job.ts_half)
        return ctx->fail(SFXG_ERR_STATE,
                         "the real series was held as FP16 for the MFMA channeliser: set keep_time_series in "
                         "sfxg_options");
      if (ctx->lp.chunks.size() != 1)
        return ctx->fail(SFXG_ERR_STATE, "the slice ran in %d chunks: stages hold the last one only",
                         (int)ctx->lp.chunks.size());
      need = (size_t)ctx->n_samples * sizeof(float);
      if (cap < need) return ctx->fail(SFXG_ERR_ARG, "buffer too small");
      HIPCHK(ctx, hipMemcpy(dst, ctx->time_series.as<float>() + (int64_t)stream_idx * ctx->lp.ts_stride, need,
                            hipMemcpyDeviceToHost));
      break;
    case SFXG_STAGE_SPECTRA:
      if (ctx->lp.chunks.size() != 1)
        return ctx->fail(SFXG_ERR_STATE, "the slice ran in %d chunks: stages hold the last one only",
                         (int)ctx->lp.chunks.size());
      need = (size_t)ctx->nseg * (ctx->nc + 1) * sizeof(float2);
      if (cap < need) return ctx->fail(SFXG_ERR_ARG, "buffer too small");
      if (ctx->lp.job.spec_prec == 0) {
        HIPCHK(ctx, hipMemcpy2D(dst, (ctx->nc + 1) * sizeof(float2),
                                ctx->spectra.as<float2>() + (int64_t)stream_idx * ctx->lp.chunk_segs * ctx->S,
                                ctx->S * sizeof(float2), (ctx->nc + 1) * sizeof(float2), ctx->nseg,
                                hipMemcpyDeviceToHost));
      } else {
        // packed FP16 / BF16 pairs (split FP16: hi and lo pairs, summed),
        // returned as float2 with the scale undone
        const int W = ctx->lp.job.spec_prec == SFXG_PREC_SPLIT_FP16 ? 2 : 1;
        const size_t row = (size_t)(ctx->nc + 1);
        std::vector<uint32_t> words((size_t)ctx->nseg * row * W);
        HIPCHK(ctx, hipMemcpy2D(words.data(), row * 4 * W,
                                ctx->spectra.as<uint32_t>() + (int64_t)stream_idx * ctx->lp.chunk_segs * ctx->S * W,
                                ctx->S * 4 * W, row * 4 * W, ctx->nseg, hipMemcpyDeviceToHost));
        const bool bf = ctx->lp.job.spec_prec == SFXG_PREC_BF16;
        const float inv = 1.0f / ctx->lp.job.spec_scale;
        auto half_to_float = [bf](uint32_t h) {
          if (bf) {
            const uint32_t b = h << 16;
            float f;
            memcpy(&f, &b, 4);
            return f;
          }
          const uint16_t u = (uint16_t)h;
          _Float16 x;
          memcpy(&x, &u, 2);
          return (float)x;
        };
        float2 *out = static_cast<float2 *>(dst);
        for (size_t k = 0; k < words.size() / W; k++) {
          float re = 0.0f, im = 0.0f;
          for (int w = W - 1; w >= 0; w--) {  // lo, then hi: the pair's sum rounds once
            re += half_to_float(words[W * k + w] & 0xffffu);
            im += half_to_float(words[W * k + w] >> 16);
          }
          out[k] = make_float2(re * inv, im * inv);
        }
      }
      break;
    case SFXG_STAGE_ACCUMULATED:
    case SFXG_STAGE_NORMALISED:
      need = (size_t)ctx->nbins * ctx->B * (ctx->nc + 1) * sizeof(float2);
      if (cap < need) return ctx->fail(SFXG_ERR_ARG, "buffer too small");
      HIPCHK(ctx, hipMemcpy(dst, stage == SFXG_STAGE_ACCUMULATED || ctx->phased ? ctx->vis.p : ctx->norm_out.p,
                            need, hipMemcpyDeviceToHost));
      break;
    case SFXG_STAGE_CHANNELS:
      need = (size_t)ctx->nbins * ctx->B * (ctx->nchan + 1) * sizeof(float2);
      if (cap < need) return ctx->fail(SFXG_ERR_ARG, "buffer too small");
      HIPCHK(ctx, hipMemcpy(dst, written_rows(ctx), need, hipMemcpyDeviceToHost));
      break;
    default:
      return ctx->fail(SFXG_ERR_ARG, "unknown stage %d", stage);
  }
  if (written) *written = need;
  return SFXG_OK;
}

int sfxg_kernel_times(sfxg_ctx *ctx, double *ms, int32_t n, int32_t reset) {
  if (!ctx || !ms) return SFXG_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (ctx->enqueued) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    collect_times(ctx);
  }
  for (int k = 0; k < SFXG_N_KERNELS && k < n; k++) ms[k] = ctx->kt_count ? ctx->kt_sum[k] / ctx->kt_count : 0.0;
  if (reset) {
    for (double &v : ctx->kt_sum) v = 0.0;
    ctx->kt_count = 0;
  }
  return SFXG_OK;
}

int sfxg_chunk_plan(const sfxg_ctx *ctx, int32_t *n_chunks, int64_t *max_segments) {
  if (!ctx || !n_chunks || !max_segments) return SFXG_ERR_ARG;
  if (!ctx->prepared) return SFXG_ERR_STATE;
  *n_chunks = (int32_t)ctx->lp.chunks.size();
  *max_segments = ctx->lp.chunk_segs;
  return SFXG_OK;
}

int sfxg_synthesize_raw(int device, void *device_dst, size_t n_bytes, int32_t bits_per_sample, int64_t first_sample,
                        int64_t delay_samples, uint64_t common_seed, uint64_t stream_seed, float rho,
                        float threshold, int64_t delay2_samples, uint64_t common2_seed, float rho2) {
  if (!device_dst || (bits_per_sample != 1 && bits_per_sample != 2) || rho < 0 || rho2 < 0 || rho + rho2 > 1)
    return SFXG_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return SFXG_ERR_DEVICE;
  if (launch_synthesize(device_dst, n_bytes, bits_per_sample, first_sample, delay_samples, common_seed, stream_seed,
                        rho, threshold, delay2_samples, common2_seed, rho2, nullptr) != hipSuccess)
    return SFXG_ERR_DEVICE;
  if (hipDeviceSynchronize() != hipSuccess) return SFXG_ERR_DEVICE;
  return SFXG_OK;
}

// ---------------------------------------------------------------------------
// Batches: the device work of several prepared contexts as one graph.
// ---------------------------------------------------------------------------
}  // extern "C"

struct sfxg_batch {
  int device = 0;
  hipStream_t stream = nullptr;           // the graph's origin stream
  std::vector<hipStream_t> branch;        // fork streams of the capture
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<sfxg_ctx *> ctxs;
  std::vector<uint64_t> gens;             // prep_gen of each context at capture
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

// A destroyed context leaves every batch that holds it (whose graph is
// dropped and captured again at the next launch).
static void leave_batches(sfxg_ctx *ctx) {
  for (sfxg_batch *b : ctx->batches) {
    (void)hipStreamSynchronize(b->stream);
    b->ctxs.erase(std::remove(b->ctxs.begin(), b->ctxs.end(), ctx), b->ctxs.end());
    b->drop_graph();
  }
  ctx->batches.clear();
  ctx->batch_member = 0;
  ctx->batch_done = nullptr;
}

extern "C" {

int sfxg_batch_create(int device, int32_t n_branches, sfxg_batch **out) {
  if (!out || n_branches < 1 || n_branches > 16) return SFXG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return SFXG_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return SFXG_ERR_DEVICE;
  std::unique_ptr<sfxg_batch> b(new sfxg_batch());
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return SFXG_ERR_DEVICE;
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) return SFXG_ERR_DEVICE;
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) return SFXG_ERR_DEVICE;
  for (int i = 0; i < n_branches; i++) {
    hipStream_t st;
    hipEvent_t ev;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return SFXG_ERR_DEVICE;
    b->branch.push_back(st);
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return SFXG_ERR_DEVICE;
    b->join.push_back(ev);
  }
  *out = b.release();
  return SFXG_OK;
}

const char *sfxg_batch_last_error(const sfxg_batch *b) { return b ? b->err.c_str() : "null batch"; }

int sfxg_batch_clear(sfxg_batch *b) {
  if (!b) return SFXG_ERR_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return SFXG_ERR_DEVICE;
  (void)hipStreamSynchronize(b->stream);
  for (sfxg_ctx *c : b->ctxs) {
    c->batch_member--;
    c->batches.erase(std::remove(c->batches.begin(), c->batches.end(), b), c->batches.end());
    if (c->batch_done == b->done) c->batch_done = nullptr;
  }
  b->ctxs.clear();
  b->gens.clear();
  b->drop_graph();
  return SFXG_OK;
}

void sfxg_batch_destroy(sfxg_batch *b) {
  if (!b) return;
  (void)sfxg_batch_clear(b);
  for (hipStream_t st : b->branch) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int sfxg_batch_add(sfxg_batch *b, sfxg_ctx *ctx) {
  if (!b || !ctx) return SFXG_ERR_ARG;
  if (ctx->device != b->device) return b->fail(SFXG_ERR_ARG, "context and batch are on different devices");
  for (sfxg_ctx *c : b->ctxs)
    if (c == ctx) return b->fail(SFXG_ERR_ARG, "context already in the batch");
  if (hipSetDevice(b->device) != hipSuccess) return SFXG_ERR_DEVICE;
  (void)hipStreamSynchronize(b->stream);
  b->ctxs.push_back(ctx);
  ctx->batch_member++;
  ctx->batches.push_back(b);
  b->drop_graph();
  return SFXG_OK;
}

// Captures the device work of every context, one fork stream per context in
// turn, into one graph.
static int batch_capture(sfxg_batch *b) {
  b->drop_graph();
  const int nb = (int)b->branch.size();
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return b->fail(SFXG_ERR_DEVICE, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  int rc = SFXG_OK;
  std::vector<bool> used(nb, false);
  if (hipEventRecord(b->fork, b->stream) != hipSuccess) rc = SFXG_ERR_DEVICE;
  for (size_t i = 0; rc == SFXG_OK && i < b->ctxs.size(); i++) {
    const int j = (int)(i % nb);
    if (!used[j]) {
      if (hipStreamWaitEvent(b->branch[j], b->fork, 0) != hipSuccess) {
        rc = SFXG_ERR_DEVICE;
        break;
      }
      used[j] = true;
    }
    rc = enqueue_slice(b->ctxs[i], b->branch[j], false);
    if (rc != SFXG_OK) b->err = "context " + std::to_string(i) + ": " + b->ctxs[i]->err;
  }
  for (int j = 0; j < nb; j++) {
    if (!used[j]) continue;
    if (hipEventRecord(b->join[j], b->branch[j]) != hipSuccess ||
        hipStreamWaitEvent(b->stream, b->join[j], 0) != hipSuccess)
      if (rc == SFXG_OK) rc = SFXG_ERR_DEVICE;
  }
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(b->stream, &g);
  if (rc != SFXG_OK) {
    if (g) (void)hipGraphDestroy(g);
    return b->err.empty() ? b->fail(rc, "capture failed") : rc;
  }
  if (e != hipSuccess) return b->fail(SFXG_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    b->drop_graph();
    return b->fail(SFXG_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  b->gens.clear();
  for (sfxg_ctx *c : b->ctxs) b->gens.push_back(c->prep_gen);
  return SFXG_OK;
}

int sfxg_batch_launch(sfxg_batch *b) {
  if (!b) return SFXG_ERR_ARG;
  if (b->ctxs.empty()) return b->fail(SFXG_ERR_STATE, "empty batch");
  if (hipSetDevice(b->device) != hipSuccess) return SFXG_ERR_DEVICE;
  bool same = b->exec != nullptr && b->gens.size() == b->ctxs.size();
  for (size_t i = 0; i < b->ctxs.size(); i++) {
    sfxg_ctx *c = b->ctxs[i];
    if (!c->prepared) return b->fail(SFXG_ERR_STATE, "context " + std::to_string(i) + " is not prepared");
    if (same && b->gens[i] != c->prep_gen) same = false;
  }
  // a replay must not overlap the previous one (same buffers)
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return b->fail(SFXG_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  if (!same) {
    const int rc = batch_capture(b);
    if (rc != SFXG_OK) return rc;
  }
  for (sfxg_ctx *c : b->ctxs)
    if (hipStreamWaitEvent(b->stream, c->prep_done, 0) != hipSuccess) return SFXG_ERR_DEVICE;
  e = hipGraphLaunch(b->exec, b->stream);
  if (e != hipSuccess) return b->fail(SFXG_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
  if (hipEventRecord(b->done, b->stream) != hipSuccess) return SFXG_ERR_DEVICE;
  for (sfxg_ctx *c : b->ctxs) {
    c->batch_done = b->done;
    c->enqueued = false;
    c->have_results = true;
    c->times_pending = false;
  }
  return SFXG_OK;
}

int sfxg_batch_synchronize(sfxg_batch *b) {
  if (!b) return SFXG_ERR_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return SFXG_ERR_DEVICE;
  const hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return b->fail(SFXG_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  return SFXG_OK;
}

// ---------------------------------------------------------------------------
// Autotuning of launch shapes
// ---------------------------------------------------------------------------
}  // extern "C"

// Minimum over reps runs (after one untimed run) of the GPU time of the work
// f enqueues on the context's stream.
template <class F>
static int time_stage(sfxg_ctx *ctx, int reps, F f, float *best_ms) {
  hipStream_t s = ctx->stream;
  int rc = f();
  if (rc != SFXG_OK) return rc;
  float best = -1.0f;
  for (int r = 0; r < reps; r++) {
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    rc = f();
    if (rc != SFXG_OK) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
    float ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    if (best < 0 || ms < best) best = ms;
  }
  *best_ms = best;
  return SFXG_OK;
}

extern "C" {

int sfxg_autotune(sfxg_ctx *ctx, int32_t reps, char *report, size_t cap) {
  if (!ctx || reps < 1) return SFXG_ERR_ARG;
  if (!ctx->prepared) return ctx->fail(SFXG_ERR_STATE, "sfxg_prepare first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  sync_slice(ctx);
  const LaunchPlan &lp = ctx->lp;
  const int nc = ctx->nc;
  std::string rep;
  char line[160];
  Tune best = ctx->tune;
  // F-engine: threads per workgroup (N_c <= 4096) on the fused F-engine or
  // the two-kernel route, or the channeliser's workgroups. Packed FP16 /
  // BF16 spectra at N_c 1024 to 16384 take the MFMA channeliser, fixed by
  // the prepared plan with its FP16 series (chan_fp32 = 1 in the table
  // selects the FP32 transform at the next preparation); with an FP32
  // series (keep_time_series) the FP32 transform at 8192 and 16384 is
  // among the candidates. The two-kernel route needs the real series in
  // HBM: not among the candidates of a plan prepared for the fused one.
  const bool small = nc <= 4096;
  const bool mf = chan_mfma_supported(nc, ctx->opt.precision);
  const bool fe = lp.fused_ok || (small && !mf);  // candidates are F-engine thread counts
  struct Cand {
    int fp32, fused, v;  // chan_fp32, fe_fused, threads or workgroups
  };
  std::vector<Cand> fc;
  if (fe) {
    for (int f = lp.fused ? 1 : 0; f <= (lp.fused_ok ? 1 : 0); f++)
      for (int v : fe_thread_candidates(nc)) fc.push_back({ctx->tune.chan_fp32, f, v});
  } else {
    if (mf)
      for (int v : chan_mfma_wgs_candidates()) fc.push_back({0, 0, v});
    // (the FP32 transform needs an FP32 series: not among the candidates
    // when the MFMA channeliser's FP16 one was prepared)
    if (!small && !lp.job.ts_half)
      for (int v : chan_wgs_candidates(nc)) fc.push_back({mf ? 1 : 0, 0, v});
  }
  if (fc.size() > 1) {
    float best_ms = -1.0f;
    Tune pick = best;
    for (const auto &c : fc) {
      Tune t = best;
      (fe ? t.fe_threads : t.chan_wgs) = c.v;
      t.chan_fp32 = c.fp32;
      if (fe) t.fe_fused = c.fused;
      float ms = 0.0f;
      const int rc = time_stage(ctx, reps, [&] { return enqueue_fengine(ctx, 0, ctx->stream, false, t); }, &ms);
      if (rc != SFXG_OK) return rc;
      const char *route = !fe ? (mf ? (c.fp32 ? " (FP32 transform)" : " (MFMA)") : "")
                              : lp.fused_ok ? (c.fused ? " (fused)" : " (delay + channeliser)") : "";
      snprintf(line, sizeof(line), "%s%s=%d: %.3f ms\n", fe ? "fe_threads" : "chan_wgs", route, c.v, ms);
      rep += line;
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        pick = t;
      }
    }
    best = pick;
  }
  // spectra of the chosen F-engine for the X-step variants
  {
    const int rc = enqueue_fengine(ctx, 0, ctx->stream, false, best);
    if (rc != SFXG_OK) return rc;
  }
  const bool mixed = ctx->opt.precision != SFXG_PREC_FP32;
  std::vector<int> xc;
  const char *xname = "";
  if (mixed && !ctx->pulsar) {
    xc = mf_channel_candidates(ctx->opt.precision, ctx->M);
    xname = "mf_channels";
  } else if (!ctx->pulsar) {
    xc = xs_waves_candidates();
    xname = "xs_waves";
  } else if (mixed) {
    xc = mfp_wgs_candidates();
    xname = "mfp_wgs";
  } else {
    xc = {0, 4};  // by bin on register tiles, or one thread per (baseline, channel)
    xname = "xs_waves";
  }
  if (xc.size() > 1) {
    float best_ms = -1.0f;
    Tune pick = best;
    for (int v : xc) {
      Tune t = best;
      if (mixed && !ctx->pulsar) t.mf_channels = v;
      else if (!mixed) t.xs_waves = v;
      else t.mfp_wgs = v;
      float ms = 0.0f;
      const int rc = time_stage(ctx, reps, [&] { return enqueue_xstep(ctx, 0, ctx->stream, t); }, &ms);
      if (rc != SFXG_OK) return rc;
      snprintf(line, sizeof(line), "%s=%d: %.3f ms\n", xname, v, ms);
      rep += line;
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        pick = t;
      }
    }
    best = pick;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[tune_key(ctx, lp.key_fused)] = best;
  }
  ctx->tune = best;
  // the trial runs overwrote the slice's buffers
  ctx->prepared = false;
  ctx->have_results = false;
  if (report && cap) {
    const size_t n = std::min(cap - 1, rep.size());
    memcpy(report, rep.data(), n);
    report[n] = 0;
  }
  return SFXG_OK;
}

int sfxg_tune_export(char *dst, size_t cap, size_t *written) {
  std::string out;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    char line[256];
    for (const auto &kv : g_tune) {
      const Tune &t = kv.second;
      snprintf(line, sizeof(line), "%s|%d|%d|%d|%d|%d|%d|%d|%d\n", kv.first.c_str(), t.fe_threads, t.chan_wgs,
               t.xs_waves, t.mf_channels, t.mfp_wgs, t.mf_ks, t.chan_fp32, t.fe_fused);
      out += line;
    }
  }
  if (written) *written = out.size() + 1;
  if (!dst || cap < out.size() + 1) return SFXG_ERR_ARG;
  memcpy(dst, out.c_str(), out.size() + 1);
  return SFXG_OK;
}

int sfxg_tune_import(const char *text) {
  if (!text) return SFXG_ERR_ARG;
  std::map<std::string, Tune> add;
  const char *p = text;
  while (*p) {
    const char *e = strchr(p, '\n');
    std::string ln = e ? std::string(p, e) : std::string(p);
    p = e ? e + 1 : p + ln.size();
    if (ln.empty()) continue;
    const size_t bar = ln.find('|');
    if (bar == std::string::npos) return SFXG_ERR_ARG;
    Tune t;
    const int nf = sscanf(ln.c_str() + bar + 1, "%d|%d|%d|%d|%d|%d|%d|%d", &t.fe_threads, &t.chan_wgs,
                          &t.xs_waves, &t.mf_channels, &t.mfp_wgs, &t.mf_ks, &t.chan_fp32, &t.fe_fused);
    if (nf < 5) return SFXG_ERR_ARG;
    add[ln.substr(0, bar)] = t;
  }
  std::lock_guard<std::mutex> g(g_tune_mu);
  for (const auto &kv : add) g_tune[kv.first] = kv.second;
  return SFXG_OK;
}

void sfxg_tune_clear(void) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune.clear();
}

}  // extern "C"
