/*
 * sfxg.h -- C ABI of the MI355X correlator-node data path.
 *
 * One context replaces the DSP objects one SFXC correlator node owns for a
 * time slice: a Bit2float_worker per stream, a Delay_correction per stream and
 * the Correlation_core (or Correlation_core_pulsar). The reference
 * interface each entry point replaces is cited next to it (paths relative
 * to the SFXC source tree, sfxc/).
 *
 * Conventions: every function returns SFXG_OK (0) or a negative SFXG_ERR_*
 * code; sfxg_last_error() then describes the failure. The caller owns all host
 * buffers. The context owns its device memory. Calls on one context are not
 * thread-safe and must come from one thread at a time, as the reference's
 * correlator-node main thread drives Delay_correction and Correlation_core
 * (src/correlator_node.cc:150-188).
 *
 * Times are int64 ticks of the reference Time class: 512 MHz clock ticks
 * since MJD 51544 (include/correlator_time.h:9-64, include/utils.h:36).
 */
#ifndef SFXG_H
#define SFXG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFXG_OK 0
#define SFXG_ERR_ARG (-1)      /* invalid argument or unsupported parameter */
#define SFXG_ERR_STREAM (-2)   /* malformed or incomplete HEADER byte stream */
#define SFXG_ERR_DEVICE (-3)   /* HIP runtime failure */
#define SFXG_ERR_STATE (-4)    /* call out of order (e.g. correlate before all streams) */
#define SFXG_ERR_NOMEM (-5)    /* device or host allocation failed */

/* Window functions: include/utils.h:209-213 */
#define SFXG_WINDOW_NONE 0
#define SFXG_WINDOW_RECT 1
#define SFXG_WINDOW_COS 2
#define SFXG_WINDOW_HAMMING 3
#define SFXG_WINDOW_HANN 4

/* Sample decode tables. REFERENCE: {-7,-2,+2,+7} and {-5,+5}
 * (src/bit2float_worker.cc:4-10). CONFIG: {-3.3359,-1,+1,+3.3359} and {-1,+1}. */
#define SFXG_LUT_REFERENCE 0
#define SFXG_LUT_CONFIG 1

/* X-step arithmetic. FP32 is the reference's arithmetic and order
 * (src/correlation_core.cc:241-273). The others run the X-step as a
 * per-channel Hermitian rank-k update on MFMA with FP32 accumulation:
 * FP16 or BF16 operands, or FP16 operands split into hi + lo parts (three
 * products). Pulsar binning takes every precision (per-bin Gram tiles on MFMA).
 * With FP16 or BF16 at fft_size_correlation 1024, 2048, 4096, 8192 or 16384 the
 * channeliser's transform also runs on MFMA with FP16 operands and FP32
 * accumulation, fed an FP16 real series (unless the tuning table's chan_fp32
 * selects the FP32 transform);
 * the visibility error stays within the FP16 bar of 1e-3 relative to the
 * FP64 pipeline. */
#define SFXG_PREC_FP32 0
#define SFXG_PREC_FP16 1
#define SFXG_PREC_BF16 2
#define SFXG_PREC_SPLIT_FP16 3

/* Stages readable with sfxg_fetch_stage (verification of each stage). */
#define SFXG_STAGE_SAMPLES 0      /* Bit2float_worker output, float[n_samples] */
#define SFXG_STAGE_TIME_SERIES 1  /* fringe-stopped real series, float[n_samples]; SFXG_ERR_STATE
                                    unless keep_time_series held it as FP32 (or the two-kernel
                                    FP32 route ran) */
#define SFXG_STAGE_SPECTRA 2      /* float2[n_segments][N_c + 1]; with SFXG_PREC_FP16 / BF16 the
                                     spectra are kept in that type (scaled by a power of two) and
                                     are returned converted back to float2 */
#define SFXG_STAGE_ACCUMULATED 3  /* phase-centre / bin accumulators, float2[n_bins][B][N_c + 1] */
#define SFXG_STAGE_NORMALISED 4   /* normalised visibilities, float2[n_bins][B][N_c + 1] */
#define SFXG_STAGE_CHANNELS 5     /* the channels integration_write stores, float2[n_bins][B][nchan + 1]:
                                     the normalised visibilities when nchan == N_c, their lag
                                     re-channelisation when nchan < N_c */

typedef struct sfxg_ctx sfxg_ctx;

/* Correlation_parameters::Station_parameters (include/control_parameters.h:132-151). */
typedef struct {
  int32_t station_number;
  int32_t station_stream;
  int32_t sample_rate;
  int32_t bandwidth;
  int64_t channel_freq;
  char sideband;      /* 'U' or 'L' */
  char polarisation;  /* 'R' or 'L' */
  int16_t reserved0;
  int32_t bits_per_sample;
  double LO_offset;
  int32_t tsys_freq;
  int32_t reserved1;
} sfxg_station_parameters;

/* Correlation_parameters (include/control_parameters.h:117-182), one
 * integration slice of one subband. */
typedef struct {
  int64_t experiment_start;
  int64_t start_time;
  int64_t stop_time;
  int64_t integration_time;
  int64_t sub_integration_time;
  int32_t number_channels;       /* output channels: fft_size_correlation, or a power of two
                                    >= 2 below it (fft_size_correlation <= 16384), which
                                    integration_write re-channelises to (see sfxg_set_mask) */
  int32_t fft_size_delaycor;      /* 256 */
  int32_t fft_size_correlation;   /* N_c, a power of two in [256, 32768] */
  int32_t integration_nr;
  int32_t slice_nr;
  int32_t slice_offset;
  int32_t sample_rate;
  int32_t bandwidth;
  int64_t channel_freq;
  char sideband;      /* 'U' or 'L' */
  char polarisation;
  int16_t reserved0;
  int32_t frequency_nr;
  int32_t cross_polarize;
  int32_t reference_station;
  int32_t window;
  int32_t n_phase_centers;
  int32_t pulsar_binning;
  int32_t n_streams;
  const sfxg_station_parameters *station_streams;
} sfxg_correlation_parameters;

/* The delay model of one stream as Delay_table_akima evaluates it
 * (src/delay_table_akima.cc:114-166): a piecewise cubic on a uniform grid
 * starting at scan_begin, coefficients c0..c3 of each piece in seconds since
 * the piece start. Clock offset and rate are folded into the delay. */
typedef struct {
  int64_t scan_begin;
  double interval;
  int32_t n_pieces;
  int32_t reserved0;
  const double *delay;      /* 4 * n_pieces, seconds */
  const double *phase;      /* 4 * n_pieces, radians; NULL means 0 */
  const double *amplitude;  /* 4 * n_pieces; NULL means 1 */
} sfxg_delay_model;

/* Pulsar_parameters::Pulsar plus the polyco in use
 * (include/control_parameters.h:73-98). */
typedef struct {
  int32_t nbins;      /* on-pulse bins; an off-pulse bin 0 is added */
  int32_t n_coef;
  double gate_start;
  double gate_stop;
  double tmid;
  double DM;
  double ref_phase;
  double ref_freq;
  double obs_freq;
  const double *coef;
} sfxg_pulsar;

typedef struct {
  int32_t lut;        /* SFXG_LUT_* */
  int32_t precision;  /* SFXG_PREC_* */
  /* Nonzero: run delay correction and the channeliser as two kernels and
   * keep the FP32 real series in HBM for SFXG_STAGE_TIME_SERIES
   * (verification). Zero (default): the two kernels too, unless the tuning
   * table picks the fused F-engine (fe_fused, N_c <= 4096, FP32 and split
   * FP16), whose series stays in LDS; the MFMA channeliser of packed FP16 /
   * BF16 spectra (N_c 1024 to 16384) gets the series as FP16 (it rounds its
   * points to FP16 anyway). FP32 and split FP16 give identical results
   * either way on the same route; the fused F-engine differs from the two
   * kernels by FP32 rounding (another radix schedule); packed FP16 / BF16
   * agree within their rounding (the FP16 series' phasors come from the
   * hardware sine and cosine). */
  int32_t keep_time_series;
  /* Spectra budget in MiB of one chunk: a slice is processed in chunks of
   * whole sub-integrations whose spectra fit it (at least one
   * sub-integration per chunk). 0: 8192. The results do not depend on it
   * (the FP32 path gives the same bytes). */
  int32_t chunk_mib;
  /* Nonzero: the context is a phased-array node, as Correlator_node builds
   * Correlation_core_phased in place of Correlation_core
   * (src/correlator_node.cc:22-47): per sub-integration the spectra of all
   * streams are summed (src/correlation_core_phased.cc:116-126) and the
   * number_ffts_in_integration / number_ffts_in_sub_integration + 1 sums
   * are written as the slice's rows, each with the header of stream 0's
   * auto-correlation (:13-50, :84-98; integration_write). No normalisation.
   * Takes FP32, no pulsar binning (src/control_parameters.cc:161-166) and
   * number_channels == fft_size_correlation. Set before sfxg_set_parameters. */
  int32_t phased_array;
} sfxg_options;

/* Creates a context on HIP device `device`. Replaces the construction of
 * Correlator_node's DSP objects (src/correlator_node.cc:37-47,213-249). */
int sfxg_create(int device, sfxg_ctx **out);
void sfxg_destroy(sfxg_ctx *ctx);
const char *sfxg_last_error(const sfxg_ctx *ctx);
int sfxg_set_options(sfxg_ctx *ctx, const sfxg_options *opt);

/* Replaces Correlator_node::set_parameters fan-out
 * (src/correlator_node.cc:310-405): Correlation_core::set_parameters
 * (src/correlation_core.cc:100-131) or Correlation_core_pulsar::set_parameters
 * (src/correlation_core_pulsar.cc:12-99) when pulsar != NULL,
 * Delay_correction::set_parameters (src/delay_correction.cc:186-259) and
 * Bit2float_worker::set_new_parameters (src/bit2float_worker.cc:235-276).
 * models: one per stream. uvw: 3 doubles per stream (or NULL for zeros). */
int sfxg_set_parameters(sfxg_ctx *ctx, const sfxg_correlation_parameters *params,
                        const sfxg_delay_model *models, const double *uvw,
                        const sfxg_pulsar *pulsar);

/* One station's raw VDIF stream for sfxg_push_vdif (multi-thread VDIF: one
 * channel per thread, Input_node_parameters with n_tracks == 0,
 * src/vdif_reader.cc:202-222). */
typedef struct {
  int32_t frame_bytes;        /* frame length with its header (dataframe_length * 8) */
  int32_t channel;            /* the stream's channel: index into thread_ids */
  int32_t n_threads;          /* 0: the stream holds one thread, whose id comes from the data */
  int32_t ref_mjd;            /* the reader's reference day (VDIF_reader ref_jday) */
  const int32_t *thread_ids;  /* thread id of every channel of the station */
  int64_t offset;             /* Input_node_parameters::offset, ticks subtracted from frame times */
} sfxg_vdif_stream;

/* Takes stream `stream_idx` of the slice as raw VDIF frames that already
 * live in device memory (device_frames, n_bytes: whole frames covering the
 * slice). Replaces, for that stream, the input node's VDIF_reader
 * (src/vdif_reader.cc:77-190: fill-pattern frames, the invalid flag, thread
 * to channel, frame times), Channel_extractor_tasklet_VDIF
 * (src/channel_extractor_tasklet_vdif.cc:25-49) and Input_node_data_writer
 * (src/input_node_data_writer.cc:105-228: the integer delay in bytes, DELAY
 * and INVALID records, one frame per block), and then the HEADER stream
 * transfer: a kernel lifts the frames' header words, the host builds the
 * stream's records from them alone, and a kernel gathers the DATA payload
 * from the frames. No sample byte crosses PCIe. The frames must stay valid
 * until this call returns; the gathered payload belongs to the context. */
int sfxg_push_vdif(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                   const sfxg_vdif_stream *vdif);

/* The phase-cal side product of a VDIF stream: Input_node_data_writer::do_phasecal
 * (src/input_node_data_writer.cc:232-306) over the frames of stream
 * `stream_idx` (as sfxg_push_vdif takes them). Every raw sample adds the
 * reference's integer value ({-7,-2,2,7} or {-5,5}) to bin (position mod
 * P), P = sample_rate / 10 kHz, of its integration; integrations start at
 * frames whose start time is a multiple of integration_ticks, and an
 * integration is written once the next one starts. dst receives one message
 * per finished integration as the node sends it to the output node, in
 * MPI_Pack's layout on one host: station, frequency number, sideband (LSB
 * 0, USB 1), polarisation (R 0, L 1) as uint8, start and duration as int64
 * ticks, P as int32, then the P sums as int32. *written holds the bytes
 * needed (also when cap is too small, which returns SFXG_ERR_ARG). */
int sfxg_vdif_phasecal(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                       const sfxg_vdif_stream *vdif, int64_t integration_ticks, uint8_t *dst, size_t cap,
                       size_t *written);

/* sfxg_vdif_phasecal without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). The
 * frames' header words, the frame table and the sums live in one static
 * device buffer of SFXG_PHASECAL_SCRATCH_BYTES per GPU (held until the process
 * ends) instead of buffers allocated per call; the output is identical.
 * Limits: 16 bytes of header words per frame, and 16 bytes per valid frame
 * plus 4 * P bytes per integration, each within the buffer; a larger call
 * returns SFXG_ERR_ARG and says what it needed. Every call on a GPU shares
 * that GPU's buffer: calls that overlap in time on one GPU (from several
 * threads or contexts) race and must be serialised by the caller. Host-side
 * bookkeeping (the frame walk) uses the C++ heap. */
#define SFXG_PHASECAL_SCRATCH_BYTES (16u << 20)
int sfxg_vdif_phasecal_static(sfxg_ctx *ctx, int32_t stream_idx, const void *device_frames, size_t n_bytes,
                              const sfxg_vdif_stream *vdif, int64_t integration_ticks, uint8_t *dst, size_t cap,
                              size_t *written);

/* The phase centres after the pointing centre of a multi-phase-centre slice
 * (params->n_phase_centers > 1): replaces the further sources of each
 * stream's Delay_table_akima (src/delay_table_akima.cc:114-136) and their
 * uvw tables. Call after sfxg_set_parameters, with n_phase_centers equal to
 * the parameters'. models: (n_phase_centers - 1) x n_streams, centre-major
 * (models[(j - 1) * n_streams + i] is stream i towards centre j); the
 * pointing centre's are those sfxg_set_parameters took. uvw: n_phase_centers
 * x n_streams x 3 doubles (NULL for zeros), replacing the uvw of
 * sfxg_set_parameters. source_nr: n_phase_centers source numbers written at
 * the head of each centre's record (NULL for 0). At every sub-integration the
 * cross-correlations go to the pointing centre as they are and to centre j
 * uv-shifted (Correlation_core::sub_integration / uvshift,
 * src/correlation_core.cc:515-592); each centre is then normalised and
 * written as a record of its own (:52-66). FP32 only. */
int sfxg_set_phase_centers(sfxg_ctx *ctx, int32_t n_phase_centers, const sfxg_delay_model *models,
                           const double *uvw, const int32_t *source_nr);

/* Mask_parameters (include/control_parameters.h:107-114) of the job: replaces
 * the correlator node's receipt of MPI_TAG_MASK_PARAMETERS
 * (src/correlator_node_controller.cc:72-77), after which every slice's
 * Correlation_parameters points at them (:55-62). They apply to slices with
 * fft_size_correlation > number_channels, whose integration_write
 * re-channelises the normalised visibilities in the lag domain
 * (src/correlation_core.cc:417-440): normalize != 0 scales every bin to unit
 * magnitude; mask (n_mask = fft_size_correlation + 1 values, or NULL / 0 for
 * ones) multiplies the bins (create_mask, :811-824); window (n_window =
 * 2 * number_channels values, or NULL / 0 for the lag window derived from the
 * job's window function, create_window :761-809) weights the lags. The values
 * are copied; sizes are checked against the slice at sfxg_prepare /
 * sfxg_correlate. Call with (0, NULL, 0, NULL, 0) to return to the defaults. */
int sfxg_set_mask(sfxg_ctx *ctx, int32_t normalize, const double *mask, int32_t n_mask, const double *window,
                  int32_t n_window);

/* Hands the complete HEADER byte stream of stream `stream_idx` (index into
 * station_streams) for this slice to the context. Replaces the reader thread
 * filling Channel_circular_input_buffer and Bit2float_worker::do_task
 * (src/correlator_node_data_reader_tasklet.cc:28-109,
 * src/bit2float_worker.cc:35-197). The bytes are parsed on the host and the
 * sample payload is copied to the device. */
int sfxg_push_stream(sfxg_ctx *ctx, int32_t stream_idx, const uint8_t *bytes, size_t n);

/* Same as sfxg_push_stream, for a stream whose sample payload already lives
 * in device memory: `control` holds the HEADER records with the payload bytes
 * of every DATA record left out (the 3-byte DATA header stays), and
 * device_payload holds those payload bytes back to back. device_payload must
 * stay valid until sfxg_synchronize returns after sfxg_correlate. */
int sfxg_push_stream_elided(sfxg_ctx *ctx, int32_t stream_idx, const uint8_t *control, size_t n,
                            const void *device_payload, size_t payload_bytes);

/* Enqueues the slice on the context's HIP stream: unpack and integer shift,
 * Delay_correction (src/delay_correction.cc:21-184) and Correlation_core
 * (src/correlation_core.cc:20-73, :241-314; pulsar: src/correlation_core_pulsar.cc:101-238).
 * Returns once the work is enqueued. */
int sfxg_correlate(sfxg_ctx *ctx);
int sfxg_synchronize(sfxg_ctx *ctx);

/* The host half of sfxg_correlate: HEADER-derived tables, buffers and the
 * slice's uploads, without the device work, which a batch then runs
 * (sfxg_batch_launch). Same reference interfaces as sfxg_correlate. */
int sfxg_prepare(sfxg_ctx *ctx);

/* Batches: the device work of several prepared contexts (slices of any
 * shape: channel counts, station counts, bit depths, precisions, pulsar or
 * not) captured into one HIP graph, the contexts' work on up to n_branches
 * concurrent graph branches, and replayed with one launch. No reference
 * counterpart: Correlator_node correlates one slice at a time
 * (src/correlator_node.cc:263-296); a batch runs the slices of several
 * correlator nodes' jobs as one launch. The graph is captured at the first
 * launch and again whenever a member context was prepared since the last
 * capture. Results are read per context (sfxg_fetch_output etc.), which waits
 * for the batch. A context may not be destroyed while it belongs to a batch,
 * and sfxg_correlate refuses a batch member. */
typedef struct sfxg_batch sfxg_batch;
int sfxg_batch_create(int device, int32_t n_branches, sfxg_batch **out);
void sfxg_batch_destroy(sfxg_batch *b);
const char *sfxg_batch_last_error(const sfxg_batch *b);
int sfxg_batch_add(sfxg_batch *b, sfxg_ctx *ctx);
int sfxg_batch_clear(sfxg_batch *b);
/* Asynchronous; waits for the previous launch of the same batch first. */
int sfxg_batch_launch(sfxg_batch *b);
int sfxg_batch_synchronize(sfxg_batch *b);

/* Size in bytes of the slice's output: the bytes Correlation_core::integration_write
 * (src/correlation_core.cc:316-477) passes to Data_writer::put_bytes, one
 * record per bin (pulsar) or phase centre. Equals the slice_size negotiated by
 * src/correlator_node.cc:386-398. */
int sfxg_output_size(sfxg_ctx *ctx, size_t *bytes);
/* Copies the output bytes (waits for the slice to finish). */
int sfxg_fetch_output(sfxg_ctx *ctx, uint8_t *dst, size_t cap, size_t *written);

/* bit_statistics::get_statistics / get_tsys of a stream
 * (src/bit_statistics.cc:33-91). Waits for the slice to finish. */
int sfxg_get_statistics(sfxg_ctx *ctx, int32_t stream_idx, int32_t levels[5], int32_t tsys[4]);

/* Correlation_core::number_of_baselines (include/correlation_core.h:50-52). */
int sfxg_number_of_baselines(sfxg_ctx *ctx, int32_t *n);

/* The tsys side product of the slice: Correlation_core::tsys_write
 * (src/correlation_core.cc:479-513) sends the output node one message per
 * stream; this writes their bytes back to back, 44 per stream in stream
 * order, in the layout MPI_Pack gives them on one host: station, frequency
 * number, sideband (LSB 0, USB 1) and polarisation (R 0, L 1) as uint8, the
 * slice start in ticks as uint64, then tsys on (low, high) and off (low,
 * high) counts as int64. */
int sfxg_fetch_tsys(sfxg_ctx *ctx, uint8_t *dst, size_t cap, size_t *written);

/* Copies one stage's device output (SFXG_STAGE_*) for verification; waits
 * for the slice. stream_idx selects the stream for per-stream stages. */
int sfxg_fetch_stage(sfxg_ctx *ctx, int32_t stage, int32_t stream_idx, void *dst, size_t cap,
                     size_t *written);

/* Average kernel durations per slice, in milliseconds, over the slices
 * correlated since the last reset, measured with HIP events on the context's
 * stream: ms[SFXG_KERNEL_*]. With the fused F-engine SFXG_KERNEL_DELAY covers
 * delay correction and channeliser and SFXG_KERNEL_CHANNELISE is 0. Waits for
 * the slice in flight. reset != 0 clears the sums afterwards. The reference's
 * counterpart is its per-component timing instrumentation
 * (src/correlation_core.cc:11-17, src/delay_correction.cc:12-18,
 * lib/common/src/monitor.h:80-173). n: capacity of ms. */
#define SFXG_KERNEL_STATS 0
#define SFXG_KERNEL_DELAY 1
#define SFXG_KERNEL_CHANNELISE 2
#define SFXG_KERNEL_XSTEP 3
#define SFXG_KERNEL_NORMALISE 4
#define SFXG_N_KERNELS 5
int sfxg_kernel_times(sfxg_ctx *ctx, double *ms, int32_t n, int32_t reset);
/* The chunks the prepared slice runs in (sfxg_options.chunk_mib): the delay,
 * channeliser and X-step kernels are launched once per chunk, so a kernel's
 * duration per launch is its sfxg_kernel_times figure / *n_chunks.
 * *max_segments: FFT segments per stream of the largest chunk. No counterpart
 * in the reference, which processes one FFT segment at a time
 * (src/delay_correction.cc:21-80). */
int sfxg_chunk_plan(const sfxg_ctx *ctx, int32_t *n_chunks, int64_t *max_segments);

/* Autotuner of launch shapes: on a prepared context (sfxg_prepare), times the
 * candidate FFT schedules of the F-engine (threads, i.e. points per thread and
 * radix stages, for N_c <= 4096, on the fused F-engine and on the delay stage
 * plus channeliser; workgroup count of the large-N channeliser)
 * and the X-step tile shapes (FP32 X-step: register tiles on 16- or 64-channel
 * blocks, or the per-baseline kernel and its waves per workgroup; channels
 * per workgroup of the MFMA X-step, workgroups of the MFMA pulsar X-step) on
 * the slice's own data, keeps the fastest of each in a process-wide table
 * keyed by the slice shape (channels, streams, precision, pulsar, fused), and
 * uses the table for every later slice of that shape. reps: timed runs per
 * candidate. report (may be NULL): text of the candidates' times. The trial
 * runs overwrite the slice's buffers: prepare or correlate it again after.
 * The reference's counterpart is the FFTW planner (src/sfxc_fft_float.cc:
 * 161,181-183, plans made with FFTW_ESTIMATE). Results change only by the
 * rounding of the FFT schedule. */
int sfxg_autotune(sfxg_ctx *ctx, int32_t reps, char *report, size_t cap);
/* The tuning table as text, one "shape|fe_threads|chan_wgs|xs_waves|mf_channels|
 * mfp_wgs|mf_ks|chan_fp32|fe_fused" line per shape (0: default; fields missing at the
 * end of a line (at least five given) are read as 0; xs_waves 1 and 2 force
 * the register tiles on 64- and 16-channel blocks, 4 and 8 the per-baseline
 * kernel; with pulsar binning in FP32, 2 forces the register tiles and 1, 4
 * or 8 the per-baseline kernel; chan_fp32 1 gives packed FP16 / BF16 spectra
 * at 1024 to 16384 channels the FP32 transform instead of the MFMA
 * one (up to 4096 its threads fe_threads sets; at 8192 and 16384
 * chan_wgs sets either channeliser's workgroups); fe_fused 1 runs the
 * fused F-engine up to 4096 channels when the series is not kept (FP32
 * spectra: FP32, split FP16, or chan_fp32 1), 0 the delay stage and the
 * channeliser as two kernels;
 * *written = bytes needed including the terminating zero. sfxg_tune_import adds such lines; sfxg_tune_clear
 * empties the table. */
int sfxg_tune_export(char *dst, size_t cap, size_t *written);
int sfxg_tune_import(const char *text);
void sfxg_tune_clear(void);

/* Synthetic 2-bit/1-bit raw station data written straight into device
 * memory (for throughput measurements on device-resident input): n_bytes
 * bytes of stream `stream_seed` starting at raw sample `first_sample`,
 * x = sqrt(rho) s[k - delay_samples] + sqrt(rho2) s2[k - delay2_samples]
 *     + sqrt(1 - rho - rho2) n_stream[k],
 * s (s2) shared by all streams with the same common_seed (common2_seed);
 * rho2 = 0 leaves the second component out. No reference counterpart. */
int sfxg_synthesize_raw(int device, void *device_dst, size_t n_bytes, int32_t bits_per_sample,
                        int64_t first_sample, int64_t delay_samples, uint64_t common_seed,
                        uint64_t stream_seed, float rho, float threshold, int64_t delay2_samples,
                        uint64_t common2_seed, float rho2);

#ifdef __cplusplus
}
#endif
#endif
