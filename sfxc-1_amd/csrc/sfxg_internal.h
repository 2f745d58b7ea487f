// Internal types shared by the host side (sfxg_host.cpp, sfxg_capi.cpp) and
// the device kernels (sfxg_kernels.hip).
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

namespace sfxg {

constexpr double kClockRate = 512000000.0;  // MAX_SAMPLE_RATE, sfxc/include/utils.h:36
constexpr int kHeaderData = 0, kHeaderInvalid = 1, kHeaderDelay = 2, kHeaderEndstream = 3;
constexpr int kCorrelatorBufferSize = 8192;  // sfxc/include/utils.h:50

// One run of consecutive output samples of a stream. src >= 0: output sample
// out_start + k is payload sample src + k; src < 0: the run is zero-filled
// (an inserted sample or invalid data).
struct Run {
  int64_t out_start;
  int64_t src;
  int64_t count;
};

// Payload bytes the reference's statistics count, in counting order:
// bytes [byte_start, byte_start + n_bytes) carry ordinals ordinal .. ordinal + n_bytes - 1.
struct CountRange {
  int64_t byte_start;
  int64_t n_bytes;
  int64_t ordinal;
};

// Bit2float_worker parameters for one slice (bit2float_worker.cc:235-313).
struct B2fParams {
  int bits_per_sample = 2;
  int sample_rate = 0;
  int base_sample_rate = 0;
  int tsys_freq = 80;
  int fft_size = 256;  // N_d
  int nfft_max = 32;
  int n_ffts_per_integration = 0;
  int delay_in_samples = 0;
};

// Result of parsing one stream's HEADER byte stream for one slice.
struct StreamPlan {
  std::vector<uint8_t> payload;       // payload bytes (empty when elided)
  const void *ext_payload = nullptr;  // caller-owned device payload (elided form)
  int64_t payload_bytes = 0;
  std::vector<Run> runs;
  std::vector<CountRange> counted;
  std::vector<std::pair<int32_t, int32_t>> invalid;  // Invalid{start, n_invalid}
  int64_t n_invalid = 0;                             // bit_statistics nInvalid
  std::vector<int32_t> chunks;                       // nfft of every output buffer
  int tsys_count0 = 0, tsys_on0 = 1, tsys_half = 0;  // tsys toggling in counted bytes
  bool ready = false;
};

// Parses a complete HEADER stream (Bit2float_worker::do_task restated at
// record granularity). Returns 0 or a negative SFXG_ERR_* code.
int parse_header_stream(const uint8_t *bytes, size_t n, bool elided, const B2fParams &p,
                        StreamPlan &out, std::string &err);

// Raw VDIF ingest (multi-thread VDIF, one channel per thread).
// One frame block as VDIF_reader::read_new_block delivers it to one channel
// (vdif_reader.cc:77-190): the payload's byte offset in the frame buffer, its
// start time in ticks and whether the frame is flagged invalid.
struct VdifBlock {
  int64_t payload_offset;
  int64_t time;
  bool invalid;
};
// A piece of DATA payload: bytes [src, src + n) of the frame buffer.
struct VdifPiece {
  int64_t src;
  int64_t n;
};
// VDIF_reader over the frames' first four header words (4 per frame):
// fill-pattern frames skipped, thread ids mapped to channels (thread_ids[c]
// is channel c's thread; none given: the first valid frame's thread is
// channel 0), times from the header (get_current_time, vdif_reader.cc:59-73).
// Appends channel `channel`'s blocks. Returns 0 or a negative SFXG_ERR_*.
int vdif_blocks(const uint32_t *words, int64_t n_frames, int frame_bytes, int sample_rate, int ref_mjd,
                int64_t offset_ticks, const std::vector<int> &thread_ids, int channel,
                std::vector<VdifBlock> &out, int &frame_payload, std::string &err);
// Input_node_data_writer::do_task over one slice (input_node_data_writer.cc:105-228)
// on frame blocks: the HEADER records with the DATA payload left out
// (control) and the payload pieces in order. delays: (tick, integer delay
// in samples) from the slice start, as Input_node_tasklet::get_delays finds them.
int vdif_write_slice(const std::vector<VdifBlock> &blocks, int frame_payload, int bits, int sample_rate,
                     int64_t slice_start, int64_t nsamples, const std::vector<std::pair<int64_t, int64_t>> &delays,
                     std::vector<uint8_t> &control, std::vector<VdifPiece> &pieces, std::string &err);

// Device side of the VDIF ingest (sfxg_vdif.hip): header words of n_frames
// frames at a stride of frame_bytes; the pieces' bytes to dst + dst_off[i].
hipError_t launch_vdif_words(const void *frames, int64_t n_frames, int frame_bytes, uint32_t *words,
                             hipStream_t s);
hipError_t launch_vdif_gather(const void *frames, const VdifPiece *pieces, const int64_t *dst_off, int n_pieces,
                              void *dst, hipStream_t s);
// Phase-cal accumulation: per valid frame its payload offset, the fold
// position of its first sample and its integration; sums[integration][P]
// (zeroed by the caller) gain the reference's integer sample values.
struct PcalFrame {
  int64_t payload_offset;
  int32_t count0;
  int32_t integration;
};
hipError_t launch_phasecal(const void *frames, const PcalFrame *table, int n_frames, int samples_per_frame, int P,
                           int bits, int32_t *sums, hipStream_t s);

// Number of correlation segments Delay_correction pushes per bit2float
// chunk (delay_correction.cc:24-36,58-69).
std::vector<int32_t> segments_per_chunk(const std::vector<int32_t> &chunks, int fft_size_delaycor,
                                        int64_t fft_rot_size, bool windowed);

// Segment counts after which Correlation_core folds its accumulators into the
// phase centre (correlation_core.cc:42-72).
std::vector<int32_t> fold_points(const std::vector<int32_t> &seg_per_chunk, int n_integration,
                                 int n_sub_integration);

// find_invalid for one baseline (correlation_core.cc:599-665).
void find_invalid(const std::vector<std::pair<int32_t, int32_t>> &inv0,
                  const std::vector<std::pair<int32_t, int32_t>> &inv1, double scale0,
                  double scale1, int64_t nflagged[2]);

// bit_statistics::get_statistics / get_tsys (bit_statistics.cc:33-91).
void bit_statistics(int bits, int64_t sample_rate, int64_t base_sample_rate, const int32_t *on,
                    const int32_t *off, int64_t n_invalid, int32_t levels[5], int32_t tsys[4]);

// ---------------------------------------------------------------------------
// device-side parameter blocks
// ---------------------------------------------------------------------------
struct DevModel {
  int64_t scan_begin;
  double interval;
  int32_t n;
  int32_t has_phase;
  int32_t has_amp;
  int32_t pad;
  const double *delay;
  const double *phase;
  const double *amp;
};

// FP64 constants of one N_d block of Delay_correction.
struct BlockConst {
  double c_term, l_term;  // fractional-delay ramp: phase c_term + k * l_term on bin k
  double phi, dphi;       // fringe phase at the block start and per-sample step
  float amp;
  float ramp_step[2];     // cos, sin of 64 l_term
  float fringe_step[2];   // cos, sin of 64 dphi
  float ramp_step16[2];   // cos, sin of 16 l_term
  float fringe_step16[2]; // cos, sin of 16 dphi
  float pad;
};

struct DevStream {
  const uint8_t *payload;
  const Run *runs;
  int64_t *block_src;        // per N_d block: the payload sample of its first sample where the
                             // block lies in one payload run, else ~(index of the run holding it)
  const CountRange *counted;
  int64_t n_runs;
  int64_t n_counted;
  int32_t bits;
  int32_t tsys_count0, tsys_on0, tsys_half;
  float lut[8];  // decode table (4 entries for 2 bit, 2 for 1 bit)
  DevModel model;
  // Delay_correction constants (delay_correction.cc:139-205)
  double center_freq;
  double mult_factor_phi;
  double start_phase;
  double LO_offset;
  double fs;         // station sample rate
  double bandwidth;
  double fl_sec;
  int64_t start_ticks;
  int64_t tps;       // ticks per sample
  int64_t fl;        // fft_length in ticks
  int64_t fl_half;   // fft_length / 2 in ticks
  int32_t oversamp;
  int32_t flip;      // sideband flip for the channeliser
  float *time_series;
  float2 *spectra;   // [n_segments][spectra_stride] float2, or packed 32-bit pairs (DevJob::spec_prec)
  float *samples;    // optional unpack output (verification)
  int32_t *hist;     // [2][256] byte histograms (tsys on, off)
  BlockConst *block_consts;  // [n_blocks]
};

// A slice is processed in chunks of whole sub-integrations; the F-engine
// kernels see one chunk: segments seg_base .. seg_base + n_segments - 1 of the
// slice, whose spectra rows (and, unfused, real series from block blk_base)
// start at the buffers' beginning.
struct DevJob {
  int32_t n_streams;
  int32_t nd;               // fft_size_delaycor
  int32_t nc;               // fft_size_correlation
  int32_t window;
  int32_t spec_prec;        // spectra stored as float2 (0) or packed FP16 / BF16 pairs (SFXG_PREC_*)
  float spec_scale;         // power-of-two scale of packed spectra
  int32_t ts_half;          // real series stored as FP16 (k_delay out, MFMA channeliser in), else FP32
  int64_t n_blocks;         // N_d blocks per stream in the slice
  int64_t n_samples;        // n_blocks * nd
  int64_t n_segments;       // segments of this chunk
  int64_t spectra_stride;   // complex elements per segment row
  int64_t seg_step;         // fft_rot_size / 2 real samples between segments
  int64_t seg_base;         // first segment of the chunk
  int64_t blk_base;         // first block of the chunk's real series (seg_base * nc / nd)
  int64_t blk_end;          // end of the chunk's blocks (k_delay)
};

// Launch shapes the autotuner chooses between; 0 selects the default.
struct Tune {
  int fe_threads = 0;   // F-engine threads per workgroup for N_c <= 4096 (fused and
                        // unfused channeliser alike: N_c / threads points per thread)
  int chan_wgs = 0;     // large-N channeliser (N_c 8192, 16384): workgroups in all
  int xs_waves = 0;     // plain X-step: 0 register tiles (16-channel blocks holding every
                        // tile when there are many tiles, else 64-channel blocks), 1 register
                        // tiles on 64-channel blocks, 2 on 16-channel blocks, else the
                        // per-baseline kernel with this many waves per workgroup
  int mf_channels = 0;  // MFMA X-step: at most this many channels per workgroup
  int mf_ks = 0;        // MFMA X-step: segments per LDS stage (16 or 32)
  int mfp_wgs = 0;      // MFMA pulsar X-step: 0 by bin, else the time-sliced kernel's workgroups
  int chan_fp32 = 0;    // packed FP16 / BF16 spectra at N_c = 16384: 0 transform on MFMA (FP16
                        // operands), 1 the FP32 transform of the other precisions
  int fe_fused = 0;     // N_c <= 4096 with an FP32 real series that is not kept: 1 the fused
                        // F-engine (series in LDS), 0 the delay stage and the channeliser as
                        // two kernels (series through HBM)
};
// Candidates of each knob for a shape (the default first).
std::vector<int> fe_thread_candidates(int nc);
std::vector<int> chan_wgs_candidates(int nc);
std::vector<int> chan_mfma_wgs_candidates();
std::vector<int> xs_waves_candidates();
// whether the plain X-step takes k_xstep_tile16 by default for this many tiles
bool xstep_tile16_default(int n_tiles);
std::vector<int> mf_channel_candidates(int precision, int M);
std::vector<int> mfp_wgs_candidates();

// kernel launchers (sfxg_kernels.hip)
hipError_t launch_unpack(const DevStream *d_streams, int n_streams, int64_t n_samples, hipStream_t s);
hipError_t launch_stats(const DevStream *d_streams, int n_streams, int64_t max_counted_bytes,
                        hipStream_t s);
// Per-block delay constants and run indices of the whole slice (once, before
// the first chunk's F-engine).
hipError_t launch_block_consts(const DevStream *d_streams, const DevJob &job, hipStream_t s);
hipError_t launch_delay(const DevStream *d_streams, const DevJob &job, const float2 *tw_nd,
                        hipStream_t s);
// Fused unpack + delay correction + channeliser (the real series stays in
// LDS) for N_c <= 4096.
bool fengine_supported(int nc);
hipError_t launch_fengine(const DevStream *d_streams, const DevJob &job, const float *window,
                          const float2 *tw_nd, const float2 *tw_split, const Tune &t, hipStream_t s);
// tw_split: [2 N_c / 128] coarse then [128] fine twiddles of exp(-2 pi i j / 2 N_c)
// (the LDS tables of the large-N channeliser).
// N_c = 32768 goes through HBM scratch of channelise_scratch_elems(nc) float2
// (0 for the LDS-resident sizes).
size_t channelise_scratch_elems(int nc);
// whether packed spectra of this precision at this N_c have the MFMA channeliser
bool chan_mfma_supported(int nc, int spec_prec);
hipError_t launch_channelise(const DevStream *d_streams, const DevJob &job, const float *window,
                             const float2 *tw_n, const float2 *tw_2n, const float2 *tw_split, float2 *scratch,
                             const Tune &t, hipStream_t s);
// Sub-integration partial sums (fold_start[0..n_k]: segment ranges) and
// their in-order fold into vis (vis == nullptr: the partials only).
// spectra rows hold segments seg_base ..; fold_start holds slice segment indices.
hipError_t launch_xstep(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base,
                        int n_ch, int M, const int32_t *fold_start, int n_k, const int32_t *bl_first,
                        const int32_t *bl_second, int n_baselines, float2 *partial, float2 *vis, const Tune &t,
                        hipStream_t s);
// Mixed-precision X-step (SFXG_PREC_FP16 / BF16 / SPLIT_FP16): per-channel
// Gram matrices on MFMA per sub-integration, formed into baselines in
// partial (n_k * n_baselines * n_ch float2) and combined into vis in fold order.
// Register-tiled FP32 X-step: tiles from xstep_tile_table (2 + 16 ints per
// tile: first streams a0, b0 and the baseline of each of the 4 x 4 pairs).
void xstep_tile_table(int M, const std::vector<int32_t> &bl_first, const std::vector<int32_t> &bl_second,
                      std::vector<int32_t> &table);
hipError_t launch_xstep_tile(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base,
                             int n_ch, int M, const int32_t *fold_start, int n_k, const int32_t *tiles, int n_tiles,
                             const int32_t *bl_first, const int32_t *bl_second, int n_baselines, float2 *partial,
                             float2 *vis, bool tile16, hipStream_t s);
// half_spectra: spectra packed in the operand type (FP16 / BF16 only).
// Whether the direct-to-LDS X-step takes split FP16 hi / lo planes for M
// streams with spectra rows stream_stride bins apart.
bool xstep_split_planes_fit(int M, int64_t stream_stride);
hipError_t launch_xstep_mfma(int precision, const void *spectra, bool half_spectra, int64_t stream_stride,
                             int64_t seg_stride,
                             int64_t seg_base, int n_ch, int M, const int32_t *fold_start, int n_k,
                             const int32_t *bl_first, const int32_t *bl_second, int n_baselines, float scale,
                             float2 *partial, float2 *vis, const Tune &t, hipStream_t s);
// segments seg_base .. seg_base + n_segments - 1 (spectra rows from seg_base),
// continuing the bins' sums in acc.
hipError_t launch_xstep_pulsar(const float2 *spectra, int64_t stream_stride, int64_t seg_stride,
                               int n_ch, int64_t seg_base, int64_t n_segments, const int32_t *bl_first,
                               const int32_t *bl_second, int n_baselines, const double *offsets,
                               double start_phase, double phase_coef, double fft_duration,
                               double gate_begin, double gate_end, int nbins, int16_t *bins, uint32_t *masks,
                               float2 *acc, const int32_t *tiles, int n_tiles, int M, bool by_bin, hipStream_t s);
// Mixed-precision X-step with pulsar binning for channels ch_base ..
// ch_base + n_ch_run - 1 of n_ch (offsets indexed by channel): binned Gram
// tiles in gram (pulsar_gram_floats(M, n_ch_run, nbins) floats, zeroed here),
// then vis[bin][b][ch] for those channels.
size_t pulsar_gram_floats(int M, int n_ch, int nbins);
// In two calls: zero the binned Gram tiles (pulsar_gram_zero, before the
// slice's first chunk), accumulate segments seg_base .. seg_base + n_segments
// - 1 of each chunk (launch_xstep_mfma_pulsar), form the baselines after the
// last chunk (pulsar_gram_baselines).
hipError_t pulsar_gram_zero(int M, int n_ch_run, int nbins, float *gram, hipStream_t s);
hipError_t launch_xstep_mfma_pulsar(int precision, const void *spectra, bool half_spectra, int64_t stream_stride,
                                    int64_t seg_stride,
                                    int ch_base, int n_ch_run, int M, int64_t seg_base, int64_t n_segments,
                                    float scale, const double *offsets, double start_phase, double phase_coef,
                                    double fft_duration, double gate_begin, double gate_end, int nbins, float *gram,
                                    int16_t *bins, uint32_t *masks, const Tune &t, hipStream_t s);
// 32-bit words of stage bin masks the by-bin kernel needs for a chunk of
// n_segments (0: the shape does not fit it); its bin table takes
// n_segments x n_ch_run int16.
size_t pulsar_bin_mask_words(int precision, int M, int n_ch_run, int nbins, int64_t n_segments);
hipError_t pulsar_gram_baselines(const float *gram, int n_ch, int ch_base, int n_ch_run, int M, int nbins,
                                 const int32_t *bl_first, const int32_t *bl_second, int n_baselines, float scale,
                                 float2 *vis, hipStream_t s);
// Zeroes `bytes` (a multiple of 4) with a kernel (no memset node in graphs).
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_auto_norms(const float2 *vis, int n_streams, int n_ch, int fft_size, int n_bins,
                             int n_baselines, double *norms, hipStream_t s);
hipError_t launch_normalise(const float2 *vis, const double *norms, const int32_t *bl_first,
                            const int32_t *bl_second, const double *flag_factor, int n_streams,
                            int n_baselines, int n_ch, int n_bins, float2 *out, hipStream_t s);
// Lag re-channelisation (correlation_core.cc:417-440, sfxg_rechan.hip): rows
// of N_c + 1 normalised bins in, rows of nchan + 1 bins out. tw holds
// exp(-2 pi i k / 2N_c) for k < 2N_c; mask N_c + 1 values; window 2 nchan.
bool rechannel_supported(int nc, int nchan);
hipError_t launch_rechannel(const float2 *in, float2 *out, int rows, int nc, int nchan, const float *mask,
                            const float *window, int normalize, const float2 *tw, hipStream_t s);
// Phased-array core (sfxg_phased.hip, correlation_core_phased.cc:13-126).
// A block is up to kPhasedBlock consecutive segments [seg0, seg1) of one
// row (absolute segment indices); a row entry lists the row's blocks
// [b0, b1) within its chunk's block table.
constexpr int kPhasedBlock = 64;
struct PhasedBlock {
  int64_t seg0, seg1;
};
struct PhasedRow {
  int32_t row, b0, b1, pad;
};
// Appends the blocks and rows of chunk [seg0, seg1) (rows of n_sub segments);
// block indices in the row entries count from the chunk's first block.
void phased_tables(int64_t seg0, int64_t seg1, int n_sub, std::vector<PhasedBlock> &blocks,
                   std::vector<PhasedRow> &rows);
// Adds the chunk's spectra (spectra[stream][seg - seg_base][ch]) to out[row][ch].
hipError_t launch_phased(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base,
                         int n_ch, int M, const PhasedBlock *blocks, int n_blocks, const PhasedRow *rows, int n_rows,
                         float2 *partial, float2 *out, hipStream_t s);
// Extra phase centres (sfxg_pc.hip, correlation_core.cc:515-592): the uv
// shift of one sub-integration, centre and baseline.
struct UvShift {
  double phi, delta;  // rotation exp(i (phi + ch * delta))
  float amp;          // 1 / weights[lag] or 1
  int32_t pad;
};
// vis[0] += partial[k] and vis[j] += shifted partial[k] for k = 0 .. n_k - 1
// in fold order; uv holds [k][j - 1][baseline]; vis is [centre][baseline][ch].
hipError_t launch_fold_combine_pc(const float2 *partial, int n_k, int n_baselines, int n_ch, int M,
                                  const UvShift *uv, int n_pc, float2 *vis, hipStream_t s);
// create_window of Correlation_core (correlation_core.cc:668-710,761-809):
// the 2 nchan-point lag window for the delay-correction window `window`
// applied over 2 fft_size points. Returns false for an unknown window.
bool lag_window(int window, int nchan, int fft_size, std::vector<float> &w);
hipError_t launch_synthesize(void *dst, size_t n_bytes, int bits, int64_t first_sample, int64_t delay_samples,
                             uint64_t common_seed, uint64_t stream_seed, float rho, float threshold,
                             int64_t delay2_samples, uint64_t common2_seed, float rho2, hipStream_t s);

}  // namespace sfxg
