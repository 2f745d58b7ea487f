// Host-side restatements that turn the reference's byte stream and chunking
// rules into the flat tables the device kernels consume.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <climits>

#include "sfxg_internal.h"
#include "../../include/sfxg.h"

namespace sfxg {

namespace {

void add_run(std::vector<Run> &runs, int64_t out_start, int64_t src, int64_t count) {
  if (count <= 0) return;
  if (!runs.empty()) {
    Run &r = runs.back();
    bool contiguous_out = r.out_start + r.count == out_start;
    bool contiguous_src = (src < 0 && r.src < 0) || (src >= 0 && r.src >= 0 && r.src + r.count == src);
    if (contiguous_out && contiguous_src) {
      r.count += count;
      return;
    }
  }
  runs.push_back(Run{out_start, src, count});
}

void add_counted(std::vector<CountRange> &c, int64_t byte_start, int64_t n, int64_t &ordinal) {
  if (n <= 0) return;
  if (!c.empty() && c.back().byte_start + c.back().n_bytes == byte_start) {
    c.back().n_bytes += n;
  } else {
    c.push_back(CountRange{byte_start, n, ordinal});
  }
  ordinal += n;
}

}  // namespace

// Bit2float_worker::do_task (sfxc/src/bit2float_worker.cc:35-197) and
// bit2float (:326-424) restated per record: instead of decoding samples, each
// SEND_DATA step records the run of payload samples it would copy and the
// bytes whose histogram it would increment. set_parameters (:278-313) gives
// the tsys phase.
int parse_header_stream(const uint8_t *bytes, size_t n, bool elided, const B2fParams &p,
                        StreamPlan &out, std::string &err) {
  out = StreamPlan();
  if (p.bits_per_sample != 1 && p.bits_per_sample != 2) {
    err = "bits_per_sample must be 1 or 2";
    return SFXG_ERR_ARG;
  }
  if (p.fft_size <= 0 || p.nfft_max <= 0 || p.n_ffts_per_integration <= 0) {
    err = "invalid slice size";
    return SFXG_ERR_ARG;
  }
  const int spb = 8 / p.bits_per_sample;
  if (p.bits_per_sample == 2) {
    if (p.tsys_freq <= 0 || p.sample_rate / (2 * p.tsys_freq) <= 0) {
      err = "invalid tsys_freq";
      return SFXG_ERR_ARG;
    }
    const int full = p.sample_rate / p.tsys_freq, half = p.sample_rate / (2 * p.tsys_freq);
    int tc = (int)floor(p.delay_in_samples + 0.5);
    tc %= full;
    while (tc < 0) tc += full;
    out.tsys_on0 = tc < half;
    tc %= half;
    tc = half - tc;
    tc /= 4;
    out.tsys_count0 = tc;
    out.tsys_half = half / 4;
  }
  enum { IDLE, SEND_INVALID, SEND_DATA } state = IDLE;
  size_t bytes_left = 0, invalid_left = 0, sample_in_byte = 0;
  int cur_delay = -1;
  int64_t read = 0;       // position in the HEADER stream
  int64_t pay = 0;        // payload index of the byte the next sample comes from
  int64_t pay_end = 0;    // payload bytes announced so far
  int64_t ordinal = 0;
  int current_fft = 0;
  while (current_fft < p.n_ffts_per_integration) {
    const int nfft = std::min(p.nfft_max, p.n_ffts_per_integration - current_fft);
    const size_t out_size = (size_t)nfft * p.fft_size;
    const int64_t base = (int64_t)current_fft * p.fft_size;
    size_t out_index = 0;
    while (out_index != out_size) {
      if (state == IDLE) {
        if ((int64_t)n - read < 3) {
          err = "HEADER stream ends before the slice is complete";
          return SFXG_ERR_STREAM;
        }
        const uint8_t header = bytes[read++];
        if (header == kHeaderData) {
          bytes_left = (size_t)bytes[read] | ((size_t)bytes[read + 1] << 8);
          read += 2;
          if (bytes_left == 0) {
            err = "DATA record of length 0";
            return SFXG_ERR_STREAM;
          }
          if (!elided) {
            if ((int64_t)n - read < (int64_t)bytes_left) {
              err = "HEADER stream ends inside a DATA record";
              return SFXG_ERR_STREAM;
            }
            out.payload.insert(out.payload.end(), bytes + read, bytes + read + bytes_left);
            read += (int64_t)bytes_left;
          }
          pay_end += (int64_t)bytes_left;
          state = SEND_DATA;
        } else if (header == kHeaderDelay) {
          const int new_delay = (int8_t)bytes[read++];
          if (cur_delay < 0) {
            if (new_delay < 0 || new_delay >= spb) {
              err = "initial DELAY outside one byte";
              return SFXG_ERR_STREAM;
            }
            sample_in_byte = (size_t)new_delay;
          } else {
            const int diff = new_delay - cur_delay;
            if (diff == -1 || diff > 1) {
              add_run(out.runs, base + (int64_t)out_index, -1, 1);  // insert a zero sample
              out_index++;
            } else {
              if (sample_in_byte != 0) {
                err = "sample removal while a partial byte is pending";
                return SFXG_ERR_STREAM;
              }
              sample_in_byte = 1;  // remove a sample
            }
          }
          cur_delay = (int8_t)new_delay;
        } else if (header == kHeaderInvalid) {
          invalid_left = (size_t)bytes[read] | ((size_t)bytes[read + 1] << 8);
          read += 2;
          out.n_invalid += (int64_t)invalid_left;
          out.invalid.push_back({(int32_t)(base + (int64_t)out_index), (int32_t)invalid_left});
          state = SEND_INVALID;
        } else {
          err = "invalid header byte in HEADER stream";
          return SFXG_ERR_STREAM;
        }
      } else if (state == SEND_INVALID) {
        const size_t free_ = out_size - out_index;
        size_t to_write = std::min(free_, invalid_left);
        if (sample_in_byte > 0) {
          const size_t drop = std::min(sample_in_byte, to_write);
          sample_in_byte -= drop;
          to_write -= drop;
          invalid_left -= drop;
        }
        if (to_write > 0) {
          add_run(out.runs, base + (int64_t)out_index, -1, (int64_t)to_write);
          invalid_left -= to_write;
          out_index += to_write;
        }
        if (invalid_left == 0) state = IDLE;
      } else {  // SEND_DATA: one bit2float call
        const size_t free_ = out_size - out_index;
        size_t samp = bytes_left * spb - sample_in_byte;
        if (samp > free_) samp = free_;
        bytes_left -= (samp + sample_in_byte) / spb;
        const int start = (int)sample_in_byte;
        const int64_t p0 = pay;
        add_run(out.runs, base + (int64_t)out_index, p0 * spb + start, (int64_t)samp);
        add_counted(out.counted, p0, 1, ordinal);  // first byte is always counted
        const int first_take = (int)std::min<size_t>(samp, (size_t)(spb - start));
        int new_sib;
        if (first_take + start != spb) {
          new_sib = start + first_take;
        } else {
          pay = p0 + 1;
          int64_t rem = (int64_t)samp - first_take;
          const int64_t nbytes = rem / spb;
          add_counted(out.counted, pay, nbytes, ordinal);
          pay += nbytes;
          rem -= nbytes * spb;
          new_sib = (int)rem;  // a partial final byte is not counted
        }
        sample_in_byte = (size_t)new_sib;
        out_index += samp;
        if (bytes_left == 0) {
          if (sample_in_byte != 0) {
            err = "DATA record ends inside a byte";
            return SFXG_ERR_STREAM;
          }
          state = IDLE;
        }
      }
    }
    out.chunks.push_back(nfft);
    current_fft += nfft;
  }
  out.payload_bytes = elided ? pay_end : (int64_t)out.payload.size();
  (void)pay_end;
  out.ready = true;
  return SFXG_OK;
}

std::vector<int32_t> segments_per_chunk(const std::vector<int32_t> &chunks, int fft_size_delaycor,
                                        int64_t fft_rot_size, bool windowed) {
  std::vector<int32_t> out;
  int64_t tbuf_start = 0, tbuf_end = 0;
  const int64_t step = fft_rot_size / 2;
  for (int32_t nbuffer : chunks) {
    int64_t nfft_cor = ((int64_t)nbuffer * fft_size_delaycor + tbuf_end - tbuf_start) / step;
    if (windowed) nfft_cor -= 1;
    tbuf_end += (int64_t)nbuffer * fft_size_delaycor;
    if (nfft_cor > 0) tbuf_start += nfft_cor * step;
    out.push_back((int32_t)std::max<int64_t>(nfft_cor, 0));
  }
  return out;
}

std::vector<int32_t> fold_points(const std::vector<int32_t> &seg_per_chunk, int n_integration,
                                 int n_sub_integration) {
  std::vector<int32_t> folds;
  int64_t current = 0, next_sub = 1;
  for (int32_t n : seg_per_chunk) {
    if (n <= 0) continue;
    current += n;
    if (current == n_integration) {
      folds.push_back((int32_t)current);
    } else if (current >= next_sub * (int64_t)n_sub_integration) {
      folds.push_back((int32_t)current);
      next_sub++;
    }
  }
  return folds;
}

namespace {
// One stream's Invalid list as a queue of half-open runs [b, e) in job-rate
// samples; (b, e) is the unconsumed rest of the run at the head.
struct InvalidQueue {
  const std::vector<std::pair<int32_t, int32_t>> &runs;
  double scale;
  size_t next = 0;
  int64_t b = 0, e = 0;
  bool live = false;
  void pop() {
    live = next < runs.size();
    if (!live) return;
    b = (int64_t)floor(runs[next].first / scale);
    e = b + (int64_t)floor(runs[next].second / scale);
    next++;
  }
};
}  // namespace

// Counting rule: a sample at which one stream of the baseline is invalid and
// the other is not is flagged for the other stream. For Invalid lists whose
// runs are disjoint and increasing, nflagged[0] = |mask_1 \ mask_0| and
// nflagged[1] = |mask_0 \ mask_1| (a brute-force per-sample test pins this).
// The lists are consumed as two queues: when the heads start together their
// common part is skipped and the head that ends first (both, if they end
// together) is used up; otherwise the head that starts first gives the
// samples before the other head's start (all of them when it ends first) to
// the other stream. A stream's own runs can overlap by the samples that a
// pending sample removal drops at the start of an INVALID record (the run
// keeps the record's length, bit2float_worker.cc:92,105-112); each run is then
// compared only with the other queue's head, which is how
// Correlation_core::find_invalid (correlation_core.cc:599-665) counts them.
void find_invalid(const std::vector<std::pair<int32_t, int32_t>> &inv0,
                  const std::vector<std::pair<int32_t, int32_t>> &inv1, double scale0,
                  double scale1, int64_t nflagged_out[2]) {
  InvalidQueue q[2] = {{inv0, scale0}, {inv1, scale1}};
  q[0].pop();
  q[1].pop();
  int64_t flagged[2] = {0, 0};
  while (q[0].live || q[1].live) {
    if (q[0].live && q[1].live && q[0].b == q[1].b) {
      const int64_t end = std::min(q[0].e, q[1].e);
      const bool used[2] = {q[0].e == end, q[1].e == end};
      for (int k = 0; k < 2; k++) {
        if (used[k]) q[k].pop();
        else q[k].b = end;
      }
      continue;
    }
    const int first = (!q[1].live || (q[0].live && q[0].b < q[1].b)) ? 0 : 1;
    InvalidQueue &a = q[first], &o = q[1 - first];
    if (!o.live || a.e < o.b) {
      flagged[1 - first] += a.e - a.b;
      a.pop();
    } else {
      flagged[1 - first] += o.b - a.b;
      a.b = o.b;
    }
  }
  nflagged_out[0] += flagged[0];
  nflagged_out[1] += flagged[1];
}

void bit_statistics(int bits, int64_t sample_rate, int64_t base_sample_rate, const int32_t *on,
                    const int32_t *off, int64_t n_invalid, int32_t levels[5], int32_t tsys[4]) {
  int stat[5] = {0, 0, 0, 0, 0};
  int ton[4] = {0, 0, 0, 0}, toff[4] = {0, 0, 0, 0};
  if (bits == 2) {
    for (int b = 0; b < 256; b++) {
      for (int k = 0; k < 4; k++) {
        const int c = (b >> (2 * k)) & 3;
        stat[c] += on[b] + off[b];
        ton[c] += on[b];
        toff[c] += off[b];
      }
    }
    tsys[0] = ton[1] + ton[2];
    tsys[1] = ton[0] + ton[3];
    tsys[2] = toff[1] + toff[2];
    tsys[3] = toff[0] + toff[3];
  } else {
    for (int b = 0; b < 256; b++)
      for (int k = 0; k < 8; k++) stat[(b >> k) & 1] += on[b] + off[b];
    tsys[0] = tsys[1] = tsys[2] = tsys[3] = 0;
  }
  stat[4] += (int)n_invalid;
  for (int k = 0; k < 5; k++) levels[k] = (int32_t)((base_sample_rate * (int64_t)stat[k]) / sample_rate);
}

// create_window (correlation_core.cc:761-809). The window functions of
// correlation_core.cc:668-692 are tabulated once per length, and each
// self-convolution (convolve, correlation_core.cc:694-710) adds the same
// products in the same order as the reference's loop.
bool lag_window(int window, int nchan, int fft_size, std::vector<float> &w) {
  const int n = 2 * nchan, m = 2 * fft_size;
  w.assign(n, 0.0f);
  if (window == SFXG_WINDOW_NONE || window == SFXG_WINDOW_RECT) {
    w[0] = 1.0f;
    for (int i = 1; i < n / 2; i++) w[i] = w[n - i] = (float)((1. - 2. * i / n) / (1. - 2. * i / m));
    w[n / 2] = 0.0f;
    return true;
  }
  if (window != SFXG_WINDOW_COS && window != SFXG_WINDOW_HAMMING && window != SFXG_WINDOW_HANN) return false;
  auto table = [window](int len) {
    std::vector<double> f(len + 1);
    for (int i = 0; i <= len; i++) {
      if (window == SFXG_WINDOW_COS) f[i] = sin(M_PI * i / len);
      else if (window == SFXG_WINDOW_HAMMING) f[i] = 0.54 - 0.46 * cos(2 * M_PI * i / len);
      else f[i] = 0.5 * (1 - cos(2 * M_PI * i / len));
    }
    return f;
  };
  const std::vector<double> fn = table(n), fm = table(m);
  auto convolve = [](const std::vector<double> &f, int len, int i) {
    i += len / 2;
    double sum = 0.0;
    for (int j = std::max(0, i - len); j <= std::min(len, i); j++) sum += f[j] * f[i - j];
    return sum;
  };
  const int ratio = m / n;
  w[0] = (float)(ratio * convolve(fn, n, n / 2) / convolve(fm, m, m / 2));
  for (int i = 1; i < n / 2; i++)
    w[i] = w[n - i] = (float)(ratio * convolve(fn, n, n / 2 + i) / convolve(fm, m, m / 2 + i));
  w[n / 2] = (float)(ratio * convolve(fn, n, n) / convolve(fm, m, (n + m) / 2));
  return true;
}

// ---------------------------------------------------------------------------
// Raw VDIF ingest
// ---------------------------------------------------------------------------

namespace {
// mjd (utils.cc:255-263): Julian day number minus 2400000.5, truncated to int.
int mjd_of(int day, int month, int year) {
  const int a = (14 - month) / 12;
  const int y = year + 4800 - a;
  const int m = month + 12 * a - 3;
  const int jdn = day + ((153 * m + 2) / 5) + 365 * y + (y / 4) - (y / 100) + (y / 400) - 32045;
  return (int)(jdn - 2400000.5);
}
constexpr uint32_t kVdifFill = 0x11223344u;
constexpr int kShrtMax = 32767;
}  // namespace

int vdif_blocks(const uint32_t *words, int64_t n_frames, int frame_bytes, int sample_rate, int ref_mjd,
                int64_t offset_ticks, const std::vector<int> &thread_ids, int channel,
                std::vector<VdifBlock> &out, int &frame_payload, std::string &err) {
  // Header fields (vdif_reader.h:27-40, little-endian bit fields):
  // w0: seconds 0-29, legacy 30, invalid 31; w1: frame in second 0-23,
  // reference epoch 24-29; w2: length in 8-byte units 0-23, log2 channels
  // 24-28; w3: thread 16-25, bits per sample - 1 26-30.
  std::vector<int> thread_map_key(thread_ids), thread_map_val;
  for (size_t c = 0; c < thread_ids.size(); c++) thread_map_val.push_back((int)c);
  auto lookup = [&](int thread) {
    for (size_t k = 0; k < thread_map_key.size(); k++)
      if (thread_map_key[k] == thread) return thread_map_val[k];
    return -1;
  };
  bool first_seen = false;
  int legacy = 0, bits_m1 = 0, log2_nchan = 0;
  frame_payload = 0;
  for (int64_t f = 0; f < n_frames; f++) {
    const uint32_t *w = words + 4 * f;
    if (w[0] == kVdifFill || w[1] == kVdifFill || w[2] == kVdifFill || w[3] == kVdifFill) continue;
    if (!first_seen) {
      first_seen = true;
      legacy = (w[0] >> 30) & 1;
      bits_m1 = (w[3] >> 26) & 0x1f;
      log2_nchan = (w[2] >> 24) & 0x1f;
      frame_payload = frame_bytes - (16 + 16 * (1 - legacy));
      if ((int64_t)(w[2] & 0xffffff) * 8 != frame_bytes) {
        err = "VDIF frame length in the header differs from frame_bytes";
        return SFXG_ERR_STREAM;
      }
      if (frame_payload <= 0) {
        err = "VDIF frame without payload";
        return SFXG_ERR_STREAM;
      }
    }
    const bool invalid = (w[0] >> 31) != 0;
    const int thread = (int)((w[3] >> 16) & 0x3ff);
    int ch = lookup(thread);
    if (invalid) {
      if (ch < 0) ch = 0;
    } else if (ch < 0) {
      // the only thread: its id comes from the data
      if (!thread_map_key.empty()) continue;
      thread_map_key.push_back(thread);
      thread_map_val.push_back(0);
      ch = 0;
    }
    if (ch != channel) continue;
    // get_current_time: seconds since the reference day plus the frame's
    // place in its second, then Time::set_time
    const uint32_t sec = w[0] & 0x3fffffff;
    const int ref_epoch = (int)((w[1] >> 24) & 0x3f);
    const int jday_epoch = mjd_of(1, 1 + 6 * (ref_epoch % 2), 2000 + ref_epoch / 2);
    const double seconds_since_reference = (double)sec - (double)((ref_mjd - jday_epoch) * 24 * 60 * 60);
    double subsec = 0;
    if (sample_rate > 0) {
      const int samples_per_frame = 8 * frame_payload / ((bits_m1 + 1) * (1 << log2_nchan));
      subsec = (double)(w[1] & 0xffffff) * samples_per_frame / sample_rate;
    }
    const int64_t t = (int64_t)((ref_mjd - 51544) * 86400 * kClockRate) +
                      (int64_t)round((seconds_since_reference + subsec) * kClockRate) - offset_ticks;
    out.push_back({f * (int64_t)frame_bytes + (16 + 16 * (1 - legacy)), t, invalid});
  }
  if (!first_seen) {
    err = "no VDIF frames";
    return SFXG_ERR_STREAM;
  }
  return SFXG_OK;
}

int vdif_write_slice(const std::vector<VdifBlock> &blocks, int frame_payload, int bits, int sample_rate,
                     int64_t slice_start, int64_t nsamples, const std::vector<std::pair<int64_t, int64_t>> &delays,
                     std::vector<uint8_t> &control, std::vector<VdifPiece> &pieces, std::string &err) {
  const int spb = 8 / bits;
  const int64_t tps = (int64_t)round(kClockRate / sample_rate);  // Time::sample_rate, ticks per sample
  const int64_t byte_length = (int64_t)round(8 * 1000000. / ((double)sample_rate * bits) * (kClockRate / 1000000));
  // Input_node_tasklet::get_delay (input_node_tasklet.cc:45-63): bytes and remaining samples
  struct Delay {
    int64_t time, bytes;
    int remaining;
  };
  std::vector<Delay> dl;
  for (const auto &d : delays) {
    const int64_t n = d.second;
    int64_t nb = (int64_t)floor((double)(n - 1) / spb);
    int64_t rem = n - nb * spb;
    if (rem * bits == 8) {
      nb++;
      rem = 0;
    }
    dl.push_back({d.first, nb, (int)rem});
  }
  if (dl.empty()) {
    err = "no delays";
    return SFXG_ERR_ARG;
  }
  auto w_delay = [&](int r) {
    control.push_back(kHeaderDelay);
    control.push_back((uint8_t)(int8_t)r);
  };
  auto w_count = [&](uint8_t type, int64_t n) {
    int64_t done = 0;
    while (done < n) {
      const int k = (int)std::min<int64_t>(n - done, kShrtMax);
      control.push_back(type);
      control.push_back((uint8_t)(k & 0xff));
      control.push_back((uint8_t)(k >> 8));
      done += k;
    }
  };
  int64_t current = slice_start;
  size_t di = 0;
  while (di + 1 < dl.size() && dl[di + 1].time <= current + byte_length) di++;
  int64_t slice_size = nsamples;
  bool sync = true;
  const int64_t block_size = frame_payload;
  auto next_delay_pos = [&](int64_t start_time) -> int64_t {
    if (di + 1 < dl.size()) return (int64_t)((double)(dl[di + 1].time - start_time) / (double)byte_length);
    return INT_MAX / 2;
  };
  for (size_t b = 0; b < blocks.size(); b++) {
    const VdifBlock &blk = blocks[b];
    int64_t byte_offset = 0;
    if (sync) {
      const int64_t dsamples = (current - blk.time) / tps;  // Time::diff_samples
      byte_offset = dsamples * bits / 8 + dl[di].bytes;
      if (byte_offset < 0) {
        err = "the slice starts before the VDIF data";
        return SFXG_ERR_STREAM;
      }
      if (byte_offset >= block_size) continue;  // data before the slice
      slice_size += dl[di].remaining;
      w_delay(dl[di].remaining);
      current -= (int64_t)dl[di].remaining * tps;
      sync = false;
    }
    int64_t index = byte_offset;
    int64_t next_invalid_pos = blk.invalid ? 0 : block_size + 1;
    const int64_t n_invalid = blk.invalid ? block_size : 0;
    bool invalid_done = !blk.invalid;
    int64_t ndp = next_delay_pos(current) + byte_offset;
    int64_t total = std::min(block_size - byte_offset, (slice_size + spb - 1) / spb);
    int64_t end_index = total + byte_offset;
    while (index < end_index) {
      if (index >= ndp) {
        di++;
        w_delay(dl[di].remaining);
        const int d_delay = dl[di].remaining - dl[di - 1].remaining;
        if (d_delay > 1 || d_delay == -1) slice_size -= 1;
        else slice_size += 1;
        ndp = next_delay_pos(current) + byte_offset;
        total = std::min(block_size - byte_offset, (slice_size + spb - 1) / spb);
        end_index = total + byte_offset;
      } else if (index >= next_invalid_pos) {
        const int64_t end_pos = std::min(next_invalid_pos + n_invalid, ndp);
        int64_t nr = std::max<int64_t>(0, end_pos - index);
        nr = std::min(nr, end_index - index);
        if (nr > 0) {
          w_count(kHeaderInvalid, nr * spb);
          index += nr;
        }
        if (end_pos == next_invalid_pos + n_invalid && !invalid_done) {
          invalid_done = true;
          next_invalid_pos = block_size + 1;
        }
      } else {
        int64_t k = std::min(ndp - index, end_index - index);
        k = std::min(next_invalid_pos - index, k);
        if (k > 0) {
          int64_t done = 0;
          while (done < k) {  // write_data: records of at most SHRT_MAX bytes
            const int n = (int)std::min<int64_t>(k - done, kShrtMax);
            control.push_back(kHeaderData);
            control.push_back((uint8_t)(n & 0xff));
            control.push_back((uint8_t)(n >> 8));
            pieces.push_back({blk.payload_offset + index + done, n});
            done += n;
          }
        }
        index += k;
      }
    }
    slice_size -= total * spb;
    current += total * spb * tps;
    if (slice_size <= 0) {
      control.push_back(kHeaderEndstream);
      return SFXG_OK;
    }
  }
  err = "the VDIF data ends before the slice";
  return SFXG_ERR_STREAM;
}

}  // namespace sfxg
