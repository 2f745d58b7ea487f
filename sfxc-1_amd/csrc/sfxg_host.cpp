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

void find_invalid(const std::vector<std::pair<int32_t, int32_t>> &inv0,
                  const std::vector<std::pair<int32_t, int32_t>> &inv1, double scale0,
                  double scale1, int64_t nflagged_out[2]) {
  const std::vector<std::pair<int32_t, int32_t>> *inv[2] = {&inv0, &inv1};
  const double scale[2] = {scale0, scale1};
  int index[2] = {0, 0};
  int nflagged[2] = {0, 0};
  int st[2] = {INT_MAX, INT_MAX}, en[2] = {INT_MAX, INT_MAX};
  const int size[2] = {(int)inv0.size(), (int)inv1.size()};
  for (int i = 0; i < 2; i++) {
    if (size[i] > 0) {
      st[i] = (int)floor((*inv[i])[0].first / scale[i]);
      en[i] = st[i] + (int)floor((*inv[i])[index[i]].second / scale[i]);
      index[i]++;
    }
  }
  int i = st[0] < st[1] ? 0 : 1;
  while (st[i] < INT_MAX) {
    if (st[0] == st[1]) {
      if (en[0] == en[1]) {
        st[0] = en[0] + 1;
        st[1] = en[1] + 1;
      } else {
        const int j = en[0] < en[1] ? 0 : 1;
        st[1 - j] = en[j];
        st[j] = en[j] + 1;
      }
    } else if (en[i] < st[1 - i]) {
      nflagged[1 - i] += en[i] - st[i];
      st[i] = en[i] + 1;
    } else {
      nflagged[1 - i] += st[1 - i] - st[i];
      st[i] = st[1 - i];
    }
    for (int j = 0; j < 2; j++) {
      if (st[j] > en[j]) {
        if (index[j] < size[j]) {
          st[j] = (int)floor((*inv[j])[index[j]].first / scale[j]);
          en[j] = st[j] + (int)floor((*inv[j])[index[j]].second / scale[j]);
          index[j]++;
        } else {
          st[j] = INT_MAX;
          en[j] = INT_MAX;
        }
      }
    }
    i = st[0] < st[1] ? 0 : 1;
  }
  nflagged_out[0] += nflagged[0];
  nflagged_out[1] += nflagged[1];
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

}  // namespace sfxg
