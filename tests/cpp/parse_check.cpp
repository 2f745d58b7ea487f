// Test harness: exposes the host-side HEADER-stream parser and chunk
// bookkeeping of libsfxg (sfxc-1_amd/csrc/sfxg_host.cpp) to Python so they
// can be checked on the CPU. Built by tests/test_host_parse.py with g++.
#include <string.h>

#include <string>
#include <vector>

#include "../../sfxc-1_amd/csrc/sfxg_internal.h"

extern "C" {

// out_runs: 3 int64 per run (out_start, src, count); out_counted: 3 int64 per
// range; out_inv: 2 int32 per entry; out_chunks: nfft per chunk.
// sizes[0..4] = {n_runs, n_counted, n_inv, n_chunks, payload_bytes};
// tsys3 = {count0, on0, half}; returns the parser's status code.
int pc_parse(const uint8_t *bytes, size_t n, int elided, const int32_t *p8, int64_t *out_runs, int64_t max_runs,
             int64_t *out_counted, int64_t max_counted, int32_t *out_inv, int64_t max_inv, int32_t *out_chunks,
             int64_t max_chunks, int64_t *sizes, int32_t *tsys3, int64_t *n_invalid, char *err, int errlen) {
  sfxg::B2fParams p;
  p.bits_per_sample = p8[0];
  p.sample_rate = p8[1];
  p.base_sample_rate = p8[2];
  p.tsys_freq = p8[3];
  p.fft_size = p8[4];
  p.nfft_max = p8[5];
  p.n_ffts_per_integration = p8[6];
  p.delay_in_samples = p8[7];
  sfxg::StreamPlan plan;
  std::string e;
  int rc = sfxg::parse_header_stream(bytes, n, elided != 0, p, plan, e);
  strncpy(err, e.c_str(), errlen - 1);
  err[errlen - 1] = 0;
  if (rc != 0) return rc;
  if ((int64_t)plan.runs.size() > max_runs || (int64_t)plan.counted.size() > max_counted ||
      (int64_t)plan.invalid.size() > max_inv || (int64_t)plan.chunks.size() > max_chunks)
    return -100;
  for (size_t i = 0; i < plan.runs.size(); i++) {
    out_runs[3 * i] = plan.runs[i].out_start;
    out_runs[3 * i + 1] = plan.runs[i].src;
    out_runs[3 * i + 2] = plan.runs[i].count;
  }
  for (size_t i = 0; i < plan.counted.size(); i++) {
    out_counted[3 * i] = plan.counted[i].byte_start;
    out_counted[3 * i + 1] = plan.counted[i].n_bytes;
    out_counted[3 * i + 2] = plan.counted[i].ordinal;
  }
  for (size_t i = 0; i < plan.invalid.size(); i++) {
    out_inv[2 * i] = plan.invalid[i].first;
    out_inv[2 * i + 1] = plan.invalid[i].second;
  }
  for (size_t i = 0; i < plan.chunks.size(); i++) out_chunks[i] = plan.chunks[i];
  sizes[0] = plan.runs.size();
  sizes[1] = plan.counted.size();
  sizes[2] = plan.invalid.size();
  sizes[3] = plan.chunks.size();
  sizes[4] = plan.payload_bytes;
  tsys3[0] = plan.tsys_count0;
  tsys3[1] = plan.tsys_on0;
  tsys3[2] = plan.tsys_half;
  *n_invalid = plan.n_invalid;
  return 0;
}

int pc_segments(const int32_t *chunks, int n, int nd, int64_t fft_rot, int windowed, int32_t *out) {
  std::vector<int32_t> c(chunks, chunks + n);
  std::vector<int32_t> s = sfxg::segments_per_chunk(c, nd, fft_rot, windowed != 0);
  for (size_t i = 0; i < s.size(); i++) out[i] = s[i];
  return (int)s.size();
}

int pc_folds(const int32_t *spc, int n, int n_int, int n_sub, int32_t *out) {
  std::vector<int32_t> c(spc, spc + n);
  std::vector<int32_t> f = sfxg::fold_points(c, n_int, n_sub);
  for (size_t i = 0; i < f.size(); i++) out[i] = f[i];
  return (int)f.size();
}

void pc_find_invalid(const int32_t *s0, const int32_t *n0, int c0, const int32_t *s1, const int32_t *n1, int c1,
                     double sc0, double sc1, int64_t *out) {
  std::vector<std::pair<int32_t, int32_t>> a, b;
  for (int i = 0; i < c0; i++) a.push_back({s0[i], n0[i]});
  for (int i = 0; i < c1; i++) b.push_back({s1[i], n1[i]});
  sfxg::find_invalid(a, b, sc0, sc1, out);
}
}
