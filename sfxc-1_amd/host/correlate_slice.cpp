// correlate_slice: a C++ host driving one correlator-node slice through the
// adapter classes, the way Correlator_node wires Bit2float_worker and
// Correlation_core (sfxc/src/correlator_node.cc:213-296).
//
// Usage: correlate_slice <slice.bin> <out.cor>
// slice.bin (little endian): sfxg_correlation_parameters fields as written by
// sfxg/slicefile.py, then per stream: sfxg_station_parameters, the delay model
// (scan_begin int64, interval f64, n int32, 4n f64) and the HEADER stream
// (uint64 length + bytes). out.cor receives the integration_write bytes.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "sfxg_adapter.h"

template <class T>
static bool rd(FILE *f, T &v) { return fread(&v, sizeof(T), 1, f) == 1; }

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s slice.bin out.cor\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  sfxg_correlation_parameters p;
  memset(&p, 0, sizeof(p));
  int32_t n_streams = 0;
  if (!rd(f, p.experiment_start) || !rd(f, p.start_time) || !rd(f, p.stop_time) || !rd(f, p.integration_time) ||
      !rd(f, p.sub_integration_time) || !rd(f, p.number_channels) || !rd(f, p.fft_size_delaycor) ||
      !rd(f, p.fft_size_correlation) || !rd(f, p.integration_nr) || !rd(f, p.sample_rate) || !rd(f, p.bandwidth) ||
      !rd(f, p.channel_freq) || !rd(f, p.sideband) || !rd(f, p.frequency_nr) || !rd(f, p.window) ||
      !rd(f, p.reference_station) || !rd(f, n_streams))
    return 3;
  p.n_phase_centers = 1;
  p.slice_offset = 1;
  p.n_streams = n_streams;
  std::vector<sfxg_station_parameters> st(n_streams);
  std::vector<sfxg_delay_model> models(n_streams);
  std::vector<std::vector<double>> coef(n_streams);
  std::vector<std::vector<uint8_t>> hdr(n_streams);
  for (int i = 0; i < n_streams; i++) {
    if (fread(&st[i], sizeof(sfxg_station_parameters), 1, f) != 1) return 3;
    sfxg_delay_model &m = models[i];
    memset(&m, 0, sizeof(m));
    if (!rd(f, m.scan_begin) || !rd(f, m.interval) || !rd(f, m.n_pieces)) return 3;
    coef[i].resize(4 * m.n_pieces);
    if (fread(coef[i].data(), sizeof(double), coef[i].size(), f) != coef[i].size()) return 3;
    m.delay = coef[i].data();
    uint64_t n = 0;
    if (!rd(f, n)) return 3;
    hdr[i].resize(n);
    if (fread(hdr[i].data(), 1, n, f) != n) return 3;
  }
  fclose(f);
  p.station_streams = st.data();
  try {
    sfxg_adapter::Context ctx(0);
    sfxg_options opt = {SFXG_LUT_CONFIG, SFXG_PREC_FP32, 0, 0};
    ctx.check(sfxg_set_options(ctx.get(), &opt), "sfxg_set_options");
    sfxg_adapter::Correlation_core core(ctx);
    core.set_parameters(p, models, std::vector<double>());
    std::vector<sfxg_adapter::Bit2float_worker> workers;
    for (int i = 0; i < n_streams; i++) workers.emplace_back(ctx, i);
    for (int i = 0; i < n_streams; i++) {
      // the reader thread delivers the stream in pieces
      const size_t piece = 65536;
      for (size_t o = 0; o < hdr[i].size(); o += piece)
        workers[i].put_bytes(hdr[i].data() + o, std::min(piece, hdr[i].size() - o));
      if (!workers[i].has_work()) throw sfxg_adapter::Error("stream without ENDSTREAM");
      workers[i].do_task();
    }
    core.do_task();
    std::vector<uint8_t> out = core.integration_write();
    FILE *o = fopen(argv[2], "wb");
    if (!o || fwrite(out.data(), 1, out.size(), o) != out.size()) return 4;
    fclose(o);
    printf("%d baselines, %zu bytes\n", core.number_of_baselines(), out.size());
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
