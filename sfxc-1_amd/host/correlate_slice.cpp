// correlate_slice: a C++ host driving one correlator-node slice through the
// drop-in classes of sfxc_dropin.h with the reference's signatures, wired
// and polled the way Correlator_node does it: per stream a reader buffer, a
// Bit2float_worker and a Delay_correction, all feeding one Correlation_core
// (src/correlator_node.cc:213-249), the slice's parameters fanned out to
// them (:310-405), then has_work / do_task until the core has finished
// (:263-296). The delay tables are built the way Delay_table::
// create_akima_spline builds them (src/delay_table_akima.cc:430-502): Akima
// splines over the delay sampled every second.
//
// Usage: correlate_slice <slice.bin> <out.cor> [precision]
// slice.bin (little endian): sfxg_correlation_parameters fields as written by
// sfxg/slicefile.py, then per stream: sfxg_station_parameters, the delay model
// (scan_begin int64, interval f64, n int32, 4n f64: the delay's cubic pieces,
// sampled here at the piece starts) and the HEADER stream (uint64 length +
// bytes). out.cor receives the bytes Correlation_core hands its Data_writer.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "sfxc_dropin.h"

template <class T>
static bool rd(FILE *f, T &v) { return fread(&v, sizeof(T), 1, f) == 1; }

// The output node's end of the slice: the bytes in a file.
class File_writer : public Data_writer {
 public:
  explicit File_writer(const char *path) : f_(fopen(path, "wb")) {
    if (!f_) throw std::runtime_error(std::string("cannot open ") + path);
  }
  ~File_writer() override { fclose(f_); }

 protected:
  size_t do_put_bytes(size_t n, const char *buf) override { return fwrite(buf, 1, n, f_); }

 private:
  FILE *f_;
};

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s slice.bin out.cor [precision]\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t t5[5];
  int32_t nchan, nd, nc, integration_nr, rate, bw, freq_nr, window, ref_station, n_streams;
  int64_t channel_freq;
  char sideband;
  if (fread(t5, sizeof(int64_t), 5, f) != 5 || !rd(f, nchan) || !rd(f, nd) || !rd(f, nc) || !rd(f, integration_nr) ||
      !rd(f, rate) || !rd(f, bw) || !rd(f, channel_freq) || !rd(f, sideband) || !rd(f, freq_nr) || !rd(f, window) ||
      !rd(f, ref_station) || !rd(f, n_streams))
    return 3;
  Correlation_parameters p;
  p.experiment_start = Time::from_ticks(t5[0]);
  p.start_time = Time::from_ticks(t5[1]);
  p.stop_time = Time::from_ticks(t5[2]);
  p.integration_time = Time::from_ticks(t5[3]);
  p.sub_integration_time = Time::from_ticks(t5[4]);
  p.number_channels = nchan;
  p.fft_size_delaycor = nd;
  p.fft_size_correlation = nc;
  p.integration_nr = integration_nr;
  p.slice_nr = integration_nr;
  p.slice_offset = 1;
  p.sample_rate = rate;
  p.bandwidth = bw;
  p.channel_freq = channel_freq;
  p.sideband = sideband;
  p.frequency_nr = freq_nr;
  p.window = window;
  p.reference_station = ref_station;
  p.n_phase_centers = 1;
  std::vector<Delay_table_akima> delays;
  std::vector<std::vector<uint8_t>> hdr(n_streams);
  for (int i = 0; i < n_streams; i++) {
    sfxg_station_parameters sp;
    if (fread(&sp, sizeof(sp), 1, f) != 1) return 3;
    Correlation_parameters::Station_parameters s;
    s.station_number = sp.station_number;
    s.station_stream = sp.station_stream;
    s.start_time = p.start_time;
    s.stop_time = p.stop_time;
    s.sample_rate = sp.sample_rate;
    s.channel_freq = sp.channel_freq;
    s.bandwidth = sp.bandwidth;
    s.sideband = sp.sideband;
    s.polarisation = sp.polarisation;
    s.bits_per_sample = sp.bits_per_sample;
    s.LO_offset = sp.LO_offset;
    s.tsys_freq = sp.tsys_freq;
    p.station_streams.push_back(s);
    int64_t scan_begin;
    double interval;
    int32_t n;
    if (!rd(f, scan_begin) || !rd(f, interval) || !rd(f, n)) return 3;
    std::vector<double> coef(4 * (size_t)n);
    if (fread(coef.data(), sizeof(double), coef.size(), f) != coef.size()) return 3;
    // the delay table's samples: the value at every piece start, and past
    // the last piece its continuation, so that the spline has the five
    // points a .del table's scan gives it at least
    const int n_pts = std::max(n + 1, 5);
    std::vector<double> t(n_pts), y(n_pts);
    for (int k = 0; k < n_pts; k++) {
      t[k] = k * interval;
      const int piece = k < n ? k : n - 1;
      const double u = (k - piece) * interval;
      const double *c = &coef[4 * piece];
      y[k] = c[0] + u * (c[1] + u * (c[2] + u * c[3]));
    }
    try {
      delays.push_back(Delay_table_akima::from_samples(Time::from_ticks(scan_begin), t, {y}, {}, {},
                                                       Time::from_ticks(scan_begin), 0, 0));
    } catch (const std::exception &e) {
      fprintf(stderr, "error: stream %d: %s\n", i, e.what());
      return 1;
    }
    uint64_t nb = 0;
    if (!rd(f, nb)) return 3;
    hdr[i].resize(nb);
    if (fread(hdr[i].data(), 1, nb, f) != nb) return 3;
  }
  fclose(f);
  try {
    sfxc_dropin::settings().options.lut = SFXG_LUT_CONFIG;
    if (argc > 3) sfxc_dropin::settings().options.precision = atoi(argv[3]);
    // Correlator_node::hook_added_data_reader, per stream
    Correlation_core core;
    std::vector<Correlator_node_types::Channel_circular_input_buffer_ptr> reader(n_streams);
    std::vector<std::unique_ptr<Bit2float_worker>> bit2float;
    std::vector<std::unique_ptr<Delay_correction>> delay;
    std::vector<bit_statistics_ptr> stats;
    for (int i = 0; i < n_streams; i++) {
      reader[i] = std::make_shared<Correlator_node_types::Channel_circular_input_buffer>();
      stats.push_back(std::make_shared<bit_statistics>());
      bit2float.emplace_back(new Bit2float_worker(i, stats[i]));
      bit2float[i]->connect_to(reader[i]);
      delay.emplace_back(new Delay_correction(i));
      delay[i]->connect_to(bit2float[i]->get_output_buffer());
      core.connect_to(i, stats[i], delay[i]->get_output_buffer());
      core.connect_to(i, bit2float[i]->get_invalid());
    }
    core.set_data_writer(std::make_shared<File_writer>(argv[2]));
    // Correlator_node::set_parameters
    std::vector<std::vector<double>> uvw(n_streams, std::vector<double>(3, 0.0));
    for (int i = 0; i < n_streams; i++) {
      bit2float[i]->set_new_parameters(p, delays[i]);
      delay[i]->set_parameters(p, delays[i]);
    }
    core.set_parameters(p, delays, uvw, 0);
    // the readers deliver the streams in pieces while the node polls
    const size_t piece = 65536;
    std::vector<size_t> sent(n_streams, 0);
    while (!core.finished()) {
      for (int i = 0; i < n_streams; i++) {
        if (sent[i] < hdr[i].size()) {
          const size_t k = std::min(piece, hdr[i].size() - sent[i]);
          reader[i]->write(hdr[i].data() + sent[i], k);
          sent[i] += k;
        }
        if (bit2float[i]->has_work()) bit2float[i]->do_task();
        if (delay[i]->has_work()) delay[i]->do_task();
      }
      if (core.has_work()) core.do_task();
      bool pending = false;
      for (int i = 0; i < n_streams; i++) pending |= sent[i] < hdr[i].size() || bit2float[i]->has_work();
      if (!pending && !core.finished() && !core.has_work())
        throw std::runtime_error("a stream ended without its ENDSTREAM record");
    }
    printf("%d baselines; levels of stream 0: %d %d %d %d %d\n", core.number_of_baselines(),
           stats[0]->get_statistics()[0], stats[0]->get_statistics()[1], stats[0]->get_statistics()[2],
           stats[0]->get_statistics()[3], stats[0]->get_statistics()[4]);
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
