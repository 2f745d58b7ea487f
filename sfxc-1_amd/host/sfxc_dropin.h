// sfxc_dropin.h -- the correlator node's DSP classes with the reference's
// names and signatures over the C ABI (include/sfxg.h), so that
// Correlator_node's wiring and polling loop (src/correlator_node.cc:213-296)
// drive the MI355X path unchanged:
//
//   Bit2float_worker   include/bit2float_worker.h:39-74
//   Delay_correction   include/delay_correction.h:39-55
//   Correlation_core   include/correlation_core.h:27-59
//   Correlation_core_pulsar::set_parameters   include/correlation_core_pulsar.h:13-16
//
// The data between them is a stream's complete HEADER byte stream for the
// slice instead of FP32 blocks: Bit2float_worker::do_task takes the bytes
// the reader wrote and queues the stream once its ENDSTREAM record arrives;
// Delay_correction::do_task passes it on; Correlation_core::do_task hands
// every stream to the library once all have arrived, correlates the slice on
// the GPU (unpack, integer shift, delay correction, channeliser, X-step,
// normalisation) and passes the integration_write bytes to the data writer.
// The statistics objects are filled from the library after the slice. Errors
// throw std::runtime_error (the reference aborts through sfxc_abort,
// src/utils.cc:44-95).
#pragma once

#include <stdio.h>

#include <algorithm>

#include "sfxc_types.h"

namespace sfxc_dropin {

// The device and options every Correlation_core of the process uses (a
// correlator node owns one GPU).
struct Library_settings {
  int device = 0;
  sfxg_options options = {SFXG_LUT_REFERENCE, SFXG_PREC_FP32, 0, 0, 0};
};
inline Library_settings &settings() {
  static Library_settings s;
  return s;
}

inline void check(sfxg_ctx *ctx, int rc, const char *what) {
  if (rc != SFXG_OK) throw std::runtime_error(std::string(what) + ": " + (ctx ? sfxg_last_error(ctx) : "failed"));
}

// The C ABI's parameter block of a slice; the station list and the models
// live in the holder.
struct Abi_parameters {
  sfxg_correlation_parameters p;
  std::vector<sfxg_station_parameters> st;
  Abi_parameters(const Correlation_parameters &c) {
    memset(&p, 0, sizeof(p));
    p.experiment_start = c.experiment_start.get_clock_ticks();
    p.start_time = c.start_time.get_clock_ticks();
    p.stop_time = c.stop_time.get_clock_ticks();
    p.integration_time = c.integration_time.get_clock_ticks();
    p.sub_integration_time = c.sub_integration_time.get_clock_ticks();
    p.number_channels = c.number_channels;
    p.fft_size_delaycor = c.fft_size_delaycor;
    p.fft_size_correlation = c.fft_size_correlation;
    p.integration_nr = c.integration_nr;
    p.slice_nr = c.slice_nr;
    p.slice_offset = c.slice_offset;
    p.sample_rate = c.sample_rate;
    p.bandwidth = c.bandwidth;
    p.channel_freq = c.channel_freq;
    p.sideband = c.sideband;
    p.polarisation = c.polarisation;
    p.frequency_nr = c.frequency_nr;
    p.cross_polarize = c.cross_polarize;
    p.reference_station = c.reference_station;
    p.window = c.window;
    p.n_phase_centers = c.n_phase_centers;
    p.pulsar_binning = c.pulsar_binning;
    for (const auto &s : c.station_streams) {
      sfxg_station_parameters q;
      memset(&q, 0, sizeof(q));
      q.station_number = s.station_number;
      q.station_stream = s.station_stream;
      q.sample_rate = s.sample_rate;
      q.bandwidth = s.bandwidth;
      q.channel_freq = s.channel_freq;
      q.sideband = s.sideband;
      q.polarisation = s.polarisation;
      q.bits_per_sample = s.bits_per_sample;
      q.LO_offset = s.LO_offset;
      q.tsys_freq = s.tsys_freq;
      st.push_back(q);
    }
    p.n_streams = (int32_t)st.size();
    p.station_streams = st.data();
  }
};

}  // namespace sfxc_dropin

class Bit2float_worker {
 public:
  typedef Correlator_node_types Types;
  typedef Types::Channel_circular_input_buffer Input_buffer;
  typedef Types::Channel_circular_input_buffer_ptr Input_buffer_ptr;
  typedef Types::Channel_queue Output_queue;
  typedef Types::Channel_queue_ptr Output_queue_ptr;
  typedef Types::Invalid Invalid;

  Bit2float_worker(int stream_nr, bit_statistics_ptr statistics_)
      : stream_nr_(stream_nr), statistics_(statistics_), out_(std::make_shared<Output_queue>()) {}

  // Moves the bytes the reader has written into the slice's HEADER stream;
  // when its ENDSTREAM record has arrived the stream goes to the output
  // queue. Returns the bytes taken.
  int64_t do_task() {
    if (!has_work()) return 0;
    int64_t taken = 0;
    std::deque<uint8_t> &in = input_->bytes;
    while (!in.empty()) {
      if (payload_ > 0) {  // DATA payload bytes, moved as one piece
        const size_t k = std::min<size_t>((size_t)payload_, in.size());
        cur_.insert(cur_.end(), in.begin(), in.begin() + k);
        in.erase(in.begin(), in.begin() + k);
        payload_ -= (int64_t)k;
        taken += (int64_t)k;
        continue;
      }
      const uint8_t b = in.front();
      in.pop_front();
      cur_.push_back(b);
      taken++;
      if (len_bytes_ > 0) {  // the 16-bit length of a DATA or INVALID record
        len_ |= (int)b << (8 * (2 - len_bytes_));
        if (--len_bytes_ == 0 && data_record_) payload_ = len_;
        continue;
      }
      if (delay_byte_) {
        delay_byte_ = false;
        continue;
      }
      switch (b) {  // a record header (src/utils.h:39-42)
        case HEADER_DATA: data_record_ = true; len_ = 0; len_bytes_ = 2; break;
        case HEADER_INVALID: data_record_ = false; len_ = 0; len_bytes_ = 2; break;
        case HEADER_DELAY: delay_byte_ = true; break;
        case HEADER_ENDSTREAM:
          out_->push(std::make_shared<std::vector<uint8_t>>(cur_.begin(), cur_.end()));
          cur_.clear();
          return taken;
        default: throw std::runtime_error("Bit2float_worker: invalid header byte in the HEADER stream");
      }
    }
    return taken;
  }
  bool has_work() { return have_parameters_ && input_ && !input_->bytes.empty(); }
  void connect_to(Input_buffer_ptr new_Input_buffer) { input_ = new_Input_buffer; }
  const char *name() { return "Bit2float_worker"; }
  Output_queue_ptr get_output_buffer() { return out_; }
  void set_new_parameters(const Correlation_parameters &parameters, Delay_table_akima &delays) {
    (void)parameters;
    (void)delays;  // the library takes the slice's parameters and delays from Correlation_core
    have_parameters_ = true;
  }
  // The library derives the invalid samples from the INVALID records itself.
  std::vector<Invalid> *get_invalid() { return &invalid_; }
  void empty_input_queue() {
    if (input_) input_->bytes.clear();
    cur_.clear();
    payload_ = 0;
    len_bytes_ = 0;
    delay_byte_ = false;
  }
  bit_statistics_ptr statistics() const { return statistics_; }
  int stream_nr() const { return stream_nr_; }

 private:
  enum : uint8_t { HEADER_DATA = 0, HEADER_INVALID = 1, HEADER_DELAY = 2, HEADER_ENDSTREAM = 3 };
  int stream_nr_;
  bit_statistics_ptr statistics_;
  Input_buffer_ptr input_;
  Output_queue_ptr out_;
  std::vector<uint8_t> cur_;
  std::vector<Invalid> invalid_;
  int64_t payload_ = 0;  // DATA payload bytes still to come
  int len_ = 0, len_bytes_ = 0;
  bool data_record_ = false, delay_byte_ = false;
  bool have_parameters_ = false;
};

class Delay_correction {
 public:
  typedef Correlator_node_types::Channel_queue Input_buffer;
  typedef Correlator_node_types::Channel_queue_ptr Input_buffer_ptr;
  typedef Correlator_node_types::Delay_queue Output_buffer;
  typedef Correlator_node_types::Delay_queue_ptr Output_buffer_ptr;

  explicit Delay_correction(int stream_nr) : stream_nr_(stream_nr), out_(std::make_shared<Output_buffer>()) {}
  Output_buffer_ptr get_output_buffer() { return out_; }
  void connect_to(Input_buffer_ptr new_input_buffer) { in_ = new_input_buffer; }
  void set_parameters(const Correlation_parameters &parameters, Delay_table_akima &delays) {
    (void)parameters;
    (void)delays;  // the delay correction runs on the GPU with the core's delay tables
  }
  // The fractional delay and fringe rotation run on the GPU in the core's
  // do_task: the stream is passed on.
  void do_task() {
    if (!has_work()) return;
    out_->push(in_->front());
    in_->pop();
  }
  bool has_work() { return in_ && !in_->empty(); }

 private:
  int stream_nr_;
  Input_buffer_ptr in_;
  Output_buffer_ptr out_;
};

class Correlation_core {
 public:
  typedef Delay_correction::Output_buffer Input_buffer;
  typedef Delay_correction::Output_buffer_ptr Input_buffer_ptr;
  typedef Correlator_node_types::Invalid Invalid;

  Correlation_core() {
    const sfxc_dropin::Library_settings &s = sfxc_dropin::settings();
    sfxc_dropin::check(nullptr, sfxg_create(s.device, &ctx_), "sfxg_create");
    sfxc_dropin::check(ctx_, sfxg_set_options(ctx_, &s.options), "sfxg_set_options");
  }
  virtual ~Correlation_core() { sfxg_destroy(ctx_); }
  Correlation_core(const Correlation_core &) = delete;
  Correlation_core &operator=(const Correlation_core &) = delete;

  // Once every stream's slice has arrived: the streams to the library, the
  // slice correlated, statistics filled, and the integration_write bytes of
  // every record (bin or phase centre) to the data writer.
  virtual void do_task() {
    if (!has_work()) return;
    for (size_t i = 0; i < inputs_.size(); i++) {
      Correlator_node_types::Slice_stream s = inputs_[i]->front();
      inputs_[i]->pop();
      sfxc_dropin::check(ctx_, sfxg_push_stream(ctx_, (int32_t)i, s->data(), s->size()), "sfxg_push_stream");
    }
    sfxc_dropin::check(ctx_, sfxg_correlate(ctx_), "sfxg_correlate");
    for (size_t i = 0; i < stats_.size(); i++) {
      if (!stats_[i]) continue;
      int32_t lv[5], ts[4];
      sfxc_dropin::check(ctx_, sfxg_get_statistics(ctx_, (int32_t)i, lv, ts), "sfxg_get_statistics");
      std::copy(lv, lv + 5, stats_[i]->statistics.begin());
      std::copy(ts, ts + 4, stats_[i]->tsys.begin());
    }
    size_t n = 0;
    sfxc_dropin::check(ctx_, sfxg_output_size(ctx_, &n), "sfxg_output_size");
    std::vector<uint8_t> out(n);
    sfxc_dropin::check(ctx_, sfxg_fetch_output(ctx_, out.data(), n, &n), "sfxg_fetch_output");
    if (writer_) writer_->put_bytes(n, reinterpret_cast<const char *>(out.data()));
    finished_ = true;
  }
  bool has_work() {
    if (finished_ || !have_parameters_ || inputs_.empty()) return false;
    for (const auto &q : inputs_)
      if (!q || q->empty()) return false;
    return true;
  }
  const char *name() { return "Correlation_core"; }
  bool finished() { return finished_; }
  bool almost_finished() { return finished_; }

  void connect_to(size_t stream, bit_statistics_ptr statistics_, Input_buffer_ptr buffer) {
    if (inputs_.size() <= stream) inputs_.resize(stream + 1);
    if (stats_.size() <= stream) stats_.resize(stream + 1);
    inputs_[stream] = buffer;
    stats_[stream] = statistics_;
  }
  void connect_to(size_t stream, std::vector<Invalid> *invalid_) { (void)stream, (void)invalid_; }

  virtual void set_parameters(const Correlation_parameters &parameters, std::vector<Delay_table_akima> &delays,
                              std::vector<std::vector<double>> &uvw, int node_nr) {
    set_parameters_impl(parameters, delays, uvw, node_nr, nullptr);
  }
  void set_data_writer(std::shared_ptr<Data_writer> writer) { writer_ = writer; }
  int number_of_baselines() {
    int32_t n = 0;
    sfxc_dropin::check(ctx_, sfxg_number_of_baselines(ctx_, &n), "sfxg_number_of_baselines");
    return n;
  }
  std::shared_ptr<Data_writer> data_writer() { return writer_; }
  // Source numbers of the phase centres' records (integration_write's
  // sourcenr), looked up by the source names of the delay tables.
  void add_source_list(const std::map<std::string, int> &sources_) { sources_list_ = sources_; }
  // Mask_parameters of the job (control_parameters.h:107-114), taken from
  // the slice's parameters at set_parameters.
  sfxg_ctx *context() { return ctx_; }

 protected:
  void set_parameters_impl(const Correlation_parameters &parameters, std::vector<Delay_table_akima> &delays,
                           std::vector<std::vector<double>> &uvw, int node_nr, const sfxg_pulsar *pulsar) {
    (void)node_nr;
    sfxc_dropin::Abi_parameters ap(parameters);
    const size_t M = parameters.station_streams.size();
    if (delays.size() < M) throw std::runtime_error("Correlation_core: one delay table per stream");
    const int n_pc = std::max(1, parameters.n_phase_centers);
    std::vector<std::vector<double>> coef(M * n_pc);
    std::vector<sfxg_delay_model> models(M * n_pc);
    for (int pc = 0; pc < n_pc; pc++)
      for (size_t i = 0; i < M; i++) models[pc * M + i] = delays[i].model(pc, coef[pc * M + i]);
    // uvw_table[stream][3 pc + k] (correlation_core.cc:356-361)
    std::vector<double> uvw_flat(3 * M * n_pc, 0.0);
    for (int pc = 0; pc < n_pc; pc++)
      for (size_t i = 0; i < M && i < uvw.size(); i++)
        for (int k = 0; k < 3 && (size_t)(3 * pc + k) < uvw[i].size(); k++) uvw_flat[(pc * M + i) * 3 + k] = uvw[i][3 * pc + k];
    if (parameters.mask_parameters) {
      const Mask_parameters &mp = *parameters.mask_parameters;
      sfxc_dropin::check(ctx_, sfxg_set_mask(ctx_, mp.normalize, mp.mask.empty() ? nullptr : mp.mask.data(),
                                             (int32_t)mp.mask.size(), mp.window.empty() ? nullptr : mp.window.data(),
                                             (int32_t)mp.window.size()),
                         "sfxg_set_mask");
    }
    sfxc_dropin::check(ctx_, sfxg_set_parameters(ctx_, &ap.p, models.data(), uvw_flat.data(), pulsar),
                       "sfxg_set_parameters");
    if (n_pc > 1) {
      std::vector<int32_t> src(n_pc, 0);
      for (int pc = 0; pc < n_pc; pc++) {
        auto it = sources_list_.find(delays[0].get_source(pc));
        src[pc] = it == sources_list_.end() ? 0 : it->second;
      }
      sfxc_dropin::check(ctx_, sfxg_set_phase_centers(ctx_, n_pc, models.data() + M, uvw_flat.data(), src.data()),
                         "sfxg_set_phase_centers");
    }
    have_parameters_ = true;
    finished_ = false;
  }

  sfxg_ctx *ctx_ = nullptr;
  std::vector<Input_buffer_ptr> inputs_;
  std::vector<bit_statistics_ptr> stats_;
  std::shared_ptr<Data_writer> writer_;
  std::map<std::string, int> sources_list_;
  bool have_parameters_ = false, finished_ = false;
};

class Correlation_core_pulsar : public Correlation_core {
 public:
  typedef Pulsar_parameters::Pulsar Pulsar;
  // The polyco whose tmid lies nearest the slice start, as
  // Correlation_core_pulsar::set_parameters picks it
  // (src/correlation_core_pulsar.cc:52-62); the library does the rest.
  void set_parameters(const Correlation_parameters &parameters, Pulsar &pulsar,
                      std::vector<Delay_table_akima> &delays, std::vector<std::vector<double>> &uvw, int node_nr) {
    if (pulsar.polyco_params.empty()) throw std::runtime_error("Correlation_core_pulsar: no polyco");
    const double start_mjd = parameters.start_time.get_mjd();
    size_t best = 0;
    for (size_t i = 1; i < pulsar.polyco_params.size(); i++)
      if (fabs(pulsar.polyco_params[i].tmid - start_mjd) < fabs(pulsar.polyco_params[best].tmid - start_mjd))
        best = i;
    const Pulsar_parameters::Polyco_params &pc = pulsar.polyco_params[best];
    sfxg_pulsar ps;
    memset(&ps, 0, sizeof(ps));
    ps.nbins = pulsar.nbins;
    ps.n_coef = (int32_t)pc.coef.size();
    ps.gate_start = pulsar.interval.start;
    ps.gate_stop = pulsar.interval.stop;
    ps.tmid = pc.tmid;
    ps.DM = pc.DM;
    ps.ref_phase = pc.ref_phase;
    ps.ref_freq = pc.ref_freq;
    ps.obs_freq = pc.obs_freq;
    ps.coef = pc.coef.data();
    set_parameters_impl(parameters, delays, uvw, node_nr, &ps);
  }
};
