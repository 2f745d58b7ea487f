// sfxg_adapter.h -- C++ front-ends with the reference's class and method
// names over the C ABI (include/sfxg.h), for hosts that drive the correlator
// node the way SFXC's Correlator_node does (sfxc/src/correlator_node.cc:263-296).
//
//   Bit2float_worker   (sfxc/include/bit2float_worker.h:39-74): collects one
//                      stream's HEADER byte stream for the slice and hands it
//                      over once complete; do_task() returns the bytes taken.
//   Correlation_core   (sfxc/include/correlation_core.h:27-59): set_parameters,
//                      do_task (runs Delay_correction and the X-step for the
//                      whole slice), finished, number_of_baselines and the
//                      integration_write bytes; set_mask_parameters takes the
//                      Mask_parameters (sfxc/include/control_parameters.h:107-114)
//                      that the node points every slice's parameters at.
// Delay_correction has no front-end: its work runs inside Correlation_core's
// do_task on the device. Errors throw sfxg_adapter::Error (the reference
// aborts through sfxc_abort, sfxc/src/utils.cc:44-95).
#pragma once

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sfxg.h"

namespace sfxg_adapter {

struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

class Context {
 public:
  explicit Context(int device) {
    if (sfxg_create(device, &ctx_) != SFXG_OK) throw Error("sfxg_create failed");
  }
  ~Context() { sfxg_destroy(ctx_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  sfxg_ctx *get() const { return ctx_; }
  void check(int rc, const char *what) const {
    if (rc != SFXG_OK) throw Error(std::string(what) + ": " + sfxg_last_error(ctx_));
  }

 private:
  sfxg_ctx *ctx_ = nullptr;
};

class Bit2float_worker {
 public:
  Bit2float_worker(Context &ctx, int stream_idx) : ctx_(ctx), stream_(stream_idx) {}
  // The reader thread's bytes for this slice (the Channel_circular_input_buffer feed).
  void put_bytes(const uint8_t *p, size_t n) { buf_.insert(buf_.end(), p, p + n); }
  // The station's raw VDIF frames in device memory instead of a HEADER
  // stream: the input node's reader and writer run in the library.
  void put_vdif(const void *device_frames, size_t n_bytes, const sfxg_vdif_stream &v) {
    ctx_.check(sfxg_push_vdif(ctx_.get(), stream_, device_frames, n_bytes, &v), "sfxg_push_vdif");
  }
  bool has_work() const { return !buf_.empty() && buf_.back() == 3 /* HEADER_ENDSTREAM */; }
  int64_t do_task() {
    ctx_.check(sfxg_push_stream(ctx_.get(), stream_, buf_.data(), buf_.size()), "sfxg_push_stream");
    const int64_t n = (int64_t)buf_.size();
    buf_.clear();
    return n;
  }

 private:
  Context &ctx_;
  int stream_;
  std::vector<uint8_t> buf_;
};

// Mask_parameters (sfxc/include/control_parameters.h:107-114).
struct Mask_parameters {
  bool normalize = false;
  std::vector<double> mask, window;
};

class Correlation_core {
 public:
  explicit Correlation_core(Context &ctx) : ctx_(ctx) {}
  void set_parameters(const sfxg_correlation_parameters &p, const std::vector<sfxg_delay_model> &models,
                      const std::vector<double> &uvw, const sfxg_pulsar *pulsar = nullptr) {
    ctx_.check(sfxg_set_parameters(ctx_.get(), &p, models.data(), uvw.empty() ? nullptr : uvw.data(), pulsar),
               "sfxg_set_parameters");
    finished_ = false;
  }
  // The further phase centres of a multi-phase-centre slice: models
  // (n - 1) x n_streams centre-major, uvw n x n_streams x 3, source_nr n.
  void set_phase_centers(int n, const std::vector<sfxg_delay_model> &models, const std::vector<double> &uvw,
                         const std::vector<int32_t> &source_nr) {
    ctx_.check(sfxg_set_phase_centers(ctx_.get(), n, models.data(), uvw.empty() ? nullptr : uvw.data(),
                                      source_nr.empty() ? nullptr : source_nr.data()),
               "sfxg_set_phase_centers");
  }
  // Correlator_node_controller's MPI_TAG_MASK_PARAMETERS handler
  // (sfxc/src/correlator_node_controller.cc:72-77).
  void set_mask_parameters(const Mask_parameters &m) {
    ctx_.check(sfxg_set_mask(ctx_.get(), m.normalize ? 1 : 0, m.mask.empty() ? nullptr : m.mask.data(),
                             (int32_t)m.mask.size(), m.window.empty() ? nullptr : m.window.data(),
                             (int32_t)m.window.size()),
               "sfxg_set_mask");
  }
  void do_task() {
    ctx_.check(sfxg_correlate(ctx_.get()), "sfxg_correlate");
    finished_ = true;
  }
  bool finished() const { return finished_; }
  int number_of_baselines() const {
    int32_t n = 0;
    ctx_.check(sfxg_number_of_baselines(ctx_.get(), &n), "sfxg_number_of_baselines");
    return n;
  }
  // The bytes integration_write passes to Data_writer::put_bytes.
  std::vector<uint8_t> integration_write() const {
    size_t n = 0;
    ctx_.check(sfxg_output_size(ctx_.get(), &n), "sfxg_output_size");
    std::vector<uint8_t> out(n);
    ctx_.check(sfxg_fetch_output(ctx_.get(), out.data(), n, &n), "sfxg_fetch_output");
    out.resize(n);
    return out;
  }
  void statistics(int stream_idx, int32_t levels[5], int32_t tsys[4]) const {
    ctx_.check(sfxg_get_statistics(ctx_.get(), stream_idx, levels, tsys), "sfxg_get_statistics");
  }

 private:
  Context &ctx_;
  bool finished_ = false;
};

}  // namespace sfxg_adapter
