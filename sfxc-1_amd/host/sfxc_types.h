// sfxc_types.h -- the correlator node's parameter and helper types as the
// drop-in classes of sfxc_dropin.h take them, for hosts built outside the
// SFXC tree. Each type keeps the reference's name, fields and the methods the
// correlator-node data path uses; inside the SFXC tree the reference's own
// definitions stand in their place (the drop-in classes use nothing else of
// them).
//
//   Time                    include/correlator_time.h:12-64 (512 MHz ticks since MJD 51544)
//   Correlation_parameters  include/control_parameters.h:117-182
//   Mask_parameters         include/control_parameters.h:107-114
//   Pulsar_parameters       include/control_parameters.h:70-105
//   Delay_table_akima       include/delay_table_akima.h:27-67, src/delay_table_akima.cc:114-166
//   bit_statistics          include/bit_statistics.h:16-34
//   Data_writer             include/data_writer.h:28,67
//
// Shared pointers are std::shared_ptr here (boost::shared_ptr in the
// reference).
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sfxg.h"

class Time {
 public:
  static constexpr double kClockRate = 512e6;  // ticks per second
  static constexpr int kReferenceMjd = 51544;
  Time() = default;
  explicit Time(double usec) : ticks_((int64_t)llround(usec * (kClockRate / 1e6))) {}
  Time(int mjd, double sec) { set_time(mjd, sec); }
  static Time from_ticks(int64_t t) {
    Time r;
    r.ticks_ = t;
    return r;
  }
  int64_t get_clock_ticks() const { return ticks_; }
  void set_clock_ticks(int64_t t) { ticks_ = t; }
  void set_time(int mjd, double sec) {
    ticks_ = (int64_t)(mjd - kReferenceMjd) * 86400 * (int64_t)kClockRate + (int64_t)llround(sec * kClockRate);
  }
  double get_mjd() const {
    const int64_t tpd = (int64_t)86400 * (int64_t)kClockRate;
    return kReferenceMjd + (double)(ticks_ / tpd) + (double)(ticks_ % tpd) / kClockRate / 86400;
  }
  double get_time_usec() const {  // microseconds since the start of the day
    const int64_t tpd = (int64_t)86400 * (int64_t)kClockRate;
    return (double)(ticks_ % tpd) / (kClockRate / 1e6);
  }
  double diff(const Time &o) const { return (double)(ticks_ - o.ticks_) / kClockRate; }  // seconds
  Time operator+(const Time &o) const { return from_ticks(ticks_ + o.ticks_); }
  Time operator-(const Time &o) const { return from_ticks(ticks_ - o.ticks_); }
  Time operator*(double v) const { return from_ticks((int64_t)llround((double)ticks_ * v)); }
  double operator/(const Time &o) const { return (double)ticks_ / (double)o.ticks_; }
  bool operator==(const Time &o) const { return ticks_ == o.ticks_; }
  bool operator!=(const Time &o) const { return ticks_ != o.ticks_; }
  bool operator<(const Time &o) const { return ticks_ < o.ticks_; }
  bool operator<=(const Time &o) const { return ticks_ <= o.ticks_; }
  bool operator>(const Time &o) const { return ticks_ > o.ticks_; }
  bool operator>=(const Time &o) const { return ticks_ >= o.ticks_; }

 private:
  int64_t ticks_ = 0;
};

class Mask_parameters {
 public:
  bool normalize = false;
  std::vector<double> mask;
  std::vector<double> window;
};

class Pulsar_parameters {
 public:
  struct Polyco_params {
    char name[11] = {0};
    double tmid = 0;       // reference time [MJD]
    double DM = 0;         // dispersion measure
    double ref_phase = 0;  // RPHASE
    double ref_freq = 0;   // F0
    int32_t data_span = 0;
    int32_t n_coef = 0;
    double obs_freq = 0;   // MHz
    std::vector<double> coef;
  };
  struct Pulsar {
    char name[11] = {0};
    int32_t nbins = 0;
    struct Interval {
      double start, stop;
    } interval = {0, 1};
    std::vector<Polyco_params> polyco_params;
  };
  std::map<std::string, Pulsar> pulsars;
};

// SFXC_WINDOW_* (include/utils.h:209-213)
enum { SFXC_WINDOW_NONE = SFXG_WINDOW_NONE, SFXC_WINDOW_RECT, SFXC_WINDOW_COS, SFXC_WINDOW_HAMMING, SFXC_WINDOW_HANN };

class Correlation_parameters {
 public:
  class Station_parameters {
   public:
    int32_t station_number = 0;
    int32_t station_stream = 0;
    Time start_time, stop_time;
    int32_t sample_rate = 0;
    int64_t channel_freq = 0;
    int32_t bandwidth = 0;
    char sideband = 'U';
    char polarisation = 'R';
    int32_t bits_per_sample = 2;
    double LO_offset = 0;
    int tsys_freq = 80;
  };
  typedef std::vector<Station_parameters> Station_list;

  Time experiment_start, start_time, stop_time, integration_time, sub_integration_time;
  int32_t number_channels = 0, fft_size_delaycor = 0, fft_size_correlation = 0;
  int32_t integration_nr = -1, slice_nr = -1, slice_offset = -1;
  int32_t sample_rate = 0;
  int64_t channel_freq = 0;
  int32_t bandwidth = 0;
  char sideband = 'n';
  int32_t frequency_nr = -1;
  char polarisation = 'n';
  bool cross_polarize = true;
  int32_t reference_station = -1;
  Station_list station_streams;
  int window = SFXC_WINDOW_RECT;
  char source[11] = {0};
  int32_t n_phase_centers = 1;
  int32_t pulsar_binning = 0;
  Pulsar_parameters *pulsar_parameters = nullptr;
  Mask_parameters *mask_parameters = nullptr;
};

// A station's delay model for a slice. The reference builds it from the
// .del table (Delay_table::create_akima_spline, src/delay_table_akima.cc:
// 430-502): GSL Akima splines of the delay, phase and amplitude over 1 s
// samples, plus the station clock (offset and rate from clock_epoch).
// from_samples restates GSL 1.9's non-periodic Akima interpolation
// (interpolation/akima.c: akima_init, akima_calc): slopes m_i of the
// segments, extended by m_-2 = 3 m_0 - 2 m_1, m_-1 = 2 m_0 - m_1 and the same
// at the far end; the derivative at a point weighs the neighbouring slopes by
// the changes of the slopes on the other side, |m_i+1 - m_i| and
// |m_i-1 - m_i-2|, and a segment whose weights both vanish is linear.
class Delay_table_akima {
 public:
  struct Spline {  // per interval i: y_i + b_i u + c_i u^2 + d_i u^3, u = x - x_i
    std::vector<double> x, y, b, c, d;
    double eval(double sec) const {
      size_t i = interval(sec);
      const double u = sec - x[i];
      return y[i] + u * (b[i] + u * (c[i] + u * d[i]));
    }
    double deriv(double sec) const {
      size_t i = interval(sec);
      const double u = sec - x[i];
      return b[i] + u * (2 * c[i] + u * 3 * d[i]);
    }
    size_t interval(double sec) const {
      size_t lo = 0, hi = x.size() - 1;  // x[lo] <= sec < x[hi] (clamped to the ends)
      while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        if (x[mid] > sec) hi = mid;
        else lo = mid;
      }
      return lo;
    }
  };
  struct Centre {
    std::string source;
    Spline delay, phase, amplitude;
  };

  Time scan_begin, interval_begin, interval_end;
  Time clock_epoch;
  double clock_offset = 0, clock_rate = 0;
  std::vector<Centre> centres;

  static Spline akima(const std::vector<double> &x, const std::vector<double> &y) {
    const size_t n = x.size();
    if (n < 5 || y.size() != n) throw std::invalid_argument("an Akima spline takes at least 5 points");
    std::vector<double> mm(n + 3);
    double *m = mm.data() + 2;  // m[-2] .. m[n]
    for (size_t i = 0; i + 1 < n; i++) m[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    m[-2] = 3.0 * m[0] - 2.0 * m[1];
    m[-1] = 2.0 * m[0] - m[1];
    m[n - 1] = 2.0 * m[n - 2] - m[n - 3];
    m[n] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
    Spline s;
    s.x = x;
    s.y = y;
    s.b.assign(n, 0.0);
    s.c.assign(n, 0.0);
    s.d.assign(n, 0.0);
    // the weighted slope at point i from the slopes around it
    auto slope_at = [&](size_t i, double weight_sum) {
      const double w_right = fabs(m[(long)i - 1] - m[(long)i - 2]) / weight_sum;
      return (1.0 - w_right) * m[(long)i - 1] + w_right * m[i];
    };
    for (size_t i = 0; i + 1 < n; i++) {
      const double sum = fabs(m[i + 1] - m[i]) + fabs(m[(long)i - 1] - m[(long)i - 2]);
      if (sum == 0.0) {  // the segment is straight
        s.b[i] = m[i];
        continue;
      }
      const double h = x[i + 1] - x[i];
      const double sum_next = fabs(m[i + 2] - m[i + 1]) + fabs(m[i] - m[(long)i - 1]);
      const double t_next = sum_next == 0.0 ? m[i] : slope_at(i + 1, sum_next);
      s.b[i] = slope_at(i, sum);
      s.c[i] = (3.0 * m[i] - 2.0 * s.b[i] - t_next) / h;
      s.d[i] = (s.b[i] + t_next - 2.0 * m[i]) / (h * h);
    }
    return s;
  }

  // One phase centre per entry of delays / phases / amplitudes (each sampled
  // at times_sec, seconds since scan_begin, 1 s apart; empty phases mean 0,
  // empty amplitudes 1).
  static Delay_table_akima from_samples(Time scan_begin, const std::vector<double> &times_sec,
                                        const std::vector<std::vector<double>> &delays,
                                        const std::vector<std::vector<double>> &phases,
                                        const std::vector<std::vector<double>> &amplitudes, Time clock_epoch,
                                        double clock_offset, double clock_rate,
                                        const std::vector<std::string> &sources = {}) {
    Delay_table_akima t;
    t.scan_begin = scan_begin;
    t.interval_begin = scan_begin + Time::from_ticks((int64_t)llround(times_sec.front() * Time::kClockRate));
    t.interval_end = scan_begin + Time::from_ticks((int64_t)llround(times_sec.back() * Time::kClockRate));
    t.clock_epoch = clock_epoch;
    t.clock_offset = clock_offset;
    t.clock_rate = clock_rate;
    for (size_t k = 0; k < delays.size(); k++) {
      Centre c;
      c.source = k < sources.size() ? sources[k] : std::string();
      c.delay = akima(times_sec, delays[k]);
      c.phase = akima(times_sec, k < phases.size() && !phases[k].empty() ? phases[k]
                                                                         : std::vector<double>(times_sec.size(), 0.0));
      c.amplitude = akima(times_sec, k < amplitudes.size() && !amplitudes[k].empty()
                                         ? amplitudes[k]
                                         : std::vector<double>(times_sec.size(), 1.0));
      t.centres.push_back(c);
    }
    return t;
  }

  int n_phase_centers() const { return (int)centres.size(); }
  const std::string &get_source(int phase_center) const { return centres[phase_center].source; }
  // Delay_table_akima::delay / rate / phase / amplitude (src/delay_table_akima.cc:114-166)
  double delay(const Time &time, int phase_center = 0) const {
    return centres[phase_center].delay.eval(time.diff(scan_begin)) + clock_offset +
           time.diff(clock_epoch) * clock_rate;
  }
  double rate(const Time &time, int phase_center = 0) const {
    return centres[phase_center].delay.deriv(time.diff(scan_begin)) + clock_rate;
  }
  double phase(const Time &time, int phase_center = 0) const {
    return centres[phase_center].phase.eval(time.diff(scan_begin));
  }
  double amplitude(const Time &time, int phase_center = 0) const {
    return centres[phase_center].amplitude.eval(time.diff(scan_begin));
  }

  // The C ABI's form of phase centre pc: the same cubic pieces on the 1 s
  // grid from the first sample, the clock folded into the delay's constant
  // and linear terms. coef receives delay, phase and amplitude coefficients
  // (3 x 4 n doubles) and must outlive the model.
  sfxg_delay_model model(int pc, std::vector<double> &coef) const {
    const Centre &c = centres[pc];
    const size_t n = c.delay.x.size() - 1;
    const double step = c.delay.x[1] - c.delay.x[0];
    for (size_t i = 1; i < n; i++)
      if (fabs((c.delay.x[i + 1] - c.delay.x[i]) - step) > 1e-9)
        throw std::invalid_argument("the C ABI takes delay samples on a uniform grid");
    coef.assign(12 * n, 0.0);
    for (size_t i = 0; i < n; i++) {
      const Spline *sp[3] = {&c.delay, &c.phase, &c.amplitude};
      for (int k = 0; k < 3; k++) {
        double *q = coef.data() + k * 4 * n + 4 * i;
        q[0] = sp[k]->y[i];
        q[1] = sp[k]->b[i];
        q[2] = sp[k]->c[i];
        q[3] = sp[k]->d[i];
      }
      const Time t_i = scan_begin + Time::from_ticks((int64_t)llround(c.delay.x[i] * Time::kClockRate));
      coef[4 * i] += clock_offset + t_i.diff(clock_epoch) * clock_rate;
      coef[4 * i + 1] += clock_rate;
    }
    sfxg_delay_model m;
    memset(&m, 0, sizeof(m));
    m.scan_begin = (scan_begin + Time::from_ticks((int64_t)llround(c.delay.x[0] * Time::kClockRate))).get_clock_ticks();
    m.interval = step;
    m.n_pieces = (int32_t)n;
    m.delay = coef.data();
    m.phase = coef.data() + 4 * n;
    m.amplitude = coef.data() + 8 * n;
    return m;
  }
};

// bit_statistics (include/bit_statistics.h:16-34): the levels of the last
// slice as integration_write stores them, and the tsys counts.
class bit_statistics {
 public:
  int *get_statistics() { return statistics.data(); }
  int *get_tsys() { return tsys.data(); }
  std::vector<int> statistics = std::vector<int>(5, 0);
  std::vector<int> tsys = std::vector<int>(4, 0);
};
typedef std::shared_ptr<bit_statistics> bit_statistics_ptr;

// Data_writer::put_bytes (include/data_writer.h:28), the output node's end.
class Data_writer {
 public:
  virtual ~Data_writer() = default;
  size_t put_bytes(size_t nBytes, const char *buff) { return do_put_bytes(nBytes, buff); }

 protected:
  virtual size_t do_put_bytes(size_t nBytes, const char *buff) = 0;
};

// Correlator_node_types (include/correlator_node_types.h:14-56), the pieces
// the wiring uses: the reader's byte buffer per stream, and the queues
// between the workers, whose elements here are a stream's complete HEADER
// byte stream for one slice.
struct Correlator_node_types {
  struct Invalid {
    int start;
    int n_invalid;
  };
  struct Channel_circular_input_buffer {
    std::deque<uint8_t> bytes;
    void write(const uint8_t *p, size_t n) { bytes.insert(bytes.end(), p, p + n); }
  };
  typedef std::shared_ptr<Channel_circular_input_buffer> Channel_circular_input_buffer_ptr;
  typedef std::shared_ptr<std::vector<uint8_t>> Slice_stream;
  struct Channel_queue {
    std::deque<Slice_stream> q;
    bool empty() const { return q.empty(); }
    void push(const Slice_stream &s) { q.push_back(s); }
    Slice_stream front() const { return q.front(); }
    void pop() { q.pop_front(); }
  };
  typedef std::shared_ptr<Channel_queue> Channel_queue_ptr;
  typedef Channel_queue Delay_queue;
  typedef Channel_queue_ptr Delay_queue_ptr;
};
