// Test harness: exposes the drop-in Delay_table_akima (sfxc-1_amd/host/
// sfxc_types.h) to Python so its Akima construction and the C ABI model it
// produces can be checked on the CPU. Built by tests/test_dropin.py with g++.
#include "../../sfxc-1_amd/host/sfxc_types.h"

extern "C" {

// Akima spline through (x, y), evaluated at xs (value, derivative).
int ak_eval(const double *x, const double *y, int n, const double *xs, int m, double *val, double *der) {
  try {
    std::vector<double> vx(x, x + n), vy(y, y + n);
    Delay_table_akima::Spline s = Delay_table_akima::akima(vx, vy);
    for (int i = 0; i < m; i++) {
      val[i] = s.eval(xs[i]);
      der[i] = s.deriv(xs[i]);
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

// A delay table from 1 s samples (times since scan_begin) with a clock, and
// its C ABI model: delay(t) at the ticks in ts, and the model's pieces
// (scan_begin, interval, n, then 4 n delay coefficients).
int ak_table(int64_t scan_begin, const double *t, const double *y, int n, int64_t clock_epoch, double clock_offset,
             double clock_rate, const int64_t *ts, int m, double *delay, double *rate, int64_t *mb, double *minterval,
             int32_t *mn, double *coef) {
  try {
    std::vector<double> vt(t, t + n), vy(y, y + n);
    Delay_table_akima d = Delay_table_akima::from_samples(Time::from_ticks(scan_begin), vt, {vy}, {}, {},
                                                          Time::from_ticks(clock_epoch), clock_offset, clock_rate);
    for (int i = 0; i < m; i++) {
      delay[i] = d.delay(Time::from_ticks(ts[i]));
      rate[i] = d.rate(Time::from_ticks(ts[i]));
    }
    std::vector<double> c;
    sfxg_delay_model md = d.model(0, c);
    *mb = md.scan_begin;
    *minterval = md.interval;
    *mn = md.n_pieces;
    for (int i = 0; i < 4 * md.n_pieces; i++) coef[i] = md.delay[i];
  } catch (...) {
    return -1;
  }
  return 0;
}
}
