/*
 * sfxc_oracle.h -- CPU restatement of the SFXC correlator-node data path.
 *
 * TEST INFRASTRUCTURE. This is the checker, not the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and the
 * shipped library never links it.
 *
 * sfxc_oracle.c is compiled twice: with OR_REAL=float (the reference's default
 * FLOAT=float build, sfxc/include/utils.h:192-204) into liboracle_f32.so and
 * with OR_REAL=double (its --enable-double build, utils.h:180-191) into
 * liboracle_f64.so. Every exported symbol carries the suffix _f32 or _f64.
 *
 * Pinning (which independent check fixes each function) is listed in the
 * header comment of sfxc_oracle.c next to each function.
 */
#ifndef SFXC_ORACLE_H
#define SFXC_ORACLE_H
#include <stdint.h>

/* Header codes of the input-node -> correlator-node byte stream
 * (sfxc/include/utils.h:39-42). */
#define OR_HEADER_DATA 0
#define OR_HEADER_INVALID 1
#define OR_HEADER_DELAY 2
#define OR_HEADER_ENDSTREAM 3

/* Window functions (sfxc/include/utils.h:209-213). */
#define OR_WINDOW_NONE 0
#define OR_WINDOW_RECT 1
#define OR_WINDOW_COS 2
#define OR_WINDOW_HAMMING 3
#define OR_WINDOW_HANN 4

/* Parameters of one Bit2float_worker for one time slice
 * (Bit2float_worker::set_new_parameters, bit2float_worker.cc:235-276). */
typedef struct {
  int32_t bits_per_sample;
  int32_t sample_rate;        /* station stream sample rate */
  int32_t base_sample_rate;   /* correlation_parameters.sample_rate */
  int32_t tsys_freq;
  int32_t fft_size;           /* fft_size_delaycor, N_d */
  int32_t nfft_max;           /* n_ffts_per_buffer */
  int32_t n_ffts_per_integration;
  int32_t delay_in_samples;   /* int(delay(start_time) * sample_rate) */
  int32_t lut;                /* 0: reference table, 1: alternative decode table */
  int32_t pad;
} or_b2f_params;

/* Piecewise-cubic model on a uniform grid, the form a GSL Akima spline over
 * a 1 s delay table takes (delay_table_akima.cc:114-166). coef[4*i+k] is the
 * k-th power coefficient of piece i in (t - t_i) seconds. */
typedef struct {
  int64_t scan_begin_ticks;   /* 512 MHz ticks (correlator_time.h:12-64) */
  double interval;            /* seconds per piece */
  int32_t n;                  /* number of pieces */
  int32_t pad;
  const double *delay;        /* seconds */
  const double *phase;        /* radians; NULL means 0 */
  const double *amp;          /* NULL means 1 */
} or_delay_model;

/* Parameters of one Delay_correction for one time slice
 * (Delay_correction::set_parameters, delay_correction.cc:186-259). */
typedef struct {
  int32_t fft_size_delaycor;    /* N_d */
  int32_t fft_size_correlation; /* N_c */
  int32_t sample_rate;          /* correlation_parameters.sample_rate */
  int32_t station_sample_rate;
  int32_t bandwidth;            /* station bandwidth, Hz */
  int32_t sideband;             /* station sideband, +1 'U', -1 'L' */
  int32_t job_sideband;         /* correlation_parameters.sideband */
  int32_t window;
  int64_t channel_freq;         /* station channel frequency, Hz */
  int64_t job_channel_freq;
  double LO_offset;
  int64_t start_ticks;
  int64_t experiment_start_ticks;
  int32_t n_ffts_per_integration;
  int32_t stride;               /* complex elements per output segment (N_c + 4) */
} or_dc_params;

/* Parameters of Correlation_core for one integration
 * (correlation_core.cc:100-171). */
typedef struct {
  int32_t n_streams;
  int32_t fft_size;                 /* fft_size_correlation, N_c */
  int32_t stride;
  int32_t n_baselines;
  int32_t number_ffts_in_integration;
  int32_t number_ffts_in_sub_integration;
  const int32_t *bl_first;          /* stream index of the first station */
  const int32_t *bl_second;
} or_core_params;

/* Pulsar gating parameters derived by Correlation_core_pulsar::set_parameters
 * (correlation_core_pulsar.cc:46-99). */
typedef struct {
  int32_t nbins;          /* pulsar.nbins + 1 */
  int32_t pad;
  double start_phase;
  double phase_coef;      /* coef[1] + 60 * ref_freq */
  double fft_duration;    /* microseconds */
  double gate_begin, gate_end;
  const double *offsets;  /* N_c + 1 dispersion offsets, pulsar periods */
} or_pulsar_params;

/* Extra phase centres of Correlation_core (correlation_core.cc:515-592):
 * what sub_integration and uvshift read besides the accumulators. */
typedef struct {
  int32_t n_phase_centers;
  int32_t sideband;              /* +1 'U', -1 'L' (correlation_parameters.sideband) */
  int32_t sample_rate;           /* correlation_parameters.sample_rate */
  int32_t pad;
  int64_t start_ticks;           /* correlation_parameters.start_time */
  int64_t channel_freq;          /* correlation_parameters.channel_freq, Hz */
  const or_delay_model *models;  /* n_phase_centers x n_streams, centre-major; centre 0 the pointing centre */
  const float *weights;          /* fft_size values of create_weights (correlation_core.cc:712-759) */
} or_pc_params;

/* Lag re-channelisation of integration_write when fft_size_correlation >
 * number_channels (correlation_core.cc:417-440): Mask_parameters
 * (control_parameters.h:107-114) and the tables create_mask / create_window
 * build from them (correlation_core.cc:761-824). */
typedef struct {
  int32_t fft_size;       /* N_c */
  int32_t nchan;          /* number_channels */
  int32_t normalize;      /* Mask_parameters::normalize */
  int32_t pad;
  const float *mask;      /* N_c + 1 values */
  const float *window;    /* 2 * nchan values */
} or_rechan_params;

#endif
