"""Pins the oracle's extra phase centres (sub_integration's uv shift,
correlation_core.cc:515-592, create_weights :712-759, Delay_table_akima::rate
delay_table_akima.cc:128-136) against numpy and closed forms. CPU only."""
import copy

import numpy as np
import pytest

from oracle import oracle as O
from sfxg import configs
from sfxg.job import (CLOCK_RATE, WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN, WINDOW_NONE, WINDOW_RECT, DelayModel,
                      PhaseCenter)


def window_fn(kind, n, i):
    i = np.asarray(i, dtype=np.float64)
    if kind in (WINDOW_NONE, WINDOW_RECT):
        return ((i >= n // 4) & (i < (3 * n) // 4)).astype(np.float64)
    if kind == WINDOW_COS:
        return np.sin(np.pi * i / n)
    if kind == WINDOW_HAMMING:
        return 0.54 - 0.46 * np.cos(2 * np.pi * i / n)
    return 0.5 * (1 - np.cos(2 * np.pi * i / n))


@pytest.mark.parametrize("kind", [WINDOW_NONE, WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN])
@pytest.mark.parametrize("n", [64, 1024])
def test_pc_weights_autocorrelation(kind, n):
    w = window_fn(kind, 2 * n - 1, np.arange(2 * n))
    r = np.correlate(w, w, "full")[2 * n - 1:]
    ref = r[:n] / r[0]
    np.testing.assert_allclose(O.pc_weights(kind, n, "f64"), ref, rtol=0, atol=1e-7)  # stored as float
    np.testing.assert_allclose(O.pc_weights(kind, n, "f32"), ref, rtol=0, atol=2e-6)


def test_rate_is_derivative():
    m = DelayModel(1000, 1.0, np.array([[1e-3, 2e-7, 3e-9, -4e-10], [1.1e-3, 2.5e-7, -1e-9, 2e-10]]))
    for sec in (0.1, 0.5, 1.3, 1.9):
        t = 1000 + int(sec * CLOCK_RATE)
        h = 1e-4
        fd = (O.eval_delay(m, t + int(h * CLOCK_RATE)) - O.eval_delay(m, t - int(h * CLOCK_RATE))) / (2 * h)
        assert abs(O.eval_rate(m, t) - fd) < 1e-12


def pc_job(extra_offsets, sub_usec=62_500, window=WINDOW_NONE, nchan=512):
    job, hdrs = configs.build_slices(seed=41, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 2.5e-6, -1.75e-6], rates=[0.0, 2e-7, -1e-7], window=window,
                                     sub_int_usec=sub_usec)[0]
    pcs = []
    for k, offs in enumerate(extra_offsets):
        ms = []
        for i, m in enumerate(job.delay_models):
            c = m.delay.copy()
            c[:, 0] += offs[i]
            ms.append(DelayModel(m.scan_begin_ticks, m.interval, c))
        pcs.append(PhaseCenter(ms, source_nr=k + 1))
    job.phase_centers = pcs
    return job, hdrs


def test_zero_shift_copies_pointing_centre():
    job, hdrs = pc_job([[0.0, 0.0, 0.0]], sub_usec=20_000)
    res = O.run_slice(job, hdrs, lut=1)
    np.testing.assert_array_equal(res["vis"][1], res["vis"][0])
    single = copy.copy(job)
    single.phase_centers = []
    res1 = O.run_slice(single, hdrs, lut=1)
    np.testing.assert_array_equal(res["vis"][0], res1["vis"][0])


@pytest.mark.parametrize("window", [WINDOW_NONE, WINDOW_HANN])
def test_one_sub_integration_rotation(window):
    """With one sub-integration the shifted centre is the pointing centre's
    cross-correlation times amplitude * exp(i (phi + k delta)), phi and delta
    from the delay differences at the integration's middle; autos are copied."""
    offs = [0.0, 3.0e-8, -2.0e-8]
    job, hdrs = pc_job([offs], window=window)
    res = O.run_slice(job, hdrs, lut=1, precision="f64")
    v0, v1 = res["vis"][0], res["vis"][1]
    M = len(job.station_streams)
    np.testing.assert_array_equal(v1[:M], v0[:M])
    n = job.number_ffts_in_integration()
    tfft = job.fft_size_correlation * round(CLOCK_RATE / job.sample_rate)
    tmid = job.start_time + int(round(tfft * (n / 2.0)))
    w = O.pc_weights(window, job.fft_size_correlation, "f64")
    k = np.arange(job.fft_size_correlation + 1)
    for b, (s1, s2) in enumerate(job.baselines()):
        if s1 == s2:
            continue
        m0, m1 = job.delay_models, job.phase_centers[0].delay_models
        dd1 = O.eval_delay(m1[s1], tmid) - O.eval_delay(m0[s1], tmid)
        dd2 = O.eval_delay(m1[s2], tmid) - O.eval_delay(m0[s2], tmid)
        r1, r2 = O.eval_rate(m0[s1], tmid), O.eval_rate(m0[s2], tmid)
        x = dd1 * (1 - r1) - dd2 * (1 - r2)
        phi = job.channel_freq * x
        phi = 2 * np.pi * (phi - np.floor(phi))
        delta = 2 * np.pi * job.sample_rate / (2.0 * job.fft_size_correlation) * x
        lag = abs(int(round((dd1 - dd2) * job.sample_rate)))
        amp = 1.0 / float(w[lag]) if lag < len(w) and w[lag] > 1e-4 else 1.0
        ref = amp * (v0[b, :, 0] + 1j * v0[b, :, 1]) * np.exp(1j * (phi + k * delta))
        got = v1[b, :, 0] + 1j * v1[b, :, 1]
        assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def lag_response(v, L, nc):
    """sum_k V[k] exp(+2 pi i k L / 2N_c) / (N_c + 1): the cross-correlation at
    a residual delay of L samples."""
    k = np.arange(nc + 1)
    return np.sum(v * np.exp(2j * np.pi * k * L / (2 * nc))) / (nc + 1)


def test_second_source_is_stopped_in_its_phase_centre():
    """A second common signal, delayed by a few samples more than the
    pointing centre's model at two stations, is phase-stopped at lag 0 in a
    phase centre whose delays are the second signal's; the pointing centre
    keeps the first signal at lag 0 and shows the second at its residual lag.
    A wrong sign of the shift's phase or slope leaves the second signal at
    twice its lag or with a phase that is not zero."""
    fs, nc = 32_000_000, 512
    offsets, rates = [0.0, 2.5e-6, -1.75e-6], [0.0, 2e-7, -1e-7]
    dlt = [0.0, 8 / fs, -6 / fs]
    job, hdrs = configs.build_slices(seed=43, n_stations=3, pols=["R"], fs=fs, bw=fs // 2, nchan=nc,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=offsets, rates=rates, window=WINDOW_NONE, sub_int_usec=15_625,
                                     extra=(0.1, [o + d for o, d in zip(offsets, dlt)], rates))[0]
    ms = []
    for i, m in enumerate(job.delay_models):
        c = m.delay.copy()
        c[:, 0] += dlt[i]
        ms.append(DelayModel(m.scan_begin_ticks, m.interval, c))
    job.phase_centers = [PhaseCenter(ms, source_nr=7)]
    res = O.run_slice(job, hdrs, lut=1)
    n0, n1 = res["norm"][0], res["norm"][1]
    for b, (s1, s2) in enumerate(job.baselines()):
        if s1 == s2:
            continue
        v0 = n0[b, :, 0] + 1j * n0[b, :, 1]
        v1 = n1[b, :, 0] + 1j * n1[b, :, 1]
        L = round((dlt[s1] - dlt[s2]) * fs)  # the second signal's residual lag at the pointing centre
        a0, a1 = lag_response(v0, 0, nc), lag_response(v1, 0, nc)
        r0, r1 = lag_response(v0, L, nc), lag_response(v1, -L, nc)
        print(b, L, abs(a0), np.angle(a0), abs(r0), abs(a1), np.angle(a1), abs(r1), abs(lag_response(v1, 2 * L, nc)))
        # pointing centre: the first signal (rho 0.05) at lag 0, the second (0.1) at L
        assert abs(np.angle(a0)) < 0.05 and 0.03 < abs(a0) < 0.06
        assert 0.06 < abs(r0) < 0.12
        # shifted centre: the second signal at lag 0 with zero phase, the first at -L
        assert abs(np.angle(a1)) < 0.05 and 0.06 < abs(a1) < 0.12
        assert 0.03 < abs(r1) < 0.06
