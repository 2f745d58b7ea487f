"""Builders for the seeded synthetic correlation jobs.

Each builder returns a list of slices; a slice is (CorrelationParameters,
[HEADER byte stream per station stream]) for one subband and one integration,
the unit the manager node hands to one correlator node
(sfxc/src/manager_node.cc:288-391).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import synth
from .job import (CLOCK_RATE, WINDOW_HANN, CorrelationParameters, DelayModel, StationStream,
                  mjd_to_ticks)

START_MJD = 60000


def build_slices(*, seed: int, n_stations: int, pols: Sequence[str], fs: int, bw: int, nchan: int,
                 seconds: float, t_int_usec: int, rho: float, bits: int, sky_freq: int,
                 offsets: Sequence[float], rates: Sequence[float], subband: int = 0,
                 window: int = WINDOW_HANN, sub_int_usec: Optional[int] = None,
                 extra: Optional[Tuple[float, Sequence[float], Sequence[float]]] = None,
                 fft_size_delaycor: Optional[int] = None, sideband: str = "U",
                 invalid: Optional[dict] = None, n_integrations: Optional[int] = None,
                 raw_out: Optional[dict] = None, margin_multiple: int = 1, pulsar=None,
                 whole_streams: bool = False) -> List[Tuple[CorrelationParameters, List[bytes]]]:
    """Synthesise the raw streams of one subband and cut them into slices.

    offsets/rates: per-station delay polynomial tau = offset + rate * t.
    extra: (rho2, offsets2, rates2) adds a second common component.
    invalid: {stream_index: [(byte_start_in_slice_raw, nbytes), ...]} marks
    input bytes invalid (relative to the first byte sent for each slice).
    raw_out: if given, receives the raw packed streams ("raw", one per stream)
    and their first sample's time ("raw_start", ticks).
    margin_multiple: the raw streams start a multiple of this many samples
    before the first slice (frame-aligned raw data for VDIF frames).
    pulsar: (Pulsar, peak, duty) injects a dispersed pulsed component
    (synth.synth_station_streams).
    whole_streams: synthesise every raw sample; otherwise only those the
    slices read (the rest of each raw stream is zero bytes).
    """
    start = mjd_to_ticks(START_MJD, 0.0)
    tps = int(round(CLOCK_RATE / fs))
    t_int_ticks = int(round(t_int_usec * (CLOCK_RATE / 1e6)))
    n_int = n_integrations or int(round(seconds * 1e6 / t_int_usec))
    total_sec = n_int * t_int_usec / 1e6
    # margin before the first slice: covers the most negative delay
    max_neg = max(0.0, -min(min(offsets), min(o + r * total_sec for o, r in zip(offsets, rates))))
    max_pos = max(0.0, max(max(offsets), max(o + r * total_sec for o, r in zip(offsets, rates))))
    spb = 8 // bits
    margin = int(math.ceil(max_neg * fs)) + 4096
    margin -= margin % (4 * spb)
    step = 4 * spb * margin_multiple // math.gcd(4 * spb, margin_multiple)
    margin = -(-margin // step) * step
    raw_start = start - margin * tps
    n_raw = margin + int(total_sec * fs) + int(math.ceil(max_pos * fs)) + 8192
    n_raw += (-n_raw) % (4096 * spb)
    seconds_model = int(math.ceil(total_sec + (margin / fs) + 2))
    st_models = [DelayModel.linear(raw_start, o + r * (raw_start - start) / CLOCK_RATE, r, seconds_model)
                 for o, r in zip(offsets, rates)]
    # shift the models so that t is seconds since `start` in the polynomial:
    # tau(t) = o + r * (t - start); pieces start at raw_start
    streams = []
    models = []
    for s in range(n_stations):
        for p in pols:
            streams.append(StationStream(station_number=s, station_stream=len(streams), sample_rate=fs,
                                         channel_freq=sky_freq, bandwidth=bw, sideband=sideband,
                                         polarisation=p, bits_per_sample=bits))
            models.append(st_models[s])
    groups = []
    # one common signal per polarisation: streams are ordered station-major
    stream_models_by_pol = {p: [] for p in pols}
    for i, st in enumerate(streams):
        stream_models_by_pol[st.polarisation].append(i)
    order = [i for p in pols for i in stream_models_by_pol[p]]
    extra_arg = None
    if extra is not None:
        rho2, o2, r2 = extra
        m2 = [DelayModel.linear(raw_start, o + r * (raw_start - start) / CLOCK_RATE, r, seconds_model)
              for o, r in zip(o2, r2)]
        extra_arg = [(rho2, [m2[streams[i].station_number] for i in order])]
    active = None
    if not whole_streams and raw_out is None:
        # samples the slices read: from the first slice's start to the last
        # one's end, moved by each stream's integer delay over that time
        t_end = start + n_int * t_int_ticks
        tt = np.linspace(start, t_end, 64).astype(np.int64)
        active = []
        for i in order:
            n = np.floor(models[i].eval(tt) * fs + 0.5)
            a0 = (start - raw_start) // tps + int(n.min()) - 8192
            a1 = (t_end - raw_start) // tps + int(n.max()) + 8192
            active.append((a0, a1 - a0))
    raw_ordered = synth.synth_station_streams(seed, [len(stream_models_by_pol[p]) for p in pols], fs, n_raw,
                                              [models[i] for i in order], raw_start, rho, bits, sky_freq, extra_arg,
                                              active=active, pulsar=pulsar)
    raw = [None] * len(streams)
    for k, i in enumerate(order):
        raw[i] = raw_ordered[k]
    slices = []
    nc = max(256, nchan)
    nd = fft_size_delaycor or min(256, nc)
    for k in range(n_int):
        t0 = start + k * t_int_ticks
        job = CorrelationParameters(
            experiment_start=start, start_time=t0, stop_time=t0 + t_int_ticks,
            integration_time=t_int_ticks,
            sub_integration_time=int(round(min(t_int_usec, sub_int_usec or 125000) * (CLOCK_RATE / 1e6))),
            number_channels=nchan, fft_size_delaycor=nd, fft_size_correlation=nc, sample_rate=fs,
            channel_freq=sky_freq, bandwidth=bw, sideband=sideband, frequency_nr=subband,
            station_streams=streams, integration_nr=k, slice_nr=k, window=window,
            delay_models=models)
        nsamp = job.n_ffts_per_integration(0) * nd
        hdrs = []
        for i in range(len(streams)):
            inv = None
            if invalid and i in invalid:
                dsamples = (t0 - raw_start) // tps
                base = dsamples * bits // 8
                inv = [(base + a, n) for a, n in invalid[i]]
            hdrs.append(synth.header_stream(raw[i], raw_start, models[i], t0, nsamp, fs, bits, inv))
        slices.append((job, hdrs))
    if raw_out is not None:
        raw_out["raw"] = raw
        raw_out["raw_start"] = raw_start
    return slices


def config0(seconds: float = 2.0, t_int_usec: int = 1_000_000, nchan: int = 1024):
    """2 stations, 1 subband, 1 pol, 16 MHz USB at 32 Ms/s, 4.974 GHz, 2-bit,
    rho 0.05, tau_2 = 1.234567 us + 1e-7 s/s * t, 1024 channels, 1 s."""
    return build_slices(seed=0, n_stations=2, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                        seconds=seconds, t_int_usec=t_int_usec, rho=0.05, bits=2, sky_freq=4_974_000_000,
                        offsets=[0.0, 1.234567e-6], rates=[0.0, 1e-7])


def config1_jobs(subband: int, n_integrations: int = 10, t_int_usec: int = 1_000_000, nchan: int = 4096,
                 n_stations: int = 8):
    """Job descriptions of configuration 1 (no sample data): 8 stations x 2
    pols, 32 MHz at 64 Ms/s, sky 6.0 GHz + 32 MHz * subband, 2-bit, 4096
    channels, 1 s integrations, delay offsets U[-5, 5] ms and rates
    U[-3e-6, 3e-6] (seed 1). Returns (jobs, raw_start_ticks)."""
    return survey_jobs(seed=1, subband=subband, n_integrations=n_integrations, t_int_usec=t_int_usec,
                       nchan=nchan, n_stations=n_stations, fs=64_000_000, bw=32_000_000,
                       sky=6_000_000_000 + 32_000_000 * subband)


def config2_jobs(subband: int, n_integrations: int = 8, t_int_usec: int = 500_000, nchan: int = 16384,
                 n_stations: int = 32):
    """Job descriptions of configuration 2 (no sample data): 32 stations x 2
    pols, 64 MHz at 128 Ms/s, 2-bit, 16384 channels, 0.5 s integrations (4 s),
    delay offsets U[-5, 5] ms and rates U[-3e-6, 3e-6] (seed 2; the
    configuration leaves them and the sky frequency open: 6.0 GHz + 64 MHz *
    subband). Returns (jobs, raw_start_ticks)."""
    return survey_jobs(seed=2, subband=subband, n_integrations=n_integrations, t_int_usec=t_int_usec,
                       nchan=nchan, n_stations=n_stations, fs=128_000_000, bw=64_000_000,
                       sky=6_000_000_000 + 64_000_000 * subband)


def config3_jobs(n_jobs: int = 48, seconds: float = 2.0, seed: int = 3):
    """Configuration 3 (a mixed batch, no sample data): n_jobs seeded jobs of
    one subband each, drawn per job: channels from {256, 1024, 4096, 16384,
    32768}, stations uniform in 4..24 (one polarisation), 1- or 2-bit,
    bandwidth from {8, 16, 32, 64} MHz (real sampling at twice that), integration
    time from {0.25, 0.5, 1, 2} s, rho uniform in [0.005, 0.05], delay offsets
    U[-5, 5] ms and rates U[-3e-6, 3e-6]; `seconds` of data per job. Returns
    a list of dicts: jobs (one per integration), raw_start, rho and the draws."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n_jobs):
        nchan = int(rng.choice([256, 1024, 4096, 16384, 32768]))
        n_st = int(rng.integers(4, 25))
        bits = int(rng.choice([1, 2]))
        bw = int(rng.choice([8, 16, 32, 64])) * 1_000_000
        t_int = float(rng.choice([0.25, 0.5, 1.0, 2.0]))
        rho = float(rng.uniform(0.005, 0.05))
        offsets = rng.uniform(-5e-3, 5e-3, n_st)
        rates = rng.uniform(-3e-6, 3e-6, n_st)
        fs = 2 * bw
        sky = 5_000_000_000 + 64_000_000 * j
        start = mjd_to_ticks(START_MJD, 0.0)
        tps = int(round(CLOCK_RATE / fs))
        margin = int(math.ceil(5.1e-3 * fs)) + 4096
        margin -= margin % 16
        raw_start = start - margin * tps
        n_int = max(1, int(round(seconds / t_int)))
        secs = int(math.ceil(n_int * t_int + margin / fs + 2))
        models = [DelayModel.linear(raw_start, o + r * (raw_start - start) / CLOCK_RATE, r, secs)
                  for o, r in zip(offsets, rates)]
        streams = [StationStream(station_number=s, station_stream=s, sample_rate=fs, channel_freq=sky, bandwidth=bw,
                                 sideband="U", polarisation="R", bits_per_sample=bits) for s in range(n_st)]
        t_int_ticks = int(round(t_int * CLOCK_RATE))
        jobs = []
        for k in range(n_int):
            t0 = start + k * t_int_ticks
            jobs.append(CorrelationParameters(
                experiment_start=start, start_time=t0, stop_time=t0 + t_int_ticks, integration_time=t_int_ticks,
                sub_integration_time=int(round(min(t_int, 0.125) * CLOCK_RATE)), number_channels=nchan,
                fft_size_delaycor=256, fft_size_correlation=nchan, sample_rate=fs, channel_freq=sky, bandwidth=bw,
                sideband="U", frequency_nr=j, station_streams=streams, integration_nr=k, slice_nr=k,
                cross_polarize=False, delay_models=models))
        out.append(dict(jobs=jobs, raw_start=raw_start, rho=rho, nchan=nchan, stations=n_st, bits=bits, fs=fs,
                        t_int=t_int))
    return out


def config4_jobs(subband: int, n_integrations: int = 20, t_int_usec: int = 1_000_000, nchan: int = 1024,
                 n_stations: int = 12):
    """Job descriptions of configuration 4 (pulsar binning, no sample data):
    12 stations x 2 pols, 16 MHz at 32 Ms/s, 2-bit, 1024 channels, 1 s
    integrations (20 s), delay offsets U[-5, 5] ms and rates U[-3e-6, 3e-6]
    (seed 4). The configuration leaves the sky frequency open: 1.4 GHz + 16
    MHz * subband. Returns (jobs, raw_start_ticks, pulsar) with the pulsar of
    pulsar4()."""
    jobs, raw_start = survey_jobs(seed=4, subband=subband, n_integrations=n_integrations, t_int_usec=t_int_usec,
                                  nchan=nchan, n_stations=n_stations, fs=32_000_000, bw=16_000_000,
                                  sky=1_400_000_000 + 16_000_000 * subband)
    for j in jobs:
        j.pulsar_binning = True
    return jobs, raw_start, pulsar4(jobs[0])


def pulsar4(job):
    """Configuration 4's pulsar: period 0.0893 s, DM 68 pc/cm^3, a linear
    ephemeris (polyco with constant terms only), 64 bins over the gate
    [0.25, 0.75) of the pulse phase (bin 0 takes the off-gate segments),
    reference frequency 1.4 GHz."""
    from .job import Pulsar, ticks_to_mjd
    return Pulsar(nbins=64, gate_start=0.25, gate_stop=0.75, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                  ref_phase=0.0, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])


def survey_jobs(*, seed, subband, n_integrations, t_int_usec, nchan, n_stations, fs, bw, sky):
    rng = np.random.default_rng(seed)
    offsets = rng.uniform(-5e-3, 5e-3, n_stations)
    rates = rng.uniform(-3e-6, 3e-6, n_stations)
    start = mjd_to_ticks(START_MJD, 0.0)
    tps = int(round(CLOCK_RATE / fs))
    margin = int(math.ceil(5.1e-3 * fs)) + 4096
    margin -= margin % 16
    raw_start = start - margin * tps
    secs = int(math.ceil(n_integrations * t_int_usec / 1e6 + margin / fs + 2))
    st_models = [DelayModel.linear(raw_start, o + r * (raw_start - start) / CLOCK_RATE, r, secs)
                 for o, r in zip(offsets, rates)]
    streams, models = [], []
    for s in range(n_stations):
        for p in ("R", "L"):
            streams.append(StationStream(station_number=s, station_stream=len(streams), sample_rate=fs,
                                         channel_freq=sky, bandwidth=bw,
                                         sideband="U", polarisation=p, bits_per_sample=2))
            models.append(st_models[s])
    t_int_ticks = int(round(t_int_usec * (CLOCK_RATE / 1e6)))
    jobs = []
    for k in range(n_integrations):
        t0 = start + k * t_int_ticks
        jobs.append(CorrelationParameters(
            experiment_start=start, start_time=t0, stop_time=t0 + t_int_ticks, integration_time=t_int_ticks,
            sub_integration_time=int(round(min(t_int_usec, 125000) * (CLOCK_RATE / 1e6))),
            number_channels=nchan, fft_size_delaycor=256, fft_size_correlation=max(256, nchan), sample_rate=fs,
            channel_freq=sky, bandwidth=bw, sideband="U", frequency_nr=subband,
            station_streams=streams, integration_nr=k, slice_nr=k, delay_models=models))
    return jobs, raw_start


def config1_subband(subband: int, seconds: float = 10.0, t_int_usec: int = 1_000_000, nchan: int = 4096,
                    n_stations: int = 8):
    """8 stations x 2 pols, subband `subband` of 8 (32 MHz at 64 Ms/s, 6.0 GHz +
    32 MHz * subband), 2-bit, rho 0.02, offsets U[-5, 5] ms, rates U[-3e-6, 3e-6]."""
    rng = np.random.default_rng(1)
    offsets = rng.uniform(-5e-3, 5e-3, n_stations)
    rates = rng.uniform(-3e-6, 3e-6, n_stations)
    return build_slices(seed=1000 + subband, n_stations=n_stations, pols=["R", "L"], fs=64_000_000,
                        bw=32_000_000, nchan=nchan, seconds=seconds, t_int_usec=t_int_usec, rho=0.02,
                        bits=2, sky_freq=6_000_000_000 + 32_000_000 * subband, offsets=list(offsets),
                        rates=list(rates), subband=subband)


def second_source(n_stations: int, seed: int = 2):
    """Configuration 2's second point source: its own delay per station, the
    first source's polynomial moved by offsets U[-50, 50] us and rates
    U[-1e-7, 1e-7] s/s (drawn after the first source's delays from the same
    seeded generator). Returns (rho2, offsets2, rates2)."""
    rng = np.random.default_rng(seed)
    offsets = rng.uniform(-5e-3, 5e-3, n_stations)
    rates = rng.uniform(-3e-6, 3e-6, n_stations)
    d_off = rng.uniform(-50e-6, 50e-6, n_stations)
    d_rate = rng.uniform(-1e-7, 1e-7, n_stations)
    return 0.005, list(offsets + d_off), list(rates + d_rate)


def config2_slices(seconds: float = 0.0078125, n_stations: int = 32, nchan: int = 16384, subband: int = 0):
    """Configuration 2 with sample data, one integration of `seconds`: 32
    stations x 2 pols, 64 MHz at 128 Ms/s, 2-bit, 16384 channels, common
    signal rho 0.01 plus the second point source (rho 0.005, second_source),
    delays of config2_jobs (seed 2), sky 6.0 GHz + 64 MHz * subband."""
    rng = np.random.default_rng(2)
    offsets = rng.uniform(-5e-3, 5e-3, n_stations)
    rates = rng.uniform(-3e-6, 3e-6, n_stations)
    return build_slices(seed=2000 + subband, n_stations=n_stations, pols=["R", "L"], fs=128_000_000,
                        bw=64_000_000, nchan=nchan, seconds=seconds, t_int_usec=int(round(seconds * 1e6)),
                        rho=0.01, bits=2, sky_freq=6_000_000_000 + 64_000_000 * subband, offsets=list(offsets),
                        rates=list(rates), subband=subband, extra=second_source(n_stations))


def config3_slices(job_index: int, seconds: float = 0.03125, n_jobs: int = 48, seed: int = 3):
    """Job `job_index` of configuration 3's draws (config3_jobs) with sample
    data, one integration of `seconds` (its sub-integration too): channels,
    stations, bit depth, bandwidth, rho and delays as drawn."""
    rng = np.random.default_rng(seed)
    for j in range(n_jobs):
        nchan = int(rng.choice([256, 1024, 4096, 16384, 32768]))
        n_st = int(rng.integers(4, 25))
        bits = int(rng.choice([1, 2]))
        bw = int(rng.choice([8, 16, 32, 64])) * 1_000_000
        float(rng.choice([0.25, 0.5, 1.0, 2.0]))
        rho = float(rng.uniform(0.005, 0.05))
        offsets = rng.uniform(-5e-3, 5e-3, n_st)
        rates = rng.uniform(-3e-6, 3e-6, n_st)
        if j == job_index:
            usec = int(round(seconds * 1e6))
            return build_slices(seed=5000 + j, n_stations=n_st, pols=["R"], fs=2 * bw, bw=bw, nchan=nchan,
                                seconds=seconds, t_int_usec=usec, sub_int_usec=usec, rho=rho, bits=bits,
                                sky_freq=5_000_000_000 + 64_000_000 * j, offsets=list(offsets), rates=list(rates),
                                subband=j)
    raise ValueError(f"job {job_index} of {n_jobs}")


def config4_slices(seconds: float = 0.0625, n_stations: int = 12, subband: int = 0, nbins: int = 64):
    """Configuration 4 with sample data, one integration of `seconds`: 12
    stations x 2 pols, 16 MHz at 32 Ms/s, 2-bit, 1024 channels, continuum rho
    0.005, delays of config4_jobs (seed 4), and the injected pulsar: period
    0.0893 s, Gaussian pulse of 5 % duty (full width at half maximum) with
    peak power 0.2, dispersed with DM 68 across the band relative to its
    observing frequency (1.4 GHz). The ephemeris (pulsar4 with the reference
    phase moved so that the integration starts at pulse phase 0.2) bins 64
    phase bins over the gate [0.25, 0.75), whose centre the pulse peaks at.
    Returns (slices, pulsar)."""
    import dataclasses
    rng = np.random.default_rng(4)
    offsets = rng.uniform(-5e-3, 5e-3, n_stations)
    rates = rng.uniform(-3e-6, 3e-6, n_stations)
    sky = 1_400_000_000 + 16_000_000 * subband
    probe, _ = survey_jobs(seed=4, subband=subband, n_integrations=1, t_int_usec=int(round(seconds * 1e6)),
                           nchan=1024, n_stations=1, fs=32_000_000, bw=16_000_000, sky=sky)
    psr = dataclasses.replace(pulsar4(probe[0]), nbins=nbins)
    # phase at the start with ref_phase 0, moved to 0.2
    from .job import REFERENCE_MJD, SECONDS_PER_DAY
    tmid_ticks = (psr.tmid - REFERENCE_MJD) * SECONDS_PER_DAY * CLOCK_RATE
    p0 = psr.ref_freq * (probe[0].start_time - tmid_ticks) / CLOCK_RATE
    psr = dataclasses.replace(psr, ref_phase=(0.2 - p0) % 1.0)
    slices = build_slices(seed=4000 + subband, n_stations=n_stations, pols=["R", "L"], fs=32_000_000,
                          bw=16_000_000, nchan=1024, seconds=seconds, t_int_usec=int(round(seconds * 1e6)),
                          rho=0.005, bits=2, sky_freq=sky, offsets=list(offsets), rates=list(rates),
                          subband=subband, pulsar=(psr, 0.2, 0.05))
    for job, _ in slices:
        job.pulsar_binning = True
    return slices, psr
