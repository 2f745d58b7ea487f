"""Job description for one correlator-node time slice.

These classes restate the fields of the reference's ``Correlation_parameters``
(sfxc/include/control_parameters.h:117-182), ``Pulsar_parameters``
(control_parameters.h:70-105) and the delay model a ``Delay_table_akima``
evaluates (sfxc/src/delay_table_akima.cc:114-166), in the flat form the C ABI
(include/sfxg.h) takes. Times are int64 ticks of the reference ``Time`` class:
512 MHz clock ticks since MJD 51544 (sfxc/include/correlator_time.h:9-64).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

CLOCK_RATE = 512_000_000  # MAX_SAMPLE_RATE, sfxc/include/utils.h:36
REFERENCE_MJD = 51544
SECONDS_PER_DAY = 86400

WINDOW_NONE, WINDOW_RECT, WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN = 0, 1, 2, 3, 4
LUT_REFERENCE, LUT_CONFIG = 0, 1
PREC_FP32, PREC_FP16, PREC_BF16, PREC_SPLIT_FP16 = 0, 1, 2, 3


def mjd_to_ticks(mjd: int, sec: float = 0.0) -> int:
    """Time::set_time (correlator_time.h:76-78)."""
    return int((mjd - REFERENCE_MJD) * SECONDS_PER_DAY * CLOCK_RATE) + int(round(sec * CLOCK_RATE))


def usec_to_ticks(usec: float) -> int:
    """Time(double usec) (sfxc/src/correlator_time.cc:14-17)."""
    return int(round(usec * (CLOCK_RATE / 1_000_000)))


def ticks_to_mjd(ticks: int) -> float:
    """Time::get_mjd (correlator_time.cc:49-54)."""
    tpd = int(SECONDS_PER_DAY * CLOCK_RATE)
    return REFERENCE_MJD + ticks // tpd + (ticks % tpd) / CLOCK_RATE / SECONDS_PER_DAY


def nr_ffts_per_integration_slice(integration_ticks: int, data_rate: int, fft_size: int) -> int:
    """Control_parameters::nr_ffts_per_integration_slice (control_parameters.h:301-307).

    The reference passes ``(int)integration_time.get_time_usec()`` which is
    converted back to a Time, so the integration time is truncated to whole
    microseconds first.
    """
    usec = int(integration_ticks / (CLOCK_RATE / 1_000_000))
    t_int = usec_to_ticks(usec)
    t_fft = usec_to_ticks(fft_size / (data_rate / 1_000_000.0))
    return int(math.floor(t_int / t_fft))


@dataclasses.dataclass
class StationStream:
    """Correlation_parameters::Station_parameters (control_parameters.h:132-151)."""

    station_number: int
    station_stream: int
    sample_rate: int
    channel_freq: int
    bandwidth: int
    sideband: str = "U"
    polarisation: str = "R"
    bits_per_sample: int = 2
    LO_offset: float = 0.0
    tsys_freq: int = 80


@dataclasses.dataclass
class DelayModel:
    """Piecewise cubic on a uniform grid: the form an Akima spline over the
    1 s delay table takes (delay_table_akima.cc:430-502). ``delay[i]`` holds
    the power coefficients of piece i in seconds since its start."""

    scan_begin_ticks: int
    interval: float
    delay: np.ndarray               # (n, 4) float64, seconds
    phase: Optional[np.ndarray] = None  # (n, 4) radians
    amp: Optional[np.ndarray] = None    # (n, 4)

    @staticmethod
    def linear(scan_begin_ticks: int, offset: float, rate: float, seconds: int,
               interval: float = 1.0) -> "DelayModel":
        n = max(1, int(math.ceil(seconds / interval)))
        c = np.zeros((n, 4))
        for i in range(n):
            c[i, 0] = offset + rate * (i * interval)
            c[i, 1] = rate
        return DelayModel(scan_begin_ticks, interval, c)

    def eval(self, ticks, which: str = "delay"):
        """Evaluate at int64 ticks (scalar or array), Horner form in FP64."""
        coef = getattr(self, which)
        t = np.asarray(ticks, dtype=np.int64)
        if coef is None:
            return np.full(t.shape, 1.0 if which == "amp" else 0.0)
        sec = (t - np.int64(self.scan_begin_ticks)).astype(np.float64) / CLOCK_RATE
        i = np.clip(np.floor(sec / self.interval).astype(np.int64), 0, coef.shape[0] - 1)
        u = sec - i * self.interval
        c = coef[i]
        return c[..., 0] + u * (c[..., 1] + u * (c[..., 2] + u * c[..., 3]))


@dataclasses.dataclass
class Pulsar:
    """The gate and polyco of Pulsar_parameters::Pulsar / Polyco_params
    (control_parameters.h:73-98) for the polyco in use."""

    nbins: int
    gate_start: float
    gate_stop: float
    tmid: float
    DM: float
    ref_phase: float
    ref_freq: float
    obs_freq: float
    coef: List[float]


@dataclasses.dataclass
class PhaseCenter:
    """A phase centre besides the pointing centre: every stream's delay
    towards it (the further splines of Delay_table_akima,
    delay_table_akima.cc:114-136), its uvw per stream and the source number its
    records carry (correlation_core.cc:44-66)."""

    delay_models: List["DelayModel"]
    uvw: Optional[np.ndarray] = None  # (n_streams, 3) metres
    source_nr: int = 0


@dataclasses.dataclass
class MaskParameters:
    """Mask_parameters (control_parameters.h:107-114): used by integration_write
    when fft_size_correlation > number_channels (correlation_core.cc:417-440).
    mask: fft_size_correlation + 1 values multiplying the spectrum (None: ones);
    window: 2 * number_channels lag weights (None: derived from the job's
    window by create_window, correlation_core.cc:761-809)."""

    normalize: bool = False
    mask: Optional[List[float]] = None
    window: Optional[List[float]] = None


@dataclasses.dataclass
class CorrelationParameters:
    """Correlation_parameters (control_parameters.h:117-182), one integration
    slice of one subband."""

    experiment_start: int
    start_time: int
    stop_time: int
    integration_time: int
    sub_integration_time: int
    number_channels: int
    fft_size_delaycor: int
    fft_size_correlation: int
    sample_rate: int
    channel_freq: int
    bandwidth: int
    sideband: str
    frequency_nr: int
    station_streams: List[StationStream]
    integration_nr: int = 0
    slice_nr: int = 0
    slice_offset: int = 1
    polarisation: str = "R"
    cross_polarize: bool = True
    reference_station: int = -1
    window: int = WINDOW_HANN
    n_phase_centers: int = 1
    pulsar_binning: bool = False
    source: str = "SYNTH"
    # one delay model per entry of station_streams (Delay_table_akima per stream)
    delay_models: List[DelayModel] = dataclasses.field(default_factory=list)
    uvw: Optional[np.ndarray] = None  # (n_streams, 3) per stream, metres
    mask: Optional[MaskParameters] = None  # Correlation_parameters::mask_parameters
    # phase centres 1 .. n - 1 (multi-phase-centre jobs); source_nr is the
    # pointing centre's source number when there are any
    phase_centers: List[PhaseCenter] = dataclasses.field(default_factory=list)
    source_nr: int = 0

    def phase_center_count(self) -> int:
        return 1 + len(self.phase_centers) if self.phase_centers else self.n_phase_centers

    # --- derived quantities (each cites the reference expression) ---------
    def n_ffts_per_integration(self, i: int = 0) -> int:
        """Delay_correction / Bit2float_worker n_ffts_per_integration
        (delay_correction.cc:244-249, bit2float_worker.cc:261-273), N_d blocks."""
        st = self.station_streams[i]
        nd, nc = self.fft_size_delaycor, self.fft_size_correlation
        if nc > nd:
            return ((nc // nd) * nr_ffts_per_integration_slice(self.integration_time, self.sample_rate, nc)
                    * st.sample_rate) // self.sample_rate
        return (nr_ffts_per_integration_slice(self.integration_time, self.sample_rate, nd)
                * st.sample_rate) // self.sample_rate

    def n_ffts_per_buffer(self, i: int = 0) -> int:
        """bit2float_worker.cc:254-257 (CORRELATOR_BUFFER_SIZE = 8192)."""
        st = self.station_streams[i]
        nd, nc = self.fft_size_delaycor, self.fft_size_correlation
        nfft_min = max(nc // nd, 1)
        return (max(8192 // nd, nfft_min) * st.sample_rate) // self.sample_rate

    def number_ffts_in_integration(self) -> int:
        """correlation_core.cc:135-142."""
        n = nr_ffts_per_integration_slice(self.integration_time, self.sample_rate, self.fft_size_correlation)
        return n - 1 if self.window != WINDOW_NONE else n

    def number_ffts_in_sub_integration(self) -> int:
        return nr_ffts_per_integration_slice(self.sub_integration_time, self.sample_rate,
                                             self.fft_size_correlation)

    def baselines(self):
        """Correlation_core::create_baselines (correlation_core.cc:149-170)."""
        m = len(self.station_streams)
        bl = [(i, i) for i in range(m)]
        for i in range(m):
            for j in range(i + 1, m):
                s1 = self.station_streams[i].station_number
                s2 = self.station_streams[j].station_number
                if s1 == s2:
                    continue
                if self.reference_station >= 0 and self.reference_station not in (s1, s2):
                    continue
                bl.append((i, j) if s1 < s2 else (j, i))
        return bl
