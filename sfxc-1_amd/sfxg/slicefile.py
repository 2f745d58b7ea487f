"""Writes one slice (parameters, delay models, HEADER streams) in the binary
form host/correlate_slice.cpp reads."""
from __future__ import annotations

import struct
from typing import Sequence

import numpy as np

from . import capi
from .job import CorrelationParameters


def write_slice(path: str, job: CorrelationParameters, headers: Sequence[bytes]) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<5q", job.experiment_start, job.start_time, job.stop_time, job.integration_time,
                            job.sub_integration_time))
        f.write(struct.pack("<4i", job.number_channels, job.fft_size_delaycor, job.fft_size_correlation,
                            job.integration_nr))
        f.write(struct.pack("<2i", job.sample_rate, job.bandwidth))
        f.write(struct.pack("<q", job.channel_freq))
        f.write(job.sideband.encode())
        f.write(struct.pack("<3i", job.frequency_nr, job.window, job.reference_station))
        f.write(struct.pack("<i", len(job.station_streams)))
        for s, m, h in zip(job.station_streams, job.delay_models, headers):
            sp = capi.StationParameters(s.station_number, s.station_stream, s.sample_rate, s.bandwidth,
                                        s.channel_freq, s.sideband.encode(), s.polarisation.encode(), 0,
                                        s.bits_per_sample, s.LO_offset, s.tsys_freq, 0)
            f.write(bytes(sp))
            f.write(struct.pack("<qdi", m.scan_begin_ticks, m.interval, m.delay.shape[0]))
            f.write(np.ascontiguousarray(m.delay, dtype="<f8").tobytes())
            f.write(struct.pack("<Q", len(h)))
            f.write(h)
