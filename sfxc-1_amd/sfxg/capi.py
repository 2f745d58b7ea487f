"""ctypes binding of the C ABI in include/sfxg.h (libsfxg.so).

The library is the product: this module only marshals a job description
(sfxg.job) into the ABI's structs. Loading fails loudly when libsfxg.so has
not been built; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from .job import CorrelationParameters, DelayModel, Pulsar

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libsfxg.so")

SFXG_OK = 0
STAGE_SAMPLES, STAGE_TIME_SERIES, STAGE_SPECTRA, STAGE_ACCUMULATED, STAGE_NORMALISED, STAGE_CHANNELS = 0, 1, 2, 3, 4, 5


class StationParameters(C.Structure):
    _fields_ = [("station_number", C.c_int32), ("station_stream", C.c_int32), ("sample_rate", C.c_int32),
                ("bandwidth", C.c_int32), ("channel_freq", C.c_int64), ("sideband", C.c_char),
                ("polarisation", C.c_char), ("reserved0", C.c_int16), ("bits_per_sample", C.c_int32),
                ("LO_offset", C.c_double), ("tsys_freq", C.c_int32), ("reserved1", C.c_int32)]


class CorrelationParametersC(C.Structure):
    _fields_ = [("experiment_start", C.c_int64), ("start_time", C.c_int64), ("stop_time", C.c_int64),
                ("integration_time", C.c_int64), ("sub_integration_time", C.c_int64),
                ("number_channels", C.c_int32), ("fft_size_delaycor", C.c_int32),
                ("fft_size_correlation", C.c_int32), ("integration_nr", C.c_int32), ("slice_nr", C.c_int32),
                ("slice_offset", C.c_int32), ("sample_rate", C.c_int32), ("bandwidth", C.c_int32),
                ("channel_freq", C.c_int64), ("sideband", C.c_char), ("polarisation", C.c_char),
                ("reserved0", C.c_int16), ("frequency_nr", C.c_int32), ("cross_polarize", C.c_int32),
                ("reference_station", C.c_int32), ("window", C.c_int32), ("n_phase_centers", C.c_int32),
                ("pulsar_binning", C.c_int32), ("n_streams", C.c_int32),
                ("station_streams", C.POINTER(StationParameters))]


class DelayModelC(C.Structure):
    _fields_ = [("scan_begin", C.c_int64), ("interval", C.c_double), ("n_pieces", C.c_int32),
                ("reserved0", C.c_int32), ("delay", C.POINTER(C.c_double)), ("phase", C.POINTER(C.c_double)),
                ("amplitude", C.POINTER(C.c_double))]


class PulsarC(C.Structure):
    _fields_ = [("nbins", C.c_int32), ("n_coef", C.c_int32), ("gate_start", C.c_double),
                ("gate_stop", C.c_double), ("tmid", C.c_double), ("DM", C.c_double), ("ref_phase", C.c_double),
                ("ref_freq", C.c_double), ("obs_freq", C.c_double), ("coef", C.POINTER(C.c_double))]


class VdifStreamC(C.Structure):
    _fields_ = [("frame_bytes", C.c_int32), ("channel", C.c_int32), ("n_threads", C.c_int32), ("ref_mjd", C.c_int32),
                ("thread_ids", C.POINTER(C.c_int32)), ("offset", C.c_int64)]


class Options(C.Structure):
    _fields_ = [("lut", C.c_int32), ("precision", C.c_int32), ("keep_time_series", C.c_int32),
                ("chunk_mib", C.c_int32), ("phased_array", C.c_int32)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (make -C sfxc-1_amd)")
        L = C.CDLL(LIB_PATH)
        L.sfxg_last_error.restype = C.c_char_p
        L.sfxg_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.sfxg_destroy.argtypes = [C.c_void_p]
        for name in ("sfxg_set_options", "sfxg_set_parameters", "sfxg_push_stream", "sfxg_push_stream_elided",
                     "sfxg_correlate", "sfxg_synchronize", "sfxg_output_size", "sfxg_fetch_output",
                     "sfxg_get_statistics", "sfxg_number_of_baselines", "sfxg_fetch_stage"):
            getattr(L, name).restype = C.c_int
        L.sfxg_last_error.argtypes = [C.c_void_p]
        L.sfxg_set_options.argtypes = [C.c_void_p, C.POINTER(Options)]
        L.sfxg_set_parameters.argtypes = [C.c_void_p, C.POINTER(CorrelationParametersC), C.POINTER(DelayModelC),
                                          C.POINTER(C.c_double), C.POINTER(PulsarC)]
        L.sfxg_push_stream.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_size_t]
        L.sfxg_push_stream_elided.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_size_t, C.c_void_p,
                                              C.c_size_t]
        L.sfxg_correlate.argtypes = [C.c_void_p]
        L.sfxg_synchronize.argtypes = [C.c_void_p]
        L.sfxg_output_size.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
        L.sfxg_fetch_output.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.sfxg_get_statistics.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.sfxg_number_of_baselines.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.sfxg_fetch_stage.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t,
                                       C.POINTER(C.c_size_t)]
        L.sfxg_kernel_times.restype = C.c_int
        L.sfxg_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.c_int32]
        L.sfxg_chunk_plan.restype = C.c_int
        L.sfxg_chunk_plan.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.sfxg_prepare.restype = C.c_int
        L.sfxg_prepare.argtypes = [C.c_void_p]
        L.sfxg_batch_create.restype = C.c_int
        L.sfxg_batch_create.argtypes = [C.c_int, C.c_int32, C.POINTER(C.c_void_p)]
        L.sfxg_batch_destroy.argtypes = [C.c_void_p]
        L.sfxg_batch_last_error.restype = C.c_char_p
        L.sfxg_batch_last_error.argtypes = [C.c_void_p]
        for name in ("sfxg_batch_add", "sfxg_batch_clear", "sfxg_batch_launch", "sfxg_batch_synchronize"):
            getattr(L, name).restype = C.c_int
        L.sfxg_batch_add.argtypes = [C.c_void_p, C.c_void_p]
        L.sfxg_batch_clear.argtypes = [C.c_void_p]
        L.sfxg_batch_launch.argtypes = [C.c_void_p]
        L.sfxg_batch_synchronize.argtypes = [C.c_void_p]
        L.sfxg_autotune.restype = C.c_int
        L.sfxg_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_size_t]
        L.sfxg_tune_export.restype = C.c_int
        L.sfxg_tune_export.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.sfxg_tune_import.restype = C.c_int
        L.sfxg_tune_import.argtypes = [C.c_char_p]
        L.sfxg_tune_clear.restype = None
        L.sfxg_tune_clear.argtypes = []
        L.sfxg_fetch_tsys.restype = C.c_int
        L.sfxg_fetch_tsys.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.sfxg_vdif_phasecal.restype = C.c_int
        L.sfxg_vdif_phasecal.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t, C.POINTER(VdifStreamC),
                                         C.c_int64, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.sfxg_vdif_phasecal_static.restype = C.c_int
        L.sfxg_vdif_phasecal_static.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t,
                                                C.POINTER(VdifStreamC), C.c_int64, C.c_void_p, C.c_size_t,
                                                C.POINTER(C.c_size_t)]
        L.sfxg_push_vdif.restype = C.c_int
        L.sfxg_push_vdif.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t, C.POINTER(VdifStreamC)]
        L.sfxg_set_phase_centers.restype = C.c_int
        L.sfxg_set_phase_centers.argtypes = [C.c_void_p, C.c_int32, C.POINTER(DelayModelC), C.POINTER(C.c_double),
                                             C.POINTER(C.c_int32)]
        L.sfxg_set_mask.restype = C.c_int
        L.sfxg_set_mask.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_int32,
                                    C.POINTER(C.c_double), C.c_int32]
        L.sfxg_synthesize_raw.restype = C.c_int
        L.sfxg_synthesize_raw.argtypes = [C.c_int, C.c_void_p, C.c_size_t, C.c_int32, C.c_int64, C.c_int64,
                                          C.c_uint64, C.c_uint64, C.c_float, C.c_float, C.c_int64, C.c_uint64,
                                          C.c_float]
        _lib = L
    return _lib


class SfxgError(RuntimeError):
    pass


class Context:
    """One correlator-node context on a HIP device."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().sfxg_create(device, C.byref(self._h))
        if rc != SFXG_OK:
            raise SfxgError(f"sfxg_create failed ({rc})")
        self._keep: List = []
        self.job: Optional[CorrelationParameters] = None
        self.nbins = 1

    def close(self):
        if self._h:
            lib().sfxg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != SFXG_OK:
            msg = lib().sfxg_last_error(self._h)
            raise SfxgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_options(self, lut: int = 0, precision: int = 0, keep_time_series: bool = False, chunk_mib: int = 0,
                    phased_array: bool = False):
        o = Options(lut, precision, int(keep_time_series), chunk_mib, int(phased_array))
        self._check(lib().sfxg_set_options(self._h, C.byref(o)), "sfxg_set_options")

    def set_parameters(self, job: CorrelationParameters, pulsar: Optional[Pulsar] = None):
        keep = []
        M = len(job.station_streams)
        sts = (StationParameters * M)()
        for i, s in enumerate(job.station_streams):
            sts[i] = StationParameters(s.station_number, s.station_stream, s.sample_rate, s.bandwidth,
                                       s.channel_freq, s.sideband.encode(), s.polarisation.encode(), 0,
                                       s.bits_per_sample, s.LO_offset, s.tsys_freq, 0)
        keep.append(sts)
        p = CorrelationParametersC(
            job.experiment_start, job.start_time, job.stop_time, job.integration_time, job.sub_integration_time,
            job.number_channels, job.fft_size_delaycor, job.fft_size_correlation, job.integration_nr,
            job.slice_nr, job.slice_offset, job.sample_rate, job.bandwidth, job.channel_freq,
            job.sideband.encode(), job.polarisation.encode(), 0, job.frequency_nr, int(job.cross_polarize),
            job.reference_station, job.window, job.phase_center_count(), int(job.pulsar_binning or pulsar is not None),
            M, sts)
        models = (DelayModelC * M)()

        def arr(a):
            if a is None:
                return C.POINTER(C.c_double)()
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        for i, m in enumerate(job.delay_models):
            models[i] = DelayModelC(m.scan_begin_ticks, m.interval, m.delay.shape[0], 0, arr(m.delay),
                                    arr(m.phase), arr(m.amp))
        keep.append(models)
        uvw = C.POINTER(C.c_double)()
        if job.uvw is not None:
            uvw = arr(np.asarray(job.uvw, dtype=np.float64).reshape(M, 3))
        ps = None
        if pulsar is not None:
            coef = arr(np.asarray(pulsar.coef, dtype=np.float64))
            ps = PulsarC(pulsar.nbins, len(pulsar.coef), pulsar.gate_start, pulsar.gate_stop, pulsar.tmid,
                         pulsar.DM, pulsar.ref_phase, pulsar.ref_freq, pulsar.obs_freq, coef)
            keep.append(ps)
        self._check(lib().sfxg_set_parameters(self._h, C.byref(p), models, uvw,
                                               C.byref(ps) if ps is not None else None), "sfxg_set_parameters")
        if job.phase_centers:
            self.set_phase_centers(job)
        self.set_mask(job.mask)
        self._keep = keep
        self.job = job
        self.nbins = pulsar.nbins + 1 if pulsar is not None else job.phase_center_count()

    def set_phase_centers(self, job: CorrelationParameters):
        """The job's phase centres after the pointing centre (sfxg_set_phase_centers)."""
        keep = []
        M = len(job.station_streams)
        n = 1 + len(job.phase_centers)
        models = (DelayModelC * ((n - 1) * M))()
        for j, pc in enumerate(job.phase_centers):
            assert len(pc.delay_models) == M
            for i, m in enumerate(pc.delay_models):
                c = np.ascontiguousarray(m.delay, dtype=np.float64).ravel()
                keep.append(c)
                models[j * M + i] = DelayModelC(m.scan_begin_ticks, m.interval, m.delay.shape[0], 0,
                                                c.ctypes.data_as(C.POINTER(C.c_double)),
                                                C.POINTER(C.c_double)(), C.POINTER(C.c_double)())
        uvw = np.zeros((n, M, 3))
        if job.uvw is not None:
            uvw[0] = np.asarray(job.uvw, dtype=np.float64).reshape(M, 3)
        for j, pc in enumerate(job.phase_centers):
            if pc.uvw is not None:
                uvw[j + 1] = np.asarray(pc.uvw, dtype=np.float64).reshape(M, 3)
        src = np.array([job.source_nr] + [pc.source_nr for pc in job.phase_centers], dtype=np.int32)
        self._check(lib().sfxg_set_phase_centers(self._h, n, models, uvw.ctypes.data_as(C.POINTER(C.c_double)),
                                                  src.ctypes.data_as(C.POINTER(C.c_int32))),
                    "sfxg_set_phase_centers")

    def set_mask(self, mask=None):
        """Mask_parameters of the job (sfxg.job.MaskParameters or None for the defaults)."""
        keep = []

        def arr(a):
            if a is None or len(a) == 0:
                return C.POINTER(C.c_double)(), 0
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double)), len(a)

        m, nm = arr(mask.mask if mask is not None else None)
        w, nw = arr(mask.window if mask is not None else None)
        norm = int(bool(mask.normalize)) if mask is not None else 0
        self._check(lib().sfxg_set_mask(self._h, norm, m, nm, w, nw), "sfxg_set_mask")

    def push_stream(self, i: int, hdr: bytes):
        self._check(lib().sfxg_push_stream(self._h, i, hdr, len(hdr)), f"sfxg_push_stream({i})")

    def push_vdif(self, i: int, device_ptr: int, nbytes: int, frame_bytes: int, channel: int, thread_ids,
                  ref_mjd: int, offset: int = 0):
        """Stream i from raw VDIF frames in device memory (sfxg_push_vdif)."""
        th = np.ascontiguousarray(thread_ids, dtype=np.int32)
        v = VdifStreamC(frame_bytes, channel, len(th), ref_mjd, th.ctypes.data_as(C.POINTER(C.c_int32)), offset)
        self._check(lib().sfxg_push_vdif(self._h, i, C.c_void_p(device_ptr), nbytes, C.byref(v)),
                    f"sfxg_push_vdif({i})")

    def vdif_phasecal(self, i: int, device_ptr: int, nbytes: int, frame_bytes: int, channel: int, thread_ids,
                      ref_mjd: int, integration_ticks: int, offset: int = 0, static: bool = False) -> bytes:
        """Phase-cal messages of stream i's VDIF frames (sfxg_vdif_phasecal, or
        sfxg_vdif_phasecal_static with static=True)."""
        name = "sfxg_vdif_phasecal_static" if static else "sfxg_vdif_phasecal"
        fn = getattr(lib(), name)
        th = np.ascontiguousarray(thread_ids, dtype=np.int32)
        v = VdifStreamC(frame_bytes, channel, len(th), ref_mjd, th.ctypes.data_as(C.POINTER(C.c_int32)), offset)
        need = C.c_size_t()
        rc = fn(self._h, i, C.c_void_p(device_ptr), nbytes, C.byref(v), integration_ticks, None, 0, C.byref(need))
        if rc != SFXG_OK and need.value == 0:
            self._check(rc, name)
        buf = (C.c_uint8 * max(1, need.value))()
        w = C.c_size_t()
        self._check(fn(self._h, i, C.c_void_p(device_ptr), nbytes, C.byref(v), integration_ticks, buf, need.value,
                       C.byref(w)), name)
        return bytes(buf)[:w.value]

    def push_stream_elided(self, i: int, control: bytes, device_ptr: int, nbytes: int):
        self._check(lib().sfxg_push_stream_elided(self._h, i, control, len(control), C.c_void_p(device_ptr),
                                                  nbytes), f"sfxg_push_stream_elided({i})")

    def correlate(self):
        self._check(lib().sfxg_correlate(self._h), "sfxg_correlate")

    def prepare(self):
        """Host half of correlate; the device work runs in a Batch."""
        self._check(lib().sfxg_prepare(self._h), "sfxg_prepare")

    def autotune(self, reps: int = 3) -> str:
        """Times the launch-shape candidates on the prepared slice and keeps the
        fastest for its shape; returns the candidates' times. The slice must be
        prepared (or correlated) again afterwards."""
        buf = C.create_string_buffer(4096)
        self._check(lib().sfxg_autotune(self._h, reps, buf, len(buf)), "sfxg_autotune")
        return buf.value.decode()

    def synchronize(self):
        self._check(lib().sfxg_synchronize(self._h), "sfxg_synchronize")

    def output(self) -> bytes:
        n = C.c_size_t()
        self._check(lib().sfxg_output_size(self._h, C.byref(n)), "sfxg_output_size")
        buf = np.zeros(n.value, dtype=np.uint8)
        w = C.c_size_t()
        self._check(lib().sfxg_fetch_output(self._h, buf.ctypes.data, n.value, C.byref(w)), "sfxg_fetch_output")
        return buf[:w.value].tobytes()

    def statistics(self, i: int):
        lv = (C.c_int32 * 5)()
        ts = (C.c_int32 * 4)()
        self._check(lib().sfxg_get_statistics(self._h, i, lv, ts), "sfxg_get_statistics")
        return np.array(lv[:], dtype=np.int32), np.array(ts[:], dtype=np.int32)

    def kernel_times(self, reset: bool = True) -> np.ndarray:
        """Average per-slice milliseconds of stats, delay (fused F-engine),
        channelise, xstep and normalise since the last reset."""
        ms = (C.c_double * 5)()
        self._check(lib().sfxg_kernel_times(self._h, ms, 5, int(reset)), "sfxg_kernel_times")
        return np.array(ms[:])

    def chunk_plan(self) -> tuple:
        """(chunks, segments per stream of the largest chunk) of the prepared
        slice: the delay, channeliser and X-step kernels run once per chunk."""
        n, segs = C.c_int32(), C.c_int64()
        self._check(lib().sfxg_chunk_plan(self._h, C.byref(n), C.byref(segs)), "sfxg_chunk_plan")
        return n.value, segs.value

    def tsys(self) -> bytes:
        """The slice's tsys_write messages (sfxg_fetch_tsys), 44 bytes per stream."""
        n = 44 * len(self.job.station_streams)
        buf = (C.c_uint8 * n)()
        w = C.c_size_t()
        self._check(lib().sfxg_fetch_tsys(self._h, buf, n, C.byref(w)), "sfxg_fetch_tsys")
        return bytes(buf)[:w.value]

    def number_of_baselines(self) -> int:
        n = C.c_int32()
        self._check(lib().sfxg_number_of_baselines(self._h, C.byref(n)), "sfxg_number_of_baselines")
        return n.value

    def stage(self, stage: int, stream: int = 0) -> np.ndarray:
        job = self.job
        nd = job.fft_size_delaycor
        nc = job.fft_size_correlation
        B = self.number_of_baselines()
        if stage in (STAGE_SAMPLES, STAGE_TIME_SERIES):
            out = np.zeros(job.n_ffts_per_integration(stream) * nd, dtype=np.float32)
        elif stage == STAGE_SPECTRA:
            out = np.zeros((job.number_ffts_in_integration(), nc + 1), dtype=np.complex64)
        elif stage == STAGE_CHANNELS:
            out = np.zeros((self.nbins, B, job.number_channels + 1), dtype=np.complex64)
        else:
            out = np.zeros((self.nbins, B, nc + 1), dtype=np.complex64)
        w = C.c_size_t()
        self._check(lib().sfxg_fetch_stage(self._h, stage, stream, out.ctypes.data, out.nbytes, C.byref(w)),
                    "sfxg_fetch_stage")
        return out


def tune_key(nchan: int, n_streams: int, precision: int, pulsar: bool, fused: bool) -> str:
    """The tuning table's shape key of a slice."""
    return f"nc={nchan} M={n_streams} prec={precision} pulsar={int(pulsar)} fused={int(fused)}"


def tune_export() -> str:
    n = C.c_size_t()
    lib().sfxg_tune_export(None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value)
    rc = lib().sfxg_tune_export(buf, n.value, C.byref(n))
    if rc != SFXG_OK:
        raise SfxgError(f"sfxg_tune_export failed ({rc})")
    return buf.value.decode()


def tune_import(text: str):
    rc = lib().sfxg_tune_import(text.encode())
    if rc != SFXG_OK:
        raise SfxgError(f"sfxg_tune_import failed ({rc})")


def tune_clear():
    lib().sfxg_tune_clear()


class Batch:
    """Prepared contexts whose device work runs as one HIP graph launch."""

    def __init__(self, device: int = 0, n_branches: int = 4):
        self._h = C.c_void_p()
        rc = lib().sfxg_batch_create(device, n_branches, C.byref(self._h))
        if rc != SFXG_OK:
            raise SfxgError(f"sfxg_batch_create failed ({rc})")
        self.members: List[Context] = []

    def _check(self, rc: int, what: str):
        if rc != SFXG_OK:
            msg = lib().sfxg_batch_last_error(self._h)
            raise SfxgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def add(self, ctx: Context):
        self._check(lib().sfxg_batch_add(self._h, ctx._h), "sfxg_batch_add")
        self.members.append(ctx)

    def clear(self):
        self._check(lib().sfxg_batch_clear(self._h), "sfxg_batch_clear")
        self.members = []

    def launch(self):
        self._check(lib().sfxg_batch_launch(self._h), "sfxg_batch_launch")

    def synchronize(self):
        self._check(lib().sfxg_batch_synchronize(self._h), "sfxg_batch_synchronize")

    def close(self):
        if self._h:
            lib().sfxg_batch_destroy(self._h)
            self._h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synthesize_raw(device: int, ptr: int, n_bytes: int, bits: int, first_sample: int, delay_samples: int,
                   common_seed: int, stream_seed: int, rho: float, threshold: float = 0.9816,
                   delay2_samples: int = 0, common2_seed: int = 0, rho2: float = 0.0):
    rc = lib().sfxg_synthesize_raw(device, C.c_void_p(ptr), n_bytes, bits, first_sample, delay_samples,
                                   common_seed, stream_seed, rho, threshold, delay2_samples, common2_seed, rho2)
    if rc != SFXG_OK:
        raise SfxgError(f"sfxg_synthesize_raw failed ({rc})")


def correlate_slice(ctx: Context, job: CorrelationParameters, headers: Sequence[bytes], lut: int = 0,
                    pulsar: Optional[Pulsar] = None, keep_time_series: bool = False, precision: int = 0,
                    chunk_mib: int = 0, phased_array: bool = False) -> bytes:
    ctx.set_options(lut=lut, precision=precision, keep_time_series=keep_time_series, chunk_mib=chunk_mib,
                    phased_array=phased_array)
    ctx.set_parameters(job, pulsar)
    for i, h in enumerate(headers):
        ctx.push_stream(i, h)
    ctx.correlate()
    return ctx.output()
