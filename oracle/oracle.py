"""Python driver of the CPU restatement (liboracle_f32.so / liboracle_f64.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. It runs the correlator-node data
path for one slice the way sfxc/src/correlator_node.cc:263-296 drives it:
Bit2float_worker -> Delay_correction per stream, then Correlation_core (or
Correlation_core_pulsar), normalisation and integration_write.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
CLOCK_RATE = 512_000_000


class B2fParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "bits_per_sample", "sample_rate", "base_sample_rate", "tsys_freq", "fft_size",
        "nfft_max", "n_ffts_per_integration", "delay_in_samples", "lut", "pad")]


class DelayModelC(C.Structure):
    _fields_ = [("scan_begin_ticks", C.c_int64), ("interval", C.c_double), ("n", C.c_int32),
                ("pad", C.c_int32), ("delay", C.POINTER(C.c_double)),
                ("phase", C.POINTER(C.c_double)), ("amp", C.POINTER(C.c_double))]


class DcParams(C.Structure):
    _fields_ = [("fft_size_delaycor", C.c_int32), ("fft_size_correlation", C.c_int32),
                ("sample_rate", C.c_int32), ("station_sample_rate", C.c_int32),
                ("bandwidth", C.c_int32), ("sideband", C.c_int32), ("job_sideband", C.c_int32),
                ("window", C.c_int32), ("channel_freq", C.c_int64), ("job_channel_freq", C.c_int64),
                ("LO_offset", C.c_double), ("start_ticks", C.c_int64),
                ("experiment_start_ticks", C.c_int64), ("n_ffts_per_integration", C.c_int32),
                ("stride", C.c_int32)]


class CoreParams(C.Structure):
    _fields_ = [("n_streams", C.c_int32), ("fft_size", C.c_int32), ("stride", C.c_int32),
                ("n_baselines", C.c_int32), ("number_ffts_in_integration", C.c_int32),
                ("number_ffts_in_sub_integration", C.c_int32),
                ("bl_first", C.POINTER(C.c_int32)), ("bl_second", C.POINTER(C.c_int32))]


class PcParamsC(C.Structure):
    _fields_ = [("n_phase_centers", C.c_int32), ("sideband", C.c_int32), ("sample_rate", C.c_int32),
                ("pad", C.c_int32), ("start_ticks", C.c_int64), ("channel_freq", C.c_int64),
                ("models", C.POINTER(DelayModelC)), ("weights", C.POINTER(C.c_float))]


class RechanParamsC(C.Structure):
    _fields_ = [("fft_size", C.c_int32), ("nchan", C.c_int32), ("normalize", C.c_int32), ("pad", C.c_int32),
                ("mask", C.POINTER(C.c_float)), ("window", C.POINTER(C.c_float))]


class PulsarParamsC(C.Structure):
    _fields_ = [("nbins", C.c_int32), ("pad", C.c_int32), ("start_phase", C.c_double),
                ("phase_coef", C.c_double), ("fft_duration", C.c_double),
                ("gate_begin", C.c_double), ("gate_end", C.c_double),
                ("offsets", C.POINTER(C.c_double))]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_threads: Optional[int] = None


def set_threads(n: Optional[int]) -> None:
    """Fixes the driver's thread count (None: the default of _workers)."""
    global _threads
    _threads = n


def _workers() -> int:
    """Threads the driver runs independent C calls on (streams, groups of
    baselines): the host cores this process may use, at most 16. The C
    calls release the GIL, and each computes what it would alone."""
    if _threads:
        return _threads
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _map(fn, items):
    items = list(items)
    if len(items) <= 1 or _workers() == 1:
        return [fn(x) for x in items]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=_workers()) as ex:
        return list(ex.map(fn, items))


def _baseline_groups(n_baselines: int):
    """Contiguous groups of baselines for the threads (each baseline's sums
    do not depend on the others')."""
    k = max(1, min(_workers() * 2, n_baselines))
    edges = np.linspace(0, n_baselines, k + 1).astype(int)
    return [(a, b) for a, b in zip(edges[:-1], edges[1:]) if b > a]


_libs: Dict[str, C.CDLL] = {}


def lib(precision: str = "f32") -> C.CDLL:
    if precision not in _libs:
        path = os.path.join(BUILD, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        _libs[precision] = C.CDLL(path)
    return _libs[precision]


def _dtype(precision):
    return np.float32 if precision == "f32" else np.float64


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _rptr(a, precision):
    return _ptr(a, C.c_float if precision == "f32" else C.c_double)


def sb_sign(c: str) -> int:
    return -1 if c == "L" else 1


# ----------------------------------------------------------------------------
def bit2float(hdr: bytes, p: dict, precision="f32"):
    """Returns dict(samples, chunks, invalid_start, invalid_n, n_invalid, counts_on, counts_off)."""
    L = lib(precision)
    fn = getattr(L, f"or_bit2float_{precision}")
    prm = B2fParams(**p)
    n_out = p["n_ffts_per_integration"] * p["fft_size"]
    out = np.zeros(n_out, dtype=_dtype(precision))
    chunks = np.zeros(n_out // max(1, p["fft_size"]) + 8, dtype=np.int32)
    max_inv = 1 << 16
    inv_s = np.zeros(max_inv, dtype=np.int32)
    inv_n = np.zeros(max_inv, dtype=np.int32)
    on = np.zeros(256, dtype=np.int32)
    off = np.zeros(256, dtype=np.int32)
    res = np.zeros(4, dtype=np.int32)
    buf = np.frombuffer(hdr, dtype=np.uint8)
    rc = fn(C.byref(prm), _ptr(buf, C.c_uint8), C.c_int64(len(buf)), _rptr(out, precision),
            _ptr(chunks, C.c_int32), C.c_int32(len(chunks)), _ptr(inv_s, C.c_int32),
            _ptr(inv_n, C.c_int32), C.c_int32(max_inv), _ptr(on, C.c_int32), _ptr(off, C.c_int32),
            _ptr(res, C.c_int32))
    if rc != 0:
        raise RuntimeError(f"or_bit2float failed: {rc}")
    return dict(samples=out, chunks=chunks[:res[0]].copy(), invalid_start=inv_s[:res[1]].copy(),
                invalid_n=inv_n[:res[1]].copy(), n_invalid=int(res[2]), bytes_read=int(res[3]),
                counts_on=on, counts_off=off)


def bit_statistics(bits, rate, base_rate, on, off, n_invalid, precision="f32"):
    L = lib(precision)
    levels = np.zeros(5, dtype=np.int32)
    tsys = np.zeros(4, dtype=np.int32)
    getattr(L, f"or_bit_statistics_{precision}")(
        C.c_int(bits), C.c_int64(rate), C.c_int64(base_rate), _ptr(on, C.c_int32),
        _ptr(off, C.c_int32), C.c_int32(n_invalid), _ptr(levels, C.c_int32), _ptr(tsys, C.c_int32))
    return levels, tsys


def fft_c2c(x: np.ndarray, sign: int, precision="f32") -> np.ndarray:
    d = _dtype(precision)
    a = np.empty(2 * len(x), dtype=d)
    a[0::2] = x.real
    a[1::2] = x.imag
    getattr(lib(precision), f"or_fft_c2c_{precision}")(_rptr(a, precision), C.c_int(len(x)), C.c_int(sign))
    return a[0::2] + 1j * a[1::2]


def rfft(x: np.ndarray, precision="f32") -> np.ndarray:
    d = _dtype(precision)
    a = np.ascontiguousarray(x, dtype=d)
    out = np.empty(2 * (len(x) // 2 + 1), dtype=d)
    getattr(lib(precision), f"or_rfft_{precision}")(_rptr(a, precision), _rptr(out, precision),
                                                    C.c_int(len(x)))
    return out[0::2] + 1j * out[1::2]


def window(kind: int, n: int, precision="f32") -> np.ndarray:
    w = np.empty(n, dtype=_dtype(precision))
    getattr(lib(precision), f"or_window_{precision}")(C.c_int(kind), C.c_int(n), _rptr(w, precision))
    return w


def _model_c(model, keep):
    def arr(a):
        if a is None:
            return C.POINTER(C.c_double)()
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        keep.append(a)
        return _ptr(a, C.c_double)
    return DelayModelC(model.scan_begin_ticks, model.interval, model.delay.shape[0], 0,
                       arr(model.delay), arr(model.phase), arr(model.amp))


def eval_delay(model, ticks: int, precision="f64") -> float:
    keep = []
    mc = _model_c(model, keep)
    f = getattr(lib(precision), f"or_eval_delay_{precision}")
    f.restype = C.c_double
    return f(C.byref(mc), C.c_int64(ticks))


def b2f_params(job, i, lut) -> dict:
    st = job.station_streams[i]
    delay0 = job.delay_models[i].eval(job.start_time)
    return dict(bits_per_sample=st.bits_per_sample, sample_rate=st.sample_rate,
                base_sample_rate=job.sample_rate, tsys_freq=st.tsys_freq,
                fft_size=job.fft_size_delaycor, nfft_max=job.n_ffts_per_buffer(i),
                n_ffts_per_integration=job.n_ffts_per_integration(i),
                delay_in_samples=int(float(delay0) * st.sample_rate), lut=lut, pad=0)


def delay_correction(job, i, samples, chunks, precision="f32", want_time_series=False):
    st = job.station_streams[i]
    stride = job.fft_size_correlation + 4
    prm = DcParams(job.fft_size_delaycor, job.fft_size_correlation, job.sample_rate, st.sample_rate,
                   st.bandwidth, sb_sign(st.sideband), sb_sign(job.sideband), job.window,
                   st.channel_freq, job.channel_freq, st.LO_offset, job.start_time,
                   job.experiment_start, job.n_ffts_per_integration(i), stride)
    keep = []
    mc = _model_c(job.delay_models[i], keep)
    nseg_max = int(len(samples) // job.fft_size_correlation) + 2
    spec = np.zeros(2 * nseg_max * stride, dtype=_dtype(precision))
    ts = np.zeros(len(samples), dtype=_dtype(precision)) if want_time_series else None
    spc = np.zeros(len(chunks), dtype=np.int32)
    s = np.ascontiguousarray(samples, dtype=_dtype(precision))
    ch = np.ascontiguousarray(chunks, dtype=np.int32)
    f = getattr(lib(precision), f"or_delay_correction_{precision}")
    f.restype = C.c_int64
    nseg = f(C.byref(prm), C.byref(mc), _rptr(s, precision), _ptr(ch, C.c_int32), C.c_int32(len(ch)),
             _rptr(spec, precision), _rptr(ts, precision) if ts is not None else None,
             _ptr(spc, C.c_int32))
    spec = spec[:2 * nseg * stride].reshape(nseg, stride, 2)
    return dict(spectra=spec, seg_per_chunk=spc, time_series=ts, nseg=int(nseg))


def _core_params(job, keep):
    bl = job.baselines()
    f = np.array([b[0] for b in bl], dtype=np.int32)
    s = np.array([b[1] for b in bl], dtype=np.int32)
    keep += [f, s]
    return CoreParams(len(job.station_streams), job.fft_size_correlation, job.fft_size_correlation + 4,
                      len(bl), job.number_ffts_in_integration(), job.number_ffts_in_sub_integration(),
                      _ptr(f, C.c_int32), _ptr(s, C.c_int32)), bl


def _core_params_range(job, a, b, keep):
    """CoreParams of baselines a .. b - 1 of the job's list."""
    cp, bl = _core_params(job, keep)
    f = np.array([x[0] for x in bl[a:b]], dtype=np.int32)
    s = np.array([x[1] for x in bl[a:b]], dtype=np.int32)
    keep += [f, s]
    cp.n_baselines = b - a
    cp.bl_first = _ptr(f, C.c_int32)
    cp.bl_second = _ptr(s, C.c_int32)
    return cp, bl


def correlate(job, spectra: List[np.ndarray], seg_per_chunk, precision="f32"):
    """or_correlate over groups of baselines on the driver's threads."""
    bl = job.baselines()
    nf = job.fft_size_correlation + 1
    arrs = [np.ascontiguousarray(s, dtype=_dtype(precision)) for s in spectra]
    ptrs = (C.POINTER(C.c_float if precision == "f32" else C.c_double) * len(arrs))(
        *[_rptr(a, precision) for a in arrs])
    vis = np.zeros((len(bl), nf, 2), dtype=_dtype(precision))
    spc = np.ascontiguousarray(seg_per_chunk, dtype=np.int32)
    fn = getattr(lib(precision), f"or_correlate_{precision}")

    def run(rng):
        a, b = rng
        keep = []
        cp, _ = _core_params_range(job, a, b, keep)
        part = np.zeros((b - a, nf, 2), dtype=_dtype(precision))
        fn(C.byref(cp), ptrs, _ptr(spc, C.c_int32), C.c_int32(len(spc)), _rptr(part, precision))
        vis[a:b] = part

    _map(run, _baseline_groups(len(bl)))
    return vis


def phased_correlate(job, spectra: List[np.ndarray], nseg: int, precision="f32"):
    """Correlation_core_phased: per sub-integration sums of all streams' spectra
    -> (number_ffts_in_integration // number_ffts_in_sub_integration + 1, N_c + 1, 2)."""
    keep = []
    cp, _ = _core_params(job, keep)
    nf = job.fft_size_correlation + 1
    nrows = job.number_ffts_in_integration() // job.number_ffts_in_sub_integration() + 1
    arrs = [np.ascontiguousarray(s, dtype=_dtype(precision)) for s in spectra]
    ptrs = (C.POINTER(C.c_float if precision == "f32" else C.c_double) * len(arrs))(
        *[_rptr(a, precision) for a in arrs])
    rows = np.zeros(2 * nrows * nf, dtype=_dtype(precision))
    f = getattr(lib(precision), f"or_phased_correlate_{precision}")
    n = f(C.byref(cp), ptrs, C.c_int64(nseg), _rptr(rows, precision))
    assert n == nrows
    return rows.reshape(nrows, nf, 2)


def eval_rate(model, ticks: int, precision="f64") -> float:
    keep = []
    m = _model_c(model, keep)
    f = getattr(lib(precision), f"or_eval_rate_{precision}")
    f.restype = C.c_double
    return f(C.byref(m), C.c_int64(ticks))


def pc_weights(window: int, n: int, precision="f32") -> np.ndarray:
    """create_weights (correlation_core.cc:712-759)."""
    w = np.zeros(n, dtype=np.float32)
    if getattr(lib(precision), f"or_pc_weights_{precision}")(C.c_int(window), C.c_int(n), _ptr(w, C.c_float)) != 0:
        raise ValueError(f"invalid window {window}")
    return w


def correlate_pc(job, spectra: List[np.ndarray], seg_per_chunk, precision="f32"):
    """Correlation_core with the job's extra phase centres:
    (n_phase_centers, n_baselines, N_c + 1, 2) accumulators."""
    keep = []
    cp, bl = _core_params(job, keep)
    nf = job.fft_size_correlation + 1
    M = len(job.station_streams)
    centres = [job.delay_models] + [pc.delay_models for pc in job.phase_centers]
    n_pc = len(centres)
    models = (DelayModelC * (n_pc * M))()
    for j, ms in enumerate(centres):
        assert len(ms) == M
        for i, m in enumerate(ms):
            models[j * M + i] = _model_c(m, keep)
    w = pc_weights(job.window, job.fft_size_correlation, precision)
    pp = PcParamsC(n_pc, sb_sign(job.sideband), job.sample_rate, 0, job.start_time, job.channel_freq, models,
                   _ptr(w, C.c_float))
    arrs = [np.ascontiguousarray(s, dtype=_dtype(precision)) for s in spectra]
    ptrs = (C.POINTER(C.c_float if precision == "f32" else C.c_double) * len(arrs))(
        *[_rptr(a, precision) for a in arrs])
    vis = np.zeros(2 * n_pc * len(bl) * nf, dtype=_dtype(precision))
    spc = np.ascontiguousarray(seg_per_chunk, dtype=np.int32)
    getattr(lib(precision), f"or_correlate_pc_{precision}")(C.byref(cp), C.byref(pp), ptrs, _ptr(spc, C.c_int32),
                                                            C.c_int32(len(spc)), _rptr(vis, precision))
    return vis.reshape(n_pc, len(bl), nf, 2)


def find_invalid(job, invalid: List[tuple]):
    """n_flagged per baseline (correlation_core.cc:599-665)."""
    bl = job.baselines()
    M = len(job.station_streams)
    out = np.zeros((len(bl), 2), dtype=np.int64)
    f = getattr(lib("f32"), "or_find_invalid_f32")
    for b in range(M, len(bl)):
        i, j = bl[b]
        s0, n0 = invalid[i]
        s1, n1 = invalid[j]
        s0 = np.ascontiguousarray(s0, dtype=np.int32)
        n0 = np.ascontiguousarray(n0, dtype=np.int32)
        s1 = np.ascontiguousarray(s1, dtype=np.int32)
        n1 = np.ascontiguousarray(n1, dtype=np.int32)
        acc = np.zeros(2, dtype=np.int64)
        sc0 = job.station_streams[i].sample_rate / job.sample_rate
        sc1 = job.station_streams[j].sample_rate / job.sample_rate
        f(_ptr(s0, C.c_int32), _ptr(n0, C.c_int32), C.c_int32(len(s0)), _ptr(s1, C.c_int32),
          _ptr(n1, C.c_int32), C.c_int32(len(s1)), C.c_double(sc0), C.c_double(sc1),
          _ptr(acc, C.c_int64))
        out[b] = acc
    return out


def normalize(job, vis, levels4, n_flagged, precision="f32"):
    keep = []
    cp, bl = _core_params(job, keep)
    nf = job.fft_size_correlation + 1
    v = np.ascontiguousarray(vis, dtype=_dtype(precision)).ravel()
    out = np.zeros(2 * len(bl) * nf, dtype=np.float32)
    l4 = np.ascontiguousarray(levels4, dtype=np.int32)
    nf1 = np.ascontiguousarray(n_flagged[:, 0], dtype=np.int64)
    nf2 = np.ascontiguousarray(n_flagged[:, 1], dtype=np.int64)
    getattr(lib(precision), f"or_normalize_{precision}")(
        C.byref(cp), _rptr(v, precision), _ptr(l4, C.c_int32), _ptr(nf1, C.c_int64),
        _ptr(nf2, C.c_int64), _ptr(out, C.c_float))
    return out.reshape(len(bl), nf, 2)


def pulsar_setup(job, psr, precision="f64"):
    nf = job.fft_size_correlation + 1
    offsets = np.zeros(nf, dtype=np.float64)
    out3 = np.zeros(3, dtype=np.float64)
    coef = np.ascontiguousarray(psr.coef, dtype=np.float64)
    from_ticks = job.start_time
    tpd = 86400 * CLOCK_RATE
    start_mjd = 51544 + from_ticks // tpd + (from_ticks % tpd) / CLOCK_RATE / 86400
    getattr(lib(precision), f"or_pulsar_setup_{precision}")(
        C.c_double(start_mjd), C.c_int(job.window), C.c_int(job.fft_size_correlation),
        C.c_int(job.sample_rate), C.c_int(sb_sign(job.sideband)), C.c_int64(job.channel_freq),
        C.c_double(psr.tmid), C.c_double(psr.DM), C.c_double(psr.ref_phase), C.c_double(psr.ref_freq),
        C.c_double(psr.obs_freq), _ptr(coef, C.c_double), C.c_int(len(coef)), _ptr(offsets, C.c_double),
        _ptr(out3, C.c_double))
    return offsets, out3


def pulsar_params_c(psr, offsets, out3, keep):
    keep.append(offsets)
    return PulsarParamsC(psr.nbins + 1, 0, out3[0], out3[1], out3[2], psr.gate_start, psr.gate_stop,
                         _ptr(offsets, C.c_double))


def pulsar_bins(job, psr, current_fft, precision="f64"):
    offsets, out3 = pulsar_setup(job, psr, precision)
    keep = []
    pp = pulsar_params_c(psr, offsets, out3, keep)
    nf = job.fft_size_correlation + 1
    bins = np.zeros(nf, dtype=np.int32)
    getattr(lib(precision), f"or_pulsar_bins_{precision}")(C.byref(pp), C.c_int(nf),
                                                          C.c_int64(current_fft), _ptr(bins, C.c_int32))
    return bins


def pulsar_correlate(job, psr, spectra, nseg, precision="f32"):
    """or_pulsar_correlate over groups of baselines on the driver's threads."""
    bl = job.baselines()
    offsets, out3 = pulsar_setup(job, psr, "f64")
    nf = job.fft_size_correlation + 1
    arrs = [np.ascontiguousarray(s, dtype=_dtype(precision)) for s in spectra]
    ptrs = (C.POINTER(C.c_float if precision == "f32" else C.c_double) * len(arrs))(
        *[_rptr(a, precision) for a in arrs])
    acc = np.zeros((psr.nbins + 1, len(bl), nf, 2), dtype=_dtype(precision))
    fn = getattr(lib(precision), f"or_pulsar_correlate_{precision}")

    def run(rng):
        a, b = rng
        keep = []
        cp, _ = _core_params_range(job, a, b, keep)
        pp = pulsar_params_c(psr, offsets, out3, keep)
        part = np.zeros((psr.nbins + 1, b - a, nf, 2), dtype=_dtype(precision))
        fn(C.byref(cp), C.byref(pp), ptrs, C.c_int64(nseg), _rptr(part, precision))
        acc[:, a:b] = part

    _map(run, _baseline_groups(len(bl)))
    return acc


def stats_header_levels(bits, levels):
    """Output_header_bitstatistics levels as integration_write stores them
    (correlation_core.cc:378-390)."""
    if bits == 2:
        return [levels[0], levels[1], levels[2], levels[3], levels[4]]
    return [0, levels[0], levels[1], 0, levels[4]]


def write_output(job, sourcenr, integration_slice, vis_f32, stats5, weights, uvw=None, bl=None):
    M = len(job.station_streams)
    if bl is None:
        bl = job.baselines()
    station = np.array([s.station_number for s in job.station_streams], dtype=np.int32)
    pol = np.array([0 if s.polarisation == "R" else 1 for s in job.station_streams], dtype=np.int32)
    if uvw is None:
        uvw = np.zeros(3 * M, dtype=np.float64)
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    st5 = np.ascontiguousarray(np.asarray(stats5, dtype=np.int32).ravel())
    f = np.array([b[0] for b in bl], dtype=np.int32)
    s = np.array([b[1] for b in bl], dtype=np.int32)
    w = np.ascontiguousarray(weights, dtype=np.int32)
    v = np.ascontiguousarray(vis_f32, dtype=np.float32).ravel()
    nchan = job.number_channels
    size = 4 + 16 + 32 * M + 24 * M + len(bl) * (8 + 8 * (nchan + 1))
    dst = np.zeros(size, dtype=np.uint8)
    fn = getattr(lib("f32"), "or_write_output_f32")
    fn.restype = C.c_int64
    n = fn(C.c_int32(sourcenr), C.c_int32(integration_slice), C.c_int(M), _ptr(station, C.c_int32),
           _ptr(pol, C.c_int32), C.c_int(0 if job.sideband == "L" else 1), C.c_int(job.frequency_nr),
           _ptr(uvw, C.c_double), _ptr(st5, C.c_int32), C.c_int(len(bl)), _ptr(f, C.c_int32),
           _ptr(s, C.c_int32), _ptr(w, C.c_int32), C.c_int(nchan), _ptr(v, C.c_float),
           _ptr(dst, C.c_uint8))
    return dst[:n].tobytes()


def irfft(X: np.ndarray, n: int, precision="f32") -> np.ndarray:
    """FFTW c2r of n real points from n/2 + 1 bins (unnormalised)."""
    d = _dtype(precision)
    a = np.empty(2 * (n // 2 + 1), dtype=d)
    a[0::2] = np.real(X)
    a[1::2] = np.imag(X)
    out = np.empty(n, dtype=d)
    getattr(lib(precision), f"or_irfft_{precision}")(_rptr(a, precision), _rptr(out, precision), C.c_int(n))
    return out


def lag_window(kind: int, nchan: int, fft_size: int, precision="f32") -> np.ndarray:
    """create_window without a user window (correlation_core.cc:761-809)."""
    w = np.empty(2 * nchan, dtype=np.float32)
    f = getattr(lib(precision), f"or_lag_window_{precision}")
    if f(C.c_int(kind), C.c_int(nchan), C.c_int(fft_size), _ptr(w, C.c_float)) != 0:
        raise ValueError(f"invalid window {kind}")
    return w


def lag_tables(job, precision="f32"):
    """create_mask / create_window (correlation_core.cc:761-824) from the
    job's Mask_parameters: (normalize, mask[N_c + 1], window[2 nchan])."""
    nc, nch = job.fft_size_correlation, job.number_channels
    mp = getattr(job, "mask", None)
    normalize = bool(mp.normalize) if mp is not None else False
    if mp is not None and mp.mask is not None and len(mp.mask) > 0:
        mask = np.asarray(mp.mask, dtype=np.float32)
        if len(mask) != nc + 1:
            raise ValueError("mask must hold fft_size_correlation + 1 values")
    else:
        mask = np.ones(nc + 1, dtype=np.float32)
    if mp is not None and mp.window is not None and len(mp.window) > 0:
        win = np.asarray(mp.window, dtype=np.float32)
        if len(win) != 2 * nch:
            raise ValueError("window must hold 2 * number_channels values")
    else:
        win = lag_window(job.window, nch, nc, precision)
    return normalize, mask, win


def rechannelize(job, norm: np.ndarray, precision="f32") -> np.ndarray:
    """integration_write's re-channelisation of normalised rows
    (correlation_core.cc:417-440): (rows, N_c + 1) -> (rows, nchan + 1, 2) float32."""
    nc, nch = job.fft_size_correlation, job.number_channels
    normalize, mask, win = lag_tables(job)
    a = np.ascontiguousarray(norm, dtype=np.float32).reshape(-1, nc + 1, 2)
    rows = a.shape[0]
    out = np.zeros((rows, nch + 1, 2), dtype=np.float32)
    rp = RechanParamsC(nc, nch, int(normalize), 0, _ptr(mask, C.c_float), _ptr(win, C.c_float))
    getattr(lib(precision), f"or_rechannelize_{precision}")(C.byref(rp), C.c_int(rows), _ptr(a, C.c_float),
                                                             _ptr(out, C.c_float))
    return out


def tsys_messages(job, stats) -> bytes:
    """Correlation_core::tsys_write (correlation_core.cc:479-513): per stream
    station, frequency number, sideband, polarisation (uint8), the slice start
    (uint64 ticks) and tsys on low/high, off low/high (int64), as MPI_Pack
    lays them out on one host."""
    import struct
    out = bytearray()
    for i, st in enumerate(job.station_streams):
        ts = stats[i][1]
        out += struct.pack("<4BQ4q", st.station_number, job.frequency_nr, 0 if job.sideband == "L" else 1,
                           0 if st.polarisation == "R" else 1, job.start_time, int(ts[0]), int(ts[1]), int(ts[2]),
                           int(ts[3]))
    return bytes(out)


# ----------------------------------------------------------------------------
# Raw VDIF: the input node's reader and writer (test-side restatement)
# ----------------------------------------------------------------------------
def _mjd(day, month, year):
    """utils.cc:255-263."""
    a = (14 - month) // 12
    y = year + 4800 - a
    m = month + 12 * a - 3
    jdn = day + (153 * m + 2) // 5 + 365 * y + y // 4 - y // 100 + y // 400 - 32045
    return int(jdn - 2400000.5)


def vdif_blocks(vdif: bytes, frame_bytes: int, thread_ids, channel: int, sample_rate: int, ref_mjd: int,
                offset: int = 0):
    """VDIF_reader::read_new_block with one frame per block
    (vdif_reader.cc:77-190, set_parameters :202-222) and
    Channel_extractor_tasklet_VDIF's routing: channel `channel`'s blocks as
    (payload bytes, start ticks, invalid)."""
    n = len(vdif) // frame_bytes
    tmap = {int(t): c for c, t in enumerate(thread_ids)}
    first = None
    out = []
    for f in range(n):
        fr = vdif[f * frame_bytes:(f + 1) * frame_bytes]
        w = np.frombuffer(fr[:16], dtype="<u4")
        if np.any(w == 0x11223344):
            continue
        if first is None:
            first = w.copy()
        legacy = (int(first[0]) >> 30) & 1
        hsize = 16 + 16 * (1 - legacy)
        fsize = frame_bytes - hsize
        invalid = (int(w[0]) >> 31) & 1
        thread = (int(w[3]) >> 16) & 0x3FF
        if invalid:
            ch = tmap.get(thread, 0)
        else:
            if thread not in tmap:
                if tmap:
                    continue
                tmap[thread] = 0
            ch = tmap[thread]
        if ch != channel:
            continue
        sec = int(w[0]) & 0x3FFFFFFF
        epoch = (int(w[1]) >> 24) & 0x3F
        jday_epoch = _mjd(1, 1 + 6 * (epoch % 2), 2000 + epoch // 2)
        seconds = float(sec) - (ref_mjd - jday_epoch) * 24 * 60 * 60
        bits = ((int(first[3]) >> 26) & 0x1F) + 1
        nch = 1 << ((int(first[2]) >> 24) & 0x1F)
        spf = 8 * fsize // (bits * nch)
        subsec = (int(w[1]) & 0xFFFFFF) * spf / sample_rate
        t = int((ref_mjd - 51544) * 86400 * float(CLOCK_RATE)) + int(round((seconds + subsec) * CLOCK_RATE)) - offset
        out.append((fr[hsize:], t, bool(invalid)))
    return out


def vdif_phasecal(vdif: bytes, frame_bytes: int, thread_ids, channel: int, job, stream: int, ref_mjd: int,
                  integration_ticks: int, offset: int = 0) -> bytes:
    """Input_node_data_writer::do_phasecal (input_node_data_writer.cc:232-306)
    over channel `channel`'s frame blocks: the messages of the finished
    integrations (station, frequency, sideband, polarisation as uint8; start
    and duration as int64 ticks; the sample count and the sums as int32)."""
    import struct
    st = job.station_streams[stream]
    fs, bits = st.sample_rate, st.bits_per_sample
    spb = 8 // bits
    P = int(fs / 10e3)
    values = np.array([-7, -2, 2, 7] if bits == 2 else [-5, 5], dtype=np.int64)  # sample_value_2 / _1
    acc = None
    count = 0
    pc_time = 0
    out = bytearray()
    for payload, t, invalid in vdif_blocks(vdif, frame_bytes, thread_ids, channel, fs, ref_mjd, offset):
        if t % integration_ticks == 0:
            if acc is not None:
                out += struct.pack("<4Bqqi", st.station_number, job.frequency_nr, 1 if st.sideband == "U" else 0,
                                   1 if st.polarisation == "L" else 0, pc_time, t - pc_time, P)
                out += acc.astype("<i4").tobytes()
            acc = np.zeros(P, dtype=np.int64)
            count = 0
            pc_time = t
        if acc is None:
            continue
        n = len(payload) * spb
        if invalid:
            count += n
            continue
        b = np.frombuffer(payload, dtype=np.uint8)
        codes = (b[:, None] >> (bits * np.arange(spb))[None, :]) & ((1 << bits) - 1)
        count %= P
        np.add.at(acc, (count + np.arange(n)) % P, values[codes.reshape(-1)])
        count += n
    return bytes(out)


def vdif_header_stream(vdif: bytes, frame_bytes: int, thread_ids, channel: int, job, stream: int, ref_mjd: int,
                       offset: int = 0) -> bytes:
    """Input_node_data_writer::do_task (input_node_data_writer.cc:105-228) over
    one slice of VDIF frame blocks: the HEADER stream channel `channel` of the
    station sends for stream `stream` of the job."""
    from sfxg import synth
    st = job.station_streams[stream]
    fs, bits = st.sample_rate, st.bits_per_sample
    spb = 8 // bits
    tps = int(round(CLOCK_RATE / fs))
    byte_length = int(round(8 * 1_000_000.0 / (fs * bits) * (CLOCK_RATE / 1_000_000)))
    nsamples = job.n_ffts_per_integration(stream) * job.fft_size_delaycor
    changes = synth.delay_changes(job.delay_models[stream], job.start_time, nsamples, fs)
    dl = []
    for t, nlev in changes:
        b, r = synth.delay_split(np.array([nlev]), bits)
        dl.append((t, int(b[0]), int(r[0])))
    blocks = vdif_blocks(vdif, frame_bytes, thread_ids, channel, fs, ref_mjd, offset)
    out = bytearray()

    def counted(kind, n):
        while n > 0:
            k = min(n, 32767)
            out.append(kind)
            out.extend(int(k).to_bytes(2, "little"))
            n -= k

    current = job.start_time
    di = 0
    while di < len(dl) - 1 and dl[di + 1][0] <= current + byte_length:
        di += 1
    slice_size = nsamples
    sync = True
    for payload, bstart, invalid in blocks:
        block_size = len(payload)
        byte_offset = 0
        if sync:
            diff = current - bstart
            dsamples = abs(diff) // tps * (1 if diff >= 0 else -1)  # C++ division truncates toward zero
            byte_offset = int(dsamples * bits / 8) if dsamples < 0 else dsamples * bits // 8
            byte_offset += dl[di][1]
            if byte_offset < 0:
                raise ValueError("slice starts before the VDIF data")
            if byte_offset >= block_size:
                continue
            slice_size += dl[di][2]
            out.extend([2, dl[di][2] & 0xFF])
            current -= dl[di][2] * tps
            sync = False

        def ndp_of():
            if di < len(dl) - 1:
                return int((dl[di + 1][0] - current) / byte_length)
            return (2**31 - 1) // 2

        index = byte_offset
        inv = [(0, block_size)] if invalid else []
        inv_i = 0
        nip = inv[0][0] if inv else block_size + 1
        ndp = ndp_of() + byte_offset
        total = min(block_size - byte_offset, (slice_size + spb - 1) // spb)
        end_index = total + byte_offset
        while index < end_index:
            if index >= ndp:
                di += 1
                out.extend([2, dl[di][2] & 0xFF])
                d = dl[di][2] - dl[di - 1][2]
                slice_size += -1 if (d > 1 or d == -1) else 1
                ndp = ndp_of() + byte_offset
                total = min(block_size - byte_offset, (slice_size + spb - 1) // spb)
                end_index = total + byte_offset
            elif index >= nip:
                nrun = inv[inv_i][1]
                end_pos = min(nip + nrun, ndp)
                nr = min(max(0, end_pos - index), end_index - index)
                if nr > 0:
                    counted(1, nr * spb)
                    index += nr
                if end_pos == nip + nrun:
                    inv_i += 1
                    nip = inv[inv_i][0] if inv_i < len(inv) else block_size + 1
            else:
                k = min(ndp - index, end_index - index, nip - index)
                done = 0
                while done < k:
                    c = min(k - done, 32767)
                    out.append(0)
                    out.extend(int(c).to_bytes(2, "little"))
                    out.extend(payload[index + done:index + done + c])
                    done += c
                index += k
        slice_size -= total * spb
        current += total * spb * tps
        if slice_size <= 0:
            out.append(3)
            return bytes(out)
    raise ValueError("the VDIF data ends before the slice")


# ----------------------------------------------------------------------------
def run_slice(job, headers: Sequence[bytes], lut: int = 0, precision: str = "f32",
              keep_intermediate: bool = False, pulsar=None, uvw=None, phased: bool = False):
    """Full correlator-node path for one slice; returns dict with the output
    bytes and intermediates. One integration per slice (the reference asserts
    (stop - start) / integration_time == 1, correlator_node.cc:327-328)."""
    M = len(job.station_streams)

    def stream(i):
        r = bit2float(headers[i], b2f_params(job, i, lut), precision)
        d = delay_correction(job, i, r["samples"], r["chunks"], precision, keep_intermediate)
        if not keep_intermediate:
            r["samples"] = None
        return r, d

    per = _map(stream, range(M))
    b2f = [r for r, _ in per]
    dc = [d for _, d in per]
    stats = []
    for i in range(M):
        st = job.station_streams[i]
        lv, ts = bit_statistics(st.bits_per_sample, st.sample_rate, job.sample_rate,
                                b2f[i]["counts_on"], b2f[i]["counts_off"], b2f[i]["n_invalid"])
        stats.append((lv, ts))
    invalid = [(b["invalid_start"], b["invalid_n"]) for b in b2f]
    nflag = find_invalid(job, invalid)
    levels4 = np.array([s[0][4] for s in stats], dtype=np.int32)
    spectra = [d["spectra"] for d in dc]
    if phased:
        # Correlation_core_phased: one row per sub-integration plus one, each
        # written as stream 0's auto-correlation (correlation_core_phased.cc:84-98)
        bl = [(0, 0)] * (job.number_ffts_in_integration() // job.number_ffts_in_sub_integration() + 1)
    else:
        bl = job.baselines()
    total_samples = job.number_ffts_in_integration() * job.fft_size_correlation
    weights = []
    for b, (i, j) in enumerate(bl):
        lv = stats[i][0]
        if i == j:
            weights.append(max(total_samples - int(lv[4]), 0))
        else:
            weights.append(max(total_samples - int(lv[4]) - int(nflag[b, 0]), 0))
    stats5 = [stats_header_levels(job.station_streams[i].bits_per_sample, stats[i][0]) for i in range(M)]
    outputs = []
    rechan = job.number_channels != job.fft_size_correlation

    def channels(norm):
        # correlation_core.cc:417-444
        return rechannelize(job, norm, precision) if rechan else norm

    if phased:
        rows = phased_correlate(job, spectra, dc[0]["nseg"], precision)
        outputs.append(write_output(job, 0, job.integration_nr, rows, stats5, weights, uvw, bl=bl))
        vis_bins = norm_bins = chan_bins = [rows]
    elif getattr(job, "phase_centers", None):
        # one record per phase centre (correlation_core.cc:52-66)
        acc = correlate_pc(job, spectra, dc[0]["seg_per_chunk"], precision)
        vis_bins, norm_bins, chan_bins = [], [], []
        sources = [job.source_nr] + [pc.source_nr for pc in job.phase_centers]
        uvws = [uvw] + [pc.uvw for pc in job.phase_centers]
        for j in range(acc.shape[0]):
            norm = normalize(job, acc[j], levels4, nflag, precision)
            chan = channels(norm)
            u = None if uvws[j] is None else np.asarray(uvws[j], dtype=np.float64).ravel()
            outputs.append(write_output(job, sources[j], job.integration_nr, chan, stats5, weights, u))
            vis_bins.append(acc[j])
            norm_bins.append(norm)
            chan_bins.append(chan)
    elif pulsar is None:
        vis = correlate(job, spectra, dc[0]["seg_per_chunk"], precision)
        norm = normalize(job, vis, levels4, nflag, precision)
        chan = channels(norm)
        outputs.append(write_output(job, 0, job.integration_nr, chan, stats5, weights, uvw))
        vis_bins = [vis]
        norm_bins = [norm]
        chan_bins = [chan]
    else:
        acc = pulsar_correlate(job, pulsar, spectra, dc[0]["nseg"], precision)
        vis_bins, norm_bins, chan_bins = [], [], []
        for b in range(pulsar.nbins + 1):
            norm = normalize(job, acc[b], levels4, nflag, precision)
            chan = channels(norm)
            outputs.append(write_output(job, b, job.integration_nr, chan, stats5, weights, uvw))
            vis_bins.append(acc[b])
            norm_bins.append(norm)
            chan_bins.append(chan)
    res = dict(output=b"".join(outputs), stats=stats, nflag=nflag, weights=weights,
               vis=vis_bins, norm=norm_bins, channels=chan_bins, seg_per_chunk=dc[0]["seg_per_chunk"],
               chunks=b2f[0]["chunks"], invalid=invalid)
    if keep_intermediate:
        res["b2f"] = b2f
        res["dc"] = dc
    return res
