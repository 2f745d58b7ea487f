"""Pins the CPU restatement (oracle/) against checks that do not retype it:
closed forms, brute force, numpy's FFT and the reference's own layout reader
formats. CPU only."""
import math
import struct

import numpy as np
import pytest

from oracle import oracle as O
from sfxg import configs, synth
from sfxg.job import (CLOCK_RATE, WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN, WINDOW_NONE, WINDOW_RECT,
                      CorrelationParameters, DelayModel, Pulsar, StationStream, mjd_to_ticks)

HD, HI, HDEL, HEND = 0, 1, 2, 3


def b2f_params(bits, n_ffts, nfft_max=4, fft_size=16, lut=0, rate=32_000_000, delay_in_samples=0, tsys=80):
    return dict(bits_per_sample=bits, sample_rate=rate, base_sample_rate=rate, tsys_freq=tsys, fft_size=fft_size,
                nfft_max=nfft_max, n_ffts_per_integration=n_ffts, delay_in_samples=delay_in_samples, lut=lut,
                pad=0)


# ---------------------------------------------------------------------------
# bit2float: decode tables (exhaustive) -- sfxc/src/bit2float_worker.cc:4-33
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bits,lut,levels", [
    (2, 0, [-7, -2, 2, 7]), (2, 1, [-3.3359, -1, 1, 3.3359]), (1, 0, [-5, 5]), (1, 1, [-1, 1])])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_lut_exhaustive(bits, lut, levels, prec):
    spb = 8 // bits
    data = bytes(range(256))
    hdr = bytes([HDEL, 0, HD, 0, 1]) + data + bytes([HEND])
    n = 256 * spb
    r = O.bit2float(hdr, b2f_params(bits, n // 16, nfft_max=n // 16, lut=lut), prec)
    exp = np.array([levels[(b >> (bits * k)) & ((1 << bits) - 1)] for b in range(256) for k in range(spb)])
    np.testing.assert_array_equal(r["samples"], exp.astype(r["samples"].dtype))
    # every byte counted once, all in the initial tsys phase (1-bit: always "on")
    assert r["counts_on"].sum() + r["counts_off"].sum() == 256


# ---------------------------------------------------------------------------
# bit2float: HEADER-stream semantics vs an independent sample-level model
# ---------------------------------------------------------------------------
def model_bit2float(hdr, bits, n_out, nfft_max, fft_size, lut_vals, tsys_count0=None, tsys_on0=True, half=None):
    """Sample-level reading of the HEADER semantics: DATA appends its samples
    after dropping a pending skip; DELAY inserts a zero (delay decrease) or
    schedules a one-sample skip (increase); INVALID appends zeros after
    dropping a pending skip. Bytes are counted once per conversion call
    (calls end at DATA-record and output-buffer boundaries) unless the call
    leaves them partly unread at its end."""
    spb = 8 // bits
    out = []
    skip = 0
    cur = None
    counted = []  # payload bytes in counting order
    i = 0
    payload_base = 0
    chunk = nfft_max * fft_size
    while len(out) < n_out:
        h = hdr[i]
        if h == HDEL:
            d = struct.unpack("b", hdr[i + 1:i + 2])[0]
            if cur is None:
                skip = d
            elif d - cur == -1 or d - cur > 1:
                out.append(0.0)
            else:
                skip = 1
            cur = d
            i += 2
        elif h == HI:
            n = hdr[i + 1] | (hdr[i + 2] << 8)
            drop = min(skip, n)
            skip -= drop
            out.extend([0.0] * (n - drop))
            i += 3
        elif h == HD:
            n = hdr[i + 1] | (hdr[i + 2] << 8)
            data = hdr[i + 3:i + 3 + n]
            samples = [lut_vals[(b >> (bits * k)) & ((1 << bits) - 1)] for b in data for k in range(spb)]
            p = skip
            skip = 0
            while p < len(samples) and len(out) < n_out:
                # one conversion call: up to the next output-buffer boundary
                room = min(chunk - (len(out) % chunk), n_out - len(out))
                q = min(len(samples), p + room)
                first = p // spb
                last = (q - 1) // spb
                counted.append(payload_base + first)
                for b in range(first + 1, last + 1):
                    if b < last or q % spb == 0:
                        counted.append(payload_base + b)
                out.extend(samples[p:q])
                p = q
            payload_base += n
            i += 3 + n
        else:
            raise AssertionError("unexpected header")
    return np.array(out[:n_out]), counted


def random_stream(rng, bits, n_out):
    spb = 8 // bits
    out = bytearray([HDEL, int(rng.integers(0, spb))])
    cur = out[1]
    # the initial sub-byte skip is consumed by a DATA record before any
    # other delay change (the input node starts every slice that way)
    n0 = int(rng.integers(1, 10))
    d0 = rng.integers(0, 256, n0, dtype=np.uint8).tobytes()
    out += bytes([HD, n0, 0]) + d0
    payload = bytearray(d0)
    produced = n0 * spb
    while produced < n_out + 64:
        kind = rng.random()
        if kind < 0.6:
            n = int(rng.integers(1, 40))
            d = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
            out += bytes([HD, n & 255, n >> 8]) + d
            payload += d
            produced += n * spb
        elif kind < 0.8:
            # delay change by +-1 sample with byte wrap, as get_delay splits it
            step = 1 if rng.random() < 0.5 else -1
            new = cur + step
            if new == spb:
                new = 0
            if new < 0:
                new = spb - 1
            out += bytes([HDEL, new])
            cur = new
            if step == 1:
                # a removal needs a byte boundary: the next record is DATA
                n = int(rng.integers(1, 20))
                d = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
                out += bytes([HD, n, 0]) + d
                payload += d
                produced += n * spb
        else:
            n = int(rng.integers(1, 50))
            out += bytes([HI, n, 0])
            produced += n
    out.append(HEND)
    return bytes(out), bytes(payload)


@pytest.mark.parametrize("bits", [1, 2])
@pytest.mark.parametrize("rate", [1.2e-4, -1.2e-4])
def test_integer_shift_closed_form(bits, rate):
    """The integer delay, independently of how the HEADER records express it:
    the correlator's output sample k at time t_k = start + k / fs is the raw
    sample recorded at t_k + tau(t_k) rounded to a sample, i.e. raw index
    k0 + k + floor(tau(t_k) fs + 0.5), for a rising delay (a sample skipped at
    each change) and a falling one (a zero inserted at each change). The
    input node places a change at a byte boundary (input_node_data_writer.cc:
    171-185) found from the change instant's distance to the slice start in
    input bytes, which the earlier changes of the slice have moved by one
    sample each (input_node_data_writer.cc:421-433): around change j the
    samples within j + one byte of it are left out of the comparison, and
    the count and places of the inserted zeros are checked there."""
    fs, nsamp = 32_000_000, 1 << 18
    tps = CLOCK_RATE // fs
    spb = 8 // bits
    rng = np.random.default_rng(11 + bits)
    raw_start = mjd_to_ticks(60000, 0.0)
    start = raw_start + 8192 * tps
    model = DelayModel.linear(raw_start, 5.0e-6, rate, 2)
    raw = rng.integers(0, 256, 2 * (nsamp // spb) + 16384, dtype=np.uint8)
    hdr = synth.header_stream(raw, raw_start, model, start, nsamp, fs, bits)
    delay0 = float(model.eval(start))
    p = b2f_params(bits, nsamp // 256, nfft_max=64, fft_size=256, lut=1, rate=fs,
                   delay_in_samples=int(delay0 * fs))
    out = O.bit2float(hdr, p, "f64")["samples"]
    levels = np.array([-3.3359, -1, 1, 3.3359] if bits == 2 else [-1.0, 1.0])
    codes = (raw[:, None] >> (bits * np.arange(spb))[None, :]) & ((1 << bits) - 1)
    decoded = levels[codes.reshape(-1)]
    k = np.arange(nsamp)
    n_k = np.floor(model.eval(start + k * tps) * fs + 0.5).astype(np.int64)
    k0 = (start - raw_start) // tps
    expect = decoded[k0 + k + n_k]
    change = np.flatnonzero(np.diff(n_k)) + 1
    assert len(change) >= 25
    near = np.zeros(nsamp, dtype=bool)
    for j, c in enumerate(change):
        near[max(0, c - j - spb - 1):c + j + spb + 2] = True
    np.testing.assert_array_equal(out[~near], expect[~near])
    assert np.mean(near) < 0.01
    # inserted zeros: one per falling change, inside that change's window; none when rising
    zeros = np.flatnonzero(out == 0.0)
    if rate < 0:
        assert len(zeros) == len(change)
        assert np.all(np.abs(zeros - change) <= np.arange(len(change)) + spb + 1)
    else:
        assert len(zeros) == 0


@pytest.mark.parametrize("bits", [1, 2])
@pytest.mark.parametrize("seed", range(6))
def test_bit2float_random_streams(bits, seed):
    rng = np.random.default_rng(seed)
    fft_size, nfft_max, n_ffts = 16, 3, 40
    n_out = fft_size * n_ffts
    hdr, payload = random_stream(rng, bits, n_out)
    lut = [-7, -2, 2, 7] if bits == 2 else [-5, 5]
    # tsys: long half-cycle, so all counted bytes fall in the first phase
    r = O.bit2float(hdr, b2f_params(bits, n_ffts, nfft_max=nfft_max, fft_size=fft_size), "f64")
    exp, counted = model_bit2float(hdr, bits, n_out, nfft_max, fft_size, lut)
    np.testing.assert_array_equal(r["samples"], exp)
    hist = np.bincount(np.frombuffer(payload, dtype=np.uint8)[counted], minlength=256) if counted else np.zeros(256)
    np.testing.assert_array_equal(r["counts_on"] + r["counts_off"], hist)
    assert list(r["chunks"]) == [min(nfft_max, n_ffts - k) for k in range(0, n_ffts, nfft_max)]


def test_bit_statistics_levels():
    # 2 bit: every byte contributes its four codes; 1 bit: eight bits
    on = np.zeros(256, dtype=np.int32)
    off = np.zeros(256, dtype=np.int32)
    on[0b11100100] = 3
    off[0x00] = 2
    lv, ts = O.bit_statistics(2, 64_000_000, 64_000_000, on, off, 5)
    np.testing.assert_array_equal(lv, [3 + 8, 3, 3, 3, 5])
    np.testing.assert_array_equal(ts, [6, 6, 0, 8])
    on1 = np.zeros(256, dtype=np.int32)
    on1[0x0F] = 2
    lv1, _ = O.bit_statistics(1, 64_000_000, 64_000_000, on1, np.zeros(256, dtype=np.int32), 0)
    np.testing.assert_array_equal(lv1, [8, 8, 0, 0, 0])


def test_tsys_phase_toggle():
    """tsys on/off half-cycles alternate every (fs / (2 f)) / 4 counted bytes
    (bit2float_worker.cc:299-312,341-345)."""
    rate, f = 4000, 100  # 20 samples per half cycle -> 5 bytes
    n_bytes = 40
    data = bytes((np.arange(n_bytes) % 256).astype(np.uint8))
    hdr = bytes([HDEL, 0, HD, n_bytes, 0]) + data + bytes([HEND])
    r = O.bit2float(hdr, b2f_params(2, n_bytes * 4 // 16, nfft_max=100, lut=0, rate=rate, tsys=f,
                                    delay_in_samples=7), "f64")
    # delay 7 samples into the 40-sample cycle: on-phase with 20 - 7 = 13
    # samples left -> 13 // 4 = 3 bytes, then alternating 5-byte blocks
    state = []
    s = True
    for k in range(n_bytes):
        state.append(s)
        if k == 2 or (k > 2 and (k - 2) % 5 == 0):
            s = not s
    on_exp = np.bincount(np.array([b for b, st in zip(data, state) if st], dtype=np.int64), minlength=256)
    np.testing.assert_array_equal(r["counts_on"], on_exp)


# ---------------------------------------------------------------------------
# FFT -- FFTW conventions of sfxc_fft_float.cc
# ---------------------------------------------------------------------------
def dft(x, sign):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 4, 8, 32, 256])
def test_fft_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        np.testing.assert_allclose(O.fft_c2c(x, sign, "f64"), dft(x, sign), rtol=0, atol=1e-11 * n)


@pytest.mark.parametrize("n", [256, 2048, 8192, 32768])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_fft_numpy(n, prec, tol):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    got = O.fft_c2c(x, -1, prec)
    assert np.max(np.abs(got - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < tol
    got_i = O.fft_c2c(x, 1, prec)
    ref_i = np.fft.ifft(x) * n
    assert np.max(np.abs(got_i - ref_i)) / np.sqrt(np.mean(np.abs(ref_i) ** 2)) < tol
    xr = rng.standard_normal(n)
    refr = np.fft.rfft(xr)
    gotr = O.rfft(xr, prec)
    assert gotr.shape == refr.shape
    assert np.max(np.abs(gotr - refr)) / np.sqrt(np.mean(np.abs(refr) ** 2)) < tol


# ---------------------------------------------------------------------------
# windows -- delay_correction.cc:299-341
# ---------------------------------------------------------------------------
def test_windows():
    n = 64
    i = np.arange(n)
    exp = {
        WINDOW_NONE: (i < n // 2).astype(float),
        WINDOW_RECT: ((i >= n // 4) & (i < 3 * n // 4)).astype(float),
        WINDOW_COS: np.sin(np.pi * i / (n - 1)),
        WINDOW_HAMMING: 0.54 - 0.46 * np.cos(2 * np.pi * i / (n - 1)),
        WINDOW_HANN: np.hanning(n),
    }
    for k, e in exp.items():
        np.testing.assert_allclose(O.window(k, n, "f64"), e, atol=1e-14)


# ---------------------------------------------------------------------------
# Delay_correction: pure tones
# ---------------------------------------------------------------------------
def chunks_for(n_blocks, nfft_max=32):
    return np.array([min(nfft_max, n_blocks - k) for k in range(0, n_blocks, nfft_max)], dtype=np.int32)


def tone_job(delay_s, rate, channel_freq, nchan=256, fs=32_000_000, bw=16_000_000, t_int_usec=4000,
             window=WINDOW_HANN):
    start = mjd_to_ticks(60000)
    st = [StationStream(0, 0, fs, channel_freq, bw)]
    model = DelayModel.linear(start, delay_s, rate, 2)
    tt = int(round(t_int_usec * CLOCK_RATE / 1e6))
    return CorrelationParameters(experiment_start=start, start_time=start, stop_time=start + tt,
                                 integration_time=tt, sub_integration_time=tt, number_channels=nchan,
                                 fft_size_delaycor=256, fft_size_correlation=nchan, sample_rate=fs,
                                 channel_freq=channel_freq, bandwidth=bw, sideband="U", frequency_nr=0,
                                 station_streams=st, window=window, delay_models=[model])


@pytest.mark.parametrize("delay_s", [0.0, 0.3e-7, 1.234567e-6, -2.71e-7])
def test_fractional_shift_tone(delay_s):
    """With the band centre at 0 Hz the fringe rotation vanishes, and a tone on
    a 256-point bin k0 comes out as 128 cos(2 pi k0 n / 256 + theta - phi_k0),
    phi_k = pi (n_int & 3) / 2 - 2 pi frac (k fs / 256 - bw / 2) / fs
    (the phase ramp of delay_correction.cc:97-130 written out)."""
    fs, bw = 32_000_000, 16_000_000
    job = tone_job(delay_s, 0.0, -bw // 2)
    n = job.n_ffts_per_integration(0) * 256
    k0, theta = 37, 0.4
    t = np.arange(n)
    x = np.cos(2 * np.pi * k0 * t / 256 + theta)
    chunks = chunks_for(n // 256)
    d = O.delay_correction(job, 0, x, chunks, "f64", want_time_series=True)
    dis = delay_s * fs
    n_int = math.floor(dis + 0.5)
    frac = dis - n_int
    phi = math.pi * (n_int & 3) / 2 - 2 * math.pi * frac * (k0 * fs / 256 - bw / 2) / fs
    exp = 128 * np.cos(2 * np.pi * k0 * t / 256 + theta - phi)
    np.testing.assert_allclose(d["time_series"], exp, atol=1e-9)


def test_fringe_stopping_tone():
    """A tone delayed at RF, x(t) = cos(2 pi f_b t - 2 pi (nu_LO + f_b) tau(t)),
    with a fast-changing delay: after integer shift, fractional-delay and
    fringe correction its spectral phase is constant over the segments. A
    sign error in either correction leaves a phase rotating at the fringe
    rate (here nu * rate = 4 kHz)."""
    fs, bw, nu = 32_000_000, 16_000_000, 4_000_000_000
    rate = 1e-6
    job = tone_job(3.3e-7, rate, nu, nchan=256, t_int_usec=2000)
    n = job.n_ffts_per_integration(0) * 256
    fb = 64 * fs / 512  # on channel 64 of the 256-channel spectrum
    m = job.delay_models[0]
    tps = CLOCK_RATE // fs
    ticks = job.start_time + np.arange(n) * tps
    blk_mid = job.start_time + (np.arange(n) // 256) * 256 * tps + (256 * tps) // 2
    n_int = np.floor(m.eval(blk_mid) * fs + 0.5)  # integer shift applied upstream
    t_raw = (ticks - job.start_time) / CLOCK_RATE + n_int / fs
    tau = m.eval(job.start_time + np.round(t_raw * CLOCK_RATE).astype(np.int64))
    x = np.cos(2 * np.pi * fb * t_raw - 2 * np.pi * (nu + fb) * tau)
    chunks = chunks_for(n // 256)
    d = O.delay_correction(job, 0, x, chunks, "f64")
    spec = d["spectra"][..., 0] + 1j * d["spectra"][..., 1]
    ph = np.angle(spec[:, 64])
    dph = np.angle(np.exp(1j * (ph - ph[0])))
    assert np.max(np.abs(dph)) < 0.02, np.max(np.abs(dph))
    assert np.abs(spec[:, 64]).min() > 0.5 * np.abs(spec[:, 64]).max()


# ---------------------------------------------------------------------------
# Correlation_core
# ---------------------------------------------------------------------------
def small_job(M=4, nchan=256, seconds_usec=8000, sub_usec=2000):
    start = mjd_to_ticks(60000)
    fs = 32_000_000
    st = [StationStream(i // 2, i, fs, 5_000_000_000, 16_000_000, polarisation="RL"[i % 2]) for i in range(M)]
    tt = int(round(seconds_usec * CLOCK_RATE / 1e6))
    return CorrelationParameters(experiment_start=start, start_time=start, stop_time=start + tt,
                                 integration_time=tt, sub_integration_time=int(round(sub_usec * CLOCK_RATE / 1e6)),
                                 number_channels=nchan, fft_size_delaycor=256, fft_size_correlation=nchan,
                                 sample_rate=fs, channel_freq=5_000_000_000, bandwidth=16_000_000, sideband="U",
                                 frequency_nr=3, station_streams=st,
                                 delay_models=[DelayModel.linear(start, 0, 0, 2)] * M)


def test_baselines_order():
    job = small_job(M=6)
    bl = job.baselines()
    assert bl[:6] == [(i, i) for i in range(6)]
    # no cross products between the two polarisations of one station
    crosses = bl[6:]
    assert all(job.station_streams[a].station_number != job.station_streams[b].station_number for a, b in crosses)
    assert len(crosses) == 6 * 5 // 2 - 3


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-5)])
def test_correlate_einsum(prec, tol):
    job = small_job()
    M, nc = 4, 256
    nseg = job.number_ffts_in_integration()
    rng = np.random.default_rng(3)
    spectra = [(rng.standard_normal((nseg, nc + 4)) + 1j * rng.standard_normal((nseg, nc + 4))) for _ in range(M)]
    arrs = [np.stack([s.real, s.imag], -1) for s in spectra]
    chunks = [1] + [2] * ((nseg - 1) // 2) + ([1] if (nseg - 1) % 2 else [])
    assert sum(chunks) == nseg
    vis = O.correlate(job, arrs, chunks, prec)
    v = vis[..., 0] + 1j * vis[..., 1]
    for b, (i, j) in enumerate(job.baselines()):
        exp = np.einsum("sk,sk->k", spectra[i][:, :nc + 1], np.conj(spectra[j][:, :nc + 1]))
        assert np.max(np.abs(v[b] - exp)) / np.sqrt(np.mean(np.abs(exp) ** 2)) < tol


def test_normalize_closed_form():
    job = small_job()
    M, nc = 4, 256
    B = len(job.baselines())
    rng = np.random.default_rng(5)
    vis = rng.standard_normal((B, nc + 1, 2)) + 2.0
    levels4 = np.array([0, 10, 0, 0], dtype=np.int32)
    nflag = np.zeros((B, 2), dtype=np.int64)
    nflag[M, 0] = 100
    out = O.normalize(job, vis, levels4, nflag, "f64")
    norms = np.maximum(vis[:M, :, 0].sum(1) / nc, 1)
    total = job.number_ffts_in_integration() * nc
    for b, (i, j) in enumerate(job.baselines()):
        if i == j:
            np.testing.assert_allclose(out[b, :, 0], vis[b, :, 0] / norms[i], rtol=1e-6)
            assert np.all(out[b, :, 1] == 0)
        else:
            N1 = 1 - nflag[b, 0] / (total - levels4[i])
            N2 = 1 - nflag[b, 1] / (total - levels4[j])
            np.testing.assert_allclose(out[b], vis[b] / np.sqrt(N1 * N2 * norms[i] * norms[j]), rtol=1e-6)


def test_find_invalid_disjoint_and_equal():
    """Disjoint invalid runs flag the other stream for their full length;
    identical runs flag nothing."""
    job = small_job(M=4)
    b = job.baselines().index((0, 2))
    none = (np.zeros(0), np.zeros(0))
    out = O.find_invalid(job, [(np.array([10, 100]), np.array([5, 20])), none,
                               (np.array([50]), np.array([7])), none])
    assert list(out[b]) == [7, 25]
    out2 = O.find_invalid(job, [(np.array([10]), np.array([5])), none, (np.array([10]), np.array([5])), none])
    assert list(out2[b]) == [0, 0]


# ---------------------------------------------------------------------------
# injected pulsar: dispersion law, then the oracle's dedispersion
# ---------------------------------------------------------------------------
def test_synth_dispersion_law():
    """The injector's dispersion against the cold-plasma law: a 0.2 ms pulse
    of noise, dispersed with DM 68 across 1400-1416 MHz, arrives in each of
    eight sub-bands at 4148.808 s DM ((f/MHz)^-2 - 1400^-2) after its time at
    1400 MHz (lower frequencies later), to 40 us (the pulse's power centroid
    per sub-band; the arrivals span 3.2 ms)."""
    import scipy.fft as sf
    fs, n = 32_000_000, 1 << 21
    rng = np.random.default_rng(3)
    t = np.arange(n) / fs
    t0 = 0.03
    a = synth.analytic(rng.standard_normal(n)) * np.exp(-0.5 * ((t - t0) / 2e-4) ** 2)
    d = synth.disperse(a, fs, 1.4e9, 68.0, 1400.0)
    D = sf.fft(d)
    f = sf.fftfreq(n, 1 / fs)
    for k in range(8):
        lo, hi = 2e6 * k, 2e6 * (k + 1)
        sub = sf.ifft(np.where((f >= lo) & (f < hi), D, 0))
        p = np.abs(sub) ** 2
        fc = 1400.0 + (lo + hi) / 2e6
        expect = t0 + 4148.808 * 68.0 * (fc ** -2 - 1400.0 ** -2)
        w = np.abs(t - expect) < 2e-3
        arrival = np.sum(t[w] * p[w]) / np.sum(p[w])  # power centroid
        assert abs(arrival - expect) < 4e-5, (k, arrival, expect)


def test_dedispersion_aligns_injected_pulse():
    """A pulse injected with DM 68 (lower frequencies later) lines up in the
    same phase bins across the band once Correlation_core_pulsar
    dedisperses it with the per-channel offsets 4148.808 DM (f^-2 -
    f_obs^-2) F (correlation_core_pulsar.cc:80-93), at the bin of the pulse's
    phase 0.5; dedispersed with the opposite sign (a planted error, given
    as -DM to the oracle) its peak moves by about seven bins across the band.
    Profile per quarter band: the mean real part of the cross-correlation
    per bin; its centroid over the bins above the median."""
    import dataclasses
    from sfxg.job import REFERENCE_MJD, SECONDS_PER_DAY, Pulsar, ticks_to_mjd
    seconds = 0.2
    fs, sky = 32_000_000, 1_400_000_000
    start = mjd_to_ticks(configs.START_MJD, 0.0)
    psr = Pulsar(nbins=64, gate_start=0.25, gate_stop=0.75, tmid=ticks_to_mjd(start) - 0.001, DM=68.0,
                 ref_phase=0.0, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    tmid_ticks = (psr.tmid - REFERENCE_MJD) * SECONDS_PER_DAY * CLOCK_RATE
    psr = dataclasses.replace(psr, ref_phase=(0.2 - psr.ref_freq * (start - tmid_ticks) / CLOCK_RATE) % 1.0)
    job, hdrs = configs.build_slices(seed=44, n_stations=2, pols=["R"], fs=fs, bw=fs // 2, nchan=1024,
                                     seconds=seconds, t_int_usec=int(seconds * 1e6), rho=0.005, bits=2,
                                     sky_freq=sky, offsets=[0.0, 1.5e-6], rates=[0.0, 1e-7],
                                     pulsar=(psr, 0.2, 0.05))[0]
    nb = psr.nbins + 1
    expect_bin = 1 + int((0.5 - psr.gate_start) * (nb - 1) / (psr.gate_stop - psr.gate_start))

    def centroids(p):
        r = O.run_slice(job, hdrs, lut=1, pulsar=p)
        v = np.array([r["norm"][b][2, :, 0] for b in range(nb)])  # cross baseline, real part
        out = []
        for q in range(4):
            prof = v[1:, 256 * q:256 * (q + 1)].mean(axis=1)
            w = np.maximum(prof - np.median(prof), 0.0)
            out.append(1 + np.sum(np.arange(nb - 1) * w) / np.sum(w))
        return np.array(out), v

    good, v = centroids(psr)
    assert np.max(good) - np.min(good) < 1.0, good
    assert np.all(np.abs(good - expect_bin) < 1.5), (good, expect_bin)
    # the pulse is there: peak bins well above the off-gate bin 0
    assert v[expect_bin, :].mean() > 5 * abs(v[0, :].mean()) + 0.02
    bad, _ = centroids(dataclasses.replace(psr, DM=-psr.DM))
    assert abs(bad[3] - bad[0]) > 4.0, bad


# ---------------------------------------------------------------------------
# pulsar bins: linear ephemeris closed form
# ---------------------------------------------------------------------------
def test_pulsar_bins_ephemeris():
    job = small_job(nchan=1024)
    f0 = 1 / 0.0893
    start_mjd = 60000.0
    psr = Pulsar(nbins=64, gate_start=0.1, gate_stop=0.7, tmid=start_mjd - 0.01, DM=68.0, ref_phase=12.25,
                 ref_freq=f0, obs_freq=5000.0, coef=[0.3, 0.0])
    fs, nc = job.sample_rate, 1024
    seg_us = nc * 1e6 / fs
    f_ch = (job.channel_freq + np.arange(nc + 1) * fs / (2 * nc)) * 1e-6  # MHz
    for s in (0, 17, 999):
        bins = O.pulsar_bins(job, psr, s)
        # pulse phase at the segment centre (windowed: one full FFT length in)
        t_min = ((start_mjd - psr.tmid) + (seg_us * (1 + s)) / 86400e6) * 1440
        phase = psr.ref_phase - math.floor(psr.ref_phase) + 60 * f0 * t_min + psr.coef[0]
        phase = phase - 4148.808 * psr.DM * (f_ch ** -2 - psr.obs_freq ** -2) * f0
        frac = phase - np.floor(phase)
        inside = (frac >= psr.gate_start) & (frac < psr.gate_stop)
        exp = np.where(inside, np.floor((frac - psr.gate_start) * 64 / 0.6).astype(int) + 1, 0)
        mism = np.nonzero(bins != exp)[0]
        # bins agree except where the two FP64 phase formulations straddle an edge
        assert len(mism) <= 2, (s, mism)


# ---------------------------------------------------------------------------
# output layout vs the struct formats of utils/print_corfile.py
# ---------------------------------------------------------------------------
def test_output_layout():
    job = small_job(M=4)
    M, nc = 4, 256
    B = len(job.baselines())
    vis = np.arange(B * (nc + 1) * 2, dtype=np.float32).reshape(B, nc + 1, 2)
    stats5 = [[1, 2, 3, 4, 5]] * M
    uvw = np.arange(3 * M, dtype=np.float64)
    out = O.write_output(job, 0, 7, vis, stats5, list(range(B)), uvw)
    o = 0
    assert struct.unpack("i", out[o:o + 4])[0] == 0
    o += 4
    ts = struct.unpack("4i", out[o:o + 16])
    assert ts == (7, B, 2, M)
    o += 16
    for k in range(2):
        h = struct.unpack("2i3d", out[o:o + 32])
        assert h[0] == k and h[2:] == tuple(uvw[6 * k:6 * k + 3])
        o += 32
    for i in range(M):
        h = struct.unpack("4B5i", out[o:o + 24])
        assert h[0] == i // 2 and h[1] == 3 and h[2] == 1 and h[3] == i % 2 and h[4:] == (1, 2, 3, 4, 5)
        o += 24
    for b, (i, j) in enumerate(job.baselines()):
        h = struct.unpack("i4B", out[o:o + 8])
        assert h[0] == b and h[1] == i // 2 and h[2] == j // 2
        assert h[3] & 3 == (i % 2) | ((j % 2) << 1) and (h[3] >> 2) & 1 == 1 and h[3] >> 3 == 3
        o += 8
        np.testing.assert_array_equal(np.frombuffer(out[o:o + 8 * (nc + 1)], dtype=np.float32),
                                      vis[b].ravel())
        o += 8 * (nc + 1)
    assert o == len(out)


# ---------------------------------------------------------------------------
# whole path: the normalised cross-correlation of two stations equals the
# 2-bit quantised correlation of rho (bivariate-normal expectation)
# ---------------------------------------------------------------------------
def quantised_corr(rho, levels, thr):
    from scipy import integrate, stats
    edges = [-np.inf, -thr, 0.0, thr, np.inf]
    # E[q(x) q(y)] / E[q^2] via the joint cell probabilities
    num = 0.0
    for a in range(4):
        for b in range(4):
            p = stats.multivariate_normal(mean=[0, 0], cov=[[1, rho], [rho, 1]]).cdf(
                [min(edges[a + 1], 9), min(edges[b + 1], 9)], lower_limit=[max(edges[a], -9), max(edges[b], -9)])
            num += levels[a] * levels[b] * p
    pm = [stats.norm.cdf(min(edges[a + 1], 9)) - stats.norm.cdf(max(edges[a], -9)) for a in range(4)]
    den = sum(l * l * p for l, p in zip(levels, pm))
    del integrate
    return num / den


@pytest.mark.slow
def test_pipeline_correlation_amplitude():
    slices = configs.config0(seconds=0.125, t_int_usec=125000)
    job, hdrs = slices[0]
    r = O.run_slice(job, hdrs, lut=1)
    v = r["norm"][0][2]
    v = v[:, 0] + 1j * v[:, 1]
    nc = job.fft_size_correlation
    # the delay model removes the geometric phase, so the correlation sits in
    # the real part; averaging the complex values avoids the noise bias of |V|
    central = v[nc // 8:7 * nc // 8]
    amp = np.mean(central.real)
    assert abs(np.mean(central.imag)) < 0.1 * amp
    exp = quantised_corr(0.05, [-3.3359, -1, 1, 3.3359], 0.9816)
    # per-channel noise ~ 1/sqrt(segments); the mean over 768 channels is
    # good to ~1 % of the expectation (fractional-delay interpolation of the
    # generator loses a little coherence, hence the one-sided margin)
    assert 0.92 * exp < amp < 1.04 * exp, (amp, exp)
    # autos normalise to a mean of ~1
    a = r["norm"][0][0][:, 0]
    assert abs(a[1:-1].mean() - 1) < 0.01
