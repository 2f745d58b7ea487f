"""Seeded synthetic station data and the input-node HEADER byte stream.

Signal model: x_i(t) = sqrt(rho) * s(t - tau_i(t)) + sqrt(1 - rho) * n_i(t),
s and n_i i.i.d. N(0, 1), quantised to 2 bits (thresholds +/- v*sigma, codes
offset binary {0:-hi, 1:-lo, 2:+lo, 3:+hi}) or 1 bit (threshold 0), packed
first-sample-in-LSB (the order sfxc/src/bit2float_worker.cc:25-31 decodes).

The HEADER stream the correlator node receives for one slice is produced by
restating the input node's integer-delay rule:
  * Input_node_tasklet::get_delay / get_delays (sfxc/src/input_node_tasklet.cc:235-282):
    n = floor(tau * fs + 0.5), split into whole bytes and remaining samples,
    change instants at sample resolution;
  * Input_node_data_writer::do_task / write_* (sfxc/src/input_node_data_writer.cc:105-228,
    388-462): DELAY(remaining) at slice start, DATA runs broken at delay changes
    (byte positions from get_next_delay_pos, :421-433) and at SHRT_MAX bytes,
    DELAY at each change, ENDSTREAM at the end.
The input node's read buffer is taken as one block covering the slice, so
DATA records are not additionally split at input-block boundaries (the
correlator node's result does not depend on where DATA records are split).
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

from .job import CLOCK_RATE, DelayModel, usec_to_ticks

HEADER_DATA, HEADER_INVALID, HEADER_DELAY, HEADER_ENDSTREAM = 0, 1, 2, 3
SHRT_MAX = 32767


# --------------------------------------------------------------------------
# integer delay rule
# --------------------------------------------------------------------------
def delay_split(n_samples: np.ndarray, bits: int):
    """Input_node_tasklet::get_delay (input_node_tasklet.cc:45-63): split an
    integer sample delay into (bytes, remaining samples)."""
    spb = 8 // bits
    n = np.asarray(n_samples, dtype=np.int64)
    nbytes = np.floor((n - 1) / spb).astype(np.int64)
    rem = n - nbytes * spb
    full = rem * bits == 8
    nbytes = np.where(full, nbytes + 1, nbytes)
    rem = np.where(full, 0, rem)
    return nbytes, rem


def integer_delay(model: DelayModel, ticks, fs: int) -> np.ndarray:
    return np.floor(model.eval(ticks) * fs + 0.5).astype(np.int64)


def delay_changes(model: DelayModel, start_ticks: int, nsamples: int, fs: int) -> List[Tuple[int, int]]:
    """Sample instants (tick, n) where floor(tau*fs + 0.5) changes inside the
    slice: what the bisection of get_delays (input_node_tasklet.cc:235-262)
    finds for a monotonic delay model (one change per sample step at most).
    Returns the list including the start delay as its first element."""
    tps = int(round(CLOCK_RATE / fs))
    out = [(start_ticks, int(integer_delay(model, start_ticks, fs)))]
    n0 = out[0][1]
    n1 = int(integer_delay(model, start_ticks + nsamples * tps, fs))
    if n0 == n1:
        return out
    step = 1 if n1 > n0 else -1
    # one bisection per integer level, all levels at once: first sample index
    # k at which n(k) has reached level m
    m = np.arange(n0 + step, n1 + step, step, dtype=np.int64)
    lo = np.zeros(m.shape, dtype=np.int64)
    hi = np.full(m.shape, nsamples, dtype=np.int64)
    while np.any(hi - lo > 1):
        mid = (lo + hi) // 2
        nm = integer_delay(model, start_ticks + mid * tps, fs)
        reached = (nm - m) * step >= 0
        hi = np.where(reached, mid, hi)
        lo = np.where(reached, lo, mid)
    for k, level in zip(hi.tolist(), m.tolist()):
        out.append((start_ticks + k * tps, level))
    return out


def header_stream(raw, raw_start_ticks: int, model: DelayModel, start_ticks: int,
                  nsamples: int, fs: int, bits: int, invalid=None, elided: bool = False):
    """Input_node_data_writer::do_task for one slice (input_node_data_writer.cc:105-228).

    raw: the station's packed byte stream whose first byte holds the sample
    at raw_start_ticks (with elided=True only its length is used: an int).
    invalid: optional list of (byte_start, nbytes) ranges of raw flagged
    invalid (the Input_buffer_element invalid list).
    elided=True leaves the payload bytes out of the DATA records and returns
    (control_bytes, first_payload_byte_in_raw, n_payload_bytes); the payload
    is then raw[first:first + n] when no invalid ranges are given.
    """
    spb = 8 // bits
    tps = int(round(CLOCK_RATE / fs))
    byte_length = usec_to_ticks(8 * 1_000_000.0 / (fs * bits))
    changes = delay_changes(model, start_ticks, nsamples, fs)
    dl = []
    for t, n in changes:
        b, r = delay_split(np.array([n]), bits)
        dl.append((t, int(b[0]), int(r[0])))
    out = bytearray()

    def w_delay(r):
        out.extend([HEADER_DELAY, r & 0xFF])

    def w_data(n, start):
        written = 0
        while written < n:
            k = min(n - written, SHRT_MAX)
            out.append(HEADER_DATA)
            out.extend(int(k).to_bytes(2, "little"))
            if not elided:
                out.extend(raw[start + written:start + written + k].tobytes())
            written += k

    def w_invalid(nsamp):
        written = 0
        while written < nsamp:
            k = min(nsamp - written, SHRT_MAX)
            out.append(HEADER_INVALID)
            out.extend(int(k).to_bytes(2, "little"))
            written += k

    current = start_ticks
    di = 0
    while di < len(dl) - 1 and dl[di + 1][0] <= current + byte_length:
        di += 1
    dsamples = (current - raw_start_ticks) // tps
    byte_offset = dsamples * bits // 8 + dl[di][1]
    if byte_offset < 0:
        raise ValueError("slice starts before the raw data")
    slice_size = nsamples + dl[di][2]
    w_delay(dl[di][2])
    current -= dl[di][2] * tps

    def next_delay_pos():
        if di < len(dl) - 1:
            pos = int((dl[di + 1][0] - current) / byte_length)
        else:
            pos = (2**31 - 1) // 2
        return max(pos, 0)

    inv = sorted(invalid or [])
    inv_i = 0
    block_size = raw if elided else len(raw)
    index = byte_offset
    ndp = next_delay_pos() + byte_offset
    total = min(block_size - byte_offset, (slice_size + spb - 1) // spb)
    end_index = total + byte_offset
    nip = inv[0][0] if inv else block_size + 1
    while index < end_index:
        if index >= ndp:
            di += 1
            w_delay(dl[di][2])
            d = dl[di][2] - dl[di - 1][2]
            slice_size += -1 if (d > 1 or d == -1) else 1
            ndp = next_delay_pos() + byte_offset
            total = min(block_size - byte_offset, (slice_size + spb - 1) // spb)
            end_index = total + byte_offset
        elif index >= nip:
            n = inv[inv_i][1]
            end_pos = min(nip + n, ndp)
            nr = max(0, end_pos - index)
            nr = min(nr, end_index - index)
            if nr > 0:
                w_invalid(nr * spb)
                index += nr
            if end_pos == nip + n:
                inv_i += 1
                nip = inv[inv_i][0] if inv_i < len(inv) else block_size + 1
        else:
            k = min(ndp - index, end_index - index, nip - index)
            w_data(k, index)
            index += k
    if index > block_size:
        raise ValueError("raw data too short for the slice")
    slice_size -= total * spb
    if slice_size > 0:
        raise ValueError("raw data too short for the slice")
    out.append(HEADER_ENDSTREAM)
    if elided:
        return bytes(out), byte_offset, index - byte_offset
    return bytes(out)


def join_elided(control: bytes, payload: np.ndarray) -> bytes:
    """Re-insert payload bytes into an elided HEADER stream (inverse of
    header_stream(..., elided=True) without invalid ranges)."""
    out = bytearray()
    pos = 0
    i = 0
    pl = payload.tobytes() if isinstance(payload, np.ndarray) else bytes(payload)
    while i < len(control):
        h = control[i]
        if h == HEADER_DATA:
            n = control[i + 1] | (control[i + 2] << 8)
            out.extend(control[i:i + 3])
            out.extend(pl[pos:pos + n])
            pos += n
            i += 3
        elif h == HEADER_DELAY:
            out.extend(control[i:i + 2])
            i += 2
        elif h == HEADER_INVALID:
            out.extend(control[i:i + 3])
            i += 3
        else:
            out.append(h)
            i += 1
    return bytes(out)


# --------------------------------------------------------------------------
# VDIF frames
# --------------------------------------------------------------------------
def vdif_epoch_mjd(ref_epoch: int) -> int:
    """MJD of the first day of a VDIF reference epoch (half years from 2000)."""
    year, month = 2000 + ref_epoch // 2, 1 + 6 * (ref_epoch % 2)
    a = (14 - month) // 12
    y = year + 4800 - a
    m = month + 12 * a - 3
    jdn = 1 + (153 * m + 2) // 5 + 365 * y + y // 4 - y // 100 + y // 400 - 32045
    return jdn - 2400001


def vdif_header(sec: int, frame: int, ref_epoch: int, frame_bytes: int, thread: int, bits: int,
                invalid: bool = False, legacy: bool = False, station: int = 0x4a42) -> bytes:
    """A VDIF data frame header (VDIF specification, 32 bytes, 16 in legacy
    mode): word 0 seconds, legacy and invalid flags; word 1 frame number in the
    second and reference epoch; word 2 frame length in 8-byte units, log2 of
    the channel count, version; word 3 station, thread, bits per sample - 1."""
    w0 = (sec & 0x3FFFFFFF) | (int(legacy) << 30) | (int(invalid) << 31)
    w1 = (frame & 0xFFFFFF) | ((ref_epoch & 0x3F) << 24)
    w2 = ((frame_bytes // 8) & 0xFFFFFF) | (0 << 24) | (0 << 29)
    w3 = (station & 0xFFFF) | ((thread & 0x3FF) << 16) | (((bits - 1) & 0x1F) << 26)
    words = [w0, w1, w2, w3] + ([] if legacy else [0, 0, 0, 0])
    return np.array(words, dtype="<u4").tobytes()


def vdif_stream(raws: Sequence[np.ndarray], thread_ids: Sequence[int], raw_start_ticks: int, fs: int, bits: int,
                payload: int, ref_epoch: int, invalid=(), fill=(), legacy: bool = False) -> bytes:
    """Multi-thread VDIF frames of one station: thread_ids[c] carries raws[c]
    (packed bytes whose first sample is at raw_start_ticks, on a frame
    boundary), one frame per thread in turn. invalid: (thread index, frame)
    pairs flagged invalid (their payload is kept); fill: frame slots written
    as fill pattern before the frame of that index."""
    tps = int(round(CLOCK_RATE / fs))
    frame_samples = payload * 8 // bits
    fps = fs // frame_samples
    assert fps * frame_samples == fs, "frames per second must be whole"
    epoch_ticks = (vdif_epoch_mjd(ref_epoch) - 51544) * 86400 * CLOCK_RATE
    t0 = raw_start_ticks - epoch_ticks
    assert t0 >= 0 and t0 % (frame_samples * tps) == 0, "raw data must start on a frame boundary"
    first = t0 // (frame_samples * tps)
    n_frames = min(len(r) for r in raws) // payload
    hdr_len = 16 if legacy else 32
    out = bytearray()
    inv = set(invalid)
    for f in range(n_frames):
        if f in set(fill):
            out.extend(np.full((payload + hdr_len) // 4, 0x11223344, dtype="<u4").tobytes())
        g = first + f
        sec, fr = g // fps, g % fps
        for c, (r, th) in enumerate(zip(raws, thread_ids)):
            out.extend(vdif_header(sec, fr, ref_epoch, payload + hdr_len, th, bits, invalid=(c, f) in inv,
                                   legacy=legacy))
            out.extend(np.asarray(r[f * payload:(f + 1) * payload], dtype=np.uint8).tobytes())
    return bytes(out)


# --------------------------------------------------------------------------
# signal synthesis
# --------------------------------------------------------------------------
def _frac_delay_fir(x: np.ndarray, frac: float, taps: int = 32) -> np.ndarray:
    """Delay x by frac samples (0 <= frac < 1) with a Blackman-windowed sinc."""
    k = np.arange(-taps // 2 + 1, taps // 2 + 1)
    h = np.sinc(k - frac) * np.blackman(taps + 2)[1:-1]
    h /= h.sum()
    return np.convolve(x, h, mode="same")


def analytic(x: np.ndarray) -> np.ndarray:
    """x + i H(x) by FFT (length padded to an FFT-friendly size; x is white
    noise, so the periodic extension is as good as any)."""
    import scipy.fft as sf
    n0 = len(x)
    n = sf.next_fast_len(n0)
    if n > n0:
        x = np.concatenate([x, np.zeros(n - n0)])
    X = sf.rfft(x, workers=-1)
    X[1:(n + 1) // 2] *= 2
    full = np.zeros(n, dtype=np.complex128)
    full[:len(X)] = X
    if n % 2 == 0:
        full[n // 2] = X[n // 2]
    return sf.ifft(full, workers=-1, overwrite_x=True)[:n0]


def delayed_signal(common_a: np.ndarray, common_start: int, n: int, delays_samples: np.ndarray,
                   fringe_cycles: np.ndarray, block: int = 4096) -> np.ndarray:
    """Real baseband sample k (k = 0 .. n-1) of a signal delayed at RF:
    Re{ s_a(k - d(k)) exp(-2 pi i nu_LO tau(k)) } (upper sideband), with
    common_a[j] the analytic baseband signal s_a at sample common_start + j,
    d(k) given per block of `block` samples (a linear delay moves by far less
    than a sample per block) and nu_LO tau(k) given per sample in cycles."""
    out = np.empty(n, dtype=np.float64)
    for b0 in range(0, n, block):
        b1 = min(n, b0 + block)
        d = delays_samples[b0 // block]
        di = int(math.floor(d))
        frac = d - di
        lo = b0 - di - common_start - 32
        seg = common_a[lo:lo + (b1 - b0) + 64]
        y = _frac_delay_fir(seg.real, frac) + 1j * _frac_delay_fir(seg.imag, frac)
        ph = np.exp(-2j * np.pi * fringe_cycles[b0:b1])
        out[b0:b1] = (y[32:32 + (b1 - b0)] * ph).real
    return out


def quantise(x: np.ndarray, bits: int, threshold: float = 0.9816) -> np.ndarray:
    if bits == 2:
        return (np.digitize(x, [-threshold, 0.0, threshold])).astype(np.uint8)
    return (x >= 0).astype(np.uint8)


def pack(codes: np.ndarray, bits: int) -> np.ndarray:
    spb = 8 // bits
    c = codes.reshape(-1, spb).astype(np.uint8)
    out = np.zeros(c.shape[0], dtype=np.uint8)
    for k in range(spb):
        out |= (c[:, k] << (bits * k)).astype(np.uint8)
    return out


def pulsar_power(ticks, psr, peak: float, duty: float):
    """Pulse power at geocentric time `ticks` (int64 array) as it arrives at
    the pulsar's observing frequency: a Gaussian in pulse phase centred on
    phase 0.5 with full width at half maximum `duty` (in periods) and height
    `peak`. The phase is the linear ephemeris Correlation_core_pulsar
    evaluates (get_phase, correlation_core_pulsar.h:40-44, with the start
    phase of correlation_core_pulsar.cc:46-64): frac(ref_phase) + c0 +
    (c1 + 60 F0) minutes since tmid; higher polyco terms must be zero."""
    from .job import REFERENCE_MJD, SECONDS_PER_DAY
    if any(c != 0.0 for c in psr.coef[2:]):
        raise ValueError("pulse injection takes a linear ephemeris")
    tmid_ticks = (psr.tmid - REFERENCE_MJD) * SECONDS_PER_DAY * CLOCK_RATE
    minutes = (np.asarray(ticks, dtype=np.float64) - tmid_ticks) / CLOCK_RATE / 60.0
    phase = (psr.ref_phase - math.floor(psr.ref_phase)) + psr.coef[0] + (psr.coef[1] + 60.0 * psr.ref_freq) * minutes
    x = phase - 0.5
    x = x - np.floor(x + 0.5)
    sigma = duty / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    return peak * np.exp(-0.5 * (x / sigma) ** 2)


def disperse(a: np.ndarray, fs: int, lo_freq: float, dm: float, obs_freq_mhz: float) -> np.ndarray:
    """Cold-plasma dispersion of an analytic baseband signal (upper sideband,
    sky frequency lo_freq + f): the component at sky frequency f is delayed
    by 4148.808 s DM ((f/MHz)^-2 - (obs/MHz)^-2), so a pulse keeps its
    arrival time at the observing frequency. Applied as the all-pass filter
    exp(-2 pi i Phi(f)) with dPhi/df equal to that delay, over one FFT of
    the whole signal (the caller pads it by the largest delay)."""
    import scipy.fft as sf
    n = len(a)
    f = sf.fftfreq(n, 1.0 / fs)
    k = 4148.808e12 * dm  # s Hz^2
    f0 = float(lo_freq)
    fobs = obs_freq_mhz * 1e6
    # Phi(f) = k (1/f0 - 1/(f0 + f) - f / fobs^2), written without cancellation
    phi = k * (f / (f0 * (f0 + f)) - f / (fobs * fobs))
    A = sf.fft(a, workers=-1)
    A *= np.exp(-2j * np.pi * (phi - np.floor(phi)))
    return sf.ifft(A, workers=-1, overwrite_x=True)


def synth_station_streams(seed: int, n_streams_per_signal: List[int], fs: int, n_samples: int,
                          delays: List[DelayModel], raw_start_ticks: int, rho: float, bits: int,
                          lo_freq: float, extra_sources=None, active=None, pulsar=None) -> List[np.ndarray]:
    """Raw packed byte streams, one per entry of `delays` (upper sideband,
    LO at lo_freq Hz).

    n_streams_per_signal groups consecutive streams sharing one common signal
    (e.g. one group per polarisation). extra_sources: list of (rho2, [DelayModel
    per stream]) adding further common components (one per group).
    active: per stream (first, count), the raw samples the slices will read;
    only those are synthesised (the rest of the stream is zero bytes).
    pulsar: (Pulsar, peak, duty) adds a pulsed component per group seen
    through each stream's own delay model: noise whose power follows
    pulsar_power at the source, dispersed with the pulsar's DM relative to its
    observing frequency; it adds to the power of the continuum and noise.
    """
    rng = np.random.default_rng(seed)
    tps = int(round(CLOCK_RATE / fs))
    pad = 1 << 12
    block = 4096
    tt_blk = raw_start_ticks + np.arange(0, n_samples + block, block) * tps

    def per_stream(m: DelayModel):
        # The model gives the delay tau(t) at geocentric time t: the station
        # records at t + tau(t) what the geocentre receives at t. The sample
        # at station time t' therefore holds the signal of t' - tau(t_g) with
        # t_g + tau(t_g) = t' (three fixed-point steps; the remainder is rate^3
        # times the delay).
        t_g = tt_blk
        for _ in range(3):
            t_g = tt_blk - np.round(m.eval(t_g) * CLOCK_RATE).astype(np.int64)
        tau_blk = m.eval(t_g)
        d_blk = tau_blk * fs
        # nu_LO * tau per sample, linear interpolation of tau inside a block
        k = np.arange(n_samples)
        tau = np.interp(k, np.arange(len(tau_blk)) * block, tau_blk)
        fr = lo_freq * tau
        return d_blk, fr - np.floor(fr)

    if active is None:
        active = [(0, n_samples)] * len(delays)
    spb = 8 // bits
    act = []
    for a0, na in active:
        a0 = max(0, (int(a0) // block) * block)
        a1 = min(n_samples, -(-(int(a0) + int(na)) // block) * block + block)
        act.append((a0, a1))
    per = [per_stream(m) for m in delays]
    extra_per = [[per_stream(m) for m in models] for _, models in (extra_sources or [])]
    # the common-signal indices j = k - d(k) (plus the FIR's reach) that the
    # active samples k of every stream and source need
    jmin, jmax = None, None
    for ds in [per] + extra_per:
        for (d, _), (a0, a1) in zip(ds, act):
            dd = d[a0 // block:a1 // block + 1]
            lo = a0 - int(math.ceil(dd.max())) - pad
            hi = a1 - int(math.floor(dd.min())) + pad
            jmin = lo if jmin is None else min(jmin, lo)
            jmax = hi if jmax is None else max(jmax, hi)
    psr_pad = 0
    if pulsar is not None:
        psr, peak, duty = pulsar
        fhi = (lo_freq + fs / 2) / 1e6
        flo = lo_freq / 1e6
        worst = 4148.808 * psr.DM * max(abs(flo ** -2 - psr.obs_freq ** -2), abs(fhi ** -2 - psr.obs_freq ** -2))
        psr_pad = int(math.ceil(worst * fs)) + pad
    out = []
    idx = 0
    for group in n_streams_per_signal:
        cstart = jmin
        clen = jmax - jmin
        common = analytic(rng.standard_normal(clen))
        extras = [(rho2, analytic(rng.standard_normal(clen)), k)
                  for k, (rho2, _) in enumerate(extra_sources or [])]
        psig = None
        if pulsar is not None:
            n_p = clen + 2 * psr_pad
            src_ticks = raw_start_ticks + (cstart - psr_pad + np.arange(n_p, dtype=np.int64)) * tps
            psig = analytic(rng.standard_normal(n_p)) * np.sqrt(pulsar_power(src_ticks, psr, peak, duty))
            psig = disperse(psig, fs, lo_freq, psr.DM, psr.obs_freq)[psr_pad:psr_pad + clen]
        for _ in range(group):
            d, fr = per[idx]
            a0, a1 = act[idx]
            n_act = a1 - a0
            x = np.sqrt(rho) * delayed_signal(common, cstart - a0, n_act, d[a0 // block:], fr[a0:a1])
            if psig is not None:
                x += delayed_signal(psig, cstart - a0, n_act, d[a0 // block:], fr[a0:a1])
            rest = 1.0 - rho
            for rho2, sig2, k in extras:
                d2, fr2 = extra_per[k][idx]
                x += np.sqrt(rho2) * delayed_signal(sig2, cstart - a0, n_act, d2[a0 // block:], fr2[a0:a1])
                rest -= rho2
            x += np.sqrt(rest) * rng.standard_normal(n_act)
            raw = np.zeros(n_samples // spb, dtype=np.uint8)
            raw[a0 // spb:a1 // spb] = pack(quantise(x, bits), bits)
            out.append(raw)
            idx += 1
    return out
