"""Pins the oracle's lag re-channelisation (integration_write when
fft_size_correlation > number_channels, correlation_core.cc:417-440, with
create_window / create_mask, correlation_core.cc:668-824) against numpy's FFT,
numpy.convolve and closed forms. CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from sfxg.job import WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN, WINDOW_NONE, WINDOW_RECT, MaskParameters

TOL = {"f64": 1e-11, "f32": 3e-6}
# The re-channelisation takes complex<float> rows (integration_normalize's
# output), so the input's own float rounding bounds the FP64 form's agreement
# with a closed form computed from unrounded values.
TOLIN = {"f64": 5e-7, "f32": 5e-6}


def hermitian_half(rng, n):
    """n/2 + 1 bins of the spectrum of a random real series of n points."""
    return np.fft.rfft(rng.standard_normal(n))


# ---------------------------------------------------------------------------
# c2r transform -- FFTW convention of sfxc_fft_float.cc:235-242
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 8, 64])
def test_irfft_bruteforce(n):
    rng = np.random.default_rng(n)
    X = hermitian_half(rng, n) + 0.3 * rng.standard_normal(n // 2 + 1) * 1j
    X[0] = X[0].real
    X[-1] = X[-1].real
    full = np.concatenate([X, np.conj(X[-2:0:-1])])
    k = np.arange(n)
    ref = (np.exp(2j * np.pi * np.outer(k, k) / n) @ full).real
    np.testing.assert_allclose(O.irfft(X, n, "f64"), ref, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("n", [512, 32768])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_irfft_numpy(n, prec):
    rng = np.random.default_rng(7 + n)
    X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    clean = X.copy()
    clean[0] = clean[0].real
    clean[-1] = clean[-1].real
    ref = np.fft.irfft(clean, n) * n
    got = O.irfft(X, n, prec)  # imaginary parts of bins 0 and n/2 are not used
    assert np.max(np.abs(got - ref)) / np.sqrt(np.mean(ref ** 2)) < TOL[prec]
    np.testing.assert_array_equal(got, O.irfft(clean, n, prec))


# ---------------------------------------------------------------------------
# create_window -- correlation_core.cc:668-710,761-809
# ---------------------------------------------------------------------------
def window_fn(kind, n, i):
    i = np.asarray(i, dtype=np.float64)
    if kind == WINDOW_RECT:
        return ((i >= n // 4) & (i < (3 * n) // 4)).astype(np.float64)
    if kind == WINDOW_COS:
        return np.sin(np.pi * i / n)
    if kind == WINDOW_HAMMING:
        return 0.54 - 0.46 * np.cos(2 * np.pi * i / n)
    return 0.5 * (1 - np.cos(2 * np.pi * i / n))


def lag_window_numpy(fn_kind, nchan, fft):
    """(m / n) conv_n(n/2 + i) / conv_m(m/2 + i), the self-convolutions taken
    with numpy.convolve over the n + 1 (m + 1) samples of the window."""
    n, m = 2 * nchan, 2 * fft
    fn = window_fn(fn_kind, n, np.arange(n + 1))
    fm = window_fn(fn_kind, m, np.arange(m + 1))
    cn, cm = np.convolve(fn, fn), np.convolve(fm, fm)
    i = np.arange(n // 2 + 1)
    half = (m // n) * cn[n + i] / cm[m + i]
    w = np.empty(n)
    w[: n // 2 + 1] = half
    w[n // 2 + 1:] = half[1: n // 2][::-1]
    return w


@pytest.mark.parametrize("nchan,fft", [(16, 64), (128, 1024), (256, 512)])
def test_lag_window_rect_is_box_autocorrelation(nchan, fft):
    # The reference's closed form for NONE / RECT is (m / n) times the ratio
    # of the autocorrelations of boxes of n / 2 and m / 2 ones (triangles).
    n, m = 2 * nchan, 2 * fft
    tn = np.correlate(np.ones(n // 2), np.ones(n // 2), "full")[n // 2 - 1:]
    tm = np.correlate(np.ones(m // 2), np.ones(m // 2), "full")[m // 2 - 1:]
    tn = np.concatenate([tn, [0.0]])
    i = np.arange(n // 2 + 1)
    half = (m // n) * tn[i] / tm[i]
    ref = np.concatenate([half, half[1:nchan][::-1]])
    for kind in (WINDOW_NONE, WINDOW_RECT):
        got = O.lag_window(kind, nchan, fft)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)
        assert got[0] == 1.0 and got[nchan] == 0.0


@pytest.mark.parametrize("kind", [WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN])
@pytest.mark.parametrize("nchan,fft", [(16, 64), (64, 1024), (512, 1024)])
def test_lag_window_convolution(kind, nchan, fft):
    ref = lag_window_numpy(kind, nchan, fft)
    got = O.lag_window(kind, nchan, fft)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-7)
    np.testing.assert_array_equal(got[1:nchan], got[2 * nchan - 1:nchan:-1])


# ---------------------------------------------------------------------------
# integration_write re-channelisation -- correlation_core.cc:417-440
# ---------------------------------------------------------------------------
class _Job:
    def __init__(self, nchan, fft, window=WINDOW_HANN, mask=None):
        self.number_channels = nchan
        self.fft_size_correlation = fft
        self.window = window
        self.mask = mask


def spectrum_of_lags(lags, fft):
    """Bins 0..fft of the 2 fft-point DFT of a real lag series given as
    {lag: value}, lags taken modulo 2 fft."""
    r = np.zeros(2 * fft)
    for l, v in lags.items():
        r[l % (2 * fft)] += v
    return np.fft.rfft(r)


def run(job, X, prec):
    a = np.stack([X.real, X.imag], axis=-1)[None].astype(np.float32)
    out = O.rechannelize(job, a, prec)[0]
    return out[:, 0].astype(np.float64) + 1j * out[:, 1]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nchan,fft", [(16, 64), (256, 1024), (1024, 4096)])
def test_band_limited_lags_sample_the_spectrum(prec, nchan, fft):
    # Lags inside (-nchan, nchan) pass the fold untouched: with unit lag
    # weights the output is the input spectrum at every (fft / nchan)-th bin.
    rng = np.random.default_rng(nchan + fft)
    lags = {l: rng.standard_normal() for l in range(-nchan + 1, nchan)}
    X = spectrum_of_lags(lags, fft)
    job = _Job(nchan, fft, mask=MaskParameters(window=[1.0] * (2 * nchan)))
    got = run(job, X, prec)
    ref = X[:: fft // nchan]
    assert got.shape == (nchan + 1,)
    assert np.max(np.abs(got - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < TOLIN[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_edge_lags_are_averaged(prec):
    nchan, fft, a, b = 32, 256, 0.75, -0.25
    X = spectrum_of_lags({nchan: a, -nchan: b}, fft)
    job = _Job(nchan, fft, mask=MaskParameters(window=[1.0] * (2 * nchan)))
    got = run(job, X, prec)
    ref = (a + b) / 2 * (-1.0) ** np.arange(nchan + 1)
    np.testing.assert_allclose(got, ref, rtol=0, atol=TOLIN[prec])


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("d", [0, 5, -7, 63])
@pytest.mark.parametrize("kind", [WINDOW_HANN, WINDOW_RECT])
def test_normalised_pure_delay(prec, d, kind):
    # Normalising a spectrum c_j exp(-2 pi i d j / 2fft) with c_j > 0 leaves a
    # unit lag at d: the output is the lag window at d times the delay's phase
    # slope over 2 nchan points.
    nchan, fft = 64, 512
    rng = np.random.default_rng(abs(d) + kind)
    j = np.arange(fft + 1)
    X = rng.uniform(0.1, 5.0, fft + 1) * np.exp(-2j * np.pi * d * j / (2 * fft))
    X[0] = X[0].real
    X[-1] = X[-1].real
    job = _Job(nchan, fft, window=kind, mask=MaskParameters(normalize=True))
    got = run(job, X, prec)
    w = O.lag_window(kind, nchan, fft).astype(np.float64)
    k = np.arange(nchan + 1)
    ref = w[d % (2 * nchan)] * np.exp(-2j * np.pi * d * k / (2 * nchan))
    np.testing.assert_allclose(got, ref, rtol=0, atol=TOLIN[prec])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mask_after_normalisation(prec):
    # The mask multiplies the normalised bins (correlation_core.cc:418-424):
    # a constant mask of 2 doubles the output even with normalisation on.
    nchan, fft = 64, 512
    rng = np.random.default_rng(3)
    X = hermitian_half(rng, 2 * fft)
    one = run(_Job(nchan, fft, mask=MaskParameters(normalize=True)), X, prec)
    two = run(_Job(nchan, fft, mask=MaskParameters(normalize=True, mask=[2.0] * (fft + 1))), X, prec)
    np.testing.assert_allclose(two, 2 * one, rtol=0, atol=TOLIN[prec] * np.max(np.abs(one)))
    # and a varying mask acts on the bins, before the lag transform
    mask = rng.uniform(0, 1, fft + 1)
    got = run(_Job(nchan, fft, mask=MaskParameters(mask=list(mask))), X, prec)
    ref = run(_Job(nchan, fft), (X * mask.astype(np.float32)).astype(np.complex64), prec)
    np.testing.assert_allclose(got, ref, rtol=0, atol=TOLIN[prec] * np.max(np.abs(ref)))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dc_nyquist_imaginary_parts_unused(prec):
    nchan, fft = 32, 256
    rng = np.random.default_rng(11)
    X = hermitian_half(rng, 2 * fft)
    Y = X.copy()
    Y[0] += 3j
    Y[-1] -= 2j
    job = _Job(nchan, fft)
    np.testing.assert_array_equal(run(job, X, prec), run(job, Y, prec))


def test_zero_row_with_normalisation():
    job = _Job(32, 128, mask=MaskParameters(normalize=True))
    got = run(job, np.zeros(129, dtype=np.complex128), "f32")
    assert np.all(got == 0)


def test_table_sizes_checked():
    with pytest.raises(ValueError):
        O.lag_tables(_Job(32, 128, mask=MaskParameters(mask=[1.0] * 128)))
    with pytest.raises(ValueError):
        O.lag_tables(_Job(32, 128, mask=MaskParameters(window=[1.0] * 32)))


# ---------------------------------------------------------------------------
# phased-array core -- correlation_core_phased.cc:13-50,101-126
# ---------------------------------------------------------------------------
class _PhasedJob:
    def __init__(self, M, fft, nffts, nsub):
        self.station_streams = list(range(M))
        self.fft_size_correlation = fft
        self._n, self._s = nffts, nsub

    def baselines(self):
        return [(i, i) for i in range(len(self.station_streams))]

    def number_ffts_in_integration(self):
        return self._n

    def number_ffts_in_sub_integration(self):
        return self._s


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nffts,nsub", [(50, 7), (48, 12), (30, 100)])
def test_phased_sums(prec, nffts, nsub):
    M, fft = 3, 64
    job = _PhasedJob(M, fft, nffts, nsub)
    rng = np.random.default_rng(nffts + nsub)
    spectra = [rng.standard_normal((nffts, fft + 4, 2)) for _ in range(M)]
    rows = O.phased_correlate(job, spectra, nffts, prec)
    nrows = nffts // nsub + 1
    assert rows.shape == (nrows, fft + 1, 2)
    x = sum(s.astype(np.float32 if prec == "f32" else np.float64).astype(np.float64) for s in spectra)
    ref = np.zeros((nrows, fft + 1, 2))
    for seg in range(nffts):
        ref[seg // nsub] += x[seg, :fft + 1]
    tol = 1e-12 if prec == "f64" else 2e-6
    assert np.max(np.abs(rows - ref)) <= tol * np.sqrt(np.mean(ref ** 2)) * 10
    if nffts % nsub == 0:
        assert np.all(rows[-1] == 0)
