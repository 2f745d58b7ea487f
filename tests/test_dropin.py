"""The drop-in classes' host side on the CPU (sfxc-1_amd/host/sfxc_types.h):
the Akima delay table against scipy's independent Akima interpolator and
the closed forms, and the C ABI model it hands the library."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ak():
    d = tempfile.mkdtemp(prefix="sfxg_ak_")
    so = os.path.join(d, "akima_check.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", os.path.join(ROOT, "tests/cpp/akima_check.cpp"),
                    "-o", so], check=True)
    return C.CDLL(so)


def P(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def evaluate(ak, x, y, xs):
    x, y, xs = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, xs))
    val = np.zeros(len(xs))
    der = np.zeros(len(xs))
    assert ak.ak_eval(P(x, C.c_double), P(y, C.c_double), len(x), P(xs, C.c_double), len(xs),
                      P(val, C.c_double), P(der, C.c_double)) == 0
    return val, der


@pytest.mark.parametrize("seed", range(5))
def test_akima_matches_scipy(ak, seed):
    """Akima's method as GSL 1.9 implements it (the reference's
    gsl_interp_akima) against scipy's Akima1DInterpolator, an independent
    implementation of the same published method: equal slopes at the knots
    and values between them, on smooth random data where no segment is
    straight."""
    from scipy.interpolate import Akima1DInterpolator
    rng = np.random.default_rng(seed)
    x = np.arange(12, dtype=np.float64)
    y = 1e-3 * np.cumsum(rng.normal(size=12)) + 1e-6 * x ** 2
    xs = np.linspace(0, 11, 500)
    val, der = evaluate(ak, x, y, xs)
    ref = Akima1DInterpolator(x, y)
    np.testing.assert_allclose(val, ref(xs), rtol=0, atol=1e-15)
    np.testing.assert_allclose(der, ref(xs, 1), rtol=0, atol=1e-14)


def test_akima_reproduces_lines_and_straight_segments(ak):
    """Linear data are reproduced exactly (every segment straight); between
    two straight stretches the spline bends only on the segments where the
    slope changes, and is linear wherever both neighbouring slope changes
    vanish."""
    x = np.arange(10, dtype=np.float64)
    y = 4.4e-3 + 2.857e-6 * x
    xs = np.linspace(0, 9, 91)
    val, der = evaluate(ak, x, y, xs)
    np.testing.assert_allclose(val, 4.4e-3 + 2.857e-6 * xs, rtol=1e-15, atol=0)
    np.testing.assert_allclose(der, 2.857e-6, rtol=1e-9)
    y2 = np.where(x < 5, x, 5 + 2 * (x - 5)).astype(np.float64)  # slope 1, then 2
    val, der = evaluate(ak, x, y2, xs)
    lin = xs <= 3
    np.testing.assert_allclose(val[lin], xs[lin], atol=1e-15)
    np.testing.assert_allclose(val[xs >= 7], 5 + 2 * (xs[xs >= 7] - 5), atol=1e-14)


def test_table_clock_and_model(ak):
    """Delay_table_akima::delay adds the station clock (offset plus rate
    since the clock epoch) to the spline; the C ABI model carries the same
    delay on the sample grid, clock folded in, so that evaluating its pieces
    gives delay() at any time."""
    CR = 512_000_000
    scan = 1_000 * CR
    t = np.arange(8, dtype=np.float64)
    y = 3e-3 + 1e-6 * t + 2e-9 * np.sin(t)
    epoch, off, rate = scan - 5 * CR, 2.5e-6, 1e-12
    ts = scan + (np.linspace(0.1, 6.9, 37) * CR).astype(np.int64)
    delay = np.zeros(len(ts))
    rt = np.zeros(len(ts))
    mb, mi, mn = C.c_int64(), C.c_double(), C.c_int32()
    coef = np.zeros(4 * 7)
    assert ak.ak_table(C.c_int64(scan), P(t, C.c_double), P(y, C.c_double), 8, C.c_int64(epoch), C.c_double(off),
                       C.c_double(rate), P(ts, C.c_int64), len(ts), P(delay, C.c_double), P(rt, C.c_double),
                       C.byref(mb), C.byref(mi), C.byref(mn), P(coef, C.c_double)) == 0
    assert mb.value == scan and mi.value == 1.0 and mn.value == 7
    sec = (ts - scan) / CR
    from scipy.interpolate import Akima1DInterpolator
    spline = Akima1DInterpolator(t, y)
    np.testing.assert_allclose(delay, spline(sec) + off + (ts - epoch) / CR * rate, rtol=0, atol=1e-17)
    np.testing.assert_allclose(rt, spline(sec, 1) + rate, rtol=0, atol=1e-16)
    c = coef.reshape(7, 4)
    i = np.floor(sec).astype(int)
    u = sec - i
    piece = c[i, 0] + u * (c[i, 1] + u * (c[i, 2] + u * c[i, 3]))
    np.testing.assert_allclose(piece, delay, rtol=0, atol=1e-17)
