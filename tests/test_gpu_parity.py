"""GPU path vs the oracle on the same seeded HEADER streams (through the C ABI)."""
import numpy as np
import pytest

from oracle import oracle as O
from sfxg import capi, configs
from sfxg.job import LUT_CONFIG, LUT_REFERENCE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c0_short():
    return configs.config0(seconds=0.25, t_int_usec=125000)


def vis_from_bytes(job, out: bytes, nbins=1):
    """Split integration_write bytes into headers and visibility arrays."""
    M = len(job.station_streams)
    nst = len({s.station_number for s in job.station_streams})
    B = len(job.baselines())
    nch = job.number_channels + 1
    one = 4 + 16 + 32 * nst + 24 * M + B * (8 + 8 * nch)
    assert len(out) == one * nbins
    heads, vis = [], []
    for b in range(nbins):
        rec = out[b * one:(b + 1) * one]
        h = 4 + 16 + 32 * nst + 24 * M
        parts = [rec[:h]]
        v = np.zeros((B, nch), dtype=np.complex64)
        o = h
        for k in range(B):
            parts.append(rec[o:o + 8])
            v[k] = np.frombuffer(rec[o + 8:o + 8 + 8 * nch], dtype=np.complex64)
            o += 8 + 8 * nch
        heads.append(b"".join(parts))
        vis.append(v)
    return heads, vis


def compare_vis(v_gpu, v_ref, tol_norm=1e-5, tol_chan=1e-5, label=""):
    """Per product: |dV| relative to the product's rms amplitude, and per
    channel relative amplitude and phase errors over channels whose amplitude
    is at least 0.1 of the rms (weaker channels sit at the accumulation noise
    floor, where a relative error has no meaning)."""
    worst = (0.0, 0.0, 0.0)
    for b in range(v_ref.shape[0]):
        r = v_ref[b].astype(np.complex128)
        g = v_gpu[b].astype(np.complex128)
        rms = np.sqrt(np.mean(np.abs(r) ** 2))
        e_norm = np.max(np.abs(g - r)) / rms
        m = np.abs(r) >= 0.1 * rms
        e_amp = np.max(np.abs(np.abs(g[m]) - np.abs(r[m])) / np.abs(r[m]))
        e_ph = np.max(np.abs(np.angle(g[m] * np.conj(r[m]))))
        worst = tuple(max(a, c) for a, c in zip(worst, (e_norm, e_amp, e_ph)))
    print(f"{label} normwise {worst[0]:.3e} amp {worst[1]:.3e} phase {worst[2]:.3e} rad")
    assert worst[0] <= tol_norm, worst
    assert worst[1] <= tol_chan, worst
    assert worst[2] <= tol_chan, worst
    return worst


@pytest.mark.parametrize("lut", [LUT_REFERENCE, LUT_CONFIG])
def test_c0_slice_stages(ctx, c0_short, lut):
    job, hdrs = c0_short[0]
    ref = O.run_slice(job, hdrs, lut=lut, keep_intermediate=True)
    out = capi.correlate_slice(ctx, job, hdrs, lut=lut)
    M = len(job.station_streams)
    for i in range(M):
        s_gpu = ctx.stage(capi.STAGE_SAMPLES, i)
        np.testing.assert_array_equal(s_gpu, ref["b2f"][i]["samples"])
        lv, ts = ctx.statistics(i)
        np.testing.assert_array_equal(lv, ref["stats"][i][0])
        np.testing.assert_array_equal(ts, ref["stats"][i][1])
        t_gpu = ctx.stage(capi.STAGE_TIME_SERIES, i)
        t_ref = ref["dc"][i]["time_series"]
        err_t = np.max(np.abs(t_gpu - t_ref)) / np.sqrt(np.mean(t_ref.astype(np.float64) ** 2))
        print("time series rel err", err_t)
        assert err_t < 2e-5
        sp = ctx.stage(capi.STAGE_SPECTRA, i)
        sp_ref = ref["dc"][i]["spectra"][:, :job.fft_size_correlation + 1, :]
        sp_ref = sp_ref[..., 0] + 1j * sp_ref[..., 1]
        err_s = np.max(np.abs(sp - sp_ref)) / np.sqrt(np.mean(np.abs(sp_ref) ** 2))
        print("spectra rel err", err_s)
        assert err_s < 2e-5
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    compare_vis(vis_g[0], vis_r[0], label=f"c0 lut={lut}")


def test_c0_all_slices(ctx, c0_short):
    for job, hdrs in c0_short:
        ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
        heads_g, vis_g = vis_from_bytes(job, out)
        heads_r, vis_r = vis_from_bytes(job, ref["output"])
        assert heads_g == heads_r
        compare_vis(vis_g[0], vis_r[0], label=f"slice {job.integration_nr}")
