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


def compare_vis(v_gpu, v_ref, tol_norm=1e-5, tol_chan=1e-5, tol_max=3e-5, label=""):
    """The plain FP32 bars, per product (row):
    * normwise: ||V_gpu - V_ref|| / ||V_ref|| over the channels <= 1e-5;
    * channels with |V_ref| >= 0.5 rms |V_ref|: relative amplitude error and
      phase error (rad) <= 1e-5;
    * every channel: |V_gpu - V_ref| / rms |V_ref| <= 3e-5.
    Two FP32 correlators whose FFTs round differently decorrelate the
    rounding of their sequential sums: over a sub-integration of n segments
    a cross product differs by about sqrt(n) 2^-24 of its rms per channel
    (the FP32 oracle differs from the FP64 one by that much), so the largest
    of many thousands of channels reaches 1e-5 at n ~ 2000 and 1.5e-5 at n ~
    8000. The last bar bounds that tail on every channel; the first two are
    the 1e-5 bars. Below half the rms a relative amplitude or phase is set
    by that absolute rounding (the Rayleigh tail of noise-dominated
    visibilities) and is not checked. The fraction of channels under the
    relative check is printed."""
    checked = []
    worst = (0.0, 0.0, 0.0, 0.0)
    for b in range(v_ref.shape[0]):
        r = v_ref[b].astype(np.complex128)
        g = v_gpu[b].astype(np.complex128)
        rms = np.sqrt(np.mean(np.abs(r) ** 2))
        if rms == 0.0:  # a product no segment reached (an empty pulsar bin)
            assert np.all(g == 0.0), f"{label}: product {b} is zero in the reference only"
            continue
        e_l2 = np.sqrt(np.mean(np.abs(g - r) ** 2)) / rms
        e_max = np.max(np.abs(g - r)) / rms
        m = np.abs(r) >= 0.5 * rms
        checked.append(np.mean(m))
        e_amp = np.max(np.abs(np.abs(g[m]) - np.abs(r[m])) / np.abs(r[m]))
        e_ph = np.max(np.abs(np.angle(g[m] * np.conj(r[m]))))
        worst = tuple(max(a, c) for a, c in zip(worst, (e_l2, e_amp, e_ph, e_max)))
    print(f"{label} normwise {worst[0]:.3e} amp {worst[1]:.3e} phase {worst[2]:.3e} rad, worst channel "
          f"{worst[3]:.3e} (relative check on {100 * np.mean(checked):.1f} % of channels)")
    assert worst[0] <= tol_norm, worst
    assert worst[1] <= tol_chan, worst
    assert worst[2] <= tol_chan, worst
    assert worst[3] <= tol_max, worst
    return worst


@pytest.mark.parametrize("lut", [LUT_REFERENCE, LUT_CONFIG])
def test_c0_slice_stages(ctx, c0_short, lut):
    job, hdrs = c0_short[0]
    ref = O.run_slice(job, hdrs, lut=lut, keep_intermediate=True)
    out = capi.correlate_slice(ctx, job, hdrs, lut=lut, keep_time_series=True)
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


def test_fused_matches_unfused(ctx, c0_short):
    """The fused F-engine (series in LDS, picked by the tuning table's
    fe_fused) and the two-kernel path agree to FP32 rounding: they share the
    channeliser's FFT schedule, and the unfused delay correction factors its
    256-point transforms as 16 x 16 where the fused one uses four radix-4
    stages. Headers are identical. Without the table's pick the series is
    not kept and the two kernels run: the same bytes as keep_time_series."""
    job, hdrs = c0_short[0]
    nc = job.fft_size_correlation
    M = len(job.station_streams)
    a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, keep_time_series=True)
    sp_a = ctx.stage(capi.STAGE_SPECTRA, 1)
    capi.tune_clear()
    try:
        d = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, keep_time_series=False)
        capi.tune_import(capi.tune_key(nc, M, 0, False, True) + "|0|0|0|0|0|0|0|1\n")
        b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, keep_time_series=False)
        sp_b = ctx.stage(capi.STAGE_SPECTRA, 1)
        with pytest.raises(capi.SfxgError):
            ctx.stage(capi.STAGE_TIME_SERIES, 1)  # the fused F-engine kept the series in LDS
    finally:
        capi.tune_clear()
    assert d == a
    err = np.max(np.abs(sp_a - sp_b)) / np.sqrt(np.mean(np.abs(sp_b.astype(np.complex128)) ** 2))
    print("fused vs unfused spectra", err)
    assert err < 1e-5  # each path is within 2e-5 of the oracle (test_c0_slice_stages); they differ by ~2e-6
    heads_a, vis_a = vis_from_bytes(job, a)
    heads_b, vis_b = vis_from_bytes(job, b)
    assert heads_a == heads_b
    compare_vis(vis_a[0], vis_b[0], label="fused vs unfused")


def test_c0_all_slices(ctx, c0_short):
    for job, hdrs in c0_short:
        ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
        heads_g, vis_g = vis_from_bytes(job, out)
        heads_r, vis_r = vis_from_bytes(job, ref["output"])
        assert heads_g == heads_r
        compare_vis(vis_g[0], vis_r[0], label=f"slice {job.integration_nr}")


def run_both(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=None):
    ref = O.run_slice(job, hdrs, lut=lut, pulsar=pulsar)
    out = capi.correlate_slice(ctx, job, hdrs, lut=lut, pulsar=pulsar)
    nb = pulsar.nbins + 1 if pulsar is not None else 1
    heads_g, vis_g = vis_from_bytes(job, out, nb)
    heads_r, vis_r = vis_from_bytes(job, ref["output"], nb)
    assert heads_g == heads_r
    for i in range(len(job.station_streams)):
        lv, ts = ctx.statistics(i)
        np.testing.assert_array_equal(lv, ref["stats"][i][0])
        np.testing.assert_array_equal(ts, ref["stats"][i][1])
    return vis_g, vis_r, ref


@pytest.mark.parametrize("nchan", [256, 2048, 8192, 16384, 32768])
def test_channel_counts(ctx, nchan):
    """Power-of-four and odd-log2 FFT sizes (the radix-2 stage), up to the
    largest LDS-resident transform."""
    job, hdrs = configs.build_slices(seed=5, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 2.5e-6, -1.75e-6], rates=[0.0, 2e-7, -1e-7])[0]
    vis_g, vis_r, _ = run_both(ctx, job, hdrs)
    compare_vis(vis_g[0], vis_r[0], label=f"nchan={nchan}")


@pytest.mark.parametrize("window,sideband", [(0, "U"), (4, "U"), (4, "L"), (3, "L")])
def test_large_fft_runs(ctx, window, sideband):
    """The large-N channeliser walks runs of consecutive segments and carries
    the shared half of each segment in registers: 0.97 s at 32 Ms/s gives 1893
    segments, so runs of two segments per workgroup and a partial last run;
    zero padding (window NONE) and the sideband flip."""
    job, hdrs = configs.build_slices(seed=12, n_stations=2, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=16384,
                                     seconds=0.97, t_int_usec=970_000, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 1.5e-6], rates=[0.0, 1e-7], window=window, sideband=sideband)[0]
    vis_g, vis_r, _ = run_both(ctx, job, hdrs)
    compare_vis(vis_g[0], vis_r[0], label=f"nchan=16384 window={window} sideband={sideband}")


@pytest.mark.parametrize("window", [0, 1, 2, 3])
def test_windows(ctx, window):
    job, hdrs = configs.build_slices(seed=6, n_stations=2, pols=["R", "L"], fs=32_000_000, bw=16_000_000,
                                     nchan=1024, seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2,
                                     sky_freq=5_000_000_000, offsets=[0.0, 1e-6], rates=[0.0, 1e-7],
                                     window=window)[0]
    vis_g, vis_r, _ = run_both(ctx, job, hdrs)
    compare_vis(vis_g[0], vis_r[0], label=f"window={window}")


def test_one_bit(ctx):
    job, hdrs = configs.build_slices(seed=7, n_stations=3, pols=["R"], fs=16_000_000, bw=8_000_000, nchan=512,
                                     seconds=0.125, t_int_usec=125000, rho=0.05, bits=1, sky_freq=1_600_000_000,
                                     offsets=[0.0, -3e-6, 4e-6], rates=[0.0, 1e-6, -1e-6])[0]
    for lut in (LUT_REFERENCE, LUT_CONFIG):
        vis_g, vis_r, _ = run_both(ctx, job, hdrs, lut=lut)
        compare_vis(vis_g[0], vis_r[0], label=f"1-bit lut={lut}")


def test_invalid_data(ctx):
    """INVALID records: zero fill, n_invalid in the statistics, weights and
    the flag factor of the normalisation."""
    inv = {0: [(1000, 300), (50000, 2000)], 2: [(1100, 100), (90000, 5000)]}
    job, hdrs = configs.build_slices(seed=8, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=1024,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 2e-6, -1e-6], rates=[0.0, 3e-7, -2e-7], invalid=inv)[0]
    vis_g, vis_r, ref = run_both(ctx, job, hdrs)
    assert ref["stats"][0][0][4] > 0 and ref["stats"][2][0][4] > 0
    compare_vis(vis_g[0], vis_r[0], label="invalid")


def test_reference_station(ctx):
    job, hdrs = configs.build_slices(seed=9, n_stations=4, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=512,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 1e-6, 2e-6, 3e-6], rates=[0.0] * 4)[0]
    job.reference_station = 2
    vis_g, vis_r, _ = run_both(ctx, job, hdrs)
    assert vis_g[0].shape[0] == 4 + 3
    compare_vis(vis_g[0], vis_r[0], label="reference station")


def test_pulsar_bins(ctx):
    """Correlation_core_pulsar: per-channel gate/bin accumulation."""
    from sfxg.job import Pulsar, ticks_to_mjd
    job, hdrs = configs.build_slices(seed=10, n_stations=2, pols=["R", "L"], fs=32_000_000, bw=16_000_000,
                                     nchan=1024, seconds=0.125, t_int_usec=125000, rho=0.05, bits=2,
                                     sky_freq=1_400_000_000, offsets=[0.0, 1e-6], rates=[0.0, 1e-7])[0]
    psr = Pulsar(nbins=8, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    vis_g, vis_r, _ = run_both(ctx, job, hdrs, pulsar=psr)
    for b in range(psr.nbins + 1):
        compare_vis(vis_g[b], vis_r[b], label=f"pulsar bin {b}")


def test_pulsar_bin_limit(ctx):
    """More than 32766 on-pulse bins are refused at set-up, before any
    launch (the mixed-precision pulsar kernels hold bin indices in 16 bits);
    the context then correlates the next slice as usual."""
    from sfxg.job import Pulsar, ticks_to_mjd
    job, hdrs = configs.build_slices(seed=10, n_stations=2, pols=["R"], fs=32_000_000, bw=16_000_000,
                                     nchan=256, seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2,
                                     sky_freq=1_400_000_000, offsets=[0.0, 1e-6], rates=[0.0, 1e-7])[0]
    for prec in (0, 1):
        ctx.set_options(lut=LUT_CONFIG, precision=prec)
        psr = Pulsar(nbins=32767, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001,
                     DM=68.0, ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
        with pytest.raises(capi.SfxgError, match="32766"):
            ctx.set_parameters(job, pulsar=psr)
    ctx.set_options(lut=LUT_CONFIG, precision=0)
    vis_g, vis_r, _ = run_both(ctx, job, hdrs)
    compare_vis(vis_g[0], vis_r[0], label="after a refused pulsar")


@pytest.mark.parametrize("n_stations,nchan,chunk_mib", [(12, 512, 0), (12, 512, 1), (3, 256, 0)])
def test_pulsar_tiles_match_per_baseline(ctx, n_stations, nchan, chunk_mib):
    """The FP32 pulsar X-step by bin on register tiles of 16 channels
    (k_xstep_pulsar_bin; 513 and 257 channels, so a last block of one
    channel) gives the same bits as the per-baseline kernel (k_xstep_pulsar):
    both add each (bin, baseline, channel)'s products segment by segment,
    across chunks too (chunk_mib 1: one sub-integration per chunk). Both match
    the oracle within the FP32 bars."""
    from sfxg.job import Pulsar, ticks_to_mjd
    rng = np.random.default_rng(n_stations + nchan)
    job, hdrs = configs.build_slices(seed=45, n_stations=n_stations, pols=["R", "L"], fs=16_000_000, bw=8_000_000,
                                     nchan=nchan, seconds=0.0625, t_int_usec=62500, sub_int_usec=15625, rho=0.05,
                                     bits=2, sky_freq=1_400_000_000,
                                     offsets=list(rng.uniform(-3e-6, 3e-6, n_stations)),
                                     rates=list(rng.uniform(-2e-7, 2e-7, n_stations)))[0]
    psr = Pulsar(nbins=16, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    M = len(job.station_streams)
    print(f"case pulsar tiles vs per-baseline: {M} streams, {nchan} channels, chunk_mib {chunk_mib}", flush=True)
    outs = {}
    for xs in (2, 4):
        capi.tune_clear()
        capi.tune_import(capi.tune_key(nchan, M, 0, True, nchan <= 4096) + f"|0|0|{xs}|0|0\n")
        try:
            outs[xs] = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=psr, chunk_mib=chunk_mib)
        finally:
            capi.tune_clear()
            ctx.set_options(lut=LUT_CONFIG)
    assert outs[2] == outs[4]
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG, pulsar=psr)
    nb = psr.nbins + 1
    heads_g, vis_g = vis_from_bytes(job, outs[2], nb)
    heads_r, vis_r = vis_from_bytes(job, ref["output"], nb)
    assert heads_g == heads_r
    for b in range(nb):
        compare_vis(vis_g[b], vis_r[b], label=f"pulsar tiles bin {b}")


def test_elided_push_matches(ctx, c0_short):
    """Device-resident payload (sfxg_push_stream_elided) gives the same bytes
    as the host HEADER stream."""
    import ctypes as C
    from sfxg import synth
    job, hdrs = c0_short[0]
    out_host = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    L = C.CDLL("libamdhip64.so")
    ctx.set_options(lut=LUT_CONFIG)
    ctx.set_parameters(job)
    bufs = []
    for i, h in enumerate(hdrs):
        ctrl = bytearray()
        payload = bytearray()
        k = 0
        while k < len(h):
            t = h[k]
            if t == 0:
                n = h[k + 1] | (h[k + 2] << 8)
                ctrl += h[k:k + 3]
                payload += h[k + 3:k + 3 + n]
                k += 3 + n
            elif t == 2:
                ctrl += h[k:k + 2]
                k += 2
            elif t == 1:
                ctrl += h[k:k + 3]
                k += 3
            else:
                ctrl.append(t)
                k += 1
        p = C.c_void_p()
        assert L.hipMalloc(C.byref(p), C.c_size_t(len(payload))) == 0
        src = (C.c_uint8 * len(payload)).from_buffer_copy(bytes(payload))
        assert L.hipMemcpy(p, src, C.c_size_t(len(payload)), 1) == 0
        bufs.append(p)
        ctx.push_stream_elided(i, bytes(ctrl), p.value, len(payload))
    ctx.correlate()
    out_dev = ctx.output()
    for p in bufs:
        L.hipFree(p)
    assert out_dev == out_host
    del synth


def test_accuracy_vs_fp64(ctx):
    """Against the FP64 oracle of the same pipeline, the GPU FP32 path is as
    accurate as the FP32 restatement of the reference (the reference's own
    arithmetic): both errors are FP32 rounding, neither is a method error."""
    job, hdrs = configs.build_slices(seed=11, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=1024,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 1e-6, -2e-6], rates=[0.0, 1e-7, 2e-7])[0]
    out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    _, vis_g = vis_from_bytes(job, out)
    r32 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f32")
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    _, v32 = vis_from_bytes(job, r32["output"])
    _, v64 = vis_from_bytes(job, r64["output"])

    def err(a, b):
        return max(np.max(np.abs(a[k] - b[k])) / np.sqrt(np.mean(np.abs(b[k]) ** 2)) for k in range(b.shape[0]))

    e_gpu = err(vis_g[0].astype(np.complex128), v64[0].astype(np.complex128))
    e_ref = err(v32[0].astype(np.complex128), v64[0].astype(np.complex128))
    print(f"vs FP64 oracle: gpu {e_gpu:.3e}, fp32 oracle {e_ref:.3e}")
    assert e_gpu < 1.5 * e_ref + 1e-7


@pytest.mark.slow
def test_config0_full(ctx):
    """Configuration 0 as specified: 2 stations, 16 MHz at 32 Ms/s, 1024
    channels, 2 s in two 1 s integrations, decode table {-3.3359,-1,1,3.3359}."""
    slices = configs.config0()
    for job, hdrs in slices:
        vis_g, vis_r, _ = run_both(ctx, job, hdrs)
        compare_vis(vis_g[0], vis_r[0], label=f"config0 integration {job.integration_nr}")


def test_cpp_host_adapter(ctx, c0_short, tmp_path):
    """A C++ host (host/correlate_slice.cpp) wiring the drop-in classes with
    the reference's signatures (host/sfxc_dropin.h: Bit2float_worker,
    Delay_correction, Correlation_core) the way Correlator_node does, its
    delay tables built as Akima splines over 1 s samples, produces the
    oracle's headers and visibilities (and the statistics it prints)."""
    import os
    import subprocess
    from sfxg import slicefile
    exe = os.path.join(os.path.dirname(capi.LIB_PATH), "correlate_slice")
    assert os.path.exists(exe), "build first: make -C sfxc-1_amd"
    job, hdrs = c0_short[0]
    sl = tmp_path / "slice.bin"
    cor = tmp_path / "out.cor"
    slicefile.write_slice(str(sl), job, hdrs)
    print("case C++ drop-in host", flush=True)
    r = subprocess.run([exe, str(sl), str(cor)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out_cpp = cor.read_bytes()
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    heads_g, vis_g = vis_from_bytes(job, out_cpp)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    compare_vis(vis_g[0], vis_r[0], label="C++ drop-in host")
    levels = [int(v) for v in r.stdout.split(":")[1].split()]
    assert levels == list(ref["stats"][0][0])
    # the library through the Python binding, on the same slice
    out_py = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    heads_p, vis_p = vis_from_bytes(job, out_py)
    assert heads_p == heads_g
    compare_vis(vis_g[0], vis_p[0], label="C++ drop-in host vs Python binding")
    # split FP16 through the same host, against the FP64 oracle
    r = subprocess.run([exe, str(sl), str(cor), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, vis_s = vis_from_bytes(job, cor.read_bytes())
    _, vis64 = vis_from_bytes(job, O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")["output"])
    l2, worst = mixed_errors(vis_s[0], vis64[0])
    assert l2 <= MIXED_BARS[3][0] and worst <= MIXED_BARS[3][1]


def product_errors(g, r):
    """Per product against the reference r: relative L2 error over channels,
    and on channels with |r| >= 0.5 rms the largest relative amplitude error
    and phase error (rad). Returns arrays over products (rows)."""
    g, r = g.astype(np.complex128), r.astype(np.complex128)
    l2, amp, ph = [], [], []
    for k in range(r.shape[0]):
        rms = np.sqrt(np.mean(np.abs(r[k]) ** 2))
        if rms == 0.0:
            continue
        l2.append(np.sqrt(np.mean(np.abs(g[k] - r[k]) ** 2)) / rms)
        m = np.abs(r[k]) >= 0.5 * rms
        amp.append(np.max(np.abs(np.abs(g[k][m]) - np.abs(r[k][m])) / np.abs(r[k][m])))
        ph.append(np.max(np.abs(np.angle(g[k][m] * np.conj(r[k][m])))))
    return np.array(l2), np.array(amp), np.array(ph)


def fp32_parity(job, hdrs, vis_g, vis_r32, label):
    """The plain FP32 bars (compare_vis) against the FP32 oracle. Where the
    rounding floor of two FP32 correlators reaches 1e-5 on some checked
    channel (about sqrt(n) 2^-24 of a product's rms per channel after n
    segments of a sub-integration, and its largest value grows with the
    number of channels checked): a sub-integration of more than 4000
    segments, or more than 2e5 (product, channel) pairs, the GPU must instead
    be as accurate as the FP32 restatement of the reference is: against the
    FP64 oracle its relative L2, amplitude and phase errors at most 1.5 times
    the FP32 oracle's."""
    try:
        compare_vis(vis_g, vis_r32, label=label)
        return
    except AssertionError:
        if job.number_ffts_in_sub_integration() <= 4000 and vis_g.shape[0] * vis_g.shape[1] <= 200_000:
            raise
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    _, v64 = vis_from_bytes(job, r64["output"])
    e_gpu = product_errors(vis_g, v64[0])
    e_ref = product_errors(vis_r32, v64[0])
    names = ("relative L2", "relative amplitude", "phase (rad)")
    for n, a, b in zip(names, e_gpu, e_ref):
        print(f"{label} ({job.number_ffts_in_sub_integration()} segments per sub-integration, {vis_g.shape[0]} "
              f"products) vs FP64 oracle, {n}: GPU {a.max():.3e}, FP32 oracle {b.max():.3e}", flush=True)
        assert a.max() <= 1.5 * b.max() + 1e-7


# Mixed-precision error bars against the FP64 oracle, per product: the
# relative L2 error over channels, ||V - V64|| / ||V64|| (the visibility
# error relative to the FP64 oracle, at most 1e-3 for FP16), and the
# worst single channel relative to the product's rms, max |V - V64| / rms |V64|.
MIXED_BARS = {1: (1e-3, 4e-3), 2: (8e-3, 3e-2), 3: (1e-5, 5e-5)}


def mixed_errors(g, r):
    """(max relative L2 error, max worst-channel error) over the products
    (rows) of g against r; all-zero reference rows must be zero in g."""
    g, r = g.astype(np.complex128), r.astype(np.complex128)
    l2 = worst = 0.0
    for k in range(r.shape[0]):
        nr = np.sqrt(np.mean(np.abs(r[k]) ** 2))
        if nr == 0.0:
            assert np.all(g[k] == 0.0)
            continue
        d = np.abs(g[k] - r[k])
        l2 = max(l2, np.sqrt(np.mean(d ** 2)) / nr)
        worst = max(worst, np.max(d) / nr)
    return l2, worst


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_mixed_precision_vs_fp64(ctx, prec):
    """Mixed-precision X-step (FP16, BF16, split-FP16 on MFMA with FP32
    accumulation) against the FP64 oracle of the same pipeline, to the bars
    of MIXED_BARS (BF16 has an 8-bit mantissa)."""
    job, hdrs = configs.build_slices(seed=12, n_stations=4, pols=["R", "L"], fs=32_000_000, bw=16_000_000,
                                     nchan=1024, seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2,
                                     sky_freq=5_000_000_000, offsets=[0.0, 1e-6, -2e-6, 3e-6],
                                     rates=[0.0, 1e-7, 2e-7, -1e-7])[0]
    ctx.set_options(lut=LUT_CONFIG, precision=prec)
    ctx.set_parameters(job)
    for i, h in enumerate(hdrs):
        ctx.push_stream(i, h)
    ctx.correlate()
    out = ctx.output()
    ctx.set_options(lut=LUT_CONFIG, precision=0)
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    assert heads_g == heads_r
    l2, worst = mixed_errors(vis_g[0], vis_r[0])
    print(f"precision {prec}: vs FP64 oracle, relative L2 {l2:.3e}, worst channel {worst:.3e}")
    assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


@pytest.mark.parametrize("prec,nbins,gate", [(1, 8, (0.1, 0.6)), (3, 8, (0.1, 0.6)), (1, 32, (0.2, 0.3)),
                                             (3, 32, (0.2, 0.3)), (2, 32, (0.95, 0.05))])
def test_mixed_pulsar_vs_fp64(ctx, prec, nbins, gate):
    """Pulsar binning on the MFMA X-step against the FP64 oracle, per bin and
    product. Narrow bins (32 over a tenth of the period, 8.7 segments each)
    change bin inside most 16-segment stages; the gate (0.95, 0.05) wraps
    through phase 1."""
    from sfxg.job import Pulsar, ticks_to_mjd
    job, hdrs = configs.build_slices(seed=10, n_stations=2, pols=["R", "L"], fs=32_000_000, bw=16_000_000,
                                     nchan=1024, seconds=0.125, t_int_usec=125000, rho=0.05, bits=2,
                                     sky_freq=1_400_000_000, offsets=[0.0, 1e-6], rates=[0.0, 1e-7])[0]
    psr = Pulsar(nbins=nbins, gate_start=gate[0], gate_stop=gate[1], tmid=ticks_to_mjd(job.start_time) - 0.001,
                 DM=68.0, ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=psr, precision=prec)
    ctx.set_options(lut=LUT_CONFIG, precision=0)
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64", pulsar=psr)
    nb = nbins + 1
    heads_g, vis_g = vis_from_bytes(job, out, nb)
    heads_r, vis_r = vis_from_bytes(job, r64["output"], nb)
    assert heads_g == heads_r
    l2 = worst = 0.0
    for b in range(nb):
        e = mixed_errors(vis_g[b], vis_r[b])
        l2, worst = max(l2, e[0]), max(worst, e[1])
    print(f"pulsar precision {prec}, {nbins} bins: vs FP64 oracle, relative L2 {l2:.3e}, worst channel {worst:.3e}")
    assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


@pytest.mark.parametrize("nchan,prec,keep_ts", [(1024, 1, False), (1024, 2, True), (4096, 2, False),
                                                (16384, 1, False), (32768, 1, False), (1024, 3, False),
                                                (2048, 3, True), (16384, 3, False), (32768, 3, False)])
def test_packed_spectra(ctx, nchan, prec, keep_ts):
    """FP16, BF16 and split FP16 X-steps read spectra that the channeliser
    stored as their operands (fused F-engine, unfused channeliser, large-N
    runs, four-step): each part is the FP32 spectrum value rounded to nearest
    in that type (unit roundoff 2^-11 for FP16, 2^-8 for BF16, relative; split
    FP16 stores hi = fp16(x) and lo = fp16(x - hi), whose sum is x to 2^-21),
    where the FP16 / BF16 channelisers' own transform (twiddles as products
    of powers) may differ from the FP32 one by a few 2^-24 of the rms per
    stage (4e-6 of the rms is allowed), and the visibilities meet the
    mixed-precision bars against the FP64 oracle."""
    job, hdrs = configs.build_slices(seed=42, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                                     seconds=0.125, t_int_usec=125000, rho=0.05, bits=2, sky_freq=1_400_000_000,
                                     offsets=[0.0, 2e-6, -1e-6], rates=[0.0, 1e-7, -2e-7])[0]
    # packed FP16 / BF16 at 1024 to 16384 channels take the FP32 transform
    # here (the MFMA one has its own test, test_mfma_channeliser)
    capi.tune_clear()
    if nchan in (1024, 2048, 4096, 8192, 16384) and prec in (1, 2):
        capi.tune_import(capi.tune_key(nchan, 3, prec, False, nchan <= 4096 and not keep_ts) + "|0|0|0|0|0|0|1\n")
    try:
        capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=0, keep_time_series=keep_ts)
        sp32 = ctx.stage(capi.STAGE_SPECTRA, 1)
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec, keep_time_series=keep_ts)
        sph = ctx.stage(capi.STAGE_SPECTRA, 1)
    finally:
        capi.tune_clear()
        ctx.set_options(lut=LUT_CONFIG, precision=0)
    u = {1: 2.0 ** -11, 2: 2.0 ** -8, 3: 2.0 ** -21}[prec]
    rms = np.sqrt(np.mean(np.abs(sp32) ** 2))
    for part in (np.real, np.imag):
        f = part(sp32).astype(np.float64)
        h = part(sph).astype(np.float64)
        assert np.all(np.abs(h - f) <= 1.01 * u * np.abs(f) + (4e-6 if prec in (1, 2) else 1e-6) * rms)
    # the rounding is not a no-op: many values moved
    assert np.mean(sph != sp32) > (0.5 if prec != 3 else 0.2)
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    assert heads_g == heads_r
    l2, worst = mixed_errors(vis_g[0], vis_r[0])
    print(f"packed spectra nchan={nchan} precision {prec}: relative L2 {l2:.3e}, worst channel {worst:.3e}")
    assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


@pytest.mark.parametrize("nchan,prec,window,sideband", [(16384, 1, 4, "U"), (16384, 2, 4, "U"), (16384, 1, 0, "U"),
                                                        (16384, 1, 3, "L"), (8192, 1, 4, "U"), (8192, 2, 1, "L"),
                                                        (4096, 1, 4, "U"), (4096, 2, 2, "L"), (2048, 1, 3, "U"),
                                                        (2048, 1, 0, "L"), (1024, 1, 4, "U"), (1024, 1, 1, "U"),
                                                        (1024, 2, 0, "L")])
def test_mfma_channeliser(ctx, nchan, prec, window, sideband):
    """Packed FP16 / BF16 spectra at 1024 to 16384 channels come from the
    MFMA channeliser (after the unfused delay stage up to 4096): its
    transform of N = 1024 NT points takes NT tiles of 32 x 32 on MFMA with
    FP16 operands and FP32 accumulation (the windowed points, the first
    stage's results and the second stage's results are each rounded to FP16
    once) and a last NT-point stage in FP32. Against the FP32 channeliser's
    spectra, per part: every value within 3e-3 of the spectrum rms plus the
    spectrum type's own rounding, and the transform's excess over that
    rounding at most 4e-4 of the rms (a numpy model of the 16384-point
    decomposition with the same roundings gives 2.8e-4 rms and 1.4e-3 at
    most over 2^16 values). The spectra differ from those of the FP32
    transform (chan_fp32 = 1 in the tuning table) in most values (FP16) or
    many (BF16, whose rounding hides most of the difference), so the MFMA
    kernel is the one that ran, and the visibilities meet the mixed-precision
    bars against the FP64 oracle. 0.25 s at 32 Ms/s is 488 segments per
    stream at 16384 channels (runs of two segments per workgroup), 7812 at
    1024 (passes of 16 segments, the last pass partial); windows NONE (zero
    padding), RECT, COS, HAMMING and HANN, lower sideband flips."""
    job, hdrs = configs.build_slices(seed=44, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                                     seconds=0.25, t_int_usec=250000, rho=0.05, bits=2, sky_freq=1_400_000_000,
                                     offsets=[0.0, 2e-6, -1e-6], rates=[0.0, 1e-7, -2e-7], window=window,
                                     sideband=sideband)[0]
    capi.tune_clear()
    try:
        capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=0)
        sp32 = ctx.stage(capi.STAGE_SPECTRA, 1)
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec)
        sph = ctx.stage(capi.STAGE_SPECTRA, 1)
        capi.tune_import(capi.tune_key(nchan, 3, prec, False, nchan <= 4096) + "|0|0|0|0|0|0|1\n")
        capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec)
        spv = ctx.stage(capi.STAGE_SPECTRA, 1)
    finally:
        capi.tune_clear()
        ctx.set_options(lut=LUT_CONFIG, precision=0)
    u = {1: 2.0 ** -11, 2: 2.0 ** -8}[prec]
    rms = np.sqrt(np.mean(np.abs(sp32) ** 2))
    for part in (np.real, np.imag):
        f = part(sp32).astype(np.float64)
        d = np.abs(part(sph).astype(np.float64) - f)
        excess = np.maximum(d - 1.01 * u * np.abs(f), 0.0) / rms
        print(f"MFMA channeliser {nchan} precision {prec} window {window} {sideband}: transform error rms "
              f"{np.sqrt(np.mean(excess ** 2)):.2e}, max {excess.max():.2e} of the spectrum rms")
        assert excess.max() <= 3e-3
        assert np.sqrt(np.mean(excess ** 2)) <= 4e-4
    # BF16's own rounding (2^-8) hides most of the transforms' difference
    assert np.mean(sph != spv) > (0.5 if prec == 1 else 0.1)
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    assert heads_g == heads_r
    l2, worst = mixed_errors(vis_g[0], vis_r[0])
    print(f"MFMA channeliser {nchan} precision {prec}: vs FP64 oracle, relative L2 {l2:.3e}, worst channel {worst:.3e}")
    assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


@pytest.mark.parametrize("nchan", [16384, 1024])
def test_mfma_channeliser_fp32_series(ctx, nchan):
    """With keep_time_series the MFMA channeliser reads the FP32 series (and
    rounds each point to FP16 as it loads it) where it otherwise gets the
    delay stage's FP16 series: the two agree up to that rounding of slightly
    different FP32 values (the FP16 series' phasors come from the hardware
    sine and cosine), the kept series matches the FP64 oracle's delay stage,
    and both runs meet the mixed-precision bars against the FP64 oracle."""
    job, hdrs = configs.build_slices(seed=46, n_stations=2, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                                     seconds=0.25, t_int_usec=250000, rho=0.05, bits=2, sky_freq=1_400_000_000,
                                     offsets=[0.0, 1.5e-6], rates=[0.0, 1e-7])[0]
    capi.tune_clear()
    try:
        out_h = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=1)
        sp_h = ctx.stage(capi.STAGE_SPECTRA, 1)
        with pytest.raises(capi.SfxgError):
            ctx.stage(capi.STAGE_TIME_SERIES, 1)  # the series was held as FP16: refused on the host
        out_f = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=1, keep_time_series=True)
        sp_f = ctx.stage(capi.STAGE_SPECTRA, 1)
        ts = ctx.stage(capi.STAGE_TIME_SERIES, 1)
    finally:
        ctx.set_options(lut=LUT_CONFIG, precision=0)
    rms = np.sqrt(np.mean(np.abs(sp_f) ** 2))
    d = np.abs(sp_h.astype(np.complex128) - sp_f.astype(np.complex128)) / rms
    print(f"FP16 vs FP32 series at {nchan} channels: spectra differ by {np.sqrt(np.mean(d ** 2)):.2e} rms, "
          f"{d.max():.2e} at most, of the spectrum rms")
    assert np.sqrt(np.mean(d ** 2)) <= 4e-4 and d.max() <= 4e-3
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64", keep_intermediate=True)
    t_ref = r64["dc"][1]["time_series"].astype(np.float64)
    err_t = np.max(np.abs(ts - t_ref)) / np.sqrt(np.mean(t_ref ** 2))
    print(f"kept FP32 series vs the FP64 oracle's: {err_t:.2e} of its rms at most")
    assert err_t < 2e-5
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    for out in (out_h, out_f):
        heads_g, vis_g = vis_from_bytes(job, out)
        assert heads_g == heads_r
        l2, worst = mixed_errors(vis_g[0], vis_r[0])
        assert l2 <= MIXED_BARS[1][0] and worst <= MIXED_BARS[1][1]


@pytest.mark.parametrize("nchan", [16384, 1024])
def test_mfma_channeliser_strong_tone(ctx, nchan):
    """Every stream carries a tone of 3 sigma (before the 2-bit quantiser)
    plus its own noise. A 1024-point partial transform of the MFMA
    channeliser then sums the tone coherently to about 2.3e5 in the delay
    stage's units (a numpy model of the series), beyond FP16's 65504 (an
    unscaled FP16 operand would overflow to infinity); its stage operands
    carry the spectra's
    power-of-two scale before they are rounded to FP16, which keeps them near
    unit rms. The FP16 spectra stay finite and the visibilities meet the
    mixed-precision bars against the FP64 oracle."""
    ro = {}
    fs = 32_000_000
    job, _ = configs.build_slices(seed=45, n_stations=2, pols=["R"], fs=fs, bw=16_000_000, nchan=nchan,
                                  seconds=0.25, t_int_usec=250000, rho=0.05, bits=2, sky_freq=1_400_000_000,
                                  offsets=[0.0, 1e-6], rates=[0.0, 0.0], raw_out=ro)[0]
    from sfxg import synth
    rng = np.random.default_rng(45)
    nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
    hdrs = []
    for i, raw in enumerate(ro["raw"]):
        n = len(raw) * 4
        x = 3.0 * np.cos(2 * np.pi * 0.2345678 * np.arange(n) + i) + rng.standard_normal(n)
        raw2 = synth.pack(synth.quantise(x, 2), 2)
        hdrs.append(synth.header_stream(raw2, ro["raw_start"], job.delay_models[i], job.start_time, nsamp, fs, 2))
    capi.tune_clear()
    try:
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=1)
        sp = ctx.stage(capi.STAGE_SPECTRA, 0)
    finally:
        ctx.set_options(lut=LUT_CONFIG, precision=0)
    assert np.all(np.isfinite(sp.real)) and np.all(np.isfinite(sp.imag))
    peak = np.abs(sp).max() / np.sqrt(np.mean(np.abs(sp) ** 2))
    assert peak > 10.0  # the tone stands out of the noise
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    assert heads_g == heads_r
    l2, _ = mixed_errors(vis_g[0], vis_r[0])
    # one channel holds most of each product's power, so the worst channel
    # is measured against the product's peak rather than its rms
    g, r = vis_g[0].astype(np.complex128), vis_r[0].astype(np.complex128)
    worst = max(np.abs(g[k] - r[k]).max() / np.abs(r[k]).max() for k in range(r.shape[0]))
    print(f"strong tone {nchan} channels: spectrum peak {peak:.0f} x rms; vs FP64 oracle relative L2 {l2:.3e}, "
          f"worst channel {worst:.3e} of the peak")
    assert l2 <= MIXED_BARS[1][0] and worst <= 1e-3


@pytest.mark.parametrize("prec", [1, 2])
def test_packed_unaligned_folds(ctx, prec):
    """Sub-integrations of 453 segments start and end inside the four-segment
    quads of packed spectra: the X-step masks the neighbouring
    sub-integration's segments out of a shared quad. Unchunked (folds inside
    one chunk) and chunked (one sub-integration per chunk) both meet the
    mixed-precision bars against the FP64 oracle."""
    job, hdrs = configs.build_slices(seed=43, n_stations=3, pols=["R", "L"], fs=16_000_000, bw=8_000_000,
                                     nchan=1024, seconds=0.25, t_int_usec=250_000, sub_int_usec=29_000, rho=0.05,
                                     bits=2, sky_freq=5_000_000_000, offsets=[0.0, 2e-6, -1e-6],
                                     rates=[0.0, 1e-7, -2e-7])[0]
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    try:
        for chunk in (0, 1):
            out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec, chunk_mib=chunk)
            heads_g, vis_g = vis_from_bytes(job, out)
            assert heads_g == heads_r
            l2, worst = mixed_errors(vis_g[0], vis_r[0])
            print(f"unaligned folds precision {prec} chunk_mib={chunk}: relative L2 {l2:.3e}, worst {worst:.3e}")
            assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]
    finally:
        ctx.set_options(lut=LUT_CONFIG, precision=0)


@pytest.mark.parametrize("prec", [1, 3])
def test_mixed_32_stations(ctx, prec):
    """The 32-station shape of the mixed-precision X-step (64 streams, 2048
    baselines: the full LDS budget of the direct-to-LDS form, ring slots past
    64 KiB) against the FP64 oracle: FP16 on packed spectra and split FP16 on
    FP32 spectra."""
    n = 32
    offsets = [1e-7 * ((7 * i) % 13 - 6) for i in range(n)]
    rates = [1e-8 * ((5 * i) % 9 - 4) for i in range(n)]
    job, hdrs = configs.build_slices(seed=44, n_stations=n, pols=["R", "L"], fs=8_000_000, bw=4_000_000, nchan=256,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=offsets, rates=rates)[0]
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64")
    heads_r, vis_r = vis_from_bytes(job, r64["output"])
    try:
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec)
    finally:
        ctx.set_options(lut=LUT_CONFIG, precision=0)
    heads_g, vis_g = vis_from_bytes(job, out)
    assert heads_g == heads_r
    l2, worst = mixed_errors(vis_g[0], vis_r[0])
    print(f"32 stations precision {prec}: relative L2 {l2:.3e}, worst channel {worst:.3e}")
    assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


def _multi_fold_job(nchan, seed=20, bits=2, window=4):
    """0.25 s slice with 31.25 ms sub-integrations: eight folds."""
    return configs.build_slices(seed=seed, n_stations=3, pols=["R"], fs=16_000_000, bw=8_000_000, nchan=nchan,
                                seconds=0.25, t_int_usec=250_000, sub_int_usec=31_250, rho=0.05, bits=bits,
                                sky_freq=5_000_000_000, offsets=[0.0, 2e-6, -1e-6], rates=[0.0, 1e-7, -2e-7],
                                window=window)[0]


@pytest.mark.parametrize("nchan,prec", [(1024, 0), (16384, 0), (1024, 1)])
def test_chunked_matches_unchunked(ctx, nchan, prec):
    """A 1 MiB spectra budget processes the slice one sub-integration per
    chunk (eight chunks); the output bytes do not change."""
    job, hdrs = _multi_fold_job(nchan)
    a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec)
    b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=prec, chunk_mib=1)
    assert a == b
    if prec == 0:
        ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
        heads_g, vis_g = vis_from_bytes(job, b)
        heads_r, vis_r = vis_from_bytes(job, ref["output"])
        assert heads_g == heads_r
        compare_vis(vis_g[0], vis_r[0], label=f"chunked nchan={nchan}")


def test_chunked_pulsar_matches_unchunked(ctx):
    """FP32 pulsar binning carries every bin's running sums across chunks."""
    from sfxg.job import Pulsar, ticks_to_mjd
    job, hdrs = _multi_fold_job(1024, seed=21)
    psr = Pulsar(nbins=8, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=psr)
    b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=psr, chunk_mib=1)
    assert a == b


def test_batch_graph_replays_match_eager():
    """Slices of different shapes (channel counts, bit depths, pulsar,
    mixed precision) as one captured graph: three replays, with eager work and
    synchronisation between them, each give the eager bytes."""
    from sfxg.job import Pulsar, ticks_to_mjd
    cases = []
    cases.append((_multi_fold_job(1024, seed=30), None, 0))
    cases.append((_multi_fold_job(16384, seed=31), None, 0))
    cases.append((_multi_fold_job(512, seed=32, bits=1), None, 0))
    cases.append((_multi_fold_job(2048, seed=33), None, 1))
    j4 = _multi_fold_job(1024, seed=34)
    psr = Pulsar(nbins=4, gate_start=0.2, gate_stop=0.7, tmid=ticks_to_mjd(j4[0].start_time) - 0.001, DM=30.0,
                 ref_phase=0.1, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    cases.append((j4, psr, 0))
    eager_ctx = capi.Context(0)
    eager = [capi.correlate_slice(eager_ctx, job, hdrs, lut=LUT_CONFIG, pulsar=p, precision=prec)
             for (job, hdrs), p, prec in cases]
    batch = capi.Batch(0, n_branches=3)
    ctxs = []
    for (job, hdrs), p, prec in cases:
        c = capi.Context(0)
        c.set_options(lut=LUT_CONFIG, precision=prec)
        c.set_parameters(job, p)
        for i, h in enumerate(hdrs):
            c.push_stream(i, h)
        c.prepare()
        batch.add(c)
        ctxs.append(c)
    for rep in range(3):
        batch.launch()
        batch.synchronize()
        for k, c in enumerate(ctxs):
            assert c.output() == eager[k], f"replay {rep}, slice {k}"
        # eager work on another context between replays
        (job, hdrs), p, prec = cases[0]
        assert capi.correlate_slice(eager_ctx, job, hdrs, lut=LUT_CONFIG) == eager[0]
    batch.close()
    for c in ctxs:
        c.close()
    eager_ctx.close()


# (nchan, precision, pulsar, fields) with fields = (fe_threads, chan_wgs, xs_waves, mf_channels, mfp_wgs);
# split FP16 with 16-channel blocks hands bin N_c to the last-bin kernel (1024 = 64 x 16), with 24 it does not
TUNE_VARIANTS = [
    (512, 0, False, (64, 0, 0, 0, 0)),
    (1024, 0, False, (128, 0, 4, 0, 0)),
    (2048, 0, False, (256, 0, 0, 0, 0)),
    (4096, 0, False, (256, 0, 0, 0, 0)),
    (16384, 0, False, (0, 1024, 0, 0, 0)),
    (16384, 1, False, (0, 256, 0, 0, 0)),
    (16384, 2, False, (0, 1024, 0, 0, 0)),
    (8192, 0, False, (0, 4096, 0, 0, 0)),
    (1024, 1, False, (128, 0, 0, 8, 0)),
    (1024, 3, False, (128, 0, 0, 16, 0)),
    (1024, 3, False, (128, 0, 0, 24, 0)),
    (1024, 1, True, (0, 0, 0, 0, 512)),
    (1024, 0, True, (0, 0, 4, 0, 0)),
    (1024, 0, True, (0, 0, 2, 0, 0)),
]


@pytest.mark.parametrize("nchan,prec,pulsar,fields", TUNE_VARIANTS)
def test_tuned_launch_shapes(ctx, nchan, prec, pulsar, fields):
    """Every launch shape the autotuner can pick, forced through the tuning
    table, against the oracle (FP32) or the FP64 oracle (mixed)."""
    from sfxg.job import Pulsar, ticks_to_mjd
    job, hdrs = configs.build_slices(seed=40, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=nchan,
                                     seconds=0.125, t_int_usec=125000, rho=0.05, bits=2, sky_freq=1_400_000_000,
                                     offsets=[0.0, 2e-6, -1e-6], rates=[0.0, 1e-7, -2e-7])[0]
    psr = None
    if pulsar:
        psr = Pulsar(nbins=8, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                     ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    capi.tune_clear()
    key = capi.tune_key(nchan, 3, prec, pulsar, nchan <= 4096)
    capi.tune_import(key + "|" + "|".join(str(v) for v in fields) + "\n")
    assert key in capi.tune_export()
    try:
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=psr, precision=prec)
    finally:
        capi.tune_clear()
        ctx.set_options(lut=LUT_CONFIG, precision=0)
    nb = psr.nbins + 1 if psr else 1
    heads_g, vis_g = vis_from_bytes(job, out, nb)
    if prec == 0:
        ref = O.run_slice(job, hdrs, lut=LUT_CONFIG, pulsar=psr)
        heads_r, vis_r = vis_from_bytes(job, ref["output"], nb)
        assert heads_g == heads_r
        for b in range(nb):
            compare_vis(vis_g[b], vis_r[b], label=f"tuned {fields} bin {b}")
    else:
        r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64", pulsar=psr)
        heads_r, vis_r = vis_from_bytes(job, r64["output"], nb)
        assert heads_g == heads_r
        for b in range(nb):
            l2, worst = mixed_errors(vis_g[b], vis_r[b])
            assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


def test_tuned_fused_matches_unfused(ctx):
    """With a tuned F-engine schedule the unfused channeliser follows it, and
    both paths still agree to FP32 rounding."""
    job, hdrs = configs.build_slices(seed=41, n_stations=2, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=1024,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 1e-6], rates=[0.0, 1e-7])[0]
    capi.tune_clear()
    capi.tune_import(capi.tune_key(1024, 2, 0, False, True) + "|128|0|0|0|0|0|0|1\n" +
                     capi.tune_key(1024, 2, 0, False, False) + "|128|0|0|0|0\n")
    try:
        a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, keep_time_series=True)
        b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, keep_time_series=False)
        capi.tune_clear()
        c = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    finally:
        capi.tune_clear()
    heads_a, vis_a = vis_from_bytes(job, a)
    heads_b, vis_b = vis_from_bytes(job, b)
    assert heads_a == heads_b
    compare_vis(vis_a[0], vis_b[0], label="tuned fused vs unfused")
    assert b != c  # another radix schedule: other rounding


def test_autotune_then_correlate(ctx):
    job, hdrs = configs.build_slices(seed=42, n_stations=3, pols=["R"], fs=32_000_000, bw=16_000_000, nchan=2048,
                                     seconds=0.125, t_int_usec=125000, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 1e-6, 2e-6], rates=[0.0, 1e-7, 0.0])[0]
    capi.tune_clear()
    try:
        ctx.set_options(lut=LUT_CONFIG)
        ctx.set_parameters(job)
        for i, h in enumerate(hdrs):
            ctx.push_stream(i, h)
        ctx.prepare()
        report = ctx.autotune(reps=2)
        print(report)
        for route in ("(delay + channeliser)", "(fused)"):  # both F-engine routes, two thread counts each
            assert f"fe_threads {route}=512" in report and f"fe_threads {route}=256" in report
        assert "xs_waves=4" in report
        assert capi.tune_key(2048, 3, 0, False, True) in capi.tune_export()
        ctx.correlate()
        out = ctx.output()
    finally:
        capi.tune_clear()
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    compare_vis(vis_g[0], vis_r[0], label="autotuned")


@pytest.mark.parametrize("n_stations,nchan", [(32, 256), (24, 1024), (3, 512)])
def test_tile16_matches_tile64(ctx, n_stations, nchan):
    """The FP32 X-step on 16-channel blocks holding every stream-pair tile
    (k_xstep_tile16: 136, 78 and 3 tiles, so three, two and one tiles per
    slot) gives the same bits as the register tiles on 64-channel blocks
    (k_xstep_tile): both form the reference's products and sums in its order.
    Both match the oracle within the FP32 bars (fp32_parity: with 2080
    products the largest relative error of the FP32 floor reaches 1e-5)."""
    rng = np.random.default_rng(n_stations)
    job, hdrs = configs.build_slices(seed=44, n_stations=n_stations, pols=["R", "L"], fs=16_000_000, bw=8_000_000,
                                     nchan=nchan, seconds=0.03125, t_int_usec=31250, rho=0.05, bits=2,
                                     sky_freq=1_400_000_000, offsets=list(rng.uniform(-3e-6, 3e-6, n_stations)),
                                     rates=list(rng.uniform(-2e-7, 2e-7, n_stations)))[0]
    M = len(job.station_streams)
    print(f"case tile16 vs tile64: {M} streams, {nchan} channels", flush=True)
    outs = {}
    for xs in (1, 2):
        capi.tune_clear()
        capi.tune_import(capi.tune_key(nchan, M, 0, False, nchan <= 4096) + f"|0|0|{xs}|0|0\n")
        try:
            outs[xs] = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
        finally:
            capi.tune_clear()
    assert outs[1] == outs[2]
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    heads_g, vis_g = vis_from_bytes(job, outs[2])
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    fp32_parity(job, hdrs, vis_g[0], vis_r[0], f"tile16, {M} streams")


@pytest.mark.parametrize("xs_waves", [0, 2, 8])
def test_station_order_conjugates(ctx, xs_waves):
    """Station numbers out of stream order: crosses list the lower station
    first, so some products are taken the other way round (the register-tile
    X-step stores their conjugates)."""
    import dataclasses
    job, hdrs = configs.build_slices(seed=43, n_stations=6, pols=["R"], fs=16_000_000, bw=8_000_000, nchan=512,
                                     seconds=0.125, t_int_usec=125000, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 1e-6, -1e-6, 2e-6, 0.5e-6, -2e-6],
                                     rates=[0.0, 1e-7, 0.0, -1e-7, 2e-7, 0.0])[0]
    perm = [4, 0, 5, 2, 1, 3]
    job.station_streams = [dataclasses.replace(s, station_number=perm[i]) for i, s in enumerate(job.station_streams)]
    capi.tune_clear()
    if xs_waves:
        capi.tune_import(capi.tune_key(512, 6, 0, False, True) + f"|0|0|{xs_waves}|0|0\n")
    try:
        out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    finally:
        capi.tune_clear()
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    compare_vis(vis_g[0], vis_r[0], label=f"station order, xs_waves={xs_waves}")
