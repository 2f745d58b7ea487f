"""The forms of Correlation_core beside the plain one, on the GPU against the
oracle through the C ABI: lag re-channelisation (integration_write with
fft_size_correlation > number_channels, correlation_core.cc:417-440, and the
Mask_parameters of control_parameters.h:107-114), the phased-array core
(Correlation_core_phased, correlation_core_phased.cc:13-126) and extra phase
centres (sub_integration / uvshift, correlation_core.cc:515-592)."""
import numpy as np
import pytest

from oracle import oracle as O
from sfxg import capi, configs
from sfxg.job import (LUT_CONFIG, WINDOW_COS, WINDOW_HAMMING, WINDOW_HANN, WINDOW_NONE, WINDOW_RECT, DelayModel,
                      MaskParameters, PhaseCenter, Pulsar, ticks_to_mjd)

from test_gpu_parity import compare_vis, vis_from_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def make_job(nc, nchan, window, seed=21, stations=3, pols=("R",), seconds=0.0625):
    job, hdrs = configs.build_slices(seed=seed, n_stations=stations, pols=list(pols), fs=32_000_000,
                                     bw=16_000_000, nchan=nc, seconds=seconds, t_int_usec=int(seconds * 1e6),
                                     rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=[0.0, 2.5e-6, -1.75e-6][:stations],
                                     rates=[0.0, 2e-7, -1e-7][:stations], window=window)[0]
    job.number_channels = nchan
    return job, hdrs


def mask_for(kind, nc, nchan, seed):
    rng = np.random.default_rng(seed)
    if kind == "band":  # flag three bands of bins
        m = np.ones(nc + 1)
        for a, b in ((0, nc // 64), (nc // 3, nc // 3 + nc // 50), (nc - nc // 40, nc + 1)):
            m[a:b] = 0.0
        return list(m)
    if kind == "random":
        return list(rng.uniform(0.5, 1.5, nc + 1))
    return None


CASES = [
    # nc, nchan, window, normalize, mask, user lag window
    (1024, 256, WINDOW_HANN, False, None, False),
    (4096, 64, WINDOW_COS, True, "band", False),
    (16384, 4096, WINDOW_RECT, False, None, True),
    (2048, 1024, WINDOW_NONE, False, None, False),
    (8192, 2, WINDOW_HAMMING, True, None, False),
    (512, 128, WINDOW_HANN, False, "random", False),
]


@pytest.mark.parametrize("nc,nchan,window,normalize,mask,user_window", CASES)
def test_rechannelisation(ctx, nc, nchan, window, normalize, mask, user_window):
    print(f"case nc={nc} nchan={nchan} window={window} normalize={normalize} mask={mask}", flush=True)
    job, hdrs = make_job(nc, nchan, window)
    win = list(O.lag_window(WINDOW_HANN, nchan, nc) ** 2) if user_window else None
    job.mask = MaskParameters(normalize=normalize, mask=mask_for(mask, nc, nchan, nc + nchan), window=win)
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    assert vis_g[0].shape[1] == nchan + 1
    # the normalised visibilities before re-channelisation, and the stored channels
    norm = ctx.stage(capi.STAGE_NORMALISED)[0]
    nr = ref["norm"][0]
    compare_vis(norm, nr[..., 0] + 1j * nr[..., 1], label="normalised")
    chan = ctx.stage(capi.STAGE_CHANNELS)[0]
    np.testing.assert_array_equal(chan, vis_g[0])
    compare_vis(vis_g[0], vis_r[0], label=f"channels nc={nc} nchan={nchan}")


def test_rechannelisation_pulsar(ctx):
    """Correlation_core_pulsar builds the same mask and window
    (correlation_core_pulsar.cc:182-185) and writes every bin through
    integration_write."""
    job, hdrs = make_job(1024, 128, WINDOW_HANN, seed=23, stations=2, pols=("R", "L"), seconds=0.125)
    job.mask = MaskParameters(normalize=False)
    psr = Pulsar(nbins=4, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG, pulsar=psr)
    out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=psr)
    nb = psr.nbins + 1
    heads_g, vis_g = vis_from_bytes(job, out, nb)
    heads_r, vis_r = vis_from_bytes(job, ref["output"], nb)
    assert heads_g == heads_r
    for b in range(nb):
        compare_vis(vis_g[b], vis_r[b], label=f"pulsar bin {b}")


def test_mask_tables_rebuilt(ctx):
    """A context reused for a job without mask parameters rebuilds its tables:
    ones for the mask and the lag window derived from the job's window."""
    job, hdrs = make_job(1024, 256, WINDOW_HANN, seed=24)
    job.mask = MaskParameters(normalize=True, mask=mask_for("band", 1024, 256, 1))
    a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    ref_a = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    job.mask = None
    b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    ref_b = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    for got, ref in ((a, ref_a), (b, ref_b)):
        _, vg = vis_from_bytes(job, got)
        _, vr = vis_from_bytes(job, ref["output"])
        compare_vis(vg[0], vr[0], label="mask reuse")
    assert a != b


def test_rechannelisation_rejects_unsupported(ctx):
    job, hdrs = make_job(1024, 256, WINDOW_HANN)
    for nchan in (384, 1, 2048):
        job.number_channels = nchan
        with pytest.raises(capi.SfxgError):
            ctx.set_parameters(job)
    job.number_channels = 256
    job.mask = MaskParameters(mask=[1.0] * 100)
    ctx.set_options(lut=LUT_CONFIG)
    ctx.set_parameters(job)
    for i, h in enumerate(hdrs):
        ctx.push_stream(i, h)
    with pytest.raises(capi.SfxgError):
        ctx.correlate()


def phased_job(nc, sub_usec, stations, pols, window=WINDOW_HANN, seed=31):
    return configs.build_slices(seed=seed, n_stations=stations, pols=list(pols), fs=32_000_000, bw=16_000_000,
                                nchan=nc, seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2,
                                sky_freq=5_000_000_000, offsets=[0.0, 2.5e-6, -1.75e-6][:stations],
                                rates=[0.0, 2e-7, -1e-7][:stations], window=window, sub_int_usec=sub_usec)[0]


@pytest.mark.parametrize("nc,sub_usec,stations,pols,window", [
    (1024, 10_000, 3, ("R",), WINDOW_HANN),     # 7 rows, the last one partial
    (4096, 62_500, 2, ("R", "L"), WINDOW_HANN),  # one row
    (2048, 15_616, 2, ("R",), WINDOW_NONE),     # 976 segments in rows of 244: four full rows, an empty fifth
])
def test_phased_array(ctx, nc, sub_usec, stations, pols, window):
    print(f"case phased nc={nc} sub={sub_usec} stations={stations} pols={pols} window={window}", flush=True)
    job, hdrs = phased_job(nc, sub_usec, stations, pols, window)
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG, phased=True)
    out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, phased_array=True)
    rows = ref["vis"][0].shape[0]
    assert ctx.number_of_baselines() == rows
    heads_g, vis_g = _phased_vis(job, out, rows)
    heads_r, vis_r = _phased_vis(job, ref["output"], rows)
    assert heads_g == heads_r
    full = [r for r in range(rows) if np.any(vis_r[r] != 0)]
    if window == WINDOW_NONE:
        assert full == list(range(rows - 1))
    for r in range(rows):
        if r not in full:
            assert np.all(vis_g[r] == 0), r
    compare_vis(vis_g[full], vis_r[full], label=f"phased nc={nc} rows={rows}")


def _phased_vis(job, out, rows):
    """integration_write bytes of a phased slice: headers and the rows."""
    import copy
    j = copy.copy(job)
    j.baselines = lambda: [(0, 0)] * rows
    heads, vis = vis_from_bytes(j, out)
    return heads, vis[0]


def test_phased_chunked_matches_unchunked(ctx):
    """Rows carry over the chunks of a slice: with a 1 MiB spectra budget
    (one sub-integration per chunk, boundaries at buffer ends, not at row
    starts) the rows agree with one chunk to FP32 rounding; the block sums
    split where a chunk ends."""
    job, hdrs = phased_job(1024, 10_000, 3, ("R",))
    a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, phased_array=True)
    b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, phased_array=True, chunk_mib=1)
    rows = ctx.number_of_baselines()
    ha, va = _phased_vis(job, a, rows)
    hb, vb = _phased_vis(job, b, rows)
    assert ha == hb
    compare_vis(vb, va, label="phased chunked")


def test_phased_rejects(ctx):
    job, hdrs = phased_job(1024, 10_000, 2, ("R",))
    ctx.set_options(lut=LUT_CONFIG, phased_array=True)
    job.number_channels = 256
    with pytest.raises(capi.SfxgError):
        ctx.set_parameters(job)
    job.number_channels = 1024
    psr = Pulsar(nbins=4, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    with pytest.raises(capi.SfxgError):
        ctx.set_parameters(job, psr)
    ctx.set_options(lut=LUT_CONFIG, precision=1, phased_array=True)
    ctx.set_parameters(job)
    for i, h in enumerate(hdrs):
        ctx.push_stream(i, h)
    with pytest.raises(capi.SfxgError):
        ctx.correlate()
    # options changed after the parameters
    ctx.set_options(lut=LUT_CONFIG, phased_array=False)
    with pytest.raises(capi.SfxgError):
        ctx.correlate()


def pc_job(nc, sub_usec, n_extra, window, seed=51, second_source=False):
    fs = 32_000_000
    offsets, rates = [0.0, 2.5e-6, -1.75e-6], [0.0, 2e-7, -1e-7]
    shifts = [[0.0, 8 / fs, -6 / fs], [0.0, -3.1e-8, 4.4e-8], [1e-8, 0.0, -2.5e-7]][:n_extra]
    extra = (0.1, [o + d for o, d in zip(offsets, shifts[0])], rates) if second_source else None
    job, hdrs = configs.build_slices(seed=seed, n_stations=3, pols=["R"], fs=fs, bw=fs // 2, nchan=nc,
                                     seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                     offsets=offsets, rates=rates, window=window, sub_int_usec=sub_usec,
                                     extra=extra)[0]
    rng = np.random.default_rng(seed)
    pcs = []
    for k, dl in enumerate(shifts):
        ms = []
        for i, m in enumerate(job.delay_models):
            c = m.delay.copy()
            c[:, 0] += dl[i]
            c[:, 1] += 1e-9 * k  # a small rate difference as well
            ms.append(DelayModel(m.scan_begin_ticks, m.interval, c))
        pcs.append(PhaseCenter(ms, uvw=rng.standard_normal((3, 3)) * 1e6, source_nr=10 + k))
    job.phase_centers = pcs
    job.source_nr = 3
    return job, hdrs


@pytest.mark.parametrize("nc,sub_usec,n_extra,window,second,nchan", [
    (1024, 10_000, 2, WINDOW_HANN, False, None),   # 6 sub-integrations, two shifted centres
    (512, 15_625, 1, WINDOW_NONE, True, None),     # a second source in the shifted centre
    (4096, 62_500, 3, WINDOW_COS, False, 1024),    # one sub-integration, re-channelised to 1024
])
def test_phase_centers(ctx, nc, sub_usec, n_extra, window, second, nchan):
    print(f"case phase centres nc={nc} sub={sub_usec} extra={n_extra} window={window} nchan={nchan}", flush=True)
    job, hdrs = pc_job(nc, sub_usec, n_extra, window, second_source=second)
    if nchan:
        job.number_channels = nchan
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    out = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    n_pc = 1 + n_extra
    heads_g, vis_g = vis_from_bytes(job, out, n_pc)
    heads_r, vis_r = vis_from_bytes(job, ref["output"], n_pc)
    assert heads_g == heads_r
    for j in range(n_pc):
        compare_vis(vis_g[j], vis_r[j], label=f"phase centre {j}")


def test_phase_centers_chunked_matches_unchunked(ctx):
    """Chunks hold whole sub-integrations, so every centre gets the same
    folds in the same order: identical bytes."""
    job, hdrs = pc_job(1024, 10_000, 2, WINDOW_HANN)
    a = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    b = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, chunk_mib=1)
    assert a == b


def test_phase_centers_reject(ctx):
    job, hdrs = pc_job(1024, 10_000, 1, WINDOW_HANN)
    # mixed precision
    with pytest.raises(capi.SfxgError):
        capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, precision=1)
    # pulsar binning with several centres
    psr = Pulsar(nbins=4, gate_start=0.1, gate_stop=0.6, tmid=ticks_to_mjd(job.start_time) - 0.001, DM=68.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    ctx.set_options(lut=LUT_CONFIG)
    with pytest.raises(capi.SfxgError):
        ctx.set_parameters(job, psr)
    # centres announced but not given
    pcs = job.phase_centers
    job.phase_centers = []
    job.n_phase_centers = 2
    ctx.set_parameters(job)
    for i, h in enumerate(hdrs):
        ctx.push_stream(i, h)
    with pytest.raises(capi.SfxgError):
        ctx.correlate()
    job.phase_centers = pcs
