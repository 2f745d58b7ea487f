"""Each workload configuration at its own shape (stations, polarisations,
channels, sample rate, bit depth, phase bins), shortened in time, through the
C ABI against the oracle on the same seeded input:

* configuration 1: 8 stations x 2 pols, 4096 channels, 64 Ms/s, delay
  offsets U[-5, 5] ms; samples and statistics bit-exact, visibilities to the
  plain FP32 bars;
* configuration 2: 32 stations x 2 pols, 16384 channels, 128 Ms/s, a second
  point source; FP16, BF16 and split FP16 against the FP64 oracle with the
  per-product relative amplitude and phase errors printed, FP32 to the plain
  bars;
* configuration 3: a seeded subset of the mixed batch's draws (32768
  channels, 1-bit, 24 stations among them) as one captured graph, every
  replay against the oracle;
* configuration 4: 12 stations x 2 pols, 1024 channels, 64 + 1 bins with the
  injected dispersed pulse; FP32 per bin to the plain bars, FP16 and split
  FP16 per bin against the FP64 oracle, also one sub-integration per chunk.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from sfxg import capi, configs
from sfxg.job import LUT_CONFIG
from test_gpu_parity import MIXED_BARS, compare_vis, fp32_parity, mixed_errors, product_errors, vis_from_bytes

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

PREC_NAME = {0: "FP32", 1: "FP16", 2: "BF16", 3: "split FP16"}


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def report(label, g, r):
    l2, amp, ph = product_errors(g, r)
    say(f"{label}: {len(l2)} products; relative L2 max {l2.max():.3e} median {np.median(l2):.3e}; "
        f"relative amplitude max {amp.max():.3e} median {np.median(amp):.3e}; "
        f"phase max {ph.max():.3e} rad median {np.median(ph):.3e} rad")


def run(ctx, job, hdrs, prec=0, pulsar=None, chunk_mib=0):
    try:
        return capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG, pulsar=pulsar, precision=prec,
                                    chunk_mib=chunk_mib)
    finally:
        ctx.set_options(lut=LUT_CONFIG, precision=0)


# ---------------------------------------------------------------------------
# configuration 1
# ---------------------------------------------------------------------------
def test_config1_shape(ctx):
    (job, hdrs), = configs.config1_subband(3, seconds=0.03125, t_int_usec=31250)
    M = len(job.station_streams)
    say(f"config 1 shape: {M} streams, {job.fft_size_correlation} channels, "
        f"{job.number_ffts_in_integration()} segments")
    assert M == 16 and job.fft_size_correlation == 4096 and job.sample_rate == 64_000_000
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG, keep_intermediate=True)
    out = run(ctx, job, hdrs)
    for i in range(M):
        np.testing.assert_array_equal(ctx.stage(capi.STAGE_SAMPLES, i), ref["b2f"][i]["samples"])
        lv, ts = ctx.statistics(i)
        np.testing.assert_array_equal(lv, ref["stats"][i][0])
        np.testing.assert_array_equal(ts, ref["stats"][i][1])
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    compare_vis(vis_g[0], vis_r[0], label="config 1 shape")


# ---------------------------------------------------------------------------
# configuration 2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    (job, hdrs), = configs.config2_slices()
    say(f"config 2 shape: {len(job.station_streams)} streams, {job.fft_size_correlation} channels, "
        f"{job.number_ffts_in_integration()} segments, {len(job.baselines())} products")
    return dict(job=job, hdrs=hdrs, r64=O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64"))


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_config2_shape(ctx, c2, prec):
    job, hdrs = c2["job"], c2["hdrs"]
    assert len(job.station_streams) == 64 and job.fft_size_correlation == 16384
    say(f"config 2 shape, {PREC_NAME[prec]}")
    out = run(ctx, job, hdrs, prec)
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, c2["r64"]["output"])
    assert heads_g == heads_r
    report(f"config 2 {PREC_NAME[prec]} vs FP64 oracle", vis_g[0], vis_r[0])
    if prec == 0:
        r32 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f32")
        _, v32 = vis_from_bytes(job, r32["output"])
        compare_vis(vis_g[0], v32[0], label="config 2 FP32 vs FP32 oracle")
    else:
        l2, worst = mixed_errors(vis_g[0], vis_r[0])
        say(f"config 2 {PREC_NAME[prec]}: relative L2 {l2:.3e}, worst channel {worst:.3e}")
        assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]


# ---------------------------------------------------------------------------
# configuration 3
# ---------------------------------------------------------------------------
C3_JOBS = [0, 3, 6, 12, 46, 2, 5]  # 32768 ch 1-bit; 32768 ch 24 st; 256 ch 22 st 1-bit; ...; 16384 ch 24 st 1-bit


@pytest.fixture(scope="module")
def c3():
    cases = []
    for j in C3_JOBS:
        (job, hdrs), = configs.config3_slices(j)
        cases.append((j, job, hdrs))
    return cases


@pytest.mark.parametrize("prec", [1, 0])
def test_config3_subset_one_graph(c3, prec):
    """The subset as one captured graph (four branches), replayed twice with
    eager work in between; every slice of every replay against the oracle
    (FP64 for FP16, FP32 for the plain path)."""
    shapes = [(j, job.fft_size_correlation, len(job.station_streams), job.station_streams[0].bits_per_sample)
              for j, job, _ in c3]
    say(f"config 3 subset, {PREC_NAME[prec]}: (job, channels, stations, bits) {shapes}")
    assert any(s[1] == 32768 for s in shapes) and any(s[3] == 1 for s in shapes) and any(s[2] == 24 for s in shapes)
    refs = []
    for j, job, hdrs in c3:
        r = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64" if prec else "f32")
        refs.append(vis_from_bytes(job, r["output"]))
    batch = capi.Batch(0, n_branches=4)
    ctxs = []
    try:
        for j, job, hdrs in c3:
            c = capi.Context(0)
            ctxs.append(c)
            c.set_options(lut=LUT_CONFIG, precision=prec)
            c.set_parameters(job)
            for i, h in enumerate(hdrs):
                c.push_stream(i, h)
            c.prepare()
            batch.add(c)
        first = None
        for rep in range(2):
            say(f"config 3 subset, {PREC_NAME[prec]}: graph replay {rep}")
            batch.launch()
            batch.synchronize()
            outs = [c.output() for c in ctxs]
            if first is None:
                first = outs
            assert outs == first
            for (j, job, _), out, (heads_r, vis_r) in zip(c3, outs, refs):
                heads_g, vis_g = vis_from_bytes(job, out)
                assert heads_g == heads_r
                if prec == 0:
                    fp32_parity(job, c3[[x[0] for x in c3].index(j)][2], vis_g[0], vis_r[0], f"config 3 job {j}")
                else:
                    l2, worst = mixed_errors(vis_g[0], vis_r[0])
                    say(f"config 3 job {j} FP16: relative L2 {l2:.3e}, worst channel {worst:.3e}")
                    assert l2 <= MIXED_BARS[1][0] and worst <= MIXED_BARS[1][1]
            # eager work on another context between the replays
            e = capi.Context(0)
            j, job, hdrs = c3[-1]
            capi.correlate_slice(e, job, hdrs, lut=LUT_CONFIG)
            e.close()
    finally:
        batch.close()
        for c in ctxs:
            c.close()


# ---------------------------------------------------------------------------
# configuration 4
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4():
    slices, psr = configs.config4_slices()
    (job, hdrs), = slices
    say(f"config 4 shape: {len(job.station_streams)} streams, {psr.nbins + 1} bins, "
        f"{job.number_ffts_in_integration()} segments")
    r64 = O.run_slice(job, hdrs, lut=LUT_CONFIG, precision="f64", pulsar=psr)
    return dict(job=job, hdrs=hdrs, psr=psr, r64=r64)


def test_config4_shape_fp32(ctx, c4):
    job, hdrs, psr = c4["job"], c4["hdrs"], c4["psr"]
    nb = psr.nbins + 1
    assert len(job.station_streams) == 24 and nb == 65
    say("config 4 shape, FP32")
    out = run(ctx, job, hdrs, 0, psr)
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG, pulsar=psr)
    heads_g, vis_g = vis_from_bytes(job, out, nb)
    heads_r, vis_r = vis_from_bytes(job, ref["output"], nb)
    assert heads_g == heads_r
    for b in range(nb):
        compare_vis(vis_g[b], vis_r[b], label=f"config 4 bin {b}")
    # the injected pulse is in the data: at the pulse's phase the same-pol
    # cross-correlations reach about 0.88 * 0.2 / 1.2 (2-bit), far above the
    # continuum of the off-gate bin
    peak = 1 + int((0.5 - psr.gate_start) * (nb - 1) / (psr.gate_stop - psr.gate_start))
    pol = [s.polarisation for s in job.station_streams]
    same = [k for k, (i, j) in enumerate(job.baselines()) if i != j and pol[i] == pol[j]]
    cross = lambda v: np.mean(np.real(v[same]))  # noqa: E731
    say(f"config 4: mean same-pol cross real part, bin {peak}: {cross(vis_g[peak]):.4f}, "
        f"off-gate bin: {cross(vis_g[0]):.4f}")
    assert cross(vis_g[peak]) > 0.1 and abs(cross(vis_g[0])) < 0.02


@pytest.mark.parametrize("prec,sub_chunks", [(1, False), (3, False), (1, True), (3, True)])
def test_config4_shape_mixed(ctx, c4, prec, sub_chunks):
    """Against the FP64 oracle per bin. sub_chunks: sub-integrations of a
    quarter of the slice and a 1 MiB spectra budget, so the binned Gram sums
    are carried across four chunks (the pulsar sums do not depend on the
    sub-integrations)."""
    job, hdrs, psr = c4["job"], c4["hdrs"], c4["psr"]
    nb = psr.nbins + 1
    if sub_chunks:
        job = dataclasses.replace(job, sub_integration_time=job.integration_time // 4)
    say(f"config 4 shape, {PREC_NAME[prec]}, chunked={sub_chunks}")
    out = run(ctx, job, hdrs, prec, psr, chunk_mib=1 if sub_chunks else 0)
    heads_g, vis_g = vis_from_bytes(job, out, nb)
    heads_r, vis_r = vis_from_bytes(job, c4["r64"]["output"], nb)
    assert heads_g == heads_r
    l2 = worst = 0.0
    for b in range(nb):
        e = mixed_errors(vis_g[b], vis_r[b])
        l2, worst = max(l2, e[0]), max(worst, e[1])
    say(f"config 4 {PREC_NAME[prec]} chunked={sub_chunks}: relative L2 {l2:.3e}, worst channel {worst:.3e}")
    assert l2 <= MIXED_BARS[prec][0] and worst <= MIXED_BARS[prec][1]
