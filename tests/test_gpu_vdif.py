"""Raw VDIF ingest on the GPU (sfxg_push_vdif) against the oracle's
restatement of the input node (VDIF_reader, Input_node_data_writer), through
the C ABI: frames in device memory, header words lifted on the device, the
HEADER records built on the host, the payload gathered on the device."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sfxg import capi, configs, synth
from sfxg.job import LUT_CONFIG

from test_gpu_parity import compare_vis, vis_from_bytes

pytestmark = pytest.mark.gpu

FS, BITS, PAYLOAD, EPOCH, REF_MJD = 32_000_000, 2, 2000, 46, 60000


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def to_device(data: bytes):
    L = C.CDLL("libamdhip64.so")
    p = C.c_void_p()
    assert L.hipMalloc(C.byref(p), C.c_size_t(len(data))) == 0
    src = (C.c_uint8 * len(data)).from_buffer_copy(data)
    assert L.hipMemcpy(p, src, C.c_size_t(len(data)), 1) == 0
    return L, p


@pytest.mark.parametrize("nchan,rates", [(1024, [0.0, 4e-7, -2e-7]), (4096, [0.0, -3e-6, 2e-6])])
def test_vdif_ingest(ctx, nchan, rates):
    print(f"case vdif nchan={nchan} rates={rates}", flush=True)
    raw = {}
    job, _ = configs.build_slices(seed=71, n_stations=3, pols=["R", "L"], fs=FS, bw=FS // 2, nchan=nchan,
                                  seconds=0.0625, t_int_usec=62500, rho=0.05, bits=BITS, sky_freq=5_000_000_000,
                                  offsets=[0.0, 2.5e-6, -1.75e-6], rates=rates, raw_out=raw,
                                  margin_multiple=PAYLOAD * 8 // BITS)[0]
    # one VDIF stream per station, one thread per polarisation; station 0 has
    # a fill frame, station 1 an invalid frame, station 2 legacy headers
    stations = sorted({s.station_number for s in job.station_streams})
    per_stream = {}
    vdifs = []
    for k, sn in enumerate(stations):
        idx = [i for i, s in enumerate(job.station_streams) if s.station_number == sn]
        threads = [3 + 2 * c for c in range(len(idx))]
        legacy = k == 2
        v = synth.vdif_stream([raw["raw"][i] for i in idx], threads, raw["raw_start"], FS, BITS, PAYLOAD, EPOCH,
                              invalid=[(1, 30)] if k == 1 else [], fill=[4] if k == 0 else [], legacy=legacy)
        fb = PAYLOAD + (16 if legacy else 32)
        vdifs.append(v)
        for c, i in enumerate(idx):
            per_stream[i] = (k, fb, threads, c)
    hdrs = [O.vdif_header_stream(vdifs[k], fb, th, c, job, i, REF_MJD)
            for i, (k, fb, th, c) in sorted(per_stream.items())]
    ref = O.run_slice(job, hdrs, lut=LUT_CONFIG)
    dev = [to_device(v) for v in vdifs]
    try:
        ctx.set_options(lut=LUT_CONFIG)
        ctx.set_parameters(job)
        for i, (k, fb, th, c) in sorted(per_stream.items()):
            ctx.push_vdif(i, dev[k][1].value, len(vdifs[k]), fb, c, th, REF_MJD)
        ctx.correlate()
        out = ctx.output()
        stats = [ctx.statistics(i) for i in range(len(job.station_streams))]
        tsys = ctx.tsys()
    finally:
        for L, p in dev:
            L.hipFree(p)
    # the same HEADER semantics as the oracle's streams pushed as bytes
    out2 = capi.correlate_slice(ctx, job, hdrs, lut=LUT_CONFIG)
    assert out == out2
    for i, (lv, ts) in enumerate(stats):
        np.testing.assert_array_equal(lv, ref["stats"][i][0])
        np.testing.assert_array_equal(ts, ref["stats"][i][1])
    assert ref["stats"][2][0][4] > 0 or ref["stats"][3][0][4] > 0  # the invalid frame reached station 1
    # the tsys side product (tsys_write) of the same slice
    assert tsys == O.tsys_messages(job, ref["stats"])
    heads_g, vis_g = vis_from_bytes(job, out)
    heads_r, vis_r = vis_from_bytes(job, ref["output"])
    assert heads_g == heads_r
    compare_vis(vis_g[0], vis_r[0], label=f"vdif nchan={nchan}")


def test_vdif_rejects(ctx):
    raw = {}
    job, _ = configs.build_slices(seed=72, n_stations=2, pols=["R"], fs=FS, bw=FS // 2, nchan=512,
                                  seconds=0.0625, t_int_usec=62500, rho=0.05, bits=BITS, sky_freq=5_000_000_000,
                                  offsets=[0.0, 1e-6], rates=[0.0, 0.0], raw_out=raw,
                                  margin_multiple=PAYLOAD * 8 // BITS)[0]
    v = synth.vdif_stream([raw["raw"][0]], [1], raw["raw_start"], FS, BITS, PAYLOAD, EPOCH)
    L, p = to_device(v)
    try:
        ctx.set_options(lut=LUT_CONFIG)
        ctx.set_parameters(job)
        with pytest.raises(capi.SfxgError):  # frame length differs from the headers'
            ctx.push_vdif(0, p.value, len(v), PAYLOAD + 40, 0, [1], REF_MJD)
        with pytest.raises(capi.SfxgError):  # the frames end before the slice
            ctx.push_vdif(0, p.value, (len(v) // (PAYLOAD + 32) // 2) * (PAYLOAD + 32), PAYLOAD + 32, 0, [1],
                          REF_MJD)
        with pytest.raises(capi.SfxgError):  # no frame of that thread
            ctx.push_vdif(0, p.value, len(v), PAYLOAD + 32, 0, [8], REF_MJD)
    finally:
        L.hipFree(p)


@pytest.mark.parametrize("bits,frames_per_integration", [(2, 8), (2, 25), (1, 6)])
def test_vdif_phasecal(ctx, bits, frames_per_integration):
    """Phase-cal sums of the raw samples on the GPU (sfxg_vdif_phasecal)
    against the oracle's do_phasecal: integer sums, identical bytes."""
    print(f"case phasecal bits={bits} frames_per_integration={frames_per_integration}", flush=True)
    raw = {}
    spf = PAYLOAD * 8 // bits
    job, _ = configs.build_slices(seed=73, n_stations=2, pols=["R", "L"], fs=FS, bw=FS // 2, nchan=512,
                                  seconds=0.0625, t_int_usec=62500, rho=0.05, bits=bits, sky_freq=5_000_000_000,
                                  offsets=[0.0, 1e-6], rates=[0.0, 0.0], raw_out=raw, margin_multiple=spf)[0]
    idx = [i for i, s in enumerate(job.station_streams) if s.station_number == job.station_streams[0].station_number]
    v = synth.vdif_stream([raw["raw"][i] for i in idx], [4, 6], raw["raw_start"], FS, bits, PAYLOAD, EPOCH,
                          invalid=[(1, 11), (0, 30)], fill=[2])
    T = frames_per_integration * spf * (512_000_000 // FS)
    L, p = to_device(v)
    try:
        ctx.set_options(lut=LUT_CONFIG)
        ctx.set_parameters(job)
        for c, i in enumerate(idx):
            got = ctx.vdif_phasecal(i, p.value, len(v), PAYLOAD + 32, c, [4, 6], REF_MJD, T)
            ref = O.vdif_phasecal(v, PAYLOAD + 32, [4, 6], c, job, i, REF_MJD, T)
            assert len(ref) > 0
            assert got == ref
    finally:
        L.hipFree(p)
