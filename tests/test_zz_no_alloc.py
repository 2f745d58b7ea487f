"""sfxg_vdif_phasecal_static (no device allocation: a static per-GPU scratch
buffer) against sfxg_vdif_phasecal, which allocates its device tables per
call, and against the oracle: two calls in a row on the same context, the
second with fewer frames, so the second call runs on the buffer the first
left behind. A call beyond the buffer's limit is refused by the host check
before any device work."""
import ctypes as C

import pytest

from oracle import oracle as O
from sfxg import capi, configs, synth
from sfxg.job import LUT_CONFIG

from test_gpu_vdif import EPOCH, FS, PAYLOAD, REF_MJD, to_device

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [2, 1])
def test_phasecal_static_matches_allocating(bits):
    print(f"case phasecal_static bits={bits}", flush=True)
    raw = {}
    spf = PAYLOAD * 8 // bits
    job, _ = configs.build_slices(seed=74, n_stations=2, pols=["R", "L"], fs=FS, bw=FS // 2, nchan=512,
                                  seconds=0.0625, t_int_usec=62500, rho=0.05, bits=bits, sky_freq=5_000_000_000,
                                  offsets=[0.0, 1e-6], rates=[0.0, 0.0], raw_out=raw, margin_multiple=spf)[0]
    idx = [i for i, s in enumerate(job.station_streams) if s.station_number == job.station_streams[0].station_number]
    v = synth.vdif_stream([raw["raw"][i] for i in idx], [4, 6], raw["raw_start"], FS, bits, PAYLOAD, EPOCH,
                          invalid=[(1, 11), (0, 30)], fill=[2])
    fb = PAYLOAD + 32
    T = 8 * spf * (512_000_000 // FS)
    ctx = capi.Context(0)
    L, p = to_device(v)
    try:
        ctx.set_options(lut=LUT_CONFIG)
        ctx.set_parameters(job)
        # the second call takes the first 60 % of the frames
        for n_bytes in (len(v), (len(v) // fb) * 6 // 10 * fb):
            for c, i in enumerate(idx):
                alloc = ctx.vdif_phasecal(i, p.value, n_bytes, fb, c, [4, 6], REF_MJD, T)
                static = ctx.vdif_phasecal(i, p.value, n_bytes, fb, c, [4, 6], REF_MJD, T, static=True)
                ref = O.vdif_phasecal(v[:n_bytes], fb, [4, 6], c, job, i, REF_MJD, T)
                assert len(ref) > 0
                assert static == alloc
                assert static == ref
    finally:
        ctx.close()
        L.hipFree(p)


def test_phasecal_static_refuses_past_its_buffer():
    """Header words of 16 bytes per frame: 32-byte frames filling 40 MiB are
    more than the 16 MiB buffer holds, so the call fails with SFXG_ERR_ARG
    and names the limit."""
    print("case phasecal_static limit", flush=True)
    job, _ = configs.build_slices(seed=75, n_stations=2, pols=["R"], fs=FS, bw=FS // 2, nchan=512,
                                  seconds=0.0625, t_int_usec=62500, rho=0.05, bits=2, sky_freq=5_000_000_000,
                                  offsets=[0.0, 1e-6], rates=[0.0, 0.0])[0]
    n = 40 << 20
    L = C.CDLL("libamdhip64.so")
    p = C.c_void_p()
    assert L.hipMalloc(C.byref(p), C.c_size_t(n)) == 0
    ctx = capi.Context(0)
    try:
        ctx.set_options(lut=LUT_CONFIG)
        ctx.set_parameters(job)
        with pytest.raises(capi.SfxgError, match="static scratch"):
            ctx.vdif_phasecal(0, p.value, n, 32, 0, [4], REF_MJD, 512_000_000, static=True)
    finally:
        ctx.close()
        L.hipFree(p)
