"""Pins the oracle's raw-VDIF input path (VDIF_reader, vdif_reader.cc:77-222,
and Input_node_data_writer on frame blocks, input_node_data_writer.cc:105-228)
against the VDIF header layout, known dates and the HEADER streams of the
synthetic input producer. CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from sfxg import configs, synth
from sfxg.job import mjd_to_ticks

FS, BITS, PAYLOAD, EPOCH = 32_000_000, 2, 2000, 46  # 8000 samples per frame, 4000 frames per second
FRAME = PAYLOAD + 32
REF_MJD = 60000


def test_header_words_known_answer():
    # VDIF specification: word 0 seconds (bits 0-29), legacy (30), invalid (31);
    # word 1 frame number (0-23), epoch (24-29); word 2 length / 8 (0-23),
    # log2 channels (24-28); word 3 station (0-15), thread (16-25), bits - 1 (26-30)
    h = np.frombuffer(synth.vdif_header(1, 5, 46, 2032, 3, 2, invalid=True), dtype="<u4")
    assert h[0] == 0x80000001 and h[1] == (5 | 46 << 24) and h[2] == 254
    assert h[3] == (0x4A42 | 3 << 16 | 1 << 26)
    assert synth.vdif_epoch_mjd(40) == 58849  # 2020-01-01
    assert synth.vdif_epoch_mjd(46) == 59945  # 2023-01-01


def test_frame_time():
    sec, frame = 10 * 86400 + 3600, 7
    fr = synth.vdif_header(sec, frame, EPOCH, FRAME, 0, BITS) + bytes(PAYLOAD)
    (payload, t, inv), = O.vdif_blocks(fr, FRAME, [0], 0, FS, REF_MJD)
    assert not inv and len(payload) == PAYLOAD
    assert t == mjd_to_ticks(59945 + 10, 3600.0 + frame / 4000.0)


@pytest.fixture(scope="module")
def vdif_job():
    raw = {}
    slices = configs.build_slices(seed=61, n_stations=2, pols=["R", "L"], fs=FS, bw=FS // 2, nchan=512,
                                  seconds=0.0625, t_int_usec=62500, rho=0.05, bits=BITS, sky_freq=5_000_000_000,
                                  offsets=[0.0, 2.5e-6], rates=[0.0, 4e-7], raw_out=raw,
                                  margin_multiple=PAYLOAD * 8 // BITS)
    return slices[0][0], slices[0][1], raw


def samples_of(job, i, hdr):
    return O.bit2float(hdr, O.b2f_params(job, i, 1))


@pytest.mark.parametrize("legacy", [False, True])
def test_writer_matches_input_producer(vdif_job, legacy):
    """One thread, no invalid frames: the VDIF path's HEADER stream decodes to
    the samples of the producer's stream for the same raw bytes (DATA records
    split at frame boundaries only)."""
    job, hdrs, raw = vdif_job
    for i in range(len(job.station_streams)):
        v = synth.vdif_stream([raw["raw"][i]], [5], raw["raw_start"], FS, BITS, PAYLOAD, EPOCH, legacy=legacy)
        fb = PAYLOAD + (16 if legacy else 32)
        h = O.vdif_header_stream(v, fb, [5], 0, job, i, REF_MJD)
        a, b = samples_of(job, i, h), samples_of(job, i, hdrs[i])
        np.testing.assert_array_equal(a["samples"], b["samples"])
        assert a["n_invalid"] == 0
        # records never cross a frame
        k, n_data = 0, []
        while h[k] != 3:
            if h[k] in (0, 1):
                n_data.append(h[k + 1] | h[k + 2] << 8)
                k += 3 + (n_data[-1] if h[k] == 0 else 0)
            else:
                k += 2
        assert max(n_data) <= PAYLOAD


def test_threads_fill_and_invalid(vdif_job):
    """Two threads with a fill frame between them: each channel gets its own
    thread's frames; an invalid frame decodes to zeros over its samples and
    counts them invalid, the rest is unchanged."""
    job, hdrs, raw = vdif_job
    i0, i1 = 0, 1  # the two polarisations of station 0
    bad = 40
    v = synth.vdif_stream([raw["raw"][i0], raw["raw"][i1]], [2, 9], raw["raw_start"], FS, BITS, PAYLOAD, EPOCH,
                          invalid=[(1, bad)], fill=[3])
    ok0 = O.vdif_header_stream(v, FRAME, [2, 9], 0, job, i0, REF_MJD)
    np.testing.assert_array_equal(samples_of(job, i0, ok0)["samples"], samples_of(job, i0, hdrs[i0])["samples"])
    h1 = O.vdif_header_stream(v, FRAME, [2, 9], 1, job, i1, REF_MJD)
    got, ref = samples_of(job, i1, h1), samples_of(job, i1, hdrs[i1])
    diff = np.nonzero(got["samples"] != ref["samples"])[0]
    assert len(diff) > 0 and np.all(got["samples"][diff] == 0)
    frame_samples = PAYLOAD * 8 // BITS
    assert diff[-1] - diff[0] < frame_samples
    assert 0 < got["n_invalid"] <= frame_samples


def test_phasecal_periodic_closed_form(vdif_job):
    """Data periodic in the phase-cal period P: every bin of an integration
    holds its value times the number of valid samples that fold onto it;
    an invalid frame takes its samples out of the count (do_phasecal,
    input_node_data_writer.cc:232-306)."""
    job, _, raw = vdif_job
    P = FS // 10_000                      # 3200 samples
    spb = 8 // BITS
    tps = 512_000_000 // FS
    S = PAYLOAD * spb
    n_frames = len(raw["raw"][0]) // PAYLOAD
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 4, P)
    k = np.arange(n_frames * S)
    pattern = codes[k % P].reshape(-1, spb)
    data = np.zeros(len(pattern), dtype=np.uint8)
    for j in range(spb):
        data |= (pattern[:, j] << (2 * j)).astype(np.uint8)
    T = 8 * S * tps  # integrations of eight frames
    # the first frame at a multiple of T opens the first integration
    f0 = next(f for f in range(n_frames) if (raw["raw_start"] + f * S * tps) % T == 0)
    bad = f0 + 3
    v = synth.vdif_stream([data], [0], raw["raw_start"], FS, BITS, PAYLOAD, EPOCH, invalid=[(0, bad)])
    msgs = O.vdif_phasecal(v, FRAME, [0], 0, job, 0, REF_MJD, T)
    rec = 24 + 4 * P
    assert len(msgs) % rec == 0 and len(msgs) // rec >= 2
    values = np.array([-7, -2, 2, 7])
    for m in range(len(msgs) // rec):
        body = msgs[m * rec:(m + 1) * rec]
        start = int(np.frombuffer(body[4:12], dtype="<i8")[0])
        dur = int(np.frombuffer(body[12:20], dtype="<i8")[0])
        n = int(np.frombuffer(body[20:24], dtype="<i4")[0])
        assert n == P and dur == T and start == raw["raw_start"] + (f0 + 8 * m) * S * tps
        sums = np.frombuffer(body[24:], dtype="<i4")
        frames = range(f0 + 8 * m, f0 + 8 * m + 8)
        pos = np.concatenate([np.arange(j * S, (j + 1) * S) for j, f in enumerate(frames) if f != bad])
        cnt = np.bincount(pos % P, minlength=P)
        # the data at fold position q is codes[(absolute sample) % P]; integrations start on multiples of P
        absq = ((f0 + 8 * m) * S) % P
        np.testing.assert_array_equal(sums, cnt * values[codes[(np.arange(P) + absq) % P]])
