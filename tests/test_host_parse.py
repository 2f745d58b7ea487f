"""Host side of libsfxg on the CPU: the HEADER-stream parser and chunk
bookkeeping (sfxc-1_amd/csrc/sfxg_host.cpp) against the oracle, and the
exported C ABI."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from sfxg import capi, configs
from test_oracle import b2f_params, random_stream

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pc():
    d = tempfile.mkdtemp(prefix="sfxg_pc_")
    so = os.path.join(d, "parse_check.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(ROOT, "tests/cpp/parse_check.cpp"),
                    os.path.join(ROOT, "sfxc-1_amd/csrc/sfxg_host.cpp"), "-o", so], check=True)
    return C.CDLL(so)


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def parse(pc, hdr, p, elided=False):
    p8 = np.array([p["bits_per_sample"], p["sample_rate"], p["base_sample_rate"], p["tsys_freq"], p["fft_size"],
                   p["nfft_max"], p["n_ffts_per_integration"], p["delay_in_samples"]], dtype=np.int32)
    mr = 1 << 16
    runs = np.zeros(3 * mr, dtype=np.int64)
    cnt = np.zeros(3 * mr, dtype=np.int64)
    inv = np.zeros(2 * mr, dtype=np.int32)
    ch = np.zeros(mr, dtype=np.int32)
    sizes = np.zeros(5, dtype=np.int64)
    tsys = np.zeros(3, dtype=np.int32)
    ninv = np.zeros(1, dtype=np.int64)
    err = C.create_string_buffer(256)
    buf = np.frombuffer(hdr, dtype=np.uint8)
    rc = pc.pc_parse(ptr(buf, C.c_uint8), C.c_size_t(len(buf)), int(elided), ptr(p8, C.c_int32),
                     ptr(runs, C.c_int64), C.c_int64(mr), ptr(cnt, C.c_int64), C.c_int64(mr), ptr(inv, C.c_int32),
                     C.c_int64(mr), ptr(ch, C.c_int32), C.c_int64(mr), ptr(sizes, C.c_int64), ptr(tsys, C.c_int32),
                     ptr(ninv, C.c_int64), err, 256)
    if rc != 0:
        raise RuntimeError(f"parse failed {rc}: {err.value}")
    n = sizes
    return dict(runs=runs[:3 * n[0]].reshape(-1, 3), counted=cnt[:3 * n[1]].reshape(-1, 3),
                invalid=inv[:2 * n[2]].reshape(-1, 2), chunks=ch[:n[3]], payload_bytes=int(n[4]), tsys=tsys,
                n_invalid=int(ninv[0]))


def payload_of(hdr):
    out = bytearray()
    i = 0
    while i < len(hdr):
        h = hdr[i]
        if h == 0:
            n = hdr[i + 1] | (hdr[i + 2] << 8)
            out += hdr[i + 3:i + 3 + n]
            i += 3 + n
        elif h == 2:
            i += 2
        elif h == 1:
            i += 3
        else:
            i += 1
    return np.frombuffer(bytes(out), dtype=np.uint8)


def expand(plan, payload, bits, lut, n_out):
    spb = 8 // bits
    out = np.full(n_out, np.nan)
    for o, src, cnt in plan["runs"]:
        if src < 0:
            out[o:o + cnt] = 0.0
        else:
            k = np.arange(src, src + cnt)
            codes = (payload[k // spb] >> ((k % spb) * bits)) & ((1 << bits) - 1)
            out[o:o + cnt] = np.asarray(lut)[codes]
    return out


def counted_hist(plan, payload):
    """Histograms (on, off) of the counted bytes with the tsys phase of each
    ordinal: the initial phase for count0 bytes, then alternating blocks of
    `half` bytes."""
    c0, on0, half = (int(v) for v in plan["tsys"])
    on = np.zeros(256, dtype=np.int64)
    off = np.zeros(256, dtype=np.int64)
    for start, n, ordv in plan["counted"]:
        g = ordv + np.arange(n)
        b = payload[start:start + n]
        st = np.full(n, on0, dtype=bool)
        if half > 0 and c0 > 0:
            k = (g - c0) // half
            st = np.where(g < c0, bool(on0), np.where((k + 1) % 2 == 1, not on0, bool(on0)))
        on += np.bincount(b[st], minlength=256)
        off += np.bincount(b[~st], minlength=256)
    return on, off


@pytest.mark.parametrize("bits", [1, 2])
@pytest.mark.parametrize("seed", range(8))
def test_parse_random_streams(pc, bits, seed):
    rng = np.random.default_rng(100 + seed)
    fft_size, nfft_max, n_ffts = 16, 3, 40
    n_out = fft_size * n_ffts
    hdr, payload = random_stream(rng, bits, n_out)
    p = b2f_params(bits, n_ffts, nfft_max=nfft_max, fft_size=fft_size, rate=4000, tsys=100,
                   delay_in_samples=int(rng.integers(0, 40)))
    ref = O.bit2float(hdr, p, "f64")
    plan = parse(pc, hdr, p)
    lut = [-7, -2, 2, 7] if bits == 2 else [-5, 5]
    np.testing.assert_array_equal(expand(plan, np.frombuffer(payload, np.uint8), bits, lut, n_out), ref["samples"])
    on, off = counted_hist(plan, np.frombuffer(payload, np.uint8))
    if bits == 1:
        on, off = on + off, np.zeros(256, dtype=np.int64)
    np.testing.assert_array_equal(on, ref["counts_on"])
    np.testing.assert_array_equal(off, ref["counts_off"])
    np.testing.assert_array_equal(plan["chunks"], ref["chunks"])
    np.testing.assert_array_equal(plan["invalid"][:, 0], ref["invalid_start"])
    np.testing.assert_array_equal(plan["invalid"][:, 1], ref["invalid_n"])
    assert plan["n_invalid"] == ref["n_invalid"]
    # the elided form (payload supplied separately) parses to the same tables
    ctrl = bytearray()
    i = 0
    while i < len(hdr):
        h = hdr[i]
        if h == 0:
            n = hdr[i + 1] | (hdr[i + 2] << 8)
            ctrl += hdr[i:i + 3]
            i += 3 + n
        elif h in (1,):
            ctrl += hdr[i:i + 3]
            i += 3
        elif h == 2:
            ctrl += hdr[i:i + 2]
            i += 2
        else:
            ctrl.append(h)
            i += 1
    plan_e = parse(pc, bytes(ctrl), p, elided=True)
    np.testing.assert_array_equal(plan_e["runs"], plan["runs"])
    np.testing.assert_array_equal(plan_e["counted"], plan["counted"])


def test_parse_synthetic_slice(pc):
    job, hdrs = configs.config0(seconds=0.0625, t_int_usec=62500)[0]
    for i in range(2):
        p = O.b2f_params(job, i, 1)
        ref = O.bit2float(hdrs[i], p, "f32")
        plan = parse(pc, hdrs[i], p)
        pl = payload_of(hdrs[i])
        got = expand(plan, pl, 2, [-3.3359, -1, 1, 3.3359], len(ref["samples"])).astype(np.float32)
        np.testing.assert_array_equal(got, ref["samples"])
        on, off = counted_hist(plan, pl)
        np.testing.assert_array_equal(on, ref["counts_on"])
        np.testing.assert_array_equal(off, ref["counts_off"])


def test_segments_and_folds(pc):
    job, hdrs = configs.config0(seconds=0.0625, t_int_usec=62500)[0]
    r = O.run_slice(job, hdrs, lut=1)
    chunks = np.ascontiguousarray(r["chunks"], dtype=np.int32)
    out = np.zeros(len(chunks), dtype=np.int32)
    n = pc.pc_segments(ptr(chunks, C.c_int32), len(chunks), 256, C.c_int64(2 * job.fft_size_correlation), 1,
                       ptr(out, C.c_int32))
    np.testing.assert_array_equal(out[:n], r["seg_per_chunk"])
    # folds: every chunk end at or past a multiple of the sub-integration
    # length, plus the integration end
    nsub = 100
    folds = np.zeros(len(chunks), dtype=np.int32)
    nf = pc.pc_folds(ptr(out, C.c_int32), n, job.number_ffts_in_integration(), nsub, ptr(folds, C.c_int32))
    cum = np.cumsum(out[out > 0])
    exp, nxt = [], 1
    for c in cum:
        if c == job.number_ffts_in_integration():
            exp.append(c)
        elif c >= nxt * nsub:
            exp.append(c)
            nxt += 1
    assert list(folds[:nf]) == exp and exp[-1] == job.number_ffts_in_integration()


def _invalid_runs(rng, own_overlaps):
    """A stream's Invalid list: runs in increasing order, adjacent or apart;
    with own_overlaps, a run may start up to 3 samples before the previous
    run's end (the samples a pending sample removal drops at the start of an
    INVALID record, which keeps its announced length)."""
    k = int(rng.integers(0, 7))
    pos = int(rng.integers(0, 30))
    st, nn = [], []
    for _ in range(k):
        n = int(rng.integers(1, 14))
        st.append(pos)
        nn.append(n)
        if own_overlaps and rng.random() < 0.4:
            pos += n - min(n, int(rng.integers(1, 4)))
        else:
            pos += n + int(rng.integers(0, 12))
    return np.array(st, dtype=np.int32), np.array(nn, dtype=np.int32)


def _flagged(fn, a, b):
    out = np.zeros(2, dtype=np.int64)
    fn(ptr(a[0], C.c_int32), ptr(a[1], C.c_int32), C.c_int32(len(a[0])), ptr(b[0], C.c_int32),
       ptr(b[1], C.c_int32), C.c_int32(len(b[0])), C.c_double(1.0), C.c_double(1.0), ptr(out, C.c_int64))
    return out


def _mask(runs, n=400):
    m = np.zeros(n, dtype=bool)
    for s, k in zip(*runs):
        m[s:s + k] = True
    return m


def test_find_invalid_counts_samples_one_stream_lacks(pc):
    """Per-sample brute force: with disjoint increasing runs (every list
    bit2float builds without a sample removal pending at an INVALID record),
    the samples flagged for stream 1 are those invalid in stream 0 and valid in
    stream 1, and the other way round. Runs of the two streams start
    together, end together, nest, touch and partly overlap. Both the library
    and the oracle's restatement of Correlation_core::find_invalid are
    checked against the masks."""
    rng = np.random.default_rng(0)
    oracle_fn = getattr(O.lib("f32"), "or_find_invalid_f32")
    kinds = set()
    for _ in range(3000):
        a, b = _invalid_runs(rng, False), _invalid_runs(rng, False)
        ma, mb = _mask(a), _mask(b)
        expect = np.array([np.sum(mb & ~ma), np.sum(ma & ~mb)], dtype=np.int64)
        np.testing.assert_array_equal(_flagged(pc.pc_find_invalid, a, b), expect)
        np.testing.assert_array_equal(_flagged(oracle_fn, a, b), expect)
        if np.any(ma & mb) and np.any(ma & ~mb) and np.any(mb & ~ma):
            kinds.add("partial")
        if len(a[0]) and len(b[0]) and a[0][0] == b[0][0]:
            kinds.add("same start")
    assert kinds == {"partial", "same start"}


@pytest.mark.parametrize("seed", range(4))
def test_find_invalid_own_overlaps_match_oracle(pc, seed):
    """Runs of one stream that overlap by the samples a pending removal drops:
    each run is compared only with the other stream's run at hand, so such
    samples can count once per run. The library counts them as the oracle's
    restatement of the reference's walk does."""
    rng = np.random.default_rng(100 + seed)
    oracle_fn = getattr(O.lib("f32"), "or_find_invalid_f32")
    differs = 0
    for _ in range(2000):
        a, b = _invalid_runs(rng, True), _invalid_runs(rng, True)
        got = _flagged(pc.pc_find_invalid, a, b)
        np.testing.assert_array_equal(got, _flagged(oracle_fn, a, b))
        ma, mb = _mask(a), _mask(b)
        differs += int(got[1] != np.sum(ma & ~mb) or got[0] != np.sum(mb & ~ma))
    assert differs > 0  # overlapping runs do leave the mask difference


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "sfxg.h")).read()
    names = set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**(sfxg_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 16
    lib = capi.lib()
    for n in sorted(names):
        assert hasattr(lib, n), n
