"""The output node's end on the CPU: slices of the oracle's integration_write
bytes (one record per pulsar bin) routed into per-bin .cor files by
Output_node_writer (sfxc-1_amd/host/sfxc_output.h) in pieces that split
records and their 4-byte file indices anywhere, then read back with the
struct formats of the reference's print_corfile.py (sfxg/corfile.py)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from sfxg import configs, corfile
from sfxg.job import LUT_CONFIG, Pulsar, ticks_to_mjd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def router():
    d = tempfile.mkdtemp(prefix="sfxg_cor_")
    exe = os.path.join(d, "cor_route")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests/cpp/cor_route.cpp"), "-o", exe], check=True)
    return exe


@pytest.fixture(scope="module")
def pulsar_slices():
    slices = configs.build_slices(seed=60, n_stations=2, pols=["R", "L"], fs=16_000_000, bw=8_000_000, nchan=256,
                                  seconds=0.125, t_int_usec=62500, rho=0.05, bits=2, sky_freq=1_400_000_000,
                                  offsets=[0.0, 1e-6], rates=[0.0, 1e-7])
    psr = Pulsar(nbins=3, gate_start=0.1, gate_stop=0.9, tmid=ticks_to_mjd(slices[0][0].start_time) - 0.001, DM=30.0,
                 ref_phase=0.3, ref_freq=1 / 0.0893, obs_freq=1400.0, coef=[0.0, 0.0])
    return slices, psr, [O.run_slice(job, hdrs, lut=LUT_CONFIG, pulsar=psr) for job, hdrs in slices]


@pytest.mark.parametrize("piece", [1, 1000, 3, 1 << 20])
def test_bins_routed_to_files(router, pulsar_slices, piece, tmp_path):
    slices, psr, refs = pulsar_slices
    nb = psr.nbins + 1
    paths = []
    for k, r in enumerate(refs):
        p = tmp_path / f"slice{k}.bin"
        p.write_bytes(r["output"])
        paths.append(str(p))
    job0 = slices[0][0]
    prefix = str(tmp_path / "out")
    subprocess.run([router, prefix, str(piece), str(nb), str(job0.number_channels)] + paths, check=True)
    for b in range(nb):
        cor = corfile.read_cor(f"{prefix}.{b}.cor")
        assert cor["header"]["header_size"] == 84 and cor["header"]["experiment"] == "N26X1"
        assert cor["header"]["number_channels"] == job0.number_channels
        assert len(cor["slices"]) == len(slices)
        raw = open(f"{prefix}.{b}.cor", "rb").read()[84:]
        rec = len(refs[0]["output"]) // nb
        assert raw == b"".join(r["output"][b * rec + 4:(b + 1) * rec] for r in refs)
        for k, (job, _) in enumerate(slices):
            s = cor["slices"][k]
            assert s["integration_slice"] == job.integration_nr
            bl = job.baselines()
            assert len(s["baselines"]) == len(bl)
            chan = refs[k]["channels"][b]
            for i, (x, y) in enumerate(bl):
                got = s["baselines"][i]
                assert (got["station1"], got["station2"]) == (job.station_streams[x].station_number,
                                                              job.station_streams[y].station_number)
                assert got["pol1"] == (job.station_streams[x].polarisation == "L")
                assert got["weight"] == refs[k]["weights"][i]
                np.testing.assert_array_equal(got["vis"], (chan[i, :, 0] + 1j * chan[i, :, 1]).astype(np.complex64))
            assert [st["levels"] for st in s["statistics"]] == [
                O.stats_header_levels(job.station_streams[i].bits_per_sample, refs[k]["stats"][i][0])
                for i in range(len(job.station_streams))]


def test_router_refuses_ragged_slices(router, tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(b"\0" * 103)
    r = subprocess.run([router, str(tmp_path / "x"), "10", "4", "16", str(p)], capture_output=True, text=True)
    assert r.returncode == 1 and "whole records" in r.stderr
