"""Reader of correlator output (.cor) files: the global header, then time
slices of a timeslice header, uvw coordinates, bit statistics and baselines,
with the struct formats of the reference's own reader
(sfxc/utils/print_corfile.py:13-30,91,107,133: '4i' timeslice, '2i3d' uvw,
'4B5i' statistics, 'i4B' baseline header, then nchan + 1 complex floats) and
the global header of sfxc/include/output_header.h:70-96."""
from __future__ import annotations

import struct
from typing import Dict, List

import numpy as np

GLOBAL_FMT = "<i32s2h5ib15s2i"  # Output_header_global, 84 bytes


def read_global(data: bytes) -> Dict:
    size = struct.unpack_from("<i", data, 0)[0]
    f = struct.unpack_from(GLOBAL_FMT, data, 0) if size >= 84 else None
    if f is None:
        raise ValueError(f"global header of {size} bytes")
    return dict(header_size=size, experiment=f[1].split(b"\0")[0].decode(), start_year=f[2], start_day=f[3],
                start_time=f[4], number_channels=f[5], integration_time=f[6], output_format_version=f[7],
                correlator_version=f[8], polarisation_type=f[9], job_nr=f[11], subjob_nr=f[12])


def read_cor(path: str) -> Dict:
    """Returns dict(global=..., slices=[dict(integration_slice, uvw, statistics, baselines)]);
    baselines: list of dict(weight, station1, station2, pol1, pol2, sideband, freq_nr, vis)."""
    data = open(path, "rb").read()
    g = read_global(data)
    nchan = g["number_channels"]
    o = g["header_size"]
    slices: List[Dict] = []
    while o < len(data):
        if len(data) - o < 16:
            raise ValueError("truncated timeslice header")
        islice, nbl, nuvw, nstat = struct.unpack_from("<4i", data, o)
        o += 16
        uvw = []
        for _ in range(nuvw):
            st, _, u, v, w = struct.unpack_from("<2i3d", data, o)
            uvw.append((st, u, v, w))
            o += 32
        stats = []
        for _ in range(nstat):
            f = struct.unpack_from("<4B5i", data, o)
            stats.append(dict(station=f[0], freq_nr=f[1], sideband=f[2], polarisation=f[3], levels=list(f[4:9])))
            o += 24
        bls = []
        for _ in range(nbl):
            weight, s1, s2, byte, _ = struct.unpack_from("<i4B", data, o)
            o += 8
            vis = np.frombuffer(data, dtype="<c8", count=nchan + 1, offset=o).copy()
            o += 8 * (nchan + 1)
            bls.append(dict(weight=weight, station1=s1, station2=s2, pol1=byte & 1, pol2=(byte >> 1) & 1,
                            sideband=(byte >> 2) & 1, freq_nr=byte >> 3, vis=vis))
        slices.append(dict(integration_slice=islice, uvw=uvw, statistics=stats, baselines=bls))
    return dict(header=g, slices=slices)
