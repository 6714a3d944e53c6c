"""NIfTI-1 (.nii.gz) and _displacements.dat I/O with the reference's byte layout.

The reference reads a 352-byte header and takes dims from offset 40 and the
datatype / bitpix from offset 70 (imageIOgzType.h:7-105); it writes the input
header back with dims and datatype patched (imageIOgzType.h:181-263).  The
displacement file is raw float32 [u; v; w] over the control grid
(imageIOgzType.h:267-276, deedsBCV0.cpp:341-349).  Arrays here are (o, n, m)
in C order, i.e. the reference's voxel order.
"""
from __future__ import annotations

import gzip
import struct

import numpy as np

_DT = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 512: np.uint16}
_CODE = {np.dtype(v): k for k, v in _DT.items()}


def make_header(shape_mno, dtype, pixdim=(1.0, 1.0, 1.0)):
    m, n, o = shape_mno
    dt = np.dtype(dtype)
    code = _CODE[dt]
    h = bytearray(352)
    struct.pack_into("<i", h, 0, 348)
    struct.pack_into("<8h", h, 40, 3, m, n, o, 1, 1, 1, 1)
    struct.pack_into("<hh", h, 70, code, dt.itemsize * 8)
    struct.pack_into("<8f", h, 76, 1.0, *pixdim, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<f", h, 108, 352.0)
    struct.pack_into("<f", h, 112, 1.0)  # scl_slope
    h[344:348] = b"n+1\x00"
    return bytes(h)


def write_nifti(path, vol, header=None, dtype=None):
    vol = np.asarray(vol)
    if dtype is not None:
        vol = vol.astype(dtype)
    o, n, m = vol.shape
    if header is None:
        header = make_header((m, n, o), vol.dtype)
    else:
        h = bytearray(header[:352])
        struct.pack_into("<5h", h, 40, 3, m, n, o, 1)
        struct.pack_into("<hh", h, 70, _CODE[vol.dtype], vol.dtype.itemsize * 8)
        header = bytes(h)
    with gzip.open(path, "wb", compresslevel=1) as f:
        f.write(header)
        f.write(np.ascontiguousarray(vol).tobytes())


def read_nifti(path):
    """Return (volume (o,n,m) in its stored dtype, 352-byte header)."""
    with gzip.open(path, "rb") as f:
        raw = f.read()
    header = raw[:352]
    dims = struct.unpack_from("<8h", header, 40)
    m, n, o, p = dims[1], dims[2], dims[3], max(dims[4], 1)
    code, bitpix = struct.unpack_from("<hh", header, 70)
    if code not in _DT:
        raise ValueError(f"Datatype {code} not supported")
    dt = np.dtype(_DT[code])
    count = m * n * o * p
    data = np.frombuffer(raw, dtype=dt, count=count, offset=352)
    vol = data.reshape((p, o, n, m))[0] if p > 1 else data.reshape((o, n, m))
    return vol.copy(), header


def write_displacements(path, flow):
    np.ascontiguousarray(flow, np.float32).tofile(path)


def read_displacements(path):
    return np.fromfile(path, dtype=np.float32)
