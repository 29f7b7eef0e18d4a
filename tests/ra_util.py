"""Minimal RA-format reader/writer used by the tests to load the reference's
fixtures and to cross-check the product's C++ RA I/O.

RA layout (src/ra.h:20-58, src/ra.cu:23-105): little-endian u64 magic
0x7961727261776172, flags, eltype, elbyte, size, ndims, dims[ndims], then the
data in column-major order (dims[0] fastest).
"""
import struct

import numpy as np

MAGIC = 0x7961727261776172
_TYPES = {(3, 4): np.float32, (3, 8): np.float64, (4, 8): np.complex64,
          (4, 16): np.complex128, (1, 4): np.int32, (1, 8): np.int64,
          (2, 4): np.uint32, (2, 8): np.uint64}


def read_header(path):
    with open(path, "rb") as f:
        magic, flags, eltype, elbyte, size, ndims = struct.unpack("<6Q", f.read(48))
        dims = struct.unpack("<%dQ" % ndims, f.read(8 * ndims))
    return dict(magic=magic, flags=flags, eltype=eltype, elbyte=elbyte,
                size=size, ndims=ndims, dims=list(dims))


def read(path):
    """Returns an ndarray with numpy shape reversed(dims) (C order), i.e.
    arr[..., i0] with dims[0] fastest -> arr.shape == dims[::-1]."""
    h = read_header(path)
    dt = _TYPES[(h["eltype"], h["elbyte"])]
    with open(path, "rb") as f:
        f.seek(48 + 8 * h["ndims"])
        data = np.frombuffer(f.read(h["size"]), dtype=dt)
    return data.reshape(h["dims"][::-1])


def write(path, arr, dims=None):
    arr = np.ascontiguousarray(arr)
    rev = {v: k for k, v in _TYPES.items()}
    eltype, elbyte = rev[arr.dtype.type]
    if dims is None:
        dims = list(arr.shape[::-1])
    with open(path, "wb") as f:
        f.write(struct.pack("<6Q", MAGIC, 0, eltype, elbyte, arr.nbytes, len(dims)))
        f.write(struct.pack("<%dQ" % len(dims), *dims))
        f.write(arr.tobytes())
