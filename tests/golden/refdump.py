"""Reader for the tagged binary files written by oracle/ref_harness.c.

Each record: u32 name length, name, one type byte ('u' u32, 'f' f32,
'd' f64, 'b' u8), u64 element count, raw little-endian data.
"""
from __future__ import annotations

import struct

import numpy as np

_DT = {"u": np.uint32, "f": np.float32, "d": np.float64, "b": np.uint8}


def read(path) -> dict:
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<I", data, pos)
        pos += 4
        name = data[pos:pos + n].decode()
        pos += n
        t = chr(data[pos])
        pos += 1
        (count,) = struct.unpack_from("<Q", data, pos)
        pos += 8
        dt = np.dtype(_DT[t])
        out[name] = np.frombuffer(data, dtype=dt, count=count, offset=pos).copy()
        pos += count * dt.itemsize
    return out
