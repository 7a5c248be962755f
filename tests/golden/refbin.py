"""Reader for the binary dumps written by oracle/ref_driver.cpp (test infrastructure).

Record layout: NUL-terminated name, one type byte ('d' f64, 'f' f32, 'i' i32,
's' bytes), int64 element count, then the raw little-endian elements.
"""
import struct

import numpy as np

_DT = {b"d": np.float64, b"f": np.float32, b"i": np.int32, b"s": np.uint8}


def read_refbin(path):
    out = {}
    with open(path, "rb") as fh:
        buf = fh.read()
    pos = 0
    while pos < len(buf):
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode()
        t = buf[end + 1:end + 2]
        (n,) = struct.unpack_from("<q", buf, end + 2)
        pos = end + 10
        dt = np.dtype(_DT[t])
        arr = np.frombuffer(buf, dtype=dt, count=n, offset=pos).copy()
        pos += n * dt.itemsize
        out[name] = arr.tobytes().decode() if t == b"s" else arr
    return out
