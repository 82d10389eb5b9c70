"""Reader/writer for the flat record files used by the golden fixtures.

A record is: 32-byte NUL-padded name, 1-byte kind ('i' int32, 'l' int64,
'd' float64), int64 count, then count little-endian values.  The reference
dump harness (oracle/ref_harness.cpp) writes this format; the fixture
scripts in this folder re-pack selected records into compressed .npz files.
"""
import struct

import numpy as np

_KINDS = {b"i": np.int32, b"l": np.int64, b"d": np.float64}


def read_records(path):
    out = {}
    with open(path, "rb") as fh:
        while True:
            head = fh.read(41)
            if not head:
                break
            if len(head) != 41:
                raise ValueError(f"truncated record header in {path}")
            name = head[:32].rstrip(b"\0").decode()
            kind = head[32:33]
            (n,) = struct.unpack("<q", head[33:41])
            dt = np.dtype(_KINDS[kind])
            buf = fh.read(n * dt.itemsize)
            if len(buf) != n * dt.itemsize:
                raise ValueError(f"truncated record {name} in {path}")
            out[name] = np.frombuffer(buf, dtype=dt).copy()
    return out
