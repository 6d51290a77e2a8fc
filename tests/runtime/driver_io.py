"""Reader of the runtime driver's output (tests/runtime/runtime_driver.cpp):
records of (int32 name length, name, int64 count, count float64)."""
import struct

import numpy as np


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            head = f.read(4)
            if not head:
                break
            n = struct.unpack("<i", head)[0]
            name = f.read(n).decode()
            count = struct.unpack("<q", f.read(8))[0]
            out[name] = np.frombuffer(f.read(8 * count), dtype="<f8").copy()
    return out
