"""Reader for the record files written by oracle/ref_driver.c.

Record format: char name[32]; int64 n; float64 data[n]; repeated.
"""
import numpy as np


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            head = f.read(32)
            if len(head) < 32:
                break
            name = head.split(b"\0", 1)[0].decode()
            n = int(np.frombuffer(f.read(8), dtype="<i8")[0])
            data = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
            out[name] = data[0] if n == 1 else data
    return out
