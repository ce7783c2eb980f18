"""Reader for the golden fixture container written by gen_golden.js.

Format (gzip of): b"WSG1", u32 LE header length, JSON header
{meta, arrays: [{name, dtype, offset, nbytes}]}, then the raw arrays.
"""
import gzip
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Golden:
    def __init__(self, name):
        path = os.path.join(HERE, name + ".wsg.gz")
        with gzip.open(path, "rb") as f:
            raw = f.read()
        if raw[:4] != b"WSG1":
            raise ValueError("bad golden magic in %s" % path)
        (hlen,) = struct.unpack("<I", raw[4:8])
        hdr = json.loads(raw[8:8 + hlen].decode())
        self.meta = hdr["meta"]
        base = 8 + hlen
        self._arrays = {}
        for a in hdr["arrays"]:
            off = base + a["offset"]
            self._arrays[a["name"]] = np.frombuffer(
                raw, dtype=np.dtype(a["dtype"]).newbyteorder("<"),
                count=a["nbytes"] // np.dtype(a["dtype"]).itemsize, offset=off).copy()

    def __getitem__(self, k):
        return self._arrays[k]

    def __contains__(self, k):
        return k in self._arrays

    def names(self):
        return list(self._arrays)


_cache = {}


def load(name):
    if name not in _cache:
        _cache[name] = Golden(name)
    return _cache[name]
