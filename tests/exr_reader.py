"""Minimal OpenEXR reader for the tests: single-part scanline files with
uncompressed HALF or FLOAT channels (the layout of the thickness frames).
Written from the published OpenEXR file layout: magic 20000630, version 2,
a header of (name, type, size, value) attributes ended by an empty name, a
line offset table, then per scanline block: y (int32), byte count (int32)
and the channels in header order, each a full row."""
import struct

import numpy as np


def read_exr(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and (version & 0xFF) == 2 and not (version & 0x1E00)
    p = 8
    attrs = {}
    while True:
        e = data.index(b"\0", p)
        name = data[p:e].decode()
        p = e + 1
        if not name:
            break
        e = data.index(b"\0", p)
        typ = data[p:e].decode()
        p = e + 1
        (size,) = struct.unpack_from("<i", data, p)
        p += 4
        attrs[name] = (typ, data[p:p + size])
        p += size
    chans = []
    raw = attrs["channels"][1]
    q = 0
    while raw[q:q + 1] != b"\0":
        e = raw.index(b"\0", q)
        nm = raw[q:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iiii", raw, e + 1)
        chans.append((nm, ptype))
        q = e + 1 + 16
    assert attrs["compression"][1] == b"\0", "only uncompressed files"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", data, p)
    out = {nm: np.empty((h, w), np.float32) for nm, _ in chans}
    for off in offsets:
        y, nbytes = struct.unpack_from("<ii", data, off)
        r = off + 8
        for nm, ptype in chans:
            dt, sz = (np.dtype("<f2"), 2) if ptype == 1 else (np.dtype("<f4"), 4)
            out[nm][y - y0] = np.frombuffer(data, dt, w, r).astype(np.float32)
            r += w * sz
        assert r - off - 8 == nbytes
    return attrs, [c for c, _ in chans], out
