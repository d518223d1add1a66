"""Minimal OpenEXR writer for the tests: single-part scanline, uncompressed,
HALF or FLOAT channels written in name order (the published file layout)."""
import struct

import numpy as np


def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def write_exr(path, channels, half=True):
    """channels: dict name -> (h, w) array."""
    names = sorted(channels)
    h, w = next(iter(channels.values())).shape
    ptype = 1 if half else 2
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iiii", ptype, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    hdr = struct.pack("<ii", 20000630, 2)
    hdr += _attr("channels", "chlist", chl)
    hdr += _attr("compression", "compression", b"\0")
    hdr += _attr("dataWindow", "box2i", box)
    hdr += _attr("displayWindow", "box2i", box)
    hdr += _attr("lineOrder", "lineOrder", b"\0")
    hdr += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    hdr += _attr("screenWindowWidth", "float", struct.pack("<f", float(w)))
    hdr += b"\0"
    dt = np.dtype("<f2") if half else np.dtype("<f4")
    line = w * len(names) * dt.itemsize
    table = len(hdr) + 8 * h
    out = bytearray(hdr)
    out += struct.pack(f"<{h}Q", *[table + y * (8 + line) for y in range(h)])
    for y in range(h):
        out += struct.pack("<ii", y, line)
        for n in names:
            out += np.asarray(channels[n][y], dtype=np.float32).astype(dt).tobytes()
    with open(path, "wb") as f:
        f.write(out)
