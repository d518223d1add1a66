"""The texture reader on malformed OpenEXR files (CPU only): truncated and
corrupted files are refused (the fields keep their defaults) and never read
outside the file or loop without end. sf_texture_fields reads the thickness
texture's B channel; bit 1 of "loaded" tells whether it was read."""
import struct

import numpy as np
import pytest

import soapbubble_amd as sb
from exr_writer import write_exr

pytestmark = pytest.mark.timeout(60)

N = 4   # n_theta: 4 x 8 fields


def _valid(tmp_path, half=True, w=6, h=3):
    rng = np.random.default_rng(5)
    ch = {c: rng.uniform(0, 1, (h, w)).astype(np.float32) for c in "BGR"}
    p = tmp_path / "ok.exr"
    write_exr(str(p), ch, half=half)
    return p.read_bytes()


def _loaded(tmp_path, data, name="bad.exr"):
    p = tmp_path / name
    p.write_bytes(bytes(data))
    f = sb.texture_fields(N, str(p), None)
    if f["loaded"] & 1:
        assert np.all(np.isfinite(f["eta"]))
    else:
        assert np.all(f["eta"] == 0.5)
    return f["loaded"] & 1


def _attr_size_at(data, name):
    """Offset of the int32 size field of attribute `name`."""
    k = data.index(name.encode() + b"\0")
    k = data.index(b"\0", k + len(name) + 1)   # end of the type name
    return k + 1


def _table_at(data):
    """Offset of the scanline offset table (after the header's final NUL)."""
    k = data.index(b"screenWindowWidth\0float\0")
    return k + len(b"screenWindowWidth\0float\0") + 4 + 4 + 1


@pytest.mark.parametrize("half", [True, False])
def test_valid_file_loads(tmp_path, half):
    assert _loaded(tmp_path, _valid(tmp_path, half)) == 1


@pytest.mark.parametrize("half", [True, False])
def test_every_truncation_is_refused(tmp_path, half):
    data = _valid(tmp_path, half)
    for n in range(len(data)):
        assert _loaded(tmp_path, data[:n]) == 0, n


@pytest.mark.parametrize("attr", ["channels", "compression", "dataWindow", "lineOrder"])
@pytest.mark.parametrize("size", [-1, -16, -(2 ** 31), 2 ** 31 - 1])
def test_bad_attribute_size_is_refused(tmp_path, attr, size):
    """A negative size used to move the read position backwards (an endless
    loop); one past the file's end is refused too."""
    data = bytearray(_valid(tmp_path))
    k = _attr_size_at(bytes(data), attr)
    data[k:k + 4] = struct.pack("<i", size)
    assert _loaded(tmp_path, data) == 0


@pytest.mark.parametrize("attr,size", [("dataWindow", 8), ("dataWindow", 0), ("compression", 0)])
def test_short_attribute_is_refused(tmp_path, attr, size):
    """An attribute shorter than its type: the bytes after it belong to the
    next attribute, and the header no longer parses as the writer meant."""
    data = bytearray(_valid(tmp_path))
    k = _attr_size_at(bytes(data), attr)
    data[k:k + 4] = struct.pack("<i", size)
    assert _loaded(tmp_path, data) == 0


def test_channel_list_cut_inside_an_entry(tmp_path):
    """The channel list's size ends inside an entry: its 16 bytes of type
    and sampling would be read past the attribute."""
    data = bytearray(_valid(tmp_path))
    k = _attr_size_at(bytes(data), "channels")
    size = struct.unpack_from("<i", data, k)[0]
    for cut in (3, 10, size - 8, size - 1):
        d = bytearray(data)
        d[k:k + 4] = struct.pack("<i", cut)
        assert _loaded(tmp_path, d) == 0, cut


def test_channel_without_terminator_in_attribute(tmp_path):
    """A channel name whose NUL lies beyond the attribute."""
    data = bytearray(_valid(tmp_path))
    k = _attr_size_at(bytes(data), "channels") + 4
    data[k + 1] = ord("x")   # "B\0" -> "Bx" joins the next bytes to the name
    assert _loaded(tmp_path, data) == 0


@pytest.mark.parametrize("ptype", [3, -1, 7])
def test_unknown_pixel_type_is_refused(tmp_path, ptype):
    data = bytearray(_valid(tmp_path))
    k = _attr_size_at(bytes(data), "channels") + 4 + 2   # after "B\0"
    data[k:k + 4] = struct.pack("<i", ptype)
    assert _loaded(tmp_path, data) == 0


@pytest.mark.parametrize("box", [(0, 0, 2 ** 31 - 1, 2), (0, 0, 5, 2 ** 31 - 1),
                                 (-2 ** 31, 0, 2 ** 31 - 1, 2), (0, 0, 5, -5),
                                 (0, 0, 600, 300)])
def test_data_window_larger_than_file_is_refused(tmp_path, box):
    """Widths and heights whose pixels could not fit in the file (including
    ones that overflow 32 bits) are refused before any allocation."""
    data = bytearray(_valid(tmp_path))
    k = _attr_size_at(bytes(data), "dataWindow") + 4
    data[k:k + 16] = struct.pack("<iiii", *box)
    assert _loaded(tmp_path, data) == 0


@pytest.mark.parametrize("off", [2 ** 64 - 1, 2 ** 64 - 8, 2 ** 63, 10 ** 9])
def test_scanline_offset_outside_file_is_refused(tmp_path, off):
    """An offset near 2^64 used to wrap the end-of-file check."""
    data = bytearray(_valid(tmp_path))
    t = _table_at(bytes(data))
    first = struct.unpack_from("<Q", data, t)[0]
    assert first == len(data) - 3 * (8 + 6 * 3 * 2)   # the writer's layout
    data[t + 8:t + 16] = struct.pack("<Q", off)
    assert _loaded(tmp_path, data) == 0


@pytest.mark.parametrize("field,value", [(0, -1), (0, 3), (1, -1), (1, 12), (1, 2 ** 31 - 1)])
def test_bad_scanline_header_is_refused(tmp_path, field, value):
    """A scanline's y outside the window, or a byte count that differs from
    the line's size (negative included)."""
    data = bytearray(_valid(tmp_path))
    t = _table_at(bytes(data))
    q = struct.unpack_from("<Q", data, t)[0]
    data[q + 4 * field:q + 4 * field + 4] = struct.pack("<i", value)
    assert _loaded(tmp_path, data) == 0


def test_random_corruption_never_crashes(tmp_path):
    """Seeded byte flips anywhere in the file: each read ends, with the
    texture refused or read into finite fields."""
    data = _valid(tmp_path)
    rng = np.random.default_rng(17)
    for trial in range(400):
        d = bytearray(data)
        for _ in range(int(rng.integers(1, 5))):
            d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
        _loaded(tmp_path, d)
