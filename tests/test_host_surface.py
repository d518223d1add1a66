"""The host surface on the CPU: init textures (initThicknessfromPic /
initAirflowfromPic, initializer.cu:50-135) through sf_texture_fields, and
the soapBubble driver's argument, config.txt and AMG-config handling up to
the point where it needs a GPU."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import soapbubble_amd as sb
from exr_reader import read_exr
from exr_writer import write_exr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "soapbubble_amd", "soapBubble")


def _resize_linear(img, w, h):
    """Independent restatement of a bilinear resize with aligned pixel
    centres and replicated borders, vectorised over the image."""
    H, W = img.shape

    def taps(n_out, n_in):
        f = (np.arange(n_out) + 0.5) * n_in / n_out - 0.5
        i0 = np.floor(f).astype(int)
        a = f - i0
        lo = i0 < 0
        i0 = np.where(lo, 0, i0)
        a = np.where(lo, 0.0, a)
        i1 = np.minimum(i0 + 1, n_in - 1)
        i0 = np.minimum(i0, n_in - 1)
        a = np.where(i0 == i1, 0.0, a)
        return i0, i1, a

    y0, y1, ay = taps(h, H)
    x0, x1, ax = taps(w, W)
    img = img.astype(np.float64)
    top = img[y0][:, x0] * (1 - ax) + img[y0][:, x1] * ax
    bot = img[y1][:, x0] * (1 - ax) + img[y1][:, x1] * ax
    return (top * (1 - ay)[:, None] + bot * ay[:, None]).astype(np.float32)


def _fields_from(B, G, R, radius):
    """The texture-to-field formulas of initializer.cu:64-67, 120-126."""
    h, w = B.shape
    eta = (B.astype(np.float64) - 0.5) * 0.3 + 0.5
    u = 0.5 * (G + np.roll(G, 1, axis=1)).astype(np.float32).astype(np.float64) * radius
    v = -0.5 * (R[:-1] + R[1:]).astype(np.float32).astype(np.float64) * radius
    return eta.ravel(), u.ravel(), v.ravel()


@pytest.mark.parametrize("src", [(24, 12), (16, 8), (100, 37)])
def test_texture_fields_match_restatement(tmp_path, src):
    W, H = src
    rng = np.random.default_rng(W)
    ch = {c: rng.uniform(-1, 1, (H, W)).astype(np.float16).astype(np.float32) for c in "BGR"}
    path = str(tmp_path / "tex.exr")
    write_exr(path, ch)
    n = 8
    radius = 0.05
    f = sb.texture_fields(n, path, path, radius)
    assert f["loaded"] == 3
    rs = {c: _resize_linear(ch[c], 2 * n, n) for c in "BGR"}
    eta, u, v = _fields_from(rs["B"], rs["G"], rs["R"], radius)
    assert np.abs(f["eta"] - eta).max() <= 1e-7
    assert np.abs(f["uair0"] - u).max() <= 1e-8
    assert np.abs(f["vair0"] - v).max() <= 1e-8
    if (W, H) == (16, 8):   # same size: the resize is the identity
        assert np.array_equal(rs["B"], ch["B"])


def test_texture_fields_missing_files_keep_defaults(tmp_path):
    f = sb.texture_fields(8, str(tmp_path / "none.exr"), None)
    assert f["loaded"] == 0
    assert np.all(f["eta"] == 0.5) and np.all(f["uair0"] == 0) and np.all(f["vair0"] == 0)


REF_INIT = "/root/reference/init"


@pytest.mark.skipif(not os.path.isdir(REF_INIT), reason="reference textures not present")
def test_reference_textures_halving():
    """The reference's 1024 x 512 textures at nTheta = 256: halving with
    aligned centres is the mean of 2 x 2 blocks, read here with the tests'
    own EXR reader."""
    _, chans, img = read_exr(os.path.join(REF_INIT, "midFrequency.exr"))
    f = sb.texture_fields(256, os.path.join(REF_INIT, "midFrequency.exr"),
                          os.path.join(REF_INIT, "velocityfield_512.exr"), 0.05)
    assert f["loaded"] == 3
    B = img["B"].astype(np.float64)
    Bh = 0.25 * (B[0::2, 0::2] + B[0::2, 1::2] + B[1::2, 0::2] + B[1::2, 1::2])
    assert np.abs(f["eta"] - ((Bh - 0.5) * 0.3 + 0.5).ravel()).max() <= 1e-6
    _, _, air = read_exr(os.path.join(REF_INIT, "velocityfield_512.exr"))
    G = air["G"].astype(np.float64)
    Gh = 0.25 * (G[0::2, 0::2] + G[0::2, 1::2] + G[1::2, 0::2] + G[1::2, 1::2])
    u = 0.5 * (Gh + np.roll(Gh, 1, axis=1)) * 0.05
    assert np.abs(f["uair0"] - u.ravel()).max() <= 1e-7


# ------------------------------------------------------------ the driver ----
def _run(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def _config(tmp_path, n_theta=32, **kw):
    vals = dict(radius=0.05, H=1e-6, U=1, cm=0.5, Gammam=1e-6, T=298.15, Ds=0.0, rm=5e-9,
                numberOfGridsInTheta=n_theta, timePerStep=0.005, timePerFrame=0.012,
                frameNumbers=3, outputDir=str(tmp_path / "out"), thicknessInit="null",
                AMGconfig=str(tmp_path / "amg.json"), blendCoeff=0.5, gpu=0)
    vals.update(kw)
    path = tmp_path / "config.txt"
    path.write_text("".join(f"{k} {v}\n" for k, v in vals.items()))
    return str(path)


def _amg(tmp_path, tol="1e-07", conv="ABSOLUTE", iters=100):
    (tmp_path / "amg.json").write_text(
        '{"config_version": 2, "solver": {"preconditioner": {"solver": "AMG", "max_iters": 1},'
        f' "solver": "PCGF", "max_iters": {iters}, "convergence": "{conv}",'
        f' "tolerance": {tol}, "norm": "L2"}}}}')


def test_cli_usage():
    r = _run()
    assert r.returncode == 255 and "config.txt" in r.stdout


def test_cli_malformed_config(tmp_path):
    p = tmp_path / "config.txt"
    p.write_text("radius 0.05\nH 0.000001\nU one\n")
    r = _run(str(p))
    assert r.returncode == 2 and "U is missing or malformed" in r.stderr
    r = _run(str(tmp_path / "absent.txt"))
    assert r.returncode == 2 and "cannot open" in r.stderr


def test_cli_rejects_grid_before_gpu(tmp_path):
    _amg(tmp_path, tol="1e-09", conv="RELATIVE_INI", iters=37)
    r = _run(_config(tmp_path, n_theta=2049))
    assert r.returncode == 3 and "n_theta must be from 4" in r.stderr
    # the outer solver's settings, not the preconditioner's max_iters = 1
    assert "tolerance 1e-09 RELATIVE, max_iters 37" in r.stdout
    r = _run(_config(tmp_path), "--precision", "single")
    assert r.returncode == 2 and "--precision fp64|mixed" in r.stderr
    r = _run(_config(tmp_path), "--precision")
    assert r.returncode == 255


# ------------------------------------------------------------ EXR output ----
def _half_away(x):
    """binary32 -> binary16 bits, round to nearest with ties away from zero
    (the reference writer's rule), binary32 subnormals to zero."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.int64)
    sign = (b >> 16) & 0x8000
    e = (b >> 23) & 0xFF
    m = b & 0x7FFFFF
    out = np.zeros_like(b)
    for k in range(b.size):
        ek, mk, s = int(e.flat[k]), int(m.flat[k]), int(sign.flat[k])
        if ek == 0:
            r = 0
        elif ek == 255:
            r = 0x7C00 | (0x200 if mk else 0)
        else:
            ne = ek - 127 + 15
            if ne >= 31:
                r = 0x7C00
            elif ne <= 0:
                sh = 14 - ne
                mm = mk | 0x800000
                r = 0 if sh > 24 else (mm >> sh) + ((mm >> (sh - 1)) & 1)
            else:
                r = ((ne << 10) | (mk >> 13)) + (1 if mk & 0x1000 else 0)
        out.flat[k] = s | r
    return out.astype(np.uint16)


def _raw_half_plane(path, w, h, ch_index):
    data = open(path, "rb").read()
    hdr_end = data.index(b"screenWindowWidth\0float\0") + len(b"screenWindowWidth\0float\0") + 8 + 1
    out = np.empty((h, w), np.uint16)
    for y in range(h):
        off = hdr_end + 8 * h + y * (8 + 6 * w) + 8 + ch_index * 2 * w
        out[y] = np.frombuffer(data, "<u2", w, off)
    return out


def test_write_image_bytes_match_layout(tmp_path):
    """Away from rounding ties the file is byte-identical to the tests'
    independent writer of the published layout."""
    h, w = 6, 10
    rng = np.random.default_rng(1)
    r, g, b = (rng.uniform(-2, 2, (h, w)).astype(np.float32) for _ in range(3))
    p1, p2 = str(tmp_path / "a.exr"), str(tmp_path / "b.exr")
    sb.write_image_exr(p1, r, g, b)
    write_exr(p2, {"R": r, "G": g, "B": b})
    assert open(p1, "rb").read() == open(p2, "rb").read()
    attrs, chans, img = read_exr(p1)
    assert chans == ["B", "G", "R"]
    assert np.frombuffer(attrs["screenWindowWidth"][1], "<f4")[0] == w


def test_write_image_half_rounding(tmp_path):
    ties = np.float32([1 + 2 ** -11, -(1 + 2 ** -11), 1 + 3 * 2 ** -11, 2 ** -15 + 2 ** -25,
                       2 ** -14 * 0.75, 65520.0, 70000.0, 1e-40, np.nan, np.inf, 0.1, 3.14159])
    plane = np.tile(ties, (2, 1))
    p = str(tmp_path / "t.exr")
    sb.write_image_exr(p, plane, plane, plane)
    got = _raw_half_plane(p, plane.shape[1], 2, 0)[0]
    assert np.array_equal(got, _half_away(ties)), (got, _half_away(ties))
    # ties round away from zero where round-to-even would not
    assert got[0] == 0x3C01 and np.float16(ties[0]).view(np.uint16) == 0x3C00


@pytest.mark.parametrize("gray", [True, False])
def test_write_image_large_scanlines(tmp_path, gray):
    """An image large enough that its scanlines are encoded by several host
    threads (1201 rows: an uneven split), one plane for all three channels
    (converted once) or three planes: every channel of every scanline holds
    the ties-away halves of its plane, and the offset table points at each
    scanline."""
    h, w = 1201, 1024
    rng = np.random.default_rng(7)
    planes = [rng.uniform(-3, 3, (h, w)).astype(np.float32) for _ in range(3)]
    if gray:
        planes = [planes[0]] * 3
    r, g, b = planes
    p = str(tmp_path / "big.exr")
    sb.write_image_exr(p, r, g, b)
    for k, plane in enumerate((b, g, r)):
        got = _raw_half_plane(p, w, h, k)
        assert np.array_equal(got, _half_away(plane.ravel()).reshape(h, w)), k
    data = open(p, "rb").read()
    hdr_end = len(data) - h * (8 + 6 * w) - 8 * h
    offs = np.frombuffer(data, "<u8", h, hdr_end)
    assert np.array_equal(offs, hdr_end + 8 * h + np.arange(h) * (8 + 6 * w))
    ys = [int(np.frombuffer(data, "<i4", 1, int(o))[0]) for o in offs[::97]]
    assert ys == list(range(0, h, 97))


@pytest.mark.parametrize("h,w", [(3, 8), (1201, 1024)])
def test_write_half_matches_float_writer(tmp_path, h, w):
    """sf_write_half_exr writes the bytes sf_write_image_exr writes for the
    gray plane whose halves it is given (1201 rows: more scanlines than one
    writev batch holds)."""
    rng = np.random.default_rng(5)
    plane = rng.uniform(-3, 3, (h, w)).astype(np.float32)
    plane.ravel()[:4] = [1 + 2 ** -11, 70000.0, 1e-40, 0.0]
    p1, p2 = str(tmp_path / "f.exr"), str(tmp_path / "h.exr")
    sb.write_image_exr(p1, plane, plane, plane)
    sb.write_half_exr(p2, _half_away(plane.ravel()).reshape(h, w))
    assert open(p1, "rb").read() == open(p2, "rb").read()


def test_write_half_rejects_bad_arguments(tmp_path):
    lib = sb.load_library()
    g = np.zeros((2, 2), np.uint16)
    ptr = g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
    assert lib.sf_write_half_exr(str(tmp_path / "x.exr").encode(), 0, 2, ptr) == 1
    assert lib.sf_write_half_exr(str(tmp_path / "x.exr").encode(), 2, 2, None) == 1
    assert lib.sf_write_half_exr(str(tmp_path / "no" / "dir" / "x.exr").encode(), 2, 2, ptr) == 1
