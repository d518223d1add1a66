"""The C-ABI library on the CPU: it loads, exports every entry point that
include/soapfilm.h declares, rejects invalid contexts before touching a GPU,
and its host-only functions (parameter derivation, synthetic fields, real
spherical harmonics) meet their definitions."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import soapbubble_amd as sb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "soapfilm.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = ctypes.CDLL(sb.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(sb.EXPORTS) == declared


def test_binding_loads_all_signatures():
    lib = sb.load_library()
    for name in sb.EXPORTS:
        assert getattr(lib, name).argtypes is not None, name


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    mod = importlib.import_module("soapbubble_amd")
    saved = mod._lib
    mod._lib = None
    try:
        with pytest.raises(RuntimeError, match="not built"):
            mod.load_library(str(tmp_path / "libsoapfilm.so"))
    finally:
        mod._lib = saved


@pytest.mark.parametrize("n_theta,code", [(3, 3), (2, 3), (0, 3), (2049, 3)])
def test_create_rejects_bad_grid(n_theta, code):
    """n_theta from 4 to 8192 whose row of 2 n_theta cells splits into at
    most 1024 equal parts of 2, 4, 8 or 16 cells (2049: rows of 4098 cells,
    2049 parts of 2, not divisible by 4)."""
    with pytest.raises(sb.SoapFilmError) as e:
        sb.SoapFilm(sb.bubble(n_theta, 0.01))
    assert e.value.code == code
    assert "n_theta must be from 4 to 8192" in str(e.value)


def test_create_rejects_bad_parameters():
    with pytest.raises(sb.SoapFilmError) as e:
        sb.SoapFilm(sb.bubble(16, -0.01))
    assert e.value.code == 1
    with pytest.raises(sb.SoapFilmError) as e:
        sb.SoapFilm([sb.bubble(16, 0.01), sb.bubble(16, 0.01, T=300.0)])
    assert e.value.code == 3


def test_null_arguments_return_status():
    lib = sb.load_library()
    out = ctypes.c_void_p()
    assert lib.sf_create(None, 1, None, ctypes.byref(out)) == 1
    assert lib.sf_step(None, 1) == 1
    assert lib.sf_destroy(None) == 0
    assert lib.sf_last_error(None)


def test_derive_params_match_oracle_and_closed_forms(oracle):
    for kw in ({}, {"radius": 2.0, "H": 5e-7, "U": 0.5}, {"evaporation": True, "gravity": True}):
        d = sb.derive_params(sb.bubble(32, 0.004, **kw))
        o = oracle.params(32, 0.004, **kw)
        for k in ("M", "Cr", "gs", "gamma_max", "eva", "dt"):
            assert d[k] == getattr(o, k), (kw, k)
    d = sb.derive_params(sb.bubble(64, 0.005))
    assert abs(d["M"] - 0.826319) < 5e-7
    assert abs(d["Cr"] - 4.679422) < 5e-7
    assert abs(d["gamma_max"] - 29.347029) < 5e-7
    assert abs(d["h"] - math.pi / 64) < 1e-16
    assert sb.derive_params(sb.bubble(64, 0.005, M_scale=1.5))["M"] == pytest.approx(1.5 * d["M"])


def test_real_ylm_orthonormal():
    """Gauss-Legendre in cos(theta) times the trapezoid rule in phi
    integrates the products of degree <= 8 exactly."""
    x, w = np.polynomial.legendre.leggauss(12)
    th = np.arccos(x)
    nph = 24
    ph = np.arange(nph) * 2 * math.pi / nph
    modes = [(l, m) for l in range(5) for m in range(-l, l + 1)]
    Y = np.array([[[sb.real_ylm(l, m, t, p) for p in ph] for t in th] for l, m in modes])
    W = w[:, None] * (2 * math.pi / nph)
    G = np.einsum("atp,btp,tp->ab", Y, Y, np.broadcast_to(W, Y.shape[1:]))
    assert np.abs(G - np.eye(len(modes))).max() < 1e-12


def test_real_ylm_closed_forms():
    t, p = 0.7, 1.9
    assert sb.real_ylm(0, 0, t, p) == pytest.approx(0.5 / math.sqrt(math.pi), abs=1e-15)
    assert sb.real_ylm(1, 0, t, p) == pytest.approx(math.sqrt(3 / (4 * math.pi)) * math.cos(t), abs=1e-15)
    c = math.sqrt(3 / (4 * math.pi)) * math.sin(t)
    assert abs(sb.real_ylm(1, 1, t, p)) == pytest.approx(c * abs(math.cos(p)), abs=1e-15)
    assert abs(sb.real_ylm(1, -1, t, p)) == pytest.approx(c * abs(math.sin(p)), abs=1e-15)


@pytest.mark.parametrize("n", [16, 64])
def test_synth_fields_properties(n):
    nph = 2 * n
    h = math.pi / n
    f = sb.synth_fields(n, 77, l_max=6, l_air=8, noise=0.0, air_amp=1.0)
    g = sb.synth_fields(n, 77, l_max=6, l_air=8, noise=0.0, air_amp=1.0)
    for k in f:
        assert np.array_equal(f[k], g[k]), k
    assert not np.array_equal(f["eta"], sb.synth_fields(n, 78, 6, 8)["eta"])
    assert abs(np.abs(f["eta"] - 1).max() - 0.1) < 1e-15
    assert abs(np.abs(f["gamma"] - 1).max() - 0.05) < 1e-15
    # discrete MAC divergence of the air field vanishes
    u = f["uair0"].reshape(n, nph)
    v = np.zeros((n + 1, nph))
    v[1:n] = f["vair0"].reshape(n - 1, nph)
    s_face = np.sin(np.arange(n + 1) * h)[:, None]
    s_c = np.sin((np.arange(n) + 0.5) * h)[:, None]
    div = ((np.roll(u, -1, axis=1) - u) + (s_face[1:] * v[1:] - s_face[:-1] * v[:-1])) / (h * s_c)
    umax = max(np.abs(u).max(), np.abs(v).max())
    assert np.abs(div).max() <= 1e-11 * umax / h
    # centre speed (face averages) normalised to air_amp
    up = 0.5 * (u + np.roll(u, -1, axis=1))
    vt = 0.5 * (v[:-1] + v[1:])
    assert abs(np.sqrt(up ** 2 + vt ** 2).max() - 1.0) < 1e-12
    f2 = sb.synth_fields(n, 77, l_max=6, l_air=8, noise=0.0, air_amp=0.4)
    assert np.allclose(f2["uair0"], 0.4 * f["uair0"], rtol=1e-14, atol=1e-16)


def test_synth_noise_bounded():
    n = 32
    a = sb.synth_fields(n, 5, noise=0.0)
    b = sb.synth_fields(n, 5, noise=0.05)
    d = b["eta"] - a["eta"]
    assert np.abs(d).max() <= 0.05 and np.abs(d).max() > 0.045
    assert abs(d.mean()) < 0.01
    assert np.array_equal(a["gamma"], b["gamma"])


@pytest.mark.parametrize("n_theta", [2, 3, 16384, 2049])
def test_create_rejects_unsupported_sizes(n_theta):
    """n_theta from 4 (rows of 8 cells) to 8192 (rows of 16384 cells, one
    block of 1024 threads with 16 cells each), with rows that split into at
    most 1024 equal parts of 2, 4, 8 or 16 cells; the check runs before any
    device work."""
    lib = sb.load_library()
    opt = sb.Options(sb.PRECISIONS["mixed"], 0, 0, 0, 1e-10, 0, 0, 0.0, 1, 0, 0, 1, 0.5, 0.0, 0, 0)
    arr = (sb.BubbleDesc * 1)(sb.bubble(n_theta, 0.01))
    out = ctypes.c_void_p()
    assert lib.sf_create(ctypes.byref(opt), 1, arr, ctypes.byref(out)) == 3
    assert "from 4 to 8192" in lib.sf_last_error(None).decode()


@pytest.mark.parametrize("kw,msg", [({"initial_guess": 3}, "initial_guess"),
                                    ({"max_outer": 17}, "max_outer")])
def test_create_rejects_bad_options(kw, msg):
    """Option checks run before any device work (no GPU needed)."""
    lib = sb.load_library()
    opt = sb.Options(sb.PRECISIONS["mixed"], 0, 0, 0, 1e-10, 0, 0, 0.0, 1, 0, 0, 1, 0.5, 0.0, 0, 0)
    for k, v in kw.items():
        setattr(opt, k, v)
    arr = (sb.BubbleDesc * 1)(sb.bubble(16, 0.01))
    out = ctypes.c_void_p()
    assert lib.sf_create(ctypes.byref(opt), 1, arr, ctypes.byref(out)) == 1
    assert msg in lib.sf_last_error(None).decode()
    with pytest.raises(KeyError):
        sb.SoapFilm(sb.bubble(16, 0.01), "mixed", initial_guess="previous")


@pytest.mark.parametrize("n_theta,kw,code,msg", [
    (512, {}, 3, "at most 262144 cells"),
    (64, {"precision": "fp64"}, 3, "SF_PREC_MIXED"),
    (64, {"scalar_advection": "bimocq"}, 3, "plain scalar advection"),
    (3, {}, 3, "n_theta must be from 4"),
])
def test_step_noalloc_rejects_before_device_work(n_theta, kw, code, msg):
    """The entry point without allocation checks its sizes and options before
    it touches a device (no GPU needed): bubbles above 2^18 cells, other
    precisions than mixed, BIMOCQ."""
    opt = sb.options(kw.get("precision", "mixed"),
                     scalar_advection=kw.get("scalar_advection", "plain"))
    with pytest.raises(sb.SoapFilmError) as e:
        sb.step_noalloc(sb.bubble(n_theta, 0.01), opt, 0.0, 1, [1, 1, 1, 1], [1, 1],
                        [1, 1, 1, 1])
    assert e.value.args[0].startswith(sb.STATUS[code]), e.value.args
    assert msg in e.value.args[0]


def test_step_noalloc_rejects_null_fields():
    with pytest.raises(sb.SoapFilmError) as e:
        sb.step_noalloc(sb.bubble(16, 0.01), sb.options("mixed"), 0.0, 1, [1, 0, 1, 1], [1, 1],
                        [1, 1, 1, 1])
    assert "null field pointer" in e.value.args[0]
