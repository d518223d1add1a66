"""Python binding of libsoapfilm.so, the MI355X soap-film step.

The compute path is the HIP library built from ``soapbubble_amd/csrc`` (see
``include/soapfilm.h``); this module only marshals arrays through its C ABI.
There is no CPU fallback: if the library is missing, importing succeeds but
every call raises ``RuntimeError``.

The surface mirrors the reference's ``Solver`` (include/solver.cuh:107-150):
``step`` is ``Solver::stepForward`` (solver.cu:260-292), ``advect`` /
``assemble`` / ``solve`` / ``update`` are its stages (bubble.cu:1171-1214,
2383-2506), fields use the reference's index convention (row-major by theta,
phi contiguous, nPhi = 2 nTheta).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsoapfilm.so")

FIELDS = {
    "u_theta": 0, "u_phi": 1, "eta": 2, "gamma": 3, "uair0": 4, "vair0": 5,
    "u_theta_star": 6, "u_phi_star": 7, "eta_star": 8, "gamma_star": 9,
    "rhs": 10, "diag": 11, "w_north": 12, "w_west": 13, "solution": 14,
}
PRECISIONS = {"fp64": 0, "mixed": 1, "fp32_adv": 2}
STATUS = {0: "SF_OK", 1: "SF_ERR_INVALID", 2: "SF_ERR_HIP", 3: "SF_ERR_UNSUPPORTED",
          4: "SF_ERR_NOMEM", 5: "SF_ERR_NOT_CONVERGED", 6: "SF_ERR_ORDER"}

# every entry point include/soapfilm.h declares
EXPORTS = (
    "sf_create", "sf_destroy", "sf_upload_field", "sf_download_field", "sf_set_time",
    "sf_set_dt", "sf_step", "sf_advect", "sf_assemble", "sf_solve", "sf_update",
    "sf_synchronize", "sf_time_advection", "sf_time_inner", "sf_get_stats", "sf_get_timing", "sf_reset_timing", "sf_last_error",
    "sf_derive_params", "sf_synth_fields", "sf_texture_fields", "sf_real_ylm",
    "sf_write_thickness_exr", "sf_write_image_exr", "sf_step_noalloc", "sf_get_tuning",
    "sf_thickness_half", "sf_write_half_exr",
)


class BubbleDesc(ctypes.Structure):
    _fields_ = [("n_theta", ctypes.c_int32), ("gravity", ctypes.c_int32),
                ("evaporation", ctypes.c_int32), ("airflow", ctypes.c_int32),
                ("radius", ctypes.c_double), ("H", ctypes.c_double), ("U", ctypes.c_double),
                ("Gamma_m", ctypes.c_double), ("T", ctypes.c_double), ("dt", ctypes.c_double),
                ("M_scale", ctypes.c_double)]


class Derived(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("M", "Cr", "gs", "gamma_max", "eva", "dt",
                                                "radius", "U", "h", "re")]


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("tol_absolute", ctypes.c_int32),
                ("tol", ctypes.c_double), ("max_outer", ctypes.c_int32),
                ("max_inner", ctypes.c_int32), ("inner_rho", ctypes.c_double),
                ("advance_time", ctypes.c_int32), ("inner_method", ctypes.c_int32),
                ("scalar_advection", ctypes.c_int32), ("multisample", ctypes.c_int32),
                ("blend", ctypes.c_double), ("reinit_threshold", ctypes.c_double),
                ("initial_guess", ctypes.c_int32), ("viscous", ctypes.c_int32),
                ("halt_on_failure", ctypes.c_int32), ("autotune", ctypes.c_int32)]


class BubbleStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("outer", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("failed_solves", ctypes.c_int32),
                ("residual", ctypes.c_double), ("max_residual", ctypes.c_double),
                ("volume", ctypes.c_double), ("broken", ctypes.c_int32),
                ("steps", ctypes.c_int32), ("time", ctypes.c_double),
                ("distortion", ctypes.c_double), ("reinits", ctypes.c_int32),
                ("no_film", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("tuned", ctypes.c_int32), ("advect_tile_width", ctypes.c_int32),
                ("nwidths", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("row_width", ctypes.c_int32 * 8), ("inner_cells", ctypes.c_int32 * 8),
                ("advect_ms", ctypes.c_double * 3), ("inner_ms", (ctypes.c_double * 3) * 8)]


class Timing(ctypes.Structure):
    _fields_ = [("advect_ms", ctypes.c_double), ("assemble_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("update_ms", ctypes.c_double),
                ("step_ms", ctypes.c_double), ("steps", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


_lib = None
_load_error = None


def load_library(path: str = LIB_PATH):
    """Load libsoapfilm.so (once) and declare its signatures."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        _load_error = f"{path} is not built (run __graft_entry__.build())"
        raise RuntimeError(_load_error)
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    vp = ctypes.c_void_p
    sig = {
        "sf_create": (ctypes.c_int, [P(Options), ctypes.c_int32, P(BubbleDesc), P(vp)]),
        "sf_destroy": (ctypes.c_int, [vp]),
        "sf_upload_field": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, dp]),
        "sf_download_field": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, dp]),
        "sf_set_time": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_double]),
        "sf_set_dt": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_double]),
        "sf_step": (ctypes.c_int, [vp, ctypes.c_int32]),
        "sf_advect": (ctypes.c_int, [vp]),
        "sf_assemble": (ctypes.c_int, [vp]),
        "sf_solve": (ctypes.c_int, [vp]),
        "sf_update": (ctypes.c_int, [vp]),
        "sf_synchronize": (ctypes.c_int, [vp]),
        "sf_time_advection": (ctypes.c_int, [vp, ctypes.c_int32, P(ctypes.c_double)]),
        "sf_time_inner": (ctypes.c_int, [vp, ctypes.c_int32, P(ctypes.c_double)]),
        "sf_get_stats": (ctypes.c_int, [vp, ctypes.c_int32, P(BubbleStats)]),
        "sf_get_timing": (ctypes.c_int, [vp, P(Timing)]),
        "sf_get_tuning": (ctypes.c_int, [vp, P(Tuning)]),
        "sf_reset_timing": (ctypes.c_int, [vp]),
        "sf_last_error": (ctypes.c_char_p, [vp]),
        "sf_derive_params": (ctypes.c_int, [P(BubbleDesc), P(Derived)]),
        "sf_synth_fields": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                           dp, dp, dp, dp]),
        "sf_texture_fields": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32,
                                             ctypes.c_double, dp, dp, dp, P(ctypes.c_int32)]),
        "sf_real_ylm": (ctypes.c_double, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_double]),
        "sf_write_thickness_exr": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_char_p]),
        "sf_thickness_half": (ctypes.c_int, [vp, ctypes.c_int32, P(ctypes.c_uint16)]),
        "sf_write_half_exr": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                             P(ctypes.c_uint16)]),
        "sf_write_image_exr": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                              P(ctypes.c_float), P(ctypes.c_float),
                                              P(ctypes.c_float)]),
        "sf_step_noalloc": (ctypes.c_int, [P(Options), P(BubbleDesc), ctypes.c_double,
                                           ctypes.c_int32, vp] + [vp] * 10 + [P(BubbleStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class SoapFilmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def bubble(n_theta: int, dt: float, *, radius: float = 1.0, H: float = 1e-6, U: float = 1.0,
           Gamma_m: float = 1e-6, T: float = 298.15, gravity: bool = False,
           evaporation: bool = False, airflow: bool = True, M_scale: float = 1.0) -> BubbleDesc:
    """A bubble descriptor with config.txt-style inputs (defaults: the
    synthetic workloads: radius 1, H 1e-6, U 1, Gamma_m 1e-6, T 298.15)."""
    return BubbleDesc(n_theta, int(gravity), int(evaporation), int(airflow), radius, H, U,
                      Gamma_m, T, dt, M_scale)


def derive_params(desc: BubbleDesc) -> dict:
    lib = load_library()
    out = Derived()
    rc = lib.sf_derive_params(ctypes.byref(desc), ctypes.byref(out))
    if rc:
        raise SoapFilmError(rc, lib.sf_last_error(None).decode())
    return {n: getattr(out, n) for n, _ in Derived._fields_}


def synth_fields(n_theta: int, seed: int, l_max: int = 4, l_air: int = 6, noise: float = 0.0,
                 air_amp: float = 1.0) -> dict:
    """Seeded synthetic eta, Gamma and initial air velocity (host only)."""
    lib = load_library()
    nph = 2 * n_theta
    out = {"eta": np.empty(n_theta * nph), "gamma": np.empty(n_theta * nph),
           "uair0": np.empty(n_theta * nph), "vair0": np.empty((n_theta - 1) * nph)}
    rc = lib.sf_synth_fields(n_theta, seed, l_max, l_air, noise, air_amp, _ptr(out["eta"]),
                             _ptr(out["gamma"]), _ptr(out["uair0"]), _ptr(out["vair0"]))
    if rc:
        raise SoapFilmError(rc, "sf_synth_fields")
    return out


def texture_fields(n_theta: int, thickness_exr: str | None = None, air_exr: str | None = None,
                   radius: float = 1.0) -> dict:
    """Fields from the reference's init textures (host only). Starts from
    the reference's defaults (eta = 0.5, air = 0) where a texture is absent
    or unreadable; "loaded" tells which textures were read."""
    lib = load_library()
    nph = 2 * n_theta
    out = {"eta": np.full(n_theta * nph, 0.5), "uair0": np.zeros(n_theta * nph),
           "vair0": np.zeros((n_theta - 1) * nph)}
    got = ctypes.c_int32(0)
    enc = lambda s: s.encode() if s else None
    rc = lib.sf_texture_fields(enc(thickness_exr), enc(air_exr), n_theta, radius,
                               _ptr(out["eta"]), _ptr(out["uair0"]), _ptr(out["vair0"]),
                               ctypes.byref(got))
    if rc:
        raise SoapFilmError(rc, "sf_texture_fields")
    out["loaded"] = got.value
    return out


def write_image_exr(path: str, r, g, b) -> None:
    """write_image (solver.cu:338-385): three (height, width) planes as a
    HALF B, G, R OpenEXR (host only)."""
    lib = load_library()
    planes = [np.ascontiguousarray(x, dtype=np.float32) for x in (r, g, b)]
    h, w = planes[0].shape
    fp = [x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for x in planes]
    rc = lib.sf_write_image_exr(path.encode(), w, h, *fp)
    if rc:
        raise SoapFilmError(rc, f"cannot write {path}")


def write_half_exr(path: str, gray) -> None:
    """write_image_exr of a gray image given as its binary16 values (a
    (height, width) uint16 array; host only)."""
    lib = load_library()
    g = np.ascontiguousarray(gray, dtype=np.uint16)
    h, w = g.shape
    rc = lib.sf_write_half_exr(path.encode(), w, h, g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    if rc:
        raise SoapFilmError(rc, f"cannot write {path}")


def real_ylm(l: int, m: int, theta: float, phi: float) -> float:
    return load_library().sf_real_ylm(l, m, theta, phi)


@dataclass
class FieldShape:
    n_theta: int

    def size(self, field: str) -> int:
        nph = 2 * self.n_theta
        if field in ("u_theta", "vair0", "u_theta_star"):
            return (self.n_theta - 1) * nph
        return self.n_theta * nph


def options(precision: str = "fp64", *, device: int = 0, use_graph: bool = False,
            tol: float = 1e-10, tol_absolute: bool = False, advance_time: bool = True,
            max_outer: int = 0, max_inner: int = 0, inner_rho: float = 0.0,
            inner_method: str = "auto", scalar_advection: str = "plain", blend: float = 0.5,
            multisample: bool = True, reinit_threshold: float = 0.0,
            initial_guess: str = "auto", viscous: bool = False,
            halt_on_failure: bool = False, autotune: bool = False) -> Options:
    """The sf_options struct of include/soapfilm.h."""
    return Options(PRECISIONS[precision], device, int(use_graph), int(tol_absolute), tol,
                   max_outer, max_inner, inner_rho, int(advance_time),
                   {"auto": 0, "cg": 1, "chebyshev": 2, "mg": 3}[inner_method],
                   {"plain": 0, "bimocq": 1}[scalar_advection], int(multisample), blend,
                   reinit_threshold,
                   {"auto": 0, "gamma_star": 1, "extrapolated": 2}[initial_guess],
                   int(viscous), int(halt_on_failure), int(autotune))


def step_noalloc(bubble: BubbleDesc, opt: Options, t_seconds: float, nsteps: int,
                 fields_in: Sequence[int], air: Sequence[int], fields_out: Sequence[int],
                 stream: int | None = None) -> dict:
    """sf_step_noalloc: nsteps steps of one bubble on device-resident float
    arrays, given as device addresses (u_theta, u_phi, eta, Gamma in and out;
    uair0, vair0). The library allocates nothing. Returns the statistics."""
    lib = load_library()
    st = BubbleStats()
    vp = ctypes.c_void_p
    rc = lib.sf_step_noalloc(ctypes.byref(opt), ctypes.byref(bubble), float(t_seconds), nsteps,
                             vp(stream) if stream else None, *[vp(int(a)) for a in fields_in],
                             *[vp(int(a)) for a in air], *[vp(int(a)) for a in fields_out],
                             ctypes.byref(st))
    if rc:
        raise SoapFilmError(rc, lib.sf_last_error(None).decode())
    return {k: getattr(st, k) for k, _ in BubbleStats._fields_ if k != "pad_"}


class SoapFilm:
    """A context: the packed state of one or more bubbles on one GPU."""

    def __init__(self, bubbles: Sequence[BubbleDesc] | BubbleDesc, precision: str = "fp64",
                 device: int = 0, use_graph: bool = False, tol: float = 1e-10,
                 tol_absolute: bool = False, advance_time: bool = True, max_outer: int = 0,
                 max_inner: int = 0, inner_rho: float = 0.0, inner_method: str = "auto",
                 scalar_advection: str = "plain", blend: float = 0.5, multisample: bool = True,
                 reinit_threshold: float = 0.0, initial_guess: str = "auto",
                 viscous: bool = False, halt_on_failure: bool = False, autotune: bool = False):
        self._lib = load_library()
        if isinstance(bubbles, BubbleDesc):
            bubbles = [bubbles]
        self.bubbles = list(bubbles)
        arr = (BubbleDesc * len(self.bubbles))(*self.bubbles)
        opt = options(precision, device=device, use_graph=use_graph, tol=tol,
                      tol_absolute=tol_absolute, advance_time=advance_time, max_outer=max_outer,
                      max_inner=max_inner, inner_rho=inner_rho, inner_method=inner_method,
                      scalar_advection=scalar_advection, blend=blend, multisample=multisample,
                      reinit_threshold=reinit_threshold, initial_guess=initial_guess,
                      viscous=viscous, halt_on_failure=halt_on_failure, autotune=autotune)
        h = ctypes.c_void_p()
        rc = self._lib.sf_create(ctypes.byref(opt), len(self.bubbles), arr, ctypes.byref(h))
        if rc:
            raise SoapFilmError(rc, self._lib.sf_last_error(None).decode())
        self._h = h
        self.precision = precision

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc:
            raise SoapFilmError(rc, self._lib.sf_last_error(self._h).decode())

    def upload(self, b: int, field: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        need = FieldShape(self.bubbles[b].n_theta).size(field)
        if a.size != need:
            raise ValueError(f"{field} needs {need} values, got {a.size}")
        self._check(self._lib.sf_upload_field(self._h, b, FIELDS[field], _ptr(a)))

    def download(self, b: int, field: str) -> np.ndarray:
        out = np.empty(FieldShape(self.bubbles[b].n_theta).size(field))
        self._check(self._lib.sf_download_field(self._h, b, FIELDS[field], _ptr(out)))
        return out

    def set_time(self, t: float, b: int = -1) -> None:
        self._check(self._lib.sf_set_time(self._h, b, t))

    def set_dt(self, dt_seconds: float, b: int = -1) -> None:
        self._check(self._lib.sf_set_dt(self._h, b, dt_seconds))

    def step(self, n: int = 1) -> None:
        self._check(self._lib.sf_step(self._h, n))

    def advect(self):
        self._check(self._lib.sf_advect(self._h))

    def assemble(self):
        self._check(self._lib.sf_assemble(self._h))

    def solve(self):
        self._check(self._lib.sf_solve(self._h))

    def update(self):
        self._check(self._lib.sf_update(self._h))

    def time_advection(self, nlaunch: int = 20) -> float:
        """Average device time (ms) of one advection launch."""
        out = ctypes.c_double()
        self._check(self._lib.sf_time_advection(self._h, nlaunch, ctypes.byref(out)))
        return out.value

    def time_inner(self, nlaunch: int = 20) -> float:
        """Average device time (ms) of one inner (Chebyshev) iteration launch;
        call between assemble() and solve()."""
        out = ctypes.c_double()
        self._check(self._lib.sf_time_inner(self._h, nlaunch, ctypes.byref(out)))
        return out.value

    def synchronize(self):
        self._check(self._lib.sf_synchronize(self._h))

    def stats(self, b: int = 0) -> dict:
        s = BubbleStats()
        self._check(self._lib.sf_get_stats(self._h, b, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in BubbleStats._fields_}

    def timing(self) -> dict:
        t = Timing()
        self._check(self._lib.sf_get_timing(self._h, ctypes.byref(t)))
        return {n: getattr(t, n) for n, _ in Timing._fields_ if n != "pad_"}

    def tuning(self) -> dict:
        """The autotuner's choices (sf_get_tuning)."""
        t = Tuning()
        self._check(self._lib.sf_get_tuning(self._h, ctypes.byref(t)))
        n = t.nwidths
        return {"tuned": bool(t.tuned), "advect_tile_width": t.advect_tile_width,
                "advect_ms": list(t.advect_ms),
                "rows": [{"n_phi": t.row_width[w], "inner_cells_per_thread": t.inner_cells[w],
                          "inner_ms": list(t.inner_ms[w])} for w in range(n)]}

    def reset_timing(self):
        self._check(self._lib.sf_reset_timing(self._h))

    def write_thickness_exr(self, b: int, path: str):
        self._check(self._lib.sf_write_thickness_exr(self._h, b, path.encode()))

    def thickness_half(self, b: int) -> np.ndarray:
        """The frame's pixels (eta H 2e5) as binary16 values, (n_theta, n_phi)."""
        n = self.bubbles[b].n_theta
        out = np.empty((n, 2 * n), dtype=np.uint16)
        self._check(self._lib.sf_thickness_half(self._h, b,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
        return out

    def load_state(self, b: int, state: dict) -> None:
        for k, v in state.items():
            self.upload(b, k, v)

    def state(self, b: int = 0) -> dict:
        return {k: self.download(b, k) for k in ("u_theta", "u_phi", "eta", "gamma")}
