"""ctypes wrapper of the FP64 CPU oracle (oracle/libsf_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the soap-film step. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product (soapbubble_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libsf_oracle.so")


class Params(ctypes.Structure):
    _fields_ = [("n_theta", ctypes.c_int32), ("gravity", ctypes.c_int32),
                ("evaporation", ctypes.c_int32), ("airflow", ctypes.c_int32),
                ("floorf_mode", ctypes.c_int32), ("tol_absolute", ctypes.c_int32),
                ("max_iters", ctypes.c_int32), ("viscous", ctypes.c_int32),
                ("dt", ctypes.c_double), ("M", ctypes.c_double), ("Cr", ctypes.c_double),
                ("gs", ctypes.c_double), ("gamma_max", ctypes.c_double),
                ("eva", ctypes.c_double), ("radius", ctypes.c_double), ("U", ctypes.c_double),
                ("time_s", ctypes.c_double), ("tol", ctypes.c_double), ("re", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("broken", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("abs_residual", ctypes.c_double),
                ("volume", ctypes.c_double)]


MAPS = ("fwd_t", "fwd_p", "bwd_t", "bwd_p", "bwdprev_t", "bwdprev_p",
        "eta_init", "eta_initlast", "eta_delta", "eta_deltalast",
        "gam_init", "gam_initlast", "gam_delta", "gam_deltalast")


class Bimocq(ctypes.Structure):
    _fields_ = [(k, ctypes.POINTER(ctypes.c_double)) for k in MAPS] + [
        ("blend", ctypes.c_double), ("threshold", ctypes.c_double),
        ("multisample", ctypes.c_int32), ("reinit_count", ctypes.c_int32),
        ("last_distortion", ctypes.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        P = ctypes.POINTER(Params)
        L.sfo_derive.argtypes = [ctypes.c_double] * 6 + [ctypes.c_int, ctypes.c_int, P]
        L.sfo_sample.restype = ctypes.c_double
        L.sfo_sample.argtypes = [P, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double]
        L.sfo_trace.argtypes = [P, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.sfo_advect.argtypes = [P] + [dp] * 8
        L.sfo_airflow.argtypes = [P] + [dp] * 4
        L.sfo_col_ind.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
        L.sfo_assemble.argtypes = [P] + [dp] * 8
        L.sfo_solve.restype = ctypes.c_int
        L.sfo_solve.argtypes = [P, dp, dp, dp, ctypes.POINTER(Stats)]
        L.sfo_spmv.argtypes = [P, dp, dp, dp]
        L.sfo_update.argtypes = [P] + [dp] * 10
        L.sfo_step.restype = ctypes.c_int
        L.sfo_step.argtypes = [P] + [dp] * 6 + [ctypes.POINTER(Stats)]
        L.sfo_visc_div.argtypes = [P, dp, dp, dp]
        L.sfo_visc_forces.argtypes = [P] + [dp] * 6
        L.sfo_volume.restype = ctypes.c_double
        L.sfo_volume.argtypes = [P, dp]
        L.sfo_num_threads.restype = ctypes.c_int
        B = ctypes.POINTER(Bimocq)
        L.sfo_lerp_coords.argtypes = [P, dp, dp, ctypes.c_double, dp]
        L.sfo_sample_mapping.argtypes = [P, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.sfo_map_update.argtypes = [P, dp, dp, ctypes.c_double, dp, dp, dp, dp]
        L.sfo_bimocq_init.argtypes = [P, dp, dp, B]
        L.sfo_advect_bimocq.argtypes = [P, dp, B, dp]
        L.sfo_accumulate.argtypes = [P, dp, dp, dp, dp, ctypes.c_int]
        L.sfo_distortion.restype = ctypes.c_double
        L.sfo_distortion.argtypes = [P, B]
        L.sfo_bimocq_reinit.argtypes = [P, dp, dp, B]
        L.sfo_bimocq_step.restype = ctypes.c_int
        L.sfo_bimocq_step.argtypes = [P] + [dp] * 6 + [B, ctypes.POINTER(Stats)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(n_theta: int, dt_seconds: float, *, radius=1.0, H=1e-6, U=1.0, Gamma_m=1e-6,
           T=298.15, gravity=False, evaporation=False, airflow=True, M_scale=1.0, tol=1e-12,
           time_s=0.0, floorf_mode=False, tol_absolute=False, max_iters=0,
           viscous=False) -> Params:
    p = Params()
    lib().sfo_derive(radius, H, U, Gamma_m, T, dt_seconds, int(gravity), int(evaporation),
                     ctypes.byref(p))
    p.n_theta = n_theta
    p.airflow = int(airflow)
    p.floorf_mode = int(floorf_mode)
    p.tol_absolute = int(tol_absolute)
    p.max_iters = max_iters
    p.tol = tol
    p.time_s = time_s
    p.M *= M_scale
    p.viscous = int(viscous)
    return p


def sizes(n_theta):
    nph = 2 * n_theta
    return n_theta * nph, (n_theta - 1) * nph


def sample(p: Params, kind: int, field: np.ndarray, th: float, ph: float) -> float:
    return lib().sfo_sample(ctypes.byref(p), kind, _p(field), th, ph)


def trace(p: Params, uth, uph, th, ph):
    out = np.zeros(2)
    lib().sfo_trace(ctypes.byref(p), _p(uth), _p(uph), th, ph, _p(out))
    return out


def advect(p: Params, st: dict) -> dict:
    N, Nt = sizes(p.n_theta)
    o = {"u_theta": np.empty(Nt), "u_phi": np.empty(N), "eta": np.empty(N), "gamma": np.empty(N)}
    lib().sfo_advect(ctypes.byref(p), _p(st["u_theta"]), _p(st["u_phi"]), _p(st["eta"]),
                     _p(st["gamma"]), _p(o["u_theta"]), _p(o["u_phi"]), _p(o["eta"]),
                     _p(o["gamma"]))
    return o


def airflow(p: Params, uair0, vair0):
    N, Nt = sizes(p.n_theta)
    u, v = np.empty(N), np.empty(Nt)
    lib().sfo_airflow(ctypes.byref(p), _p(uair0), _p(vair0), _p(u), _p(v))
    return u, v


def col_ind(p: Params):
    N, _ = sizes(p.n_theta)
    c = np.empty(5 * N, dtype=np.int32)
    lib().sfo_col_ind(ctypes.byref(p), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return c


def assemble(p: Params, st: dict, uair, vair):
    """Returns (val[5N], rhs[N], gamma*) where gamma* carries the broken-cell
    override of the assembly."""
    N, _ = sizes(p.n_theta)
    gam = st["gamma"].copy()
    val, rhs = np.empty(5 * N), np.empty(N)
    lib().sfo_assemble(ctypes.byref(p), _p(st["u_theta"]), _p(st["u_phi"]), _p(gam),
                       _p(st["eta"]), _p(uair), _p(vair), _p(val), _p(rhs))
    return val, rhs, gam


def solve(p: Params, val, rhs, x0):
    x = x0.copy()
    s = Stats()
    rc = lib().sfo_solve(ctypes.byref(p), _p(val), _p(rhs), _p(x), ctypes.byref(s))
    return x, s, rc


def spmv(p: Params, val, x):
    y = np.empty_like(x)
    lib().sfo_spmv(ctypes.byref(p), _p(val), _p(x), _p(y))
    return y


def update(p: Params, st_star: dict, gam_next, uair, vair) -> dict:
    N, Nt = sizes(p.n_theta)
    g = gam_next.copy()
    o = {"u_theta": np.empty(Nt), "u_phi": np.empty(N), "eta": np.empty(N)}
    lib().sfo_update(ctypes.byref(p), _p(st_star["u_theta"]), _p(st_star["u_phi"]),
                     _p(st_star["eta"]), _p(st_star["gamma"]), _p(g), _p(uair), _p(vair),
                     _p(o["u_theta"]), _p(o["u_phi"]), _p(o["eta"]))
    o["gamma"] = g
    return o


def step(p: Params, st: dict, uair0, vair0, nsteps: int = 1, dt_seconds: float | None = None,
         advance_time: bool = True):
    """Advances the state dict in place; returns the list of per-step stats."""
    out = []
    for _ in range(nsteps):
        s = Stats()
        rc = lib().sfo_step(ctypes.byref(p), _p(st["u_theta"]), _p(st["u_phi"]), _p(st["eta"]),
                            _p(st["gamma"]), _p(uair0), _p(vair0), ctypes.byref(s))
        out.append((rc, s.iters, s.rel_residual, s.volume, s.broken))
        if advance_time:
            p.time_s += dt_seconds if dt_seconds is not None else p.dt * p.radius
    return out


def visc_div(p: Params, uth, uph):
    """divergenceKernel (bubble.cu:1217-1269): div(u) at the centres."""
    N, _ = sizes(p.n_theta)
    d = np.empty(N)
    lib().sfo_visc_div(ctypes.byref(p), _p(uth), _p(uph), _p(d))
    return d


def visc_forces(p: Params, uth, uph, eta, div):
    """(fu, fv) of computeforcevelphiKernel_viscous / computeforcevelthetaKernel."""
    N, Nt = sizes(p.n_theta)
    fu, fv = np.empty(N), np.empty(Nt)
    lib().sfo_visc_forces(ctypes.byref(p), _p(uth), _p(uph), _p(eta), _p(div), _p(fu), _p(fv))
    return fu, fv


def volume(p: Params, eta) -> float:
    return lib().sfo_volume(ctypes.byref(p), _p(eta))


def num_threads() -> int:
    return lib().sfo_num_threads()


# ---------------------------------------------------------------- BIMOCQ ----
class BimocqState:
    """Owns the BIMOCQ arrays of one bubble (numpy) and the ctypes view. The
    reference swaps array pointers at reinitialisation; the arrays are read
    back through the struct after every call."""

    def __init__(self, p: Params, eta, gam, blend=0.5, threshold=4.0, multisample=True):
        N = p.n_theta * 2 * p.n_theta
        self.arr = {k: np.zeros(N) for k in MAPS}
        self.s = Bimocq(*[_p(self.arr[k]) for k in MAPS], blend, threshold, int(multisample), 0, 0.0)
        lib().sfo_bimocq_init(ctypes.byref(p), _p(eta), _p(gam), ctypes.byref(self.s))
        self.N = N

    def view(self, k):
        """The array currently bound to field k (after pointer swaps)."""
        ptr = getattr(self.s, k)
        return np.ctypeslib.as_array(ptr, shape=(self.N,))

    @property
    def reinit_count(self):
        return self.s.reinit_count

    @property
    def distortion(self):
        return self.s.last_distortion


def lerp_coords(p: Params, a, b, alpha):
    out = np.zeros(2)
    lib().sfo_lerp_coords(ctypes.byref(p), _p(np.asarray(a, float)), _p(np.asarray(b, float)),
                          alpha, _p(out))
    return out


def sample_mapping(p: Params, mt, mp, rx, ry):
    out = np.zeros(2)
    lib().sfo_sample_mapping(ctypes.byref(p), _p(mt), _p(mp), rx, ry, _p(out))
    return out


def map_update(p: Params, uth, uph, dt, mt, mp):
    ot, op = np.empty_like(mt), np.empty_like(mp)
    lib().sfo_map_update(ctypes.byref(p), _p(uth), _p(uph), dt, _p(mt), _p(mp), _p(ot), _p(op))
    return ot, op


def advect_bimocq(p: Params, eta, b: BimocqState):
    out = np.empty_like(eta)
    lib().sfo_advect_bimocq(ctypes.byref(p), _p(eta), ctypes.byref(b.s), _p(out))
    return out


def bimocq_reinit(p: Params, st: dict, b: BimocqState):
    lib().sfo_bimocq_reinit(ctypes.byref(p), _p(st["eta"]), _p(st["gamma"]), ctypes.byref(b.s))


def bimocq_step(p: Params, st: dict, uair0, vair0, b: BimocqState, nsteps: int = 1,
                advance_time: bool = True):
    out = []
    for _ in range(nsteps):
        s = Stats()
        rc = lib().sfo_bimocq_step(ctypes.byref(p), _p(st["u_theta"]), _p(st["u_phi"]),
                                   _p(st["eta"]), _p(st["gamma"]), _p(uair0), _p(vair0),
                                   ctypes.byref(b.s), ctypes.byref(s))
        out.append((rc, s.iters, s.rel_residual, b.distortion, b.reinit_count))
        if advance_time:
            p.time_s += p.dt * p.radius
    return out
