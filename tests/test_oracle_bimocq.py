"""Pins of the oracle's BIMOCQ scalar advection (bubble.cu:512-633,
967-1013, 1133-1168, 2081-2114, 2180-2280, 2509-2564) on the CPU: the
spherical coordinate interpolation against its closed forms, identity maps,
and the reduction of the whole BIMOCQ step to the plain step when nothing
moves."""
import math

import numpy as np
import pytest

from cases import copy_state, make_case


def _xyz(p, pt):
    h = math.pi / p.n_theta
    t, f = pt[0] * h, pt[1] * h
    return np.array([math.sin(t) * math.cos(f), math.sin(t) * math.sin(f), math.cos(t)])


def test_lerp_coords_closed_forms(oracle):
    p = oracle.params(16, 0.01)
    a, b = np.array([3.3, 7.1]), np.array([5.9, 11.4])
    assert np.allclose(oracle.lerp_coords(p, a, b, 0.0), a, atol=1e-12)
    assert np.allclose(oracle.lerp_coords(p, a, b, 1.0), b, atol=1e-12)
    # the midpoint of a great-circle arc is the normalised sum of its ends
    mid = oracle.lerp_coords(p, a, b, 0.5)
    want = _xyz(p, a) + _xyz(p, b)
    assert np.allclose(_xyz(p, mid), want / np.linalg.norm(want), atol=1e-13)
    # equal angles along the arc
    q = oracle.lerp_coords(p, a, b, 0.3)
    ang = lambda u, v: math.acos(min(1.0, float(np.dot(u, v))))
    assert abs(ang(_xyz(p, a), _xyz(p, q)) - 0.3 * ang(_xyz(p, a), _xyz(p, b))) <= 1e-12
    # coincident points: the reference returns `from`
    assert np.array_equal(oracle.lerp_coords(p, a, a, 0.4), a)


def _identity(n):
    j, i = np.meshgrid(np.arange(n), np.arange(2 * n), indexing="ij")
    return (j + 0.5).ravel().astype(float), i.ravel().astype(float)


def test_identity_mapping(oracle):
    n = 16
    p = oracle.params(n, 0.01)
    mt, mp = _identity(n)
    for j, i in [(0, 0), (3, 7), (15, 31), (8, 0)]:
        got = oracle.sample_mapping(p, mt, mp, j + 0.5, float(i))
        assert abs(got[0] - (j + 0.5)) <= 1e-12 and abs((got[1] - i + n) % (2 * n) - n) <= 1e-11
    # between nodes the great-circle interpolation bows toward the pole by
    # O(h) cells (the chord of a latitude circle): small, and halving with h
    errs = []
    for n in (16, 32, 64):
        p = oracle.params(n, 0.01)
        mt, mp = _identity(n)
        e = 0.0
        for x, y in [(n / 3 + 0.5, 5.5), (n / 2 - 0.25, 17.25), (n / 4 + 0.75, 40.5)]:
            got = oracle.sample_mapping(p, mt, mp, x, y)
            e = max(e, abs(got[0] - x))
        errs.append(e)
    assert errs[0] < 0.05 and errs[0] / errs[1] > 1.7 and errs[1] / errs[2] > 1.7, errs


def test_map_update_zero_velocity(oracle):
    n = 16
    nph = 2 * n
    p = oracle.params(n, 0.01)
    mt, mp = _identity(n)
    ot, op = oracle.map_update(p, np.zeros((n - 1) * nph), np.zeros(n * nph), p.dt, mt, mp)
    assert np.abs(ot - mt).max() <= 1e-12
    assert np.abs((op - mp + n) % nph - n).max() <= 1e-11


def _still_state(n, seed):
    """Constant Gamma, zero velocity and no air: nothing moves, the film only
    evaporates and the solve returns Gamma*."""
    st, ua, va = make_case(n, seed, u_scale=0.0)
    st["gamma"][:] = 1.1
    return st, ua, va


def test_bimocq_step_reduces_to_plain_when_still(oracle):
    """With nothing moving the maps stay the identity, the single-sample
    BIMOCQ value of eta is init + accumulated changes = the current eta, and
    the whole step equals the plain step."""
    n = 16
    st, ua, va = _still_state(n, 4)
    kw = dict(airflow=False, evaporation=True)
    p1, p2 = oracle.params(n, 0.01, **kw), oracle.params(n, 0.01, **kw)
    a, b = copy_state(st), copy_state(st)
    bm = oracle.BimocqState(p1, a["eta"], a["gamma"], multisample=False)
    oracle.bimocq_step(p1, a, ua, va, bm, 3)
    oracle.step(p2, b, ua, va, 3)
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        d = np.abs(a[k] - b[k]).max()
        assert d <= 1e-12 * max(1.0, np.abs(b[k]).max()), (k, d)
    assert bm.reinit_count == 0
    assert b["eta"].mean() < st["eta"].mean()   # it evaporated


def test_reinit_when_still(oracle):
    """A reinitialisation every step (threshold below zero) with nothing
    moving: the velocity is untouched, a constant Gamma keeps its value to
    the reference's float clamp, eta stays within the bounds of its 3 x 3
    neighbourhood, the maps return to the identity and the changes to zero."""
    n = 16
    nph = 2 * n
    st, ua, va = _still_state(n, 6)
    p = oracle.params(n, 0.01, airflow=False)
    a = copy_state(st)
    bm = oracle.BimocqState(p, a["eta"], a["gamma"], threshold=-1.0)
    before = a["eta"].copy()
    oracle.bimocq_step(p, a, ua, va, bm, 1)
    assert bm.reinit_count == 1
    assert np.abs(a["u_phi"]).max() == 0 and np.abs(a["u_theta"]).max() == 0
    assert np.abs(a["gamma"] - np.float32(1.1)).max() <= 1e-15
    e = before.reshape(n, nph)
    lo, hi = e.copy(), e.copy()
    for t in (-1, 0, 1):
        for s in (-1, 0, 1):
            sh = np.roll(np.roll(e, s, axis=1), t, axis=0)
            lo, hi = np.minimum(lo, sh), np.maximum(hi, sh)
    inner = a["eta"].reshape(n, nph)[1:-1]
    assert np.all(inner >= lo[1:-1].astype(np.float32) - 1e-7)
    assert np.all(inner <= hi[1:-1].astype(np.float32) + 1e-7)
    mt, mp = _identity(n)
    assert np.array_equal(bm.view("fwd_t"), mt) and np.array_equal(bm.view("bwd_p"), mp)
    assert np.all(bm.view("eta_delta") == 0) and np.array_equal(bm.view("eta_init"), a["eta"])


def test_accumulate_constant_change(oracle):
    n = 16
    nph = 2 * n
    p = oracle.params(n, 0.01)
    mt, mp = _identity(n)
    delta = np.full(n * nph, 0.25)
    oracle.lib().sfo_accumulate(oracle.ctypes.byref(p), oracle._p(mt), oracle._p(mp),
                                oracle._p(np.full(n * nph, -0.125)), oracle._p(delta), 1)
    assert np.abs(delta - 0.125).max() <= 1e-15


def test_distortion_and_reinit_moving(oracle):
    """A moving film: the maps drift apart, distortion grows step by step,
    and a reinitialisation resets the maps to the identity and the changes
    to zero."""
    n = 32
    st, ua, va = make_case(n, 3, u_scale=0.8)
    p = oracle.params(n, 0.005)
    bm = oracle.BimocqState(p, st["eta"], st["gamma"], threshold=0.01)
    r = oracle.bimocq_step(p, st, ua, va, bm, 5)
    ds = [x[3] for x in r]
    assert ds[0] > 0 and ds[1] > ds[0]
    assert bm.reinit_count >= 1
    assert np.array_equal(bm.view("eta_init"), st["eta"]) or bm.distortion <= 0.01


def test_multisample_is_a_bounded_average(oracle):
    """MULTISAMPLE (weights 1/8 at the four quarter-cell diagonals, 1/2 at
    the centre) through identity maps: constants are kept, other fields are
    averaged within their own bounds."""
    n = 16
    st, ua, va = _still_state(n, 5)
    p = oracle.params(n, 0.01)
    bm = oracle.BimocqState(p, st["eta"], st["gamma"])
    out = oracle.advect_bimocq(p, st["eta"], bm)
    assert out.min() >= st["eta"].min() - 1e-12 and out.max() <= st["eta"].max() + 1e-12
    assert 0 < np.abs(out - st["eta"]).max() < 0.05
    c = np.full_like(st["eta"], 0.8)
    bc = oracle.BimocqState(p, c, st["gamma"])
    assert np.abs(oracle.advect_bimocq(p, c, bc) - 0.8).max() <= 1e-14
