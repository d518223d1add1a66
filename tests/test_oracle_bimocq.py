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


# A meridian flow (u_theta = omega, u_phi = 0): every trace runs along its
# meridian, a great circle, by s = omega dt / h cells (the trace pin of
# tests/test_oracle_samplers.py), and a map lookup between two points of one
# meridian is exact along it (equal angles), with no phi fraction. So k map
# updates from the identity give bwd(x) = x - k s (updates with +dt) and
# fwd(x) = x + k s (-dt) exactly, on the rows the polar rows' cross-pole
# velocity samples have not reached (one row more per update).
_MER = dict(n=32, dt=0.01, omega=1.3)


def _meridian_maps(oracle, k):
    n, dt, omega = _MER["n"], _MER["dt"], _MER["omega"]
    nph = 2 * n
    p = oracle.params(n, dt)
    uth, uph = np.full((n - 1) * nph, omega), np.zeros(n * nph)
    mt, mp = _identity(n)
    maps = []
    bt, bp, ft, fp = mt, mp, mt, mp
    for _ in range(k):
        bt, bp = oracle.map_update(p, uth, uph, dt, bt, bp)
        ft, fp = oracle.map_update(p, uth, uph, -dt, ft, fp)
        maps.append((bt, bp, ft, fp))
    return p, omega * dt / (math.pi / n), maps


def test_maps_follow_a_meridian_flow(oracle):
    """updateMappingKernel on a moving flow against its closed form: a dt of
    the wrong sign, or the map sampled at x instead of the traced point,
    fails it."""
    n = _MER["n"]
    nph = 2 * n
    p, s, maps = _meridian_maps(oracle, 3)
    mt, mp = _identity(n)
    for k, (bt, bp, ft, fp) in enumerate(maps, start=1):
        rows = slice((k + 1) * nph, (n - 1 - k) * nph)
        assert np.abs(bt[rows] - (mt[rows] - k * s)).max() <= 1e-11, k
        assert np.abs(ft[rows] - (mt[rows] + k * s)).max() <= 1e-11, k
        for q in (bp, fp):
            assert np.abs((q[rows] - mp[rows] + n) % nph - n).max() <= 1e-10, k
    # and the two maps are inverse there: the distortion measure's terms vanish
    bt, bp, ft, fp = maps[-1]
    for j in range(6, n - 6):
        back = oracle.sample_mapping(p, bt, bp, ft[j * nph + 5], fp[j * nph + 5])
        assert abs(back[0] - (j + 0.5)) <= 1e-11


def test_accumulate_through_a_shifted_forward_map(oracle):
    """accumulateChangesCentered pulls the step's change back through the
    forward map: with fwd(x) = x + 2 s and a change linear in theta,
    c_j = a j + b, each interior row gains exactly a (j + 2 s) + b (the
    backward map in its place gives a (j - 2 s) + b)."""
    n = _MER["n"]
    nph = 2 * n
    p, s, maps = _meridian_maps(oracle, 2)
    ft, fp = maps[1][2], maps[1][3]
    a, b = 0.03, 0.2
    change = np.repeat(a * np.arange(n) + b, nph)
    delta = np.full(n * nph, 0.5)
    oracle.lib().sfo_accumulate(oracle.ctypes.byref(p), oracle._p(ft), oracle._p(fp),
                                oracle._p(change), oracle._p(delta), 0)
    d = delta.reshape(n, nph)
    for j in range(3, n - 3):
        assert np.abs(d[j] - (0.5 + a * (j + 2 * s) + b)).max() <= 1e-12, j


def test_bimocq_eta_through_shifted_maps(oracle):
    """advectionCenteredBimocq through the backward maps of a meridian flow
    (current interval bwd(x) = x - 2 s, previous interval x - s) with the
    four fields linear in theta and distinct: the single-sample eta is
    (1 - bl)(init_last(x - 3 s) + delta(x - 2 s) + delta_last(x - 3 s)) +
    bl (init(x - 2 s) + delta(x - 2 s)) exactly, so a swapped map, field or
    blend weight fails it."""
    n = _MER["n"]
    nph = 2 * n
    p, s, maps = _meridian_maps(oracle, 2)
    bl = 0.3
    ones = np.ones(n * nph)
    bm = oracle.BimocqState(p, ones, ones, blend=bl, multisample=False)
    lin = {"eta_init": (0.02, 1.0), "eta_initlast": (-0.01, 0.8),
           "eta_delta": (0.005, 0.1), "eta_deltalast": (0.003, -0.05)}
    for k, (a, b) in lin.items():
        bm.view(k)[:] = np.repeat(a * np.arange(n) + b, nph)
    bm.view("bwd_t")[:], bm.view("bwd_p")[:] = maps[1][0], maps[1][1]
    bm.view("bwdprev_t")[:], bm.view("bwdprev_p")[:] = maps[0][0], maps[0][1]
    out = oracle.advect_bimocq(p, ones, bm).reshape(n, nph)
    f = lambda k, x: lin[k][0] * x + lin[k][1]     # at row coordinate x
    for j in range(4, n - 4):
        ip, lp = j - 2 * s, j - 3 * s
        want = (1 - bl) * (f("eta_initlast", lp) + f("eta_delta", ip) + f("eta_deltalast", lp)) \
            + bl * (f("eta_init", ip) + f("eta_delta", ip))
        assert np.abs(out[j] - want).max() <= 1e-12, (j, out[j][0], want)


def test_reinit_error_correction_through_shifted_maps(oracle):
    """correctCentered1/2 (bubble.cu:2180-2280) through the maps of a
    meridian flow (fwd(x) = x + 2 s, bwd(x) = x - 2 s): with the current
    field c, the accumulated change delta and the initial field init linear
    in theta, every sample is exact, so the error estimate is the closed form
    e = (sum_q w_q c(fwd(x + d_q)) - delta - init) / 2 and the corrected
    field is c - sum_q w_q e(bwd(x + d_q)) (the five points d_q: quarter-cell
    diagonals of weight 1/8 and the centre, 1/2). The map lookups at the
    quarter-cell points bow toward the pole by a few thousandths of a cell
    (a great-circle arc between two points of one latitude); their values
    come from sampleMapping, pinned by its own closed forms. init is chosen
    as c(x + 2 s) - delta - 2 eps with eps of one sign and below a third of
    c's row step,
    so e is about eps, non-zero, and the correction stays inside the 3 x 3
    clamp: an error estimate of the wrong sign (or a map swapped for the
    other) moves the result by about 2 eps."""
    n = _MER["n"]
    nph = 2 * n
    p, s, maps = _meridian_maps(oracle, 2)
    bt, bp, ft, fp = maps[1]
    rows = np.arange(n, dtype=float)
    W = (0.125, 0.125, 0.125, 0.125, 0.5)
    DX = (-0.25, 0.25, -0.25, 0.25, 0.0)
    DY = (-0.25, -0.25, 0.25, 0.25, 0.0)
    spec = {"eta": (0.02, 1.0, (0.004, 0.05), (0.003, 0.0001)),
            "gamma": (-0.015, 1.2, (-0.002, 0.03), (-0.002, -0.00005))}
    st = {}
    ones = np.ones(n * nph)
    bm = oracle.BimocqState(p, ones, ones)
    bm.view("fwd_t")[:], bm.view("fwd_p")[:] = ft, fp
    bm.view("bwd_t")[:], bm.view("bwd_p")[:] = bt, bp
    for k, (a, b, (da, db), (ea, eb)) in spec.items():
        c = a * rows + b
        delta = da * rows + db
        eps = ea + eb * rows
        init = a * (rows + 2 * s) + b - delta - 2 * eps
        st[k] = np.repeat(c, nph)
        pre = "eta" if k == "eta" else "gam"
        bm.view(pre + "_delta")[:] = np.repeat(delta, nph)
        bm.view(pre + "_init")[:] = np.repeat(init, nph)
    col = 5
    fwd_rows = np.array([[oracle.sample_mapping(p, ft, fp, j + 0.5 + DX[q], col + DY[q])[0]
                          for q in range(5)] for j in range(n)])
    bwd_rows = np.array([[oracle.sample_mapping(p, bt, bp, j + 0.5 + DX[q], col + DY[q])[0]
                          for q in range(5)] for j in range(n)])
    want = {}
    for k, (a, b, (da, db), (ea, eb)) in spec.items():
        c = a * rows + b
        delta = da * rows + db
        init = a * (rows + 2 * s) + b - delta - 2 * (ea + eb * rows)
        te = (np.asarray(W) * (a * (fwd_rows - 0.5) + b)).sum(axis=1)
        e = 0.5 * (te - delta - init)
        assert np.abs(e[4:n - 4]).min() > 1e-4          # the correction is not zero
        corr = (np.asarray(W) * np.interp(bwd_rows - 0.5, rows, e)).sum(axis=1)
        want[k] = c - corr
        inner = slice(6, n - 6)
        assert np.all(np.abs(corr[inner]) < 0.5 * abs(a))   # inside the 3 x 3 clamp
    full = {"u_theta": np.zeros((n - 1) * nph), "u_phi": np.zeros(n * nph),
            "eta": st["eta"], "gamma": st["gamma"]}
    oracle.bimocq_reinit(p, full, bm)
    for k in ("eta", "gamma"):
        got = full[k].reshape(n, nph)
        for j in range(6, n - 6):
            # the reference clamps in float: the result is a float
            assert np.abs(got[j] - want[k][j]).max() <= 2e-7 * abs(want[k][j]), (k, j)
