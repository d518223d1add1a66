"""Pins of the oracle's samplers and great-circle trace (CPU only).

The samplers (sampleCentered, sampleVPhi, sampleVTheta, bubble.cu:178-376)
are checked against an independent ghost-row restatement: the field is
extended by one row beyond each pole whose values are the phi + pi mirror of
the first/last row (negated for the velocity components, whose unit vectors
flip across the pole), and a sample is plain bilinear interpolation on that
extended periodic grid. The reference's branch structure (validateCoord's
reflection, the polar-row branches, the half-weight polar caps of u_theta)
must reduce to exactly that.

The trace (traceGreatCircle, bubble.cu:416-506) is checked against closed
forms: zero velocity, uniform flow along the equator and along a meridian,
and O(h^2) convergence to the exact rotation of a solid-body flow.
"""
import math

import numpy as np
import pytest

CENTRED, UPHI, UTHETA = 0, 1, 2


def ghost_sample(kind, f, nth, rx, ry):
    """Bilinear sample on the pole-extended grid at index-space (rx, ry),
    0 <= rx <= nth. f is (rows, nph)."""
    nph = 2 * nth
    f = f.reshape(-1, nph)
    sign = 1.0 if kind == CENTRED else -1.0
    mirror = lambda row: sign * np.roll(row, -nth)   # value at column i is row[i + nth]
    if kind == UTHETA:
        pos = [-1.0] + [j + 1.0 for j in range(nth - 1)] + [nth + 1.0]
        rows = [mirror(f[0])] + [f[j] for j in range(nth - 1)] + [mirror(f[nth - 2])]
        col_off = 0.0
    else:
        pos = [-0.5] + [j + 0.5 for j in range(nth)] + [nth + 0.5]
        rows = [mirror(f[0])] + [f[j] for j in range(nth)] + [mirror(f[nth - 1])]
        col_off = 0.0 if kind == CENTRED else -0.5
    k = int(np.searchsorted(pos, rx, side="right")) - 1
    k = min(max(k, 0), len(pos) - 2)
    w = (rx - pos[k]) / (pos[k + 1] - pos[k])
    c = ry - col_off
    c0 = math.floor(c)
    a = c - c0
    i0, i1 = c0 % nph, (c0 + 1) % nph
    lo = (1 - a) * rows[k][i0] + a * rows[k][i1]
    hi = (1 - a) * rows[k + 1][i0] + a * rows[k + 1][i1]
    return (1 - w) * lo + w * hi


def _field(kind, nth, rng):
    nph = 2 * nth
    rows = nth - 1 if kind == UTHETA else nth
    return rng.uniform(-1.0, 1.0, rows * nph)


def _points(nth, rng, n=400):
    nph = 2 * nth
    rx = np.concatenate([rng.uniform(0.0, nth, n),
                         np.arange(0, 2 * nth + 1) * 0.5,            # nodes and faces
                         [0.0, 1e-9, 0.25, 0.5, 0.75, 1.0, nth - 1.0, nth - 0.5, nth - 1e-9, nth]])
    ry = np.concatenate([rng.uniform(-nph / 2, 1.5 * nph, n),
                         rng.integers(-nph // 2, nph + nph // 2, 2 * nth + 1) * 0.5,
                         rng.uniform(0, nph, 10)])
    return rx, ry


@pytest.mark.parametrize("kind", [CENTRED, UPHI, UTHETA])
@pytest.mark.parametrize("nth", [8, 16])
def test_sampler_matches_ghost_grid(oracle, kind, nth):
    rng = np.random.default_rng(100 * kind + nth)
    p = oracle.params(nth, 0.01)
    f = _field(kind, nth, rng)
    rx, ry = _points(nth, rng)
    worst = 0.0
    for x, y in zip(rx, ry):
        if kind == UTHETA and x == 0.0:
            continue   # the north pole itself: test_vtheta_at_north_pole
        got = oracle.sample(p, kind, f, float(x), float(y))
        ref = ghost_sample(kind, f, nth, float(x), float(y))
        worst = max(worst, abs(got - ref))
        assert abs(got - ref) <= 1e-13, (kind, x, y, got, ref)
    assert worst <= 1e-13


@pytest.mark.parametrize("kind", [CENTRED, UPHI, UTHETA])
def test_sampler_constant_and_periodic(oracle, kind):
    nth = 16
    nph = 2 * nth
    p = oracle.params(nth, 0.01)
    rows = nth - 1 if kind == UTHETA else nth
    rng = np.random.default_rng(7)
    if kind == CENTRED:
        f = np.full(rows * nph, 0.37)
        for x, y in zip(rng.uniform(0, nth, 200), rng.uniform(-nph / 2, 1.5 * nph, 200)):
            assert abs(oracle.sample(p, kind, f, x, y) - 0.37) <= 1e-15
    f = _field(kind, nth, rng)
    for x, y in zip(rng.uniform(0, nth, 100), rng.uniform(0, nph / 2, 100)):
        a = oracle.sample(p, kind, f, x, y)
        assert abs(a - oracle.sample(p, kind, f, x, y + nph)) <= 1e-13
        assert abs(a - oracle.sample(p, kind, f, x, y - nph / 2 + nph / 2)) <= 1e-13


@pytest.mark.parametrize("kind", [CENTRED, UPHI, UTHETA])
def test_sampler_exact_at_nodes(oracle, kind):
    nth = 8
    nph = 2 * nth
    p = oracle.params(nth, 0.01)
    f = _field(kind, nth, np.random.default_rng(3)).reshape(-1, nph)
    xo, yo = {CENTRED: (0.5, 0.0), UPHI: (0.5, -0.5), UTHETA: (1.0, 0.0)}[kind]
    for j in range(f.shape[0]):
        for i in range(nph):
            got = oracle.sample(p, kind, f.ravel(), j + xo, i + yo)
            assert got == f[j, i], (j, i)


def test_sampler_pole_mirror(oracle):
    """A sample taken through the pole (theta < 0, reflected to -theta at
    phi + pi) equals the sample at the mirrored point, negated for the
    velocity components only when the reference reflects the point itself."""
    nth = 16
    p = oracle.params(nth, 0.01)
    f = _field(CENTRED, nth, np.random.default_rng(5))
    for x, y in [(0.2, 3.3), (0.05, 7.0), (0.45, 30.9)]:
        a = oracle.sample(p, CENTRED, f, x, y)
        b = ghost_sample(CENTRED, f, nth, x, y)
        assert abs(a - b) <= 1e-14


def test_vtheta_at_north_pole(oracle):
    """At theta = 0 exactly, sampleVTheta reflects rawId.x - 1 = -1 to +1 and
    takes the interior branch on rows 1 and 2 with weight 0: the value is
    row 1 at phi + pi, negated (bubble.cu:248-312). Pinned by the text of the
    reference only; the trace reaches it only for a point exactly on the
    pole."""
    nth = 8
    nph = 2 * nth
    p = oracle.params(nth, 0.01)
    f = _field(UTHETA, nth, np.random.default_rng(4)).reshape(nth - 1, nph)
    for y in (0.0, 2.5, 13.75):
        c = y + nth
        c0 = math.floor(c)
        a = c - c0
        want = -((1 - a) * f[1, c0 % nph] + a * f[1, (c0 + 1) % nph])
        assert abs(oracle.sample(p, UTHETA, f.ravel(), 0.0, y) - want) <= 1e-15


def test_sampler_nan_skip(oracle):
    """bubbleLerp (bubble.cu:153-160): a NaN operand is replaced by the other
    operand of that interpolation."""
    nth = 8
    nph = 2 * nth
    p = oracle.params(nth, 0.01)
    f = _field(CENTRED, nth, np.random.default_rng(9)).reshape(nth, nph)
    f[3, 5] = np.nan
    # interior point between rows 3, 4 and columns 5, 6
    x, y = 3.5 + 0.3, 5.25
    got = oracle.sample(p, CENTRED, f.ravel(), x, y)
    lo = f[3, 6]                       # NaN skipped in the phi interpolation of row 3
    hi = 0.75 * f[4, 5] + 0.25 * f[4, 6]
    assert abs(got - (0.7 * lo + 0.3 * hi)) <= 1e-15
    # both operands NaN: the interpolation stays NaN and the theta lerp skips it
    f[3, 6] = np.nan
    got = oracle.sample(p, CENTRED, f.ravel(), x, y)
    assert abs(got - hi) <= 1e-15


# ------------------------------------------------------------- the trace ----
def _sizes(nth):
    nph = 2 * nth
    return (nth - 1) * nph, nth * nph


def test_trace_zero_velocity(oracle):
    nth = 16
    p = oracle.params(nth, 0.01)
    nt, nc = _sizes(nth)
    uth, uph = np.zeros(nt), np.zeros(nc)
    for x, y in [(0.5, 0.0), (3.7, 11.2), (15.5, 31.0)]:
        assert np.array_equal(oracle.trace(p, uth, uph, x, y), [x, y])


@pytest.mark.parametrize("omega", [0.8, -1.7])
def test_trace_along_equator(oracle, omega):
    """u_phi = omega everywhere, u_theta = 0: a point on the equator moves
    along the equator, a great circle, by omega dt."""
    nth = 32
    dt = 0.05
    p = oracle.params(nth, dt)
    h = math.pi / nth
    nt, nc = _sizes(nth)
    uth, uph = np.zeros(nt), np.full(nc, omega)
    for y in [3.0, 17.25, 60.5]:
        got = oracle.trace(p, uth, uph, nth / 2, y)
        assert abs(got[0] - nth / 2) <= 1e-12
        want = (y - omega * dt / h) % (2 * nth)
        assert abs((got[1] - want + nth) % (2 * nth) - nth) <= 1e-11, (got, want)


def test_trace_along_meridian(oracle):
    """u_theta = omega everywhere, u_phi = 0: a point away from the poles
    moves along its meridian by omega dt."""
    nth = 32
    dt, omega = 0.05, 1.3
    p = oracle.params(nth, dt)
    h = math.pi / nth
    nt, nc = _sizes(nth)
    uth, uph = np.full(nt, omega), np.zeros(nc)
    for x, y in [(nth / 2, 5.0), (10.0, 40.0), (22.5, 7.0)]:
        got = oracle.trace(p, uth, uph, x, y)
        assert abs(got[0] - (x - omega * dt / h)) <= 1e-11
        # phi comes back from atan2, in (-nth, nth]
        assert abs((got[1] - y + nth) % (2 * nth) - nth) <= 1e-11


def _solid_body(nth, axis, omega):
    """MAC velocity fields of the rotation u = omega axis x r."""
    nph = 2 * nth
    h = math.pi / nth
    a = np.asarray(axis, float) / np.linalg.norm(axis)

    def comp(th, ph, which):
        r = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        u = omega * np.cross(a[:, None, None], r, axis=0)
        if which == "theta":
            e = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
        else:
            e = np.stack([-np.sin(ph), np.cos(ph), np.zeros_like(ph)])
        return (u * e).sum(axis=0)

    jt, it = np.meshgrid(np.arange(nth - 1), np.arange(nph), indexing="ij")
    uth = comp((jt + 1.0) * h, it * h, "theta")
    jp, ip = np.meshgrid(np.arange(nth), np.arange(nph), indexing="ij")
    uph = comp((jp + 0.5) * h, (ip - 0.5) * h, "phi")
    return uth.ravel(), uph.ravel(), a


def _rotate(v, a, ang):
    return (v * math.cos(ang) + np.cross(a, v) * math.sin(ang) + a * np.dot(a, v) * (1 - math.cos(ang)))


def test_trace_solid_body_second_order(oracle):
    """On the great circle perpendicular to the rotation axis the departure
    point is the start point rotated by -omega dt; the only error is the
    bilinear interpolation of the velocity, O(h^2)."""
    dt, omega = 0.04, 1.0
    axis = (0.3, -0.5, 0.8)
    errs = []
    for nth in (16, 32, 64):
        h = math.pi / nth
        p = oracle.params(nth, dt)
        uth, uph, a = _solid_body(nth, axis, omega)
        # points on the circle a . r = 0, away from the poles
        b1 = np.cross(a, [0.0, 0.0, 1.0]); b1 /= np.linalg.norm(b1)
        b2 = np.cross(a, b1)
        worst = 0.0
        for s in np.linspace(0.1, 2 * math.pi, 7, endpoint=False):
            r = b1 * math.cos(s) + b2 * math.sin(s)
            th, ph = math.acos(r[2]), math.atan2(r[1], r[0])
            got = oracle.trace(p, uth, uph, th / h, ph / h)
            want = _rotate(r, a, -omega * dt)
            gth, gph = got * h
            g = np.array([math.sin(gth) * math.cos(gph), math.sin(gth) * math.sin(gph), math.cos(gth)])
            worst = max(worst, float(np.linalg.norm(g - want)))
        errs.append(worst)
    assert errs[-1] < 1e-4
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs
