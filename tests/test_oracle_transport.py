"""Pins of the oracle's great-circle velocity advection (CPU only).

advectionVSphere{Theta,Phi}Kernel (bubble.cu:639-904) advance along a great
circle and parallel-transport the sampled velocity back into the start
frame through the frames (u_, v_ = w x u_) of the circle. Both are
rotations: the point x moved by the arc s along the unit direction u_ is
R(n, s) x with n = x x u_, and parallel transport along a great circle is
that same rotation applied to a tangent vector. The pins here state the
method through rotation matrices (Rodrigues' formula) rather than through
the circle frames, and check two closed forms:

* the departure point of the frozen-field trace (traceGreatCircle,
  bubble.cu:416-506) for a solid-body rotation about a tilted axis is the
  start point rotated back about that axis; the midpoint method's local
  error is third order in dt at points off the rotation's equator, where the
  flow's stream lines are small circles (the binormal term of the
  transported midpoint velocity is non-zero there);
* the advected velocity at a face is the field at that departure point
  parallel-transported along the great circle from it to the face; the
  scheme approaches it at third order in dt.

A sign error in either transport (the midpoint's v-frame cross(midx,
uCircleMid), or the departure's cross(px, uCircleP)) turns the orders into
2 and 1.
"""
import math

import numpy as np
import pytest

from test_oracle_samplers import UPHI, UTHETA, _solid_body, ghost_sample


def rodrigues(axis, ang):
    """Rotation by `ang` about the unit vector `axis` (Rodrigues' formula)."""
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    K = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + math.sin(ang) * K + (1.0 - math.cos(ang)) * (K @ K)


def unit(th, ph):
    return np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])


def e_theta(th, ph):
    return np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])


def e_phi(th, ph):
    return np.array([-math.sin(ph), math.cos(ph), 0.0])


def angles(x):
    return math.acos(min(1.0, max(-1.0, float(x[2])))), math.atan2(float(x[1]), float(x[0]))


def velocity_at(uth, uph, nth, x):
    """Tangent velocity vector at the unit vector x from the bilinear
    samples of the MAC fields (the ghost-row restatement of sampleVTheta /
    sampleVPhi)."""
    h = math.pi / nth
    th, ph = angles(x)
    vt = ghost_sample(UTHETA, uth, nth, th / h, ph / h)
    vp = ghost_sample(UPHI, uph, nth, th / h, ph / h)
    return vt * e_theta(th, ph) + vp * e_phi(th, ph)


def advect_face(uth, uph, nth, dt, th, ph):
    """The great-circle midpoint advection of the velocity at the face point
    (th, ph), restated with rotation matrices: the start velocity u, its
    half-step point m = R(n, -|u| dt / 2) x (n = x x u / |u|), the velocity
    there rotated back by the inverse rotation, the full step p along that
    direction, and the velocity at p rotated back to x. Returns the vector
    at x, or the speed |u| when it is not positive (what the reference
    writes then)."""
    x = unit(th, ph)
    u = velocity_at(uth, uph, nth, x)
    s = float(np.linalg.norm(u))
    if not s > 0:
        return s
    R = rodrigues(np.cross(x, u / s), -0.5 * s * dt)
    u = R.T @ velocity_at(uth, uph, nth, R @ x)
    s = float(np.linalg.norm(u))
    if not s > 0:
        return s
    R = rodrigues(np.cross(x, u / s), -s * dt)
    return R.T @ velocity_at(uth, uph, nth, R @ x)


def restated_advection(uth, uph, nth, dt):
    h = math.pi / nth
    nph = 2 * nth
    out_t = np.empty((nth - 1) * nph)
    out_p = np.empty(nth * nph)
    for j in range(nth - 1):
        for i in range(nph):
            th, ph = (j + 1.0) * h, i * h
            v = advect_face(uth, uph, nth, dt, th, ph)
            out_t[j * nph + i] = v if np.isscalar(v) else float(v @ e_theta(th, ph))
    for j in range(nth):
        for i in range(nph):
            th, ph = (j + 0.5) * h, (i - 0.5) * h
            v = advect_face(uth, uph, nth, dt, th, ph)
            out_p[j * nph + i] = v if np.isscalar(v) else float(v @ e_phi(th, ph))
    return out_t, out_p


_AXIS = np.array([0.4, -0.3, 0.85]) / np.linalg.norm([0.4, -0.3, 0.85])


def _fields(kind, nth, seed):
    nph = 2 * nth
    if kind == "random":
        rng = np.random.default_rng(seed)
        return rng.uniform(-1.0, 1.0, (nth - 1) * nph), rng.uniform(-1.0, 1.0, nth * nph)
    uth, uph, _ = _solid_body(nth, _AXIS, 1.3)
    return uth, uph


@pytest.mark.parametrize("kind", ["random", "tilted"])
@pytest.mark.parametrize("nth,dt", [(8, 0.1), (16, 0.05), (16, 0.15)])
def test_velocity_advection_matches_rotation_restatement(oracle, kind, nth, dt):
    uth, uph = _fields(kind, nth, 10 * nth + 1)
    nph = 2 * nth
    st = {"u_theta": uth, "u_phi": uph, "eta": np.ones(nth * nph), "gamma": np.ones(nth * nph)}
    got = oracle.advect(oracle.params(nth, dt), st)
    want_t, want_p = restated_advection(uth, uph, nth, dt)
    assert np.abs(got["u_theta"] - want_t).max() <= 1e-12
    assert np.abs(got["u_phi"] - want_p).max() <= 1e-12
    # the binormal (v-frame) component is not negligible in these flows
    diff = np.abs(got["u_phi"] - uph).max()
    assert diff > 1e-3


def _off_equator_points(rng, n, axis):
    pts = []
    while len(pts) < n:
        r = rng.normal(size=3)
        r /= np.linalg.norm(r)
        c = abs(float(r @ axis))
        if 0.3 < c < 0.9 and abs(r[2]) < 0.9:
            pts.append(r)
    return pts


def test_trace_tilted_rotation_third_order(oracle):
    """traceGreatCircle for u = axis x r (|axis| = 1) at points between 17
    and 73 degrees from the rotation's equator: the departure point is
    R(axis, -dt) x up to the midpoint method's O(dt^3) local error
    (measured: 5.1e-4, 6.5e-5, 8.4e-6 at dt = 0.2, 0.1, 0.05 on 128 x 256;
    the transported midpoint velocity with its binormal sign flipped gives
    O(dt^2): 2.0e-2, 5.0e-3, 1.2e-3)."""
    nth = 128
    h = math.pi / nth
    uth, uph, a = _solid_body(nth, _AXIS, 1.0)
    pts = _off_equator_points(np.random.default_rng(1), 40, a)
    errs = []
    for dt in (0.2, 0.1, 0.05):
        p = oracle.params(nth, dt)
        worst = 0.0
        for r in pts:
            th, ph = angles(r)
            got = oracle.trace(p, uth, uph, th / h, ph / h) * h
            worst = max(worst, float(np.linalg.norm(unit(*got) - rodrigues(a, -dt) @ r)))
        errs.append(worst)
    assert errs[2] < 2e-5, errs
    assert errs[0] / errs[1] > 6.5 and errs[1] / errs[2] > 6.5, errs


def test_velocity_advection_transport_third_order(oracle):
    """advectionVSphere{Theta,Phi}Kernel for u = axis x r at faces between
    17 and 73 degrees from the rotation's equator, against the velocity of
    the sampled field at the exact departure point p = R(axis, -dt) x,
    parallel-transported along the great circle from p to x (the rotation
    about p x x by the angle between them): third order in dt (measured
    4.0e-4, 5.0e-5, 5.8e-6 at dt = 0.2, 0.1, 0.05 on 128 x 256; a flipped
    midpoint transport gives order 2, 8.9e-4 at dt = 0.05, a flipped
    departure transport order 1, 2.5e-2)."""
    nth = 128
    nph = 2 * nth
    h = math.pi / nth
    uth, uph, a = _solid_body(nth, _AXIS, 1.0)
    st = {"u_theta": uth, "u_phi": uph, "eta": np.ones(nth * nph), "gamma": np.ones(nth * nph)}
    rng = np.random.default_rng(3)
    faces = []
    while len(faces) < 60:
        kind = int(rng.integers(2))
        j, i = int(rng.integers(nth - 1)), int(rng.integers(nph))
        th, ph = ((j + 1.0) * h, i * h) if kind == 0 else ((j + 0.5) * h, (i - 0.5) * h)
        r = unit(th, ph)
        if 0.3 < abs(float(r @ a)) < 0.9 and abs(r[2]) < 0.9:
            faces.append((kind, j, i, th, ph))
    errs = []
    for dt in (0.2, 0.1, 0.05):
        out = oracle.advect(oracle.params(nth, dt), st)
        worst = 0.0
        for kind, j, i, th, ph in faces:
            x = unit(th, ph)
            p = rodrigues(a, -dt) @ x
            ax = np.cross(p, x)
            s = float(np.linalg.norm(ax))
            v = rodrigues(ax / s, math.asin(min(1.0, s))) @ velocity_at(uth, uph, nth, p)
            if kind == 0:
                got, want = out["u_theta"][j * nph + i], float(v @ e_theta(th, ph))
            else:
                got, want = out["u_phi"][j * nph + i], float(v @ e_phi(th, ph))
            worst = max(worst, abs(got - want))
        errs.append(worst)
    assert errs[2] < 2e-5, errs
    assert errs[0] / errs[1] > 6.5 and errs[1] / errs[2] > 6.5, errs
