"""Pins of the oracle's VISCOUS stages (CPU only): divergenceKernel and the
two viscous force kernels of bubble.cu:1217-1269, 1511-1816 against closed
forms and invariants, and their plumbing into the assembly and the update
against the plain (gravity) step they reduce to when Re^-1 = 0."""
import numpy as np
import pytest

from cases import make_case


def _grids(n):
    h = np.pi / n
    nph = 2 * n
    th_c = (np.arange(n) + 0.5) * h            # centres, u_phi rows
    th_f = (np.arange(n - 1) + 1.0) * h        # u_theta rows
    ph_c = np.arange(nph) * h                  # centres, u_theta columns
    ph_u = (np.arange(nph) - 0.5) * h          # u_phi columns
    return h, nph, th_c, th_f, ph_c, ph_u


def _div_errors(oracle, n):
    """Errors of the discrete divergence against its continuous form for
    u_theta = sin(theta) (term (1/sin) d(sin v)/dtheta = 2 cos theta) and
    u_phi = sin(theta) sin(phi) (term du/dphi without 1/sin, as the
    reference writes it: sin(theta) cos(phi))."""
    h, nph, th_c, th_f, ph_c, ph_u = _grids(n)
    p = oracle.params(n, 0.01)
    uth = np.repeat(np.sin(th_f), nph)
    uph = np.zeros(n * nph)
    d = oracle.visc_div(p, uth, uph).reshape(n, nph)
    inner = slice(2, n - 2)
    e1 = np.max(np.abs(d[inner] - 2 * np.cos(th_c[inner])[:, None]))
    uth0 = np.zeros((n - 1) * nph)
    uph2 = (np.sin(th_c)[:, None] * np.sin(ph_u)[None, :]).ravel()
    d2 = oracle.visc_div(p, uth0, uph2).reshape(n, nph)
    ref2 = np.sin(th_c)[:, None] * np.cos(ph_c)[None, :]
    e2 = np.max(np.abs(d2 - ref2))
    return e1, e2


def test_visc_div_closed_forms_second_order(oracle):
    a = _div_errors(oracle, 32)
    b = _div_errors(oracle, 64)
    for ea, eb in zip(a, b):
        assert eb < 2e-3
        assert ea / eb > 3.5           # O(h^2): a factor ~4 per halving of h


def _rigid_rotation(n, omega=0.3):
    """Rotation about the polar axis: u_phi = omega sin(theta), u_theta = 0 --
    a Killing field: no strain, no divergence."""
    h, nph, th_c, th_f, ph_c, ph_u = _grids(n)
    uph = np.repeat(omega * np.sin(th_c), nph)
    uth = np.zeros((n - 1) * nph)
    return uth, uph


def _force_max(oracle, n, re_scale):
    uth, uph = _rigid_rotation(n)
    p = oracle.params(n, 0.01, gravity=True)
    p.re *= re_scale
    h, nph, th_c, th_f, ph_c, _ = _grids(n)
    # a smooth, non-uniform thickness so that the thickness-gradient terms
    # (which multiply the strain) take part
    eta = (0.7 + 0.1 * np.sin(th_c)[:, None] * np.cos(ph_c)[None, :]).ravel()
    d = oracle.visc_div(p, uth, uph)
    fu, fv = oracle.visc_forces(p, uth, uph, eta, d)
    fv_visc = fv.reshape(n - 1, nph) - (p.gs * np.sin(th_f))[:, None]
    return np.max(np.abs(d)), np.max(np.abs(fu)), np.max(np.abs(fv_visc))


def test_visc_forces_vanish_for_rigid_rotation(oracle):
    """The strain of a rigid rotation is zero, so the viscous forces are
    discretisation error only, falling at second order (Re^-1 scaled up so
    that the error stays above rounding and below the +-0.1 clamp)."""
    d32, fu32, fv32 = _force_max(oracle, 32, 100.0)
    d64, fu64, fv64 = _force_max(oracle, 64, 100.0)
    assert d32 < 1e-12 and d64 < 1e-12
    assert fu64 < 1e-2 and fv64 < 1e-2
    assert fu32 / fu64 > 3.0 and fv32 / fv64 > 3.0


def test_visc_forces_without_viscosity_are_gravity(oracle):
    """Re^-1 = 0: fu = 0 and fv = g sin(theta) on every face, whatever the
    flow and thickness (including the sin(theta) < 0.1 rows)."""
    n = 16
    st, _, _ = make_case(n, 5, u_scale=0.7)
    p = oracle.params(n, 0.01, gravity=True)
    p.re = 0.0
    d = oracle.visc_div(p, st["u_theta"], st["u_phi"])
    fu, fv = oracle.visc_forces(p, st["u_theta"], st["u_phi"], st["eta"], d)
    h, nph, th_c, th_f, _, _ = _grids(n)
    assert np.all(fu == 0.0)
    assert np.allclose(fv.reshape(n - 1, nph), (p.gs * np.sin(th_f))[:, None], rtol=0, atol=1e-15)


def test_visc_forces_clamp_and_polar_rows(oracle):
    """Large Re^-1 saturates the force at the float clamp +-0.1 (plus gravity
    in fv); rows with sin(theta) < 0.1 carry no viscous force."""
    n = 32
    st, _, _ = make_case(n, 9, l_max=6, u_scale=1.0)
    p = oracle.params(n, 0.01, gravity=True)
    p.re = 1e3
    d = oracle.visc_div(p, st["u_theta"], st["u_phi"])
    fu, fv = oracle.visc_forces(p, st["u_theta"], st["u_phi"], st["eta"], d)
    h, nph, th_c, th_f, _, _ = _grids(n)
    fu = fu.reshape(n, nph)
    fvv = fv.reshape(n - 1, nph) - (p.gs * np.sin(th_f))[:, None]
    lim = np.float32(0.1)
    assert np.max(np.abs(fu)) <= lim and np.max(np.abs(fvv)) <= lim + 1e-12
    assert np.isclose(np.max(np.abs(fu)), lim)
    polar_c = np.sin(th_c) < 0.1
    polar_f = np.sin(th_f) < 0.1
    assert np.all(fu[polar_c] == 0.0)
    assert np.all(np.abs(fvv[polar_f]) < 1e-15)


def test_visc_nan_on_empty_faces(oracle):
    """A face between two cells of zero thickness gets a NaN force, as the
    reference writes it (isinf(2 / (eta + eta)))."""
    n = 16
    st, _, _ = make_case(n, 3, u_scale=0.5)
    nph = 2 * n
    eta = st["eta"].reshape(n, nph).copy()
    eta[5, 7:9] = 0.0
    eta[6, 7] = 0.0
    p = oracle.params(n, 0.01, gravity=True)
    d = oracle.visc_div(p, st["u_theta"], st["u_phi"])
    fu, fv = oracle.visc_forces(p, st["u_theta"], st["u_phi"], eta.ravel(), d)
    assert np.isnan(fu.reshape(n, nph)[5, 8])        # faces (5,7)|(5,8)
    assert np.isnan(fv.reshape(n - 1, nph)[5, 7])    # faces (5,7)|(6,7)
    assert np.isfinite(fu.reshape(n, nph)[5, 7])


@pytest.mark.parametrize("steps", [1, 3])
def test_viscous_step_without_viscosity_is_the_gravity_step(oracle, steps):
    """With Re^-1 = 0 and no air flow, the VISCOUS assembly term (fv = g
    sin(theta), fu = 0) is the gravity term and the viscous updates are the
    plain ones: the two steps agree to rounding."""
    n, dt = 16, 0.005
    st0, ua, va = make_case(n, 21, l_max=4, u_scale=0.5)
    out = []
    for visc in (False, True):
        p = oracle.params(n, dt, gravity=True, airflow=False, viscous=visc)
        p.re = 0.0
        st = {k: v.copy() for k, v in st0.items()}
        oracle.step(p, st, ua, va, steps)
        out.append(st)
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        a, b = out[0][k], out[1][k]
        assert np.max(np.abs(a - b)) <= 1e-11 * max(1.0, np.max(np.abs(a))), k


def test_viscous_step_differs_with_viscosity(oracle):
    """The live path: with the reference's Re^-1 the step moves away from the
    plain one, by more than rounding and by less than the state's size."""
    n, dt = 16, 0.005
    st0, ua, va = make_case(n, 22, l_max=4, u_scale=0.5)
    out = []
    for visc in (False, True):
        p = oracle.params(n, dt, gravity=True, airflow=False, viscous=visc)
        st = {k: v.copy() for k, v in st0.items()}
        oracle.step(p, st, ua, va, 2)
        out.append(st)
    dif = np.max(np.abs(out[0]["u_theta"] - out[1]["u_theta"]))
    assert 1e-10 < dif < 0.5 * np.max(np.abs(out[0]["u_theta"]))
