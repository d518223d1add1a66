"""GPU parity of the VISCOUS path (header.cuh:99, off in the reference) against
the oracle's restatement (needs an MI355X): steps with the viscous forces in
the assembly and the velocity updates, at the reference's Re^-1 and at a
larger one that drives the +-0.1 clamp."""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import copy_state, load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


# dt scales with the radius so that the non-dimensional step stays 0.004;
# at radius 0.01 (Re^-1 = 0.0125) the +-0.1 clamp acts, and the reference's
# float clamp turns last-bit differences of its argument into float-ulp
# steps of the force, hence 1e-8 there.
@pytest.mark.parametrize("prec,radius,tol", [("fp64", 1.0, 1e-9), ("fp64", 0.01, 1e-8),
                                             ("mixed", 1.0, 5e-5)])
def test_viscous_steps(oracle, prec, radius, tol):
    n, dt, steps = 32, 0.004 * radius, 8
    _say(f"viscous {prec} radius {radius} {n}x{2 * n} {steps} steps")
    st, ua, va = make_case(n, 31, l_max=4, u_scale=0.5)
    p = oracle.params(n, dt, radius=radius, gravity=True, viscous=True)
    ref = copy_state(st)
    oracle.step(p, ref, ua, va, steps)
    desc = sb.bubble(n, dt, radius=radius, gravity=True)
    assert abs(sb.derive_params(desc)["re"] - p.re) <= 1e-15 * p.re
    with sb.SoapFilm(desc, prec, viscous=True) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        got = ctx.state(0)
        s = ctx.stats(0)
        assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < tol, (k, e)


def test_viscous_changes_the_step(oracle):
    """The option is live: the viscous and the plain GPU steps differ."""
    n, dt, steps = 32, 0.00004, 4
    st, ua, va = make_case(n, 32, l_max=4, u_scale=0.5)
    out = {}
    for visc in (False, True):
        with sb.SoapFilm(sb.bubble(n, dt, radius=0.01, gravity=True), "fp64",
                         viscous=visc) as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(steps)
            out[visc] = ctx.state(0)
    assert rel_l2(out[True]["u_theta"], out[False]["u_theta"]) > 1e-6


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("mixed", 5e-5)])
def test_viscous_fine_grid(oracle, prec, tol):
    """256 x 512 (rows of 4 cells a thread in the row kernels), graph replay
    for the mixed path."""
    n, dt, steps = 256, 0.002, 4
    _say(f"viscous {prec} {n}x{2 * n} {steps} steps")
    st, ua, va = make_case(n, 33, l_max=8, u_scale=0.5)
    p = oracle.params(n, dt, gravity=True, viscous=True)
    ref = copy_state(st)
    oracle.step(p, ref, ua, va, steps)
    with sb.SoapFilm(sb.bubble(n, dt, gravity=True), prec, viscous=True,
                     use_graph=prec == "mixed", tol=1e-12 if prec == "fp64" else 1e-10) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        got = ctx.state(0)
        s = ctx.stats(0)
        assert s["failed_solves"] == 0
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < tol, (k, e)
