"""Grids whose row length is not a power of two (needs an MI355X), as the
reference lays out any numberOfGridsInTheta (solver.cu:225-240, main.cu:28):

- 200 x 400 (rows of 4 cells a thread, 100 interface unknowns per row, a
  non-power-of-two reduced line system) and 384 x 768 (192 unknowns), in
  FP64 and mixed precision, eager and graph-replayed;
- 101 x 202 (odd n_theta: 2 cells a thread, 101 unknowns; cell tiles wider
  than the row) and the smallest such grid, 6 x 12;
- 1000 x 2000 and 1536 x 3072 (rows of 4 and 8 cells a thread for the
  Chebyshev iterations, 500 and 384 unknowns);
- a packed ensemble mixing power-of-two and other bubbles (the advection
  then takes its general phi wrap for all of them);
every field against the FP64 oracle, which has no size restriction."""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import copy_state, load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u_theta", "u_phi", "eta", "gamma")


def _say(msg):
    print(msg, flush=True)


def _check(oracle, n, dt, prec, graph, steps, tol, seed=7):
    _say(f"general size {n}x{2 * n} {prec} graph={graph} steps={steps}")
    st, ua, va = make_case(n, seed, l_max=min(8, n // 2), u_scale=0.5)
    ref = copy_state(st)
    oracle.step(oracle.params(n, dt), ref, ua, va, steps)
    with sb.SoapFilm(sb.bubble(n, dt), prec, use_graph=graph) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        ctx.synchronize()
        got = ctx.state(0)
        s = ctx.stats(0)
    assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, s
    for k in FIELDS:
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.2e}")
        assert e < tol, (n, prec, k, e)


@pytest.mark.parametrize("n", [200, 384])
@pytest.mark.parametrize("prec,graph,tol", [("fp64", False, 1e-9), ("mixed", False, 5e-5),
                                            ("mixed", True, 5e-5)])
def test_non_power_of_two(oracle, n, prec, graph, tol):
    _check(oracle, n, 0.003, prec, graph, 10, tol)


@pytest.mark.parametrize("n", [101, 6])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("mixed", 5e-5)])
def test_odd_and_tiny(oracle, n, prec, tol):
    _check(oracle, n, 0.005, prec, False, 8, tol)


@pytest.mark.parametrize("n,steps", [(1000, 3), (1536, 2)])
def test_wide_non_power_of_two_rows(oracle, n, steps):
    _check(oracle, n, 0.001, "mixed", True, steps, 5e-5)
    _check(oracle, n, 0.001, "fp64", False, 1, 1e-9)


def test_mixed_sizes_ensemble(oracle):
    sizes = [(64, 0.005), (200, 0.003), (101, 0.004), (6, 0.01), (128, 0.004)]
    _say("ensemble of power-of-two and other sizes, mixed graph")
    cases = []
    for k, (n, dt) in enumerate(sizes):
        st, ua, va = make_case(n, 40 + k, l_max=min(8, n // 2), u_scale=0.5)
        cases.append((n, dt, st, ua, va))
    steps = 8
    with sb.SoapFilm([sb.bubble(n, dt) for n, dt, *_ in cases], "mixed", use_graph=True) as ctx:
        for k, (n, dt, st, ua, va) in enumerate(cases):
            load(ctx, k, st, ua, va)
        ctx.step(steps)
        ctx.synchronize()
        got = [ctx.state(k) for k in range(len(cases))]
        stats = [ctx.stats(k) for k in range(len(cases))]
    for k, (n, dt, st, ua, va) in enumerate(cases):
        assert stats[k]["failed_solves"] == 0, (k, stats[k])
        ref = copy_state(st)
        oracle.step(oracle.params(n, dt), ref, ua, va, steps)
        e = max(rel_l2(got[k][f], ref[f]) for f in FIELDS)
        _say(f"  bubble {k} ({n}x{2 * n}): worst field rel L2 {e:.2e}")
        assert e < 5e-5, (k, n, e)
