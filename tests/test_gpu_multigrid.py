"""The geometric-multigrid preconditioner of the inner PCG (needs an MI355X):
zebra phi-line relaxation with semi-coarsening in theta, V(1,1) cycles.

- The stiff regime of the reference's own config.txt: radius 0.05, 256 x 512,
  timePerStep 0.005 s (a non-dimensional step of 0.1), where the line-Jacobi
  preconditioned operator's spectral bound is 0.991 (the oracle's PCG needs
  about 150 iterations): every solve to 1e-10 in a few V-cycles, the
  captured-graph schedule, every field against the oracle.
- The synthetic 64 x 128 workload in FP64 and mixed precision.
- A packed ensemble with grids whose level hierarchy ends early (200 rows:
  200, 100, 50, 25; 101 rows: no coarser level), where the coarsest level
  takes symmetric zebra passes.
"""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import copy_state, load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u_theta", "u_phi", "eta", "gamma")


def _say(msg):
    print(msg, flush=True)


def _run(descs, cases, prec, steps, graph, **kw):
    with sb.SoapFilm(descs, prec, use_graph=graph, inner_method="mg", **kw) as ctx:
        for k, (st, ua, va) in enumerate(cases):
            load(ctx, k, st, ua, va)
        ctx.step(steps)
        ctx.synchronize()
        return [ctx.state(k) for k in range(len(cases))], [ctx.stats(k) for k in range(len(cases))]


@pytest.mark.parametrize("prec,graph,tol", [("mixed", True, 5e-5), ("mixed", False, 5e-5),
                                            ("fp64", False, 1e-9)])
def test_stiff_reference_config(oracle, prec, graph, tol):
    n, dt, radius, steps = 256, 0.005, 0.05, 6
    _say(f"stiff {n}x{2 * n} radius {radius} dt {dt} s, multigrid, {prec} graph={graph}")
    st, ua, va = make_case(n, 21, l_max=8, u_scale=0.3)
    got, stats = _run([sb.bubble(n, dt, radius=radius)], [(st, ua, va)], prec, steps, graph)
    s = stats[0]
    _say(f"  iterations of the last solve {s['iterations']}, refinements {s['outer']}, "
         f"max residual {s['max_residual']:.2e}")
    assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, s
    # a handful of V-cycles per solve, where the line-Jacobi PCG needs ~150
    assert s["iterations"] <= 12, s
    ref = copy_state(st)
    oracle.step(oracle.params(n, dt, radius=radius), ref, ua, va, steps)
    for k in FIELDS:
        e = rel_l2(got[0][k], ref[k])
        _say(f"  {k}: rel L2 {e:.2e}")
        assert e < tol, (k, e)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("mixed", 5e-5)])
def test_synthetic_small(oracle, prec, tol):
    n, dt, steps = 64, 0.005, 10
    st, ua, va = make_case(n, 22, u_scale=0.5)
    got, stats = _run([sb.bubble(n, dt)], [(st, ua, va)], prec, steps, prec == "mixed")
    assert stats[0]["failed_solves"] == 0, stats[0]
    ref = copy_state(st)
    oracle.step(oracle.params(n, dt), ref, ua, va, steps)
    for k in FIELDS:
        assert rel_l2(got[0][k], ref[k]) < tol, k


def test_short_hierarchies_ensemble(oracle):
    sizes = [(200, 0.003), (101, 0.004), (64, 0.005), (6, 0.01)]
    _say("multigrid ensemble 200, 101, 64, 6")
    cases = [make_case(n, 50 + k, l_max=min(8, n // 2), u_scale=0.5) for k, (n, _) in enumerate(sizes)]
    steps = 8
    got, stats = _run([sb.bubble(n, dt) for n, dt in sizes], cases, "mixed", steps, True)
    for k, ((n, dt), (st, ua, va)) in enumerate(zip(sizes, cases)):
        assert stats[k]["failed_solves"] == 0, (k, stats[k])
        ref = copy_state(st)
        oracle.step(oracle.params(n, dt), ref, ua, va, steps)
        e = max(rel_l2(got[k][f], ref[f]) for f in FIELDS)
        _say(f"  bubble {k} ({n}x{2 * n}): worst field rel L2 {e:.2e}")
        assert e < 5e-5, (k, n, e)


@pytest.mark.parametrize("graph", [True, False])
def test_automatic_switch_in_the_stiff_regime(oracle, graph):
    """The automatic inner method starts with Chebyshev; the first solve at
    the stiff config needs far more iterations than a captured step holds,
    so the context continues with multigrid (later solves: a handful of
    V-cycles), eager or under graph replay, and still matches the oracle."""
    n, dt, radius, steps = 256, 0.005, 0.05, 8
    _say(f"automatic inner method, stiff {n}x{2 * n}, graph={graph}")
    st, ua, va = make_case(n, 23, l_max=8, u_scale=0.3)
    with sb.SoapFilm(sb.bubble(n, dt, radius=radius), "mixed", use_graph=graph) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        ctx.synchronize()
        got = ctx.state(0)
        s = ctx.stats(0)
    _say(f"  last solve: {s['iterations']} inner iterations, {s['outer']} refinements")
    assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, s
    assert s["iterations"] <= 12, s
    ref = copy_state(st)
    oracle.step(oracle.params(n, dt, radius=radius), ref, ua, va, steps)
    for k in FIELDS:
        assert rel_l2(got[k], ref[k]) < 5e-5, k
