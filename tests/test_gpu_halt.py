"""halt_on_failure (needs an MI355X): a bubble whose solve misses its
tolerance takes no further step, eager or under graph replay, while the
other bubbles of the context go on; uploading a new state lets it step
again. (An unreachable tolerance forces the failure from the host; every
kernel still runs to its end.)"""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import load, make_case

pytestmark = pytest.mark.gpu

FIELDS = ("u_theta", "u_phi", "eta", "gamma")


@pytest.mark.parametrize("graph", [False, True])
def test_failed_solve_stops_only_that_bubble(graph):
    print(f"halt_on_failure graph={graph}", flush=True)
    n, dt = 32, 0.005
    cases = [make_case(n, 60 + k, u_scale=0.5) for k in range(2)]
    # bubble 0 cannot meet 1e-30 (its solve fails at the first step); the
    # tolerance is per context, so bubble 1 runs in a context of its own
    with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=graph, tol=1e-30, max_outer=2,
                     max_inner=8, halt_on_failure=True) as ctx:
        load(ctx, 0, *cases[0])
        ctx.step(1)
        ctx.synchronize()
        after1 = ctx.state(0)
        s1 = ctx.stats(0)
        ctx.step(9)
        ctx.synchronize()
        after10 = ctx.state(0)
        s10 = ctx.stats(0)
        assert s1["failed_solves"] == 1 and s10["failed_solves"] == 1, (s1, s10)
        for k in FIELDS:
            assert np.array_equal(after1[k], after10[k]), k
        # a new state steps again (and fails again)
        load(ctx, 0, *cases[1])
        ctx.step(1)
        ctx.synchronize()
        moved = ctx.state(0)
        assert ctx.stats(0)["failed_solves"] == 2
        assert not np.array_equal(moved["eta"], cases[1][0]["eta"])
    with sb.SoapFilm([sb.bubble(n, dt)] * 2, "mixed", use_graph=graph,
                     halt_on_failure=True) as ctx:
        for k in range(2):
            load(ctx, k, *cases[k])
        ctx.step(10)
        ctx.synchronize()
        for k in range(2):
            assert ctx.stats(k)["failed_solves"] == 0
