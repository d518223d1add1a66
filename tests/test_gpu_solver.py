"""The Gamma solve's initial guess and the adaptive inner-iteration schedule of
the captured step (needs an MI355X). Neither changes what the solve
converges to: every solve still meets the FP64 relative residual 1e-10, so
states agree to the accumulated rounding of the stored fields."""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u_theta", "u_phi", "eta", "gamma")


def _say(msg):
    print(msg, flush=True)


# Two solves that both meet the residual bound differ by up to about the
# condition number times 1e-10 in Gamma; u carries its finite differences
# (a factor 1/h more), hence 1e-8 for FP64.
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-8), ("mixed", 2e-6)])
def test_initial_guess_changes_only_the_iterate(prec, tol):
    n, dt, steps = 128, 0.004, 30
    _say(f"initial guess {prec} {n}x{2 * n} {steps} steps")
    st, ua, va = make_case(n, 7, l_max=8, u_scale=0.5)
    out = {}
    for guess in ("gamma_star", "extrapolated"):
        with sb.SoapFilm(sb.bubble(n, dt), prec, initial_guess=guess) as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(steps)
            s = ctx.stats(0)
            assert s["failed_solves"] == 0, guess
            assert s["max_residual"] <= 1e-10, guess
            out[guess] = (ctx.state(0), s)
    for k in FIELDS:
        assert rel_l2(out["extrapolated"][0][k], out["gamma_star"][0][k]) < tol, k
    # the extrapolated iterate starts closer: no more inner iterations
    assert out["extrapolated"][1]["iterations"] <= out["gamma_star"][1]["iterations"]


def test_adaptive_schedule_matches_eager():
    """Graph replays whose inner schedule shrinks as the guess improves agree
    with host-driven eager stepping, and a state upload that drops the stored
    corrections mid-run is absorbed without a failed solve."""
    n, dt = 128, 0.004
    _say(f"adaptive schedule {n}x{2 * n}")
    st, ua, va = make_case(n, 11, l_max=8, u_scale=0.5)
    out = []
    for use_graph in (False, True):
        with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=use_graph) as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(40)
            ctx.synchronize()
            mid = ctx.state(0)
            # re-upload Gamma: the stored solve corrections are dropped
            ctx.upload(0, "gamma", mid["gamma"])
            ctx.step(24)
            ctx.synchronize()
            s = ctx.stats(0)
            assert s["failed_solves"] == 0, use_graph
            assert s["max_residual"] <= 1e-10, use_graph
            out.append(ctx.state(0))
    for k in FIELDS:
        assert rel_l2(out[1][k], out[0][k]) < 2e-6, k


def test_adaptive_schedule_ensemble():
    """A packed ensemble of uneven bubbles under graph replay: the shared
    schedule follows the slowest bubble and no solve fails."""
    n_list = [32, 64, 128]
    dts = [0.008, 0.006, 0.004]
    _say("adaptive schedule ensemble")
    descs = [sb.bubble(n, dt) for n, dt in zip(n_list, dts)]
    cases = [make_case(n, 20 + q, l_max=6, u_scale=0.5) for q, n in enumerate(n_list)]
    with sb.SoapFilm(descs, "mixed", use_graph=True) as ctx:
        for b, (st, ua, va) in enumerate(cases):
            load(ctx, b, st, ua, va)
        ctx.step(48)
        ctx.synchronize()
        for b in range(len(n_list)):
            s = ctx.stats(b)
            assert s["failed_solves"] == 0, b
            assert s["max_residual"] <= 1e-10, b
            assert np.isfinite(ctx.state(b)["gamma"]).all()


def test_last_refinement_applies_the_line_solve():
    """The captured step's last refinement runs no inner iteration: when the
    refinements before it miss the tolerance it corrects x by the line solve
    z = M^-1 r of the residual before it. From Gamma* (no extrapolated
    guess) one inner iteration a refinement cannot reach 1e-10, so the second
    refinement runs in every step and must lower the residual the step
    reports (by about the line-Jacobi bound mu of the grid, below 1/2),
    against the same step with one refinement. `outer` counts the FP64
    residual evaluations of the last solve, the assembly's included."""
    n, dt, steps = 64, 0.005, 8
    _say(f"last refinement {n}x{2 * n}")
    st, ua, va = make_case(n, 13, l_max=6, u_scale=0.5)
    res = {}
    for outer in (1, 2):
        with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=True, max_inner=1,
                         max_outer=outer, initial_guess="gamma_star") as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(steps)
            ctx.synchronize()
            s = ctx.stats(0)
            assert np.isfinite(ctx.state(0)["gamma"]).all()
        _say(f"  max_outer {outer}: residual {s['residual']:.3e}, outer {s['outer']}, "
             f"failed {s['failed_solves']}")
        assert s["outer"] == outer + 1
        res[outer] = (s["residual"], s["failed_solves"])
    assert res[1][1] == steps               # one refinement of one iteration: every solve misses
    assert 0.0 < res[2][0] < 0.5 * res[1][0]
    assert res[2][1] <= res[1][1]


def test_unreachable_tolerance_fails_and_stays_finite():
    """A tolerance the refinement cannot reach: every solve is counted as
    failed and the state stays finite. The FP32 inner solve stops where it
    stagnates at its rounding floor (|r|^2 falling by less than 2 over eight
    iterations) instead of spending its whole budget, so the calibrated
    schedule stays short enough to capture."""
    n, dt = 32, 0.008
    _say(f"unreachable tolerance {n}x{2 * n}")
    st, ua, va = make_case(n, 17, l_max=4, u_scale=0.5)
    with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=True, tol=1e-20,
                     inner_rho=1e-30) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(3)
        s = ctx.stats(0)
        _say(f"  residual {s['residual']:.3e}, iterations {s['iterations']}, failed {s['failed_solves']}")
        assert s["failed_solves"] == 3
        assert np.isfinite(s["residual"]) and 1e-20 < s["residual"] < 1e-9
        assert s["iterations"] < 4 * 48
        ctx.step(2)
        s = ctx.stats(0)
        assert s["failed_solves"] == 5
        for k in FIELDS:
            assert np.isfinite(ctx.state(0)[k]).all(), k


def test_long_chebyshev_schedule_runs_host_driven():
    """Chebyshev chosen explicitly in the stiff regime of the reference's
    config.txt (256 x 512, radius 0.05, 0.005 s: line-Jacobi bound 0.991)
    runs up to its 200 iterations a refinement, more than a captured step
    holds (48): the steps run host-driven and the state stays finite. (Its
    FP32 iteration stalls near 5e-7 there, which is why the automatic inner
    method switches to multigrid in this regime: tests/test_gpu_multigrid.py.)"""
    n, dt, radius, steps = 256, 0.005, 0.05, 2
    _say(f"explicit Chebyshev, stiff {n}x{2 * n}")
    st, ua, va = make_case(n, 18, l_max=8, u_scale=0.3)
    with sb.SoapFilm(sb.bubble(n, dt, radius=radius), "mixed", use_graph=True,
                     inner_method="chebyshev") as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        ctx.synchronize()
        got = ctx.state(0)
        s = ctx.stats(0)
    _say(f"  iterations {s['iterations']}, failed {s['failed_solves']}, max residual {s['max_residual']:.2e}")
    assert s["iterations"] > 48
    assert np.isfinite(s["max_residual"]) and s["max_residual"] < 1e-3
    for k in FIELDS:
        assert np.isfinite(got[k]).all(), k


def test_set_dt_under_graph_replay(oracle):
    """A step size change between captured runs (the frame loop's partial
    steps): the stored guess and the inner schedule are dropped and set up
    again, every solve still meets the tolerance, and the states follow the
    host-driven run and the oracle with the same step sequence."""
    n, dt1, dt2 = 64, 0.004, 0.008
    _say(f"set_dt under graph replay {n}x{2 * n}")
    st, ua, va = make_case(n, 19, l_max=6, u_scale=0.5)
    out = []
    for use_graph in (False, True):
        with sb.SoapFilm(sb.bubble(n, dt1), "mixed", use_graph=use_graph) as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(12)
            ctx.set_dt(dt2)
            ctx.step(12)
            ctx.set_dt(dt1)
            ctx.step(6)
            ctx.synchronize()
            s = ctx.stats(0)
            assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, use_graph
            out.append(ctx.state(0))
    ref = {k: v.copy() for k, v in st.items()}
    t = 0.0
    for dt, steps in ((dt1, 12), (dt2, 12), (dt1, 6)):
        p = oracle.params(n, dt, time_s=t)
        oracle.step(p, ref, ua, va, steps)
        t = p.time_s
    for k in FIELDS:
        assert rel_l2(out[1][k], out[0][k]) < 2e-6, k
        e = rel_l2(out[1][k], ref[k])
        _say(f"  {k}: rel L2 vs oracle {e:.3e}")
        assert e < 5e-5, k


@pytest.mark.parametrize("n,prec,graph", [(64, "mixed", False), (64, "fp64", False),
                                          (1024, "mixed", False), (1024, "mixed", True)])
def test_inner_timing_leaves_the_step_unchanged(n, prec, graph):
    """sf_time_inner between the assembly and the solve (the measurement of
    the inner launch) changes nothing the step computes: the state after the
    step is bitwise the one of the same step without it, and so are the
    solve statistics. The solve's convergence tests read sums accumulated by
    atomics in no fixed order, but no test lies within rounding of its
    bound here, so the decisions and hence every value are the same. With
    graph replays before it (a captured context after its host-driven
    steps), the timing follows the assembly with no host call in between
    that waits for the stream: its save and restore of the control blocks
    must be ordered with the queued assembly."""
    dt = 0.004 if n == 64 else 0.001
    _say(f"inner timing {prec} {n}x{2 * n} graph={graph}")
    st, ua, va = make_case(n, 23, l_max=8, u_scale=0.5)
    out = []
    for timed in (False, True):
        with sb.SoapFilm(sb.bubble(n, dt), prec, inner_method="chebyshev",
                         use_graph=graph) as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(10 if graph else 2)
            ctx.advect()
            ctx.assemble()
            if timed:
                ms = ctx.time_inner(5)
                _say(f"  inner launch {ms * 1e3:.2f} us")
                assert ms > 0
            ctx.solve()
            ctx.update()
            ctx.step(1)
            ctx.synchronize()
            out.append((ctx.state(0), ctx.stats(0)))
    for k in FIELDS:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    for k in ("iterations", "outer", "failed_solves"):
        assert out[0][1][k] == out[1][1][k], k


def test_inner_timing_checks_its_place():
    """sf_time_inner outside its place in the step (SF_ERR_ORDER), or with
    the PCG inner method (SF_ERR_INVALID), is refused with a status and
    leaves the context usable."""
    n = 32
    st, ua, va = make_case(n, 3, l_max=4, u_scale=0.5)
    with sb.SoapFilm(sb.bubble(n, 0.008), "mixed") as ctx:
        load(ctx, 0, st, ua, va)
        with pytest.raises(sb.SoapFilmError) as e:
            ctx.time_inner(3)
        assert e.value.code == 6
        ctx.step(1)
    with sb.SoapFilm(sb.bubble(n, 0.008), "mixed", inner_method="cg") as ctx:
        load(ctx, 0, st, ua, va)
        ctx.advect()
        ctx.assemble()
        with pytest.raises(sb.SoapFilmError) as e:
            ctx.time_inner(3)
        assert e.value.code == 1
        ctx.solve()
        ctx.update()


def test_graph_cache_eviction_changes_nothing(monkeypatch):
    """A one-executable graph cache (the test hook SF_GRAPH_CACHE_MAX=1:
    every capture evicts the previous executable after the stream drains)
    against the default cache, through schedule changes and step-size
    changes: the same kernels run in the same order, so the states agree
    bitwise."""
    n, dt1, dt2 = 64, 0.004, 0.008
    _say(f"graph cache eviction {n}x{2 * n}")
    st, ua, va = make_case(n, 23, l_max=6, u_scale=0.5)
    out = []
    for cache in (None, "1"):
        if cache is None:
            monkeypatch.delenv("SF_GRAPH_CACHE_MAX", raising=False)
        else:
            monkeypatch.setenv("SF_GRAPH_CACHE_MAX", cache)
        with sb.SoapFilm(sb.bubble(n, dt1), "mixed", use_graph=True) as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(20)
            ctx.set_dt(dt2)
            ctx.step(10)
            ctx.set_dt(dt1)
            ctx.step(10)
            ctx.synchronize()
            s = ctx.stats(0)
            assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, cache
            out.append(ctx.state(0))
    for k in FIELDS:
        assert np.array_equal(out[0][k], out[1][k]), k


def _line_solve_rows(diag, w_west, r, n, nph):
    """z = M^-1 r for the periodic phi-line blocks M of the operator:
    diag z_i - wW_i z_(i-1) - wW_(i+1) z_(i+1) = r_i on every theta row."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl

    blocks = []
    for j in range(n):
        d = diag[j * nph:(j + 1) * nph]
        ww = w_west[j * nph:(j + 1) * nph]
        we = np.roll(ww, -1)
        m = sps.diags([d, -ww[1:], -we[:-1]], [0, -1, 1], shape=(nph, nph), format="lil")
        m[0, nph - 1] = -ww[0]
        m[nph - 1, 0] = -we[nph - 1]
        blocks.append(m.tocsr())
    return spl.spsolve(sps.block_diag(blocks, format="csc"), r)


@pytest.mark.parametrize("n,dt", [(64, 0.004), (256, 0.002), (1024, 0.001), (100, 0.003),
                                  (200, 0.002)])
def test_spare_refinement_forms_its_correction(n, dt):
    """The captured step's last refinement of no inner iteration corrects
    x_1 by z = M^-1 r_1, the phi-line solve of the FP32 residual before it,
    which its residual launch forms itself on its rows and one row either
    side. Two contexts take the same replayed step from the same state: one
    refinement of one Chebyshev launch (its solution is x_1), and the same
    plus the spare refinement (x_2). x_2 - x_1 must be M^-1 r_1, with r_1 =
    b - A x_1 in FP64 (A assembled by the oracle from the step's advected
    fields, b the library's), rounded to FP32, and M the phi-line blocks of
    the stored FP32 operator; the spare must have run (three FP64 residual
    evaluations). Rows of 128 cells pack several to a block, 512-cell rows
    two, 2048-cell rows one; rows of 200 and 400 cells (50 and 100 line
    unknowns) take the general line solve and the non-power-of-two phi
    wrap."""
    from oracle import oracle as O

    nph = 2 * n
    _say(f"spare refinement {n}x{nph}")
    st, ua, va = make_case(n, 29, l_max=8, u_scale=0.5)
    names = ("solution", "rhs", "diag", "w_west", "u_theta_star", "u_phi_star", "eta_star",
             "gamma_star")
    got = {}
    for outer in (1, 2):
        with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=True, inner_method="chebyshev",
                         max_inner=1, max_outer=outer, initial_guess="gamma_star") as ctx:
            load(ctx, 0, st, ua, va)
            ctx.step(3)
            # the same state again: the next step replays the captured graph
            load(ctx, 0, st, ua, va)
            ctx.step(1)
            ctx.synchronize()
            got[outer] = {f: ctx.download(0, f) for f in names}
            got[outer]["outer"] = ctx.stats(0)["outer"]
            got[outer]["residual"] = ctx.stats(0)["residual"]
    assert got[2]["outer"] == 3, got[2]["outer"]
    for f in names[1:]:
        assert np.array_equal(got[1][f], got[2][f]), f
    g = got[1]
    x1, x2 = g["solution"], got[2]["solution"]
    star = {k: g[k + "_star"] for k in ("u_theta", "u_phi", "eta", "gamma")}
    p = O.params(n, dt)
    N, Nt = O.sizes(n)
    val, _, _ = O.assemble(p, star, np.zeros(N), np.zeros(Nt))
    r = (g["rhs"] - O.spmv(p, val, x1)).astype(np.float32).astype(np.float64)
    z = _line_solve_rows(g["diag"], g["w_west"], r, n, nph)
    e = rel_l2(x2 - x1, z)
    _say(f"  |r_1| / |b| {np.linalg.norm(r) / np.linalg.norm(g['rhs']):.2e}, "
         f"rel L2 of x_2 - x_1 against M^-1 r_1: {e:.2e}")
    assert np.linalg.norm(z) > 0
    assert e < 1e-3
    # the residual of x_2 that the launch evaluated: A x_2 takes the
    # correction of the rows either side and of the neighbouring threads
    res = np.linalg.norm(g["rhs"] - O.spmv(p, val, x2)) / np.linalg.norm(g["rhs"])
    _say(f"  residual of x_2: library {got[2]['residual']:.4e}, oracle operator {res:.4e}")
    assert abs(got[2]["residual"] - res) <= 1e-3 * res
