"""The autotuner (sf_options.autotune; needs an MI355X): at the first
sf_step call after 16 steps it times the advection's cell tiles and each row
width's Chebyshev launch geometry on the context's own data, keeps the
fastest and captures the step again; whatever it picks, the steps still
match the oracle."""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import copy_state, load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u_theta", "u_phi", "eta", "gamma")


def test_autotuned_ensemble_matches_oracle(oracle):
    sizes = [(64, 0.005), (128, 0.004), (256, 0.002), (200, 0.003)]
    cases = [make_case(n, 70 + k, l_max=8, u_scale=0.5) for k, (n, _) in enumerate(sizes)]
    steps = 24
    with sb.SoapFilm([sb.bubble(n, dt) for n, dt in sizes], "mixed", use_graph=True,
                     autotune=True) as ctx:
        for k, c in enumerate(cases):
            load(ctx, k, *c)
        ctx.step(16)
        ctx.step(steps - 16)
        ctx.synchronize()
        got = [ctx.state(k) for k in range(len(sizes))]
        stats = [ctx.stats(k) for k in range(len(sizes))]
        tun = ctx.tuning()
    print(f"tuning: {tun}", flush=True)
    assert tun["tuned"] and tun["advect_tile_width"] in (256, 128, 64)
    # one launch group: rows of 128, 256, 400 and 512 cells share 4 cells a
    # thread (the group is reported by its widest row)
    assert [r["n_phi"] for r in tun["rows"]] == [512]
    for r in tun["rows"]:
        assert r["inner_cells_per_thread"] in (2, 4, 8, 16)
        assert max(r["inner_ms"]) > 0
    for k, ((n, dt), (st, ua, va)) in enumerate(zip(sizes, cases)):
        assert stats[k]["failed_solves"] == 0, (k, stats[k])
        ref = copy_state(st)
        oracle.step(oracle.params(n, dt), ref, ua, va, steps)
        e = max(rel_l2(got[k][f], ref[f]) for f in FIELDS)
        assert e < 5e-5, (k, n, e)


def test_autotune_headline_size():
    """The 1024 x 2048 workload: the tuner's timings and choices (printed),
    and the steps meet the solver tolerance."""
    n, dt = 1024, 0.001
    f = sb.synth_fields(n, 2000, 64, 12, 0.05, 1.0)
    nph = 2 * n
    st = {"u_theta": np.zeros((n - 1) * nph), "u_phi": np.zeros(n * nph),
          "eta": f["eta"], "gamma": f["gamma"]}
    with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=True, autotune=True) as ctx:
        load(ctx, 0, st, f["uair0"], f["vair0"])
        ctx.step(20)
        ctx.step(12)
        ctx.synchronize()
        s = ctx.stats(0)
        tun = ctx.tuning()
    print(f"tuning 1024x2048: {tun}", flush=True)
    assert tun["tuned"]
    assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, s
