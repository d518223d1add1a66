"""BIMOCQ scalar advection on the GPU against the oracle's restatement
(bubble.cu:512-633, 967-1013, 2081-2114, 2180-2280, 2509-2564;
solver.cu:260-292)."""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import copy_state, load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run_pair(oracle, n, dt, steps, prec, threshold, seed=3, multisample=True, blend=0.5):
    st, ua, va = make_case(n, seed, u_scale=0.8)
    p = oracle.params(n, dt)
    ref = copy_state(st)
    bm = oracle.BimocqState(p, ref["eta"], ref["gamma"], blend=blend, threshold=threshold,
                            multisample=multisample)
    hist = oracle.bimocq_step(p, ref, ua, va, bm, steps)
    with sb.SoapFilm(sb.bubble(n, dt), prec, tol=1e-12 if prec == "fp64" else 1e-10,
                     scalar_advection="bimocq", blend=blend, multisample=multisample,
                     reinit_threshold=threshold) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        got = ctx.state(0)
        s = ctx.stats(0)
    return ref, got, s, bm, hist


@pytest.mark.parametrize("threshold", [4.0, 0.01])
def test_bimocq_fp64_matches_oracle(oracle, threshold):
    _say(f"BIMOCQ fp64 32x64, 6 steps, reinit threshold {threshold}")
    ref, got, s, bm, hist = _run_pair(oracle, 32, 0.005, 6, "fp64", threshold)
    assert s["failed_solves"] == 0
    assert s["reinits"] == bm.reinit_count
    if threshold < 1:
        assert bm.reinit_count >= 1
    assert abs(s["distortion"] - bm.distortion) <= 1e-6 * max(1.0, bm.distortion)
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-9, (k, e)


def test_bimocq_single_sample_fp64(oracle):
    _say("BIMOCQ fp64 single sample, blend 0.3")
    ref, got, s, bm, _ = _run_pair(oracle, 16, 0.01, 4, "fp64", 0.02, seed=7, multisample=False,
                                   blend=0.3)
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        assert rel_l2(got[k], ref[k]) < 1e-9, k


def test_bimocq_mixed_matches_oracle(oracle):
    _say("BIMOCQ mixed 64x128, 6 steps")
    ref, got, s, bm, _ = _run_pair(oracle, 64, 0.004, 6, "mixed", 0.02)
    assert s["failed_solves"] == 0
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-4, (k, e)


def test_bimocq_ensemble_and_graph(oracle):
    """Two bubbles of different sizes with BIMOCQ, graph replay against the
    oracle run per bubble."""
    _say("BIMOCQ ensemble 16 + 32, mixed, graph")
    specs = [(16, 0.006, 11), (32, 0.004, 12)]
    steps = 5
    descs, inputs, refs = [], [], []
    for n, dt, seed in specs:
        st, ua, va = make_case(n, seed, u_scale=0.7)
        descs.append(sb.bubble(n, dt))
        inputs.append((st, ua, va))
        p = oracle.params(n, dt)
        ref = copy_state(st)
        bm = oracle.BimocqState(p, ref["eta"], ref["gamma"], threshold=0.01)
        oracle.bimocq_step(p, ref, ua, va, bm, steps)
        refs.append(ref)
    with sb.SoapFilm(descs, "mixed", use_graph=True, scalar_advection="bimocq",
                     reinit_threshold=0.01) as ctx:
        for b, (st, ua, va) in enumerate(inputs):
            load(ctx, b, st, ua, va)
        ctx.step(steps)
        for b, ref in enumerate(refs):
            got = ctx.state(b)
            for k in ("u_theta", "u_phi", "eta", "gamma"):
                e = rel_l2(got[k], ref[k])
                assert e < 1e-4, (b, k, e)


# At 256 x 512 (h = 0.012 rad) map values lie well within 0.02 rad of their
# cell and of each other, so the map arithmetic runs its series branches
# (the spherical interpolation as a f + b t, cell-relative index angles);
# the coarse grids above mostly take the acos/atan2 branches. Threshold 1e-4
# cells forces reinitialisations at this resolution.
@pytest.mark.parametrize("threshold", [4.0, 1e-4])
def test_bimocq_fine_grid_fp64(oracle, threshold):
    _say(f"BIMOCQ fp64 256x512, 3 steps, reinit threshold {threshold}")
    ref, got, s, bm, _ = _run_pair(oracle, 256, 0.002, 3, "fp64", threshold, seed=21)
    assert s["failed_solves"] == 0
    assert s["reinits"] == bm.reinit_count
    if threshold < 1:
        assert bm.reinit_count >= 1
    assert abs(s["distortion"] - bm.distortion) <= 1e-6 * max(1.0, bm.distortion)
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-9, (k, e)


def test_bimocq_fine_grid_mixed(oracle):
    _say("BIMOCQ mixed 256x512, 4 steps")
    ref, got, s, bm, _ = _run_pair(oracle, 256, 0.002, 4, "mixed", 4.0, seed=22)
    assert s["failed_solves"] == 0
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-4, (k, e)


@pytest.mark.parametrize("threshold", [4.0, 0.01])
def test_bimocq_smallest_grid_fp64(oracle, threshold):
    """4 x 8: every row is a polar row; maps cross the poles."""
    _say(f"BIMOCQ fp64 4x8, 6 steps, reinit threshold {threshold}")
    ref, got, s, bm, _ = _run_pair(oracle, 4, 0.01, 6, "fp64", threshold, seed=23)
    assert s["failed_solves"] == 0
    assert s["reinits"] == bm.reinit_count
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-9, (k, e)


@pytest.mark.parametrize("use_graph", [True, False])
def test_bimocq_eta_upload_restarts_maps(oracle, use_graph):
    """Uploading eta between steps restarts BIMOCQ from the uploaded field
    (identity maps, initial fields = the upload), also where the steps
    replay captured graphs: with the previous solve's Gamma* as the initial
    guess the second sf_step call replays graphs from its first step."""
    _say(f"BIMOCQ mixed 32x64, 6 steps, eta upload, 6 steps, graph {use_graph}")
    n, dt, steps = 32, 0.005, 6
    st, ua, va = make_case(n, 31, u_scale=0.8)
    eta2 = sb.synth_fields(n, 32, 4, 6, 0.0, 1.0)["eta"]
    p = oracle.params(n, dt)
    ref = copy_state(st)
    bm = oracle.BimocqState(p, ref["eta"], ref["gamma"], threshold=0.01)
    oracle.bimocq_step(p, ref, ua, va, bm, steps)
    ref["eta"] = eta2.copy()
    bm = oracle.BimocqState(p, ref["eta"], ref["gamma"], threshold=0.01)
    oracle.bimocq_step(p, ref, ua, va, bm, steps)
    with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=use_graph, initial_guess="gamma_star",
                     scalar_advection="bimocq", reinit_threshold=0.01) as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(steps)
        ctx.upload(0, "eta", eta2)
        ctx.step(steps)
        got = ctx.state(0)
        s = ctx.stats(0)
    assert s["failed_solves"] == 0
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-4, (k, e)
