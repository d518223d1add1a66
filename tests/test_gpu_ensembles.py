"""Packed ensembles shaped like the batched workloads (needs an MI355X): bubbles
of uneven resolution, each with its own seed, air-flow amplitude and
Marangoni scale, advanced together under graph replay (mixed precision),
every bubble's fields against its own FP64 oracle run."""
import numpy as np
import pytest

import soapbubble_amd as sb
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(oracle, groups, steps, l_max, seed0):
    rng = np.random.default_rng(seed0)
    descs, cases = [], []
    b = 0
    for count, n, dt in groups:
        for _ in range(count):
            amp = rng.uniform(0.2, 1.0)
            ms = rng.uniform(0.5, 2.0)
            f = sb.synth_fields(n, seed0 + 16 * b, min(l_max, n // 2), 12, 0.0, amp)
            st = {"u_theta": np.zeros((n - 1) * 2 * n), "u_phi": np.zeros(n * 2 * n),
                  "eta": f["eta"], "gamma": f["gamma"]}
            descs.append(sb.bubble(n, dt, M_scale=ms))
            cases.append((n, dt, ms, st, f["uair0"], f["vair0"]))
            b += 1
    with sb.SoapFilm(descs, "mixed", use_graph=True) as ctx:
        for k, (n, dt, ms, st, ua, va) in enumerate(cases):
            ctx.load_state(k, st)
            ctx.upload(k, "uair0", ua)
            ctx.upload(k, "vair0", va)
        ctx.step(steps)
        ctx.synchronize()
        got = [ctx.state(k) for k in range(len(cases))]
        stats = [ctx.stats(k) for k in range(len(cases))]
    worst = 0.0
    for k, (n, dt, ms, st, ua, va) in enumerate(cases):
        assert stats[k]["failed_solves"] == 0 and stats[k]["max_residual"] <= 1e-10, k
        p = oracle.params(n, dt, M_scale=ms)
        ref = {key: v.copy() for key, v in st.items()}
        oracle.step(p, ref, ua, va, steps)
        e = max(rel_l2(got[k][key], ref[key]) for key in ("u_theta", "u_phi", "eta", "gamma"))
        _say(f"  bubble {k} ({n}x{2 * n}, M x{ms:.2f}): worst field rel L2 {e:.2e}")
        worst = max(worst, e)
    return worst


def test_uneven_ensemble_per_bubble():
    """Eight bubbles each at 64x128, 128x256 and 256x512 (dt 0.005, 0.004,
    0.002), 20 steps."""
    from oracle import oracle
    _say("ensemble 8 x (64, 128, 256)")
    worst = _run(oracle, [(8, 64, 0.005), (8, 128, 0.004), (8, 256, 0.002)], 20, 16, 3000)
    assert worst < 5e-5


def test_large_uneven_ensemble_per_bubble():
    """One 1024x2048 bubble (l <= 128 modes, dt 0.001) with sixteen 128x256
    ones (dt 0.004), 6 steps."""
    from oracle import oracle
    _say("ensemble 1024 + 16 x 128")
    worst = _run(oracle, [(1, 1024, 0.001), (16, 128, 0.004)], 6, 128, 4000)
    assert worst < 5e-5


def test_c4_batch_per_bubble():
    """The large uneven batch itself: four 2048x4096 bubbles (l <= 128 modes,
    dt 0.0005) and sixty-four 128x256 ones (dt 0.004) in one packed context,
    3 steps (the oracle needs about a second a step for each large bubble on
    16 host threads, so the batch's 50 steps are cut to 3 here); every
    bubble's every field against its own oracle run."""
    from oracle import oracle
    _say("c4 batch 4 x 2048 + 64 x 128")
    worst = _run(oracle, [(4, 2048, 0.0005), (64, 128, 0.004)], 3, 128, 4000)
    assert worst < 5e-5


def test_c3_batch_per_bubble():
    """The uneven batch itself: thirty-two bubbles each at 64x128, 128x256
    and 256x512 (dt 0.005, 0.004, 0.002) in one packed context (one launch
    per stage, narrow rows several to a block), 10 steps; every bubble's
    every field against its own oracle run."""
    from oracle import oracle
    _say("c3 batch 32 x (64, 128, 256)")
    worst = _run(oracle, [(32, 64, 0.005), (32, 128, 0.004), (32, 256, 0.002)], 10, 16, 3000)
    assert worst < 5e-5
