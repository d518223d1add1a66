"""The reference-resolution workload (needs an MI355X): one bubble 256 x 512,
dt = 0.002, 500 steps, eta0/Gamma0 from l <= 16 seeded modes, air from an
l <= 12 stream function; FP32 advection with the FP64 solve (precision
fp32_adv) and the mixed path, every field against the FP64 oracle after
the last step, and the film-volume drift of both."""
import time

import numpy as np
import pytest

import soapbubble_amd as sb
from cases import copy_state, load, make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu

N_THETA, DT, STEPS = 256, 0.002, 500


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c1_reference(oracle):
    st, ua, va = make_case(N_THETA, 100, l_max=16, l_air=12)
    p = oracle.params(N_THETA, DT)
    ref = copy_state(st)
    t0 = time.perf_counter()
    oracle.step(p, ref, ua, va, STEPS)
    _say(f"oracle {STEPS} steps {N_THETA}x{2 * N_THETA}: {time.perf_counter() - t0:.1f} s")
    return st, ua, va, ref, oracle.volume(p, st["eta"]), oracle.volume(p, ref["eta"])


# FP32 storage of the fields (rounding 6e-8 a step) integrated over 500
# steps of a flow that stretches the film: 1e-4 relative L2 bounds what the
# two single-precision paths reach against the FP64 restatement.
@pytest.mark.parametrize("prec", ["fp32_adv", "mixed"])
def test_c1_all_fields_against_oracle(c1_reference, prec):
    st, ua, va, ref, v0, v_ref = c1_reference
    _say(f"c1 {prec}")
    with sb.SoapFilm(sb.bubble(N_THETA, DT), prec, use_graph=prec == "mixed") as ctx:
        load(ctx, 0, st, ua, va)
        ctx.step(STEPS)
        ctx.synchronize()
        got = ctx.state(0)
        s = ctx.stats(0)
    assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10
    for k in ("u_theta", "u_phi", "eta", "gamma"):
        e = rel_l2(got[k], ref[k])
        _say(f"  {k}: rel L2 {e:.3e}")
        assert e < 1e-4, (k, e)
    drift_ref = (v_ref - v0) / v0
    drift_gpu = (s["volume"] - v0) / v0
    _say(f"  volume drift: oracle {drift_ref:.3e}, GPU {drift_gpu:.3e}")
    assert abs(drift_gpu - drift_ref) < 1e-5
