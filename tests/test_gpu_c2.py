"""The mixed-precision workload at its acceptance point (needs an MI355X):
one 1024x2048 bubble, dt 0.001, eta with spherical-harmonic degree <= 64 plus
5 % uniform noise, air degree <= 12, from rest; 100 steps of the mixed path
(FP32 fields and advection, FP32 inner solve inside FP64 refinement) replayed
as hipGraphs, against 100 steps of the FP64 oracle: eta and Gamma within
1e-4 relative L2, every solve at a relative residual of at most 1e-10."""
import numpy as np
import pytest

import soapbubble_amd as sb
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_c2_hundred_steps_against_oracle(oracle):
    n, dt, steps = 1024, 0.001, 100
    f = sb.synth_fields(n, 2000, 64, 12, 0.05, 1.0)
    nph = 2 * n
    st = {"u_theta": np.zeros((n - 1) * nph), "u_phi": np.zeros(n * nph),
          "eta": f["eta"], "gamma": f["gamma"]}
    print(f"c2 {n}x{nph} mixed graph, {steps} steps", flush=True)
    with sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=True) as ctx:
        ctx.load_state(0, st)
        ctx.upload(0, "uair0", f["uair0"])
        ctx.upload(0, "vair0", f["vair0"])
        ctx.step(steps)
        ctx.synchronize()
        got = ctx.state(0)
        s = ctx.stats(0)
    assert s["failed_solves"] == 0 and s["max_residual"] <= 1e-10, s
    ref = {k: v.copy() for k, v in st.items()}
    oracle.step(oracle.params(n, dt), ref, f["uair0"], f["vair0"], steps)
    errs = {k: rel_l2(got[k], ref[k]) for k in ("eta", "gamma", "u_theta", "u_phi")}
    print(f"  rel L2 after {steps} steps: {errs}", flush=True)
    assert errs["eta"] < 1e-4 and errs["gamma"] < 1e-4, errs
    assert errs["u_theta"] < 1e-4 and errs["u_phi"] < 1e-4, errs
