"""Shared seeded inputs of the parity tests."""
import numpy as np

import soapbubble_amd as sb


def make_case(n_theta, seed, l_max=4, l_air=6, noise=0.0, u_scale=0.0, air_amp=1.0):
    """Synthetic state (u_theta, u_phi, eta, Gamma) and air fields. With
    u_scale > 0 the initial velocity is u_scale times the air field, which
    gives the advection stages a non-trivial flow."""
    f = sb.synth_fields(n_theta, seed, l_max, l_air, noise, air_amp)
    st = {"u_theta": u_scale * f["vair0"], "u_phi": u_scale * f["uair0"],
          "eta": f["eta"].copy(), "gamma": f["gamma"].copy()}
    return st, f["uair0"], f["vair0"]


def copy_state(st):
    return {k: v.copy() for k, v in st.items()}


def load(ctx, b, st, ua, va):
    ctx.load_state(b, st)
    ctx.upload(b, "uair0", ua)
    ctx.upload(b, "vair0", va)


def csr_parts(val):
    """Split the reference's 5-entry rows (up, left, centre, right, down) into
    the north/west face weights and the diagonal."""
    v = val.reshape(-1, 5)
    return -v[:, 0], -v[:, 1], v[:, 2], -v[:, 3], -v[:, 4]
