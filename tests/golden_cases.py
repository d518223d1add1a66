"""Loader of the golden fixtures under tests/golden/ (see make_golden.py)."""
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("u_theta", "u_phi", "eta", "gamma")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def load(name):
    """(case dict, inputs (state, uair0, vair0), {steps: state})"""
    from cases import make_case
    case = manifest()["cases"][name]
    st, ua, va = make_case(case["n_theta"], case["seed"], u_scale=case["u_scale"])
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        outs = {s: {k: z[f"s{s}_{k}"].copy() for k in FIELDS} for s in case["steps"]}
    return case, (st, ua, va), outs
