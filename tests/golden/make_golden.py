"""Writes the golden fixtures of the soap-film step: oracle states after 1, 3
and 20 steps from seeded synthetic inputs.

    python tests/golden/make_golden.py

Inputs are regenerated from their seeds by soapbubble_amd.synth_fields (a
host-only function of the library); the manifest records the sha256 of the
inputs so a change of the generator is caught, and of every stored output.
Outputs are FP64, little-endian, in the reference's index convention
(row-major by theta, phi contiguous).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

FIELDS = ("u_theta", "u_phi", "eta", "gamma")

# name: n_theta, dt (s), seed, initial velocity scale, steps recorded
CASES = {
    "g8": dict(n_theta=8, dt=0.01, seed=16, u_scale=0.8, steps=(1, 3)),
    "g16": dict(n_theta=16, dt=0.005, seed=32, u_scale=0.8, steps=(1, 3)),
    "c0": dict(n_theta=64, dt=0.005, seed=0, u_scale=0.0, steps=(20,)),
}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def inputs(case):
    from cases import make_case
    return make_case(case["n_theta"], case["seed"], u_scale=case["u_scale"])


def main():
    from cases import copy_state
    from oracle import oracle as O

    manifest = {"solver_tol": 1e-12, "cases": {}}
    for name, case in CASES.items():
        st, ua, va = inputs(case)
        entry = dict(case, steps=list(case["steps"]),
                     inputs={k: sha(v) for k, v in dict(st, uair0=ua, vair0=va).items()},
                     outputs={})
        p = O.params(case["n_theta"], case["dt"], tol=1e-12)
        cur = copy_state(st)
        arrays = {}
        done = 0
        for s in case["steps"]:
            O.step(p, cur, ua, va, s - done)
            done = s
            for k in FIELDS:
                key = f"s{s}_{k}"
                arrays[key] = cur[k].astype("<f8")
                entry["outputs"][key] = sha(arrays[key])
        np.savez(os.path.join(HERE, f"{name}.npz"), **arrays)
        manifest["cases"][name] = entry
        print(name, "written", flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
