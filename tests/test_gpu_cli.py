"""The soapBubble driver end to end on the GPU: config.txt, init textures,
the frame loop with partial steps and frame-start time (bubble.cu:2654-2696)
and the EXR frames, against the same frame loop run on the oracle."""
import os
import subprocess

import numpy as np
import pytest

import soapbubble_amd as sb
from exr_reader import read_exr
from exr_writer import write_exr
from test_host_surface import CLI, _amg, _config

pytestmark = pytest.mark.gpu


def _smooth(n, seed):
    rng = np.random.default_rng(seed)
    j, i = np.meshgrid(np.arange(n), np.arange(2 * n), indexing="ij")
    th, ph = (j + 0.5) * np.pi / n, i * np.pi / n
    f = np.zeros_like(th)
    for _ in range(4):
        a, l, m = rng.uniform(-1, 1), rng.integers(1, 4), rng.integers(0, 3)
        f += a * np.cos(l * th) * np.cos(m * ph + rng.uniform(0, 6.28))
    return f / np.abs(f).max()


@pytest.mark.parametrize("advection,viscous,precision", [("plain", False, "fp64"),
                                                         ("bimocq", False, "fp64"),
                                                         ("plain", True, "fp64"),
                                                         ("plain", False, "mixed"),
                                                         ("bimocq", False, "mixed")])
def test_cli_frames_match_oracle(tmp_path, oracle, advection, viscous, precision):
    """Mixed precision replays each frame's whole steps as captured graphs
    (the partial last step of a frame runs host-driven)."""
    print(f"soapBubble 32x64 {advection} viscous={viscous} {precision}, 3 frames of 0.012 s at "
          "dt 0.005 s", flush=True)
    n, radius, H, dt, DT, frames = 32, 0.05, 1e-6, 0.005, 0.012, 3
    thick = str(tmp_path / "thickness.exr")
    air = str(tmp_path / "air.exr")
    write_exr(thick, {c: 0.5 + 0.3 * _smooth(n, 1) for c in "BGR"})
    write_exr(air, {"B": np.zeros((n, 2 * n)), "G": _smooth(n, 2), "R": _smooth(n, 3)})
    _amg(tmp_path, tol="1e-12" if precision == "fp64" else "1e-10", conv="RELATIVE")
    cfg = _config(tmp_path, n_theta=n, radius=radius, H=H, timePerStep=dt, timePerFrame=DT,
                  frameNumbers=frames, thicknessInit=thick)
    r = subprocess.run([CLI, cfg, "--precision", precision, "--advection", advection, "--air", air,
                        "--viscous", "on" if viscous else "off"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = tmp_path / "out"

    # the same frame loop on the oracle
    f = sb.texture_fields(n, thick, air, radius)
    st = {"u_theta": np.zeros((n - 1) * 2 * n), "u_phi": np.zeros(2 * n * n),
          "eta": f["eta"].copy(), "gamma": np.ones(2 * n * n)}
    kw = dict(radius=radius, H=H, gravity=True, evaporation=True, tol=1e-12, viscous=viscous)
    bm = None
    if advection == "bimocq":
        # blendCoeff 0.5 of the config, five samples, threshold 4 cells
        bm = oracle.BimocqState(oracle.params(n, dt, **kw), st["eta"], st["gamma"], blend=0.5)

    def one(p):
        if bm is None:
            oracle.step(p, st, f["uair0"], f["vair0"], 1, advance_time=False)
        else:
            oracle.bimocq_step(p, st, f["uair0"], f["vair0"], bm, 1, advance_time=False)

    frames_ref = [st["eta"].copy()]
    T = 0.0
    for i in range(1, frames):
        t0 = T
        p = oracle.params(n, dt, time_s=t0, **kw)
        while T + dt <= i * DT:
            one(p)
            T += dt
        if T + 1e-7 < i * DT:
            one(oracle.params(n, i * DT - T, time_s=t0, **kw))
        T = i * DT
        frames_ref.append(st["eta"].copy())

    for k, eta in enumerate(frames_ref):
        _, chans, img = read_exr(str(out / f"frame{k:04d}.exr"))
        want = (eta * H * 2e5).astype(np.float32).astype(np.float16).astype(np.float32)
        got = img["B"].ravel()
        diff = got != want
        # a half rounding boundary may fall between the GPU's and the
        # oracle's values (they agree to ~1e-12); then they differ by one ulp
        ulp = np.abs(np.spacing(want.astype(np.float16)).astype(np.float32))
        assert np.all(np.abs(got - want)[diff] <= ulp[diff] * 1.01), k
        assert diff.mean() < 0.01, (k, diff.mean())


def test_cli_broken_film_stops(tmp_path):
    """A film with no cell of positive thickness is broken at the first
    frame written (write_thickness_img, solver.cu:568-575): the loop stops
    before any step, as the reference's (bubble.cu:2680-2683)."""
    n = 16
    thick = str(tmp_path / "thickness.exr")
    write_exr(thick, {c: np.full((n, 2 * n), 0.5 - 0.5 / 0.3) for c in "BGR"})   # eta = 0
    _amg(tmp_path, tol="1e-10", conv="RELATIVE")
    cfg = _config(tmp_path, n_theta=n, radius=0.05, H=1e-6, timePerStep=0.005,
                  timePerFrame=0.01, frameNumbers=4, thicknessInit=thick)
    r = subprocess.run([CLI, cfg, "--precision", "mixed"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Film is broken." in r.stderr
    assert "Frames completed: 1" in r.stdout
    assert not (tmp_path / "out" / "frame0001.exr").exists()


def test_cli_steps_replay_graphs(tmp_path):
    """With mixed precision the driver's frames replay captured graphs
    whatever the AMG config's max_iters (the cap of AmgX's PCG, 100 here):
    the inner schedule stays automatic. The solve trace shows the captures
    and a schedule within what a captured step holds, and the driver
    reports its frame-loop time."""
    n, dt = 64, 0.001 * 0.05
    print(f"soapBubble {n}x{2 * n} mixed, 4 frames of 10 steps, solve trace", flush=True)
    _amg(tmp_path, tol="1e-07", conv="ABSOLUTE", iters=100)
    cfg = _config(tmp_path, n_theta=n, radius=0.05, H=1e-6, timePerStep=dt,
                  timePerFrame=10 * dt, frameNumbers=4)
    env = dict(os.environ, SF_SOLVE_TRACE="1")
    r = subprocess.run([CLI, cfg, "--precision", "mixed", "--advection", "plain"],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Frames completed: 4" in r.stdout
    assert "max_iters 100" in r.stdout
    assert "captured 4 steps" in r.stderr, r.stderr[-2000:]
    scheds = [int(l.split(":")[1].split()[0]) for l in r.stderr.splitlines()
              if l.startswith("[sf solve] graph schedule:")]
    assert scheds and max(scheds) <= 48, scheds
    assert float(r.stdout.split("Frame loop seconds:")[1].split()[0]) > 0
