"""Benchmark of the soap-film step on one MI355X.

Workload (default): one bubble at nTheta x nPhi = 1024 x 2048 (2,097,152
cells), dt = 0.001, seeded init with spherical-harmonic degree <= 64 plus 5%
uniform noise on eta, air stream function degree <= 12; mixed precision
(FP32 fields and advection, an FP32 Chebyshev-accelerated phi-line Jacobi
inner solve inside FP64 iterative refinement to a relative residual of
1e-10), the steps replayed as captured hipGraphs (four steps a launch).

One "step" = advection of u_theta, u_phi, eta, Gamma + assembly + implicit
solve + updates. The metric is cell-steps/s over the timed steps.

Also reported on the same line: the plain FP64 GPU path on the same input
(timed after the same warm-up), the error of the mixed path against the
FP64 path and against the oracle after 100 steps, the fraction of HBM peak
of the inner-iteration kernel, the largest share of the step (roofline),
and of the advection kernel (roofline_advect), the FP64 CPU oracle's
throughput over those 100 steps (cpu_baseline), and the packed ensembles of
96 and 68 bubbles (the second with its error against the oracle per
bubble).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress of the benchmark's legs on stderr (the JSON line is stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def make_inputs(n_theta, seed, l_max, l_air, noise):
    import soapbubble_amd as sb

    f = sb.synth_fields(n_theta, seed, l_max, l_air, noise, 1.0)
    nph = 2 * n_theta
    st = {"u_theta": np.zeros((n_theta - 1) * nph), "u_phi": np.zeros(n_theta * nph),
          "eta": f["eta"], "gamma": f["gamma"]}
    return st, f["uair0"], f["vair0"]


def new_ctx(prec, n_theta, dt, st, ua, va, use_graph):
    import soapbubble_amd as sb

    ctx = sb.SoapFilm(sb.bubble(n_theta, dt), prec, use_graph=use_graph)
    ctx.load_state(0, st)
    ctx.upload(0, "uair0", ua)
    ctx.upload(0, "vair0", va)
    return ctx


def timed_run(ctx, warmup, steps):
    ctx.step(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    return time.perf_counter() - t0


def advect_bytes(n_theta, esz):
    """Algorithmic advection bytes per bubble-step: read u_theta, u_phi, eta,
    Gamma once and write them once."""
    nph = 2 * n_theta
    return esz * (2 * (n_theta - 1) * nph + 2 * n_theta * nph + 4 * n_theta * nph)


def inner_bytes(n_theta, esz):
    """Algorithmic bytes of one inner (Chebyshev) launch past the first:
    read e_k, e_(k-1), the FP32 residual and the operator (diagonal, west
    and north face weights) once, write e_(k+1) once. The north/south rows
    of e_k and of the north weights are re-reads of rows other blocks read."""
    return esz * 7 * n_theta * 2 * n_theta


def frame_loop(n_theta, advection, frames=(3, 13), steps_per_frame=10, tol="1e-07",
               convergence="ABSOLUTE", dt=0.001):
    """The soapBubble driver end to end (config.txt, init textures, the frame
    loop of Bubble::run with its per-step broken-film check, one EXR frame a
    frame), mixed precision, the reference's AMG settings (PCGF_F.json:
    ABSOLUTE 1e-7, or the given tolerance), radius 0.05 m and the non-dimensional
    step dt (timePerStep = dt x radius / U). Two runs of different frame counts; the difference of
    the frame-loop times the driver reports, per frame, leaves out the
    process start (HIP initialisation, textures, context set-up) and the
    loop's first frames (calibration, graph instantiation). Textures: smooth
    synthetic images written as the reference's HALF B, G, R EXR."""
    import shutil
    import subprocess
    import tempfile

    import soapbubble_amd as sb

    exe = os.path.join(ROOT, "soapbubble_amd", "soapBubble")
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="sf_frames_")
    try:
        h, w = 512, 1024
        y, x = np.meshgrid(np.linspace(0, 1, h), np.linspace(0, 1, w), indexing="ij")
        th = 0.5 + 0.2 * np.sin(2 * np.pi * x) * np.sin(np.pi * y) + 0.05 * np.cos(6 * np.pi * x)
        sb.write_image_exr(os.path.join(d, "thick.exr"), th, th, th)
        sb.write_image_exr(os.path.join(d, "air.exr"), 0.5 * np.cos(2 * np.pi * y),
                           0.5 * np.sin(2 * np.pi * x) * np.sin(np.pi * y), 0 * x)
        with open(os.path.join(d, "amg.json"), "w") as f:
            f.write('{"config_version": 2, "solver": {"preconditioner": {"solver": "AMG", '
                    '"max_iters": 1}, "solver": "PCGF", "max_iters": 100, "convergence": '
                    f'"{convergence}", "tolerance": {tol}, "norm": "L2"}}}}')
        radius = 0.05
        dt_s = dt * radius               # U = 1
        el, loop = {}, {}
        for fr in frames:
            vals = dict(radius=radius, H=1e-6, U=1, cm=0.5, Gammam=1e-6, T=298.15, Ds=0.0,
                        rm=5e-9, numberOfGridsInTheta=n_theta, timePerStep=dt_s,
                        timePerFrame=steps_per_frame * dt_s, frameNumbers=fr,
                        outputDir=os.path.join(d, f"out{fr}"),
                        thicknessInit=os.path.join(d, "thick.exr"),
                        AMGconfig=os.path.join(d, "amg.json"), blendCoeff=0.5, gpu=0)
            cfg = os.path.join(d, f"config{fr}.txt")
            with open(cfg, "w") as f:
                f.write("".join(f"{k} {v}\n" for k, v in vals.items()))
            t0 = time.perf_counter()
            r = subprocess.run([exe, cfg, "--precision", "mixed", "--advection", advection,
                                "--air", os.path.join(d, "air.exr")],
                               capture_output=True, text=True, timeout=300)
            el[fr] = time.perf_counter() - t0
            if r.returncode != 0 or f"Frames completed: {fr}" not in r.stdout:
                raise RuntimeError(f"soapBubble failed ({r.returncode}): {r.stdout} {r.stderr}")
            loop[fr] = float(r.stdout.split("Frame loop seconds:")[1].split()[0])
        a, b = frames
        per = (loop[b] - loop[a]) / (b - a)
        return {"workload": f"soapBubble {n_theta}x{2 * n_theta}, {advection} advection, "
                            f"{steps_per_frame} steps of dt {dt:g} a frame, EXR frames, "
                            f"{convergence} {tol}",
                "frames": list(frames), "elapsed_s": [el[a], el[b]],
                "loop_s": [loop[a], loop[b]],
                "ms_per_frame": per * 1e3, "ms_per_step": per * 1e3 / steps_per_frame,
                "cell_steps_per_s": n_theta * 2 * n_theta * steps_per_frame / per}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def ensemble_specs(groups, l_max, seed0):
    """Per-bubble descriptors and fields of a packed ensemble: groups of
    (count, n_theta, dt); bubble b has seed seed0 + 16 b, air amplitude
    U[0.2, 1.0) and Marangoni number M times U[0.5, 2.0)."""
    import soapbubble_amd as sb

    rng = np.random.default_rng(seed0)
    specs = []
    b = 0
    for count, n, dt in groups:
        for _ in range(count):
            amp = rng.uniform(0.2, 1.0)
            ms = rng.uniform(0.5, 2.0)
            f = sb.synth_fields(n, seed0 + 16 * b, l_max, 12, 0.0, amp)
            specs.append({"n": n, "dt": dt, "M_scale": ms, "fields": f})
            b += 1
    return specs


# the packed ensembles of uneven resolution: (count, n_theta, dt) groups
C3_GROUPS = [(32, 64, 0.005), (32, 128, 0.004), (32, 256, 0.002)]
C4_GROUPS = [(4, 2048, 0.0005), (64, 128, 0.004)]


def ensemble_ctx(specs, use_graph=True):
    import soapbubble_amd as sb

    ctx = sb.SoapFilm([sb.bubble(q["n"], q["dt"], M_scale=q["M_scale"]) for q in specs], "mixed",
                      use_graph=use_graph)
    for k, q in enumerate(specs):
        n, f = q["n"], q["fields"]
        nph = 2 * n
        ctx.load_state(k, {"u_theta": np.zeros((n - 1) * nph), "u_phi": np.zeros(n * nph),
                           "eta": f["eta"], "gamma": f["gamma"]})
        ctx.upload(k, "uair0", f["uair0"])
        ctx.upload(k, "vair0", f["vair0"])
    return ctx


def ensemble(name, groups, steps, warmup, l_max, seed0, err_steps=0):
    """A packed ensemble (one context, one launch per stage for all bubbles,
    graph replay), timed over `steps` steps after `warmup`. With err_steps,
    a second run of err_steps steps from the same state is compared bubble
    by bubble with the FP64 oracle (relative L2 of every field)."""
    specs = ensemble_specs(groups, l_max, seed0)
    cells = sum(q["n"] * 2 * q["n"] for q in specs)
    ctx = ensemble_ctx(specs)
    el = timed_run(ctx, warmup, steps)
    fails = sum(ctx.stats(k)["failed_solves"] for k in range(len(specs)))
    ctx.close()
    out = {"workload": name, "bubbles": len(specs), "cells": cells, "steps": steps,
           "warmup": warmup, "value": cells * steps / el, "unit": "cell-steps/s",
           "ms_per_step": el * 1e3 / steps, "failed_solves": fails}
    if err_steps:
        from oracle import oracle as O

        ctx = ensemble_ctx(specs)
        ctx.step(err_steps)
        per = []
        for k, q in enumerate(specs):
            g = ctx.state(k)
            f, n = q["fields"], q["n"]
            nph = 2 * n
            st = {"u_theta": np.zeros((n - 1) * nph), "u_phi": np.zeros(n * nph),
                  "eta": f["eta"].copy(), "gamma": f["gamma"].copy()}
            O.step(O.params(n, q["dt"], M_scale=q["M_scale"]), st, f["uair0"], f["vair0"],
                   err_steps)
            per.append({kk: rel_l2(g[kk], st[kk]) for kk in ("eta", "gamma", "u_theta", "u_phi")})
        ctx.close()
        worst = {kk: max(e[kk] for e in per) for kk in per[0]}
        out["err_vs_oracle"] = {
            "steps": err_steps, "worst": worst,
            "per_bubble": [{"n_theta": q["n"], "eta": float(f"{e['eta']:.3g}"),
                            "gamma": float(f"{e['gamma']:.3g}"),
                            "u": float(f"{max(e['u_theta'], e['u_phi']):.3g}")}
                           for q, e in zip(specs, per)]}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def stiff_leg(warmup, steps, err_steps, with_oracle):
    """The reference's own config.txt regime: radius 0.05 m, 256 x 512,
    timePerStep 0.005 s (a non-dimensional step of 0.1), where the
    line-Jacobi preconditioned operator has a spectral bound of 0.991; mixed
    precision with the geometric-multigrid preconditioned inner PCG (graph
    replay), with the Chebyshev inner solve (whose schedule is too long to
    capture: host-driven steps) and with the automatic choice (Chebyshev,
    switching to multigrid after the first solve shows the schedule too
    long), ms per step of each, and the multigrid path's error against the
    oracle after err_steps steps."""
    import soapbubble_amd as sb

    n, dt, radius = 256, 0.005, 0.05
    f = sb.synth_fields(n, 2100, 16, 12, 0.0, 1.0)
    nph = 2 * n
    st = {"u_theta": np.zeros((n - 1) * nph), "u_phi": np.zeros(n * nph),
          "eta": f["eta"], "gamma": f["gamma"]}
    out = {"workload": f"one bubble {n}x{nph}, radius {radius} m, dt {dt} s (config.txt), mixed",
           "steps": steps, "warmup": warmup}
    for name, method in (("multigrid", "mg"), ("chebyshev", "chebyshev"), ("automatic", "auto")):
        ctx = sb.SoapFilm(sb.bubble(n, dt, radius=radius), "mixed", use_graph=True,
                          inner_method=method)
        ctx.load_state(0, st)
        ctx.upload(0, "uair0", f["uair0"])
        ctx.upload(0, "vair0", f["vair0"])
        el = timed_run(ctx, warmup, steps)
        s = ctx.stats(0)
        ctx.close()
        out[name] = {"ms_per_step": el * 1e3 / steps, "cell_steps_per_s": n * nph * steps / el,
                     "last_inner_iterations": s["iterations"], "last_refinements": s["outer"],
                     "failed_solves": s["failed_solves"], "max_residual": s["max_residual"]}
    if with_oracle and err_steps:
        from oracle import oracle as O

        ctx = sb.SoapFilm(sb.bubble(n, dt, radius=radius), "mixed", use_graph=True,
                          inner_method="mg")
        ctx.load_state(0, st)
        ctx.upload(0, "uair0", f["uair0"])
        ctx.upload(0, "vair0", f["vair0"])
        ctx.step(err_steps)
        g = ctx.state(0)
        ctx.close()
        ref = {k: v.copy() for k, v in st.items()}
        O.step(O.params(n, dt, radius=radius), ref, f["uair0"], f["vair0"], err_steps)
        out["multigrid"]["err_vs_oracle"] = {"steps": err_steps}
        out["multigrid"]["err_vs_oracle"].update(
            {k: rel_l2(g[k], ref[k]) for k in ("eta", "gamma", "u_theta", "u_phi")})
    return out


def copy_calibration(mib):
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counter
    readings against the bytes it must move give the counters' scale."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 4 * n, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # the GPU's clocks take about 20 ms of sustained load to ramp after an
    # idle period; 50 steps (the first three host-driven: guess warm-up and
    # schedule calibration) cover it
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n-theta", type=int, default=1024)
    ap.add_argument("--dt", type=float, default=0.001)
    ap.add_argument("--fp64-steps", type=int, default=20)
    ap.add_argument("--oracle-steps", type=int, default=100,
                    help="steps of the FP64 oracle (CPU baseline and the headline's error)")
    ap.add_argument("--c4-err-steps", type=int, default=5,
                    help="steps of the 68-bubble ensemble compared with the oracle per bubble")
    ap.add_argument("--c3-err-steps", type=int, default=10,
                    help="steps of the 96-bubble ensemble compared with the oracle per bubble")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-bimocq", action="store_true")
    ap.add_argument("--no-viscous", action="store_true")
    ap.add_argument("--viscous-steps", type=int, default=20)
    ap.add_argument("--no-ensembles", action="store_true")
    ap.add_argument("--no-frames", action="store_true")
    ap.add_argument("--no-stiff", action="store_true")
    ap.add_argument("--bimocq-steps", type=int, default=10)
    ap.add_argument("--profile", choices=("graph", "eager", "inner", "c4", "c4-eager"),
                    default=None,
                    help="run only the headline steps (for rocprofv3), graph or eager launches; "
                         "inner: after `warmup` eager steps, one assembly and `steps` + 1 "
                         "back-to-back inner (Chebyshev) launches; c4, c4-eager: the steps of "
                         "the 4 x 2048x4096 + 64 x 128x256 ensemble (35.7M cells, past the "
                         "256 MB last-level cache), graph or eager launches")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run only a device-to-device copy of MIB MiB three times (torch), the "
                         "reference kernel that scales the profiler's HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(bubbles would shard across GPUs without communication).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import soapbubble_amd as sb

    n, dt = args.n_theta, args.dt
    nph = 2 * n
    cells = n * nph
    seed = 2000
    st0, ua, va = make_inputs(n, seed, 64, 12, 0.05)

    if args.profile == "inner":
        ctx = new_ctx("mixed", n, dt, st0, ua, va, False)
        ctx.step(args.warmup)
        ctx.advect()
        ctx.assemble()
        ms = ctx.time_inner(args.steps)
        ctx.close()
        print(json.dumps({"profile": "inner", "launches": args.steps + 1, "warmup": args.warmup,
                          "inner_launch_ms": ms}))
        return
    if args.profile in ("c4", "c4-eager"):
        specs = ensemble_specs(C4_GROUPS, 128, 4000)
        ctx = ensemble_ctx(specs, args.profile == "c4")
        el = timed_run(ctx, args.warmup, args.steps)
        fails = sum(ctx.stats(k)["failed_solves"] for k in range(len(specs)))
        ctx.close()
        print(json.dumps({"profile": args.profile, "steps": args.steps, "warmup": args.warmup,
                          "cells": sum(2 * q["n"] * q["n"] for q in specs),
                          "ms_per_step": el * 1e3 / args.steps, "failed_solves": fails}))
        return
    if args.profile:
        ctx = new_ctx("mixed", n, dt, st0, ua, va, args.profile == "graph")
        el = timed_run(ctx, args.warmup, args.steps)
        s = ctx.stats(0)
        ctx.close()
        print(json.dumps({"profile": args.profile, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": el * 1e3 / args.steps, "failed_solves": s["failed_solves"]}))
        return

    log("headline: mixed, graph")
    # headline: mixed precision, hipGraph per step
    ctx = new_ctx("mixed", n, dt, st0, ua, va, True)
    elapsed = timed_run(ctx, args.warmup, args.steps)
    mixed_state = ctx.state(0)
    mstats = ctx.stats(0)
    value = cells * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    steps_total = args.warmup + args.steps
    ctx.close()

    log("stage times, advection and inner launch timings")
    # per-stage device times of the same kernels, eager launches, from HIP
    # events on the library's stream
    ctx = new_ctx("mixed", n, dt, st0, ua, va, False)
    ctx.step(2)
    # 20 ms of advection launches first: the context set-up leaves the GPU
    # idle long enough for its clocks to drop
    ctx.time_advection(450)
    ctx.reset_timing()
    nst = 5
    ctx.step(nst)
    tm = ctx.timing()
    # the advection kernel alone: back-to-back launches between two events
    adv_ms = ctx.time_advection(100)
    # the inner iteration launch alone (the kernel with the largest share
    # of the step: several launches per step), the same way
    ctx.advect()
    ctx.assemble()
    ctx.time_inner(1400)
    inner_ms = ctx.time_inner(200)
    ctx.solve()
    ctx.update()
    ctx.close()
    adv_bytes = advect_bytes(n, 4)
    adv_gbs = adv_bytes / (adv_ms * 1e-3) / 1e9
    in_bytes = inner_bytes(n, 4)
    in_gbs = in_bytes / (inner_ms * 1e-3) / 1e9
    # the kernels' names as the profiler prints them (template arguments: the
    # cells a thread, packed rows; power-of-two rows)
    in_kernel = f"k_cheb<float, {8 if nph >= 2048 else 4}, false>"
    adv_kernel = f"k_advect<float, {'true' if nph & (nph - 1) == 0 else 'false'}>"
    stage_ms = {k: round(tm[k] / nst, 4) for k in ("advect_ms", "assemble_ms", "solve_ms",
                                                     "update_ms", "step_ms")}

    log("plain FP64 path")
    # plain FP64 path on the same input
    fp64 = None
    err_fp64 = None
    if not args.no_fp64:
        # the same warm-up as the mixed leg, so both are timed with the
        # GPU's clocks ramped
        ctx = new_ctx("fp64", n, dt, st0, ua, va, False)
        w64 = min(args.warmup, steps_total)
        k64 = min(args.fp64_steps, steps_total - w64)
        el64 = timed_run(ctx, w64, k64)
        fp64 = cells * k64 / el64
        if w64 + k64 < steps_total:
            ctx.step(steps_total - w64 - k64)
        ref_state = ctx.state(0)
        ctx.close()
        err_fp64 = {k: rel_l2(mixed_state[k], ref_state[k]) for k in ("eta", "gamma",
                                                                     "u_theta", "u_phi")}

    log("oracle (CPU baseline) and the headline path after the same steps")
    # FP64 CPU oracle over the workload's first `oracle_steps` steps (100:
    # the acceptance point of the mixed path's error) from the same state:
    # its time is the CPU baseline, its state the reference of the headline
    # path (mixed, graph replay) after the same steps
    cpu = None
    err_oracle = None
    if not args.no_cpu:
        from oracle import oracle as O

        p = O.params(n, dt)
        st = {k: v.copy() for k, v in st0.items()}
        k = args.oracle_steps
        t0 = time.perf_counter()
        O.step(p, st, ua, va, k)
        el = time.perf_counter() - t0
        cpu = {"value": cells * k / el, "unit": "cell-steps/s", "cores": O.num_threads(),
               "cpu_model": cpu_model(), "kind": "port",
               "sample": f"the first {k} steps of the same {n}x{nph} workload, FP64 OpenMP oracle"}
        ctx = new_ctx("mixed", n, dt, st0, ua, va, True)
        ctx.step(k)
        g = ctx.state(0)
        gs = ctx.stats(0)
        ctx.close()
        v0 = O.volume(p, st0["eta"])
        err_oracle = {"steps": k, "path": "mixed, hipGraph replay"}
        err_oracle.update({kk: rel_l2(g[kk], st[kk]) for kk in ("eta", "gamma", "u_theta",
                                                                 "u_phi")})
        err_oracle["volume_drift"] = {"gpu": gs["volume"] / v0 - 1.0,
                                      "oracle": O.volume(p, st["eta"]) / v0 - 1.0}

    log("BIMOCQ")
    # the reference's default scalar advection (BIMOCQ, five samples per
    # point, blendCoeff 0.5) on the same input: mixed with a graph, and the
    # FP64 path for the error
    bimocq = None
    if not args.no_bimocq:
        nb_steps = args.bimocq_steps
        ctx = sb.SoapFilm(sb.bubble(n, dt), "mixed", use_graph=True, scalar_advection="bimocq")
        ctx.load_state(0, st0)
        ctx.upload(0, "uair0", ua)
        ctx.upload(0, "vair0", va)
        elb = timed_run(ctx, args.warmup, nb_steps)
        gb = ctx.state(0)
        sbm = ctx.stats(0)
        ctx.close()
        ctx = sb.SoapFilm(sb.bubble(n, dt), "fp64", scalar_advection="bimocq")
        ctx.load_state(0, st0)
        ctx.upload(0, "uair0", ua)
        ctx.upload(0, "vair0", va)
        ctx.step(args.warmup + nb_steps)
        rb = ctx.state(0)
        ctx.close()
        bimocq = {"value": cells * nb_steps / elb, "unit": "cell-steps/s",
                  "ms_per_step": elb * 1e3 / nb_steps, "steps": nb_steps, "warmup": args.warmup,
                  "distortion": sbm["distortion"], "reinits": sbm["reinits"],
                  "failed_solves": sbm["failed_solves"],
                  "err_vs_fp64_gpu": {k: rel_l2(gb[k], rb[k]) for k in ("eta", "gamma",
                                                                        "u_theta", "u_phi")}}

    log("VISCOUS")
    # the VISCOUS option (gravity on, so the viscous forces enter the
    # assembly): mixed with a graph, and the FP64 path for the error
    viscous = None
    if not args.no_viscous:
        nv_steps = args.viscous_steps
        out = {}
        for prec, graph in (("mixed", True), ("fp64", False)):
            ctx = sb.SoapFilm(sb.bubble(n, dt, gravity=True), prec, use_graph=graph,
                              viscous=True)
            ctx.load_state(0, st0)
            ctx.upload(0, "uair0", ua)
            ctx.upload(0, "vair0", va)
            el = timed_run(ctx, args.warmup, nv_steps)
            out[prec] = (el, ctx.state(0), ctx.stats(0))
            ctx.close()
        el, gv, sv = out["mixed"]
        rv = out["fp64"][1]
        viscous = {"value": cells * nv_steps / el, "unit": "cell-steps/s",
                   "ms_per_step": el * 1e3 / nv_steps, "steps": nv_steps,
                   "warmup": args.warmup, "gravity": True,
                   "fp64_ms_per_step": out["fp64"][0] * 1e3 / nv_steps,
                   "failed_solves": sv["failed_solves"], "max_residual": sv["max_residual"],
                   "err_vs_fp64_gpu": {k: rel_l2(gv[k], rv[k]) for k in ("eta", "gamma",
                                                                         "u_theta", "u_phi")}}

    log("ensembles")
    # packed ensembles of uneven resolution (batched form), mixed, graph
    ensembles = None
    if not args.no_ensembles:
        ensembles = [
            ensemble("32 x 64x128 (dt 0.005) + 32 x 128x256 (0.004) + 32 x 256x512 (0.002)",
                     C3_GROUPS, 200, args.warmup, 16, 3000,
                     err_steps=0 if args.no_cpu else args.c3_err_steps),
            ensemble("4 x 2048x4096 (dt 0.0005) + 64 x 128x256 (0.004)",
                     C4_GROUPS, 50, args.warmup, 128, 4000,
                     err_steps=0 if args.no_cpu else args.c4_err_steps),
        ]

    # the stiff regime of the reference's config.txt: multigrid and Chebyshev
    stiff = None
    if not args.no_stiff:
        log("stiff config.txt regime: multigrid and Chebyshev")
        stiff = stiff_leg(10, 20, 5, not args.no_cpu)

    # the soapBubble driver's frame loop, plain and BIMOCQ (the reference's
    # default scalar advection)
    frames = None
    if not args.no_frames:
        log("frame loop")
        # the reference's AMG settings, and the headline's tolerance (its
        # steps against the headline step: the frame loop's own cost)
        frames = [frame_loop(n, "plain"), frame_loop(n, "bimocq"),
                  frame_loop(n, "plain", tol="1e-10", convergence="RELATIVE"),
                  # the reference's config.txt itself: 256x512, timePerStep
                  # 0.005 s at radius 0.05 (dt 0.1, the stiff regime: the
                  # automatic inner method switches to multigrid), 4 steps a
                  # frame, BIMOCQ (its default advection)
                  frame_loop(256, "bimocq", steps_per_frame=4, dt=0.1)]

    # HBM bytes per launch measured by rocprofv3 for this workload
    traffic = adv_traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_1024x2048_mixed.json")
    if n == 1024 and os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        adv_traffic = tj["kernels"].get(adv_kernel, {}).get("traffic_bytes")
        il = tj.get("inner_launch", {})
        traffic = il.get("traffic_bytes") if il.get("kernel") == in_kernel else None

    out = {
        "metric": "cell-steps/s (advect+implicit solve) at stated err vs FP64 oracle; % HBM peak",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spherical harmonics + noise)",
        "config": {"workload": f"one bubble {n}x{nph}, dt={dt}, mixed precision, hipGraph step",
                   "n_theta": n, "n_phi": nph, "cells": cells, "dt": dt, "seed": seed,
                   "precision": "fp32 fields and advection; system assembled in fp64, "
                                "fp32 Chebyshev-accelerated phi-line Jacobi inside fp64 "
                                "iterative refinement to rel. residual 1e-10"},
        "roofline": {"bound": "hbm", "kernel": in_kernel, "achieved": in_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": in_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": in_bytes, "avg_ms": inner_ms},
        "roofline_advect": {"bound": "hbm", "kernel": adv_kernel, "achieved": adv_gbs,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": adv_gbs / HBM_PEAK_GBS, "traffic": adv_traffic,
                            "bytes_per_launch": adv_bytes, "avg_ms": adv_ms},
        "cpu_baseline": cpu,
        "stage_ms": stage_ms,
        "solver": {"max_residual": mstats["max_residual"], "failed_solves": mstats["failed_solves"],
                   "last_iterations": mstats["iterations"], "last_outer": mstats["outer"]},
        "fp64_plain_value": fp64,
        "speedup_vs_fp64": (value / fp64) if fp64 else None,
        "err_vs_fp64_gpu": err_fp64,
        "err_vs_oracle": err_oracle,
        "bimocq": bimocq,
        "viscous": viscous,
        "ensembles": ensembles,
        "stiff": stiff,
        "frame_loop": frames,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
