"""The fixed-input PBF substep of bench.py, for rocprofv3: the same scene, coefficients and
launch shape, timed the same way, without the bench's other sections.

One untimed substep at the default launch shape comes first; then the given shape runs
`--warmup` untimed and `--steps` timed fixed-input substeps (restore the input, one
substep), and a third window of the same substeps with per-kernel HIP events on the model's
stream.  Prints one JSON line: ms per substep and each profiler entry's launches and mean
launch time, to set beside the rocprofv3 trace of the same command.

A Clouds workload (clouds-65k: the cumulus case) runs the Clouds model's steps from the
input instead, as bench.py's Clouds section does: `--warmup` untimed steps, then `--steps`
timed steps from the input, then the same with per-kernel HIP events; `--flags` (if given)
picks its walk: LAUNCH_QUAD (64) a quad of lanes per particle, 0 one lane.  The boids workloads
(boids-8k: the 8,192-boid flock; boids-131k: the reference app's 131,072 boids with its
default rules and dt 0.1) do the same for bench.py's boids runs, at `--block` and `--flags`.
The learned workload (learned-1m) trains bench.py's network (2M teacher samples from seeds 101
and 102 of the 1M dam break, stable coefficients, 200 epochs), sets the fallback threshold to
`--tau-fraction` of the plain 4-iteration substep's post-correction max |rho/rho0 - 1| on the
1M input, and times the learned substep (one learned iteration, `--fallback-iters` fallback
iterations) the same fixed-input way.

    rocprofv3 --kernel-trace --stats -d OUT -o run -- \\
        python3 -m realtimeparticles_amd.profile_step --workload pbf-1m --block 64 --flags 13 \\
        --queue 16 --sort-bits 10
"""
from __future__ import annotations

import argparse
import json
import time

from . import capi, scenes


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="pbf-1m")
    ap.add_argument("--mapping", default="literal", choices=["literal", "stable"])
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--flags", type=int, default=None,
                    help="RTP_LAUNCH_* bits (default 13 for the fluids, the model's own for "
                         "the Clouds model)")
    ap.add_argument("--queue", type=int, default=16)
    ap.add_argument("--sort-bits", type=int, default=10)
    ap.add_argument("--tau-fraction", type=float, default=0.8)
    ap.add_argument("--fallback-iters", type=int, default=4)
    a = ap.parse_args(argv)
    if a.workload.startswith("clouds"):
        return clouds_main(a)
    if a.workload.startswith("boids"):
        return boids_main(a)
    if a.flags is None:
        a.flags = 13
    if a.workload == "learned-1m":
        return learned_main(a)

    sc = scenes.pbf_scene(a.workload)
    m = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    m.set_params(**sc.params(a.mapping))
    m.set_nb_jacobi_iters(sc.iters)
    m.load_state(sc.pos, sc.vel)
    m.save_checkpoint()
    m.update(1)  # the default launch shape
    m.set_tuning(a.block, a.flags, a.sort_bits, a.queue)

    def fixed_steps(k):
        for _ in range(k):
            m.restore_checkpoint()
            m.update(1)

    fixed_steps(a.warmup)
    m.sync()
    t0 = time.perf_counter()
    fixed_steps(a.steps)
    m.sync()
    elapsed = time.perf_counter() - t0
    m.enable_profiling(True)
    m.reset_profile()
    fixed_steps(a.steps)
    m.sync()
    prof = m.profile()
    m.close()
    print(json.dumps({
        "workload": a.workload, "mapping": a.mapping, "particles": sc.n,
        "tuning": {"block": a.block, "flags": a.flags, "queue": a.queue, "sort_bits": a.sort_bits},
        "ms_per_step": round(elapsed * 1e3 / a.steps, 5),
        "kernels": {k: {"launches": v[0], "mean_launch_ms": round(v[1] / max(v[0], 1), 5)}
                    for k, v in sorted(prof.items())}}), flush=True)


def learned_main(a):
    from . import learned
    feat, target = learned.teacher_data(lambda s: scenes.pbf_scene("pbf-1m", seed=s), [101, 102],
                                        mapping="stable", teacher_iters=4, substeps=1)
    mlp, _, _ = learned.train(feat, target, epochs=200, batch=65536, lr=2e-3, seed=1)
    sc = scenes.pbf_scene("pbf-1m")
    m = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    m.set_params(**sc.params("stable"))
    m.set_nb_jacobi_iters(sc.iters)
    m.load_state(sc.pos, sc.vel)
    m.save_checkpoint()
    m.set_tuning(a.block, a.flags, a.sort_bits, a.queue)
    # the threshold: a fraction of the plain substep's post-correction max error here
    m.set_residual_pass(True)
    m.restore_checkpoint()
    m.update(1)
    plain_max = float(m.residual_error()[0])
    m.set_residual_pass(False)
    tau = a.tau_fraction * plain_max
    m.set_warm_start(mlp, 1, a.fallback_iters, tau)

    def fixed_steps(k):
        for _ in range(k):
            m.restore_checkpoint()
            m.update(1)

    fixed_steps(a.warmup)
    m.sync()
    t0 = time.perf_counter()
    fixed_steps(a.steps)
    m.sync()
    elapsed = time.perf_counter() - t0
    sizes = m.fallback_sizes()
    m.enable_profiling(True)
    m.reset_profile()
    fixed_steps(a.steps)
    m.sync()
    prof = m.profile()
    m.set_warm_start(None)
    m.close()
    mlp.close()
    print(json.dumps({
        "workload": a.workload, "particles": sc.n, "tau": tau, "plain4_max_error": plain_max,
        "fallback_iters": a.fallback_iters, "fallback_sizes": list(sizes),
        "tuning": {"block": a.block, "flags": a.flags, "queue": a.queue, "sort_bits": a.sort_bits},
        "ms_per_step": round(elapsed * 1e3 / a.steps, 5),
        "kernels": {k: {"launches": v[0], "mean_launch_ms": round(v[1] / max(v[0], 1), 5)}
                    for k, v in sorted(prof.items())}}), flush=True)


def clouds_main(a):
    sc = scenes.cloud_scene(a.workload)
    m = capi.Clouds(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    if a.flags is not None:
        m.set_quad(bool(a.flags & capi.LAUNCH_QUAD))
    m.set_params(sc.params(), sc.cloud_params())
    m.set_nb_jacobi_iters(sc.iters)

    def from_input(k):
        m.load_state(sc.pos, sc.vel)
        m.update(k)
        m.sync()

    from_input(a.warmup)
    m.load_state(sc.pos, sc.vel)
    m.sync()
    t0 = time.perf_counter()
    m.update(a.steps)
    m.sync()
    elapsed = time.perf_counter() - t0
    m.enable_profiling(True)
    from_input(a.steps)
    prof = m.profile()
    m.close()
    print(json.dumps({
        "workload": a.workload, "particles": sc.n, "steps": a.steps,
        "ms_per_step": round(elapsed * 1e3 / a.steps, 5),
        "kernels": {k: {"launches": v[0], "mean_launch_ms": round(v[1] / max(v[0], 1), 5)}
                    for k, v in sorted(prof.items())}}), flush=True)


def boids_main(a):
    if a.workload == "boids-8k":
        sc, prm = scenes.boids_flock(8192), None
    elif a.workload == "boids-131k":
        sc, prm = scenes.boids_flock(131072, seed=3, box=10.0, res=30), (0.5, 1.6, 1.6, 1.45, 0.1)
    else:
        raise SystemExit(f"unknown workload {a.workload}")
    m = capi.Boids(sc.max_n, sc.n, sc.box, sc.res)
    if prm is None:
        m.set_params(sc.velocity_scale, sc.alignment, sc.separation, sc.cohesion)
        m.set_time_step(sc.time_step)
    else:
        m.set_params(*prm[:4])
        m.set_time_step(prm[4])
    m.set_tuning(a.block, a.flags if a.flags is not None else 0)

    def from_input(k):
        m.load_state(sc.pos, sc.vel)
        m.update(k)
        m.sync()

    from_input(a.warmup)
    m.load_state(sc.pos, sc.vel)
    m.sync()
    t0 = time.perf_counter()
    m.update(a.steps)
    m.sync()
    elapsed = time.perf_counter() - t0
    m.enable_profiling(True)
    from_input(a.steps)
    prof = m.profile()
    m.close()
    print(json.dumps({
        "workload": a.workload, "boids": sc.n, "steps": a.steps,
        "tuning": {"block": a.block, "flags": a.flags},
        "ms_per_step": round(elapsed * 1e3 / a.steps, 5),
        "kernels": {k: {"launches": v[0], "mean_launch_ms": round(v[1] / max(v[0], 1), 5)}
                    for k, v in sorted(prof.items())}}), flush=True)


if __name__ == "__main__":
    main()
