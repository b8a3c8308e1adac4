"""Learned warm-start for the PBF solver: teacher data, training and evaluation.

The network (include/rtp.h, csrc/mlp.hip) maps eight per-particle inputs of the first
density pass to a lambda.  Its training target is the lambda a particle accumulates over
`teacher_iters` plain Jacobi iterations (the sum of every iteration's lambda).  Applied in
one correction step, the sum over the solver's default four iterations leaves a lower mean
density error than the four plain iterations themselves; sums over many more iterations
overshoot (every particle at the free surface keeps a lambda of the same sign while its
displacement saturates), so the target stays at that count.  A learned substep runs
`learned_iters` iterations whose first uses the network's lambda; with fallback, every
particle whose post-correction |rho/rho0 - 1| exceeds tau, and every particle within h of
one, then runs plain iterations.  The particles within h of that set (its halo) recompute
their density and lambda in each of those iterations, so the set yields to a compressed
neighbour, but keep their positions, as every other particle does.

Everything here runs on the GPU through the C ABI: teacher substeps, the training loop
(forward, backward and Adam kernels) and the evaluation.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import capi, scenes


@dataclasses.dataclass
class Normalization:
    mu: np.ndarray
    sigma: np.ndarray
    out_scale: float


def teacher_data(scene_fn, seeds, mapping: str = "literal", teacher_iters: int = 4,
                 substeps: int = 2, keep: float = 1.0, rng_seed: int = 0):
    """(features, targets) from teacher substeps along plain-solver trajectories.

    For every seed the scene starts from its input; at each of `substeps` substeps the
    state is checkpointed, a teacher substep records features and accumulated lambda, the
    checkpoint is restored and the trajectory advances by one plain substep.  `keep`
    subsamples the rows.
    """
    rng = np.random.default_rng(rng_seed)
    feats, targets = [], []
    for seed in seeds:
        sc = scene_fn(seed)
        m = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
        m.set_params(**sc.params(mapping))
        m.set_nb_jacobi_iters(sc.iters)
        m.load_state(sc.pos, sc.vel)
        for _ in range(substeps):
            m.save_checkpoint()
            f, t = m.record_teacher(teacher_iters)
            m.restore_checkpoint()
            m.update(1)
            ok = np.isfinite(f).all(1) & np.isfinite(t)
            sel = ok & (rng.random(t.shape[0]) < keep)
            feats.append(f[sel])
            targets.append(t[sel])
        m.close()
    return np.concatenate(feats), np.concatenate(targets)


def fit_normalization(feat: np.ndarray, target: np.ndarray) -> Normalization:
    mu = feat.mean(0).astype(np.float32)
    sigma = feat.std(0).astype(np.float32)
    sigma[sigma < 1e-12] = 1.0
    scale = float(target.std()) or 1.0
    return Normalization(mu, sigma, scale)


def train(feat: np.ndarray, target: np.ndarray, epochs: int = 20, batch: int = 65536,
          lr: float = 2e-3, seed: int = 0, mlp: capi.Mlp | None = None):
    """Train a network on (feat, target) with the GPU kernels; returns (mlp, loss history)."""
    norm = fit_normalization(feat, target)
    mlp = mlp or capi.Mlp(seed=seed)
    mlp.set_normalization(norm.mu, norm.sigma, norm.out_scale)
    hist = mlp.train(feat, target, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, hist, norm


@dataclasses.dataclass
class Variant:
    name: str
    particle_updates_per_s: float
    ms_per_step: float
    residual_max: float
    residual_mean: float
    residual_compression: float
    fallback_set: tuple = (0, 0, 0)  # (seeds, set, halo) of the fallback, learned runs


def evaluate(sc: scenes.FluidScene, mapping: str, runs, steps: int = 20, warmup: int = 3,
             tune: bool = False):
    """Times fixed-input substeps (restore the input, one substep) for each run spec, then
    runs one more with the post-correction density pass to read its error.  The timed
    substeps carry that pass only where the fallback gate needs it.  With `tune`, the
    launch shape is autotuned once on the plain solver and used for every run.

    runs: list of (name, mlp or None, iterations, fallback_iters, tau).
    """
    out = []
    m = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    m.set_params(**sc.params(mapping))
    m.load_state(sc.pos, sc.vel)
    m.save_checkpoint()
    if tune:
        m.set_nb_jacobi_iters(sc.iters)
        m.autotune(trial_steps=3)
    for name, mlp, iters, fb, tau in runs:
        if mlp is None:
            m.set_warm_start(None)
            m.set_nb_jacobi_iters(iters)
        else:
            m.set_nb_jacobi_iters(sc.iters)
            m.set_warm_start(mlp, iters, fb, tau)
        for _ in range(warmup):
            m.restore_checkpoint()
            m.update(1)
        m.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.restore_checkpoint()
            m.update(1)
        m.sync()
        el = time.perf_counter() - t0
        m.set_residual_pass(True)
        m.restore_checkpoint()
        m.update(1)
        r = m.residual_error()
        m.set_residual_pass(False)
        fset = m.fallback_sizes() if (mlp is not None and fb > 0) else (0, 0, 0)
        out.append(Variant(name, sc.n * steps / el, el * 1e3 / steps, float(r[0]), float(r[1]),
                           float(r[2]), fset))
    m.set_warm_start(None)
    m.close()
    return out
