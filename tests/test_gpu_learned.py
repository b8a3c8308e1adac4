"""GPU tests of the learned warm-start.

* the BF16 matrix-core network against the FP64 oracle (oracle/mlp_oracle.py): forward
  values and the gradients of the training loss;
* the teacher features and accumulated-lambda targets against the C oracle (exact);
* the learned substep with its per-particle fallback against the oracle's composition of
  the reference kernels (given the network's lambda), bit for bit, at a small size and at
  the 1,048,576-particle dam break;
* what the trained network gains over the same iteration budget without it.
"""
import dataclasses

import numpy as np
import pytest

from helpers import _eq, compare_pbf, log, oracle_teacher
from oracle import mlp_oracle as MO
from oracle import oracle as O
from realtimeparticles_amd import capi, learned, scenes

pytestmark = pytest.mark.gpu


def rand_net(seed, n):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, MO.N_PARAMS).astype(np.float32)
    feat = rng.normal(0, 1.5, (n, 8)).astype(np.float32)
    mu = rng.normal(0, 0.2, 8).astype(np.float32)
    sigma = rng.uniform(0.5, 2.0, 8).astype(np.float32)
    return p, feat, mu, sigma


@pytest.mark.parametrize("n", [1, 64, 1000, 4099])
def test_mlp_forward_matches_fp64_oracle(n):
    p, feat, mu, sigma = rand_net(n, n)
    m = capi.Mlp()
    m.set_params(p)
    m.set_normalization(mu, sigma, 1.3)
    got = m.forward(feat).astype(np.float64)
    ref, _ = MO.forward(p, feat, mu.astype(np.float64), sigma.astype(np.float64), 1.3, bf16_ops=True)
    err = np.abs(got - ref)
    log(f"forward n={n}: max abs err {err.max():.3e}, max |y| {np.abs(ref).max():.3f}")
    assert np.all(err <= 1e-2 * (1.0 + np.abs(ref)))
    # and the bf16 network stays close to the full-precision one
    ref64, _ = MO.forward(p, feat, mu, sigma, 1.3, bf16_ops=False)
    assert np.abs(got - ref64).max() <= 0.1 * (1.0 + np.abs(ref64).max())
    m.close()


def test_mlp_gradients_match_fp64_oracle():
    p, feat, mu, sigma = rand_net(7, 300)
    target = np.random.default_rng(8).normal(0, 2.0, 300).astype(np.float32)
    m = capi.Mlp()
    m.set_params(p)
    m.set_normalization(mu, sigma, 2.0)
    g, loss = m.gradients(feat, target)
    gr, lr = MO.gradients(p, feat, target, mu, sigma, 2.0, bf16_ops=True)
    assert loss == pytest.approx(lr, rel=1e-2)
    for name, (a, b) in {"W1": (0, 512), "b1": (512, 576), "W2": (576, 4672), "b2": (4672, 4736),
                         "w3": (4736, 4800), "b3": (4800, 4801)}.items():
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / max(np.linalg.norm(gr[a:b]), 1e-12)
        log(f"grad {name}: rel err {rel:.2e}")
        # measured on the MI355X: at most 7e-4 (W1); bf16 operands are rounded the same way
        # on both sides, so what remains is FP32 against FP64 accumulation
        assert rel < 1e-3, name
    # the parameters are unchanged by a gradient evaluation
    assert np.array_equal(m.params(), p)
    m.close()


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(3)
    feat = rng.normal(0, 1, (65536, 8)).astype(np.float32)
    target = (np.sin(feat[:, 0]) + 0.5 * feat[:, 1] * feat[:, 2] - 0.3 * feat[:, 3] ** 2
              + 0.1 * feat[:, 7]).astype(np.float32)
    mlp, hist, _ = learned.train(feat, target, epochs=40, batch=4096, lr=3e-3, seed=1)
    log(f"training loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.1 * hist[0]
    pred = mlp.forward(feat[:2000])
    resid = np.mean((pred - target[:2000]) ** 2) / np.var(target)
    assert resid < 0.1
    mlp.close()


def small_scene(seed):
    return scenes.dam_break((24, 16, 16), (4, 2, 2), (40, 20, 20), seed=seed, name="pbf-small")


@pytest.mark.parametrize("mapping", ["literal", "stable"])
def test_teacher_features_and_targets_match_oracle(mapping):
    sc = small_scene(21)
    prm = sc.params(mapping)
    gpu = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    gpu.set_params(**prm)
    gpu.load_state(sc.pos, sc.vel)
    feat, target = gpu.record_teacher(5)
    rf, rt = oracle_teacher(sc, prm, 5)
    _eq(feat, rf, f"features ({mapping})")
    _eq(target, rt, f"accumulated lambda ({mapping})")
    gpu.close()


def _small_net(seed):
    """A random network with a small output, so the learned lambda stays near the scale of
    the plain one."""
    p, _, _, _ = rand_net(seed, 1)
    mlp = capi.Mlp()
    mlp.set_params(p * 0.1)
    mlp.set_normalization(np.zeros(8, np.float32), np.ones(8, np.float32), 1e-4)
    return mlp


def _learned_run(sc, prm, mlp, learned_iters, fb, tau, tuning=None, residual=False):
    """The GPU model after one learned substep, and the oracle's substep from the same
    input with the network's lambda on the oracle's own features."""
    gpu = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    gpu.set_params(**prm)
    gpu.set_nb_jacobi_iters(4)
    if tuning:
        gpu.set_tuning(*tuning)
    gpu.load_state(sc.pos, sc.vel)
    gpu.set_warm_start(mlp, learned_iters, fb, tau)
    gpu.set_residual_pass(residual)
    gpu.update(1)
    orc = O.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**prm), 4,
                         gravity=sc.gravity)
    orc.upload(sc.pos, sc.vel)
    act = orc.learned_step(mlp.forward, learned_iters, fb, tau)
    return gpu, orc, act


def _check_set(gpu, act, n):
    """The GPU's fallback sizes and flags (1 set, 2 halo) against the oracle's (2 seed,
    1 other member, 3 halo)."""
    member = (act[:n] == 1) | (act[:n] == 2)
    halo = act[:n] == 3
    assert gpu.fallback_sizes() == (int((act == 2).sum()), int(member.sum()), int(halo.sum()))
    flags = gpu.read_field(capi.FIELD_FALLBACK_SET, n)
    _eq(flags == 1, member, "fallback set")
    _eq(flags == 2, halo, "fallback halo")
    return flags


def _threshold_between(sc, prm, mlp):
    """A threshold between the mean and the max residual error of learned-1, so that some
    particles but not all seed the fallback set."""
    gpu = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    gpu.set_params(**prm)
    gpu.load_state(sc.pos, sc.vel)
    gpu.set_warm_start(mlp, 1, 0, 0.0)
    gpu.set_residual_pass(True)
    gpu.update(1)
    r = gpu.residual_error()
    gpu.close()
    return float(r[1] + 0.5 * (r[0] - r[1]))


@pytest.mark.parametrize("fb", [3, 4])
@pytest.mark.parametrize("shape", ["default", "tuned"])
@pytest.mark.parametrize("gate", ["none", "some", "all"])
def test_learned_substep_matches_oracle(gate, shape, fb):
    """Learned-1 with 3 or 4 fallback iterations (the iterations alternate between two
    position buffers; an odd count copies the set back) against the oracle's learned
    substep, bit for bit on every field: no particle above tau (the substep equals one
    without fallback), some, and every particle (the fallback is plain iterations)."""
    sc = small_scene(22)
    prm = sc.params("stable")
    mlp = _small_net(5)
    tau = {"none": 1e30, "all": -1.0}.get(gate) or _threshold_between(sc, prm, mlp)
    tuning = (64, capi.LAUNCH_XCD_MAP | capi.LAUNCH_RESORT | capi.LAUNCH_NB_LIST, 0, 16) \
        if shape == "tuned" else None
    gpu, orc, act = _learned_run(sc, prm, mlp, 1, fb, tau, tuning)
    log(f"gate {gate} (tau {tau:.4g}), {shape}: oracle set {int(((act == 1) | (act == 2)).sum())} "
        f"of {sc.n}, {int((act == 2).sum())} seeds, halo {int((act == 3).sum())}; "
        f"GPU {gpu.fallback_sizes()}")
    compare_pbf(gpu, orc, sc.n, 1)
    _check_set(gpu, act, sc.n)
    seeds, members = gpu.fallback_count()
    if gate == "none":
        assert members == 0
        ref, _, _ = _learned_run(sc, prm, mlp, 1, 0, 0.0, tuning)
        _eq(gpu.read_state()[0], ref.read_state()[0], "no particle above tau: no fallback")
        ref.close()
    elif gate == "all":
        assert seeds == members == sc.n
    else:
        assert 0 < seeds < members < sc.n
    gpu.close()
    mlp.close()


def _within_h(q, mask, h):
    out = np.zeros(q.shape[0], bool)
    for i in np.flatnonzero(mask):
        d = q[i] - q
        out |= np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) < h
    return out


@pytest.mark.parametrize("role", ["set", "halo"])
def test_fallback_takes_sorted_element_0(role):
    """Sorted element 0 is in no cell list (Q1), so no walk of another particle meets it.  A
    clump moved near the dam block's corner particle (input 0, which sorts to slot 0) carries
    larger residual errors than that corner.  "set": the clump at 0.7 h, tau between the
    corner's error and the clump's: slot 0 is no seed but lies within h of one.  "halo": the
    clump at 1.3 h, tau such that slot 0 is outside the set but within h of a member (found by
    brute force on the probe's positions).  Either way slot 0 gets its flag by its own walk,
    on the GPU as in the oracle, and the substep stays the oracle's bit for bit."""
    base = small_scene(22)
    pos = base.pos.copy()
    rng = np.random.default_rng(0)
    off, spread, k = ((0.07, 0.02, 0.02), 0.004, 8) if role == "set" else ((0.13, 0.03, 0.0), 0.01, 16)
    pos[-k:, :3] = (pos[0, :3] + np.array(off, np.float32)
                    + rng.uniform(-spread, spread, (k, 3)).astype(np.float32))
    sc = dataclasses.replace(base, pos=pos)
    prm = sc.params("stable")
    mlp = _small_net(5)
    probe = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    probe.set_params(**prm)
    probe.load_state(sc.pos, sc.vel)
    probe.set_warm_start(mlp, 1, 0, 0.0)
    probe.set_residual_pass(True)
    probe.update(1)
    assert probe.read_grid()[1][0] == 0  # the corner particle is sorted element 0
    rho0 = np.float32(prm["restDensity"])
    err = np.abs(probe.read_field(capi.FIELD_RESIDUAL_DENSITY) / rho0 - np.float32(1.0))
    q = probe.read_field(capi.FIELD_PRED_POS)[:, :3]
    probe.close()
    d = q[0] - q
    r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    h = np.float32(sc.box[0]) / np.float32(sc.res[0])  # the cell size is the effect radius
    if role == "set":
        near = (r < np.float32(0.8) * h) & (np.arange(sc.n) > 0)
        assert near.any() and float(err[near].max()) > float(err[0])
        tau = float((float(err[near].max()) + float(err[0])) / 2)
    else:
        lo = max(float(err[0]), float(err[r < h].max()))
        cands = np.unique(err[err > lo])
        tau = None
        for i in range(cands.size - 1, 0, -1):
            t = float((cands[i] + cands[i - 1]) / 2)
            member = _within_h(q, err > t, h) | (err > t)
            if not member[0] and (member & (r < np.float32(0.9) * h)).any():
                tau = t
                break
        assert tau is not None, "no tau leaves element 0 in the halo"
    gpu, orc, act = _learned_run(sc, prm, mlp, 1, 3, tau)
    log(f"element 0 ({role}): error {err[0]:.4g}, tau {tau:.4g}; sizes {gpu.fallback_sizes()}")
    compare_pbf(gpu, orc, sc.n, 1)
    flags = _check_set(gpu, act, sc.n)
    if role == "set":
        assert act[0] == 1 and flags[0] == 1
    else:
        assert act[0] == 3 and flags[0] == 2
    gpu.close()
    mlp.close()


def test_fallback_residual_pass_changes_nothing():
    """The diagnostic post-correction pass after the fallback reads the state and leaves it
    alone, and the fallback lowers the error it measures."""
    sc = small_scene(22)
    prm = sc.params("stable")
    mlp = _small_net(5)
    tau = _threshold_between(sc, prm, mlp)
    a, _, _ = _learned_run(sc, prm, mlp, 1, 3, tau, residual=False)
    b, _, _ = _learned_run(sc, prm, mlp, 1, 3, tau, residual=True)
    c, _, _ = _learned_run(sc, prm, mlp, 1, 0, tau, residual=True)
    _eq(b.read_state()[0], a.read_state()[0], "diagnostic pass changes nothing")
    rb, rc = b.residual_error(), c.residual_error()
    log(f"residual without fallback {rc}, with {rb}")
    assert rb[1] < rc[1]
    for m in (a, b, c):
        m.close()
    mlp.close()


def test_learned_substep_1m_matches_oracle():
    """Config E at its size: the 1,048,576-particle dam break (128x128x64 lattice), learned-1
    with 3 fallback iterations at the benchmark's launch shape.  The network's lambda on the
    oracle's features against the FP64 network oracle (within the bf16 network's tolerance),
    and everything downstream of it -- positions, velocities, densities, lambdas, predicted
    positions, the grid and the fallback set -- bit for bit against the oracle."""
    sc = scenes.pbf_scene("pbf-1m")
    prm = sc.params("stable")
    mlp = _small_net(9)
    tau = _threshold_between(sc, prm, mlp)
    captured = {}

    def network(feat):
        captured["feat"] = feat
        out = mlp.forward(feat)
        captured["out"] = out
        return out

    tuning = (64, capi.LAUNCH_XCD_MAP | capi.LAUNCH_RESORT | capi.LAUNCH_NB_LIST, 0, 16)
    gpu = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
    gpu.set_params(**prm)
    gpu.set_nb_jacobi_iters(4)
    gpu.set_tuning(*tuning)
    gpu.load_state(sc.pos, sc.vel)
    gpu.set_warm_start(mlp, 1, 3, tau)
    gpu.update(1)
    orc = O.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**prm), 4,
                         gravity=sc.gravity)
    orc.upload(sc.pos, sc.vel)
    log("pbf-1m learned substep: oracle")
    act = orc.learned_step(network, 1, 3, tau)
    log(f"pbf-1m learned substep: tau {tau:.4g}, fallback set "
        f"{int(((act == 1) | (act == 2)).sum())} particles, {int((act == 2).sum())} seeds, halo "
        f"{int((act == 3).sum())}; GPU {gpu.fallback_sizes()}")
    compare_pbf(gpu, orc, sc.n, 1)
    _check_set(gpu, act, sc.n)
    assert 0 < int((act == 2).sum()) < sc.n and (act == 3).any()
    p = mlp.params()
    ref, _ = MO.forward(p.astype(np.float64), captured["feat"].astype(np.float64),
                        np.zeros(8), np.ones(8), 1e-4, bf16_ops=True)
    err = np.abs(captured["out"].astype(np.float64) - ref)
    log(f"pbf-1m network lambda: max abs err {err.max():.3e} over {err.shape[0]} rows")
    assert np.all(err <= 1e-2 * (1e-4 + np.abs(ref)))
    gpu.close()
    mlp.close()


def test_trained_network_beats_same_budget_without_it():
    # teacher data from seeds disjoint from the evaluation seed
    feat, target = learned.teacher_data(small_scene, seeds=[31, 32, 33, 34, 35, 36],
                                        mapping="stable", teacher_iters=4, substeps=3)
    mlp, hist, _ = learned.train(feat, target, epochs=80, batch=4096, lr=2e-3, seed=2)
    log(f"teacher rows {feat.shape[0]}, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    sc = small_scene(1)
    res = learned.evaluate(sc, "stable", [
        ("plain-1", None, 1, 0, 0.0),
        ("plain-2", None, 2, 0, 0.0),
        ("plain-4", None, 4, 0, 0.0),
        ("learned-1", mlp, 1, 0, 0.0),
        ("learned-2", mlp, 2, 0, 0.0),
    ], steps=3, warmup=2)
    r = {v.name: v for v in res}
    for v in res:
        log(f"{v.name}: residual max {v.residual_max:.4f} mean {v.residual_mean:.5f} "
            f"max compression {v.residual_compression:.4f}")
    # one learned iteration does better than one and than two plain ones
    assert r["learned-1"].residual_mean < r["plain-2"].residual_mean < r["plain-1"].residual_mean
    # repeated fixed-input substeps give the same residual: the features carry no state
    again = learned.evaluate(sc, "stable", [("learned-1", mlp, 1, 0, 0.0)], steps=2, warmup=0)
    assert again[0].residual_mean == r["learned-1"].residual_mean
    mlp.close()


def test_thread_order_keeps_learned_bits():
    """The learned substep (features, network, fallback, residual pass) gives the same bits
    with the neighbour passes' threads re-ordered by current cell and with survivor lists
    (the second learned iteration records and replays them)."""
    sc = small_scene(23)
    p, _, _, _ = rand_net(6, 1)
    mlp = capi.Mlp()
    mlp.set_params(p * 0.1)
    mlp.set_normalization(np.zeros(8, np.float32), np.ones(8, np.float32), 1e-4)
    out = []
    for flags in (0, capi.LAUNCH_RESORT, capi.LAUNCH_RESORT | capi.LAUNCH_NB_LIST):
        m = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
        m.set_params(**sc.params("stable"))
        m.load_state(sc.pos, sc.vel)
        m.set_tuning(128, flags, 0, 32)
        m.set_warm_start(mlp, 2, 2, 0.05)
        m.set_residual_pass(True)
        m.update(3)
        out.append((m.read_state(), m.read_field(capi.FIELD_LAMBDA), m.residual_error(),
                    m.fallback_count()))
        m.close()
    log(f"fallback set of the last substep: {out[0][3]}")
    assert 0 < out[0][3][0]
    for o in out[1:]:
        _eq(o[0][0], out[0][0][0], "positions")
        _eq(o[0][1], out[0][0][1], "velocities")
        _eq(o[1], out[0][1], "lambda")
        _eq(o[2], out[0][2], "residual")
        assert o[3] == out[0][3]
    mlp.close()

