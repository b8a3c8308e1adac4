"""GPU tests of the learned warm-start.

* the BF16 matrix-core network against the FP64 oracle (oracle/mlp_oracle.py): forward
  values and the gradients of the training loss;
* the teacher features and accumulated-lambda targets against the C oracle (exact);
* the fallback gate, and what the trained network gains over the same iteration budget
  without it.
"""
import numpy as np
import pytest

from helpers import _eq, log, oracle_teacher
from oracle import mlp_oracle as MO
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
        assert rel < 5e-2, name
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


def test_fallback_gate():
    sc = small_scene(22)
    prm = sc.params("stable")
    p, _, mu, sigma = rand_net(5, 1)
    mlp = capi.Mlp()
    mlp.set_params(p * 0.1)
    mlp.set_normalization(np.zeros(8, np.float32), np.ones(8, np.float32), 1e-4)

    def run(fb, tau, diag=False):
        m = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=sc.gravity)
        m.set_params(**prm)
        m.load_state(sc.pos, sc.vel)
        m.set_warm_start(mlp, 1, fb, tau)
        m.set_residual_pass(diag)
        m.update(1)
        out = m.read_state()[0], m.residual_error()
        m.close()
        return out

    p0, r0 = run(0, 0.0, diag=True)
    p1, r1 = run(3, 1e30)   # gate closed: identical to no fallback
    _eq(p1, p0, "gate closed")
    p2, r2 = run(3, -1.0)   # gate open: three more plain iterations
    assert not np.array_equal(p2, p0)
    # the gate's pass precedes the fallback and matches the diagnostic pass bit for bit
    # (a deterministic reduction)
    assert np.array_equal(r1, r0) and np.array_equal(r2, r0)
    # with the diagnostic pass on, the error left after the fallback iterations is read
    p5, r5 = run(3, -1.0, diag=True)
    _eq(p5, p2, "diagnostic pass changes nothing")
    log(f"residual without fallback {r0}, after three fallback iterations {r5}")
    assert r5[1] < r0[1]
    # a threshold just below / above the measured mean decides the gate
    p3, _ = run(3, float(r0[1]) * 0.999)
    _eq(p3, p2, "gate open at tau < mean")
    p4, _ = run(3, float(r0[1]) * 1.001)
    _eq(p4, p0, "gate closed at tau > mean")
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
        m.set_warm_start(mlp, 2, 2, -1.0)
        m.set_residual_pass(True)
        m.update(3)
        out.append((m.read_state(), m.read_field(capi.FIELD_LAMBDA), m.residual_error()))
        m.close()
    for o in out[1:]:
        _eq(o[0][0], out[0][0][0], "positions")
        _eq(o[0][1], out[0][0][1], "velocities")
        _eq(o[1], out[0][1], "lambda")
        _eq(o[2], out[0][2], "residual")
    mlp.close()

