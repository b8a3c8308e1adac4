"""Pins of the FP64 network oracle: central finite differences of the loss for every
parameter group, and the bfloat16 rounding helper against hand-picked values."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding():
    # bfloat16 keeps 7 fraction bits: neighbours of 1.0 are 1 - 2^-8 and 1 + 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5],
                 np.float32)
    got = M.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                    # exact tie -> even neighbour 1.0
    assert got[2] == 1.0                    # below the tie
    assert got[3] == 1.0 + 2 ** -7          # above the tie
    assert got[4] == 1.0 + 2 ** -6          # tie between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert got[5] == -2.5


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, M.N_PARAMS)
    feat = rng.normal(0, 1, (37, 8))
    target = rng.normal(0, 1, 37)
    mu = rng.normal(0, 0.1, 8)
    sigma = rng.uniform(0.5, 2.0, 8)
    g, _ = M.gradients(p, feat, target, mu, sigma, 1.7)
    eps = 1e-6
    idx = np.concatenate([np.arange(0, 512, 37), np.arange(512, 576, 7), np.arange(576, 4672, 301),
                          np.arange(4672, 4736, 9), np.arange(4736, 4800, 11), [4800]])
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        pm = p.copy()
        pm[i] -= eps
        lp = M.gradients(pp, feat, target, mu, sigma, 1.7)[1]
        lm = M.gradients(pm, feat, target, mu, sigma, 1.7)[1]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gelu_grad_is_derivative():
    z = np.linspace(-6, 6, 241)
    fd = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(z), fd, atol=1e-8)
