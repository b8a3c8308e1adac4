"""Pins of the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests or golden vectors, so every oracle function is pinned here by
something independent of its own formula: the normalisation integrals of the SPH kernels,
finite differences, libm, numpy's stable sort, brute-force run scans and conservation laws.
"""
import ctypes
import math

import numpy as np
import pytest

from realtimeparticles_amd import scenes


def f4(x, y, z):
    return np.array([x, y, z, 0.0], np.float32)


def consts(O, box=(4, 2, 2), res=(40, 20, 20), cap=100, g=9.8):
    return O.make_consts(box, res, cap, g)


# ------------------------------------------------------------------ constants (Q11)

def test_program_constants(oracle):
    c = consts(oracle)
    h = np.float32(0.1)
    assert np.float32(c.effectRadius) == h
    assert np.float32(c.cellSize) == h
    assert list(c.absWall) == [2.0, 1.0, 1.0]
    assert c.numCells == 16000
    # Each constant goes float -> "%.10f" -> float: the 10-decimal text must parse back.
    assert oracle.lib().orc_float_roundtrip(np.float32(0.1)) == np.float32(0.1)
    # poly6 and spiky normalisations to float32 precision (~1e-6 relative).
    poly6 = 315.0 / (64.0 * math.pi * float(h) ** 9)
    spiky = 15.0 / (math.pi * float(h) ** 6)
    assert abs(c.poly6Coeff / poly6 - 1) < 2e-6
    assert abs(c.spikyCoeff / spiky - 1) < 2e-6
    assert c.maxVel == 30.0
    assert np.float32(c.effectRadiusSq) == np.float32(0.01)


def test_pown_matches_libm(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-1, 1, 2000), rng.uniform(-0.01, 0.01, 500)]).astype(np.float32)
    for n, tol in ((2, 1), (3, 2), (4, 3)):
        for x in xs:
            got = np.float32(L.orc_pown(float(x), n))
            ref = np.float32(np.float64(x) ** n)
            ulp = np.spacing(np.abs(ref)) if ref != 0 else np.float32(1e-45)
            assert abs(float(got) - float(ref)) <= tol * float(ulp)
    assert L.orc_pown(3.0, 0) == 1.0


# ------------------------------------------------------------------ SPH kernels (a7, a8)

def test_poly6_normalisation_and_value_at_zero(oracle):
    c = consts(oracle)
    L = oracle.lib()
    h = float(np.float32(c.effectRadius))
    w0 = L.orc_poly6(f4(0, 0, 0).ctypes.data, ctypes.byref(c))
    assert w0 == pytest.approx(c.poly6Coeff * np.float32(h * h) ** 3, rel=1e-6)
    # integral of W over the ball of radius h is 1 (midpoint rule on a 64^3 lattice)
    m = 64
    s = 2 * h / m
    ax = -h + (np.arange(m) + 0.5) * s
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.zeros((m ** 3, 4), np.float32)
    pts[:, 0], pts[:, 1], pts[:, 2] = X.ravel(), Y.ravel(), Z.ravel()
    vals = np.array([L.orc_poly6(pts[i].ctypes.data, ctypes.byref(c)) for i in range(0, m ** 3)],
                    np.float64)
    assert vals.sum() * s ** 3 == pytest.approx(1.0, abs=2e-3)
    # W vanishes at and beyond h
    assert L.orc_poly6(f4(h, 0, 0).ctypes.data, ctypes.byref(c)) == 0.0
    assert L.orc_poly6(f4(0, 0, 1.5 * h).ctypes.data, ctypes.byref(c)) == 0.0


def test_grad_spiky_is_gradient_of_spiky(oracle):
    c = consts(oracle)
    L = oracle.lib()
    h = float(np.float32(c.effectRadius))
    spiky = float(c.spikyCoeff)
    rng = np.random.default_rng(5)
    out = np.zeros(4, np.float32)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = rng.uniform(0.05, 0.95) * h
        v = (d * r).astype(np.float32)
        vec = np.array([v[0], v[1], v[2], 0], np.float32)
        L.orc_grad_spiky(vec.ctypes.data, ctypes.byref(c), out.ctypes.data)
        # central differences of W(v) = spiky * (h - |v|)^3 in float64
        fd = np.zeros(3)
        eps = 1e-7
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            wp = spiky * (h - np.linalg.norm(vec[:3].astype(np.float64) + e)) ** 3
            wm = spiky * (h - np.linalg.norm(vec[:3].astype(np.float64) - e)) ** 3
            fd[a] = (wp - wm) / (2 * eps)
        assert np.allclose(out[:3], fd, rtol=1e-3, atol=1e-3 * np.abs(fd).max())
        assert out[3] == 0.0
    # zero vector and beyond the support give exact zeros
    L.orc_grad_spiky(f4(0, 0, 0).ctypes.data, ctypes.byref(c), out.ctypes.data)
    assert not out.any()
    L.orc_grad_spiky(f4(2 * h, 0, 0).ctypes.data, ctypes.byref(c), out.ctypes.data)
    assert not out.any()


def test_spiky_normalisation(oracle):
    # integral of spiky*(h-r)^3 over the ball is 1: 4*pi*spiky*h^6/60
    c = consts(oracle)
    h = float(np.float32(c.effectRadius))
    assert 4 * math.pi * c.spikyCoeff * h ** 6 / 60.0 == pytest.approx(1.0, rel=2e-6)


def test_art_pressure_at_dq_is_minus_k(oracle):
    c = consts(oracle)
    L = oracle.lib()
    p = oracle.fluid_params(artPressureRadius=0.2, artPressureCoeff=0.1, artPressureExp=4)
    h = np.float32(c.effectRadius)
    dq = np.float32(np.float32(0.2) * h)
    v = f4(dq, 0, 0)
    assert L.orc_art_pressure(v.ctypes.data, ctypes.byref(p), ctypes.byref(c)) == pytest.approx(-0.1, rel=1e-6)
    v2 = f4(0, 0, 0)
    ratio = 1.0 / (1 - 0.04) ** 3
    assert L.orc_art_pressure(v2.ctypes.data, ctypes.byref(p), ctypes.byref(c)) == pytest.approx(-0.1 * ratio ** 4, rel=1e-5)
    p.isArtPressureEnabled = 0
    assert L.orc_art_pressure(v.ctypes.data, ctypes.byref(p), ctypes.byref(c)) == 0.0


# ------------------------------------------------------------------ grid (a1-a6)

def test_cell_hash_special_cases(oracle):
    c = consts(oracle)
    L = oracle.lib()
    out = np.zeros(3, np.uint32)

    def cell(x, y, z):
        L.orc_cell3d(f4(x, y, z).ctypes.data, ctypes.byref(c), out.ctypes.data)
        return tuple(int(v) for v in out)

    assert cell(-2.0, -1.0, -1.0) == (0, 0, 0)
    assert cell(2.0, 1.0, 1.0) == (40, 20, 20)          # exactly on +W: index == res
    assert cell(5.0, -7.0, 3.0) == (40, 0, 20)          # beyond the walls: clamped
    assert cell(-1.95, -0.95, -0.95) == (0, 0, 0)
    assert cell(np.nan, 0.0, 0.0)[0] == 0               # clamp(NaN) -> lower wall
    # a point inside cell k along x lands in k: (k+0.5)*h - W
    for k in range(40):
        assert cell(np.float32((k + 0.5) * 0.1 - 2.0), 0, 0)[0] == k


def test_cell_hash_brackets_point(oracle):
    c = consts(oracle, box=(8, 8, 4), res=(80, 80, 40))
    rng = np.random.default_rng(11)
    pos = np.zeros((5000, 4), np.float32)
    pos[:, :3] = rng.uniform(-5, 5, (5000, 3)).astype(np.float32)
    key = np.zeros(5000, np.uint32)
    oracle.lib().orc_fill_cell_ids(pos.ctypes.data, key.ctypes.data, 5000, ctypes.byref(c))
    wall = np.array([4.0, 4.0, 2.0])
    q = np.clip(pos[:, :3].astype(np.float64), -wall, wall) + wall
    cx = key // (80 * 40)
    cy = (key // 40) % 80
    cz = key % 40
    cells = np.stack([cx, cy, cz], 1).astype(np.float64)
    inside = (q < np.array([8.0, 8.0, 4.0]) - 1e-5).all(1)
    # exact real interval with a one-ulp allowance for the float division
    assert (cells[inside] * 0.1 <= q[inside] + 1e-6).all()
    assert (q[inside] < (cells[inside] + 1) * 0.1 + 1e-6).all()


def test_radix_sort_is_stable_sort(oracle):
    rng = np.random.default_rng(2)
    for n, hi in ((512, 10), (4096, 1 << 20), (66048, 1 << 32)):
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        k = keys.copy()
        idx = np.zeros(n, np.uint32)
        assert oracle.lib().orc_radix_sort(k.ctypes.data, idx.ctypes.data, n) == 0
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(idx, ref.astype(np.uint32))
        assert np.array_equal(k, keys[ref])
    # RadixSort.cpp:34-35 precondition
    k = np.zeros(100, np.uint32)
    assert oracle.lib().orc_radix_sort(k.ctypes.data, k.ctypes.data, 100) == -1


def brute_start_end(keys, n, nc, cap, key_len):
    se = np.tile(np.array([1, 0], np.int64), (nc, 1))
    i = 0
    while i < n:
        j = i
        while j + 1 < n and keys[j + 1] == keys[i]:
            j += 1
        k = int(keys[i])
        # the run may continue past n into the tail (only if tail keys equal; they do not)
        if k < nc:
            start = i if i > 0 else 1
            end = j
            if j == n - 1 and n < key_len and keys[n] == k:
                end = None  # fillEndCell writes no end for this run
            se[k, 0] = start if i > 0 else se[k, 0]
            if end is not None:
                se[k, 1] = end
            if se[k, 1] > se[k, 0]:
                se[k, 1] = se[k, 0] + min(se[k, 1] - se[k, 0], cap)
        i = j + 1
    return se


def test_start_end_against_brute_force(oracle):
    c = consts(oracle, box=(2, 2, 2), res=(4, 4, 4), cap=5)
    L = oracle.lib()
    nc = 64
    rng = np.random.default_rng(9)
    for trial in range(20):
        n = int(rng.integers(1, 200))
        max_n = n + 8
        keys = np.sort(rng.integers(0, nc + 3, n)).astype(np.uint32)  # some keys >= NC
        if trial % 3 == 0:
            keys[: min(12, n)] = keys[0]  # long first run: Q1 + cap
        full = np.concatenate([keys, (2 * nc + np.arange(n, max_n)).astype(np.uint32)])
        se = np.zeros((nc, 2), np.uint32)
        L.orc_reset_start_end(se.ctypes.data, ctypes.byref(c))
        L.orc_fill_start(full.ctypes.data, se.ctypes.data, n, ctypes.byref(c))
        L.orc_fill_end(full.ctypes.data, se.ctypes.data, n, max_n, ctypes.byref(c))
        L.orc_adjust_end(se.ctypes.data, ctypes.byref(c))
        ref = brute_start_end(full, n, nc, 5, max_n)
        assert np.array_equal(se.astype(np.int64), ref), trial


# ------------------------------------------------------------------ PBF invariants

def lattice_state(O, jitter=0.0, iters=1, **over):
    sc = scenes.dam_break((16, 16, 16), (4, 4, 4), (40, 40, 40), jitter=jitter)
    # centre the block so that it is far from every wall
    sc.pos[:, :3] += np.float32(1.2)
    prm = sc.params()
    prm.update(over)
    f = O.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**prm), iters, gravity=0.0)
    f.upload(sc.pos, sc.vel)
    return sc, f


def interior_mask(sc, pos, margin=2.5):
    lo = pos[:, :3].min(0) + margin * 0.05
    hi = pos[:, :3].max(0) - margin * 0.05
    return ((pos[:, :3] > lo) & (pos[:, :3] < hi)).all(1)


def test_rest_density_of_lattice_interior(oracle):
    sc, f = lattice_state(oracle, iters=1)
    f.step(1)
    m = interior_mask(sc, f.pred[: sc.n])
    assert m.sum() > 500
    rho = f.density[: sc.n][m].astype(np.float64)
    assert np.allclose(rho, sc.rest_density, rtol=2e-5)
    # hydrostatic interior: C ~ 0 so lambda ~ 0
    lam = f.lam[: sc.n][m]
    assert np.abs(lam).max() < 1e-7


def test_correction_sums_to_zero(oracle):
    # Pair terms (lambda_i + lambda_j + s_corr(r_ij)) grad W(p_i - p_j) are antisymmetric.
    # The particle sorted first is left out of its own cell (Q1); an isolated particle
    # placed in cell 0 takes that slot so every real pair is seen from both sides.
    sc, f = lattice_state(oracle, jitter=0.2, iters=1)
    f.pos[0, :3] = (-1.95, -1.95, -1.95)
    oracle.lib().orc_reset_cell_ids(f.key.ctypes.data, f.max_n, ctypes.byref(f.c))
    f.step(1)
    corr = f.corr[: sc.n, :3].astype(np.float64)
    scale = np.abs(corr).max()
    assert scale > 0
    assert np.abs(corr.sum(0)).max() < 1e-4 * scale * np.sqrt(sc.n)


def test_pbf_step_keeps_w_zero_and_finite(oracle):
    sc = scenes.pbf_scene("pbf-65k")
    f = oracle.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, oracle.fluid_params(**sc.params()),
                            4, gravity=9.8)
    f.upload(sc.pos, sc.vel)
    f.step(2)
    assert np.isfinite(f.pos[: sc.n]).all() and np.isfinite(f.vel[: sc.n]).all()
    assert not f.pos[: sc.n, 3].any() and not f.vel[: sc.n, 3].any()
    # grid invariants after a step: keys sorted, idx a permutation, tail untouched
    assert (np.diff(f.key.astype(np.int64)) >= 0).all()
    assert np.array_equal(np.sort(f.idx), np.arange(sc.max_n))
    assert np.array_equal(f.idx[sc.n:], np.arange(sc.n, sc.max_n))


# ------------------------------------------------------------------ boids (a22-a24)

def boids(O, pos, vel, dt=0.01, **prm):
    n = pos.shape[0]
    p = O.BoidsParams(prm.get("vs", 2.0), prm.get("al", 1.0), prm.get("se", 1.5), prm.get("co", 1.0))
    b = O.OracleBoids((2, 2, 2), (20, 20, 20), 512, n, p, time_step=dt)
    b.upload(pos, vel)
    return b


def test_isolated_boid_has_no_acceleration(oracle):
    pos = np.array([[0.05, 0.05, 0.05, 0], [0.7, 0.7, 0.7, 0]], np.float32)
    vel = np.array([[1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    b = boids(oracle, pos, vel)
    b.step(1)
    assert not b.acc[:2].any()
    # speed already in [0.2 vmax, vmax] keeps its value
    assert np.allclose(np.linalg.norm(b.vel[:2, :3], axis=1), 1.0, rtol=1e-6)


def test_two_boid_rules(oracle):
    # boid A at origin-ish, B at distance 0.05 along +x; same velocity direction (+y)
    pos = np.array([[0.01, 0.01, 0.01, 0], [0.06, 0.01, 0.01, 0]], np.float32)
    vel = np.array([[0, 1, 0, 0], [0, 1, 0, 0]], np.float32)
    b = boids(oracle, pos, vel, al=1.0, se=1.5, co=1.0, vs=2.0)
    b.step(1)
    # after the sort A is element 0 -> Q1 removes it from its cell, so only A sees B
    # A's acceleration: alignment +y*2, separation -x*2*1.5, cohesion +x*2
    accs = {tuple(np.round(p[:3], 4)): a for p, a in zip(b.pos[:2], b.acc[:2])}
    a_acc = b.acc[0] if b.pos[0, 0] < b.pos[1, 0] else b.acc[1]
    assert a_acc[0] == pytest.approx(-3.0 + 2.0, abs=1e-6)
    assert a_acc[1] == pytest.approx(2.0, abs=1e-6)
    assert a_acc[2] == pytest.approx(0.0, abs=1e-6)
    assert accs is not None


def test_boid_speed_bounds_and_wall_bounce(oracle):
    rng = np.random.default_rng(1)
    n = 256
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.99, 0.99, (n, 3))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(size=(n, 3)) * 3
    b = boids(oracle, pos, vel, dt=0.01, vs=2.0)
    for _ in range(20):
        b.step(1)
        sp = np.linalg.norm(b.vel[:n, :3].astype(np.float64), axis=1)
        at_wall = (np.abs(b.pos[:n, :3]) == 1.0).any(1)
        free = ~at_wall
        assert (sp[free] >= 0.4 * (1 - 1e-6)).all() and (sp[free] <= 2.0 * (1 + 1e-6)).all()
        assert (np.abs(b.pos[:n, :3]) <= 1.0).all()


def test_wall_bounce_halves_and_reverses(oracle):
    pos = np.array([[0.995, 0.0, 0.0, 0]], np.float32)
    vel = np.array([[1.5, 0.0, 0.0, 0]], np.float32)
    b = boids(oracle, pos, vel, dt=0.01, vs=2.0)
    b.step(1)
    assert b.pos[0, 0] == 1.0
    assert b.vel[0, 0] == pytest.approx(-0.75, rel=1e-6)


def test_cyclic_wall_wraps_without_touching_velocity(oracle):
    pos = np.array([[0.995, 0.0, 0.0, 0], [0.0, -0.998, 0.0, 0], [0.3, 0.2, 0.1, 0]], np.float32)
    vel = np.array([[1.5, 0.0, 0.0, 0], [0.0, -0.5, 0.0, 0], [0.5, 0.0, 0.0, 0]], np.float32)
    b = boids(oracle, pos, vel, dt=0.01, vs=2.0)
    b.set_cyclic(True)
    b.step(1)
    # past +W in x: re-enters at -W; past -W in y: at +W; the velocity is kept
    got = {tuple(p[:3]) for p in b.pos[:3]}
    assert (-1.0, 0.0, 0.0) in got and (0.0, 1.0, 0.0) in got
    assert np.float32(0.305) in {np.float32(p[0]) for p in b.pos[:3]}
    speeds = sorted(np.round(np.linalg.norm(b.vel[:3, :3], axis=1), 6))
    assert speeds == [0.5, 0.5, 1.5]


def test_target_rule_closed_form(oracle):
    L = oracle.lib()
    pos = np.array([f4(0, 0, 0), f4(0, 0, 0), f4(0, 0, 0), f4(3, 0, 0)], np.float32)
    acc = np.zeros((4, 4), np.float32)
    acc[1] = (0.25, 0.5, 0, 0)
    tgt = (ctypes.c_float * 3)(1.0, 0.0, 0.0)
    L.orc_bd_add_target(pos.ctypes.data, tgt, 2.5, 1, acc.ctypes.data, 3)
    # distance 1: min(1.3 / 1, 1.4 * MAX_STEERING = 0.7) = 0.7 towards the target
    assert acc[0].tolist() == pytest.approx([0.7, 0, 0, 0])
    assert acc[1].tolist() == pytest.approx([0.95, 0.5, 0, 0])
    tgt2 = (ctypes.c_float * 3)(0.0, 2.0, 0.0)
    a2 = np.zeros((4, 4), np.float32)
    L.orc_bd_add_target(pos.ctypes.data, tgt2, 2.5, -1, a2.ctypes.data, 4)
    # distance 2: 1.3 / 2 = 0.65 per unit, repelled (sign -1); the boid at (3, 0, 0) is
    # sqrt(13) away, outside the radius
    assert a2[0].tolist() == pytest.approx([0, -1.3, 0, 0])
    assert not a2[3].any()


def test_2d_rules_see_only_the_yz_plane(oracle):
    # a planar flock at x = 0 in a cubic grid: the 2-D rules equal the 3-D ones
    rng = np.random.default_rng(4)
    n = 300
    pos = np.zeros((n, 4), np.float32)
    pos[:, 1:3] = rng.uniform(-0.9, 0.9, (n, 2))
    vel = np.zeros((n, 4), np.float32)
    vel[:, 1:3] = rng.normal(size=(n, 2))
    a, b = boids(oracle, pos, vel), boids(oracle, pos, vel)
    b.set_dimension(2)
    for _ in range(5):
        a.step(1)
        b.step(1)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.acc, b.acc)
    assert a.acc[:n, 1:].any()
    # The 2-D rules visit the x layer RX/2 whatever the boid's own x.  A (layer 10) and B
    # (layer 9) are 0.004 apart: the 3-D rules let each see the other; the 2-D rules let B
    # see A but not A see B.  The third boid takes sorted slot 0 (Q1).
    pos2 = np.array([[0.002, 0.01, 0.01, 0], [-0.002, 0.01, 0.01, 0], [-0.5, 0.5, 0.5, 0]],
                    np.float32)
    vel2 = np.array([[0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 1, 0]], np.float32)
    c3, c2 = boids(oracle, pos2, vel2, dt=1e-6), boids(oracle, pos2, vel2, dt=1e-6)
    c2.set_dimension(2)
    c3.step(1)
    c2.step(1)

    def acc_of(b, x_sign):
        k = [i for i in range(3) if np.sign(b.pos[i, 0]) == x_sign and abs(b.pos[i, 0]) < 0.01]
        return b.acc[k[0]]
    assert acc_of(c3, 1).any() and acc_of(c3, -1).any()
    assert not acc_of(c2, 1).any() and acc_of(c2, -1).any()


# ------------------------------------------------------------------ rendering side of a frame

def test_camera_distance_special_cases(oracle):
    L = oracle.lib()
    pos = np.array([f4(3, 4, 0), f4(0, 0, 12), f4(1, 2, 2), f4(0, 0, 0),
                    f4(2e4, 0, 0), f4(np.inf, np.inf, np.inf), f4(0, 0, 1e-5)], np.float32)
    f32x3 = ctypes.c_float * 3
    cam = f32x3(0.0, 0.0, 0.0)
    out = np.zeros(len(pos), np.uint32)
    L.orc_fill_camera_dist(pos.ctypes.data, cam, out.ctypes.data, len(pos))
    # 1e6 - 100 * |p - cam|, truncated; clamped at 0 far away and for infinite positions;
    # distances under half a float ulp of 1e6 / 100 leave the key at the reset value
    assert list(out) == [999500, 998800, 999700, 1000000, 0, 0, 1000000]
    cam = f32x3(1.0, 2.0, 2.0)
    L.orc_fill_camera_dist(pos.ctypes.data, cam, out.ctypes.data, 4)
    assert out[2] == 1000000 and out[3] == 999700
    reset = np.zeros(1024, np.uint32)
    L.orc_reset_camera_dist(reset.ctypes.data, 1024)
    assert (reset == 1000000).all()


def test_fluid_colour_branches(oracle):
    L = oracle.lib()
    p = oracle.fluid_params(restDensity=400.0)
    dens = np.array([400.0, 0.0, 540.0, 200.0, 600.0], np.float32)
    col = np.zeros((5, 4), np.float32)
    L.orc_fld_fill_color(dens.ctypes.data, ctypes.byref(p), col.ctypes.data, 5)
    blue = np.array([0.0, 0.1, 1.0, 0.5], np.float32)
    light = np.array([0.7, 0.7, 1.0, 0.5], np.float32)
    dark = np.array([0.0, 0.0, 0.8, 0.5], np.float32)
    assert np.array_equal(col[0], blue)                 # at rest density: blue exactly
    # constraint 1 (empty): blue + (light - blue) / 0.35
    assert np.allclose(col[1], blue + (light - blue) / 0.35, rtol=1e-6)
    # constraint -0.35 (compressed by 35 %): dark blue
    assert np.allclose(col[2], dark, atol=1e-6)
    # constraint 0.5 towards light blue, -0.5 past dark blue; alpha stays 0.5
    assert np.allclose(col[3], blue + 0.5 * (light - blue) / 0.35, rtol=1e-6)
    assert np.allclose(col[4], blue - 0.5 * (blue - dark) / 0.35, rtol=1e-6)
    assert (col[:, 3] == 0.5).all()
    bc = np.zeros((3, 4), np.float32)
    L.orc_bd_fill_color(bc.ctypes.data, 3)
    assert (bc == np.array([1.0, 0.02, 0.02, 0.5], np.float32)).all()


def test_detector_against_brute_force(oracle):
    c = consts(oracle)
    rng = np.random.default_rng(5)
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-2.2, 2.2, (n, 3)).astype(np.float32) * np.float32([1, 0.5, 0.5])
    pos[:8, :3] = [[2, 1, 1], [-2, -1, -1], [2, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5],
                   [-5, 0, 0], [0.05, 0.05, 0.05]]
    det = np.full((int(c.numCells), 8), 7.0, np.float32)
    oracle.lib().orc_reset_detector(det.ctypes.data, ctypes.byref(c))
    assert not det.any()
    oracle.lib().orc_fill_detector(pos.ctypes.data, det.ctypes.data, n, ctypes.byref(c))
    w = np.array([2.0, 1.0, 1.0], np.float32)
    cs = np.float32(c.cellSize)
    ijk = np.floor((np.clip(pos[:, :3], -w, w) + w) / cs).astype(np.int64)
    cell = (ijk[:, 0] * 20 + ijk[:, 1]) * 20 + ijk[:, 2]
    want = np.zeros(int(c.numCells), bool)
    want[cell[cell < c.numCells]] = True
    assert np.array_equal(det.any(1), want)
    assert set(np.unique(det)) <= {0.0, 1.0} and (det.min(1) == det.max(1)).all()


def _camera_keys(pos, cam):
    d = pos[:, :3] - np.asarray(cam, np.float32)
    ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    return np.maximum(np.float32(1e6) - ln * np.float32(100.0), np.float32(0)).astype(np.uint32)


def test_render_frame_is_physics_then_stable_camera_sort(oracle):
    """A frame with rendering = the physics-only frame, then every buffer permuted by the
    stable argsort of the camera keys (over maxN, with the reset key past n)."""
    sc, f = lattice_state(oracle, jitter=0.2, iters=2)
    _, g = lattice_state(oracle, jitter=0.2, iters=2)
    cam = (7.0, 3.0, -5.0)
    f.enable_render(cam)
    f.step(1)
    g.step(1)
    keys = np.full(sc.max_n, 1000000, np.uint32)
    keys[: sc.n] = _camera_keys(g.pos[: sc.n], cam)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(f.idx, perm)
    assert np.array_equal(f.render.dist, keys[perm])
    for a, b in ((f.pos, g.pos), (f.vel, g.vel), (f.pred, g.pred)):
        assert np.array_equal(a, b[perm])
    col = np.tile(np.float32([0.0, 0.1, 1.0, 0.0]), (sc.max_n, 1))
    oracle.lib().orc_fld_fill_color(g.density.ctypes.data, ctypes.byref(g.p), col.ctypes.data, sc.n)
    assert np.array_equal(f.render.col, col[perm])
    det = np.zeros_like(f.render.det)
    oracle.lib().orc_fill_detector(g.pos.ctypes.data, det.ctypes.data, sc.n, ctypes.byref(g.c))
    assert np.array_equal(f.render.det, det)
    # the closest particle is drawn last; the inactive tail stays in place
    assert np.array_equal(perm[sc.n:], np.arange(sc.n, sc.max_n))
    # a paused frame sorts again: the same keys, already in order, give the identity
    f.paused = True
    pos1 = f.pos.copy()
    f.step(1)
    assert np.array_equal(f.idx, np.arange(sc.max_n)) and np.array_equal(f.pos, pos1)
