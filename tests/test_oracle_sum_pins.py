"""Pins of the oracle's neighbour sums and per-particle updates (fluids.cl:62-439).

Every check here is independent of the oracle's own code: closed forms worked out by hand
for two particles, invariants the sums must keep (uniform velocity, momentum, a rigid
rotation), and a grid-free all-pairs evaluation in float64 written from the SPH kernels'
textbook definitions (Mueller et al. 2003; the PBF constraint of Macklin and Mueller 2013):

    W(r)      = 315 / (64 pi h^9) (h^2 - r^2)^3                     r < h
    grad W(v) = -45 / (pi h^6) (h - r)^2 v / r                      0 < r < h
    lambda_i  = -C_i / ((sum_j |grad W_ij|^2 + |sum_j grad W_ij|^2) / rho0^2 + eps)
    dp_i      = sum_j (lambda_i + lambda_j + s_corr(v_ij)) grad W(v_ij) / rho0
    omega_i   = sum_j (v_j - v_i) x grad W(p_i - p_j)
    N_i       = sum_j |omega_j| grad W(p_i - p_j);  v_i += eps_v (N_i/|N_i| x omega_i) dt
    xsph_i    = v_i + c sum_j (v_j - v_i) W(p_i - p_j)

The oracle walks the uniform grid (27 cells, sorted order, float32); the checks here loop
over all pairs in float64.  A dropped term, a wrong sign, a swapped operand of a cross
product or a wrong cell walk changes the result by far more than the float32 rounding the
tolerances allow for.
"""
import ctypes
import math

import numpy as np
import pytest

BOX = (2, 2, 2)
RES = (20, 20, 20)


def f4(x, y, z, w=0.0):
    return np.array([x, y, z, w], np.float32)


class Grid:
    """Particles sorted into the oracle's grid: keys, stable sort over maxN, cell table.

    Particle 0 of the input is a sentinel placed alone in cell 0 so that it takes sorted
    slot 0, the slot fillStartCell leaves out of its cell (grid.cl:95-108); every other
    particle then appears in its cell's list and sees all of its neighbours.
    """

    def __init__(self, O, pos, box=BOX, res=RES, cap=100):
        self.O = O
        L = O.lib()
        self.c = O.make_consts(box, res, cap, 9.8)
        n = pos.shape[0]
        self.n = n
        self.max_n = max(512, -(-n // 512) * 512)
        full = np.zeros((self.max_n, 4), np.float32)
        full[:] = (np.inf, np.inf, np.inf, 0)
        full[:n] = pos
        key = np.zeros(self.max_n, np.uint32)
        L.orc_reset_cell_ids(key.ctypes.data, self.max_n, ctypes.byref(self.c))
        L.orc_fill_cell_ids(full.ctypes.data, key.ctypes.data, n, ctypes.byref(self.c))
        idx = np.zeros(self.max_n, np.uint32)
        assert L.orc_radix_sort(key.ctypes.data, idx.ctypes.data, self.max_n) == 0
        self.idx = idx
        self.pred = np.ascontiguousarray(full[idx])
        self.se = np.zeros((int(self.c.numCells), 2), np.uint32)
        L.orc_reset_start_end(self.se.ctypes.data, ctypes.byref(self.c))
        L.orc_fill_start(key.ctypes.data, self.se.ctypes.data, n, ctypes.byref(self.c))
        L.orc_fill_end(key.ctypes.data, self.se.ctypes.data, n, self.max_n, ctypes.byref(self.c))
        L.orc_adjust_end(self.se.ctypes.data, ctypes.byref(self.c))
        assert idx[0] == 0, "the sentinel must take sorted slot 0"

    def sorted_f4(self, a):
        full = np.zeros((self.max_n, 4), np.float32)
        full[: self.n] = a
        return np.ascontiguousarray(full[self.idx])

    def sorted_f(self, a):
        full = np.zeros(self.max_n, np.float32)
        full[: self.n] = a
        return np.ascontiguousarray(full[self.idx])

    def unsort(self, a):
        out = np.empty_like(a[: self.n])
        out[self.idx[: self.n]] = a[: self.n]
        return out

    # ---- oracle kernels over the sorted arrays, results in input order
    def density(self, p):
        d = np.zeros(self.max_n, np.float32)
        self.O.lib().orc_fld_density(self.pred.ctypes.data, self.se.ctypes.data, ctypes.byref(p),
                                     d.ctypes.data, self.n, ctypes.byref(self.c))
        return d

    def lam(self, p, dens_sorted):
        lam = np.zeros(self.max_n, np.float32)
        self.O.lib().orc_fld_lambda(self.pred.ctypes.data, dens_sorted.ctypes.data,
                                    self.se.ctypes.data, ctypes.byref(p), lam.ctypes.data, self.n,
                                    ctypes.byref(self.c))
        return self.unsort(lam)

    def correction(self, p, lam):
        ls = self.sorted_f(lam)
        corr = np.zeros((self.max_n, 4), np.float32)
        self.O.lib().orc_fld_correction(ls.ctypes.data, self.se.ctypes.data, self.pred.ctypes.data,
                                        ctypes.byref(p), corr.ctypes.data, self.n,
                                        ctypes.byref(self.c))
        return self.unsort(corr)

    def vorticity(self, p, vel):
        vs = self.sorted_f4(vel)
        out = np.zeros((self.max_n, 4), np.float32)
        self.O.lib().orc_fld_vorticity(self.pred.ctypes.data, self.se.ctypes.data, vs.ctypes.data,
                                       ctypes.byref(p), out.ctypes.data, self.n, ctypes.byref(self.c))
        return self.unsort(out)

    def confinement(self, p, vort, vel):
        ws = self.sorted_f4(vort)
        vs = self.sorted_f4(vel)
        self.O.lib().orc_fld_confinement(self.pred.ctypes.data, self.se.ctypes.data, ws.ctypes.data,
                                         ctypes.byref(p), vs.ctypes.data, self.n,
                                         ctypes.byref(self.c))
        return self.unsort(vs)

    def xsph(self, p, vel):
        vs = self.sorted_f4(vel)
        out = np.zeros((self.max_n, 4), np.float32)
        self.O.lib().orc_fld_xsph(self.pred.ctypes.data, self.se.ctypes.data, vs.ctypes.data,
                                  ctypes.byref(p), out.ctypes.data, self.n, ctypes.byref(self.c))
        return self.unsort(out)


# ---------------------------------------------------------------- float64 SPH, all pairs

H = float(np.float32(0.1))


def w64(r):
    r = np.asarray(r, np.float64)
    return np.where(r < H, 315.0 / (64.0 * math.pi * H ** 9) * (H * H - r * r) ** 3, 0.0)


def gradw64(v):
    """Spiky gradient of W(v) with respect to v (float64), zero at 0 and beyond h."""
    v = np.asarray(v, np.float64)
    r = np.linalg.norm(v, axis=-1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        g = -45.0 / (math.pi * H ** 6) * (H - r) ** 2 * v / r
    return np.where((r > 0) & (r < H), g, 0.0)


def pairs(pos):
    """v_ij = p_i - p_j over all pairs (i, j), float64, shape (n, n, 3)."""
    p = pos[:, :3].astype(np.float64)
    return p[:, None, :] - p[None, :, :]


def density64(pos):
    return w64(np.linalg.norm(pairs(pos), axis=-1)).sum(1)


def lambda64(pos, rho0, eps):
    v = pairs(pos)
    g = gradw64(v)
    c = density64(pos) / rho0 - 1.0
    sum_sq = (g * g).sum(-1).sum(1)
    sum_g = g.sum(1)
    return -c / ((sum_sq + (sum_g * sum_g).sum(-1)) / rho0 ** 2 + eps)


def s_corr64(v, k, n, dq):
    return -k * (w64(np.linalg.norm(v, axis=-1)) / w64(dq * H)) ** n


def correction64(pos, lam, rho0, k=0.0, n=4, dq=0.2):
    v = pairs(pos)
    s = lam[:, None] + lam[None, :] + (s_corr64(v, k, n, dq) if k else 0.0)
    return (s[..., None] * gradw64(v)).sum(1) / rho0


def vorticity64(pos, vel):
    dv = vel[None, :, :3].astype(np.float64) - vel[:, None, :3].astype(np.float64)
    return np.cross(dv, gradw64(pairs(pos))).sum(1)


def xsph64(pos, vel, c):
    dv = vel[None, :, :3].astype(np.float64) - vel[:, None, :3].astype(np.float64)
    w = w64(np.linalg.norm(pairs(pos), axis=-1))
    return vel[:, :3].astype(np.float64) + c * (dv * w[..., None]).sum(1)


def confinement64(pos, vort, vel, eps_v, dt):
    wn = np.linalg.norm(vort[:, :3].astype(np.float64), axis=1)
    nv = (wn[None, :, None] * gradw64(pairs(pos))).sum(1)
    ln = np.linalg.norm(nv, axis=1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        nn = np.where(ln > 0, nv / ln, 0.0)
    return vel[:, :3].astype(np.float64) + eps_v * np.cross(nn, vort[:, :3].astype(np.float64)) * dt


def cluster(n=64, seed=0, side=0.17, centre=(0.31, -0.17, 0.23)):
    """Sentinel at the (-,-,-) corner, then n particles uniform in a cube of the given side:
    interior particles with 30-50 neighbours and a free surface all around."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n + 1, 4), np.float32)
    pos[0] = f4(-0.995, -0.995, -0.995)
    pos[1:, :3] = (np.asarray(centre) + rng.uniform(-side / 2, side / 2, (n, 3))).astype(np.float32)
    return pos


def params(O, **kw):
    base = dict(restDensity=8078.2, relaxCFM=600.0, timeStep=np.float32(1 / 120),
                isArtPressureEnabled=0, isVorticityConfEnabled=1, vorticityConfCoeff=2e-4,
                xsphViscosityCoeff=0.01, artPressureCoeff=0.1, artPressureExp=4,
                artPressureRadius=0.2)
    base.update(kw)
    return O.fluid_params(**base)


def rel_err(got, want):
    scale = np.abs(want).max()
    return np.abs(np.asarray(got, np.float64) - want).max() / scale


# ---------------------------------------------------------------- density (a12)

def test_density_matches_all_pairs_float64(oracle):
    for seed in range(3):
        pos = cluster(64, seed)
        g = Grid(oracle, pos)
        dens = g.unsort(g.density(params(oracle)))
        want = density64(pos[1:])
        # each particle's sum of 20-60 float32 terms: a few ulps of the largest term
        assert np.abs(dens[1:] - want).max() <= 2e-6 * want.max()
        nb = (np.linalg.norm(pairs(pos[1:]), axis=-1) < H).sum(1)
        assert nb.max() >= 30 and nb.min() <= 20  # interior and surface both present


def test_density_sees_neighbours_across_cell_faces(oracle):
    # two particles 0.6h apart straddling a cell face, and the same pair inside one cell
    for a, b in (((0.095, 0.0, 0.0), (0.155, 0.0, 0.0)), ((0.01, 0.01, 0.01), (0.07, 0.01, 0.01)),
                 ((0.1, 0.1, 0.1), (0.05, 0.05, 0.1))):
        pos = np.array([f4(-0.995, -0.995, -0.995), f4(*a), f4(*b)], np.float32)
        g = Grid(oracle, pos)
        d = g.unsort(g.density(params(oracle)))
        r = np.linalg.norm(np.subtract(a, b))
        want = w64(0.0) + w64(r)
        assert d[1] == pytest.approx(want, rel=1e-6) and d[2] == pytest.approx(want, rel=1e-6)


# ---------------------------------------------------------------- lambda (a13)

@pytest.mark.parametrize("eps", [600.0, 1.0, 1e-3])
def test_lambda_two_particle_closed_form(oracle, eps):
    r = 0.05
    pos = np.array([f4(-0.995, -0.995, -0.995), f4(0.2, 0.3, 0.4), f4(0.2 + r, 0.3, 0.4)],
                   np.float32)
    rho0 = 8078.2
    p = params(oracle, relaxCFM=eps, restDensity=rho0)
    g = Grid(oracle, pos)
    dens_s = g.density(p)
    lam = g.lam(p, dens_s)
    # rho = W(0) + W(r); |grad W(r)| = 45/(pi h^6) (h - r)^2; the pair term appears once in
    # sum |grad W|^2 and once in |sum grad W|^2 (the self term's gradient is zero)
    rho = w64(0.0) + w64(r)
    gm = 45.0 / (math.pi * H ** 6) * (H - r) ** 2
    want = -(rho / rho0 - 1.0) / (2.0 * gm * gm / rho0 ** 2 + eps)
    assert lam[1] == pytest.approx(want, rel=2e-6)
    assert lam[2] == pytest.approx(want, rel=2e-6)
    # a dropped |sum grad W|^2, rho0 in place of rho0^2, or no eps each miss by far more
    # (eps is negligible at 1e-3, where its absence cannot show)
    wrongs = [-(rho / rho0 - 1) / (gm * gm / rho0 ** 2 + eps),
              -(rho / rho0 - 1) / (2 * gm * gm / rho0 + eps)]
    if eps >= 1.0:
        wrongs.append(-(rho / rho0 - 1) / (2 * gm * gm / rho0 ** 2))
    for wrong in wrongs:
        assert abs(lam[1] / wrong - 1) > 1e-3
    # density deficit (C < 0) gives a positive lambda
    assert lam[1] > 0


def test_lambda_matches_all_pairs_float64(oracle):
    for seed, eps in ((1, 600.0), (2, 10.0), (3, 1e-2)):
        pos = cluster(64, seed)
        rho0 = 6000.0
        p = params(oracle, relaxCFM=eps, restDensity=rho0)
        g = Grid(oracle, pos)
        lam = g.lam(p, g.density(p))
        want = lambda64(pos[1:], rho0, eps)
        assert rel_err(lam[1:], want) < 2e-5, seed
        # free-surface particles carry a large |sum grad W|^2 term
        gsum = gradw64(pairs(pos[1:])).sum(1)
        assert (np.linalg.norm(gsum, axis=1) > 1e5).any()


# ---------------------------------------------------------------- delta p (a14, a8)

def test_correction_two_particle_closed_form(oracle):
    r = 0.04
    pa, pb = (0.2, 0.3, 0.4), (0.2 + 0.6 * r, 0.3 + 0.8 * r, 0.4)
    pos = np.array([f4(-0.995, -0.995, -0.995), f4(*pa), f4(*pb)], np.float32)
    rho0 = 8078.2
    g = Grid(oracle, pos)
    lam = np.array([0.0, 1.5e-3, -4e-4], np.float32)
    v = pos[1, :3].astype(np.float64) - pos[2, :3].astype(np.float64)
    for k in (0.0, 0.1):
        p = params(oracle, restDensity=rho0, isArtPressureEnabled=int(k > 0), artPressureCoeff=k)
        corr = g.correction(p, lam)
        s = 1.5e-3 - 4e-4 + (s_corr64(v, k, 4, 0.2) if k else 0.0)
        want_a = s * gradw64(v) / rho0
        assert np.allclose(corr[1, :3], want_a, rtol=1e-5, atol=1e-6 * np.abs(want_a).max())
        assert np.allclose(corr[2, :3], -want_a, rtol=1e-5, atol=1e-6 * np.abs(want_a).max())
        assert corr[1, 3] == 0.0
        # a positive pair coefficient pulls a towards b; s_corr (negative) pushes apart
        toward_b = float(np.dot(corr[1, :3], -v))
        assert (toward_b > 0) == (s > 0)
    # with k = 0.1 at r = 0.4 h, s_corr = -0.1 (W(0.4h)/W(0.2h))^4 dominates the lambdas
    assert s_corr64(v, 0.1, 4, 0.2) < -(1.5e-3 - 4e-4)


def test_art_pressure_exponent_and_radius(oracle):
    L = oracle.lib()
    c = oracle.make_consts(BOX, RES, 100, 9.8)
    for n_exp in (1, 2, 4, 6):
        for dq in (0.1, 0.2, 0.3):
            p = params(oracle, isArtPressureEnabled=1, artPressureCoeff=0.05, artPressureExp=n_exp,
                       artPressureRadius=dq)
            for r in (0.0, 0.03, 0.07):
                v = f4(r, 0, 0)
                got = L.orc_art_pressure(v.ctypes.data, ctypes.byref(p), ctypes.byref(c))
                want = -0.05 * (w64(r) / w64(dq * H)) ** n_exp
                assert got == pytest.approx(want, rel=1e-5)


def test_correction_matches_all_pairs_float64(oracle):
    rng = np.random.default_rng(7)
    for seed, k in ((4, 0.0), (5, 0.1), (6, 0.001)):
        pos = cluster(64, seed)
        rho0 = 8078.2
        lam = np.zeros(65, np.float32)
        lam[1:] = rng.normal(0, 1e-3, 64).astype(np.float32)
        p = params(oracle, restDensity=rho0, isArtPressureEnabled=int(k > 0), artPressureCoeff=k)
        g = Grid(oracle, pos)
        corr = g.correction(p, lam)
        want = correction64(pos[1:], lam[1:].astype(np.float64), rho0, k)
        assert rel_err(corr[1:, :3], want) < 2e-5, seed


# ---------------------------------------------------------------- vorticity (a17)

def test_vorticity_of_uniform_velocity_is_zero(oracle):
    pos = cluster(64, 8)
    vel = np.tile(f4(0.3, -1.2, 0.7), (65, 1))
    g = Grid(oracle, pos)
    w = g.vorticity(params(oracle), vel)
    assert not w.any()


def test_vorticity_two_particle_closed_form(oracle):
    pos = np.array([f4(-0.995, -0.995, -0.995), f4(0.2, 0.3, 0.4), f4(0.23, 0.33, 0.42)], np.float32)
    vel = np.array([f4(0, 0, 0), f4(0.1, 0.0, 0.5), f4(-0.4, 0.9, 0.2)], np.float32)
    g = Grid(oracle, pos)
    w = g.vorticity(params(oracle), vel)
    v = pos[1, :3].astype(np.float64) - pos[2, :3].astype(np.float64)
    dv = vel[2, :3].astype(np.float64) - vel[1, :3].astype(np.float64)
    want_a = np.cross(dv, gradw64(v))
    # for b both factors change sign: the same vector
    assert np.allclose(w[1, :3], want_a, rtol=1e-5, atol=1e-6 * np.abs(want_a).max())
    assert np.allclose(w[2, :3], want_a, rtol=1e-5, atol=1e-6 * np.abs(want_a).max())
    # the operands' order matters: the reversed cross product is the negative
    assert not np.allclose(w[1, :3], np.cross(gradw64(v), dv), rtol=1e-2)
    assert w[1, 3] == 0.0


def test_vorticity_of_rigid_rotation_is_antiparallel_to_spin(oracle):
    # v = Omega x p: sum_j (v_j - v_i) x grad_i W(p_i - p_j) = -sum_j f(r) (r r^T - r^2 I) Omega / r
    # with f > 0, i.e. -(2/3) sum_j f(r) r Omega on an isotropic neighbourhood: antiparallel
    # to Omega (the sum uses the gradient with respect to p_i)
    rng = np.random.default_rng(9)
    pos = np.zeros((1 + 15 ** 3, 4), np.float32)
    pos[0] = f4(-0.995, -0.995, -0.995)
    ax = (np.arange(15) - 7) * 0.05
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    pos[1:, :3] = (np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + 0.1
                   + rng.uniform(-0.002, 0.002, (15 ** 3, 3))).astype(np.float32)
    omega = np.array([0.3, -0.5, 0.8])
    vel = np.zeros_like(pos)
    vel[1:, :3] = np.cross(omega, pos[1:, :3].astype(np.float64)).astype(np.float32)
    g = Grid(oracle, pos)
    w = g.vorticity(params(oracle), vel)[1:, :3].astype(np.float64)
    inner = (np.abs(pos[1:, :3] - 0.1) < 0.2).all(1)
    cosang = (w[inner] @ omega) / (np.linalg.norm(w[inner], axis=1) * np.linalg.norm(omega))
    assert (cosang < -0.999).all()
    assert np.allclose(w, vorticity64(pos[1:], vel[1:]), rtol=0,
                       atol=2e-5 * np.abs(w).max())


# ---------------------------------------------------------------- confinement (a18)

def test_confinement_without_vorticity_keeps_velocity(oracle):
    pos = cluster(64, 10)
    rng = np.random.default_rng(10)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(size=(65, 3)).astype(np.float32)
    g = Grid(oracle, pos)
    # N = 0 everywhere: normalize(0) = 0, so v is unchanged bit for bit
    out = g.confinement(params(oracle), np.zeros_like(pos), vel)
    assert np.array_equal(out, vel)


def test_confinement_two_particle_direction(oracle):
    pos = np.array([f4(-0.995, -0.995, -0.995), f4(0.2, 0.3, 0.4), f4(0.26, 0.3, 0.4)], np.float32)
    vort = np.array([f4(0, 0, 0), f4(0.0, 0.0, 2.0), f4(0.0, 5.0, 0.0)], np.float32)
    vel = np.array([f4(0, 0, 0), f4(0.1, 0.2, 0.3), f4(-0.3, 0.0, 0.1)], np.float32)
    eps_v, dt = 2e-4, float(np.float32(1 / 120))
    g = Grid(oracle, pos)
    out = g.confinement(params(oracle, vorticityConfCoeff=eps_v, timeStep=dt), vort, vel)
    # N_a = |omega_b| grad W(p_a - p_b) points from a towards b (+x): N/|N| = +x;
    # +x cross omega_a = (1,0,0) x (0,0,2) = (0,-2,0)
    assert out[1, :3] == pytest.approx(vel[1, :3] + eps_v * dt * np.array([0, -2.0, 0]), rel=1e-6,
                                       abs=1e-9)
    # N_b points towards a (-x): (-1,0,0) x (0,5,0) = (0,0,-5)
    assert out[2, :3] == pytest.approx(vel[2, :3] + eps_v * dt * np.array([0, 0, -5.0]), rel=1e-6,
                                       abs=1e-9)
    want = confinement64(pos[1:], vort[1:], vel[1:], eps_v, dt)
    assert np.allclose(out[1:, :3], want, rtol=1e-6, atol=1e-9)


def test_confinement_matches_all_pairs_float64(oracle):
    rng = np.random.default_rng(11)
    pos = cluster(64, 11)
    vort = np.zeros_like(pos)
    vort[1:, :3] = rng.normal(size=(64, 3)).astype(np.float32)
    vel = np.zeros_like(pos)
    vel[1:, :3] = rng.normal(size=(64, 3)).astype(np.float32)
    eps_v, dt = 0.05, float(np.float32(1 / 120))
    g = Grid(oracle, pos)
    out = g.confinement(params(oracle, vorticityConfCoeff=eps_v, timeStep=dt), vort, vel)
    want = confinement64(pos[1:], vort[1:], vel[1:], eps_v, dt)
    dvel = out[1:, :3].astype(np.float64) - vel[1:, :3]
    dwant = want - vel[1:, :3]
    # the increment is read back through v + dv in float32: allow two ulps of v
    tol = 1e-5 * np.abs(dwant).max() + 2 * np.spacing(np.abs(vel[1:, :3])).astype(np.float64)
    assert (np.abs(dvel - dwant) <= tol).all()
    assert np.abs(dvel).max() > 1e-4


# ---------------------------------------------------------------- XSPH (a19)

def test_xsph_keeps_uniform_velocity(oracle):
    pos = cluster(64, 12)
    vel = np.tile(f4(0.3, -1.2, 0.7), (65, 1))
    g = Grid(oracle, pos)
    out = g.xsph(params(oracle, xsphViscosityCoeff=0.3), vel)
    assert np.array_equal(out, vel)


def test_xsph_two_particle_closed_form_and_momentum(oracle):
    pos = np.array([f4(-0.995, -0.995, -0.995), f4(0.2, 0.3, 0.4), f4(0.23, 0.34, 0.4)], np.float32)
    vel = np.array([f4(0, 0, 0), f4(1.0, 0.0, -2.0), f4(0.0, 3.0, 0.5)], np.float32)
    c = 0.01
    g = Grid(oracle, pos)
    out = g.xsph(params(oracle, xsphViscosityCoeff=c), vel)
    w = w64(0.05)
    dv = vel[2, :3].astype(np.float64) - vel[1, :3]
    assert out[1, :3] == pytest.approx(vel[1, :3] + c * w * dv, rel=1e-6)
    assert out[2, :3] == pytest.approx(vel[2, :3] - c * w * dv, rel=1e-6)
    # the pair exchange conserves the sum of velocities
    assert (out[1, :3].astype(np.float64) + out[2, :3]) == pytest.approx(
        vel[1, :3].astype(np.float64) + vel[2, :3], abs=1e-5)


def test_xsph_matches_all_pairs_float64_and_conserves_momentum(oracle):
    rng = np.random.default_rng(13)
    pos = cluster(64, 13)
    vel = np.zeros_like(pos)
    vel[1:, :3] = rng.normal(size=(64, 3)).astype(np.float32)
    c = 1e-4
    g = Grid(oracle, pos)
    out = g.xsph(params(oracle, xsphViscosityCoeff=c), vel)
    want = xsph64(pos[1:], vel[1:], c)
    d_got = out[1:, :3].astype(np.float64) - vel[1:, :3]
    d_want = want - vel[1:, :3]
    assert np.allclose(d_got, d_want, rtol=0, atol=2e-3 * np.abs(d_want).max())
    assert np.abs(d_got.sum(0)).max() < 1e-3 * np.abs(d_got).max() * 8


# ---------------------------------------------------------------- per-particle updates (a10, a11, a16)

def test_predict_position_hand_values(oracle):
    L = oracle.lib()
    c = oracle.make_consts((4, 2, 2), (40, 20, 20), 100, 9.8)
    p = params(oracle, timeStep=0.01)
    pos = np.array([f4(1.0, 2.0, 3.0), f4(-0.5, 0.25, 0.0)], np.float32)
    vel = np.array([f4(0.5, -1.0, 2.0), f4(0.0, 0.0, 0.0, 0.0)], np.float32)
    pred = np.zeros_like(pos)
    L.orc_fld_predict(pos.ctypes.data, vel.ctypes.data, ctypes.byref(p), pred.ctypes.data, 2,
                      ctypes.byref(c))
    # pred = pos + (vel + (0, -g, 0) dt) dt; vel is not touched
    assert pred[0].tolist() == pytest.approx([1.005, 2.0 - 0.01098, 3.02, 0.0], rel=1e-6)
    assert pred[1].tolist() == pytest.approx([-0.5, 0.25 - 0.00098, 0.0, 0.0], rel=1e-6)
    assert vel[0].tolist() == [0.5, -1.0, 2.0, 0.0]


def test_boundary_clamp_is_asymmetric(oracle):
    L = oracle.lib()
    c = oracle.make_consts((4, 2, 2), (40, 20, 20), 100, 9.8)
    pred = np.array([f4(5, 5, 5, 3.0), f4(-5, -5, -5, -1.0), f4(0.5, -0.25, 0.75, 0.0),
                     f4(1.95, -0.995, 0.95)], np.float32)
    L.orc_fld_boundary(pred.ctypes.data, 4, ctypes.byref(c))
    # walls at -W + 0.01 and +W - 0.1 (fluids.cl:435-439); w is clamped to 0
    assert pred[0].tolist() == pytest.approx([1.9, 0.9, 0.9, 0.0], rel=1e-7)
    assert pred[1].tolist() == pytest.approx([-1.99, -0.99, -0.99, 0.0], rel=1e-7)
    assert pred[2].tolist() == [0.5, -0.25, 0.75, 0.0]
    assert pred[3].tolist() == pytest.approx([1.9, -0.99, 0.9, 0.0], rel=1e-7)


def test_update_velocity_divides_by_dt_plus_eps_and_clamps(oracle):
    L = oracle.lib()
    c = oracle.make_consts((4, 2, 2), (40, 20, 20), 100, 9.8)
    p = params(oracle, timeStep=0.01)
    pos = np.array([f4(0, 0, 0, 0), f4(1, 1, 1, 0)], np.float32)
    pred = np.array([f4(0.1, -1.0, 1e-4, 0.5), f4(1.0, 1.0, 1.0, 0.0)], np.float32)
    vel = np.zeros_like(pos)
    L.orc_fld_update_vel(pred.ctypes.data, pos.ctypes.data, ctypes.byref(p), vel.ctypes.data, 2,
                         ctypes.byref(c))
    dt = float(np.float32(0.01)) + 1e-8
    # 0.1 / (dt + 1e-8) = 9.99999..., distinguishable from 0.1 / dt = 10 in float32
    assert vel[0, 0] == pytest.approx(0.1 / dt, rel=1e-7)
    assert vel[0, 0] != np.float32(10.0)
    assert vel[0, 1] == -30.0          # -99.99... clamped to -MAX_VEL
    assert vel[0, 2] == pytest.approx(1e-4 / dt, rel=1e-6)
    assert vel[0, 3] == 30.0           # the w lane is clamped too
    assert not vel[1].any()
