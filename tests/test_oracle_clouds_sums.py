"""Pins of the oracle's Clouds neighbour sums and per-particle updates (clouds.cl:257-966).

The checks are independent of the oracle's code.  They use a grid-free all-pairs evaluation
in float64, written from the textbook SPH kernels (Mueller et al. 2003), the PBF constraint
(Macklin and Mueller 2013), and the temperature constraint as clouds.cl states it.  They
run on a cluster that straddles the periodic x and z walls.  Each pair enters once, with
its minimum-image vector: x and z are periodic with period 2W, and y has no images.
Closed forms cover the per-particle updates.

    lambda_i  = -(rho_i/rho0 - 1) / ((sum_j |g_ij|^2 + |sum_j g_ij|^2) / rho0^2 + eps)
    dp_i      = sum_j (lambda_i + lambda_j + s_corr(v_ij)) g_ij / rho0
    omega_i   = sum_j (v_j - v_i) x g_ij;   N_i = sum_j |omega_j| g_ij
    xsph_i    = v_i + c sum_j (v_j - v_i) W_ij
    lap_i     = sum_j (T_i - T_j) (v_ij . g_ij) / (|v_ij|^2 + 1e-8) / rho0_c
    d_ij      = (v_ij . g_ij) / (|v_ij|^2 rho0_c + 1e-8)
    lamT_i    = -lap_i / (sum_j d_ij^2 + (sum_j d_ij)^2 + eps_c)
    corrT_i   = sum_j (lamT_i + lamT_j) d_ij

Here g_ij is the spiky gradient at v_ij = p_i - p_j (taken at its image).  A missing image,
a dropped term, a wrong sign or a swapped operand moves the results by much more than the
float32 rounding the tolerances allow for.
"""
import ctypes
import math

import numpy as np
import pytest

BOX = (4.0, 4.0, 4.0)
RES = (12, 12, 12)  # h = 1/3
PERIOD = 4.0  # 2W in x and z


class Cluster:
    """A sentinel alone in cell 0 takes sorted slot 0, which fillStartCell leaves out of its
    cell (grid.cl:95-108).  After it come n particles in a cube of the given side, centred
    on the (+W, +W) corner of x and z, so each one sees neighbours through the x image, the
    z image and both.  The oracle's grid, sort and cell table are built over them."""

    def __init__(self, O, n=90, seed=0, side=0.62, y0=0.0, fluid=None, cloud=None):
        self.O = O
        rng = np.random.default_rng(seed)
        pos = np.zeros((n + 1, 4), np.float32)
        pos[0, :3] = (-1.99, -1.99, -1.99)
        u = rng.uniform(-side / 2, side / 2, (n, 3))
        x = np.mod(2.0 + u[:, 0] + 2.0, PERIOD) - 2.0
        z = np.mod(2.0 + u[:, 2] + 2.0, PERIOD) - 2.0
        pos[1:, 0], pos[1:, 1], pos[1:, 2] = x, y0 + u[:, 1], z
        self.n = n + 1
        self.max_n = 512
        self.fluid = fluid if fluid is not None else O.fluid_params(restDensity=450.0)
        self.cloud = cloud if cloud is not None else O.cloud_params(restDensity=450.0)
        self.f = O.OracleClouds(BOX, RES, self.max_n, self.n, self.fluid, self.cloud, 1)
        self.f.upload(pos, np.zeros_like(pos))
        L, f, c, pt = O.lib(), self.f, ctypes.byref(self.f.c), O.ptr
        L.orc_fill_cell_ids(pt(f.pos), pt(f.key), self.n, c)
        assert L.orc_radix_sort(pt(f.key), pt(f.idx), self.max_n) == 0
        assert f.idx[0] == 0, "the sentinel must take sorted slot 0"
        self.pred = np.ascontiguousarray(f.pos[f.idx])
        L.orc_reset_start_end(pt(f.se), c)
        L.orc_fill_start(pt(f.key), pt(f.se), self.n, c)
        L.orc_fill_end(pt(f.key), pt(f.se), self.n, self.max_n, c)
        L.orc_adjust_end(pt(f.se), c)
        self.q = self.pred[1: self.n, :3].astype(np.float64)  # sorted, sentinel left out

    # arrays over the sorted slots [1, n) <-> the oracle's maxN buffers
    def f1(self, a):
        out = np.zeros(self.max_n, np.float32)
        out[1: self.n] = a
        return out

    def f4(self, a):
        out = np.zeros((self.max_n, 4), np.float32)
        out[1: self.n, :3] = a
        return out

    def call(self, name, *args):
        getattr(self.O.lib(), name)(*args, self.n, ctypes.byref(self.f.c))

    def pairs(self):
        """Minimum-image v_ij = p_i - p_j over the cluster, float64, (n, n, 3)."""
        v = self.q[:, None, :] - self.q[None, :, :]
        for k in (0, 2):
            v[..., k] -= PERIOD * np.round(v[..., k] / PERIOD)
        return v


def w64(v, h):
    r2 = (np.asarray(v) ** 2).sum(-1)
    return np.where(r2 < h * h, 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r2) ** 3, 0.0)


def gradw64(v, h):
    r = np.linalg.norm(v, axis=-1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        g = -45.0 / (math.pi * h ** 6) * (h - r) ** 2 * v / r
    return np.where((r > 0) & (r < h), g, 0.0)


def rel_err(got, want):
    return np.abs(np.asarray(got, np.float64) - want).max() / np.abs(want).max()


def _wraps(cl):
    """How many pairs within h see each other only through an image."""
    v = cl.q[:, None, :] - cl.q[None, :, :]
    r_direct = np.linalg.norm(v, axis=-1)
    r_image = np.linalg.norm(cl.pairs(), axis=-1)
    h = float(cl.f.c.effectRadius)
    return int(((r_image < h) & (r_direct >= h)).sum())


# ------------------------------------------------------------------------- PBF constraint

@pytest.mark.parametrize("seed,eps", [(0, 600.0), (1, 10.0), (2, 0.05)])
def test_lambda_matches_all_pairs_with_images(oracle, seed, eps):
    O = oracle
    cl = Cluster(O, seed=seed, fluid=O.fluid_params(restDensity=300.0, relaxCFM=eps))
    h = float(cl.f.c.effectRadius)
    assert _wraps(cl) > 100
    dens = np.zeros(cl.max_n, np.float32)
    lam = np.zeros(cl.max_n, np.float32)
    pt = O.ptr
    cl.call("orc_cld_density", pt(cl.pred), pt(cl.f.se), ctypes.byref(cl.fluid), pt(dens))
    cl.call("orc_cld_lambda", pt(cl.pred), pt(dens), pt(cl.f.se), ctypes.byref(cl.fluid), pt(lam))
    v = cl.pairs()
    rho = w64(v, h).sum(1)
    assert rel_err(dens[1: cl.n], rho) < 2e-6
    g = gradw64(v, h)
    sum_sq = (g * g).sum(-1).sum(1) + (g.sum(1) ** 2).sum(-1)
    rho0 = 300.0
    want = -(dens[1: cl.n].astype(np.float64) / rho0 - 1.0) / (sum_sq / rho0 ** 2 + eps)
    assert rel_err(lam[1: cl.n], want) < 2e-5
    # without the images, or without |sum g|^2, the values are far off
    v0 = cl.q[:, None, :] - cl.q[None, :, :]
    g0 = gradw64(v0, h)
    no_img = -(dens[1: cl.n] / rho0 - 1.0) / (((g0 * g0).sum(-1).sum(1) + (g0.sum(1) ** 2).sum(-1))
                                            / rho0 ** 2 + eps)
    no_sum = -(dens[1: cl.n] / rho0 - 1.0) / ((g * g).sum(-1).sum(1) / rho0 ** 2 + eps)
    if eps < 100:
        assert rel_err(no_img, want) > 1e-2 and rel_err(no_sum, want) > 1e-2


@pytest.mark.parametrize("k", [0.0, 0.1])
def test_correction_matches_all_pairs_with_images(oracle, k):
    O = oracle
    fluid = O.fluid_params(restDensity=300.0, isArtPressureEnabled=int(k > 0), artPressureCoeff=k,
                           artPressureExp=4, artPressureRadius=0.2)
    cl = Cluster(O, seed=3, fluid=fluid)
    h = float(cl.f.c.effectRadius)
    rng = np.random.default_rng(3)
    lam = rng.normal(0, 1e-3, cl.n - 1)
    corr = np.zeros((cl.max_n, 4), np.float32)
    pt = O.ptr
    lam_s = cl.f1(lam)
    cl.call("orc_cld_correction", pt(lam_s), pt(cl.f.se), pt(cl.pred), ctypes.byref(fluid), pt(corr))
    v = cl.pairs()
    lam32 = lam.astype(np.float32).astype(np.float64)
    s = lam32[:, None] + lam32[None, :]
    if k:
        s = s - k * (w64(v, h) / w64(np.array([0.2 * h, 0, 0]), h)) ** 4
    want = (s[..., None] * gradw64(v, h)).sum(1) / 300.0
    assert rel_err(corr[1: cl.n, :3], want) < 2e-5
    assert not corr[:, 3].any()


def test_predict_and_update_vel_closed_forms(oracle):
    O = oracle
    q = O.cloud_params(timeStep=0.02)
    n = 5
    pos = np.array([[0.1, 0.2, 0.3, 0], [-1.0, 1.5, 0.0, 0], [0, 0, 0, 0], [1, 1, 1, 0],
                    [0.5, -0.5, 0.25, 0]], np.float32)
    vel = np.array([[1.0, 2.0, 3.0, 0], [-0.5, 0.0, 0.25, 0], [0, 0, 0, 0], [40, -50, 0.1, 0],
                    [0.3, 0.3, 0.3, 0]], np.float32)
    buoy = np.array([0.5, -9.81, 0.0, 100.0, 1.25], np.float32)
    pred = np.zeros_like(pos)
    tot = np.zeros_like(pos)
    O.lib().orc_cld_predict(O.ptr(pos), O.ptr(vel), O.ptr(buoy), ctypes.byref(q), O.ptr(pred),
                            O.ptr(tot), n)
    dt = np.float32(0.02)
    pv = vel.astype(np.float64).copy()
    pv[:, 1] += buoy.astype(np.float64) * dt
    assert np.allclose(tot, pv * dt, rtol=1e-6, atol=1e-9)
    assert np.allclose(pred, pos + pv * dt, rtol=1e-6, atol=1e-7)
    # updateVel: clamp(totCorr / (dt + 1e-8), +-30), from the fluid parameters' time step
    fluid = O.fluid_params(timeStep=0.02)
    out = np.zeros_like(pos)
    c = O.make_consts(BOX, RES, 100, 9.81)
    O.lib().orc_cld_update_vel(O.ptr(tot), ctypes.byref(fluid), O.ptr(out), n, ctypes.byref(c))
    want = np.clip(tot.astype(np.float64) / (0.02 + 1e-8), -30.0, 30.0)
    assert np.allclose(out, want, rtol=1e-6, atol=1e-9)
    assert out[3, 0] == 30.0 and out[3, 1] == -30.0  # the clamp, on the 4th particle


# --------------------------------------------------------------- vorticity, confinement, XSPH

def test_vorticity_confinement_xsph_match_all_pairs_with_images(oracle):
    O = oracle
    fluid = O.fluid_params(restDensity=300.0, vorticityConfCoeff=4e-4, xsphViscosityCoeff=0.01,
                           timeStep=0.01)
    cl = Cluster(O, seed=5, fluid=fluid)
    h = float(cl.f.c.effectRadius)
    rng = np.random.default_rng(5)
    vel = rng.normal(0, 0.5, (cl.n - 1, 3)).astype(np.float32)
    pt = O.ptr
    v = cl.pairs()
    g = gradw64(v, h)
    vel64 = vel.astype(np.float64)
    dv = vel64[None, :, :] - vel64[:, None, :]
    # vorticity
    vort = np.zeros((cl.max_n, 4), np.float32)
    vel_s = cl.f4(vel)
    cl.call("orc_cld_vorticity", pt(cl.pred), pt(cl.f.se), pt(vel_s), ctypes.byref(fluid), pt(vort))
    want_w = np.cross(dv, g).sum(1)
    assert rel_err(vort[1: cl.n, :3], want_w) < 2e-5
    assert not np.allclose(vort[1: cl.n, :3], np.cross(g, dv).sum(1), rtol=1e-2)
    # confinement on that vorticity: v += eps (N/|N| x omega) dt, in place
    w32 = vort[1: cl.n, :3].astype(np.float64)
    nvec = (np.linalg.norm(w32, axis=1)[None, :, None] * g).sum(1)
    nn = nvec / np.linalg.norm(nvec, axis=1, keepdims=True)
    want_v = vel64 + 4e-4 * np.cross(nn, w32) * 0.01
    out = cl.f4(vel)
    cl.call("orc_cld_confinement", pt(cl.pred), pt(cl.f.se), pt(vort), ctypes.byref(fluid), pt(out))
    dv_got = out[1: cl.n, :3].astype(np.float64) - vel64
    assert rel_err(dv_got, want_v - vel64) < 1e-2  # the increment is ~1e-6 of |v|: float32 noise
    assert np.allclose(out[1: cl.n, :3], want_v, rtol=0, atol=2e-7)
    # XSPH
    xs = np.zeros((cl.max_n, 4), np.float32)
    cl.call("orc_cld_xsph", pt(cl.pred), pt(cl.f.se), pt(vel_s), ctypes.byref(fluid), pt(xs))
    want_x = vel64 + 0.01 * (dv * w64(v, h)[..., None]).sum(1)
    assert rel_err(xs[1: cl.n, :3] - vel, want_x - vel64) < 2e-5


def test_confinement_increment_direction_two_particles_across_the_wall(oracle):
    """Two particles facing each other through the x wall: N_a points from a to b's image."""
    O = oracle
    fluid = O.fluid_params(vorticityConfCoeff=4e-4, timeStep=0.01)
    c = O.make_consts(BOX, RES, 100, 9.81)
    f = O.OracleClouds(BOX, RES, 512, 3, fluid, O.cloud_params(), 1)
    pos = np.array([[-1.99, -1.99, -1.99, 0], [1.95, 0.1, 0.1, 0], [-1.9, 0.1, 0.1, 0]], np.float32)
    f.upload(pos, np.zeros_like(pos))
    L, pt = O.lib(), O.ptr
    L.orc_fill_cell_ids(pt(f.pos), pt(f.key), 3, ctypes.byref(c))
    assert L.orc_radix_sort(pt(f.key), pt(f.idx), 512) == 0
    pred = np.ascontiguousarray(f.pos[f.idx])
    L.orc_reset_start_end(pt(f.se), ctypes.byref(c))
    L.orc_fill_start(pt(f.key), pt(f.se), 3, ctypes.byref(c))
    L.orc_fill_end(pt(f.key), pt(f.se), 3, 512, ctypes.byref(c))
    L.orc_adjust_end(pt(f.se), ctypes.byref(c))
    a = int(np.flatnonzero(f.idx[:3] == 1)[0])
    b = int(np.flatnonzero(f.idx[:3] == 2)[0])
    vort = np.zeros((512, 4), np.float32)
    vort[a, :3] = (0.0, 0.0, 2.0)
    vort[b, :3] = (0.0, 5.0, 0.0)
    vel = np.zeros((512, 4), np.float32)
    L.orc_cld_confinement(pt(pred), pt(f.se), pt(vort), ctypes.byref(fluid), pt(vel), 3, ctypes.byref(c))
    # a at x = 1.95 sees b's image at 2.1 (+x): +x cross (0,0,2) = (0,-2,0)
    assert vel[a, :3] == pytest.approx([0.0, -2.0 * 4e-4 * 0.01, 0.0], rel=1e-6, abs=1e-12)
    # b sees a's image at -2.05 (-x): -x cross (0,5,0) = (0,0,-5)
    assert vel[b, :3] == pytest.approx([0.0, 0.0, -5.0 * 4e-4 * 0.01], rel=1e-6, abs=1e-12)


# ----------------------------------------------------------------- temperature constraint

@pytest.mark.parametrize("seed,cfm", [(6, 600.0), (7, 1e-3)])
def test_temperature_constraint_matches_all_pairs_with_images(oracle, seed, cfm):
    O = oracle
    cloud = O.cloud_params(restDensity=350.0, relaxCFM=cfm)
    cl = Cluster(O, seed=seed, cloud=cloud)
    h = float(cl.f.c.effectRadius)
    rng = np.random.default_rng(seed)
    temp = (280.0 + rng.normal(0, 5.0, cl.n - 1)).astype(np.float32)
    pt = O.ptr
    lap = np.zeros(cl.max_n, np.float32)
    lamT = np.zeros(cl.max_n, np.float32)
    corT = np.zeros(cl.max_n, np.float32)
    temp_s = cl.f1(temp)
    cl.call("orc_cld_laplacian_temp", pt(cl.pred), pt(temp_s), pt(cl.f.se), ctypes.byref(cloud),
            pt(lap))
    cl.call("orc_cld_lambda_temp", pt(cl.pred), pt(lap), pt(cl.f.se), ctypes.byref(cloud), pt(lamT))
    cl.call("orc_cld_correction_temp", pt(lamT), pt(cl.f.se), pt(cl.pred), ctypes.byref(cloud),
            pt(corT))
    v = cl.pairs()
    g = gradw64(v, h)
    vg = (v * g).sum(-1)
    vv = (v * v).sum(-1)
    t64 = temp.astype(np.float64)
    want_lap = ((t64[:, None] - t64[None, :]) * vg / (vv + 1e-8)).sum(1) / 350.0
    assert rel_err(lap[1: cl.n], want_lap) < 2e-5
    d = vg / (vv * 350.0 + 1e-8)
    lap32 = lap[1: cl.n].astype(np.float64)
    want_lam = -lap32 / ((d * d).sum(1) + d.sum(1) ** 2 + cfm)
    assert rel_err(lamT[1: cl.n], want_lam) < 2e-5
    l32 = lamT[1: cl.n].astype(np.float64)
    want_cor = ((l32[:, None] + l32[None, :]) * d).sum(1)
    assert rel_err(corT[1: cl.n], want_cor) < 5e-5
    # correctTemperature applies 0.3 of the correction
    t = cl.f1(temp)
    O.lib().orc_cld_correct_temp(pt(corT), pt(t), cl.n)
    assert np.allclose(t[1: cl.n], temp + 0.3 * corT[1: cl.n], rtol=1e-7, atol=0)
