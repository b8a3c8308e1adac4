"""GPU parity of the Clouds model (physics/ocl/Clouds.cpp:503-610, clouds.cl) against the
oracle, bit for bit on every buffer the step writes: positions, velocities, predicted and
total-correction positions, temperature, vapour and cloud density, buoyancy, particle ids,
fluid density and lambda, the temperature Laplacian and its lambda, vorticity, and the grid.
"""
import numpy as np
import pytest

from helpers import _eq, log
from oracle import oracle as O
from realtimeparticles_amd import capi, scenes

pytestmark = pytest.mark.gpu


def test_pinned_exp_matches_oracle():
    """The device's exp (double and float) and saturationVaporDensity against the CPU's
    evaluation of the same algorithm, over the whole double range and the model's
    temperatures."""
    rng = np.random.default_rng(5)
    x64 = np.concatenate([rng.uniform(-745.2, 709.8, 200000), rng.uniform(-30, 30, 200000),
                          [0.0, -0.0, np.inf, -np.inf, np.nan, 709.78, -745.13, -708.4, 1e-300]])
    x32 = np.concatenate([rng.uniform(-104, 89, 200000).astype(np.float32),
                          rng.uniform(200, 330, 200000).astype(np.float32),
                          np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 223.0, 313.0, 29.5, 1e-30],
                                   np.float32)])
    y64, y32, s32 = capi.selfcheck_exp(x64, x32)
    L = O.lib()
    r64 = np.array([L.orc_exp64(float(v)) for v in x64])
    r32 = np.array([L.orc_exp32(float(v)) for v in x32], np.float32)
    rs = np.array([L.orc_cld_svd(float(v)) for v in x32], np.float32)
    _eq(y64, r64, "exp (double)")
    _eq(y32, r32, "exp (float)")
    _eq(s32, rs, "saturation vapour density")


def make_pair(sc, iters=None, fluid=None, cloud=None, gravity=None, quad=None):
    """quad: the walk passes with a quad of lanes per particle (True), one lane (False), or
    the model's default for the scene's size (None: quads at these sizes)."""
    fl = sc.params() if fluid is None else fluid
    cl = sc.cloud_params() if cloud is None else cloud
    it = sc.iters if iters is None else iters
    g = sc.gravity if gravity is None else gravity
    gpu = capi.Clouds(sc.max_n, sc.n, sc.box, sc.res, gravity=g)
    if quad is not None:
        gpu.set_quad(quad)
    gpu.set_params(fl, cl)
    gpu.set_nb_jacobi_iters(it)
    gpu.load_state(sc.pos, sc.vel)
    orc = O.OracleClouds(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**fl),
                         O.cloud_params(**cl), it, gravity=g)
    orc.upload(sc.pos, sc.vel)
    return gpu, orc


FIELDS = [
    (capi.FIELD_TEMPERATURE, "temp"), (capi.FIELD_VAPOR_DENSITY, "vapor"),
    (capi.FIELD_CLOUD_DENSITY, "cloud"), (capi.FIELD_BUOYANCY, "buoyancy"),
    (capi.FIELD_PARTICLE_ID, "partID"), (capi.FIELD_DENSITY, "density"),
    (capi.FIELD_LAMBDA, "lambda_"), (capi.FIELD_LAPLACIAN_TEMP, "lap"),
    (capi.FIELD_LAMBDA_TEMP, "lambdaT"), (capi.FIELD_PRED_POS, "pred"),
    (capi.FIELD_TOT_CORR_POS, "totCorr"), (capi.FIELD_VORTICITY, "vort"),
]


def compare(gpu, orc, n, what, grid=True, fields=FIELDS):
    pos, vel = gpu.read_state(n)
    _eq(pos, orc.pos[:n], f"{what}: p_pos")
    _eq(vel, orc.vel[:n], f"{what}: p_vel")
    for field, name in fields:
        _eq(gpu.read_field(field, n), getattr(orc, name)[:n], f"{what}: {name}")
    if grid:
        keys, idx, se = gpu.read_grid()
        _eq(keys, orc.key, f"{what}: p_cellID")
        _eq(idx, orc.idx, f"{what}: sort index")
        _eq(se, orc.se, f"{what}: c_startEndPartID")


def test_initial_state_matches_oracle():
    sc = scenes.cloud_scene("clouds-65k")
    gpu, orc = make_pair(sc)
    m = sc.max_n
    # every slot of the initial fields (cld_initTemperature / cld_initVaporDensity over maxN)
    _eq(gpu.read_field(capi.FIELD_TEMPERATURE, m), orc.temp, "initial temperature")
    _eq(gpu.read_field(capi.FIELD_VAPOR_DENSITY, m), orc.vapor, "initial vapour density")
    _eq(gpu.read_field(capi.FIELD_PARTICLE_ID, m), orc.partID, "particle ids")
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_cumulus_ten_steps_bit_exact(quad):
    """The app's cumulus case (65,536 particles in the lowest quarter of a 10 x 20 x 10 box,
    grid 30 x 60 x 30) with initCloudsJson's parameters, 2 Jacobi iterations; the walk passes
    with a quad of lanes per particle and with one lane."""
    sc = scenes.cloud_scene("clouds-65k")
    gpu, orc = make_pair(sc, quad=quad)
    for step in range(1, 11):
        gpu.update(1)
        orc.step(1)
        log(f"clouds-65k step {step}")
        compare(gpu, orc, sc.n, f"step {step}", grid=step % 5 == 0)
    gpu.close()


@pytest.mark.parametrize("variant", ["no-smoothing", "no-vorticity", "iters-0", "iters-1",
                                     "iters-3", "dim-2", "homogeneous", "strong-heat"])
def test_cloud_variants_bit_exact(variant):
    name = "clouds-65k-homogeneous" if variant == "homogeneous" else "clouds-65k"
    sc = scenes.cloud_scene(name, seed=23)
    fl, cl, iters = sc.params(), sc.cloud_params(), sc.iters
    if variant == "no-smoothing":
        cl["isTempSmoothingEnabled"] = 0
    elif variant == "no-vorticity":
        fl["isVorticityConfEnabled"] = 0
    elif variant.startswith("iters-"):
        iters = int(variant.split("-")[1])
    elif variant == "dim-2":
        fl["dim"] = 2
        cl["dim"] = 2
    elif variant == "strong-heat":
        # fast warming and strong phase transitions: clouds form within the run
        cl.update(groundHeatCoeff=400.0, buoyancyCoeff=2.0, phaseTransitionRate=15.0,
                  initVaporDensityCoeff=1.5, latentHeatCoeff=0.1)
    gpu, orc = make_pair(sc, iters=iters, fluid=fl, cloud=cl)
    for step in range(1, 5):
        gpu.update(1)
        orc.step(1)
        log(f"clouds {variant} step {step}")
        fields = FIELDS if fl["isVorticityConfEnabled"] else FIELDS[:-1]
        if iters == 0:
            fields = [f for f in fields if f[1] not in ("density", "lambda_")]
        compare(gpu, orc, sc.n, f"{variant} step {step}", grid=step == 4, fields=fields)
    if variant == "strong-heat":
        assert orc.cloud[: sc.n].max() > 0.0
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_thirty_one_iterations_bit_exact(quad):
    """31 Jacobi iterations: the lambda of iterations 29 and 30 has no flag bit left, so their
    corrections walk instead of replaying the density passes' survivor lists."""
    sc = scenes.cloud_scene("clouds-65k", seed=47)
    sc.pos, sc.n = sc.pos[:8192].copy(), 8192
    sc.vel = np.zeros_like(sc.pos)
    gpu, orc = make_pair(sc, iters=31, quad=quad)
    for step in range(1, 3):
        gpu.update(1)
        orc.step(1)
        log(f"31 iterations, step {step}")
        compare(gpu, orc, sc.n, f"step {step}", grid=(step == 2))
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_cumulus_particle_count_change(quad):
    """setNbParticles keeps the whole-buffer sort of the stale keys, as for the fluids."""
    sc = scenes.cloud_scene("clouds-65k", seed=29)
    gpu, orc = make_pair(sc, quad=quad)
    gpu.update(2)
    orc.step(2)
    n2 = sc.n - 3000
    gpu.set_nb_particles(n2)
    orc.s.n = n2
    for step in range(3, 6):
        gpu.update(1)
        orc.step(1)
        compare(gpu, orc, n2, f"shrunk step {step}", grid=True)
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_cumulus_hundred_steps(quad):
    """100 steps: the state stays the oracle's, bit for bit, after the clouds start rising."""
    sc = scenes.cloud_scene("clouds-65k", seed=31)
    gpu, orc = make_pair(sc, quad=quad)
    gpu.update(100)
    orc.step(100)
    compare(gpu, orc, sc.n, "step 100")
    pos, _ = gpu.read_state()
    assert np.isfinite(pos).all()
    log(f"clouds-65k after 100 steps: mean height {pos[:, 1].mean():.3f}, "
        f"max temperature {gpu.read_field(capi.FIELD_TEMPERATURE).max():.2f}")
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_capped_cell_and_wall_clusters_bit_exact(quad):
    """400 particles in one cell (the cap of 100 keeps 101 of them in its list, so its z-column
    is not one contiguous range), clusters pressed against the periodic x and z walls and the
    y walls (wrapped neighbours with +-2W shifts), among the cumulus particles."""
    sc = scenes.cloud_scene("clouds-65k", seed=37)
    rng = np.random.default_rng(11)
    extra = []
    # one interior cell of the 30 x 60 x 30 grid (cell size 1/3): x in [0, 1/3), y in [-8, -7.67)
    c = np.zeros((400, 4), np.float32)
    c[:, 0] = rng.uniform(0.01, 0.32, 400)
    c[:, 1] = rng.uniform(-7.99, -7.68, 400)
    c[:, 2] = rng.uniform(0.01, 0.32, 400)
    extra.append(c)
    for axis, side in ((0, 1), (0, -1), (2, 1), (2, -1), (1, -1)):
        w = np.zeros((300, 4), np.float32)
        w[:, :3] = rng.uniform([-4.9, -9.9, -4.9], [4.9, -5.1, 4.9], (300, 3))
        half = sc.box[axis] / 2
        w[:, axis] = side * rng.uniform(half - 0.3, half - 1e-4, 300)
        extra.append(w)
    pos = np.concatenate([sc.pos[:60000]] + extra)
    sc.pos, sc.vel, sc.n = pos, np.zeros_like(pos), pos.shape[0]
    gpu, orc = make_pair(sc, quad=quad)
    for step in range(1, 5):
        gpu.update(1)
        orc.step(1)
        log(f"capped cell and wall clusters, step {step}")
        compare(gpu, orc, sc.n, f"step {step}", grid=True)
        if step == 1:
            _, _, se = gpu.read_grid()
            assert (se[:, 1].astype(np.int64) - se[:, 0] == 100).any()  # a capped cell
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_dense_ball_overflows_the_survivor_lists_bit_exact(quad):
    """600 particles in a ball of radius 0.2 around a cell corner (about 75 per cell, under
    the cap of 100): each keeps several hundred neighbours, more than a survivor list holds
    (128), so the passes that replay lists walk for those particles."""
    sc = scenes.cloud_scene("clouds-65k", seed=41)
    rng = np.random.default_rng(13)
    d = rng.normal(size=(600, 3))
    d *= (0.2 * rng.uniform(0.0, 1.0, (600, 1)) ** (1.0 / 3.0)) / np.linalg.norm(d, axis=1, keepdims=True)
    ball = np.zeros((600, 4), np.float32)
    ball[:, :3] = np.array([1.0, -6.0, 1.0]) + d  # (1, -6, 1) is a corner of cells of size 1/3
    pos = np.concatenate([sc.pos[:64000], ball])
    sc.pos, sc.vel, sc.n = pos, np.zeros_like(pos), pos.shape[0]
    gpu, orc = make_pair(sc, quad=quad)
    for step in range(1, 4):
        gpu.update(1)
        orc.step(1)
        log(f"dense ball, step {step}")
        compare(gpu, orc, sc.n, f"step {step}", grid=(step == 3))
    _, _, se = gpu.read_grid()
    assert (se[:, 1].astype(np.int64) - se[:, 0] + 1).max() <= 100  # no cell was capped
    gpu.close()


@pytest.mark.parametrize("quad", [True, False])
def test_infinite_velocity_spreads_nan_bit_exact(quad):
    """One particle with an infinite velocity: its predicted position is infinite, so the
    gradient terms of its neighbours turn NaN and so do the lambda, vorticity and confined
    velocity buffers the replaying passes multiply.  Those passes then walk instead of
    replaying their lists (a candidate left out of a list need not contribute +-0 with
    non-finite data), and every buffer still equals the oracle's, NaN for NaN."""
    sc = scenes.cloud_scene("clouds-65k", seed=43)
    sc.vel = np.zeros_like(sc.pos)
    sc.vel[1234, 0] = np.inf
    gpu, orc = make_pair(sc, quad=quad)
    for step in range(1, 3):
        gpu.update(1)
        orc.step(1)
        log(f"infinite velocity, step {step}")
        compare(gpu, orc, sc.n, f"step {step}", grid=(step == 2))
    assert np.isnan(gpu.read_field(capi.FIELD_LAMBDA)).any()
    gpu.close()
