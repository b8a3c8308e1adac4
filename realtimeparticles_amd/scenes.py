"""Seeded synthetic scenes for the PBF dam break and the 3-D boids flock.

Every workload in the benchmark and the test-suite is generated here, so the HIP path and
the CPU oracle always start from the same bytes.  Random numbers come from SplitMix64
(counter-based, so any slice of a stream can be generated independently) mapped to 24-bit
uniform floats in [0, 1).

Scene sizes:
  boids-8k      8,192 boids in a 2^3 box, 20^3 grid (cell = vision radius 0.1)
  pbf-65k       64x32x32 jittered lattice, 4x2x2 box, 40x20x20 grid, h = 0.1
  pbf-1m        128x128x64 lattice, 8x8x4 box, 80x80x40 grid
  pbf-4m        256x128x128 lattice, 16x8x8 box, 160x80x80 grid
Each PBF lattice fills 80 % of every box edge from the (-x,-y,-z) corner, spacing 0.5 h,
jitter U[-0.05 h, +0.05 h] per axis, velocity 0 (the dam break of
physics/ocl/Fluids.cpp:341-345 with the lattice of utils/Geometry.cpp:198-229).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start..start+count-1 of the SplitMix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform24(seed: int, start: int, count: int) -> np.ndarray:
    """Uniform float32 in [0, 1) with 24 random bits (exact in float32)."""
    z = splitmix64(seed, start, count) >> np.uint64(40)
    return (z.astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)


def lattice_rest_density(h: float) -> float:
    """Closed-form poly6 density of a simple-cubic lattice of spacing h/2, unit mass.

    Neighbours with r < h sit at |offset|^2 in {0, 1, 2, 3} (h/2)^2 with multiplicities
    1, 6, 12, 8, so rho0 = 315/(64 pi h^3) * (1 + 6(3/4)^3 + 12(1/2)^3 + 8(1/4)^3)
    = 315 * 5.15625 / (64 pi h^3).
    """
    s = 1.0 + 6 * 0.75 ** 3 + 12 * 0.5 ** 3 + 8 * 0.25 ** 3
    return 315.0 * s / (64.0 * math.pi * h ** 3)


@dataclasses.dataclass
class FluidScene:
    name: str
    box: tuple
    res: tuple
    lattice: tuple
    n: int
    max_n: int
    pos: np.ndarray  # (n, 4) float32
    vel: np.ndarray  # (n, 4) float32
    rest_density: float
    gravity: float = 9.8
    iters: int = 4
    time_step: float = 1.0 / 120.0

    def params(self, mapping: str = "literal") -> dict:
        """FluidKernelInputs values for this scene.

        "literal": the dam-break parameter set mapped one to one (rest density = lattice
        density, relaxation 600, s_corr k=0.1 n=4 dq=0.2h, XSPH c=0.01, vorticity 2e-4).
        With the reference's kernels (unit mass, XSPH summing (v_j - v_i) W without a
        1/rho_j factor) these values blow the lattice apart within the first substep and
        collapse the fluid into point clumps after a few more.
        "stable": a variant for diagnostics only -- XSPH and vorticity coefficients divided
        by rho0 (the missing 1/rho_j) and the reference's own s_corr coefficient 0.001.
        """
        rho0 = float(np.float32(self.rest_density))
        p = dict(restDensity=rho0, relaxCFM=600.0,
                 timeStep=float(np.float32(self.time_step)), dim=3,
                 isArtPressureEnabled=1, artPressureRadius=0.2, artPressureCoeff=0.1,
                 artPressureExp=4, isVorticityConfEnabled=1, vorticityConfCoeff=0.0002,
                 xsphViscosityCoeff=0.01)
        if mapping == "stable":
            p.update(artPressureCoeff=0.001, vorticityConfCoeff=float(np.float32(0.0002 / rho0)),
                     xsphViscosityCoeff=float(np.float32(0.01 / rho0)))
        elif mapping != "literal":
            raise ValueError(f"unknown parameter mapping {mapping!r}")
        return p


PBF_SIZES = {
    # name: (lattice, box, res)
    "pbf-65k": ((64, 32, 32), (4, 2, 2), (40, 20, 20)),
    "pbf-1m": ((128, 128, 64), (8, 8, 4), (80, 80, 40)),
    "pbf-4m": ((256, 128, 128), (16, 8, 8), (160, 80, 80)),
}


def dam_break(lattice, box, res, seed: int = 1, name: str = "pbf", jitter: float = 0.05,
              extra: int = 512) -> FluidScene:
    """Jittered lattice in the (-x,-y,-z) corner; spacing h/2, jitter +-jitter*h."""
    lx, ly, lz = (int(v) for v in lattice)
    h = float(box[0]) / float(res[0])
    s = 0.5 * h
    n = lx * ly * lz
    ix, iy, iz = np.meshgrid(np.arange(lx), np.arange(ly), np.arange(lz), indexing="ij")
    base = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    corner = -0.5 * np.asarray(box, np.float64)
    u = uniform24(seed, 0, 3 * n).reshape(n, 3).astype(np.float64)
    xyz = corner + (base + 0.5) * s + (2.0 * u - 1.0) * (jitter * h)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = xyz.astype(np.float32)
    vel = np.zeros((n, 4), np.float32)
    return FluidScene(name=name, box=tuple(float(b) for b in box), res=tuple(int(r) for r in res),
                      lattice=(lx, ly, lz), n=n, max_n=n + extra, pos=pos, vel=vel,
                      rest_density=lattice_rest_density(h))


def pbf_scene(name: str, seed: int = 1) -> FluidScene:
    lattice, box, res = PBF_SIZES[name]
    return dam_break(lattice, box, res, seed=seed, name=name)


def batch_scenes(count: int = 256, seed: int = 11, n_lo: int = 2048, n_hi: int = 65536,
                 h: float = 0.1) -> list[FluidScene]:
    """Independent dam breaks of uneven sizes for the batched solver.

    Per scene k (draws k*8 .. k*8+3 of the SplitMix64 stream of `seed`):
      target count N = exp(U[ln n_lo, ln n_hi]) (log-uniform),
      box aspect ratios x:z and y:z in U[1, 4], |gravity| in U[5, 15];
    the lattice (spacing h/2) is N^(1/3)-scaled to those ratios (each side at least 4),
    the box is the smallest whole number of h-cells per axis that the lattice fills to at
    most 80 %, and the jitter stream is seed * 1000003 + k.  Every particle is active.
    """
    u = uniform24(seed, 0, 8 * count).reshape(count, 8).astype(np.float64)
    out = []
    for k in range(count):
        target = math.exp(math.log(n_lo) + u[k, 0] * (math.log(n_hi) - math.log(n_lo)))
        ax, ay = 1.0 + 3.0 * u[k, 1], 1.0 + 3.0 * u[k, 2]
        g = 5.0 + 10.0 * u[k, 3]
        lz = max(4, int(round((target / (ax * ay)) ** (1.0 / 3.0))))
        lat = (max(4, int(round(ax * lz))), max(4, int(round(ay * lz))), lz)
        res = tuple(max(4, int(math.ceil(l * 0.5 / 0.8))) for l in lat)
        box = tuple(float(np.float32(r * h)) for r in res)
        sc = dam_break(lat, box, res, seed=seed * 1000003 + k, name=f"batch-{k}", extra=0)
        sc.gravity = float(np.float32(g))
        out.append(sc)
    return out


@dataclasses.dataclass
class BoidsScene:
    name: str
    box: tuple
    res: tuple
    n: int
    max_n: int
    pos: np.ndarray
    vel: np.ndarray
    velocity_scale: float = 2.0
    alignment: float = 1.0
    cohesion: float = 1.0
    separation: float = 1.5
    time_step: float = 0.01


def boids_flock(n: int = 8192, seed: int = 7, box: float = 2.0, res: int = 20,
                extra: int = 512) -> BoidsScene:
    """Positions U[-box/2, box/2)^3; velocity = uniform unit direction x U[0.5, 1.0]."""
    u = uniform24(seed, 0, 6 * n).reshape(n, 6).astype(np.float64)
    half = 0.5 * box
    xyz = -half + u[:, :3] * box
    zc = 2.0 * u[:, 3] - 1.0
    phi = 2.0 * math.pi * u[:, 4]
    rxy = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
    d = np.stack([rxy * np.cos(phi), rxy * np.sin(phi), zc], axis=1)
    speed = 0.5 + 0.5 * u[:, 5]
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    pos[:, :3] = xyz.astype(np.float32)
    vel[:, :3] = (d * speed[:, None]).astype(np.float32)
    return BoidsScene(name=f"boids-{n}", box=(box, box, box), res=(res, res, res), n=n,
                      max_n=n + extra, pos=pos, vel=vel)


@dataclasses.dataclass
class CloudScene:
    """A Clouds model case (physics/ocl/Clouds.cpp:401-500): particles at random positions in
    a box region, velocity 0, with the parameter tree's defaults (Clouds.cpp:63-98)."""
    name: str
    box: tuple
    res: tuple
    n: int
    max_n: int
    pos: np.ndarray
    vel: np.ndarray
    iters: int = 2
    gravity: float = 9.81

    def params(self) -> dict:
        """FluidKernelInputs of initCloudsJson's "Fluids" tree (Clouds.cpp:65-87)."""
        return dict(restDensity=450.0, relaxCFM=600.0, timeStep=float(np.float32(0.010)), dim=3,
                    isArtPressureEnabled=1, artPressureRadius=float(np.float32(0.006)),
                    artPressureCoeff=float(np.float32(0.001)), artPressureExp=4,
                    isVorticityConfEnabled=1, vorticityConfCoeff=float(np.float32(0.0004)),
                    xsphViscosityCoeff=float(np.float32(0.0001)))

    def cloud_params(self) -> dict:
        """CloudKernelInputs after transferJsonInputsToModel (Clouds.cpp:283-303): the "Clouds"
        tree (:88-97), rest density, time step and relaxation taken from the "Fluids" tree, and
        the initial vapour coefficient of Clouds.hpp:45."""
        f = lambda v: float(np.float32(v))  # noqa: E731
        return dict(dim=3, timeStep=f(0.010), restDensity=450.0, groundHeatCoeff=10.0,
                    buoyancyCoeff=f(0.10), gravCoeff=f(0.0005), adiabaticLapseRate=5.0,
                    phaseTransitionRate=f(0.3485), latentHeatCoeff=f(0.07),
                    isTempSmoothingEnabled=1, relaxCFM=600.0, initVaporDensityCoeff=0.75,
                    windCoeff=1.0)


CLOUD_SIZES = {
    # name: (particles, box, res, fill region as fractions of the box: y from the floor)
    # the app's cumulus case: 65,536 particles in the lowest quarter of a 10 x 20 x 10 box
    # (Clouds.cpp:463-468 with ParticleSystemApp.cpp:300-304's doubled y)
    "clouds-65k": (65536, (10.0, 20.0, 10.0), (30, 60, 30), 0.25),
    # the homogeneous case: the whole box
    "clouds-65k-homogeneous": (65536, (10.0, 20.0, 10.0), (30, 60, 30), 1.0),
}


def cloud_scene(name: str = "clouds-65k", seed: int = 17, max_n: int = 131072) -> CloudScene:
    """Positions uniform in x, z over the box and in y over the fill region from the floor
    (the reference draws them with rand(); a SplitMix64 stream here), velocity 0."""
    n, box, res, frac = CLOUD_SIZES[name]
    u = uniform24(seed, 0, 3 * n).reshape(n, 3).astype(np.float32)
    lo = np.array([-box[0] / 2, -box[1] / 2, -box[2] / 2], np.float32)
    hi = np.array([box[0] / 2, -box[1] / 2 + frac * box[1], box[2] / 2], np.float32)
    vec = hi - lo
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = u * vec + lo  # Geometry.cpp:229-237: rand()/RAND_MAX * vec + start, in float
    vel = np.zeros((n, 4), np.float32)
    return CloudScene(name=name, box=box, res=res, n=n, max_n=max_n, pos=pos, vel=vel)
