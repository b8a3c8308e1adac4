"""ctypes binding of the CPU oracle (oracle/rtp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
CPU-baseline leg of ``bench.py``.  The HIP library never routes through this module.

The oracle restates the reference's OpenCL kernels in C; this module owns the numpy
arrays that stand for the reference's OpenCL buffers (physics/ocl/Fluids.cpp:128-150,
physics/ocl/Boids.cpp:122-139) and drives ``orc_fluids_update`` / ``orc_boids_update``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "librtp_oracle.so")

_u32 = ctypes.c_uint32
_f32 = ctypes.c_float
_vp = ctypes.c_void_p


class Consts(ctypes.Structure):
    _fields_ = [
        ("effectRadius", _f32),
        ("effectRadiusSq", _f32),
        ("absWall", _f32 * 3),
        ("res", ctypes.c_int32 * 3),
        ("cellSize", _f32),
        ("numCells", _u32),
        ("maxPartsInCell", _u32),
        ("poly6Coeff", _f32),
        ("spikyCoeff", _f32),
        ("maxVel", _f32),
        ("gravity", _f32),
    ]


class FluidParams(ctypes.Structure):
    """physics/ocl/kernels/define.cl:13-25 (FluidParams, 44 bytes)."""

    _fields_ = [
        ("restDensity", _f32),
        ("relaxCFM", _f32),
        ("timeStep", _f32),
        ("dim", _u32),
        ("isArtPressureEnabled", _u32),
        ("artPressureRadius", _f32),
        ("artPressureCoeff", _f32),
        ("artPressureExp", _u32),
        ("isVorticityConfEnabled", _u32),
        ("vorticityConfCoeff", _f32),
        ("xsphViscosityCoeff", _f32),
    ]


class CloudParams(ctypes.Structure):
    """clouds.cl:18-32 CloudParams == Clouds.hpp:14-48 CloudKernelInputs (52 bytes)."""

    _fields_ = [
        ("dim", _u32),
        ("timeStep", _f32),
        ("restDensity", _f32),
        ("groundHeatCoeff", _f32),
        ("buoyancyCoeff", _f32),
        ("gravCoeff", _f32),
        ("adiabaticLapseRate", _f32),
        ("phaseTransitionRate", _f32),
        ("latentHeatCoeff", _f32),
        ("isTempSmoothingEnabled", _u32),
        ("relaxCFM", _f32),
        ("initVaporDensityCoeff", _f32),
        ("windCoeff", _f32),
    ]


class BoidsParams(ctypes.Structure):
    """physics/ocl/kernels/boids.cl:12-17 (BoidsRuleParams)."""

    _fields_ = [
        ("velocityScale", _f32),
        ("alignmentScale", _f32),
        ("separationScale", _f32),
        ("cohesionScale", _f32),
    ]


class _FluidsState(ctypes.Structure):
    _fields_ = [
        ("c", Consts),
        ("p", FluidParams),
        ("maxN", _u32),
        ("n", _u32),
        ("iters", _u32),
        ("pos", _vp), ("vel", _vp), ("pred", _vp), ("corr", _vp), ("vort", _vp),
        ("velIn", _vp), ("tmp", _vp),
        ("density", _vp), ("lambda_", _vp),
        ("key", _vp), ("idx", _vp),
        ("se", _vp),
    ]


class _CloudsState(ctypes.Structure):
    _fields_ = [("c", Consts), ("p", FluidParams), ("q", CloudParams), ("maxN", _u32), ("n", _u32),
                ("iters", _u32)] + [(k, _vp) for k in (
                    "pos", "vel", "pred", "totCorr", "corr", "vort", "velIn", "tmp", "temp",
                    "tempIn", "buoyancy", "vapor", "vaporIn", "cloud", "cloudIn", "gen", "partID",
                    "density", "lambda_", "lap", "lambdaT", "corrT", "ftmp", "key", "idx", "se")]


class _BoidsState(ctypes.Structure):
    _fields_ = [
        ("c", Consts),
        ("p", BoidsParams),
        ("timeStep", _f32),
        ("maxN", _u32),
        ("n", _u32),
        ("pos", _vp), ("vel", _vp), ("acc", _vp), ("tmp", _vp),
        ("key", _vp), ("idx", _vp),
        ("se", _vp),
        ("dim", ctypes.c_int32), ("cyclic", ctypes.c_int32), ("targetOn", ctypes.c_int32),
        ("target", _f32 * 3), ("targetRadius", _f32), ("targetSign", ctypes.c_int32),
    ]


class _Render(ctypes.Structure):
    _fields_ = [
        ("cameraPos", _f32 * 3),
        ("col", _vp), ("cameraDist", _vp), ("detector", _vp),
    ]


class _RenderBuffers:
    """p_col, p_cameraDist and c_partDetector of one model, reset as at model reset
    (Fluids.cpp:209-215, Boids.cpp:265-272)."""

    def __init__(self, max_n: int, c, camera, col_init):
        self.col = np.empty((max_n, 4), np.float32)
        self.col[:] = col_init
        self.dist = np.zeros(max_n, np.uint32)
        self.det = np.zeros((int(c.numCells), 8), np.float32)
        self.r = _Render()
        self.r.col, self.r.cameraDist, self.r.detector = ptr(self.col), ptr(self.dist), ptr(self.det)
        self.set_camera(camera)
        lib().orc_reset_camera_dist(ptr(self.dist), max_n)

    def set_camera(self, camera) -> None:
        for k in range(3):
            self.r.cameraPos[k] = float(camera[k])


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no fast-math, no FP contraction)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.orc_float_roundtrip.restype = _f32
        L.orc_float_roundtrip.argtypes = [_f32]
        L.orc_make_consts.argtypes = [P(_f32), P(ctypes.c_int32), _u32, _f32, P(Consts)]
        L.orc_pown.restype = _f32
        L.orc_pown.argtypes = [_f32, _u32]
        L.orc_poly6.restype = _f32
        L.orc_poly6.argtypes = [_vp, P(Consts)]
        L.orc_grad_spiky.argtypes = [_vp, P(Consts), _vp]
        L.orc_art_pressure.restype = _f32
        L.orc_art_pressure.argtypes = [_vp, P(FluidParams), P(Consts)]
        L.orc_cell3d.argtypes = [_vp, P(Consts), _vp]
        L.orc_reset_cell_ids.argtypes = [_vp, _u32, P(Consts)]
        L.orc_fill_cell_ids.argtypes = [_vp, _vp, _u32, P(Consts)]
        L.orc_reset_start_end.argtypes = [_vp, P(Consts)]
        L.orc_fill_start.argtypes = [_vp, _vp, _u32, P(Consts)]
        L.orc_fill_end.argtypes = [_vp, _vp, _u32, _u32, P(Consts)]
        L.orc_adjust_end.argtypes = [_vp, P(Consts)]
        L.orc_radix_sort.restype = ctypes.c_int
        L.orc_radix_sort.argtypes = [_vp, _vp, _u32]
        L.orc_permute_f4.argtypes = [_vp, _vp, _vp, _u32]
        for name in ("orc_fld_density",):
            getattr(L, name).argtypes = [_vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_predict.argtypes = [_vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_boundary.argtypes = [_vp, _u32, P(Consts)]
        L.orc_fld_lambda.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_correction.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_correct_pos.argtypes = [_vp, _vp, _u32]
        L.orc_fld_update_vel.argtypes = [_vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_vorticity.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_confinement.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_xsph.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fld_update_pos.argtypes = [_vp, _vp, _u32]
        L.orc_fld_features.argtypes = [_vp, _vp, _vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_fallback_set.argtypes = [_vp, _vp, _vp, _f32, _f32, _vp, _u32, P(Consts)]
        L.orc_bd_rules3d.argtypes = [_vp, _vp, _vp, P(BoidsParams), _vp, _u32, P(Consts)]
        L.orc_bd_update_vel.argtypes = [_vp, _f32, _f32, _vp, _u32]
        L.orc_bd_update_pos_wall.argtypes = [_vp, _f32, _vp, _u32, P(Consts)]
        L.orc_bd_rules2d.argtypes = [_vp, _vp, _vp, P(BoidsParams), _vp, _u32, P(Consts)]
        L.orc_bd_add_target.argtypes = [_vp, P(_f32), _f32, ctypes.c_int32, _vp, _u32]
        L.orc_bd_update_pos_periodic.argtypes = [_vp, _f32, _vp, _u32, P(Consts)]
        L.orc_fluids_update.restype = ctypes.c_int
        L.orc_fluids_update.argtypes = [P(_FluidsState)]
        L.orc_boids_update.restype = ctypes.c_int
        L.orc_boids_update.argtypes = [P(_BoidsState)]
        L.orc_reset_camera_dist.argtypes = [_vp, _u32]
        L.orc_fill_camera_dist.argtypes = [_vp, P(_f32), _vp, _u32]
        L.orc_fld_fill_color.argtypes = [_vp, P(FluidParams), _vp, _u32]
        L.orc_bd_fill_color.argtypes = [_vp, _u32]
        L.orc_reset_detector.argtypes = [_vp, P(Consts)]
        L.orc_fill_detector.argtypes = [_vp, _vp, _u32, P(Consts)]
        L.orc_fluids_frame.restype = ctypes.c_int
        L.orc_fluids_frame.argtypes = [P(_FluidsState), P(_Render), ctypes.c_int]
        L.orc_boids_frame.restype = ctypes.c_int
        L.orc_boids_frame.argtypes = [P(_BoidsState), P(_Render), ctypes.c_int]
        L.orc_exp64.restype = ctypes.c_double
        L.orc_exp64.argtypes = [ctypes.c_double]
        L.orc_exp32.restype = _f32
        L.orc_exp32.argtypes = [_f32]
        L.orc_cld_ext_heat.restype = _f32
        L.orc_cld_ext_heat.argtypes = [_f32, P(Consts)]
        L.orc_cld_env_temp.restype = _f32
        L.orc_cld_env_temp.argtypes = [_f32, P(Consts)]
        L.orc_cld_svd.restype = _f32
        L.orc_cld_svd.argtypes = [_f32]
        L.orc_cld_init_temp.argtypes = [_vp, _vp, _u32, P(Consts)]
        L.orc_cld_init_vapor.argtypes = [P(CloudParams), _vp, _vp, _u32]
        L.orc_cld_heat_ground.argtypes = [_vp, _vp, P(CloudParams), _vp, _u32, P(Consts)]
        L.orc_cld_buoyancy.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _u32, P(Consts)]
        L.orc_cld_adiabatic.argtypes = [_vp, _vp, P(CloudParams), _vp, _u32]
        L.orc_cld_generate.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _u32]
        L.orc_cld_phase.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _vp, _u32]
        L.orc_cld_latent.argtypes = [_vp, _vp, P(CloudParams), _vp, _u32]
        L.orc_cld_predict.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _vp, _u32]
        L.orc_cld_correct_temp.argtypes = [_vp, _vp, _u32]
        L.orc_permute_f1.argtypes = [_vp, _vp, _vp, _u32]
        L.orc_cld_boundary.argtypes = [_vp, _u32, P(Consts)]
        for name in ("orc_cld_density",):
            getattr(L, name).argtypes = [_vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_cld_lambda.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_cld_correction.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_cld_laplacian_temp.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _u32, P(Consts)]
        L.orc_cld_lambda_temp.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _u32, P(Consts)]
        L.orc_cld_correction_temp.argtypes = [_vp, _vp, _vp, P(CloudParams), _vp, _u32, P(Consts)]
        L.orc_cld_vorticity.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_cld_confinement.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_cld_xsph.argtypes = [_vp, _vp, _vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_cld_update_pos.argtypes = [_vp, P(CloudParams), _vp, _u32, P(Consts)]
        L.orc_cld_update_vel.argtypes = [_vp, P(FluidParams), _vp, _u32, P(Consts)]
        L.orc_clouds_update.restype = ctypes.c_int
        L.orc_clouds_update.argtypes = [P(_CloudsState)]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def make_consts(box, res, max_parts_in_cell: int, gravity: float = 9.81) -> Consts:
    c = Consts()
    b = (_f32 * 3)(*[float(x) for x in box])
    r = (ctypes.c_int32 * 3)(*[int(x) for x in res])
    lib().orc_make_consts(b, r, int(max_parts_in_cell), float(gravity), ctypes.byref(c))
    return c


def fluid_params(**kw) -> FluidParams:
    """FluidKernelInputs defaults (physics/ocl/Fluids.hpp:17-32), overridable by keyword."""
    p = FluidParams(restDensity=450.0, relaxCFM=600.0, timeStep=0.010, dim=3,
                    isArtPressureEnabled=1, artPressureRadius=0.006, artPressureCoeff=0.001,
                    artPressureExp=4, isVorticityConfEnabled=1, vorticityConfCoeff=0.0004,
                    xsphViscosityCoeff=0.0001)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


class OracleFluids:
    """Fluids::update (physics/ocl/Fluids.cpp:400-471) on numpy-owned buffers."""

    def __init__(self, box, res, max_n: int, n: int, params: FluidParams, iters: int,
                 gravity: float = 9.81, max_parts_in_cell: int = 100):
        self.c = make_consts(box, res, max_parts_in_cell, gravity)
        nc = int(self.c.numCells)
        self.max_n, self.n = int(max_n), int(n)
        f4 = lambda: np.zeros((self.max_n, 4), np.float32)  # noqa: E731
        self.pos, self.vel, self.pred, self.corr = f4(), f4(), f4(), f4()
        self.vort, self.vel_in, self.tmp = f4(), f4(), f4()
        self.density = np.zeros(self.max_n, np.float32)
        self.lam = np.zeros(self.max_n, np.float32)
        self.key = np.zeros(self.max_n, np.uint32)
        self.idx = np.zeros(self.max_n, np.uint32)
        self.se = np.zeros((nc, 2), np.uint32)
        self.render = None
        self.paused = False
        self.s = _FluidsState()
        self.s.c = self.c
        self.s.maxN, self.s.n, self.s.iters = self.max_n, self.n, int(iters)
        self.set_params(params)
        for name, arr in (("pos", self.pos), ("vel", self.vel), ("pred", self.pred),
                          ("corr", self.corr), ("vort", self.vort), ("velIn", self.vel_in),
                          ("tmp", self.tmp), ("density", self.density), ("lambda_", self.lam),
                          ("key", self.key), ("idx", self.idx), ("se", self.se)):
            setattr(self.s, name, ptr(arr))

    def set_params(self, params: FluidParams) -> None:
        self.p = params
        self.s.p = params

    def set_iters(self, iters: int) -> None:
        self.s.iters = int(iters)

    def upload(self, pos: np.ndarray, vel: np.ndarray) -> None:
        """Fluids::initFluidsParticles (Fluids.cpp:383-395) + resetCellIDs (:214)."""
        m = pos.shape[0]
        self.pos[:] = (np.inf, np.inf, np.inf, 0.0)
        self.vel[:] = 0.0
        self.pos[:m] = pos
        self.vel[:m] = vel
        self.pred[:] = self.pos
        lib().orc_reset_cell_ids(ptr(self.key), self.max_n, ctypes.byref(self.c))

    def enable_render(self, camera) -> None:
        """The rendering steps of every frame (Fluids.cpp:458-468), with p_col at its
        initial colour (Fluids.cpp:394) and the camera distances reset (:215)."""
        self.render = _RenderBuffers(self.max_n, self.c, camera, (0.0, 0.1, 1.0, 0.0))

    def step(self, nsteps: int = 1) -> None:
        for _ in range(nsteps):
            if self.render is not None:
                rc = lib().orc_fluids_frame(ctypes.byref(self.s), ctypes.byref(self.render.r),
                                            int(self.paused))
            elif self.paused:
                continue
            else:
                rc = lib().orc_fluids_update(ctypes.byref(self.s))
            if rc:
                raise RuntimeError(f"orc_fluids_update failed ({rc})")


    def learned_step(self, network, learned_iters: int, fallback_iters: int = 0,
                     tau: float = 0.0) -> np.ndarray:
        """One learned substep (no reference counterpart), composed of the reference's
        kernels in the order of Fluids::update (physics/ocl/Fluids.cpp:400-471):

        - predict, cell keys, sort, permutation, cell ranges (:412-424);
        - learned_iters Jacobi iterations (:427-439) whose first takes its lambda from
          network(features) -- the 8 inputs of the first density pass (orc_fld_features);
        - with fallback_iters > 0: the post-correction density, the fallback set and its
          halo (orc_fallback_set at tau) and fallback_iters Jacobi iterations that recompute
          the density and lambda of the set and the halo and correct only the set's
          particles, with every other particle's lambda 0 (its density is kept; an empty set
          changes nothing);
        - the velocity update, vorticity confinement, XSPH and commit (:442-457).

        Returns the fallback set's flags: 2 seed, 1 other member, 3 halo, 0 outside (all 0
        without fallback)."""
        L, c, p = lib(), ctypes.byref(self.c), ctypes.byref(self.p)
        n, max_n = self.n, self.max_n
        L.orc_fld_predict(ptr(self.pos), ptr(self.vel), p, ptr(self.pred), n, c)
        L.orc_fill_cell_ids(ptr(self.pred), ptr(self.key), n, c)
        if L.orc_radix_sort(ptr(self.key), ptr(self.idx), max_n):
            raise RuntimeError("orc_radix_sort failed")
        for buf in (self.pos, self.vel, self.pred):
            L.orc_permute_f4(ptr(self.idx), ptr(buf), ptr(self.tmp), max_n)
        L.orc_reset_start_end(ptr(self.se), c)
        L.orc_fill_start(ptr(self.key), ptr(self.se), n, c)
        L.orc_fill_end(ptr(self.key), ptr(self.se), n, max_n, c)
        L.orc_adjust_end(ptr(self.se), c)
        for it in range(learned_iters):
            L.orc_fld_boundary(ptr(self.pred), n, c)
            L.orc_fld_density(ptr(self.pred), ptr(self.se), p, ptr(self.density), n, c)
            L.orc_fld_lambda(ptr(self.pred), ptr(self.density), ptr(self.se), p, ptr(self.lam),
                             n, c)
            if it == 0:
                feat = np.zeros((n, 8), np.float32)
                L.orc_fld_features(ptr(self.pred), ptr(self.density), ptr(self.se),
                                   ptr(self.lam), ptr(self.vel), p, ptr(feat), n, c)
                self.lam[:n] = np.asarray(network(feat), np.float32)
            L.orc_fld_correction(ptr(self.lam), ptr(self.se), ptr(self.pred), p, ptr(self.corr),
                                 n, c)
            L.orc_fld_correct_pos(ptr(self.corr), ptr(self.pred), n)
        act = np.zeros(n, np.uint8)
        if fallback_iters > 0:
            res = np.zeros(max_n, np.float32)
            L.orc_fld_density(ptr(self.pred), ptr(self.se), p, ptr(res), n, c)
            L.orc_fallback_set(ptr(self.pred), ptr(self.se), ptr(res), float(self.p.restDensity),
                               float(tau), ptr(act), n, c)
            on = (act == 1) | (act == 2)
            known = act > 0  # the set and its halo
            dens = np.zeros(max_n, np.float32)
            lam = np.zeros(max_n, np.float32)
            for _ in range(fallback_iters if on.any() else 0):
                L.orc_fld_boundary(ptr(self.pred), n, c)
                L.orc_fld_density(ptr(self.pred), ptr(self.se), p, ptr(dens), n, c)
                L.orc_fld_lambda(ptr(self.pred), ptr(dens), ptr(self.se), p, ptr(lam), n, c)
                self.density[:n][known] = dens[:n][known]
                self.lam[:n] = np.where(known, lam[:n], np.float32(0.0))
                L.orc_fld_correction(ptr(self.lam), ptr(self.se), ptr(self.pred), p,
                                     ptr(self.corr), n, c)
                moved = self.pred.copy()
                L.orc_fld_correct_pos(ptr(self.corr), ptr(moved), n)
                self.pred[:n][on] = moved[:n][on]
        L.orc_fld_update_vel(ptr(self.pred), ptr(self.pos), p, ptr(self.vel), n, c)
        if self.p.isVorticityConfEnabled:
            L.orc_fld_vorticity(ptr(self.pred), ptr(self.se), ptr(self.vel), p, ptr(self.vort),
                                n, c)
            L.orc_fld_confinement(ptr(self.pred), ptr(self.se), ptr(self.vort), p, ptr(self.vel),
                                  n, c)
            self.vel_in[:] = self.vel
            L.orc_fld_xsph(ptr(self.pred), ptr(self.se), ptr(self.vel_in), p, ptr(self.vel), n, c)
        L.orc_fld_update_pos(ptr(self.pred), ptr(self.pos), n)
        return act


def cloud_params(**kw) -> CloudParams:
    """CloudKernelInputs defaults (physics/ocl/Clouds.hpp:14-48), overridable by keyword."""
    q = CloudParams(dim=3, timeStep=0.01, restDensity=400.0, groundHeatCoeff=10.0,
                    buoyancyCoeff=0.10, gravCoeff=0.0005, adiabaticLapseRate=5.0,
                    phaseTransitionRate=0.3485, latentHeatCoeff=0.07, isTempSmoothingEnabled=1,
                    relaxCFM=600.0, initVaporDensityCoeff=0.75, windCoeff=1.0)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


class OracleClouds:
    """Clouds::update (physics/ocl/Clouds.cpp:503-610, physics part) on numpy-owned buffers."""

    def __init__(self, box, res, max_n: int, n: int, params: FluidParams, cloud: CloudParams,
                 iters: int, gravity: float = 9.81, max_parts_in_cell: int = 100):
        self.c = make_consts(box, res, max_parts_in_cell, gravity)
        nc = int(self.c.numCells)
        self.max_n, self.n = int(max_n), int(n)
        f4 = lambda: np.zeros((self.max_n, 4), np.float32)  # noqa: E731
        f1 = lambda: np.zeros(self.max_n, np.float32)  # noqa: E731
        self.bufs = {}
        for k in ("pos", "vel", "pred", "totCorr", "corr", "vort", "velIn", "tmp"):
            self.bufs[k] = f4()
        for k in ("temp", "tempIn", "buoyancy", "vapor", "vaporIn", "cloud", "cloudIn", "gen",
                  "partID", "density", "lambda_", "lap", "lambdaT", "corrT", "ftmp"):
            self.bufs[k] = f1()
        self.bufs["key"] = np.zeros(self.max_n, np.uint32)
        self.bufs["idx"] = np.zeros(self.max_n, np.uint32)
        self.bufs["se"] = np.zeros((nc, 2), np.uint32)
        self.s = _CloudsState()
        self.s.c = self.c
        self.s.maxN, self.s.n, self.s.iters = self.max_n, self.n, int(iters)
        self.set_params(params, cloud)
        for k, arr in self.bufs.items():
            setattr(self.s, k, ptr(arr))

    def __getattr__(self, name):
        bufs = self.__dict__.get("bufs")
        if bufs is not None and name in bufs:
            return bufs[name]
        if bufs is not None and name == "lam":
            return bufs["lambda_"]
        raise AttributeError(name)

    def set_params(self, params: FluidParams, cloud: CloudParams) -> None:
        self.p, self.q = params, cloud
        self.s.p, self.s.q = params, cloud

    def set_iters(self, iters: int) -> None:
        self.s.iters = int(iters)

    def upload(self, pos: np.ndarray, vel: np.ndarray) -> None:
        """Clouds::initCloudsParticles (Clouds.cpp:401-500) from given positions and velocities:
        the tail at +inf, cloud density 0, particle ids, then cld_initTemperature and
        cld_initVaporDensity over maxN, and resetCellIDs (:396)."""
        m = pos.shape[0]
        b = self.bufs
        b["pos"][:] = (np.inf, np.inf, np.inf, 0.0)
        b["vel"][:] = 0.0
        b["pos"][:m] = pos
        b["vel"][:m] = vel
        b["cloud"][:] = 0.0
        b["partID"][:] = np.arange(self.max_n, dtype=np.float32)
        L = lib()
        L.orc_cld_init_temp(ptr(b["pos"]), ptr(b["temp"]), self.max_n, ctypes.byref(self.c))
        L.orc_cld_init_vapor(ctypes.byref(self.q), ptr(b["temp"]), ptr(b["vapor"]), self.max_n)
        L.orc_reset_cell_ids(ptr(b["key"]), self.max_n, ctypes.byref(self.c))

    def step(self, nsteps: int = 1) -> None:
        for _ in range(nsteps):
            rc = lib().orc_clouds_update(ctypes.byref(self.s))
            if rc:
                raise RuntimeError(f"orc_clouds_update failed ({rc})")


class OracleBoids:
    """Boids::update (physics/ocl/Boids.cpp:323-384) on numpy-owned buffers."""

    def __init__(self, box, res, max_n: int, n: int, params: BoidsParams, time_step: float = 0.1,
                 max_parts_in_cell: int = 3000):
        self.c = make_consts(box, res, max_parts_in_cell, 9.81)
        nc = int(self.c.numCells)
        self.max_n, self.n = int(max_n), int(n)
        f4 = lambda: np.zeros((self.max_n, 4), np.float32)  # noqa: E731
        self.pos, self.vel, self.acc, self.tmp = f4(), f4(), f4(), f4()
        self.key = np.zeros(self.max_n, np.uint32)
        self.idx = np.zeros(self.max_n, np.uint32)
        self.se = np.zeros((nc, 2), np.uint32)
        self.render = None
        self.paused = False
        self.s = _BoidsState()
        self.s.c = self.c
        self.s.p = params
        self.s.timeStep = float(time_step)
        self.s.maxN, self.s.n = self.max_n, self.n
        self.s.dim = 3
        for name, arr in (("pos", self.pos), ("vel", self.vel), ("acc", self.acc),
                          ("tmp", self.tmp), ("key", self.key), ("idx", self.idx), ("se", self.se)):
            setattr(self.s, name, ptr(arr))

    def upload(self, pos: np.ndarray, vel: np.ndarray) -> None:
        """Boids::initBoidsParticles (Boids.cpp:310-318) + resetCellIDs (:271)."""
        m = pos.shape[0]
        self.pos[:] = (np.inf, np.inf, np.inf, 0.0)
        self.vel[:] = (np.inf, np.inf, np.inf, 0.0)
        self.acc[:] = 0.0
        self.pos[:m] = pos
        self.vel[:m] = vel
        lib().orc_reset_cell_ids(ptr(self.key), self.max_n, ctypes.byref(self.c))

    def set_dimension(self, dim: int) -> None:
        self.s.dim = int(dim)

    def set_cyclic(self, flag: bool) -> None:
        self.s.cyclic = int(bool(flag))

    def set_target(self, active: bool, pos=(0.0, 0.0, 0.0), radius: float = 2.0,
                   sign: int = 1) -> None:
        """Target::pos / radiusEffect / signEffect of the next frames (Target.hpp)."""
        self.s.targetOn = int(bool(active))
        for k in range(3):
            self.s.target[k] = float(pos[k])
        self.s.targetRadius = float(radius)
        self.s.targetSign = int(sign)

    def enable_render(self, camera) -> None:
        """The rendering steps of every frame (Boids.cpp:375-381), with p_col filled by
        bd_fillBoidsColor and the camera distances reset (Boids.cpp:267-272)."""
        self.render = _RenderBuffers(self.max_n, self.c, camera, (1.0, 0.02, 0.02, 0.5))

    def step(self, nsteps: int = 1) -> None:
        for _ in range(nsteps):
            if self.render is not None:
                rc = lib().orc_boids_frame(ctypes.byref(self.s), ctypes.byref(self.render.r),
                                           int(self.paused))
            elif self.paused:
                continue
            else:
                rc = lib().orc_boids_update(ctypes.byref(self.s))
            if rc:
                raise RuntimeError(f"orc_boids_update failed ({rc})")
