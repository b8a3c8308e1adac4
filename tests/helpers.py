"""Shared helpers for the GPU parity tests: build matched GPU / oracle runs from one scene."""
from __future__ import annotations

import sys

import numpy as np

from oracle import oracle as O
from realtimeparticles_amd import capi, scenes


def log(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def make_pbf_pair(sc: scenes.FluidScene, iters: int = 4, gravity: float | None = None,
                  max_parts_in_cell: int = 100, **over):
    """A GPU Fluids model and an oracle run, both loaded with the scene's state."""
    g = sc.gravity if gravity is None else gravity
    prm = sc.params()
    prm.update(over)
    gpu = capi.Fluids(sc.max_n, sc.n, sc.box, sc.res, gravity=g,
                      max_parts_in_cell=max_parts_in_cell)
    gpu.set_params(**prm)
    gpu.set_nb_jacobi_iters(iters)
    gpu.load_state(sc.pos, sc.vel)
    orc = O.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**prm), iters,
                         gravity=g, max_parts_in_cell=max_parts_in_cell)
    orc.upload(sc.pos, sc.vel)
    return gpu, orc


def compare_pbf(gpu: capi.Fluids, orc: O.OracleFluids, n: int, step: int, grid: bool = True):
    """Assert bit-equality (as values) of the state both runs leave after an update."""
    pos, vel = gpu.read_state(n)
    where = f"step {step}"
    _eq(pos[:, :3], orc.pos[:n, :3], f"pos {where}")
    _eq(vel[:, :3], orc.vel[:n, :3], f"vel {where}")
    _eq(gpu.read_field(capi.FIELD_DENSITY, n), orc.density[:n], f"density {where}")
    _eq(gpu.read_field(capi.FIELD_LAMBDA, n), orc.lam[:n], f"lambda {where}")
    _eq(gpu.read_field(capi.FIELD_PRED_POS, n)[:, :3], orc.pred[:n, :3], f"predPos {where}")
    if grid:
        keys, idx, se = gpu.read_grid()
        _eq(keys, orc.key, f"cellIDs {where}")
        _eq(idx, orc.idx, f"sort index {where}")
        _eq(se, orc.se, f"startEnd {where}")


def _eq(a: np.ndarray, b: np.ndarray, what: str) -> None:
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if np.array_equal(a, b, equal_nan=a.dtype.kind == "f"):
        return
    same = ((a == b) | (np.isnan(a) & np.isnan(b))) if a.dtype.kind == "f" else (a == b)
    bad = np.argwhere(~same)
    if a.dtype.kind == "f":
        diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
        rel = diff / np.maximum(np.abs(b.astype(np.float64)), 1e-30)
        first = bad[0] if len(bad) else None
        raise AssertionError(f"{what}: {len(bad)} mismatches, max abs {diff.max():.3e}, "
                             f"max rel {np.nanmax(rel):.3e}, first at {first}: "
                             f"{a[tuple(first)] if first is not None else ''} vs "
                             f"{b[tuple(first)] if first is not None else ''}")
    raise AssertionError(f"{what}: {len(bad)} mismatches, first at {bad[0]}: "
                         f"{a[tuple(bad[0])]} vs {b[tuple(bad[0])]}")


def oracle_grid(pos: np.ndarray, max_n: int, box, res, cap: int):
    """Reference grid of one batch of positions: resetCellIDs, fillCellIDs, sort, start/end."""
    import ctypes
    L = O.lib()
    c = O.make_consts(box, res, cap, 9.81)
    n = pos.shape[0]
    buf = np.zeros((max_n, 4), np.float32)
    buf[:n] = pos
    key = np.zeros(max_n, np.uint32)
    idx = np.zeros(max_n, np.uint32)
    L.orc_reset_cell_ids(key.ctypes.data, max_n, ctypes.byref(c))
    L.orc_fill_cell_ids(buf.ctypes.data, key.ctypes.data, n, ctypes.byref(c))
    if max_n % 512 == 0:
        assert L.orc_radix_sort(key.ctypes.data, idx.ctypes.data, max_n) == 0
    else:  # the reference requires max_n % 512 == 0; its result is the stable sort
        order = np.argsort(key, kind="stable").astype(np.uint32)
        key[:] = key[order]
        idx[:] = order
    se = np.zeros((int(c.numCells), 2), np.uint32)
    L.orc_reset_start_end(se.ctypes.data, ctypes.byref(c))
    L.orc_fill_start(key.ctypes.data, se.ctypes.data, n, ctypes.byref(c))
    L.orc_fill_end(key.ctypes.data, se.ctypes.data, n, max_n, ctypes.byref(c))
    L.orc_adjust_end(se.ctypes.data, ctypes.byref(c))
    return key, idx, se


def oracle_teacher(sc, prm, iters, return_grid=False):
    """Features and summed lambda of one teacher substep, restated step by step on the CPU
    (and the substep's grid: keys, sort index, start/end, with return_grid)."""
    import ctypes
    L = O.lib()
    f = O.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**prm), iters, gravity=sc.gravity)
    f.upload(sc.pos, sc.vel)
    c, p, n, pt = ctypes.byref(f.c), ctypes.byref(f.p), sc.n, O.ptr
    L.orc_fld_predict(pt(f.pos), pt(f.vel), p, pt(f.pred), n, c)
    L.orc_fill_cell_ids(pt(f.pred), pt(f.key), n, c)
    assert L.orc_radix_sort(pt(f.key), pt(f.idx), f.max_n) == 0
    for b in (f.pos, f.vel, f.pred):
        L.orc_permute_f4(pt(f.idx), pt(b), pt(f.tmp), f.max_n)
    L.orc_reset_start_end(pt(f.se), c)
    L.orc_fill_start(pt(f.key), pt(f.se), n, c)
    L.orc_fill_end(pt(f.key), pt(f.se), n, f.max_n, c)
    L.orc_adjust_end(pt(f.se), c)
    feat = np.zeros((n, 8), np.float32)
    acc = None
    for k in range(iters):
        L.orc_fld_boundary(pt(f.pred), n, c)
        L.orc_fld_density(pt(f.pred), pt(f.se), p, pt(f.density), n, c)
        L.orc_fld_lambda(pt(f.pred), pt(f.density), pt(f.se), p, pt(f.lam), n, c)
        if k == 0:
            L.orc_fld_features(pt(f.pred), pt(f.density), pt(f.se), pt(f.lam), pt(f.vel), p,
                               pt(feat), n, c)
            acc = f.lam[:n].copy()
        else:
            acc = (acc + f.lam[:n]).astype(np.float32)
        L.orc_fld_correction(pt(f.lam), pt(f.se), pt(f.pred), p, pt(f.corr), n, c)
        L.orc_fld_correct_pos(pt(f.corr), pt(f.pred), n)
    if return_grid:
        return feat, acc, (f.key.copy(), f.idx.copy(), f.se.copy())
    return feat, acc


def compare_fluids_render(gpu, orc, what):
    """Every buffer of a fluids model with the rendering steps on, over maxN."""
    m = orc.max_n
    pos, vel = gpu.read_state(m)
    _eq(pos, orc.pos, f"{what}: p_pos")
    _eq(vel, orc.vel, f"{what}: p_vel")
    _eq(gpu.read_field(capi.FIELD_PRED_POS, m), orc.pred, f"{what}: p_predPos")
    _eq(gpu.read_color(), orc.render.col, f"{what}: p_col")
    _eq(gpu.read_camera_dist(), orc.render.dist, f"{what}: p_cameraDist")
    keys, idx, se = gpu.read_grid()
    _eq(keys, orc.key, f"{what}: p_cellID")
    _eq(idx, orc.idx, f"{what}: RadixSortIndices")
    _eq(se, orc.se, f"{what}: c_startEndPartID")
    _eq(gpu.read_detector(), orc.render.det, f"{what}: c_partDetector")
    n = gpu.n
    _eq(gpu.read_field(capi.FIELD_DENSITY, n), orc.density[:n], f"{what}: p_density")
