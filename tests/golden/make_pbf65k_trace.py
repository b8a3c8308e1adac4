"""Writes pbf65k_trace.json: the oracle's run of the 65,536-particle dam break.

    python tests/golden/make_pbf65k_trace.py [--mapping literal|stable ...]

The 64x32x32 jittered lattice in the 4x2x2 box (grid 40x20x20, h = 0.1, 4 Jacobi
iterations, substep dt = 1/120, gravity 9.8) run by the CPU oracle for 300 frames (600
substeps), the reference's Fluids::update order (physics/ocl/Fluids.cpp:400-471).  After
every substep it records max, mean and max-compression of |rho/rho0 - 1| over the active
particles, from the density the last Jacobi iteration leaves (p_density), reduced in
float64; every 30 substeps the SHA-256 of the positions and velocities (float32 x 4 per
particle).  The GPU test compares its own run with both.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "pbf65k_trace.json")
SUBSTEPS = 600
EVERY = 30


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def density_error(density: np.ndarray, rho0: float) -> list[float]:
    rel = density.astype(np.float64) / np.float64(np.float32(rho0)) - 1.0
    return [float(np.abs(rel).max()), float(np.abs(rel).mean()), float(max(rel.max(), 0.0))]


def run(mapping: str) -> dict:
    from oracle import oracle as O
    from realtimeparticles_amd import scenes
    O.build()
    sc = scenes.pbf_scene("pbf-65k")
    prm = sc.params(mapping)
    orc = O.OracleFluids(sc.box, sc.res, sc.max_n, sc.n, O.fluid_params(**prm), sc.iters,
                         gravity=sc.gravity)
    orc.upload(sc.pos, sc.vel)
    errs, hashes = [], []
    t0 = time.time()
    for step in range(1, SUBSTEPS + 1):
        orc.step(1)
        errs.append(density_error(orc.density[: sc.n], prm["restDensity"]))
        if step % EVERY == 0:
            hashes.append({"substep": step, "pos": digest(orc.pos[: sc.n]),
                           "vel": digest(orc.vel[: sc.n])})
            print(f"{mapping}: substep {step}, max |rho/rho0 - 1| {errs[-1][0]:.4f}, "
                  f"{time.time() - t0:.0f} s", flush=True)
    return {"params": prm, "density_error": errs, "state_sha256": hashes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mapping", nargs="+", default=["literal", "stable"])
    args = ap.parse_args()
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    data.setdefault("scene", "pbf-65k: 64x32x32 lattice, box 4x2x2, grid 40x20x20, seed 1")
    data.setdefault("substeps", SUBSTEPS)
    data.setdefault("columns", ["max", "mean", "max_compression"])
    for m in args.mapping:
        data[m] = run(m)
        with open(OUT, "w") as f:
            json.dump(data, f, separators=(",", ":"))
            f.write("\n")


if __name__ == "__main__":
    main()
