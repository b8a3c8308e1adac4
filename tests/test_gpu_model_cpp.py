"""The C++ Physics::Model path (CreateModel -> reset -> update), driven like the reference
app with its own defaults (box 10^3, grid 30^3, 131,072 particles, FluidKernelInputs and
boids JSON defaults), compared for exact equality with the oracle replaying the same steps."""
import os
import subprocess

import numpy as np
import pytest

from helpers import _eq, log
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "model_demo")


def run_demo(kind, tmp_path, steps):
    out = tmp_path / f"{kind}.bin"
    r = subprocess.run([DEMO, kind, str(out), str(steps)], capture_output=True, text=True,
                       timeout=120)
    log(r.stdout + r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(out, np.uint8)
    n, st, max_n = np.frombuffer(raw[:12].tobytes(), np.uint32)
    arr = np.frombuffer(raw[12:].tobytes(), np.float32).reshape(4, int(n), 4)
    return int(n), int(max_n), arr


def test_cpp_fluids_dam_matches_oracle(tmp_path):
    n, max_n, (pos0, vel0, pos1, vel1) = run_demo("fluids", tmp_path, 3)
    assert n == 131072 and max_n == 131072
    orc = O.OracleFluids((10, 10, 10), (30, 30, 30), max_n, n, O.fluid_params(), 2, gravity=9.81)
    orc.upload(pos0, vel0)
    orc.step(3)
    _eq(pos1[:, :3], orc.pos[:n, :3], "pos")
    _eq(vel1[:, :3], orc.vel[:n, :3], "vel")


def test_cpp_boids_matches_oracle(tmp_path):
    n, max_n, (pos0, vel0, pos1, vel1) = run_demo("boids", tmp_path, 3)
    assert n == 65536
    # initBoidsParticles loads the positions into the velocity buffer too
    _eq(vel0, pos0, "initial velocity")
    orc = O.OracleBoids((10, 10, 10), (30, 30, 30), max_n, n,
                        O.BoidsParams(0.5, 1.6, 1.6, 1.45), time_step=0.1)
    orc.upload(pos0, vel0)
    orc.step(3)
    _eq(pos1[:, :3], orc.pos[:n, :3], "pos")
    _eq(vel1[:, :3], orc.vel[:n, :3], "vel")
