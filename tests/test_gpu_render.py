"""GPU parity of the rendering side of a frame (Fluids.cpp:458-468, Boids.cpp:375-381).

With the rendering steps on, every update ends with the grid detector, the fluid colour and
the camera-distance sort, which permutes the particle buffers and so changes the next
update's within-cell order.  Each case runs the oracle's whole frame beside the GPU and
compares, after every update, every buffer over maxN exactly: positions, velocities,
p_predPos / p_acc, p_col, p_cameraDist, the cell keys, the sort permutation (the camera
sort's, as RadixSortIndices holds after the reference's frame), the cell table and the
detector.  The runs cover a camera that moves, paused updates (which still sort), a particle
count that shrinks (stale keys in the sorted range) and pairs of substeps replayed from a
hipGraph.
"""
import numpy as np
import pytest

from helpers import _eq, compare_fluids_render as compare_fluids, log, make_pbf_pair
from oracle import oracle as O
from realtimeparticles_amd import capi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("flags", [0, capi.LAUNCH_RESORT | capi.LAUNCH_XCD_MAP | capi.LAUNCH_NB_LIST])
def test_fluids_render_frames_bit_exact(flags):
    sc = scenes.pbf_scene("pbf-65k")
    gpu, orc = make_pbf_pair(sc, iters=sc.iters, **sc.params("stable"))
    gpu.set_tuning(64 if flags else 256, flags, 0, 16)
    cam = (6.0, 3.0, -5.0)
    gpu.set_render(True)
    gpu.set_camera(cam)
    orc.enable_render(cam)
    log(f"render fluids (flags {flags}): {sc.n} particles, camera {cam}")
    _eq(gpu.read_detector(), _reset_detector(orc), "detector at reset")
    for f in range(1, 4):
        gpu.update(1)
        orc.step(1)
        compare_fluids(gpu, orc, f"frame {f}")
    # pairs of substeps from the graph, with the camera moving between updates
    cam = (-1.0, 0.5, 0.25)  # inside the box: keys near FAR_DIST
    gpu.set_camera(cam)
    orc.render.set_camera(cam)
    gpu.update(2)
    orc.step(2)
    compare_fluids(gpu, orc, "frames 4-5 (graph)")
    # paused: the camera moves, the sort runs, the physics does not
    gpu.pause(True)
    orc.paused = True
    cam = (0.0, -4.0, 9.0)
    gpu.set_camera(cam)
    orc.render.set_camera(cam)
    gpu.update(1)
    orc.step(1)
    compare_fluids(gpu, orc, "paused frame 6")
    gpu.update(3)
    orc.step(3)
    compare_fluids(gpu, orc, "paused frames 7-9")
    gpu.pause(False)
    orc.paused = False
    # fewer particles: keys of the dropped ones stay in both sorts' ranges
    m = sc.n - 1000
    gpu.set_nb_particles(m)
    orc.s.n = m
    for f in (10, 11):
        gpu.update(1)
        orc.step(1)
        compare_fluids(gpu, orc, f"frame {f} (n = {m})")
    gpu.update(2)
    orc.step(2)
    compare_fluids(gpu, orc, f"frames 12-13 (n = {m}, graph)")
    gpu.close()


def _reset_detector(orc):
    import ctypes
    det = np.zeros_like(orc.render.det)
    O.lib().orc_fill_detector(orc.pos.ctypes.data, det.ctypes.data, orc.s.n, ctypes.byref(orc.c))
    return det


def test_fluids_render_off_leaves_physics_unchanged():
    """Switching the rendering steps off again returns to the physics-only frame."""
    sc = scenes.pbf_scene("pbf-65k")
    gpu, orc = make_pbf_pair(sc, iters=2)
    gpu.set_render(True)
    gpu.set_render(False)
    gpu.update(3)
    orc.step(3)
    pos, vel = gpu.read_state(orc.max_n)
    _eq(pos, orc.pos, "p_pos")
    _eq(vel, orc.vel, "p_vel")
    gpu.close()


def compare_boids(gpu, orc, what):
    m = orc.max_n
    pos, vel = gpu.read_state(m)
    _eq(pos, orc.pos, f"{what}: p_pos")
    _eq(vel, orc.vel, f"{what}: p_vel")
    _eq(gpu.read_acc(m), orc.acc, f"{what}: p_acc")
    _eq(gpu.read_color(), orc.render.col, f"{what}: p_col")
    _eq(gpu.read_camera_dist(), orc.render.dist, f"{what}: p_cameraDist")
    keys, idx, se = gpu.read_grid()
    _eq(keys, orc.key, f"{what}: p_cellID")
    _eq(idx, orc.idx, f"{what}: RadixSortIndices")
    _eq(se, orc.se, f"{what}: c_startEndPartID")
    _eq(gpu.read_detector(), orc.render.det, f"{what}: c_partDetector")


def test_boids_render_frames_bit_exact():
    sc = scenes.boids_flock(8192, seed=12)
    gpu = capi.Boids(sc.max_n, sc.n, sc.box, sc.res)
    gpu.set_params(sc.velocity_scale, sc.alignment, sc.separation, sc.cohesion)
    gpu.set_time_step(sc.time_step)
    gpu.load_state(sc.pos, sc.vel)
    p = O.BoidsParams(sc.velocity_scale, sc.alignment, sc.separation, sc.cohesion)
    orc = O.OracleBoids(sc.box, sc.res, sc.max_n, sc.n, p, time_step=sc.time_step)
    orc.upload(sc.pos, sc.vel)
    cam = (0.3, 2.5, 1.5)
    gpu.set_render(True)
    gpu.set_camera(cam)
    orc.enable_render(cam)
    log(f"render boids: {sc.n} boids, camera {cam}")
    for f in range(1, 31):
        gpu.update(1)
        orc.step(1)
        if f % 10 == 0 or f < 3:
            compare_boids(gpu, orc, f"frame {f}")
    gpu.pause(True)
    orc.paused = True
    cam = (-3.0, 0.0, 0.0)
    gpu.set_camera(cam)
    orc.render.set_camera(cam)
    gpu.update(2)
    orc.step(2)
    compare_boids(gpu, orc, "paused frames 31-32")
    gpu.pause(False)
    orc.paused = False
    gpu.set_nb_particles(6000)
    orc.s.n = 6000
    for f in range(33, 38):
        gpu.update(1)
        orc.step(1)
    compare_boids(gpu, orc, "frame 37 (n = 6000)")
    gpu.close()


def test_fluids_render_full_and_empty():
    """Every slot active (n = maxN: no FAR_DIST tail) and no particle at all."""
    sc = scenes.dam_break((16, 16, 16), (4, 2, 2), (40, 20, 20), seed=14, extra=0)
    gpu, orc = make_pbf_pair(sc, iters=2, **sc.params("stable"))
    cam = (3.0, 2.0, 1.0)
    gpu.set_render(True)
    gpu.set_camera(cam)
    orc.enable_render(cam)
    for f in range(1, 4):
        gpu.update(1)
        orc.step(1)
        compare_fluids(gpu, orc, f"full frame {f}")
    gpu.set_nb_particles(0)
    orc.s.n = 0
    gpu.update(2)
    orc.step(2)
    compare_fluids(gpu, orc, "empty frames")
    gpu.close()


def test_render_and_variant_argument_errors():
    """Bad arguments of the rendering and boids-variant entry points fail with an error and
    leave the handle usable."""
    sc = scenes.boids_flock(1024, seed=2)
    b = capi.Boids(sc.max_n, sc.n, sc.box, sc.res)
    b.load_state(sc.pos, sc.vel)
    with pytest.raises(capi.RtpError, match="enable the rendering steps"):
        b.set_camera((0.0, 0.0, 0.0))
    with pytest.raises(capi.RtpError, match="never enabled"):
        b.read_color()
    with pytest.raises(capi.RtpError, match="dimension"):
        b.set_dimension(4)
    with pytest.raises(capi.RtpError, match="boundary"):
        capi._check(capi.lib().rtp_boids_set_boundary(b._h, 7), "rtp_boids_set_boundary")
    with pytest.raises(capi.RtpError, match="target"):
        capi._check(capi.lib().rtp_boids_set_target(b._h, 1, None, 1.0, 1), "rtp_boids_set_target")
    with pytest.raises(capi.RtpError, match="flags"):
        b.set_tuning(64, capi.LAUNCH_RESORT)
    with pytest.raises(capi.RtpError, match="blockSize"):
        b.set_tuning(96, 0)
    b.update(2)  # still steps
    f = capi.Fluids(4096, 4096, (4, 2, 2), (40, 20, 20))
    with pytest.raises(capi.RtpError, match="cells"):
        f.set_render(True)
        capi._check(capi.lib().rtp_fluids_read_field(f._h, capi.FIELD_DETECTOR,
                                                     np.zeros(8 * 20000, np.float32).ctypes.data,
                                                     20000), "read_field")
    f.close()
    b.close()


def test_device_render_pointers():
    sc = scenes.dam_break((16, 16, 16), (4, 2, 2), (40, 20, 20), seed=15, extra=512)
    gpu, _ = make_pbf_pair(sc, iters=1)
    with pytest.raises(capi.RtpError, match="never enabled"):
        gpu.device_render()
    gpu.set_render(True)
    col0, dist0, det0 = gpu.device_render()
    assert col0 and dist0 and det0
    gpu.update(1)
    col1, dist1, det1 = gpu.device_render()
    assert det1 == det0
    # a paused update only sorts, so p_col moves to the other buffer of its pair
    gpu.pause(True)
    gpu.update(1)
    col2, _, _ = gpu.device_render()
    assert col2 != col1
    gpu.close()
