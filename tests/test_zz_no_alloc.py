"""rtp_grid_build_noalloc: the grid build on caller-owned device buffers with no allocation.

Its outputs must equal rtp_grid_build's (which the other grid tests compare with the oracle)
bit for bit, over the whole of maxN and the whole cell table, on two calls in a row: the
1,048,576-particle dam break (multi-pass sort) with special positions mixed in, on a stream
the test creates, then a smaller scene in another box (one-workgroup sort) on the null
stream.  The outputs start filled with garbage, so every word must be written by the call.
"""
import ctypes

import numpy as np
import pytest

from realtimeparticles_amd import capi, scenes

H2D, D2H = 1, 2


class _Hip:
    """The HIP runtime librtp.so itself links (so device pointers are shared with it)."""

    def __init__(self):
        capi.lib()
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.L = L
        self.ptrs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self._ok(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)), "hipMalloc")
        self.ptrs.append(p.value)
        self._ok(self.L.hipMemcpy(p.value, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        self._ok(self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._ok(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def sync(self, s: int) -> None:
        self._ok(self.L.hipStreamSynchronize(s), "hipStreamSynchronize")

    def close(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _garbage(shape) -> np.ndarray:
    return np.full(shape, 0xA5A5A5A5, np.uint32)


def _noalloc_build(hip, pos, max_n, box, res, stream):
    n = pos.shape[0]
    nc = int(res[0]) * int(res[1]) * int(res[2])
    pos_d = hip.upload(np.ascontiguousarray(pos, np.float32))
    keys_d = hip.upload(_garbage(max_n))
    idx_d = hip.upload(_garbage(max_n))
    se_d = hip.upload(_garbage((nc, 2)))
    capi.grid_build_noalloc(pos_d, n, max_n, box, res, keys_d, idx_d, se_d, stream=stream)
    hip.sync(stream)
    return (hip.download(keys_d, _garbage(max_n)), hip.download(idx_d, _garbage(max_n)),
            hip.download(se_d, _garbage((nc, 2))))


def _special_positions(pos, box):
    """Overwrite a few positions with NaN, +-inf, the walls and points beyond them."""
    w = np.asarray(box, np.float32) / 2
    pos = pos.copy()
    pos[3, :3] = np.nan
    pos[10, :3] = [np.inf, -np.inf, 0.0]
    pos[17, :3] = w
    pos[18, :3] = -w
    pos[19, :3] = w * 3
    pos[20, :3] = [w[0], -w[1], 0.0]
    return pos


@pytest.mark.gpu
def test_grid_build_noalloc_matches_grid_build_twice():
    if capi.device_count() < 1:
        pytest.skip("needs a HIP device")
    hip = _Hip()
    s = hip.stream()
    try:
        big = scenes.pbf_scene("pbf-1m")
        pos1 = _special_positions(big.pos, big.box)
        print(f"noalloc call 1: {pos1.shape[0]} particles, maxN {big.max_n}, grid {big.res}",
              flush=True)
        assert big.max_n <= capi.NOALLOC_MAX_PARTICLES
        got = _noalloc_build(hip, pos1, big.max_n, big.box, big.res, s)
        want = capi.grid_build(pos1, big.max_n, big.box, big.res)
        for name, a, b in zip(("keys", "idx", "startEnd"), got, want):
            assert np.array_equal(a, b), f"call 1: {name} differs from rtp_grid_build"

        small = scenes.dam_break((20, 18, 16), (3, 2, 2), (30, 20, 20), seed=5, name="noalloc-small")
        pos2 = _special_positions(small.pos, small.box)
        print(f"noalloc call 2: {pos2.shape[0]} particles, maxN {small.max_n}, grid {small.res}",
              flush=True)
        assert pos2.shape[0] <= 8192 < pos1.shape[0]
        got = _noalloc_build(hip, pos2, small.max_n, small.box, small.res, 0)
        want = capi.grid_build(pos2, small.max_n, small.box, small.res)
        for name, a, b in zip(("keys", "idx", "startEnd"), got, want):
            assert np.array_equal(a, b), f"call 2: {name} differs from rtp_grid_build"
    finally:
        hip.sync(s)
        hip.L.hipStreamDestroy(s)
        hip.close()


def test_grid_build_noalloc_rejects_sizes_beyond_its_array():
    d = capi._desc(capi.NOALLOC_MAX_PARTICLES + 1, 16, (2, 2, 2), (20, 20, 20), 0.0, 100, 0)
    rc = capi.lib().rtp_grid_build_noalloc(ctypes.byref(d), 1, 1, 1, 1, None)
    assert rc == -1
    assert b"exceeds" in capi.lib().rtp_last_error()
