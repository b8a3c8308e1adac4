"""GPU parity of the uniform grid: cell hash, stable radix sort, cell start/end (bit-exact)."""
import numpy as np
import pytest

from helpers import _eq, log, oracle_grid
from realtimeparticles_amd import capi, scenes

pytestmark = pytest.mark.gpu


# 4096 / 4097 and 8192 / 8193 straddle the one-workgroup sort's chunk sizes and its limit;
# 2^18, 2^20 and 2^22 the multi-launch sort's subtile sizes (4, 8, 16 and 32 keys per lane)
@pytest.mark.parametrize("n", [1, 63, 64, 2047, 2048, 2049, 4096, 4097, 8191, 8192, 8193, 65536,
                               100003, (1 << 18) - 1, 1 << 18, (1 << 20) - 1, 1 << 20,
                               (1 << 22) - 1, 1 << 22])
@pytest.mark.parametrize("hi", [2, 1 << 9, 1 << 18, (1 << 32) - 1])
def test_sort_pairs_is_stable_sort(n, hi):
    log(f"sort_pairs n={n} hi={hi}")
    rng = np.random.default_rng(n ^ hi)
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    out, perm = capi.sort_pairs(keys, max_key=hi - 1 if hi > 1 else 1)
    ref = np.argsort(keys, kind="stable")
    _eq(perm, ref.astype(np.uint32), "perm")
    _eq(out, keys[ref], "keys")


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8, 9, 10, 11])
@pytest.mark.parametrize("n", [8193, 300001])
def test_sort_pairs_every_digit_width(n, bits):
    """Each digit width of the multi-launch sort (the widths the launch tuning picks from),
    over keys of 2 NC + maxN size, as the stale tail of a shrunk model produces."""
    log(f"sort_pairs n={n} bits={bits}")
    rng = np.random.default_rng(bits * 7 + n)
    hi = 2 * 256000 + n
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    out, perm = capi.sort_pairs(keys, max_key=hi - 1, digit_bits=bits)
    ref = np.argsort(keys, kind="stable")
    _eq(perm, ref.astype(np.uint32), "perm")
    _eq(out, keys[ref], "keys")


def test_sort_pairs_nearly_sorted_input():
    # the next step's keys are the previous order with a few cells changed
    rng = np.random.default_rng(4)
    keys = np.sort(rng.integers(0, 259240, 1 << 20)).astype(np.uint32)
    flip = rng.choice(keys.size, 5000, replace=False)
    keys[flip] = rng.integers(0, 259240, 5000).astype(np.uint32)
    out, perm = capi.sort_pairs(keys, max_key=259239)
    ref = np.argsort(keys, kind="stable")
    _eq(perm, ref.astype(np.uint32), "perm")


def special_positions(n, box, rng):
    w = np.asarray(box, np.float32) / 2
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = (rng.uniform(-1.1, 1.1, (n, 3)) * w).astype(np.float32)
    k = n // 8
    pos[:k, 0] = w[0]                                   # exactly on +wall: index == res
    pos[k:2 * k, 1] = -w[1]
    pos[2 * k:3 * k, :3] = np.round(pos[2 * k:3 * k, :3], 1)  # on cell faces
    pos[3 * k, :3] = np.nan
    pos[3 * k + 1, :3] = np.inf
    return pos


@pytest.mark.parametrize("name", ["pbf-65k", "pbf-1m"])
def test_grid_build_matches_oracle(name):
    sc = scenes.pbf_scene(name)
    rng = np.random.default_rng(17)
    for trial, pos in enumerate([sc.pos, special_positions(sc.n, sc.box, rng)]):
        log(f"grid_build {name} trial {trial}")
        keys, idx, se = capi.grid_build(pos, sc.max_n, sc.box, sc.res, 100)
        rk, ri, rs = oracle_grid(pos, sc.max_n, sc.box, sc.res, 100)
        _eq(keys, rk, "cellIDs")
        _eq(idx, ri, "sort index")
        _eq(se, rs, "startEnd")


@pytest.mark.parametrize("cap", [1, 3, 100])
@pytest.mark.parametrize("max_extra", [1, 512, 700])
def test_grid_cap_and_first_cell(cap, max_extra):
    # crowded cells (runs longer than the cap), one cell holding sorted element 0 (Q1),
    # and max_n values that are not multiples of 512
    log(f"grid cap={cap} extra={max_extra}")
    rng = np.random.default_rng(cap * 1000 + max_extra)
    box, res = (2.0, 2.0, 2.0), (8, 8, 8)
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    pos[:400, :3] = rng.uniform(-1, -0.76, (400, 3)).astype(np.float32)   # crowd cell 0
    pos[400:900, :3] = rng.uniform(0.3, 0.45, (500, 3)).astype(np.float32)
    max_n = n + max_extra
    keys, idx, se = capi.grid_build(pos, max_n, box, res, cap)
    rk, ri, rs = oracle_grid(pos, max_n, box, res, cap)
    _eq(keys, rk, "cellIDs")
    _eq(idx, ri, "sort index")
    _eq(se, rs, "startEnd")


def test_grid_n_equals_max_n():
    # fillEndCell reads key[n]; with n == max_n that read is past the buffer and counts as
    # a different key
    rng = np.random.default_rng(8)
    n = 4096
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    keys, idx, se = capi.grid_build(pos, n, (2, 2, 2), (10, 10, 10), 100)
    rk, ri, rs = oracle_grid(pos, n, (2, 2, 2), (10, 10, 10), 100)
    _eq(keys, rk, "cellIDs")
    _eq(se, rs, "startEnd")
