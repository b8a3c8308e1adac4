"""The C ABI library: it loads, exports exactly what include/rtp.h declares, and refuses to
run without a GPU instead of falling back to a CPU path."""
import ctypes
import os
import re
import subprocess

import pytest

from realtimeparticles_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rtp.h")


def declared_in_header():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(rtp_[a-z0-9_]+)\s*\(", text, re.M)))


def exported_by_library():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return sorted({line.split()[-1] for line in out.splitlines() if " T rtp_" in line})


def test_library_exports_every_declared_symbol():
    decl = declared_in_header()
    assert len(decl) >= 40
    exp = exported_by_library()
    assert decl == exp
    L = capi.lib()
    for name in decl:
        assert hasattr(L, name)
    # the Python binding covers the whole ABI
    assert sorted(capi.declared_symbols()) == decl


def test_abi_version_and_error_text():
    L = capi.lib()
    assert L.rtp_abi_version() == 1
    assert isinstance(L.rtp_last_error(), bytes)


def test_struct_layouts_match_reference_sizes():
    # FluidParams (define.cl:13-25) is 44 bytes, BoidsRuleParams 16 bytes
    assert ctypes.sizeof(capi.FluidParams) == 44
    assert ctypes.sizeof(capi.BoidsParams) == 16
    assert ctypes.sizeof(capi.ModelDesc) == 44  # 2 u32 + 3 f32 + 3 i32 + u32 + f32 + i32


def test_argument_checks_do_not_need_a_device():
    L = capi.lib()
    assert L.rtp_fluids_create(None, None) == -1
    assert b"NULL" in L.rtp_last_error()
    assert L.rtp_fluids_set_params(None, None) == -1
    assert L.rtp_boids_update(None, 1) == -1


def test_clouds_slot_limit_is_checked_before_the_device():
    """A Clouds survivor-list entry packs the particle index into 28 bits: the model refuses
    2^28 slots or more with an argument error, before it looks for a device."""
    d = capi.ModelDesc()
    d.maxNbParticles = 1 << 28
    d.currNbParticles = 0
    for k in range(3):
        d.boxSize[k] = 10.0
        d.gridRes[k] = 30
    h = ctypes.c_void_p()
    assert capi.lib().rtp_clouds_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"2^28" in capi.lib().rtp_last_error()


@pytest.mark.skipif(capi.device_count() > 0, reason="a GPU is visible")
def test_no_cpu_fallback_without_gpu():
    with pytest.raises(capi.RtpError, match="no HIP device"):
        capi.Fluids(1024, 512, (2, 2, 2), (20, 20, 20))
    with pytest.raises(capi.RtpError):
        capi.sort_pairs([3, 1, 2])
