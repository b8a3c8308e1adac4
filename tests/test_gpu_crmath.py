"""The short correctly rounded square root and division of the neighbour passes
(csrc/rtp_crmath.hpp) against the compiler's general correctly rounded sequences, on the GPU.

The neighbour passes are bit-exact against the oracle only if these return exactly what
sqrtf and '/' return.  Square root: every non-negative bit pattern (0, denormals, normals,
+inf, NaNs).  Division: every one of the 2^23 divisor significands against 2,047 numerator
significands of both signs and +-0, at the corners of the operand domain
(d in [2^-40, 2^40], |n| in [2^-90, 2^50], |n|/d in [2^-120, 2^90]) and at the extremes the
passes produce (gradient numerators up to 2^25 over r down to 2^-27; poly6 values down to
2^-84 over wdq up to 2^30).  The full 2^46 significand pairs at one exponent pair run with
python -m realtimeparticles_amd.selfcheck --exhaustive.
"""
import pytest

from helpers import log
from realtimeparticles_amd import capi

pytestmark = pytest.mark.gpu


def test_sqrt_every_non_negative_float():
    bad, first = capi.selfcheck_sqrt(0, 0x80000000)
    log(f"sqrt: {bad} mismatches over 2^31 patterns")
    assert bad == 0, f"first mismatching pattern {first:#010x}"


@pytest.mark.parametrize("n_exp,d_exp", [
    (0, 0), (-90, -40), (-90, 30), (50, -40), (50, 40), (-90, 0), (0, -40), (0, 40),
    (25, -27), (-84, 30), (22, -30), (-30, -4),
])
def test_division_all_divisor_significands(n_exp, d_exp):
    bad, first = capi.selfcheck_division(n_exp, d_exp, n_sig_step=4099)
    log(f"division 2^{n_exp} / 2^{d_exp}: {bad} mismatches")
    assert bad == 0, f"first mismatching divisor significand {first:#08x}"


def test_selfcheck_arguments():
    with pytest.raises(capi.RtpError):
        capi.selfcheck_division(0, 0, 5, 4)
    with pytest.raises(capi.RtpError):
        capi.selfcheck_division(0, 200)
