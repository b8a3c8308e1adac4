"""Run the on-device checks of the short correctly rounded square root and division
(csrc/rtp_crmath.hpp) against the compiler's general sequences.

    python -m realtimeparticles_amd.selfcheck              # what tests/test_gpu_crmath.py runs
    python -m realtimeparticles_amd.selfcheck --exhaustive # every pair of significands at
                                                            # 2^0 / 2^0, about two minutes
"""
from __future__ import annotations

import argparse
import sys
import time

from . import capi


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--exhaustive", action="store_true")
    ap.add_argument("--chunks", type=int, default=64)
    a = ap.parse_args(argv)
    t0 = time.time()
    bad, first = capi.selfcheck_sqrt(0, 0x80000000)
    print(f"sqrt over every non-negative pattern: {bad} mismatches ({time.time() - t0:.1f} s)",
          flush=True)
    total = bad
    if a.exhaustive:
        step = (1 << 23) // a.chunks
        for k in range(a.chunks):
            lo, hi = k * step, (k + 1) * step if k + 1 < a.chunks else 1 << 23
            bad, first = capi.selfcheck_division(0, 0, lo, hi, 1)
            total += bad
            print(f"division 2^0/2^0, divisor significands [{lo:#x}, {hi:#x}) x all numerator "
                  f"significands: {bad} mismatches ({time.time() - t0:.0f} s)", flush=True)
    else:
        for ne, de in ((0, 0), (-90, -40), (-90, 30), (50, -40), (50, 40)):
            bad, first = capi.selfcheck_division(ne, de, n_sig_step=4099)
            total += bad
            print(f"division 2^{ne}/2^{de}: {bad} mismatches", flush=True)
    print("total mismatches:", total, flush=True)
    return 0 if total == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
