"""The 96x96x192, 64-frequency precision study against the FP64 oracle.

meteorscattering_amd.precision runs the solve (2000 steps, 30-80 MHz,
collocated current) in the all-FP64 GPU path, the FP32 path (FP64 NTFF
sums), the FP32 path with compensated FP32 NTFF sums on FP32 matrix tiles
(MSC_NTFF_ACC=32) and the FP64 path with those sums; here the reference is
the oracle's FP64 run of the same solve.  Every GPU result's RCS (bistatic,
two 1-degree principal cuts, directions within 40 dB of each cut's peak; and
monostatic) must be within 0.1 dB of the oracle's at every frequency, and
the all-FP64 path's NTFF payload within 1e-10.  The study's figures are
printed as one JSON line.
"""
from __future__ import annotations

import json
import sys
import time

import pytest

from meteorscattering_amd import precision

pytestmark = pytest.mark.gpu


def test_precision_study_against_the_fp64_oracle(oracle_lib):
    steps = 2000
    p, wp, nu = precision.workload(steps)
    t0 = time.perf_counter()
    o = oracle_lib.Oracle(p, wp, nu)
    while o.t < steps:
        o.step(min(500, steps - o.t))
        print(f"oracle step {o.t}: {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    ref = {"fields": [o.field(c) for c in range(9)], "ntff": o.ntff(), "seconds": time.perf_counter() - t0}
    o.close()
    out = precision.study(steps, ref, "FP64 oracle (CPU)")
    print("PRECISION_STUDY " + json.dumps(out), flush=True)
    assert out["gpu_fp64"]["ntff_max_rel_error"] < 1e-10
    for name in ("gpu_fp64", "gpu_fp32", "gpu_fp32_compensated_fp32_sums", "gpu_fp64_compensated_fp32_sums"):
        r = out[name]
        assert r["bistatic_db_error_max"] < 0.1 and r["monostatic_db_error_max"] < 0.1, (name, r)
