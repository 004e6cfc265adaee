"""Precision study on the 96x96x192 geometry with 64 NTFF frequencies.

    python -m meteorscattering_amd.precision [--steps 2000] [--json out.json]

Runs the same solve through the FP32 path (fields, J and CPML in FP32;
plasma coefficients and NTFF sums in FP64) and the all-FP64 path, both on
the GPU, and reports:
  - max relative error of each field kind (E, H, J) at the end, relative to
    the FP64 field's peak;
  - the NTFF payload's max relative error;
  - per frequency, the bistatic RCS error in dB over two principal-plane
    cuts at 1 degree (phi cut theta = 90 deg, phi 0..180; theta cut phi = 0,
    theta 0..180), over directions where the FP64 RCS is within 40 dB of its
    cut maximum, and the monostatic (backscatter) error.
The FP64 GPU path itself is pinned to the FP64 oracle at small sizes by
tests/test_gpu_parity.py.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs, rcs
from .solver import Solver


def run(p, wp, nu, precision: int) -> dict:
    t0 = time.perf_counter()
    with Solver(p, wp, nu, precision=precision) as s:
        s.step(p["tsteps"])
        s.sync()
        el = time.perf_counter() - t0
        fields = [s.field(c) for c in range(9)]
        ntff = s.ntff()
    return {"fields": fields, "ntff": ntff, "seconds": el}


def study(steps: int = 2000) -> dict:
    p, wp, nu = configs.ntff_m64_96x96x192(tsteps=steps)
    # the reference's node-averaged Lee-Kalluri drive grows without bound
    # over 2000 steps; the study runs the collocated form
    p.update(j_collocated=1, numfiles=1, savefields=0)
    r64 = run(p, wp, nu, 64)
    r32 = run(p, wp, nu, 32)
    out = {"grid": [p["nx"], p["ny"], p["nz"]], "steps": steps, "nntf": p["nntf"],
           "fntf_mhz": [round(f / 1e6, 4) for f in p["fntf"]],
           "seconds": {"fp64": round(r64["seconds"], 2), "fp32": round(r32["seconds"], 2)}}
    kinds = {}
    for name, comps in (("E", range(0, 3)), ("H", range(3, 6)), ("J", range(6, 9))):
        peak = max(np.abs(r64["fields"][c]).max() for c in comps)
        err = max(np.abs(r32["fields"][c] - r64["fields"][c]).max() for c in comps)
        kinds[name] = float(err / peak) if peak > 0 else 0.0
    out["field_max_rel_error"] = kinds
    a, b = r32["ntff"], r64["ntff"]
    out["ntff_max_rel_error"] = float(np.abs(a - b).max() / np.abs(b).max())
    t0 = time.perf_counter()
    e = rcs.db_errors(a, b, p)
    out["rcs_seconds"] = round(time.perf_counter() - t0, 1)
    out["bistatic_db_error_per_freq"] = [round(x, 5) for x in e["bistatic_db_per_freq"]]
    out["bistatic_db_error_max"] = round(e["bistatic_db_max"], 5)
    out["monostatic_db_error_per_freq"] = [round(x, 6) for x in e["monostatic_db_per_freq"]]
    out["monostatic_dbsm_fp64"] = [round(x, 3) for x in e["monostatic_dbsm_reference"]]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    out = study(args.steps)
    txt = json.dumps(out)
    if args.json:
        with open(args.json, "w") as f:
            f.write(txt + "\n")
    print(txt, flush=True)


if __name__ == "__main__":
    sys.exit(main())
