"""Precision study on the 96x96x192 geometry with 64 NTFF frequencies, or
on the 400x400x800 headline solve.

    python -m meteorscattering_amd.precision [--steps 2000] [--json out.json]
    python -m meteorscattering_amd.precision --workload headline   # 6000 steps

The same solve (30-80 MHz, collocated Lee-Kalluri current) runs through
  - the all-FP64 GPU path (fields, J, CPML FP64; FP64 NTFF sums);
  - the FP32 GPU path (fields, J and CPML FP32; plasma coefficients FP64;
    FP64 NTFF sums on FP64 matrix tiles) - the default;
  - the FP32 GPU path with compensated FP32 NTFF sums (MSC_NTFF_ACC=32:
    FP32 samples and phase table on FP32 matrix tiles, FP32 accumulation
    over each K-step block, Neumaier-compensated FP32 running sums);
  - the FP64 GPU path with compensated FP32 NTFF sums (the accumulation's
    own error, free of the FP32 fields'; not on the headline solve).
Each run is compared with a reference run: the all-FP64 GPU path here, or
the FP64 oracle's run when the caller passes it (tests/test_gpu_precision.py
does), reporting
  - the max relative error of each field kind (E, H, J) at the end,
    relative to the reference's end-time peak of that kind (late-time fields
    are small after the pulse has left, which inflates these);
  - the NTFF payload's max relative error;
  - per frequency, the bistatic RCS error in dB over two principal-plane
    cuts at 1 degree, over directions within 40 dB of each cut's peak, and
    the monostatic (backscatter, meteorNTF.m) error.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import configs, rcs
from .solver import Solver


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(p, wp, nu, precision: int, acc: int) -> dict:
    old = os.environ.get("MSC_NTFF_ACC")
    os.environ["MSC_NTFF_ACC"] = str(acc)
    try:
        t0 = time.perf_counter()
        with Solver(p, wp, nu, precision=precision) as s:
            left = p["tsteps"]
            while left > 0:  # in chunks, with a progress line each
                n = min(left, 1000)
                s.step(n)
                s.sync()
                left -= n
                log(f"FP{precision} fields, NTFF sums {acc}: {p['tsteps'] - left} of {p['tsteps']} steps")
            el = time.perf_counter() - t0
            fields = [s.field(c) for c in range(9)]
            ntff = s.ntff()
    finally:
        if old is None:
            os.environ.pop("MSC_NTFF_ACC", None)
        else:
            os.environ["MSC_NTFF_ACC"] = old
    return {"fields": fields, "ntff": ntff, "seconds": el}


def compare(run: dict, ref: dict, e: dict) -> dict:
    """Field and payload errors of `run` against `ref`; `e` its RCS errors
    (rcs.db_errors)."""
    kinds = {}
    for name, comps in (("E", range(0, 3)), ("H", range(3, 6)), ("J", range(6, 9))):
        peak = max(np.abs(ref["fields"][c]).max() for c in comps)
        err = max(np.abs(run["fields"][c] - ref["fields"][c]).max() for c in comps)
        kinds[name] = float(err / peak) if peak > 0 else 0.0
    a, b = run["ntff"], ref["ntff"]
    return {"field_max_rel_error": kinds,
            "ntff_max_rel_error": float(np.abs(a - b).max() / np.abs(b).max()),
            "bistatic_db_error_per_freq": [round(x, 6) for x in e["bistatic_db_per_freq"]],
            "bistatic_db_error_max": round(e["bistatic_db_max"], 6),
            "monostatic_db_error_per_freq": [round(x, 6) for x in e["monostatic_db_per_freq"]],
            "monostatic_db_error_max": round(e["monostatic_db_max"], 6),
            "seconds": round(run["seconds"], 2)}


def workload(steps: int = 2000, which: str = "m64"):
    """`m64`: the 96x96x192 trail with 64 frequencies; `headline`: the
    400x400x800 trail (12-cell CPML, 3-D media, nu(z), one frequency) over
    its whole 6000 steps by default."""
    if which == "headline":
        p, wp, nu = configs.trail_400x400x800(tsteps=steps)
        p.update(numfiles=1, savefields=0)
        return p, wp, nu
    p, wp, nu = configs.ntff_m64_96x96x192(tsteps=steps)
    # the reference's node-averaged Lee-Kalluri drive grows without bound
    # over 2000 steps; the study runs the collocated form
    p.update(j_collocated=1, numfiles=1, savefields=0)
    return p, wp, nu


def study(steps: int = 2000, reference: dict | None = None, reference_name: str | None = None,
          which: str = "m64") -> dict:
    """`reference`: {"fields": [9 arrays], "ntff": payload, "seconds": s} of
    the same workload (workload(steps, which)); None measures against the
    all-FP64 GPU path."""
    p, wp, nu = workload(steps, which)
    runs = {
        "gpu_fp64": run_gpu(p, wp, nu, 64, 64),
        "gpu_fp32": run_gpu(p, wp, nu, 32, 64),
        "gpu_fp32_compensated_fp32_sums": run_gpu(p, wp, nu, 32, 32),
    }
    if which != "headline":
        runs["gpu_fp64_compensated_fp32_sums"] = run_gpu(p, wp, nu, 64, 32)
    log("GPU runs done: " + ", ".join(f"{k} {v['seconds']:.1f} s" for k, v in runs.items()))
    if reference is None:
        reference, reference_name = runs["gpu_fp64"], "all-FP64 GPU path"
    ref = reference
    names = [n for n, r in runs.items() if r is not ref]
    t0 = time.perf_counter()
    errs = rcs.db_errors([runs[n]["ntff"] for n in names], ref["ntff"], p, log=log)
    out = {"grid": [p["nx"], p["ny"], p["nz"]], "steps": steps, "nntf": p["nntf"],
           "fntf_mhz": [round(f / 1e6, 4) for f in p["fntf"]], "reference": reference_name,
           "reference_seconds": round(ref["seconds"], 1), "rcs_seconds": round(time.perf_counter() - t0, 1),
           "monostatic_dbsm_reference": [round(x, 3) for x in errs[0]["monostatic_dbsm_reference"]]}
    for name, e in zip(names, errs):
        out[name] = compare(runs[name], ref, e)
        log(f"{name}: {json.dumps({k: v for k, v in out[name].items() if 'per_freq' not in k})}")
    a, b = runs["gpu_fp32_compensated_fp32_sums"]["ntff"], runs["gpu_fp32"]["ntff"]
    out["compensated_fp32_vs_fp64_sums_ntff_rel"] = float(np.abs(a - b).max() / np.abs(b).max())
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["m64", "headline"], default="m64")
    ap.add_argument("--steps", type=int, default=None, help="default 2000 (m64), 6000 (headline)")
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    steps = args.steps or (6000 if args.workload == "headline" else 2000)
    out = study(steps, which=args.workload)
    txt = json.dumps(out)
    if args.json:
        with open(args.json, "w") as f:
            f.write(txt + "\n")
    print(txt, flush=True)


if __name__ == "__main__":
    sys.exit(main())
