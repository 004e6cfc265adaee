"""Benchmark of the fused FDTD step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 400x400x800|96x96x192|96x96x192_m64] [--no-cpu]
    python bench.py --workload batch [--requests 48]          # 48-trail sweep, server vs serial
    python bench.py --workload serve --rate R [--requests N]  # Poisson stream, server vs serial

A step is one time step of the whole grid: the fused Yee H + plasma J + Yee
E kernels (with CPML, TF/SF), the NTFF running DFT and the probe/cadence
kernels.  Steps run in pairs: the two-step kernel fdtd_pair advances the
region 6 cells clear of every face by two steps per launch, the one-step
kernel the CPML slabs and the rest around it; the roofline object covers
fdtd_pair (the dominant kernel) on the per-cell-update basis (80 B where a
tile carries the plasma current J, 48 B where eps0 wp^2 is zero on the whole
tile and J stays zero), with the HBM bytes it actually moves per launch
(hbm_traffic) beside it, and the whole step.  The timed region ends with a
flush of the NTFF samples buffered since the last contraction.  The workload is the 400x400x800 FP32 single solve (128M cells,
12-cell CPML, Lee-Kalluri magnetized plasma, trail tilted 30 deg, nu(z)
with a 7 km scale height), synthetic media from a seed, all inputs resident
in HBM before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s (% HBM peak); RCS solves/hr + p50/p99 vs serial GPU solves"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, public spec
ALGO_BYTES_FP32 = 80           # E,H,J read+write (72 B) + eps0 wp^2, nu (8 B) per cell-update
ALGO_BYTES_FP64 = 160
# Cells of tiles where eps0 wp^2 is zero everywhere skip J (it stays zero)
# and the media: 48 B (E, H read+write) there, 80 B where J is carried
NOJ_BYTES_FP32 = 48
NOJ_BYTES_FP64 = 96


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, wp, nu, steps: int) -> dict:
    """The FP64 oracle (C restatement of the reference, OpenMP) on the host
    cores, on the same grid, timed over `steps` steps (set-up excluded)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = pyoracle.Oracle(p, wp, nu)
    o.step(1)  # first-touch of every array
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    ncell = p["nx"] * p["ny"] * p["nz"]
    return {"value": round(ncell * steps / dt / 1e9, 5), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps of the same {p['nx']}x{p['ny']}x{p['nz']} grid, FP64, after 1 warm step; "
                      f"{dt:.1f} s"}


def rcs_check() -> dict:
    """RCS error of the FP32 path, in dB: the 96x96x192 trail (collocated
    current, 16 frequencies 30-80 MHz) over 1200 steps in FP32 and in the
    all-FP64 path (which the tests pin to the FP64 oracle at 1e-10)."""
    from meteorscattering_amd import configs, rcs
    from meteorscattering_amd.solver import Solver

    p, wp, nu = configs.trail_96x96x192(tsteps=1200)
    p.update(j_collocated=1, nntf=16, fntf=list(np.linspace(30e6, 80e6, 16)), savefields=0, numfiles=1)
    pays = {}
    for prec in (32, 64):
        with Solver(p, wp, nu, precision=prec) as g:
            g.step(p["tsteps"])
            pays[prec] = g.ntff()
    e = rcs.db_errors(pays[32], pays[64], p)
    return {"bistatic_db_max": round(e["bistatic_db_max"], 5), "monostatic_db_max": round(e["monostatic_db_max"], 5),
            "workload": "96x96x192 trail, 16 frequencies 30-80 MHz, 1200 steps; bistatic over two 1-degree "
                        "principal cuts within 40 dB of each cut's peak",
            "against": "all-FP64 GPU path (pinned to the FP64 oracle at 1e-10 by the tests)"}


def traffic_from_profile(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_launch")


def batch_or_serve(args, copy_gbps) -> dict:
    """The trail sweep (--workload batch: every trail submitted at once) or
    the request stream (--workload serve: Poisson arrivals), on the
    multi-solve server, against
    the same trails run one at a time through the single-solve path.  The
    synthetic media are built before any timing; both timings include each
    solve's device set-up and NTFF read-back.  The serial baseline's serving
    latencies follow from its measured per-solve times and the same arrival
    times (one-at-a-time, first come first served)."""
    from meteorscattering_amd import configs, workloads

    t0 = time.perf_counter()
    if args.workload == "batch":
        trails = configs.trail_sweep(seed=args.seed, count=args.requests)
        wname = f"trail sweep: {args.requests} trails, nx,ny U[64,256], nz U[128,512], 1500-4000 steps, M=1"
    else:
        trails = configs.request_stream(seed=args.seed, count=args.requests)
        wname = (f"request stream: {args.requests} requests, Poisson {args.rate}/s, 64^3..256x256x1024 "
                 "log-uniform, 1500-4000 steps, M=16")
    log(f"{wname}: media built in {time.perf_counter() - t0:.1f} s")
    updates = workloads.cell_updates(trails)
    table = None
    if args.autotune:
        # one tiling per size class, tuned on each class's median request;
        # the server and the serial baseline both use the table
        from meteorscattering_amd import autotune
        t1 = time.perf_counter()
        table = autotune.tune_classes(trails, [4 << 20, 16 << 20], precision=args.precision)
        log(f"autotuned {len(table)} size classes in {time.perf_counter() - t1:.1f} s: {table}")
    gang = None
    if args.gang_tune:
        from meteorscattering_amd import autotune
        t1 = time.perf_counter()
        gtable, g0, g1 = autotune.tune_gang(trails, [4 << 20, 16 << 20], precision=args.precision, log=log)
        gang = {"table": [{"max_cells": int(m), **t} for m, t in gtable], "default_batch_s": round(g0, 3),
                "tuned_batch_s": round(g1, 3), "tuning_s": round(time.perf_counter() - t1, 1),
                "steps_per_trail": 400}
        log(f"gang-tuned in {gang['tuning_s']} s: {gtable} ({g0:.3f} -> {g1:.3f} s per tuning batch)")
    ser = workloads.run_serial(trails, precision=args.precision, tilings=table)
    log(f"serial: {ser['wall']:.2f} s busy")
    stable = [(g["max_cells"], {k: g[k] for k in ("ty", "ty_b", "chunk", "chunk2")}) for g in gang["table"]] \
        if gang else table
    if args.workload == "batch":
        res = workloads.run_batch(trails, precision=args.precision, tilings=stable)
        arrivals = np.zeros(len(trails))
    else:
        arrivals = workloads.poisson_arrivals(args.rate, len(trails), args.seed)
        res = workloads.run_serve(trails, arrivals, precision=args.precision, tilings=stable)
    same = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(ser["ntff"], res["ntff"]))
    nonfinite = sum(int(not np.all(np.isfinite(a))) for a in res["ntff"])
    lat = res["latency"] if "latency" in res else np.array([i["t_done"] - i["t_submit"] for i in res["info"]])
    slat = workloads.fifo_latencies(arrivals, ser["times"])
    sph = len(trails) / res["wall"] * 3600
    # the serial baseline under the same arrivals: one solve at a time, first
    # come first served, so its last result comes at the last arrival's
    # finish (its busy time when every request arrives at once, as in batch)
    serial_span = max(float(ser["wall"]), float(np.max(arrivals + slat)))
    sph_serial = len(trails) / serial_span * 3600
    log(f"serial under the same arrivals: {serial_span:.2f} s ({sph_serial:.0f} solves/hr)")
    log(f"{args.workload}: {res['wall']:.2f} s ({sph:.0f} solves/hr), x{sph / sph_serial:.2f} vs serial; "
        f"identical results: {same}")
    q = lambda a, f: round(float(np.percentile(a, f)), 3)
    return {
        "metric": METRIC, "value": round(sph, 1), "unit": "RCS solves/hr", "n_gpus": 1,
        "steps": int(sum(p["tsteps"] for p, _, _ in trails)), "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64", "data": "synthetic",
        "config": {"workload": wname, "seed": args.seed, "requests": len(trails),
                   "cells_total": int(sum(p["nx"] * p["ny"] * p["nz"] for p, _, _ in trails)),
                   "tilings": [{"max_cells": int(m), **t} for m, t in table] if table else "defaults"},
        "gcell_updates_per_s": round(updates / res["wall"] / 1e9, 3),
        "latency_s": {"p50": q(lat, 50), "p99": q(lat, 99), "max": q(lat, 100)},
        "serial_gpu": {"solves_per_hr": round(sph_serial, 1), "wall_s": round(serial_span, 3),
                       "busy_s": round(ser["wall"], 3),
                       "gcell_updates_per_s": round(updates / serial_span / 1e9, 3),
                       "latency_s": {"p50": q(slat, 50), "p99": q(slat, 99), "max": q(slat, 100)}},
        "server_setup_s": round(res["setup"], 3) if "setup" in res else None,
        "speedup_vs_serial": round(sph / sph_serial, 3),
        "results_identical_to_serial": bool(same),
        "gang_tuning": gang,
        "solves_with_nonfinite_ntff": nonfinite,
        "measured_stream_copy_gbps": round(copy_gbps, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="400x400x800")
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rcs", action="store_true", help="skip the FP32 RCS error record")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--rate", type=float, default=2.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--requests", type=int, default=48, help="batch/serve: number of trails")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--autotune", action="store_true",
                    help="batch/serve: tune one tiling per size class (autotune.tune_classes) first")
    ap.add_argument("--gang-tune", action="store_true",
                    help="batch/serve: tune the server's size-class tilings inside the gang (autotune.tune_gang) "
                         "first; the serial baseline keeps its own defaults")
    ap.add_argument("--copy-only", action="store_true",
                    help="run only the streaming-copy reference (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only one GPU is used in this build (asked for {args.gpus}); "
              "the solve is single-GPU and multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    from meteorscattering_amd import configs
    from meteorscattering_amd.solver import Solver, stream_copy_gbps

    copy_bytes = 2 << 30
    copy_gbps = stream_copy_gbps(copy_bytes, 20)
    log(f"stream copy {copy_bytes >> 20} MiB: {copy_gbps:.0f} GB/s (read + write)")
    if args.copy_only:
        print(json.dumps({"stream_copy_bytes": copy_bytes, "stream_copy_gbps": round(copy_gbps, 1)}))
        return

    if args.workload in ("batch", "serve"):
        print(json.dumps(batch_or_serve(args, copy_gbps)), flush=True)
        return

    t0 = time.perf_counter()
    if args.workload == "400x400x800":
        p, wp, nu = configs.trail_400x400x800(seed=0)
        wname = "400x400x800_fp32_lk_cpml12"
    elif args.workload == "96x96x192":
        p, wp, nu = configs.trail_96x96x192()
        wname = "96x96x192_fp32_lk_cpml10"
    elif args.workload == "96x96x192_m64":
        p, wp, nu = configs.ntff_m64_96x96x192()
        wname = "96x96x192_fp32_lk_cpml10_m64"
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    if args.precision == 64:
        wname = wname.replace("fp32", "fp64")
    total = args.warmup + args.steps
    if total > p["tsteps"]:
        p["tsteps"] = total
        p["numfiles"] = min(p["numfiles"], total)
    ncell = p["nx"] * p["ny"] * p["nz"]
    log(f"workload {wname}: {ncell} cells, media built in {time.perf_counter() - t0:.1f} s")

    s = Solver(p, wp, nu, precision=args.precision)
    log(f"solver resident: {s.device_bytes / 2**30:.2f} GiB on device")
    s.step(args.warmup)
    s.sync()
    s.set_timing(True)
    t1 = time.perf_counter()
    s.step(args.steps)
    # the NTFF samples buffered since the last K-step contraction are
    # contracted inside the timed region, so it holds at least one
    # contraction whatever the step count
    s.flush_ntff()
    s.sync()
    el = time.perf_counter() - t1
    kms = s.kernel_ms
    tm = s.timing()
    # fdtd_pair's memory traffic alone (same tiles, loads and stores, no
    # arithmetic; the solve's state is unchanged), after the timed region
    pattern_ms = s.pattern_probe(10) if tm["pair_launches"] > 0 else None
    s.close()
    ms_per_step = el * 1e3 / args.steps
    value = ncell * args.steps / el / 1e9
    per_update = ALGO_BYTES_FP32 if args.precision == 32 else ALGO_BYTES_FP64
    noj = NOJ_BYTES_FP32 if args.precision == 32 else NOJ_BYTES_FP64
    jb = per_update - noj
    algo_step = noj * ncell + jb * tm["step_j_cells"]
    f_step = tm["step_j_cells"] / ncell
    span = tm["span_ms"]  # elapsed GPU time per step, all kernels
    step_achieved = algo_step / (span * 1e-3) / 1e9 if span > 0 else None
    if tm["pair_launches"] > 0:
        # two cell-updates per cell of the region per launch
        kernel, kname, kernel_ms = "fdtd_pair", "fdtd_pair (two-step region)", tm["pair_ms"]
        algo = 2 * (noj * tm["pair_cells"] + jb * tm["pair_j_cells"])
        f_pair = tm["pair_j_cells"] / tm["pair_cells"]
        basis = (f"({noj} + {jb} f) B per cell-update, f = {f_pair:.4f} the fraction of the region's cells in "
                 f"tiles that carry J, x 2 updates x {tm['pair_cells']} cells of the two-step region; each "
                 f"cell's fields are read and written once per launch, so the measured bytes (traffic) are "
                 f"about half of this")
    else:
        kernel, kname, kernel_ms = "fdtd_step", "fdtd_step (all one-step launches of a step)", kms
        algo = algo_step
        basis = f"({noj} + {jb} f) B per cell-update, f = {f_step:.4f}, x {ncell} cells"
    achieved = algo / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    traffic = traffic_from_profile(kernel, wname)
    hbm = None
    if traffic and kernel_ms > 0:
        hbm_gbs = traffic / (kernel_ms * 1e-3) / 1e9
        hbm = {"bytes_per_launch": traffic, "achieved": round(hbm_gbs, 1), "frac": round(hbm_gbs / HBM_PEAK_GBS, 4),
               "frac_of_copy": round(hbm_gbs / copy_gbps, 4),
               "basis": "HBM bytes the launch actually moves (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                        "calibrated on a copy of known size; profiles/) over this run's launch duration"}
    pattern = None
    if pattern_ms:
        # the launch's minimum bytes: each cell's E, H (and J, media where the
        # tile carries J) read once and written once
        minb = noj * tm["pair_cells"] + jb * tm["pair_j_cells"]
        pattern = {"ms_per_launch": round(pattern_ms, 4), "min_bytes_per_launch": minb,
                   "achieved_min_bytes": round(minb / (pattern_ms * 1e-3) / 1e9, 1),
                   "fdtd_pair_min_bytes": round(minb / (kernel_ms * 1e-3) / 1e9, 1),
                   "fdtd_pair_frac_of_pattern": round(pattern_ms / kernel_ms, 4),
                   "basis": "pair_pattern: fdtd_pair's tiles, loads (tile and 2-cell halo, one plane ahead) and "
                            "owned stores with no arithmetic, timed after the timed region; bandwidth on the "
                            "launch's minimum bytes (each value read and written once)"}
        log(f"fdtd_pair memory pattern alone: {pattern_ms:.3f} ms/launch ({pattern['achieved_min_bytes']:.0f} GB/s "
            f"on minimum bytes); fdtd_pair {kernel_ms:.3f} ms ({pattern['fdtd_pair_frac_of_pattern']:.2f} of it)")
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, GPU span {span:.3f} ms/step, "
        f"{value:.2f} Gcell/s; {kname} {kernel_ms:.3f} ms/launch, {achieved:.0f} GB/s algorithmic"
        + (f", {hbm['achieved']:.0f} GB/s of HBM traffic" if hbm else ""))
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64", "data": "synthetic",
        "config": {"workload": f"fused Yee+Lee-Kalluri plasma step, {wname}", "nx": p["nx"], "ny": p["ny"],
                   "nz": p["nz"], "pmllen": p["pmllen"], "ntff_freqs": p["nntf"],
                   "media": ("Gaussian trail tilted 30 deg, nu(z) 7 km scale height, seed 0"
                             if args.workload == "400x400x800" else "Gaussian trail along z, n0 1e12, r0 1.5 m"),
                   "j_collocated": p.get("j_collocated", 0),
                   "ntff_sums": "compensated FP32" if os.environ.get("MSC_NTFF_ACC") == "32" else "FP64"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic,
                     "achieved_is": "algorithmic bytes per launch / launch duration: an effective bandwidth, "
                                    "which exceeds the HBM traffic when a launch advances its cells two steps",
                     "hbm_traffic": hbm,
                     "kernel": kname, "kernel_ms": round(kernel_ms, 4),
                     "launches": tm["pair_launches"] if kernel == "fdtd_pair" else args.steps,
                     "cells_per_launch": tm["pair_cells"] if kernel == "fdtd_pair" else ncell,
                     "j_cells_per_launch": tm["pair_j_cells"] if kernel == "fdtd_pair" else tm["step_j_cells"],
                     "updates_per_cell": 2 if kernel == "fdtd_pair" else 1,
                     "algorithmic_bytes_per_launch": algo, "algorithmic_basis": basis,
                     "step": {"gpu_ms_per_step": round(span, 4),
                              "achieved": round(step_achieved, 1) if step_achieved else None,
                              "frac": round(step_achieved / HBM_PEAK_GBS, 4) if step_achieved else None,
                              "basis": f"({noj} + {jb} f) B per cell-update, f = {f_step:.4f}, x {ncell} cells "
                                       f"per step over the elapsed GPU time of all the step's kernels (HIP "
                                       f"events, first to last)"},
                     "measured_stream_copy_gbps": round(copy_gbps, 1),
                     "frac_of_copy": round(achieved / copy_gbps, 4) if achieved else None,
                     "pattern": pattern},
    }
    if not args.no_rcs:
        t2 = time.perf_counter()
        out["rcs_db_err"] = rcs_check()
        log(f"rcs check: {out['rcs_db_err']} ({time.perf_counter() - t2:.1f} s)")
    if not args.no_cpu:
        t2 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(p, wp, nu, args.cpu_steps)
        log(f"cpu baseline: {out['cpu_baseline']} ({time.perf_counter() - t2:.1f} s incl. set-up)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
