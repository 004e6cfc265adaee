"""Tile-shape autotuner (``msc_autotune``) and per-size-class tables.

    from meteorscattering_amd import autotune
    tiling, ms = autotune.tune(p, wp, nu)               # one solve
    table = autotune.tune_classes(requests, bounds)     # server size classes, each on one solve
    table = autotune.tune_gang(requests, bounds)        # server size classes, inside the gang

The library times candidate tilings on the solve itself (a coordinate
search from the defaults over chunk2, chunk, ty and ty_b; each candidate a
fresh solver, 2 warm-up steps, then `steps` timed steps) and keeps the
fastest.  Tilings change speed only: every cell is computed by the same code
whatever the tiling, so results are bit-identical.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .server import Server


def default() -> dict:
    lib = _abi.load_library()
    t = _abi.Tiling()
    lib.msc_tiling_default(C.byref(t))
    return _abi.tiling_dict(t)


def tune(p: dict, wp: np.ndarray, nu: np.ndarray, precision: int = 32, steps: int = 16) -> tuple[dict, float]:
    """Fastest tiling for this solve and its elapsed GPU time per step (ms)."""
    lib = _abi.load_library()
    s = _abi.params_to_struct(p)
    wp = np.ascontiguousarray(wp, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    best = _abi.Tiling()
    ms = C.c_double(0.0)
    rc = lib.msc_autotune(C.byref(s), _abi.dptr(wp), _abi.dptr(nu), precision, int(steps), C.byref(best),
                          C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"libmsc error {rc}: {lib.msc_last_error().decode()}")
    return _abi.tiling_dict(best), float(ms.value)


def tune_classes(requests: list[tuple[dict, np.ndarray, np.ndarray]], bounds: list[int], precision: int = 32,
                 steps: int = 16) -> list[tuple[int, dict]]:
    """One tiling per size class: classes are cell counts up to each of
    `bounds` (increasing; the last class takes the rest), each tuned on its
    median-sized request.  Returns [(max_cells, tiling), ...] for
    Server.set_tilings; classes without a request are left out."""
    cells = [r[0]["nx"] * r[0]["ny"] * r[0]["nz"] for r in requests]
    edges = list(bounds) + [max(cells + [0])]
    table = []
    lo = 0
    for hi in edges:
        members = sorted((c, k) for k, c in enumerate(cells) if lo < c <= hi)
        lo = hi
        if not members:
            continue
        p, wp, nu = requests[members[len(members) // 2][1]]
        tiling, _ = tune(p, wp, nu, precision, steps)
        table.append((hi, tiling))
    return table


def _class_of(cells: int, bounds: list[int]) -> int:
    for c, b in enumerate(bounds):
        if cells <= b:
            return c
    return len(bounds)


def gang_time(srv: Server, requests, table, steps: int) -> float:
    """Seconds on the server clock for every request's first `steps` steps,
    all submitted at once to `srv` (which must hold no request) with size-class
    table `table`."""
    srv.set_tilings(table)
    t0 = srv.clock
    ids = [srv.submit(dict(p, tsteps=min(steps, p["tsteps"]), numfiles=1), wp, nu) for p, wp, nu in requests]
    srv.run_until_done(4)
    t1 = max(srv.info(i)["t_done"] for i in ids)
    for i in ids:
        if srv.info(i)["state_name"] != "done":
            raise RuntimeError(f"request {i} during gang tuning: {srv.info(i)}")
        srv.release(i)
    return t1 - t0


def tune_gang(requests: list[tuple[dict, np.ndarray, np.ndarray]], bounds: list[int], precision: int = 32,
              steps: int = 400, chunk2s=(32, 48, 64), chunks=(8, 12, 16), tys=(7, 15),
              budget_bytes: int = 0, log=None) -> tuple[list[tuple[int, dict]], float, float]:
    """Tilings by size class tuned inside the gang: a coordinate search per
    class (over chunk2, chunk, ty, ty_b, from the defaults), each candidate
    table timed on one batch of every request's first `steps` steps in one
    server (fill 0), keeping a candidate only when the batch is 0.5% faster.
    Every candidate tile is at least as large as the default's (two-step
    chunks of 32 planes or more, one-step chunks of 8 or more, 7 or 15
    rows), so no gang holds more workgroups than the default tiling's.
    Returns the table (for Server.set_tilings), the default table's batch
    time and the tuned one's (seconds).  Tilings change speed only, never
    results."""
    if min(chunk2s) < 32 or min(chunks) < 8 or min(tys) < 7:
        raise ValueError("gang tuning keeps tiles at least as large as the defaults (chunk2 >= 32, chunk >= 8, "
                         "rows >= 7)")
    cells = [p["nx"] * p["ny"] * p["nz"] for p, _, _ in requests]
    edges = list(bounds) + [max(cells)]
    present = sorted({_class_of(c, bounds) for c in cells})
    cur = {c: default() for c in present}

    def table_of(t):
        return [(edges[c], dict(t[c])) for c in present]

    with Server(precision=precision, budget_bytes=budget_bytes, fill_tiles=0) as srv:
        gang_time(srv, requests, table_of(cur), min(steps, 40))  # warm the server and the kernels
        base = best = gang_time(srv, requests, table_of(cur), steps)
        if log:
            log(f"gang tuning: defaults {best:.3f} s")
        for c in present:
            for field, values in (("chunk2", chunk2s), ("chunk", chunks), ("ty", tys), ("ty_b", tys)):
                start = cur[c][field]
                for v in values:
                    if v == start:
                        continue
                    cand = {k: dict(t) for k, t in cur.items()}
                    cand[c][field] = v
                    sec = gang_time(srv, requests, table_of(cand), steps)
                    if log:
                        log(f"gang tuning: class <= {edges[c]} {field}={v}: {sec:.3f} s")
                    if sec < best * 0.995:
                        cur, best = cand, sec
    return table_of(cur), base, best
