"""Tile-shape autotuner (``msc_autotune``) and per-size-class tables.

    from meteorscattering_amd import autotune
    tiling, ms = autotune.tune(p, wp, nu)               # one solve
    table = autotune.tune_classes(requests, bounds)     # server size classes

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
