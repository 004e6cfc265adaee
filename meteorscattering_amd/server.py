"""Python binding of the multi-solve server (msc_server_* in include/msc.h).

Counterpart of the reference's batch layer (setup/meteorbatch.m,
setup/meteorrandom.m: one PBS job per run): here every run is a request to
one server that keeps many solves resident on the GPU and advances them
together.  No CPU fallback: without libmsc.so or a HIP device, construction
raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


def _check(rc: int, lib) -> None:
    if rc != 0:
        raise RuntimeError(f"libmsc error {rc}: {lib.msc_last_error().decode()}")


def footprint(p: dict, precision: int = 32) -> int:
    """Device bytes one solve of parameters `p` occupies (0: invalid)."""
    lib = _abi.load_library()
    s = _abi.params_to_struct(p)
    return int(lib.msc_solve_footprint(C.byref(s), precision))


def ntff_doubles(p: dict) -> int:
    M2 = 2 * p["nntf"]
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    return 2 * M2 + 8 * M2 * (ny * nz + nx * nz + nx * ny)


class Server:
    """Many solves on one GPU; see include/msc.h for the contract."""

    def __init__(self, precision: int = 32, budget_bytes: int = 0, fill_tiles: int | None = None,
                 graphs: bool = True):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        _check(self.lib.msc_server_create(precision, int(budget_bytes), C.byref(h)), self.lib)
        self._h = h
        if fill_tiles is not None or not graphs:
            _check(self.lib.msc_server_configure(h, 32768 if fill_tiles is None else int(fill_tiles), int(graphs)),
                   self.lib)
        self.precision = precision
        self.params: dict[int, dict] = {}
        self._maps: dict[int, tuple] = {}

    def set_tilings(self, classes: list[tuple[int, dict]]) -> None:
        """Tilings by size class, [(max_cells, tiling dict), ...] in
        increasing max_cells: a solve of n cells takes the first class with
        n <= max_cells (the last otherwise).  Only while nothing is queued
        or resident."""
        n = len(classes)
        arr = (_abi.Tiling * max(n, 1))(*[_abi.tiling_struct(t) for _, t in classes])
        mc = (C.c_int64 * max(n, 1))(*[int(m) for m, _ in classes])
        _check(self.lib.msc_server_set_tilings(self._h, arr, mc, n), self.lib)

    def submit(self, p: dict, wp: np.ndarray, nu: np.ndarray) -> int:
        s = _abi.params_to_struct(p)
        wp = np.ascontiguousarray(wp, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        rid = C.c_int64()
        _check(self.lib.msc_server_submit(self._h, C.byref(s), _abi.dptr(wp), _abi.dptr(nu), C.byref(rid)), self.lib)
        self.params[rid.value] = dict(p)
        # the worker reads the maps in place until the request is prepared
        self._maps[rid.value] = (wp, nu)
        return rid.value

    def advance(self, max_iterations: int = 1) -> int:
        """Run up to `max_iterations` gang iterations; returns solves completed."""
        n = C.c_int()
        _check(self.lib.msc_server_advance(self._h, int(max_iterations), C.byref(n)), self.lib)
        for rid in [r for r in self._maps if self._state(r) > _abi.REQ_QUEUED]:
            del self._maps[rid]
        return n.value

    def _state(self, rid: int) -> int:
        inf = _abi.RequestInfo()
        _check(self.lib.msc_server_info(self._h, rid, C.byref(inf)), self.lib)
        return inf.state

    def info(self, rid: int) -> dict:
        inf = _abi.RequestInfo()
        _check(self.lib.msc_server_info(self._h, rid, C.byref(inf)), self.lib)
        d = {f: getattr(inf, f) for f, _ in _abi.RequestInfo._fields_}
        d["state_name"] = _abi.REQ_STATES[inf.state]
        if inf.state in (_abi.REQ_FAILED, _abi.REQ_CANCELLED):
            d["error"] = self.lib.msc_last_error().decode()
        return d

    def ntff(self, rid: int) -> np.ndarray:
        out = np.empty(ntff_doubles(self.params[rid]))
        _check(self.lib.msc_server_get_ntff(self._h, rid, _abi.dptr(out), out.size), self.lib)
        return out

    def probes(self, rid: int) -> np.ndarray:
        p = self.params[rid]
        n = 6 * p["tsteps"] * p["nprobes"]
        out = np.empty(max(n, 1))
        _check(self.lib.msc_server_get_probes(self._h, rid, _abi.dptr(out), n), self.lib)
        return out[:n]

    def cancel(self, rid: int) -> None:
        """Stop a request that has steps left (msc_server_cancel)."""
        _check(self.lib.msc_server_cancel(self._h, rid), self.lib)

    def set_time_limit(self, seconds: float) -> None:
        """Fail requests that run longer than `seconds` once in the gang (0: no limit)."""
        _check(self.lib.msc_server_set_time_limit(self._h, float(seconds)), self.lib)

    def release(self, rid: int) -> None:
        _check(self.lib.msc_server_release(self._h, rid), self.lib)
        self.params.pop(rid, None)

    def stats(self) -> dict:
        st = _abi.ServerStats()
        _check(self.lib.msc_server_get_stats(self._h, C.byref(st)), self.lib)
        return {f: getattr(st, f) for f, _ in _abi.ServerStats._fields_}

    @property
    def clock(self) -> float:
        return self.lib.msc_server_clock(self._h)

    def run_until_done(self, chunk: int = 64) -> None:
        """Advance until nothing is queued, resident or draining."""
        while True:
            self.advance(chunk)
            st = self.stats()
            if not (st["queued"] or st["ready"] or st["running"] or st["draining"]):
                return

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.msc_server_destroy(self._h)
            self._h = None
        self._maps = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
