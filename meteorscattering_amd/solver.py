"""Python binding of the solver handle (msc_solver_* in include/msc.h).

Every call goes through libmsc.so (HIP kernels for gfx950).  There is no
CPU fallback: without the library or a HIP device, construction raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


def _check(rc: int, lib) -> None:
    if rc != 0:
        raise RuntimeError(f"libmsc error {rc}: {lib.msc_last_error().decode()}")


class Solver:
    """One FDTD solve resident on the GPU."""

    def __init__(self, p: dict, wp: np.ndarray, nu: np.ndarray, precision: int = 32, tiling: dict | None = None):
        """`tiling`: {ty, ty_b, chunk, chunk2} (``msc_tiling``, e.g. from
        autotune.tune); None takes the defaults."""
        self.lib = _abi.load_library()
        self.p = dict(p)
        self._s = _abi.params_to_struct(p)
        wp = np.ascontiguousarray(wp, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        h = C.c_void_p()
        if tiling is None:
            rc = self.lib.msc_solver_create(C.byref(self._s), _abi.dptr(wp), _abi.dptr(nu), precision, C.byref(h))
        else:
            t = _abi.tiling_struct(tiling)
            rc = self.lib.msc_solver_create_tiled(C.byref(self._s), _abi.dptr(wp), _abi.dptr(nu), precision,
                                                  C.byref(t), C.byref(h))
        _check(rc, self.lib)
        self._h = h
        self.precision = precision
        self.shape = (p["nx"], p["ny"], p["nz"])

    def step(self, n: int = 1) -> None:
        _check(self.lib.msc_solver_step(self._h, int(n)), self.lib)

    def sync(self) -> None:
        _check(self.lib.msc_solver_sync(self._h), self.lib)

    @property
    def t(self) -> int:
        return self.lib.msc_solver_time(self._h)

    def field(self, comp: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(self.lib.msc_solver_get_field(self._h, comp, _abi.dptr(out)), self.lib)
        return out

    def field_noalloc(self, comp: int, out: np.ndarray | None = None) -> np.ndarray:
        """field() through the library's fixed device scratch
        (msc_solver_get_field_noalloc: no device allocation)."""
        if out is None:
            out = np.empty(self.shape, dtype=np.float64)
        _check(self.lib.msc_solver_get_field_noalloc(self._h, comp, _abi.dptr(out), out.size), self.lib)
        return out

    def ntff(self) -> np.ndarray:
        n = self.lib.msc_solver_ntff_count(self._h)
        out = np.empty(n, dtype=np.float64)
        _check(self.lib.msc_solver_get_ntff(self._h, _abi.dptr(out), n), self.lib)
        return out

    def flush_ntff(self) -> None:
        """Contract the NTFF samples buffered since the last K-step block
        (asynchronous; ntff() does it itself)."""
        _check(self.lib.msc_solver_flush_ntff(self._h), self.lib)

    def probes(self) -> np.ndarray:
        n = 6 * self.p["tsteps"] * self.p["nprobes"]
        out = np.empty(max(n, 1), dtype=np.float64)
        _check(self.lib.msc_solver_get_probes(self._h, _abi.dptr(out), n), self.lib)
        return out[:n]

    def nonfinite(self) -> bool:
        """True once a NaN or infinite value has reached the NTFF surface."""
        v = C.c_int()
        _check(self.lib.msc_solver_nonfinite(self._h, C.byref(v)), self.lib)
        return bool(v.value)

    @property
    def nframes(self) -> int:
        return self.lib.msc_solver_nframes(self._h)

    def frame(self, f: int):
        n = self.lib.msc_solver_frame_count(self._h)
        em = np.zeros(1)
        if self.p.get("savefields", 1):
            e = np.empty(n)
            h = np.empty(n)
            _check(self.lib.msc_solver_get_frame(self._h, f, _abi.dptr(e), _abi.dptr(h), _abi.dptr(em)), self.lib)
            return e, h, float(em[0])
        _check(self.lib.msc_solver_get_frame(self._h, f, None, None, _abi.dptr(em)), self.lib)
        return None, None, float(em[0])

    def max_abs_ex(self) -> float:
        v = C.c_double()
        _check(self.lib.msc_solver_max_abs_ex(self._h, C.byref(v)), self.lib)
        return v.value

    @property
    def device_bytes(self) -> int:
        return self.lib.msc_solver_device_bytes(self._h)

    def set_timing(self, on: bool = True) -> None:
        _check(self.lib.msc_solver_set_timing(self._h, 1 if on else 0), self.lib)

    @property
    def kernel_ms(self) -> float:
        return self.lib.msc_solver_kernel_ms(self._h)

    def pattern_probe(self, reps: int = 10) -> float:
        """Mean ms of `reps` launches of the two-step kernel's memory traffic
        with no arithmetic over this solve's region (msc_solver_pattern_probe);
        the solve's state is unchanged."""
        v = C.c_double()
        _check(self.lib.msc_solver_pattern_probe(self._h, int(reps), C.byref(v)), self.lib)
        return v.value

    def timing(self) -> dict:
        """Timing of the last step() call by launch kind (``msc_timing``)."""
        t = _abi.Timing()
        _check(self.lib.msc_solver_get_timing(self._h, C.byref(t)), self.lib)
        return {f: getattr(t, f) for f, _ in t._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.msc_solver_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_copy_gbps(nbytes: int, iters: int = 10) -> float:
    """Attainable device bandwidth (read + write) of a plain streaming copy."""
    lib = _abi.load_library()
    v = C.c_double()
    _check(lib.msc_stream_copy(int(nbytes), int(iters), C.byref(v)), lib)
    return v.value


def run_dir(path: str, precision: int = 32) -> None:
    """The drop-in process (msc_run_dir): read, solve and write in `path`."""
    lib = _abi.load_library()
    _check(lib.msc_run_dir(path.encode(), precision), lib)
