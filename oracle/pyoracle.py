"""ctypes wrapper of oracle/liboracle.so (the FP64 CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always to check the HIP path's results or as the CPU baseline,
never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from meteorscattering_amd._abi import MscParams, params_to_struct  # noqa: E402

LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        P = C.POINTER
        dp = P(C.c_double)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [P(MscParams), dp, dp, C.c_char_p, C.c_int]
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_time.restype = C.c_int
        L.orc_time.argtypes = [vp]
        L.orc_get_field.argtypes = [vp, C.c_int, dp]
        L.orc_ntff_count.restype = C.c_size_t
        L.orc_ntff_count.argtypes = [vp]
        L.orc_get_ntff.argtypes = [vp, dp]
        L.orc_probe_count.restype = C.c_size_t
        L.orc_probe_count.argtypes = [vp]
        L.orc_get_probes.argtypes = [vp, dp]
        L.orc_frame_count.restype = C.c_size_t
        L.orc_frame_count.argtypes = [vp]
        L.orc_get_frames.argtypes = [vp, dp, dp]
        L.orc_max_abs_ex.restype = C.c_double
        L.orc_max_abs_ex.argtypes = [vp]
        L.orc_get_media.argtypes = [vp, dp, dp]
        L.orc_lk_matrices.argtypes = [C.c_double, C.c_double, dp, dp, dp]
        L.orc_cpml_tables.argtypes = [vp, dp, dp, dp, dp]
        L.orc_source_tables.argtypes = [vp, dp, dp]
        L.orc_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One FP64 CPU solve with the reference's semantics."""

    def __init__(self, p: dict, wp: np.ndarray, nu: np.ndarray):
        self.p = dict(p)
        self._s = params_to_struct(p)
        self._wp = np.ascontiguousarray(wp, dtype=np.float64)
        self._nu = np.ascontiguousarray(nu, dtype=np.float64)
        err = C.create_string_buffer(256)
        self._h = lib().orc_create(C.byref(self._s), _dp(self._wp), _dp(self._nu), err, 256)
        if not self._h:
            raise ValueError(f"oracle: {err.value.decode()}")
        self.shape = (p["nx"], p["ny"], p["nz"])

    def step(self, n: int = 1) -> None:
        lib().orc_step(self._h, int(n))

    @property
    def t(self) -> int:
        return lib().orc_time(self._h)

    def field(self, comp: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        lib().orc_get_field(self._h, comp, _dp(out))
        return out

    def ntff(self) -> np.ndarray:
        out = np.empty(lib().orc_ntff_count(self._h), dtype=np.float64)
        lib().orc_get_ntff(self._h, _dp(out))
        return out

    def probes(self) -> np.ndarray:
        out = np.empty(lib().orc_probe_count(self._h), dtype=np.float64)
        lib().orc_get_probes(self._h, _dp(out))
        return out

    def frames(self) -> tuple[np.ndarray, np.ndarray]:
        n = lib().orc_frame_count(self._h)
        e = np.empty(n, dtype=np.float64)
        h = np.empty(n, dtype=np.float64)
        lib().orc_get_frames(self._h, _dp(e), _dp(h))
        return e, h

    def max_abs_ex(self) -> float:
        return lib().orc_max_abs_ex(self._h)

    def media(self) -> tuple[np.ndarray, np.ndarray]:
        """(wpe, nue): the extruded plasma and collision frequencies."""
        wpe = np.empty(self.shape, dtype=np.float64)
        nue = np.empty(self.shape, dtype=np.float64)
        lib().orc_get_media(self._h, _dp(wpe), _dp(nue))
        return wpe, nue

    def cpml_tables(self):
        L = self.p.get("pmllen") or 10
        arrs = [np.empty(2 * L) for _ in range(4)]
        lib().orc_cpml_tables(self._h, *[_dp(a) for a in arrs])
        return arrs

    def source_tables(self):
        a = np.empty(self.p["tsteps"])
        b = np.empty(self.p["tsteps"])
        lib().orc_source_tables(self._h, _dp(a), _dp(b))
        return a, b

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def lk_matrices(nu: float, dt: float, wce) -> tuple[np.ndarray, np.ndarray]:
    A = np.empty(9)
    K = np.empty(9)
    w = np.asarray(wce, dtype=np.float64)
    lib().orc_lk_matrices(float(nu), float(dt), _dp(w), _dp(A), _dp(K))
    return A.reshape(3, 3), K.reshape(3, 3)


def run_like_reference(p: dict, wp: np.ndarray, nu: np.ndarray) -> dict:
    """Whole solve with the reference's output cadence; returns the payloads
    the reference writes (NTF payload, probes, E/H frames, log maxima)."""
    o = Oracle(p, wp, nu)
    T = p["tsteps"]
    cad = T // p["numfiles"]
    ef, hf, emax = [], [], []
    for t in range(T):
        o.step(1)
        if t % cad == 0:
            emax.append(o.max_abs_ex())
            if p["savefields"]:
                e, h = o.frames()
                ef.append(e)
                hf.append(h)
    out = {"ntff": o.ntff(), "probes": o.probes(), "emax": np.array(emax),
           "eframes": np.array(ef), "hframes": np.array(hf)}
    o.close()
    return out
