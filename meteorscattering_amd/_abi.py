"""ctypes mirror of include/msc.h and the loader for the native library.

The product library is ``meteorscattering_amd/libmsc.so`` (HIP kernels for
gfx950 plus the C ABI).  Loading fails loudly when it has not been built:
there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

MSC_MAX_PROBES = 64
MSC_MAX_NTF = 256

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmsc.so")

# field component ids (msc.h)
EX, EY, EZ, HX, HY, HZ, JX, JY, JZ = range(9)
COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Jx", "Jy", "Jz")


class MscParams(C.Structure):
    """``msc_params``: inputs.dat fields in file order, B0.dat, extension record."""

    _fields_ = [
        ("dopml", C.c_int32),
        ("savefields", C.c_int32),
        ("f0", C.c_double),
        ("fsig", C.c_double),
        ("sourcetype", C.c_int32),
        ("sourcepol", C.c_int32),
        ("gridfactor", C.c_double),
        ("dt", C.c_double),
        ("tsteps", C.c_int32),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("metsize", C.c_double * 2),
        ("wpmax", C.c_double),
        ("numax", C.c_double),
        ("metloc", C.c_double * 3),
        ("metangle", C.c_double),
        ("mettype", C.c_int32),
        ("numfiles", C.c_int32),
        ("nprobes", C.c_int32),
        ("probex", C.c_int32 * MSC_MAX_PROBES),
        ("probey", C.c_int32 * MSC_MAX_PROBES),
        ("probez", C.c_int32 * MSC_MAX_PROBES),
        ("nntf", C.c_int32),
        ("fntf", C.c_double * MSC_MAX_NTF),
        ("doplate", C.c_int32),
        ("platesize", C.c_double),
        ("doB0", C.c_int32),
        ("B0", C.c_double * 3),
        ("pmllen", C.c_int32),
        ("media3d", C.c_int32),
        ("nu_scale_height", C.c_double),
        ("src_tau", C.c_double),
        ("j_collocated", C.c_int32),
    ]


def params_to_struct(p: dict) -> MscParams:
    s = MscParams()
    for name, ctype in MscParams._fields_:
        if name not in p:
            continue
        v = p[name]
        if hasattr(ctype, "_length_"):
            arr = getattr(s, name)
            vals = list(np.atleast_1d(v))
            if len(vals) > ctype._length_:
                raise ValueError(f"{name}: {len(vals)} values exceed the ABI limit {ctype._length_}")
            for n, x in enumerate(vals):
                arr[n] = x
        else:
            setattr(s, name, v)
    return s


def struct_to_params(s: MscParams) -> dict:
    p = {}
    for name, ctype in MscParams._fields_:
        v = getattr(s, name)
        if hasattr(ctype, "_length_"):
            v = list(v)
        p[name] = v
    npb, nf = p["nprobes"], p["nntf"]
    for k in ("probex", "probey", "probez"):
        p[k] = [int(x) for x in p[k][:npb]]
    p["fntf"] = [float(x) for x in p["fntf"][:nf]]
    return p


_lib = None


def load_library(path: str | None = None) -> C.CDLL:
    """Load libmsc.so; raise if it is missing (no silent fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"native library {path} is not built; run __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950) first")
    lib = C.CDLL(path)
    _declare(lib)
    if path == LIB_PATH:
        _lib = lib
    return lib


# request states (msc.h)
REQ_QUEUED, REQ_READY, REQ_RUNNING, REQ_DRAINING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(7)
REQ_STATES = ("queued", "ready", "running", "draining", "done", "failed", "cancelled")


class RequestInfo(C.Structure):
    """``msc_request_info``."""

    _fields_ = [
        ("state", C.c_int32),
        ("t", C.c_int32),
        ("tsteps", C.c_int32),
        ("device_bytes", C.c_size_t),
        ("t_submit", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
    ]


class Tiling(C.Structure):
    """``msc_tiling``."""

    _fields_ = [
        ("ty", C.c_int32),
        ("ty_b", C.c_int32),
        ("chunk", C.c_int32),
        ("chunk2", C.c_int32),
    ]


def tiling_struct(t: dict) -> Tiling:
    return Tiling(int(t["ty"]), int(t["ty_b"]), int(t["chunk"]), int(t["chunk2"]))


def tiling_dict(t: Tiling) -> dict:
    return {f: int(getattr(t, f)) for f, _ in t._fields_}


class Timing(C.Structure):
    """``msc_timing``."""

    _fields_ = [
        ("steps", C.c_int32),
        ("pair_launches", C.c_int32),
        ("step_ms", C.c_double),
        ("pair_ms", C.c_double),
        ("pair_cells", C.c_int64),
        ("span_ms", C.c_double),
        ("pair_j_cells", C.c_int64),
        ("step_j_cells", C.c_int64),
    ]


class ServerStats(C.Structure):
    """``msc_server_stats``."""

    _fields_ = [
        ("queued", C.c_int32),
        ("ready", C.c_int32),
        ("running", C.c_int32),
        ("draining", C.c_int32),
        ("arena_bytes", C.c_size_t),
        ("arena_used", C.c_size_t),
        ("iterations", C.c_longlong),
        ("solve_steps", C.c_longlong),
    ]


class LayoutReport(C.Structure):
    """``msc_layout_report``."""

    _fields_ = [
        ("cells", C.c_longlong),
        ("one_step_tiles", C.c_longlong),
        ("outer_tiles", C.c_longlong),
        ("pair_tiles", C.c_longlong),
        ("region_cells", C.c_longlong),
        ("max_list_tiles", C.c_longlong),
    ]


def _declare(lib: C.CDLL) -> None:
    P = C.POINTER
    vp = C.c_void_p
    dp = P(C.c_double)
    sig = {
        "msc_params_default": (None, [P(MscParams)]),
        "msc_params_validate": (C.c_int, [P(MscParams)]),
        "msc_last_error": (C.c_char_p, []),
        "msc_version": (C.c_char_p, []),
        "msc_read_run_dir": (C.c_int, [C.c_char_p, P(MscParams), P(dp), P(dp)]),
        "msc_write_run_dir": (C.c_int, [C.c_char_p, P(MscParams), dp, dp]),
        "msc_free": (None, [vp]),
        "msc_solver_create": (C.c_int, [P(MscParams), dp, dp, C.c_int, P(vp)]),
        "msc_tiling_default": (None, [P(Tiling)]),
        "msc_solver_create_tiled": (C.c_int, [P(MscParams), dp, dp, C.c_int, P(Tiling), P(vp)]),
        "msc_autotune": (C.c_int, [P(MscParams), dp, dp, C.c_int, C.c_int, P(Tiling), dp]),
        "msc_solver_step": (C.c_int, [vp, C.c_int]),
        "msc_solver_sync": (C.c_int, [vp]),
        "msc_solver_time": (C.c_int, [vp]),
        "msc_solver_get_field": (C.c_int, [vp, C.c_int, dp]),
        "msc_solver_get_field_noalloc": (C.c_int, [vp, C.c_int, dp, C.c_size_t]),
        "msc_solver_ntff_count": (C.c_size_t, [vp]),
        "msc_solver_get_ntff": (C.c_int, [vp, dp, C.c_size_t]),
        "msc_solver_flush_ntff": (C.c_int, [vp]),
        "msc_solver_get_probes": (C.c_int, [vp, dp, C.c_size_t]),
        "msc_solver_nonfinite": (C.c_int, [vp, P(C.c_int)]),
        "msc_solver_frame_count": (C.c_size_t, [vp]),
        "msc_solver_nframes": (C.c_int, [vp]),
        "msc_solver_get_frame": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "msc_solver_max_abs_ex": (C.c_int, [vp, dp]),
        "msc_solver_device_bytes": (C.c_size_t, [vp]),
        "msc_solver_set_timing": (C.c_int, [vp, C.c_int]),
        "msc_solver_kernel_ms": (C.c_double, [vp]),
        "msc_solver_get_timing": (C.c_int, [vp, P(Timing)]),
        "msc_solver_destroy": (None, [vp]),
        "msc_run_dir": (C.c_int, [C.c_char_p, C.c_int]),
        "msc_stream_copy": (C.c_int, [C.c_size_t, C.c_int, dp]),
        "msc_solver_pattern_probe": (C.c_int, [vp, C.c_int, dp]),
        "msc_solve_footprint": (C.c_size_t, [P(MscParams), C.c_int]),
        "msc_layout_check": (C.c_int, [P(MscParams), C.c_int, C.c_int, P(Tiling), P(LayoutReport)]),
        "msc_server_create": (C.c_int, [C.c_int, C.c_size_t, P(vp)]),
        "msc_server_submit": (C.c_int, [vp, P(MscParams), dp, dp, P(C.c_int64)]),
        "msc_server_advance": (C.c_int, [vp, C.c_int, P(C.c_int)]),
        "msc_server_info": (C.c_int, [vp, C.c_int64, P(RequestInfo)]),
        "msc_server_get_ntff": (C.c_int, [vp, C.c_int64, dp, C.c_size_t]),
        "msc_server_get_probes": (C.c_int, [vp, C.c_int64, dp, C.c_size_t]),
        "msc_server_release": (C.c_int, [vp, C.c_int64]),
        "msc_server_cancel": (C.c_int, [vp, C.c_int64]),
        "msc_server_set_time_limit": (C.c_int, [vp, C.c_double]),
        "msc_server_get_stats": (C.c_int, [vp, P(ServerStats)]),
        "msc_server_clock": (C.c_double, [vp]),
        "msc_server_configure": (C.c_int, [vp, C.c_int, C.c_int]),
        "msc_server_set_tilings": (C.c_int, [vp, P(Tiling), P(C.c_int64), C.c_int]),
        "msc_server_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def exported_symbols(header: str) -> list[str]:
    """Function names declared in a C header (for the export check)."""
    import re

    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(msc_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))
