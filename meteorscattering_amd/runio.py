"""Run-directory files of the meteor FDTD solver, written and read in numpy.

Input side: the counterpart of setup/meteorrun.m:22-114 (grid and time-step
rule, inputs.dat, B0.dat, wp.dat, nu.dat).  Output side: readers for the
files the solver writes (meteor-fork.cpp:1616-1738) as analysis/meteorNTF.m
and analysis/plotMresults.m read them.

All byte layouts are native little-endian int32 / float64.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

# analysis/loadconstants.m:3-10 and setup/loadconstants.m:3-11
U0_MATLAB = 4 * math.pi * 1e-7
E0_MATLAB = 8.845e-12
VP_MATLAB = math.sqrt(1 / (U0_MATLAB * E0_MATLAB))
ETA0_MATLAB = math.sqrt(U0_MATLAB / E0_MATLAB)
# constants of the C solver, meteor-fork.cpp:12-19
QE = 1.602177e-19
E0 = 8.85419e-12
U0 = 1.2566e-6
ME = 9.10939e-31
C_LIGHT = 3.0e8

EXT_MAGIC = b"MSCX"
EXT_VERSION = 2  # v1 lacks j_collocated


def default_params() -> dict:
    """setup/meteordefaults.m:1-92 as a parameter dict (plus doB0/B0)."""
    nx = ny = nz = 201
    f0 = 150e6
    return dict(
        dopml=1, savefields=1, f0=f0, fsig=200.0, sourcetype=2, sourcepol=3,
        gridfactor=200.0, dt=0.0, tsteps=0, nx=nx, ny=ny, nz=nz,
        metsize=[0.5, 0.5], wpmax=1.0, numax=1e7, metloc=[0.0, 0.0, 0.0],
        metangle=0.0, mettype=0, numfiles=100, nprobes=1,
        probex=[(nx + 1) // 2], probey=[(ny + 1) // 2], probez=[(nz + 1) // 2],
        nntf=15, fntf=list(np.linspace(0.9, 1.1, 15) * f0), doplate=0, platesize=1.0,
        doB0=0, B0=[0.0, 0.0, 5e-5],
        pmllen=0, media3d=0, nu_scale_height=0.0, src_tau=0.0, j_collocated=0,
    )


def grid_spacing(p: dict) -> float:
    """dx = (c/f0)/gridfactor, meteor-fork.cpp:119-132 (c = 3e8)."""
    return (C_LIGHT / p["f0"]) / p["gridfactor"]


def meteorrun_dt(dx: float, wpmax: float, gridfactor: float) -> float:
    """Time step rule of setup/meteorrun.m:33-34."""
    stab = 1 / math.sqrt(1 + (wpmax * math.pi / gridfactor / 2) ** 2)
    return dx / VP_MATLAB / math.sqrt(3) * stab * 0.8


def meteorrun_tsteps(nx: int, dx: float, dt: float, tfactor: float) -> int:
    """setup/meteorrun.m:35."""
    return int(round(nx * dx / VP_MATLAB / dt * tfactor))


def ext_is_default(p: dict) -> bool:
    return (int(p.get("pmllen", 0)) in (0, 10) and not p.get("media3d", 0)
            and not p.get("nu_scale_height", 0.0) and not p.get("src_tau", 0.0)
            and not p.get("j_collocated", 0))


def encode_inputs_dat(p: dict, with_doB0: bool = True, with_ext: bool | None = None) -> bytes:
    """inputs.dat bytes: the order of setup/meteorrun.m:45-69 (= the reader
    meteor-fork.cpp:84-115), then doB0 (the fork reads it; meteorrun.m never
    wrote it), then this build's optional extension record."""
    out = bytearray()
    i32 = lambda v: out.extend(struct.pack("<i", int(v)))
    f64 = lambda v: out.extend(struct.pack("<d", float(v)))
    i32(p["dopml"]); i32(p["savefields"]); f64(p["f0"]); f64(p["fsig"])
    i32(p["sourcetype"]); i32(p["sourcepol"]); f64(p["gridfactor"]); f64(p["dt"])
    i32(p["tsteps"]); i32(p["nx"]); i32(p["ny"]); i32(p["nz"])
    for v in p["metsize"][:2]:
        f64(v)
    f64(p["wpmax"]); f64(p["numax"])
    for v in p["metloc"][:3]:
        f64(v)
    f64(p["metangle"]); i32(p["mettype"]); i32(p["numfiles"]); i32(p["nprobes"])
    for key in ("probex", "probey", "probez"):
        vals = list(p[key])[: p["nprobes"]]
        assert len(vals) == p["nprobes"], key
        for v in vals:
            i32(v)
    i32(p["nntf"])
    fn = list(p["fntf"])[: p["nntf"]]
    assert len(fn) == p["nntf"]
    for v in fn:
        f64(v)
    i32(p["doplate"]); f64(p["platesize"])
    if with_doB0:
        i32(p.get("doB0", 0))
        if with_ext is None:
            with_ext = not ext_is_default(p)
        if with_ext:
            out.extend(EXT_MAGIC)
            i32(EXT_VERSION); i32(p.get("pmllen", 0)); i32(p.get("media3d", 0))
            f64(p.get("nu_scale_height", 0.0)); f64(p.get("src_tau", 0.0))
            i32(p.get("j_collocated", 0))
    return bytes(out)


def decode_inputs_dat(buf: bytes) -> dict:
    """Inverse of encode_inputs_dat; a missing doB0 reads as 0."""
    p = default_params()
    off = 0

    def take(fmt):
        nonlocal off
        n = struct.calcsize(fmt)
        if off + n > len(buf):
            raise EOFError(f"inputs.dat truncated at byte {off}")
        v = struct.unpack_from(fmt, buf, off)
        off += n
        return v

    p["dopml"], = take("<i"); p["savefields"], = take("<i"); p["f0"], = take("<d"); p["fsig"], = take("<d")
    p["sourcetype"], = take("<i"); p["sourcepol"], = take("<i"); p["gridfactor"], = take("<d"); p["dt"], = take("<d")
    p["tsteps"], = take("<i"); p["nx"], p["ny"], p["nz"] = take("<3i")
    p["metsize"] = list(take("<2d")); p["wpmax"], = take("<d"); p["numax"], = take("<d")
    p["metloc"] = list(take("<3d")); p["metangle"], = take("<d"); p["mettype"], = take("<i")
    p["numfiles"], = take("<i"); p["nprobes"], = take("<i")
    n = p["nprobes"]
    p["probex"] = list(take(f"<{n}i")); p["probey"] = list(take(f"<{n}i")); p["probez"] = list(take(f"<{n}i"))
    p["nntf"], = take("<i")
    p["fntf"] = list(take(f"<{p['nntf']}d"))
    p["doplate"], = take("<i"); p["platesize"], = take("<d")
    p["doB0"] = 0
    p.update(pmllen=0, media3d=0, nu_scale_height=0.0, src_tau=0.0, j_collocated=0)
    if off + 4 <= len(buf):
        p["doB0"], = take("<i")
        if buf[off:off + 4] == EXT_MAGIC:
            off += 4
            ver, = take("<i")
            if ver not in (1, EXT_VERSION):
                raise ValueError(f"unknown extension record version {ver}")
            p["pmllen"], p["media3d"] = take("<2i")
            p["nu_scale_height"], p["src_tau"] = take("<2d")
            if ver >= 2:
                p["j_collocated"], = take("<i")
    return p


def write_run_dir(d: str, p: dict, wp: np.ndarray, nu: np.ndarray) -> None:
    """Write inputs.dat, B0.dat, wp.dat and nu.dat.

    ``wp``/``nu`` are given in the solver's index semantics: the C reader
    fills ``wp2[i][j]`` row-major from the file (meteor-fork.cpp:340-347), so
    the file is ``wp.ravel(order='C')``.  (MATLAB's fwrite of wp(i,j) is
    column-major, so the solver's wp2[i][j] is MATLAB's wp(j+1,i+1) when
    Nx == Ny; use :func:`matlab_map_to_solver` for maps built the MATLAB way.)
    """
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "inputs.dat"), "wb") as f:
        f.write(encode_inputs_dat(p))
    with open(os.path.join(d, "B0.dat"), "wb") as f:
        f.write(np.asarray(p.get("B0", [0, 0, 0]), dtype="<f8").tobytes())
    shape = (p["nx"], p["ny"], p["nz"]) if p.get("media3d") else (p["nx"], p["ny"])
    for name, arr in (("wp.dat", wp), ("nu.dat", nu)):
        a = np.asarray(arr, dtype="<f8")
        if a.shape != shape:
            raise ValueError(f"{name}: shape {a.shape} != {shape}")
        with open(os.path.join(d, name), "wb") as f:
            f.write(np.ascontiguousarray(a).tobytes())


def read_run_dir(d: str) -> tuple[dict, np.ndarray, np.ndarray]:
    with open(os.path.join(d, "inputs.dat"), "rb") as f:
        p = decode_inputs_dat(f.read())
    p["B0"] = list(np.fromfile(os.path.join(d, "B0.dat"), dtype="<f8", count=3))
    shape = (p["nx"], p["ny"], p["nz"]) if p["media3d"] else (p["nx"], p["ny"])
    n = int(np.prod(shape))
    wp = np.fromfile(os.path.join(d, "wp.dat"), dtype="<f8", count=n).reshape(shape)
    nu = np.fromfile(os.path.join(d, "nu.dat"), dtype="<f8", count=n).reshape(shape)
    return p, wp, nu


def matlab_map_to_solver(wp_matlab: np.ndarray) -> np.ndarray:
    """Bytes MATLAB's fwrite(wp) produces, viewed as the solver's wp2[Nx][Ny]."""
    nx, ny = wp_matlab.shape
    return np.asarray(wp_matlab, dtype=np.float64).ravel(order="F").reshape(nx, ny)


def meteorrun_wp_map(p: dict) -> np.ndarray:
    """MATLAB wp(i,j) of setup/meteorrun.m:83-96 (rad/s), before imrotate.

    metangle != 0 (imrotate, :99) is not reproduced: the MATLAB builtin is
    absent here, so such maps are parity unpinned and rejected.
    """
    if p.get("metangle", 0.0):
        raise NotImplementedError("metangle != 0 needs MATLAB imrotate (not available)")
    dx = grid_spacing(p)
    nx, ny = p["nx"], p["ny"]
    x = -(nx - 1) / 2 * dx + dx * np.arange(nx)
    y = -(ny - 1) / 2 * dx + dx * np.arange(ny)
    ms = p["metsize"]
    ml = p["metloc"]
    sx = np.where(x < ml[0], ms[0], ms[1])
    w0 = p["wpmax"] * (2 * math.pi * p["f0"])
    return w0 * np.exp(-((x[:, None] - ml[0]) ** 2) / sx[:, None] ** 2) * np.exp(-((y[None, :] - ml[1]) ** 2) / ms[0] ** 2)


# ---------------------------------------------------------------- outputs

def ntf_face_names() -> list[str]:
    return ["My_front", "Mz_front", "Jy_front", "Jz_front", "My_back", "Mz_back", "Jy_back", "Jz_back",
            "Mx_top", "Mz_top", "Jx_top", "Jz_top", "Mx_bot", "Mz_bot", "Jx_bot", "Jz_bot",
            "Mx_left", "My_left", "Jx_left", "Jy_left", "Mx_right", "My_right", "Jx_right", "Jy_right"]


def parse_ntf(buf: bytes, nx: int, ny: int, nz: int) -> dict:
    """output_NTF.dat (meteor-fork.cpp:1677-1714; reader meteorNTF.m:12-57).

    Face arrays come back shaped [n1][n2][2M] as the solver stored them:
    front/back [Ny][Nz], top/bottom [Nx][Nz], left/right [Nx][Ny]."""
    nntf, ntfb = struct.unpack_from("<2i", buf, 0)
    off = 8
    M2 = 2 * nntf
    rd = lambda n: np.frombuffer(buf, dtype="<f8", count=n, offset=off)
    out = {"nntf": nntf, "ntfb": ntfb}
    out["fntf"] = rd(nntf).copy(); off += 8 * nntf
    out["Eydft"] = rd(M2).copy(); off += 8 * M2
    out["Ezdft"] = rd(M2).copy(); off += 8 * M2
    names = ntf_face_names()
    shapes = [(ny, nz)] * 8 + [(nx, nz)] * 8 + [(nx, ny)] * 8
    for name, (a, b) in zip(names, shapes):
        n = a * b * M2
        out[name] = rd(n).reshape(a, b, M2).copy()
        off += 8 * n
    if off != len(buf):
        raise ValueError(f"output_NTF.dat: {len(buf) - off} trailing bytes")
    return out


def read_ntf(path: str, nx: int, ny: int, nz: int) -> dict:
    with open(path, "rb") as f:
        return parse_ntf(f.read(), nx, ny, nz)


def encode_ntf(nntf: int, ntfb: int, fntf, payload: np.ndarray) -> bytes:
    """output_NTF.dat bytes from the msc_solver_get_ntff payload."""
    return struct.pack("<2i", nntf, ntfb) + np.asarray(fntf, "<f8").tobytes() + np.asarray(payload, "<f8").tobytes()


def read_probe(path: str, tsteps: int) -> dict:
    """Probe.dat (meteor-fork.cpp:1719-1729)."""
    buf = open(path, "rb").read()
    n, = struct.unpack_from("<i", buf, 0)
    idx = np.frombuffer(buf, dtype="<i4", count=3 * n, offset=4).reshape(3, n)
    data = np.frombuffer(buf, dtype="<f8", count=6 * tsteps * n, offset=4 + 12 * n).reshape(6, tsteps, n)
    return {"nprobes": n, "index": idx.copy(), "data": data.copy()}


def frame_doubles(nx: int, ny: int, nz: int, with_inc: bool = True) -> int:
    return 2 * (nx * ny + nx * nz + ny * nz) + (2 * nx if with_inc else 0)


def read_frames(path: str, nx: int, ny: int, nz: int, with_inc: bool = True) -> np.ndarray:
    """output_E.dat / output_H.dat (with_inc) or output_J.dat: [nframes][frame]."""
    a = np.fromfile(path, dtype="<f8")
    n = frame_doubles(nx, ny, nz, with_inc)
    if a.size % n:
        raise ValueError(f"{path}: size {a.size} not a multiple of frame size {n}")
    return a.reshape(-1, n)


def split_frame(fr: np.ndarray, nx: int, ny: int, nz: int) -> dict:
    a, b, c = 2 * nx * ny, 2 * nx * nz, 2 * ny * nz
    out = {"slicez": fr[:a].reshape(2, nx, ny), "slicey": fr[a:a + b].reshape(2, nx, nz),
           "slicex": fr[a + b:a + b + c].reshape(2, ny, nz)}
    if fr.size > a + b + c:
        out["inc1"] = fr[a + b + c:a + b + c + nx]
        out["inc2"] = fr[a + b + c + nx:a + b + c + 2 * nx]
    return out
