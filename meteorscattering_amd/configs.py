"""Synthetic workloads: Gaussian meteor trails on the reference's grid.

Each builder returns (params, wp, nu) in the solver's conventions (see
runio.write_run_dir).  The trail has a Gaussian radial electron density
n(r) = n0 exp(-r^2/r0^2), i.e. plasma frequency wp(r) = wp0 exp(-r^2/(2 r0^2))
with wp0 = sqrt(n0 q^2 / (eps0 m_e)).  Incidence "theta off broadside" is
realised by rotating the trail axis and B by theta in the x-z plane while
the plane wave keeps travelling along +x, so the reference's TF/SF and
NTFF geometry is unchanged (the media become 3-D, media3d=1).
"""
from __future__ import annotations

import math

import numpy as np

QE = 1.602177e-19
ME = 9.10939e-31
E0 = 8.85419e-12
C0 = 3.0e8


def plasma_frequency(n0: float) -> float:
    return math.sqrt(n0 * QE * QE / (E0 * ME))


def b_field(angle_from_z_deg: float, magnitude: float = 5e-5, tilt_deg: float = 0.0) -> list[float]:
    """B at `angle_from_z_deg` from +z in the x-z plane, then rotated by
    `tilt_deg` about y (the same rotation applied to the trail)."""
    a = math.radians(angle_from_z_deg + tilt_deg)
    return [magnitude * math.sin(a), 0.0, magnitude * math.cos(a)]


def base_params(n: tuple[int, int, int], dx: float, tsteps: int, f0: float = 50e6, pmllen: int = 10,
                courant: float = 0.99) -> dict:
    nx, ny, nz = n
    gridfactor = (C0 / f0) / dx
    dt = courant * dx / (C0 * math.sqrt(3))
    c = ((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2)
    return dict(
        dopml=1, savefields=1, f0=f0, fsig=12.0, sourcetype=3, sourcepol=1, gridfactor=gridfactor, dt=dt,
        tsteps=tsteps, nx=nx, ny=ny, nz=nz, metsize=[1e30, 1e30], wpmax=0.0, numax=0.0,
        metloc=[0.0, 0.0, 0.0], metangle=0.0, mettype=0, numfiles=10, nprobes=1,
        probex=[c[0]], probey=[c[1]], probez=[c[2]], nntf=1, fntf=[f0], doplate=0, platesize=1.0,
        doB0=1, B0=b_field(60.0), pmllen=pmllen, media3d=0, nu_scale_height=0.0, src_tau=0.0,
    )


def axis(n: int, dx: float) -> np.ndarray:
    """Grid coordinates, meteor-fork.cpp:137-141 (integer (N-1)/2)."""
    return -((n - 1) // 2) * dx + dx * np.arange(n)


def trail_2d(p: dict, n0: float, r0: float, nu0: float) -> tuple[np.ndarray, np.ndarray]:
    """Trail along z: 2-D maps extruded by the solver (metsize[0] huge = no taper)."""
    dx = (C0 / p["f0"]) / p["gridfactor"]
    x, y = axis(p["nx"], dx), axis(p["ny"], dx)
    r2 = x[:, None] ** 2 + y[None, :] ** 2
    wp = plasma_frequency(n0) * np.exp(-r2 / (2 * r0 * r0))
    return wp, np.full_like(wp, nu0)


def trail_3d(p: dict, n0: float, r0: float, nu0: float, tilt_deg: float) -> tuple[np.ndarray, np.ndarray]:
    """Trail through the origin along (sin t, 0, cos t): full 3-D media."""
    dx = (C0 / p["f0"]) / p["gridfactor"]
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    x, y, z = axis(nx, dx), axis(ny, dx), axis(nz, dx)
    t = math.radians(tilt_deg)
    ax, az = math.sin(t), math.cos(t)
    wp0 = plasma_frequency(n0)
    wp = np.empty((nx, ny, nz), dtype=np.float64)
    zz = z[None, :]
    yy2 = (y ** 2)[:, None]
    for i in range(nx):
        along = x[i] * ax + zz * az
        r2 = x[i] ** 2 + yy2 + zz ** 2 - along ** 2
        wp[i] = wp0 * np.exp(-r2 / (2 * r0 * r0))
    nu = np.full((nx, ny, nz), nu0, dtype=np.float64)
    return wp, nu


def trail_96x96x192(tsteps: int = 2000) -> tuple[dict, np.ndarray, np.ndarray]:
    """96x96x192, dx 0.5 m, 10-cell CPML, Courant 0.99/sqrt(3), Ricker at
    50 MHz (nearest reference waveform to a Gaussian derivative), Ez, n0 1e12,
    r0 1.5 m, nu 1e5, B 5e-5 T at 60 deg from z, single-frequency NTFF."""
    p = base_params((96, 96, 192), 0.5, tsteps)
    wp, nu = trail_2d(p, 1e12, 1.5, 1e5)
    return p, wp, nu


def trail_400x400x800(seed: int = 0, tsteps: int = 6000, media3d: bool = True) -> tuple[dict, np.ndarray, np.ndarray]:
    """400x400x800 (128M cells), dx 0.25 m, 12-cell CPML, 6000 steps; n0
    log-uniform in [1e11,1e14], r0 uniform in [0.5,4] m from `seed`; nu
    decays along z with a 7 km scale height; 30 deg off broadside."""
    rng = np.random.default_rng(seed)
    n0 = 10 ** rng.uniform(11, 14)
    r0 = rng.uniform(0.5, 4.0)
    p = base_params((400, 400, 800), 0.25, tsteps, pmllen=12)
    p.update(nu_scale_height=7000.0, B0=b_field(60.0, tilt_deg=30.0), j_collocated=1)
    p["_trail"] = dict(n0=n0, r0=r0, nu0=1e5, tilt_deg=30.0)
    if media3d:
        p["media3d"] = 1
        wp, nu = trail_3d(p, n0, r0, 1e5, 30.0)
    else:
        wp, nu = trail_2d(p, n0, r0, 1e5)
    return p, wp, nu


def ntff_m64_96x96x192(tsteps: int = 2000) -> tuple[dict, np.ndarray, np.ndarray]:
    """The 96x96x192 trail with 64 NTFF frequencies from 30 to 80 MHz."""
    p, wp, nu = trail_96x96x192(tsteps)
    p.update(nntf=64, fntf=list(np.linspace(30e6, 80e6, 64)))
    return p, wp, nu


def random_trail(rng: np.random.Generator, nx: int, ny: int, nz: int, dx: float = 0.5,
                 steps_range=(1500, 4000), nntf: int = 1, maps: bool = True) -> tuple[dict, np.ndarray, np.ndarray]:
    """One trail of the sweep and the request stream: n0 log-uniform [1e10,1e15], r0 U[0.3,5] m,
    nu log-uniform [1e3,1e7], B angle U[0,90] deg, incidence U[0,60] deg.
    With maps=False the media are not built (wp, nu are None; trail_media(p)
    builds them later): the draws, and so every later trail, are the same."""
    n0 = 10 ** rng.uniform(10, 15)
    r0 = rng.uniform(0.3, 5.0)
    nu0 = 10 ** rng.uniform(3, 7)
    bang = rng.uniform(0, 90)
    inc = rng.uniform(0, 60)
    ncell = nx * ny * nz
    # more steps for bigger grids, as the reference's tfactor rule scales with Nx
    frac = (math.log(ncell) - math.log(64 ** 3)) / (math.log(256 * 256 * 512) - math.log(64 ** 3))
    steps = int(round(steps_range[0] + min(max(frac, 0.0), 1.0) * (steps_range[1] - steps_range[0])))
    # dt as setup/meteorrun.m:33-34: Courant 0.8 times the plasma factor below
    p = base_params((nx, ny, nz), dx, steps, courant=0.8)
    # j_collocated: the reference's node-averaged Lee-Kalluri drive grows
    # without bound (see msc_params.j_collocated); these runs use the fix
    p.update(B0=b_field(bang, tilt_deg=inc), media3d=1, numfiles=1, savefields=0, j_collocated=1,
             nntf=nntf, fntf=list(np.linspace(40e6, 60e6, nntf)) if nntf > 1 else [50e6])
    # keep the explicit plasma coupling stable: the reference's own rule
    # (setup/meteorrun.m:33-34) shrinks dt with the peak plasma frequency
    wpm = plasma_frequency(n0) / (2 * math.pi * p["f0"])
    stab = 1 / math.sqrt(1 + (wpm * math.pi / p["gridfactor"] / 2) ** 2)
    p["dt"] *= stab
    p["_trail"] = dict(n0=n0, r0=r0, nu0=nu0, tilt_deg=inc, b_angle=bang)
    if not maps:
        return p, None, None
    wp, nu = trail_media(p)
    return p, wp, nu


def trail_media(p: dict) -> tuple[np.ndarray, np.ndarray]:
    """The 3-D media of a random trail from its drawn parameters."""
    t = p["_trail"]
    return trail_3d(p, t["n0"], t["r0"], t["nu0"], t["tilt_deg"])


def trail_sweep(seed: int = 0, count: int = 48, maps: bool = True):
    """A sweep of random trails (48 by default), nx,ny ~ U{64..256}, nz ~ U{128..512}, dx 0.5 m."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny = (int(v) for v in rng.integers(64, 257, size=2))
        nz = int(rng.integers(128, 513))
        out.append(random_trail(rng, nx, ny, nz, maps=maps))
    return out


def request_stream(seed: int = 0, count: int = 64, nntf: int = 16, maps: bool = True):
    """Serving requests: the sweep's physics, grids log-uniform in cell count
    from 64^3 to 256x256x1024 (nx = ny = 64*4^s, nz = 64*16^s, s ~ U[0,1]),
    16-frequency NTFF (40-60 MHz)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        s = rng.uniform(0.0, 1.0)
        nxy = int(round(64 * 4 ** s))
        nz = int(round(64 * 16 ** s))
        out.append(random_trail(rng, nxy, nxy, nz, nntf=nntf, maps=maps))
    return out
