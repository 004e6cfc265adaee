"""Small cubic run configurations used for golden fixtures and parity tests.

Cubic grids only: the reference fills z[k] for k < Nx only
(meteor-fork.cpp:137-141), so only cubic grids have defined media there.
Every case sets numfiles so the cadence tsteps/numfiles is valid.
"""
from __future__ import annotations

import math

import numpy as np

QE = 1.602177e-19
ME = 9.10939e-31
E0 = 8.85419e-12


def plasma_freq(n0: float) -> float:
    return math.sqrt(n0 * QE * QE / (E0 * ME))


def base(n: int = 44, tsteps: int = 160) -> dict:
    f0 = 50e6
    gridfactor = 12.0  # dx = 0.5 m
    dx = 3e8 / f0 / gridfactor
    dt = 0.99 * dx / (3e8 * math.sqrt(3))
    c = (n + 1) // 2
    return dict(
        dopml=1, savefields=1, f0=f0, fsig=12.0, sourcetype=3, sourcepol=1,
        gridfactor=gridfactor, dt=dt, tsteps=tsteps, nx=n, ny=n, nz=n,
        metsize=[4.0, 4.0], wpmax=0.5, numax=0.0, metloc=[0.0, 0.0, 1.0],
        metangle=0.0, mettype=0, numfiles=4, nprobes=2,
        probex=[c, c - 3], probey=[c, c + 2], probez=[c, c - 1],
        nntf=2, fntf=[45e6, 55e6], doplate=0, platesize=1.0, doB0=1,
        B0=[5e-5 * math.sin(math.radians(60)), 0.0, 5e-5 * math.cos(math.radians(60))],
        pmllen=0, media3d=0, nu_scale_height=0.0, src_tau=0.0,
    )


def trail_maps(p: dict, n0: float = 1e13, r0: float = 1.5, nu0: float = 2e6,
               centre=(0.0, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    """Gaussian radial trail along z: n(r) = n0 exp(-r^2/r0^2), so
    wp(r) = wp0 exp(-r^2/(2 r0^2)); uniform collision frequency.  Maps are in
    the solver's [nx][ny] index semantics (grid coordinates of
    meteor-fork.cpp:137-141)."""
    dx = 3e8 / p["f0"] / p["gridfactor"]
    nx, ny = p["nx"], p["ny"]
    x = -((nx - 1) // 2) * dx + dx * np.arange(nx)
    y = -((ny - 1) // 2) * dx + dx * np.arange(ny)
    r2 = (x[:, None] - centre[0]) ** 2 + (y[None, :] - centre[1]) ** 2
    wp = plasma_freq(n0) * np.exp(-r2 / (2 * r0 * r0))
    nu = np.full((nx, ny), nu0)
    return wp, nu


def case(name: str) -> tuple[dict, np.ndarray, np.ndarray]:
    p = base()
    if name == "lk_ricker_ez":
        wp, nu = trail_maps(p)
    elif name == "di_modgauss_ey":
        p.update(doB0=0, sourcetype=2, sourcepol=2, nntf=1, fntf=[50e6])
        wp, nu = trail_maps(p, n0=3e13, nu0=5e6)
    elif name == "plate_ricker_45":
        p.update(doplate=1, platesize=4.0, sourcepol=3, doB0=0, nntf=1, fntf=[50e6])
        wp, nu = trail_maps(p)
    elif name == "sphere_sin_rhcp":
        p.update(doplate=2, platesize=3.0, sourcetype=1, sourcepol=4, doB0=0, numfiles=5)
        wp, nu = trail_maps(p)
    elif name == "lk_nopml_lhcp":
        p.update(dopml=0, sourcepol=5, nntf=3, fntf=[40e6, 50e6, 60e6])
        wp, nu = trail_maps(p, n0=5e12, r0=2.0, nu0=1e5, centre=(0.5, -0.5))
    elif name == "lk_overdense_offset":
        p = base(n=46, tsteps=150)
        p.update(sourcepol=1, metloc=[0.0, 0.0, -2.0], metsize=[3.0, 3.0])
        wp, nu = trail_maps(p, n0=2e14, r0=1.0, nu0=1e7, centre=(0.5, 1.0))
    else:
        raise KeyError(name)
    return p, wp, nu


CASES = ["lk_ricker_ez", "di_modgauss_ey", "plate_ricker_45", "sphere_sin_rhcp",
         "lk_nopml_lhcp", "lk_overdense_offset"]
