"""Compact, comparison-ready summaries of solver outputs.

A fixture keeps some arrays in full (probes, incident lines, the NTF face
regions analysis/meteorNTF.m carves) and, for every output array, seeded
random projections: sum(w_r * a) for 4 weight vectors w_r ~ U(-1, 1), plus
sum(|a|) and max|a|.  A single wrong element of size d moves a projection
by about d, so a mismatch anywhere in an array is caught.
"""
from __future__ import annotations

import zlib

import numpy as np

NPROJ = 4


def weights(name: str, n: int) -> np.ndarray:
    seed = zlib.crc32(name.encode()) & 0xFFFFFFFF
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=(NPROJ, n))


def project(name: str, a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64).ravel()
    w = weights(name, a.size)
    return np.concatenate([w @ a, [np.abs(a).sum(), np.abs(a).max() if a.size else 0.0]])


def compare_projection(name: str, a: np.ndarray, ref: np.ndarray, rtol: float) -> float:
    """Max projection error relative to sum|ref|; returns it."""
    got = project(name, a)
    scale = max(ref[NPROJ], 1e-300)
    return float(np.max(np.abs(got - ref)) / scale)


def carve_ranges(nx: int, ny: int, nz: int, ntfb: int):
    """Index ranges meteorNTF.m:64-196 keeps (0-based, inclusive start,
    exclusive stop, per-axis extents): ntfb .. N-ntfb (+1)."""
    return {"x": (ntfb, nx - ntfb + 1), "y": (ntfb, ny - ntfb + 1), "z": (ntfb, nz - ntfb + 1)}
