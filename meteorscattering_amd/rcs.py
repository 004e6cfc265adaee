"""Radar cross section from the NTFF running DFTs (output_NTF.dat).

Counterpart of analysis/meteorNTF.m:3-363 plus its closed-form references
(analysis/sphereRCSvsFreq.m:3-31, the plate formula meteorNTF.m:346-350).

Samples and weights follow meteorNTF.m:64-196: on every face the tangential
axis on which a current sample sits on grid nodes is kept from ntfb to
N-ntfb inclusive with half weight at both ends, the staggered axis from
ntfb to N-ntfb-1.  The sums are complexified as i*sin_sum + cos_sum
(:199-230), i.e. phasors of an e^{-i w t} convention, so the radiation
integral carries exp(-i k r.r').  Backscatter (meteorNTF's only direction)
is r = -x, theta_hat = -z, phi_hat = -y (:307-317), sigma by Taflove eq.
8.37 (:332).  Unlike meteorNTF.m (which slices every face with Nx, so it
assumes a cubic grid) the extents are per axis, identical on cubic grids.

Yee positions (grid-index units): Ex (i+1/2, j, k), Ey (i, j+1/2, k),
Ez (i, j, k+1/2), Hx (i, j+1/2, k+1/2), Hy (i+1/2, j, k+1/2),
Hz (i+1/2, j+1/2, k); H samples are averaged across the face normal
(meteor-fork.cpp:973-1023), so every sample lies on its face plane.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import special

from . import runio

# analysis/loadconstants.m:3-10
U0 = 4 * math.pi * 1e-7
E0 = 8.845e-12
VP = math.sqrt(1 / (U0 * E0))
ETA0 = math.sqrt(U0 / E0)

# For each of the 24 arrays: (face axis, current kind, vector component,
# staggering along the face's two tangential axes: True = node axis).
# Face tangential axes: x-faces (j, k), y-faces (i, k), z-faces (i, j).
_LAYOUT = {
    "My_front": (0, "M", 1, (True, False)), "Mz_front": (0, "M", 2, (False, True)),
    "Jy_front": (0, "J", 1, (False, True)), "Jz_front": (0, "J", 2, (True, False)),
    "Mx_top": (1, "M", 0, (True, False)), "Mz_top": (1, "M", 2, (False, True)),
    "Jx_top": (1, "J", 0, (False, True)), "Jz_top": (1, "J", 2, (True, False)),
    "Mx_left": (2, "M", 0, (True, False)), "My_left": (2, "M", 1, (False, True)),
    "Jx_left": (2, "J", 0, (False, True)), "Jy_left": (2, "J", 1, (True, False)),
}


def _layout(name: str):
    base = name.replace("_back", "_front").replace("_bot", "_top").replace("_right", "_left")
    face_axis, kind, comp, nodes = _LAYOUT[base]
    high = name.endswith(("_back", "_top", "_right"))
    return face_axis, kind, comp, nodes, high


def _axis_samples(n: int, b: int, node: bool, d: float):
    """Indices, positions (m) and weights of one tangential axis."""
    if node:
        idx = np.arange(b, n - b + 1)
        w = np.ones(idx.size)
        w[0] = w[-1] = 0.5
        pos = idx * d
    else:
        idx = np.arange(b, n - b)
        w = np.ones(idx.size)
        pos = (idx + 0.5) * d
    return idx, pos, w


def surface_samples(ntf: dict, p: dict):
    """Yield (kind, component, complex samples [M, S], weights*dA [S], positions [S,3])."""
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    d = (runio.C_LIGHT / p["f0"]) / p["gridfactor"]
    b = ntf["ntfb"]
    dims = (nx, ny, nz)
    for name in runio.ntf_face_names():
        face_axis, kind, comp, nodes, high = _layout(name)
        tang = [a for a in range(3) if a != face_axis]
        (i1, p1, w1) = _axis_samples(dims[tang[0]], b, nodes[0], d)
        (i2, p2, w2) = _axis_samples(dims[tang[1]], b, nodes[1], d)
        arr = ntf[name]  # [n1][n2][2M]
        sub = arr[np.ix_(i1, i2)]  # [S1][S2][2M]
        z = 1j * sub[..., 0::2] + sub[..., 1::2]  # [S1][S2][M]
        w = np.outer(w1, w2) * d * d
        pos = np.zeros((i1.size, i2.size, 3))
        pos[..., face_axis] = (dims[face_axis] - b if high else b) * d
        pos[..., tang[0]] = p1[:, None]
        pos[..., tang[1]] = p2[None, :]
        yield kind, comp, z.reshape(-1, z.shape[-1]).T, w.ravel(), pos.reshape(-1, 3)


def _direction(theta, phi):
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    r = np.stack([st * cp, st * sp, ct], -1)
    th = np.stack([ct * cp, ct * sp, -st], -1)
    ph = np.stack([-sp, cp, np.zeros_like(sp)], -1)
    return r, th, ph


def source_phasors(ntf: dict):
    ez = 1j * ntf["Ezdft"][0::2] + ntf["Ezdft"][1::2]
    ey = 1j * ntf["Eydft"][0::2] + ntf["Eydft"][1::2]
    return ey, ez


def bistatic(ntf, p: dict, theta, phi, log=None):
    """sigma (m^2) at directions (theta, phi) [rad] for every NTF frequency.

    Returns arrays shaped [n_directions, M]: 'sig' (total, Taflove 8.37),
    'e_theta2' = |L_phi + eta0 N_theta|^2, 'e_phi2' = |L_theta - eta0 N_phi|^2.
    `ntf` may also be a list of payloads of one grid and frequency set (e.g.
    the same solve in several precisions): they share the phase factors, and
    a list of results comes back.  `log`, if given, is called with a line
    after each surface array.
    """
    many = isinstance(ntf, (list, tuple))
    ntfs = list(ntf) if many else [ntf]
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64))
    theta, phi = np.broadcast_arrays(theta, phi)
    f = np.asarray(ntfs[0]["fntf"], dtype=np.float64)
    k = 2 * math.pi * f / VP
    rhat, thhat, phhat = _direction(theta, phi)
    nd, M, P = theta.size, f.size, len(ntfs)
    L = np.zeros((P, nd, M, 3), dtype=np.complex128)
    N = np.zeros((P, nd, M, 3), dtype=np.complex128)
    # evenly spaced frequencies: exp(-i k_m p) = exp(-i k_0 p) exp(-i dk p)^m,
    # one complex product per sample and frequency instead of an exp
    even = M > 2 and np.allclose(np.diff(k), k[1] - k[0], rtol=1e-12, atol=0)
    for parts in zip(*[surface_samples(n, p) for n in ntfs]):
        kind, comp, _, w, pos = parts[0]
        proj = rhat @ pos.T  # [nd, S]
        zw = np.stack([z for _, _, z, _, _ in parts], axis=-1) * w[None, :, None]  # [M, S, P]
        acc = L if kind == "M" else N
        if even:
            ph = np.exp(-1j * k[0] * proj)
            step = np.exp(-1j * (k[1] - k[0]) * proj)
            for m in range(M):
                acc[:, :, m, comp] += (ph @ zw[m]).T
                ph *= step
        else:
            for m in range(M):
                acc[:, :, m, comp] += (np.exp(-1j * k[m] * proj) @ zw[m]).T
        if log:
            log(f"rcs: {kind}{'xyz'[comp]} surface array of {pos.shape[0]} samples done")
    out = []
    for q, n in enumerate(ntfs):
        Lth = np.einsum("dmc,dc->dm", L[q], thhat)
        Lph = np.einsum("dmc,dc->dm", L[q], phhat)
        Nth = np.einsum("dmc,dc->dm", N[q], thhat)
        Nph = np.einsum("dmc,dc->dm", N[q], phhat)
        ey, ez = source_phasors(n)
        pinc = (np.abs(ez) ** 2 + np.abs(ey) ** 2) / 2 / ETA0
        a = np.abs(Lph + ETA0 * Nth) ** 2
        c = np.abs(Lth - ETA0 * Nph) ** 2
        sig = k ** 2 / (8 * math.pi * ETA0) * (a + c) / np.abs(pinc)
        out.append({"f": f, "k": k, "sig": sig, "e_theta2": a, "e_phi2": c, "theta": theta, "phi": phi})
    return out if many else out[0]


def monostatic(ntf: dict, p: dict) -> dict:
    """Backscatter as analysis/meteorNTF.m computes it (r = -x, th = -z, ph = -y).

    Returns sig (Taflove 8.37, :332) and the polarised ratios thth, phth,
    phph, thph of :326-329, per frequency."""
    f = np.asarray(ntf["fntf"], dtype=np.float64)
    k = 2 * math.pi * f / VP
    M = f.size
    Lv = np.zeros((M, 3), dtype=np.complex128)
    Nv = np.zeros((M, 3), dtype=np.complex128)
    for kind, comp, z, w, pos in surface_samples(ntf, p):
        ph = np.exp(1j * k[:, None] * pos[None, :, 0])  # meteorNTF's exp(1j k dist), dist = x'
        s = np.sum(z * ph * w[None, :], axis=1)
        (Lv if kind == "M" else Nv)[:, comp] += s
    Lth, Lph, Nth, Nph = -Lv[:, 2], -Lv[:, 1], -Nv[:, 2], -Nv[:, 1]
    ey, ez = source_phasors(ntf)
    eth, eph = -ez, -ey
    pinc = (np.abs(eth) ** 2 + np.abs(eph) ** 2) / 2 / ETA0
    a = np.abs(Lph + ETA0 * Nth) ** 2
    c = np.abs(Lth - ETA0 * Nph) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        out = {
            "f": f, "k": k,
            "thth": k ** 2 / (4 * math.pi) * a / np.abs(eth) ** 2,
            "phth": k ** 2 / (4 * math.pi) * c / np.abs(eth) ** 2,
            "phph": k ** 2 / (4 * math.pi) * c / np.abs(eph) ** 2,
            "thph": k ** 2 / (4 * math.pi) * a / np.abs(eph) ** 2,
            "sig": k ** 2 / (8 * math.pi * ETA0) * (a + c) / np.abs(pinc),
        }
    return out


def db(sig) -> np.ndarray:
    """dBsm = 10 log10(sigma / 1 m^2)."""
    return 10 * np.log10(np.asarray(sig))


def principal_cuts(step_deg: float = 1.0):
    """Two principal-plane cuts: theta = 90 deg with phi from 0 to 180 deg,
    then phi = 0 with theta from 0 to 180 deg (181 directions each at 1 deg).
    Returns (theta, phi) in radians and the slices of the two cuts."""
    a = np.radians(np.arange(0.0, 180.0 + 1e-9, step_deg))
    n = a.size
    th = np.concatenate([np.full(n, math.pi / 2), a])
    ph = np.concatenate([a, np.zeros(n)])
    return th, ph, (slice(0, n), slice(n, 2 * n))


def payload_to_ntf(payload: np.ndarray, p: dict) -> dict:
    """The msc_solver_get_ntff payload as the face dict parse_ntf returns."""
    b = (p.get("pmllen") or 10) + 5
    return runio.parse_ntf(runio.encode_ntf(p["nntf"], b, p["fntf"], payload), p["nx"], p["ny"], p["nz"])


def db_errors(payload, reference: np.ndarray, p: dict, floor_db: float = 40.0, step_deg: float = 1.0, log=None):
    """RCS of `payload` against that of `reference` (same solve), in dB.

    Per frequency: the largest |dB difference| of the bistatic RCS over the
    two principal cuts, taken over directions where the reference RCS is
    within `floor_db` of its own maximum on that cut (deep nulls carry no
    meaningful dB figure), and the monostatic (meteorNTF.m backscatter)
    difference.  `payload` may be a list of payloads of the same solve: a
    list of results comes back, the phase factors computed once for all."""
    many = isinstance(payload, (list, tuple))
    pays = list(payload) if many else [payload]
    ntfs = [payload_to_ntf(a, p) for a in pays] + [payload_to_ntf(reference, p)]
    th, ph, cuts = principal_cuts(step_deg)
    dbs = [db(r["sig"]) for r in bistatic(ntfs, p, th, ph, log=log)]
    dr = dbs[-1]
    mr = db(monostatic(ntfs[-1], p)["sig"])
    out = []
    for da, a in zip(dbs[:-1], ntfs[:-1]):
        per_f = []
        for m in range(da.shape[1]):
            worst = 0.0
            for cut in cuts:
                ref = dr[cut, m]
                keep = ref > ref.max() - floor_db
                worst = max(worst, float(np.abs(da[cut, m] - ref)[keep].max()))
            per_f.append(worst)
        ma = db(monostatic(a, p)["sig"])
        out.append({"bistatic_db_per_freq": per_f, "bistatic_db_max": max(per_f),
                    "monostatic_db_per_freq": [float(x) for x in np.abs(ma - mr)],
                    "monostatic_db_max": float(np.abs(ma - mr).max()),
                    "monostatic_dbsm_reference": [float(x) for x in mr]})
    return out if many else out[0]


def mie_pec_backscatter(f, radius: float, terms: int = 100, c: float = 3e8) -> np.ndarray:
    """PEC-sphere backscatter RCS (m^2), the series of
    analysis/sphereRCSvsFreq.m:3-31 with spherical Bessel/Hankel functions."""
    f = np.atleast_1d(np.asarray(f, dtype=np.float64))
    out = np.empty(f.size)
    for q, fr in enumerate(f):
        lam = c / fr
        ka = 2 * math.pi * radius / lam
        # the reference sums 100 terms; beyond Wiscombe's ka + 4 ka^(1/3) + 2
        # they vanish in exact arithmetic but y_n overflows, so stop there
        nmax = min(terms, int(ka + 4 * ka ** (1 / 3) + 2) + 8)
        n = np.arange(1, nmax + 1)
        jn = special.spherical_jn(n, ka)
        yn = special.spherical_yn(n, ka)
        jn1 = special.spherical_jn(n - 1, ka)
        yn1 = special.spherical_yn(n - 1, ka)
        hn = jn - 1j * yn  # h^(2)
        hn1 = jn1 - 1j * yn1
        a = jn / hn
        b = (ka * jn1 - n * jn) / (ka * hn1 - n * hn)
        out[q] = lam ** 2 / math.pi * abs(np.sum((-1.0) ** n * (n + 0.5) * (b - a))) ** 2
    return out


def plate_broadside(f, size: float) -> np.ndarray:
    """Square plate of edge `size` at normal incidence, 4 pi A^2 / lambda^2
    (analysis/meteorNTF.m:347)."""
    f = np.asarray(f, dtype=np.float64)
    return 4 * math.pi * size ** 4 * f ** 2 / VP ** 2
