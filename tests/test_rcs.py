"""RCS post-processing (meteorscattering_amd/rcs.py, counterpart of
analysis/meteorNTF.m) pinned by closed forms and by FDTD-vs-Mie.

MATLAB is absent, so meteorNTF.m cannot be run.  These checks pin it:
- the Mie series against its Rayleigh and optical limits;
- surface weights against the exact face areas and centroids;
- the bistatic form at backscatter against the monostatic form;
- a vacuum run giving no scattered field;
- the reference's sigma=1e7 sphere (run by the FP64 oracle) against the Mie
  series at 70-100 MHz, where the FDTD is resolved (mistakes such as a
  flipped J sign or a conjugated phasor move these by > 3 dB).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cases
from meteorscattering_amd import configs, rcs, runio


def test_mie_limits():
    a = 1.0
    # Rayleigh: sigma / (pi a^2) -> 9 (ka)^4
    ka = 0.02
    f = ka * 3e8 / (2 * math.pi * a)
    s = rcs.mie_pec_backscatter([f], a)[0] / (math.pi * a * a)
    assert abs(s / (9 * ka ** 4) - 1) < 2e-3
    # optical: sigma -> pi a^2
    f = 60 * 3e8 / (2 * math.pi * a)
    s = rcs.mie_pec_backscatter([f], a, terms=150)[0] / (math.pi * a * a)
    assert abs(s - 1) < 0.05
    # first resonance peak near ka = 1 (~3.6 pi a^2)
    ks = np.linspace(0.8, 1.2, 81)
    s = rcs.mie_pec_backscatter(ks * 3e8 / (2 * math.pi * a), a) / (math.pi * a * a)
    assert 3.5 < s.max() < 3.8


def _fake_ntf(p, fill=0.0):
    nx, ny, nz, M = p["nx"], p["ny"], p["nz"], p["nntf"]
    out = {"nntf": M, "ntfb": (p.get("pmllen") or 10) + 5, "fntf": np.asarray(p["fntf"], float),
           "Eydft": np.zeros(2 * M), "Ezdft": np.zeros(2 * M)}
    shapes = [(ny, nz)] * 8 + [(nx, nz)] * 8 + [(nx, ny)] * 8
    for name, sh in zip(runio.ntf_face_names(), shapes):
        out[name] = np.full(sh + (2 * M,), fill)
    return out


def test_surface_weights_are_face_areas_and_centred():
    p = configs.base_params((50, 57, 64), 0.5, 10)
    ntf = _fake_ntf(p)
    b = ntf["ntfb"]
    d = 0.5
    ext = [(p["nx"] - 2 * b) * d, (p["ny"] - 2 * b) * d, (p["nz"] - 2 * b) * d]
    seen = 0
    for kind, comp, z, w, pos in rcs.surface_samples(ntf, p):
        axis = int(np.argmax(np.ptp(pos, axis=0) == 0))  # the constant coordinate is the face normal
        tang = [a for a in range(3) if a != axis]
        assert abs(w.sum() - ext[tang[0]] * ext[tang[1]]) < 1e-9
        centre = (np.array([p["nx"], p["ny"], p["nz"]]) / 2.0) * d
        cen = (pos * w[:, None]).sum(0) / w.sum()
        np.testing.assert_allclose(cen[tang], centre[tang], atol=1e-9)
        seen += 1
    assert seen == 24


def test_constant_current_integral():
    """A uniform unit M_y on the front face (cos part 1) gives L_y = face area
    (backscatter phase is constant on an x-face)."""
    p = configs.base_params((48, 48, 48), 0.5, 10)
    p.update(nntf=1, fntf=[50e6])
    ntf = _fake_ntf(p)
    ntf["My_front"][..., 1] = 1.0
    ntf["Ezdft"][1] = 1.0  # unit source so Pinc > 0
    b = ntf["ntfb"]
    area = ((48 - 2 * b) * 0.5) ** 2
    k = 2 * math.pi * 50e6 / rcs.VP
    r = rcs.monostatic(ntf, p)
    # |L_phi|^2 = area^2, Pinc = 1/(2 eta0): sigma = k^2/(8 pi eta0) area^2 * 2 eta0
    assert abs(r["sig"][0] / (k * k / (4 * math.pi) * area * area) - 1) < 1e-12


@pytest.fixture(scope="module")
def sphere_ntf(oracle_lib):
    n = 72
    p = configs.base_params((n, n, n), 0.1, 900, f0=100e6)
    p.update(doplate=2, platesize=0.6, doB0=0, sourcepol=1, nntf=3, fntf=[70e6, 85e6, 100e6],
             numfiles=1, savefields=0)
    o = oracle_lib.Oracle(p, np.zeros((n, n)), np.zeros((n, n)))
    o.step(p["tsteps"])
    pay = o.ntff()
    o.close()
    return p, runio.parse_ntf(runio.encode_ntf(3, 15, p["fntf"], pay), n, n, n)


def test_sphere_fdtd_matches_mie(sphere_ntf):
    p, ntf = sphere_ntf
    got = rcs.db(rcs.monostatic(ntf, p)["sig"])
    mie = rcs.db(rcs.mie_pec_backscatter(p["fntf"], 0.6))
    assert np.all(np.abs(got - mie) < 1.35), (got, mie)


def test_bistatic_at_backscatter_equals_monostatic(sphere_ntf):
    p, ntf = sphere_ntf
    m = rcs.monostatic(ntf, p)["sig"]
    b = rcs.bistatic(ntf, p, math.pi / 2, math.pi)["sig"][0]
    np.testing.assert_allclose(b, m, rtol=1e-12)


def test_bistatic_mirror_symmetry_of_sphere(sphere_ntf):
    """A sphere centred on the grid with z-polarised incidence along +x has a
    pattern symmetric under y -> -y (phi -> -phi)."""
    p, ntf = sphere_ntf
    th = np.array([math.pi / 2, math.pi / 3, 2 * math.pi / 3])
    ph = np.array([2.5, 2.0, 2.8])
    a = rcs.bistatic(ntf, p, th, ph)["sig"]
    b = rcs.bistatic(ntf, p, th, -ph)["sig"]
    # the reference masks every staggered component at integer positions, so
    # the discretised sphere is symmetric only to about 0.25 dB
    np.testing.assert_allclose(rcs.db(a), rcs.db(b), atol=0.3)


def test_vacuum_has_no_scattered_field(oracle_lib, sphere_ntf):
    n = 48
    p = configs.base_params((n, n, n), 0.5, 300)
    p.update(doB0=0, numfiles=1, savefields=0, nntf=2, fntf=[40e6, 50e6])
    o = oracle_lib.Oracle(p, np.zeros((n, n)), np.zeros((n, n)))
    o.step(300)
    ntf = runio.parse_ntf(runio.encode_ntf(2, 15, p["fntf"], o.ntff()), n, n, n)
    s = rcs.monostatic(ntf, p)["sig"]
    assert np.all(s < 1e-20)


def test_even_frequency_recurrence_equals_direct_phases():
    """bistatic() builds exp(-i k_m r.r') by recurrence for evenly spaced
    frequencies; one frequency at a time takes the direct exp."""
    p = configs.base_params((48, 50, 52), 0.5, 10)
    p.update(nntf=5, fntf=list(np.linspace(40e6, 60e6, 5)))
    ntf = _fake_ntf(p)
    rng = np.random.default_rng(3)
    for name in runio.ntf_face_names():
        ntf[name] = rng.standard_normal(ntf[name].shape)
    ntf["Ezdft"] = rng.standard_normal(10)
    th = np.linspace(0.1, 3.0, 7)
    ph = np.linspace(-3.0, 3.0, 7)
    many = rcs.bistatic(ntf, p, th, ph)["sig"]
    for m in range(5):
        one = dict(ntf, fntf=ntf["fntf"][m:m + 1], Ezdft=ntf["Ezdft"][2 * m:2 * m + 2],
                   Eydft=ntf["Eydft"][2 * m:2 * m + 2])
        for name in runio.ntf_face_names():
            one[name] = ntf[name][..., 2 * m:2 * m + 2]
        np.testing.assert_allclose(rcs.bistatic(one, p, th, ph)["sig"][:, 0], many[:, m], rtol=1e-10)


def test_db_errors_closed_form(oracle_lib):
    """rcs.db_errors: a payload against itself is 0 dB; scaling every face
    sum by c (the source DFT unchanged) scales L and N by c and sigma by c^2,
    20 log10(c) dB in every direction and at every frequency."""
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(tsteps=120, nntf=3, fntf=[45e6, 50e6, 55e6])
    o = oracle_lib.Oracle(p, wp, nu)
    o.step(120)
    a = o.ntff()
    o.close()
    e0 = rcs.db_errors(a, a, p)
    assert e0["bistatic_db_max"] == 0.0 and e0["monostatic_db_max"] == 0.0
    c = 1.0 + 3e-3
    b = a.copy()
    b[4 * p["nntf"]:] *= c
    e = rcs.db_errors(b, a, p)
    want = 20 * math.log10(c)
    np.testing.assert_allclose(e["bistatic_db_per_freq"], want, rtol=1e-9)
    np.testing.assert_allclose(e["monostatic_db_per_freq"], want, rtol=1e-9)
