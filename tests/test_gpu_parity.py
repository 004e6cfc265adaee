"""HIP path (libmsc.so) against the FP64 oracle on the same inputs.

Field error of a kind (E, H or J): max |gpu - oracle| over its three
components divided by the running peak max |oracle| of that kind.
Tolerances:
  FP64 path: 1e-10 - same algorithm, differing only in rounding (reciprocal
             multiplies, FMA contraction, device exp);
  FP32 path: 2e-5 over the 80-160-step horizons below (FP32 fields, J and
             CPML; FP64 plasma coefficients and NTFF sums); NTFF payload
             2e-5 of its largest entry.
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import cases
import summary
from meteorscattering_amd import _abi, runio

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def msc():
    from meteorscattering_amd.solver import Solver, run_dir

    return Solver, run_dir


def relerr(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def worst(a, b):
    d = np.abs(a - b)
    idx = np.unravel_index(np.argmax(d), d.shape)
    return f"at {tuple(int(x) for x in idx)}: gpu={a[idx]!r} oracle={b[idx]!r}"


def run_pair(Solver, oracle_lib, p, wp, nu, precision, nsteps, check_every, tol, label, stride=1):
    """Step GPU and oracle together, `stride` steps per call.  Error of a
    field kind (E, H or J) is max |gpu - oracle| over its three components
    divided by the running peak max |oracle| of that kind over the run so
    far."""
    g = Solver(p, wp, nu, precision=precision)
    o = oracle_lib.Oracle(p, wp, nu)
    report = []
    peak = [0.0, 0.0, 0.0]
    errs = [0.0, 0.0, 0.0]
    for t in range(stride - 1, nsteps + stride - 1, stride):
        t = min(t, nsteps - 1)
        n = t + 1 - g.t
        g.step(n)
        o.step(n)
        if (t + 1) % check_every and t + 1 != nsteps:
            continue
        pairs = [(g.field(c), o.field(c)) for c in range(9)]
        for kind in range(3):
            peak[kind] = max(peak[kind], max(np.abs(b).max() for a, b in pairs[3 * kind:3 * kind + 3]))
        for c, (a, b) in enumerate(pairs):
            kind = c // 3
            if peak[kind] == 0:
                continue
            e = float(np.abs(a - b).max() / peak[kind])
            errs[kind] = max(errs[kind], e)
            if e > tol:
                report.append(f"{label} step {t} {_abi.COMPONENTS[c]} rel {e:.3e} {worst(a, b)}")
        if report:
            break
    print(f"{label}: max-rel field error E {errs[0]:.2e} H {errs[1]:.2e} J {errs[2]:.2e}")
    return g, o, report


STABLE = ["lk_ricker_ez", "di_modgauss_ey", "plate_ricker_45", "sphere_sin_rhcp", "lk_nopml_lhcp"]


@pytest.mark.parametrize("name", STABLE + ["lk_overdense_offset"])
def test_fp64_fields_match_oracle_every_step(msc, oracle_lib, name):
    Solver, _ = msc
    p, wp, nu = cases.case(name)
    g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, 64, p["tsteps"], 10, 1e-10, name)
    assert not rep, "\n".join(rep[:12])
    # NTFF payload and probes
    a, b = g.ntff(), o.ntff()
    assert relerr(a, b) < 1e-10
    np.testing.assert_allclose(g.probes(), o.probes(), rtol=0, atol=1e-10 * np.abs(o.probes()).max())


@pytest.mark.parametrize("name", STABLE)
def test_fp32_fields_match_oracle(msc, oracle_lib, name):
    Solver, _ = msc
    p, wp, nu = cases.case(name)
    g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, 32, p["tsteps"], 20, 2e-5, name)
    assert not rep, "\n".join(rep[:12])
    a, b = g.ntff(), o.ntff()
    assert relerr(a, b) < 2e-5


def test_fp64_frames_probes_and_log_maxima(msc, oracle_lib):
    Solver, _ = msc
    p, wp, nu = cases.case("lk_ricker_ez")
    r = oracle_lib.run_like_reference(p, wp, nu)
    g = Solver(p, wp, nu, precision=64)
    g.step(p["tsteps"])
    g.sync()
    assert g.nframes == len(r["eframes"])
    for f in range(g.nframes):
        e, h, em = g.frame(f)
        assert relerr(e, r["eframes"][f]) < 1e-10
        assert relerr(h, r["hframes"][f]) < 1e-10
        assert abs(em - r["emax"][f]) <= 1e-10 * max(r["emax"].max(), 1e-300)
    np.testing.assert_allclose(g.probes(), r["probes"], rtol=0, atol=1e-10 * np.abs(r["probes"]).max())


def test_noncubic_grid_matches_oracle(msc, oracle_lib):
    """Non-cubic grids (the reference's z coordinates are undefined there)."""
    Solver, _ = msc
    p, _, _ = cases.case("lk_ricker_ez")
    p.update(nx=46, ny=50, nz=57, tsteps=120, probex=[23, 20], probey=[25, 27], probez=[28, 26])
    wp, nu = cases.trail_maps(p, n0=2e13, r0=1.5, nu0=1e6, centre=(0.5, -0.25))
    for prec, tol in ((64, 1e-10), (32, 2e-5)):
        g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, prec, p["tsteps"], 30, tol, f"noncubic{prec}")
        assert not rep, "\n".join(rep[:12])


def test_interior_tiles_with_cpml_match_oracle(msc, oracle_lib, monkeypatch):
    """A grid large enough that tiles clear of every CPML slab (and of the
    slab halos) run the interior variant while the rest run the CPML variant:
    z = 160 gives middle 31-cell z tiles clear of both z slabs, 8-plane x
    chunks and 7-row y tiles leave interior chunks and rows inside 60 x 48."""
    Solver, _ = msc
    monkeypatch.setenv("MSC_CHUNK", "8")
    monkeypatch.setenv("MSC_TY", "7")
    p, _, _ = cases.case("lk_ricker_ez")
    p.update(nx=60, ny=48, nz=160, tsteps=100, nprobes=2, probex=[30, 20], probey=[24, 30], probez=[80, 120])
    wp, nu = cases.trail_maps(p, n0=2e13, r0=1.5, nu0=1e6, centre=(0.5, -0.25))
    for prec, tol in ((64, 1e-10), (32, 2e-5)):
        g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, prec, p["tsteps"], 25, tol, f"interior{prec}")
        assert not rep, "\n".join(rep[:12])
        assert relerr(g.ntff(), o.ntff()) < tol


def test_tilings_are_bit_identical(msc, oracle_lib, monkeypatch):
    """Every cell is computed by the same code whatever the tile heights of
    the interior and CPML-boundary launches (3, 7, 15 rows) or the x-chunk length (halo
    cells are recomputed from read-only data).  The grid has interior
    tiles (z = 160 leaves a middle z-tile clear of both z slabs)."""
    Solver, _ = msc
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(tsteps=80, nx=58, ny=61, nz=160, probex=[29], probey=[30], probez=[80], nprobes=1)
    wp, nu = cases.trail_maps(p)
    outs = []
    for ty, tyb, chunk in ((7, 3, 8), (3, 7, 5), (3, 3, 1), (15, 7, 47), (7, 15, 16)):
        monkeypatch.setenv("MSC_TY", str(ty))
        monkeypatch.setenv("MSC_TY_B", str(tyb))
        monkeypatch.setenv("MSC_CHUNK", str(chunk))
        g = Solver(p, wp, nu, precision=32)
        g.step(80)
        outs.append((np.stack([g.field(c) for c in range(9)]), g.ntff()))
        g.close()
    for f, n in outs[1:]:
        assert np.array_equal(f, outs[0][0])
        assert np.array_equal(n, outs[0][1])


# Two-step grid: with L = 10 the two-step region is [16,44) x [16,34) x
# [31,186) (z bounds on the 31-cell one-step z tiles); it holds the TF/SF x
# and y faces (20, 40; 20, 30) and the high z face (185, the region's last
# layer), the plate or sphere, the trail, and probes in its two rim layers.
PAIR_GRID = dict(nx=60, ny=50, nz=205, nprobes=3, probex=[30, 17, 43], probey=[25, 33, 16],
                 probez=[100, 32, 184])


def pair_case(name, jcol, tsteps):
    p, _, _ = cases.case(name)
    p.update(PAIR_GRID, tsteps=tsteps, j_collocated=jcol, numfiles=3)
    wp, nu = cases.trail_maps(p, n0=2e13, r0=1.5, nu0=1e6, centre=(0.5, -0.25))
    return p, wp, nu


def snapshot(g):
    frames = [g.frame(f) for f in range(g.nframes)]
    return [g.field(c) for c in range(9)] + [g.ntff(), g.probes()] + [np.concatenate(
        [np.ravel(x) for x in fr]) for fr in frames]


@pytest.mark.parametrize("name,jcol,chunk2", [("lk_ricker_ez", 0, 32), ("lk_ricker_ez", 1, 32),
                                              ("di_modgauss_ey", 0, 32), ("plate_ricker_45", 0, 32),
                                              ("sphere_sin_rhcp", 0, 32), ("lk_nopml_lhcp", 0, 32),
                                              ("lk_ricker_ez", 0, 3), ("lk_nopml_lhcp", 0, 1)])
@pytest.mark.parametrize("precision", [32, 64])
def test_two_step_tiles_are_bit_identical(msc, monkeypatch, name, jcol, chunk2, precision):
    """Two-step mode (fdtd_pair on the region clear of the CPML slabs and the
    NTFF planes, one-step launches around it; x chunks of `chunk2` planes; stepped
    in uneven chunks so single steps, frame steps and pairs mix) gives the
    one-step path's fields, NTFF payload, probes and frames bit for bit."""
    Solver, _ = msc
    p, wp, nu = pair_case(name, jcol, 90)
    monkeypatch.setenv("MSC_CHUNK2", str(chunk2))
    outs = []
    for pair, chunks in ((0, [90]), (1, [1, 2, 5, 13, 30, 39])):
        monkeypatch.setenv("MSC_PAIR", str(pair))
        g = Solver(p, wp, nu, precision=precision)
        for c in chunks:
            g.step(c)
        outs.append(snapshot(g))
        g.close()
    labels = list(_abi.COMPONENTS) + ["ntff", "probes"] + [f"frame{f}" for f in range(len(outs[0]) - 11)]
    bad = [f"{lab}: {worst(a, b)}" for lab, a, b in zip(labels, outs[1], outs[0]) if not np.array_equal(a, b)]
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("precision", [32, 64])
def test_crowded_probes_give_the_same_bits(msc, monkeypatch, precision):
    """Probes whose 3x3x3 boxes overlap inside the two-step region (three in
    one (y, z) column, a fourth beside them) need step-t+1 write-back from
    cells in up to four boxes; two-step mode gives the one-step path's
    fields, probes and NTFF payload bit for bit."""
    Solver, _ = msc
    p, wp, nu = pair_case("lk_ricker_ez", 0, 60)
    p.update(nprobes=4, probex=[24, 26, 28, 27], probey=[25, 25, 25, 26], probez=[100, 100, 100, 101])
    outs = []
    for pair in (0, 1):
        monkeypatch.setenv("MSC_PAIR", str(pair))
        with Solver(p, wp, nu, precision=precision) as g:
            for c in (1, 2, 5, 13, 39):
                g.step(c)
            outs.append(snapshot(g))
    bad = [i for i, (a, b) in enumerate(zip(outs[1], outs[0])) if not np.array_equal(a, b)]
    assert not bad, bad


@pytest.mark.parametrize("name", ["lk_ricker_ez", "lk_nopml_lhcp", "plate_ricker_45"])
@pytest.mark.parametrize("pair,precision", [(1, 32), (0, 32), (1, 64), (0, 64)])
def test_lane_widths_give_the_same_bits(msc, monkeypatch, name, pair, precision):
    """One-step tiles of 64 z lanes (63 owned z cells) and of 32 lanes (two
    rows per wave, 31 cells; the default of single solves and server gangs,
    whose two-step regions then start and end on multiples of 31) give the
    same fields, NTFF payload and probes bit for bit."""
    Solver, _ = msc
    p, wp, nu = pair_case(name, 0, 61)
    p.update(savefields=0)
    monkeypatch.setenv("MSC_PAIR", str(pair))
    outs = []
    for lanes in (64, 32):
        monkeypatch.setenv("MSC_ZLANES", str(lanes))
        with Solver(p, wp, nu, precision=precision) as g:
            for c in (1, 6, 23, 31):
                g.step(c)
            outs.append(snapshot(g))
    labels = list(_abi.COMPONENTS) + ["ntff", "probes"]
    bad = [f"{lab}: {worst(a, b)}" for lab, a, b in zip(labels, outs[1], outs[0]) if not np.array_equal(a, b)]
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("pair", [1, 0])
def test_concurrent_lists_give_the_same_bits(msc, monkeypatch, pair):
    """The one-step lists of a step launched one per list (MSC_FUSE_LISTS=0)
    on 0 (one after another), 1 and 3 auxiliary streams give the same
    fields, NTFF payload, probes and frames bit for bit: they write disjoint
    cells and read only the current set."""
    Solver, _ = msc
    p, wp, nu = pair_case("lk_ricker_ez", 0, 61)
    monkeypatch.setenv("MSC_PAIR", str(pair))
    monkeypatch.setenv("MSC_FUSE_LISTS", "0")
    outs = []
    for ns in (0, 1, 3):
        monkeypatch.setenv("MSC_LIST_STREAMS", str(ns))
        with Solver(p, wp, nu, precision=32) as g:
            for c in (1, 6, 23, 31):
                g.step(c)
            outs.append(snapshot(g))
    for o in outs[1:]:
        bad = [i for i, (a, b) in enumerate(zip(o, outs[0])) if not np.array_equal(a, b)]
        assert not bad, bad


@pytest.mark.parametrize("name,ty,lanes,closed", [("lk_ricker_ez", 7, 32, 0), ("lk_ricker_ez", 3, 32, 0),
                                                   ("lk_ricker_ez", 15, 64, 0), ("lk_ricker_ez", 7, 64, 1),
                                                   ("di_modgauss_ey", 7, 64, 0), ("plate_ricker_45", 7, 32, 0),
                                                   ("lk_nopml_lhcp", 7, 32, 0)])
@pytest.mark.parametrize("pair", [0, 1])
@pytest.mark.parametrize("precision", [32, 64])
def test_fused_lists_give_the_same_bits(msc, monkeypatch, name, ty, lanes, closed, pair, precision):
    """MSC_FUSE_LISTS=1 (the five one-step lists of a step in one launch of
    fdtd_steps, each tile running its list's CPML variant, the default below
    16 MB per field array) gives the fields, NTFF payload, probes and frames of
    MSC_FUSE_LISTS=0 (one launch per list) bit for bit, for
    tiles of 3, 7 and 15 rows, 32 lanes (with slice steps) and 64, the three
    current models, and with the two-step region on and off."""
    Solver, _ = msc
    p, wp, nu = pair_case(name, 0, 61)
    if lanes == 64:
        p.update(savefields=0)
    monkeypatch.setenv("MSC_PAIR", str(pair))
    monkeypatch.setenv("MSC_ZLANES", str(lanes))
    monkeypatch.setenv("MSC_TY", str(ty))
    monkeypatch.setenv("MSC_TY_B", str(ty))
    if closed:
        monkeypatch.setenv("MSC_LK_CLOSED", "1")
    outs = []
    for fuse in (0, 1):
        monkeypatch.setenv("MSC_FUSE_LISTS", str(fuse))
        with Solver(p, wp, nu, precision=precision) as g:
            for c in (1, 6, 23, 31):
                g.step(c)
            outs.append(snapshot(g))
    labels = list(_abi.COMPONENTS) + ["ntff", "probes"] + [f"frame {f}" for f in range(len(outs[0]) - 11)]
    bad = [f"{lab}: {worst(a, b)}" for lab, a, b in zip(labels, outs[1], outs[0]) if not np.array_equal(a, b)]
    assert not bad, "\n".join(bad)


def test_lane_width_is_checked(msc, monkeypatch):
    """MSC_ZLANES other than 32 or 64 is refused at creation, as is 64 with
    slice capture (those tiles carry no slice code)."""
    Solver, _ = msc
    p, wp, nu = pair_case("lk_ricker_ez", 0, 10)
    monkeypatch.setenv("MSC_ZLANES", "48")
    with pytest.raises(RuntimeError, match="MSC_ZLANES"):
        Solver(p, wp, nu, precision=32)
    monkeypatch.setenv("MSC_ZLANES", "64")
    p.update(savefields=1)
    with pytest.raises(RuntimeError, match="slice"):
        Solver(p, wp, nu, precision=32)


@pytest.mark.parametrize("pair", [0, 1])
@pytest.mark.parametrize("precision", [32, 64])
def test_current_free_tiles_change_nothing(msc, monkeypatch, pair, precision):
    """Tiles where eps0 wp^2 is zero on every cell skip J and the media; a
    narrow trail tapered in z leaves most tiles current-free.  Fields, NTFF
    payload and probes equal those of the run that updates J everywhere
    (zeros may differ in sign only)."""
    Solver, _ = msc
    p, _, _ = pair_case("lk_ricker_ez", 0, 60)
    p.update(metsize=[2.0, 2.0])
    wp, nu = cases.trail_maps(p, n0=5e13, r0=0.35, nu0=1e6, centre=(0.5, -0.25))
    monkeypatch.setenv("MSC_PAIR", str(pair))
    outs = []
    for everywhere in (True, False):
        if everywhere:
            monkeypatch.setenv("MSC_CURRENT_EVERYWHERE", "1")
        else:
            monkeypatch.delenv("MSC_CURRENT_EVERYWHERE", raising=False)
        g = Solver(p, wp, nu, precision=precision)
        g.step(60)
        outs.append(snapshot(g))
        jcells = g.timing()["step_j_cells"]
        g.close()
    ncell = p["nx"] * p["ny"] * p["nz"]
    assert 0 < jcells < ncell * 0.9, jcells  # tiles skip J (FP64 underflows later: more carry it)
    labels = list(_abi.COMPONENTS) + ["ntff", "probes"] + [f"frame{f}" for f in range(len(outs[0]) - 11)]
    bad = [f"{lab}: {worst(a, b)}" for lab, a, b in zip(labels, outs[1], outs[0]) if not np.array_equal(a, b)]
    assert not bad, "\n".join(bad)
    # the J fields are nonzero somewhere: the trail carries a current
    assert np.abs(outs[1][6]).max() > 0


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 2e-5)])
def test_two_step_tiles_match_oracle(msc, oracle_lib, monkeypatch, precision, tol):
    """Two-step mode against the oracle, 10 steps per call."""
    Solver, _ = msc
    monkeypatch.setenv("MSC_PAIR", "1")
    p, wp, nu = pair_case("lk_ricker_ez", 0, 100)
    g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, precision, p["tsteps"], 20, tol, f"pair{precision}",
                         stride=10)
    assert not rep, "\n".join(rep[:12])
    assert relerr(g.ntff(), o.ntff()) < tol
    assert relerr(g.probes(), o.probes()) < tol


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 2e-5)])
def test_lee_kalluri_closed_form_variant_matches_oracle(msc, oracle_lib, monkeypatch, precision, tol):
    """The closed-form coefficient kernel (used when the polynomial would not
    be exact, forced here) against the oracle."""
    Solver, _ = msc
    monkeypatch.setenv("MSC_LK_CLOSED", "1")
    p, wp, nu = cases.case("lk_nopml_lhcp")
    g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, precision, p["tsteps"], 40, tol, f"lkclosed{precision}")
    assert not rep, "\n".join(rep[:12])


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 2e-5)])
def test_collocated_current_extension_matches_oracle(msc, oracle_lib, precision, tol):
    """j_collocated = 1 (E at J's own edge; not in the reference) on a grid
    with interior tiles, against the oracle's same extension."""
    Solver, _ = msc
    p, _, _ = cases.case("lk_ricker_ez")
    p.update(nx=60, ny=48, nz=140, tsteps=120, j_collocated=1, nprobes=1, probex=[30], probey=[24], probez=[70])
    wp, nu = cases.trail_maps(p, n0=5e13, r0=1.5, nu0=1e5, centre=(0.5, -0.25))
    g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, precision, p["tsteps"], 30, tol, f"jcol{precision}")
    assert not rep, "\n".join(rep[:12])
    assert relerr(g.ntff(), o.ntff()) < tol


def test_lee_kalluri_zero_collision_and_zero_field_limits(msc, oracle_lib):
    """nu = 0 in part of the grid and B = 0: finite limits (the reference
    divides by zero there), GPU polynomial path against the oracle's limits."""
    Solver, _ = msc
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(tsteps=100)
    nu = nu.copy()
    nu[:, : nu.shape[1] // 2] = 0.0
    for B in ([4.33e-5, 0.0, 2.5e-5], [0.0, 0.0, 0.0]):
        p["B0"] = B
        g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, 64, p["tsteps"], 50, 1e-10, f"limits{B}")
        assert not rep, "\n".join(rep[:12])


def test_vacuum_tfsf_leakage_fp32(msc):
    Solver, _ = msc
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(tsteps=150, doB0=0, sourcepol=4)
    g = Solver(p, wp * 0, nu, precision=32)
    L, n = 10, p["nx"]
    b1, b2 = L + 10, n - L - 10
    worst_in = worst_out = 0.0
    for t in range(15):
        g.step(10)
        for c in (0, 1, 2):
            f = g.field(c)
            worst_in = max(worst_in, np.abs(f[b1 + 1:b2 - 1, b1 + 1:b2 - 1, b1 + 1:b2 - 1]).max())
            m = np.ones(f.shape, bool)
            m[b1 - 2:b2 + 2, b1 - 2:b2 + 2, b1 - 2:b2 + 2] = False
            worst_out = max(worst_out, np.abs(f[m]).max())
    assert worst_in > 0.2
    assert worst_out < 1e-6 * worst_in


@pytest.mark.parametrize("name", ["lk_ricker_ez", "sphere_sin_rhcp"])
def test_dropin_run_dir_matches_reference_fixture(msc, name, tmp_path):
    """The `meteor` executable in a run directory written from the fixture's
    inputs produces files matching the reference's (FP64 path)."""
    f = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    p = json.loads(str(f["params"]))
    runio.write_run_dir(str(tmp_path), p, f["wp"], f["nu"])
    exe = os.path.join(os.path.dirname(_abi.LIB_PATH), "meteor")
    env = dict(os.environ, MSC_PRECISION="64")
    r = subprocess.run([exe], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    ntf = runio.read_ntf(str(tmp_path / "output_NTF.dat"), nx, ny, nz)
    b = int(f["ntf_header"][1])
    assert ntf["ntfb"] == b and ntf["nntf"] == p["nntf"]
    for fn in runio.ntf_face_names():
        a = ntf[fn]
        assert summary.compare_projection(fn, a, f["proj_" + fn], 0) < 1e-10, fn
        np.testing.assert_allclose(a[b:a.shape[0] - b + 1, b:a.shape[1] - b + 1], f["carve_" + fn],
                                   rtol=0, atol=1e-10 * max(np.abs(f["carve_" + fn]).max(), 1e-300))
    probe = runio.read_probe(str(tmp_path / "Probe.dat"), p["tsteps"])
    np.testing.assert_allclose(probe["data"], f["probes"], rtol=0, atol=1e-10 * np.abs(f["probes"]).max())
    ef = runio.read_frames(str(tmp_path / "output_E.dat"), nx, ny, nz)
    hf = runio.read_frames(str(tmp_path / "output_H.dat"), nx, ny, nz)
    jf = runio.read_frames(str(tmp_path / "output_J.dat"), nx, ny, nz, with_inc=False)
    assert ef.shape[0] == int(f["nframes"][0]) and jf.shape[0] == int(f["nframes"][2])
    assert np.abs(jf).max() == 0.0
    for kind, frames in (("E", ef), ("H", hf)):
        for n, fr in enumerate(frames):
            parts = runio.split_frame(fr, nx, ny, nz)
            for key in ("slicez", "slicey", "slicex"):
                assert summary.compare_projection(f"{kind}{n}_{key}", parts[key], f[f"proj_{kind}{n}_{key}"], 0) < 1e-10
            np.testing.assert_array_equal(np.concatenate([parts["inc1"], parts["inc2"]]), f[f"{kind}{n}_inc"])
    log = open(tmp_path / "log.txt").read()
    assert "All done!" in log
    emax = [float(x) for x in __import__("re").findall(r"Maximum abs\(Ex\) is ([^\s]+)", log)]
    np.testing.assert_allclose(emax, f["emax_log"], rtol=2e-5, atol=1e-12)


def test_timing_record_counts_launch_kinds(msc, monkeypatch):
    """msc_solver_get_timing: in two-step mode a 10-step call runs pairs,
    reports the region's cell counts and a positive elapsed time; with
    MSC_PAIR=0 it reports no two-step launches, and neither does the default
    here: the region's 6 two-step tiles are far fewer than 1.5 per CU, so a
    single solve runs it one-step."""
    Solver, _ = msc
    p, wp, nu = pair_case("lk_ricker_ez", 0, 40)
    for pair in (1, 0, None):
        if pair is None:
            monkeypatch.delenv("MSC_PAIR", raising=False)
        else:
            monkeypatch.setenv("MSC_PAIR", str(pair))
        with Solver(p, wp, nu, precision=32) as g:
            g.step(2)  # past the first frame step
            g.set_timing(True)
            g.step(10)
            g.sync()
            tm = g.timing()
        assert tm["steps"] == 10 and tm["span_ms"] > 0 and tm["step_ms"] > 0
        if pair:
            assert tm["pair_launches"] >= 4 and tm["pair_ms"] > 0
            assert tm["pair_cells"] == (44 - 16) * (34 - 16) * (186 - 31)
            assert 0 < tm["pair_j_cells"] <= tm["pair_cells"]
        else:
            assert tm["pair_launches"] == 0 and tm["pair_cells"] == 0


@pytest.mark.parametrize("precision", [32, 64])
def test_default_two_step_mode_follows_the_region_size(msc, monkeypatch, precision):
    """With MSC_PAIR unset, a single solve runs two-step tiles when its
    region gives at least 1.5 fdtd_pair tiles per CU: 200x200x300 with L = 10
    has a region [16,184)^2 x [31,279) of 6 x 14 x 5 = 420 tiles in FP32
    (60 owned z cells per tile) and 6 x 14 x 9 = 756 in FP64 (28 owned z
    cells), the 60x50x205 grid 1 x 2 x 3 = 6 and 1 x 2 x 6 = 12."""
    import ctypes

    Solver, _ = msc
    monkeypatch.delenv("MSC_PAIR", raising=False)
    # the CU count from the HIP runtime libmsc.so uses (attribute 63 is
    # hipDeviceAttributeMultiprocessorCount in ROCm 7.2's hip_runtime_api.h)
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    assert hip.hipDeviceGetAttribute(ctypes.byref(n), 63, 0) == 0
    cus = n.value
    assert cus > 0
    zt = {32: (5, 3), 64: (9, 6)}[precision]
    for (nx, ny, nz), tiles in (((200, 200, 300), 6 * 14 * zt[0]), ((60, 50, 205), 1 * 2 * zt[1])):
        p, _, _ = cases.case("lk_ricker_ez")
        p.update(nx=nx, ny=ny, nz=nz, tsteps=12, numfiles=1, nprobes=1, probex=[nx // 2], probey=[ny // 2],
                 probez=[nz // 2])
        wp, nu = cases.trail_maps(p, n0=2e13, r0=1.5, nu0=1e6, centre=(0.5, -0.25))
        with Solver(p, wp, nu, precision=precision) as g:
            g.step(2)
            g.set_timing(True)
            g.step(6)
            g.sync()
            tm = g.timing()
        expect = tiles >= 3 * cus // 2
        assert (tm["pair_launches"] > 0) == expect, (nx, ny, nz, tiles, cus, tm["pair_launches"])


@pytest.mark.parametrize("nntf", [8, 16, 24, 100])
@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 2e-5)])
def test_ntff_matrix_core_contraction_matches_oracle(msc, oracle_lib, nntf, precision, tol):
    """With 2M >= 16 frequency rows the NTFF contraction runs on the matrix
    cores (FP64 MFMA tiles): 8, 16 and 24 frequencies take passes of 1, 2
    and 4 tiles; 100 frequencies = 200 rows, a pass of 8 tiles and one of 5
    with a partial tile.  75 steps close K-step blocks (16-slot Phi blocks,
    the last one partial) mid-run and at the end."""
    Solver, _ = msc
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(tsteps=75, nntf=nntf, fntf=list(np.linspace(30e6, 80e6, nntf)))
    g = Solver(p, wp, nu, precision=precision)
    o = oracle_lib.Oracle(p, wp, nu)
    g.step(75)
    o.step(75)
    e = relerr(g.ntff(), o.ntff())
    print(f"M={nntf} FP{precision}: NTFF rel err {e:.2e}")
    assert e < tol
    g.close()
    o.close()


def test_autotuned_and_explicit_tilings_give_the_same_bits(msc, monkeypatch):
    """msc_autotune returns a valid tiling; solvers created with it and with
    an unusual explicit tiling give the default solver's results bit for bit
    (tilings change speed only; two-step tiles forced on this small grid)."""
    Solver, _ = msc
    monkeypatch.setenv("MSC_PAIR", "1")
    from meteorscattering_amd import autotune
    p, wp, nu = pair_case("lk_ricker_ez", 0, 30)
    tuned, ms = autotune.tune(p, wp, nu, precision=32, steps=6)
    assert tuned["ty"] in (3, 7, 15) and tuned["ty_b"] in (3, 7, 15) and tuned["chunk"] >= 1 and tuned["chunk2"] >= 1
    assert ms > 0
    outs = []
    for tiling in (None, tuned, {"ty": 3, "ty_b": 15, "chunk": 5, "chunk2": 7}):
        g = Solver(p, wp, nu, precision=32, tiling=tiling)
        g.step(30)
        outs.append(snapshot(g))
        g.close()
    for o in outs[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(o, outs[0]))


@pytest.mark.parametrize("nntf", [2, 16])
@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 2e-5)])
def test_ntff_read_mid_run_matches_oracle(msc, oracle_lib, nntf, precision, tol):
    """NTFF payloads read at steps inside K-step blocks (the buffered samples
    are flushed into the sums, the source DFT is summed through the step),
    twice in one block, at a block end and at the end of the run: each equals
    the oracle's running DFT after the same step, and reading does not change
    the later sums."""
    Solver, _ = msc
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(nntf=nntf, fntf=list(np.linspace(40e6, 60e6, nntf)))
    g = Solver(p, wp, nu, precision=precision)
    o = oracle_lib.Oracle(p, wp, nu)
    for t in (45, 50, 64, 101, p["tsteps"]):
        g.step(t - g.t)
        o.step(t - o.t)
        e = relerr(g.ntff(), o.ntff())
        assert e < tol, (t, e)
    g.close()
    o.close()


@pytest.mark.parametrize("nntf", [2, 16, 100])
@pytest.mark.parametrize("precision", [32, 64])
def test_ntff_compensated_fp32_sums_match_oracle(msc, oracle_lib, monkeypatch, nntf, precision):
    """MSC_NTFF_ACC=32: FP32 samples and the FP32 phase table on FP32 matrix
    tiles (FP32 accumulation over each K-step block), block sums added to
    compensated FP32 (sum, compensation) pairs.  Read mid-block and at the
    end, the payload matches the oracle's FP64 running DFT to 2e-5 of its
    largest entry (2, 16 and 100 frequencies: one partial row tile, one full
    tile, two row groups)."""
    Solver, _ = msc
    monkeypatch.setenv("MSC_NTFF_ACC", "32")
    p, wp, nu = cases.case("lk_ricker_ez")
    p.update(nntf=nntf, fntf=list(np.linspace(30e6, 80e6, nntf)))
    g = Solver(p, wp, nu, precision=precision)
    o = oracle_lib.Oracle(p, wp, nu)
    worst = 0.0
    for t in (50, 101, p["tsteps"]):
        g.step(t - g.t)
        o.step(t - o.t)
        e = relerr(g.ntff(), o.ntff())
        worst = max(worst, e)
        assert e < 2e-5, (t, e)
    print(f"compensated FP32 sums, M={nntf} FP{precision}: NTFF rel err {worst:.2e}")
    g.close()
    o.close()


@pytest.mark.parametrize("pair", ["0", "1"])
@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 2e-5)])
def test_headline_extensions_on_a_small_grid_match_oracle(msc, oracle_lib, monkeypatch, pair, precision, tol):
    """Every extension the 400x400x800 workload uses, on a grid the oracle
    runs in seconds: 12-cell CPML (slab offsets, TF/SF faces at L + 10, NTFF
    planes at L + 5, the two-step region from L + 6), 3-D media of a trail
    tilted 30 deg, nu(z) with a 9 m scale height and the collocated
    Lee-Kalluri current; one-step and two-step tiles."""
    from meteorscattering_amd import configs

    Solver, _ = msc
    monkeypatch.setenv("MSC_PAIR", pair)
    p = configs.base_params((64, 56, 220), 0.5, 130, pmllen=12)
    p.update(nu_scale_height=9.0, B0=configs.b_field(60.0, tilt_deg=30.0), j_collocated=1, media3d=1,
             numfiles=3, nprobes=2, probex=[32, 21], probey=[28, 30], probez=[110, 40])
    wp, nu = configs.trail_3d(p, 3e13, 1.5, 2e6, 30.0)
    g, o, rep = run_pair(Solver, oracle_lib, p, wp, nu, precision, 120, 20, tol,
                         f"extensions pair={pair} FP{precision}", stride=10)
    assert not rep, "\n".join(rep[:12])
    assert relerr(g.ntff(), o.ntff()) < tol
    assert relerr(g.probes(), o.probes()) < tol
    if pair == "1":
        g.set_timing(True)
        g.step(10)
        g.sync()
        assert g.timing()["pair_launches"] > 0


@pytest.mark.parametrize("precision", [32, 64])
def test_pattern_probe_leaves_the_solve_unchanged(msc, monkeypatch, precision):
    """msc_solver_pattern_probe (fdtd_pair's loads and owned stores with no
    arithmetic) stores every value it reads back where it read it: the
    fields right after it, and every step after that, are bit-identical to
    a solve that never ran it."""
    Solver, _ = msc
    p, wp, nu = pair_case("lk_ricker_ez", 0, 61)
    p.update(savefields=0)
    monkeypatch.setenv("MSC_PAIR", "1")
    outs = []
    for probe in (False, True):
        with Solver(p, wp, nu, precision=precision) as g:
            g.step(17)
            if probe:
                assert g.pattern_probe(3) > 0
            g.sync()
            a = snapshot(g)
            g.step(21)
            outs.append((a, snapshot(g)))
    for x, y in zip(outs[0], outs[1]):
        assert all(np.array_equal(u, v) for u, v in zip(x, y))
