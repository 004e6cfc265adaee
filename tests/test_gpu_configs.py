"""The benchmark workloads, unchanged, on the HIP path against the FP64 oracle.

Each workload comes from meteorscattering_amd.configs with its own grid,
CPML thickness, media, current model and tiling rules; only the number of
steps run here is cut to what the oracle finishes in seconds.

  400x400x800 (12-cell CPML, tilted 3-D trail, nu(z) with a 7 km scale
      height, collocated Lee-Kalluri current): 60 steps, so that the Ricker
      pulse peaks at the TF/SF face x = 22, which lies inside the two-step
      region [18, 382) where the tilted trail crosses it (z 50-135); the
      default tiling runs fdtd_pair there;
  96x96x192 (reference form: 10-cell CPML, node-averaged Lee-Kalluri): 200
      steps, one-step tiles by default and two-step tiles forced;
  96x96x192 with 64 NTFF frequencies: 130 steps (K = 64: two contracted
      blocks and a flushed partial one), NTFF payload on the matrix cores;
  sweep and request-stream trails (random 3-D trails, stability-limited
      dt, 1-16 frequencies): oracle parity over 240 steps, and whole solves
      through the server equal to the same solves alone, with finite NTFF;
  RCS in dB: the 96x96x192 geometry with 16 frequencies over 1200 steps,
      FP32 against the oracle's FP64 NTFF, within 0.1 dB.

Field error of a kind (E, H or J): max |gpu - oracle| over its three
components divided by the peak |oracle| of that kind so far, at checkpoints
through the run.  Tolerances: FP64 1e-10; FP32 2e-5 (NTFF payload: of its
largest entry).
"""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest

from meteorscattering_amd import configs, rcs
from meteorscattering_amd.server import Server
from meteorscattering_amd.solver import Solver

pytestmark = pytest.mark.gpu

TOL = {64: 1e-10, 32: 2e-5}


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def oracle_states(oracle_lib, p, wp, nu, checkpoints):
    """The oracle's fields at each checkpoint, and its NTFF payload and probes
    at the last."""
    t0 = time.perf_counter()
    o = oracle_lib.Oracle(p, wp, nu)
    fields = []
    for t in checkpoints:
        o.step(t - o.t)
        fields.append([o.field(c) for c in range(9)])
    st = {"checkpoints": list(checkpoints), "fields": fields, "ntff": o.ntff(), "probes": o.probes()}
    o.close()
    say(f"oracle {p['nx']}x{p['ny']}x{p['nz']} {checkpoints[-1]} steps: {time.perf_counter() - t0:.1f} s")
    return st


def check_run(g, st, precision, label):
    """Step GPU solver `g` through the oracle's checkpoints.  At each, the
    error of a field kind is max |gpu - oracle| over its components divided
    by the peak |oracle| of that kind over the checkpoints so far (after the
    pulse has left a small grid, its late-time fields are small); the NTFF
    payload and probes are compared at the end."""
    tol = TOL[precision]
    peak = [0.0, 0.0, 0.0]
    worst = [0.0, 0.0, 0.0]
    for t, ref in zip(st["checkpoints"], st["fields"]):
        g.step(t - g.t)
        for kind in range(3):
            comps = range(3 * kind, 3 * kind + 3)
            peak[kind] = max(peak[kind], max(np.abs(ref[c]).max() for c in comps))
            if peak[kind] > 0:
                err = max(np.abs(g.field(c) - ref[c]).max() for c in comps)
                worst[kind] = max(worst[kind], float(err / peak[kind]))
    n = rel(g.ntff(), st["ntff"])
    pr = rel(g.probes(), st["probes"]) if st["probes"].size else 0.0
    say(f"{label} FP{precision}: E {worst[0]:.2e} H {worst[1]:.2e} J {worst[2]:.2e} NTFF {n:.2e} "
        f"probes {pr:.2e}")
    assert max(worst) < tol, (label, worst)
    assert n < tol, (label, n)
    assert pr < tol, (label, pr)


# ---- 400x400x800 ----------------------------------------------------------

STEPS_400 = 60


@pytest.fixture(scope="module")
def trail_400(oracle_lib):
    p, wp, nu = configs.trail_400x400x800()
    st = oracle_states(oracle_lib, p, wp, nu, [STEPS_400])
    # the pulse and the trail's current have both arrived
    assert np.abs(st["fields"][-1][2]).max() > 1e-3 and np.abs(st["fields"][-1][8]).max() > 0
    return p, wp, nu, st


@pytest.mark.parametrize("precision", [32, 64])
def test_400x400x800_matches_oracle_with_two_step_region(trail_400, precision):
    """The headline workload as bench.py runs it (default tiling: the
    two-step region holds 4,836 fdtd_pair tiles) against the oracle."""
    p, wp, nu, st = trail_400
    with Solver(p, wp, nu, precision=precision) as g:
        g.step(1)  # t = 0 writes a frame: one-step everywhere
        g.set_timing(True)
        g.step(STEPS_400 - 1)
        g.sync()
        tm = g.timing()
        assert tm["pair_launches"] == (STEPS_400 - 2) // 2 and tm["pair_cells"] > 0.7 * 400 * 400 * 800, tm
        assert tm["pair_j_cells"] > 0, tm
        check_run(g, st, precision, "400x400x800")


# ---- 96x96x192 -------------------------------------------------------------

@pytest.fixture(scope="module")
def trail_96(oracle_lib):
    p, wp, nu = configs.trail_96x96x192()
    return p, wp, nu, oracle_states(oracle_lib, p, wp, nu, [40, 80, 120, 160, 200])


@pytest.mark.parametrize("precision,pair", [(32, None), (64, None), (32, "1"), (64, "1")])
def test_96x96x192_reference_form_matches_oracle(trail_96, monkeypatch, precision, pair):
    p, wp, nu, st = trail_96
    if pair is None:
        monkeypatch.delenv("MSC_PAIR", raising=False)
    else:
        monkeypatch.setenv("MSC_PAIR", pair)
    with Solver(p, wp, nu, precision=precision) as g:
        g.step(2)
        g.set_timing(True)
        g.step(18)
        g.sync()
        assert (g.timing()["pair_launches"] > 0) == (pair == "1")
        check_run(g, st, precision, f"96x96x192 pair={pair}")


@pytest.fixture(scope="module")
def trail_96_m64(oracle_lib):
    p, wp, nu = configs.ntff_m64_96x96x192()
    return p, wp, nu, oracle_states(oracle_lib, p, wp, nu, [65, 130])


@pytest.mark.parametrize("precision", [32, 64])
def test_96x96x192_64_frequencies_matches_oracle(trail_96_m64, precision):
    """M = 64: the contraction runs on the matrix cores (FP64 MFMA), two
    64-step blocks plus a flushed block of 2 steps."""
    p, wp, nu, st = trail_96_m64
    with Solver(p, wp, nu, precision=precision) as g:
        check_run(g, st, precision, "96x96x192 M=64")


# ---- sweep and request-stream trails ----------------------------------------

def _smallest(trails, n):
    return sorted(trails, key=lambda t: t[0]["nx"] * t[0]["ny"] * t[0]["nz"])[:n]


@pytest.fixture(scope="module")
def mix():
    sweep = configs.trail_sweep(maps=False)
    stream = configs.request_stream(count=48, maps=False)
    picks = _smallest(sweep, 3) + [sorted(sweep, key=lambda t: t[0]["nz"])[len(sweep) // 2]] + _smallest(stream, 2)
    return [(p,) + configs.trail_media(p) for p, _, _ in picks]


@pytest.mark.parametrize("which", [0, 4])
@pytest.mark.parametrize("precision", [32, 64])
def test_sweep_and_stream_trails_match_oracle(oracle_lib, mix, which, precision):
    """A sweep trail (M = 1) and a request-stream trail (M = 16): n0 up to
    1e15 m^-3, nu up to 1e7 1/s, the stability-limited dt, 3-D media."""
    p, wp, nu = mix[which]
    st = oracle_states(oracle_lib, p, wp, nu, [40, 80, 120, 160, 200, 240])
    with Solver(p, wp, nu, precision=precision) as g:
        check_run(g, st, precision, f"trail {p['nx']}x{p['ny']}x{p['nz']} M={p['nntf']}")


def test_sweep_and_stream_trails_in_the_server_equal_alone(mix, monkeypatch):
    """Whole solves (1,500-4,000 steps) of six workload trails, one of them
    mid-size with a two-step region (paired whatever its size:
    MSC_SERVER_PAIR_MIN=0), all resident in one server: every NTFF payload
    and probe history equals the same solve alone, bit for bit, and is
    finite."""
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")
    alone = []
    for p, wp, nu in mix:
        with Solver(p, wp, nu, precision=32) as g:
            g.step(p["tsteps"])
            alone.append((g.ntff(), g.probes()))
    with Server(precision=32, fill_tiles=0) as srv:
        ids = [srv.submit(p, wp, nu) for p, wp, nu in mix]
        srv.run_until_done()
        for (p, _, _), rid, (ntff, probes) in zip(mix, ids, alone):
            inf = srv.info(rid)
            assert inf["state_name"] == "done", inf
            a = srv.ntff(rid)
            assert np.all(np.isfinite(a)) and np.abs(a).max() > 0
            np.testing.assert_array_equal(a, ntff)
            np.testing.assert_array_equal(srv.probes(rid), probes)
        say(f"server: {len(mix)} workload trails equal alone; stats {srv.stats()}")


# ---- RCS in dB ---------------------------------------------------------------

def test_rcs_fp32_within_0p1_db_of_the_fp64_oracle(oracle_lib):
    """96x96x192 trail (collocated current, 16 frequencies 30-80 MHz), 1200
    steps: bistatic RCS over two 1-degree principal cuts (directions within
    40 dB of each cut's peak) and the monostatic RCS, from the FP32 HIP NTFF
    and from the oracle's FP64 NTFF, agree within 0.1 dB per frequency."""
    p, wp, nu = configs.trail_96x96x192(tsteps=1200)
    p.update(j_collocated=1, nntf=16, fntf=list(np.linspace(30e6, 80e6, 16)), savefields=0, numfiles=1)
    t0 = time.perf_counter()
    o = oracle_lib.Oracle(p, wp, nu)
    o.step(1200)
    ref = o.ntff()
    o.close()
    say(f"oracle 1200 steps M=16: {time.perf_counter() - t0:.1f} s")
    with Solver(p, wp, nu, precision=32) as g:
        g.step(1200)
        a = g.ntff()
    e = rcs.db_errors(a, ref, p)
    say(f"RCS FP32 vs oracle: bistatic {e['bistatic_db_max']:.2e} dB, monostatic {e['monostatic_db_max']:.2e} dB; "
        f"monostatic dBsm {np.round(e['monostatic_dbsm_reference'], 2).tolist()}")
    assert e["bistatic_db_max"] < 0.1 and e["monostatic_db_max"] < 0.1
