"""Multi-solve server (msc_server_*): results in a mix equal results alone.

The server advances every resident solve with ganged launches (tiles of
all solves in one launch, one argument block per solve).  Each solve's
arithmetic is its single-solve arithmetic, so the NTFF payload and the
probes must be bit-identical to msc_solver_* on the same inputs, whatever
else is resident, whichever current model the neighbours use (DI,
Lee-Kalluri polynomial and closed form are launched as separate groups),
and whether the arena made the request wait.
"""
from __future__ import annotations

import numpy as np
import pytest

import cases
from meteorscattering_amd.server import Server, footprint
from meteorscattering_amd.solver import Solver

pytestmark = pytest.mark.gpu

MIX = ["lk_ricker_ez", "di_modgauss_ey", "plate_ricker_45", "lk_nopml_lhcp", "lk_overdense_offset"]


def _case(name, **kw):
    p, wp, nu = cases.case(name)
    p = dict(p, savefields=0, **kw)
    return p, wp, nu


def _alone(p, wp, nu, precision=32):
    with Solver(p, wp, nu, precision=precision) as s:
        s.step(p["tsteps"])
        s.sync()
        return s.ntff(), s.probes(), footprint(p, precision)


@pytest.fixture(scope="module")
def alone():
    return {n: _alone(*_case(n)) for n in MIX}


def test_mix_equals_alone(alone):
    with Server(precision=32, budget_bytes=4 << 30) as srv:
        ids = {}
        for n in MIX:
            ids[n] = srv.submit(*_case(n))
        srv.run_until_done()
        for n in MIX:
            inf = srv.info(ids[n])
            assert inf["state_name"] == "done", (n, inf)
            assert 0 < inf["t_submit"] <= inf["t_start"] <= inf["t_done"]
            assert inf["t"] == inf["tsteps"]
            ntff, probes, nbytes = alone[n]
            assert inf["device_bytes"] >= nbytes
            np.testing.assert_array_equal(srv.ntff(ids[n]), ntff, err_msg=n)
            np.testing.assert_array_equal(srv.probes(ids[n]), probes, err_msg=n)
        st = srv.stats()
        assert st["arena_used"] == 0
        # every gang iteration advanced several solves at once
        assert st["solve_steps"] == sum(cases.case(n)[0]["tsteps"] for n in MIX)
        assert st["iterations"] < st["solve_steps"]


def test_full_arena_queues_in_order_and_rejects_oversize(alone):
    sizes = {n: footprint(_case(n)[0]) for n in MIX}
    budget = max(sizes.values()) + min(sizes.values()) // 2  # one resident solve at a time
    with Server(precision=32, budget_bytes=budget) as srv:
        big = dict(_case("lk_ricker_ez")[0], nx=120, ny=120, nz=120)
        bid = srv.submit(big, *_big_maps(big))
        ids = [srv.submit(*_case(n)) for n in MIX]
        srv.run_until_done()
        inf = srv.info(bid)
        assert inf["state_name"] == "failed" and "arena" in inf["error"]
        starts = []
        for n, rid in zip(MIX, ids):
            inf = srv.info(rid)
            assert inf["state_name"] == "done"
            starts.append(inf["t_start"])
            ntff, probes, _ = alone[n]
            np.testing.assert_array_equal(srv.ntff(rid), ntff, err_msg=n)
            np.testing.assert_array_equal(srv.probes(rid), probes, err_msg=n)
        assert starts == sorted(starts)  # first come, first served
        # a second round on the same arena reuses the freed blocks
        rid = srv.submit(*_case("di_modgauss_ey"))
        srv.run_until_done()
        np.testing.assert_array_equal(srv.ntff(rid), alone["di_modgauss_ey"][0])
        srv.release(rid)
        with pytest.raises(RuntimeError):
            srv.info(rid)


def _big_maps(p):
    return np.zeros((p["nx"], p["ny"])), np.zeros((p["nx"], p["ny"]))


def test_fp64_mix_equals_alone():
    names = ["lk_ricker_ez", "di_modgauss_ey"]
    with Server(precision=64, budget_bytes=2 << 30) as srv:
        ids = [srv.submit(*_case(n)) for n in names]
        srv.run_until_done()
        for n, rid in zip(names, ids):
            ntff, probes, _ = _alone(*_case(n), precision=64)
            np.testing.assert_array_equal(srv.ntff(rid), ntff, err_msg=n)
            np.testing.assert_array_equal(srv.probes(rid), probes, err_msg=n)


def test_savefields_is_rejected():
    p, wp, nu = cases.case("lk_ricker_ez")
    with Server(precision=32, budget_bytes=1 << 30) as srv:
        with pytest.raises(RuntimeError, match="savefields"):
            srv.submit(dict(p, savefields=1), wp, nu)


@pytest.mark.parametrize("fill,graphs", [(1, True), (0, False), (400, True)])
def test_scheduling_settings_do_not_change_results(alone, fill, graphs):
    """fill 1 steps one solve at a time (the oldest), fill 0 every resident
    solve, fill 400 a few small ones at a time with pauses and resumptions;
    with and without captured graphs the results are the single-solve bits."""
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=fill, graphs=graphs) as srv:
        ids = {n: srv.submit(*_case(n)) for n in MIX}
        srv.run_until_done()
        for n in MIX:
            np.testing.assert_array_equal(srv.ntff(ids[n]), alone[n][0], err_msg=n)
            np.testing.assert_array_equal(srv.probes(ids[n]), alone[n][1], err_msg=n)
        if fill == 1:
            st = srv.stats()
            assert st["iterations"] == st["solve_steps"]


def _two_step_case(name, tsteps, **kw):
    """60x50x205 grid: in the server (32-lane one-step tiles) a two-step
    region [16,44) x [16,34) x [31,186) with a probe in its rim; alone the
    grid runs one-step, and the bits match."""
    p, _, _ = cases.case(name)
    p.update(nx=60, ny=50, nz=205, savefields=0, tsteps=tsteps, nprobes=2, probex=[30, 17], probey=[25, 33],
             probez=[100, 64], **kw)
    wp, nu = cases.trail_maps(p, n0=2e13, r0=1.5, nu0=1e6, centre=(0.5, -0.25))
    return p, wp, nu


TWO_STEP_MIX = {
    "lk61": lambda: _two_step_case("lk_ricker_ez", 61),
    "di40": lambda: _two_step_case("di_modgauss_ey", 40),
    "jcol33": lambda: _two_step_case("lk_ricker_ez", 33, j_collocated=1),
    "small50": lambda: _case("lk_ricker_ez", tsteps=50),  # no two-step region
}


@pytest.mark.parametrize("fill,graphs", [(0, True), (1, False), (5000, True)])
def test_two_step_gangs_equal_alone(fill, graphs, monkeypatch):
    """Solves with a two-step region advance two steps per gang iteration
    (the last one alone when their step count is odd) beside one-step
    solves, with MSC_SERVER_PAIR_MIN=0 so that gangs of these small regions
    pair; fill 1 steps one solve at a time, 5000 a few at a time.  The
    NTFF payloads and probes equal the bits of single solves run one-step
    (these regions are too small for a single solve's two-step rule) with
    64-lane tiles."""
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    monkeypatch.setenv("MSC_ZLANES", "64")
    alone = {k: _alone(*v) for k, v in mix.items()}
    monkeypatch.delenv("MSC_ZLANES")
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")  # these small regions pair in any gang
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=fill, graphs=graphs) as srv:
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)
        st = srv.stats()
        assert st["solve_steps"] == sum(v[0]["tsteps"] for v in mix.values())
        assert st["iterations"] < st["solve_steps"]
        if fill == 1:  # one solve per iteration: two-step solves take half the iterations (plus an odd last step)
            expect = sum((v[0]["tsteps"] + 1) // 2 if k != "small50" else v[0]["tsteps"] for k, v in mix.items())
            assert st["iterations"] == expect


@pytest.mark.parametrize("fuse,graphs,precision", [(1, True, 32), (1, False, 32), (0, True, 32), (1, True, 64)])
def test_one_launch_gang_lists_equal_alone(fuse, graphs, precision, monkeypatch):
    """MSC_SERVER_FUSE=1 runs a gang's one-step lists as one launch per
    current model and tile rows (fdtd_steps, the default for gangs of few
    one-step cells), 0 as one launch per list and variant; with two-step
    and one-step solves of two current models in one gang, captured or
    launched directly, both give the single-solve bits."""
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    mix["di_small"] = _case("di_modgauss_ey", tsteps=40)
    alone = {k: _alone(*v, precision=precision) for k, v in mix.items()}
    monkeypatch.setenv("MSC_SERVER_FUSE", str(fuse))
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")  # both halves of an iteration
    with Server(precision=precision, budget_bytes=4 << 30, fill_tiles=0, graphs=graphs) as srv:
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)


@pytest.mark.parametrize("fuse", [1, 0])
def test_gangs_of_small_regions_run_one_step(fuse, monkeypatch):
    """By default a gang pairs only when its solves' regions give at least
    1.5 two-step workgroups per CU together.  With fill 1 each gang is one
    of these small solves, whose region does not, so every iteration
    advances it by one step (iterations = solve steps, where
    test_two_step_gangs_equal_alone's paired gangs take about half), and
    the results are the single-solve bits."""
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    alone = {k: _alone(*v) for k, v in mix.items()}
    monkeypatch.delenv("MSC_SERVER_PAIR_MIN", raising=False)
    monkeypatch.setenv("MSC_SERVER_FUSE", str(fuse))
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=1) as srv:
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)
        st = srv.stats()
        assert st["iterations"] == st["solve_steps"] == sum(v[0]["tsteps"] for v in mix.values())


def test_timing_mode_keeps_the_bits_and_reports(capfd, monkeypatch):
    """MSC_SERVER_TIMING=1 (direct launches with events between the parts of
    each half) gives the single-solve bits and, when the server closes,
    prints the GPU time of the one-step lists, the two-step launch, the
    NTFF samples and probes, and the contractions."""
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    alone = {k: _alone(*v) for k, v in mix.items()}
    monkeypatch.setenv("MSC_SERVER_TIMING", "1")
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=0) as srv:
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)
    err = capfd.readouterr().err
    assert "server timing:" in err
    for part in ("one-step lists", "two-step launch", "NTFF samples + probes"):
        assert part in err, err


def test_fp64_two_step_gangs_equal_alone(monkeypatch):
    """FP64 gangs with two-step tiles (16 rows x 32 lanes, 12 x 28 owned) on
    the 60x50x205 grids (paired with MSC_SERVER_PAIR_MIN=0) equal FP64
    single solves (which run these small regions one-step) bit for bit."""
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    alone = {k: _alone(*v, precision=64) for k, v in mix.items()}
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")
    with Server(precision=64, budget_bytes=4 << 30, fill_tiles=1) as srv:
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)
        st = srv.stats()
        # one solve per iteration: the two-step solves take half the iterations
        expect = sum((v[0]["tsteps"] + 1) // 2 if k != "small50" else v[0]["tsteps"] for k, v in mix.items())
        assert st["iterations"] == expect


def test_server_refuses_auxiliary_list_streams(monkeypatch):
    """The server launches its gangs on one stream: MSC_LIST_STREAMS other
    than 0 is refused at creation (the single-solve setting of that name
    does not carry over), and 0 is accepted."""
    monkeypatch.setenv("MSC_LIST_STREAMS", "2")
    with pytest.raises(RuntimeError, match="MSC_LIST_STREAMS"):
        Server(precision=32, budget_bytes=1 << 30)
    monkeypatch.setenv("MSC_LIST_STREAMS", "0")
    with Server(precision=32, budget_bytes=1 << 30) as srv:
        assert srv.stats()["arena_bytes"] >= 1 << 30


def test_size_class_tilings_equal_alone(monkeypatch):
    """Tilings by size class (small grids 3/15-row tiles and 5/7-plane
    chunks, large ones 15/3 rows and 11/13 planes): the gang launches carry
    several tile heights at once (two-step gangs: MSC_SERVER_PAIR_MIN=0);
    results equal the single-solve bits.  The table can change only while
    the server holds no request."""
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    alone = {k: _alone(*v) for k, v in mix.items()}
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")
    small = {"ty": 3, "ty_b": 15, "chunk": 5, "chunk2": 7}
    large = {"ty": 15, "ty_b": 3, "chunk": 11, "chunk2": 13}
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=0) as srv:
        srv.set_tilings([(100_000, small), (10**12, large)])
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        with pytest.raises(RuntimeError):
            srv.set_tilings([])
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)


def test_compensated_fp32_ntff_sums_in_the_server_equal_alone(monkeypatch):
    """With MSC_NTFF_ACC=32 the gang's contractions run the compensated FP32
    kernel on each solve's own sums (two-step gangs: MSC_SERVER_PAIR_MIN=0):
    payloads equal the single solves'."""
    monkeypatch.setenv("MSC_NTFF_ACC", "32")
    monkeypatch.setenv("MSC_SERVER_PAIR_MIN", "0")
    mix = {k: f() for k, f in TWO_STEP_MIX.items()}
    alone = {k: _alone(*v) for k, v in mix.items()}
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=0) as srv:
        ids = {k: srv.submit(*v) for k, v in mix.items()}
        srv.run_until_done()
        for k in mix:
            np.testing.assert_array_equal(srv.ntff(ids[k]), alone[k][0], err_msg=k)
            np.testing.assert_array_equal(srv.probes(ids[k]), alone[k][1], err_msg=k)


def _unstable(**kw):
    """Plasma frequency 1e4 times the trail's: omega_p dt ~ 1700, far past the
    explicit coupling's stability limit, so the FP32 fields overflow to
    infinity and NaN within a few dozen steps of the pulse reaching it."""
    p, wp, nu = _case("lk_ricker_ez", **kw)
    return p, wp * 1e4, nu


def test_cancel_mid_flight_leaves_co_resident_results_bit_identical(alone):
    names = ["lk_ricker_ez", "di_modgauss_ey", "lk_nopml_lhcp"]
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=0) as srv:
        ids = {n: srv.submit(*_case(n)) for n in names}
        while srv.info(ids["di_modgauss_ey"])["t"] < 6:  # set-up is asynchronous
            srv.advance(1)
        srv.cancel(ids["di_modgauss_ey"])
        srv.run_until_done()
        inf = srv.info(ids["di_modgauss_ey"])
        assert inf["state_name"] == "cancelled" and 0 < inf["t"] < inf["tsteps"], inf
        for n in ("lk_ricker_ez", "lk_nopml_lhcp"):
            assert srv.info(ids[n])["state_name"] == "done"
            np.testing.assert_array_equal(srv.ntff(ids[n]), alone[n][0], err_msg=n)
            np.testing.assert_array_equal(srv.probes(ids[n]), alone[n][1], err_msg=n)
        assert srv.stats()["arena_used"] == 0
        with pytest.raises(RuntimeError, match="no steps left"):
            srv.cancel(ids["lk_ricker_ez"])
        with pytest.raises(RuntimeError, match="not done"):
            srv.ntff(ids["di_modgauss_ey"])
        srv.release(ids["di_modgauss_ey"])


def test_cancel_a_queued_request(alone):
    """With room for one solve at a time, a request cancelled while it waits
    never runs; the one ahead of it is unaffected."""
    sizes = [footprint(_case(n)[0]) for n in ("lk_ricker_ez", "lk_nopml_lhcp")]
    with Server(precision=32, budget_bytes=max(sizes) + min(sizes) // 2) as srv:
        a = srv.submit(*_case("lk_ricker_ez"))
        b = srv.submit(*_case("lk_nopml_lhcp"))
        srv.cancel(b)
        srv.run_until_done()
        assert srv.info(b)["state_name"] == "cancelled" and srv.info(b)["t"] == 0
        np.testing.assert_array_equal(srv.ntff(a), alone["lk_ricker_ez"][0])


def test_time_limit_fails_requests_that_run_too_long(alone):
    with Server(precision=32, budget_bytes=4 << 30) as srv:
        srv.set_time_limit(1e-6)
        rid = srv.submit(*_case("lk_ricker_ez"))
        srv.run_until_done()
        inf = srv.info(rid)
        assert inf["state_name"] == "failed" and "time limit" in inf["error"], inf
        assert inf["t"] < inf["tsteps"]
        srv.set_time_limit(0)
        rid = srv.submit(*_case("lk_ricker_ez"))
        srv.run_until_done()
        np.testing.assert_array_equal(srv.ntff(rid), alone["lk_ricker_ez"][0])
        assert srv.stats()["arena_used"] == 0
        with pytest.raises(RuntimeError):
            srv.set_time_limit(-1)


def test_non_finite_fields_fail_the_request_not_its_neighbours(alone):
    """An unstable solve's fields overflow: the watchdog fails it (mid-flight
    or at completion) while a co-resident solve completes with the bits it
    has alone; a single solve reports the same flag."""
    with Solver(*_unstable(), precision=32) as g:
        g.step(120)
        assert g.nonfinite()
    with Solver(*_case("lk_ricker_ez"), precision=32) as g:
        g.step(120)
        assert not g.nonfinite()
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=0) as srv:
        bad = srv.submit(*_unstable())
        good = srv.submit(*_case("lk_ricker_ez"))
        srv.run_until_done()
        inf = srv.info(bad)
        assert inf["state_name"] == "failed" and "non-finite" in inf["error"], inf
        assert srv.info(good)["state_name"] == "done"
        np.testing.assert_array_equal(srv.ntff(good), alone["lk_ricker_ez"][0])
        p, wp, nu = _case("lk_ricker_ez")
        wp = wp.copy()
        wp[3, 4] = np.nan
        with pytest.raises(RuntimeError, match="not finite"):
            srv.submit(p, wp, nu)


def test_gang_tuning_returns_a_table_that_keeps_the_bits(alone):
    """autotune.tune_gang times candidate size-class tables inside one
    server's gang (short batches of every request) and returns a table no
    smaller than the defaults; solves under it equal their solo bits."""
    from meteorscattering_amd import autotune

    names = ["lk_ricker_ez", "di_modgauss_ey", "lk_nopml_lhcp"]
    reqs = [_case(n) for n in names]
    table, t0, t1 = autotune.tune_gang(reqs, [80_000], steps=24, chunk2s=(32, 48), chunks=(8, 16), tys=(7, 15),
                                       budget_bytes=4 << 30)
    assert table and all(t["chunk"] >= 8 and t["chunk2"] >= 32 and t["ty"] >= 7 for _, t in table)
    assert 0 < t1 <= t0
    with pytest.raises(ValueError):
        autotune.tune_gang(reqs, [80_000], chunks=(4, 8))
    with Server(precision=32, budget_bytes=4 << 30, fill_tiles=0) as srv:
        srv.set_tilings(table)
        ids = {n: srv.submit(*_case(n)) for n in names}
        srv.run_until_done()
        for n in names:
            np.testing.assert_array_equal(srv.ntff(ids[n]), alone[n][0], err_msg=n)
