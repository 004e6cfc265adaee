"""Batch and serving drivers over the multi-solve server, with the serial
single-solve baseline they are compared against.

Reference counterpart: setup/meteorbatch.m and setup/meteorrandom.m, which
submit one job per run (setup/meteorrun.m:154-183); every job runs the
whole time loop in its own process, one after another on a node.
"""
from __future__ import annotations

import time

import numpy as np

from .server import Server
from .solver import Solver


def cell_updates(trails) -> int:
    return sum(p["nx"] * p["ny"] * p["nz"] * p["tsteps"] for p, _, _ in trails)


def class_tiling(table, p: dict):
    """The tiling a server with size-class table `table` gives solve `p`."""
    if not table:
        return None
    n = p["nx"] * p["ny"] * p["nz"]
    for m, t in table:
        if n <= m:
            return t
    return table[-1][1]


def run_serial(trails, precision: int = 32, tilings=None) -> dict:
    """One solve at a time through msc_solver_*: create (plan, media upload),
    all steps, NTFF read-back, destroy.  Returns the per-solve wall times.
    `tilings`: a size-class table as for Server.set_tilings (None: defaults)."""
    times, ntffs = [], []
    t0 = time.perf_counter()
    for p, wp, nu in trails:
        t = time.perf_counter()
        with Solver(p, wp, nu, precision=precision, tiling=class_tiling(tilings, p)) as s:
            s.step(p["tsteps"])
            ntffs.append(s.ntff())
        times.append(time.perf_counter() - t)
    return {"wall": time.perf_counter() - t0, "times": np.array(times), "ntff": ntffs}


def run_batch(trails, precision: int = 32, budget: int = 0, chunk: int = 4, tilings=None) -> dict:
    """Every trail submitted at once to one server that steps every resident
    solve each iteration (throughput setting).  `wall` runs from before the
    server exists to the last result; `setup` is the part of it spent creating
    the server (its arena allocation: a fraction of a second on fresh device
    memory, several seconds when a previous server in the same process has
    just released a large arena)."""
    t0 = time.perf_counter()
    with Server(precision=precision, budget_bytes=budget, fill_tiles=0) as srv:
        setup = time.perf_counter() - t0
        if tilings:
            srv.set_tilings(tilings)
        ids = [srv.submit(p, wp, nu) for p, wp, nu in trails]
        srv.run_until_done(chunk)
        wall = time.perf_counter() - t0
        infos = [srv.info(i) for i in ids]
        ntffs = [srv.ntff(i) for i in ids]
        st = srv.stats()
    return {"wall": wall, "setup": setup, "info": infos, "ntff": ntffs, "stats": st}


def fifo_latencies(arrivals: np.ndarray, service: np.ndarray) -> np.ndarray:
    """Latencies of a one-at-a-time server taking requests in arrival order:
    finish_i = max(arrival_i, finish_(i-1)) + service_i."""
    fin = np.empty_like(arrivals)
    prev = 0.0
    for i, (a, s) in enumerate(zip(arrivals, service)):
        prev = max(a, prev) + s
        fin[i] = prev
    return fin - arrivals


def poisson_arrivals(rate: float, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def run_serve(trails, arrivals: np.ndarray, precision: int = 32, budget: int = 0, chunk: int = 2,
              tilings=None) -> dict:
    """Requests submitted at their arrival times (server clock) while the
    server keeps stepping the resident solves; latency = completion - arrival."""
    n = len(trails)
    with Server(precision=precision, budget_bytes=budget) as srv:
        if tilings:
            srv.set_tilings(tilings)
        base = srv.clock
        ids = [None] * n
        nxt = 0
        while True:
            now = srv.clock - base
            while nxt < n and arrivals[nxt] <= now:
                p, wp, nu = trails[nxt]
                ids[nxt] = srv.submit(p, wp, nu)
                nxt += 1
            srv.advance(chunk)
            st = srv.stats()
            idle = not (st["queued"] or st["ready"] or st["running"] or st["draining"])
            if idle and nxt == n:
                break
            if idle:
                time.sleep(max(0.0, arrivals[nxt] - (srv.clock - base)))
        infos = [srv.info(i) for i in ids]
        ntffs = [srv.ntff(i) for i in ids]
        wall = max(inf["t_done"] for inf in infos) - base
    lat = np.array([inf["t_done"] - base - a for inf, a in zip(infos, arrivals)])
    return {"wall": wall, "latency": lat, "info": infos, "ntff": ntffs}
