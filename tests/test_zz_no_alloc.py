"""msc_solver_get_field_noalloc (no device allocation; a fixed device
scratch shared by the solvers of a device) against msc_solver_get_field, bit
for bit, through the C ABI: every component of a solve whose grid needs
several passes through the 2^20-value scratch, then, in the next call, of a
smaller solve; an undersized output is refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cases
from meteorscattering_amd import _abi
from meteorscattering_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _solver(n, steps, precision):
    p, _, _ = cases.case("lk_ricker_ez")
    nx, ny, nz = n
    p.update(nx=nx, ny=ny, nz=nz, tsteps=steps + 4, probex=[nx // 2], probey=[ny // 2], probez=[nz // 2],
             nprobes=1, numfiles=1)
    wp, nu = cases.trail_maps(p, n0=3e13, r0=1.5, nu0=1e6, centre=(0.5, -0.25))
    g = Solver(p, wp, nu, precision=precision)
    g.step(steps)
    return g


@pytest.mark.parametrize("precision", [32, 64])
def test_noalloc_read_back_equals_get_field(precision):
    lib = _abi.load_library()
    # 3.3 M cells: four passes (the last partial); then 44^3 in one pass
    for n, steps in (((120, 100, 276), 70), ((44, 44, 44), 60)):
        g = _solver(n, steps, precision)
        N = n[0] * n[1] * n[2]
        for comp in range(9):
            want = g.field(comp)
            out = np.full(n, np.nan)
            rc = lib.msc_solver_get_field_noalloc(g._h, comp, _abi.dptr(out), out.size)
            assert rc == 0, lib.msc_last_error().decode()
            assert np.array_equal(out, want), (n, comp)
            if comp >= 6:
                assert np.abs(want).max() > 0  # the trail carries a current
        small = np.empty(N - 1)
        rc = lib.msc_solver_get_field_noalloc(g._h, 0, _abi.dptr(small), C.c_size_t(small.size))
        assert rc != 0 and "grid has" in lib.msc_last_error().decode()
        g.close()
