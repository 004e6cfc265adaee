"""CPU tests of the C-ABI library: it loads, exports every symbol the public
header declares, and its host-side entry points (parameter validation,
run-directory I/O) agree with the reference's byte layouts.  No compute
call needs a GPU here."""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import cases
from meteorscattering_amd import _abi, runio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "msc.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.fail("libmsc.so not built: run __graft_entry__.build()")
    return _abi.load_library()


def test_every_declared_symbol_is_exported(lib):
    names = _abi.exported_symbols(HEADER)
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)  # resolvable through the dynamic loader


def test_params_struct_layout_matches_header(lib):
    """The ctypes mirror has the C struct's size: msc_params_default writes
    exactly sizeof(msc_params) bytes into a guarded buffer."""
    size = C.sizeof(_abi.MscParams)
    buf = (C.c_ubyte * (size + 64))()
    for n in range(size, size + 64):
        buf[n] = 0xAB
    lib.msc_params_default(C.cast(buf, C.POINTER(_abi.MscParams)))
    assert all(buf[n] == 0xAB for n in range(size, size + 64))
    p = _abi.struct_to_params(_abi.MscParams.from_buffer(buf))
    assert p["nx"] == 201 and p["f0"] == 150e6 and p["nntf"] == 15 and p["sourcepol"] == 3


def _run_case():
    p, wp, nu = cases.case("lk_ricker_ez")
    return p, wp, nu


def test_c_reader_reads_python_writer(lib, tmp_path):
    p, wp, nu = _run_case()
    runio.write_run_dir(str(tmp_path), p, wp, nu)
    s = _abi.MscParams()
    a, b = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    rc = lib.msc_read_run_dir(str(tmp_path).encode(), C.byref(s), C.byref(a), C.byref(b))
    assert rc == 0, lib.msc_last_error()
    q = _abi.struct_to_params(s)
    for k in ("dopml", "savefields", "f0", "fsig", "sourcetype", "sourcepol", "gridfactor", "dt", "tsteps",
              "nx", "ny", "nz", "numfiles", "nprobes", "nntf", "doplate", "platesize", "doB0"):
        assert q[k] == p[k], k
    assert q["probex"] == p["probex"] and q["fntf"] == p["fntf"]
    assert q["B0"] == pytest.approx(p["B0"], abs=0)
    n = p["nx"] * p["ny"]
    np.testing.assert_array_equal(np.ctypeslib.as_array(a, (n,)), wp.ravel())
    np.testing.assert_array_equal(np.ctypeslib.as_array(b, (n,)), nu.ravel())
    lib.msc_free(a)
    lib.msc_free(b)


def test_c_writer_bytes_equal_python_writer(lib, tmp_path):
    p, wp, nu = _run_case()
    p = dict(p, pmllen=11, nu_scale_height=7000.0, j_collocated=1)
    d1, d2 = tmp_path / "c", tmp_path / "py"
    d1.mkdir()
    runio.write_run_dir(str(d2), p, wp, nu)
    s = _abi.params_to_struct(p)
    rc = lib.msc_write_run_dir(str(d1).encode(), C.byref(s), _abi.dptr(np.ascontiguousarray(wp)),
                               _abi.dptr(np.ascontiguousarray(nu)))
    assert rc == 0, lib.msc_last_error()
    for f in ("inputs.dat", "B0.dat", "wp.dat", "nu.dat"):
        assert (d1 / f).read_bytes() == (d2 / f).read_bytes(), f
    q = runio.decode_inputs_dat((d1 / "inputs.dat").read_bytes())
    assert q["pmllen"] == 11 and q["nu_scale_height"] == 7000.0 and q["j_collocated"] == 1
    # the C reader takes the record back, and still reads a version-1 record
    s2 = _abi.MscParams()
    a, b = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    assert lib.msc_read_run_dir(str(d1).encode(), C.byref(s2), C.byref(a), C.byref(b)) == 0
    assert s2.j_collocated == 1 and s2.pmllen == 11
    lib.msc_free(a)
    lib.msc_free(b)
    v1 = bytearray((d1 / "inputs.dat").read_bytes()[:-4])
    k = v1.find(runio.EXT_MAGIC)
    v1[k + 4:k + 8] = struct.pack("<i", 1)
    (d1 / "inputs.dat").write_bytes(bytes(v1))
    assert lib.msc_read_run_dir(str(d1).encode(), C.byref(s2), C.byref(a), C.byref(b)) == 0
    assert s2.j_collocated == 0 and s2.pmllen == 11 and runio.decode_inputs_dat(bytes(v1))["j_collocated"] == 0
    lib.msc_free(a)
    lib.msc_free(b)


def test_missing_doB0_reads_as_zero(lib, tmp_path):
    """setup/meteorrun.m:44-70 never writes doB0; the fork reads past EOF."""
    p, wp, nu = _run_case()
    runio.write_run_dir(str(tmp_path), p, wp, nu)
    (tmp_path / "inputs.dat").write_bytes(runio.encode_inputs_dat(p, with_doB0=False))
    s = _abi.MscParams()
    a, b = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    assert lib.msc_read_run_dir(str(tmp_path).encode(), C.byref(s), C.byref(a), C.byref(b)) == 0
    assert s.doB0 == 0
    lib.msc_free(a)
    lib.msc_free(b)


def test_truncated_and_invalid_inputs_fail_cleanly(lib, tmp_path):
    p, wp, nu = _run_case()
    runio.write_run_dir(str(tmp_path), p, wp, nu)
    full = (tmp_path / "inputs.dat").read_bytes()
    s = _abi.MscParams()
    a, b = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    (tmp_path / "inputs.dat").write_bytes(full[:60])
    assert lib.msc_read_run_dir(str(tmp_path).encode(), C.byref(s), C.byref(a), C.byref(b)) == -2
    assert b"truncated" in lib.msc_last_error()
    # numfiles > tsteps: the reference divides by zero (meteor-fork.cpp:1616)
    q = dict(p, numfiles=p["tsteps"] + 1)
    (tmp_path / "inputs.dat").write_bytes(runio.encode_inputs_dat(q))
    assert lib.msc_read_run_dir(str(tmp_path).encode(), C.byref(s), C.byref(a), C.byref(b)) == -1
    assert b"numfiles" in lib.msc_last_error()
    # short wp.dat
    (tmp_path / "inputs.dat").write_bytes(full)
    (tmp_path / "wp.dat").write_bytes(wp.tobytes()[:100])
    assert lib.msc_read_run_dir(str(tmp_path).encode(), C.byref(s), C.byref(a), C.byref(b)) == -2
    # probe on index 0 (the probe average reads index-1)
    bad = _abi.params_to_struct(dict(p, probex=[0, 5]))
    assert lib.msc_params_validate(C.byref(bad)) == -1


def test_solver_create_validates_before_touching_the_device(lib):
    p, wp, nu = _run_case()
    bad = _abi.params_to_struct(dict(p, nx=10))
    h = C.c_void_p()
    rc = lib.msc_solver_create(C.byref(bad), _abi.dptr(wp), _abi.dptr(nu), 32, C.byref(h))
    assert rc != 0 and not h.value
    rc = lib.msc_solver_create(C.byref(_abi.params_to_struct(p)), _abi.dptr(wp), _abi.dptr(nu), 16, C.byref(h))
    assert rc == -1 and b"precision" in lib.msc_last_error()


def test_inputs_dat_layout_matches_reference_reader_offsets():
    """Field offsets of inputs.dat as the reference reads them (meteor-fork.cpp:84-115)."""
    p, _, _ = _run_case()
    b = runio.encode_inputs_dat(p)
    assert struct.unpack_from("<2i", b, 0) == (p["dopml"], p["savefields"])
    assert struct.unpack_from("<d", b, 8)[0] == p["f0"]
    assert struct.unpack_from("<3i", b, 52) == (p["nx"], p["ny"], p["nz"])
    np_ = p["nprobes"]
    off = 4 * 2 + 8 * 2 + 4 * 2 + 8 * 2 + 4 * 4 + 8 * 2 + 8 * 2 + 8 * 3 + 8 + 4 * 3
    assert off == 140
    assert struct.unpack_from("<i", b, off - 4)[0] == np_
    off += 12 * np_
    assert struct.unpack_from("<i", b, off)[0] == p["nntf"]
    off += 4 + 8 * p["nntf"]
    assert struct.unpack_from("<i", b, off)[0] == p["doplate"]
    assert struct.unpack_from("<i", b, off + 12)[0] == p["doB0"]
    assert len(b) == off + 16


def test_ntf_file_roundtrip():
    p, _, _ = _run_case()
    nx, ny, nz, M = p["nx"], p["ny"], p["nz"], p["nntf"]
    n = 4 * M + 8 * 2 * M * (ny * nz + nx * nz + nx * ny)
    pay = np.random.default_rng(1).standard_normal(n)
    b = runio.encode_ntf(M, 15, p["fntf"], pay)
    d = runio.parse_ntf(b, nx, ny, nz)
    assert d["ntfb"] == 15 and list(d["fntf"]) == p["fntf"]
    np.testing.assert_array_equal(d["Eydft"], pay[:2 * M])
    np.testing.assert_array_equal(d["My_front"].ravel(), pay[4 * M:4 * M + 2 * M * ny * nz])


def test_copied_executable_still_finds_the_library(tmp_path):
    """setup/meteorrun.m:162 copies the executable alone into the run
    directory; the dynamic loader must still resolve libmsc.so."""
    exe = os.path.join(os.path.dirname(_abi.LIB_PATH), "meteor")
    if not os.path.exists(exe):
        pytest.fail("meteor not built: run __graft_entry__.build()")
    dst = tmp_path / "meteor"
    dst.write_bytes(open(exe, "rb").read())
    dst.chmod(0o755)
    out = subprocess.run(["ldd", str(dst)], capture_output=True, text=True).stdout
    line = [ln for ln in out.splitlines() if "libmsc.so" in ln]
    assert line and "not found" not in line[0], out


def test_solve_footprint_is_host_only_and_exact_in_form(lib):
    """msc_solve_footprint plans the block without a device: FP32 holds the
    29 N-sized arrays (three rotating sets of E, H and J, two media arrays),
    the 18 CPML psi slabs (12 for H, 6 for E; 2L x n x n each on a cubic
    grid) and the NTFF sample buffer (K <= 64 slots of 4 values per face
    point), plus under 4 MB of NTFF sums, frames and tables; FP64 nearly
    twice as much; invalid parameters give 0."""
    p, _, _ = _run_case()
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    n = nx * ny * nz
    L = p.get("pmllen") or 10
    f32 = lib.msc_solve_footprint(C.byref(_abi.params_to_struct(p)), 32)
    f64 = lib.msc_solve_footprint(C.byref(_abi.params_to_struct(p)), 64)
    base = 4 * (29 * n + 18 * 2 * L * nx * ny)
    u = 4 * 2 * (ny * nz + nx * nz + nx * ny)
    assert base < f32 < base + 4 * 64 * u + (4 << 20)
    assert 1.6 < f64 / f32 < 2.0  # the FP64 NTFF sums and tables are common to both
    assert lib.msc_solve_footprint(C.byref(_abi.params_to_struct(dict(p, nx=10))), 32) == 0
    assert lib.msc_solve_footprint(C.byref(_abi.params_to_struct(p)), 16) == 0


def test_non_finite_media_are_refused_before_any_device_work():
    """msc_solver_create checks the parameters and every media value before
    it looks for a device: NaN or infinity in wp.dat/nu.dat is refused."""
    import ctypes as C
    import math

    from meteorscattering_amd import configs

    lib = _abi.load_library()
    p = configs.base_params((44, 44, 44), 0.5, 40)
    wp = np.ones((44, 44)) * 1e7
    nu = np.ones((44, 44)) * 1e5
    for arr, val in ((wp, math.nan), (nu, math.inf)):
        bad = arr.copy()
        bad[5, 7] = val
        a, b = (bad, nu) if arr is wp else (wp, bad)
        s = _abi.params_to_struct(p)
        h = C.c_void_p()
        rc = lib.msc_solver_create(C.byref(s), _abi.dptr(a), _abi.dptr(b), 32, C.byref(h))
        assert rc == -1 and b"not finite" in lib.msc_last_error()


def test_gang_layouts_own_every_cell_once(lib):
    """msc_layout_check over every trail of the 48-trail sweep and of the
    48-request stream as a server gang lays them out: one-step tiles of 32
    and 64 lanes, FP32 and FP64 two-step tiles, the default tiling and a
    short-chunk table (4-plane one-step chunks, 3-row CPML tiles, 16-plane
    two-step chunks).  Each descriptor must decode inside its tables and its
    workgroup, carry the CPML axes its cells meet, and each step's tiles must
    own every cell once.  The gang sizes (workgroups of one half-iteration
    with every trail resident) are printed for the record."""
    from meteorscattering_amd import configs

    short = _abi.Tiling(7, 3, 4, 16)
    trails = [p for p, _, _ in configs.trail_sweep(count=48, maps=False)]
    trails += [p for p, _, _ in configs.request_stream(count=48, maps=False)]
    cases = [(32, 32, None), (32, 64, None), (64, 32, None), (32, 32, short), (32, 64, short)]
    for prec, zl, til in cases:
        one = outer = pair = 0
        for n, p in enumerate(trails):
            if prec == 64 and n % 4:
                continue  # FP64 changes only the two-step tiles: every fourth trail
            rep = _abi.LayoutReport()
            rc = lib.msc_layout_check(C.byref(_abi.params_to_struct(p)), prec, zl,
                                      C.byref(til) if til is not None else None, C.byref(rep))
            assert rc == 0, (n, prec, zl, lib.msc_last_error().decode())
            assert rep.cells == p["nx"] * p["ny"] * p["nz"]
            one += rep.one_step_tiles
            outer += rep.outer_tiles if rep.pair_tiles else rep.one_step_tiles
            pair += rep.pair_tiles
        print(f"fp{prec} {zl} lanes {'short' if til else 'default'}: one-step tiling {one} workgroups; "
              f"two-step gang: {outer} one-step + {pair} fdtd_pair workgroups")
    # what the check refuses
    p = trails[0]
    assert lib.msc_layout_check(C.byref(_abi.params_to_struct(p)), 32, 48, None, None) != 0
    assert lib.msc_layout_check(C.byref(_abi.params_to_struct(p)), 16, 32, None, None) != 0
