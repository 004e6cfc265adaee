"""Write golden fixtures from the reference solver built from its own source.

Run where /root/reference exists (the fixtures it writes are data and travel
with the tests; the reference itself does not):

    make -C oracle ref
    python tests/golden/make_golden.py

For each case in cases.py it writes a run directory with the project's own
writer (meteorscattering_amd.runio; inputs.dat carries the fork's doB0),
runs oracle/_ref/meteor-fork there (stack VLAs need `ulimit -s unlimited`),
parses output_NTF.dat, Probe.dat, output_{E,H,J}.dat and log.txt, and stores
inputs plus compact summaries in tests/golden/<case>.npz.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from meteorscattering_amd import runio  # noqa: E402
import cases  # noqa: E402
import summary  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "meteor-fork")


def run_reference(rundir: str) -> None:
    cmd = f"ulimit -s unlimited && cd '{rundir}' && OMP_NUM_THREADS=4 '{REF_BIN}'"
    r = subprocess.run(["bash", "-c", cmd], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed ({r.returncode}): {r.stderr[-2000:]}")


def parse_log_emax(text: str) -> list[float]:
    return [float(m) for m in re.findall(r"Maximum abs\(Ex\) is ([^\s]+)", text)]


def make(name: str) -> str:
    p, wp, nu = cases.case(name)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    with tempfile.TemporaryDirectory() as d:
        runio.write_run_dir(d, p, wp, nu)
        run_reference(d)
        ntf = runio.read_ntf(os.path.join(d, "output_NTF.dat"), nx, ny, nz)
        probe = runio.read_probe(os.path.join(d, "Probe.dat"), p["tsteps"])
        ef = runio.read_frames(os.path.join(d, "output_E.dat"), nx, ny, nz)
        hf = runio.read_frames(os.path.join(d, "output_H.dat"), nx, ny, nz)
        jf = runio.read_frames(os.path.join(d, "output_J.dat"), nx, ny, nz, with_inc=False)
        log = open(os.path.join(d, "log.txt")).read()
    out = {
        "params": np.array(json.dumps(p)),
        "wp": wp, "nu": nu,
        "ntf_header": np.array([ntf["nntf"], ntf["ntfb"]], dtype=np.int64),
        "fntf": ntf["fntf"], "Eydft": ntf["Eydft"], "Ezdft": ntf["Ezdft"],
        "probes": probe["data"], "probe_index": probe["index"],
        "emax_log": np.array(parse_log_emax(log)),
        "nframes": np.array([ef.shape[0], hf.shape[0], jf.shape[0]]),
        "jframes_absmax": np.array(np.abs(jf).max() if jf.size else 0.0),
    }
    b = ntf["ntfb"]
    for fname in runio.ntf_face_names():
        a = ntf[fname]
        out["proj_" + fname] = summary.project(fname, a)
        out["carve_" + fname] = a[b:a.shape[0] - b + 1, b:a.shape[1] - b + 1, :]
    for kind, frames in (("E", ef), ("H", hf)):
        for n, fr in enumerate(frames):
            parts = runio.split_frame(fr, nx, ny, nz)
            for key in ("slicez", "slicey", "slicex"):
                out[f"proj_{kind}{n}_{key}"] = summary.project(f"{kind}{n}_{key}", parts[key])
            out[f"{kind}{n}_inc"] = np.concatenate([parts["inc1"], parts["inc2"]])
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


def main(argv):
    names = argv[1:] or cases.CASES
    if not os.path.exists(REF_BIN):
        raise SystemExit(f"{REF_BIN} missing: run `make -C oracle ref` where /root/reference exists")
    for n in names:
        path = make(n)
        print(f"{n}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv)
