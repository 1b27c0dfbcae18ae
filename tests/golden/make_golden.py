"""Write the golden fixtures: seeded synthetic inputs run through the
reference program compiled from its own sources (oracle/_ref/fm3dt_parallel,
built by `make -C oracle ref`) under `mpirun -np 9` (its hard-coded 2x2x2
layout plus the writer rank).

Each fixture (tests/golden/<name>.npz) holds the inputs themselves (grid
header, vertex model, bottom faces, receivers) and the reference outputs
(rtravel.out values and frechet.out rows), plus the SHA-256 of the input
files as written, so tests never need the reference at run time.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "parallel-fast-marching-method_amd"))

from fmm3d import io, synth  # noqa: E402

REF_EXE = os.path.join(REPO, "oracle", "_ref", "fm3dt_parallel")
MPIRUN = "/opt/conda/bin/mpirun"

# (name, nodes, depth km, span deg, vertices, amplitude, nsrc, nrc, seed, fom)
CASES = [
    ("tele_31x33x35_mixed", (31, 33, 35), 400.0, (10.0, 10.0), (10, 11, 12), 0.05, 8, 20, 11, 1),
    ("tele_41x45x39_mixed", (41, 45, 39), 500.0, (12.0, 11.0), (9, 10, 11), 0.08, 6, 30, 7, 1),
    ("tele_25x27x29_first", (25, 27, 29), 300.0, (8.0, 9.0), (8, 8, 9), 0.05, 6, 16, 5, 0),
]


def make(name, nn, depth, span, nv, amp, nsrc, nrc, seed, fom, keep=None):
    g, velv, srcs, rcv = synth.tele_case(nn, depth, span, nv, amp, nsrc, nrc, seed)
    d = keep or tempfile.mkdtemp(prefix="fmm3d_golden_")
    info = synth.write_case(d, g, velv, srcs, rcv, fom=fom, cfd=1)
    r = subprocess.run([MPIRUN, "-np", "9", REF_EXE], cwd=d, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"{name}: reference failed\n{r.stdout}\n{r.stderr}")
    tr = io.read_rtravel(os.path.join(d, "rtravel.out"))
    rows = io.read_frechet(os.path.join(d, "frechet.out"))
    rid = np.array([r_[0] for r_ in rows], dtype=np.int32)
    rlen = np.array([len(r_[1]) for r_ in rows], dtype=np.int64)
    rcols = np.concatenate([r_[1] for r_ in rows]).astype(np.int32)
    rvals = np.concatenate([r_[2] for r_ in rows]).astype(np.float32)
    # the receivers exactly as the reference read them back from receivers.dat
    rcv_rad = io.read_receivers(os.path.join(d, "receivers.dat"))[0]
    out = os.path.join(HERE, name + ".npz")
    np.savez_compressed(
        out,
        grid=np.array(json.dumps(dataclasses.asdict(g))),
        velv=velv,
        faces=info["faces"],
        rcv_rad=rcv_rad,
        fom=np.int32(fom),
        rtravel=tr,
        frechet_rid=rid,
        frechet_len=rlen,
        frechet_cols=rcols,
        frechet_vals=rvals,
        sha256=np.array(json.dumps(info["sha256"])),
    )
    if keep is None:
        shutil.rmtree(d, ignore_errors=True)
    print(f"{name}: {len(tr)} times, {len(rows)} rows, {rcols.size} entries -> {out}")


def main():
    if not os.path.exists(REF_EXE):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    for c in CASES:
        make(*c)


if __name__ == "__main__":
    main()
