"""The resident inversion loop (fm3dt_invert) against sub.sh's sequence of
separate forward runs (fm3dt_gpu), with a stub subinv that rescales the
model by an amount that depends on the iteration it reads from subiter.in.
Both must leave the same files: rtimes.dat (the first pass), subiter.in,
and the last pass's rtravel.out and frechet.out (times within 1e-5 s,
Frechet columns equal and values within 1e-5 relative).  The last pass is
also checked against the oracle: the reference's partitioned solve
(or_fmm_tele_part) on the model subinv left, whose receiver times the
drop-in's default solve reproduces bit for bit."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from fmm3d import build, io, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STUB = r'''
import sys
it = int(open("subiter.in").read().split()[0])
lines = open("gridc.vtx").read().split("\n")
head, vals = lines[:6], [float(v) for v in lines[6:] if v.strip()]
f = 1.0 + 0.02 * it
with open("gridc.vtx", "w") as out:
    out.write("\n".join(head) + "\n")
    for v in vals:
        out.write(repr(v * f) + "\n")
with open("stub.log", "a") as log:
    log.write(f"{it}\n")
'''


def _case(d):
    g, velv, srcs, rcv = synth.tele_case((31, 33, 35), 400.0, (10.0, 10.0), (10, 11, 12), 0.05, 4, 10, 31)
    info = synth.write_case(str(d), g, velv, srcs, rcv, fom=1, cfd=1)
    shutil.copy(d / "gridc.vtx", d / "gridi.vtx")
    (d / "stub.py").write_text(STUB)
    return info["faces"]


def _frechet_close(a, b):
    ra, rb = io.read_frechet(str(a)), io.read_frechet(str(b))
    assert len(ra) == len(rb)
    for (ia, ca, va), (ib, cb, vb) in zip(ra, rb):
        assert ia == ib and np.array_equal(ca, cb)
        assert np.all(np.abs(va - vb) <= 1e-5 * np.maximum(np.abs(va), 1e-6))


def test_resident_loop_matches_subsh_sequence(tmp_path):
    if not (os.path.exists(build.EXE) and os.path.exists(build.INVERT_EXE)):
        pytest.skip("Fortran hosts not built")
    stub = f"{sys.executable} stub.py"
    a, b = tmp_path / "resident", tmp_path / "subsh"
    a.mkdir(); b.mkdir()
    faces = _case(a)
    _case(b)
    print("fm3dt_invert, 2 iterations", flush=True)
    r = subprocess.run([build.INVERT_EXE, "2", stub], cwd=str(a), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Running Inversion...") == 2 and r.stdout.count("Parallel FMM Done") == 3
    # sub.sh's sequence, one process per forward pass
    print("fm3dt_gpu x 3 with the stub between", flush=True)
    shutil.copy(b / "gridi.vtx", b / "gridc.vtx")
    (b / "subiter.in").write_text("1\n")
    run = lambda: subprocess.run([build.EXE], cwd=str(b), capture_output=True, text=True, timeout=120)
    assert run().returncode == 0
    shutil.copy(b / "rtravel.out", b / "rtimes.dat")
    first = io.read_rtravel(str(b / "rtravel.out"))
    for it in (1, 2):
        assert subprocess.run(stub.split(), cwd=str(b), timeout=60).returncode == 0
        assert run().returncode == 0
        (b / "subiter.in").write_text(f"{it + 1}\n")
    assert (a / "stub.log").read_text() == "1\n2\n" == (b / "stub.log").read_text()
    assert (a / "subiter.in").read_text().split() == ["3"]
    assert (a / "gridc.vtx").read_text() == (b / "gridc.vtx").read_text()
    ta, tb = io.read_rtravel(str(a / "rtimes.dat")), io.read_rtravel(str(b / "rtimes.dat"))
    assert np.abs(ta - tb).max() <= 1e-5
    la, lb = io.read_rtravel(str(a / "rtravel.out")), io.read_rtravel(str(b / "rtravel.out"))
    assert np.abs(la - lb).max() <= 1e-5
    # the model changed, and so did the times: the loop re-gridded each pass
    assert np.abs(la - first).max() > 1e-3
    _frechet_close(a / "frechet.out", b / "frechet.out")
    # the last pass against the oracle on the final model
    g, velv = io.read_vtx(str(a / "gridc.vtx"))
    rcv = io.read_receivers(str(a / "receivers.dat"))[0]
    want = np.concatenate([O.srtimes(g, O.fmm_tele_part(g, velv, f, 1)[0], rcv) for f in faces])
    assert np.array_equal(la, want), float(np.abs(la - want).max())


def test_resident_loop_stops_when_subinv_fails(tmp_path):
    if not os.path.exists(build.INVERT_EXE):
        pytest.skip("fm3dt_invert not built")
    _case(tmp_path)
    r = subprocess.run([build.INVERT_EXE, "3", "false"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "failed" in r.stdout
    assert (tmp_path / "rtimes.dat").exists() and (tmp_path / "subiter.in").read_text().split() == ["1"]
