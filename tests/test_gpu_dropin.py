"""The Fortran drop-in host (fm3dt_gpu) in a directory laid out like the
reference's working directory: it must read the same files and write
rtravel.out and frechet.out in the reference's formats, with the values the
library computes -- by default the reference's own partitioned result, so
on the golden inputs its files carry the reference program's outputs."""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

import goldens as G
from fmm3d import build, io, lib, synth

pytestmark = pytest.mark.gpu


def test_dropin_writes_the_reference_file_contract(tmp_path):
    if not os.path.exists(build.EXE):
        pytest.skip("fm3dt_gpu not built")
    g, velv, srcs, rcv = synth.tele_case((31, 33, 35), 400.0, (10.0, 10.0), (10, 11, 12), 0.05, 5, 12, 21)
    info = synth.write_case(str(tmp_path), g, velv, srcs, rcv, fom=1, cfd=1)
    env = dict(os.environ, FMM3D_BATCH="2", FMM3D_DOMAIN="whole")
    r = subprocess.run([build.EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    tr = io.read_rtravel(str(tmp_path / "rtravel.out")).reshape(5, 12)
    rows = io.read_frechet(str(tmp_path / "frechet.out"))
    assert len(rows) == 5 * 12
    rr = io.read_receivers(str(tmp_path / "receivers.dat"))[0]
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_tele(info["faces"], lib.ORDER_MIXED, lib.PREC_FP64)
        t = c.receivers(rr)
        rp, cols, vals = c.frechet(rr)
    assert np.array_equal(tr, t)
    for k, (rid, rc, rv) in enumerate(rows):
        assert rid == k % 12 + 1
        assert np.array_equal(rc, cols[rp[k]:rp[k + 1]]) and np.array_equal(rv, vals[rp[k]:rp[k + 1]])
    # byte for byte the file sequential unformatted writes give
    ref = tmp_path / "frechet_ref.out"
    io.write_frechet(str(ref), [(k % 12 + 1, cols[rp[k]:rp[k + 1]], vals[rp[k]:rp[k + 1]]) for k in range(len(rows))])
    assert (tmp_path / "frechet.out").read_bytes() == ref.read_bytes()


def test_dropin_rejects_inconsistent_array_counts(tmp_path):
    if not os.path.exists(build.EXE):
        pytest.skip("fm3dt_gpu not built")
    g, velv, srcs, rcv = synth.tele_case((11, 12, 13), 100.0, (2.0, 2.0), (6, 6, 7), 0.05, 2, 4, 3)
    synth.write_case(str(tmp_path), g, velv, srcs, rcv)
    with open(tmp_path / "receivers.dat") as f:
        lines = f.read().split("\n")
    lines[0] = "2"
    with open(tmp_path / "receivers.dat", "w") as f:
        f.write("\n".join(lines))
    r = subprocess.run([build.EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert "inconsistent" in r.stdout


def _golden_cases():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(G.GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return {c[0]: c for c in mod.CASES}


@pytest.mark.parametrize("name", G.names())
def test_dropin_default_reproduces_the_reference_files(tmp_path, name):
    """The golden's inputs regenerated from their seed (same SHA-256 as the
    files the reference read), fm3dt_gpu run with its defaults: rtravel.out
    equals the reference's values exactly, frechet.out rows within 1e-4 of
    the reference's (entries under 8e-6 may differ in presence: the
    reference cuts each rank's partial row at 1e-6 before summing)."""
    if not os.path.exists(build.EXE):
        pytest.skip("fm3dt_gpu not built")
    gd = G.load(name)
    _, nn, depth, span, nv, amp, nsrc, nrc, seed, fom = _golden_cases()[name]
    g, velv, srcs, rcv = synth.tele_case(nn, depth, span, nv, amp, nsrc, nrc, seed)
    info = synth.write_case(str(tmp_path), g, velv, srcs, rcv, fom=fom, cfd=1)
    assert info["sha256"] == gd.sha256
    env = dict(os.environ, FMM3D_BATCH="3")
    env.pop("FMM3D_DOMAIN", None)
    env.pop("FMM3D_PRECISION", None)
    r = subprocess.run([build.EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    tr = io.read_rtravel(str(tmp_path / "rtravel.out")).reshape(nsrc, nrc)
    assert np.array_equal(tr, gd.rtravel)
    rows = io.read_frechet(str(tmp_path / "frechet.out"))
    assert len(rows) == len(gd.rows)
    for (rid, rc, rv), (gid, gc, gv) in zip(rows, gd.rows):
        assert rid == gid
        rel, odd = G.compare_rows(G.row_dict(rc, rv), G.row_dict(gc, gv), small=8e-6)
        assert rel <= 1e-4 and not odd, (rid, rel, odd)
