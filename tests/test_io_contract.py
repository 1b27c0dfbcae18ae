"""The file contract: readers/writers round trip and read the reference's
own input files (fm3dt_parallel.f90:6082-6126, 6275-6332; output formats
F:6503-6599, 7133-7176)."""
import os

import numpy as np
import pytest

from fmm3d import io, synth

REF = "/root/reference"


def test_case_round_trip(tmp_path):
    g, velv, srcs, rcv = synth.tele_case((11, 12, 13), 100.0, (2.0, 2.0), (6, 6, 7), 0.05, 3, 5, 3)
    info = synth.write_case(str(tmp_path), g, velv, srcs, rcv)
    cfg = io.read_in(str(tmp_path / "fm3dt_parallel.in"))
    assert (cfg.gridv, cfg.itimes, cfg.fom, cfg.cfd, cfg.rtravel, cfg.frechet) == \
        ("gridc.vtx", "itimes.out", 1, 1, "rtravel.out", "frechet.out")
    g2, v2 = io.read_vtx(str(tmp_path / "gridc.vtx"))
    assert g2 == g
    assert np.array_equal(v2, velv)
    r = io.read_receivers(str(tmp_path / "receivers.dat"))[0]
    assert np.array_equal(r, synth.receivers_rad(rcv))
    f = io.read_itimes(str(tmp_path / "itimes.out"), 3, g.nnx, g.nnz)
    assert np.array_equal(f, info["faces"])
    assert len(io.read_sources(str(tmp_path / "sources.dat"))[0]) == 3


def test_frechet_and_rtravel_round_trip(tmp_path):
    rows = [(1, np.array([3, 7, 9], np.int32), np.array([-1e-3, -2.5e-2, -4e-6], np.float32)),
            (2, np.array([], np.int32), np.array([], np.float32)),
            (3, np.array([11], np.int32), np.array([-0.5], np.float32))]
    p = str(tmp_path / "frechet.out")
    io.write_frechet(p, rows)
    back = io.read_frechet(p)
    assert len(back) == 3
    for (a, b, c), (d, e, f) in zip(rows, back):
        assert a == d and np.array_equal(b, e) and np.array_equal(c, f)
    # 16 bytes per record: 4-byte marker, int32, float32, 4-byte marker
    assert os.path.getsize(p) == 16 * (3 + 4)
    t = np.array([355.2345, 601.5, 1e-3], np.float32)
    io.write_rtravel(str(tmp_path / "r.out"), t)
    assert np.array_equal(io.read_rtravel(str(tmp_path / "r.out")), t)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference tree not present")
def test_reads_reference_inputs():
    cfg = io.read_in(os.path.join(REF, "fm3dt_parallel.in"))
    assert cfg.fom == 1 and cfg.cfd == 1 and cfg.earth == 6371.0
    g, v = io.read_vtx(os.path.join(REF, "gridc.vtx"))
    assert (g.nvr, g.nvx, g.nvz, g.nnr, g.nnx, g.nnz) == (13, 45, 50, 31, 110, 140)
    assert v.shape == (52, 47, 15) and v.min() > 0
    r = io.read_receivers(os.path.join(REF, "receivers.dat"))[0]
    assert r.shape == (20, 3)
    s = io.read_sources(os.path.join(REF, "sources.dat"))[0]
    assert len(s) == 163
    t = io.read_rtravel(os.path.join(REF, "rtimes.dat"))
    assert t.size == 163 * 20 and 300 < t.min() < t.max() < 1200
