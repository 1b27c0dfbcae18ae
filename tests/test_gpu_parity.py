"""HIP path (through the C ABI) against the CPU oracle on the golden inputs.

Tolerances: float64 fields within 1e-7 s of the oracle's heap march (the
relaxation stops once a sweep moves no value by more than 1e-14 of itself;
at these sizes the fields agree to ~1e-11 s); receiver times computed from the same field are
bit-identical; Frechet rows from the same field within 1e-6 relative
(ray steps use the device's sin()); float32 fields and receiver times within
1 ms of float64.
"""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def ctx_for(gd):
    c = lib.Context(gd.g)
    c.set_model(gd.velv)
    return c


@pytest.mark.parametrize("name", G.names())
def test_fp64_fields_match_oracle(name):
    gd = G.load(name)
    veln = O.gridder(gd.g, gd.velv)
    with ctx_for(gd) as c:
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP64)
        for s in range(gd.faces.shape[0]):
            T = O.fmm_tele(gd.g, veln, gd.faces[s], gd.fom)
            d = float(np.abs(c.field(s) - T).max())
            assert d <= 1e-7, (s, d)


@pytest.mark.parametrize("name", G.names())
def test_receivers_and_frechet_match_oracle(name):
    gd = G.load(name)
    veln = O.gridder(gd.g, gd.velv)
    nrc = len(gd.rcv)
    with ctx_for(gd) as c:
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP64)
        tr = c.receivers(gd.rcv)
        rp, cols, vals = c.frechet(gd.rcv)
        assert rp.shape == (gd.faces.shape[0] * nrc + 1,)
        for s in range(gd.faces.shape[0]):
            F = c.field(s)
            assert np.array_equal(tr[s], O.srtimes(gd.g, F, gd.rcv))
            orp, oc, ov, ok, _ = O.rpaths(gd.g, veln, F, gd.rcv)
            for i in range(nrc):
                r = s * nrc + i
                a = G.row_dict(cols[rp[r]:rp[r + 1]], vals[rp[r]:rp[r + 1]])
                b = G.row_dict(oc[orp[i]:orp[i + 1]], ov[orp[i]:orp[i + 1]])
                # columns come out ascending
                assert np.all(np.diff(cols[rp[r]:rp[r + 1]]) > 0)
                rel, odd = G.compare_rows(a, b, small=1.1e-6)
                assert rel <= 1e-6 and not odd, (s, i, rel, odd)


@pytest.mark.parametrize("name", G.names())
def test_fp32_within_one_millisecond(name):
    gd = G.load(name)
    with ctx_for(gd) as c:
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP64)
        t64 = c.receivers(gd.rcv)
        f64 = [c.field(s) for s in range(gd.faces.shape[0])]
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP32)
        t32 = c.receivers(gd.rcv)
        assert np.abs(t32.astype(np.float64) - t64).max() <= 1e-3
        for s in range(gd.faces.shape[0]):
            assert np.abs(c.field(s) - f64[s]).max() <= 1e-3


def test_batch_invariance():
    """Each source solved alone and inside a batch gives the same result."""
    gd = G.load(G.names()[1])
    with ctx_for(gd) as c:
        for prec, tol in ((lib.PREC_FP64, 1e-7), (lib.PREC_FP32, 1e-4)):
            c.solve_tele(gd.faces, gd.fom, prec)
            batch = [c.field(s) for s in range(gd.faces.shape[0])]
            for s in (0, 3):
                c.solve_tele(gd.faces[s:s + 1], gd.fom, prec)
                assert float(np.abs(c.field(0) - batch[s]).max()) <= tol


def test_edges_and_errors():
    gd = G.load(G.names()[0])
    veln = O.gridder(gd.g, gd.velv)
    c = lib.Context(gd.g)
    with pytest.raises(lib.FmmError) as e:
        c.solve_tele(gd.faces[:1])
    assert e.value.code == lib.E_STATE
    c.set_model(gd.velv)
    with pytest.raises(lib.FmmError) as e:
        c.L.fmm3d_solve_tele(c._h, 1, np.ascontiguousarray(gd.faces[:1]), 7, 0, 0) and c._check(-1)
    # Part of the bottom face unknown (<= 0): those nodes are solved for.
    # The heap march reaches them only through nodes it pops (the face nodes
    # other than the seed never trigger updates, fm3dt_parallel.f90:
    # 6930-6958), so its order shapes the result there; the check is the
    # oracle's relaxation fixed point, the same discrete system (the
    # single-phase and three-phase relaxations agree on it).
    face = gd.faces[0].copy()
    face[:5, :7] = 0.0
    face[-3:, -2:] = -1.0
    c.solve_tele(face)
    F = c.field(0)
    R1, _ = O.relax1_tele(gd.g, veln, face, 1)
    assert float(np.abs(F - R1).max()) <= 1e-7
    # receivers outside the grid: NaN time, empty Frechet row
    rcv = gd.rcv[:3].copy()
    rcv[1, 1] += 1.0          # far south of the grid
    tr = c.receivers(rcv)
    assert np.isnan(tr[0, 1]) and np.isfinite(tr[0, 0]) and np.isfinite(tr[0, 2])
    rp, cols, vals = c.frechet(rcv)
    assert rp[2] == rp[1] and rp[1] > 0 and rp[3] > rp[2]
    # a too-small buffer reports the size needed
    n = int(rp[-1])
    rp2 = np.zeros(4, np.int64)
    rc = c.L.fmm3d_frechet(c._h, 3, np.ascontiguousarray(rcv), rp2, np.zeros(8, np.int32), np.zeros(8, np.float32), 8)
    assert rc == lib.E_CAPACITY and rp2[-1] == n
    c.close()


def test_pinned_frechet_rows_equal_the_copied_ones():
    """fmm3d_frechet_pinned returns the rows of fmm3d_frechet bit for bit;
    a second call (fewer receivers) reuses the context's arrays."""
    gd = G.load(G.names()[1])
    with ctx_for(gd) as c:
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP32)
        for rcv in (gd.rcv, gd.rcv[:5]):
            rp, cols, vals = c.frechet(rcv)
            prp, pcols, pvals = c.frechet(rcv, pinned=True)
            assert np.array_equal(rp, prp)
            assert np.array_equal(cols, pcols) and np.array_equal(vals, pvals)
