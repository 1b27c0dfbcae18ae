"""Teleseismic sources staged in device memory (fmm3d_stage_tele_sides) and
solved from there (fmm3d_solve_tele_staged) give the fields and receiver
times of the one-call solves on the same faces, solve after solve; a solve
of another kind re-stages, the partitioned solve and a fresh context leave
nothing staged."""
import numpy as np
import pytest

from fmm3d import lib, synth

pytestmark = pytest.mark.gpu


def _case():
    g, velv, tele, rcv = synth.tele_case((41, 45, 39), 500.0, (12.0, 11.0), (9, 10, 11), 0.08, 5, 20, 7)
    parts = [synth.tele_faces(g, s) for s in tele]
    F, X, Z, S = (np.stack([p[k] for p in parts]) for k in range(4))
    return g, velv, F, X, Z, S, synth.receivers_rad(rcv)


@pytest.mark.parametrize("sides", [True, False])
def test_staged_solves_match_one_call_solves(sides):
    g, velv, F, X, Z, S, rcv = _case()
    with lib.Context(g) as d, lib.Context(g) as c:
        d.set_model(velv)
        c.set_model(velv)
        if sides:
            c.stage_tele(F, X, Z, S)
        else:
            c.stage_tele(F)
        for prec, tol in ((lib.PREC_FP64, 1e-9), (lib.PREC_FP32, 1e-3), (lib.PREC_FP32, 1e-3)):
            if sides:
                d.solve_tele_sides(F, X, Z, S, lib.ORDER_MIXED, prec)
            else:
                d.solve_tele(F, lib.ORDER_MIXED, prec)
            c.solve_tele_staged(lib.ORDER_MIXED, prec)
            assert c.nsrc == F.shape[0]
            for s in range(F.shape[0]):
                assert float(np.abs(c.field(s) - d.field(s)).max()) <= tol, (sides, prec, s)
            assert float(np.abs(c.receivers(rcv) - d.receivers(rcv)).max()) <= tol
            rp, cols, vals = c.frechet(rcv)
            rp2, cols2, vals2 = d.frechet(rcv)
            assert rp[-1] > 0 and abs(int(rp[-1]) - int(rp2[-1])) <= max(10, int(rp2[-1]) // 1000)


def test_staging_is_replaced_and_checked():
    g, velv, F, X, Z, S, rcv = _case()
    with lib.Context(g) as c:
        c.set_model(velv)
        with pytest.raises(lib.FmmError):
            c.solve_tele_staged(lib.ORDER_MIXED, lib.PREC_FP32)     # nothing staged
        c.stage_tele(F, X, Z, S)
        # a one-call solve of two sources stages them in place of the five
        c.solve_tele(F[:2], lib.ORDER_MIXED, lib.PREC_FP64)
        ref = [c.field(s) for s in range(2)]
        c.solve_tele_staged(lib.ORDER_MIXED, lib.PREC_FP64)
        assert c.nsrc == 2
        for s in range(2):
            assert float(np.abs(c.field(s) - ref[s]).max()) <= 1e-9
        c.solve_tele(F[:2], lib.ORDER_MIXED, lib.PREC_FP64, partition=True)
        with pytest.raises(lib.FmmError):
            c.solve_tele_staged(lib.ORDER_MIXED, lib.PREC_FP64)     # the partitioned solve leaves none
        bad = F.copy()
        bad[1, 3, 4] = np.nan
        with pytest.raises(lib.FmmError):
            c.stage_tele(bad)
