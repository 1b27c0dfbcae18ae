"""201^3 teleseismic workload (0-600 km, 15x15 deg, 25^3 vertices, +-3 %):
the float64 HIP solve against the oracle's heap march for one source, and
the float32 batched solve against float64 for a batch, at the full size."""
import numpy as np
import pytest

from fmm3d import lib, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    g, velv, srcs, rcv = synth.config_201(nsrc=8, nrc=300, seed=2)
    faces = np.stack([synth.tele_bottom_face(g, s) for s in srcs])
    return g, velv, faces, synth.receivers_rad(rcv)


def test_fp64_matches_oracle_at_full_size(case):
    g, velv, faces, rcv = case
    veln = O.gridder(g, velv)
    T = O.fmm_tele(g, veln, faces[0], 1)
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_tele(faces[:1], lib.ORDER_MIXED, lib.PREC_FP64)
        F = c.field(0)
        # the relaxation stops once a sweep moves no value by more than 1e-14 of
        # itself; over ~10^3 dependent hops that leaves a few 1e-8 s at most
        d = float(np.abs(F - T).max())
        assert d <= 1e-6, d
        # invariants: the bottom face is kept, nothing precedes the face minimum
        assert np.array_equal(F[-1], faces[0])
        assert F.min() >= faces[0].min()


def test_fp32_batch_within_one_millisecond(case):
    g, velv, faces, rcv = case
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP64)
        t64 = c.receivers(rcv)
        c.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP32)
        t32 = c.receivers(rcv)
    assert np.isfinite(t64).all()
    assert np.abs(t32.astype(np.float64) - t64).max() <= 1e-3
