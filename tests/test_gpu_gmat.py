"""fmm3d_gmat_* (the Frechet matrix resident on the device with subinv's
products) against the oracle's restatement of subinv (float32 sums in its
order): bit for bit on the reference's own rows, on a random matrix of the
201^3 workload's size (15 000 rows x 19 683 vertices), and on the rows the
GPU itself traces; malformed rows are refused."""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib
from oracle import oracle as O
from test_oracle_gmat import golden_matrix, random_matrix

pytestmark = pytest.mark.gpu


def _check(rp, cols, vals, ncols, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=len(rp) - 1).astype(np.float32)
    a = rng.normal(size=ncols).astype(np.float32)
    cd = rng.uniform(0.5, 2.0, size=len(rp) - 1).astype(np.float32)
    with lib.FrechetMatrix(rp, cols, vals, ncols) as G_:
        assert np.array_equal(G_.gtx(d), O.gmat_gtx(rp, cols, vals, ncols, d))
        assert np.array_equal(G_.gx(a, cd), O.gmat_gx(rp, cols, vals, a, cd))
        assert np.array_equal(G_.gx(a), O.gmat_gx(rp, cols, vals, a))


def test_gmat_products_equal_subinv_order():
    _check(*golden_matrix(), 1)
    _check(*random_matrix(15000, 19683, 250, 5), 2)
    # rows traced on the device
    gd = G.load("tele_31x33x35_mixed")
    with lib.Context(gd.g) as c:
        c.set_model(gd.velv)
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP32)
        rp, cols, vals = c.frechet(gd.rcv)
    _check(rp, cols, vals, (gd.g.nvr + 2) * (gd.g.nvx + 2) * (gd.g.nvz + 2), 3)


def test_gmat_rejects_malformed_rows():
    rp = np.array([0, 2], np.int64)
    with pytest.raises(lib.FmmError) as e:
        lib.FrechetMatrix(rp, np.array([5, 3], np.int32), np.ones(2, np.float32), 10)
    assert e.value.code == lib.E_ARG and "ascend" in str(e.value)
    with pytest.raises(lib.FmmError) as e:
        lib.FrechetMatrix(rp, np.array([3, 11], np.int32), np.ones(2, np.float32), 10)
    assert e.value.code == lib.E_ARG
