"""fmm3d_receivers_noalloc -- receiver times staged in the library's fixed
device array, with no device allocation -- against fmm3d_receivers,
through the C ABI: bit for bit on two calls in a row, the second with a
smaller input (fewer sources resident, fewer receivers, float32 fields),
and an input beyond the array's 8 MiB is refused with E_ARG."""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib

pytestmark = pytest.mark.gpu


def test_noalloc_receivers_equal_receivers():
    gd = G.load(G.names()[1])
    with lib.Context(gd.g) as c:
        c.set_model(gd.velv)
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP64)
        rcv = gd.rcv.copy()
        rcv[3, 1] += 1.0          # one receiver outside the grid: NaN in both
        a = c.receivers(rcv)
        b = c.receivers_noalloc(rcv)
        assert a.shape == b.shape == (gd.faces.shape[0], len(rcv))
        assert np.array_equal(a, b, equal_nan=True) and np.isnan(b[:, 3]).all()
        # second call, smaller: 3 sources in float32, 7 receivers
        c.solve_tele(gd.faces[:3], gd.fom, lib.PREC_FP32)
        a2 = c.receivers(gd.rcv[:7])
        b2 = c.receivers_noalloc(gd.rcv[:7])
        assert b2.shape == (3, 7) and np.array_equal(a2, b2)
        # beyond the maximum: 3 sources x 400k receivers = 14.4 MB of staging
        with pytest.raises(lib.FmmError) as e:
            c.receivers_noalloc(np.tile(gd.rcv[:1], (400000, 1)))
        assert e.value.code == lib.E_ARG and "maximum" in str(e.value)
        # the context is still usable afterwards
        assert np.array_equal(c.receivers_noalloc(gd.rcv[:7]), a2)
