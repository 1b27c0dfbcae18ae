"""The oracle's restatement of the reference's model slicer
(gmtplot/gmtslicet.f90, or_slice) against the reference program's own
output, compiled from its source (tests/golden/make_golden_slice.py), and
against what the B-spline fixes: a constant model slices to that constant
(partition of unity), a model linear in the vertex depth index gives a
depth slice constant over the plane and N-S / E-W slices that change along
depth only."""
import json
import os

import numpy as np
import pytest

import goldens as G
from fmm3d.synth import GridDesc
from oracle import oracle as O

FIX = os.path.join(G.GOLDEN, "slice_31x33x35.npz")


def _fixture():
    z = np.load(FIX, allow_pickle=False)
    return z, GridDesc(**json.loads(str(z["grid"])))


@pytest.mark.skipif(not os.path.exists(FIX), reason="slice fixture not generated")
def test_oracle_slices_equal_the_reference_output():
    z, g = _fixture()
    for (kind, _, d1, d2), key, pos in zip(z["slices"], ("depth", "ns", "ew"), z["pos"]):
        v, (n1, n2) = O.slice_velocity(g, z["vtrue"], z["vref"], int(kind), float(pos), int(d1), int(d2))
        assert v.size == z[key].size == n1 * n2
        assert np.array_equal(v, z[key]), (key, float(np.abs(v - z[key]).max()))


@pytest.mark.skipif(not os.path.exists(FIX), reason="slice fixture not generated")
def test_slices_reproduce_constant_and_linear_models():
    z, g = _fixture()
    const = np.full_like(z["vtrue"], 6.5)
    for kind, pos in zip((0, 1, 2), z["pos"]):
        v, _ = O.slice_velocity(g, const, None, kind, float(pos), 3, 4)
        assert np.abs(v - 6.5).max() < 1e-13
    # linear in the vertex depth index (r fastest in the vertex array)
    kr = np.arange(g.nvr + 2, dtype=np.float64)
    lin = np.broadcast_to(5.0 + 0.25 * kr, z["vtrue"].shape).copy()
    v, (n1, n2) = O.slice_velocity(g, lin, None, 0, float(z["pos"][0]), 3, 4)
    assert np.ptp(v) < 1e-12                      # one depth: one value
    for kind in (1, 2):
        v, (n1, n2) = O.slice_velocity(g, lin, None, kind, float(z["pos"][kind]), 3, 4)
        # rows: the lateral axis; along a row the depth axis from its far
        # (deepest) end, as gmtslicet writes it; the vertex index, and so
        # the model, grows with depth, so each row decreases
        a = v.reshape(n2, n1)
        assert np.ptp(a, axis=0).max() < 1e-12    # the same at every lateral position
        assert np.all(np.diff(a[0]) < 0)
    # a slice beyond the grid is refused
    assert O.slice_velocity(g, const, None, 0, g.rgor + 50.0, 3, 4) is None
