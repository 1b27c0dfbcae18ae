"""Closed-form checks of the oracle's numerics (the pieces no golden pins
on its own): the octant quadratic, the B-spline gridder, the heap march on
fields the discrete scheme represents exactly, and the equivalence of the
heap march with the relaxation fixed point the HIP solver computes."""
import ctypes as C
import math

import numpy as np
import pytest

import goldens as G
from fmm3d import synth
from oracle import oracle as O


def test_octant_known_answers():
    h = [3.0, 8.0, 7.0]
    s = 0.125
    # one first-order neighbour: T = t1 + h*s
    assert O.octant_solve([1, 0, 0], [10.0, 0, 0], [0, 0, 0], h, s) == pytest.approx(10.0 + 3.0 * s, abs=1e-15)
    assert O.octant_solve([0, 0, 1], [0, 0, 10.0], [0, 0, 0], h, s) == pytest.approx(10.0 + 7.0 * s, abs=1e-15)
    # one second-order neighbour: (3T - 4t1 + t2)/(2h) = s
    t1, t2 = 10.0, 9.7
    assert O.octant_solve([2, 0, 0], [t1, 0, 0], [t2, 0, 0], h, s) == pytest.approx((4 * t1 - t2 + 2 * 3.0 * s) / 3, abs=1e-13)
    # two equal first-order neighbours at equal spacing: T = t + h*s/sqrt(2)
    hh = [5.0, 5.0, 5.0]
    assert O.octant_solve([1, 1, 0], [4.0, 4.0, 0], [0, 0, 0], hh, s) == pytest.approx(4.0 + 5.0 * s / math.sqrt(2), abs=1e-14)
    # three: T = t + h*s/sqrt(3)
    assert O.octant_solve([1, 1, 1], [4.0, 4.0, 4.0], [0, 0, 0], hh, s) == pytest.approx(4.0 + 5.0 * s / math.sqrt(3), abs=1e-14)
    # the solution satisfies the discrete equation for arbitrary causal data
    t = [10.0, 10.3, 10.2]
    T = O.octant_solve([1, 1, 1], t, [0, 0, 0], h, s)
    res = sum(((T - ti) / hi) ** 2 for ti, hi in zip(t, h))
    assert res == pytest.approx(s * s, rel=1e-12)
    # mixed orders, plugged back into the stencil
    t1 = [10.0, 10.3, 10.2]
    t2 = [9.6, 0.0, 9.9]
    T = O.octant_solve([2, 1, 2], t1, t2, h, s)
    d = [(3 * T - 4 * t1[0] + t2[0]) / (2 * h[0]), (T - t1[1]) / h[1], (3 * T - 4 * t1[2] + t2[2]) / (2 * h[2])]
    assert sum(x * x for x in d) == pytest.approx(s * s, rel=1e-10)
    # negative discriminant is clamped: the minimiser of the quadratic form
    T = O.octant_solve([1, 1, 0], [0.0, 100.0, 0], [0, 0, 0], [1.0, 1.0, 1.0], 0.1)
    assert T == pytest.approx(50.0, abs=1e-12)


def _grid(nn=(13, 14, 15)):
    return synth.make_grid(nn, 120.0, (3.0, 3.5), nv=(7, 8, 9))


def test_gridder_partition_of_unity_and_linear_reproduction():
    g = _grid()
    v = np.full((g.nvz + 2, g.nvx + 2, g.nvr + 2), 6.5)
    veln = O.gridder(g, v)
    assert np.abs(veln - 6.5).max() < 1e-13
    # a field linear in depth is reproduced exactly by the cubic B-spline
    k = np.arange(g.nvr + 2)
    hv = g.gor - (k - 1) * g.dvr
    v = np.ascontiguousarray(np.broadcast_to(5.0 - 0.01 * hv, (g.nvz + 2, g.nvx + 2, g.nvr + 2)))
    veln = O.gridder(g, v)
    h = np.array([g.rgor - (i - 1) * g.dnr for i in range(1, g.nnr + 1)])
    assert np.abs(veln[2:-1, 2:-1, 2:-1] - (5.0 - 0.01 * h)[:, None, None]).max() < 1e-12


def test_vertical_plane_wave_is_exact():
    """Homogeneous medium, uniform bottom face: t = t0 + (h - h_bot)*s."""
    g = _grid()
    v = np.full((g.nvz + 2, g.nvx + 2, g.nvr + 2), 8.0)
    veln = O.gridder(g, v)
    face = np.full((g.nnx, g.nnz), 500.0)
    h = np.array([g.rgor - (i - 1) * g.dnr for i in range(1, g.nnr + 1)])
    for fom in (0, 1):
        T = O.fmm_tele(g, veln, face, fom)
        ex = 500.0 + (h - h[-1])[:, None, None] / 8.0
        assert np.abs(T - ex).max() < 1e-10


def test_heap_march_equals_relaxation_fixed_point():
    """The three relaxation phases the HIP solver runs (first order with the
    value-based predicate; second order with the stencil choice frozen from
    the previous field, twice) reach the heap march's field bit for bit."""
    L = O.lib()
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    up = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    L.or_fim_sweep_pred.argtypes = [C.POINTER(O.OrGrid), dp, up, C.c_int, dp, dp, dp, C.c_int]
    L.or_fim_sweep_pred.restype = C.c_double
    g, velv, srcs, rcv = synth.tele_case((21, 23, 25), 300.0, (8.0, 9.0), (8, 8, 9), 0.06, 3, 8, 9)
    veln = O.gridder(g, velv)
    og = O.or_grid(g)
    fixed = np.zeros((g.nnr, g.nnx, g.nnz), np.uint8)
    fixed[-1] = 1

    def run(face, fom, init, pred):
        t = init.copy()
        t[-1] = face
        for _ in range(2000):
            prev = t.copy()
            L.or_fim_sweep_pred(C.byref(og), veln, fixed, fom, t, t if pred is None else pred, t, 1)
            if np.array_equal(prev, t):
                return t
        raise AssertionError("no convergence")

    for s in srcs:
        face = synth.tele_bottom_face(g, s)
        A = run(face, 0, np.full(fixed.shape, np.inf), None)
        assert np.array_equal(A, O.fmm_tele(g, veln, face, 0))
        B = run(face, 1, A, A)
        Cc = run(face, 1, B, B)
        assert np.array_equal(Cc, O.fmm_tele(g, veln, face, 1))
        assert np.array_equal(O.relax_tele(g, veln, face, 1), Cc)


def test_single_phase_column_relaxation_equals_heap_march():
    """The column strategy's discrete system -- the value-based mixed-order
    update ordered by the node's own value, unsupported values dropped,
    relaxed layer by layer from the bottom with four line passes per layer
    -- reaches the heap march's field bit for bit on every golden source,
    first and mixed order."""
    for name in G.names():
        gd = G.load(name)
        veln = O.gridder(gd.g, gd.velv)
        for s in range(gd.faces.shape[0]):
            R, n = O.relax1_tele(gd.g, veln, gd.faces[s], gd.fom)
            assert n >= 1
            assert np.array_equal(R, O.fmm_tele(gd.g, veln, gd.faces[s], gd.fom)), (name, s)


def test_single_phase_relaxation_matches_three_phase_with_face_holes():
    """With unknown nodes in the bottom face the heap march's order shapes
    the result; both relaxations reach the same fixed point there."""
    gd = G.load(G.names()[0])
    veln = O.gridder(gd.g, gd.velv)
    face = gd.faces[0].copy()
    face[:5, :7] = 0.0
    face[-3:, -2:] = -1.0
    R1, _ = O.relax1_tele(gd.g, veln, face, 1)
    assert np.array_equal(R1, O.relax_tele(gd.g, veln, face, 1))
    assert np.isfinite(R1).all()


def test_flow_order_gauss_seidel_equals_heap_march_in_any_order():
    """The hyperplane strategy's sweep -- Gauss-Seidel in (walk layer, theta
    along fx, phi along fz) order, one update per node -- reaches the heap
    march's field bit for bit on the golden sources, whichever lateral order
    it walks in (the fixed point does not depend on the order; the number of
    sweeps does: two for most sources when the order follows the wave, more
    where the wave's lateral direction changes sign inside the model)."""
    best = []
    for name in G.names():
        gd = G.load(name)
        veln = O.gridder(gd.g, gd.velv)
        for s in range(0, gd.faces.shape[0], 2):
            face = gd.faces[s]
            H = O.fmm_tele(gd.g, veln, face, gd.fom)
            fixed = np.zeros((gd.g.nnr, gd.g.nnx, gd.g.nnz), np.uint8)
            tfix = np.zeros(fixed.shape)
            fixed[-1] = face > 0
            tfix[-1] = face
            counts = []
            for fx, fz in ((1, 1), (-1, 1), (1, -1), (-1, -1)):
                R, n = O.gs_flow(gd.g, veln, fixed, tfix, gd.fom, fx, fz)
                assert n >= 2, (name, s, fx, fz, n)
                assert np.array_equal(R, H), (name, s, fx, fz)
                counts.append(n)
            best.append(min(counts))
    assert sorted(best)[len(best) // 2] == 2 and max(best) <= 20, best
    # a single sweep from a partial field stops early and says so
    gd = G.load(G.names()[0])
    veln = O.gridder(gd.g, gd.velv)
    fixed = np.zeros((gd.g.nnr, gd.g.nnx, gd.g.nnz), np.uint8)
    fixed[-1] = 1
    tfix = np.zeros(fixed.shape)
    tfix[-1] = gd.faces[0]
    R1, n1 = O.gs_flow(gd.g, veln, fixed, tfix, gd.fom, 1, -1, sweeps=1)
    assert n1 == -1 and np.isfinite(R1).all()
    R2, n2 = O.gs_flow(gd.g, veln, fixed, tfix, gd.fom, 1, -1, sweeps=1, start=R1)
    R2b, _ = O.gs_flow(gd.g, veln, fixed, tfix, gd.fom, 1, -1, sweeps=2)
    assert np.array_equal(R2, R2b)
