"""Hyperplane column walk (strategy 2) against the oracle.

While every wait on an upstream patch succeeds, one sweep of the hyperplane
walk is one Gauss-Seidel sweep in flow order -- walk layers bottom to top,
theta along the slot's lateral flow, then phi along it -- which is the
oracle's gs_flow.  Float64 fields are compared with it after one sweep and
after two (the second reverses a weak axis), and at convergence with the
relaxation tolerance at 0 (the fixed point: the heap march's field), both
with hyperplane sweeps throughout and with the column walk taking over
after two.
Tolerance 1e-9 s: a node that read a neighbour from the wrong sweep is off
by seconds or infinite.  Float32 within 1 ms of float64.
"""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def flow_code(face):
    """The solver's lateral flow code of a bottom face: signs of the
    least-squares (theta, phi) index gradient of its known times; the
    weaker axis gets +-2 when under a quarter of the other."""
    nx, nz = face.shape
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    m = face > 0
    A = np.c_[np.ones(m.sum()), ix[m], iz[m]]
    gx, gz = np.linalg.lstsq(A, face[m], rcond=None)[0][1:]
    code = [1 if gx >= 0 else -1, 1 if gz >= 0 else -1]
    if abs(gx) < 0.25 * abs(gz):
        code[0] *= 2
    elif abs(gz) < 0.25 * abs(gx):
        code[1] *= 2
    return code


def sweep_dir(f, k):
    s = 1 if f > 0 else -1
    return -s if abs(f) == 2 and k % 2 == 0 else s


def oracle_sweeps(gd, veln, face, n):
    fixed = np.zeros((gd.g.nnr, gd.g.nnx, gd.g.nnz), np.uint8)
    tfix = np.zeros(fixed.shape)
    fixed[-1] = face > 0
    tfix[-1] = face
    code = flow_code(face)
    t = None
    for k in range(1, n + 1):
        t, _ = O.gs_flow(gd.g, veln, fixed, tfix, gd.fom, sweep_dir(code[0], k), sweep_dir(code[1], k),
                         sweeps=1, start=t)
    return t


def plane_ctx(gd, chunk=4):
    c = lib.Context(gd.g)
    c.set_model(gd.velv)
    c.set_option("strategy", 2)
    c.set_option("plane_chunk", chunk)
    c.set_option("tol_fp64", 0.0)
    return c


@pytest.mark.parametrize("name", G.names())
@pytest.mark.parametrize("chunk", [2, 4, 8])
def test_sweeps_are_flow_order_gauss_seidel(name, chunk):
    gd = G.load(name)
    veln = O.gridder(gd.g, gd.velv)
    print(f"plane sweeps: {name} chunk {chunk}", flush=True)
    with plane_ctx(gd, chunk) as c:
        for n in (1, 2):
            c.set_option("max_sweeps", n)
            c.solve_tele(gd.faces, gd.fom, lib.PREC_FP64)
            for s in range(gd.faces.shape[0]):
                ref = oracle_sweeps(gd, veln, gd.faces[s], n)
                F = c.field(s)
                assert np.array_equal(np.isfinite(F), np.isfinite(ref)), (n, s)
                fin = np.isfinite(ref)
                d = float(np.abs(F[fin] - ref[fin]).max())
                assert d <= 1e-9, (n, s, d)


@pytest.mark.parametrize("name", G.names())
@pytest.mark.parametrize("plane_sweeps", [0, 2])
def test_fixed_point_is_the_heap_march(name, plane_sweeps):
    """All sweeps hyperplane (0), or two and then the column walk (2)."""
    gd = G.load(name)
    veln = O.gridder(gd.g, gd.velv)
    print(f"plane fixed point: {name} plane_sweeps {plane_sweeps}", flush=True)
    with plane_ctx(gd) as c:
        c.set_option("plane_sweeps", plane_sweeps)
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP64)
        st = c.stats()
        for s in range(gd.faces.shape[0]):
            H = O.fmm_tele(gd.g, veln, gd.faces[s], gd.fom)
            d = float(np.abs(c.field(s) - H).max())
            assert d <= 1e-9, (s, d)
        f64 = [c.field(s) for s in range(gd.faces.shape[0])]
        t64 = c.receivers(gd.rcv)
        c.solve_tele(gd.faces, gd.fom, lib.PREC_FP32)
        t32 = c.receivers(gd.rcv)
        assert np.abs(t32.astype(np.float64) - t64).max() <= 1e-3
        for s in range(gd.faces.shape[0]):
            assert np.abs(c.field(s) - f64[s]).max() <= 1e-3
    print("  sweeps", st["sweeps"], "updates", st["node_updates"], flush=True)
