"""ctypes binding of the CPU oracle (fmm3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never loads
it.

Array conventions (see the C file header): fields (nnr, nnx, nnz) phi
fastest; veln (nnr+3, nnx+3, nnz+3) on nodes -1..n+1; vertex arrays
(nvz+2, nvx+2, nvr+2) r fastest; receivers (nrc, 3) height km, colat rad,
lon rad.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OrGrid(C.Structure):
    _fields_ = [("nvr", C.c_int), ("nvx", C.c_int), ("nvz", C.c_int),
                ("gor", C.c_double), ("gox", C.c_double), ("goz", C.c_double),
                ("dvr", C.c_double), ("dvx", C.c_double), ("dvz", C.c_double),
                ("nnr", C.c_int), ("nnx", C.c_int), ("nnz", C.c_int),
                ("rgor", C.c_double), ("rgox", C.c_double), ("rgoz", C.c_double),
                ("dnr", C.c_double), ("dnx", C.c_double), ("dnz", C.c_double),
                ("earth", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        L.or_gridder.argtypes = [C.POINTER(OrGrid), dp, dp]
        L.or_gridder.restype = None
        L.or_fmm_tele.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, dp]
        L.or_fmm_tele.restype = C.c_long
        L.or_fmm_tele_part.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, dp, C.c_int]
        L.or_fmm_tele_part.restype = C.c_int
        L.or_relax_tele_part.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, dp, C.c_int]
        L.or_relax_tele_part.restype = C.c_int
        L.or_relax_tele.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, dp]
        L.or_relax_tele.restype = C.c_long
        L.or_relax1_tele.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, dp]
        L.or_relax1_tele.restype = C.c_long
        L.or_fmm_fixed.argtypes = [C.POINTER(OrGrid), dp, up, dp, C.c_int, dp]
        L.or_fmm_fixed.restype = C.c_long
        L.or_relax1_fixed.argtypes = [C.POINTER(OrGrid), dp, up, dp, C.c_int, C.c_int, dp]
        L.or_relax1_fixed.restype = C.c_long
        L.or_gs_flow.argtypes = [C.POINTER(OrGrid), dp, up, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.or_gs_flow.restype = C.c_long
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.or_point_region.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, up, dp]
        L.or_point_region.restype = C.c_long
        L.or_fim_sweep.argtypes = [C.POINTER(OrGrid), dp, up, C.c_int, dp, dp]
        L.or_fim_sweep.restype = C.c_double
        L.or_srtimes.argtypes = [C.POINTER(OrGrid), dp, C.c_int, dp, fp]
        L.or_srtimes.restype = C.c_int
        L.or_rpaths.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_int, dp, lp, ip, fp, C.c_long, ip, ip]
        L.or_rpaths.restype = C.c_long
        L.or_slice.argtypes = [C.POINTER(OrGrid), dp, dp, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, dp,
                               C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_slice.restype = C.c_int
        f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.or_gmat_gtx.argtypes = [C.c_long, C.c_long, lp, ip, f32, f32, f32]
        L.or_gmat_gtx.restype = None
        L.or_gmat_gx.argtypes = [C.c_long, lp, ip, f32, f32, C.c_void_p, f32]
        L.or_gmat_gx.restype = None
        L.or_octant_solve.argtypes = [ip, dp, dp, dp, C.c_double]
        L.or_octant_solve.restype = C.c_double
        _lib = L
    return _lib


def or_grid(g) -> OrGrid:
    """GridDesc (degrees) -> OrGrid (radians, reference conversion)."""
    return OrGrid(g.nvr, g.nvx, g.nvz, g.gor, g.gox, g.goz, g.dvr, g.dvx, g.dvz,
                  g.nnr, g.nnx, g.nnz, g.rgor, g.rgox, g.rgoz, g.dnr, g.dnx, g.dnz, g.earth)


def gridder(g, velv: np.ndarray) -> np.ndarray:
    out = np.zeros((g.nnr + 3, g.nnx + 3, g.nnz + 3))
    lib().or_gridder(C.byref(or_grid(g)), np.ascontiguousarray(velv, dtype=np.float64), out)
    return out


def fmm_tele(g, veln: np.ndarray, face: np.ndarray, fom: int = 1) -> np.ndarray:
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    lib().or_fmm_tele(C.byref(or_grid(g)), veln, np.ascontiguousarray(face, dtype=np.float64), fom, ttn)
    return ttn


def fmm_tele_part(g, velv: np.ndarray, face: np.ndarray, fom: int = 1, max_pass: int = 1000):
    """The reference's 2x2x2-partitioned solve; returns (ttn, passes)."""
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    n = lib().or_fmm_tele_part(C.byref(or_grid(g)), np.ascontiguousarray(velv, dtype=np.float64),
                               np.ascontiguousarray(face, dtype=np.float64), fom, ttn, max_pass)
    return ttn, n


def relax_tele_part(g, velv: np.ndarray, face: np.ndarray, fom: int = 1, max_pass: int = 1000):
    """The partitioned solve in the GPU's parallel form (pop-order block
    relaxation, face updates in dataflow rounds); returns (ttn, passes)."""
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    n = lib().or_relax_tele_part(C.byref(or_grid(g)), np.ascontiguousarray(velv, dtype=np.float64),
                                 np.ascontiguousarray(face, dtype=np.float64), fom, ttn, max_pass)
    return ttn, n


def relax_tele(g, veln: np.ndarray, face: np.ndarray, fom: int = 1) -> np.ndarray:
    """Three-phase relaxation fixed point (the brick-sweep strategy's system); unknown = inf."""
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    n = lib().or_relax_tele(C.byref(or_grid(g)), veln, np.ascontiguousarray(face, dtype=np.float64), fom, ttn)
    if n < 0:
        raise RuntimeError("or_relax_tele: no convergence")
    return ttn


def fmm_fixed(g, veln, fixed, tfix, fom=1):
    """Heap march from a general fixed node set (every fixed node alive)."""
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    lib().or_fmm_fixed(C.byref(or_grid(g)), veln, np.ascontiguousarray(fixed, dtype=np.uint8),
                       np.ascontiguousarray(tfix, dtype=np.float64), fom, ttn)
    return ttn


def relax1_fixed(g, veln, fixed, tfix, fom=1, alternate=True):
    """Single-phase column relaxation from a general fixed set; (field, sweeps)."""
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    n = lib().or_relax1_fixed(C.byref(or_grid(g)), veln, np.ascontiguousarray(fixed, dtype=np.uint8),
                              np.ascontiguousarray(tfix, dtype=np.float64), fom, int(alternate), ttn)
    if n < 0:
        raise RuntimeError("or_relax1_fixed: no convergence")
    return ttn, n


def gs_flow(g, veln, fixed, tfix, fom=1, fx=1, fz=1, down=False, sweeps=0, start=None):
    """Gauss-Seidel sweeps in flow order (walk layer, theta along fx, phi
    along fz) from the fixed set (or from the field `start`); `sweeps` > 0
    stops after that many.  Returns (field, n): n = the sweep that changed
    nothing, or -(sweeps run)."""
    if start is None:
        ttn = np.where(np.asarray(fixed, bool), tfix, np.inf).astype(np.float64)
    else:
        ttn = np.array(start, dtype=np.float64, order="C")
    n = lib().or_gs_flow(C.byref(or_grid(g)), veln, np.ascontiguousarray(fixed, dtype=np.uint8),
                         fom, int(fx), int(fz), int(down), int(sweeps), ttn)
    return ttn, n


def point_region(g, velv, src, refine=5, radius=10, fom=1, solve=0):
    """Refined neighbourhood of a point source (height km, colat rad, lon rad):
    returns (lo, hi, fixed, tfix) -- the coarse fixed set the refined solve
    hands to the coarse grid.  solve 0: heap march, 1: column relaxation."""
    lo = np.zeros(3, np.int32); hi = np.zeros(3, np.int32)
    fixed = np.zeros((g.nnr, g.nnx, g.nnz), np.uint8)
    tfix = np.zeros((g.nnr, g.nnx, g.nnz))
    n = lib().or_point_region(C.byref(or_grid(g)), np.ascontiguousarray(velv, dtype=np.float64),
                              np.ascontiguousarray(src, dtype=np.float64), refine, radius, fom, solve,
                              lo, hi, fixed, tfix)
    if n < 0:
        raise ValueError("point source outside the grid")
    return lo, hi, fixed, tfix


def point_source(g, velv, veln, src, refine=5, radius=10, fom=1):
    """Point-source field by the heap march: refined neighbourhood, then the
    coarse grid from the transferred nodes."""
    lo, hi, fixed, tfix = point_region(g, velv, src, refine, radius, fom, 0)
    return fmm_fixed(g, veln, fixed, tfix, fom)


def relax1_tele(g, veln: np.ndarray, face: np.ndarray, fom: int = 1):
    """Single-phase value-based relaxation in column-walk order (the column
    strategy's system) to its fixed point; returns (field, sweeps)."""
    ttn = np.zeros((g.nnr, g.nnx, g.nnz))
    n = lib().or_relax1_tele(C.byref(or_grid(g)), veln, np.ascontiguousarray(face, dtype=np.float64), fom, ttn)
    if n < 0:
        raise RuntimeError("or_relax1_tele: no convergence")
    return ttn, n


def fim_sweep(g, veln, fixed, t_in, fom=1):
    out = np.empty_like(t_in)
    d = lib().or_fim_sweep(C.byref(or_grid(g)), veln, np.ascontiguousarray(fixed, dtype=np.uint8),
                           fom, np.ascontiguousarray(t_in), out)
    return out, d


def srtimes(g, ttn: np.ndarray, rcv: np.ndarray) -> np.ndarray:
    out = np.zeros(len(rcv), dtype=np.float32)
    lib().or_srtimes(C.byref(or_grid(g)), np.ascontiguousarray(ttn), len(rcv),
                     np.ascontiguousarray(rcv, dtype=np.float64), out)
    return out


def rpaths(g, veln: np.ndarray, ttn: np.ndarray, rcv: np.ndarray, cap: int = 0, src=None):
    """Returns (row_ptr int64 [nrc+1], cols int32, vals float32, ok int32, steps int32).
    src: point-source position (height km, colat rad, lon rad) or None."""
    nrc = len(rcv)
    if cap <= 0:
        cap = nrc * g.nvert if hasattr(g, "nvert") else nrc * 100000
    rp = np.zeros(nrc + 1, dtype=np.int64)
    cols = np.zeros(cap, dtype=np.int32)
    vals = np.zeros(cap, dtype=np.float32)
    ok = np.zeros(nrc, dtype=np.int32)
    steps = np.zeros(nrc, dtype=np.int32)
    if src is None:
        n = lib().or_rpaths(C.byref(or_grid(g)), veln, np.ascontiguousarray(ttn), nrc,
                            np.ascontiguousarray(rcv, dtype=np.float64), rp, cols, vals, cap, ok, steps)
    else:
        L = lib()
        L.or_rpaths_point.argtypes = list(L.or_rpaths.argtypes) + [np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
        L.or_rpaths_point.restype = C.c_long
        n = L.or_rpaths_point(C.byref(or_grid(g)), veln, np.ascontiguousarray(ttn), nrc,
                              np.ascontiguousarray(rcv, dtype=np.float64), rp, cols, vals, cap, ok, steps,
                              np.ascontiguousarray(src, dtype=np.float64))
    if n < 0:
        raise RuntimeError("or_rpaths: capacity too small")
    return rp, cols[:n].copy(), vals[:n].copy(), ok, steps


def octant_solve(ord3, t1, t2, h, slown) -> float:
    return lib().or_octant_solve(np.asarray(ord3, dtype=np.int32), np.asarray(t1, dtype=np.float64),
                                 np.asarray(t2, dtype=np.float64), np.asarray(h, dtype=np.float64), slown)


def slice_velocity(g, vtrue, vref, kind, where, d1, d2):
    """gmtslicet's velocity-perturbation slice (kind 0 depth at height
    `where` km, 1 N-S at longitude `where` deg, 2 E-W at latitude `where`
    deg; dicing d1, d2): the values in gmtslicet's output order, and
    (n1, n2).  None when a tap falls outside the arrays."""
    deg = np.array([g.rgoxd, g.dnxd, g.rgozd, g.dnzd], np.float64)
    if kind == 0:
        n = ((g.nnx - 1) * d1 + 1) * ((g.nnz - 1) * d2 + 1)
    elif kind == 1:
        n = ((g.nnr - 3) * d1 + 1) * ((g.nnx - 1) * d2 + 1)
    else:
        n = ((g.nnr - 3) * d1 + 1) * ((g.nnz - 1) * d2 + 1)
    out = np.zeros(n)
    n1, n2 = C.c_int(), C.c_int()
    vr = None if vref is None else np.ascontiguousarray(vref, dtype=np.float64)
    rc = lib().or_slice(C.byref(or_grid(g)), deg, np.ascontiguousarray(vtrue, dtype=np.float64),
                        None if vr is None else vr.ctypes.data, kind, float(where), d1, d2, out,
                        C.byref(n1), C.byref(n2))
    if rc != 0:
        return None
    return out, (n1.value, n2.value)


def gmat_gtx(row_ptr, cols, vals, ncols, x):
    """subinv's G^T x (float32, its transpose order)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    y = np.zeros(ncols, np.float32)
    lib().or_gmat_gtx(len(rp) - 1, ncols, rp, np.ascontiguousarray(cols, dtype=np.int32),
                      np.ascontiguousarray(vals, dtype=np.float32), np.ascontiguousarray(x, dtype=np.float32), y)
    return y


def gmat_gx(row_ptr, cols, vals, x, cd=None):
    """subinv's (G x) / cd (float32, row order)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
    y = np.zeros(len(rp) - 1, np.float32)
    c = None if cd is None else np.ascontiguousarray(cd, dtype=np.float32)
    lib().or_gmat_gx(len(rp) - 1, rp, np.ascontiguousarray(cols, dtype=np.int32),
                     np.ascontiguousarray(vals, dtype=np.float32), np.ascontiguousarray(x, dtype=np.float32),
                     None if c is None else c.ctypes.data, y)
    return y
