"""ctypes binding of libfmm3d_hip.so (the C ABI declared in include/fmm3d.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

from .synth import GridDesc

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfmm3d_hip.so")

OK = 0
E_ARG, E_DEVICE, E_NOMEM, E_NOCONV, E_STATE, E_CAPACITY, E_TIMEOUT = -1, -2, -3, -4, -5, -6, -7
SRC_TELE, SRC_POINT = 0, 1
ORDER_FIRST, ORDER_MIXED = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
WALK_ALTERNATE = 1
REFERENCE_PARTITION = 2

EXPORTS = ("fmm3d_create", "fmm3d_destroy", "fmm3d_last_error", "fmm3d_set_model_vtx",
           "fmm3d_solve_fixed", "fmm3d_solve_point", "fmm3d_solve_tele_sides", "fmm3d_solve_mixed",
           "fmm3d_stage_tele_sides", "fmm3d_solve_tele_staged",
           "fmm3d_server_create", "fmm3d_server_submit", "fmm3d_server_wait", "fmm3d_server_result",
           "fmm3d_server_stats", "fmm3d_server_destroy", "fmm3d_server_rows",
           "fmm3d_solve_tele", "fmm3d_receivers", "fmm3d_receivers_noalloc", "fmm3d_frechet", "fmm3d_frechet_pinned", "fmm3d_get_field", "fmm3d_get_passes",
           "fmm3d_get_stats", "fmm3d_set_option", "fmm3d_slice_velocity", "fmm3d_slice_last_error",
           "fmm3d_gmat_create", "fmm3d_gmat_gtx", "fmm3d_gmat_gx", "fmm3d_gmat_last_error", "fmm3d_gmat_destroy")


class GridDescC(C.Structure):
    _fields_ = [("nvr", C.c_int), ("nvx", C.c_int), ("nvz", C.c_int),
                ("gor", C.c_double), ("goxd", C.c_double), ("gozd", C.c_double),
                ("dvr", C.c_double), ("dvxd", C.c_double), ("dvzd", C.c_double),
                ("nnr", C.c_int), ("nnx", C.c_int), ("nnz", C.c_int),
                ("rgor", C.c_double), ("rgoxd", C.c_double), ("rgozd", C.c_double),
                ("dnr", C.c_double), ("dnxd", C.c_double), ("dnzd", C.c_double),
                ("earth", C.c_double)]


class StatsC(C.Structure):
    _fields_ = [("node_updates", C.c_int64), ("sweeps", C.c_int32), ("tile_r", C.c_int32),
                ("tile_x", C.c_int32), ("tile_z", C.c_int32), ("solve_ms", C.c_double),
                ("kernel_ms", C.c_double), ("kernel_launches", C.c_int64),
                ("phase_sweeps", C.c_int32 * 6), ("brick_visits", C.c_int64),
                ("node_visits", C.c_int64)]


class RequestC(C.Structure):
    _fields_ = [("grid", GridDescC), ("velv", C.c_void_p), ("model_id", C.c_uint64), ("kind", C.c_int),
                ("nsrc", C.c_int), ("faces", C.c_void_p), ("src", C.c_void_p), ("refine", C.c_int),
                ("radius", C.c_int), ("nrc", C.c_int), ("rcv", C.c_void_p), ("order", C.c_int),
                ("precision", C.c_int), ("rows", C.c_int)]


class ServerInfoC(C.Structure):
    _fields_ = [("requests_done", C.c_int64), ("sources_done", C.c_int64), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("source_solves_per_s", C.c_double), ("hbm_bytes", C.c_double),
                ("models_resident", C.c_int), ("sources_in_flight", C.c_int)]


class FmmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load(path: Optional[str] = None):
    """Load (once) and prototype the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built; run fmm3d.build.build_hip() (or __graft_entry__.build())")
    L = C.CDLL(p)
    vp = C.c_void_p
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    L.fmm3d_create.argtypes = [C.POINTER(GridDescC), C.c_int, C.POINTER(vp)]
    L.fmm3d_destroy.argtypes = [vp]
    L.fmm3d_last_error.argtypes = [vp]
    L.fmm3d_last_error.restype = C.c_char_p
    L.fmm3d_set_model_vtx.argtypes = [vp, dp]
    L.fmm3d_solve_tele.argtypes = [vp, C.c_int, dp, C.c_int, C.c_int, C.c_int]
    L.fmm3d_solve_tele_sides.argtypes = [vp, C.c_int, dp, dp, dp, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                         C.c_int, C.c_int, C.c_int]
    L.fmm3d_stage_tele_sides.argtypes = [vp, C.c_int, dp, C.c_void_p, C.c_void_p, C.c_void_p]
    L.fmm3d_solve_tele_staged.argtypes = [vp, C.c_int, C.c_int]
    L.fmm3d_server_create.argtypes = [C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
    L.fmm3d_server_submit.argtypes = [vp, C.POINTER(RequestC), C.POINTER(C.c_int64)]
    L.fmm3d_server_wait.argtypes = [vp, C.c_int64, C.c_double]
    L.fmm3d_server_result.argtypes = [vp, C.c_int64, fp, C.POINTER(C.c_double)]
    L.fmm3d_server_rows.argtypes = [vp, C.c_int64, lp, ip, fp, C.c_int64]
    L.fmm3d_server_stats.argtypes = [vp, C.POINTER(ServerInfoC)]
    L.fmm3d_server_destroy.argtypes = [vp]
    L.fmm3d_solve_fixed.argtypes = [vp, C.c_int, lp, lp, dp, C.c_int, C.c_int, C.c_int]
    L.fmm3d_solve_point.argtypes = [vp, C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    # nullable arrays go as plain addresses
    L.fmm3d_solve_mixed.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int]
    L.fmm3d_receivers.argtypes = [vp, C.c_int, dp, fp]
    L.fmm3d_receivers_noalloc.argtypes = [vp, C.c_int, dp, fp]
    L.fmm3d_frechet.argtypes = [vp, C.c_int, dp, lp, ip, fp, C.c_int64]
    L.fmm3d_frechet_pinned.argtypes = [vp, C.c_int, dp, C.POINTER(C.POINTER(C.c_int64)),
                                       C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_float)),
                                       C.POINTER(C.c_int64)]
    L.fmm3d_get_field.argtypes = [vp, C.c_int, dp]
    L.fmm3d_slice_velocity.argtypes = [C.POINTER(GridDescC), C.c_int, dp, vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                       dp, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fmm3d_slice_last_error.restype = C.c_char_p
    fp32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.fmm3d_gmat_create.argtypes = [C.c_int, C.c_int64, C.c_int64, lp, ip, fp32, C.POINTER(vp)]
    L.fmm3d_gmat_gtx.argtypes = [vp, fp32, fp32]
    L.fmm3d_gmat_gx.argtypes = [vp, fp32, vp, fp32]
    L.fmm3d_gmat_last_error.argtypes = [vp]
    L.fmm3d_gmat_last_error.restype = C.c_char_p
    L.fmm3d_gmat_destroy.argtypes = [vp]
    L.fmm3d_get_passes.argtypes = [vp, C.c_int, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
    L.fmm3d_get_stats.argtypes = [vp, C.POINTER(StatsC)]
    L.fmm3d_set_option.argtypes = [vp, C.c_char_p, C.c_double]
    for f in EXPORTS:
        if f not in ("fmm3d_last_error", "fmm3d_slice_last_error", "fmm3d_gmat_last_error"):
            getattr(L, f).restype = C.c_int
    _lib = L
    return L


def grid_c(g: GridDesc) -> GridDescC:
    return GridDescC(g.nvr, g.nvx, g.nvz, g.gor, g.goxd, g.gozd, g.dvr, g.dvxd, g.dvzd,
                     g.nnr, g.nnx, g.nnz, g.rgor, g.rgoxd, g.rgozd, g.dnr, g.dnxd, g.dnzd, g.earth)


class Context:
    """One fmm3d_ctx: a grid, a velocity model and the resident fields of the
    last solve on one device.  Mirrors the stages of PROGRAM fmmin3d
    (fm3dt_parallel.f90:6015-6662): gridder, per-source solve, srtimes,
    rays/Frechet rows."""

    def __init__(self, g: GridDesc, device: int = 0):
        self.L = load()
        self.g = g
        self._h = C.c_void_p()
        rc = self.L.fmm3d_create(C.byref(grid_c(g)), device, C.byref(self._h))
        if rc != OK:
            raise FmmError(rc, self.L.fmm3d_last_error(None).decode())
        self.nsrc = 0
        self._staged = 0

    def _check(self, rc: int):
        if rc != OK:
            raise FmmError(rc, self.L.fmm3d_last_error(self._h).decode())

    def close(self):
        if self._h:
            self.L.fmm3d_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, velv: np.ndarray) -> None:
        self._check(self.L.fmm3d_set_model_vtx(self._h, np.ascontiguousarray(velv, dtype=np.float64)))

    def solve_tele(self, faces: np.ndarray, order: int = ORDER_MIXED, precision: int = PREC_FP64,
                   partition: bool = False) -> None:
        """Bottom-face teleseismic solves; partition=True reproduces the
        reference's own 2x2x2 partitioned result (float64 only)."""
        faces = np.ascontiguousarray(faces, dtype=np.float64)
        if faces.ndim == 2:
            faces = faces[None]
        if faces.shape[1:] != (self.g.nnx, self.g.nnz):
            raise ValueError(f"faces must be (nsrc, {self.g.nnx}, {self.g.nnz}), got {faces.shape}")
        flags = REFERENCE_PARTITION if partition else 0
        self._check(self.L.fmm3d_solve_tele(self._h, faces.shape[0], faces, order, precision, flags))
        self.nsrc = faces.shape[0]
        self._staged = 0 if partition else self.nsrc   # a solve stages its own sources

    def passes(self) -> np.ndarray:
        """Outer passes per source of the last partitioned solve."""
        out = np.zeros(self.nsrc, np.int32)
        self._check(self.L.fmm3d_get_passes(self._h, self.nsrc, out))
        return out

    def solve_tele_sides(self, faces: np.ndarray, xface: np.ndarray, zface: np.ndarray, side: np.ndarray,
                         order: int = ORDER_MIXED, precision: int = PREC_FP64) -> None:
        """Bottom faces (nsrc, nnx, nnz) plus one theta face (nsrc, nnr, nnz) and
        one phi face (nsrc, nnr, nnx) per source; side (nsrc, 2): which theta
        and phi face (-1 none, 0 low index, 1 high index)."""
        faces = np.ascontiguousarray(faces, dtype=np.float64)
        xface = np.ascontiguousarray(xface, dtype=np.float64)
        zface = np.ascontiguousarray(zface, dtype=np.float64)
        side = np.ascontiguousarray(side, dtype=np.int32)
        n = faces.shape[0]
        g = self.g
        if (faces.shape != (n, g.nnx, g.nnz) or xface.shape != (n, g.nnr, g.nnz) or zface.shape != (n, g.nnr, g.nnx)
                or side.shape != (n, 2)):
            raise ValueError("solve_tele_sides: face arrays of inconsistent shapes")
        self._check(self.L.fmm3d_solve_tele_sides(self._h, n, faces, xface, zface, side, order, precision, 0))
        self.nsrc = n
        self._staged = n

    def stage_tele(self, faces: np.ndarray, xface: Optional[np.ndarray] = None, zface: Optional[np.ndarray] = None,
                   side: Optional[np.ndarray] = None) -> None:
        """Stage teleseismic sources in device memory (fmm3d_stage_tele_sides):
        bottom faces (nsrc, nnx, nnz), and optionally the side faces and sides
        as solve_tele_sides takes them; solve_tele_staged solves them."""
        faces = np.ascontiguousarray(faces, dtype=np.float64)
        n = faces.shape[0]
        g = self.g
        if faces.shape != (n, g.nnx, g.nnz):
            raise ValueError("stage_tele: faces must be (nsrc, nnx, nnz)")
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        if xface is not None:
            xface = np.ascontiguousarray(xface, dtype=np.float64)
            zface = np.ascontiguousarray(zface, dtype=np.float64)
            side = np.ascontiguousarray(side, dtype=np.int32)
            if xface.shape != (n, g.nnr, g.nnz) or zface.shape != (n, g.nnr, g.nnx) or side.shape != (n, 2):
                raise ValueError("stage_tele: face arrays of inconsistent shapes")
        self._check(self.L.fmm3d_stage_tele_sides(self._h, n, faces, ptr(xface), ptr(zface), ptr(side)))
        self._staged = n

    def solve_tele_staged(self, order: int = ORDER_MIXED, precision: int = PREC_FP64) -> None:
        """Solve the sources staged by stage_tele (fmm3d_solve_tele_staged)."""
        self._check(self.L.fmm3d_solve_tele_staged(self._h, order, precision))
        self.nsrc = self._staged

    def solve_fixed(self, sets, order: int = ORDER_MIXED, precision: int = PREC_FP64, alternate: bool = False) -> None:
        """sets: one (node_index array, time array) pair per source."""
        offs = np.zeros(len(sets) + 1, np.int64)
        for i, (n, _) in enumerate(sets):
            offs[i + 1] = offs[i] + len(n)
        node = np.ascontiguousarray(np.concatenate([np.asarray(n, np.int64) for n, _ in sets]) if sets else np.zeros(0, np.int64))
        tv = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float64) for _, t in sets]) if sets else np.zeros(0))
        if node.size == 0:
            node = np.zeros(1, np.int64); tv = np.zeros(1)
        self._check(self.L.fmm3d_solve_fixed(self._h, len(sets), offs, node, tv, order, precision,
                                             WALK_ALTERNATE if alternate else 0))
        self.nsrc = len(sets)

    def solve_point(self, src: np.ndarray, refine: int = 5, radius: int = 10, order: int = ORDER_MIXED,
                    precision: int = PREC_FP64) -> None:
        """src: (nsrc, 3) height km, colatitude rad, longitude rad."""
        src = np.ascontiguousarray(np.atleast_2d(src), dtype=np.float64)
        self._check(self.L.fmm3d_solve_point(self._h, src.shape[0], src, refine, radius, order, precision, 0))
        self.nsrc = src.shape[0]

    def solve_mixed(self, faces=None, sides=None, src=None, radius=None, refine: int = 5,
                    order: int = ORDER_MIXED, precision: int = PREC_FP64) -> None:
        """One batch of teleseismic sources (bottom faces (nt, nnx, nnz) and,
        optionally, sides = (xface, zface, side) as for solve_tele_sides) and
        point sources (src (np, 3), radius (np,) nodes): fields 0 .. nt-1 are
        the teleseismic ones, nt .. nt+np-1 the point ones."""
        g = self.g
        keep = []

        def addr(a, dtype):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)
        nt = 0 if faces is None else int(np.asarray(faces).shape[0])
        npt = 0 if src is None else int(np.atleast_2d(src).shape[0])
        if faces is not None and np.asarray(faces).shape != (nt, g.nnx, g.nnz):
            raise ValueError("solve_mixed: faces must be (nsrc, nnx, nnz)")
        xf = zf = sd = None
        if sides is not None:
            xface, zface, side = sides
            if (np.shape(xface) != (nt, g.nnr, g.nnz) or np.shape(zface) != (nt, g.nnr, g.nnx)
                    or np.shape(side) != (nt, 2)):
                raise ValueError("solve_mixed: side arrays of inconsistent shapes")
            xf, zf, sd = addr(xface, np.float64), addr(zface, np.float64), addr(side, np.int32)
        if npt and (radius is None or np.shape(radius) != (npt,)):
            raise ValueError("solve_mixed: need one radius per point source")
        self._check(self.L.fmm3d_solve_mixed(
            self._h, nt, addr(faces, np.float64), xf, zf, sd, npt,
            addr(None if src is None else np.atleast_2d(src), np.float64), refine,
            addr(radius, np.int32), order, precision, 0))
        self.nsrc = nt + npt
        if nt > 0:
            self._staged = nt

    def field(self, src: int) -> np.ndarray:
        out = np.zeros((self.g.nnr, self.g.nnx, self.g.nnz))
        self._check(self.L.fmm3d_get_field(self._h, src, out))
        return out

    def receivers(self, rcv: np.ndarray) -> np.ndarray:
        rcv = np.ascontiguousarray(rcv, dtype=np.float64)
        out = np.zeros(self.nsrc * len(rcv), dtype=np.float32)
        self._check(self.L.fmm3d_receivers(self._h, len(rcv), rcv, out))
        return out.reshape(self.nsrc, len(rcv))

    def receivers_noalloc(self, rcv: np.ndarray) -> np.ndarray:
        """fmm3d_receivers_noalloc: the same times, staged in the library's
        fixed device array (at most 8 MiB of coordinates and times)."""
        rcv = np.ascontiguousarray(rcv, dtype=np.float64)
        out = np.zeros(self.nsrc * len(rcv), dtype=np.float32)
        self._check(self.L.fmm3d_receivers_noalloc(self._h, len(rcv), rcv, out))
        return out.reshape(self.nsrc, len(rcv))

    def frechet(self, rcv: np.ndarray, cap: int = 0, pinned: bool = False) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(row_ptr int64 [nsrc*nrc+1], cols int32, vals float32).

        The arrays are sized from the previous call's entry count per row
        (rows of one model and receiver set change little between calls),
        retried once at the exact size when short; cols and vals are views
        of the first row_ptr[-1] entries.  pinned=True returns views of the
        context's pinned host arrays (fmm3d_frechet_pinned), valid until the
        next pinned call on this context."""
        rcv = np.ascontiguousarray(rcv, dtype=np.float64)
        nrows = self.nsrc * len(rcv)
        if pinned:
            prp = C.POINTER(C.c_int64)(); pc = C.POINTER(C.c_int32)(); pv = C.POINTER(C.c_float)()
            n = C.c_int64()
            self._check(self.L.fmm3d_frechet_pinned(self._h, len(rcv), rcv, C.byref(prp), C.byref(pc),
                                                    C.byref(pv), C.byref(n)))
            rp = np.ctypeslib.as_array(prp, shape=(nrows + 1,))
            if n.value == 0:
                return rp, np.zeros(0, np.int32), np.zeros(0, np.float32)
            return rp, np.ctypeslib.as_array(pc, shape=(n.value,)), np.ctypeslib.as_array(pv, shape=(n.value,))
        if cap <= 0:
            per = getattr(self, "_fr_per_row", 512.0)
            cap = max(1024, int(nrows * per * 1.05) + 1024)
        for _ in range(2):
            rp = np.empty(nrows + 1, dtype=np.int64)
            cols = np.empty(cap, dtype=np.int32)
            vals = np.empty(cap, dtype=np.float32)
            rc = self.L.fmm3d_frechet(self._h, len(rcv), rcv, rp, cols, vals, cap)
            if rc == E_CAPACITY:
                cap = int(rp[-1])
                continue
            self._check(rc)
            n = int(rp[-1])
            self._fr_per_row = max(1.0, n / max(nrows, 1))
            return rp, cols[:n], vals[:n]
        raise FmmError(E_CAPACITY, "frechet capacity retry failed")

    def set_option(self, key: str, value: float) -> None:
        self._check(self.L.fmm3d_set_option(self._h, key.encode(), float(value)))

    def stats(self) -> dict:
        s = StatsC()
        self._check(self.L.fmm3d_get_stats(self._h, C.byref(s)))
        d = {f: getattr(s, f) for f, _ in StatsC._fields_}
        d["phase_sweeps"] = list(d["phase_sweeps"])
        return d


class Server:
    """The serving loop (fmm3d_server_*): submit requests, wait, fetch results."""

    def __init__(self, device: int = 0, hbm_budget: float = 64e9, max_slots: int = 64):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.fmm3d_server_create(device, float(hbm_budget), max_slots, C.byref(h))
        if rc != OK:
            raise FmmError(rc, "fmm3d_server_create failed")
        self._h = h
        self._keep = {}

    def submit(self, g, velv, model_id: int, *, faces=None, src=None, refine: int = 5, radius: int = 10,
               rcv=None, order: int = ORDER_MIXED, precision: int = PREC_FP32, rows: bool = False) -> int:
        velv = np.ascontiguousarray(velv, dtype=np.float64)
        rcv = np.ascontiguousarray(np.zeros((0, 3)) if rcv is None else rcv, dtype=np.float64)
        r = RequestC()
        r.grid = grid_c(g)
        r.velv = velv.ctypes.data
        r.model_id = model_id
        if faces is not None:
            faces = np.ascontiguousarray(faces, dtype=np.float64)
            if faces.ndim == 2:
                faces = faces[None]
            r.kind, r.nsrc, r.faces = SRC_TELE, faces.shape[0], faces.ctypes.data
            keep = (velv, rcv, faces)
        else:
            src = np.ascontiguousarray(np.atleast_2d(src), dtype=np.float64)
            r.kind, r.nsrc, r.src = SRC_POINT, src.shape[0], src.ctypes.data
            keep = (velv, rcv, src)
        r.refine, r.radius = refine, radius
        r.nrc, r.rcv = rcv.shape[0], rcv.ctypes.data if rcv.size else None
        r.order, r.precision = order, precision
        r.rows = 1 if rows else 0
        t = C.c_int64()
        rc = self.L.fmm3d_server_submit(self._h, C.byref(r), C.byref(t))
        if rc != OK:
            raise FmmError(rc, "fmm3d_server_submit rejected the request")
        self._keep[t.value] = (r.nsrc, r.nrc)
        return t.value

    def rows(self, ticket: int):
        """Frechet rows of a finished request submitted with rows=True:
        (row_ptr, cols, vals) as Context.frechet; call before result()."""
        nsrc, nrc = self._keep[ticket]
        rp = np.zeros(nsrc * nrc + 1, np.int64)
        cap = nsrc * nrc * 512 + 1024
        for _ in range(2):
            cols = np.zeros(cap, np.int32)
            vals = np.zeros(cap, np.float32)
            rc = self.L.fmm3d_server_rows(self._h, ticket, rp, cols, vals, cap)
            if rc == OK:
                n = int(rp[-1])
                return rp, cols[:n], vals[:n]
            if rc != E_CAPACITY:
                raise FmmError(rc, "fmm3d_server_rows failed")
            cap = int(rp[-1])
        raise FmmError(E_CAPACITY, "fmm3d_server_rows capacity retry failed")

    def wait(self, ticket: int, timeout_s: float = 60.0) -> bool:
        rc = self.L.fmm3d_server_wait(self._h, ticket, timeout_s)
        if rc not in (OK, E_TIMEOUT):
            raise FmmError(rc, "fmm3d_server_wait failed")
        return rc == OK

    def result(self, ticket: int):
        nsrc, nrc = self._keep.pop(ticket)
        out = np.zeros((nsrc, nrc), np.float32)
        lat = C.c_double()
        rc = self.L.fmm3d_server_result(self._h, ticket, out if out.size else np.zeros(1, np.float32), C.byref(lat))
        if rc != OK:
            raise FmmError(rc, f"request {ticket} failed")
        return out, lat.value

    def stats(self) -> dict:
        s = ServerInfoC()
        self.L.fmm3d_server_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerInfoC._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self.L.fmm3d_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def slice_velocity(g: GridDesc, velv: np.ndarray, velv_ref: Optional[np.ndarray], kind: int, where: float,
                   dice1: int, dice2: int, device: int = 0):
    """fmm3d_slice_velocity: the values of one gmtslicet slice in its write
    order, and (n1, n2)."""
    L = load()
    velv = np.ascontiguousarray(velv, dtype=np.float64)
    ref = None if velv_ref is None else np.ascontiguousarray(velv_ref, dtype=np.float64)
    n1, n2 = C.c_int(), C.c_int()
    rc = L.fmm3d_slice_velocity(C.byref(grid_c(g)), device, velv, None if ref is None else ref.ctypes.data, kind,
                                float(where), dice1, dice2, np.zeros(1), 0, C.byref(n1), C.byref(n2))
    if rc != E_CAPACITY:
        raise FmmError(rc, L.fmm3d_slice_last_error().decode())
    out = np.zeros(n1.value * n2.value)
    rc = L.fmm3d_slice_velocity(C.byref(grid_c(g)), device, velv, None if ref is None else ref.ctypes.data, kind,
                                float(where), dice1, dice2, out, out.size, C.byref(n1), C.byref(n2))
    if rc != OK:
        raise FmmError(rc, L.fmm3d_slice_last_error().decode())
    return out, (n1.value, n2.value)


class FrechetMatrix:
    """fmm3d_gmat_*: the Frechet matrix resident on the device, with
    subinv's products G^T x and (G x) / cd."""

    def __init__(self, row_ptr, cols, vals, ncols: int, device: int = 0):
        self.L = load()
        self.rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.nrows, self.ncols = len(self.rp) - 1, int(ncols)
        self._h = C.c_void_p()
        rc = self.L.fmm3d_gmat_create(device, self.nrows, self.ncols, self.rp,
                                      np.ascontiguousarray(cols, dtype=np.int32),
                                      np.ascontiguousarray(vals, dtype=np.float32), C.byref(self._h))
        if rc != OK:
            raise FmmError(rc, self.L.fmm3d_gmat_last_error(None).decode())

    def gtx(self, x) -> np.ndarray:
        y = np.zeros(self.ncols, np.float32)
        rc = self.L.fmm3d_gmat_gtx(self._h, np.ascontiguousarray(x, dtype=np.float32), y)
        if rc != OK:
            raise FmmError(rc, self.L.fmm3d_gmat_last_error(self._h).decode())
        return y

    def gx(self, x, cd=None) -> np.ndarray:
        y = np.zeros(self.nrows, np.float32)
        c = None if cd is None else np.ascontiguousarray(cd, dtype=np.float32)
        rc = self.L.fmm3d_gmat_gx(self._h, np.ascontiguousarray(x, dtype=np.float32),
                                  None if c is None else c.ctypes.data, y)
        if rc != OK:
            raise FmmError(rc, self.L.fmm3d_gmat_last_error(self._h).decode())
        return y

    def close(self):
        if self._h:
            self.L.fmm3d_gmat_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
