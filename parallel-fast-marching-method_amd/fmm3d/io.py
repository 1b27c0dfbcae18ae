"""Readers and writers for the fm3dt_parallel file contract.

Inputs (fm3dt_parallel.f90:6082-6126, 6176-6190, 6208-6219, 6275-6332):
``fm3dt_parallel.in`` (18 fixed-position lines), the ``.vtx`` vertex model,
``sources.dat``, ``receivers.dat`` and the binary ``itimes.out`` (nsrc
consecutive float64 bottom faces, phi fastest, no record markers).

Outputs (fm3dt_parallel.f90:6503-6599, 7133-7176): ``rtravel.out`` (one
list-directed REAL(4) per line, ordered by array, source, receiver) and
``frechet.out`` (Fortran unformatted sequential: per row a header record
``(int32 nnz, float32 receiver_id)`` then ``nnz`` records ``(int32 column,
float32 value)``; every record is framed by 4-byte length markers).

Degrees are converted with the reference's float32-rounded pi
(fm3dt_parallel.f90:61), so cell indices computed from them agree with the
reference's FLOOR calls.
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import List, Tuple

import numpy as np

from .synth import GridDesc, PI_REF


@dataclasses.dataclass
class RunConfig:
    """The 18-line ``fm3dt_parallel.in`` (fm3dt_parallel.f90:6082-6102)."""
    sources: str = "sources.dat"
    itimes: str = "itimes.out"
    receivers: str = "receivers.dat"
    gridv: str = "gridc.vtx"
    earth: float = 6371.0
    fom: int = 1
    ltfr: int = 1
    cslat: float = 0.4
    cslong: float = 0.4
    fsrt: int = 1
    rtravel: str = "rtravel.out"
    cfd: int = 1
    frechet: str = "frechet.out"


def _word(line: str) -> str:
    # FORMAT(a20): the first 20 characters, then the leading token
    return line[:20].split()[0] if line[:20].split() else ""


def read_in(path: str) -> RunConfig:
    with open(path) as f:
        L = f.read().split("\n")
    c = RunConfig()
    c.sources = _word(L[3])
    c.itimes = _word(L[4])
    c.receivers = _word(L[5])
    c.gridv = _word(L[6])
    c.earth = float(L[7].split()[0])
    c.fom = int(L[8].split()[0])
    c.ltfr = int(L[9].split()[0])
    c.cslat, c.cslong = (float(x) for x in L[10].split()[:2])
    c.fsrt = int(L[14].split()[0])
    c.rtravel = _word(L[15])
    c.cfd = int(L[16].split()[0])
    c.frechet = _word(L[17])
    return c


def read_vtx(path: str, earth: float = 6371.0) -> Tuple[GridDesc, np.ndarray]:
    """Returns the grid descriptor and the vertex array (nvz+2, nvx+2, nvr+2),
    r fastest (fm3dt_parallel.f90:6103-6126).  Each record's first value is
    the velocity; anything after it on the line is ignored, as list-directed
    input does."""
    with open(path) as f:
        lines = f.read().split("\n")
    h = [ln.split() for ln in lines[:6]]
    nvr, nvx, nvz = (int(x) for x in h[0][:3])
    gor, goxd, gozd = (float(x) for x in h[1][:3])
    dvr, dvxd, dvzd = (float(x) for x in h[2][:3])
    nnr, nnx, nnz = (int(x) for x in h[3][:3])
    rgor, rgoxd, rgozd = (float(x) for x in h[4][:3])
    dnr, dnxd, dnzd = (float(x) for x in h[5][:3])
    g = GridDesc(nvr, nvx, nvz, gor, goxd, gozd, dvr, dvxd, dvzd,
                 nnr, nnx, nnz, rgor, rgoxd, rgozd, dnr, dnxd, dnzd, earth)
    n = (nvr + 2) * (nvx + 2) * (nvz + 2)
    vals = []
    for ln in lines[6:]:
        t = ln.split()
        if t:
            vals.append(float(t[0].replace("D", "E").replace("d", "e")))
            if len(vals) == n:
                break
    if len(vals) != n:
        raise ValueError(f"{path}: expected {n} vertex values, found {len(vals)}")
    return g, np.array(vals, dtype=np.float64).reshape(nvz + 2, nvx + 2, nvr + 2)


def read_sources(path: str) -> List[List[Tuple[float, float, float, str]]]:
    """Per receiver array, the list of (lat, lon, depth, phase).  The solver
    reads but never uses these (fm3dt_parallel.f90:6275-6285)."""
    with open(path) as f:
        tok = [ln for ln in f.read().split("\n")]
    pos = 0
    nra = int(tok[pos].split()[0]); pos += 1
    out = []
    for _ in range(nra):
        nsrc = int(tok[pos].split()[0]); pos += 1
        arr = []
        for _ in range(nsrc):
            la, lo, dp = (float(x) for x in tok[pos].split()[:3]); pos += 1
            ph = tok[pos].strip(); pos += 1
            arr.append((la, lo, dp, ph))
        out.append(arr)
    return out


def read_receivers(path: str) -> List[np.ndarray]:
    """Per receiver array an (nrc, 3) array of (height km, colat rad, lon rad),
    converted as fm3dt_parallel.f90:6287-6315."""
    with open(path) as f:
        tok = f.read().split("\n")
    pos = 0
    nra = int(tok[pos].split()[0]); pos += 1
    out = []
    for _ in range(nra):
        nrc = int(tok[pos].split()[0]); pos += 1
        a = np.zeros((nrc, 3))
        for i in range(nrc):
            h, la, lo = (float(x) for x in tok[pos].split()[:3]); pos += 1
            a[i] = (h, (90.0 - la) * PI_REF / 180.0, lo * PI_REF / 180.0)
        out.append(a)
    return out


def read_itimes(path: str, nsrc: int, nnx: int, nnz: int, first: int = 0) -> np.ndarray:
    """``nsrc`` bottom faces starting at face ``first``: (nsrc, nnx, nnz) float64."""
    n = nnx * nnz
    a = np.fromfile(path, dtype="<f8", count=nsrc * n, offset=first * n * 8)
    if a.size != nsrc * n:
        raise ValueError(f"{path}: short read ({a.size} of {nsrc * n} values)")
    return a.reshape(nsrc, nnx, nnz)


def write_rtravel(path: str, times: np.ndarray, append: bool = False) -> None:
    """One REAL(4) per line.  The text is the shortest repr with a leading
    blank; flang's own list-directed writer may pick a different last digit
    of equal length, so consumers compare the values, not the bytes."""
    with open(path, "a" if append else "w") as f:
        for v in np.asarray(times, dtype=np.float32):
            f.write(" " + np.format_float_positional(v, unique=True, trim="-") + "\n")


def read_rtravel(path: str) -> np.ndarray:
    with open(path) as f:
        return np.array([float(x) for x in f.read().split()], dtype=np.float32)


_REC = struct.Struct("<iifi")  # marker, int32, float32, marker


def write_frechet(path: str, rows: List[Tuple[int, np.ndarray, np.ndarray]], append: bool = False) -> None:
    """rows: (receiver_id, cols int32, vals float32) in output order."""
    with open(path, "ab" if append else "wb") as f:
        for rid, cols, vals in rows:
            f.write(_REC.pack(8, len(cols), float(rid), 8))
            buf = np.empty(len(cols), dtype=[("m0", "<i4"), ("c", "<i4"), ("v", "<f4"), ("m1", "<i4")])
            buf["m0"] = 8
            buf["c"] = cols
            buf["v"] = vals
            buf["m1"] = 8
            f.write(buf.tobytes())


def read_frechet(path: str) -> List[Tuple[int, np.ndarray, np.ndarray]]:
    """Inverse of write_frechet; also reads the reference's own output."""
    raw = np.fromfile(path, dtype=[("m0", "<i4"), ("c", "<i4"), ("v", "<f4"), ("m1", "<i4")])
    if raw.size and not (np.all(raw["m0"] == 8) and np.all(raw["m1"] == 8)):
        raise ValueError(f"{path}: unexpected record markers")
    rows = []
    i = 0
    while i < raw.size:
        n = int(raw["c"][i])
        rid = int(round(float(raw["v"][i])))
        rows.append((rid, raw["c"][i + 1:i + 1 + n].copy(), raw["v"][i + 1:i + 1 + n].copy()))
        i += 1 + n
    return rows
