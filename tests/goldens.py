"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import dataclasses
import glob
import json
import os

import numpy as np

from fmm3d.synth import GridDesc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@dataclasses.dataclass
class Golden:
    name: str
    g: GridDesc
    velv: np.ndarray
    faces: np.ndarray
    rcv: np.ndarray
    fom: int
    rtravel: np.ndarray          # (nsrc, nrc) float32, reference output
    rows: list                   # per (source, receiver): (rid, cols, vals), reference output
    sha256: dict


def names():
    """The small goldens (inputs stored in the fixture); the 201^3 fixture
    regenerates its inputs and loads through load_tele201."""
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "tele_*.npz")))


def load(name: str) -> Golden:
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    g = GridDesc(**json.loads(str(z["grid"])))
    nsrc = z["faces"].shape[0]
    rows = []
    off = 0
    for rid, n in zip(z["frechet_rid"], z["frechet_len"]):
        rows.append((int(rid), z["frechet_cols"][off:off + n], z["frechet_vals"][off:off + n]))
        off += n
    return Golden(name, g, z["velv"], z["faces"], z["rcv_rad"], int(z["fom"]),
                  z["rtravel"].reshape(nsrc, -1), rows, json.loads(str(z["sha256"])))


def row_dict(cols, vals):
    """Row as {column: value}, summing duplicate columns (the reference
    writer can leave the last merged column unsorted and unsummed)."""
    d = {}
    for c, v in zip(np.asarray(cols).tolist(), np.asarray(vals).tolist()):
        d[c] = d.get(c, 0.0) + v
    return d


def compare_rows(a: dict, b: dict, small: float):
    """(max |a-b| / max |b|, columns present in only one row with |v| > small)."""
    keys = set(a) | set(b)
    if not keys:
        return 0.0, []
    va = np.array([a.get(k, 0.0) for k in keys])
    vb = np.array([b.get(k, 0.0) for k in keys])
    rel = float(np.max(np.abs(va - vb)) / max(np.max(np.abs(vb)), 1e-30))
    odd = [k for k in set(a) ^ set(b) if abs(a.get(k, 0.0) + b.get(k, 0.0)) > small]
    return rel, odd


TELE201 = os.path.join(GOLDEN, "tele201_ref.npz")


def has_tele201() -> bool:
    return os.path.exists(TELE201)


@dataclasses.dataclass
class Golden201:
    g: GridDesc
    velv: np.ndarray
    faces: np.ndarray            # (nsrc, nnx, nnz) bottom faces
    rcv: np.ndarray              # (nrc, 3) as the reference read them
    rtravel: np.ndarray          # (nsrc, nrc) float32, reference output
    summary: list                # per row: dict(rid, nnz, colsum, valsum, l2)
    full_rows: int               # receivers per source whose full rows are kept
    full: dict                   # (source, receiver) -> {column: value}


def load_tele201() -> Golden201:
    """The 201^3 fixture (tests/golden/make_golden201.py): the inputs are
    regenerated from the seed and must hash to what the reference read."""
    import tempfile
    import shutil
    from fmm3d import synth
    z = np.load(TELE201, allow_pickle=False)
    nsrc = int(z["nsrc"])
    g, velv, srcs, rcv_deg = synth.config_201(nsrc=50, nrc=int(z["rtravel"].shape[1]), seed=int(z["seed"]))
    srcs = srcs[:nsrc]
    d = tempfile.mkdtemp(prefix="fmm3d_t201_")
    try:
        info = synth.write_case(d, g, velv, srcs, rcv_deg, fom=int(z["fom"]), cfd=1)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    want = json.loads(str(z["sha256"]))
    if info["sha256"] != want:
        raise RuntimeError(f"regenerated 201^3 inputs differ from the fixture's: {info['sha256']} vs {want}")
    summary = [dict(rid=int(a), nnz=int(b), colsum=int(c), valsum=float(v), l2=float(l))
               for a, b, c, v, l in zip(z["row_rid"], z["row_nnz"], z["row_colsum"], z["row_valsum"], z["row_l2"])]
    nrc = z["rtravel"].shape[1]
    fr = int(z["full_rows"])
    full = {}
    off = 0
    k = 0
    for s in range(nsrc):
        for i in range(fr):
            n = int(z["full_len"][k])
            full[(s, i)] = row_dict(z["full_cols"][off:off + n], z["full_vals"][off:off + n])
            off += n
            k += 1
    return Golden201(g, velv, info["faces"], z["rcv_rad"], z["rtravel"], summary, fr, full)
