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
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


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
