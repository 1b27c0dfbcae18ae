"""Write the oracle fixture of the 321^3 precision study
(tests/golden/precision321/oracle.npz).

Model: 321^3 nodes over 0-600 km and 15x15 deg, a 33^3 B-spline vertex grid
holding the 1-D gradient times a +10 % dipping slab (60 km, 45 deg) and ten
+-5 % Gaussian blobs, seed 4 (fmm3d.synth.config_321); 64 teleseismic plane
waves injected on the bottom face and the two side faces they enter
through; 500 surface receivers.  The CPU oracle (oracle/fmm3d_oracle.c: the
float64 value-based relaxation from the fixed boundary set to its fixed
point, srtimes and rpaths) solves the sources in SOURCES; the fixture keeps
their receiver times, full Frechet rows of the first NROWS receivers and,
for every receiver, the row's entry count, column sum and value 2-norm.

The relaxation, not the heap march, is the oracle here: a neighbour enters
a node's stencil only while its time is below the node's (the predicate of
fm3dt_parallel.f90:1124, 1149), whereas the heap march takes every alive
neighbour.  With bottom faces alone the two agree bit for bit; side faces
carry 1-D times that can be later than the 3-D arrival beside them, and
the heap march then builds stencils on those later (downwind) fixed nodes.

The inputs are regenerated from the seed by the tests; the SHA-256 of the
vertex model and of the boundary times is stored so a test can tell that
they did.

Run from the repository root:  python tests/golden/make_precision321.py
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "parallel-fast-marching-method_amd"), REPO]

from fmm3d import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

SOURCES = (0, 21, 42)
NROWS = 40
OUT = os.path.join(HERE, "precision321", "oracle.npz")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    g, velv, srcs, rcv = synth.config_321(nsrc=64, nrc=500, seed=4)
    rr = synth.receivers_rad(rcv)
    veln = O.gridder(g, velv)
    times, nnz, csum, vnorm = [], [], [], []
    rcols, rvals, rlen = [], [], []
    bsha = []
    for s in SOURCES:
        t0 = time.time()
        idx, tv = synth.tele_boundary_nodes(g, srcs[s])
        bsha.append(sha(tv))
        fixed = np.zeros((g.nnr, g.nnx, g.nnz), np.uint8)
        tfix = np.zeros(fixed.shape)
        fixed.flat[idx] = 1
        tfix.flat[idx] = tv
        T, sweeps = O.relax1_fixed(g, veln, fixed, tfix, 1)
        assert sweeps > 0, "the oracle relaxation must converge"
        times.append(O.srtimes(g, T, rr))
        rp, cols, vals, ok, _ = O.rpaths(g, veln, T, rr)
        assert ok.all(), "every oracle ray must reach the boundary"
        for r in range(len(rr)):
            c, v = cols[rp[r]:rp[r + 1]], vals[rp[r]:rp[r + 1]]
            nnz.append(len(c))
            csum.append(int(c.astype(np.int64).sum()))
            vnorm.append(float(np.sqrt((v.astype(np.float64) ** 2).sum())))
            if r < NROWS:
                rcols.append(c)
                rvals.append(v)
                rlen.append(len(c))
        print(f"source {s}: {time.time() - t0:.1f} s, {sweeps} sweeps, times {T.min():.3f}..{T.max():.3f} s", flush=True)
        del T
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    np.savez_compressed(
        OUT, sources=np.array(SOURCES, np.int32), nrows=np.int32(NROWS),
        times=np.array(times, np.float32), nnz=np.array(nnz, np.int32),
        colsum=np.array(csum, np.int64), vnorm=np.array(vnorm, np.float64),
        row_len=np.array(rlen, np.int32), row_cols=np.concatenate(rcols).astype(np.int32),
        row_vals=np.concatenate(rvals).astype(np.float32),
        velv_sha=sha(velv), bound_sha=np.array(bsha), rcv_sha=sha(rr))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
