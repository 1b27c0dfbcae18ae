"""The oracle's restatement of subinv's sparse products (or_gmat_gtx,
or_gmat_gx: float32 sums in subinv's order) against an independent float64
computation (scipy.sparse) on the reference's own Frechet rows (the golden
frechet.out) and on a random matrix: equal to float32 rounding."""
import numpy as np
from scipy.sparse import csr_matrix

import goldens as G
from oracle import oracle as O


def golden_matrix():
    gd = G.load("tele_41x45x39_mixed")
    rows = [G.row_dict(c, v) for _, c, v in gd.rows]
    rp = np.zeros(len(rows) + 1, np.int64)
    cols, vals = [], []
    for i, d in enumerate(rows):
        ks = sorted(d)
        cols += ks
        vals += [d[k] for k in ks]
        rp[i + 1] = rp[i] + len(ks)
    ncols = (gd.g.nvr + 2) * (gd.g.nvx + 2) * (gd.g.nvz + 2)
    return rp, np.array(cols, np.int32), np.array(vals, np.float32), ncols


def random_matrix(nrows=300, ncols=2000, per=40, seed=3):
    rng = np.random.default_rng(seed)
    rp = np.zeros(nrows + 1, np.int64)
    cols = []
    for i in range(nrows):
        c = np.sort(rng.choice(ncols, per, replace=False)) + 1
        cols.append(c)
        rp[i + 1] = rp[i] + per
    cols = np.concatenate(cols).astype(np.int32)
    vals = rng.normal(size=cols.size).astype(np.float32)
    return rp, cols, vals, ncols


def test_products_match_float64_on_reference_rows_and_random():
    rng = np.random.default_rng(1)
    for rp, cols, vals, ncols in (golden_matrix(), random_matrix()):
        A = csr_matrix((vals.astype(np.float64), cols.astype(np.int64) - 1, rp), shape=(len(rp) - 1, ncols))
        d = rng.normal(size=len(rp) - 1).astype(np.float32)
        a = rng.normal(size=ncols).astype(np.float32)
        cd = rng.uniform(0.5, 2.0, size=len(rp) - 1).astype(np.float32)
        y = O.gmat_gtx(rp, cols, vals, ncols, d)
        ref = A.T @ d.astype(np.float64)
        assert np.abs(y - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-30) * 10
        z = O.gmat_gx(rp, cols, vals, a, cd)
        ref = (A @ a.astype(np.float64)) / cd
        assert np.abs(z - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-30) * 10
        assert np.array_equal(O.gmat_gx(rp, cols, vals, a), np.float32(1.0) * O.gmat_gx(rp, cols, vals, a, np.ones_like(cd)))
