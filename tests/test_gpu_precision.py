"""Precision study at 321^3 (0-600 km, 15x15 deg, 33^3 vertices, a +10 %
dipping slab and ten +-5 % blobs, seed 4): 64 teleseismic plane waves
injected on the bottom face and the two side faces they enter through,
500 surface receivers.

* float64 on the device against the CPU oracle's value-based relaxation
  (tests/golden/precision321/oracle.npz, sources 0, 21 and 42; with side
  faces the heap march solves a different system, see
  tests/golden/make_precision321.py): receiver times within one float32
  ulp (the output is float32; times are ~470-700 s),
  Frechet rows of the first 40 receivers with the same columns and values
  within 1e-4 relative, and every row's entry count, column sum and value
  2-norm (1e-4 relative);
* float32 updates against float64 on the device, all 64 x 500 receivers:
  max |dt| <= 1 ms.  float32 fields are stored relative to each source's
  plane-wave reference (offset + 1-D vertical profile + lateral plane).
  Rays traced on the float32 field follow its gradient; the per-row
  relative 2-norm error against the float64 rows is printed and held to a
  median of 1e-5, at most 1 % of rows above 1e-4 and a maximum of 5e-2
  (measured: median 3.5e-6, p99 3.3e-5, 0.4 % above 1e-4, maximum 2.6e-2 --
  a ray that crosses a cell edge where the two fields order differently).
"""
import hashlib
import os

import numpy as np
import pytest

from fmm3d import lib, synth

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "precision321", "oracle.npz")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def row_errors(rp_a, ca, va, rp_b, cb, vb, ncol):
    """Per row: relative 2-norm of (b - a) against a, and whether both rows
    hold the same columns.  Rows are CSR (row_ptr, column, value)."""
    nr = len(rp_a) - 1
    ra = np.repeat(np.arange(nr), np.diff(rp_a))
    rb = np.repeat(np.arange(nr), np.diff(rp_b))
    keys = np.concatenate([ra * ncol + ca, rb * ncol + cb])
    w = np.concatenate([-va.astype(np.float64), vb.astype(np.float64)])
    uk, inv = np.unique(keys, return_inverse=True)
    diff = np.bincount(inv, weights=w, minlength=len(uk))
    both = np.bincount(inv, minlength=len(uk)) == 2
    urow = uk // ncol
    num = np.bincount(urow, weights=diff * diff, minlength=nr)
    den = np.bincount(ra, weights=va.astype(np.float64) ** 2, minlength=nr)
    only = np.bincount(urow, weights=(~both).astype(np.float64), minlength=nr)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.sqrt(num / den)
    return rel, only == 0


@pytest.fixture(scope="module")
def study():
    g, velv, srcs, rcv = synth.config_321(nsrc=64, nrc=500, seed=4)
    rr = synth.receivers_rad(rcv)
    parts = [synth.tele_faces(g, s) for s in srcs]
    F, X, Z, S = (np.stack([p[k] for p in parts]) for k in range(4))
    ref = np.load(FIX, allow_pickle=False)
    assert _sha(velv) == str(ref["velv_sha"]) and _sha(rr) == str(ref["rcv_sha"])
    for k, s in enumerate(ref["sources"]):
        _, tv = synth.tele_boundary_nodes(g, srcs[int(s)])
        assert _sha(tv) == str(ref["bound_sha"][k])
    out = {}
    with lib.Context(g) as c:
        c.set_model(velv)
        for prec in (lib.PREC_FP64, lib.PREC_FP32):
            print(f"precision {prec}: solving 64 sources at 321^3", flush=True)
            c.solve_tele_sides(F, X, Z, S, lib.ORDER_MIXED, prec)
            st = c.stats()
            t = c.receivers(rr)
            rp, cols, vals = c.frechet(rr)
            out[prec] = (t, rp.copy(), cols.copy(), vals.copy(), st)
            print(f"  solve {st['solve_ms']:.0f} ms, {st['sweeps']} sweeps, {rp[-1]} Frechet entries", flush=True)
    return g, ref, out


def test_fp64_receivers_match_oracle(study):
    g, ref, out = study
    t64 = out[lib.PREC_FP64][0]
    for k, s in enumerate(ref["sources"]):
        o = ref["times"][k].astype(np.float32)
        d = np.abs(t64[int(s)] - o)
        assert np.isfinite(o).all() and (d <= np.spacing(o)).all(), (int(s), float(d.max()))


def test_fp64_frechet_rows_match_oracle(study):
    g, ref, out = study
    _, rp, cols, vals, _ = out[lib.PREC_FP64]
    nrc, nrows = 500, int(ref["nrows"])
    off = np.concatenate([[0], np.cumsum(ref["row_len"])])
    worst = 0.0
    for k, s in enumerate(ref["sources"]):
        for i in range(nrc):
            r = int(s) * nrc + i
            c, v = cols[rp[r]:rp[r + 1]], vals[rp[r]:rp[r + 1]].astype(np.float64)
            j = k * nrc + i
            assert abs(len(c) - int(ref["nnz"][j])) <= 2, (int(s), i)
            vn = float(np.sqrt((v * v).sum()))
            assert abs(vn - ref["vnorm"][j]) <= 1e-4 * ref["vnorm"][j], (int(s), i, vn, ref["vnorm"][j])
            if len(c) == int(ref["nnz"][j]):
                assert int(c.astype(np.int64).sum()) == int(ref["colsum"][j]), (int(s), i)
            if i < nrows:
                q = k * nrows + i
                oc, ov = ref["row_cols"][off[q]:off[q + 1]], ref["row_vals"][off[q]:off[q + 1]].astype(np.float64)
                a = dict(zip(c.tolist(), v.tolist()))
                b = dict(zip(oc.tolist(), ov.tolist()))
                for col in set(a) ^ set(b):
                    # a column on one side only sits at the 1e-6 keep threshold
                    assert abs(a.get(col, b.get(col))) <= 1.1e-6, (int(s), i, col)
                scale = max(abs(x) for x in b.values())
                rel = max(abs(a.get(col, 0.0) - b.get(col, 0.0)) for col in set(a) | set(b)) / scale
                worst = max(worst, rel)
                assert rel <= 1e-4, (int(s), i, rel)
    print(f"float64 rows vs oracle: worst relative error {worst:.2e}")


def test_fp32_receivers_within_one_millisecond(study):
    g, ref, out = study
    t64, t32 = out[lib.PREC_FP64][0], out[lib.PREC_FP32][0]
    assert np.isfinite(t64).all() and np.isfinite(t32).all()
    d = np.abs(t32.astype(np.float64) - t64)
    print(f"float32 vs float64 receivers: max |dt| {d.max():.3e} s, rms {np.sqrt((d * d).mean()):.3e} s")
    for k, s in enumerate(ref["sources"]):
        do = np.abs(t32[int(s)].astype(np.float64) - ref["times"][k])
        print(f"  source {int(s)} vs oracle: max |dt| {do.max():.3e} s, rms {np.sqrt((do * do).mean()):.3e} s")
    assert d.max() <= 1e-3


def test_fp32_frechet_error_is_stated(study):
    g, ref, out = study
    _, rp64, c64, v64, _ = out[lib.PREC_FP64]
    _, rp32, c32, v32, _ = out[lib.PREC_FP32]
    ncol = (g.nvr + 2) * (g.nvx + 2) * (g.nvz + 2) + 1
    rel, same = row_errors(rp64, c64, v64, rp32, c32, v32, ncol)
    rel = rel[np.isfinite(rel)]
    print(f"float32 vs float64 Frechet rows: {len(rel)} rows, same columns in {int(same.sum())}, "
          f"relative 2-norm error median {np.median(rel):.2e} p99 {np.percentile(rel, 99):.2e} "
          f"max {rel.max():.2e}, rows above 1e-4: {(rel > 1e-4).mean():.3f}")
    # float32 fields are stored against each source's plane-wave reference:
    # 99 % of rows within 1e-4; a ray whose path crosses a cell edge where the
    # two fields order differently takes another path (the rare large rows)
    assert np.median(rel) <= 1e-5 and (rel > 1e-4).mean() <= 0.01 and rel.max() <= 5e-2
