"""Point sources (refined neighbourhood) and side-face injection on the GPU,
through the C ABI, against the CPU restatement (tests/test_oracle_sources.py
pins that restatement by closed forms)."""
import numpy as np
import pytest

from fmm3d import lib, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cfg0():
    g, velv, _, rcv = synth.tele_case((61, 61, 61), 400.0, (10.0, 10.0), (12, 12, 12), 0.05, 1, 20, 1)
    srcs = np.array([[-20.0, g.rgox + 30.37 * g.dnx, g.rgoz + 30.61 * g.dnz],     # centred, 20 km
                     [-55.0, g.rgox + 4.2 * g.dnx, g.rgoz + 50.9 * g.dnz],        # near two sides
                     [-3.0, g.rgox + 12.5 * g.dnx, g.rgoz + 0.3 * g.dnz]])        # shallow, on a side
    return g, velv, srcs, synth.receivers_rad(rcv)


@pytest.fixture(scope="module")
def cfg0():
    g, velv, srcs, rcv = _cfg0()
    veln = O.gridder(g, velv)
    ref = [O.point_source(g, velv, veln, s, 5, 10, 1) for s in srcs]
    return g, velv, veln, srcs, rcv, ref


def test_point_sources_fp64_match_oracle(cfg0):
    g, velv, veln, srcs, rcv, ref = cfg0
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_point(srcs, 5, 10, lib.ORDER_MIXED, lib.PREC_FP64)
        tr = c.receivers(rcv)
        for s in range(len(srcs)):
            F = c.field(s)
            assert float(np.abs(F - ref[s]).max()) <= 1e-6, s
            assert np.array_equal(tr[s], O.srtimes(g, F, rcv))


def test_point_source_rays_match_oracle(cfg0):
    g, velv, veln, srcs, rcv, ref = cfg0
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_point(srcs[:1], 5, 10, lib.ORDER_MIXED, lib.PREC_FP64)
        rp, cols, vals = c.frechet(rcv)
        orp, oc, ov, ok, steps = O.rpaths(g, veln, c.field(0), rcv, src=srcs[0])
        assert ok.all() and steps.max() < 1000
        for i in range(len(rcv)):
            a = dict(zip(cols[rp[i]:rp[i + 1]].tolist(), vals[rp[i]:rp[i + 1]].tolist()))
            b = dict(zip(oc[orp[i]:orp[i + 1]].tolist(), ov[orp[i]:orp[i + 1]].tolist()))
            ks = sorted(set(a) | set(b))
            va = np.array([a.get(k, 0.0) for k in ks]); vb = np.array([b.get(k, 0.0) for k in ks])
            assert np.abs(va - vb).max() <= 1e-6 * np.abs(vb).max(), i


def test_point_order_keeps_the_field(cfg0):
    """Patches ordered outward from the source (point_order 1, the default)
    and unordered (0) reach the same fixed point, for the refined blocks and
    the coarse solve; both match the oracle."""
    g, velv, veln, srcs, rcv, ref = cfg0
    fields, sweeps = {}, {}
    with lib.Context(g) as c:
        c.set_model(velv)
        for po in (0, 1):
            c.set_option("point_order", po)
            c.solve_point(srcs, 5, 10, lib.ORDER_MIXED, lib.PREC_FP64)
            fields[po] = [c.field(s) for s in range(len(srcs))]
            sweeps[po] = c.stats()["sweeps"]
        with pytest.raises(lib.FmmError) as e:
            c.set_option("point_order", 2)
        assert e.value.code == lib.E_ARG
    print(f"coarse sweeps: unordered {sweeps[0]}, outward {sweeps[1]}")
    for s in range(len(srcs)):
        assert float(np.abs(fields[1][s] - fields[0][s]).max()) <= 1e-9, s
        assert float(np.abs(fields[1][s] - ref[s]).max()) <= 1e-6, s


def test_point_workers_and_context_pool(cfg0):
    """One worker or several, first call or with the refined contexts pooled
    from earlier calls: the same fields."""
    g, velv, veln, srcs, rcv, ref = cfg0
    fields = []
    with lib.Context(g) as c:
        c.set_model(velv)
        for nw in (1, 3, 3, 16):
            c.set_option("point_workers", nw)
            c.solve_point(srcs, 5, 10, lib.ORDER_MIXED, lib.PREC_FP64)
            fields.append([c.field(s) for s in range(len(srcs))])
        for bad in (0, 17):
            with pytest.raises(lib.FmmError) as e:
                c.set_option("point_workers", bad)
            assert e.value.code == lib.E_ARG
    for run in fields[1:]:
        for s in range(len(srcs)):
            assert float(np.abs(run[s] - fields[0][s]).max()) <= 1e-9, s


def test_point_window_and_group_size_keep_the_field(cfg0):
    """Layer windows on or off, refined blocks in one group or one per
    group, float64 refined sweeps from a float32 solve or from the start:
    the same fixed point, and the oracle's."""
    g, velv, veln, srcs, rcv, ref = cfg0
    runs = {}
    with lib.Context(g) as c:
        c.set_model(velv)
        for pw, mn, warm in ((1, 128, 1), (0, 128, 1), (1, 1, 1), (1, 128, 0)):
            c.set_option("point_window", pw)
            c.set_option("point_group_mnodes", mn)
            c.set_option("point_warm", warm)
            c.solve_point(srcs, 5, 10, lib.ORDER_MIXED, lib.PREC_FP64)
            runs[(pw, mn, warm)] = [c.field(s) for s in range(len(srcs))]
        for key, bad in (("point_window", 2), ("point_group_mnodes", 0), ("point_warm", -1)):
            with pytest.raises(lib.FmmError) as e:
                c.set_option(key, bad)
            assert e.value.code == lib.E_ARG
    base = runs[(1, 128, 1)]
    for key, run in runs.items():
        for s in range(len(srcs)):
            assert float(np.abs(run[s] - base[s]).max()) <= 1e-9, (key, s)
            assert float(np.abs(run[s] - ref[s]).max()) <= 1e-6, (key, s)


def test_point_sources_fp32_within_one_millisecond(cfg0):
    g, velv, veln, srcs, rcv, ref = cfg0
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_point(srcs, 5, 10, lib.ORDER_MIXED, lib.PREC_FP32)
        for s in range(len(srcs)):
            assert float(np.abs(c.field(s) - ref[s]).max()) <= 1e-3, s


def test_point_sources_fp32_receivers_201():
    """32 local sources (5-80 km deep, refinement 5 over 5 nodes) on a 201^3
    grid (0-600 km, 15x15 deg, 25^3 vertices, +-5 %): float32 receiver times
    within 1 ms of float64."""
    g = synth.make_grid((201, 201, 201), 600.0, (15.0, 15.0), nv=(25, 25, 25))
    velv = synth.make_velv(g, 0.05, 5)
    src = np.stack([synth.point_source_rad(g, s) for s in synth.local_sources(g, 32, 5)])
    rcv = synth.receivers_rad(synth.receivers(g, 100, 5))
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_point(src, 5, 5, lib.ORDER_MIXED, lib.PREC_FP64)
        t64 = c.receivers(rcv).astype(np.float64)
        c.solve_point(src, 5, 5, lib.ORDER_MIXED, lib.PREC_FP32)
        t32 = c.receivers(rcv).astype(np.float64)
    d = float(np.nanmax(np.abs(t32 - t64)))
    print(f"float32 vs float64 point-source receivers at 201^3: max |dt| {d:.3e} s")
    assert np.isfinite(t64).all() and d <= 1e-3


def test_side_face_injection_matches_relaxation():
    """Bottom face plus incoming side faces with 1-D tau-p times: the GPU field
    equals the oracle's relaxation of the same fixed set (float64), and the
    float32 solve stays within 1 ms."""
    g, velv, srcs, rcv = synth.tele_case((31, 33, 35), 400.0, (10.0, 10.0), (10, 11, 12), 0.05, 3, 20, 11)
    veln = O.gridder(g, velv)
    sets = [synth.tele_boundary_nodes(g, s, True) for s in srcs]
    refs = []
    for idx, tv in sets:
        fixed = np.zeros(g.nnr * g.nnx * g.nnz, np.uint8)
        fixed[idx] = 1
        tf = np.zeros(fixed.shape)
        tf[idx] = tv
        R, _ = O.relax1_fixed(g, veln, fixed.reshape(g.nnr, g.nnx, g.nnz), tf.reshape(g.nnr, g.nnx, g.nnz), 1, False)
        refs.append(R)
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_fixed(sets, lib.ORDER_MIXED, lib.PREC_FP64)
        for s in range(len(srcs)):
            assert float(np.abs(c.field(s) - refs[s]).max()) <= 1e-7, s
        c.solve_fixed(sets, lib.ORDER_MIXED, lib.PREC_FP32)
        for s in range(len(srcs)):
            assert float(np.abs(c.field(s) - refs[s]).max()) <= 1e-3, s


def test_source_errors():
    g, velv, srcs, rcv = _cfg0()
    with lib.Context(g) as c:
        c.set_model(velv)
        bad = srcs[:1].copy()
        bad[0, 1] += 1.0                       # far outside in colatitude
        with pytest.raises(lib.FmmError) as e:
            c.solve_point(bad)
        assert e.value.code == lib.E_ARG
        with pytest.raises(lib.FmmError) as e:
            c.solve_fixed([(np.array([g.nnr * g.nnx * g.nnz]), np.array([1.0]))])
        assert e.value.code == lib.E_ARG
        with pytest.raises(lib.FmmError) as e:
            c.solve_fixed([(np.array([5]), np.array([np.nan]))])
        assert e.value.code == lib.E_ARG


def test_side_faces_as_face_arrays_equal_the_fixed_set():
    """fmm3d_solve_tele_sides (bottom + one theta and one phi face as arrays)
    relaxes the same fixed set as fmm3d_solve_fixed with the node lists of
    tele_boundary_nodes: float64 fields within 1e-9 s (different patch
    orders, same fixed point), float32 within 1 ms of them."""
    g, velv, srcs, rcv = synth.tele_case((31, 33, 35), 400.0, (10.0, 10.0), (10, 11, 12), 0.05, 4, 12, 41)
    parts = [synth.tele_faces(g, s) for s in srcs]
    faces = np.stack([p[0] for p in parts]); xf = np.stack([p[1] for p in parts])
    zf = np.stack([p[2] for p in parts]); side = np.stack([p[3] for p in parts])
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_fixed([synth.tele_boundary_nodes(g, s) for s in srcs], lib.ORDER_MIXED, lib.PREC_FP64)
        ref = [c.field(k) for k in range(len(srcs))]
        c.solve_tele_sides(faces, xf, zf, side, lib.ORDER_MIXED, lib.PREC_FP64)
        for k in range(len(srcs)):
            assert float(np.abs(c.field(k) - ref[k]).max()) <= 1e-9
        c.solve_tele_sides(faces, xf, zf, side, lib.ORDER_MIXED, lib.PREC_FP32)
        for k in range(len(srcs)):
            assert float(np.abs(c.field(k) - ref[k]).max()) <= 1e-3
        # a face left out (-1) is not injected: the result is the bottom-only solve
        none = np.full_like(side, -1)
        c.solve_tele_sides(faces, xf, zf, none, lib.ORDER_MIXED, lib.PREC_FP64)
        a = c.field(0)
        c.solve_tele(faces[:1], lib.ORDER_MIXED, lib.PREC_FP64)
        assert float(np.abs(a - c.field(0)).max()) <= 1e-9
        bad = side.copy(); bad[0, 0] = 2
        with pytest.raises(lib.FmmError) as e:
            c.solve_tele_sides(faces, xf, zf, bad)
        assert e.value.code == lib.E_ARG


def test_mixed_batch_equals_single_kind_solves():
    """fmm3d_solve_mixed: teleseismic sources (bottom and side faces) and
    point sources of different refinement radii in one batch give, slot by
    slot, the fields of the single-kind calls (float64 within 1e-9 s: the
    same fixed points), the same receiver times and Frechet rows (rays of
    the point slots end at their sources), and float32 within 1 ms."""
    g, velv, tele, rcv = synth.tele_case((61, 61, 61), 400.0, (10.0, 10.0), (12, 12, 12), 0.05, 3, 20, 1)
    rcv = synth.receivers_rad(rcv)
    parts = [synth.tele_faces(g, s) for s in tele]
    F, X, Z, S = (np.stack([p[k] for p in parts]) for k in range(4))
    pts = np.array([[-20.0, g.rgox + 30.37 * g.dnx, g.rgoz + 30.61 * g.dnz],
                    [-55.0, g.rgox + 14.2 * g.dnx, g.rgoz + 40.9 * g.dnz]])
    rad = np.array([3, 7])
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_tele_sides(F, X, Z, S, lib.ORDER_MIXED, lib.PREC_FP64)
        ref = [c.field(k) for k in range(3)]
        ref_t = list(c.receivers(rcv))
        rp, cols, vals = c.frechet(rcv)
        ref_rows = [(cols[rp[i]:rp[i + 1]], vals[rp[i]:rp[i + 1]]) for i in range(len(rp) - 1)]
        for k in range(2):
            c.solve_point(pts[k:k + 1], 5, int(rad[k]), lib.ORDER_MIXED, lib.PREC_FP64)
            ref.append(c.field(0))
            ref_t.append(c.receivers(rcv)[0])
            rp, cols, vals = c.frechet(rcv)
            ref_rows += [(cols[rp[i]:rp[i + 1]], vals[rp[i]:rp[i + 1]]) for i in range(len(rp) - 1)]
        c.solve_mixed(F, (X, Z, S), pts, rad, 5, lib.ORDER_MIXED, lib.PREC_FP64)
        assert c.nsrc == 5
        for k in range(5):
            d = float(np.abs(c.field(k) - ref[k]).max())
            assert d <= 1e-9, (k, d)
        t = c.receivers(rcv)
        for k in range(5):
            assert np.abs(t[k] - ref_t[k]).max() <= np.spacing(np.abs(ref_t[k]).max()), k
        rp, cols, vals = c.frechet(rcv)
        assert len(rp) - 1 == len(ref_rows)
        for r, (rc_, rv) in enumerate(ref_rows):
            a = dict(zip(cols[rp[r]:rp[r + 1]].tolist(), vals[rp[r]:rp[r + 1]].astype(np.float64).tolist()))
            b = dict(zip(rc_.tolist(), rv.astype(np.float64).tolist()))
            assert set(a) == set(b), r
            scale = max([abs(v) for v in b.values()] + [1e-30])
            assert max([abs(a[q] - b[q]) for q in b] + [0.0]) <= 1e-6 * scale, r
        c.solve_mixed(F, (X, Z, S), pts, rad, 5, lib.ORDER_MIXED, lib.PREC_FP32)
        t32 = c.receivers(rcv)
        assert np.abs(t32.astype(np.float64) - t).max() <= 1e-3
        # one kind alone through the mixed call
        c.solve_tele(F[2:3], lib.ORDER_MIXED, lib.PREC_FP64)
        bottom = c.field(0)
        c.solve_mixed(faces=F, precision=lib.PREC_FP64)
        assert c.nsrc == 3 and float(np.abs(c.field(2) - bottom).max()) <= 1e-9
        c.solve_mixed(src=pts, radius=rad, precision=lib.PREC_FP64)
        assert float(np.abs(c.field(1) - ref[4]).max()) <= 1e-9
        with pytest.raises(lib.FmmError) as e:
            c.solve_mixed(src=pts, radius=np.array([3, 0]))
        assert e.value.code == lib.E_ARG
        with pytest.raises(lib.FmmError) as e:
            c.solve_mixed()
        assert e.value.code == lib.E_ARG
