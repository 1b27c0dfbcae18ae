"""The batched and serving configurations at their real shapes.

Batched: 256 sources in one fmm3d_solve_mixed call on the 201^3 model --
192 plane waves (bottom and entry side faces, seed 3) and 64 local sources
(5-80 km deep, refinement 5 over 5-20 nodes, seed 3).  Float32 receivers
within 1 ms of float64; sampled slots of each kind equal their single-kind
solves (float64 fields within 1e-8 s); point sources give the same float64
fields with and without the float32 warm start of their refined blocks.

Serving: requests over 65^3, 129^3, 201^3 and 257^3 grids (+-5 % models),
teleseismic and local, in flight together; each request's receiver times
are those it gets alone (float32 within 1 ms, float64 within 2e-5 s -- the
float32 output rounding), and its Frechet rows (float64) are the rows of
the same sources solved as a batch.
"""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib, synth

pytestmark = pytest.mark.gpu


def test_config3_mixed_256_sources_201():
    g, velv, _, rcv_deg = synth.config_201(nsrc=1, nrc=300, seed=2)
    rcv = synth.receivers_rad(rcv_deg)
    tele = synth.tele_sources(g, 192, 3)
    loc = synth.local_sources(g, 64, 3)
    parts = [synth.tele_faces(g, s) for s in tele]
    F, X, Z, S = (np.stack([p[k] for p in parts]) for k in range(4))
    src = np.stack([synth.point_source_rad(g, s) for s in loc])
    rad = np.array([s.radius_nodes for s in loc], np.int32)
    with lib.Context(g) as c:
        c.set_model(velv)
        print("mixed 256, float64", flush=True)
        c.solve_mixed(F, (X, Z, S), src, rad, 5, lib.ORDER_MIXED, lib.PREC_FP64)
        assert c.nsrc == 256
        t64 = c.receivers(rcv)
        picks_t, picks_p = (0, 97, 191), (0, 31, 63)
        f64 = {k: c.field(k) for k in picks_t}
        f64.update({192 + q: c.field(192 + q) for q in picks_p})
        print("mixed 256, float32", flush=True)
        c.solve_mixed(F, (X, Z, S), src, rad, 5, lib.ORDER_MIXED, lib.PREC_FP32)
        t32 = c.receivers(rcv)
        err = float(np.nanmax(np.abs(t32.astype(np.float64) - t64)))
        print(f"float32 vs float64 receivers: {err:.3e} s", flush=True)
        assert err <= 1e-3
        # sampled slots against their single-kind solves
        for k in picks_t:
            c.solve_tele_sides(F[k:k + 1], X[k:k + 1], Z[k:k + 1], S[k:k + 1], lib.ORDER_MIXED, lib.PREC_FP64)
            d = float(np.abs(c.field(0) - f64[k]).max())
            assert d <= 1e-8, (k, d)
        for q in picks_p:
            for warm in (1, 0):
                c.set_option("point_warm", warm)
                c.solve_point(src[q:q + 1], 5, int(rad[q]), lib.ORDER_MIXED, lib.PREC_FP64)
                d = float(np.nanmax(np.abs(c.field(0) - f64[192 + q])))
                print(f"point slot {q} (radius {rad[q]}) alone, warm start {warm}: {d:.2e} s", flush=True)
                assert d <= 1e-8, (q, warm, d)
        c.set_option("point_warm", 1)


def _grid(n, seed):
    g = synth.make_grid((n, n, n), 600.0, (15.0, 15.0), nv=(20, 20, 20))
    velv = synth.make_velv(g, 0.05, seed)
    return g, velv


def test_config5_serving_uneven_grids_alone_and_in_the_mix():
    cases = []
    for k, n in enumerate((65, 129, 201, 257)):
        g, velv = _grid(n, 5 + k)
        rcv = synth.receivers_rad(synth.receivers(g, 40, 5 + k))
        tele = synth.tele_sources(g, 8, 5 + k)
        faces = np.stack([synth.tele_bottom_face(g, s) for s in tele])
        loc = synth.local_sources(g, 4, 5 + k)
        pts = np.stack([synth.point_source_rad(g, s) for s in loc])
        cases.append(dict(n=n, g=g, velv=velv, rcv=rcv, faces=faces, pts=pts, mid=100 + k))
    with lib.Server(hbm_budget=64e9, max_slots=16) as srv:
        tickets = []
        for c in cases:
            tickets.append((srv.submit(c["g"], c["velv"], c["mid"], faces=c["faces"], rcv=c["rcv"],
                                       precision=lib.PREC_FP32), c, "tele", lib.PREC_FP32))
            tickets.append((srv.submit(c["g"], c["velv"], c["mid"], src=c["pts"], refine=5, radius=5, rcv=c["rcv"],
                                       precision=lib.PREC_FP32), c, "point", lib.PREC_FP32))
        for c in cases[1:3]:
            tickets.append((srv.submit(c["g"], c["velv"], c["mid"], faces=c["faces"][:3], rcv=c["rcv"],
                                       precision=lib.PREC_FP64, rows=True), c, "tele64", lib.PREC_FP64))
        got = []
        for t, c, kind, prec in tickets:
            assert srv.wait(t, 600.0), (c["n"], kind)
            rows = srv.rows(t) if kind == "tele64" else None
            times, lat = srv.result(t)
            got.append((c, kind, prec, times, rows, lat))
        st = srv.stats()
        print(f"served {st['requests_done']} requests, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms", flush=True)
        assert st["requests_done"] == len(tickets) and st["sources_in_flight"] == 0
    for c, kind, prec, times, rows, lat in got:
        with lib.Context(c["g"]) as ctx:
            ctx.set_model(c["velv"])
            if kind == "point":
                ctx.solve_point(c["pts"], 5, 5, lib.ORDER_MIXED, prec)
            else:
                ctx.solve_tele(c["faces"][:3] if kind == "tele64" else c["faces"], lib.ORDER_MIXED, prec)
            ref = ctx.receivers(c["rcv"])
            tol = 1e-3 if prec == lib.PREC_FP32 else 2e-5
            d = float(np.nanmax(np.abs(times.astype(np.float64) - ref)))
            print(f"{c['n']}^3 {kind}: in the mix vs alone {d:.2e} s, latency {lat:.0f} ms", flush=True)
            assert d <= tol, (c["n"], kind, d)
            if rows is not None:
                rp, cols, vals = ctx.frechet(c["rcv"])
                srp, scols, svals = rows
                assert len(srp) == len(rp)
                for r in range(len(rp) - 1):
                    a = G.row_dict(scols[srp[r]:srp[r + 1]], svals[srp[r]:srp[r + 1]])
                    b = G.row_dict(cols[rp[r]:rp[r + 1]], vals[rp[r]:rp[r + 1]])
                    rel, odd = G.compare_rows(a, b, small=1.1e-6)
                    assert rel <= 1e-4 and not odd, (c["n"], r, rel, odd)


def test_server_rows_need_the_flag():
    a = G.load("tele_31x33x35_mixed")
    with lib.Server(hbm_budget=8e9, max_slots=4) as srv:
        t = srv.submit(a.g, a.velv, 1, faces=a.faces[:2], rcv=a.rcv)
        assert srv.wait(t, 120.0)
        with pytest.raises(lib.FmmError) as e:
            srv.rows(t)
        assert e.value.code == lib.E_STATE
        srv.result(t)
