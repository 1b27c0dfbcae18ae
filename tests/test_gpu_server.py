"""The serving loop: requests of different grids, models, source kinds and
precisions in flight together give the same receiver times as each alone
(float64 within 2e-5 s -- float32 receiver output rounding --, float32
within 1 ms); sources are admitted as slots free up; the HBM budget makes a
second model wait for the first to go idle and be evicted."""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib, synth

pytestmark = pytest.mark.gpu


def _solo(g, velv, rcv, faces=None, src=None, precision=lib.PREC_FP32):
    with lib.Context(g) as c:
        c.set_model(velv)
        if faces is not None:
            c.solve_tele(faces, lib.ORDER_MIXED, precision)
        else:
            c.solve_point(src, 5, 10, lib.ORDER_MIXED, precision)
        return c.receivers(rcv)


def _cases():
    a = G.load("tele_31x33x35_mixed")
    b = G.load("tele_41x45x39_mixed")
    g, velv, _, rcv = synth.tele_case((61, 61, 61), 400.0, (10.0, 10.0), (12, 12, 12), 0.05, 1, 20, 1)
    src = np.array([[-20.0, g.rgox + 30.37 * g.dnx, g.rgoz + 30.61 * g.dnz],
                    [-55.0, g.rgox + 14.2 * g.dnx, g.rgoz + 40.9 * g.dnz]])
    return [
        dict(g=a.g, velv=a.velv, mid=1, faces=a.faces[:4], rcv=a.rcv),
        dict(g=b.g, velv=b.velv, mid=2, faces=b.faces[:3], rcv=b.rcv),
        dict(g=g, velv=velv, mid=3, src=src, rcv=synth.receivers_rad(rcv)),
    ]


def test_results_alone_and_in_the_mix():
    cases = _cases()
    with lib.Server(hbm_budget=16e9, max_slots=8) as S:
        tickets = []
        for prec in (lib.PREC_FP32, lib.PREC_FP64):
            for k in cases:
                t = S.submit(k["g"], k["velv"], k["mid"], faces=k.get("faces"), src=k.get("src"), rcv=k["rcv"], precision=prec)
                tickets.append((t, k, prec))
        got = []
        for t, k, prec in tickets:
            assert S.wait(t, 300.0)
            g, lat = S.result(t)
            got.append(g)
            assert lat > 0.0
        st = S.stats()
        assert st["requests_done"] == len(tickets) and st["sources_in_flight"] == 0
        assert 0.0 < st["p50_ms"] <= st["p99_ms"]
    # the solo solves once the server is gone: a solve that captures its
    # sweeps into a hipGraph must not overlap the server's worker threads
    for g, (t, k, prec) in zip(got, tickets):
        ref = _solo(k["g"], k["velv"], k["rcv"], faces=k.get("faces"), src=k.get("src"), precision=prec)
        tol = 1e-3 if prec == lib.PREC_FP32 else 2e-5
        assert np.nanmax(np.abs(g.astype(np.float64) - ref)) <= tol, (k["mid"], prec)


def test_sources_join_as_slots_free_up():
    a = G.load("tele_31x33x35_mixed")
    with lib.Server(hbm_budget=16e9, max_slots=2) as S:
        ts = [S.submit(a.g, a.velv, 7, faces=a.faces[i % 8:(i % 8) + 2], rcv=a.rcv) for i in range(6)]
        outs = []
        for t in ts:
            assert S.wait(t, 300.0)
            outs.append(S.result(t)[0])
        st = S.stats()
        assert st["sources_done"] == 12 and st["models_resident"] == 1
    ref = _solo(a.g, a.velv, a.rcv, faces=a.faces[0:2])
    assert np.nanmax(np.abs(outs[0].astype(np.float64) - ref)) <= 1e-3


def test_budget_makes_a_second_model_wait_for_eviction():
    a = G.load("tele_31x33x35_mixed")
    b = G.load("tele_41x45x39_mixed")
    # room for one model context (its refined sub-grid allowance dominates), not two
    with lib.Server(hbm_budget=100e6, max_slots=4) as S:
        t1 = S.submit(a.g, a.velv, 1, faces=a.faces[:2], rcv=a.rcv)
        t2 = S.submit(b.g, b.velv, 2, faces=b.faces[:2], rcv=b.rcv)
        assert S.wait(t1, 300.0) and S.wait(t2, 300.0)
        S.result(t1)
        S.result(t2)
        assert S.stats()["models_resident"] == 1


def test_bad_requests_are_rejected():
    a = G.load("tele_31x33x35_mixed")
    with lib.Server(hbm_budget=1e9, max_slots=2) as S:
        with pytest.raises(lib.FmmError):
            S.submit(a.g, a.velv, 1, faces=np.zeros((0, a.g.nnx, a.g.nnz)), rcv=a.rcv)
    # a model that cannot fit the budget even alone fails instead of waiting
    with lib.Server(hbm_budget=1e6, max_slots=2) as S:
        t = S.submit(a.g, a.velv, 1, faces=a.faces[:1], rcv=a.rcv)
        assert S.wait(t, 60.0)
        with pytest.raises(lib.FmmError) as e:
            S.result(t)
        assert e.value.code == lib.E_NOMEM


def test_point_preparation_does_not_hold_up_other_requests():
    """A large point request's refined blocks are solved beside the sweeps:
    a teleseismic request submitted right after it finishes first, and
    both give their solo results."""
    a = G.load("tele_31x33x35_mixed")
    g, velv, _, rcv = synth.tele_case((61, 61, 61), 400.0, (10.0, 10.0), (12, 12, 12), 0.05, 1, 20, 1)
    rcv = synth.receivers_rad(rcv)
    rng = np.random.default_rng(3)
    pts = np.stack([[-rng.uniform(5.0, 60.0), g.rgox + rng.uniform(10, 50) * g.dnx, g.rgoz + rng.uniform(10, 50) * g.dnz]
                    for _ in range(24)])
    with lib.Server(hbm_budget=16e9, max_slots=8) as S:
        tp = S.submit(g, velv, 1, src=pts, refine=5, radius=8, rcv=rcv)
        tt = S.submit(a.g, a.velv, 2, faces=a.faces[:2], rcv=a.rcv)
        assert S.wait(tt, 300.0)
        assert not S.wait(tp, 0.0), "the point request finished before the teleseismic one"
        got_t, _ = S.result(tt)
        assert S.wait(tp, 600.0)
        got_p, _ = S.result(tp)
    assert np.nanmax(np.abs(got_t.astype(np.float64) - _solo(a.g, a.velv, a.rcv, faces=a.faces[:2]))) <= 1e-3
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_point(pts, 5, 8, lib.ORDER_MIXED, lib.PREC_FP32)
        ref = c.receivers(rcv)
    assert np.nanmax(np.abs(got_p.astype(np.float64) - ref)) <= 1e-3
