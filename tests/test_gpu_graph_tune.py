"""Sweep chunks replayed from a captured hipGraph give the fields the direct
launches give, solve after solve with other work in between; the autotuner
picks one of its candidates and its solve matches the untuned one."""
import numpy as np
import pytest

import goldens as G
from fmm3d import lib, synth

pytestmark = pytest.mark.gpu


def test_graph_replays_match_direct_launches():
    gd = G.load("tele_41x45x39_mixed")
    with lib.Context(gd.g) as d, lib.Context(gd.g) as c:
        d.set_model(gd.velv)
        c.set_model(gd.velv)
        d.set_option("graph", 0)
        c.set_option("graph", 1)
        for rep in range(3):
            for prec, tol in ((lib.PREC_FP64, 1e-7), (lib.PREC_FP32, 1e-3)):
                faces = gd.faces[rep:rep + 4]
                d.solve_tele(faces, gd.fom, prec)
                c.solve_tele(faces, gd.fom, prec)
                for s in range(faces.shape[0]):
                    assert float(np.abs(c.field(s) - d.field(s)).max()) <= tol, (rep, prec, s)
                # eager work between replays
                c.receivers(gd.rcv)


def test_autotuner_choice_and_result():
    g, velv, srcs, rcv = synth.config_201(nsrc=8, nrc=50, seed=2)
    g2, velv2, _, _ = synth.tele_case((61, 61, 61), 400.0, (10.0, 10.0), (12, 12, 12), 0.05, 1, 20, 3)
    faces = np.stack([synth.tele_bottom_face(g2, s) for s in synth.tele_sources(g2, 6, 3)])
    with lib.Context(g2) as c, lib.Context(g2) as d:
        c.set_model(velv2)
        d.set_model(velv2)
        c.set_option("autotune", 1)
        c.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP64)
        st = c.stats()
        assert (st["tile_x"], st["tile_z"], st["tile_r"]) in {(8, 8, 16), (8, 8, 8), (8, 16, 16), (16, 8, 16),
                                                             (16, 16, 16), (16, 16, 8), (8, 16, 8), (16, 8, 8)}
        d.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP64)
        for s in range(faces.shape[0]):
            assert float(np.abs(c.field(s) - d.field(s)).max()) <= 1e-7
        # the cached choice is reused without re-tuning
        c.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP64)
        assert c.stats()["tile_x"] == st["tile_x"] and c.stats()["tile_r"] == st["tile_r"]


@pytest.mark.parametrize("every", [1, 3, 8])
def test_check_interval_and_per_source_exit(every):
    """Convergence checks every `every` sweeps (graph chunks of that many
    launches; a source whose last sweep changed nothing leaves the sweeps at
    each check) reach the fields of the default 4-sweep chunks: float64
    within 1e-9 s, float32 within 1 ms.  The batch mixes sources that
    converge at different sweeps: plane waves and point sources."""
    gd = G.load("tele_41x45x39_mixed")
    g = gd.g
    pts = np.array([[-30.0, g.rgox + 20.4 * g.dnx, g.rgoz + 17.2 * g.dnz]])
    with lib.Context(g) as d, lib.Context(g) as c:
        d.set_model(gd.velv)
        c.set_model(gd.velv)
        c.set_option("check_every", every)
        for prec, tol in ((lib.PREC_FP64, 1e-9), (lib.PREC_FP32, 1e-3)):
            d.solve_mixed(gd.faces[:5], None, pts, np.array([4]), 5, gd.fom, prec)
            c.solve_mixed(gd.faces[:5], None, pts, np.array([4]), 5, gd.fom, prec)
            for s in range(6):
                assert float(np.abs(c.field(s) - d.field(s)).max()) <= tol, (every, prec, s)
        with pytest.raises(lib.FmmError):
            c.set_option("check_every", 9)


@pytest.mark.parametrize("verify,narrow", [(1, 0), (1, 1), (2, -1)])
def test_fixed_point_check_keeps_the_fields(verify, narrow):
    """The read-only fixed-point check after sweep `verify` (sources with no
    node that would move leave the sweeps; with narrowing the next sweep
    relaxes only the moving patches) reaches the fields of the sweeps
    without it: float64 within 1e-9 s, float32 within 1 ms; plane waves with
    side faces and a point source in one batch."""
    g, velv, tele, rcv = synth.tele_case((41, 45, 39), 500.0, (12.0, 11.0), (9, 10, 11), 0.08, 5, 20, 7)
    parts = [synth.tele_faces(g, s) for s in tele]
    F, X, Z, S = (np.stack([p[k] for p in parts]) for k in range(4))
    pts = np.array([[-30.0, g.rgox + 20.4 * g.dnx, g.rgoz + 17.2 * g.dnz]])
    with lib.Context(g) as d, lib.Context(g) as c:
        d.set_model(velv)
        c.set_model(velv)
        d.set_option("verify", 0)
        c.set_option("verify", verify)
        c.set_option("verify_patches", narrow)
        for prec, tol in ((lib.PREC_FP64, 1e-9), (lib.PREC_FP32, 1e-3)):
            d.solve_mixed(F, (X, Z, S), pts, np.array([4]), 5, lib.ORDER_MIXED, prec)
            c.solve_mixed(F, (X, Z, S), pts, np.array([4]), 5, lib.ORDER_MIXED, prec)
            for s in range(6):
                assert float(np.abs(c.field(s) - d.field(s)).max()) <= tol, (verify, narrow, prec, s)
        with pytest.raises(lib.FmmError):
            c.set_option("verify_patches", 2)
