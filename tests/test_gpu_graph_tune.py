"""Sweep chunks replayed from a captured hipGraph give the fields the direct
launches give, solve after solve with other work in between; the autotuner
times its candidates at the first solve of a kind, caches the choice in its
file and reuses it."""
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


def test_graph_follows_the_check_interval():
    """One context changing its check interval between solves: each solve
    replays a graph of its own chunk length (a graph captured for 2 sweeps
    replayed for a 4-sweep chunk would skip sweeps); fields equal those of
    direct launches."""
    gd = G.load("tele_41x45x39_mixed")
    with lib.Context(gd.g) as d, lib.Context(gd.g) as c:
        d.set_model(gd.velv)
        c.set_model(gd.velv)
        d.set_option("graph", 0)
        c.set_option("graph", 1)
        for every in (2, 4, 2, 3):
            for prec, tol in ((lib.PREC_FP64, 1e-7), (lib.PREC_FP32, 1e-3)):
                c.set_option("check_every", every)
                d.set_option("check_every", every)
                d.solve_tele(gd.faces, gd.fom, prec)
                c.solve_tele(gd.faces, gd.fom, prec)
                for s in range(gd.faces.shape[0]):
                    assert float(np.abs(c.field(s) - d.field(s)).max()) <= tol, (every, prec, s)


def _tune_case(n=129, nsrc=8, seed=3):
    g = synth.make_grid((n, n, n), 600.0, (15.0, 15.0), nv=(20, 20, 20))
    velv = synth.make_velv(g, 0.05, seed)
    faces = np.stack([synth.tele_bottom_face(g, s) for s in synth.tele_sources(g, nsrc, seed)])
    return g, velv, faces


def test_autotuner_times_caches_and_reuses(tmp_path, monkeypatch):
    """The first float32 solve of a kind times the candidates and writes its
    choice to the tuning cache; a new context reads it back (same
    configuration, the same fields bit for bit); a choice written to the
    cache by hand is applied as it stands."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("FMM3D_TUNE_CACHE", str(cache))
    g, velv, faces = _tune_case()
    cands = {(8, 8, 16), (8, 8, 8), (8, 16, 16), (16, 8, 16), (16, 16, 16)}
    with lib.Context(g) as c:
        c.set_model(velv)
        c.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP32)
        st = c.stats()
        fields = [c.field(s) for s in range(faces.shape[0])]
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and " = " in lines[0], lines
    key, val = lines[0].split(" = ")
    strategy, px, pz, passes, check = map(int, val.split())
    print(f"tuned: {lines[0]}", flush=True)
    assert f"{g.nnr}x{g.nnx}x{g.nnz}" in key
    assert strategy in (1, 2) and check in (2, 4)
    if strategy == 1:
        assert (px, pz, passes) in cands and (st["tile_x"], st["tile_z"], st["tile_r"]) == (px, pz, passes)
    with lib.Context(g) as d:
        d.set_model(velv)
        d.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP32)
        assert (d.stats()["tile_x"], d.stats()["tile_z"], d.stats()["tile_r"]) == (st["tile_x"], st["tile_z"], st["tile_r"])
        for s in range(faces.shape[0]):
            assert np.array_equal(d.field(s), fields[s]), s
    assert len(cache.read_text().splitlines()) == 1   # read, not re-tuned
    # a hand-written choice for the same key (the last line of a key wins)
    with open(cache, "a") as f:
        f.write(f"{key} = 1 16 16 8 2\n")
    with lib.Context(g) as e:
        e.set_model(velv)
        e.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP32)
        assert (e.stats()["tile_x"], e.stats()["tile_z"], e.stats()["tile_r"]) == (16, 16, 8)
        for s in range(faces.shape[0]):
            assert float(np.abs(e.field(s) - fields[s]).max()) <= 1e-3, s


def test_caller_choice_is_not_tuned(tmp_path, monkeypatch):
    """A configuration the caller sets is used as it stands: no timing, no
    cache entry; small batches keep the default untimed."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("FMM3D_TUNE_CACHE", str(cache))
    g, velv, faces = _tune_case()
    with lib.Context(g) as c:
        c.set_model(velv)
        c.set_option("column_x", 16)
        c.solve_tele(faces, lib.ORDER_MIXED, lib.PREC_FP64)
        assert c.stats()["tile_x"] == 16
    with lib.Context(g) as d:
        d.set_model(velv)
        d.solve_tele(faces[:4], lib.ORDER_MIXED, lib.PREC_FP64)
        assert (d.stats()["tile_x"], d.stats()["tile_z"], d.stats()["tile_r"]) == (8, 8, 16)
    assert not cache.exists()


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


@pytest.mark.parametrize("verify,narrow,filt", [(1, 0, 1), (1, 1, 1), (1, 1, 0), (2, -1, 1)])
def test_fixed_point_check_keeps_the_fields(verify, narrow, filt):
    """The read-only fixed-point check after sweep `verify` (sources with no
    node that would move leave the sweeps; with narrowing the next sweep
    relaxes only the moving patches; with the filter, after a first sweep,
    only nodes an input relaxed after them could move are evaluated) reaches
    the fields of the sweeps without it: float64 within 1e-9 s, float32
    within 1 ms; plane waves with side faces and a point source in one
    batch."""
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
        c.set_option("verify_filter", filt)
        for prec, tol in ((lib.PREC_FP64, 1e-9), (lib.PREC_FP32, 1e-3)):
            d.solve_mixed(F, (X, Z, S), pts, np.array([4]), 5, lib.ORDER_MIXED, prec)
            c.solve_mixed(F, (X, Z, S), pts, np.array([4]), 5, lib.ORDER_MIXED, prec)
            for s in range(6):
                assert float(np.abs(c.field(s) - d.field(s)).max()) <= tol, (verify, narrow, filt, prec, s)
        with pytest.raises(lib.FmmError):
            c.set_option("verify_patches", 2)
        with pytest.raises(lib.FmmError):
            c.set_option("verify_filter", 2)
