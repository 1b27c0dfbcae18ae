"""fmm3d_slice_velocity (the GPU form of gmtplot/gmtslicet.f90) against the
oracle's restatement, which equals the reference program's output bit for
bit: float64 values identical on the golden model and on the 201^3
workload's model; the drop-in gmtslice_gpu writes the reference's files
(values equal, bounds files byte for byte)."""
import importlib.util
import json
import os
import subprocess

import numpy as np
import pytest

import goldens as G
from fmm3d import build, lib, synth
from fmm3d.synth import GridDesc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIX = os.path.join(G.GOLDEN, "slice_31x33x35.npz")


def _fixture():
    z = np.load(FIX, allow_pickle=False)
    return z, GridDesc(**json.loads(str(z["grid"])))


def test_gpu_slices_equal_oracle_and_reference():
    z, g = _fixture()
    for (kind, _, d1, d2), key, pos in zip(z["slices"], ("depth", "ns", "ew"), z["pos"]):
        v, n = lib.slice_velocity(g, z["vtrue"], z["vref"], int(kind), float(pos), int(d1), int(d2))
        assert np.array_equal(v, z[key]), (key, float(np.abs(v - z[key]).max()))
    # the 201^3 workload's model, no reference model
    g2, velv, _, _ = synth.config_201(nsrc=1, nrc=10, seed=2)
    for kind, pos in ((0, -300.0), (1, g2.rgozd + 7.3), (2, g2.rgoxd - 5.9)):
        v, n = lib.slice_velocity(g2, velv, None, kind, pos, 4, 5)
        w, m = O.slice_velocity(g2, velv, None, kind, pos, 4, 5)
        assert n == m and np.array_equal(v, w), (kind, float(np.abs(v - w).max()))
    with pytest.raises(lib.FmmError) as e:
        lib.slice_velocity(g, z["vtrue"], None, 0, g.rgor + 50.0, 3, 3)
    assert e.value.code == lib.E_ARG


def test_dropin_slicer_writes_the_reference_files(tmp_path):
    if not os.path.exists(build.SLICE_EXE):
        pytest.skip("gmtslice_gpu not built")
    spec = importlib.util.spec_from_file_location("make_golden_slice", os.path.join(G.GOLDEN, "make_golden_slice.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    z, g = _fixture()
    synth.write_vtx(str(tmp_path / "gridc.vtx"), g, z["vtrue"])
    synth.write_vtx(str(tmp_path / "gridi.vtx"), g, z["vref"])
    mk.write_in(str(tmp_path), [float(p) for p in z["pos"]])
    r = subprocess.run([build.SLICE_EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for fn, key in (("grid2dvd.z", "depth"), ("grid2dvns.z", "ns"), ("grid2dvew.z", "ew")):
        with open(tmp_path / fn) as f:
            vals = np.array([float(x) for x in f.read().split()])
        assert np.array_equal(vals, z[key]), fn
    bounds = json.loads(str(z["bounds"]))
    for fn, text in bounds.items():
        assert (tmp_path / fn).read_text() == text, fn
