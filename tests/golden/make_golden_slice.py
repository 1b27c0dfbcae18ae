"""Write the model-slice golden fixture: a seeded synthetic model (the true
model, +-5 %, and a reference model, the same 1-D gradient unperturbed) run
through the reference's slicer compiled from its own source
(oracle/_ref/gmtslicet, `make -C oracle ref`) with depth, N-S and E-W
slices enabled.  The fixture holds the inputs, the slice parameters and
the reference's output files (values and the bounds files' text).

Run from the repository root:  python tests/golden/make_golden_slice.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "parallel-fast-marching-method_amd"))

from fmm3d import synth  # noqa: E402

REF_EXE = os.path.join(REPO, "oracle", "_ref", "gmtslicet")
OUT = os.path.join(HERE, "slice_31x33x35.npz")

# (kind, position, dicing 1, dicing 2): depth at height -140 km, N-S at a
# longitude 3.3 deg east of the grid's west edge, E-W at a latitude 4.1 deg
# south of its north edge
SLICES = [(0, -140.0, 6, 6), (1, 3.3, 6, 6), (2, -4.1, 6, 9)]

IN_TEMPLATE = """cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c Input file names
cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
gridc.vtx                   c: True model velocity vertices
gridi.vtx                   c: Reference model velocity vertices
ccccccccccccccccccccccccccccccccccccccccccccccc
c Slice parameters
ccccccccccccccccccccccccccccccccccccccccccccccc
1                           c: Extract depth slice? (0=no, 1=yes)
{dep!r}                      c: Depth of slice
bounddp.gmt                 c: GMT plotting bounds for depth slice
1                           c: Extract N-S slice? (0=no, 1=yes)
{lon!r}                      c: Longitude of slice
boundns.gmt                 c: GMT plotting bounds for N-S slice
1                           c: Extract E-W slice? (0=no, 1=yes)
{lat!r}                      c: Latitude of slice
boundew.gmt                 c: GMT plotting bounds for E-W slice
ccccccccccccccccccccccccccccccccccccccccccccccc
c Velocity grid parameters
ccccccccccccccccccccccccccccccccccccccccccccccc
1                           c: Plot depth slice? (0=no,1=yes)
{d0a}   {d0b}                       c: Dicing in theta,phi
grid2dvd.z                  c: GMT output file for depth slice
1                           c: Plot N-S slice? (0=no,1=yes)
{d1a}   {d1b}                       c: Dicing in r,theta
grid2dvns.z                 c: GMT output file for N-S slice
1                           c: Plot E-W slice? (0=no,1=yes)
{d2a}   {d2b}                       c: Dicing in r,phi
grid2dvew.z                 c: GMT output file for E-W slice
"""


def case():
    """The fixture's grid, true and reference vertex models, and the three
    slice positions in gmtslicet's units (height km, lon deg, lat deg)."""
    g = synth.make_grid((31, 33, 35), 400.0, (10.0, 10.0), nv=(10, 11, 12))
    vtrue = synth.make_velv(g, 0.05, 11)
    vref = synth.make_velv(g, 0.0, 11)
    pos = [SLICES[0][1], g.rgozd + SLICES[1][1], g.rgoxd + SLICES[2][1]]
    return g, vtrue, vref, pos


def write_in(d, pos):
    with open(os.path.join(d, "gmtslicet.in"), "w") as f:
        f.write(IN_TEMPLATE.format(dep=pos[0], lon=pos[1], lat=pos[2], d0a=SLICES[0][2], d0b=SLICES[0][3],
                                   d1a=SLICES[1][2], d1b=SLICES[1][3], d2a=SLICES[2][2], d2b=SLICES[2][3]))


def main():
    if not os.path.exists(REF_EXE):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    g, vtrue, vref, pos = case()
    d = tempfile.mkdtemp(prefix="fmm3d_slice_")
    try:
        synth.write_vtx(os.path.join(d, "gridc.vtx"), g, vtrue)
        synth.write_vtx(os.path.join(d, "gridi.vtx"), g, vref)
        write_in(d, pos)
        r = subprocess.run([REF_EXE], cwd=d, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"gmtslicet failed\n{r.stdout}\n{r.stderr}")
        vals = {}
        for fn in ("grid2dvd.z", "grid2dvns.z", "grid2dvew.z"):
            with open(os.path.join(d, fn)) as f:
                vals[fn] = np.array([float(x) for x in f.read().split()])
        bounds = {}
        for fn in ("bounddp.gmt", "boundns.gmt", "boundew.gmt"):
            with open(os.path.join(d, fn)) as f:
                bounds[fn] = f.read()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    np.savez_compressed(
        OUT,
        grid=np.array(json.dumps(dataclasses.asdict(g))),
        vtrue=vtrue, vref=vref, pos=np.array(pos), slices=np.array(SLICES),
        depth=vals["grid2dvd.z"], ns=vals["grid2dvns.z"], ew=vals["grid2dvew.z"],
        bounds=np.array(json.dumps(bounds)),
    )
    print(f"slice golden: {[len(v) for v in vals.values()]} values -> {OUT}")


if __name__ == "__main__":
    main()
