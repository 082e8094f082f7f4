"""Writes couette_positions.npz from the hyStrath couette_N2-O2 tutorial's
particle snapshot (run/hyStrath/dsmcFoam+/couette_N2-O2/backup-5/lagrangian/
dsmc/positions: 47,583 lines "(x y z) cell", written by OpenFOAM's
Cloud<passiveParticle> on the 5x100x1 block mesh of constant/polyMesh).

The file is read as text; nothing from the reference is imported or run.
Usage: python tests/golden/make_couette_fixture.py [/root/reference]
"""
import os
import re
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
path = os.path.join(ref, "run/hyStrath/dsmcFoam+/couette_N2-O2/backup-5/lagrangian/dsmc/positions")
pat = re.compile(r"^\(([-0-9.eE+]+) ([-0-9.eE+]+) ([-0-9.eE+]+)\) (\d+)\s*$")
xyz, cells = [], []
for line in open(path):
    m = pat.match(line)
    if m:
        xyz.append([float(m.group(1)), float(m.group(2)), float(m.group(3))])
        cells.append(int(m.group(4)))
xyz = np.array(xyz, np.float64)
cells = np.array(cells, np.int16)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "couette_positions.npz")
np.savez_compressed(out, xyz=xyz, cell=cells)
print(f"{len(cells)} parcels -> {out}")
