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

# The first HEAD parcels' velocity, rotational energy, species and vibrational
# level (U, ERot, typeId, vibLevel of the same snapshot), parsed here with
# plain regular expressions, independently of hystrath_amd.lagrangian.  U and
# ERot are kept in single precision, as the engine holds them.
HEAD = 8192
d = os.path.dirname(path)


def entries(name):
    txt = open(os.path.join(d, name)).read()
    body = txt[txt.index("\n(", txt.index("FoamFile")) + 2:]
    return body.split("\n")


U = []
for line in entries("U"):
    m = re.match(r"^\(([-0-9.eE+]+) ([-0-9.eE+]+) ([-0-9.eE+]+)\)\s*$", line)
    if m:
        U.append([float(m.group(k)) for k in (1, 2, 3)])
    if len(U) == HEAD:
        break
erot = []
for line in entries("ERot"):
    if re.match(r"^[-0-9.eE+]+\s*$", line):
        erot.append(float(line))
    if len(erot) == HEAD:
        break
tid = []
for line in entries("typeId"):
    if re.match(r"^\d+\s*$", line):
        tid.append(int(line))
    if len(tid) == HEAD:
        break
vib = []
for line in entries("vibLevel"):
    m = re.match(r"^1\((\d+)\)\s*$", line)
    if m:
        vib.append(int(m.group(1)))
    if len(vib) == HEAD:
        break
out2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "couette_cloud_head.npz")
np.savez_compressed(out2, xyz=xyz[:HEAD], cell=cells[:HEAD].astype(np.int32), U=np.array(U, np.float64).astype(np.float32),
                    erot=np.array(erot, np.float64).astype(np.float32), type=np.array(tid, np.uint8), vib=np.array(vib, np.uint16))
print(f"{HEAD} parcels with U, ERot, typeId, vibLevel -> {out2}")

# The whole snapshot's U, ERot, typeId and vibLevel, the cell state
# dsmcSigmaTcRMax and the long-average fields the tutorial wrote at t = 5 s
# (sampled from t = 0.5 s: system/fieldPropertiesDict resetAtOutputUntilTime
# 0.5, 450,000 steps of 1e-5 s): rhoN, U, Ttra, Trot, Tvib, Tov and p of the
# N2, O2 and mixture field models, internal cells only.
full = {}
for name, pat_, conv in (("U", r"^\(([-0-9.eE+]+) ([-0-9.eE+]+) ([-0-9.eE+]+)\)\s*$", None),
                         ("ERot", r"^([-0-9.eE+]+)\s*$", None), ("typeId", r"^(\d+)\s*$", None),
                         ("vibLevel", r"^1\((\d+)\)\s*$", None)):
    vals = []
    for line in entries(name):
        m = re.match(pat_, line)
        if m:
            vals.append([float(g) for g in m.groups()])
    full[name] = np.array(vals)
assert all(len(v) == len(cells) for v in full.values())


def internal_field(path_):
    txt = open(path_).read()
    i = txt.index("internalField")
    j = txt.index("(", i)
    n = int(txt[txt.index("\n", i):j].strip())
    body = txt[j + 1:txt.index("\n)", j)]
    if "List<vector>" in txt[i:j]:
        v = np.array([[float(x) for x in m] for m in re.findall(r"\(([-0-9.eE+]+) ([-0-9.eE+]+) ([-0-9.eE+]+)\)", body)])
    else:
        v = np.array([float(x) for x in body.split()])
    assert len(v) == n, (path_, len(v), n)
    return v


b5 = os.path.dirname(os.path.dirname(d))
fld = {}
for group in ("N2", "O2", "mixture"):
    for q in ("rhoN", "U", "Ttra", "Trot", "Tvib", "Tov", "p"):
        fld[f"{q}_{group}"] = internal_field(os.path.join(b5, f"{q}_{group}"))
sig = internal_field(os.path.join(b5, "dsmcSigmaTcRMax"))
out3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "couette_backup5.npz")
np.savez_compressed(out3, U=full["U"].astype(np.float32), erot=full["ERot"][:, 0].astype(np.float32),
                    type=full["typeId"][:, 0].astype(np.uint8), vib=full["vibLevel"][:, 0].astype(np.uint16),
                    sigma_tcr_max=sig, **fld)
print(f"{len(cells)} parcels, dsmcSigmaTcRMax and {len(fld)} long-average fields -> {out3}")
