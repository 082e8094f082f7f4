"""Writes heatbath_probes.npz from the hyStrath heatBath-5species tutorial's
probe time series (run/hyStrath/dsmcFoam+/heatBath-5species/gnuplot/
solution/{rhoN_N2,rhoN_O2,rhoN_NO,rhoN_N,rhoN_O,Ttra_mixture,Trot_mixture,
Tvib_mixture,Tov_mixture}: one value per time step of 1e-9 s, instantaneous
dsmcVolFields of the single cell, "time value" per line after '#' headers).

The reference's run: nEquivalentParticles 100 (about 154,000 parcels),
10,000 steps.  Kept: every step of the first 1,000, then every tenth.
The files are read as text; nothing from the reference is imported or run.
Usage: python tests/golden/make_heatbath_fixture.py [/root/reference]
"""
import os
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
base = os.path.join(ref, "run/hyStrath/dsmcFoam+/heatBath-5species/gnuplot/solution")
names = ["rhoN_N2", "rhoN_O2", "rhoN_NO", "rhoN_N", "rhoN_O",
         "Ttra_mixture", "Trot_mixture", "Tvib_mixture", "Tov_mixture"]
cols = {}
time = None
for nm in names:
    t, v = [], []
    for line in open(os.path.join(base, nm)):
        s = line.split()
        if not s or s[0].startswith("#"):
            continue
        t.append(float(s[0]))
        v.append(float(s[1]))
    t = np.array(t)
    if time is None:
        time = t
    assert np.array_equal(time, t), nm
    cols[nm] = np.array(v)
step = np.rint(time / 1e-9).astype(np.int64)
keep = (step <= 1000) | (step % 10 == 0)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "heatbath_probes.npz")
np.savez_compressed(out, step=step[keep], **{k: v[keep] for k, v in cols.items()})
print(f"{keep.sum()} of {len(step)} steps -> {out}")
