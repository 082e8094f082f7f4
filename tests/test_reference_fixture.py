"""The engine and the oracle against the reference's own long-average fields.

run/hyStrath/dsmcFoam+/couette_N2-O2/backup-5 holds a full particle snapshot
of the tutorial at t = 5 s and the fields that dsmcFoam+ averaged from
t = 0.5 s to 5 s (450,000 steps): number density, velocity and the
translational, rotational, vibrational and overall temperatures of N2, O2
and the mixture.  tests/golden/make_couette_fixture.py parsed them into
couette_backup5.npz.  Starting from the snapshot (already steady), the
engine must sample fields whose batch-means 95% confidence band holds the
reference's values at the rate a 95% band should.

Reference behaviour: the whole of dsmcCloud::evolve for LB-VHS N2/O2 between
diffuse walls (dsmcCloud.C:819-926), and dsmcVolFields.C:1242-1790 for the
fields.  The reference's long averages are themselves a single realisation;
with 450,000 steps their noise is well below the band's.
"""
import os

import numpy as np
import pytest
from scipy import stats

import oracle_bind as ob
from hystrath_amd import cases, fields

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GROUPS = (("N2", [0]), ("O2", [1]), ("mixture", None))
SCALARS = ("rhoN", "Ttra", "Trot", "Tvib", "Tov")


def snapshot():
    pos = np.load(os.path.join(GOLDEN, "couette_positions.npz"))
    b = np.load(os.path.join(GOLDEN, "couette_backup5.npz"))
    xyz = pos["xyz"]
    d = dict(px=xyz[:, 0].copy(), py=xyz[:, 1].copy(), pz=xyz[:, 2].copy(), cell=pos["cell"].astype(np.int32),
             ux=b["U"][:, 0].copy(), uy=b["U"][:, 1].copy(), uz=b["U"][:, 2].copy(), erot=b["erot"].copy(),
             type=b["type"].copy(), vib=b["vib"].reshape(1, -1).copy())
    return d, b


def field_set(samples, nsteps, case):
    """{(group, quantity): per-cell array} for the compared quantities; U is
    its x component (the shear direction)."""
    out = {}
    for g, ids in GROUPS:
        f = fields.vol_fields(samples, nsteps, case, type_ids=ids)
        for q in SCALARS:
            out[(g, q)] = f[q]
        out[(g, "Ux")] = f["U"][0]
    return out


def reference_set(b):
    out = {}
    for g, _ in GROUPS:
        for q in SCALARS:
            out[(g, q)] = b[f"{q}_{g}"]
        out[(g, "Ux")] = b[f"U_{g}"][:, 0]
    return out


def test_oracle_rows_match_the_reference_long_averages():
    """The oracle from the snapshot, 200 steps to settle and 800 sampled.
    Averaged over the five cells of each row (x is periodic), density and the
    translational, rotational and overall temperatures lie within 2.5% of the
    reference in every row for the mixture (3.5% for one species) and within
    0.5% on average over the rows; the shear velocity within 8 m/s on average
    (range 16-282 m/s).  The margins are those of 800 correlated steps of 95
    parcels a cell; a wrong wall temperature or wall velocity, or rotation
    left out of the exchange, misses them by far.  The vibrational
    temperature is left out here: with Zv of order 1e5 at these temperatures
    it moves over far more steps than this test takes (the GPU test below
    covers it)."""
    d, b = snapshot()
    case = cases.couette_n2o2(capacity=80000)
    st = ob.OracleState(case)
    st.load(d, sigma=b["sigma_tcr_max"], rem=np.zeros(case.ncells))
    st.evolve(200)
    st.samples[:] = 0
    st.s.n_sampled = 0
    n = 800
    st.evolve(n)
    got = field_set(st.sample_array(), n, case)
    ref = reference_set(b)
    for key, r in ref.items():
        g = got[key].reshape(100, 5).mean(1)
        rr = r.reshape(100, 5).mean(1)
        if key[1] == "Ux":
            assert np.abs(g - rr).mean() < 8.0, key
        elif key[1] in ("rhoN", "Ttra", "Trot", "Tov"):
            rel = g / rr - 1
            assert np.abs(rel).max() < (0.025 if key[0] == "mixture" else 0.035), (key, np.abs(rel).max())
            assert abs(rel.mean()) < 0.005, (key, rel.mean())


@pytest.mark.gpu
def test_engine_fields_hold_the_reference_in_their_95_percent_band():
    """20 batches of 5,000 steps after 2,000 steps to settle.  A batch is
    longer than the slowest fluctuation of this flow (heat and momentum
    diffuse across the 1 m gap in about 3,000 steps: L^2 / nu with nu about
    37 m^2/s), so batch means are close to independent.  Per cell and
    quantity the batch means give a Student-t 95% band; over the 500 cells of
    each (group, quantity) the reference's value lies outside it in at most
    9% of the cells (5% expected plus four binomial standard deviations) and
    no cell is off by more than 6.5 standard errors.  The vibrational
    temperature relaxes over many more steps (Zv of order 1e5): its standard
    errors are widened by the lag-one autocorrelation of the batch means.  A
    bias shared by all cells is tested on the domain average, whose batch
    means carry the whole-gap fluctuations: the reference's domain average
    lies in its 99.9% band (density, fixed by the parcel count, within
    1e-4)."""
    from hystrath_amd.engine import Engine

    d, b = snapshot()
    case = cases.couette_n2o2(capacity=80000)
    ref = reference_set(b)
    nb, m = 20, 5000
    batches = {k: [] for k in ref}
    with Engine(case) as eng:
        eng.upload(d)
        eng.set_cell_state(b["sigma_tcr_max"], np.zeros(case.ncells))
        eng.evolve(2000)
        for _ in range(nb):
            eng.reset_samples()
            eng.evolve(m)
            s, n = eng.samples()
            assert n == m
            for k, v in field_set(s, n, case).items():
                batches[k].append(v)
        c = eng.counters()
        assert c["track_overflow"] == 0 and c["tile_overflow"] == 0 and c["capacity_overflow"] == 0
    t95 = stats.t.ppf(0.975, nb - 1)
    t999 = stats.t.ppf(0.9995, nb - 1)
    report, failures = [], []
    for k, r in ref.items():
        x = np.array(batches[k])
        mean = x.mean(0)
        dev = x - mean
        # successive batch means of the vibrational temperature stay correlated
        # (Zv of order 1e5): the lag-one autocorrelation, pooled over the cells,
        # widens the band by the AR(1) factor (1 + rho1) / (1 - rho1)
        rho1 = float(np.clip((dev[1:] * dev[:-1]).sum() / max((dev * dev).sum(), 1e-300), 0.0, 0.9))
        sd = x.std(0, ddof=1) * np.sqrt((1 + rho1) / (1 - rho1))
        # the reference's value is itself an average, over 450,000 steps: its
        # standard error, that of one batch scaled to its length, adds to ours
        se = sd * np.sqrt(1.0 / nb + m / 450000.0)
        z = (r - mean) / se
        out = np.mean(np.abs(z) > t95)
        g = x.mean(1)
        gd = g - g.mean()
        g1 = float(np.clip((gd[1:] * gd[:-1]).sum() / max((gd * gd).sum(), 1e-300), 0.0, 0.9))
        gse = g.std(ddof=1) / np.sqrt(nb) * np.sqrt((1 + g1) / (1 - g1))
        rel = g.mean() / r.mean() - 1
        # the domain-average density of this closed gap is fixed by the parcel
        # count: it has no spread to test against, only the value
        domain_ok = abs(rel) < 1e-4 if gse < 1e-9 * abs(g.mean()) else abs(r.mean() - g.mean()) <= t999 * gse
        report.append(f"{k}: rho1 {rho1:.2f}, outside {out:.3f}, max |z| {np.abs(z).max():.2f}, "
                      f"domain rel {rel:+.5f}, domain ok {domain_ok}")
        if out > 0.05 + 4 * np.sqrt(0.05 * 0.95 / len(r)) or np.abs(z).max() > 6.5 or not domain_ok:
            failures.append(k)
    print("\n".join(report))
    assert not failures, failures
