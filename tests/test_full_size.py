"""The workloads at their stated sizes on the GPU.

- Mach-10 cylinder, 400x400 cells, about 4M parcels: after the bow shock has
  started to form, three steps against the oracle, bit-exact.  At this size
  the cell scan's look-back runs over many more than 64 predecessor blocks,
  tiles are the full 1,536 / 2,048 parcels and the stagnation cells are
  heavy (more than 96 parcels), so their tiles are ordered first.
- Periodic argon box, 20^3 cells, 20 parcels a cell (160k), bimodal 300 K /
  1000 K start, 2,000 steps: energy conserved, T = 650 K, and the sampled
  collision rate equals the VHS equilibrium frequency (Bird 4.74) within 1%.
- The 64-case batch of uneven argon and N2 cylinder cases (10k to 2M parcels)
  through hs_dsmc_evolve_batch: no device-side bound trips, and three of its
  engines equal their oracles.

Reference behaviour: dsmcCloud::evolve (dsmcCloud.C:819-926) and the parts it
calls, as in the smaller parity tests.
"""
import math

import numpy as np
import pytest

import helpers as H
import oracle_bind as ob
from hystrath_amd import abi, cases
from hystrath_amd.engine import Engine, evolve_batch

pytestmark = pytest.mark.gpu


def _compare(eng, st, label, samples=True):
    H.assert_same_particles(eng.download(), st.particles(), label)
    sig, rem = eng.cell_state()
    np.testing.assert_array_equal(sig, st.sigma, err_msg=f"{label}: sigmaTcRMax")
    np.testing.assert_array_equal(rem, st.rem, err_msg=f"{label}: remainder")
    if samples:
        np.testing.assert_array_equal(eng.samples()[0], st.sample_array(), err_msg=f"{label}: samples")


def test_cylinder_400x400_steps_match_the_oracle():
    case = cases.cylinder_flow(nx=400, ny=400, ppc=25, dt=2.5e-7, seed=424242,
                               flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL)
    with Engine(case) as eng:
        H.fill_cylinder(eng, case)
        eng.evolve(1200)
        c = eng.counters()
        assert c["track_overflow"] == 0 and c["tile_overflow"] == 0 and c["capacity_overflow"] == 0
        d = eng.download()
        n = len(d["px"])
        assert n > 3_800_000
        counts = np.bincount(d["cell"], minlength=case.ncells)
        assert (counts > 96).sum() > 50, "no heavy cells: the stagnation region has not formed"
        assert case.ncells > 64 * 2048  # more than 64 look-back predecessors in the cell scan
        gt, ct = eng.tiles()
        assert (gt, ct) == (1536, 2048)
        st = ob.OracleState(case)
        sig, rem = eng.cell_state()
        st.load(d, sigma=sig, rem=rem)
        st.inflow[:] = eng.inflow()[: st.inflow.size]
        st.step = eng.step
        eng.reset_samples()
        for k in range(3):
            eng.evolve(1)
            st.evolve(1)
            _compare(eng, st, f"cylinder400/step{k}", samples=False)
        got, want = eng.samples()[0], st.sample_array()
        bad = np.argwhere(got != want)
        if len(bad):
            cnt = np.bincount(st.particles()["cell"], minlength=case.ncells)
            rows = [(int(k), int(c), float(got[k, c]), float(want[k, c]), int(cnt[c])) for k, c in bad[:12]]
            raise AssertionError(f"{len(bad)} sample sums differ (slot, cell, engine, oracle, parcels now): {rows}")
        np.testing.assert_array_equal(eng.inflow()[: st.inflow.size], st.inflow)


def test_box_20cubed_relaxes_to_650K_at_the_bird_collision_rate():
    case = cases.periodic_box(n_side=20, ppc=20, seed=20251203)
    m = case.species[0].mass
    with Engine(case) as eng:
        H.fill_box(eng, case)
        p0 = eng.download()
        assert len(p0["px"]) == 160_000

        def temp(p):
            u = np.stack([p["ux"], p["uy"], p["uz"]], 1).astype(np.float64)
            return m * ((u - u.mean(0)) ** 2).sum() / (3 * abi.HS_KB * len(u))

        T0 = temp(p0)
        assert abs(T0 - 650.0) < 0.01 * 650
        eng.evolve(1000)
        eng.reset_samples()
        eng.evolve(1000)
        p1 = eng.download()
        T1 = temp(p1)
        assert abs(T1 / T0 - 1) < 1e-5  # energy conserved up to FP32 velocity storage
        # the two populations have mixed: one temperature, a Maxwellian
        u = np.stack([p1["ux"], p1["uy"], p1["uz"]], 1).astype(np.float64)
        kurt = ((u - u.mean(0)) ** 4).mean(0) / ((u - u.mean(0)) ** 2).mean(0) ** 2
        assert np.all(np.abs(kurt - 3.0) < 0.05), kurt
        s, n = eng.samples()
        tot = s.sum(axis=1)
        rate = 2 * tot[abi.HS_S_NCOLL] / (tot[abi.species_slot(0, abi.HS_SS_N)] * case.cfg.dt)
        d, w = case.species[0].diameter, case.species[0].omega
        nu = 4 * d * d * case.n_density * math.sqrt(math.pi * abi.HS_KB * 273.0 / m) * (T1 / 273.0) ** (1 - w)
        assert abs(rate / nu - 1) < 0.01, rate / nu
        c = eng.counters()
        assert c["track_overflow"] == 0 and c["tile_overflow"] == 0 and c["capacity_overflow"] == 0


def test_batch_of_64_cases_runs_clean_and_matches_its_oracles():
    cs = cases.cylinder_batch(64)
    engs = []
    try:
        for c in cs:
            e = Engine(c)
            H.fill_cylinder(e, c)
            engs.append(e)
        sizes = [e.count() for e in engs]
        assert sum(sizes) > 25_000_000
        order = np.argsort(sizes)
        picks = [int(order[0]), int(order[1]), int(order[len(order) // 2])]
        oracles = {}
        for k in picks:
            st = ob.OracleState(cs[k])
            H.fill_cylinder(st, cs[k])
            oracles[k] = st
        evolve_batch(engs, 6)
        for k in picks:
            oracles[k].evolve(6)
            _compare(engs[k], oracles[k], f"batch64/case{k}")
        evolve_batch(engs, 40)
        for e in engs:
            c = e.counters()
            assert c["track_overflow"] == 0 and c["tile_overflow"] == 0 and c["capacity_overflow"] == 0
            assert e.step == 46
    finally:
        for e in engs:
            e.close()
