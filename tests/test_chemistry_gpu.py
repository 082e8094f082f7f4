"""QK chemistry on the GPU against the oracle, bit for bit.

The reacting collide (one thread per cell, the reaction branch of
noTimeCounter.C:250-300 with dissociationQK, exchangeQK and
dissociationExchangeQK) must leave every parcel, the dissociation products
(placed after the sorted population in cell order, candidate order within
the cell, dsmcCloud::addNewParcel at dsmcCloud.C:396-432), sigmaTcRMax, the
remainders, the cell sums (which include the step's products,
dsmcCloud.C:904-907) and the counters equal to the oracle's
(tests/test_chemistry.py pins the oracle to the reference's heatBath-5species
run).  Cases: the tutorial's one-cell heat bath at fewer parcels, the same
box cut into 27 cells, and reacting air past the cylinder with freestream
inflow (whose next-step inflow is appended after the products), each
through eager steps and graph-replayed pairs; and the stage entry points.
"""
import numpy as np
import pytest

import helpers as H
import oracle_bind as ob
from hystrath_amd import abi, cases
from hystrath_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _fill_bath(obj, c):
    obj.fill(c.init_n, c.init_T)


def _fill_air(obj, c):
    # a hot start: the domain at 30,000 K, so reactions begin at once
    obj.fill(c.n_in, 30000.0, u=(c.U_inf, 0.0, 0.0))


def _pair(c, fill):
    eng = Engine(c)
    st = ob.OracleState(c)
    fill(eng, c)
    fill(st, c)
    return eng, st


def _check(eng, st, label):
    H.assert_same_particles(eng.download(), st.particles(), label)
    sig, rem = eng.cell_state()
    np.testing.assert_array_equal(sig, st.sigma, err_msg=f"{label}: sigmaTcRMax")
    np.testing.assert_array_equal(rem, st.rem, err_msg=f"{label}: remainder")


def _check_end(eng, st, steps, label):
    s, n = eng.samples()
    np.testing.assert_array_equal(s, st.sample_array(), err_msg=f"{label}: cell sums")
    assert n == steps
    c, oc = eng.counters(), st.counters()
    assert (c["candidates"], c["collisions"], c["deleted"], c["injected"], c["wall_hits"]) == tuple(oc[:5]), label
    assert c["capacity_overflow"] == 0 and c["tile_overflow"] == 0


@pytest.mark.parametrize("which", ["bath1", "bath27", "air"])
def test_reacting_steps_match_the_oracle(which):
    hf = abi.HS_FLAG_SAMPLE | abi.HS_FLAG_HEATFLUX | abi.HS_FLAG_COLL_SEP
    c, fill, calls = {
        "bath1": (cases.heat_bath_5species(fn=2000.0, flags=hf), _fill_bath, (1, 1, 1, 4, 3)),
        "bath27": (cases.heat_bath_5species(fn=400.0, n_side=3), _fill_bath, (1, 2, 5)),
        "air": (cases.air_cylinder(nx=24, ny=24, ppc=12, seed=3, flags=hf), _fill_air, (1, 1, 3, 4, 2)),
    }[which]
    eng, st = _pair(c, fill)
    n0 = st.n
    total = 0
    print(f"[chem] {which}: {n0} parcels, calls {calls}", flush=True)
    for r, n in enumerate(calls):
        eng.evolve(n)
        st.evolve(n)
        total += n
        _check(eng, st, f"{which}/call{r}")
    _check_end(eng, st, total, which)
    ty = st.particles()["type"]
    # dissociations happened (N or O parcels), and in the air case the inflow
    if which == "air":
        assert st.counters()[3] > 0
        np.testing.assert_array_equal(eng.inflow()[: st.inflow.size], st.inflow)
    else:
        assert st.n > n0
    assert np.count_nonzero(ty >= 3) > 10
    eng.close()


@pytest.mark.parametrize("fn", [700.0, 210.0])
def test_one_large_reacting_cell_matches_the_oracle(fn):
    """The tutorial's single cell at about 22,000 parcels (Fn 700) and 73,000
    (Fn 210): more parcels than the LDS tiles and the per-parcel conflict
    stamps hold, so the wave-cooperative NTC shares stamps between parcels,
    and at 73,000 builds 32-bit sub-cell lists; the cell's sums with its
    products are taken by a wave."""
    c = cases.heat_bath_5species(fn=fn)
    eng, st = _pair(c, _fill_bath)
    print(f"[chem] one cell, {st.n} parcels", flush=True)
    for r, n in enumerate((1, 2)):
        eng.evolve(n)
        st.evolve(n)
        _check(eng, st, f"fn{fn}/call{r}")
    _check_end(eng, st, 3, f"fn{fn}")
    eng.close()


def test_stage_entry_points_with_reactions():
    """hs_dsmc_sort + hs_dsmc_collide (buildCellOccupancy, then
    noTimeCounter::collide with its reactions) equal the oracle's sort and
    collide, products included; the step counter advances separately."""
    c = cases.heat_bath_5species(fn=1000.0, n_side=2)
    eng, st = _pair(c, _fill_bath)
    for k in range(3):
        eng.collide(sample=True)
        st.sort()
        st.collide(sample=True)
        _check(eng, st, f"stage{k}")
        eng.step = eng.step + 1
        st.step = st.step + 1
    np.testing.assert_array_equal(eng.samples()[0], st.sample_array())
    assert st.n > int(sum(c.init_n) * 1e-15 / c.cfg.fn)
    eng.close()


def test_heat_bath_tutorial_at_full_size_follows_the_reference_run():
    """The heatBath-5species tutorial as the reference ran it (one cell,
    nEquivalentParticles 100, about 154,000 parcels) for 200 steps on the
    engine: species number densities within the counting noise of both runs
    (4 sigma plus 1%) of the reference's probe series
    (tests/golden/heatbath_probes.npz), and the mixture's translational and
    rotational temperatures, from this step's cell sums, within 2-3%."""
    import os
    from hystrath_amd import fields
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "heatbath_probes.npz"))
    c = cases.heat_bath_5species(fn=100.0)
    eng = Engine(c)
    eng.fill(c.init_n, c.init_T)
    V = 1e-15
    done = 0
    for k in (50, 100, 150, 200):
        eng.evolve(k - 1 - done)
        before = eng.samples()[0].copy()
        eng.evolve(1)
        done = k
        f = fields.vol_fields(eng.samples()[0] - before, 1, c)
        row = int(np.searchsorted(fx["step"], k))
        counts = np.bincount(eng.download()["type"], minlength=5)
        for i, nm in enumerate(["N2", "O2", "NO", "N", "O"]):
            ref = fx["rhoN_" + nm][row]
            n_ref = ref * V / 100.0
            if n_ref < 50:
                continue
            tol = 4.0 * np.sqrt(1.0 / max(counts[i], 1) + 1.0 / n_ref) + 0.01
            assert abs(counts[i] * c.cfg.fn / V / ref - 1.0) < tol, (k, nm)
        for key, rel in (("Ttra", 0.02), ("Trot", 0.03)):
            assert abs(float(f[key][0]) / fx[key + "_mixture"][row] - 1.0) < rel, (k, key)
        print(f"[chem] heat bath step {k}: {eng.count()} parcels", flush=True)
    assert eng.counters()["capacity_overflow"] == 0
    eng.close()


def test_reacting_engines_stepped_together_match_their_oracles():
    """hs_dsmc_evolve_batch with two reacting cases (the heat bath in one cell
    and cut into 8, other seeds): each ends equal to its own oracle."""
    from hystrath_amd.engine import evolve_batch
    cs = [cases.heat_bath_5species(fn=3000.0, seed=5), cases.heat_bath_5species(fn=1500.0, n_side=2, seed=6)]
    pairs = [_pair(c, _fill_bath) for c in cs]
    engs = [e for e, _ in pairs]
    for n in (3, 4):
        evolve_batch(engs, n)
        for k, (e, st) in enumerate(pairs):
            st.evolve(n)
            _check(e, st, f"batch/case{k}/+{n}")
    for k, (e, st) in enumerate(pairs):
        _check_end(e, st, 7, f"batch/case{k}")
        e.close()


def test_reacting_steps_with_zv_2008_match_the_oracle():
    """The heat bath cut into 27 cells with inverseZvFormulation 2008: cell
    temperatures (some zero) set on the engine and the oracle, the pairs that
    do not react relax their vibration with them; bit-identical steps."""
    c = cases.heat_bath_5species(fn=400.0, n_side=3, seed=9)
    c.cfg.inv_zv_form = 1
    eng, st = _pair(c, _fill_bath)
    t = np.linspace(2000.0, 20000.0, c.ncells)
    t[::4] = 0.0
    eng.set_cell_temperature(t)
    st.set_cell_temperature(t)
    for r, n in enumerate((1, 2, 3)):
        eng.evolve(n)
        st.evolve(n)
        _check(eng, st, f"zv2008/call{r}")
    _check_end(eng, st, 6, "zv2008")
    eng.close()
