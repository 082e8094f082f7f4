"""GPU parity: every stage of the HIP engine against the CPU oracle on the same
seeded inputs.  Integer/index state (cells, counts, order) and FP64 positions
must be bit-exact; the engine and the oracle evaluate the same IEEE operation
sequence, so velocities are required bit-exact as well (stricter than the
1e-6 relative bar).  Sizes are kept small enough for the oracle to finish in
seconds."""
import numpy as np
import pytest

from hystrath_amd import abi, cases
from hystrath_amd.engine import Engine, EngineError, evolve_batch
import helpers as H
import oracle_bind as ob

pytestmark = pytest.mark.gpu


def _pair(case, filler):
    eng = Engine(case)
    st = ob.OracleState(case)
    filler(eng, case)
    filler(st, case)
    return eng, st


def _check_state(eng, st, label):
    H.assert_same_particles(eng.download(), st.particles(), label)
    sig, rem = eng.cell_state()
    np.testing.assert_array_equal(sig, st.sigma, err_msg=f"{label}: sigmaTcRMax")
    np.testing.assert_array_equal(rem, st.rem, err_msg=f"{label}: collisionSelectionRemainder")


@pytest.mark.parametrize("which", ["box", "cylinder", "couette"])
def test_fill_parity(which):
    case, filler = {
        "box": (H.small_box(), H.fill_box),
        "cylinder": (H.small_cylinder(), H.fill_cylinder),
        "couette": (cases.couette_n2o2(), H.fill_couette),
    }[which]
    eng, st = _pair(case, filler)
    assert st.n > 0
    _check_state(eng, st, f"fill/{which}")
    eng.close()


@pytest.mark.parametrize("model", [abi.HS_COLL_VHS, abi.HS_COLL_VSS])
def test_sigma_and_binary_collide_argon(model):
    case = H.small_box(collision_model=model)
    case.species[0].alpha = 1.4
    case.cfg.species[0].alpha = 1.4
    eng = Engine(case)
    rng = np.random.default_rng(3)
    n = 4096
    tp = np.zeros(n, np.int32)
    up = rng.normal(0, 400, (n, 3))
    uq = rng.normal(0, 400, (n, 3))
    uq[:8] = up[:8]  # zero relative speed
    s_gpu = eng.sigma_tcr(tp, tp, up, uq)
    s_cpu = ob.sigma_tcr(case, tp, tp, up, uq)
    np.testing.assert_array_equal(s_gpu, s_cpu)
    uni = rng.uniform(0, 1, (n, 8))
    g = eng.binary_collide(tp, tp, up, uq, uni)
    c = ob.binary_collide(case, tp, tp, up, uq, uni)
    assert g[7] == 0 and c[7] == 0
    np.testing.assert_array_equal(g[0], c[0])
    np.testing.assert_array_equal(g[1], c[1])
    np.testing.assert_array_equal(g[6], c[6])
    m = case.species[0].mass
    p0 = m * (up + uq)
    p1 = m * (g[0] + g[1])
    scale = m * (np.abs(up) + np.abs(uq)).sum(axis=1, keepdims=True)
    assert np.max(np.abs(p1 - p0) / scale) < 1e-14
    e0 = 0.5 * m * ((up ** 2).sum(1) + (uq ** 2).sum(1))
    e1 = 0.5 * m * ((g[0] ** 2).sum(1) + (g[1] ** 2).sum(1))
    assert np.max(np.abs(e1 - e0) / e0) < 1e-14
    eng.close()


@pytest.mark.parametrize("model,speed", [(abi.HS_COLL_LB_VHS, 2500.0), (abi.HS_COLL_LB_VSS, 2500.0),
                                         (abi.HS_COLL_LB_VHS, 15000.0)])
def test_binary_collide_lb(model, speed):
    """Larsen-Borgnakke pairs against the oracle.  At 15 km/s per component
    most pairs reach vibrational levels of 256 and more, past the engine's
    table of 1/Zv, where it is computed in the collision."""
    case = cases.couette_n2o2()
    case.cfg.collision_model = model
    eng = Engine(case)
    rng = np.random.default_rng(4)
    n = 4096
    tp = rng.integers(0, 2, n).astype(np.int32)
    tq = rng.integers(0, 2, n).astype(np.int32)
    up = rng.normal(0, speed, (n, 3))
    uq = rng.normal(0, speed, (n, 3))
    ep = rng.exponential(5e-20, n)
    eq = rng.exponential(5e-20, n)
    vp = np.zeros((n, 3), np.int32)
    vq = np.zeros((n, 3), np.int32)
    vp[:, 0] = rng.integers(0, 4, n)
    vq[:, 0] = rng.integers(0, 4, n)
    uni = rng.uniform(0, 1, (n, 256))
    if speed > 10000.0:
        th = np.array([case.species[0].theta_v[0], case.species[1].theta_v[0]])
        mr = np.array([s.mass for s in case.species[:2]])
        mr = mr[tp] * mr[tq] / (mr[tp] + mr[tq])
        imax = 0.5 * mr * ((up - uq) ** 2).sum(1) / (abi.HS_KB * th[tp])
        assert (imax >= 256).mean() > 0.3
    g = eng.binary_collide(tp, tq, up, uq, uni, ep, eq, vp, vq)
    c = ob.binary_collide(case, tp, tq, up, uq, uni, ep, eq, vp, vq)
    assert g[7] == 0 and c[7] == 0
    for k in range(7):
        np.testing.assert_array_equal(g[k], c[k], err_msg=f"output {k}")
    # total energy (translational + rotational + vibrational) conserved
    mass = np.array([case.species[0].mass, case.species[1].mass])
    th = np.array([case.species[0].theta_v[0], case.species[1].theta_v[0]])
    kb = abi.HS_KB

    def etot(u1, u2, e1, e2, v1, v2):
        return (0.5 * mass[tp] * (u1 ** 2).sum(1) + 0.5 * mass[tq] * (u2 ** 2).sum(1) + e1 + e2
                + v1[:, 0] * kb * th[tp] + v2[:, 0] * kb * th[tq])

    E0 = etot(up, uq, ep, eq, vp, vq)
    E1 = etot(g[0], g[1], g[2], g[3], g[4], g[5])
    assert np.max(np.abs(E1 - E0) / E0) < 1e-13
    P0 = mass[tp][:, None] * up + mass[tq][:, None] * uq
    P1 = mass[tp][:, None] * g[0] + mass[tq][:, None] * g[1]
    scale = (mass[tp][:, None] * np.abs(up) + mass[tq][:, None] * np.abs(uq)).sum(1)
    assert np.max(np.abs(P1 - P0).sum(1) / scale) < 1e-14
    eng.close()


@pytest.mark.parametrize("which", ["box", "cylinder", "couette"])
def test_move_sort_collide_stages(which):
    case, filler = {
        "box": (H.small_box(), H.fill_box),
        "cylinder": (H.small_cylinder(), H.fill_cylinder),
        "couette": (cases.couette_n2o2(), H.fill_couette),
    }[which]
    eng, st = _pair(case, filler)
    eng.move()
    st.move()
    _check_state(eng, st, f"{which}/move")
    eng.sort()
    st.sort()
    _check_state(eng, st, f"{which}/sort")
    counts_eng = np.bincount(eng.download()["cell"], minlength=case.ncells)
    np.testing.assert_array_equal(counts_eng, st.count)
    eng.collide(sample=True)
    st.collide(sample=True)
    _check_state(eng, st, f"{which}/collide")
    s_eng, n_eng = eng.samples()
    np.testing.assert_array_equal(s_eng, st.sample_array())
    assert n_eng == st.s.n_sampled == 1
    c = eng.counters()
    assert c["candidates"] == st.counters()[0]
    assert c["collisions"] == st.counters()[1]
    eng.close()


def _n2_box(flags):
    return cases.periodic_box(n_side=6, L=0.03, ppc=30, seed=12, species=[cases.make_species("N2")],
                              collision_model=abi.HS_COLL_LB_VHS, flags=flags)


@pytest.mark.parametrize("which,steps", [("box", 6), ("cylinder", 8), ("couette", 5), ("couette_hf", 4),
                                         ("n2_box_hf", 5)])
def test_evolve_parity(which, steps):
    """n2_box_hf: one molecular species (rotation and vibration, LB-VHS) with
    the heat-flux sums."""
    hf = abi.HS_FLAG_SAMPLE | abi.HS_FLAG_HEATFLUX | abi.HS_FLAG_COLL_SEP
    case, filler = {
        "n2_box_hf": (_n2_box(hf), H.fill_box),
        "box": (H.small_box(), H.fill_box),
        "cylinder": (H.small_cylinder(), H.fill_cylinder),
        "couette": (cases.couette_n2o2(), H.fill_couette),
        "couette_hf": (cases.couette_n2o2(flags=hf), H.fill_couette),
    }[which]
    eng, st = _pair(case, filler)
    for k in range(steps):
        eng.evolve(1)
        st.evolve(1)
        _check_state(eng, st, f"{which}/step{k}")
    s_eng, n_eng = eng.samples()
    np.testing.assert_array_equal(s_eng, st.sample_array())
    assert n_eng == steps
    if st.inflow is not None:
        np.testing.assert_array_equal(eng.inflow()[: st.inflow.size], st.inflow)
    c = eng.counters()
    oc = st.counters()
    assert (c["candidates"], c["collisions"], c["deleted"], c["injected"], c["wall_hits"]) == tuple(oc[:5])
    assert c["track_overflow"] == 0 and c["tile_overflow"] == 0 and c["capacity_overflow"] == 0
    assert eng.step == st.step == steps
    eng.close()


def test_zv_2008_with_cell_temperatures_matches_the_oracle():
    """inverseZvFormulation 2008 on the couette case: the cells' overall
    temperatures (zero in some cells, so those use the collision temperature)
    set on the engine and the oracle, changed between calls as a write does;
    every step bit-identical.  An engine of another formulation refuses them."""
    case = cases.couette_n2o2(flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_HEATFLUX)
    case.cfg.inv_zv_form = 1
    eng, st = _pair(case, H.fill_couette)
    rng = np.random.default_rng(8)
    for call in range(3):
        t = rng.uniform(800.0, 9000.0, case.ncells)
        t[rng.random(case.ncells) < 0.3] = 0.0
        eng.set_cell_temperature(t)
        st.set_cell_temperature(t)
        for k in range(2):
            eng.evolve(1)
            st.evolve(1)
            _check_state(eng, st, f"2008/call{call}/step{k}")
    np.testing.assert_array_equal(eng.samples()[0], st.sample_array())
    c, oc = eng.counters(), st.counters()
    assert (c["candidates"], c["collisions"]) == tuple(oc[:2])
    eng.close()
    other = Engine(cases.couette_n2o2())
    with pytest.raises(EngineError, match="inverseZvFormulation 2008"):
        other.set_cell_temperature(np.zeros(other.case.ncells))
    other.close()


@pytest.mark.parametrize("ppc,steps", [(600, 4), (5000, 2), (15000, 2)])
def test_dense_and_oversize_cell_parity(ppc, steps):
    """A single dense cell: 600 parcels run the wave-cooperative NTC (many
    candidates per step) in LDS; 5000 exceed an LDS tile and run it on the
    parcels in global memory with a stamp per parcel; 15000 exceed the stamps
    too, and parcels share them.  All must equal the candidate-by-candidate
    oracle."""
    case = cases.periodic_box(n_side=1, L=0.01, ppc=ppc, seed=17)
    eng, st = _pair(case, H.fill_box)
    assert st.n == ppc
    for k in range(steps):
        eng.evolve(1)
        st.evolve(1)
        _check_state(eng, st, f"dense{ppc}/step{k}")
    np.testing.assert_array_equal(eng.samples()[0], st.sample_array())
    assert eng.counters()["collisions"] > 50 * steps
    eng.close()


def test_sparse_wide_tiles_parity():
    """Two parcels per cell: a particle-balanced tile spans more cells than one
    sub-tile of 256, so the sub-tile loop and its boundaries are exercised."""
    case = cases.periodic_box(n_side=16, ppc=2, seed=23)
    eng, st = _pair(case, H.fill_box)
    for k in range(4):
        eng.evolve(1)
        st.evolve(1)
        _check_state(eng, st, f"sparse/step{k}")
    np.testing.assert_array_equal(eng.samples()[0], st.sample_array())
    eng.close()


@pytest.mark.parametrize("which", ["cylinder", "couette"])
def test_graph_replayed_steps_match_the_oracle(which):
    """evolve(n >= 3) replays a captured HIP graph of two steps (one per
    starting buffer) for all but its last step.  Every replay is checked
    against the oracle: calls of three steps (one replay, one eager step, with
    the inflow of the cylinder case injected on the side stream inside the
    graph) alternate between the two captured graphs; longer calls replay one
    graph several times; toggling the timing flag drops the graphs (timed
    steps run eagerly) and the next call captures them again."""
    case, filler = {
        "cylinder": (H.small_cylinder(), H.fill_cylinder),
        "couette": (cases.couette_n2o2(), H.fill_couette),
    }[which]
    eng, st = _pair(case, filler)
    flags = case.cfg.flags
    total = 0
    for r, n in enumerate((3, 3, 3, 3, 9, 1, 7, 4)):
        if r == 5:
            eng.set_flags(flags | abi.HS_FLAG_TIMING)
        if r == 6:
            eng.set_flags(flags)
        eng.evolve(n)
        st.evolve(n)
        total += n
        _check_state(eng, st, f"{which}/call{r}")
    s_eng, n_eng = eng.samples()
    np.testing.assert_array_equal(s_eng, st.sample_array())
    assert n_eng == total
    c, oc = eng.counters(), st.counters()
    assert (c["candidates"], c["collisions"], c["deleted"], c["injected"], c["wall_hits"]) == tuple(oc[:5])
    assert c["tile_overflow"] == 0 and c["capacity_overflow"] == 0
    if which == "cylinder":
        assert c["injected"] > 0
    assert eng.step == st.step == total
    eng.close()


@pytest.mark.parametrize("which,tiles", [("cylinder", (0, 0)), ("cylinder", (64, 64)), ("cylinder", (1536, 2048)),
                                         ("couette", (64, 96)), ("couette", (1536, 1536))])
def test_tilings_give_the_same_steps(which, tiles):
    """Parcels per gather and collide workgroup change no result: the engine
    matches the oracle with the automatic choice (small populations: 384 and
    256), the smallest tiles and the full-device ones."""
    case, filler = {
        "cylinder": (H.small_cylinder(), H.fill_cylinder),
        "couette": (cases.couette_n2o2(), H.fill_couette),
    }[which]
    eng, st = _pair(case, filler)
    eng.set_tiles(*tiles)
    got = eng.tiles()
    if tiles == (0, 0):
        assert got == (384, 256), got   # ~36k parcels on 256 compute units
    else:
        assert got == tiles
    for k in range(4):
        eng.evolve(1)
        st.evolve(1)
        _check_state(eng, st, f"{which}/{tiles}/step{k}")
    np.testing.assert_array_equal(eng.samples()[0], st.sample_array())
    eng.evolve(5)   # graph replays with these tiles
    st.evolve(5)
    _check_state(eng, st, f"{which}/{tiles}/after5")
    # new tile sizes between calls: the captured graphs are dropped and taken again
    eng.set_tiles(128, 64)
    eng.evolve(6)
    st.evolve(6)
    _check_state(eng, st, f"{which}/{tiles}/retiled")
    for bad in ((32, 0), (0, 4096), (-1, 0)):
        with pytest.raises(EngineError, match="tile sizes"):
            eng.set_tiles(*bad)
    eng.close()


def test_empty_domain_fills_from_the_inflow_like_the_oracle():
    """No initial parcels: the freestream inflow fills the cylinder case from
    empty.  Graph-replayed and single steps equal the oracle's, and a closed
    box with no parcels steps without creating or sampling any."""
    case = H.small_cylinder()
    eng, st = Engine(case), ob.OracleState(case)
    assert eng.count() == 0
    for n in (7, 1, 1, 4):
        eng.evolve(n)
        st.evolve(n)
        _check_state(eng, st, f"from-empty/{n}")
    assert eng.count() > 0
    np.testing.assert_array_equal(eng.samples()[0], st.sample_array())
    c, oc = eng.counters(), st.counters()
    assert (c["candidates"], c["collisions"], c["deleted"], c["injected"], c["wall_hits"]) == tuple(oc[:5])
    eng.close()
    box = H.small_box()
    with Engine(box) as e:
        e.evolve(5)
        e.evolve(1)
        assert e.count() == 0
        s, nsamp = e.samples()
        assert nsamp == 6 and not s.any()
        assert all(v == 0 for k, v in e.counters().items() if k != "particle_steps")


def test_inflow_beyond_capacity_is_dropped_and_counted():
    """A capacity of 1,500 parcels under a Mach-10 inflow of ~160 parcels a
    step: the population stops at the capacity, the parcels that did not fit
    are counted, and every stored parcel stays a valid one."""
    base = H.small_cylinder()
    case = cases.cylinder_flow(nx=60, ny=60, ppc=10, dt=2.5e-7 * 400 / 60, seed=5, capacity=1500,
                               flags=base.cfg.flags)
    with Engine(case) as e:
        for n in (9, 1, 6):
            e.evolve(n)
        c = e.counters()
        assert e.count() <= 1500
        assert c["capacity_overflow"] > 0
        assert c["injected"] - c["deleted"] == e.count()
        p = e.download()
        assert ((p["cell"] >= 0) & (p["cell"] < case.ncells)).all()
        assert np.isfinite(p["ux"]).all() and np.isfinite(p["uy"]).all()
        assert ((p["px"] >= -0.35) & (p["px"] <= 0.65) & (p["py"] >= -0.5) & (p["py"] <= 0.5)).all()


def test_batched_engines_match_the_oracle():
    """hs_dsmc_evolve_batch over three different cases at once (the cylinder
    with its inflow, the LB-VHS couette, a closed VHS box): every engine ends
    each call bit-identical to the oracle of its own case, over odd and even
    step counts; a batch naming one engine twice is rejected."""
    built = [(H.small_cylinder(), H.fill_cylinder), (cases.couette_n2o2(), H.fill_couette),
             (H.small_box(), H.fill_box)]
    pairs = [_pair(c, f) for c, f in built]
    engs = [e for e, _ in pairs]
    total = 0
    for n in (5, 1, 4):
        evolve_batch(engs, n)
        total += n
        for k, (e, st) in enumerate(pairs):
            st.evolve(n)
            _check_state(e, st, f"batch/engine{k}/+{n}")
    for e, st in pairs:
        s_eng, n_eng = e.samples()
        np.testing.assert_array_equal(s_eng, st.sample_array())
        assert n_eng == total and e.step == total
    with pytest.raises(EngineError, match="twice"):
        evolve_batch([engs[0], engs[1], engs[0]], 1)
    evolve_batch([], 3)
    for e in engs:
        e.close()


def test_autotuner_picks_a_measured_tiling_that_changes_no_result():
    """hs_dsmc_autotune steps the engine under each candidate tiling and
    keeps the fastest: the steps it takes are ordinary steps (the engine still
    equals the oracle advanced by as many steps), the choice is a candidate
    pair, and a second engine of the same population class starts from it."""
    case = H.small_cylinder(n=60, ppc=10)
    eng, st = _pair(case, H.fill_cylinder)
    s0 = eng.step
    g, c, ms = eng.autotune(steps_per_trial=3)
    assert g in (256, 512, 1024, 1536) and c in (256, 512, 1024, 1536, 2048) and ms > 0
    taken = eng.step - s0
    assert taken > 0 and taken % 5 == 0
    st.evolve(taken)
    _check_state(eng, st, "autotune")
    assert eng.tiles() == (g, c)
    eng2, st2 = _pair(case, H.fill_cylinder)
    assert eng2.tiles() == (g, c)
    eng2.evolve(3)
    st2.evolve(3)
    _check_state(eng2, st2, "autotuned tiles")
    eng.close()
    eng2.close()


def test_fill_cells_parity():
    """hs_dsmc_fill_cells (a population from per-cell fields: the warm
    start's initialisation) against the oracle's, bit-exact: counts,
    positions, velocities, internal energies, sigmaTcRMax and remainders."""
    for case in (H.small_cylinder(n=40, ppc=10), cases.couette_n2o2()):
        nc, ns = case.ncells, case.cfg.n_species
        rng = np.random.default_rng(8)
        base = np.array(getattr(case, "init_n", [getattr(case, "n_inf", 1e20)]))
        n_cell = (base[:, None] * rng.uniform(0.3, 3.0, (ns, nc))).ravel()
        u_cell = rng.normal(0, 800, (3, nc))
        t_cell = rng.uniform(100.0, 5000.0, nc)
        eng = Engine(case)
        st = ob.OracleState(case)
        eng.fill_cells(n_cell, u_cell, t_cell, fill_id=5)
        st.fill_cells(n_cell, u_cell, t_cell, fill_id=5)
        assert st.n > 0
        _check_state(eng, st, f"fill_cells/{case.name}")
        eng.evolve(2)
        st.evolve(2)
        _check_state(eng, st, f"fill_cells/{case.name}/2 steps")
        eng.close()
