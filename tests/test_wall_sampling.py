"""Wall-face sampling (boundaryMeasurements, dsmcPatchBoundary.C:263-483):
the oracle pinned by geometry closed forms, exact energy/momentum accounting
and equilibrium kinetic theory; the engine against the oracle on the GPU."""
import math

import numpy as np
import pytest

from hystrath_amd import abi, cases, fields
from tests import oracle_bind as ob
from tests.helpers import assert_same_particles

KB = abi.HS_KB


# ---- geometry --------------------------------------------------------------
def test_arc_in_rect_closed_forms():
    R = 0.1
    L, _ = ob.arc_in_rect(0.0, 0.0, R, -1.0, 1.0, -1.0, 1.0)
    assert L == pytest.approx(2 * math.pi * R, rel=1e-14)
    L, n = ob.arc_in_rect(0.0, 0.0, R, 0.0, 1.0, 0.0, 1.0)          # first quadrant
    assert L == pytest.approx(0.5 * math.pi * R, rel=1e-14)
    np.testing.assert_allclose(n, [-math.sqrt(0.5), -math.sqrt(0.5)], rtol=1e-13)
    L, n = ob.arc_in_rect(0.0, 0.0, R, -1.0, 1.0, 0.0, 1.0)         # upper half
    assert L == pytest.approx(math.pi * R, rel=1e-14)
    np.testing.assert_allclose(n, [0.0, -1.0], atol=1e-13)
    L, n = ob.arc_in_rect(0.0, 0.0, R, 0.05, 1.0, -1.0, 1.0)        # cap x >= R/2: arc 2*acos(1/2)*R
    assert L == pytest.approx(2 * math.acos(0.5) * R, rel=1e-13)
    np.testing.assert_allclose(n, [-1.0, 0.0], atol=1e-13)
    assert ob.arc_in_rect(0.0, 0.0, R, 0.2, 0.3, 0.2, 0.3)[0] == 0.0  # disjoint
    assert ob.arc_in_rect(0.0, 0.0, R, -0.01, 0.01, -0.01, 0.01)[0] == 0.0  # inside the disk


def test_arc_in_rect_against_brute_force():
    rng = np.random.default_rng(4)
    R, cx, cy = 0.1, 0.013, -0.007
    th = (np.arange(2_000_000) + 0.5) * (2 * math.pi / 2_000_000)
    px, py = cx + R * np.cos(th), cy + R * np.sin(th)
    for _ in range(40):
        x0, y0 = rng.uniform(-0.15, 0.12, 2)
        w, h = rng.uniform(0.005, 0.1, 2)
        inside = (px >= x0) & (px <= x0 + w) & (py >= y0) & (py <= y0 + h)
        L, _ = ob.arc_in_rect(cx, cy, R, x0, x0 + w, y0, y0 + h)
        assert L == pytest.approx(inside.mean() * 2 * math.pi * R, abs=4e-6 * R)


def test_wall_face_layout():
    case = cases.wall_channel(nx=4, ny=20)
    f = ob.wall_faces(case)
    # xmin then xmax, each with ny*nz faces, y index fastest
    assert list(f["side"]) == [0] * 20 + [1] * 20
    assert list(f["cell"][:20]) == [4 * j for j in range(20)]
    assert list(f["cell"][20:]) == [4 * j + 3 for j in range(20)]
    d = case.cfg.delta[0]
    np.testing.assert_allclose(f["area"], d * d, rtol=1e-15)
    np.testing.assert_array_equal(f["normal"][:20], [[-1.0, 0.0, 0.0]] * 20)
    np.testing.assert_array_equal(f["normal"][20:], [[1.0, 0.0, 0.0]] * 20)
    # a cylinder: deletion sides carry no wall faces; its faces cover 2 pi R dz
    cyl = cases.cylinder_flow(nx=200, ny=200, ppc=2)
    g = ob.wall_faces(cyl)
    assert set(g["side"]) == {abi.HS_WALL_CYLINDER}
    # (grid lines at +-R: the slivers of ~2e-9 m left out are below 1e-7 of the length)
    assert g["area"].sum() == pytest.approx(2 * math.pi * 0.1 * cyl.cfg.delta[2], rel=1e-7)
    cell = g["cell"]
    assert (np.diff(cell) > 0).all()
    assert (g["area"] > 1e-6 * cyl.cfg.delta[0] * cyl.cfg.delta[2]).all()
    # every face cell holds fluid
    full = cyl.cfg.delta[0] * cyl.cfg.delta[1] * cyl.cfg.delta[2]
    v = cyl.volumes[cell]
    assert ((v > 0) & (v <= full * (1 + 1e-12))).all()
    # off-grid radius: the faces cover 2 pi R dz to round-off
    cyl2 = cases.cylinder_flow(nx=97, ny=89, R=0.0937, ppc=2)
    g2 = ob.wall_faces(cyl2)
    assert g2["area"].sum() == pytest.approx(2 * math.pi * 0.0937 * cyl2.cfg.delta[2], rel=1e-12)
    # normals point into the body: toward the centre from the face cell
    cx, cy, _ = cyl.cell_centres()
    assert ((g["normal"][:, 0] * cx[cell] + g["normal"][:, 1] * cy[cell]) < 0).all()


# ---- oracle: exact accounting -------------------------------------------------
def _energy_momentum(case, d):
    sp = case.species
    t = d["type"].astype(int) if "type" in d else np.zeros(len(d["px"]), int)
    m = np.array([s.mass for s in sp])[t]
    U = np.stack([d["ux"], d["uy"], d["uz"]]).astype(np.float64)
    E = 0.5 * m * (U * U).sum(axis=0)
    if d.get("erot") is not None:
        E = E + d["erot"].astype(np.float64)
    if d.get("vib") is not None:
        for k in range(d["vib"].shape[0]):
            th = np.array([s.theta_v[k] if s.n_vib_modes > k else 0.0 for s in sp])[t]
            E = E + d["vib"][k].astype(np.float64) * KB * th
    return E.sum(), (m * U).sum(axis=1)


@pytest.mark.parametrize("coll", [abi.HS_COLL_NONE, abi.HS_COLL_VHS])
def test_wall_sums_account_for_all_energy_and_momentum(coll):
    """sum_f q_f A_f dt = Fn (E_gas before - after) and the same for fD and
    momentum: the walls are the only sink, so any error of sign, area, dt or
    mass in the wall sums breaks these identities."""
    case = cases.wall_channel(collision_model=coll)
    o = ob.OracleState(case)
    o.fill([case.n_density], case.T_gas, exact=True)
    E0, P0 = _energy_momentum(case, o.particles())
    n0 = o.n
    o.evolve(40)
    assert o.n == n0
    E1, P1 = _energy_momentum(case, o.particles())
    W = o.wall_array()
    f = ob.wall_faces(case)
    Adt = f["area"] * case.cfg.dt
    q_tot = (W[abi.HS_WS_Q] * Adt).sum()
    fd_tot = (W[abi.HS_WS_FDX:abi.HS_WS_FDZ + 1] * Adt).sum(axis=1)
    dE = case.cfg.fn * (E0 - E1)
    dP = case.cfg.fn * (P0 - P1)
    assert abs(dE) > 1e-3 * case.cfg.fn * E0            # the walls did exchange energy
    assert q_tot == pytest.approx(dE, rel=1e-6)
    np.testing.assert_allclose(fd_tot, dP, rtol=1e-6, atol=1e-6 * np.abs(dP).max())
    assert o.counters()[4] > 300


def test_specular_walls_exchange_no_energy_or_tangential_momentum():
    case = cases.wall_channel(lo_model=abi.HS_PATCH_SPECULAR, hi_model=abi.HS_PATCH_SPECULAR)
    o = ob.OracleState(case)
    o.fill([case.n_density], case.T_gas, exact=True)
    o.evolve(20)
    W = o.wall_array()
    assert (W[abi.HS_WS_Q] == 0.0).all()
    assert (W[abi.HS_WS_FDY] == 0.0).all() and (W[abi.HS_WS_FDZ] == 0.0).all()
    assert (W[abi.HS_WS_FDX][:20] < 0).all() and (W[abi.HS_WS_FDX][20:] > 0).all()  # pushes outward
    assert W[abi.HS_WS_RHON].min() > 0


def test_equilibrium_wall_state_is_the_gas_state():
    """Diffuse walls at the gas temperature: the flux-weighted wall density is
    n, p = n k T, Ttra = T, no shear, no net heat flux (kinetic theory)."""
    T, n = 500.0, 1e21
    case = cases.wall_channel(ny=10, T_gas=T, T_lo=T, T_hi=T, u_hi=(0.0, 0.0, 0.0), ppc=200, n_density=n,
                              collision_model=abi.HS_COLL_NONE)
    o = ob.OracleState(case)
    o.fill([n], T, exact=True)
    o.evolve(300)
    faces = ob.wall_faces(case)
    W = o.wall_array()
    nsteps = o.s.n_sampled
    assert nsteps == 300
    # pool each wall's faces for statistics
    for sl in (slice(0, 10), slice(10, 20)):
        Wp = W[:, sl].sum(axis=1, keepdims=True)
        fp = {k: v[sl][:1] for k, v in faces.items()}
        wf = fields.wall_fields(Wp / 10.0, nsteps, case, fp)
        assert wf["rhoN"][0] == pytest.approx(n, rel=0.04)
        assert wf["p"][0] == pytest.approx(n * KB * T, rel=0.04)
        assert wf["Ttra"][0] == pytest.approx(T, rel=0.04)
        assert wf["tau"][0] < 0.05 * wf["p"][0]
        # one-way energy flux 2kT * n sqrt(kT/(2 pi m)); the net is a few % of it
        m = case.species[0].mass
        one_way = 2 * KB * T * n * math.sqrt(KB * T / (2 * math.pi * m))
        assert abs(wf["q"][0]) < 0.06 * one_way


# ---- engine vs oracle ---------------------------------------------------------------
def _parity(case, fill, steps):
    from hystrath_amd import engine
    o = ob.OracleState(case)
    fill(o)
    with engine.Engine(case) as e:
        fill(e)
        e.evolve(steps)
        o.evolve(steps)
        assert_same_particles(e.download(), o.particles(), case.name)
        fe, fo = e.wall_faces(), ob.wall_faces(case)
        for k in ("side", "cell", "area", "normal"):
            np.testing.assert_array_equal(fe[k], fo[k])
        We, ns = e.wall_samples()
        Wo = o.wall_array()
        assert ns == o.s.n_sampled == steps
        assert np.abs(Wo).max() > 0
        # FP64 atomics add in another order: equal to a relative 1e-10 of each row's scale
        scale = np.abs(Wo).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(We - Wo) <= 1e-10 * scale).all()
        return We, Wo


@pytest.mark.gpu
def test_wall_sampling_parity_channel():
    case = cases.wall_channel()
    _parity(case, lambda x: x.fill([case.n_density], case.T_gas, exact=True), 12)


@pytest.mark.gpu
def test_wall_sampling_parity_cylinder():
    case = cases.cylinder_flow(nx=60, ny=60, ppc=10, dt=2.5e-7 * 400 / 60, seed=5,
                               flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL)
    _parity(case, lambda x: x.fill([case.n_inf], case.T_inf, u=(case.U_inf, 0.0, 0.0)), 15)


@pytest.mark.gpu
def test_wall_sampling_parity_couette_lb():
    case = cases.couette_n2o2(flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL)
    _parity(case, lambda x: x.fill(case.init_n, case.init_T), 4)


# ---- face tracker (dsmcFaceTracker.C:124-193) -----------------------------------------
def _face_planes(cfg):
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    off = [0, (nx + 1) * ny * nz]
    off.append(off[1] + nx * (ny + 1) * nz)
    off.append(off[2] + nx * ny * (nz + 1))
    return off


def _face(cfg, off, a, k, i, j, kk):
    nn = (cfg.nx, cfg.ny, cfg.nz)
    idx = (i, j, kk)
    a1, a2 = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
    return off[a] + k + (nn[a] + 1) * (idx[a1] + nn[a1] * idx[a2])


def _cell_counts(case, d):
    c = d["cell"]
    return np.bincount(c[c >= 0], minlength=case.ncells)


def test_face_tracker_cell_balance_and_wall_signs():
    """Every cell away from the cyclic sides gains exactly its signed
    internal-face flux; each wall hit counts -1 (the reflected parcel moves
    against the outward normal)."""
    case = cases.wall_channel(nx=6, ny=12, flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_FACE_FLUX)
    cfg = case.cfg
    o = ob.OracleState(case)
    o.fill([case.n_density], case.T_gas, exact=True)
    N0 = _cell_counts(case, o.particles())
    o.evolve(30)
    N1 = _cell_counts(case, o.particles())
    F = o.flux.sum(axis=0)
    off = _face_planes(cfg)
    assert o.n_faces == off[3]
    for j in range(1, cfg.ny - 1):
        for i in range(cfg.nx):
            c = i + cfg.nx * j
            inflow = 0
            if i > 0:
                inflow += F[_face(cfg, off, 0, i, i, j, 0)]
            if i < cfg.nx - 1:
                inflow -= F[_face(cfg, off, 0, i + 1, i, j, 0)]
            inflow += F[_face(cfg, off, 1, j, i, j, 0)] - F[_face(cfg, off, 1, j + 1, i, j, 0)]
            assert N1[c] - N0[c] == inflow, (i, j)
    walls = [_face(cfg, off, 0, 0, 0, j, 0) for j in range(cfg.ny)] + \
            [_face(cfg, off, 0, cfg.nx, cfg.nx - 1, j, 0) for j in range(cfg.ny)]
    assert (F[walls] <= 0).all()
    assert -F[walls].sum() == o.counters()[4] > 0
    # z is empty: never crossed; the cyclic y planes count unsigned
    assert (F[off[2]:off[3]] == 0).all()
    cyc = [_face(cfg, off, 1, k, i, k if k < cfg.ny else cfg.ny - 1, 0) for k in (0, cfg.ny) for i in range(cfg.nx)]
    assert (F[cyc] >= 0).all() and F[cyc].sum() > 0


def test_face_tracker_deletion_and_cylinder():
    """Outer deletion faces count +1 per deleted parcel; the cylinder's wall
    faces -1 per hit (the only wall)."""
    case = cases.cylinder_flow(nx=40, ny=40, ppc=6, dt=2.5e-7 * 10, seed=3,
                               flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_FACE_FLUX)
    o = ob.OracleState(case)
    o.fill([case.n_inf], case.T_inf, u=(case.U_inf, 0.0, 0.0))
    o.evolve(12)
    F = o.flux.sum(axis=0)
    off = _face_planes(case.cfg)
    nw = len(ob.wall_faces(case)["side"])
    assert o.n_faces == off[3] + nw
    cnt = o.counters()
    assert -F[off[3]:].sum() == cnt[4] > 0
    assert (F[off[3]:] <= 0).all()
    nx, ny = case.cfg.nx, case.cfg.ny
    outer = [_face(case.cfg, off, 0, k, 0, j, 0) for k in (0, nx) for j in range(ny)] + \
            [_face(case.cfg, off, 1, k, i, 0, 0) for k in (0, ny) for i in range(nx)]
    assert F[outer].sum() == cnt[2] > 0
    # per species: mass flux is mass times the count (one species here)
    assert o.flux.shape == (1, o.n_faces)


@pytest.mark.gpu
def test_face_tracker_parity():
    from hystrath_amd import engine
    for case, fill, steps in (
        (cases.wall_channel(nx=6, ny=12, flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_FACE_FLUX | abi.HS_FLAG_WALL),
         lambda x, c: x.fill([c.n_density], c.T_gas, exact=True), 10),
        (cases.cylinder_flow(nx=60, ny=60, ppc=10, dt=2.5e-7 * 400 / 60, seed=5,
                             flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_FACE_FLUX),
         lambda x, c: x.fill([c.n_inf], c.T_inf, u=(c.U_inf, 0.0, 0.0)), 10),
        (cases.couette_n2o2(flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_FACE_FLUX), lambda x, c: x.fill(c.init_n, c.init_T),
         4),
    ):
        o = ob.OracleState(case)
        fill(o, case)
        with engine.Engine(case) as e:
            fill(e, case)
            e.evolve(steps)
            o.evolve(steps)
            assert_same_particles(e.download(), o.particles(), case.name)
            Fe, ns = e.face_flux()
            assert ns == steps
            np.testing.assert_array_equal(Fe, o.flux)
            assert np.abs(Fe).sum() > 0
            e.reset_samples()
            assert not e.face_flux()[0].any()
