"""Case set-up for the structured-block engine: species tables, meshes with
exact cut-cell volumes, particle SoA containers and the standard workloads.

The species values are the ones the reference tutorials use in
constant/dsmcProperties (e.g. run/hyStrath/dsmcFoam+/couette_N2-O2/
constant/dsmcProperties:48-78); argon follows the workload definition
(m = 6.63e-26 kg, d = 4.17e-10 m, omega = 0.81, Tref = 273 K).
"""
import math

import numpy as np

from . import abi

KB = abi.HS_KB

SPECIES = {
    "Ar": dict(mass=6.63e-26, diameter=4.17e-10, omega=0.81, alpha=1.0),
    # couette_N2-O2/constant/dsmcProperties
    "N2": dict(mass=46.5e-27, diameter=4.17e-10, omega=0.74, alpha=1.36, rot_dof=2.0,
               n_vib_modes=1, theta_v=[3371.0], zref=[52560.0], tref_zv=[3371.0],
               theta_d=113500.0),
    "O2": dict(mass=53.12e-27, diameter=4.07e-10, omega=0.77, alpha=1.4, rot_dof=2.0,
               n_vib_modes=1, theta_v=[2256.0], zref=[17900.0], tref_zv=[2256.0],
               theta_d=59500.0),
    # heatBath-5species/constant/dsmcProperties
    "NO": dict(mass=49.81e-27, diameter=4.2e-10, omega=0.79, alpha=1.0, rot_dof=2.0,
               n_vib_modes=1, theta_v=[2719.0], zref=[1400.0], tref_zv=[2719.0],
               theta_d=75500.0),
    "N": dict(mass=23.25e-27, diameter=3.00e-10, omega=0.8, alpha=1.0),
    "O": dict(mass=26.56e-27, diameter=3.00e-10, omega=0.8, alpha=1.0),
}

# heatBath-5species/system/chemReactDict, species (N2 O2 NO N O): reaction
# model, reactants, dissociation products per reactant, and for the exchange
# models (products, heatOfReactionExchange K, aCoeff, bCoeff)
HEATBATH_REACTIONS = [
    ("dissociationQK", ("O2", "N2"), (("O", "O"), ("N", "N")), None),
    ("dissociationExchangeQK", ("O2", "N"), (("O", "O"), ()), (("NO", "O"), 19693.6, 0.1, 0.1)),
    ("dissociationQK", ("O2", "O2"), (("O", "O"), ("O", "O")), None),
    ("dissociationQK", ("O2", "O"), (("O", "O"), ()), None),
    ("dissociationQK", ("O2", "NO"), (("O", "O"), ("N", "O")), None),
    ("dissociationQK", ("N2", "N2"), (("N", "N"), ("N", "N")), None),
    ("dissociationQK", ("N2", "N"), (("N", "N"), ()), None),
    ("dissociationExchangeQK", ("N2", "O"), (("N", "N"), ()), (("NO", "N"), -37482.0, 0.15, 0.15)),
    ("dissociationQK", ("N2", "NO"), (("N", "N"), ("N", "O")), None),
    ("dissociationExchangeQK", ("NO", "N"), (("N", "O"), ()), (("N2", "O"), 37482.0, 0.033, 0.8)),
    ("dissociationExchangeQK", ("NO", "O"), (("N", "O"), ()), (("O2", "N"), -19693.6, 0.085, 0.65)),
    ("dissociationQK", ("NO", "NO"), (("N", "O"), ("N", "O")), None),
]


def make_species(name=None, **kw):
    d = dict(SPECIES[name]) if name else {}
    d.update(kw)
    s = abi.hs_species()
    s.mass = d["mass"]
    s.diameter = d["diameter"]
    s.omega = d["omega"]
    s.alpha = d.get("alpha", 1.0)
    s.rot_dof = d.get("rot_dof", 0.0)
    s.n_vib_modes = d.get("n_vib_modes", 0)
    s.charge = d.get("charge", 0)
    for m, v in enumerate(d.get("theta_v", [])):
        s.theta_v[m] = v
    for m, v in enumerate(d.get("zref", [])):
        s.zref[m] = v
    for m, v in enumerate(d.get("tref_zv", [])):
        s.tref_zv[m] = v
    s.theta_d = d.get("theta_d", 0.0)
    return s


# --------------------------------------------------------------------------
# cut-cell geometry
# --------------------------------------------------------------------------
def _disk_rect_area(x0, x1, y0, y1, R):
    """Exact area of [x0,x1]x[y0,y1] intersected with the disk |p| <= R."""
    a, b = max(x0, -R), min(x1, R)
    if a >= b:
        return 0.0
    pts = {a, b}
    for y in (y0, y1):
        if abs(y) < R:
            w = math.sqrt(R * R - y * y)
            for x in (-w, w):
                if a < x < b:
                    pts.add(x)
    pts = sorted(pts)

    def G(x):  # antiderivative of sqrt(R^2 - x^2), well conditioned near |x| = R
        x = min(max(x, -R), R)
        s = math.sqrt(max((R - x) * (R + x), 0.0))
        return 0.5 * (x * s + R * R * math.atan2(x, s))

    area = 0.0
    for lo, hi in zip(pts[:-1], pts[1:]):
        xm = 0.5 * (lo + hi)
        h = math.sqrt(max(R * R - xm * xm, 0.0))
        top_is_h = h < y1
        bot_is_h = -h > y0
        top = h if top_is_h else y1
        bot = -h if bot_is_h else y0
        if top <= bot:
            continue
        # integrand = top - bot with top/bot either a constant or +-h(x)
        c = (0.0 if top_is_h else y1) - (0.0 if bot_is_h else y0)
        k = (1.0 if top_is_h else 0.0) + (1.0 if bot_is_h else 0.0)
        area += c * (hi - lo) + k * (G(hi) - G(lo))
    return min(max(area, 0.0), (x1 - x0) * (y1 - y0))


def cell_volumes(nx, ny, nz, origin, delta, cyl=None):
    """Fluid volume of every cell (blockMesh order); cyl = (cx, cy, R) removes
    the disk (axis z) from each cell exactly."""
    full = delta[0] * delta[1] * delta[2]
    vol = np.full(nx * ny * nz, full, dtype=np.float64)
    if cyl is None:
        return vol
    cx, cy, R = cyl
    for j in range(ny):
        y0 = origin[1] + j * delta[1] - cy
        y1 = origin[1] + (j + 1) * delta[1] - cy
        ny_ = 0.0 if y0 <= 0.0 <= y1 else min(abs(y0), abs(y1))
        for i in range(nx):
            x0 = origin[0] + i * delta[0] - cx
            x1 = origin[0] + (i + 1) * delta[0] - cx
            nx_ = 0.0 if x0 <= 0.0 <= x1 else min(abs(x0), abs(x1))
            if nx_ * nx_ + ny_ * ny_ >= R * R:
                continue
            cut = _disk_rect_area(x0, x1, y0, y1, R)
            frac = 1.0 - cut / (delta[0] * delta[1])
            v = 0.0 if frac < 1e-12 else frac * full
            for k in range(nz):
                vol[i + nx * (j + ny * k)] = v
    return vol


# --------------------------------------------------------------------------
# configuration
# --------------------------------------------------------------------------
class Case:
    """An hs_config plus the host arrays it points to (kept alive here)."""

    def __init__(self, nx, ny, nz, origin, delta, species, patches, *, collision_model=abi.HS_COLL_VHS,
                 fn=1.0, dt=1e-6, seed=1234, capacity=1 << 20, flags=abi.HS_FLAG_SAMPLE, cyl=None,
                 cyl_patch=None, tref=273.0, z_rot=5.0, z_vib_fixed=0.0, z_el=500.0, inv_zv_form=2,
                 inflow=None, name="case"):
        self.name = name
        self.volumes = cell_volumes(nx, ny, nz, origin, delta, cyl)
        c = abi.hs_config()
        c.nx, c.ny, c.nz = nx, ny, nz
        for a in range(3):
            c.origin[a] = origin[a]
            c.delta[a] = delta[a]
        c.cell_volume = abi.ptr(self.volumes, abi.C.c_double)
        for s, p in enumerate(patches):
            c.patch[s] = p
        if cyl is not None:
            c.cyl_enabled = 1
            c.cyl_center[0], c.cyl_center[1], c.cyl_radius = cyl
            c.cyl_patch = cyl_patch
        c.n_species = len(species)
        for i, s in enumerate(species):
            c.species[i] = s
        c.collision_model = collision_model
        c.tref = tref
        c.z_rot = z_rot
        c.z_vib_fixed = z_vib_fixed
        c.z_el = z_el
        c.inv_zv_form = inv_zv_form
        if inflow:
            for i, nd in enumerate(inflow["n"]):
                c.inflow_n[i] = nd
            c.inflow_t = inflow["T"]
            c.inflow_trot = inflow.get("Trot", inflow["T"])
            c.inflow_tvib = inflow.get("Tvib", inflow["T"])
            for a in range(3):
                c.inflow_u[a] = inflow["U"][a]
        c.fn = fn
        c.dt = dt
        c.seed = seed
        c.capacity = capacity
        c.flags = flags
        self.cfg = c
        self.species = species

    def set_reactions(self, table, names, allow_splitting=True):
        """chemReactDict's reactions: entries (reactionModel, reactants,
        dissociationProducts, exchange) as in HEATBATH_REACTIONS, with
        exchange = (exchangeProducts, heatOfReactionExchange, aCoeff, bCoeff)
        or None; names are the typeIdList."""
        arr = (abi.hs_reaction * max(len(table), 1))()
        for k, (model, reac, diss, exch) in enumerate(table):
            r = arr[k]
            r.model = abi.REACTION_MODELS[model]
            for i, nm in enumerate(reac):
                r.reactants[i] = names.index(nm)
            for i in range(2):
                prods = diss[i] if i < len(diss) else ()
                for j in range(2):
                    r.diss_products[i][j] = names.index(prods[j]) if prods else -1
            if exch is not None:
                prods, heat, a, b = exch
                r.exch_products[0] = names.index(prods[0])
                r.exch_products[1] = names.index(prods[1])
                r.heat_exchange, r.a_coeff, r.b_coeff = heat, a, b
            else:
                r.exch_products[0] = r.exch_products[1] = -1
            r.allow_splitting = 1 if allow_splitting else 0
        self._reactions = arr
        self.cfg.reactions = arr if table else None
        self.cfg.n_reactions = len(table)

    @property
    def ncells(self):
        return self.cfg.nx * self.cfg.ny * self.cfg.nz

    @property
    def has_rot(self):
        return any(s.rot_dof > 0 for s in self.species)

    @property
    def vib_modes(self):
        return max([s.n_vib_modes for s in self.species] + [0])

    def cell_centres(self):
        c = self.cfg
        i = np.arange(c.nx)
        j = np.arange(c.ny)
        k = np.arange(c.nz)
        K, J, I = np.meshgrid(k, j, i, indexing="ij")
        return (c.origin[0] + (I.ravel() + 0.5) * c.delta[0],
                c.origin[1] + (J.ravel() + 0.5) * c.delta[1],
                c.origin[2] + (K.ravel() + 0.5) * c.delta[2])


def patch(model, temperature=0.0, velocity=(0.0, 0.0, 0.0), inflow=0):
    p = abi.hs_patch()
    p.model = model
    p.inflow = inflow
    p.temperature = temperature
    for a in range(3):
        p.velocity[a] = velocity[a]
    return p


class Particles:
    """Host structure-of-arrays with the engine's layout."""

    def __init__(self, capacity, vib_modes=0, rot=False, multi=False):
        self.capacity = capacity
        self.vib_modes = vib_modes
        self.px = np.zeros(capacity, np.float64)
        self.py = np.zeros(capacity, np.float64)
        self.pz = np.zeros(capacity, np.float64)
        self.ux = np.zeros(capacity, np.float32)
        self.uy = np.zeros(capacity, np.float32)
        self.uz = np.zeros(capacity, np.float32)
        self.cell = np.full(capacity, -1, np.int32)
        self.erot = np.zeros(capacity, np.float32) if rot else None
        self.vib = np.zeros(vib_modes * capacity, np.uint16) if vib_modes else None
        self.type = np.zeros(capacity, np.uint8) if multi else None

    @classmethod
    def for_case(cls, case, capacity=None):
        return cls(capacity or case.cfg.capacity, case.vib_modes, case.has_rot, case.cfg.n_species > 1)

    def struct(self):
        s = abi.hs_particles()
        s.px = abi.ptr(self.px, abi.C.c_double)
        s.py = abi.ptr(self.py, abi.C.c_double)
        s.pz = abi.ptr(self.pz, abi.C.c_double)
        s.ux = abi.ptr(self.ux, abi.C.c_float)
        s.uy = abi.ptr(self.uy, abi.C.c_float)
        s.uz = abi.ptr(self.uz, abi.C.c_float)
        s.cell = abi.ptr(self.cell, abi.C.c_int32)
        s.erot = abi.ptr(self.erot, abi.C.c_float)
        s.vib = abi.ptr(self.vib, abi.C.c_uint16)
        s.type = abi.ptr(self.type, abi.C.c_uint8)
        return s

    def view(self, n):
        """dict of the first n entries (copies)."""
        d = dict(px=self.px[:n].copy(), py=self.py[:n].copy(), pz=self.pz[:n].copy(),
                 ux=self.ux[:n].copy(), uy=self.uy[:n].copy(), uz=self.uz[:n].copy(),
                 cell=self.cell[:n].copy())
        if self.erot is not None:
            d["erot"] = self.erot[:n].copy()
        if self.vib is not None:
            d["vib"] = self.vib.reshape(self.vib_modes, self.capacity)[:, :n].copy()
        if self.type is not None:
            d["type"] = self.type[:n].copy()
        return d

    def load(self, d):
        n = len(d["px"])
        for k in ("px", "py", "pz", "ux", "uy", "uz", "cell"):
            getattr(self, k)[:n] = d[k]
        if self.erot is not None and "erot" in d:
            self.erot[:n] = d["erot"]
        if self.vib is not None and "vib" in d:
            self.vib.reshape(self.vib_modes, self.capacity)[:, :n] = d["vib"]
        if self.type is not None and "type" in d:
            self.type[:n] = d["type"]
        return n


# --------------------------------------------------------------------------
# standard workloads
# --------------------------------------------------------------------------
def periodic_box(n_side=20, L=0.1, ppc=20, n_density=1e20, T_hot=1000.0, dt=5.770e-6,
                 seed=2024, capacity=None, flags=abi.HS_FLAG_SAMPLE,
                 collision_model=abi.HS_COLL_VHS, species=None):
    """3D periodic argon box: n_side^3 cells of side L/n_side, exactly ppc
    simulators per cell, VHS argon (omega 0.81, d 4.17e-10, Tref 273)."""
    d = L / n_side
    V = d ** 3
    fn = n_density * V / ppc
    cyc = patch(abi.HS_PATCH_CYCLIC)
    sp = species or [make_species("Ar")]
    nc = n_side ** 3
    cap = capacity or int(nc * ppc * 1.3) + 1024
    case = Case(n_side, n_side, n_side, (0.0, 0.0, 0.0), (d, d, d), sp, [cyc] * 6,
                collision_model=collision_model, fn=fn, dt=dt, seed=seed, capacity=cap,
                flags=flags, name="periodic_box")
    case.n_density = n_density
    case.T_hot = T_hot
    return case


def cylinder_flow(nx=400, ny=400, Lx=1.0, Ly=1.0, x_cyl=0.35, R=0.1, n_inf=4.2e20, T_inf=200.0,
                  mach=10.0, T_wall=500.0, ppc=25, dt=2.5e-7, seed=7, capacity=None,
                  flags=abi.HS_FLAG_SAMPLE, species=None, collision_model=abi.HS_COLL_VHS):
    """2D Mach-10 argon flow over a cylinder of radius R: a Cartesian block
    of nx*ny cells (one cell thick in z, empty), the cylinder cut out exactly,
    diffuse isothermal wall at T_wall, freestream inflow + deletion on the
    four outer sides.  species (default argon) may be a molecule; its
    internal temperatures at the inflow equal T_inf."""
    ar = species or make_species("Ar")
    dx, dy = Lx / nx, Ly / ny
    dz = dx
    gamma = (5.0 + ar.rot_dof) / (3.0 + ar.rot_dof)
    a_inf = math.sqrt(gamma * KB * T_inf / ar.mass)
    U = mach * a_inf
    fn = n_inf * dx * dy * dz / ppc
    outer = patch(abi.HS_PATCH_DELETION, inflow=1)
    empty = patch(abi.HS_PATCH_EMPTY)
    wall = patch(abi.HS_PATCH_DIFFUSE, temperature=T_wall)
    origin = (-x_cyl, -0.5 * Ly, 0.0)
    cap = capacity or int(nx * ny * ppc * 1.6) + (1 << 16)
    case = Case(nx, ny, 1, origin, (dx, dy, dz), [ar], [outer, outer, outer, outer, empty, empty],
                collision_model=collision_model, fn=fn, dt=dt, seed=seed, capacity=cap, flags=flags,
                cyl=(0.0, 0.0, R), cyl_patch=wall,
                inflow=dict(n=[n_inf], T=T_inf, U=(U, 0.0, 0.0)), name="cylinder_flow")
    case.n_inf, case.T_inf, case.U_inf = n_inf, T_inf, U
    return case


def air_cylinder(nx=1000, ny=1000, Lx=0.1, Ly=0.1, x_cyl=0.035, R=0.01, n_inf=1e21, T_inf=220.0, U=7000.0,
                 T_wall=1000.0, ppc=30, dt=None, seed=41, capacity=None, flags=abi.HS_FLAG_SAMPLE):
    """Reacting five-species air (N2 O2 NO N O, the species and the twelve QK
    reactions of the heatBath-5species tutorial) at U past a diffuse
    isothermal cylinder of radius R: a Cartesian block of nx*ny cells, one
    cell thick in z (empty), the cylinder cut out exactly, freestream inflow
    of 79% N2 and 21% O2 at T_inf and deletion on the four outer sides;
    LB-VHS with Zrot 5 and Bird's Zv.  The defaults are 7 km/s, 220 K,
    1e21 m^-3 air on 10^6 cells with 30 parcels per freestream cell."""
    names = ["N2", "O2", "NO", "N", "O"]
    sp = [make_species(n) for n in names]
    dx, dy = Lx / nx, Ly / ny
    dz = dx
    dt = dt or 0.25 * dx / U
    fn = n_inf * dx * dy * dz / ppc
    outer = patch(abi.HS_PATCH_DELETION, inflow=1)
    empty = patch(abi.HS_PATCH_EMPTY)
    wall = patch(abi.HS_PATCH_DIFFUSE, temperature=T_wall)
    origin = (-x_cyl, -0.5 * Ly, 0.0)
    cap = capacity or int(nx * ny * ppc * 2.0) + (1 << 16)
    n_in = [0.79 * n_inf, 0.21 * n_inf, 0.0, 0.0, 0.0]
    case = Case(nx, ny, 1, origin, (dx, dy, dz), sp, [outer, outer, outer, outer, empty, empty],
                collision_model=abi.HS_COLL_LB_VHS, fn=fn, dt=dt, seed=seed, capacity=cap, flags=flags,
                cyl=(0.0, 0.0, R), cyl_patch=wall, inflow=dict(n=n_in, T=T_inf, U=(U, 0.0, 0.0)),
                name="air_cylinder")
    case.set_reactions(HEATBATH_REACTIONS, names)
    case.names = names
    case.n_in, case.T_inf, case.U_inf = n_in, T_inf, U
    return case


def couette_n2o2(seed=11, capacity=200000, flags=abi.HS_FLAG_SAMPLE):
    """The couette_N2-O2 tutorial: system/blockMeshDict (5x100x1 cells on
    [0,0.05]x[0,1]x[0,0.01], cyclic x, walls y, empty z), system/boundariesDict
    (lowerWall diffuse 2000 K at rest, upperWall diffuse 3000 K moving at
    (300 0 0)), constant/dsmcProperties (LB-VHS N2/O2, Zrot 5,
    inverseZvFormulation pre-2008, nEquivalentParticles 429576195535.3) and
    system/controlDict deltaT 1e-5."""
    sp = [make_species("N2"), make_species("O2")]
    cyc = patch(abi.HS_PATCH_CYCLIC)
    lower = patch(abi.HS_PATCH_DIFFUSE, temperature=2000.0)
    upper = patch(abi.HS_PATCH_DIFFUSE, temperature=3000.0, velocity=(300.0, 0, 0))
    empty = patch(abi.HS_PATCH_EMPTY)
    case = Case(5, 100, 1, (0.0, 0.0, 0.0), (0.01, 0.01, 0.01), sp, [cyc, cyc, lower, upper, empty, empty],
                collision_model=abi.HS_COLL_LB_VHS, fn=429576195535.3, dt=1e-5, seed=seed,
                capacity=capacity, flags=flags, inv_zv_form=0, name="couette_N2-O2")
    # system/dsmcInitialiseDict
    case.init_n = [3.2320494711701479e19, 8.591523910705456e18]
    case.init_T = 2000.0
    return case


def heat_bath_5species(fn=100.0, seed=31, capacity=None, flags=abi.HS_FLAG_SAMPLE, n_side=1, allow_splitting=True):
    """The heatBath-5species tutorial: one (1e-5 m)^3 cell with specular
    walls (system/blockMeshDict, boundariesDict), LB-VHS with every
    relaxation collision number 1, nEquivalentParticles 100, typeIdList
    (N2 O2 NO N O) (constant/dsmcProperties), the twelve QK reactions of
    system/chemReactDict, deltaT 1e-9 (system/controlDict) and a fill of
    N2 1.21753030168e22 and O2 3.23647295384e21 m^-3 at 30,000 K in every
    mode (system/dsmcInitialiseDict).  fn scales the parcel count (the
    tutorial's is 100, about 154,000 parcels); n_side > 1 cuts the same box
    into n_side^3 cells."""
    names = ["N2", "O2", "NO", "N", "O"]
    sp = [make_species(n) for n in names]
    wall = patch(abi.HS_PATCH_SPECULAR)
    L = 1e-5
    d = L / n_side
    n_tot = 1.21753030168e22 + 3.23647295384e21
    cap = capacity or int(1.9 * n_tot * L ** 3 / fn) + 4096
    case = Case(n_side, n_side, n_side, (0.0, 0.0, 0.0), (d, d, d), sp, [wall] * 6,
                collision_model=abi.HS_COLL_LB_VHS, fn=fn, dt=1e-9, seed=seed, capacity=cap, flags=flags,
                z_rot=1.0, z_vib_fixed=1.0, z_el=1.0, name="heatBath-5species")
    case.set_reactions(HEATBATH_REACTIONS, names, allow_splitting=allow_splitting)
    case.names = names
    case.init_n = [1.21753030168e22, 3.23647295384e21, 0.0, 0.0, 0.0]
    case.init_T = 30000.0
    return case


def wall_channel(nx=4, ny=20, L=0.02, n_density=1e21, T_gas=600.0, T_lo=300.0, T_hi=1000.0, u_hi=(0.0, 200.0, 0.0),
                 lo_model=abi.HS_PATCH_DIFFUSE, hi_model=abi.HS_PATCH_DIFFUSE, ppc=50, dt=2e-7, seed=31,
                 collision_model=abi.HS_COLL_VHS, species=None, flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL):
    """Argon between two walls normal to x (xmin at T_lo at rest, xmax at
    T_hi moving with u_hi), cyclic in y, empty in z: the heat-transfer and
    shear set-up that exercises boundaryMeasurements."""
    d = L / nx
    sp = species or [make_species("Ar")]
    cyc = patch(abi.HS_PATCH_CYCLIC)
    empty = patch(abi.HS_PATCH_EMPTY)
    lo = patch(lo_model, temperature=T_lo)
    hi = patch(hi_model, temperature=T_hi, velocity=u_hi)
    fn = n_density * d ** 3 / ppc
    cap = int(nx * ny * ppc * 1.5) + 1024
    case = Case(nx, ny, 1, (0.0, 0.0, 0.0), (d, d, d), sp, [lo, hi, cyc, cyc, empty, empty],
                collision_model=collision_model, fn=fn, dt=dt, seed=seed, capacity=cap, flags=flags,
                name="wall_channel")
    case.n_density, case.T_gas = n_density, T_gas
    return case


def cylinder_batch(count=64, seed=20251203, d_ref=4.17e-10, species=("Ar", "N2")):
    """count independent cylinder cases of uneven size: argon (VHS) and N2
    (LB-VHS, rotation and vibration) in turn, Mach uniform in [5, 25],
    Knudsen number (mean free path over the cylinder's diameter) log-uniform
    in [0.01, 1], square meshes of 50 to 300 cells a side (drawn with density
    proportional to the side, so larger meshes are as common as their share
    of cells) and 4 to 22 parcels per freestream cell: 10k to 2M parcels a
    case, about 37M in all for 64 cases.  The time step keeps the freestream
    at a quarter of a cell per step."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        name = species[k % len(species)]
        sp = make_species(name)
        n_side = int(round(math.sqrt(rng.uniform(50.0 ** 2, 300.0 ** 2))))
        mach = float(rng.uniform(5.0, 25.0))
        kn = float(10.0 ** rng.uniform(-2.0, 0.0))
        ppc = float(rng.uniform(4.0, 22.0))
        n_inf = 1.0 / (math.sqrt(2.0) * math.pi * d_ref * d_ref * kn * 0.2)
        T_inf = 200.0
        gamma = (5.0 + sp.rot_dof) / (3.0 + sp.rot_dof)
        U = mach * math.sqrt(gamma * KB * T_inf / sp.mass)
        dt = 0.25 * (1.0 / n_side) / U
        case = cylinder_flow(nx=n_side, ny=n_side, n_inf=n_inf, T_inf=T_inf, mach=mach, ppc=ppc, dt=dt,
                             seed=seed + k, capacity=int(n_side * n_side * ppc * 2.5) + (1 << 16),
                             flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL, species=sp,
                             collision_model=abi.HS_COLL_LB_VHS if sp.rot_dof > 0 else abi.HS_COLL_VHS)
        case.batch_params = dict(species=name, n_side=n_side, mach=round(mach, 3), kn=round(kn, 5),
                                 ppc=round(ppc, 2))
        out.append(case)
    return out
