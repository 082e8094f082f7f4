"""ctypes binding of the CPU oracle (oracle/_build/libdsmc_oracle.so).

Test infrastructure: only the tests, the smoke check and the benchmark's
CPU-baseline leg use it.
"""
import ctypes as C
import os

import numpy as np

from hystrath_amd import abi
from hystrath_amd.cases import Particles

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# HS_ORACLE_LIB points the tests at another build of the same source (used to
# check, on the CPU, that the tests catch a mistake planted in a copy)
LIB_PATH = os.environ.get("HS_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "libdsmc_oracle.so")

OR_RNG_MOVE, OR_RNG_COLLIDE, OR_RNG_INJECT, OR_RNG_FILL_COUNT, OR_RNG_FILL_GEN, OR_RNG_REMAINDER = range(1, 7)


class or_state(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("capacity", C.c_int64),
        ("p", abi.hs_particles), ("tmp", abi.hs_particles),
        ("sigma_tcr_max", C.POINTER(C.c_double)),
        ("remainder", C.POINTER(C.c_double)),
        ("samples", C.POINTER(C.c_double)),
        ("n_sampled", C.c_int64),
        ("inflow_acc", C.POINTER(C.c_double)),
        ("count", C.POINTER(C.c_int32)),
        ("start", C.POINTER(C.c_int32)),
        ("step", C.c_uint32),
        ("counters", C.c_int64 * 8),
        ("wall", C.POINTER(C.c_double)),
        ("face_flux", C.POINTER(C.c_int64)),
        ("t_overall", C.POINTER(C.c_double)),
    ]


_libs = {}


def lib(path=None):
    """The oracle library at path (default: the single-thread checker),
    loaded once."""
    path = path or LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        P = C.POINTER
        cfgp = P(abi.hs_config)
        stp = P(or_state)
        L.or_philox4x32.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.or_uniform_from.argtypes = [C.c_uint32, C.c_uint32]
        L.or_uniform_from.restype = C.c_double
        for f in ("or_log", "or_exp"):
            getattr(L, f).argtypes = [C.c_double]
            getattr(L, f).restype = C.c_double
        L.or_pow.argtypes = [C.c_double, C.c_double]
        L.or_pow.restype = C.c_double
        L.or_sincos2pi.argtypes = [C.c_double, P(C.c_double), P(C.c_double)]
        L.or_stream_draws.argtypes = [C.c_uint64, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P(C.c_double)]
        L.or_stream_gauss.argtypes = L.or_stream_draws.argtypes
        L.or_sample_slots.argtypes = [cfgp]
        L.or_sample_slots.restype = C.c_int32
        L.or_inflow_face_count.argtypes = [cfgp]
        L.or_inflow_face_count.restype = C.c_int64
        L.or_sigma_tcr.argtypes = [cfgp, C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double), P(C.c_double)]
        L.or_binary_collide.argtypes = [cfgp, C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double),
                                        P(C.c_double), P(C.c_double), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                        C.c_int32, P(C.c_int32), P(C.c_double)]
        L.or_fill.argtypes = [cfgp, stp, P(C.c_double), C.c_double, C.c_double, C.c_double, P(C.c_double), C.c_int32, C.c_uint32]
        L.or_fill_cells.argtypes = [cfgp, stp, P(C.c_double), P(C.c_double), P(C.c_double), C.c_uint32]
        L.or_inject.argtypes = [cfgp, stp]
        L.or_move.argtypes = [cfgp, stp, C.c_int64]
        L.or_sort.argtypes = [cfgp, stp]
        L.or_collide.argtypes = [cfgp, stp, C.c_int32]
        L.or_evolve.argtypes = [cfgp, stp, C.c_int32]
        L.or_wall_faces.argtypes = [cfgp, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double)]
        L.or_wall_faces.restype = C.c_int64
        L.or_arc_in_rect.argtypes = [C.c_double] * 7 + [P(C.c_double)]
        L.or_arc_in_rect.restype = C.c_double
        L.or_face_count.argtypes = [cfgp]
        L.or_face_count.restype = C.c_int64
        L.or_ntc_pairs.argtypes = [cfgp, C.c_int64, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double),
                                   C.c_uint32, C.c_int64, P(C.c_int32), P(C.c_int32)]
        L.or_react_pairs.argtypes = [cfgp, C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double),
                                     P(C.c_double), P(C.c_double), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                     C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                     P(C.c_int32)]
        L.or_threads.restype = C.c_int
        L.or_set_threads.argtypes = [C.c_int]
        _libs[path] = L
    return _libs[path]


def dptr(a):
    return abi.ptr(a, C.c_double)


def iptr(a):
    return abi.ptr(a, C.c_int32)


# ---- primitives ----------------------------------------------------------
def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32(c, k, o)
    return list(o)


def stream(seed, purpose, ident, step, sub, n):
    out = np.zeros(n, np.float64)
    lib().or_stream_draws(seed, purpose, ident, step, sub, n, dptr(out))
    return out


def stream_gauss(seed, purpose, ident, step, sub, n):
    out = np.zeros(n, np.float64)
    lib().or_stream_gauss(seed, purpose, ident, step, sub, n, dptr(out))
    return out


def sincos2pi(u):
    s, c = C.c_double(), C.c_double()
    lib().or_sincos2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


def sigma_tcr(case, tp, tq, up, uq):
    n = len(tp)
    out = np.zeros(n)
    lib().or_sigma_tcr(C.byref(case.cfg), n, iptr(np.ascontiguousarray(tp, np.int32)),
                       iptr(np.ascontiguousarray(tq, np.int32)), dptr(np.ascontiguousarray(up, np.float64)),
                       dptr(np.ascontiguousarray(uq, np.float64)), dptr(out))
    return out


def binary_collide(case, tp, tq, up, uq, uniforms, erot_p=None, erot_q=None, vib_p=None, vib_q=None, t_macro=None):
    """Returns (up', uq', erot_p', erot_q', vib_p', vib_q', used) in FP64.
    t_macro: each pair's cell overall temperature (inverseZvFormulation 2008)."""
    n = len(tp)
    up = np.array(up, np.float64).reshape(n, 3).copy()
    uq = np.array(uq, np.float64).reshape(n, 3).copy()
    ep = None if erot_p is None else np.array(erot_p, np.float64).copy()
    eq = None if erot_q is None else np.array(erot_q, np.float64).copy()
    vp = None if vib_p is None else np.array(vib_p, np.int32).reshape(n, abi.HS_MAX_VIB_MODES).copy()
    vq = None if vib_q is None else np.array(vib_q, np.int32).reshape(n, abi.HS_MAX_VIB_MODES).copy()
    uni = np.ascontiguousarray(uniforms, np.float64)
    nu = uni.shape[1]
    used = np.zeros(n, np.int32)
    rc = lib().or_binary_collide(C.byref(case.cfg), n, iptr(np.ascontiguousarray(tp, np.int32)),
                                 iptr(np.ascontiguousarray(tq, np.int32)), dptr(up), dptr(uq), dptr(ep), dptr(eq),
                                 iptr(vp), iptr(vq), dptr(uni), nu, iptr(used),
                                 None if t_macro is None else dptr(np.ascontiguousarray(t_macro, np.float64)))
    return up, uq, ep, eq, vp, vq, used, rc


def react_pairs(case, tp, tq, up, uq, erot_p, erot_q, vib_p, vib_q, uniforms):
    """The accepted collision of each pair, its QK reaction first (the
    oracle's or_react_pairs).  Returns a dict of the pairs' new states and
    the products."""
    n = len(tp)
    m = abi.HS_MAX_VIB_MODES
    tp = np.array(tp, np.int32).copy()
    tq = np.array(tq, np.int32).copy()
    up = np.array(up, np.float64).reshape(n, 3).copy()
    uq = np.array(uq, np.float64).reshape(n, 3).copy()
    ep = np.array(erot_p, np.float64).copy()
    eq = np.array(erot_q, np.float64).copy()
    vp = np.array(vib_p, np.int32).reshape(n, m).copy()
    vq = np.array(vib_q, np.int32).reshape(n, m).copy()
    uni = np.ascontiguousarray(uniforms, np.float64).reshape(n, -1)
    used = np.zeros(n, np.int32)
    reacted = np.zeros(n, np.int32)
    ptype = np.zeros(n, np.int32)
    pu = np.zeros((n, 3), np.float64)
    pfrom = np.zeros(n, np.int32)
    rc = lib().or_react_pairs(C.byref(case.cfg), n, iptr(tp), iptr(tq), dptr(up), dptr(uq), dptr(ep), dptr(eq),
                              iptr(vp), iptr(vq), dptr(uni), uni.shape[1], iptr(used), iptr(reacted), iptr(ptype),
                              dptr(pu), iptr(pfrom))
    return dict(tp=tp, tq=tq, up=up, uq=uq, erot_p=ep, erot_q=eq, vib_p=vp, vib_q=vq, used=used, reacted=reacted,
                prod_type=ptype, prod_u=pu, prod_from_q=pfrom, rc=rc)


def ntc_pairs(case, cell, pos, step, ncand):
    """The (P, Q) candidate pairs the oracle's NTC draws for candidates
    0..ncand-1 of `cell` holding parcels at pos[k] (sorted order)."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    px, py, pz = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    p = np.zeros(ncand, np.int32)
    q = np.zeros(ncand, np.int32)
    rc = lib().or_ntc_pairs(C.byref(case.cfg), cell, len(pos), dptr(px), dptr(py), dptr(pz), step, ncand,
                            iptr(p), iptr(q))
    assert rc == 0
    return p, q


# ---- whole-state driver --------------------------------------------------
class OracleState:
    """An or_state with numpy storage, mirroring the engine's state."""

    def __init__(self, case, capacity=None, lib_path=None):
        self.case = case
        self.L = lib(lib_path)
        cap = capacity or case.cfg.capacity
        self.cap = cap
        self.P = Particles.for_case(case, cap)
        self.T = Particles.for_case(case, cap)
        nc = case.ncells
        self.sigma = np.zeros(nc)
        self.rem = np.zeros(nc)
        self.nslots = self.L.or_sample_slots(C.byref(case.cfg))
        self.samples = np.zeros(self.nslots * nc)
        nf = self.L.or_inflow_face_count(C.byref(case.cfg))
        self.nfaces = nf
        self.inflow = np.zeros(max(case.cfg.n_species * nf, 1)) if nf else None
        self.count = np.zeros(nc, np.int32)
        self.start = np.zeros(nc + 1, np.int32)
        self.n_wall = self.L.or_wall_faces(C.byref(case.cfg), None, None, None, None)
        self.wall = np.zeros(max(case.cfg.n_species * abi.HS_WS_COUNT * self.n_wall, 1))
        self.n_faces = self.L.or_face_count(C.byref(case.cfg))
        self.flux = np.zeros((case.cfg.n_species, self.n_faces), np.int64)
        s = or_state()
        s.n = 0
        s.capacity = cap
        s.p = self.P.struct()
        s.tmp = self.T.struct()
        s.sigma_tcr_max = dptr(self.sigma)
        s.remainder = dptr(self.rem)
        s.samples = dptr(self.samples)
        s.inflow_acc = dptr(self.inflow) if nf else None
        s.count = iptr(self.count)
        s.start = iptr(self.start)
        s.wall = dptr(self.wall)
        s.face_flux = self.flux.ctypes.data_as(C.POINTER(C.c_int64))
        self.t_overall = np.zeros(nc)
        s.t_overall = dptr(self.t_overall)
        self.s = s

    def set_cell_temperature(self, t_overall):
        self.t_overall[:] = np.asarray(t_overall, np.float64)

    @property
    def n(self):
        return self.s.n

    @property
    def step(self):
        return self.s.step

    @step.setter
    def step(self, v):
        self.s.step = v

    def counters(self):
        return list(self.s.counters)

    def fill(self, n_density, T, u=(0.0, 0.0, 0.0), Trot=None, Tvib=None, exact=False, fill_id=0):
        nd = np.zeros(abi.HS_MAX_SPECIES)
        nd[: len(n_density)] = n_density
        uu = np.array(u, np.float64)
        self.L.or_fill(C.byref(self.case.cfg), C.byref(self.s), dptr(nd), T, Trot if Trot is not None else T,
                      Tvib if Tvib is not None else T, dptr(uu), 1 if exact else 0, fill_id)

    def fill_cells(self, n_cell, u_cell, t_cell, fill_id=0):
        n = np.ascontiguousarray(n_cell, np.float64).ravel()
        u = np.ascontiguousarray(u_cell, np.float64).ravel()
        t = np.ascontiguousarray(t_cell, np.float64).ravel()
        self.L.or_fill_cells(C.byref(self.case.cfg), C.byref(self.s), dptr(n), dptr(u), dptr(t), fill_id)

    def load(self, d, sigma=None, rem=None):
        self.s.n = self.P.load(d)
        if sigma is not None:
            self.sigma[:] = sigma
        if rem is not None:
            self.rem[:] = rem

    def particles(self):
        return self.P.view(self.s.n)

    def inject(self):
        self.L.or_inject(C.byref(self.case.cfg), C.byref(self.s))

    def move(self, n_old=None):
        self.L.or_move(C.byref(self.case.cfg), C.byref(self.s), self.s.n if n_old is None else n_old)

    def sort(self):
        self.L.or_sort(C.byref(self.case.cfg), C.byref(self.s))

    def collide(self, sample=True):
        self.L.or_collide(C.byref(self.case.cfg), C.byref(self.s), 1 if sample else 0)

    def evolve(self, nsteps):
        self.L.or_evolve(C.byref(self.case.cfg), C.byref(self.s), nsteps)

    def sample_array(self):
        return self.samples.reshape(self.nslots, self.case.ncells)

    def wall_array(self):
        rows = self.case.cfg.n_species * abi.HS_WS_COUNT
        return self.wall[: rows * self.n_wall].reshape(rows, self.n_wall)


def wall_faces(case):
    n = lib().or_wall_faces(C.byref(case.cfg), None, None, None, None)
    side = np.zeros(n, np.int32)
    cell = np.zeros(n, np.int32)
    area = np.zeros(n)
    normal = np.zeros((n, 3))
    lib().or_wall_faces(C.byref(case.cfg), iptr(side), iptr(cell), dptr(area), dptr(normal))
    return dict(side=side, cell=cell, area=area, normal=normal)


def arc_in_rect(cx, cy, r, x0, x1, y0, y1):
    nrm = np.zeros(2)
    length = lib().or_arc_in_rect(cx, cy, r, x0, x1, y0, y1, dptr(nrm))
    return length, nrm
