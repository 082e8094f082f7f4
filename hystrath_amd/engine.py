"""Python handle on the MI355X engine through its C ABI (include/hystrath_dsmc.h).

This is the host-side mirror of dsmcCloud's per-timestep API used by the
tests, the benchmark and the case driver.  The engine has no CPU path: if the
HIP library is missing or no GPU is visible, construction raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .cases import Particles

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhystrath_dsmc.so")

# every entry point the header declares (checked by the ABI test)
EXPORTS = [
    "hs_dsmc_sample_slots", "hs_dsmc_create", "hs_dsmc_destroy", "hs_dsmc_last_error", "hs_dsmc_create_error",
    "hs_dsmc_device_count", "hs_dsmc_upload_particles", "hs_dsmc_download_particles", "hs_dsmc_particle_count",
    "hs_dsmc_set_cell_state", "hs_dsmc_get_cell_state", "hs_dsmc_set_cell_temperature", "hs_dsmc_inflow_face_count",
    "hs_dsmc_set_inflow_accumulators", "hs_dsmc_get_inflow_accumulators", "hs_dsmc_set_step", "hs_dsmc_get_step",
    "hs_dsmc_fill", "hs_dsmc_fill_cells", "hs_dsmc_fill_noalloc", "hs_dsmc_evolve", "hs_dsmc_inject", "hs_dsmc_move", "hs_dsmc_sort", "hs_dsmc_collide",
    "hs_dsmc_advance_step", "hs_dsmc_synchronize", "hs_dsmc_read_samples", "hs_dsmc_reset_samples", "hs_dsmc_set_samples",
    "hs_dsmc_read_counters", "hs_dsmc_reset_counters", "hs_dsmc_kernel_times", "hs_dsmc_binary_collide",
    "hs_dsmc_sigma_tcr", "hs_dsmc_device_copy", "hs_dsmc_set_flags", "hs_dsmc_wall_face_count",
    "hs_dsmc_wall_faces", "hs_dsmc_read_wall_samples", "hs_dsmc_face_count", "hs_dsmc_read_face_flux",
    "hs_dsmc_set_tiles", "hs_dsmc_get_tiles", "hs_dsmc_autotune", "hs_dsmc_evolve_batch",
    "hs_dsmc_batch_create", "hs_dsmc_batch_evolve", "hs_dsmc_batch_destroy", "hs_dsmc_batch_last_error",
]

STAGES = ["inject", "move", "scan", "scatter", "collide", "gather"]
COUNTERS = ["candidates", "collisions", "deleted", "injected", "wall_hits", "track_overflow", "tile_overflow",
            "capacity_overflow", "particle_steps"]

_lib = None


class EngineError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"HIP engine library not built: {path} (run `python -m hystrath_amd.build`)")
    L = C.CDLL(path)
    P = C.POINTER
    eng = C.c_void_p
    cfgp = P(abi.hs_config)
    sig = {
        "hs_dsmc_sample_slots": ([cfgp], C.c_int32),
        "hs_dsmc_create": ([cfgp, P(eng)], C.c_int),
        "hs_dsmc_destroy": ([eng], None),
        "hs_dsmc_last_error": ([eng], C.c_char_p),
        "hs_dsmc_create_error": ([], C.c_char_p),
        "hs_dsmc_device_count": ([P(C.c_int)], C.c_int),
        "hs_dsmc_upload_particles": ([eng, P(abi.hs_particles), C.c_int64], C.c_int),
        "hs_dsmc_download_particles": ([eng, P(abi.hs_particles), C.c_int64, P(C.c_int64)], C.c_int),
        "hs_dsmc_particle_count": ([eng, P(C.c_int64)], C.c_int),
        "hs_dsmc_set_cell_state": ([eng, P(C.c_double), P(C.c_double)], C.c_int),
        "hs_dsmc_set_cell_temperature": ([eng, P(C.c_double)], C.c_int),
        "hs_dsmc_get_cell_state": ([eng, P(C.c_double), P(C.c_double)], C.c_int),
        "hs_dsmc_inflow_face_count": ([eng, P(C.c_int64)], C.c_int),
        "hs_dsmc_set_inflow_accumulators": ([eng, P(C.c_double)], C.c_int),
        "hs_dsmc_get_inflow_accumulators": ([eng, P(C.c_double)], C.c_int),
        "hs_dsmc_set_step": ([eng, C.c_uint32], C.c_int),
        "hs_dsmc_get_step": ([eng, P(C.c_uint32)], C.c_int),
        "hs_dsmc_fill": ([eng, P(C.c_double), C.c_double, C.c_double, C.c_double, P(C.c_double), C.c_int32,
                          C.c_uint32], C.c_int),
        "hs_dsmc_fill_cells": ([eng, P(C.c_double), P(C.c_double), P(C.c_double), C.c_uint32], C.c_int),
        "hs_dsmc_fill_noalloc": ([eng, P(C.c_double), C.c_double, C.c_double, C.c_double, P(C.c_double),
                                  C.c_int32, C.c_uint32], C.c_int),
        "hs_dsmc_evolve": ([eng, C.c_int32], C.c_int),
        "hs_dsmc_inject": ([eng], C.c_int),
        "hs_dsmc_move": ([eng], C.c_int),
        "hs_dsmc_sort": ([eng], C.c_int),
        "hs_dsmc_collide": ([eng, C.c_int32], C.c_int),
        "hs_dsmc_advance_step": ([eng], C.c_int),
        "hs_dsmc_synchronize": ([eng], C.c_int),
        "hs_dsmc_read_samples": ([eng, P(C.c_double), C.c_int64, P(C.c_int64)], C.c_int),
        "hs_dsmc_reset_samples": ([eng], C.c_int),
        "hs_dsmc_set_samples": ([eng, P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, C.c_int64], C.c_int),
        "hs_dsmc_read_counters": ([eng, P(C.c_int64), C.c_int32], C.c_int),
        "hs_dsmc_reset_counters": ([eng], C.c_int),
        "hs_dsmc_kernel_times": ([eng, P(C.c_double), P(C.c_int64), C.c_int32], C.c_int),
        "hs_dsmc_binary_collide": ([eng, C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double),
                                    P(C.c_double), P(C.c_double), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                    C.c_int32, P(C.c_int32)], C.c_int),
        "hs_dsmc_sigma_tcr": ([eng, C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double),
                               P(C.c_double)], C.c_int),
        "hs_dsmc_device_copy": ([C.c_int64, C.c_int32, P(C.c_double)], C.c_int),
        "hs_dsmc_set_flags": ([eng, C.c_int32], C.c_int),
        "hs_dsmc_set_tiles": ([eng, C.c_int32, C.c_int32], C.c_int),
        "hs_dsmc_evolve_batch": ([P(C.c_void_p), C.c_int32, C.c_int32], C.c_int),
        "hs_dsmc_batch_create": ([P(C.c_void_p), C.c_int32, P(C.c_void_p)], C.c_int),
        "hs_dsmc_batch_evolve": ([C.c_void_p, C.c_int32], C.c_int),
        "hs_dsmc_batch_destroy": ([C.c_void_p], None),
        "hs_dsmc_batch_last_error": ([C.c_void_p], C.c_char_p),
        "hs_dsmc_get_tiles": ([eng, P(C.c_int32), P(C.c_int32)], C.c_int),
        "hs_dsmc_autotune": ([eng, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_double)], C.c_int),
        "hs_dsmc_wall_face_count": ([eng, P(C.c_int64)], C.c_int),
        "hs_dsmc_wall_faces": ([eng, P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double)], C.c_int),
        "hs_dsmc_read_wall_samples": ([eng, P(C.c_double), C.c_int64, P(C.c_int64)], C.c_int),
        "hs_dsmc_face_count": ([eng, P(C.c_int64)], C.c_int),
        "hs_dsmc_read_face_flux": ([eng, P(C.c_int64), C.c_int64, P(C.c_int64)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def device_count():
    n = C.c_int(0)
    load_library().hs_dsmc_device_count(C.byref(n))
    return n.value


def device_copy_ms(nbytes, iters=20):
    """Mean ms of a plain device copy of nbytes (read nbytes + write nbytes)."""
    L = load_library()
    ms = C.c_double()
    rc = L.hs_dsmc_device_copy(nbytes, iters, C.byref(ms))
    if rc != 0:
        raise EngineError(f"device copy failed ({rc}): {L.hs_dsmc_create_error().decode()}")
    return ms.value


def evolve_batch(engines, nsteps=1):
    """hs_dsmc_evolve_batch: nsteps of every engine, interleaved on the device."""
    L = load_library()
    hs = (C.c_void_p * len(engines))(*[e.h for e in engines])
    rc = L.hs_dsmc_evolve_batch(hs, len(engines), nsteps)
    if rc != 0:
        msgs = [L.hs_dsmc_last_error(e.h).decode() for e in engines if e.h]
        raise EngineError(f"evolve_batch failed ({rc}): {'; '.join(m for m in msgs if m) or L.hs_dsmc_create_error().decode()}")


class Batch:
    """hs_dsmc_batch: the engines stepped with one launch per stage over all
    of them.  The engines must outlive the batch."""

    def __init__(self, engines):
        self.L = load_library()
        self.engines = list(engines)
        hs = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        rc = self.L.hs_dsmc_batch_create(hs, len(self.engines), C.byref(h))
        if rc != 0:
            raise EngineError(f"batch_create failed ({rc}): {self.L.hs_dsmc_batch_last_error(None).decode()}")
        self.h = h

    def evolve(self, nsteps=1):
        rc = self.L.hs_dsmc_batch_evolve(self.h, nsteps)
        if rc != 0:
            raise EngineError(f"batch_evolve failed ({rc}): {self.L.hs_dsmc_batch_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.hs_dsmc_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _d(a):
    return abi.ptr(a, C.c_double)


def _i(a):
    return abi.ptr(a, C.c_int32)


class Engine:
    """One engine on the current HIP device (dsmcCloud's particle path)."""

    def __init__(self, case):
        self.L = load_library()
        self.case = case
        h = C.c_void_p()
        rc = self.L.hs_dsmc_create(C.byref(case.cfg), C.byref(h))
        if rc != 0:
            raise EngineError(f"hs_dsmc_create failed ({rc}): {self.L.hs_dsmc_create_error().decode()}")
        self.h = h
        self.nslots = self.L.hs_dsmc_sample_slots(C.byref(case.cfg))
        nf = C.c_int64()
        self.L.hs_dsmc_inflow_face_count(h, C.byref(nf))
        self.nfaces = nf.value

    def _ck(self, rc, what):
        if rc != 0:
            raise EngineError(f"{what} failed ({rc}): {self.L.hs_dsmc_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.hs_dsmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state ------------------------------------------------------------
    def upload(self, d):
        """d: dict of SoA arrays as returned by Particles.view()."""
        n = len(d["px"])
        p = abi.hs_particles()
        keep = {}
        for k, ct in (("px", C.c_double), ("py", C.c_double), ("pz", C.c_double), ("ux", C.c_float),
                      ("uy", C.c_float), ("uz", C.c_float), ("cell", C.c_int32)):
            keep[k] = np.ascontiguousarray(d[k])
            setattr(p, k, abi.ptr(keep[k], ct))
        if "erot" in d:
            keep["erot"] = np.ascontiguousarray(d["erot"], np.float32)
            p.erot = abi.ptr(keep["erot"], C.c_float)
        if "vib" in d:
            keep["vib"] = np.ascontiguousarray(d["vib"], np.uint16)
            p.vib = abi.ptr(keep["vib"], C.c_uint16)
        if "type" in d:
            keep["type"] = np.ascontiguousarray(d["type"], np.uint8)
            p.type = abi.ptr(keep["type"], C.c_uint8)
        self._ck(self.L.hs_dsmc_upload_particles(self.h, C.byref(p), n), "upload")

    def count(self):
        n = C.c_int64()
        self._ck(self.L.hs_dsmc_particle_count(self.h, C.byref(n)), "particle_count")
        return n.value

    def download(self):
        n = self.count()
        buf = Particles(max(n, 1), self.case.vib_modes, self.case.has_rot, self.case.cfg.n_species > 1)
        p = buf.struct()
        got = C.c_int64()
        self._ck(self.L.hs_dsmc_download_particles(self.h, C.byref(p), max(n, 1), C.byref(got)), "download")
        d = buf.view(got.value)
        if buf.vib is not None:  # the ABI packs vib planes with stride n
            d["vib"] = buf.vib[: buf.vib_modes * got.value].reshape(buf.vib_modes, got.value).copy()
        return d

    def set_cell_state(self, sigma=None, rem=None):
        s = None if sigma is None else np.ascontiguousarray(sigma, np.float64)
        r = None if rem is None else np.ascontiguousarray(rem, np.float64)
        self._ck(self.L.hs_dsmc_set_cell_state(self.h, _d(s), _d(r)), "set_cell_state")

    def set_cell_temperature(self, t_overall):
        """Each cell's overall temperature for inverseZvFormulation 2008."""
        t = np.ascontiguousarray(t_overall, np.float64)
        if t.shape != (self.case.ncells,):
            raise EngineError(f"set_cell_temperature: {t.shape} values for {self.case.ncells} cells")
        self._ck(self.L.hs_dsmc_set_cell_temperature(self.h, _d(t)), "set_cell_temperature")

    def cell_state(self):
        nc = self.case.ncells
        s, r = np.zeros(nc), np.zeros(nc)
        self._ck(self.L.hs_dsmc_get_cell_state(self.h, _d(s), _d(r)), "get_cell_state")
        return s, r

    def set_inflow(self, acc):
        a = np.ascontiguousarray(acc, np.float64)
        self._ck(self.L.hs_dsmc_set_inflow_accumulators(self.h, _d(a)), "set_inflow")

    def inflow(self):
        a = np.zeros(max(self.case.cfg.n_species * self.nfaces, 1))
        self._ck(self.L.hs_dsmc_get_inflow_accumulators(self.h, _d(a)), "get_inflow")
        return a

    @property
    def step(self):
        s = C.c_uint32()
        self._ck(self.L.hs_dsmc_get_step(self.h, C.byref(s)), "get_step")
        return s.value

    @step.setter
    def step(self, v):
        self._ck(self.L.hs_dsmc_set_step(self.h, v), "set_step")

    def fill(self, n_density, T, u=(0.0, 0.0, 0.0), Trot=None, Tvib=None, exact=False, fill_id=0):
        nd = np.zeros(abi.HS_MAX_SPECIES)
        nd[: len(n_density)] = n_density
        uu = np.array(u, np.float64)
        self._ck(self.L.hs_dsmc_fill(self.h, _d(nd), T, T if Trot is None else Trot, T if Tvib is None else Tvib,
                                     _d(uu), 1 if exact else 0, fill_id), "fill")

    def fill_cells(self, n_cell, u_cell, t_cell, fill_id=0):
        """Fill with per-cell number densities [species, cell], velocities
        [3, cell] and temperatures [cell]."""
        n = np.ascontiguousarray(n_cell, np.float64).ravel()
        u = np.ascontiguousarray(u_cell, np.float64).ravel()
        t = np.ascontiguousarray(t_cell, np.float64).ravel()
        nc, ns = self.case.ncells, self.case.cfg.n_species
        if n.size != ns * nc or u.size != 3 * nc or t.size != nc:
            raise EngineError("fill_cells: arrays must be [species, cells], [3, cells] and [cells]")
        self._ck(self.L.hs_dsmc_fill_cells(self.h, _d(n), _d(u), _d(t), fill_id), "fill_cells")

    # ---- stepping ----------------------------------------------------------
    def evolve(self, nsteps=1):
        self._ck(self.L.hs_dsmc_evolve(self.h, nsteps), "evolve")

    def inject(self):
        self._ck(self.L.hs_dsmc_inject(self.h), "inject")

    def move(self):
        self._ck(self.L.hs_dsmc_move(self.h), "move")

    def sort(self):
        self._ck(self.L.hs_dsmc_sort(self.h), "sort")

    def collide(self, sample=True):
        self._ck(self.L.hs_dsmc_collide(self.h, 1 if sample else 0), "collide")

    def set_flags(self, flags):
        self._ck(self.L.hs_dsmc_set_flags(self.h, flags), "set_flags")
        self.case.cfg.flags = flags

    def set_tiles(self, gather_parcels=0, collide_parcels=0):
        """Parcels per gather / collide workgroup; 0 chooses from the population."""
        self._ck(self.L.hs_dsmc_set_tiles(self.h, gather_parcels, collide_parcels), "set_tiles")

    def tiles(self):
        g, c = C.c_int32(), C.c_int32()
        self._ck(self.L.hs_dsmc_get_tiles(self.h, C.byref(g), C.byref(c)), "get_tiles")
        return g.value, c.value

    def autotune(self, steps_per_trial=8):
        """Times the candidate tilings on the live state (the steps are steps
        of the run); returns (gather, collide, ms per step) of the fastest."""
        g, c, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._ck(self.L.hs_dsmc_autotune(self.h, steps_per_trial, C.byref(g), C.byref(c), C.byref(ms)), "autotune")
        return g.value, c.value, ms.value

    def synchronize(self):
        self._ck(self.L.hs_dsmc_synchronize(self.h), "synchronize")

    # ---- measurements -------------------------------------------------------
    def samples(self):
        out = np.zeros(self.nslots * self.case.ncells)
        ns = C.c_int64()
        self._ck(self.L.hs_dsmc_read_samples(self.h, _d(out), out.size, C.byref(ns)), "read_samples")
        return out.reshape(self.nslots, self.case.ncells), ns.value

    def reset_samples(self):
        self._ck(self.L.hs_dsmc_reset_samples(self.h), "reset_samples")

    def set_samples(self, cell_sums, nsteps, wall_sums=None):
        """Restores cumulative cell (and wall) sums and their step count."""
        c = np.ascontiguousarray(cell_sums, np.float64).ravel()
        w = None if wall_sums is None else np.ascontiguousarray(wall_sums, np.float64).ravel()
        self._ck(self.L.hs_dsmc_set_samples(self.h, _d(c), c.size, _d(w) if w is not None else None,
                                            0 if w is None else w.size, nsteps), "set_samples")

    def wall_faces(self):
        """dict of side, cell, area, normal[n,3] of the wall faces"""
        n = C.c_int64()
        self._ck(self.L.hs_dsmc_wall_face_count(self.h, C.byref(n)), "wall_face_count")
        n = n.value
        side = np.zeros(n, np.int32)
        cell = np.zeros(n, np.int32)
        area = np.zeros(n)
        normal = np.zeros((n, 3))
        self._ck(self.L.hs_dsmc_wall_faces(self.h, _i(side), _i(cell), _d(area), _d(normal)), "wall_faces")
        return dict(side=side, cell=cell, area=area, normal=normal)

    def face_flux(self):
        """([species, face] int64 parcel fluxes of the face tracker, sampled steps)"""
        n = C.c_int64()
        self._ck(self.L.hs_dsmc_face_count(self.h, C.byref(n)), "face_count")
        out = np.zeros((self.case.cfg.n_species, n.value), np.int64)
        ns = C.c_int64()
        self._ck(self.L.hs_dsmc_read_face_flux(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)), out.size,
                                               C.byref(ns)), "read_face_flux")
        return out, ns.value

    def wall_samples(self):
        """([species*HS_WS_COUNT + slot, face] sums, sampled steps)"""
        n = C.c_int64()
        self._ck(self.L.hs_dsmc_wall_face_count(self.h, C.byref(n)), "wall_face_count")
        rows = self.case.cfg.n_species * abi.HS_WS_COUNT
        out = np.zeros(rows * max(n.value, 1))
        ns = C.c_int64()
        self._ck(self.L.hs_dsmc_read_wall_samples(self.h, _d(out), out.size, C.byref(ns)), "read_wall_samples")
        return out[: rows * n.value].reshape(rows, n.value), ns.value

    def counters(self):
        out = np.zeros(len(COUNTERS), np.int64)
        self._ck(self.L.hs_dsmc_read_counters(self.h, abi.ptr(out, C.c_int64), len(COUNTERS)), "read_counters")
        return dict(zip(COUNTERS, out.tolist()))

    def reset_counters(self):
        self._ck(self.L.hs_dsmc_reset_counters(self.h), "reset_counters")

    def kernel_times(self, reset=False):
        ms = np.zeros(len(STAGES))
        nl = np.zeros(len(STAGES), np.int64)
        self._ck(self.L.hs_dsmc_kernel_times(self.h, _d(ms), abi.ptr(nl, C.c_int64), len(STAGES)), "kernel_times")
        if reset:
            self._ck(self.L.hs_dsmc_kernel_times(self.h, None, None, -1), "kernel_times reset")
        return {k: (float(ms[i]), int(nl[i])) for i, k in enumerate(STAGES)}

    # ---- test hooks --------------------------------------------------------
    def binary_collide(self, tp, tq, up, uq, uniforms, erot_p=None, erot_q=None, vib_p=None, vib_q=None):
        n = len(tp)
        up = np.array(up, np.float64).reshape(n, 3).copy()
        uq = np.array(uq, np.float64).reshape(n, 3).copy()
        ep = None if erot_p is None else np.array(erot_p, np.float64).copy()
        eq = None if erot_q is None else np.array(erot_q, np.float64).copy()
        vp = None if vib_p is None else np.array(vib_p, np.int32).reshape(n, abi.HS_MAX_VIB_MODES).copy()
        vq = None if vib_q is None else np.array(vib_q, np.int32).reshape(n, abi.HS_MAX_VIB_MODES).copy()
        uni = np.ascontiguousarray(uniforms, np.float64)
        used = np.zeros(n, np.int32)
        rc = self.L.hs_dsmc_binary_collide(self.h, n, _i(np.ascontiguousarray(tp, np.int32)),
                                           _i(np.ascontiguousarray(tq, np.int32)), _d(up), _d(uq), _d(ep), _d(eq),
                                           _i(vp), _i(vq), _d(uni), uni.shape[1], _i(used))
        return up, uq, ep, eq, vp, vq, used, rc

    def sigma_tcr(self, tp, tq, up, uq):
        n = len(tp)
        out = np.zeros(n)
        self._ck(self.L.hs_dsmc_sigma_tcr(self.h, n, _i(np.ascontiguousarray(tp, np.int32)),
                                          _i(np.ascontiguousarray(tq, np.int32)),
                                          _d(np.ascontiguousarray(up, np.float64)),
                                          _d(np.ascontiguousarray(uq, np.float64)), _d(out)), "sigma_tcr")
        return out
