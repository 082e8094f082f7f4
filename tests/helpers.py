"""Shared helpers for the parity tests: paired engine/oracle states and
exact comparisons."""
import numpy as np

from hystrath_amd import abi, cases

FIELDS = ("px", "py", "pz", "ux", "uy", "uz", "cell", "erot", "vib", "type")


def assert_same_particles(got, want, label=""):
    assert len(got["px"]) == len(want["px"]), f"{label}: count {len(got['px'])} != {len(want['px'])}"
    for k in FIELDS:
        if k not in want and k not in got:
            continue
        a, b = np.asarray(got[k]), np.asarray(want[k])
        assert a.shape == b.shape, f"{label}:{k} shape {a.shape} != {b.shape}"
        if a.dtype.kind == "f":
            bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        else:
            bad = a != b
        if bad.any():
            i = int(np.argmax(bad.reshape(-1)))
            raise AssertionError(f"{label}:{k} differs at {int(bad.sum())} entries, first flat index {i}: "
                                 f"{a.reshape(-1)[i]!r} vs {b.reshape(-1)[i]!r}")


def small_box(n_side=6, ppc=20, seed=99, collision_model=abi.HS_COLL_VHS, flags=abi.HS_FLAG_SAMPLE):
    return cases.periodic_box(n_side=n_side, L=0.1 * n_side / 20, ppc=ppc, seed=seed,
                              collision_model=collision_model, flags=flags)


def small_cylinder(n=60, ppc=10, seed=5, flags=abi.HS_FLAG_SAMPLE):
    return cases.cylinder_flow(nx=n, ny=n, ppc=ppc, dt=2.5e-7 * 400 / n, seed=seed, flags=flags)


def fill_box(obj, case):
    nd = case.n_density
    obj.fill([nd / 2], 300.0, exact=True, fill_id=0)
    obj.fill([nd / 2], case.T_hot, exact=True, fill_id=1)


def fill_cylinder(obj, case):
    obj.fill([case.n_inf], case.T_inf, u=(case.U_inf, 0.0, 0.0), fill_id=0)


def fill_couette(obj, case):
    obj.fill(case.init_n, case.init_T, fill_id=0)
