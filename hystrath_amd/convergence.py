"""Time to converged fields: how long a run takes until its running-average
fields are within a tolerance of a long average.

The reference answers "how long to sample" by a fixed schedule
(system/fieldPropertiesDict: timeOption / resetAtOutputUntilTime, read by
timeDataMeas.C:71-104) and the user's decision; this module measures it.
Error is the relative L2 norm over the cells the long average sampled with at least
`min_parcels` parcels on average:

  density, translational and internal temperature:  ||f - f_long|| / ||f_long||
  velocity components:                               ||U_a - U_a,long|| / ||U_long||

(velocity components are scaled by the long average's speed, not by their own norm,
so that a component that is zero over most of the domain does not measure
noise against nothing).
"""
import numpy as np

from . import fields

FIELDS = ("rho", "Ux", "Uy", "Ttra", "Tint")


def field_set(samples, nsteps, case):
    """The compared fields of the mixture from cumulative sample sums:
    number density, the x and y velocity components, the translational
    temperature and (for molecules) the internal temperature, the
    degree-of-freedom weighted mean of T_rot and T_vib."""
    f = fields.vol_fields(samples, nsteps, case)
    out = {"rho": f["rhoN"], "Ux": f["U"][0], "Uy": f["U"][1], "Ttra": f["Ttra"], "nmean": f["dsmcNMean"]}
    if case.has_rot or case.vib_modes:
        S = fields.species_sums(samples, case)
        zrot = sum(case.species[s].rot_dof * S[s]["N"] for s in range(case.cfg.n_species))
        n = sum(S[s]["N"] for s in range(case.cfg.n_species))
        zr = np.where(n > 0, zrot / np.where(n > 0, n, 1.0), 0.0)
        zv = f["zetaVib"]
        den = zr + zv
        out["Tint"] = np.where(den > 1e-15, (zr * f["Trot"] + zv * f["Tvib"]) / np.where(den > 1e-15, den, 1.0), 0.0)
    return out


def errors(run, long_avg, min_parcels=1.0):
    """Per-field relative L2 errors of run against the long average (dicts from
    field_set)."""
    mask = long_avg["nmean"] >= min_parcels
    speed = np.sqrt(np.sum(long_avg["Ux"][mask] ** 2 + long_avg["Uy"][mask] ** 2))
    out = {}
    for k in FIELDS:
        if k not in long_avg or k not in run:
            continue
        diff = np.linalg.norm(run[k][mask] - long_avg[k][mask])
        if k in ("Ux", "Uy"):
            out[k] = float(diff / speed) if speed > 0 else float("inf")
        else:
            ref = np.linalg.norm(long_avg[k][mask])
            out[k] = float(diff / ref) if ref > 0 else float("inf")
    return out


def converged(err, tol):
    return all(v <= tol for v in err.values())
