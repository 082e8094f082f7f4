"""dsmcVolFields: macroscopic fields from the engine's cumulative per-species
sampling sums (dsmcVolFields.C:1242-1290 every step, 1365-1790 at write).

A field model is a list of species (its `typeIds`, e.g. ["N2"] or the
mixture); its sums are the sums over those species.  Every cell field is an
FP64 array over cells.  Cells with no sample are 0 (as the reference sets
them, dsmcVolFields.C:1279-1289).
"""
import math

import numpy as np

from . import abi

KB = abi.HS_KB
NA = 6.02214086e23  # per mol (CODATA 2014, the OpenFOAM v1706 era constant)
SMALL = 1e-15
VSMALL = 1e-300
GREAT = 1e15


def species_sums(samples, case):
    """Per-species views of the sample array [slot][cell]."""
    ns = case.cfg.n_species
    out = []
    for s in range(ns):
        b = abi.HS_S_GLOBAL_COUNT + s * abi.HS_SS_COUNT
        d = {
            "N": samples[b + abi.HS_SS_N],
            "MU": samples[b + abi.HS_SS_MUX:b + abi.HS_SS_MUZ + 1],
            "MCC": samples[b + abi.HS_SS_MCC],
            "EROT": samples[b + abi.HS_SS_EROT],
            "EVIB": samples[b + abi.HS_SS_EVIB0:b + abi.HS_SS_EVIB0 + abi.HS_MAX_VIB_MODES],
        }
        if case.cfg.flags & abi.HS_FLAG_HEATFLUX:
            h = abi.HS_S_GLOBAL_COUNT + ns * abi.HS_SS_COUNT + s * abi.HS_HF_COUNT
            d["HF"] = samples[h:h + abi.HS_HF_COUNT]
        out.append(d)
    return out


def vol_fields(samples, nsteps, case, type_ids=None, tref_mfp=273.0):
    """The fields of one dsmcVolFields model over the species `type_ids`
    (indices; default all = mixture).  nsteps is nTimeSteps_."""
    cfg = case.cfg
    sp = case.species
    ids = list(range(cfg.n_species)) if type_ids is None else list(type_ids)
    S = species_sums(samples, case)
    V = case.volumes
    fn = cfg.fn
    dt = cfg.dt
    nc = case.ncells
    T = max(nsteps, 1)
    dsmcN = sum(S[s]["N"] for s in ids)
    dsmcM = sum(sp[s].mass * S[s]["N"] for s in ids)
    mom = sum(S[s]["MU"] for s in ids)
    mcc = sum(S[s]["MCC"] for s in ids)
    erot = sum(S[s]["EROT"] for s in ids)
    zrot = sum(sp[s].rot_dof * S[s]["N"] for s in ids)
    nCum, mCum = fn * dsmcN, fn * dsmcM
    ok = (dsmcN > 1e-3) & (V > 0)
    Vs = np.where(V > 0, V, 1.0)
    f = {}
    with np.errstate(divide="ignore", invalid="ignore"):
        rhoN = np.where(ok, nCum / (T * Vs), 0.0)
        rhoM = np.where(ok, mCum / (T * Vs), 0.0)
        U = np.where(ok, fn * mom / np.where(ok, mCum, 1.0), 0.0)
        lke_mean = 0.5 * fn * mcc / (Vs * T)
        Ttra = np.where(ok, 2.0 / (3.0 * KB * np.where(ok, rhoN, 1.0))
                        * (lke_mean - 0.5 * rhoM * (U * U).sum(axis=0)), 0.0)
        f.update(rhoN=rhoN, rhoM=rhoM, U=U, Ttra=Ttra, p=rhoN * KB * Ttra, dsmcNMean=np.where(ok, dsmcN / T, 0.001))
        # rotation
        Trot = np.where(zrot > SMALL, 2.0 * erot / (KB * np.where(zrot > SMALL, zrot, 1.0)), 0.0)
        zeta_rot_tot = np.where(dsmcN > SMALL, zrot / np.where(dsmcN > SMALL, dsmcN, 1.0), 0.0)
        # vibration (dsmcVolFields.C:1440-1490): per species, the last excited
        # mode's zeta*T over the species' total zeta
        Tvib = np.zeros(nc)
        zvib = np.zeros(nc)
        mol = np.zeros(nc)
        for s in ids:
            ns_ = S[s]["N"]
            zeta_s = np.zeros(nc)
            zeta_t = np.zeros(nc)
            for mo in range(sp[s].n_vib_modes):
                th = sp[s].theta_v[mo]
                ev = S[s]["EVIB"][mo]
                good = (ev > VSMALL) & (ns_ > SMALL)
                imean = np.where(good, ev / (KB * th * np.where(good, ns_, 1.0)), 0.0)
                good &= imean > SMALL
                lf = np.where(good, np.log(1.0 + 1.0 / np.where(good, imean, 1.0)), 0.0)
                tv = np.where(good, th / np.where(good, lf, 1.0), 0.0)
                zm = np.where(good, 2.0 * imean * lf, 0.0)
                zeta_s += zm
                zeta_t = np.where(good, zm * tv, zeta_t)
            has = zeta_s > SMALL
            tv_s = np.where(has, zeta_t / np.where(has, zeta_s, 1.0), 0.0)
            nsc = fn * ns_
            mol += np.where(has, nsc, 0.0)
            Tvib += np.where(has, nsc * tv_s, 0.0)
            zvib += np.where(has, nsc * zeta_s, 0.0)
        m = mol > SMALL
        Tvib = np.where(m, Tvib / np.where(m, mol, 1.0), 0.0)
        zvib = np.where(m, zvib / np.where(m, mol, 1.0), 0.0)
        Tov = (3.0 * Ttra + zeta_rot_tot * Trot + zvib * Tvib) / (3.0 + zeta_rot_tot + zvib)
        f.update(Trot=Trot, Tvib=Tvib, Tov=Tov, zetaVib=zvib)
        # Mach number
        molar_mass = np.zeros(nc)
        cv = np.zeros(nc)
        cp = np.zeros(nc)
        good = (dsmcN > SMALL) & (Ttra > SMALL)
        for s in ids:
            X = np.where(good, fn * S[s]["N"] / np.where(good, nCum, 1.0), 0.0)
            molar_mass += X * sp[s].mass
            cv += X * (3.0 + sp[s].rot_dof)
            cp += X * (5.0 + sp[s].rot_dof)
        gamma = np.where(good, cp / np.where(good, cv, 1.0), 0.0)
        a = np.sqrt(np.where(good, gamma * KB / np.where(good, molar_mass, 1.0) * Ttra, 1.0))
        f["Ma"] = np.where(good, np.sqrt((U * U).sum(axis=0)) / a, 0.0)
        # mean free path (Bird 4.76) and mean collision rate (Bird 4.74)
        mfp = np.zeros(nc)
        mcr = np.zeros(nc)
        hot = Ttra > 1.0
        for s in ids:
            smfp = np.zeros(nc)
            smcr = np.zeros(nc)
            for r in ids:
                dpq = 0.5 * (sp[s].diameter + sp[r].diameter)
                om = 0.5 * (sp[s].omega + sp[r].omega)
                mr = sp[s].mass * sp[r].mass / (sp[s].mass + sp[r].mass)
                nq = fn * S[r]["N"] / (Vs * T)
                g = hot & (S[r]["N"] > SMALL)
                Tt = np.where(g, Ttra, tref_mfp)
                smfp += np.where(g, math.pi * dpq * dpq * nq * (tref_mfp / Tt) ** (om - 0.5)
                                 * math.sqrt(1.0 + sp[s].mass / sp[r].mass), 0.0)
                smcr += np.where(g, 2.0 * math.sqrt(math.pi) * dpq * dpq * nq * (Tt / tref_mfp) ** (1.0 - om)
                                 * math.sqrt(2.0 * KB * tref_mfp / mr), 0.0)
            smfp = np.where(smfp > SMALL, 1.0 / np.where(smfp > SMALL, smfp, 1.0), 0.0)
            w = np.where(nCum > SMALL, fn * S[s]["N"] / np.where(nCum > SMALL, nCum, 1.0), 0.0)
            mfp += np.where(hot & (rhoN > SMALL), smfp * w, 0.0)
            mcr += np.where(hot & (rhoN > SMALL), smcr * w, 0.0)
        f["mfp"] = np.where(mfp < SMALL, GREAT, mfp)
        f["meanCollisionRate"] = mcr
        f["mct"] = np.where(mcr > SMALL, 1.0 / np.where(mcr > SMALL, mcr, 1.0), GREAT)
        ncoll = samples[abi.HS_S_NCOLL]
        f["measuredCollisionRate"] = np.where(nCum > SMALL, ncoll * fn / (np.where(nCum > SMALL, nCum, 1.0) * dt), 0.0)
        sep = samples[abi.HS_S_COLLSEP]
        f["meanCollisionSeparation"] = np.where(ncoll > SMALL, sep / np.where(ncoll > SMALL, ncoll, 1.0), GREAT)
        # pressure tensor and heat flux (measureHeatFluxShearStress)
        if cfg.flags & abi.HS_FLAG_HEATFLUX:
            H = sum(S[s]["HF"] for s in ids)
            c = np.where(dsmcN > SMALL, rhoN / np.where(dsmcN > SMALL, dsmcN, 1.0), 0.0)
            Ux, Uy, Uz = U
            P = np.zeros((3, 3, nc))
            P[0, 0] = c * (H[0] - dsmcM * Ux * Ux)
            P[0, 1] = P[1, 0] = c * (H[1] - dsmcM * Ux * Uy)
            P[0, 2] = P[2, 0] = c * (H[2] - dsmcM * Ux * Uz)
            P[1, 1] = c * (H[3] - dsmcM * Uy * Uy)
            P[1, 2] = P[2, 1] = c * (H[4] - dsmcM * Uy * Uz)
            P[2, 2] = c * (H[5] - dsmcM * Uz * Uz)
            q = np.zeros((3, nc))
            for i, (mccu, eu, Ui) in enumerate(((H[6], H[9], Ux), (H[7], H[10], Uy), (H[8], H[11], Uz))):
                q[i] = c * (0.5 * mccu - 0.5 * mcc * Ui + eu - H[12] * Ui) - (P[i, 0] * Ux + P[i, 1] * Uy + P[i, 2] * Uz)
            f["pressureTensor"] = P
            f["heatFluxVector"] = q
    return f


def global_temperature(samples, case, type_ids=None):
    """Translational and rotational temperature of all sampled parcels
    pooled over the whole mesh."""
    cfg = case.cfg
    sp = case.species
    ids = list(range(cfg.n_species)) if type_ids is None else list(type_ids)
    S = species_sums(samples, case)
    N = sum(S[s]["N"].sum() for s in ids)
    M = sum(sp[s].mass * S[s]["N"].sum() for s in ids)
    mom = sum(S[s]["MU"].sum(axis=1) for s in ids)
    mcc = sum(S[s]["MCC"].sum() for s in ids)
    Ttr = (mcc - (mom * mom).sum() / M) / (3.0 * KB * N)
    z = sum(sp[s].rot_dof * S[s]["N"].sum() for s in ids)
    er = sum(S[s]["EROT"].sum() for s in ids)
    Trot = 2.0 * er / (KB * z) if z > 0 else 0.0
    return Ttr, Trot, N


def wall_fields(wall_sums, nsteps, case, faces, type_ids=None):
    """Boundary values of one dsmcVolFields model on the wall faces
    (dsmcVolFields.C:1890-2126) from the engine's wall-face sums
    [species*HS_WS_COUNT + slot][face].  `faces` is the engine's wall-face
    list (side, cell, area, normal).  Returns per-face arrays: rhoN, rhoM,
    U[3], Ttra, Trot, Tvib, Tov, fD[3], p (= fD.n), tau (tangential |fD|)
    and q (wall heat flux, positive into the wall)."""
    cfg = case.cfg
    sp = case.species
    ids = list(range(cfg.n_species)) if type_ids is None else list(type_ids)
    nf = wall_sums.shape[1]
    W = [wall_sums[s * abi.HS_WS_COUNT:(s + 1) * abi.HS_WS_COUNT] for s in range(cfg.n_species)]
    T = max(nsteps, 1)
    fn = cfg.fn
    rhoNBF = sum(W[s][abi.HS_WS_RHON] for s in ids)
    rhoMBF = sum(sp[s].mass * W[s][abi.HS_WS_RHON] for s in ids)
    keBF = sum(W[s][abi.HS_WS_KE] for s in ids)
    momBF = sum(W[s][abi.HS_WS_MOMX:abi.HS_WS_MOMZ + 1] for s in ids)
    erotBF = sum(W[s][abi.HS_WS_EROT] for s in ids)
    zrotBF = sum(sp[s].rot_dof * W[s][abi.HS_WS_RHON] for s in ids)
    qBF = sum(W[s][abi.HS_WS_Q] for s in ids)
    fDBF = sum(W[s][abi.HS_WS_FDX:abi.HS_WS_FDZ + 1] for s in ids)
    f = {}
    with np.errstate(divide="ignore", invalid="ignore"):
        rhoN = rhoNBF * fn / T
        rhoM = rhoMBF * fn / T
        lke = keBF * fn / T
        ok = rhoM > VSMALL
        U = np.where(ok, momBF / np.where(ok, rhoMBF, 1.0), 0.0)
        Ttra = np.where(ok, 2.0 / (3.0 * KB * np.where(ok, rhoN, 1.0)) * (lke - 0.5 * rhoM * (U * U).sum(axis=0)),
                        0.0)
        zeta_rot_tot = np.where(rhoNBF > SMALL, zrotBF / np.where(rhoNBF > SMALL, rhoNBF, 1.0), 0.0)
        Trot = np.where(zrotBF > SMALL, 2.0 * erotBF / (KB * np.where(zrotBF > SMALL, zrotBF, 1.0)), 0.0)
        Tvib = np.zeros(nf)
        zvib = np.zeros(nf)
        mol = np.zeros(nf)
        for s in ids:
            ns_ = W[s][abi.HS_WS_RHON]
            zeta_s = np.zeros(nf)
            zt = np.zeros(nf)
            for mo in range(sp[s].n_vib_modes):
                th = sp[s].theta_v[mo]
                good = ns_ > SMALL
                imean = np.where(good, W[s][abi.HS_WS_EVIB0 + mo] / (KB * th * np.where(good, ns_, 1.0)), 0.0)
                good &= imean > SMALL
                lf = np.where(good, np.log(1.0 + 1.0 / np.where(good, imean, 1.0)), 0.0)
                tv = np.where(good, th / np.where(good, lf, 1.0), 0.0)
                zm = np.where(good, 2.0 * imean * lf, 0.0)
                zeta_s += zm
                zt += zm * tv
            has = zeta_s > SMALL
            mol += np.where(has, ns_, 0.0)
            Tvib += np.where(has, ns_ * zt / np.where(has, zeta_s, 1.0), 0.0)
            zvib += np.where(has, ns_ * zeta_s, 0.0)
        m = mol > SMALL
        Tvib = np.where(m, Tvib / np.where(m, mol, 1.0), 0.0)
        zvib = np.where(m, zvib / np.where(m, mol, 1.0), 0.0)
        Tov = (3.0 * Ttra + zeta_rot_tot * Trot + zvib * Tvib) / (3.0 + zeta_rot_tot + zvib)
    fD = fDBF / T
    n = faces["normal"].T
    p = (fD * n).sum(axis=0)
    tau = np.sqrt(np.maximum((fD * fD).sum(axis=0) - p * p, 0.0))
    f.update(rhoN=rhoN, rhoM=rhoM, U=U, Ttra=Ttra, Trot=Trot, Tvib=Tvib, Tov=Tov, fD=fD, p=p, tau=tau, q=qBF / T)
    return f
