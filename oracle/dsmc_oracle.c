/*
 * dsmc_oracle.c -- CPU restatement of dsmcFoam+'s per-timestep particle
 * path.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP engine
 * and the CPU baseline of the benchmark.  The product never links this.
 *
 * Reference files followed (paths relative to the hyStrath tree):
 *   NTC selection      src/lagrangian/dsmc/collisionPartnerSelection/derived/noTimeCounter/noTimeCounter.C:81-339
 *   randomLabel        src/lagrangian/dsmc/clouds/dsmcCloud.C:1015-1040
 *   VHS                src/lagrangian/dsmc/collisions/derived/VariableHardSphere/VariableHardSphere.C:80-228
 *   VSS                .../VariableSoftSphere/VariableSoftSphere.C:78-247
 *   Larsen-Borgnakke   .../LarsenBorgnakkeVariableHardSphere/LarsenBorgnakkeVariableHardSphere.C:110-279
 *   LB samplers        src/lagrangian/dsmc/clouds/dsmcCloud.C:1181-1244,1327-1531,1534-1656
 *   move               src/lagrangian/dsmc/parcels/dsmcParcel.C:38-148 and the tracker
 *                      src/lagrangian/basic/particle/particleTemplates.C:727-1241,1525-1571
 *   walls              .../dsmcDiffuseWallPatch/dsmcDiffuseWallPatch.C:67-139,
 *                      .../basic/dsmcPatchBoundary/dsmcPatchBoundary.C:216-261,
 *                      .../dsmcSpecularWallPatch/dsmcSpecularWallPatch.C:47-61,
 *                      .../dsmcDeletionPatch/dsmcDeletionPatch.C:140-150
 *   inflow             .../dsmcFreeStreamInflowPatch/dsmcFreeStreamInflowPatch.C:91-364
 *   cell occupancy     src/lagrangian/dsmc/clouds/dsmcCloud.C:63-74
 *   sampling sums      src/lagrangian/dsmc/macroscopicProperties/derived/combined/dsmcVolFields/dsmcVolFields.C:1115-1239
 *   mesh fill          src/lagrangian/dsmc/initialiseDsmcParcels/derived/dsmcMeshFill/dsmcMeshFill.C:68-~230
 *
 * Deliberate, documented differences from the reference (all of them change
 * which random numbers are used, never the distributions):
 *   - Random numbers come from Philox4x32-10 streams keyed by (seed, purpose,
 *     entity, step) instead of OpenFOAM's single sequential Random; each
 *     NTC collision candidate has its own stream (cell, step, candidate).
 *   - Gaussians are Box-Muller pairs (second value cached) instead of the
 *     polar method; arguments of log are 1-u so they are never 0.
 *   - A freestream particle's face position is uniform on the rectangular
 *     face from two draws instead of triangle selection + randomPoint.
 * The mesh is a uniform Cartesian block; the tet-walk of the reference
 * reduces to axis-aligned face crossings on it, plus an analytic cylinder.
 */
#include "dsmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
/* The OpenMP build (oracle/Makefile target mt) is the benchmark's all-core CPU
 * baseline: parcels and cells are shared out among threads, and sums that
 * threads share are updated atomically.  The default build has no OpenMP
 * and is the deterministic single-thread checker. */
#define OR_ATOMIC _Pragma("omp atomic")
#else
#define OR_ATOMIC
#endif

#define OR_VSMALL 1.0e-300
#define OR_SMALL 1.0e-15
#define OR_ROOTVSMALL 1.0e-150
#define OR_PI 3.14159265358979323846
#define OR_TWO_PI 6.28318530717958647692
#define OR_MAX_TRACK 1000
#define OR_MAX_REDRAW 100000
#define OR_MAX_CAND (1 << 22)
#define OR_MAX_FILL_TRIES 100000

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                         */
/* ------------------------------------------------------------------------ */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53 random bits -> [0,1) exactly */
double or_uniform_from(uint32_t hi, uint32_t lo)
{
    const uint64_t x = (((uint64_t)hi << 32) | (uint64_t)lo) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

typedef struct {
    uint32_t key[2];
    uint32_t ctr[3];
    uint32_t blk;
    int have_u, have_g;
    double spare_u, spare_g;
    const double* inj;       /* injected uniforms (test hook) or NULL */
    int inj_n, inj_used;
} or_rng;

static void rng_init(or_rng* r, uint64_t seed, int purpose, uint32_t id, uint32_t step, uint32_t sub)
{
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ctr[0] = id;
    r->ctr[1] = step;
    r->ctr[2] = ((uint32_t)purpose << 24) | (sub & 0xFFFFFFu);
    r->blk = 0;
    r->have_u = 0;
    r->have_g = 0;
    r->spare_u = 0.0;
    r->spare_g = 0.0;
    r->inj = NULL;
    r->inj_n = 0;
    r->inj_used = 0;
}

static double rng_u(or_rng* r)
{
    if (r->inj) {
        const int k = r->inj_used++;
        return k < r->inj_n ? r->inj[k] : 0.5;
    }
    if (r->have_u) { r->have_u = 0; return r->spare_u; }
    uint32_t c[4] = { r->ctr[0], r->ctr[1], r->ctr[2], r->blk };
    uint32_t o[4];
    or_philox4x32(c, r->key, o);
    r->blk += 1;
    r->spare_u = or_uniform_from(o[2], o[3]);
    r->have_u = 1;
    return or_uniform_from(o[0], o[1]);
}

/* ------------------------------------------------------------------------ */
/* Deterministic FP64 math from IEEE +,-,*,/ only                           */
/* ------------------------------------------------------------------------ */
#define OR_LN2_HI 6.93147180369123816490e-01 /* 0x3fe62e42fee00000 */
#define OR_LN2_LO 1.90821492927058770002e-10 /* 0x3dea39ef35793c76 */
#define OR_INV_LN2 1.44269504088896338700e+00

static double bits_to_d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t d_to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* log(x): x = m*2^k, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 * s = (m-1)/(m+1), series to s^23 (|s| <= 0.1716, truncation < 1e-19). */
double or_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    int k = 0;
    uint64_t b = d_to_bits(x);
    if (((b >> 52) & 0x7ff) == 0) { x = x * 18014398509481984.0; b = d_to_bits(x); k = -54; }
    k += (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_to_d((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; k += 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double t = 1.0 / 23.0;
    t = 1.0 / 21.0 + s2 * t;
    t = 1.0 / 19.0 + s2 * t;
    t = 1.0 / 17.0 + s2 * t;
    t = 1.0 / 15.0 + s2 * t;
    t = 1.0 / 13.0 + s2 * t;
    t = 1.0 / 11.0 + s2 * t;
    t = 1.0 / 9.0 + s2 * t;
    t = 1.0 / 7.0 + s2 * t;
    t = 1.0 / 5.0 + s2 * t;
    t = 1.0 / 3.0 + s2 * t;
    const double lm = (2.0 * s) + (2.0 * s) * (s2 * t);
    const double dk = (double)k;
    return dk * OR_LN2_HI + (dk * OR_LN2_LO + lm);
}

/* exp(x) = 2^k e^r, |r| <= ln2/2, Taylor to r^14 (truncation < 1e-18). */
double or_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    const double kd = floor(x * OR_INV_LN2 + 0.5);
    const double r = (x - kd * OR_LN2_HI) - kd * OR_LN2_LO;
    double p = 1.0 / 87178291200.0;
    p = 1.0 / 6227020800.0 + r * p;
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    int k = (int)kd;
    if (k < -1000) { p = p * 0x1p-511; k += 511; }
    if (k > 1000) { p = p * 0x1p511; k -= 511; }
    return p * bits_to_d((uint64_t)(k + 1023) << 52);
}

double or_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    return or_exp(y * or_log(x));
}

/* sin and cos of 2*pi*u for u in [0,1): quadrant from round(4u), remainder
 * d = u - q/4 is exact, r = 2*pi*d in [-pi/4, pi/4]; Taylor to r^17 / r^18. */
void or_sincos2pi(double u, double* s_out, double* c_out)
{
    const double q = floor(u * 4.0 + 0.5);
    const double d = u - q * 0.25;
    const double r = d * OR_TWO_PI;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;
    ps = -1.0 / 1307674368000.0 + r2 * ps;
    ps = 1.0 / 6227020800.0 + r2 * ps;
    ps = -1.0 / 39916800.0 + r2 * ps;
    ps = 1.0 / 362880.0 + r2 * ps;
    ps = -1.0 / 5040.0 + r2 * ps;
    ps = 1.0 / 120.0 + r2 * ps;
    ps = -1.0 / 6.0 + r2 * ps;
    const double sr = r + (r * r2) * ps;
    double pc = -1.0 / 6402373705728000.0;
    pc = 1.0 / 20922789888000.0 + r2 * pc;
    pc = -1.0 / 87178291200.0 + r2 * pc;
    pc = 1.0 / 479001600.0 + r2 * pc;
    pc = -1.0 / 3628800.0 + r2 * pc;
    pc = 1.0 / 40320.0 + r2 * pc;
    pc = -1.0 / 720.0 + r2 * pc;
    pc = 1.0 / 24.0 + r2 * pc;
    pc = -0.5 + r2 * pc;
    const double cr = 1.0 + r2 * pc;
    switch (((int)q) & 3) {
    case 0: *s_out = sr; *c_out = cr; break;
    case 1: *s_out = cr; *c_out = -sr; break;
    case 2: *s_out = -sr; *c_out = -cr; break;
    default: *s_out = -cr; *c_out = sr; break;
    }
}

/* one standard normal; Box-Muller pair, the second value cached */
static double rng_gauss(or_rng* r)
{
    if (r->have_g) { r->have_g = 0; return r->spare_g; }
    const double u1 = rng_u(r);
    const double u2 = rng_u(r);
    const double rad = sqrt(-2.0 * or_log(1.0 - u1));
    double sn, cs;
    or_sincos2pi(u2, &sn, &cs);
    r->spare_g = rad * sn;
    r->have_g = 1;
    return rad * cs;
}

void or_stream_draws(uint64_t seed, int purpose, uint32_t id, uint32_t step,
                     uint32_t sub, int n, double* out)
{
    or_rng r;
    rng_init(&r, seed, purpose, id, step, sub);
    for (int i = 0; i < n; ++i) out[i] = rng_u(&r);
}

void or_stream_gauss(uint64_t seed, int purpose, uint32_t id, uint32_t step,
                     uint32_t sub, int n, double* out)
{
    or_rng r;
    rng_init(&r, seed, purpose, id, step, sub);
    for (int i = 0; i < n; ++i) out[i] = rng_gauss(&r);
}

/* dsmcCloud::randomLabel (dsmcCloud.C:1015-1040) */
static int rng_label(or_rng* r, int a, int b)
{
    if (a == b) return a;
    const int start = a < b ? a : b;
    const int end = a < b ? b : a;
    for (;;) {
        const int v = start + (int)(rng_u(r) * (double)(end - start + 1));
        if (v != end + 1) return v;
    }
}

/* ------------------------------------------------------------------------ */
/* derived constants                                                        */
/* ------------------------------------------------------------------------ */
/* one reaction of chemReactDict with its constants (see derive) */
typedef struct {
    int model, r[2], dprod[2][2];
    int xmol, xatom;       /* exchange products: the molecule, the atom          */
    int pos_mol;           /* exchangeQK::posMolReactant_ (exchangeQK.C:52-60)   */
    int split;             /* allowSplitting                                      */
    double heat_diss[2];   /* k thetaD of each reactant (dissociationQK.C:57-60)  */
    double heat_x;         /* heatOfReactionExchange k (exchangeQK.C:387-391)     */
    double a_dash, b;      /* Bird's QK exchange constants (exchangeQK.C:195-200) */
    double om_x;           /* omega of the exchange's redistribution (exchangeQK.C:303-308) */
} or_rx;

typedef struct {
    int ns;
    int n_react;
    int react_of[HS_MAX_SPECIES][HS_MAX_SPECIES];  /* pairAddressing_, -1 none */
    or_rx rx[HS_MAX_REACTIONS];
    double sig_coef[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double ref_term[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double omega_pq[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double alpha_pq[HS_MAX_SPECIES][HS_MAX_SPECIES];
    /* Bird 2010 Zv: [species][partner][mode] Zref*(thetaD/TrefZv)^-omegaPQ */
    double zv_b[HS_MAX_SPECIES][HS_MAX_SPECIES][HS_MAX_VIB_MODES];
    double zv_pow2[HS_MAX_SPECIES][HS_MAX_VIB_MODES];
    int empty[3];
    double centre[3];
    int any_rot, any_vib, multi;
    int max_modes;
} or_derived;

/* a species' kind (dsmcParcelI.H:153-196): 0 electron, 10/11 atom,
 * 20/21 diatomic, 30/31 polyatomic molecule (neutral / ion) */
static int species_kind(const hs_species* sp)
{
    if (sp->charge == -1) return 0;
    const int base = sp->n_vib_modes == 0 ? 10 : (sp->n_vib_modes == 1 ? 20 : 30);
    return base + (sp->charge == 1 ? 1 : 0);
}

/* the reactions' constants and the pair table of dsmcReactions::
 * initialConfiguration (dsmcReactions.C:139-160) with each model's
 * tryReactMolecules (dissociationQK.C:430-455, exchangeQK.C:420-443) */
static void derive_reactions(const hs_config* cfg, or_derived* d)
{
    for (int p = 0; p < HS_MAX_SPECIES; ++p)
        for (int q = 0; q < HS_MAX_SPECIES; ++q) d->react_of[p][q] = -1;
    d->n_react = (cfg->reactions && cfg->n_reactions > 0) ? cfg->n_reactions : 0;
    if (d->n_react > HS_MAX_REACTIONS) d->n_react = HS_MAX_REACTIONS;
    for (int i = 0; i < d->n_react; ++i) {
        const hs_reaction* R = &cfg->reactions[i];
        or_rx* x = &d->rx[i];
        memset(x, 0, sizeof(*x));
        x->model = R->model;
        x->split = R->allow_splitting != 0;
        for (int k = 0; k < 2; ++k) {
            x->r[k] = R->reactants[k];
            x->dprod[k][0] = R->diss_products[k][0];
            x->dprod[k][1] = R->diss_products[k][1];
            x->heat_diss[k] = HS_KB * cfg->species[R->reactants[k]].theta_d;
        }
        x->pos_mol = -1;
        x->xmol = x->xatom = -1;
        if (R->model != HS_REACT_DISSOCIATION_QK) {
            for (int k = 0; k < 2; ++k)
                if (species_kind(&cfg->species[R->reactants[k]]) >= 20) x->pos_mol = k;
            for (int k = 0; k < 2; ++k) {
                const int pid = R->exch_products[k];
                const int kind = species_kind(&cfg->species[pid]);
                if (kind >= 20) x->xmol = pid;
                else if (kind == 10 || kind == 11) x->xatom = pid;
            }
            x->heat_x = R->heat_exchange * HS_KB;
            x->b = R->b_coeff;
            const double om = 0.5 * (cfg->species[R->reactants[0]].omega + cfg->species[R->reactants[1]].omega);
            const double chib = 2.5 - om;
            x->a_dash = R->a_coeff * (pow(chib, R->b_coeff) * exp(lgamma(chib)) / exp(lgamma(chib + R->b_coeff)));
            if (x->xatom >= 0) x->om_x = 0.5 * (cfg->species[x->xatom].omega + cfg->species[x->xatom].omega);
        }
        for (int p = 0; p < cfg->n_species; ++p)
            for (int q = p; q < cfg->n_species; ++q) {
                const int pid = p == x->r[0] ? 0 : (p == x->r[1] ? 1 : -1);
                const int qid = q == x->r[0] ? 0 : (q == x->r[1] ? 1 : -1);
                if (pid < 0 || qid < 0) continue;
                int ok;
                if (R->model == HS_REACT_EXCHANGE_QK) ok = pid != qid;
                else ok = (pid == qid && x->r[0] == x->r[1]) || (pid != qid && x->r[0] != x->r[1]);
                if (ok) { d->react_of[p][q] = i; d->react_of[q][p] = i; }
            }
    }
}

static void derive(const hs_config* cfg, or_derived* d)
{
    memset(d, 0, sizeof(*d));
    d->ns = cfg->n_species;
    for (int p = 0; p < d->ns; ++p) {
        const hs_species* sp = &cfg->species[p];
        if (sp->rot_dof > 0) d->any_rot = 1;
        if (sp->n_vib_modes > 0) d->any_vib = 1;
        if (sp->n_vib_modes > d->max_modes) d->max_modes = sp->n_vib_modes;
        for (int m = 0; m < sp->n_vib_modes; ++m)
            d->zv_pow2[p][m] = pow(sp->theta_d / sp->tref_zv[m], 1.0 / 3.0) - 1.0;
        for (int q = 0; q < d->ns; ++q) {
            const hs_species* sq = &cfg->species[q];
            const double dpq = 0.5 * (sp->diameter + sq->diameter);
            const double om = 0.5 * (sp->omega + sq->omega);
            const double mr = sp->mass * sq->mass / (sp->mass + sq->mass);
            d->sig_coef[p][q] = OR_PI * dpq * dpq / exp(lgamma(2.5 - om));
            d->ref_term[p][q] = 2.0 * HS_KB * cfg->tref / mr;
            d->omega_pq[p][q] = om;
            d->alpha_pq[p][q] = 0.5 * (sp->alpha + sq->alpha);
            for (int m = 0; m < sp->n_vib_modes; ++m)
                d->zv_b[p][q][m] = sp->zref[m] * pow(sp->theta_d / sp->tref_zv[m], -om);
        }
    }
    d->multi = d->ns > 1;
    derive_reactions(cfg, d);
    const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
    for (int a = 0; a < 3; ++a) {
        d->empty[a] = cfg->patch[2 * a].model == HS_PATCH_EMPTY;
        d->centre[a] = cfg->origin[a] + 0.5 * ((double)nn[a] * cfg->delta[a]);
    }
}

int32_t or_sample_slots(const hs_config* cfg)
{
    int32_t n = HS_S_GLOBAL_COUNT + cfg->n_species * HS_SS_COUNT;
    if (cfg->flags & HS_FLAG_HEATFLUX) n += cfg->n_species * HS_HF_COUNT;
    return n;
}

/* ------------------------------------------------------------------------ */
/* inflow faces: sides in hs_side order, faces in blockMesh order on each   */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t cell; int8_t axis, dir; int32_t a1i, a2i; } or_face;

static int64_t list_inflow_faces(const hs_config* cfg, or_face* out)
{
    int64_t nf = 0;
    const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
    for (int side = 0; side < 6; ++side) {
        if (!cfg->patch[side].inflow) continue;
        const int axis = side / 2, dir = (side & 1) ? 1 : -1;
        const int a1 = axis == 0 ? 1 : 0;
        const int a2 = axis == 2 ? 1 : 2;
        for (int j2 = 0; j2 < nn[a2]; ++j2)
            for (int j1 = 0; j1 < nn[a1]; ++j1) {
                int idx[3];
                idx[axis] = dir > 0 ? nn[axis] - 1 : 0;
                idx[a1] = j1;
                idx[a2] = j2;
                const int32_t c = idx[0] + cfg->nx * (idx[1] + cfg->ny * idx[2]);
                if (!(cfg->cell_volume[c] > 0.0)) continue;
                if (out) {
                    out[nf].cell = c; out[nf].axis = (int8_t)axis; out[nf].dir = (int8_t)dir;
                    out[nf].a1i = j1; out[nf].a2i = j2;
                }
                ++nf;
            }
    }
    return nf;
}

int64_t or_inflow_face_count(const hs_config* cfg) { return list_inflow_faces(cfg, NULL); }

/* ------------------------------------------------------------------------ */
/* equipartition samplers (dsmcCloud.C:1181-1244)                           */
/* ------------------------------------------------------------------------ */
static double equip_erot(or_rng* r, double T, double dof)
{
    if (dof < OR_SMALL) return 0.0;
    if (dof < 2.0 + OR_SMALL && dof > 2.0 - OR_SMALL)
        return -or_log(1.0 - rng_u(r)) * HS_KB * T;
    const double a = 0.5 * dof - 1.0;
    double ratio, P;
    int guard = 0;
    do {
        ratio = 10.0 * rng_u(r);
        P = or_pow(ratio / a, a) * or_exp(a - ratio);
    } while (P < rng_u(r) && ++guard < OR_MAX_REDRAW);
    return ratio * HS_KB * T;
}

static void equip_vib(or_rng* r, double T, const hs_species* sp, uint16_t* lv, int64_t stride)
{
    for (int m = 0; m < sp->n_vib_modes; ++m) {
        const int l = (int)(-or_log(1.0 - rng_u(r)) * T / sp->theta_v[m]);
        lv[m * stride] = (uint16_t)(l > 65535 ? 65535 : l);
    }
}

/* ------------------------------------------------------------------------ */
/* sigmaTcR and binary collisions                                           */
/* ------------------------------------------------------------------------ */
static double sigma_tcr(const or_derived* d, int tp, int tq, const double up[3], const double uq[3])
{
    const double dx = up[0] - uq[0], dy = up[1] - uq[1], dz = up[2] - uq[2];
    const double cr2 = dx * dx + dy * dy + dz * dz;
    const double cr = sqrt(cr2);
    if (cr < OR_VSMALL) return 0.0;
    const double om = d->omega_pq[tp][tq];
    return d->sig_coef[tp][tq] * or_pow(d->ref_term[tp][tq] / cr2, om - 0.5) * cr;
}

/* postCollisionVelocities, VHS (VariableHardSphere.C:191-228) or VSS
 * (VariableSoftSphere.C:189-247).  cr < 0 means |UP-UQ|.  In the LB-VSS path
 * (rescale=1) the pre-collision relative-velocity components are scaled to
 * magnitude cr so Bird eq. 2.22 conserves the redistributed energy. */
static void post_coll_vel(const hs_config* cfg, const or_derived* d, int vss, int rescale,
                          int tp, int tq, double up[3], double uq[3], double cr, or_rng* r)
{
    const double mp = cfg->species[tp].mass, mq = cfg->species[tq].mass;
    if (cr < 0.0) {
        const double dx = up[0] - uq[0], dy = up[1] - uq[1], dz = up[2] - uq[2];
        cr = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const double msum = mp + mq;
    double ucm[3];
    for (int c = 0; c < 3; ++c) ucm[c] = (mp * up[c] + mq * uq[c]) / msum;
    double rel[3];
    if (!vss) {
        const double cos_t = 2.0 * rng_u(r) - 1.0;
        const double sin_t = sqrt(1.0 - cos_t * cos_t);
        double sp, cp;
        or_sincos2pi(rng_u(r), &sp, &cp);
        rel[0] = cr * cos_t;
        rel[1] = cr * (sin_t * cp);
        rel[2] = cr * (sin_t * sp);
    } else {
        double c0 = up[0] - uq[0], c1 = up[1] - uq[1], c2 = up[2] - uq[2];
        if (rescale) {
            const double cm = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            if (cm > OR_VSMALL) { const double f = cr / cm; c0 *= f; c1 *= f; c2 *= f; }
        }
        const double cos_t = 2.0 * or_pow(rng_u(r), 1.0 / d->alpha_pq[tp][tq]) - 1.0;
        const double sin_t = sqrt(1.0 - cos_t * cos_t);
        double sp, cp;
        or_sincos2pi(rng_u(r), &sp, &cp);
        const double D = sqrt(c1 * c1 + c2 * c2);
        if (D > OR_VSMALL) {
            rel[0] = cos_t * c0 + sin_t * sp * D;
            rel[1] = cos_t * c1 + sin_t * (cr * c2 * cp - c0 * c1 * sp) / D;
            rel[2] = cos_t * c2 - sin_t * (cr * c1 * cp + c0 * c2 * sp) / D;
        } else {
            /* relative velocity along x: Bird 2.22 is singular, the
             * deflection is isotropic about x */
            rel[0] = cos_t * c0;
            rel[1] = cr * (sin_t * cp);
            rel[2] = cr * (sin_t * sp);
        }
    }
    for (int c = 0; c < 3; ++c) {
        up[c] = ucm[c] + rel[c] * mq / msum;
        uq[c] = ucm[c] - rel[c] * mp / msum;
    }
}

/* dsmcCloud::postCollisionVibrationalEnergyLevel (dsmcCloud.C:1390-1531);
 * post = postReaction: the level is always redrawn, with no Zv test.
 * t_macro: the cell's overall temperature, which inverseZvFormulation 2008
 * (inv_zv_form 1) uses in Zv when it exceeds SMALL (dsmcCloud.C:1440-1455,
 * fields().overallT); otherwise, and in the other formulations, the
 * quantised collision temperature */
static int vib_level(const hs_config* cfg, const or_derived* d, int t, int partner, int mode,
                     int level, int imax, double ec, double om, int post, or_rng* r, double t_macro)
{
    const hs_species* sp = &cfg->species[t];
    const double theta_v = sp->theta_v[mode];
    double inv_zv = 1.0;
    if (post) {
        int idash, guard = 0;
        double func;
        do {
            idash = rng_label(r, 0, imax);
            const double evib = (double)idash * HS_KB * theta_v;
            func = or_pow(1.0 - evib / ec, 1.5 - om);
        } while (func < rng_u(r) && ++guard < OR_MAX_REDRAW);
        return idash;
    }
    if (cfg->z_vib_fixed == 0.0) {
        const double T = (cfg->inv_zv_form == 1 && t_macro > OR_SMALL) ? t_macro
                                                                       : (double)imax * theta_v / (3.5 - om);
        const double pow1 = or_pow(sp->theta_d / T, 1.0 / 3.0) - 1.0;
        const double zvp1 = or_pow(sp->theta_d / T, om);
        const double zvp2 = or_pow(d->zv_b[t][partner][mode], pow1 / d->zv_pow2[t][mode]);
        const double zv = zvp1 * zvp2;
        inv_zv = cfg->inv_zv_form == 2 ? 1.0 / (5.0 * zv) : 1.0 / zv;
    } else {
        inv_zv = 1.0 / cfg->z_vib_fixed;
    }
    int idash = level;
    if (inv_zv > rng_u(r)) {
        double func;
        int guard = 0;
        do {
            idash = rng_label(r, 0, imax);
            const double evib = (double)idash * HS_KB * theta_v;
            func = or_pow(1.0 - evib / ec, 1.5 - om);
        } while (func < rng_u(r) && ++guard < OR_MAX_REDRAW);
    }
    return idash;
}

/* dsmcCloud::postCollisionRotationalEnergy (dsmcCloud.C:1327-1387) */
static double rot_ratio(double dof, double chib, or_rng* r)
{
    if (dof == 2.0) return 1.0 - or_pow(rng_u(r), 1.0 / chib);
    const double a1 = 0.5 * dof - 1.0, b1 = chib - 1.0;
    if (a1 < OR_SMALL && b1 < OR_SMALL) return rng_u(r);
    double ratio, P;
    int guard = 0;
    do {
        ratio = rng_u(r);
        if (a1 < OR_SMALL) P = or_pow(1.0 - ratio, b1);
        else if (b1 < OR_SMALL) P = or_pow(1.0 - ratio, a1);
        else P = or_pow((a1 + b1) * ratio / a1, a1) * or_pow((a1 + b1) * (1.0 - ratio) / b1, b1);
    } while (P < rng_u(r) && ++guard < OR_MAX_REDRAW);
    return ratio;
}

/* LarsenBorgnakkeVariableHardSphere::redistribute (LB-VHS.C:150-279) for a
 * species with a single electronic level; post = postReaction */
static void lb_redistribute(const hs_config* cfg, const or_derived* d, int t, int partner,
                            double* et, double* erot, int32_t* vib, double om, int post, or_rng* r,
                            double t_macro)
{
    const hs_species* sp = &cfg->species[t];
    if (sp->charge == -1) return;
    /* electronic mode: one level, the Liechty loop takes exactly one draw */
    if (1.0 / cfg->z_el > rng_u(r)) (void)rng_u(r);
    if (sp->n_vib_modes > 0) {
        double pre[HS_MAX_VIB_MODES];
        for (int m = 0; m < sp->n_vib_modes; ++m) pre[m] = (double)vib[m] * HS_KB * sp->theta_v[m];
        for (int m = 0; m < sp->n_vib_modes; ++m) {
            const double ec = *et + pre[m];
            const int imax = (int)(ec / (HS_KB * sp->theta_v[m]));
            if (imax > 0) {
                vib[m] = vib_level(cfg, d, t, partner, m, vib[m], imax, ec, om, post, r, t_macro);
                *et = ec - (double)vib[m] * HS_KB * sp->theta_v[m];
            }
        }
    }
    if (sp->rot_dof > 0) {
        const double pre = *erot;
        if (1.0 / cfg->z_rot > rng_u(r)) {
            const double ec = *et + pre;
            const double ratio = rot_ratio(sp->rot_dof, 2.5 - om, r);
            *erot = ratio * ec;
            *et = ec - *erot;
        }
    }
}

/* BinaryCollisionModel::collide for the configured model */
static void binary_collide(const hs_config* cfg, const or_derived* d, int tp, int tq,
                           double up[3], double uq[3], double* erp, double* erq,
                           int32_t* vp, int32_t* vq, or_rng* r, double t_macro)
{
    const int model = cfg->collision_model;
    if (model == HS_COLL_VHS || model == HS_COLL_VSS) {
        post_coll_vel(cfg, d, model == HS_COLL_VSS, 0, tp, tq, up, uq, -1.0, r);
        return;
    }
    /* Larsen-Borgnakke (LB-VHS.C:110-147) */
    const double mp = cfg->species[tp].mass, mq = cfg->species[tq].mass;
    const double mr = mp * mq / (mp + mq);
    const double dx = up[0] - uq[0], dy = up[1] - uq[1], dz = up[2] - uq[2];
    const double cr2 = dx * dx + dy * dy + dz * dz;
    double et = 0.5 * mr * cr2;
    const double om = d->omega_pq[tp][tq];
    lb_redistribute(cfg, d, tp, tq, &et, erp, vp, om, 0, r, t_macro);
    lb_redistribute(cfg, d, tq, tp, &et, erq, vq, om, 0, r, t_macro);
    const double cr = sqrt(2.0 * et / mr);
    post_coll_vel(cfg, d, model == HS_COLL_LB_VSS, 1, tp, tq, up, uq, cr, r);
}

/* ------------------------------------------------------------------------ */
/* quantum-kinetic chemistry (src/lagrangian/dsmc/reactions)                 */
/* ------------------------------------------------------------------------ */
/* one parcel's state during a reaction */
typedef struct {
    int t;
    double u[3];
    double er;
    int32_t v[HS_MAX_VIB_MODES];
} or_pstate;

static int is_lb(const hs_config* cfg)
{
    return cfg->collision_model == HS_COLL_LB_VHS || cfg->collision_model == HS_COLL_LB_VSS;
}

/* eVib_tot (dsmcParcelI.H:389-398) */
static double evib_total(const hs_config* cfg, const or_pstate* a)
{
    const hs_species* sp = &cfg->species[a->t];
    double e = 0.0;
    for (int m = 0; m < sp->n_vib_modes; ++m) e += (double)a->v[m] * HS_KB * sp->theta_v[m];
    return e;
}

static void become(or_pstate* a, int t)
{
    a->t = t;
    a->er = 0.0;
    for (int m = 0; m < HS_MAX_VIB_MODES; ++m) a->v[m] = 0;
}

/* dissociationQK::testDissociation (dissociationQK.C:199-238) */
static void test_dissociation(const hs_config* cfg, const or_pstate* a, double et, int* mode, double* ec,
                              double* total, double* prob)
{
    const hs_species* sp = &cfg->species[a->t];
    const int kind = species_kind(sp);
    if (kind != 20 && kind != 30) return;
    for (int m = 0; m < sp->n_vib_modes; ++m) {
        const double theta_v = sp->theta_v[m];
        const double ev = (double)a->v[m] * HS_KB * theta_v;
        const int id = (int)(sp->theta_d / theta_v);
        *ec = et + ev;
        const int imax = (int)(*ec / (HS_KB * theta_v));
        if (imax > id) {
            *prob = 1.0;
            *total += *prob;
            *mode = m;
            break;
        }
    }
}

/* exchangeQK::testExchange (exchangeQK.C:180-267); a is the molecule */
static void test_exchange(const hs_config* cfg, const or_rx* x, const or_pstate* a, double et, double om,
                          double* ec, double* total, double* prob)
{
    const hs_species* sp = &cfg->species[a->t];
    const double chib = 2.5 - om;
    const double tcoll = et / (HS_KB * chib);
    double ea = x->a_dash * or_pow(tcoll / 273.0, x->b) * fabs(x->heat_x);
    double sum = 1.0;
    if (x->heat_x < 0.0) ea -= x->heat_x;
    int m = 0;
    do {
        const double kth = HS_KB * sp->theta_v[m];
        const double ev = (double)a->v[m] * HS_KB * sp->theta_v[m];
        *ec = et + ev;
        if (*ec > ea) {
            if (ea > kth) {
                sum = 0.0;
                const int ia = (int)(*ec / kth);
                for (int i = 0; i <= ia; ++i)
                    sum += or_pow(1.0 - (double)i * HS_KB * sp->theta_v[m] / *ec, 1.5 - om);
            }
            *prob = or_pow(1.0 - ea / *ec, 1.5 - om) / sum;
            m = sp->n_vib_modes;
        }
        m += 1;
    } while (m < sp->n_vib_modes);
    *total += *prob;
}

/* VariableHardSphere::postReactionVelocities (VariableHardSphere.C:231-262):
 * ua enters as the pair's centre-of-mass velocity */
static void post_reaction_vel(const hs_config* cfg, int ta, int tb, double ua[3], double ub[3], double cr,
                              or_rng* r)
{
    const double ma = cfg->species[ta].mass, mb = cfg->species[tb].mass;
    const double cos_t = 2.0 * rng_u(r) - 1.0;
    const double sin_t = sqrt(1.0 - cos_t * cos_t);
    double sp, cp;
    or_sincos2pi(rng_u(r), &sp, &cp);
    const double rel[3] = { cr * cos_t, cr * (sin_t * cp), cr * (sin_t * sp) };
    for (int c = 0; c < 3; ++c) {
        ub[c] = ua[c] - rel[c] * ma / (ma + mb);
        ua[c] = ua[c] + rel[c] * mb / (ma + mb);
    }
}

/* dissociationQK::dissociateParticleByPartner (dissociationQK.C:240-378): a
 * (reactant nr) dissociates in vibrational mode `mode`, b is the partner;
 * a becomes the first product and *prod the second.  Returns 1 when the
 * split happens (allowSplitting) */
static int dissociate(const hs_config* cfg, const or_derived* d, const or_rx* x, or_pstate* a, or_pstate* b,
                      int nr, int mode, double ec, or_rng* r, or_pstate* prod)
{
    if (!x->split) return 0;
    ec -= x->heat_diss[nr];
    const double om = d->omega_pq[a->t][b->t];
    if (is_lb(cfg)) lb_redistribute(cfg, d, b->t, a->t, &ec, &b->er, b->v, om, 1, r, 0.0);
    const double ma = cfg->species[a->t].mass, mb = cfg->species[b->t].mass;
    const double mr = ma * mb / (ma + mb);
    const double cr = sqrt(2.0 * ec / mr);
    post_coll_vel(cfg, d, 0, 0, a->t, b->t, a->u, b->u, cr, r);
    const int t1 = x->dprod[nr][0], t2 = x->dprod[nr][1];
    const double m1 = cfg->species[t1].mass, m2 = cfg->species[t2].mass;
    const double mrp = m1 * m2 / (m1 + m2);
    const double evt = evib_total(cfg, a);
    const double evm = (double)a->v[mode] * HS_KB * cfg->species[a->t].theta_v[mode];
    const double evn = evt - evm;
    const double left = a->er + evn + 0.0;  /* + the single electronic level's energy, 0 */
    const double crp = sqrt(2.0 * left / mrp);
    become(prod, t2);
    post_reaction_vel(cfg, t1, t2, a->u, prod->u, crp, r);
    become(a, t1);
    return 1;
}

/* exchangeQK::exchange (exchangeQK.C:269-370): a is the molecule, b the atom */
static int exchange(const hs_config* cfg, const or_derived* d, const or_rx* x, or_pstate* a, or_pstate* b,
                    or_rng* r)
{
    if (!x->split) return 0;
    const double ma = cfg->species[a->t].mass, mb = cfg->species[b->t].mass;
    const double mr = ma * mb / (ma + mb);
    const double dx = a->u[0] - b->u[0], dy = a->u[1] - b->u[1], dz = a->u[2] - b->u[2];
    double et = 0.5 * mr * (dx * dx + dy * dy + dz * dz);
    double ucm[3];
    for (int c = 0; c < 3; ++c) ucm[c] = (ma * a->u[c] + mb * b->u[c]) / (ma + mb);
    const double mpx = cfg->species[x->xatom].mass, mqx = cfg->species[x->xmol].mass;
    const double mrx = mpx * mqx / (mpx + mqx);
    const double evt = evib_total(cfg, a);
    et += a->er + evt + 0.0 + 0.0 + x->heat_x;  /* + both electronic energies, 0 */
    become(a, x->xatom);
    become(b, x->xmol);
    if (is_lb(cfg)) lb_redistribute(cfg, d, b->t, a->t, &et, &b->er, b->v, x->om_x, 1, r, 0.0);
    const double crx = sqrt(2.0 * et / mrx);
    for (int c = 0; c < 3; ++c) a->u[c] = ucm[c];
    post_reaction_vel(cfg, x->xatom, x->xmol, a->u, b->u, crx, r);
    return 1;
}

/* dsmcReaction::decreasing_sort_indices (dsmcReaction.C:176-205): n draws,
 * then the indices by probability, equal probabilities by their draw, both
 * decreasing (insertion sort: the order is unique unless two draws tie) */
static void decreasing_order(const double* v, int n, or_rng* r, int* idx)
{
    double rnd[3];
    for (int i = 0; i < n; ++i) {
        rnd[i] = rng_u(r);
        idx[i] = i;
    }
    for (int i = 1; i < n; ++i) {
        const int y = idx[i];
        int j = i - 1;
        while (j >= 0 && (v[y] > v[idx[j]] || (v[y] == v[idx[j]] && rnd[y] > rnd[idx[j]]))) {
            idx[j + 1] = idx[j];
            --j;
        }
        idx[j + 1] = y;
    }
}

/* kinetic energy of the relative motion, 1/2 mr |ua - ub|^2 */
static double rel_energy(const hs_config* cfg, const or_pstate* a, const or_pstate* b)
{
    const double ma = cfg->species[a->t].mass, mb = cfg->species[b->t].mass;
    const double mr = ma * mb / (ma + mb);
    const double dx = a->u[0] - b->u[0], dy = a->u[1] - b->u[1], dz = a->u[2] - b->u[2];
    return 0.5 * mr * (dx * dx + dy * dy + dz * dz);
}

/* The reaction of an accepted pair with a reaction model (noTimeCounter.C:
 * 250-300): returns 1 if it reacted (relax_ is off and no binary collision
 * follows).  A dissociation sets *prod to the new parcel and *from_q to
 * whether Q (1) or P (0) dissociated.  p is the candidate P, q the
 * candidate Q; the models swap them when P is not their first reactant. */
static int react_pair(const hs_config* cfg, const or_derived* d, int ri, or_pstate* P, or_pstate* Q, or_rng* r,
                      or_pstate* prod, int* from_q)
{
    const or_rx* x = &d->rx[ri];
    or_pstate *p = P, *q = Q;
    prod->t = -1;
    if (x->model == HS_REACT_EXCHANGE_QK) {
        /* exchangeQK::reaction (exchangeQK.C:457-517) */
        const int kp = species_kind(&cfg->species[p->t]);
        if (kp == 10 || kp == 11) { p = Q; q = P; }
        const double om = d->omega_pq[p->t][q->t];
        const double et = rel_energy(cfg, p, q);
        double total = 0.0, prob = 0.0, ec = 0.0;
        test_exchange(cfg, x, p, et, om, &ec, &total, &prob);
        if (total > rng_u(r)) return exchange(cfg, d, x, p, q, r);
        return 0;
    }
    /* dissociationQK::reaction (dissociationQK.C:469-600) and
     * dissociationExchangeQK::reaction (dissociationExchangeQK.C:98-270) */
    if (p->t != x->r[0]) { p = Q; q = P; }
    const int q_is_Q = q == Q;
    const int three = x->model == HS_REACT_DISSOCIATION_EXCHANGE_QK;
    const double om = d->omega_pq[p->t][q->t];
    const double et = rel_energy(cfg, p, q);
    double total = 0.0, prob[3] = { 0.0, 0.0, 0.0 }, ec[3] = { 0.0, 0.0, 0.0 };
    int mode_p = -1, mode_q = -1;
    test_dissociation(cfg, p, et, &mode_p, &ec[0], &total, &prob[0]);
    if (three || p->t != q->t) test_dissociation(cfg, q, et, &mode_q, &ec[1], &total, &prob[1]);
    if (three) {
        if (x->pos_mol == 0) test_exchange(cfg, x, p, et, om, &ec[2], &total, &prob[2]);
        else test_exchange(cfg, x, q, et, om, &ec[2], &total, &prob[2]);
    }
    if (!(total > rng_u(r))) return 0;
    const int n = three ? 3 : 2;
    double np[3];
    int idx[3];
    for (int i = 0; i < n; ++i) np[i] = prob[i] / total;
    decreasing_order(np, n, r, idx);
    double cum = 0.0;
    for (int k = 0; k < n; ++k) {
        const int i = idx[k];
        if (!(np[i] > OR_SMALL)) break;
        cum += np[i];
        if (cum > rng_u(r)) {
            if (i == 0) {
                *from_q = !q_is_Q;
                return dissociate(cfg, d, x, p, q, 0, mode_p, ec[0], r, prod);
            }
            if (i == 1) {
                *from_q = q_is_Q;
                return dissociate(cfg, d, x, q, p, 1, mode_q, ec[1], r, prod);
            }
            return x->pos_mol == 0 ? exchange(cfg, d, x, p, q, r) : exchange(cfg, d, x, q, p, r);
        }
    }
    return 0;
}

/* The accepted collision of one pair (noTimeCounter.C:250-300): its reaction
 * when the pair has a model, else (or when it did not react) the binary
 * collision.  Returns 1 if it reacted. */
static int collide_or_react(const hs_config* cfg, const or_derived* d, or_pstate* P, or_pstate* Q, or_rng* r,
                            or_pstate* prod, int* from_q, double t_macro)
{
    prod->t = -1;
    const int ri = d->n_react ? d->react_of[P->t][Q->t] : -1;
    if (ri >= 0 && react_pair(cfg, d, ri, P, Q, r, prod, from_q)) return 1;
    binary_collide(cfg, d, P->t, Q->t, P->u, Q->u, &P->er, &Q->er, P->v, Q->v, r, t_macro);
    return 0;
}

int or_sigma_tcr(const hs_config* cfg, int64_t n, const int32_t* tp, const int32_t* tq,
                 const double* up, const double* uq, double* out)
{
    or_derived d;
    derive(cfg, &d);
    for (int64_t k = 0; k < n; ++k) out[k] = sigma_tcr(&d, tp[k], tq[k], up + 3 * k, uq + 3 * k);
    return 0;
}

int or_binary_collide(const hs_config* cfg, int64_t n, const int32_t* tp, const int32_t* tq,
                      double* up, double* uq, double* erot_p, double* erot_q,
                      int32_t* vib_p, int32_t* vib_q, const double* uniforms,
                      int32_t n_uniform, int32_t* used_out, const double* t_macro)
{
    or_derived d;
    derive(cfg, &d);
    int rc = 0;
    for (int64_t k = 0; k < n; ++k) {
        or_rng r;
        rng_init(&r, 0, 0, 0, 0, 0);
        r.inj = uniforms + k * n_uniform;
        r.inj_n = n_uniform;
        double erp = erot_p ? erot_p[k] : 0.0, erq = erot_q ? erot_q[k] : 0.0;
        int32_t* vp = vib_p ? vib_p + k * HS_MAX_VIB_MODES : NULL;
        int32_t* vq = vib_q ? vib_q + k * HS_MAX_VIB_MODES : NULL;
        int32_t zp[HS_MAX_VIB_MODES] = { 0 }, zq[HS_MAX_VIB_MODES] = { 0 };
        binary_collide(cfg, &d, tp[k], tq[k], up + 3 * k, uq + 3 * k, &erp, &erq,
                       vp ? vp : zp, vq ? vq : zq, &r, t_macro ? t_macro[k] : 0.0);
        if (erot_p) erot_p[k] = erp;
        if (erot_q) erot_q[k] = erq;
        if (used_out) used_out[k] = r.inj_used;
        if (r.inj_used > n_uniform) rc = HS_E_OVERFLOW;
    }
    return rc;
}

int or_react_pairs(const hs_config* cfg, int64_t n, int32_t* tp, int32_t* tq, double* up, double* uq,
                   double* erot_p, double* erot_q, int32_t* vib_p, int32_t* vib_q, const double* uniforms,
                   int32_t n_uniform, int32_t* used_out, int32_t* reacted_out, int32_t* prod_type,
                   double* prod_u, int32_t* prod_from_q)
{
    or_derived d;
    derive(cfg, &d);
    int rc = 0;
    for (int64_t k = 0; k < n; ++k) {
        or_rng r;
        rng_init(&r, 0, 0, 0, 0, 0);
        r.inj = uniforms + k * n_uniform;
        r.inj_n = n_uniform;
        or_pstate a = { tp[k], { up[3 * k], up[3 * k + 1], up[3 * k + 2] }, erot_p[k], { 0 } };
        or_pstate b = { tq[k], { uq[3 * k], uq[3 * k + 1], uq[3 * k + 2] }, erot_q[k], { 0 } };
        for (int m = 0; m < HS_MAX_VIB_MODES; ++m) {
            a.v[m] = vib_p[k * HS_MAX_VIB_MODES + m];
            b.v[m] = vib_q[k * HS_MAX_VIB_MODES + m];
        }
        or_pstate prod;
        int from_q = 0;
        const int reacted = collide_or_react(cfg, &d, &a, &b, &r, &prod, &from_q, 0.0);
        tp[k] = a.t;
        tq[k] = b.t;
        for (int c = 0; c < 3; ++c) {
            up[3 * k + c] = a.u[c];
            uq[3 * k + c] = b.u[c];
            prod_u[3 * k + c] = prod.t >= 0 ? prod.u[c] : 0.0;
        }
        erot_p[k] = a.er;
        erot_q[k] = b.er;
        for (int m = 0; m < HS_MAX_VIB_MODES; ++m) {
            vib_p[k * HS_MAX_VIB_MODES + m] = a.v[m];
            vib_q[k * HS_MAX_VIB_MODES + m] = b.v[m];
        }
        reacted_out[k] = reacted;
        prod_type[k] = prod.t;
        prod_from_q[k] = prod.t >= 0 ? from_q : -1;
        used_out[k] = r.inj_used;
        if (r.inj_used > n_uniform) rc = HS_E_OVERFLOW;
    }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* move (dsmcParcel::move + tracker on the Cartesian block)                 */
/* ------------------------------------------------------------------------ */
static double face_coord(const hs_config* cfg, int a, int idx)
{
    return cfg->origin[a] + (double)idx * cfg->delta[a];
}

/* dsmcDiffuseWallPatch::performDiffuseReflection (dsmcDiffuseWallPatch.C:67-139)
 * with dsmcPatchBoundary::calculateWallUnitVectors (dsmcPatchBoundary.C:216-261).
 * nw is the wall unit normal pointing out of the gas. */
static void diffuse_reflect(const hs_config* cfg, const hs_patch* pt, const double nw[3], int t,
                            double U[3], float* erot, uint16_t* vib, int64_t vstride, or_rng* r)
{
    const hs_species* sp = &cfg->species[t];
    double und = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
    double ut[3] = { U[0] - und * nw[0], U[1] - und * nw[1], U[2] - und * nw[2] };
    int guard = 0;
    while (sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]) < OR_SMALL && guard++ < 100) {
        U[0] = U[0] * (0.8 + 0.2 * rng_u(r));
        U[1] = U[1] * (0.8 + 0.2 * rng_u(r));
        U[2] = U[2] * (0.8 + 0.2 * rng_u(r));
        und = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
        ut[0] = U[0] - und * nw[0]; ut[1] = U[1] - und * nw[1]; ut[2] = U[2] - und * nw[2];
    }
    const double mut = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
    double tw1[3] = { ut[0] / mut, ut[1] / mut, ut[2] / mut };
    if (!(mut > 0.0)) { tw1[0] = nw[1]; tw1[1] = nw[2]; tw1[2] = nw[0]; }
    const double tw2[3] = { nw[1] * tw1[2] - nw[2] * tw1[1],
                            nw[2] * tw1[0] - nw[0] * tw1[2],
                            nw[0] * tw1[1] - nw[1] * tw1[0] };
    const double g1 = rng_gauss(r);
    const double g2 = rng_gauss(r);
    const double u = rng_u(r);
    const double T = pt->temperature;
    const double sc = sqrt(HS_KB * T / sp->mass);
    double arg = 1.0 - u;
    if (arg < OR_VSMALL) arg = OR_VSMALL;
    const double rn = sqrt(-2.0 * or_log(arg));
    for (int c = 0; c < 3; ++c)
        U[c] = sc * ((g1 * tw1[c] + g2 * tw2[c]) - rn * nw[c]) + pt->velocity[c];
    if (erot) *erot = (float)equip_erot(r, T, sp->rot_dof);
    if (vib) equip_vib(r, T, sp, vib, vstride);
}

/* ------------------------------------------------------------------------ */
/* wall-face sampling (dsmcPatchBoundary::measurePropertiesBeforeControl /  */
/* AfterControl, dsmcPatchBoundary.C:263-483)                               */
/* ------------------------------------------------------------------------ */
static int cmp_double(const void* a, const void* b)
{
    const double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

double or_arc_in_rect(double cx, double cy, double r, double x0, double x1, double y0, double y1, double n_out[2])
{
    /* the circle's crossings of the four lines split [0, 2pi) into arcs that
     * lie wholly inside or wholly outside the rectangle */
    double th[10];
    int k = 0;
    th[k++] = 0.0;
    th[k++] = OR_TWO_PI;
    const double xs[2] = { x0, x1 }, ys[2] = { y0, y1 };
    for (int j = 0; j < 2; ++j) {
        const double u = (xs[j] - cx) / r;
        if (u > -1.0 && u < 1.0) {
            const double a = acos(u);
            th[k++] = a;
            th[k++] = OR_TWO_PI - a;
        }
        const double v = (ys[j] - cy) / r;
        if (v > -1.0 && v < 1.0) {
            double a = asin(v);
            th[k++] = OR_PI - a;
            if (a < 0.0) a += OR_TWO_PI;
            th[k++] = a;
        }
    }
    qsort(th, (size_t)k, sizeof(double), cmp_double);
    double len = 0.0, nx = 0.0, ny = 0.0;
    for (int j = 0; j + 1 < k; ++j) {
        const double dth = th[j + 1] - th[j];
        if (!(dth > 0.0)) continue;
        const double m = 0.5 * (th[j] + th[j + 1]);
        const double px = cx + r * cos(m), py = cy + r * sin(m);
        if (px >= x0 && px <= x1 && py >= y0 && py <= y1) {
            len += r * dth;
            nx -= r * dth * cos(m);
            ny -= r * dth * sin(m);
        }
    }
    if (n_out) {
        const double nm = sqrt(nx * nx + ny * ny);
        n_out[0] = nm > 0.0 ? nx / nm : 0.0;
        n_out[1] = nm > 0.0 ? ny / nm : 0.0;
    }
    return len;
}

static int side_measured(const hs_config* cfg, int side)
{
    const int m = cfg->patch[side].model;
    return m == HS_PATCH_DIFFUSE || m == HS_PATCH_SPECULAR;
}

/* wall-face layout: side faces first (first tangential index fastest), then
 * one face per cell crossed by the cylinder wall */
typedef struct {
    int64_t n;
    int64_t off[6];    /* first face of each measured side, -1 otherwise */
    int32_t* cyl_face; /* [nc] or NULL */
    double* area;      /* [n] */
    int64_t cyl_first; /* first cylinder face in the wall-face list */
    int measure;       /* wall sums on */
    /* face tracker (dsmcFaceTracker) */
    int64_t* flux;     /* [ns][n_faces] or NULL */
    int64_t n_faces, face_off[4];
} or_wall;

/* sign() of OpenFOAM: +1 for x >= 0 */
static int of_sign(double x) { return x >= 0.0 ? 1 : -1; }

/* face-tracker face of plane k on axis a at tangential cell indices of idx */
static int64_t mesh_face(const hs_config* cfg, const or_wall* w, int a, int k, const int idx[3])
{
    const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
    const int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
    return w->face_off[a] + k + (int64_t)(nn[a] + 1) * (idx[a1] + (int64_t)nn[a1] * idx[a2]);
}

static void track_face(const or_wall* w, int t, int64_t face, int64_t v)
{
    OR_ATOMIC w->flux[(int64_t)t * w->n_faces + face] += v;
}

int64_t or_face_count(const hs_config* cfg)
{
    const int64_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    int64_t n = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
    if (cfg->cyl_enabled) {
        const int64_t nw = or_wall_faces(cfg, NULL, NULL, NULL, NULL);
        int64_t ns = 0;
        for (int sd = 0; sd < 6; ++sd)
            if (cfg->patch[sd].model == HS_PATCH_DIFFUSE || cfg->patch[sd].model == HS_PATCH_SPECULAR) {
                const int axis = sd / 2;
                const int64_t nn[3] = { nx, ny, nz };
                ns += nn[axis == 0 ? 1 : 0] * nn[axis == 2 ? 1 : 2];
            }
        n += nw - ns;
    }
    return n;
}

int64_t or_wall_faces(const hs_config* cfg, int32_t* side, int32_t* cell, double* area, double* normal)
{
    const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
    int64_t f = 0;
    for (int sd = 0; sd < 6; ++sd) {
        if (!side_measured(cfg, sd)) continue;
        const int axis = sd / 2, dir = (sd & 1) ? 1 : -1;
        const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
        for (int j2 = 0; j2 < nn[a2]; ++j2)
            for (int j1 = 0; j1 < nn[a1]; ++j1, ++f) {
                int idx[3];
                idx[axis] = dir > 0 ? nn[axis] - 1 : 0;
                idx[a1] = j1;
                idx[a2] = j2;
                if (side) side[f] = sd;
                if (cell) cell[f] = idx[0] + cfg->nx * (idx[1] + cfg->ny * idx[2]);
                if (area) area[f] = cfg->delta[a1] * cfg->delta[a2];
                if (normal) {
                    normal[3 * f] = normal[3 * f + 1] = normal[3 * f + 2] = 0.0;
                    normal[3 * f + axis] = (double)dir;
                }
            }
    }
    if (cfg->cyl_enabled) {
        for (int k = 0; k < cfg->nz; ++k)
            for (int j = 0; j < cfg->ny; ++j)
                for (int i = 0; i < cfg->nx; ++i) {
                    const double x0 = face_coord(cfg, 0, i), x1 = face_coord(cfg, 0, i + 1);
                    const double y0 = face_coord(cfg, 1, j), y1 = face_coord(cfg, 1, j + 1);
                    double nw[2];
                    const double len = or_arc_in_rect(cfg->cyl_center[0], cfg->cyl_center[1], cfg->cyl_radius,
                                                      x0, x1, y0, y1, nw);
                    /* arcs shorter than 1e-6 of the cell (slivers where the circle
                     * grazes a grid line) carry no face */
                    if (!(len > 1e-6 * (cfg->delta[0] < cfg->delta[1] ? cfg->delta[0] : cfg->delta[1]))) continue;
                    if (side) side[f] = 6;
                    if (cell) cell[f] = i + cfg->nx * (j + cfg->ny * k);
                    if (area) area[f] = len * cfg->delta[2];
                    if (normal) { normal[3 * f] = nw[0]; normal[3 * f + 1] = nw[1]; normal[3 * f + 2] = 0.0; }
                    ++f;
                }
    }
    return f;
}

static void wall_layout(const hs_config* cfg, or_wall* w)
{
    memset(w, 0, sizeof(*w));
    const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
    int64_t f = 0;
    for (int sd = 0; sd < 6; ++sd) {
        w->off[sd] = -1;
        if (!side_measured(cfg, sd)) continue;
        const int axis = sd / 2;
        const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
        w->off[sd] = f;
        f += (int64_t)nn[a1] * nn[a2];
    }
    w->cyl_first = f;
    w->n = or_wall_faces(cfg, NULL, NULL, NULL, NULL);
    w->area = (double*)malloc((size_t)(w->n > 0 ? w->n : 1) * sizeof(double));
    int32_t* cell = (int32_t*)malloc((size_t)(w->n > 0 ? w->n : 1) * sizeof(int32_t));
    or_wall_faces(cfg, NULL, cell, w->area, NULL);
    if (cfg->cyl_enabled) {
        const int64_t nc = (int64_t)cfg->nx * cfg->ny * cfg->nz;
        w->cyl_face = (int32_t*)malloc((size_t)nc * sizeof(int32_t));
        for (int64_t c = 0; c < nc; ++c) w->cyl_face[c] = -1;
        for (int64_t g = f; g < w->n; ++g) w->cyl_face[cell[g]] = (int32_t)g;
    }
    free(cell);
}

static void wall_free(or_wall* w)
{
    free(w->area);
    free(w->cyl_face);
}

/* the flux-weighted terms of one side of a hit (before or after the
 * interaction, dsmcPatchBoundary.C:263-357 and :359-426); returns the
 * parcel's total energy */
static double wall_side_terms(const hs_config* cfg, const or_wall* w, double* sums, int64_t f, int t,
                              const double nw[3], const double U[3], const float* erot, const uint16_t* vib,
                              int64_t vstride)
{
    const hs_species* sp = &cfg->species[t];
    const double m = sp->mass, adt = w->area[f] * cfg->dt;
    const double un = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
    double dn = fabs(un) * adt;
    if (dn < OR_SMALL) dn = OR_SMALL;
    const double wt = 1.0 / dn;
    const double ke = 0.5 * m * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    const int64_t st = w->n;
    double* base = sums + (int64_t)t * HS_WS_COUNT * st + f;
    OR_ATOMIC base[HS_WS_RHON * st] += wt;
    OR_ATOMIC base[HS_WS_KE * st] += ke * wt;
    for (int c = 0; c < 3; ++c) {
        OR_ATOMIC base[(HS_WS_MOMX + c) * st] += m * (U[c] - un * nw[c]) * wt;
    }
    double e = ke;
    if (sp->rot_dof > 0) {
        const double er = (double)*erot;
        OR_ATOMIC base[HS_WS_EROT * st] += er * wt;
        e += er;
    }
    for (int k = 0; k < sp->n_vib_modes; ++k) {
        const double ev = (double)vib[k * vstride] * HS_KB * sp->theta_v[k];
        OR_ATOMIC base[(HS_WS_EVIB0 + k) * st] += ev * wt;
        e += ev;
    }
    return e;
}

/* heat flux and force density of one hit (dsmcPatchBoundary.C:428-482) */
static void wall_exchange_terms(const hs_config* cfg, const or_wall* w, double* sums, int64_t f, int t, double ea,
                                const double Ua[3], double eb, const double Ub[3])
{
    const double m = cfg->species[t].mass;
    const double cq = cfg->fn / (w->area[f] * cfg->dt);
    const int64_t st = w->n;
    double* base = sums + (int64_t)t * HS_WS_COUNT * st + f;
    OR_ATOMIC base[HS_WS_Q * st] += cq * (ea - eb);
    for (int c = 0; c < 3; ++c) {
        OR_ATOMIC base[(HS_WS_FDX + c) * st] += cq * (m * Ua[c] - m * Ub[c]);
    }
}

static void move_one(const hs_config* cfg, const or_derived* d, or_state* s, const or_wall* w, int64_t i,
                     int is_new)
{
    or_rng r;
    rng_init(&r, cfg->seed, OR_RNG_MOVE, (uint32_t)i, s->step, 0);
    hs_particles* P = &s->p;
    double t = cfg->dt;
    if (is_new) {
        const double sf = rng_u(&r);
        t = (1.0 - sf) * cfg->dt;
    }
    double p[3] = { P->px[i], P->py[i], P->pz[i] };
    double U[3] = { (double)P->ux[i], (double)P->uy[i], (double)P->uz[i] };
    int32_t c = P->cell[i];
    if (c < 0) return;
    int idx[3] = { c % cfg->nx, (c / cfg->nx) % cfg->ny, c / (cfg->nx * cfg->ny) };
    const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
    const int t_id = P->type ? P->type[i] : 0;
    float* erot = P->erot ? &P->erot[i] : NULL;
    uint16_t* vib = (P->vib && cfg->species[t_id].n_vib_modes > 0) ? &P->vib[i] : NULL;
    for (int a = 0; a < 3; ++a)
        if (d->empty[a]) p[a] = d->centre[a];
    int keep = 1, iter = 0;
    while (keep && t > OR_ROOTVSMALL) {
        if (++iter > OR_MAX_TRACK) {
            OR_ATOMIC s->counters[5] += 1;
            break;
        }
        double ut[3], e[3];
        for (int a = 0; a < 3; ++a) {
            ut[a] = d->empty[a] ? 0.0 : U[a];
            e[a] = p[a] + t * ut[a];
        }
        double lam = 1.0;
        int axis = -1, dir = 0;
        for (int a = 0; a < 3; ++a) {
            const double den = e[a] - p[a];
            if (ut[a] > 0.0 && den > 0.0) {
                double l = (face_coord(cfg, a, idx[a] + 1) - p[a]) / den;
                if (l < 0.0) l = 0.0;
                if (l < lam) { lam = l; axis = a; dir = 1; }
            } else if (ut[a] < 0.0 && den < 0.0) {
                double l = (face_coord(cfg, a, idx[a]) - p[a]) / den;
                if (l < 0.0) l = 0.0;
                if (l < lam) { lam = l; axis = a; dir = -1; }
            }
        }
        int wall = 0;
        if (cfg->cyl_enabled) {
            const double dx = e[0] - p[0], dy = e[1] - p[1];
            const double rx = p[0] - cfg->cyl_center[0], ry = p[1] - cfg->cyl_center[1];
            const double b = rx * dx + ry * dy;
            if (b < 0.0) {
                const double aa = dx * dx + dy * dy;
                const double cc = rx * rx + ry * ry - cfg->cyl_radius * cfg->cyl_radius;
                const double disc = b * b - aa * cc;
                if (disc >= 0.0) {
                    const double q = sqrt(disc) - b;
                    double sh = cc / q;
                    if (sh < 0.0) sh = 0.0;
                    if (sh < lam) { lam = sh; wall = 1; }
                }
            }
        }
        if (!wall && axis < 0) {
            p[0] = e[0]; p[1] = e[1]; p[2] = e[2];
            t = 0.0;
            break;
        }
        for (int b = 0; b < 3; ++b) p[b] = p[b] + lam * (e[b] - p[b]);
        t = t - t * lam;
        if (wall) {
            OR_ATOMIC s->counters[4] += 1;
            double nx_ = cfg->cyl_center[0] - p[0], ny_ = cfg->cyl_center[1] - p[1];
            const double nm = sqrt(nx_ * nx_ + ny_ * ny_);
            const double nw[3] = { nx_ / nm, ny_ / nm, 0.0 };
            const int32_t wf = w->cyl_face ? w->cyl_face[idx[0] + cfg->nx * (idx[1] + cfg->ny * idx[2])] : -1;
            const int32_t f = w->measure ? wf : -1;
            double e_pre = 0.0;
            const double u_pre[3] = { U[0], U[1], U[2] };
            if (f >= 0) e_pre = wall_side_terms(cfg, w, s->wall, f, t_id, nw, U, erot, vib, s->capacity);
            if (cfg->cyl_patch.model == HS_PATCH_DIFFUSE) {
                diffuse_reflect(cfg, &cfg->cyl_patch, nw, t_id, U, erot, vib, s->capacity, &r);
            } else {
                const double und = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
                if (und > 0.0)
                    for (int b = 0; b < 3; ++b) U[b] = U[b] - 2.0 * und * nw[b];
            }
            if (f >= 0) {
                const double e_post = wall_side_terms(cfg, w, s->wall, f, t_id, nw, U, erot, vib, s->capacity);
                wall_exchange_terms(cfg, w, s->wall, f, t_id, e_pre, u_pre, e_post, U);
            }
            if (w->flux && wf >= 0)
                track_face(w, t_id, w->face_off[3] + (wf - w->cyl_first),
                           of_sign(U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2]));
            continue;
        }
        p[axis] = face_coord(cfg, axis, dir > 0 ? idx[axis] + 1 : idx[axis]);
        const int ni = idx[axis] + dir;
        if (ni >= 0 && ni < nn[axis]) {
            /* internal face: Sf points +axis, U crosses it along dir */
            if (w->flux) track_face(w, t_id, mesh_face(cfg, w, axis, dir > 0 ? idx[axis] + 1 : idx[axis], idx), dir);
            idx[axis] = ni;
            continue;
        }
        const int sd = 2 * axis + (dir > 0 ? 1 : 0);
        const hs_patch* pt = &cfg->patch[sd];
        if (pt->model == HS_PATCH_CYCLIC) {
            /* counted, unsigned, on the coupled face of the other side */
            if (w->flux) track_face(w, t_id, mesh_face(cfg, w, axis, dir > 0 ? 0 : nn[axis], idx), 1);
            p[axis] = face_coord(cfg, axis, dir > 0 ? 0 : nn[axis]);
            idx[axis] = dir > 0 ? 0 : nn[axis] - 1;
            continue;
        }
        if (pt->model == HS_PATCH_DELETION) {
            if (w->flux)
                track_face(w, t_id, mesh_face(cfg, w, axis, dir > 0 ? nn[axis] : 0, idx), of_sign(U[axis] * (double)dir));
            keep = 0;
            continue;
        }
        /* diffuse or specular wall (an empty plane is never crossed) */
        OR_ATOMIC s->counters[4] += 1;
        double nw[3] = { 0.0, 0.0, 0.0 };
        nw[axis] = (double)dir;
        int64_t f = -1;
        if (w->measure && w->off[sd] >= 0) {
            const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
            f = w->off[sd] + idx[a1] + (int64_t)nn[a1] * idx[a2];
        }
        double e_pre = 0.0;
        const double u_pre[3] = { U[0], U[1], U[2] };
        if (f >= 0) e_pre = wall_side_terms(cfg, w, s->wall, f, t_id, nw, U, erot, vib, s->capacity);
        if (pt->model == HS_PATCH_DIFFUSE)
            diffuse_reflect(cfg, pt, nw, t_id, U, erot, vib, s->capacity, &r);
        else if (U[axis] * (double)dir > 0.0)
            U[axis] = -U[axis];
        if (f >= 0) {
            const double e_post = wall_side_terms(cfg, w, s->wall, f, t_id, nw, U, erot, vib, s->capacity);
            wall_exchange_terms(cfg, w, s->wall, f, t_id, e_pre, u_pre, e_post, U);
        }
        if (w->flux)
            track_face(w, t_id, mesh_face(cfg, w, axis, dir > 0 ? nn[axis] : 0, idx), of_sign(U[axis] * (double)dir));
    }
    P->px[i] = p[0]; P->py[i] = p[1]; P->pz[i] = p[2];
    P->ux[i] = (float)U[0]; P->uy[i] = (float)U[1]; P->uz[i] = (float)U[2];
    if (!keep) {
        P->cell[i] = -1;
        OR_ATOMIC s->counters[2] += 1;
    }
    else P->cell[i] = idx[0] + cfg->nx * (idx[1] + cfg->ny * idx[2]);
}

/* the wall-sum and face-tracker context of a run of move steps */
static void move_context(const hs_config* cfg, or_state* s, or_wall* w)
{
    const int sample = (cfg->flags & HS_FLAG_SAMPLE) != 0;
    const int measure = s->wall && (cfg->flags & HS_FLAG_WALL) && sample;
    const int track = s->face_flux && (cfg->flags & HS_FLAG_FACE_FLUX) && sample;
    if (measure || track) wall_layout(cfg, w);
    else memset(w, 0, sizeof(*w));
    w->measure = measure;
    w->flux = track ? s->face_flux : NULL;
    const int64_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    w->face_off[0] = 0;
    w->face_off[1] = (nx + 1) * ny * nz;
    w->face_off[2] = w->face_off[1] + nx * (ny + 1) * nz;
    w->face_off[3] = w->face_off[2] + nx * ny * (nz + 1);
    w->n_faces = or_face_count(cfg);
}

static void move_all(const hs_config* cfg, const or_derived* d, or_state* s, const or_wall* w, int64_t n_old)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096)
#endif
    for (int64_t i = 0; i < s->n; ++i) move_one(cfg, d, s, w, i, i >= n_old);
}

int or_move(const hs_config* cfg, or_state* s, int64_t n_old)
{
    or_derived d;
    derive(cfg, &d);
    or_wall w;
    move_context(cfg, s, &w);
    move_all(cfg, &d, s, &w, n_old);
    wall_free(&w);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* cell occupancy: stable counting sort (dsmcCloud.C:63-74)                 */
/* ------------------------------------------------------------------------ */
static void copy_particle(const hs_config* cfg, const or_derived* d, const hs_particles* src,
                          int64_t i, int64_t src_stride, hs_particles* dst, int64_t j,
                          int64_t dst_stride)
{
    dst->px[j] = src->px[i]; dst->py[j] = src->py[i]; dst->pz[j] = src->pz[i];
    dst->ux[j] = src->ux[i]; dst->uy[j] = src->uy[i]; dst->uz[j] = src->uz[i];
    dst->cell[j] = src->cell[i];
    if (src->erot && dst->erot) dst->erot[j] = src->erot[i];
    if (src->type && dst->type) dst->type[j] = src->type[i];
    if (src->vib && dst->vib)
        for (int m = 0; m < d->max_modes; ++m) dst->vib[m * dst_stride + j] = src->vib[m * src_stride + i];
    (void)cfg;
}

#ifdef _OPENMP
/* the same stable counting sort with the parcels cut into one contiguous
 * block per thread: a block's parcels of a cell go after those of the blocks
 * before it, so the order is the serial one */
static void sort_threads(const hs_config* cfg, const or_derived* d, or_state* s, int64_t nc)
{
    const int T = omp_get_max_threads();
    int32_t* cnt = (int32_t*)calloc((size_t)T * nc, sizeof(int32_t));
    const int64_t n = s->n;
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const int64_t lo = n * t / T, hi = n * (t + 1) / T;
        int32_t* mine = cnt + (int64_t)t * nc;
        for (int64_t i = lo; i < hi; ++i)
            if (s->p.cell[i] >= 0) mine[s->p.cell[i]] += 1;
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t c = 0; c < nc; ++c) {
            int32_t sum = 0;
            for (int u = 0; u < T; ++u) {
                const int32_t k = cnt[(int64_t)u * nc + c];
                cnt[(int64_t)u * nc + c] = sum;
                sum += k;
            }
            s->count[c] = sum;
        }
#pragma omp single
        {
            s->start[0] = 0;
            for (int64_t c = 0; c < nc; ++c) s->start[c + 1] = s->start[c] + s->count[c];
        }
        for (int64_t i = lo; i < hi; ++i) {
            const int32_t c = s->p.cell[i];
            if (c < 0) continue;
            copy_particle(cfg, d, &s->p, i, s->capacity, &s->tmp, s->start[c] + mine[c]++, s->capacity);
        }
#pragma omp barrier
        const int64_t m = s->start[nc];
#pragma omp for schedule(static)
        for (int64_t j = 0; j < m; ++j) copy_particle(cfg, d, &s->tmp, j, s->capacity, &s->p, j, s->capacity);
    }
    s->n = s->start[nc];
    free(cnt);
}
#endif

int or_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_sort(const hs_config* cfg, or_state* s)
{
    or_derived d;
    derive(cfg, &d);
    const int64_t nc = (int64_t)cfg->nx * cfg->ny * cfg->nz;
#ifdef _OPENMP
    if (omp_get_max_threads() > 1) {
        sort_threads(cfg, &d, s, nc);
        return 0;
    }
#endif
    memset(s->count, 0, sizeof(int32_t) * nc);
    for (int64_t i = 0; i < s->n; ++i)
        if (s->p.cell[i] >= 0) s->count[s->p.cell[i]] += 1;
    s->start[0] = 0;
    for (int64_t c = 0; c < nc; ++c) s->start[c + 1] = s->start[c] + s->count[c];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * nc);
    memcpy(cur, s->start, sizeof(int32_t) * nc);
    for (int64_t i = 0; i < s->n; ++i) {
        const int32_t c = s->p.cell[i];
        if (c < 0) continue;
        copy_particle(cfg, &d, &s->p, i, s->capacity, &s->tmp, cur[c]++, s->capacity);
    }
    free(cur);
    s->n = s->start[nc];
    for (int64_t j = 0; j < s->n; ++j) copy_particle(cfg, &d, &s->tmp, j, s->capacity, &s->p, j, s->capacity);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* NTC collisions (noTimeCounter.C:81-339) + sampling (dsmcVolFields.C)     */
/* ------------------------------------------------------------------------ */
/* dsmcVolFields::calculateField (dsmcVolFields.C:1115-1239): per species,
 * the parcels of the cell in order */
/* the sampling kernel's variant: with internal energies or several species
 * (the engine's "general" kernels) or without */
static int sums_general(const hs_config* cfg)
{
    if (cfg->n_species > 1) return 1;
    for (int i = 0; i < cfg->n_species; ++i)
        if (cfg->species[i].rot_dof > 0 || cfg->species[i].n_vib_modes > 0) return 1;
    return 0;
}

static void sample_cell(const hs_config* cfg, const or_derived* d, or_state* s, int64_t c,
                        int32_t b, int32_t nC, int64_t b2, int32_t n2, int64_t nc)
{
    const int ns = cfg->n_species;
    const hs_particles* P = &s->p;
    /* the sums are formed in the order of the sampling kernel: a cell of more
     * than 64 parcels by a wave, where lane l adds the terms of parcels l,
     * l+64, ... in order and the 64 partial sums are combined pairwise across
     * lane distances 32, 16, 8, 4, 2, 1; a smaller cell by four lanes, where
     * lane r adds parcels r, r+4, ... and the partial sums are combined as
     * (p0 + p2) + (p1 + p3), except in the kernel variant with internal
     * energies, where one lane adds its parcels in order.  With reactions a
     * cell that made products this step is summed over its parcels and then
     * its products (b2 .. b2+n2-1): by a wave as above when there are more
     * than 64 of them, else by one lane in order; the other cells as above. */
    const int lanes = (d->n_react && n2 > 0) ? (nC + n2 > 64 ? 64 : 1)
                                            : (nC > 64 ? 64 : (sums_general(cfg) ? 1 : 4));
    const int hf = (cfg->flags & HS_FLAG_HEATFLUX) != 0;
    enum { NQ = HS_SS_COUNT + HS_HF_COUNT };
    const int nq = hf ? NQ : HS_SS_COUNT;
    for (int sp_id = 0; sp_id < ns; ++sp_id) {
        const hs_species* sp = &cfg->species[sp_id];
        const double m = sp->mass;
        double part[NQ][64];
        for (int j = 0; j < nq; ++j)
            for (int l = 0; l < lanes; ++l) part[j][l] = 0.0;
        for (int32_t k = 0; k < nC + n2; ++k) {
            const int64_t i = k < nC ? b + k : b2 + (k - nC);
            const int t = P->type ? P->type[i] : 0;
            if (t != sp_id) continue;
            double q[NQ];
            const double ux = P->ux[i], uy = P->uy[i], uz = P->uz[i];
            const double lke = m * (ux * ux + uy * uy + uz * uz);
            const double er = P->erot ? (double)P->erot[i] : 0.0;
            double ev = 0.0;
            for (int j = HS_SS_EVIB0; j < HS_SS_COUNT; ++j) q[j] = 0.0;
            q[HS_SS_N] = 1.0;
            q[HS_SS_MUX] = m * ux;
            q[HS_SS_MUY] = m * uy;
            q[HS_SS_MUZ] = m * uz;
            q[HS_SS_MCC] = lke;
            q[HS_SS_EROT] = er;
            for (int mo = 0; mo < sp->n_vib_modes; ++mo) {
                const double e = (double)P->vib[mo * s->capacity + i] * HS_KB * sp->theta_v[mo];
                q[HS_SS_EVIB0 + mo] = e;
                ev += e;
            }
            double* h = q + HS_SS_COUNT;
            if (hf) {
                const double eint = er + ev;
                h[0] = m * (ux * ux); h[1] = m * ux * uy; h[2] = m * ux * uz;
                h[3] = m * (uy * uy); h[4] = m * uy * uz; h[5] = m * (uz * uz);
                h[6] = lke * ux; h[7] = lke * uy; h[8] = lke * uz;
                h[9] = eint * ux; h[10] = eint * uy; h[11] = eint * uz; h[12] = eint;
            }
            const int lane = (int)(k & (lanes - 1));
            for (int j = 0; j < nq; ++j) part[j][lane] += q[j];
        }
        for (int o = lanes / 2; o > 0; o >>= 1)
            for (int j = 0; j < nq; ++j) {
                double nxt[64];
                for (int l = 0; l < lanes; ++l) nxt[l] = part[j][l] + part[j][l ^ o];
                memcpy(part[j], nxt, sizeof(double) * (size_t)lanes);
            }
        const int64_t base = HS_S_GLOBAL_COUNT + (int64_t)sp_id * HS_SS_COUNT;
        for (int k = 0; k < HS_SS_COUNT; ++k) s->samples[(base + k) * nc + c] += part[k][0];
        if (hf) {
            const int64_t hb = HS_S_GLOBAL_COUNT + (int64_t)ns * HS_SS_COUNT + (int64_t)sp_id * HS_HF_COUNT;
            for (int k = 0; k < HS_HF_COUNT; ++k) s->samples[(hb + k) * nc + c] += part[HS_SS_COUNT + k][0];
        }
    }
}

/* octant sub-cells of one cell (noTimeCounter.C:114-138): parcel k is in
 * sub-cell pos(x - xc) + 2 pos(y - yc) + 4 pos(z - zc) with OpenFOAM's
 * pos(s) = (s >= 0); scn = members per sub-cell, sco = their offsets in
 * sub_list, which lists each sub-cell's members in parcel order */
static void subcell_tables(int32_t nC, const double* px, const double* py, const double* pz,
                           const double cc[3], int32_t* which, int32_t scn[8], int32_t sco[9],
                           int32_t* sub_list)
{
    for (int q = 0; q < 8; ++q) scn[q] = 0;
    for (int32_t k = 0; k < nC; ++k) {
        const int sc = (px[k] - cc[0] >= 0.0 ? 1 : 0) + 2 * (py[k] - cc[1] >= 0.0 ? 1 : 0)
                     + 4 * (pz[k] - cc[2] >= 0.0 ? 1 : 0);
        which[k] = sc;
        scn[sc] += 1;
    }
    sco[0] = 0;
    for (int q = 0; q < 8; ++q) sco[q + 1] = sco[q] + scn[q];
    int32_t fill[8];
    memcpy(fill, sco, sizeof(fill));
    for (int32_t k = 0; k < nC; ++k) sub_list[fill[which[k]]++] = k;
}

/* one candidate pair (noTimeCounter.C:157-198): P uniform over the cell; Q
 * from P's sub-cell when it has more than one member, else from the whole
 * cell, redrawn until Q != P */
static void draw_pair(or_rng* r, int32_t nC, const int32_t* which, const int32_t scn[8],
                      const int32_t sco[9], const int32_t* sub_list, int* p_out, int* q_out)
{
    const int pp = rng_label(r, 0, nC - 1);
    const int sc = which[pp];
    const int nsc = scn[sc];
    int qq = -1, guard = 0;
    if (nsc > 1) {
        do { qq = sub_list[sco[sc] + rng_label(r, 0, nsc - 1)]; }
        while (pp == qq && ++guard < OR_MAX_REDRAW);
    } else {
        do { qq = rng_label(r, 0, nC - 1); } while (pp == qq && ++guard < OR_MAX_REDRAW);
    }
    *p_out = pp;
    *q_out = qq;
}

int or_ntc_pairs(const hs_config* cfg, int64_t cell, int32_t nC, const double* px, const double* py,
                 const double* pz, uint32_t step, int64_t ncand, int32_t* p_out, int32_t* q_out)
{
    if (nC < 2 || nC > 1 << 20) return -1;
    const int64_t c = cell;
    const int ci = (int)(c % cfg->nx), cj = (int)((c / cfg->nx) % cfg->ny), ck = (int)(c / ((int64_t)cfg->nx * cfg->ny));
    const double cc[3] = { cfg->origin[0] + ((double)ci + 0.5) * cfg->delta[0],
                           cfg->origin[1] + ((double)cj + 0.5) * cfg->delta[1],
                           cfg->origin[2] + ((double)ck + 0.5) * cfg->delta[2] };
    int32_t* which = (int32_t*)malloc(sizeof(int32_t) * nC);
    int32_t* sub_list = (int32_t*)malloc(sizeof(int32_t) * nC);
    int32_t scn[8], sco[9];
    subcell_tables(nC, px, py, pz, cc, which, scn, sco, sub_list);
    for (int64_t k = 0; k < ncand; ++k) {
        or_rng r;
        rng_init(&r, cfg->seed, OR_RNG_COLLIDE, (uint32_t)c, step, (uint32_t)k);
        int pp, qq;
        draw_pair(&r, nC, which, scn, sco, sub_list, &pp, &qq);
        p_out[k] = pp;
        q_out[k] = qq;
    }
    free(which);
    free(sub_list);
    return 0;
}

/* writes a parcel's state back after its collision; after a reaction its
 * type and every vibrational plane are rewritten too */
static void store_state(const hs_config* cfg, const or_derived* d, or_state* s, int64_t i, const or_pstate* a,
                        int reacted)
{
    hs_particles* P = &s->p;
    P->ux[i] = (float)a->u[0]; P->uy[i] = (float)a->u[1]; P->uz[i] = (float)a->u[2];
    if (P->erot) P->erot[i] = (float)a->er;
    const int nm = reacted ? d->max_modes : cfg->species[a->t].n_vib_modes;
    for (int m = 0; m < nm; ++m) P->vib[m * s->capacity + i] = (uint16_t)a->v[m];
    if (reacted && P->type) P->type[i] = (uint8_t)a->t;
}

/* dsmcCloud::addNewParcel (dsmcCloud.C:396-432) for a dissociation product:
 * at the dissociating parcel's position and cell, with no internal energy.
 * Products go after the sorted population in the order they are made (cell
 * by cell, candidate by candidate); one beyond the capacity is dropped and
 * counted in counters[7] */
static void append_product(const hs_config* cfg, const or_derived* d, or_state* s, int64_t parent, int32_t c,
                           const or_pstate* prod, int64_t cnt[3])
{
    (void)cfg;
    (void)cnt;
    if (s->n >= s->capacity) {
        s->counters[7] += 1;
        return;
    }
    hs_particles* P = &s->p;
    const int64_t j = s->n++;
    P->px[j] = P->px[parent]; P->py[j] = P->py[parent]; P->pz[j] = P->pz[parent];
    P->ux[j] = (float)prod->u[0]; P->uy[j] = (float)prod->u[1]; P->uz[j] = (float)prod->u[2];
    P->cell[j] = c;
    if (P->erot) P->erot[j] = 0.0f;
    for (int m = 0; m < d->max_modes; ++m) P->vib[m * s->capacity + j] = 0;
    if (P->type) P->type[j] = (uint8_t)prod->t;
}

/* one cell of or_collide; cnt gets candidates, collisions and clamps */
static void collide_cell(const hs_config* cfg, const or_derived* d, or_state* s, int64_t c, int32_t sample,
                         int64_t nc, int active, int32_t* which, int32_t* sub_list, int64_t cnt[3])
{
    hs_particles* P = &s->p;
    const int32_t b = s->start[c];
    const int32_t nC = s->start[c + 1] - b;
    const int64_t n_before = s->n;  /* this cell's products start here */
    double ncoll = 0.0, sep = 0.0;
    if (active && nC > 1) {
        const int ci = (int)(c % cfg->nx), cj = (int)((c / cfg->nx) % cfg->ny), ck = (int)(c / ((int64_t)cfg->nx * cfg->ny));
        const double cc[3] = { cfg->origin[0] + ((double)ci + 0.5) * cfg->delta[0],
                               cfg->origin[1] + ((double)cj + 0.5) * cfg->delta[1],
                               cfg->origin[2] + ((double)ck + 0.5) * cfg->delta[2] };
        int32_t scn[8], sco[9];
        subcell_tables(nC, &P->px[b], &P->py[b], &P->pz[b], cc, which, scn, sco, sub_list);
        const double smax = s->sigma_tcr_max[c];
        const double t_macro = s->t_overall ? s->t_overall[c] : 0.0;
        const double sel = s->remainder[c]
            + 0.5 * nC * (nC - 1) * cfg->fn * smax * cfg->dt / cfg->cell_volume[c];
        int64_t ncand = (int64_t)sel;
        s->remainder[c] = sel - (double)ncand;
        if (ncand > OR_MAX_CAND) { ncand = OR_MAX_CAND; cnt[2] += 1; }
        cnt[0] += ncand;
        for (int64_t k = 0; k < ncand; ++k) {
            /* every candidate draws from its own stream (cell, step, k) */
            or_rng r;
            rng_init(&r, cfg->seed, OR_RNG_COLLIDE, (uint32_t)c, s->step, (uint32_t)k);
            int pp, qq;
            draw_pair(&r, nC, which, scn, sco, sub_list, &pp, &qq);
            if (pp == qq) continue;
            const int64_t ip = b + pp, iq = b + qq;
            const int tp = P->type ? P->type[ip] : 0, tq = P->type ? P->type[iq] : 0;
            if (cfg->species[tp].charge == -1 && cfg->species[tq].charge == -1) continue;
            double up[3] = { P->ux[ip], P->uy[ip], P->uz[ip] };
            double uq[3] = { P->ux[iq], P->uy[iq], P->uz[iq] };
            const double stcr = sigma_tcr(d, tp, tq, up, uq);
            if (stcr > s->sigma_tcr_max[c]) s->sigma_tcr_max[c] = stcr;
            if (stcr / smax > rng_u(&r)) {
                or_pstate a = { tp, { up[0], up[1], up[2] }, P->erot ? P->erot[ip] : 0.0, { 0 } };
                or_pstate b = { tq, { uq[0], uq[1], uq[2] }, P->erot ? P->erot[iq] : 0.0, { 0 } };
                for (int m = 0; m < cfg->species[tp].n_vib_modes; ++m) a.v[m] = P->vib[m * s->capacity + ip];
                for (int m = 0; m < cfg->species[tq].n_vib_modes; ++m) b.v[m] = P->vib[m * s->capacity + iq];
                or_pstate prod;
                int from_q = 0;
                const int reacted = collide_or_react(cfg, d, &a, &b, &r, &prod, &from_q, t_macro);
                store_state(cfg, d, s, ip, &a, reacted);
                store_state(cfg, d, s, iq, &b, reacted);
                if (prod.t >= 0) append_product(cfg, d, s, from_q ? iq : ip, (int32_t)c, &prod, cnt);
                ncoll += 1.0;
                if (cfg->flags & HS_FLAG_COLL_SEP) {
                    const double ddx = P->px[ip] - P->px[iq], ddy = P->py[ip] - P->py[iq], ddz = P->pz[ip] - P->pz[iq];
                    sep += sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                }
                cnt[1] += 1;
            }
        }
    }
    if (sample && s->samples) {
        sample_cell(cfg, d, s, c, b, nC, n_before, (int32_t)(s->n - n_before), nc);
        s->samples[(int64_t)HS_S_NCOLL * nc + c] += ncoll;
        s->samples[(int64_t)HS_S_COLLSEP * nc + c] += sep;
    }
}

int or_collide(const hs_config* cfg, or_state* s, int32_t sample)
{
    or_derived d;
    derive(cfg, &d);
    const int64_t nc = (int64_t)cfg->nx * cfg->ny * cfg->nz;
    const int active = cfg->collision_model != HS_COLL_NONE;
    int32_t max_nc = 1;
    for (int64_t c = 0; c < nc; ++c)
        if (s->start[c + 1] - s->start[c] > max_nc) max_nc = s->start[c + 1] - s->start[c];
    /* with reactions the cells run in order: their products are appended
     * cell by cell */
#ifdef _OPENMP
#pragma omp parallel if (d.n_react == 0)
#endif
    {
        int32_t* sub_list = (int32_t*)malloc(sizeof(int32_t) * max_nc);
        int32_t* which = (int32_t*)malloc(sizeof(int32_t) * max_nc);
        int64_t cnt[3] = { 0, 0, 0 };
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t c = 0; c < nc; ++c) collide_cell(cfg, &d, s, c, sample, nc, active, which, sub_list, cnt);
        OR_ATOMIC s->counters[0] += cnt[0];
        OR_ATOMIC s->counters[1] += cnt[1];
        OR_ATOMIC s->counters[6] += cnt[2];
        free(sub_list);
        free(which);
    }
    if (sample && s->samples) s->n_sampled += 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* freestream inflow (dsmcFreeStreamInflowPatch.C:91-364)                   */
/* ------------------------------------------------------------------------ */
static void new_particle_internal(const hs_config* cfg, or_state* s, int64_t j, int t, double T_rot,
                                  double T_vib, or_rng* r)
{
    const hs_species* sp = &cfg->species[t];
    if (s->p.erot) s->p.erot[j] = (float)equip_erot(r, T_rot, sp->rot_dof);
    if (s->p.vib && sp->n_vib_modes > 0) equip_vib(r, T_vib, sp, &s->p.vib[j], s->capacity);
    if (s->p.type) s->p.type[j] = (uint8_t)t;
}

int or_inject(const hs_config* cfg, or_state* s)
{
    or_derived d;
    derive(cfg, &d);
    const int64_t nf = list_inflow_faces(cfg, NULL);
    if (nf == 0) return 0;
    or_face* faces = (or_face*)malloc(sizeof(or_face) * nf);
    list_inflow_faces(cfg, faces);
    const int ns = cfg->n_species;
    const double sqrt_pi = sqrt(OR_PI);
    /* accumulation (Bird eq. 4.22) */
    for (int sp = 0; sp < ns; ++sp) {
        if (!(cfg->inflow_n[sp] > 0.0)) continue;
        const double mass = cfg->species[sp].mass;
        const double mps = sqrt(2.0 * HS_KB * cfg->inflow_t / mass);
        for (int64_t f = 0; f < nf; ++f) {
            const int a = faces[f].axis;
            const int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
            const double fa = cfg->delta[a1] * cfg->delta[a2];
            /* U . (-Sf/|Sf|): inward normal component */
            const double s_cos = (cfg->inflow_u[a] * (double)(-faces[f].dir)) / mps;
            s->inflow_acc[sp * nf + f] +=
                (fa * cfg->inflow_n[sp] * cfg->dt * mps
                 * (exp(-(s_cos * s_cos)) + sqrt_pi * s_cos * (1.0 + erf(s_cos))))
                / (2.0 * sqrt_pi * cfg->fn);
        }
    }
    /* insertion */
    int64_t n = s->n;
    for (int64_t f = 0; f < nf; ++f) {
        const int a = faces[f].axis, dir = faces[f].dir;
        const int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
        const int nn[3] = { cfg->nx, cfg->ny, cfg->nz };
        for (int sp = 0; sp < ns; ++sp) {
            if (!(cfg->inflow_n[sp] > 0.0)) continue;
            or_rng r;
            rng_init(&r, cfg->seed, OR_RNG_INJECT, (uint32_t)(f * ns + sp), s->step, 0);
            double* acc = &s->inflow_acc[sp * nf + f];
            int ni = (int)(*acc);
            if (ni < 0) ni = 0;
            if ((*acc - (double)ni) > rng_u(&r)) ni += 1;
            *acc -= (double)ni;
            const double mass = cfg->species[sp].mass;
            const double mps = sqrt(2.0 * HS_KB * cfg->inflow_t / mass);
            const double un_face = cfg->inflow_u[a] * (double)(-dir);
            const double s_cos = un_face / mps;
            const double coef_a = s_cos + sqrt(s_cos * s_cos + 2.0);
            const double coef_b = 0.5 * (1.0 + s_cos * (s_cos - sqrt(s_cos * s_cos + 2.0)));
            double scaling = 3.0;
            if (s_cos < -3.0) scaling = fabs(s_cos) + 1.0;
            const double sc = sqrt(HS_KB * cfg->inflow_t / mass);
            for (int k = 0; k < ni; ++k) {
                if (n >= s->capacity) { s->counters[6] += 1; break; }
                /* parcel k of this face and species draws from its own stream */
                or_rng rp;
                rng_init(&rp, cfg->seed, OR_RNG_INJECT, (uint32_t)(f * ns + sp), s->step, 1u + (uint32_t)k);
                double pos[3];
                const double ua = rng_u(&rp), ub = rng_u(&rp);
                pos[a] = face_coord(cfg, a, dir > 0 ? nn[a] : 0);
                pos[a1] = d.empty[a1] ? d.centre[a1]
                                      : face_coord(cfg, a1, faces[f].a1i) + ua * cfg->delta[a1];
                pos[a2] = d.empty[a2] ? d.centre[a2]
                                      : face_coord(cfg, a2, faces[f].a2i) + ub * cfg->delta[a2];
                double un;
                if (fabs(un_face) > OR_VSMALL) {
                    double P;
                    int guard = 0;
                    do {
                        const double unt = scaling * (2.0 * rng_u(&rp) - 1.0);
                        un = unt + s_cos;
                        if (un < 0.0) P = -1.0;
                        else P = 2.0 * un / coef_a * or_exp(coef_b - unt * unt);
                    } while (P < rng_u(&rp) && ++guard < OR_MAX_REDRAW);
                } else {
                    un = sqrt(-or_log(1.0 - rng_u(&rp)));
                }
                const double g1 = rng_gauss(&rp), g2 = rng_gauss(&rp);
                double U[3];
                U[a] = mps * un * (double)(-dir);
                U[a1] = sc * g1 + cfg->inflow_u[a1];
                U[a2] = sc * g2 + cfg->inflow_u[a2];
                s->p.px[n] = pos[0]; s->p.py[n] = pos[1]; s->p.pz[n] = pos[2];
                s->p.ux[n] = (float)U[0]; s->p.uy[n] = (float)U[1]; s->p.uz[n] = (float)U[2];
                s->p.cell[n] = faces[f].cell;
                if (s->p.vib)
                    for (int m = 0; m < d.max_modes; ++m) s->p.vib[m * s->capacity + n] = 0;
                new_particle_internal(cfg, s, n, sp, cfg->inflow_trot, cfg->inflow_tvib, &rp);
                s->counters[3] += 1;
                ++n;
            }
        }
    }
    s->n = n;
    free(faces);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* initial fill (dsmcMeshFill.C:68-~230)                                    */
/* ------------------------------------------------------------------------ */
/* the fill of dsmcMeshFill.C:68-~230 with, where the per-cell arrays are
 * given, each cell's own number densities nd_c[species][cell], velocity
 * u_c[axis][cell] and temperature t_c[cell] (internal temperatures equal to
 * it); sigmaTcRMax then follows each cell's temperature */
static int fill_general(const hs_config* cfg, or_state* s, const double* nd, const double* nd_c, double T,
                        const double* t_c, double Trot, double Tvib, const double u[3], const double* u_c,
                        int32_t exact, uint32_t fill_id)
{
    or_derived d;
    derive(cfg, &d);
    const int64_t nc = (int64_t)cfg->nx * cfg->ny * cfg->nz;
    const int ns = cfg->n_species;
    int64_t n = s->n;
    for (int64_t c = 0; c < nc; ++c) {
        const double V = cfg->cell_volume[c];
        if (!(V > 0.0)) continue;
        const int idx[3] = { (int)(c % cfg->nx), (int)((c / cfg->nx) % cfg->ny), (int)(c / ((int64_t)cfg->nx * cfg->ny)) };
        const double Tc = t_c ? t_c[c] : T;
        const double uc[3] = { u_c ? u_c[c] : u[0], u_c ? u_c[nc + c] : u[1], u_c ? u_c[2 * nc + c] : u[2] };
        for (int sp = 0; sp < ns; ++sp) {
            const double nds = nd_c ? nd_c[(int64_t)sp * nc + c] : nd[sp];
            if (!(nds > 0.0)) continue;
            or_rng rc, rg;
            rng_init(&rc, cfg->seed, OR_RNG_FILL_COUNT, (uint32_t)c, fill_id, (uint32_t)sp);
            rng_init(&rg, cfg->seed, OR_RNG_FILL_GEN, (uint32_t)c, fill_id, (uint32_t)sp);
            const double req = nds * V / cfg->fn;
            int ni;
            if (exact) ni = (int)floor(req + 0.5);
            else { ni = (int)req; if ((req - (double)ni) > rng_u(&rc)) ni += 1; }
            const hs_species* spp = &cfg->species[sp];
            const double sc = sqrt(HS_KB * Tc / spp->mass);
            for (int k = 0; k < ni; ++k) {
                if (n >= s->capacity) { s->counters[6] += 1; break; }
                double pos[3] = { 0.0, 0.0, 0.0 };
                int ok = 0;
                for (int tries = 0; tries < OR_MAX_FILL_TRIES && !ok; ++tries) {
                    for (int a = 0; a < 3; ++a) {
                        const double ua = rng_u(&rg);
                        pos[a] = d.empty[a] ? d.centre[a] : face_coord(cfg, a, idx[a]) + ua * cfg->delta[a];
                    }
                    ok = 1;
                    if (cfg->cyl_enabled) {
                        const double rx = pos[0] - cfg->cyl_center[0], ry = pos[1] - cfg->cyl_center[1];
                        if (rx * rx + ry * ry < cfg->cyl_radius * cfg->cyl_radius) ok = 0;
                    }
                }
                /* a failed rejection keeps its slot as a deleted parcel
                 * (cell -1) so the population layout does not depend on it */
                const double g0 = rng_gauss(&rg), g1 = rng_gauss(&rg), g2 = rng_gauss(&rg);
                s->p.px[n] = pos[0]; s->p.py[n] = pos[1]; s->p.pz[n] = pos[2];
                s->p.ux[n] = (float)(sc * g0 + uc[0]);
                s->p.uy[n] = (float)(sc * g1 + uc[1]);
                s->p.uz[n] = (float)(sc * g2 + uc[2]);
                s->p.cell[n] = ok ? (int32_t)c : -1;
                if (s->p.vib)
                    for (int m = 0; m < d.max_modes; ++m) s->p.vib[m * s->capacity + n] = 0;
                new_particle_internal(cfg, s, n, sp, t_c ? Tc : Trot, t_c ? Tc : Tvib, &rg);
                ++n;
            }
        }
    }
    s->n = n;
    /* sigmaTcRMax = sigmaT * most probable speed of the most abundant species */
    for (int64_t c = 0; c < nc; ++c) {
        int most = 0;
        for (int sp = 1; sp < ns; ++sp) {
            const double a = nd_c ? nd_c[(int64_t)sp * nc + c] : nd[sp];
            const double b = nd_c ? nd_c[(int64_t)most * nc + c] : nd[most];
            if (a > b) most = sp;
        }
        const hs_species* sm = &cfg->species[most];
        const double Tc = t_c ? t_c[c] : T;
        s->sigma_tcr_max[c] = (OR_PI * sm->diameter * sm->diameter) * sqrt(2.0 * HS_KB * Tc / sm->mass);
        or_rng r;
        rng_init(&r, cfg->seed, OR_RNG_REMAINDER, (uint32_t)c, fill_id, 0);
        s->remainder[c] = rng_u(&r);
    }
    return 0;
}

int or_fill(const hs_config* cfg, or_state* s, const double* nd, double T, double Trot,
            double Tvib, const double u[3], int32_t exact, uint32_t fill_id)
{
    return fill_general(cfg, s, nd, NULL, T, NULL, Trot, Tvib, u, NULL, exact, fill_id);
}

int or_fill_cells(const hs_config* cfg, or_state* s, const double* nd_cell, const double* u_cell,
                  const double* t_cell, uint32_t fill_id)
{
    static const double zero3[3] = { 0.0, 0.0, 0.0 };
    return fill_general(cfg, s, NULL, nd_cell, 0.0, t_cell, 0.0, 0.0, zero3, u_cell, 0, fill_id);
}

/* ------------------------------------------------------------------------ */
/* dsmcCloud::evolve                                                        */
/* ------------------------------------------------------------------------ */
int or_evolve(const hs_config* cfg, or_state* s, int32_t nsteps)
{
    or_derived d;
    derive(cfg, &d);
    or_wall w;
    move_context(cfg, s, &w);
    for (int32_t k = 0; k < nsteps; ++k) {
        const int64_t n_old = s->n;
        if (s->inflow_acc) or_inject(cfg, s);
        move_all(cfg, &d, s, &w, n_old);
        or_sort(cfg, s);
        or_collide(cfg, s, (cfg->flags & HS_FLAG_SAMPLE) ? 1 : 0);
        s->step += 1;
    }
    wall_free(&w);
    return 0;
}
