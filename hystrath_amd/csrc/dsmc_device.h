// dsmc_device.h -- device-side building blocks of the MI355X DSMC engine:
// the parameter block, the counter-based random streams, FP64 math built from
// IEEE +,-,*,/ and sqrt only (so every compiler rounds it identically), and
// the collision physics of dsmcFoam+:
//   VariableHardSphere.C:80-228, VariableSoftSphere.C:78-247,
//   LarsenBorgnakkeVariableHardSphere.C:110-279, dsmcCloud.C:1015-1040,
//   dsmcCloud.C:1181-1244,1327-1531.
// Compile with -ffp-contract=off: each a*b+c below is two roundings.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hystrath_dsmc.h"

#define HS_VSMALL 1.0e-300
#define HS_SMALL 1.0e-15
#define HS_ROOTVSMALL 1.0e-150
#define HS_PI 3.14159265358979323846
#define HS_TWO_PI 6.28318530717958647692
#define HS_MAX_TRACK 1000
#define HS_MAX_REDRAW 100000
#define HS_MAX_CAND (1 << 22)
#define HS_MAX_FILL_TRIES 100000

// random-stream purposes (top byte of counter word 2)
enum { RNG_MOVE = 1, RNG_COLLIDE = 2, RNG_INJECT = 3, RNG_FILL_COUNT = 4,
       RNG_FILL_GEN = 5, RNG_REMAINDER = 6 };

// counters[] slots
enum { CNT_CAND = 0, CNT_COLL = 1, CNT_DEL = 2, CNT_INJ = 3, CNT_WALL = 4,
       CNT_TRACK_OVF = 5, CNT_TILE_OVF = 6, CNT_CAP_OVF = 7, CNT_PSTEPS = 8, CNT_N = 9 };

// One reaction of chemReactDict with its constants (the QK models of
// src/lagrangian/dsmc/reactions/derived).
struct DevReaction {
    int model;             // hs_reaction_model
    int r[2];              // reactant typeIds
    int dprod[2][2];       // dissociation products of each reactant, -1: none
    int xmol, xatom;       // exchange products: the molecule, the atom
    int pos_mol;           // exchangeQK::posMolReactant_ (exchangeQK.C:52-60)
    int split;             // allowSplitting
    double heat_diss[2];   // k thetaD of each reactant (dissociationQK.C:57-60)
    double heat_x;         // heatOfReactionExchange k (exchangeQK.C:387-391)
    double a_dash, b;      // Bird's QK exchange constants (exchangeQK.C:195-200)
    double om_x;           // omega of the exchange's redistribution (exchangeQK.C:303-308)
};

// Everything a kernel needs about the case, resident in device memory and
// read through uniform (scalar) loads.
struct DevParams {
    int nx, ny, nz, nc;
    double rnx, rnxy;  // 1/nx and 1/(nx*ny), for cell_ijk
    double origin[3], delta[3];
    int empty[3];
    double centre[3];
    int patch_model[6];
    double patch_T[6];
    double patch_U[6][3];
    int cyl;
    double cyl_c[2], cyl_r;
    int cyl_model;
    double cyl_T;
    double cyl_U[3];
    int ns, max_modes;
    double mass[HS_MAX_SPECIES], rotdof[HS_MAX_SPECIES];
    int nvib[HS_MAX_SPECIES], charge[HS_MAX_SPECIES];
    double theta_v[HS_MAX_SPECIES][HS_MAX_VIB_MODES];
    double sig_coef[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double ref_term[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double omega_pq[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double alpha_pq[HS_MAX_SPECIES][HS_MAX_SPECIES];
    double theta_d[HS_MAX_SPECIES];
    double zv_b[HS_MAX_SPECIES][HS_MAX_SPECIES][HS_MAX_VIB_MODES];
    double zv_pow2[HS_MAX_SPECIES][HS_MAX_VIB_MODES];
    int coll_model;
    double z_rot, z_el, z_vib_fixed;
    int inv_zv_form;
    const double* t_overall;    // [nc] cells' overall temperature (inverseZvFormulation 2008 only)
    const double* zv_tab;       // 1/Zv by (species, partner, mode, imax < HS_ZV_TAB) from the collision
                                // temperature, or null (see inv_zv_collision)
    double fn, dt;
    uint32_t seed_lo, seed_hi;
    int flags, nslots;
    int has_erot, has_vib, has_type;
    int64_t capacity;
    // freestream inflow
    double inflow_t, inflow_trot, inflow_tvib;
    double inflow_u[3];
    double inflow_n[HS_MAX_SPECIES];
    int n_inflow_faces;
    // wall-face sampling (boundaryMeasurements), HS_FLAG_WALL
    int n_wall;                 // wall faces
    int wall_off[6];            // first face of each measured side, -1 otherwise
    const double* wall_area;    // [n_wall]
    const int32_t* cyl_face;    // [nc] wall face of the cylinder in each cell, -1 if none
    double* wall_sum;           // [ns*HS_WS_COUNT][n_wall]
    int cyl_first;              // first cylinder face in the wall-face list
    // face tracker (dsmcFaceTracker), HS_FLAG_FACE_FLUX
    unsigned long long* face_flux;  // [ns][n_faces] two's-complement int64 counts
    int64_t n_faces;
    int64_t face_off[4];        // first face of the x, y, z planes; cylinder faces
    // chemistry (dsmcReactions)
    int n_react;
    int kind[HS_MAX_SPECIES];   // dsmcParcelI.H:153-196: 0 electron, 10/11 atom, 20/21, 30/31 molecule
    int react_of[HS_MAX_SPECIES][HS_MAX_SPECIES];  // pairAddressing_: reaction of a pair, -1 none
    DevReaction rx[HS_MAX_REACTIONS];
};

// face-tracker face of plane k on axis a at the tangential indices of idx
__device__ __forceinline__ int64_t mesh_face(const DevParams& P, int a, int k, const int idx[3]) {
    const int j1 = a == 0 ? idx[1] : idx[0], j2 = a == 2 ? idx[1] : idx[2];
    const int n1 = a == 0 ? P.ny : P.nx;
    const int na = a == 0 ? P.nx : (a == 1 ? P.ny : P.nz);
    return P.face_off[a] + k + (int64_t)(na + 1) * (j1 + (int64_t)n1 * j2);
}

__device__ __forceinline__ void track_face(const DevParams& P, int t, int64_t face, long long v) {
    atomicAdd(&P.face_flux[(int64_t)t * P.n_faces + face], (unsigned long long)v);
}

// Device view of one particle buffer (structure of arrays).
struct DevParticles {
    double* px; double* py; double* pz;
    float* ux; float* uy; float* uz;
    int32_t* cell;
    float* erot;     // nullptr when no species rotates
    uint16_t* vib;   // [mode][capacity] or nullptr
    uint8_t* type;   // nullptr for one species
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double uniform53(uint32_t hi, uint32_t lo) {
    const uint64_t x = (((uint64_t)hi << 32) | (uint64_t)lo) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// FP64 math from IEEE +,-,*,/ only (same operation order as any host build)
// ---------------------------------------------------------------------------
#define HS_LN2_HI 6.93147180369123816490e-01
#define HS_LN2_LO 1.90821492927058770002e-10
#define HS_INV_LN2 1.44269504088896338700e+00

__device__ __forceinline__ double dm_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return __longlong_as_double(0x7FF8000000000000ll);
    if (x == 0.0) return -__longlong_as_double(0x7FF0000000000000ll);
    if (x == __longlong_as_double(0x7FF0000000000000ll)) return x;
    int k = 0;
    uint64_t b = (uint64_t)__double_as_longlong(x);
    if (((b >> 52) & 0x7ff) == 0) { x = x * 18014398509481984.0; b = (uint64_t)__double_as_longlong(x); k = -54; }
    k += (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
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
    return dk * HS_LN2_HI + (dk * HS_LN2_LO + lm);
}

__device__ __forceinline__ double dm_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return __longlong_as_double(0x7FF0000000000000ll);
    if (x < -745.1332191019412) return 0.0;
    const double kd = floor(x * HS_INV_LN2 + 0.5);
    const double r = (x - kd * HS_LN2_HI) - kd * HS_LN2_LO;
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
    return p * __longlong_as_double((long long)((uint64_t)(k + 1023) << 52));
}

__device__ __forceinline__ double dm_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    return dm_exp(y * dm_log(x));
}

__device__ __forceinline__ void dm_sincos2pi(double u, double* s_out, double* c_out) {
    const double q = floor(u * 4.0 + 0.5);
    const double d = u - q * 0.25;
    const double r = d * HS_TWO_PI;
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

// ---------------------------------------------------------------------------
// Random streams.  PhiloxStream: key = seed, counter = (id, step,
// purpose<<24 | sub, block); each block yields two doubles.  InjStream
// serves caller-provided uniforms (test hook).
// ---------------------------------------------------------------------------
struct PhiloxStream {
    uint32_t k0, k1, c0, c1, c2, blk;
    double spare_u, spare_g;
    bool have_u, have_g;
    __device__ __forceinline__ void init(const DevParams& P, int purpose, uint32_t id, uint32_t step, uint32_t sub) {
        k0 = P.seed_lo; k1 = P.seed_hi;
        c0 = id; c1 = step; c2 = ((uint32_t)purpose << 24) | (sub & 0xFFFFFFu);
        blk = 0; have_u = false; have_g = false; spare_u = 0.0; spare_g = 0.0;
    }
    __device__ __forceinline__ double u() {
        if (have_u) { have_u = false; return spare_u; }
        uint32_t o[4];
        philox4x32_10(c0, c1, c2, blk, k0, k1, o);
        blk += 1;
        spare_u = uniform53(o[2], o[3]);
        have_u = true;
        return uniform53(o[0], o[1]);
    }
    __device__ __forceinline__ double gauss() {
        if (have_g) { have_g = false; return spare_g; }
        const double u1 = u();
        const double u2 = u();
        const double rad = sqrt(-2.0 * dm_log(1.0 - u1));
        double sn, cs;
        dm_sincos2pi(u2, &sn, &cs);
        spare_g = rad * sn;
        have_g = true;
        return rad * cs;
    }
};

struct InjStream {
    const double* src;
    int n, used;
    __device__ __forceinline__ double u() {
        const int k = used++;
        return k < n ? src[k] : 0.5;
    }
};

// dsmcCloud::randomLabel (dsmcCloud.C:1015-1040)
template <class R>
__device__ __forceinline__ int random_label(R& r, int a, int b) {
    if (a == b) return a;
    const int start = a < b ? a : b;
    const int end = a < b ? b : a;
    for (int guard = 0; guard < HS_MAX_REDRAW; ++guard) {
        const int v = start + (int)(r.u() * (double)(end - start + 1));
        if (v != end + 1) return v;
    }
    return end;
}

// ---------------------------------------------------------------------------
// equipartition samplers (dsmcCloud.C:1181-1244)
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ double equip_erot(R& r, double T, double dof) {
    if (dof < HS_SMALL) return 0.0;
    if (dof < 2.0 + HS_SMALL && dof > 2.0 - HS_SMALL) return -dm_log(1.0 - r.u()) * HS_KB * T;
    const double a = 0.5 * dof - 1.0;
    double ratio, P;
    int guard = 0;
    do {
        ratio = 10.0 * r.u();
        P = dm_pow(ratio / a, a) * dm_exp(a - ratio);
    } while (P < r.u() && ++guard < HS_MAX_REDRAW);
    return ratio * HS_KB * T;
}

template <class R>
__device__ __forceinline__ int equip_vib_level(R& r, double T, double theta_v) {
    const int l = (int)(-dm_log(1.0 - r.u()) * T / theta_v);
    return l > 65535 ? 65535 : l;
}

// ---------------------------------------------------------------------------
// sigmaTcR (VariableHardSphere.C:80-121); pair constants precomputed on host
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sigma_tcr(const DevParams& P, int tp, int tq, const double up[3],
                                            const double uq[3]) {
    const double dx = up[0] - uq[0], dy = up[1] - uq[1], dz = up[2] - uq[2];
    const double cr2 = dx * dx + dy * dy + dz * dz;
    const double cr = sqrt(cr2);
    if (cr < HS_VSMALL) return 0.0;
    const double om = P.omega_pq[tp][tq];
    return P.sig_coef[tp][tq] * dm_pow(P.ref_term[tp][tq] / cr2, om - 0.5) * cr;
}

// postCollisionVelocities: VHS (VariableHardSphere.C:191-228) or VSS
// (VariableSoftSphere.C:189-247); cr < 0 means |UP-UQ|; rescale scales the
// pre-collision relative velocity to |cr| (LB-VSS energy conservation).
template <class R>
__device__ __forceinline__ void post_coll_vel(const DevParams& P, bool vss, bool rescale, int tp, int tq,
                                              double up[3], double uq[3], double cr, R& r) {
    const double mp = P.mass[tp], mq = P.mass[tq];
    if (cr < 0.0) {
        const double dx = up[0] - uq[0], dy = up[1] - uq[1], dz = up[2] - uq[2];
        cr = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const double msum = mp + mq;
    double ucm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) ucm[c] = (mp * up[c] + mq * uq[c]) / msum;
    double rel[3];
    if (!vss) {
        const double cos_t = 2.0 * r.u() - 1.0;
        const double sin_t = sqrt(1.0 - cos_t * cos_t);
        double sp, cp;
        dm_sincos2pi(r.u(), &sp, &cp);
        rel[0] = cr * cos_t;
        rel[1] = cr * (sin_t * cp);
        rel[2] = cr * (sin_t * sp);
    } else {
        double c0 = up[0] - uq[0], c1 = up[1] - uq[1], c2 = up[2] - uq[2];
        if (rescale) {
            const double cm = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            if (cm > HS_VSMALL) { const double f = cr / cm; c0 *= f; c1 *= f; c2 *= f; }
        }
        const double cos_t = 2.0 * dm_pow(r.u(), 1.0 / P.alpha_pq[tp][tq]) - 1.0;
        const double sin_t = sqrt(1.0 - cos_t * cos_t);
        double sp, cp;
        dm_sincos2pi(r.u(), &sp, &cp);
        const double D = sqrt(c1 * c1 + c2 * c2);
        if (D > HS_VSMALL) {
            rel[0] = cos_t * c0 + sin_t * sp * D;
            rel[1] = cos_t * c1 + sin_t * (cr * c2 * cp - c0 * c1 * sp) / D;
            rel[2] = cos_t * c2 - sin_t * (cr * c1 * cp + c0 * c2 * sp) / D;
        } else {
            rel[0] = cos_t * c0;
            rel[1] = cr * (sin_t * cp);
            rel[2] = cr * (sin_t * sp);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        up[c] = ucm[c] + rel[c] * mq / msum;
        uq[c] = ucm[c] - rel[c] * mp / msum;
    }
}

// 1/Zv (dsmcCloud.C:1427-1508) from the quantised collision temperature of
// level imax, for inverseZvFormulation pre-2008 and the default (1/(5 Zv));
// the engine tabulates it for imax < HS_ZV_TAB with this same code, so a
// table entry and the value computed here are the same double.
#define HS_ZV_TAB 256
__device__ __forceinline__ double inv_zv_collision(const DevParams& P, int t, int partner, int mode, int imax,
                                                   double om) {
    const double T = (double)imax * P.theta_v[t][mode] / (3.5 - om);
    const double pow1 = dm_pow(P.theta_d[t] / T, 1.0 / 3.0) - 1.0;
    const double zvp1 = dm_pow(P.theta_d[t] / T, om);
    const double zvp2 = dm_pow(P.zv_b[t][partner][mode], pow1 / P.zv_pow2[t][mode]);
    const double zv = zvp1 * zvp2;
    return P.inv_zv_form == 2 ? 1.0 / (5.0 * zv) : 1.0 / zv;
}

// dsmcCloud::postCollisionVibrationalEnergyLevel (dsmcCloud.C:1390-1531);
// POST = postReaction: the level is always redrawn, with no Zv test.
// cell: the collision's cell, whose overall temperature inverseZvFormulation
// 2008 (inv_zv_form 1) uses in Zv when it exceeds SMALL (dsmcCloud.C:
// 1440-1455); otherwise (or with cell -1) the quantised collision
// temperature is used.
template <bool POST = false, class R>
__device__ int vib_level(const DevParams& P, int t, int partner, int mode, int level, int imax,
                         double ec, double om, R& r, int cell = -1) {
    const double theta_v = P.theta_v[t][mode];
    if (POST) {
        int idash, guard = 0;
        double func;
        do {
            idash = random_label(r, 0, imax);
            const double evib = (double)idash * HS_KB * theta_v;
            func = dm_pow(1.0 - evib / ec, 1.5 - om);
        } while (func < r.u() && ++guard < HS_MAX_REDRAW);
        return idash;
    }
    double inv_zv = 1.0;
    if (P.z_vib_fixed == 0.0) {
        double tm = 0.0;
        if (P.inv_zv_form == 1 && cell >= 0) tm = P.t_overall[cell];
        if (tm > HS_SMALL) {
            // inverseZvFormulation 2008 with the cell's overall temperature
            const double pow1 = dm_pow(P.theta_d[t] / tm, 1.0 / 3.0) - 1.0;
            const double zvp1 = dm_pow(P.theta_d[t] / tm, om);
            const double zvp2 = dm_pow(P.zv_b[t][partner][mode], pow1 / P.zv_pow2[t][mode]);
            inv_zv = 1.0 / (zvp1 * zvp2);
        } else if (P.zv_tab && imax < HS_ZV_TAB) {
            inv_zv = P.zv_tab[((t * P.ns + partner) * HS_MAX_VIB_MODES + mode) * HS_ZV_TAB + imax];
        } else {
            inv_zv = inv_zv_collision(P, t, partner, mode, imax, om);
        }
    } else {
        inv_zv = 1.0 / P.z_vib_fixed;
    }
    int idash = level;
    if (inv_zv > r.u()) {
        double func;
        int guard = 0;
        do {
            idash = random_label(r, 0, imax);
            const double evib = (double)idash * HS_KB * theta_v;
            func = dm_pow(1.0 - evib / ec, 1.5 - om);
        } while (func < r.u() && ++guard < HS_MAX_REDRAW);
    }
    return idash;
}

// dsmcCloud::postCollisionRotationalEnergy (dsmcCloud.C:1327-1387)
template <class R>
__device__ double rot_ratio(double dof, double chib, R& r) {
    if (dof == 2.0) return 1.0 - dm_pow(r.u(), 1.0 / chib);
    const double a1 = 0.5 * dof - 1.0, b1 = chib - 1.0;
    if (a1 < HS_SMALL && b1 < HS_SMALL) return r.u();
    double ratio, Pp;
    int guard = 0;
    do {
        ratio = r.u();
        if (a1 < HS_SMALL) Pp = dm_pow(1.0 - ratio, b1);
        else if (b1 < HS_SMALL) Pp = dm_pow(1.0 - ratio, a1);
        else Pp = dm_pow((a1 + b1) * ratio / a1, a1) * dm_pow((a1 + b1) * (1.0 - ratio) / b1, b1);
    } while (Pp < r.u() && ++guard < HS_MAX_REDRAW);
    return ratio;
}

// LarsenBorgnakkeVariableHardSphere::redistribute (LB-VHS.C:150-279) for a
// species with one electronic level
// INTERNAL false: every species has neither rotational nor vibrational
// modes, and only the electronic test draw remains.  POST = postReaction.
template <bool INTERNAL, bool POST = false, class R>
__device__ void lb_redistribute(const DevParams& P, int t, int partner, double* et, double* erot,
                                int* vib, double om, R& r, int cell = -1) {
    if (P.charge[t] == -1) return;
    if (1.0 / P.z_el > r.u()) (void)r.u();
    if (!INTERNAL) return;
    const int nm = P.nvib[t];
    if (nm > 0) {
        double pre[HS_MAX_VIB_MODES];
        for (int m = 0; m < nm; ++m) pre[m] = (double)vib[m] * HS_KB * P.theta_v[t][m];
        for (int m = 0; m < nm; ++m) {
            const double ec = *et + pre[m];
            const int imax = (int)(ec / (HS_KB * P.theta_v[t][m]));
            if (imax > 0) {
                vib[m] = vib_level<POST>(P, t, partner, m, vib[m], imax, ec, om, r, cell);
                *et = ec - (double)vib[m] * HS_KB * P.theta_v[t][m];
            }
        }
    }
    if (P.rotdof[t] > 0) {
        const double pre = *erot;
        if (1.0 / P.z_rot > r.u()) {
            const double ec = *et + pre;
            const double ratio = rot_ratio(P.rotdof[t], 2.5 - om, r);
            *erot = ratio * ec;
            *et = ec - *erot;
        }
    }
}

// BinaryCollisionModel::collide for the configured model (INTERNAL as for
// lb_redistribute)
template <bool INTERNAL = true, class R>
__device__ void binary_collide(const DevParams& P, int tp, int tq, double up[3], double uq[3],
                               double* erp, double* erq, int* vp, int* vq, R& r, int cell = -1) {
    const int model = P.coll_model;
    if (model == HS_COLL_VHS || model == HS_COLL_VSS) {
        post_coll_vel(P, model == HS_COLL_VSS, false, tp, tq, up, uq, -1.0, r);
        return;
    }
    const double mp = P.mass[tp], mq = P.mass[tq];
    const double mr = mp * mq / (mp + mq);
    const double dx = up[0] - uq[0], dy = up[1] - uq[1], dz = up[2] - uq[2];
    const double cr2 = dx * dx + dy * dy + dz * dz;
    double et = 0.5 * mr * cr2;
    const double om = P.omega_pq[tp][tq];
    lb_redistribute<INTERNAL>(P, tp, tq, &et, erp, vp, om, r, cell);
    lb_redistribute<INTERNAL>(P, tq, tp, &et, erq, vq, om, r, cell);
    const double cr = sqrt(2.0 * et / mr);
    post_coll_vel(P, model == HS_COLL_LB_VSS, true, tp, tq, up, uq, cr, r);
}

// ---------------------------------------------------------------------------
// quantum-kinetic chemistry (src/lagrangian/dsmc/reactions)
// ---------------------------------------------------------------------------
// one parcel's state during a reaction
struct PState {
    int t;
    double u[3];
    double er;
    int v[HS_MAX_VIB_MODES];
};

__device__ __forceinline__ bool is_lb(const DevParams& P) {
    return P.coll_model == HS_COLL_LB_VHS || P.coll_model == HS_COLL_LB_VSS;
}

// eVib_tot (dsmcParcelI.H:389-398)
__device__ __forceinline__ double evib_total(const DevParams& P, const PState& a) {
    double e = 0.0;
    for (int m = 0; m < P.nvib[a.t]; ++m) e += (double)a.v[m] * HS_KB * P.theta_v[a.t][m];
    return e;
}

__device__ __forceinline__ void become(PState& a, int t) {
    a.t = t;
    a.er = 0.0;
#pragma unroll
    for (int m = 0; m < HS_MAX_VIB_MODES; ++m) a.v[m] = 0;
}

// 1/2 mr |ua - ub|^2
__device__ __forceinline__ double rel_energy(const DevParams& P, const PState& a, const PState& b) {
    const double ma = P.mass[a.t], mb = P.mass[b.t];
    const double mr = ma * mb / (ma + mb);
    const double dx = a.u[0] - b.u[0], dy = a.u[1] - b.u[1], dz = a.u[2] - b.u[2];
    return 0.5 * mr * (dx * dx + dy * dy + dz * dz);
}

// dissociationQK::testDissociation (dissociationQK.C:199-238)
__device__ __forceinline__ void test_dissociation(const DevParams& P, const PState& a, double et, int* mode,
                                                  double* ec, double* total, double* prob) {
    const int kind = P.kind[a.t];
    if (kind != 20 && kind != 30) return;
    for (int m = 0; m < P.nvib[a.t]; ++m) {
        const double theta_v = P.theta_v[a.t][m];
        const double ev = (double)a.v[m] * HS_KB * theta_v;
        const int id = (int)(P.theta_d[a.t] / theta_v);
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

// exchangeQK::testExchange (exchangeQK.C:180-267); a is the molecule
__device__ __forceinline__ void test_exchange(const DevParams& P, const DevReaction& x, const PState& a, double et,
                                              double om, double* ec, double* total, double* prob) {
    const double chib = 2.5 - om;
    const double tcoll = et / (HS_KB * chib);
    double ea = x.a_dash * dm_pow(tcoll / 273.0, x.b) * fabs(x.heat_x);
    double sum = 1.0;
    if (x.heat_x < 0.0) ea -= x.heat_x;
    const int nm = P.nvib[a.t];
    int m = 0;
    do {
        const double th = P.theta_v[a.t][m];
        const double kth = HS_KB * th;
        const double ev = (double)a.v[m] * HS_KB * th;
        *ec = et + ev;
        if (*ec > ea) {
            if (ea > kth) {
                sum = 0.0;
                const int ia = (int)(*ec / kth);
                for (int i = 0; i <= ia; ++i) sum += dm_pow(1.0 - (double)i * HS_KB * th / *ec, 1.5 - om);
            }
            *prob = dm_pow(1.0 - ea / *ec, 1.5 - om) / sum;
            m = nm;
        }
        m += 1;
    } while (m < nm);
    *total += *prob;
}

// VariableHardSphere::postReactionVelocities (VariableHardSphere.C:231-262):
// ua enters as the pair's centre-of-mass velocity
template <class R>
__device__ __forceinline__ void post_reaction_vel(const DevParams& P, int ta, int tb, double ua[3], double ub[3],
                                                  double cr, R& r) {
    const double ma = P.mass[ta], mb = P.mass[tb];
    const double cos_t = 2.0 * r.u() - 1.0;
    const double sin_t = sqrt(1.0 - cos_t * cos_t);
    double sp, cp;
    dm_sincos2pi(r.u(), &sp, &cp);
    const double rel[3] = { cr * cos_t, cr * (sin_t * cp), cr * (sin_t * sp) };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ub[c] = ua[c] - rel[c] * ma / (ma + mb);
        ua[c] = ua[c] + rel[c] * mb / (ma + mb);
    }
}

// dissociationQK::dissociateParticleByPartner (dissociationQK.C:240-378): a
// (reactant nr) dissociates in vibrational mode `mode`, b is the partner; a
// becomes the first product and prod the second.  True when the split
// happens (allowSplitting).
template <class R>
__device__ bool dissociate(const DevParams& P, const DevReaction& x, PState& a, PState& b, int nr, int mode,
                           double ec, R& r, PState& prod) {
    if (!x.split) return false;
    ec -= x.heat_diss[nr];
    const double om = P.omega_pq[a.t][b.t];
    if (is_lb(P)) lb_redistribute<true, true>(P, b.t, a.t, &ec, &b.er, b.v, om, r);
    const double ma = P.mass[a.t], mb = P.mass[b.t];
    const double mr = ma * mb / (ma + mb);
    const double cr = sqrt(2.0 * ec / mr);
    post_coll_vel(P, false, false, a.t, b.t, a.u, b.u, cr, r);
    const int t1 = x.dprod[nr][0], t2 = x.dprod[nr][1];
    const double m1 = P.mass[t1], m2 = P.mass[t2];
    const double mrp = m1 * m2 / (m1 + m2);
    const double evt = evib_total(P, a);
    const double evm = (double)a.v[mode] * HS_KB * P.theta_v[a.t][mode];
    const double evn = evt - evm;
    const double left = a.er + evn + 0.0;  // + the single electronic level's energy, 0
    const double crp = sqrt(2.0 * left / mrp);
    become(prod, t2);
    post_reaction_vel(P, t1, t2, a.u, prod.u, crp, r);
    become(a, t1);
    return true;
}

// exchangeQK::exchange (exchangeQK.C:269-370): a is the molecule, b the atom
template <class R>
__device__ bool exchange(const DevParams& P, const DevReaction& x, PState& a, PState& b, R& r) {
    if (!x.split) return false;
    const double ma = P.mass[a.t], mb = P.mass[b.t];
    double et = rel_energy(P, a, b);
    double ucm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) ucm[c] = (ma * a.u[c] + mb * b.u[c]) / (ma + mb);
    const double mpx = P.mass[x.xatom], mqx = P.mass[x.xmol];
    const double mrx = mpx * mqx / (mpx + mqx);
    const double evt = evib_total(P, a);
    et += a.er + evt + 0.0 + 0.0 + x.heat_x;  // + both electronic energies, 0
    become(a, x.xatom);
    become(b, x.xmol);
    if (is_lb(P)) lb_redistribute<true, true>(P, b.t, a.t, &et, &b.er, b.v, x.om_x, r);
    const double crx = sqrt(2.0 * et / mrx);
#pragma unroll
    for (int c = 0; c < 3; ++c) a.u[c] = ucm[c];
    post_reaction_vel(P, x.xatom, x.xmol, a.u, b.u, crx, r);
    return true;
}

// dsmcReaction::decreasing_sort_indices (dsmcReaction.C:176-205): n draws,
// then the indices by probability, equal probabilities by their draw, both
// decreasing (insertion sort: the order is unique unless two draws tie)
template <class R>
__device__ __forceinline__ void decreasing_order(const double* v, int n, R& r, int* idx) {
    double rnd[3];
    for (int i = 0; i < n; ++i) {
        rnd[i] = r.u();
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

// The reaction of an accepted pair with a reaction model (noTimeCounter.C:
// 250-300): true if it reacted (relax_ is off and no binary collision
// follows).  A dissociation sets prod (prod.t >= 0) and from_q (Q split).
// The models swap P and Q when P is not their first reactant.
template <class R>
__device__ bool react_pair(const DevParams& P, int ri, PState& Pp, PState& Qp, R& r, PState& prod, int* from_q) {
    const DevReaction& x = P.rx[ri];
    PState* p = &Pp;
    PState* q = &Qp;
    if (x.model == HS_REACT_EXCHANGE_QK) {
        // exchangeQK::reaction (exchangeQK.C:457-517)
        const int kp = P.kind[p->t];
        if (kp == 10 || kp == 11) { p = &Qp; q = &Pp; }
        const double om = P.omega_pq[p->t][q->t];
        const double et = rel_energy(P, *p, *q);
        double total = 0.0, prob = 0.0, ec = 0.0;
        test_exchange(P, x, *p, et, om, &ec, &total, &prob);
        if (total > r.u()) return exchange(P, x, *p, *q, r);
        return false;
    }
    // dissociationQK::reaction (dissociationQK.C:469-600) and
    // dissociationExchangeQK::reaction (dissociationExchangeQK.C:98-270)
    if (p->t != x.r[0]) { p = &Qp; q = &Pp; }
    const bool q_is_Q = q == &Qp;
    const bool three = x.model == HS_REACT_DISSOCIATION_EXCHANGE_QK;
    const double om = P.omega_pq[p->t][q->t];
    const double et = rel_energy(P, *p, *q);
    double total = 0.0, prob[3] = { 0.0, 0.0, 0.0 }, ec[3] = { 0.0, 0.0, 0.0 };
    int mode_p = -1, mode_q = -1;
    test_dissociation(P, *p, et, &mode_p, &ec[0], &total, &prob[0]);
    if (three || p->t != q->t) test_dissociation(P, *q, et, &mode_q, &ec[1], &total, &prob[1]);
    if (three) {
        if (x.pos_mol == 0) test_exchange(P, x, *p, et, om, &ec[2], &total, &prob[2]);
        else test_exchange(P, x, *q, et, om, &ec[2], &total, &prob[2]);
    }
    if (!(total > r.u())) return false;
    const int n = three ? 3 : 2;
    double np[3];
    int idx[3];
    for (int i = 0; i < n; ++i) np[i] = prob[i] / total;
    decreasing_order(np, n, r, idx);
    double cum = 0.0;
    for (int k = 0; k < n; ++k) {
        const int i = idx[k];
        if (!(np[i] > HS_SMALL)) break;
        cum += np[i];
        if (cum > r.u()) {
            if (i == 0) {
                *from_q = q_is_Q ? 0 : 1;
                return dissociate(P, x, *p, *q, 0, mode_p, ec[0], r, prod);
            }
            if (i == 1) {
                *from_q = q_is_Q ? 1 : 0;
                return dissociate(P, x, *q, *p, 1, mode_q, ec[1], r, prod);
            }
            return x.pos_mol == 0 ? exchange(P, x, *p, *q, r) : exchange(P, x, *q, *p, r);
        }
    }
    return false;
}

// The accepted collision of one pair with chemistry on (noTimeCounter.C:
// 250-300): its reaction when the pair has a model, else (or when it did not
// react) the binary collision.  True if it reacted.
template <class R>
__device__ __attribute__((noinline)) bool collide_or_react(const DevParams& P, PState& a, PState& b, R& r,
                                                           PState& prod, int* from_q, int cell) {
    prod.t = -1;
    const int ri = P.react_of[a.t][b.t];
    if (ri >= 0 && react_pair(P, ri, a, b, r, prod, from_q)) return true;
    binary_collide<true>(P, a.t, b.t, a.u, b.u, &a.er, &b.er, a.v, b.v, r, cell);
    return false;
}

// ---------------------------------------------------------------------------
// walls
// ---------------------------------------------------------------------------
// q = n / d and n - q*d for 0 <= n < 2^31, 0 < d < 2^31, from the double
// reciprocal r = 1/d: the product n*r is within one of n/d (relative error
// below 2^-51), so one correction step makes the quotient exact.
__device__ __forceinline__ int div_exact(int n, int d, double r, int* rem) {
    int q = (int)((double)n * r);
    long long m = (long long)n - (long long)q * d;
    if (m < 0) { q -= 1; m += d; }
    else if (m >= d) { q += 1; m -= d; }
    *rem = (int)m;
    return q;
}

// The block indices (i, j, k) of cell c = i + nx*(j + ny*k).
__device__ __forceinline__ void cell_ijk(const DevParams& P, int c, int ijk[3]) {
    int r;
    ijk[2] = div_exact(c, P.nx * P.ny, P.rnxy, &r);
    ijk[1] = div_exact(r, P.nx, P.rnx, &ijk[0]);
}

__device__ __forceinline__ double face_coord(const DevParams& P, int a, int idx) {
    return P.origin[a] + (double)idx * P.delta[a];
}

// dsmcDiffuseWallPatch::performDiffuseReflection (dsmcDiffuseWallPatch.C:67-139)
// + dsmcPatchBoundary::calculateWallUnitVectors (dsmcPatchBoundary.C:216-261);
// nw points out of the gas.  Internal energies re-sampled at the wall T.
__device__ void diffuse_reflect(const DevParams& P, double T, const double uw[3], const double nw[3], int t,
                                double U[3], float* erot, uint16_t* vib, int64_t vstride,
                                PhiloxStream& r) {
    double und = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
    double ut[3] = { U[0] - und * nw[0], U[1] - und * nw[1], U[2] - und * nw[2] };
    int guard = 0;
    while (sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]) < HS_SMALL && guard++ < 100) {
        U[0] = U[0] * (0.8 + 0.2 * r.u());
        U[1] = U[1] * (0.8 + 0.2 * r.u());
        U[2] = U[2] * (0.8 + 0.2 * r.u());
        und = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
        ut[0] = U[0] - und * nw[0]; ut[1] = U[1] - und * nw[1]; ut[2] = U[2] - und * nw[2];
    }
    const double mut = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
    double tw1[3] = { ut[0] / mut, ut[1] / mut, ut[2] / mut };
    if (!(mut > 0.0)) { tw1[0] = nw[1]; tw1[1] = nw[2]; tw1[2] = nw[0]; }
    const double tw2[3] = { nw[1] * tw1[2] - nw[2] * tw1[1],
                            nw[2] * tw1[0] - nw[0] * tw1[2],
                            nw[0] * tw1[1] - nw[1] * tw1[0] };
    const double g1 = r.gauss();
    const double g2 = r.gauss();
    const double u = r.u();
    const double sc = sqrt(HS_KB * T / P.mass[t]);
    double arg = 1.0 - u;
    if (arg < HS_VSMALL) arg = HS_VSMALL;
    const double rn = sqrt(-2.0 * dm_log(arg));
#pragma unroll
    for (int c = 0; c < 3; ++c) U[c] = sc * ((g1 * tw1[c] + g2 * tw2[c]) - rn * nw[c]) + uw[c];
    if (erot) *erot = (float)equip_erot(r, T, P.rotdof[t]);
    if (vib)
        for (int m = 0; m < P.nvib[t]; ++m) vib[m * vstride] = (uint16_t)equip_vib_level(r, T, P.theta_v[t][m]);
}

// ---------------------------------------------------------------------------
// wall-face sampling: dsmcPatchBoundary::measurePropertiesBeforeControl /
// AfterControl (dsmcPatchBoundary.C:263-483) for one hit on wall face f; nw
// is the unit normal pointing out of the gas.  wall_side_terms adds the
// flux-weighted terms of one side of the interaction and returns the
// parcel's total energy; the momentum is m*U.
// ---------------------------------------------------------------------------
__device__ double wall_side_terms(const DevParams& P, int64_t f, int t, const double nw[3], const double U[3],
                                  const float* erot, const uint16_t* vib, int64_t vstride) {
    const double m = P.mass[t], adt = P.wall_area[f] * P.dt;
    const double un = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
    double dn = fabs(un) * adt;
    if (dn < HS_SMALL) dn = HS_SMALL;
    const double w = 1.0 / dn;
    const double ke = 0.5 * m * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    const int64_t st = P.n_wall;
    double* base = P.wall_sum + (int64_t)t * HS_WS_COUNT * st + f;
    unsafeAtomicAdd(base + HS_WS_RHON * st, w);
    unsafeAtomicAdd(base + HS_WS_KE * st, ke * w);
#pragma unroll
    for (int c = 0; c < 3; ++c) unsafeAtomicAdd(base + (HS_WS_MOMX + c) * st, m * (U[c] - un * nw[c]) * w);
    double e = ke;
    if (P.rotdof[t] > 0) {
        const double er = (double)*erot;
        unsafeAtomicAdd(base + HS_WS_EROT * st, er * w);
        e += er;
    }
    const int nv = P.nvib[t];
#pragma unroll
    for (int k = 0; k < HS_MAX_VIB_MODES; ++k) {
        if (k < nv) {
            const double ev = (double)vib[k * vstride] * HS_KB * P.theta_v[t][k];
            unsafeAtomicAdd(base + (HS_WS_EVIB0 + k) * st, ev * w);
            e += ev;
        }
    }
    return e;
}

// q and fD of one hit from the energies and velocities before (a) and after (b)
__device__ __forceinline__ void wall_exchange_terms(const DevParams& P, int64_t f, int t, double ea, const double Ua[3],
                                                    double eb, const double Ub[3]) {
    const double m = P.mass[t];
    const double cq = P.fn / (P.wall_area[f] * P.dt);
    const int64_t st = P.n_wall;
    double* base = P.wall_sum + (int64_t)t * HS_WS_COUNT * st + f;
    unsafeAtomicAdd(base + HS_WS_Q * st, cq * (ea - eb));
#pragma unroll
    for (int c = 0; c < 3; ++c) unsafeAtomicAdd(base + (HS_WS_FDX + c) * st, cq * (m * Ua[c] - m * Ub[c]));
}
