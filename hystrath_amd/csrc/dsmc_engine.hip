// dsmc_engine.hip -- the C ABI of include/hystrath_dsmc.h: engine lifetime,
// device memory, state transfer and the per-step launch sequence.
//
// One engine = one HIP stream on the device that was current at create time.
// Particle state is two structure-of-arrays buffers; every step moves the
// current buffer in place and the collide kernel writes the cell-sorted,
// collided population into the other one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dsmc_kernels.h"

namespace {

thread_local std::string g_create_error;

enum { T_INJECT = 0, T_MOVE, T_SCAN, T_SCATTER, T_COLLIDE, T_GATHER, T_NSTAGE };
constexpr int kMaxTimedPairs = 65536;

struct TimedPair {
    int stage;
    hipEvent_t a, b;
};

}  // namespace

static int scratch_on_device(int dev, void** out);

struct hs_dsmc_engine {
    hs_config cfg;
    std::vector<double> vol_h;
    std::vector<hs_reaction> reactions;  // chemReactDict, copied at create
    DevProducts prod = {};               // dissociation products (chemistry only)
    int device = 0;
    int nc = 0, ns = 0, general = 0, nslots = 0;
    int has_erot = 0, has_vib = 0, has_type = 0, max_modes = 0;
    int64_t cap = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;          // the next step's inflow, overlapped with this step's sort and collide
    hipEvent_t ev_scanned = nullptr;     // this step's cell scan is done (n_sorted known)
    hipEvent_t ev_injected = nullptr;    // the next step's inflow is written
    DevParams hp;
    DevParams* dp = nullptr;
    DevParticles buf[2];
    int cur = 0;
    std::vector<void*> allocs;
    double* d_vol = nullptr;
    double* d_t_overall = nullptr;   // inverseZvFormulation 2008: cells' overall temperature
    double* d_zv_tab = nullptr;      // 1/Zv by species, partner, mode and top level (LB)
    double* d_sigma = nullptr;
    double* d_rem = nullptr;
    double* d_samples = nullptr;
    int32_t* d_count = nullptr;
    int32_t* d_start = nullptr;
    int32_t* d_slot = nullptr;
    int32_t* d_perm = nullptr;
    uint16_t* d_big_list = nullptr;
    uint8_t* d_big_sub = nullptr;
    int32_t* d_tile_begin = nullptr;    // gather tiles
    int32_t* d_ctile_begin = nullptr;   // collide tiles
    int32_t* d_tile_order = nullptr;   // collide workgroup -> tile, heavy tiles first
    int32_t* d_tile_heavy = nullptr;
    WallQueue wq = { nullptr, nullptr, nullptr };  // flights stopped at walls
    int tile_parcels = 1536;     // parcels per gather workgroup (particle-balanced tiles)
    int collide_parcels = 2048;  // parcels per collide workgroup
    int tile_max[2] = { 1536, 2048 };   // gather, collide: the tile sizes of a full device
    int tile_req[2] = { 0, 0 };         // hs_dsmc_set_tiles; 0: chosen from the population
    int ncu = 256;
    int64_t n_hint = 0;  // population the host last saw (fill, upload, scalar reads)
    DevFaces faces;
    int32_t* d_face_off = nullptr;
    // wall faces (boundaryMeasurements)
    std::vector<int32_t> wall_side, wall_cell;
    std::vector<double> wall_area, wall_normal;
    double* d_wall_area = nullptr;
    int32_t* d_cyl_face = nullptr;
    double* d_wall_sum = nullptr;
    unsigned long long* d_face_flux = nullptr;
    DevScalars* d_S = nullptr;
    ScanWs ws;
    unsigned long long* d_lb_status = nullptr;  // look-back status words of the cell scan
    unsigned long long* d_lb_ticket = nullptr;  // its block tickets
    // evolve's steady-state step pair (move, sort+collide with the next
    // step's inflow ahead, twice) captured once per starting buffer: the
    // pair returns the buffers to where it found them, so one graph replays
    // every pair that starts on that buffer
    hipGraphExec_t pair_graph[2] = { nullptr, nullptr };
    int64_t pair_launches[T_NSTAGE] = { 0, 0, 0, 0, 0, 0 };
    bool graphs = true;  // HS_DSMC_GRAPHS=0 runs every step eagerly
    int64_t n_sampled = 0;
    int move_grid = 1024;     // grid-stride launches over the parcels (scatter, hist-only move)
    int flight_grid[2] = { 1024, 1024 };  // the moving launch, without / with the face tracker:
                                          // as many blocks as fit on the device at once
    std::string err;
    // timing
    std::vector<TimedPair> timed;
    std::vector<hipEvent_t> event_pool;
    int64_t launches[T_NSTAGE] = { 0, 0, 0, 0, 0, 0 };
    double closed_ms[T_NSTAGE] = { 0, 0, 0, 0, 0, 0 };
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
#define HS_TRY(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return fail(eng, HS_E_HIP, std::string(#expr) + ": " +       \
                                                             hipGetErrorString(e_));     \
    } while (0)

static int fail(hs_dsmc_engine* eng, int code, const std::string& msg) {
    if (eng) eng->err = msg;
    else g_create_error = msg;
    return code;
}

template <class T>
static int dev_alloc(hs_dsmc_engine* eng, T** p, int64_t count) {
    if (count <= 0) { *p = nullptr; return 0; }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)count * sizeof(T));
    if (e != hipSuccess) return fail(eng, HS_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    eng->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

static int check_launch(hs_dsmc_engine* eng, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(eng, HS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

static hipEvent_t take_event(hs_dsmc_engine* eng) {
    if (!eng->event_pool.empty()) {
        hipEvent_t e = eng->event_pool.back();
        eng->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct StageTimer {
    hs_dsmc_engine* eng;
    int stage;
    hipStream_t st;
    hipEvent_t a = nullptr;
    StageTimer(hs_dsmc_engine* e, int s, hipStream_t on = nullptr) : eng(e), stage(s), st(on ? on : e->stream) {
        eng->launches[s] += 1;
        if ((eng->cfg.flags & HS_FLAG_TIMING) && (int)eng->timed.size() < kMaxTimedPairs) {
            a = take_event(eng);
            if (a) (void)hipEventRecord(a, st);
        }
    }
    ~StageTimer() {
        if (!a) return;
        hipEvent_t b = take_event(eng);
        if (!b) return;
        (void)hipEventRecord(b, st);
        eng->timed.push_back({ stage, a, b });
    }
};

static DevParticles& cur_buf(hs_dsmc_engine* eng) { return eng->buf[eng->cur]; }
static DevParticles& nxt_buf(hs_dsmc_engine* eng) { return eng->buf[1 - eng->cur]; }

// inflow faces in the order: sides (xmin,xmax,ymin,ymax,zmin,zmax) with an
// inflow model, then the two tangential indices (second fastest)
struct HostFace {
    int32_t cell;
    int8_t axis, dir;
    int32_t a1i, a2i;
};

static std::vector<HostFace> list_faces(const hs_config& c, const std::vector<double>& vol) {
    std::vector<HostFace> out;
    const int nn[3] = { c.nx, c.ny, c.nz };
    for (int side = 0; side < 6; ++side) {
        if (!c.patch[side].inflow) continue;
        const int axis = side / 2, dir = (side & 1) ? 1 : -1;
        const int a1 = axis == 0 ? 1 : 0;
        const int a2 = axis == 2 ? 1 : 2;
        for (int j2 = 0; j2 < nn[a2]; ++j2)
            for (int j1 = 0; j1 < nn[a1]; ++j1) {
                int idx[3];
                idx[axis] = dir > 0 ? nn[axis] - 1 : 0;
                idx[a1] = j1;
                idx[a2] = j2;
                const int32_t cell = idx[0] + c.nx * (idx[1] + c.ny * idx[2]);
                if (!(vol[cell] > 0.0)) continue;
                out.push_back({ cell, (int8_t)axis, (int8_t)dir, j1, j2 });
            }
    }
    return out;
}

// Length of the circle (cx, cy, r) inside [x0,x1]x[y0,y1] and its
// length-weighted unit normal pointing to the centre: the circle's crossings
// of the four lines split [0, 2pi) into arcs wholly inside or outside.
static double arc_in_rect(double cx, double cy, double r, double x0, double x1, double y0, double y1, double n_out[2]) {
    const double pi = 3.14159265358979323846, two_pi = 6.28318530717958647692;
    std::vector<double> th = { 0.0, two_pi };
    const double xs[2] = { x0, x1 }, ys[2] = { y0, y1 };
    for (int j = 0; j < 2; ++j) {
        const double u = (xs[j] - cx) / r;
        if (u > -1.0 && u < 1.0) {
            const double a = acos(u);
            th.push_back(a);
            th.push_back(two_pi - a);
        }
        const double v = (ys[j] - cy) / r;
        if (v > -1.0 && v < 1.0) {
            double a = asin(v);
            th.push_back(pi - a);
            if (a < 0.0) a += two_pi;
            th.push_back(a);
        }
    }
    std::sort(th.begin(), th.end());
    double len = 0.0, nx = 0.0, ny = 0.0;
    for (size_t j = 0; j + 1 < th.size(); ++j) {
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
    const double nm = sqrt(nx * nx + ny * ny);
    n_out[0] = nm > 0.0 ? nx / nm : 0.0;
    n_out[1] = nm > 0.0 ? ny / nm : 0.0;
    return len;
}

// Wall faces: every face of each diffuse or specular side (first tangential
// index fastest), then one face per cell that the cylinder wall crosses.
static void list_wall_faces(hs_dsmc_engine* eng, int wall_off[6]) {
    const hs_config& c = eng->cfg;
    const int nn[3] = { c.nx, c.ny, c.nz };
    auto coord = [&](int a, int i) { return c.origin[a] + (double)i * c.delta[a]; };
    for (int side = 0; side < 6; ++side) {
        wall_off[side] = -1;
        const int m = c.patch[side].model;
        if (m != HS_PATCH_DIFFUSE && m != HS_PATCH_SPECULAR) continue;
        wall_off[side] = (int)eng->wall_side.size();
        const int axis = side / 2, dir = (side & 1) ? 1 : -1;
        const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
        for (int j2 = 0; j2 < nn[a2]; ++j2)
            for (int j1 = 0; j1 < nn[a1]; ++j1) {
                int idx[3];
                idx[axis] = dir > 0 ? nn[axis] - 1 : 0;
                idx[a1] = j1;
                idx[a2] = j2;
                eng->wall_side.push_back(side);
                eng->wall_cell.push_back(idx[0] + c.nx * (idx[1] + c.ny * idx[2]));
                eng->wall_area.push_back(c.delta[a1] * c.delta[a2]);
                double n[3] = { 0.0, 0.0, 0.0 };
                n[axis] = (double)dir;
                eng->wall_normal.insert(eng->wall_normal.end(), n, n + 3);
            }
    }
    if (!c.cyl_enabled) return;
    for (int k = 0; k < c.nz; ++k)
        for (int j = 0; j < c.ny; ++j)
            for (int i = 0; i < c.nx; ++i) {
                double nw[2];
                const double len = arc_in_rect(c.cyl_center[0], c.cyl_center[1], c.cyl_radius, coord(0, i),
                                               coord(0, i + 1), coord(1, j), coord(1, j + 1), nw);
                // arcs shorter than 1e-6 of the cell (slivers where the circle
                // grazes a grid line) carry no face
                if (!(len > 1e-6 * std::min(c.delta[0], c.delta[1]))) continue;
                eng->wall_side.push_back(6);
                eng->wall_cell.push_back(i + c.nx * (j + c.ny * k));
                eng->wall_area.push_back(len * c.delta[2]);
                eng->wall_normal.insert(eng->wall_normal.end(), { nw[0], nw[1], 0.0 });
            }
}

// a species' kind (dsmcParcelI.H:153-196): 0 electron, 10/11 atom, 20/21
// diatomic, 30/31 polyatomic molecule (neutral / ion)
static int species_kind(const hs_species& sp) {
    if (sp.charge == -1) return 0;
    const int base = sp.n_vib_modes == 0 ? 10 : (sp.n_vib_modes == 1 ? 20 : 30);
    return base + (sp.charge == 1 ? 1 : 0);
}

// whether reaction x applies to the pair (p, q): tryReactMolecules of
// dissociationQK.C:430-455 (dissociation models) and exchangeQK.C:420-443
static bool reaction_applies(const hs_reaction& x, int p, int q) {
    const int pid = p == x.reactants[0] ? 0 : (p == x.reactants[1] ? 1 : -1);
    const int qid = q == x.reactants[0] ? 0 : (q == x.reactants[1] ? 1 : -1);
    if (pid < 0 || qid < 0) return false;
    if (x.model == HS_REACT_EXCHANGE_QK) return pid != qid;
    const bool same = x.reactants[0] == x.reactants[1];
    return (pid == qid && same) || (pid != qid && !same);
}

// the checks of dsmcReaction / dissociationQK / exchangeQK::setProperties
// (dissociationQK.C:33-197, exchangeQK.C:44-178) and the one-model-per-pair
// rule of dsmcReactions.C:139-160
static int validate_reactions(const hs_config* c) {
    if (c->n_reactions == 0) return 0;
    if (c->n_reactions < 0 || c->n_reactions > HS_MAX_REACTIONS)
        return fail(nullptr, HS_E_INVALID, "n_reactions out of range");
    if (!c->reactions) return fail(nullptr, HS_E_INVALID, "reactions is NULL");
    if (c->collision_model != HS_COLL_VHS && c->collision_model != HS_COLL_LB_VHS)
        return fail(nullptr, HS_E_UNSUPPORTED,
                    "chemical reactions need VariableHardSphere or LarsenBorgnakkeVariableHardSphere scattering");
    const int ns = c->n_species;
    auto kind = [&](int t) { return species_kind(c->species[t]); };
    auto is_mol = [&](int t) { return kind(t) == 20 || kind(t) == 30; };
    for (int i = 0; i < c->n_reactions; ++i) {
        const hs_reaction& x = c->reactions[i];
        if (x.model < HS_REACT_DISSOCIATION_QK || x.model > HS_REACT_DISSOCIATION_EXCHANGE_QK)
            return fail(nullptr, HS_E_INVALID, "unknown reactionModel");
        for (int r = 0; r < 2; ++r)
            if (x.reactants[r] < 0 || x.reactants[r] >= ns)
                return fail(nullptr, HS_E_INVALID, "Cannot find type id of a reactant");
        if (x.model != HS_REACT_EXCHANGE_QK) {
            if (!is_mol(x.reactants[0]) && !is_mol(x.reactants[1]))
                return fail(nullptr, HS_E_INVALID, "None of the reactants is a molecule");
            for (int r = 0; r < 2; ++r) {
                const int p0 = x.diss_products[r][0], p1 = x.diss_products[r][1];
                if (p0 < 0 && p1 < 0) {
                    if (is_mol(x.reactants[r]))
                        return fail(nullptr, HS_E_INVALID, "a molecule reactant should have dissociation products");
                    continue;
                }
                if (!is_mol(x.reactants[r]))
                    return fail(nullptr, HS_E_INVALID,
                                "a reactant that is not a molecule should have no dissociation products");
                if (p0 < 0 || p1 < 0 || p0 >= ns || p1 >= ns)
                    return fail(nullptr, HS_E_INVALID, "There should be 2 dissociation products of known type ids");
                if (kind(x.reactants[r]) == 30)
                    return fail(nullptr, HS_E_INVALID,
                                "Dissociation product of a polyatomic molecule must be a diatomic/polyatomic molecule");
                if (kind(p0) != 10 || kind(p1) != 10)
                    return fail(nullptr, HS_E_INVALID, "Dissociation product of a diatomic molecule must be an atom");
            }
        }
        if (x.model != HS_REACT_DISSOCIATION_QK) {
            bool mol = false, atom = false;
            for (int r = 0; r < 2; ++r) {
                const int k = kind(x.reactants[r]);
                if (k >= 20) mol = true;
                else if (k == 10 || k == 11) atom = true;
                else return fail(nullptr, HS_E_INVALID, "an exchange reactant must be an atom or a molecule");
            }
            if (!mol) return fail(nullptr, HS_E_INVALID, "None of the exchange reactants is a molecule");
            if (!atom) return fail(nullptr, HS_E_INVALID, "None of the exchange reactants is an atom");
            mol = atom = false;
            for (int r = 0; r < 2; ++r) {
                const int t = x.exch_products[r];
                if (t < 0 || t >= ns) return fail(nullptr, HS_E_INVALID, "Cannot find type id of an exchange product");
                const int k = kind(t);
                if (k >= 20) mol = true;
                else if (k == 10 || k == 11) atom = true;
                else return fail(nullptr, HS_E_INVALID, "an exchange product must be an atom or a molecule");
            }
            if (!mol || !atom)
                return fail(nullptr, HS_E_INVALID, "the exchange products must be one molecule and one atom");
        }
    }
    for (int p = 0; p < ns; ++p)
        for (int q = p; q < ns; ++q) {
            int found = 0;
            for (int i = 0; i < c->n_reactions; ++i) found += reaction_applies(c->reactions[i], p, q) ? 1 : 0;
            if (found > 1)
                return fail(nullptr, HS_E_INVALID, "There is more than one reaction model specified for a typeId pair");
        }
    return 0;
}

static int validate(const hs_config* c) {
    if (!c) return fail(nullptr, HS_E_INVALID, "config is NULL");
    if (c->nx <= 0 || c->ny <= 0 || c->nz <= 0) return fail(nullptr, HS_E_INVALID, "mesh dimensions must be positive");
    const int64_t nc = (int64_t)c->nx * c->ny * c->nz;
    if (nc > (int64_t)1 << 30) return fail(nullptr, HS_E_INVALID, "too many cells");
    for (int a = 0; a < 3; ++a)
        if (!(c->delta[a] > 0.0)) return fail(nullptr, HS_E_INVALID, "cell size must be positive");
    if (!c->cell_volume) return fail(nullptr, HS_E_INVALID, "cell_volume is NULL");
    if (c->n_species < 1 || c->n_species > HS_MAX_SPECIES) return fail(nullptr, HS_E_INVALID, "n_species out of range");
    for (int s = 0; s < c->n_species; ++s) {
        const hs_species& sp = c->species[s];
        if (!(sp.mass > 0.0) || !(sp.diameter > 0.0)) return fail(nullptr, HS_E_INVALID, "species mass/diameter must be positive");
        if (sp.n_vib_modes < 0 || sp.n_vib_modes > HS_MAX_VIB_MODES)
            return fail(nullptr, HS_E_INVALID, "nVibrationalModes out of range");
        if (sp.charge < -1 || sp.charge > 1)
            return fail(nullptr, HS_E_INVALID, "Charge value should be 0 for neutrals, 1 for ions, or -1 for electrons");
        for (int m = 0; m < sp.n_vib_modes; ++m)
            if (!(sp.theta_v[m] > 0.0)) return fail(nullptr, HS_E_INVALID, "characteristicVibrationalTemperature must be positive");
    }
    if (c->collision_model < HS_COLL_NONE || c->collision_model > HS_COLL_LB_VSS)
        return fail(nullptr, HS_E_INVALID, "unknown BinaryCollisionModel");
    if (c->collision_model >= HS_COLL_LB_VHS) {
        if (!(c->z_rot > 0.0) || !(c->z_el > 0.0))
            return fail(nullptr, HS_E_INVALID, "relaxation collision numbers must be positive");
        if (c->inv_zv_form < 0 || c->inv_zv_form > 2) return fail(nullptr, HS_E_INVALID, "inverseZvFormulation out of range");
    }
    const int nn[3] = { c->nx, c->ny, c->nz };
    for (int a = 0; a < 3; ++a) {
        const int m0 = c->patch[2 * a].model, m1 = c->patch[2 * a + 1].model;
        if (m0 < HS_PATCH_CYCLIC || m0 > HS_PATCH_EMPTY || m1 < HS_PATCH_CYCLIC || m1 > HS_PATCH_EMPTY)
            return fail(nullptr, HS_E_INVALID, "unknown patch model");
        if ((m0 == HS_PATCH_CYCLIC) != (m1 == HS_PATCH_CYCLIC))
            return fail(nullptr, HS_E_INVALID, "cyclic patches must come in pairs");
        if ((m0 == HS_PATCH_EMPTY) != (m1 == HS_PATCH_EMPTY))
            return fail(nullptr, HS_E_INVALID, "empty patches must come in pairs");
        if (m0 == HS_PATCH_EMPTY && nn[a] != 1) return fail(nullptr, HS_E_INVALID, "an empty direction must be one cell thick");
        if ((c->patch[2 * a].inflow || c->patch[2 * a + 1].inflow) && m0 == HS_PATCH_EMPTY)
            return fail(nullptr, HS_E_INVALID, "inflow on an empty patch");
    }
    if (c->cyl_enabled) {
        if (!(c->cyl_radius > 0.0)) return fail(nullptr, HS_E_INVALID, "cylinder radius must be positive");
        if (c->cyl_patch.model != HS_PATCH_DIFFUSE && c->cyl_patch.model != HS_PATCH_SPECULAR)
            return fail(nullptr, HS_E_INVALID, "cylinder wall must be diffuse or specular");
    }
    if (c->capacity <= 0 || c->capacity >= ((int64_t)1 << 31) - 1) return fail(nullptr, HS_E_INVALID, "capacity out of range");
    if (!(c->fn > 0.0) || !(c->dt > 0.0)) return fail(nullptr, HS_E_INVALID, "nEquivalentParticles and deltaT must be positive");
    return validate_reactions(c);
}

static void derive(hs_dsmc_engine* eng) {
    const hs_config& c = eng->cfg;
    DevParams& P = eng->hp;
    memset(&P, 0, sizeof(P));
    P.nx = c.nx; P.ny = c.ny; P.nz = c.nz;
    P.nc = eng->nc;
    P.rnx = 1.0 / (double)c.nx;
    P.rnxy = 1.0 / ((double)c.nx * (double)c.ny);
    const int nn[3] = { c.nx, c.ny, c.nz };
    for (int a = 0; a < 3; ++a) {
        P.origin[a] = c.origin[a];
        P.delta[a] = c.delta[a];
        P.empty[a] = c.patch[2 * a].model == HS_PATCH_EMPTY;
        P.centre[a] = c.origin[a] + 0.5 * ((double)nn[a] * c.delta[a]);
    }
    for (int s = 0; s < 6; ++s) {
        P.patch_model[s] = c.patch[s].model;
        P.patch_T[s] = c.patch[s].temperature;
        for (int a = 0; a < 3; ++a) P.patch_U[s][a] = c.patch[s].velocity[a];
    }
    P.cyl = c.cyl_enabled;
    P.cyl_c[0] = c.cyl_center[0];
    P.cyl_c[1] = c.cyl_center[1];
    P.cyl_r = c.cyl_radius;
    P.cyl_model = c.cyl_patch.model;
    P.cyl_T = c.cyl_patch.temperature;
    for (int a = 0; a < 3; ++a) P.cyl_U[a] = c.cyl_patch.velocity[a];
    P.ns = c.n_species;
    for (int p = 0; p < P.ns; ++p) {
        const hs_species& sp = c.species[p];
        P.mass[p] = sp.mass;
        P.rotdof[p] = sp.rot_dof;
        P.nvib[p] = sp.n_vib_modes;
        P.charge[p] = sp.charge;
        P.theta_d[p] = sp.theta_d;
        if (sp.n_vib_modes > P.max_modes) P.max_modes = sp.n_vib_modes;
        for (int m = 0; m < sp.n_vib_modes; ++m) {
            P.theta_v[p][m] = sp.theta_v[m];
            P.zv_pow2[p][m] = pow(sp.theta_d / sp.tref_zv[m], 1.0 / 3.0) - 1.0;
        }
        for (int q = 0; q < P.ns; ++q) {
            const hs_species& sq = c.species[q];
            const double dpq = 0.5 * (sp.diameter + sq.diameter);
            const double om = 0.5 * (sp.omega + sq.omega);
            const double mr = sp.mass * sq.mass / (sp.mass + sq.mass);
            P.sig_coef[p][q] = 3.14159265358979323846 * dpq * dpq / exp(lgamma(2.5 - om));
            P.ref_term[p][q] = 2.0 * HS_KB * c.tref / mr;
            P.omega_pq[p][q] = om;
            P.alpha_pq[p][q] = 0.5 * (sp.alpha + sq.alpha);
            for (int m = 0; m < sp.n_vib_modes; ++m) P.zv_b[p][q][m] = sp.zref[m] * pow(sp.theta_d / sp.tref_zv[m], -om);
        }
    }
    P.coll_model = c.collision_model;
    P.z_rot = c.z_rot;
    P.z_el = c.z_el;
    P.z_vib_fixed = c.z_vib_fixed;
    P.inv_zv_form = c.inv_zv_form;
    P.fn = c.fn;
    P.dt = c.dt;
    P.seed_lo = (uint32_t)c.seed;
    P.seed_hi = (uint32_t)(c.seed >> 32);
    P.flags = c.flags;
    P.nslots = eng->nslots;
    P.has_erot = eng->has_erot;
    P.has_vib = eng->has_vib;
    P.has_type = eng->has_type;
    P.capacity = eng->cap;
    P.inflow_t = c.inflow_t;
    P.inflow_trot = c.inflow_trot;
    P.inflow_tvib = c.inflow_tvib;
    for (int a = 0; a < 3; ++a) P.inflow_u[a] = c.inflow_u[a];
    for (int s = 0; s < HS_MAX_SPECIES; ++s) P.inflow_n[s] = c.inflow_n[s];
    // chemistry: the reactions' constants (dissociationQK.C:57-60,
    // exchangeQK.C:195-200,303-308,387-391) and the pair table
    P.n_react = (int)eng->reactions.size();
    for (int p = 0; p < HS_MAX_SPECIES; ++p) {
        P.kind[p] = p < P.ns ? species_kind(c.species[p]) : 0;
        for (int q = 0; q < HS_MAX_SPECIES; ++q) P.react_of[p][q] = -1;
    }
    for (int i = 0; i < P.n_react; ++i) {
        const hs_reaction& R = eng->reactions[i];
        DevReaction& x = P.rx[i];
        x.model = R.model;
        x.split = R.allow_splitting != 0;
        for (int k = 0; k < 2; ++k) {
            x.r[k] = R.reactants[k];
            x.dprod[k][0] = R.diss_products[k][0];
            x.dprod[k][1] = R.diss_products[k][1];
            x.heat_diss[k] = HS_KB * c.species[R.reactants[k]].theta_d;
        }
        x.pos_mol = -1;
        x.xmol = x.xatom = -1;
        if (R.model != HS_REACT_DISSOCIATION_QK) {
            for (int k = 0; k < 2; ++k)
                if (species_kind(c.species[R.reactants[k]]) >= 20) x.pos_mol = k;
            for (int k = 0; k < 2; ++k) {
                const int t = R.exch_products[k];
                const int kd = species_kind(c.species[t]);
                if (kd >= 20) x.xmol = t;
                else if (kd == 10 || kd == 11) x.xatom = t;
            }
            x.heat_x = R.heat_exchange * HS_KB;
            x.b = R.b_coeff;
            const double om = 0.5 * (c.species[R.reactants[0]].omega + c.species[R.reactants[1]].omega);
            const double chib = 2.5 - om;
            x.a_dash = R.a_coeff * (pow(chib, R.b_coeff) * exp(lgamma(chib)) / exp(lgamma(chib + R.b_coeff)));
            x.om_x = 0.5 * (c.species[x.xatom].omega + c.species[x.xatom].omega);
        }
        for (int p = 0; p < P.ns; ++p)
            for (int q = p; q < P.ns; ++q)
                if (reaction_applies(R, p, q)) { P.react_of[p][q] = i; P.react_of[q][p] = i; }
    }
}

static int alloc_particles(hs_dsmc_engine* eng, DevParticles& b) {
    const int64_t cap = eng->cap;
    int rc;
    if ((rc = dev_alloc(eng, &b.px, cap)) || (rc = dev_alloc(eng, &b.py, cap)) || (rc = dev_alloc(eng, &b.pz, cap)) ||
        (rc = dev_alloc(eng, &b.ux, cap)) || (rc = dev_alloc(eng, &b.uy, cap)) || (rc = dev_alloc(eng, &b.uz, cap)) ||
        (rc = dev_alloc(eng, &b.cell, cap)))
        return rc;
    b.erot = nullptr;
    b.vib = nullptr;
    b.type = nullptr;
    if (eng->has_erot && (rc = dev_alloc(eng, &b.erot, cap))) return rc;
    if (eng->has_vib && (rc = dev_alloc(eng, &b.vib, cap * eng->max_modes))) return rc;
    if (eng->has_type && (rc = dev_alloc(eng, &b.type, cap))) return rc;
    return 0;
}

// ---------------------------------------------------------------------------
// lifetime
// ---------------------------------------------------------------------------
extern "C" const char* hs_dsmc_create_error(void) { return g_create_error.c_str(); }

extern "C" const char* hs_dsmc_last_error(const hs_dsmc_engine* eng) {
    return eng ? eng->err.c_str() : g_create_error.c_str();
}

extern "C" int hs_dsmc_device_count(int* n) {
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
    if (n) *n = k;
    return 0;
}

extern "C" int32_t hs_dsmc_sample_slots(const hs_config* cfg) {
    int32_t n = HS_S_GLOBAL_COUNT + cfg->n_species * HS_SS_COUNT;
    if (cfg->flags & HS_FLAG_HEATFLUX) n += cfg->n_species * HS_HF_COUNT;
    return n;
}

static void drop_graphs(hs_dsmc_engine* eng) {
    for (auto& g : eng->pair_graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

extern "C" void hs_dsmc_destroy(hs_dsmc_engine* eng) {
    if (!eng) return;
    (void)hipSetDevice(eng->device);
    if (eng->stream) (void)hipStreamSynchronize(eng->stream);
    if (eng->side) (void)hipStreamSynchronize(eng->side);
    drop_graphs(eng);
    if (eng->ev_scanned) (void)hipEventDestroy(eng->ev_scanned);
    if (eng->ev_injected) (void)hipEventDestroy(eng->ev_injected);
    for (auto& t : eng->timed) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : eng->event_pool) (void)hipEventDestroy(e);
    for (void* p : eng->allocs) (void)hipFree(p);
    if (eng->side) (void)hipStreamDestroy(eng->side);
    if (eng->stream) (void)hipStreamDestroy(eng->stream);
    delete eng;
}

extern "C" int hs_dsmc_create(const hs_config* cfg, hs_dsmc_engine** out) {
    g_create_error.clear();
    if (!out) return fail(nullptr, HS_E_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, HS_E_HIP, "no HIP device available (the engine has no CPU path)");
    hs_dsmc_engine* eng = new hs_dsmc_engine();
    eng->cfg = *cfg;
    if (const char* g = getenv("HS_DSMC_GRAPHS")) eng->graphs = std::strcmp(g, "0") != 0;
    (void)hipGetDevice(&eng->device);
    eng->nc = cfg->nx * cfg->ny * cfg->nz;
    eng->ns = cfg->n_species;
    eng->cap = cfg->capacity;
    eng->vol_h.assign(cfg->cell_volume, cfg->cell_volume + eng->nc);
    eng->cfg.cell_volume = eng->vol_h.data();
    if (cfg->n_reactions > 0) eng->reactions.assign(cfg->reactions, cfg->reactions + cfg->n_reactions);
    eng->cfg.reactions = eng->reactions.empty() ? nullptr : eng->reactions.data();
    for (int s = 0; s < eng->ns; ++s) {
        if (cfg->species[s].rot_dof > 0) eng->has_erot = 1;
        if (cfg->species[s].n_vib_modes > 0) eng->has_vib = 1;
        if (cfg->species[s].n_vib_modes > eng->max_modes) eng->max_modes = cfg->species[s].n_vib_modes;
    }
    eng->has_type = eng->ns > 1;
    eng->general = eng->has_erot || eng->has_vib || eng->has_type;
    // collide tiles sized to the LDS chunk of the kernel variant
    eng->collide_parcels = eng->general ? 1536 : 2048;
    eng->tile_max[1] = eng->collide_parcels;
    eng->nslots = hs_dsmc_sample_slots(cfg);
    derive(eng);
    auto bail = [&](int code) {
        g_create_error = eng->err;
        hs_dsmc_destroy(eng);
        return code;
    };
    {
        // the fixed scratch of hs_dsmc_fill_noalloc on this device
        void* sp = nullptr;
        if (scratch_on_device(eng->device, &sp)) return bail(fail(eng, HS_E_HIP, "hipGetSymbolAddress failed"));
    }
    if (hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&eng->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&eng->ev_scanned, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&eng->ev_injected, hipEventDisableTiming) != hipSuccess)
        return bail(fail(eng, HS_E_HIP, "hipStreamCreate / hipEventCreate failed"));
    const int nc = eng->nc;
    if ((rc = dev_alloc(eng, &eng->dp, 1)) || (rc = alloc_particles(eng, eng->buf[0])) ||
        (rc = alloc_particles(eng, eng->buf[1])) || (rc = dev_alloc(eng, &eng->d_vol, nc)) ||
        (rc = dev_alloc(eng, &eng->d_sigma, nc)) || (rc = dev_alloc(eng, &eng->d_rem, nc)) ||
        (rc = dev_alloc(eng, &eng->d_samples, (int64_t)eng->nslots * nc)) ||
        (rc = dev_alloc(eng, &eng->d_count, nc)) || (rc = dev_alloc(eng, &eng->d_start, (int64_t)nc + 1)) ||
        (rc = dev_alloc(eng, &eng->d_slot, eng->cap)) || (rc = dev_alloc(eng, &eng->d_perm, eng->cap)) ||
        (rc = dev_alloc(eng, &eng->d_big_list, eng->cap)) || (rc = dev_alloc(eng, &eng->d_big_sub, eng->cap)) ||
        (rc = dev_alloc(eng, &eng->d_tile_begin, eng->cap / 64 + 4)) ||
        (rc = dev_alloc(eng, &eng->d_ctile_begin, eng->cap / 64 + 4)) ||
        (rc = dev_alloc(eng, &eng->d_tile_order, eng->cap / 64 + 4)) || 
        (rc = dev_alloc(eng, &eng->d_tile_heavy, eng->cap / 64 + 4)) ||
        (rc = dev_alloc(eng, &eng->wq.i, eng->cap)) || (rc = dev_alloc(eng, &eng->wq.t, eng->cap)) ||
        (rc = dev_alloc(eng, &eng->wq.code_iter, eng->cap)) ||
        (rc = dev_alloc(eng, &eng->d_S, 1)))
        return bail(rc);
    if (!eng->reactions.empty()) {
        DevProducts& pr = eng->prod;
        const int64_t cap = eng->cap;
        if ((rc = dev_alloc(eng, &pr.cell, cap)) || (rc = dev_alloc(eng, &pr.rank, cap)) ||
            (rc = dev_alloc(eng, &pr.parent, cap)) || (rc = dev_alloc(eng, &pr.type, cap)) ||
            (rc = dev_alloc(eng, &pr.ux, cap)) || (rc = dev_alloc(eng, &pr.uy, cap)) ||
            (rc = dev_alloc(eng, &pr.uz, cap)) || (rc = dev_alloc(eng, &pr.count, nc)) ||
            (rc = dev_alloc(eng, &pr.off, (int64_t)nc + 1)) || (rc = dev_alloc(eng, &pr.list, cap)) ||
            (rc = dev_alloc(eng, &pr.bigcells, nc)))
            return bail(rc);
        launch_fill_i32(pr.count, 0, nc, eng->stream);
        launch_fill_i32(pr.off, 0, (int64_t)nc + 1, eng->stream);
    }
    // inflow faces
    std::vector<HostFace> hf = list_faces(eng->cfg, eng->vol_h);
    eng->faces.n = (int32_t)hf.size();
    const int nf = eng->faces.n, ns = eng->ns;
    // scan workspace sized for the largest scanned array (cells, faces*ns, cells*ns)
    int64_t scan_n = (int64_t)nc * ns;
    if ((int64_t)nf * ns > scan_n) scan_n = (int64_t)nf * ns;
    eng->ws.nb_max = (int32_t)((scan_n + 2047) / 2048 + 1);
    if ((rc = dev_alloc(eng, &eng->ws.bsum, eng->ws.nb_max)) ||
        (rc = dev_alloc(eng, &eng->d_lb_status, scan_cells_blocks(nc))) || (rc = dev_alloc(eng, &eng->d_lb_ticket, 1)))
        return bail(rc);
    launch_fill_i64(reinterpret_cast<int64_t*>(eng->d_lb_status), 0, scan_cells_blocks(nc), eng->stream);
    launch_fill_i64(reinterpret_cast<int64_t*>(eng->d_lb_ticket), 0, 1, eng->stream);
    if (nf > 0) {
        std::vector<int32_t> fc(nf), a1(nf), a2(nf);
        std::vector<int8_t> ax(nf), dr(nf);
        std::vector<double> inc((size_t)ns * nf, 0.0);
        const double sqrt_pi = sqrt(3.14159265358979323846);
        for (int f = 0; f < nf; ++f) {
            fc[f] = hf[f].cell; ax[f] = hf[f].axis; dr[f] = hf[f].dir; a1[f] = hf[f].a1i; a2[f] = hf[f].a2i;
        }
        for (int s = 0; s < ns; ++s) {
            if (!(cfg->inflow_n[s] > 0.0)) continue;
            const double mass = cfg->species[s].mass;
            const double mps = sqrt(2.0 * HS_KB * cfg->inflow_t / mass);
            for (int f = 0; f < nf; ++f) {
                const int a = hf[f].axis;
                const int a1i = a == 0 ? 1 : 0, a2i = a == 2 ? 1 : 2;
                const double fa = cfg->delta[a1i] * cfg->delta[a2i];
                const double s_cos = (cfg->inflow_u[a] * (double)(-hf[f].dir)) / mps;
                inc[(size_t)s * nf + f] = (fa * cfg->inflow_n[s] * cfg->dt * mps
                                           * (exp(-(s_cos * s_cos)) + sqrt_pi * s_cos * (1.0 + erf(s_cos))))
                                          / (2.0 * sqrt_pi * cfg->fn);
            }
        }
        // ni <= floor(acc + inc) + 1 with acc < 1 after every step
        double max_new = 0.0;
        for (double x : inc) max_new += std::floor(x) + 2.0;
        if (max_new > (double)eng->cap) max_new = (double)eng->cap;
        eng->faces.max_new = (int32_t)max_new;
        DevFaces& F = eng->faces;
        if ((rc = dev_alloc(eng, &F.cell, nf)) || (rc = dev_alloc(eng, &F.axis, nf)) ||
            (rc = dev_alloc(eng, &F.dir, nf)) || (rc = dev_alloc(eng, &F.a1i, nf)) || (rc = dev_alloc(eng, &F.a2i, nf)) ||
            (rc = dev_alloc(eng, &F.inc, (int64_t)ns * nf)) || (rc = dev_alloc(eng, &F.acc, (int64_t)ns * nf)) ||
            (rc = dev_alloc(eng, &F.n_new, (int64_t)ns * nf)) || (rc = dev_alloc(eng, &F.off, (int64_t)ns * nf + 1)))
            return bail(rc);
        if (hipMemcpy(F.cell, fc.data(), nf * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(F.axis, ax.data(), nf, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(F.dir, dr.data(), nf, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(F.a1i, a1.data(), nf * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(F.a2i, a2.data(), nf * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(F.inc, inc.data(), inc.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(eng, HS_E_HIP, "hipMemcpy of inflow faces failed"));
        launch_fill_f64(F.acc, 0.0, (int64_t)ns * nf, eng->stream);
    }
    // wall faces
    list_wall_faces(eng, eng->hp.wall_off);
    const int64_t nw = (int64_t)eng->wall_side.size();
    eng->hp.n_wall = (int)nw;
    if (nw > 0) {
        if ((rc = dev_alloc(eng, &eng->d_wall_area, nw)) ||
            (rc = dev_alloc(eng, &eng->d_wall_sum, (int64_t)ns * HS_WS_COUNT * nw)))
            return bail(rc);
        if (hipMemcpy(eng->d_wall_area, eng->wall_area.data(), nw * 8, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(eng, HS_E_HIP, "hipMemcpy of wall faces failed"));
        launch_fill_f64(eng->d_wall_sum, 0.0, (int64_t)ns * HS_WS_COUNT * nw, eng->stream);
        if (cfg->cyl_enabled) {
            std::vector<int32_t> cf(nc, -1);
            for (int64_t f = 0; f < nw; ++f)
                if (eng->wall_side[f] == 6) cf[eng->wall_cell[f]] = (int32_t)f;
            if ((rc = dev_alloc(eng, &eng->d_cyl_face, nc))) return bail(rc);
            if (hipMemcpy(eng->d_cyl_face, cf.data(), (size_t)nc * 4, hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(eng, HS_E_HIP, "hipMemcpy of cylinder faces failed"));
        }
    }
    eng->hp.wall_area = eng->d_wall_area;
    eng->hp.cyl_face = eng->d_cyl_face;
    eng->hp.wall_sum = eng->d_wall_sum;
    eng->hp.cyl_first = (int)nw;
    for (int64_t f = 0; f < nw; ++f)
        if (eng->wall_side[f] == 6) { eng->hp.cyl_first = (int)f; break; }
    // face tracker
    {
        const int64_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
        eng->hp.face_off[0] = 0;
        eng->hp.face_off[1] = (nx + 1) * ny * nz;
        eng->hp.face_off[2] = eng->hp.face_off[1] + nx * (ny + 1) * nz;
        eng->hp.face_off[3] = eng->hp.face_off[2] + nx * ny * (nz + 1);
        eng->hp.n_faces = eng->hp.face_off[3] + (nw - eng->hp.cyl_first);
        if ((rc = dev_alloc(eng, &eng->d_face_flux, (int64_t)ns * eng->hp.n_faces))) return bail(rc);
        launch_fill_i64((int64_t*)eng->d_face_flux, 0, (int64_t)ns * eng->hp.n_faces, eng->stream);
        eng->hp.face_flux = eng->d_face_flux;
    }
    // inverseZvFormulation 2008: the cells' overall temperature, zero (the
    // collision temperature is used) until hs_dsmc_set_cell_temperature
    eng->hp.t_overall = nullptr;
    if (cfg->collision_model >= HS_COLL_LB_VHS && cfg->inv_zv_form == 1) {
        if ((rc = dev_alloc(eng, &eng->d_t_overall, nc))) return bail(rc);
        launch_fill_f64(eng->d_t_overall, 0.0, nc, eng->stream);
        eng->hp.t_overall = eng->d_t_overall;
    }
    // Larsen-Borgnakke with a computed Zv: its values at the collision
    // temperatures of the levels below HS_ZV_TAB, filled once below
    eng->hp.zv_tab = nullptr;
    if (cfg->collision_model >= HS_COLL_LB_VHS && cfg->z_vib_fixed == 0.0 && eng->hp.max_modes > 0) {
        if ((rc = dev_alloc(eng, &eng->d_zv_tab, (int64_t)ns * ns * HS_MAX_VIB_MODES * HS_ZV_TAB))) return bail(rc);
        eng->hp.zv_tab = eng->d_zv_tab;
    }
    if (hipMemcpy(eng->dp, &eng->hp, sizeof(DevParams), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(eng->d_vol, eng->vol_h.data(), (size_t)nc * 8, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(eng, HS_E_HIP, "hipMemcpy of parameters failed"));
    if (eng->d_zv_tab) launch_zv_table(eng->dp, eng->d_zv_tab, ns, eng->stream);
    DevScalars s0;
    memset(&s0, 0, sizeof(s0));
    if (hipMemcpy(eng->d_S, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(eng, HS_E_HIP, "hipMemcpy of scalars failed"));
    launch_fill_i32(eng->d_count, 0, nc, eng->stream);
    launch_fill_i32(eng->d_tile_heavy, 0, eng->cap / 64 + 4, eng->stream);
    launch_fill_f64(eng->d_sigma, 0.0, nc, eng->stream);
    launch_fill_f64(eng->d_rem, 0.0, nc, eng->stream);
    launch_fill_f64(eng->d_samples, 0.0, (int64_t)eng->nslots * nc, eng->stream);
    int ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, eng->device) == hipSuccess) ncu = prop.multiProcessorCount;
    eng->ncu = ncu;
    eng->move_grid = ncu * 8;
    for (int tr = 0; tr < 2; ++tr) eng->flight_grid[tr] = ncu * move_blocks_per_cu(tr);
    if (check_launch(eng, "init fills") || hipStreamSynchronize(eng->stream) != hipSuccess)
        return bail(fail(eng, HS_E_HIP, "initialisation failed"));
    *out = eng;
    return 0;
}

// ---------------------------------------------------------------------------
// state transfer
// ---------------------------------------------------------------------------
static int read_scalars(hs_dsmc_engine* eng, DevScalars* s) {
    HS_TRY(hipMemcpyAsync(s, eng->d_S, sizeof(DevScalars), hipMemcpyDeviceToHost, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    eng->n_hint = s->n;
    return 0;
}

extern "C" int hs_dsmc_upload_particles(hs_dsmc_engine* eng, const hs_particles* p, int64_t n) {
    if (!eng || !p) return fail(eng, HS_E_INVALID, "NULL argument");
    if (n < 0 || n > eng->cap) return fail(eng, HS_E_CAPACITY, "particle count exceeds capacity");
    if (!p->px || !p->py || !p->pz || !p->ux || !p->uy || !p->uz || !p->cell)
        return fail(eng, HS_E_INVALID, "position, velocity and cell arrays are required");
    if ((eng->has_erot && !p->erot) || (eng->has_vib && !p->vib) || (eng->has_type && !p->type))
        return fail(eng, HS_E_INVALID, "erot/vib/type arrays required by the species table are missing");
    for (int64_t i = 0; i < n; ++i) {
        const int32_t c = p->cell[i];
        if (c < -1 || c >= eng->nc) return fail(eng, HS_E_INVALID, "cell index out of range");
        if (eng->has_type && p->type[i] >= eng->ns) return fail(eng, HS_E_INVALID, "typeId out of range");
    }
    HS_TRY(hipSetDevice(eng->device));
    DevParticles& b = cur_buf(eng);
    hipStream_t st = eng->stream;
    if (n > 0) {
        HS_TRY(hipMemcpyAsync(b.px, p->px, n * 8, hipMemcpyHostToDevice, st));
        HS_TRY(hipMemcpyAsync(b.py, p->py, n * 8, hipMemcpyHostToDevice, st));
        HS_TRY(hipMemcpyAsync(b.pz, p->pz, n * 8, hipMemcpyHostToDevice, st));
        HS_TRY(hipMemcpyAsync(b.ux, p->ux, n * 4, hipMemcpyHostToDevice, st));
        HS_TRY(hipMemcpyAsync(b.uy, p->uy, n * 4, hipMemcpyHostToDevice, st));
        HS_TRY(hipMemcpyAsync(b.uz, p->uz, n * 4, hipMemcpyHostToDevice, st));
        HS_TRY(hipMemcpyAsync(b.cell, p->cell, n * 4, hipMemcpyHostToDevice, st));
        if (eng->has_erot) HS_TRY(hipMemcpyAsync(b.erot, p->erot, n * 4, hipMemcpyHostToDevice, st));
        if (eng->has_type) HS_TRY(hipMemcpyAsync(b.type, p->type, n, hipMemcpyHostToDevice, st));
        if (eng->has_vib)
            for (int m = 0; m < eng->max_modes; ++m)
                HS_TRY(hipMemcpyAsync(b.vib + m * eng->cap, p->vib + m * n, n * 2, hipMemcpyHostToDevice, st));
    }
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    s.n = (int32_t)n;
    s.n_old = (int32_t)n;
    eng->n_hint = n;
    HS_TRY(hipMemcpyAsync(eng->d_S, &s, sizeof(s), hipMemcpyHostToDevice, st));
    HS_TRY(hipStreamSynchronize(st));
    return 0;
}

extern "C" int hs_dsmc_particle_count(hs_dsmc_engine* eng, int64_t* n_out) {
    if (!eng || !n_out) return fail(eng, HS_E_INVALID, "NULL argument");
    HS_TRY(hipSetDevice(eng->device));
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    *n_out = s.n;
    return 0;
}

extern "C" int hs_dsmc_download_particles(hs_dsmc_engine* eng, hs_particles* p, int64_t cap, int64_t* n_out) {
    if (!eng || !p) return fail(eng, HS_E_INVALID, "NULL argument");
    HS_TRY(hipSetDevice(eng->device));
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    const int64_t n = s.n;
    if (n_out) *n_out = n;
    if (n > cap) return fail(eng, HS_E_CAPACITY, "host buffer too small for the population");
    DevParticles& b = cur_buf(eng);
    hipStream_t st = eng->stream;
    if (n > 0) {
        if (p->px) HS_TRY(hipMemcpyAsync(p->px, b.px, n * 8, hipMemcpyDeviceToHost, st));
        if (p->py) HS_TRY(hipMemcpyAsync(p->py, b.py, n * 8, hipMemcpyDeviceToHost, st));
        if (p->pz) HS_TRY(hipMemcpyAsync(p->pz, b.pz, n * 8, hipMemcpyDeviceToHost, st));
        if (p->ux) HS_TRY(hipMemcpyAsync(p->ux, b.ux, n * 4, hipMemcpyDeviceToHost, st));
        if (p->uy) HS_TRY(hipMemcpyAsync(p->uy, b.uy, n * 4, hipMemcpyDeviceToHost, st));
        if (p->uz) HS_TRY(hipMemcpyAsync(p->uz, b.uz, n * 4, hipMemcpyDeviceToHost, st));
        if (p->cell) HS_TRY(hipMemcpyAsync(p->cell, b.cell, n * 4, hipMemcpyDeviceToHost, st));
        if (p->erot && eng->has_erot) HS_TRY(hipMemcpyAsync(p->erot, b.erot, n * 4, hipMemcpyDeviceToHost, st));
        if (p->type && eng->has_type) HS_TRY(hipMemcpyAsync(p->type, b.type, n, hipMemcpyDeviceToHost, st));
        if (p->vib && eng->has_vib)
            for (int m = 0; m < eng->max_modes; ++m)
                HS_TRY(hipMemcpyAsync(p->vib + m * n, b.vib + m * eng->cap, n * 2, hipMemcpyDeviceToHost, st));
    }
    HS_TRY(hipStreamSynchronize(st));
    return 0;
}

extern "C" int hs_dsmc_set_cell_state(hs_dsmc_engine* eng, const double* sigma, const double* rem) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    if (sigma) HS_TRY(hipMemcpyAsync(eng->d_sigma, sigma, (size_t)eng->nc * 8, hipMemcpyHostToDevice, eng->stream));
    if (rem) HS_TRY(hipMemcpyAsync(eng->d_rem, rem, (size_t)eng->nc * 8, hipMemcpyHostToDevice, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_set_cell_temperature(hs_dsmc_engine* eng, const double* t_overall) {
    if (!eng || !t_overall) return fail(eng, HS_E_INVALID, "NULL argument");
    if (!eng->d_t_overall)
        return fail(eng, HS_E_INVALID,
                    "cell temperatures are read only by inverseZvFormulation 2008 with a Larsen-Borgnakke model");
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipMemcpyAsync(eng->d_t_overall, t_overall, (size_t)eng->nc * 8, hipMemcpyHostToDevice, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_get_cell_state(hs_dsmc_engine* eng, double* sigma, double* rem) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    if (sigma) HS_TRY(hipMemcpyAsync(sigma, eng->d_sigma, (size_t)eng->nc * 8, hipMemcpyDeviceToHost, eng->stream));
    if (rem) HS_TRY(hipMemcpyAsync(rem, eng->d_rem, (size_t)eng->nc * 8, hipMemcpyDeviceToHost, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_inflow_face_count(hs_dsmc_engine* eng, int64_t* n) {
    if (!eng || !n) return fail(eng, HS_E_INVALID, "NULL argument");
    *n = eng->faces.n;
    return 0;
}

extern "C" int hs_dsmc_set_inflow_accumulators(hs_dsmc_engine* eng, const double* acc) {
    if (!eng || !acc) return fail(eng, HS_E_INVALID, "NULL argument");
    if (eng->faces.n == 0) return 0;
    for (int64_t k = 0; k < (int64_t)eng->ns * eng->faces.n; ++k)
        if (!(acc[k] >= -1.0 && acc[k] < 1.0))
            return fail(eng, HS_E_INVALID, "inflow accumulators must lie in [-1, 1)");
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipMemcpyAsync(eng->faces.acc, acc, (size_t)eng->ns * eng->faces.n * 8, hipMemcpyHostToDevice, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_get_inflow_accumulators(hs_dsmc_engine* eng, double* acc) {
    if (!eng || !acc) return fail(eng, HS_E_INVALID, "NULL argument");
    if (eng->faces.n == 0) return 0;
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipMemcpyAsync(acc, eng->faces.acc, (size_t)eng->ns * eng->faces.n * 8, hipMemcpyDeviceToHost, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_set_step(hs_dsmc_engine* eng, uint32_t step) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    s.step = step;
    HS_TRY(hipMemcpyAsync(eng->d_S, &s, sizeof(s), hipMemcpyHostToDevice, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_get_step(hs_dsmc_engine* eng, uint32_t* step) {
    if (!eng || !step) return fail(eng, HS_E_INVALID, "NULL argument");
    HS_TRY(hipSetDevice(eng->device));
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    *step = s.step;
    return 0;
}

// ---------------------------------------------------------------------------
// initial fill
// ---------------------------------------------------------------------------
// The fill's device scratch: per-(cell, species) counts and append offsets,
// the number densities and the drift velocity.
struct FillScratch {
    int32_t* cnt;
    int32_t* off;
    double* dn;
    double* du;
};

static int fill_checks(hs_dsmc_engine* eng, const double* n, double t, const double u[3]) {
    if (!eng || !n || !u) return fail(eng, HS_E_INVALID, "NULL argument");
    if (!(t >= 0.0)) return fail(eng, HS_E_INVALID, "temperature must be non-negative");
    // species-independent expected count guards the capacity
    double expect = 0.0;
    for (int s = 0; s < eng->ns; ++s) {
        if (n[s] < 0.0) return fail(eng, HS_E_INVALID, "number density must be non-negative");
        for (int c = 0; c < eng->nc; ++c) expect += n[s] * eng->vol_h[c] / eng->cfg.fn;
    }
    if (expect > (double)eng->cap) return fail(eng, HS_E_CAPACITY, "fill would exceed the particle capacity");
    return 0;
}

// Every scratch entry the kernels read is written first in this call: dn and
// du by the copies, cnt by k_fill_count (all nc*ns entries), off by the scan.
// fc and sig_cell (device arrays, or NULL) give per-cell inputs.
static int fill_run(hs_dsmc_engine* eng, const double* n, double t, double trot, double tvib, const double u[3],
                    int32_t exact, uint32_t fill_id, const FillScratch& w, FillCells fc = { nullptr, nullptr, nullptr },
                    const double* sig_cell = nullptr) {
    const int nc = eng->nc, ns = eng->ns;
    hipStream_t st = eng->stream;
    double nh[HS_MAX_SPECIES] = { 0 };
    for (int s = 0; s < ns; ++s) nh[s] = n[s];
    if (hipMemcpyAsync(w.dn, nh, sizeof(nh), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(w.du, u, 24, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(eng, HS_E_HIP, "hipMemcpy in fill failed");
    launch_fill_count(eng->dp, w.dn, exact, fill_id, eng->d_vol, w.cnt, nc, ns, fc, st);
    launch_scan(w.cnt, w.off, nc * ns, SCAN_APPEND, eng->ws, eng->d_S, eng->cap, st);
    launch_fill_gen(eng->dp, w.dn, t, trot, tvib, w.du, fill_id, w.cnt, w.off, nc, ns, cur_buf(eng), eng->d_S, fc,
                    st);
    launch_set_nold(eng->d_S, st);
    int most = 0;
    for (int s = 1; s < ns; ++s)
        if (n[s] > n[most]) most = s;
    const hs_species& sm = eng->cfg.species[most];
    const double smax = (3.14159265358979323846 * sm.diameter * sm.diameter) * sqrt(2.0 * HS_KB * t / sm.mass);
    launch_fill_cells(eng->dp, smax, fill_id, eng->d_sigma, eng->d_rem, nc, sig_cell, st);
    int rc = check_launch(eng, "fill");
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(eng, HS_E_HIP, "fill failed");
    DevScalars sc;
    if (!rc) rc = read_scalars(eng, &sc);
    return rc;
}

extern "C" int hs_dsmc_fill(hs_dsmc_engine* eng, const double* n, double t, double trot, double tvib,
                            const double u[3], int32_t exact, uint32_t fill_id) {
    int rc = fill_checks(eng, n, t, u);
    if (rc) return rc;
    HS_TRY(hipSetDevice(eng->device));
    const size_t m = (size_t)eng->nc * eng->ns;
    FillScratch w = { nullptr, nullptr, nullptr, nullptr };
    if (hipMalloc(&w.cnt, m * 4) != hipSuccess || hipMalloc(&w.off, (m + 1) * 4) != hipSuccess ||
        hipMalloc(&w.dn, HS_MAX_SPECIES * 8) != hipSuccess || hipMalloc(&w.du, 3 * 8) != hipSuccess)
        rc = fail(eng, HS_E_HIP, "hipMalloc in fill failed");
    else
        rc = fill_run(eng, n, t, trot, tvib, u, exact, fill_id, w);
    (void)hipFree(w.cnt);
    (void)hipFree(w.off);
    (void)hipFree(w.dn);
    (void)hipFree(w.du);
    return rc;
}

// The fill with each cell's own state: used to start a population from
// fields (the warm start of a case from predicted long-average fields).
extern "C" int hs_dsmc_fill_cells(hs_dsmc_engine* eng, const double* n_cell, const double* u_cell,
                                  const double* t_cell, uint32_t fill_id) {
    if (!eng || !n_cell || !u_cell || !t_cell) return fail(eng, HS_E_INVALID, "NULL argument");
    const int nc = eng->nc, ns = eng->ns;
    double expect = 0.0;
    std::vector<double> sig((size_t)nc);
    for (int c = 0; c < nc; ++c) {
        if (!(t_cell[c] >= 0.0) || !std::isfinite(t_cell[c]))
            return fail(eng, HS_E_INVALID, "fill_cells: temperatures must be finite and non-negative");
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(u_cell[(size_t)a * nc + c]))
                return fail(eng, HS_E_INVALID, "fill_cells: velocities must be finite");
        int most = 0;
        for (int s = 0; s < ns; ++s) {
            const double v = n_cell[(size_t)s * nc + c];
            if (!(v >= 0.0) || !std::isfinite(v))
                return fail(eng, HS_E_INVALID, "fill_cells: number densities must be finite and non-negative");
            expect += v * eng->vol_h[c] / eng->cfg.fn;
            if (v > n_cell[(size_t)most * nc + c]) most = s;
        }
        const hs_species& sm = eng->cfg.species[most];
        sig[c] = (3.14159265358979323846 * sm.diameter * sm.diameter) * sqrt(2.0 * HS_KB * t_cell[c] / sm.mass);
    }
    if (expect > (double)eng->cap) return fail(eng, HS_E_CAPACITY, "fill would exceed the particle capacity");
    HS_TRY(hipSetDevice(eng->device));
    const size_t m = (size_t)nc * ns;
    FillScratch w = { nullptr, nullptr, nullptr, nullptr };
    double *dn = nullptr, *du = nullptr, *dt = nullptr, *ds = nullptr;
    int rc = 0;
    const double n0[HS_MAX_SPECIES] = { 0 };
    const double u0[3] = { 0.0, 0.0, 0.0 };
    hipStream_t st = eng->stream;
    if (hipMalloc(&w.cnt, m * 4) != hipSuccess || hipMalloc(&w.off, (m + 1) * 4) != hipSuccess ||
        hipMalloc(&w.dn, HS_MAX_SPECIES * 8) != hipSuccess || hipMalloc(&w.du, 3 * 8) != hipSuccess ||
        hipMalloc(&dn, m * 8) != hipSuccess || hipMalloc(&du, (size_t)3 * nc * 8) != hipSuccess ||
        hipMalloc(&dt, (size_t)nc * 8) != hipSuccess || hipMalloc(&ds, (size_t)nc * 8) != hipSuccess) {
        rc = fail(eng, HS_E_HIP, "hipMalloc in fill_cells failed");
    } else if (hipMemcpyAsync(dn, n_cell, m * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(du, u_cell, (size_t)3 * nc * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(dt, t_cell, (size_t)nc * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(ds, sig.data(), (size_t)nc * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        rc = fail(eng, HS_E_HIP, "hipMemcpy in fill_cells failed");
    } else {
        rc = fill_run(eng, n0, 0.0, 0.0, 0.0, u0, 0, fill_id, w, FillCells{ dn, du, dt }, ds);
    }
    for (void* p : { (void*)w.cnt, (void*)w.off, (void*)w.dn, (void*)w.du, (void*)dn, (void*)du, (void*)dt, (void*)ds })
        if (p) (void)hipFree(p);
    return rc;
}

// ---------------------------------------------------------------------------
// the fill without allocation: its scratch in a fixed device array
// ---------------------------------------------------------------------------
// Layout: dn [0, 64), du [64, 88), cnt from 256, off after cnt rounded up to
// 256 bytes.  HS_FILL_NOALLOC_MAX_PAIRS cell-species pairs fit.
constexpr size_t kScratchBytes = 256 + 4 * (size_t)HS_FILL_NOALLOC_MAX_PAIRS + 256 + 4 * ((size_t)HS_FILL_NOALLOC_MAX_PAIRS + 1);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = { nullptr };
}  // namespace

// The device address of this device's copy of `scratch`, resolved once per
// device with that device current (the array's name is not a device address
// in host code).  0 on success.
static int scratch_on_device(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int hs_dsmc_fill_noalloc(hs_dsmc_engine* eng, const double* n, double t, double trot, double tvib,
                                    const double u[3], int32_t exact, uint32_t fill_id) {
    int rc = fill_checks(eng, n, t, u);
    if (rc) return rc;
    const int64_t m = (int64_t)eng->nc * eng->ns;
    if (m > HS_FILL_NOALLOC_MAX_PAIRS)
        return fail(eng, HS_E_INVALID, "hs_dsmc_fill_noalloc: " + std::to_string(eng->nc) + " cells x " +
                    std::to_string(eng->ns) + " species exceed its fixed scratch of " +
                    std::to_string(HS_FILL_NOALLOC_MAX_PAIRS) + " cell-species pairs (use hs_dsmc_fill)");
    HS_TRY(hipSetDevice(eng->device));
    unsigned char* base = nullptr;
    if (scratch_on_device(eng->device, (void**)&base))
        return fail(eng, HS_E_HIP, "hipGetSymbolAddress of the fill scratch failed");
    const size_t cnt_bytes = ((4 * (size_t)m) + 255) / 256 * 256;
    FillScratch w;
    w.dn = (double*)base;
    w.du = (double*)(base + 64);
    w.cnt = (int32_t*)(base + 256);
    w.off = (int32_t*)(base + 256 + cnt_bytes);
    return fill_run(eng, n, t, trot, tvib, u, exact, fill_id, w);
}

// ---------------------------------------------------------------------------
// the time step
// ---------------------------------------------------------------------------
// ahead = 0: this step's inflow, appended to the current population.
// ahead = 1: the next step's inflow on the side stream, appended after this
// step's sorted population in the buffer that becomes current at the swap
// (the sort and collide write only below n_sorted); commit installs it.
static void enqueue_inject(hs_dsmc_engine* eng, int ahead) {
    if (eng->faces.n == 0) return;
    hipStream_t st = ahead ? eng->side : eng->stream;
    StageTimer tm(eng, T_INJECT, st);
    launch_inject_count(eng->dp, eng->faces, eng->ns, eng->d_S, ahead, st);
    launch_scan(eng->faces.n_new, eng->faces.off, eng->faces.n * eng->ns, ahead ? SCAN_APPEND_NEXT : SCAN_APPEND,
                eng->ws, eng->d_S, eng->cap, st);
    launch_inject_gen(eng->dp, eng->faces, eng->ns, ahead ? nxt_buf(eng) : cur_buf(eng), eng->d_S, ahead, st);
}

static void enqueue_move(hs_dsmc_engine* eng, int do_move, int do_hist) {
    StageTimer tm(eng, T_MOVE);
    const int track = (eng->cfg.flags & HS_FLAG_FACE_FLUX) && (eng->cfg.flags & HS_FLAG_SAMPLE);
    launch_move(eng->dp, cur_buf(eng), eng->d_S, eng->d_count, eng->d_slot, eng->wq, do_move, do_hist, track,
                do_move ? eng->flight_grid[track] : eng->move_grid, eng->stream);
}

// inject_next: the next step's inflow runs on the side stream, concurrently
// with this step's scatter, gather and collide (it needs only n_sorted).
static void enqueue_sort_collide(hs_dsmc_engine* eng, int do_collide, int do_sample, int advance, int inject_next) {
    {
        StageTimer tm(eng, T_SCAN);
        launch_scan_cells(eng->d_count, eng->d_start, eng->nc, eng->d_lb_status, eng->d_lb_ticket, eng->d_S,
                          eng->tile_parcels, eng->d_tile_begin, eng->collide_parcels, eng->d_ctile_begin,
                          eng->d_tile_heavy, eng->stream);
    }
    inject_next = inject_next && advance && eng->faces.n > 0;
    // with chemistry the products go after the sorted population, so the next
    // step's inflow is appended after them, at the end of this step
    const bool react = !eng->reactions.empty();
    const bool inject_after = react && inject_next;
    if (react) inject_next = 0;
    if (inject_next) {
        (void)hipEventRecord(eng->ev_scanned, eng->stream);
        (void)hipStreamWaitEvent(eng->side, eng->ev_scanned, 0);
        enqueue_inject(eng, 1);
        (void)hipEventRecord(eng->ev_injected, eng->side);
    }
    {
        StageTimer tm(eng, T_SCATTER);
        launch_scatter(cur_buf(eng), eng->d_S, eng->d_start, eng->d_slot, eng->d_perm, eng->d_tile_heavy,
                       eng->collide_parcels, eng->d_tile_order, eng->move_grid, eng->stream);
    }
    {
        StageTimer tm(eng, T_GATHER);
        launch_gather(eng->dp, cur_buf(eng), nxt_buf(eng), eng->d_start, eng->d_perm, eng->d_tile_begin,
                      eng->tile_parcels, eng->cap, eng->d_S, eng->d_big_sub, eng->general, eng->d_slot, eng->stream);
    }
    if ((do_collide || do_sample) && react) {
        StageTimer tm(eng, T_COLLIDE);
        launch_collide_react(eng->dp, nxt_buf(eng), eng->d_start, eng->d_ctile_begin, eng->d_tile_order,
                             eng->collide_parcels, eng->d_big_sub, eng->d_sigma, eng->d_rem, eng->d_samples,
                             eng->d_vol, eng->d_S, eng->prod, do_collide, do_sample, eng->nc, eng->ns, eng->d_big_list,
                             eng->ws, eng->cap, eng->stream);
    } else if (do_collide || do_sample) {
        StageTimer tm(eng, T_COLLIDE);
        launch_collide(eng->dp, nxt_buf(eng), eng->d_start, eng->d_ctile_begin, eng->d_tile_heavy, eng->d_tile_order,
                       eng->collide_parcels, eng->cap, eng->d_big_sub, eng->d_sigma, eng->d_rem, eng->d_samples,
                       eng->d_vol, eng->d_S, do_collide, do_sample, eng->general, eng->nc, eng->d_big_list,
                       eng->stream);
    }
    if (inject_next) (void)hipStreamWaitEvent(eng->stream, eng->ev_injected, 0);
    launch_commit(eng->d_S, advance, inject_next, eng->stream);
    eng->cur = 1 - eng->cur;
    if (inject_after) enqueue_inject(eng, 0);
    if (do_sample) eng->n_sampled += 1;
}

// Gather and collide tiles: whole cells grouped up to pt parcels per
// workgroup.  A full device takes the largest tiles; a small population is
// cut into smaller ones (halving, down to 256 parcels) until it makes about
// four workgroups per compute unit.  The tiling changes no result: the random
// streams belong to cells and candidates, and each cell is sampled whole.
// Tile sizes the autotuner measured, per kernel variant (internal energies or
// not) and population class (log2 of the parcel count): later engines of the
// same class on any device of the process start from them.
namespace {
std::mutex g_tune_mu;
struct TuneKey {
    int general, pop_log2;
    bool operator<(const TuneKey& o) const {
        return general != o.general ? general < o.general : pop_log2 < o.pop_log2;
    }
};
std::map<TuneKey, std::pair<int, int>> g_tuned;
int pop_class(int64_t n) {
    int k = 0;
    while (n > 1 && k < 62) { n >>= 1; ++k; }
    return k;
}
}  // namespace

static bool tuned_tiles(const hs_dsmc_engine* eng, int* g, int* c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(TuneKey{ eng->general, pop_class(eng->n_hint) });
    if (it == g_tuned.end()) return false;
    *g = it->second.first;
    *c = it->second.second;
    return true;
}

static int tile_for(const hs_dsmc_engine* eng, int which) {
    if (eng->tile_req[which] > 0) return eng->tile_req[which];
    int tg = 0, tc = 0;
    if (tuned_tiles(eng, &tg, &tc)) return which == 0 ? tg : tc;
    const int64_t per = eng->n_hint / (4 * (int64_t)eng->ncu);
    int pt = eng->tile_max[which];
    while (pt / 2 >= 256 && pt > per) pt /= 2;
    return pt;
}

static int choose_tiles(hs_dsmc_engine* eng) {
    const int g = tile_for(eng, 0), c = tile_for(eng, 1);
    if (g == eng->tile_parcels && c == eng->collide_parcels) return 0;
    HS_TRY(hipStreamSynchronize(eng->stream));
    drop_graphs(eng);  // captured with the old tile sizes
    eng->tile_parcels = g;
    eng->collide_parcels = c;
    return 0;
}

// Captures the steady-state step pair starting on the current buffer.  The
// capture runs nothing; the host-side step bookkeeping it did is undone.
static int capture_pair(hs_dsmc_engine* eng, int collide, int sample) {
    int64_t launches[T_NSTAGE];
    std::memcpy(launches, eng->launches, sizeof(launches));
    const int64_t n_sampled = eng->n_sampled;
    const int cur = eng->cur;
    HS_TRY(hipStreamBeginCapture(eng->stream, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < 2; ++j) {
        enqueue_move(eng, 1, 1);
        enqueue_sort_collide(eng, collide, sample, 1, 1);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(eng->stream, &g);
    for (int s = 0; s < T_NSTAGE; ++s) eng->pair_launches[s] = eng->launches[s] - launches[s];
    std::memcpy(eng->launches, launches, sizeof(launches));
    eng->n_sampled = n_sampled;
    eng->cur = cur;
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(eng, HS_E_HIP, std::string("capturing the step pair: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&eng->pair_graph[cur], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        eng->pair_graph[cur] = nullptr;
        return fail(eng, HS_E_HIP, std::string("instantiating the step pair: ") + hipGetErrorString(ei));
    }
    return 0;
}

// One engine's way through an evolve of nsteps: evolve_begin enqueues the
// first step's inflow (later steps' inflow is injected ahead, in the step
// before); each evolve_advance enqueues the next graph-replayed pair, or a
// single eager step at the end.
struct EvolveRun {
    hs_dsmc_engine* eng;
    int nsteps, k, collide, sample;
    bool graphs;
};

static int evolve_begin(EvolveRun& r, hs_dsmc_engine* eng, int nsteps) {
    r.eng = eng;
    r.nsteps = nsteps;
    r.k = 0;
    HS_TRY(hipSetDevice(eng->device));
    if (int rc = choose_tiles(eng)) return rc;
    r.sample = (eng->cfg.flags & HS_FLAG_SAMPLE) ? 1 : 0;
    r.collide = eng->cfg.collision_model != HS_COLL_NONE;
    // per-stage timing records events between the launches: eager steps only
    r.graphs = eng->graphs && !(eng->cfg.flags & HS_FLAG_TIMING);
    if (nsteps > 0) enqueue_inject(eng, 0);
    return 0;
}

static int evolve_advance(EvolveRun& r) {
    hs_dsmc_engine* eng = r.eng;
    HS_TRY(hipSetDevice(eng->device));
    // steps k and k+1 both inject ahead; the last step does not
    if (r.graphs && r.k + 2 < r.nsteps) {
        if (!eng->pair_graph[eng->cur])
            if (int rc = capture_pair(eng, r.collide, r.sample)) return rc;
        HS_TRY(hipGraphLaunch(eng->pair_graph[eng->cur], eng->stream));
        for (int s = 0; s < T_NSTAGE; ++s) eng->launches[s] += eng->pair_launches[s];
        eng->n_sampled += 2 * r.sample;
        r.k += 2;
        return 0;
    }
    enqueue_move(eng, 1, 1);
    enqueue_sort_collide(eng, r.collide, r.sample, 1, r.k + 1 < r.nsteps);
    r.k += 1;
    return 0;
}

extern "C" int hs_dsmc_evolve(hs_dsmc_engine* eng, int32_t nsteps) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    if (nsteps < 0) return fail(eng, HS_E_INVALID, "nsteps must be non-negative");
    EvolveRun r;
    if (int rc = evolve_begin(r, eng, nsteps)) return rc;
    while (r.k < nsteps)
        if (int rc = evolve_advance(r)) return rc;
    return check_launch(eng, "evolve");
}

extern "C" int hs_dsmc_evolve_batch(hs_dsmc_engine* const* engines, int32_t count, int32_t nsteps) {
    if (!engines || count < 0) return fail(nullptr, HS_E_INVALID, "engines is NULL or count negative");
    for (int32_t e = 0; e < count; ++e) {
        if (!engines[e]) return fail(nullptr, HS_E_INVALID, "NULL engine in the batch");
        for (int32_t f = 0; f < e; ++f)
            if (engines[f] == engines[e]) return fail(engines[e], HS_E_INVALID, "an engine appears twice in the batch");
    }
    if (nsteps < 0) return fail(count ? engines[0] : nullptr, HS_E_INVALID, "nsteps must be non-negative");
    std::vector<EvolveRun> runs((size_t)count);
    for (int32_t e = 0; e < count; ++e)
        if (int rc = evolve_begin(runs[e], engines[e], nsteps)) return rc;
    // step-major: every engine's next pair is enqueued before any engine's
    // following one, so the engines' streams keep the device busy together
    for (int done = 0; done < count;) {
        done = 0;
        for (auto& r : runs) {
            if (r.k >= nsteps) { ++done; continue; }
            if (int rc = evolve_advance(r)) return rc;
        }
    }
    for (int32_t e = 0; e < count; ++e)
        if (int rc = check_launch(engines[e], "evolve_batch")) return rc;
    return 0;
}

// ---------------------------------------------------------------------------
// batches: one launch per stage over many engines
// ---------------------------------------------------------------------------
// The engines of a batch are grouped by kernel variant (internal energies or
// not, face tracker on or not); a step launches each stage once per group,
// over the workgroups of all the group's cases, on the group's stream.  Each
// workgroup finds its case in a table; the kernels run the single-engine
// bodies on the case's own buffers, so a batched step equals the engines' own
// steps.  The graph of a two-step pair is captured per buffer parity (the
// cases' buffer indices relative to the batch's base, the same for all).
namespace {
enum { BT_INJ_COUNT = 0, BT_INJ_SCAN, BT_INJ_GEN, BT_MOVE, BT_WALL, BT_SCAN, BT_SCATTER, BT_GATHER, BT_COLLIDE, BT_N };

struct BatchGroup {
    int general = 0, track = 0;
    std::vector<int> members;          // engine indices, largest capacity first
    std::vector<BatchCase> h_cases;
    std::vector<BatchBlock> h_tab;
    int off[BT_N] = {}, len[BT_N] = {};
    BatchCase* d_cases = nullptr;
    BatchBlock* d_tab = nullptr;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t ev_scanned = nullptr, ev_injected = nullptr, ev_join = nullptr;
};

int batch_key(const hs_dsmc_engine* e) {
    const int track = (e->cfg.flags & HS_FLAG_FACE_FLUX) && (e->cfg.flags & HS_FLAG_SAMPLE);
    return e->general * 2 + track;
}
}  // namespace

struct hs_dsmc_batch {
    std::vector<hs_dsmc_engine*> engines;
    std::vector<int> base_cur;       // each engine's current buffer at parity 0
    std::vector<int> keys;           // the variant of each engine when the groups were built
    std::vector<BatchGroup> groups;
    std::vector<hipEvent_t> ev_engine;  // each engine's stream has reached the batch
    hipEvent_t ev_fork = nullptr, ev_done = nullptr;
    hipGraphExec_t pair_graph[2] = { nullptr, nullptr };
    int device = 0;
    bool graphs = true;
    std::string err;
};

static int bfail(hs_dsmc_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_create_error = msg;
    return code;
}

#define BTRY(expr)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return bfail(b, HS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void batch_drop_graphs(hs_dsmc_batch* b) {
    for (auto& g : b->pair_graph)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

static void free_group(BatchGroup& g) {
    if (g.d_cases) (void)hipFree(g.d_cases);
    if (g.d_tab) (void)hipFree(g.d_tab);
    for (hipEvent_t ev : { g.ev_scanned, g.ev_injected, g.ev_join })
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t s : { g.stream, g.side })
        if (s) (void)hipStreamDestroy(s);
    g = BatchGroup();
}

extern "C" void hs_dsmc_batch_destroy(hs_dsmc_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (auto& g : b->groups)
        if (g.stream) (void)hipStreamSynchronize(g.stream);
    batch_drop_graphs(b);
    for (auto& g : b->groups) free_group(g);
    for (hipEvent_t ev : b->ev_engine)
        if (ev) (void)hipEventDestroy(ev);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    delete b;
}

extern "C" const char* hs_dsmc_batch_last_error(const hs_dsmc_batch* b) {
    return b ? b->err.c_str() : g_create_error.c_str();
}

static BatchCase batch_case_of(hs_dsmc_batch* b, int e) {
    hs_dsmc_engine* eng = b->engines[e];
    BatchCase c;
    std::memset(&c, 0, sizeof(c));
    c.P = eng->dp;
    c.buf[0] = eng->buf[0];
    c.buf[1] = eng->buf[1];
    c.cur0 = b->base_cur[e];
    c.nc = eng->nc;
    c.ns = eng->ns;
    c.pt_g = eng->tile_max[0];
    c.pt_c = eng->tile_max[1];
    c.do_collide = eng->cfg.collision_model != HS_COLL_NONE;
    c.do_sample = (eng->cfg.flags & HS_FLAG_SAMPLE) ? 1 : 0;
    c.with_faces = eng->faces.n > 0;
    c.cap = eng->cap;
    c.S = eng->d_S;
    c.count = eng->d_count;
    c.start = eng->d_start;
    c.slot = eng->d_slot;
    c.perm = eng->d_perm;
    c.sub = eng->d_big_sub;
    c.big_list = eng->d_big_list;
    c.tile_g = eng->d_tile_begin;
    c.tile_c = eng->d_ctile_begin;
    c.heavy = eng->d_tile_heavy;
    c.order = eng->d_tile_order;
    c.sigma = eng->d_sigma;
    c.rem = eng->d_rem;
    c.samples = eng->d_samples;
    c.vol = eng->d_vol;
    c.lb_status = eng->d_lb_status;
    c.lb_ticket = eng->d_lb_ticket;
    c.wq = eng->wq;
    c.faces = eng->faces;
    return c;
}

// The workgroups each case takes in each stage: the single-engine launch
// sizes, except the grid-stride move and scatter, which take one workgroup
// per 2048 parcels of capacity up to the single-engine grid.
static void batch_tables(hs_dsmc_batch* b, BatchGroup& g) {
    std::vector<BatchBlock> t[BT_N];
    for (int k = 0; k < (int)g.members.size(); ++k) {
        const hs_dsmc_engine* eng = b->engines[g.members[k]];
        const BatchCase& c = g.h_cases[k];
        auto add = [&](int stage, int64_t n) {
            for (int64_t x = 0; x < n; ++x) t[stage].push_back(BatchBlock{ k, (int32_t)x, (int32_t)n, 0 });
        };
        const int ne = c.faces.n * c.ns;
        if (ne > 0) {
            add(BT_INJ_COUNT, (ne + 255) / 256);
            add(BT_INJ_SCAN, 1);
            if (c.faces.max_new > 0) add(BT_INJ_GEN, (c.faces.max_new + 255) / 256);
        }
        const int64_t per = (eng->cap + 2047) / 2048;
        const int64_t gm = std::max<int64_t>(1, std::min<int64_t>(per, eng->flight_grid[g.track]));
        add(BT_MOVE, gm);
        add(BT_WALL, std::max<int64_t>(1, gm / 8));
        add(BT_SCAN, scan_cells_blocks(eng->nc));
        add(BT_SCATTER, std::max<int64_t>(1, std::min<int64_t>(per, eng->move_grid)));
        add(BT_GATHER, eng->cap / c.pt_g + 2);
        if (c.do_collide || c.do_sample) add(BT_COLLIDE, eng->cap / c.pt_c + 2);
    }
    g.h_tab.clear();
    for (int s = 0; s < BT_N; ++s) {
        g.off[s] = (int)g.h_tab.size();
        g.len[s] = (int)t[s].size();
        g.h_tab.insert(g.h_tab.end(), t[s].begin(), t[s].end());
    }
}

// (Re)builds the groups and their device tables when the engines' variants,
// buffers or settings changed since the last evolve; returns the parity of
// the next step in *par.
static int batch_layout(hs_dsmc_batch* b, int* par) {
    const int n = (int)b->engines.size();
    std::vector<int> keys(n);
    for (int e = 0; e < n; ++e) keys[e] = batch_key(b->engines[e]);
    bool same_par = true;
    const int p0 = b->engines[0]->cur ^ b->base_cur[0];
    for (int e = 1; e < n; ++e)
        if ((b->engines[e]->cur ^ b->base_cur[e]) != p0) same_par = false;
    if (keys != b->keys || !same_par) {
        // new groups (or buffers out of step with each other): rebase at parity 0
        for (auto& g : b->groups) {
            if (g.stream) BTRY(hipStreamSynchronize(g.stream));
            free_group(g);
        }
        b->groups.clear();
        batch_drop_graphs(b);
        for (int e = 0; e < n; ++e) b->base_cur[e] = b->engines[e]->cur;
        b->keys = keys;
        std::map<int, int> gi;
        for (int e = 0; e < n; ++e) {
            if (!gi.count(keys[e])) {
                gi[keys[e]] = (int)b->groups.size();
                b->groups.emplace_back();
                BatchGroup& g = b->groups.back();
                g.general = keys[e] / 2;
                g.track = keys[e] % 2;
                BTRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
                BTRY(hipStreamCreateWithFlags(&g.side, hipStreamNonBlocking));
                BTRY(hipEventCreateWithFlags(&g.ev_scanned, hipEventDisableTiming));
                BTRY(hipEventCreateWithFlags(&g.ev_injected, hipEventDisableTiming));
                BTRY(hipEventCreateWithFlags(&g.ev_join, hipEventDisableTiming));
            }
            b->groups[gi[keys[e]]].members.push_back(e);
        }
        for (auto& g : b->groups)
            std::stable_sort(g.members.begin(), g.members.end(),
                             [&](int x, int y) { return b->engines[x]->cap > b->engines[y]->cap; });
    }
    *par = b->engines[0]->cur ^ b->base_cur[0];
    for (auto& g : b->groups) {
        std::vector<BatchCase> cs;
        for (int e : g.members) cs.push_back(batch_case_of(b, e));
        const bool changed = cs.size() != g.h_cases.size() || !g.d_cases ||
                             std::memcmp(cs.data(), g.h_cases.data(), cs.size() * sizeof(BatchCase)) != 0;
        if (!changed) continue;
        // the pair graphs replay on the first group's stream and read every
        // group's tables: all streams drain before any table is replaced
        for (auto& h : b->groups)
            if (h.stream) BTRY(hipStreamSynchronize(h.stream));
        batch_drop_graphs(b);
        g.h_cases = cs;
        batch_tables(b, g);
        if (g.d_cases) { BTRY(hipFree(g.d_cases)); g.d_cases = nullptr; }
        if (g.d_tab) { BTRY(hipFree(g.d_tab)); g.d_tab = nullptr; }
        BTRY(hipMalloc(&g.d_cases, cs.size() * sizeof(BatchCase)));
        BTRY(hipMalloc(&g.d_tab, std::max<size_t>(1, g.h_tab.size()) * sizeof(BatchBlock)));
        BTRY(hipMemcpy(g.d_cases, cs.data(), cs.size() * sizeof(BatchCase), hipMemcpyHostToDevice));
        if (!g.h_tab.empty())
            BTRY(hipMemcpy(g.d_tab, g.h_tab.data(), g.h_tab.size() * sizeof(BatchBlock), hipMemcpyHostToDevice));
    }
    return 0;
}

static const BatchBlock* btab(const BatchGroup& g, int s) { return g.d_tab + g.off[s]; }

static void batch_inject(BatchGroup& g, int par, int ahead, hipStream_t st) {
    launch_batch_inject(g.d_cases, btab(g, BT_INJ_COUNT), g.len[BT_INJ_COUNT], btab(g, BT_INJ_SCAN),
                        g.len[BT_INJ_SCAN], btab(g, BT_INJ_GEN), g.len[BT_INJ_GEN], par, ahead, st);
}

// One step of a group at parity par on its stream (the next step's inflow on
// its side stream, between the cell scan and the commit).
static void batch_step(BatchGroup& g, int par, int inject_next) {
    hipStream_t st = g.stream;
    launch_batch_move(g.d_cases, btab(g, BT_MOVE), g.len[BT_MOVE], btab(g, BT_WALL), g.len[BT_WALL], par, g.track,
                      st);
    launch_batch_scan_cells(g.d_cases, btab(g, BT_SCAN), g.len[BT_SCAN], st);
    inject_next = inject_next && g.len[BT_INJ_SCAN] > 0;
    if (inject_next) {
        (void)hipEventRecord(g.ev_scanned, st);
        (void)hipStreamWaitEvent(g.side, g.ev_scanned, 0);
        batch_inject(g, par, 1, g.side);
        (void)hipEventRecord(g.ev_injected, g.side);
    }
    launch_batch_scatter(g.d_cases, btab(g, BT_SCATTER), g.len[BT_SCATTER], par, st);
    launch_batch_gather(g.d_cases, btab(g, BT_GATHER), g.len[BT_GATHER], par, g.general, st);
    launch_batch_collide(g.d_cases, btab(g, BT_COLLIDE), g.len[BT_COLLIDE], par, g.general, st);
    if (inject_next) (void)hipStreamWaitEvent(st, g.ev_injected, 0);
    launch_batch_commit(g.d_cases, (int)g.members.size(), 1, inject_next, st);
}

// the other groups' streams follow the first group's (fork) / it waits for them (join)
static void batch_fork(hs_dsmc_batch* b) {
    (void)hipEventRecord(b->ev_fork, b->groups[0].stream);
    for (size_t k = 1; k < b->groups.size(); ++k) (void)hipStreamWaitEvent(b->groups[k].stream, b->ev_fork, 0);
}
static void batch_join(hs_dsmc_batch* b) {
    for (size_t k = 1; k < b->groups.size(); ++k) {
        (void)hipEventRecord(b->groups[k].ev_join, b->groups[k].stream);
        (void)hipStreamWaitEvent(b->groups[0].stream, b->groups[k].ev_join, 0);
    }
}

static int batch_capture(hs_dsmc_batch* b, int par) {
    hipStream_t s0 = b->groups[0].stream;
    BTRY(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    batch_fork(b);
    for (int j = 0; j < 2; ++j)
        for (auto& g : b->groups) batch_step(g, par ^ j, 1);
    batch_join(b);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s0, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return bfail(b, HS_E_HIP, std::string("capturing the batch step pair: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&b->pair_graph[par], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        b->pair_graph[par] = nullptr;
        return bfail(b, HS_E_HIP, std::string("instantiating the batch step pair: ") + hipGetErrorString(ei));
    }
    return 0;
}

extern "C" int hs_dsmc_batch_create(hs_dsmc_engine* const* engines, int32_t count, hs_dsmc_batch** out) {
    if (!out) return bfail(nullptr, HS_E_INVALID, "NULL output pointer");
    *out = nullptr;
    if (!engines || count < 1) return bfail(nullptr, HS_E_INVALID, "a batch needs at least one engine");
    for (int32_t e = 0; e < count; ++e) {
        if (!engines[e]) return bfail(nullptr, HS_E_INVALID, "NULL engine in the batch");
        if (engines[e]->device != engines[0]->device)
            return bfail(nullptr, HS_E_INVALID, "the engines of a batch must be on one device");
        for (int32_t f = 0; f < e; ++f)
            if (engines[f] == engines[e]) return bfail(nullptr, HS_E_INVALID, "an engine appears twice in the batch");
        if (!engines[e]->reactions.empty())
            return bfail(nullptr, HS_E_UNSUPPORTED, "engines with chemical reactions cannot join a one-launch batch");
    }
    hs_dsmc_batch* b = new hs_dsmc_batch();
    b->engines.assign(engines, engines + count);
    b->base_cur.assign(count, 0);
    b->device = engines[0]->device;
    b->graphs = engines[0]->graphs;
    auto bail = [&](hipError_t e, const char* what) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(e);
        hs_dsmc_batch_destroy(b);
        return HS_E_HIP;
    };
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return bail(e, "hipSetDevice");
    b->ev_engine.assign(count, nullptr);
    for (int32_t k = 0; k < count; ++k)
        if ((e = hipEventCreateWithFlags(&b->ev_engine[k], hipEventDisableTiming)) != hipSuccess)
            return bail(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming)) != hipSuccess)
        return bail(e, "hipEventCreate");
    *out = b;
    return 0;
}

// The batched step is switched off: its first run on the device ended in a
// crash of the host process inside the first evolve (in the HIP runtime, with
// the host code clean under AddressSanitizer against a CPU stand-in of the
// runtime), and the cause is not identified.  The one construct it uses that
// the single-engine path does not is a stream capture that forks from its
// origin to several streams, each forking again to its side stream.  Until
// the cause is shown the call fails; hs_dsmc_evolve_batch steps batches.
static const bool kBatchEvolveEnabled = false;

extern "C" int hs_dsmc_batch_evolve(hs_dsmc_batch* b, int32_t nsteps) {
    if (!b) return bfail(nullptr, HS_E_INVALID, "NULL batch");
    if (nsteps < 0) return bfail(b, HS_E_INVALID, "nsteps must be non-negative");
    if (nsteps == 0) return 0;
    if (!kBatchEvolveEnabled)
        return bfail(b, HS_E_INVALID,
                     "hs_dsmc_batch_evolve is switched off (its first device run crashed the host process; "
                     "cause not identified): use hs_dsmc_evolve_batch");
    BTRY(hipSetDevice(b->device));
    int par = 0;
    if (int rc = batch_layout(b, &par)) return rc;
    // after every engine's earlier work
    for (size_t e = 0; e < b->engines.size(); ++e) {
        BTRY(hipEventRecord(b->ev_engine[e], b->engines[e]->stream));
        BTRY(hipStreamWaitEvent(b->groups[0].stream, b->ev_engine[e], 0));
    }
    batch_fork(b);
    for (auto& g : b->groups) batch_inject(g, par, 0, g.stream);  // the first step's inflow
    batch_join(b);
    int k = 0;
    while (k < nsteps) {
        if (b->graphs && k + 2 < nsteps) {
            if (!b->pair_graph[par])
                if (int rc = batch_capture(b, par)) return rc;
            BTRY(hipGraphLaunch(b->pair_graph[par], b->groups[0].stream));
            k += 2;
            continue;
        }
        batch_fork(b);
        for (auto& g : b->groups) batch_step(g, par, k + 1 < nsteps);
        batch_join(b);
        par ^= 1;
        k += 1;
    }
    BTRY(hipGetLastError());
    // the engines' later work follows the batch
    BTRY(hipEventRecord(b->ev_done, b->groups[0].stream));
    for (hs_dsmc_engine* eng : b->engines) {
        BTRY(hipStreamWaitEvent(eng->stream, b->ev_done, 0));
        eng->cur ^= (nsteps & 1);
        if (eng->cfg.flags & HS_FLAG_SAMPLE) eng->n_sampled += nsteps;
    }
    return 0;
}

extern "C" int hs_dsmc_inject(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    enqueue_inject(eng, 0);
    return check_launch(eng, "inject");
}

extern "C" int hs_dsmc_move(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    enqueue_move(eng, 1, 0);
    launch_clear_queue(eng->d_S, eng->stream);  // no commit follows a lone move
    return check_launch(eng, "move");
}

extern "C" int hs_dsmc_sort(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    if (int rc = choose_tiles(eng)) return rc;
    enqueue_move(eng, 0, 1);
    enqueue_sort_collide(eng, 0, 0, 0, 0);
    return check_launch(eng, "sort");
}

extern "C" int hs_dsmc_collide(hs_dsmc_engine* eng, int32_t sample) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    if (int rc = choose_tiles(eng)) return rc;
    enqueue_move(eng, 0, 1);
    enqueue_sort_collide(eng, eng->cfg.collision_model != HS_COLL_NONE, sample ? 1 : 0, 0, 0);
    return check_launch(eng, "collide");
}

extern "C" int hs_dsmc_advance_step(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    uint32_t s = 0;
    if (int rc = hs_dsmc_get_step(eng, &s)) return rc;
    return hs_dsmc_set_step(eng, s + 1);
}

extern "C" int hs_dsmc_set_flags(hs_dsmc_engine* eng, int32_t flags) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    if ((flags & HS_FLAG_HEATFLUX) && !(eng->cfg.flags & HS_FLAG_HEATFLUX))
        return fail(eng, HS_E_INVALID, "heat-flux sums can only be enabled at create time");
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipStreamSynchronize(eng->stream));
    drop_graphs(eng);  // captured with the old flags' kernels and stages
    eng->cfg.flags = flags;
    eng->hp.flags = flags;
    HS_TRY(hipMemcpyAsync(eng->dp, &eng->hp, sizeof(DevParams), hipMemcpyHostToDevice, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_set_tiles(hs_dsmc_engine* eng, int32_t gather_parcels, int32_t collide_parcels) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    if (gather_parcels < 0 || collide_parcels < 0 || (gather_parcels && gather_parcels < 64) ||
        (collide_parcels && collide_parcels < 64) || gather_parcels > eng->tile_max[0] ||
        collide_parcels > eng->tile_max[1])
        return fail(eng, HS_E_INVALID, "tile sizes must be 0 (automatic) or between 64 and the full-device size");
    HS_TRY(hipSetDevice(eng->device));
    eng->tile_req[0] = gather_parcels;
    eng->tile_req[1] = collide_parcels;
    return 0;
}

// Times steps of the live state under each candidate tiling (gather and
// collide parcels per workgroup, powers of two from 256 and the full-device
// sizes, skipping tilings of fewer workgroups than compute units) and keeps
// the fastest.  The steps taken are ordinary steps of the run: the tiling
// changes no result.
extern "C" int hs_dsmc_autotune(hs_dsmc_engine* eng, int32_t steps_per_trial, int32_t* gather_out,
                                int32_t* collide_out, double* ms_out) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    if (steps_per_trial < 1 || steps_per_trial > 10000)
        return fail(eng, HS_E_INVALID, "autotune: steps_per_trial must be in [1, 10000]");
    HS_TRY(hipSetDevice(eng->device));
    int64_t n = 0;
    if (int rc0 = hs_dsmc_particle_count(eng, &n)) return rc0;
    std::vector<int> gs, cs;
    for (int v : { 256, 512, 1024, 1536 })
        if (v <= eng->tile_max[0] && (v == 256 || n / v >= eng->ncu)) gs.push_back(v);
    for (int v : { 256, 512, 1024, 1536, 2048 })
        if (v <= eng->tile_max[1] && (v == 256 || n / v >= eng->ncu)) cs.push_back(v);
    hipEvent_t a = nullptr, b = nullptr;
    HS_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(eng, HS_E_HIP, "hipEventCreate failed");
    }
    const int saved[2] = { eng->tile_req[0], eng->tile_req[1] };
    double best = 1e300;
    int bg = 0, bc = 0, rc = 0;
    for (int g : gs) {
        for (int c : cs) {
            eng->tile_req[0] = g;
            eng->tile_req[1] = c;
            // two steps capture the graphs of this tiling, then the timed ones
            if ((rc = hs_dsmc_evolve(eng, 2))) break;
            if (hipEventRecord(a, eng->stream) != hipSuccess) { rc = fail(eng, HS_E_HIP, "hipEventRecord"); break; }
            if ((rc = hs_dsmc_evolve(eng, steps_per_trial))) break;
            float ms = 0.0f;
            if (hipEventRecord(b, eng->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                rc = fail(eng, HS_E_HIP, "timing a tiling failed");
                break;
            }
            const double per = ms / steps_per_trial;
            if (per < best) { best = per; bg = g; bc = c; }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    eng->tile_req[0] = saved[0];
    eng->tile_req[1] = saved[1];
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[TuneKey{ eng->general, pop_class(eng->n_hint) }] = { bg, bc };
    }
    if (gather_out) *gather_out = bg;
    if (collide_out) *collide_out = bc;
    if (ms_out) *ms_out = best;
    return 0;
}

extern "C" int hs_dsmc_get_tiles(hs_dsmc_engine* eng, int32_t* gather_parcels, int32_t* collide_parcels) {
    if (!eng || !gather_parcels || !collide_parcels) return fail(eng, HS_E_INVALID, "NULL argument");
    *gather_parcels = tile_for(eng, 0);
    *collide_parcels = tile_for(eng, 1);
    return 0;
}

extern "C" int hs_dsmc_synchronize(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// sampling, counters, timing
// ---------------------------------------------------------------------------
extern "C" int hs_dsmc_read_samples(hs_dsmc_engine* eng, double* out, int64_t n_doubles, int64_t* nsteps_out) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    const int64_t need = (int64_t)eng->nslots * eng->nc;
    if (out && n_doubles < need) return fail(eng, HS_E_INVALID, "sample buffer too small");
    HS_TRY(hipSetDevice(eng->device));
    if (out) HS_TRY(hipMemcpyAsync(out, eng->d_samples, need * 8, hipMemcpyDeviceToHost, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    if (nsteps_out) *nsteps_out = eng->n_sampled;
    return 0;
}

extern "C" int hs_dsmc_reset_samples(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    launch_fill_f64(eng->d_samples, 0.0, (int64_t)eng->nslots * eng->nc, eng->stream);
    if (eng->hp.n_wall > 0)
        launch_fill_f64(eng->d_wall_sum, 0.0, (int64_t)eng->ns * HS_WS_COUNT * eng->hp.n_wall, eng->stream);
    launch_fill_i64((int64_t*)eng->d_face_flux, 0, (int64_t)eng->ns * eng->hp.n_faces, eng->stream);
    eng->n_sampled = 0;
    return check_launch(eng, "reset samples");
}

extern "C" int hs_dsmc_face_count(hs_dsmc_engine* eng, int64_t* n_faces) {
    if (!eng || !n_faces) return fail(eng, HS_E_INVALID, "NULL argument");
    *n_faces = eng->hp.n_faces;
    return 0;
}

extern "C" int hs_dsmc_read_face_flux(hs_dsmc_engine* eng, int64_t* out, int64_t n, int64_t* nsteps_out) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    const int64_t need = (int64_t)eng->ns * eng->hp.n_faces;
    if (out && n < need) return fail(eng, HS_E_INVALID, "face flux buffer too small");
    HS_TRY(hipSetDevice(eng->device));
    if (out) HS_TRY(hipMemcpyAsync(out, eng->d_face_flux, need * 8, hipMemcpyDeviceToHost, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    if (nsteps_out) *nsteps_out = eng->n_sampled;
    return 0;
}

extern "C" int hs_dsmc_wall_face_count(hs_dsmc_engine* eng, int64_t* n_faces) {
    if (!eng || !n_faces) return fail(eng, HS_E_INVALID, "NULL argument");
    *n_faces = eng->hp.n_wall;
    return 0;
}

extern "C" int hs_dsmc_wall_faces(hs_dsmc_engine* eng, int32_t* side, int32_t* cell, double* area, double* normal) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    const size_t n = eng->wall_side.size();
    if (side) std::copy(eng->wall_side.begin(), eng->wall_side.end(), side);
    if (cell) std::copy(eng->wall_cell.begin(), eng->wall_cell.end(), cell);
    if (area) std::copy(eng->wall_area.begin(), eng->wall_area.end(), area);
    if (normal) std::copy(eng->wall_normal.begin(), eng->wall_normal.begin() + 3 * n, normal);
    return 0;
}

extern "C" int hs_dsmc_read_wall_samples(hs_dsmc_engine* eng, double* out, int64_t n_doubles, int64_t* nsteps_out) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    const int64_t need = (int64_t)eng->ns * HS_WS_COUNT * eng->hp.n_wall;
    if (out && n_doubles < need) return fail(eng, HS_E_INVALID, "wall sample buffer too small");
    HS_TRY(hipSetDevice(eng->device));
    if (out && need > 0)
        HS_TRY(hipMemcpyAsync(out, eng->d_wall_sum, need * 8, hipMemcpyDeviceToHost, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    if (nsteps_out) *nsteps_out = eng->n_sampled;
    return 0;
}

extern "C" int hs_dsmc_set_samples(hs_dsmc_engine* eng, const double* cell_sums, int64_t n_cell,
                                   const double* wall_sums, int64_t n_wall, int64_t nsteps) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    const int64_t need_c = (int64_t)eng->nslots * eng->nc;
    const int64_t need_w = (int64_t)eng->ns * HS_WS_COUNT * eng->hp.n_wall;
    if (!cell_sums || n_cell != need_c)
        return fail(eng, HS_E_INVALID, "set_samples: cell sums must hold exactly " + std::to_string(need_c) + " values");
    if (wall_sums && n_wall != need_w)
        return fail(eng, HS_E_INVALID, "set_samples: wall sums must hold exactly " + std::to_string(need_w) + " values");
    if (nsteps < 0) return fail(eng, HS_E_INVALID, "set_samples: negative step count");
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipMemcpyAsync(eng->d_samples, cell_sums, need_c * 8, hipMemcpyHostToDevice, eng->stream));
    if (need_w > 0) {
        if (wall_sums)
            HS_TRY(hipMemcpyAsync(eng->d_wall_sum, wall_sums, need_w * 8, hipMemcpyHostToDevice, eng->stream));
        else
            launch_fill_f64(eng->d_wall_sum, 0.0, need_w, eng->stream);
    }
    HS_TRY(hipStreamSynchronize(eng->stream));
    eng->n_sampled = nsteps;
    return check_launch(eng, "set samples");
}

extern "C" int hs_dsmc_read_counters(hs_dsmc_engine* eng, int64_t* out, int32_t n) {
    if (!eng || !out) return fail(eng, HS_E_INVALID, "NULL argument");
    HS_TRY(hipSetDevice(eng->device));
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    for (int k = 0; k < n && k < CNT_N; ++k) out[k] = (int64_t)s.counters[k];
    return 0;
}

extern "C" int hs_dsmc_reset_counters(hs_dsmc_engine* eng) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    DevScalars s;
    if (int rc = read_scalars(eng, &s)) return rc;
    memset(s.counters, 0, sizeof(s.counters));
    HS_TRY(hipMemcpyAsync(eng->d_S, &s, sizeof(s), hipMemcpyHostToDevice, eng->stream));
    HS_TRY(hipStreamSynchronize(eng->stream));
    return 0;
}

extern "C" int hs_dsmc_kernel_times(hs_dsmc_engine* eng, double* ms, int64_t* launches, int32_t n) {
    if (!eng) return fail(eng, HS_E_INVALID, "NULL engine");
    HS_TRY(hipSetDevice(eng->device));
    HS_TRY(hipStreamSynchronize(eng->stream));
    for (auto& t : eng->timed) {
        float x = 0.0f;
        if (hipEventElapsedTime(&x, t.a, t.b) == hipSuccess) eng->closed_ms[t.stage] += x;
        eng->event_pool.push_back(t.a);
        eng->event_pool.push_back(t.b);
    }
    eng->timed.clear();
    for (int k = 0; k < n && k < T_NSTAGE; ++k) {
        if (ms) ms[k] = eng->closed_ms[k];
        if (launches) launches[k] = eng->launches[k];
    }
    if (n < 0) {  // negative n: reset
        for (int k = 0; k < T_NSTAGE; ++k) { eng->closed_ms[k] = 0.0; eng->launches[k] = 0; }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// test hooks
// ---------------------------------------------------------------------------
extern "C" int hs_dsmc_binary_collide(hs_dsmc_engine* eng, int64_t n, const int32_t* type_p, const int32_t* type_q,
                                      double* up, double* uq, double* erot_p, double* erot_q, int32_t* vib_p,
                                      int32_t* vib_q, const double* uniforms, int32_t n_uniform, int32_t* used_out) {
    if (!eng || !type_p || !type_q || !up || !uq || !uniforms || n_uniform <= 0)
        return fail(eng, HS_E_INVALID, "NULL argument");
    if (n <= 0) return 0;
    for (int64_t k = 0; k < n; ++k)
        if (type_p[k] < 0 || type_p[k] >= eng->ns || type_q[k] < 0 || type_q[k] >= eng->ns)
            return fail(eng, HS_E_INVALID, "typeId out of range");
    HS_TRY(hipSetDevice(eng->device));
    hipStream_t st = eng->stream;
    hipError_t ce = hipSuccess;
    auto ck = [&](hipError_t x) { if (ce == hipSuccess) ce = x; };
    int32_t *dtp, *dtq, *dvp = nullptr, *dvq = nullptr, *dused;
    double *dup, *duq, *dep = nullptr, *deq = nullptr, *duni;
    HS_TRY(hipMalloc(&dtp, n * 4));
    HS_TRY(hipMalloc(&dtq, n * 4));
    HS_TRY(hipMalloc(&dup, n * 24));
    HS_TRY(hipMalloc(&duq, n * 24));
    HS_TRY(hipMalloc(&duni, (size_t)n * n_uniform * 8));
    HS_TRY(hipMalloc(&dused, n * 4));
    if (erot_p) HS_TRY(hipMalloc(&dep, n * 8));
    if (erot_q) HS_TRY(hipMalloc(&deq, n * 8));
    if (vib_p) HS_TRY(hipMalloc(&dvp, n * 4 * HS_MAX_VIB_MODES));
    if (vib_q) HS_TRY(hipMalloc(&dvq, n * 4 * HS_MAX_VIB_MODES));
    ck(hipMemcpyAsync(dtp, type_p, n * 4, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(dtq, type_q, n * 4, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(dup, up, n * 24, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(duq, uq, n * 24, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(duni, uniforms, (size_t)n * n_uniform * 8, hipMemcpyHostToDevice, st));
    if (dep) ck(hipMemcpyAsync(dep, erot_p, n * 8, hipMemcpyHostToDevice, st));
    if (deq) ck(hipMemcpyAsync(deq, erot_q, n * 8, hipMemcpyHostToDevice, st));
    if (dvp) ck(hipMemcpyAsync(dvp, vib_p, n * 4 * HS_MAX_VIB_MODES, hipMemcpyHostToDevice, st));
    if (dvq) ck(hipMemcpyAsync(dvq, vib_q, n * 4 * HS_MAX_VIB_MODES, hipMemcpyHostToDevice, st));
    launch_binary_collide(eng->dp, n, dtp, dtq, dup, duq, dep, deq, dvp, dvq, duni, n_uniform, dused, st);
    int rc = check_launch(eng, "binary collide");
    ck(hipMemcpyAsync(up, dup, n * 24, hipMemcpyDeviceToHost, st));
    ck(hipMemcpyAsync(uq, duq, n * 24, hipMemcpyDeviceToHost, st));
    if (dep) ck(hipMemcpyAsync(erot_p, dep, n * 8, hipMemcpyDeviceToHost, st));
    if (deq) ck(hipMemcpyAsync(erot_q, deq, n * 8, hipMemcpyDeviceToHost, st));
    if (dvp) ck(hipMemcpyAsync(vib_p, dvp, n * 4 * HS_MAX_VIB_MODES, hipMemcpyDeviceToHost, st));
    if (dvq) ck(hipMemcpyAsync(vib_q, dvq, n * 4 * HS_MAX_VIB_MODES, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> used(n);
    ck(hipMemcpyAsync(used.data(), dused, n * 4, hipMemcpyDeviceToHost, st));
    ck(hipStreamSynchronize(st));
    if (ce != hipSuccess && !rc) rc = fail(eng, HS_E_HIP, std::string("binary collide: ") + hipGetErrorString(ce));
    if (used_out) memcpy(used_out, used.data(), n * 4);
    for (int64_t k = 0; k < n && !rc; ++k)
        if (used[k] > n_uniform) rc = fail(eng, HS_E_OVERFLOW, "a pair consumed more uniforms than provided");
    (void)hipFree(dtp); (void)hipFree(dtq); (void)hipFree(dup); (void)hipFree(duq); (void)hipFree(duni); (void)hipFree(dused);
    if (dep) (void)hipFree(dep);
    if (deq) (void)hipFree(deq);
    if (dvp) (void)hipFree(dvp);
    if (dvq) (void)hipFree(dvq);
    return rc;
}

extern "C" int hs_dsmc_sigma_tcr(hs_dsmc_engine* eng, int64_t n, const int32_t* type_p, const int32_t* type_q,
                                 const double* up, const double* uq, double* out) {
    if (!eng || !type_p || !type_q || !up || !uq || !out) return fail(eng, HS_E_INVALID, "NULL argument");
    if (n <= 0) return 0;
    for (int64_t k = 0; k < n; ++k)
        if (type_p[k] < 0 || type_p[k] >= eng->ns || type_q[k] < 0 || type_q[k] >= eng->ns)
            return fail(eng, HS_E_INVALID, "typeId out of range");
    HS_TRY(hipSetDevice(eng->device));
    hipStream_t st = eng->stream;
    hipError_t ce = hipSuccess;
    auto ck = [&](hipError_t x) { if (ce == hipSuccess) ce = x; };
    int32_t *dtp, *dtq;
    double *dup, *duq, *dout;
    HS_TRY(hipMalloc(&dtp, n * 4));
    HS_TRY(hipMalloc(&dtq, n * 4));
    HS_TRY(hipMalloc(&dup, n * 24));
    HS_TRY(hipMalloc(&duq, n * 24));
    HS_TRY(hipMalloc(&dout, n * 8));
    ck(hipMemcpyAsync(dtp, type_p, n * 4, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(dtq, type_q, n * 4, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(dup, up, n * 24, hipMemcpyHostToDevice, st));
    ck(hipMemcpyAsync(duq, uq, n * 24, hipMemcpyHostToDevice, st));
    launch_sigma_tcr(eng->dp, n, dtp, dtq, dup, duq, dout, st);
    int rc = check_launch(eng, "sigma_tcr");
    ck(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, st));
    ck(hipStreamSynchronize(st));
    if (ce != hipSuccess && !rc) rc = fail(eng, HS_E_HIP, std::string("sigma_tcr: ") + hipGetErrorString(ce));
    (void)hipFree(dtp); (void)hipFree(dtq); (void)hipFree(dup); (void)hipFree(duq); (void)hipFree(dout);
    return rc;
}

// ---------------------------------------------------------------------------
// device copy: achievable HBM bandwidth of a plain read+write stream
// ---------------------------------------------------------------------------
extern "C" int hs_dsmc_device_copy(int64_t bytes, int32_t iters, double* ms_out) {
    if (bytes <= 0 || bytes % 32 != 0 || iters <= 0 || !ms_out) return fail(nullptr, HS_E_INVALID, "bad copy size");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(nullptr, HS_E_HIP, "device copy: allocation failed");
    } else {
        int dev = 0, ncu = 256;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            ncu = prop.multiProcessorCount;
        launch_fill_f64((double*)a, 1.0, bytes / 8, st);
        launch_copy(a, b, bytes, ncu * 16, st);  // warm
        (void)hipEventRecord(e0, st);
        for (int k = 0; k < iters; ++k) launch_copy(a, b, bytes, ncu * 16, st);
        (void)hipEventRecord(e1, st);
        float ms = 0.0f;
        if (hipStreamSynchronize(st) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(nullptr, HS_E_HIP, "device copy failed");
        *ms_out = (double)ms / iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
