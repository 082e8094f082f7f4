// dsmc_kernels.hip -- the per-timestep kernels of the MI355X DSMC engine.
//
// One time step (dsmcCloud::evolve, dsmcCloud.C:819-926) is
//   inject_count -> scan(append) -> inject_gen        freestream inflow
//   move (+ per-cell histogram via atomics)           Cloud<dsmcParcel>::move
//   scan(cells)                                       cell start offsets
//   scatter                                           slot -> old index
//   collide                                           stable gather into cell
//                                                     order, NTC + binary
//                                                     collisions, sampling
// All counts live on the device so a step can be replayed without host
// round trips.  Every kernel checks its indices against the device-side
// counts; loops that draw random numbers are bounded.
#include "dsmc_kernels.h"

#define BLOCK 256
#define WAVE 64

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    return x;
}

// exclusive scan over a 256-thread block; returns the exclusive prefix of v
// and the block total.
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
    __shared__ int ws[BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int incl = wave_incl_scan(v, lane);
    if (lane == WAVE - 1) ws[wid] = incl;
    __syncthreads();
    int woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / WAVE; ++w) {
        const int x = ws[w];
        if (w < wid) woff += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return incl - v + woff;
}

__device__ __forceinline__ void add_counter(DevScalars* S, int k, unsigned long long v) {
    v = wave_sum_u64(v);
    if ((threadIdx.x & (WAVE - 1)) == 0 && v) atomicAdd(&S->counters[k], v);
}

// ---------------------------------------------------------------------------
// exclusive scan (reduce -> scan block sums -> apply), 2048 items per block
// ---------------------------------------------------------------------------
#define SCAN_ITEMS 8
#define SCAN_TILE (BLOCK * SCAN_ITEMS)

// publishes the scan total as k_scan_apply does
__device__ __forceinline__ void scan_publish(int32_t* out, int32_t n, int total, int mode, DevScalars* S,
                                             int64_t capacity) {
    out[n] = total;
    S->total = total;
    if (mode == SCAN_CELLS) {
        S->n_sorted = total;
    } else if (mode == SCAN_APPEND) {
        const int64_t nn = (int64_t)S->n + total;
        S->n_old = S->n;
        S->n = (int32_t)(nn > capacity ? capacity : nn);
        if (nn > capacity) atomicAdd(&S->counters[CNT_CAP_OVF], (unsigned long long)(nn - capacity));
    } else if (mode == SCAN_APPEND_NEXT) {
        // the next step's inflow, appended after the sorted population while
        // this step's sort and collide still use n and n_old (commit installs it)
        const int64_t nn = (int64_t)S->n_sorted + total;
        S->n_next = (int32_t)(nn > capacity ? capacity : nn);
        if (nn > capacity) atomicAdd(&S->counters[CNT_CAP_OVF], (unsigned long long)(nn - capacity));
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const int32_t* __restrict__ in, int32_t n,
                                                       int32_t* __restrict__ bsum) {
    const int base = blockIdx.x * SCAN_TILE;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int i = base + k * BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    int tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(BLOCK) void k_scan_blocks(int32_t* __restrict__ bsum, int32_t nb,
                                                       DevScalars* __restrict__ S) {
    int carry = 0;
    for (int base = 0; base < nb; base += BLOCK) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int tot;
        const int ex = block_excl_scan(v, &tot);
        if (i < nb) bsum[i] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) S->total = carry;
}

__global__ __launch_bounds__(BLOCK) void k_scan_apply(int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                      int32_t n, const int32_t* __restrict__ bsum, int mode,
                                                      DevScalars* __restrict__ S, int64_t capacity) {
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_excl_scan(s, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int i = base + k;
        if (i < n) {
            out[i] = ex;
            if (mode == SCAN_CELLS) in[i] = 0;
        }
        ex += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) scan_publish(out, n, S->total, mode, S, capacity);
}

// The whole scan in one workgroup (short arrays): tiles in order with a carry.
__device__ __forceinline__ void scan_single_body(int32_t* __restrict__ in, int32_t* __restrict__ out, int32_t n,
                                                 int mode, DevScalars* __restrict__ S, int64_t capacity) {
    int carry = 0;
    for (int tile = 0; tile < n; tile += SCAN_TILE) {
        const int base = tile + threadIdx.x * SCAN_ITEMS;
        int v[SCAN_ITEMS];
        int s = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const int i = base + k;
            v[k] = i < n ? in[i] : 0;
            s += v[k];
        }
        int tot;
        int ex = block_excl_scan(s, &tot) + carry;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const int i = base + k;
            if (i < n) {
                out[i] = ex;
                if (mode == SCAN_CELLS) in[i] = 0;
            }
            ex += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) scan_publish(out, n, carry, mode, S, capacity);
}

__global__ __launch_bounds__(BLOCK) void k_scan_single(int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                       int32_t n, int mode, DevScalars* __restrict__ S,
                                                       int64_t capacity) {
    scan_single_body(in, out, n, mode, S, capacity);
}

void launch_scan(const int32_t* in, int32_t* out, int32_t n, ScanMode mode, ScanWs& ws, DevScalars* S,
                 int64_t capacity, hipStream_t st) {
    const int nb = n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 1;
    if (nb <= 8) {
        k_scan_single<<<1, BLOCK, 0, st>>>(const_cast<int32_t*>(in), out, n, (int)mode, S, capacity);
        return;
    }
    k_scan_reduce<<<nb, BLOCK, 0, st>>>(in, n, ws.bsum);
    k_scan_blocks<<<1, BLOCK, 0, st>>>(ws.bsum, nb, S);
    k_scan_apply<<<nb, BLOCK, 0, st>>>(const_cast<int32_t*>(in), out, n, ws.bsum, (int)mode, S, capacity);
}

// ---------------------------------------------------------------------------
// the cell scan in one pass (decoupled look-back): counts -> start offsets,
// counts cleared, n_sorted published, and the gather and collide tilings
// marked from the same values
// ---------------------------------------------------------------------------
// A block's status word: epoch << 34 | flag << 32 | value, flag 1 = the
// block's aggregate, 2 = the inclusive prefix through the block.  A word from
// an earlier epoch (an earlier scan) reads as not ready, so nothing is cleared
// between scans.  Every launch over the same cells takes gridDim.x tickets from
// one counter, so ticket t is block t % gridDim.x of scan t / gridDim.x: the
// block and the epoch come from the device alone, and a captured launch
// replays correctly.
#define HEAVY_CELL 96   // parcels: a cell this full marks its collide tile heavy
#define LB_AGG 1ull
#define LB_INCL 2ull
#define LB_SPIN_MAX (1 << 24)

__device__ __forceinline__ unsigned long long lb_word(uint32_t epoch, unsigned long long flag, int32_t v) {
    return ((unsigned long long)epoch << 34) | (flag << 32) | (unsigned long long)(uint32_t)v;
}

// The gather and collide tiles starting between cells i and i+1 (the parcels
// of cell i occupy [e, e + c)): tile q of size pt starts at the first cell
// whose start offset is >= q*pt.
__device__ __forceinline__ void tile_marks_after(int i, int32_t e, int32_t c, int pt, int32_t* __restrict__ tb) {
    const int q0 = e / pt, q1 = (e + c) / pt;
    for (int q = q0 + 1; q <= q1; ++q) tb[q] = i + 1;
}

// gx: the launch's workgroups over these cells (scan_cells_blocks(n))
__device__ __forceinline__ void scan_cells_body(int32_t* __restrict__ cnt, int32_t* __restrict__ start, int32_t n,
                                                unsigned long long* __restrict__ status,
                                                unsigned long long* __restrict__ ticket,
                                                DevScalars* __restrict__ S, int pt_g, int32_t* __restrict__ tile_g,
                                                int pt_c, int32_t* __restrict__ tile_c, int32_t* __restrict__ heavy,
                                                const int gx) {
    __shared__ int s_blk;
    __shared__ uint32_t s_epoch;
    __shared__ int s_excl;
    // logical block = order of start, so every predecessor has started
    if (threadIdx.x == 0) {
        const unsigned long long t = atomicAdd(ticket, 1ull);
        s_blk = (int)(t % (unsigned long long)gx);
        s_epoch = (uint32_t)((t / (unsigned long long)gx) % 0x3fffffffull) + 1;  // 1 .. 2^30 - 1: never a cleared word's 0
    }
    __syncthreads();
    const int b = s_blk;
    const uint32_t epoch = s_epoch;
    const int i0 = b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;  // SCAN_ITEMS consecutive cells per thread
    int v[SCAN_ITEMS];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (i0 + k < n) ? cnt[i0 + k] : 0;
        sum += v[k];
    }
    int tot;
    int ex = block_excl_scan(sum, &tot);
    if (threadIdx.x < WAVE) {
        // wave 0 looks back over 64 predecessors at a time: the nearest one
        // with its inclusive prefix ends the walk (blocks below 0 read as an
        // inclusive prefix of 0)
        const int lane = threadIdx.x;
        if (lane == 0) atomicExch(&status[b], lb_word(epoch, b == 0 ? LB_INCL : LB_AGG, tot));
        int excl = 0, j = b - 1, spins = 0;
        while (j >= 0) {
            const int jj = j - lane;
            const unsigned long long w = jj >= 0 ? __hip_atomic_load(&status[jj], __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT)
                                                 : lb_word(epoch, LB_INCL, 0);
            const bool ready = (uint32_t)(w >> 34) == epoch;
            const bool incl = ready && ((w >> 32) & 3ull) == LB_INCL;
            const unsigned long long im = __ballot(incl), nr = __ballot(!ready);
            const int fi = im ? __ffsll((long long)im) - 1 : WAVE;
            const unsigned long long need = fi >= WAVE - 1 ? ~0ull : ((2ull << fi) - 1ull);
            if (nr & need) {
                if (++spins > LB_SPIN_MAX) {  // a predecessor never published: report, do not hang
                    if (lane == 0) atomicAdd(&S->counters[CNT_TILE_OVF], 1ull);
                    break;
                }
                continue;
            }
            int val = lane <= fi ? (int32_t)(uint32_t)w : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, WAVE);
            excl += val;
            if (fi < WAVE) break;
            j -= WAVE;
        }
        if (lane == 0) {
            if (b > 0) atomicExch(&status[b], lb_word(epoch, LB_INCL, excl + tot));
            s_excl = excl;
        }
    }
    __syncthreads();
    ex += s_excl;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int i = i0 + k;
        if (i < n) {
            start[i] = ex;
            cnt[i] = 0;
            tile_marks_after(i, ex, v[k], pt_g, tile_g);
            tile_marks_after(i, ex, v[k], pt_c, tile_c);
            if (v[k] > HEAVY_CELL) heavy[ex / pt_c] = 1;
            if (i == n - 1) {
                const int total = ex + v[k];
                start[n] = total;
                S->total = total;
                S->n_sorted = total;
                tile_g[0] = 0;
                tile_c[0] = 0;
                tile_g[total / pt_g + 1] = n;
                tile_c[total / pt_c + 1] = n;
            }
        }
        ex += v[k];
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_cells(int32_t* __restrict__ cnt, int32_t* __restrict__ start, int32_t n,
                                                      unsigned long long* __restrict__ status,
                                                      unsigned long long* __restrict__ ticket,
                                                      DevScalars* __restrict__ S, int pt_g,
                                                      int32_t* __restrict__ tile_g, int pt_c,
                                                      int32_t* __restrict__ tile_c, int32_t* __restrict__ heavy) {
    scan_cells_body(cnt, start, n, status, ticket, S, pt_g, tile_g, pt_c, tile_c, heavy, (int)gridDim.x);
}

void launch_scan_cells(int32_t* cnt, int32_t* start, int32_t n, unsigned long long* status,
                       unsigned long long* ticket, DevScalars* S, int pt_g, int32_t* tile_g,
                       int pt_c, int32_t* tile_c, int32_t* heavy, hipStream_t st) {
    k_scan_cells<<<scan_cells_blocks(n), BLOCK, 0, st>>>(cnt, start, n, status, ticket, S, pt_g, tile_g, pt_c,
                                                          tile_c, heavy);
}

int scan_cells_blocks(int32_t n) { return n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 1; }

// ---------------------------------------------------------------------------
// generic fills (graph-safe replacements for hipMemsetAsync)
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_i64(int64_t* p, int64_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
static int grid_for(int64_t n, int cap = 4096) {
    int64_t g = (n + BLOCK - 1) / BLOCK;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
    k_fill_i32<<<grid_for(n), BLOCK, 0, st>>>(p, v, n);
}
void launch_fill_f64(double* p, double v, int64_t n, hipStream_t st) {
    k_fill_f64<<<grid_for(n), BLOCK, 0, st>>>(p, v, n);
}
void launch_fill_i64(int64_t* p, int64_t v, int64_t n, hipStream_t st) {
    k_fill_i64<<<grid_for(n), BLOCK, 0, st>>>(p, v, n);
}

// ---------------------------------------------------------------------------
// new-particle internal state (equipartition at the given temperatures)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void new_internal(const DevParams& P, DevParticles& b, int64_t j, int t, double Trot,
                                             double Tvib, PhiloxStream& r) {
    if (b.erot) b.erot[j] = (float)equip_erot(r, Trot, P.rotdof[t]);
    if (b.vib)
        for (int m = 0; m < P.nvib[t]; ++m)
            b.vib[m * P.capacity + j] = (uint16_t)equip_vib_level(r, Tvib, P.theta_v[t][m]);
    if (b.type) b.type[j] = (uint8_t)t;
}

// ---------------------------------------------------------------------------
// freestream inflow (dsmcFreeStreamInflowPatch.C:91-364)
// ---------------------------------------------------------------------------
// ahead = 1: the next step's inflow, drawn before this step's commit
// advances the step counter.
__device__ __forceinline__ void inject_count_body(const DevParams* __restrict__ Pp, const DevFaces& F, int ns,
                                                  DevScalars* __restrict__ S, int ahead, const int bx) {
    const DevParams& P = *Pp;
    const int e = bx * BLOCK + threadIdx.x;
    if (e >= F.n * ns) return;
    const int f = e / ns, s = e - f * ns;
    if (!(P.inflow_n[s] > 0.0)) { F.n_new[e] = 0; return; }
    double* acc = &F.acc[s * F.n + f];
    double a = *acc + F.inc[s * F.n + f];
    int ni = (int)a;
    if (ni < 0) ni = 0;
    PhiloxStream r;
    r.init(P, RNG_INJECT, (uint32_t)e, S->step + (uint32_t)ahead, 0);
    if ((a - (double)ni) > r.u()) ni += 1;
    a -= (double)ni;
    *acc = a;
    F.n_new[e] = ni;
}

__global__ __launch_bounds__(BLOCK) void k_inject_count(const DevParams* __restrict__ Pp, DevFaces F, int ns,
                                                        DevScalars* __restrict__ S, int ahead) {
    inject_count_body(Pp, F, ns, S, ahead, (int)blockIdx.x);
}

// Parcel j of this step belongs to the entry e (face*ns + species) with
// off[e] <= j < off[e+1] and is its (j-off[e])-th; it is written at idx.
// off: the entries' offsets (LDS copy or global).
__device__ void inject_one(const DevParams& P, const DevFaces& F, const int32_t* off, int ns, uint32_t step,
                           DevParticles& b,
                           const DevScalars* S, int j, int64_t idx) {
    const int ne = F.n * ns;
    int lo = 0, hi = ne;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= j) lo = mid;
        else hi = mid;
    }
    const int e = lo;
    const int k = j - off[e];
    const int f = e / ns, s = e - f * ns;
    PhiloxStream r;
    r.init(P, RNG_INJECT, (uint32_t)e, step, 1u + (uint32_t)k);
    const int a = F.axis[f], dir = F.dir[f];
    const int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
    const int nn[3] = { P.nx, P.ny, P.nz };
    const double mass = P.mass[s];
    const double mps = sqrt(2.0 * HS_KB * P.inflow_t / mass);
    const double un_face = P.inflow_u[a] * (double)(-dir);
    const double s_cos = un_face / mps;
    const double coef_a = s_cos + sqrt(s_cos * s_cos + 2.0);
    const double coef_b = 0.5 * (1.0 + s_cos * (s_cos - sqrt(s_cos * s_cos + 2.0)));
    double scaling = 3.0;
    if (s_cos < -3.0) scaling = fabs(s_cos) + 1.0;
    const double sc = sqrt(HS_KB * P.inflow_t / mass);
    double pos[3];
    const double ua = r.u(), ub = r.u();
    pos[a] = face_coord(P, a, dir > 0 ? nn[a] : 0);
    pos[a1] = P.empty[a1] ? P.centre[a1] : face_coord(P, a1, F.a1i[f]) + ua * P.delta[a1];
    pos[a2] = P.empty[a2] ? P.centre[a2] : face_coord(P, a2, F.a2i[f]) + ub * P.delta[a2];
    double un;
    if (fabs(un_face) > HS_VSMALL) {
        double Pa;
        int guard = 0;
        do {
            const double unt = scaling * (2.0 * r.u() - 1.0);
            un = unt + s_cos;
            if (un < 0.0) Pa = -1.0;
            else Pa = 2.0 * un / coef_a * dm_exp(coef_b - unt * unt);
        } while (Pa < r.u() && ++guard < HS_MAX_REDRAW);
    } else {
        un = sqrt(-dm_log(1.0 - r.u()));
    }
    const double g1 = r.gauss(), g2 = r.gauss();
    double U[3];
    U[a] = mps * un * (double)(-dir);
    U[a1] = sc * g1 + P.inflow_u[a1];
    U[a2] = sc * g2 + P.inflow_u[a2];
    b.px[idx] = pos[0]; b.py[idx] = pos[1]; b.pz[idx] = pos[2];
    b.ux[idx] = (float)U[0]; b.uy[idx] = (float)U[1]; b.uz[idx] = (float)U[2];
    b.cell[idx] = F.cell[f];
    if (b.vib)
        for (int m = 0; m < P.max_modes; ++m) b.vib[m * P.capacity + idx] = 0;
    new_internal(P, b, idx, s, P.inflow_trot, P.inflow_tvib, r);
}

// One thread per new parcel; the entry offsets are searched in an LDS copy
// when they fit.
#define INJ_LDS_OFF 4096
__device__ __forceinline__ void inject_gen_body(const DevParams* __restrict__ Pp, const DevFaces& F, int ns,
                                                DevParticles& b, DevScalars* __restrict__ S, int ahead, const int bx) {
    __shared__ int32_t soff[INJ_LDS_OFF];
    const DevParams& P = *Pp;
    const int ne = F.n * ns;
    const bool lds = ne + 1 <= INJ_LDS_OFF;
    if (lds) {
        for (int q = threadIdx.x; q <= ne; q += BLOCK) soff[q] = F.off[q];
        __syncthreads();
    }
    const int32_t* off = lds ? soff : F.off;
    const int j = bx * BLOCK + threadIdx.x;
    const int64_t idx = (int64_t)(ahead ? S->n_sorted : S->n_old) + j;
    const bool mine = j < off[ne] && idx < P.capacity;
    if (mine) inject_one(P, F, off, ns, S->step + (uint32_t)ahead, b, S, j, idx);
    // every lane reaches the wave reduction of the counter
    add_counter(S, CNT_INJ, mine ? 1ull : 0ull);
}

__global__ __launch_bounds__(BLOCK) void k_inject_gen(const DevParams* __restrict__ Pp, DevFaces F, int ns,
                                                      DevParticles b, DevScalars* __restrict__ S, int ahead) {
    inject_gen_body(Pp, F, ns, b, S, ahead, (int)blockIdx.x);
}

void launch_inject_count(const DevParams* P, const DevFaces& F, int ns, DevScalars* S, int ahead, hipStream_t st) {
    const int n = F.n * ns;
    if (n == 0) return;
    k_inject_count<<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(P, F, ns, S, ahead);
}

void launch_inject_gen(const DevParams* P, const DevFaces& F, int ns, DevParticles b, DevScalars* S, int ahead,
                       hipStream_t st) {
    if (F.n * ns == 0 || F.max_new <= 0) return;
    k_inject_gen<<<(F.max_new + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(P, F, ns, b, S, ahead);
}

// ---------------------------------------------------------------------------
// move (dsmcParcel::move, dsmcParcel.C:38-148, on the Cartesian block) with
// the cell histogram of buildCellOccupancy fused in
// ---------------------------------------------------------------------------
struct MoveCounts {
    unsigned long long del, wall, ovf;
};

// State of one parcel's flight through the block.
struct Flight {
    double p[3], U[3];
    int idx[3];
    double t;      // flight time left
    int iter;      // tracking iterations so far
    bool keep;     // false: deleted
    bool vch;      // the velocity changed (a wall interaction)
};

// Wall interaction at p (code: side 0-5 of the block, 6 = the cylinder):
// dsmcDiffuseWallPatch / dsmcSpecularWallPatch with the boundaryMeasurements
// and face-tracker terms.
__device__ void wall_hit(const DevParams& P, Flight& s, int code, int t_id, float* erot, uint16_t* vib,
                         PhiloxStream& r, bool measure, bool track, MoveCounts& mc) {
    mc.wall += 1;
    s.vch = true;
    double* U = s.U;
    const int* idx = s.idx;
    if (code == 6) {
        const double nx_ = P.cyl_c[0] - s.p[0], ny_ = P.cyl_c[1] - s.p[1];
        const double nm = sqrt(nx_ * nx_ + ny_ * ny_);
        const double nw[3] = { nx_ / nm, ny_ / nm, 0.0 };
        const int32_t wf = (measure || track) ? P.cyl_face[idx[0] + P.nx * (idx[1] + P.ny * idx[2])] : -1;
        const int32_t f = measure ? wf : -1;
        double e_pre = 0.0, u_pre[3] = { U[0], U[1], U[2] };
        if (f >= 0) e_pre = wall_side_terms(P, f, t_id, nw, U, erot, vib, P.capacity);
        if (P.cyl_model == HS_PATCH_DIFFUSE) {
            diffuse_reflect(P, P.cyl_T, P.cyl_U, nw, t_id, U, erot, vib, P.capacity, r);
        } else {
            const double und = U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2];
            if (und > 0.0)
                for (int q = 0; q < 3; ++q) U[q] = U[q] - 2.0 * und * nw[q];
        }
        if (f >= 0) {
            const double e_post = wall_side_terms(P, f, t_id, nw, U, erot, vib, P.capacity);
            wall_exchange_terms(P, f, t_id, e_pre, u_pre, e_post, U);
        }
        if (track && wf >= 0)
            track_face(P, t_id, P.face_off[3] + (wf - P.cyl_first),
                       (U[0] * nw[0] + U[1] * nw[1] + U[2] * nw[2]) >= 0.0 ? 1 : -1);
        return;
    }
    const int axis = code / 2, dir = (code & 1) ? 1 : -1;
    const int model = P.patch_model[code];
    double nw[3] = { 0.0, 0.0, 0.0 };
    nw[axis] = (double)dir;
    int64_t f = -1;
    if (measure && P.wall_off[code] >= 0) {
        const int j1 = axis == 0 ? idx[1] : idx[0], j2 = axis == 2 ? idx[1] : idx[2];
        const int n1 = axis == 0 ? P.ny : P.nx;
        f = P.wall_off[code] + j1 + (int64_t)n1 * j2;
    }
    double e_pre = 0.0, u_pre[3] = { U[0], U[1], U[2] };
    if (f >= 0) e_pre = wall_side_terms(P, f, t_id, nw, U, erot, vib, P.capacity);
    if (model == HS_PATCH_DIFFUSE)
        diffuse_reflect(P, P.patch_T[code], P.patch_U[code], nw, t_id, U, erot, vib, P.capacity, r);
    else if (U[axis] * (double)dir > 0.0)
        U[axis] = -U[axis];
    if (f >= 0) {
        const double e_post = wall_side_terms(P, f, t_id, nw, U, erot, vib, P.capacity);
        wall_exchange_terms(P, f, t_id, e_pre, u_pre, e_post, U);
    }
    if (track) {
        const int nn_a = axis == 0 ? P.nx : (axis == 1 ? P.ny : P.nz);
        track_face(P, t_id, mesh_face(P, axis, dir > 0 ? nn_a : 0, idx), U[axis] * (double)dir >= 0.0 ? 1 : -1);
    }
}

// Flight of one parcel (dsmcParcel::move, dsmcParcel.C:38-148, on the
// Cartesian block): face crossings, cyclic and deletion sides, walls.
// DEFER: stops at the first wall whose interaction draws random numbers or
// takes wall sums (a diffuse side, the cylinder, a specular side while
// sampling) with p on the wall and returns its code (side 0-5, 6 cylinder);
// otherwise every wall is handled here.  Returns -1 when the flight is over.
template <bool DEFER>
__device__ int fly(const DevParams& P, Flight& s, int t_id, float* erot, uint16_t* vib, PhiloxStream& r,
                   bool measure, bool track, MoveCounts& mc) {
    const int nn[3] = { P.nx, P.ny, P.nz };
    double* p = s.p;
    double* U = s.U;
    int* idx = s.idx;
    while (s.keep && s.t > HS_ROOTVSMALL) {
        if (++s.iter > HS_MAX_TRACK) { mc.ovf += 1; break; }
        const double t = s.t;
        // Which faces the straight flight to e crosses.  The fraction to the
        // face ahead on axis a is l = num/den (clamped at 0) and the face is
        // crossed when l < 1; for den > 0 that is exactly num < den (for den < 0,
        // num > den), since a correctly rounded quotient of two positive
        // doubles is below 1 exactly when the dividend is below the divisor.
        // The divisions are taken only for the faces actually crossed.
        double ut[3], e[3], den[3], num[3];
        bool cr[3];
        bool any = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ut[a] = P.empty[a] ? 0.0 : U[a];
            e[a] = p[a] + t * ut[a];
            den[a] = e[a] - p[a];
            const bool up = ut[a] > 0.0 && den[a] > 0.0;
            const bool dn = ut[a] < 0.0 && den[a] < 0.0;
            num[a] = face_coord(P, a, up ? idx[a] + 1 : idx[a]) - p[a];
            cr[a] = (up && num[a] < den[a]) || (dn && num[a] > den[a]);
            any = any || cr[a];
        }
        // the cylinder: its hit fraction cc/qc (clamped at 0, qc > 0) is below 1
        // exactly when cc < qc
        bool hit = false;
        double cc = 0.0, qc = 1.0;
        if (P.cyl) {
            const double dx = den[0], dy = den[1];
            const double rx = p[0] - P.cyl_c[0], ry = p[1] - P.cyl_c[1];
            const double bb = rx * dx + ry * dy;
            if (bb < 0.0) {
                const double aa = dx * dx + dy * dy;
                cc = rx * rx + ry * ry - P.cyl_r * P.cyl_r;
                const double disc = bb * bb - aa * cc;
                if (disc >= 0.0) {
                    qc = sqrt(disc) - bb;
                    hit = cc < qc;
                }
            }
        }
        if (!any && !hit) {
            p[0] = e[0]; p[1] = e[1]; p[2] = e[2];
            s.t = 0.0;
            break;
        }
        double lam = 1.0;
        int axis = -1, dir = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (cr[a]) {
                double l = num[a] / den[a];
                if (l < 0.0) l = 0.0;
                if (l < lam) { lam = l; axis = a; dir = ut[a] > 0.0 ? 1 : -1; }
            }
        }
        bool wall = false;
        if (hit) {
            double sh = cc / qc;
            if (sh < 0.0) sh = 0.0;
            if (sh < lam) { lam = sh; wall = true; }
        }
        if (!wall && axis < 0) {  // not reached: a flagged face or hit always sets one
            p[0] = e[0]; p[1] = e[1]; p[2] = e[2];
            s.t = 0.0;
            break;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) p[q] = p[q] + lam * (e[q] - p[q]);
        s.t = t - t * lam;
        if (wall) {
            if (DEFER) return 6;
            wall_hit(P, s, 6, t_id, erot, vib, r, measure, track, mc);
            continue;
        }
        p[axis] = face_coord(P, axis, dir > 0 ? idx[axis] + 1 : idx[axis]);
        const int ni = idx[axis] + dir;
        if (ni >= 0 && ni < nn[axis]) {
            // internal face: Sf points +axis, U crosses it along dir
            if (track) track_face(P, t_id, mesh_face(P, axis, dir > 0 ? idx[axis] + 1 : idx[axis], idx), dir);
            idx[axis] = ni;
            continue;
        }
        const int side = 2 * axis + (dir > 0 ? 1 : 0);
        const int model = P.patch_model[side];
        if (model == HS_PATCH_CYCLIC) {
            // counted, unsigned, on the coupled face of the other side
            if (track) track_face(P, t_id, mesh_face(P, axis, dir > 0 ? 0 : nn[axis], idx), 1);
            p[axis] = face_coord(P, axis, dir > 0 ? 0 : nn[axis]);
            idx[axis] = dir > 0 ? 0 : nn[axis] - 1;
        } else if (model == HS_PATCH_DELETION) {
            if (track)
                track_face(P, t_id, mesh_face(P, axis, dir > 0 ? nn[axis] : 0, idx),
                           U[axis] * (double)dir >= 0.0 ? 1 : -1);
            s.keep = false;
        } else if (DEFER && (model == HS_PATCH_DIFFUSE || measure)) {
            return side;
        } else if (DEFER) {
            // specular side without wall sums (an empty plane is never crossed)
            mc.wall += 1;
            s.vch = true;
            if (U[axis] * (double)dir > 0.0) U[axis] = -U[axis];
            if (track)
                track_face(P, t_id, mesh_face(P, axis, dir > 0 ? nn[axis] : 0, idx),
                           U[axis] * (double)dir >= 0.0 ? 1 : -1);
        } else {
            wall_hit(P, s, side, t_id, erot, vib, r, measure, track, mc);
        }
    }
    return -1;
}

// End of a flight: the new state; returns the new cell (-1: deleted).
__device__ __forceinline__ int32_t land(const DevParams& P, DevParticles& b, int64_t i, const Flight& s,
                                        MoveCounts& mc) {
    b.px[i] = s.p[0]; b.py[i] = s.p[1]; b.pz[i] = s.p[2];
    if (s.vch) { b.ux[i] = (float)s.U[0]; b.uy[i] = (float)s.U[1]; b.uz[i] = (float)s.U[2]; }
    const int32_t cn = s.keep ? s.idx[0] + P.nx * (s.idx[1] + P.ny * s.idx[2]) : -1;
    if (!s.keep) mc.del += 1;
    b.cell[i] = cn;
    return cn;
}

__device__ __forceinline__ void load_pu(const DevParticles& b, int64_t i, Flight& s) {
    s.p[0] = b.px[i]; s.p[1] = b.py[i]; s.p[2] = b.pz[i];
    s.U[0] = (double)b.ux[i]; s.U[1] = (double)b.uy[i]; s.U[2] = (double)b.uz[i];
}

__device__ __forceinline__ void set_cell(const DevParams& P, int32_t c, Flight& s) {
    cell_ijk(P, c, s.idx);
    s.keep = true;
    s.vch = false;
}

// The move stream of parcel i: its draws in order are the step fraction of a
// new parcel (dsmcParcel.C:52-59), then the wall interactions.
__device__ __forceinline__ void move_stream(const DevParams& P, PhiloxStream& r, int64_t i, uint32_t step,
                                            bool is_new, double* t0) {
    r.init(P, RNG_MOVE, (uint32_t)i, step, 0);
    double t = P.dt;
    if (is_new) {
        const double sf = r.u();
        t = (1.0 - sf) * P.dt;
    }
    *t0 = t;
}

// Histogram slot of each particle.  Particles are nearly cell-sorted, so the
// lanes of a wave form runs of equal cells: one atomic per run returns the
// run's base slot and each lane adds its rank in the run.
__device__ __forceinline__ void hist_slots(int32_t c, bool valid, int64_t i, int lane, int32_t* __restrict__ count,
                                           int32_t* __restrict__ slot) {
    const int32_t prev = __shfl_up(c, 1, WAVE);
    const bool head = (lane == 0) || (prev != c);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
    const int leader = 63 - __clzll((long long)(heads & upto));
    const unsigned long long lead_mask = leader == WAVE - 1 ? ~0ull : ((2ull << leader) - 1ull);
    const unsigned long long above = heads & ~lead_mask;
    int base = 0;
    if (head && c >= 0) {
        const int next = above ? __ffsll((long long)above) - 1 : WAVE;
        base = atomicAdd(&count[c], next - lane);
    }
    base = __shfl(base, leader, WAVE);
    if (valid && c >= 0) slot[i] = base + (lane - leader);
}

// move (Cloud<dsmcParcel>::move) with the cell histogram of
// buildCellOccupancy fused in.  Flights that reach a wall needing random
// draws or wall sums are queued for k_move_wall, which keeps this kernel's
// registers (and so its occupancy) those of the free flight.  TRACK: the face
// tracker is on (HS_FLAG_FACE_FLUX with sampling); without it the tracker's
// code is left out of the kernel.
template <bool TRACK>
__device__ __forceinline__ void move_body(const DevParams* __restrict__ Pp, DevParticles& b,
                                          DevScalars* __restrict__ S, int32_t* __restrict__ count,
                                          int32_t* __restrict__ slot, const WallQueue& Q, int do_move, int do_hist,
                                          const int bx, const int gx) {
    const DevParams& P = *Pp;
    const int64_t n = S->n, n_old = S->n_old;
    const uint32_t step = S->step;
    const int lane = threadIdx.x & (WAVE - 1);
    const bool measure = (P.flags & HS_FLAG_WALL) && (P.flags & HS_FLAG_SAMPLE) && P.n_wall > 0;
    const bool track = TRACK;
    const unsigned long long lt = lane == 0 ? 0ull : ((1ull << lane) - 1ull);
    MoveCounts mc = { 0, 0, 0 };
    // the next parcel's cell, position and velocity are loaded while this
    // one flies
    const int64_t stride = (int64_t)gx * BLOCK;
    int32_t nx_c = -1;
    double nx_p[3] = { 0.0, 0.0, 0.0 };
    float nx_u[3] = { 0.0f, 0.0f, 0.0f };
    auto fetch = [&](int64_t j) {
        if (j < n) {
            nx_c = b.cell[j];
            if (do_move) {
                nx_p[0] = b.px[j]; nx_p[1] = b.py[j]; nx_p[2] = b.pz[j];
                nx_u[0] = b.ux[j]; nx_u[1] = b.uy[j]; nx_u[2] = b.uz[j];
            }
        }
    };
    fetch(bx * (int64_t)BLOCK + threadIdx.x);
    // wave-uniform trip count: every lane of a wave sees the same bases
    for (int64_t base = bx * (int64_t)BLOCK + (threadIdx.x - lane); base < n; base += stride) {
        const int64_t i = base + lane;
        const bool valid = i < n;
        int32_t c = -1;
        int code = -1;
        Flight s;
        if (valid) {
            c = nx_c;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s.p[a] = nx_p[a];
                s.U[a] = (double)nx_u[a];
            }
        }
        fetch(i + stride);
        if (valid) {
            if (do_move && c >= 0) {
                PhiloxStream r;
                move_stream(P, r, i, step, i >= n_old, &s.t);
                set_cell(P, c, s);
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if (P.empty[a]) s.p[a] = P.centre[a];
                s.iter = 0;
                const int t_id = b.type ? (int)b.type[i] : 0;
                code = fly<true>(P, s, t_id, nullptr, nullptr, r, measure, track, mc);
                if (code >= 0) {
                    // stopped at a wall: position on it, cell unchanged until the wall pass
                    b.px[i] = s.p[0]; b.py[i] = s.p[1]; b.pz[i] = s.p[2];
                    const int32_t cw = s.idx[0] + P.nx * (s.idx[1] + P.ny * s.idx[2]);
                    b.cell[i] = cw;
                    c = -1;
                } else {
                    c = land(P, b, i, s, mc);
                }
            }
        }
        if (do_move) {
            const unsigned long long dm = __ballot(code >= 0);
            if (dm) {
                int qb = 0;
                if (lane == __ffsll((long long)dm) - 1) qb = atomicAdd(&S->qn, __popcll(dm));
                qb = __shfl(qb, __ffsll((long long)dm) - 1, WAVE);
                if (code >= 0) {
                    const int k = qb + __popcll(dm & lt);
                    Q.i[k] = (int32_t)i;
                    Q.t[k] = s.t;
                    Q.code_iter[k] = code | (s.iter << 8);
                }
            }
        }
        if (do_hist) {
            if (valid && c >= P.nc) { b.cell[i] = -1; c = -1; }
            hist_slots(c, valid, i, lane, count, slot);
        }
    }
    if (do_move && bx == 0 && threadIdx.x == 0) atomicAdd(&S->counters[CNT_PSTEPS], (unsigned long long)n);
    add_counter(S, CNT_DEL, mc.del);
    add_counter(S, CNT_WALL, mc.wall);
    add_counter(S, CNT_TRACK_OVF, mc.ovf);
}

template <bool TRACK>
__global__ __launch_bounds__(BLOCK) void k_move(const DevParams* __restrict__ Pp, DevParticles b,
                                                DevScalars* __restrict__ S, int32_t* __restrict__ count,
                                                int32_t* __restrict__ slot, WallQueue Q, int do_move, int do_hist) {
    move_body<TRACK>(Pp, b, S, count, slot, Q, do_move, do_hist, (int)blockIdx.x, (int)gridDim.x);
}

// The queued flights: the wall interaction they stopped at, the rest of the
// flight (every further wall handled here), the cell histogram.  The queue is
// cleared afterwards by k_commit (or k_clear_queue).
template <bool TRACK>
__device__ __forceinline__ void move_wall_body(const DevParams* __restrict__ Pp, DevParticles& b,
                                               DevScalars* __restrict__ S, int32_t* __restrict__ count,
                                               int32_t* __restrict__ slot, const WallQueue& Q, int do_hist,
                                               const int bx, const int gx) {
    const DevParams& P = *Pp;
    const int64_t nq = S->qn, n_old = S->n_old;
    const uint32_t step = S->step;
    const int lane = threadIdx.x & (WAVE - 1);
    const bool measure = (P.flags & HS_FLAG_WALL) && (P.flags & HS_FLAG_SAMPLE) && P.n_wall > 0;
    const bool track = TRACK;
    MoveCounts mc = { 0, 0, 0 };
    for (int64_t base = bx * (int64_t)BLOCK + (threadIdx.x - lane); base < nq; base += (int64_t)gx * BLOCK) {
        const int64_t k = base + lane;
        const bool valid = k < nq;
        int32_t c = -1;
        int64_t i = 0;
        if (valid) {
            i = Q.i[k];
            const int ci = Q.code_iter[k];
            Flight s;
            PhiloxStream r;
            double t0;
            move_stream(P, r, i, step, i >= n_old, &t0);
            load_pu(b, i, s);
            set_cell(P, b.cell[i], s);
            s.t = Q.t[k];
            s.iter = ci >> 8;
            const int t_id = b.type ? (int)b.type[i] : 0;
            float* erot = b.erot ? &b.erot[i] : nullptr;
            uint16_t* vib = (b.vib && P.nvib[t_id] > 0) ? &b.vib[i] : nullptr;
            wall_hit(P, s, ci & 0xFF, t_id, erot, vib, r, measure, track, mc);
            fly<false>(P, s, t_id, erot, vib, r, measure, track, mc);
            c = land(P, b, i, s, mc);
        }
        if (do_hist) hist_slots(c, valid, i, lane, count, slot);
    }
    add_counter(S, CNT_DEL, mc.del);
    add_counter(S, CNT_WALL, mc.wall);
    add_counter(S, CNT_TRACK_OVF, mc.ovf);
}

template <bool TRACK>
__global__ __launch_bounds__(BLOCK) void k_move_wall(const DevParams* __restrict__ Pp, DevParticles b,
                                                     DevScalars* __restrict__ S, int32_t* __restrict__ count,
                                                     int32_t* __restrict__ slot, WallQueue Q, int do_hist) {
    move_wall_body<TRACK>(Pp, b, S, count, slot, Q, do_hist, (int)blockIdx.x, (int)gridDim.x);
}

__global__ void k_clear_queue(DevScalars* S) { S->qn = 0; }
void launch_clear_queue(DevScalars* S, hipStream_t st) { k_clear_queue<<<1, 1, 0, st>>>(S); }

void launch_move(const DevParams* P, DevParticles b, DevScalars* S, int32_t* count, int32_t* slot,
                 const WallQueue& Q, int do_move, int do_hist, int track, int grid, hipStream_t st) {
    const int gw = grid / 8 > 0 ? grid / 8 : 1;
    if (track) {
        k_move<true><<<grid, BLOCK, 0, st>>>(P, b, S, count, slot, Q, do_move, do_hist);
        if (do_move) k_move_wall<true><<<gw, BLOCK, 0, st>>>(P, b, S, count, slot, Q, do_hist);
    } else {
        k_move<false><<<grid, BLOCK, 0, st>>>(P, b, S, count, slot, Q, do_move, do_hist);
        if (do_move) k_move_wall<false><<<gw, BLOCK, 0, st>>>(P, b, S, count, slot, Q, do_hist);
    }
}

// Workgroups of k_move resident per CU (its register use sets the occupancy).
int move_blocks_per_cu(int track) {
    int nb = 0;
    const hipError_t e = track ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_move<true>, BLOCK, 0)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_move<false>, BLOCK, 0);
    if (e != hipSuccess || nb < 1) nb = 4;
    return nb;
}

__global__ void k_set_nold(DevScalars* S) { S->n_old = S->n; }
void launch_set_nold(DevScalars* S, hipStream_t st) { k_set_nold<<<1, 1, 0, st>>>(S); }

// ---------------------------------------------------------------------------
// scatter: perm[start[cell] + slot] = old index
// ---------------------------------------------------------------------------
// Also builds the order in which the collide workgroups take their tiles:
// heavy tiles first (they run longest), the others after.  Tiles are
// independent, so the order does not change any result.  heavy[] comes from
// the cell scan and is cleared here.
__device__ __forceinline__ void scatter_body(const int32_t* __restrict__ cell, DevScalars* __restrict__ S,
                                             const int32_t* __restrict__ start, const int32_t* __restrict__ slot,
                                             int32_t* __restrict__ perm, int32_t* __restrict__ heavy, int pt,
                                             int32_t* __restrict__ order, const int bx, const int gx) {
    const int64_t n = S->n;  // pre-sort population (the cell scan writes n_sorted)
    const int nt = (int)((int64_t)S->n_sorted / pt) + 1;
    for (int q = bx * BLOCK + threadIdx.x; q < nt; q += gx * BLOCK) {
        if (heavy[q]) {
            heavy[q] = 0;
            order[atomicAdd(&S->ord[0], 1)] = q;
        } else {
            order[nt - 1 - atomicAdd(&S->ord[1], 1)] = q;
        }
    }
    for (int64_t i = bx * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)gx * BLOCK) {
        const int32_t c = cell[i];
        if (c >= 0) perm[start[c] + slot[i]] = (int32_t)i;
    }
}

__global__ __launch_bounds__(BLOCK) void k_scatter(const int32_t* __restrict__ cell, DevScalars* __restrict__ S,
                                                   const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ slot, int32_t* __restrict__ perm,
                                                   int32_t* __restrict__ heavy, int pt, int32_t* __restrict__ order) {
    scatter_body(cell, S, start, slot, perm, heavy, pt, order, (int)blockIdx.x, (int)gridDim.x);
}

void launch_scatter(DevParticles b, DevScalars* S, const int32_t* start, const int32_t* slot, int32_t* perm,
                    int32_t* heavy, int pt, int32_t* order, int grid, hipStream_t st) {
    k_scatter<<<grid, BLOCK, 0, st>>>(b.cell, S, start, slot, perm, heavy, pt, order);
}

// ---------------------------------------------------------------------------
// collide: stable gather into cell order, NTC, binary collisions, sampling
// (noTimeCounter.C:81-339, dsmcVolFields.C:1115-1239)
// ---------------------------------------------------------------------------
#define CT 256           // collide workgroup: 4 waves
#define CT_CELLS 256     // cells per sub-tile: one thread per cell
#define DENSE_CAND 32    // cells with more candidates use the wave-cooperative NTC

struct CellCtx {
    unsigned long long cand, colls, ovf;
};

// Position of the k-th (from 0) set bit of m.
__device__ __forceinline__ int select_bit(unsigned long long m, int k) {
    int pos = 0;
    const int c0 = __popc((uint32_t)m);
    if (k >= c0) { k -= c0; m >>= 32; pos += 32; }
    uint32_t w = (uint32_t)m;
    int c = __popc(w & 0xFFFFu);
    if (k >= c) { k -= c; w >>= 16; pos += 16; }
    c = __popc(w & 0xFFu);
    if (k >= c) { k -= c; w >>= 8; pos += 8; }
    c = __popc(w & 0xFu);
    if (k >= c) { k -= c; w >>= 4; pos += 4; }
    c = __popc(w & 0x3u);
    if (k >= c) { k -= c; w >>= 2; pos += 2; }
    c = (int)(w & 1u);
    if (k >= c) pos += 1;
    return pos;
}

// The members of each sub-cell of one cell, in ascending index order: cnt(q)
// and member(q, k), the k-th member's index in the cell.
// CellTab: the LDS table of a cell -- for at most 64 parcels, a membership
// bit mask per sub-cell; for more, (offset | count << 16) into the member list.
struct CellTab {
    const unsigned long long* m;
    const uint16_t* slist;
    int ss;
    bool small;
    __device__ __forceinline__ int cnt(int q) const {
        return small ? __popcll(m[q]) : (int)((m[q] >> 16) & 0xFFFFull);
    }
    __device__ __forceinline__ int member(int q, int k) const {
        return small ? select_bit(m[q], k) : (int)slist[ss + (int)(m[q] & 0xFFFFull) + k];
    }
};

// CellTab32: the table of a cell of any size in global memory: per sub-cell
// (offset | count << 32) into a 32-bit member list.
struct CellTab32 {
    const unsigned long long* m;
    const int32_t* slist;
    __device__ __forceinline__ int cnt(int q) const { return (int)(m[q] >> 32); }
    __device__ __forceinline__ int member(int q, int k) const { return slist[(int)(m[q] & 0xFFFFFFFFull) + k]; }
};

// Candidate count of noTimeCounter.C:142-153 from the cell's sigmaTcRMax,
// remainder and volume; stores the new remainder.
__device__ __forceinline__ long long ntc_candidates_of(const DevParams& P, int c, int nC, double smax, double rem0,
                                                      double v, double* rem, CellCtx& cx) {
    const double sel = rem0 + 0.5 * nC * (nC - 1) * P.fn * smax * P.dt / v;
    long long ncand = (long long)sel;
    rem[c] = sel - (double)ncand;
    if (ncand > HS_MAX_CAND) { ncand = HS_MAX_CAND; cx.ovf += 1; }
    return ncand;
}

__device__ __forceinline__ long long ntc_candidates(const DevParams& P, int c, int nC, const double* sigma,
                                                   double* rem, const double* vol, double* smax_out, CellCtx& cx) {
    const double smax = sigma[c];
    *smax_out = smax;
    return ntc_candidates_of(P, c, nC, smax, rem[c], vol[c], rem, cx);
}

// Draws the pair of candidate k (its own stream) and reports whether it may
// collide: noTimeCounter.C:157-228.
template <bool GEN, class Tab>
__device__ __forceinline__ bool draw_pair(const DevParams& P, PhiloxStream& r, int nC, const uint8_t* sub, int ss,
                                          const Tab& tab, const uint8_t* ty, int* pp_out, int* qq_out) {
    const int pp = random_label(r, 0, nC - 1);
    const int q = sub[ss + pp];
    const int nsc = tab.cnt(q);
    int qq = -1, guard = 0;
    if (nsc > 1) {
        do { qq = tab.member(q, random_label(r, 0, nsc - 1)); } while (pp == qq && ++guard < HS_MAX_REDRAW);
    } else {
        do { qq = random_label(r, 0, nC - 1); } while (pp == qq && ++guard < HS_MAX_REDRAW);
    }
    *pp_out = pp;
    *qq_out = qq;
    if (pp == qq) return false;
    const int tp = GEN ? (int)ty[ss + pp] : 0, tq = GEN ? (int)ty[ss + qq] : 0;
    return !(P.charge[tp] == -1 && P.charge[tq] == -1);
}

// The accepted collision of candidate (pp, qq) in cell c:
// BinaryCollisionModel::collide.
template <bool GEN>
__device__ __forceinline__ void do_collision(const DevParams& P, PhiloxStream& r, int ip, int iq, double up[3],
                                             double uq[3], float* vx, float* vy, float* vz, float* er, uint16_t* vb,
                                             int vstride, const uint8_t* ty, int c) {
    const int tp = GEN ? (int)ty[ip] : 0, tq = GEN ? (int)ty[iq] : 0;
    double erp = 0.0, erq = 0.0;
    int vp[HS_MAX_VIB_MODES] = { 0, 0, 0 }, vq[HS_MAX_VIB_MODES] = { 0, 0, 0 };
    if (GEN) {
        if (P.has_erot) { erp = (double)er[ip]; erq = (double)er[iq]; }
        for (int m = 0; m < P.nvib[tp]; ++m) vp[m] = vb[m * vstride + ip];
        for (int m = 0; m < P.nvib[tq]; ++m) vq[m] = vb[m * vstride + iq];
    }
    binary_collide<GEN>(P, tp, tq, up, uq, &erp, &erq, vp, vq, r, c);
    vx[ip] = (float)up[0]; vy[ip] = (float)up[1]; vz[ip] = (float)up[2];
    vx[iq] = (float)uq[0]; vy[iq] = (float)uq[1]; vz[iq] = (float)uq[2];
    if (GEN) {
        if (P.has_erot) { er[ip] = (float)erp; er[iq] = (float)erq; }
        for (int m = 0; m < P.nvib[tp]; ++m) vb[m * vstride + ip] = (uint16_t)vp[m];
        for (int m = 0; m < P.nvib[tq]; ++m) vb[m * vstride + iq] = (uint16_t)vq[m];
    }
}

// Where a reacting collide stages the products it makes (see DevProducts).
struct ProdOut {
    DevScalars* S;
    DevProducts pr;
};

// Stages a dissociation product: cell c, its rank among the cell's products
// of this step (candidate order), the sorted index of the parcel that split.
__device__ __forceinline__ void stage_product(const DevParams& P, const ProdOut& po, int c, int rank,
                                              int64_t parent, const PState& prod) {
    const int slot = atomicAdd(&po.S->nprod, 1);
    if (slot >= P.capacity) return;  // more than the capacity: dropped, counted at the commit
    po.pr.cell[slot] = c;
    po.pr.rank[slot] = rank;
    po.pr.parent[slot] = (int32_t)parent;
    po.pr.type[slot] = (uint8_t)prod.t;
    po.pr.ux[slot] = (float)prod.u[0];
    po.pr.uy[slot] = (float)prod.u[1];
    po.pr.uz[slot] = (float)prod.u[2];
}

// The accepted collision of candidate (ip, iq) with chemistry: its QK
// reaction when the pair has a model, else the binary collision (noTimeCounter.C:
// 250-300).  After a reaction the types and every vibrational plane of both
// parcels are rewritten.  prod.t >= 0: a product was made; from_q: Q split.
__device__ __forceinline__ void react_collision(const DevParams& P, PhiloxStream& r, int ip, int iq, double up[3],
                                                double uq[3], float* vx, float* vy, float* vz, float* er,
                                                uint16_t* vb, int vstride, uint8_t* ty, PState& prod, int* from_q,
                                                int c) {
    PState a, b;
    a.t = ty[ip];
    b.t = ty[iq];
#pragma unroll
    for (int c = 0; c < 3; ++c) { a.u[c] = up[c]; b.u[c] = uq[c]; }
    a.er = er ? (double)er[ip] : 0.0;
    b.er = er ? (double)er[iq] : 0.0;
#pragma unroll
    for (int m = 0; m < HS_MAX_VIB_MODES; ++m) { a.v[m] = 0; b.v[m] = 0; }
    for (int m = 0; m < P.nvib[a.t]; ++m) a.v[m] = vb[m * vstride + ip];
    for (int m = 0; m < P.nvib[b.t]; ++m) b.v[m] = vb[m * vstride + iq];
    const bool reacted = collide_or_react(P, a, b, r, prod, from_q, c);
    vx[ip] = (float)a.u[0]; vy[ip] = (float)a.u[1]; vz[ip] = (float)a.u[2];
    vx[iq] = (float)b.u[0]; vy[iq] = (float)b.u[1]; vz[iq] = (float)b.u[2];
    if (er) { er[ip] = (float)a.er; er[iq] = (float)b.er; }
    const int np = reacted ? P.max_modes : P.nvib[a.t];
    const int nq = reacted ? P.max_modes : P.nvib[b.t];
    for (int m = 0; m < np; ++m) vb[m * vstride + ip] = (uint16_t)a.v[m];
    for (int m = 0; m < nq; ++m) vb[m * vstride + iq] = (uint16_t)b.v[m];
    if (reacted) { ty[ip] = (uint8_t)a.t; ty[iq] = (uint8_t)b.t; }
}

// NTC of one dense cell by a whole wave, 64 candidates per batch.  A
// candidate's pair comes from its own stream alone, so each lane draws its
// pair once; only the acceptance test and the collision need the velocities.
// A candidate is evaluated as soon as every earlier candidate of the batch
// that shares one of its parcels has been: each round, every pending lane
// stamps its two parcels (LDS atomicMin of the lane id), and a lane that is
// the earliest pending claimant of both is ready.  Ready lanes share no
// parcel, so they run together, and each sees exactly the velocities the
// candidate-by-candidate loop (noTimeCounter.C:157-302) gives it.  The rounds
// per batch are the depth of its parcel-sharing chains.  The sub-cell lists
// and tables are built and the stamps cleared beforehand; they are clear
// again on return.
// REACT: with chemistry; each batch stages its products with their ranks in
// candidate order (a prefix over the batch's lanes), counted in *nprod_out.
// FENCE: the parcel arrays are in global memory; a workgroup-scope fence after
// every round makes one round's writes visible to the next round's lanes.
// The stamp of parcel x is stamp[(ss + x) & smask]: with a mask smaller than
// the cell, parcels share stamps, which only holds a candidate back a round
// more (the earliest pending candidate is always ready, and none runs before
// an earlier one that shares a parcel with it).
template <bool GEN, bool REACT = false, bool FENCE = false, class Tab = CellTab>
__device__ void ntc_wave(const DevParams& P, int c, int ss, int nC, long long ncand, double smax, float* vx, float* vy,
                         float* vz, const uint8_t* sub, const Tab& tab, float* er,
                         uint16_t* vb, int vstride, const uint8_t* ty, int32_t* stamp, double* sigma, uint32_t step,
                         const DevParticles& pos, int64_t gs, CellCtx& cx, uint32_t& ncoll_out, double& sep_out,
                         const ProdOut* po = nullptr, int* nprod_out = nullptr, int smask = 0x7FFFFFFF) {
    int nprod = 0;
    const int lane = threadIdx.x & (WAVE - 1);
    const bool sep_on = (P.flags & HS_FLAG_COLL_SEP) != 0;
    double smax_cur = smax;
    double sep = 0.0;
    unsigned long long ncoll_i = 0;
    for (long long k0 = 0; k0 < ncand; k0 += WAVE) {
        const long long k = k0 + lane;
        int pp = -1, qq = -1;
        bool pend = false;
        PhiloxStream r;
        if (k < ncand) {
            r.init(P, RNG_COLLIDE, (uint32_t)c, step, (uint32_t)k);
            pend = draw_pair<GEN>(P, r, nC, sub, ss, tab, ty, &pp, &qq);
        }
        bool doit = false;
        double sm = 0.0, sep_l = 0.0;
        PState prod;
        prod.t = -1;
        int fq = 0;
        while (__ballot(pend)) {
            if (pend) {
                atomicMin(&stamp[(ss + pp) & smask], lane);
                atomicMin(&stamp[(ss + qq) & smask], lane);
            }
            __builtin_amdgcn_wave_barrier();
            const bool ready = pend && stamp[(ss + pp) & smask] == lane && stamp[(ss + qq) & smask] == lane;
            __builtin_amdgcn_wave_barrier();
            if (pend) {
                stamp[(ss + pp) & smask] = 0x7FFFFFFF;
                stamp[(ss + qq) & smask] = 0x7FFFFFFF;
            }
            if (ready) {
                const int ip = ss + pp, iq = ss + qq;
                const int tp = GEN ? (int)ty[ip] : 0, tq = GEN ? (int)ty[iq] : 0;
                double up[3] = { (double)vx[ip], (double)vy[ip], (double)vz[ip] };
                double uq[3] = { (double)vx[iq], (double)vy[iq], (double)vz[iq] };
                const double stcr = sigma_tcr(P, tp, tq, up, uq);
                sm = stcr;
                if (stcr / smax > r.u()) {
                    if (REACT)
                        react_collision(P, r, ip, iq, up, uq, vx, vy, vz, er, vb, vstride, const_cast<uint8_t*>(ty),
                                        prod, &fq, c);
                    else
                        do_collision<GEN>(P, r, ip, iq, up, uq, vx, vy, vz, er, vb, vstride, ty, c);
                    doit = true;
                    if (sep_on) {
                        const int64_t gp = gs + pp, gq = gs + qq;
                        const double ddx = pos.px[gp] - pos.px[gq], ddy = pos.py[gp] - pos.py[gq],
                                     ddz = pos.pz[gp] - pos.pz[gq];
                        sep_l = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                    }
                }
                pend = false;
            }
            if (FENCE) __threadfence_block();
            __builtin_amdgcn_wave_barrier();
        }
        if (REACT) {
            const unsigned long long pm = __ballot(prod.t >= 0);
            if (prod.t >= 0)
                stage_product(P, *po, c, nprod + __popcll(pm & ((1ull << lane) - 1ull)), gs + (fq ? qq : pp), prod);
            nprod += __popcll(pm);
        }
        // running maximum of sigmaTcR over the batch
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sm = fmax(sm, __shfl_xor(sm, o, WAVE));
        if (sm > smax_cur) smax_cur = sm;
        const unsigned long long dm = __ballot(doit);
        ncoll_i += (unsigned long long)__popcll(dm);
        if (sep_on) {
            // add in candidate order, as the sequential loop does
            for (int j = 0; j < WAVE; ++j) {
                const double v = __shfl(sep_l, j, WAVE);
                if ((dm >> j) & 1ull) sep += v;
            }
        }
    }
    if (lane == 0) {
        sigma[c] = smax_cur;
        cx.cand += (unsigned long long)ncand;
        cx.colls += ncoll_i;
    }
    ncoll_out = (uint32_t)ncoll_i;
    sep_out = sep;
    if (REACT) *nprod_out = nprod;
}

// dsmcVolFields::calculateField per-cell sums for one cell and one species
// s: the species' parcels of the cell in order (dsmcVolFields.C:1115-1239),
// by one thread.  The kernel variant with internal energies samples its
// cells of at most 64 parcels this way, one thread per (cell, species), so a
// workgroup's threads share the cells' species.  Each sum has one writer;
// the cumulative sums take non-returning atomic adds (the same IEEE addition
// as a read-add-write, without waiting for the read).
template <bool GEN>
__device__ __forceinline__ void sample_cell_species(const DevParams& P, int c, int ss, int nC, int s, const float* vx,
                                                    const float* vy, const float* vz, const float* er,
                                                    const uint16_t* vb, int vstride, const uint8_t* ty,
                                                    double* samples, int nc) {
    const bool hf = (P.flags & HS_FLAG_HEATFLUX) != 0;
    const double m = P.mass[s];
    double aN = 0.0, aX = 0.0, aY = 0.0, aZ = 0.0, aC = 0.0, aR = 0.0;
    double ev[HS_MAX_VIB_MODES] = { 0.0, 0.0, 0.0 };
    const int nvib = GEN ? P.nvib[s] : 0;
#pragma unroll 4
    for (int k = 0; k < nC; ++k) {
        const int i = ss + k;
        if (GEN && (int)ty[i] != s) continue;
        const double ux = (double)vx[i], uy = (double)vy[i], uz = (double)vz[i];
        aN += 1.0;
        aX += m * ux;
        aY += m * uy;
        aZ += m * uz;
        aC += m * (ux * ux + uy * uy + uz * uz);
        if (GEN && P.has_erot) aR += (double)er[i];
        for (int mo = 0; mo < nvib; ++mo) ev[mo] += (double)vb[mo * vstride + i] * HS_KB * P.theta_v[s][mo];
    }
    // no parcel of the species: its sums are +0, and adding +0 changes no
    // cumulative sum (they start at +0 and never become -0)
    if (aN == 0.0) return;
    const int64_t base = HS_S_GLOBAL_COUNT + (int64_t)s * HS_SS_COUNT;
    unsafeAtomicAdd(&samples[(base + HS_SS_N) * nc + c], aN);
    unsafeAtomicAdd(&samples[(base + HS_SS_MUX) * nc + c], aX);
    unsafeAtomicAdd(&samples[(base + HS_SS_MUY) * nc + c], aY);
    unsafeAtomicAdd(&samples[(base + HS_SS_MUZ) * nc + c], aZ);
    unsafeAtomicAdd(&samples[(base + HS_SS_MCC) * nc + c], aC);
    if (GEN && P.has_erot) unsafeAtomicAdd(&samples[(base + HS_SS_EROT) * nc + c], aR);
    for (int mo = 0; mo < nvib; ++mo) unsafeAtomicAdd(&samples[(base + HS_SS_EVIB0 + mo) * nc + c], ev[mo]);
    if (hf) {
        double h[HS_HF_COUNT];
        for (int q = 0; q < HS_HF_COUNT; ++q) h[q] = 0.0;
        for (int k = 0; k < nC; ++k) {
            const int i = ss + k;
            if (GEN && (int)ty[i] != s) continue;
            const double ux = (double)vx[i], uy = (double)vy[i], uz = (double)vz[i];
            const double lke = m * (ux * ux + uy * uy + uz * uz);
            const double erv = (GEN && P.has_erot) ? (double)er[i] : 0.0;
            double evs = 0.0;
            for (int mo = 0; mo < nvib; ++mo) evs += (double)vb[mo * vstride + i] * HS_KB * P.theta_v[s][mo];
            const double eint = erv + evs;
            h[0] += m * (ux * ux); h[1] += m * ux * uy; h[2] += m * ux * uz;
            h[3] += m * (uy * uy); h[4] += m * uy * uz; h[5] += m * (uz * uz);
            h[6] += lke * ux; h[7] += lke * uy; h[8] += lke * uz;
            h[9] += eint * ux; h[10] += eint * uy; h[11] += eint * uz; h[12] += eint;
        }
        const int64_t hb = HS_S_GLOBAL_COUNT + (int64_t)P.ns * HS_SS_COUNT + (int64_t)s * HS_HF_COUNT;
        for (int q = 0; q < HS_HF_COUNT; ++q) unsafeAtomicAdd(&samples[(hb + q) * nc + c], h[q]);
    }
}

// The same sums (the kernel variant without internal energies) for a cell of
// at most 64 parcels by a group of four consecutive lanes (r = lane mod 4):
// lane r adds the terms of parcels
// r, r+4, r+8, ... in order, and the partial sums are combined as
// (p0 + p2) + (p1 + p3) (every lane of the group ends with that value).  The
// oracle forms these sums in the same order.  Lane r = 0 adds them, with the
// cell's collision count and separation, to the cumulative sums; each cell
// has one writer, and the adds are non-returning atomics (the same IEEE
// addition as a read-add-write, without waiting for the read).
__device__ __forceinline__ double quad_sum(double v) {
    v = v + __shfl_xor(v, 2, WAVE);
    return v + __shfl_xor(v, 1, WAVE);
}

template <bool GEN>
__device__ __forceinline__ void sample_cell_quad(const DevParams& P, int c, int ss, int nC, int r, const float* vx,
                                                 const float* vy, const float* vz, const float* er,
                                                 const uint16_t* vb, int vstride, const uint8_t* ty, double* samples,
                                                 int nc, double ncoll, double sep) {
    if (r == 0 && ncoll != 0.0) {
        unsafeAtomicAdd(&samples[(int64_t)HS_S_NCOLL * nc + c], ncoll);
        unsafeAtomicAdd(&samples[(int64_t)HS_S_COLLSEP * nc + c], sep);
    }
    const int ns = GEN ? P.ns : 1;
    const bool hf = (P.flags & HS_FLAG_HEATFLUX) != 0;
    for (int s = 0; s < ns; ++s) {
        const double m = P.mass[s];
        double aN = 0.0, aX = 0.0, aY = 0.0, aZ = 0.0, aC = 0.0, aR = 0.0;
        double ev[HS_MAX_VIB_MODES] = { 0.0, 0.0, 0.0 };
        const int nvib = GEN ? P.nvib[s] : 0;
        for (int k = r; k < nC; k += 4) {
            const int i = ss + k;
            if (GEN && (int)ty[i] != s) continue;
            const double ux = (double)vx[i], uy = (double)vy[i], uz = (double)vz[i];
            aN += 1.0;
            aX += m * ux;
            aY += m * uy;
            aZ += m * uz;
            aC += m * (ux * ux + uy * uy + uz * uz);
            if (GEN && P.has_erot) aR += (double)er[i];
            for (int mo = 0; mo < nvib; ++mo) ev[mo] += (double)vb[mo * vstride + i] * HS_KB * P.theta_v[s][mo];
        }
        aN = quad_sum(aN);
        aX = quad_sum(aX);
        aY = quad_sum(aY);
        aZ = quad_sum(aZ);
        aC = quad_sum(aC);
        if (GEN && P.has_erot) aR = quad_sum(aR);
        for (int mo = 0; mo < nvib; ++mo) ev[mo] = quad_sum(ev[mo]);
        const int64_t base = HS_S_GLOBAL_COUNT + (int64_t)s * HS_SS_COUNT;
        if (r == 0 && aN != 0.0) {
            unsafeAtomicAdd(&samples[(base + HS_SS_N) * nc + c], aN);
            unsafeAtomicAdd(&samples[(base + HS_SS_MUX) * nc + c], aX);
            unsafeAtomicAdd(&samples[(base + HS_SS_MUY) * nc + c], aY);
            unsafeAtomicAdd(&samples[(base + HS_SS_MUZ) * nc + c], aZ);
            unsafeAtomicAdd(&samples[(base + HS_SS_MCC) * nc + c], aC);
            if (GEN && P.has_erot) unsafeAtomicAdd(&samples[(base + HS_SS_EROT) * nc + c], aR);
            for (int mo = 0; mo < nvib; ++mo) unsafeAtomicAdd(&samples[(base + HS_SS_EVIB0 + mo) * nc + c], ev[mo]);
        }
        if (hf) {
            double h[HS_HF_COUNT];
            for (int q = 0; q < HS_HF_COUNT; ++q) h[q] = 0.0;
            for (int k = r; k < nC; k += 4) {
                const int i = ss + k;
                if (GEN && (int)ty[i] != s) continue;
                const double ux = (double)vx[i], uy = (double)vy[i], uz = (double)vz[i];
                const double lke = m * (ux * ux + uy * uy + uz * uz);
                const double erv = (GEN && P.has_erot) ? (double)er[i] : 0.0;
                double evs = 0.0;
                for (int mo = 0; mo < nvib; ++mo) evs += (double)vb[mo * vstride + i] * HS_KB * P.theta_v[s][mo];
                const double eint = erv + evs;
                h[0] += m * (ux * ux); h[1] += m * ux * uy; h[2] += m * ux * uz;
                h[3] += m * (uy * uy); h[4] += m * uy * uz; h[5] += m * (uz * uz);
                h[6] += lke * ux; h[7] += lke * uy; h[8] += lke * uz;
                h[9] += eint * ux; h[10] += eint * uy; h[11] += eint * uz; h[12] += eint;
            }
            for (int q = 0; q < HS_HF_COUNT; ++q) h[q] = quad_sum(h[q]);
            if (r == 0 && aN != 0.0) {
                const int64_t hb = HS_S_GLOBAL_COUNT + (int64_t)P.ns * HS_SS_COUNT + (int64_t)s * HS_HF_COUNT;
                for (int q = 0; q < HS_HF_COUNT; ++q) unsafeAtomicAdd(&samples[(hb + q) * nc + c], h[q]);
            }
        }
    }
}

// The same sums for a cell of more than 64 parcels, by one wave: lane l adds
// the terms of parcels l, l+64, l+128, ... in order, and the 64 partial sums
// are combined pairwise across lane distances 32, 16, 8, 4, 2, 1 (every lane
// ends with the same value: IEEE addition is commutative).  The oracle forms
// these sums in the same order.  nColls and the separation are not added.
__device__ __forceinline__ double wave_tree_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, WAVE);
    return v;
}

// b2, n2: a second range of the cell's items after its nC parcels (the
// products it made this step), taken in the same lane order.
template <bool GEN>
__device__ void sample_cell_wave(const DevParams& P, int c, int ss, int nC, const float* vx, const float* vy,
                                 const float* vz, const float* er, const uint16_t* vb, int vstride, const uint8_t* ty,
                                 double* samples, int nc, int b2 = 0, int n2 = 0) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int ns = GEN ? P.ns : 1;
    const bool hf = (P.flags & HS_FLAG_HEATFLUX) != 0;
    for (int s = 0; s < ns; ++s) {
        const double m = P.mass[s];
        double aN = 0.0, aX = 0.0, aY = 0.0, aZ = 0.0, aC = 0.0, aR = 0.0;
        double ev[HS_MAX_VIB_MODES] = { 0.0, 0.0, 0.0 };
        double h[HS_HF_COUNT];
        for (int q = 0; q < HS_HF_COUNT; ++q) h[q] = 0.0;
        const int nvib = GEN ? P.nvib[s] : 0;
        for (int k = lane; k < nC + n2; k += WAVE) {
            const int i = k < nC ? ss + k : b2 + (k - nC);
            if (GEN && (int)ty[i] != s) continue;
            const double ux = (double)vx[i], uy = (double)vy[i], uz = (double)vz[i];
            const double lke = m * (ux * ux + uy * uy + uz * uz);
            aN += 1.0;
            aX += m * ux;
            aY += m * uy;
            aZ += m * uz;
            aC += lke;
            const double erv = (GEN && P.has_erot) ? (double)er[i] : 0.0;
            if (GEN && P.has_erot) aR += erv;
            double evs = 0.0;
            for (int mo = 0; mo < nvib; ++mo) {
                const double e = (double)vb[mo * vstride + i] * HS_KB * P.theta_v[s][mo];
                ev[mo] += e;
                evs += e;
            }
            if (hf) {
                const double eint = erv + evs;
                h[0] += m * (ux * ux); h[1] += m * ux * uy; h[2] += m * ux * uz;
                h[3] += m * (uy * uy); h[4] += m * uy * uz; h[5] += m * (uz * uz);
                h[6] += lke * ux; h[7] += lke * uy; h[8] += lke * uz;
                h[9] += eint * ux; h[10] += eint * uy; h[11] += eint * uz; h[12] += eint;
            }
        }
        aN = wave_tree_sum(aN);
        if (aN == 0.0) continue;  // no parcel of the species: adding zeros changes no sum
        aX = wave_tree_sum(aX);
        aY = wave_tree_sum(aY);
        aZ = wave_tree_sum(aZ);
        aC = wave_tree_sum(aC);
        if (GEN && P.has_erot) aR = wave_tree_sum(aR);
        for (int mo = 0; mo < nvib; ++mo) ev[mo] = wave_tree_sum(ev[mo]);
        if (hf)
            for (int q = 0; q < HS_HF_COUNT; ++q) h[q] = wave_tree_sum(h[q]);
        if (lane == 0) {
            const int64_t base = HS_S_GLOBAL_COUNT + (int64_t)s * HS_SS_COUNT;
            unsafeAtomicAdd(&samples[(base + HS_SS_N) * nc + c], aN);
            unsafeAtomicAdd(&samples[(base + HS_SS_MUX) * nc + c], aX);
            unsafeAtomicAdd(&samples[(base + HS_SS_MUY) * nc + c], aY);
            unsafeAtomicAdd(&samples[(base + HS_SS_MUZ) * nc + c], aZ);
            unsafeAtomicAdd(&samples[(base + HS_SS_MCC) * nc + c], aC);
            if (GEN && P.has_erot) unsafeAtomicAdd(&samples[(base + HS_SS_EROT) * nc + c], aR);
            for (int mo = 0; mo < nvib; ++mo) unsafeAtomicAdd(&samples[(base + HS_SS_EVIB0 + mo) * nc + c], ev[mo]);
            if (hf) {
                const int64_t hb = HS_S_GLOBAL_COUNT + (int64_t)P.ns * HS_SS_COUNT + (int64_t)s * HS_HF_COUNT;
                for (int q = 0; q < HS_HF_COUNT; ++q) unsafeAtomicAdd(&samples[(hb + q) * nc + c], h[q]);
            }
        }
    }
}

__device__ __forceinline__ int subcell_of(const DevParams& P, int c, double x, double y, double z) {
    int ijk[3];
    cell_ijk(P, c, ijk);
    const int ci = ijk[0], cj = ijk[1], ck = ijk[2];
    const double cc0 = P.origin[0] + ((double)ci + 0.5) * P.delta[0];
    const double cc1 = P.origin[1] + ((double)cj + 0.5) * P.delta[1];
    const double cc2 = P.origin[2] + ((double)ck + 0.5) * P.delta[2];
    return (x - cc0 >= 0.0 ? 1 : 0) + 2 * (y - cc1 >= 0.0 ? 1 : 0) + 4 * (z - cc2 >= 0.0 ? 1 : 0);
}

// Rank of key k among key[ss..ee) (LDS, 16-byte aligned array, reads may run
// up to 3 entries past ee inside the array): 4 keys per 128-bit read.  Only
// the first and last 4-key windows can hold keys outside [ss, ee); the
// windows between are compared without bounds tests.
__device__ __forceinline__ int rank_in(const int32_t* key, int ss, int ee, int k) {
    const int q0 = ss & ~3, q1 = (ee - 1) & ~3;  // first and last window
    int r = 0;
    {
        const int4 v = *reinterpret_cast<const int4*>(key + q0);
        r += (q0 >= ss && q0 < ee && v.x < k) ? 1 : 0;
        r += (q0 + 1 >= ss && q0 + 1 < ee && v.y < k) ? 1 : 0;
        r += (q0 + 2 >= ss && q0 + 2 < ee && v.z < k) ? 1 : 0;
        r += (q0 + 3 >= ss && q0 + 3 < ee && v.w < k) ? 1 : 0;
    }
#pragma unroll 4
    for (int q = q0 + 4; q < q1; q += 4) {
        const int4 v = *reinterpret_cast<const int4*>(key + q);
        r += (v.x < k ? 1 : 0) + (v.y < k ? 1 : 0) + (v.z < k ? 1 : 0) + (v.w < k ? 1 : 0);
    }
    if (q1 > q0) {
        const int4 v = *reinterpret_cast<const int4*>(key + q1);
        r += (v.x < k) ? 1 : 0;
        r += (q1 + 1 < ee && v.y < k) ? 1 : 0;
        r += (q1 + 2 < ee && v.z < k) ? 1 : 0;
        r += (q1 + 3 < ee && v.w < k) ? 1 : 0;
    }
    return r;
}

// Sub-cell member lists of one cell (ss, nC) by a whole wave: counts and
// ascending fills from ballots over 64-parcel windows; tab[q] gets
// (offset | count << 16) of sub-cell q.
template <class IDX = uint16_t>
__device__ __forceinline__ void subcell_lists_wave(const uint8_t* sub, IDX* slist, int ss, int nC,
                                                   unsigned long long* tab) {
    // 16-bit lists: (offset | count << 16); 32-bit: (offset | count << 32)
    const int shift = sizeof(IDX) == 2 ? 16 : 32;
    const int lane = threadIdx.x & (WAVE - 1);
    const unsigned long long lt = lane == 0 ? 0ull : ((1ull << lane) - 1ull);
    // a window's lanes of sub-cell qq, from three ballots of the sub-cell's bits
    auto members = [](int qq, unsigned long long v, unsigned long long b0, unsigned long long b1,
                      unsigned long long b2) {
        return v & ((qq & 1) ? b0 : ~b0) & ((qq & 2) ? b1 : ~b1) & ((qq & 4) ? b2 : ~b2);
    };
    int cnt[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    for (int base = 0; base < nC; base += WAVE) {
        const int j = base + lane;
        const int q = j < nC ? (int)sub[ss + j] : 0;
        const unsigned long long v = __ballot(j < nC);
        const unsigned long long b0 = __ballot(q & 1), b1 = __ballot(q & 2), b2 = __ballot(q & 4);
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) cnt[qq] += __popcll(members(qq, v, b0, b1, b2));
    }
    int run[8];
    {
        int acc = 0;
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) {
            run[qq] = acc;
            if (lane == qq) tab[qq] = (unsigned long long)acc | ((unsigned long long)cnt[qq] << shift);
            acc += cnt[qq];
        }
    }
    for (int base = 0; base < nC; base += WAVE) {
        const int j = base + lane;
        const int q = j < nC ? (int)sub[ss + j] : 0;
        const unsigned long long v = __ballot(j < nC);
        const unsigned long long b0 = __ballot(q & 1), b1 = __ballot(q & 2), b2 = __ballot(q & 4);
        int at = 0;
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) {
            const unsigned long long m = members(qq, v, b0, b1, b2);
            if (q == qq) at = run[qq] + __popcll(m & lt);
            run[qq] += __popcll(m);
        }
        if (j < nC) slist[ss + at] = (IDX)j;
    }
}

template <bool GEN, int PMAX, bool REACT = false>
struct TileLds {
    alignas(16) int32_t key[PMAX];  // old index of each slot; then the NTC conflict stamps
    alignas(16) uint8_t loc[PMAX];  // sub-tile cell of each slot (filled 4 bytes at a time)
    alignas(16) uint8_t sub[PMAX];
    uint16_t slist[PMAX];    // cells of more than 64 parcels: their parcels by sub-cell, ascending
    float vx[PMAX], vy[PMAX], vz[PMAX];
    float er[GEN ? PMAX : 1];
    uint16_t vb[GEN ? HS_MAX_VIB_MODES * PMAX : 1];
    uint8_t ty[GEN ? PMAX : 1];
    int32_t segS[CT_CELLS + 1];
    unsigned long long smask[CT_CELLS][8];  // sub-cell tables (CellTab)
    double sep[CT_CELLS];
    double smax[CT_CELLS];   // sigmaTcRMax at the start of the step
    int32_t ncand[CT_CELLS];
    int32_t dense[CT_CELLS];
    int32_t ndense;
    int32_t dnext;           // next dense cell a wave takes
    int32_t big[CT_CELLS];   // cells of more than 64 parcels
    int32_t nbig;
    // candidate rounds of the sparse cells
    int32_t rn[CT_CELLS];    // candidates handled by the rounds (0 for dense cells)
    int32_t rpfx[CT_CELLS];  // first slot of each cell
    unsigned long long rsmax[CT_CELLS];  // running max of sigmaTcR (bits of a non-negative double)
    uint32_t rcoll[CT_CELLS];
    double gsep[CT];         // collision separation of each slot of a pass
    uint8_t gacc[CT];        // the slot's candidate collided
    // chemistry: products made per cell, and each slot's count of earlier products in a pass
    int32_t pcount[REACT ? CT_CELLS : 1];
    int32_t gpre[REACT ? CT : 1];
};

// NTC of the sparse cells of a chunk (at most DENSE_CAND candidates each),
// candidate-parallel over the workgroup.  The candidates of all cells, in cell
// then candidate order, take the workgroup's slots CT at a time (a pass); each
// slot draws its pair once from the candidate's own stream.  Rounds then run
// as in ntc_wave: every pending slot stamps its two parcels (LDS atomicMin of
// the slot), a slot that is the earliest pending claimant of both is ready,
// and the ready slots -- which share no parcel -- evaluate and collide
// against exactly the velocities the sequential loop of noTimeCounter.C:157-302
// gives them.  A pass ends when all its slots are done; the next pass (the
// later candidates) starts after it.
// REACT: with chemistry; at the end of each pass every product is staged
// with its rank in its cell's candidate order (the cell's products of earlier
// passes plus a prefix over the pass's slots), and L.pcount counts them.
template <bool GEN, int PMAX, bool REACT = false>
__device__ void ntc_rounds(TileLds<GEN, PMAX, REACT>& L, const DevParams& P, int c0, int a, int b, int S0,
                           uint32_t step, const DevParticles& pos, CellCtx& cx, const ProdOut* po = nullptr) {
    const int t = threadIdx.x;
    const bool own = t >= a && t < b;
    const bool sep_on = (P.flags & HS_FLAG_COLL_SEP) != 0;
    int total;
    const int first = block_excl_scan(own ? L.rn[t] : 0, &total);
    if (own) L.rpfx[t] = first;
    __syncthreads();
    for (int g0 = 0; g0 < total; g0 += CT) {
        const int g = g0 + t;
        int l = a, ss = 0, pp = -1, qq = -1;
        bool pend = false, doit = false;
        double sep_l = 0.0;
        PhiloxStream r;
        PState prod;
        prod.t = -1;
        int fq = 0;
        if (g < total) {
            // the last cell whose first slot is <= g
            int lo = a, hi = b - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (L.rpfx[mid] <= g) lo = mid;
                else hi = mid - 1;
            }
            l = lo;
            const int k = g - L.rpfx[l];
            ss = L.segS[l] - S0;
            const int nC = L.segS[l + 1] - L.segS[l];
            r.init(P, RNG_COLLIDE, (uint32_t)(c0 + l), step, (uint32_t)k);
            const CellTab tab = { L.smask[l], L.slist, ss, nC <= 64 };
            pend = draw_pair<GEN>(P, r, nC, L.sub, ss, tab, L.ty, &pp, &qq);
        }
        for (;;) {
            if (pend) {
                atomicMin(&L.key[ss + pp], g);
                atomicMin(&L.key[ss + qq], g);
            }
            __syncthreads();
            const bool ready = pend && L.key[ss + pp] == g && L.key[ss + qq] == g;
            __syncthreads();
            if (pend) {
                L.key[ss + pp] = 0x7FFFFFFF;
                L.key[ss + qq] = 0x7FFFFFFF;
            }
            if (ready) {
                const int ip = ss + pp, iq = ss + qq;
                const int tp = GEN ? (int)L.ty[ip] : 0, tq = GEN ? (int)L.ty[iq] : 0;
                double up[3] = { (double)L.vx[ip], (double)L.vy[ip], (double)L.vz[ip] };
                double uq[3] = { (double)L.vx[iq], (double)L.vy[iq], (double)L.vz[iq] };
                const double stcr = sigma_tcr(P, tp, tq, up, uq);
                atomicMax(&L.rsmax[l], (unsigned long long)__double_as_longlong(stcr));
                if (stcr / L.smax[l] > r.u()) {
                    if (REACT)
                        react_collision(P, r, ip, iq, up, uq, L.vx, L.vy, L.vz, L.er, L.vb, PMAX, L.ty, prod, &fq,
                                        c0 + l);
                    else
                        do_collision<GEN>(P, r, ip, iq, up, uq, L.vx, L.vy, L.vz, L.er, L.vb, PMAX, L.ty, c0 + l);
                    atomicAdd(&L.rcoll[l], 1u);
                    doit = true;
                    if (sep_on) {
                        const int64_t gp = (int64_t)S0 + ip, gq = (int64_t)S0 + iq;
                        const double ddx = pos.px[gp] - pos.px[gq], ddy = pos.py[gp] - pos.py[gq],
                                     ddz = pos.pz[gp] - pos.pz[gq];
                        sep_l = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                    }
                }
                pend = false;
            }
            if (!__syncthreads_or(pend)) break;
        }
        if (REACT) {
            // the slots of one cell are consecutive and in candidate order
            const int made = prod.t >= 0 ? 1 : 0;
            int ptot;
            const int ex = block_excl_scan(made, &ptot);
            L.gpre[t] = ex;
            __syncthreads();
            if (made) {
                const int first = max(L.rpfx[l], g0) - g0;
                stage_product(P, *po, c0 + l, L.pcount[l] + ex - L.gpre[first], (int64_t)S0 + ss + (fq ? qq : pp),
                              prod);
            }
            __syncthreads();
            if (made) atomicAdd(&L.pcount[l], 1);
            __syncthreads();
        }
        if (sep_on) {
            // each cell adds its slots' separations in candidate order
            L.gsep[t] = sep_l;
            L.gacc[t] = doit ? 1 : 0;
            __syncthreads();
            if (own) {
                const int beg = max(L.rpfx[t], g0), end = min(L.rpfx[t] + L.rn[t], g0 + CT);
                for (int q = beg; q < end; ++q)
                    if (L.gacc[q - g0]) L.sep[t] += L.gsep[q - g0];
            }
            __syncthreads();
        }
    }
    (void)cx;
}

// ---------------------------------------------------------------------------
// k_gather: the stable cell order of buildCellOccupancy (dsmcCloud.C:63-74).
// Each parcel's rank among the parcels of its cell by old index gives its
// slot; the population is copied to the other buffer in that order together
// with each parcel's sub-cell byte (noTimeCounter.C:126-138).  Small LDS and
// few registers: many workgroups per CU hide the scattered reads.
// ---------------------------------------------------------------------------
#define PMAX_SORT 4096    // parcels per LDS chunk of the gather

struct GatherLds {
    alignas(16) int32_t key[PMAX_SORT];  // old index of each slot of the chunk
    uint8_t loc[PMAX_SORT];              // sub-tile cell of each slot
    int32_t segS[CT_CELLS + 1];
};

__device__ __forceinline__ void copy_internal(const DevParams& P, const DevParticles& src, DevParticles& dst, int k,
                                              int64_t gd) {
    const int64_t cap = P.capacity;
    if (dst.erot) dst.erot[gd] = src.erot[k];
    if (dst.vib)
        for (int m = 0; m < P.max_modes; ++m) dst.vib[m * cap + gd] = src.vib[m * cap + k];
    if (dst.type) dst.type[gd] = src.type[k];
}

// Ranks in a cell larger than a chunk (keys = its parcels' old indices, all
// distinct).  When the keys span at most 16 times the parcel count, a bitmap
// over the key range and the prefix of its words' bit counts go to the
// cell's own stretch of `scratch` (the histogram slots, free once the scatter
// has read them), and a key's rank is the prefix of its word plus the bits
// below it in the word.  Returns false (nothing built) otherwise.
struct BigRanks {
    const uint32_t* bm;
    const uint32_t* pre;
    int kmin;
    __device__ __forceinline__ int rank(int k) const {
        const int d = k - kmin;
        return (int)pre[d >> 5] + __popc(bm[d >> 5] & ((1u << (d & 31)) - 1u));
    }
};

__device__ __forceinline__ bool big_cell_ranks(const int32_t* __restrict__ perm, int32_t* __restrict__ scratch, int S0,
                                               int nC, BigRanks* out) {
    __shared__ int s_min, s_max;
    const int t = threadIdx.x;
    if (t == 0) { s_min = 0x7FFFFFFF; s_max = -1; }
    __syncthreads();
    int lo = 0x7FFFFFFF, hi = -1;
    for (int j = t; j < nC; j += CT) {
        const int k = perm[S0 + j];
        lo = min(lo, k);
        hi = max(hi, k);
    }
    atomicMin(&s_min, lo);
    atomicMax(&s_max, hi);
    __syncthreads();
    const int kmin = s_min;
    const int64_t range = (int64_t)s_max - kmin + 1;
    const int nw = (int)((range + 31) / 32);
    __syncthreads();  // every thread has read s_min, s_max
    if (2 * (int64_t)nw > nC) return false;
    uint32_t* bm = reinterpret_cast<uint32_t*>(scratch + S0);
    uint32_t* pre = bm + nw;
    for (int w = t; w < nw; w += CT) bm[w] = 0u;
    __threadfence_block();
    __syncthreads();
    for (int j = t; j < nC; j += CT) {
        const int d = perm[S0 + j] - kmin;
        atomicOr(&bm[d >> 5], 1u << (d & 31));
    }
    __threadfence_block();
    __syncthreads();
    int running = 0;
    for (int w0 = 0; w0 < nw; w0 += CT) {
        const int w = w0 + t;
        const int v = w < nw ? __popc(bm[w]) : 0;
        int tot;
        const int ex = block_excl_scan(v, &tot);
        if (w < nw) pre[w] = (uint32_t)(running + ex);
        running += tot;
    }
    __threadfence_block();
    __syncthreads();
    out->bm = bm;
    out->pre = pre;
    out->kmin = kmin;
    return true;
}

template <bool GEN>
__device__ void gather_cells(GatherLds& L, const DevParams& P, int c0, int ncell, const DevParticles& src,
                             DevParticles& dst, const int32_t* __restrict__ start, const int32_t* __restrict__ perm,
                             uint8_t* __restrict__ sub_out, int32_t* __restrict__ scratch) {
    const int t = threadIdx.x;
    if (t < ncell) L.segS[t] = start[c0 + t];
    if (t == 0) L.segS[ncell] = start[c0 + ncell];
    __syncthreads();
    int a = 0;
    while (a < ncell) {
        const int S0 = L.segS[a];
        const bool ok = t >= a && t < ncell && L.segS[t + 1] - S0 <= PMAX_SORT;
        const int b = a + __syncthreads_count(ok);
        if (b == a) {
            // one cell larger than a chunk: ranks from global memory
            const int nC = L.segS[a + 1] - S0;
            const int c = c0 + a;
            BigRanks br;
            const bool bitmap = big_cell_ranks(perm, scratch, S0, nC, &br);
            for (int j = t; j < nC; j += CT) {
                const int k = perm[S0 + j];
                int rank = 0;
                if (bitmap) {
                    rank = br.rank(k);
                } else {
                    for (int q = 0; q < nC; ++q) rank += (perm[S0 + q] < k) ? 1 : 0;
                }
                const int64_t gd = (int64_t)S0 + rank;
                const double x = src.px[k], y = src.py[k], z = src.pz[k];
                // positions and cells are read again only by the next step's
                // move: streamed past the caches, which keep what the collide
                // reads (velocities, sub-cell bytes)
                __builtin_nontemporal_store(x, &dst.px[gd]); __builtin_nontemporal_store(y, &dst.py[gd]);
                __builtin_nontemporal_store(z, &dst.pz[gd]);
                dst.ux[gd] = src.ux[k]; dst.uy[gd] = src.uy[k]; dst.uz[gd] = src.uz[k];
                __builtin_nontemporal_store(c, &dst.cell[gd]);
                if (GEN) copy_internal(P, src, dst, k, gd);
                sub_out[gd] = (uint8_t)subcell_of(P, c, x, y, z);
            }
            __syncthreads();
            a += 1;
            continue;
        }
        const int cnt = L.segS[b] - S0;
        if (t >= a && t < b)
            for (int j = L.segS[t] - S0; j < L.segS[t + 1] - S0; ++j) L.loc[j] = (uint8_t)t;
#pragma unroll 4
        for (int j = t; j < cnt; j += CT) L.key[j] = perm[S0 + j];
        __syncthreads();
        // four parcels per thread with their loads in flight together
        for (int j0 = t; j0 < cnt; j0 += 4 * CT) {
            int kk[4], dd[4], cc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u * CT;
                kk[u] = -1;
                dd[u] = 0;
                cc[u] = 0;
                if (j < cnt) {
                    const int l = L.loc[j];
                    const int ss = L.segS[l] - S0, ee = L.segS[l + 1] - S0;
                    kk[u] = L.key[j];
                    dd[u] = ss + rank_in(L.key, ss, ee, kk[u]);
                    cc[u] = c0 + l;
                }
            }
            double x[4], y[4], z[4];
            float ux[4], uy[4], uz[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = kk[u] >= 0 ? kk[u] : kk[0];
                x[u] = src.px[k]; y[u] = src.py[k]; z[u] = src.pz[k];
                ux[u] = src.ux[k]; uy[u] = src.uy[k]; uz[u] = src.uz[k];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (kk[u] < 0) continue;
                const int64_t gd = (int64_t)S0 + dd[u];
                __builtin_nontemporal_store(x[u], &dst.px[gd]); __builtin_nontemporal_store(y[u], &dst.py[gd]);
                __builtin_nontemporal_store(z[u], &dst.pz[gd]);
                dst.ux[gd] = ux[u]; dst.uy[gd] = uy[u]; dst.uz[gd] = uz[u];
                __builtin_nontemporal_store(cc[u], &dst.cell[gd]);
                if (GEN) copy_internal(P, src, dst, kk[u], gd);
                sub_out[gd] = (uint8_t)subcell_of(P, cc[u], x[u], y[u], z[u]);
            }
        }
        __syncthreads();
        a = b;
    }
}

template <bool GEN>
__device__ __forceinline__ void gather_body(const DevParams* __restrict__ Pp, DevParticles& src, DevParticles& dst,
                                            const int32_t* __restrict__ start, const int32_t* __restrict__ perm,
                                            const int32_t* __restrict__ tile_begin, int pt,
                                            const DevScalars* __restrict__ S, uint8_t* __restrict__ sub_out,
                                            const int q, int32_t* __restrict__ scratch) {
    __shared__ GatherLds L;
    const DevParams& P = *Pp;
    // the tile bounds are read beside the population, not after it (q + 1 is
    // within the tile array, which has capacity / 64 + 4 entries, for every
    // workgroup of the launch); a tile past the population is not used
    const int cb = tile_begin[q], ce = tile_begin[q + 1];
    if ((int64_t)q * pt > (int64_t)S->n_sorted) return;
    for (int c0 = cb; c0 < ce; c0 += CT_CELLS)
        gather_cells<GEN>(L, P, c0, min(CT_CELLS, ce - c0), src, dst, start, perm, sub_out, scratch);
}

template <bool GEN>
__global__ __launch_bounds__(CT) void k_gather(const DevParams* __restrict__ Pp, DevParticles src, DevParticles dst,
                                               const int32_t* __restrict__ start, const int32_t* __restrict__ perm,
                                               const int32_t* __restrict__ tile_begin, int pt,
                                               const DevScalars* __restrict__ S, uint8_t* __restrict__ sub_out,
                                               int32_t* __restrict__ scratch) {
    gather_body<GEN>(Pp, src, dst, start, perm, tile_begin, pt, S, sub_out, (int)blockIdx.x, scratch);
}

// ---------------------------------------------------------------------------
// collide with QK chemistry (dsmcReactions): the one-pass sums of a cell and
// its products in global memory
// ---------------------------------------------------------------------------
// The cell's sums without the heat-flux terms in one pass: parcels
// b .. b+nC-1, then b2 .. b2+n2-1, each added to its species' accumulators
// (every species' sums keep the parcel order: the same values as a pass per
// species).  MAXS bounds the species count.
template <int MAXS>
__device__ __forceinline__ void sample_cell_react(const DevParams& P, const DevParticles& d, int64_t b, int nC,
                                                  int64_t b2, int n2, double* __restrict__ samples, int nc, int c) {
    const int64_t cap = P.capacity;
    double acc[MAXS][HS_SS_COUNT];
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
#pragma unroll
        for (int q = 0; q < HS_SS_COUNT; ++q) acc[s][q] = 0.0;
    for (int k = 0; k < nC + n2; ++k) {
        const int64_t i = k < nC ? b + k : b2 + (k - nC);
        const int t = (int)d.type[i];
        const double ux = (double)d.ux[i], uy = (double)d.uy[i], uz = (double)d.uz[i];
        const double er = d.erot ? (double)d.erot[i] : 0.0;
        int lv[HS_MAX_VIB_MODES] = { 0, 0, 0 };
        for (int mo = 0; mo < P.nvib[t]; ++mo) lv[mo] = d.vib[mo * cap + i];
#pragma unroll
        for (int s = 0; s < MAXS; ++s) {
            if (t != s) continue;
            const double m = P.mass[s];
            acc[s][HS_SS_N] += 1.0;
            acc[s][HS_SS_MUX] += m * ux;
            acc[s][HS_SS_MUY] += m * uy;
            acc[s][HS_SS_MUZ] += m * uz;
            acc[s][HS_SS_MCC] += m * (ux * ux + uy * uy + uz * uz);
            acc[s][HS_SS_EROT] += er;
#pragma unroll
            for (int mo = 0; mo < HS_MAX_VIB_MODES; ++mo)
                if (mo < P.nvib[s]) acc[s][HS_SS_EVIB0 + mo] += (double)lv[mo] * HS_KB * P.theta_v[s][mo];
        }
    }
#pragma unroll
    for (int s = 0; s < MAXS; ++s) {
        if (s >= P.ns) break;
        if (acc[s][HS_SS_N] == 0.0) continue;  // no parcel of the species: adding zeros changes no sum
        const int64_t base = HS_S_GLOBAL_COUNT + (int64_t)s * HS_SS_COUNT;
#pragma unroll
        for (int q = 0; q < HS_SS_COUNT; ++q) unsafeAtomicAdd(&samples[(base + q) * nc + c], acc[s][q]);
    }
}

// ---------------------------------------------------------------------------
// k_collide: NTC, collisions and sampling on the sorted population
// ---------------------------------------------------------------------------
// One cell with more parcels than an LDS chunk holds, in global memory: the
// first wave runs the wave-cooperative NTC (ntc_wave) on the cell's parcels
// in place.  Its conflict stamps are the chunk's LDS arrays (lds_ints 32-bit
// words, unused while this cell is handled): one per parcel when the cell
// fits, else shared by parcels whose index agrees modulo a power of two
// (which only delays candidates).  The wave builds the sub-cell lists: 16-bit
// up to 65,535 parcels; beyond, with chemistry, 32-bit (any size), and
// without it the cell is counted as overflow and not collided.  Then the
// first wave sums the cell; with chemistry a cell that made products is
// summed after they are placed (k_sample_react).
template <bool GEN, bool REACT = false>
__device__ __forceinline__ void collide_oversize(const DevParams& P, int c, int S0, int nC, DevParticles& dst,
                                              double* __restrict__ sigma, double* __restrict__ rem,
                                              double* __restrict__ samples, const double* __restrict__ vol,
                                              uint32_t step, int do_collide, int do_sample, int nc, uint16_t* big_list,
                                              uint8_t* sub, CellCtx& cx, const ProdOut* po, int32_t* lds,
                                              int lds_ints) {
    const int64_t cap = P.capacity;
    const int t = threadIdx.x;
    float* gvx = dst.ux + S0;
    float* gvy = dst.uy + S0;
    float* gvz = dst.uz + S0;
    float* ger = GEN && dst.erot ? dst.erot + S0 : nullptr;
    uint16_t* gvb = GEN && dst.vib ? dst.vib + S0 : nullptr;
    uint8_t* gty = GEN ? dst.type + S0 : nullptr;
    __shared__ unsigned long long s_tab[8];
    __shared__ double s_smax;
    __shared__ long long s_ncand;
    __shared__ int s_nprod;
    double ncoll = 0.0, sep = 0.0;
    const bool lists16 = nC <= 65535;
    if (t == 0) s_nprod = 0;
    if (do_collide && nC > 1 && (lists16 || REACT)) {
        int hash = 1;
        while (hash * 2 <= lds_ints) hash *= 2;
        const int nstamp = nC <= lds_ints ? nC : hash;
        const int smask = nC <= lds_ints ? 0x7FFFFFFF : hash - 1;
        if (t < WAVE) {
            if (lists16) subcell_lists_wave<uint16_t>(sub + S0, big_list + S0, 0, nC, s_tab);
            else subcell_lists_wave<int32_t>(sub + S0, po->pr.list + S0, 0, nC, s_tab);
        }
        for (int j = t; j < nstamp; j += CT) lds[j] = 0x7FFFFFFF;
        if (t == 0) {
            double smax;
            s_ncand = ntc_candidates(P, c, nC, sigma, rem, vol, &smax, cx);
            s_smax = smax;
        }
        __syncthreads();
        if (t < WAVE) {
            uint32_t nco = 0;
            double sp = 0.0;
            int np = 0;
            if (lists16) {
                const CellTab tab = { s_tab, big_list + S0, 0, false };
                ntc_wave<GEN, REACT, true>(P, c, 0, nC, s_ncand, s_smax, gvx, gvy, gvz, sub + S0, tab, ger, gvb,
                                           (int)cap, gty, lds, sigma, step, dst, S0, cx, nco, sp, po, &np, smask);
            } else {
                const CellTab32 tab = { s_tab, po->pr.list + S0 };
                ntc_wave<GEN, REACT, true, CellTab32>(P, c, 0, nC, s_ncand, s_smax, gvx, gvy, gvz, sub + S0, tab, ger,
                                                      gvb, (int)cap, gty, lds, sigma, step, dst, S0, cx, nco, sp, po,
                                                      &np, smask);
            }
            ncoll = (double)nco;
            sep = sp;
            if (REACT && t == 0) s_nprod = np;
        }
    } else if (t == 0 && !lists16 && !REACT) {
        cx.ovf += 1;
    }
    if (t == 0) {
        if (REACT) po->pr.count[c] = s_nprod;
        if (do_sample && ncoll != 0.0) {
            unsafeAtomicAdd(&samples[(int64_t)HS_S_NCOLL * nc + c], ncoll);
            unsafeAtomicAdd(&samples[(int64_t)HS_S_COLLSEP * nc + c], sep);
        }
    }
    __threadfence_block();
    __syncthreads();
    // the cell's sums by the first wave
    const bool summed_later = REACT && s_nprod > 0;
    if (do_sample && !summed_later && (REACT || lists16) && t < WAVE)
        sample_cell_wave<GEN>(P, c, 0, nC, gvx, gvy, gvz, ger, gvb, (int)cap, gty, samples, nc);
    __syncthreads();
}

// Collide and sample the cells c0 .. c0+ncell-1 (ncell <= CT_CELLS) in LDS
// chunks of at most PMAX parcels.
// REACT: with chemistry (ProdOut po): product counts per cell, types
// written back, and a cell that made products left to k_sample_react.
template <bool GEN, int PMAX, bool REACT = false>
__device__ void collide_cells(TileLds<GEN, PMAX, REACT>& L, const DevParams& P, int c0, int ncell, DevParticles& dst,
                              const int32_t* __restrict__ start, uint8_t* __restrict__ sub,
                              double* __restrict__ sigma, double* __restrict__ rem, double* __restrict__ samples,
                              const double* __restrict__ vol, uint32_t step, int do_collide, int do_sample, int nc,
                              uint16_t* big_list, CellCtx& cx, const ProdOut* po = nullptr) {
    const int t = threadIdx.x;
    const int wid = t / WAVE, lane = t & (WAVE - 1);
    const int64_t cap = P.capacity;
    if (t < ncell) L.segS[t] = start[c0 + t];
    if (t == 0) {
        L.segS[ncell] = start[c0 + ncell];
        L.nbig = 0;
    }
    __syncthreads();
    int a = 0;
    while (a < ncell) {
        const int S0 = L.segS[a];
        const bool ok = t >= a && t < ncell && L.segS[t + 1] - S0 <= PMAX;
        const int b = a + __syncthreads_count(ok);
        if (b == a) {
            collide_oversize<GEN, REACT>(P, c0 + a, S0, L.segS[a + 1] - S0, dst, sigma, rem, samples, vol, step,
                                         do_collide, do_sample, nc, big_list, sub, cx, po, L.key,
                                         (int)((reinterpret_cast<char*>(&L.segS[0]) - reinterpret_cast<char*>(&L.key[0]))
                                               / sizeof(int32_t)));
            a += 1;
            continue;
        }
        const int cnt = L.segS[b] - S0;
        // slot -> cell map; each cell's sigmaTcRMax, remainder and volume; the
        // sorted velocities and sub-cells of the chunk (contiguous)
        double pf_sig = 0.0, pf_rem = 0.0, pf_vol = 1.0;
        if (t >= a && t < b) {
            if (do_collide) {
                pf_sig = sigma[c0 + t];
                pf_rem = rem[c0 + t];
                pf_vol = vol[c0 + t];
            }
            {
                // the cell's slot range, four bytes per store where it covers whole words
                int j = L.segS[t] - S0;
                const int e = L.segS[t + 1] - S0;
                for (; j < e && (j & 3); ++j) L.loc[j] = (uint8_t)t;
                const uint32_t w4 = 0x01010101u * (uint32_t)t;
                for (; j + 4 <= e; j += 4) *reinterpret_cast<uint32_t*>(&L.loc[j]) = w4;
                for (; j < e; ++j) L.loc[j] = (uint8_t)t;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) L.smask[t][q] = 0ull;
            if (L.segS[t + 1] - L.segS[t] > 64) L.big[atomicAdd(&L.nbig, 1)] = t;
        }
        // the chunk's velocities and sub-cells: a batch of loads is issued
        // before any of its LDS stores, so their latencies overlap.  With
        // internal energies a batch is one parcel's velocity, sub-cell,
        // rotational energy, vibrational levels and species (measured: batches
        // of 2 parcels were no faster, of 4 slower, from the NTC's spills)
        if (GEN) {
            constexpr int LG = 1;
            for (int j0 = t; j0 < cnt; j0 += LG * CT) {
                float bx[LG], by[LG], bz[LG], be[LG];
                uint8_t bs[LG], bt[LG];
                uint16_t bv[LG][HS_MAX_VIB_MODES];
#pragma unroll
                for (int u = 0; u < LG; ++u) {
                    const int j = j0 + u * CT;
                    if (j < cnt) {
                        const int64_t gj = (int64_t)S0 + j;
                        bx[u] = dst.ux[gj]; by[u] = dst.uy[gj]; bz[u] = dst.uz[gj];
                        bs[u] = sub[gj];
                        be[u] = dst.erot ? dst.erot[gj] : 0.0f;
#pragma unroll
                        for (int m = 0; m < HS_MAX_VIB_MODES; ++m)
                            if (m < P.max_modes) bv[u][m] = dst.vib[m * cap + gj];
                        bt[u] = dst.type ? dst.type[gj] : (uint8_t)0;
                    }
                }
#pragma unroll
                for (int u = 0; u < LG; ++u) {
                    const int j = j0 + u * CT;
                    if (j < cnt) {
                        L.vx[j] = bx[u]; L.vy[j] = by[u]; L.vz[j] = bz[u];
                        L.sub[j] = bs[u];
                        L.er[j] = be[u];
#pragma unroll
                        for (int m = 0; m < HS_MAX_VIB_MODES; ++m)
                            if (m < P.max_modes) L.vb[m * PMAX + j] = bv[u][m];
                        L.ty[j] = bt[u];
                    }
                }
            }
        } else {
            constexpr int LB = 8;
            for (int j0 = t; j0 < cnt; j0 += LB * CT) {
                float bx[LB], by[LB], bz[LB];
                uint8_t bs[LB];
#pragma unroll
                for (int u = 0; u < LB; ++u) {
                    const int j = j0 + u * CT;
                    if (j < cnt) {
                        const int64_t gj = (int64_t)S0 + j;
                        bx[u] = dst.ux[gj]; by[u] = dst.uy[gj]; bz[u] = dst.uz[gj];
                        bs[u] = sub[gj];
                    }
                }
#pragma unroll
                for (int u = 0; u < LB; ++u) {
                    const int j = j0 + u * CT;
                    if (j < cnt) {
                        L.vx[j] = bx[u]; L.vy[j] = by[u]; L.vz[j] = bz[u];
                        L.sub[j] = bs[u];
                    }
                }
            }
        }
        if (t == 0) {
            L.ndense = 0;
            L.dnext = 0;
        }
        __syncthreads();
        // sub-cell tables in parallel: a parcel of a cell of at most 64 sets
        // its bit in its sub-cell's mask; in a larger cell it finds its place
        // among the parcels of its sub-cell (ascending index).  Stamps cleared.
        for (int j = t; j < cnt; j += CT) {
            const int l = L.loc[j];
            const int ss = L.segS[l] - S0, ee = L.segS[l + 1] - S0;
            if (ee - ss <= 64) atomicOr(&L.smask[l][L.sub[j]], 1ull << (j - ss));
            L.key[j] = 0x7FFFFFFF;
        }
        // cells of more than 64 parcels: one wave each builds the member
        // lists with ballots over 64-parcel windows
        for (int bi = wid; bi < L.nbig; bi += CT / WAVE)
            subcell_lists_wave(L.sub, L.slist, L.segS[L.big[bi]] - S0,
                               L.segS[L.big[bi] + 1] - L.segS[L.big[bi]], L.smask[L.big[bi]]);
        // candidates per cell (noTimeCounter.C:142-153)
        if (t >= a && t < b) {
            const int c = c0 + t;
            const int nC = L.segS[t + 1] - L.segS[t];
            int nsparse = 0;
            L.rcoll[t] = 0u;
            L.sep[t] = 0.0;
            if (REACT) L.pcount[t] = 0;
            if (do_collide && nC > 1) {
                const double smax = pf_sig;
                const long long ncand = ntc_candidates_of(P, c, nC, smax, pf_rem, pf_vol, rem, cx);
                L.smax[t] = smax;
                L.rsmax[t] = (unsigned long long)__double_as_longlong(smax);
                L.ncand[t] = (int32_t)ncand;
                if (ncand > DENSE_CAND) {
                    L.dense[atomicAdd(&L.ndense, 1)] = t;
                } else {
                    nsparse = (int)ncand;
                    cx.cand += (unsigned long long)ncand;
                }
            }
            L.rn[t] = nsparse;
        }
        __syncthreads();
        // dense cells: one wave each.  In the variant with internal energies
        // (without chemistry) the cells with most candidates go first, each
        // wave taking the next cell when it finishes one; otherwise wave w
        // takes cells w, w+4, ... (measured faster there)
        constexpr bool LPT = GEN && !REACT;
        const int nd = L.ndense;
        if (LPT && nd > CT / WAVE && nd <= 64) {
            if (t == 0) {
                for (int x = 1; x < nd; ++x) {
                    const int v = L.dense[x], k = L.ncand[v];
                    int y = x - 1;
                    for (; y >= 0 && L.ncand[L.dense[y]] < k; --y) L.dense[y + 1] = L.dense[y];
                    L.dense[y + 1] = v;
                }
            }
            __syncthreads();
        }
        for (int dw = wid;; dw += CT / WAVE) {
            int di = dw;
            if (LPT) {
                if (lane == 0) di = atomicAdd(&L.dnext, 1);
                di = __shfl(di, 0, WAVE);
            }
            if (di >= nd) break;
            const int l = L.dense[di];
            const int ss = L.segS[l] - S0, nC = L.segS[l + 1] - L.segS[l];
            const CellTab tab = { L.smask[l], L.slist, ss, nC <= 64 };
            uint32_t ncoll;
            double sep;
            int np = 0;
            ntc_wave<GEN, REACT>(P, c0 + l, ss, nC, L.ncand[l], L.smax[l], L.vx, L.vy, L.vz, L.sub, tab, L.er, L.vb,
                                 PMAX, L.ty, L.key, sigma, step, dst, (int64_t)S0 + ss, cx, ncoll, sep, po, &np);
            if (lane == 0) {
                L.rcoll[l] = ncoll;
                L.sep[l] = sep;
                if (REACT) L.pcount[l] = np;
            }
        }
        // sparse cells: candidate-parallel rounds over the workgroup
        ntc_rounds<GEN, PMAX, REACT>(L, P, c0, a, b, S0, step, dst, cx, po);
        if (t >= a && t < b && L.rn[t] > 0) {
            sigma[c0 + t] = __longlong_as_double((long long)L.rsmax[t]);
            cx.colls += L.rcoll[t];
        }
        __syncthreads();
        if (REACT && t >= a && t < b) po->pr.count[c0 + t] = L.pcount[t];
        if (do_sample) {
            // cells of up to 64 parcels: four lanes each (one thread per species
            // in the variant with internal energies); larger ones: one wave each
            const int nbig = L.nbig;
            if (GEN) {
                // one thread per (cell, species); a cell that made products is
                // summed with them once they are placed
                const int ns = P.ns;
                const int nw = (b - a) * ns;
                for (int w = t; w < nw; w += CT) {
                    const int l = a + w / ns, sp = w - (w / ns) * ns;
                    const int ss = L.segS[l] - S0, nC = L.segS[l + 1] - L.segS[l];
                    if (sp == 0 && L.rcoll[l] != 0u) {
                        unsafeAtomicAdd(&samples[(int64_t)HS_S_NCOLL * nc + c0 + l], (double)L.rcoll[l]);
                        unsafeAtomicAdd(&samples[(int64_t)HS_S_COLLSEP * nc + c0 + l], L.sep[l]);
                    }
                    if (nC <= 64 && !(REACT && L.pcount[l] > 0))
                        sample_cell_species<GEN>(P, c0 + l, ss, nC, sp, L.vx, L.vy, L.vz, L.er, L.vb, PMAX, L.ty,
                                                 samples, nc);
                }
            } else {
                const int r = t & 3;
                for (int l = a + (t >> 2); l < b; l += CT / 4) {
                    const int ss = L.segS[l] - S0, nC = L.segS[l + 1] - L.segS[l];
                    if (nC <= 64) {
                        sample_cell_quad<GEN>(P, c0 + l, ss, nC, r, L.vx, L.vy, L.vz, L.er, L.vb, PMAX, L.ty,
                                              samples, nc, (double)L.rcoll[l], L.sep[l]);
                    } else if (r == 0 && L.rcoll[l] != 0u) {
                        unsafeAtomicAdd(&samples[(int64_t)HS_S_NCOLL * nc + c0 + l], (double)L.rcoll[l]);
                        unsafeAtomicAdd(&samples[(int64_t)HS_S_COLLSEP * nc + c0 + l], L.sep[l]);
                    }
                }
            }
            for (int bi = wid; bi < nbig; bi += CT / WAVE) {
                const int l = L.big[bi];
                if (REACT && L.pcount[l] > 0) continue;
                sample_cell_wave<GEN>(P, c0 + l, L.segS[l] - S0, L.segS[l + 1] - L.segS[l], L.vx, L.vy, L.vz, L.er,
                                      L.vb, PMAX, L.ty, samples, nc);
            }
        }
        for (int j = t; j < cnt; j += CT) {
            const int64_t gd = (int64_t)S0 + j;
            dst.ux[gd] = L.vx[j]; dst.uy[gd] = L.vy[j]; dst.uz[gd] = L.vz[j];
            if (GEN) {
                if (dst.erot) dst.erot[gd] = L.er[j];
                if (dst.vib)
                    for (int m = 0; m < P.max_modes; ++m) dst.vib[m * cap + gd] = L.vb[m * PMAX + j];
                if (REACT) dst.type[gd] = L.ty[j];
            }
        }
        // every wave has read the big-cell list before it is cleared
        __syncthreads();
        if (t == 0) L.nbig = 0;
        __syncthreads();
        a = b;
    }
}

// Tile q of the particle-balanced decomposition owns the cells whose start
// offset lies in [q*PT, (q+1)*PT): tile_begin[q] is its first cell.  Two
// tilings from one pass: the gather's (pt_g parcels) and the collide's (pt_c).
// Also flags the collide tiles holding a cell of more than HEAVY_CELL parcels
// (heavy[] cleared beforehand).
__device__ __forceinline__ void tile_marks(const int32_t* __restrict__ start, int i, int nc, int pt,
                                           int32_t* __restrict__ tile_begin, int* q0_out) {
    const int q0 = start[i - 1] / pt, q1 = start[i] / pt;
    for (int q = q0 + 1; q <= q1; ++q) tile_begin[q] = i;
    if (i == nc) tile_begin[q1 + 1] = nc;
    *q0_out = q0;
}

__global__ __launch_bounds__(BLOCK) void k_tiles(const int32_t* __restrict__ start, int nc, int pt_g,
                                                 int32_t* __restrict__ tile_begin_g, int pt_c,
                                                 int32_t* __restrict__ tile_begin_c, int32_t* __restrict__ heavy) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i > nc) return;
    if (i == 0) { tile_begin_g[0] = 0; tile_begin_c[0] = 0; return; }
    int qg, qc;
    tile_marks(start, i, nc, pt_g, tile_begin_g, &qg);
    tile_marks(start, i, nc, pt_c, tile_begin_c, &qc);
    if (start[i] - start[i - 1] > HEAVY_CELL) heavy[qc] = 1;
}

template <bool GEN, int PMAX, bool REACT = false>
__device__ __forceinline__ void collide_body(const DevParams* __restrict__ Pp, DevParticles& dst,
                                             const int32_t* __restrict__ start, const int32_t* __restrict__ tile_begin,
                                             const int32_t* __restrict__ order, int pt, uint8_t* __restrict__ sub,
                                             double* __restrict__ sigma, double* __restrict__ rem,
                                             double* __restrict__ samples, const double* __restrict__ vol,
                                             DevScalars* __restrict__ S, int do_collide, int do_sample, int nc,
                                             uint16_t* big_list, const int bx, const ProdOut* po = nullptr) {
    __shared__ TileLds<GEN, PMAX, REACT> L;
    const DevParams& P = *Pp;
    // the tile order is read beside the population (every workgroup of the
    // launch is within the order array); its entry is used only by a
    // workgroup inside the population, whose entry the scatter wrote
    const int q_any = order[bx];
    const int64_t nsort = S->n_sorted;
    if ((int64_t)bx * pt > nsort) return;
    const int q = q_any;
    const int cb = tile_begin[q], ce = tile_begin[q + 1];
    if (cb >= ce) return;
    const uint32_t step = S->step;
    CellCtx cx = { 0, 0, 0 };
    for (int c0 = cb; c0 < ce; c0 += CT_CELLS)
        collide_cells<GEN, PMAX, REACT>(L, P, c0, min(CT_CELLS, ce - c0), dst, start, sub, sigma, rem, samples, vol,
                                        step, do_collide, do_sample, nc, big_list, cx, po);
    add_counter(S, CNT_CAND, cx.cand);
    add_counter(S, CNT_COLL, cx.colls);
    add_counter(S, CNT_TILE_OVF, cx.ovf);
}

template <bool GEN, int PMAX>
__global__ __launch_bounds__(CT, 2) void k_collide(const DevParams* __restrict__ Pp, DevParticles dst,
                                                   const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ tile_begin,
                                                   const int32_t* __restrict__ order, int pt,
                                                   uint8_t* __restrict__ sub, double* __restrict__ sigma,
                                                   double* __restrict__ rem, double* __restrict__ samples,
                                                   const double* __restrict__ vol, DevScalars* __restrict__ S,
                                                   int do_collide, int do_sample, int nc, uint16_t* big_list) {
    collide_body<GEN, PMAX>(Pp, dst, start, tile_begin, order, pt, sub, sigma, rem, samples, vol, S, do_collide,
                            do_sample, nc, big_list, (int)blockIdx.x);
}

// the collide with QK chemistry: the variant with internal energies, whose
// accepted pairs react (or collide), staging their products
template <int PMAX>
__global__ __launch_bounds__(CT, 2) void k_collide_chem(const DevParams* __restrict__ Pp, DevParticles dst,
                                                        const int32_t* __restrict__ start,
                                                        const int32_t* __restrict__ tile_begin,
                                                        const int32_t* __restrict__ order, int pt,
                                                        uint8_t* __restrict__ sub, double* __restrict__ sigma,
                                                        double* __restrict__ rem, double* __restrict__ samples,
                                                        const double* __restrict__ vol, DevScalars* __restrict__ S,
                                                        int do_collide, int do_sample, int nc, uint16_t* big_list,
                                                        DevProducts pr) {
    const ProdOut po = { S, pr };
    collide_body<true, PMAX, true>(Pp, dst, start, tile_begin, order, pt, sub, sigma, rem, samples, vol, S, do_collide,
                                   do_sample, nc, big_list, (int)blockIdx.x, &po);
}

void launch_tiles(const int32_t* start, int nc, int pt_g, int32_t* tile_begin_g, int pt_c, int32_t* tile_begin_c,
                  int32_t* heavy, hipStream_t st) {
    k_tiles<<<(nc + 1 + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(start, nc, pt_g, tile_begin_g, pt_c, tile_begin_c, heavy);
}

#define PMAX_MONO 2304
#define PMAX_GEN 1520

void launch_gather(const DevParams* P, DevParticles src, DevParticles dst, const int32_t* start, const int32_t* perm,
                   const int32_t* tile_begin, int pt, int64_t capacity, const DevScalars* S, uint8_t* sub,
                   int general, int32_t* scratch, hipStream_t st) {
    const int nt = (int)(capacity / pt) + 2;
    if (general)
        k_gather<true><<<nt, CT, 0, st>>>(P, src, dst, start, perm, tile_begin, pt, S, sub, scratch);
    else
        k_gather<false><<<nt, CT, 0, st>>>(P, src, dst, start, perm, tile_begin, pt, S, sub, scratch);
}

void launch_collide(const DevParams* P, DevParticles dst, const int32_t* start, const int32_t* tile_begin,
                    int32_t* heavy, int32_t* order, int pt, int64_t capacity, uint8_t* sub, double* sigma,
                    double* rem, double* samples, const double* vol, DevScalars* S, int do_collide, int do_sample,
                    int general, int nc, uint16_t* big_list, hipStream_t st) {
    const int nt = (int)(capacity / pt) + 2;
    (void)heavy;  // the order was built by k_scatter
    if (general)
        k_collide<true, PMAX_GEN><<<nt, CT, 0, st>>>(P, dst, start, tile_begin, order, pt, sub, sigma, rem, samples,
                                                      vol, S, do_collide, do_sample, nc, big_list);
    else
        k_collide<false, PMAX_MONO><<<nt, CT, 0, st>>>(P, dst, start, tile_begin, order, pt, sub, sigma, rem,
                                                        samples, vol, S, do_collide, do_sample, nc, big_list);
}

// The table of 1/Zv from the collision temperature (inv_zv_collision) for
// every species, partner, vibrational mode and top level 1 <= imax < HS_ZV_TAB.
__global__ __launch_bounds__(BLOCK) void k_zv_table(const DevParams* __restrict__ Pp, double* __restrict__ tab) {
    const DevParams& P = *Pp;
    const int n = P.ns * P.ns * HS_MAX_VIB_MODES * HS_ZV_TAB;
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
        const int imax = e % HS_ZV_TAB;
        int r = e / HS_ZV_TAB;
        const int mode = r % HS_MAX_VIB_MODES;
        r /= HS_MAX_VIB_MODES;
        const int partner = r % P.ns, t = r / P.ns;
        tab[e] = (imax > 0 && mode < P.nvib[t]) ? inv_zv_collision(P, t, partner, mode, imax, P.omega_pq[t][partner])
                                                : 0.0;
    }
}

void launch_zv_table(const DevParams* P, double* tab, int ns, hipStream_t st) {
    const int n = ns * ns * HS_MAX_VIB_MODES * HS_ZV_TAB;
    k_zv_table<<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(P, tab);
}

// the staged products after the sorted population, in (cell, rank) order
// (dsmcCloud::addNewParcel, dsmcCloud.C:396-432): at the dissociating
// parcel's position, with no internal energy; beyond the capacity dropped
__global__ __launch_bounds__(BLOCK) void k_place_products(const DevParams* __restrict__ Pp, DevParticles dst,
                                                          const DevScalars* __restrict__ S, DevProducts pr) {
    const DevParams& P = *Pp;
    const int64_t cap = P.capacity;
    const int64_t base = S->n_sorted;
    const int64_t np = S->nprod < cap ? S->nprod : cap;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < np; e += (int64_t)gridDim.x * BLOCK) {
        const int c = pr.cell[e];
        const int64_t j = base + pr.off[c] + pr.rank[e];
        if (j >= cap) continue;
        const int64_t par = pr.parent[e];
        dst.px[j] = dst.px[par]; dst.py[j] = dst.py[par]; dst.pz[j] = dst.pz[par];
        dst.ux[j] = pr.ux[e]; dst.uy[j] = pr.uy[e]; dst.uz[j] = pr.uz[e];
        dst.cell[j] = c;
        if (dst.erot) dst.erot[j] = 0.0f;
        for (int m = 0; m < P.max_modes; ++m) dst.vib[m * cap + j] = 0;
        if (dst.type) dst.type[j] = pr.type[e];
    }
}

// The cell's sums (dsmcVolFields.C:1115-1239) after its reactions: per
// species, its parcels and then the products it made this step, one thread
// adding them in order (the oracle's order for such a cell).  The collide
// has summed the cells that made no products.
template <int MAXS>
__global__ __launch_bounds__(BLOCK) void k_sample_react(const DevParams* __restrict__ Pp, DevParticles dst,
                                                        const int32_t* __restrict__ start,
                                                        DevScalars* __restrict__ S, DevProducts pr,
                                                        double* __restrict__ samples, int nc) {
    const DevParams& P = *Pp;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    const bool hf = (P.flags & HS_FLAG_HEATFLUX) != 0;
    if (pr.count[c] == 0) return;
    const int64_t cap = P.capacity;
    const int64_t b = start[c];
    const int nC = start[c + 1] - start[c];
    const int64_t b2 = (int64_t)S->n_sorted + pr.off[c];
    int n2 = pr.count[c];
    if (b2 + n2 > cap) n2 = b2 < cap ? (int)(cap - b2) : 0;
    if (nC + n2 > 64) {
        // summed by a wave (k_sample_react_big)
        pr.bigcells[atomicAdd(&S->nbigprod, 1)] = c;
        return;
    }
    if (!hf) {
        sample_cell_react<MAXS>(P, dst, b, nC, b2, n2, samples, nc, c);
        return;
    }
    for (int s = 0; s < P.ns; ++s) {
        const double m = P.mass[s];
        const int nvib = P.nvib[s];
        double aN = 0.0, aX = 0.0, aY = 0.0, aZ = 0.0, aC = 0.0, aR = 0.0;
        double ev[HS_MAX_VIB_MODES] = { 0.0, 0.0, 0.0 };
        double h[HS_HF_COUNT];
        for (int q = 0; q < HS_HF_COUNT; ++q) h[q] = 0.0;
        for (int k = 0; k < nC + n2; ++k) {
            const int64_t i = k < nC ? b + k : b2 + (k - nC);
            if ((int)dst.type[i] != s) continue;
            const double ux = (double)dst.ux[i], uy = (double)dst.uy[i], uz = (double)dst.uz[i];
            const double lke = m * (ux * ux + uy * uy + uz * uz);
            const double er = dst.erot ? (double)dst.erot[i] : 0.0;
            aN += 1.0;
            aX += m * ux;
            aY += m * uy;
            aZ += m * uz;
            aC += lke;
            aR += er;
            double evs = 0.0;
            for (int mo = 0; mo < nvib; ++mo) {
                const double e = (double)dst.vib[mo * cap + i] * HS_KB * P.theta_v[s][mo];
                ev[mo] += e;
                evs += e;
            }
            if (hf) {
                const double eint = er + evs;
                h[0] += m * (ux * ux); h[1] += m * ux * uy; h[2] += m * ux * uz;
                h[3] += m * (uy * uy); h[4] += m * uy * uz; h[5] += m * (uz * uz);
                h[6] += lke * ux; h[7] += lke * uy; h[8] += lke * uz;
                h[9] += eint * ux; h[10] += eint * uy; h[11] += eint * uz; h[12] += eint;
            }
        }
        const int64_t base = HS_S_GLOBAL_COUNT + (int64_t)s * HS_SS_COUNT;
        unsafeAtomicAdd(&samples[(base + HS_SS_N) * nc + c], aN);
        unsafeAtomicAdd(&samples[(base + HS_SS_MUX) * nc + c], aX);
        unsafeAtomicAdd(&samples[(base + HS_SS_MUY) * nc + c], aY);
        unsafeAtomicAdd(&samples[(base + HS_SS_MUZ) * nc + c], aZ);
        unsafeAtomicAdd(&samples[(base + HS_SS_MCC) * nc + c], aC);
        unsafeAtomicAdd(&samples[(base + HS_SS_EROT) * nc + c], aR);
        for (int mo = 0; mo < HS_MAX_VIB_MODES; ++mo)
            unsafeAtomicAdd(&samples[(base + HS_SS_EVIB0 + mo) * nc + c], ev[mo]);
        if (hf) {
            const int64_t hb = HS_S_GLOBAL_COUNT + (int64_t)P.ns * HS_SS_COUNT + (int64_t)s * HS_HF_COUNT;
            for (int q = 0; q < HS_HF_COUNT; ++q) unsafeAtomicAdd(&samples[(hb + q) * nc + c], h[q]);
        }
    }
}

// The sums of a cell of more than 64 items (parcels and products) that made
// products, by one wave in the lane order of sample_cell_wave; the cells
// listed by k_sample_react, one wave each over a fixed grid.
__global__ __launch_bounds__(BLOCK) void k_sample_react_big(const DevParams* __restrict__ Pp, DevParticles dst,
                                                            const int32_t* __restrict__ start,
                                                            const DevScalars* __restrict__ S, DevProducts pr,
                                                            double* __restrict__ samples, int nc) {
    const DevParams& P = *Pp;
    const int64_t cap = P.capacity;
    const int nw = gridDim.x * (BLOCK / WAVE);
    const int nlist = S->nbigprod;
    for (int w = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE; w < nlist; w += nw) {
        const int c = pr.bigcells[w];
        const int b = start[c];
        const int nC = start[c + 1] - b;
        const int64_t b2 = (int64_t)S->n_sorted + pr.off[c];
        int n2 = pr.count[c];
        if (b2 + n2 > cap) n2 = b2 < cap ? (int)(cap - b2) : 0;
        sample_cell_wave<true>(P, c, b, nC, dst.ux, dst.uy, dst.uz, dst.erot, dst.vib, (int)cap, dst.type, samples, nc,
                               (int)b2, n2);
    }
}

// the population after this step's products; the staging counters cleared
__global__ void k_products_commit(DevScalars* S, const int32_t* __restrict__ off, int nc, int64_t cap) {
    const int64_t nn = (int64_t)S->n_sorted + off[nc];
    S->n_sorted = (int32_t)(nn > cap ? cap : nn);
    if (nn > cap) atomicAdd(&S->counters[CNT_CAP_OVF], (unsigned long long)(nn - cap));
    S->nprod = 0;
    S->nbigprod = 0;
}

void launch_collide_react(const DevParams* P, DevParticles dst, const int32_t* start, const int32_t* tile_begin,
                          int32_t* order, int pt, uint8_t* sub, double* sigma, double* rem, double* samples,
                          const double* vol, DevScalars* S, DevProducts pr, int do_collide, int do_sample, int nc,
                          int ns, uint16_t* big_list, ScanWs& ws, int64_t capacity, hipStream_t st) {
    const int nt = (int)(capacity / pt) + 2;
    k_collide_chem<PMAX_GEN><<<nt, CT, 0, st>>>(P, dst, start, tile_begin, order, pt, sub, sigma, rem, samples, vol, S,
                                                 do_collide, do_sample, nc, big_list, pr);
    launch_scan(pr.count, pr.off, nc, SCAN_PLAIN, ws, S, capacity, st);
    k_place_products<<<grid_for(capacity, 1024), BLOCK, 0, st>>>(P, dst, S, pr);
    const int g = (nc + BLOCK - 1) / BLOCK;
    if (do_sample) {
        if (ns <= 5) k_sample_react<5><<<g, BLOCK, 0, st>>>(P, dst, start, S, pr, samples, nc);
        else k_sample_react<HS_MAX_SPECIES><<<g, BLOCK, 0, st>>>(P, dst, start, S, pr, samples, nc);
        k_sample_react_big<<<grid_for(nc, 1024), BLOCK, 0, st>>>(P, dst, start, S, pr, samples, nc);
    }
    k_products_commit<<<1, 1, 0, st>>>(S, pr.off, nc, capacity);
}

// publish the sorted population (with the next step's inflow when it was
// injected ahead); optionally advance the step counter
__device__ __forceinline__ void commit_body(DevScalars* S, int advance, int with_next) {
    S->n = with_next ? S->n_next : S->n_sorted;
    S->n_old = S->n_sorted;
    S->ord[0] = 0;
    S->ord[1] = 0;
    S->qn = 0;  // the move's wall queue, consumed by k_move_wall
    if (advance) S->step += 1;
}
__global__ void k_commit(DevScalars* S, int advance, int with_next) { commit_body(S, advance, with_next); }
void launch_commit(DevScalars* S, int advance, int with_next, hipStream_t st) {
    k_commit<<<1, 1, 0, st>>>(S, advance, with_next);
}

// ---------------------------------------------------------------------------
// initial fill (dsmcMeshFill.C:68-~230)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_fill_count(const DevParams* __restrict__ Pp, const double* __restrict__ nd,
                                                      int exact, uint32_t fill_id, const double* __restrict__ vol,
                                                      int32_t* __restrict__ cnt, int nc, int ns, FillCells fc) {
    const DevParams& P = *Pp;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= nc * ns) return;
    const int c = e / ns, s = e - c * ns;
    const double V = vol[c];
    const double nds = fc.nd ? fc.nd[(int64_t)s * nc + c] : nd[s];
    if (!(V > 0.0) || !(nds > 0.0)) { cnt[e] = 0; return; }
    const double req = nds * V / P.fn;
    int ni;
    if (exact) {
        ni = (int)floor(req + 0.5);
    } else {
        PhiloxStream r;
        r.init(P, RNG_FILL_COUNT, (uint32_t)c, fill_id, (uint32_t)s);
        ni = (int)req;
        if ((req - (double)ni) > r.u()) ni += 1;
    }
    cnt[e] = ni;
}

__global__ __launch_bounds__(BLOCK) void k_fill_gen(const DevParams* __restrict__ Pp, const double* __restrict__ nd,
                                                    double T, double Trot, double Tvib, const double* __restrict__ u3,
                                                    uint32_t fill_id, const int32_t* __restrict__ cnt,
                                                    const int32_t* __restrict__ off, int nc, int ns, DevParticles b,
                                                    DevScalars* __restrict__ S, FillCells fc) {
    const DevParams& P = *Pp;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= nc * ns) return;
    const int ni = cnt[e];
    if (ni == 0) return;
    const int c = e / ns, s = e - c * ns;
    int idx[3];
    cell_ijk(P, c, idx);
    PhiloxStream r;
    r.init(P, RNG_FILL_GEN, (uint32_t)c, fill_id, (uint32_t)s);
    const double Tc = fc.t ? fc.t[c] : T;
    const double uc0 = fc.u ? fc.u[c] : u3[0], uc1 = fc.u ? fc.u[nc + c] : u3[1],
                 uc2 = fc.u ? fc.u[2 * (int64_t)nc + c] : u3[2];
    const double sc = sqrt(HS_KB * Tc / P.mass[s]);
    const int64_t base = (int64_t)S->n_old + off[e];
    for (int k = 0; k < ni; ++k) {
        const int64_t j = base + k;
        if (j >= P.capacity) break;
        double pos[3] = { 0.0, 0.0, 0.0 };
        bool ok = false;
        for (int tries = 0; tries < HS_MAX_FILL_TRIES && !ok; ++tries) {
            for (int a = 0; a < 3; ++a) {
                const double ua = r.u();
                pos[a] = P.empty[a] ? P.centre[a] : face_coord(P, a, idx[a]) + ua * P.delta[a];
            }
            ok = true;
            if (P.cyl) {
                const double rx = pos[0] - P.cyl_c[0], ry = pos[1] - P.cyl_c[1];
                if (rx * rx + ry * ry < P.cyl_r * P.cyl_r) ok = false;
            }
        }
        const double g0 = r.gauss(), g1 = r.gauss(), g2 = r.gauss();
        b.px[j] = pos[0]; b.py[j] = pos[1]; b.pz[j] = pos[2];
        b.ux[j] = (float)(sc * g0 + uc0);
        b.uy[j] = (float)(sc * g1 + uc1);
        b.uz[j] = (float)(sc * g2 + uc2);
        b.cell[j] = ok ? c : -1;
        if (b.vib)
            for (int m = 0; m < P.max_modes; ++m) b.vib[m * P.capacity + j] = 0;
        new_internal(P, b, j, s, fc.t ? Tc : Trot, fc.t ? Tc : Tvib, r);
    }
}

__global__ __launch_bounds__(BLOCK) void k_fill_cells(const DevParams* __restrict__ Pp, double smax, uint32_t fill_id,
                                                      double* __restrict__ sigma, double* __restrict__ rem, int nc,
                                                      const double* __restrict__ sig_cell) {
    const DevParams& P = *Pp;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    sigma[c] = sig_cell ? sig_cell[c] : smax;
    PhiloxStream r;
    r.init(P, RNG_REMAINDER, (uint32_t)c, fill_id, 0);
    rem[c] = r.u();
}

void launch_fill_count(const DevParams* P, const double* nd, int exact, uint32_t fill_id, const double* vol,
                       int32_t* cnt, int nc, int ns, FillCells fc, hipStream_t st) {
    const int n = nc * ns;
    k_fill_count<<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(P, nd, exact, fill_id, vol, cnt, nc, ns, fc);
}

void launch_fill_gen(const DevParams* P, const double* nd, double T, double Trot, double Tvib, const double* u3,
                     uint32_t fill_id, const int32_t* cnt, const int32_t* off, int nc, int ns, DevParticles b,
                     DevScalars* S, FillCells fc, hipStream_t st) {
    const int n = nc * ns;
    k_fill_gen<<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(P, nd, T, Trot, Tvib, u3, fill_id, cnt, off, nc, ns, b, S,
                                                          fc);
}

void launch_fill_cells(const DevParams* P, double smax, uint32_t fill_id, double* sigma, double* rem, int nc,
                       const double* sig_cell, hipStream_t st) {
    k_fill_cells<<<(nc + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(P, smax, fill_id, sigma, rem, nc, sig_cell);
}

// ---------------------------------------------------------------------------
// test hooks
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_binary_collide(const DevParams* __restrict__ Pp, int64_t n,
                                                          const int32_t* tp, const int32_t* tq, double* up, double* uq,
                                                          double* erp, double* erq, int32_t* vp, int32_t* vq,
                                                          const double* uni, int32_t nu, int32_t* used) {
    const DevParams& P = *Pp;
    const int64_t k = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (k >= n) return;
    InjStream r;
    r.src = uni + k * nu;
    r.n = nu;
    r.used = 0;
    double a[3] = { up[3 * k], up[3 * k + 1], up[3 * k + 2] };
    double b[3] = { uq[3 * k], uq[3 * k + 1], uq[3 * k + 2] };
    double ep = erp ? erp[k] : 0.0, eq = erq ? erq[k] : 0.0;
    int va[HS_MAX_VIB_MODES] = { 0, 0, 0 }, vbq[HS_MAX_VIB_MODES] = { 0, 0, 0 };
    if (vp)
        for (int m = 0; m < HS_MAX_VIB_MODES; ++m) va[m] = vp[k * HS_MAX_VIB_MODES + m];
    if (vq)
        for (int m = 0; m < HS_MAX_VIB_MODES; ++m) vbq[m] = vq[k * HS_MAX_VIB_MODES + m];
    binary_collide(P, tp[k], tq[k], a, b, &ep, &eq, va, vbq, r);
    for (int c = 0; c < 3; ++c) { up[3 * k + c] = a[c]; uq[3 * k + c] = b[c]; }
    if (erp) erp[k] = ep;
    if (erq) erq[k] = eq;
    if (vp)
        for (int m = 0; m < HS_MAX_VIB_MODES; ++m) vp[k * HS_MAX_VIB_MODES + m] = va[m];
    if (vq)
        for (int m = 0; m < HS_MAX_VIB_MODES; ++m) vq[k * HS_MAX_VIB_MODES + m] = vbq[m];
    used[k] = r.used;
}

__global__ __launch_bounds__(BLOCK) void k_sigma_tcr(const DevParams* __restrict__ Pp, int64_t n, const int32_t* tp,
                                                     const int32_t* tq, const double* up, const double* uq,
                                                     double* out) {
    const DevParams& P = *Pp;
    const int64_t k = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (k >= n) return;
    const double a[3] = { up[3 * k], up[3 * k + 1], up[3 * k + 2] };
    const double b[3] = { uq[3 * k], uq[3 * k + 1], uq[3 * k + 2] };
    out[k] = sigma_tcr(P, tp[k], tq[k], a, b);
}

void launch_binary_collide(const DevParams* P, int64_t n, const int32_t* tp, const int32_t* tq, double* up,
                           double* uq, double* erp, double* erq, int32_t* vp, int32_t* vq, const double* uni,
                           int32_t nu, int32_t* used, hipStream_t st) {
    if (n <= 0) return;
    k_binary_collide<<<(int)((n + BLOCK - 1) / BLOCK), BLOCK, 0, st>>>(P, n, tp, tq, up, uq, erp, erq, vp, vq, uni,
                                                                      nu, used);
}

void launch_sigma_tcr(const DevParams* P, int64_t n, const int32_t* tp, const int32_t* tq, const double* up,
                      const double* uq, double* out, hipStream_t st) {
    if (n <= 0) return;
    k_sigma_tcr<<<(int)((n + BLOCK - 1) / BLOCK), BLOCK, 0, st>>>(P, n, tp, tq, up, uq, out);
}

// ---------------------------------------------------------------------------
// plain device copy (HBM bandwidth and counter calibration)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_copy(const double4* __restrict__ a, double4* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) b[i] = a[i];
}
void launch_copy(const void* a, void* b, int64_t bytes, int grid, hipStream_t st) {
    k_copy<<<grid, BLOCK, 0, st>>>((const double4*)a, (double4*)b, bytes / 32);
}

// ---------------------------------------------------------------------------
// batched stages: every workgroup reads its case and its index within that
// case's launch from the table (one scalar load), then runs the same body as
// the single-engine kernel, so a batched step equals the engines' own steps
// ---------------------------------------------------------------------------
// The tables are read through the constant address space: the loads are
// invariant for the kernel, so the compiler may reload a descriptor field
// rather than hold it in registers across the stage (as with kernel arguments).
#if defined(__HIP_DEVICE_COMPILE__)
#define HS_CONST_AS __attribute__((address_space(4)))
#else
#define HS_CONST_AS
#endif
__device__ __forceinline__ const HS_CONST_AS BatchCase& batch_case(const BatchCase* __restrict__ C,
                                                                   const BatchBlock* __restrict__ T, int& bx,
                                                                   int& gx) {
    const HS_CONST_AS BatchBlock* Tc = (const HS_CONST_AS BatchBlock*)T;
    const HS_CONST_AS BatchCase* Cc = (const HS_CONST_AS BatchCase*)C;
    const BatchBlock t = Tc[blockIdx.x];
    bx = t.bx;
    gx = t.gx;
    return Cc[t.c];
}

__global__ __launch_bounds__(BLOCK) void kb_inject_count(const BatchCase* __restrict__ C,
                                                         const BatchBlock* __restrict__ T, int ahead) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    const DevFaces F = c.faces;
    inject_count_body(c.P, F, c.ns, c.S, ahead, bx);
}

__global__ __launch_bounds__(BLOCK) void kb_inject_scan(const BatchCase* __restrict__ C,
                                                        const BatchBlock* __restrict__ T, int ahead) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    scan_single_body(c.faces.n_new, c.faces.off, c.faces.n * c.ns, ahead ? SCAN_APPEND_NEXT : SCAN_APPEND, c.S,
                     c.cap);
}

__global__ __launch_bounds__(BLOCK) void kb_inject_gen(const BatchCase* __restrict__ C,
                                                       const BatchBlock* __restrict__ T, int par, int ahead) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    const DevFaces F = c.faces;
    DevParticles b = c.buf[(c.cur0 ^ par ^ ahead) & 1];
    inject_gen_body(c.P, F, c.ns, b, c.S, ahead, bx);
}

// the occupancy of the single-engine k_move (at most 80 registers without the
// face tracker)
template <bool TRACK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TRACK ? 5 : 6, 8)))
void kb_move(const BatchCase* __restrict__ C, const BatchBlock* __restrict__ T, int par) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    DevParticles b = c.buf[(c.cur0 ^ par) & 1];
    const WallQueue Q = c.wq;
    move_body<TRACK>(c.P, b, c.S, c.count, c.slot, Q, 1, 1, bx, gx);
}

template <bool TRACK>
__global__ __launch_bounds__(BLOCK) void kb_move_wall(const BatchCase* __restrict__ C,
                                                      const BatchBlock* __restrict__ T, int par) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    DevParticles b = c.buf[(c.cur0 ^ par) & 1];
    const WallQueue Q = c.wq;
    move_wall_body<TRACK>(c.P, b, c.S, c.count, c.slot, Q, 1, bx, gx);
}

__global__ __launch_bounds__(BLOCK) void kb_scan_cells(const BatchCase* __restrict__ C,
                                                       const BatchBlock* __restrict__ T) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    (void)bx;
    scan_cells_body(c.count, c.start, c.nc, c.lb_status, c.lb_ticket, c.S, c.pt_g, c.tile_g, c.pt_c, c.tile_c,
                    c.heavy, gx);
}

__global__ __launch_bounds__(BLOCK) void kb_scatter(const BatchCase* __restrict__ C, const BatchBlock* __restrict__ T,
                                                    int par) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    scatter_body(c.buf[(c.cur0 ^ par) & 1].cell, c.S, c.start, c.slot, c.perm, c.heavy, c.pt_c, c.order, bx, gx);
}

template <bool GEN>
__global__ __launch_bounds__(CT) void kb_gather(const BatchCase* __restrict__ C, const BatchBlock* __restrict__ T,
                                                int par) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    DevParticles src = c.buf[(c.cur0 ^ par) & 1];
    DevParticles dst = c.buf[(c.cur0 ^ par ^ 1) & 1];
    gather_body<GEN>(c.P, src, dst, c.start, c.perm, c.tile_g, c.pt_g, c.S, c.sub, bx, c.slot);
}

template <bool GEN, int PMAX>
__global__ __launch_bounds__(CT, 2) void kb_collide(const BatchCase* __restrict__ C,
                                                    const BatchBlock* __restrict__ T, int par) {
    int bx, gx;
    const HS_CONST_AS BatchCase& c = batch_case(C, T, bx, gx);
    DevParticles dst = c.buf[(c.cur0 ^ par ^ 1) & 1];
    collide_body<GEN, PMAX>(c.P, dst, c.start, c.tile_c, c.order, c.pt_c, c.sub, c.sigma, c.rem, c.samples, c.vol,
                            c.S, c.do_collide, c.do_sample, c.nc, c.big_list, bx);
}

__global__ void kb_commit(const BatchCase* __restrict__ C, int ncase, int advance, int inject_next) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ncase) commit_body(C[e].S, advance, inject_next && C[e].with_faces);
}

void launch_batch_inject(const BatchCase* C, const BatchBlock* t_count, int n_count, const BatchBlock* t_scan,
                         int n_scan, const BatchBlock* t_gen, int n_gen, int par, int ahead, hipStream_t st) {
    if (n_count > 0) kb_inject_count<<<n_count, BLOCK, 0, st>>>(C, t_count, ahead);
    if (n_scan > 0) kb_inject_scan<<<n_scan, BLOCK, 0, st>>>(C, t_scan, ahead);
    if (n_gen > 0) kb_inject_gen<<<n_gen, BLOCK, 0, st>>>(C, t_gen, par, ahead);
}

void launch_batch_move(const BatchCase* C, const BatchBlock* t_move, int n_move, const BatchBlock* t_wall,
                       int n_wall, int par, int track, hipStream_t st) {
    if (n_move <= 0) return;
    if (track) {
        kb_move<true><<<n_move, BLOCK, 0, st>>>(C, t_move, par);
        kb_move_wall<true><<<n_wall, BLOCK, 0, st>>>(C, t_wall, par);
    } else {
        kb_move<false><<<n_move, BLOCK, 0, st>>>(C, t_move, par);
        kb_move_wall<false><<<n_wall, BLOCK, 0, st>>>(C, t_wall, par);
    }
}

void launch_batch_scan_cells(const BatchCase* C, const BatchBlock* t, int n, hipStream_t st) {
    if (n > 0) kb_scan_cells<<<n, BLOCK, 0, st>>>(C, t);
}

void launch_batch_scatter(const BatchCase* C, const BatchBlock* t, int n, int par, hipStream_t st) {
    if (n > 0) kb_scatter<<<n, BLOCK, 0, st>>>(C, t, par);
}

void launch_batch_gather(const BatchCase* C, const BatchBlock* t, int n, int par, int general, hipStream_t st) {
    if (n <= 0) return;
    if (general) kb_gather<true><<<n, CT, 0, st>>>(C, t, par);
    else kb_gather<false><<<n, CT, 0, st>>>(C, t, par);
}

void launch_batch_collide(const BatchCase* C, const BatchBlock* t, int n, int par, int general, hipStream_t st) {
    if (n <= 0) return;
    if (general) kb_collide<true, PMAX_GEN><<<n, CT, 0, st>>>(C, t, par);
    else kb_collide<false, PMAX_MONO><<<n, CT, 0, st>>>(C, t, par);
}

void launch_batch_commit(const BatchCase* C, int ncase, int advance, int inject_next, hipStream_t st) {
    if (ncase > 0) kb_commit<<<(ncase + 63) / 64, 64, 0, st>>>(C, ncase, advance, inject_next);
}
