// dsmc_kernels.h -- host-side launchers of the engine's HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dsmc_device.h"

// Inflow face table (device arrays, one entry per inflow face).
struct DevFaces {
    int32_t* cell;
    int8_t* axis;
    int8_t* dir;
    int32_t* a1i;
    int32_t* a2i;
    double* inc;       // [species][face] per-step accumulator increment
    double* acc;       // [species][face] accumulatedParcelsToInsert_
    int32_t* n_new;    // [face*ns + s] parcels to insert this step
    int32_t* off;      // exclusive scan of n_new
    int32_t n;
    int32_t max_new;   // bound on parcels inserted in one step (launch size)
};

// Device scalars of one engine (one allocation).
struct DevScalars {
    int32_t n;        // current particle count
    int32_t n_old;    // count before this step's injection
    uint32_t step;    // time-step index (random-stream counter)
    int32_t total;    // scratch: last scan total
    int32_t n_sorted; // population after the cell scan (published by commit)
    int32_t n_next;   // n_sorted plus the next step's inflow when injected ahead (installed by commit)
    int32_t ord[2];   // heavy / light tile counters of the collide order (cleared by commit)
    int32_t qn;       // flights queued for the wall pass of the move (cleared by commit)
    int32_t nprod;    // dissociation products staged by the reacting collide (cleared after placement)
    int32_t nbigprod; // cells of more than 64 items with products, listed for their wave sums (cleared likewise)
    unsigned long long counters[CNT_N];
};

// scan workspace for up to n elements
struct ScanWs {
    int32_t* bsum;
    int32_t nb_max;
};

// what the scan writes besides the exclusive prefix
enum ScanMode {
    SCAN_PLAIN = 0,     // out[n] = total
    SCAN_CELLS = 1,     // out[n] = total, zero the input, S.n_sorted = total
    SCAN_APPEND = 2,    // S.n_old = S.n, S.n = min(S.n + total, capacity)
    SCAN_APPEND_NEXT = 3  // S.n_next = min(S.n_sorted + total, capacity): the next step's inflow
};

void launch_scan(const int32_t* in, int32_t* out, int32_t n, ScanMode mode, ScanWs& ws, DevScalars* S,
                 int64_t capacity, hipStream_t st);
void launch_inject_count(const DevParams* P, const DevFaces& F, int ns, DevScalars* S, int ahead, hipStream_t st);
void launch_inject_gen(const DevParams* P, const DevFaces& F, int ns, DevParticles b, DevScalars* S, int ahead,
                       hipStream_t st);
// Flights of the move stopped at a wall, for its wall pass: parcel index,
// flight time left, wall code | tracking iterations << 8 (capacity entries).
struct WallQueue {
    int32_t* i;
    double* t;
    int32_t* code_iter;
};
// move (+ the flights queued at walls) and/or the cell histogram
void launch_move(const DevParams* P, DevParticles b, DevScalars* S, int32_t* count, int32_t* slot,
                 const WallQueue& Q, int do_move, int do_hist, int track, int grid, hipStream_t st);
void launch_set_nold(DevScalars* S, hipStream_t st);
int move_blocks_per_cu(int track);
void launch_scatter(DevParticles b, DevScalars* S, const int32_t* start, const int32_t* slot, int32_t* perm,
                    int32_t* heavy, int pt, int32_t* order, int grid, hipStream_t st);
// particle-balanced tiles of pt parcels (tile_begin) and their heavy flags
// (clear before the first use; the collide order clears the flags it reads)
int scan_cells_blocks(int32_t n);
void launch_scan_cells(int32_t* cnt, int32_t* start, int32_t n, unsigned long long* status,
                       unsigned long long* ticket, DevScalars* S, int pt_g, int32_t* tile_g,
                       int pt_c, int32_t* tile_c, int32_t* heavy, hipStream_t st);
void launch_tiles(const int32_t* start, int nc, int pt_g, int32_t* tile_begin_g, int pt_c, int32_t* tile_begin_c,
                  int32_t* heavy, hipStream_t st);
// stable cell order: rank + gather of src into dst in cell order, sub-cell byte per parcel;
// scratch: capacity 32-bit words free during the gather (the histogram slots)
void launch_gather(const DevParams* P, DevParticles src, DevParticles dst, const int32_t* start, const int32_t* perm,
                   const int32_t* tile_begin, int pt, int64_t capacity, const DevScalars* S, uint8_t* sub,
                   int general, int32_t* scratch, hipStream_t st);
// NTC, collisions and sampling on the sorted population in dst
void launch_collide(const DevParams* P, DevParticles dst, const int32_t* start, const int32_t* tile_begin,
                    int32_t* heavy, int32_t* order, int pt, int64_t capacity, uint8_t* sub, double* sigma,
                    double* rem, double* samples, const double* vol, DevScalars* S, int do_collide, int do_sample,
                    int general, int nc, uint16_t* big_list, hipStream_t st);
// Parcels made by dissociations in one collide (dsmcCloud::addNewParcel):
// staged in any order with their cell, their rank among the cell's
// products and the dissociating parcel, then placed after the sorted
// population in (cell, rank) order.  cap entries each.
struct DevProducts {
    int32_t* cell;
    int32_t* rank;
    int32_t* parent;  // sorted index of the parcel that dissociated
    uint8_t* type;
    float* ux;
    float* uy;
    float* uz;
    int32_t* count;   // [nc] products per cell
    int32_t* off;     // [nc+1] exclusive prefix of count
    int32_t* list;    // [cap] 32-bit sub-cell member lists of cells of more than 65,535 parcels
    int32_t* bigcells;  // [nc] cells of more than 64 items with products (their sums by a wave)
};
// NTC with QK chemistry on the sorted population in dst (the collide tiles
// of launch_collide, the variant with internal energies); then the products'
// prefix, placement, the sums of the cells that made products (parcels, then
// products) and the population update.
void launch_collide_react(const DevParams* P, DevParticles dst, const int32_t* start, const int32_t* tile_begin,
                          int32_t* order, int pt, uint8_t* sub, double* sigma, double* rem, double* samples,
                          const double* vol, DevScalars* S, DevProducts pr, int do_collide, int do_sample, int nc,
                          int ns, uint16_t* big_list, ScanWs& ws, int64_t capacity, hipStream_t st);
void launch_clear_queue(DevScalars* S, hipStream_t st);
void launch_commit(DevScalars* S, int advance, int with_next, hipStream_t st);
// per-cell fill inputs (NULL: the uniform values): number densities
// [species][cell], velocities [axis][cell], temperatures [cell]
struct FillCells {
    const double* nd;
    const double* u;
    const double* t;
};
void launch_fill_count(const DevParams* P, const double* nd, int exact, uint32_t fill_id, const double* vol,
                       int32_t* cnt, int nc, int ns, FillCells fc, hipStream_t st);
void launch_fill_gen(const DevParams* P, const double* nd, double T, double Trot, double Tvib, const double* u3,
                     uint32_t fill_id, const int32_t* cnt, const int32_t* off, int nc, int ns, DevParticles b,
                     DevScalars* S, FillCells fc, hipStream_t st);
void launch_fill_cells(const DevParams* P, double smax, uint32_t fill_id, double* sigma, double* rem, int nc,
                       const double* sig_cell, hipStream_t st);
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);
void launch_fill_f64(double* p, double v, int64_t n, hipStream_t st);
void launch_fill_i64(int64_t* p, int64_t v, int64_t n, hipStream_t st);
// 1/Zv by (species, partner, mode, imax < HS_ZV_TAB) into tab (ns*ns*HS_MAX_VIB_MODES*HS_ZV_TAB)
void launch_zv_table(const DevParams* P, double* tab, int ns, hipStream_t st);
void launch_binary_collide(const DevParams* P, int64_t n, const int32_t* tp, const int32_t* tq, double* up,
                           double* uq, double* erp, double* erq, int32_t* vp, int32_t* vq, const double* uni,
                           int32_t nu, int32_t* used, hipStream_t st);
void launch_sigma_tcr(const DevParams* P, int64_t n, const int32_t* tp, const int32_t* tq, const double* up,
                      const double* uq, double* out, hipStream_t st);
void launch_copy(const void* a, void* b, int64_t bytes, int grid, hipStream_t st);

// ---------------------------------------------------------------------------
// batched launches: one launch per stage over many engines (cases)
// ---------------------------------------------------------------------------
// One case's device state as the stage kernels take it.  buf[cur0 ^ par] is
// the current population of a step at parity par, the other its next buffer.
struct BatchCase {
    const DevParams* P;
    DevParticles buf[2];
    int32_t cur0;
    int32_t nc, ns;
    int32_t pt_g, pt_c;         // gather and collide tile parcels
    int32_t do_collide, do_sample, with_faces;
    int64_t cap;
    DevScalars* S;
    int32_t* count;
    int32_t* start;
    int32_t* slot;
    int32_t* perm;
    uint8_t* sub;
    uint16_t* big_list;
    int32_t* tile_g;
    int32_t* tile_c;
    int32_t* heavy;
    int32_t* order;
    double* sigma;
    double* rem;
    double* samples;
    const double* vol;
    unsigned long long* lb_status;
    unsigned long long* lb_ticket;
    WallQueue wq;
    DevFaces faces;
};

// Workgroup w of a batched launch is workgroup bx of gx of case c.
struct BatchBlock {
    int32_t c, bx, gx, pad;
};

// The stages of a step over the cases of C (tables T_* of n_* workgroups;
// a stage with n = 0 launches nothing).  general: the cases carry internal
// energies; track: the face tracker is on.
void launch_batch_inject(const BatchCase* C, const BatchBlock* t_count, int n_count, const BatchBlock* t_scan,
                         int n_scan, const BatchBlock* t_gen, int n_gen, int par, int ahead, hipStream_t st);
void launch_batch_move(const BatchCase* C, const BatchBlock* t_move, int n_move, const BatchBlock* t_wall,
                       int n_wall, int par, int track, hipStream_t st);
void launch_batch_scan_cells(const BatchCase* C, const BatchBlock* t, int n, hipStream_t st);
void launch_batch_scatter(const BatchCase* C, const BatchBlock* t, int n, int par, hipStream_t st);
void launch_batch_gather(const BatchCase* C, const BatchBlock* t, int n, int par, int general, hipStream_t st);
void launch_batch_collide(const BatchCase* C, const BatchBlock* t, int n, int par, int general, hipStream_t st);
void launch_batch_commit(const BatchCase* C, int ncase, int advance, int inject_next, hipStream_t st);
