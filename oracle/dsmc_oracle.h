/*
 * dsmc_oracle.h -- CPU restatement of hyStrath dsmcFoam+'s per-timestep
 * particle path, used ONLY as test infrastructure (parity checker and the
 * CPU baseline of the benchmark).  Nothing in the product links it.
 *
 * Each function cites the reference file:line it follows.  The random
 * stream is a counter-based Philox4x32-10 (Salmon et al., SC'11) keyed by
 * (seed, purpose, entity, step) so that the GPU engine can replay it exactly;
 * transcendental functions are this file's own FP64 routines built from
 * IEEE +,-,*,/ and sqrt only, so results are bit-reproducible across
 * compilers.  Build with -ffp-contract=off.
 */
#ifndef DSMC_ORACLE_H
#define DSMC_ORACLE_H

#include <stdint.h>
#include "../include/hystrath_dsmc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* random-stream purposes (Philox counter word 2, top byte) */
enum { OR_RNG_MOVE = 1, OR_RNG_COLLIDE = 2, OR_RNG_INJECT = 3,
       OR_RNG_FILL_COUNT = 4, OR_RNG_FILL_GEN = 5, OR_RNG_REMAINDER = 6 };

typedef struct or_state {
    int64_t n, capacity;
    hs_particles p;          /* current population (capacity-sized arrays) */
    hs_particles tmp;        /* sort scratch (capacity-sized)               */
    double* sigma_tcr_max;   /* [nc]                                        */
    double* remainder;       /* [nc]                                        */
    double* samples;         /* [slots][nc] or NULL                         */
    int64_t n_sampled;       /* nTimeSteps_                                  */
    double* inflow_acc;      /* [n_species][n_inflow_faces] or NULL          */
    int32_t* count;          /* [nc]                                        */
    int32_t* start;          /* [nc+1]                                      */
    uint32_t step;
    int64_t counters[8];     /* candidates, collisions, deleted, injected,
                                wall hits, track overflow, -, - */
    double* wall;            /* [n_species*HS_WS_COUNT][wall faces] or NULL;
                                summed by or_move when HS_FLAG_WALL and
                                HS_FLAG_SAMPLE are set                    */
    int64_t* face_flux;      /* [n_species][or_face_count] or NULL; counted
                                by or_move with HS_FLAG_FACE_FLUX and
                                HS_FLAG_SAMPLE                            */
    double* t_overall;       /* [nc] or NULL: each cell's overall temperature
                                for inverseZvFormulation 2008 (0 until set) */
} or_state;

/* primitives (exported for pinning tests) */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_uniform_from(uint32_t hi, uint32_t lo);
double or_log(double x);
double or_exp(double x);
double or_pow(double x, double y);
void or_sincos2pi(double u, double* s, double* c);
/* n draws of stream (seed, purpose, id, step, sub) into out */
void or_stream_draws(uint64_t seed, int purpose, uint32_t id, uint32_t step,
                     uint32_t sub, int n, double* out);
void or_stream_gauss(uint64_t seed, int purpose, uint32_t id, uint32_t step,
                     uint32_t sub, int n, double* out);

int32_t or_sample_slots(const hs_config* cfg);
int64_t or_inflow_face_count(const hs_config* cfg);

/* sigmaTcR (VariableHardSphere.C:80-121) */
int or_sigma_tcr(const hs_config* cfg, int64_t n, const int32_t* tp, const int32_t* tq,
                 const double* up, const double* uq, double* out);
/* BinaryCollisionModel::collide with injected uniforms (see header of the
 * C ABI, hs_dsmc_binary_collide) */
int or_binary_collide(const hs_config* cfg, int64_t n, const int32_t* tp, const int32_t* tq,
                      double* up, double* uq, double* erot_p, double* erot_q,
                      int32_t* vib_p, int32_t* vib_q, const double* uniforms,
                      int32_t n_uniform, int32_t* used_out, const double* t_macro);

/* The accepted collision of n independent pairs with an injected uniform
 * stream (pair k reads uniforms[k*n_uniform ..]): the pair's QK reaction
 * when chemReactDict gives it a model (dissociationQK.C:469-600,
 * exchangeQK.C:457-517, dissociationExchangeQK.C:98-270), else or when it
 * does not react the binary collision (noTimeCounter.C:250-300).  Types,
 * velocities [3k+c], rotational energies and vibrational levels
 * [k*HS_MAX_VIB_MODES+m] are in/out; reacted_out[k] = 1 if it reacted;
 * prod_type[k] = the dissociation product's typeId or -1, prod_u its
 * velocity, prod_from_q[k] = 1 if Q dissociated (0: P, -1: no product). */
int or_react_pairs(const hs_config* cfg, int64_t n, int32_t* tp, int32_t* tq, double* up, double* uq,
                   double* erot_p, double* erot_q, int32_t* vib_p, int32_t* vib_q, const double* uniforms,
                   int32_t n_uniform, int32_t* used_out, int32_t* reacted_out, int32_t* prod_type,
                   double* prod_u, int32_t* prod_from_q);

/* stages on an or_state */
int or_fill(const hs_config* cfg, or_state* s, const double* n, double t, double trot,
            double tvib, const double u[3], int32_t exact_per_cell, uint32_t fill_id);
/* the fill with per-cell number densities [species][cell], velocities
 * [axis][cell] and temperatures [cell] (see hs_dsmc_fill_cells) */
int or_fill_cells(const hs_config* cfg, or_state* s, const double* nd_cell, const double* u_cell,
                  const double* t_cell, uint32_t fill_id);
int or_inject(const hs_config* cfg, or_state* s);
int or_move(const hs_config* cfg, or_state* s, int64_t n_old);
/* wall faces of boundaryMeasurements in the engine's order; any output may be
 * NULL; returns the face count */
int64_t or_wall_faces(const hs_config* cfg, int32_t* side, int32_t* cell, double* area, double* normal);
/* faces of the face tracker: mesh faces by axis, then the cylinder's wall faces */
int64_t or_face_count(const hs_config* cfg);
/* length of the circle (cx, cy, r) inside [x0,x1]x[y0,y1]; n_out (may be
 * NULL) gets the length-weighted unit normal pointing to the centre */
double or_arc_in_rect(double cx, double cy, double r, double x0, double x1, double y0, double y1, double n_out[2]);
/* threads the stages use: 1, or OpenMP's count in the threaded build */
int or_threads(void);
/* sets the thread count of the threaded build (no effect in the checker) */
void or_set_threads(int n);
int or_sort(const hs_config* cfg, or_state* s);
int or_collide(const hs_config* cfg, or_state* s, int32_t sample);
/* the NTC candidate pairs (P, Q) that or_collide draws for candidates
 * 0..ncand-1 of `cell` at `step`, the cell holding the nC parcels at
 * (px, py, pz) in sorted order (noTimeCounter.C:114-198); test hook */
int or_ntc_pairs(const hs_config* cfg, int64_t cell, int32_t nC, const double* px, const double* py,
                 const double* pz, uint32_t step, int64_t ncand, int32_t* p_out, int32_t* q_out);
int or_evolve(const hs_config* cfg, or_state* s, int32_t nsteps);

#ifdef __cplusplus
}
#endif
#endif
