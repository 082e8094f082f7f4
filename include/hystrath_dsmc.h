/*
 * hystrath_dsmc.h -- C ABI of the MI355X particle engine for dsmcFoam+.
 *
 * This is the drop-in boundary for dsmcCloud's per-timestep particle path
 * (src/lagrangian/dsmc/clouds/dsmcCloud.C:819-926).  Every entry point takes
 * plain pointers and sizes; no C++ or torch types cross it.  One engine owns
 * one HIP stream and all of its device memory; calls on one engine must not
 * overlap, independent engines are thread-safe.
 *
 * Every function returns 0 on success and a negative HS_E* code on failure;
 * hs_dsmc_last_error() then gives a message.  Host buffers passed in are only
 * borrowed for the duration of the call.
 *
 * Mesh: a uniform Cartesian hex block (what a single-block blockMeshDict
 * "hex (...) (nx ny nz) simpleGrading (1 1 1)" produces), cells numbered
 * i + nx*(j + ny*k) as blockMesh numbers them, six outer patches and an
 * optional embedded cylinder wall (axis z).  Cell volumes are an input so
 * cut cells around the cylinder carry their fluid volume.
 */
#ifndef HYSTRATH_DSMC_H
#define HYSTRATH_DSMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HS_MAX_SPECIES   8
#define HS_MAX_VIB_MODES 3

/* Boltzmann constant used everywhere (CODATA 2014, the value of the
 * OpenFOAM-v1706 era physicoChemical::k). */
#define HS_KB 1.38064852e-23

/* ---- error codes -------------------------------------------------------- */
#define HS_OK             0
#define HS_E_INVALID    (-1)   /* invalid argument / configuration          */
#define HS_E_CAPACITY   (-2)   /* particle capacity exceeded                */
#define HS_E_HIP        (-3)   /* HIP runtime error                         */
#define HS_E_OVERFLOW   (-4)   /* a device-side bound tripped (see flags)    */
#define HS_E_UNSUPPORTED (-5)  /* model/option not implemented              */

/* ---- runtime-selection enums (dictionary names in comments) ------------- */
/* BinaryCollisionModel (BinaryCollisionModel.C:59-90) */
enum hs_collision_model {
    HS_COLL_NONE   = 0,  /* NoBinaryCollision                   */
    HS_COLL_VHS    = 1,  /* VariableHardSphere                  */
    HS_COLL_VSS    = 2,  /* VariableSoftSphere                  */
    HS_COLL_LB_VHS = 3,  /* LarsenBorgnakkeVariableHardSphere   */
    HS_COLL_LB_VSS = 4   /* LarsenBorgnakkeVariableSoftSphere   */
};

/* dsmcPatchBoundaries / dsmcCyclicBoundaries / empty patches */
enum hs_patch_model {
    HS_PATCH_CYCLIC   = 0,  /* cyclic patch pair (particleTemplates.C:1525)  */
    HS_PATCH_SPECULAR = 1,  /* dsmcSpecularWallPatch                          */
    HS_PATCH_DIFFUSE  = 2,  /* dsmcDiffuseWallPatch                           */
    HS_PATCH_DELETION = 3,  /* dsmcDeletionPatch                              */
    HS_PATCH_EMPTY    = 4   /* empty (reduced-D direction, no tracking)       */
};

/* config.flags */
#define HS_FLAG_SAMPLE        0x1  /* dsmcVolFields sampling every step      */
#define HS_FLAG_HEATFLUX      0x2  /* measureHeatFluxShearStress sums         */
#define HS_FLAG_COLL_SEP      0x4  /* collision separation (meanCollisionSeparation) */
#define HS_FLAG_TIMING        0x8  /* per-kernel HIP event timing             */
#define HS_FLAG_WALL         0x10  /* boundaryMeasurements at diffuse/specular walls */
#define HS_FLAG_FACE_FLUX    0x20  /* dsmcFaceTracker parcel fluxes through every face */

/* sides of the block, patch[] order */
enum hs_side { HS_XMIN = 0, HS_XMAX = 1, HS_YMIN = 2, HS_YMAX = 3, HS_ZMIN = 4, HS_ZMAX = 5 };

/* dsmcParcel::constantProperties (dsmcParcelI.H:37-200) */
typedef struct hs_species {
    double mass;                       /* mass                                 */
    double diameter;                   /* diameter                             */
    double omega;                      /* omega (viscosity index)              */
    double alpha;                      /* alpha (VSS), default 1               */
    double rot_dof;                    /* rotationalDegreesOfFreedom           */
    int32_t n_vib_modes;               /* nVibrationalModes (<= HS_MAX_VIB_MODES) */
    int32_t charge;                    /* charge (-1 electron, 0, 1)           */
    double theta_v[HS_MAX_VIB_MODES];  /* characteristicVibrationalTemperature */
    double zref[HS_MAX_VIB_MODES];     /* Zref                                 */
    double tref_zv[HS_MAX_VIB_MODES];  /* referenceTempForZref                 */
    double theta_d;                    /* dissociationTemperature              */
} hs_species;

/* one patch: model + its Properties sub-dictionary + optional
 * dsmcFreeStreamInflowPatch general boundary on the same patch */
typedef struct hs_patch {
    int32_t model;                     /* hs_patch_model                       */
    int32_t inflow;                    /* 1: dsmcFreeStreamInflowPatch here     */
    double temperature;                /* dsmcDiffuseWallPatchProperties/temperature */
    double velocity[3];                /* dsmcDiffuseWallPatchProperties/velocity    */
} hs_patch;

/* ---- chemical reactions (dsmcReactions, system/chemReactDict) ----------- */
/* dsmcReaction::New selects by reactionModel (dsmcReaction.C:106-135); the
 * quantum-kinetic models of src/lagrangian/dsmc/reactions/derived: */
enum hs_reaction_model {
    HS_REACT_DISSOCIATION_QK = 1,          /* dissociationQK (dissociationQK.C:199-600)              */
    HS_REACT_EXCHANGE_QK = 2,              /* exchangeQK (exchangeQK.C:180-517)                      */
    HS_REACT_DISSOCIATION_EXCHANGE_QK = 3  /* dissociationExchangeQK (dissociationExchangeQK.C:98-270) */
};
#define HS_MAX_REACTIONS 32

/* One entry of chemReactDict's reactions list.  A species' kind follows
 * dsmcParcelI.H:153-196: charge -1 electron, else atom (no vibrational
 * mode), diatomic (one) or polyatomic molecule.  At most one reaction per
 * unordered pair of typeIds (dsmcReactions.C:139-160).  Requires the
 * VariableHardSphere or LarsenBorgnakkeVariableHardSphere collision model. */
typedef struct hs_reaction {
    int32_t model;                 /* hs_reaction_model                              */
    int32_t reactants[2];          /* reactants (typeIds)                             */
    int32_t diss_products[2][2];   /* dissociationProducts: two typeIds per reactant,
                                      -1 -1 for the empty list of an atom            */
    int32_t exch_products[2];      /* exchangeProducts: a molecule and an atom, either order */
    double heat_exchange;          /* heatOfReactionExchange (K; positive: exothermic) */
    double a_coeff, b_coeff;       /* aCoeff, bCoeff of the exchange                  */
    int32_t allow_splitting;       /* allowSplitting (the reference's default is yes) */
} hs_reaction;

typedef struct hs_config {
    /* ---- mesh ---- */
    int32_t nx, ny, nz;
    double origin[3];                  /* lower corner                          */
    double delta[3];                   /* cell size per axis                    */
    const double* cell_volume;         /* nx*ny*nz fluid volumes (host, borrowed at create) */
    hs_patch patch[6];                 /* hs_side order                         */
    int32_t cyl_enabled;               /* embedded cylinder wall, axis z        */
    double cyl_center[2];
    double cyl_radius;
    hs_patch cyl_patch;                /* SPECULAR or DIFFUSE                   */

    /* ---- species table ---- */
    int32_t n_species;
    hs_species species[HS_MAX_SPECIES];

    /* ---- BinaryCollisionModel + Coeffs ---- */
    int32_t collision_model;           /* hs_collision_model                    */
    double tref;                       /* <Model>Coeffs/Tref (default 273)      */
    double z_rot;                      /* rotationalRelaxationCollisionNumber (5) */
    double z_vib_fixed;                /* vibrationalRelaxationCollisionNumber (0 = Bird 2010 Zv) */
    double z_el;                       /* electronicRelaxationCollisionNumber (500) */
    int32_t inv_zv_form;               /* inverseZvFormulation: 0 pre-2008, 1 2008 (the cells' overall
                                          temperature, hs_dsmc_set_cell_temperature), 2 default */

    /* ---- dsmcFreeStreamInflowPatch properties (shared by inflow patches) ---- */
    double inflow_n[HS_MAX_SPECIES];   /* numberDensities per species           */
    double inflow_t;                   /* translationalTemperature              */
    double inflow_trot;                /* rotationalTemperature                 */
    double inflow_tvib;                /* vibrationalTemperature                */
    double inflow_u[3];                /* velocity                              */

    /* ---- numerics ---- */
    double fn;                         /* nEquivalentParticles                  */
    double dt;                         /* deltaT                                */
    uint64_t seed;                     /* seedNumber                            */
    int64_t capacity;                  /* max simulator particles               */
    int32_t flags;                     /* HS_FLAG_*                             */

    /* ---- chemReactDict (host, borrowed at create) ---- */
    const hs_reaction* reactions;      /* n_reactions entries, or NULL          */
    int32_t n_reactions;               /* 0: no chemistry                       */
} hs_config;

/* Per-particle state, structure of arrays.  erot is used when any species has
 * rotational DoF, vib (n_vib_modes planes of length n) when any species has
 * vibrational modes, type when n_species > 1; pass NULL otherwise. */
typedef struct hs_particles {
    double* px; double* py; double* pz;    /* FP64 positions                    */
    float*  ux; float*  uy; float*  uz;    /* FP32 velocities                   */
    int32_t* cell;                         /* owning cell                       */
    float*  erot;                          /* rotational energy (J)             */
    uint16_t* vib;                         /* vibLevel, plane-major [mode][n]   */
    uint8_t* type;                         /* typeId                            */
} hs_particles;

/* Per-cell sampling accumulators (dsmcVolFields.C:1115-1239), cumulative
 * over sampled steps, plane-major FP64 arrays [slot][cell].  Sums are kept
 * per species so every dsmcVolFields model (its typeIds subset, e.g. "N2" or
 * "mixture") is a sum over its species.  Slots 0..1 are per cell: */
enum hs_sample_slot {
    HS_S_NCOLL = 0,    /* nColls (cellMeasurements), all species           */
    HS_S_COLLSEP = 1,  /* collision separation sum                          */
    HS_S_GLOBAL_COUNT = 2
};
/* then species s occupies slots HS_S_GLOBAL_COUNT + s*HS_SS_COUNT + k: */
enum hs_species_slot {
    HS_SS_N = 0,       /* dsmcNSpeciesCum: simulator parcels                */
    HS_SS_MUX, HS_SS_MUY, HS_SS_MUZ, /* sum m U                              */
    HS_SS_MCC,         /* sum m |U|^2 (linear KE without the 1/2)           */
    HS_SS_EROT,        /* sum ERot                                          */
    HS_SS_EVIB0,       /* sum of vibrational energy per mode (3 slots)      */
    HS_SS_COUNT = HS_SS_EVIB0 + HS_MAX_VIB_MODES
};
/* and with HS_FLAG_HEATFLUX, after all species blocks, species s has 13
 * slots: m uu, m uv, m uw, m vv, m vw, m ww, m|U|^2 u, m|U|^2 v, m|U|^2 w,
 * Eint u, Eint v, Eint w, Eint (measureHeatFluxShearStress,
 * dsmcVolFields.C:1190-1210). */
#define HS_HF_COUNT 13
int32_t hs_dsmc_sample_slots(const hs_config* cfg);

/* Wall-face sampling with HS_FLAG_WALL: the per-hit sums of
 * dsmcPatchBoundary::measurePropertiesBeforeControl / AfterControl
 * (dsmcPatchBoundary.C:263-483) that dsmcVolFields accumulates into its
 * boundary fields (dsmcVolFields.C:1292-1356), cumulative over sampled steps,
 * plane-major FP64 [species*HS_WS_COUNT + slot][wall face].  Every hit adds
 * a "before" and an "after" term weighted by w = 1/max(|U.n| A dt, 1e-15)
 * (A = face area, n = unit normal pointing out of the gas); q and fD add
 * Fn*(pre - post)/(dt A).  Wall faces are measured on diffuse and specular
 * walls (dsmcDiffuseWallPatch.C:167, dsmcSpecularWallPatch.C:81), not on
 * deletion patches (dsmcDeletionPatch.C:115). */
enum hs_wall_slot {
    HS_WS_RHON = 0,    /* sum w                      (speciesRhoNBF)          */
    HS_WS_KE,          /* sum 1/2 m|U|^2 w           (speciesLinearKEBF; speciesMccBF = 2x) */
    HS_WS_MOMX, HS_WS_MOMY, HS_WS_MOMZ, /* sum m U_t w, U_t = U - (U.n)n (speciesMomentumBF) */
    HS_WS_EROT,        /* sum ERot w                 (speciesErotBF)          */
    HS_WS_EVIB0,       /* sum eVib_m w per mode, 3 slots (speciesEvibModBF)   */
    HS_WS_Q = HS_WS_EVIB0 + HS_MAX_VIB_MODES, /* sum Fn (E_pre - E_post)/(dt A) (speciesqBF) */
    HS_WS_FDX, HS_WS_FDY, HS_WS_FDZ,          /* sum Fn m (U_pre - U_post)/(dt A) (speciesfDBF) */
    HS_WS_COUNT
};

/* ---- engine lifetime ---------------------------------------------------- */
typedef struct hs_dsmc_engine hs_dsmc_engine;

/* Replaces dsmcCloud::dsmcCloud(Time&, word, dynamicFvMesh&, bool)
 * (dsmcCloud.C:586-691): species table, models, sigmaTcRMax, remainder. */
int hs_dsmc_create(const hs_config* cfg, hs_dsmc_engine** out);
void hs_dsmc_destroy(hs_dsmc_engine* eng);
const char* hs_dsmc_last_error(const hs_dsmc_engine* eng);
/* message of the last failed hs_dsmc_create (thread-local) */
const char* hs_dsmc_create_error(void);
int hs_dsmc_device_count(int* n);

/* ---- state transfer ----------------------------------------------------- */
/* Replaces dsmcParcel::readFields (dsmcParcelIO.C:133-337) for the SoA. */
int hs_dsmc_upload_particles(hs_dsmc_engine* eng, const hs_particles* p, int64_t n);
/* Replaces dsmcParcel::writeFields (dsmcParcelIO.C:338-519).  n_out gets the
 * particle count; fails with HS_E_CAPACITY if it exceeds cap. */
int hs_dsmc_download_particles(hs_dsmc_engine* eng, hs_particles* p, int64_t cap, int64_t* n_out);
int hs_dsmc_particle_count(hs_dsmc_engine* eng, int64_t* n_out);
/* sigmaTcRMax_ and collisionSelectionRemainder_ (dsmcCloud.H:144-147);
 * either pointer may be NULL. */
int hs_dsmc_set_cell_state(hs_dsmc_engine* eng, const double* sigma_tcr_max, const double* remainder);
int hs_dsmc_get_cell_state(hs_dsmc_engine* eng, double* sigma_tcr_max, double* remainder);
/* Each cell's overall temperature (nc values), which inverseZvFormulation
 * 2008 puts in the vibrational collision number in place of the collision
 * temperature wherever it exceeds SMALL: replaces fields().overallT(cellI)
 * (dsmcCloud.C:1440-1455), the Tov of the first dsmcVolFields
 * (dsmcFieldProperties.C:235-244, dsmcVolFields.H:244-247), which that field
 * recomputes at each write (dsmcVolFields.C:1365, 1500-1508).  Zero at
 * creation; HS_E_INVALID unless the engine has inv_zv_form 1 and a
 * Larsen-Borgnakke model. */
int hs_dsmc_set_cell_temperature(hs_dsmc_engine* eng, const double* t_overall);
/* freestream inflow accumulators accumulatedParcelsToInsert_
 * (dsmcFreeStreamInflowPatch.C:132-140), [species][inflow face]; values
 * lie in [-1, 1) between steps and the setter rejects others */
int hs_dsmc_inflow_face_count(hs_dsmc_engine* eng, int64_t* n_faces);
int hs_dsmc_set_inflow_accumulators(hs_dsmc_engine* eng, const double* acc);
int hs_dsmc_get_inflow_accumulators(hs_dsmc_engine* eng, double* acc);
/* time-step index used as the random-stream counter */
int hs_dsmc_set_step(hs_dsmc_engine* eng, uint32_t step);
int hs_dsmc_get_step(hs_dsmc_engine* eng, uint32_t* step);

/* ---- initial state ------------------------------------------------------ */
/* Replaces dsmcMeshFill::setInitialConfiguration (dsmcMeshFill.C:68-~230):
 * Poisson-rounded uniform fill of every cell with a Maxwellian at
 * (t, trot, tvib, u) and number densities n[species]; when exact_per_cell
 * is non-zero each cell gets exactly round(n*V/Fn) particles per species.
 * Appends to the current population.  Also resets sigmaTcRMax to
 * sigmaT*sqrt(2kT/m) of the most abundant species and draws the remainder
 * (dsmcCloud.C:105-116). */
int hs_dsmc_fill(hs_dsmc_engine* eng, const double* n, double t, double trot,
                 double tvib, const double u[3], int32_t exact_per_cell,
                 uint32_t fill_id);

/* The fill with each cell's own state: n_cell[species*ncells + cell] number
 * densities, u_cell[axis*ncells + cell] velocities and t_cell[cell]
 * temperatures (translational, rotational and vibrational alike); counts
 * are Poisson-rounded as in hs_dsmc_fill and sigmaTcRMax starts at each
 * cell's sigmaT*sqrt(2kT/m) of its most abundant species.  The reference
 * initialises uniform states (dsmcMeshFill.C:68-~230) or maps whole
 * solutions (mapFields); this is the start of a population from predicted
 * fields, the learned warm start's particle initialisation. */
int hs_dsmc_fill_cells(hs_dsmc_engine* eng, const double* n_cell, const double* u_cell,
                       const double* t_cell, uint32_t fill_id);

/* The same fill with no allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops re-seeding populations, harnesses that forbid
 * hipMalloc).  Its scratch (per cell-species counts and offsets) lives in a
 * fixed device array of the library, one per GPU, held until the process
 * ends; it holds at most HS_FILL_NOALLOC_MAX_PAIRS cell-species pairs
 * (nx*ny*nz*n_species), beyond which the call fails with HS_E_INVALID.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU, from
 * engines used on different threads, race.  Results equal hs_dsmc_fill's. */
#define HS_FILL_NOALLOC_MAX_PAIRS 500000
int hs_dsmc_fill_noalloc(hs_dsmc_engine* eng, const double* n, double t, double trot,
                         double tvib, const double u[3], int32_t exact_per_cell,
                         uint32_t fill_id);

/* ---- the timestep ------------------------------------------------------- */
/* Replaces dsmcCloud::evolve (dsmcCloud.C:819-926) nsteps times: freestream
 * injection (dsmcFreeStreamInflowPatch::controlParcelsBeforeMove), move
 * (Cloud<dsmcParcel>::move, Cloud.C:202-466), buildCellOccupancy
 * (dsmcCloud.C:63-74), noTimeCounter::collide (noTimeCounter.C:81-339) and
 * dsmcVolFields::calculateField sums (dsmcVolFields.C:1071-1290). */
int hs_dsmc_evolve(hs_dsmc_engine* eng, int32_t nsteps);
/* Evolves several independent engines nsteps each, as hs_dsmc_evolve would
 * one after the other, with every engine's result the same as when it runs
 * alone.  The engines' steps are enqueued interleaved on their own streams, so
 * small cases share the device.  No counterpart in the reference, which runs
 * one case per process: the batch form for parameter sweeps and ensembles.
 * HS_E_INVALID for a NULL or repeated engine; on a failure the message is in
 * the failing engine (NULL engine: hs_dsmc_create_error), and engines before
 * it in the batch may have advanced further. */
int hs_dsmc_evolve_batch(hs_dsmc_engine* const* engines, int32_t count, int32_t nsteps);
/* A batch of engines stepped with one launch per stage: every stage of a step
 * runs as one kernel launch over the workgroups of all the batch's cases (per
 * kernel variant: cases with internal energies and cases with the face
 * tracker on take their own launches), so small cases fill the device
 * together instead of one after another.  Every engine's result is the same
 * as when it runs alone.  No counterpart in the reference, which runs one
 * case per process: the batch form for parameter sweeps and ensembles.
 * The engines must be on one device and outlive the batch; between batch
 * evolves they may be used on their own (their streams are ordered with the
 * batch's).  Errors: hs_dsmc_batch_last_error (NULL batch: the create error).
 * hs_dsmc_batch_evolve is switched off in this build and fails with
 * HS_E_INVALID for nsteps > 0: its first device run crashed the host process
 * and the cause is not identified; hs_dsmc_evolve_batch steps batches. */
typedef struct hs_dsmc_batch hs_dsmc_batch;
int hs_dsmc_batch_create(hs_dsmc_engine* const* engines, int32_t count, hs_dsmc_batch** out);
int hs_dsmc_batch_evolve(hs_dsmc_batch* batch, int32_t nsteps);
void hs_dsmc_batch_destroy(hs_dsmc_batch* batch);
const char* hs_dsmc_batch_last_error(const hs_dsmc_batch* batch);
/* Stage entry points (same kernels, for tests and custom loops). */
int hs_dsmc_inject(hs_dsmc_engine* eng);                 /* controlParcelsBeforeMove */
int hs_dsmc_move(hs_dsmc_engine* eng);                   /* Cloud<dsmcParcel>::move  */
int hs_dsmc_sort(hs_dsmc_engine* eng);                   /* buildCellOccupancy       */
int hs_dsmc_collide(hs_dsmc_engine* eng, int32_t sample); /* noTimeCounter::collide (+ sampling) */
int hs_dsmc_advance_step(hs_dsmc_engine* eng);           /* step counter += 1 */
int hs_dsmc_synchronize(hs_dsmc_engine* eng);
/* Switch HS_FLAG_SAMPLE / HS_FLAG_COLL_SEP / HS_FLAG_TIMING / HS_FLAG_WALL /
 * HS_FLAG_FACE_FLUX between steps
 * (the sampling schedule of timeDataMeas, timeDataMeas.C:71-104);
 * HS_FLAG_HEATFLUX can only be kept as it was at create time. */
int hs_dsmc_set_flags(hs_dsmc_engine* eng, int32_t flags);
/* Parcels per gather / collide workgroup (whole cells are grouped up to this
 * many).  No counterpart in the reference, whose cells are processed one by
 * one: a tuning knob for small populations.  0 (the default) chooses from the
 * population last seen by the host (fill, upload, counts): the full-device
 * tiles (1536 / 2048, 1536 with internal energy) for large populations,
 * halved down to 256 until there are about four workgroups per compute unit.
 * Results do not depend on the tiling.  hs_dsmc_get_tiles reports the sizes
 * the next step uses. */
int hs_dsmc_set_tiles(hs_dsmc_engine* eng, int32_t gather_parcels, int32_t collide_parcels);
int hs_dsmc_get_tiles(hs_dsmc_engine* eng, int32_t* gather_parcels, int32_t* collide_parcels);
/* The autotuner of the batched mode; no counterpart in the reference.  Steps
 * the engine (2 + steps_per_trial steps per candidate, ordinary steps of the
 * run, since the tiling changes no result) under each candidate tiling --
 * gather and collide parcels per workgroup of 256, 512, 1024, 1536 and 2048,
 * up to the full-device sizes, skipping tilings with fewer workgroups than
 * compute units -- and times them with HIP events.  The fastest pair is
 * returned and kept for every engine of the process with the same kernel
 * variant (internal energies or not) and population class (log2 of the
 * parcel count) that has no tiles set by hs_dsmc_set_tiles; this engine's
 * own setting is left as it was. */
int hs_dsmc_autotune(hs_dsmc_engine* eng, int32_t steps_per_trial, int32_t* gather_out,
                     int32_t* collide_out, double* ms_out);

/* ---- sampling ----------------------------------------------------------- */
/* dsmcVolFields cumulative sums, plane-major [slot][cell]; nsteps_out gets
 * the number of sampled steps (nTimeSteps_). */
int hs_dsmc_read_samples(hs_dsmc_engine* eng, double* out, int64_t n_doubles, int64_t* nsteps_out);
/* dsmcVolFields::resetField (dsmcVolFields.C:2270-2290); clears the cell
 * and wall-face sums and the face fluxes */
int hs_dsmc_reset_samples(hs_dsmc_engine* eng);
/* Restores cumulative sums and their step count, as a restarted run needs
 * them: dsmcVolFields reads its accumulated data back when it continues an
 * average (averagingAcrossManyRuns, dsmcVolFields.C:1047-1063).  cell_sums is
 * [slot][cell] as hs_dsmc_read_samples gives it (exactly slots*cells values);
 * wall_sums as hs_dsmc_read_wall_samples gives it, or NULL to clear the wall
 * sums.  The face fluxes are left as they are. */
int hs_dsmc_set_samples(hs_dsmc_engine* eng, const double* cell_sums, int64_t n_cell,
                        const double* wall_sums, int64_t n_wall, int64_t nsteps);
/* Wall faces (the boundary faces of boundaryMeasurements, boundaryMeasurements.H:74-90):
 * first every face of each diffuse or specular side, sides in hs_side order,
 * faces ordered by the two tangential cell indices (x before y before z, the
 * first one fastest); then one face per cell that the embedded cylinder wall
 * crosses (arc longer than 1e-6 of the cell size), in cell order, with
 * area = arc length inside the cell times the cell's z size and the
 * area-weighted normal.  Any output may be NULL:
 * side[f] = hs_side or 6 for the cylinder, cell[f] = adjacent cell,
 * area[f], normal[3*f+a] = unit normal pointing out of the gas. */
int hs_dsmc_wall_face_count(hs_dsmc_engine* eng, int64_t* n_faces);
int hs_dsmc_wall_faces(hs_dsmc_engine* eng, int32_t* side, int32_t* cell, double* area, double* normal);
/* [species*HS_WS_COUNT + slot][wall face] sums; nsteps_out as above */
int hs_dsmc_read_wall_samples(hs_dsmc_engine* eng, double* out, int64_t n_doubles, int64_t* nsteps_out);
/* dsmcFaceTracker::trackFaceTransition (dsmcFaceTracker.C:124-193), with
 * HS_FLAG_FACE_FLUX (on sampled steps): parcelIdFlux per species and face,
 * cumulative since the last hs_dsmc_reset_samples (the reference clears it
 * every step, dsmcFaceTracker.C:86-97; a consumer sums it over steps).
 * Every face crossing adds sign(U.Sf): Sf points +axis on internal faces
 * and out of the domain on boundary faces; U is the velocity after a wall
 * or deletion interaction; a cyclic crossing adds +1 on the coupled face of
 * the other side.  massIdFlux is mass[species] times the count.  Faces:
 * for axis a (x, y, z) the planes k = 0..n_a, face id
 * off_a + k + (n_a+1)*(j1 + n_a1*j2) with (j1, j2) the tangential cell
 * indices as for wall faces, off_x = 0, off_y = (nx+1)*ny*nz,
 * off_z = off_y + nx*(ny+1)*nz; then the cylinder's wall faces in wall-face
 * order.  out is int64 [species][face]. */
int hs_dsmc_face_count(hs_dsmc_engine* eng, int64_t* n_faces);
int hs_dsmc_read_face_flux(hs_dsmc_engine* eng, int64_t* out, int64_t n, int64_t* nsteps_out);
/* counters since the last reset: [0] candidates, [1] collisions,
 * [2] deleted, [3] injected, [4] wall hits, [5] tracking-iteration
 * overflows, [6] oversize cells / candidate clamps, [7] particles dropped at
 * capacity, [8] particle-steps moved */
int hs_dsmc_read_counters(hs_dsmc_engine* eng, int64_t* out, int32_t n);
int hs_dsmc_reset_counters(hs_dsmc_engine* eng);
/* per-kernel device time (ms, summed over launches since last reset) and
 * launch counts, in the order inject, move, scan, scatter, collide, gather
 * (the stable cell-order gather, before collide); requires HS_FLAG_TIMING */
int hs_dsmc_kernel_times(hs_dsmc_engine* eng, double* ms, int64_t* launches, int32_t n);

/* ---- calibration -------------------------------------------------------- */
/* No counterpart in the reference: times a plain device-to-device copy of
 * `bytes` (multiple of 32) on the current device, iters times, and returns
 * the mean milliseconds per copy.  Used to state the achievable HBM
 * bandwidth beside the nominal one and to scale profiler byte counters. */
int hs_dsmc_device_copy(int64_t bytes, int32_t iters, double* ms_out);

/* ---- binary collision test hook ----------------------------------------- */
/* BinaryCollisionModel::collide (VariableHardSphere.C:125-228,
 * VariableSoftSphere.C:189-247, LarsenBorgnakkeVariableHardSphere.C:110-279)
 * for n independent pairs with an injected uniform stream: pair k reads
 * uniforms[k*n_uniform .. k*n_uniform + n_uniform) in the reference's draw
 * order and computes in FP64.  Velocities are [3*k+c], FP64 in/out.
 * erot/vib may be NULL for monatomic species.  used_out[k] = draws consumed
 * (an error if a pair needs more than n_uniform). */
int hs_dsmc_binary_collide(hs_dsmc_engine* eng, int64_t n, const int32_t* type_p,
                           const int32_t* type_q, double* up, double* uq,
                           double* erot_p, double* erot_q, int32_t* vib_p,
                           int32_t* vib_q, const double* uniforms,
                           int32_t n_uniform, int32_t* used_out);
/* sigmaTcR (VariableHardSphere.C:80-121) for n pairs, FP64. */
int hs_dsmc_sigma_tcr(hs_dsmc_engine* eng, int64_t n, const int32_t* type_p,
                      const int32_t* type_q, const double* up, const double* uq,
                      double* out);

#ifdef __cplusplus
}
#endif
#endif /* HYSTRATH_DSMC_H */
