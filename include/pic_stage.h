/*
 * pic_stage.h -- C ABI of the MI355X PIC particle stage.
 *
 * The stage replaces, for the Boris-Leapfrog / HDG-Poisson time stepper
 * (reference: src/timedisc/timedisc_TimeStepPoissonByBorisLeapfrog.f90:32-286),
 * the four particle routines the stepper calls:
 *
 *   Deposition()                    src/particles/pic/deposition/pic_depo.f90:944-1018
 *   InterpolateFieldToParticle()    src/particles/pic/interpolation/pic_interpolation.f90:248-322
 *   push loop (inline in stepper)   timedisc_TimeStepPoissonByBorisLeapfrog.f90:128-198
 *   PerformTracking()               src/particles/tracking/particle_tracking.f90:32-70
 *                                   (TriaTracking, particle_triatracking.f90:63-484, and
 *                                    RefMapping, particle_reftracking.f90:30-414)
 *
 * In the reference these are Fortran module procedures that communicate
 * through module arrays (PartState, PEM, PDM, U_N, PS_N, FieldAtParticle).
 * Here the same arrays cross the boundary explicitly: plain pointers with the
 * reference's Fortran memory layout, so a Fortran host binds them with
 * ISO_C_BINDING and passes its own arrays unchanged:
 *
 *   mesh: ElemInfo(6,nElems), SideInfo(5,nSides), NodeCoords(3,nNodes),
 *         GlobalNodeIDs(nNodes), BCType(4,nBCs)   (HOPR file arrays)
 *   U_N(iElem)%E(1:3,0:N,0:N,0:N)        -> double E[nElems][(N+1)^3][3]
 *   PS_N(iElem)%PartSource(1:4,0:N,...)  -> double PS[nElems][(N+1)^3][4]
 *   PartState(1:6,nPart)                 -> double PartState[nPart][6]
 *   FieldAtParticle(1:6,nPart)           -> double F[nPart][6]
 *   PEM%GlobalElemID(nPart)              -> int32_t elem[nPart]  (1-based; 0 = unknown)
 *   PartSpecies(nPart)                   -> int32_t species[nPart] (1-based)
 *   PDM%IsNewPart / ParticleInside       -> uint8_t flags
 *
 * The library owns its device memory; host buffers stay the caller's and are
 * copied in and out.  All calls are synchronous on the handle's HIP stream
 * unless the name says _async.  Every entry point returns 0 on success and a
 * negative PIC_E* code on failure; pic_last_error() returns the message (the
 * reference aborts instead, src/globals/globals.f90:435).  A handle must not
 * be used from two threads at once (the reference procedures are not
 * reentrant either).
 */
#ifndef PIC_STAGE_H
#define PIC_STAGE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PIC_OK = 0,
  PIC_EINVAL = -1,     /* invalid argument or unsupported configuration   */
  PIC_EHIP = -2,       /* HIP runtime error                               */
  PIC_ECAPACITY = -3,  /* more particles than the handle was created for  */
  PIC_ESTATE = -4,     /* call out of order (e.g. no particles set)       */
  PIC_ETRACK = -5      /* tracking hit a state the reference aborts on    */
};

enum { PIC_TRACK_REFMAPPING = 1, PIC_TRACK_TRIA = 3 };         /* TrackingMethod, src/piclas.h:357-359 */
enum { PIC_DEPO_NONE = 0, PIC_DEPO_CVWM = 1, PIC_DEPO_SF = 2 }; /* PIC-Deposition-Type                  */
enum { PIC_BC_OPEN = 1, PIC_BC_REFLECTIVE = 2, PIC_BC_PERIODIC = 3 }; /* PartBound%TargetBoundCond      */
enum { PIC_PREC_FP64 = 0, PIC_PREC_MIXED = 1 };
/* dominant deposition kernel, reported in pic_stats.deposit_kernel */
enum {
  PIC_KERNEL_NONE = 0,
  PIC_KERNEL_CVWM = 1,          /* k_cvwm_elem                                   */
  PIC_KERNEL_SF_VALU = 2,       /* k_sf_gather: shape function on the vector ALUs */
  PIC_KERNEL_SF_MATRIX = 3,     /* k_sf_gather_mfma: sums on the matrix cores     */
  PIC_KERNEL_SF_CARTESIAN = 4,  /* k_sf_gather_cart: axis-aligned tensor grids    */
  PIC_KERNEL_SF_MIXED = 5       /* k_sf_gather_mixed(_cart): FP32 sums           */
};

/* Mesh in HOPR layout, as the reference's mesh reader holds it
 * (src/mesh/mesh_readin.f90; index constants src/piclas.h:136-176). */
typedef struct {
  int32_t nElems, nSides, nNodes;
  int32_t NGeo;                   /* geometry degree: 1 linear, 2 curved (RefMapping) */
  int32_t N;                      /* DG polynomial degree of E and PartSource     */
  const int32_t *ElemInfo;        /* [nElems][6]                                  */
  const int32_t *SideInfo;        /* [nSides][5]                                  */
  const double  *NodeCoords;      /* [nNodes][3]                                  */
  const int32_t *GlobalNodeIDs;   /* [nNodes] 1-based unique node ids             */
  int32_t nBCs;
  const int32_t *BCType;          /* [nBCs][4] (type, curve, state, alpha)        */
  const int32_t *PartBoundCond;   /* [nBCs] PIC_BC_*  (Part-BoundaryX-Condition)  */
  int32_t nPeriodicVectors;
  const double  *PeriodicVectors; /* [nPeriodicVectors][3]  (GEO%PeriodicVectors) */
} pic_mesh_desc;

/* Run-time parameters the reference reads from parameter.ini. */
typedef struct {
  int32_t tracking;               /* TrackingMethod: PIC_TRACK_TRIA or _REFMAPPING */
  int32_t deposition;             /* PIC-Deposition-Type                          */
  double  sf_radius;              /* PIC-shapefunction-radius                     */
  int32_t sf_alpha;               /* PIC-shapefunction-alpha                      */
  int32_t refmapping_guess;       /* RefMappingGuess                              */
  double  refmapping_eps;         /* RefMappingEps                                */
  double  external_field[6];      /* PIC-externalField * PIC-scaleexternalField   */
  int32_t nSpecies;               /* <= 8                                         */
  double  charge[8];              /* Part-SpeciesX-ChargeIC                       */
  double  mass[8];                /* Part-SpeciesX-MassIC                         */
  double  mpf[8];                 /* Part-SpeciesX-MacroParticleFactor            */
  int32_t precision;              /* PIC_PREC_*                                   */
  int64_t max_particles;          /* Part-maxParticleNumber                       */
  int32_t device;                 /* HIP device ordinal                           */
} pic_params;

typedef struct {
  int64_t n_particles;            /* particles currently inside the domain        */
  int64_t n_lost;                 /* particles removed by tracking (cumulative)    */
  int64_t n_steps;                /* time steps performed (cumulative)            */
  double  ms_deposit;             /* device time in deposition kernels            */
  double  ms_push_track;          /* device time in interpolate+push+track        */
  double  ms_sort;                /* device time in the element sort              */
  int32_t n_deposit_launches;     /* launches of the dominant deposition kernel   */
  double  ms_deposit_main;        /* device time of that kernel alone             */
  int32_t deposit_kernel;         /* that kernel: PIC_KERNEL_*                    */
  double  ms_push_main;           /* device time of the push kernel alone (eager
                                     steps; interpolation, push and the test in
                                     the last element)                           */
} pic_stats;

typedef struct pic_handle pic_handle;

/* Set-up / tear-down (InitParticleMesh, InitializeDeposition, InitPIC). */
int  pic_create(const pic_mesh_desc *mesh, const pic_params *par, pic_handle **out);
int  pic_destroy(pic_handle *h);
const char *pic_last_error(void);

/* Field: U_N%E (host array).  pic_set_field_device takes a device pointer of
 * the same layout and copies device-to-device on the handle's stream. */
int  pic_set_field(pic_handle *h, const double *E);
int  pic_set_field_device(pic_handle *h, const double *E_dev);

/* Particles (PartState, PEM%GlobalElemID, PartSpecies, PDM%IsNewPart).
 * elem may be NULL or hold zeros: the stage then locates the particles
 * (SinglePointToElement, src/particles/tracking/particle_localization.f90:81-203). */
int  pic_set_particles(pic_handle *h, int64_t n, const double *PartState, const int32_t *elem,
                       const int32_t *species, const uint8_t *is_new);

/* Particle emission on the device (initial insertion,
 * particles/emission/particle_position_and_velocity.f90: SetParticlePosition
 * with SpaceIC 'cuboid' or 'sin_deviation', SetParticleVelocity with
 * 'constant' or 'maxwell_lpn'): positions and velocities are formed on the
 * device, located there (SinglePointToElement) and set as the stage's
 * particles, as pic_set_particles with IsNewPart = T.  Particle ids follow
 * the inits in order (init 0's particles first); particles no element holds
 * are kept as not inside.  Random numbers: Philox4x32-10 keyed by seed with
 * the counter (particle id, draw), so results depend on (seed, id) only; the
 * reference's RANDOM_NUMBER stream is not reproduced. */
enum { PIC_SPACE_CUBOID = 1, PIC_SPACE_SIN_DEVIATION = 2 };   /* Part-SpeciesX-InitY-SpaceIC        */
enum { PIC_VELO_CONSTANT = 1, PIC_VELO_MAXWELL_LPN = 2 };     /* Part-SpeciesX-InitY-velocityDistribution */
typedef struct {
  int32_t species;                /* 1-based                                       */
  int32_t space;                  /* PIC_SPACE_*                                   */
  int64_t particle_number;        /* ParticleNumber (cuboid)                       */
  double  base_point[3];          /* BasePointIC (cuboid)                          */
  double  base_vector1[3];        /* BaseVector1IC (cuboid)                        */
  double  base_vector2[3];        /* BaseVector2IC (cuboid)                        */
  double  cuboid_height;          /* CuboidHeightIC (cuboid)                       */
  int32_t max_particle_number[3]; /* maxParticleNumber-x/y/z (sin_deviation)       */
  double  amplitude;              /* Amplitude (sin_deviation)                     */
  double  wave_number;            /* WaveNumber (sin_deviation)                    */
  int32_t velocity;               /* PIC_VELO_*                                    */
  double  velo_ic;                /* VeloIC                                        */
  double  velo_vec[3];            /* VeloVecIC                                     */
  double  temperature;            /* MWTemperatureIC                               */
} pic_emission;
int  pic_emit_particles(pic_handle *h, int32_t n_inits, const pic_emission *inits, uint64_t seed, int64_t *n_emitted);
/* pic_emit_particles without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (serving loops, harnesses that forbid hipMalloc).  The device staging
 * lives in a fixed array the library holds on every GPU it runs on (set up by
 * pic_create): at most 524,288 particles per call (40.4 MB per GPU); larger
 * emissions fail with PIC_ECAPACITY.  Results are those of pic_emit_particles.
 * The staging is shared by every handle on one GPU: calls that overlap on one
 * GPU (from two threads) race. */
int  pic_emit_particles_noalloc(pic_handle *h, int32_t n_inits, const pic_emission *inits, uint64_t seed,
                                int64_t *n_emitted);

/* Particle state in the reference's state-file layout (element-sorted):
 *   PartInt  [nElems][2]  first and last offset of each element's particles
 *                         (the element holds rows first .. last-1)
 *   PartData [n][PIC_PART_DATA_SIZE]  x y z vx vy vz species (as a real)
 * pic_get_particle_data replaces FillParticleData
 * (src/io_hdf5/hdf5_output_particle.f90:1514-1620): particles inside, elements
 * ascending, caller order within an element.  *n_out receives their number
 * (also when cap is too small: the call then fails with PIC_ECAPACITY).
 * pic_restart_particles replaces ParticleRestart's particle part
 * (src/particles/restart/particle_restart.f90:134-440, after ParticleReadin,
 * particle_readin.f90:503-637): the stored element is checked
 * (ParticleInsideQuad; RefMapping: |xi| <= ElemEpsOneCell), particles that
 * fail it are located again (SinglePointToElement) and counted in
 * *n_missing; those no element holds are set as not inside and counted in
 * *n_lost.  A species ID <= 0 is rejected (the reference aborts).  The HDF5
 * file is read and written by the host (piclas_amd/state_file.py). */
#define PIC_PART_DATA_SIZE 7
int  pic_get_particle_data(pic_handle *h, int64_t cap, double *PartData, int64_t *PartInt, int64_t *n_out);
int  pic_restart_particles(pic_handle *h, int64_t n, const double *PartData, const int64_t *PartInt,
                           int64_t *n_missing, int64_t *n_lost);

/* The four stage routines, by reference name. */
int  pic_deposition(pic_handle *h);                          /* Deposition()                 */
int  pic_interpolate_field_to_particle(pic_handle *h);       /* InterpolateFieldToParticle() */
int  pic_push_boris_leapfrog(pic_handle *h, double dt);      /* stepper push loop            */
int  pic_perform_tracking(pic_handle *h);                    /* PerformTracking()            */

/* nsteps of TimeStepPoissonByBorisLeapfrog with the field held fixed:
 * Deposition; LastPartPos <- PartState; interpolate; push; track. */
int  pic_time_step(pic_handle *h, double dt, int32_t nsteps);

/* Results, in the caller's original particle order. */
int  pic_get_particles(pic_handle *h, double *PartState, int32_t *elem, uint8_t *inside);
int  pic_get_last_part_pos(pic_handle *h, double *LastPartPos);
int  pic_get_field_at_particle(pic_handle *h, double *FieldAtParticle);
/* PartPosRef(1:3,:) of RefMapping (particle_vars.f90:52): reference coordinates
 * of every particle in its element, caller order. */
int  pic_get_part_pos_ref(pic_handle *h, double *PartPosRef);
int  pic_get_part_source(pic_handle *h, double *PS);
/* Device pointer to PS[nElems][(N+1)^3][4], valid until the next deposition. */
int  pic_part_source_device(pic_handle *h, const double **PS_dev);
/* CalcDepositedCharge (src/particles/pic/analyze/pic_analyze.f90:136-210) on the
 * current source: PartCharge[0] = sum_e sum_ijk w_i w_j w_k rho / sJ,
 * PartCharge[1] = |Q_particles - PartCharge[0]|, PartCharge[2] = PartCharge[1] / Q_particles
 * (0 when Q_particles is ALMOSTZERO), Q_particles = sum over particles inside of ChargeIC * MPF. */
int  pic_deposited_charge(pic_handle *h, double *PartCharge);
/* CalcPotentialEnergy (src/analyze/analyzefield_tools.f90:99-228, HDG with PP_nVar=1):
 * WEl = eps0/2 sum_e sum_ijk w_i w_j w_k / sJ |E|^2 of the field last set. */
int  pic_field_energy(pic_handle *h, double *WEl);
/* Self-consistent steps through a fixed linear field operator, and their
 * adjoint (no counterpart in the reference, where HDG() supplies E; these
 * differentiate the particle stage).  G[3*nUniqueNodes][nElems] (row
 * 3*(GlobalNodeID-1)+c) maps the element means of rho,
 * rbar_e = sum_ijk w_i w_j w_k rho_e,ijk / 8, to E at the unique corner nodes;
 * E at the Gauss points is their trilinear interpolation.  Shape-function
 * deposition, TriaTracking, FP64, periodic meshes.
 *   pic_sc_time_step: nsteps of deposit, field, interpolate, push, track; then,
 *                     if loss is given, the field energy (CalcPotentialEnergy)
 *                     of the field of the final deposition.
 *   pic_adjoint:      K such steps and the loss, and the gradients of the loss
 *                     with respect to the positions and velocities at entry
 *                     (caller order, [nPart][3]); the states are recomputed
 *                     from checkpoints every checkpoint_every steps.  Leaves the
 *                     stage in the state after K steps. */
int  pic_set_field_operator(pic_handle *h, const double *G);
int  pic_sc_time_step(pic_handle *h, double dt, int32_t nsteps, double *loss);
int  pic_adjoint(pic_handle *h, double dt, int32_t K, int32_t checkpoint_every, double *loss, double *dLdx0,
                 double *dLdv0);

/* Execution tuning (no counterpart in the reference; results do not depend on
 * it beyond round-off in the order of the shape-function sums):
 *   use_graph  pic_time_step replays a captured two-step HIP graph (no host
 *              synchronisation between steps; per-stage timers then report the
 *              replays as push/track time)
 *   sf_kernel  0 default, 1 shape-function sums on the vector ALUs, 2 on the
 *              matrix cores (N = 3)
 *   sf_persistent  1: the matrix-core shape-function sums run as persistent
 *              waves, as many as the device holds at once, that take target
 *              elements from atomic counters (one per XCD range of
 *              elements) instead of one wave per target element; for meshes
 *              whose elements hold very different particle counts.  Each
 *              target is summed by one wave in the same order either way, so
 *              the result is bit-identical.
 *   sf_shared_loads  1: on meshes that are not axis-aligned tensor grids, the
 *              four waves of a workgroup take four consecutive target
 *              elements and stage each record of the union of their source
 *              lists once in LDS for the four (more culls per wave, and no
 *              skipping of the record blocks whose bounding box cannot reach
 *              a target, which the default gather does); bit-identical as
 *              well.  Takes precedence over sf_persistent there. */
typedef struct {
  int32_t use_graph;
  int32_t sf_kernel;
  int32_t sf_persistent;
  int32_t sf_shared_loads;
} pic_tuning;
int  pic_set_tuning(pic_handle *h, const pic_tuning *t);
int  pic_get_stats(pic_handle *h, pic_stats *st);
int  pic_reset_stats(pic_handle *h);
int  pic_synchronize(pic_handle *h);

#ifdef __cplusplus
}
#endif
#endif
