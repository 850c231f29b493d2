/*
 * pic_oracle.h -- CPU oracle for the PIC particle stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, operation-for-operation
 * restatement of the reference's particle stage (field interpolation, the
 * Boris-Leapfrog push, TriaTracking and RefMapping tracking on linear and
 * curved hexahedra, cell_volweight_mean and shape_function
 * deposition, and the mesh tables those routines read).  The tests compare
 * the HIP implementation in piclas_amd/ against it, and it is the CPU
 * baseline timed beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Conventions follow the reference's Fortran arrays:
 *   ElemInfo[nElems][6], SideInfo[nSides][5]   HOPR mesh arrays
 *                                               (src/piclas.h:136-176)
 *   NodeCoords[nNodes][3], GlobalNodeIDs[nNodes]
 *   U_N%E(1:3,0:N,0:N,0:N)     -> E[nElems][(N+1)^3][3]   (i fastest)
 *   PS_N%PartSource(1:4,...)   -> PS[nElems][(N+1)^3][4]
 *   PartState(1:6,nPart)       -> PartState[nPart][6]
 *   element ids at the API are 1-based global ids (0 = unknown / lost).
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t nElems, nSides, nNodes;
  int32_t NGeo;                  /* geometry degree of NodeCoords (1 = linear hex) */
  int32_t N;                     /* DG degree of E and PartSource                  */
  const int32_t *ElemInfo;       /* [nElems][6]                                    */
  const int32_t *SideInfo;       /* [nSides][5]                                    */
  const double  *NodeCoords;     /* [nNodes][3]                                    */
  const int32_t *GlobalNodeIDs;  /* [nNodes], 1-based unique node ids              */
  int32_t nBCs;
  const int32_t *BCType;         /* [nBCs][4] HOPR BCType rows (type,curve,state,alpha) */
  const int32_t *PartBoundCond;  /* [nBCs] 1 open, 2 reflective, 3 periodic        */
  int32_t nPeriodicVectors;
  const double  *PeriodicVectors;/* [nPeriodicVectors][3]                          */
} om_mesh_desc;

enum { OM_TRACK_REFMAPPING = 1, OM_TRACK_TRIA = 3 };
enum { OM_DEPO_NONE = 0, OM_DEPO_CVWM = 1, OM_DEPO_SF = 2 };

typedef struct {
  int32_t tracking;              /* OM_TRACK_*                                     */
  int32_t deposition;            /* OM_DEPO_*                                      */
  double  sf_radius;             /* PIC-shapefunction-radius                       */
  int32_t sf_alpha;              /* PIC-shapefunction-alpha                        */
  int32_t refmapping_guess;      /* RefMappingGuess (1..4)                         */
  double  refmapping_eps;        /* RefMappingEps                                  */
  double  external_field[6];     /* PIC-externalField * PIC-scaleexternalField     */
  int32_t nSpecies;
  double  charge[8];             /* Species%ChargeIC                               */
  double  mass[8];               /* Species%MassIC                                 */
  double  mpf[8];                /* Species%MacroParticleFactor                    */
} om_params;

typedef struct om_state om_state;

om_state *om_create(const om_mesh_desc *mesh, const om_params *par);
void      om_destroy(om_state *s);
const char *om_last_error(void);
void      om_set_threads(om_state *s, int nthreads);

int  om_set_field(om_state *s, const double *E);
int  om_set_particles(om_state *s, int32_t n, const double *PartState,
                      const int32_t *elem, const int32_t *species, const uint8_t *is_new);
/* SinglePointToElement for every particle whose element is 0 */
int  om_locate_particles(om_state *s);

/* The stage, by reference name. */
int  om_deposition(om_state *s);                     /* Deposition()                  */
int  om_interpolate_field_to_particle(om_state *s);  /* InterpolateFieldToParticle()  */
int  om_push_boris_leapfrog(om_state *s, double dt); /* push loop of the 508 stepper  */
int  om_perform_tracking(om_state *s);               /* PerformTracking()             */
/* nsteps of: Deposition, LastPartPos <- PartState, interpolate, push, track */
int  om_time_step(om_state *s, double dt, int32_t nsteps);

int32_t om_get_nparticles(const om_state *s);
int  om_get_particles(const om_state *s, double *PartState, int32_t *elem, uint8_t *inside);
int  om_get_last_part_pos(const om_state *s, double *LastPartPos);
int  om_get_field_at_particle(const om_state *s, double *FieldAtParticle);
int  om_get_part_source(const om_state *s, double *PS);
int64_t om_get_lost_count(const om_state *s);
/* Threaded shape-function deposit only: wall seconds spent zeroing and
 * summing the per-thread PartSource copies, over all calls. */
double om_get_private_reduce_seconds(const om_state *s);
/* How often the TriaTracking walk took the multi-crossing branch
 * (particle_triatracking.f90:309-405), summed over all calls. */
int64_t om_get_multi_crossings(const om_state *s);
/* Test only: plant a known mistake so that a test can show its checks catch
 * it.  1: the multi-crossing tie-break takes the side crossed last instead of
 * first.  0 restores the restatement. */
void om_plant_mistake(om_state *s, int which);
/* Test only: record up to cap multi-crossing events of the TriaTracking walk
 * (0 switches the log off).  One record is OM_TB_REC doubles: particle index,
 * 1-based element, number of candidate sides, the chosen local side and
 * triangle (0, 0 when the particle was lost), LastPartPos(3), PartState(3),
 * then per candidate: local side, triangle, flags (bit 0: not excluded by
 * DoneLastElem, bit 1: the last position lies on the element's side of the
 * triangle, i.e. ParticleThroughSideLastPosCheck held). */
#define OM_TB_REC 29
int  om_set_tiebreak_log(om_state *s, int64_t cap);
int64_t om_get_tiebreak_log(const om_state *s, double *out);  /* returns the number of records */

/* Derived tables, for checking the product's own mesh producers. */
int  om_get_elem_xgp(const om_state *s, double *xgp);        /* [nElems][(N+1)^3][3] */
int  om_get_sj(const om_state *s, double *sJ);               /* [nElems][(N+1)^3]    */
int  om_get_elem_side_node_id(const om_state *s, int32_t *id); /* [nElems][6][4] 0-based node */
int  om_get_concave(const om_state *s, uint8_t *c);          /* [nElems][6]          */
int  om_get_node_volume(const om_state *s, double *nv);      /* [nUniqueNodes]       */
int32_t om_get_n_unique_nodes(const om_state *s);
int  om_get_elem_bary(const om_state *s, double *bary, double *radius);
int  om_get_xieta_basis(const om_state *s, double *basis, double *slen); /* [nElems][6][3], [nElems][6] */
int  om_get_part_pos_ref(const om_state *s, double *PartPosRef);         /* [nPart][3] (RefMapping) */
int  om_get_refmapping_tables(const om_state *s, uint8_t *ElemCurved, double *ElemEpsOneCell); /* [nElems] */

/* Reference-coordinate mapping of single points (GetPositionInRefElem). */
int  om_position_in_ref_elem(om_state *s, int32_t n, const double *x, const int32_t *elem,
                             double *xi, uint8_t *success);
/* Field evaluation at reference positions (EvaluateFieldAtRefPos) */
int  om_eval_field_at_ref_pos(om_state *s, int32_t n, const double *xi, const int32_t *elem,
                              double *U);
/* GetEMFieldDW (inverse-distance weights to the Gauss points) at physical
 * positions, whatever RefElemNewton would say: for pinning the loop exit. */
int  om_em_field_dw(om_state *s, int32_t n, const double *x, const int32_t *elem, double *U);
/* RefMapping boundary tracking tables (CartesianPeriodic = F meshes):
 * om_get_n_bc_sides: boundary sides (0 when every boundary is periodic);
 * om_get_bc_sides: per boundary side its 0-based HOPR side row, SideType,
 *   Bezier control points [nBC][NGeo+1][NGeo+1][3] (p, q, xyz), origin and radius;
 * om_get_bc_elem_lists: CSR of every element's sides in range (offsets
 *   [nElems+1], boundary side index, distance) in the order of the lists;
 * om_bc_side_intersection: intersection of the segment last -> x with one
 *   boundary side as ParticleBCTracking computes it (alpha = -1: no hit). */
int32_t om_get_n_bc_sides(const om_state *s);
int  om_get_bc_sides(const om_state *s, int32_t *side, uint8_t *type, double *bez, double *origin, double *radius);
int32_t om_get_bc_list_size(const om_state *s);
int  om_get_bc_elem_lists(const om_state *s, int32_t *off, int32_t *bcside, double *dist);
int  om_bc_side_intersection(om_state *s, int32_t b, const double *last, const double *x, double *alpha_xi_eta);

/* Analysis: integral of rho (CalcDepositedCharge) and electric field energy (CalcPotentialEnergy) */
double om_deposited_charge(const om_state *s);
double om_particle_charge(const om_state *s);
double om_field_energy(const om_state *s);

/* Self-consistent loop through a fixed linear field operator (adjoint workload):
 * G[3*nUniqueNodes][nElems] maps element-mean rho to E at the unique corner
 * nodes; E at the Gauss points is their trilinear interpolation. */
int  om_sc_set_operator(om_state *s, const double *G);
int  om_sc_field_from_source(om_state *s);
/* Test only: overwrite PartSource (as a deposition would leave it) and read
 * back the field at the Gauss points (as om_sc_field_from_source leaves it). */
int  om_set_part_source(om_state *s, const double *PS);
int  om_get_field(const om_state *s, double *E);
int  om_sc_forward(om_state *s, double dt, int32_t K, double *loss);

/* 1D node / basis helpers (basis.f90), exposed for closed-form checks */
void om_legendre_gauss_nodes(int32_t N, double *x, double *w);
void om_cheby_gauss_lobatto_nodes(int32_t N, double *x);
void om_barycentric_weights(int32_t N, const double *x, double *wb);
void om_lagrange_polys(double x, int32_t N, const double *xGP, const double *wb, double *L);

#ifdef __cplusplus
}
#endif
#endif
