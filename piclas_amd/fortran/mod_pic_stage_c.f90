!==================================================================================================================================
! MOD_PIC_Stage_C: Fortran host binding of the MI355X particle stage (libpicstage.so, include/pic_stage.h).
!
! The reference's Boris-Leapfrog stepper (src/timedisc/timedisc_TimeStepPoissonByBorisLeapfrog.f90:32-286) calls
!   Deposition()                 src/particles/pic/deposition/pic_depo.f90:944
!   InterpolateFieldToParticle() src/particles/pic/interpolation/pic_interpolation.f90:248
!   (inline Boris push)          timedisc_TimeStepPoissonByBorisLeapfrog.f90:128-198
!   PerformTracking()            src/particles/tracking/particle_tracking.f90:32
! on module arrays.  This module keeps those names and the reference's array shapes
! (PartState(1:6,:), U_N%E(1:3,0:N,0:N,0:N), PS_N%PartSource(1:4,0:N,0:N,0:N)) and forwards every call through
! ISO_C_BINDING to the C ABI, in the same way the reference binds its own C helpers (src/globals/globals.f90:116-121).
! The particle arrays live on the GPU between calls; PICStage_SetParticles / PICStage_GetParticles move them.
!==================================================================================================================================
MODULE MOD_PIC_Stage_C
USE ISO_C_BINDING
IMPLICIT NONE
PRIVATE

INTEGER(C_INT),PARAMETER,PUBLIC :: PIC_TRACK_REFMAPPING=1, PIC_TRACK_TRIA=3, PIC_DEPO_NONE=0, PIC_DEPO_CVWM=1, PIC_DEPO_SF=2
INTEGER(C_INT),PARAMETER,PUBLIC :: PIC_BC_OPEN=1, PIC_BC_REFLECTIVE=2, PIC_BC_PERIODIC=3

TYPE, BIND(C), PUBLIC :: tPicMeshDesc
  INTEGER(C_INT32_T) :: nElems, nSides, nNodes, NGeo, N
  TYPE(C_PTR)        :: ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs
  INTEGER(C_INT32_T) :: nBCs
  TYPE(C_PTR)        :: BCType, PartBoundCond
  INTEGER(C_INT32_T) :: nPeriodicVectors
  TYPE(C_PTR)        :: PeriodicVectors
END TYPE tPicMeshDesc

TYPE, BIND(C), PUBLIC :: tPicParams
  INTEGER(C_INT32_T) :: tracking, deposition
  REAL(C_DOUBLE)     :: sf_radius
  INTEGER(C_INT32_T) :: sf_alpha, refmapping_guess
  REAL(C_DOUBLE)     :: refmapping_eps
  REAL(C_DOUBLE)     :: external_field(6)
  INTEGER(C_INT32_T) :: nSpecies
  REAL(C_DOUBLE)     :: charge(8), mass(8), mpf(8)
  INTEGER(C_INT32_T) :: precision
  INTEGER(C_INT64_T) :: max_particles
  INTEGER(C_INT32_T) :: device
END TYPE tPicParams

!> one emission init (Part-SpeciesX-InitY-*), see pic_emit_particles
TYPE, BIND(C), PUBLIC :: tPicEmission
  INTEGER(C_INT32_T) :: species, space
  INTEGER(C_INT64_T) :: particle_number
  REAL(C_DOUBLE)     :: base_point(3), base_vector1(3), base_vector2(3), cuboid_height
  INTEGER(C_INT32_T) :: max_particle_number(3)
  REAL(C_DOUBLE)     :: amplitude, wave_number
  INTEGER(C_INT32_T) :: velocity
  REAL(C_DOUBLE)     :: velo_ic, velo_vec(3), temperature
END TYPE tPicEmission

INTEGER(C_INT32_T),PARAMETER,PUBLIC :: PIC_SPACE_CUBOID = 1, PIC_SPACE_SIN_DEVIATION = 2
INTEGER(C_INT32_T),PARAMETER,PUBLIC :: PIC_VELO_CONSTANT = 1, PIC_VELO_MAXWELL_LPN = 2

INTERFACE
  INTEGER(C_INT) FUNCTION pic_create(mesh, par, h) BIND(C, NAME='pic_create')
    IMPORT :: C_INT, C_PTR, tPicMeshDesc, tPicParams
    TYPE(tPicMeshDesc),INTENT(IN) :: mesh
    TYPE(tPicParams),INTENT(IN)   :: par
    TYPE(C_PTR),INTENT(OUT)       :: h
  END FUNCTION pic_create
  INTEGER(C_INT) FUNCTION pic_destroy(h) BIND(C, NAME='pic_destroy')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: h
  END FUNCTION pic_destroy
  TYPE(C_PTR) FUNCTION pic_last_error() BIND(C, NAME='pic_last_error')
    IMPORT :: C_PTR
  END FUNCTION pic_last_error
  INTEGER(C_INT) FUNCTION pic_set_field(h, E) BIND(C, NAME='pic_set_field')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: h
    REAL(C_DOUBLE),INTENT(IN)  :: E(*)
  END FUNCTION pic_set_field
  INTEGER(C_INT) FUNCTION pic_set_particles(h, n, PartState, elem, species, is_new) BIND(C, NAME='pic_set_particles')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T, C_INT32_T, C_INT8_T
    TYPE(C_PTR),VALUE              :: h
    INTEGER(C_INT64_T),VALUE       :: n
    REAL(C_DOUBLE),INTENT(IN)      :: PartState(*)
    INTEGER(C_INT32_T),INTENT(IN)  :: elem(*), species(*)
    INTEGER(C_INT8_T),INTENT(IN)   :: is_new(*)
  END FUNCTION pic_set_particles
  INTEGER(C_INT) FUNCTION pic_h(h) BIND(C, NAME='pic_deposition')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: h
  END FUNCTION pic_h
  INTEGER(C_INT) FUNCTION pic_interp(h) BIND(C, NAME='pic_interpolate_field_to_particle')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: h
  END FUNCTION pic_interp
  INTEGER(C_INT) FUNCTION pic_push(h, dt) BIND(C, NAME='pic_push_boris_leapfrog')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE    :: h
    REAL(C_DOUBLE),VALUE :: dt
  END FUNCTION pic_push
  INTEGER(C_INT) FUNCTION pic_track(h) BIND(C, NAME='pic_perform_tracking')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: h
  END FUNCTION pic_track
  INTEGER(C_INT) FUNCTION pic_time_step(h, dt, nsteps) BIND(C, NAME='pic_time_step')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
    TYPE(C_PTR),VALUE        :: h
    REAL(C_DOUBLE),VALUE     :: dt
    INTEGER(C_INT32_T),VALUE :: nsteps
  END FUNCTION pic_time_step
  INTEGER(C_INT) FUNCTION pic_get_particles(h, PartState, elem, inside) BIND(C, NAME='pic_get_particles')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T, C_INT8_T
    TYPE(C_PTR),VALUE               :: h
    REAL(C_DOUBLE),INTENT(OUT)      :: PartState(*)
    INTEGER(C_INT32_T),INTENT(OUT)  :: elem(*)
    INTEGER(C_INT8_T),INTENT(OUT)   :: inside(*)
  END FUNCTION pic_get_particles
  INTEGER(C_INT) FUNCTION pic_get_part_source(h, PS) BIND(C, NAME='pic_get_part_source')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: h
    REAL(C_DOUBLE),INTENT(OUT) :: PS(*)
  END FUNCTION pic_get_part_source
  INTEGER(C_INT) FUNCTION pic_get_field_at_particle(h, F) BIND(C, NAME='pic_get_field_at_particle')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: h
    REAL(C_DOUBLE),INTENT(OUT) :: F(*)
  END FUNCTION pic_get_field_at_particle
  INTEGER(C_INT) FUNCTION pic_get_part_pos_ref(h, PartPosRef) BIND(C, NAME='pic_get_part_pos_ref')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: h
    REAL(C_DOUBLE),INTENT(OUT) :: PartPosRef(*)
  END FUNCTION pic_get_part_pos_ref
  INTEGER(C_INT) FUNCTION pic_deposited_charge(h, PartCharge) BIND(C, NAME='pic_deposited_charge')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: h
    REAL(C_DOUBLE),INTENT(OUT) :: PartCharge(3)
  END FUNCTION pic_deposited_charge
  INTEGER(C_INT) FUNCTION pic_field_energy(h, WEl) BIND(C, NAME='pic_field_energy')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: h
    REAL(C_DOUBLE),INTENT(OUT) :: WEl
  END FUNCTION pic_field_energy
  INTEGER(C_INT) FUNCTION pic_emit_particles(h, n_inits, inits, seed, n_emitted) BIND(C, NAME='pic_emit_particles')
    IMPORT :: C_INT, C_PTR, C_INT32_T, C_INT64_T, tPicEmission
    TYPE(C_PTR),VALUE                 :: h
    INTEGER(C_INT32_T),VALUE          :: n_inits
    TYPE(tPicEmission),INTENT(IN)     :: inits(*)
    INTEGER(C_INT64_T),VALUE          :: seed
    INTEGER(C_INT64_T),INTENT(OUT)    :: n_emitted
  END FUNCTION pic_emit_particles
  INTEGER(C_INT) FUNCTION pic_get_particle_data(h, cap, PartData, PartInt, n_out) BIND(C, NAME='pic_get_particle_data')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T
    TYPE(C_PTR),VALUE               :: h
    INTEGER(C_INT64_T),VALUE        :: cap
    REAL(C_DOUBLE),INTENT(OUT)      :: PartData(*)
    INTEGER(C_INT64_T),INTENT(OUT)  :: PartInt(*)
    INTEGER(C_INT64_T),INTENT(OUT)  :: n_out
  END FUNCTION pic_get_particle_data
  INTEGER(C_INT) FUNCTION pic_restart_particles(h, n, PartData, PartInt, n_missing, n_lost) &
      BIND(C, NAME='pic_restart_particles')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T
    TYPE(C_PTR),VALUE               :: h
    INTEGER(C_INT64_T),VALUE        :: n
    REAL(C_DOUBLE),INTENT(IN)       :: PartData(*)
    INTEGER(C_INT64_T),INTENT(IN)   :: PartInt(*)
    INTEGER(C_INT64_T),INTENT(OUT)  :: n_missing, n_lost
  END FUNCTION pic_restart_particles
  INTEGER(C_SIZE_T) FUNCTION c_strlen(s) BIND(C, NAME='strlen')
    IMPORT :: C_SIZE_T, C_PTR
    TYPE(C_PTR),VALUE :: s
  END FUNCTION c_strlen
END INTERFACE

TYPE(C_PTR),SAVE :: stage = C_NULL_PTR

PUBLIC :: InitPICStage, FinalizePICStage, PICStage_SetField, PICStage_SetParticles, PICStage_GetParticles
PUBLIC :: Deposition, InterpolateFieldToParticle, PushBorisLeapfrog, PerformTracking, TimeStepPoissonByBorisLeapfrog
PUBLIC :: PICStage_GetPartSource, PICStage_GetFieldAtParticle, PICStage_GetPartPosRef
PUBLIC :: CalcDepositedCharge, CalcPotentialEnergy, PICStage_EmitParticles
PUBLIC :: PICStage_FillParticleData, PICStage_ParticleRestart

CONTAINS

SUBROUTINE Check(rc, what)
! Reference semantics: a fatal problem aborts (src/globals/globals.f90:435)
INTEGER(C_INT),INTENT(IN)   :: rc
CHARACTER(LEN=*),INTENT(IN) :: what
CHARACTER(KIND=C_CHAR),POINTER :: msg(:)
INTEGER :: n
IF (rc.EQ.0) RETURN
n = INT(c_strlen(pic_last_error()))
CALL C_F_POINTER(pic_last_error(), msg, [n])
WRITE(*,'(A,A,I0,A,*(A))') what, ' failed (', rc, '): ', msg(1:n)
ERROR STOP 'MOD_PIC_Stage_C'
END SUBROUTINE Check

!> InitParticleMesh / InitializeDeposition / InitPIC equivalent: hand the HOPR mesh arrays and the parameters over.
SUBROUTINE InitPICStage(ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs, BCType, PartBoundCond, PeriodicVectors, NGeo, N, par)
INTEGER(C_INT32_T),TARGET,INTENT(IN) :: ElemInfo(:,:), SideInfo(:,:), GlobalNodeIDs(:), BCType(:,:), PartBoundCond(:)
REAL(C_DOUBLE),TARGET,INTENT(IN)     :: NodeCoords(:,:), PeriodicVectors(:,:)
INTEGER,INTENT(IN)                   :: NGeo, N
TYPE(tPicParams),INTENT(IN)          :: par
TYPE(tPicMeshDesc) :: m
m%nElems = SIZE(ElemInfo,2); m%nSides = SIZE(SideInfo,2); m%nNodes = SIZE(NodeCoords,2)
m%NGeo = NGeo; m%N = N
m%ElemInfo = C_LOC(ElemInfo); m%SideInfo = C_LOC(SideInfo)
m%NodeCoords = C_LOC(NodeCoords); m%GlobalNodeIDs = C_LOC(GlobalNodeIDs)
m%nBCs = SIZE(BCType,2); m%BCType = C_LOC(BCType); m%PartBoundCond = C_LOC(PartBoundCond)
m%nPeriodicVectors = SIZE(PeriodicVectors,2)
IF (m%nPeriodicVectors.GT.0) THEN
  m%PeriodicVectors = C_LOC(PeriodicVectors)
ELSE
  m%PeriodicVectors = C_NULL_PTR
END IF
CALL Check(pic_create(m, par, stage), 'pic_create')
END SUBROUTINE InitPICStage

SUBROUTINE FinalizePICStage()
IF (C_ASSOCIATED(stage)) CALL Check(pic_destroy(stage), 'pic_destroy')
stage = C_NULL_PTR
END SUBROUTINE FinalizePICStage

!> U_N(iElem)%E(1:3,0:N,0:N,0:N) for all elements, stored contiguously as E(1:3,0:N,0:N,0:N,1:nElems)
SUBROUTINE PICStage_SetField(E)
REAL(C_DOUBLE),INTENT(IN) :: E(:,:,:,:,:)
CALL Check(pic_set_field(stage, E), 'pic_set_field')
END SUBROUTINE PICStage_SetField

!> initial particles emitted on the device (SetParticlePosition / SetParticleVelocity
!> for every init), located there; nEmitted = particles inside the mesh
SUBROUTINE PICStage_EmitParticles(inits, seed, nEmitted)
TYPE(tPicEmission),INTENT(IN)   :: inits(:)
INTEGER(C_INT64_T),INTENT(IN)   :: seed
INTEGER(C_INT64_T),INTENT(OUT)  :: nEmitted
CALL Check(pic_emit_particles(stage, INT(SIZE(inits),C_INT32_T), inits, seed, nEmitted), 'pic_emit_particles')
END SUBROUTINE PICStage_EmitParticles

!> FillParticleData: the particles inside, element-sorted as the state file holds them;
!> PartData(1:7,1:locnPart) = position, velocity, species; PartInt(1:2,1:nElems) = first and last offset
SUBROUTINE PICStage_FillParticleData(PartData, PartInt, locnPart)
REAL(C_DOUBLE),INTENT(OUT)      :: PartData(:,:)
INTEGER(C_INT64_T),INTENT(OUT)  :: PartInt(:,:)
INTEGER(C_INT64_T),INTENT(OUT)  :: locnPart
CALL Check(pic_get_particle_data(stage, INT(SIZE(PartData,2),C_INT64_T), PartData, PartInt, locnPart), &
           'pic_get_particle_data')
END SUBROUTINE PICStage_FillParticleData

!> ParticleRestart from the state file's PartData(1:7,:) and PartInt(1:2,1:nElems): stored elements
!> checked on the device, particles that fail the check located again
SUBROUTINE PICStage_ParticleRestart(PartData, PartInt, NbrOfMissingParticles, NbrOfLostParticles)
REAL(C_DOUBLE),INTENT(IN)       :: PartData(:,:)
INTEGER(C_INT64_T),INTENT(IN)   :: PartInt(:,:)
INTEGER(C_INT64_T),INTENT(OUT)  :: NbrOfMissingParticles, NbrOfLostParticles
CALL Check(pic_restart_particles(stage, INT(SIZE(PartData,2),C_INT64_T), PartData, PartInt, &
                                 NbrOfMissingParticles, NbrOfLostParticles), 'pic_restart_particles')
END SUBROUTINE PICStage_ParticleRestart

!> PartState(1:6,1:n), PEM%GlobalElemID (0: locate), PartSpecies, PDM%IsNewPart
SUBROUTINE PICStage_SetParticles(PartState, GlobalElemID, PartSpecies, IsNewPart)
REAL(C_DOUBLE),INTENT(IN)      :: PartState(:,:)
INTEGER(C_INT32_T),INTENT(IN)  :: GlobalElemID(:), PartSpecies(:)
LOGICAL,INTENT(IN)             :: IsNewPart(:)
INTEGER(C_INT8_T) :: isnew(SIZE(IsNewPart))
isnew = MERGE(1_C_INT8_T, 0_C_INT8_T, IsNewPart)
CALL Check(pic_set_particles(stage, INT(SIZE(PartState,2),C_INT64_T), PartState, GlobalElemID, PartSpecies, isnew), &
           'pic_set_particles')
END SUBROUTINE PICStage_SetParticles

!> PartState, PEM%GlobalElemID and PDM%ParticleInside in the caller's particle order
SUBROUTINE PICStage_GetParticles(PartState, GlobalElemID, ParticleInside)
REAL(C_DOUBLE),INTENT(OUT)     :: PartState(:,:)
INTEGER(C_INT32_T),INTENT(OUT) :: GlobalElemID(:)
LOGICAL,INTENT(OUT)            :: ParticleInside(:)
INTEGER(C_INT8_T) :: ins(SIZE(ParticleInside))
CALL Check(pic_get_particles(stage, PartState, GlobalElemID, ins), 'pic_get_particles')
ParticleInside = ins.NE.0
END SUBROUTINE PICStage_GetParticles

!> PS_N(iElem)%PartSource(1:4,0:N,0:N,0:N) for all elements as PartSource(1:4,0:N,0:N,0:N,1:nElems)
SUBROUTINE PICStage_GetPartSource(PartSource)
REAL(C_DOUBLE),INTENT(OUT) :: PartSource(:,:,:,:,:)
CALL Check(pic_get_part_source(stage, PartSource), 'pic_get_part_source')
END SUBROUTINE PICStage_GetPartSource

SUBROUTINE PICStage_GetFieldAtParticle(FieldAtParticle)
REAL(C_DOUBLE),INTENT(OUT) :: FieldAtParticle(:,:)
CALL Check(pic_get_field_at_particle(stage, FieldAtParticle), 'pic_get_field_at_particle')
END SUBROUTINE PICStage_GetFieldAtParticle

!> PartPosRef(1:3,nPart) of RefMapping (src/particles/particle_vars.f90:52)
SUBROUTINE PICStage_GetPartPosRef(PartPosRef)
REAL(C_DOUBLE),INTENT(OUT) :: PartPosRef(:,:)
CALL Check(pic_get_part_pos_ref(stage, PartPosRef), 'pic_get_part_pos_ref')
END SUBROUTINE PICStage_GetPartPosRef

!> PartCharge(1:3) of src/particles/pic/analyze/pic_analyze.f90:136-210: deposited charge, absolute and relative error
SUBROUTINE CalcDepositedCharge(PartCharge)
REAL(C_DOUBLE),INTENT(OUT) :: PartCharge(3)
CALL Check(pic_deposited_charge(stage, PartCharge), 'CalcDepositedCharge')
END SUBROUTINE CalcDepositedCharge

!> WEl of src/analyze/analyzefield_tools.f90:99-228 for the field last set
SUBROUTINE CalcPotentialEnergy(WEl)
REAL(C_DOUBLE),INTENT(OUT) :: WEl
CALL Check(pic_field_energy(stage, WEl), 'CalcPotentialEnergy')
END SUBROUTINE CalcPotentialEnergy

SUBROUTINE Deposition()
CALL Check(pic_h(stage), 'Deposition')
END SUBROUTINE Deposition

SUBROUTINE InterpolateFieldToParticle()
CALL Check(pic_interp(stage), 'InterpolateFieldToParticle')
END SUBROUTINE InterpolateFieldToParticle

!> LastPartPos <- PartState (stepper lines 109-110) followed by the push loop (lines 128-198)
SUBROUTINE PushBorisLeapfrog(dt)
REAL(C_DOUBLE),INTENT(IN) :: dt
CALL Check(pic_push(stage, dt), 'PushBorisLeapfrog')
END SUBROUTINE PushBorisLeapfrog

SUBROUTINE PerformTracking()
CALL Check(pic_track(stage), 'PerformTracking')
END SUBROUTINE PerformTracking

!> nsteps of the 508 stepper with a fixed field: Deposition, interpolate, push, track
SUBROUTINE TimeStepPoissonByBorisLeapfrog(dt, nsteps)
REAL(C_DOUBLE),INTENT(IN) :: dt
INTEGER,INTENT(IN)        :: nsteps
CALL Check(pic_time_step(stage, dt, INT(nsteps,C_INT32_T)), 'TimeStepPoissonByBorisLeapfrog')
END SUBROUTINE TimeStepPoissonByBorisLeapfrog

END MODULE MOD_PIC_Stage_C
