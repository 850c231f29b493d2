!==================================================================================================================================
! pic_stage_driver: a 508-shaped Fortran host loop over MOD_PIC_Stage_C.
!
! Reads a case file (stream, little endian) holding the HOPR mesh arrays, the parameters, U_N%E and the particles, runs
! nSteps of Deposition / InterpolateFieldToParticle / push / PerformTracking (mode 0: stage by stage as the reference's
! TimeStepPoissonByBorisLeapfrog calls them; mode 1: the fused call), deposits once more and writes PartState,
! PEM%GlobalElemID, PDM%ParticleInside, PS_N%PartSource, PartCharge(1:3) and WEl to the result file.
!
! usage: pic_stage_driver <case.bin> <result.bin>
!==================================================================================================================================
PROGRAM pic_stage_driver
USE ISO_C_BINDING
USE MOD_PIC_Stage_C
IMPLICIT NONE
CHARACTER(LEN=1024)            :: inpath, outpath
INTEGER(C_INT32_T)             :: hdr(16)
INTEGER                        :: nElems, nSides, nNodes, NGeo, N, nBCs, nPV, nPart, nSteps, mode, nSpecies, iStep, nDof
INTEGER(C_INT32_T),ALLOCATABLE :: ElemInfo(:,:), SideInfo(:,:), GlobalNodeIDs(:), BCType(:,:), PartBoundCond(:)
INTEGER(C_INT32_T),ALLOCATABLE :: GlobalElemID(:), PartSpecies(:)
REAL(C_DOUBLE)                 :: PartCharge(3), WEl
REAL(C_DOUBLE),ALLOCATABLE     :: NodeCoords(:,:), PeriodicVectors(:,:), E(:,:,:,:,:), PartState(:,:), PS(:,:,:,:,:)
LOGICAL,ALLOCATABLE            :: IsNewPart(:), ParticleInside(:)
INTEGER(C_INT8_T),ALLOCATABLE  :: insideOut(:)
REAL(C_DOUBLE)                 :: dt, rpar(8)
REAL(C_DOUBLE),ALLOCATABLE     :: spec(:,:)
TYPE(tPicParams)               :: par

CALL GET_COMMAND_ARGUMENT(1, inpath)
CALL GET_COMMAND_ARGUMENT(2, outpath)
OPEN(10, FILE=TRIM(inpath), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
READ(10) hdr
nElems=hdr(1); nSides=hdr(2); nNodes=hdr(3); NGeo=hdr(4); N=hdr(5); nBCs=hdr(6); nPV=hdr(7); nPart=hdr(8)
nSteps=hdr(9); mode=hdr(10); nSpecies=hdr(13)
nDof=(N+1)**3
ALLOCATE(ElemInfo(6,nElems), SideInfo(5,nSides), NodeCoords(3,nNodes), GlobalNodeIDs(nNodes), BCType(4,nBCs))
ALLOCATE(PartBoundCond(nBCs), PeriodicVectors(3,MAX(nPV,1)), E(3,0:N,0:N,0:N,nElems), spec(3,nSpecies))
ALLOCATE(PartState(6,nPart), GlobalElemID(nPart), PartSpecies(nPart), IsNewPart(nPart), ParticleInside(nPart))
ALLOCATE(insideOut(nPart), PS(4,0:N,0:N,0:N,nElems))
READ(10) ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs, BCType, PartBoundCond
IF (nPV.GT.0) READ(10) PeriodicVectors(:,1:nPV)
READ(10) dt, rpar        ! rpar = sf_radius, refmapping_eps, externalField(1:6)
READ(10) spec            ! ChargeIC, MassIC, MacroParticleFactor per species
READ(10) E
READ(10) PartState, GlobalElemID, PartSpecies
CLOSE(10)
IsNewPart = .FALSE.

par%tracking = MERGE(PIC_TRACK_REFMAPPING, PIC_TRACK_TRIA, hdr(14).EQ.PIC_TRACK_REFMAPPING)
par%deposition = hdr(11)
par%sf_alpha = hdr(12)
par%sf_radius = rpar(1)
par%refmapping_guess = MAX(hdr(15), 1)
par%refmapping_eps = rpar(2)
par%external_field = rpar(3:8)
par%nSpecies = nSpecies
par%charge = 0.; par%mass = 1.; par%mpf = 1.
par%charge(1:nSpecies) = spec(1,:)
par%mass(1:nSpecies)   = spec(2,:)
par%mpf(1:nSpecies)    = spec(3,:)
par%precision = 0
par%max_particles = nPart
par%device = 0

IF (nPV.GT.0) THEN
  CALL InitPICStage(ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs, BCType, PartBoundCond, PeriodicVectors(:,1:nPV), NGeo, N, par)
ELSE
  CALL InitPICStage(ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs, BCType, PartBoundCond, PeriodicVectors(:,1:0), NGeo, N, par)
END IF
CALL PICStage_SetField(E)
CALL PICStage_SetParticles(PartState, GlobalElemID, PartSpecies, IsNewPart)
DO iStep = 1, nSteps
  IF (mode.EQ.0) THEN
    CALL Deposition()
    ! (HDG(time,iter) would run here)
    CALL InterpolateFieldToParticle()
    CALL PushBorisLeapfrog(dt)
    CALL PerformTracking()
  ELSE
    CALL TimeStepPoissonByBorisLeapfrog(dt, 1)
  END IF
END DO
CALL Deposition()
CALL PICStage_GetParticles(PartState, GlobalElemID, ParticleInside)
CALL PICStage_GetPartSource(PS)
CALL CalcDepositedCharge(PartCharge)
CALL CalcPotentialEnergy(WEl)
CALL FinalizePICStage()
insideOut = MERGE(1_C_INT8_T, 0_C_INT8_T, ParticleInside)
OPEN(11, FILE=TRIM(outpath), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
WRITE(11) PartState, GlobalElemID, insideOut, PS, PartCharge, WEl
CLOSE(11)
WRITE(*,'(A,I0,A,I0,A)') 'pic_stage_driver: ', nSteps, ' steps of ', nPart, ' particles done'
END PROGRAM pic_stage_driver
