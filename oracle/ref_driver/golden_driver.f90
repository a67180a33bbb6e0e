! Golden-vector driver: runs the reference TEBD modules on one job file and
! prints every quantity the parity tests compare.  It is test infrastructure:
! it is compiled only against the reference sources (recipe: oracle/Makefile,
! target `ref`), its binary lives in oracle/_ref/, and it never ships.
!
! The driver follows the reference's own mains:
!   model setup      Case_Studies/Heisenberg_Wrapper/XXDynamics.f90:62-101
!   state + gates    XXDynamics.f90:104-132 (AllocateGamLam, ProductStateMPD,
!                    AllocateProp, ConstructPropagators)
!   time loop        XXDynamics.f90:151-178 (SVDInit, TrotterStep)
!   imaginary time   propagation_module.f90:656-662,741-743 (dt=-i*dtau,
!                    TrotterStep then CanonicalFormAll)
! Site-dependent fields use the reference's half-split helpers
! (Hamiltonian_tools_module.f90:816-886), exactly as the uniform-field branch
! of HamiltonianHeisenberg (heisenberg_module.f90:158-167) does.
!
! Job file (list-directed):
!   model L param chi         model 1=Heisenberg(param=2*spin) 2=BH JVector
!                             (param=maxFilling) 3=Fermi-Hubbard (param=maxFilling)
!   c1 c2 c3                  Heisenberg: Jx Jy Jz ; BH: U0 mu0 V0 ; FH: t U0 mu0
!   field(1:nf)               Heisenberg: h_i (nf=L) ; BH: J_i (nf=L-1) ; FH: none
!   dtre dtim nsteps canon    complex dt passed to ConstructPropagators,
!                             canon=1 runs CanonicalFormAll after each step
!   (re,im) coefficients of the product state, d per site, L sites
!   mode                      0 = trajectory, 1 = single TwoSiteOp on a given state,
!                             2 = ImagTimeProp (propagation_module.f90:636-760) with chi = chiMax,
!                             3 = trajectory with the Trotter order and boundary of the next line
!   [mode 3: trotterOrder BoundaryCond  (2 or 5, 'O' or 'P'); periodic adds H(L), the L-1 edge]
!   mode 4 = trajectory, then the final state written with the reference's checkpoint
!            writers (io_module RecordLambdas / RecordGammas) in both kinds, 'S' and 'B',
!            to ckpt_lambdas_S.dat, ckpt_gammas_S.dat, ckpt_lambdas_B.dat, ckpt_gammas_B.dat
!            in the working directory
!   mode 5 = trajectory (trotterOrder BoundaryCond line as mode 3), then EvaluateMeasures
!            (observables_module.f90:2286-2526) with local and averaged Sz (spins) or n,
!            the correlator of the bond two-site operator Sz x Sz (spins) or n x n, for
!            fermions the fermi-phase correlator c+_up x c_up, and entropies of level 2
!   mode 6 = number-conserving trajectory: the product state's labels (AllocateLabel,
!            ProductStateLabels, Hamiltonian_tools_module.f90:124-178), then TrotterStepNC
!            (propagation_module.f90:890-1017, TwoSiteOpNC local_operations_module.f90:1845-1964);
!            the next line: totNum intDegFree (0: number = state index - 1, 1: the model's
!            Conserv); the labels of every link are printed with the final state
!   mode 7 = number-conserving imaginary time: the product state's labels as in mode 6, then
!            ImagTimePropNC (propagation_module.f90:1986-2121: chi stages, TrotterStepNC and
!            CanonicalFormAllNC :1943-1983 every step); next lines: totNum intDegFree, then
!            chiIn dtITP maxITPsteps stepsForJudge convCriterion1 convCriterion2; the final
!            state and every link's labels are printed
!   mode 8 = number-conserving trajectory as mode 6 under the Trotter order and boundary of
!            the next line (as mode 3; TrotterStep5thOrderNC, TrotterStep2ndOrderPBCNC,
!            TrotterStep5thOrderPBCNC, propagation_module.f90:1019-1942, with SwappingNC
!            local_operations_module.f90:1966-2082); for a periodic Bose-Hubbard chain a
!            further line holds the edge tunnelling J_L; then the line totNum intDegFree
!   [mode 1: link, then Gamma(chi,d,chi) re im for each site, Lambda(chi) for L+1 links]
!   [mode 2: chiIn dtITP maxITPsteps stepsForJudge convCriterion1 convCriterion2]
PROGRAM golden_driver
USE system_parameters
USE TEBDtools_module
USE Hamiltonian_tools_module
USE heisenberg_module
USE bose_hubbard_module
USE fermi_hubbard_module
USE local_operations_module
USE observables_module
USE propagation_module
USE io_module, ONLY: RecordLambdas, RecordGammas, openUnit
IMPLICIT NONE
TYPE(tensor), POINTER :: Gammas(:)
TYPE(vector), POINTER :: Lambdas(:)
TYPE(matrix), POINTER :: H(:), U(:)
COMPLEX(KIND=rKind), ALLOCATABLE :: carray(:,:)
REAL(KIND=rKind), ALLOCATABLE :: field(:), expl(:), expn(:), buf(:)
REAL(KIND=rKind) :: c1, c2, c3, dtre, dtim, err, energy, totTrunc
REAL(KIND=rKind), ALLOCATABLE :: sz(:,:), nop(:,:)
INTEGER :: model, L, param, chi, nsteps, canon, i, j, k, step, nf, mode, link, chiIn
INTEGER :: a, b, dd, nH, nfc
TYPE(measure) :: Meas
TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
INTEGER :: intdeg
CHARACTER :: bcc
CHARACTER(len=256) :: jobfile

CALL GET_COMMAND_ARGUMENT(1, jobfile)
OPEN(UNIT=11, FILE=TRIM(jobfile), STATUS='OLD', ACTION='READ')
READ(11,*) model, L, param, chi
READ(11,*) c1, c2, c3
systemSize = L
chiMax = chi
print_switch = .FALSE.
BoundaryCond = 'O'
trotterOrder = 2

IF (model == 1) THEN
  spin = 0.5_rKind*param
  spinSize = FLOOR(2.0_rKind*spin+1.0_rKind)
  maxFilling = spinSize
  localSize = heisenbergLocalDim()
  CALL CreateHeisenbergOps()
  nf = L
ELSE IF (model == 2) THEN
  maxFilling = param
  localSize = maxFilling+1
  CALL CreateFieldOps()
  nf = L-1
ELSE
  spin = 0.5_rKind
  spinSize = 2
  maxFilling = param
  localSize = spinSFermiLocalDim()
  CALL CreateFermiSops()
  nf = 0
END IF
dd = localSize
ALLOCATE(field(MAX(nf,L)))
field = 0.0_rKind
IF (nf > 0) READ(11,*) (field(i), i=1,nf)
READ(11,*) dtre, dtim, nsteps, canon
ALLOCATE(carray(dd, L), buf(2*dd))
DO i=1,L
  READ(11,*) (buf(k), k=1,2*dd)
  DO k=1,dd
    carray(k,i) = CMPLX(buf(2*k-1), buf(2*k), KIND=rKind)
  END DO
END DO
READ(11,*) mode
nH = L-1
IF (mode == 3 .OR. mode == 5 .OR. mode == 8) THEN
  READ(11,*) trotterOrder, bcc
  BoundaryCond = bcc
  IF (BoundaryCond == 'P') nH = L
  IF (mode == 8 .AND. BoundaryCond == 'P' .AND. model == 2) READ(11,*) field(L)
END IF

CALL AllocateOps(H, nH, dd*dd)
IF (model == 1) THEN
  CALL HamiltonianHeisenberg(H, c1, c2, c3)
  DO i=1,L-1
    H(i)%m = H(i)%m + field(i)*HamiLeft(Sz_opS) + field(i+1)*HamiRight(Sz_opS)
  END DO
  H(1)%m = H(1)%m + field(1)*HamiLeft(Sz_opS)
  H(L-1)%m = H(L-1)%m + field(L)*HamiRight(Sz_opS)
ELSE IF (model == 2) THEN
  CALL HamiltonianBoseHubbardJVector(H, field, c1, c2, c3)
ELSE
  CALL HamiltonianHubbard(H, c1, c2, c3)
END IF

! one-site operators for the observables: Sz (spins) or n (bosons, fermions)
ALLOCATE(sz(dd,dd), nop(dd,dd))
sz = 0.0_rKind
nop = 0.0_rKind
IF (model == 1) THEN
  sz = Sz_opS%mr
  DO k=1,dd
    nop(k,k) = k-1
  END DO
ELSE IF (model == 2) THEN
  nop = MATMUL(TRANSPOSE(a_op%mr), a_op%mr)
ELSE
  sz = Sz_opS%mr
  nop = ntot_opS%mr
END IF

PRINT '(A,3I6)', 'DIMS', L, dd, chi
DO i=1,nH
  PRINT '(A,I6)', 'H', i
  DO a=1,dd*dd
    PRINT '(64ES26.17E3)', (REAL(H(i)%m(a,b)), AIMAG(H(i)%m(a,b)), b=1,dd*dd)
  END DO
END DO
PRINT '(A)', 'SZOP'
DO a=1,dd
  PRINT '(16ES26.17E3)', (sz(a,b), b=1,dd)
END DO
PRINT '(A)', 'NOP'
DO a=1,dd
  PRINT '(16ES26.17E3)', (nop(a,b), b=1,dd)
END DO
IF (model == 3) THEN
  DO k=1,spinSize
    PRINT '(A,I3)', 'AOP', k
    DO a=1,dd
      PRINT '(16ES26.17E3)', (a_opS(k)%mr(a,b), b=1,dd)
    END DO
  END DO
END IF

CALL AllocateProp(U)
CALL ConstructPropagators(H, U, CMPLX(dtre, dtim, KIND=rKind))
DO i=1,SIZE(U)
  PRINT '(A,I6)', 'U', i
  DO a=1,dd*dd
    PRINT '(64ES26.17E3)', (REAL(U(i)%m(a,b)), AIMAG(U(i)%m(a,b)), b=1,dd*dd)
  END DO
END DO

CALL AllocateGamLam(Gammas, Lambdas, chi)
CALL ProductStateMPD(Gammas, Lambdas, carray)
CALL SVDInit(chi)

IF (mode == 1) THEN
  ! single two-site update on an explicit state read from the job file
  READ(11,*) link
  DEALLOCATE(buf)
  ALLOCATE(buf(2*chi*dd*chi))
  DO i=1,L
    READ(11,*) (buf(k), k=1,2*chi*dd*chi)
    k = 0
    DO b=1,chi
      DO j=1,dd
        DO a=1,chi
          Gammas(i)%t(a,j,b) = CMPLX(buf(2*k+1), buf(2*k+2), KIND=rKind)
          k = k+1
        END DO
      END DO
    END DO
  END DO
  DO i=1,L+1
    READ(11,*) (Lambdas(i)%v(k), k=1,chi)
  END DO
  CALL TwoSiteOp(link, U(link)%m, Gammas, Lambdas, err)
  PRINT '(A,ES26.17E3)', 'TRUNC', err
  CALL PrintState()
  STOP
END IF

ALLOCATE(expl(L), expn(L))
IF (mode == 2) THEN
  ! imaginary-time ground state through the reference's own driver routine; it prints its
  ! convergence judgements ("ITP step j ...", "Percent difference ...") itself
  READ(11,*) chiIn, dtITP, maxITPsteps, stepsForJudge, convCriterion1, convCriterion2
  CALL SVDFinish()
  CALL ImagTimeProp(H, Gammas, Lambdas, chiIn)
  CALL PrintObs(0, 0.0_rKind)
  CALL PrintState()
  STOP
END IF
IF (mode == 7) THEN
  READ(11,*) totNum, intdeg
  READ(11,*) chiIn, dtITP, maxITPsteps, stepsForJudge, convCriterion1, convCriterion2
  CALL SVDFinish()
  CALL AllocateLabel(LabelLeft, LabelRight, chi)
  IF (intdeg == 1) THEN
    CALL ProductStateLabels(LabelLeft, LabelRight, carray, 1)
    PRINT '(A,64I4)', 'CONSERV', (Conserv%vi(k), k=1,dd)
    CALL ImagTimePropNC(H, Gammas, Lambdas, LabelLeft, LabelRight, chiIn, 1)
  ELSE
    CALL ProductStateLabels(LabelLeft, LabelRight, carray)
    CALL ImagTimePropNC(H, Gammas, Lambdas, LabelLeft, LabelRight, chiIn)
  END IF
  CALL PrintObs(0, 0.0_rKind)
  CALL PrintState()
  DO i=1,L+1
    PRINT '(A,I6)', 'LABELL', i
    PRINT '(1024I7)', (LabelLeft(i)%vi(k), k=1,chi)
    PRINT '(A,I6)', 'LABELR', i
    PRINT '(1024I7)', (LabelRight(i)%vi(k), k=1,chi)
  END DO
  STOP
END IF
IF (mode == 6 .OR. mode == 8) THEN
  READ(11,*) totNum, intdeg
  CALL SVDFinish()  ! the NC update allocates its own SVD workspace per block (SVDInitNC)
  CALL AllocateLabel(LabelLeft, LabelRight, chi)
  IF (intdeg == 1) THEN
    CALL ProductStateLabels(LabelLeft, LabelRight, carray, 1)
  ELSE
    CALL ProductStateLabels(LabelLeft, LabelRight, carray)
  END IF
  IF (intdeg == 1) PRINT '(A,64I4)', 'CONSERV', (Conserv%vi(k), k=1,dd)
  CALL PrintObs(0, 0.0_rKind)
  DO step=1,nsteps
    IF (intdeg == 1) THEN
      CALL TrotterStepNC(U, Gammas, Lambdas, LabelLeft, LabelRight, err, 1)
    ELSE
      CALL TrotterStepNC(U, Gammas, Lambdas, LabelLeft, LabelRight, err)
    END IF
    CALL PrintObs(step, err)
  END DO
  CALL PrintState()
  DO i=1,L+1
    PRINT '(A,I6)', 'LABELL', i
    PRINT '(1024I7)', (LabelLeft(i)%vi(k), k=1,chi)
    PRINT '(A,I6)', 'LABELR', i
    PRINT '(1024I7)', (LabelRight(i)%vi(k), k=1,chi)
  END DO
  STOP
END IF
totTrunc = 0.0_rKind
CALL PrintObs(0, 0.0_rKind)
DO step=1,nsteps
  CALL TrotterStep(U, Gammas, Lambdas, err)
  IF (canon == 1) CALL CanonicalFormAll(Gammas, Lambdas)
  totTrunc = totTrunc + err
  CALL PrintObs(step, err)
END DO
CALL PrintState()
IF (mode == 5) THEN
  nfc = 0
  IF (model == 3) nfc = 1
  CALL AllocateMeasures(Meas, 1, 1, 1, nfc, 2)
  IF (model == 2) THEN
    Meas%local(1)%Op = nop
    Meas%avg(1)%Op = nop
    Meas%corr(1)%Op = tensorProd(nop, nop)
  ELSE
    Meas%local(1)%Op = sz
    Meas%avg(1)%Op = sz
    Meas%corr(1)%Op = tensorProd(sz, sz)
  END IF
  IF (model == 3) Meas%fermicorr(1)%Op = tensorProd(TRANSPOSE(a_opS(1)%mr), a_opS(1)%mr)
  CALL EvaluateMeasures(Meas, Gammas, Lambdas, H)
  PRINT '(A,ES26.17E3)', 'MEN', Meas%en
  PRINT '(A,ES26.17E3)', 'MQME', Meas%ent%qme
  PRINT '(A,256ES26.17E3)', 'MLOCAL', (REAL(Meas%local(1)%value(i)), AIMAG(Meas%local(1)%value(i)), i=1,L)
  PRINT '(A,2ES26.17E3)', 'MAVG', REAL(Meas%avg(1)%value), AIMAG(Meas%avg(1)%value)
  PRINT '(A,256ES26.17E3)', 'MVN', (Meas%ent%vN(i), i=1,L)
  PRINT '(A,256ES26.17E3)', 'MCHAIN', (Meas%ent%chain(i), i=1,L+1)
  DO i=1,L
    PRINT '(A,I6)', 'MTB', i
    PRINT '(256ES26.17E3)', (Meas%ent%tbvN(i,j), j=1,L)
  END DO
  DO i=1,L
    PRINT '(A,I6)', 'MCORR', i
    PRINT '(512ES26.17E3)', (REAL(Meas%corr(1)%value(i,j)), AIMAG(Meas%corr(1)%value(i,j)), j=1,L)
  END DO
  IF (nfc > 0) THEN
    DO i=1,L
      PRINT '(A,I6)', 'MFCORR', i
      PRINT '(512ES26.17E3)', (REAL(Meas%fermicorr(1)%value(i,j)), AIMAG(Meas%fermicorr(1)%value(i,j)), j=1,L)
    END DO
  END IF
END IF
IF (mode == 4) THEN
  CALL openUnit('ckpt_lambdas_S.dat', 21)
  CALL RecordLambdas(21, Lambdas, 'S')
  CLOSE(21)
  CALL openUnit('ckpt_gammas_S.dat', 22)
  CALL RecordGammas(22, Gammas, 'S')
  CLOSE(22)
  CALL openUnit('ckpt_lambdas_B.dat', 23)
  CALL RecordLambdas(23, Lambdas, 'B')
  CLOSE(23)
  CALL openUnit('ckpt_gammas_B.dat', 24)
  CALL RecordGammas(24, Gammas, 'B')
  CLOSE(24)
END IF

CONTAINS

SUBROUTINE PrintObs(st, tr)
INTEGER, INTENT(IN) :: st
REAL(KIND=rKind), INTENT(IN) :: tr
INTEGER :: ii
  CALL TotalEnergy(energy, H, Gammas, Lambdas)
  CALL OneSiteExpVal(expl, sz, Gammas, Lambdas)
  CALL OneSiteExpVal(expn, nop, Gammas, Lambdas)
  PRINT '(A,I6,2ES26.17E3)', 'STEP', st, energy, tr
  PRINT '(A,256ES26.17E3)', 'SZ', (expl(ii), ii=1,L)
  PRINT '(A,256ES26.17E3)', 'N', (expn(ii), ii=1,L)
  PRINT '(A,256ES26.17E3)', 'S', (ChainEntropy(ii, Lambdas), ii=2,L)
END SUBROUTINE PrintObs

SUBROUTINE PrintState()
INTEGER :: ii, aa, jj, bb
  DO ii=1,L+1
    PRINT '(A,I6)', 'LAMBDA', ii
    PRINT '(1024ES26.17E3)', (Lambdas(ii)%v(aa), aa=1,chi)
  END DO
  DO ii=1,L
    PRINT '(A,I6)', 'GAMMA', ii
    PRINT '(4096ES26.17E3)', (((REAL(Gammas(ii)%t(aa,jj,bb)), AIMAG(Gammas(ii)%t(aa,jj,bb)), &
         aa=1,chi), jj=1,dd), bb=1,chi)
  END DO
END SUBROUTINE PrintState

END PROGRAM golden_driver
