! Fortran parity test of the iso_c_binding shim: the reference's own TrotterStep
! / CanonicalFormAll (propagation_module.f90:42-114,617-634) and the HIP-backed
! ones from tebd_hip_propagation run the same chains; the reference's
! TotalEnergy, OneSiteExpVal and ChainEntropy (observables_module.f90) judge both.
! Prints one line per case and stops with code 1 when an observable differs by
! more than 1e-9.
PROGRAM shim_vs_reference
USE system_parameters
USE TEBDtools_module
USE Hamiltonian_tools_module
USE heisenberg_module
USE fermi_hubbard_module
USE local_operations_module, ONLY: SVDInit, SVDFinish
USE observables_module
USE propagation_module, ONLY: TrotterStep, CanonicalFormAll
USE tebd_hip_propagation, ONLY: TrotterStepHIP => TrotterStep, CanonicalFormAllHIP => CanonicalFormAll, &
                                SVDInitHIP => SVDInit, SVDFinishHIP => SVDFinish, tebd_hip_resident, &
                                tebd_hip_to_host, tebd_hip_mode
IMPLICIT NONE
REAL(KIND=rKind) :: worst
LOGICAL :: ok
ok = .TRUE.
print_switch = .FALSE.
BoundaryCond = 'O'
trotterOrder = 2

CALL run_case('xxz L=16 chi=16 real time', 1, 16, 16, CMPLX(0.02_rKind, 0.0_rKind, KIND=rKind), 30, .FALSE., .FALSE., worst)
ok = ok .AND. worst < 1e-9_rKind
CALL run_case('xxz L=16 chi=16 resident', 1, 16, 16, CMPLX(0.02_rKind, 0.0_rKind, KIND=rKind), 30, .FALSE., .TRUE., worst)
ok = ok .AND. worst < 1e-9_rKind
CALL run_case('fermi-hubbard L=8 chi=32', 3, 8, 32, CMPLX(0.02_rKind, 0.0_rKind, KIND=rKind), 20, .FALSE., .FALSE., worst)
ok = ok .AND. worst < 1e-9_rKind
CALL run_case('xxz L=10 chi=8 imaginary time', 1, 10, 8, CMPLX(0.0_rKind, -0.05_rKind, KIND=rKind), 20, .TRUE., &
              .FALSE., worst)
ok = ok .AND. worst < 1e-9_rKind
IF (.NOT. ok) STOP 1

CONTAINS

SUBROUTINE run_case(name, model, L, chi, dt, nsteps, canon, resident, worst)
  CHARACTER(len=*), INTENT(IN) :: name
  INTEGER, INTENT(IN) :: model, L, chi, nsteps
  COMPLEX(KIND=rKind), INTENT(IN) :: dt
  LOGICAL, INTENT(IN) :: canon, resident
  REAL(KIND=rKind), INTENT(OUT) :: worst
  TYPE(tensor), POINTER :: GA(:), GB(:)
  TYPE(vector), POINTER :: LA(:), LB(:)
  TYPE(matrix), POINTER :: H(:), U(:)
  COMPLEX(KIND=rKind), ALLOCATABLE :: carray(:, :)
  REAL(KIND=rKind), ALLOCATABLE :: szA(:), szB(:), hfield(:)
  REAL(KIND=rKind) :: eA, eB, tA, tB
  INTEGER :: i, k
  systemSize = L
  chiMax = chi
  IF (model == 1) THEN
    spin = 0.5_rKind
    spinSize = 2
    maxFilling = spinSize
    localSize = heisenbergLocalDim()
    CALL CreateHeisenbergOps()
  ELSE
    spin = 0.5_rKind
    spinSize = 2
    maxFilling = 2
    localSize = spinSFermiLocalDim()
    CALL CreateFermiSops()
  END IF
  CALL AllocateOps(H, L - 1, localSize * localSize)
  ALLOCATE(hfield(L))
  DO i = 1, L
    hfield(i) = 0.3_rKind * SIN(1.7_rKind * i)
  END DO
  IF (model == 1) THEN
    CALL HamiltonianHeisenberg(H, 1.0_rKind, 1.0_rKind, 0.7_rKind)
    DO i = 1, L - 1
      H(i)%m = H(i)%m + hfield(i) * HamiLeft(Sz_opS) + hfield(i + 1) * HamiRight(Sz_opS)
    END DO
    H(1)%m = H(1)%m + hfield(1) * HamiLeft(Sz_opS)
    H(L - 1)%m = H(L - 1)%m + hfield(L) * HamiRight(Sz_opS)
  ELSE
    CALL HamiltonianHubbard(H, 1.0_rKind, 4.0_rKind, 0.0_rKind)
  END IF
  CALL AllocateProp(U)
  CALL ConstructPropagators(H, U, dt)
  ALLOCATE(carray(localSize, L))
  carray = 0.0_rKind
  DO i = 1, L
    IF (canon) THEN
      carray(:, i) = 1.0_rKind / SQRT(1.0_rKind * localSize)
    ELSE IF (model == 1) THEN
      carray(1 + MOD(i + 1, 2), i) = 1.0_rKind
    ELSE
      carray(2 + MOD(i + 1, 2), i) = 1.0_rKind
    END IF
  END DO
  CALL AllocateGamLam(GA, LA, chi)
  CALL AllocateGamLam(GB, LB, chi)
  CALL ProductStateMPD(GA, LA, carray)
  CALL ProductStateMPD(GB, LB, carray)
  CALL SVDInit(chi)
  tebd_hip_resident = resident
  CALL SVDInitHIP(chi)
  DO k = 1, nsteps
    CALL TrotterStep(U, GA, LA, tA)
    IF (canon) CALL CanonicalFormAll(GA, LA)
    CALL TrotterStepHIP(U, GB, LB, tB)
    IF (canon) CALL CanonicalFormAllHIP(GB, LB)
  END DO
  IF (resident) CALL tebd_hip_to_host(GB, LB)
  ALLOCATE(szA(L), szB(L))
  CALL TotalEnergy(eA, H, GA, LA)
  CALL TotalEnergy(eB, H, GB, LB)
  CALL OneSiteExpVal(szA, Sz_opS%mr, GA, LA)
  CALL OneSiteExpVal(szB, Sz_opS%mr, GB, LB)
  worst = MAX(ABS(eA - eB), MAXVAL(ABS(szA - szB)))
  DO i = 2, L
    worst = MAX(worst, ABS(ChainEntropy(i, LA) - ChainEntropy(i, LB)))
  END DO
  PRINT '(A,A,A,ES10.3,A,ES12.5,A,ES12.5)', 'shim ', name, ': max observable difference ', worst, &
        '  energy ', eA, ' vs ', eB
  CALL SVDFinishHIP()
  CALL SVDFinish()
  tebd_hip_resident = .FALSE.
END SUBROUTINE run_case

END PROGRAM shim_vs_reference
