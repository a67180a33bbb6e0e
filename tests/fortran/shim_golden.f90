! Runs a golden job through the reference's call surface as the iso_c_binding shim
! provides it (SVDInit, TrotterStep, CanonicalFormAll, SVDFinish, ImagTimeProp, and the
! number-conserving TrotterStepNC and ImagTimePropNC from
! tebd_hip_propagation) and writes the final state, for tests/test_fortran_shim.py to
! compare with the reference's own outputs (tests/golden).
!
! Usage: shim_golden JOB OUT
!   JOB (stream, little-endian): int32 L, d, chi, nsteps, canon, resident, mode, kind;
!       kind 0 (Trotter steps): (L-1) gates of d^2 x d^2 complex(8), column-major;
!       kind 1 (ImagTimeProp, chi = chiMax): int32 chiIn, maxITPsteps, stepsForJudge,
!       real(8) dtITP, convCriterion1, convCriterion2, then (L-1) bond Hamiltonians;
!       kind 2 (TrotterStepNC): int32 totNum, intDegFree (0/1), Conserv(d), then the
!       LabelLeft and LabelRight of the L+1 links (chi int32 each), then the gates as kind 0;
!       kind 3 (ImagTimePropNC, chi = chiMax): the settings of kind 1, then those of kind 2,
!       then (L-1) bond Hamiltonians;
!       kind 4 (TrotterStepNC under another schedule): int32 trotterOrder, periodic (0/1) and
!       the number of gates of the table, then kind 2 with that many gates;
!       then L Gammas of chi x d x chi complex(8); L+1 lambdas of chi real(8)
!   OUT: real(8) total truncation error; L Gammas; L+1 lambdas (same layout); for kinds 2,
!        3 and 4 then the L+1 LabelLeft and L+1 LabelRight (chi int32 each)
PROGRAM shim_golden
USE, INTRINSIC :: iso_c_binding
USE system_parameters, ONLY: rKind, systemSize, localSize, trotterOrder, BoundaryCond, chiMax, dtITP, &
                             maxITPsteps, stepsForJudge, convCriterion1, convCriterion2
USE system_parameters, ONLY: totNum
USE TEBDtools_module, ONLY: tensor, vector, matrix, vectorInt
USE Hamiltonian_tools_module, ONLY: Conserv
USE tebd_hip_propagation, ONLY: SVDInit, SVDFinish, TrotterStep, CanonicalFormAll, ImagTimeProp, &
                                TrotterStepNC, ImagTimePropNC, tebd_hip_resident, tebd_hip_to_host, &
                                tebd_hip_mode
IMPLICIT NONE
INTEGER(c_int32_t) :: hdr(8), itp(3)
REAL(c_double) :: itpr(3)
INTEGER :: L, d, chi, nsteps, i, k, uin, uout, chiIn
TYPE(tensor), POINTER :: Gammas(:)
TYPE(vector), POINTER :: Lambdas(:)
TYPE(matrix), POINTER :: Udt(:)
TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
INTEGER(c_int32_t) :: nch(2), sch(3)
INTEGER :: ng
REAL(KIND=rKind) :: truncerr, total
CHARACTER(len=1024) :: fin, fout

CALL get_command_argument(1, fin)
CALL get_command_argument(2, fout)
OPEN(newunit=uin, file=TRIM(fin), access='stream', form='unformatted', status='old')
READ(uin) hdr
L = hdr(1); d = hdr(2); chi = hdr(3); nsteps = hdr(4)
systemSize = L
localSize = d
trotterOrder = 2
BoundaryCond = 'O'
tebd_hip_resident = hdr(6) /= 0
tebd_hip_mode = INT(hdr(7), c_int)
IF (hdr(8) == 1 .OR. hdr(8) == 3) THEN
  READ(uin) itp, itpr
  chiIn = itp(1)
  maxITPsteps = itp(2)
  stepsForJudge = itp(3)
  dtITP = itpr(1)
  convCriterion1 = itpr(2)
  convCriterion2 = itpr(3)
  chiMax = chi
END IF
ng = L - 1
IF (hdr(8) == 4) THEN
  READ(uin) sch
  trotterOrder = sch(1)
  IF (sch(2) /= 0) BoundaryCond = 'P'
  ng = sch(3)
END IF
IF (hdr(8) == 2 .OR. hdr(8) == 3 .OR. hdr(8) == 4) THEN
  READ(uin) nch
  totNum = nch(1)
  ALLOCATE(Conserv%vi(d))
  READ(uin) Conserv%vi
  ALLOCATE(LabelLeft(L + 1), LabelRight(L + 1))
  DO i = 1, L + 1
    ALLOCATE(LabelLeft(i)%vi(chi))
    READ(uin) LabelLeft(i)%vi
  END DO
  DO i = 1, L + 1
    ALLOCATE(LabelRight(i)%vi(chi))
    READ(uin) LabelRight(i)%vi
  END DO
END IF
ALLOCATE(Udt(ng), Gammas(L), Lambdas(L + 1))
DO i = 1, ng
  ALLOCATE(Udt(i)%m(d * d, d * d))
  READ(uin) Udt(i)%m
END DO
DO i = 1, L
  ALLOCATE(Gammas(i)%t(chi, d, chi))
  READ(uin) Gammas(i)%t
END DO
DO i = 1, L + 1
  ALLOCATE(Lambdas(i)%v(chi))
  READ(uin) Lambdas(i)%v
END DO
CLOSE(uin)

total = 0.0_rKind
IF (hdr(8) == 1) THEN
  ! Udt holds the bond Hamiltonians of the job here
  CALL ImagTimeProp(Udt, Gammas, Lambdas, chiIn)
ELSE IF (hdr(8) == 3) THEN
  IF (nch(2) /= 0) THEN
    CALL ImagTimePropNC(Udt, Gammas, Lambdas, LabelLeft, LabelRight, chiIn, 1)
  ELSE
    CALL ImagTimePropNC(Udt, Gammas, Lambdas, LabelLeft, LabelRight, chiIn)
  END IF
ELSE IF (hdr(8) == 2 .OR. hdr(8) == 4) THEN
  CALL SVDInit(chi)
  DO k = 1, nsteps
    IF (nch(2) /= 0) THEN
      CALL TrotterStepNC(Udt, Gammas, Lambdas, LabelLeft, LabelRight, truncerr, 1)
    ELSE
      CALL TrotterStepNC(Udt, Gammas, Lambdas, LabelLeft, LabelRight, truncerr)
    END IF
    total = total + truncerr
  END DO
  IF (tebd_hip_resident) CALL tebd_hip_to_host(Gammas, Lambdas)
ELSE
  CALL SVDInit(chi)
  DO k = 1, nsteps
    CALL TrotterStep(Udt, Gammas, Lambdas, truncerr)
    total = total + truncerr
    IF (hdr(5) /= 0) CALL CanonicalFormAll(Gammas, Lambdas)
  END DO
  IF (tebd_hip_resident) CALL tebd_hip_to_host(Gammas, Lambdas)
END IF
CALL SVDFinish()

OPEN(newunit=uout, file=TRIM(fout), access='stream', form='unformatted', status='replace')
WRITE(uout) total
DO i = 1, L
  WRITE(uout) Gammas(i)%t
END DO
DO i = 1, L + 1
  WRITE(uout) Lambdas(i)%v
END DO
IF (hdr(8) == 2 .OR. hdr(8) == 3 .OR. hdr(8) == 4) THEN
  DO i = 1, L + 1
    WRITE(uout) LabelLeft(i)%vi
  END DO
  DO i = 1, L + 1
    WRITE(uout) LabelRight(i)%vi
  END DO
END IF
CLOSE(uout)
PRINT '(A,I0,A,I0)', 'shim_golden: steps ', nsteps, ' sites ', L
END PROGRAM shim_golden
