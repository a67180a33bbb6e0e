! The reference's propagation call surface, backed by libtebdhip.so.
!
! Same names and argument lists as the reference:
!   SVDInit(chi), SVDFinish()                 local_operations_module.f90:84-153
!   TwoSiteOp(link, Op2, Gammas, Lambdas, truncerr)
!                                             local_operations_module.f90:46-48,463-494
!   TrotterStep(Udt, Gammas, Lambdas, totalTruncerr)
!                                             propagation_module.f90:42-114
!   CanonicalFormAll(Gammas, Lambdas)         propagation_module.f90:617-634
!   ImagTimeProp(H, Gammas, Lambdas, chiIn[, intDegFree]), ImagTimePropSpin(H, Gammas,
!   Lambdas, chiIn)                           propagation_module.f90:636-876
!   ConstructPropagators(H, U, dt)            TEBDtools_module.f90:97-99,1079-1101
!   TwoSiteOpNC(link, Op2, Gammas, Lambdas, LabelLeft, LabelRight, truncerr[, intDegFree]),
!   TrotterStepNC(Udt, Gammas, Lambdas, LabelLeft, LabelRight, totalTruncerr[, intDegFree])
!                                             local_operations_module.f90:1845-1964,
!                                             propagation_module.f90:890-1241
!   CanonicalFormAllNC(Gammas, Lambdas, LabelLeft, LabelRight[, intDegFree]),
!   ImagTimePropNC(H, Gammas, Lambdas, LabelLeft, LabelRight, chiIn[, intDegFree])
!                                             propagation_module.f90:1943-2121
! on the reference's own types (TYPE(tensor), TYPE(vector), TYPE(matrix),
! TEBDtools_module.f90:31-45) and globals (systemSize, localSize, trotterOrder,
! BoundaryCond, and for imaginary time chiMax, dtITP, maxITPsteps, stepsForJudge,
! convCriterion1/2 from system_parameters; for number conservation totNum, the labels'
! TYPE(vectorInt) and, with intDegFree, Conserv from Hamiltonian_tools_module).  Build it after the reference's
! system_parameters and TEBDtools_module and link libtebdhip.so; a program that
! also needs the reference's own procedures of the same names renames one set
! on USE (e.g. USE tebd_hip_propagation, TrotterStepHIP => TrotterStep).
!
! Like the reference's module-global SVD workspace, the device context is module
! state: SVDInit(chi) creates it for chains of systemSize sites with local
! dimension localSize and bond dimension chi, SVDFinish() frees it.
!
! Host arrays stay authoritative by default: every call uploads Gammas/Lambdas,
! runs on the GPU and downloads the result.  With tebd_hip_resident = .TRUE. the
! state stays on the device between calls; call tebd_hip_to_host(Gammas,
! Lambdas) before reading it on the host and tebd_hip_to_device after changing it.
MODULE tebd_hip_propagation
USE, INTRINSIC :: iso_c_binding
USE system_parameters, ONLY: rKind, systemSize, localSize, trotterOrder, BoundaryCond, chiMax, dtITP, &
                             maxITPsteps, stepsForJudge, convCriterion1, convCriterion2, totNum
USE TEBDtools_module, ONLY: tensor, vector, matrix, vectorInt
USE Hamiltonian_tools_module, ONLY: Conserv
USE tebd_hip
IMPLICIT NONE
PRIVATE
PUBLIC :: SVDInit, SVDFinish, TwoSiteOp, TrotterStep, CanonicalFormAll, ImagTimeProp, ImagTimePropSpin, &
          ConstructPropagators, TwoSiteOpNC, TrotterStepNC, CanonicalFormAllNC, ImagTimePropNC
PUBLIC :: tebd_hip_resident, tebd_hip_mode, tebd_hip_to_host, tebd_hip_to_device, tebd_hip_context

TYPE(c_ptr), SAVE :: ctx = c_null_ptr
INTEGER, SAVE :: ctx_chi = 0
LOGICAL, SAVE :: on_device = .FALSE.
LOGICAL, SAVE :: tebd_hip_resident = .FALSE.
INTEGER(c_int), SAVE :: tebd_hip_mode = TEBD_MODE_FP64

CONTAINS

FUNCTION tebd_hip_context() RESULT(p)
  TYPE(c_ptr) :: p
  p = ctx
END FUNCTION tebd_hip_context

SUBROUTINE SVDInit(chi)
  INTEGER, INTENT(IN) :: chi
  CHARACTER(kind=c_char) :: bc
  IF (c_associated(ctx)) CALL SVDFinish()
  bc = BoundaryCond
  CALL tebd_check(ctx, tebd_create(ctx, INT(systemSize, c_int), INT(localSize, c_int), INT(chi, c_int), &
                  INT(trotterOrder, c_int), bc, tebd_hip_mode), 'SVDInit')
  ctx_chi = chi
  on_device = .FALSE.
END SUBROUTINE SVDInit

SUBROUTINE SVDFinish()
  IF (c_associated(ctx)) CALL tebd_destroy(ctx)
  ctx = c_null_ptr
  ctx_chi = 0
  on_device = .FALSE.
END SUBROUTINE SVDFinish

SUBROUTINE tebd_hip_to_device(Gammas, Lambdas)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  TYPE(c_ptr), ALLOCATABLE, TARGET :: gp(:), lp(:)
  INTEGER :: i
  IF (.NOT. c_associated(ctx)) STOP 'tebd_hip: SVDInit has not been called'
  ALLOCATE(gp(SIZE(Gammas)), lp(SIZE(Lambdas)))
  DO i = 1, SIZE(Gammas)
    gp(i) = c_loc(Gammas(i)%t(1, 1, 1))
  END DO
  DO i = 1, SIZE(Lambdas)
    lp(i) = c_loc(Lambdas(i)%v(1))
  END DO
  CALL tebd_check(ctx, tebd_upload(ctx, c_loc(gp), c_loc(lp), INT(SIZE(Gammas(1)%t, 1), c_int)), 'upload')
  on_device = .TRUE.
END SUBROUTINE tebd_hip_to_device

SUBROUTINE tebd_hip_to_host(Gammas, Lambdas)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  TYPE(c_ptr), ALLOCATABLE, TARGET :: gp(:), lp(:)
  INTEGER :: i
  ALLOCATE(gp(SIZE(Gammas)), lp(SIZE(Lambdas)))
  DO i = 1, SIZE(Gammas)
    gp(i) = c_loc(Gammas(i)%t(1, 1, 1))
  END DO
  DO i = 1, SIZE(Lambdas)
    lp(i) = c_loc(Lambdas(i)%v(1))
  END DO
  CALL tebd_check(ctx, tebd_download(ctx, c_loc(gp), c_loc(lp), INT(SIZE(Gammas(1)%t, 1), c_int)), 'download')
END SUBROUTINE tebd_hip_to_host

SUBROUTINE before_call(Gammas, Lambdas)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  IF (.NOT. c_associated(ctx)) CALL SVDInit(SIZE(Gammas(1)%t, 1))
  IF (.NOT. (tebd_hip_resident .AND. on_device)) CALL tebd_hip_to_device(Gammas, Lambdas)
END SUBROUTINE before_call

SUBROUTINE after_call(Gammas, Lambdas)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  IF (.NOT. tebd_hip_resident) CALL tebd_hip_to_host(Gammas, Lambdas)
END SUBROUTINE after_call

SUBROUTINE TwoSiteOp(link, Op2, Gammas, Lambdas, truncerr)
  INTEGER, INTENT(IN) :: link
  COMPLEX(KIND=rKind), INTENT(IN), TARGET :: Op2(:, :)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  REAL(KIND=rKind), INTENT(INOUT) :: truncerr
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: op(:, :)
  REAL(c_double) :: t
  CALL before_call(Gammas, Lambdas)
  ALLOCATE(op(SIZE(Op2, 1), SIZE(Op2, 2)))
  op = Op2
  CALL tebd_check(ctx, tebd_two_site_op(ctx, INT(link, c_int), c_loc(op), t), 'TwoSiteOp')
  truncerr = t
  CALL after_call(Gammas, Lambdas)
END SUBROUTINE TwoSiteOp

SUBROUTINE TrotterStep(Udt, Gammas, Lambdas, totalTruncerr)
  TYPE(matrix), POINTER :: Udt(:)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  REAL(KIND=rKind), INTENT(INOUT) :: totalTruncerr
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: u(:, :, :)
  REAL(c_double) :: t
  INTEGER :: i, nb, d2
  CALL before_call(Gammas, Lambdas)
  nb = SIZE(Udt)  ! the propagator table of the schedule (trotterOrder, BoundaryCond)
  d2 = localSize * localSize
  ALLOCATE(u(d2, d2, nb))
  DO i = 1, nb
    u(:, :, i) = Udt(i)%m
  END DO
  CALL tebd_check(ctx, tebd_set_gates(ctx, c_loc(u), INT(nb, c_int)), 'TrotterStep (gates)')
  CALL tebd_check(ctx, tebd_trotter_step(ctx, t), 'TrotterStep')
  totalTruncerr = t
  CALL after_call(Gammas, Lambdas)
END SUBROUTINE TrotterStep

! Number conservation on, with the labels of the call, before a number-conserving update;
! the number of local state i is Conserv%vi(i) with intDegFree, i - 1 without.
SUBROUTINE nc_before(LabelLeft, LabelRight, intDegFree)
  TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
  INTEGER, INTENT(IN), OPTIONAL :: intDegFree
  INTEGER(c_int), ALLOCATABLE, TARGET :: cv(:), ll(:, :), lr(:, :)
  INTEGER :: i, chi
  chi = SIZE(LabelLeft(1)%vi)
  IF (PRESENT(intDegFree)) THEN
    ALLOCATE(cv(localSize))
    cv = Conserv%vi(1:localSize)
    CALL tebd_check(ctx, tebd_set_conserved(ctx, c_loc(cv), INT(totNum, c_int)), 'number conservation')
  ELSE
    CALL tebd_check(ctx, tebd_set_conserved(ctx, c_null_ptr, INT(totNum, c_int)), 'number conservation')
  END IF
  ALLOCATE(ll(chi, systemSize + 1), lr(chi, systemSize + 1))
  DO i = 1, systemSize + 1
    ll(:, i) = LabelLeft(i)%vi
    lr(:, i) = LabelRight(i)%vi
  END DO
  CALL tebd_check(ctx, tebd_set_labels(ctx, c_loc(ll), c_loc(lr), INT(chi, c_int)), 'labels')
END SUBROUTINE nc_before

! The updated labels back into LabelLeft / LabelRight, and number conservation off.
SUBROUTINE nc_after(LabelLeft, LabelRight)
  TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
  INTEGER(c_int), ALLOCATABLE, TARGET :: ll(:, :), lr(:, :)
  INTEGER :: i, chi
  chi = SIZE(LabelLeft(1)%vi)
  ALLOCATE(ll(chi, systemSize + 1), lr(chi, systemSize + 1))
  CALL tebd_check(ctx, tebd_get_labels(ctx, c_loc(ll), c_loc(lr), INT(chi, c_int)), 'labels')
  DO i = 1, systemSize + 1
    LabelLeft(i)%vi = ll(:, i)
    LabelRight(i)%vi = lr(:, i)
  END DO
  CALL tebd_check(ctx, tebd_set_conserved(ctx, c_null_ptr, -1_c_int), 'number conservation')
END SUBROUTINE nc_after

SUBROUTINE TwoSiteOpNC(link, Op2, Gammas, Lambdas, LabelLeft, LabelRight, truncerr, intDegFree)
  INTEGER, INTENT(IN) :: link
  COMPLEX(KIND=rKind), INTENT(IN) :: Op2(:, :)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
  REAL(KIND=rKind), INTENT(INOUT) :: truncerr
  INTEGER, INTENT(IN), OPTIONAL :: intDegFree
  CALL before_call(Gammas, Lambdas)
  CALL nc_before(LabelLeft, LabelRight, intDegFree)
  CALL TwoSiteOp(link, Op2, Gammas, Lambdas, truncerr)
  CALL nc_after(LabelLeft, LabelRight)
END SUBROUTINE TwoSiteOpNC

SUBROUTINE TrotterStepNC(Udt, Gammas, Lambdas, LabelLeft, LabelRight, totalTruncerr, intDegFree)
  TYPE(matrix), POINTER :: Udt(:)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
  REAL(KIND=rKind), INTENT(INOUT) :: totalTruncerr
  INTEGER, INTENT(IN), OPTIONAL :: intDegFree
  CALL before_call(Gammas, Lambdas)
  CALL nc_before(LabelLeft, LabelRight, intDegFree)
  CALL TrotterStep(Udt, Gammas, Lambdas, totalTruncerr)
  CALL nc_after(LabelLeft, LabelRight)
END SUBROUTINE TrotterStepNC

SUBROUTINE CanonicalFormAll(Gammas, Lambdas)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  CALL before_call(Gammas, Lambdas)
  CALL tebd_check(ctx, tebd_canonical_form_all(ctx), 'CanonicalFormAll')
  CALL after_call(Gammas, Lambdas)
END SUBROUTINE CanonicalFormAll

SUBROUTINE CanonicalFormAllNC(Gammas, Lambdas, LabelLeft, LabelRight, intDegFree)
  TYPE(tensor), POINTER :: Gammas(:)
  TYPE(vector), POINTER :: Lambdas(:)
  TYPE(vectorInt), POINTER :: LabelLeft(:), LabelRight(:)
  INTEGER, INTENT(IN), OPTIONAL :: intDegFree
  CALL before_call(Gammas, Lambdas)
  CALL nc_before(LabelLeft, LabelRight, intDegFree)
  CALL CanonicalFormAll(Gammas, Lambdas)
  CALL nc_after(LabelLeft, LabelRight)
END SUBROUTINE CanonicalFormAllNC

! The propagator table of the reference's ConstructPropagators for trotterOrder and
! BoundaryCond; U must be allocated as AllocateProp does (one entry per table slot).
SUBROUTINE ConstructPropagators(H, U, dt)
  TYPE(matrix), POINTER :: H(:), U(:)
  COMPLEX(KIND=rKind), INTENT(IN) :: dt
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: hb(:, :, :), ub(:, :, :)
  INTEGER :: i, nb, nu, d2
  INTEGER(c_int) :: rc
  CHARACTER(kind=c_char) :: bc
  nb = SIZE(H)
  nu = SIZE(U)
  d2 = SIZE(H(1)%m, 1)
  bc = BoundaryCond
  ALLOCATE(hb(d2, d2, nb), ub(d2, d2, nu))
  DO i = 1, nb
    hb(:, :, i) = H(i)%m
  END DO
  rc = tebd_construct_propagators(c_loc(hb), INT(systemSize, c_int), INT(localSize, c_int), &
                                  INT(trotterOrder, c_int), bc, REAL(dt, c_double), AIMAG(dt), c_loc(ub))
  IF (rc /= nu) STOP 'ConstructPropagators: propagator table size does not match trotterOrder / BoundaryCond'
  DO i = 1, nu
    U(i)%m = ub(:, :, i)
  END DO
END SUBROUTINE ConstructPropagators

! Imaginary-time ground-state search with the reference's chi stages (chiIn, then chiMax)
! and its convergence test on every bond's largest lambda, run on the device.  On return
! GammasOuter / LambdasOuter are reallocated at chiMax and hold the result, as the
! reference leaves them.  intDegFree only selects what the reference prints at each
! judgement and is not used.
SUBROUTINE ImagTimeProp(H, GammasOuter, LambdasOuter, chiIn, intDegFree)
  TYPE(matrix), POINTER :: H(:)
  TYPE(tensor), POINTER :: GammasOuter(:)
  TYPE(vector), POINTER :: LambdasOuter(:)
  INTEGER, INTENT(INOUT) :: chiIn
  INTEGER, INTENT(IN), OPTIONAL :: intDegFree
  IF (.NOT. c_associated(ctx) .OR. ctx_chi /= chiMax) CALL SVDInit(chiMax)
  CALL tebd_hip_to_device(GammasOuter, LambdasOuter)
  CALL itp_device(H, chiIn, 'ImagTimeProp')
  CALL realloc_state(GammasOuter, LambdasOuter)
  CALL tebd_hip_to_host(GammasOuter, LambdasOuter)
  IF (PRESENT(intDegFree)) CONTINUE
END SUBROUTINE ImagTimeProp

! ImagTimePropNC: the same stages with TrotterStepNC and CanonicalFormAllNC on the device.
! On return the Gammas, Lambdas and labels are reallocated at chiMax and hold the result,
! as the reference leaves them.  intDegFree selects Conserv as the on-site numbers.
SUBROUTINE ImagTimePropNC(H, GammasOuter, LambdasOuter, LabelLeftOuter, LabelRightOuter, chiIn, intDegFree)
  TYPE(matrix), POINTER :: H(:)
  TYPE(tensor), POINTER :: GammasOuter(:)
  TYPE(vector), POINTER :: LambdasOuter(:)
  TYPE(vectorInt), POINTER :: LabelLeftOuter(:), LabelRightOuter(:)
  INTEGER, INTENT(INOUT) :: chiIn
  INTEGER, INTENT(IN), OPTIONAL :: intDegFree
  INTEGER :: i
  IF (.NOT. c_associated(ctx) .OR. ctx_chi /= chiMax) CALL SVDInit(chiMax)
  CALL tebd_hip_to_device(GammasOuter, LambdasOuter)
  CALL nc_before(LabelLeftOuter, LabelRightOuter, intDegFree)
  CALL itp_device(H, chiIn, 'ImagTimePropNC')
  CALL realloc_state(GammasOuter, LambdasOuter)
  DO i = 1, SIZE(LabelLeftOuter)
    DEALLOCATE(LabelLeftOuter(i)%vi, LabelRightOuter(i)%vi)
    ALLOCATE(LabelLeftOuter(i)%vi(chiMax), LabelRightOuter(i)%vi(chiMax))
  END DO
  CALL tebd_hip_to_host(GammasOuter, LambdasOuter)
  CALL nc_after(LabelLeftOuter, LabelRightOuter)
END SUBROUTINE ImagTimePropNC

! tebd_imag_time_prop with the bond Hamiltonians H and the globals of system_parameters.
SUBROUTINE itp_device(H, chiIn, what)
  TYPE(matrix), POINTER :: H(:)
  INTEGER, INTENT(IN) :: chiIn
  CHARACTER(len=*), INTENT(IN) :: what
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: hb(:, :, :)
  INTEGER(c_int), TARGET :: steps(4)
  INTEGER :: i, nb, d2
  nb = SIZE(H)
  d2 = SIZE(H(1)%m, 1)
  ALLOCATE(hb(d2, d2, nb))
  DO i = 1, nb
    hb(:, :, i) = H(i)%m
  END DO
  CALL tebd_check(ctx, tebd_imag_time_prop(ctx, c_loc(hb), INT(nb, c_int), REAL(dtITP, c_double), &
                  INT(chiIn, c_int), INT(maxITPsteps, c_int), INT(stepsForJudge, c_int), &
                  REAL(convCriterion1, c_double), REAL(convCriterion2, c_double), c_loc(steps)), what)
END SUBROUTINE itp_device

! Gammas and Lambdas reallocated at chiMax (the reference's AllocateGamLam at the last stage).
SUBROUTINE realloc_state(GammasOuter, LambdasOuter)
  TYPE(tensor), POINTER :: GammasOuter(:)
  TYPE(vector), POINTER :: LambdasOuter(:)
  INTEGER :: i
  DO i = 1, SIZE(GammasOuter)
    DEALLOCATE(GammasOuter(i)%t)
    ALLOCATE(GammasOuter(i)%t(chiMax, localSize, chiMax))
  END DO
  DO i = 1, SIZE(LambdasOuter)
    DEALLOCATE(LambdasOuter(i)%v)
    ALLOCATE(LambdasOuter(i)%v(chiMax))
  END DO
END SUBROUTINE realloc_state

! ImagTimePropSpin differs from ImagTimeProp only in what it prints at each judgement.
SUBROUTINE ImagTimePropSpin(H, GammasOuter, LambdasOuter, chiIn)
  TYPE(matrix), POINTER :: H(:)
  TYPE(tensor), POINTER :: GammasOuter(:)
  TYPE(vector), POINTER :: LambdasOuter(:)
  INTEGER, INTENT(INOUT) :: chiIn
  CALL ImagTimeProp(H, GammasOuter, LambdasOuter, chiIn)
END SUBROUTINE ImagTimePropSpin

END MODULE tebd_hip_propagation
