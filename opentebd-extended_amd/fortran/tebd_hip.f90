! iso_c_binding interface to libtebdhip.so (C declarations: include/tebd_hip.h).
!
! Module tebd_hip is self-contained: plain C pointers, integers and reals.
! Module tebd_hip_propagation (tebd_hip_propagation.f90) sits on top of it and
! offers the reference's own call surface on the reference's derived types.
MODULE tebd_hip
USE, INTRINSIC :: iso_c_binding
IMPLICIT NONE

INTEGER(c_int), PARAMETER :: TEBD_OK = 0
INTEGER(c_int), PARAMETER :: TEBD_MODE_FP64 = 0
INTEGER(c_int), PARAMETER :: TEBD_MODE_MIXED = 1
INTEGER(c_int), PARAMETER :: TEBD_MODE_BATCHED = 2
INTEGER(c_int), PARAMETER :: TEBD_MODE_LEARNED_RANK = 4
INTEGER(c_int), PARAMETER :: TEBD_MODE_SPLIT = 8

INTERFACE
  INTEGER(c_int) FUNCTION tebd_create(ctx, L, d, chi_max, order, bc, mode) BIND(C, name="tebd_create")
    IMPORT :: c_ptr, c_int, c_char
    TYPE(c_ptr) :: ctx
    INTEGER(c_int), VALUE :: L, d, chi_max, order
    CHARACTER(kind=c_char), VALUE :: bc
    INTEGER(c_int), VALUE :: mode
  END FUNCTION tebd_create

  ! nchain independent chains of L sites in one context (disorder realisations)
  INTEGER(c_int) FUNCTION tebd_create_ensemble(ctx, nchain, L, d, chi_max, order, bc, mode) &
      BIND(C, name="tebd_create_ensemble")
    IMPORT :: c_ptr, c_int, c_char
    TYPE(c_ptr) :: ctx
    INTEGER(c_int), VALUE :: nchain, L, d, chi_max, order
    CHARACTER(kind=c_char), VALUE :: bc
    INTEGER(c_int), VALUE :: mode
  END FUNCTION tebd_create_ensemble

  SUBROUTINE tebd_destroy(ctx) BIND(C, name="tebd_destroy")
    IMPORT :: c_ptr
    TYPE(c_ptr), VALUE :: ctx
  END SUBROUTINE tebd_destroy

  INTEGER(c_int) FUNCTION tebd_set_gates(ctx, U, nb) BIND(C, name="tebd_set_gates")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, U
    INTEGER(c_int), VALUE :: nb
  END FUNCTION tebd_set_gates

  INTEGER(c_int) FUNCTION tebd_upload(ctx, gamma, lambda, chi_in) BIND(C, name="tebd_upload")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, gamma, lambda
    INTEGER(c_int), VALUE :: chi_in
  END FUNCTION tebd_upload

  INTEGER(c_int) FUNCTION tebd_download(ctx, gamma, lambda, chi_out) BIND(C, name="tebd_download")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, gamma, lambda
    INTEGER(c_int), VALUE :: chi_out
  END FUNCTION tebd_download

  INTEGER(c_int) FUNCTION tebd_trotter_step(ctx, truncerr) BIND(C, name="tebd_trotter_step")
    IMPORT :: c_ptr, c_int, c_double
    TYPE(c_ptr), VALUE :: ctx
    REAL(c_double) :: truncerr
  END FUNCTION tebd_trotter_step

  INTEGER(c_int) FUNCTION tebd_two_site_op(ctx, link, op2, truncerr) BIND(C, name="tebd_two_site_op")
    IMPORT :: c_ptr, c_int, c_double
    TYPE(c_ptr), VALUE :: ctx, op2
    INTEGER(c_int), VALUE :: link
    REAL(c_double) :: truncerr
  END FUNCTION tebd_two_site_op

  INTEGER(c_int) FUNCTION tebd_construct_propagators(H, L, d, order, bc, dt_re, dt_im, U) &
      BIND(C, name="tebd_construct_propagators")
    IMPORT :: c_ptr, c_int, c_double, c_char
    TYPE(c_ptr), VALUE :: H, U
    INTEGER(c_int), VALUE :: L, d, order
    CHARACTER(kind=c_char), VALUE :: bc
    REAL(c_double), VALUE :: dt_re, dt_im
  END FUNCTION tebd_construct_propagators

  INTEGER(c_int) FUNCTION tebd_imag_time_prop(ctx, H, nh, dtau, chi_in, max_steps, steps_for_judge, conv1, conv2, &
                                              steps) BIND(C, name="tebd_imag_time_prop")
    IMPORT :: c_ptr, c_int, c_double
    TYPE(c_ptr), VALUE :: ctx, H, steps
    INTEGER(c_int), VALUE :: nh, chi_in, max_steps, steps_for_judge
    REAL(c_double), VALUE :: dtau, conv1, conv2
  END FUNCTION tebd_imag_time_prop

  INTEGER(c_int) FUNCTION tebd_set_conserved(ctx, conserv, tot_num) BIND(C, name="tebd_set_conserved")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, conserv
    INTEGER(c_int), VALUE :: tot_num
  END FUNCTION tebd_set_conserved

  INTEGER(c_int) FUNCTION tebd_set_labels(ctx, label_left, label_right, chi) BIND(C, name="tebd_set_labels")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, label_left, label_right
    INTEGER(c_int), VALUE :: chi
  END FUNCTION tebd_set_labels

  INTEGER(c_int) FUNCTION tebd_get_labels(ctx, label_left, label_right, chi) BIND(C, name="tebd_get_labels")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, label_left, label_right
    INTEGER(c_int), VALUE :: chi
  END FUNCTION tebd_get_labels

  INTEGER(c_int) FUNCTION tebd_canonical_form_all(ctx) BIND(C, name="tebd_canonical_form_all")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx
  END FUNCTION tebd_canonical_form_all

  INTEGER(c_int) FUNCTION tebd_set_cutoff(ctx, cutoff) BIND(C, name="tebd_set_cutoff")
    IMPORT :: c_ptr, c_int, c_double
    TYPE(c_ptr), VALUE :: ctx
    REAL(c_double), VALUE :: cutoff
  END FUNCTION tebd_set_cutoff

  INTEGER(c_int) FUNCTION tebd_get_bond_dims(ctx, chi) BIND(C, name="tebd_get_bond_dims")
    IMPORT :: c_ptr, c_int
    TYPE(c_ptr), VALUE :: ctx, chi
  END FUNCTION tebd_get_bond_dims

  TYPE(c_ptr) FUNCTION tebd_last_error(ctx) BIND(C, name="tebd_last_error")
    IMPORT :: c_ptr
    TYPE(c_ptr), VALUE :: ctx
  END FUNCTION tebd_last_error

  INTEGER(c_size_t) FUNCTION c_strlen(s) BIND(C, name="strlen")
    IMPORT :: c_ptr, c_size_t
    TYPE(c_ptr), VALUE :: s
  END FUNCTION c_strlen
END INTERFACE

CONTAINS

! Message of the context's last failed call, as a Fortran string.
FUNCTION tebd_error_message(ctx) RESULT(msg)
  TYPE(c_ptr), INTENT(IN) :: ctx
  CHARACTER(len=:), ALLOCATABLE :: msg
  TYPE(c_ptr) :: p
  CHARACTER(kind=c_char), POINTER :: chars(:)
  INTEGER :: n, i
  p = tebd_last_error(ctx)
  n = INT(c_strlen(p))
  CALL c_f_pointer(p, chars, [n])
  ALLOCATE(CHARACTER(len=n) :: msg)
  DO i = 1, n
    msg(i:i) = chars(i)
  END DO
END FUNCTION tebd_error_message

! Reference-compatible fatal behaviour: a failed call stops the program.
SUBROUTINE tebd_check(ctx, rc, what)
  TYPE(c_ptr), INTENT(IN) :: ctx
  INTEGER(c_int), INTENT(IN) :: rc
  CHARACTER(len=*), INTENT(IN) :: what
  IF (rc /= TEBD_OK) THEN
    IF (c_associated(ctx)) THEN
      PRINT *, 'tebd_hip: ', what, ' failed (', rc, '): ', tebd_error_message(ctx)
    ELSE
      PRINT *, 'tebd_hip: ', what, ' failed (', rc, ')'
    END IF
    STOP 'tebd_hip: fatal error'
  END IF
END SUBROUTINE tebd_check

END MODULE tebd_hip
