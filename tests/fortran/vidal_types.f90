! Test harness only: the two reference modules the iso_c_binding shim
! (opentebd-extended_amd/fortran/tebd_hip_propagation.f90) binds to, declared with the
! reference's names, kinds and components so the shim compiles unchanged without the
! reference's sources:
!   system_parameters: rKind, systemSize, localSize, trotterOrder, BoundaryCond, the
!                      imaginary-time settings chiMax, dtITP, maxITPsteps, stepsForJudge,
!                      convCriterion1/2, and totNum (Case_Studies/*/system_parameters.f90:16-58)
!   TEBDtools_module:  TYPE vector / matrix / tensor / vectorInt
!                      (VidalCore_v_2.0/TEBDtools_module.f90:31-49)
!   Hamiltonian_tools_module: the on-site numbers Conserv (HamiltonianCore_v_2.0/
!                      Hamiltonian_tools_module.f90:57)
! Only these declarations; none of the reference's procedures.  Host code that uses the
! shim in earnest compiles it against the reference's own modules instead.
MODULE system_parameters
IMPLICIT NONE
INTEGER, PARAMETER :: rKind = SELECTED_REAL_KIND(p=15, r=30)
! (values are set by the harness from each job)
INTEGER :: systemSize = 0
INTEGER :: localSize = 0
INTEGER :: trotterOrder = 2
CHARACTER :: BoundaryCond = 'O'
INTEGER :: chiMax = 0
REAL(KIND=rKind) :: dtITP = 0.0_rKind
INTEGER :: stepsForJudge = 1
INTEGER :: maxITPsteps = 0
REAL(KIND=rKind) :: convCriterion1 = 0.0_rKind
REAL(KIND=rKind) :: convCriterion2 = 0.0_rKind
INTEGER :: totNum = 0
END MODULE system_parameters

MODULE TEBDtools_module
USE system_parameters, ONLY: rKind
IMPLICIT NONE
TYPE vector
  REAL(KIND=rKind), POINTER :: v(:)
END TYPE vector
TYPE matrix
  COMPLEX(KIND=rKind), POINTER :: m(:, :)
END TYPE matrix
TYPE tensor
  COMPLEX(KIND=rKind), POINTER :: t(:, :, :)
END TYPE tensor
TYPE vectorInt
  INTEGER, POINTER :: vi(:)
END TYPE vectorInt
END MODULE TEBDtools_module

MODULE Hamiltonian_tools_module
USE TEBDtools_module, ONLY: vectorInt
IMPLICIT NONE
TYPE(vectorInt) :: Conserv
END MODULE Hamiltonian_tools_module
