!=======================================================================
! common_letkf -- drop-in replacement of the reference module of the same
! name (common/common_letkf.f90).  Same module name, same public
! parameter nbv, same letkf_core dummy arguments (common_letkf.f90:50-64);
! the local solve runs on the GPU through the C ABI of include/lekf.h
! (lekf_core_f64).  Absent OPTIONAL arguments become NULL pointers,
! LOGICAL becomes a C int, and a nonzero status stops the program with the
! reference's code (2: eigensolver failure as common_mtx.f90:61-64,75-78;
! 99: inconsistent sizes).
!=======================================================================
MODULE common_letkf
  USE ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: nbv, letkf_core

  ! ensemble size of the reference drivers (common/common_letkf.f90:28)
  INTEGER, PARAMETER :: nbv = 20

  INTERFACE
    INTEGER(C_INT) FUNCTION lekf_core_f64(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, parm_infl, &
         & trans, transm, pao, rdiag_wloc, minfl) BIND(C, NAME="lekf_core_f64")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      INTEGER(C_INT), VALUE :: ne, nobs, nobsl
      REAL(C_DOUBLE), INTENT(IN) :: hdxb(*), rdiag(*), rloc(*), dep(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: parm_infl
      REAL(C_DOUBLE), INTENT(OUT) :: trans(*)
      TYPE(C_PTR), VALUE :: transm, pao
      INTEGER(C_INT), VALUE :: rdiag_wloc
      TYPE(C_PTR), VALUE :: minfl
    END FUNCTION lekf_core_f64
  END INTERFACE

CONTAINS

SUBROUTINE letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, parm_infl, trans, transm, pao, &
     & rdiag_wloc, minfl)
  INTEGER, INTENT(IN) :: ne
  INTEGER, INTENT(IN) :: nobs
  INTEGER, INTENT(IN) :: nobsl
  REAL(C_DOUBLE), INTENT(IN) :: hdxb(1:nobs, 1:ne)
  REAL(C_DOUBLE), INTENT(IN) :: rdiag(1:nobs)
  REAL(C_DOUBLE), INTENT(IN) :: rloc(1:nobs)
  REAL(C_DOUBLE), INTENT(IN) :: dep(1:nobs)
  REAL(C_DOUBLE), INTENT(INOUT) :: parm_infl
  REAL(C_DOUBLE), INTENT(OUT) :: trans(ne, ne)
  REAL(C_DOUBLE), INTENT(OUT), OPTIONAL, TARGET :: transm(ne)
  REAL(C_DOUBLE), INTENT(OUT), OPTIONAL, TARGET :: pao(ne, ne)
  LOGICAL, INTENT(IN), OPTIONAL :: rdiag_wloc
  REAL(C_DOUBLE), INTENT(IN), OPTIONAL, TARGET :: minfl

  TYPE(C_PTR) :: p_transm, p_pao, p_minfl
  INTEGER(C_INT) :: wl, st
  REAL(C_DOUBLE) :: dummy(1)

  p_transm = C_NULL_PTR
  p_pao = C_NULL_PTR
  p_minfl = C_NULL_PTR
  IF (PRESENT(transm)) p_transm = C_LOC(transm)
  IF (PRESENT(pao)) p_pao = C_LOC(pao)
  IF (PRESENT(minfl)) p_minfl = C_LOC(minfl)
  wl = 0
  IF (PRESENT(rdiag_wloc)) THEN
    IF (rdiag_wloc) wl = 1
  END IF
  dummy = 0.0d0
  IF (nobs > 0) THEN
    st = lekf_core_f64(INT(ne, C_INT), INT(nobs, C_INT), INT(nobsl, C_INT), hdxb, rdiag, rloc, &
         & dep, parm_infl, trans, p_transm, p_pao, wl, p_minfl)
  ELSE
    st = lekf_core_f64(INT(ne, C_INT), 1_C_INT, 0_C_INT, dummy, dummy, dummy, dummy, parm_infl, &
         & trans, p_transm, p_pao, wl, p_minfl)
  END IF
  SELECT CASE (st)
  CASE (0, 4)
    RETURN
  CASE (2)
    WRITE(6, '(A)') '!!! ERROR (letkf_core): eigensolver failure on the GPU'
    STOP 2
  CASE (99)
    WRITE(6, '(A,3I8)') '!!! ERROR (letkf_core): inconsistent sizes ne,nobs,nobsl=', ne, nobs, nobsl
    STOP 99
  CASE DEFAULT
    WRITE(6, '(A,I6)') '!!! ERROR (letkf_core): GPU library status ', st
    STOP 1
  END SELECT
END SUBROUTINE letkf_core

END MODULE common_letkf
