!=======================================================================
! omp_core_driver -- letkf_core called from OpenMP threads, the way the
! reference's analysis loops call it (scale/letkf/letkf_tools.f90:175 and
! speedy/letkf/letkf_tools.f90:133, !$OMP PARALLEL DO over grid points),
! through the drop-in common_letkf.
!
!   omp_core_driver <in> <out>
!
! <in>  (stream, native): ncase, ne, nobs (int32); then per case nobsl
!       (int32), hdxb(nobs, ne) column-major, rdiag(nobs), rloc(nobs),
!       dep(nobs), parm_infl (float64; rows past nobsl unused)
! <out> per case: trans(ne, ne), transm(ne), pao(ne, ne), parm_infl
! The cases are solved in one PARALLEL DO over the cases (dynamic schedule,
! OMP_NUM_THREADS threads).
!=======================================================================
PROGRAM omp_core_driver
  USE ISO_C_BINDING
  USE common_letkf, ONLY: letkf_core
  IMPLICIT NONE
  CHARACTER(LEN=512) :: fin, fout
  INTEGER(C_INT) :: hdr(3)
  INTEGER :: ncase, ne, nobs, c
  INTEGER(C_INT), ALLOCATABLE :: nl(:)
  REAL(C_DOUBLE), ALLOCATABLE :: hdxb(:, :, :), rdiag(:, :), rloc(:, :), dep(:, :), parm(:)
  REAL(C_DOUBLE), ALLOCATABLE :: trans(:, :, :), transm(:, :), pao(:, :, :)

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(10, FILE=TRIM(fin), FORM='UNFORMATTED', ACCESS='STREAM', STATUS='OLD')
  READ(10) hdr
  ncase = hdr(1)
  ne = hdr(2)
  nobs = hdr(3)
  ALLOCATE(nl(ncase), hdxb(nobs, ne, ncase), rdiag(nobs, ncase), rloc(nobs, ncase))
  ALLOCATE(dep(nobs, ncase), parm(ncase))
  ALLOCATE(trans(ne, ne, ncase), transm(ne, ncase), pao(ne, ne, ncase))
  DO c = 1, ncase
    READ(10) nl(c)
    READ(10) hdxb(:, :, c), rdiag(:, c), rloc(:, c), dep(:, c), parm(c)
  END DO
  CLOSE(10)

!$OMP PARALLEL DO SCHEDULE(DYNAMIC)
  DO c = 1, ncase
    CALL letkf_core(ne, nobs, INT(nl(c)), hdxb(:, :, c), rdiag(:, c), rloc(:, c), dep(:, c), &
         & parm(c), trans(:, :, c), transm(:, c), pao(:, :, c))
  END DO
!$OMP END PARALLEL DO

  OPEN(11, FILE=TRIM(fout), FORM='UNFORMATTED', ACCESS='STREAM', STATUS='REPLACE')
  DO c = 1, ncase
    WRITE(11) trans(:, :, c), transm(:, c), pao(:, :, c), parm(c)
  END DO
  CLOSE(11)
END PROGRAM omp_core_driver
