!=======================================================================
! ref_bind.f90 -- C-callable entry points into the REFERENCE Fortran
! (common_letkf / common_mtx built from /root/reference by oracle/Makefile).
!
! TEST INFRASTRUCTURE: the oracle's "reference" leg.  Used by the parity
! tests, the golden-vector generator and the CPU-baseline leg of bench.py.
! Never linked into the product.
!
! ref_letkf_core   -> common/common_letkf.f90:50 (letkf_core), with the
!                     OPTIONAL arguments selected by presence flags
! ref_mtx_eigen    -> common/common_mtx.f90:41 (mtx_eigen, imode=1)
! ref_solve_apply  -> the per-point analysis loop of the drivers
!                     (speedy/letkf/letkf_tools.f90:150-170,
!                     scale/letkf/letkf_tools.f90:175 OpenMP pattern):
!                     gather the local rows, call the reference
!                     letkf_core, apply Xa = xmean + Xb' * trans.
! ref_l96_rk4      -> lorenz96/model/lorenz96.f90:93-128 (tinteg_rk4, the
!                     reference model with its nx = 40), dt and force set
! ref_com_gamma    -> common/common.f90:861-912 (com_gamma)
! ref_com_distll_1 -> common/common.f90:401-424 (com_distll_1)
!=======================================================================
MODULE ref_bind
  USE ISO_C_BINDING
  USE common, ONLY: r_size, com_gamma, com_distll_1
  USE common_mtx, ONLY: mtx_eigen
  USE common_letkf, ONLY: letkf_core
  USE lorenz96, ONLY: tinteg_rk4, l96_nx => nx, l96_dt => dt, l96_force => force
  IMPLICIT NONE
CONTAINS

SUBROUTINE ref_l96_rk4(kt, dt, force, xin, xout) BIND(C, NAME="ref_l96_rk4")
  INTEGER(C_INT), VALUE :: kt
  REAL(C_DOUBLE), VALUE :: dt, force
  REAL(C_DOUBLE), INTENT(IN) :: xin(l96_nx)
  REAL(C_DOUBLE), INTENT(OUT) :: xout(l96_nx)
  l96_dt = dt
  l96_force = force
  CALL tinteg_rk4(kt, xin, xout)
END SUBROUTINE ref_l96_rk4

REAL(C_DOUBLE) FUNCTION ref_com_gamma(x) BIND(C, NAME="ref_com_gamma")
  REAL(C_DOUBLE), VALUE :: x
  REAL(r_size) :: xx, ga
  xx = x
  CALL com_gamma(xx, ga)
  ref_com_gamma = ga
END FUNCTION ref_com_gamma

REAL(C_DOUBLE) FUNCTION ref_com_distll_1(alon, alat, blon, blat) BIND(C, NAME="ref_com_distll_1")
  REAL(C_DOUBLE), VALUE :: alon, alat, blon, blat
  REAL(r_size) :: d
  CALL com_distll_1(alon, alat, blon, blat, d)
  ref_com_distll_1 = d
END FUNCTION ref_com_distll_1

INTEGER(C_INT) FUNCTION ref_l96_nx() BIND(C, NAME="ref_l96_nx")
  ref_l96_nx = l96_nx
END FUNCTION ref_l96_nx

SUBROUTINE ref_letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, parm, trans, &
     & transm, has_transm, pao, has_pao, rdiag_wloc, minfl, has_minfl) BIND(C, NAME="ref_letkf_core")
  INTEGER(C_INT), VALUE :: ne, nobs, nobsl, has_transm, has_pao, rdiag_wloc, has_minfl
  REAL(C_DOUBLE), INTENT(IN) :: hdxb(nobs, ne), rdiag(nobs), rloc(nobs), dep(nobs)
  REAL(C_DOUBLE), INTENT(INOUT) :: parm
  REAL(C_DOUBLE), INTENT(OUT) :: trans(ne, ne), transm(ne), pao(ne, ne)
  REAL(C_DOUBLE), INTENT(IN) :: minfl
  LOGICAL :: wl
  REAL(r_size) :: p
  wl = (rdiag_wloc /= 0)
  p = parm
  IF (has_transm /= 0 .AND. has_pao /= 0) THEN
    IF (has_minfl /= 0) THEN
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, transm=transm, pao=pao, rdiag_wloc=wl, minfl=minfl)
    ELSE
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, transm=transm, pao=pao, rdiag_wloc=wl)
    END IF
  ELSE IF (has_transm /= 0) THEN
    IF (has_minfl /= 0) THEN
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, transm=transm, rdiag_wloc=wl, minfl=minfl)
    ELSE
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, transm=transm, rdiag_wloc=wl)
    END IF
  ELSE IF (has_pao /= 0) THEN
    IF (has_minfl /= 0) THEN
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, pao=pao, rdiag_wloc=wl, minfl=minfl)
    ELSE
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, pao=pao, rdiag_wloc=wl)
    END IF
  ELSE
    IF (has_minfl /= 0) THEN
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, rdiag_wloc=wl, minfl=minfl)
    ELSE
      CALL letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, p, trans, rdiag_wloc=wl)
    END IF
  END IF
  parm = p
END SUBROUTINE ref_letkf_core

SUBROUTINE ref_mtx_eigen(n, a, eival, eivec, nrank) BIND(C, NAME="ref_mtx_eigen")
  INTEGER(C_INT), VALUE :: n
  REAL(C_DOUBLE), INTENT(IN) :: a(n, n)
  REAL(C_DOUBLE), INTENT(OUT) :: eival(n), eivec(n, n)
  INTEGER(C_INT), INTENT(OUT) :: nrank
  INTEGER :: nr
  CALL mtx_eigen(1, n, a, eival, eivec, nr)
  nrank = nr
END SUBROUTINE ref_mtx_eigen

!-----------------------------------------------------------------------
! Batched reference analysis over packed local problems.
!   npt points; point b uses local obs loc_idx(loc_off(b)+1 : loc_off(b+1))
!   (0-based indices into the obs table) with localisation loc_rho.
!   hdxf(nobs, ne), rdiag(nobs), dep(nobs): obs table (column-major)
!   xb(ne, nrow): per point nv rows of the background ensemble, rows of
!   point b are row_of(b)*nv .. +nv-1 (0-based), each row = ne members.
!   xa(ne, nrow) out.  infl(npt) inout.  fixed_infl > 0 resets the
!   inflation after every call as the L96 driver does (letkf.f90:256).
!   Returns the number of points whose letkf_core call was made.
!-----------------------------------------------------------------------
INTEGER(C_INT) FUNCTION ref_solve_apply(npt, ne, nobs, maxp, loc_off, loc_idx, loc_rho, &
     & hdxf, rdiag, dep, nv, xb, xa, infl, fixed_infl) BIND(C, NAME="ref_solve_apply")
  INTEGER(C_INT), VALUE :: npt, ne, nobs, maxp, nv
  INTEGER(C_LONG), INTENT(IN) :: loc_off(npt + 1)
  INTEGER(C_INT), INTENT(IN) :: loc_idx(*)
  REAL(C_DOUBLE), INTENT(IN) :: loc_rho(*)
  REAL(C_DOUBLE), INTENT(IN) :: hdxf(nobs, ne), rdiag(nobs), dep(nobs)
  REAL(C_DOUBLE), INTENT(IN) :: xb(ne, nv * npt)
  REAL(C_DOUBLE), INTENT(OUT) :: xa(ne, nv * npt)
  REAL(C_DOUBLE), INTENT(INOUT) :: infl(npt)
  REAL(C_DOUBLE), VALUE :: fixed_infl
  REAL(r_size), ALLOCATABLE :: hl(:, :), rl(:), rd(:), dl(:)
  REAL(r_size) :: trans(ne, ne), parm, xm
  INTEGER :: b, q, p, o, r, m, kk, ld
  INTEGER(C_LONG) :: base
  ld = MAX(maxp, 1)
  !$OMP PARALLEL PRIVATE(hl, rl, rd, dl, trans, parm, xm, b, q, p, o, r, m, kk, base)
  ALLOCATE(hl(ld, ne), rl(ld), rd(ld), dl(ld))
  !$OMP DO SCHEDULE(DYNAMIC)
  DO b = 1, npt
    base = loc_off(b)
    p = INT(loc_off(b + 1) - base)
    DO q = 1, p
      o = loc_idx(base + q) + 1
      hl(q, :) = hdxf(o, :)
      rl(q) = loc_rho(base + q)
      rd(q) = rdiag(o)
      dl(q) = dep(o)
    END DO
    parm = infl(b)
    CALL letkf_core(ne, ld, p, hl, rd, rl, dl, parm, trans)
    IF (fixed_infl > 0.0d0) parm = fixed_infl
    infl(b) = parm
    DO r = 1, nv
      xm = xb(1, (b - 1) * nv + r)
      DO m = 2, ne
        xm = xm + xb(m, (b - 1) * nv + r)
      END DO
      xm = xm / REAL(ne, r_size)
      DO m = 1, ne
        xa(m, (b - 1) * nv + r) = xm
        DO kk = 1, ne
          xa(m, (b - 1) * nv + r) = xa(m, (b - 1) * nv + r) + (xb(kk, (b - 1) * nv + r) - xm) * trans(kk, m)
        END DO
      END DO
    END DO
  END DO
  !$OMP END DO
  DEALLOCATE(hl, rl, rd, dl)
  !$OMP END PARALLEL
  ref_solve_apply = npt
END FUNCTION ref_solve_apply

INTEGER(C_INT) FUNCTION ref_max_threads() BIND(C, NAME="ref_max_threads")
  !$ USE omp_lib
  ref_max_threads = 1
  !$ ref_max_threads = omp_get_max_threads()
END FUNCTION ref_max_threads

END MODULE ref_bind
