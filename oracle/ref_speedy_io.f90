!=======================================================================
! ref_speedy_io.f90 -- C-callable entry points into the REFERENCE SPEEDY
! file routines (common_speedy / common_obs_speedy built from
! /root/reference by oracle/Makefile).
!
! TEST INFRASTRUCTURE: pins the package's SPEEDY file reader/writer
! against the reference's own.  Never linked into the product.
!
! ref_spd_grid      -> nlon, nlat, nlev, nv3d, nv2d of
!                      speedy/common/common_speedy.f90:18-22
! ref_spd_read_grd  -> read_grd   (speedy/common/common_speedy.f90:163-194)
! ref_spd_write_grd -> write_grd  (speedy/common/common_speedy.f90:226-257)
! ref_spd_get_nobs  -> get_nobs   (speedy/common/common_obs_speedy.f90:400-456)
! ref_spd_read_obs  -> read_obs   (speedy/common/common_obs_speedy.f90:458-502)
! ref_spd_read_obs2 -> read_obs2  (speedy/common/common_obs_speedy.f90:504-552)
! ref_spd_write_obs2-> write_obs2 (speedy/common/common_obs_speedy.f90:554-602)
! ref_spd_obsope1   -> phys2ijk, the range rules of obsope
!                      (speedy/letkf/obsope.f90:56-89) and Trans_XtoY
!                      (speedy/common/common_obs_speedy.f90:30-282) for one
!                      obs on one member's fields (T30 grid)
!
! Files are native-endian here: the byte order of the reference build
! (-byteswapio) is a 4-byte-word swap of the whole file, checked apart.
!=======================================================================
MODULE ref_speedy_io
  USE ISO_C_BINDING
  USE common, ONLY: r_size
  USE common_speedy, ONLY: nlon, nlat, nlev, nv3d, nv2d, read_grd, write_grd, &
       & set_common_speedy, calc_pfull, iv2d_ps
  USE common_obs_speedy, ONLY: get_nobs, read_obs, read_obs2, write_obs2, phys2ijk, &
       & Trans_XtoY, id_u_obs, id_v_obs, id_ps_obs
  IMPLICIT NONE
CONTAINS

FUNCTION fname_of(c, n) RESULT(f)
  INTEGER(C_INT), INTENT(IN) :: n
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  CHARACTER(LEN=n) :: f
  INTEGER :: i
  DO i = 1, n
    f(i:i) = c(i)
  END DO
END FUNCTION fname_of

SUBROUTINE ref_spd_grid(dims) BIND(C, NAME="ref_spd_grid")
  INTEGER(C_INT), INTENT(OUT) :: dims(5)
  dims = (/ nlon, nlat, nlev, nv3d, nv2d /)
END SUBROUTINE ref_spd_grid

SUBROUTINE ref_spd_read_grd(c, n, v3d, v2d) BIND(C, NAME="ref_spd_read_grd")
  INTEGER(C_INT), VALUE :: n
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  REAL(C_DOUBLE), INTENT(OUT) :: v3d(nlon, nlat, nlev, nv3d), v2d(nlon, nlat, nv2d)
  CALL read_grd(fname_of(c, n), v3d, v2d)
END SUBROUTINE ref_spd_read_grd

SUBROUTINE ref_spd_write_grd(c, n, v3d, v2d) BIND(C, NAME="ref_spd_write_grd")
  INTEGER(C_INT), VALUE :: n
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  REAL(C_DOUBLE), INTENT(IN) :: v3d(nlon, nlat, nlev, nv3d), v2d(nlon, nlat, nv2d)
  CALL write_grd(fname_of(c, n), v3d, v2d)
END SUBROUTINE ref_spd_write_grd

INTEGER(C_INT) FUNCTION ref_spd_get_nobs(c, n, nrec) BIND(C, NAME="ref_spd_get_nobs")
  INTEGER(C_INT), VALUE :: n, nrec
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  INTEGER :: nn
  CALL get_nobs(fname_of(c, n), nrec, nn)
  ref_spd_get_nobs = nn
END FUNCTION ref_spd_get_nobs

SUBROUTINE ref_spd_read_obs(c, n, nn, elem, rlon, rlat, rlev, odat, oerr) &
    BIND(C, NAME="ref_spd_read_obs")
  INTEGER(C_INT), VALUE :: n, nn
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  REAL(C_DOUBLE), INTENT(OUT) :: elem(nn), rlon(nn), rlat(nn), rlev(nn), odat(nn), oerr(nn)
  CALL read_obs(fname_of(c, n), nn, elem, rlon, rlat, rlev, odat, oerr)
END SUBROUTINE ref_spd_read_obs

SUBROUTINE ref_spd_read_obs2(c, n, nn, elem, rlon, rlat, rlev, odat, oerr, ohx, oqc) &
    BIND(C, NAME="ref_spd_read_obs2")
  INTEGER(C_INT), VALUE :: n, nn
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  REAL(C_DOUBLE), INTENT(OUT) :: elem(nn), rlon(nn), rlat(nn), rlev(nn), odat(nn), oerr(nn)
  REAL(C_DOUBLE), INTENT(OUT) :: ohx(nn)
  INTEGER(C_INT), INTENT(OUT) :: oqc(nn)
  INTEGER :: q(nn)
  CALL read_obs2(fname_of(c, n), nn, elem, rlon, rlat, rlev, odat, oerr, ohx, q)
  oqc = q
END SUBROUTINE ref_spd_read_obs2

SUBROUTINE ref_spd_write_obs2(c, n, nn, elem, rlon, rlat, rlev, odat, oerr, ohx, oqc) &
    BIND(C, NAME="ref_spd_write_obs2")
  INTEGER(C_INT), VALUE :: n, nn
  CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(n)
  REAL(C_DOUBLE), INTENT(IN) :: elem(nn), rlon(nn), rlat(nn), rlev(nn), odat(nn), oerr(nn)
  REAL(C_DOUBLE), INTENT(IN) :: ohx(nn)
  INTEGER(C_INT), INTENT(IN) :: oqc(nn)
  INTEGER :: q(nn)
  q = oqc
  CALL write_obs2(fname_of(c, n), nn, elem, rlon, rlat, rlev, odat, oerr, ohx, q)
END SUBROUTINE ref_spd_write_obs2

! status: 0 the obs is used (yobs set), 1 x out of range, 2 y out of range,
! 3 z out of range, 4 a ps obs below -100 (the reference's CYCLE branches)
SUBROUTINE ref_spd_obsope1(v3d, v2d, elem, rlon, rlat, rlev, odat, ri, rj, rk, yobs, status) &
    BIND(C, NAME="ref_spd_obsope1")
  REAL(C_DOUBLE), INTENT(IN) :: v3d(nlon, nlat, nlev, nv3d), v2d(nlon, nlat, nv2d)
  REAL(C_DOUBLE), VALUE :: elem, rlon, rlat, rlev, odat
  REAL(C_DOUBLE), INTENT(OUT) :: ri, rj, rk, yobs
  INTEGER(C_INT), INTENT(OUT) :: status
  REAL(r_size) :: p_full(nlon, nlat, nlev)
  REAL(r_size) :: phi0z(nlon, nlat)
  LOGICAL, SAVE :: ready = .FALSE.
  IF (.NOT. ready) THEN
    ! set_common_speedy reads the orography from unit 21 (only its own phi0
    ! uses it): a scratch record of zeros
    phi0z = 0.0d0
    OPEN(21, STATUS='SCRATCH', FORM='UNFORMATTED', ACCESS='SEQUENTIAL')
    WRITE(21) phi0z
    REWIND(21)
    CALL set_common_speedy
    CLOSE(21)
    ready = .TRUE.
  END IF
  CALL calc_pfull(nlon, nlat, v2d(:, :, iv2d_ps), p_full)
  ri = 0.0d0
  rj = 0.0d0
  rk = 0.0d0
  yobs = 0.0d0
  status = 0
  CALL phys2ijk(p_full, elem, rlon, rlat, rlev, ri, rj, rk)
  IF (CEILING(ri) < 2 .OR. nlon + 1 < CEILING(ri)) THEN
    status = 1
    RETURN
  END IF
  IF (CEILING(rj) < 2 .OR. nlat < CEILING(rj)) THEN
    status = 2
    RETURN
  END IF
  IF (CEILING(rk) > nlev) THEN
    status = 3
    RETURN
  END IF
  IF (CEILING(rk) < 2 .AND. NINT(elem) /= id_ps_obs) THEN
    IF (NINT(elem) == id_u_obs .OR. NINT(elem) == id_v_obs) THEN
      rk = 1.00001d0
    ELSE
      status = 3
      RETURN
    END IF
  END IF
  IF (NINT(elem) == id_ps_obs .AND. odat < -100.0d0) THEN
    status = 4
    RETURN
  END IF
  CALL Trans_XtoY(elem, ri, rj, rk, v3d, v2d, p_full, yobs)
END SUBROUTINE ref_spd_obsope1

END MODULE ref_speedy_io
