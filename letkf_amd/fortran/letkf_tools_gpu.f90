!=======================================================================
! letkf_tools_gpu -- das_letkf-level Fortran host API over the C ABI of
! include/lekf.h.  The ensemble stays resident on the GPU between calls
! (one handle per grid, created on first use).
!
!   das_letkf_global  das_letkf + set_letkf_obs of the global drivers
!                     (speedy/letkf/letkf_tools.f90:45-267,
!                      speedy/letkf/letkf_obs.f90:48-335): gues3d(nij,nlev,ne,nv3d),
!                     gues2d(nij,ne,nv2d) -> anal3d, anal2d; raw obs records
!                     with H(x) of every member; inflation field in/out.
!   das_letkf_l96     the Lorenz-96 analysis loop (lorenz96/letkf/letkf.f90:216-265).
!
! A nonzero library status stops the program with the reference's code.
!=======================================================================
MODULE letkf_tools_gpu
  USE ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: das_letkf_global, das_letkf_l96, letkf_gpu_release

  TYPE, BIND(C) :: lekf_global_cfg_t
    INTEGER(C_INT) :: nlon, nlat, nlev
    TYPE(C_PTR) :: lon, lat, sig
    INTEGER(C_INT) :: ne, nv3d, nv2d, ips, iq, q_top_lev
    REAL(C_DOUBLE) :: sigma_obs, sigma_obsv, gross_error
    REAL(C_DOUBLE) :: var_local(7)
    INTEGER(C_INT) :: prec, infl_mode
    REAL(C_DOUBLE) :: infl_value
    INTEGER(C_INT) :: nobs_max
  END TYPE

  TYPE, BIND(C) :: lekf_l96_cfg_t
    INTEGER(C_INT) :: nx, ne
    REAL(C_DOUBLE) :: xlocal
    INTEGER(C_INT) :: prec, infl_mode
    REAL(C_DOUBLE) :: infl_value
  END TYPE

  INTERFACE
    TYPE(C_PTR) FUNCTION lekf_global_create(cfg, status) BIND(C, NAME="lekf_global_create")
      IMPORT :: C_PTR, C_INT, lekf_global_cfg_t
      TYPE(lekf_global_cfg_t), INTENT(IN) :: cfg
      INTEGER(C_INT), INTENT(OUT) :: status
    END FUNCTION
    TYPE(C_PTR) FUNCTION lekf_l96_create(cfg, status) BIND(C, NAME="lekf_l96_create")
      IMPORT :: C_PTR, C_INT, lekf_l96_cfg_t
      TYPE(lekf_l96_cfg_t), INTENT(IN) :: cfg
      INTEGER(C_INT), INTENT(OUT) :: status
    END FUNCTION
    SUBROUTINE lekf_destroy(h) BIND(C, NAME="lekf_destroy")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE
    INTEGER(C_INT) FUNCTION lekf_set_ensemble(h, g3, g2) BIND(C, NAME="lekf_set_ensemble")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: g3(*), g2(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION lekf_set_infl(h, i3, i2) BIND(C, NAME="lekf_set_infl")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: i3(*), i2(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION lekf_set_obs_global(h, nobs, elm, lon, lat, lev, dat, err, qc, hxf) &
         & BIND(C, NAME="lekf_set_obs_global")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: nobs
      REAL(C_DOUBLE), INTENT(IN) :: elm(*), lon(*), lat(*), lev(*), dat(*), err(*), hxf(*)
      TYPE(C_PTR), VALUE :: qc
    END FUNCTION
    INTEGER(C_INT) FUNCTION lekf_set_obs_l96(h, ny, gidx, y, err, hxf) BIND(C, NAME="lekf_set_obs_l96")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: ny
      INTEGER(C_INT), INTENT(IN) :: gidx(*)
      REAL(C_DOUBLE), INTENT(IN) :: y(*), err(*), hxf(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION lekf_das_run(h) BIND(C, NAME="lekf_das_run")
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION lekf_get_analysis(h, a3, a2) BIND(C, NAME="lekf_get_analysis")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: a3(*), a2(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION lekf_get_infl(h, i3, i2) BIND(C, NAME="lekf_get_infl")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: i3(*), i2(*)
    END FUNCTION
  END INTERFACE

  TYPE(C_PTR), SAVE :: h_global = C_NULL_PTR
  TYPE(C_PTR), SAVE :: h_l96 = C_NULL_PTR
  INTEGER, SAVE :: g_shape(6) = 0
  INTEGER, SAVE :: l_shape(2) = 0
  REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: g_lon(:), g_lat(:), g_sig(:)

CONTAINS

SUBROUTINE stop_on(st, where)
  INTEGER(C_INT), INTENT(IN) :: st
  CHARACTER(*), INTENT(IN) :: where
  IF (st == 0 .OR. st == 4) RETURN
  WRITE(6, '(3A,I6)') '!!! ERROR (', where, '): GPU library status ', st
  SELECT CASE (st)
  CASE (2)
    STOP 2
  CASE (99)
    STOP 99
  CASE DEFAULT
    STOP 1
  END SELECT
END SUBROUTINE stop_on

SUBROUTINE das_letkf_global(nlon, nlat, nlev, lon, lat, sig, ne, nv3d, nv2d, gues3d, gues2d, &
     & anal3d, anal2d, nobs, obselm, obslon, obslat, obslev, obsdat, obserr, obshxf, &
     & sigma_obs, sigma_obsv, gross_error, infl3d, prec)
  INTEGER, INTENT(IN) :: nlon, nlat, nlev, ne, nv3d, nv2d, nobs
  REAL(C_DOUBLE), INTENT(IN) :: lon(nlon), lat(nlat), sig(nlev)
  REAL(C_DOUBLE), INTENT(IN) :: gues3d(nlon*nlat, nlev, ne, nv3d), gues2d(nlon*nlat, ne, nv2d)
  REAL(C_DOUBLE), INTENT(OUT) :: anal3d(nlon*nlat, nlev, ne, nv3d), anal2d(nlon*nlat, ne, nv2d)
  REAL(C_DOUBLE), INTENT(IN) :: obselm(nobs), obslon(nobs), obslat(nobs), obslev(nobs)
  REAL(C_DOUBLE), INTENT(IN) :: obsdat(nobs), obserr(nobs), obshxf(nobs, ne)
  REAL(C_DOUBLE), INTENT(IN) :: sigma_obs, sigma_obsv, gross_error
  REAL(C_DOUBLE), INTENT(INOUT) :: infl3d(nlon*nlat, nlev)
  INTEGER, INTENT(IN), OPTIONAL :: prec
  TYPE(lekf_global_cfg_t) :: cfg
  INTEGER(C_INT) :: st
  REAL(C_DOUBLE) :: i2(nlon*nlat, MAX(nv2d, 1))
  INTEGER :: shp(6)

  shp = (/ nlon, nlat, nlev, ne, nv3d, nv2d /)
  IF (C_ASSOCIATED(h_global) .AND. ANY(shp /= g_shape)) THEN
    CALL lekf_destroy(h_global)
    h_global = C_NULL_PTR
  END IF
  IF (.NOT. C_ASSOCIATED(h_global)) THEN
    IF (ALLOCATED(g_lon)) DEALLOCATE(g_lon, g_lat, g_sig)
    ALLOCATE(g_lon(nlon), g_lat(nlat), g_sig(nlev))
    g_lon = lon
    g_lat = lat
    g_sig = sig
    cfg%nlon = nlon
    cfg%nlat = nlat
    cfg%nlev = nlev
    cfg%lon = C_LOC(g_lon)
    cfg%lat = C_LOC(g_lat)
    cfg%sig = C_LOC(g_sig)
    cfg%ne = ne
    cfg%nv3d = nv3d
    cfg%nv2d = nv2d
    cfg%ips = 0
    cfg%iq = 3
    cfg%q_top_lev = 5          ! speedy/letkf/letkf_tools.f90:172
    cfg%sigma_obs = sigma_obs
    cfg%sigma_obsv = sigma_obsv
    cfg%gross_error = gross_error
    cfg%var_local = 1.0d0
    cfg%prec = 0
    IF (PRESENT(prec)) cfg%prec = prec
    cfg%infl_mode = 0
    cfg%infl_value = 1.01d0
    cfg%nobs_max = 0
    h_global = lekf_global_create(cfg, st)
    CALL stop_on(st, 'lekf_global_create')
    g_shape = shp
  END IF
  i2 = 1.0d0
  CALL stop_on(lekf_set_ensemble(h_global, gues3d, gues2d), 'lekf_set_ensemble')
  CALL stop_on(lekf_set_infl(h_global, infl3d, i2), 'lekf_set_infl')
  CALL stop_on(lekf_set_obs_global(h_global, INT(nobs, C_INT), obselm, obslon, obslat, obslev, &
       & obsdat, obserr, C_NULL_PTR, obshxf), 'lekf_set_obs_global')
  CALL stop_on(lekf_das_run(h_global), 'lekf_das_run')
  CALL stop_on(lekf_get_analysis(h_global, anal3d, anal2d), 'lekf_get_analysis')
  CALL stop_on(lekf_get_infl(h_global, infl3d, i2), 'lekf_get_infl')
END SUBROUTINE das_letkf_global

SUBROUTINE das_letkf_l96(nx, ne, xf, xa, ny, gidx, y, err, hxf, xlocal, infl)
  INTEGER, INTENT(IN) :: nx, ne, ny
  REAL(C_DOUBLE), INTENT(IN) :: xf(nx, ne)
  REAL(C_DOUBLE), INTENT(OUT) :: xa(nx, ne)
  INTEGER(C_INT), INTENT(IN) :: gidx(ny)      ! 0-based grid index of each obs
  REAL(C_DOUBLE), INTENT(IN) :: y(ny), err(ny), hxf(ny, ne), xlocal, infl
  TYPE(lekf_l96_cfg_t) :: cfg
  INTEGER(C_INT) :: st
  REAL(C_DOUBLE) :: dummy(1)

  IF (C_ASSOCIATED(h_l96) .AND. (l_shape(1) /= nx .OR. l_shape(2) /= ne)) THEN
    CALL lekf_destroy(h_l96)
    h_l96 = C_NULL_PTR
  END IF
  IF (.NOT. C_ASSOCIATED(h_l96)) THEN
    cfg%nx = nx
    cfg%ne = ne
    cfg%xlocal = xlocal
    cfg%prec = 0
    cfg%infl_mode = 1          ! fixed multiplicative inflation (lorenz96/letkf/letkf.f90:256)
    cfg%infl_value = infl
    h_l96 = lekf_l96_create(cfg, st)
    CALL stop_on(st, 'lekf_l96_create')
    l_shape = (/ nx, ne /)
  END IF
  dummy = 0.0d0
  CALL stop_on(lekf_set_ensemble(h_l96, xf, dummy), 'lekf_set_ensemble')
  CALL stop_on(lekf_set_obs_l96(h_l96, INT(ny, C_INT), gidx, y, err, hxf), 'lekf_set_obs_l96')
  CALL stop_on(lekf_das_run(h_l96), 'lekf_das_run')
  CALL stop_on(lekf_get_analysis(h_l96, xa, dummy), 'lekf_get_analysis')
END SUBROUTINE das_letkf_l96

SUBROUTINE letkf_gpu_release()
  IF (C_ASSOCIATED(h_global)) CALL lekf_destroy(h_global)
  IF (C_ASSOCIATED(h_l96)) CALL lekf_destroy(h_l96)
  h_global = C_NULL_PTR
  h_l96 = C_NULL_PTR
END SUBROUTINE letkf_gpu_release

END MODULE letkf_tools_gpu
