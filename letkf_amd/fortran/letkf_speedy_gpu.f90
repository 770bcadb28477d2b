!=======================================================================
! letkf_obs / letkf_tools -- drop-in replacements of the SPEEDY LETKF's
! modules of the same names (speedy/letkf/letkf_obs.f90:1-41 and
! speedy/letkf/letkf_tools.f90:1-51) over the GPU library.
!
! The SPEEDY main program keeps its sequence (speedy/letkf/letkf.f90:63-135):
!
!   CALL set_letkf_obs                          obs records + every member's
!                                               H(x) (the files obsope wrote)
!   CALL das_letkf(gues3d,gues2d,anal3d,anal2d) same dummy arguments as
!                                               speedy/letkf/letkf_tools.f90:45-51
!
! set_letkf_obs_records takes the records in memory (the reference reads
! them with read_obs2 from obs01NNN.dat, one file per member); departures,
! the gross-error check (speedy/letkf/letkf_obs.f90:146-174) and the
! lat/lon binning (:243-313) run on the GPU inside das_letkf.  The grid is
! SPEEDY's T30 grid (speedy/common/common_speedy.f90:64-123) unless
! set_grid gives another.  Inflation follows the reference's cov_infl_mul
! (speedy/letkf/letkf_tools.f90:22-24,113-146): > 0 a constant field with
! 1.01 at the top level, <= 0 the field infl_mul3d (the reference's
! infl_mul.grd), which the analysis updates.
!=======================================================================
MODULE letkf_obs
  USE ISO_C_BINDING
  IMPLICIT NONE
  PUBLIC

  INTEGER, SAVE :: nobs = 0
  ! localisation scales and the gross-error factor: the reference's values
  ! (speedy/letkf/letkf_obs.f90:24-27,50), settable by the caller
  REAL(C_DOUBLE), SAVE :: sigma_obs = 1000.0d3
  REAL(C_DOUBLE), SAVE :: sigma_obsv = 0.1d0
  REAL(C_DOUBLE), SAVE :: gross_error = 10.0d0
  REAL(C_DOUBLE), ALLOCATABLE, SAVE :: obselm(:), obslon(:), obslat(:), obslev(:)
  REAL(C_DOUBLE), ALLOCATABLE, SAVE :: obsdat(:), obserr(:)
  REAL(C_DOUBLE), ALLOCATABLE, SAVE :: obshxf(:,:)   ! H(x) of every member (nobs, nbv)

CONTAINS

SUBROUTINE set_letkf_obs_records(elm, lon, lat, lev, dat, err, hxf)
  REAL(C_DOUBLE), INTENT(IN) :: elm(:), lon(:), lat(:), lev(:), dat(:), err(:), hxf(:,:)
  nobs = SIZE(elm)
  IF (ALLOCATED(obselm)) DEALLOCATE(obselm, obslon, obslat, obslev, obsdat, obserr, obshxf)
  ALLOCATE(obselm(nobs), obslon(nobs), obslat(nobs), obslev(nobs), obsdat(nobs), obserr(nobs))
  ALLOCATE(obshxf(nobs, SIZE(hxf, 2)))
  obselm = elm
  obslon = lon
  obslat = lat
  obslev = lev
  obsdat = dat
  obserr = err
  obshxf = hxf
END SUBROUTINE set_letkf_obs_records

END MODULE letkf_obs

MODULE letkf_tools
  USE ISO_C_BINDING
  USE letkf_obs
  USE letkf_tools_gpu, ONLY: das_letkf_global
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: das_letkf, set_grid, cov_infl_mul, infl_mul3d

  ! > 0: globally constant inflation; <= 0: the field infl_mul3d, updated
  REAL(C_DOUBLE), SAVE :: cov_infl_mul = -1.01d0
  REAL(C_DOUBLE), ALLOCATABLE, SAVE :: infl_mul3d(:,:)   ! (nij, nlev)

  INTEGER, SAVE :: nlon = 96, nlat = 48
  REAL(C_DOUBLE), ALLOCATABLE, SAVE :: glon(:), glat(:), gsig(:)

CONTAINS

! grid of the analysis: longitudes (deg, ascending in [0, 360)), latitudes
! (deg, ascending), sigma levels
SUBROUTINE set_grid(lon, lat, sig)
  REAL(C_DOUBLE), INTENT(IN) :: lon(:), lat(:), sig(:)
  IF (ALLOCATED(glon)) DEALLOCATE(glon, glat, gsig)
  nlon = SIZE(lon)
  nlat = SIZE(lat)
  ALLOCATE(glon(nlon), glat(nlat), gsig(SIZE(sig)))
  glon = lon
  glat = lat
  gsig = sig
END SUBROUTINE set_grid

! SPEEDY's T30 grid (speedy/common/common_speedy.f90:64-123): 3.75-degree
! longitudes, the Gaussian latitudes to three decimals, seven sigma levels
SUBROUTINE default_grid()
  INTEGER :: i
  REAL(C_DOUBLE), PARAMETER :: lat30(24) = (/ &
       & 1.856d0, 5.567d0, 9.278d0, 12.989d0, 16.700d0, 20.411d0, 24.122d0, 27.833d0, &
       & 31.545d0, 35.256d0, 38.967d0, 42.678d0, 46.389d0, 50.099d0, 53.810d0, 57.521d0, &
       & 61.232d0, 64.942d0, 68.652d0, 72.362d0, 76.070d0, 79.777d0, 83.479d0, 87.159d0 /)
  REAL(C_DOUBLE), PARAMETER :: sig7(7) = (/ 0.95d0, 0.835d0, 0.685d0, 0.51d0, 0.34d0, &
       & 0.2d0, 0.08d0 /)
  REAL(C_DOUBLE) :: lon(96), lat(48)
  DO i = 1, 96
    lon(i) = 3.75d0 * REAL(i - 1, C_DOUBLE)
  END DO
  DO i = 1, 24
    lat(24 + i) = lat30(i)
    lat(25 - i) = -lat30(i)
  END DO
  CALL set_grid(lon, lat, sig7)
END SUBROUTINE default_grid

SUBROUTINE das_letkf(gues3d, gues2d, anal3d, anal2d)
  REAL(C_DOUBLE), INTENT(INOUT) :: gues3d(:,:,:,:)   ! (nij, nlev, nbv, nv3d)
  REAL(C_DOUBLE), INTENT(INOUT) :: gues2d(:,:,:)     ! (nij, nbv, nv2d)
  REAL(C_DOUBLE), INTENT(OUT) :: anal3d(:,:,:,:)
  REAL(C_DOUBLE), INTENT(OUT) :: anal2d(:,:,:)
  REAL(C_DOUBLE), ALLOCATABLE :: infl3d(:,:)
  INTEGER :: nij, nlev, nbv, nv3d, nv2d

  IF (.NOT. ALLOCATED(glon)) CALL default_grid()
  nij = SIZE(gues3d, 1)
  nlev = SIZE(gues3d, 2)
  nbv = SIZE(gues3d, 3)
  nv3d = SIZE(gues3d, 4)
  nv2d = SIZE(gues2d, 3)
  IF (nij /= nlon * nlat .OR. nlev /= SIZE(gsig) .OR. SIZE(gues2d, 1) /= nij .OR. &
       & SIZE(gues2d, 2) /= nbv .OR. SIZE(obshxf, 2) /= nbv) THEN
    WRITE(6, '(A)') '!!! ERROR (das_letkf): array shapes do not match the grid'
    STOP 99
  END IF
  ALLOCATE(infl3d(nij, nlev))
  IF (cov_infl_mul > 0.0d0) THEN
    ! speedy/letkf/letkf_tools.f90:113-121
    infl3d = cov_infl_mul
    infl3d(:, nlev) = 1.01d0
  ELSE
    IF (.NOT. ALLOCATED(infl_mul3d)) THEN
      WRITE(6, '(A)') '!!! ERROR (das_letkf): cov_infl_mul <= 0 needs the field infl_mul3d'
      STOP 99
    END IF
    infl3d = infl_mul3d
  END IF
  CALL das_letkf_global(nlon, nlat, nlev, glon, glat, gsig, nbv, nv3d, nv2d, gues3d, gues2d, &
       & anal3d, anal2d, nobs, obselm, obslon, obslat, obslev, obsdat, obserr, obshxf, &
       & sigma_obs, sigma_obsv, gross_error, infl3d)
  IF (cov_infl_mul <= 0.0d0) infl_mul3d = infl3d
  DEALLOCATE(infl3d)
END SUBROUTINE das_letkf

END MODULE letkf_tools
