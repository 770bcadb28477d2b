!=======================================================================
! speedy_driver -- the SPEEDY LETKF main program's analysis sequence
! (speedy/letkf/letkf.f90:79-109) through the drop-in modules letkf_obs and
! letkf_tools: set the obs, then CALL das_letkf(gues3d,gues2d,anal3d,anal2d).
!
! usage: speedy_driver <input> <output>
!   input (stream, column-major):
!     nlon, nlat, nlev, nbv, nv3d, nv2d, nobs (int32)
!     lon(nlon), lat(nlat), sig(nlev), gues3d(nlon*nlat,nlev,nbv,nv3d),
!     gues2d(nlon*nlat,nbv,nv2d), elm, lon, lat, lev, dat, err (nobs each),
!     hxf(nobs,nbv), sigma_obs, sigma_obsv, gross_error, cov_infl_mul,
!     infl3d(nlon*nlat,nlev)                                 (float64)
!   output (stream): anal3d, anal2d, then the inflation field after the
!     analysis (cov_infl_mul <= 0; the input field otherwise)
!=======================================================================
PROGRAM speedy_driver
  USE ISO_C_BINDING
  USE letkf_obs
  USE letkf_tools
  USE letkf_tools_gpu, ONLY: letkf_gpu_release
  IMPLICIT NONE
  CHARACTER(512) :: fin, fout
  INTEGER(C_INT) :: nlon_, nlat_, nlev, nbv, nv3d, nv2d, nob
  REAL(C_DOUBLE), ALLOCATABLE :: lon(:), lat(:), sig(:)
  REAL(C_DOUBLE), ALLOCATABLE :: gues3d(:,:,:,:), gues2d(:,:,:), anal3d(:,:,:,:), anal2d(:,:,:)
  REAL(C_DOUBLE), ALLOCATABLE :: elm(:), olon(:), olat(:), olev(:), dat(:), err(:), hxf(:,:)
  REAL(C_DOUBLE), ALLOCATABLE :: infl(:,:)
  REAL(C_DOUBLE) :: s(4)
  INTEGER :: nij

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) nlon_, nlat_, nlev, nbv, nv3d, nv2d, nob
  nij = nlon_ * nlat_
  ALLOCATE(lon(nlon_), lat(nlat_), sig(nlev))
  ALLOCATE(gues3d(nij,nlev,nbv,nv3d), gues2d(nij,nbv,nv2d))
  ALLOCATE(anal3d(nij,nlev,nbv,nv3d), anal2d(nij,nbv,nv2d))
  ALLOCATE(elm(nob), olon(nob), olat(nob), olev(nob), dat(nob), err(nob), hxf(nob,nbv))
  ALLOCATE(infl(nij,nlev))
  READ(10) lon, lat, sig, gues3d, gues2d, elm, olon, olat, olev, dat, err, hxf, s, infl
  CLOSE(10)
  CALL set_grid(lon, lat, sig)
  sigma_obs = s(1)
  sigma_obsv = s(2)
  gross_error = s(3)
  cov_infl_mul = s(4)
  IF (cov_infl_mul <= 0.0d0) THEN
    ALLOCATE(infl_mul3d(nij,nlev))
    infl_mul3d = infl
  END IF
  CALL set_letkf_obs_records(elm, olon, olat, olev, dat, err, hxf)
  CALL das_letkf(gues3d, gues2d, anal3d, anal2d)
  IF (cov_infl_mul <= 0.0d0) infl = infl_mul3d
  CALL letkf_gpu_release()
  OPEN(11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(11) anal3d, anal2d, infl
  CLOSE(11)
END PROGRAM speedy_driver
