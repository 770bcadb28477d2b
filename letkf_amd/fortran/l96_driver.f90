!=======================================================================
! l96_driver -- one Lorenz-96 analysis step through the Fortran host side.
!
! The per-point loop is the reference's (lorenz96/letkf/letkf.f90:216-265,
! nwindow = 1, fixed localisation, fixed inflation) with USE common_letkf
! resolving to the GPU drop-in; the same step is then run through the
! das-level module (letkf_tools_gpu::das_letkf_l96).
!
! usage: l96_driver <input> <output>
!   input  (stream): nx, ne, ny (int32), xf(nx,ne), gidx(ny) (int32, 0-based),
!                    y(ny), err(ny), hxf(ny,ne)  (float64, column-major)
!   output (stream): xa(nx,ne) from the letkf_core loop, then xa(nx,ne) from
!                    das_letkf_l96
!=======================================================================
PROGRAM l96_driver
  USE ISO_C_BINDING
  USE common_letkf, ONLY: letkf_core
  USE letkf_tools_gpu, ONLY: das_letkf_l96, letkf_gpu_release
  IMPLICIT NONE
  CHARACTER(512) :: fin, fout
  INTEGER(C_INT) :: nx, ne, ny
  REAL(C_DOUBLE), ALLOCATABLE :: xf(:,:), xa(:,:), xa2(:,:), y(:), err(:), hxf(:,:)
  REAL(C_DOUBLE), ALLOCATABLE :: hdxf(:,:), d(:), hm(:), dxf(:,:), xm(:)
  REAL(C_DOUBLE), ALLOCATABLE :: hloc(:,:), rloc(:), rdiag(:), dloc(:), trans(:,:)
  INTEGER(C_INT), ALLOCATABLE :: gidx(:)
  REAL(C_DOUBLE) :: parm, dist, xmaxloc
  REAL(C_DOUBLE), PARAMETER :: xlocal = 3.0d0, infl = 1.05d0
  INTEGER :: ix, i, k, m, nyl

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) nx, ne, ny
  ALLOCATE(xf(nx,ne), xa(nx,ne), xa2(nx,ne), y(ny), err(ny), hxf(ny,ne), gidx(ny))
  ALLOCATE(hdxf(ny,ne), d(ny), hm(ny), dxf(nx,ne), xm(nx))
  ALLOCATE(hloc(ny,ne), rloc(ny), rdiag(ny), dloc(ny), trans(ne,ne))
  READ(10) xf, gidx, y, err, hxf
  CLOSE(10)
  ! ensemble mean and perturbations (letkf.f90:153-165, 185-214)
  DO ix = 1, nx
    xm(ix) = xf(ix,1)
    DO m = 2, ne
      xm(ix) = xm(ix) + xf(ix,m)
    END DO
    xm(ix) = xm(ix) / REAL(ne, C_DOUBLE)
    dxf(ix,:) = xf(ix,:) - xm(ix)
  END DO
  DO i = 1, ny
    hm(i) = hxf(i,1)
    DO m = 2, ne
      hm(i) = hm(i) + hxf(i,m)
    END DO
    hm(i) = hm(i) / REAL(ne, C_DOUBLE)
    hdxf(i,:) = hxf(i,:) - hm(i)
    d(i) = y(i) - hm(i)
  END DO
  xmaxloc = xlocal * 2.0d0 * SQRT(10.0d0/3.0d0)
  DO ix = 1, nx
    nyl = 0
    DO i = 1, ny
      dist = REAL(MIN(ABS(gidx(i)+1-ix), nx-ABS(gidx(i)+1-ix)), C_DOUBLE)
      IF (dist < xmaxloc) THEN
        nyl = nyl + 1
        dloc(nyl) = d(i)
        rdiag(nyl) = err(i)**2
        rloc(nyl) = EXP(-0.5d0 * (dist/xlocal)**2)
        hloc(nyl,:) = hdxf(i,:)
        IF (rloc(nyl) < 0.0001d0) nyl = nyl - 1
      END IF
    END DO
    parm = infl
    CALL letkf_core(ne, ny, nyl, hloc, rdiag, rloc, dloc, parm, trans)
    DO m = 1, ne
      xa(ix,m) = xm(ix)
      DO k = 1, ne
        xa(ix,m) = xa(ix,m) + dxf(ix,k) * trans(k,m)
      END DO
    END DO
  END DO
  CALL das_letkf_l96(nx, ne, xf, xa2, ny, gidx, y, err, hxf, xlocal, infl)
  CALL letkf_gpu_release()
  OPEN(11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(11) xa, xa2
  CLOSE(11)
END PROGRAM l96_driver
