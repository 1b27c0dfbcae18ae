!-----------------------------------------------------------------------
! gmtslice_gpu -- drop-in replacement for the reference's model slicer
! (gmtplot/gmtslicet.f90).  Reads gmtslicet.in and the two .vtx models it
! names from the current directory in the same way, writes the same
! bounds files (plot extents, value counts, vertex counts) and the same
! velocity-perturbation slice files (one list-directed value per line, in
! the same order); the B-spline evaluation runs on the GPU through
! fmm3d_slice_velocity (include/fmm3d.h).  FMM3D_DEVICE selects the device
! (default 0).
!-----------------------------------------------------------------------
PROGRAM gmtslice_gpu
   USE ISO_C_BINDING
   USE fmm3d_c
   IMPLICIT NONE
   INTEGER :: extrds, extrnss, extrews, pvgd, pvgns, pvgew, ddt, ddp, dnsr, dnst, dewr, dewp
   INTEGER :: nnr, nnt, nnp, rnnr, rnnt, rnnp, idm1, idm2, idm3, isw, i, ios, dev, id1, id2
   REAL(C_DOUBLE) :: gor, got, gop, rgsr, rgst, rgsp, rgor, rgot, rgop, rrgsr, rrgst, rrgsp
   REAL(C_DOUBLE) :: sldep, slns, slew, lft, rgt, btm, top, rd1, rd2, rd3
   CHARACTER(LEN=30) :: ifilev, ifilevr, ofiledb, ofilensb, ofileewb, ofiledv, ofilensv, ofileewv
   CHARACTER(LEN=32) :: env
   REAL(C_DOUBLE), ALLOCATABLE, TARGET :: velv(:), vref(:)
   TYPE(fmm3d_grid_desc) :: g
   LOGICAL :: have_ref

   OPEN(UNIT=10, FILE='gmtslicet.in', STATUS='old')
   READ(10, *)
   READ(10, *)
   READ(10, *)
   READ(10, '(a26)') ifilev
   READ(10, '(a26)') ifilevr
   READ(10, *)
   READ(10, *)
   READ(10, *)
   READ(10, *) extrds
   READ(10, *) sldep
   READ(10, '(a22)') ofiledb
   READ(10, *) extrnss
   READ(10, *) slns
   READ(10, '(a22)') ofilensb
   READ(10, *) extrews
   READ(10, *) slew
   READ(10, '(a22)') ofileewb

   OPEN(UNIT=20, FILE=ifilev, STATUS='old')
   READ(20, *) nnr, nnt, nnp
   READ(20, *) gor, got, gop
   READ(20, *) rgsr, rgst, rgsp
   READ(20, *) rnnr, rnnt, rnnp
   READ(20, *) rgor, rgot, rgop
   READ(20, *) rrgsr, rrgst, rrgsp
   ALLOCATE(velv((nnr + 2) * (nnt + 2) * (nnp + 2)), vref((nnr + 2) * (nnt + 2) * (nnp + 2)))
   DO i = 1, SIZE(velv)
      READ(20, *) velv(i)
   END DO
   CLOSE(20)

   READ(10, *)
   READ(10, *)
   READ(10, *)
   READ(10, *) pvgd
   READ(10, *) ddt, ddp
   READ(10, '(a22)') ofiledv
   READ(10, *) pvgns
   READ(10, *) dnsr, dnst
   READ(10, '(a22)') ofilensv
   READ(10, *) pvgew
   READ(10, *) dewr, dewp
   READ(10, '(a22)') ofileewv
   CLOSE(10)

   IF (extrds == 1) THEN
      lft = rgop
      rgt = rgop + (rnnp - 1) * rrgsp
      btm = rgot - (rnnt - 1) * rrgst
      top = rgot
      OPEN(UNIT=50, FILE=ofiledb, STATUS='unknown')
      WRITE(50, '(f16.11)') lft
      WRITE(50, '(f16.11)') rgt
      WRITE(50, '(f16.11)') btm
      WRITE(50, '(f16.11)') top
      id1 = (rnnp - 1) * ddp + 1
      id2 = (rnnt - 1) * ddt + 1
      WRITE(50, *) id1
      WRITE(50, *) id2
   END IF
   IF (extrnss == 1) THEN
      rgt = rgot - (rnnt - 1) * rrgst
      lft = rgot
      btm = rgor - (rnnr - 1) * rrgsr
      top = rgor
      OPEN(UNIT=60, FILE=ofilensb, STATUS='unknown')
      WRITE(60, '(f16.11)') rgt
      WRITE(60, '(f16.11)') lft
      WRITE(60, '(f16.11)') btm
      WRITE(60, '(f16.11)') top
      id1 = (rnnt - 1) * dnst + 1
      id2 = (rnnr - 3) * dnsr + 1
      WRITE(60, *) id1
      WRITE(60, *) id2
   END IF
   IF (extrews == 1) THEN
      lft = rgop
      rgt = rgop + (rnnp - 1) * rrgsp
      btm = rgor - (rnnr - 1) * rrgsr
      top = rgor
      OPEN(UNIT=70, FILE=ofileewb, STATUS='unknown')
      WRITE(70, '(f16.11)') lft
      WRITE(70, '(f16.11)') rgt
      WRITE(70, '(f16.11)') btm
      WRITE(70, '(f16.11)') top
      id1 = (rnnp - 1) * dewp + 1
      id2 = (rnnr - 3) * dewr + 1
      WRITE(70, *) id1
      WRITE(70, *) id2
   END IF

   ! the velocity slices are perturbations from the reference model (its
   ! header must match; the reference reports a mismatch and goes on)
   have_ref = .FALSE.
   IF (pvgd == 1 .OR. pvgns == 1 .OR. pvgew == 1) THEN
      isw = 0
      OPEN(UNIT=20, FILE=ifilevr, STATUS='old')
      READ(20, *) idm1, idm2, idm3
      IF (idm1 /= nnr .OR. idm2 /= nnt .OR. idm3 /= nnp) isw = 1
      READ(20, *) rd1, rd2, rd3
      IF (rd1 /= gor .OR. rd2 /= got .OR. rd3 /= gop) isw = 1
      READ(20, *) rd1, rd2, rd3
      IF (rd1 /= rgsr .OR. rd2 /= rgst .OR. rd3 /= rgsp) isw = 1
      IF (isw == 1) THEN
         WRITE(6, *) 'ERROR! Actual velocity grid and reference'
         WRITE(6, *) 'velocity grid have different dimensions or'
         WRITE(6, *) 'different numbers of grid points!'
         WRITE(6, *) 'TERMINATING PROGRAM!!!'
      END IF
      READ(20, *)
      READ(20, *)
      READ(20, *)
      DO i = 1, SIZE(vref)
         READ(20, *) vref(i)
      END DO
      CLOSE(20)
      have_ref = .TRUE.
   END IF

   dev = 0
   CALL GET_ENVIRONMENT_VARIABLE('FMM3D_DEVICE', env, STATUS=ios)
   IF (ios == 0) READ(env, *) dev
   g%nvr = nnr; g%nvx = nnt; g%nvz = nnp
   g%gor = gor; g%goxd = got; g%gozd = gop
   g%dvr = rgsr; g%dvxd = rgst; g%dvzd = rgsp
   g%nnr = rnnr; g%nnx = rnnt; g%nnz = rnnp
   g%rgor = rgor; g%rgoxd = rgot; g%rgozd = rgop
   g%dnr = rrgsr; g%dnxd = rrgst; g%dnzd = rrgsp
   g%earth = 0.0
   IF (pvgd == 1) CALL slice_out(0, sldep, ddt, ddp, ofiledv)
   IF (pvgns == 1) CALL slice_out(1, slns, dnsr, dnst, ofilensv)
   IF (pvgew == 1) CALL slice_out(2, slew, dewr, dewp, ofileewv)

   IF (extrds == 1) THEN
      WRITE(50, *) nnp
      WRITE(50, *) nnt
      CLOSE(50)
   END IF
   IF (extrnss == 1) THEN
      WRITE(60, *) nnt
      WRITE(60, *) nnr
      CLOSE(60)
   END IF
   IF (extrews == 1) THEN
      WRITE(70, *) nnp
      WRITE(70, *) nnr
      CLOSE(70)
   END IF
   STOP
CONTAINS
   SUBROUTINE slice_out(kind, where, d1, d2, fname)
      INTEGER, INTENT(IN) :: kind, d1, d2
      REAL(C_DOUBLE), INTENT(IN) :: where
      CHARACTER(LEN=*), INTENT(IN) :: fname
      REAL(C_DOUBLE), ALLOCATABLE :: out(:)
      REAL(C_DOUBLE) :: probe(1)
      INTEGER(C_INT) :: rc, n1, n2
      INTEGER :: k
      TYPE(C_PTR) :: refp
      refp = C_NULL_PTR
      IF (have_ref) refp = C_LOC(vref)
      rc = fmm3d_slice_velocity(g, INT(dev, C_INT), velv, refp, INT(kind, C_INT), where, INT(d1, C_INT), &
                                INT(d2, C_INT), probe, 0_C_INT64_T, n1, n2)
      IF (rc /= -6) CALL slice_check(rc)
      ALLOCATE(out(INT(n1, 8) * n2))
      rc = fmm3d_slice_velocity(g, INT(dev, C_INT), velv, refp, INT(kind, C_INT), where, INT(d1, C_INT), &
                                INT(d2, C_INT), out, INT(SIZE(out), C_INT64_T), n1, n2)
      CALL slice_check(rc)
      OPEN(UNIT=30, FILE=fname, STATUS='unknown')
      DO k = 1, SIZE(out)
         WRITE(30, *) out(k)
      END DO
      CLOSE(30)
      DEALLOCATE(out)
   END SUBROUTINE slice_out

   SUBROUTINE slice_check(rc)
      INTEGER(C_INT), INTENT(IN) :: rc
      TYPE(C_PTR) :: msg
      CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
      INTEGER :: n, j
      CHARACTER(LEN=512) :: text
      IF (rc == 0) RETURN
      msg = fmm3d_slice_last_error()
      n = INT(c_strlen(msg))
      CALL C_F_POINTER(msg, chars, [n])
      text = ' '
      DO j = 1, MIN(n, 512)
         text(j:j) = chars(j)
      END DO
      WRITE(6, *) 'ERROR in fmm3d_slice_velocity (status ', rc, '): ', TRIM(text)
      STOP 1
   END SUBROUTINE slice_check
END PROGRAM gmtslice_gpu
