!-----------------------------------------------------------------------
! fm3dt_core -- shared parts of the Fortran hosts (fm3dt_gpu, fm3dt_invert):
! the ISO_C_BINDING interface of libfmm3d_hip.so (include/fmm3d.h), the
! reading of fm3dt_parallel.in and the .vtx model, and one forward pass
! over the reference's files, as PROGRAM fmmin3d does
! (fm3dt_parallel.f90:6015-6662).
!-----------------------------------------------------------------------
MODULE fmm3d_c
   USE ISO_C_BINDING
   IMPLICIT NONE
   TYPE, BIND(C) :: fmm3d_grid_desc
      INTEGER(C_INT) :: nvr, nvx, nvz
      REAL(C_DOUBLE) :: gor, goxd, gozd, dvr, dvxd, dvzd
      INTEGER(C_INT) :: nnr, nnx, nnz
      REAL(C_DOUBLE) :: rgor, rgoxd, rgozd, dnr, dnxd, dnzd
      REAL(C_DOUBLE) :: earth
   END TYPE fmm3d_grid_desc
   INTERFACE
      INTEGER(C_INT) FUNCTION fmm3d_create(g, device, ctx) BIND(C, NAME='fmm3d_create')
         IMPORT :: C_INT, C_PTR, fmm3d_grid_desc
         TYPE(fmm3d_grid_desc), INTENT(IN) :: g
         INTEGER(C_INT), VALUE :: device
         TYPE(C_PTR) :: ctx
      END FUNCTION
      INTEGER(C_INT) FUNCTION fmm3d_destroy(ctx) BIND(C, NAME='fmm3d_destroy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
      END FUNCTION
      TYPE(C_PTR) FUNCTION fmm3d_last_error(ctx) BIND(C, NAME='fmm3d_last_error')
         IMPORT :: C_PTR
         TYPE(C_PTR), VALUE :: ctx
      END FUNCTION
      INTEGER(C_INT) FUNCTION fmm3d_set_model_vtx(ctx, velv) BIND(C, NAME='fmm3d_set_model_vtx')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         REAL(C_DOUBLE), INTENT(IN) :: velv(*)
      END FUNCTION
      INTEGER(C_INT) FUNCTION fmm3d_solve_tele(ctx, nsrc, faces, order, prec, flags) BIND(C, NAME='fmm3d_solve_tele')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: nsrc, order, prec, flags
         REAL(C_DOUBLE), INTENT(IN) :: faces(*)
      END FUNCTION
      INTEGER(C_INT) FUNCTION fmm3d_receivers(ctx, nrc, rcv, t_out) BIND(C, NAME='fmm3d_receivers')
         IMPORT :: C_INT, C_PTR, C_DOUBLE, C_FLOAT
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: nrc
         REAL(C_DOUBLE), INTENT(IN) :: rcv(*)
         REAL(C_FLOAT) :: t_out(*)
      END FUNCTION
      INTEGER(C_INT) FUNCTION fmm3d_frechet(ctx, nrc, rcv, row_ptr, cols, vals, cap) BIND(C, NAME='fmm3d_frechet')
         IMPORT :: C_INT, C_PTR, C_DOUBLE, C_FLOAT, C_INT32_T, C_INT64_T
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: nrc
         REAL(C_DOUBLE), INTENT(IN) :: rcv(*)
         INTEGER(C_INT64_T) :: row_ptr(*)
         INTEGER(C_INT32_T) :: cols(*)
         REAL(C_FLOAT) :: vals(*)
         INTEGER(C_INT64_T), VALUE :: cap
      END FUNCTION
      INTEGER(C_INT) FUNCTION fmm3d_slice_velocity(g, device, velv, velv_ref, kind, where, dice1, dice2, &
                                                  out, cap, n1, n2) BIND(C, NAME='fmm3d_slice_velocity')
         IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT64_T, fmm3d_grid_desc
         TYPE(fmm3d_grid_desc), INTENT(IN) :: g
         INTEGER(C_INT), VALUE :: device, kind, dice1, dice2
         REAL(C_DOUBLE), INTENT(IN) :: velv(*)
         TYPE(C_PTR), VALUE :: velv_ref
         REAL(C_DOUBLE), VALUE :: where
         REAL(C_DOUBLE) :: out(*)
         INTEGER(C_INT64_T), VALUE :: cap
         INTEGER(C_INT) :: n1, n2
      END FUNCTION
      TYPE(C_PTR) FUNCTION fmm3d_slice_last_error() BIND(C, NAME='fmm3d_slice_last_error')
         IMPORT :: C_PTR
      END FUNCTION
      INTEGER(C_SIZE_T) FUNCTION c_strlen(s) BIND(C, NAME='strlen')
         IMPORT :: C_SIZE_T, C_PTR
         TYPE(C_PTR), VALUE :: s
      END FUNCTION
   END INTERFACE
CONTAINS
   SUBROUTINE check(rc, ctx, what)
      INTEGER(C_INT), INTENT(IN) :: rc
      TYPE(C_PTR), INTENT(IN) :: ctx
      CHARACTER(LEN=*), INTENT(IN) :: what
      TYPE(C_PTR) :: msg
      CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
      INTEGER :: n, i
      CHARACTER(LEN=512) :: text
      IF (rc == 0) RETURN
      msg = fmm3d_last_error(ctx)
      n = INT(c_strlen(msg))
      CALL C_F_POINTER(msg, chars, [n])
      text = ' '
      DO i = 1, MIN(n, 512)
         text(i:i) = chars(i)
      END DO
      WRITE(6,*) 'ERROR in ', what, ' (status ', rc, '): ', TRIM(text)
      STOP 1
   END SUBROUTINE check
END MODULE fmm3d_c

MODULE fm3dt_run
   USE ISO_C_BINDING
   USE fmm3d_c
   IMPLICIT NONE
   REAL(KIND=C_DOUBLE), PARAMETER :: pi = 3.1415926535898   ! default-real literal, as the reference
   TYPE run_config
      CHARACTER(LEN=20) :: sources, itimes, receivers, gridv, rtravel, frechet
      REAL(C_DOUBLE) :: earth
      INTEGER :: fom, fsrt, cfd
      INTEGER :: dev, batch, prec, flags
   END TYPE run_config
CONTAINS
   ! fm3dt_parallel.in (fm3dt_parallel.f90:6082-6102) and the environment:
   ! FMM3D_DEVICE (default 0), FMM3D_PRECISION=fp32 for the float32 solver
   ! (default float64, the reference's arithmetic), FMM3D_BATCH (sources per
   ! device call, default 64).  In float64 the default solve reproduces the
   ! reference's own output, its 2x2x2 partitioned result
   ! (FMM3D_REFERENCE_PARTITION); FMM3D_DOMAIN=whole solves the grid as one
   ! domain instead (the float32 solver always does).
   SUBROUTINE read_config(cfg)
      TYPE(run_config), INTENT(OUT) :: cfg
      CHARACTER(LEN=20) :: cdum
      CHARACTER(LEN=32) :: env
      INTEGER :: ltfr, ios
      REAL(C_DOUBLE) :: cslat, cslong
      OPEN(UNIT=10, FILE='fm3dt_parallel.in', STATUS='old')
      READ(10, '(a20)') cdum
      READ(10, '(a20)') cdum
      READ(10, '(a20)') cdum
      READ(10, '(a20)') cfg%sources
      READ(10, '(a20)') cfg%itimes
      READ(10, '(a20)') cfg%receivers
      READ(10, '(a20)') cfg%gridv
      READ(10, *) cfg%earth
      READ(10, *) cfg%fom
      READ(10, *) ltfr
      READ(10, *) cslat, cslong
      READ(10, '(a20)') cdum
      READ(10, '(a20)') cdum
      READ(10, '(a20)') cdum
      READ(10, *) cfg%fsrt
      READ(10, '(a20)') cfg%rtravel
      READ(10, *) cfg%cfd
      READ(10, '(a20)') cfg%frechet
      CLOSE(10)
      cfg%dev = 0
      CALL GET_ENVIRONMENT_VARIABLE('FMM3D_DEVICE', env, STATUS=ios)
      IF (ios == 0) READ(env, *) cfg%dev
      cfg%batch = 64
      CALL GET_ENVIRONMENT_VARIABLE('FMM3D_BATCH', env, STATUS=ios)
      IF (ios == 0) READ(env, *) cfg%batch
      cfg%prec = 0
      CALL GET_ENVIRONMENT_VARIABLE('FMM3D_PRECISION', env, STATUS=ios)
      IF (ios == 0 .AND. TRIM(env) == 'fp32') cfg%prec = 1
      cfg%flags = 0
      IF (cfg%prec == 0) cfg%flags = 2   ! FMM3D_REFERENCE_PARTITION
      CALL GET_ENVIRONMENT_VARIABLE('FMM3D_DOMAIN', env, STATUS=ios)
      IF (ios == 0 .AND. TRIM(env) == 'whole') cfg%flags = 0
   END SUBROUTINE read_config

   ! the .vtx model (fm3dt_parallel.f90:6103-6126): header verbatim, values r fastest
   SUBROUTINE read_model(cfg, g, velv)
      TYPE(run_config), INTENT(IN) :: cfg
      TYPE(fmm3d_grid_desc), INTENT(OUT) :: g
      REAL(C_DOUBLE), ALLOCATABLE, INTENT(INOUT) :: velv(:)
      INTEGER :: i
      OPEN(UNIT=10, FILE=cfg%gridv, STATUS='old')
      READ(10, *) g%nvr, g%nvx, g%nvz
      READ(10, *) g%gor, g%goxd, g%gozd
      READ(10, *) g%dvr, g%dvxd, g%dvzd
      READ(10, *) g%nnr, g%nnx, g%nnz
      READ(10, *) g%rgor, g%rgoxd, g%rgozd
      READ(10, *) g%dnr, g%dnxd, g%dnzd
      g%earth = cfg%earth
      IF (ALLOCATED(velv)) DEALLOCATE(velv)
      ALLOCATE(velv((g%nvr + 2) * (g%nvx + 2) * (g%nvz + 2)))
      DO i = 1, SIZE(velv)
         READ(10, *) velv(i)
      END DO
      CLOSE(10)
   END SUBROUTINE read_model

   ! One forward pass with the model already on the device: every source of
   ! every receiver array, bottom faces from itimes.out, receiver times to
   ! rtravel.out and Frechet rows to frechet.out (fm3dt_parallel.f90:
   ! 6176-6600).
   SUBROUTINE fmm_pass(cfg, ctx, g)
      TYPE(run_config), INTENT(IN) :: cfg
      TYPE(C_PTR), INTENT(IN) :: ctx
      TYPE(fmm3d_grid_desc), INTENT(IN) :: g
      CHARACTER(LEN=8) :: pht
      INTEGER :: nra, idum, nsrc, nrc, i, j, k, ii, s0, nb, ios, itu
      INTEGER(C_INT64_T) :: cap, r, nrows
      REAL(C_DOUBLE) :: la, lo, dp
      REAL(C_DOUBLE), ALLOCATABLE :: faces(:), rcv(:)
      REAL(C_FLOAT), ALLOCATABLE :: tt(:), vals(:)
      INTEGER(C_INT64_T), ALLOCATABLE :: rp(:)
      INTEGER(C_INT32_T), ALLOCATABLE :: cols(:)
      INTEGER(4), ALLOCATABLE :: recs(:)
      INTEGER(8) :: face_rec, w
      OPEN(UNIT=60, FILE=cfg%sources, STATUS='old')
      READ(60, *) nra
      IF (cfg%fsrt == 1) THEN
         OPEN(UNIT=70, FILE=cfg%receivers, STATUS='old')
         READ(70, *) idum
         IF (idum /= nra) THEN
            WRITE(6, *) 'ERROR!!!'
            WRITE(6, *) 'Source and receiver files are'
            WRITE(6, *) 'inconsistent!!!!'
            WRITE(6, *) 'First line of each file should'
            WRITE(6, *) 'be identical!!!'
            WRITE(6, *) 'TERMINATING PROGRAM!!!'
            STOP
         END IF
      END IF
      ! frechet.out holds sequential unformatted records; they are assembled in
      ! memory (4-byte length, 8-byte payload, 4-byte length: the layout a
      ! sequential WRITE produces) and written a batch at a time through
      ! stream access, since one WRITE per entry dominates a pass otherwise
      IF (cfg%cfd == 1) OPEN(UNIT=50, FILE=cfg%frechet, ACCESS='stream', FORM='unformatted', STATUS='replace')
      OPEN(NEWUNIT=itu, FILE=cfg%itimes, ACCESS='stream', FORM='unformatted', STATUS='old')
      face_rec = 0

      DO ii = 1, nra
         READ(60, *) nsrc
         DO i = 1, nsrc                      ! source positions are read, not used (F:6275-6285)
            READ(60, *) la, lo, dp
            READ(60, '(a8)') pht
         END DO
         READ(70, *) nrc
         ALLOCATE(rcv(3 * nrc))
         DO i = 1, nrc                       ! (height, colat, lon) as F:6293-6312
            READ(70, *) rcv(3 * i - 2), la, lo
            rcv(3 * i - 1) = (90.0 - la) * pi / 180.0
            rcv(3 * i) = lo * pi / 180.0
         END DO
         IF (ii == 1) THEN
            OPEN(UNIT=20, FILE=cfg%rtravel, STATUS='replace')
         ELSE
            OPEN(UNIT=20, FILE=cfg%rtravel, POSITION='append')
         END IF
         DO s0 = 1, nsrc, cfg%batch
            nb = MIN(cfg%batch, nsrc - s0 + 1)
            ALLOCATE(faces(INT(g%nnx, 8) * g%nnz * nb), tt(nrc * nb))
            READ(itu, POS=1 + face_rec * 8 * INT(g%nnx, 8) * g%nnz) faces
            face_rec = face_rec + nb
            CALL check(fmm3d_solve_tele(ctx, INT(nb, C_INT), faces, INT(cfg%fom, C_INT), INT(cfg%prec, C_INT), &
                                        INT(cfg%flags, C_INT)), &
                       ctx, 'fmm3d_solve_tele')
            IF (cfg%fsrt == 1) THEN
               CALL check(fmm3d_receivers(ctx, INT(nrc, C_INT), rcv, tt), ctx, 'fmm3d_receivers')
               DO i = 1, nrc * nb
                  WRITE(20, *) tt(i)
               END DO
            END IF
            IF (cfg%cfd == 1) THEN
               nrows = INT(nrc, 8) * nb
               ALLOCATE(rp(nrows + 1))
               cap = nrows * 512
               DO k = 1, 2
                  ALLOCATE(cols(MAX(cap, 1_8)), vals(MAX(cap, 1_8)))
                  ios = fmm3d_frechet(ctx, INT(nrc, C_INT), rcv, rp, cols, vals, cap)
                  IF (ios /= -6) EXIT
                  cap = rp(nrows + 1)
                  DEALLOCATE(cols, vals)
               END DO
               CALL check(INT(ios, C_INT), ctx, 'fmm3d_frechet')
               ! per row a (count, receiver) record, then a (column, value) record per entry
               ALLOCATE(recs(4 * (nrows + rp(nrows + 1))))
               w = 0
               DO r = 1, nrows
                  j = INT(MOD(r - 1, INT(nrc, 8))) + 1
                  recs(w + 1) = 8
                  recs(w + 2) = INT(rp(r + 1) - rp(r), 4)
                  recs(w + 3) = TRANSFER(REAL(j, 4), 0_4)
                  recs(w + 4) = 8
                  w = w + 4
                  DO i = INT(rp(r)) + 1, INT(rp(r + 1))
                     recs(w + 1) = 8
                     recs(w + 2) = cols(i)
                     recs(w + 3) = TRANSFER(vals(i), 0_4)
                     recs(w + 4) = 8
                     w = w + 4
                  END DO
               END DO
               WRITE(50) recs(1:w)
               DEALLOCATE(rp, cols, vals, recs)
            END IF
            DEALLOCATE(faces, tt)
         END DO
         CLOSE(20)
         DEALLOCATE(rcv)
      END DO
      CLOSE(60)
      IF (cfg%fsrt == 1) CLOSE(70)
      IF (cfg%cfd == 1) CLOSE(50)
      CLOSE(itu)
   END SUBROUTINE fmm_pass
END MODULE fm3dt_run
