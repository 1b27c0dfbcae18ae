!-----------------------------------------------------------------------
! fm3dt_invert -- the inversion loop of sub.sh (sub.sh:1-40) with the
! forward solver resident on one GPU.  sub.sh starts a new 9-rank MPI job
! for every forward pass; here the run configuration, the grid and the
! device context (its source slots, tuned kernel choices and captured
! graphs) are set up once, and each pass re-reads only the model subinv
! rewrote and re-grids it on the device.  The files and their order are
! sub.sh's:
!
!   cp gridi.vtx gridc.vtx; subiter.in <- 1
!   forward pass; cp rtravel.out rtimes.dat
!   N times: subinv; forward pass; subiter.in <- iteration + 1
!
! Usage: fm3dt_invert [iterations [subinv command]]   (defaults: 5, ./subinv)
! The forward pass reads the model fm3dt_parallel.in names (gridc.vtx in
! sub.sh's layout); subinv may change its values, not its header.  subinv
! runs as a child process; a non-zero exit status stops the loop.
! Environment: as fm3dt_gpu (FMM3D_DEVICE, FMM3D_PRECISION, FMM3D_BATCH).
!-----------------------------------------------------------------------
PROGRAM fm3dt_invert
   USE ISO_C_BINDING
   USE fmm3d_c
   USE fm3dt_run
   IMPLICIT NONE
   TYPE(run_config) :: cfg
   TYPE(fmm3d_grid_desc) :: g, g2
   TYPE(C_PTR) :: ctx
   REAL(C_DOUBLE), ALLOCATABLE :: velv(:)
   CHARACTER(LEN=512) :: arg, subinv
   INTEGER :: niter, it, es, cs

   niter = 5
   subinv = './subinv'
   IF (COMMAND_ARGUMENT_COUNT() >= 1) THEN
      CALL GET_COMMAND_ARGUMENT(1, arg)
      READ(arg, *) niter
   END IF
   IF (COMMAND_ARGUMENT_COUNT() >= 2) CALL GET_COMMAND_ARGUMENT(2, subinv)

   CALL copy_file('gridi.vtx', 'gridc.vtx')
   CALL write_iter(1)
   CALL read_config(cfg)
   CALL read_model(cfg, g, velv)
   ctx = C_NULL_PTR
   CALL check(fmm3d_create(g, INT(cfg%dev, C_INT), ctx), C_NULL_PTR, 'fmm3d_create')
   CALL check(fmm3d_set_model_vtx(ctx, velv), ctx, 'fmm3d_set_model_vtx')
   WRITE(6, '(a)') 'Running Parallel FMM...'
   CALL fmm_pass(cfg, ctx, g)
   WRITE(6, '(a)') 'Parallel FMM Done'
   CALL copy_file(cfg%rtravel, 'rtimes.dat')
   DO it = 1, niter
      WRITE(6, '(a)') 'Running Inversion...'
      CALL EXECUTE_COMMAND_LINE(TRIM(subinv), WAIT=.TRUE., EXITSTAT=es, CMDSTAT=cs)
      IF (cs /= 0 .OR. es /= 0) THEN
         WRITE(6, *) 'ERROR: ', TRIM(subinv), ' failed (status ', cs, es, ')'
         CALL check(fmm3d_destroy(ctx), C_NULL_PTR, 'fmm3d_destroy')
         STOP 1
      END IF
      WRITE(6, '(a)') 'Inversion Done'
      CALL read_model(cfg, g2, velv)
      IF (g2%nvr /= g%nvr .OR. g2%nvx /= g%nvx .OR. g2%nvz /= g%nvz .OR. g2%nnr /= g%nnr .OR. &
          g2%nnx /= g%nnx .OR. g2%nnz /= g%nnz .OR. g2%gor /= g%gor .OR. g2%goxd /= g%goxd .OR. &
          g2%gozd /= g%gozd .OR. g2%dvr /= g%dvr .OR. g2%dvxd /= g%dvxd .OR. g2%dvzd /= g%dvzd .OR. &
          g2%rgor /= g%rgor .OR. g2%rgoxd /= g%rgoxd .OR. g2%rgozd /= g%rgozd .OR. g2%dnr /= g%dnr .OR. &
          g2%dnxd /= g%dnxd .OR. g2%dnzd /= g%dnzd) THEN
         WRITE(6, *) 'ERROR: the model header changed between passes'
         CALL check(fmm3d_destroy(ctx), C_NULL_PTR, 'fmm3d_destroy')
         STOP 1
      END IF
      CALL check(fmm3d_set_model_vtx(ctx, velv), ctx, 'fmm3d_set_model_vtx')
      WRITE(6, '(a)') 'Running Parallel FMM...'
      CALL fmm_pass(cfg, ctx, g)
      WRITE(6, '(a)') 'Parallel FMM Done'
      CALL write_iter(it + 1)
   END DO
   CALL check(fmm3d_destroy(ctx), C_NULL_PTR, 'fmm3d_destroy')

CONTAINS
   ! a byte-for-byte copy (sub.sh's cp)
   SUBROUTINE copy_file(a, b)
      CHARACTER(LEN=*), INTENT(IN) :: a, b
      CHARACTER(LEN=1), ALLOCATABLE :: buf(:)
      INTEGER :: ua, ub
      INTEGER(8) :: n
      OPEN(NEWUNIT=ua, FILE=TRIM(a), ACCESS='stream', FORM='unformatted', STATUS='old')
      INQUIRE(UNIT=ua, SIZE=n)
      ALLOCATE(buf(MAX(n, 1_8)))
      IF (n > 0) READ(ua) buf(1:n)
      CLOSE(ua)
      OPEN(NEWUNIT=ub, FILE=TRIM(b), ACCESS='stream', FORM='unformatted', STATUS='replace')
      IF (n > 0) WRITE(ub) buf(1:n)
      CLOSE(ub)
   END SUBROUTINE copy_file

   ! subiter.in holds the iteration count subinv reads (sub.sh's print -u4 $COUNT)
   SUBROUTINE write_iter(n)
      INTEGER, INTENT(IN) :: n
      INTEGER :: u
      OPEN(NEWUNIT=u, FILE='subiter.in', STATUS='replace')
      WRITE(u, '(I0)') n
      CLOSE(u)
   END SUBROUTINE write_iter
END PROGRAM fm3dt_invert
