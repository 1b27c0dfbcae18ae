!-----------------------------------------------------------------------
! fm3dt_gpu -- drop-in replacement for `mpirun -np 9 ./fm3dt_parallel`.
!
! Reads the same files from the current directory (fm3dt_parallel.in and
! the files it names: the .vtx model, sources.dat, receivers.dat and the
! binary itimes.out) and writes the same outputs (rtravel.out, one
! list-directed REAL(4) per line; frechet.out, unformatted sequential
! records (column, value) after a (count, receiver) header per row), as
! PROGRAM fmmin3d does (fm3dt_parallel.f90:6015-6662).  The solve, the
! receiver interpolation and the rays/Frechet rows run on one GPU through
! the C ABI of libfmm3d_hip.so (include/fmm3d.h); the file handling is in
! fm3dt_core.f90.
!
! Environment: FMM3D_DEVICE (default 0), FMM3D_PRECISION=fp32 for the
! float32 solver (default float64, the reference's arithmetic),
! FMM3D_BATCH (sources per device call, default 64).
!-----------------------------------------------------------------------
PROGRAM fm3dt_gpu
   USE ISO_C_BINDING
   USE fmm3d_c
   USE fm3dt_run
   IMPLICIT NONE
   TYPE(run_config) :: cfg
   TYPE(fmm3d_grid_desc) :: g
   TYPE(C_PTR) :: ctx
   REAL(C_DOUBLE), ALLOCATABLE :: velv(:)

   CALL read_config(cfg)
   CALL read_model(cfg, g, velv)
   ctx = C_NULL_PTR
   CALL check(fmm3d_create(g, INT(cfg%dev, C_INT), ctx), C_NULL_PTR, 'fmm3d_create')
   CALL check(fmm3d_set_model_vtx(ctx, velv), ctx, 'fmm3d_set_model_vtx')
   CALL fmm_pass(cfg, ctx, g)
   CALL check(fmm3d_destroy(ctx), C_NULL_PTR, 'fmm3d_destroy')
END PROGRAM fm3dt_gpu
