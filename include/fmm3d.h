/*
 * fmm3d.h -- C ABI of libfmm3d_hip.so, the MI355X forward solver for the
 * fm3dt_parallel file contract (3-D spherical eikonal solve, B-spline
 * gridder, receiver times, rays with Frechet rows).
 *
 * The reference has no library interface: its boundary is a process and a
 * set of files (sub.sh:24,33 runs `mpirun -np 9 ./fm3dt_parallel` in a
 * directory holding fm3dt_parallel.in, the .vtx model, sources.dat,
 * receivers.dat and itimes.out; it writes rtravel.out and frechet.out).
 * Each entry point below replaces one stage of that program; the comment
 * names the reference code it stands for.  Plain pointers and sizes only;
 * host buffers are caller-owned; device memory is owned by the context.
 *
 * Conventions
 *   - Grid descriptors carry the six .vtx header lines verbatim (degrees);
 *     the library converts with the reference's float32-rounded pi
 *     (fm3dt_parallel.f90:61, 6110-6117).
 *   - Host field arrays are (nnr, nnx, nnz) with phi fastest: the memory
 *     order of the reference's ttn(iz,ix,ir).
 *   - Vertex arrays are (nvz+2, nvx+2, nvr+2) with r fastest: the .vtx
 *     record order.  Frechet columns are the 1-based r-fastest vertex index.
 *   - Receivers are (nrc, 3) rows of (height km, colatitude rad, longitude
 *     rad), converted as fm3dt_parallel.f90:6310-6311.
 *   - Every call returns FMM3D_OK (0) or a negative FMM3D_E_* code and never
 *     aborts; fmm3d_last_error() gives the message.  A context is not
 *     re-entrant: use one per host thread.
 */
#ifndef FMM3D_H
#define FMM3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMM3D_OK            0
#define FMM3D_E_ARG        -1   /* bad argument or inconsistent sizes       */
#define FMM3D_E_DEVICE     -2   /* HIP runtime / device error               */
#define FMM3D_E_NOMEM      -3   /* device allocation failed                 */
#define FMM3D_E_NOCONV     -4   /* solver hit its iteration cap             */
#define FMM3D_E_STATE      -5   /* call out of order (e.g. no model yet)    */
#define FMM3D_E_CAPACITY   -6   /* caller's output buffer too small         */
#define FMM3D_E_TIMEOUT    -7   /* server: the request is not done yet      */

/* solver options */
#define FMM3D_ORDER_FIRST   0   /* first-order stencils (.in line 9 = 0)    */
#define FMM3D_ORDER_MIXED   1   /* mixed first/second order (.in line 9 = 1) */
#define FMM3D_PREC_FP64     0   /* float64 node updates                     */
#define FMM3D_PREC_FP32     1   /* float32 updates on times offset by the source's boundary minimum */
#define FMM3D_WALK_ALTERNATE 1  /* flag: column walk alternates bottom-up / top-down (point sources) */
#define FMM3D_REFERENCE_PARTITION 2  /* fmm3d_solve_tele flag: the reference's 2x2x2 partitioned solve */

typedef struct {
    int nvr, nvx, nvz;          /* vertex counts, border excluded          */
    double gor, goxd, gozd;     /* vertex origin: height km, lat deg (N edge), lon deg (W edge) */
    double dvr, dvxd, dvzd;     /* vertex spacing: km, deg, deg            */
    int nnr, nnx, nnz;          /* propagation nodes                       */
    double rgor, rgoxd, rgozd;  /* node origin (same units)                */
    double dnr, dnxd, dnzd;     /* node spacing                            */
    double earth;               /* earth radius km (.in line 8)            */
} fmm3d_grid_desc;

typedef struct {
    int64_t node_updates;       /* node updates in the last solve call     */
    int32_t sweeps;             /* relaxation sweeps (max over sources)    */
    int32_t tile_r, tile_x, tile_z;  /* tile shape used                    */
    double solve_ms;            /* device time of the last solve (events)  */
    double kernel_ms;           /* device time inside the relaxation kernel */
    int64_t kernel_launches;    /* relaxation-kernel launches              */
    int32_t phase_sweeps[6];    /* sweeps per phase: A, B, C, C, C, -      */
    int64_t brick_visits;       /* bricks / patch columns relaxed (summed over sweeps) */
    int64_t node_visits;        /* column walk: free nodes staged (summed over sweeps) */
} fmm3d_stats;

typedef struct fmm3d_ctx fmm3d_ctx;

/* Create a context on HIP device `device` for one grid.  Replaces the
 * grid read and decomposition of fm3dt_parallel.f90:6103-6171. */
int fmm3d_create(const fmm3d_grid_desc *grid, int device, fmm3d_ctx **out);

/* Release the context and all device memory it owns. */
int fmm3d_destroy(fmm3d_ctx *ctx);

/* Message of the last failed call on `ctx` (or of the last failed create
 * when ctx is NULL).  Valid until the next call on the same context. */
const char *fmm3d_last_error(const fmm3d_ctx *ctx);

/* Upload a vertex model and run the cubic B-spline gridder on the device
 * (gridder, fm3dt_parallel.f90:6989-7038).  velv: (nvz+2)(nvx+2)(nvr+2)
 * float64, r fastest. */
int fmm3d_set_model_vtx(fmm3d_ctx *ctx, const double *velv);

/* Teleseismic solves, one per source, with bottom-face injection: node
 * values > 0 are fixed, the rest unknown (start_fmm + local_fmm + the
 * outer relaxation, fm3dt_parallel.f90:6371-6448, 6894-6961, 78-155).
 * faces: nsrc records of nnx*nnz float64, phi fastest (itimes.out).
 * order: FMM3D_ORDER_*; precision: FMM3D_PREC_*.  The nsrc fields stay
 * resident in the context (replacing those of the previous solve) for the
 * receiver and Frechet calls.
 * flags: 0 solves the upwind system on the whole grid as one domain (the
 * field a single-rank heap march gives); FMM3D_REFERENCE_PARTITION
 * (float64 only) reproduces the reference program's own result: its 2x2x2
 * blocks with stencils ending at block faces, face relaxation from two-deep
 * ghosts, the per-block march from the reclassified state, and passes until
 * the largest change is at most the float32 literal 1e-5
 * (fm3dt_parallel.f90:157-416, 708-710, 1100-2825, 6127-6171, 6397-6435). */
int fmm3d_solve_tele(fmm3d_ctx *ctx, int nsrc, const double *faces, int order,
                     int precision, int flags);

/* Teleseismic solves with the plane wave injected on the bottom face and on
 * one theta and one phi side face per source (the faces it enters through),
 * all from 1-D boundary times.  No counterpart in the reference, which
 * reads bottom faces only (itimes.out, fm3dt_parallel.f90:6894-6961); the
 * discrete system and the stop rule are those of fmm3d_solve_tele.
 *   faces: nsrc bottom faces, nnx*nnz float64, phi fastest (itimes.out);
 *   xface: nsrc theta faces, nnr*nnz float64 (ir = 0 at the top), phi fastest;
 *   zface: nsrc phi faces, nnr*nnx float64, theta fastest;
 *   side:  2*nsrc ints, per source (theta side, phi side): -1 none, 0 the
 *          face at index 0, 1 the face at index nnx-1 (nnz-1).
 * Values > 0 are fixed, the rest unknown, on every face.  flags: 0. */
int fmm3d_solve_tele_sides(fmm3d_ctx *ctx, int nsrc, const double *faces, const double *xface,
                           const double *zface, const int *side, int order, int precision, int flags);

/* The same solves in two steps, for callers that solve the same sources
 * more than once (several models of an inversion, repeated precisions) or
 * keep their inputs resident: fmm3d_stage_tele_sides copies the faces (as
 * fmm3d_solve_tele_sides takes them; xface = NULL for bottom faces only)
 * into device memory the context owns and fits each source's plane-wave
 * reference and lateral flow there and then; fmm3d_solve_tele_staged
 * solves the staged sources, with the fields and results of
 * fmm3d_solve_tele_sides on the same faces.  No counterpart in the
 * reference (its faces come from itimes.out on every run,
 * fm3dt_parallel.f90:6894-6961).  A solve of any kind on the context
 * stages its own sources in place of them; the partitioned solve leaves
 * none staged (FMM3D_E_STATE). */
int fmm3d_stage_tele_sides(fmm3d_ctx *ctx, int nsrc, const double *faces, const double *xface,
                           const double *zface, const int *side);
int fmm3d_solve_tele_staged(fmm3d_ctx *ctx, int order, int precision);

/* Solves from general fixed (Dirichlet) node sets, one per source: the
 * generalisation of the bottom-face injection above to any nodes, e.g.
 * bottom and side faces of a teleseismic plane wave or the refined
 * neighbourhood of a point source.  No counterpart in the reference, which
 * injects bottom faces only (fm3dt_parallel.f90:6894-6961); the node update
 * and the relaxed discrete system are the same.  Source s owns entries
 * offs[s] .. offs[s+1]-1 of node (0-based index (ir*nnx + ix)*nnz + iz)
 * and time (seconds, finite, >= 0).  flags: 0, or FMM3D_WALK_ALTERNATE
 * when waves travel downward too (point sources). */
int fmm3d_solve_fixed(fmm3d_ctx *ctx, int nsrc, const int64_t *offs, const int64_t *node,
                      const double *time, int order, int precision, int flags);

/* Point (local) sources with a refined source neighbourhood; no
 * counterpart in the reference (teleseismic only).  src: nsrc records of
 * (height km, colatitude rad, longitude rad).  Around each source the
 * coarse cell holding it, widened by `radius` nodes per side, is refined by
 * `refine` per axis; the sub-grid gets B-spline velocities, straight-ray
 * times on the 8 corners of the fine cell holding the source and a float64
 * solve; the coarse nodes it covers then seed the coarse solve
 * (FMM3D_WALK_ALTERNATE).  Up to "point_workers" (default 4) refined
 * blocks are solved at once, largest first, on worker threads with their
 * own streams; their contexts stay with ctx, pooled by block shape.
 * Returns FMM3D_E_ARG for a source outside the grid.  flags: 0. */
int fmm3d_solve_point(fmm3d_ctx *ctx, int nsrc, const double *src, int refine, int radius,
                      int order, int precision, int flags);

/* One batch of teleseismic and point sources relaxed together: slots
 * 0 .. ntele-1 hold the teleseismic sources (faces, xface, zface, side as
 * for fmm3d_solve_tele_sides; xface null for bottom faces only) and slots
 * ntele .. ntele+npoint-1 the point sources (src as for fmm3d_solve_point,
 * source q refined by `refine` over radius[q] nodes).  Teleseismic slots
 * walk upward in flow order, point slots alternate their walk, and every
 * slot leaves the sweeps once a sweep changed none of its nodes; the
 * result per slot is that of the single-kind calls.  No counterpart in the
 * reference (one teleseismic source at a time, fm3dt_parallel.f90:6371).
 * flags: 0. */
int fmm3d_solve_mixed(fmm3d_ctx *ctx, int ntele, const double *faces, const double *xface,
                      const double *zface, const int *side, int npoint, const double *src,
                      int refine, const int *radius, int order, int precision, int flags);

/* Receiver times for every resident source (srtimes,
 * fm3dt_parallel.f90:7040-7248): t_out is nsrc*nrc float32, source-major.
 * A receiver outside the grid yields NaN. */
int fmm3d_receivers(fmm3d_ctx *ctx, int nrc, const double *rcv, float *t_out);

/* fmm3d_receivers without any device allocation: the receiver coordinates
 * and the times are staged in a fixed device array that the library holds
 * on each device it runs on until the process ends (found at context
 * creation), so nrc*3*8 bytes (rounded up to 256) + nsrc*nrc*4 bytes must
 * not exceed 8 MiB, else FMM3D_E_ARG.  Same results as fmm3d_receivers.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every call on a device shares that
 * array, so calls that overlap on one GPU (two host threads, two contexts)
 * race; serialise them. */
int fmm3d_receivers_noalloc(fmm3d_ctx *ctx, int nrc, const double *rcv, float *t_out);

/* Rays and Frechet rows for every resident source and receiver
 * (parallel_rpaths + the writer's row assembly, fm3dt_parallel.f90:
 * 7250-7777, 6503-6599).  Rows are source-major then receiver order;
 * row_ptr has nsrc*nrc+1 entries; a receiver outside the grid gets an
 * empty row.  Entries with |value| >= 1e-6 are kept, columns ascending.
 * Returns FMM3D_E_CAPACITY (with row_ptr[nsrc*nrc] = entries needed) if
 * cap is too small. */
int fmm3d_frechet(fmm3d_ctx *ctx, int nrc, const double *rcv, int64_t *row_ptr,
                  int32_t *cols, float *vals, int64_t cap);

/* fmm3d_frechet with the rows in pinned host arrays owned by the context:
 * on success *row_ptr (nsrc*nrc+1 entries), *cols and *vals (*nnz entries)
 * point into them and stay valid until the next fmm3d_frechet_pinned call on
 * ctx or fmm3d_destroy.  The device-to-host copies then go straight into the
 * result; no counterpart in the reference (a performance variant of the row
 * assembly above). */
int fmm3d_frechet_pinned(fmm3d_ctx *ctx, int nrc, const double *rcv, const int64_t **row_ptr,
                         const int32_t **cols, const float **vals, int64_t *nnz);

/* Copy resident field `src` to the host as float64 (nnr, nnx, nnz), phi
 * fastest (the reference's ttn after the final exchange, F:6439-6448). */
int fmm3d_get_field(fmm3d_ctx *ctx, int src, double *ttn);

/* Outer passes per source of the last FMM3D_REFERENCE_PARTITION solve of
 * nsrc sources (the reference's `count`, F:6398-6430); FMM3D_E_STATE after
 * any other solve. */
int fmm3d_get_passes(const fmm3d_ctx *ctx, int nsrc, int *passes);

/* Statistics of the last solve. */
int fmm3d_get_stats(const fmm3d_ctx *ctx, fmm3d_stats *out);

/* Solver tuning knobs (no counterpart in the reference, which has none):
 *   "brick_layers"  layers per brick (>= 4)
 *   "patch_x", "patch_z"  lateral patch of a brick (supported pairs: 8x32,
 *                   4x32, 4x16, 8x16, 16x16, 4x64, 2x32, 4x8)
 *   "lateral_iters" lateral relaxations per layer
 *   "tol_fp64", "tol_fp32"  relative change below which a sweep counts as
 *                   settled
 *   "direction"     0: sweep directions alternate; 1: per brick, along time
 *   "strategy"      0: brick sweeps in three phases; 1: column walk (default);
 *                   2: hyperplane walk (8 x 8 patches, lane (i, j) relaxes
 *                   layer s - i - j at step s: one Gauss-Seidel sweep in flow
 *                   order per launch, one update per node) for batches of
 *                   teleseismic sources; batches holding point sources and
 *                   the serving loop use the column walk
 *   "plane_chunk"   hyperplane walk: steps per staging chunk (2, 4 or 8; default 4)
 *   "plane_sweeps"  hyperplane walk: sweeps before the column walk takes over
 *                   (0, the default: hyperplane sweeps throughout)
 *   "column_x", "column_z"  column-walk patch (8, 16, 32 or 64 each; supported
 *                   pairs 8x8, 8x16, 16x8, 16x16, 8x32, 32x8, 16x32, 32x16,
 *                   32x32, 16x64, 64x16, 8x64, 64x8)
 *   "column_passes" line passes per layer at most (1..16, default 16)
 *   "ordered"       column walk: 1 (default) patches start each layer after
 *                   their upstream neighbours (along the source's lateral
 *                   flow) finished it; 0 no ordering
 *   "ordered_sweeps" column walk: only sweeps up to this one wait on their
 *                   upstream patches (default 0: every sweep)
 *   "point_order"   sources whose walk alternates (point sources, and
 *                   fmm3d_solve_fixed with FMM3D_WALK_ALTERNATE): 1 (default)
 *                   their patches start outward from the one holding the
 *                   earliest fixed node and wait on their neighbour towards
 *                   it, as ordered sources wait along their flow; 0 no
 *                   ordering.  Refined-block solves follow the setting
 *   "point_workers" threads solving refined blocks at once (1..16, default 4)
 *   "point_group_mnodes" refined blocks solved together as one group: at
 *                   most this many million nodes of the group's grid
 *                   (1..1024, default 128)
 *   "point_window"  alternating float64 walks, and the float32 warm-up of
 *                   refined blocks: 1 (default) a patch walks only the
 *                   layers within two of one that changed, in it or a
 *                   lateral neighbour, in the sweep before or so far in
 *                   this one, and on while its own walk changes layers;
 *                   0 whole columns
 *   "point_warm"    refined blocks: 1 (default) a float32 solve first, then
 *                   float64 sweeps from its field to the float64 fixed
 *                   point; 0 float64 from the start
 *   "spin_limit"    polls before such a wait is given up (default 20000)
 *   "check_every"   column walk: sweeps between convergence checks (1..8,
 *                   default 4); at a check, every source whose last sweep
 *                   changed nothing leaves the remaining sweeps
 *   "graph"         1 (default): replay each chunk of check_every sweeps
 *                   from a captured hipGraph (batch solves; the serving
 *                   engine launches directly)
 *   "autotune"      1 (default): the first column-walk solve of a kind
 *                   (device, grid, precision, order, walk mode, batch size
 *                   class) takes its configuration from the tuning cache, or
 *                   times candidates on the real batch -- the current
 *                   configuration, patch shapes 8x8/8x16/16x8/16x16 with 16
 *                   passes and 8x8 with 8, a 2-sweep check interval, the
 *                   hyperplane walk -- and keeps one only when at least 5 %
 *                   faster; the choice goes to the cache (FMM3D_TUNE_CACHE,
 *                   else $XDG_CACHE_HOME/fmm3d/tune.txt, else
 *                   ~/.cache/fmm3d/tune.txt; FMM3D_TUNE_CACHE="" for none).
 *                   Batches under 8 sources or grids under 2 M nodes keep
 *                   the current configuration untimed, and so does a context
 *                   whose strategy, column_x/column_z, column_passes or
 *                   check_every the caller set (stats.tile_x/tile_z/tile_r
 *                   report the configuration used); 0: off
 *   "max_sweeps"    column walk: stop after this many sweeps (0 = at
 *                   convergence); a diagnostic, the field is then approximate
 *   "timing"        1: time every relaxation launch with events (stats.kernel_ms);
 *                   the sweeps are then launched directly, not from a graph
 *   "verify"        column walk: after this sweep (default 1; 0 never) one
 *                   read-only pass evaluates every node's update on the field
 *                   as it stands: a source with no node that would move is
 *                   at its fixed point and leaves the sweeps
 *   "verify_patches" 1: the check also narrows the next sweep to the patches
 *                   with such a node and their neighbours (teleseismic,
 *                   flow-ordered sources); 0: not; -1 (default): in float32
 *                   only
 *   "verify_filter" 1 (default): the check after a source's first sweep
 *                   evaluates only nodes that an input relaxed after them
 *                   (the layers walked later, lateral neighbours outside
 *                   the patch not waited for) could move, i.e. where such
 *                   an input is not above the node; 0 every node
 *                   (FMM3D_VERIFY_FILTER)
 *   "tele_reference" float32 teleseismic solves: 1 (default) store each
 *                   node against the source's plane-wave reference; 0 the
 *                   plain offset from the boundary minimum (FMM3D_TELE_REF)
 *   "tol_floor"     fields with a reference: a change counts for the sweep
 *                   stop above tol_fp32 * max(|value|, tol_floor) (default
 *                   64 s; FMM3D_TOL_FLOOR)
 *   "ref_qbits"     fields with a reference: updated values rounded to
 *                   multiples of 2^-ref_qbits s (20..30, default 22; 0 not;
 *                   FMM3D_REF_QBITS)
 *   "pass_tol"      fields with a reference: a layer's line passes stop
 *                   after one that moves no node by more than this many
 *                   seconds, and the fixed-point check counts only such
 *                   moves (0..1e-3, default 5e-7; 0 exact; FMM3D_PASS_TOL)
 *   "layer_guess"   fields with a reference: a node the walk reaches for
 *                   the first time starts its layer's passes from the stored
 *                   values of the nodes walked just before it in its column
 *                   (the reference carries the vertical travel time, so they
 *                   are close), extrapolated by the polynomial through the
 *                   last n of them: 4 (default, the cubic), 3, 2 (the
 *                   line), 1 (the node before), 0 from infinity
 *                   (FMM3D_LAYER_GUESS)
 * Returns FMM3D_E_ARG for an unknown key or a value out of range. */
int fmm3d_set_option(fmm3d_ctx *ctx, const char *key, double value);

/* Velocity-perturbation slice of a vertex model, on the device: the
 * reference's slicer (gmtplot/gmtslicet.f90:4-540).  The vertex model velv
 * minus velv_ref (or velv itself when velv_ref is NULL), both
 * (nvz+2)(nvx+2)(nvr+2) r fastest, is B-splined onto the propagation grid
 * of `grid` and one plane of it is sampled with a second B-spline diced
 * dice1 x dice2 per cell:
 *   kind 0: depth slice at height `where` km, dicing (theta, phi);
 *   kind 1: N-S slice at longitude `where` deg, dicing (r, theta);
 *   kind 2: E-W slice at latitude `where` deg, dicing (r, phi).
 * n1 x n2 values go to out in the order gmtslicet writes them (outer loop
 * over the second axis, inner over the first from its far end);
 * FMM3D_E_CAPACITY (with *n1, *n2 set) when cap is smaller.  FMM3D_E_ARG
 * when the slice or a B-spline tap falls outside the grids (gmtslicet
 * reads out of bounds there).  No context: the call owns its device memory
 * and stream for its duration.  fmm3d_slice_last_error() gives the message
 * of the calling thread's last failed slice. */
int fmm3d_slice_velocity(const fmm3d_grid_desc *grid, int device, const double *velv, const double *velv_ref,
                         int kind, double where, int dice1, int dice2, double *out, int64_t cap, int *n1, int *n2);
const char *fmm3d_slice_last_error(void);

/* The Frechet matrix G on the device and the sparse products of subinv's
 * subspace step (subinv.f90:612-620 G^T d; 759-790 and 829-850 G a / Cd,
 * then G^T of it).  subinv itself stays on the CPU; these are for an
 * inversion loop that keeps G resident.  G: nrows rows (traveltimes) in
 * CSR, row_ptr[0..nrows], 1-based columns (vertices, frechet.out's) below
 * ncols and ascending within a row, float32 values.  The transpose is
 * built on the device with the rows of each column ascending, as subinv
 * builds it (subinv.f90:457-475).  Products sum in float32 entry by entry
 * in subinv's order, so they are subinv's values:
 *   fmm3d_gmat_gtx: y[ncols] = G^T x[nrows]
 *   fmm3d_gmat_gx:  y[nrows] = (G x[ncols]) / cd[nrows] (cd may be NULL: 1)
 * fmm3d_gmat_last_error(NULL) gives the message of the calling thread's
 * last failed fmm3d_gmat_create. */
typedef struct fmm3d_gmat fmm3d_gmat;
int fmm3d_gmat_create(int device, int64_t nrows, int64_t ncols, const int64_t *row_ptr, const int32_t *cols,
                      const float *vals, fmm3d_gmat **out);
int fmm3d_gmat_gtx(fmm3d_gmat *g, const float *x, float *y);
int fmm3d_gmat_gx(fmm3d_gmat *g, const float *x, const float *cd, float *y);
const char *fmm3d_gmat_last_error(const fmm3d_gmat *g);
int fmm3d_gmat_destroy(fmm3d_gmat *g);

/* ---------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, which runs one source after
 * another per process, fm3dt_parallel.f90:6371).  A server owns one device
 * and a scheduler thread.  Requests -- sources on a grid and velocity model,
 * with receivers -- are copied at submission.  Models stay resident (one
 * context per grid, model id and precision), and requests are admitted by
 * HBM footprint: a new model's context gets as many source slots as the
 * budget leaves.  Sources join a running batch one slot at a time between
 * sweep chunks (continuous batching), converged sources leave it, and
 * contexts of different grids sweep concurrently on their own streams.
 * ------------------------------------------------------------------------- */
#define FMM3D_SRC_TELE      0   /* bottom faces (itimes.out records)        */
#define FMM3D_SRC_POINT     1   /* point sources with a refined neighbourhood */

typedef struct fmm3d_server fmm3d_server;

typedef struct {
    fmm3d_grid_desc grid;
    const double *velv;         /* vertex model, (nvz+2)(nvx+2)(nvr+2), r fastest */
    uint64_t model_id;          /* requests with equal grid, model id and precision share one resident model */
    int kind;                   /* FMM3D_SRC_TELE or FMM3D_SRC_POINT */
    int nsrc;
    const double *faces;        /* TELE: nsrc records of nnx*nnz, phi fastest */
    const double *src;          /* POINT: nsrc records (height km, colatitude rad, longitude rad) */
    int refine, radius;         /* POINT: refinement factor and radius in nodes */
    int nrc;
    const double *rcv;          /* nrc records (height km, colatitude rad, longitude rad) */
    int order, precision;       /* FMM3D_ORDER_*, FMM3D_PREC_* */
    int rows;                   /* 1: also trace the rays and keep the Frechet rows (fmm3d_server_rows) */
} fmm3d_request;

typedef struct {
    int64_t requests_done;
    int64_t sources_done;
    double p50_ms, p99_ms;      /* request latency, submission to completion */
    double source_solves_per_s; /* over the span from the first submission to the last completion */
    double hbm_bytes;           /* device bytes held by resident models and slots */
    int models_resident;
    int sources_in_flight;
} fmm3d_server_info;

/* hbm_budget: device bytes the server may hold; max_slots: source slots per
 * model context at most. */
int fmm3d_server_create(int device, double hbm_budget, int max_slots, fmm3d_server **out);
/* Copies the request; *ticket names it for wait/result. */
int fmm3d_server_submit(fmm3d_server *srv, const fmm3d_request *req, int64_t *ticket);
/* FMM3D_OK once done (or failed: result reports it), FMM3D_E_TIMEOUT if not
 * within timeout_s seconds (0: poll). */
int fmm3d_server_wait(fmm3d_server *srv, int64_t ticket, double timeout_s);
/* Frechet rows of a finished request submitted with rows = 1, in the layout
 * of fmm3d_frechet (source-major then receiver order; row_ptr has
 * nsrc*nrc+1 entries; columns ascending): the reference's rows for every
 * source and receiver (parallel_rpaths + the writer's merge,
 * fm3dt_parallel.f90:7250-7777, 6503-6599).  FMM3D_E_CAPACITY with
 * row_ptr[nsrc*nrc] = entries needed when cap is too small; FMM3D_E_STATE
 * for a request without rows.  Call before fmm3d_server_result, which
 * releases the ticket. */
int fmm3d_server_rows(fmm3d_server *srv, int64_t ticket, int64_t *row_ptr, int32_t *cols, float *vals, int64_t cap);
/* Receiver times (nsrc*nrc float32, source-major) and the request's latency;
 * releases the ticket.  Returns the request's status. */
int fmm3d_server_result(fmm3d_server *srv, int64_t ticket, float *times, double *latency_ms);
int fmm3d_server_stats(fmm3d_server *srv, fmm3d_server_info *out);
/* Drains nothing: pending requests are dropped, the thread joins. */
int fmm3d_server_destroy(fmm3d_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* FMM3D_H */
