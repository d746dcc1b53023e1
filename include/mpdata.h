/* Device-resident MPDATA interface (extensions beside the drop-in smolar.h).
 *
 * No counterpart in the reference except where noted.  A grid handle keeps the face
 * Courant numbers, the field and all scratch in HBM; steps are enqueued on the
 * library's stream for the current device and run without host round trips.
 *
 * Boundary modes:
 *   MPD_FRAME     the reference's semantics (core/src/smolar.c): the outer one-cell
 *                 frame of the grid is held at its initial values and only the interior
 *                 is advanced; arrays are nx*ny*nz including the frame.
 *   MPD_PERIODIC  every axis wraps; every cell is advanced.  The donor and
 *                 antidiffusive formulas are the reference's interior formulas.
 * Options: iord >= 1 passes per step (1 = donor cell); nonoscillatory limiter
 * (Smolarkiewicz & Grabowski 1990) on the corrective passes, PERIODIC only; a
 * per-step emission increment q added after the last pass (c <- A(c) + q), PERIODIC
 * only.
 *
 * Arrays are FP64, cell index k*nx*ny + j*nx + i.  PERIODIC face arrays use the same
 * layout (U[i,j,k] is the x-face between cells i and i+1 mod nx).  FRAME face arrays
 * use the reference's layouts (uu.c:122, vv.c:133, ww.c:299).
 *
 * Every function returns MPD_OK (0) or an MPD_ERR_* code; mpd_error_string() gives the
 * calling thread's last message.  Nothing aborts the process.
 * Link: -lsmolar_hip.
 */
#pragma once
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPD_OK 0
#define MPD_ERR_ARG 1         /* invalid argument: sizes, NULL pointers, options */
#define MPD_ERR_DEVICE 2      /* HIP runtime failure or no usable gfx950 device */
#define MPD_ERR_NOMEM 3       /* device or host allocation failed */
#define MPD_ERR_UNSUPPORTED 4 /* option combination not provided */
#define MPD_ERR_NUMERIC 5     /* a health check found a non-finite or negative value */

#define MPD_FRAME 0
#define MPD_PERIODIC 1

#define MPD_ARITH_EXACT 0 /* unfused FP64, IEEE division: bit-identical to the reference order */
#define MPD_ARITH_FAST 1  /* contracted multiply-adds, reciprocal division: <= 1e-12 per step */

const char* mpd_error_string(void);
/* Number of visible HIP devices (0 without a GPU); MPD_OK even when 0. */
int mpd_device_count(int* n);

/* ---------------------------------------------------------------- single grid --- */
typedef struct mpd_grid_s* mpd_grid;

int mpd_create(mpd_grid* g, int nx, int ny, int nz, int boundary, int iord, int nonosc);
int mpd_destroy(mpd_grid g);
/* Face Courant numbers from cell-centred winds at two time levels, spacings and dt,
 * computed on the device (uu.c:132-134, vv.c:141-145, ww.c:305-311).  Host arrays. */
int mpd_set_winds(mpd_grid g, const double* dx, const double* dy, const double* dz,
                  double dt, const double* u0, const double* u1, const double* v0,
                  const double* v1, const double* w0, const double* w1);
/* Face Courant numbers given directly (host arrays in the boundary mode's layout). */
int mpd_set_courant(mpd_grid g, const double* U, const double* V, const double* W);
int mpd_get_courant(mpd_grid g, double* U, double* V, double* W);
int mpd_set_field(mpd_grid g, const double* c);
int mpd_get_field(mpd_grid g, double* c);
/* FRAME mode: the previous pass's field of the last step (the reference's *c0 on
 * return from smolar()). */
int mpd_get_previous_field(mpd_grid g, double* c);
/* Per-step emission increment (host array, nx*ny*nz), or NULL to clear. */
int mpd_set_emission(mpd_grid g, const double* q);
/* The emission increment set by mpd_set_emission or mpd_generate_blobs (host array,
 * nx*ny*nz); MPD_ERR_ARG when none is set. */
int mpd_get_emission(mpd_grid g, double* q);
/* Enqueue nsteps steps; returns without waiting. */
int mpd_step(mpd_grid g, int nsteps);
int mpd_synchronize(mpd_grid g);
/* Device address of the current field (valid until the next mpd_step). */
int mpd_field_device_ptr(mpd_grid g, double** dptr);
/* Runs nsteps steps bracketed by HIP events on the launch stream; *ms = elapsed time. */
int mpd_step_timed(mpd_grid g, int nsteps, double* ms);
/* Runs nsteps steps recording an event after every kernel launch (PERIODIC grids);
 * ms[i] = average duration of the i-th kernel of a step (i < max), *n = kernels per
 * step. */
int mpd_step_kernel_times(mpd_grid g, int nsteps, double* ms, int max, int* n);
/* Device copy of n doubles timed over reps launches (*ms per launch): the plain
 * streaming kernel against which memory-traffic counters are calibrated. */
int mpd_calibration_copy(size_t n, int reps, double* ms);
/* Mass (sum of c), minimum and maximum of the field, reduced on the device: the
 * failure sentinel for long runs (MPD_ERR_NUMERIC if any value is non-finite). */
int mpd_field_stats(mpd_grid g, double* mass, double* cmin, double* cmax);
/* Step sentinel for long runs, checked on the device without host synchronisation:
 * after every `every` steps of mpd_step (0 = off, the default) the field is scanned for
 * non-finite values (and negative ones when nonnegative is set); the first step count,
 * since this call, after which one was seen is kept.  Arming resets it.  (Steps are then
 * launched one by one: no graph replay.) */
int mpd_set_sentinel(mpd_grid g, int every, int nonnegative);
/* Waits for the grid's work; MPD_OK with *first_bad_step = -1 while the sentinel saw
 * nothing, else MPD_ERR_NUMERIC with the step count (first_bad_step may be NULL). */
int mpd_sentinel_check(mpd_grid g, long long* first_bad_step);
/* Kernel-path selection.  PERIODIC grids: 0 = automatic (fused), 1 = staged (one
 * kernel per stage, intermediates in HBM), 2 = fused (two kernels per corrective pass:
 * antidiffusive velocities, then limiter and limited donor pass), 3 = one kernel per
 * corrective pass (nonoscillatory grids; the donor field and the antidiffusive
 * velocities stay on chip), 4 = its single-plane form (nonoscillatory grids of one
 * z-plane; other grids take 2); 0 picks 4 for nonoscillatory grids of one z-plane, 3 for
 * those of at least 8 z-planes and 2 otherwise; grids with more than 2^29 cells per
 * z-plane always take the staged path.  FRAME grids: 1 = the per-face
 * antidiffusive kernel and a separate donor pass, otherwise the tiled kernel that forms
 * the donor field and the velocities together (results are identical). */
int mpd_set_path(mpd_grid g, int path);
/* Arithmetic of the PERIODIC fused kernels: MPD_ARITH_FAST (default) or
 * MPD_ARITH_EXACT.  FRAME grids and the staged path always compute exactly. */
int mpd_set_arith(mpd_grid g, int arith);
/* Tile shape of the fused kernels (0,0 = the default 32x8); one of the compiled shapes
 * 32x8, 16x16, 64x8, 16x8, 32x16.  Sets it for both kernels of a pass (the limiter
 * kernel's own shape, if any, is cleared). */
int mpd_set_tile(mpd_grid g, int tile_x, int tile_y);
/* Tile shape and z-chunk of the limiter kernel alone (the second fused kernel of each
 * nonoscillatory pass); 0,0,0 = those of the first kernel (default); zc 0 = chosen from
 * the grid size.  The results do not depend on the launch shapes. */
int mpd_set_limiter_tile(mpd_grid g, int tile_x, int tile_y, int zc);
/* The launch settings: v[0..2] tile_x, tile_y, zc of the first kernel, v[3..5] those of
 * the limiter kernel (0 = default / same as the first kernel). */
int mpd_get_tiles(mpd_grid g, int* v);
/* Whether mpd_step replays blocks of 8 steps as captured HIP graphs (PERIODIC grids on
 * the fused path; default off: on MI355X the replays measure no faster than launching
 * the kernels, whose gaps are the GPU's own dispatch latency).  The results are
 * identical either way. */
int mpd_set_graph(mpd_grid g, int on);
/* Planes per workgroup of the fused kernels (0 = chosen from the grid size). */
int mpd_set_zchunk(mpd_grid g, int zc);
/* Planes per workgroup of the one-kernel-per-pass kernel (0 = chosen from the grid
 * size: the longest chunk that still gives every CU a workgroup). */
int mpd_set_pass_zchunk(mpd_grid g, int zc);
/* The z-chunk the one-kernel-per-pass kernel uses (the set value or the chosen one). */
int mpd_get_pass_zchunk(mpd_grid g, int* zc);
/* The kernel path the grid's steps take: 1 staged (or donor only), 2 two kernels per
 * corrective pass, 3 one kernel per corrective pass (automatic for nonoscillatory
 * PERIODIC grids of at least 8 z-planes), 4 its single-plane form (automatic for those
 * of one z-plane).  FRAME grids: 1 per-face kernels, 2 tiled. */
int mpd_get_path(mpd_grid g, int* path);
/* Times every tile shape x z-chunk length on the grid itself (one warm-up step and nsteps
 * timed steps each, each kernel between events of its own), then times whole steps of the
 * leading pairings (the two fastest shapes of each of the two fused kernels, and the
 * fastest shape for both together; three interleaved rounds) and keeps the fastest.  It
 * reports the first kernel's shape (mpd_get_tiles gives both).  The field buffers are
 * restored, and the trial steps neither capture graphs nor count for the step sentinel,
 * so the grid's state is unchanged.  Results are cached per process by shape and
 * options, and in the file named by the environment variable MPD_TUNE_CACHE when set.
 * PERIODIC grids with iord >= 2 on the fused path only (MPD_ERR_UNSUPPORTED otherwise);
 * a grid whose steps take one kernel per corrective pass (mpd_get_path 3) has nothing to
 * tune: it returns MPD_OK with the two-kernel settings unchanged. */
int mpd_autotune(mpd_grid g, int nsteps, int* tile_x, int* tile_y, int* zc);

/* One PERIODIC step c_out = A(c_in) (+ q) on caller-owned DEVICE arrays that allocates
 * no device memory: no counterpart in the reference.  For callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It computes what mpd_step computes on a PERIODIC grid with the same Courant numbers
 * U, V, W (periodic face layout), iord (2..64), limiter option and emission q (NULL:
 * none), in fast arithmetic on the fused path, and writes c_out (distinct from c_in).
 * The fused path's scratch is a fixed device array of 256 MiB per device, held until the
 * process ends: at most 8 Mi cells at iord 2, 4 Mi cells at iord >= 3 (MPD_ERR_ARG
 * beyond); z-planes of at most 2^29 cells.  Kernels are enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) and the call returns without waiting.  The first
 * call on a device looks up that device's scratch address.  Every call on a device
 * shares its scratch array: calls that overlap on one device race. */
int mpd_step_periodic_noalloc(int nx, int ny, int nz, int iord, int nonosc, const double* U,
                              const double* V, const double* W, const double* q,
                              const double* c_in, double* c_out, void* stream);

/* ------------------------------------------------------------ generators --- */
#define MPD_MAX_MODES 8
#define MPD_MAX_BLOBS 64
#define MPD_TARGET_FIELD 0
#define MPD_TARGET_EMISSION 1

/* Cell-centred wind omega*(axis x (x - centre)) + curl(A) with
 * A_d(x) = sum_m amp[d][m] sin(2 pi (k[d][m][0] x/nx + k[d][m][1] y/ny + k[d][m][2] z/nz)
 *                              + phase[d][m]),
 * cell centres at integer coordinates.  Face Courant numbers follow with dx = dt = 1 and
 * are scaled so that max |face Courant number| = cmax. */
typedef struct {
    double axis[3], centre[3], omega;
    int nmodes;
    double k[3][MPD_MAX_MODES][3];
    double amp[3][MPD_MAX_MODES], phase[3][MPD_MAX_MODES];
    double cmax;
} mpd_wind_spec;
int mpd_generate_wind(mpd_grid g, const mpd_wind_spec* w);
/* Sum of nblobs Gaussians amp[q] exp(-|x - centre_q|^2 / (2 sigma[q]^2)) (centres as
 * x,y,z triples; periodic = 1 adds the 26 neighbouring periodic images) written to the
 * field (MPD_TARGET_FIELD) or the emission increment (MPD_TARGET_EMISSION). */
int mpd_generate_blobs(mpd_grid g, int target, int nblobs, const double* centres,
                       const double* sigma, const double* amp, int periodic);

/* ------------------------------------------------------------- adjoint --- */
/* PERIODIC grids with iord = 2 (limiter on or off); the face Courant numbers are fixed
 * parameters.  One step is c_out = A(c_in) + q.
 *
 * One adjoint step on host arrays: lin = A'(c)^T lout, and lq += lout when lq is not
 * NULL.  Every kink follows the branch the forward step takes. */
int mpd_adjoint_step(mpd_grid g, const double* c, const double* lout, double* lin, double* lq);
/* How the adjoint sums the antidiffusive-velocity contributions: 0 (default) gathers
 * them in an LDS-tiled kernel -- deterministic, bit-reproducible; 1 scatters with
 * hardware FP64 atomics -- last bits vary run to run. */
int mpd_set_adjoint_scatter(mpd_grid g, int atomic);
/* Kernels of the adjoint's limiter stages: 1 (default) one LDS-tiled kernel, 0 two
 * per-cell passes through device arrays (the same arithmetic, for cross-checks). */
int mpd_set_adjoint_tiled(mpd_grid g, int tiled);
/* Device memory mpd_adjoint_run may use for keeping forward records (each step's state
 * and intermediates, 60 bytes per cell and step with the limiter, 40 without) of whole
 * checkpoint segments from its forward sweep, beyond the one segment record it always
 * holds: the segments kept this way are not recomputed in the backward sweep.  0
 * (default): only the last segment is kept; -1: as many segments as the free device
 * memory holds, less an 8 GiB reserve.  The gradient is the same bit for bit. */
int mpd_set_adjoint_tape(mpd_grid g, long long max_bytes);
/* The tape budget set above and the number of segment records the last mpd_adjoint_run
 * with gradients kept from its forward sweep (either pointer may be NULL). */
int mpd_get_adjoint_tape(mpd_grid g, long long* max_bytes, int* kept_segments);
/* Forward run of nsteps steps from the grid's field recording the receptor values every
 * sample_every steps: out[(s-1)*nrec + r] = c^{s*sample_every}[rec_r]. */
int mpd_sample_receptors(mpd_grid g, int nsteps, int nrec, const int* receptors,
                         int sample_every, double* out);
/* Receptor-misfit loss and gradient.  From the grid's field c^0 and its emission q (zero
 * if unset), c^{n+1} = A(c^n) + q for n < nsteps, and
 *   J = 1/2 sum_{s=1}^{nsteps/sample_every} sum_r (c^{s*sample_every}[rec_r] - obs[(s-1)*nrec + r])^2
 * (rec: distinct cell indices).  The forward keeps c^n every checkpoint_every steps and
 * the full record of the last segments (mpd_set_adjoint_tape); the backward sweep
 * recomputes each other segment from its checkpoint and applies the adjoint steps in
 * reverse.  grad_c0 = dJ/dc^0 and grad_q = dJ/dq (host arrays, may be NULL;
 * both NULL: loss only).  times (may be NULL, 3 doubles): forward and backward
 * milliseconds, and the part of the backward spent recomputing checkpoint segments.  The
 * forward runs in exact arithmetic; on return the grid holds c^nsteps. */
int mpd_adjoint_run(mpd_grid g, int nsteps, int checkpoint_every, int nrec,
                    const int* receptors, int sample_every, const double* obs, double* J,
                    double* grad_c0, double* grad_q, double* times);

/* ------------------------------------------------------------- ensembles --- */
typedef struct mpd_batch_s* mpd_batch;

#define MPD_BATCH_PER_INSTANCE 0 /* one launch sequence per instance (baseline) */
#define MPD_BATCH_FUSED 1        /* one launch per kernel stage for all instances */
#define MPD_BATCH_GRAPH 2        /* the MPD_BATCH_FUSED time loop replayed as a hipGraph */

/* ninst periodic instances; dims[3i..3i+2] = nx, ny, nz of instance i; iord >= 2. */
int mpd_batch_create(mpd_batch* b, int ninst, const int* dims, int iord, int nonosc);
int mpd_batch_destroy(mpd_batch b);
/* Handle of instance i for mpd_set_courant / mpd_set_winds / mpd_set_field /
 * mpd_get_field / mpd_set_emission (owned by the batch: never mpd_destroy it, and step
 * it only through mpd_batch_step). */
int mpd_batch_instance(mpd_batch b, int i, mpd_grid* g);
int mpd_batch_cells(mpd_batch b, long long* cells);
int mpd_batch_set_arith(mpd_batch b, int arith);
/* Tile shape, z-chunk and grid kind (persistent = 1: one resident wave of workgroups
 * striding over the work items; 0: one workgroup per item), for both kernels of a pass
 * (clears the limiter kernel's own shape). */
int mpd_batch_set_tile(mpd_batch b, int tile_x, int tile_y, int zc, int persistent);
/* Tile shape and z-chunk of the limiter kernel alone (0, 0, 0: those of the first
 * kernel); the results do not depend on it. */
int mpd_batch_set_limiter_tile(mpd_batch b, int tile_x, int tile_y, int zc);
/* v[0..3]: tile_x, tile_y, zc, persistent of the first kernel; v[4..6]: the limiter
 * kernel's tile_x, tile_y, zc (0 = the first kernel's). */
int mpd_batch_get_tiles(mpd_batch b, int* v);
/* Kernel path of the batched launches: 0 automatic (one kernel per corrective pass when
 * the ensemble has 2^22 cells or more, as for single grids), 2 two kernels per pass, 3
 * one kernel per pass.  Per-instance launches follow each instance's own path.  The
 * results agree to rounding (bit for bit in exact arithmetic). */
int mpd_batch_set_path(mpd_batch b, int path);
/* The path the batched launches take (2 or 3). */
int mpd_batch_get_path(mpd_batch b, int* path);
/* z-chunk of the one-kernel batched launches (0: 32 planes, at most each instance's
 * height). */
int mpd_batch_set_pass_zchunk(mpd_batch b, int zc);
/* Times every tile shape x z-chunk x grid kind over nsteps steps on the batch itself
 * (fields restored afterwards), applies and returns the fastest (any output may be
 * NULL).  A batch on one kernel per pass has one tile shape (32 x 16, persistent 0):
 * there the batched launches' z-chunk is timed (16 .. 128 planes, up to the tallest
 * instance) and returned in zc.  Cached per process; also in the file named by
 * MPD_TUNE_CACHE when set. */
int mpd_batch_autotune(mpd_batch b, int nsteps, int* tile_x, int* tile_y, int* zc,
                       int* persistent);
int mpd_batch_step(mpd_batch b, int nsteps, int mode);
/* As mpd_batch_step, bracketed by HIP events (table upload and graph capture happen
 * before the first event). */
int mpd_batch_step_timed(mpd_batch b, int nsteps, int mode, double* ms);

#ifdef __cplusplus
}
#endif
