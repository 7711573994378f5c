/*
 * le2d.h -- extensions of the le_core.h ABI for a device-resident solver.
 *
 * None of these has a counterpart in the reference's simple/le_core.h; they
 * exist because the fields now live in HBM between calls.  Include after (or
 * instead of) le_core.h; `real` follows USE_DOUBLE as there.
 *
 * Coherence of task->grid (the host mirror):
 *   - le_init_task zeroes the mirror; the first step uploads it.
 *   - le_set_ball, le_save_task, le_free_task and le_sync_host bring the
 *     mirror up to date before touching it.
 *   - After writing task->grid directly, call le_mark_host_dirty so the next
 *     step uploads it; before reading task->grid directly, call le_sync_host.
 *   - In LE2D_SYNC_COHERENT mode every step ends with a download, so the
 *     mirror is always current (at PCIe cost).
 *
 * Return values: 0 on success, otherwise one of LE2D_ERR_*; le2d_last_error()
 * describes the last failure on the calling thread.
 */
#ifndef LE2D_H
#define LE2D_H

#include <stddef.h>

#include "le_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LE2D_OK 0
#define LE2D_ERR_ARG 1
#define LE2D_ERR_DEVICE 2
#define LE2D_ERR_NOMEM 3
#define LE2D_ERR_STATE 4
#define LE2D_ERR_BUDGET 5
#define LE2D_ERR_BUSY 6

/* When le_step returns.  WAIT (default): after the step has finished on the
 * GPU.  ASYNC: after the launches are queued.  COHERENT: after the step has
 * finished and task->grid holds the result.  The environment variable
 * LE2D_SYNC=wait|async|coherent sets the default for new tasks. */
#define LE2D_SYNC_WAIT 0
#define LE2D_SYNC_ASYNC 1
#define LE2D_SYNC_COHERENT 2

const char *le2d_last_error(void);

/* The flux limiter this library was built with (make LIMITER=...): the
 * reference chooses it at compile time too (simple/le_core.c:36-43, `#define
 * limiter limiter_superbee`), and superbee is its choice and the default. */
#define LE2D_LIMITER_SUPERBEE 0
#define LE2D_LIMITER_MINMOD 1
#define LE2D_LIMITER_CIR 2
int le2d_limiter(void);

/* Device memory.  Every device allocation the library makes (state buffers,
 * batch and server arenas, tile and material tables, reduction partials)
 * goes through `alloc`, and is returned through `release` of the allocator
 * that made it, whatever is installed by then.  alloc returns 0 and a device
 * pointer in *ptr, or nonzero to refuse (the library call then fails with
 * LE2D_ERR_NOMEM and leaves its object as it was).  Both NULL: hipMalloc /
 * hipFree (the default).  Each of liblecore_d.so and liblecore_f.so keeps
 * its own setting.  No counterpart in the reference, which mallocs host
 * memory and does not check the result (simple/le_core.c:118-121). */
typedef int (*le2d_alloc_fn)(void **ptr, size_t bytes, void *ctx);
typedef void (*le2d_free_fn)(void *ptr, void *ctx);
int le2d_set_device_allocator(le2d_alloc_fn alloc, le2d_free_fn release, void *ctx);

int le2d_set_sync_mode(le_task *task, int mode);
/* Block until all queued work of the task has finished. */
int le2d_wait(le_task *task);
/* Copy the device state into task->grid if the device is newer. */
int le_sync_host(le_task *task);
/* Declare that task->grid was written directly; the next step uploads it. */
int le_mark_host_dirty(le_task *task);
/* `steps` full steps (x then y), one completion wait at the end. */
int le_step_n(le_task *task, int steps);

/* Launch shape: rows per x-sweep workgroup, rows per y-sweep segment,
 * columns per y-sweep workgroup (64, 128, 256), rows per segment of the fused
 * step, lanes per fused-step workgroup (64, 128, 256; it writes that many
 * columns less four).  Zero keeps the current value.  Results do not depend
 * on the shape. */
int le2d_set_tiles(le_task *task, int x_rows, int y_seg, int y_block, int xy_seg, int xy_block);
int le2d_get_tiles(le_task *task, int *x_rows, int *y_seg, int *y_block, int *xy_seg,
                   int *xy_block);

/* Two full steps per pass: with seg > 0, le_step_n advances clamped-edge
 * tasks (with or without a material plane) two steps per kernel pass while
 * at least two remain, `seg` rows per segment; each pass reads and writes
 * every node once for two steps.  Bit-identical to single steps.  seg = 0:
 * one step per pass.  le_step always takes one step. */
int le2d_set_step_pairs(le_task *task, int seg);
int le2d_get_step_pairs(le_task *task, int *seg);
/* Lanes per two-step workgroup, 64, 128 or 256 (it writes that many columns
 * less eight).  The default is 64 lanes (one wavefront), or 256 for FP32 grids
 * of about 8192^2 and more. */
int le2d_set_step_pair_window(le_task *task, int lanes);
int le2d_get_step_pair_window(le_task *task, int *lanes);
/* With 64-lane windows: on = 1 runs the two stages of a pass on two
 * wavefronts of a 128-lane workgroup (each holds one stage's state, so more
 * are resident), 0 on one wavefront.  Bit-identical either way; the tuner
 * chooses between them with the window width. */
int le2d_set_step_pair_split(le_task *task, int on);
int le2d_get_step_pair_split(le_task *task, int *on);
/* Time the two-step pass at segment heights 16 .. 160 (4 .. 160 on grids too
 * small to fill the device with 16-row segments) and windows of 64 (one or
 * two wavefronts, le2d_set_step_pair_split), 128 and 256 lanes on the task's
 * grid and keep the fastest.  The state does not
 * advance: each candidate reads the current state and writes the idle
 * buffer.  le_step_n does this once per
 * grid shape and process on its first call of 8 or more steps, unless the
 * caller has set a height (le2d_set_step_pairs), a material plane is set or
 * the environment has LE2D_TUNE=0.  A task with a material plane keeps the
 * default window and height; with 64-lane windows its first such call times
 * the one- and two-wavefront forms instead and keeps the faster. */
int le2d_tune_step_pairs(le_task *task, int *seg);
/* Run `passes` two-step passes, timing each launch with HIP events; mean ms
 * of one pass (two steps).  Needs seg > 0 and an ST_AOS task. */
int le2d_time_pairs(le_task *task, int passes, double *ms_pass);

/* Run `steps` steps as separate x- and y-sweep launches, timing each launch
 * with HIP events on the task's stream; mean ms of one x- and one y-sweep. */
int le2d_time_sweeps(le_task *task, int steps, double *ms_x, double *ms_y);
/* Run `steps` fused steps (the path le_step takes), timing each launch; mean
 * ms of one step. */
int le2d_time_steps(le_task *task, int steps, double *ms_step);

/* Streaming device-to-device copy of `bytes`, `iters` times; mean ms per
 * copy.  Measures the achievable HBM bandwidth on this device. */
int le2d_copy_probe(size_t bytes, int iters, double *ms);

/* Device bytes held by the task (both state buffers). */
size_t le2d_device_bytes(le_task *task);

/* Raw restart file of the full state: a 128-byte header (magic "LE2DRAW1",
 * sizeof(real), nx, ny, stype, then dt, h and the eight material constants as
 * doubles) and the five fields as planes of nx*ny little-endian reals.
 * le2d_restore needs a task of the same size and real type; it also restores
 * dt, h and the material.  A dump/restore round trip is bit-exact. */
int le2d_dump(le_task *task, const char *path);
int le2d_restore(le_task *task, const char *path);

/* Heterogeneous media: `count` materials (c1, c2, rho triples, doubles) and
 * one material index per node (ids[i + j*nx] < count; NULL = all material 0).
 * From then on le_step, le_step_n, le_step_soa and le_step_cf update every
 * node entirely in its own material: the stencil of a node is projected on
 * that node's eigenvectors, advected with its Courant numbers (dt*c/h) and
 * projected back with its constants.  The reference has one material per
 * task and defines no interface rule; with all ids equal this is the
 * reference scheme exactly (bit-identical in FP64).  The constants are formed
 * in FP64 and rounded once to `real`; task->mat is not used for stepping
 * while a plane is set.  count = 0 removes the plane.  Single sweeps
 * (le_step_x, le_step_y, ...) are refused while a plane is set. */
int le2d_set_material_plane(le_task *task, int count, const double *c1_c2_rho,
                            const unsigned char *ids);
/* Total elastic energy in FP64: the sum over nodes of
 * rho/2 |v|^2 + (sxx+syy)^2 / (8 (lambda+mu)) + ((sxx-syy)^2/2 + 2 sxy^2) / (4 mu),
 * times h.x*h.y, with mu = rho c2^2 and lambda = rho c1^2 - 2 mu of each
 * node's material (task->mat without a plane). */
int le2d_energy(le_task *task, double *e);

/* A read-only view of a device state: the current buffer, element size (4 or
 * 8), grid size, row pitch and plane stride in elements.  Valid until the
 * next step or free of its owner.  Views of tasks from liblecore_d.so and
 * liblecore_f.so (and of batch instances) can be compared with each other. */
typedef struct {
  const void *base;
  int elem, nx, ny;
  long pitch, plane;
} le2d_view;

int le2d_task_view(le_task *task, le2d_view *v);
/* out = {||a - ref||_2, ||ref||_2, max|a - ref|, max|ref|} over all five
 * fields, accumulated in FP64 on the device. */
int le2d_diff_norms(const le2d_view *a, const le2d_view *ref, double out[4]);
/* le2d_diff_norms without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under graph capture
 * set-up, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-workgroup partials live in a fixed 32 KiB device array of the library
 * (one per GPU, held until the process ends) instead of a buffer allocated
 * per call; its size does not depend on the grid, so any grid is accepted.
 * Results are bit-identical to le2d_diff_norms.  Calls that overlap on one
 * GPU (from several threads) share that array and race: serialise them. */
int le2d_diff_norms_noalloc(const le2d_view *a, const le2d_view *ref, double out[4]);

/* ---------------------------------------------------------------------------
 * Batched instances: many independent grids of uneven sizes and homogeneous
 * materials, advanced together by one kernel launch per step or per two
 * steps (replayed from a hipGraph).  No counterpart in the reference, which steps
 * one task at a time.  Element type per batch: real_size 4 (FP32 storage and
 * arithmetic, constants formed in FP64 and rounded once) or 8 (FP64,
 * bit-identical to stepping each grid alone through le_step).  Edges clamp
 * as in le_step.  Typical use:
 *   b = le2d_batch_create(4); le2d_batch_add(b, ...) per instance;
 *   le2d_batch_commit(b); [le2d_batch_autotune(b, ...)];
 *   le2d_batch_set_ball / le2d_batch_load per instance;
 *   le2d_batch_step(b, steps); le2d_batch_read(b, i, host) per instance. */
typedef struct le2d_batch le2d_batch;

le2d_batch *le2d_batch_create(int real_size);
void le2d_batch_destroy(le2d_batch *b);
/* Returns the instance index, or a negative error code. */
int le2d_batch_add(le2d_batch *b, int nx, int ny, double dt, double hx, double hy, double c1,
                   double c2, double rho);
int le2d_batch_count(le2d_batch *b);
/* Allocates every instance (zeroed fields) and builds the tile table. */
int le2d_batch_commit(le2d_batch *b);
size_t le2d_batch_device_bytes(le2d_batch *b);
/* Disk initial condition on the device, with le_set_ball's arithmetic. */
int le2d_batch_set_ball(le2d_batch *b, int inst, double cx, double cy, double r, double s);
/* AOS host arrays of nx*ny nodes of five values of the batch's real type. */
int le2d_batch_load(le2d_batch *b, int inst, const void *aos);
int le2d_batch_read(le2d_batch *b, int inst, void *aos);
/* `steps` steps of every instance: the batched launch (graph replay) or, for
 * comparison, the plain single-grid launches per instance that le_step_n
 * makes (two steps per launch while two remain, then one). */
int le2d_batch_step(le2d_batch *b, int steps);
int le2d_batch_step_plain(le2d_batch *b, int steps);
int le2d_batch_wait(le2d_batch *b);
/* Launch shape: window width 64/128/256 lanes and rows per segment (the
 * launch form unchanged); le2d_batch_set_launch also picks the form:
 * LE2D_BATCH_TILES, one workgroup per tile; LE2D_BATCH_PERSISTENT, resident
 * workgroups walk the tile table; LE2D_BATCH_PAIRS, one workgroup per tile
 * and two steps per launch (an odd step count ends with one single-step
 * launch); LE2D_BATCH_PAIRS_SPLIT, the same with 64-lane windows whose two
 * stages run on two wavefronts.  All forms give the same results. */
#define LE2D_BATCH_TILES 0
#define LE2D_BATCH_PERSISTENT 1
#define LE2D_BATCH_PAIRS 2
#define LE2D_BATCH_PAIRS_SPLIT 3
int le2d_batch_set_shape(le2d_batch *b, int fb, int seg);
int le2d_batch_set_launch(le2d_batch *b, int fb, int seg, int form);
int le2d_batch_get_launch(le2d_batch *b, int *form);
int le2d_batch_get_shape(le2d_batch *b, int *fb, int *seg, int *ntiles, int *blocks);
/* Mean ms per step of `steps` batched (plain = 0) or per-instance (plain = 1)
 * steps, HIP events on the batch's stream. */
int le2d_batch_time(le2d_batch *b, int steps, int plain, double *ms);
/* Times each window width x segment height on the batch and keeps the
 * fastest (remembered per element size, instance count and node total).
 * Advances the instances by an even number of steps. */
int le2d_batch_autotune(le2d_batch *b, int trials, int *best_fb, int *best_seg, double *best_ms);
/* View of instance `inst`'s current state (see le2d_view). */
int le2d_batch_view(le2d_batch *b, int inst, le2d_view *v);

/* ---------------------------------------------------------------------------
 * Serving loop: simulation requests admitted by HBM budget and slots, fields
 * resident on the device, the steps of all live requests batched into one
 * launch per precision per step as they arrive (continuous batching).  No
 * counterpart in the reference.  A request's result does not depend on the
 * other requests in flight (bit-identical alone and in any mix). */
typedef struct {
  int nx, ny;        /* grid size (>= 3) */
  int real_size;     /* 4: FP32 tier, 8: FP64 */
  int steps;         /* full steps to run */
  double dt, hx, hy; /* time step, spacing */
  double c1, c2, rho;   /* homogeneous material */
  double cx, cy, r, s;  /* disk source: sxx = syy = s where |x - c| < r */
  int keep_result;   /* 1: keep the final state until le2d_server_result */
} le2d_request;

typedef struct {
  long long completed, queued;
  int live, peak_live;
  double node_updates;   /* node-updates executed so far */
  double busy_seconds;   /* wall time the engine spent in chunks */
  size_t bytes_in_use, peak_bytes;
  long long cancelled;   /* requests cancelled (queued, live or kept results) */
  long long rejected;    /* submits refused with LE2D_ERR_BUSY */
} le2d_server_stats;

typedef struct le2d_server le2d_server;

/* hbm_budget: device bytes for request fields (one arena, reserved now);
 * max_live: concurrent requests; plain = 1: serial baseline (one slot, the
 * single-grid kernel). */
le2d_server *le2d_server_create(size_t hbm_budget, int max_live, int plain);
/* Launch shape of the batched steps and the longest run of steps between
 * admission points. */
int le2d_server_set_shape(le2d_server *srv, int fb, int seg, int max_chunk);
/* Largest rows per segment of the two-step launches (default 64; the server
 * picks the height for its live set, shorter when few requests are live);
 * 0: one step per launch.  Results do not change. */
int le2d_server_set_pairs(le2d_server *srv, int pair_seg);
/* Lanes per two-step workgroup, 64, 128 or 256; 0 (the default): 256 for
 * FP32 requests and 64 for FP64.  Results do not change. */
int le2d_server_set_pair_window(le2d_server *srv, int lanes);
/* With 64-lane two-step windows, on = 1 runs each pass's two stages on two
 * wavefronts (default 0).  Results do not change. */
int le2d_server_set_pair_split(le2d_server *srv, int on);
/* Admission.  max_queue: requests that may wait (0, the default: no limit);
 * at the limit le2d_server_submit returns -LE2D_ERR_BUSY and the caller
 * retries later (back-pressure).  Admission takes queued requests in
 * arrival order while slots are free; among the first `window` (default 64)
 * a request that does not fit the budget now is passed by later ones that
 * do, until it has been passed `max_pass` times (default 16; 0: strict
 * arrival order), after which admission waits for it. */
int le2d_server_set_admission(le2d_server *srv, int max_queue, int window, int max_pass);
/* Queue a request; returns its id (>= 0) or a negative error code
 * (-LE2D_ERR_BUDGET when it can never fit, -LE2D_ERR_BUSY when the queue is
 * at its limit). */
long long le2d_server_submit(le2d_server *srv, const le2d_request *req);
/* Cancel a request: a queued one leaves the queue at once; a live one stops
 * at the end of the current chunk of steps and its memory returns to the
 * budget; a finished one's kept result is discarded.  le2d_server_wait on a
 * cancelled request returns LE2D_ERR_STATE.  LE2D_ERR_STATE also when the
 * request has already been released or cancelled. */
int le2d_server_cancel(le2d_server *srv, long long id);
/* Block until request `id` has finished or was cancelled (timeout_s < 0: no
 * limit). */
int le2d_server_wait(le2d_server *srv, long long id, double timeout_s);
/* Block until every submitted request has finished. */
int le2d_server_drain(le2d_server *srv, double timeout_s);
/* Copy a kept result into an AOS host array (nx*ny nodes of five reals of
 * the request's type) and release its device memory.  Only one call per
 * result succeeds; a concurrent second call returns LE2D_ERR_STATE. */
int le2d_server_result(le2d_server *srv, long long id, void *aos);
/* Submission, admission and completion times (seconds, steady clock). */
int le2d_server_times(le2d_server *srv, long long id, double *t_submit, double *t_admit,
                      double *t_done);
int le2d_server_get_stats(le2d_server *srv, le2d_server_stats *out);
/* Stops the engine after its current chunk; requests still queued or live
 * are abandoned.  No other call on `srv` may be in progress or follow. */
void le2d_server_destroy(le2d_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* LE2D_H */
