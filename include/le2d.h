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

/* When le_step returns.  WAIT (default): after the step has finished on the
 * GPU.  ASYNC: after the launches are queued.  COHERENT: after the step has
 * finished and task->grid holds the result.  The environment variable
 * LE2D_SYNC=wait|async|coherent sets the default for new tasks. */
#define LE2D_SYNC_WAIT 0
#define LE2D_SYNC_ASYNC 1
#define LE2D_SYNC_COHERENT 2

const char *le2d_last_error(void);

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

#ifdef __cplusplus
}
#endif

#endif /* LE2D_H */
