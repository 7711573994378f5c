/*
 * le_core.h -- C ABI of the MI355X split-step elasticity solver.
 *
 * Drop-in for the reference's simple/le_core.h: the same type names, field
 * order, storage constants and entry points, so a caller written against the
 * reference (simple/main.c) compiles and links against liblecore_d.so
 * (-DUSE_DOUBLE) or liblecore_f.so unchanged.  The fields live in HBM between
 * calls; `task->grid` is the host mirror (see le2d.h for the coherence rules).
 *
 * Layouts are ABI, checked by tests/test_abi.py:
 *   double: sizeof(le_task) 112, offsets dt 0, h 8, n 24, mat 32, stype 96,
 *           grid 104; le_material 64 B; le_node 40 B.
 *   float:  sizeof(le_task) 64, offsets 0, 4, 12, 20, 52, 56.
 */
#ifndef LE_CORE_H
#define LE_CORE_H

#include <sys/time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cycle counter and wall clock used by the reference's driver around its step
 * loop (reference simple/le_core.h:13-26). */
static __inline__ unsigned long getCC(void)
{
	unsigned lo, hi;
	__asm__ __volatile__("rdtsc" : "=a"(lo), "=d"(hi));
	return ((unsigned long)hi << 32) | lo;
}

static __inline__ double timer(void)
{
	struct timeval tv;
	gettimeofday(&tv, 0);
	return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* Storage of task->grid (reference simple/le_core.h:34-35). */
#define ST_SOA 0 /* five planes vx|vy|sxx|sxy|syy, each nx*ny, i fastest */
#define ST_AOS 1 /* nodes of five reals, node (i, j) at index i + j*nx   */

/* Arithmetic type, chosen at compile time as in the reference
 * (simple/le_core.h:37-61). */
#ifdef USE_DOUBLE
typedef double real;
#else
typedef float real;
#endif

typedef int int_t;

typedef struct { real x, y; } le_vec2;          /* spacing, centres          */
typedef struct { real xx, xy, yy; } le_smatrix2; /* symmetric 2-D stress      */
typedef struct { int_t x, y; } le_point2;        /* node counts               */

/* One grid node: velocity then stress (simple/le_core.h:156-159). */
typedef struct {
	le_vec2 v;
	le_smatrix2 s;
} le_node;

/* The four Riemann invariants with non-zero speed (simple/le_core.h:164-166). */
typedef struct { real w1, w2, w3, w4; } le_w;

/* Homogeneous isotropic material and the constants the sweeps use
 * (simple/le_core.h:175-186): irhoc1 = 1/(c1 rho), irhoc2 = 1/(c2 rho),
 * rhoc1 = c1 rho, rhoc2 = c2 rho, rhoc3 = rho c1 lambda / (lambda + 2 mu). */
typedef struct {
	real c1, c2, rho;
	real irhoc1, irhoc2, rhoc1, rhoc2, rhoc3;
} le_material;

/* Problem description and host mirror of the fields (simple/le_core.h:191-212). */
typedef struct {
	real dt;
	le_vec2 h;
	le_point2 n;
	le_material mat;
	int stype;
	real *grid;
} le_task;

/* replaces simple/le_core.h:217 (defined simple/le_core.c:171-187) */
void le_init_material(const real c1, const real c2, const real rho, le_material *m);

/* replaces simple/le_core.h:220 (simple/le_core.c:114-123).  Allocates the
 * zeroed host mirror; device memory is taken at the first step. */
void le_init_task(le_task *task, const real dt, const le_vec2 h, const le_material mat,
                  const le_point2 n, const int stype);

/* replaces simple/le_core.h:223 (simple/le_core.c:125-128).  Releases the host
 * mirror and the device state. */
void le_free_task(le_task *task);

/* replaces simple/le_core.h:228 (simple/le_core.c:71-94): sxx = syy = s at
 * nodes strictly inside the disk of radius r around c. */
void le_set_ball(le_task *t, const le_vec2 c, const real r, const real s);

/* replaces simple/le_core.h:237 (simple/le_core.c:130-169): legacy VTK of |v|.
 * Returns 0 on success, 1 if the file cannot be opened. */
int le_save_task(le_task *task, const char *file);

/* replaces simple/le_core.h:242 (simple/le_core.c:605-620): one x-sweep then
 * one y-sweep, on the GPU. */
void le_step(le_task *task);

/* replaces simple/le_core.h:248, which the reference declares but never
 * defines: le_step_x followed by le_step_y_cf(task, cfs). */
void le_step_cf(le_task *task, const int cfs);

/* replaces simple/le_core.h:254 (simple/le_core.c:623-627): the step on SOA
 * storage, including that path's y-edge (rows past the end read a stale slot
 * of its 5-row ring instead of the last row). */
void le_step_soa(le_task *task);

/* Single sweeps, linkable in the reference though not declared in its header
 * (simple/le_core.c:435, :503, :540, :367, :259). */
void le_step_x(le_task *task);
void le_step_y(le_task *task);
void le_step_y_cf(le_task *task, const int_t cfs);
void le_soa_step_x(le_task *task);
void le_soa_step_y(le_task *task);

#ifdef __cplusplus
}
#endif

#endif /* LE_CORE_H */
