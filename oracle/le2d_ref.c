/*
 * le2d_ref.c -- CPU oracle for the split-step elasticity sweeps.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker or the CPU baseline.
 *
 * It restates, node by node and in the reference's own operation order, the
 * arithmetic of the reference solver's AOS path (simple/le_core.c).  Each
 * function cites the lines it follows.  Compiled twice:
 *   -DORC_REAL=double  -> liboracle_d.so  (the FP64 parity oracle)
 *   -DORC_REAL=float   -> liboracle_f.so  (the reference's float build, whose
 *                          double literals promote parts of the arithmetic)
 * The build must not contract a*b+c into FMA (-ffp-contract=off) and must not
 * vectorise across nodes in a way that reorders sums (there are no sums here).
 *
 * Storage is addressed through (node stride, field stride) so the same code
 * walks the AOS grid (5 reals per node, simple/le_core.c:13-14) and the SOA
 * planes (simple/le_core.c:16-21).  Field order: vx, vy, sxx, sxy, syy
 * (simple/le_core.h:156-159).
 *
 * Parity pins (tests/test_oracle.py): golden vectors written by the compiled
 * reference (tests/golden/), closed forms for the material and the ball, the
 * textbook superbee values, the zero-slope fixed point, the c=1 exact shift,
 * conservation, y-independence and brute-force 1-D advection.  The
 * heterogeneous sweeps (orc_sweep_mat, orc_step_mat) are pinned by their
 * reduction to orc_run with one material or identical ones, by node-wise
 * equality with the homogeneous sweep in each node's own material, and by
 * the normal-incidence reflection and transmission coefficients of P and S
 * pulses at an interface.  The minmod and CIR builds (ORC_LIMITER, below;
 * tests/test_limiters.py) are pinned by the limiters' definitions, the
 * binomial spreading of a pulse under first-order upwind (CIR), minmod's
 * reduction to Lax-Wendroff and Beam-Warming on concave and convex data, and
 * the total-variation bound of every limiter.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real_t;

#define F_VX 0
#define F_VY 1
#define F_SXX 2
#define F_SXY 3
#define F_SYY 4

/* y-sweep edge handling.  ORC_EDGE_CLAMP: positions past the last row read the
 * last row (simple/le_core.c:521,534).  ORC_EDGE_RING: positions past the last
 * row read the stale slot of a 5-row ring that is no longer refilled, as the
 * SOA y-sweep (simple/le_core.c:310-349) and le_step_y_cf (:574-594) do. */
#define ORC_EDGE_CLAMP 0
#define ORC_EDGE_RING 1

/* Material constants, in the order of simple/le_core.h:175-186. */
typedef struct {
	real_t c1, c2, rho;
	real_t irhoc1, irhoc2, rhoc1, rhoc2, rhoc3;
} orc_material;

/* simple/le_core.c:171-187.  Literals stay double, as in the reference, so the
 * float build keeps the reference's promotions. */
void orc_material_init(real_t c1, real_t c2, real_t rho, orc_material *m)
{
	m->c1 = c1;
	m->c2 = c2;
	m->rho = rho;
	m->irhoc1 = 1.0 / (c1 * rho);
	m->irhoc2 = 1.0 / (c2 * rho);
	m->rhoc1 = c1 * rho;
	m->rhoc2 = c2 * rho;
	{
		real_t mu = rho * c2 * c2;
		real_t lam = rho * c1 * c1 - 2.0 * mu;
		m->rhoc3 = rho * c1 * lam / (lam + 2.0 * mu);
	}
}

/* Ternary min/max of simple/le_core.c:31-33 (not fmin/fmax: the NaN and
 * signed-zero behaviour differs). */
static inline real_t orc_min(real_t a, real_t b) { return a > b ? b : a; }
static inline real_t orc_max(real_t a, real_t b) { return a > b ? a : b; }

/* Superbee limiter, simple/le_core.c:38,43:
 * max(0, max(min(1, 2r), min(2, r))), with 2.0 a double literal. */
real_t orc_superbee(real_t r)
{
	return orc_max(0.0, orc_max(orc_min(1.0, 2.0 * r), orc_min(2.0, r)));
}

/* The limiter is a build-time choice, as in the reference (simple/le_core.c:
 * 36-43, `#define limiter limiter_superbee`): ORC_LIMITER 0 superbee (the
 * reference's and the default), 1 minmod max(0, min(1, r)) (:36), 2 CIR, the
 * constant 0.0 (:37). */
#ifndef ORC_LIMITER
#define ORC_LIMITER 0
#endif
int orc_limiter_kind(void) { return ORC_LIMITER; }

real_t orc_limiter(real_t r)
{
#if ORC_LIMITER == 1
	return orc_max(0.0, orc_min(1.0, r));
#elif ORC_LIMITER == 2
	(void)r;
	return 0.0;
#else
	return orc_superbee(r);
#endif
}

/* Second-order TVD update of one value, simple/le_core.c:48-69.
 * a = u[i-2], b = u[i-1], u = u[i], d = u[i+1] along the upwind direction. */
real_t orc_tvd2(real_t c, real_t a, real_t b, real_t u, real_t d)
{
	real_t num, den, r, rm, fp, fm;

	num = u - b;
	den = d - u;
	num += 1e-6;
	den += 1e-6;
	r = num / den;

	num = b - a;
	den = u - b;
	num += 1e-6;
	den += 1e-6;
	rm = num / den;

	fp = u + orc_limiter(r) / 2.0 * (1.0 - c) * (d - u);
	fm = b + orc_limiter(rm) / 2.0 * (1.0 - c) * (u - b);
	return u - c * (fp - fm);
}

/* Four Riemann invariants of one node for a sweep along axis `ax`
 * (0 = x, simple/le_core.c:192-204; 1 = y, :206-218).  The along-axis
 * velocity / normal stress pair and the tangential velocity / shear pair. */
static void orc_invariants(const orc_material *m, const real_t *node, long fs, int ax,
                           real_t w[4])
{
	const real_t vn = node[(ax == 0 ? F_VX : F_VY) * fs];
	const real_t sn = node[(ax == 0 ? F_SXX : F_SYY) * fs] * m->irhoc1;
	const real_t vt = node[(ax == 0 ? F_VY : F_VX) * fs];
	const real_t st = node[F_SXY * fs] * m->irhoc2;
	w[0] = vn - sn;
	w[1] = vn + sn;
	w[2] = vt - st;
	w[3] = vt + st;
}

/* Back-transform of the increments, simple/le_core.c:221-249. */
static void orc_apply(const orc_material *m, real_t *node, long fs, int ax, const real_t d[4])
{
	const real_t e1 = 0.5 * d[0];
	const real_t e2 = 0.5 * d[1];
	const real_t e3 = 0.5 * d[2];
	const real_t e4 = 0.5 * d[3];
	const int vn = ax == 0 ? F_VX : F_VY, vt = ax == 0 ? F_VY : F_VX;
	const int sn = ax == 0 ? F_SXX : F_SYY, so = ax == 0 ? F_SYY : F_SXX;

	node[vn * fs] += e1 + e2;
	node[vt * fs] += e3 + e4;
	node[sn * fs] += (e2 - e1) * m->rhoc1;
	node[so * fs] += (e2 - e1) * m->rhoc3;
	node[F_SXY * fs] += m->rhoc2 * (e4 - e3);
}

/*
 * One sweep along one line of n nodes (simple/le_core.c:474-498 for x,
 * :516-535 for y).  `line` points at node 0 of the line, `ns` is the stride
 * between consecutive nodes of the line and `fs` the stride between fields.
 * The invariants of all n nodes are formed from pre-sweep values first, so the
 * in-place update cannot feed back (the reference achieves the same with a
 * window filled ahead of the write).  win[p] holds position p-2.
 */
static void orc_line(const orc_material *m, real_t *line, int n, long ns, long fs, int ax,
                     real_t k1, real_t k2, int edge, real_t (*win)[4])
{
	int p, i, q;
	for (p = 0; p < n; p++)
		orc_invariants(m, line + (long)p * ns, fs, ax, win[p + 2]);
	/* left: positions -2, -1 repeat node 0 (simple/le_core.c:486,521) */
	for (q = 0; q < 4; q++) win[0][q] = win[1][q] = win[2][q];
	/* right: positions n, n+1.  Clamp repeats node n-1.  The ring's slot for
	 * position p was last filled with position p-5 (five slots, refills stop
	 * after position n-1), which is node max(p-5, 0). */
	for (p = n; p < n + 2; p++) {
		const int src = (edge == ORC_EDGE_CLAMP) ? n - 1 : (p - 5 > 0 ? p - 5 : 0);
		for (q = 0; q < 4; q++) win[p + 2][q] = win[src + 2][q];
	}
	for (i = 0; i < n; i++) {
		real_t *w_2 = win[i], *w_1 = win[i + 1], *w0 = win[i + 2], *w1 = win[i + 3],
		       *w2 = win[i + 4];
		real_t d[4];
		/* simple/le_core.c:251-257: speeds +c1, -c1, +c2, -c2 */
		d[0] = orc_tvd2(k1, w_2[0], w_1[0], w0[0], w1[0]) - w0[0];
		d[1] = orc_tvd2(k1, w2[1], w1[1], w0[1], w_1[1]) - w0[1];
		d[2] = orc_tvd2(k2, w_2[2], w_1[2], w0[2], w1[2]) - w0[2];
		d[3] = orc_tvd2(k2, w2[3], w1[3], w0[3], w_1[3]) - w0[3];
		orc_apply(m, line + (long)i * ns, fs, ax, d);
	}
}

/* Grid layout descriptor: AOS (node stride 5, field stride 1) or SOA planes
 * (node stride 1, field stride nx*ny). */
static void orc_strides(int nx, int ny, int soa, long *node_stride, long *field_stride)
{
	if (soa) {
		*node_stride = 1;
		*field_stride = (long)nx * ny;
	} else {
		*node_stride = 5;
		*field_stride = 1;
	}
}

/* x-sweep, simple/le_core.c:435-501 (AOS) and :367-433 (SOA). */
int orc_sweep_x(real_t *g, int nx, int ny, int soa, real_t dt, real_t hx, const orc_material *m)
{
	const real_t k1 = dt * m->c1 / hx;
	const real_t k2 = dt * m->c2 / hx;
	long ns, fs;
	int j;
	if (nx < 3 || ny < 3) return 1;
	orc_strides(nx, ny, soa, &ns, &fs);
#pragma omp parallel for schedule(static)
	for (j = 0; j < ny; j++) {
		real_t (*win)[4] = malloc(sizeof(real_t[4]) * (size_t)(nx + 4));
		orc_line(m, g + (long)j * nx * ns, nx, ns, fs, 0, k1, k2, ORC_EDGE_CLAMP, win);
		free(win);
	}
	return 0;
}

/* y-sweep, simple/le_core.c:503-538 (clamp) or the ring edge of the SOA
 * y-sweep :259-364 and le_step_y_cf :540-602. */
int orc_sweep_y(real_t *g, int nx, int ny, int soa, real_t dt, real_t hy, const orc_material *m,
                int edge)
{
	const real_t k1 = dt * m->c1 / hy;
	const real_t k2 = dt * m->c2 / hy;
	long ns, fs;
	int i;
	if (nx < 3 || ny < 3) return 1;
	orc_strides(nx, ny, soa, &ns, &fs);
#pragma omp parallel for schedule(static)
	for (i = 0; i < nx; i++) {
		real_t (*win)[4] = malloc(sizeof(real_t[4]) * (size_t)(ny + 4));
		orc_line(m, g + (long)i * ns, ny, (long)nx * ns, fs, 1, k1, k2, edge, win);
		free(win);
	}
	return 0;
}

/* One full step: x then y (simple/le_core.c:605-620 AOS, :623-627 SOA). */
int orc_step(real_t *g, int nx, int ny, int soa, real_t dt, real_t hx, real_t hy,
             const orc_material *m)
{
	int rc = orc_sweep_x(g, nx, ny, soa, dt, hx, m);
	if (rc) return rc;
	return orc_sweep_y(g, nx, ny, soa, dt, hy, m, soa ? ORC_EDGE_RING : ORC_EDGE_CLAMP);
}

/* Repeated steps (the step loop of simple/main.c:57-63). */
int orc_run(real_t *g, int nx, int ny, int soa, real_t dt, real_t hx, real_t hy,
            const orc_material *m, int steps)
{
	int s, rc = 0;
	for (s = 0; s < steps && !rc; s++) rc = orc_step(g, nx, ny, soa, dt, hx, hy, m);
	return rc;
}

/* Pressure disk, simple/le_core.c:71-94: sxx = syy = s where the distance of
 * node (h.x*i, h.y*j) from c is strictly below r.  sqrt is libm's, in double,
 * of a sum formed in real_t. */
long orc_set_ball(real_t *g, int nx, int ny, int soa, real_t hx, real_t hy, real_t cx, real_t cy,
                  real_t r, real_t s)
{
	long ns, fs, hits = 0;
	int i, j;
	orc_strides(nx, ny, soa, &ns, &fs);
	for (i = 0; i < nx; i++) {
		for (j = 0; j < ny; j++) {
			const real_t px = hx * i, py = hy * j;
			const real_t dx = px - cx, dy = py - cy;
			if (sqrt(dx * dx + dy * dy) < r) {
				real_t *node = g + ((long)i + (long)j * nx) * ns;
				node[F_SXX * fs] = s;
				node[F_SYY * fs] = s;
				hits++;
			}
		}
	}
	return hits;
}

/* Legacy-VTK writer of |v|, simple/le_core.c:99-169: ASCII header, then one
 * big-endian float32 per node, rows j outer, i inner. */
int orc_save_vtk(const real_t *g, int nx, int ny, int soa, real_t hx, real_t hy, const char *path)
{
	long ns, fs;
	int i, j;
	FILE *fp = fopen(path, "w");
	if (!fp) return 1;
	orc_strides(nx, ny, soa, &ns, &fs);
	fprintf(fp, "# vtk DataFile Version 3.0\n");
	fprintf(fp, "Created by le_save_task\n");
	fprintf(fp, "BINARY\n");
	fprintf(fp, "DATASET STRUCTURED_POINTS\n");
	fprintf(fp, "DIMENSIONS %d %d 1\n", nx, ny);
	fprintf(fp, "SPACING %f %f 0.0\n", (double)hx, (double)hy);
	fprintf(fp, "ORIGIN 0.0 0.0 0.0\n");
	fprintf(fp, "POINT_DATA %d\n", (int)((unsigned)nx * (unsigned)ny));
	fprintf(fp, "SCALARS v float 1\n");
	fprintf(fp, "LOOKUP_TABLE v_table\n");
	for (j = 0; j < ny; j++) {
		for (i = 0; i < nx; i++) {
			const real_t *node = g + ((long)i + (long)j * nx) * ns;
			const real_t vx = node[F_VX * fs], vy = node[F_VY * fs];
			const float v = (float)sqrt(vx * vx + vy * vy);
			uint32_t bits;
			unsigned char be[4];
			memcpy(&bits, &v, 4);
			be[0] = (unsigned char)(bits >> 24);
			be[1] = (unsigned char)(bits >> 16);
			be[2] = (unsigned char)(bits >> 8);
			be[3] = (unsigned char)bits;
			fwrite(be, 1, 4, fp);
		}
	}
	fclose(fp);
	return 0;
}

/*
 * Heterogeneous media (no counterpart in the reference, which has one
 * material per task): every node is updated in its own material.  For the
 * sweep along `ax`, node i of a line with material m = ids[i] projects the
 * five stencil nodes i-2 .. i+2 (edge handling as orc_line) on m's
 * invariants, applies the reference's tvd2 with m's Courant numbers
 * dt*c/h and projects back with m's constants.  With one material this is
 * orc_line operation for operation.  mats[k] holds material k's constants
 * (orc_material_init).
 */
static void orc_line_mat(const orc_material *mats, const unsigned char *ids, long is, real_t *line,
                         int n, long ns, long fs, int ax, real_t dt, real_t h, int edge,
                         real_t *save)
{
	int i, q, p;
	/* pre-sweep copy of the line: save[5*p + f] */
	for (p = 0; p < n; p++)
		for (q = 0; q < 5; q++) save[5 * p + q] = line[(long)p * ns + q * fs];
	for (i = 0; i < n; i++) {
		const orc_material *m = &mats[ids[(long)i * is]];
		const real_t k1 = dt * m->c1 / h, k2 = dt * m->c2 / h;
		real_t w[5][4], d[4];
		for (q = 0; q < 5; q++) {
			int pos = i - 2 + q, src;
			if (pos < 0) src = 0;
			else if (pos < n) src = pos;
			else src = (edge == ORC_EDGE_CLAMP) ? n - 1 : (pos - 5 > 0 ? pos - 5 : 0);
			orc_invariants(m, save + 5 * src, 1, ax, w[q]);
		}
		d[0] = orc_tvd2(k1, w[0][0], w[1][0], w[2][0], w[3][0]) - w[2][0];
		d[1] = orc_tvd2(k1, w[4][1], w[3][1], w[2][1], w[1][1]) - w[2][1];
		d[2] = orc_tvd2(k2, w[0][2], w[1][2], w[2][2], w[3][2]) - w[2][2];
		d[3] = orc_tvd2(k2, w[4][3], w[3][3], w[2][3], w[1][3]) - w[2][3];
		orc_apply(m, line + (long)i * ns, fs, ax, d);
	}
}

/* One sweep along `ax` (0 = x with clamped ends, 1 = y with `edge`) of the
 * heterogeneous scheme on an AOS or SOA grid; ids[i + j*nx] < count;
 * c1c2rho holds count (c1, c2, rho) triples. */
int orc_sweep_mat(real_t *g, int nx, int ny, int soa, real_t dt, real_t h, int count,
                  const double *c1c2rho, const unsigned char *ids, int ax, int edge)
{
	long ns, fs;
	int l, k;
	orc_material *mats;
	if (nx < 3 || ny < 3 || count < 1 || ax < 0 || ax > 1) return 1;
	mats = malloc(sizeof(orc_material) * (size_t)count);
	for (k = 0; k < count; k++)
		orc_material_init(c1c2rho[3 * k], c1c2rho[3 * k + 1], c1c2rho[3 * k + 2], &mats[k]);
	orc_strides(nx, ny, soa, &ns, &fs);
	if (ax == 0) {
#pragma omp parallel for schedule(static)
		for (l = 0; l < ny; l++) {
			real_t *save = malloc(sizeof(real_t) * 5 * (size_t)nx);
			orc_line_mat(mats, ids + (long)l * nx, 1, g + (long)l * nx * ns, nx, ns, fs, 0, dt,
			             h, ORC_EDGE_CLAMP, save);
			free(save);
		}
	} else {
#pragma omp parallel for schedule(static)
		for (l = 0; l < nx; l++) {
			real_t *save = malloc(sizeof(real_t) * 5 * (size_t)ny);
			orc_line_mat(mats, ids + l, nx, g + (long)l * ns, ny, (long)nx * ns, fs, 1, dt, h,
			             edge, save);
			free(save);
		}
	}
	free(mats);
	return 0;
}

/* One step (x then y) of the heterogeneous scheme. */
int orc_step_mat(real_t *g, int nx, int ny, int soa, real_t dt, real_t hx, real_t hy, int count,
                 const double *c1c2rho, const unsigned char *ids, int edge)
{
	int rc = orc_sweep_mat(g, nx, ny, soa, dt, hx, count, c1c2rho, ids, 0, ORC_EDGE_CLAMP);
	if (rc) return rc;
	return orc_sweep_mat(g, nx, ny, soa, dt, hy, count, c1c2rho, ids, 1, edge);
}

/* Size of real_t, so a loader can check which build it opened. */
int orc_real_size(void) { return (int)sizeof(real_t); }
