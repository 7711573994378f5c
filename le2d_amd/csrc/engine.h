// engine.h -- internal interface of the MI355X sweep engine.
//
// A grid lives in HBM as five pitched planes (vx, vy, sxx, sxy, syy), each
// row starting on a 256-byte boundary, in a ping-pong pair: a sweep reads one
// buffer and writes the other.  The C ABI (le_core_api.cpp) and the batched
// and serving layers sit on top of this header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

// Flux limiter of the TVD scheme, a build-time choice as in the reference
// (simple/le_core.c:36-43, `#define limiter limiter_superbee`): 0 superbee
// (the reference's choice and the default), 1 minmod, 2 CIR (no limiting:
// first-order upwind).  The Makefile's LIMITER=superbee|minmod|cir sets it.
#ifndef LE2D_LIMITER
#define LE2D_LIMITER 0
#endif
#if LE2D_LIMITER < 0 || LE2D_LIMITER > 2
#error "LE2D_LIMITER must be 0 (superbee), 1 (minmod) or 2 (cir)"
#endif

namespace le2d {

// y-edge handling.  CLAMP: positions past the last row read the last row
// (reference simple/le_core.c:534).  RING: they read the stale slot of the
// 5-row ring of the SOA y-sweep (simple/le_core.c:310-349) and le_step_y_cf
// (:574-594), i.e. row max(p-5, 0) for position p >= ny.
enum Edge { EDGE_CLAMP = 0, EDGE_RING = 1 };

// Constants of one sweep in the arithmetic type T.  k = dt*c/h is the
// Courant number of the +-c1 (k1) and +-c2 (k2) invariants; hk = (1 - k)/2.
template <typename T>
struct SweepCoef {
  T k1, k2, hk1, hk2;
  T irhoc1, irhoc2, rhoc1, rhoc2, rhoc3;
};

// Launch shape of the sweeps (chosen per grid by the autotuner).
struct Tiles {
  int x_rows;   // rows per x-sweep workgroup
  int y_seg;    // rows per y-sweep segment (each segment re-reads 4 halo rows)
  int y_block;  // columns per y-sweep workgroup (64, 128 or 256)
  int xy_seg;   // rows per segment of the fused x+y step
  int xy_block; // lanes per fused-step workgroup (64, 128, 256); writes xy_block-4 columns
  int pair_seg; // rows per segment of the two-step pass (0: one step per pass)
  int pair_block; // lanes per two-step workgroup (64, 128 or 256); writes pair_block-8 columns
  int pair_split = 0; // 64-lane two-step windows: the two stages on two wavefronts
};

struct DevGrid {
  int elem = 0;        // bytes per value: 4 or 8
  int nx = 0, ny = 0;
  long pitch = 0;      // elements between rows (of one field)
  long plane = 0;      // elements between the fields of a row
  void *buf[2] = {nullptr, nullptr};
  int cur = 0;         // buffer that holds the current state
  hipStream_t stream = nullptr;
  int device = 0;
  Tiles tiles{4, 128, 256, 64, 128, 0, 256};
};

// Error reporting: every function returns 0 on success or a nonzero code and
// records a message retrievable with last_error().
enum Status {
  OK = 0,
  ERR_ARG = 1,
  ERR_DEVICE = 2,
  ERR_NOMEM = 3,
  ERR_STATE = 4,
  ERR_BUDGET = 5,
  ERR_BUSY = 6
};
int set_error(int code, const char *fmt, ...);
const char *last_error();

// Aborts with a message when no usable HIP device is present.
int require_device();

// Every device allocation of the library goes through dev_alloc / dev_free:
// the caller's allocator when one is installed (le2d_set_device_allocator),
// else hipMalloc / hipFree.  An allocation is freed by the allocator that
// made it, whatever is installed by then.  dev_alloc returns OK, or
// ERR_NOMEM with *p = nullptr and the error message set (`what` names the
// allocation).
int dev_alloc(void **p, size_t bytes, const char *what);
void dev_free(void *p);

// Layout of one state buffer: field q of node (i, j) at q*plane + j*pitch + i.
void grid_layout(int elem, int nx, int ny, long *pitch, long *plane);
size_t grid_bytes(int elem, int nx, int ny);  // device bytes of one state buffer
int grid_alloc(DevGrid *g, int elem, int nx, int ny, hipStream_t stream);
void grid_free(DevGrid *g);

// Host <-> device transfers of the current state.  AOS host arrays hold
// nx*ny nodes of five values; SOA host arrays hold five planes of nx*ny.
int grid_upload(DevGrid *g, const void *host, int soa);
int grid_download(DevGrid *g, void *host, int soa);

// One sweep of the current state into the other buffer, then swap.
template <typename T>
int sweep_x(DevGrid *g, const SweepCoef<T> &c);
template <typename T>
int sweep_y(DevGrid *g, const SweepCoef<T> &c, int edge);
// One full step (x-sweep then y-sweep) in a single pass.
template <typename T>
int step_xy(DevGrid *g, const SweepCoef<T> &cx, const SweepCoef<T> &cy, int edge);
// Two full steps in a single pass (clamped edges; tiles.pair_seg rows per
// segment, which must be > 0).
template <typename T>
int step2_xy(DevGrid *g, const SweepCoef<T> &cx, const SweepCoef<T> &cy);

// Streaming copy of `bytes` between two device buffers, `iters` times; mean
// milliseconds per copy (HIP events).  The reference point for achievable
// HBM bandwidth and for scaling hardware byte counters.
int copy_probe(size_t bytes, int iters, double *ms);

Tiles default_tiles(int elem, int nx, int ny);
// Workgroups of the two-step kernel (element size elem, fb lanes) resident on
// the current device at once (occupancy API x compute units); 0 if unknown.
long pair_resident(int elem, int fb);
// The same for the single-step kernel (clamped edges).
long step_resident(int elem, int fb);
// Segment height of the two-step pass for an nx x ny grid in windows of fb
// lanes, from the occupancy model in default_tiles.
int pair_seg_for(int elem, int nx, int ny, int fb);
// One grid row (five padded field rows) fits the 31-bit range of a buffer
// resource, which the row loads and stores address through.
bool row_fits(int elem, int nx);

// Heterogeneous media.  A material plane assigns each node one of `count`
// materials (c1, c2, rho); see media.hip for the scheme.
struct MatTile {
  int strip, j0, j1, mat;  // mat >= 0: the tile's one material; -1: mixed
  int seg;                 // segment index (sets the march direction)
};

struct MatPlane {
  bool active = false;
  int count = 0;
  std::vector<double> c1c2rho;     // 3 per material
  std::vector<unsigned char> ids;  // host copy, nx*ny
  unsigned char *d_ids = nullptr;  // device copy, rows of idp bytes
  long idp = 0;
  void *d_cx = nullptr, *d_cy = nullptr;  // SweepCoef<T> per material
  double *d_etab = nullptr;               // (rho, lambda, mu) per material
  MatTile *d_homog = nullptr, *d_mixed = nullptr;
  int n_homog = 0, n_mixed = 0;
  int built_fb = -1, built_seg = -1, built_edge = -1;
  // tables of the two-step pass (strips of the two-step window less 8 columns)
  MatTile *d_homog2 = nullptr, *d_mixed2 = nullptr;
  int n_homog2 = 0, n_mixed2 = 0;
  int built_pair_seg = -1, built_pair_fb = -1;
  // the interface tiles run on a second stream beside the single-material
  // ones (forked from and joined back into the grid's stream)
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

int mat_set(DevGrid *g, MatPlane *m, int count, const double *c1c2rho, const unsigned char *ids,
            double dt, double hx, double hy);
void mat_release(MatPlane *m);
template <typename T>
int step_xy_mat(DevGrid *g, MatPlane *m, int edge);
// Two steps in one pass (clamped edges, tiles.pair_seg rows per segment).
template <typename T>
int step2_xy_mat(DevGrid *g, MatPlane *m);

// FP64 norms of a - b over all fields: out = {||a-b||_2, ||b||_2,
// max|a-b|, max|b|}.  a and b are device states (planes of pitch p*, plane
// stride q*, element size e*) of the same nx x ny grid.
int diff_norms(const void *a, int ea, long pa, long qa, const void *b, int eb, long pb, long qb,
               int nx, int ny, hipStream_t stream, double out[4]);
// The same norms with no allocation: partials in a fixed device array of the
// library (one per device; concurrent calls on one device race).
int diff_norms_noalloc(const void *a, int ea, long pa, long qa, const void *b, int eb, long pb,
                       long qb, int nx, int ny, double out[4]);

// Sum over the grid of the elastic energy density in FP64 (per-node material
// from the plane if active, else the given homogeneous material).
int energy(DevGrid *g, const MatPlane *m, double rho, double c1, double c2, double *out);

}  // namespace le2d
