// batch_dev.h -- kernels of batched launches (many independent grids per
// launch), shared by the fixed batch (batch.hip) and the serving loop
// (serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_dev.h"

namespace le2d {

// One instance of a batched launch.  The state of step s is in buf[par ^ (s & 1)]
// where s counts the launches of the current table (par fixes each
// instance's own buffer parity when the table was built).
template <typename T>
struct InstDesc {
  T *buf[2];
  int nx, ny;
  long pitch, plane;
  SweepCoef<T> cx, cy;
  int par;
};

struct TileJob {
  int inst, strip, j0, j1;
  int seg;  // segment index within the instance (sets the march direction)
};

// Persistent over the tile table.  Workgroup b runs on XCD b % 8 (dispatch
// is round-robin); each XCD walks its own contiguous eighth of the table, so
// concurrently running workgroups of one XCD process neighbouring strips.
template <typename T, int FB>
__global__ __launch_bounds__(FB, 3) void k_step_batch(const InstDesc<T> *__restrict__ inst,
                                                   const TileJob *__restrict__ tiles, int ntiles,
                                                   int step_parity) {
  __shared__ W4<T> sw[FB];
  __shared__ W4<T> sf[FB];
  const int per = gridDim.x >> 3;  // gridDim.x is a multiple of 8
  const int xcd = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int lo = (int)((long)ntiles * xcd / 8), hi = (int)((long)ntiles * (xcd + 1) / 8);
  for (int k = lo + r; k < hi; k += per) {
    const TileJob tj = tiles[k];
    const InstDesc<T> &d = inst[tj.inst];
    const int p = d.par ^ step_parity;
    march_tile_alt<T, EDGE_CLAMP, FB>(d.buf[p], d.buf[p ^ 1], d.nx, d.ny, d.pitch, d.plane,
                                      tj.strip, tj.j0, tj.j1, tj.seg, d.cx, d.cy, sw, sf);
  }
}

// One workgroup per tile: the table in XCD-aware order (xcd_tile), so each
// XCD runs a contiguous eighth of it, as the persistent form walks it.
template <typename T, int FB>
__global__ __launch_bounds__(FB, 3) void k_step_batch_tiles(const InstDesc<T> *__restrict__ inst,
                                                          const TileJob *__restrict__ tiles,
                                                          int step_parity) {
  __shared__ W4<T> sw[FB];
  __shared__ W4<T> sf[FB];
  const TileJob tj = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const InstDesc<T> &d = inst[tj.inst];
  const int p = d.par ^ step_parity;
  march_tile_alt<T, EDGE_CLAMP, FB>(d.buf[p], d.buf[p ^ 1], d.nx, d.ny, d.pitch, d.plane,
                                    tj.strip, tj.j0, tj.j1, tj.seg, d.cx, d.cy, sw, sf);
}

// One workgroup per tile of the two-step table (strips of FB-8 columns): each
// launch advances every instance by two steps (march_pair).
template <typename T, int FB>
__global__ __launch_bounds__(FB, 2) void k_step_batch_pairs(const InstDesc<T> *__restrict__ inst,
                                                          const TileJob *__restrict__ tiles,
                                                          int step_parity) {
  __shared__ W4<T> sw[pair_lds<FB>()];
  __shared__ W4<T> sf[pair_lds<FB>()];
  const TileJob tj = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const InstDesc<T> &d = inst[tj.inst];
  const int p = d.par ^ step_parity;
  march_pair<T, FB>(d.buf[p], d.buf[p ^ 1], d.nx, d.ny, d.pitch, d.plane, tj.strip, tj.j0, tj.j1,
                    d.cx, d.cy, sw, sf);
}

// The same with 64-lane windows whose two stages run on two wavefronts of
// a 128-lane workgroup (march_pair_split).
template <typename T>
__global__ __launch_bounds__(128, 3) void k_step_batch_pairs_split(
    const InstDesc<T> *__restrict__ inst, const TileJob *__restrict__ tiles, int step_parity) {
  __shared__ W4<T> sw[128];
  __shared__ W4<T> sf[128];
  __shared__ T ring[2 * kSplitRows * 5 * 64];
  const TileJob tj = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const InstDesc<T> &d = inst[tj.inst];
  const int p = d.par ^ step_parity;
  march_pair_split<T, sizeof(T) == 4>(d.buf[p], d.buf[p ^ 1], d.nx, d.ny, d.pitch, d.plane,
                                      tj.strip, tj.j0, tj.j1, d.cx, d.cy, sw, sf, ring);
}

// Disk initial condition with the host le_set_ball's arithmetic: the node
// position and offset in T, the distance as the double square root of the
// T-rounded sum of squares, strict comparison with r (simple/le_core.c:71-94).
template <typename T>
__global__ void k_set_ball(T *__restrict__ buf, int nx, int ny, long pitch, long plane, T hx, T hy,
                           T cx, T cy, T r, T s) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long)nx * ny) return;
  const int j = (int)(n / nx), i = (int)(n - (long)j * nx);
  const T px = hx * (T)i, py = hy * (T)j;
  const T dx = px - cx, dy = py - cy;
  const T d2 = dx * dx + dy * dy;
  if (::sqrt((double)d2) < (double)r) {
    buf[2 * plane + (long)j * pitch + i] = s;
    buf[4 * plane + (long)j * pitch + i] = s;
  }
}

// Material and sweep constants in double (the reference's le_init_material
// and le_step_x/y expressions), rounded once to T.
template <typename T>
inline void coef_from_material(double c1, double c2, double rho, double dt, double hx, double hy,
                               SweepCoef<T> *cx, SweepCoef<T> *cy) {
  const double irhoc1 = 1.0 / (c1 * rho), irhoc2 = 1.0 / (c2 * rho);
  const double rhoc1 = c1 * rho, rhoc2 = c2 * rho;
  const double mu = rho * c2 * c2;
  const double la = rho * c1 * c1 - 2.0 * mu;
  const double rhoc3 = rho * c1 * la / (la + 2.0 * mu);
  for (int a = 0; a < 2; ++a) {
    const double h = a == 0 ? hx : hy;
    const double k1 = dt * c1 / h, k2 = dt * c2 / h;
    SweepCoef<T> &c = a == 0 ? *cx : *cy;
    c.k1 = (T)k1;
    c.k2 = (T)k2;
    c.hk1 = (T)((1.0 - k1) * 0.5);
    c.hk2 = (T)((1.0 - k2) * 0.5);
    c.irhoc1 = (T)irhoc1;
    c.irhoc2 = (T)irhoc2;
    c.rhoc1 = (T)rhoc1;
    c.rhoc2 = (T)rhoc2;
    c.rhoc3 = (T)rhoc3;
  }
}

// Persistent grid for a batched launch: resident workgroups on the whole
// device (a multiple of 8), at most one per 8 tiles rounded up.
template <typename T, int FB>
inline int persistent_blocks(int ntiles) {
  int dev = 0, cus = 256, per_cu = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_step_batch<T, FB>, FB, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  long want = (long)cus * per_cu;
  want = want < ((long)ntiles + 7) / 8 * 8 ? want : ((long)ntiles + 7) / 8 * 8;
  return (int)(want / 8 * 8 < 8 ? 8 : want / 8 * 8);
}

}  // namespace le2d
