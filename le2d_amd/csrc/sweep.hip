// sweep.hip -- the x-sweep, y-sweep and fused-step kernels for gfx950 and
// their launchers.  Numerics and the fused-step tile march: sweep_dev.h.
#include <map>
#include <mutex>
#include <tuple>

#include "sweep_dev.h"

namespace le2d {

// ---------------------------------------------------------------------------
// x-sweep.  A workgroup of BX lanes owns BX consecutive nodes of a row and
// walks `rows` rows.  Per row: coalesced loads of the five planes, invariants
// to LDS (with a 2-node clamped halo each side), one face per lane (lane 0
// also the segment's left face) to LDS, then the node update and coalesced
// stores.  LDS is double-buffered by row parity, so a row costs two barriers.
template <typename T, int BX>
__global__ __launch_bounds__(BX) void k_sweep_x(const T *__restrict__ src, T *__restrict__ dst,
                                                int nx, int ny, long pitch, long plane,
                                                int rows, SweepCoef<T> c) {
  __shared__ W4<T> sw[2][BX + 4];
  __shared__ W4<T> sf[2][BX + 1];
  const int t = threadIdx.x;
  const int x0 = blockIdx.x * BX;
  const int x = x0 + t;
  const int xc = min(x, nx - 1);
  // lanes 0..3 also fetch the halo nodes x0-2, x0-1, x0+BX, x0+BX+1 (clamped)
  const int hx = (t < 2) ? max(x0 - 2 + t, 0) : min(x0 + BX + (t - 2), nx - 1);
  const int hs = (t < 2) ? t : BX + t;
  const int j0 = blockIdx.y * rows;
  const int j1 = min(j0 + rows, ny);
  const unsigned boff = (unsigned)xc * (unsigned)sizeof(T);
  const unsigned hoff = (unsigned)hx * (unsigned)sizeof(T);
  const unsigned woff = (unsigned)x * (unsigned)sizeof(T);
  int b = 0;
  for (int j = j0; j < j1; ++j, b ^= 1) {
    const T *r = src + (long)j * pitch;
    Node<T> u = load_row_node(r, plane, boff);
    const W4<T> w = invariants(u.vx, u.sxx, u.vy, u.sxy, c);
    sw[b][t + 2] = w;
    if (t < 4) {
      const Node<T> h = load_row_node(r, plane, hoff);
      sw[b][hs] = invariants(h.vx, h.sxx, h.vy, h.sxy, c);
    }
    __syncthreads();
    // face x+1/2 -> sf[t+1]; lane 0 also face x0-1/2 -> sf[0]
    sf[b][t + 1] = faces(sw[b][t + 1], sw[b][t + 2], sw[b][t + 3], sw[b][t + 4], c);
    if (t == 0) sf[b][0] = faces(sw[b][0], sw[b][1], sw[b][2], sw[b][3], c);
    __syncthreads();
    if (x < nx) {
      apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w, sf[b][t], sf[b][t + 1], c);
      store_row_node(dst + (long)j * pitch, plane, woff, u);
    }
  }
}

// ---------------------------------------------------------------------------
// y-sweep.  Each lane owns one column and marches down a segment of `seg`
// rows with a register window of invariants (positions j-1 .. j+2) and the
// raw nodes of rows j .. j+2; every step reads one new row and writes one
// row, so a wavefront's accesses are contiguous row pieces (no strided
// loads, no transpose).  The face flux at j+1/2 computed in step j is reused
// as the inflow face in step j+1.  Segments re-read two rows above and below.
template <typename T, int EDGE>
__global__ __launch_bounds__(256) void k_sweep_y(const T *__restrict__ src, T *__restrict__ dst,
                                                 int nx, int ny, long pitch, long plane, int seg,
                                                 SweepCoef<T> c) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;  // no barriers below
  const int j0 = blockIdx.y * seg;
  const int j1 = min(j0 + seg, ny);
  const unsigned boff = (unsigned)x * (unsigned)sizeof(T);
#define YROW(p) load_row_node(src + (long)ypos<EDGE>((p), ny) * pitch, plane, boff)

  // prologue: positions j0-2 .. j0+2
  const W4<T> wm2 = inv_y(YROW(j0 - 2), c);
  W4<T> wm1 = inv_y(YROW(j0 - 1), c);
  Node<T> u0 = YROW(j0);
  Node<T> u1 = YROW(j0 + 1);
  Node<T> u2 = YROW(j0 + 2);
  W4<T> w0 = inv_y(u0, c), w1 = inv_y(u1, c), w2 = inv_y(u2, c);
  W4<T> fin = faces(wm2, wm1, w0, w1, c);  // face j0-1/2

  for (int j = j0; j < j1; ++j) {
    // issue the read of position j+3 before this row's arithmetic (reading
    // a row further ahead measured no faster)
    const Node<T> un = YROW(j + 3);
    const W4<T> fout = faces(wm1, w0, w1, w2, c);  // face j+1/2
    Node<T> u = u0;
    apply(u.vy, u.vx, u.syy, u.sxx, u.sxy, w0, fin, fout, c);
    store_row_node(dst + (long)j * pitch, plane, boff, u);
    fin = fout;
    wm1 = w0;
    w0 = w1;
    w1 = w2;
    u0 = u1;
    u1 = u2;
    u2 = un;
    w2 = inv_y(un, c);
  }
#undef YROW
}

// ---------------------------------------------------------------------------
// Fused step over one grid (layout and march: march_tile in sweep_dev.h).
template <typename T, int EDGE, int FB>
__global__ __launch_bounds__(FB, 3) void k_step_xy(const T *__restrict__ src, T *__restrict__ dst,
                                                int nx, int ny, long pitch, long plane, int seg,
                                                int nstrips, SweepCoef<T> cx, SweepCoef<T> cy) {
  __shared__ W4<T> sw[FB];
  __shared__ W4<T> sf[FB];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int strip = tile % nstrips, segi = tile / nstrips;
  const int j0 = segi * seg;
  const int j1 = min(j0 + seg, ny);
  march_tile_alt<T, EDGE, FB>(src, dst, nx, ny, pitch, plane, strip, j0, j1, segi, cx, cy, sw, sf);
}

// Two fused steps over one grid (march_pair in sweep_dev.h).  Occupancy 2:
// the two stages hold twice the march state of k_step_xy.
template <typename T, int FB>
__global__ __launch_bounds__(FB, 2) void k_step2_xy(const T *__restrict__ src, T *__restrict__ dst,
                                                 int nx, int ny, long pitch, long plane, int seg,
                                                 int nstrips, SweepCoef<T> cx, SweepCoef<T> cy) {
  __shared__ W4<T> sw[pair_lds<FB>()];
  __shared__ W4<T> sf[pair_lds<FB>()];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int strip = tile % nstrips, segi = tile / nstrips;
  const int j0 = segi * seg;
  const int j1 = min(j0 + seg, ny);
  march_pair<T, FB>(src, dst, nx, ny, pitch, plane, strip, j0, j1, cx, cy, sw, sf);
}

// Two fused steps, the stages on two wavefronts (march_pair_split): 128
// lanes per workgroup, each wavefront one 64-lane window of one stage; three
// workgroups' waves per SIMD (four spill in FP64).  HOLD: FP32 keeps each
// slot's y-invariants in registers.
template <typename T, bool HOLD>
__global__ __launch_bounds__(128, 3) void k_step2_split(
    const T *__restrict__ src, T *__restrict__ dst, int nx, int ny, long pitch, long plane,
    int seg, int nstrips, SweepCoef<T> cx, SweepCoef<T> cy) {
  __shared__ W4<T> sw[128];
  __shared__ W4<T> sf[128];
  __shared__ T ring[2 * kSplitRows * 5 * 64];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int strip = tile % nstrips, segi = tile / nstrips;
  const int j0 = segi * seg;
  const int j1 = min(j0 + seg, ny);
  march_pair_split<T, HOLD>(src, dst, nx, ny, pitch, plane, strip, j0, j1, cx, cy, sw, sf, ring);
}

// ---------------------------------------------------------------------------
// Launchers.

template <typename T, int FB>
static void launch_step_xy(DevGrid *g, const SweepCoef<T> &cx, const SweepCoef<T> &cy, int edge,
                           int seg, unsigned nblocks, int nstrips) {
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
  if (edge == EDGE_RING)
    hipLaunchKernelGGL((k_step_xy<T, EDGE_RING, FB>), dim3(nblocks), dim3(FB), 0, g->stream, src,
                       dst, g->nx, g->ny, g->pitch, g->plane, seg, nstrips, cx, cy);
  else
    hipLaunchKernelGGL((k_step_xy<T, EDGE_CLAMP, FB>), dim3(nblocks), dim3(FB), 0, g->stream, src,
                       dst, g->nx, g->ny, g->pitch, g->plane, seg, nstrips, cx, cy);
}

template <typename T>
int step_xy(DevGrid *g, const SweepCoef<T> &cx, const SweepCoef<T> &cy, int edge) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "step_xy: element size mismatch");
  const int fb = g->tiles.xy_block == 64 || g->tiles.xy_block == 256 ? g->tiles.xy_block : 128;
  const int nstrips = (g->nx + fb - 5) / (fb - 4);
  const int seg = max(1, g->tiles.xy_seg);
  const long nsegs = (g->ny + seg - 1) / seg;
  const long nblocks = nsegs * nstrips;
  if (nblocks > 0x7fffffffL) return set_error(ERR_ARG, "step_xy: grid too large");
  if (fb == 64)
    launch_step_xy<T, 64>(g, cx, cy, edge, seg, (unsigned)nblocks, nstrips);
  else if (fb == 128)
    launch_step_xy<T, 128>(g, cx, cy, edge, seg, (unsigned)nblocks, nstrips);
  else
    launch_step_xy<T, 256>(g, cx, cy, edge, seg, (unsigned)nblocks, nstrips);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ERR_DEVICE, "step_xy launch: %s", hipGetErrorString(e));
  g->cur ^= 1;
  return OK;
}

template <typename T, int FB>
static int launch_step2_xy(DevGrid *g, const SweepCoef<T> &cx, const SweepCoef<T> &cy, int seg) {
  const int nstrips = (g->nx + FB - 9) / (FB - 8);
  const long nsegs = (g->ny + seg - 1) / seg;
  const long nblocks = nsegs * nstrips;
  if (nblocks > 0x7fffffffL) return set_error(ERR_ARG, "step2_xy: grid too large");
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
  if (FB == 64 && g->tiles.pair_split)
    hipLaunchKernelGGL((k_step2_split<T, sizeof(T) == 4>),
                       dim3((unsigned)nblocks), dim3(128), 0,
                       g->stream, src, dst, g->nx, g->ny, g->pitch, g->plane, seg, nstrips, cx,
                       cy);
  else
    hipLaunchKernelGGL((k_step2_xy<T, FB>), dim3((unsigned)nblocks), dim3(FB), 0, g->stream, src,
                       dst, g->nx, g->ny, g->pitch, g->plane, seg, nstrips, cx, cy);
  return OK;
}

template <typename T>
int step2_xy(DevGrid *g, const SweepCoef<T> &cx, const SweepCoef<T> &cy) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "step2_xy: element size mismatch");
  const int seg = g->tiles.pair_seg;
  if (seg < 1) return set_error(ERR_ARG, "step2_xy: no pair segment height");
  const int fb = g->tiles.pair_block;
  const int rc = fb == 64    ? launch_step2_xy<T, 64>(g, cx, cy, seg)
                 : fb == 128 ? launch_step2_xy<T, 128>(g, cx, cy, seg)
                             : launch_step2_xy<T, 256>(g, cx, cy, seg);
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ERR_DEVICE, "step2_xy launch: %s", hipGetErrorString(e));
  g->cur ^= 1;
  return OK;
}

template <typename T, int FB, bool PAIR>
static int occupancy_of() {
  int n = 0;
  hipError_t e;
  if constexpr (PAIR)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step2_xy<T, FB>, FB, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step_xy<T, EDGE_CLAMP, FB>, FB, 0);
  return e == hipSuccess ? n : 0;
}

// Workgroups per CU x CUs of the current device, cached per device.
template <bool PAIR>
static long resident_of(int elem, int fb) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, long> cache;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(dev, elem, fb);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per = 0;
  if (elem == 8)
    per = fb == 64    ? occupancy_of<double, 64, PAIR>()
          : fb == 128 ? occupancy_of<double, 128, PAIR>()
                      : occupancy_of<double, 256, PAIR>();
  else
    per = fb == 64    ? occupancy_of<float, 64, PAIR>()
          : fb == 128 ? occupancy_of<float, 128, PAIR>()
                      : occupancy_of<float, 256, PAIR>();
  const long r = (long)per * cus;
  cache[key] = r;
  return r;
}

long pair_resident(int elem, int fb) { return resident_of<true>(elem, fb); }
long step_resident(int elem, int fb) { return resident_of<false>(elem, fb); }

template int step2_xy<double>(DevGrid *, const SweepCoef<double> &, const SweepCoef<double> &);
template int step2_xy<float>(DevGrid *, const SweepCoef<float> &, const SweepCoef<float> &);
template int step_xy<double>(DevGrid *, const SweepCoef<double> &, const SweepCoef<double> &, int);
template int step_xy<float>(DevGrid *, const SweepCoef<float> &, const SweepCoef<float> &, int);

template <typename T>
int sweep_x(DevGrid *g, const SweepCoef<T> &c) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "sweep_x: element size mismatch");
  constexpr int BX = 256;
  // gridDim.y is at most 65535
  const int rows = max(max(1, g->tiles.x_rows), (g->ny + 65534) / 65535);
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
  dim3 grid((g->nx + BX - 1) / BX, (g->ny + rows - 1) / rows);
  hipLaunchKernelGGL((k_sweep_x<T, BX>), grid, dim3(BX), 0, g->stream, src, dst, g->nx, g->ny,
                     g->pitch, g->plane, rows, c);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ERR_DEVICE, "sweep_x launch: %s", hipGetErrorString(e));
  g->cur ^= 1;
  return OK;
}

template <typename T>
int sweep_y(DevGrid *g, const SweepCoef<T> &c, int edge) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "sweep_y: element size mismatch");
  const int bx = g->tiles.y_block == 64 || g->tiles.y_block == 128 ? g->tiles.y_block : 256;
  const int seg = max(max(1, g->tiles.y_seg), (g->ny + 65534) / 65535);
  dim3 grid((g->nx + bx - 1) / bx, (g->ny + seg - 1) / seg);
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
  if (edge == EDGE_RING)
    hipLaunchKernelGGL((k_sweep_y<T, EDGE_RING>), grid, dim3(bx), 0, g->stream, src, dst, g->nx,
                       g->ny, g->pitch, g->plane, seg, c);
  else
    hipLaunchKernelGGL((k_sweep_y<T, EDGE_CLAMP>), grid, dim3(bx), 0, g->stream, src, dst, g->nx,
                       g->ny, g->pitch, g->plane, seg, c);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ERR_DEVICE, "sweep_y launch: %s", hipGetErrorString(e));
  g->cur ^= 1;
  return OK;
}

template int sweep_x<double>(DevGrid *, const SweepCoef<double> &);
template int sweep_x<float>(DevGrid *, const SweepCoef<float> &);
template int sweep_y<double>(DevGrid *, const SweepCoef<double> &, int);
template int sweep_y<float>(DevGrid *, const SweepCoef<float> &, int);

}  // namespace le2d
