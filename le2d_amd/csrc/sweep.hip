// sweep.hip -- the x- and y-sweep kernels for gfx950 and their launchers.
//
// Numerics (reference simple/le_core.c:29-69, :192-257): per sweep each node's
// state u = (vx, vy, sxx, sxy, syy) is projected on four Riemann invariants
// with speeds +c1, -c1, +c2, -c2; each invariant is advanced by the
// superbee-limited second-order TVD scheme with clamped edges; the increment
// is projected back.  All invariants are formed from pre-sweep values.
//
// Face formulation.  The reference evaluates, per node and invariant, the two
// face fluxes f(i+1/2) and f(i-1/2) (tvd2, :48-69).  The flux a node computes
// for its right face is, operand for operand, the one its right neighbour
// computes for its left face, so each face is evaluated once here and shared:
// one division, one limiter and one flux per face and invariant instead of
// two.  The node update then applies u - k*(f_out - f_in) - u in the
// reference's order, which keeps FP64 results bit-identical when the
// compiler does not contract multiply-adds (-ffp-contract=off).
//
// Layout: five planes, element (i, j) of plane q at q*plane + j*pitch + i.
#include "engine.h"

namespace le2d {

// ---------------------------------------------------------------------------
// Scalar pieces.

// Superbee limiter with the reference's ternary min/max (simple/le_core.c:31-38):
// max(0, max(min(1, 2r), min(2, r))).
template <typename T>
__device__ __forceinline__ T superbee(T r) {
  const T r2 = T(2) * r;
  const T a = (T(1) > r2) ? r2 : T(1);
  const T b = (T(2) > r) ? r : T(2);
  const T m = (a > b) ? a : b;
  return (T(0) > m) ? T(0) : m;
}

// Limited flux through the face between upwind node u0 and downwind node u1;
// um is the node upwind of u0 (simple/le_core.c:50-66 with eps = 1e-6).
template <typename T>
__device__ __forceinline__ T face_flux(T um, T u0, T u1, T omk) {
  const T eps = T(1e-6);
  const T r = ((u0 - um) + eps) / ((u1 - u0) + eps);
  return u0 + ((superbee(r) * T(0.5)) * omk) * (u1 - u0);
}

// Increment of one invariant from its outflow and inflow face fluxes
// (simple/le_core.c:68 and :253-256): (u - k*(out - in)) - u.
template <typename T>
__device__ __forceinline__ T face_delta(T u, T k, T out, T in) {
  return (u - k * (out - in)) - u;
}

template <typename T>
struct W4 {
  T a, b, c, d;  // speeds +c1, -c1, +c2, -c2
};

// Invariants for a sweep along the axis whose velocity is vn and whose normal
// stress is sn (simple/le_core.c:192-218).
template <typename T>
__device__ __forceinline__ W4<T> invariants(T vn, T sn, T vt, T sxy, const SweepCoef<T> &c) {
  const T n = sn * c.irhoc1;
  const T m = sxy * c.irhoc2;
  W4<T> w;
  w.a = vn - n;
  w.b = vn + n;
  w.c = vt - m;
  w.d = vt + m;
  return w;
}

// Faces of the four invariants between positions p and p+1, given the
// invariants at p-1, p, p+1, p+2.
template <typename T>
__device__ __forceinline__ W4<T> faces(const W4<T> &wm, const W4<T> &w0, const W4<T> &w1,
                                       const W4<T> &w2, const SweepCoef<T> &c) {
  W4<T> f;
  f.a = face_flux(wm.a, w0.a, w1.a, c.omk1);  // +c1: upwind is the left side
  f.b = face_flux(w2.b, w1.b, w0.b, c.omk1);  // -c1: upwind is the right side
  f.c = face_flux(wm.c, w0.c, w1.c, c.omk2);
  f.d = face_flux(w2.d, w1.d, w0.d, c.omk2);
  return f;
}

// Back-transform of the increments (simple/le_core.c:221-249).  vn/sn are
// the along-axis velocity and normal stress, vt/so the tangential velocity
// and the other normal stress.
template <typename T>
__device__ __forceinline__ void apply(T &vn, T &vt, T &sn, T &so, T &sxy, const W4<T> &w,
                                      const W4<T> &fl, const W4<T> &fr,
                                      const SweepCoef<T> &c) {
  const T e1 = T(0.5) * face_delta(w.a, c.k1, fr.a, fl.a);
  const T e2 = T(0.5) * face_delta(w.b, c.k1, fl.b, fr.b);
  const T e3 = T(0.5) * face_delta(w.c, c.k2, fr.c, fl.c);
  const T e4 = T(0.5) * face_delta(w.d, c.k2, fl.d, fr.d);
  vn = vn + (e1 + e2);
  vt = vt + (e3 + e4);
  sn = sn + (e2 - e1) * c.rhoc1;
  so = so + (e2 - e1) * c.rhoc3;
  sxy = sxy + c.rhoc2 * (e4 - e3);
}

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
  int b = 0;
  for (int j = j0; j < j1; ++j, b ^= 1) {
    const T *r = src + (long)j * pitch;
    T vx = r[xc], vy = r[plane + xc], sxx = r[2 * plane + xc], sxy = r[3 * plane + xc],
      syy = r[4 * plane + xc];
    const W4<T> w = invariants(vx, sxx, vy, sxy, c);
    sw[b][t + 2] = w;
    if (t < 4) {
      sw[b][hs] = invariants(r[hx], r[2 * plane + hx], r[plane + hx], r[3 * plane + hx], c);
    }
    __syncthreads();
    // face x+1/2 -> sf[t+1]; lane 0 also face x0-1/2 -> sf[0]
    sf[b][t + 1] = faces(sw[b][t + 1], sw[b][t + 2], sw[b][t + 3], sw[b][t + 4], c);
    if (t == 0) sf[b][0] = faces(sw[b][0], sw[b][1], sw[b][2], sw[b][3], c);
    __syncthreads();
    if (x < nx) {
      apply(vx, vy, sxx, syy, sxy, w, sf[b][t], sf[b][t + 1], c);
      T *o = dst + (long)j * pitch + x;
      o[0] = vx;
      o[plane] = vy;
      o[2 * plane] = sxx;
      o[3 * plane] = sxy;
      o[4 * plane] = syy;
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
template <int EDGE>
__device__ __forceinline__ int ypos(int p, int ny) {
  if (p < 0) return 0;
  if (p < ny) return p;
  return EDGE == EDGE_CLAMP ? ny - 1 : max(p - 5, 0);
}

template <typename T>
struct Node {
  T vx, vy, sxx, sxy, syy;
};

template <typename T>
__device__ __forceinline__ Node<T> load_node(const T *__restrict__ p, long plane) {
  Node<T> n;
  n.vx = p[0];
  n.vy = p[plane];
  n.sxx = p[2 * plane];
  n.sxy = p[3 * plane];
  n.syy = p[4 * plane];
  return n;
}

template <typename T>
__device__ __forceinline__ W4<T> inv_y(const Node<T> &n, const SweepCoef<T> &c) {
  return invariants(n.vy, n.syy, n.vx, n.sxy, c);
}

template <typename T, int EDGE>
__global__ __launch_bounds__(256) void k_sweep_y(const T *__restrict__ src, T *__restrict__ dst,
                                                 int nx, int ny, long pitch, long plane, int seg,
                                                 SweepCoef<T> c) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;  // no barriers below
  const int j0 = blockIdx.y * seg;
  const int j1 = min(j0 + seg, ny);
  const T *col = src + x;
  T *out = dst + x;

  // prologue: positions j0-2 .. j0+2
  const W4<T> wm2 = inv_y(load_node(col + (long)ypos<EDGE>(j0 - 2, ny) * pitch, plane), c);
  W4<T> wm1 = inv_y(load_node(col + (long)ypos<EDGE>(j0 - 1, ny) * pitch, plane), c);
  Node<T> u0 = load_node(col + (long)ypos<EDGE>(j0, ny) * pitch, plane);
  Node<T> u1 = load_node(col + (long)ypos<EDGE>(j0 + 1, ny) * pitch, plane);
  Node<T> u2 = load_node(col + (long)ypos<EDGE>(j0 + 2, ny) * pitch, plane);
  W4<T> w0 = inv_y(u0, c), w1 = inv_y(u1, c), w2 = inv_y(u2, c);
  W4<T> fin = faces(wm2, wm1, w0, w1, c);  // face j0-1/2

  for (int j = j0; j < j1; ++j) {
    // issue the read of position j+3 before this row's arithmetic
    const Node<T> un = load_node(col + (long)ypos<EDGE>(j + 3, ny) * pitch, plane);
    const W4<T> fout = faces(wm1, w0, w1, w2, c);  // face j+1/2
    Node<T> u = u0;
    apply(u.vy, u.vx, u.syy, u.sxx, u.sxy, w0, fin, fout, c);
    T *o = out + (long)j * pitch;
    o[0] = u.vx;
    o[plane] = u.vy;
    o[2 * plane] = u.sxx;
    o[3 * plane] = u.sxy;
    o[4 * plane] = u.syy;
    fin = fout;
    wm1 = w0;
    w0 = w1;
    w1 = w2;
    u0 = u1;
    u1 = u2;
    u2 = un;
    w2 = inv_y(un, c);
  }
}

// ---------------------------------------------------------------------------
// Launchers.

template <typename T>
int sweep_x(DevGrid *g, const SweepCoef<T> &c) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "sweep_x: element size mismatch");
  constexpr int BX = 256;
  // gridDim.y is at most 65535
  const int rows = max(max(1, g->tiles.x_rows), (g->ny + 65534) / 65535);
  dim3 grid((g->nx + BX - 1) / BX, (g->ny + rows - 1) / rows);
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
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
