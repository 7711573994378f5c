// sweep_dev.h -- device numerics of the sweeps, shared by every kernel.
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
// Layout: field q (vx, vy, sxx, sxy, syy) of node (i, j) at q*plane + j*pitch + i;
// the fields of a row lie side by side (plane = padded row, pitch = 5 rows).
#pragma once

#include "engine.h"

namespace le2d {

// ---------------------------------------------------------------------------
// Scalar pieces.

// Superbee limiter with the reference's ternary min/max (simple/le_core.c:31-38):
// max(0, max(min(1, 2r), min(2, r))).
template <typename T>
__device__ __forceinline__ T superbee(T r) {
  // max(0, max(min(1, 2r), min(2, r))) = max(c, min(r, 2)) = med3(r, c, 2)
  // with c = clamp(2r, 0, 1): for r < 0, c = 0 and min(r, 2) = r < 0; for
  // r >= 0, c = min(2r, 1).  The clamp is the add's output modifier.
  if constexpr (sizeof(T) == 4) {
    // FP32 tier (held to an error bound against FP64, not to bits): two
    // instructions.  A NaN r (0/0: both slopes exactly -eps) gives 0 here,
    // the first-order flux, where the reference's formula gives 2.
    const float c = __builtin_fminf(__builtin_fmaxf(r + r, 0.0f), 1.0f);
#if LE2D_FP32_NAN2
    return __builtin_fmaxf(c, __builtin_fminf(r, 2.0f));  // NaN: max(0, 2) = 2
#else
    return __builtin_amdgcn_fmed3f(r, c, 2.0f);
#endif
  } else {
    // FP64: max(c, min(r, 2)) is the reference's value bit for bit for every
    // r (NaN and infinities included: the clamp sends NaN to 0, min(NaN, 2)
    // is 2) except r = -0, where this gives +0 and the reference's ternaries
    // keep -0.  No result bit depends on that sign.  The limiter value
    // reaches the node only through the face flux u + (phi*hk)*du, where a
    // zero phi changes at most the sign of a zero flux, and through the
    // increment (u - k*(f_out - f_in)) - u, which is the same for fluxes that
    // differ only in the sign of a zero: f_out - f_in is then the same nonzero
    // value or a zero, and (u - (+-0)) - u is +0 for every finite u (NaN for
    // infinite u either way).
    const double c = __builtin_fmin(__builtin_fmax(r + r, 0.0), 1.0);
    return __builtin_fmax(c, __builtin_fmin(r, 2.0));
  }
}

// The build's flux limiter (LE2D_LIMITER).  minmod is the reference's
// ternary form max(0, min(1, r)) (simple/le_core.c:36), so NaN gives 1 and
// -0 stays -0 in both types; CIR is 0 (:37).  Both are options beside the
// default superbee.
template <typename T>
__device__ __forceinline__ T limiter(T r) {
#if LE2D_LIMITER == 1
  const T m = T(1) > r ? r : T(1);
  return T(0) > m ? T(0) : m;
#elif LE2D_LIMITER == 2
  (void)r;
  return T(0);
#else
  return superbee(r);
#endif
}

// Limited flux through the face between upwind node u0 and downwind node u1;
// um is the node upwind of u0 (simple/le_core.c:50-66 with eps = 1e-6).
// The reference forms ((phi / 2) * (1 - k)) * (u1 - u0); phi / 2 is exact for
// every value phi takes (superbee: 0, 1, 2, r in (1,2), 2r in (0,1); minmod:
// 0, 1, r in [0,1); CIR: 0; r is a normal number: its numerator is 0 or at
// least about 2^-73 in magnitude), so (phi / 2) * (1 - k) rounds the same
// real number as phi * hk with the exact constant hk = (1 - k) / 2, one
// multiply fewer per face.
// Slope ratio.  FP64: IEEE division, as the reference.  FP32 tier: the
// hardware reciprocal (about 1 ulp) times the numerator; that tier is held to
// an error bound against FP64, not to bit parity, and the correctly rounded
// FP32 division sequence costs several times more issue slots.
// Build options of the FP32 tier's accuracy/cost trade-off (defaults in
// brackets): LE2D_FP32_RATIO [0] the reciprocal times the numerator, 1 the
// same refined by one Newton step on the quotient, 2 IEEE division;
// LE2D_FP32_NAN2 [0] a NaN ratio gives the limiter 0, 1 gives 2 as the
// reference's formula does.
#ifndef LE2D_FP32_RATIO
#define LE2D_FP32_RATIO 0
#endif
#ifndef LE2D_FP32_NAN2
#define LE2D_FP32_NAN2 0
#endif
__device__ __forceinline__ double slope_ratio(double n, double d) { return n / d; }
__device__ __forceinline__ float slope_ratio(float n, float d) {
#if LE2D_FP32_RATIO == 2
  return n / d;
#elif LE2D_FP32_RATIO == 1
  const float rc = __builtin_amdgcn_rcpf(d);
  const float q = n * rc;
  return __builtin_fmaf(__builtin_fmaf(-d, q, n), rc, q);
#else
  return n * __builtin_amdgcn_rcpf(d);
#endif
}

template <typename T>
__device__ __forceinline__ T face_flux(T um, T u0, T u1, T hk) {
  const T eps = T(1e-6);
  const T du = u1 - u0;
  const T r = slope_ratio((u0 - um) + eps, du + eps);
  if constexpr (sizeof(T) == 4)
    return __builtin_fmaf(limiter(r) * hk, du, u0);  // FP32 tier: fused
  else
    return u0 + (limiter(r) * hk) * du;
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
  W4<T> w;
  if constexpr (sizeof(T) == 4) {
    // FP32 tier: one fused multiply-add per invariant (the negations on the
    // uniform factors, so no lane value needs negating)
    const T ni1 = -c.irhoc1, ni2 = -c.irhoc2;
    w.a = __builtin_fmaf(sn, ni1, vn);
    w.b = __builtin_fmaf(sn, c.irhoc1, vn);
    w.c = __builtin_fmaf(sxy, ni2, vt);
    w.d = __builtin_fmaf(sxy, c.irhoc2, vt);
  } else {
    const T n = sn * c.irhoc1;
    const T m = sxy * c.irhoc2;
    w.a = vn - n;
    w.b = vn + n;
    w.c = vt - m;
    w.d = vt + m;
  }
  return w;
}

// Faces of the four invariants between positions p and p+1, given the
// invariants at p-1, p, p+1, p+2.
template <typename T>
__device__ __forceinline__ W4<T> faces(const W4<T> &wm, const W4<T> &w0, const W4<T> &w1,
                                       const W4<T> &w2, const SweepCoef<T> &c) {
  W4<T> f;
  f.a = face_flux(wm.a, w0.a, w1.a, c.hk1);  // +c1: upwind is the left side
  f.b = face_flux(w2.b, w1.b, w0.b, c.hk1);  // -c1: upwind is the right side
  f.c = face_flux(wm.c, w0.c, w1.c, c.hk2);
  f.d = face_flux(w2.d, w1.d, w0.d, c.hk2);
  return f;
}

// Back-transform of the increments (simple/le_core.c:221-249).  vn/sn are
// the along-axis velocity and normal stress, vt/so the tangential velocity
// and the other normal stress.
template <typename T>
__device__ __forceinline__ void apply(T &vn, T &vt, T &sn, T &so, T &sxy, const W4<T> &w,
                                      const W4<T> &fl, const W4<T> &fr,
                                      const SweepCoef<T> &c) {
  if constexpr (sizeof(T) == 4) {
    // FP32 tier: the increment (u - k*(out - in)) - u is k*(in - out); the
    // back-transform's factor 1/2 and the stress factors are folded into
    // uniform constants (hoisted), each field takes one fused multiply-add
    const T h1 = T(0.5) * c.k1, h2 = T(0.5) * c.k2;
    const T da = fl.a - fr.a, db = fr.b - fl.b, dc = fl.c - fr.c, dd = fr.d - fl.d;
    const T t1 = db - da, t2 = dd - dc;
    vn = __builtin_fmaf(h1, da + db, vn);
    vt = __builtin_fmaf(h2, dc + dd, vt);
    sn = __builtin_fmaf(h1 * c.rhoc1, t1, sn);
    so = __builtin_fmaf(h1 * c.rhoc3, t1, so);
    sxy = __builtin_fmaf(h2 * c.rhoc2, t2, sxy);
  } else {
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
}

// ---------------------------------------------------------------------------
// Rows and nodes.

// Row of window position p along y: clamp (le_step) or the ring edge
// (le_step_soa / le_step_y_cf).
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

// c ? a : b field by field.  A conditional on two Node lvalues is copied
// through a selected address, which keeps an array holding either operand
// in scratch memory; selecting the values keeps it in registers.
template <typename T>
__device__ __forceinline__ Node<T> pick_node(bool c, const Node<T> &a, const Node<T> &b) {
  Node<T> n;
  n.vx = c ? a.vx : b.vx;
  n.vy = c ? a.vy : b.vy;
  n.sxx = c ? a.sxx : b.sxx;
  n.sxy = c ? a.sxy : b.sxy;
  n.syy = c ? a.syy : b.syy;
  return n;
}

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

// Row loads and stores go through a buffer resource whose base is the row
// (wavefront-uniform, built in SGPRs): the lane's byte offset within the field
// row is the VGPR offset and the field's plane offset the SGPR offset, so a
// load or store costs no per-lane address arithmetic (a global access with a
// 64-bit per-lane address costs one 64-bit add each).  The resource covers
// the row's five padded field rows (row_fits keeps that under 2^31 bytes);
// every offset used is inside it.  Word 3 is the gfx9 raw-buffer format
// (32-bit data, no swizzle).
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const T *row, long plane) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(row), (short)0,
                                           (int)(5 * plane * (long)sizeof(T)), 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff, T v) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

template <typename T>
__device__ __forceinline__ Node<T> load_row_node(const T *__restrict__ row, long plane,
                                                 unsigned boff) {
  const __amdgpu_buffer_rsrc_t r = row_rsrc(row, plane);
  const int ps = (int)(plane * (long)sizeof(T));
  Node<T> n;
  n.vx = buf_ld<T>(r, boff, 0);
  n.vy = buf_ld<T>(r, boff, ps);
  n.sxx = buf_ld<T>(r, boff, 2 * ps);
  n.sxy = buf_ld<T>(r, boff, 3 * ps);
  n.syy = buf_ld<T>(r, boff, 4 * ps);
  return n;
}

template <typename T>
__device__ __forceinline__ void store_row_node(T *__restrict__ row, long plane, unsigned boff,
                                               const Node<T> &u) {
  const __amdgpu_buffer_rsrc_t r = row_rsrc(row, plane);
  const int ps = (int)(plane * (long)sizeof(T));
  buf_st<T>(r, boff, 0, u.vx);
  buf_st<T>(r, boff, ps, u.vy);
  buf_st<T>(r, boff, 2 * ps, u.sxx);
  buf_st<T>(r, boff, 3 * ps, u.sxy);
  buf_st<T>(r, boff, 4 * ps, u.syy);
}

template <typename T>
__device__ __forceinline__ W4<T> inv_y(const Node<T> &n, const SweepCoef<T> &c) {
  return invariants(n.vy, n.syy, n.vx, n.sxy, c);
}

// ---------------------------------------------------------------------------
// Fused step: the x-sweep and the y-sweep of one time step in one pass over
// HBM (5 reads + 5 writes per node-update instead of twice that).
//
// A workgroup of FB lanes owns a window of FB consecutive columns starting two
// columns left of its strip of FB-4 output columns, and a segment of `seg`
// rows.  Lane t reads column x0-2+t (clamped into the grid), so the two halo
// columns on each side of the strip come from ordinary lanes: no extra loads,
// no divergent work.  Walking positions p = j0-2 .. j1+1 down the strip, each
// position's row (ypos-mapped) is read once and x-swept through LDS (every
// face of the strip in one pass over lanes 2..FB-2); each lane's post-x node
// then feeds the register y-march of k_sweep_y.  Only rows [j0, j1) of the
// FB-4 inner columns are written.  The y-march consumes x-swept values of the
// rows just outside the segment, recomputed here from the unswept input,
// which is exactly what the y-sweep of the two-kernel sequence reads.
//
// Workgroups are renumbered so that the tiles one XCD receives (the hardware
// deals workgroups to the 8 XCDs round-robin) are consecutive strips of the
// same segments: strips that share halo columns share an L2.
__device__ __forceinline__ int xcd_tile(int L, int N) {
  const int per = N >> 3, rem = N & 7, xcd = L & 7, k = L >> 3;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

template <typename T>
__device__ __forceinline__ void store_node(T *p, long plane, const Node<T> &u) {
  p[0] = u.vx;
  p[plane] = u.vy;
  p[2 * plane] = u.sxx;
  p[3 * plane] = u.sxy;
  p[4 * plane] = u.syy;
}

// Synchronisation of a window's LDS exchange: a workgroup barrier, or for a
// one-wavefront window (FB = 64) a wavefront-scope release/acquire fence
// pair around a wave barrier.  A wavefront's lanes run in lockstep and its
// LDS operations complete in order, so no wait and no s_barrier is needed;
// the workgroup-scope fence of __syncthreads would wait for every LDS write
// to complete and keep the compiler from overlapping independent exchanges.
template <int FB>
__device__ __forceinline__ void window_sync() {
  if constexpr (FB <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// x-sweep of one window row; returns the lane's post-x node (meaningful on
// lanes 2..FB-3).  sw[t] holds lane t's invariants, sf[t] the face between
// lanes t-1 and t.
template <typename T, int FB>
__device__ __forceinline__ Node<T> window_xsweep(Node<T> u, W4<T> *sw, W4<T> *sf, int t,
                                                 const SweepCoef<T> &c) {
  const W4<T> w = invariants(u.vx, u.sxx, u.vy, u.sxy, c);
  sw[t] = w;
  window_sync<FB>();
  if (t >= 2 && t <= FB - 2) sf[t] = faces(sw[t - 2], sw[t - 1], sw[t], sw[t + 1], c);
  window_sync<FB>();
  if (t >= 2 && t <= FB - 3) apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w, sf[t], sf[t + 1], c);
  return u;
}

// One tile of the fused step: the window of strip `strip` (lane t reads
// column strip*(FB-4)-2+t, clamped; lanes 2..FB-3 write) over output rows
// [j0, j1), marched downwards (DIR = +1, j increasing) or upwards (DIR = -1).
// Callers alternate the direction between neighbouring segments of a strip,
// so the two segments that read the halo rows around a shared boundary do so
// at about the same time (both start there, or both end there) and the
// second read hits L2.  Each face is the same operands and operations in
// either direction, so the result does not depend on DIR.  Used by the
// single-grid, material and batched kernels alike.
template <typename T, int EDGE, int FB, int DIR>
__device__ __forceinline__ void march_tile(const T *__restrict__ src, T *__restrict__ dst, int nx,
                                           int ny, long pitch, long plane, int strip, int j0,
                                           int j1, const SweepCoef<T> &cx,
                                           const SweepCoef<T> &cy, W4<T> *sw, W4<T> *sf) {
  const int t = threadIdx.x;
  const int x = strip * (FB - 4) - 2 + t;
  const unsigned boff = (unsigned)min(max(x, 0), nx - 1) * (unsigned)sizeof(T);
  const unsigned woff = (unsigned)max(x, 0) * (unsigned)sizeof(T);
  const bool writer = t >= 2 && t <= FB - 3 && x < nx;
#define ROW(p) load_row_node(src + (long)ypos<EDGE>((p), ny) * pitch, plane, boff)
#define XS(n) window_xsweep<T, FB>((n), sw, sf, t, cx)
  // Slots P0/P1/P2 hold the x-swept nodes of positions j, j+DIR, j+2*DIR
  // (their y-invariants are re-formed when needed, 6 operations each, rather
  // than held), FI is the face on the side already passed and NC the raw row
  // of position j+3*DIR.  Row j is written, then position j+3*DIR is x-swept
  // into P0 while NN receives the raw row of position j+4*DIR.  Unrolled six
  // times (slot period 3, face and prefetch period 2) so no register is
  // copied between steps.  wb keeps the invariants of position j-DIR (the
  // components its faces still read).
  Node<T> A, B, C, N0, N1;
  W4<T> F0, F1, wb;
  int j;
  if constexpr (DIR > 0) {
    const W4<T> w2 = inv_y(XS(ROW(j0 - 2)), cy);
    wb = inv_y(XS(ROW(j0 - 1)), cy);
    A = XS(ROW(j0));
    B = XS(ROW(j0 + 1));
    C = XS(ROW(j0 + 2));
    N0 = ROW(j0 + 3);
    F0 = faces(w2, wb, inv_y(A, cy), inv_y(B, cy), cy);  // face j0-1 | j0
    j = j0;
  } else {
    const W4<T> w2 = inv_y(XS(ROW(j1 + 1)), cy);
    wb = inv_y(XS(ROW(j1)), cy);
    A = XS(ROW(j1 - 1));
    B = XS(ROW(j1 - 2));
    C = XS(ROW(j1 - 3));
    N0 = ROW(j1 - 4);
    F0 = faces(inv_y(B, cy), inv_y(A, cy), wb, w2, cy);  // face j1-1 | j1
    j = j1 - 1;
  }
#define STEP(P0, P1, P2, FI, FO, NC, NN)                                     \
  {                                                                           \
    NN = ROW(j + 4 * DIR);                                                    \
    const W4<T> w0 = inv_y(P0, cy);                                           \
    if constexpr (DIR > 0) {                                                  \
      FO = faces(wb, w0, inv_y(P1, cy), inv_y(P2, cy), cy); /* j | j+1 */     \
      apply(P0.vy, P0.vx, P0.syy, P0.sxx, P0.sxy, w0, FI, FO, cy);            \
    } else {                                                                  \
      FO = faces(inv_y(P2, cy), inv_y(P1, cy), w0, wb, cy); /* j-1 | j */     \
      apply(P0.vy, P0.vx, P0.syy, P0.sxx, P0.sxy, w0, FO, FI, cy);            \
    }                                                                         \
    if (writer) store_row_node(dst + (long)j * pitch, plane, woff, P0);      \
    j += DIR;                                                                 \
    if (DIR > 0 ? j >= j1 : j < j0) break;                                    \
    if constexpr (DIR > 0) {                                                  \
      wb.a = w0.a;                                                            \
      wb.c = w0.c;                                                            \
    } else {                                                                  \
      wb.b = w0.b;                                                            \
      wb.d = w0.d;                                                            \
    }                                                                         \
    P0 = XS(NC);                                                              \
  }
  for (;;) {
    STEP(A, B, C, F0, F1, N0, N1)
    STEP(B, C, A, F1, F0, N1, N0)
    STEP(C, A, B, F0, F1, N0, N1)
    STEP(A, B, C, F1, F0, N1, N0)
    STEP(B, C, A, F0, F1, N0, N1)
    STEP(C, A, B, F1, F0, N1, N0)
  }
#undef STEP
#undef XS
#undef ROW
}

// Direction by segment index: even segments march down, odd ones up.
template <typename T, int EDGE, int FB>
__device__ __forceinline__ void march_tile_alt(const T *__restrict__ src, T *__restrict__ dst,
                                               int nx, int ny, long pitch, long plane, int strip,
                                               int j0, int j1, int segi,
                                               const SweepCoef<T> &cx, const SweepCoef<T> &cy,
                                               W4<T> *sw, W4<T> *sf) {
  if (segi & 1)
    march_tile<T, EDGE, FB, -1>(src, dst, nx, ny, pitch, plane, strip, j0, j1, cx, cy, sw, sf);
  else
    march_tile<T, EDGE, FB, 1>(src, dst, nx, ny, pitch, plane, strip, j0, j1, cx, cy, sw, sf);
}

// ---------------------------------------------------------------------------
// Two full steps per pass.
//
// A workgroup of FB lanes holds columns x0-4 .. x0+FB-5 and writes the FB-8
// columns x0 .. x0+FB-9 of rows [j0, j1) at time n+2.  Stage 1 is the fused
// step of time n (x-sweep of each raw row, register y-march) and emits the
// time-(n+1) rows j0-2 .. j1+1; stage 2 x-sweeps each of them as it appears
// and runs a second y-march that emits the time-(n+2) rows.  Raw rows
// j0-4 .. j1+5 are read once (clamped at the grid edges, as the reference
// clamps), so a pass moves 5 reads + 5 writes per node for two steps.
//
// Edges.  Each x-sweep reads its neighbour lanes through indices clamped
// into the lanes that hold grid columns, so both stages see the reference's
// column clamp whatever lanes outside the grid hold.  Stage-2 positions above
// row 0 or below row ny-1 take that edge row's stage-2 value (the reference's
// row clamp of the time-(n+1) state).  The operands and operations of every
// update are those of two single steps, so the result is bit-identical.
// sw and sf hold pair_lds<FB>() entries each (see there).

// Lane t+1's value on lane t (lane 63 receives 0): a wavefront shift of the
// data-parallel primitives (DPP wave_shl:1), no LDS round trip.
__device__ __forceinline__ int lane_next(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_next(float v) {
  return __builtin_bit_cast(float, lane_next(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double lane_next(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)lane_next((int)(unsigned)b);
  const unsigned hi = (unsigned)lane_next((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ W4<T> lane_next(const W4<T> &v) {
  W4<T> r;
  r.a = lane_next(v.a);
  r.b = lane_next(v.b);
  r.c = lane_next(v.c);
  r.d = lane_next(v.d);
  return r;
}
// Lane t-1's value on lane t (lane 0 receives 0): DPP wave_shr:1.
__device__ __forceinline__ float lane_prev(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}

// FP32 one-wavefront x-sweep faces with shared slopes.  The limiter's
// numerator (u0 - um) + eps of a face is, operand for operand, the
// denominator du + eps of the face upwind of it: the face on the left for
// the invariants moving to +x (a, c), the face on the right for those moving
// to -x (b, d).  Lane t forms its face's du and du + eps once from lanes nb.y
// and nb.z (du = w1 - w0 for a, c and w0 - w1 for b, d), and takes the
// numerators from lanes t-1 and t+1 by lane shifts: two subtractions per
// face and invariant instead of four, the same operations on the same
// operands.  Lanes 0 and 63 receive shifted zeros; no output reads their
// faces.
__device__ __forceinline__ W4<float> faces_shared(const W4<float> &w0, const W4<float> &w1,
                                                  const SweepCoef<float> &c) {
  const float eps = 1e-6f;
  const float da = w1.a - w0.a, db = w0.b - w1.b, dc = w1.c - w0.c, dd = w0.d - w1.d;
  const float sa = da + eps, sb = db + eps, sc = dc + eps, sd = dd + eps;
  W4<float> f;
  f.a = __builtin_fmaf(limiter(slope_ratio(lane_prev(sa), sa)) * c.hk1, da, w0.a);
  f.b = __builtin_fmaf(limiter(slope_ratio(lane_next(sb), sb)) * c.hk1, db, w1.b);
  f.c = __builtin_fmaf(limiter(slope_ratio(lane_prev(sc), sc)) * c.hk2, dc, w0.c);
  f.d = __builtin_fmaf(limiter(slope_ratio(lane_next(sd), sd)) * c.hk2, dd, w1.d);
  return f;
}

// x-sweep of one window row with clamped neighbour lanes: the face between
// lanes t-1 and t reads lanes nb = (t-2, t-1, t, t+1), each clamped into the
// lanes holding grid columns.  Meaningful on lanes 2..FB-3.
// FP32: every lane forms a face and applies (lanes outside 2..FB-3 produce
// values no output reads; their LDS indices stay in the window), so the
// x-sweep has no divergent branch (1.9% faster two-step pass).  FP64 keeps
// the branches: without them the one-wave pass spills.
template <typename T, int FB>
__device__ __forceinline__ Node<T> window_xsweep_cl(Node<T> u, W4<T> *sw, W4<T> *sf, int t,
                                                    const int4 &nb, const SweepCoef<T> &c) {
  const W4<T> w = invariants(u.vx, u.sxx, u.vy, u.sxy, c);
  sw[t] = w;
  window_sync<FB>();
  if constexpr (FB == 64) {
    // One-wavefront window: lane t+1's face reaches lane t by a lane shift
    // (one LDS round trip fewer; FP64 8192^2 two-step pass 1.5% faster,
    // 2048^2 3-6%).  FP64 forms faces on lanes 2..FB-2 only (branch-free it
    // spills); the others shift zeros that no apply reads.
    if constexpr (sizeof(T) == 4) {
      const W4<T> f = faces_shared(sw[nb.y], sw[nb.z], c);
      apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w, f, lane_next(f), c);
    } else {
      W4<T> f = {T(0), T(0), T(0), T(0)};
      if (t >= 2 && t <= FB - 2) f = faces(sw[nb.x], sw[nb.y], sw[nb.z], sw[nb.w], c);
      const W4<T> fr = lane_next(f);
      if (t >= 2 && t <= FB - 3) apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w, f, fr, c);
    }
    // the next x-sweep of this window rewrites sw only after every lane read it
    window_sync<FB>();
  } else if constexpr (sizeof(T) == 4) {
    sf[t] = faces(sw[nb.x], sw[nb.y], sw[nb.z], sw[nb.w], c);
    window_sync<FB>();
    apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w, sf[t], sf[min(t + 1, FB - 1)], c);
  } else {
    if (t >= 2 && t <= FB - 2) sf[t] = faces(sw[nb.x], sw[nb.y], sw[nb.z], sw[nb.w], c);
    window_sync<FB>();
    if (t >= 2 && t <= FB - 3) apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w, sf[t], sf[t + 1], c);
  }
  return u;
}

// LDS entries (W4) of each of march_pair's two arrays: a one-wavefront window
// gives the stage-2 x-sweeps their own half (no barrier orders the two
// exchanges of a step, so separate arrays let them overlap); wider windows,
// whose barriers serialise the exchanges anyway, share one (measured 2.8%
// faster for 256 lanes in FP64).
template <int FB>
constexpr int pair_lds() { return FB <= 64 ? 2 * FB : FB; }

template <typename T, int FB>
__device__ __forceinline__ void march_pair(const T *__restrict__ src, T *__restrict__ dst, int nx,
                                           int ny, long pitch, long plane, int strip, int j0,
                                           int j1, const SweepCoef<T> &cx,
                                           const SweepCoef<T> &cy, W4<T> *sw, W4<T> *sf) {
  const int t = threadIdx.x;
  const int x0 = strip * (FB - 8);
  const int x = x0 - 4 + t;
  const unsigned boff = (unsigned)min(max(x, 0), nx - 1) * (unsigned)sizeof(T);
  const unsigned woff = (unsigned)max(x, 0) * (unsigned)sizeof(T);
  const bool writer = t >= 4 && t <= FB - 5 && x < nx;
  // lanes of grid column 0 and nx-1 (clamped into the window)
  const int tlo = max(0, 4 - x0), thi = min(FB - 1, nx + 3 - x0);
  const int4 nb = make_int4(min(max(t - 2, tlo), thi), min(max(t - 1, tlo), thi),
                            min(max(t, tlo), thi), min(max(t + 1, tlo), thi));
  constexpr int PAIR_LDS2 = pair_lds<FB>() / FB - 1;  // 1: stage 2 has its own half
#define ROW(p) load_row_node(src + (long)ypos<EDGE_CLAMP>((p), ny) * pitch, plane, boff)
#define XS(n) window_xsweep_cl<T, FB>((n), sw, sf, t, nb, cx)
#define XS2(n) window_xsweep_cl<T, FB>((n), sw + PAIR_LDS2 * FB, sf + PAIR_LDS2 * FB, t, nb, cx)
  // Stage 1: slots A/B/C hold the x-swept time-n rows q, q+1, q+2, F0/F1 the
  // faces, wb the invariants of row q-1, N0/N1 the raw rows in flight (as in
  // march_tile).  Stage 2: slots A2/B2/C2 hold the x-swept time-(n+1) rows
  // r, r+1, r+2 (r = q-2), G0/G1 its faces, vb the invariants of row r-1.
  // HOLD (FP32, which has the registers): the y-invariants of every slot are
  // kept (WA.., VA..) instead of re-formed each step.
  constexpr bool HOLD = sizeof(T) == 4;
  Node<T> A, B, C, N0, N1, A2, B2, C2, Xm2, Xm1;
  W4<T> F0, F1, wb, G0, G1, vb, WA, WB, WC, VA, VB, VC;
  int q = j0 - 2;
  {
    const W4<T> w2 = inv_y(XS(ROW(q - 2)), cy);
    wb = inv_y(XS(ROW(q - 1)), cy);
    A = XS(ROW(q));
    B = XS(ROW(q + 1));
    C = XS(ROW(q + 2));
    N0 = ROW(q + 3);
    WA = inv_y(A, cy);
    WB = inv_y(B, cy);
    if constexpr (HOLD) WC = inv_y(C, cy);
    F0 = faces(w2, wb, WA, WB, cy);  // face q-1 | q
  }
  // Stage-1 step: time-(n+1) row q (in P0), its x-sweep into X, then the next
  // raw row x-swept into P0.
#define S1(P0, P1, P2, W0, W1, W2, FI, FO, NC, NN, X)                        \
  {                                                                           \
    NN = ROW(q + 4);                                                          \
    const W4<T> w0 = HOLD ? W0 : inv_y(P0, cy);                               \
    FO = faces(wb, w0, HOLD ? W1 : inv_y(P1, cy), HOLD ? W2 : inv_y(P2, cy), cy); \
    apply(P0.vy, P0.vx, P0.syy, P0.sxx, P0.sxy, w0, FI, FO, cy);              \
    wb.a = w0.a;                                                              \
    wb.c = w0.c;                                                              \
    X = XS2(P0);                                                              \
    P0 = XS(NC);                                                              \
    if constexpr (HOLD) W0 = inv_y(P0, cy);                                   \
    ++q;                                                                      \
  }
  S1(A, B, C, WA, WB, WC, F0, F1, N0, N1, Xm2)  // q = j0-2
  S1(B, C, A, WB, WC, WA, F1, F0, N1, N0, Xm1)  // q = j0-1
  S1(C, A, B, WC, WA, WB, F0, F1, N0, N1, A2)   // q = j0
  S1(A, B, C, WA, WB, WC, F1, F0, N1, N0, B2)   // q = j0+1
  if (j0 + 1 >= ny) B2 = A2;                    // row clamp of the time-(n+1) state
  if (j0 - 1 < 0) Xm1 = A2;
  if (j0 - 2 < 0) Xm2 = Xm1;
  vb = inv_y(Xm1, cy);
  VA = inv_y(A2, cy);
  VB = inv_y(B2, cy);
  G0 = faces(inv_y(Xm2, cy), vb, VA, VB, cy);  // face j0-1 | j0
  int r = j0;
  // One stage-1 step (row q = r+2 of time n+1 into Q2, clamped past the last
  // row), then one stage-2 step writing row r of time n+2.
#define S2(P0, P1, P2, W0, W1, W2, FI, FO, NC, NN, Q0, Q1, Q2, V0, V1, V2, GI, GO) \
  {                                                                           \
    NN = ROW(q + 4);                                                          \
    const W4<T> w0 = HOLD ? W0 : inv_y(P0, cy);                               \
    FO = faces(wb, w0, HOLD ? W1 : inv_y(P1, cy), HOLD ? W2 : inv_y(P2, cy), cy); \
    apply(P0.vy, P0.vx, P0.syy, P0.sxx, P0.sxy, w0, FI, FO, cy);              \
    wb.a = w0.a;                                                              \
    wb.c = w0.c;                                                              \
    {                                                                         \
      const Node<T> xn = XS2(P0);                                             \
      if (q >= ny) {                                                          \
        Q2 = Q1;                                                              \
        if constexpr (HOLD) V2 = V1;                                          \
      } else {                                                                \
        Q2 = xn;                                                              \
        if constexpr (HOLD) V2 = inv_y(Q2, cy);                               \
      }                                                                       \
    }                                                                         \
    P0 = XS(NC);                                                              \
    if constexpr (HOLD) W0 = inv_y(P0, cy);                                   \
    ++q;                                                                      \
    const W4<T> v0 = HOLD ? V0 : inv_y(Q0, cy);                               \
    GO = faces(vb, v0, HOLD ? V1 : inv_y(Q1, cy), HOLD ? V2 : inv_y(Q2, cy), cy); \
    apply(Q0.vy, Q0.vx, Q0.syy, Q0.sxx, Q0.sxy, v0, GI, GO, cy);              \
    if (writer) store_row_node(dst + (long)r * pitch, plane, woff, Q0);      \
    if (++r >= j1) break;                                                     \
    vb.a = v0.a;                                                              \
    vb.c = v0.c;                                                              \
  }
  // stage 1 continues at phase 4 of its six-step slot cycle, stage 2 at 0
  for (;;) {
    S2(B, C, A, WB, WC, WA, F0, F1, N0, N1, A2, B2, C2, VA, VB, VC, G0, G1)
    S2(C, A, B, WC, WA, WB, F1, F0, N1, N0, B2, C2, A2, VB, VC, VA, G1, G0)
    S2(A, B, C, WA, WB, WC, F0, F1, N0, N1, C2, A2, B2, VC, VA, VB, G0, G1)
    S2(B, C, A, WB, WC, WA, F1, F0, N1, N0, A2, B2, C2, VA, VB, VC, G1, G0)
    S2(C, A, B, WC, WA, WB, F0, F1, N0, N1, B2, C2, A2, VB, VC, VA, G0, G1)
    S2(A, B, C, WA, WB, WC, F1, F0, N1, N0, C2, A2, B2, VC, VA, VB, G1, G0)
  }
#undef S2
#undef S1
#undef XS2
#undef XS
#undef ROW
}

// ---------------------------------------------------------------------------
// Two full steps per pass with the two stages on two wavefronts.
//
// The window, segment, edge rules and operations of march_pair with a
// 64-lane window, but stage 1 (x-sweep of the raw rows, y-march to time
// n+1) runs on wavefront 0 and stage 2 (x-sweep of the time-(n+1) rows,
// y-march to time n+2) on wavefront 1 of a 128-lane workgroup.  Each time-
// (n+1) row passes from the first to the second through an LDS ring of 2*RB
// slots, one workgroup barrier per RB rows: wavefront 0 writes row k into slot
// k % (2*RB) and arrives at barrier k / RB after the last row of that group
// (or the last row of all); wavefront 1 passes barrier b before it reads rows
// b*RB .. b*RB+RB-1, and passes barrier b+1 only after reading them, while
// wavefront 0 fills the other half of the ring; so a slot is rewritten only
// after its row was read.  Both wavefronts pass exactly ceil((j1 - j0 + 4) /
// RB) barriers (stage-1 rows j0-2 .. j1+1).  Each wavefront holds one stage's
// march state, about
// half of march_pair's registers, so more of them are resident per SIMD; each
// x-sweep's LDS exchange stays within its wavefront (wavefront-scope
// ordering) in its own half of sw / sf.  The result is march_pair's, bit for
// bit.  sw and sf hold 128 entries each, ring 2*RB x 5 x 64 values.
__device__ __forceinline__ void split_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename T>
__device__ __forceinline__ void ring_put(T *slot, int t, const Node<T> &u) {
  slot[t] = u.vx;
  slot[64 + t] = u.vy;
  slot[128 + t] = u.sxx;
  slot[192 + t] = u.sxy;
  slot[256 + t] = u.syy;
}

template <typename T>
__device__ __forceinline__ Node<T> ring_get(const T *slot, int t) {
  Node<T> u;
  u.vx = slot[t];
  u.vy = slot[64 + t];
  u.sxx = slot[128 + t];
  u.sxy = slot[192 + t];
  u.syy = slot[256 + t];
  return u;
}

// Rows per barrier of the ring (RB): 2 and 3 measured no faster than 1.
constexpr int kSplitRows = 1;

template <typename T, bool HOLD, int RB = kSplitRows>
__device__ __forceinline__ void march_pair_split(const T *__restrict__ src, T *__restrict__ dst,
                                                 int nx, int ny, long pitch, long plane,
                                                 int strip, int j0, int j1,
                                                 const SweepCoef<T> &cx, const SweepCoef<T> &cy,
                                                 W4<T> *sw, W4<T> *sf, T *ring) {
  constexpr int FB = 64;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int t = (int)(threadIdx.x & 63);
  const int x0 = strip * (FB - 8);
  const int x = x0 - 4 + t;
  const int tlo = max(0, 4 - x0), thi = min(FB - 1, nx + 3 - x0);
  const int4 nb = make_int4(min(max(t - 2, tlo), thi), min(max(t - 1, tlo), thi),
                            min(max(t, tlo), thi), min(max(t + 1, tlo), thi));
  const int rows = j1 - j0 + 4;  // stage-1 rows j0-2 .. j1+1
  constexpr int RS = 2 * RB;     // ring slots
  sw += FB * wave;
  sf += FB * wave;
#define XS(n) window_xsweep_cl<T, FB>((n), sw, sf, t, nb, cx)
  if (wave == 0) {
    // Stage 1 (as march_pair's): slots A/B/C hold the x-swept time-n rows
    // q, q+1, q+2, F0/F1 the faces, wb the invariants of row q-1, N0/N1 the
    // raw rows in flight; HOLD keeps each slot's y-invariants (WA..).
    const unsigned boff = (unsigned)min(max(x, 0), nx - 1) * (unsigned)sizeof(T);
#define ROW(p) load_row_node(src + (long)ypos<EDGE_CLAMP>((p), ny) * pitch, plane, boff)
    Node<T> A, B, C, N0, N1;
    W4<T> F0, F1, wb, WA, WB, WC;
    int q = j0 - 2, k = 0;
    {
      const W4<T> w2 = inv_y(XS(ROW(q - 2)), cy);
      wb = inv_y(XS(ROW(q - 1)), cy);
      A = XS(ROW(q));
      B = XS(ROW(q + 1));
      C = XS(ROW(q + 2));
      N0 = ROW(q + 3);
      WA = inv_y(A, cy);
      WB = inv_y(B, cy);
      if constexpr (HOLD) WC = inv_y(C, cy);
      F0 = faces(w2, wb, WA, WB, cy);  // face q-1 | q
    }
#define P1(P0, P1, P2, W0, W1, W2, FI, FO, NC, NN)                           \
  {                                                                           \
    NN = ROW(q + 4);                                                          \
    const W4<T> w0 = HOLD ? W0 : inv_y(P0, cy);                               \
    FO = faces(wb, w0, HOLD ? W1 : inv_y(P1, cy), HOLD ? W2 : inv_y(P2, cy), cy); \
    apply(P0.vy, P0.vx, P0.syy, P0.sxx, P0.sxy, w0, FI, FO, cy);              \
    wb.a = w0.a;                                                              \
    wb.c = w0.c;                                                              \
    ring_put(ring + (k % RS) * 5 * FB, t, P0);                                \
    if (k % RB == RB - 1 || k == rows - 1) split_sync();                      \
    if (++k >= rows) break;                                                   \
    P0 = XS(NC);                                                              \
    if constexpr (HOLD) W0 = inv_y(P0, cy);                                   \
    ++q;                                                                      \
  }
    for (;;) {
      P1(A, B, C, WA, WB, WC, F0, F1, N0, N1)
      P1(B, C, A, WB, WC, WA, F1, F0, N1, N0)
      P1(C, A, B, WC, WA, WB, F0, F1, N0, N1)
      P1(A, B, C, WA, WB, WC, F1, F0, N1, N0)
      P1(B, C, A, WB, WC, WA, F0, F1, N0, N1)
      P1(C, A, B, WC, WA, WB, F1, F0, N1, N0)
    }
#undef P1
#undef ROW
  } else {
    // Stage 2: slots A2/B2/C2 hold the x-swept time-(n+1) rows r, r+1, r+2,
    // G0/G1 the faces, vb the invariants of row r-1.  Positions above row 0
    // or below row ny-1 take that edge row's value (march_pair's clamps).
    const unsigned woff = (unsigned)max(x, 0) * (unsigned)sizeof(T);
    const bool writer = t >= 4 && t <= FB - 5 && x < nx;
    Node<T> A2, B2, C2, Xm2, Xm1;
    W4<T> G0, G1, vb, VA, VB, VC;
    int k = 0;
    // barrier before each group of RB rows, then row k from its slot
#define GET() (k % RB == 0 ? split_sync() : (void)0, XS(ring_get(ring + ((k++) % RS) * 5 * FB, t)))
    Xm2 = GET();  // row j0-2
    Xm1 = GET();  // row j0-1
    A2 = GET();   // row j0
    B2 = GET();   // row j0+1
    if (j0 + 1 >= ny) B2 = A2;
    if (j0 - 1 < 0) Xm1 = A2;
    if (j0 - 2 < 0) Xm2 = Xm1;
    vb = inv_y(Xm1, cy);
    VA = inv_y(A2, cy);
    VB = inv_y(B2, cy);
    G0 = faces(inv_y(Xm2, cy), vb, VA, VB, cy);  // face j0-1 | j0
    int r = j0, q = j0 + 2;
#define P2(Q0, Q1, Q2, V0, V1, V2, GI, GO)                                    \
  {                                                                           \
    {                                                                         \
      const Node<T> xn = GET();                                               \
      if (q >= ny) {                                                          \
        Q2 = Q1;                                                              \
        if constexpr (HOLD) V2 = V1;                                          \
      } else {                                                                \
        Q2 = xn;                                                              \
        if constexpr (HOLD) V2 = inv_y(Q2, cy);                               \
      }                                                                       \
    }                                                                         \
    ++q;                                                                      \
    const W4<T> v0 = HOLD ? V0 : inv_y(Q0, cy);                               \
    GO = faces(vb, v0, HOLD ? V1 : inv_y(Q1, cy), HOLD ? V2 : inv_y(Q2, cy), cy); \
    apply(Q0.vy, Q0.vx, Q0.syy, Q0.sxx, Q0.sxy, v0, GI, GO, cy);              \
    if (writer) store_row_node(dst + (long)r * pitch, plane, woff, Q0);      \
    if (++r >= j1) break;                                                     \
    vb.a = v0.a;                                                              \
    vb.c = v0.c;                                                              \
  }
    for (;;) {
      P2(A2, B2, C2, VA, VB, VC, G0, G1)
      P2(B2, C2, A2, VB, VC, VA, G1, G0)
      P2(C2, A2, B2, VC, VA, VB, G0, G1)
      P2(A2, B2, C2, VA, VB, VC, G1, G0)
      P2(B2, C2, A2, VB, VC, VA, G0, G1)
      P2(C2, A2, B2, VC, VA, VB, G1, G0)
    }
#undef P2
#undef GET
  }
#undef XS
}

// ---------------------------------------------------------------------------
// Heterogeneous media: node-local march for tiles that straddle a material
// interface.
//
// The reference has one material per task (simple/le_core.h:168-186) and no
// interface rule.  Here every node is updated entirely in its own material:
// the five stencil nodes of a sweep are projected on that node's
// eigenvectors, advected with that node's Courant numbers and projected back
// with that node's constants.  Both faces of a node are therefore evaluated
// by the node itself (no sharing across an interface); with one material the
// operands and operations are exactly those of march_tile, so the bits agree.
// `ids` holds one material index per node (pitched rows of `idp` bytes);
// cxt/cyt are the per-material constants of the x- and y-sweep.
template <typename T, int EDGE, int FB>
__device__ __forceinline__ void march_tile_mixed(const T *__restrict__ src, T *__restrict__ dst,
                                                 const unsigned char *__restrict__ ids, long idp,
                                                 int nx, int ny, long pitch, long plane, int strip,
                                                 int j0, int j1,
                                                 const SweepCoef<T> *__restrict__ cxt,
                                                 const SweepCoef<T> *__restrict__ cyt,
                                                 Node<T> *sn) {
  const int t = threadIdx.x;
  const int x = strip * (FB - 4) - 2 + t;
  const int xc = min(max(x, 0), nx - 1);
  const unsigned boff = (unsigned)xc * (unsigned)sizeof(T);
  const unsigned woff = (unsigned)max(x, 0) * (unsigned)sizeof(T);
  const bool inner = t >= 2 && t <= FB - 3;
  const bool writer = inner && x < nx;

  // x-sweep of the row of position p; returns this lane's post-x node.
  auto xrow = [&](int p) -> Node<T> {
    const int row = ypos<EDGE>(p, ny);
    Node<T> u = load_row_node(src + (long)row * pitch, plane, boff);
    const int m = ids[(long)row * idp + xc];
    sn[t] = u;
    window_sync<FB>();
    if (inner) {
      const SweepCoef<T> c = cxt[m];
      W4<T> w[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const Node<T> &v = sn[t - 2 + q];
        w[q] = invariants(v.vx, v.sxx, v.vy, v.sxy, c);
      }
      const W4<T> fl = faces(w[0], w[1], w[2], w[3], c);
      const W4<T> fr = faces(w[1], w[2], w[3], w[4], c);
      apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w[2], fl, fr, c);
    }
    window_sync<FB>();
    return u;
  };

  // y-march over positions; P[q] holds the post-x node of position j-2+q
  Node<T> P[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) P[q] = xrow(j0 - 2 + q);
  for (int j = j0;;) {
    const SweepCoef<T> c = cyt[ids[(long)j * idp + xc]];
    W4<T> w[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) w[q] = inv_y(P[q], c);
    const W4<T> fl = faces(w[0], w[1], w[2], w[3], c);
    const W4<T> fr = faces(w[1], w[2], w[3], w[4], c);
    Node<T> u = P[2];
    apply(u.vy, u.vx, u.syy, u.sxx, u.sxy, w[2], fl, fr, c);
    if (writer) store_row_node(dst + (long)j * pitch, plane, woff, u);
    if (++j >= j1) break;
#pragma unroll
    for (int q = 0; q < 4; ++q) P[q] = P[q + 1];
    P[4] = xrow(j + 2);
  }
}

// Two full steps of the node-local scheme over a tile that straddles a
// material interface: the strip geometry, edge rules and stage order of
// march_pair, with every node's update in its own material (as
// march_tile_mixed).  Rows are held in shift registers: few tiles take this
// path.  sn has FB entries.
template <typename T, int FB>
__device__ __forceinline__ void march_pair_mixed(const T *__restrict__ src, T *__restrict__ dst,
                                                 const unsigned char *__restrict__ ids, long idp,
                                                 int nx, int ny, long pitch, long plane, int strip,
                                                 int j0, int j1,
                                                 const SweepCoef<T> *__restrict__ cxt,
                                                 const SweepCoef<T> *__restrict__ cyt,
                                                 Node<T> *sn) {
  const int t = threadIdx.x;
  const int x0 = strip * (FB - 8);
  const int x = x0 - 4 + t;
  const int xc = min(max(x, 0), nx - 1);
  const unsigned boff = (unsigned)xc * (unsigned)sizeof(T);
  const unsigned woff = (unsigned)max(x, 0) * (unsigned)sizeof(T);
  const bool writer = t >= 4 && t <= FB - 5 && x < nx;
  const int tlo = max(0, 4 - x0), thi = min(FB - 1, nx + 3 - x0);
  int nb[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) nb[k] = min(max(t - 2 + k, tlo), thi);
  const bool inner = t >= 2 && t <= FB - 3;

  // node-local x-sweep of row `row` (u: this lane's node of it)
  auto xsweep = [&](Node<T> u, int row) -> Node<T> {
    const int m = ids[(long)row * idp + xc];
    sn[t] = u;
    window_sync<FB>();
    if (inner) {
      const SweepCoef<T> c = cxt[m];
      W4<T> w[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const Node<T> &v = sn[nb[k]];
        w[k] = invariants(v.vx, v.sxx, v.vy, v.sxy, c);
      }
      const W4<T> fl = faces(w[0], w[1], w[2], w[3], c);
      const W4<T> fr = faces(w[1], w[2], w[3], w[4], c);
      apply(u.vx, u.vy, u.sxx, u.syy, u.sxy, w[2], fl, fr, c);
    }
    window_sync<FB>();
    return u;
  };
  // node-local y-update of row j from the x-swept rows j-2 .. j+2
  auto yupdate = [&](const Node<T> *P, int j) -> Node<T> {
    const SweepCoef<T> c = cyt[ids[(long)j * idp + xc]];
    W4<T> w[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = inv_y(P[k], c);
    const W4<T> fl = faces(w[0], w[1], w[2], w[3], c);
    const W4<T> fr = faces(w[1], w[2], w[3], w[4], c);
    Node<T> u = P[2];
    apply(u.vy, u.vx, u.syy, u.sxx, u.sxy, w[2], fl, fr, c);
    return u;
  };
  auto raw = [&](int p) -> Node<T> {
    return load_row_node(src + (long)ypos<EDGE_CLAMP>(p, ny) * pitch, plane, boff);
  };

  // stage 1: P[k] = x-swept time-n row at position q-2+k (clamped rows)
  Node<T> P[5], Q[5];
  int q = j0 - 2;
#pragma unroll
  for (int k = 0; k < 5; ++k) P[k] = xsweep(raw(q - 2 + k), ypos<EDGE_CLAMP>(q - 2 + k, ny));
  // time-(n+1) row q x-swept for stage 2 (meaningful for 0 <= q < ny)
  auto next = [&]() -> Node<T> {
    const int row = min(max(q, 0), ny - 1);
    const Node<T> y1 = yupdate(P, row);
    const Node<T> x2 = xsweep(y1, row);
#pragma unroll
    for (int k = 0; k < 4; ++k) P[k] = P[k + 1];
    P[4] = xsweep(raw(q + 3), ypos<EDGE_CLAMP>(q + 3, ny));
    ++q;
    return x2;
  };
  // stage 2: Q[k] = x-swept time-(n+1) row at position r-2+k, clamped
#pragma unroll
  for (int k = 0; k < 4; ++k) Q[k] = next();  // positions j0-2 .. j0+1
  Q[3] = pick_node(j0 + 1 >= ny, Q[2], Q[3]);
  Q[1] = pick_node(j0 - 1 < 0, Q[2], Q[1]);
  Q[0] = pick_node(j0 - 2 < 0, Q[1], Q[0]);
  for (int r = j0; r < j1; ++r) {
    const Node<T> xn = next();  // position r+2
    Q[4] = pick_node(r + 2 >= ny, Q[3], xn);
    const Node<T> u = yupdate(Q, r);
    if (writer) store_row_node(dst + (long)r * pitch, plane, woff, u);
#pragma unroll
    for (int k = 0; k < 4; ++k) Q[k] = Q[k + 1];
  }
}

}  // namespace le2d
