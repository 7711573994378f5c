// media.hip -- heterogeneous media (a per-node material plane) for the fused
// step, and the FP64 energy reduction.
//
// A material plane splits the step's tiles in two tables on the host, once:
// tiles whose whole read window (the 256-, 128- or 64-column window and the
// ypos-mapped rows j0-2 .. j1+1) holds a single material run march_tile with
// that material's constants; the few tiles that straddle an interface run
// march_tile_mixed (node-local scheme, sweep_dev.h).  Both kernels read the
// same input buffer and write disjoint output tiles.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "sweep_dev.h"

namespace le2d {

template <typename T, int EDGE, int FB>
__global__ __launch_bounds__(FB, 3) void k_step_mat_homog(const T *__restrict__ src,
                                                       T *__restrict__ dst, int nx, int ny,
                                                       long pitch, long plane,
                                                       const MatTile *__restrict__ tiles,
                                                       const SweepCoef<T> *__restrict__ cxt,
                                                       const SweepCoef<T> *__restrict__ cyt) {
  __shared__ W4<T> sw[FB];
  __shared__ W4<T> sf[FB];
  const MatTile tj = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const SweepCoef<T> cx = cxt[tj.mat], cy = cyt[tj.mat];
  march_tile_alt<T, EDGE, FB>(src, dst, nx, ny, pitch, plane, tj.strip, tj.j0, tj.j1, tj.seg, cx,
                              cy, sw, sf);
}

template <typename T, int EDGE, int FB>
__global__ __launch_bounds__(FB) void k_step_mat_mixed(const T *__restrict__ src,
                                                       T *__restrict__ dst, int nx, int ny,
                                                       long pitch, long plane,
                                                       const MatTile *__restrict__ tiles,
                                                       const unsigned char *__restrict__ ids,
                                                       long idp,
                                                       const SweepCoef<T> *__restrict__ cxt,
                                                       const SweepCoef<T> *__restrict__ cyt) {
  __shared__ Node<T> sn[FB];
  const MatTile tj = tiles[blockIdx.x];
  march_tile_mixed<T, EDGE, FB>(src, dst, ids, idp, nx, ny, pitch, plane, tj.strip, tj.j0, tj.j1,
                                cxt, cyt, sn);
}

// Two-step pass: single-material tiles run march_pair with their material's
// constants, the others the node-local two-step march; both in windows of FB
// lanes (strips of FB-8 columns), the grid's two-step window width.
template <typename T, int FB>
__global__ __launch_bounds__(FB, 2) void k_step2_mat_homog(const T *__restrict__ src,
                                                         T *__restrict__ dst, int nx, int ny,
                                                         long pitch, long plane,
                                                         const MatTile *__restrict__ tiles,
                                                         const SweepCoef<T> *__restrict__ cxt,
                                                         const SweepCoef<T> *__restrict__ cyt) {
  __shared__ W4<T> sw[pair_lds<FB>()];
  __shared__ W4<T> sf[pair_lds<FB>()];
  const MatTile tj = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const SweepCoef<T> cx = cxt[tj.mat], cy = cyt[tj.mat];
  march_pair<T, FB>(src, dst, nx, ny, pitch, plane, tj.strip, tj.j0, tj.j1, cx, cy, sw, sf);
}

// The same for 64-lane windows with the two stages on two wavefronts
// (march_pair_split).
template <typename T>
__global__ __launch_bounds__(128, 3) void k_step2_mat_homog_split(
    const T *__restrict__ src, T *__restrict__ dst, int nx, int ny, long pitch, long plane,
    const MatTile *__restrict__ tiles, const SweepCoef<T> *__restrict__ cxt,
    const SweepCoef<T> *__restrict__ cyt) {
  __shared__ W4<T> sw[128];
  __shared__ W4<T> sf[128];
  __shared__ T ring[2 * kSplitRows * 5 * 64];
  const MatTile tj = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const SweepCoef<T> cx = cxt[tj.mat], cy = cyt[tj.mat];
  march_pair_split<T, sizeof(T) == 4>(src, dst, nx, ny, pitch, plane, tj.strip, tj.j0, tj.j1, cx,
                                      cy, sw, sf, ring);
}

// Interface tiles of the two-step pass, launched beside the single-material
// kernel.  At 64 lanes that is the two-wavefront form, three waves of at most
// 168 VGPRs per SIMD: held to the same bound, a wave of this kernel starts as
// soon as one neighbouring wave ends (128 lanes keep the bound too, without
// spilling).  At 256 lanes two waves per SIMD keep it free of spills.
template <typename T, int FB>
__global__ __launch_bounds__(FB, FB == 256 ? 2 : 3) void k_step2_mat_mixed(
    const T *__restrict__ src, T *__restrict__ dst, int nx, int ny, long pitch, long plane,
    const MatTile *__restrict__ tiles, const unsigned char *__restrict__ ids, long idp,
    const SweepCoef<T> *__restrict__ cxt, const SweepCoef<T> *__restrict__ cyt) {
  __shared__ Node<T> sn[FB];
  const MatTile tj = tiles[blockIdx.x];
  march_pair_mixed<T, FB>(src, dst, ids, idp, nx, ny, pitch, plane, tj.strip, tj.j0, tj.j1, cxt,
                          cyt, sn);
}

// Energy density in FP64 (plane strain, isotropic):
//   rho/2 |v|^2 + (sxx+syy)^2 / (8 (lambda+mu)) + ((sxx-syy)^2/2 + 2 sxy^2) / (4 mu)
// summed over the grid; one FP64 partial per workgroup, summed on the host in
// workgroup order (deterministic).  mtab holds (rho, lambda, mu) per material.
template <typename T>
__global__ __launch_bounds__(256) void k_energy(const T *__restrict__ g, int nx, int ny, long pitch,
                                                long plane, const unsigned char *__restrict__ ids,
                                                long idp, const double *__restrict__ mtab,
                                                double *__restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const long n = (long)nx * ny;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long j = k / nx, i = k - j * nx;
    const long o = j * pitch + i;
    const double vx = g[o], vy = g[plane + o], sxx = g[2 * plane + o], sxy = g[3 * plane + o],
                 syy = g[4 * plane + o];
    const int m = ids ? ids[j * idp + i] : 0;
    const double rho = mtab[3 * m], lam = mtab[3 * m + 1], mu = mtab[3 * m + 2];
    const double tr = sxx + syy, dv = sxx - syy;
    acc += 0.5 * rho * (vx * vx + vy * vy) + tr * tr / (8.0 * (lam + mu)) +
           (0.5 * dv * dv + 2.0 * sxy * sxy) / (4.0 * mu);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Differences of two states in FP64: per workgroup the partial sums of
// (a-b)^2 and b^2 and the maxima of |a-b| and |b| over all five fields.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_diff_norms(const TA *__restrict__ a, long pa, long qa,
                                                    const TB *__restrict__ b, long pb, long qb,
                                                    int nx, int ny, double *__restrict__ partial) {
  __shared__ double red[4][256];
  double s2 = 0, r2 = 0, sm = 0, rm = 0;
  const long n = (long)nx * ny;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long j = k / nx, i = k - j * nx;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double x = (double)a[q * qa + j * pa + i], y = (double)b[q * qb + j * pb + i];
      const double d = x - y;
      s2 += d * d;
      r2 += y * y;
      sm = fmax(sm, fabs(d));
      rm = fmax(rm, fabs(y));
    }
  }
  const int t = threadIdx.x;
  red[0][t] = s2;
  red[1][t] = r2;
  red[2][t] = sm;
  red[3][t] = rm;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] += red[0][t + s];
      red[1][t] += red[1][t + s];
      red[2][t] = fmax(red[2][t], red[2][t + s]);
      red[3][t] = fmax(red[3][t], red[3][t + s]);
    }
    __syncthreads();
  }
  if (t == 0)
    for (int c = 0; c < 4; ++c) partial[4 * blockIdx.x + c] = red[c][0];
}

// Workgroups of the norm kernel (it strides over any grid).
constexpr int kNormBlocks = 1024;

int diff_norms(const void *a, int ea, long pa, long qa, const void *b, int eb, long pb, long qb,
               int nx, int ny, hipStream_t stream, double out[4]) {
  const int blocks = kNormBlocks;
  double *partial = nullptr;
  const int arc = dev_alloc((void **)&partial, 4 * blocks * sizeof(double), "diff_norms");
  if (arc) return arc;
  dim3 g(blocks), bl(256);
  if (ea == 8 && eb == 8)
    hipLaunchKernelGGL((k_diff_norms<double, double>), g, bl, 0, stream, (const double *)a, pa, qa,
                       (const double *)b, pb, qb, nx, ny, partial);
  else if (ea == 4 && eb == 8)
    hipLaunchKernelGGL((k_diff_norms<float, double>), g, bl, 0, stream, (const float *)a, pa, qa,
                       (const double *)b, pb, qb, nx, ny, partial);
  else if (ea == 8 && eb == 4)
    hipLaunchKernelGGL((k_diff_norms<double, float>), g, bl, 0, stream, (const double *)a, pa, qa,
                       (const float *)b, pb, qb, nx, ny, partial);
  else
    hipLaunchKernelGGL((k_diff_norms<float, float>), g, bl, 0, stream, (const float *)a, pa, qa,
                       (const float *)b, pb, qb, nx, ny, partial);
  std::vector<double> h(4 * blocks);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  dev_free(partial);
  if (e != hipSuccess) return set_error(ERR_DEVICE, "diff_norms: %s", hipGetErrorString(e));
  double s2 = 0, r2 = 0, sm = 0, rm = 0;
  for (int k = 0; k < blocks; ++k) {
    s2 += h[4 * k];
    r2 += h[4 * k + 1];
    sm = std::max(sm, h[4 * k + 2]);
    rm = std::max(rm, h[4 * k + 3]);
  }
  out[0] = std::sqrt(s2);
  out[1] = std::sqrt(r2);
  out[2] = sm;
  out[3] = rm;
  return OK;
}

// ---------------------------------------------------------------------------
// diff_norms without allocation.  The per-workgroup partials live in a
// file-scope device array (one copy per device, held until the process
// ends) instead of a hipMalloc'd buffer, and a second kernel sums them on
// the device in the order the host loop of diff_norms uses, so the results
// are bit-identical to diff_norms.  The array's size does not depend on the
// grid (the norm kernel strides over any grid with a fixed workgroup count).

constexpr int kNormScratch = 4 * kNormBlocks + 4;  // partials, then the four totals
static __device__ __align__(256) double scratch[kNormScratch];

// Sums of the partials in workgroup order (sum of squares) and their maxima:
// one lane, the arithmetic of diff_norms' host loop.
__global__ __launch_bounds__(64) void k_sum_partials(double *__restrict__ sc, int blocks) {
  if (threadIdx.x != 0) return;
  double s2 = 0, r2 = 0, sm = 0, rm = 0;
  for (int k = 0; k < blocks; ++k) {
    s2 += sc[4 * k];
    r2 += sc[4 * k + 1];
    sm = sm < sc[4 * k + 2] ? sc[4 * k + 2] : sm;  // std::max(sm, v)
    rm = rm < sc[4 * k + 3] ? sc[4 * k + 3] : rm;
  }
  sc[4 * blocks] = s2;
  sc[4 * blocks + 1] = r2;
  sc[4 * blocks + 2] = sm;
  sc[4 * blocks + 3] = rm;
}

// Device address of `scratch` on each device, looked up once per device
// with that device current.  The host-side name `scratch` is only a handle:
// kernels receive the looked-up address.
static std::mutex g_scratch_mu;
static double *g_scratch[64] = {};

static int scratch_on(int dev, double **p) {
  if (dev < 0 || dev >= 64) return set_error(ERR_DEVICE, "diff_norms_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q)
      return set_error(ERR_DEVICE, "diff_norms_noalloc: scratch address: %s",
                       hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(q);
  }
  *p = g_scratch[dev];
  return OK;
}

int diff_norms_noalloc(const void *a, int ea, long pa, long qa, const void *b, int eb, long pb,
                       long qb, int nx, int ny, double out[4]) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(ERR_DEVICE, "diff_norms_noalloc: device");
  double *sc = nullptr;
  int rc = scratch_on(dev, &sc);
  if (rc) return rc;
  const dim3 g(kNormBlocks), bl(256);
  if (ea == 8 && eb == 8)
    hipLaunchKernelGGL((k_diff_norms<double, double>), g, bl, 0, nullptr, (const double *)a, pa,
                       qa, (const double *)b, pb, qb, nx, ny, sc);
  else if (ea == 4 && eb == 8)
    hipLaunchKernelGGL((k_diff_norms<float, double>), g, bl, 0, nullptr, (const float *)a, pa, qa,
                       (const double *)b, pb, qb, nx, ny, sc);
  else if (ea == 8 && eb == 4)
    hipLaunchKernelGGL((k_diff_norms<double, float>), g, bl, 0, nullptr, (const double *)a, pa, qa,
                       (const float *)b, pb, qb, nx, ny, sc);
  else
    hipLaunchKernelGGL((k_diff_norms<float, float>), g, bl, 0, nullptr, (const float *)a, pa, qa,
                       (const float *)b, pb, qb, nx, ny, sc);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, nullptr, sc, kNormBlocks);
  double tot[4];
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpy(tot, sc + 4 * kNormBlocks, sizeof tot, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_error(ERR_DEVICE, "diff_norms_noalloc: %s", hipGetErrorString(e));
  out[0] = std::sqrt(tot[0]);
  out[1] = std::sqrt(tot[1]);
  out[2] = tot[2];
  out[3] = tot[3];
  return OK;
}

// ---------------------------------------------------------------------------

static int mcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) return set_error(ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return OK;
}

void mat_release(MatPlane *m) {
  if (m->aux) {
    (void)hipStreamSynchronize(m->aux);
    (void)hipStreamDestroy(m->aux);
  }
  for (hipEvent_t e : {m->ev_fork, m->ev_join})
    if (e) (void)hipEventDestroy(e);
  m->aux = nullptr;
  m->ev_fork = m->ev_join = nullptr;
  for (void *p : {(void *)m->d_ids, m->d_cx, m->d_cy, (void *)m->d_homog, (void *)m->d_mixed,
                  (void *)m->d_homog2, (void *)m->d_mixed2, (void *)m->d_etab})
    dev_free(p);
  m->d_ids = nullptr;
  m->d_cx = m->d_cy = nullptr;
  m->d_homog = m->d_mixed = m->d_homog2 = m->d_mixed2 = nullptr;
  m->d_etab = nullptr;
  m->n_homog = m->n_mixed = m->n_homog2 = m->n_mixed2 = 0;
  m->built_fb = m->built_seg = m->built_edge = m->built_pair_seg = m->built_pair_fb = -1;
}

// Material and sweep constants in double (le_init_material and the sweeps'
// k = dt*c/h), rounded once to T.
template <typename T>
static void mat_coef(double c1, double c2, double rho, double dt, double h, SweepCoef<T> *c) {
  const double mu = rho * c2 * c2;
  const double la = rho * c1 * c1 - 2.0 * mu;
  const double k1 = dt * c1 / h, k2 = dt * c2 / h;
  c->k1 = (T)k1;
  c->k2 = (T)k2;
  c->hk1 = (T)((1.0 - k1) * 0.5);
  c->hk2 = (T)((1.0 - k2) * 0.5);
  c->irhoc1 = (T)(1.0 / (c1 * rho));
  c->irhoc2 = (T)(1.0 / (c2 * rho));
  c->rhoc1 = (T)(c1 * rho);
  c->rhoc2 = (T)(c2 * rho);
  c->rhoc3 = (T)(rho * c1 * la / (la + 2.0 * mu));
}

template <typename T>
static int upload_coefs(MatPlane *m, double dt, double hx, double hy) {
  std::vector<SweepCoef<T>> cx(m->count), cy(m->count);
  std::vector<double> et(3 * m->count);
  for (int k = 0; k < m->count; ++k) {
    const double c1 = m->c1c2rho[3 * k], c2 = m->c1c2rho[3 * k + 1], rho = m->c1c2rho[3 * k + 2];
    mat_coef<T>(c1, c2, rho, dt, hx, &cx[k]);
    mat_coef<T>(c1, c2, rho, dt, hy, &cy[k]);
    et[3 * k] = rho;
    et[3 * k + 2] = rho * c2 * c2;
    et[3 * k + 1] = rho * c1 * c1 - 2.0 * et[3 * k + 2];
  }
  const size_t b = m->count * sizeof(SweepCoef<T>);
  int rc = OK;
  if (!m->d_cx) rc = dev_alloc(&m->d_cx, b, "materials");
  if (!rc && !m->d_cy) rc = dev_alloc(&m->d_cy, b, "materials");
  if (!rc && !m->d_etab) rc = dev_alloc((void **)&m->d_etab, et.size() * sizeof(double), "materials");
  if (rc) return rc;
  rc = mcheck(hipMemcpy(m->d_cx, cx.data(), b, hipMemcpyHostToDevice), "materials");
  if (!rc) rc = mcheck(hipMemcpy(m->d_cy, cy.data(), b, hipMemcpyHostToDevice), "materials");
  if (!rc)
    rc = mcheck(hipMemcpy(m->d_etab, et.data(), et.size() * sizeof(double), hipMemcpyHostToDevice),
                "materials");
  return rc;
}

int mat_set(DevGrid *g, MatPlane *m, int count, const double *c1c2rho, const unsigned char *ids,
            double dt, double hx, double hy) {
  if (count < 1 || count > 256) return set_error(ERR_ARG, "materials: count %d", count);
  for (int k = 0; k < 3 * count; ++k)
    if (!(c1c2rho[k] > 0)) return set_error(ERR_ARG, "materials: c1, c2, rho must be > 0");
  const long n = (long)g->nx * g->ny;
  if (ids)
    for (long k = 0; k < n; ++k)
      if (ids[k] >= count) return set_error(ERR_ARG, "materials: id %d >= count %d", ids[k], count);
  mat_release(m);
  m->count = count;
  m->c1c2rho.assign(c1c2rho, c1c2rho + 3 * count);
  m->ids.assign(n, 0);
  if (ids) m->ids.assign(ids, ids + n);
  m->idp = (g->nx + 255) / 256 * 256;
  int rc = dev_alloc((void **)&m->d_ids, (size_t)m->idp * g->ny, "materials: id plane");
  if (rc) return rc;
  rc = mcheck(hipMemcpy2D(m->d_ids, m->idp, m->ids.data(), g->nx, g->nx, g->ny,
                              hipMemcpyHostToDevice),
                  "materials: ids");
  if (rc) return rc;
  rc = g->elem == 8 ? upload_coefs<double>(m, dt, hx, hy) : upload_coefs<float>(m, dt, hx, hy);
  if (rc) return rc;
  m->active = true;
  return OK;
}

// Split the tiles of a strip width and segment height into single-material
// and mixed ones by their whole read window: the strip's `width` output
// columns with `ch` halo columns on each side, and rows j0-rh .. j1-1+rh
// (edge-mapped).
static void split_tiles(const MatPlane *m, int nx, int ny, int width, int ch, int seg, int rh,
                        int edge, std::vector<MatTile> *homog, std::vector<MatTile> *mixed) {
  const int strips = (nx + width - 1) / width;
  std::vector<int> rows;
  for (int j0 = 0; j0 < ny; j0 += seg) {
    const int j1 = std::min(j0 + seg, ny);
    rows.clear();
    for (int p = j0 - rh; p <= j1 - 1 + rh; ++p) {
      int r = p < 0 ? 0 : (p < ny ? p : (edge == EDGE_CLAMP ? ny - 1 : std::max(p - 5, 0)));
      rows.push_back(r);
    }
    for (int s = 0; s < strips; ++s) {
      const int xa = std::max(s * width - ch, 0), xb = std::min(s * width + width - 1 + ch, nx - 1);
      const unsigned char first = m->ids[(long)rows[0] * nx + xa];
      bool one = true;
      for (size_t q = 0; q < rows.size() && one; ++q) {
        const unsigned char *r = m->ids.data() + (long)rows[q] * nx;
        for (int x = xa; x <= xb; ++x)
          if (r[x] != first) {
            one = false;
            break;
          }
      }
      (one ? homog : mixed)->push_back({s, j0, j1, one ? (int)first : -1, j0 / seg});
    }
  }
}

static int upload_tiles(const std::vector<MatTile> &v, MatTile **d, int *n) {
  dev_free(*d);
  *d = nullptr;
  *n = 0;
  if (v.empty()) return OK;
  int rc = dev_alloc((void **)d, v.size() * sizeof(MatTile), "materials: tiles");
  if (rc) return rc;
  rc = mcheck(hipMemcpy(*d, v.data(), v.size() * sizeof(MatTile), hipMemcpyHostToDevice),
                  "materials: tiles");
  if (rc) return rc;
  *n = (int)v.size();
  return OK;
}

// Tables of the current single-step launch shape.
static int mat_tiles(DevGrid *g, MatPlane *m, int edge) {
  const int fb = g->tiles.xy_block == 64 || g->tiles.xy_block == 128 ? g->tiles.xy_block : 256;
  const int seg = std::max(1, g->tiles.xy_seg);
  if (m->built_fb == fb && m->built_seg == seg && m->built_edge == edge) return OK;
  std::vector<MatTile> homog, mixed;
  split_tiles(m, g->nx, g->ny, fb - 4, 2, seg, 2, edge, &homog, &mixed);
  int rc = upload_tiles(homog, &m->d_homog, &m->n_homog);
  if (!rc) rc = upload_tiles(mixed, &m->d_mixed, &m->n_mixed);
  if (rc) return rc;
  m->built_fb = fb;
  m->built_seg = seg;
  m->built_edge = edge;
  return OK;
}

// The interface tiles are few and march slowly (both faces of every node),
// so run after the single-material tiles they would leave the device mostly
// idle: 0.44 ms of a 6.5 ms two-step pass on 16384^2.  They are launched
// first, on a second stream forked from the grid's stream, and joined back
// before the next launch on it.  Both launches read the same input and write
// disjoint tiles.
static int fork_aux(DevGrid *g, MatPlane *m) {
  if (!m->aux) {
    if (hipStreamCreateWithFlags(&m->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_join, hipEventDisableTiming) != hipSuccess)
      return set_error(ERR_DEVICE, "materials: second stream");
  }
  if (hipEventRecord(m->ev_fork, g->stream) != hipSuccess ||
      hipStreamWaitEvent(m->aux, m->ev_fork, 0) != hipSuccess)
    return set_error(ERR_DEVICE, "materials: fork");
  return OK;
}

static int join_aux(DevGrid *g, MatPlane *m) {
  if (hipEventRecord(m->ev_join, m->aux) != hipSuccess ||
      hipStreamWaitEvent(g->stream, m->ev_join, 0) != hipSuccess)
    return set_error(ERR_DEVICE, "materials: join");
  return OK;
}

static int pair_lanes(const DevGrid *g) {
  const int fb = g->tiles.pair_block;
  return fb == 64 || fb == 128 ? fb : 256;
}

// Tables of the two-step pass: strips of (window - 8) columns read 4 halo
// columns and 4 halo rows on each side.
static int mat_tiles2(DevGrid *g, MatPlane *m) {
  const int seg = g->tiles.pair_seg, fb = pair_lanes(g);
  if (m->built_pair_seg == seg && m->built_pair_fb == fb) return OK;
  std::vector<MatTile> homog, mixed;
  split_tiles(m, g->nx, g->ny, fb - 8, 4, seg, 4, EDGE_CLAMP, &homog, &mixed);
  int rc = upload_tiles(homog, &m->d_homog2, &m->n_homog2);
  if (!rc) rc = upload_tiles(mixed, &m->d_mixed2, &m->n_mixed2);
  if (rc) return rc;
  m->built_pair_seg = seg;
  m->built_pair_fb = fb;
  return OK;
}

template <typename T, int FB>
static int launch_mat2(DevGrid *g, MatPlane *m) {
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
  const SweepCoef<T> *cx = static_cast<const SweepCoef<T> *>(m->d_cx);
  const SweepCoef<T> *cy = static_cast<const SweepCoef<T> *>(m->d_cy);
  if (m->n_mixed2) {
    int rc = fork_aux(g, m);
    if (rc) return rc;
    hipLaunchKernelGGL((k_step2_mat_mixed<T, FB>), dim3(m->n_mixed2), dim3(FB), 0, m->aux, src,
                       dst, g->nx, g->ny, g->pitch, g->plane, m->d_mixed2, m->d_ids, m->idp, cx,
                       cy);
  }
  if (m->n_homog2 && FB == 64 && g->tiles.pair_split)
    hipLaunchKernelGGL((k_step2_mat_homog_split<T>), dim3(m->n_homog2), dim3(128), 0, g->stream,
                       src, dst, g->nx, g->ny, g->pitch, g->plane, m->d_homog2, cx, cy);
  else if (m->n_homog2)
    hipLaunchKernelGGL((k_step2_mat_homog<T, FB>), dim3(m->n_homog2), dim3(FB), 0, g->stream, src,
                       dst, g->nx, g->ny, g->pitch, g->plane, m->d_homog2, cx, cy);
  return m->n_mixed2 ? join_aux(g, m) : OK;
}

template <typename T, int EDGE, int FB>
static int launch_mat(DevGrid *g, MatPlane *m) {
  const T *src = static_cast<const T *>(g->buf[g->cur]);
  T *dst = static_cast<T *>(g->buf[g->cur ^ 1]);
  const SweepCoef<T> *cx = static_cast<const SweepCoef<T> *>(m->d_cx);
  const SweepCoef<T> *cy = static_cast<const SweepCoef<T> *>(m->d_cy);
  if (m->n_mixed) {
    int rc = fork_aux(g, m);
    if (rc) return rc;
    hipLaunchKernelGGL((k_step_mat_mixed<T, EDGE, FB>), dim3(m->n_mixed), dim3(FB), 0, m->aux,
                       src, dst, g->nx, g->ny, g->pitch, g->plane, m->d_mixed, m->d_ids, m->idp,
                       cx, cy);
  }
  if (m->n_homog)
    hipLaunchKernelGGL((k_step_mat_homog<T, EDGE, FB>), dim3(m->n_homog), dim3(FB), 0, g->stream,
                       src, dst, g->nx, g->ny, g->pitch, g->plane, m->d_homog, cx, cy);
  return m->n_mixed ? join_aux(g, m) : OK;
}

template <typename T, int EDGE>
static int launch_mat_fb(DevGrid *g, MatPlane *m) {
  if (m->built_fb == 64) return launch_mat<T, EDGE, 64>(g, m);
  if (m->built_fb == 128) return launch_mat<T, EDGE, 128>(g, m);
  return launch_mat<T, EDGE, 256>(g, m);
}

template <typename T>
int step_xy_mat(DevGrid *g, MatPlane *m, int edge) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "step_xy_mat: element size mismatch");
  if (!m->active) return set_error(ERR_STATE, "step_xy_mat: no material plane");
  int rc = mat_tiles(g, m, edge);
  if (rc) return rc;
  rc = edge == EDGE_RING ? launch_mat_fb<T, EDGE_RING>(g, m) : launch_mat_fb<T, EDGE_CLAMP>(g, m);
  if (rc) return rc;
  rc = mcheck(hipGetLastError(), "step_xy_mat launch");
  if (rc) return rc;
  g->cur ^= 1;
  return OK;
}

template int step_xy_mat<double>(DevGrid *, MatPlane *, int);
template int step_xy_mat<float>(DevGrid *, MatPlane *, int);

template <typename T>
int step2_xy_mat(DevGrid *g, MatPlane *m) {
  if (g->elem != (int)sizeof(T)) return set_error(ERR_ARG, "step2_xy_mat: element size mismatch");
  if (!m->active) return set_error(ERR_STATE, "step2_xy_mat: no material plane");
  if (g->tiles.pair_seg < 1) return set_error(ERR_ARG, "step2_xy_mat: no pair segment height");
  int rc = mat_tiles2(g, m);
  if (rc) return rc;
  const int fb = pair_lanes(g);
  rc = fb == 64    ? launch_mat2<T, 64>(g, m)
       : fb == 128 ? launch_mat2<T, 128>(g, m)
                   : launch_mat2<T, 256>(g, m);
  if (rc) return rc;
  rc = mcheck(hipGetLastError(), "step2_xy_mat launch");
  if (rc) return rc;
  g->cur ^= 1;
  return OK;
}

template int step2_xy_mat<double>(DevGrid *, MatPlane *);
template int step2_xy_mat<float>(DevGrid *, MatPlane *);

int energy(DevGrid *g, const MatPlane *m, double rho, double c1, double c2, double *out) {
  double *etab = nullptr;
  double *partial = nullptr;
  const int blocks = 2048;
  const double *tab = m && m->active ? m->d_etab : nullptr;
  if (!tab) {
    const double mu = rho * c2 * c2;
    const double host[3] = {rho, rho * c1 * c1 - 2.0 * mu, mu};
    const int arc = dev_alloc((void **)&etab, sizeof host, "energy: table");
    if (arc) return arc;
    if (hipMemcpy(etab, host, sizeof host, hipMemcpyHostToDevice) != hipSuccess) {
      dev_free(etab);
      return set_error(ERR_DEVICE, "energy: table");
    }
    tab = etab;
  }
  {
    const int arc = dev_alloc((void **)&partial, blocks * sizeof(double), "energy");
    if (arc) {
      dev_free(etab);
      return arc;
    }
  }
  const unsigned char *ids = m && m->active ? m->d_ids : nullptr;
  const long idp = m && m->active ? m->idp : 0;
  if (g->elem == 8)
    hipLaunchKernelGGL(k_energy<double>, dim3(blocks), dim3(256), 0, g->stream,
                       static_cast<const double *>(g->buf[g->cur]), g->nx, g->ny, g->pitch,
                       g->plane, ids, idp, tab, partial);
  else
    hipLaunchKernelGGL(k_energy<float>, dim3(blocks), dim3(256), 0, g->stream,
                       static_cast<const float *>(g->buf[g->cur]), g->nx, g->ny, g->pitch, g->plane,
                       ids, idp, tab, partial);
  std::vector<double> h(blocks);
  int rc = mcheck(hipGetLastError(), "energy launch");
  if (!rc) rc = mcheck(hipMemcpyAsync(h.data(), partial, blocks * sizeof(double),
                                      hipMemcpyDeviceToHost, g->stream), "energy");
  if (!rc) rc = mcheck(hipStreamSynchronize(g->stream), "energy");
  dev_free(partial);
  dev_free(etab);
  if (rc) return rc;
  double s = 0.0;
  for (double v : h) s += v;
  *out = s;
  return OK;
}

}  // namespace le2d
