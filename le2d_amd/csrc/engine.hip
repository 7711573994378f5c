// engine.hip -- device memory, host<->device transfers and error state of
// the sweep engine.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <mutex>
#include <unordered_map>

#include "engine.h"
#include "le2d.h"

namespace le2d {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *last_error() { return g_err; }

int require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    fprintf(stderr,
            "le2d: no HIP device is available (%s); this library runs the sweeps on the GPU "
            "only and has no CPU path\n",
            e != hipSuccess ? hipGetErrorString(e) : "device count 0");
    abort();
  }
  return OK;
}

// Device allocator.  Allocations made by an installed allocator are recorded
// with its free function, so each is returned to the allocator that made it.
namespace {
struct Allocator {
  le2d_alloc_fn alloc = nullptr;
  le2d_free_fn release = nullptr;
  void *ctx = nullptr;
};
struct Owner {
  le2d_free_fn release;
  void *ctx;
};
std::mutex g_alloc_mu;
Allocator g_alloc;
std::unordered_map<void *, Owner> g_owned;
}  // namespace

int dev_alloc(void **p, size_t bytes, const char *what) {
  *p = nullptr;
  Allocator a;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    a = g_alloc;
  }
  if (a.alloc) {
    void *q = nullptr;
    if (a.alloc(&q, bytes, a.ctx) != 0 || !q)
      return set_error(ERR_NOMEM, "%s: the installed allocator refused %zu bytes", what, bytes);
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    g_owned[q] = {a.release, a.ctx};
    *p = q;
    return OK;
  }
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(ERR_NOMEM, "%s: hipMalloc(%zu) failed: %s", what, bytes,
                     hipGetErrorString(e));
  }
  return OK;
}

}  // namespace le2d

extern "C" int le2d_set_device_allocator(le2d_alloc_fn alloc, le2d_free_fn release, void *ctx) {
  if ((alloc == nullptr) != (release == nullptr))
    return le2d::set_error(le2d::ERR_ARG, "set_device_allocator: give both functions or neither");
  std::lock_guard<std::mutex> lk(le2d::g_alloc_mu);
  le2d::g_alloc = {alloc, release, alloc ? ctx : nullptr};
  return le2d::OK;
}

namespace le2d {

void dev_free(void *p) {
  if (!p) return;
  Owner o{nullptr, nullptr};
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_owned.find(p);
    if (it != g_owned.end()) {
      o = it->second;
      g_owned.erase(it);
    }
  }
  if (o.release)
    o.release(p, o.ctx);
  else
    (void)hipFree(p);
}

// One grid row of one field, padded to 256 bytes so every row starts aligned.
static long field_pitch(int nx, int elem) {
  const long per = 256 / elem;
  return ((long)nx + per - 1) / per * per;
}

// Row-interleaved fields: the five fields of grid row j are stored side by
// side (vx, vy, sxx, sxy, syy), each padded to 256 bytes, and the rows follow
// one another.  The workgroups of a sweep that work on the same rows then
// read and write one contiguous region instead of five regions a plane
// apart, which keeps the DRAM access stream local.
void grid_layout(int elem, int nx, int ny, long *pitch, long *plane) {
  (void)ny;
  const long fp = field_pitch(nx, elem);
  *pitch = 5 * fp;
  *plane = fp;
}

bool row_fits(int elem, int nx) {
  return nx >= 1 && 5 * field_pitch(nx, elem) * (long)elem <= 0x7fffffffL;
}

size_t grid_bytes(int elem, int nx, int ny) {
  return (size_t)5 * field_pitch(nx, elem) * ny * elem;
}

int grid_alloc(DevGrid *g, int elem, int nx, int ny, hipStream_t stream) {
  if ((elem != 4 && elem != 8) || nx < 3 || ny < 3)
    return set_error(ERR_ARG, "grid_alloc: need elem 4|8 and nx, ny >= 3 (got %d, %d, %d)", elem,
                     nx, ny);
  if (!row_fits(elem, nx)) return set_error(ERR_ARG, "grid_alloc: row of %d nodes too long", nx);
  g->elem = elem;
  g->nx = nx;
  g->ny = ny;
  grid_layout(elem, nx, ny, &g->pitch, &g->plane);
  g->stream = stream;
  g->cur = 0;
  (void)hipGetDevice(&g->device);
  const size_t bytes = grid_bytes(elem, nx, ny);
  for (int k = 0; k < 2; ++k) {
    const int rc = dev_alloc(&g->buf[k], bytes, "grid_alloc");
    if (rc) {
      grid_free(g);
      return rc;
    }
  }
  g->tiles = default_tiles(elem, nx, ny);
  return OK;
}

void grid_free(DevGrid *g) {
  for (int k = 0; k < 2; ++k) {
    dev_free(g->buf[k]);
    g->buf[k] = nullptr;
  }
}

// AOS (node-contiguous, unpitched) <-> the device layout.  One lane per node;
// the AOS side is read/written as 5 consecutive values per lane.
template <typename T>
__global__ void k_aos_to_planes(const T *__restrict__ aos, T *__restrict__ pl, int nx, int ny,
                                long pitch, long plane) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long)nx * ny) return;
  const long j = n / nx, i = n - j * nx;
  const T *s = aos + 5 * n;
  T *d = pl + j * pitch + i;
#pragma unroll
  for (int q = 0; q < 5; ++q) d[q * plane] = s[q];
}

template <typename T>
__global__ void k_planes_to_aos(const T *__restrict__ pl, T *__restrict__ aos, int nx, int ny,
                                long pitch, long plane) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long)nx * ny) return;
  const long j = n / nx, i = n - j * nx;
  const T *s = pl + j * pitch + i;
  T *d = aos + 5 * n;
#pragma unroll
  for (int q = 0; q < 5; ++q) d[q] = s[q * plane];
}

static int check(hipError_t e, const char *what) {
  if (e != hipSuccess) return set_error(ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return OK;
}

// The other ping-pong buffer stages the AOS image (it holds at least 5*nx*ny
// values and carries no state between steps).
int grid_upload(DevGrid *g, const void *host, int soa) {
  const size_t es = g->elem;
  const size_t row = (size_t)g->nx * es;
  void *cur = g->buf[g->cur];
  if (soa) {
    for (int q = 0; q < 5; ++q) {
      const char *h = static_cast<const char *>(host) + (size_t)q * g->nx * g->ny * es;
      char *d = static_cast<char *>(cur) + (size_t)q * g->plane * es;
      int rc = check(hipMemcpy2DAsync(d, g->pitch * es, h, row, row, g->ny, hipMemcpyHostToDevice,
                                      g->stream),
                     "upload planes");
      if (rc) return rc;
    }
    return check(hipStreamSynchronize(g->stream), "upload sync");
  }
  void *stage = g->buf[g->cur ^ 1];
  const size_t bytes = (size_t)g->nx * g->ny * 5 * es;
  int rc = check(hipMemcpyAsync(stage, host, bytes, hipMemcpyHostToDevice, g->stream), "upload");
  if (rc) return rc;
  const long n = (long)g->nx * g->ny;
  const int bs = 256;
  const unsigned nb = (unsigned)((n + bs - 1) / bs);
  if (g->elem == 8)
    hipLaunchKernelGGL(k_aos_to_planes<double>, dim3(nb), dim3(bs), 0, g->stream,
                       static_cast<const double *>(stage), static_cast<double *>(cur), g->nx, g->ny,
                       g->pitch, g->plane);
  else
    hipLaunchKernelGGL(k_aos_to_planes<float>, dim3(nb), dim3(bs), 0, g->stream,
                       static_cast<const float *>(stage), static_cast<float *>(cur), g->nx, g->ny,
                       g->pitch, g->plane);
  rc = check(hipGetLastError(), "aos_to_planes launch");
  if (rc) return rc;
  return check(hipStreamSynchronize(g->stream), "upload sync");
}

int grid_download(DevGrid *g, void *host, int soa) {
  const size_t es = g->elem;
  const size_t row = (size_t)g->nx * es;
  void *cur = g->buf[g->cur];
  if (soa) {
    for (int q = 0; q < 5; ++q) {
      char *h = static_cast<char *>(host) + (size_t)q * g->nx * g->ny * es;
      const char *d = static_cast<const char *>(cur) + (size_t)q * g->plane * es;
      int rc = check(hipMemcpy2DAsync(h, row, d, g->pitch * es, row, g->ny, hipMemcpyDeviceToHost,
                                      g->stream),
                     "download planes");
      if (rc) return rc;
    }
    return check(hipStreamSynchronize(g->stream), "download sync");
  }
  void *stage = g->buf[g->cur ^ 1];
  const long n = (long)g->nx * g->ny;
  const int bs = 256;
  const unsigned nb = (unsigned)((n + bs - 1) / bs);
  if (g->elem == 8)
    hipLaunchKernelGGL(k_planes_to_aos<double>, dim3(nb), dim3(bs), 0, g->stream,
                       static_cast<const double *>(cur), static_cast<double *>(stage), g->nx, g->ny,
                       g->pitch, g->plane);
  else
    hipLaunchKernelGGL(k_planes_to_aos<float>, dim3(nb), dim3(bs), 0, g->stream,
                       static_cast<const float *>(cur), static_cast<float *>(stage), g->nx, g->ny,
                       g->pitch, g->plane);
  int rc = check(hipGetLastError(), "planes_to_aos launch");
  if (rc) return rc;
  rc = check(hipMemcpyAsync(host, stage, (size_t)n * 5 * es, hipMemcpyDeviceToHost, g->stream),
             "download");
  if (rc) return rc;
  return check(hipStreamSynchronize(g->stream), "download sync");
}

static int best_segment(long strips, int ny, long resident, int halo, int cap);

Tiles default_tiles(int elem, int nx, int ny) {
  Tiles t;
  // one row per x-sweep workgroup (the most workgroups in flight): 8192^2
  // FP64 0.895 ms against 1.026 at 4 rows, FP32 0.426 against 0.457
  t.x_rows = 1;
  t.pair_split = 0;
  {
    // fused step: 256-lane windows (252 output columns) and ~64 workgroups
    // per CU in segments of at least 16 rows (measured best on 8192^2: 16-24
    // rows; short segments keep the halo rows of neighbouring segments in L2).
    // Grids under 4096 columns take one-wave windows (60 columns, no
    // barriers between waves): 2048^2 FP64 50.5 against 45.9 Gnode-updates/s,
    // FP32 144 against 137.
    t.xy_block = nx >= 4096 ? 256 : 64;
    const long strips = (nx + t.xy_block - 5) / (t.xy_block - 4);
    const long want = (16384 + strips - 1) / strips;
    long seg = (ny + want - 1) / want;
    if (seg < 16) seg = 16;
    if (seg > ny) seg = ny;
    t.xy_seg = (int)seg;
    // grids that 16-row segments leave short of one round of resident
    // workgroups: a step lasts one workgroup's march (4 halo rows, then the
    // segment), so the occupancy model of the two-step pass picks a shorter
    // segment (512^2 FP64 16.5 Gnode-updates/s at 2-4 rows against 11.5 at
    // 16; 1024^2 39.6 at 12 against 33.9)
    long resident = step_resident(elem, t.xy_block);
    if (resident <= 0) resident = 768;
    if (strips * ((ny + 15) / 16) < resident)
      t.xy_seg = best_segment(strips, ny, resident, 4, 16);
  }
  t.y_block = nx >= 8 * 256 ? 256 : (nx >= 8 * 128 ? 128 : 64);
  // about 512 workgroups (two per CU), each marching at least 32 rows:
  // longer marches re-read fewer halo rows (8192^2 FP64 1.019 ms at 512
  // rows against 1.059 at 128, FP32 0.510 against 0.540; 2048^2 keeps 32)
  const long cols = (nx + t.y_block - 1) / t.y_block;
  long want = (512 + cols - 1) / cols;
  long seg = (ny + want - 1) / want;
  if (seg < 32) seg = 32;
  if (seg > ny) seg = ny;
  t.y_seg = (int)seg;
  // two-step pass (le_step_n, paired le_step calls), measured on MI355X:
  // FP64 takes one-wave windows (64 lanes, 56 columns, no barriers between
  // waves; 8192^2 at 80-row segments 90 Gnode-updates/s against 85 with 128
  // or 256 lanes).  FP32, issue- rather than latency-bound, keeps the 3% halo
  // of 256-lane windows on grids 4096 or more columns wide (8192^2 229
  // against 210; 2048^2 137 against 153 with 64 lanes).
  t.pair_block = elem == 8 || nx < 4096 ? 64 : 256;
  t.pair_seg = pair_seg_for(elem, nx, ny, t.pair_block);
  return t;
}

// The candidate segment height (2 .. cap rows) that maximises (fill of the
// last round of `resident` workgroups) x (share of a segment's rows that are
// output rather than the `halo` re-read rows).
static int best_segment(long strips, int ny, long resident, int halo, int cap) {
  int best = 2;
  double best_score = -1;
  for (int s : {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 80}) {
    if (s > cap || (s > ny && s != 2)) break;
    const long blocks = strips * ((ny + s - 1) / s);
    const long rounds = (blocks + resident - 1) / resident;
    const double fill = (double)blocks / (double)(rounds * resident);
    const double rows = std::min(s, ny);
    const double score = fill * rows / (rows + halo);
    if (score > best_score + 1e-9) {
      best_score = score;
      best = s;
    }
  }
  return best;
}

// Segment height of the two-step pass: the candidate height that maximises
// (fill of the last round of resident workgroups) x (share of a segment's
// rows that are output rather than the 8 re-read halo rows).  The grid's
// workgroups run in rounds of `resident`; a last round that is mostly empty
// leaves the device idle.  Taller segments than 80 rows (64-lane windows)
// or 64 (wider) measured slower than the model says and are not taken.
// Measured optima it reproduces (Gnode-updates/s): FP64 64 lanes, 80 rows
// on 8192^2 (90.2; 86.4 at 160), 8192 x 1024 (75.2; 66.7 at 32) and 4096^2
// (79.7), 48 on 2048^2 (61.0); FP32 256 lanes, 64 rows on 8192^2 (229.4;
// 223.5 at 128) and 16384^2 (231.1; 223.2 at 160).  On grids whose
// workgroups all fit in one round the pass lasts one workgroup's march, so
// short segments win: FP64 12 rows on 1024^2 (43.2; 36.6 at 16), 4 on 512^2
// (18.2; 13.4 at 16), 2 on 256^2 (7.3; 3.5 at 16).
int pair_seg_for(int elem, int nx, int ny, int fb) {
  long resident = pair_resident(elem, fb);
  if (resident <= 0) resident = 2048;
  const long strips = (nx + fb - 9) / (fb - 8);
  return best_segment(strips, ny, resident, 8, fb == 64 ? 80 : 64);
}

// One 16-byte vector per lane, one pass: the fastest copy shape measured on
// MI355X (about 6.2 TB/s for 4 GiB), against ~4.7 TB/s for grid-stride loops.
typedef unsigned v4u __attribute__((ext_vector_type(4)));

__global__ void k_copy16(const v4u *__restrict__ src, v4u *__restrict__ dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int copy_probe(size_t bytes, int iters, double *ms) {
  require_device();
  if (bytes < 16 || iters < 1) return set_error(ERR_ARG, "copy_probe: bytes %zu iters %d", bytes, iters);
  void *a = nullptr, *b = nullptr;
  int rc = dev_alloc(&a, bytes, "copy_probe");
  if (!rc) rc = dev_alloc(&b, bytes, "copy_probe");
  if (rc) {
    dev_free(a);
    return rc;
  }
  (void)hipMemset(a, 1, bytes);
  const long n = (long)(bytes / 16);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, 0, (const v4u *)a, (v4u *)b, n);
  (void)hipEventRecord(e0, 0);
  for (int k = 0; k < iters; ++k)
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, 0, (const v4u *)a, (v4u *)b, n);
  (void)hipEventRecord(e1, 0);
  hipError_t e = hipEventSynchronize(e1);
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  if (e != hipSuccess) return set_error(ERR_DEVICE, "copy_probe: %s", hipGetErrorString(e));
  *ms = t / iters;
  return OK;
}

}  // namespace le2d
