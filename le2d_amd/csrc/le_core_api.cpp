// le_core_api.cpp -- the le_core.h C ABI over the HIP sweep engine.
//
// Compiled twice: with -DUSE_DOUBLE into liblecore_d.so and without into
// liblecore_f.so, exactly as the reference builds `real` (simple/Makefile:6-7).
// Host-side arithmetic (material constants, disk test, VTK magnitudes) is the
// reference's, operation for operation; the sweeps run on the GPU.
//
// Device state is found through a table keyed by the host mirror pointer
// (task->grid), because le_task has no spare field and callers may copy the
// struct by value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "le2d.h"
#include "le_core.h"

namespace {

using le2d::DevGrid;
using le2d::SweepCoef;

struct TaskState {
  DevGrid g;
  le2d::MatPlane mat;       // per-node materials, when set
  bool dev_valid = false;   // device buffers hold a state
  bool host_dirty = true;   // mirror newer than device
  bool dev_newer = false;   // device newer than mirror
  bool pinned = false;      // mirror from hipHostMalloc
  bool owned = true;        // mirror allocated by le_init_task
  int mode = LE2D_SYNC_WAIT;
  hipStream_t stream = nullptr;
  bool pairs_fixed = false;  // segment height set by the caller or tuned
  bool mat_split_tuned = false;  // material plane: one or two wavefronts chosen
  // A step accepted by le_step but not yet run, so that it and the next
  // le_step share one two-step pass; run before anything else touches the
  // device state (flush_pending), with the coefficients it was called with.
  bool pending = false;
  SweepCoef<real> pend_cx, pend_cy;
};

std::mutex g_mu;
std::unordered_map<const real *, TaskState *> g_tasks;
// tuned two-step segment heights per (element size, nx, ny), for the process
std::mutex g_pair_mu;
std::map<std::tuple<int, int, int>, std::tuple<int, int, int>> g_pair_seg;  // rows, lanes, split

int env_mode() {
  const char *s = getenv("LE2D_SYNC");
  if (!s) return LE2D_SYNC_WAIT;
  if (!strcmp(s, "async")) return LE2D_SYNC_ASYNC;
  if (!strcmp(s, "coherent")) return LE2D_SYNC_COHERENT;
  return LE2D_SYNC_WAIT;
}

bool have_device() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

[[noreturn]] void die(const char *what, int rc) {
  fprintf(stderr, "le2d: %s failed (%d): %s\n", what, rc, le2d::last_error());
  abort();
}

TaskState *find(const le_task *t) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tasks.find(t->grid);
  if (it != g_tasks.end()) return it->second;
  // A task whose mirror was not allocated by le_init_task: adopt it.
  TaskState *s = new TaskState();
  s->owned = false;
  s->mode = env_mode();
  g_tasks[t->grid] = s;
  return s;
}

int flush_pending(TaskState *s);

// Make the device hold the current state (allocate on first use, upload
// host edits, run a deferred step unless keep_pending).
int ensure_device(const le_task *t, TaskState *s, bool keep_pending = false) {
  le2d::require_device();
  if (!s->dev_valid) {
    if (!s->stream) {
      hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
      if (e != hipSuccess)
        return le2d::set_error(LE2D_ERR_DEVICE, "stream: %s", hipGetErrorString(e));
    }
    int rc = le2d::grid_alloc(&s->g, (int)sizeof(real), t->n.x, t->n.y, s->stream);
    if (rc) return rc;
    s->dev_valid = true;
    s->host_dirty = true;
  }
  if (s->host_dirty) {
    s->pending = false;  // the uploaded mirror replaces the device state
    int rc = le2d::grid_upload(&s->g, t->grid, t->stype == ST_SOA);
    if (rc) return rc;
    s->host_dirty = false;
    s->dev_newer = false;
  }
  return keep_pending ? LE2D_OK : flush_pending(s);
}

// Run a deferred le_step on its own (single-step pass).
int flush_pending(TaskState *s) {
  if (!s->pending) return LE2D_OK;
  s->pending = false;
  const int rc = le2d::step_xy<real>(&s->g, s->pend_cx, s->pend_cy, le2d::EDGE_CLAMP);
  if (rc) return rc;
  s->dev_newer = true;
  return LE2D_OK;
}

int sync_host(le_task *t, TaskState *s) {
  if (s->host_dirty) return LE2D_OK;  // the mirror is the newest state (uploaded next)
  int rc = flush_pending(s);
  if (rc) return rc;
  if (!s->dev_valid || !s->dev_newer) return LE2D_OK;
  rc = le2d::grid_download(&s->g, t->grid, t->stype == ST_SOA);
  if (rc) return rc;
  s->dev_newer = false;
  return LE2D_OK;
}

// Sweep constants as the reference forms them (simple/le_core.c:468-469,
// :511-512): k = dt*c/h in `real`; the (1 - k) of tvd2 (:65-66) is formed in
// double and halved exactly.
SweepCoef<real> coef(const le_task *t, real h) {
  SweepCoef<real> c;
  c.k1 = t->dt * t->mat.c1 / h;
  c.k2 = t->dt * t->mat.c2 / h;
  c.hk1 = (real)(1.0 - c.k1) * (real)0.5;
  c.hk2 = (real)(1.0 - c.k2) * (real)0.5;
  c.irhoc1 = t->mat.irhoc1;
  c.irhoc2 = t->mat.irhoc2;
  c.rhoc1 = t->mat.rhoc1;
  c.rhoc2 = t->mat.rhoc2;
  c.rhoc3 = t->mat.rhoc3;
  return c;
}

enum { DO_X = 1, DO_Y = 2 };

// Two-step segment height and window width for this grid: each candidate is
// timed out of place -- the pass reads the current state and writes the idle
// buffer (scratch between steps), and the buffer swap is undone -- so the
// state does not advance.  After a warm-up, every candidate is timed over 4
// passes, the best three again over 12 (single short timings of neighbouring
// candidates differ by less than their noise); the fastest is kept, and
// remembered for the grid shape.
int tune_pairs(le_task *t, TaskState *s) {
  const auto key = std::make_tuple((int)sizeof(real), s->g.nx, s->g.ny);
  {
    std::lock_guard<std::mutex> lk(g_pair_mu);
    auto it = g_pair_seg.find(key);
    if (it != g_pair_seg.end()) {
      s->g.tiles.pair_seg = std::get<0>(it->second);
      s->g.tiles.pair_block = std::get<1>(it->second);
      s->g.tiles.pair_split = std::get<2>(it->second);
      s->pairs_fixed = true;
      return LE2D_OK;
    }
  }
  const SweepCoef<real> cx = coef(t, t->h.x), cy = coef(t, t->h.y);
  const int cand[] = {4, 8, 12, 16, 24, 32, 48, 64, 80, 96, 128, 160};
  hipEvent_t ev[2];
  for (auto &e : ev) (void)hipEventCreate(&e);
  int rc = LE2D_OK;
  // mean ms of one pass at (block, seg) after one untimed pass; < 0 on error
  auto time_one = [&](int block, int split, int seg, int reps) -> float {
    s->g.tiles.pair_seg = seg;
    s->g.tiles.pair_block = block;
    s->g.tiles.pair_split = split;
    for (int k = 0; k <= reps && !rc; ++k) {
      if (k == 1) (void)hipEventRecord(ev[0], s->stream);
      rc = le2d::step2_xy<real>(&s->g, cx, cy);
      s->g.cur ^= 1;  // undo the swap: the state has not moved
    }
    (void)hipEventRecord(ev[1], s->stream);
    if (rc || hipEventSynchronize(ev[1]) != hipSuccess) {
      if (!rc) rc = le2d::set_error(LE2D_ERR_DEVICE, "step pairs: tuning pass failed");
      return -1.0f;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    return ms / reps;
  };
  struct Cand {
    float ms;
    int block, split, seg;
  };
  // warm-up at the default shape, about 25 ms: after host-side work the
  // device clocks are still low and would make the first candidates look slow
  {
    const int b0 = s->g.tiles.pair_block, p0 = s->g.tiles.pair_split, s0 = s->g.tiles.pair_seg;
    const float est = time_one(b0, p0, s0, 1);
    if (est > 0) {
      const int reps = std::max(1, std::min(64, (int)(25.0f / est)));
      (void)time_one(b0, p0, s0, reps);
    }
  }
  std::vector<Cand> res;
  // windows: 256, 128 and 64 lanes, and 64-lane windows with the two stages
  // on two wavefronts (k_step2_split)
  const int blocks[4][2] = {{256, 0}, {128, 0}, {64, 0}, {64, 1}};
  for (const auto &bs : blocks) {
    const int block = bs[0], split = bs[1];
    // segments under 16 rows only for grids that 16-row segments leave short
    // of one round of resident workgroups (there a pass lasts one
    // workgroup's march)
    const long strips = (s->g.nx + block - 9) / (block - 8);
    const long resident = le2d::pair_resident((int)sizeof(real), block);
    const bool small = strips * ((s->g.ny + 15) / 16) < (resident > 0 ? resident : 2048);
    for (int seg : cand) {
      if (seg < 16 && !small) continue;
      if (seg > s->g.ny && seg != cand[0]) break;  // taller segments are all one
      const float ms = time_one(block, split, seg, 4);
      if (ms < 0) break;
      res.push_back({ms, block, split, seg});
    }
  }
  Cand best{1e30f, 256, 0, 64};
  if (!rc && !res.empty()) {
    std::sort(res.begin(), res.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    if (res.size() > 3) res.resize(3);
    for (const Cand &c : res) {
      const float ms = time_one(c.block, c.split, c.seg, 12);
      if (ms < 0) break;
      if (ms < best.ms) best = {ms, c.block, c.split, c.seg};
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->g.tiles.pair_seg = best.seg;
  s->g.tiles.pair_block = best.block;
  s->g.tiles.pair_split = best.split;
  s->pairs_fixed = true;
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_pair_mu);
  g_pair_seg[key] = std::make_tuple(best.seg, best.block, best.split);
  return LE2D_OK;
}

// Material planes keep their default window and segment height (other
// shapes would rebuild the tile tables on the host), but with 64-lane windows
// the two forms of the pass share the tables: time each over three passes,
// out of place (each pass reads the current state and writes the idle
// buffer; the swap is undone), and keep the faster.
int tune_mat_split(TaskState *s) {
  hipEvent_t ev[2];
  for (auto &e : ev) (void)hipEventCreate(&e);
  int rc = LE2D_OK, pick = 0;
  float best = 1e30f;
  for (int split : {0, 1}) {
    s->g.tiles.pair_split = split;
    for (int k = 0; k <= 3 && !rc; ++k) {
      if (k == 1) (void)hipEventRecord(ev[0], s->stream);
      rc = le2d::step2_xy_mat<real>(&s->g, &s->mat);
      if (!rc) s->g.cur ^= 1;  // the state has not moved
    }
    (void)hipEventRecord(ev[1], s->stream);
    if (rc || hipEventSynchronize(ev[1]) != hipSuccess) break;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (ms < best) {
      best = ms;
      pick = split;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->g.tiles.pair_split = pick;
  s->mat_split_tuned = true;
  return rc;
}

bool tuning_enabled() {
  const char *e = getenv("LE2D_TUNE");
  return !(e && !strcmp(e, "0"));
}

// LE2D_PAIR_CALLS=0: every le_step runs its own step before returning.
bool pairing_enabled() {
  static const bool on = [] {
    const char *e = getenv("LE2D_PAIR_CALLS");
    return !(e && !strcmp(e, "0"));
  }();
  return on;
}

bool same_coef(const SweepCoef<real> &a, const SweepCoef<real> &b) {
  return !memcmp(&a, &b, sizeof a);
}

// Queue sweeps on the task's stream, then honour the sync mode.  defer (one
// le_step): an odd step of the clamped full step is held back (pending) and
// runs with the next le_step as one two-step pass -- the same operations, so
// the same bits -- or alone before any other use of the device state.  Not
// in the host-coherent mode, whose mirror is current after every call, nor
// with a material plane.
int run(le_task *t, int what, int edge, int steps, bool defer = false) {
  TaskState *s = find(t);
  int rc = ensure_device(t, s, true);
  if (rc) return rc;
  const SweepCoef<real> cx = coef(t, t->h.x), cy = coef(t, t->h.y);
  const bool pairable = what == (DO_X | DO_Y) && edge == le2d::EDGE_CLAMP &&
                        s->g.tiles.pair_seg > 0 && !s->mat.active;
  if (s->pending) {
    // a held step joins this call's steps only with the same coefficients
    if (pairable && same_coef(s->pend_cx, cx) && same_coef(s->pend_cy, cy)) {
      s->pending = false;
      ++steps;
    } else if ((rc = flush_pending(s))) {
      return rc;
    }
  }
  if (defer && pairable && s->mode != LE2D_SYNC_COHERENT && (steps & 1) && pairing_enabled()) {
    s->pending = true;
    s->pend_cx = cx;
    s->pend_cy = cy;
    --steps;
  }
  if (s->mat.active && what != (DO_X | DO_Y))
    return le2d::set_error(LE2D_ERR_STATE,
                           "single sweeps are not available with a material plane; use le_step");
  int k = 0;
  if (what == (DO_X | DO_Y) && edge == le2d::EDGE_CLAMP && s->g.tiles.pair_seg > 0 &&
      !s->pairs_fixed && !s->mat.active && steps >= 8 && tuning_enabled()) {
    rc = tune_pairs(t, s);
    if (rc) return rc;
  }
  if (what == (DO_X | DO_Y) && edge == le2d::EDGE_CLAMP && s->g.tiles.pair_seg > 0 &&
      s->g.tiles.pair_block == 64 && !s->pairs_fixed && s->mat.active && !s->mat_split_tuned &&
      steps >= 8 && tuning_enabled()) {
    rc = tune_mat_split(s);
    if (rc) return rc;
  }
  if (what == (DO_X | DO_Y) && edge == le2d::EDGE_CLAMP && s->g.tiles.pair_seg > 0) {
    // two steps per pass while at least two remain
    for (; k + 1 < steps; k += 2) {
      rc = s->mat.active ? le2d::step2_xy_mat<real>(&s->g, &s->mat)
                         : le2d::step2_xy<real>(&s->g, cx, cy);
      if (rc) return rc;
    }
  }
  for (; k < steps; ++k) {
    if (what == (DO_X | DO_Y)) {
      // one pass per step
      rc = s->mat.active ? le2d::step_xy_mat<real>(&s->g, &s->mat, edge)
                         : le2d::step_xy<real>(&s->g, cx, cy, edge);
      if (rc) return rc;
      continue;
    }
    if (what & DO_X) {
      rc = le2d::sweep_x<real>(&s->g, cx);
      if (rc) return rc;
    }
    if (what & DO_Y) {
      rc = le2d::sweep_y<real>(&s->g, cy, edge);
      if (rc) return rc;
    }
  }
  if (steps > 0) s->dev_newer = true;
  if (s->mode == LE2D_SYNC_ASYNC) return LE2D_OK;
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return le2d::set_error(LE2D_ERR_DEVICE, "step: %s", hipGetErrorString(e));
  if (s->mode == LE2D_SYNC_COHERENT) return sync_host(t, s);
  return LE2D_OK;
}

void run_or_die(le_task *t, int what, int edge, const char *name, bool defer = false) {
  int rc = run(t, what, edge, 1, defer);
  if (rc) die(name, rc);
}

}  // namespace

extern "C" {

// simple/le_core.c:171-187, same expressions in `real` with double literals.
void le_init_material(const real c1, const real c2, const real rho, le_material *m) {
  m->c1 = c1;
  m->c2 = c2;
  m->rho = rho;
  m->irhoc1 = 1.0 / (c1 * rho);
  m->irhoc2 = 1.0 / (c2 * rho);
  m->rhoc1 = c1 * rho;
  m->rhoc2 = c2 * rho;
  const real mu = rho * c2 * c2;
  const real la = rho * c1 * c1 - 2.0 * mu;
  m->rhoc3 = rho * c1 * la / (la + 2.0 * mu);
}

// simple/le_core.c:114-123.  The mirror is pinned when a device exists so
// uploads and downloads run at DMA speed.
void le_init_task(le_task *task, const real dt, const le_vec2 h, const le_material mat,
                  const le_point2 n, const int stype) {
  task->dt = dt;
  task->h = h;
  task->mat = mat;
  task->n = n;
  task->stype = stype;
  const size_t bytes = sizeof(le_node) * (size_t)n.x * (size_t)n.y;
  void *p = nullptr;
  bool pinned = false;
  if (bytes && have_device() && hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess) {
    pinned = true;
  } else {
    p = malloc(bytes ? bytes : 1);
  }
  if (!p) {
    fprintf(stderr, "le2d: le_init_task: cannot allocate %zu bytes\n", bytes);
    abort();
  }
  memset(p, 0, bytes);
  task->grid = static_cast<real *>(p);
  TaskState *s = new TaskState();
  s->pinned = pinned;
  s->mode = env_mode();
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tasks.find(task->grid);
  if (it != g_tasks.end()) delete it->second;
  g_tasks[task->grid] = s;
}

// simple/le_core.c:125-128.
void le_free_task(le_task *task) {
  TaskState *s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tasks.find(task->grid);
    if (it != g_tasks.end()) {
      s = it->second;
      g_tasks.erase(it);
    }
  }
  if (s) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    le2d::mat_release(&s->mat);
    le2d::grid_free(&s->g);
    if (s->stream) (void)hipStreamDestroy(s->stream);
  }
  if (!s || s->owned) {
    if (s && s->pinned)
      (void)hipHostFree(task->grid);
    else
      free(task->grid);
  }
  delete s;
}

// simple/le_core.c:71-94.  Loop order and arithmetic as the reference; the
// distance is libm's double sqrt of a sum formed in `real`.
void le_set_ball(le_task *t, const le_vec2 c, const real r, const real s) {
  TaskState *st = find(t);
  if (sync_host(t, st)) die("le_set_ball (download)", LE2D_ERR_DEVICE);
  const long nx = t->n.x, ny = t->n.y;
  for (int i = 0; i < t->n.x; i++) {
    for (int j = 0; j < t->n.y; j++) {
      const le_vec2 x = {t->h.x * i, t->h.y * j};
      const le_vec2 d = {x.x - c.x, x.y - c.y};
      if (::sqrt((double)(d.x * d.x + d.y * d.y)) < r) {
        if (t->stype == ST_AOS) {
          le_node *nd = reinterpret_cast<le_node *>(t->grid) + (i + j * nx);
          nd->s.xx = s;
          nd->s.yy = s;
        } else if (t->stype == ST_SOA) {
          t->grid[2 * nx * ny + i + j * nx] = s;
          t->grid[4 * nx * ny + i + j * nx] = s;
        } else {
          assert(0);
        }
      }
    }
  }
  st->host_dirty = true;
}

// |v| of rows [j0, j1) as big-endian floats (simple/le_core.c:99-169: the
// speed sqrt(vx*vx + vy*vy) formed in real, rounded to float, written
// byte-swapped, row by row).
static void vtk_rows(const le_task *t, long j0, long j1, unsigned char *out) {
  const long nx = t->n.x, ny = t->n.y;
  for (long j = j0; j < j1; j++) {
    for (long i = 0; i < nx; i++) {
      real vx, vy;
      if (t->stype == ST_AOS) {
        const le_node *nd = reinterpret_cast<const le_node *>(t->grid) + (i + j * nx);
        vx = nd->v.x;
        vy = nd->v.y;
      } else {
        vx = t->grid[i + j * nx];
        vy = t->grid[nx * ny + i + j * nx];
      }
      const float v = (float)::sqrt((double)(vx * vx + vy * vy));
      uint32_t bits;
      memcpy(&bits, &v, 4);
      out[0] = (unsigned char)(bits >> 24);
      out[1] = (unsigned char)(bits >> 16);
      out[2] = (unsigned char)(bits >> 8);
      out[3] = (unsigned char)bits;
      out += 4;
    }
  }
}

// simple/le_core.c:99-169: header lines, then |v| as big-endian float32,
// rows j outer, nodes i inner.
int le_save_task(le_task *t, const char *file) {
  TaskState *st = find(t);
  int rc = sync_host(t, st);
  if (rc) die("le_save_task (download)", rc);
  FILE *fp = fopen(file, "w");
  if (fp == NULL) {
    perror("Failed to open file");
    return 1;
  }
  fprintf(fp, "# vtk DataFile Version 3.0\n");
  fprintf(fp, "Created by le_save_task\n");
  fprintf(fp, "BINARY\n");
  fprintf(fp, "DATASET STRUCTURED_POINTS\n");
  fprintf(fp, "DIMENSIONS %d %d 1\n", t->n.x, t->n.y);
  fprintf(fp, "SPACING %f %f 0.0\n", (double)t->h.x, (double)t->h.y);
  fprintf(fp, "ORIGIN 0.0 0.0 0.0\n");
  fprintf(fp, "POINT_DATA %d\n", (int)((unsigned)t->n.x * (unsigned)t->n.y));
  fprintf(fp, "SCALARS v float 1\n");
  fprintf(fp, "LOOKUP_TABLE v_table\n");
  const long nx = t->n.x, ny = t->n.y;
  const long row_bytes = nx * 4;
  // rows in chunks of about 64 MB, each chunk converted by up to 16 threads
  // and written with one fwrite
  const long chunk = std::min(ny, std::max(1L, (64L << 20) / row_bytes));
  unsigned char *buf = static_cast<unsigned char *>(malloc((size_t)(chunk * row_bytes)));
  if (!buf) {
    fprintf(stderr, "le_save_task: cannot allocate a buffer of %ld bytes\n", chunk * row_bytes);
    fclose(fp);
    return 1;
  }
  const long threads = std::max(1L, std::min(16L, (long)std::thread::hardware_concurrency()));
  for (long j0 = 0; j0 < ny; j0 += chunk) {
    const long j1 = std::min(ny, j0 + chunk), rows = j1 - j0;
    const long parts = rows * nx < (1L << 18) ? 1 : std::min(threads, rows);
    std::vector<std::thread> pool;
    // part k: rows [rows*k/parts, rows*(k+1)/parts) of the chunk; part 0 on
    // the caller's thread
    for (long k = 1; k < parts; ++k) {
      const long a = rows * k / parts, b = rows * (k + 1) / parts;
      try {
        pool.emplace_back(vtk_rows, t, j0 + a, j0 + b, buf + a * row_bytes);
      } catch (...) {  // no thread: the caller's thread converts the rest
        vtk_rows(t, j0 + a, j1, buf + a * row_bytes);
        break;
      }
    }
    vtk_rows(t, j0, j0 + rows / parts, buf);
    for (std::thread &th : pool) th.join();
    fwrite(buf, 1, (size_t)(rows * row_bytes), fp);
  }
  free(buf);
  fclose(fp);
  return 0;
}

// simple/le_core.c:435 (asserts AOS, :437).
void le_step_x(le_task *task) {
  assert(task->stype == ST_AOS);
  run_or_die(task, DO_X, le2d::EDGE_CLAMP, "le_step_x");
}

// simple/le_core.c:503 (asserts AOS, :505).
void le_step_y(le_task *task) {
  assert(task->stype == ST_AOS);
  run_or_die(task, DO_Y, le2d::EDGE_CLAMP, "le_step_y");
}

// simple/le_core.c:605-620.  Consecutive calls are run in pairs (see run).
void le_step(le_task *task) {
  assert(task->stype == ST_AOS);
  run_or_die(task, DO_X | DO_Y, le2d::EDGE_CLAMP, "le_step", true);
}

// simple/le_core.c:540-602: column blocks of width cfs; the result does not
// depend on cfs, only on the ring edge of its 5-row window.
void le_step_y_cf(le_task *task, const int_t cfs) {
  assert(task->stype == ST_AOS);
  assert(cfs > 0 && cfs <= task->n.x);
  run_or_die(task, DO_Y, le2d::EDGE_RING, "le_step_y_cf");
}

// Declared at simple/le_core.h:248, never defined by the reference; the
// natural composition of its pieces.
void le_step_cf(le_task *task, const int cfs) {
  assert(task->stype == ST_AOS);
  assert(cfs > 0 && cfs <= task->n.x);
  run_or_die(task, DO_X, le2d::EDGE_CLAMP, "le_step_cf");
  run_or_die(task, DO_Y, le2d::EDGE_RING, "le_step_cf");
}

// simple/le_core.c:367-433.
void le_soa_step_x(le_task *task) {
  assert(task->stype == ST_SOA);
  run_or_die(task, DO_X, le2d::EDGE_CLAMP, "le_soa_step_x");
}

// simple/le_core.c:259-364.
void le_soa_step_y(le_task *task) {
  assert(task->stype == ST_SOA);
  run_or_die(task, DO_Y, le2d::EDGE_RING, "le_soa_step_y");
}

// simple/le_core.c:623-627.
void le_step_soa(le_task *task) {
  assert(task->stype == ST_SOA);
  run_or_die(task, DO_X | DO_Y, le2d::EDGE_RING, "le_step_soa");
}

// ---------------------------------------------------------------------------
// le2d.h extensions.

const char *le2d_last_error(void) { return le2d::last_error(); }

int le2d_set_sync_mode(le_task *task, int mode) {
  if (mode < LE2D_SYNC_WAIT || mode > LE2D_SYNC_COHERENT)
    return le2d::set_error(LE2D_ERR_ARG, "sync mode %d", mode);
  TaskState *s = find(task);
  int rc = flush_pending(s);
  if (rc) return rc;
  s->mode = mode;
  if (mode == LE2D_SYNC_COHERENT && s->stream) {
    if (hipStreamSynchronize(s->stream) != hipSuccess)
      return le2d::set_error(LE2D_ERR_DEVICE, "sync mode: stream");
    return sync_host(task, s);
  }
  return LE2D_OK;
}

int le2d_wait(le_task *task) {
  TaskState *s = find(task);
  int rc = flush_pending(s);
  if (rc) return rc;
  if (!s->stream) return LE2D_OK;
  hipError_t e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? LE2D_OK
                         : le2d::set_error(LE2D_ERR_DEVICE, "wait: %s", hipGetErrorString(e));
}

int le_sync_host(le_task *task) { return sync_host(task, find(task)); }

int le_mark_host_dirty(le_task *task) {
  find(task)->host_dirty = true;
  return LE2D_OK;
}

int le_step_n(le_task *task, int steps) {
  if (steps < 0) return le2d::set_error(LE2D_ERR_ARG, "steps %d", steps);
  const int edge = task->stype == ST_SOA ? le2d::EDGE_RING : le2d::EDGE_CLAMP;
  return run(task, DO_X | DO_Y, edge, steps);
}

int le2d_set_tiles(le_task *task, int x_rows, int y_seg, int y_block, int xy_seg, int xy_block) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  auto width_ok = [](int w) { return w == 0 || w == 64 || w == 128 || w == 256; };
  if (x_rows < 0 || y_seg < 0 || xy_seg < 0 || !width_ok(y_block) || !width_ok(xy_block))
    return le2d::set_error(LE2D_ERR_ARG, "tiles %d %d %d %d %d", x_rows, y_seg, y_block, xy_seg,
                           xy_block);
  if (x_rows) s->g.tiles.x_rows = x_rows;
  if (y_seg) s->g.tiles.y_seg = y_seg;
  if (y_block) s->g.tiles.y_block = y_block;
  if (xy_seg) s->g.tiles.xy_seg = xy_seg;
  if (xy_block) s->g.tiles.xy_block = xy_block;
  return LE2D_OK;
}

int le2d_get_tiles(le_task *task, int *x_rows, int *y_seg, int *y_block, int *xy_seg,
                   int *xy_block) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  *x_rows = s->g.tiles.x_rows;
  *y_seg = s->g.tiles.y_seg;
  *y_block = s->g.tiles.y_block;
  *xy_seg = s->g.tiles.xy_seg;
  *xy_block = s->g.tiles.xy_block;
  return LE2D_OK;
}

int le2d_time_steps(le_task *task, int steps, double *ms_step) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (steps < 1) return le2d::set_error(LE2D_ERR_ARG, "steps %d", steps);
  const int edge = task->stype == ST_SOA ? le2d::EDGE_RING : le2d::EDGE_CLAMP;
  const SweepCoef<real> cx = coef(task, task->h.x), cy = coef(task, task->h.y);
  hipEvent_t ev[2];
  for (auto &e : ev) (void)hipEventCreate(&e);
  double tot = 0;
  for (int k = 0; k < steps && !rc; ++k) {
    (void)hipEventRecord(ev[0], s->stream);
    rc = le2d::step_xy<real>(&s->g, cx, cy, edge);
    (void)hipEventRecord(ev[1], s->stream);
    (void)hipEventSynchronize(ev[1]);
    float a = 0;
    (void)hipEventElapsedTime(&a, ev[0], ev[1]);
    tot += a;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->dev_newer = true;
  if (rc) return rc;
  *ms_step = tot / steps;
  return LE2D_OK;
}

int le2d_set_step_pairs(le_task *task, int seg) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (seg < 0) return le2d::set_error(LE2D_ERR_ARG, "step pairs: segment %d", seg);
  s->g.tiles.pair_seg = seg;
  s->pairs_fixed = true;
  return LE2D_OK;
}

int le2d_tune_step_pairs(le_task *task, int *seg) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (s->g.tiles.pair_seg < 1) s->g.tiles.pair_seg = 64;
  s->pairs_fixed = false;
  {
    std::lock_guard<std::mutex> lk(g_pair_mu);
    g_pair_seg.erase(std::make_tuple((int)sizeof(real), s->g.nx, s->g.ny));
  }
  rc = tune_pairs(task, s);
  if (rc) return rc;
  if (seg) *seg = s->g.tiles.pair_seg;
  return LE2D_OK;
}

int le2d_set_step_pair_window(le_task *task, int lanes) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (lanes != 64 && lanes != 128 && lanes != 256)
    return le2d::set_error(LE2D_ERR_ARG, "step pairs: window %d", lanes);
  s->g.tiles.pair_block = lanes;
  s->pairs_fixed = true;
  return LE2D_OK;
}

int le2d_set_step_pair_split(le_task *task, int on) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (on != 0 && on != 1) return le2d::set_error(LE2D_ERR_ARG, "step pairs: split %d", on);
  s->g.tiles.pair_split = on;
  s->pairs_fixed = true;
  return LE2D_OK;
}

int le2d_get_step_pair_split(le_task *task, int *on) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  *on = s->g.tiles.pair_split;
  return LE2D_OK;
}

int le2d_get_step_pair_window(le_task *task, int *lanes) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  *lanes = s->g.tiles.pair_block;
  return LE2D_OK;
}

int le2d_get_step_pairs(le_task *task, int *seg) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  *seg = s->g.tiles.pair_seg;
  return LE2D_OK;
}

int le2d_time_pairs(le_task *task, int passes, double *ms_pass) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (passes < 1 || s->g.tiles.pair_seg < 1 || task->stype != ST_AOS)
    return le2d::set_error(LE2D_ERR_ARG, "time pairs: %d passes, segment %d", passes,
                           s->g.tiles.pair_seg);
  const SweepCoef<real> cx = coef(task, task->h.x), cy = coef(task, task->h.y);
  hipEvent_t ev[2];
  for (auto &e : ev) (void)hipEventCreate(&e);
  double tot = 0;
  for (int k = 0; k < passes && !rc; ++k) {
    (void)hipEventRecord(ev[0], s->stream);
    rc = le2d::step2_xy<real>(&s->g, cx, cy);
    (void)hipEventRecord(ev[1], s->stream);
    (void)hipEventSynchronize(ev[1]);
    float a = 0;
    (void)hipEventElapsedTime(&a, ev[0], ev[1]);
    tot += a;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->dev_newer = true;
  if (rc) return rc;
  *ms_pass = tot / passes;
  return LE2D_OK;
}

int le2d_set_material_plane(le_task *task, int count, const double *c1_c2_rho,
                            const unsigned char *ids) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (count == 0) {
    le2d::mat_release(&s->mat);
    s->mat.active = false;
    return LE2D_OK;
  }
  s->mat_split_tuned = false;
  return le2d::mat_set(&s->g, &s->mat, count, c1_c2_rho, ids, (double)task->dt,
                       (double)task->h.x, (double)task->h.y);
}

int le2d_energy(le_task *task, double *e) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  double sum = 0.0;
  rc = le2d::energy(&s->g, &s->mat, (double)task->mat.rho, (double)task->mat.c1,
                    (double)task->mat.c2, &sum);
  if (rc) return rc;
  *e = sum * (double)task->h.x * (double)task->h.y;
  return LE2D_OK;
}

int le2d_task_view(le_task *task, le2d_view *v) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  v->base = s->g.buf[s->g.cur];
  v->elem = s->g.elem;
  v->nx = s->g.nx;
  v->ny = s->g.ny;
  v->pitch = s->g.pitch;
  v->plane = s->g.plane;
  return LE2D_OK;
}

int le2d_diff_norms(const le2d_view *a, const le2d_view *ref, double out[4]) {
  if (!a || !ref || a->nx != ref->nx || a->ny != ref->ny)
    return le2d::set_error(LE2D_ERR_ARG, "diff_norms: views of different grids");
  (void)hipDeviceSynchronize();  // the views' owners may have work queued on their streams
  return le2d::diff_norms(a->base, a->elem, a->pitch, a->plane, ref->base, ref->elem, ref->pitch,
                          ref->plane, a->nx, a->ny, nullptr, out);
}

int le2d_diff_norms_noalloc(const le2d_view *a, const le2d_view *ref, double out[4]) {
  if (!a || !ref || a->nx != ref->nx || a->ny != ref->ny)
    return le2d::set_error(LE2D_ERR_ARG, "diff_norms_noalloc: views of different grids");
  if ((a->elem != 4 && a->elem != 8) || (ref->elem != 4 && ref->elem != 8) || a->nx < 1 ||
      a->ny < 1)
    return le2d::set_error(LE2D_ERR_ARG, "diff_norms_noalloc: bad view");
  (void)hipDeviceSynchronize();  // the views' owners may have work queued on their streams
  return le2d::diff_norms_noalloc(a->base, a->elem, a->pitch, a->plane, ref->base, ref->elem,
                                  ref->pitch, ref->plane, a->nx, a->ny, out);
}

int le2d_limiter(void) { return LE2D_LIMITER; }

int le2d_copy_probe(size_t bytes, int iters, double *ms) {
  return le2d::copy_probe(bytes, iters, ms);
}

int le2d_time_sweeps(le_task *task, int steps, double *ms_x, double *ms_y) {
  TaskState *s = find(task);
  int rc = ensure_device(task, s);
  if (rc) return rc;
  if (steps < 1) return le2d::set_error(LE2D_ERR_ARG, "steps %d", steps);
  const int edge = task->stype == ST_SOA ? le2d::EDGE_RING : le2d::EDGE_CLAMP;
  const SweepCoef<real> cx = coef(task, task->h.x), cy = coef(task, task->h.y);
  hipEvent_t ev[3];
  for (auto &e : ev) (void)hipEventCreate(&e);
  double tx = 0, ty = 0;
  for (int k = 0; k < steps && !rc; ++k) {
    (void)hipEventRecord(ev[0], s->stream);
    rc = le2d::sweep_x<real>(&s->g, cx);
    (void)hipEventRecord(ev[1], s->stream);
    if (!rc) rc = le2d::sweep_y<real>(&s->g, cy, edge);
    (void)hipEventRecord(ev[2], s->stream);
    (void)hipEventSynchronize(ev[2]);
    float a = 0, b = 0;
    (void)hipEventElapsedTime(&a, ev[0], ev[1]);
    (void)hipEventElapsedTime(&b, ev[1], ev[2]);
    tx += a;
    ty += b;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->dev_newer = true;
  if (rc) return rc;
  *ms_x = tx / steps;
  *ms_y = ty / steps;
  return LE2D_OK;
}

size_t le2d_device_bytes(le_task *task) {
  TaskState *s = find(task);
  return s->dev_valid ? 2 * le2d::grid_bytes(s->g.elem, s->g.nx, s->g.ny) : 0;
}

// Raw state file: a 128-byte header (magic, sizeof(real), nx, ny, stype, then
// dt, h.x, h.y and the eight material constants as doubles) followed by the
// five fields as planes vx|vy|sxx|sxy|syy of nx*ny little-endian reals, node
// (i, j) at i + j*nx.  Restoring it reproduces the dumped state bit for bit.
namespace {
const char kMagic[8] = {'L', 'E', '2', 'D', 'R', 'A', 'W', '1'};
struct RawHeader {
  char magic[8];
  int32_t real_size, nx, ny, stype;
  double p[11];  // dt, h.x, h.y, c1, c2, rho, irhoc1, irhoc2, rhoc1, rhoc2, rhoc3
  char pad[128 - 24 - 88];
};
static_assert(sizeof(RawHeader) == 128, "raw header is 128 bytes");
}  // namespace

int le2d_dump(le_task *task, const char *path) {
  TaskState *s = find(task);
  int rc = sync_host(task, s);
  if (rc) return rc;
  FILE *fp = fopen(path, "wb");
  if (!fp) return le2d::set_error(LE2D_ERR_ARG, "dump: cannot open %s", path);
  RawHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, kMagic, 8);
  h.real_size = (int32_t)sizeof(real);
  h.nx = task->n.x;
  h.ny = task->n.y;
  h.stype = task->stype;
  const le_material &m = task->mat;
  const double p[11] = {task->dt, task->h.x, task->h.y, m.c1, m.c2, m.rho,
                        m.irhoc1, m.irhoc2, m.rhoc1, m.rhoc2, m.rhoc3};
  memcpy(h.p, p, sizeof p);
  bool ok = fwrite(&h, sizeof h, 1, fp) == 1;
  const size_t n = (size_t)task->n.x * task->n.y;
  if (task->stype == ST_SOA) {
    ok = ok && fwrite(task->grid, sizeof(real), 5 * n, fp) == 5 * n;
  } else {
    real *buf = static_cast<real *>(malloc(n * sizeof(real)));
    for (int q = 0; q < 5 && ok && buf; ++q) {
      for (size_t k = 0; k < n; ++k) buf[k] = task->grid[5 * k + q];
      ok = fwrite(buf, sizeof(real), n, fp) == n;
    }
    ok = ok && buf;
    free(buf);
  }
  ok = (fclose(fp) == 0) && ok;
  return ok ? LE2D_OK : le2d::set_error(LE2D_ERR_ARG, "dump: short write to %s", path);
}

int le2d_restore(le_task *task, const char *path) {
  TaskState *s = find(task);
  FILE *fp = fopen(path, "rb");
  if (!fp) return le2d::set_error(LE2D_ERR_ARG, "restore: cannot open %s", path);
  RawHeader h;
  if (fread(&h, sizeof h, 1, fp) != 1 || memcmp(h.magic, kMagic, 8) != 0) {
    fclose(fp);
    return le2d::set_error(LE2D_ERR_ARG, "restore: %s is not a raw state file", path);
  }
  if (h.real_size != (int32_t)sizeof(real) || h.nx != task->n.x || h.ny != task->n.y) {
    fclose(fp);
    return le2d::set_error(LE2D_ERR_ARG,
                           "restore: file holds %dx%d of %d-byte reals, task is %dx%d of %d",
                           h.nx, h.ny, h.real_size, task->n.x, task->n.y, (int)sizeof(real));
  }
  const size_t n = (size_t)task->n.x * task->n.y;
  real *buf = static_cast<real *>(malloc(5 * n * sizeof(real)));
  bool ok = buf && fread(buf, sizeof(real), 5 * n, fp) == 5 * n;
  fclose(fp);
  if (!ok) {
    free(buf);
    return le2d::set_error(LE2D_ERR_ARG, "restore: %s is truncated", path);
  }
  task->dt = (real)h.p[0];
  task->h.x = (real)h.p[1];
  task->h.y = (real)h.p[2];
  real *mp[8] = {&task->mat.c1, &task->mat.c2, &task->mat.rho, &task->mat.irhoc1,
                 &task->mat.irhoc2, &task->mat.rhoc1, &task->mat.rhoc2, &task->mat.rhoc3};
  for (int k = 0; k < 8; ++k) *mp[k] = (real)h.p[3 + k];
  if (task->stype == ST_SOA) {
    memcpy(task->grid, buf, 5 * n * sizeof(real));
  } else {
    for (int q = 0; q < 5; ++q)
      for (size_t k = 0; k < n; ++k) task->grid[5 * k + q] = buf[q * n + k];
  }
  free(buf);
  s->host_dirty = true;
  s->dev_newer = false;
  s->pending = false;
  return LE2D_OK;
}

}  // extern "C"
