// batch.hip -- many independent grids of uneven sizes and materials advanced
// by one launch per step (persistent, or one workgroup per tile) or per two
// steps, captured as a hipGraph.
//
// Every instance is a ping-pong pair of state buffers (the single-grid
// layout) carved from one arena.  A step is one launch of k_step_batch over
// a table of tiles (instance, strip, row segment) of all instances; each
// tile runs the fused-step march of sweep_dev.h with its instance's pointers
// and constants, so a batched instance is bit-identical to the same grid
// stepped alone (FP64) or agrees to FP32 rounding (FP32).  All instances
// advance together and share one buffer parity.
//
// The launch shape -- window width (64, 128 or 256 lanes), rows per segment,
// and the launch form -- is chosen by le2d_batch_autotune, which times candidate shapes
// on the batch itself and keeps the fastest (per element size, instance
// count and node total, for the life of the process).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "batch_dev.h"
#include "le2d.h"

namespace le2d {

// ---------------------------------------------------------------------------

namespace {

struct Inst {
  int nx, ny;
  double dt, hx, hy, c1, c2, rho;
  long pitch, plane;
  size_t off[2];
};

// Launch shape: window width, rows per segment, and the launch form.
enum { FORM_TILES = 0, FORM_PERSISTENT = 1, FORM_PAIRS = 2, FORM_PAIRS_SPLIT = 3 };
inline bool two_step(int form) { return form == FORM_PAIRS || form == FORM_PAIRS_SPLIT; }
struct Shape {
  int fb, seg;
  int form;  // FORM_TILES: one workgroup per tile; FORM_PERSISTENT: resident
             // workgroups walk the table; FORM_PAIRS: one workgroup per tile of
             // the two-step table, two steps per launch (fb 64, 128 or 256);
             // FORM_PAIRS_SPLIT: the same with 64-lane windows whose two
             // stages run on two wavefronts (march_pair_split)
};

std::mutex g_tune_mu;
std::map<std::tuple<int, int, long>, Shape> g_tune;  // (elem, instances, Mnodes) -> shape

template <typename T>
void inst_coef(const Inst &in, SweepCoef<T> *cx, SweepCoef<T> *cy) {
  coef_from_material<T>(in.c1, in.c2, in.rho, in.dt, in.hx, in.hy, cx, cy);
}

int hcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) return set_error(ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return OK;
}

}  // namespace

struct BatchImpl {
  int elem = 8;
  std::vector<Inst> inst;
  bool committed = false;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  void *d_inst = nullptr;
  TileJob *d_tiles = nullptr;
  TileJob *d_tiles1 = nullptr;  // FORM_PAIRS: single-step table for odd remainders
  int ntiles1 = 0;
  int ntiles = 0;
  Shape shape{256, 64, FORM_PAIRS};
  int blocks = 0;
  int parity = 0;
  hipStream_t stream = nullptr;
  // graph[p]: two steps starting from buffer parity p (replays leave p)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  long steps_done = 0;

  void drop_graphs() {
    for (int p = 0; p < 2; ++p) {
      if (exec[p]) (void)hipGraphExecDestroy(exec[p]);
      if (graph[p]) (void)hipGraphDestroy(graph[p]);
      exec[p] = nullptr;
      graph[p] = nullptr;
    }
  }

  ~BatchImpl() {
    drop_graphs();
    if (stream) (void)hipStreamSynchronize(stream);
    dev_free(d_tiles);
    dev_free(d_tiles1);
    dev_free(d_inst);
    dev_free(arena);
    if (stream) (void)hipStreamDestroy(stream);
  }

  long total_nodes() const {
    long n = 0;
    for (const Inst &in : inst) n += (long)in.nx * in.ny;
    return n;
  }

  DevGrid view(int i) {
    DevGrid g;
    const Inst &in = inst[i];
    g.elem = elem;
    g.nx = in.nx;
    g.ny = in.ny;
    g.pitch = in.pitch;
    g.plane = in.plane;
    g.buf[0] = arena + in.off[0];
    g.buf[1] = arena + in.off[1];
    g.cur = parity;
    g.stream = stream;
    g.tiles = default_tiles(elem, in.nx, in.ny);
    return g;
  }

  template <typename T>
  int upload_desc() {
    std::vector<InstDesc<T>> h(inst.size());
    for (size_t i = 0; i < inst.size(); ++i) {
      const Inst &in = inst[i];
      h[i].buf[0] = reinterpret_cast<T *>(arena + in.off[0]);
      h[i].buf[1] = reinterpret_cast<T *>(arena + in.off[1]);
      h[i].nx = in.nx;
      h[i].ny = in.ny;
      h[i].pitch = in.pitch;
      h[i].plane = in.plane;
      inst_coef<T>(in, &h[i].cx, &h[i].cy);
      h[i].par = 0;
    }
    if (!d_inst) {
      const int rc = dev_alloc(&d_inst, h.size() * sizeof(InstDesc<T>), "batch: descriptor table");
      if (rc) return rc;
    }
    return hcheck(hipMemcpy(d_inst, h.data(), h.size() * sizeof(InstDesc<T>), hipMemcpyHostToDevice),
                  "batch: descriptors");
  }

  // Tile table: instances in order, then segments, then strips, so
  // neighbouring entries share halo columns.  `width` output columns per
  // strip.
  int upload_table(int width, int seg_rows, TileJob **d, int *n) {
    std::vector<TileJob> h;
    for (size_t i = 0; i < inst.size(); ++i) {
      const Inst &in = inst[i];
      const int strips = (in.nx + width - 1) / width;
      const int seg = std::max(1, std::min(seg_rows, in.ny));
      for (int j0 = 0; j0 < in.ny; j0 += seg)
        for (int st = 0; st < strips; ++st)
          h.push_back({(int)i, st, j0, std::min(j0 + seg, in.ny), j0 / seg});
    }
    dev_free(*d);
    *d = nullptr;
    const int rc = dev_alloc((void **)d, std::max<size_t>(1, h.size()) * sizeof(TileJob),
                             "batch: tile table");
    if (rc) return rc;
    *n = (int)h.size();
    return hcheck(hipMemcpy(*d, h.data(), h.size() * sizeof(TileJob), hipMemcpyHostToDevice),
                  "batch: tiles");
  }

  int build_tiles(Shape s) {
    if (s.form == FORM_PAIRS && s.fb != 64 && s.fb != 128) s.fb = 256;
    if (s.form == FORM_PAIRS_SPLIT) s.fb = 64;
    int rc = upload_table(two_step(s.form) ? s.fb - 8 : s.fb - 4, s.seg, &d_tiles, &ntiles);
    if (rc) return rc;
    if (two_step(s.form)) {
      rc = upload_table(s.fb - 4, s.seg, &d_tiles1, &ntiles1);
      if (rc) return rc;
    }
    shape = s;
    // persistent grid: resident workgroups on the whole device; else one
    // workgroup per tile
    if (s.form != FORM_PERSISTENT)
      blocks = ntiles;
    else if (elem == 8)
      blocks = s.fb == 64 ? persistent_blocks<double, 64>(ntiles)
                          : (s.fb == 128 ? persistent_blocks<double, 128>(ntiles)
                                         : persistent_blocks<double, 256>(ntiles));
    else
      blocks = s.fb == 64 ? persistent_blocks<float, 64>(ntiles)
                          : (s.fb == 128 ? persistent_blocks<float, 128>(ntiles)
                                         : persistent_blocks<float, 256>(ntiles));
    drop_graphs();
    return OK;
  }

  template <typename T, int FB>
  void launch_t(int par) {
    if (ntiles <= 0) return;
    const InstDesc<T> *d = static_cast<const InstDesc<T> *>(d_inst);
    if (shape.form == FORM_PERSISTENT)
      hipLaunchKernelGGL((k_step_batch<T, FB>), dim3(blocks), dim3(FB), 0, stream, d, d_tiles,
                         ntiles, par);
    else if (shape.form == FORM_PAIRS_SPLIT)
      hipLaunchKernelGGL((k_step_batch_pairs_split<T>), dim3(ntiles), dim3(128), 0, stream, d,
                         d_tiles, par);
    else if (shape.form == FORM_PAIRS)
      hipLaunchKernelGGL((k_step_batch_pairs<T, FB>), dim3(ntiles), dim3(FB), 0, stream, d,
                         d_tiles, par);
    else
      hipLaunchKernelGGL((k_step_batch_tiles<T, FB>), dim3(ntiles), dim3(FB), 0, stream, d,
                         d_tiles, par);
  }

  // FORM_PAIRS: one single step (the odd remainder of a step count), same
  // window width.
  template <typename T, int FB>
  void launch_single_t(int par) {
    hipLaunchKernelGGL((k_step_batch_tiles<T, FB>), dim3(ntiles1), dim3(FB), 0, stream,
                       static_cast<const InstDesc<T> *>(d_inst), d_tiles1, par);
  }

  void launch_single(int par) {
    if (ntiles1 <= 0) return;
    if (elem == 8) {
      if (shape.fb == 64) launch_single_t<double, 64>(par);
      else if (shape.fb == 128) launch_single_t<double, 128>(par);
      else launch_single_t<double, 256>(par);
    } else {
      if (shape.fb == 64) launch_single_t<float, 64>(par);
      else if (shape.fb == 128) launch_single_t<float, 128>(par);
      else launch_single_t<float, 256>(par);
    }
  }

  void launch(int par) {
    if (elem == 8) {
      if (shape.fb == 64) launch_t<double, 64>(par);
      else if (shape.fb == 128) launch_t<double, 128>(par);
      else launch_t<double, 256>(par);
    } else {
      if (shape.fb == 64) launch_t<float, 64>(par);
      else if (shape.fb == 128) launch_t<float, 128>(par);
      else launch_t<float, 256>(par);
    }
  }

  // Two steps (parity p then p^1) captured as a graph.
  int ensure_graph(int p) {
    if (exec[p]) return OK;
    int rc = hcheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
    if (rc) return rc;
    launch(p);
    launch(p ^ 1);
    rc = hcheck(hipStreamEndCapture(stream, &graph[p]), "end capture");
    if (rc) return rc;
    return hcheck(hipGraphInstantiate(&exec[p], graph[p], nullptr, nullptr, 0), "instantiate");
  }

  int step(int steps, bool use_graph) {
    if (!committed) return set_error(ERR_STATE, "batch: not committed");
    const int spl = two_step(shape.form) ? 2 : 1;  // steps per launch
    int k = 0;
    if (use_graph && steps >= 2 * spl) {
      int rc = ensure_graph(parity);  // two launches: the parity comes back
      if (rc) return rc;
      for (; k + 2 * spl <= steps; k += 2 * spl) {
        rc = hcheck(hipGraphLaunch(exec[parity], stream), "graph launch");
        if (rc) return rc;
      }
    }
    for (; k + spl <= steps; k += spl) {
      launch(parity);
      parity ^= 1;
    }
    for (; k < steps; ++k) {  // odd remainder of two-step launches
      launch_single(parity);
      parity ^= 1;
    }
    steps_done += steps;
    return hcheck(hipGetLastError(), "batch step");
  }
};

}  // namespace le2d

using le2d::BatchImpl;

struct le2d_batch {
  BatchImpl impl;
};

extern "C" {

le2d_batch *le2d_batch_create(int real_size) {
  if (real_size != 4 && real_size != 8) {
    le2d::set_error(le2d::ERR_ARG, "batch: real_size %d", real_size);
    return nullptr;
  }
  le2d::require_device();
  le2d_batch *b = new le2d_batch();
  b->impl.elem = real_size;
  if (hipStreamCreateWithFlags(&b->impl.stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    le2d::set_error(le2d::ERR_DEVICE, "batch: stream");
    return nullptr;
  }
  return b;
}

void le2d_batch_destroy(le2d_batch *b) { delete b; }

int le2d_batch_add(le2d_batch *b, int nx, int ny, double dt, double hx, double hy, double c1,
                   double c2, double rho) {
  if (!b || b->impl.committed) return -le2d::set_error(le2d::ERR_STATE, "batch: add after commit");
  if (nx < 3 || ny < 3 || !le2d::row_fits(b->impl.elem, nx) || !(dt > 0) || !(hx > 0) ||
      !(hy > 0) || !(c1 > 0) || !(c2 > 0) || !(rho > 0))
    return -le2d::set_error(le2d::ERR_ARG, "batch: bad instance %dx%d", nx, ny);
  le2d::Inst in{};
  in.nx = nx;
  in.ny = ny;
  in.dt = dt;
  in.hx = hx;
  in.hy = hy;
  in.c1 = c1;
  in.c2 = c2;
  in.rho = rho;
  b->impl.inst.push_back(in);
  return (int)b->impl.inst.size() - 1;
}

int le2d_batch_count(le2d_batch *b) { return b ? (int)b->impl.inst.size() : 0; }

size_t le2d_batch_device_bytes(le2d_batch *b) { return b ? b->impl.arena_bytes : 0; }

int le2d_batch_commit(le2d_batch *b) {
  if (!b) return le2d::ERR_ARG;
  BatchImpl &B = b->impl;
  if (B.committed) return le2d::OK;
  if (B.inst.empty()) return le2d::set_error(le2d::ERR_ARG, "batch: empty");
  size_t off = 0;
  for (auto &in : B.inst) {
    le2d::grid_layout(B.elem, in.nx, in.ny, &in.pitch, &in.plane);
    const size_t bytes = le2d::grid_bytes(B.elem, in.nx, in.ny);
    for (int k = 0; k < 2; ++k) {
      in.off[k] = off;
      off += (bytes + 4095) / 4096 * 4096;
    }
  }
  (void)hipStreamSynchronize(B.stream);
  le2d::dev_free(B.arena);  // a commit that failed part way left these
  le2d::dev_free(B.d_inst);
  B.d_inst = nullptr;
  B.arena = nullptr;
  B.arena_bytes = 0;
  int rc = le2d::dev_alloc((void **)&B.arena, off, "batch: arena");
  if (rc) return rc;
  B.arena_bytes = off;
  rc = le2d::hcheck(hipMemsetAsync(B.arena, 0, off, B.stream), "batch: clear");
  if (rc) return rc;
  rc = B.elem == 8 ? B.upload_desc<double>() : B.upload_desc<float>();
  if (rc) return rc;
  // default shape until autotuned: the cached winner for this batch, else
  // the two-step form in one-wave windows (the autotuner's choice in FP64,
  // and within a few percent of it in FP32 on mixed batches) with the
  // segment height that maximises (fill of the last round of resident
  // workgroups) x (share of rows that are output rather than halo) over all
  // instances
  le2d::Shape s{64, 64, le2d::FORM_PAIRS};
  {
    long resident = le2d::pair_resident(B.elem, 64);
    if (resident <= 0) resident = 2048;
    double best_score = -1;
    for (int seg : {8, 16, 32, 64, 128}) {
      long blocks = 0;
      for (const auto &in : B.inst) blocks += (long)((in.nx + 55) / 56) * ((in.ny + seg - 1) / seg);
      const long rounds = (blocks + resident - 1) / resident;
      const double score = (double)blocks / (double)(rounds * resident) * seg / (seg + 8);
      if (score > best_score + 1e-9) {
        best_score = score;
        s.seg = seg;
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(le2d::g_tune_mu);
    auto it = le2d::g_tune.find({B.elem, (int)B.inst.size(), B.total_nodes() >> 20});
    if (it != le2d::g_tune.end()) s = it->second;
  }
  rc = B.build_tiles(s);
  if (rc) return rc;
  B.committed = true;
  return le2d::hcheck(hipStreamSynchronize(B.stream), "batch: commit");
}

int le2d_batch_set_ball(le2d_batch *b, int i, double cx, double cy, double r, double s) {
  if (!b || !b->impl.committed || i < 0 || i >= (int)b->impl.inst.size())
    return le2d::set_error(le2d::ERR_ARG, "batch: set_ball instance %d", i);
  BatchImpl &B = b->impl;
  const le2d::Inst &in = B.inst[i];
  const long n = (long)in.nx * in.ny;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  void *buf = B.arena + in.off[B.parity];
  if (B.elem == 8)
    hipLaunchKernelGGL(le2d::k_set_ball<double>, dim3(blocks), dim3(256), 0, B.stream,
                       (double *)buf, in.nx, in.ny, in.pitch, in.plane, in.hx, in.hy, cx, cy, r, s);
  else
    hipLaunchKernelGGL(le2d::k_set_ball<float>, dim3(blocks), dim3(256), 0, B.stream, (float *)buf,
                       in.nx, in.ny, in.pitch, in.plane, (float)in.hx, (float)in.hy, (float)cx,
                       (float)cy, (float)r, (float)s);
  return le2d::hcheck(hipGetLastError(), "batch: set_ball");
}

int le2d_batch_load(le2d_batch *b, int i, const void *aos) {
  if (!b || !b->impl.committed || i < 0 || i >= (int)b->impl.inst.size())
    return le2d::set_error(le2d::ERR_ARG, "batch: load instance %d", i);
  le2d::DevGrid g = b->impl.view(i);
  return le2d::grid_upload(&g, aos, 0);
}

int le2d_batch_read(le2d_batch *b, int i, void *aos) {
  if (!b || !b->impl.committed || i < 0 || i >= (int)b->impl.inst.size())
    return le2d::set_error(le2d::ERR_ARG, "batch: read instance %d", i);
  int rc = le2d::hcheck(hipStreamSynchronize(b->impl.stream), "batch: read sync");
  if (rc) return rc;
  le2d::DevGrid g = b->impl.view(i);
  return le2d::grid_download(&g, aos, 0);
}

int le2d_batch_step(le2d_batch *b, int steps) {
  if (!b || steps < 0) return le2d::set_error(le2d::ERR_ARG, "batch: step %d", steps);
  return b->impl.step(steps, true);
}

int le2d_batch_step_plain(le2d_batch *b, int steps) {
  if (!b || !b->impl.committed || steps < 0)
    return le2d::set_error(le2d::ERR_ARG, "batch: plain step");
  BatchImpl &B = b->impl;
  // what stepping each grid alone through le_step_n does: two steps per
  // launch while two remain (its default segment height), then one
  for (int k = 0; k < steps;) {
    const bool pair = k + 1 < steps;
    for (size_t i = 0; i < B.inst.size(); ++i) {
      le2d::DevGrid g = B.view((int)i);
      int rc;
      if (B.elem == 8) {
        le2d::SweepCoef<double> cx, cy;
        le2d::inst_coef<double>(B.inst[i], &cx, &cy);
        rc = pair ? le2d::step2_xy<double>(&g, cx, cy)
                  : le2d::step_xy<double>(&g, cx, cy, le2d::EDGE_CLAMP);
      } else {
        le2d::SweepCoef<float> cx, cy;
        le2d::inst_coef<float>(B.inst[i], &cx, &cy);
        rc = pair ? le2d::step2_xy<float>(&g, cx, cy)
                  : le2d::step_xy<float>(&g, cx, cy, le2d::EDGE_CLAMP);
      }
      if (rc) return rc;
    }
    B.parity ^= 1;
    k += pair ? 2 : 1;
  }
  B.steps_done += steps;
  return le2d::OK;
}

int le2d_batch_wait(le2d_batch *b) {
  if (!b) return le2d::ERR_ARG;
  return le2d::hcheck(hipStreamSynchronize(b->impl.stream), "batch: wait");
}

int le2d_batch_set_launch(le2d_batch *b, int fb, int seg, int form) {
  if (!b || !b->impl.committed || (fb != 64 && fb != 128 && fb != 256) || seg < 1 ||
      form < le2d::FORM_TILES || form > le2d::FORM_PAIRS_SPLIT)
    return le2d::set_error(le2d::ERR_ARG, "batch: shape %d %d form %d", fb, seg, form);
  return b->impl.build_tiles({fb, seg, form});
}

int le2d_batch_get_launch(le2d_batch *b, int *form) {
  if (!b || !form) return le2d::ERR_ARG;
  *form = b->impl.shape.form;
  return le2d::OK;
}

int le2d_batch_set_shape(le2d_batch *b, int fb, int seg) {
  if (!b) return le2d::ERR_ARG;
  return le2d_batch_set_launch(b, fb, seg, b->impl.shape.form);
}

int le2d_batch_get_shape(le2d_batch *b, int *fb, int *seg, int *ntiles, int *blocks) {
  if (!b) return le2d::ERR_ARG;
  *fb = b->impl.shape.fb;
  *seg = b->impl.shape.seg;
  *ntiles = b->impl.ntiles;
  *blocks = b->impl.blocks;
  return le2d::OK;
}

// Time `steps` steps of the batched launch (graph) or the per-instance loop;
// mean ms per step, HIP events on the batch stream.
int le2d_batch_time(le2d_batch *b, int steps, int plain, double *ms) {
  if (!b || steps < 1) return le2d::ERR_ARG;
  BatchImpl &B = b->impl;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, B.stream);
  int rc = plain ? le2d_batch_step_plain(b, steps) : le2d_batch_step(b, steps);
  (void)hipEventRecord(e1, B.stream);
  hipError_t e = hipEventSynchronize(e1);
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (e != hipSuccess) return le2d::set_error(le2d::ERR_DEVICE, "batch time: %s", hipGetErrorString(e));
  *ms = t / steps;
  return le2d::OK;
}

// Times every candidate shape and keeps the fastest (per step): the
// single-step forms at every window width and segment height, and the
// two-step form at every segment height.  Stepping the batch is part of
// tuning: the instances advance by an even number of steps per candidate,
// so callers tune before loading their initial state or account for it.
int le2d_batch_autotune(le2d_batch *b, int trials, int *best_fb, int *best_seg, double *best_ms) {
  if (!b || !b->impl.committed || trials < 1) return le2d::ERR_ARG;
  BatchImpl &B = b->impl;
  const int fbs[3] = {64, 128, 256};
  const int segs[4] = {16, 32, 64, 128};
  std::vector<le2d::Shape> cand;
  for (int form : {le2d::FORM_TILES, le2d::FORM_PERSISTENT})
    for (int fb : fbs)
      for (int seg : segs) cand.push_back({fb, seg, form});
  for (int fb : fbs)
    for (int seg : segs) cand.push_back({fb, seg, le2d::FORM_PAIRS});
  for (int seg : segs) cand.push_back({64, seg, le2d::FORM_PAIRS_SPLIT});
  le2d::Shape best{B.shape};
  double best_t = 1e30;
  for (const le2d::Shape &c : cand) {
    int rc = B.build_tiles(c);
    if (rc) return rc;
    const int spl = le2d::two_step(c.form) ? 2 : 1;
    rc = B.step(2 * spl, true);  // captures the graph, warms up
    if (rc) return rc;
    double ms = 0;
    rc = le2d_batch_time(b, 2 * spl * trials, 0, &ms);
    if (rc) return rc;
    if (ms < best_t) {
      best_t = ms;
      best = c;
    }
  }
  int rc = B.build_tiles(best);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(le2d::g_tune_mu);
    le2d::g_tune[{B.elem, (int)B.inst.size(), B.total_nodes() >> 20}] = best;
  }
  if (best_fb) *best_fb = best.fb;
  if (best_seg) *best_seg = best.seg;
  if (best_ms) *best_ms = best_t;
  return le2d::OK;
}

int le2d_batch_view(le2d_batch *b, int i, le2d_view *v) {
  if (!b || !b->impl.committed || i < 0 || i >= (int)b->impl.inst.size())
    return le2d::set_error(le2d::ERR_ARG, "batch: view instance %d", i);
  const BatchImpl &B = b->impl;
  const le2d::Inst &in = B.inst[i];
  v->base = B.arena + in.off[B.parity];
  v->elem = B.elem;
  v->nx = in.nx;
  v->ny = in.ny;
  v->pitch = in.pitch;
  v->plane = in.plane;
  return le2d::OK;
}

}  // extern "C"
