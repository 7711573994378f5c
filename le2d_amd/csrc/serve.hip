// serve.hip -- a serving loop for simulation requests: admission by HBM
// budget and slots, fields resident on the device, continuous batching of the
// live requests' steps.
//
// A request is one grid (size, material, disk source, step count, FP32 or
// FP64).  An engine thread runs chunks: it admits queued requests in arrival
// order while the budget (a device arena sub-allocated first-fit) and the
// slot count allow -- a request that does not fit yet may be passed by later
// ones that do, a bounded number of times -- initialises their fields on
// the device, then launches
// `chunk` steps in which every live request advances by one step per launch
// -- one k_step_batch_pairs launch (one workgroup per tile, two steps) per
// precision per two steps, and a k_step_batch_tiles launch for an odd last
// step, FP32 and FP64 on their own streams so they overlap.  The chunk ends when the first live
// request reaches its step count (or after at most `max_chunk` steps); that
// request completes there, its memory returns to the arena and waiting
// requests are admitted.  The arithmetic of a request does not depend on
// which other requests share its launches, so its result is bit-identical
// alone and in any mix.
//
// With plain = 1 the server is the serial baseline: one slot, and each
// request runs with the single-grid kernels as le_step_n uses them (k_step2_xy
// while two steps remain, then k_step_xy) and their default tiles.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_dev.h"
#include "le2d.h"

namespace le2d {
namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) in 64 KiB units, coalescing on free.
class Arena {
 public:
  explicit Arena(size_t size) { free_[0] = size; }
  bool alloc(size_t bytes, size_t *off) {
    bytes = (bytes + kUnit - 1) / kUnit * kUnit;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + kUnit - 1) / kUnit * kUnit;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }

 private:
  static constexpr size_t kUnit = 65536;
  std::map<size_t, size_t> free_;
  size_t used_ = 0;
};

// States in order: a request is finished from DONE on.  RELEASING: a
// result download holds the request (a second le2d_server_result or a
// cancel finds it busy).  CANCELLED: removed before or while it ran, or its
// kept result discarded.
enum ReqState { QUEUED = 0, LIVE = 1, DONE = 2, RELEASING = 3, RELEASED = 4, CANCELLED = 5 };

struct Req {
  le2d_request spec;
  int state = QUEUED;
  int passed = 0;       // later requests admitted ahead of this one
  bool cancel = false;  // cancelled while live: dropped at the end of the chunk
  double t_submit = 0, t_admit = 0, t_done = 0;
  size_t off = 0, bytes = 0;
  long pitch = 0, plane = 0;
  int cur = 0;
  int left = 0;
};

}  // namespace

struct Server {
  size_t budget;
  int max_live;
  int plain;
  int max_chunk = 16;
  // queue limit (0: none; submit returns -LE2D_ERR_BUSY at the limit), the
  // queued requests admission looks at, and how often one that does not
  // fit may be passed by later ones before admission waits for it
  int max_queue = 0;
  int window = 64, max_pass = 16;
  int fb = 256, seg = 16;
  int pair_seg = 64;  // two-step launches: largest segment height; 0: one step per launch
  // lanes per two-step workgroup (64, 128 or 256); 0: 256 for FP32 requests,
  // which are issue-bound, 64 (one wavefront, no barriers between waves)
  // for FP64, which are latency-bound
  int pair_fb = 0;
  // 64-lane two-step windows with the two stages on two wavefronts
  int pair_split = 0;
  // The settings above are written by API calls (under mu); the engine
  // thread launches with this copy, taken under mu whenever it rebuilds the
  // tables, so a setting that changes mid-chunk never meets tables built
  // for another.
  struct Launch {
    int fb = 256, seg = 16, pair_seg = 64, pair_fb = 0, pair_split = 0;
  } act;
  int pair_lanes(int si) const { return act.pair_fb ? act.pair_fb : (si == 1 ? 64 : 256); }
  char *arena_dev = nullptr;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<long long> queue;
  std::unordered_map<long long, Req> reqs;
  std::vector<long long> live;
  long long next_id = 0;
  bool stopping = false;
  bool failed = false;
  char fail_msg[256] = "";
  hipStream_t stream[2] = {nullptr, nullptr};  // [0] FP32, [1] FP64
  void *d_inst[2] = {nullptr, nullptr};
  TileJob *d_tiles[2] = {nullptr, nullptr};
  TileJob *d_tiles2[2] = {nullptr, nullptr};  // two-step table
  size_t cap_inst[2] = {0, 0}, cap_tiles[2] = {0, 0}, cap_tiles2[2] = {0, 0};
  int ntiles[2] = {0, 0}, ntiles2[2] = {0, 0};
  // the requests of the current tables, by steps left (most first), and the
  // number of single-step / two-step tiles of the first k of them
  std::vector<Req *> order[2];
  std::vector<int> pre1[2], pre2[2];
  bool dirty = true;
  // statistics
  long long completed = 0, cancelled = 0, rejected = 0;
  double node_updates = 0, busy_s = 0;
  size_t peak_bytes = 0;
  int peak_live = 0;
  std::thread th;

  Server(size_t b, int ml, int p) : budget(b), max_live(ml), plain(p), arena(b) {}

  static size_t req_bytes(const le2d_request &q) {
    return 2 * ((grid_bytes(q.real_size, q.nx, q.ny) + 65535) / 65536 * 65536);
  }

  void fail(const char *what) {
    std::lock_guard<std::mutex> lk(mu);
    failed = true;
    snprintf(fail_msg, sizeof fail_msg, "%s: %s", what, last_error());
    cv_done.notify_all();
  }

  // Admit queued requests in arrival order while slots are free (caller
  // holds mu).  Among the first `window` queued requests, one that does not
  // fit the arena now is passed by later ones that do, until it has been
  // passed `max_pass` times; from then on admission waits for it, so no
  // request waits for ever behind a stream of smaller ones.
  void admit_locked() {
    int seen = 0;
    for (auto it = queue.begin(); it != queue.end() && (int)live.size() < max_live &&
                                  seen < window;
         ++seen) {
      Req &r = reqs[*it];
      size_t off = 0;
      if (!arena.alloc(r.bytes, &off)) {
        if (r.passed >= max_pass) break;
        ++it;
        continue;
      }
      for (auto jt = queue.begin(); jt != it; ++jt) ++reqs[*jt].passed;
      r.off = off;
      r.state = LIVE;
      r.t_admit = now_s();
      r.cur = 0;
      r.left = r.spec.steps;
      live.push_back(*it);
      it = queue.erase(it);
      dirty = true;
      init_fields(r);
    }
    peak_bytes = std::max(peak_bytes, arena.used());
    peak_live = std::max(peak_live, (int)live.size());
  }

  void init_fields(Req &r) {
    const le2d_request &q = r.spec;
    grid_layout(q.real_size, q.nx, q.ny, &r.pitch, &r.plane);
    const int si = q.real_size == 8 ? 1 : 0;
    char *base = arena_dev + r.off;
    (void)hipMemsetAsync(base, 0, r.bytes, stream[si]);
    const long n = (long)q.nx * q.ny;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (q.real_size == 8)
      hipLaunchKernelGGL(k_set_ball<double>, dim3(nb), dim3(256), 0, stream[si], (double *)base,
                         q.nx, q.ny, r.pitch, r.plane, q.hx, q.hy, q.cx, q.cy, q.r, q.s);
    else
      hipLaunchKernelGGL(k_set_ball<float>, dim3(nb), dim3(256), 0, stream[si], (float *)base,
                         q.nx, q.ny, r.pitch, r.plane, (float)q.hx, (float)q.hy, (float)q.cx,
                         (float)q.cy, (float)q.r, (float)q.s);
  }

  char *buf(const Req &r, int k) { return arena_dev + r.off + k * (r.bytes / 2); }

  // Device copy of a tile table, grown as needed (in stream order: the
  // stream is idle whenever the tables are rebuilt).
  int put_tiles(int si, const std::vector<TileJob> &tiles, TileJob **d, size_t *cap) {
    if (tiles.size() * sizeof(TileJob) > *cap) {
      dev_free(*d);
      *d = nullptr;
      *cap = 0;
      const size_t want = std::max<size_t>(tiles.size() * sizeof(TileJob) * 2, 65536);
      const int rc = dev_alloc((void **)d, want, "server: tiles");
      if (rc) return rc;
      *cap = want;
    }
    if (!tiles.empty() && hipMemcpyAsync(*d, tiles.data(), tiles.size() * sizeof(TileJob),
                                         hipMemcpyHostToDevice, stream[si]) != hipSuccess)
      return set_error(ERR_DEVICE, "server: tiles");
    return OK;
  }

  // Two-step segment height for the live set of one precision: the height,
  // at most pair_seg rows, that maximises (fill of the last round of
  // resident workgroups) x (share of rows that are output rather than halo)
  // over all live requests.  Many live requests keep tall segments; the few
  // of a light load or of the tail of a burst take short ones, as a small
  // grid does.
  int live_pair_seg(int si, const std::vector<Req *> &set) const {
    const int lanes = pair_lanes(si);
    long resident = pair_resident(si == 1 ? 8 : 4, lanes);
    if (resident <= 0) resident = 2048;
    int best = std::min(act.pair_seg, 2);
    double best_score = -1;
    for (int s : {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 80, 96, 128}) {
      if (s > act.pair_seg) break;
      long blocks = 0;
      for (const Req *r : set)
        blocks += (long)((r->spec.nx + lanes - 9) / (lanes - 8)) * ((r->spec.ny + s - 1) / s);
      if (blocks <= 0) break;
      const long rounds = (blocks + resident - 1) / resident;
      const double score = (double)blocks / (double)(rounds * resident) * s / (s + 8);
      if (score > best_score + 1e-9) {
        best_score = score;
        best = s;
      }
    }
    return best;
  }

  template <typename T>
  int build_tables(int si, const std::vector<Req *> &set) {
    std::vector<InstDesc<T>> desc(set.size());
    std::vector<TileJob> tiles, tiles2;
    const int seg2 = act.pair_seg > 0 ? live_pair_seg(si, set) : 0;
    pre1[si].assign(1, 0);
    pre2[si].assign(1, 0);
    for (size_t i = 0; i < set.size(); ++i) {
      const Req &r = *set[i];
      const le2d_request &q = r.spec;
      InstDesc<T> &d = desc[i];
      d.buf[0] = reinterpret_cast<T *>(buf(r, 0));
      d.buf[1] = reinterpret_cast<T *>(buf(r, 1));
      d.nx = q.nx;
      d.ny = q.ny;
      d.pitch = r.pitch;
      d.plane = r.plane;
      d.par = r.cur;
      coef_from_material<T>(q.c1, q.c2, q.rho, q.dt, q.hx, q.hy, &d.cx, &d.cy);
      const int strips = (q.nx + act.fb - 5) / (act.fb - 4);
      const int sg = std::max(1, std::min(act.seg, q.ny));
      for (int j0 = 0; j0 < q.ny; j0 += sg)
        for (int s = 0; s < strips; ++s)
          tiles.push_back({(int)i, s, j0, std::min(j0 + sg, q.ny), j0 / sg});
      if (act.pair_seg > 0) {  // strips of lanes - 8 output columns (k_step_batch_pairs)
        const int lanes = pair_lanes(si);
        const int strips2 = (q.nx + lanes - 9) / (lanes - 8);
        const int sg2 = std::max(1, std::min(seg2, q.ny));
        for (int j0 = 0; j0 < q.ny; j0 += sg2)
          for (int s = 0; s < strips2; ++s)
            tiles2.push_back({(int)i, s, j0, std::min(j0 + sg2, q.ny), j0 / sg2});
      }
      pre1[si].push_back((int)tiles.size());
      pre2[si].push_back((int)tiles2.size());
    }
    if (desc.size() * sizeof(InstDesc<T>) > cap_inst[si]) {
      dev_free(d_inst[si]);
      d_inst[si] = nullptr;
      cap_inst[si] = 0;
      const size_t want = std::max<size_t>(desc.size() * sizeof(InstDesc<T>) * 2, 4096);
      const int rc = dev_alloc(&d_inst[si], want, "server: descriptors");
      if (rc) return rc;
      cap_inst[si] = want;
    }
    // tables are read by launches queued on this stream: copy in stream order
    if (!desc.empty() && hipMemcpyAsync(d_inst[si], desc.data(), desc.size() * sizeof(InstDesc<T>),
                                        hipMemcpyHostToDevice, stream[si]) != hipSuccess)
      return set_error(ERR_DEVICE, "server: tables");
    int rc = put_tiles(si, tiles, &d_tiles[si], &cap_tiles[si]);
    if (!rc) rc = put_tiles(si, tiles2, &d_tiles2[si], &cap_tiles2[si]);
    if (!rc && hipStreamSynchronize(stream[si]) != hipSuccess)
      rc = set_error(ERR_DEVICE, "server: tables");
    if (rc) return rc;
    ntiles[si] = (int)tiles.size();
    ntiles2[si] = (int)tiles2.size();
    return OK;
  }

  // One step over single-step tiles [first, last) of one precision.
  template <typename T>
  void launch_batch(int si, int step_parity, int first, int last) {
    if (last <= first) return;
    const InstDesc<T> *d = static_cast<const InstDesc<T> *>(d_inst[si]);
    const dim3 grid((unsigned)(last - first));
    const TileJob *tj = d_tiles[si] + first;
    if (act.fb == 64)
      hipLaunchKernelGGL((k_step_batch_tiles<T, 64>), grid, dim3(64), 0, stream[si], d, tj,
                         step_parity);
    else if (act.fb == 128)
      hipLaunchKernelGGL((k_step_batch_tiles<T, 128>), grid, dim3(128), 0, stream[si], d, tj,
                         step_parity);
    else
      hipLaunchKernelGGL((k_step_batch_tiles<T, 256>), grid, dim3(256), 0, stream[si], d, tj,
                         step_parity);
  }

  // Two steps over the first `count` two-step tiles of one precision.
  template <typename T>
  void launch_pairs(int si, int step_parity, int count) {
    if (count <= 0) return;
    const dim3 grid((unsigned)count);
    const InstDesc<T> *d = static_cast<const InstDesc<T> *>(d_inst[si]);
    const int lanes = pair_lanes(si);
    if (lanes == 64 && act.pair_split)
      hipLaunchKernelGGL((k_step_batch_pairs_split<T>), grid, dim3(128), 0, stream[si], d,
                         d_tiles2[si], step_parity);
    else if (lanes == 64)
      hipLaunchKernelGGL((k_step_batch_pairs<T, 64>), grid, dim3(64), 0, stream[si], d,
                         d_tiles2[si], step_parity);
    else if (lanes == 128)
      hipLaunchKernelGGL((k_step_batch_pairs<T, 128>), grid, dim3(128), 0, stream[si], d,
                         d_tiles2[si], step_parity);
    else
      hipLaunchKernelGGL((k_step_batch_pairs<T, 256>), grid, dim3(256), 0, stream[si], d,
                         d_tiles2[si], step_parity);
  }

  // Requests of the current tables of precision si with at least k steps
  // this chunk (the tables are ordered by steps left, most first).
  int count_at_least(int si, int k, int chunk) const {
    int n = 0;
    for (const Req *r : order[si])
      if (std::min(r->left, chunk) >= k) ++n;
      else break;
    return n;
  }

  // The launches of one chunk.  Each live request takes min(left, chunk)
  // steps: two-step launch p covers the requests with at least 2p+2 steps
  // this chunk (a prefix of the tables), and a single-step launch beside it
  // the requests with exactly 2p+1 (the next run of them).  Every request
  // in launch p has taken exactly p launches before it in this chunk, so
  // its buffer parity is its table parity xor (p & 1).  Without two-step
  // launches, launch s covers the requests with at least s+1 steps.
  template <typename T>
  void launch_chunk(int si, int chunk) {
    if (order[si].empty()) return;
    if (act.pair_seg > 0) {
      for (int p = 0; 2 * p < chunk; ++p) {
        const int n2 = 2 * p + 2 <= chunk ? count_at_least(si, 2 * p + 2, chunk) : 0;
        const int n1 = count_at_least(si, 2 * p + 1, chunk);
        launch_pairs<T>(si, p & 1, pre2[si][n2]);
        launch_batch<T>(si, p & 1, pre1[si][n2], pre1[si][n1]);
      }
    } else {
      for (int s = 0; s < chunk; ++s)
        launch_batch<T>(si, s & 1, 0, pre1[si][count_at_least(si, s + 1, chunk)]);
    }
  }

  // The serial baseline's launch on the one live request: the single-grid
  // kernels as le_step_n uses them, two steps (two = true) or one.
  int plain_step(Req &r, bool two) {
    const le2d_request &q = r.spec;
    DevGrid g;
    g.elem = q.real_size;
    g.nx = q.nx;
    g.ny = q.ny;
    g.pitch = r.pitch;
    g.plane = r.plane;
    g.buf[0] = buf(r, 0);
    g.buf[1] = buf(r, 1);
    g.cur = r.cur;
    g.stream = stream[q.real_size == 8 ? 1 : 0];
    g.tiles = default_tiles(g.elem, g.nx, g.ny);
    if (q.real_size == 8) {
      SweepCoef<double> cx, cy;
      coef_from_material<double>(q.c1, q.c2, q.rho, q.dt, q.hx, q.hy, &cx, &cy);
      return two ? step2_xy<double>(&g, cx, cy) : step_xy<double>(&g, cx, cy, EDGE_CLAMP);
    }
    SweepCoef<float> cx, cy;
    coef_from_material<float>(q.c1, q.c2, q.rho, q.dt, q.hx, q.hy, &cx, &cy);
    return two ? step2_xy<float>(&g, cx, cy) : step_xy<float>(&g, cx, cy, EDGE_CLAMP);
  }

  void run() {
    if (hipSetDevice(0) != hipSuccess) {
      fail("server: device");
      return;
    }
    for (;;) {
      std::vector<Req *> set[2];
      int chunk = 0;
      bool rebuild = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        // Admission runs on every wake: a submit (or a released result) that
        // arrives while nothing is live only queues, and would otherwise sit
        // there until the next unrelated wake.
        for (;;) {
          if (stopping) return;
          admit_locked();
          if (!live.empty()) break;
          cv_work.wait(lk);
        }
        // requests that finish inside a chunk stop there (launch_chunk), so
        // a chunk runs max_chunk steps (even, with two-step launches) unless
        // every live request has fewer left
        int most = 0;
        for (long long id : live) {
          Req &r = reqs[id];
          most = std::max(most, r.left);
          set[r.spec.real_size == 8 ? 1 : 0].push_back(&r);
        }
        if (dirty) {
          act = {fb, seg, pair_seg, pair_fb, pair_split};
          rebuild = true;
          dirty = false;
        }
        chunk = std::min(max_chunk, most);
        if (!plain && act.pair_seg > 0 && chunk >= 2) chunk &= ~1;
      }
      const double t0 = now_s();
      int rc = OK;
      if (plain) {
        Req &r = set[0].empty() ? *set[1][0] : *set[0][0];
        for (int s = 0; s < chunk && !rc;) {
          const bool two = s + 1 < chunk;
          rc = plain_step(r, two);  // reads buf[r.cur], writes the other
          r.cur ^= 1;
          s += two ? 2 : 1;
        }
      } else {
        if (rebuild) {
          for (int si = 0; si < 2; ++si) {
            std::stable_sort(set[si].begin(), set[si].end(),
                             [](const Req *a, const Req *b) { return a->left > b->left; });
            order[si] = set[si];
          }
          if (!set[0].empty()) rc = build_tables<float>(0, set[0]);
          if (!rc && !set[1].empty()) rc = build_tables<double>(1, set[1]);
        }
        if (!rc) {
          launch_chunk<float>(0, chunk);
          launch_chunk<double>(1, chunk);
        }
        if (!rc && hipGetLastError() != hipSuccess) rc = set_error(ERR_DEVICE, "server: launch");
      }
      for (int si = 0; si < 2 && !rc; ++si)
        if (hipStreamSynchronize(stream[si]) != hipSuccess) rc = set_error(ERR_DEVICE, "server: sync");
      if (rc) {
        fail("server step");
        return;
      }
      const double t1 = now_s();
      std::lock_guard<std::mutex> lk(mu);
      busy_s += t1 - t0;
      std::vector<long long> still;
      bool stale = false;  // a table parity no longer names the current buffer
      for (long long id : live) {
        Req &r = reqs[id];
        const int taken = std::min(r.left, chunk);
        if (!plain) {
          const int n = act.pair_seg > 0 ? taken / 2 + (taken & 1) : taken;  // launches taken
          r.cur ^= n & 1;
          stale |= (n & 1) && r.left > taken;
        }
        r.left -= taken;
        node_updates += (double)r.spec.nx * r.spec.ny * taken;
        if (r.cancel) {  // cancelled while live: its memory returns now
          r.state = CANCELLED;
          r.t_done = t1;
          ++cancelled;
          dirty = true;
          arena.release(r.off, r.bytes);
        } else if (r.left == 0) {
          r.state = DONE;
          r.t_done = t1;
          ++completed;
          dirty = true;
          if (!r.spec.keep_result) {
            arena.release(r.off, r.bytes);
            r.state = RELEASED;
          }
        } else {
          still.push_back(id);
        }
      }
      if (still.size() != live.size()) {
        live.swap(still);
        dirty = true;
        cv_done.notify_all();
      }
      // tables hold each request's parity at build time: rebuild after an
      // odd number of launches for any request that stays
      if (stale) dirty = true;
    }
  }
};

}  // namespace le2d

using le2d::Server;

struct le2d_server {
  Server *s;
};

extern "C" {

le2d_server *le2d_server_create(size_t hbm_budget, int max_live, int plain) {
  le2d::require_device();
  if (hbm_budget < (1u << 20) || max_live < 1) {
    le2d::set_error(le2d::ERR_ARG, "server: budget %zu, slots %d", hbm_budget, max_live);
    return nullptr;
  }
  Server *s = new Server(hbm_budget, plain ? 1 : max_live, plain);
  if (le2d::dev_alloc((void **)&s->arena_dev, hbm_budget, "server: arena")) {
    delete s;
    return nullptr;
  }
  for (int k = 0; k < 2; ++k)
    if (hipStreamCreateWithFlags(&s->stream[k], hipStreamNonBlocking) != hipSuccess) {
      le2d::set_error(le2d::ERR_DEVICE, "server: stream");
      if (k == 1) (void)hipStreamDestroy(s->stream[0]);
      le2d::dev_free(s->arena_dev);
      delete s;
      return nullptr;
    }
  s->th = std::thread([s] { s->run(); });
  le2d_server *h = new le2d_server;
  h->s = s;
  return h;
}

int le2d_server_set_pairs(le2d_server *h, int pair_seg) {
  if (!h || pair_seg < 0) return le2d::set_error(le2d::ERR_ARG, "server: pairs");
  std::lock_guard<std::mutex> lk(h->s->mu);
  h->s->pair_seg = pair_seg;
  h->s->dirty = true;
  return le2d::OK;
}

int le2d_server_set_pair_window(le2d_server *h, int lanes) {
  if (!h || (lanes != 0 && lanes != 64 && lanes != 128 && lanes != 256))
    return le2d::set_error(le2d::ERR_ARG, "server: pair window %d", lanes);
  std::lock_guard<std::mutex> lk(h->s->mu);
  h->s->pair_fb = lanes;
  h->s->dirty = true;
  return le2d::OK;
}

int le2d_server_set_pair_split(le2d_server *h, int on) {
  if (!h || (on != 0 && on != 1)) return le2d::set_error(le2d::ERR_ARG, "server: pair split %d", on);
  std::lock_guard<std::mutex> lk(h->s->mu);
  h->s->pair_split = on;
  h->s->dirty = true;
  return le2d::OK;
}

int le2d_server_set_admission(le2d_server *h, int max_queue, int window, int max_pass) {
  if (!h || max_queue < 0 || window < 1 || max_pass < 0)
    return le2d::set_error(le2d::ERR_ARG, "server: admission %d %d %d", max_queue, window,
                           max_pass);
  std::lock_guard<std::mutex> lk(h->s->mu);
  h->s->max_queue = max_queue;
  h->s->window = window;
  h->s->max_pass = max_pass;
  h->s->cv_work.notify_one();
  return le2d::OK;
}

int le2d_server_set_shape(le2d_server *h, int fb, int seg, int max_chunk) {
  if (!h || (fb != 64 && fb != 128 && fb != 256) || seg < 1 || max_chunk < 1)
    return le2d::set_error(le2d::ERR_ARG, "server: shape");
  std::lock_guard<std::mutex> lk(h->s->mu);
  h->s->fb = fb;
  h->s->seg = seg;
  h->s->max_chunk = max_chunk;
  h->s->dirty = true;
  return le2d::OK;
}

long long le2d_server_submit(le2d_server *h, const le2d_request *q) {
  if (!h || !q) return -le2d::ERR_ARG;
  if ((q->real_size != 4 && q->real_size != 8) || q->nx < 3 || q->ny < 3 ||
      !le2d::row_fits(q->real_size, q->nx) || q->steps < 1 ||
      !(q->dt > 0) || !(q->hx > 0) || !(q->hy > 0) || !(q->c1 > 0) || !(q->c2 > 0) ||
      !(q->rho > 0))
    return -le2d::set_error(le2d::ERR_ARG, "server: bad request");
  Server *s = h->s;
  const size_t bytes = Server::req_bytes(*q);
  if (bytes > s->budget)
    return -le2d::set_error(le2d::ERR_BUDGET, "server: request needs %zu bytes, budget %zu", bytes,
                            s->budget);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->failed) return -le2d::set_error(le2d::ERR_DEVICE, "%s", s->fail_msg);
  if (s->max_queue > 0 && (int)s->queue.size() >= s->max_queue) {
    ++s->rejected;
    return -le2d::set_error(le2d::ERR_BUSY, "server: %d requests queued (the limit)",
                            (int)s->queue.size());
  }
  const long long id = s->next_id++;
  le2d::Req &r = s->reqs[id];
  r.spec = *q;
  r.bytes = bytes;
  r.t_submit = le2d::now_s();
  s->queue.push_back(id);
  s->cv_work.notify_one();
  return id;
}

int le2d_server_wait(le2d_server *h, long long id, double timeout_s) {
  if (!h) return le2d::ERR_ARG;
  Server *s = h->s;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return le2d::set_error(le2d::ERR_ARG, "server: no request %lld", id);
  // submits insert while this waits: a rehash moves iterators, not elements
  const le2d::Req *r = &it->second;
  const auto done = [&] { return s->failed || r->state >= le2d::DONE; };
  if (timeout_s < 0)
    s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
    return le2d::set_error(le2d::ERR_STATE, "server: request %lld not done", id);
  if (s->failed) return le2d::set_error(le2d::ERR_DEVICE, "%s", s->fail_msg);
  if (r->state == le2d::CANCELLED)
    return le2d::set_error(le2d::ERR_STATE, "server: request %lld was cancelled", id);
  return le2d::OK;
}

int le2d_server_drain(le2d_server *h, double timeout_s) {
  if (!h) return le2d::ERR_ARG;
  Server *s = h->s;
  std::unique_lock<std::mutex> lk(s->mu);
  const auto idle = [&] { return s->failed || (s->queue.empty() && s->live.empty()); };
  if (timeout_s < 0)
    s->cv_done.wait(lk, idle);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle))
    return le2d::set_error(le2d::ERR_STATE, "server: not drained");
  return s->failed ? le2d::set_error(le2d::ERR_DEVICE, "%s", s->fail_msg) : le2d::OK;
}

int le2d_server_result(le2d_server *h, long long id, void *aos) {
  if (!h) return le2d::ERR_ARG;
  Server *s = h->s;
  le2d::Req r;
  {
    // claim the result: a second call (or a cancel) finds it RELEASING and
    // fails, so the region is downloaded and released exactly once
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.state != le2d::DONE || !it->second.spec.keep_result)
      return le2d::set_error(le2d::ERR_STATE, "server: no kept result for %lld", id);
    it->second.state = le2d::RELEASING;
    r = it->second;
  }
  le2d::DevGrid g;
  g.elem = r.spec.real_size;
  g.nx = r.spec.nx;
  g.ny = r.spec.ny;
  g.pitch = r.pitch;
  g.plane = r.plane;
  g.buf[0] = s->buf(r, 0);
  g.buf[1] = s->buf(r, 1);
  g.cur = r.cur;
  hipStream_t st;
  int rc = le2d::OK;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    rc = le2d::set_error(le2d::ERR_DEVICE, "server: result stream");
  } else {
    g.stream = st;
    rc = le2d::grid_download(&g, aos, 0);
    (void)hipStreamDestroy(st);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  le2d::Req &rr = s->reqs[id];
  if (rc) {  // the result stays kept; the call may be repeated
    rr.state = le2d::DONE;
    return rc;
  }
  s->arena.release(rr.off, rr.bytes);
  rr.state = le2d::RELEASED;
  s->cv_work.notify_one();
  return le2d::OK;
}

int le2d_server_cancel(le2d_server *h, long long id) {
  if (!h) return le2d::ERR_ARG;
  Server *s = h->s;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return le2d::set_error(le2d::ERR_ARG, "server: no request %lld", id);
  le2d::Req &r = it->second;
  switch (r.state) {
    case le2d::QUEUED: {
      auto q = std::find(s->queue.begin(), s->queue.end(), id);
      if (q != s->queue.end()) s->queue.erase(q);
      r.state = le2d::CANCELLED;
      r.t_done = le2d::now_s();
      ++s->cancelled;
      s->cv_done.notify_all();
      s->cv_work.notify_one();
      return le2d::OK;
    }
    case le2d::LIVE:
      r.cancel = true;  // the engine drops it and frees its memory after the chunk
      return le2d::OK;
    case le2d::DONE:
      if (r.spec.keep_result) {  // discard the kept result
        s->arena.release(r.off, r.bytes);
        r.state = le2d::CANCELLED;
        ++s->cancelled;
        s->cv_work.notify_one();
        return le2d::OK;
      }
      break;
    default:
      break;
  }
  return le2d::set_error(le2d::ERR_STATE, "server: request %lld cannot be cancelled (state %d)",
                         id, r.state);
}

int le2d_server_times(le2d_server *h, long long id, double *t_submit, double *t_admit,
                      double *t_done) {
  if (!h) return le2d::ERR_ARG;
  std::lock_guard<std::mutex> lk(h->s->mu);
  auto it = h->s->reqs.find(id);
  if (it == h->s->reqs.end()) return le2d::set_error(le2d::ERR_ARG, "server: no request %lld", id);
  *t_submit = it->second.t_submit;
  *t_admit = it->second.t_admit;
  *t_done = it->second.t_done;
  return it->second.state >= le2d::DONE && it->second.state != le2d::CANCELLED ? le2d::OK
                                                                               : le2d::ERR_STATE;
}

int le2d_server_get_stats(le2d_server *h, le2d_server_stats *out) {
  if (!h || !out) return le2d::ERR_ARG;
  Server *s = h->s;
  std::lock_guard<std::mutex> lk(s->mu);
  out->completed = s->completed;
  out->queued = (long long)s->queue.size();
  out->live = (int)s->live.size();
  out->peak_live = s->peak_live;
  out->node_updates = s->node_updates;
  out->busy_seconds = s->busy_s;
  out->bytes_in_use = s->arena.used();
  out->peak_bytes = s->peak_bytes;
  out->cancelled = s->cancelled;
  out->rejected = s->rejected;
  return le2d::OK;
}

void le2d_server_destroy(le2d_server *h) {
  if (!h) return;
  Server *s = h->s;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    s->queue.clear();
    s->cv_work.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  for (int k = 0; k < 2; ++k) {
    if (s->stream[k]) (void)hipStreamSynchronize(s->stream[k]);
    le2d::dev_free(s->d_inst[k]);
    le2d::dev_free(s->d_tiles[k]);
    le2d::dev_free(s->d_tiles2[k]);
    if (s->stream[k]) (void)hipStreamDestroy(s->stream[k]);
  }
  le2d::dev_free(s->arena_dev);
  delete s;
  delete h;
}

}  // extern "C"
