// Serving loop: a request queue in front of the solver with admission by
// HBM footprint and continuous batching.
//
// One scheduler thread owns every HIP call.  Device memory for requests comes
// from one resident arena allocated when the server starts (first-fit
// sub-allocation on the host), so steady-state serving never calls hipMalloc
// or hipFree.  Plans and twiddle tables are cached per grid shape and stay
// resident.  Pending requests are grouped by size class; a group is
// dispatched as soon as its buffers fit in the arena and a stream is free, so
// requests that arrive while others run are admitted as memory frees up and
// their kernels overlap with the groups already on the device.  Every
// request's solve uses only its own buffers and the same kernels and
// configurations whatever else is running, so its result is bit-identical
// alone and in a mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "ff_arena.hpp"
#include "ff_solver.hpp"

using namespace ff;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Req {
  uint64_t id = 0;
  ff_request r{};
  ff_field field{};
  bool has_field = false;
  double t_submit = 0, t_dispatch = 0, t_done = 0;
  size_t off_d = 0, off_p = 0, off_s = 0, off_n = 0;
  bool admitted = false;  // off_d / off_p (and off_s / off_n if has_s / has_n) hold arena blocks
  bool has_s = false, has_n = false;
  size_t bytes = 0, scratch = 0, normbytes = 0;
  int slot = -1;  // pinned host slot receiving the six norm sums
  ff_status status = FF_OK;
  double norms[3] = {0, 0, 0};
  bool done = false;
  int cls = 0;
};

struct Group {
  std::vector<Req*> reqs;
  int stream = 0;
  hipEvent_t ev = nullptr;
  double work = 0;          // estimated cost of the group's solves (see work_of)
  std::vector<int> tables;  // per-instance table slots of its grouped solves
  bool large = false;       // FF_SERVE_LARGE_ALONE: runs with the GPU to itself
};

// Estimated cost of a request's solve, in point-passes: five passes per
// solve, FP32 + refinement adds ten per step at about half the bytes.
double work_of(const ff_request& r) {
  const double N = double(r.n[0]) * r.n[1] * r.n[2];
  return r.precision == FF_FP32_REFINE ? N * (2.5 + 7.5 * r.refine_steps) : N * 5.0;
}

// FF_SERVE_LARGE_ALONE: the work above which a group runs with the GPU to
// itself, and below which a group may still run beside it (point-passes;
// 6e7 is FP64 up to about 230^3, FP32 + 2 steps up to about 150^3).
constexpr double kLargeWork = 3e8;
constexpr double kSmallWork = 6e7;

double group_work(const std::vector<Req*>& reqs, int max_batch) {
  double w = 0;
  for (size_t i = 0; i < reqs.size() && (int)i < max_batch; ++i) w += work_of(reqs[i]->r);
  return w;
}

}  // namespace

struct ff_server {
  ff_server_config cfg{};
  int device = 0;
  void* arena_mem = nullptr;
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<int> stream_load;
  std::vector<double> stream_work;  // estimated work queued on each stream
  int large_inflight = 0;           // FF_SERVE_LARGE_ALONE groups on the device
  std::vector<hipEvent_t> free_events;
  // plans by (extents, precision, refinement steps, lo, hi), with the tick of
  // their last use; at most kMaxPlans are kept (least recently used evicted)
  std::map<std::tuple<int, int, int, int, int, double, double, double, double, double, double>,
           std::pair<ff_plan*, uint64_t>>
      plans;
  uint64_t plan_tick = 0;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<Req*> incoming;
  std::vector<Req*> pending;
  std::vector<Group> inflight;
  std::map<uint64_t, Req*> all;
  // pinned host slots (6 doubles each) the norm sums of in-flight requests
  // are copied into; owned by the scheduler thread
  double* pinned = nullptr;
  std::vector<int> free_slots;
  // per-instance tables of grouped solves: pinned staging and device copies,
  // kTableEntries entries per slot, a slot held until its group completes;
  // owned by the scheduler thread
  ff::PassInst* table_host = nullptr;
  ff::PassInst* table_dev = nullptr;
  std::vector<int> free_tables;
  std::atomic<bool> stop{false};
  std::thread worker;
  uint64_t next_id = 1;
  // statistics (under mu); arena_peak: the scheduler thread's arena peak,
  // published under mu (the arena itself belongs to that thread)
  std::vector<double> latencies;
  size_t arena_peak = 0;
  double first_submit = 0, last_done = 0, points = 0;
  long long completed = 0, rejected = 0, dispatched_groups = 0, grouped_solves = 0;
  std::string fatal;
};

namespace {

// Pinned host slots for in-flight norm results, and the device work area of
// one request's two norm reductions.
constexpr int kNormSlots = 4096;
// Grouped solves: at most kGroupMax requests per grouped launch sequence, and
// table slots of kTableEntries per-instance entries (launches x requests:
// 85 launches for FP32 + 8 refinement steps, 16 requests).
constexpr int kGroupMax = 16;
constexpr int kTableEntries = 96 * kGroupMax;
// Only grids below 128^3 points are grouped: a single solve of a larger grid
// already fills the GPU, and grouping its launches measured 7 % slower at
// 128^3 (48^3 FP64 requests: 2.4x the throughput, 64^3 FP32 + 2: 2.8x).
constexpr long long kGroupPoints = 128LL * 128 * 128;
// Cached plans (host structures only).
constexpr size_t kMaxPlans = 64;
size_t norm_bytes() { return 2 * sizeof(double) * kDiffDoubles; }

int size_class(const ff_request& r) { return ff::size_class(r.precision, r.refine_steps, r.n); }

// Device work area of a request's solve beyond density and potential: the
// FP32 refinement work area, then the work array of grids whose z rows run
// in four steps (0: none).
size_t refine_bytes(const ff_request& r) {
  const size_t N = (size_t)r.n[0] * r.n[1] * r.n[2];
  const size_t Nh = (size_t)r.n[0] * r.n[1] * (r.n[2] / 2 + 1);
  return r.precision == FF_FP32_REFINE ? align_up(Nh * 16 + align_up(N * 4, 256) + 256, 256) : 0;
}
size_t work_bytes(const ff_request& r) {
  return refine_bytes(r) + (z_rows_four_step(r.n, r.precision) ? z_work_bytes(r.n) : 0);
}

ff_plan* plan_for(ff_server* s, const Req& q, ff_status& st) {
  // The domain of a field request is the field's own; a host-array request
  // is on the benchmark's grid, [-pi, pi) per axis (FishFry.hpp:108-139).
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    const double d = q.has_field ? q.field.d[a] : 2.0 * M_PI / q.r.n[a];
    lo[a] = q.has_field ? q.field.lo[a] : -M_PI;
    hi[a] = q.has_field ? lo[a] + d * (q.r.n[a] - 1) : M_PI - d;
  }
  const auto key = std::make_tuple(q.r.n[0], q.r.n[1], q.r.n[2], q.r.precision, q.r.refine_steps, lo[0], lo[1],
                                   lo[2], hi[0], hi[1], hi[2]);
  const uint64_t tick = ++s->plan_tick;
  auto it = s->plans.find(key);
  if (it != s->plans.end()) {
    it->second.second = tick;
    return it->second.first;
  }
  // Server plans own no device memory (their twiddle tables are the resident
  // per-device cache's and each request binds its refinement work area from
  // the arena), so evicting one whose copies still have kernels in flight is
  // safe.
  if (s->plans.size() >= kMaxPlans) {
    auto lru = s->plans.begin();
    for (auto e = s->plans.begin(); e != s->plans.end(); ++e)
      if (e->second.second < lru->second.second) lru = e;
    plan_destroy(lru->second.first);
    s->plans.erase(lru);
  }
  ff_options o{q.r.precision, q.r.refine_steps, 0, 0};
  ff_plan* p = nullptr;
  st = plan_create(q.r.n, lo, hi, &o, &p, /*own_scratch=*/false);
  if (st != FF_OK) return nullptr;
  s->plans[key] = {p, tick};
  return p;
}

// Frees the request's memory, then publishes its result (after which a
// waiter may delete it).  count: include it in the latency statistics.
void finish(ff_server* s, Req* q, ff_status st, bool count) {
  if (q->slot >= 0) {
    s->free_slots.push_back(q->slot);
    q->slot = -1;
  }
  if (q->admitted) {
    if (q->has_n) s->arena.release(q->off_n);
    if (q->has_s) s->arena.release(q->off_s);
    s->arena.release(q->off_d);
    s->arena.release(q->off_p);
    q->admitted = false;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  q->status = st;
  q->t_done = now_s();
  if (count) {
    s->latencies.push_back(q->t_done - q->t_submit);
    s->points += double(q->r.n[0]) * q->r.n[1] * q->r.n[2];
    s->last_done = std::max(s->last_done, q->t_done);
    s->completed++;
  }
  q->done = true;
}

// Requests solved by one plan configuration: same extents, precision and
// refinement steps (their domains may differ: the filter constants are
// per-instance fields).
bool same_solve(const ff_request& a, const ff_request& b) {
  return a.n[0] == b.n[0] && a.n[1] == b.n[1] && a.n[2] == b.n[2] && a.precision == b.precision &&
         a.refine_steps == b.refine_steps;
}
bool groupable(const ff_request& r) { return (long long)r.n[0] * r.n[1] * r.n[2] < kGroupPoints; }

struct Item {
  Req* q;
  ff_plan p;  // the cached plan bound to the request's work areas
  double* d;
  double* phi;
  ff_status st;
};

// Solves the requests mem of items as grouped launches on hs: each one's
// solve is recorded (nothing runs), and when all recordings agree except in
// their per-instance fields, every pass step runs once for all of them with a
// per-instance table (as the batch's grouped mode does).  False (nothing
// enqueued) when they do not agree, a launch is not recordable or no table
// slot is free; the caller then solves them one by one.
bool grouped_solve(ff_server* s, std::vector<Item>& items, const std::vector<size_t>& mem, hipStream_t hs, Group& g) {
  if (mem.size() < 2 || s->free_tables.empty()) return false;
  std::vector<std::vector<PassLaunch>> rec(mem.size());
  for (size_t j = 0; j < mem.size(); ++j) {
    Item& it = items[mem[j]];
    set_launch_recorder(&rec[j]);
    const ff_status st = solve_enqueue(&it.p, it.d, it.phi, hs, nullptr);
    set_launch_recorder(nullptr);
    if (st != FF_OK || rec[j].empty() || rec[j].size() != rec[0].size()) {
      clear_error();
      return false;
    }
  }
  const size_t nl = rec[0].size(), m = mem.size();
  if (nl * m > (size_t)kTableEntries) return false;
  for (size_t j = 1; j < m; ++j)
    for (size_t k = 0; k < nl; ++k)
      if (!launch_same(rec[0][k], rec[j][k])) return false;
  const int slot = s->free_tables.back();
  PassInst* host = s->table_host + (size_t)slot * kTableEntries;
  PassInst* dev = s->table_dev + (size_t)slot * kTableEntries;
  for (size_t k = 0; k < nl; ++k)
    for (size_t j = 0; j < m; ++j) host[k * m + j] = launch_inst_fields(rec[j][k]);
  ff_status st = hip_check(hipMemcpyAsync(dev, host, nl * m * sizeof(PassInst), hipMemcpyHostToDevice, hs),
                           "grouped solve table upload");
  if (st != FF_OK) return false;  // nothing launched yet: solve one by one
  s->free_tables.pop_back();
  g.tables.push_back(slot);  // the upload and the launches read it until the group completes
  for (size_t k = 0; k < nl && st == FF_OK; ++k) {
    PassLaunch r = rec[0][k];
    r.l.stream = hs;
    r.l.grid.y = (unsigned)m;
    if (r.kind == 0) r.col.inst = dev + k * m;
    else r.row.inst = dev + k * m;
    st = hip_check(replay_launch(r), "grouped solve launch");
  }
  for (size_t i : mem) items[i].st = st;
  if (st == FF_OK) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->grouped_solves += (long long)m;
  }
  return true;
}

// Admit and dispatch one size class group on stream k: inputs, then the
// solves (requests of one shape, precision and step count together as grouped
// launches), then outputs and norms.
void dispatch(ff_server* s, std::vector<Req*>& group, int k) {
  hipStream_t hs = s->streams[k];
  unsigned char* base = static_cast<unsigned char*>(s->arena_mem);
  Group g;
  g.stream = k;
  std::vector<Item> items;
  for (Req* q : group) {
    ff_status st = FF_OK;
    ff_plan* cached = plan_for(s, *q, st);
    if (!cached) {
      finish(s, q, st, false);
      continue;
    }
    Item it{q, *cached, nullptr, nullptr, FF_OK};  // shares twiddles; the work areas are the request's own
    if (q->has_s) {
      if (it.p.opt.precision == FF_FP32_REFINE) plan_bind_scratch(&it.p, base + q->off_s);
      if (it.p.fz.n) it.p.work = base + q->off_s + refine_bytes(q->r);  // four-step z rows
    }
    it.d = reinterpret_cast<double*>(base + q->off_d);
    it.phi = reinterpret_cast<double*>(base + q->off_p);
    if (q->has_field) {
      st = field_init(&q->field, it.d, hs);
    } else {
      st = hip_check(hipMemcpyAsync(it.d, q->r.f_host, sizeof(double) * q->r.n[0] * q->r.n[1] * q->r.n[2],
                                    hipMemcpyHostToDevice, hs),
                     "request input copy");
    }
    if (st != FF_OK) {
      // work already enqueued for this request may still use its blocks
      (void)hipStreamSynchronize(hs);
      finish(s, q, st, false);
      continue;
    }
    items.push_back(it);
  }
  std::vector<char> solved(items.size(), 0);
  for (size_t a = 0; a < items.size(); ++a) {
    if (solved[a]) continue;
    std::vector<size_t> mem{a};
    for (size_t b = a + 1; b < items.size() && mem.size() < (size_t)kGroupMax && groupable(items[a].q->r); ++b)
      if (!solved[b] && same_solve(items[a].q->r, items[b].q->r)) mem.push_back(b);
    if (!grouped_solve(s, items, mem, hs, g))
      for (size_t i : mem) items[i].st = solve_enqueue(&items[i].p, items[i].d, items[i].phi, hs, nullptr);
    for (size_t i : mem) solved[i] = 1;
  }
  for (Item& it : items) {
    Req* q = it.q;
    ff_status st = it.st;
    if (st == FF_OK && q->r.phi_host)
      st = hip_check(hipMemcpyAsync(q->r.phi_host, it.phi, sizeof(double) * q->r.n[0] * q->r.n[1] * q->r.n[2],
                                    hipMemcpyDeviceToHost, hs),
                     "request output copy");
    if (st == FF_OK && q->has_n) {
      // error norms against the exact potential, reduced on the request's own
      // stream into its arena block and copied to a pinned host slot, so
      // completion needs no synchronization
      double* w = reinterpret_cast<double*>(base + q->off_n);
      st = field_diff_enqueue(&q->field, it.phi, w, hs);
      if (st == FF_OK) st = field_diff_enqueue(&q->field, nullptr, w + kDiffDoubles, hs);
      if (st == FF_OK && !s->free_slots.empty()) {
        q->slot = s->free_slots.back();
        s->free_slots.pop_back();
        double* h = s->pinned + 6 * q->slot;
        st = hip_check(hipMemcpyAsync(h, w + kDiffResult, 3 * sizeof(double), hipMemcpyDeviceToHost, hs), "norms copy");
        if (st == FF_OK)
          st = hip_check(hipMemcpyAsync(h + 3, w + kDiffDoubles + kDiffResult, 3 * sizeof(double),
                                        hipMemcpyDeviceToHost, hs),
                         "norms copy");
      }
    }
    if (st != FF_OK) {
      // work already enqueued for this request may still use its blocks
      (void)hipStreamSynchronize(hs);
      finish(s, q, st, false);
      continue;
    }
    q->t_dispatch = now_s();
    g.reqs.push_back(q);
  }
  if (g.reqs.empty()) {
    if (!g.tables.empty()) (void)hipStreamSynchronize(hs);  // their launches may still read them
    for (int t : g.tables) s->free_tables.push_back(t);
    return;
  }
  if (!s->free_events.empty()) {
    g.ev = s->free_events.back();
    s->free_events.pop_back();
  } else if (hipEventCreateWithFlags(&g.ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamSynchronize(hs);
    for (Req* q : g.reqs) finish(s, q, FF_ERR_HIP, false);
    for (int t : g.tables) s->free_tables.push_back(t);
    return;
  }
  (void)hipEventRecord(g.ev, hs);
  s->stream_load[k]++;
  for (Req* q : g.reqs) g.work += work_of(q->r);
  g.large = (s->cfg.flags & FF_SERVE_LARGE_ALONE) && g.work > kLargeWork;
  if (g.large) s->large_inflight++;
  s->stream_work[k] += g.work;
  s->inflight.push_back(g);
  std::lock_guard<std::mutex> lk(s->mu);
  s->dispatched_groups++;
}

void complete(ff_server* s, Group& g) {
  for (Req* q : g.reqs) {
    ff_status st = FF_OK;
    if (q->has_n) {
      double e[3], z[3];
      if (q->slot >= 0) {  // copied by the stream before the group's event
        const double* h = s->pinned + 6 * q->slot;
        for (int k = 0; k < 3; ++k) {
          e[k] = h[k];
          z[k] = h[3 + k];
        }
      } else {  // no pinned slot was free: read the completed sums directly
        const double* w = reinterpret_cast<const double*>(static_cast<unsigned char*>(s->arena_mem) + q->off_n);
        st = hip_check(hipMemcpy(e, w + kDiffResult, sizeof e, hipMemcpyDeviceToHost), "norms read");
        if (st == FF_OK)
          st = hip_check(hipMemcpy(z, w + kDiffDoubles + kDiffResult, sizeof z, hipMemcpyDeviceToHost), "norms read");
      }
      if (st == FF_OK) {
        q->norms[0] = e[0] / z[0];
        q->norms[1] = std::sqrt(e[1] / z[1]);
        q->norms[2] = e[2] / z[2];
      }
    }
    finish(s, q, st, true);
  }
  for (int t : g.tables) s->free_tables.push_back(t);
  if (g.large) s->large_inflight--;
  s->stream_load[g.stream]--;
  s->stream_work[g.stream] = std::max(0.0, s->stream_work[g.stream] - g.work);
  s->free_events.push_back(g.ev);
  std::lock_guard<std::mutex> lk(s->mu);
  s->cv_done.notify_all();
}

void worker_main(ff_server* s) {
  (void)hipSetDevice(s->device);
  const double window = s->cfg.batch_window_us * 1e-6;
  while (true) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->incoming.empty() && s->pending.empty() && s->inflight.empty()) {
        if (s->stop) break;
        s->cv_submit.wait_for(lk, std::chrono::milliseconds(5));
      }
      while (!s->incoming.empty()) {
        s->pending.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
    }
    // completions
    for (size_t i = 0; i < s->inflight.size();) {
      if (hipEventQuery(s->inflight[i].ev) == hipSuccess) {
        Group g = s->inflight[i];
        s->inflight.erase(s->inflight.begin() + i);
        complete(s, g);
      } else {
        ++i;
      }
    }
    // admission and dispatch, oldest class first
    const size_t before = s->pending.size();
    const double t = now_s();
    std::map<int, std::vector<Req*>> classes;
    for (Req* q : s->pending) classes[q->cls].push_back(q);
    std::vector<std::pair<double, int>> order;
    for (auto& c : classes) order.push_back({c.second.front()->t_submit, c.first});
    std::sort(order.begin(), order.end());
    const bool alone = (s->cfg.flags & FF_SERVE_LARGE_ALONE) != 0;
    for (auto& oc : order) {
      std::vector<Req*>& reqs = classes[oc.second];
      const double w = alone ? group_work(reqs, s->cfg.max_batch) : 0.0;
      // beside a large group only small ones run (a larger group waits
      // for it to complete; newer small classes go ahead meanwhile)
      if (alone && s->large_inflight && w >= kSmallWork) continue;
      const bool full = (int)reqs.size() >= s->cfg.max_batch;
      const bool due = t - reqs.front()->t_submit >= window || s->inflight.empty();
      if (!full && !due) continue;
      // a large group waits for the device to drain, and holds back newer
      // classes meanwhile (so that it is not starved)
      if (alone && w > kLargeWork && !s->inflight.empty()) break;
      // the stream with the least work queued among those with fewer than two
      // groups, so a small request does not wait behind a large one
      int k = -1;
      for (int j = 0; j < (int)s->streams.size(); ++j)
        if (s->stream_load[j] < 2 && (k < 0 || s->stream_work[j] < s->stream_work[k])) k = j;
      if (k < 0) break;  // every stream has two groups queued
      std::vector<Req*> group;
      for (Req* q : reqs) {
        if ((int)group.size() >= s->cfg.max_batch) break;
        size_t od, op, os = 0;
        if (!s->arena.alloc(q->bytes, od)) continue;
        if (!s->arena.alloc(q->bytes, op)) {
          s->arena.release(od);
          continue;
        }
        if (q->scratch && !s->arena.alloc(q->scratch, os)) {
          s->arena.release(od);
          s->arena.release(op);
          continue;
        }
        size_t on = 0;
        if (q->normbytes && !s->arena.alloc(q->normbytes, on)) {
          if (q->scratch) s->arena.release(os);
          s->arena.release(od);
          s->arena.release(op);
          continue;
        }
        q->off_d = od;
        q->off_p = op;
        q->off_s = os;
        q->off_n = on;
        q->has_s = q->scratch != 0;
        q->has_n = q->normbytes != 0;
        q->admitted = true;
        group.push_back(q);
      }
      if (group.empty()) continue;  // waits for memory to free up
      for (Req* q : group) s->pending.erase(std::find(s->pending.begin(), s->pending.end(), q));
      dispatch(s, group, k);
    }
    // nothing could be admitted (memory or streams busy) or only in-flight
    // work remains: poll again shortly instead of spinning
    if (!s->inflight.empty() && (s->pending.empty() || s->pending.size() == before))
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->arena_peak = s->arena.peak();
      s->cv_done.notify_all();
    }
  }
}

}  // namespace

extern "C" {

ff_status ff_server_create(const ff_server_config* cfg, ff_server** out) {
  clear_error();
  if (!out) return set_error(FF_ERR_INVALID, "ff_server_create: null output");
  *out = nullptr;
  ff_server* s = new (std::nothrow) ff_server();
  if (!s) return set_error(FF_ERR_NOMEM, "host allocation failed");
  ff_server_config c{size_t(8) << 30, 4, 16, 200.0, 0};
  if (cfg) c = *cfg;
  if (c.streams < 1 || c.streams > 16 || c.max_batch < 1 || c.hbm_budget < (size_t(1) << 20) ||
      c.batch_window_us < 0 || (c.flags & ~FF_SERVE_LARGE_ALONE)) {
    delete s;
    return set_error(FF_ERR_INVALID,
                     "ff_server_create: streams 1..16, max_batch >= 1, budget >= 1 MiB, flags FF_SERVE_LARGE_ALONE or 0");
  }
  s->cfg = c;
  ff_status st = hip_check(hipGetDevice(&s->device), "hipGetDevice");
  if (st == FF_OK) st = hip_check(hipMalloc(&s->arena_mem, c.hbm_budget), "serving arena hipMalloc");
  for (int k = 0; k < c.streams && st == FF_OK; ++k) {
    hipStream_t hs = nullptr;
    st = hip_check(hipStreamCreateWithFlags(&hs, hipStreamNonBlocking), "serving stream");
    if (st == FF_OK) {
      s->streams.push_back(hs);
      s->stream_load.push_back(0);
      s->stream_work.push_back(0.0);
    }
  }
  if (st != FF_OK) {
    for (hipStream_t hs : s->streams) (void)hipStreamDestroy(hs);
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    delete s;
    return st;
  }
  s->arena.reset(c.hbm_budget);
  st = hip_check(hipHostMalloc((void**)&s->pinned, sizeof(double) * 6 * kNormSlots, hipHostMallocDefault),
                 "serving pinned slots");
  if (st != FF_OK) {
    for (hipStream_t hs : s->streams) (void)hipStreamDestroy(hs);
    (void)hipFree(s->arena_mem);
    delete s;
    return st;
  }
  for (int k = kNormSlots - 1; k >= 0; --k) s->free_slots.push_back(k);
  // table slots: two in-flight groups per stream, and two more
  const int ntables = 2 * c.streams + 2;
  const size_t tbytes = (size_t)ntables * kTableEntries * sizeof(PassInst);
  st = hip_check(hipHostMalloc((void**)&s->table_host, tbytes, hipHostMallocDefault), "serving table staging");
  if (st == FF_OK) st = hip_check(hipMalloc((void**)&s->table_dev, tbytes), "serving table hipMalloc");
  if (st != FF_OK) {
    for (hipStream_t hs : s->streams) (void)hipStreamDestroy(hs);
    (void)hipFree(s->arena_mem);
    (void)hipHostFree(s->pinned);
    if (s->table_host) (void)hipHostFree(s->table_host);
    delete s;
    return st;
  }
  for (int t = ntables - 1; t >= 0; --t) s->free_tables.push_back(t);
  s->worker = std::thread(worker_main, s);
  *out = s;
  return FF_OK;
}

size_t ff_server_footprint(const ff_request* r) {
  if (!r) return 0;
  for (int a = 0; a < 3; ++a)
    if (r->n[a] < 2) return 0;
  const int ones[3] = {1, 1, 1};
  size_t b = 2 * align_up(reference_bytes(r->n, ones), 256);
  b += align_up(work_bytes(*r), 256);
  if (r->want_norms && r->field) b += align_up(norm_bytes(), 256);
  return b;
}

ff_status ff_server_submit(ff_server* s, const ff_request* r, uint64_t* id) {
  clear_error();
  if (!s || !r || !id) return set_error(FF_ERR_INVALID, "ff_server_submit: null argument");
  if ((r->field == nullptr) == (r->f_host == nullptr))
    return set_error(FF_ERR_INVALID, "ff_server_submit: give exactly one of field and f_host");
  if (r->precision != FF_FP64 && r->precision != FF_FP32_REFINE)
    return set_error(FF_ERR_INVALID, "ff_server_submit: bad precision");
  if (r->refine_steps < 0 || r->refine_steps > 8) return set_error(FF_ERR_INVALID, "ff_server_submit: bad steps");
  for (int a = 0; a < 3; ++a)
    if (r->n[a] < 2) return set_error(FF_ERR_INVALID, "ff_server_submit: extents must be >= 2");
  if (r->field) {
    ff_status st = field_check(r->field);
    if (st != FF_OK) return st;
    for (int a = 0; a < 3; ++a)
      if (r->field->n[a] != r->n[a]) return set_error(FF_ERR_INVALID, "ff_server_submit: field extents differ");
  }
  const size_t need = ff_server_footprint(r);
  if (need > s->arena.size()) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->rejected++;
    return set_error(FF_ERR_REJECTED, "ff_server_submit: request larger than the serving HBM budget");
  }
  Req* q = new (std::nothrow) Req();
  if (!q) return set_error(FF_ERR_NOMEM, "host allocation failed");
  q->r = *r;
  if (r->field) {
    q->field = *r->field;
    q->has_field = true;
    q->r.field = &q->field;
  }
  const int ones[3] = {1, 1, 1};
  q->bytes = reference_bytes(r->n, ones);
  q->scratch = work_bytes(*r);
  if (r->want_norms && r->field) q->normbytes = norm_bytes();
  q->cls = size_class(*r);
  q->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    q->id = s->next_id++;
    if (!s->first_submit) s->first_submit = q->t_submit;
    s->all[q->id] = q;
    s->incoming.push_back(q);
  }
  s->cv_submit.notify_one();
  *id = q->id;
  return FF_OK;
}

ff_status ff_server_wait(ff_server* s, uint64_t id, double timeout_s, ff_result* out) {
  clear_error();
  if (!s || !out) return set_error(FF_ERR_INVALID, "ff_server_wait: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  // looked up again after every wake: another thread waiting on the same id
  // may have taken the result (and freed the request) meanwhile
  auto it = s->all.find(id);
  while (it != s->all.end() && !it->second->done) {
    const bool timed_out = s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout;
    it = s->all.find(id);
    if (timed_out && it != s->all.end() && !it->second->done)
      return set_error(FF_ERR_TIMEOUT, "ff_server_wait: timed out");
  }
  if (it == s->all.end()) return set_error(FF_ERR_NOT_FOUND, "ff_server_wait: unknown request id");
  Req* q = it->second;
  out->status = q->status;
  out->latency_s = q->t_done - q->t_submit;
  // a request that failed before dispatch spent all its time queued
  out->queue_s = (q->t_dispatch > 0 ? q->t_dispatch : q->t_done) - q->t_submit;
  out->points = (long long)q->r.n[0] * q->r.n[1] * q->r.n[2];
  for (int k = 0; k < 3; ++k) out->norms[k] = q->norms[k];
  s->all.erase(it);
  delete q;
  const ff_status rs = (ff_status)out->status;
  return rs == FF_OK ? FF_OK : set_error(rs, "request failed");
}

ff_status ff_server_stats(ff_server* s, ff_serving_stats* out) {
  clear_error();
  if (!s || !out) return set_error(FF_ERR_INVALID, "ff_server_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double qv) {
    if (l.empty()) return 0.0;
    const double pos = qv * double(l.size() - 1);
    const size_t a = (size_t)pos;
    const size_t b = std::min(a + 1, l.size() - 1);
    return l[a] + (pos - double(a)) * (l[b] - l[a]);
  };
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->groups = s->dispatched_groups;
  out->grouped_solves = s->grouped_solves;
  out->p50_s = pct(0.5);
  out->p99_s = pct(0.99);
  out->points_per_s = (s->last_done > s->first_submit) ? s->points / (s->last_done - s->first_submit) : 0.0;
  out->arena_bytes = s->arena.size();
  out->arena_peak_bytes = s->arena_peak;
  return FF_OK;
}

ff_status ff_server_destroy(ff_server* s) {
  if (!s) return FF_OK;
  s->stop = true;
  s->cv_submit.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (hipStream_t hs : s->streams) (void)hipStreamSynchronize(hs);
  for (auto& kv : s->plans) plan_destroy(kv.second.first);
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  for (hipStream_t hs : s->streams) (void)hipStreamDestroy(hs);
  if (s->arena_mem) (void)hipFree(s->arena_mem);
  if (s->pinned) (void)hipHostFree(s->pinned);
  if (s->table_host) (void)hipHostFree(s->table_host);
  if (s->table_dev) (void)hipFree(s->table_dev);
  for (auto& kv : s->all) delete kv.second;
  delete s;
  return FF_OK;
}

}  // extern "C"
