// Served clustering: a request loop with continuous batching over the batched
// entry point.  No counterpart in the reference, which clusters one event per
// synchronous Clusterer::make_clusters call (core/detail/Clusterer.hpp:134-217).
//
// Device memory is fixed at creation and stays resident: `slots` batch slots,
// each with a stream, a plan (the grow-only scratch) reserved for its batch
// capacity, device input/output buffers and pinned host staging.  The capacity
// per slot is the configured one, lowered until all slots fit the device
// budget; a request larger than one slot's capacity is refused at submission.
//
// A scheduler thread forms a batch whenever a slot is free: the oldest pending
// request fixes the batch's class (dimension, parameters, kernel, metric,
// wrapping) and every pending request of that class is taken in arrival order
// while the batch holds at most half the slot's capacity (a single request may
// fill the slot).  With small_batch_points set, small requests (at most that
// many points) form batches of their own of at most that many points (more when
// many of them wait), so that a small request never waits for a large one of its
// batch, and one slot is kept for them while the others hold large batches.  Requests that arrive while slots are busy wait in
// the queue and join the next batch (continuous batching); the slots' streams
// overlap one batch's copies with another's kernels.  A waiter thread per slot
// blocks on the slot's completion event, scatters the outputs to the callers'
// buffers and releases the slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "clue_abi.h"

namespace clue_hip {
void set_error(const std::string& msg);  // the engine's clue_last_error_message
// scratch of a plan for batched calls of up to npoints points in max_events events
clue_status reserve_batches(clue_plan_t plan, int32_t ndim, int32_t npoints, int32_t max_events);
}

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  uint64_t ticket = 0;
  clue_request req{};
  int64_t t_submit = 0, t_launch = 0, t_done = 0;
  int32_t batch_events = 0;
  clue_status status = CLUE_OK;
  std::string error;
  bool done = false;
};

bool same_class(const clue_request& a, const clue_request& b) {
  if (a.ndim != b.ndim) return false;
  if (std::memcmp(&a.params, &b.params, sizeof(a.params)) != 0) return false;
  if (a.kernel.type != b.kernel.type || std::memcmp(a.kernel.params, b.kernel.params, sizeof(a.kernel.params)))
    return false;
  if (a.metric.type != b.metric.type || std::memcmp(a.metric.params, b.metric.params, sizeof(a.metric.params)))
    return false;
  return std::memcmp(a.wrapped, b.wrapped, sizeof(a.wrapped)) == 0;
}

struct Slot {
  hipStream_t stream = nullptr;
  clue_plan_t plan = nullptr;
  float* d_in = nullptr;     // (max_dim + 1) columns of `cap` floats: coordinates, then weights
  int32_t* d_out = nullptr;  // 4 columns of `cap`: cluster, is_seed, nearest_higher, rho
  float* h_in = nullptr;     // pinned, same layouts
  int32_t* h_out = nullptr;
  hipEvent_t done_ev = nullptr;
  std::vector<std::shared_ptr<Request>> batch;  // non-empty while the slot is busy
  std::vector<int64_t> offsets;
  bool busy = false;
  bool launched = false;
  bool large = false;  // the batch holds more than small_batch_points points
  std::condition_variable cv;
  std::thread waiter;
};

}  // namespace

struct clue_server_s {
  int32_t device = 0;
  clue_server_config cfg{};
  int64_t cap = 0;  // points per batch
  size_t device_bytes = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_sched;
  std::condition_variable cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread scheduler;
  clue_server_stats stats{};
};

namespace {

clue_status serr(clue_status st, const std::string& msg) {
  clue_hip::set_error(msg);
  return st;
}

void finish(clue_server_s* s, Slot& slot, clue_status st, const std::string& msg) {
  const int64_t t = now_ns();
  // outputs: scatter the batch columns to every request's host buffers
  if (st == CLUE_OK) {
    const int64_t cap = s->cap;
    for (size_t k = 0; k < slot.batch.size(); ++k) {
      const clue_request& r = slot.batch[k]->req;
      const int64_t o = slot.offsets[k];
      const size_t bytes = static_cast<size_t>(r.n) * 4;
      std::memcpy(r.cluster_out, slot.h_out + o, bytes);
      if (r.is_seed_out) std::memcpy(r.is_seed_out, slot.h_out + cap + o, bytes);
      if (r.nearest_higher_out) std::memcpy(r.nearest_higher_out, slot.h_out + 2 * cap + o, bytes);
      if (r.rho_out) std::memcpy(r.rho_out, slot.h_out + 3 * cap + o, bytes);
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto& r : slot.batch) {
    r->t_done = t;
    r->status = st;
    r->error = msg;
    r->done = true;
    s->stats.events_done += 1;
    s->stats.points_done += r->req.n;
  }
  slot.batch.clear();
  slot.busy = false;
  slot.launched = false;
  s->cv_done.notify_all();
  s->cv_sched.notify_all();
}

// Host staging, copies in, the batched call, copies out, completion event.
clue_status launch(clue_server_s* s, Slot& slot) {
  const auto& batch = slot.batch;
  const clue_request& r0 = batch.front()->req;
  const int32_t D = r0.ndim;
  const int64_t cap = s->cap;
  const int32_t maxd = s->cfg.max_dim;
  slot.offsets.assign(batch.size() + 1, 0);
  for (size_t k = 0; k < batch.size(); ++k) slot.offsets[k + 1] = slot.offsets[k] + batch[k]->req.n;
  const int64_t n = slot.offsets.back();
  for (size_t k = 0; k < batch.size(); ++k) {
    const clue_request& r = batch[k]->req;
    const int64_t o = slot.offsets[k];
    for (int d = 0; d < D; ++d)
      std::memcpy(slot.h_in + d * cap + o, r.coords + static_cast<int64_t>(d) * r.n, static_cast<size_t>(r.n) * 4);
    std::memcpy(slot.h_in + maxd * cap + o, r.weights, static_cast<size_t>(r.n) * 4);
  }
  for (int d = 0; d < D; ++d)
    if (hipMemcpyAsync(slot.d_in + d * cap, slot.h_in + d * cap, n * 4, hipMemcpyHostToDevice, slot.stream) !=
        hipSuccess)
      return serr(CLUE_ERR_RUNTIME, "copy to device failed");
  if (hipMemcpyAsync(slot.d_in + maxd * cap, slot.h_in + maxd * cap, n * 4, hipMemcpyHostToDevice, slot.stream) !=
      hipSuccess)
    return serr(CLUE_ERR_RUNTIME, "copy to device failed");
  const float* cp[CLUE_MAX_DIM];
  for (int d = 0; d < D; ++d) cp[d] = slot.d_in + d * cap;
  bool any_seed = false, any_nh = false, any_rho = false;
  for (auto& r : batch) {
    any_seed = any_seed || r->req.is_seed_out;
    any_nh = any_nh || r->req.nearest_higher_out;
    any_rho = any_rho || r->req.rho_out;
  }
  clue_status st = clue_make_clusters_batched(
      slot.plan, slot.stream, &r0.params, D, static_cast<int32_t>(batch.size()), slot.offsets.data(), cp,
      slot.d_in + maxd * cap, &r0.kernel, &r0.metric, r0.wrapped, slot.d_out, any_seed ? slot.d_out + cap : nullptr,
      any_rho ? reinterpret_cast<float*>(slot.d_out + 3 * cap) : nullptr, any_nh ? slot.d_out + 2 * cap : nullptr);
  if (st != CLUE_OK) return st;
  for (int c = 0; c < 4; ++c) {
    if ((c == 1 && !any_seed) || (c == 2 && !any_nh) || (c == 3 && !any_rho)) continue;
    if (hipMemcpyAsync(slot.h_out + c * cap, slot.d_out + c * cap, n * 4, hipMemcpyDeviceToHost, slot.stream) !=
        hipSuccess)
      return serr(CLUE_ERR_RUNTIME, "copy to host failed");
  }
  if (hipEventRecord(slot.done_ev, slot.stream) != hipSuccess) return serr(CLUE_ERR_RUNTIME, "event record failed");
  return CLUE_OK;
}

void scheduler_loop(clue_server_s* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    Slot* slot = nullptr;
    bool small_only = false;
    const int64_t small = s->cfg.small_batch_points;
    const int large_slots = s->slots.size() > 1 ? static_cast<int>(s->slots.size()) - 1 : 1;
    s->cv_sched.wait(lk, [&] {
      slot = nullptr;
      small_only = false;
      int large_busy = 0;
      for (auto& sl : s->slots) {
        if (!sl->busy) {
          if (!slot) slot = sl.get();
        } else if (sl->large) {
          ++large_busy;
        }
      }
      if (s->pending.empty()) return s->stop;
      if (!slot) return false;
      if (small <= 0 || large_busy < large_slots) return true;
      // every other slot holds a large batch: this one is kept for small requests
      small_only = true;
      for (auto& r : s->pending)
        if (r->req.n <= small) return true;
      return false;
    });
    if (s->pending.empty()) break;  // stopping, nothing left to launch
    // batch: the oldest request's class (the oldest small request's when only
    // small ones may start), in arrival order, up to half a slot's capacity (a
    // single request may fill the slot) or small_batch_points for small
    // requests: under load, several batches then share the GPU, and one batch's
    // copies overlap another's kernels
    auto first_it = s->pending.begin();
    if (small_only)
      while ((*first_it)->req.n > small) ++first_it;
    const clue_request first = (*first_it)->req;
    const bool small_first = small > 0 && first.n <= small;
    int64_t target = small_first ? small : s->cap / 2;
    if (small_first && !small_only) {
      // many small requests waiting (saturation): one large batch of them costs less
      int64_t waiting = 0;
      for (auto& r : s->pending)
        if (r->req.n <= small && same_class(r->req, first)) waiting += r->req.n;
      if (waiting > small * static_cast<int64_t>(s->slots.size())) target = s->cap / 2;
    }
    int64_t pts = 0;
    for (auto it = s->pending.begin(); it != s->pending.end();) {
      if (static_cast<int32_t>(slot->batch.size()) >= s->cfg.max_batch_events) break;
      const int64_t n = (*it)->req.n;
      const bool size_ok = small <= 0 || ((n <= small) == small_first);
      if (same_class((*it)->req, first) && size_ok && pts + n <= s->cap &&
          (slot->batch.empty() || pts + n <= target)) {
        pts += (*it)->req.n;
        slot->batch.push_back(*it);
        it = s->pending.erase(it);
      } else {
        ++it;
      }
    }
    slot->busy = true;
    slot->large = small <= 0 || pts > small;
    s->stats.batches += 1;
    s->stats.max_batch_events = std::max<int64_t>(s->stats.max_batch_events, slot->batch.size());
    s->stats.max_batch_points = std::max<int64_t>(s->stats.max_batch_points, pts);
    lk.unlock();
    const int64_t t = now_ns();
    for (auto& r : slot->batch) {
      r->t_launch = t;
      r->batch_events = static_cast<int32_t>(slot->batch.size());
    }
    const clue_status st = launch(s, *slot);
    if (st != CLUE_OK) {
      const std::string msg = clue_last_error_message();
      (void)hipStreamSynchronize(slot->stream);
      finish(s, *slot, st, msg);
      lk.lock();
      continue;
    }
    lk.lock();
    slot->launched = true;
    slot->cv.notify_all();
  }
}

void waiter_loop(clue_server_s* s, Slot* slot) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    slot->cv.wait(lk, [&] { return slot->launched || (s->stop && !slot->busy && s->pending.empty()); });
    if (!slot->launched) break;
    lk.unlock();
    const hipError_t e = hipEventSynchronize(slot->done_ev);
    finish(s, *slot, e == hipSuccess ? CLUE_OK : CLUE_ERR_RUNTIME,
           e == hipSuccess ? std::string() : std::string("batch failed: ") + hipGetErrorString(e));
    lk.lock();
  }
}

size_t slot_bytes(clue_plan_t plan, int64_t cap, int32_t maxd) {
  size_t b = 0;
  (void)clue_plan_device_bytes(plan, &b);
  return b + static_cast<size_t>(cap) * 4 * (maxd + 1) + static_cast<size_t>(cap) * 16;
}

void destroy_slots(clue_server_s* s) {
  for (auto& sl : s->slots) {
    if (sl->plan) clue_plan_destroy(sl->plan);
    if (sl->d_in) (void)hipFree(sl->d_in);
    if (sl->d_out) (void)hipFree(sl->d_out);
    if (sl->h_in) (void)hipHostFree(sl->h_in);
    if (sl->h_out) (void)hipHostFree(sl->h_out);
    if (sl->done_ev) (void)hipEventDestroy(sl->done_ev);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
  }
  s->slots.clear();
}

}  // namespace

extern "C" {

clue_status clue_server_create(int32_t device, const clue_server_config* config, clue_server_t* server) {
  if (!server) return serr(CLUE_ERR_INVALID_ARGUMENT, "server is NULL");
  *server = nullptr;
  clue_server_config cfg{};
  if (config) cfg = *config;
  if (cfg.device_budget_bytes == 0) cfg.device_budget_bytes = size_t{32} << 30;
  if (cfg.slots <= 0) cfg.slots = 4;
  if (cfg.max_batch_points <= 0) cfg.max_batch_points = 8 << 20;
  if (cfg.max_batch_events <= 0) cfg.max_batch_events = 4096;
  if (cfg.max_dim <= 0) cfg.max_dim = 3;
  if (cfg.max_dim > CLUE_MAX_DIM) return serr(CLUE_ERR_INVALID_ARGUMENT, "max_dim above 10");
  if (hipSetDevice(device) != hipSuccess) return serr(CLUE_ERR_INVALID_ARGUMENT, "invalid device");
  auto s = std::make_unique<clue_server_s>();
  s->device = device;
  s->cfg = cfg;
  // The capacity per slot: the configured one, lowered until every slot's
  // plan scratch and buffers fit the budget.
  int64_t cap = cfg.max_batch_points;
  clue_plan_t probe = nullptr;
  if (clue_plan_create(device, &probe) != CLUE_OK) return serr(CLUE_ERR_RUNTIME, clue_last_error_message());
  while (true) {
    if (clue_hip::reserve_batches(probe, cfg.max_dim, static_cast<int32_t>(cap), cfg.max_batch_events) != CLUE_OK) {
      clue_plan_destroy(probe);
      return serr(CLUE_ERR_OUT_OF_MEMORY, clue_last_error_message());
    }
    const size_t per_slot = slot_bytes(probe, cap, cfg.max_dim);
    if (per_slot * static_cast<size_t>(cfg.slots) <= cfg.device_budget_bytes) break;
    clue_plan_destroy(probe);
    probe = nullptr;
    const int64_t fit = static_cast<int64_t>(static_cast<double>(cap) * 0.95 * static_cast<double>(cfg.device_budget_bytes) /
                                             static_cast<double>(per_slot * cfg.slots));
    if (fit < 1024) return serr(CLUE_ERR_OUT_OF_RANGE, "the device budget does not hold one slot of 1024 points");
    cap = std::min(cap - 1, fit);
    if (clue_plan_create(device, &probe) != CLUE_OK) return serr(CLUE_ERR_RUNTIME, clue_last_error_message());
  }
  clue_plan_destroy(probe);
  s->cap = cap;
  for (int k = 0; k < cfg.slots; ++k) {
    auto sl = std::make_unique<Slot>();
    bool ok = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&sl->done_ev, hipEventDisableTiming) == hipSuccess &&
              hipMalloc(&sl->d_in, static_cast<size_t>(cap) * 4 * (cfg.max_dim + 1)) == hipSuccess &&
              hipMalloc(&sl->d_out, static_cast<size_t>(cap) * 16) == hipSuccess &&
              hipHostMalloc(&sl->h_in, static_cast<size_t>(cap) * 4 * (cfg.max_dim + 1), hipHostMallocDefault) ==
                  hipSuccess &&
              hipHostMalloc(&sl->h_out, static_cast<size_t>(cap) * 16, hipHostMallocDefault) == hipSuccess;
    if (ok) {
      // touch the pinned staging once here, so that no request pays its first-touch faults
      std::memset(sl->h_in, 0, static_cast<size_t>(cap) * 4 * (cfg.max_dim + 1));
      std::memset(sl->h_out, 0, static_cast<size_t>(cap) * 16);
    }
    ok = ok && clue_plan_create(device, &sl->plan) == CLUE_OK &&
         clue_hip::reserve_batches(sl->plan, cfg.max_dim, static_cast<int32_t>(cap), cfg.max_batch_events) ==
             CLUE_OK;
    if (ok) s->device_bytes += slot_bytes(sl->plan, cap, cfg.max_dim);
    s->slots.push_back(std::move(sl));
    if (!ok) {
      destroy_slots(s.get());
      return serr(CLUE_ERR_OUT_OF_MEMORY, "could not allocate the server's slots");
    }
  }
  s->stats.slot_capacity_points = cap;
  s->stats.device_bytes = static_cast<int64_t>(s->device_bytes);
  clue_server_s* raw = s.get();
  for (auto& sl : s->slots) sl->waiter = std::thread(waiter_loop, raw, sl.get());
  s->scheduler = std::thread(scheduler_loop, raw);
  *server = s.release();
  return CLUE_OK;
}

clue_status clue_server_submit(clue_server_t s, const clue_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket) return serr(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (req->ndim < 1 || req->ndim > s->cfg.max_dim)
    return serr(CLUE_ERR_INVALID_ARGUMENT, "dimension outside the server's range");
  if (req->n <= 0 || !req->coords || !req->weights || !req->cluster_out)
    return serr(CLUE_ERR_INVALID_ARGUMENT, "a request needs points, weights and a cluster output");
  if (req->n > s->cap) return serr(CLUE_ERR_OUT_OF_RANGE, "request larger than a batch slot of the device budget");
  auto r = std::make_shared<Request>();
  r->req = *req;
  r->t_submit = now_ns();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return serr(CLUE_ERR_INVALID_ARGUMENT, "server is stopping");
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->live[r->ticket] = r;
  s->pending.push_back(r);
  s->stats.events_submitted += 1;
  s->cv_sched.notify_all();
  return CLUE_OK;
}

clue_status clue_server_wait(clue_server_t s, uint64_t ticket, clue_request_timing* timing) {
  if (!s) return serr(CLUE_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return serr(CLUE_ERR_INVALID_ARGUMENT, "unknown ticket");
  std::shared_ptr<Request> r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  s->live.erase(ticket);
  if (timing) {
    timing->submit_ns = r->t_submit;
    timing->launch_ns = r->t_launch;
    timing->done_ns = r->t_done;
    timing->batch_events = r->batch_events;
  }
  if (r->status != CLUE_OK) return serr(r->status, r->error);
  return CLUE_OK;
}

clue_status clue_server_get_stats(clue_server_t s, clue_server_stats* stats) {
  if (!s || !stats) return serr(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *stats = s->stats;
  return CLUE_OK;
}

clue_status clue_server_destroy(clue_server_t s) {
  if (!s) return CLUE_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_sched.notify_all();
  }
  if (s->scheduler.joinable()) s->scheduler.join();  // everything pending is launched
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& sl : s->slots) sl->cv.notify_all();
  }
  for (auto& sl : s->slots) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_done.wait(lk, [&] { return !sl->busy; });
      sl->cv.notify_all();
    }
    if (sl->waiter.joinable()) sl->waiter.join();
  }
  (void)hipSetDevice(s->device);
  destroy_slots(s);
  delete s;
  return CLUE_OK;
}

}  // extern "C"
