// Serving loop of libebi_hip: continuous batching of concurrent queries over
// HBM-resident columns.  No counterpart in lemolatoon/ebi, whose Decoder is
// single-threaded and synchronous (ebi/src/api/decoder.rs:202-439); every
// query runs the same plan as the synchronous entry points, so its result
// is the same alone and in the mix.
//
// Threads:
//  * the caller's threads submit queries (FIFO queue) and wait on tickets;
//  * the scheduler thread admits queued queries in arrival order while an
//    in-flight slot and the memory for the query's scratch are free (device
//    and pinned arenas carved once at create), issues every query admitted in
//    one pass as one group on the next stream (one batched decode launch for
//    the group), and polls the groups' completion events;
//  * kFinishers finisher threads (each with its own copy stream) copy
//    materialized records into the callers' buffers, fold per-chunk
//    partials, release the arenas and complete the tickets, so the
//    scheduler never blocks on a host copy and one large result does not
//    hold up the completions behind it.
// Streams keep the default priority.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/ebi_hip.h"
#include "column.hpp"
#include "internal.hpp"
#include "plan.hpp"

namespace ebi {
namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, cap) with coalescing frees.
class Arena {
 public:
  void reset(uint64_t cap, uint64_t align) {
    align_ = align;
    cap_ = cap;
    free_.clear();
    if (cap) free_[0] = cap;
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t len, uint64_t *off) {
    len = std::max<uint64_t>((len + align_ - 1) / align_ * align_, align_);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      *off = it->first;
      const uint64_t rest = it->second - len, start = it->first + len;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += len;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t len) {
    len = std::max<uint64_t>((len + align_ - 1) / align_ * align_, align_);
    used_ -= len;
    auto it = free_.emplace(off, len).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t cap() const { return cap_; }
  uint64_t peak() const { return peak_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t cap_ = 0, align_ = 256, used_ = 0, peak_ = 0;
};

struct Group;

struct Ticket {
  uint64_t id = 0;
  Plan plan;
  int status = EBI_PENDING;
  bool uses_column = false;  // counted in plan.c->users until it completes
  bool claimed = false;      // a serve_wait call is waiting on it
  Clock::time_point t_submit, t_done;
  uint64_t dev_off = 0, dev_len = 0, pin_off = 0, pin_len = 0;
  Group *group = nullptr;
};

struct Group {
  PlanGroup pg;
  std::vector<Ticket *> tickets;
  hipEvent_t done = nullptr;
  uint64_t dev_off = 0, pin_off = 0, len = 0;
  size_t remaining = 0;
};

}  // namespace
}  // namespace ebi

using namespace ebi;

struct ebi_hip_server {
  ebi_hip_ctx *ctx = nullptr;
  int device = 0;
  int cu_count = 256;
  ebi_serve_opts opts{};
  uint8_t *dev_base = nullptr;
  uint8_t *pin_base = nullptr;
  Arena dev, pin;
  std::vector<hipStream_t> streams;
  std::vector<hipStream_t> copy_streams;  // one per finisher
  std::vector<hipEvent_t> free_events;
  std::mutex mu;
  std::condition_variable cv_sched, cv_fin, cv_done;
  std::deque<Ticket *> pending;
  std::unordered_map<uint64_t, Ticket *> tickets;
  std::vector<Group *> groups;     // on the device
  std::deque<Ticket *> finish_q;   // device work complete, host finish pending
  std::thread sched;
  std::vector<std::thread> finishers;
  bool stop = false, sched_done = false;
  uint64_t next_id = 1;
  uint32_t inflight = 0, rr = 0;
  ebi_serve_stats stats{};
};

namespace {

// Completes `t` (lock held).
void complete(ebi_hip_server *s, Ticket *t, int status) {
  if (t->uses_column) {
    t->plan.c->users.fetch_sub(1);
    t->uses_column = false;
  }
  t->status = status;
  t->t_done = Clock::now();
  s->stats.completed++;
  if (status == EBI_REJECTED) s->stats.rejected++;
}

void release_ticket_memory(ebi_hip_server *s, Ticket *t) {
  if (t->dev_len) s->dev.release(t->dev_off, t->dev_len);
  if (t->pin_len) s->pin.release(t->pin_off, t->pin_len);
  t->dev_len = t->pin_len = 0;
}

// Admits queued tickets into one group (lock held).  Returns nullptr if none.
Group *admit(ebi_hip_server *s) {
  if (s->pending.empty() || s->inflight >= s->opts.max_inflight) return nullptr;
  const uint64_t gb = group_block_bytes(s->opts.max_batch);
  uint64_t g_dev = 0, g_pin = 0;
  if (!s->dev.alloc(gb, &g_dev)) return nullptr;
  if (!s->pin.alloc(gb, &g_pin)) {
    s->dev.release(g_dev, gb);
    return nullptr;
  }
  Group *g = new Group();
  g->dev_off = g_dev;
  g->pin_off = g_pin;
  g->len = gb;
  while (!s->pending.empty() && s->inflight < s->opts.max_inflight && g->tickets.size() < s->opts.max_batch) {
    Ticket *t = s->pending.front();
    const PlanNeed need = plan_need(t->plan);
    if (need.dev + gb > s->dev.cap() || need.pin + gb > s->pin.cap()) {
      s->pending.pop_front();  // can never fit: reject, keep admitting
      complete(s, t, EBI_REJECTED);
      s->cv_done.notify_all();
      continue;
    }
    uint64_t d = 0, p = 0;
    if (!s->dev.alloc(need.dev, &d)) break;  // wait for memory, in arrival order
    if (!s->pin.alloc(need.pin, &p)) {
      s->dev.release(d, need.dev);
      break;
    }
    s->pending.pop_front();
    t->dev_off = d;
    t->dev_len = need.dev;
    t->pin_off = p;
    t->pin_len = need.pin;
    t->group = g;
    g->tickets.push_back(t);
    g->pg.plans.push_back(&t->plan);
    s->inflight++;
  }
  if (g->tickets.empty()) {
    s->dev.release(g_dev, gb);
    s->pin.release(g_pin, gb);
    delete g;
    return nullptr;
  }
  g->remaining = g->tickets.size();
  s->stats.batches++;
  s->stats.max_inflight = std::max<uint64_t>(s->stats.max_inflight, s->inflight);
  s->stats.device_bytes_peak = s->dev.peak();
  if (s->free_events.empty()) {
    hipEvent_t e;
    hipEventCreateWithFlags(&e, hipEventDisableTiming);
    s->free_events.push_back(e);
  }
  g->done = s->free_events.back();
  s->free_events.pop_back();
  return g;
}

// Issues a group on its stream (no lock: touches only the group's memory).
hipError_t issue(ebi_hip_server *s, Group *g, hipStream_t st) {
  group_bind(g->pg, s->dev_base + g->dev_off, s->pin_base + g->pin_off, s->opts.max_batch);
  for (Ticket *t : g->tickets) {
    plan_bind(t->plan, s->dev_base + t->dev_off, s->pin_base + t->pin_off);
    plan_stage(t->plan);
  }
  hipError_t e = group_enqueue(g->pg, st, s->cu_count);
  if (e == hipSuccess) e = hipEventRecord(g->done, st);
  return e;
}

void scheduler(ebi_hip_server *s) {
  hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    Group *g = admit(s);
    if (g) {
      hipStream_t st = s->streams[s->rr++ % s->streams.size()];
      lk.unlock();
      const hipError_t e = issue(s, g, st);
      lk.lock();
      if (e != hipSuccess) {
        for (Ticket *t : g->tickets) {
          t->plan.status = EBI_HIP;
          t->plan.error = hipGetErrorString(e);
        }
        hipEventRecord(g->done, st);  // completes after whatever was queued
      }
      s->groups.push_back(g);
      continue;  // admit more before polling
    }
    // poll the groups on the device
    bool progressed = false;
    for (size_t i = 0; i < s->groups.size();) {
      Group *q = s->groups[i];
      const hipError_t qe = hipEventQuery(q->done);
      if (qe == hipErrorNotReady) {
        ++i;
        continue;
      }
      // any other result than success is a device or runtime error (it is
      // sticky, so the event would never report completion): the group's
      // queries fail with it, and the finishers skip their result copies
      const uint32_t flags = qe == hipSuccess ? q->pg.h_flags[0] : 0u;
      for (Ticket *t : q->tickets) {
        if (qe != hipSuccess && !t->plan.status) {
          t->plan.status = EBI_HIP;
          t->plan.error = hipGetErrorString(qe);
        }
        if (flags && !t->plan.status) t->plan.status = map_dev_err(flags);
        s->finish_q.push_back(t);
      }
      s->free_events.push_back(q->done);
      q->done = nullptr;
      s->groups.erase(s->groups.begin() + i);
      progressed = true;
    }
    if (progressed) s->cv_fin.notify_all();
    if (s->stop && s->pending.empty() && s->groups.empty()) break;
    if (!s->groups.empty()) {
      // device work outstanding: short poll interval
      s->cv_sched.wait_for(lk, std::chrono::microseconds(20));
    } else {
      // Nothing on the device and nothing admissible now (admit() just ran
      // under this lock): sleep until a submit, a completion that frees a
      // slot or memory, or stop.  An unconditional wait: a predicate that is
      // already true would return without releasing the lock, and the
      // finisher could then never free the memory admission waits for.
      s->cv_sched.wait(lk);
    }
  }
  s->sched_done = true;
  s->cv_fin.notify_all();
}

constexpr int kFinishers = 4;

void finisher(ebi_hip_server *s, hipStream_t copy_stream) {
  hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_fin.wait(lk, [&] { return !s->finish_q.empty() || (s->sched_done && s->inflight == 0); });
    if (s->finish_q.empty()) {
      if (s->sched_done && s->inflight == 0) break;
      continue;
    }
    Ticket *t = s->finish_q.front();
    s->finish_q.pop_front();
    lk.unlock();
    int status = t->plan.status;
    if (!status) {
      hipError_t e = hipSuccess;
      if (plan_enqueue_values(t->plan, copy_stream, &e) && e == hipSuccess) e = hipStreamSynchronize(copy_stream);
      status = e != hipSuccess ? EBI_HIP : plan_finish(t->plan);
    }
    lk.lock();
    release_ticket_memory(s, t);
    Group *g = t->group;
    if (g && --g->remaining == 0) {
      s->dev.release(g->dev_off, g->len);
      s->pin.release(g->pin_off, g->len);
      delete g;
    }
    t->group = nullptr;
    s->inflight--;
    complete(s, t, status);
    s->cv_done.notify_all();
    s->cv_sched.notify_one();
    if (s->sched_done && s->inflight == 0) s->cv_fin.notify_all();  // the other finishers end too
  }
}

}  // namespace

extern "C" int ebi_hip_serve_create(ebi_hip_ctx *ctx, const ebi_serve_opts *opts, ebi_hip_server **out) {
  if (!ctx || !out) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  ebi_hip_server *s = new (std::nothrow) ebi_hip_server();
  if (!s) return EBI_OOM;
  if (opts) s->opts = *opts;
  if (!s->opts.device_budget) s->opts.device_budget = 8ull << 30;
  if (!s->opts.pinned_budget) s->opts.pinned_budget = 256ull << 20;
  if (!s->opts.max_inflight) s->opts.max_inflight = 256;
  if (!s->opts.n_streams) s->opts.n_streams = 4;
  if (!s->opts.max_batch) s->opts.max_batch = 64;
  s->opts.n_streams = std::min<uint32_t>(s->opts.n_streams, 16);
  s->opts.max_batch = std::min<uint32_t>(s->opts.max_batch, 4096);  // fills of a group: 3 per plan, one grid row each
  s->ctx = ctx;
  s->device = ctx->device;
  s->cu_count = ctx->cu_count;
  auto bail = [&](int code, hipError_t e, const char *what) {
    for (hipStream_t st : s->streams) hipStreamDestroy(st);
    for (hipStream_t st : s->copy_streams) hipStreamDestroy(st);
    if (s->dev_base) hipFree(s->dev_base);
    if (s->pin_base) hipHostFree(s->pin_base);
    delete s;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    return fail(ctx, code, what);
  };
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return bail(EBI_HIP, e, "serve: hipSetDevice");
  if ((e = hipMalloc((void **)&s->dev_base, s->opts.device_budget)) != hipSuccess)
    return bail(EBI_OOM, e, "serve: device arena");
  if ((e = hipHostMalloc((void **)&s->pin_base, s->opts.pinned_budget, hipHostMallocDefault)) != hipSuccess)
    return bail(EBI_OOM, e, "serve: pinned arena");
  s->dev.reset(s->opts.device_budget, kAlign);
  s->pin.reset(s->opts.pinned_budget, kAlign);
  for (uint32_t i = 0; i < s->opts.n_streams; ++i) {
    hipStream_t st;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return bail(EBI_HIP, e, "serve: stream");
    s->streams.push_back(st);
  }
  for (int i = 0; i < kFinishers; ++i) {
    hipStream_t st;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
      return bail(EBI_HIP, e, "serve: copy stream");
    s->copy_streams.push_back(st);
  }
  s->sched = std::thread(scheduler, s);
  for (hipStream_t st : s->copy_streams) s->finishers.emplace_back(finisher, s, st);
  *out = s;
  return EBI_OK;
}

extern "C" int ebi_hip_serve_submit(ebi_hip_server *s, const ebi_query *q, size_t n, uint64_t *tickets) {
  if (!s || (!q && n) || (!tickets && n)) return EBI_INVALID_ARGUMENT;
  std::vector<Ticket *> made;
  made.reserve(n);
  const Clock::time_point now = Clock::now();
  for (size_t i = 0; i < n; ++i) {
    Ticket *t = new (std::nothrow) Ticket();
    if (!t) {
      for (Ticket *m : made) delete m;
      return EBI_OOM;
    }
    t->t_submit = now;
    t->status = plan_init(t->plan, q[i]);  // host-side validation only
    if (t->status == EBI_OK) {
      t->status = EBI_PENDING;
      t->plan.c->users.fetch_add(1);  // the column must stay loaded until it completes
      t->uses_column = true;
    }
    made.push_back(t);
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      for (Ticket *m : made) {
        if (m->uses_column) m->plan.c->users.fetch_sub(1);
        delete m;
      }
      return EBI_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < n; ++i) {
      Ticket *t = made[i];
      t->id = s->next_id++;
      tickets[i] = t->id;
      s->tickets[t->id] = t;
      s->stats.submitted++;
      if (t->status == EBI_PENDING) {
        s->pending.push_back(t);
      } else {
        t->t_done = now;
        s->stats.completed++;
      }
    }
  }
  s->cv_sched.notify_one();
  s->cv_done.notify_all();
  return EBI_OK;
}

extern "C" int ebi_hip_serve_poll(ebi_hip_server *s, uint64_t ticket, int *status) {
  if (!s) return EBI_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return EBI_INVALID_ARGUMENT;
  if (it->second->status == EBI_PENDING) return EBI_PENDING;
  if (status) *status = it->second->status;
  return EBI_OK;
}

extern "C" int ebi_hip_serve_wait(ebi_hip_server *s, uint64_t ticket, int64_t timeout_ms, int *status,
                                  double *latency_ms) {
  if (!s) return EBI_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return EBI_INVALID_ARGUMENT;
  Ticket *t = it->second;
  // one waiter per ticket: the waiter that returns EBI_OK releases it, so a
  // second concurrent waiter would be left holding a freed ticket
  if (t->claimed) return EBI_INVALID_ARGUMENT;
  t->claimed = true;
  auto done = [&] { return t->status != EBI_PENDING; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
    t->claimed = false;
    return EBI_PENDING;
  }
  if (status) *status = t->status;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(t->t_done - t->t_submit).count();
  s->tickets.erase(ticket);
  delete t;
  return EBI_OK;
}

extern "C" int ebi_hip_serve_stats(const ebi_hip_server *s, ebi_serve_stats *st) {
  if (!s || !st) return EBI_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(const_cast<ebi_hip_server *>(s)->mu);
  *st = s->stats;
  return EBI_OK;
}

extern "C" int ebi_hip_serve_destroy(ebi_hip_server *s) {
  if (!s) return EBI_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  if (s->sched.joinable()) s->sched.join();
  s->cv_fin.notify_all();
  for (std::thread &f : s->finishers)
    if (f.joinable()) f.join();
  for (auto &kv : s->tickets) {
    if (kv.second->uses_column) kv.second->plan.c->users.fetch_sub(1);
    delete kv.second;
  }
  for (hipStream_t st : s->streams) hipStreamDestroy(st);
  for (hipStream_t st : s->copy_streams) hipStreamDestroy(st);
  for (hipEvent_t e : s->free_events) hipEventDestroy(e);
  hipFree(s->dev_base);
  hipHostFree(s->pin_base);
  delete s;
  return EBI_OK;
}
