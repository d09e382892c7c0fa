// Continuous-batching grad_run server behind the C ABI (qg_server_*, qg_submit,
// qg_poll, qg_wait_any).
//
// A server owns one device arena (a single allocation made at creation) and
// `slots` worker threads, each with its own context (HIP stream) and one
// McClean plan that it re-lays out for every request's shape
// (qg_plan_reshape): after the plans' tables have grown to the largest shapes
// seen, a shape change allocates nothing and frees nothing on the device.
// Requests are admitted by footprint: a request waits in the queue until a
// worker is idle and the arena can hold its workspace (qg_mcclean_workspace_
// bytes, the library's own layout rule); queued requests that fit may
// overtake a head-of-queue request that does not at most `backfill_limit`
// times, after which the workers wait for the head (no starvation).  Kernels
// of different slots run on different streams, so requests overlap on the
// device.  The memory mode of a request depends on its shape only (the
// per-pass history when it needs at most history_max_bytes, else checkpoints
// every ckpt_every layers), so a request's result is bitwise the same alone
// and in any mix.
//
// QG_SERVER_SHORTEST_FIRST replaces the arrival order by the estimated work
// of a request (its algorithmic bytes, (6 L + 3) 2^n amplitudes): the cheapest
// queued request that fits starts first.  A request that later arrivals have
// overtaken backfill_limit times is served in arrival order before any other,
// and while it does not fit nothing else starts (aging: no starvation).
//
// QG_SERVER_SIMULATE runs the same admission and scheduling without a device:
// a request sleeps for a time proportional to its footprint and returns
// E = n and grad = L everywhere.  It exists so the scheduler can be tested
// on hosts without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/qradient_amd.h"

namespace qg {
int set_error(int code, const std::string& msg);  // qg_api.cpp: the thread's qg_last_error
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit sub-allocator of [0, size) with coalescing; blocks are multiples
// of kAlign bytes.
class Arena {
 public:
  static constexpr size_t kAlign = 64 * 1024;
  void reset(size_t size) {
    size_ = size / kAlign * kAlign;
    free_.clear();
    if (size_) free_[0] = size_;
    used_ = peak_ = 0;
  }
  static size_t round(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
  bool alloc(size_t bytes, size_t* off) {
    const size_t b = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < b) continue;
      *off = it->first;
      const size_t rest = it->second - b;
      const size_t at = it->first + b;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += b;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    const size_t b = round(bytes);
    used_ -= b;
    auto it = free_.emplace(off, b).first;
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
  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0, peak_ = 0;
};

struct Req {
  int64_t ticket = 0;
  int n = 0, L = 0;
  std::vector<int32_t> axes;
  std::vector<double> angles;
  std::vector<qg_term> terms;
  int ckpt = QG_HISTORY;  // the plan's ckpt_every argument
  int mode = 0;           // reported: 0 history, k > 0 checkpoints
  size_t footprint = 0, offset = 0;
  double cost = 0.0;   // estimated work: (6 L + 3) 2^n, in amplitudes moved
  int overtaken = 0;   // later arrivals started before this request
  bool done = false;
  int status = QG_OK;
  std::string err;
  double E = 0.0;
  std::vector<double> grad;
  Clock::time_point t_submit, t_start, t_end;
  double dev_start = 0.0, dev_end = 0.0;
  int slot = -1;
};

struct Slot {
  hipStream_t stream = nullptr;
  qg_ctx* ctx = nullptr;
  qg_plan* plan = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  std::thread th;
};

}  // namespace

struct qg_server {
  int device = 0, dtype = QG_C64, flags = 0, backfill_limit = 8, ckpt_every = 8;
  size_t history_max = 0;
  char* arena_base = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  hipEvent_t ref_event = nullptr;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
  std::deque<int64_t> finished;  // done and not yet handed out by qg_wait_any
  int64_t next_ticket = 1;
  int bypassed = 0;
  int running = 0;
  bool closing = false;
  uint64_t completed = 0;
  bool simulate() const { return (flags & QG_SERVER_SIMULATE) != 0; }
  bool shortest_first() const { return (flags & QG_SERVER_SHORTEST_FIRST) != 0; }
};

namespace {

int footprint_of(qg_server* s, int n, int L, int* ckpt, int* mode, size_t* bytes) {
  qg_ctx* ctx = s->slots.empty() ? nullptr : s->slots[0].ctx;
  size_t hb = 0;
  int chosen = 0;
  if (qg_mcclean_workspace_bytes(ctx, n, L, s->dtype, QG_HISTORY, &hb, &chosen) == QG_OK && hb <= s->history_max) {
    *ckpt = QG_HISTORY;
    *mode = 0;
    *bytes = hb;
    return QG_OK;
  }
  const int r = qg_mcclean_workspace_bytes(ctx, n, L, s->dtype, s->ckpt_every, bytes, &chosen);
  if (r != QG_OK) return r;
  *ckpt = s->ckpt_every;
  *mode = chosen;
  return QG_OK;
}

// Shortest-first order: the oldest request that has been overtaken
// backfill_limit times, else the cheapest that fits (ties: the earlier).
int pick_shortest(qg_server* s) {
  for (size_t i = 0; i < s->queue.size(); ++i) {
    Req* r = s->queue[i];
    if (r->overtaken < s->backfill_limit) continue;
    return s->arena.alloc(r->footprint, &r->offset) ? (int)i : -1;  // else hold capacity for it
  }
  std::vector<int> order(s->queue.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return s->queue[a]->cost < s->queue[b]->cost; });
  for (int i : order)
    if (s->arena.alloc(s->queue[i]->footprint, &s->queue[i]->offset)) return i;
  return -1;
}

// Index of the queued request to run next, its arena block taken; -1 if none
// may run now.  The requests queued before it count one more overtaking.
// (Lock held.)
int pick(qg_server* s) {
  int pick_i = -1;
  if (s->shortest_first()) {
    pick_i = pick_shortest(s);
  } else {
    for (size_t i = 0; i < s->queue.size(); ++i) {
      Req* r = s->queue[i];
      if (s->arena.alloc(r->footprint, &r->offset)) {
        s->bypassed = i > 0 ? s->bypassed + 1 : 0;
        pick_i = (int)i;
        break;
      }
      if (i == 0 && s->bypassed >= s->backfill_limit) return -1;  // hold capacity for the head
    }
  }
  for (int i = 0; i < pick_i; ++i) s->queue[i]->overtaken++;
  return pick_i;
}

int run_on_device(qg_server* s, Slot& sl, Req* r) {
  int rc = QG_OK;
  if (!sl.plan)
    rc = qg_mcclean_plan(sl.ctx, r->n, r->L, s->dtype, QG_ENT_CNOT, r->ckpt, &sl.plan);
  else
    rc = qg_plan_reshape(sl.plan, r->n, r->L, r->ckpt);
  if (rc != QG_OK) return rc;
  if (r->footprint > 0) {
    rc = qg_plan_set_workspace(sl.plan, s->arena_base + r->offset, Arena::round(r->footprint));
    if (rc != QG_OK) return rc;
  }
  if (hipEventRecord(sl.ev_start, sl.stream) != hipSuccess) return qg::set_error(QG_ERR_HIP, "event record failed");
  rc = qg_mcclean_grad_run(sl.plan, r->axes.data(), r->angles.data(), r->terms.data(), (int)r->terms.size(), nullptr,
                           &r->E, r->grad.data());
  if (rc != QG_OK) return rc;
  if (hipEventRecord(sl.ev_end, sl.stream) != hipSuccess || hipEventSynchronize(sl.ev_end) != hipSuccess)
    return qg::set_error(QG_ERR_HIP, "event record failed");
  float a = 0.f, b = 0.f;
  if (hipEventElapsedTime(&a, s->ref_event, sl.ev_start) == hipSuccess &&
      hipEventElapsedTime(&b, s->ref_event, sl.ev_end) == hipSuccess) {
    r->dev_start = a;
    r->dev_end = b;
  }
  return QG_OK;
}

void simulate_run(qg_server* s, Req* r) {
  r->dev_start = ms_between(s->t0, Clock::now());
  // 20 microseconds plus 1 microsecond per MiB of workspace
  const double us = 20.0 + (double)r->footprint / (1 << 20);
  std::this_thread::sleep_for(std::chrono::microseconds((int64_t)us));
  r->E = (double)r->n;
  std::fill(r->grad.begin(), r->grad.end(), (double)r->L);
  r->dev_end = ms_between(s->t0, Clock::now());
}

void worker(qg_server* s, int si) {
  Slot& sl = s->slots[si];
  if (!s->simulate()) (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    int i = -1;
    s->cv_work.wait(lk, [&] { return (i = pick(s)) >= 0 || (s->closing && s->queue.empty()); });
    if (i < 0) return;
    Req* r = s->queue[i];
    s->queue.erase(s->queue.begin() + i);
    r->slot = si;
    ++s->running;
    lk.unlock();
    r->t_start = Clock::now();
    int rc = QG_OK;
    if (s->simulate()) {
      simulate_run(s, r);
    } else {
      rc = run_on_device(s, sl, r);
      if (rc != QG_OK) r->err = qg_last_error();
    }
    r->t_end = Clock::now();
    lk.lock();
    r->status = rc;
    r->done = true;
    --s->running;
    s->arena.release(r->offset, r->footprint);
    ++s->completed;
    s->finished.push_back(r->ticket);
    s->cv_done.notify_all();
    s->cv_work.notify_all();  // memory and a worker are free again
  }
}

int validate(int n, int L, const int32_t* axes, const double* angles, const qg_term* terms, int nterms) {
  if (n < 1 || n > 32) return qg::set_error(QG_ERR_INVALID, "qubit number must be in [1, 32]");
  if (L < 1) return qg::set_error(QG_ERR_INVALID, "layer number must be >= 1");
  if (!axes || !angles || (nterms > 0 && !terms) || nterms < 0) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  for (int k = 0; k < L * n; ++k) {
    if (axes[k] < 0 || axes[k] > 2) return qg::set_error(QG_ERR_INVALID, "Invalid axis " + std::to_string(axes[k]));
    if (!std::isfinite(angles[k])) return qg::set_error(QG_ERR_INVALID, "non-finite angle");
  }
  for (int k = 0; k < nterms; ++k) {
    const qg_term& t = terms[k];
    if (t.i < 0 || t.i >= n) return qg::set_error(QG_ERR_INVALID, "observable: qubit index out of range");
    if (!std::isfinite(t.w)) return qg::set_error(QG_ERR_INVALID, "observable: non-finite weight");
    if (t.kind == QG_TERM_ZZ && (t.j <= t.i || t.j >= n))
      return qg::set_error(QG_ERR_INVALID, "observable: zz terms need i < j < n (upper triangle, observable.py:56-64)");
    if (t.kind < QG_TERM_X || t.kind > QG_TERM_ZZ) return qg::set_error(QG_ERR_INVALID, "observable: unknown term kind");
  }
  return QG_OK;
}

void destroy_slots(qg_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.plan) qg_plan_destroy(sl.plan);
    if (sl.ctx) qg_ctx_destroy(sl.ctx);
    if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
    if (sl.ev_end) (void)hipEventDestroy(sl.ev_end);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  s->slots.clear();
}

}  // namespace

extern "C" int qg_server_create(int device, int slots, int dtype, size_t arena_bytes, size_t history_max_bytes,
                                int ckpt_every, int backfill_limit, int flags, qg_server** out) {
  if (!out) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  if (slots < 1 || slots > 64) return qg::set_error(QG_ERR_INVALID, "slots must be in [1, 64]");
  if (dtype != QG_C128 && dtype != QG_C64) return qg::set_error(QG_ERR_INVALID, "dtype must be QG_C128 or QG_C64");
  if (ckpt_every < 1) return qg::set_error(QG_ERR_INVALID, "ckpt_every must be >= 1");
  if (backfill_limit < 0) return qg::set_error(QG_ERR_INVALID, "backfill_limit must be >= 0");
  if (arena_bytes < Arena::kAlign) return qg::set_error(QG_ERR_INVALID, "arena_bytes must be at least 64 KiB");
  std::unique_ptr<qg_server> s(new qg_server());
  s->device = device;
  s->dtype = dtype;
  s->flags = flags;
  s->history_max = history_max_bytes;
  s->ckpt_every = ckpt_every;
  s->backfill_limit = backfill_limit;
  s->t0 = Clock::now();
  s->slots.resize(slots);
  if (!s->simulate()) {
    if (hipSetDevice(device) != hipSuccess) return qg::set_error(QG_ERR_INVALID, "qg_server_create: invalid device");
    void* base = nullptr;
    const hipError_t e = hipMalloc(&base, arena_bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return qg::set_error(QG_ERR_NOMEM, "qg_server_create: arena of " + std::to_string(arena_bytes) +
                                             " bytes: " + hipGetErrorString(e));
    }
    s->arena_base = (char*)base;
    for (Slot& sl : s->slots) {
      int rc = QG_OK;
      if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreate(&sl.ev_start) != hipSuccess || hipEventCreate(&sl.ev_end) != hipSuccess)
        rc = qg::set_error(QG_ERR_HIP, "qg_server_create: stream or event creation failed");
      if (rc == QG_OK) rc = qg_ctx_create(device, sl.stream, &sl.ctx);
      if (rc != QG_OK) {
        destroy_slots(s.get());
        (void)hipFree(s->arena_base);
        return rc;
      }
    }
    // device timestamps of requests are taken against this event
    if (hipEventCreate(&s->ref_event) != hipSuccess || hipEventRecord(s->ref_event, s->slots[0].stream) != hipSuccess ||
        hipEventSynchronize(s->ref_event) != hipSuccess) {
      destroy_slots(s.get());
      (void)hipFree(s->arena_base);
      return qg::set_error(QG_ERR_HIP, "qg_server_create: reference event failed");
    }
  }
  s->arena.reset(arena_bytes);
  qg_server* raw = s.release();
  for (int i = 0; i < slots; ++i) raw->slots[i].th = std::thread(worker, raw, i);
  *out = raw;
  return QG_OK;
}

extern "C" int qg_server_footprint(qg_server* s, int n, int L, size_t* bytes, int* ckpt_every) {
  if (!s || !bytes) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  int ck = 0, mode = 0;
  const int rc = footprint_of(s, n, L, &ck, &mode, bytes);
  if (rc == QG_OK && ckpt_every) *ckpt_every = mode;
  return rc;
}

extern "C" int qg_submit(qg_server* s, int n, int L, const int32_t* axes, const double* angles, const qg_term* terms,
                         int nterms, int64_t* ticket) {
  if (!s || !ticket) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  int rc = validate(n, L, axes, angles, terms, nterms);
  if (rc != QG_OK) return rc;
  std::unique_ptr<Req> r(new Req());
  r->n = n;
  r->L = L;
  rc = footprint_of(s, n, L, &r->ckpt, &r->mode, &r->footprint);
  if (rc != QG_OK) return rc;
  r->cost = (6.0 * L + 3.0) * std::ldexp(1.0, n);
  if (Arena::round(r->footprint) > s->arena.size())
    return qg::set_error(QG_ERR_NOMEM, "request needs " + std::to_string(r->footprint) +
                                           " bytes of workspace, the server's arena holds " +
                                           std::to_string(s->arena.size()));
  r->axes.assign(axes, axes + (size_t)L * n);
  r->angles.assign(angles, angles + (size_t)L * n);
  r->terms.assign(terms, terms + nterms);
  r->grad.assign((size_t)L * n, 0.0);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->closing) return qg::set_error(QG_ERR_INVALID, "server is closing");
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->queue.push_back(r.get());
  s->reqs.emplace(r->ticket, std::move(r));
  s->cv_work.notify_all();
  return QG_OK;
}

extern "C" int qg_wait_any(qg_server* s, int timeout_ms, int64_t* ticket) {
  if (!s || !ticket) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return !s->finished.empty(); };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
  if (s->finished.empty()) return QG_ERR_PENDING;
  *ticket = s->finished.front();
  s->finished.pop_front();
  return QG_OK;
}

extern "C" int qg_poll(qg_server* s, int64_t ticket, int timeout_ms, double* E_out, double* grad_out,
                       qg_request_info* info) {
  if (!s) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return qg::set_error(QG_ERR_INVALID, "unknown or already collected ticket");
  Req* r = it->second.get();
  auto done = [&] { return r->done; };
  if (timeout_ms < 0) s->cv_done.wait(lk, done);
  else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
  if (!r->done) return QG_ERR_PENDING;
  if (info) {
    info->queue_ms = ms_between(r->t_submit, r->t_start);
    info->run_ms = ms_between(r->t_start, r->t_end);
    info->latency_ms = ms_between(r->t_submit, r->t_end);
    info->device_start_ms = r->dev_start;
    info->device_end_ms = r->dev_end;
    info->slot = r->slot;
    info->ckpt_every = r->mode;
    info->footprint = r->footprint;
  }
  const int status = r->status;
  const std::string err = r->err;
  if (status == QG_OK) {
    if (E_out) *E_out = r->E;
    if (grad_out) std::memcpy(grad_out, r->grad.data(), sizeof(double) * r->grad.size());
  }
  auto f = std::find(s->finished.begin(), s->finished.end(), ticket);
  if (f != s->finished.end()) s->finished.erase(f);
  s->reqs.erase(it);
  if (status != QG_OK) return qg::set_error(status, err);
  return QG_OK;
}

extern "C" int qg_server_stats(qg_server* s, qg_server_counters* out) {
  if (!s || !out) return qg::set_error(QG_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  out->queued = (int64_t)s->queue.size();
  out->running = s->running;
  out->completed = (int64_t)s->completed;
  out->arena_bytes = s->arena.size();
  out->arena_in_use = s->arena.used();
  out->arena_peak = s->arena.peak();
  return QG_OK;
}

extern "C" int qg_server_destroy(qg_server* s) {
  if (!s) return QG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->closing = true;
  }
  s->cv_work.notify_all();
  for (Slot& sl : s->slots)
    if (sl.th.joinable()) sl.th.join();  // workers finish the queue first
  s->cv_done.notify_all();
  if (!s->simulate()) {
    (void)hipSetDevice(s->device);
    destroy_slots(s);
    if (s->ref_event) (void)hipEventDestroy(s->ref_event);
    (void)hipFree(s->arena_base);
  }
  delete s;
  return QG_OK;
}
