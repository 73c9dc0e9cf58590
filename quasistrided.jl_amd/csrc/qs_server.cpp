// A continuous-batching request loop over resident operands and plans.
//
// No counterpart in the reference (one contraction per call, single thread,
// src/planning/plan.jl:12-13).  Clients submit contractions whose operands
// already live in HBM, each against a plan built once for its layout (the
// plan's offset tables are shared by every request of that layout; the
// storages and coefficients are snapshotted at submit).  A worker thread
// admits requests first-come first-served while the bytes they hold stay
// under the budget, packs what is admitted into a grouped launch on one of four
// streams (a batch per free stream: requests arriving while batches run
// start on another stream, so their work overlaps on the device), and
// retires batches as their events complete.  Every member computes its K sum
// unsplit with the tile class of its own shape, so a request's result is the
// same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "qs_plan.h"

namespace qs {

using Clock = std::chrono::steady_clock;

struct Request {
  GemmArgs g;
  qs_plan *plan = nullptr;
  int cls = TILE_64x64;
  int64_t bytes = 0, flops = 0;
  double al[2] = {1, 0}, be[2] = {0, 0};
  Clock::time_point t_submit, t_done;
  bool done = false;
};

// In-flight batches, one stream each: a new batch can start while up to
// three others run, so small requests do not queue behind two large ones.
static constexpr int kSlots = 4;
// Grouped mode: a request of at least this much work runs in a batch of its
// own (a batch retires when its last member does, so small members would wait
// for it), and large batches hold at most kSlots - 1 streams; small requests
// may pass a large one that waits for a stream.
static constexpr int64_t kLargeFlops = int64_t(4) << 30;

}  // namespace qs

struct qs_server {
  int64_t budget = 0;
  int max_batch = 64;
  int grouped = 1;
  int slots = qs::kSlots;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, qs::Request> reqs;
  uint64_t next_id = 1;
  int64_t in_flight_bytes = 0, peak_bytes = 0;
  hipStream_t streams[qs::kSlots] = {};
  hipEvent_t events[qs::kSlots] = {};
  qs::BatchBuffers bufs[qs::kSlots];
  bool busy[qs::kSlots] = {};
  bool large[qs::kSlots] = {};   // the slot's batch is one large request
  std::vector<uint64_t> slot_ids[qs::kSlots];
  std::thread worker;
  bool stop = false;
  std::string error;
  // statistics
  std::vector<double> lat_ms;
  int64_t submitted = 0, completed = 0, rejected = 0, batches = 0, batched = 0, flops_done = 0;
  qs::Clock::time_point first_submit, last_done;
  bool any = false;
};

namespace qs {

static void retire(qs_server *s, int slot) {
  const auto now = Clock::now();
  for (uint64_t id : s->slot_ids[slot]) {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) continue;
    Request &r = it->second;
    r.done = true;
    r.t_done = now;
    s->in_flight_bytes -= r.bytes;
    s->lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r.t_submit).count());
    s->completed++;
    s->flops_done += r.flops;
    s->last_done = now;
  }
  s->slot_ids[slot].clear();
  s->busy[slot] = false;
  s->large[slot] = false;
}

static void worker_loop(qs_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // Retire finished batches.
    bool retired = false;
    for (int i = 0; i < s->slots; ++i)
      if (s->busy[i] && hipEventQuery(s->events[i]) == hipSuccess) {
        retire(s, i);
        retired = true;
      }
    if (retired) s->cv_done.notify_all();
    // Admit into a free slot.
    int slot = -1;
    for (int i = 0; i < s->slots; ++i)
      if (!s->busy[i]) { slot = i; break; }
    std::vector<uint64_t> batch;
    bool batch_large = false;
    if (slot >= 0 && !s->grouped) {
      while (!s->pending.empty() && (int)batch.size() < s->max_batch) {
        Request &r = s->reqs[s->pending.front()];
        if (s->in_flight_bytes + r.bytes > s->budget) break;   // first-come first-served
        s->in_flight_bytes += r.bytes;
        s->peak_bytes = std::max(s->peak_bytes, s->in_flight_bytes);
        batch.push_back(s->pending.front());
        s->pending.pop_front();
      }
    } else if (slot >= 0) {
      int large_busy = 0;
      for (int i = 0; i < s->slots; ++i) large_busy += s->busy[i] && s->large[i];
      for (auto it = s->pending.begin(); it != s->pending.end() && (int)batch.size() < s->max_batch;) {
        Request &r = s->reqs[*it];
        const bool big = r.flops >= kLargeFlops;
        if (big && !batch.empty()) break;   // never shares a batch; the batch so far starts now
        if (big && large_busy >= s->slots - 1) {   // waits for a stream; smaller requests behind it may pass
          ++it;
          continue;
        }
        if (s->in_flight_bytes + r.bytes > s->budget) break;   // first-come first-served within the budget
        s->in_flight_bytes += r.bytes;
        s->peak_bytes = std::max(s->peak_bytes, s->in_flight_bytes);
        batch.push_back(*it);
        it = s->pending.erase(it);
        if (big) {
          batch_large = true;
          break;
        }
      }
    }
    if (!batch.empty()) {
      std::vector<GemmArgs> args;
      std::vector<int> cls;
      std::vector<double> coef;
      std::vector<Request *> solo;
      for (uint64_t id : batch) {
        Request &r = s->reqs[id];
        if (!s->grouped || !grouped_member(r.g)) {
          solo.push_back(&r);
          continue;
        }
        args.push_back(r.g);
        cls.push_back(r.cls);
        coef.insert(coef.end(), {r.al[0], r.al[1], r.be[0], r.be[1]});
      }
      s->slot_ids[slot] = batch;
      s->busy[slot] = true;
      s->large[slot] = batch_large;
      s->batches++;
      s->batched += (int64_t)batch.size();
      lk.unlock();
      qs_status st = launch_member_batch(args, cls, coef, s->bufs[slot], s->streams[slot]);
      for (Request *r : solo) {   // one launch each, K unsplit (no shared workspace)
        if (st != QS_OK) break;
        hipError_t e = hipSuccess;
        if (r->al[0] == 0 && r->al[1] == 0) {
          if (r->g.beta_case != 1) e = launch_scale(r->g, s->streams[slot]);
        } else {
          e = launch_gemm(r->g, s->streams[slot]);
        }
        if (e != hipSuccess) st = fail(QS_ERR_DEVICE, std::string("serving: launch: ") + hipGetErrorString(e));
      }
      (void)hipEventRecord(s->events[slot], s->streams[slot]);
      lk.lock();
      if (st != QS_OK && s->error.empty()) s->error = qs_last_error();
      continue;
    }
    if (s->stop && s->pending.empty()) {
      bool idle = true;
      for (int i = 0; i < s->slots; ++i) idle = idle && !s->busy[i];
      if (idle) break;
    }
    bool any_busy = false;
    for (int i = 0; i < s->slots; ++i) any_busy = any_busy || s->busy[i];
    if (any_busy)
      s->cv_work.wait_for(lk, std::chrono::microseconds(20));   // poll the in-flight events
    else
      s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
  }
}

static double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace qs

using namespace qs;

extern "C" {

qs_status qs_server_create(int64_t budget_bytes, int32_t max_batch, int32_t grouped, qs_server **out) {
  if (!out) return fail(QS_ERR_ARGUMENT, "qs_server_create: null output");
  *out = nullptr;
  if (budget_bytes <= 0 || max_batch < 1) return fail(QS_ERR_ARGUMENT, "qs_server_create: budget and max_batch must be positive");
  qs_server *s = new qs_server();
  s->budget = budget_bytes;
  s->max_batch = max_batch;
  s->grouped = grouped ? 1 : 0;
  s->slots = grouped ? kSlots : 1;   // per-request launches run in order on one stream
  if (!grouped) s->max_batch = 1;
  (void)hipGetDevice(&s->device);
  for (int i = 0; i < s->slots; ++i) {
    if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->events[i], hipEventDisableTiming) != hipSuccess) {
      qs_server_destroy(s);
      return fail(QS_ERR_DEVICE, "qs_server_create: streams");
    }
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return QS_OK;
}

qs_status qs_server_submit(qs_server *s, qs_plan *p, const void *a, int64_t a_base, int64_t a_len, const void *b,
                           int64_t b_base, int64_t b_len, void *c, int64_t c_base, int64_t c_len, const double *alpha,
                           const double *beta, uint64_t *ticket) {
  if (!s || !p || !ticket) return fail(QS_ERR_ARGUMENT, "qs_server_submit: null argument");
  if (p->desc.memory != QS_MEM_DEVICE) return fail(QS_ERR_ARGUMENT, "qs_server_submit: device-memory plans only");
  if (p->released) return fail(QS_ERR_ARGUMENT, "qs_server_submit: the plan was released");
  if (p->path != QS_PATH_GEMM && p->path != QS_PATH_SKINNY)
    return fail(QS_ERR_ARGUMENT, "qs_server_submit: the plan does not run the tiled GEMM path");
  if (!a) { a = p->desc.a; a_base = p->desc.a_base; a_len = p->desc.a_len; }
  if (!b) { b = p->desc.b; b_base = p->desc.b_base; b_len = p->desc.b_len; }
  if (!c) { c = p->desc.c; c_base = p->desc.c_base; c_len = p->desc.c_len; }
  const int64_t alo = a_base + p->mA_lo + p->kA_lo, ahi = a_base + p->mA_hi + p->kA_hi;
  const int64_t blo = b_base + p->kB_lo + p->nB_lo, bhi = b_base + p->kB_hi + p->nB_hi;
  const int64_t clo = c_base + p->mC_lo + p->nC_lo, chi = c_base + p->mC_hi + p->nC_hi;
  if (alo < 0 || ahi >= a_len || blo < 0 || bhi >= b_len || clo < 0 || chi >= c_len)
    return fail(QS_ERR_BOUNDS, "qs_server_submit: operand spans exceed their storages");
  Request r;
  r.plan = p;
  r.g = make_args(p, a, a_base, b, b_base, c, c_base);
  r.g.k_splits = 1;
  r.g.k_chunk = p->K;
  r.g.partials = nullptr;
  const int ct = p->desc.compute_type;
  r.cls = grouped_tile(p->M, p->N, ct);
  r.al[0] = alpha ? alpha[0] : 1.0;
  r.al[1] = alpha && is_complex(ct) ? alpha[1] : 0.0;
  r.be[0] = beta ? beta[0] : 0.0;
  r.be[1] = beta && is_complex(ct) ? beta[1] : 0.0;
  r.g.alpha_re = r.al[0]; r.g.alpha_im = r.al[1]; r.g.beta_re = r.be[0]; r.g.beta_im = r.be[1];
  r.g.beta_case = (r.be[0] == 0 && r.be[1] == 0) ? 0 : (r.be[0] == 1 && r.be[1] == 0) ? 1 : 2;
  r.bytes = (ahi - alo + 1) * elem_bytes(p->desc.a_type) + (bhi - blo + 1) * elem_bytes(p->desc.b_type) +
            (chi - clo + 1) * elem_bytes(p->desc.c_type);
  r.flops = (is_complex(ct) ? 8 : 2) * p->M * p->N * p->K;
  if (r.bytes > s->budget) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->rejected++;
    return fail(QS_ERR_MEMORY, "qs_server_submit: the request alone exceeds the memory budget");
  }
  r.t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(QS_ERR_ARGUMENT, "qs_server_submit: server is stopping");
    const uint64_t id = s->next_id++;
    if (!s->any) { s->first_submit = r.t_submit; s->any = true; }
    s->reqs.emplace(id, r);
    s->pending.push_back(id);
    s->submitted++;
    *ticket = id;
  }
  s->cv_work.notify_one();
  return QS_OK;
}

qs_status qs_server_poll(qs_server *s, uint64_t ticket, int32_t *done, double *latency_ms) {
  if (!s || !done) return fail(QS_ERR_ARGUMENT, "qs_server_poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(QS_ERR_ARGUMENT, "qs_server_poll: unknown ticket");
  *done = it->second.done ? 1 : 0;
  if (it->second.done) {
    if (latency_ms)
      *latency_ms = std::chrono::duration<double, std::milli>(it->second.t_done - it->second.t_submit).count();
    s->reqs.erase(it);
  }
  return QS_OK;
}

qs_status qs_server_wait(qs_server *s, uint64_t ticket, double *latency_ms) {
  if (!s) return fail(QS_ERR_ARGUMENT, "qs_server_wait: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(QS_ERR_ARGUMENT, "qs_server_wait: unknown ticket");
  s->cv_done.wait(lk, [&] { return s->reqs[ticket].done || !s->error.empty(); });
  if (!s->error.empty()) return fail(QS_ERR_DEVICE, "serving: " + s->error);
  Request &r = s->reqs[ticket];
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  s->reqs.erase(ticket);
  return QS_OK;
}

qs_status qs_server_drain(qs_server *s) {
  if (!s) return fail(QS_ERR_ARGUMENT, "qs_server_drain: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    bool idle = s->pending.empty();
    for (int i = 0; i < s->slots; ++i) idle = idle && !s->busy[i];
    return idle || !s->error.empty();
  });
  if (!s->error.empty()) return fail(QS_ERR_DEVICE, "serving: " + s->error);
  return QS_OK;
}

qs_status qs_server_stats_get(qs_server *s, qs_server_stats *out) {
  if (!s || !out) return fail(QS_ERR_ARGUMENT, "qs_server_stats_get: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->batches = s->batches;
  out->mean_batch = s->batches ? (double)s->batched / (double)s->batches : 0.0;
  out->p50_ms = percentile(s->lat_ms, 0.50);
  out->p90_ms = percentile(s->lat_ms, 0.90);
  out->p99_ms = percentile(s->lat_ms, 0.99);
  out->max_ms = s->lat_ms.empty() ? 0.0 : *std::max_element(s->lat_ms.begin(), s->lat_ms.end());
  double sum = 0;
  for (double v : s->lat_ms) sum += v;
  out->mean_ms = s->lat_ms.empty() ? 0.0 : sum / (double)s->lat_ms.size();
  out->span_s = s->any && s->completed ? std::chrono::duration<double>(s->last_done - s->first_submit).count() : 0.0;
  out->throughput_rps = out->span_s > 0 ? (double)s->completed / out->span_s : 0.0;
  out->tflops = out->span_s > 0 ? (double)s->flops_done / out->span_s / 1e12 : 0.0;
  out->peak_bytes = s->peak_bytes;
  out->budget_bytes = s->budget;
  return QS_OK;
}

qs_status qs_server_reset_stats(qs_server *s) {
  if (!s) return fail(QS_ERR_ARGUMENT, "qs_server_reset_stats: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_ms.clear();
  s->submitted = s->completed = s->rejected = s->batches = s->batched = s->flops_done = 0;
  s->any = false;
  s->peak_bytes = s->in_flight_bytes;
  return QS_OK;
}

qs_status qs_server_destroy(qs_server *s) {
  if (!s) return QS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (int i = 0; i < kSlots; ++i) {
    if (s->streams[i]) {
      (void)hipStreamSynchronize(s->streams[i]);
      (void)hipStreamDestroy(s->streams[i]);
    }
    if (s->events[i]) (void)hipEventDestroy(s->events[i]);
    BatchBuffers &B = s->bufs[i];
    if (B.d_items) (void)dev_free(B.d_items);
    if (B.h_items) (void)pinned_free(B.h_items);
    if (B.d_coef) (void)dev_free(B.d_coef);
    if (B.h_coef) (void)pinned_free(B.h_coef);
  }
  delete s;
  return QS_OK;
}

}  // extern "C"
