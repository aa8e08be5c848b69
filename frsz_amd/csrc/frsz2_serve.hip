// FRSZ2 serving runtime: a request loop that keeps compressed buffers
// resident in HBM and runs compress, decompress and random-access-read
// requests with continuous batching.
//
// * Memory: one device pool of hbm_budget bytes (a single hipMalloc) with a
//   first-fit allocator on the host; every request's output (a compressed
//   stream or decoded values) is carved from it, so the pool is the
//   admission budget.
// * Admission: a scheduler thread walks the queue in arrival order and
//   admits requests while their outputs fit the pool.  A request that reads
//   a stream (decompress, gather) waits until that stream has been written;
//   it does not hold up requests behind it.  A request whose output does not
//   fit stops admission until memory frees (so large requests are not
//   starved).
// * Continuous batching: whenever one of `slots` HIP streams is free, every
//   admissible request goes into one batch launch on that stream (the
//   grouped kernels of frsz2_batch_launch: one per class present; up to 8
//   descriptors travel in the kernel arguments, more are copied from pinned
//   host memory).  Up to
//   `slots` launches run on the device at once; requests that arrive while
//   they run are admitted into the next launch as slots and memory free.
// * Large requests (at least `large_bytes` of HBM traffic) are batched only
//   with other large requests: a launch completes as a whole, so a small
//   request batched with a large one would wait for all of the large one's
//   bytes.  Launches of large requests hold at most slots - 1 slots at once,
//   so a small request always finds a slot that no large request holds.
// * Completion: the scheduler polls each slot's event; requests complete,
//   their outputs stay resident until released (or right away when the
//   request asked to discard them), and a stream stays resident until the
//   last request reading it has completed.
//
// Each request computes exactly what the single-array entry point computes
// (the batch kernel runs the same tile code), so a result is the same alone
// and in any mix.  No counterpart in the reference.
#include "batch_internal.hpp"
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace frsz2 {
namespace {

using clk = std::chrono::steady_clock;

constexpr std::uint64_t POOL_ALIGN = 256;

// first-fit allocator of byte ranges of the device pool
class pool_allocator
{
public:
  void init(std::uint64_t size)
  {
    free_.clear();
    sizes_.clear();
    free_[0] = size;
    used_ = peak_ = 0;
  }

  bool alloc(std::uint64_t bytes, std::uint64_t* off)
  {
    bytes = (std::max<std::uint64_t>(bytes, 1) + POOL_ALIGN - 1) / POOL_ALIGN * POOL_ALIGN;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const std::uint64_t rest = it->second - bytes;
        const std::uint64_t next = it->first + bytes;
        free_.erase(it);
        if (rest) {
          free_[next] = rest;
        }
        sizes_[*off] = bytes;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }

  void release(std::uint64_t off)
  {
    const auto s = sizes_.find(off);
    if (s == sizes_.end()) {
      return;
    }
    const std::uint64_t bytes = s->second;
    sizes_.erase(s);
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

  std::uint64_t used() const { return used_; }
  std::uint64_t peak() const { return peak_; }

private:
  std::map<std::uint64_t, std::uint64_t> free_;
  std::unordered_map<std::uint64_t, std::uint64_t> sizes_;
  std::uint64_t used_ = 0;
  std::uint64_t peak_ = 0;
};

enum rstate
{
  QUEUED = 0,
  RUNNING = 1,
  DONE = 2,
  FAILED = 3
};

struct record
{
  frsz2_request r{};
  std::uint64_t source = 0; // ticket of the stream read (decompress, gather)
  int state = QUEUED;
  int status = FRSZ2_OK;
  double t_arrival = 0, t_admit = 0, t_launch = 0, t_done = 0;
  std::uint64_t out_off = 0;
  std::uint64_t out_bytes = 0;
  bool out_live = false;
  int refs = 0;                      // client hold + readers not yet completed
  bool client_hold = false;
  bool waited = false;
  frsz2_result* sink = nullptr;      // replay: where the result goes
};

struct slot
{
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bjob* h_desc = nullptr; // pinned
  bjob* d_desc = nullptr;
  std::vector<std::uint64_t> tickets;
  std::vector<frsz2_job> jobs;
  std::vector<bjob> desc;
  std::vector<char> tiled;
  bool busy = false;
  bool large = false; // the launch holds one large request
};

std::uint64_t
output_bytes(const frsz2_request& r)
{
  const std::uint64_t vs = r.dtype == FRSZ2_FLOAT32 ? 4 : 8;
  switch (r.op) {
    case FRSZ2_OP_COMPRESS:
      return layout::make(r.bits, r.block, r.n).total_bytes();
    case FRSZ2_OP_DECOMPRESS:
      return r.n * vs;
    default:
      return r.m * vs;
  }
}

} // namespace
} // namespace frsz2

struct frsz2_server
{
  frsz2_server_config cfg{};
  std::uint8_t* pool = nullptr;
  frsz2::pool_allocator alloc;
  std::vector<frsz2::slot> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::list<std::uint64_t> queue;
  std::unordered_map<std::uint64_t, frsz2::record> recs;
  std::uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  frsz2::clk::time_point epoch;
  // statistics
  std::uint64_t launches = 0;
  std::uint64_t launched_requests = 0;
  std::uint64_t completed = 0;
  std::uint64_t max_queue = 0;
  int replay_pending = 0;
  frsz2::clk::time_point last_progress;
};

namespace frsz2 {
namespace {

double
seconds_since(const frsz2_server* S, clk::time_point t)
{
  return std::chrono::duration<double>(t - S->epoch).count();
}

void
free_output(frsz2_server* S, record& q)
{
  if (q.out_live) {
    S->alloc.release(q.out_off);
    q.out_live = false;
  }
}

// drop a record once nothing refers to it any more
void
maybe_erase(frsz2_server* S, std::uint64_t ticket)
{
  auto it = S->recs.find(ticket);
  if (it == S->recs.end()) {
    return;
  }
  record& q = it->second;
  if ((q.state == DONE || q.state == FAILED) && !q.out_live && q.refs <= 0 && (q.waited || q.sink)) {
    S->recs.erase(it);
  }
}

void
unref(frsz2_server* S, std::uint64_t ticket)
{
  auto it = S->recs.find(ticket);
  if (it == S->recs.end()) {
    return;
  }
  record& q = it->second;
  if (--q.refs <= 0 && (q.state == DONE || q.state == FAILED)) {
    free_output(S, q);
    maybe_erase(S, ticket);
  }
}

void
fill_result(frsz2_server* S, const record& q, frsz2_result* res)
{
  res->status = q.status;
  res->t_arrival = q.t_arrival;
  res->t_admit = q.t_admit;
  res->t_launch = q.t_launch;
  res->t_done = q.t_done;
  res->out = q.out_live ? S->pool + q.out_off : nullptr;
  res->out_bytes = q.out_live ? q.out_bytes : 0;
}

void
finish(frsz2_server* S, std::uint64_t ticket, int status, double now)
{
  auto it = S->recs.find(ticket);
  if (it == S->recs.end()) {
    return;
  }
  record& q = it->second;
  q.state = status == FRSZ2_OK ? DONE : FAILED;
  q.status = status;
  q.t_done = now;
  ++S->completed;
  if (q.r.op != FRSZ2_OP_COMPRESS && q.source) {
    unref(S, q.source);
  }
  if (q.state == FAILED || q.refs <= 0) {
    free_output(S, q);
  }
  if (q.sink) {
    fill_result(S, q, q.sink);
    --S->replay_pending;
  }
  maybe_erase(S, ticket);
}

void
complete_slot(frsz2_server* S, slot& sl, int status)
{
  const double now = seconds_since(S, clk::now());
  for (const std::uint64_t t : sl.tickets) {
    finish(S, t, status, now);
  }
  sl.tickets.clear();
  sl.busy = false;
}

// HBM bytes a request moves: its values and its stream (a gather: index,
// value and the words it reads per element)
std::uint64_t
traffic_bytes(const frsz2_request& r, std::uint64_t out_bytes)
{
  const std::uint64_t vs = r.dtype == FRSZ2_FLOAT32 ? 4 : 8;
  if (r.op == FRSZ2_OP_GATHER) {
    return r.m * (8 + 3 * vs);
  }
  const std::uint64_t stream = r.op == FRSZ2_OP_COMPRESS ? out_bytes : layout::make(r.bits, r.block, r.n).total_bytes();
  return r.n * vs + stream;
}

// admits queued requests into sl (called with the lock held); returns true
// if anything was admitted.  A launch holds only small or only large
// requests, whichever its first admitted request is; allow_large: the slot
// may take large requests.
bool
admit(frsz2_server* S, slot& sl, bool allow_large)
{
  sl.jobs.clear();
  sl.tickets.clear();
  sl.large = false;
  const double now = seconds_since(S, clk::now());
  for (auto it = S->queue.begin(); it != S->queue.end();) {
    if (static_cast<int>(sl.jobs.size()) >= S->cfg.max_batch) {
      break;
    }
    record& q = S->recs[*it];
    const bool large = traffic_bytes(q.r, q.out_bytes) >= S->cfg.large_bytes;
    if (large ? !allow_large || (!sl.jobs.empty() && !sl.large) : sl.large) {
      ++it; // the other kind of launch takes it
      continue;
    }
    const record* src = nullptr;
    if (q.r.op != FRSZ2_OP_COMPRESS) {
      auto s = S->recs.find(q.source);
      if (s == S->recs.end() || s->second.state == FAILED || (s->second.state != QUEUED && !s->second.out_live)) {
        // the stream it reads is gone
        const std::uint64_t t = *it;
        it = S->queue.erase(it);
        q.state = FAILED;
        finish(S, t, FRSZ2_EARG, now);
        continue;
      }
      if (s->second.state != DONE) {
        ++it; // stream not written yet (its compress still queued, as a large
              // request waiting for a slot, or running)
        continue;
      }
      src = &s->second;
    }
    std::uint64_t off = 0;
    if (!S->alloc.alloc(q.out_bytes, &off)) {
      break; // wait for memory, keep arrival order
    }
    q.out_off = off;
    q.out_live = true;
    q.state = RUNNING;
    q.t_admit = now;
    frsz2_job j{};
    j.op = q.r.op;
    j.dtype = q.r.dtype;
    j.bits = q.r.bits;
    j.block = q.r.block;
    j.n = q.r.n;
    if (q.r.op == FRSZ2_OP_COMPRESS) {
      j.in = q.r.values;
      j.out = S->pool + off;
      j.out_bytes = q.out_bytes;
    } else {
      j.in = S->pool + src->out_off;
      j.in_bytes = src->out_bytes;
      j.out = S->pool + off;
      j.out_bytes = q.out_bytes;
      j.idx = q.r.idx;
      j.m = q.r.m;
    }
    sl.jobs.push_back(j);
    sl.tickets.push_back(*it);
    sl.large = large;
    it = S->queue.erase(it);
  }
  return !sl.jobs.empty();
}

// enqueue the slot's batch on its stream (lock not held)
int
launch_slot(frsz2_server* S, slot& sl)
{
  sl.desc.clear();
  const std::uint64_t tiles = build_descriptors(sl.jobs.data(), sl.jobs.size(), sl.desc, sl.tiled);
  const int nd = static_cast<int>(sl.desc.size());
  if (nd > 0) {
    const batch_schedule sc{ S->cfg.schedule, S->cfg.workgroups, S->cfg.xcd_order };
    hipError_t e = hipSuccess;
    if (!descriptors_inline(sc, nd, tiles)) {
      std::copy(sl.desc.begin(), sl.desc.end(), sl.h_desc);
      e = hipMemcpyAsync(sl.d_desc, sl.h_desc, nd * sizeof(bjob), hipMemcpyHostToDevice, sl.stream);
      if (e != hipSuccess) {
        return hip_status(e, "hipMemcpyAsync");
      }
    }
    e = launch_batch(sl.d_desc, sl.desc.data(), nd, tiles, sc, sl.stream);
    if (e != hipSuccess) {
      return hip_status(e, "batch launch");
    }
  }
  for (std::size_t i = 0; i < sl.jobs.size(); ++i) {
    if (!sl.tiled[i]) {
      const int rc = run_plain(sl.jobs[i], sl.stream);
      if (rc) {
        return rc;
      }
    }
  }
  return hip_status(hipEventRecord(sl.done, sl.stream), "hipEventRecord");
}

void
scheduler(frsz2_server* S)
{
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    bool any_busy = false;
    bool progressed = false;
    for (slot& sl : S->slots) {
      if (!sl.busy) {
        continue;
      }
      const hipError_t e = hipEventQuery(sl.done);
      if (e == hipSuccess) {
        complete_slot(S, sl, FRSZ2_OK);
        progressed = true;
      } else if (e != hipErrorNotReady) {
        set_error(FRSZ2_EHIP, std::string("batch failed: ") + hipGetErrorString(e));
        complete_slot(S, sl, FRSZ2_EHIP);
        progressed = true;
      } else {
        any_busy = true;
      }
    }
    if (progressed) {
      S->cv.notify_all();
    }
    if (!S->queue.empty()) {
      int large_busy = 0;
      for (const slot& sl : S->slots) {
        large_busy += sl.busy && sl.large ? 1 : 0;
      }
      const int large_cap = S->cfg.slots > 1 ? S->cfg.slots - 1 : 1;
      for (slot& sl : S->slots) {
        if (sl.busy || S->queue.empty()) {
          continue;
        }
        if (!admit(S, sl, large_busy < large_cap)) {
          break;
        }
        large_busy += sl.large ? 1 : 0;
        sl.busy = true;
        any_busy = true;
        ++S->launches;
        S->launched_requests += sl.jobs.size();
        lk.unlock();
        const int rc = launch_slot(S, sl);
        const double t = seconds_since(S, clk::now());
        lk.lock();
        for (const std::uint64_t tk : sl.tickets) {
          auto it = S->recs.find(tk);
          if (it != S->recs.end()) {
            it->second.t_launch = t;
          }
        }
        if (rc != FRSZ2_OK) {
          (void)hipStreamSynchronize(sl.stream);
          complete_slot(S, sl, rc);
          S->cv.notify_all();
        }
        progressed = true;
      }
    }
    if (progressed) {
      S->last_progress = clk::now();
    } else if (!any_busy && !S->queue.empty() && S->cfg.stall_ms > 0 &&
               clk::now() - S->last_progress > std::chrono::milliseconds(S->cfg.stall_ms)) {
      // nothing runs and nothing has completed for stall_ms, and nothing can
      // be admitted (the pool is held by outputs that are not being
      // released): every queued request that has itself waited stall_ms
      // fails at once; younger ones keep waiting until they have
      const double now = seconds_since(S, clk::now());
      const double limit = 1e-3 * S->cfg.stall_ms;
      bool failed = false;
      for (auto it = S->queue.begin(); it != S->queue.end();) {
        const std::uint64_t t = *it;
        if (now - S->recs[t].t_arrival < limit) {
          ++it;
          continue;
        }
        it = S->queue.erase(it);
        finish(S, t, FRSZ2_ECAPACITY, now);
        failed = true;
      }
      if (failed) {
        set_error(FRSZ2_ECAPACITY, "request could not be admitted: HBM pool held by unreleased outputs");
        S->cv.notify_all();
        continue;
      }
    }
    if (S->stop && S->queue.empty() && !any_busy) {
      break;
    }
    if (!progressed) {
      if (any_busy) {
        // work on the device: poll again soon
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      } else {
        S->cv.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
}

// extra_refs: holds on the output besides the client's (replay: the trace's
// readers of a compress); count_reader: a decompress / gather takes a hold on
// its stream (false in replay, where the stream's readers were counted up
// front)
int
submit_locked(frsz2_server* S, const frsz2_request& r, std::uint64_t source, int extra_refs, bool count_reader,
              frsz2_result* sink, double arrival, std::uint64_t* ticket)
{
  const std::uint64_t t = S->next_ticket++;
  record& q = S->recs[t];
  q.r = r;
  q.t_arrival = arrival;
  q.sink = sink;
  q.source = source;
  q.client_hold = !(r.flags & FRSZ2_REQ_DISCARD);
  q.refs = (q.client_hold ? 1 : 0) + extra_refs;
  q.out_bytes = output_bytes(r);
  if (ticket) {
    *ticket = t;
  }
  // validation (the same checks as the single-array calls)
  int rc = check_params(r.dtype, r.bits, r.block);
  if (rc == FRSZ2_OK && r.op != FRSZ2_OP_COMPRESS && r.op != FRSZ2_OP_DECOMPRESS && r.op != FRSZ2_OP_GATHER) {
    rc = set_error(FRSZ2_EARG, "unknown request op");
  }
  if (rc == FRSZ2_OK && r.op == FRSZ2_OP_COMPRESS && r.n > 0 &&
      (!r.values || (reinterpret_cast<std::uintptr_t>(r.values) & (r.dtype == FRSZ2_FLOAT32 ? 3u : 7u)))) {
    rc = set_error(FRSZ2_EARG, "null or misaligned input");
  }
  if (rc == FRSZ2_OK && r.op == FRSZ2_OP_GATHER && r.m > 0 &&
      (!r.idx || (reinterpret_cast<std::uintptr_t>(r.idx) & 7u))) {
    rc = set_error(FRSZ2_EARG, "null or misaligned indices");
  }
  if (rc == FRSZ2_OK && q.out_bytes > S->cfg.hbm_budget) {
    rc = set_error(FRSZ2_ECAPACITY, "request output exceeds the HBM budget");
  }
  if (rc == FRSZ2_OK && r.op != FRSZ2_OP_COMPRESS) {
    auto s = S->recs.find(source);
    if (s == S->recs.end() || s->second.r.op != FRSZ2_OP_COMPRESS || s->second.state == FAILED ||
        (s->second.state == DONE && !s->second.out_live) || s->second.r.dtype != r.dtype ||
        s->second.r.bits != r.bits || s->second.r.block != r.block || s->second.r.n != r.n) {
      rc = set_error(FRSZ2_EARG, "source is not a resident stream with the same dtype, bits, block and n");
    } else if (count_reader) {
      ++s->second.refs;
    }
  }
  if (rc != FRSZ2_OK) {
    if (r.op != FRSZ2_OP_COMPRESS) {
      if (!count_reader && source) {
        unref(S, source); // the hold counted up front for this reader
      }
      q.source = 0;
    }
    q.client_hold = false;
    q.refs = 0;
    q.waited = true; // the caller gets the error now and will not wait
    finish(S, t, rc, arrival);
    return rc;
  }
  S->queue.push_back(t);
  S->max_queue = std::max<std::uint64_t>(S->max_queue, S->queue.size());
  return FRSZ2_OK;
}

void
destroy_server(frsz2_server* S)
{
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) {
    S->worker.join();
  }
  for (slot& sl : S->slots) {
    if (sl.stream) {
      (void)hipStreamSynchronize(sl.stream);
      (void)hipStreamDestroy(sl.stream);
    }
    if (sl.done) {
      (void)hipEventDestroy(sl.done);
    }
    if (sl.h_desc) {
      (void)hipHostFree(sl.h_desc);
    }
    if (sl.d_desc) {
      (void)hipFree(sl.d_desc);
    }
  }
  if (S->pool) {
    (void)hipFree(S->pool);
  }
  delete S;
}

} // namespace
} // namespace frsz2

extern "C" {

int
frsz2_server_create(const frsz2_server_config* cfg, frsz2_server** out)
{
  using namespace frsz2;
  clear_error();
  if (!out) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  *out = nullptr;
  auto* S = new frsz2_server();
  if (cfg) {
    S->cfg = *cfg;
  }
  if (S->cfg.hbm_budget == 0) {
    S->cfg.hbm_budget = std::uint64_t{ 64 } << 30;
  }
  if (S->cfg.slots <= 0) {
    S->cfg.slots = 4;
  }
  if (S->cfg.max_batch <= 0) {
    S->cfg.max_batch = 1024;
  }
  if (S->cfg.schedule != FRSZ2_SCHEDULE_PERSISTENT) {
    S->cfg.schedule = FRSZ2_SCHEDULE_TILES;
  }
  if (S->cfg.workgroups < 0) {
    S->cfg.workgroups = 0;
  }
  S->cfg.xcd_order = S->cfg.xcd_order ? 1 : 0;
  if (S->cfg.stall_ms == 0) {
    S->cfg.stall_ms = 10000;
  }
  if (S->cfg.large_bytes == 0) {
    S->cfg.large_bytes = std::uint64_t{ 32 } << 20;
  }
  hipError_t e = hipMalloc(&S->pool, S->cfg.hbm_budget);
  if (e != hipSuccess) {
    S->pool = nullptr;
    destroy_server(S);
    return hip_status(e, "hipMalloc (HBM pool)");
  }
  S->alloc.init(S->cfg.hbm_budget);
  S->slots.resize(static_cast<std::size_t>(S->cfg.slots));
  for (slot& sl : S->slots) {
    if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_desc), S->cfg.max_batch * sizeof(bjob),
                           hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&sl.d_desc, S->cfg.max_batch * sizeof(bjob))) != hipSuccess) {
      destroy_server(S);
      return hip_status(e, "server setup");
    }
  }
  S->epoch = clk::now();
  S->last_progress = S->epoch;
  S->worker = std::thread(scheduler, S);
  *out = S;
  return FRSZ2_OK;
}

int
frsz2_server_submit(frsz2_server* S, const frsz2_request* req, uint64_t* ticket)
{
  using namespace frsz2;
  clear_error();
  if (!S || !req || !ticket) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  int rc;
  {
    std::lock_guard<std::mutex> g(S->mu);
    rc = submit_locked(S, *req, req->source, 0, true, nullptr, seconds_since(S, clk::now()), ticket);
  }
  S->cv.notify_all();
  return rc;
}

int
frsz2_server_wait(frsz2_server* S, uint64_t ticket, frsz2_result* res)
{
  using namespace frsz2;
  clear_error();
  if (!S || !res) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    auto it = S->recs.find(ticket);
    if (it == S->recs.end()) {
      return set_error(FRSZ2_EARG, "unknown ticket");
    }
    record& q = it->second;
    if (q.state == DONE || q.state == FAILED) {
      fill_result(S, q, res);
      q.waited = true;
      const int st = q.status;
      maybe_erase(S, ticket);
      return st == FRSZ2_OK ? FRSZ2_OK : set_error(st, "request failed");
    }
    S->cv.wait(lk);
  }
}

int
frsz2_server_release(frsz2_server* S, uint64_t ticket)
{
  using namespace frsz2;
  clear_error();
  if (!S) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->recs.find(ticket);
  if (it == S->recs.end() || !it->second.client_hold) {
    return set_error(FRSZ2_EARG, "ticket holds no output");
  }
  it->second.client_hold = false;
  it->second.waited = true;
  unref(S, ticket);
  S->cv.notify_all(); // memory may have freed for a waiting request
  return FRSZ2_OK;
}

int
frsz2_server_replay(frsz2_server* S, const frsz2_request* reqs, size_t count, frsz2_result* results,
                    uint64_t* tickets)
{
  using namespace frsz2;
  clear_error();
  if (!S || (count > 0 && (!reqs || !results))) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  // streams stay resident until their last reader in the trace completes
  std::vector<int> readers(count, 0);
  for (size_t i = 0; i < count; ++i) {
    if (reqs[i].op != FRSZ2_OP_COMPRESS) {
      const uint64_t s = reqs[i].source;
      if (s >= i || reqs[s].op != FRSZ2_OP_COMPRESS) {
        return set_error(FRSZ2_EARG, "request " + std::to_string(i) + ": source must be an earlier compress");
      }
      ++readers[s];
    }
  }
  std::vector<uint64_t> tk(count, 0);
  const auto t0 = clk::now();
  const double base = seconds_since(S, t0);
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->replay_pending += static_cast<int>(count);
  }
  for (size_t i = 0; i < count; ++i) {
    const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(reqs[i].arrival));
    for (;;) {
      const auto now = clk::now();
      if (now >= due) {
        break;
      }
      if (due - now > std::chrono::microseconds(200)) {
        std::this_thread::sleep_for(std::chrono::microseconds(50));
      } else {
        std::this_thread::yield();
      }
    }
    {
      std::lock_guard<std::mutex> g(S->mu);
      const uint64_t src = reqs[i].op == FRSZ2_OP_COMPRESS ? 0 : tk[reqs[i].source];
      (void)submit_locked(S, reqs[i], src, readers[i], false, &results[i], base + reqs[i].arrival, &tk[i]);
    }
    S->cv.notify_all();
  }
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv.wait(lk, [&] { return S->replay_pending <= 0; });
  for (size_t i = 0; i < count; ++i) {
    results[i].t_arrival -= base;
    results[i].t_admit -= base;
    results[i].t_launch -= base;
    results[i].t_done -= base;
    if (tickets) {
      tickets[i] = (reqs[i].flags & FRSZ2_REQ_DISCARD) ? 0 : tk[i];
    }
  }
  return FRSZ2_OK;
}

int
frsz2_server_get_stats(frsz2_server* S, frsz2_server_stats* st)
{
  frsz2::clear_error();
  if (!S || !st) {
    return frsz2::set_error(FRSZ2_EARG, "null pointer");
  }
  std::lock_guard<std::mutex> g(S->mu);
  st->launches = S->launches;
  st->launched_requests = S->launched_requests;
  st->completed = S->completed;
  st->pool_used = S->alloc.used();
  st->pool_peak = S->alloc.peak();
  st->max_queue = S->max_queue;
  return FRSZ2_OK;
}

int
frsz2_server_destroy(frsz2_server* S)
{
  frsz2::clear_error();
  if (S) {
    frsz2::destroy_server(S);
  }
  return FRSZ2_OK;
}

} // extern "C"
