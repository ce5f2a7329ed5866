// serve.cpp -- request server: compress / decompress requests of mixed sizes
// are admitted as memory frees up and batched, continuously, into the
// batched launches (batch.hip).
//
// A worker thread owns two launch slots.  Each slot has its own stream and
// workspace and two preallocated batches (encode, decode) with pinned
// descriptor staging.  Whenever a slot is free and requests are waiting, the
// worker admits waiting requests in arrival order while their outputs fit
// the device-memory arena and the slot's batch limits, uploads one batch plan
// per kind, launches the batched encode and the batched decode and records an
// event.  Requests that arrive meanwhile form the next batch on the other
// slot, so consecutive batches overlap on the device.  Each request's output
// (frame or elements) lives in the arena until the client releases it.
//
// The codec arithmetic is unchanged: a compressed request yields the frame
// mans_frame_encode writes for it alone (byte planes, 64 lanes), a
// decompressed request the elements mans_frame_decode returns.
#include <hip/hip_runtime.h>

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

#include "../../include/mans.h"

// batch slots (capi.hip)
struct mans_batch;
mans_status batch_slot_create(mans_batch** out, const mans_params* p, bool decode, uint32_t cap_bufs,
                              uint64_t cap_bytes, mans_workspace* ws);
mans_status batch_slot_plan(mans_batch* b, const mans_params* p, const mans_buffer* bufs, int32_t count,
                            hipStream_t s);
const uint32_t* batch_slot_bad(const mans_batch* b);
uint32_t batch_decode_error_bits();
const uint32_t* workspace_error_word(const mans_workspace* ws);
mans_status workspace_clear_error(mans_workspace* ws, hipStream_t s);

namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kSlotBufs = 4096;             // requests per batch and kind
constexpr uint64_t kSlotBytes = 16ull << 30;     // uncompressed bytes per batch and kind
constexpr uint64_t kGranule = 256;

uint64_t round256(uint64_t v) { return (v + kGranule - 1) / kGranule * kGranule; }

// First-fit allocator over one device arena (offsets in bytes).
class Arena {
 public:
  void init(uint64_t bytes) {
    free_.clear();
    if (bytes) free_[0] = bytes;
    total_ = bytes;
    used_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t& off) {
    bytes = round256(std::max<uint64_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      off = it->first;
      const uint64_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = round256(std::max<uint64_t>(bytes, 1));
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
  uint64_t total() const { return total_; }
  uint64_t used() const { return used_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t total_ = 0, used_ = 0;
};

struct Req {
  uint64_t id = 0;
  mans_request r{};
  uint64_t out_off = 0, out_cap = 0, out_size = 0;
  int32_t status = MANS_OK;
  int state = 0;  // 0 waiting, 1 in flight, 2 done
  Clock::time_point t_submit, t_launch, t_done;
};

struct Slot {
  hipStream_t stream = nullptr;
  mans_workspace* ws = nullptr;
  mans_batch* enc = nullptr;
  mans_batch* dec = nullptr;
  uint64_t* d_sizes = nullptr;
  uint64_t* h_sizes = nullptr;  // pinned
  uint32_t* h_bad = nullptr;    // pinned: per-request error flags of the decode batch
  uint32_t* h_err = nullptr;    // pinned: the slot workspace's error word after the launch
  hipEvent_t done = nullptr;
  std::vector<Req*> enc_reqs, dec_reqs;
  bool busy = false;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct mans_server {
  mans_params p{};
  int device = 0;
  uint8_t* arena_base = nullptr;
  Arena arena;
  Slot slot[2];
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> waiting;
  std::unordered_map<uint64_t, Req*> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  uint64_t batches = 0;
};

namespace {

uint64_t out_bytes_for(const mans_params& p, const mans_request& r) {
  if (r.kind == MANS_REQ_COMPRESS) {
    mans_params q = p;
    q.elem_bytes = r.elem_bytes;
    q.mode = MANS_MODE_PLANES;
    return mans_frame_max_size(r.n_elems, &q);
  }
  return r.n_elems * (uint64_t)r.elem_bytes;
}

// Launches the requests admitted into `sl` (called without the lock; the
// slot's requests are not touched by other threads while it is busy).
void launch_slot(mans_server* S, Slot& sl) {
  std::vector<mans_buffer> eb, db;
  for (Req* q : sl.enc_reqs) {
    mans_buffer m{};
    m.d_in = q->r.d_in;
    m.d_out = S->arena_base + q->out_off;
    m.out_capacity = q->out_cap;
    m.n_elems = q->r.n_elems;
    m.elem_bytes = q->r.elem_bytes;
    eb.push_back(m);
  }
  for (Req* q : sl.dec_reqs) {
    mans_buffer m{};
    m.d_in = q->r.d_in;
    m.in_bytes = q->r.in_bytes;
    m.d_out = S->arena_base + q->out_off;
    m.out_capacity = q->out_cap;
    m.n_elems = q->r.n_elems;
    m.elem_bytes = q->r.elem_bytes;
    db.push_back(m);
  }
  mans_status st = MANS_OK;
  if (!eb.empty()) {
    st = batch_slot_plan(sl.enc, &S->p, eb.data(), (int32_t)eb.size(), sl.stream);
    if (st == MANS_OK) st = mans_batch_encode_run(sl.enc, sl.d_sizes, sl.stream);
    if (st == MANS_OK &&
        hipMemcpyAsync(sl.h_sizes, sl.d_sizes, eb.size() * 8, hipMemcpyDeviceToHost, sl.stream) != hipSuccess)
      st = MANS_ERR_HIP;
    if (st != MANS_OK)
      for (Req* q : sl.enc_reqs) q->status = st;
  }
  st = MANS_OK;
  if (!db.empty()) {
    st = batch_slot_plan(sl.dec, &S->p, db.data(), (int32_t)db.size(), sl.stream);
    if (st == MANS_OK) st = mans_batch_decode_run(sl.dec, sl.stream);
    if (st == MANS_OK && hipMemcpyAsync(sl.h_bad, batch_slot_bad(sl.dec), db.size() * 4, hipMemcpyDeviceToHost,
                                        sl.stream) != hipSuccess)
      st = MANS_ERR_HIP;
    if (st != MANS_OK)
      for (Req* q : sl.dec_reqs) q->status = st;
  }
  // the error word comes back with the results: finishing the slot needs no
  // further copy or synchronisation
  if (hipMemcpyAsync(sl.h_err, workspace_error_word(sl.ws), 4, hipMemcpyDeviceToHost, sl.stream) != hipSuccess) {
    for (Req* q : sl.enc_reqs) q->status = MANS_ERR_HIP;
    for (Req* q : sl.dec_reqs) q->status = MANS_ERR_HIP;
  }
  (void)hipEventRecord(sl.done, sl.stream);
}

// Marks the requests of a finished slot done (called with the lock held).
// A malformed frame fails only its own decompress request (the decode
// batch's per-buffer flags); the slot's sticky error word fails the
// compress requests only for bits an encode launch raises.
void finish_slot(mans_server* S, Slot& sl) {
  const uint32_t flags = *sl.h_err;
  // reset before the slot's next launch (same stream)
  // (were the reset to fail, the next launch's compress requests would fail
  // too: never a wrong success)
  if (flags != 0) (void)workspace_clear_error(sl.ws, sl.stream);
  const bool enc_bad = (flags & ~batch_decode_error_bits()) != 0;
  const auto now = Clock::now();
  for (size_t i = 0; i < sl.enc_reqs.size(); ++i) {
    Req* q = sl.enc_reqs[i];
    q->out_size = sl.h_sizes[i];
    if (q->status == MANS_OK && enc_bad) q->status = MANS_ERR_CORRUPT;
    q->t_done = now;
    q->state = 2;
  }
  for (size_t i = 0; i < sl.dec_reqs.size(); ++i) {
    Req* q = sl.dec_reqs[i];
    q->out_size = q->r.n_elems * (uint64_t)q->r.elem_bytes;
    if (q->status == MANS_OK && sl.h_bad[i] != 0) q->status = MANS_ERR_CORRUPT;
    q->t_done = now;
    q->state = 2;
  }
  sl.enc_reqs.clear();
  sl.dec_reqs.clear();
  sl.busy = false;
}

// Admits waiting requests into `sl` in arrival order (lock held).  Stops at
// the first request whose output does not fit the arena (it waits for
// memory, later requests wait behind it) or that would overflow the batch.
bool admit(mans_server* S, Slot& sl) {
  uint64_t enc_bytes = 0, dec_bytes = 0;
  const auto now = Clock::now();
  while (!S->waiting.empty()) {
    Req* q = S->waiting.front();
    const bool enc = q->r.kind == MANS_REQ_COMPRESS;
    const uint64_t in = q->r.n_elems * (uint64_t)q->r.elem_bytes;
    std::vector<Req*>& list = enc ? sl.enc_reqs : sl.dec_reqs;
    uint64_t& bytes = enc ? enc_bytes : dec_bytes;
    if (list.size() >= kSlotBufs || (!list.empty() && bytes + in > kSlotBytes)) break;
    uint64_t off = 0;
    if (!S->arena.alloc(q->out_cap, off)) break;
    q->out_off = off;
    q->state = 1;
    q->t_launch = now;
    list.push_back(q);
    bytes += in;
    S->waiting.pop_front();
  }
  return !sl.enc_reqs.empty() || !sl.dec_reqs.empty();
}

void worker_loop(mans_server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    bool progress = false;
    for (Slot& sl : S->slot) {
      if (sl.busy && hipEventQuery(sl.done) == hipSuccess) {
        finish_slot(S, sl);
        progress = true;
      }
    }
    if (progress) S->cv_done.notify_all();
    for (Slot& sl : S->slot) {
      if (sl.busy || S->waiting.empty()) continue;
      if (!admit(S, sl)) continue;
      sl.busy = true;
      ++S->batches;
      lk.unlock();
      launch_slot(S, sl);
      lk.lock();
      progress = true;
    }
    const bool any_busy = S->slot[0].busy || S->slot[1].busy;
    if (S->stop && S->waiting.empty() && !any_busy) break;
    if (!progress) {
      if (any_busy) {
        // poll the in-flight batches without holding the lock
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        S->cv_work.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
  }
}

}  // namespace

extern "C" {

mans_status mans_server_create(mans_server** out, const mans_params* p, uint64_t memory_budget, int32_t device) {
  if (!out || !p || p->lanes != 64) return MANS_ERR_INVALID_ARG;
  *out = nullptr;
  if (p->prob_bits < 9 || p->prob_bits > 11) return MANS_ERR_PRECISION;
  if (hipSetDevice(device) != hipSuccess) return MANS_ERR_NO_DEVICE;
  mans_server* S = new mans_server();
  S->p = *p;
  S->device = device;
  size_t free0 = 0, total0 = 0;
  (void)hipMemGetInfo(&free0, &total0);
  mans_status st = MANS_OK;
  for (Slot& sl : S->slot) {
    if (st != MANS_OK) break;
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      st = MANS_ERR_HIP;
      break;
    }
    st = mans_workspace_create(&sl.ws, device);
    if (st == MANS_OK) st = batch_slot_create(&sl.enc, p, false, kSlotBufs, kSlotBytes, sl.ws);
    if (st == MANS_OK) st = batch_slot_create(&sl.dec, p, true, kSlotBufs, kSlotBytes, sl.ws);
    if (st == MANS_OK && (hipMalloc(&sl.d_sizes, kSlotBufs * 8) != hipSuccess ||
                          hipHostMalloc((void**)&sl.h_sizes, kSlotBufs * 8, hipHostMallocDefault) != hipSuccess ||
                          hipHostMalloc((void**)&sl.h_bad, kSlotBufs * 4, hipHostMallocDefault) != hipSuccess ||
                          hipHostMalloc((void**)&sl.h_err, 4, hipHostMallocDefault) != hipSuccess))
      st = MANS_ERR_HIP;
  }
  if (st == MANS_OK) {
    // the budget covers the arena and the server's own state
    size_t free1 = 0, total1 = 0;
    (void)hipMemGetInfo(&free1, &total1);
    const uint64_t own = free0 > free1 ? (uint64_t)(free0 - free1) : 0;
    if (memory_budget <= own + (64ull << 20)) {
      st = MANS_ERR_BUDGET;
    } else {
      const uint64_t arena = (memory_budget - own) / kGranule * kGranule;
      if (hipMalloc(&S->arena_base, arena) != hipSuccess) st = MANS_ERR_BUDGET;
      else S->arena.init(arena);
    }
  }
  if (st != MANS_OK) {
    mans_server_destroy(S);
    return st;
  }
  S->worker = std::thread(worker_loop, S);
  *out = S;
  return MANS_OK;
}

// Checks one request and builds its record (not yet queued).
static mans_status make_req(mans_server* S, const mans_request* r, Req** out) {
  if (r->kind != MANS_REQ_COMPRESS && r->kind != MANS_REQ_DECOMPRESS) return MANS_ERR_INVALID_ARG;
  if (r->elem_bytes != 1 && r->elem_bytes != 2 && r->elem_bytes != 4) return MANS_ERR_INVALID_ARG;
  if ((r->n_elems > 0 || r->kind == MANS_REQ_DECOMPRESS) && !r->d_in) return MANS_ERR_INVALID_ARG;
  if (((uintptr_t)r->d_in & 15u) != 0) return MANS_ERR_ALIGNMENT;
  if (r->n_elems > 0xffffffffull) return MANS_ERR_TOO_LARGE;
  const uint64_t cap = out_bytes_for(S->p, *r);
  if (cap > S->arena.total() || r->n_elems * (uint64_t)r->elem_bytes > kSlotBytes) return MANS_ERR_BUDGET;
  Req* q = new Req();
  q->r = *r;
  q->out_cap = cap;
  *out = q;
  return MANS_OK;
}

mans_status mans_server_submit(mans_server* S, const mans_request* r, uint64_t* id) {
  if (!S || !r || !id) return MANS_ERR_INVALID_ARG;
  return mans_server_submit_many(S, r, 1, id);
}

mans_status mans_server_submit_many(mans_server* S, const mans_request* r, int32_t count, uint64_t* ids) {
  if (!S || count < 0 || (count > 0 && (!r || !ids))) return MANS_ERR_INVALID_ARG;
  std::vector<Req*> qs((size_t)count, nullptr);
  for (int32_t i = 0; i < count; ++i) {
    const mans_status st = make_req(S, r + i, &qs[(size_t)i]);
    if (st != MANS_OK) {
      for (Req* q : qs) delete q;
      return st;
    }
  }
  const auto now = Clock::now();
  {
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) {
      for (Req* q : qs) delete q;
      return MANS_ERR_INVALID_ARG;
    }
    for (int32_t i = 0; i < count; ++i) {
      Req* q = qs[(size_t)i];
      q->t_submit = now;
      q->id = S->next_id++;
      S->reqs[q->id] = q;
      S->waiting.push_back(q);
      ids[i] = q->id;
    }
  }
  S->cv_work.notify_one();
  return MANS_OK;
}

mans_status mans_server_wait(mans_server* S, uint64_t id, int32_t timeout_ms, mans_result* res) {
  if (!S || !res) return MANS_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return MANS_ERR_INVALID_ARG;
  Req* q = it->second;
  auto ready = [&] { return q->state == 2; };
  if (timeout_ms < 0) S->cv_done.wait(lk, ready);
  else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return MANS_ERR_BUSY;
  res->status = q->status;
  res->reserved = 0;
  res->d_out = S->arena_base + q->out_off;
  res->out_bytes = q->out_size;
  res->queue_ms = ms_between(q->t_submit, q->t_launch);
  res->latency_ms = ms_between(q->t_submit, q->t_done);
  return MANS_OK;
}

mans_status mans_server_release(mans_server* S, uint64_t id) {
  if (!S) return MANS_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || it->second->state != 2) return MANS_ERR_INVALID_ARG;
    Req* q = it->second;
    S->arena.release(q->out_off, q->out_cap);
    S->reqs.erase(it);
    delete q;
  }
  S->cv_work.notify_one();  // memory freed: waiting requests may fit now
  return MANS_OK;
}

mans_status mans_server_stats(mans_server* S, uint64_t* arena_bytes, uint64_t* arena_used, uint64_t* batches,
                              uint64_t* waiting) {
  if (!S) return MANS_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(S->mu);
  if (arena_bytes) *arena_bytes = S->arena.total();
  if (arena_used) *arena_used = S->arena.used();
  if (batches) *batches = S->batches;
  if (waiting) *waiting = S->waiting.size();
  return MANS_OK;
}

mans_status mans_server_destroy(mans_server* S) {
  if (!S) return MANS_OK;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  (void)hipSetDevice(S->device);
  (void)hipDeviceSynchronize();
  for (Slot& sl : S->slot) {
    if (sl.enc) mans_batch_destroy(sl.enc);
    if (sl.dec) mans_batch_destroy(sl.dec);
    if (sl.ws) mans_workspace_destroy(sl.ws);
    if (sl.d_sizes) (void)hipFree(sl.d_sizes);
    if (sl.h_sizes) (void)hipHostFree(sl.h_sizes);
    if (sl.h_bad) (void)hipHostFree(sl.h_bad);
    if (sl.h_err) (void)hipHostFree(sl.h_err);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  if (S->arena_base) (void)hipFree(S->arena_base);
  for (auto& kv : S->reqs) delete kv.second;
  delete S;
  return MANS_OK;
}

}  // extern "C"
