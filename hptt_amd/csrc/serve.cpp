// serve.cpp -- see serve.h.
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "device.h"

namespace hpttx {

namespace {
constexpr uint64_t kGranule = 256;
uint64_t round_up(uint64_t x, uint64_t g) { return (x + g - 1) / g * g; }
const auto g_t0 = std::chrono::steady_clock::now();
}  // namespace

// ---------------------------------------------------------------------------
bool Arena::init(uint64_t bytes, std::string &err) {
  cap_ = round_up(bytes, kGranule);
  if (!cap_) return true;
  if (hipMalloc(&base_, cap_) != hipSuccess) {
    (void)hipGetLastError();
    err = "arena allocation failed";
    base_ = nullptr;
    cap_ = 0;
    return false;
  }
  free_.clear();
  free_[0] = cap_;
  used_ = 0;
  return true;
}

void Arena::destroy() {
  if (base_) (void)hipFree(base_);
  base_ = nullptr;
  cap_ = used_ = 0;
  free_.clear();
}

bool Arena::alloc(uint64_t bytes, uint64_t &off) {
  bytes = round_up(std::max<uint64_t>(bytes, 1), kGranule);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= bytes) {
      off = it->first;
      const uint64_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      return true;
    }
  }
  return false;
}

void Arena::free(uint64_t off, uint64_t bytes) {
  bytes = round_up(std::max<uint64_t>(bytes, 1), kGranule);
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

// ---------------------------------------------------------------------------
uint64_t settle_requests(std::vector<Request *> &reqs, int status, const std::string &why,
                         double t, ServerStats &st, std::vector<double> &lat,
                         std::deque<Request *> &finished, Arena &arena) {
  uint64_t bytes = 0;
  for (Request *r : reqs) {
    bytes += r->bytes;
    r->t_done = t;
    if (status == kOk) {
      r->state = kDone;
      r->status = kOk;
      st.completed++;
      st.bytes_done += r->bytes;
      lat.push_back(t - r->t_submit);
    } else {
      r->state = kFailed;
      r->status = status;
      r->error = why;
      st.failed++;
      if (r->arena_out) arena.free(r->arena_off, r->arena_len);
      r->arena_out = false;
    }
    finished.push_back(r);
  }
  if (status == kOk && !reqs.empty()) st.last_done = t;
  reqs.clear();
  return bytes;
}

// ---------------------------------------------------------------------------
double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - g_t0).count();
}

int Server::start(const ServerConfig &cfg, std::string &err) {
  cfg_ = cfg;
  if (cfg_.slots < 1) cfg_.slots = 1;
  if (cfg_.slots > 8) cfg_.slots = 8;
  if (cfg_.max_batch < 1) cfg_.max_batch = 1;
  if (hipGetDevice(&device_) != hipSuccess) { err = "no HIP device"; return kEHip; }
  {
    int rc = require_target_device(err);
    if (rc != kOk) return rc;
  }
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) { err = "hipMemGetInfo failed"; return kEHip; }
  budget_ = cfg_.inflight_bytes ? cfg_.inflight_bytes : (uint64_t)fr / 4;
  if (cfg_.arena_bytes) {
    if (cfg_.arena_bytes > (uint64_t)fr) { err = "arena larger than free HBM"; return kENoMem; }
    if (!arena_.init(cfg_.arena_bytes, err)) return kENoMem;
  }
  slots_.resize(cfg_.slots);
  for (Slot &s : slots_) {
    hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    // Blocking-sync events: the completer can wait on one without spinning.
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming | hipEventBlockingSync);
    if (e == hipSuccess) e = hipMalloc(&s.d_ent, sizeof(BatchEntry) * cfg_.max_batch);
    if (e == hipSuccess) e = hipMalloc(&s.d_begins, sizeof(uint32_t) * cfg_.max_batch);
    if (e == hipSuccess) e = hipHostMalloc(&s.h_ent, sizeof(BatchEntry) * cfg_.max_batch, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&s.h_begins, sizeof(uint32_t) * cfg_.max_batch, hipHostMallocDefault);
    for (size_t i = 0; i + 1 < kBatchLanes && e == hipSuccess; ++i) {
      hipStream_t ss = nullptr;
      e = hipStreamCreateWithFlags(&ss, hipStreamNonBlocking);
      if (e == hipSuccess) s.side.push_back(ss);
    }
    for (size_t i = 0; i < kBatchLanes && e == hipSuccess; ++i) {
      hipEvent_t ev = nullptr;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) s.fork.push_back(ev);
    }
    if (e != hipSuccess) {
      err = std::string("server slot setup: ") + hipGetErrorString(e);
      stop();
      return e == hipErrorOutOfMemory ? kENoMem : kEHip;
    }
  }
  stop_ = false;
  dispatcher_ = std::thread([this] { dispatch_loop(); });
  completer_ = std::thread([this] { complete_loop(); });
  return kOk;
}

void Server::stop() {
  {
    std::unique_lock<std::mutex> lk(mu_);
    // From here on submit refuses; requests still queued fail (waiters are
    // woken); launches in flight finish first.
    stop_ = true;
    std::vector<Request *> dropped(pending_.begin(), pending_.end());
    pending_.clear();
    settle_requests(dropped, kEInval, "server stopped before the request was admitted", now(),
                    stats_, lat_, finished_, arena_);
    progress_++;
    cv_done_.notify_all();
    cv_done_.wait(lk, [&] {
      for (const Slot &s2 : slots_)
        if (s2.busy) return false;
      return true;
    });
  }
  cv_dispatch_.notify_all();
  cv_complete_.notify_all();
  if (dispatcher_.joinable()) dispatcher_.join();
  if (completer_.joinable()) completer_.join();
  std::lock_guard<std::mutex> lk(mu_);
  for (Slot &s : slots_) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.d_ent) (void)hipFree(s.d_ent);
    if (s.d_begins) (void)hipFree(s.d_begins);
    if (s.h_ent) (void)hipHostFree(s.h_ent);
    if (s.h_begins) (void)hipHostFree(s.h_begins);
    for (hipStream_t ss : s.side) {
      (void)hipStreamSynchronize(ss);
      (void)hipStreamDestroy(ss);
    }
    for (hipEvent_t ev : s.fork) (void)hipEventDestroy(ev);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s = Slot();
  }
  slots_.clear();
  arena_.destroy();
  for (auto &kv : all_) delete kv.second;
  all_.clear();
  finished_.clear();
  cv_done_.notify_all();
}

int Server::plan_cached(const BatchItemSpec &spec, std::shared_ptr<const PlannedItem> &q,
                        std::string &err) {
  // Plans depend on shape, perm, padding, update / complex-coefficient flags
  // and the pointers' 16- and 128-byte alignment (vectors; streaming), not
  // on the coefficient values.
  const bool cc = spec.cc && (spec.coef.ai != 0.0 || spec.coef.bi != 0.0);
  const bool upd = cc ? (std::fabs(spec.coef.br) >= 1e-16 || std::fabs(spec.coef.bi) >= 1e-16)
                      : std::fabs(spec.coef.br) >= 1e-16;
  const uint32_t es = dtype_esize(spec.dtype);
  // A binary key (this runs once per request, so no formatting).
  std::string key;
  key.reserve(20 + 16 * (size_t)kMaxRank);
  auto put = [&key](uint32_t v) { key.append(reinterpret_cast<const char *>(&v), sizeof v); };
  put((uint32_t)spec.dtype);
  put(spec.rank);
  put((upd ? 1u : 0u) | (cc ? 2u : 0u) |
      ((((uintptr_t)spec.A | (uintptr_t)spec.B) % 128) == 0 ? 4u : 0u));
  put(es ? align_vec(spec.A, es, 16 / es) : 0);
  put(es ? align_vec(spec.B, es, 16 / es) : 0);
  for (uint32_t d = 0; d < spec.rank && d < (uint32_t)kMaxRank; ++d) {
    put(spec.in_size[d]);
    put(spec.perm[d]);
    put(spec.in_outer[d]);
    put(spec.out_outer[d]);
  }
  {
    std::lock_guard<std::mutex> lk(plan_mu_);
    auto it = plans_.find(key);
    if (it != plans_.end()) {
      q = it->second;
      return kOk;
    }
  }
  auto fresh = std::make_shared<PlannedItem>();
  int rc = plan_item(spec, TuneParams{}, *fresh, err);
  if (rc != kOk) return rc;
  q = fresh;
  std::lock_guard<std::mutex> lk(plan_mu_);
  if (plans_.size() >= kPlanCacheEntries) plans_.clear();   // requests keep theirs
  plans_[key] = q;
  return kOk;
}

int Server::make_request(const BatchItemSpec &spec_in, bool server_out, Request *&out,
                         std::string &err) {
  out = nullptr;
  BatchItemSpec spec = spec_in;
  if (server_out) {
    if (!arena_.capacity()) { err = "server has no output arena"; return kEInval; }
    spec.B = arena_.base();   // stands in for a 256-byte-aligned arena slot while planning
  }
  std::shared_ptr<const PlannedItem> q;
  int rc = plan_cached(spec, q, err);
  if (rc != kOk) return rc;
  if (server_out && q->update) { err = "server-allocated outputs need beta below 1e-16"; return kEInval; }
  Request *r = new Request();
  r->spec = spec;
  r->plan = q;
  r->bytes = q->p.N * q->p.esize * (uint64_t)(q->update ? 3 : 2);
  r->arena_out = server_out;
  r->arena_len = server_out ? q->p.spanB * q->p.esize : 0;
  if (server_out && round_up(r->arena_len, kGranule) > arena_.capacity()) {
    delete r;
    err = "output larger than the server's arena";
    return kENoMem;
  }
  out = r;
  return kOk;
}

// Queue planned requests (mu_ held); false once the server has stopped.
bool Server::enqueue_locked(Request *r, uint64_t &ticket) {
  if (stop_) return false;
  r->ticket = next_ticket_++;
  r->t_submit = now();
  if (stats_.first_submit < 0) stats_.first_submit = r->t_submit;
  stats_.submitted++;
  all_[r->ticket] = r;
  pending_.push_back(r);
  ticket = r->ticket;
  return true;
}

int Server::submit(const BatchItemSpec &spec, bool server_out, uint64_t &ticket, std::string &err) {
  Request *r = nullptr;
  int rc = make_request(spec, server_out, r, err);
  if (rc != kOk) return rc;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (!enqueue_locked(r, ticket)) {
      delete r;
      err = "server stopped";
      return kEInval;
    }
  }
  cv_dispatch_.notify_one();
  return kOk;
}

int Server::submit_many(const BatchItemSpec *specs, const bool *server_out, size_t count,
                        uint64_t *tickets, size_t &queued, std::string &err) {
  queued = 0;
  // Planned and queued in chunks that double in size (128, 256, ...), the
  // dispatcher woken after each: the GPU starts on the first chunk while the
  // rest are planned, and each chunk's kernels outlast the planning of the
  // next.
  std::vector<Request *> reqs;
  int rc = kOk;
  size_t i = 0, chunk = kSubmitFirstChunk;
  while (i < count && rc == kOk) {
    reqs.clear();
    for (; i < count && reqs.size() < chunk; ++i) {
      Request *r = nullptr;
      rc = make_request(specs[i], server_out[i], r, err);
      if (rc != kOk) break;
      reqs.push_back(r);
    }
    size_t q = 0;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (Request *r : reqs) {
        if (!enqueue_locked(r, tickets[queued])) break;
        ++queued;
        ++q;
      }
    }
    if (q) cv_dispatch_.notify_one();
    if (q < reqs.size()) {
      for (size_t k = q; k < reqs.size(); ++k) delete reqs[k];
      if (rc == kOk) {
        err = "server stopped";
        rc = kEInval;
      }
    }
    chunk *= 2;
  }
  return rc;
}

uint64_t Server::live_budget_locked() {
  if (cfg_.inflight_bytes) return budget_ = cfg_.inflight_bytes;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget_ = (uint64_t)fr / 4;
  else (void)hipGetLastError();   // keep the last budget
  return budget_;
}

uint64_t Server::budget() {
  std::lock_guard<std::mutex> lk(mu_);
  return budget_;
}

void Server::dispatch_loop() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  bool blocked = false;
  uint64_t seen_done = 0;
  for (;;) {
    cv_dispatch_.wait(lk, [&] {
      if (stop_) return true;
      if (pending_.empty()) return false;
      if (blocked && progress_ == seen_done) return false;
      for (const Slot &s : slots_)
        if (!s.busy) return true;
      return false;
    });
    if (stop_) break;
    blocked = false;
    Slot *slot = nullptr;
    for (Slot &s : slots_)
      if (!s.busy) {
        slot = &s;
        break;
      }
    // Admit in arrival order while the in-flight byte budget (from the HBM
    // free now, so that allocations made outside the server shrink it) and
    // the arena allow; a request is always admitted when nothing else is in
    // flight.
    const uint64_t budget = live_budget_locked();
    std::vector<Request *> batch;
    uint64_t add = 0;
    while (!pending_.empty() && batch.size() < cfg_.max_batch) {
      Request *r = pending_.front();
      const bool idle = inflight_ + add == 0;
      if (!idle && inflight_ + add + r->bytes > budget) break;
      if (r->arena_out) {
        uint64_t off = 0;
        if (!arena_.alloc(r->arena_len, off)) {
          if (idle && arena_.used() == 0) {
            pending_.pop_front();
            std::vector<Request *> one{r};
            settle_requests(one, kENoMem, "output does not fit the server's arena", now(), stats_,
                            lat_, finished_, arena_);
            progress_++;
            cv_done_.notify_all();
            continue;
          }
          break;
        }
        r->arena_off = off;
        r->spec.B = arena_.base() + off;
      }
      pending_.pop_front();
      batch.push_back(r);
      add += r->bytes;
    }
    if (batch.empty()) {
      blocked = true;
      seen_done = progress_;
      continue;
    }
    inflight_ += add;
    slot->busy = true;
    slot->reqs = batch;
    slot->bytes = add;
    slot->seq = launch_seq_++;
    stats_.launches++;
    lk.unlock();

    // One launch per (element type, kernel variant, LDS class), as for
    // batches (or per element type and LDS class with HPTTX_BATCH_MODE=
    // merged); within one, largest first.
    const bool merged = batch_merged_default(false);
    for (Request *r : batch) {
      BatchEntry probe;
      make_entry(*r->plan, r->spec, probe);
      r->variant = batch_group_key(probe.variant, merged);
      r->lds_class = merged ? 0u : batch_lds_class(probe.lds_bytes);
    }
    std::stable_sort(batch.begin(), batch.end(), [](const Request *a, const Request *b) {
      if (a->spec.dtype != b->spec.dtype) return a->spec.dtype < b->spec.dtype;
      if (a->variant != b->variant) return a->variant < b->variant;
      if (a->lds_class != b->lds_class) return a->lds_class < b->lds_class;
      return a->plan->p.N > b->plan->p.N;
    });
    struct Group {
      int dtype;
      uint32_t variant, lds_class, first, count, total, lds;
    };
    std::vector<Group> groups;
    hipError_t he = hipSuccess;
    for (uint32_t i = 0; i < batch.size(); ++i) {
      Request *r = batch[i];
      if (groups.empty() || groups.back().dtype != r->spec.dtype ||
          groups.back().variant != r->variant || groups.back().lds_class != r->lds_class)
        groups.push_back(Group{r->spec.dtype, r->variant, r->lds_class, i, 0, 0, 0});
      Group &g = groups.back();
      BatchEntry &e = slot->h_ent[i];
      make_entry(*r->plan, r->spec, e);
      if ((uint64_t)g.total + e.ntiles > 0xFFFFFFFFull) he = hipErrorInvalidValue;
      e.tile_begin = g.total;
      slot->h_begins[i] = g.total;
      g.total += e.ntiles;
      g.lds = std::max(g.lds, batch_lds_bytes(e.lds_bytes));
      g.count++;
      r->t_launch = now();
    }
    const size_t n = batch.size();
    if (he == hipSuccess)
      he = hipMemcpyAsync(slot->d_ent, slot->h_ent, sizeof(BatchEntry) * n, hipMemcpyHostToDevice,
                          slot->stream);
    if (he == hipSuccess)
      he = hipMemcpyAsync(slot->d_begins, slot->h_begins, sizeof(uint32_t) * n,
                          hipMemcpyHostToDevice, slot->stream);
    if (groups.size() > kMaxSlotGroups) he = hipErrorInvalidValue;
    // Several groups: fork them over the slot's lanes (the groups' tiles,
    // largest group first, to the lane with the fewest so far) and join the
    // lanes back onto the slot's stream, as a batch exec does.  The tables
    // were copied on the slot's stream, before the fork.
    const size_t lanes = groups.size() > 1 ? 1 + slot->side.size() : 1;
    if (lanes > 1 && he == hipSuccess) {
      he = hipEventRecord(slot->fork[0], slot->stream);
      for (size_t i = 0; i + 1 < lanes && he == hipSuccess; ++i)
        he = hipStreamWaitEvent(slot->side[i], slot->fork[0], 0);
    }
    std::vector<size_t> order(groups.size());
    for (size_t gi = 0; gi < order.size(); ++gi) order[gi] = gi;
    std::sort(order.begin(), order.end(),
              [&](size_t x, size_t y) { return groups[x].total > groups[y].total; });
    std::vector<uint64_t> load(lanes, 0);
    for (size_t gi : order) {
      if (he != hipSuccess) break;
      const Group &g = groups[gi];
      size_t best = 0;
      for (size_t l = 1; l < lanes; ++l)
        if (load[l] < load[best]) best = l;
      load[best] += g.total;
      hipStream_t ls = best == 0 ? slot->stream : slot->side[best - 1];
      he = launch_batch_group(g.dtype, g.variant, slot->d_ent + g.first, slot->d_begins + g.first,
                              nullptr, g.count, g.total, g.lds,
                              batch_grid(g.dtype, g.variant, g.lds, g.total, batch_persistent()),
                              ls);
    }
    for (size_t i = 0; i + 1 < lanes && he == hipSuccess; ++i) {
      he = hipEventRecord(slot->fork[1 + i], slot->side[i]);
      if (he == hipSuccess) he = hipStreamWaitEvent(slot->stream, slot->fork[1 + i], 0);
    }
    if (he == hipSuccess) he = hipEventRecord(slot->done, slot->stream);
    lk.lock();
    if (he != hipSuccess) {
      // Nothing (more) was enqueued for this slot: fail its requests once the
      // stream has drained what was.
      const std::string why = std::string("launch: ") + hipGetErrorString(he);
      lk.unlock();
      (void)hipStreamSynchronize(slot->stream);
      for (hipStream_t ss : slot->side) (void)hipStreamSynchronize(ss);
      lk.lock();
      settle_requests(slot->reqs, kEHip, why, now(), stats_, lat_, finished_, arena_);
      progress_++;
      inflight_ -= slot->bytes;
      slot->busy = false;
      cv_done_.notify_all();
      continue;
    }
    for (Request *r : slot->reqs) r->state = kRunning;
    slot->seq |= 1ull << 63;   // launched: the completer may poll its event
    cv_complete_.notify_one();
  }
}

void Server::complete_loop() {
  (void)hipSetDevice(device_);
  // Poll the launched slots' events for up to kSpin (a timed sleep would add
  // the kernel timer's slack, tens of microseconds, to every completion),
  // then block on the oldest launch's event (blocking-sync: no host core is
  // held while long launches run) and poll again.
  constexpr double kSpin = 200e-6;
  double spin_from = -1;
  for (;;) {
    std::vector<Slot *> live;
    {
      std::unique_lock<std::mutex> lk(mu_);
      auto any_busy = [&] {
        for (const Slot &s : slots_)
          if (s.busy) return true;
        return false;
      };
      cv_complete_.wait(lk, [&] {
        if (stop_ && !any_busy()) return true;
        for (const Slot &s : slots_)
          if (s.busy && (s.seq >> 63)) return true;
        return false;
      });
      for (Slot &s : slots_)
        if (s.busy && (s.seq >> 63)) live.push_back(&s);
      if (live.empty()) {
        if (stop_ && !any_busy()) return;
        continue;
      }
    }
    std::vector<std::pair<Slot *, hipError_t>> done;
    for (Slot *s : live) {
      const hipError_t e = hipEventQuery(s->done);
      if (e != hipErrorNotReady) done.push_back({s, e});
    }
    if (done.empty()) {
      const double t = now();
      if (spin_from < 0) spin_from = t;
      if (t - spin_from < kSpin) {
        std::this_thread::yield();
        continue;
      }
      Slot *oldest = live[0];
      for (Slot *s : live)
        if ((s->seq & ~(1ull << 63)) < (oldest->seq & ~(1ull << 63))) oldest = s;
      (void)hipEventSynchronize(oldest->done);   // its status is read by the next query
      continue;
    }
    spin_from = -1;
    const double t = now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto &sd : done) {
        Slot *s = sd.first;
        if (sd.second == hipSuccess) {
          settle_requests(s->reqs, kOk, std::string(), t, stats_, lat_, finished_, arena_);
        } else {
          (void)hipGetLastError();
          settle_requests(s->reqs, kEHip,
                          std::string("launch failed: ") + hipGetErrorString(sd.second), t,
                          stats_, lat_, finished_, arena_);
        }
        progress_++;
        inflight_ -= s->bytes;
        s->busy = false;
        s->seq = 0;
      }
    }
    cv_dispatch_.notify_one();
    cv_done_.notify_all();
  }
}

size_t Server::poll(Request *out, size_t max) {
  std::lock_guard<std::mutex> lk(mu_);
  size_t n = 0;
  while (n < max && !finished_.empty()) {
    out[n++] = *finished_.front();
    finished_.pop_front();
  }
  return n;
}

bool Server::wait(uint64_t ticket, double timeout_s, Request &out) {
  std::unique_lock<std::mutex> lk(mu_);
  auto pred = [&] {
    auto it = all_.find(ticket);
    return it == all_.end() || it->second->state >= kDone;
  };
  if (timeout_s < 0) cv_done_.wait(lk, pred);
  else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return false;
  auto it = all_.find(ticket);
  if (it == all_.end()) return false;
  out = *it->second;
  return true;
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] {
    if (!pending_.empty()) return false;
    for (const Slot &s : slots_)
      if (s.busy) return false;
    return true;
  });
}

int Server::release(uint64_t ticket) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = all_.find(ticket);
  if (it == all_.end()) return kEInval;
  Request *r = it->second;
  if (r->state < kDone) return kEInval;
  if (r->arena_out) arena_.free(r->arena_off, r->arena_len);
  for (auto f = finished_.begin(); f != finished_.end(); ++f)
    if (*f == r) {
      finished_.erase(f);
      break;
    }
  all_.erase(it);
  delete r;
  progress_++;
  cv_dispatch_.notify_one();
  return kOk;
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

void Server::latencies(std::vector<double> &out) {
  std::lock_guard<std::mutex> lk(mu_);
  out = lat_;
}

}  // namespace hpttx
