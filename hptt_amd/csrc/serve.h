// serve.h -- continuous-batching transposition server.
//
// No counterpart in the reference (a library call per transposition).  A
// request is one transposition (the arguments of one plan creation plus its
// data pointers).  Requests queue on the host; a dispatcher thread admits
// them in arrival order while the bytes in flight stay within a budget taken
// from the HBM free at that moment and an output arena (for requests that
// let the server allocate B) has room, packs the admitted requests into
// launches of the batch kernel (one per element type, kernel variant and LDS
// class) and issues them on the next free stream slot -- so requests that
// arrive while others run are admitted as slots and memory free up, and
// their launches overlap on the device.  A completer thread polls each
// slot's event for a short while, then blocks on the oldest one, and stamps
// completion times.  Plans are cached by shape and descriptor tables stay
// resident per slot.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.h"

namespace hpttx {

// Distinct (element type, variant) launches one slot can hold: every variant
// code of every type.
constexpr size_t kMaxSlotGroups = 512;

struct ServerConfig {
  uint64_t inflight_bytes = 0;   // budget of algorithmic bytes in flight (0: a quarter of the
                                 // HBM free at each admission round)
  uint64_t arena_bytes = 0;      // server-allocated outputs (0: none)
  uint32_t slots = 3;            // launches in flight at once (one stream each)
  uint32_t max_batch = 256;      // requests per launch
};

enum ReqState : int { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
  uint64_t ticket = 0;
  BatchItemSpec spec{};
  std::shared_ptr<const PlannedItem> plan;   // shared with the plan cache
  uint64_t bytes = 0;            // algorithmic bytes
  uint64_t arena_off = 0;        // output in the arena (when server-allocated)
  uint64_t arena_len = 0;
  bool arena_out = false;
  int state = kQueued;
  int status = 0;
  uint32_t variant = 0;          // batch-kernel variant of its plan
  uint32_t lds_class = 0;        // LDS need in 8 KiB classes (launch grouping)
  double t_submit = 0, t_launch = 0, t_done = 0;
  std::string error;             // why it failed (state kFailed)
};

// First-fit allocator over one device allocation (256-byte granules).
class Arena {
 public:
  bool init(uint64_t bytes, std::string &err);
  void destroy();
  bool alloc(uint64_t bytes, uint64_t &off);
  void free(uint64_t off, uint64_t bytes);
  char *base() const { return base_; }
  uint64_t capacity() const { return cap_; }
  uint64_t used() const { return used_; }

 private:
  char *base_ = nullptr;
  uint64_t cap_ = 0, used_ = 0;
  std::map<uint64_t, uint64_t> free_;   // offset -> length
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // Side streams of the slot: a launch round's groups are forked over the
  // slot's stream and these (kBatchLanes in all, as a batch's are) and
  // joined back before `done`.  fork[0] is recorded on the slot's stream,
  // fork[1 + i] on side[i].
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> fork;
  BatchEntry *d_ent = nullptr;
  uint32_t *d_begins = nullptr;
  BatchEntry *h_ent = nullptr;     // pinned staging
  uint32_t *h_begins = nullptr;    // pinned staging
  std::vector<Request *> reqs;
  uint64_t bytes = 0;
  bool busy = false;
  uint64_t seq = 0;
};

struct ServerStats {
  uint64_t submitted = 0, completed = 0, failed = 0, launches = 0;
  uint64_t bytes_done = 0;
  double first_submit = -1, last_done = -1;
};

// Completion bookkeeping shared by the completer (a slot's event reached or
// failed) and the dispatcher (a launch that could not be enqueued), kept
// free of HIP calls so that it can be tested on a host without a GPU.
// status kOk marks every request done (latency recorded, bytes counted);
// any other status marks it failed with that status and `why`.  Arena
// outputs of failed requests are returned to `arena`.  Each request goes to
// `finished`; the return value is the requests' algorithmic bytes.
uint64_t settle_requests(std::vector<Request *> &reqs, int status, const std::string &why,
                         double t, ServerStats &st, std::vector<double> &lat,
                         std::deque<Request *> &finished, Arena &arena);

class Server {
 public:
  int start(const ServerConfig &cfg, std::string &err);
  void stop();
  int submit(const BatchItemSpec &spec, bool server_out, uint64_t &ticket, std::string &err);
  // Requests planned first, then queued together under one lock with one
  // wake-up of the dispatcher (which then finds them all pending).  Stops at
  // the first request that fails to plan: *queued of them are queued, with
  // their tickets in tickets[0 .. *queued).
  int submit_many(const BatchItemSpec *specs, const bool *server_out, size_t count,
                  uint64_t *tickets, size_t &queued, std::string &err);
  // Completed (done or failed) tickets not yet reported by poll.
  size_t poll(Request *out, size_t max);
  bool wait(uint64_t ticket, double timeout_s, Request &out);
  void drain();
  int release(uint64_t ticket);
  ServerStats stats();
  void latencies(std::vector<double> &out);
  double now() const;
  // In-flight byte budget used by the last admission round.
  uint64_t budget();
  char *arena_base() const { return arena_.base(); }
  const ServerConfig &config() const { return cfg_; }

 private:
  void dispatch_loop();
  void complete_loop();
  // Budget for the next admission round (mu_ held): the configured one, or
  // a quarter of the HBM free right now.
  uint64_t live_budget_locked();
  int plan_cached(const BatchItemSpec &spec, std::shared_ptr<const PlannedItem> &q,
                  std::string &err);
  int make_request(const BatchItemSpec &spec, bool server_out, Request *&out, std::string &err);
  bool enqueue_locked(Request *r, uint64_t &ticket);

  ServerConfig cfg_;
  uint64_t budget_ = 0;
  std::mutex mu_;
  std::condition_variable cv_dispatch_, cv_complete_, cv_done_;
  std::deque<Request *> pending_;
  std::unordered_map<uint64_t, Request *> all_;
  std::deque<Request *> finished_;     // not yet polled
  std::vector<Slot> slots_;
  Arena arena_;
  uint64_t inflight_ = 0;
  uint64_t next_ticket_ = 1;
  uint64_t launch_seq_ = 0;
  uint64_t progress_ = 0;            // completions, failures and releases so far
  bool stop_ = false;
  ServerStats stats_;
  std::vector<double> lat_;
  std::thread dispatcher_, completer_;
  // Plans by shape (about 1.5 KB each); the cache is emptied when full.
  static constexpr size_t kPlanCacheEntries = 16384;
  // submit_many's first chunk of requests planned before the dispatcher is woken.
  static constexpr size_t kSubmitFirstChunk = 128;
  std::mutex plan_mu_;
  std::unordered_map<std::string, std::shared_ptr<const PlannedItem>> plans_;
  int device_ = 0;
};

}  // namespace hpttx
