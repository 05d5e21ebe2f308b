// serve_capi.cpp -- the extern "C" serving interface declared in include/hptt_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "hptt_amd.h"
#include "serve.h"

using namespace hpttx;

struct hpttx_server {
  Server s;
};

namespace hpttx {
void set_last_error(const std::string &m);
int batch_item_to_spec(const hpttx_batch_item &it, uint32_t k, BatchItemSpec &s, bool allow_null_out);
}  // namespace hpttx

namespace {
void fill_completion(const Request &r, hpttx_completion *c) {
  c->ticket = r.ticket;
  c->state = r.state;
  c->status = r.status;
  c->out = r.spec.B;
  c->t_submit = r.t_submit;
  c->t_launch = r.t_launch;
  c->t_done = r.t_done;
  c->bytes = r.bytes;
}
}  // namespace

extern "C" {

int hpttx_server_create(const hpttx_server_config *config, hpttx_server **out) {
  if (!out) { set_last_error("server out-pointer is null"); return HPTTX_EINVAL; }
  *out = nullptr;
  ServerConfig cfg;
  if (config) {
    cfg.inflight_bytes = config->inflight_bytes;
    cfg.arena_bytes = config->arena_bytes;
    if (config->slots) cfg.slots = config->slots;
    if (config->max_batch) cfg.max_batch = config->max_batch;
  }
  hpttx_server *h = new (std::nothrow) hpttx_server();
  if (!h) { set_last_error("out of host memory"); return HPTTX_ENOMEM; }
  std::string err;
  int rc = h->s.start(cfg, err);
  if (rc != kOk) {
    delete h;
    set_last_error(err);
    return rc;
  }
  *out = h;
  return HPTTX_OK;
}

int hpttx_server_submit(hpttx_server *h, const hpttx_batch_item *req, uint64_t *ticket) {
  if (!h || !req || !ticket) { set_last_error("null argument"); return HPTTX_EINVAL; }
  BatchItemSpec spec;
  int rc = batch_item_to_spec(*req, 0, spec, true);
  if (rc != HPTTX_OK) return rc;
  std::string err;
  rc = h->s.submit(spec, req->out == nullptr, *ticket, err);
  if (rc != kOk) set_last_error(err);
  return rc;
}

int hpttx_server_submit_many(hpttx_server *h, uint32_t count, const hpttx_batch_item *reqs,
                             uint64_t *tickets, uint32_t *queued) {
  if (!h || !queued || (count && (!reqs || !tickets))) {
    set_last_error("null argument");
    return HPTTX_EINVAL;
  }
  *queued = 0;
  std::vector<BatchItemSpec> specs(count);
  std::unique_ptr<bool[]> own(new bool[count ? count : 1]);
  for (uint32_t i = 0; i < count; ++i) {
    tickets[i] = 0;
    own[i] = reqs[i].out == nullptr;
    int rc = batch_item_to_spec(reqs[i], i, specs[i], true);
    if (rc != HPTTX_OK) return rc;   // an argument error queues none
  }
  std::string err;
  size_t n = 0;
  int rc = h->s.submit_many(specs.data(), own.get(), count, tickets, n, err);
  *queued = (uint32_t)n;
  if (rc != kOk) set_last_error(err);
  return rc;
}

int hpttx_server_poll(hpttx_server *h, hpttx_completion *out, uint32_t max, uint32_t *n) {
  if (!h || !n || (max && !out)) { set_last_error("null argument"); return HPTTX_EINVAL; }
  std::vector<Request> tmp(max);
  size_t got = h->s.poll(tmp.data(), max);
  for (size_t i = 0; i < got; ++i) {
    fill_completion(tmp[i], &out[i]);
    // a failed request's reason becomes this thread's last error
    if (tmp[i].state == kFailed) set_last_error(tmp[i].error);
  }
  *n = (uint32_t)got;
  return HPTTX_OK;
}

int hpttx_server_wait(hpttx_server *h, uint64_t ticket, double timeout_s, hpttx_completion *out) {
  if (!h) { set_last_error("null server"); return HPTTX_EINVAL; }
  Request r;
  if (!h->s.wait(ticket, timeout_s, r)) {
    set_last_error("unknown ticket or timeout");
    return HPTTX_EINVAL;
  }
  if (out) fill_completion(r, out);
  if (r.state == kFailed) set_last_error(r.error);
  return HPTTX_OK;
}

int hpttx_server_drain(hpttx_server *h) {
  if (!h) { set_last_error("null server"); return HPTTX_EINVAL; }
  h->s.drain();
  return HPTTX_OK;
}

int hpttx_server_release(hpttx_server *h, uint64_t ticket) {
  if (!h) { set_last_error("null server"); return HPTTX_EINVAL; }
  int rc = h->s.release(ticket);
  if (rc != kOk) set_last_error("unknown or unfinished ticket");
  return rc;
}

int hpttx_server_get_stats(hpttx_server *h, hpttx_server_stats *st) {
  if (!h || !st) { set_last_error("null argument"); return HPTTX_EINVAL; }
  ServerStats s = h->s.stats();
  std::vector<double> lat;
  h->s.latencies(lat);
  std::memset(st, 0, sizeof(*st));
  st->submitted = s.submitted;
  st->completed = s.completed;
  st->failed = s.failed;
  st->launches = s.launches;
  st->bytes_done = s.bytes_done;
  st->first_submit = s.first_submit;
  st->last_done = s.last_done;
  st->inflight_budget = h->s.budget();
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      size_t i = (size_t)std::ceil(p * lat.size()) - 1;
      return lat[std::min(i, lat.size() - 1)];
    };
    st->p50_latency = pct(0.50);
    st->p99_latency = pct(0.99);
    double sum = 0;
    for (double v : lat) sum += v;
    st->mean_latency = sum / lat.size();
  }
  return HPTTX_OK;
}

double hpttx_server_now(hpttx_server *h) { return h ? h->s.now() : 0.0; }

void hpttx_server_destroy(hpttx_server *h) {
  if (!h) return;
  h->s.stop();
  delete h;
}

}  // extern "C"
