// mi_server.cpp -- serving engine of libmi_hgp.so: continuous batching of
// ERI and Coulomb/exchange requests on one GPU.  No counterpart in the
// reference, which evaluates one Basis per blocking call.
//
// Model.  Requests are queued in arrival order.  mi_server_step is the whole
// engine, non-blocking, called by the host loop as often as it likes:
//   1. retire: every batch whose completion event has fired releases its plan
//      (stream-ordered frees) and its device bytes;
//   2. admit: while fewer than max_inflight batches are on the device, the
//      head of the queue is admitted if its device bytes fit the budget (it
//      waits otherwise: first come, first served).  ERI requests at the head
//      with equal options are merged, up to max_batch and the budget, into one
//      batched plan (one launch per quartet class over all of them), executed
//      into an engine-owned buffer and copied into each caller's device
//      buffer; a Coulomb/exchange request is a batch of its own.
//   3. each batch runs on its own stream, so batches admitted while others
//      run overlap with them on the device.
// Device memory of a batch (plan buffers, HRR scratch, merged output) is
// allocated stream-ordered on the batch stream and freed the same way, so
// neither admission nor retirement waits for other batches.  Shell-pair data
// stays resident for the life of its batch; the Boys grid and op tables are
// resident for the life of the server (its context).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

#include "mi_internal.h"

namespace {

struct Request {
  int64_t id = 0;
  int kind = MI_REQ_ERI;
  int state = MI_REQ_QUEUED;
  // private copy of the basis
  std::vector<double> xyz, pexp, pcoef, pnorm, scale;
  std::vector<int32_t> L, poff;
  mi_basis b{};
  mi_opts o{};
  const double* D = nullptr;
  double *out = nullptr, *J = nullptr, *K = nullptr;
  int64_t n_unique = 0;
  size_t bytes = 0;  // device bytes when admitted alone (plan + output)
};

struct Batch {
  mi_plan* plan = nullptr;
  int slot = -1;  // stream index
  hipEvent_t done = nullptr;
  double* buf = nullptr;  // merged ERI output (engine-owned)
  size_t bytes = 0;
  std::vector<int64_t> reqs;
};

bool same_opts(const mi_opts& a, const mi_opts& b) {
  return a.order == b.order && a.precision == b.precision && a.tau == b.tau && a.schwarz_cut == b.schwarz_cut;
}

}  // namespace

struct mi_server {
  mi_ctx* ctx = nullptr;
  mi_server_opts opts{};
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Request> reqs;
  std::vector<Batch> inflight;
  int64_t next_id = 1;
  size_t in_use = 0, peak = 0;
  int64_t nbatches = 0, ndone = 0, nfailed = 0;
  std::string err;
};

namespace {

int serr(mi_server* s, int code, const std::string& msg) {
  s->err = msg;
  return code;
}

// Device bytes a request needs alone: its plan (built on the host only) plus
// the merged-output share.
int request_bytes(mi_server* s, Request& r) {
  mi_plan* P = nullptr;
  int rc = mi_plan_build(s->ctx, &r.b, 1, &r.o, MI_BUILD_DEFER_UPLOAD | MI_BUILD_NO_BOUNDS, &P);
  if (rc) return serr(s, rc, s->ctx->err);
  r.n_unique = P->n_unique;
  r.bytes = P->device_bytes + (r.kind == MI_REQ_ERI ? sizeof(double) * (size_t)P->n_unique : 0);
  mi_plan_destroy(P);
  return MI_OK;
}

void fail(mi_server* s, Request& r, const std::string& why) {
  r.state = MI_REQ_FAILED;
  ++s->nfailed;
  s->err = why;
}

void retire(mi_server* s) {
  for (size_t i = 0; i < s->inflight.size();) {
    Batch& b = s->inflight[i];
    const hipError_t q = hipEventQuery(b.done);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    for (int64_t id : b.reqs) {
      Request& r = s->reqs[id];
      if (q == hipSuccess) {
        r.state = MI_REQ_DONE;
        ++s->ndone;
      } else {
        fail(s, r, std::string("batch failed on the device: ") + hipGetErrorString(q));
      }
    }
    hipStream_t st = s->streams[b.slot];
    if (b.buf) hipFreeAsync(b.buf, st);
    mi_plan_destroy(b.plan);  // stream-ordered frees on st
    hipEventDestroy(b.done);
    s->in_use -= b.bytes;
    s->slot_busy[b.slot] = false;
    s->inflight.erase(s->inflight.begin() + i);
  }
}

// Launch one batch of requests ids (all ERI with equal options, or one J/K)
// on stream slot.  Returns MI_OK or an error (requests then marked failed).
int launch(mi_server* s, const std::vector<int64_t>& ids, int slot) {
  hipStream_t st = s->streams[slot];
  Request& r0 = s->reqs[ids[0]];
  std::vector<mi_basis> bases;
  for (int64_t id : ids) bases.push_back(s->reqs[id].b);
  Batch b;
  b.slot = slot;
  b.reqs = ids;
  // No bounds pass at build time: that pass reads the bounds back to the host
  // (a synchronisation), and step must not block.  Mixed and screened batches
  // run their Schwarz pass inside the execute, on the batch stream.
  int rc = mi_plan_build(s->ctx, bases.data(), (int)bases.size(), &r0.o, MI_BUILD_DEFER_UPLOAD | MI_BUILD_NO_BOUNDS,
                         &b.plan);
  if (rc == MI_OK) rc = mi_plan_upload(b.plan, st);
  if (rc == MI_OK && r0.kind == MI_REQ_ERI) {
    const int64_t len = b.plan->n_unique;
    if (hipMallocAsync((void**)&b.buf, sizeof(double) * (size_t)len, st) != hipSuccess) {
      rc = mi_set_err(s->ctx, MI_ENOMEM, "batch output of %lld doubles", (long long)len);
    } else {
      rc = mi_plan_run(b.plan, b.buf, len, st, nullptr, nullptr);
      for (size_t m = 0; rc == MI_OK && m < ids.size(); ++m) {
        const Request& r = s->reqs[ids[m]];
        const int64_t off = b.plan->mol_off[m], n = b.plan->mol_off[m + 1] - off;
        if (hipMemcpyAsync(r.out, b.buf + off, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st) != hipSuccess)
          rc = mi_set_err(s->ctx, MI_EDEVICE, "result copy failed");
      }
    }
  } else if (rc == MI_OK) {
    JKArgs jk{r0.D, r0.J, r0.K};
    rc = mi_plan_run(b.plan, nullptr, 0, st, nullptr, &jk);
  }
  if (rc == MI_OK && hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess)
    rc = mi_set_err(s->ctx, MI_EDEVICE, "hipEventCreate failed");
  if (rc == MI_OK && hipEventRecord(b.done, st) != hipSuccess) rc = mi_set_err(s->ctx, MI_EDEVICE, "event record");
  if (rc != MI_OK) {
    for (int64_t id : ids) fail(s, s->reqs[id], s->ctx->err);
    if (b.buf) hipFreeAsync(b.buf, st);
    if (b.plan) mi_plan_destroy(b.plan);
    if (b.done) hipEventDestroy(b.done);
    return rc;
  }
  b.bytes = b.plan->device_bytes + (b.buf ? sizeof(double) * (size_t)b.plan->n_unique : 0);
  for (int64_t id : ids) s->reqs[id].state = MI_REQ_RUNNING;
  s->in_use += b.bytes;
  s->peak = std::max(s->peak, s->in_use);
  s->slot_busy[slot] = true;
  ++s->nbatches;
  s->inflight.push_back(b);
  return MI_OK;
}

void admit(mi_server* s) {
  const size_t budget = (size_t)s->opts.hbm_budget;
  while (!s->queue.empty() && (int)s->inflight.size() < s->opts.max_inflight) {
    const int64_t head = s->queue.front();
    Request& r = s->reqs[head];
    if (r.bytes > budget) {  // can never fit
      s->queue.pop_front();
      fail(s, r, "request needs more device memory than the server budget");
      continue;
    }
    if (s->in_use + r.bytes > budget) break;  // wait for memory
    int slot = 0;
    while (s->slot_busy[slot]) ++slot;
    std::vector<int64_t> ids{head};
    s->queue.pop_front();
    if (r.kind == MI_REQ_ERI) {
      size_t bytes = r.bytes;
      while (!s->queue.empty() && (int)ids.size() < s->opts.max_batch) {
        Request& n = s->reqs[s->queue.front()];
        if (n.kind != MI_REQ_ERI || !same_opts(n.o, r.o) || s->in_use + bytes + n.bytes > budget) break;
        bytes += n.bytes;
        ids.push_back(n.id);
        s->queue.pop_front();
      }
    }
    launch(s, ids, slot);
  }
}

}  // namespace

extern "C" {

int mi_server_create(int32_t device, const mi_server_opts* opts, mi_server** out) {
  if (!out) return MI_EINVAL;
  *out = nullptr;
  mi_server* s = new mi_server();
  int rc = mi_create(device, &s->ctx);
  if (rc) {
    delete s;
    return rc;
  }
  mi_ctx_set_streams(s->ctx, 1);  // a batch is one stream; concurrency comes from batches
  s->opts.max_inflight = opts ? opts->max_inflight : 4;
  s->opts.max_batch = opts ? opts->max_batch : 64;
  s->opts.hbm_budget = opts ? opts->hbm_budget : (int64_t)32 << 30;
  if (s->opts.max_inflight < 1 || s->opts.max_inflight > 8 || s->opts.max_batch < 1 || s->opts.hbm_budget <= 0) {
    mi_destroy(s->ctx);
    delete s;
    return MI_EINVAL;
  }
  s->streams.assign(s->opts.max_inflight, nullptr);
  s->slot_busy.assign(s->opts.max_inflight, false);
  for (auto& st : s->streams)
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (auto& t : s->streams)
        if (t) hipStreamDestroy(t);
      mi_destroy(s->ctx);
      delete s;
      return MI_EDEVICE;
    }
  *out = s;
  return MI_OK;
}

void mi_server_destroy(mi_server* s) {
  if (!s) return;
  for (auto& st : s->streams) hipStreamSynchronize(st);
  retire(s);
  for (auto& st : s->streams) hipStreamDestroy(st);
  mi_destroy(s->ctx);
  delete s;
}

const char* mi_server_last_error(const mi_server* s) { return s ? s->err.c_str() : "null server"; }

int mi_server_submit(mi_server* s, const mi_basis* basis, const mi_opts* opts, int32_t kind, const double* D,
                     double* out, double* J, double* K, int64_t* id) {
  if (!s || !basis || !id) return MI_EINVAL;
  if (kind != MI_REQ_ERI && kind != MI_REQ_JK) return serr(s, MI_EINVAL, "unknown request kind");
  if (kind == MI_REQ_ERI && !out) return serr(s, MI_EINVAL, "ERI request without an output buffer");
  if (kind == MI_REQ_JK && (!D || !J || !K)) return serr(s, MI_EINVAL, "J/K request without D, J or K");
  Request r;
  r.id = s->next_id++;
  r.kind = kind;
  const int ns = basis->nshell;
  if (ns <= 0 || !basis->shell_prim_off) return serr(s, MI_EINVAL, "basis has no shells");
  const int np = basis->shell_prim_off[ns];
  r.xyz.assign(basis->shell_xyz, basis->shell_xyz + 3 * ns);
  r.L.assign(basis->shell_l, basis->shell_l + ns);
  r.poff.assign(basis->shell_prim_off, basis->shell_prim_off + ns + 1);
  r.pexp.assign(basis->prim_exp, basis->prim_exp + np);
  r.pcoef.assign(basis->prim_coef, basis->prim_coef + np);
  r.pnorm.assign(basis->prim_norm, basis->prim_norm + np);
  r.scale.assign(basis->bf_scale, basis->bf_scale + basis->nbf);
  r.b = *basis;
  r.b.shell_xyz = r.xyz.data();
  r.b.shell_l = r.L.data();
  r.b.shell_prim_off = r.poff.data();
  r.b.prim_exp = r.pexp.data();
  r.b.prim_coef = r.pcoef.data();
  r.b.prim_norm = r.pnorm.data();
  r.b.bf_scale = r.scale.data();
  if (opts) r.o = *opts;
  r.o.order = kind == MI_REQ_JK ? MI_ORDER_JK : (opts ? opts->order : MI_ORDER_CHRR);
  r.o.stream = nullptr;
  r.D = D;
  r.out = out;
  r.J = J;
  r.K = K;
  int rc = request_bytes(s, r);
  if (rc) return rc;
  const int64_t rid = r.id;
  s->reqs.emplace(rid, std::move(r));
  Request& q = s->reqs[rid];  // the vectors moved; re-point the basis
  q.b.shell_xyz = q.xyz.data();
  q.b.shell_l = q.L.data();
  q.b.shell_prim_off = q.poff.data();
  q.b.prim_exp = q.pexp.data();
  q.b.prim_coef = q.pcoef.data();
  q.b.prim_norm = q.pnorm.data();
  q.b.bf_scale = q.scale.data();
  s->queue.push_back(rid);
  *id = rid;
  return MI_OK;
}

int mi_server_step(mi_server* s) {
  if (!s) return MI_EINVAL;
  retire(s);
  admit(s);
  return MI_OK;
}

int mi_server_poll(mi_server* s, int64_t id, int32_t* state) {
  if (!s || !state) return MI_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return serr(s, MI_EINVAL, "unknown request id");
  *state = it->second.state;
  return MI_OK;
}

int mi_server_release(mi_server* s, int64_t id) {
  if (!s) return MI_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return serr(s, MI_EINVAL, "unknown request id");
  if (it->second.state != MI_REQ_DONE && it->second.state != MI_REQ_FAILED)
    return serr(s, MI_EINVAL, "request still queued or running");
  s->reqs.erase(it);
  return MI_OK;
}

int mi_server_drain(mi_server* s) {
  if (!s) return MI_EINVAL;
  while (!s->queue.empty() || !s->inflight.empty()) {
    admit(s);
    if (!s->inflight.empty()) hipEventSynchronize(s->inflight.front().done);
    retire(s);
  }
  return s->nfailed ? MI_EDEVICE : MI_OK;
}

int mi_server_stats(const mi_server* s, int64_t* st, int32_t n) {
  if (!s || !st) return MI_EINVAL;
  int64_t running = 0;
  for (const Batch& b : s->inflight) running += (int64_t)b.reqs.size();
  const int64_t v[8] = {(int64_t)s->queue.size(), running,          s->ndone, s->nfailed, s->nbatches,
                        (int64_t)s->in_use,      (int64_t)s->peak, (int64_t)s->inflight.size()};
  for (int i = 0; i < n && i < 8; ++i) st[i] = v[i];
  return MI_OK;
}

}  // extern "C"
