// mi_server.cpp -- serving engine of libmi_hgp.so: continuous batching of
// ERI and Coulomb/exchange requests on one GPU.  No counterpart in the
// reference, which evaluates one Basis per blocking call.
//
// Model.  Requests are queued in arrival order.  mi_server_step is the whole
// engine on the caller's thread, non-blocking, called as often as the host
// loop likes:
//   1. retire: every launched batch whose completion event has fired
//      releases its plan (stream-ordered frees) and its device bytes;
//   2. admit: while fewer than max_inflight batches are on the device, the
//      head of the queue (once sized) is admitted if its device bytes fit the
//      budget (it waits otherwise: first come, first served).  ERI requests
//      at the head with equal options are merged, up to max_batch and the
//      budget, into one batched plan (one launch per quartet class over all
//      of them) that writes each molecule straight into its caller's buffer;
//      a Coulomb/exchange request is a batch of its own;
//   3. each batch runs on its own stream, so batches admitted while others
//      run overlap with them on the device.
// The host work of a request -- sizing its device footprint at submit, and
// for an admitted batch the plan build (with the Schwarz bounds pass that
// orders pairs by bound and leaves screened quartets out of the tasks), the
// upload and the launch -- runs on a pool of worker threads, each with its
// own context, so step never waits for it and plan builds of several batches
// overlap each other and the device.  Device memory of a batch is allocated
// stream-ordered on the batch stream and freed the same way.  Shell-pair data
// stays resident for the life of its batch; the Boys grid and op tables are
// resident for the life of the server (its contexts).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
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
  size_t bytes = 0;               // device bytes of its plan alone (set by a worker)
  std::atomic<int> sized{0};      // 0: sizing pending, 1: sized, -1: sizing failed
  std::string why;                // sizing error
};

// Batch states: a worker builds, uploads and launches it (PREPARING), then it
// runs on its stream until its event fires (LAUNCHED), or its host work
// failed (BROKEN).
enum { B_PREPARING = 0, B_LAUNCHED = 1, B_BROKEN = 2 };

struct Batch {
  mi_plan* plan = nullptr;
  int slot = -1;  // stream index
  hipEvent_t done = nullptr;
  size_t bytes = 0;  // reserved at admission
  std::vector<int64_t> reqs;
  std::atomic<int> state{B_PREPARING};
  std::string why;
};

// Worker pool: jobs run on threads that each own a context (plans are built
// and launched there; a context is used by one thread at a time).
struct Pool {
  std::vector<std::thread> threads;
  std::vector<mi_ctx*> ctxs;
  std::deque<std::function<void(mi_ctx*)>> jobs;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  void push(std::function<void(mi_ctx*)> f) {
    {
      std::lock_guard<std::mutex> g(mu);
      jobs.push_back(std::move(f));
    }
    cv.notify_one();
  }
  void run(mi_ctx* c) {
    hipSetDevice(c->device);
    for (;;) {
      std::function<void(mi_ctx*)> f;
      {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] { return stop || !jobs.empty(); });
        if (jobs.empty()) return;  // stop requested and nothing left
        f = std::move(jobs.front());
        jobs.pop_front();
      }
      f(c);
    }
  }
};

bool same_opts(const mi_opts& a, const mi_opts& b) {
  return a.order == b.order && a.precision == b.precision && a.tau == b.tau && a.schwarz_cut == b.schwarz_cut &&
         a.method == b.method;
}

}  // namespace

#define MI_SERVER_WORKERS 8

struct mi_server {
  mi_ctx* ctx = nullptr;
  mi_server_opts opts{};
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  std::list<std::unique_ptr<Batch>> inflight;
  Pool pool;
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

// Device bytes a request needs alone (its plan, built on the host only; the
// outputs are the caller's buffers).  Runs on a worker.
void size_request(mi_ctx* c, Request* r) {
  mi_plan* P = nullptr;
  const int rc = mi_plan_build(c, &r->b, 1, &r->o, MI_BUILD_DEFER_UPLOAD | MI_BUILD_NO_BOUNDS, &P);
  if (rc) {
    r->why = c->err;
    r->sized.store(-1);
    return;
  }
  r->n_unique = P->n_unique;
  r->bytes = P->device_bytes;
  mi_plan_destroy(P);
  r->sized.store(1);
}

void fail(mi_server* s, Request& r, const std::string& why) {
  r.state = MI_REQ_FAILED;
  ++s->nfailed;
  s->err = why;
}

void retire(mi_server* s) {
  for (auto it = s->inflight.begin(); it != s->inflight.end();) {
    Batch& b = **it;
    const int st = b.state.load();
    if (st == B_PREPARING) {
      ++it;
      continue;
    }
    hipError_t q = hipSuccess;
    if (st == B_LAUNCHED) {
      q = hipEventQuery(b.done);
      if (q == hipErrorNotReady) {
        ++it;
        continue;
      }
    }
    for (int64_t id : b.reqs) {
      Request& r = *s->reqs[id];
      if (st == B_LAUNCHED && q == hipSuccess) {
        r.state = MI_REQ_DONE;
        ++s->ndone;
      } else {
        fail(s, r, st == B_BROKEN ? b.why : std::string("batch failed on the device: ") + hipGetErrorString(q));
      }
    }
    if (b.plan) mi_plan_destroy(b.plan);  // stream-ordered frees on the batch stream
    if (b.done) hipEventDestroy(b.done);
    s->in_use -= b.bytes;
    s->slot_busy[b.slot] = false;
    it = s->inflight.erase(it);
  }
}

// Build, upload and launch one batch (runs on a worker with context c): all
// ERI requests with equal options, or one J/K request, on stream st.
void launch_batch(mi_server* s, mi_ctx* c, Batch* b, std::vector<Request*> rs, hipStream_t st) {
  std::vector<mi_basis> bases;
  for (Request* r : rs) bases.push_back(r->b);
  const Request& r0 = *rs[0];
  // a screening or mixed plan runs its bounds pass here (synchronous on this
  // worker's bounds stream), so pairs are ordered by bound and screened
  // quartets never enter a task
  int rc = mi_plan_build(c, bases.data(), (int)bases.size(), &r0.o, MI_BUILD_DEFER_UPLOAD, &b->plan);
  if (rc == MI_OK && r0.kind == MI_REQ_ERI) {
    std::vector<double*> outs;
    for (Request* r : rs) outs.push_back(r->out);
    rc = mi_plan_set_mol_outputs(b->plan, outs.data());
  }
  if (rc == MI_OK) rc = mi_plan_upload(b->plan, st);
  if (rc == MI_OK) {
    if (r0.kind == MI_REQ_ERI) {
      rc = mi_plan_run(b->plan, rs[0]->out, b->plan->n_unique, st, nullptr, nullptr);
    } else {
      JKArgs jk{r0.D, r0.J, r0.K};
      rc = mi_plan_run(b->plan, nullptr, 0, st, nullptr, &jk);
    }
  }
  if (rc == MI_OK && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess)
    rc = mi_set_err(c, MI_EDEVICE, "hipEventCreate failed");
  if (rc == MI_OK && hipEventRecord(b->done, st) != hipSuccess) rc = mi_set_err(c, MI_EDEVICE, "event record");
  if (rc != MI_OK) {
    b->why = c->err;
    hipStreamSynchronize(st);  // nothing of this batch may still run when its plan goes
    b->state.store(B_BROKEN);
    return;
  }
  b->state.store(B_LAUNCHED);
}

void admit(mi_server* s) {
  const size_t budget = (size_t)s->opts.hbm_budget;
  while (!s->queue.empty() && (int)s->inflight.size() < s->opts.max_inflight) {
    const int64_t head = s->queue.front();
    Request& r = *s->reqs[head];
    const int sized = r.sized.load();
    if (sized == 0) break;  // a worker is still sizing the head
    if (sized < 0 || r.bytes > budget) {  // can never run
      s->queue.pop_front();
      fail(s, r, sized < 0 ? r.why : "request needs more device memory than the server budget");
      continue;
    }
    if (s->in_use + r.bytes > budget) break;  // wait for memory
    int slot = 0;
    while (s->slot_busy[slot]) ++slot;
    std::vector<Request*> rs{&r};
    size_t bytes = r.bytes;
    s->queue.pop_front();
    if (r.kind == MI_REQ_ERI) {
      while (!s->queue.empty() && (int)rs.size() < s->opts.max_batch) {
        Request& n = *s->reqs[s->queue.front()];
        if (n.sized.load() != 1 || n.kind != MI_REQ_ERI || !same_opts(n.o, r.o) || s->in_use + bytes + n.bytes > budget)
          break;
        bytes += n.bytes;
        rs.push_back(&n);
        s->queue.pop_front();
      }
    }
    auto bp = std::make_unique<Batch>();
    Batch* b = bp.get();
    b->slot = slot;
    b->bytes = bytes;
    for (Request* q : rs) {
      b->reqs.push_back(q->id);
      q->state = MI_REQ_RUNNING;
    }
    s->in_use += bytes;
    s->peak = std::max(s->peak, s->in_use);
    s->slot_busy[slot] = true;
    ++s->nbatches;
    s->inflight.push_back(std::move(bp));
    hipStream_t st = s->streams[slot];
    s->pool.push([s, b, rs, st](mi_ctx* c) { launch_batch(s, c, b, rs, st); });
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
  bool ok = true;
  for (auto& st : s->streams)
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) ok = false;
  // worker contexts (one per thread: plan building writes a context's error text)
  for (int w = 0; ok && w < MI_SERVER_WORKERS; ++w) {
    mi_ctx* c = nullptr;
    if (mi_create(device, &c) != MI_OK) {
      ok = false;
      break;
    }
    mi_ctx_set_streams(c, 1);
    s->pool.ctxs.push_back(c);
  }
  if (!ok) {
    for (auto& t : s->streams)
      if (t) hipStreamDestroy(t);
    for (mi_ctx* c : s->pool.ctxs) mi_destroy(c);
    mi_destroy(s->ctx);
    delete s;
    return MI_EDEVICE;
  }
  for (mi_ctx* c : s->pool.ctxs) s->pool.threads.emplace_back([s, c] { s->pool.run(c); });
  *out = s;
  return MI_OK;
}

void mi_server_destroy(mi_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->pool.mu);
    s->pool.stop = true;
  }
  s->pool.cv.notify_all();
  for (auto& t : s->pool.threads) t.join();  // queued jobs finish first
  for (auto& st : s->streams) hipStreamSynchronize(st);
  retire(s);
  for (auto& st : s->streams) hipStreamDestroy(st);
  for (mi_ctx* c : s->pool.ctxs) mi_destroy(c);
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
  auto rp = std::make_unique<Request>();
  Request& r = *rp;
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
  const int64_t rid = r.id;
  Request* rq = rp.get();
  s->reqs.emplace(rid, std::move(rp));
  s->pool.push([rq](mi_ctx* c) { size_request(c, rq); });  // the object stays put until released
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
  *state = it->second->state;
  return MI_OK;
}

int mi_server_release(mi_server* s, int64_t id) {
  if (!s) return MI_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return serr(s, MI_EINVAL, "unknown request id");
  if (it->second->state != MI_REQ_DONE && it->second->state != MI_REQ_FAILED)
    return serr(s, MI_EINVAL, "request still queued or running");
  s->reqs.erase(it);
  return MI_OK;
}

int mi_server_drain(mi_server* s) {
  if (!s) return MI_EINVAL;
  while (!s->queue.empty() || !s->inflight.empty()) {
    admit(s);
    retire(s);
    bool launched = false;
    for (auto& b : s->inflight)
      if (b->state.load() == B_LAUNCHED) {
        hipEventSynchronize(b->done);
        launched = true;
        break;
      }
    if (!launched) std::this_thread::sleep_for(std::chrono::microseconds(50));  // workers still preparing
  }
  return s->nfailed ? MI_EDEVICE : MI_OK;
}

int mi_server_stats(const mi_server* s, int64_t* st, int32_t n) {
  if (!s || !st) return MI_EINVAL;
  int64_t running = 0;
  for (const auto& b : s->inflight) running += (int64_t)b->reqs.size();
  const int64_t v[8] = {(int64_t)s->queue.size(), running,          s->ndone, s->nfailed, s->nbatches,
                        (int64_t)s->in_use,      (int64_t)s->peak, (int64_t)s->inflight.size()};
  for (int i = 0; i < n && i < 8; ++i) st[i] = v[i];
  return MI_OK;
}

}  // extern "C"
