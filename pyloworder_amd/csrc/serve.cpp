// Native request loop (pc_serve_*): continuous batching of POD, projection
// and reconstruction requests on one GPU under an HBM admission budget, with
// POD results kept resident as bases.
//
// No reference counterpart: pyLOM runs one POD per script.  Each request is
// one library call (pc_tsqr_svd with fused centring, pc_project,
// pc_reconstruct) on a slot's own context and stream, so its result is the
// one that call gives alone.
//
//   * Admission is FIFO with a bounded skip-ahead: a request is admitted as
//     soon as a slot is idle and its output bytes fit the budget, so requests
//     that arrive while others run join at once.  Idle bases (no pinned
//     users) are evicted least recently used first to make room.
//   * Slots are worker threads, each with a HIP stream of default priority and
//     a pc_ctx whose workspace is reserved up front for the largest request
//     the server accepts (no workspace reallocation while serving).
//   * The output budget is one device allocation made at creation (an
//     arena).  A request's output blocks are carved from it, best fit over
//     address-ordered free ranges, when the request is admitted, and go back
//     (coalesced with their neighbours) on release or eviction.  Serving
//     itself calls neither hipMalloc nor hipFree, and both synchronise or
//     stall the device.  Admission is therefore allocation: a request whose
//     blocks do not fit evicts idle bases, least recently used first, until
//     they do or none is left.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

#include "arena.hpp"
#include "common.hpp"

using pcserve::Arena;

int pc_ws_reserve(pc_ctx* ctx, int64_t bytes);

namespace {

struct Basis {
  void* U = nullptr;
  void* S = nullptr;
  void* VT = nullptr;
  int64_t m = 0, n = 0;
  int dtype = PC_F32;
  int64_t nbytes = 0;
  double last_used = 0.0;
  int pins = 0;
};

struct Req {
  pc_request q{};
  int64_t ticket = -1;
  int status = PC_PENDING;
  int64_t basis = -1;
  void* blk[3] = {nullptr, nullptr, nullptr};  // a POD's U, S, VT, reserved at admission
  void* out = nullptr;                         // a projection's / reconstruction's output
  hipEvent_t ready = nullptr;                  // X's producer at submission
  int64_t out_rows = 0, out_cols = 0;
  int64_t nbytes = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int slot = -1;
  std::string err;
};

int64_t esize(int dtype) { return dtype == PC_F64 ? 8 : 4; }


}  // namespace

struct pc_server {
  int device = 0;
  int nslots = 0;
  int64_t budget = 0;
  int skip_ahead = 0;
  int precision = PC_PREC_MIXED;
  int center = PC_CENTER_MEAN;
  int64_t m_max = 0, n_max = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req*> pending;
  std::map<int64_t, Req*> reqs;
  std::map<int64_t, Basis> bases;
  std::vector<int> idle;
  std::vector<Req*> work;
  std::vector<std::thread> threads;
  std::vector<pc_ctx*> ctx;
  std::vector<hipStream_t> streams;
  // outputs and bases, carved from one allocation of the output budget
  // (guarded by mu, like the rest of the server's state)
  void* arena_mem = nullptr;
  Arena arena;
  int64_t inflight_bytes = 0, resident_bytes = 0;
  int64_t next_ticket = 0, next_basis = 0;
  int64_t n_completed = 0, n_rejected = 0, n_evictions = 0, n_inflight = 0;
  std::vector<hipEvent_t> events;  // free input-ready events
  bool stop = false;
  std::string err;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int serve_fail(pc_server* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

// precision of a POD solve: the server's, except that the mixed kernels take
// float32 and n <= 256 (wider or float64 requests run FP64 MFMA)
int pod_precision(const pc_server* s, const pc_request& q) {
  if (s->precision == PC_PREC_MIXED && (q.dtype != PC_F32 || q.n > 256)) return PC_PREC_FP64;
  return s->precision;
}

// byte sizes of a request's output blocks (a POD's U, S, VT stay resident
// as its basis; a projection or reconstruction has one block)
int block_sizes(const pc_request& q, const Basis* b, int64_t sz[3]) {
  if (q.kind == PC_REQ_POD) {
    const int64_t e = esize(q.dtype);
    sz[0] = q.m * q.n * e;
    sz[1] = q.n * e;
    sz[2] = q.n * q.n * e;
    return 3;
  }
  const int64_t k = q.modes > 0 && q.modes < b->n ? q.modes : b->n;
  sz[0] = q.kind == PC_REQ_PROJECT ? k * q.n * esize(b->dtype) : b->m * b->n * esize(b->dtype);
  return 1;
}

// arena bytes a request's outputs take (whole granules)
int64_t bytes_needed(const pc_request& q, const Basis* b) {
  int64_t sz[3];
  const int nb = block_sizes(q, b, sz);
  int64_t t = 0;
  for (int i = 0; i < nb; i++) t += Arena::round(sz[i]);
  return t;
}

// evict the least recently used idle basis (no queued or running users);
// false when every resident basis is pinned
bool evict_one(pc_server* s) {
  int64_t victim = -1;
  double oldest = 0;
  for (auto& kv : s->bases)
    if (kv.second.pins == 0 && (victim < 0 || kv.second.last_used < oldest)) {
      victim = kv.first;
      oldest = kv.second.last_used;
    }
  if (victim < 0) return false;
  Basis& b = s->bases[victim];
  s->arena.release(b.U);
  s->arena.release(b.S);
  s->arena.release(b.VT);
  s->resident_bytes -= b.nbytes;
  s->bases.erase(victim);
  s->n_evictions++;
  return true;
}

// carve a request's output blocks from the arena, evicting idle bases until
// they fit; false (nothing held) when they cannot fit yet
bool reserve(pc_server* s, Req* r, const Basis* b) {
  int64_t sz[3];
  const int nb = block_sizes(r->q, b, sz);
  for (;;) {
    void* p[3] = {nullptr, nullptr, nullptr};
    int got = 0;
    while (got < nb && (p[got] = s->arena.alloc(sz[got])) != nullptr) got++;
    if (got == nb) {
      if (r->q.kind == PC_REQ_POD)
        for (int i = 0; i < 3; i++) r->blk[i] = p[i];
      else
        r->out = p[0];
      return true;
    }
    for (int i = 0; i < got; i++) s->arena.release(p[i]);
    if (!evict_one(s)) return false;
  }
}

void finish_rejected(pc_server* s, Req* r, int code, const std::string& msg) {
  if (r->ready) {
    s->events.push_back(r->ready);
    r->ready = nullptr;
  }
  r->status = code;
  r->err = msg;
  r->t_done = s->now();
  s->n_rejected++;
}

// Requests that use a resident basis may pass the skip-ahead window: they pin
// their basis from submission on, so a head request waiting for that memory
// would otherwise wait on requests queued behind it.
void admit(pc_server* s) {
  size_t i = 0;
  while (!s->idle.empty() && i < s->pending.size()) {
    Req* r = s->pending[i];
    if (int(i) > s->skip_ahead && r->q.kind == PC_REQ_POD) {
      i++;
      continue;
    }
    Basis* b = nullptr;
    if (r->q.kind != PC_REQ_POD) b = &s->bases[r->q.basis];  // pinned at submit: still resident
    const int64_t need = bytes_needed(r->q, b);
    if (need > s->arena.size) {
      s->pending.erase(s->pending.begin() + i);
      if (b) b->pins--;
      finish_rejected(s, r, PC_ERR_NOMEM, "request outputs exceed the server budget");
      continue;
    }
    if (!reserve(s, r, b)) {
      i++;
      continue;
    }
    s->pending.erase(s->pending.begin() + i);
    r->nbytes = need;
    r->t_admit = s->now();
    s->inflight_bytes += need;
    const int slot = s->idle.back();
    s->idle.pop_back();
    r->slot = slot;
    s->work[slot] = r;
    s->n_inflight++;
  }
  s->cv.notify_all();
}

// run one request on a slot (outside the server lock); returns a pc_status
int run_request(pc_server* s, int slot, Req* r, const Basis* b, Basis* made) {
  pc_ctx* ctx = s->ctx[slot];
  hipStream_t st = s->streams[slot];
  const pc_request& q = r->q;
  // X may still be in flight on its producer's stream
  if (r->ready && hipStreamWaitEvent(st, r->ready, 0) != hipSuccess) {
    r->err = "hipStreamWaitEvent on the input's stream failed";
    return PC_ERR_HIP;
  }
  if (q.kind == PC_REQ_POD) {
    const int64_t e = esize(q.dtype);
    void *U = r->blk[0], *S = r->blk[1], *VT = r->blk[2];
    int rc = pc_tsqr_svd(ctx, pc_dtype(q.dtype), q.X, q.ldx, q.m, q.n, s->center, U, q.n, S, VT, q.n,
                         pod_precision(s, q));
    if (rc == PC_OK && hipStreamSynchronize(st) != hipSuccess) rc = PC_ERR_HIP;
    if (rc != PC_OK) {
      r->err = pc_last_error(ctx);
      (void)hipStreamSynchronize(st);
      return rc;
    }
    made->U = U;
    made->S = S;
    made->VT = VT;
    made->m = q.m;
    made->n = q.n;
    made->dtype = q.dtype;
    made->nbytes = r->nbytes;
    return PC_OK;
  }
  const int64_t k = q.modes > 0 && q.modes < b->n ? q.modes : b->n;
  const int64_t e = esize(b->dtype);
  int rc;
  if (q.kind == PC_REQ_PROJECT) {
    if (q.dtype != b->dtype || q.m != b->m) {
      r->err = "project: X must have the basis' rows and element type";
      return PC_ERR_SHAPE;
    }
    r->out_rows = k;
    r->out_cols = q.n;
    // projections and reconstructions run FP64 MFMA on the stored data
    rc = pc_project(ctx, pc_dtype(b->dtype), b->U, b->n, q.X, q.ldx, b->m, k, q.n, r->out, q.n, PC_PREC_FP64);
  } else {
    r->out_rows = b->m;
    r->out_cols = b->n;
    rc = pc_reconstruct(ctx, pc_dtype(b->dtype), b->U, b->n, b->S, b->VT, b->n, b->m, k, b->n, r->out, b->n,
                        PC_PREC_FP64);
  }
  if (rc == PC_OK && hipStreamSynchronize(st) != hipSuccess) rc = PC_ERR_HIP;
  if (rc != PC_OK) {
    r->err = pc_last_error(ctx);
    (void)hipStreamSynchronize(st);
  }
  return rc;
}

void worker(pc_server* s, int slot) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv.wait(lk, [&] { return s->work[slot] != nullptr || s->stop; });
    Req* r = s->work[slot];
    if (!r) return;  // stopped with nothing assigned
    const Basis* b = r->q.kind == PC_REQ_POD ? nullptr : &s->bases[r->q.basis];
    lk.unlock();
    Basis made;
    const int rc = run_request(s, slot, r, b, &made);
    lk.lock();
    r->t_done = s->now();
    s->inflight_bytes -= r->nbytes;
    if (r->q.kind != PC_REQ_POD) {
      Basis& bb = s->bases[r->q.basis];
      bb.pins--;
      bb.last_used = r->t_done;
    }
    if (rc != PC_OK) {
      // the stream was synchronised: the blocks are free to reuse
      for (void*& p : r->blk) s->arena.release(p), p = nullptr;
      s->arena.release(r->out);
      r->out = nullptr;
    } else if (r->q.kind == PC_REQ_POD) {
      for (void*& p : r->blk) p = nullptr;  // now the basis' blocks
      made.last_used = r->t_done;
      r->basis = s->next_basis++;
      s->bases[r->basis] = made;
      s->resident_bytes += made.nbytes;
    } else if (rc == PC_OK) {
      // a projection / reconstruction output is counted until released
      s->resident_bytes += r->nbytes;
    }
    if (r->ready) {
      s->events.push_back(r->ready);  // the slot stream was synchronised: the wait is over
      r->ready = nullptr;
    }
    r->status = rc;
    s->work[slot] = nullptr;
    s->idle.push_back(slot);
    s->n_inflight--;
    s->n_completed++;
    admit(s);
  }
}

}  // namespace

extern "C" int pc_serve_create(int device, int slots, int64_t budget_bytes, int skip_ahead, int precision, int center,
                               int64_t m_max, int64_t n_max, pc_server** out) {
  if (!out) return PC_ERR_INVALID;
  *out = nullptr;
  if (slots < 1 || slots > 64 || budget_bytes <= 0 || skip_ahead < 0 || m_max < 1 || n_max < 1 ||
      (precision != PC_PREC_FP64 && precision != PC_PREC_MIXED) || center < 0 || center > 2)
    return PC_ERR_INVALID;
  pc_server* s = new pc_server();
  s->device = device;
  s->nslots = slots;
  s->budget = budget_bytes;
  s->skip_ahead = skip_ahead;
  s->precision = precision;
  s->center = center;
  s->m_max = m_max;
  s->n_max = n_max;
  s->t0 = std::chrono::steady_clock::now();
  s->work.assign(slots, nullptr);
  auto fail = [&](int code) {
    for (pc_ctx* c : s->ctx) (void)pc_ctx_destroy(c);
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    delete s;
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(PC_ERR_HIP);

  // one context per slot, its workspace reserved for the largest request of
  // either precision a POD may take (a reservation now, not a reallocation
  // in the middle of serving)
  int64_t ws = pc_workspace_query(0, PC_F32, m_max, n_max, 0, PC_PREC_FP64);
  if (precision == PC_PREC_MIXED && n_max > 0) {
    const int64_t wm = pc_workspace_query(0, PC_F32, m_max, n_max < 256 ? n_max : 256, 0, PC_PREC_MIXED);
    if (wm > ws) ws = wm;
  }
  const int64_t wd = pc_workspace_query(0, PC_F64, m_max, n_max, 0, PC_PREC_FP64);
  if (wd > ws) ws = wd;
  if (ws < 0) return fail(PC_ERR_SHAPE);
  // the budget covers the slots' workspaces too: what is left is for outputs
  // and resident bases
  s->budget = budget_bytes - int64_t(slots) * ws;
  if (s->budget <= 0) return fail(PC_ERR_NOMEM);
  // the output budget is reserved now, as one allocation
  if (hipMalloc(&s->arena_mem, size_t(s->budget)) != hipSuccess) {
    (void)hipGetLastError();
    s->arena_mem = nullptr;
    return fail(PC_ERR_NOMEM);
  }
  s->arena.init(s->arena_mem, s->budget);
  for (int i = 0; i < slots; i++) {
    pc_ctx* c = nullptr;
    hipStream_t st = nullptr;
    if (pc_ctx_create(device, 0, 0, &c) != PC_OK) return fail(PC_ERR_HIP);
    s->ctx.push_back(c);
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail(PC_ERR_HIP);
    s->streams.push_back(st);
    if (pc_set_stream(c, st) != PC_OK) return fail(PC_ERR_HIP);
    if (pc_ws_reserve(c, ws) != PC_OK) return fail(PC_ERR_NOMEM);
  }
  for (int i = slots - 1; i >= 0; i--) s->idle.push_back(i);
  for (int i = 0; i < slots; i++) s->threads.emplace_back(worker, s, i);
  *out = s;
  return PC_OK;
}

extern "C" int pc_serve_destroy(pc_server* s) {
  if (!s) return PC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // queued requests are cancelled; those running finish
    for (Req* r : s->pending) {
      if (r->q.kind != PC_REQ_POD) s->bases[r->q.basis].pins--;
      if (r->ready) s->events.push_back(r->ready);
      r->ready = nullptr;
      r->status = PC_ERR_INVALID;
      r->err = "server destroyed";
    }
    s->pending.clear();
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& t : s->threads) t.join();
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : s->reqs) delete kv.second;
  (void)hipFree(s->arena_mem);  // outputs and bases
  for (hipEvent_t ev : s->events) (void)hipEventDestroy(ev);
  for (pc_ctx* c : s->ctx) (void)pc_ctx_destroy(c);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  delete s;
  return PC_OK;
}

extern "C" int pc_serve_submit(pc_server* s, const pc_request* q, int64_t* ticket) {
  if (!s || !q || !ticket) return PC_ERR_INVALID;
  if (q->kind < PC_REQ_POD || q->kind > PC_REQ_RECONSTRUCT) return serve_fail(s, PC_ERR_INVALID, "bad request kind");
  if (q->kind != PC_REQ_RECONSTRUCT) {
    if (!q->X || q->m < 1 || q->n < 1 || q->ldx < q->n) return serve_fail(s, PC_ERR_INVALID, "bad input matrix");
    if (q->dtype != PC_F32 && q->dtype != PC_F64) return serve_fail(s, PC_ERR_UNSUPPORTED, "real float32 / float64 only");
  }
  if (q->kind == PC_REQ_POD && (q->m > s->m_max || q->n > s->n_max || q->m < q->n))
    return serve_fail(s, PC_ERR_SHAPE, "POD request beyond the server's (m_max, n_max) or m < n");
  if (q->kind == PC_REQ_PROJECT && q->n > s->n_max)
    return serve_fail(s, PC_ERR_SHAPE, "project request wider than n_max");
  std::lock_guard<std::mutex> lk(s->mu);
  Req* r = new Req();
  r->q = *q;
  r->ticket = s->next_ticket++;
  r->t_submit = s->now();
  s->reqs[r->ticket] = r;
  *ticket = r->ticket;
  if (q->kind != PC_REQ_POD) {
    auto it = s->bases.find(q->basis);
    if (it == s->bases.end()) {
      finish_rejected(s, r, PC_ERR_INVALID, "basis is not resident");
      return PC_OK;  // the request exists and reports its error
    }
    it->second.pins++;  // a basis with queued or running users is never evicted
  }
  if (q->kind != PC_REQ_RECONSTRUCT) {
    // mark what X's producer has queued so far; the slot waits for it (the
    // event lives on the server's device, made current for the call)
    int prev = s->device;
    (void)hipGetDevice(&prev);
    if (prev != s->device) (void)hipSetDevice(s->device);
    struct Restore {
      int dev, cur;
      ~Restore() {
        if (dev != cur) (void)hipSetDevice(dev);
      }
    } restore{prev, s->device};
    if (s->events.empty()) {
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        if (q->kind != PC_REQ_POD) s->bases[q->basis].pins--;
        finish_rejected(s, r, PC_ERR_HIP, "hipEventCreate failed");
        return PC_OK;
      }
      s->events.push_back(ev);
    }
    r->ready = s->events.back();
    s->events.pop_back();
    if (hipEventRecord(r->ready, static_cast<hipStream_t>(q->stream)) != hipSuccess) {
      if (q->kind != PC_REQ_POD) s->bases[q->basis].pins--;
      finish_rejected(s, r, PC_ERR_HIP, "hipEventRecord on the request's stream failed");
      return PC_OK;
    }
  }
  s->pending.push_back(r);
  admit(s);
  return PC_OK;
}

extern "C" int pc_serve_wait(pc_server* s, int64_t ticket, double timeout_s, pc_result* out) {
  if (!s) return PC_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return serve_fail(s, PC_ERR_INVALID, "unknown ticket");
  Req* r = it->second;
  auto done = [&] { return r->status != PC_PENDING; };
  if (timeout_s < 0) s->cv.wait(lk, done);
  else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return PC_PENDING;
  if (out) {
    out->status = r->status;
    out->basis = r->basis;
    out->out = r->out;
    out->out_rows = r->out_rows;
    out->out_cols = r->out_cols;
    out->out_ld = r->out_cols;
    out->t_submit = r->t_submit;
    out->t_admit = r->t_admit;
    out->t_done = r->t_done;
  }
  if (r->status != PC_OK) s->err = r->err;
  return PC_OK;
}

extern "C" int pc_serve_release(pc_server* s, int64_t ticket) {
  if (!s) return PC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return serve_fail(s, PC_ERR_INVALID, "unknown ticket");
  Req* r = it->second;
  if (r->status == PC_PENDING) return serve_fail(s, PC_ERR_INVALID, "request still pending");
  if (r->out) {
    s->arena.release(r->out);
    s->resident_bytes -= r->nbytes;
  }
  delete r;
  s->reqs.erase(it);
  admit(s);
  return PC_OK;
}

extern "C" int pc_serve_drain(pc_server* s, double timeout_s) {
  if (!s) return PC_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] { return s->pending.empty() && s->n_inflight == 0; };
  if (timeout_s < 0) s->cv.wait(lk, idle);
  else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), idle)) return PC_PENDING;
  return PC_OK;
}

extern "C" int pc_serve_basis(pc_server* s, int64_t basis, int pin, pc_basis_view* out) {
  if (!s || pin < -1 || pin > 1) return PC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->bases.find(basis);
  if (it == s->bases.end()) return serve_fail(s, PC_ERR_INVALID, "basis is not resident");
  Basis& b = it->second;
  if (pin < 0 && b.pins <= 0) return serve_fail(s, PC_ERR_INVALID, "basis is not pinned");
  b.pins += pin;
  if (pin < 0) admit(s);
  if (out) {
    out->U = b.U;
    out->S = b.S;
    out->VT = b.VT;
    out->m = b.m;
    out->n = b.n;
    out->dtype = b.dtype;
  }
  return PC_OK;
}

extern "C" int pc_serve_bases(pc_server* s, int64_t* ids, int64_t cap, int64_t* count) {
  if (!s || !count || cap < 0 || (cap > 0 && !ids)) return PC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  int64_t k = 0;
  for (auto& kv : s->bases)
    if (k < cap) ids[k++] = kv.first;
  *count = int64_t(s->bases.size());
  return PC_OK;
}

extern "C" int pc_serve_stats(pc_server* s, int64_t* v, int n) {
  if (!s || !v || n < 0) return PC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t all[9] = {s->n_completed, s->n_rejected, s->n_evictions, s->inflight_bytes,
                          s->resident_bytes, int64_t(s->pending.size()), s->n_inflight, int64_t(s->bases.size()),
                          s->budget};
  for (int i = 0; i < n && i < 9; i++) v[i] = all[i];
  return PC_OK;
}

extern "C" const char* pc_serve_last_error(const pc_server* s) { return s ? s->err.c_str() : "null server"; }
