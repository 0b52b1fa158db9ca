// Context, workspace and error plumbing of the POD core C ABI.
#include <cstdarg>
#include "common.hpp"

static thread_local std::string g_noctx_err;

int pc_fail(pc_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  else g_noctx_err = buf;
  return code;
}

extern "C" const char* pc_version(void) { return "pyloworder_amd podcore 0.1 (gfx950)"; }

extern "C" int pc_ctx_create(int device, int64_t budget, uint32_t flags, pc_ctx** out) {
  if (!out) return pc_fail(nullptr, PC_ERR_INVALID, "pc_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return pc_fail(nullptr, PC_ERR_HIP, "pc_ctx_create: no HIP device visible");
  if (device < 0 || device >= ndev)
    return pc_fail(nullptr, PC_ERR_INVALID, "pc_ctx_create: device %d out of range", device);
  pc_ctx* c = new pc_ctx();
  c->device = device;
  c->budget = budget;
  c->flags = flags;
  if (hipSetDevice(device) != hipSuccess) { delete c; return pc_fail(nullptr, PC_ERR_HIP, "hipSetDevice failed"); }
  if (hipMalloc(&c->dflag, 64 * sizeof(int)) != hipSuccess) { delete c; return pc_fail(nullptr, PC_ERR_HIP, "hipMalloc flags failed"); }
  if (hipHostMalloc(&c->hflag, 64 * sizeof(int)) != hipSuccess) {
    (void)hipFree(c->dflag); delete c; return pc_fail(nullptr, PC_ERR_HIP, "hipHostMalloc failed");
  }
  if (pc_scratch_attach(device) != PC_OK) {
    (void)hipFree(c->dflag); (void)hipHostFree(c->hflag); delete c;
    return pc_fail(nullptr, PC_ERR_HIP, "pc_ctx_create: address of the static scratch not available");
  }
  *out = c;
  return PC_OK;
}

extern "C" int pc_ctx_destroy(pc_ctx* ctx) {
  if (!ctx) return PC_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  for (int i = 0; i < pc_ctx::kGroupStreams; i++) {
    if (ctx->sub[i]) {
      (void)pc_ctx_destroy(ctx->sub[i]);
      ctx->sub[i] = nullptr;
    }
  }
  for (int i = 0; i < pc_ctx::kGroupStreams; i++) {
    if (ctx->gs[i]) {
      (void)hipStreamSynchronize(ctx->gs[i]);
      (void)hipStreamDestroy(ctx->gs[i]);
    }
  }
  for (hipEvent_t e : ctx->gev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  for (hipEvent_t e : ctx->sev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->dflag) (void)hipFree(ctx->dflag);
  if (ctx->hflag) (void)hipHostFree(ctx->hflag);
  delete ctx;
  return PC_OK;
}

extern "C" int pc_set_stream(pc_ctx* ctx, void* s) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_set_stream: ctx is NULL");
  ctx->stream = reinterpret_cast<hipStream_t>(s);
  return PC_OK;
}

extern "C" const char* pc_last_error(const pc_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_noctx_err.c_str();
}

extern "C" int64_t pc_workspace_bytes(const pc_ctx* ctx) { return ctx ? ctx->ws_bytes : 0; }

extern "C" int pc_synchronize(pc_ctx* ctx) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_synchronize: ctx is NULL");
  PC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PC_OK;
}

int pc_ws_reserve(pc_ctx* ctx, int64_t bytes) {
  if (bytes <= ctx->ws_bytes) return PC_OK;
  if (ctx->ws_fixed)
    return pc_fail(ctx, PC_ERR_NOMEM, "the call needs %lld bytes of workspace, its fixed scratch holds %lld",
                   (long long)bytes, (long long)ctx->ws_bytes);
  // an entry point holding the front of the buffer (ws_skip > 0) reserves its
  // nested needs up front: growing here would free memory still in use
  if (ctx->ws_skip > 0 && ctx->ws)
    return pc_fail(ctx, PC_ERR_INVALID, "internal: workspace must grow (%lld -> %lld bytes) while its front is held",
                   (long long)ctx->ws_bytes, (long long)bytes);
  if (ctx->budget > 0 && bytes > ctx->budget)
    return pc_fail(ctx, PC_ERR_NOMEM, "workspace of %lld bytes exceeds the HBM budget of %lld bytes",
                   (long long)bytes, (long long)ctx->budget);
  // the old buffer may still be in use by queued work on the stream
  PC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->ws) {
    PC_HIP(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  int64_t want = round_up(bytes + (bytes >> 3), int64_t(1) << 20);
  if (ctx->budget > 0 && want > ctx->budget) want = bytes;
  hipError_t e = hipMalloc(&ctx->ws, want);
  if (e != hipSuccess)
    return pc_fail(ctx, PC_ERR_NOMEM, "hipMalloc(%lld) for workspace failed: %s", (long long)want,
                   hipGetErrorString(e));
  ctx->ws_bytes = want;
  return PC_OK;
}

extern "C" int pc_truncation_residual(const double* S, int64_t n, double r, int64_t* N) {
  // pyLOM/vmmath/truncation.py:17-39.  r > 0: first i with ||S[i:]||/||S|| < r;
  // r <= 0: first i+1 with sum(S[:i+1])/sum(S) > |r|.
  if (!S || !N || n < 0) return pc_fail(nullptr, PC_ERR_INVALID, "pc_truncation_residual: bad argument");
  int64_t cnt = 0;
  if (r > 0) {
    // tail norms from the back so each ||S[i:]|| costs O(1)
    std::vector<double> tail(n + 1, 0.0);
    for (int64_t i = n - 1; i >= 0; --i) tail[i] = tail[i + 1] + S[i] * S[i];
    double norm = sqrt(tail[0]);
    for (int64_t i = 0; i < n; ++i) {
      double acc = sqrt(tail[i]) / norm;
      if (acc < r) break;
      ++cnt;
    }
  } else {
    double ar = fabs(r), sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum += S[i];
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      acc += S[i] / sum;
      ++cnt;
      if (acc > ar) break;
    }
  }
  *N = cnt;
  return PC_OK;
}

// ---------------------------------------------------------------------------
// per-stage timing: events are recorded on the context stream around each
// streaming pass; pc_profile_read synchronises and folds them in.
// ---------------------------------------------------------------------------
int pc_prof_begin(pc_ctx* ctx, hipEvent_t* ev) {
  *ev = nullptr;
  if (!ctx->prof) return 0;
  if (hipEventCreate(ev) != hipSuccess) { *ev = nullptr; return 0; }
  (void)hipEventRecord(*ev, ctx->stream);
  return 1;
}

void pc_prof_end(pc_ctx* ctx, int stage, hipEvent_t ev0) {
  if (!ctx->prof || !ev0) return;
  hipEvent_t ev1;
  if (hipEventCreate(&ev1) != hipSuccess) return;
  (void)hipEventRecord(ev1, ctx->stream);
  ctx->prof_pending.push_back({stage, {ev0, ev1}});
}

extern "C" int pc_profile(pc_ctx* ctx, int enable) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_profile: ctx is NULL");
  ctx->prof = enable ? 1 : 0;
  for (int i = 0; i < 8; i++) { ctx->prof_ms[i] = 0; ctx->prof_n[i] = 0; }
  return PC_OK;
}

extern "C" int pc_profile_read(pc_ctx* ctx, double* ms, int64_t* counts, int nstages) {
  if (!ctx || !ms || !counts) return pc_fail(ctx, PC_ERR_INVALID, "pc_profile_read: NULL argument");
  PC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& e : ctx->prof_pending) {
    float t = 0;
    if (hipEventElapsedTime(&t, e.second.first, e.second.second) == hipSuccess) {
      ctx->prof_ms[e.first] += t;
      ctx->prof_n[e.first] += 1;
    }
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  ctx->prof_pending.clear();
  for (int i = 0; i < nstages && i < 8; i++) { ms[i] = ctx->prof_ms[i]; counts[i] = ctx->prof_n[i]; }
  return PC_OK;
}
