// Shared definitions of the POD core library (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <mutex>

#include "../../include/podcore.h"

#define PC_HD __host__ __device__ __forceinline__
#define PC_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// complex FP64/FP32 scalars (interleaved re, im — numpy's complex layout)
// ---------------------------------------------------------------------------
struct cdouble {
  double x, y;
  cdouble() = default;
  PC_HD cdouble(double a) : x(a), y(0) {}
  PC_HD cdouble(double a, double b) : x(a), y(b) {}
};
struct cfloat {
  float x, y;
};
PC_HD cdouble operator+(cdouble a, cdouble b) { return cdouble(a.x + b.x, a.y + b.y); }
PC_HD cdouble operator-(cdouble a, cdouble b) { return cdouble(a.x - b.x, a.y - b.y); }
PC_HD cdouble operator-(cdouble a) { return cdouble(-a.x, -a.y); }
PC_HD cdouble operator*(cdouble a, cdouble b) { return cdouble(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
PC_HD cdouble operator*(double s, cdouble a) { return cdouble(s * a.x, s * a.y); }
PC_HD cdouble operator*(cdouble a, double s) { return cdouble(s * a.x, s * a.y); }
PC_HD cdouble operator/(cdouble a, double s) { return cdouble(a.x / s, a.y / s); }
PC_HD cdouble& operator+=(cdouble& a, cdouble b) { a.x += b.x; a.y += b.y; return a; }
PC_HD cdouble& operator-=(cdouble& a, cdouble b) { a.x -= b.x; a.y -= b.y; return a; }

// scalar traits used by the small dense kernels (real double or cdouble)
PC_HD double conj_(double a) { return a; }
PC_HD cdouble conj_(cdouble a) { return cdouble(a.x, -a.y); }
PC_HD double abs2_(double a) { return a * a; }
PC_HD double abs2_(cdouble a) { return a.x * a.x + a.y * a.y; }
PC_HD double real_(double a) { return a; }
PC_HD double real_(cdouble a) { return a.x; }
PC_HD double cmul_re(double a, double b) { return a * b; }
// fma-style accumulate conj(a)*b
PC_HD double dotc_(double a, double b) { return a * b; }
PC_HD cdouble dotc_(cdouble a, cdouble b) { return cdouble(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
// 1/a without forming |a|^2 (which underflows for |a| < 1e-154): Smith's division
PC_HD double inv_(double a) { return 1.0 / a; }
PC_HD cdouble inv_(cdouble a) {
  if (fabs(a.x) >= fabs(a.y)) {
    double r = a.y / a.x, d = a.x + a.y * r;
    return cdouble(1.0 / d, -r / d);
  }
  double r = a.x / a.y, d = a.x * r + a.y;
  return cdouble(r / d, -1.0 / d);
}
// unit-modulus phase of a non-zero scalar, computed by division (no reciprocal overflow)
PC_HD double phase_(double c, double g) { return c / g; }
PC_HD cdouble phase_(cdouble c, double g) { return cdouble(c.x / g, c.y / g); }
// total order key for ranking: NaN sorts as -inf (descending ranks) ...
PC_HD double key_desc(double v) { return v == v ? v : -__builtin_huge_val(); }
// ... and as +inf (ascending ranks)
PC_HD double key_asc(double v) { return v == v ? v : __builtin_huge_val(); }
template <class T> PC_HD T mk_(double re, double im);
template <> PC_HD double mk_<double>(double re, double) { return re; }
template <> PC_HD cdouble mk_<cdouble>(double re, double im) { return cdouble(re, im); }

// Batched launches of the small-stage kernels: blockIdx.z selects the problem
// and each pointer moves by a fixed byte stride per problem (zb = 0 and
// gridDim.z = 1 for a single problem).  Null pointers stay null.
template <class P> PC_DEV P* zoff(P* p, int64_t zb) {
  return p ? (P*)((const char*)p + int64_t(blockIdx.z) * zb) : p;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct pc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t budget = 0;     // HBM budget for the workspace, bytes (0 = unlimited)
  uint32_t flags = 0;
  char* ws = nullptr;     // device workspace
  int64_t ws_bytes = 0;
  int64_t ws_skip = 0;    // bytes at the front of ws held by the calling entry point
  bool ws_fixed = false;  // ws is borrowed for one call (pc_tsqr_svd_noalloc): never reallocated
  int* dflag = nullptr;   // device status words (error flags, ranks)
  int* hflag = nullptr;   // pinned host mirror
  int* gstat = nullptr;   // grouped solve: where the current problem leaves its status words
  // grouped solve: internal streams (default priority) the problems are spread
  // over, and the fork / join events; created on first use
  static constexpr int kGroupStreams = 16;  // upper bound; see group_streams()
  hipStream_t gs[kGroupStreams] = {};
  hipEvent_t gev[kGroupStreams + 1] = {};
  // one sub-context per internal stream (own workspace and status words),
  // driven by its own host thread during a grouped solve
  pc_ctx* sub[kGroupStreams] = {};
  // set on those sub-contexts: the small stage takes its one-launch kernels
  // (same results), since launches, not one CU's step time, bound it when
  // many solves share the GPU
  bool concurrent = false;
  // side stream (default priority) and fork / join events for the small
  // stage's off-path triangular inverses of a lone solve; created on first use
  hipStream_t side = nullptr;
  hipEvent_t sev[2] = {};
  std::string err;
  // launch-geometry tuning of the product pass (pc_autotune)
  int autotune = 0;
  // optional per-stage timing (pc_profile): events around each streaming pass
  int prof = 0;
  double prof_ms[8] = {0};
  int64_t prof_n[8] = {0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
};

// stage ids of pc_profile_read
enum { PCP_GRAM = 0, PCP_ZGRAM = 1, PCP_GEMM = 2, PCP_SMALL = 3, PCP_OTHER = 4 };
int pc_prof_begin(pc_ctx* ctx, hipEvent_t* ev);
void pc_prof_end(pc_ctx* ctx, int stage, hipEvent_t ev0);

// status helpers --------------------------------------------------------------
int pc_fail(pc_ctx* ctx, int code, const char* fmt, ...);
#define PC_HIP(ctx, call)                                                        \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return pc_fail(ctx, PC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)
#define PC_CHECK(ctx, expr)                                                      \
  do {                                                                           \
    int rc_ = (expr);                                                            \
    if (rc_ != PC_OK) return rc_;                                                \
  } while (0)

// grows ctx->ws to at least `bytes` (sync + realloc when growing)
int pc_ws_reserve(pc_ctx* ctx, int64_t bytes);
// takes, once per device and with that device current, the address of the
// library's static device scratch (solve.hip)
int pc_scratch_attach(int device);

// bump allocator over the workspace (256-byte aligned)
struct WsCursor {
  char* base;
  int64_t off = 0;
  int64_t cap;
  template <class T> T* take(int64_t count) {
    off = (off + 255) & ~int64_t(255);
    T* p = reinterpret_cast<T*>(base + off);
    off += count * int64_t(sizeof(T));
    return p;
  }
};

PC_HD int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
PC_HD int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// number of row chunks a streaming pass over m rows uses.  It depends only
// on (m, rows-per-tile) so that every reduction order — and thus every result —
// is a function of the problem alone, never of what else runs on the GPU.
inline int stream_chunks(int64_t m, int tile_rows) {
  int64_t tiles = ceil_div(m, tile_rows);
  int64_t per = ceil_div(tiles, 1024);  // at most 1024 chunks
  if (per < 4) per = 4;                 // at least 4 tiles per chunk
  int64_t chunks = ceil_div(tiles, per);
  return chunks < 1 ? 1 : int(chunks);
}

// element sizes
inline int64_t dtype_bytes(pc_dtype t) {
  switch (t) {
    case PC_F32: return 4;
    case PC_F64: return 8;
    case PC_C64: return 8;
    case PC_C128: return 16;
  }
  return 0;
}
inline bool dtype_complex(pc_dtype t) { return t == PC_C64 || t == PC_C128; }
inline bool dtype_single(pc_dtype t) { return t == PC_F32 || t == PC_C64; }
