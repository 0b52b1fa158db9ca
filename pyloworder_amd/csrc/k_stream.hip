// Streaming (HBM-bound, non-MFMA) kernels: row statistics, centering,
// vecmat, transpose, difference norms.
#include "common.hpp"

namespace {

template <class T> PC_DEV double to_d(T v) { return double(v); }

// One wave per row: temporal mean (pyLOM/vmmath/averaging.py:17-29) and
// population variance (averaging.py:46-59), both accumulated in FP64.
template <class T, class TO = T>
__global__ __launch_bounds__(256) void k_row_stats(const T* __restrict__ X, int64_t ldx, int64_t m,
                                                   int64_t n, TO* __restrict__ mean, TO* __restrict__ var) {
  int lane = threadIdx.x & 63;
  int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= m) return;
  const T* x = X + row * ldx;
  double s = 0.0;
  for (int64_t j = lane; j < n; j += 64) s += to_d(x[j]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  double mu = s / double(n);
  if (var) {
    double v = 0.0;
    for (int64_t j = lane; j < n; j += 64) { double d = to_d(x[j]) - mu; v += d * d; }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) var[row] = TO(v / double(n));
  }
  if (lane == 0 && mean) mean[row] = TO(mu);
}

// Y = (X - mean) [/ var]   (averaging.py:31-44 and 61-75)
template <class T>
__global__ __launch_bounds__(256) void k_row_center(const T* X, int64_t ldx, int64_t m, int64_t n,
                                                    const T* __restrict__ mean, const T* __restrict__ var,
                                                    T* Y, int64_t ldy) {
  int lane = threadIdx.x & 63;
  int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= m) return;
  T mu = mean[row];
  T vv = var ? var[row] : T(1);
  const T* x = X + row * ldx;
  T* y = Y + row * ldy;
  if (var) {
    for (int64_t j = lane; j < n; j += 64) y[j] = (x[j] - mu) / vv;
  } else {
    for (int64_t j = lane; j < n; j += 64) y[j] = x[j] - mu;
  }
}

// C[i,:] = v[i] * A[i,:]  (pyLOM/vmmath/maths.py:112-129); T = element type,
// V = type of v (real for real data).
template <class T, class V>
__global__ __launch_bounds__(256) void k_vecmat(const V* __restrict__ v, const T* A, int64_t lda, int64_t m,
                                                int64_t n, T* C, int64_t ldc) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int64_t total = m * n;
  for (; idx < total; idx += int64_t(gridDim.x) * 256) {
    int64_t i = idx / n, j = idx - i * n;
    C[i * ldc + j] = A[i * lda + j] * v[i];
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_vecmat_c(const T* __restrict__ v, const T* A, int64_t lda, int64_t m,
                                                  int64_t n, T* C, int64_t ldc) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int64_t total = m * n;
  for (; idx < total; idx += int64_t(gridDim.x) * 256) {
    int64_t i = idx / n, j = idx - i * n;
    T a = A[i * lda + j], s = v[i];
    T c;
    c.x = a.x * s.x - a.y * s.y;
    c.y = a.x * s.y + a.y * s.x;
    C[i * ldc + j] = c;
  }
}

// B = A^T through a 32x32 LDS tile (maths.py:17-29)
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ A, int64_t lda, int64_t m, int64_t n,
                                                   T* __restrict__ B, int64_t ldb) {
  // 32 x 32 tiles, flattened (row tiles fastest) and taken grid-stride, so
  // neither dimension is limited by the grid
  __shared__ T tile[32][33];
  const int64_t tm = ceil_div(m, 32), tiles = tm * ceil_div(n, 32);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int64_t tt = blockIdx.x; tt < tiles; tt += gridDim.x) {
    const int64_t bi = (tt % tm) * 32, bj = (tt / tm) * 32;
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      int64_t i = bi + r, j = bj + tx;
      if (i < m && j < n) tile[r][tx] = A[i * lda + j];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      int64_t j = bj + r, i = bi + tx;
      if (i < m && j < n) B[j * ldb + i] = tile[tx][r];
    }
  }
}

// partial sums of |A-B|^2 and |A|^2 per block (RMSE / energy)
template <class T>
__global__ __launch_bounds__(256) void k_diff_norms(const T* A, int64_t lda, const T* B, int64_t ldb, int64_t m,
                                                    int64_t n, double* part) {
  double s1 = 0.0, s2 = 0.0;
  int64_t total = m * n;
  for (int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x; idx < total; idx += int64_t(gridDim.x) * 256) {
    int64_t i = idx / n, j = idx - i * n;
    double a = double(A[i * lda + j]), b = double(B[i * ldb + j]);
    s1 += (a - b) * (a - b);
    s2 += a * a;
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { r1[threadIdx.x] += r1[threadIdx.x + o]; r2[threadIdx.x] += r2[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = r1[0]; part[2 * blockIdx.x + 1] = r2[0]; }
}

__global__ void k_sum_pairs(const double* part, int nb, double* out) {
  // deterministic serial-per-lane reduction of a small number of partials
  __shared__ double r1[256], r2[256];
  double s1 = 0, s2 = 0;
  for (int i = threadIdx.x; i < nb; i += 256) { s1 += part[2 * i]; s2 += part[2 * i + 1]; }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { r1[threadIdx.x] += r1[threadIdx.x + o]; r2[threadIdx.x] += r2[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = r1[0]; out[1] = r2[0]; }
}

int check_mat(pc_ctx* ctx, const void* p, int64_t ld, int64_t m, int64_t n, const char* what) {
  if (m < 0 || n < 0) return pc_fail(ctx, PC_ERR_SHAPE, "%s: negative size", what);
  if (m > 0 && n > 0 && !p) return pc_fail(ctx, PC_ERR_INVALID, "%s: NULL pointer", what);
  if (m > 0 && ld < n) return pc_fail(ctx, PC_ERR_INVALID, "%s: leading dimension %lld < %lld", what, (long long)ld, (long long)n);
  return PC_OK;
}

}  // namespace

// FP64 statistics of float32 rows (the variance-normalised mixed path)
// Dataset.X row interleave (pyLOM/dataset.py:252-292): output row
// r = p nf + c (point p, field component c) is row idim + ndim_f p of field f,
// where component c is component idim of field f in the given field order.
constexpr int kMaxFields = 32;
struct FieldSet {
  const void* src[kMaxFields];
  int64_t ld[kMaxFields];
  int field_of[kMaxFields * 8];  // component -> field
  int idim_of[kMaxFields * 8];   // component -> its index inside the field
  int ndim[kMaxFields];
};
template <class T>
__global__ __launch_bounds__(256) void k_interleave(FieldSet fs, int nf, int64_t npoints, int64_t nt, T* X,
                                                    int64_t ldx) {
  const int64_t rows = npoints * nf;
  for (int64_t r = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); r < rows; r += int64_t(gridDim.x) * 4) {
    const int64_t p = r / nf;
    const int c = int(r - p * nf);
    const int f = fs.field_of[c];
    const T* s = static_cast<const T*>(fs.src[f]) + (int64_t(fs.idim_of[c]) + int64_t(fs.ndim[f]) * p) * fs.ld[f];
    T* o = X + r * ldx;
    for (int64_t j = threadIdx.x & 63; j < nt; j += 64) o[j] = s[j];
  }
}

void launch_row_stats_f32(hipStream_t st, const float* X, int64_t ldx, int64_t m, int64_t n, double* mean,
                          double* var) {
  dim3 g = dim3(unsigned(ceil_div(m, 4)));
  hipLaunchKernelGGL((k_row_stats<float, double>), g, dim3(256), 0, st, X, ldx, m, n, mean, var);
}

void launch_row_stats_f64(hipStream_t st, const double* X, int64_t ldx, int64_t m, int64_t n, double* mean,
                          double* var) {
  dim3 g = dim3(unsigned(ceil_div(m, 4)));
  hipLaunchKernelGGL((k_row_stats<double, double>), g, dim3(256), 0, st, X, ldx, m, n, mean, var);
}

int pc_check_mat(pc_ctx* ctx, const void* p, int64_t ld, int64_t m, int64_t n, const char* what) {
  return check_mat(ctx, p, ld, m, n, what);
}

extern "C" int pc_row_stats(pc_ctx* ctx, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int64_t n,
                            void* mean, void* var) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_row_stats: ctx is NULL");
  PC_CHECK(ctx, check_mat(ctx, X, ldx, m, n, "pc_row_stats X"));
  if (n <= 0) return pc_fail(ctx, PC_ERR_SHAPE, "pc_row_stats: n must be positive");
  if (m == 0) return PC_OK;
  dim3 g = dim3(unsigned(ceil_div(m, 4)));
  if (dt == PC_F64)
    hipLaunchKernelGGL(k_row_stats<double>, g, dim3(256), 0, ctx->stream, (const double*)X, ldx, m, n, (double*)mean, (double*)var);
  else if (dt == PC_F32)
    hipLaunchKernelGGL(k_row_stats<float>, g, dim3(256), 0, ctx->stream, (const float*)X, ldx, m, n, (float*)mean, (float*)var);
  else
    return pc_fail(ctx, PC_ERR_UNSUPPORTED, "pc_row_stats: real dtypes only");
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

extern "C" int pc_row_center(pc_ctx* ctx, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int64_t n,
                             const void* mean, const void* var, void* Y, int64_t ldy) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_row_center: ctx is NULL");
  PC_CHECK(ctx, check_mat(ctx, X, ldx, m, n, "pc_row_center X"));
  PC_CHECK(ctx, check_mat(ctx, Y, ldy, m, n, "pc_row_center Y"));
  if (m > 0 && !mean) return pc_fail(ctx, PC_ERR_INVALID, "pc_row_center: mean is NULL");
  if (m == 0 || n == 0) return PC_OK;
  dim3 g = dim3(unsigned(ceil_div(m, 4)));
  if (dt == PC_F64)
    hipLaunchKernelGGL(k_row_center<double>, g, dim3(256), 0, ctx->stream, (const double*)X, ldx, m, n,
                       (const double*)mean, (const double*)var, (double*)Y, ldy);
  else if (dt == PC_F32)
    hipLaunchKernelGGL(k_row_center<float>, g, dim3(256), 0, ctx->stream, (const float*)X, ldx, m, n,
                       (const float*)mean, (const float*)var, (float*)Y, ldy);
  else
    return pc_fail(ctx, PC_ERR_UNSUPPORTED, "pc_row_center: real dtypes only");
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

extern "C" int pc_vecmat(pc_ctx* ctx, pc_dtype dt, const void* v, const void* A, int64_t lda, int64_t m,
                         int64_t n, void* C, int64_t ldc) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_vecmat: ctx is NULL");
  PC_CHECK(ctx, check_mat(ctx, A, lda, m, n, "pc_vecmat A"));
  PC_CHECK(ctx, check_mat(ctx, C, ldc, m, n, "pc_vecmat C"));
  if (m == 0 || n == 0) return PC_OK;
  if (!v) return pc_fail(ctx, PC_ERR_INVALID, "pc_vecmat: v is NULL");
  int64_t blocks = ceil_div(m * n, 256);
  if (blocks > 65536) blocks = 65536;
  dim3 g = dim3(unsigned(blocks));
  switch (dt) {
    case PC_F64: hipLaunchKernelGGL((k_vecmat<double, double>), g, dim3(256), 0, ctx->stream, (const double*)v, (const double*)A, lda, m, n, (double*)C, ldc); break;
    case PC_F32: hipLaunchKernelGGL((k_vecmat<float, float>), g, dim3(256), 0, ctx->stream, (const float*)v, (const float*)A, lda, m, n, (float*)C, ldc); break;
    case PC_C64: hipLaunchKernelGGL((k_vecmat_c<cfloat>), g, dim3(256), 0, ctx->stream, (const cfloat*)v, (const cfloat*)A, lda, m, n, (cfloat*)C, ldc); break;
    case PC_C128: hipLaunchKernelGGL((k_vecmat_c<cdouble>), g, dim3(256), 0, ctx->stream, (const cdouble*)v, (const cdouble*)A, lda, m, n, (cdouble*)C, ldc); break;
  }
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

// Y[i][j] += alpha X[i][j] over an m x n block (complex data as interleaved reals)
template <class T>
__global__ __launch_bounds__(256) void k_axpy(T alpha, const T* __restrict__ X, int64_t ldx, T* __restrict__ Y,
                                              int64_t ldy, int64_t m, int64_t n) {
  const int64_t tot = m * n;
  for (int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x; idx < tot; idx += int64_t(gridDim.x) * 256) {
    const int64_t i = idx / n, j = idx - i * n;
    Y[i * ldy + j] += alpha * X[i * ldx + j];
  }
}

extern "C" int pc_axpy(pc_ctx* ctx, pc_dtype dt, double alpha, const void* X, int64_t ldx, void* Y, int64_t ldy,
                       int64_t m, int64_t n) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_axpy: ctx is NULL");
  PC_CHECK(ctx, check_mat(ctx, X, ldx, m, n, "pc_axpy X"));
  PC_CHECK(ctx, check_mat(ctx, Y, ldy, m, n, "pc_axpy Y"));
  if (m == 0 || n == 0) return PC_OK;
  const int w = dtype_complex(dt) ? 2 : 1;  // complex rows as 2n reals
  int64_t blocks = ceil_div(m * n * w, 256);
  if (blocks > 65536) blocks = 65536;
  dim3 g = dim3(unsigned(blocks));
  if (dt == PC_F64 || dt == PC_C128)
    hipLaunchKernelGGL(k_axpy<double>, g, dim3(256), 0, ctx->stream, alpha, (const double*)X, w * ldx, (double*)Y,
                       w * ldy, m, w * n);
  else
    hipLaunchKernelGGL(k_axpy<float>, g, dim3(256), 0, ctx->stream, float(alpha), (const float*)X, w * ldx,
                       (float*)Y, w * ldy, m, w * n);
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

extern "C" int pc_transpose(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, int64_t m, int64_t n,
                            void* B, int64_t ldb) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_transpose: ctx is NULL");
  PC_CHECK(ctx, check_mat(ctx, A, lda, m, n, "pc_transpose A"));
  PC_CHECK(ctx, check_mat(ctx, B, ldb, n, m, "pc_transpose B"));
  if (m == 0 || n == 0) return PC_OK;
  dim3 g(unsigned(std::min<int64_t>(ceil_div(m, 32) * ceil_div(n, 32), int64_t(1) << 20)));
  switch (dt) {
    case PC_F64: hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, ctx->stream, (const double*)A, lda, m, n, (double*)B, ldb); break;
    case PC_F32: hipLaunchKernelGGL(k_transpose<float>, g, dim3(256), 0, ctx->stream, (const float*)A, lda, m, n, (float*)B, ldb); break;
    case PC_C64: hipLaunchKernelGGL(k_transpose<cfloat>, g, dim3(256), 0, ctx->stream, (const cfloat*)A, lda, m, n, (cfloat*)B, ldb); break;
    case PC_C128: hipLaunchKernelGGL(k_transpose<cdouble>, g, dim3(256), 0, ctx->stream, (const cdouble*)A, lda, m, n, (cdouble*)B, ldb); break;
  }
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

extern "C" int pc_diff_norms(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, const void* B, int64_t ldb,
                             int64_t m, int64_t n, double* out) {
  if (!ctx || !out) return pc_fail(ctx, PC_ERR_INVALID, "pc_diff_norms: NULL argument");
  PC_CHECK(ctx, check_mat(ctx, A, lda, m, n, "pc_diff_norms A"));
  PC_CHECK(ctx, check_mat(ctx, B, ldb, m, n, "pc_diff_norms B"));
  out[0] = out[1] = 0.0;
  if (m == 0 || n == 0) return PC_OK;
  // complex data: treat each (re, im) pair as two reals; |a-b|^2 sums match
  int64_t mm = m, nn = n, la = lda, lb = ldb;
  pc_dtype base = dt;
  if (dt == PC_C64) { base = PC_F32; nn *= 2; la *= 2; lb *= 2; }
  if (dt == PC_C128) { base = PC_F64; nn *= 2; la *= 2; lb *= 2; }
  const int nb = 1024;
  PC_CHECK(ctx, pc_ws_reserve(ctx, ctx->ws_skip + (2 * nb + 2) * 8));
  double* part = reinterpret_cast<double*>(ctx->ws + ctx->ws_skip);
  double* res = part + 2 * nb;
  if (base == PC_F64)
    hipLaunchKernelGGL(k_diff_norms<double>, dim3(nb), dim3(256), 0, ctx->stream, (const double*)A, la, (const double*)B, lb, mm, nn, part);
  else
    hipLaunchKernelGGL(k_diff_norms<float>, dim3(nb), dim3(256), 0, ctx->stream, (const float*)A, la, (const float*)B, lb, mm, nn, part);
  hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, ctx->stream, part, nb, res);
  PC_HIP(ctx, hipGetLastError());
  PC_HIP(ctx, hipMemcpyAsync(ctx->hflag, res, 16, hipMemcpyDeviceToHost, ctx->stream));
  PC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  memcpy(out, ctx->hflag, 16);
  return PC_OK;
}

extern "C" int pc_dataset_x(pc_ctx* ctx, pc_dtype dt, int64_t nfields, const void* const* src, const int64_t* ndim,
                            const int64_t* ld, int64_t npoints, int64_t nt, void* X, int64_t ldx) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "pc_dataset_x: ctx is NULL");
  if (!src || !ndim || !ld || !X) return pc_fail(ctx, PC_ERR_INVALID, "pc_dataset_x: NULL argument");
  if (nfields <= 0 || nfields > kMaxFields) return pc_fail(ctx, PC_ERR_SHAPE, "pc_dataset_x: 1..%d fields", kMaxFields);
  if (dt != PC_F32 && dt != PC_F64) return pc_fail(ctx, PC_ERR_UNSUPPORTED, "pc_dataset_x: real fields only");
  FieldSet fs;
  int nc = 0;
  for (int64_t f = 0; f < nfields; f++) {
    if (!src[f] || ndim[f] < 1 || nc + ndim[f] > kMaxFields * 8 || ld[f] < nt)
      return pc_fail(ctx, PC_ERR_INVALID, "pc_dataset_x: field %lld: bad pointer, ndim or leading dimension",
                     (long long)f);
    fs.src[f] = src[f];
    fs.ld[f] = ld[f];
    fs.ndim[f] = int(ndim[f]);
    for (int d = 0; d < ndim[f]; d++) {
      fs.field_of[nc] = int(f);
      fs.idim_of[nc] = d;
      nc++;
    }
  }
  PC_CHECK(ctx, check_mat(ctx, X, ldx, npoints * nc, nt, "pc_dataset_x X"));
  if (npoints == 0 || nt == 0) return PC_OK;
  PC_HIP(ctx, hipSetDevice(ctx->device));
  const dim3 g(unsigned(std::min<int64_t>(ceil_div(npoints * nc, 4), 1 << 20)));
  if (dt == PC_F64)
    hipLaunchKernelGGL(k_interleave<double>, g, dim3(256), 0, ctx->stream, fs, nc, npoints, nt, (double*)X, ldx);
  else
    hipLaunchKernelGGL(k_interleave<float>, g, dim3(256), 0, ctx->stream, fs, nc, npoints, nt, (float*)X, ldx);
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}
