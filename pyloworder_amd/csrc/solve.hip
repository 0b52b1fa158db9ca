// Orchestration of the POD core: tall-skinny SVD / QR pipelines, randomized
// SVD, products and grouped solves.  All device work is queued on the
// context stream; the only host waits inside a solve are the convergence
// checks between batches of block-Jacobi sweeps (several block pairs only,
// skipped under stream capture) and the final status read.
//
// Tall-skinny SVD of Y (m x n, optionally row-centred), "two-pass
// Gram-Cholesky" (replaces Demmel TSQR + svd(R) of pyLOM/vmmath/svd.py:49-252):
//   P1  G1 = Y^H Y                       (MFMA streaming pass, fused centring)
//       R1 = chol(G1 + tau I), W1 = R1^-1 (FP64; W1 rounded to fp16 when mixed)
//   P2  G2 = (Y W1)^H (Y W1)             (MFMA streaming pass)
//       P^T G2 P = R2^H R2 (pivoted, rank r), M = R2 P^T W1^-1 (r x n)
//       one-sided Jacobi: M J = H,  sigma = column norms, VT = J^H
//       W = W1 P R2^-1 U_M               (U_M = H sigma^-1)
//   P3  U = Y W                          (MFMA streaming pass)
// plus an orthonormal completion of U for numerically null singular values.
#include <algorithm>
#include <atomic>
#include <thread>
#include <array>
#include <map>

#include "common.hpp"
#include "kernels.hpp"
#include "k_small.hip"
#include "k_dense.hip"
#include "k_mwg.hip"

int pc_check_mat(pc_ctx* ctx, const void* p, int64_t ld, int64_t m, int64_t n, const char* what);

namespace {

constexpr double kShiftF64 = 1e-13;  // tau / max(diag G1), FP64 path (G1 rounding ~ n u: 3e-14 at n = 256)
constexpr double kShiftMx = 1e-6;    // tau / trace(G1), mixed path
constexpr double kTolF64 = 1e-15;    // pivoted Cholesky rank tolerance (relative to max diag)
constexpr double kTolMx = 1e-7;
constexpr double kNullF64 = 1e-11;   // sigma <= tol * sigma_max counts as null: the FP64 path resolves
                                     // a null direction of Y to ~1e-13 sigma_max (rounding through R1)
constexpr double kNullMx = 1e-6;

inline unsigned nblk(int64_t count, int per = 256) { return unsigned(ceil_div(count, per)); }

constexpr int kBjacMaxSweeps = 48;

// dst[0..3] = src ? src[0..3] : 0 (the four status words of a solve)
__global__ void k_status_copy(const int* src, int* dst, int64_t zs = 0, int64_t zd = 0) {
  src = zoff(src, zs);
  dst = zoff(dst, zd);
  if (threadIdx.x < 4) dst[threadIdx.x] = src ? src[threadIdx.x] : 0;
}
constexpr double kJacTolMx = 1e-8;  // Jacobi orthogonality tolerance, mixed path

// One-sided block Jacobi on H (rows x cols, column-major) with J accumulated:
// one launch per round-robin step of a sweep, a check launch per sweep.  After
// convergence every remaining launch exits immediately.
// Block layout of the one-sided Jacobi of a rows x cols problem: real FP64
// with rows <= 256 takes Gram-driven MFMA steps on 16-column blocks, other
// problems column-pair rotations staged in LDS.  b = block width, nb2 = even
// number of blocks (>= 2); nb2 == 2 means one block pair holds every column.
// Gram-driven steps for real problems
template <class T>
void jacobi_blocks(int rows, int cols, bool& gram_steps, int& b, int& nb2) {
  gram_steps = std::is_same<T, double>::value;
  b = gram_steps ? GB : bjac_width(std::max(rows, cols), int(sizeof(T)));
  nb2 = int(ceil_div(cols, b));
  nb2 += nb2 & 1;
  if (nb2 < 2) nb2 = 2;
}

template <class T>
void launch_block_jacobi(pc_ctx* ctx, T* H, int ldh, int rows, int cols, T* J, int ldj, double* sig, int* state,
                         double* dstate, int* status, double min_tol, int nz = 1, int64_t zb = 0) {
  hipStream_t st = ctx->stream;
  bool gram_steps;
  int b, nb2;
  // a context that shares the GPU with other streams (grouped solve) takes
  // the one-launch steps (they hold the pair's rows in LDS: up to GROWS):
  // there the launch count costs more than the one-CU step time the split
  // steps save.  Both forms give the same H and J.
  const bool fused = ctx->concurrent && rows <= GROWS;
  jacobi_blocks<T>(rows, cols, gram_steps, b, nb2);
  // row slices of the apply launch
  const int nsl = std::max(kBjacSlices, int(ceil_div(std::max(rows, cols), 32)));
  hipLaunchKernelGGL((k_bjac_init<T>), dim3(1, 1, nz), dim3(1024), 0, st, H, ldh, rows, cols, J, ldj, state, dstate,
                     kBjacMaxSweeps + 1, zb);
  // one block pair (all columns in one workgroup): each launch iterates to
  // convergence on its own, so a short fixed run of sweeps is queued and the
  // host never waits; several pairs: sweeps in host-checked batches
  const bool one_pair = nb2 == 2;
  int max_inner = one_pair ? 60 : gram_steps ? 1 : 3;
  // relative orthogonality test |h_p^H h_q| <= tol ||h_p|| ||h_q||: the dot
  // product itself carries ~eps sqrt(rows) rounding, so tol grows with rows
  double tol = std::max(2.220446049250313e-16 * std::max(rows, 16), min_tol);
  // Sweeps are queued in batches: after the first batch the host reads the
  // done flag and stops queueing once the iteration has converged (typically
  // after ~10 sweeps at n = 256).  Under stream capture no host read is
  // possible, so every sweep is queued; converged sweeps then return at once.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(st, &cap);
  const bool host_checks = cap == hipStreamCaptureStatusNone && !one_pair;
  constexpr int kFirstBatch = 8, kNextBatch = 2, kOnePairSweeps = 6;
  int next_check = kFirstBatch;
  const int sweeps = one_pair ? kOnePairSweeps : kBjacMaxSweeps;
  for (int s = 0; s < sweeps; s++) {
    if (host_checks && s == next_check) {
      next_check += kNextBatch;
      if (hipMemcpyAsync(ctx->hflag + 8, state, sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
          hipStreamSynchronize(st) == hipSuccess && ctx->hflag[8])
        break;
    }
    for (int step = 0; step < nb2 - 1; step++) {
      if constexpr (std::is_same<T, double>::value) {
        if (gram_steps) {
          if (fused) {
            hipLaunchKernelGGL(k_bjac_gstep<true>, dim3(nb2 / 2 * 8), dim3(1024), 0, st, H, ldh, rows, cols, J, ldj,
                               nb2, step, s, state, dstate, tol, max_inner);
            continue;
          }
          hipLaunchKernelGGL(k_bjac_gstep<false>, dim3(nb2 / 2 * 8), dim3(1024), 0, st, H, ldh, rows, cols, J, ldj,
                             nb2, step, s, state, dstate, tol, max_inner);
          hipLaunchKernelGGL(k_bjac_apply, dim3(nb2 / 2 * nsl), dim3(256), 0, st, H, ldh, rows, cols, J, ldj, nb2,
                             step, s, state, (const double*)dstate, nsl);
          continue;
        }
      }
      if (one_pair && rows <= 64 && 2 * b <= 32) {
        hipLaunchKernelGGL((k_bjac_small<T>), dim3(1, 1, nz), dim3(64), 0, st, H, ldh, rows, cols, J, ldj, b, s, state,
                           dstate, tol, max_inner, zb);
        continue;
      }
      hipLaunchKernelGGL((k_bjac_step<T>), dim3(nb2 / 2), dim3(1024), 0, st, H, ldh, rows, cols, J, ldj, b, nb2, step,
                         s, state, dstate, tol, max_inner);
    }
    if (!(std::is_same<T, double>::value && gram_steps && !fused))  // else the sweep closes in k_bjac_apply
      hipLaunchKernelGGL(k_bjac_check, dim3(1, 1, nz), dim3(64), 0, st, state, s, zb);
  }
  hipLaunchKernelGGL((k_bjac_norms<T>), dim3(1, 1, nz), dim3(1024), 0, st, H, ldh, rows, cols, sig, state, status, zb);
}


// Small dense factorisations: blocked MFMA kernels for real n <= 256 (one
// workgroup), one launch per step over many workgroups for larger single
// problems (k_mwg.hip), the generic single-workgroup kernels otherwise
// (complex n <= 128, batched groups).
template <class T>
inline bool use_fast(int n, int nz) { return std::is_same<T, double>::value && n <= dense::NMAX && nz == 1; }
inline bool use_mw(int n, int nz) { return n > 128 && nz == 1; }

// work: n x n of T, dsc: 1 double, flag: 1 int (multi-workgroup path only)
template <class T>
void run_chol_shift(hipStream_t st, const T* G, int n, double shift, T* R, int* status, int nz = 1, int64_t zb = 0,
                    T* work = nullptr, double* dsc = nullptr, int* flag = nullptr) {
  if constexpr (std::is_same<T, double>::value)
    if (use_fast<T>(n, nz)) {
      hipLaunchKernelGGL(k_chol_fast, dim3(1), dim3(NT), 0, st, G, n, n, shift, R, n, status);
      return;
    }
  if (use_mw(n, nz) && work && dsc && flag) {
    hipLaunchKernelGGL((k_chol_mw_tau<T>), dim3(1), dim3(1024), 0, st, G, n, n, shift, dsc, flag);
    hipLaunchKernelGGL((k_chol_mw_copy<T>), dim3(mw_grid(int64_t(n) * n, MW)), dim3(MW), 0, st, G, n, n,
                       (const double*)dsc, work);
    for (int k = 0; k < n - 1; k++)
      hipLaunchKernelGGL((k_chol_mw_step<T>), dim3(mw_grid(n - k - 1, MW / 64)), dim3(MW), 0, st, work, n, k,
                         (const double*)dsc, flag);
    hipLaunchKernelGGL((k_chol_mw_finish<T>), dim3(mw_grid(int64_t(n) * n, MW)), dim3(MW), 0, st, (const T*)work, n,
                       R, n, (const int*)flag);
    // the last pivot, and any attempt that failed, is checked by the generic kernel
    hipLaunchKernelGGL((k_chol_check_last<T>), dim3(1), dim3(64), 0, st, (const T*)work, n, (const double*)dsc, flag);
    hipLaunchKernelGGL((k_chol_shift<T>), dim3(1), dim3(NT), 0, st, G, n, n, shift, R, n, status, int64_t(0),
                       (const int*)flag);
    return;
  }
  hipLaunchKernelGGL((k_chol_shift<T>), dim3(1, 1, nz), dim3(NT), 0, st, G, n, n, shift, R, n, status, zb,
                     (const int*)nullptr);
}
template <class T>
void run_trinv(hipStream_t st, const T* R, int n, T* W, int nz = 1, int64_t zb = 0) {
  if constexpr (std::is_same<T, double>::value)
    if (use_fast<T>(n, nz)) {
      hipLaunchKernelGGL(k_trinv_cols, dim3(n), dim3(64), 0, st, R, n, n, W, n);
      return;
    }
  if (use_mw(n, nz)) {  // columns spread over the workgroups
    hipLaunchKernelGGL((k_trinv_upper<T>), dim3(unsigned(ceil_div(n, NT / 64))), dim3(NT), 0, st, R, n, n, W, n,
                       int64_t(0));
    return;
  }
  hipLaunchKernelGGL((k_trinv_upper<T>), dim3(1, 1, nz), dim3(NT), 0, st, R, n, n, W, n, zb);
}
// state: 1 int (multi-workgroup path)
template <class T>
void run_pchol(hipStream_t st, const T* G, int n, double tol, T* A, T* R, int* perm, int* rank, int* status,
               int nz = 1, int64_t zb = 0, int* state = nullptr) {
  if constexpr (std::is_same<T, double>::value)
    if (use_fast<T>(n, nz)) {
      hipLaunchKernelGGL(k_pchol_fast, dim3(1), dim3(NT), 0, st, G, n, n, tol, A, R, n, perm, rank, status);
      return;
    }
  if (use_mw(n, nz) && state) {
    hipLaunchKernelGGL((k_pchol_mw_copy<T>), dim3(mw_grid(int64_t(n) * n, MW)), dim3(MW), 0, st, G, n, n, A, R, n,
                       perm, rank, state);
    for (int k = 0; k < n; k++) {
      hipLaunchKernelGGL((k_pchol_mw_pivot<T>), dim3(1), dim3(1024), 0, st, G, n, A, n, k, tol, R, n, perm, rank,
                         state, status);
      if (k + 1 < n)
        hipLaunchKernelGGL((k_pchol_mw_update<T>), dim3(mw_grid(n - k - 1, MW / 64)), dim3(MW), 0, st, A, n, k,
                           (const T*)R, n, (const int*)state);
    }
    return;
  }
  hipLaunchKernelGGL((k_pchol<T>), dim3(1, 1, nz), dim3(NT), 0, st, G, n, n, tol, A, R, n, perm, rank, status, zb);
}
// square Householder QR (n x n), Q and R explicit; work: 2 n^2 + n of T.
// fused: one launch (k_house_qr_fast) instead of the panel / strip sequence,
// for contexts that share the GPU with other streams; same Q and R.
template <class T>
void run_house_qr_sq(hipStream_t st, const T* A, int n, T* Q, T* R, T* work, bool fused = false) {
  if constexpr (std::is_same<T, double>::value)
    if (n <= dense::NMAX && fused) {
      hipLaunchKernelGGL(k_house_qr_fast, dim3(1), dim3(NT), 0, st, A, n, n, Q, n, R, n, work);
      return;
    }
  if constexpr (std::is_same<T, double>::value)
    if (n <= dense::NMAX) {
      // panels on one workgroup, the block-reflector updates on one
      // workgroup per 32-column strip
      const int npan = (n + dense::PW - 1) / dense::PW;
      for (int pp = 0; pp < npan; pp++) {
        hipLaunchKernelGGL(k_hqr_panel, dim3(1), dim3(NT), 0, st, A, n, n, pp, work);
        const int nt = n - std::min(n, (pp + 1) * dense::PW);
        if (nt > 0) hipLaunchKernelGGL(k_hqr_strip, dim3((nt + 31) / 32), dim3(256), 0, st, n, pp, 0, work, Q, n);
      }
      hipLaunchKernelGGL(k_hqr_rq, dim3(nblk(int64_t(n) * n)), dim3(256), 0, st, n, (const double*)work, R, n, Q, n);
      for (int pp = npan - 1; pp >= 0; pp--)
        hipLaunchKernelGGL(k_hqr_strip, dim3((n - pp * dense::PW + 31) / 32), dim3(256), 0, st, n, pp, 1, work, Q,
                           n);
      return;
    }
  if (use_mw(n, 1)) {
    T* Wc = work;
    T* Qc = work + int64_t(n) * n;
    T* tau = Qc + int64_t(n) * n;
    hipLaunchKernelGGL((k_hqr_init<T>), dim3(mw_grid(2 * int64_t(n) * n, MW)), dim3(MW), 0, st, A, n, n, n, n, Wc, Qc);
    for (int k = 0; k < n; k++) {
      hipLaunchKernelGGL((k_hqr_reflect<T>), dim3(1), dim3(MW), 0, st, Wc, n, k, tau);
      if (k + 1 < n)
        hipLaunchKernelGGL((k_hqr_apply<T>), dim3(mw_grid(n - k - 1, MW / 64)), dim3(MW), 0, st, Wc, n, n, k,
                           (const T*)tau);
    }
    for (int k = n - 1; k >= 0; k--)
      hipLaunchKernelGGL((k_hqr_qapply<T>), dim3(mw_grid(n - k, MW / 64)), dim3(MW), 0, st, (const T*)Wc, n, n, k,
                         (const T*)tau, Qc);
    hipLaunchKernelGGL((k_hqr_out<T>), dim3(mw_grid(2 * int64_t(n) * n, MW)), dim3(MW), 0, st, (const T*)Wc,
                       (const T*)Qc, n, n, n, Q, n, R, n);
    return;
  }
  hipLaunchKernelGGL((k_house_qr<T>), dim3(1), dim3(NT), 0, st, A, n, n, n, Q, n, R, n, work);
}
// workspace of the small dense stage (T elements): house QR and null completion
inline int64_t dense_work(int n) { return 3 * int64_t(n) * n + 64 + n + 1024 * int64_t((n + 31) / 32); }

// output element type for (compute type T, real storage type ST)
template <class T, class ST> struct VtType;
template <> struct VtType<double, double> { typedef double type; };
template <> struct VtType<double, float> { typedef float type; };
template <> struct VtType<cdouble, double> { typedef cdouble type; };
template <> struct VtType<cdouble, float> { typedef cfloat type; };

template <class T, class TS> PC_DEV T load_elem(const TS* X, int64_t off, int j);
template <> PC_DEV double load_elem<double, double>(const double* X, int64_t off, int j) { return X[off + j]; }
template <> PC_DEV double load_elem<double, float>(const float* X, int64_t off, int j) { return double(X[off + j]); }
template <> PC_DEV cdouble load_elem<cdouble, cdouble>(const cdouble* X, int64_t off, int j) { return X[off + j]; }
template <> PC_DEV cdouble load_elem<cdouble, cfloat>(const cfloat* X, int64_t off, int j) {
  return cdouble(X[off + j].x, X[off + j].y);
}
PC_DEV void add_elem(double* p, double v) { *p += v; }
PC_DEV void add_elem(float* p, double v) { *p = float(double(*p) + v); }
PC_DEV void add_elem(cdouble* p, cdouble v) { p->x += v.x; p->y += v.y; }
PC_DEV void add_elem(cfloat* p, cdouble v) { p->x = float(p->x + v.x); p->y = float(p->y + v.y); }

// nnull[0] = n - rank: the columns of Q the null-space completion builds
__global__ void k_null_count(const int* rank, int n, int* nnull) {
  if (threadIdx.x == 0) nnull[0] = n - rank[0];
}

// QR output: R (n x n upper, rows >= rank zero); marks the n - rank null columns of Q
template <class T, class TO>
__global__ __launch_bounds__(256) void k_store_r(const T* Rm, int ldi, int n, TO* R, int64_t ldr, const int* rank,
                                                 int* nnull) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int rk = rank[0];
  if (idx == 0) nnull[0] = n - rk;
  if (idx >= int64_t(n) * n) return;
  int i = int(idx / n), j = int(idx % n);
  T v = (i <= j && i < rk) ? Rm[int64_t(i) * ldi + j] : T(0.0);
  store_vt(R + int64_t(i) * ldr + j, v);
}


// QR sign convention: R := D^H R and Q := Q D with D = diag(R_ii / |R_ii|)
// (1 where R_ii = 0), so that diag(R) is real and non-negative; Q R is unchanged
template <class T>
__global__ __launch_bounds__(256) void k_qr_phase(const T* Rm, int ldr, int n, T* ph) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const T d = Rm[int64_t(i) * ldr + i];
  const double a = sqrt(abs2_(d));
  ph[i] = a > 0.0 ? phase_(d, a) : T(1.0);
}
template <class T>
__global__ __launch_bounds__(256) void k_qr_apply_phase(T* Rm, int ldr, T* Qm, int ldq, int n, const T* ph) {
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * n) return;
  const int i = int(idx / n), j = int(idx % n);
  Rm[int64_t(i) * ldr + j] = conj_(ph[i]) * Rm[int64_t(i) * ldr + j];
  Qm[int64_t(i) * ldq + j] = Qm[int64_t(i) * ldq + j] * ph[j];
}

// ---------------------------------------------------------------------------
// LAPACK's sign convention for the QR (numpy.linalg.qr = ?geqrf + ?orgqr, the
// leaf QR of the reference's tsqr, pyLOM/vmmath/svd.py:28-36, 62).  ?larfg
// takes beta = -sign(Re alpha) ||x|| at step k, so R_H = D R+ and Q_H = Q+ D,
// where Q+ R+ is the QR with a positive real diagonal and D = diag(+-1).  The
// Householder reflectors of A are those of Q+ (column k of A is R+_kk > 0 times
// column k of Q+ plus earlier columns), and the alpha of step k of the
// Householder QR of an orthonormal Q+ is the k-th pivot of the LU of its top
// n x n block Q1 with the diagonal shifted by s_k = -sign(Re pivot_k)
// (Ballard, Demmel, Grigori, Jacquelin, Nguyen, Solomonik, "Reconstructing
// Householder vectors from tall-skinny QR", IPDPS 2014).  So D_k = s_k, read
// from the first n rows of Q alone.  A column already reduced (|alpha| = 1,
// ||x|| = 0) takes no reflection in ?larfg (tau = 0, beta = alpha): then
// D_k = sign(alpha).
// ---------------------------------------------------------------------------
// Q1[i][:] = row i of the final Q (i < gridDim.x): y_i W (K = n inputs, cols
// outputs) plus the null-space fix of row i (nnull may be NULL: no fix)
template <class T, class TS>
__global__ __launch_bounds__(256) void k_q1_rows(const TS* X, int64_t ldx, int n, int cols, const T* W, int ldw,
                                                 const int* nnull, const int* jsel, const T* cnull, T* Q1) {
  const int i = blockIdx.x;
  HIP_DYNAMIC_SHARED(char, q1_smem)  // n values of T
  T* yrow = reinterpret_cast<T*>(q1_smem);
  for (int j = threadIdx.x; j < n; j += 256) yrow[j] = load_elem<T, TS>(X, int64_t(i) * ldx, j);
  __syncthreads();
  const int d = nnull ? nnull[0] : 0, g = cols - d;
  for (int k = threadIdx.x; k < cols; k += 256) {
    T u = T(0.0);
    for (int j = 0; j < n; j++) u += yrow[j] * W[int64_t(j) * ldw + k];
    if (k >= g)
      for (int a = 0; a < d; a++)
        if (jsel[a] == i) u += cnull[int64_t(a) * d + (k - g)];
    Q1[int64_t(i) * cols + k] = u;
  }
}

// D (n signs) from Q1 (n x n, overwritten by its shifted LU); one workgroup.
// The shift makes every pivot at least 1 in modulus, so no pivoting is needed.
template <class T>
__global__ __launch_bounds__(1024) void k_lapack_signs(T* Q1, int n, T* D) {
  constexpr double kReduced = 64 * 2.220446049250313e-16;
  for (int k = 0; k < n; k++) {
    const T a = Q1[int64_t(k) * n + k];
    const double re = real_(a);
    const T s = T(re >= 0.0 ? -1.0 : 1.0);
    if (threadIdx.x == 0) {
      const bool reduced = fabs(sqrt(abs2_(a)) - 1.0) <= kReduced && abs2_(a - T(re)) <= kReduced * kReduced;
      D[k] = reduced ? T(re >= 0.0 ? 1.0 : -1.0) : s;
    }
    const T ip = inv_(a - s);
    const int w = n - k - 1;
    for (int64_t idx = threadIdx.x; idx < int64_t(w) * w; idx += 1024) {
      const int i = k + 1 + int(idx / w), j = k + 1 + int(idx % w);
      Q1[int64_t(i) * n + j] -= Q1[int64_t(i) * n + k] * ip * Q1[int64_t(k) * n + j];
    }
    __syncthreads();
  }
}

// R := D R (rows), W := W D (columns), null fix entries of column g + b by D_{g+b}
template <class T>
__global__ __launch_bounds__(256) void k_apply_signs(const T* D, int n, T* R, int ldr, T* W, int ldw,
                                                     const int* nnull, T* cnull) {
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * n) return;
  const int i = int(idx / n), j = int(idx % n);
  R[int64_t(i) * ldr + j] = R[int64_t(i) * ldr + j] * real_(D[i]);
  W[int64_t(i) * ldw + j] = W[int64_t(i) * ldw + j] * real_(D[j]);
  const int d = nnull[0], g = n - d;
  if (idx < int64_t(d) * d) cnull[idx] = cnull[idx] * real_(D[g + int(idx % d)]);
}

template <class T> struct SmallBufs {
  T *G1, *R1, *W1, *R1h, *G2, *Awk, *R2, *Hc, *J, *UM, *R2i, *tmp, *Wf;
  double* sig;
  int *perm, *order, *rank, *nnull, *status;
  int *perm3, *rank3;  // pivoted Cholesky of M^H M (Jacobi preconditioner)
  int* jstate;       // block Jacobi: done flag + per-sweep rotation flags
  double* jdstate;   // block Jacobi: negligible squared column norm
  double* dsc;       // multi-workgroup Cholesky: the shift
  double* Wpad;  // NB x NB real (embedded) operand for the FP64 streaming kernels
  // null completion
  double* lev;
  T* urows;
  int* jsel;
  T* cnull;
  // wide problems: Z = Yc W1 (m x nr FP64, ld ldz), in U when U is FP64
  double* Zw;
  int64_t ldz;
  // mixed operands
  _Float16 *Wh, *Wl;
  float* wscale;
  double* Wr;
};

struct Plan {
  pc_dtype dt;
  bool cplx;
  bool mixed;
  int n;       // logical columns
  int nr;      // real columns of the streaming view (2n for complex)
  int NB;      // padded real width of the streaming kernels
  int64_t m;
  int tr;      // rows per tile of the P1 kernel
  int chunks;
  int64_t tpc;
  int tr3, chunks3;  // P2/P3 kernels (mixed uses other tiles)
  int64_t tpc3;
  int chunks2;       // mixed P2 (row tiles of tr3, at most 512 chunks)
  int64_t tpc2;
  int nparts;  // partial slabs per P1 launch
  bool wide;   // 256 < real columns <= 2048: stripe-pair FP64 kernels, Z materialised
};

constexpr int kWideMaxReal = 2048;            // real columns of the widest problem
constexpr int64_t kWideSlabBytes = int64_t(256) << 20;  // partial Gram slabs of a wide pass

int make_plan(pc_ctx* ctx, pc_dtype dt, int64_t m, int64_t n, int precision, Plan& p) {
  p.dt = dt;
  p.cplx = dtype_complex(dt);
  p.mixed = precision == PC_PREC_MIXED;
  if (p.mixed && !dtype_single(dt))
    return pc_fail(ctx, PC_ERR_UNSUPPORTED, "mixed precision needs float32 or complex64 storage");
  if (precision != PC_PREC_MIXED && precision != PC_PREC_FP64)
    return pc_fail(ctx, PC_ERR_INVALID, "unknown precision %d", precision);
  p.n = int(n);
  p.nr = p.cplx ? 2 * int(n) : int(n);
  p.m = m;
  if (p.mixed) {
    p.NB = mx2_nb(p.nr);
    p.tr = mx2_gram_rows(p.nr);
    p.tr3 = mx2_tile_rows(p.nr);
  } else {
    p.NB = f64_nb(p.nr);
    p.tr = f64_tile_rows();
    p.tr3 = f64_tile_rows();
  }
  p.wide = false;
  if (p.NB < 0 && p.mixed)
    return pc_fail(ctx, PC_ERR_UNSUPPORTED, "mixed precision takes n <= %s (n = %lld); use FP64",
                   p.cplx ? "128 complex" : "256", (long long)n);
  if (p.NB < 0) {
    if (p.nr > kWideMaxReal)
      return pc_fail(ctx, PC_ERR_SHAPE, "n = %lld is larger than this build supports (%s)", (long long)n,
                     p.cplx ? "complex: 1024" : "2048");
    // wide: columns in stripes of 128, Gram slabs capped at kWideSlabBytes (the
    // chunk count still depends on (m, n) only)
    p.wide = true;
    p.NB = int(round_up(p.nr, wide_stripe()));
    p.tr = p.tr3 = wide_tile_rows();
    const int64_t cap = std::max<int64_t>(1, kWideSlabBytes / (int64_t(p.NB) * p.NB * 8));
    p.chunks = int(std::min<int64_t>(stream_chunks(m, p.tr), cap));
    p.tpc = ceil_div(ceil_div(m, p.tr), p.chunks);
    p.chunks3 = stream_chunks(m, p.tr3);
    p.tpc3 = ceil_div(ceil_div(m, p.tr3), p.chunks3);
    p.chunks2 = p.chunks3;
    p.tpc2 = p.tpc3;
    p.nparts = p.chunks;
    return PC_OK;
  }
  p.chunks = stream_chunks(m, p.tr);
  // mixed P1 and P2: at most 512 chunks (two workgroups per CU in turn for
  // long inputs), half the partial slabs to write and sum
  auto cap512 = [](int64_t tiles, int c) {
    const int c512 = int(ceil_div(tiles, ceil_div(tiles, 512)));
    return c512 < c ? c512 : c;
  };
  if (p.mixed) p.chunks = cap512(ceil_div(m, p.tr), p.chunks);
  p.tpc = ceil_div(ceil_div(m, p.tr), p.chunks);
  p.chunks3 = stream_chunks(m, p.tr3);
  p.tpc3 = ceil_div(ceil_div(m, p.tr3), p.chunks3);
  p.chunks2 = p.mixed ? cap512(ceil_div(m, p.tr3), p.chunks3) : p.chunks3;
  p.tpc2 = ceil_div(ceil_div(m, p.tr3), p.chunks2);
  p.nparts = p.chunks;
  return PC_OK;
}

template <class T>
int64_t carve(WsCursor& c, const Plan& p, SmallBufs<T>& b, double** mean, double** var, double** part,
              int64_t part_slabs) {
  int n = p.n;
  int64_t nn = int64_t(n) * n;
  *mean = c.take<double>(p.m);
  *var = c.take<double>(p.m);
  *part = c.take<double>(part_slabs * int64_t(p.NB) * p.NB);
  b.G1 = c.take<T>(nn); b.R1 = c.take<T>(nn); b.W1 = c.take<T>(nn); b.R1h = c.take<T>(nn);
  b.G2 = c.take<T>(nn); b.Awk = c.take<T>(nn); b.R2 = c.take<T>(nn); b.Hc = c.take<T>(nn);
  b.J = c.take<T>(nn); b.UM = c.take<T>(nn); b.R2i = c.take<T>(nn); b.tmp = c.take<T>(nn); b.Wf = c.take<T>(nn);
  b.sig = c.take<double>(n);
  b.perm = c.take<int>(n); b.order = c.take<int>(n); b.rank = c.take<int>(4); b.nnull = c.take<int>(4);
  b.status = c.take<int>(4);
  b.perm3 = c.take<int>(n); b.rank3 = c.take<int>(4);
  b.jstate = c.take<int>(kBjacMaxSweeps + 8);
  b.jdstate = c.take<double>(bjac_dstate_len(n));  // [0] negligible norm, then the Gram steps' pair rotations
  b.dsc = c.take<double>(4);
  b.Wpad = c.take<double>(int64_t(p.NB) * p.NB);
  int ncand = int(std::min<int64_t>(p.m, 4 * int64_t(n) + 16));
  b.lev = c.take<double>(ncand);
  b.urows = c.take<T>(int64_t(ncand) * n);
  b.jsel = c.take<int>(n);
  b.cnull = c.take<T>(nn);
  b.Wh = c.take<_Float16>(int64_t(p.NB) * p.NB);
  b.Wl = c.take<_Float16>(int64_t(p.NB) * p.NB);
  b.wscale = c.take<float>(p.NB);
  b.Wr = c.take<double>(int64_t(p.NB) * p.NB);
  b.Zw = nullptr;
  b.ldz = p.nr;
  if (p.wide && dtype_single(p.dt)) b.Zw = c.take<double>(p.m * int64_t(p.nr));
  return c.off;
}

// ---------------------------------------------------------------------------
// null completion kernels
// ---------------------------------------------------------------------------
// leverage of candidate rows: u_c = y_c W[:, :g] (g = n - nnull), lev_c = ||u_c||^2
template <class T, class TS>
__global__ __launch_bounds__(256) void k_null_lev(const TS* X, int64_t ldx, int64_t m, int n, int center,
                                                  const double* mean, const double* var, const T* W, int ldw,
                                                  const int* nnull, int ncand, double* lev, T* urows, int64_t zb = 0, int64_t zx = 0) {
  X = zoff(X, zx);
  mean = zoff(mean, zb);
  var = zoff(var, zb);
  W = zoff(W, zb);
  nnull = zoff(nnull, zb);
  lev = zoff(lev, zb);
  urows = zoff(urows, zb);
  int d = nnull[0];
  if (d == 0) return;
  int c = blockIdx.x;
  int64_t row = ncand >= m ? c : (int64_t(c) * m) / ncand;
  HIP_DYNAMIC_SHARED(char, lev_smem)  // n values of T
  T* yrow = reinterpret_cast<T*>(lev_smem);
  for (int j = threadIdx.x; j < n; j += 256) {
    T v = load_elem<T, TS>(X, row * ldx, j);
    if (center == 3) {
      v = v - load_elem<T, TS>(X, row * ldx, 0);  // mixed path: rows shifted by their first element
    } else if (center) {
      v = v - T(mean[row]);
      if (center == 2) v = v * (1.0 / var[row]);
    }
    yrow[j] = v;
  }
  __syncthreads();
  int g = n - d;
  double s = 0;
  for (int k = threadIdx.x; k < n; k += 256) {
    T u = T(0.0);
    if (k < g)
      for (int j = 0; j < n; j++) u += yrow[j] * W[int64_t(j) * ldw + k];
    urows[int64_t(c) * n + k] = u;
    s += abs2_(u);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) lev[c] = sh[0];
}


template <class T>
void launch_null_lev(hipStream_t st, const Plan& p, const void* A, int64_t lda, int center, const double* mean,
                     const double* var, SmallBufs<T>& b, int ncand, int nz = 1, int64_t zb = 0, int64_t zx = 0);
template <>
void launch_null_lev<double>(hipStream_t st, const Plan& p, const void* A, int64_t lda, int center, const double* mean,
                             const double* var, SmallBufs<double>& b, int ncand, int nz,
                             int64_t zb, int64_t zx) {
  if (p.dt == PC_F64)
    hipLaunchKernelGGL((k_null_lev<double, double>), dim3(ncand, 1, nz), dim3(256), size_t(p.n) * sizeof(double), st, (const double*)A, lda, p.m, p.n,
                       center, mean, var, b.Wf, p.n, b.nnull, ncand, b.lev, b.urows, zb, zx);
  else
    hipLaunchKernelGGL((k_null_lev<double, float>), dim3(ncand, 1, nz), dim3(256), size_t(p.n) * sizeof(double), st, (const float*)A, lda, p.m, p.n,
                       center, mean, var, b.Wf, p.n, b.nnull, ncand, b.lev, b.urows, zb, zx);
}
template <>
void launch_null_lev<cdouble>(hipStream_t st, const Plan& p, const void* A, int64_t lda, int center, const double* mean,
                              const double* var, SmallBufs<cdouble>& b, int ncand, int nz,
                              int64_t zb, int64_t zx) {
  if (p.dt == PC_C128)
    hipLaunchKernelGGL((k_null_lev<cdouble, cdouble>), dim3(ncand, 1, nz), dim3(256), size_t(p.n) * sizeof(cdouble), st, (const cdouble*)A, lda, p.m,
                       p.n, center, mean, var, b.Wf, p.n, b.nnull, ncand, b.lev, b.urows, zb, zx);
  else
    hipLaunchKernelGGL((k_null_lev<cdouble, cfloat>), dim3(ncand, 1, nz), dim3(256), size_t(p.n) * sizeof(cdouble), st, (const cfloat*)A, lda, p.m, p.n,
                       center, mean, var, b.Wf, p.n, b.nnull, ncand, b.lev, b.urows, zb, zx);
}

template <class T>
void launch_q1_rows(hipStream_t st, const Plan& p, const void* A, int64_t lda, SmallBufs<T>& b);
template <>
void launch_q1_rows<double>(hipStream_t st, const Plan& p, const void* A, int64_t lda, SmallBufs<double>& b) {
  if (p.dt == PC_F64)
    hipLaunchKernelGGL((k_q1_rows<double, double>), dim3(p.n), dim3(256), size_t(p.n) * sizeof(double), st, (const double*)A, lda, p.n, p.n, b.Wf, p.n,
                       (const int*)b.nnull, (const int*)b.jsel, (const double*)b.cnull, b.tmp);
  else
    hipLaunchKernelGGL((k_q1_rows<double, float>), dim3(p.n), dim3(256), size_t(p.n) * sizeof(double), st, (const float*)A, lda, p.n, p.n, b.Wf, p.n,
                       (const int*)b.nnull, (const int*)b.jsel, (const double*)b.cnull, b.tmp);
}
template <>
void launch_q1_rows<cdouble>(hipStream_t st, const Plan& p, const void* A, int64_t lda, SmallBufs<cdouble>& b) {
  if (p.dt == PC_C128)
    hipLaunchKernelGGL((k_q1_rows<cdouble, cdouble>), dim3(p.n), dim3(256), size_t(p.n) * sizeof(cdouble), st, (const cdouble*)A, lda, p.n, p.n, b.Wf,
                       p.n, (const int*)b.nnull, (const int*)b.jsel, (const cdouble*)b.cnull, b.tmp);
  else
    hipLaunchKernelGGL((k_q1_rows<cdouble, cfloat>), dim3(p.n), dim3(256), size_t(p.n) * sizeof(cdouble), st, (const cfloat*)A, lda, p.n, p.n, b.Wf,
                       p.n, (const int*)b.nnull, (const int*)b.jsel, (const cdouble*)b.cnull, b.tmp);
}

// select the d lowest-leverage candidates, build the completion:
//   Md = I - U_J U_J^H = R^H R,  C = R^-1,  W[:, g:] = -W[:, :g] U_J^H C,
//   jsel[a] = row index j_a, cnull[a][b] = C[a][b]  (U[j_a][g+b] += C[a][b])
inline size_t null_build_lds(int n, int ncand) { return size_t(ncand) * sizeof(double) + size_t(n) * sizeof(int); }
template <class T>
__global__ __launch_bounds__(1024) void k_null_build(int64_t m, int n, const int* nnull, int ncand, const double* lev,
                                                     const T* urows, T* W, int ldw, int* jsel, T* cnull, T* work,
                                                     int* status, int64_t zb = 0) {
  nnull = zoff(nnull, zb);
  lev = zoff(lev, zb);
  urows = zoff(urows, zb);
  W = zoff(W, zb);
  jsel = zoff(jsel, zb);
  cnull = zoff(cnull, zb);
  work = zoff(work, zb);
  status = zoff(status, zb);
  int d = nnull[0];
  if (d == 0) return;
  int g = n - d;
  // dynamic LDS: the candidates' sort keys (ncand doubles), then the picks
  // (d <= n indices); null_build_lds() gives its size
  HIP_DYNAMIC_SHARED(double, keys)
  int* pick = reinterpret_cast<int*>(keys + ncand);
  for (int c = threadIdx.x; c < ncand; c += 1024) keys[c] = key_asc(lev[c]);
  __syncthreads();
  // rank candidates by leverage (ties by index), keep the first d
  for (int c = threadIdx.x; c < ncand; c += 1024) {
    double v = keys[c];
    int r = 0;
    for (int e = 0; e < ncand; e++) {
      double u = keys[e];
      r += (u < v) || (u == v && e < c);
    }
    if (r < d) pick[r] = c;
  }
  __syncthreads();
  for (int a = threadIdx.x; a < d; a += 1024) {
    int c = pick[a];
    c = c < 0 ? 0 : (c >= ncand ? ncand - 1 : c);
    jsel[a] = int(ncand >= m ? c : (int64_t(c) * m) / ncand);
  }
  // Md (d x d) in work, row-major
  T* Md = work;
  T* Rd = work + int64_t(d) * d;
  for (int64_t idx = threadIdx.x; idx < int64_t(d) * d; idx += 1024) {
    int a = int(idx / d), b = int(idx % d);
    const T* ua = urows + int64_t(pick[a]) * n;
    const T* ub = urows + int64_t(pick[b]) * n;
    T s = T(0.0);
    for (int k = 0; k < g; k++) s += ua[k] * conj_(ub[k]);
    Md[idx] = T(a == b ? 1.0 : 0.0) - s;
  }
  __syncthreads();
  // Cholesky Md = Rd^H Rd (upper), in place into Rd
  for (int64_t idx = threadIdx.x; idx < int64_t(d) * d; idx += 1024) {
    int a = int(idx / d), b = int(idx % d);
    Rd[idx] = a <= b ? Md[idx] : T(0.0);
  }
  __syncthreads();
  __shared__ double piv;
  for (int k = 0; k < d; k++) {
    if (threadIdx.x == 0) {
      double v = real_(Rd[int64_t(k) * d + k]);
      if (!(v > 1e-12)) { status[0] |= 4; v = 1e-12; }
      piv = sqrt(v);
    }
    __syncthreads();
    for (int j = k + 1 + threadIdx.x; j < d; j += 1024) Rd[int64_t(k) * d + j] = Rd[int64_t(k) * d + j] * (1.0 / piv);
    if (threadIdx.x == 0) Rd[int64_t(k) * d + k] = T(piv);
    __syncthreads();
    int w = d - k - 1;
    for (int64_t idx = threadIdx.x; idx < int64_t(w) * w; idx += 1024) {
      int i = k + 1 + int(idx / w), j = k + 1 + int(idx % w);
      if (i <= j) Rd[int64_t(i) * d + j] -= conj_(Rd[int64_t(k) * d + i]) * Rd[int64_t(k) * d + j];
    }
    __syncthreads();
  }
  // C = Rd^-1 (upper) by column back substitution into Md
  for (int b = threadIdx.x; b < d; b += 1024) {
    for (int a = d - 1; a >= 0; a--) {
      T s = T(a == b ? 1.0 : 0.0);
      for (int k = a + 1; k <= b; k++) s -= Rd[int64_t(a) * d + k] * Md[int64_t(k) * d + b];
      T raa = Rd[int64_t(a) * d + a];
      Md[int64_t(a) * d + b] = a > b ? T(0.0) : s * conj_(raa) * (1.0 / abs2_(raa));
    }
  }
  __syncthreads();
  for (int64_t idx = threadIdx.x; idx < int64_t(d) * d; idx += 1024) cnull[idx] = Md[idx];
  // T2 = U_J^H C  (g x d), then W[:, g+b] = -W[:, :g] T2[:, b]
  T* T2 = Rd;  // reuse (g x d fits: d*d may be < g*d, so use tail of work)
  T2 = work + 2 * int64_t(d) * d;
  for (int64_t idx = threadIdx.x; idx < int64_t(g) * d; idx += 1024) {
    int k = int(idx / d), b = int(idx % d);
    T s = T(0.0);
    for (int a = 0; a < d; a++) s += conj_(urows[int64_t(pick[a]) * n + k]) * Md[int64_t(a) * d + b];
    T2[idx] = s;
  }
  __syncthreads();
  for (int64_t idx = threadIdx.x; idx < int64_t(n) * d; idx += 1024) {
    int j = int(idx / d), b = int(idx % d);
    T s = T(0.0);
    for (int k = 0; k < g; k++) s += W[int64_t(j) * ldw + k] * T2[int64_t(k) * d + b];
    W[int64_t(j) * ldw + g + b] = -s;
  }
}

template <class T, class TO>
__global__ __launch_bounds__(256) void k_null_fix(const int* nnull, int64_t m, int n, const int* jsel, const T* cnull, TO* U,
                                                  int64_t ldu, int64_t zb = 0, int64_t zc = 0) {
  nnull = zoff(nnull, zb);
  jsel = zoff(jsel, zb);
  cnull = zoff(cnull, zb);
  U = zoff(U, zc);
  int d = nnull[0];
  int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= d * d) return;
  int a = idx / d, b = idx % d;
  int g = n - d;
  int64_t row = jsel[a];
  if (row < 0 || row >= m || g + b >= n) return;
  TO* p = U + row * ldu + g + b;
  add_elem(p, cnull[idx]);
}

// set unit diagonal on the rows >= rank (keeps R2 invertible)
template <class T>
__global__ void k_pad_identity(T* R, int ldr, int n, const int* rank, int64_t zb = 0) {
  R = zoff(R, zb);
  rank = zoff(rank, zb);
  int r = rank[0];
  for (int i = r + threadIdx.x; i < n; i += blockDim.x) R[int64_t(i) * ldr + i] = T(1.0);
}

// real n x nc FP64 -> padded NB x NCB operand
__global__ __launch_bounds__(256) void k_pad_real(const double* W, int ldw, int n, int nc, double* P, int NB, int NCB) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(NB) * NCB) return;
  int i = int(idx / NCB), j = int(idx % NCB);
  P[idx] = (i < n && j < nc) ? W[int64_t(i) * ldw + j] : 0.0;
}

__global__ __launch_bounds__(256) void k_embed_pad(const cdouble* W, int ldw, int n, int nc, double* P, int NB, int NCB, int64_t zb = 0) {
  W = zoff(W, zb);
  P = zoff(P, zb);
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(NB) * NCB) return;
  int i = int(idx / NCB), j = int(idx % NCB);
  double v = 0.0;
  if (i < 2 * n && j < 2 * nc) {
    cdouble w = W[int64_t(i >> 1) * ldw + (j >> 1)];
    int pi = i & 1, pj = j & 1;
    v = (pi == 0) ? (pj == 0 ? w.x : w.y) : (pj == 0 ? -w.y : w.x);
  }
  P[idx] = v;
}

// M <- M C, C = I - 1 1^T / n  (M n x n, column-major: M[i][j] = Hc[j * ld + i])
template <class T>
__global__ __launch_bounds__(256) void k_center_right(T* Hc, int ld, int n, int64_t zb = 0) {
  Hc = zoff(Hc, zb);
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  T s = T(0.0);
#pragma unroll 16
  for (int j = 0; j < n; j++) s += Hc[int64_t(j) * ld + i];  // loads ahead, sum in order
  s = s * (1.0 / n);
#pragma unroll 16
  for (int j = 0; j < n; j++) Hc[int64_t(j) * ld + i] -= s;
}

// mixed: Wr (real FP64 NB x NB, values of the rounded operand) -> T n x n
__global__ __launch_bounds__(256) void k_unpad_real(const double* P, int NB, int n, double* W) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * n) return;
  int i = int(idx / n), j = int(idx % n);
  W[idx] = P[int64_t(i) * NB + j];
}
__global__ __launch_bounds__(256) void k_unembed(const double* P, int NB, int n, cdouble* W, int64_t zb = 0) {
  P = zoff(P, zb);
  W = zoff(W, zb);
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * n) return;
  int a = int(idx / n), b = int(idx % n);
  W[idx] = cdouble(P[int64_t(2 * a) * NB + 2 * b], P[int64_t(2 * a) * NB + 2 * b + 1]);
}

}  // namespace

// ---------------------------------------------------------------------------
// core pipeline
// ---------------------------------------------------------------------------
enum CoreMode { CORE_SVD = 0, CORE_QR = 1, CORE_RAND = 2 };

// Randomized QR / SVD through the factor of the two Gram passes (mixed
// precision, real float32 storage): A = Q_Y M with Q_Y = A W1[:, perm] R2^-1
// orthonormal and M (n x n) FP64.  The reference's algorithm
// (pyLOM/vmmath/svd.py:120-144, 254-273) uses A only through A Omega, A^T Q
// and Q^T A; with A = Q_Y M every sketch is Q_Y times the sketch of M, every
// tsqr factor of Q_Y Y_M is Q_Y times that of Y_M, and Q^T A = Q_M^T M.  So
// randomized_svd(A) = (Q_Y U_M, S_M, V_M) with (U_M, S_M, V_M) the randomized
// SVD of M, same Omega and power steps: computed in FP64 on the device,
// then U = A (W1[:, perm] R2^-1 U_M) in one product pass.  The QR's signs
// are LAPACK's on the m x r Q, read from its first r rows.
struct RandArgs {
  int64_t r;
  int q;
  uint64_t seed;
  const void* omega;                       // n x r (ld r), storage type, or NULL: drawn from seed
  void* Q; int64_t ldq; void* B; int64_t ldb;                 // randomized_qr outputs (or NULL)
  void* U; int64_t ldu; void* S; void* VT; int64_t ldvt;      // randomized_svd outputs (or NULL)
};
static int randomized_core(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, int64_t m, int64_t n, int64_t r,
                           int q, uint64_t seed, const void* omega, void* Qout, int64_t ldq, void* Bout, int64_t ldb,
                           void* U, int64_t ldu, void* S, void* VT, int64_t ldvt, int precision);
static int64_t rand_need(pc_dtype dt, int64_t m, int64_t n, int64_t r, bool draw);
// bytes of the buffers core_rand_tail takes from the cursor
inline int64_t rand_tail_bytes(int64_t n, int64_t r) {
  const int64_t e[] = {n * r * 4, n * r * 8, n * r * 8, r * n * 8, r * 8, n * n * 8, n * r * 8, r * r * 8, r * 8};
  int64_t t = 0;
  for (int64_t b : e) t += round_up(b, 256) + 256;
  return t;
}

template <class TS>
__global__ __launch_bounds__(256) void k_widen(const TS* in, int64_t count, double* out) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < count) out[i] = double(in[i]);
}
// out[i][j] = in[i][j] * (D ? D[i] : 1) for rows x cols, narrowed to the storage type
template <class TO>
__global__ __launch_bounds__(256) void k_narrow_rows(const double* in, int ldi, int rows, int cols, const double* D,
                                                     TO* out, int64_t ldo) {
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(rows) * cols) return;
  const int i = int(idx / cols), j = int(idx % cols);
  out[int64_t(i) * ldo + j] = TO(in[int64_t(i) * ldi + j] * (D ? D[i] : 1.0));
}
// W[:, k] *= D_k (k < cols)
__global__ __launch_bounds__(256) void k_scale_cols(double* W, int ldw, int rows, int cols, const double* D) {
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(rows) * cols) return;
  W[int64_t(idx / cols) * ldw + idx % cols] *= D[idx % cols];
}

// status word of a solve -> error code and message
static int status_error(pc_ctx* ctx, int s0) {
  if (s0 & 1) return pc_fail(ctx, PC_ERR_LINALG, "non-finite values in the input (first Gram pass is not finite)");
  if (s0 & 16) return pc_fail(ctx, PC_ERR_LINALG, "second Gram pass is not finite");
  if (s0 & 2) return pc_fail(ctx, PC_ERR_LINALG, "Jacobi SVD did not converge");
  if (s0 & 4) return pc_fail(ctx, PC_ERR_LINALG, "null-space completion failed (rows of high leverage)");
  if (s0 & 8) return pc_fail(ctx, PC_ERR_LINALG, "non-finite singular values (overflow in the input scale?)");
  return PC_OK;
}

struct CoreOut {
  void* U; int64_t ldu;    // SVD: U ; QR: Q
  void* S;                 // SVD
  void* VT; int64_t ldvt;  // SVD: VT ; QR: R (ld = ldvt)
};

template <class T>
static int launch_streams(pc_ctx* ctx, const Plan& p, const void* A, int64_t lda, int center, double* mean,
                          double* var, double* part, SmallBufs<T>& b, int pass, const PassBatch& bz = PassBatch()) {
  hipStream_t st = ctx->stream;
  int rc = 0;
  pc_dtype sdt = p.cplx ? (p.dt == PC_C64 ? PC_F32 : PC_F64) : p.dt;
  int64_t ldr = p.cplx ? 2 * lda : lda;
  hipEvent_t ev0;
  pc_prof_begin(ctx, &ev0);
  if (p.wide) {
    // P1: FP64 row statistics, then the stripe-pair Gram of Yc; P2: Z = Yc W1
    // written once, then its stripe-pair Gram
    if (pass == 1 && center) {
      if (sdt == PC_F64) launch_row_stats_f64(st, (const double*)A, ldr, p.m, p.nr, mean, var);
      else launch_row_stats_f32(st, (const float*)A, ldr, p.m, p.nr, mean, var);
    }
    if (pass == 1) {
      rc = launch_wgram_f64(st, sdt, A, ldr, p.m, p.nr, center, mean, var, p.chunks, p.tpc, part, p.NB);
    } else {
      rc = launch_wgemm_f64(st, sdt, A, ldr, p.m, p.nr, center, mean, var, b.Wpad, p.NB, p.nr, p.chunks3, p.tpc3,
                            b.Zw, b.ldz, 1);
      if (rc == 0)
        rc = launch_wgram_f64(st, PC_F64, b.Zw, b.ldz, p.m, p.nr, 0, nullptr, nullptr, p.chunks, p.tpc, part, p.NB);
    }
  } else if (p.mixed) {
    // center 1: the kernels shift each row by its first element and P1 records
    // that shift in `mean`; center 2: exact FP64 statistics first
    if (pass == 1 && center == 2) {
      if (bz.nz != 1) return pc_fail(ctx, PC_ERR_INVALID, "batched passes take no variance scaling");
      launch_row_stats_f32(st, (const float*)A, ldr, p.m, p.nr, mean, var);
    }
    if (pass == 1)
      rc = launch_gram_mx2(st, (const float*)A, ldr, p.m, p.nr, center, mean, var, p.chunks, p.tpc, part, bz);
    else
      rc = launch_zgram_mx2(st, (const float*)A, ldr, p.m, p.nr, center, mean, var, b.Wh, b.wscale, p.chunks2,
                            p.tpc2, part, bz, /*upper: W1 = R1^-1, zero-padded*/ true);
  } else {
    if (pass == 1)
      rc = launch_gram_f64(st, sdt, A, ldr, p.m, p.nr, center, 0, nullptr, nullptr, center ? mean : nullptr,
                           center == 2 ? var : nullptr, nullptr, p.chunks, p.tpc, part, bz);
    else
      rc = launch_gram_f64(st, sdt, A, ldr, p.m, p.nr, center, 1, mean, var, nullptr, nullptr, b.Wpad, p.chunks,
                           p.tpc, part, bz, /*upper: W1 = R1^-1, zero-padded*/ true);
  }
  if (rc != 0) return pc_fail(ctx, PC_ERR_SHAPE, "streaming kernel has no variant for n = %d", p.n);
  PC_HIP(ctx, hipGetLastError());
  pc_prof_end(ctx, pass == 1 ? PCP_GRAM : PCP_ZGRAM, ev0);
  return PC_OK;
}

template <class T>
static void reduce_gram(hipStream_t st, const Plan& p, const double* part, int nparts, T* G, int nz = 1,
                        int64_t zb = 0);
// the chunk partials are summed in place into chunk 0 (k_reduce_slab), then
// chunk 0 is unpacked into the symmetric / Hermitian n x n Gram
static void reduce_slab(hipStream_t st, const Plan& p, const double* part, int nparts, int nz, int64_t zb) {
  const int64_t e = int64_t(p.NB) * p.NB;
  hipLaunchKernelGGL(k_reduce_slab, dim3(unsigned(ceil_div(e, 64)), 1, nz), dim3(1024), 0, st,
                     const_cast<double*>(part), nparts, e, p.NB, p.nr, zb);
}
template <>
void reduce_gram<double>(hipStream_t st, const Plan& p, const double* part, int nparts, double* G, int nz,
                         int64_t zb) {
  reduce_slab(st, p, part, nparts, nz, zb);
  hipLaunchKernelGGL(k_reduce_gram_real, dim3(nblk(int64_t(p.n) * p.n), 1, nz), dim3(256), 0, st, part, 1,
                     int64_t(p.NB) * p.NB, p.NB, p.n, G, p.n, zb);
}
template <>
void reduce_gram<cdouble>(hipStream_t st, const Plan& p, const double* part, int nparts, cdouble* G, int nz,
                          int64_t zb) {
  reduce_slab(st, p, part, nparts, nz, zb);
  hipLaunchKernelGGL(k_reduce_gram_cplx, dim3(nblk(int64_t(p.n) * p.n), 1, nz), dim3(256), 0, st, part, 1,
                     int64_t(p.NB) * p.NB, p.NB, p.n, G, p.n, zb);
}

template <class T>
static void to_operand(hipStream_t st, const Plan& p, const T* W, int nc, SmallBufs<T>& b, bool split, int nz = 1,
                       int64_t zb = 0);
// FP64 path: padded / embedded FP64 operand; mixed path: fp16 hi (and lo) + scales,
// with Wr the FP64 value of the rounded hi part (used as the exact preconditioner)
template <>
void to_operand<double>(hipStream_t st, const Plan& p, const double* W, int nc, SmallBufs<double>& b, bool split,
                        int nz, int64_t zb) {
  if (nz != 1) return;  // batched solves are complex only
  if (!p.mixed) {
    hipLaunchKernelGGL(k_pad_real, dim3(nblk(int64_t(p.NB) * p.NB)), dim3(256), 0, st, W, p.n, p.n, nc, b.Wpad, p.NB,
                       p.NB);
  } else {
    hipLaunchKernelGGL(k_pad_real, dim3(nblk(int64_t(p.NB) * p.NB)), dim3(256), 0, st, W, p.n, p.n, nc, b.Wpad, p.NB,
                       p.NB);
    launch_pack_w_mx(st, b.Wpad, p.NB, p.NB, p.NB, p.NB, p.NB, b.Wh, split ? b.Wl : nullptr, b.wscale, b.Wr, p.NB);
  }
}
template <>
void to_operand<cdouble>(hipStream_t st, const Plan& p, const cdouble* W, int nc, SmallBufs<cdouble>& b, bool split,
                         int nz, int64_t zb) {
  hipLaunchKernelGGL(k_embed_pad, dim3(nblk(int64_t(p.NB) * p.NB), 1, nz), dim3(256), 0, st, W, p.n, p.n, nc, b.Wpad,
                     p.NB, p.NB, zb);
  if (p.mixed)
    launch_pack_w_mx(st, b.Wpad, p.NB, p.NB, p.NB, p.NB, p.NB, b.Wh, split ? b.Wl : nullptr, b.wscale, b.Wr, p.NB, nz,
                     zb);
}

template <class T>
static void rounded_back(hipStream_t st, const Plan& p, SmallBufs<T>& b, T* out, int nz = 1, int64_t zb = 0);
template <> void rounded_back<double>(hipStream_t st, const Plan& p, SmallBufs<double>& b, double* out, int nz,
                                      int64_t zb) {
  if (nz != 1) return;  // batched solves are complex only
  hipLaunchKernelGGL(k_unpad_real, dim3(nblk(int64_t(p.n) * p.n)), dim3(256), 0, st, b.Wr, p.NB, p.n, out);
}
template <> void rounded_back<cdouble>(hipStream_t st, const Plan& p, SmallBufs<cdouble>& b, cdouble* out, int nz,
                                       int64_t zb) {
  hipLaunchKernelGGL(k_unembed, dim3(nblk(int64_t(p.n) * p.n), 1, nz), dim3(256), 0, st, b.Wr, p.NB, p.n, out, zb);
}

// P3: U = Y W (the operand prepared by to_operand), then the null-space fix of U
// ---------------------------------------------------------------------------
// Launch-geometry tuner of the product pass.  U = Y W is row-local: how its
// row tiles are spread over workgroups (the chunk count) changes its speed and
// never its result, so it is the parameter tuned.  With tuning enabled on a
// context (pc_autotune), the first product pass of a new (dtype, m, n,
// precision) times every candidate chunk count twice on the real operands
// (rewriting the same U) and keeps the fastest, second timing, in a
// process-wide table that later solves of that shape use on any context.
// Batched and captured launches only read the table.
// ---------------------------------------------------------------------------
namespace tune {
std::mutex mu;
std::map<std::array<int64_t, 4>, int> table;
std::array<int64_t, 4> key(const Plan& p) { return {int64_t(p.dt), p.m, int64_t(p.n), p.mixed ? 1 : 0}; }
}  // namespace tune

template <class T, class ST>
static int launch_p3(pc_ctx* ctx, const Plan& p0, const void* A, int64_t lda, int center, double* mean, double* var,
                     SmallBufs<T>& b, CoreOut out, const PassBatch& bz = PassBatch()) {
  hipStream_t st = ctx->stream;
  pc_dtype sdt = p0.cplx ? (p0.dt == PC_C64 ? PC_F32 : PC_F64) : p0.dt;
  int64_t ldr = p0.cplx ? 2 * lda : lda;
  int64_t ldu = p0.cplx ? 2 * out.ldu : out.ldu;
  Plan p = p0;
  auto launch = [&](int chunks, int64_t tpc) {
    if (p.wide)
      return launch_wgemm_f64(st, sdt, A, ldr, p.m, p.nr, center, mean, var, b.Wpad, p.NB, p.nr, chunks, tpc, out.U,
                              ldu, 0);
    if (p.mixed)
      return launch_gemm_mx2(st, (const float*)A, ldr, p.m, p.nr, center, mean, var, b.Wh, b.Wl, b.wscale, p.nr,
                             chunks, tpc, (float*)out.U, ldu, bz);
    return launch_gemm_f64(st, sdt, A, ldr, p.m, p.nr, center, mean, var, b.Wpad, p.nr, chunks, tpc, out.U, ldu,
                           bz);
  };
  const int64_t tiles = ceil_div(p.m, p.tr3);
  auto geometry = [&](int chunks) {
    p.tpc3 = ceil_div(tiles, int64_t(chunks));
    p.chunks3 = int(ceil_div(tiles, p.tpc3));
  };
  {
    std::unique_lock<std::mutex> lk(tune::mu);
    auto it = tune::table.find(tune::key(p0));
    if (it != tune::table.end()) {
      geometry(it->second);
    } else if (ctx->autotune && bz.nz == 1) {
      lk.unlock();
      hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
      (void)hipStreamIsCapturing(st, &cap);
      if (cap == hipStreamCaptureStatusNone) {
        hipEvent_t e0, e1;
        PC_HIP(ctx, hipEventCreate(&e0));
        PC_HIP(ctx, hipEventCreate(&e1));
        int best = p0.chunks3;
        float best_ms = 0.0f;
        const int cand[] = {p0.chunks3, 256, 512, 768, 1024, 1536, 2048, 3072};
        for (int c : cand) {
          if (c > tiles) continue;
          geometry(c);
          float ms = 0.0f;
          for (int rep = 0; rep < 2; rep++) {
            PC_HIP(ctx, hipEventRecord(e0, st));
            const int rc = launch(p.chunks3, p.tpc3);
            if (rc != 0) return pc_fail(ctx, PC_ERR_SHAPE, "no gemm variant for n = %d (launcher code %d)", p.n, rc);
            PC_HIP(ctx, hipEventRecord(e1, st));
            PC_HIP(ctx, hipEventSynchronize(e1));
            PC_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
          }
          if (best_ms == 0.0f || ms < best_ms) {
            best_ms = ms;
            best = p.chunks3;
          }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        geometry(best);
        std::lock_guard<std::mutex> g(tune::mu);
        tune::table[tune::key(p0)] = best;
      }
    }
  }
  int rc;
  hipEvent_t ev3;
  pc_prof_begin(ctx, &ev3);
  rc = launch(p.chunks3, p.tpc3);
  if (rc != 0) return pc_fail(ctx, PC_ERR_SHAPE, "no gemm variant for n = %d (launcher code %d)", p.n, rc);
  pc_prof_end(ctx, PCP_GEMM, ev3);
  hipLaunchKernelGGL((k_null_fix<T, typename VtType<T, ST>::type>), dim3(nblk(int64_t(p.n) * p.n), 1, bz.nz),
                     dim3(256), 0, st, b.nnull, p.m, p.n, b.jsel, b.cnull, (typename VtType<T, ST>::type*)out.U,
                     out.ldu, bz.zb, bz.zc);
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

template <class T, class ST>
static int core_rand_tail(pc_ctx* ctx, const Plan& p, const void* A, int64_t lda, double* mean, double* var,
                          SmallBufs<double>& b, WsCursor& cur, const RandArgs& ra);

// Work of a lone solve that nothing on the main stream reads until a later
// point runs on the context's side stream: fork() orders the side stream
// after the main stream's work so far, join() orders the main stream after
// the side work.  A fork without its join is joined when the guard goes out
// of scope, on error returns too.  Grouped members (already concurrent), the
// allocation-free entry point (borrowed workspace: it makes no stream
// either) and contexts whose side stream cannot be made run everything on
// the main stream.  Fork and join are capture-compatible.
struct SideWork {
  pc_ctx* ctx;
  hipStream_t st;
  bool pending = false;
  bool use = false;
  SideWork(pc_ctx* c, hipStream_t s) : ctx(c), st(s) {
    if (c->concurrent || c->ws_fixed) return;
    if (!c->side) {
      // the stream and events are made outside any capture of st
      hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return;
      }
    }
    if (!c->side && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      c->side = nullptr;
      return;
    }
    for (hipEvent_t& e : c->sev)
      if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {  // made with the stream
        (void)hipGetLastError();
        e = nullptr;
        return;
      }
    use = true;
  }
  // the stream the off-path work goes to
  hipStream_t fork() {
    if (!use) return st;
    if (hipEventRecord(ctx->sev[0], st) != hipSuccess || hipStreamWaitEvent(ctx->side, ctx->sev[0], 0) != hipSuccess) {
      (void)hipGetLastError();
      use = false;
      return st;
    }
    pending = true;
    return ctx->side;
  }
  void join() {
    if (!pending) return;
    pending = false;
    (void)hipEventRecord(ctx->sev[1], ctx->side);
    (void)hipStreamWaitEvent(st, ctx->sev[1], 0);
  }
  ~SideWork() { join(); }
};

template <class T, class ST>
static int core_t(pc_ctx* ctx, const Plan& p, const void* A, int64_t lda, int center, int mode, CoreOut out,
                  const RandArgs* ra = nullptr) {
  hipStream_t st = ctx->stream;
  int n = p.n;
  int64_t nn = int64_t(n) * n;
  // workspace
  SmallBufs<T> b;
  double *mean, *var, *part;
  int64_t slabs = std::max(p.nparts, p.mixed ? p.chunks3 : p.chunks);
  int64_t base = ctx->ws_skip;
  {
    WsCursor probe{nullptr, 0, 0};
    int64_t need = carve<T>(probe, p, b, &mean, &var, &part, slabs) + dense_work(n) * int64_t(sizeof(T));
    // inside a grouped call the status words live in the workspace: it must
    // not move, so the group reserves the largest problem's need up front
    if (ctx->gstat && base + need + 4096 > ctx->ws_bytes)
      return pc_fail(ctx, PC_ERR_INVALID, "grouped workspace smaller than a member's need");
    if (mode == CORE_RAND)  // the factor's randomized SVD and the product pass, nested in this front
      need += rand_tail_bytes(n, ra->r) + rand_need(PC_F64, n, n, ra->r, false) + 4096;
    PC_CHECK(ctx, pc_ws_reserve(ctx, base + need + 4096));
  }
  WsCursor cur{ctx->ws + base, 0, ctx->ws_bytes - base};
  carve<T>(cur, p, b, &mean, &var, &part, slabs);
  T* nwork = cur.take<T>(dense_work(n));
  // status words cleared by a kernel (no memset node when the solve is
  // captured into a graph)
  hipLaunchKernelGGL(k_status_copy, dim3(1), dim3(64), 0, st, nullptr, b.status);

  // ---- P1
  PC_CHECK(ctx, launch_streams<T>(ctx, p, A, lda, center, mean, var, part, b, 1));
  reduce_gram<T>(st, p, part, p.nparts, b.G1);
  // rank[1] / rank[2]: state words of the multi-workgroup factorisations
  run_chol_shift<T>(st, b.G1, n, p.mixed ? kShiftMx : kShiftF64, b.R1, b.status, 1, 0, b.Awk, b.dsc, b.rank + 2);
  run_trinv<T>(st, b.R1, n, b.W1);
  to_operand<T>(st, p, b.W1, n, b, false);
  if (p.mixed) rounded_back<T>(st, p, b, b.W1);  // W1 := the operand actually applied
  // (W1^-1 stays on the main stream: beside P2 its workgroups, with
  // 133 KB of LDS each, would take whole CUs from P2)
  run_trinv<T>(st, b.W1, n, b.R1h);
  SideWork side(ctx, st);
  // ---- P2
  if (p.wide && !b.Zw) {  // FP64 storage: Z lives in U until P3 overwrites it
    b.Zw = static_cast<double*>(out.U);
    b.ldz = p.cplx ? 2 * out.ldu : out.ldu;
  }
  PC_CHECK(ctx, launch_streams<T>(ctx, p, A, lda, center, mean, var, part, b, 2));
  reduce_gram<T>(st, p, part, p.mixed ? p.chunks2 : p.chunks, b.G2);
  run_pchol<T>(st, b.G2, n, p.mixed ? kTolMx : kTolF64, b.Awk, b.R2, b.perm, b.rank, b.status, 1, 0, b.rank + 1);
  // M = R2 (rows < rank) * R1h[perm, :], stored column-major into Hc
  dim3 gg(unsigned(ceil_div(n, 32)), unsigned(ceil_div(n, 32)));
  hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.R2, n, (const int*)nullptr,
                     b.R1h, n, (const int*)b.perm, T(0.0), b.Hc, n, 1);
  // mixed path: the passes saw Y - a 1^T (a = row shift); Y C = (Y - a 1^T) C, so the
  // exact centring is M <- M C
  if (p.mixed && center)
    hipLaunchKernelGGL((k_center_right<T>), dim3(nblk(n)), dim3(256), 0, st, b.Hc, n, n);
  hipLaunchKernelGGL((k_pad_identity<T>), dim3(1), dim3(256), 0, st, b.R2, n, n, b.rank);
  // R2i = R2^-1 is read first in W, after the Jacobi: it runs beside the
  // preconditioner and the Jacobi (no kernel between writes R2 or R2i)
  run_trinv<T>(side.fork(), b.R2, n, b.R2i);
  if (mode == CORE_RAND) {
    side.join();
    if constexpr (std::is_same<T, double>::value && std::is_same<ST, float>::value)
      return core_rand_tail(ctx, p, A, lda, mean, var, b, cur, *ra);
    else
      return pc_fail(ctx, PC_ERR_UNSUPPORTED, "randomized through the factor: float32 storage only");
  }
  int nc = n;
  bool gram_steps_;
  int jb_, jnb2_;
  jacobi_blocks<T>(n, n, gram_steps_, jb_, jnb2_);
  if (mode == CORE_SVD && jnb2_ == 2) {
    // every column fits one block pair: the Jacobi launch iterates to
    // convergence on its own, so M (Hc, column-major) is rotated directly and
    // V = J; no preconditioner is needed
    launch_block_jacobi<T>(ctx, b.Hc, n, n, n, b.J, n, b.sig, b.jstate, b.jdstate, b.status,
                           p.mixed ? kJacTolMx : 0.0);
    hipLaunchKernelGGL(k_sort_desc, dim3(1), dim3(1024), 0, st, b.sig, n, b.order);
    hipLaunchKernelGGL((k_svd_finish<T, ST, typename VtType<T, ST>::type>), dim3(nblk(nn)), dim3(256), 0, st, b.Hc,
                       n, n, n, b.sig, b.order, b.J, n, p.mixed ? kNullMx : kNullF64, (const int*)b.rank,
                       (ST*)out.S, (typename VtType<T, ST>::type*)out.VT, out.ldvt, b.UM, n, b.nnull);
  } else if (mode == CORE_SVD) {
    // Right preconditioner V0 = P3 Q4 (Drmac-Veselic): P3^T (M^H M) P3 = R3^H R3
    // (pivoted Cholesky), R3^H = Q4 R4 (Householder).  One-sided Jacobi on
    // M V0 takes the same rotations as on the lower-triangular R4^H and
    // converges in a few sweeps; V0 only has to be unitary, so R3 coming from
    // the Gram is fine.  Buffers: tmp = M (row-major), G1 = M^H M then R4,
    // R1 = R3, Awk = R3^H then V0, UM = Q4.
    hipLaunchKernelGGL((k_copy2d<T>), dim3(nblk(nn)), dim3(256), 0, st, b.Hc, n, n, n, b.tmp, n, 1);
    hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 2, 0, n, n, n, T(1.0), b.tmp, n, (const int*)nullptr,
                       b.tmp, n, (const int*)nullptr, T(0.0), b.G1, n, 0);
    run_pchol<T>(st, b.G1, n, 1e-30, b.Awk, b.R1, b.perm3, b.rank3, b.status, 1, 0, b.rank3 + 1);
    hipLaunchKernelGGL((k_copy2d<T>), dim3(nblk(nn)), dim3(256), 0, st, b.R1, n, n, n, b.Awk, n, 2);
    run_house_qr_sq<T>(st, b.Awk, n, b.UM, b.G1, nwork, ctx->concurrent);
    hipLaunchKernelGGL((k_permute_rows<T>), dim3(nblk(nn)), dim3(256), 0, st, b.UM, n, n, (const int*)b.perm3, b.Awk,
                       n);
    hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.tmp, n, (const int*)nullptr,
                       b.Awk, n, (const int*)nullptr, T(0.0), b.Hc, n, 1);
    // mixed precision: the factor carries ~1e-7 relative data error, so the
    // columns need only be orthogonal to 1e-8 (sigma errs by O(tol^2))
    launch_block_jacobi<T>(ctx, b.Hc, n, n, n, b.J, n, b.sig, b.jstate, b.jdstate, b.status,
                           p.mixed ? kJacTolMx : 0.0);
    // V = V0 J (column-major like J)
    hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 1, n, n, n, T(1.0), b.Awk, n, (const int*)nullptr,
                       b.J, n, (const int*)nullptr, T(0.0), b.tmp, n, 1);
    hipLaunchKernelGGL(k_sort_desc, dim3(1), dim3(1024), 0, st, b.sig, n, b.order);
    hipLaunchKernelGGL((k_svd_finish<T, ST, typename VtType<T, ST>::type>), dim3(nblk(nn)), dim3(256), 0, st, b.Hc,
                       n, n, n, b.sig, b.order, b.tmp, n, p.mixed ? kNullMx : kNullF64, (const int*)b.rank,
                       (ST*)out.S, (typename VtType<T, ST>::type*)out.VT, out.ldvt, b.UM, n, b.nnull);
  } else {
    // QR of M (n x n, rows >= rank zero): M = Q_M R_M;  R output, UM := Q_M
    hipLaunchKernelGGL((k_copy2d<T>), dim3(nblk(nn)), dim3(256), 0, st, b.Hc, n, n, n, b.tmp, n, 1);
    run_house_qr_sq<T>(st, b.tmp, n, b.UM, b.J, nwork, ctx->concurrent);
    // diag(R) real and non-negative (the phases go to Q_M's columns); tmp is free here
    hipLaunchKernelGGL((k_qr_phase<T>), dim3(nblk(n)), dim3(256), 0, st, b.J, n, n, b.tmp);
    hipLaunchKernelGGL((k_qr_apply_phase<T>), dim3(nblk(nn)), dim3(256), 0, st, b.J, n, b.UM, n, n,
                       (const T*)b.tmp);
    // R is stored after the LAPACK signs are known (below); the null-space
    // completion needs the count of null columns now
    hipLaunchKernelGGL(k_null_count, dim3(1), dim3(64), 0, st, (const int*)b.rank, n, b.nnull);
  }
  // W = W1[:, perm] * (R2^-1 * UM)
  side.join();
  hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.R2i, n, (const int*)nullptr,
                     b.UM, n, (const int*)nullptr, T(0.0), b.tmp, n, 0);
  hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.W1, n, (const int*)b.perm,
                     b.tmp, n, (const int*)nullptr, T(0.0), b.Wf, n, 0);
  // null completion (kernels return at once when nnull == 0)
  int ncand = int(std::min<int64_t>(p.m, 4 * int64_t(n) + 16));
  launch_null_lev<T>(st, p, A, lda, (p.mixed && center == 1) ? 3 : center, mean, var, b, ncand);
  hipLaunchKernelGGL((k_null_build<T>), dim3(1), dim3(1024), null_build_lds(n, ncand), st, p.m, n, b.nnull, ncand, b.lev, b.urows, b.Wf, n,
                     b.jsel, b.cnull, nwork, b.status);
  if (mode == CORE_QR) {
    // LAPACK's signs (numpy.linalg.qr): D from the first n rows of Q, then
    // R = D R+, Q = Q+ D through W and the null fix; Q1 in tmp, D in G2
    launch_q1_rows<T>(st, p, A, lda, b);
    hipLaunchKernelGGL((k_lapack_signs<T>), dim3(1), dim3(1024), 0, st, b.tmp, n, b.G2);
    hipLaunchKernelGGL((k_apply_signs<T>), dim3(nblk(nn)), dim3(256), 0, st, (const T*)b.G2, n, b.J, n, b.Wf, n,
                       (const int*)b.nnull, b.cnull);
    hipLaunchKernelGGL((k_store_r<T, typename VtType<T, ST>::type>), dim3(nblk(nn)), dim3(256), 0, st, b.J, n, n,
                       (typename VtType<T, ST>::type*)out.VT, out.ldvt, b.rank, b.nnull);
  }
  // ---- P3
  to_operand<T>(st, p, b.Wf, nc, b, true);
  PC_CHECK(ctx, (launch_p3<T, ST>(ctx, p, A, lda, center, mean, var, b, out)));
  if (ctx->gstat) {  // grouped solve: statuses are read once for the whole group
    hipLaunchKernelGGL(k_status_copy, dim3(1), dim3(64), 0, st, b.status, ctx->gstat);
    PC_HIP(ctx, hipGetLastError());
    return PC_OK;
  }
  // status check: one small D2H copy, synchronous (errors must surface here)
  PC_HIP(ctx, hipMemcpyAsync(ctx->hflag, b.status, 16, hipMemcpyDeviceToHost, st));
  PC_HIP(ctx, hipStreamSynchronize(st));
  return status_error(ctx, ctx->hflag[0]);
}

// ---------------------------------------------------------------------------
// Batched solve of a uniform group: every problem has the same shape, leading
// dimensions and centring, and A, U, S and VT sit at constant strides (the
// per-frequency blocks of a spectral POD).  Every kernel of the solve, the
// streaming passes included, is one launch over all problems (blockIdx.z =
// problem, workspace slices zb bytes apart).  Kernels and arguments are those
// of core_t's one-block-pair path, so each result is bit-identical to a lone
// solve of that problem.
// ---------------------------------------------------------------------------
struct BatchDesc {
  int64_t zA, zU, zS, zV;  // byte strides of A, U, S and VT between consecutive problems
};

inline int64_t byte_diff(const void* a, const void* b) {
  return int64_t(reinterpret_cast<const char*>(a) - reinterpret_cast<const char*>(b));
}

// complex SVD problems whose Jacobi is one block pair on one wavefront
// (k_bjac_small), uniform as described above
static bool batch_eligible(pc_dtype dt, const pc_problem* pr, int64_t count, BatchDesc& d) {
  if (count < 2 || count > 65535 || !dtype_complex(dt)) return false;
  const pc_problem& q = pr[0];
  bool gs;
  int jb, jnb2;
  jacobi_blocks<cdouble>(int(q.n), int(q.n), gs, jb, jnb2);
  if (!(jnb2 == 2 && q.n <= 64 && 2 * jb <= 32)) return false;
  d.zA = byte_diff(pr[1].A, q.A);
  d.zU = byte_diff(pr[1].U, q.U);
  d.zS = byte_diff(pr[1].S, q.S);
  d.zV = byte_diff(pr[1].VT, q.VT);
  if (d.zU == 0 || d.zS == 0 || d.zV == 0) return false;  // outputs must not alias
  for (int64_t i = 1; i < count; i++) {
    const pc_problem& r = pr[i];
    if (r.m != q.m || r.n != q.n || r.lda != q.lda || r.ldu != q.ldu || r.ldvt != q.ldvt || r.center != q.center)
      return false;
    if (byte_diff(r.A, q.A) != i * d.zA || byte_diff(r.U, q.U) != i * d.zU || byte_diff(r.S, q.S) != i * d.zS ||
        byte_diff(r.VT, q.VT) != i * d.zV)
      return false;
  }
  return true;
}

// Runs the batch on ctx->stream: problem i's workspace slice starts at
// ctx->ws + base + i zb, its four status words are copied to gstat + 4 i.
template <class T, class ST>
static int core_batch_t(pc_ctx* ctx, const Plan& p, const pc_problem* pr, int64_t count, const BatchDesc& d,
                        int64_t base, int64_t zb, int* gstat) {
  hipStream_t st = ctx->stream;
  const int n = p.n;
  const int64_t nn = int64_t(n) * n;
  const int nz = int(count);
  const int center = pr[0].center;
  const int64_t lda = pr[0].lda;
  const int64_t slabs = std::max(p.nparts, p.mixed ? p.chunks3 : p.chunks);
  SmallBufs<T> b;
  double *mean, *var, *part;
  WsCursor cur{ctx->ws + base, 0, zb};
  carve<T>(cur, p, b, &mean, &var, &part, slabs);
  T* nwork = cur.take<T>(dense_work(n));
  const PassBatch pb{nz, d.zA, zb, d.zU};
  const unsigned z = unsigned(nz);
  hipLaunchKernelGGL(k_status_copy, dim3(1, 1, z), dim3(64), 0, st, nullptr, b.status, int64_t(0), zb);
  // ---- P1
  PC_CHECK(ctx, launch_streams<T>(ctx, p, pr[0].A, lda, center, mean, var, part, b, 1, pb));
  reduce_gram<T>(st, p, part, p.nparts, b.G1, nz, zb);
  run_chol_shift<T>(st, b.G1, n, p.mixed ? kShiftMx : kShiftF64, b.R1, b.status, nz, zb);
  run_trinv<T>(st, b.R1, n, b.W1, nz, zb);
  to_operand<T>(st, p, b.W1, n, b, false, nz, zb);
  if (p.mixed) rounded_back<T>(st, p, b, b.W1, nz, zb);
  run_trinv<T>(st, b.W1, n, b.R1h, nz, zb);
  // ---- P2
  PC_CHECK(ctx, launch_streams<T>(ctx, p, pr[0].A, lda, center, mean, var, part, b, 2, pb));
  reduce_gram<T>(st, p, part, p.mixed ? p.chunks2 : p.chunks, b.G2, nz, zb);
  run_pchol<T>(st, b.G2, n, p.mixed ? kTolMx : kTolF64, b.Awk, b.R2, b.perm, b.rank, b.status, nz, zb);
  dim3 gg(unsigned(ceil_div(n, 32)), unsigned(ceil_div(n, 32)), z);
  hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.R2, n, (const int*)nullptr,
                     b.R1h, n, (const int*)b.perm, T(0.0), b.Hc, n, 1, zb);
  if (p.mixed && center)
    hipLaunchKernelGGL((k_center_right<T>), dim3(nblk(n), 1, z), dim3(256), 0, st, b.Hc, n, n, zb);
  hipLaunchKernelGGL((k_pad_identity<T>), dim3(1, 1, z), dim3(256), 0, st, b.R2, n, n, b.rank, zb);
  run_trinv<T>(st, b.R2, n, b.R2i, nz, zb);
  launch_block_jacobi<T>(ctx, b.Hc, n, n, n, b.J, n, b.sig, b.jstate, b.jdstate, b.status,
                         p.mixed ? kJacTolMx : 0.0, nz, zb);
  hipLaunchKernelGGL(k_sort_desc, dim3(1, 1, z), dim3(1024), 0, st, b.sig, n, b.order, zb);
  hipLaunchKernelGGL((k_svd_finish<T, ST, typename VtType<T, ST>::type>), dim3(nblk(nn), 1, z), dim3(256), 0, st,
                     b.Hc, n, n, n, b.sig, b.order, b.J, n, p.mixed ? kNullMx : kNullF64, (const int*)b.rank,
                     (ST*)pr[0].S, (typename VtType<T, ST>::type*)pr[0].VT, pr[0].ldvt, b.UM, n, b.nnull, zb, d.zS,
                     d.zV);
  hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.R2i, n, (const int*)nullptr,
                     b.UM, n, (const int*)nullptr, T(0.0), b.tmp, n, 0, zb);
  hipLaunchKernelGGL((k_gemm_small<T>), gg, dim3(256), 0, st, 0, 0, n, n, n, T(1.0), b.W1, n, (const int*)b.perm,
                     b.tmp, n, (const int*)nullptr, T(0.0), b.Wf, n, 0, zb);
  const int ncand = int(std::min<int64_t>(p.m, 4 * int64_t(n) + 16));
  launch_null_lev<T>(st, p, pr[0].A, lda, (p.mixed && center == 1) ? 3 : center, mean, var, b, ncand, nz, zb, d.zA);
  hipLaunchKernelGGL((k_null_build<T>), dim3(1, 1, z), dim3(1024), null_build_lds(n, ncand), st, p.m, n, b.nnull, ncand, b.lev, b.urows,
                     b.Wf, n, b.jsel, b.cnull, nwork, b.status, zb);
  to_operand<T>(st, p, b.Wf, n, b, true, nz, zb);
  // ---- P3 (and the null-space fix of each U)
  CoreOut o0{pr[0].U, pr[0].ldu, pr[0].S, pr[0].VT, pr[0].ldvt};
  PC_CHECK(ctx, (launch_p3<T, ST>(ctx, p, pr[0].A, lda, center, mean, var, b, o0, pb)));
  hipLaunchKernelGGL(k_status_copy, dim3(1, 1, z), dim3(64), 0, st, b.status, gstat, zb, int64_t(16));
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

static int core_batch(pc_ctx* ctx, pc_dtype dt, const Plan& p, const pc_problem* pr, int64_t count,
                      const BatchDesc& d, int64_t base, int64_t zb, int* gstat) {
  if (dt == PC_C64) return core_batch_t<cdouble, float>(ctx, p, pr, count, d, base, zb, gstat);
  return core_batch_t<cdouble, double>(ctx, p, pr, count, d, base, zb, gstat);
}

static int core_dispatch(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, int64_t m, int64_t n, int center,
                         int mode, CoreOut out, int precision) {
  Plan p;
  PC_CHECK(ctx, make_plan(ctx, dt, m, n, precision, p));
  switch (dt) {
    case PC_F64: return core_t<double, double>(ctx, p, A, lda, center, mode, out);
    case PC_F32: return core_t<double, float>(ctx, p, A, lda, center, mode, out);
    case PC_C128: return core_t<cdouble, double>(ctx, p, A, lda, center, mode, out);
    case PC_C64: return core_t<cdouble, float>(ctx, p, A, lda, center, mode, out);
  }
  return pc_fail(ctx, PC_ERR_INVALID, "bad dtype");
}

static int check_common(pc_ctx* ctx, pc_dtype dt, int64_t m, int64_t n) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "ctx is NULL");
  if (dt < PC_F32 || dt > PC_C128) return pc_fail(ctx, PC_ERR_INVALID, "bad dtype %d", int(dt));
  if (n <= 0 || m <= 0) return pc_fail(ctx, PC_ERR_SHAPE, "empty matrix (m=%lld, n=%lld)", (long long)m, (long long)n);
  if (m < n) return pc_fail(ctx, PC_ERR_SHAPE, "tall-skinny solve needs m >= n (m=%lld, n=%lld)", (long long)m, (long long)n);
  if (n > (dtype_complex(dt) ? kWideMaxReal / 2 : kWideMaxReal))
    return pc_fail(ctx, PC_ERR_SHAPE, "n = %lld > %d is not supported", (long long)n,
                   dtype_complex(dt) ? kWideMaxReal / 2 : kWideMaxReal);
  PC_HIP(ctx, hipSetDevice(ctx->device));
  return PC_OK;
}

extern "C" int pc_tsqr_svd(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, int64_t m, int64_t n, int center,
                           void* U, int64_t ldu, void* S, void* VT, int64_t ldvt, int precision) {
  PC_CHECK(ctx, check_common(ctx, dt, m, n));
  PC_CHECK(ctx, pc_check_mat(ctx, A, lda, m, n, "pc_tsqr_svd A"));
  PC_CHECK(ctx, pc_check_mat(ctx, U, ldu, m, n, "pc_tsqr_svd U"));
  PC_CHECK(ctx, pc_check_mat(ctx, VT, ldvt, n, n, "pc_tsqr_svd VT"));
  if (!S) return pc_fail(ctx, PC_ERR_INVALID, "pc_tsqr_svd: S is NULL");
  if (center < 0 || center > 2) return pc_fail(ctx, PC_ERR_INVALID, "bad center mode %d", center);
  if (center && dtype_complex(dt)) return pc_fail(ctx, PC_ERR_UNSUPPORTED, "row centring is for real data");
  CoreOut o{U, ldu, S, VT, ldvt};
  return core_dispatch(ctx, dt, A, lda, m, n, center, CORE_SVD, o, precision);
}

extern "C" int pc_tsqr(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, int64_t m, int64_t n, void* Q,
                       int64_t ldq, void* R, int64_t ldr, int precision) {
  PC_CHECK(ctx, check_common(ctx, dt, m, n));
  PC_CHECK(ctx, pc_check_mat(ctx, A, lda, m, n, "pc_tsqr A"));
  PC_CHECK(ctx, pc_check_mat(ctx, Q, ldq, m, n, "pc_tsqr Q"));
  PC_CHECK(ctx, pc_check_mat(ctx, R, ldr, n, n, "pc_tsqr R"));
  CoreOut o{Q, ldq, nullptr, R, ldr};
  return core_dispatch(ctx, dt, A, lda, m, n, 0, CORE_QR, o, precision);
}

#include "ops.inc"
#include "dmd.inc"

// ---------------------------------------------------------------------------
// Allocation-free tall-skinny SVD: the workspace of pc_tsqr_svd taken from a
// file-scope device array instead of the context's (grown with hipMalloc).
// Each GPU holds its own copy of `scratch` at its own address; the address is
// taken once per device at context creation (pc_scratch_attach) and only that
// address is handed to kernels.  Calls on one GPU share the array, so calls
// that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int64_t kScratchBytes = int64_t(128) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static void* g_scratch_ptr[kScratchDevices] = {};
static std::mutex g_scratch_mu;

int pc_scratch_attach(int device) {
  if (device < 0 || device >= kScratchDevices) return PC_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return PC_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PC_ERR_HIP;
  g_scratch_ptr[device] = ptr;
  return PC_OK;
}

extern "C" int pc_tsqr_svd_noalloc(pc_ctx* ctx, pc_dtype dt, const void* A, int64_t lda, int64_t m, int64_t n,
                                   int center, void* U, int64_t ldu, void* S, void* VT, int64_t ldvt, int precision) {
  PC_CHECK(ctx, check_common(ctx, dt, m, n));
  PC_CHECK(ctx, pc_check_mat(ctx, A, lda, m, n, "pc_tsqr_svd_noalloc A"));
  PC_CHECK(ctx, pc_check_mat(ctx, U, ldu, m, n, "pc_tsqr_svd_noalloc U"));
  PC_CHECK(ctx, pc_check_mat(ctx, VT, ldvt, n, n, "pc_tsqr_svd_noalloc VT"));
  if (!S) return pc_fail(ctx, PC_ERR_INVALID, "pc_tsqr_svd_noalloc: S is NULL");
  if (center < 0 || center > 2) return pc_fail(ctx, PC_ERR_INVALID, "bad center mode %d", center);
  if (center && dtype_complex(dt)) return pc_fail(ctx, PC_ERR_UNSUPPORTED, "row centring is for real data");
  if (ctx->gstat) return pc_fail(ctx, PC_ERR_INVALID, "pc_tsqr_svd_noalloc: not inside a grouped solve");
  Plan p;
  PC_CHECK(ctx, make_plan(ctx, dt, m, n, precision, p));
  const int64_t need = p.cplx ? core_need<cdouble>(p) : core_need<double>(p);
  if (need > kScratchBytes)
    return pc_fail(ctx, PC_ERR_NOMEM,
                   "pc_tsqr_svd_noalloc: m = %lld, n = %lld needs %lld bytes of scratch, more than its %lld "
                   "(pc_tsqr_svd takes it)", (long long)m, (long long)n, (long long)need, (long long)kScratchBytes);
  void* sp;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    sp = ctx->device >= 0 && ctx->device < kScratchDevices ? g_scratch_ptr[ctx->device] : nullptr;
  }
  if (!sp) return pc_fail(ctx, PC_ERR_HIP, "pc_tsqr_svd_noalloc: no static scratch on device %d", ctx->device);
  // the context's own workspace stays where it is (work queued on it is
  // unaffected); this call runs on the static scratch and never grows it
  char* ws0 = ctx->ws;
  const int64_t bytes0 = ctx->ws_bytes, skip0 = ctx->ws_skip;
  const bool fixed0 = ctx->ws_fixed;
  ctx->ws = static_cast<char*>(sp);
  ctx->ws_bytes = kScratchBytes;
  ctx->ws_skip = 0;
  ctx->ws_fixed = true;
  CoreOut o{U, ldu, S, VT, ldvt};
  const int rc = core_dispatch(ctx, dt, A, lda, m, n, center, CORE_SVD, o, precision);
  ctx->ws = ws0;
  ctx->ws_bytes = bytes0;
  ctx->ws_skip = skip0;
  ctx->ws_fixed = fixed0;
  return rc;
}

extern "C" int pc_autotune(pc_ctx* ctx, int enable) {
  if (!ctx || enable < 0 || enable > 1) return pc_fail(ctx, PC_ERR_INVALID, "pc_autotune: bad arguments");
  ctx->autotune = enable;
  return PC_OK;
}

extern "C" int pc_autotune_query(pc_dtype dtype, int64_t m, int64_t n, int precision, int64_t* chunks) {
  if (!chunks) return PC_ERR_INVALID;
  std::lock_guard<std::mutex> g(tune::mu);
  auto it = tune::table.find({int64_t(dtype), m, n, precision == PC_PREC_MIXED ? 1 : 0});
  if (it == tune::table.end()) return 0;
  *chunks = it->second;
  return 1;
}

extern "C" int pc_autotune_clear(void) {
  std::lock_guard<std::mutex> g(tune::mu);
  tune::table.clear();
  return PC_OK;
}
