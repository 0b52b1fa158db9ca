// Small dense factorisations of wide problems (n > 256) spread over the GPU:
// one launch per elimination step, the step's updates spread over many
// workgroups (one wave per row or column).  The single-workgroup kernels of
// k_small.hip serialise every step on one CU, which costs seconds at
// n = 1000; here a step costs one launch and its share of L2 bandwidth.
// Included by solve.hip after k_small.hip (shares its helpers).
//
//   Householder QR  : work copy column-major (coalesced column reads);
//                     per column k: k_hqr_reflect (one workgroup: ?larfg's
//                     beta = -sign(Re alpha) ||x||, tau, v), k_hqr_apply
//                     (every trailing column j > k: w = v^H a_j,
//                     a_j -= conj(tau) w v); Q = H_0 ... H_{k-1} [I; 0]
//                     accumulated backward by k_hqr_qapply.
//   Shifted Cholesky: right-looking on the upper triangle with row k kept
//                     unscaled during the steps (A[i][j] -= conj(A[k][i])
//                     A[k][j] / A[k][k]), scaled at the end.  A pivot below
//                     tau / 2 flags the attempt; the single-workgroup kernel
//                     then redoes it with a growing shift (rare: only when
//                     rounding makes G + tau I indefinite).
//   Pivoted Cholesky: per step k_pchol_pivot (one workgroup: argmax of the
//                     remaining diagonal, symmetric swap, row k of R) and
//                     k_pchol_update (trailing Schur complement); it stops at
//                     the first pivot <= tol max(diag G) like k_pchol.
namespace {

constexpr int MW = 256;  // threads per workgroup (4 waves)

// ---- Householder QR ---------------------------------------------------------
template <class T>
__global__ __launch_bounds__(MW) void k_hqr_init(const T* A, int lda, int rows, int cols, int kmax, T* Wc, T* Qc) {
  const int64_t tot = int64_t(rows) * cols, totq = int64_t(rows) * kmax;
  for (int64_t idx = int64_t(blockIdx.x) * MW + threadIdx.x; idx < tot + totq; idx += int64_t(gridDim.x) * MW) {
    if (idx < tot) {
      const int j = int(idx / rows), i = int(idx % rows);
      Wc[idx] = A[int64_t(i) * lda + j];
    } else {
      const int64_t q = idx - tot;
      const int j = int(q / rows), i = int(q % rows);
      Qc[q] = T(i == j ? 1.0 : 0.0);
    }
  }
}

template <class T>
__global__ __launch_bounds__(MW) void k_hqr_reflect(T* Wc, int rows, int k, T* tau) {
  __shared__ double sh[MW / 64];
  __shared__ T scale_s;
  T* x = Wc + int64_t(k) * rows;
  double s = 0.0;
  for (int i = k + 1 + threadIdx.x; i < rows; i += MW) s += abs2_(x[i]);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = 0.0;
    for (int q = 0; q < MW / 64; q++) ss += sh[q];
    const T al = x[k];
    const double re = real_(al);
    if (ss == 0.0 && abs2_(al - T(re)) == 0.0) {  // already reduced: no reflection (tau = 0)
      tau[k] = T(0.0);
      scale_s = T(0.0);
      x[k] = T(re);
    } else {
      const double beta = -(re >= 0 ? 1.0 : -1.0) * sqrt(abs2_(al) + ss);
      tau[k] = (T(beta) - al) * (1.0 / beta);
      scale_s = inv_(al - T(beta));
      x[k] = T(beta);
    }
  }
  __syncthreads();
  const T sc = scale_s;
  if (!(real_(sc) == 0.0 && abs2_(sc) == 0.0))
    for (int i = k + 1 + threadIdx.x; i < rows; i += MW) x[i] = x[i] * sc;
}

// a_j -= conj(tau_k) (v^H a_j) v for the trailing columns j in (k, cols), one wave each
template <class T>
__global__ __launch_bounds__(MW) void k_hqr_apply(T* Wc, int rows, int cols, int k, const T* tau) {
  const int j = k + 1 + blockIdx.x * (MW / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= cols) return;
  const T* v = Wc + int64_t(k) * rows;
  T* a = Wc + int64_t(j) * rows;
  T w = T(0.0);
  for (int i = k + lane; i < rows; i += 64) w += dotc_(i == k ? T(1.0) : v[i], a[i]);
  w = group_sum_t(w, 64);
  const T f = conj_(tau[k]) * w;
  for (int i = k + lane; i < rows; i += 64) a[i] -= (i == k ? T(1.0) : v[i]) * f;
}

// Q := H_k Q on the columns j in [k, kmax), one wave each (backward accumulation)
template <class T>
__global__ __launch_bounds__(MW) void k_hqr_qapply(const T* Wc, int rows, int kmax, int k, const T* tau, T* Qc) {
  const int j = k + blockIdx.x * (MW / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= kmax) return;
  const T* v = Wc + int64_t(k) * rows;
  T* q = Qc + int64_t(j) * rows;
  T w = T(0.0);
  for (int i = k + lane; i < rows; i += 64) w += dotc_(i == k ? T(1.0) : v[i], q[i]);
  w = group_sum_t(w, 64);
  const T f = tau[k] * w;
  for (int i = k + lane; i < rows; i += 64) q[i] -= (i == k ? T(1.0) : v[i]) * f;
}

template <class T>
__global__ __launch_bounds__(MW) void k_hqr_out(const T* Wc, const T* Qc, int rows, int cols, int kmax, T* Q, int ldq,
                                                T* R, int ldr) {
  const int64_t totr = int64_t(cols) * cols, totq = int64_t(rows) * kmax;
  for (int64_t idx = int64_t(blockIdx.x) * MW + threadIdx.x; idx < totr + totq; idx += int64_t(gridDim.x) * MW) {
    if (idx < totr) {
      const int i = int(idx / cols), j = int(idx % cols);
      R[int64_t(i) * ldr + j] = (i <= j && i < rows) ? Wc[int64_t(j) * rows + i] : T(0.0);
    } else {
      const int64_t q = idx - totr;
      const int i = int(q / kmax), j = int(q % kmax);
      Q[int64_t(i) * ldq + j] = Qc[int64_t(j) * rows + i];
    }
  }
}

// ---- shifted Cholesky -------------------------------------------------------
// tau = shift_rel max(diag G) (1 if not positive), flag = 0
template <class T>
__global__ __launch_bounds__(1024) void k_chol_mw_tau(const T* G, int ldg, int n, double shift_rel, double* tau,
                                                      int* flag) {
  __shared__ double sh[32];
  double dmax = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) dmax = fmax(dmax, real_(G[int64_t(i) * ldg + i]));
  dmax = block_max(dmax, sh);
  if (threadIdx.x == 0) {
    double t = shift_rel * (isfinite(dmax) ? dmax : 1.0);
    tau[0] = t > 0.0 ? t : 1.0;
    flag[0] = isfinite(dmax) ? 0 : 1;
  }
}
// A = upper(G) + tau I (lower part zero)
template <class T>
__global__ __launch_bounds__(MW) void k_chol_mw_copy(const T* G, int ldg, int n, const double* tau, T* A) {
  const int64_t tot = int64_t(n) * n;
  for (int64_t idx = int64_t(blockIdx.x) * MW + threadIdx.x; idx < tot; idx += int64_t(gridDim.x) * MW) {
    const int i = int(idx / n), j = int(idx % n);
    T v = i <= j ? G[int64_t(i) * ldg + j] : T(0.0);
    if (i == j) v = v + T(tau[0]);
    A[idx] = v;
  }
}
// step k: rows i in (k, n), one wave each: A[i][j] -= conj(A[k][i]) A[k][j] / A[k][k], j >= i
template <class T>
__global__ __launch_bounds__(MW) void k_chol_mw_step(T* A, int n, int k, const double* tau, int* flag) {
  if (flag[0]) return;
  const double d = real_(A[int64_t(k) * n + k]);
  if (!(d >= 0.5 * tau[0]) || !isfinite(d)) {  // rounding made G + tau I indefinite at this scale
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 1;
    return;
  }
  const int i = k + 1 + blockIdx.x * (MW / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n) return;
  const T* rk = A + int64_t(k) * n;
  const T f = conj_(rk[i]) * (1.0 / d);
  T* ri = A + int64_t(i) * n;
  for (int j = i + lane; j < n; j += 64) ri[j] -= f * rk[j];
}
// the steps check the pivots of rows 0..n-2; the last one is checked here
template <class T>
__global__ void k_chol_check_last(const T* A, int n, const double* tau, int* flag) {
  if (threadIdx.x != 0 || flag[0]) return;
  const double d = real_(A[int64_t(n - 1) * n + (n - 1)]);
  if (!(d >= 0.5 * tau[0]) || !isfinite(d)) flag[0] = 1;
}
// R[k][j] = A[k][j] / sqrt(A[k][k]) (j >= k), 0 below; nothing when the attempt failed
template <class T>
__global__ __launch_bounds__(MW) void k_chol_mw_finish(const T* A, int n, T* R, int ldr, const int* flag) {
  if (flag[0]) return;
  const int64_t tot = int64_t(n) * n;
  for (int64_t idx = int64_t(blockIdx.x) * MW + threadIdx.x; idx < tot; idx += int64_t(gridDim.x) * MW) {
    const int i = int(idx / n), j = int(idx % n);
    R[int64_t(i) * ldr + j] = j >= i ? A[idx] * (1.0 / sqrt(real_(A[int64_t(i) * n + i]))) : T(0.0);
  }
}

// ---- pivoted Cholesky ------------------------------------------------------------
// A = G (full), R = 0, perm = identity, rank = n, state[0] = 0 (running)
template <class T>
__global__ __launch_bounds__(MW) void k_pchol_mw_copy(const T* G, int ldg, int n, T* A, T* R, int ldr, int* perm,
                                                      int* rank, int* state) {
  const int64_t tot = int64_t(n) * n;
  for (int64_t idx = int64_t(blockIdx.x) * MW + threadIdx.x; idx < tot; idx += int64_t(gridDim.x) * MW) {
    const int i = int(idx / n), j = int(idx % n);
    A[idx] = G[int64_t(i) * ldg + j];
    R[int64_t(i) * ldr + j] = T(0.0);
    if (j == 0) perm[i] = i;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) { rank[0] = n; state[0] = 0; }
}
template <class T>
__global__ __launch_bounds__(1024) void k_pchol_mw_pivot(const T* G, int ldg, T* A, int n, int k, double tol, T* R,
                                                         int ldr, int* perm, int* rank, int* state, int* status) {
  if (state[0]) return;
  __shared__ double sh[32];
  __shared__ int sidx[32];
  __shared__ int piv_s;
  __shared__ double pval_s;
  double dmax = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) dmax = fmax(dmax, real_(G[int64_t(i) * ldg + i]));
  dmax = block_max(dmax, sh);
  if (!(dmax > 0.0) || !isfinite(dmax)) {
    if (threadIdx.x == 0) { rank[0] = 0; state[0] = 1; if (!isfinite(dmax)) atomicOr(status, 16); }
    return;
  }
  double best = -1.0;
  int bi = k;
  for (int i = k + threadIdx.x; i < n; i += 1024) {
    const double d = real_(A[int64_t(i) * n + i]);
    if (d > best) { best = d; bi = i; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o);
    const int oi = __shfl_xor(bi, o);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  if ((threadIdx.x & 63) == 0) { sh[threadIdx.x >> 6] = best; sidx[threadIdx.x >> 6] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = -1.0;
    int ii = k;
    for (int q = 0; q < 16; q++)
      if (sh[q] > b || (sh[q] == b && sidx[q] < ii)) { b = sh[q]; ii = sidx[q]; }
    piv_s = ii;
    pval_s = b;
  }
  __syncthreads();
  const int p = piv_s;
  const double pv = pval_s;
  if (!(pv > tol * dmax)) {
    if (threadIdx.x == 0) { rank[0] = k; state[0] = 1; }
    return;
  }
  if (p != k) {
    for (int j = threadIdx.x; j < n; j += 1024) {
      const T t = A[int64_t(k) * n + j]; A[int64_t(k) * n + j] = A[int64_t(p) * n + j]; A[int64_t(p) * n + j] = t;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 1024) {
      const T t = A[int64_t(i) * n + k]; A[int64_t(i) * n + k] = A[int64_t(i) * n + p]; A[int64_t(i) * n + p] = t;
    }
    for (int i = threadIdx.x; i < k; i += 1024) {
      const T t = R[int64_t(i) * ldr + k]; R[int64_t(i) * ldr + k] = R[int64_t(i) * ldr + p]; R[int64_t(i) * ldr + p] = t;
    }
    if (threadIdx.x == 0) { const int t = perm[k]; perm[k] = perm[p]; perm[p] = t; }
    __syncthreads();
  }
  const double rkk = sqrt(pv);
  for (int j = k + threadIdx.x; j < n; j += 1024)
    R[int64_t(k) * ldr + j] = j == k ? T(rkk) : A[int64_t(k) * n + j] * (1.0 / rkk);
}
// trailing Schur complement: rows i in (k, n), one wave each
template <class T>
__global__ __launch_bounds__(MW) void k_pchol_mw_update(T* A, int n, int k, const T* R, int ldr, const int* state) {
  if (state[0]) return;
  const int i = k + 1 + blockIdx.x * (MW / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n) return;
  const T* rk = R + int64_t(k) * ldr;
  const T f = conj_(rk[i]);
  T* ai = A + int64_t(i) * n;
  for (int j = k + 1 + lane; j < n; j += 64) ai[j] -= f * rk[j];
}

inline unsigned mw_grid(int64_t items, int per) {
  const int64_t g = ceil_div(items, per);
  return unsigned(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace
