// Small dense FP64 kernels (n <= 512) of the POD core: reduction of the
// per-chunk Gram partials, shifted and pivoted Cholesky, triangular inverse,
// small products, one-sided Jacobi SVD, Householder QR and the completion of
// U for numerically null singular values.  Real (double) and complex
// (cdouble) instantiations.  All matrices are row-major unless a kernel says
// column-major.  Single-workgroup kernels use 1024 threads and keep their
// operands in global memory (L2-resident at these sizes).
#include "common.hpp"
#include "small.hpp"

namespace {

constexpr int NT = 1024;

PC_DEV double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
// sum inside aligned groups of `g` lanes (g power of two <= 64)
PC_DEV double group_sum(double v, int g) {
  for (int o = g >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <class T> PC_DEV T group_sum_t(T v, int g);
template <> PC_DEV double group_sum_t<double>(double v, int g) { return group_sum(v, g); }
template <> PC_DEV cdouble group_sum_t<cdouble>(cdouble v, int g) {
  return cdouble(group_sum(v.x, g), group_sum(v.y, g));
}

// block-wide reduction helpers (1024 threads)
PC_DEV double block_sum(double v, double* sh) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; i++) r += sh[i];
  __syncthreads();
  return r;
}
PC_DEV double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = -1e308;
  int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; i++) r = fmax(r, sh[i]);
  __syncthreads();
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Gram partial reduction: G = sum over chunks of the upper blocks, mirrored.
// Real: G[i][j] = sum_c P[c][i][j] (i <= j), G[j][i] = G[i][j].
// Complex fold (input is the Gram of the interleaved real view, size 2n):
//   Re G_ab = P[2a][2b] + P[2a+1][2b+1], Im G_ab = P[2a][2b+1] - P[2a+1][2b].
// Summation over chunks is in a fixed order (deterministic).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_gram_real(const double* __restrict__ P, int chunks, int64_t pstride,
                                                         int NB, int n, double* G, int ldg) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int64_t tot = int64_t(n) * n;
  if (idx >= tot) return;
  int i = int(idx / n), j = int(idx % n);
  if (i > j) return;
  double s = 0.0;
  const double* p = P + int64_t(i) * NB + j;
  for (int c = 0; c < chunks; c++) s += p[c * pstride];
  G[int64_t(i) * ldg + j] = s;
  G[int64_t(j) * ldg + i] = s;
}

__global__ __launch_bounds__(256) void k_reduce_gram_cplx(const double* __restrict__ P, int chunks, int64_t pstride,
                                                         int NB, int n, cdouble* G, int ldg) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  int64_t tot = int64_t(n) * n;
  if (idx >= tot) return;
  int a = int(idx / n), b = int(idx % n);
  if (a > b) return;
  double rr = 0, ii = 0, ri = 0, ir = 0;
  const int64_t o_rr = int64_t(2 * a) * NB + 2 * b, o_ii = int64_t(2 * a + 1) * NB + 2 * b + 1;
  const int64_t o_ri = int64_t(2 * a) * NB + 2 * b + 1, o_ir = int64_t(2 * a + 1) * NB + 2 * b;
  for (int c = 0; c < chunks; c++) {
    const double* p = P + c * pstride;
    rr += p[o_rr]; ii += p[o_ii]; ri += p[o_ri]; ir += p[o_ir];
  }
  cdouble g(rr + ii, ri - ir);
  G[int64_t(a) * ldg + b] = g;
  G[int64_t(b) * ldg + a] = conj_(g);
}

// ---------------------------------------------------------------------------
// Shifted Cholesky: R upper with R^H R = G + tau I, tau = shift_rel * max G_ii.
// Unblocked right-looking in place on a copy (A = R on exit, lower part zero).
// status[0] set to 1 on a non-finite or non-positive pivot.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_chol_shift(const T* G, int ldg, int n, double shift_rel, T* R, int ldr,
                                                   int* status) {
  __shared__ double sh[32];
  __shared__ double piv;
  double dmax = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) dmax = fmax(dmax, real_(G[int64_t(i) * ldg + i]));
  dmax = block_max(dmax, sh);
  double tau = shift_rel * dmax;
  for (int64_t idx = threadIdx.x; idx < int64_t(n) * n; idx += NT) {
    int i = int(idx / n), j = int(idx % n);
    T v = i <= j ? G[int64_t(i) * ldg + j] : T(0.0);
    if (i == j) v = v + T(tau);
    R[int64_t(i) * ldr + j] = v;
  }
  __syncthreads();
  for (int k = 0; k < n; k++) {
    if (threadIdx.x == 0) {
      double d = real_(R[int64_t(k) * ldr + k]);
      if (!(d > 0.0) || !isfinite(d)) { status[0] = 1; d = 1.0; }
      piv = sqrt(d);
    }
    __syncthreads();
    double p = piv;
    T* rk = R + int64_t(k) * ldr;
    for (int j = k + 1 + threadIdx.x; j < n; j += NT) rk[j] = rk[j] * (1.0 / p);
    if (threadIdx.x == 0) rk[k] = T(p);
    __syncthreads();
    int w = n - k - 1;
    int64_t cnt = int64_t(w) * w;
    for (int64_t idx = threadIdx.x; idx < cnt; idx += NT) {
      int i = k + 1 + int(idx / w), j = k + 1 + int(idx % w);
      if (i <= j) R[int64_t(i) * ldr + j] -= conj_(rk[i]) * rk[j];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Inverse of an upper-triangular matrix: W = R^{-1}; column j solved by the
// waves in turn (back substitution, dot products over the solved part).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_trinv_upper(const T* R, int ldr, int n, T* W, int ldw) {
  for (int64_t idx = threadIdx.x; idx < int64_t(n) * n; idx += NT) {
    int i = int(idx / n), j = int(idx % n);
    if (i > j) W[int64_t(i) * ldw + j] = T(0.0);
  }
  int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = wv; j < n; j += NT / 64) {
    // W[:, j]: W[j][j] = 1/R[j][j]; W[i][j] = -(sum_{k=i+1..j} R[i][k] W[k][j]) / R[i][i]
    T rjj = R[int64_t(j) * ldr + j];
    T wjj = inv_(rjj);
    if (lane == 0) W[int64_t(j) * ldw + j] = wjj;
    __threadfence_block();
    for (int i = j - 1; i >= 0; i--) {
      T s = T(0.0);
      for (int k = i + 1 + lane; k <= j; k += 64) s += R[int64_t(i) * ldr + k] * W[int64_t(k) * ldw + j];
      s = group_sum_t(s, 64);
      T rii = R[int64_t(i) * ldr + i];
      T v = -(s * inv_(rii));
      if (lane == 0) W[int64_t(i) * ldw + j] = v;
      __threadfence_block();
    }
  }
}

// ---------------------------------------------------------------------------
// Pivoted Cholesky with rank detection on a working copy A (n x n, upper used):
// P^T G P = R^H R, R upper-trapezoidal with rows >= rank exactly zero (the
// caller pads a unit diagonal later, before inverting).  perm[k] = the
// original index in pivot position k.  Stops when the largest remaining
// Schur diagonal is <= tol * max(diag G).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_pchol(const T* G, int ldg, int n, double tol, T* A, T* R, int ldr, int* perm,
                                              int* rank_out, int* status) {
  __shared__ double sh[32];
  __shared__ int sidx[32];
  __shared__ int piv_s;
  __shared__ double pval_s;
  for (int64_t idx = threadIdx.x; idx < int64_t(n) * n; idx += NT) {
    int i = int(idx / n), j = int(idx % n);
    A[int64_t(i) * n + j] = G[int64_t(i) * ldg + j];
    R[int64_t(i) * ldr + j] = T(0.0);
  }
  for (int i = threadIdx.x; i < n; i += NT) perm[i] = i;
  __syncthreads();
  double dmax = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) dmax = fmax(dmax, real_(A[int64_t(i) * n + i]));
  dmax = block_max(dmax, sh);
  if (!(dmax > 0.0) || !isfinite(dmax)) {
    if (threadIdx.x == 0) { rank_out[0] = 0; if (!isfinite(dmax)) status[0] = 1; }
    return;
  }
  double thr = tol * dmax;
  int rank = n;
  for (int k = 0; k < n; k++) {
    // argmax of the remaining diagonal
    double best = -1.0;
    int bi = k;
    for (int i = k + threadIdx.x; i < n; i += NT) {
      double d = real_(A[int64_t(i) * n + i]);
      if (d > best) { best = d; bi = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      double ob = __shfl_xor(best, o);
      int oi = __shfl_xor(bi, o);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[w] = best; sidx[w] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = -1.0; int ii = k;
      for (int q = 0; q < NT / 64; q++)
        if (sh[q] > b || (sh[q] == b && sidx[q] < ii)) { b = sh[q]; ii = sidx[q]; }
      piv_s = ii; pval_s = b;
    }
    __syncthreads();
    int p = piv_s;
    double pv = pval_s;
    if (!(pv > thr)) { rank = k; break; }
    if (p != k) {
      // symmetric swap of rows/cols k and p in A (full matrix kept), columns of R
      for (int j = threadIdx.x; j < n; j += NT) {
        T t = A[int64_t(k) * n + j]; A[int64_t(k) * n + j] = A[int64_t(p) * n + j]; A[int64_t(p) * n + j] = t;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += NT) {
        T t = A[int64_t(i) * n + k]; A[int64_t(i) * n + k] = A[int64_t(i) * n + p]; A[int64_t(i) * n + p] = t;
      }
      for (int i = threadIdx.x; i < k; i += NT) {
        T t = R[int64_t(i) * ldr + k]; R[int64_t(i) * ldr + k] = R[int64_t(i) * ldr + p]; R[int64_t(i) * ldr + p] = t;
      }
      if (threadIdx.x == 0) { int t = perm[k]; perm[k] = perm[p]; perm[p] = t; }
      __syncthreads();
    }
    double rkk = sqrt(pv);
    for (int j = k + threadIdx.x; j < n; j += NT) {
      T v = j == k ? T(rkk) : A[int64_t(k) * n + j] * (1.0 / rkk);
      R[int64_t(k) * ldr + j] = v;
    }
    __syncthreads();
    int wdt = n - k - 1;
    int64_t cnt = int64_t(wdt) * wdt;
    const T* rk = R + int64_t(k) * ldr;
    for (int64_t idx = threadIdx.x; idx < cnt; idx += NT) {
      int i = k + 1 + int(idx / wdt), j = k + 1 + int(idx % wdt);
      A[int64_t(i) * n + j] -= conj_(rk[i]) * rk[j];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) rank_out[0] = rank;
}

// ---------------------------------------------------------------------------
// Small product C = alpha * op(A) op(B) + beta * C with optional gathers:
//   op(A)[i][k] uses column acol[k] of A (op = none) when acol != NULL,
//   op(B)[k][j] uses row brow[k] of B (op = none) when brow != NULL.
// opA/opB: 0 none, 1 transpose, 2 conjugate transpose.  If ctrans, C is
// stored transposed (C^T, column-major result).  32x32 tiles, 256 threads.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_gemm_small(int opA, int opB, int M, int N, int K, T alpha, const T* A,
                                                    int lda, const int* acol, const T* B, int ldb, const int* brow,
                                                    T beta, T* C, int ldc, int ctrans) {
  __shared__ T As[32][33];
  __shared__ T Bs[32][33];
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty 0..7
  int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  T acc[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int r = ty; r < 32; r += 8) {
      int i = i0 + r, k = k0 + tx;
      T v = T(0.0);
      if (i < M && k < K) {
        if (opA == 0) v = A[int64_t(i) * lda + (acol ? acol[k] : k)];
        else { v = A[int64_t(k) * lda + i]; if (opA == 2) v = conj_(v); }
      }
      As[r][tx] = v;
      int kk = k0 + r, j = j0 + tx;
      T u = T(0.0);
      if (kk < K && j < N) {
        if (opB == 0) u = B[int64_t(brow ? brow[kk] : kk) * ldb + j];
        else { u = B[int64_t(j) * ldb + kk]; if (opB == 2) u = conj_(u); }
      }
      Bs[r][tx] = u;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 32; k++) {
      T b = Bs[k][tx];
      for (int q = 0; q < 4; q++) acc[q] += As[ty + 8 * q][k] * b;
    }
    __syncthreads();
  }
  for (int q = 0; q < 4; q++) {
    int i = i0 + ty + 8 * q, j = j0 + tx;
    if (i < M && j < N) {
      int64_t off = ctrans ? int64_t(j) * ldc + i : int64_t(i) * ldc + j;
      T old = C[off];
      T v = alpha * acc[q];
      if (!(real_(beta) == 0.0 && abs2_(beta) == 0.0)) v += beta * old;
      C[off] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// One-sided Jacobi SVD on the columns of a column-major matrix H (rows x cols,
// ld = ldh) with accumulation of the unitary J (cols x cols, column-major,
// ld = ldj): H J -> columns mutually orthogonal.  Round-robin parallel order;
// a sweep with no rotation ends the iteration.  On exit sig[k] = ||H[:,k]||.
// status[0] |= 2 when max_sweeps were used up.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_jacobi(T* H, int ldh, int rows, int cols, T* J, int ldj, double* sig,
                                               int* status, int max_sweeps, double tol, int init_j) {
  __shared__ int pos[1024];
  __shared__ double sh[32];
  __shared__ int rot_flag;
  int ncp = cols + (cols & 1);          // even number of players (a dummy if odd)
  int npairs = ncp / 2;
  if (init_j) {
    for (int64_t idx = threadIdx.x; idx < int64_t(cols) * cols; idx += NT) {
      int i = int(idx / cols), j = int(idx % cols);
      J[int64_t(j) * ldj + i] = T(i == j ? 1.0 : 0.0);
    }
  }
  for (int i = threadIdx.x; i < ncp; i += NT) pos[i] = i;
  // threads per pair: power of two, <= 64
  int g = 64;
  while (g > 1 && g * npairs > NT) g >>= 1;
  int gid = threadIdx.x / g, gl = threadIdx.x % g;
  int groups = NT / g;
  // A column whose squared norm is below (eps sqrt(rows))^2 ||H||_F^2 is zero to
  // working precision: it takes part in no rotation.  Without this test the
  // null columns of a rank-deficient H keep rotating among themselves, shrink
  // towards underflow and the iteration never stops.
  double fro2 = 0.0;
  for (int64_t idx = threadIdx.x; idx < int64_t(cols) * rows; idx += NT) {
    int k = int(idx / rows), i = int(idx % rows);
    fro2 += abs2_(H[int64_t(k) * ldh + i]);
  }
  fro2 = block_sum(fro2, sh);
  const double negl = 2.220446049250313e-16 * 2.220446049250313e-16 * double(rows) * fro2;
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; sweep++) {
    if (threadIdx.x == 0) rot_flag = 0;
    __syncthreads();
    for (int step = 0; step < ncp - 1; step++) {
      for (int pr = gid; pr < npairs; pr += groups) {
        int p = pos[pr], q = pos[ncp - 1 - pr];
        if (p >= cols || q >= cols) continue;
        if (p > q) { int t = p; p = q; q = t; }
        T* hp = H + int64_t(p) * ldh;
        T* hq = H + int64_t(q) * ldh;
        double a = 0, b = 0;
        T c = T(0.0);
        for (int i = gl; i < rows; i += g) {
          T x = hp[i], y = hq[i];
          a += abs2_(x);
          b += abs2_(y);
          c += dotc_(x, y);
        }
        a = group_sum(a, g);
        b = group_sum(b, g);
        c = group_sum_t(c, g);
        double gm = sqrt(abs2_(c));
        if (!(a > negl) || !(b > negl)) continue;
        if (!(gm > tol * sqrt(a) * sqrt(b))) continue;
        // phase e = c / |c| (by division: 1/gm could overflow); real rotation on (h_p, conj(e) h_q)
        T e = phase_(c, gm);
        double zeta = (b - a) / (2.0 * gm);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        T ec = conj_(e);
        for (int i = gl; i < rows; i += g) {
          T x = hp[i], y = hq[i] * ec;
          hp[i] = x * cs - y * sn;
          hq[i] = x * sn + y * cs;
        }
        T* jp = J + int64_t(p) * ldj;
        T* jq = J + int64_t(q) * ldj;
        for (int i = gl; i < cols; i += g) {
          T x = jp[i], y = jq[i] * ec;
          jp[i] = x * cs - y * sn;
          jq[i] = x * sn + y * cs;
        }
        if (gl == 0) rot_flag = 1;
      }
      __syncthreads();
      // round robin: keep pos[0], rotate the rest by one
      if (threadIdx.x == 0) {
        int last = pos[ncp - 1];
        for (int i = ncp - 1; i > 1; i--) pos[i] = pos[i - 1];
        pos[1] = last;
      }
      __syncthreads();
    }
    int any = rot_flag;
    __syncthreads();
    if (!any) break;
  }
  if (sweep >= max_sweeps && threadIdx.x == 0) atomicOr(status, 2);
  // column norms
  int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = wv; k < cols; k += NT / 64) {
    double s = 0;
    const T* hk = H + int64_t(k) * ldh;
    for (int i = lane; i < rows; i += 64) s += abs2_(hk[i]);
    s = wave_sum(s);
    if (lane == 0) {
      sig[k] = sqrt(s);
      if (!isfinite(s)) atomicOr(status, 8);
    }
  }
}

// ---------------------------------------------------------------------------
// order[k] = index of the k-th largest sig (stable; ties by index)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_sort_desc(const double* sig, int n, int* order) {
  // NaN keys rank as -inf, ties by index: the ranks always form a permutation,
  // so every order[] entry is written with a valid column index.
  for (int i = threadIdx.x; i < n; i += NT) {
    double v = key_desc(sig[i]);
    int rnk = 0;
    for (int j = 0; j < n; j++) {
      double u = key_desc(sig[j]);
      rnk += (u > v) || (u == v && j < i);
    }
    order[rnk] = i;
  }
}

// ---------------------------------------------------------------------------
// Householder QR of a small row-major matrix A (rows x cols, rows <= 512):
// A = Q R; on exit R (cols x cols, upper; rows >= rows zero) and Q (rows x
// min(rows, cols)) explicitly formed.  LAPACK's convention: beta = -sign(alpha)
// * ||x||, so R's diagonal carries the sign the reference's geqrf gives.
// work: rows*cols (A copy) + cols (tau) + rows (v) of T.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NT) void k_house_qr(const T* A, int lda, int rows, int cols, T* Q, int ldq, T* R, int ldr,
                                                 T* work) {
  __shared__ double sh[32];
  __shared__ T alpha_s;
  __shared__ double beta_s;
  T* Wm = work;                                 // rows x cols copy
  T* tau = work + int64_t(rows) * cols;         // cols
  int kmax = rows < cols ? rows : cols;
  for (int64_t idx = threadIdx.x; idx < int64_t(rows) * cols; idx += NT) {
    int i = int(idx / cols), j = int(idx % cols);
    Wm[idx] = A[int64_t(i) * lda + j];
  }
  __syncthreads();
  for (int k = 0; k < kmax; k++) {
    double s = 0;
    for (int i = k + 1 + threadIdx.x; i < rows; i += NT) s += abs2_(Wm[int64_t(i) * cols + k]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
      T al = Wm[int64_t(k) * cols + k];
      double anorm = sqrt(abs2_(al) + s);
      double re = real_(al);
      double beta = -(re >= 0 ? 1.0 : -1.0) * anorm;
      if (s == 0.0 && abs2_(al - T(re)) == 0.0) {
        tau[k] = T(0.0);
        beta_s = re;
        alpha_s = T(1.0);
      } else {
        // tau = (beta - alpha) / beta ; v = x / (alpha - beta), v0 = 1
        tau[k] = (T(beta) - al) * (1.0 / beta);
        T d = al - T(beta);
        alpha_s = inv_(d);  // 1/(alpha - beta)
        beta_s = beta;
      }
    }
    __syncthreads();
    T sc = alpha_s;
    T tk = tau[k];
    for (int i = k + 1 + threadIdx.x; i < rows; i += NT) Wm[int64_t(i) * cols + k] = Wm[int64_t(i) * cols + k] * sc;
    __syncthreads();
    if (threadIdx.x == 0) Wm[int64_t(k) * cols + k] = T(beta_s);
    // apply H^H = I - conj(tau) v v^H to the trailing columns
    int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = k + 1 + wv; j < cols; j += NT / 64) {
      T w = T(0.0);
      for (int i = k + lane; i < rows; i += 64) {
        T v = i == k ? T(1.0) : Wm[int64_t(i) * cols + k];
        w += dotc_(v, Wm[int64_t(i) * cols + j]);
      }
      w = group_sum_t(w, 64);
      T f = conj_(tk) * w;
      for (int i = k + lane; i < rows; i += 64) {
        T v = i == k ? T(1.0) : Wm[int64_t(i) * cols + k];
        Wm[int64_t(i) * cols + j] -= v * f;
      }
    }
    __syncthreads();
  }
  // R
  for (int64_t idx = threadIdx.x; idx < int64_t(cols) * cols; idx += NT) {
    int i = int(idx / cols), j = int(idx % cols);
    R[int64_t(i) * ldr + j] = (i <= j && i < rows) ? Wm[int64_t(i) * cols + j] : T(0.0);
  }
  // Q = H_0 H_1 ... H_{kmax-1} applied to the first kmax columns of I (backward)
  for (int64_t idx = threadIdx.x; idx < int64_t(rows) * kmax; idx += NT) {
    int i = int(idx / kmax), j = int(idx % kmax);
    Q[int64_t(i) * ldq + j] = T(i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int k = kmax - 1; k >= 0; k--) {
    T tk = tau[k];
    int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = k + wv; j < kmax; j += NT / 64) {
      T w = T(0.0);
      for (int i = k + lane; i < rows; i += 64) {
        T v = i == k ? T(1.0) : Wm[int64_t(i) * cols + k];
        w += dotc_(v, Q[int64_t(i) * ldq + j]);
      }
      w = group_sum_t(w, 64);
      T f = tk * w;
      for (int i = k + lane; i < rows; i += 64) {
        T v = i == k ? T(1.0) : Wm[int64_t(i) * cols + k];
        Q[int64_t(i) * ldq + j] -= v * f;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Copy / layout helpers
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_copy2d(const T* A, int lda, int rows, int cols, T* B, int ldb, int trans) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(rows) * cols) return;
  int i = int(idx / cols), j = int(idx % cols);
  T v = A[int64_t(i) * lda + j];
  if (trans) B[int64_t(j) * ldb + i] = v;
  else B[int64_t(i) * ldb + j] = v;
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T* A, int64_t count, T v) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx < count) A[idx] = v;
}

// complex n x n -> real 2n x 2n embedding acting on interleaved rows
__global__ __launch_bounds__(256) void k_embed_cplx(const cdouble* W, int ldw, int n, int nc, double* E, int lde) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * nc) return;
  int a = int(idx / nc), b = int(idx % nc);
  cdouble w = W[int64_t(a) * ldw + b];
  E[int64_t(2 * a) * lde + 2 * b] = w.x;
  E[int64_t(2 * a) * lde + 2 * b + 1] = w.y;
  E[int64_t(2 * a + 1) * lde + 2 * b] = -w.y;
  E[int64_t(2 * a + 1) * lde + 2 * b + 1] = w.x;
}

// ---------------------------------------------------------------------------
// SVD finalisation.  Inputs (column-major, sorted by `order`):
//   Hc (rows x n) = M J after Jacobi, sig, J (n x n col-major).
// Outputs: S (n, real, float or double), VT rows = J columns^H (ldvt),
//   UM (n x n row-major, FP64): UM[i][k] = Hc[i][order[k]] / sig  (zero for
//   null columns and rows >= rows),
//   nnull[0] = number of columns with sig <= null_tol * sig_max (they sit at
//   the end after sorting), ranks >= rank_in also count as null.
// ---------------------------------------------------------------------------
template <class T, class ST, class VT_>
__global__ __launch_bounds__(256) void k_svd_finish(const T* Hc, int ldh, int rows, int n, const double* sig,
                                                    const int* order, const T* J, int ldj, double null_tol,
                                                    const int* rank_in, ST* S, VT_* VT, int64_t ldvt, T* UM, int ldum,
                                                    int* nnull) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double smax = sig[order[0]];
  int rk = rank_in ? rank_in[0] : n;
  if (idx == 0) {
    int c = 0;
    for (int k = 0; k < n; k++) {
      double s = sig[order[k]];
      if (!(s > null_tol * smax) || k >= rk) c++;
    }
    nnull[0] = c;
  }
  if (idx < n) {
    double s = sig[order[idx]];
    bool isnull = !(s > null_tol * smax) || idx >= rk;
    S[idx] = ST(isnull && idx >= rk ? 0.0 : s);
  }
  if (idx >= int64_t(n) * n) return;
  int k = int(idx / n), j = int(idx % n);
  int ok = order[k];
  ok = ok < 0 ? 0 : (ok >= n ? n - 1 : ok);  // order is a permutation; the clamp only guards
  // VT[k][j] = conj(J[j][ok])
  T v = conj_(J[int64_t(ok) * ldj + j]);
  store_vt(VT + int64_t(k) * ldvt + j, v);
  // UM[j][k] (row j of the r-space, column k)
  double s = sig[ok];
  bool isnull = !(s > null_tol * smax) || k >= rk;
  T u = (j < rows && !isnull) ? Hc[int64_t(ok) * ldh + j] * (1.0 / s) : T(0.0);
  UM[int64_t(j) * ldum + k] = u;
}

// explicit instantiation helpers are in small_api.hip
