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
// Chunk sum of the partial slabs, in place into chunk 0, for the entries
// i <= j < nr of the NB x NB slab.  A workgroup owns 64 consecutive entries
// and 16 chunk lanes; lane k sums chunks k, k+16, ... in four interleaved
// accumulators combined in a fixed order, then lane 0 adds the 16 lane sums in
// order: the summation order depends only on (chunks, NB), so results are
// deterministic.  Entry e is read and written only by its own workgroup.
__global__ __launch_bounds__(1024) void k_reduce_slab(double* P, int chunks, int64_t pstride, int NB, int nr, int64_t zb = 0) {
  P = zoff(P, zb);
  __shared__ double red[16][65];
  const int le = threadIdx.x & 63, k = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + le;
  const int i = e / NB, j = e - i * NB;
  const bool use = e < NB * NB && i <= j && j < nr;
  double s = 0.0;
  if (use) {
    const double* p = P + e;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = k;
    for (; c + 48 < chunks; c += 64) {
      s0 += p[int64_t(c) * pstride];
      s1 += p[int64_t(c + 16) * pstride];
      s2 += p[int64_t(c + 32) * pstride];
      s3 += p[int64_t(c + 48) * pstride];
    }
    for (; c < chunks; c += 16) s0 += p[int64_t(c) * pstride];
    s = (s0 + s1) + (s2 + s3);
  }
  red[k][le] = s;
  __syncthreads();
  if (k == 0 && use) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; q++) t += red[q][le];
    P[e] = t;
  }
}

__global__ __launch_bounds__(256) void k_reduce_gram_real(const double* __restrict__ P, int chunks, int64_t pstride,
                                                         int NB, int n, double* G, int ldg, int64_t zb = 0) {
  P = zoff(P, zb);
  G = zoff(G, zb);
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
                                                         int NB, int n, cdouble* G, int ldg, int64_t zb = 0) {
  P = zoff(P, zb);
  G = zoff(G, zb);
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
  cdouble g(rr + ii, a == b ? 0.0 : ri - ir);  // the diagonal is real
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
                                                   int* status, int64_t zb = 0, const int* only_if = nullptr) {
  if (only_if && !only_if[0]) return;  // fallback of the multi-workgroup factorisation: it succeeded
  G = zoff(G, zb);
  R = zoff(R, zb);
  status = zoff(status, zb);
  __shared__ double sh[32];
  __shared__ double piv;
  __shared__ int restart;
  double dmax = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) dmax = fmax(dmax, real_(G[int64_t(i) * ldg + i]));
  dmax = block_max(dmax, sh);
  if (!isfinite(dmax)) {
    if (threadIdx.x == 0) status[0] = 1;
    dmax = 1.0;
  }
  double tau = shift_rel * dmax;
  if (!(tau > 0.0)) tau = 1.0;
  // For G + tau I >= tau I every Cholesky pivot is >= tau.  A pivot below
  // tau/2 means rounding made G indefinite at this scale (the Gram of
  // mean-removed data is singular): restart with a larger shift.  The factor
  // only preconditions the next pass, whose operand is inverted exactly, so the
  // shift changes the preconditioner, never the result.
  for (int attempt = 0; attempt < 20; attempt++) {
    for (int64_t idx = threadIdx.x; idx < int64_t(n) * n; idx += NT) {
      int i = int(idx / n), j = int(idx % n);
      T v = i <= j ? G[int64_t(i) * ldg + j] : T(0.0);
      if (i == j) v = v + T(tau);
      R[int64_t(i) * ldr + j] = v;
    }
    if (threadIdx.x == 0) restart = 0;
    __syncthreads();
    for (int k = 0; k < n; k++) {
      if (threadIdx.x == 0) {
        double d = real_(R[int64_t(k) * ldr + k]);
        if (!isfinite(d)) { status[0] = 1; d = tau; }
        else if (d < 0.5 * tau) { restart = 1; d = tau; }
        piv = sqrt(d);
      }
      __syncthreads();
      if (restart) break;
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
    int again = restart;
    __syncthreads();
    if (!again) return;
    tau *= 8.0;
  }
  if (threadIdx.x == 0) status[0] = 1;
}

// ---------------------------------------------------------------------------
// Inverse of an upper-triangular matrix: W = R^{-1}; column j solved by the
// waves in turn (back substitution, dot products over the solved part).
// ---------------------------------------------------------------------------
// With a grid of several workgroups the columns are spread over them (each
// column is one wave's independent back substitution: same results).
template <class T>
__global__ __launch_bounds__(NT) void k_trinv_upper(const T* R, int ldr, int n, T* W, int ldw, int64_t zb = 0) {
  R = zoff(R, zb);
  W = zoff(W, zb);
  for (int64_t idx = int64_t(blockIdx.x) * NT + threadIdx.x; idx < int64_t(n) * n; idx += int64_t(gridDim.x) * NT) {
    int i = int(idx / n), j = int(idx % n);
    if (i > j) W[int64_t(i) * ldw + j] = T(0.0);
  }
  int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = blockIdx.x * (NT / 64) + wv; j < n; j += gridDim.x * (NT / 64)) {
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
                                              int* rank_out, int* status, int64_t zb = 0) {
  G = zoff(G, zb);
  A = zoff(A, zb);
  R = zoff(R, zb);
  perm = zoff(perm, zb);
  rank_out = zoff(rank_out, zb);
  status = zoff(status, zb);
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
    if (threadIdx.x == 0) { rank_out[0] = 0; if (!isfinite(dmax)) atomicOr(status, 16); }
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
                                                    T beta, T* C, int ldc, int ctrans, int64_t zb = 0) {
  A = zoff(A, zb);
  acol = zoff(acol, zb);
  B = zoff(B, zb);
  brow = zoff(brow, zb);
  C = zoff(C, zb);
  __shared__ T As[32][33];
  __shared__ T Bs[32][33];
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty 0..7
  int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  T acc[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  // this thread's elements of the next 32-wide k slab of A and B, loaded
  // while the current slab's products run
  T va[4], ub[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = ty + 8 * q;
      int i = i0 + r, k = k0 + tx;
      T v = T(0.0);
      if (i < M && k < K) {
        if (opA == 0) v = A[int64_t(i) * lda + (acol ? acol[k] : k)];
        else { v = A[int64_t(k) * lda + i]; if (opA == 2) v = conj_(v); }
      }
      va[q] = v;
      int kk = k0 + r, j = j0 + tx;
      T u = T(0.0);
      if (kk < K && j < N) {
        if (opB == 0) u = B[int64_t(brow ? brow[kk] : kk) * ldb + j];
        else { u = B[int64_t(j) * ldb + kk]; if (opB == 2) u = conj_(u); }
      }
      ub[q] = u;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      As[ty + 8 * q][tx] = va[q];
      Bs[ty + 8 * q][tx] = ub[q];
    }
    __syncthreads();
    if (k0 + 32 < K) fetch(k0 + 32);
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
// Block one-sided Jacobi (multi-workgroup).  The columns of H (rows x cols,
// column-major) and of the accumulated unitary J are cut into nb2 blocks of b
// columns.  One launch of k_bjac_step processes one round-robin step: each
// workgroup takes a pair of blocks, stages their 2b columns of H and J in LDS,
// orthogonalises them there (one wave per column pair) and writes them back.
// k_bjac_check closes a sweep: when no workgroup rotated, state[0] (done) is
// set and every later launch returns at once.  No host synchronisation and no
// inter-workgroup waiting is involved; the stream orders the launches.
//   state: [0] done, [1 + s] rotated-in-sweep-s flags;  dstate[0] = negligible
//   squared column norm (eps^2 rows ||H||_F^2).
// ---------------------------------------------------------------------------
constexpr int BJ_LDS = 131072;

template <class T>
__global__ __launch_bounds__(NT) void k_bjac_init(const T* H, int ldh, int rows, int cols, T* J, int ldj, int* state,
                                                  double* dstate, int nflags, int64_t zb = 0) {
  H = zoff(H, zb);
  J = zoff(J, zb);
  state = zoff(state, zb);
  dstate = zoff(dstate, zb);
  __shared__ double sh[32];
  for (int i = threadIdx.x; i < nflags; i += NT) state[i] = 0;
  for (int64_t idx = threadIdx.x; idx < int64_t(cols) * cols; idx += NT) {
    int j = int(idx / cols), i = int(idx % cols);  // consecutive threads, consecutive addresses
    J[int64_t(j) * ldj + i] = T(i == j ? 1.0 : 0.0);
  }
  double f = 0.0;
  for (int64_t idx = threadIdx.x; idx < int64_t(cols) * rows; idx += NT) {
    int k = int(idx / rows), i = int(idx % rows);
    f += abs2_(H[int64_t(k) * ldh + i]);
  }
  f = block_sum(f, sh);
  if (threadIdx.x == 0) dstate[0] = 2.220446049250313e-16 * 2.220446049250313e-16 * double(rows) * f;
}

// player sitting at slot k of step s in the circle schedule of N players
PC_DEV int rr_player(int k, int s, int N) { return k == 0 ? 0 : 1 + ((k - 1 + s) % (N - 1)); }

template <class T>
__global__ __launch_bounds__(NT) void k_bjac_step(T* H, int ldh, int rows, int cols, T* J, int ldj, int b, int nb2,
                                                  int step, int sweep, int* state, const double* dstate, double tol,
                                                  int max_inner) {
  __shared__ __align__(16) char smem[BJ_LDS];
  __shared__ int rot_any, rot_sweep;
  if (state[0]) return;
  const int c2 = 2 * b;
  T* Hs = reinterpret_cast<T*>(smem);
  T* Js = Hs + int64_t(c2) * rows;
  int p = blockIdx.x;
  int bi = rr_player(p, step, nb2), bj = rr_player(nb2 - 1 - p, step, nb2);
  double negl = dstate[0];
  // stage the 2b columns (columns beyond `cols` are empty)
  for (int64_t idx = threadIdx.x; idx < int64_t(c2) * rows; idx += NT) {
    int lc = int(idx / rows), i = int(idx % rows);
    int gc = lc < b ? bi * b + lc : bj * b + (lc - b);
    Hs[idx] = gc < cols ? H[int64_t(gc) * ldh + i] : T(0.0);
  }
  for (int64_t idx = threadIdx.x; idx < int64_t(c2) * cols; idx += NT) {
    int lc = int(idx / cols), i = int(idx % cols);
    int gc = lc < b ? bi * b + lc : bj * b + (lc - b);
    Js[idx] = gc < cols ? J[int64_t(gc) * ldj + i] : T(0.0);
  }
  if (threadIdx.x == 0) rot_any = 0;
  __syncthreads();
  int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int last_rot = 0;
  for (int inner = 0; inner < max_inner; inner++) {
    if (threadIdx.x == 0) rot_sweep = 0;
    __syncthreads();
    for (int s = 0; s < c2 - 1; s++) {
      for (int pr = wv; pr < b; pr += NT / 64) {
        int u = rr_player(pr, s, c2), v = rr_player(c2 - 1 - pr, s, c2);
        if (u > v) { int t = u; u = v; v = t; }
        int gu = u < b ? bi * b + u : bj * b + (u - b);
        int gv = v < b ? bi * b + v : bj * b + (v - b);
        if (gu >= cols || gv >= cols) continue;
        T* hu = Hs + int64_t(u) * rows;
        T* hv = Hs + int64_t(v) * rows;
        double a = 0, bb = 0;
        T c = T(0.0);
        for (int i = lane; i < rows; i += 64) {
          T x = hu[i], y = hv[i];
          a += abs2_(x);
          bb += abs2_(y);
          c += dotc_(x, y);
        }
        a = wave_sum(a);
        bb = wave_sum(bb);
        c = group_sum_t(c, 64);
        double gm = sqrt(abs2_(c));
        if (!(a > negl) || !(bb > negl)) continue;
        if (!(gm > tol * sqrt(a) * sqrt(bb))) continue;
        T e = phase_(c, gm);
        double zeta = (bb - a) / (2.0 * gm);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        T ec = conj_(e);
        for (int i = lane; i < rows; i += 64) {
          T x = hu[i], y = hv[i] * ec;
          hu[i] = x * cs - y * sn;
          hv[i] = x * sn + y * cs;
        }
        T* ju = Js + int64_t(u) * cols;
        T* jv = Js + int64_t(v) * cols;
        for (int i = lane; i < cols; i += 64) {
          T x = ju[i], y = jv[i] * ec;
          ju[i] = x * cs - y * sn;
          jv[i] = x * sn + y * cs;
        }
        if (lane == 0) { rot_sweep = 1; rot_any = 1; }
      }
      __syncthreads();
    }
    int any = rot_sweep;
    __syncthreads();
    last_rot = any;
    if (!any) break;
  }
  // write back
  for (int64_t idx = threadIdx.x; idx < int64_t(c2) * rows; idx += NT) {
    int lc = int(idx / rows), i = int(idx % rows);
    int gc = lc < b ? bi * b + lc : bj * b + (lc - b);
    if (gc < cols) H[int64_t(gc) * ldh + i] = Hs[idx];
  }
  for (int64_t idx = threadIdx.x; idx < int64_t(c2) * cols; idx += NT) {
    int lc = int(idx / cols), i = int(idx % cols);
    int gc = lc < b ? bi * b + lc : bj * b + (lc - b);
    if (gc < cols) J[int64_t(gc) * ldj + i] = Js[idx];
  }
  // a single block pair reports whether its last inner sweep still rotated;
  // several pairs report any rotation (the sweep is then not yet converged)
  int flag = nb2 == 2 ? last_rot : rot_any;
  if (threadIdx.x == 0 && flag) atomicOr(&state[1 + sweep], 1);
}

// close sweep `sweep`: no workgroup reported a rotation -> done
__global__ void k_bjac_check(int* state, int sweep, int64_t zb = 0) {
  state = zoff(state, zb);
  if (threadIdx.x == 0 && !state[0] && state[1 + sweep] == 0) state[0] = 1;
}

// ---------------------------------------------------------------------------
// Block step of the real one-sided Jacobi driven by the Gram of the block
// pair (rows <= 256, blocks of GB = 16 columns).  Per launch and workgroup:
//   1. stage the pair's 32 columns of H in LDS;
//   2. G = Hs^T Hs with v_mfma_f64_16x16x4f64 (split over the rows, partial
//      sums added in a fixed order);
//   3. if no pair (u, v) of G has |G_uv| > tol sqrt(G_uu G_vv) the columns are
//      already orthogonal to the test's accuracy: nothing is written and the
//      sweep flag stays clear;
//   4. otherwise cyclic two-sided Jacobi sweeps on G (16 disjoint rotations per
//      round) accumulate the orthogonal Q, and H[:, pair] <- Hs Q and
//      J[:, pair] <- J[:, pair] Q are formed with MFMA.
// The rotation of columns (u, v) is the one-sided Jacobi rotation of the
// columns themselves (same angle formula); the test in 3. always uses a Gram
// freshly computed from the current columns, so convergence is decided by
// the one-sided criterion.
// ---------------------------------------------------------------------------
constexpr int GB = 16;
constexpr int GC = 2 * GB;
constexpr int GLD = GC + 1;
constexpr int GROWS = 256;
// dstate layout of the Gram-driven steps: [0] negligible squared norm, then
// from kBjacQOff the nb2 / 2 pair rotations (GC x GC each) and their flags
constexpr int kBjacQOff = 4;
constexpr int kBjacDstate = kBjacQOff + (GROWS / GB / 2) * GC * GC + GROWS / GB / 2;
// dstate length for an n-column problem (at least kBjacDstate)
inline int64_t bjac_dstate_len(int n) {
  int64_t nb2 = (int64_t(n) + GB - 1) / GB;
  nb2 += nb2 & 1;
  const int64_t len = kBjacQOff + (nb2 / 2) * GC * GC + nb2 / 2;
  return len > kBjacDstate ? len : kBjacDstate;
}

typedef double d4s __attribute__((ext_vector_type(4)));
PC_DEV d4s mfma64s(double a, double b, d4s c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// FUSED: the step also forms H[:, pair] Q and J[:, pair] Q itself, on its one
// workgroup (one launch per step, for solves that share the GPU with other
// streams; rows <= GROWS); otherwise it stores Q for k_bjac_apply.  Both
// form the same products in the same k order, so H and J do not depend on
// the choice.  Unfused, the pair's Gram is summed over row chunks of GROWS
// staged in turn, so any number of rows is taken (one chunk up to GROWS).
template <bool FUSED>
__global__ __launch_bounds__(NT) void k_bjac_gstep(double* H, int ldh, int rows, int cols, double* J, int ldj,
                                                   int nb2, int step, int sweep, int* state, double* dstate,
                                                   double tol, int max_inner) {
  __shared__ double Hs[GC * (GROWS + 1)];
  __shared__ double G[GC * GLD], G2[GC * GLD];
  __shared__ double Q[GC * GLD], Q2[GC * GLD];
  __shared__ double Pp[3 * 4 * 256];
  __shared__ int need;
  const int RP = (rows + 15) & ~15;
  const int ld = RP + 1;
  // the launch holds 8 workgroups per pair and only every eighth works:
  // workgroup x runs on XCD x mod 8, so the working ones share one L2
  if (blockIdx.x & 7) return;
  const int p = blockIdx.x >> 3;
  const int bi = rr_player(p, step, nb2), bj = rr_player(nb2 - 1 - p, step, nb2);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
  auto gcol = [&](int lc) { return lc < GB ? bi * GB + lc : bj * GB + (lc - GB); };
  // Stage row chunk c (rows c GROWS ... of the pair's columns) into Hs with
  // leading dimension RPc + 1.  Every load is issued before the first LDS
  // store, and those of chunk 0 together with the done flag (a converged
  // iteration discards them): false when the iteration has converged.
  const int nch = FUSED ? 1 : (rows + GROWS - 1) / GROWS;
  auto chunk_rows = [&](int c) { const int rc = rows - c * GROWS; return rc < GROWS ? rc : GROWS; };
  auto stage = [&](int c) {
    constexpr int KL = GC * GROWS / NT;
    const int rc = chunk_rows(c), RPc = (rc + 15) & ~15, ldc = RPc + 1;
    const int64_t r0 = int64_t(c) * GROWS;
    double hv[KL];
#pragma unroll
    for (int k = 0; k < KL; k++) {
      const int idx = threadIdx.x + k * NT, lc = idx / RPc, i = idx - lc * RPc, gc = gcol(lc);
      hv[k] = (idx < GC * RPc && gc < cols && i < rc) ? H[int64_t(gc) * ldh + r0 + i] : 0.0;
    }
    if (c == 0 && state[0]) return false;
#pragma unroll
    for (int k = 0; k < KL; k++) {
      const int idx = threadIdx.x + k * NT, lc = idx / RPc, i = idx - lc * RPc;
      if (idx < GC * RPc) Hs[lc * ldc + i] = hv[k];
    }
    return true;
  };
  if (!stage(0)) return;
  const double negl = dstate[0];
  if (threadIdx.x == 0) need = 0;
  __syncthreads();
  // G = Hs^T Hs, symmetric: waves w < 12 form the upper 16 x 16 tiles
  // t = w >> 2 ((0,0), (0,1), (1,1)) over row quarter q = w & 3 of each
  // chunk, and tile (0,1) is mirrored into (1,0)
  {
    const int t = w >> 2, q = w & 3, bI = t == 2 ? 1 : 0, bJ = t == 0 ? 0 : 1;
    d4s acc = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < nch; c++) {
      if (c > 0) {
        __syncthreads();  // the previous chunk is read
        stage(c);
        __syncthreads();
      }
      const int RPc = (chunk_rows(c) + 15) & ~15, ldc = RPc + 1;
      const int kq = RPc / 4;
      const double* ca = Hs + (16 * bI + lr) * ldc;
      const double* cb = Hs + (16 * bJ + lr) * ldc;
      if (t < 3) {
        // the quarter's operands (kq / 4 <= GROWS / 16 k-steps) are read from
        // LDS before the products, so the products do not wait on each read
        constexpr int KS = GROWS / 16;
        double xa[KS], xb[KS];
#pragma unroll
        for (int k = 0; k < KS; k++) {
          const int k0 = q * kq + 4 * k;
          xa[k] = 4 * k < kq ? ca[k0 + lk] : 0.0;
          xb[k] = 4 * k < kq ? cb[k0 + lk] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < KS; k++)
          if (4 * k < kq) acc = mfma64s(xa[k], xb[k], acc);
      }
    }
    if (t < 3 && q > 0)
      for (int r = 0; r < 4; r++) Pp[((q - 1) * 3 + t) * 256 + r * 64 + l] = acc[r];
    __syncthreads();
    if (t < 3 && q == 0) {
      for (int q2 = 0; q2 < 3; q2++)
        for (int r = 0; r < 4; r++) acc[r] += Pp[(q2 * 3 + t) * 256 + r * 64 + l];
      for (int r = 0; r < 4; r++) {
        G[(16 * bI + 4 * r + lk) * GLD + 16 * bJ + lr] = acc[r];
        if (t == 1) G[(16 + lr) * GLD + 4 * r + lk] = acc[r];
      }
    }
  }
  __syncthreads();
  // The first step of a sweep rotates every pair of its 2 GB columns (GC - 1
  // rounds); later steps only the GB^2 pairs across the two blocks (GB
  // rounds): every block meets its own columns in the first step, so a sweep
  // still rotates each pair of columns once, in n - 1 rounds in all.
  const bool full = step == 0;
  for (int idx = threadIdx.x; idx < GC * GC; idx += NT) {
    int u = idx / GC, v = idx - u * GC;
    Q[u * GLD + v] = u == v ? 1.0 : 0.0;
    if (u < v && (full || (u < GB && v >= GB))) {
      double a = G[u * GLD + u], bb = G[v * GLD + v], c = G[u * GLD + v];
      if (a > negl && bb > negl && fabs(c) > tol * sqrt(a) * sqrt(bb)) need = 1;
    }
  }
  __syncthreads();
  double* qout = dstate + kBjacQOff + int64_t(p) * GC * GC;
  double* qflag = dstate + kBjacQOff + int64_t(nb2 / 2) * GC * GC + p;
  if (!need) {
    if (!FUSED && threadIdx.x == 0) *qflag = 0.0;  // k_bjac_apply leaves this pair as it is
    return;
  }
  if (threadIdx.x == 0) atomicOr(&state[1 + sweep], 1);
  // k-th pair (u < v) of round s
  auto pair_of = [&](int k, int s, int& u, int& v) {
    if (full) {
      u = rr_player(k, s, GC);
      v = rr_player(GC - 1 - k, s, GC);
      if (u > v) { int t = u; u = v; v = t; }
    } else {
      u = k;
      v = GB + (k + s) % GB;
    }
  };
  // Each round applies GB disjoint rotations as G <- R^T G R and Q <- Q R:
  // GB threads form the rotations; thread (bp, bq) of the first GB * GB then
  // updates the 2 x 2 blocks of G and Q at rows {u_bp, v_bp} and columns
  // {u_bq, v_bq} from the old buffers into the other ones.  The rotation of
  // pair (u, v) zeroes G_uv: column u of R is cs e_u - sn e_v, column v is
  // sn e_u + cs e_v, tan(2 theta) = 2 G_uv / (G_vv - G_uu), |theta| <= pi/4.
  // Square roots and reciprocals come from the hardware estimates refined by
  // Newton steps (one for the angle, two for cs, which fixes cs^2 + sn^2 = 1
  // to rounding); the angle itself only steers convergence.
  static_assert(GB * GB <= NT, "one thread per 2 x 2 block of the pair's Gram");
  __shared__ double rc[GB], rs[GB];
  const int bp = threadIdx.x / GB, bq = threadIdx.x % GB;
  const int rounds = full ? GC - 1 : GB;
  for (int inner = 0; inner < max_inner; inner++) {
    int rotated = 0;
    for (int s = 0; s < rounds; s++) {
      const double* Gc = (s & 1) ? G2 : G;
      double* Gn = (s & 1) ? G : G2;
      const double* Qc = (s & 1) ? Q2 : Q;
      double* Qn = (s & 1) ? Q : Q2;
      if (threadIdx.x < GB) {
        int u, v;
        pair_of(threadIdx.x, s, u, v);
        const double a = Gc[u * GLD + u], bb = Gc[v * GLD + v], c = Gc[u * GLD + v];
        const double d = bb - a, e = 2.0 * fabs(c);
        const double mde = fmax(fabs(d), e);
        const int k = mde > 0.0 ? ilogb(mde) : 0;
        const double ds = ldexp(d, -k), es = ldexp(e, -k);   // max(|ds|, es) in [1, 2)
        const double x = ds * ds + es * es;
        double y = __builtin_amdgcn_rsq(x);
        y = y * (1.5 - 0.5 * x * y * y);
        const double den = fabs(ds) + x * y;                  // |d| + hypot(d, e), scaled
        double r = __builtin_amdgcn_rcp(den);
        r = r * (2.0 - den * r);
        const double t = es * r;                              // tan(theta) in [0, 1]
        const double z = 1.0 + t * t;
        double cs = __builtin_amdgcn_rsq(z);
        cs = cs * (1.5 - 0.5 * z * cs * cs);
        cs = cs * (1.5 - 0.5 * z * cs * cs);
        double sn = cs * t;
        if ((d < 0) != (c < 0)) sn = -sn;
        const bool act = a > negl && bb > negl && fabs(c) > tol * sqrt(a) * sqrt(bb);
        rc[threadIdx.x] = act ? cs : 1.0;
        rs[threadIdx.x] = act ? sn : 0.0;
        rotated |= act;
      }
      __syncthreads();
      if (threadIdx.x < GB * GB) {
        int up, vp, uq, vq;
        pair_of(bp, s, up, vp);
        pair_of(bq, s, uq, vq);
        const double cp = rc[bp], sp = rs[bp], cq = rc[bq], sq = rs[bq];
        const double g_uu = Gc[up * GLD + uq], g_uv = Gc[up * GLD + vq];
        const double g_vu = Gc[vp * GLD + uq], g_vv = Gc[vp * GLD + vq];
        const double q_uu = Qc[up * GLD + uq], q_uv = Qc[up * GLD + vq];
        const double q_vu = Qc[vp * GLD + uq], q_vv = Qc[vp * GLD + vq];
        // columns by pair bq, then rows by pair bp
        const double x_uu = cq * g_uu - sq * g_uv, x_uv = sq * g_uu + cq * g_uv;
        const double x_vu = cq * g_vu - sq * g_vv, x_vv = sq * g_vu + cq * g_vv;
        Gn[up * GLD + uq] = cp * x_uu - sp * x_vu;
        Gn[up * GLD + vq] = cp * x_uv - sp * x_vv;
        Gn[vp * GLD + uq] = sp * x_uu + cp * x_vu;
        Gn[vp * GLD + vq] = sp * x_uv + cp * x_vv;
        Qn[up * GLD + uq] = cq * q_uu - sq * q_uv;
        Qn[up * GLD + vq] = sq * q_uu + cq * q_uv;
        Qn[vp * GLD + uq] = cq * q_vu - sq * q_vv;
        Qn[vp * GLD + vq] = sq * q_vu + cq * q_vv;
      }
      __syncthreads();
    }
    // an odd number of rounds leaves the result in the second buffers
    if (rounds & 1) {
      for (int idx = threadIdx.x; idx < GC * GC; idx += NT) {
        int u = idx / GC, v = idx - u * GC;
        G[u * GLD + v] = G2[u * GLD + v];
        Q[u * GLD + v] = Q2[u * GLD + v];
      }
    }
    if (!__syncthreads_or(rotated)) break;
  }
  if constexpr (!FUSED) {
    // the pair's rotation Q (row-major) for k_bjac_apply, which forms
    // H[:, pair] Q and J[:, pair] Q over many workgroups
    for (int idx = threadIdx.x; idx < GC * GC; idx += NT) qout[idx] = Q[(idx / GC) * GLD + idx % GC];
    if (threadIdx.x == 0) *qflag = 1.0;
    return;
  }
  // H[:, pair] = Hs Q  (row blocks of 16 rows, one wave each)
  for (int rb = w; rb < RP / 16; rb += NT / 64) {
    for (int cB = 0; cB < 2; cB++) {
      d4s acc = {0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < GC; k0 += 4)
        acc = mfma64s(Hs[(k0 + lk) * ld + 16 * rb + lr], Q[(k0 + lk) * GLD + 16 * cB + lr], acc);
      const int gc = gcol(16 * cB + lr);
      for (int r = 0; r < 4; r++) {
        const int row = 16 * rb + 4 * r + lk;
        if (row < rows && gc < cols) H[int64_t(gc) * ldh + row] = acc[r];
      }
    }
  }
  // J[:, pair] = J[:, pair] Q  (J is cols x cols, column-major)
  const int CP = (cols + 15) & ~15;
  for (int rb = w; rb < CP / 16; rb += NT / 64) {
    double av[GC / 4];
    const int row = 16 * rb + lr;
    for (int kk = 0; kk < GC / 4; kk++) {
      const int gc = gcol(4 * kk + lk);
      av[kk] = (gc < cols && row < cols) ? J[int64_t(gc) * ldj + row] : 0.0;
    }
    d4s acc[2];
    for (int cB = 0; cB < 2; cB++) {
      acc[cB] = d4s{0.0, 0.0, 0.0, 0.0};
      for (int kk = 0; kk < GC / 4; kk++) acc[cB] = mfma64s(av[kk], Q[(4 * kk + lk) * GLD + 16 * cB + lr], acc[cB]);
    }
    for (int cB = 0; cB < 2; cB++) {
      const int gc = gcol(16 * cB + lr);
      for (int r = 0; r < 4; r++) {
        const int orow = 16 * rb + 4 * r + lk;
        if (orow < cols && gc < cols) J[int64_t(gc) * ldj + orow] = acc[cB][r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Second half of a block step of k_bjac_gstep: H[:, pair] <- H[:, pair] Q and
// J[:, pair] <- J[:, pair] Q with the pair's rotation Q from dstate.  One
// workgroup per (pair, 32-row slice), so the FP64 MFMA work of the step is
// spread over 8 (nb2 / 2) workgroups per pair instead of one; every slice
// reads and writes only its own rows, and the products are the ones the
// single-workgroup version formed (same k order).  nsl slices per pair
// cover max(rows, cols) (kBjacSlices up to GROWS).
// ---------------------------------------------------------------------------
constexpr int kBjacSlices = GROWS / 32;
__global__ __launch_bounds__(256) void k_bjac_apply(double* H, int ldh, int rows, int cols, double* J, int ldj,
                                                    int nb2, int step, int sweep, int* state, const double* dstate,
                                                    int nsl) {
  __shared__ double Qs[GC * GLD];
  const int p = blockIdx.x / nsl, sl = blockIdx.x % nsl;
  const int bi = rr_player(p, step, nb2), bj = rr_player(nb2 - 1 - p, step, nb2);
  auto gcol = [&](int lc) { return lc < GB ? bi * GB + lc : bj * GB + (lc - GB); };
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
  // wave w: matrix w >> 1 (0 H, 1 J), 16-row block w & 1 of the slice, both
  // 16-column blocks of the pair.  Every global read (flags, Q, the slice)
  // is issued before the first test, so the launch waits on memory once.
  double* M = (w >> 1) ? J : H;
  const int ld = (w >> 1) ? ldj : ldh, nr = (w >> 1) ? cols : rows;
  const int row = 32 * sl + 16 * (w & 1);
  const double* qin = dstate + kBjacQOff + int64_t(p) * GC * GC;
  double av[GC / 4], qv[GC * GC / 256];
#pragma unroll
  for (int kk = 0; kk < GC / 4; kk++) {
    const int gc = gcol(4 * kk + lk);
    av[kk] = (gc < cols && row + lr < nr) ? M[int64_t(gc) * ld + row + lr] : 0.0;
  }
#pragma unroll
  for (int t = 0; t < GC * GC / 256; t++) qv[t] = qin[threadIdx.x + 256 * t];
  const int done = state[0];
  const double rot = dstate[kBjacQOff + int64_t(nb2 / 2) * GC * GC + p];
  // the last step of a sweep also closes it (k_bjac_check's test): every
  // step's k_bjac_gstep has set its rotation flag by now
  if (step == nb2 - 2 && blockIdx.x == 0 && threadIdx.x == 0 && !done && state[1 + sweep] == 0) state[0] = 1;
  if (done || rot == 0.0) return;
#pragma unroll
  for (int t = 0; t < GC * GC / 256; t++) {
    const int idx = threadIdx.x + 256 * t;
    Qs[(idx / GC) * GLD + idx % GC] = qv[t];
  }
  __syncthreads();
  double q0[GC / 4], q1[GC / 4];
#pragma unroll
  for (int kk = 0; kk < GC / 4; kk++) {
    q0[kk] = Qs[(4 * kk + lk) * GLD + lr];
    q1[kk] = Qs[(4 * kk + lk) * GLD + 16 + lr];
  }
  d4s acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
  for (int kk = 0; kk < GC / 4; kk++) {
    acc[0] = mfma64s(av[kk], q0[kk], acc[0]);
    acc[1] = mfma64s(av[kk], q1[kk], acc[1]);
  }
#pragma unroll
  for (int cB = 0; cB < 2; cB++) {
    const int gc = gcol(16 * cB + lr);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int orow = row + 4 * r + lk;
      if (orow < nr && gc < cols) M[int64_t(gc) * ld + orow] = acc[cB][r];
    }
  }
}

// One-sided Jacobi of a small problem held by a single block pair (2b <= 32
// columns, rows <= 64), on one wavefront: 4 lanes per column pair, every
// pair of a round in parallel, 4-lane shuffle reductions, and wave-level
// ordering between rounds instead of workgroup barriers.  Inner sweeps run
// until one rotates nothing (at most max_inner); state[1 + sweep] is set when
// the last inner sweep still rotated, as k_bjac_step does for one pair.
template <class T>
__global__ __launch_bounds__(64) void k_bjac_small(T* H, int ldh, int rows, int cols, T* J, int ldj, int b, int sweep,
                                                   int* state, const double* dstate, double tol, int max_inner, int64_t zb = 0) {
  H = zoff(H, zb);
  J = zoff(J, zb);
  state = zoff(state, zb);
  dstate = zoff(dstate, zb);
  __shared__ T Hs[32 * 64];
  __shared__ T Js[32 * 32];
  if (state[0]) return;
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  const int c2 = 2 * b, l = threadIdx.x;
  const double negl = dstate[0];
  for (int idx = l; idx < c2 * rows; idx += 64) {
    int lc = idx / rows, i = idx - lc * rows;
    Hs[idx] = lc < cols ? H[int64_t(lc) * ldh + i] : T(0.0);
  }
  for (int idx = l; idx < c2 * cols; idx += 64) {
    int lc = idx / cols, i = idx - lc * cols;
    Js[idx] = lc < cols ? J[int64_t(lc) * ldj + i] : T(0.0);
  }
  wave_sync();
  const int pr = l >> 2, sub = l & 3;
  int last_rot = 0;
  for (int inner = 0; inner < max_inner; inner++) {
    unsigned long long rot = 0;
    for (int s = 0; s < c2 - 1; s++) {
      int u = 0, v = 0;
      if (pr < b) {
        u = rr_player(pr, s, c2);
        v = rr_player(c2 - 1 - pr, s, c2);
        if (u > v) { int t = u; u = v; v = t; }
      }
      const bool valid = pr < b && v < cols;  // the same for the 4 lanes of a pair
      T* hu = Hs + int64_t(u) * rows;
      T* hv = Hs + int64_t(v) * rows;
      double a = 0.0, bb = 0.0;
      T c = T(0.0);
      if (valid) {
        for (int i = sub; i < rows; i += 4) {
          T x = hu[i], y = hv[i];
          a += abs2_(x);
          bb += abs2_(y);
          c += dotc_(x, y);
        }
      }
      a = group_sum(a, 4);
      bb = group_sum(bb, 4);
      c = group_sum_t(c, 4);
      int act = 0;
      if (valid) {
        double gm = sqrt(abs2_(c));
        if (a > negl && bb > negl && gm > tol * sqrt(a) * sqrt(bb)) {
          T e = phase_(c, gm);
          double zeta = (bb - a) / (2.0 * gm);
          double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
          T ec = conj_(e);
          for (int i = sub; i < rows; i += 4) {
            T x = hu[i], y = hv[i] * ec;
            hu[i] = x * cs - y * sn;
            hv[i] = x * sn + y * cs;
          }
          T* ju = Js + int64_t(u) * cols;
          T* jv = Js + int64_t(v) * cols;
          for (int i = sub; i < cols; i += 4) {
            T x = ju[i], y = jv[i] * ec;
            ju[i] = x * cs - y * sn;
            jv[i] = x * sn + y * cs;
          }
          act = 1;
        }
      }
      rot |= __ballot(act);
      wave_sync();
    }
    last_rot = rot != 0;
    if (!rot) break;
  }
  for (int idx = l; idx < c2 * rows; idx += 64) {
    int lc = idx / rows, i = idx - lc * rows;
    if (lc < cols) H[int64_t(lc) * ldh + i] = Hs[idx];
  }
  for (int idx = l; idx < c2 * cols; idx += 64) {
    int lc = idx / cols, i = idx - lc * cols;
    if (lc < cols) J[int64_t(lc) * ldj + i] = Js[idx];
  }
  if (l == 0 && last_rot) atomicOr(&state[1 + sweep], 1);
}

// rows of the orthogonal preconditioner: V[perm[k]][:] = Q[k][:]
template <class T>
__global__ __launch_bounds__(256) void k_permute_rows(const T* Qm, int ldq, int n, const int* perm, T* V, int ldv) {
  int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * n) return;
  int k = int(idx / n), j = int(idx % n);
  int pk = perm[k];
  pk = pk < 0 ? 0 : (pk >= n ? n - 1 : pk);  // perm is a permutation; the clamp only guards
  V[int64_t(pk) * ldv + j] = Qm[int64_t(k) * ldq + j];
}

// column norms, non-finite and non-convergence flags
template <class T>
__global__ __launch_bounds__(NT) void k_bjac_norms(const T* H, int ldh, int rows, int cols, double* sig,
                                                   const int* state, int* status, int64_t zb = 0) {
  H = zoff(H, zb);
  sig = zoff(sig, zb);
  state = zoff(state, zb);
  status = zoff(status, zb);
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
  if (threadIdx.x == 0 && !state[0]) atomicOr(status, 2);
}

// block width of the block Jacobi for an n x n problem with element size esz:
// the 2b staged columns of H and J must fit BJ_LDS; one pair when they all fit
inline int bjac_width(int n, int esz) {
  int b = (n + 1) / 2;
  while (int64_t(2) * b * (2 * n) * esz > BJ_LDS) b = (b + 1) / 2;
  return b < 1 ? 1 : b;
}

// ---------------------------------------------------------------------------
// order[k] = index of the k-th largest sig (stable; ties by index)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_sort_desc(const double* sig, int n, int* order, int64_t zb = 0) {
  sig = zoff(sig, zb);
  order = zoff(order, zb);
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
  if (trans == 2) v = conj_(v);  // 2: conjugate transpose
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
                                                    int* nnull, int64_t zb = 0, int64_t zs = 0, int64_t zv = 0) {
  Hc = zoff(Hc, zb);
  sig = zoff(sig, zb);
  order = zoff(order, zb);
  J = zoff(J, zb);
  rank_in = zoff(rank_in, zb);
  UM = zoff(UM, zb);
  nnull = zoff(nnull, zb);
  S = zoff(S, zs);
  VT = zoff(VT, zv);
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
