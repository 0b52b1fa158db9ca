// FP64 MFMA streaming kernels for wide problems (256 < real columns <= 2048):
// the tall-skinny SVD of n columns beyond what one register-resident Gram
// tile holds.  Columns are cut into stripes of W = 128.
//
//   k_wgram_f64 : one (row chunk, stripe pair I <= J) per workgroup:
//                 G[I, J] = Yc_I^T Yc_J over the chunk's rows, 32-row tiles of
//                 both stripes staged in LDS, 64 16x16 output blocks (8 per
//                 wave).  The chunk's partial Gram lands in its NBp x NBp slab
//                 at (I W, J W); only blocks with I <= J are written, the
//                 reductions read the upper triangle only.
//   k_wgemm_f64 : one (row chunk, 128-column output slice) per workgroup:
//                 C[:, c0:c0+128] = Yc W[:, c0:c0+128], the K dimension in
//                 chunks of 128 staged in LDS, W read from L2.
// Yc is the row-centred input (center 1: minus the row mean; 2: also divided
// by the row variance; the FP64 statistics come from a prior pass), so the
// centring costs nothing extra.  Element types: TS the real storage type
// (complex data is read as its interleaved real view).
//
// MFMA lane layouts (v_mfma_f64_16x16x4f64, wave64):
//   A (16x4): lane l holds A[l % 16][l / 16]; B (4x16): B[l / 16][l % 16];
//   D (16x16), reg r: D[4 r + l / 16][l % 16].
#include "common.hpp"
#include "kernels.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int WS = 128;       // stripe / slice width
constexpr int WTR = 32;       // rows per tile
constexpr int WNT = 512;      // threads per workgroup (8 waves)

PC_DEV d4 mfma64(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// centred value of x in row `row` (FP64 statistics)
PC_DEV double centred(double x, int center, const double* mean, const double* var, int64_t row) {
  if (center == 0) return x;
  const double y = x - mean[row];
  return center == 2 ? y / var[row] : y;
}

template <class TS>
__global__ __launch_bounds__(WNT) void k_wgram_f64(const TS* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                                                   const double* __restrict__ mean, const double* __restrict__ var,
                                                   int kstr, int64_t tpc, double* __restrict__ part, int NBp) {
  __shared__ double S[WTR][2 * WS + 2];
  // pair p -> (I, J), I <= J, row-major over the upper triangle of kstr stripes
  int p = blockIdx.y, I = 0;
  while (p >= kstr - I) { p -= kstr - I; I++; }
  const int J = I + p;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, kr = l >> 4, cc = l & 15;
  d4 acc[8];
#pragma unroll
  for (int s = 0; s < 8; s++) acc[s] = d4{0, 0, 0, 0};
  const int64_t ntiles = ceil_div(m, WTR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    __syncthreads();
    for (int idx = threadIdx.x; idx < WTR * 2 * WS; idx += WNT) {
      const int r = idx / (2 * WS), c = idx % (2 * WS);
      const int col = c < WS ? I * WS + c : J * WS + (c - WS);
      const int64_t row = t * WTR + r;
      double v = 0.0;
      if (row < m && col < n) v = centred(double(X[row * ldx + col]), center, mean, var, row);
      S[r][c] = v;
    }
    __syncthreads();
#pragma unroll 2
    for (int kk = 0; kk < WTR / 4; kk++) {
      const double* srow = S[kk * 4 + kr];
#pragma unroll
      for (int s = 0; s < 8; s++) {
        const int b = w * 8 + s, bi = b >> 3, bj = b & 7;
        acc[s] = mfma64(srow[bi * 16 + cc], srow[WS + bj * 16 + cc], acc[s]);
      }
    }
  }
  double* P = part + int64_t(blockIdx.x) * NBp * NBp;
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int b = w * 8 + s, bi = b >> 3, bj = b & 7;
#pragma unroll
    for (int r = 0; r < 4; r++)
      P[int64_t(I * WS + bi * 16 + 4 * r + kr) * NBp + J * WS + bj * 16 + cc] = acc[s][r];
  }
}

template <class TS, class TO>
__global__ __launch_bounds__(WNT) void k_wgemm_f64(const TS* __restrict__ X, int64_t ldx, int64_t m, int K, int center,
                                                   const double* __restrict__ mean, const double* __restrict__ var,
                                                   const double* __restrict__ Wm, int64_t ldw, int N, int64_t tpc,
                                                   TO* __restrict__ C, int64_t ldc) {
  __shared__ double S[WTR][WS + 2];
  const int c0 = blockIdx.y * WS;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, kr = l >> 4, cc = l & 15;
  const int64_t ntiles = ceil_div(m, WTR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    d4 acc[2] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    for (int k0 = 0; k0 < K; k0 += WS) {
      __syncthreads();
      for (int idx = threadIdx.x; idx < WTR * WS; idx += WNT) {
        const int r = idx / WS, c = idx % WS;
        const int64_t row = t * WTR + r;
        const int col = k0 + c;
        double v = 0.0;
        if (row < m && col < K) v = centred(double(X[row * ldx + col]), center, mean, var, row);
        S[r][c] = v;
      }
      __syncthreads();
      // wave w forms both 16-row blocks (slot s = row block) of column
      // block w, so each W element it loads feeds two products
      const int kw = K - k0 < WS ? K - k0 : WS;
      const int col = c0 + w * 16 + cc;
      for (int kk = 0; kk < (kw + 3) / 4; kk++) {
        const int k = k0 + kk * 4 + kr;
        const double* wrow = Wm + int64_t(k < K ? k : 0) * ldw;
        const double bv = (k < K && col < N) ? wrow[col] : 0.0;
#pragma unroll
        for (int s = 0; s < 2; s++) acc[s] = mfma64(S[s * 16 + cc][kk * 4 + kr], bv, acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int R = s, Jb = w;
      const int col = c0 + Jb * 16 + cc;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t row = t * WTR + R * 16 + 4 * r + kr;
        if (row < m && col < N) C[row * ldc + col] = TO(acc[s][r]);
      }
    }
  }
}

// P[chunk] (PBp x QBp) += A^T B over the chunk's rows, A m x p, B m x q: one
// (row chunk, stripe I of A, stripe J of B) per workgroup
template <class TA, class TB>
__global__ __launch_bounds__(WNT) void k_wxgram_f64(const TA* __restrict__ A, int64_t lda, int p,
                                                    const TB* __restrict__ B, int64_t ldb, int q, int64_t m, int kq,
                                                    int64_t tpc, double* __restrict__ part, int PBp, int QBp) {
  __shared__ double S[WTR][2 * WS + 2];
  const int I = blockIdx.y / kq, J = blockIdx.y % kq;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, kr = l >> 4, cc = l & 15;
  d4 acc[8];
#pragma unroll
  for (int s = 0; s < 8; s++) acc[s] = d4{0, 0, 0, 0};
  const int64_t ntiles = ceil_div(m, WTR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    __syncthreads();
    for (int idx = threadIdx.x; idx < WTR * 2 * WS; idx += WNT) {
      const int r = idx / (2 * WS), c = idx % (2 * WS);
      const int64_t row = t * WTR + r;
      double v = 0.0;
      if (c < WS) {
        const int col = I * WS + c;
        if (row < m && col < p) v = double(A[row * lda + col]);
      } else {
        const int col = J * WS + (c - WS);
        if (row < m && col < q) v = double(B[row * ldb + col]);
      }
      S[r][c] = v;
    }
    __syncthreads();
#pragma unroll 2
    for (int kk = 0; kk < WTR / 4; kk++) {
      const double* srow = S[kk * 4 + kr];
#pragma unroll
      for (int s = 0; s < 8; s++) {
        const int b = w * 8 + s, bi = b >> 3, bj = b & 7;
        acc[s] = mfma64(srow[bi * 16 + cc], srow[WS + bj * 16 + cc], acc[s]);
      }
    }
  }
  double* P = part + int64_t(blockIdx.x) * PBp * QBp;
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int b = w * 8 + s, bi = b >> 3, bj = b & 7;
#pragma unroll
    for (int r = 0; r < 4; r++)
      P[int64_t(I * WS + bi * 16 + 4 * r + kr) * QBp + J * WS + bj * 16 + cc] = acc[s][r];
  }
}

}  // namespace

int wide_stripe() { return WS; }
int wide_tile_rows() { return WTR; }

int launch_wgram_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int n, int center,
                     const double* mean, const double* var, int chunks, int64_t tpc, double* part, int NBp) {
  const int kstr = int(ceil_div(n, WS));
  if (kstr * WS > NBp) return -1;
  const dim3 g(unsigned(chunks), unsigned(kstr * (kstr + 1) / 2));
  if (dt == PC_F64)
    hipLaunchKernelGGL((k_wgram_f64<double>), g, dim3(WNT), 0, st, (const double*)X, ldx, m, n, center, mean, var,
                       kstr, tpc, part, NBp);
  else if (dt == PC_F32)
    hipLaunchKernelGGL((k_wgram_f64<float>), g, dim3(WNT), 0, st, (const float*)X, ldx, m, n, center, mean, var,
                       kstr, tpc, part, NBp);
  else
    return -1;
  return 0;
}

// C (m x N, ldc) = Yc (m x K) W (K x N, ldw FP64); C in FP64 (out_f64) or the storage type
int launch_wgemm_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int K, int center,
                     const double* mean, const double* var, const double* W, int64_t ldw, int N, int chunks,
                     int64_t tpc, void* C, int64_t ldc, int out_f64) {
  const dim3 g(unsigned(chunks), unsigned(ceil_div(N, WS)));
  if (dt == PC_F64)
    hipLaunchKernelGGL((k_wgemm_f64<double, double>), g, dim3(WNT), 0, st, (const double*)X, ldx, m, K, center, mean,
                       var, W, ldw, N, tpc, (double*)C, ldc);
  else if (dt == PC_F32 && out_f64)
    hipLaunchKernelGGL((k_wgemm_f64<float, double>), g, dim3(WNT), 0, st, (const float*)X, ldx, m, K, center, mean,
                       var, W, ldw, N, tpc, (double*)C, ldc);
  else if (dt == PC_F32)
    hipLaunchKernelGGL((k_wgemm_f64<float, float>), g, dim3(WNT), 0, st, (const float*)X, ldx, m, K, center, mean,
                       var, W, ldw, N, tpc, (float*)C, ldc);
  else
    return -1;
  return 0;
}

// partial slabs (chunks x PBp x QBp) of A^T B, A m x p and B m x q of one real
// storage type; PBp, QBp multiples of the stripe width
int launch_wxgram_f64(hipStream_t st, pc_dtype dt, const void* A, int64_t lda, int p, const void* B, int64_t ldb,
                      int q, int64_t m, int chunks, int64_t tpc, double* part, int PBp, int QBp) {
  const int kp = int(ceil_div(p, WS)), kq = int(ceil_div(q, WS));
  if (kp * WS > PBp || kq * WS > QBp || int64_t(kp) * kq > 65535) return -1;
  const dim3 g(unsigned(chunks), unsigned(kp * kq));
  if (dt == PC_F64)
    hipLaunchKernelGGL((k_wxgram_f64<double, double>), g, dim3(WNT), 0, st, (const double*)A, lda, p,
                       (const double*)B, ldb, q, m, kq, tpc, part, PBp, QBp);
  else if (dt == PC_F32)
    hipLaunchKernelGGL((k_wxgram_f64<float, float>), g, dim3(WNT), 0, st, (const float*)A, lda, p, (const float*)B,
                       ldb, q, m, kq, tpc, part, PBp, QBp);
  else
    return -1;
  return 0;
}
