// SPOD front end: Welch blocks of the spectral POD (pyLOM/SPOD/wrapper.py:39-76).
// Power-of-two block lengths use a radix-2 FFT in LDS (k_spod_fft), other
// lengths the direct windowed DFT (k_spod_blocks).
// For block b (time samples b*step .. b*step+nDFT-1), frequency f < nf and row i
//   Q[f][i][b] = s_f * sum_t (x_i(b*step + t) - mu_i) w_t exp(-2 pi i f t / nDFT)
// with w the Hamming window of pyLOM/vmmath/fft.py:17-27 (0.54 - 0.46 cos(2 pi t /
// (nDFT - 1))), s_f = winWeight / nDFT / sqrt(nBlks), doubled for 0 < f < nf-1
// (wrapper.py:73-74 and the 1/sqrt(nBlks) of wrapper.py:82).  Q[f] is the
// M x nBlks complex64 matrix of frequency f with contiguous rows, i.e. exactly
// the operand the per-frequency tall-skinny SVD reads.
#include <cmath>
#include <type_traits>
#include <vector>
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int SPR = 8;   // rows per workgroup
constexpr int SPRR = 4;  // rows per thread task
constexpr int TC = 512;  // samples per LDS chunk
constexpr int NDFT_MAX = 4096;

// window and twiddle tables: w_t, cos(2 pi k / nDFT), sin(2 pi k / nDFT), FP64-accurate
__global__ __launch_bounds__(256) void k_spod_tables(int nDFT, float* win, float* twc, float* tws) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nDFT) return;
  win[t] = float(0.54 - 0.46 * cospi(2.0 * t / double(nDFT - 1)));
  twc[t] = float(cospi(2.0 * t / double(nDFT)));
  tws[t] = float(sinpi(2.0 * t / double(nDFT)));
}

// One workgroup: SPR rows x one block.  Samples are staged through LDS in
// chunks of TC; each thread owns (frequency, 4-row group) tasks, processed in
// rounds of 256 so every thread reaches the same barriers.
__global__ __launch_bounds__(256) void k_spod_blocks(const float* __restrict__ X, int64_t ldx, int64_t M,
                                                     const double* __restrict__ mean, const float* __restrict__ win,
                                                     const float* __restrict__ twc, const float* __restrict__ tws,
                                                     int nDFT, int step, int nBlks, int nf,
                                                     const float* __restrict__ fscale, float2* __restrict__ Q) {
  extern __shared__ float sm[];
  float* ys = sm;              // [SPR][TC] windowed samples of the current chunk
  float* tc = ys + SPR * TC;   // [nDFT] cos table
  float* ts = tc + nDFT;       // [nDFT] sin table
  const int64_t row0 = int64_t(blockIdx.x) * SPR;
  const int b = blockIdx.y;
  const int64_t t0 = int64_t(b) * step;
  for (int t = threadIdx.x; t < nDFT; t += 256) {
    tc[t] = twc[t];
    ts[t] = tws[t];
  }
  constexpr int NG = SPR / SPRR;
  const int ntask = nf * NG;
  const int rounds = (ntask + 255) / 256;
  const bool single = nDFT <= TC;
  auto stage = [&](int c0) {
    const int len = min(TC, nDFT - c0);
    for (int idx = threadIdx.x; idx < SPR * TC; idx += 256) {
      const int r = idx / TC, t = idx - r * TC;
      const int64_t row = row0 + r;
      float y = 0.0f;
      if (row < M && t < len) {
        double x = double(X[row * ldx + t0 + c0 + t]);
        if (mean) x -= mean[row];
        y = float(x) * win[c0 + t];
      }
      ys[idx] = y;
    }
  };
  if (single) stage(0);
  for (int rd = 0; rd < rounds; rd++) {
    const int task = rd * 256 + int(threadIdx.x);
    const bool active = task < ntask;
    const int f = active ? task / NG : 0, g = task % NG;
    float re[SPRR], im[SPRR];
#pragma unroll
    for (int q = 0; q < SPRR; q++) re[q] = im[q] = 0.0f;
    int k = 0;  // f t mod nDFT
    for (int c0 = 0; c0 < nDFT; c0 += TC) {
      if (!single) {
        __syncthreads();  // the previous chunk is consumed
        stage(c0);
      }
      __syncthreads();
      const int len = min(TC, nDFT - c0);
      if (active) {
        for (int t = 0; t < len; t++) {
          const float c = tc[k], s = ts[k];
#pragma unroll
          for (int q = 0; q < SPRR; q++) {
            const float y = ys[(g * SPRR + q) * TC + t];
            re[q] += y * c;
            im[q] -= y * s;
          }
          k += f;
          if (k >= nDFT) k -= nDFT;
        }
      }
    }
    if (active) {
      const float sc = fscale[f];
#pragma unroll
      for (int q = 0; q < SPRR; q++) {
        const int64_t row = row0 + g * SPRR + q;
        if (row < M) Q[(int64_t(f) * M + row) * nBlks + b] = make_float2(re[q] * sc, im[q] * sc);
      }
    }
  }
}

// Power-of-two nDFT: radix-2 Stockham FFT in LDS.  A workgroup takes 2 NF
// rows of one block; rows 2p and 2p+1 form one complex sequence
// z = y_a + i y_b, and after the FFT Z the two real spectra are
//   Y_a[f] = (Z[f] + conj Z[N-f]) / 2,   Y_b[f] = (Z[f] - conj Z[N-f]) / 2i.
// Stage with half-size m (m = 1, 2, ..., N/2), butterfly j < N/2, k = j mod m:
//   a = A[j], b = A[j + N/2] w^(k N / 2m);  B[2j - k] = a + b, B[2j - k + m] = a - b.
// Twiddles w^q = exp(-2 pi i q / N) come from the FP64-accurate tables.
// Workgroup x runs on XCD x mod 8 (round-robin dispatch); the blocks of one
// row group are given to consecutive workgroups of the same XCD, so that
// their writes Q[f][row][0..nBlks) (nBlks consecutive complex values) merge
// in that XCD's L2 and the overlapping input samples are read once from HBM.
template <int NF>
__global__ __launch_bounds__(256) void k_spod_fft(const float* __restrict__ X, int64_t ldx, int64_t M,
                                                  const double* __restrict__ mean, const float* __restrict__ win,
                                                  const float* __restrict__ twc, const float* __restrict__ tws,
                                                  int nDFT, int logN, int step, int nBlks, int nf,
                                                  const float* __restrict__ fscale, float2* __restrict__ Q) {
  extern __shared__ float2 fsm[];
  float2* A = fsm;                 // [NF][N]
  float2* B = A + NF * nDFT;       // [NF][N]
  float2* tw = B + NF * nDFT;      // [N / 2]
  const int N = nDFT, H = nDFT / 2;
  const int64_t kx = int64_t(blockIdx.x) >> 3;
  const int64_t rg = (kx / nBlks) * 8 + (blockIdx.x & 7);
  const int b = int(kx % nBlks);
  const int64_t row0 = rg * (2 * NF);
  if (row0 >= M) return;  // padding of the grid to whole XCD rounds (the whole workgroup leaves)
  const int64_t t0 = int64_t(b) * step;
  for (int q = threadIdx.x; q < H; q += 256) tw[q] = make_float2(twc[q], -tws[q]);
  for (int idx = threadIdx.x; idx < NF * N; idx += 256) {
    const int p = idx / N, t = idx - p * N;
    const int64_t ra = row0 + 2 * p, rb = ra + 1;
    float ya = 0.0f, yb = 0.0f;
    if (ra < M) {
      double x = double(X[ra * ldx + t0 + t]);
      if (mean) x -= mean[ra];
      ya = float(x) * win[t];
    }
    if (rb < M) {
      double x = double(X[rb * ldx + t0 + t]);
      if (mean) x -= mean[rb];
      yb = float(x) * win[t];
    }
    A[idx] = make_float2(ya, yb);
  }
  __syncthreads();
  for (int s = 0, m = 1; s < logN; s++, m <<= 1) {
    const int tstride = N >> (s + 1);  // N / 2m
    for (int idx = threadIdx.x; idx < NF * H; idx += 256) {
      const int p = idx / H, j = idx - p * H;
      const int k = j & (m - 1);
      const float2 a = A[p * N + j], bv = A[p * N + j + H], w = tw[k * tstride];
      const float2 wb = make_float2(w.x * bv.x - w.y * bv.y, w.x * bv.y + w.y * bv.x);
      B[p * N + 2 * j - k] = make_float2(a.x + wb.x, a.y + wb.y);
      B[p * N + 2 * j - k + m] = make_float2(a.x - wb.x, a.y - wb.y);
    }
    __syncthreads();
    float2* t = A; A = B; B = t;
  }
  for (int idx = threadIdx.x; idx < NF * nf; idx += 256) {
    const int p = idx / nf, f = idx - p * nf;
    const float2 z = A[p * N + f], zr = A[p * N + ((N - f) & (N - 1))];
    const float sc = fscale[f];
    // Y_a = (z + conj zr) / 2,  Y_b = (z - conj zr) / 2i
    const float2 ya = make_float2(0.5f * (z.x + zr.x), 0.5f * (z.y - zr.y));
    const float2 yb = make_float2(0.5f * (z.y + zr.y), -0.5f * (z.x - zr.x));
    const int64_t ra = row0 + 2 * p, rb = ra + 1;
    if (ra < M) Q[(int64_t(f) * M + ra) * nBlks + b] = make_float2(ya.x * sc, ya.y * sc);
    if (rb < M) Q[(int64_t(f) * M + rb) * nBlks + b] = make_float2(yb.x * sc, yb.y * sc);
  }
}

}  // namespace

extern "C" int pc_spod_blocks(pc_ctx* ctx, const float* X, int64_t ldx, int64_t M, int64_t N, int64_t nDFT,
                              int64_t nolap, int remove_mean, void* Q) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "ctx is NULL");
  if (!X || !Q) return pc_fail(ctx, PC_ERR_INVALID, "pc_spod_blocks: NULL pointer");
  if (M <= 0 || N <= 0) return pc_fail(ctx, PC_ERR_SHAPE, "pc_spod_blocks: empty matrix");
  if (ldx < N) return pc_fail(ctx, PC_ERR_INVALID, "pc_spod_blocks: ldx < N");
  if (nDFT < 2 || nDFT > NDFT_MAX || nDFT > N)
    return pc_fail(ctx, PC_ERR_SHAPE, "pc_spod_blocks: need 2 <= nDFT <= min(N, 4096) (nDFT=%lld)", (long long)nDFT);
  if (nolap < 0 || nolap >= nDFT) return pc_fail(ctx, PC_ERR_SHAPE, "pc_spod_blocks: need 0 <= nolap < nDFT");
  const int64_t step = nDFT - nolap;
  const int64_t nBlks = (N - nolap) / step;
  if (nBlks < 1) return pc_fail(ctx, PC_ERR_SHAPE, "pc_spod_blocks: no complete block");
  const int64_t nf = (nDFT + 1) / 2 + 1;  // ceil(nDFT / 2) + 1 (wrapper.py:67)
  PC_HIP(ctx, hipSetDevice(ctx->device));
  // workspace: mean (M doubles), tables (3 nDFT floats), fscale (nf floats)
  const int64_t bytes = round_up(M * 8, 256) + round_up(3 * nDFT * 4, 256) + round_up(nf * 4, 256);
  PC_CHECK(ctx, pc_ws_reserve(ctx, ctx->ws_skip + bytes));
  WsCursor cur{ctx->ws + ctx->ws_skip, 0, ctx->ws_bytes - ctx->ws_skip};
  double* mean = cur.take<double>(M);
  float* win = cur.take<float>(nDFT);
  float* twc = cur.take<float>(nDFT);
  float* tws = cur.take<float>(nDFT);
  float* fsc = cur.take<float>(nf);
  hipStream_t st = ctx->stream;
  if (remove_mean) launch_row_stats_f32(st, X, ldx, M, N, mean, nullptr);
  hipLaunchKernelGGL(k_spod_tables, dim3(unsigned(ceil_div(nDFT, 256))), dim3(256), 0, st, int(nDFT), win, twc, tws);
  // winWeight = 1 / mean(window) (wrapper.py:55), from the same formula in FP64
  double wsum = 0.0;
  for (int64_t t = 0; t < nDFT; t++) wsum += 0.54 - 0.46 * std::cos(2.0 * M_PI * double(t) / double(nDFT - 1));
  const double base = (double(nDFT) / wsum) / double(nDFT) / std::sqrt(double(nBlks));
  std::vector<float> hs(static_cast<size_t>(nf));
  for (int64_t f = 0; f < nf; f++) hs[size_t(f)] = float(base * ((f > 0 && f < nf - 1) ? 2.0 : 1.0));
  PC_HIP(ctx, hipMemcpyAsync(fsc, hs.data(), size_t(nf) * 4, hipMemcpyHostToDevice, st));
  int logN = 0;
  while ((int64_t(1) << logN) < nDFT) logN++;
  const bool pow2 = (int64_t(1) << logN) == nDFT;
  // LDS of the FFT: 2 NF N + N/2 complex, kept within 64 KB
  auto fft = [&](auto nfc) {
    constexpr int NF = decltype(nfc)::value;
    const size_t lds = (size_t(2 * NF) * size_t(nDFT) + size_t(nDFT / 2)) * sizeof(float2);
    const int64_t groups = round_up(ceil_div(M, 2 * NF), 8);
    dim3 g = dim3(unsigned(groups * nBlks));
    hipLaunchKernelGGL(k_spod_fft<NF>, g, dim3(256), lds, st, X, ldx, M, remove_mean ? mean : nullptr, win, twc,
                       tws, int(nDFT), logN, int(step), int(nBlks), int(nf), fsc, (float2*)Q);
  };
  if (pow2 && nDFT <= 512) fft(std::integral_constant<int, 4>{});
  else if (pow2 && nDFT == 1024) fft(std::integral_constant<int, 2>{});
  else if (pow2 && nDFT == 2048) fft(std::integral_constant<int, 1>{});
  else {
    const size_t lds = (size_t(SPR) * TC + 2 * size_t(nDFT)) * 4;
    dim3 g = dim3(unsigned(ceil_div(M, SPR)), unsigned(nBlks));
    hipLaunchKernelGGL(k_spod_blocks, g, dim3(256), lds, st, X, ldx, M, remove_mean ? mean : nullptr, win, twc, tws,
                       int(nDFT), int(step), int(nBlks), int(nf), fsc, (float2*)Q);
  }
  PC_HIP(ctx, hipGetLastError());
  // the host fscale buffer must outlive the copy
  PC_HIP(ctx, hipStreamSynchronize(st));
  return PC_OK;
}
