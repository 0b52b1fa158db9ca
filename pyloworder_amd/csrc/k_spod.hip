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
#include <algorithm>
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

// Power-of-two nDFT: radix-2 Stockham FFT in LDS.  A workgroup owns one row
// pair and runs through the blocks in chunks of BC; rows a, b form one
// complex sequence z = y_a + i y_b, and after the FFT Z the two real spectra
// are
//   Y_a[f] = (Z[f] + conj Z[N-f]) / 2,   Y_b[f] = (Z[f] - conj Z[N-f]) / 2i.
// Stage with half-size m (m = 1, 2, ..., N/2), butterfly j < N/2, k = j mod m:
//   a = A[j], b = A[j + N/2] w^(k N / 2m);  B[2j - k] = a + b, B[2j - k + m] = a - b.
// Twiddles w^q = exp(-2 pi i q / N) come from the FP64-accurate tables.  The
// spectra of a chunk are staged in LDS as [f][row][block], so that each
// frequency's two rows leave as one contiguous run Q[f][ra][b0..b1),
// Q[f][ra+1][b0..b1) (consecutive rows are adjacent in Q).
__global__ __launch_bounds__(256) void k_spod_fft(const float* __restrict__ X, int64_t ldx, int64_t M,
                                                  const double* __restrict__ mean, const float* __restrict__ win,
                                                  const float* __restrict__ twc, const float* __restrict__ tws,
                                                  int nDFT, int logN, int step, int nBlks, int nf, int BC, int P,
                                                  const float* __restrict__ fscale, float2* __restrict__ Q) {
  extern __shared__ float2 fsm[];
  const int N = nDFT, H = nDFT / 2;
  float2* A = fsm;              // [P][N]: P blocks transformed side by side
  float2* B = A + P * N;        // [P][N]
  float2* tw = B + P * N;       // [N / 2]
  float2* stg = tw + H;         // [nf][2][BC]
  // Row pairs are split into 8 contiguous ranges, one per XCD (workgroup x
  // runs on XCD x mod 8), so neighbouring pairs, whose output runs share
  // cache lines, are written through the same L2.
  const int64_t nWG = gridDim.x, per = nWG >> 3, rem = nWG & 7;
  const int64_t xcd = blockIdx.x & 7, xi = blockIdx.x >> 3;
  const int64_t pair = nWG < 8 ? int64_t(blockIdx.x) : xcd * per + (xcd < rem ? xcd : rem) + xi;
  const int64_t ra = pair * 2, rb = ra + 1;
  const bool va = ra < M, vb = rb < M;
  const double mua = (mean && va) ? mean[ra] : 0.0, mub = (mean && vb) ? mean[rb] : 0.0;
  for (int q = threadIdx.x; q < H; q += 256) tw[q] = make_float2(twc[q], -tws[q]);
  // Samples of the next pass are fetched into registers while the current
  // pass is transformed (at most P N / 256 <= 8 per thread and row).
  constexpr int EMAX = 8;
  float xa[EMAX], xb[EMAX];
  auto fetch = [&](int gb, int np) {
#pragma unroll
    for (int u = 0; u < EMAX; u++) {
      const int e = threadIdx.x + u * 256;
      if (e < np * N) {
        const int pi = e / N, t = e - pi * N;
        const int64_t t0 = int64_t(gb + pi) * step + t;
        xa[u] = va ? X[ra * ldx + t0] : 0.0f;
        xb[u] = vb ? X[rb * ldx + t0] : 0.0f;
      }
    }
  };
  fetch(0, nBlks < P ? nBlks : P);
  for (int b0 = 0; b0 < nBlks; b0 += BC) {
    const int nb = nBlks - b0 < BC ? nBlks - b0 : BC;
    for (int bp = 0; bp < nb; bp += P) {
      const int np = nb - bp < P ? nb - bp : P;  // blocks in this pass
#pragma unroll
      for (int u = 0; u < EMAX; u++) {
        const int e = threadIdx.x + u * 256;
        if (e < np * N) {
          const int t = e & (N - 1);
          A[e] = make_float2(float(double(xa[u]) - mua) * win[t], float(double(xb[u]) - mub) * win[t]);
        }
      }
      __syncthreads();
      {  // next pass: the rest of this chunk, else the start of the next one
        int gb = b0 + bp + P, cnt = nb - bp - P;
        if (cnt <= 0) {
          gb = b0 + BC;
          cnt = nBlks - gb < BC ? nBlks - gb : BC;
        }
        if (cnt > 0) fetch(gb, cnt < P ? cnt : P);
      }
      float2* src = A;
      float2* dst = B;
      for (int s = 0, m = 1; s < logN; s++, m <<= 1) {
        const int tstride = N >> (s + 1);  // N / 2m
        for (int jj = threadIdx.x; jj < np * H; jj += 256) {
          const int pi = jj / H, j = jj - pi * H, k = j & (m - 1);
          const float2* sp = src + pi * N;
          float2* dp = dst + pi * N;
          const float2 a = sp[j], bv = sp[j + H], w = tw[k * tstride];
          const float2 wb = make_float2(w.x * bv.x - w.y * bv.y, w.x * bv.y + w.y * bv.x);
          dp[2 * j - k] = make_float2(a.x + wb.x, a.y + wb.y);
          dp[2 * j - k + m] = make_float2(a.x - wb.x, a.y - wb.y);
        }
        __syncthreads();
        float2* t = src; src = dst; dst = t;
      }
      for (int e = threadIdx.x; e < np * nf; e += 256) {
        const int pi = e / nf, f = e - pi * nf, bi = bp + pi;
        const float2 z = src[pi * N + f], zr = src[pi * N + ((N - f) & (N - 1))];
        const float sc = fscale[f];
        // Y_a = (z + conj zr) / 2,  Y_b = (z - conj zr) / 2i
        stg[(f * 2 + 0) * BC + bi] = make_float2(0.5f * (z.x + zr.x) * sc, 0.5f * (z.y - zr.y) * sc);
        stg[(f * 2 + 1) * BC + bi] = make_float2(0.5f * (z.y + zr.y) * sc, -0.5f * (z.x - zr.x) * sc);
      }
      __syncthreads();  // A is reloaded by the next pass
    }
    // frequency f: rows ra, ra + 1 x blocks b0 .. b0 + nb, one contiguous run
    // when both rows exist (Q[f][row][b] with row stride nBlks)
    const int run = 2 * nb;
    for (int idx = threadIdx.x; idx < nf * run; idx += 256) {
      const int f = idx / run, e = idx - f * run, r = e / nb, bi = e - r * nb;
      const int64_t row = ra + r;
      if (row < M) Q[(int64_t(f) * M + row) * nBlks + b0 + bi] = stg[(f * 2 + r) * BC + bi];
    }
    __syncthreads();  // the staging area is refilled by the next chunk
  }
}

// SPOD modes: P[(j M + i) nf + k] = Re U[order[k]][i][j], the real parts of
// the left singular vectors in pyLOM's column-major flattening with the
// frequencies in sorted order (pyLOM/SPOD/wrapper.py:84, 90-94).  U holds nf
// complex64 M x nB matrices.  One workgroup per tile of I points x KC
// frequencies, staged through LDS: each U row segment is read contiguously
// (I nB values per frequency) and each P row segment written contiguously
// (whole rows when all frequencies fit one tile).
template <class TO>
__global__ __launch_bounds__(256) void k_spod_modes(const float2* __restrict__ U, int64_t M, int nB, int nf,
                                                    const int64_t* __restrict__ order, int I, int KC,
                                                    TO* __restrict__ P) {
  extern __shared__ float sre[];  // [KC][I nB]
  const int64_t i0 = int64_t(blockIdx.x) * I;
  const int k0 = int(blockIdx.y) * KC;
  const int ni = int(M - i0 < I ? M - i0 : I);
  const int nk = nf - k0 < KC ? nf - k0 : KC;
  const int run = ni * nB, ld = I * nB;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  // loads: wave w takes frequencies w, w + 4, ..., eight at a time so that
  // eight independent loads are in flight per lane; one contiguous segment
  // per frequency
  constexpr int UF = 8;
  for (int kb = w; kb < nk; kb += 4 * UF) {
    const float2* src[UF];
#pragma unroll
    for (int u = 0; u < UF; u++) {
      const int kl = kb + 4 * u;
      int64_t ok = kl < nk ? order[k0 + kl] : 0;
      ok = ok < 0 ? 0 : (ok >= nf ? nf - 1 : ok);  // order is a permutation; the clamp only guards
      src[u] = U + (ok * M + i0) * nB;
    }
    for (int e = l; e < run; e += 64) {
      float v[UF];
#pragma unroll
      for (int u = 0; u < UF; u++) v[u] = kb + 4 * u < nk ? src[u][e].x : 0.0f;
#pragma unroll
      for (int u = 0; u < UF; u++)
        if (kb + 4 * u < nk) sre[(kb + 4 * u) * ld + e] = v[u];
    }
  }
  __syncthreads();
  // stores: wave w takes output rows (j, il) = w, w + 4, ...; lanes cover the
  // row's frequencies
  for (int r = w; r < run; r += 4) {
    const int j = r / ni, il = r - j * ni;
    TO* dst = P + (int64_t(j) * M + i0 + il) * nf + k0;
    const float* col = sre + il * nB + j;
    for (int kl = l; kl < nk; kl += 64) dst[kl] = TO(col[kl * ld]);
  }
}

}  // namespace

extern "C" int pc_spod_modes(pc_ctx* ctx, const void* U, int64_t M, int64_t nB, int64_t nf, const int64_t* order,
                             void* P, int out_f64) {
  if (!ctx) return pc_fail(nullptr, PC_ERR_INVALID, "ctx is NULL");
  if (!U || !order || !P) return pc_fail(ctx, PC_ERR_INVALID, "pc_spod_modes: NULL pointer");
  if (M <= 0 || nB <= 0 || nf <= 0) return pc_fail(ctx, PC_ERR_SHAPE, "pc_spod_modes: empty shape");
  constexpr int64_t kLds = 12288;  // staged floats per workgroup (48 KB)
  if (nB > kLds || nf > (int64_t(1) << 30))
    return pc_fail(ctx, PC_ERR_SHAPE, "pc_spod_modes: nBlks = %lld is larger than supported", (long long)nB);
  // all frequencies in one chunk when they fit (whole P rows written
  // contiguously), else balanced chunks of at most 64
  int64_t KC = nf;
  if (KC * nB > kLds) {
    const int64_t nch = ceil_div(nf, 64);
    KC = ceil_div(nf, nch);
    if (KC * nB > kLds) KC = std::max<int64_t>(1, kLds / nB);
  }
  const int64_t I = std::max<int64_t>(1, kLds / (KC * nB));
  const dim3 g(unsigned(ceil_div(M, I)), unsigned(ceil_div(nf, KC)));
  const size_t lds = size_t(KC * I * nB) * sizeof(float);
  hipStream_t st = ctx->stream;
  if (out_f64)
    hipLaunchKernelGGL((k_spod_modes<double>), g, dim3(256), lds, st, (const float2*)U, M, int(nB), int(nf), order,
                       int(I), int(KC), (double*)P);
  else
    hipLaunchKernelGGL((k_spod_modes<float>), g, dim3(256), lds, st, (const float2*)U, M, int(nB), int(nf), order,
                       int(I), int(KC), (float*)P);
  PC_HIP(ctx, hipGetLastError());
  return PC_OK;
}

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
  // LDS of the FFT: A, B (P blocks of N complex each, P = 512 / N so that the
  // 256 threads all have a butterfly), twiddles (N/2) and the staging of BC
  // blocks' spectra (nf x 2 x BC complex), kept within 56 KB
  if (pow2 && nDFT <= 2048) {
    int64_t P = nDFT >= 512 ? 1 : 512 / nDFT;
    P = P > nBlks ? nBlks : P;
    const int64_t fixed = (2 * P * nDFT + nDFT / 2) * int64_t(sizeof(float2));
    int64_t bc = (57344 - fixed) / (nf * 2 * int64_t(sizeof(float2)));
    bc = bc < 1 ? 1 : (bc > nBlks ? nBlks : bc);
    const size_t lds = size_t(fixed + nf * 2 * bc * int64_t(sizeof(float2)));
    hipLaunchKernelGGL(k_spod_fft, dim3(unsigned(ceil_div(M, 2))), dim3(256), lds, st, X, ldx, M,
                       remove_mean ? mean : nullptr, win, twc, tws, int(nDFT), logN, int(step), int(nBlks), int(nf),
                       int(bc), int(P), fsc, (float2*)Q);
  }
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
