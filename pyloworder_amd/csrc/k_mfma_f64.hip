// FP64 MFMA streaming kernels (v_mfma_f64_16x16x4f64) of the plain POD path.
//
// A pass over the tall matrix A (m x n, row-major, n <= NB) is split into
// `chunks` contiguous row ranges, one workgroup each (8 waves).  Each
// workgroup stages tiles of TR = 32 full rows in LDS (so the row-local
// temporal mean can be removed while loading), and
//   gram  : accumulates the upper 16x16 blocks of Y^T Y in registers,
//   zgram : forms Z = Y W (W n x n, FP64, from L2) into LDS, then Z^T Z,
//   gemm  : writes C = Y W (W n x nc) straight to HBM,
//   xgram : accumulates A^T B for two tall operands (projection, sketches).
// Partial Grams (one NB x NB slab per chunk, upper blocks) are reduced in a
// fixed order afterwards, so results never depend on scheduling.
//
// MFMA lane layouts used (wave64, verified on gfx950):
//   A operand (16x4):  lane l holds A[l % 16][l / 16]
//   B operand (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16), reg r:  D[4 * r + l / 16][l % 16]
#include "common.hpp"
#include "kernels.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int TR = 32;       // rows per tile
constexpr int NTH = 512;     // threads per workgroup
constexpr int PADD = 2;      // LDS row padding (doubles)

PC_DEV d4 mfma64(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

template <class TS> PC_DEV double ld_d(const TS* p) { return double(*p); }

// Load rows [row0, row0+TR) of X (n valid columns) into S[TR][NB+PADD], removing
// the row mean (center>=1) and dividing by the variance (center==2).  mode_in:
// 0 compute the statistics here (and store them when mean_out != NULL),
// 1 read them from mean_in / var_in.
template <class TS, int NB>
PC_DEV void load_tile(const TS* __restrict__ X, int64_t ldx, int64_t m, int n, int64_t row0, int center,
                      int stats_in, const double* __restrict__ mean_in, const double* __restrict__ var_in,
                      double* __restrict__ mean_out, double* __restrict__ var_out, double (*S)[NB + PADD]) {
  constexpr int TPR = NTH / TR;           // 16 threads per row
  constexpr int PER = (NB + TPR - 1) / TPR;
  int t = threadIdx.x;
  int r = t / TPR, sub = t % TPR;
  int64_t row = row0 + r;
  bool rv = row < m;
  double v[PER];
  const TS* xr = X + (rv ? row : 0) * ldx;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int c = sub + q * TPR;
    v[q] = (rv && c < n) ? ld_d(xr + c) : 0.0;
  }
  if (center) {
    double mu, var = 1.0;
    if (stats_in) {
      mu = rv ? mean_in[row] : 0.0;
      if (center == 2) var = rv ? var_in[row] : 1.0;
    } else {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < PER; q++) s += v[q];
      for (int o = TPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
      mu = s / double(n);
      if (center == 2) {
        double s2 = 0.0;
#pragma unroll
        for (int q = 0; q < PER; q++) {
          int c = sub + q * TPR;
          double d = c < n ? v[q] - mu : 0.0;
          s2 += d * d;
        }
        for (int o = TPR / 2; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
        var = s2 / double(n);
      }
      if (rv && sub == 0) {
        if (mean_out) mean_out[row] = mu;
        if (var_out && center == 2) var_out[row] = var;
      }
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      int c = sub + q * TPR;
      if (c < n && rv) v[q] = center == 2 ? (v[q] - mu) / var : v[q] - mu;
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int c = sub + q * TPR;
    if (c < NB) S[r][c] = v[q];
  }
}

// (I, J) of the b-th upper block in row-major triangular order of nb x nb
PC_DEV void tri_block(int b, int nb, int& I, int& J) {
  int i = 0, off = 0;
  while (off + (nb - i) <= b) { off += nb - i; i++; }
  I = i;
  J = i + (b - off);
}

// Accumulate the Gram of the tile S into the slot accumulators.
template <int NB, int SL>
PC_DEV void gram_tile(double (*S)[NB + PADD], const int* bI, const int* bJ, int nslot, d4* acc) {
  int l = threadIdx.x & 63;
  int kr = l >> 4, cc = l & 15;
#pragma unroll
  for (int kk = 0; kk < TR / 4; kk++) {
    const double* srow = S[kk * 4 + kr];
#pragma unroll
    for (int s = 0; s < SL; s++) {
      if (s < nslot) {
        double a = srow[bI[s] * 16 + cc];
        double b = srow[bJ[s] * 16 + cc];
        acc[s] = mfma64(a, b, acc[s]);
      }
    }
  }
}

// Z = S W for the tile (TR x NB times NB x NB), result into Zs.  Slot s of
// wave w holds row block R and 16-column block J of Z.  With an upper
// triangular W (upper != 0) block J only needs the k-steps kk < 4 (J + 1)
// (rows < 16 (J + 1)): the k range is cut into segments, segment f running
// the slots s >= f up to slot f's end, so the loops carry no per-step
// conditions.  The products skipped are exact zeros: Z does not change.
// Waves 2k and 2k + 1 share a SIMD.  The column blocks are listed per wave
// pair as J_q = 4 q + (q odd ? 3 - k : k), i.e. k, 7 - k, 8 + k, 15 - k at
// NB = 256, so every SIMD does the same share of the k-steps.  With 2 or 4
// slots a wave takes both row blocks of its column blocks (wave 2k: J_0 and
// J_3, wave 2k + 1: J_1 and J_2 at NB = 256; J_0 and J_1 at NB = 128), so
// each W element it loads feeds two products; otherwise slot s takes row
// block w & 1 of J_s.  Slots are ascending in J.  Which wave forms a block
// does not change its products.
template <int NB>
PC_DEV void zmul_tile(double (*S)[NB + PADD], const double* __restrict__ W, double (*Zs)[NB + PADD], int upper) {
  constexpr int nb = NB / 16;
  constexpr int NBLK = 2 * nb;                    // TR/16 = 2 row blocks
  constexpr int SLZ = (NBLK + 7) / 8;
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int kr = l >> 4, cc = l & 15;
  d4 acc[SLZ];
#pragma unroll
  for (int s = 0; s < SLZ; s++) acc[s] = d4{0, 0, 0, 0};
  auto zblk = [&](int s, int& R, int& J) {
    const int k = w >> 1;
    if constexpr (NBLK == 8 * SLZ && (SLZ == 2 || SLZ == 4)) {
      const int h = w & 1, i = s >> 1;
      const int q = SLZ == 4 ? (h ? 1 + i : 3 * i) : h;
      R = s & 1;
      J = 4 * q + ((q & 1) ? 3 - k : k);
      return true;
    } else {
      const int b = w + 8 * s;
      R = b & 1;
      J = NBLK % 8 == 0 ? 4 * s + ((s & 1) ? 3 - k : k) : b >> 1;
      return b < NBLK;
    }
  };
  auto step = [&](int kk, int first) {
    const double* wrow = W + int64_t(kk * 4 + kr) * NB;
#pragma unroll
    for (int s = 0; s < SLZ; s++) {
      int R, J;
      if (zblk(s, R, J) && s >= first) {
        double a = S[R * 16 + cc][kk * 4 + kr];
        double bb = wrow[J * 16 + cc];
        acc[s] = mfma64(a, bb, acc[s]);
      }
    }
  };
  if (NBLK % 8 == 0 && upper) {
    int kbeg = 0;
#pragma unroll
    for (int f = 0; f < SLZ; f++) {
      int Rf, Jf;
      zblk(f, Rf, Jf);
      const int kend = __builtin_amdgcn_readfirstlane(4 * (Jf + 1));
#pragma unroll 4
      for (int kk = kbeg; kk < kend; kk++) step(kk, f);
      kbeg = kend;
    }
  } else {
#pragma unroll 4
    for (int kk = 0; kk < NB / 4; kk++) step(kk, 0);
  }
#pragma unroll
  for (int s = 0; s < SLZ; s++) {
    int R, J;
    if (zblk(s, R, J)) {
#pragma unroll
      for (int r = 0; r < 4; r++) Zs[R * 16 + 4 * r + kr][J * 16 + cc] = acc[s][r];
    }
  }
}

// ---------------------------------------------------------------------------
// gram / zgram kernel
// ---------------------------------------------------------------------------
template <class TS, int NB, bool ZMODE>
__global__ __launch_bounds__(NTH) void k_gram_f64(const TS* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                                                  int stats_in, const double* mean_in, const double* var_in,
                                                  double* mean_out, double* var_out, const double* __restrict__ W,
                                                  int64_t tiles_per_chunk, double* __restrict__ part, int upper,
                                                  int64_t zx = 0, int64_t zb = 0) {
  X = zoff(X, zx);
  mean_in = zoff(mean_in, zb);
  var_in = zoff(var_in, zb);
  mean_out = zoff(mean_out, zb);
  var_out = zoff(var_out, zb);
  W = zoff(W, zb);
  part = zoff(part, zb);
  constexpr int nb = NB / 16;
  constexpr int NBLK = nb * (nb + 1) / 2;
  constexpr int SL = (NBLK + 7) / 8;
  __shared__ double S[TR][NB + PADD];
  __shared__ double Zs[ZMODE ? TR : 1][ZMODE ? NB + PADD : 1];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int bI[SL], bJ[SL];
  int nslot = 0;
#pragma unroll
  for (int s = 0; s < SL; s++) {
    int b = w + 8 * s;
    if (b < NBLK) { tri_block(b, nb, bI[s], bJ[s]); nslot = s + 1; }
    else { bI[s] = 0; bJ[s] = 0; }
  }
  d4 acc[SL];
#pragma unroll
  for (int s = 0; s < SL; s++) acc[s] = d4{0, 0, 0, 0};
  int64_t ntiles = ceil_div(m, TR);
  int64_t t0 = int64_t(blockIdx.x) * tiles_per_chunk;
  int64_t t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    __syncthreads();
    load_tile<TS, NB>(X, ldx, m, n, t * TR, center, stats_in, mean_in, var_in, mean_out, var_out, S);
    __syncthreads();
    if constexpr (ZMODE) {
      zmul_tile<NB>(S, W, Zs, upper);
      __syncthreads();
      gram_tile<NB, SL>(Zs, bI, bJ, nslot, acc);
    } else {
      gram_tile<NB, SL>(S, bI, bJ, nslot, acc);
    }
  }
  double* P = part + int64_t(blockIdx.x) * NB * NB;
  int kr = l >> 4, cc = l & 15;
#pragma unroll
  for (int s = 0; s < SL; s++) {
    if (s < nslot) {
#pragma unroll
      for (int r = 0; r < 4; r++) P[int64_t(bI[s] * 16 + 4 * r + kr) * NB + bJ[s] * 16 + cc] = acc[s][r];
    }
  }
}

// ---------------------------------------------------------------------------
// gemm kernel: C[m x nc] = Yc W, Y centred as the stats say, W NB x NCB FP64.
// ---------------------------------------------------------------------------
template <class TS, class TO, int NB, int NCB>
__global__ __launch_bounds__(NTH) void k_gemm_f64(const TS* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                                                  const double* mean_in, const double* var_in,
                                                  const double* __restrict__ W, int nc, int64_t tiles_per_chunk,
                                                  TO* __restrict__ C, int64_t ldc, int64_t zx = 0, int64_t zb = 0, int64_t zc = 0) {
  X = zoff(X, zx);
  mean_in = zoff(mean_in, zb);
  var_in = zoff(var_in, zb);
  W = zoff(W, zb);
  C = zoff(C, zc);
  constexpr int ncb = NCB / 16;
  constexpr int NBLK = 2 * ncb;
  constexpr int SLZ = (NBLK + 7) / 8;
  __shared__ double S[TR][NB + PADD];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int kr = l >> 4, cc = l & 15;
  // Slot s of wave w: row block R, 16-column block J.  With an even number of
  // slots per wave every wave takes both row blocks of SLZ / 2 column blocks,
  // so each W element it loads feeds two products (it fed one, and the two
  // waves of a column block loaded it twice).  The products of a block do not
  // depend on the wave that forms them.
  auto slot_blk = [&](int s, int& R, int& J) {
    if constexpr (SLZ % 2 == 0 && NBLK == 8 * SLZ) {
      R = s & 1;
      J = (w * SLZ + s) >> 1;
      return true;
    } else {
      const int b = w + 8 * s;
      R = b & 1;
      J = b >> 1;
      return b < NBLK;
    }
  };
  int64_t ntiles = ceil_div(m, TR);
  int64_t t0 = int64_t(blockIdx.x) * tiles_per_chunk;
  int64_t t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    __syncthreads();
    load_tile<TS, NB>(X, ldx, m, n, t * TR, center, 1, mean_in, var_in, nullptr, nullptr, S);
    __syncthreads();
    d4 acc[SLZ];
#pragma unroll
    for (int s = 0; s < SLZ; s++) acc[s] = d4{0, 0, 0, 0};
#pragma unroll 4
    for (int kk = 0; kk < NB / 4; kk++) {
      const double* wrow = W + int64_t(kk * 4 + kr) * NCB;
#pragma unroll
      for (int s = 0; s < SLZ; s++) {
        int R, J;
        if (slot_blk(s, R, J)) acc[s] = mfma64(S[R * 16 + cc][kk * 4 + kr], wrow[J * 16 + cc], acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < SLZ; s++) {
      int R, J;
      if (slot_blk(s, R, J)) {
        int col = J * 16 + cc;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int64_t row = t * TR + R * 16 + 4 * r + kr;
          if (row < m && col < nc) C[row * ldc + col] = TO(acc[s][r]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// cross Gram: P[chunk] += A^T B, A m x p (p <= PB), B m x q (q <= QB)
// ---------------------------------------------------------------------------
template <class TA, class TB, int PB, int QB>
__global__ __launch_bounds__(NTH) void k_xgram_f64(const TA* __restrict__ A, int64_t lda, int p,
                                                   const TB* __restrict__ B, int64_t ldb, int q, int64_t m,
                                                   int64_t tiles_per_chunk, double* __restrict__ part) {
  constexpr int pb = PB / 16, qb = QB / 16;
  constexpr int NBLK = pb * qb;
  constexpr int SL = (NBLK + 7) / 8;
  __shared__ double SA[TR][PB + PADD];
  __shared__ double SB[TR][QB + PADD];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int kr = l >> 4, cc = l & 15;
  d4 acc[SL];
#pragma unroll
  for (int s = 0; s < SL; s++) acc[s] = d4{0, 0, 0, 0};
  int64_t ntiles = ceil_div(m, TR);
  int64_t t0 = int64_t(blockIdx.x) * tiles_per_chunk;
  int64_t t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    __syncthreads();
    for (int idx = threadIdx.x; idx < TR * PB; idx += NTH) {
      int r = idx / PB, c = idx % PB;
      int64_t row = t * TR + r;
      SA[r][c] = (row < m && c < p) ? double(A[row * lda + c]) : 0.0;
    }
    for (int idx = threadIdx.x; idx < TR * QB; idx += NTH) {
      int r = idx / QB, c = idx % QB;
      int64_t row = t * TR + r;
      SB[r][c] = (row < m && c < q) ? double(B[row * ldb + c]) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TR / 4; kk++) {
#pragma unroll
      for (int s = 0; s < SL; s++) {
        int b = w + 8 * s;
        if (b < NBLK) {
          int I = b / qb, J = b % qb;
          acc[s] = mfma64(SA[kk * 4 + kr][I * 16 + cc], SB[kk * 4 + kr][J * 16 + cc], acc[s]);
        }
      }
    }
  }
  double* P = part + int64_t(blockIdx.x) * PB * QB;
#pragma unroll
  for (int s = 0; s < SL; s++) {
    int b = w + 8 * s;
    if (b < NBLK) {
      int I = b / qb, J = b % qb;
#pragma unroll
      for (int r = 0; r < 4; r++) P[int64_t(I * 16 + 4 * r + kr) * QB + J * 16 + cc] = acc[s][r];
    }
  }
}

template <class F>
int dispatch_nb(int n, F&& f) {
  if (n <= 16) return f(std::integral_constant<int, 16>{});
  if (n <= 32) return f(std::integral_constant<int, 32>{});
  if (n <= 64) return f(std::integral_constant<int, 64>{});
  if (n <= 128) return f(std::integral_constant<int, 128>{});
  if (n <= 256) return f(std::integral_constant<int, 256>{});
  return -1;
}

}  // namespace

int f64_nb(int n) { return n <= 16 ? 16 : n <= 32 ? 32 : n <= 64 ? 64 : n <= 128 ? 128 : n <= 256 ? 256 : -1; }
int f64_tile_rows() { return TR; }

// P1 / P2 launchers -----------------------------------------------------------
template <class TS>
static int launch_gram_t(hipStream_t st, const TS* X, int64_t ldx, int64_t m, int n, int center, int stats_in,
                         const double* mean_in, const double* var_in, double* mean_out, double* var_out,
                         const double* W, int chunks, int64_t tpc, double* part, const PassBatch& bz, bool upper) {
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    const dim3 g(unsigned(chunks), 1, unsigned(bz.nz));
    if (W)
      hipLaunchKernelGGL((k_gram_f64<TS, NB, true>), g, dim3(NTH), 0, st, X, ldx, m, n, center, stats_in, mean_in,
                         var_in, mean_out, var_out, W, tpc, part, upper ? 1 : 0, bz.zx, bz.zb);
    else
      hipLaunchKernelGGL((k_gram_f64<TS, NB, false>), g, dim3(NTH), 0, st, X, ldx, m, n, center, stats_in, mean_in,
                         var_in, mean_out, var_out, W, tpc, part, 0, bz.zx, bz.zb);
    return 0;
  });
}

int launch_gram_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int n, int center,
                    int stats_in, const double* mean_in, const double* var_in, double* mean_out, double* var_out,
                    const double* W, int chunks, int64_t tpc, double* part, const PassBatch& bz, bool upper) {
  if (dt == PC_F64)
    return launch_gram_t<double>(st, (const double*)X, ldx, m, n, center, stats_in, mean_in, var_in, mean_out, var_out,
                                 W, chunks, tpc, part, bz, upper);
  return launch_gram_t<float>(st, (const float*)X, ldx, m, n, center, stats_in, mean_in, var_in, mean_out, var_out, W,
                              chunks, tpc, part, bz, upper);
}

template <class TS, class TO>
static int launch_gemm_t(hipStream_t st, const TS* X, int64_t ldx, int64_t m, int n, int center,
                         const double* mean_in, const double* var_in, const double* W, int nc, int chunks,
                         int64_t tpc, TO* C, int64_t ldc, const PassBatch& bz) {
  int ncb = f64_nb(nc);
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    return dispatch_nb(ncb, [&](auto ncc) {
      constexpr int NCB = decltype(ncc)::value;
      hipLaunchKernelGGL((k_gemm_f64<TS, TO, NB, NCB>), dim3(unsigned(chunks), 1, unsigned(bz.nz)), dim3(NTH), 0, st,
                         X, ldx, m, n, center, mean_in, var_in, W, nc, tpc, C, ldc, bz.zx, bz.zb, bz.zc);
      return 0;
    });
  });
}

int launch_gemm_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int n, int center,
                    const double* mean_in, const double* var_in, const double* W, int nc, int chunks, int64_t tpc,
                    void* C, int64_t ldc, const PassBatch& bz) {
  if (dt == PC_F64)
    return launch_gemm_t<double, double>(st, (const double*)X, ldx, m, n, center, mean_in, var_in, W, nc, chunks, tpc,
                                         (double*)C, ldc, bz);
  return launch_gemm_t<float, float>(st, (const float*)X, ldx, m, n, center, mean_in, var_in, W, nc, chunks, tpc,
                                     (float*)C, ldc, bz);
}

template <class T>
static int launch_xgram_t(hipStream_t st, const T* A, int64_t lda, int p, const T* B, int64_t ldb, int q, int64_t m,
                          int chunks, int64_t tpc, double* part) {
  int pb = f64_nb(p), qb = f64_nb(q);
  if (pb < 0 || qb < 0) return -1;
  if (int64_t(pb) * qb > 128 * 256) return -1;  // register budget: at most 128 blocks
  return dispatch_nb(pb, [&](auto pc_) {
    constexpr int PB = decltype(pc_)::value;
    return dispatch_nb(qb, [&](auto qc_) {
      constexpr int QB = decltype(qc_)::value;
      if constexpr (PB * QB <= 128 * 256) {
        hipLaunchKernelGGL((k_xgram_f64<T, T, PB, QB>), dim3(chunks), dim3(NTH), 0, st, A, lda, p, B, ldb, q, m, tpc,
                           part);
        return 0;
      } else {
        return -1;
      }
    });
  });
}

int launch_xgram_f64(hipStream_t st, pc_dtype dt, const void* A, int64_t lda, int p, const void* B, int64_t ldb, int q,
                     int64_t m, int chunks, int64_t tpc, double* part) {
  if (dt == PC_F64)
    return launch_xgram_t<double>(st, (const double*)A, lda, p, (const double*)B, ldb, q, m, chunks, tpc, part);
  return launch_xgram_t<float>(st, (const float*)A, lda, p, (const float*)B, ldb, q, m, chunks, tpc, part);
}
