// Mixed-precision streaming kernels of the float32 POD path (fp16 split on
// v_mfma_f32_32x32x16_f16, FP32 accumulation inside a workgroup, FP64 across
// workgroups).  A float32 value x is carried as hi = fp16(x 2^e) and
// lo = fp16(x 2^e - hi) (22 significant bits) with power-of-two scales that are
// undone exactly.  One workgroup of 8 waves per row chunk.
//
//   P1 k_gram_mx2  : G = Ys^T Ys, products hi.hi + hi.lo + lo.hi; one running
//                    scale per workgroup.  Each wave loads one 16-row k-step
//                    straight into fragment order: a lane reads V consecutive
//                    floats (float4 for n > 64) of 8 rows, and the V values go
//                    to V different 32-column blocks (block b = g V + u holds
//                    the physical columns g 32V + V c + u, c = 0..31).  The
//                    Gram is written back with that permutation undone.
//   P2 k_zgram_mx2 : Z = Ys W1 (W1 already rounded to fp16: hi.W + lo.W), then
//                    G2 = Z^T Z (three products).
//   P3 k_gemm_mx2  : C = Ys W (hi.hi + hi.lo + lo.hi), per-row scales of Ys,
//                    per-column scales of W.
// In P2/P3 every load instruction reads whole rows (float4 per lane); the fp16
// pieces go to LDS in A-fragment order, and each wave keeps the B fragments of
// its own 32 output columns (W) in registers for the whole kernel.
//
// Row shift: Ys = Y - a 1^T with a_i = x_i0 (center = 1) or the exact
// temporal mean (center = 2, then also divided by the variance).  Any a gives
// (Y - a 1^T) C = Y C for the centring projector C = I - 1 1^T / n, so the
// exact centring is applied to the small n x n factor; the shift only removes
// the large common part before the fp16 split, and x - x_i0 is exact in
// float32 whenever x and x_i0 are within a factor two.
//
// MFMA lane layouts (wave64, gfx950), hl = lane / 32, c = lane % 32:
//   A (32x16): lane holds A[c][8 hl + t], t = 0..7
//   B (16x32): lane holds B[8 hl + t][c]
//   D (32x32): reg r holds D[8 (r / 4) + 4 hl + r % 4][c]
#include "common.hpp"
#include "kernels.hpp"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NT2 = 512;  // 8 waves
constexpr int NW2 = NT2 / 64;
constexpr int EUN = 1 << 20;  // running exponent not set yet

PC_DEV f16v mfma(h8 a, h8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }

// e with |x| 2^e < 2^14 for |x| <= mx (mx > 0)
PC_DEV int scale_exp(float mx) {
  int E;
  frexpf(mx, &E);
  return 14 - E;
}
PC_DEV float group_max(float v, int g) {
  for (int o = g >> 1; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
PC_DEV void split(float x, _Float16& h, _Float16& lo) {
  h = (_Float16)x;
  lo = (_Float16)(x - (float)h);
}

// (I, J) of the idx-th upper 32x32 block, rows taken in pairs (p, NBLK-1-p)
// so that each wave's consecutive slots mostly share their A operand
template <int NBLK>
PC_DEV bool upper_slot(int idx, int& I, int& J) {
#pragma unroll
  for (int p = 0; p < (NBLK + 1) / 2; p++) {
    const int a = p, b = NBLK - 1 - p;
    if (idx < NBLK - a) { I = a; J = a + idx; return true; }
    idx -= NBLK - a;
    if (b != a) {
      if (idx < NBLK - b) { I = b; J = b + idx; return true; }
      idx -= NBLK - b;
    }
  }
  I = J = 0;
  return false;
}

// running workgroup exponent: a larger tile maximum lowers it and rescales acc
template <int S>
PC_DEV int update_scale(float tmax, int e_run, f16v* acc) {
  if (!(tmax > 0.0f)) return e_run;
  const int et = scale_exp(tmax);
  if (e_run == EUN) return et;
  if (et < e_run) {
    const float f = ldexpf(1.0f, 2 * (et - e_run));
#pragma unroll
    for (int s = 0; s < S; s++) acc[s] *= f;
    return et;
  }
  return e_run;
}

// shifted value of x in row `row` (see the file comment); a is x_i0 for center 1
PC_DEV float shifted(float x, float a, int center, int64_t row, const double* mean, const double* var) {
  if (center == 1) return x - a;
  if (center == 2) return float((double(x) - mean[row]) * (1.0 / var[row]));
  return x;
}

// Gram accumulation of one k-step from fragments F[(blk * 2 + part) * 64 + lane]
template <int S>
PC_DEV void gram_step(const h8* __restrict__ F, const int* sI, const int* sJ, const bool* sv, f16v* acc, int l) {
  int curI = -1;
  h8 ah = {}, al = {};
#pragma unroll
  for (int s = 0; s < S; s++) {
    if (sv[s]) {
      if (sI[s] != curI) {
        curI = sI[s];
        ah = F[(curI * 2 + 0) * 64 + l];
        al = F[(curI * 2 + 1) * 64 + l];
      }
      h8 bh, bl;
      if (sJ[s] == curI) { bh = ah; bl = al; }
      else { bh = F[(sJ[s] * 2 + 0) * 64 + l]; bl = F[(sJ[s] * 2 + 1) * 64 + l]; }
      acc[s] = mfma(ah, bh, acc[s]);
      acc[s] = mfma(ah, bl, acc[s]);
      acc[s] = mfma(al, bh, acc[s]);
    }
  }
}

// partial slab (upper triangle of NB x NB, FP64); prow/pcol map block + position
// to the physical row/column.  Each physical upper entry is written once.
template <int NB, int S, class Map>
PC_DEV void write_upper(const f16v* acc, const int* sI, const int* sJ, const bool* sv, double inv, double* P, int l,
                        Map phys) {
  const int hl = l >> 5, c = l & 31;
#pragma unroll
  for (int s = 0; s < S; s++) {
    if (!sv[s]) continue;
    const int I = sI[s], J = sJ[s];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int i = 8 * (r >> 2) + 4 * hl + (r & 3);
      const int pr = phys(I, i), pc = phys(J, c);
      if (I == J && pr > pc) continue;
      const int lo = pr < pc ? pr : pc, hi = pr < pc ? pc : pr;
      P[int64_t(lo) * NB + hi] = double(acc[s][r]) * inv;
    }
  }
}

// ---------------------------------------------------------------------------
// P1
// ---------------------------------------------------------------------------
template <int NB, bool VEC>
__global__ __launch_bounds__(NT2, 1) void k_gram_mx2(const float* __restrict__ X, int64_t ldx, int64_t m, int n,
                                                     int center, double* __restrict__ mean_io,
                                                     const double* __restrict__ var, int64_t tpc,
                                                     double* __restrict__ part) {
  constexpr int V = NB >= 128 ? 4 : NB / 32;
  constexpr int NG = NB / (32 * V);
  constexpr int NBLK = NB / 32;
  constexpr int TR = 16 * NW2;
  constexpr int NU = NBLK * (NBLK + 1) / 2;
  constexpr int S = (NU + NW2 - 1) / NW2;
  __shared__ h8 F[NW2 * NBLK * 2 * 64];
  __shared__ float red[NW2];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  int sI[S], sJ[S];
  bool sv[S];
#pragma unroll
  for (int s = 0; s < S; s++) sv[s] = upper_slot<NBLK>(w * S + s, sI[s], sJ[s]);
  f16v acc[S];
#pragma unroll
  for (int s = 0; s < S; s++) acc[s] = f16v{};
  int e_run = EUN;
  const int64_t ntiles = ceil_div(m, TR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  float v[NG][8][V];
  auto load = [&](int64_t tile) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int64_t row = tile * TR + 16 * w + 8 * hl + t;
      const bool rv = row < m;
      const float* xr = X + (rv ? row : 0) * ldx;
#pragma unroll
      for (int g = 0; g < NG; g++) {
        const int col = g * 32 * V + V * c;
        if (VEC && rv && col + V <= n) {
          if constexpr (V == 4) {
            f4 q = *reinterpret_cast<const f4*>(xr + col);
            v[g][t][0] = q[0]; v[g][t][1] = q[1]; v[g][t][2] = q[2]; v[g][t][3] = q[3];
          } else {
#pragma unroll
            for (int u = 0; u < V; u++) v[g][t][u] = xr[col + u];
          }
        } else {
#pragma unroll
          for (int u = 0; u < V; u++) v[g][t][u] = (rv && col + u < n) ? xr[col + u] : 0.0f;
        }
      }
    }
  };
  if (t0 < t1) load(t0);
  for (int64_t tile = t0; tile < t1; tile++) {
    // shift and tile maximum
    float mx = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int64_t row = tile * TR + 16 * w + 8 * hl + t;
      const bool rv = row < m;
      const float a = __shfl(v[0][t][0], hl * 32);  // x_i0 sits in lane c = 0 of this half
      if (center == 1 && rv && c == 0 && mean_io) mean_io[row] = double(a);
#pragma unroll
      for (int g = 0; g < NG; g++)
#pragma unroll
        for (int u = 0; u < V; u++) {
          const int col = g * 32 * V + V * c + u;
          float y = (rv && col < n) ? shifted(v[g][t][u], a, center, rv ? row : 0, mean_io, var) : 0.0f;
          v[g][t][u] = y;
          mx = fmaxf(mx, fabsf(y));
        }
    }
    mx = group_max(mx, 64);
    if (l == 0) red[w] = mx;
    __syncthreads();
    float tmax = red[0];
#pragma unroll
    for (int q = 1; q < NW2; q++) tmax = fmaxf(tmax, red[q]);
    e_run = update_scale<S>(tmax, e_run, acc);
    const float sc = ldexpf(1.0f, e_run == EUN ? 0 : e_run);
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
      for (int u = 0; u < V; u++) {
        h8 hi, lo;
#pragma unroll
        for (int t = 0; t < 8; t++) {
          _Float16 h, o;
          split(v[g][t][u] * sc, h, o);
          hi[t] = h;
          lo[t] = o;
        }
        const int blk = g * V + u;
        F[((w * NBLK + blk) * 2 + 0) * 64 + l] = hi;
        F[((w * NBLK + blk) * 2 + 1) * 64 + l] = lo;
      }
    __syncthreads();
    if (tile + 1 < t1) load(tile + 1);  // in flight during the products
#pragma unroll
    for (int ks = 0; ks < NW2; ks++) gram_step<S>(F + ks * NBLK * 2 * 64, sI, sJ, sv, acc, l);
    __syncthreads();
  }
  const double inv = ldexp(1.0, e_run == EUN ? 0 : -2 * e_run);
  write_upper<NB, S>(acc, sI, sJ, sv, inv, part + int64_t(blockIdx.x) * NB * NB, l,
                     [](int blk, int pos) { return (blk / V) * 32 * V + V * pos + (blk % V); });
}

// ---------------------------------------------------------------------------
// P2 / P3 row tiles: 32 rows (64 for n <= 32), each wave loads whole rows with
// one float4 per lane
// ---------------------------------------------------------------------------
template <int NB>
struct RowTile {
  static constexpr int LPR = NB / 4;        // lanes per row
  static constexpr int RPI = 64 / LPR;      // rows per load instruction
  static constexpr int TR = NW2 * RPI > 32 ? NW2 * RPI : 32;  // rows per tile
  static constexpr int NR = TR / 32;        // 32-row blocks per tile
  static constexpr int RPW = TR / NW2;      // rows per wave
  static constexpr int NLD = RPW / RPI;     // load instructions per wave
  static constexpr int KSW = NB / 16;       // k-steps of the products
  static constexpr int FRAGB = 64 * 16 + 16;  // bytes per fragment (padded against bank conflicts)
  static constexpr int FBYTES = NR * KSW * 2 * FRAGB;
};

template <int NB, bool VEC>
PC_DEV void rows_load(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int64_t tile, f4* v) {
  using RT = RowTile<NB>;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int col = 4 * (l % RT::LPR);
#pragma unroll
  for (int i = 0; i < RT::NLD; i++) {
    const int64_t row = tile * RT::TR + w * RT::RPW + i * RT::RPI + l / RT::LPR;
    const bool rv = row < m;
    const float* xr = X + (rv ? row : 0) * ldx;
    if (VEC && rv && col + 4 <= n) {
      v[i] = *reinterpret_cast<const f4*>(xr + col);
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) v[i][u] = (rv && col + u < n) ? xr[col + u] : 0.0f;
    }
  }
}

// shift, per-row scale (rexp), split, store A fragments: Fs[((R * KSW + ks) * 2 + part)]
template <int NB>
PC_DEV void rows_split(f4* v, int64_t m, int n, int center, const double* mean, const double* var, int64_t tile,
                       char* Fs, int* rexp) {
  using RT = RowTile<NB>;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int j = l % RT::LPR, col = 4 * j;
#pragma unroll
  for (int i = 0; i < RT::NLD; i++) {
    const int rt = w * RT::RPW + i * RT::RPI + l / RT::LPR;  // row in the tile
    const int64_t row = tile * RT::TR + rt;
    const bool rv = row < m;
    const float a = __shfl(v[i][0], (l / RT::LPR) * RT::LPR);  // x_i0 (lane holding column 0 of this row)
    float mx = 0.0f;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      float y = (rv && col + u < n) ? shifted(v[i][u], a, center, rv ? row : 0, mean, var) : 0.0f;
      v[i][u] = y;
      mx = fmaxf(mx, fabsf(y));
    }
    mx = group_max(mx, RT::LPR);
    const int e = mx > 0.0f ? scale_exp(mx) : 0;
    if (j == 0) rexp[rt] = e;
    const float sc = ldexpf(1.0f, e);
    _Float16 hi[4], lo[4];
#pragma unroll
    for (int u = 0; u < 4; u++) split(v[i][u] * sc, hi[u], lo[u]);
    // columns 4j..4j+3 = k-slots t0..t0+3 of k-step ks, half hl
    const int R = rt >> 5, cr = rt & 31, ks = j >> 2, hlw = (j >> 1) & 1, t0 = 4 * (j & 1);
    const int off = (cr + 32 * hlw) * 16 + t0 * 2;
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    h4 vh = {hi[0], hi[1], hi[2], hi[3]}, vl = {lo[0], lo[1], lo[2], lo[3]};
    *reinterpret_cast<h4*>(Fs + ((R * RT::KSW + ks) * 2 + 0) * RT::FRAGB + off) = vh;
    *reinterpret_cast<h4*>(Fs + ((R * RT::KSW + ks) * 2 + 1) * RT::FRAGB + off) = vl;
  }
}

PC_DEV h8 frag(const char* Fs, int idx, int fragb, int l) { return *reinterpret_cast<const h8*>(Fs + idx * fragb + l * 16); }

// ---------------------------------------------------------------------------
// P3: C = Ys W
// ---------------------------------------------------------------------------
template <int NB, int NCB, bool VEC>
__global__ __launch_bounds__(NT2, 1) void k_gemm_mx2(const float* __restrict__ X, int64_t ldx, int64_t m, int n,
                                                     int center, const double* __restrict__ mean,
                                                     const double* __restrict__ var, const h8* __restrict__ Wh,
                                                     const h8* __restrict__ Wl, const float* __restrict__ wscale,
                                                     int nc, int64_t tpc, float* __restrict__ C, int64_t ldc) {
  using RT = RowTile<NB>;
  constexpr int KSW = RT::KSW, NCBLK = NCB / 32;
  __shared__ __align__(16) char Fs[RT::FBYTES];
  __shared__ int rexp[RT::TR];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  const int J = w;
  h8 bh[KSW], bl[KSW];
  float isc = 1.0f;
  if (J < NCBLK) {
#pragma unroll
    for (int ks = 0; ks < KSW; ks++) {
      bh[ks] = Wh[int64_t(ks * NCBLK + J) * 64 + l];
      bl[ks] = Wl[int64_t(ks * NCBLK + J) * 64 + l];
    }
    isc = 1.0f / wscale[J * 32 + c];
  }
  const int64_t ntiles = ceil_div(m, RT::TR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  f4 v[RT::NLD];
  if (t0 < t1) rows_load<NB, VEC>(X, ldx, m, n, t0, v);
  for (int64_t tile = t0; tile < t1; tile++) {
    rows_split<NB>(v, m, n, center, mean, var, tile, Fs, rexp);
    __syncthreads();
    if (tile + 1 < t1) rows_load<NB, VEC>(X, ldx, m, n, tile + 1, v);
    if (J < NCBLK) {
      const int col = J * 32 + c;
#pragma unroll
      for (int R = 0; R < RT::NR; R++) {
        f16v acc = {};
#pragma unroll
        for (int ks = 0; ks < KSW; ks++) {
          h8 ah = frag(Fs, (R * KSW + ks) * 2 + 0, RT::FRAGB, l);
          h8 al = frag(Fs, (R * KSW + ks) * 2 + 1, RT::FRAGB, l);
          acc = mfma(ah, bh[ks], acc);
          acc = mfma(ah, bl[ks], acc);
          acc = mfma(al, bh[ks], acc);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int rt = 32 * R + 8 * (r >> 2) + 4 * hl + (r & 3);
          const int64_t row = tile * RT::TR + rt;
          if (row < m && col < nc) C[row * ldc + col] = ldexpf(acc[r], -rexp[rt]) * isc;
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// P2: Z = Ys W1 (fp16 W1, per-column scale), G2 = Z^T Z
// ---------------------------------------------------------------------------
template <int NB, bool VEC>
__global__ __launch_bounds__(NT2, 1) void k_zgram_mx2(const float* __restrict__ X, int64_t ldx, int64_t m, int n,
                                                      int center, const double* __restrict__ mean,
                                                      const double* __restrict__ var, const h8* __restrict__ Wf,
                                                      const float* __restrict__ wscale, int64_t tpc,
                                                      double* __restrict__ part) {
  using RT = RowTile<NB>;
  constexpr int KSW = RT::KSW, NBLK = NB / 32;
  constexpr int NU = NBLK * (NBLK + 1) / 2;
  constexpr int S = (NU + NW2 - 1) / NW2;
  constexpr int KSG = RT::TR / 16;  // Gram k-steps per tile
  __shared__ __align__(16) char Fs[RT::FBYTES];
  __shared__ h8 F2[KSG * NBLK * 2 * 64];
  __shared__ int rexp[RT::TR];
  __shared__ float red[NW2];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  const int J = w;
  h8 bw[KSW];
  float isc = 1.0f;
  if (J < NBLK) {
#pragma unroll
    for (int ks = 0; ks < KSW; ks++) bw[ks] = Wf[int64_t(ks * NBLK + J) * 64 + l];
    isc = 1.0f / wscale[J * 32 + c];
  }
  int sI[S], sJ[S];
  bool sv[S];
#pragma unroll
  for (int s = 0; s < S; s++) sv[s] = upper_slot<NBLK>(w * S + s, sI[s], sJ[s]);
  f16v acc[S];
#pragma unroll
  for (int s = 0; s < S; s++) acc[s] = f16v{};
  int e_run = EUN;
  const int64_t ntiles = ceil_div(m, RT::TR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  f4 v[RT::NLD];
  if (t0 < t1) rows_load<NB, VEC>(X, ldx, m, n, t0, v);
  for (int64_t tile = t0; tile < t1; tile++) {
    rows_split<NB>(v, m, n, center, mean, var, tile, Fs, rexp);
    __syncthreads();
    if (tile + 1 < t1) rows_load<NB, VEC>(X, ldx, m, n, tile + 1, v);
    f16v z[RT::NR];
    float zmx = 0.0f;
#pragma unroll
    for (int R = 0; R < RT::NR; R++) {
      z[R] = f16v{};
      if (J < NBLK) {
#pragma unroll
        for (int ks = 0; ks < KSW; ks++) {
          z[R] = mfma(frag(Fs, (R * KSW + ks) * 2 + 0, RT::FRAGB, l), bw[ks], z[R]);
          z[R] = mfma(frag(Fs, (R * KSW + ks) * 2 + 1, RT::FRAGB, l), bw[ks], z[R]);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int rt = 32 * R + 8 * (r >> 2) + 4 * hl + (r & 3);
          const float zz = ldexpf(z[R][r], -rexp[rt]) * isc;
          z[R][r] = zz;
          zmx = fmaxf(zmx, fabsf(zz));
        }
      }
    }
    zmx = group_max(zmx, 64);
    if (l == 0) red[w] = zmx;
    __syncthreads();
    float tmax = red[0];
#pragma unroll
    for (int q = 1; q < NW2; q++) tmax = fmaxf(tmax, red[q]);
    e_run = update_scale<S>(tmax, e_run, acc);
    const float sc = ldexpf(1.0f, e_run == EUN ? 0 : e_run);
    if (J < NBLK) {
      // D registers 0..7 / 8..15 of a 32-row block are two 16-row Gram k-steps
#pragma unroll
      for (int R = 0; R < RT::NR; R++)
#pragma unroll
        for (int half = 0; half < 2; half++) {
          h8 hi, lo;
#pragma unroll
          for (int t = 0; t < 8; t++) {
            _Float16 h, o;
            split(z[R][half * 8 + t] * sc, h, o);
            hi[t] = h;
            lo[t] = o;
          }
          const int ks = 2 * R + half;
          F2[((ks * NBLK + J) * 2 + 0) * 64 + l] = hi;
          F2[((ks * NBLK + J) * 2 + 1) * 64 + l] = lo;
        }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KSG; ks++) gram_step<S>(F2 + ks * NBLK * 2 * 64, sI, sJ, sv, acc, l);
    __syncthreads();
  }
  const double inv = ldexp(1.0, e_run == EUN ? 0 : -2 * e_run);
  write_upper<NB, S>(acc, sI, sJ, sv, inv, part + int64_t(blockIdx.x) * NB * NB, l,
                     [](int blk, int pos) { return blk * 32 + pos; });
}

template <class F>
int dispatch_nb(int n, F&& f) {
  if (n <= 32) return f(std::integral_constant<int, 32>{});
  if (n <= 64) return f(std::integral_constant<int, 64>{});
  if (n <= 128) return f(std::integral_constant<int, 128>{});
  if (n <= 256) return f(std::integral_constant<int, 256>{});
  return -1;
}

bool vec_ok(const float* X, int64_t ldx) { return (reinterpret_cast<uintptr_t>(X) & 15) == 0 && (ldx & 3) == 0; }

}  // namespace

int mx2_nb(int n) { return n <= 32 ? 32 : n <= 64 ? 64 : n <= 128 ? 128 : n <= 256 ? 256 : -1; }
int mx2_gram_rows() { return 16 * NW2; }
int mx2_tile_rows(int n) {
  return dispatch_nb(n, [](auto nbc) { return RowTile<decltype(nbc)::value>::TR; });
}

int launch_gram_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, double* mean_io,
                    const double* var, int chunks, int64_t tpc, double* part) {
  const bool vec = vec_ok(X, ldx);
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    if (vec)
      hipLaunchKernelGGL((k_gram_mx2<NB, true>), dim3(chunks), dim3(NT2), 0, st, X, ldx, m, n, center, mean_io, var,
                         tpc, part);
    else
      hipLaunchKernelGGL((k_gram_mx2<NB, false>), dim3(chunks), dim3(NT2), 0, st, X, ldx, m, n, center, mean_io,
                         var, tpc, part);
    return 0;
  });
}

int launch_zgram_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean,
                     const double* var, const _Float16* Wf, const float* wscale, int chunks, int64_t tpc,
                     double* part) {
  const bool vec = vec_ok(X, ldx);
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    if (vec)
      hipLaunchKernelGGL((k_zgram_mx2<NB, true>), dim3(chunks), dim3(NT2), 0, st, X, ldx, m, n, center, mean, var,
                         (const h8*)Wf, wscale, tpc, part);
    else
      hipLaunchKernelGGL((k_zgram_mx2<NB, false>), dim3(chunks), dim3(NT2), 0, st, X, ldx, m, n, center, mean, var,
                         (const h8*)Wf, wscale, tpc, part);
    return 0;
  });
}

int launch_gemm_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean,
                    const double* var, const _Float16* Wh, const _Float16* Wl, const float* wscale, int nc,
                    int chunks, int64_t tpc, float* C, int64_t ldc) {
  const int ncb = mx2_nb(nc);
  if (ncb < 0) return -1;
  const bool vec = vec_ok(X, ldx);
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    return dispatch_nb(ncb, [&](auto ncc) {
      constexpr int NCB = decltype(ncc)::value;
      if (vec)
        hipLaunchKernelGGL((k_gemm_mx2<NB, NCB, true>), dim3(chunks), dim3(NT2), 0, st, X, ldx, m, n, center, mean,
                           var, (const h8*)Wh, (const h8*)Wl, wscale, nc, tpc, C, ldc);
      else
        hipLaunchKernelGGL((k_gemm_mx2<NB, NCB, false>), dim3(chunks), dim3(NT2), 0, st, X, ldx, m, n, center, mean,
                           var, (const h8*)Wh, (const h8*)Wl, wscale, nc, tpc, C, ldc);
      return 0;
    });
  });
}
