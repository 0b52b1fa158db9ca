// Mixed-precision MFMA streaming kernels (v_mfma_f32_32x32x16_f16) of the
// float32 POD path.  Operands are float32 data split into two fp16 pieces,
// hi = fp16(x 2^e) and lo = fp16(x 2^e - hi) (22 significant bits), with
// power-of-two scales that are undone exactly:
//   gram  (P1): Y^T Y, products hi.hi + hi.lo + lo.hi, one running scale per
//               workgroup (accumulators rescaled when a larger tile arrives);
//   zgram (P2): Z = Y W1 (hi.W + lo.W, W1 already rounded to fp16 and used as
//               such), then Z^T Z with the three split products;
//   gemm  (P3): C = Y W with hi.hi + hi.lo + lo.hi; per-row scales of Y and
//               per-column scales of W.
// Accumulation is FP32 inside a workgroup's row chunk and FP64 across chunks.
// Row-local centring (temporal mean / variance) is fused into the loads.
//
// MFMA lane layouts (wave64, verified on gfx950), with hl = lane / 32, c = lane % 32:
//   A (32x16): lane holds A[c][8 hl + t],   t = 0..7
//   B (16x32): lane holds B[8 hl + t][c]
//   D (32x32): reg r holds D[8 (r / 4) + 4 hl + r % 4][c]
// The Gram operands only need A and B to agree on which physical row each
// k-slot stands for, so D registers 0..7 / 8..15 of a 32-row block serve
// directly as the two 16-row k-steps of the next Gram.
#include <climits>
#include "common.hpp"
#include "kernels.hpp"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
constexpr int NTH = 256;  // 4 waves
constexpr int TRM = 64;   // rows per tile
constexpr int EUNSET = 1 << 20;

PC_DEV f16v mfma16(h8 a, h8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }

// exponent e with |x| * 2^e < 2^14 for every |x| <= mx (mx > 0)
PC_DEV int scale_exp(float mx) {
  int E;
  frexpf(mx, &E);  // mx = f 2^E, f in [0.5, 1)
  return 14 - E;
}

PC_DEV void split8(const float* v, float sc, h8& hi, h8& lo) {
#pragma unroll
  for (int t = 0; t < 8; t++) {
    float x = v[t] * sc;
    _Float16 h = (_Float16)x;
    hi[t] = h;
    lo[t] = (_Float16)(x - (float)h);
  }
}

PC_DEV float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Row-pair assignment of the upper 32x32 blocks of an nb x nb Gram to 4 waves:
// wave p owns block rows p and nb-1-p (when distinct), i.e. nb+1 blocks.
struct Slots {
  int I1, I2, cnt1, nslot;
  PC_DEV void init(int w, int nb) {
    int npair = (nb + 1) / 2;
    if (w < npair) {
      I1 = w;
      I2 = nb - 1 - w;
      cnt1 = nb - I1;
      nslot = cnt1 + (I2 != I1 ? nb - I2 : 0);
    } else {
      I1 = I2 = 0;
      cnt1 = nslot = 0;
    }
  }
  PC_DEV void ij(int s, int nb, int& I, int& J) const {
    if (s < cnt1) { I = I1; J = I1 + s; }
    else { I = I2; J = I2 + (s - cnt1); }
  }
};

// Gram accumulation over KS k-steps of fragments F[(ks * nb + I) * 2 + part][lane]
template <int nb, int SL, int KS>
PC_DEV void gram_frags(const h8* F, const Slots& sl, f16v* acc, int l) {
  if (sl.nslot == 0) return;
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    h8 a1h = F[((ks * nb + sl.I1) * 2 + 0) * 64 + l];
    h8 a1l = F[((ks * nb + sl.I1) * 2 + 1) * 64 + l];
    h8 a2h = F[((ks * nb + sl.I2) * 2 + 0) * 64 + l];
    h8 a2l = F[((ks * nb + sl.I2) * 2 + 1) * 64 + l];
#pragma unroll
    for (int s = 0; s < SL; s++) {
      if (s < sl.nslot) {
        int I, J;
        sl.ij(s, nb, I, J);
        bool first = s < sl.cnt1;
        h8 ah = first ? a1h : a2h;
        h8 al = first ? a1l : a2l;
        h8 bh = F[((ks * nb + J) * 2 + 0) * 64 + l];
        h8 bl = F[((ks * nb + J) * 2 + 1) * 64 + l];
        acc[s] = mfma16(ah, bh, acc[s]);
        acc[s] = mfma16(ah, bl, acc[s]);
        acc[s] = mfma16(al, bh, acc[s]);
      }
    }
  }
}

template <int nb, int SL>
PC_DEV void write_slab(const f16v* acc, const Slots& sl, double inv, double* P, int l) {
  constexpr int NB = nb * 32;
  int hl = l >> 5, c = l & 31;
#pragma unroll
  for (int s = 0; s < SL; s++) {
    if (s < sl.nslot) {
      int I, J;
      sl.ij(s, nb, I, J);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        int row = I * 32 + 8 * (r >> 2) + 4 * hl + (r & 3);
        P[int64_t(row) * NB + J * 32 + c] = double(acc[s][r]) * inv;
      }
    }
  }
}

// WG-uniform running scale update; returns the new exponent and rescales acc.
template <int SL>
PC_DEV int update_scale(float tmax, int e_run, f16v* acc) {
  if (!(tmax > 0.0f)) return e_run;
  int et = scale_exp(tmax);
  if (e_run == EUNSET) return et;
  if (et < e_run) {
    float f = ldexpf(1.0f, 2 * (et - e_run));
#pragma unroll
    for (int s = 0; s < SL; s++) acc[s] *= f;
    return et;
  }
  return e_run;
}

// ---------------------------------------------------------------------------
// P1: G = Yc^T Yc, fused row centring; writes mean / var when asked.
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(NTH) void k_gram_mx(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                                                 double* __restrict__ mean_out, double* __restrict__ var_out,
                                                 int64_t tpc, double* __restrict__ part) {
  constexpr int nb = NB / 32;
  constexpr int KS = TRM / 16;  // 4 k-steps: wave w loads k-step w
  constexpr int SL = nb + 1;
  __shared__ h8 F[KS * nb * 2 * 64];
  __shared__ float red[4];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  Slots sl;
  sl.init(w, nb);
  f16v acc[SL];
#pragma unroll
  for (int s = 0; s < SL; s++) acc[s] = f16v{};
  int e_run = EUNSET;
  int64_t ntiles = ceil_div(m, TRM);
  int64_t t0 = int64_t(blockIdx.x) * tpc;
  int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    int64_t rbase = t * TRM + w * 16 + 8 * hl;
    float v[nb][8];
#pragma unroll
    for (int tt = 0; tt < 8; tt++) {
      int64_t row = rbase + tt;
      bool rv = row < m;
      const float* xr = X + (rv ? row : 0) * ldx;
#pragma unroll
      for (int I = 0; I < nb; I++) {
        int col = I * 32 + c;
        v[I][tt] = (rv && col < n) ? xr[col] : 0.0f;
      }
    }
    if (center) {
#pragma unroll
      for (int tt = 0; tt < 8; tt++) {
        double s = 0.0;
#pragma unroll
        for (int I = 0; I < nb; I++) s += double(v[I][tt]);
        for (int o = 16; o > 0; o >>= 1) s += __shfl_xor(s, o);
        double mu = s / double(n);
        double var = 1.0;
        if (center == 2) {
          double s2 = 0.0;
#pragma unroll
          for (int I = 0; I < nb; I++) {
            double d = (I * 32 + c < n) ? double(v[I][tt]) - mu : 0.0;
            s2 += d * d;
          }
          for (int o = 16; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
          var = s2 / double(n);
        }
        int64_t row = rbase + tt;
        if (row < m && c == 0) {
          if (mean_out) mean_out[row] = mu;
          if (var_out && center == 2) var_out[row] = var;
        }
        bool rv = row < m;
#pragma unroll
        for (int I = 0; I < nb; I++) {
          int col = I * 32 + c;
          double d = double(v[I][tt]) - mu;
          if (center == 2) d /= var;
          v[I][tt] = (rv && col < n) ? float(d) : 0.0f;
        }
      }
    }
    float mx = 0.0f;
#pragma unroll
    for (int I = 0; I < nb; I++)
#pragma unroll
      for (int tt = 0; tt < 8; tt++) mx = fmaxf(mx, fabsf(v[I][tt]));
    mx = wave_max(mx);
    if (l == 0) red[w] = mx;
    __syncthreads();
    float tmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    e_run = update_scale<SL>(tmax, e_run, acc);
    float sc = ldexpf(1.0f, e_run == EUNSET ? 0 : e_run);
#pragma unroll
    for (int I = 0; I < nb; I++) {
      h8 hi, lo;
      split8(v[I], sc, hi, lo);
      F[((w * nb + I) * 2 + 0) * 64 + l] = hi;
      F[((w * nb + I) * 2 + 1) * 64 + l] = lo;
    }
    __syncthreads();
    gram_frags<nb, SL, KS>(F, sl, acc, l);
    __syncthreads();
  }
  double inv = ldexp(1.0, e_run == EUNSET ? 0 : -2 * e_run);
  write_slab<nb, SL>(acc, sl, inv, part + int64_t(blockIdx.x) * NB * NB, l);
}

// ---------------------------------------------------------------------------
// load a 64-row tile of centred Y as A-operand fragments with per-row scales:
// wave w loads row block R = w / 2, k-steps [(w & 1) KSW/2, (w & 1 + 1) KSW/2).
// F[((R * KSW + ks) * 2 + part) * 64 + lane]; rexp[64] receives row exponents.
// ---------------------------------------------------------------------------
template <int NB>
PC_DEV void load_rows_split(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                            const double* __restrict__ mean_in, const double* __restrict__ var_in, int64_t row0,
                            h8* F, int* rmaxi, int* rexp) {
  constexpr int KSW = NB / 16;
  constexpr int KH = KSW / 2 > 0 ? KSW / 2 : 1;
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  int R = w >> 1;
  int kbeg = (w & 1) * KH;
  int64_t row = row0 + R * 32 + c;
  bool rv = row < m;
  float mu = 0.0f, ivar = 1.0f;
  double mud = 0.0, vard = 1.0;
  if (center && rv) {
    mud = mean_in[row];
    if (center == 2) vard = var_in[row];
  }
  (void)mu; (void)ivar;
  const float* xr = X + (rv ? row : 0) * ldx;
  float v[KH][8];
  float mx = 0.0f;
#pragma unroll
  for (int j = 0; j < KH; j++) {
    int ks = kbeg + j;
#pragma unroll
    for (int tt = 0; tt < 8; tt++) {
      int col = ks * 16 + 8 * hl + tt;
      float x = 0.0f;
      if (rv && col < n && ks < KSW) {
        double d = double(xr[col]);
        if (center) { d -= mud; if (center == 2) d /= vard; }
        x = float(d);
      }
      v[j][tt] = x;
      mx = fmaxf(mx, fabsf(x));
    }
  }
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  if (hl == 0) atomicMax(&rmaxi[R * 32 + c], __float_as_int(mx));
  __syncthreads();
  float rm = __int_as_float(rmaxi[R * 32 + c]);
  int e = rm > 0.0f ? scale_exp(rm) : 0;
  if (hl == 0 && (w & 1) == 0) rexp[R * 32 + c] = e;
  float sc = ldexpf(1.0f, e);
#pragma unroll
  for (int j = 0; j < KH; j++) {
    int ks = kbeg + j;
    if (ks < KSW) {
      h8 hi, lo;
      split8(v[j], sc, hi, lo);
      F[((R * KSW + ks) * 2 + 0) * 64 + l] = hi;
      F[((R * KSW + ks) * 2 + 1) * 64 + l] = lo;
    }
  }
}

// C block pair (rows R = 0, 1 of the tile; column block J) = Ysplit * W (B frags)
template <int NB, bool WLO>
PC_DEV void gemm_block(const h8* F, const h8* __restrict__ Wh, const h8* __restrict__ Wl, int ncb, int J, f16v& a0,
                       f16v& a1, int l) {
  constexpr int KSW = NB / 16;
  a0 = f16v{};
  a1 = f16v{};
#pragma unroll 4
  for (int ks = 0; ks < KSW; ks++) {
    h8 bh = Wh[int64_t(ks * ncb + J) * 64 + l];
    h8 y0h = F[((0 * KSW + ks) * 2 + 0) * 64 + l];
    h8 y0l = F[((0 * KSW + ks) * 2 + 1) * 64 + l];
    h8 y1h = F[((1 * KSW + ks) * 2 + 0) * 64 + l];
    h8 y1l = F[((1 * KSW + ks) * 2 + 1) * 64 + l];
    a0 = mfma16(y0h, bh, a0);
    a1 = mfma16(y1h, bh, a1);
    a0 = mfma16(y0l, bh, a0);
    a1 = mfma16(y1l, bh, a1);
    if constexpr (WLO) {
      h8 bl = Wl[int64_t(ks * ncb + J) * 64 + l];
      a0 = mfma16(y0h, bl, a0);
      a1 = mfma16(y1h, bl, a1);
    }
  }
}

// ---------------------------------------------------------------------------
// P2: Z = Yc W1 (W1 fp16, per-column scale), G2 = Z^T Z
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(NTH) void k_zgram_mx(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                                                  const double* __restrict__ mean_in, const double* __restrict__ var_in,
                                                  const h8* __restrict__ Wf, const float* __restrict__ wscale,
                                                  int64_t tpc, double* __restrict__ part) {
  constexpr int nb = NB / 32;
  constexpr int KSW = NB / 16;
  constexpr int KS = TRM / 16;  // Gram k-steps per tile
  constexpr int SL = nb + 1;
  constexpr int JW = (nb + 3) / 4;  // column blocks per wave
  constexpr int FSZ = (2 * KSW * 2 > KS * nb * 2 ? 2 * KSW * 2 : KS * nb * 2) * 64;
  __shared__ h8 F[FSZ];
  __shared__ int rmaxi[64];
  __shared__ int rexp[64];
  __shared__ float red[4];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  Slots sl;
  sl.init(w, nb);
  f16v acc[SL];
#pragma unroll
  for (int s = 0; s < SL; s++) acc[s] = f16v{};
  int e_run = EUNSET;
  int64_t ntiles = ceil_div(m, TRM);
  int64_t t0 = int64_t(blockIdx.x) * tpc;
  int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    if (threadIdx.x < 64) rmaxi[threadIdx.x] = 0;
    __syncthreads();
    load_rows_split<NB>(X, ldx, m, n, center, mean_in, var_in, t * TRM, F, rmaxi, rexp);
    __syncthreads();
    // Z blocks of this wave, kept in registers until the tile scale is known
    f16v z[JW][2];
    float zmx = 0.0f;
#pragma unroll
    for (int jj = 0; jj < JW; jj++) {
      int J = w + 4 * jj;
      if (J < nb) {
        gemm_block<NB, false>(F, Wf, nullptr, nb, J, z[jj][0], z[jj][1], l);
        float isc = 1.0f / wscale[J * 32 + c];
#pragma unroll
        for (int R = 0; R < 2; R++)
#pragma unroll
          for (int r = 0; r < 16; r++) {
            int rowD = R * 32 + 8 * (r >> 2) + 4 * hl + (r & 3);
            float zz = ldexpf(z[jj][R][r], -rexp[rowD]) * isc;
            z[jj][R][r] = zz;
            zmx = fmaxf(zmx, fabsf(zz));
          }
      } else {
        z[jj][0] = f16v{};
        z[jj][1] = f16v{};
      }
    }
    zmx = wave_max(zmx);
    if (l == 0) red[w] = zmx;
    __syncthreads();  // also: every wave is done reading the Y fragments
    float tmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    e_run = update_scale<SL>(tmax, e_run, acc);
    float sc = ldexpf(1.0f, e_run == EUNSET ? 0 : e_run);
#pragma unroll
    for (int jj = 0; jj < JW; jj++) {
      int J = w + 4 * jj;
      if (J < nb) {
#pragma unroll
        for (int R = 0; R < 2; R++)
#pragma unroll
          for (int half = 0; half < 2; half++) {
            float vv[8];
#pragma unroll
            for (int tt = 0; tt < 8; tt++) vv[tt] = z[jj][R][half * 8 + tt];
            h8 hi, lo;
            split8(vv, sc, hi, lo);
            int ks = R * 2 + half;
            F[((ks * nb + J) * 2 + 0) * 64 + l] = hi;
            F[((ks * nb + J) * 2 + 1) * 64 + l] = lo;
          }
      }
    }
    __syncthreads();
    gram_frags<nb, SL, KS>(F, sl, acc, l);
    __syncthreads();
  }
  double inv = ldexp(1.0, e_run == EUNSET ? 0 : -2 * e_run);
  write_slab<nb, SL>(acc, sl, inv, part + int64_t(blockIdx.x) * NB * NB, l);
}

// ---------------------------------------------------------------------------
// P3: C = Yc W, W = Wh + Wl (fp16 pieces, per-column scale), C float32
// ---------------------------------------------------------------------------
template <int NB, int NCB>
__global__ __launch_bounds__(NTH) void k_gemm_mx(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int center,
                                                 const double* __restrict__ mean_in, const double* __restrict__ var_in,
                                                 const h8* __restrict__ Wh, const h8* __restrict__ Wl,
                                                 const float* __restrict__ wscale, int nc, int64_t tpc,
                                                 float* __restrict__ C, int64_t ldc) {
  constexpr int KSW = NB / 16;
  constexpr int ncb = NCB / 32;
  __shared__ h8 F[2 * KSW * 2 * 64];
  __shared__ int rmaxi[64];
  __shared__ int rexp[64];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  int64_t ntiles = ceil_div(m, TRM);
  int64_t t0 = int64_t(blockIdx.x) * tpc;
  int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    if (threadIdx.x < 64) rmaxi[threadIdx.x] = 0;
    __syncthreads();
    load_rows_split<NB>(X, ldx, m, n, center, mean_in, var_in, t * TRM, F, rmaxi, rexp);
    __syncthreads();
    for (int J = w; J < ncb; J += 4) {
      f16v a0, a1;
      gemm_block<NB, true>(F, Wh, Wl, ncb, J, a0, a1, l);
      int col = J * 32 + c;
      float isc = 1.0f / wscale[col];
#pragma unroll
      for (int R = 0; R < 2; R++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
          int rowD = R * 32 + 8 * (r >> 2) + 4 * hl + (r & 3);
          int64_t row = t * TRM + rowD;
          float val = ldexpf(R == 0 ? a0[r] : a1[r], -rexp[rowD]) * isc;
          if (row < m && col < nc) C[row * ldc + col] = val;
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// W (FP64 NB x NCB, ld) -> B fragments (hi, lo) with per-column exponents;
// Wr = hi * 2^-f (the operand the hi-only products apply).  One workgroup
// per 32-column block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack_w(const double* __restrict__ W, int ldw, int NB, int NCB, h8* Wh, h8* Wl,
                                                float* wscale, double* Wr, int ldwr) {
  int J = blockIdx.x;
  int ncb = NCB / 32;
  __shared__ float fsc[32];
  // column maxima
  if (threadIdx.x < 32) {
    int col = J * 32 + threadIdx.x;
    double mx = 0.0;
    for (int i = 0; i < NB; i++) mx = fmax(mx, fabs(W[int64_t(i) * ldw + col]));
    int e = mx > 0.0 ? scale_exp(float(mx)) : 0;
    if (mx > 0.0 && float(mx) * ldexpf(1.0f, e) >= 16384.0f) e -= 1;  // rounding of mx to float
    e = e > 120 ? 120 : (e < -120 ? -120 : e);
    float s = ldexpf(1.0f, e);
    fsc[threadIdx.x] = s;
    wscale[col] = s;
  }
  __syncthreads();
  int KSW = NB / 16;
  for (int idx = threadIdx.x; idx < KSW * 64; idx += 256) {
    int ks = idx / 64, l = idx % 64;
    int hl = l >> 5, c = l & 31;
    int col = J * 32 + c;
    float s = fsc[c];
    h8 hi, lo;
    for (int t = 0; t < 8; t++) {
      int k = ks * 16 + 8 * hl + t;
      double x = W[int64_t(k) * ldw + col] * double(s);
      _Float16 h = (_Float16)float(x);
      hi[t] = h;
      lo[t] = (_Float16)float(x - double(float(h)));
      if (Wr) Wr[int64_t(k) * ldwr + col] = double(float(h)) / double(s);
    }
    Wh[int64_t(ks * ncb + J) * 64 + l] = hi;
    if (Wl) Wl[int64_t(ks * ncb + J) * 64 + l] = lo;
  }
}

template <class F>
int dispatch_mx(int n, F&& f) {
  if (n <= 32) return f(std::integral_constant<int, 32>{});
  if (n <= 64) return f(std::integral_constant<int, 64>{});
  if (n <= 128) return f(std::integral_constant<int, 128>{});
  if (n <= 256) return f(std::integral_constant<int, 256>{});
  return -1;
}

}  // namespace

int mx_nb(int n) { return n <= 32 ? 32 : n <= 64 ? 64 : n <= 128 ? 128 : n <= 256 ? 256 : -1; }
int mx_gram_rows() { return TRM; }
int mx_gemm_rows() { return TRM; }

int launch_gram_mx(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, double* mean_out,
                   double* var_out, int chunks, int64_t tpc, double* part) {
  return dispatch_mx(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    hipLaunchKernelGGL(k_gram_mx<NB>, dim3(chunks), dim3(NTH), 0, st, X, ldx, m, n, center, mean_out, var_out, tpc,
                       part);
    return 0;
  });
}

int launch_zgram_mx(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean_in,
                    const double* var_in, const _Float16* Wf, const float* wscale, int chunks, int64_t tpc,
                    double* part) {
  return dispatch_mx(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    hipLaunchKernelGGL(k_zgram_mx<NB>, dim3(chunks), dim3(NTH), 0, st, X, ldx, m, n, center, mean_in, var_in,
                       (const h8*)Wf, wscale, tpc, part);
    return 0;
  });
}

int launch_gemm_mx(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean_in,
                   const double* var_in, const _Float16* Wh, const _Float16* Wl, const float* wscale, int nc,
                   int chunks, int64_t tpc, float* C, int64_t ldc) {
  int ncb = mx_nb(nc);
  if (ncb < 0) return -1;
  return dispatch_mx(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    return dispatch_mx(ncb, [&](auto ncc) {
      constexpr int NCB = decltype(ncc)::value;
      hipLaunchKernelGGL((k_gemm_mx<NB, NCB>), dim3(chunks), dim3(NTH), 0, st, X, ldx, m, n, center, mean_in, var_in,
                         (const h8*)Wh, (const h8*)Wl, wscale, nc, tpc, C, ldc);
      return 0;
    });
  });
}

int launch_pack_w_mx(hipStream_t st, const double* W, int ldw, int n, int nc, int NB, int NCB, _Float16* Wh,
                     _Float16* Wl, float* wscale, double* Wr, int ldwr) {
  (void)n; (void)nc;
  hipLaunchKernelGGL(k_pack_w, dim3(NCB / 32), dim3(256), 0, st, W, ldw, NB, NCB, (h8*)Wh, (h8*)Wl, wscale, Wr, ldwr);
  return 0;
}
