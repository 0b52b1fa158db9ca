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
// In P2/P3 every load instruction reads whole rows (four columns per lane: one
// float4, two float2 or four floats, by the alignment of the rows); the fp16
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

// Workgroup barrier for LDS traffic only.  __syncthreads() also fences global
// memory, which on this ISA waits for every outstanding global load and store
// (vmcnt(0)) and would serialise the prefetch of the next tile with the
// products of the current one.  The asm "memory" clobber keeps the compiler
// from moving LDS accesses across the barrier.
PC_DEV void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Raw buffer access (range-checked by the hardware: loads beyond num_records
// return 0, stores beyond it are dropped), so edge tiles need no branches.
// 0x00020000: dword 3 of the gfx9 buffer descriptor (32-bit data format).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
PC_DEV rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

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
// maximum of v over the 64 lanes of the wave, returned as a wave-uniform
// value: DPP steps within each row of 16 lanes (quad swaps, half-row and row
// mirrors), the two row broadcasts that carry the running maximum into lane
// 63, then a read of lane 63
PC_DEV float wave_max64(float v) {
  auto dpp = [](float x, auto ctrl, auto rows) {
    const int xi = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(xi, xi, decltype(ctrl)::value,
                                                                  decltype(rows)::value, 0xF, false));
  };
  v = fmaxf(v, dpp(v, std::integral_constant<int, 0xB1>{}, std::integral_constant<int, 0xF>{}));   // quad [1,0,3,2]
  v = fmaxf(v, dpp(v, std::integral_constant<int, 0x4E>{}, std::integral_constant<int, 0xF>{}));   // quad [2,3,0,1]
  v = fmaxf(v, dpp(v, std::integral_constant<int, 0x141>{}, std::integral_constant<int, 0xF>{}));  // row_half_mirror
  v = fmaxf(v, dpp(v, std::integral_constant<int, 0x140>{}, std::integral_constant<int, 0xF>{}));  // row_mirror
  v = fmaxf(v, dpp(v, std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{}));  // row_bcast:15
  v = fmaxf(v, dpp(v, std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{}));  // row_bcast:31
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
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

// Contiguous range [start, start + cnt) of upper-block indices that wave w
// accumulates.  Waves 2k and 2k + 1 share a SIMD, so the NU % NW2 blocks
// beyond an even share go first to the even waves (one each), then to the
// odd ones: every SIMD's pair then holds NU / 4 blocks, give or take one.
// Which wave accumulates a block does not change its products.
template <int NU>
PC_DEV void wave_slots(int w, int& start, int& cnt) {
  constexpr int base = NU / NW2, rem = NU % NW2;
  auto extra = [](int v) { return ((v & 1) ? NW2 / 2 + (v >> 1) : (v >> 1)) < rem ? 1 : 0; };
  start = w * base;
  for (int v = 0; v < w; v++) start += extra(v);
  cnt = base + extra(w);
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

// shifted value of x (see the file comment): a = x_i0 for CENTER 1; for
// CENTER 2, mu / iv are the row mean and 1 / variance (FP64)
template <int CENTER>
PC_DEV float shifted(float x, float a, double mu, double iv) {
  if constexpr (CENTER == 1) return x - a;
  else if constexpr (CENTER == 2) return float((double(x) - mu) * iv);
  else return x;
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
template <int NB, int VEC, int CENTER>
__global__ __launch_bounds__(NT2, 1) void k_gram_mx2(const float* __restrict__ X, int64_t ldx, int64_t m, int n,
                                                     double* __restrict__ mean_io,
                                                     const double* __restrict__ var, int64_t tpc,
                                                     double* __restrict__ part, int64_t zx = 0, int64_t zb = 0) {
  X = zoff(X, zx);
  mean_io = zoff(mean_io, zb);
  var = zoff(var, zb);
  part = zoff(part, zb);
  constexpr int V = NB >= 128 ? 4 : NB / 32;
  constexpr int NG = NB / (32 * V);  // column groups of 32 V columns
  constexpr int NBLK = NB / 32;
  constexpr int KS = NW2 / NG;       // k-steps per tile: wave w loads k-step w % KS of group w / KS
  constexpr int TR = 16 * KS;
  constexpr int NU = NBLK * (NBLK + 1) / 2;
  constexpr int S = (NU + NW2 - 1) / NW2;
  __shared__ h8 F[KS * NBLK * 2 * 64];
  __shared__ float red[NW2];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  const int wk = w % KS, wg = w / KS;
  int sI[S], sJ[S];
  bool sv[S];
  {
    int s0, sc;
    wave_slots<NU>(w, s0, sc);
#pragma unroll
    for (int s = 0; s < S; s++) sv[s] = upper_slot<NBLK>(s0 + s, sI[s], sJ[s]) && s < sc;
  }
  f16v acc[S];
#pragma unroll
  for (int s = 0; s < S; s++) acc[s] = f16v{};
  int e_run = EUN;
  const int64_t ntiles = ceil_div(m, TR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  float v[8][V];
  float a0[8];  // x_i0 of the 8 rows (CENTER 1)
  // branch-free loads through a per-tile buffer resource (rows >= m read 0; the
  // columns >= n are masked where the values are used, so no instruction waits
  // on the loads until the tile is processed)
  auto load = [&](int64_t tile) {
    const int64_t row0 = tile * TR;
    const int64_t nrow = m - row0 < TR ? m - row0 : TR;
    const rsrc_t rs = make_rsrc(X + row0 * ldx, uint32_t(((nrow - 1) * ldx + n) * 4));
    const int col = wg * 32 * V + V * c;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int rt = 16 * wk + 8 * hl + t;
      if constexpr (VEC == 1 && V == 4) {
        const uint32_t off = uint32_t((rt * ldx + (col < n ? col : 0)) * 4);
        f4 q = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
#pragma unroll
        for (int u = 0; u < 4; u++) v[t][u] = q[u];
      } else {
#pragma unroll
        for (int u = 0; u < V; u++)
          v[t][u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                  rs, uint32_t((rt * ldx + (col + u < n ? col + u : 0)) * 4), 0, 0));
      }
      if constexpr (CENTER == 1)
        a0[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, uint32_t(rt * ldx * 4), 0, 0));
    }
  };
  if (t0 < t1) load(t0);
  for (int64_t tile = t0; tile < t1; tile++) {
    // shift and tile maximum
    float mx = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int64_t row = tile * TR + 16 * wk + 8 * hl + t;
      const bool rv = row < m;
      double mu = 0.0, iv = 1.0;
      if constexpr (CENTER == 2) {
        mu = mean_io[rv ? row : 0];
        iv = 1.0 / var[rv ? row : 0];
      }
      const float a = CENTER == 1 ? a0[t] : 0.0f;
#pragma unroll
      for (int u = 0; u < V; u++) {
        const int col = wg * 32 * V + V * c + u;
        float y = (rv && col < n) ? shifted<CENTER>(v[t][u], a, mu, iv) : 0.0f;
        v[t][u] = y;
        mx = fmaxf(mx, fabsf(y));
      }
    }
    mx = group_max(mx, 64);
    if (l == 0) red[w] = mx;
    lds_sync();
    float tmax = red[0];
#pragma unroll
    for (int q = 1; q < NW2; q++) tmax = fmaxf(tmax, red[q]);
    e_run = update_scale<S>(tmax, e_run, acc);
    const float sc = ldexpf(1.0f, e_run == EUN ? 0 : e_run);
#pragma unroll
    for (int u = 0; u < V; u++) {
      h8 hi, lo;
#pragma unroll
      for (int t = 0; t < 8; t++) {
        _Float16 h, o;
        split(v[t][u] * sc, h, o);
        hi[t] = h;
        lo[t] = o;
      }
      const int blk = wg * V + u;
      F[((wk * NBLK + blk) * 2 + 0) * 64 + l] = hi;
      F[((wk * NBLK + blk) * 2 + 1) * 64 + l] = lo;
    }
    lds_sync();
    load(tile + 1 < t1 ? tile + 1 : t1 - 1);  // in flight during the products (clamped: straight-line code)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) gram_step<S>(F + ks * NBLK * 2 * 64, sI, sJ, sv, acc, l);
    lds_sync();
  }
  const double inv = ldexp(1.0, e_run == EUN ? 0 : -2 * e_run);
  write_upper<NB, S>(acc, sI, sJ, sv, inv, part + int64_t(blockIdx.x) * NB * NB, l,
                     [](int blk, int pos) { return (blk / V) * 32 * V + V * pos + (blk % V); });
}

// ---------------------------------------------------------------------------
// P2 / P3 row tiles: 32 rows (64 for n <= 32), each wave loads whole rows, four
// columns per lane
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

// rows of one tile through a per-tile buffer resource: rows >= m read 0; the
// columns >= n are masked in rows_split
template <int NB, int VEC>
PC_DEV void rows_load(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int64_t tile, f4* v) {
  using RT = RowTile<NB>;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int col = 4 * (l % RT::LPR);
  const int64_t row0 = tile * RT::TR;
  const int64_t nrow = m - row0 < RT::TR ? m - row0 : RT::TR;
  const rsrc_t rs = make_rsrc(X + row0 * ldx, uint32_t(((nrow - 1) * ldx + n) * 4));
#pragma unroll
  for (int i = 0; i < RT::NLD; i++) {
    const int rt = w * RT::RPW + i * RT::RPI + l / RT::LPR;
    const uint32_t off = uint32_t((rt * ldx + (col < n ? col : 0)) * 4);
    if constexpr (VEC == 1) {
      typedef int i4 __attribute__((ext_vector_type(4)));
      i4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
      v[i] = __builtin_bit_cast(f4, q);
    } else if constexpr (VEC == 2) {
      // two 8-byte global loads (columns col, col + 1 and col + 2, col + 3),
      // predicated on the row: n is even, so a pair is wholly inside or wholly
      // outside the valid columns, and a pair outside reads column 0 (masked
      // in rows_split); rows >= m read 0 as through the buffer resource
      const float* rp = X + (row0 + rt) * ldx;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int c0 = col + 2 * h;
        float2 q = make_float2(0.0f, 0.0f);
        if (rt < nrow) q = *reinterpret_cast<const float2*>(rp + (c0 < n ? c0 : 0));
        v[i][2 * h] = q.x;
        v[i][2 * h + 1] = q.y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++)
        v[i][u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                rs, uint32_t((rt * ldx + (col + u < n ? col + u : 0)) * 4), 0, 0));
    }
  }
}

template <int NB, int CENTER>
PC_DEV void split_row(f4& vi, int i, int64_t m, int n, const double* mean, const double* var, int64_t tile, char* Fs,
                      int* rexp) {
  using RT = RowTile<NB>;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int j = l % RT::LPR, col = 4 * j;
  const int rt = w * RT::RPW + i * RT::RPI + l / RT::LPR;  // row in the tile
  const int64_t row = tile * RT::TR + rt;
  const bool rv = row < m;
  // x_i0 (lane holding column 0 of this row; lane 0 when a row fills the wave)
  float a;
  if constexpr (RT::LPR == 64) a = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vi[0]), 0));
  else a = __shfl(vi[0], (l / RT::LPR) * RT::LPR);
  double mu = 0.0, iv = 1.0;
  if constexpr (CENTER == 2) {
    mu = mean[rv ? row : 0];
    iv = 1.0 / var[rv ? row : 0];
  }
  float mx = 0.0f;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    float y = (rv && col + u < n) ? shifted<CENTER>(vi[u], a, mu, iv) : 0.0f;
    vi[u] = y;
    mx = fmaxf(mx, fabsf(y));
  }
  if constexpr (RT::LPR == 64) mx = wave_max64(mx);
  else mx = group_max(mx, RT::LPR);
  const int e = mx > 0.0f ? scale_exp(mx) : 0;
  rexp[rt] = e;  // every lane of the row holds the same e (no branch)
  const float sc = ldexpf(1.0f, e);
  _Float16 hi[4], lo[4];
#pragma unroll
  for (int u = 0; u < 4; u++) split(vi[u] * sc, hi[u], lo[u]);
  // columns 4j..4j+3 = k-slots t0..t0+3 of k-step ks, half hl
  const int R = rt >> 5, cr = rt & 31, ks = j >> 2, hlw = (j >> 1) & 1, t0 = 4 * (j & 1);
  const int off = (cr + 32 * hlw) * 16 + t0 * 2;
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  h4 vh = {hi[0], hi[1], hi[2], hi[3]}, vl = {lo[0], lo[1], lo[2], lo[3]};
  *reinterpret_cast<h4*>(Fs + ((R * RT::KSW + ks) * 2 + 0) * RT::FRAGB + off) = vh;
  *reinterpret_cast<h4*>(Fs + ((R * RT::KSW + ks) * 2 + 1) * RT::FRAGB + off) = vl;
}

// shift, per-row scale (rexp), split, store A fragments: Fs[((R * KSW + ks) * 2 + part)]
template <int NB, int CENTER>
PC_DEV void rows_split(f4* v, int64_t m, int n, const double* mean, const double* var, int64_t tile, char* Fs,
                       int* rexp) {
#pragma unroll
  for (int i = 0; i < RowTile<NB>::NLD; i++) split_row<NB, CENTER>(v[i], i, m, n, mean, var, tile, Fs, rexp);
}

PC_DEV h8 frag(const char* Fs, int idx, int fragb, int l) { return *reinterpret_cast<const h8*>(Fs + idx * fragb + l * 16); }

// LDS-staged row tiles (NB = 256, float4 rows).  buffer_load_dwordx4 ... lds
// moves 16 bytes per lane from the buffer into LDS at M0 + 16 * lane, so one
// instruction per row lays the tile out in row-major order.  The loads are
// issued from inline asm: the compiler then neither counts them nor waits for
// them (its own waits before LDS accesses would drain the pipeline), so the
// kernel waits itself with s_waitcnt vmcnt(N), N counting the vector memory
// operations issued after the ones waited for (they complete in issue
// order).  No compiler-visible vector memory load may be issued between a
// stage load and its wait.  M0 is saved and restored around the load.
PC_DEV void lds_dma16(rsrc_t rs, uint32_t voff, uint32_t lds_addr) {
  uint32_t save;
  asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\tbuffer_load_dwordx4 %2, %3, 0 offen lds\n\ts_mov_b32 m0, %0"
               : "=&s"(save)
               : "s"(lds_addr), "v"(voff), "s"(rs)
               : "memory");
}
template <int N>
PC_DEV void wait_vm() {
  static_assert(N >= 0 && N <= 63, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// A wave loads and later reads only its own RPW rows, so a stage needs no
// barrier, only the wave's own wait.  Rows >= m read 0 through the buffer
// resource, like rows_load.
template <int NB>
PC_DEV void rows_stage(const float* __restrict__ X, int64_t ldx, int64_t m, int n, int64_t tile, char* stage) {
  using RT = RowTile<NB>;
  static_assert(RT::LPR == 64 && RT::RPI == 1, "one row per load instruction");
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), l = threadIdx.x & 63;
  const int col = 4 * l;
  const int64_t row0 = tile * RT::TR;
  const int64_t nrow = m - row0 < RT::TR ? m - row0 : RT::TR;
  const rsrc_t rs = make_rsrc(X + row0 * ldx, uint32_t(((nrow - 1) * ldx + n) * 4));
  const uint32_t base = uint32_t(reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) char*)stage));
#pragma unroll
  for (int i = 0; i < RT::NLD; i++) {
    const int rt = w * RT::RPW + i;
    lds_dma16(rs, uint32_t((rt * ldx + (col < n ? col : 0)) * 4), base + uint32_t(rt * NB * 4));
  }
}
template <int NB>
PC_DEV void rows_unstage(const char* stage, f4* v) {
  using RT = RowTile<NB>;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < RT::NLD; i++) v[i] = *reinterpret_cast<const f4*>(stage + (w * RT::RPW + i) * NB * 4 + l * 16);
}


// ---------------------------------------------------------------------------
// P3: C = Ys W
// ---------------------------------------------------------------------------
template <int NB, int VEC, int CENTER, bool STAGED = false>
__global__ __launch_bounds__(NT2, 1) void k_gemm_mx2(const float* __restrict__ X, int64_t ldx, int64_t m, int n,
                                                     const double* __restrict__ mean,
                                                     const double* __restrict__ var, const h8* __restrict__ Wh,
                                                     const h8* __restrict__ Wl, const float* __restrict__ wscale,
                                                     int nc, int64_t tpc, float* __restrict__ C, int64_t ldc, int64_t zx = 0, int64_t zb = 0, int64_t zc = 0) {
  X = zoff(X, zx);
  mean = zoff(mean, zb);
  var = zoff(var, zb);
  Wh = zoff(Wh, zb);
  Wl = zoff(Wl, zb);
  wscale = zoff(wscale, zb);
  C = zoff(C, zc);
  using RT = RowTile<NB>;
  constexpr int KSW = RT::KSW, NCBLK = NB / 32;  // output columns padded like the input
  __shared__ __align__(16) char Fs[RT::FBYTES];
  __shared__ int rexp[RT::TR];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  // wave w computes column block J = w mod NCBLK for the 32-row blocks
  // R = w / NCBLK, + NW2 / NCBLK, ... of each tile: with fewer column blocks
  // than waves, the waves sharing a column block split the tile's rows.  Every
  // wave owns a column block when NB = 256: keep that test out of the code.
  constexpr int RSTEP = NCBLK >= NW2 ? 1 : NW2 / NCBLK;
  const int J = NCBLK >= NW2 ? w : w % NCBLK;
  const int R0 = NCBLK >= NW2 ? 0 : w / NCBLK;
  const bool own = NCBLK >= NW2 || R0 < RT::NR;
  h8 bh[KSW], bl[KSW];
  float isc = 1.0f;
  if (own) {
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
  // stores of one 32-row block: rows >= m fall beyond num_records, columns >=
  // nc get an offset beyond it; the hardware drops both (no branch, so the
  // wait for the next tile's loads never has to cover these stores)
  auto store_tile = [&](int64_t tile, const f16v& acc, const int* rx) {
    const int col = J * 32 + c;
    const int64_t row0 = tile * RT::TR;
    const int64_t nrow = m - row0 < RT::TR ? m - row0 : RT::TR;
    const rsrc_t cs = make_rsrc(C + row0 * ldc, uint32_t(((nrow - 1) * ldc + nc) * 4));
    // lane offset of register 0; the other registers add a wave-uniform row
    // step (the empty asm keeps the base from being hoisted as 16 offsets)
    uint32_t off0 = uint32_t((4 * hl * ldc + col) * 4);
    asm volatile("" : "+v"(off0));
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int rt = 8 * (r >> 2) + 4 * hl + (r & 3);
      const uint32_t off = col < nc ? off0 + uint32_t((8 * (r >> 2) + (r & 3)) * ldc * 4) : 0xFFFFFFFFu;
      // (the builtin's data operand is a 32-bit integer: pass the float's bits)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, ldexpf(acc[r], -rx[rt]) * isc), cs, off, 0, 0);
    }
  };
  // the tile's products and stores (A fragments in F, row exponents in rx)
  auto products = [&](int64_t tile, const char* F, const int* rx) {
    if (own) {
      const int col = J * 32 + c;
      // stores of this tile: rows >= m fall beyond num_records, columns >= nc get
      // an offset beyond it; the hardware drops both (no branch, so the wait for
      // the next tile's loads never has to cover these stores)
      const int64_t row0 = tile * RT::TR;
      const int64_t nrow = m - row0 < RT::TR ? m - row0 : RT::TR;
      const rsrc_t cs = make_rsrc(C + row0 * ldc, uint32_t(((nrow - 1) * ldc + nc) * 4));
#pragma unroll
      for (int R = R0; R < RT::NR; R += RSTEP) {
        f16v acc = {};
#pragma unroll
        for (int ks = 0; ks < KSW; ks++) {
          h8 ah = frag(F, (R * KSW + ks) * 2 + 0, RT::FRAGB, l);
          h8 al = frag(F, (R * KSW + ks) * 2 + 1, RT::FRAGB, l);
          acc = mfma(ah, bh[ks], acc);
          acc = mfma(ah, bl[ks], acc);
          acc = mfma(al, bh[ks], acc);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int rt = 32 * R + 8 * (r >> 2) + 4 * hl + (r & 3);
          const uint32_t off = col < nc ? uint32_t((rt * ldc + col) * 4) : 0xFFFFFFFFu;
          // (the builtin's data operand is a 32-bit integer: pass the float's bits)
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, ldexpf(acc[r], -rx[rt]) * isc), cs, off, 0,
                                                0);
        }
      }
    }
  };
  if constexpr (STAGED) {
    static_assert(VEC == 1 && NB == 256 && CENTER != 2, "staged pipeline: 256 float4-aligned columns");
    // Software pipeline over LDS: while the products of tile t run on
    // fragments Fa / Fb (alternating), the same waves split tile t + 1 into
    // the other fragment buffer, so the matrix cores and the vector ALU work
    // side by side and a tile costs one barrier.  Tiles arrive through two
    // LDS stages (St0 / St1: tile t0 + j lands in stage j mod 2); a stage is
    // refilled with the tile two ahead as soon as its rows are in registers.
    // Every wave owns one 32-column block here, so an iteration issues
    // exactly NLD loads and ST = 16 stores per wave, and the wait for tile
    // t + 1 leaves the later operations outstanding: NLD loads + ST stores
    // before the second iteration, NLD + 2 ST afterwards.  (The mean / variance loads
    // of CENTER 2 would be compiler-visible loads inside the loop: that case
    // takes the register pipeline below.)
    constexpr int NLD = RT::NLD, ST = 16, SB = RT::TR * NB * 4;
    static_assert(NCBLK == NW2 && RT::NR == 1, "one 32-row block and one column block per wave");
    __shared__ __align__(16) char St0[SB], St1[SB], Fb[RT::FBYTES];
    __shared__ int rxb[RT::TR];
    char* Fa = Fs;
    int* rxa = rexp;
    // the W fragments are loop invariant: settle their loads before the stage
    // loads so that no wait for them falls inside the loop
    __builtin_amdgcn_s_waitcnt(0);
    if (t0 >= t1) return;
    auto clamp = [&](int64_t t) { return t < t1 ? t : t1 - 1; };
    rows_stage<NB>(X, ldx, m, n, t0, St0);
    rows_stage<NB>(X, ldx, m, n, clamp(t0 + 1), St1);
    wait_vm<NLD>();
    {
      f4 v[NLD];
      rows_unstage<NB>(St0, v);
      rows_split<NB, CENTER>(v, m, n, mean, var, t0, Fa, rxa);
    }
    lds_sync();
    rows_stage<NB>(X, ldx, m, n, clamp(t0 + 2), St0);
    // iteration k: products of tile t (fragments F, exponents rx), split of
    // tile t + 1 (stage S) into (Fn, rxn), refill of S with tile t + 3
    auto body = [&](int64_t tile, int k, const char* F, const int* rx, char* S, char* Fn, int* rxn) {
      if (k == 0) wait_vm<NLD>();
      else if (k == 1) wait_vm<NLD + ST>();
      else wait_vm<NLD + 2 * ST>();
      f4 v[NLD];
      rows_unstage<NB>(S, v);
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // the stage is read before it is refilled
      rows_stage<NB>(X, ldx, m, n, clamp(tile + 3), S);
      // products of tile t in four chunks of four k-steps (two accumulators,
      // alternate k-steps), each followed by the split of one row of tile
      // t + 1: the chunks are kept in this order, so the vector work of one
      // wave meets the matrix work of the other wave on the SIMD
      const int64_t nt = clamp(tile + 1);
      f16v acc0 = {}, acc1 = {};
#pragma unroll
      for (int q = 0; q < NLD; q++) {
#pragma unroll
        for (int kk = 0; kk < KSW / NLD; kk++) {
          const int ks = q * (KSW / NLD) + kk;
          const h8 ah = frag(F, ks * 2 + 0, RT::FRAGB, l);
          const h8 al = frag(F, ks * 2 + 1, RT::FRAGB, l);
          f16v& acc = (ks & 1) ? acc1 : acc0;
          acc = mfma(ah, bh[ks], acc);
          acc = mfma(ah, bl[ks], acc);
          acc = mfma(al, bh[ks], acc);
        }
        __builtin_amdgcn_sched_barrier(0);
        split_row<NB, CENTER>(v[q], q, m, n, mean, var, nt, Fn, rxn);
        __builtin_amdgcn_sched_barrier(0);
      }
      store_tile(tile, acc0 + acc1, rx);
      lds_sync();
    };
    int64_t tile = t0;
    int k = 0;
    for (; tile + 1 < t1; tile += 2, k += 2) {
      body(tile, k, Fa, rxa, St1, Fb, rxb);
      body(tile + 1, k + 1, Fb, rxb, St0, Fa, rxa);
    }
    if (tile < t1) body(tile, k, Fa, rxa, St1, Fb, rxb);
    // no LDS write may still be pending when the workgroup ends
    wait_vm<0>();
    return;
  }
  // two tiles in flight per wave: buffer va / vb alternate, and the load issued
  // while a tile is processed is for the tile after next (always issued,
  // clamped to the chunk's last tile, so the code stays straight-line)
  f4 va[RT::NLD], vb[RT::NLD];
  if (t0 < t1) {
    rows_load<NB, VEC>(X, ldx, m, n, t0, va);
    rows_load<NB, VEC>(X, ldx, m, n, t0 + 1 < t1 ? t0 + 1 : t0, vb);
  }
  // the W fragments are loop invariant: settle every prologue load here so that
  // no wait for them is placed inside the loop (it would also wait for the
  // prefetch of the next tiles)
  __builtin_amdgcn_s_waitcnt(0);
  auto body = [&](int64_t tile, f4* v) {
    rows_split<NB, CENTER>(v, m, n, mean, var, tile, Fs, rexp);
    lds_sync();
    rows_load<NB, VEC>(X, ldx, m, n, tile + 2 < t1 ? tile + 2 : t1 - 1, v);
    products(tile, Fs, rexp);
    lds_sync();
  };
  int64_t tile = t0;
  for (; tile + 1 < t1; tile += 2) {
    body(tile, va);
    body(tile + 1, vb);
  }
  if (tile < t1) body(tile, va);
}

// ---------------------------------------------------------------------------
// P2: Z = Ys W1 (fp16 W1, per-column scale), G2 = Z^T Z
// ---------------------------------------------------------------------------
template <int NB, int VEC, int CENTER>
__global__ __launch_bounds__(NT2, 1) void k_zgram_mx2(const float* __restrict__ X, int64_t ldx, int64_t m, int n,
                                                      const double* __restrict__ mean,
                                                      const double* __restrict__ var, const h8* __restrict__ Wf,
                                                      const float* __restrict__ wscale, int64_t tpc,
                                                      double* __restrict__ part, int upper, int64_t zx = 0,
                                                      int64_t zb = 0) {
  X = zoff(X, zx);
  mean = zoff(mean, zb);
  var = zoff(var, zb);
  Wf = zoff(Wf, zb);
  wscale = zoff(wscale, zb);
  part = zoff(part, zb);
  using RT = RowTile<NB>;
  constexpr int KSW = RT::KSW, NBLK = NB / 32;
  constexpr int NU = NBLK * (NBLK + 1) / 2;
  constexpr int S = (NU + NW2 - 1) / NW2;
  constexpr int KSG = RT::TR / 16;  // Gram k-steps per tile
  // Y fragments and row exponents alternate between two buffers: while the
  // Z products of tile t read one, the waves split tile t + 1 into the other
  // (one row between consecutive chunks of products), so the vector ALU work
  // of one wave meets the matrix work of the other wave on each SIMD, and a
  // tile costs two barriers.  Every product and split is the one the
  // unpipelined order formed, so the results do not change.
  __shared__ __align__(16) char Fs[2][RT::FBYTES];
  __shared__ h8 F2[KSG * NBLK * 2 * 64];
  __shared__ int rexp[2][RT::TR];
  __shared__ float red[NW2];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, hl = l >> 5, c = l & 31;
  // Column block of Z this wave forms.  With an upper triangular W1 (upper
  // != 0: W1 = R1^-1, zero-padded) block J only needs the k-steps of rows
  // < 32 (J + 1), a share J + 1 of NBLK; when every wave owns a block, waves
  // 2k and 2k + 1 take blocks k and NBLK - 1 - k, so each pair does the same
  // share.  The skipped products are exact zeros: results do not change.
  const int J = __builtin_amdgcn_readfirstlane(NBLK == NW2 && upper ? ((w & 1) ? NBLK - 1 - (w >> 1) : (w >> 1)) : w);
  const int kend = upper ? (2 * J + 2 < KSW ? 2 * J + 2 : KSW) : KSW;
  const bool own = NBLK >= NW2 || J < NBLK;  // see k_gemm_mx2
  h8 bw[KSW];
  float isc = 1.0f;
  if (own) {
#pragma unroll
    for (int ks = 0; ks < KSW; ks++) bw[ks] = Wf[int64_t(ks * NBLK + J) * 64 + l];
    isc = 1.0f / wscale[J * 32 + c];
  }
  int sI[S], sJ[S];
  bool sv[S];
  {
    int s0, sc;
    wave_slots<NU>(w, s0, sc);
#pragma unroll
    for (int s = 0; s < S; s++) sv[s] = upper_slot<NBLK>(s0 + s, sI[s], sJ[s]) && s < sc;
  }
  f16v acc[S];
#pragma unroll
  for (int s = 0; s < S; s++) acc[s] = f16v{};
  int e_run = EUN;
  const int64_t ntiles = ceil_div(m, RT::TR);
  const int64_t t0 = int64_t(blockIdx.x) * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  auto clamp = [&](int64_t t) { return t < t1 ? t : t1 - 1; };
  // one tile in registers (the rest are taken by W1 and the Gram
  // accumulators); a load is always issued, clamped to the chunk's last tile
  f4 va[RT::NLD];
  if (t0 < t1) rows_load<NB, VEC>(X, ldx, m, n, t0, va);
  // the W fragments are loop invariant: settle every prologue load here so that
  // no wait for them is placed inside the loop (it would also wait for the
  // prefetch of the next tiles)
  __builtin_amdgcn_s_waitcnt(0);
  if (t0 < t1) {
    rows_split<NB, CENTER>(va, m, n, mean, var, t0, Fs[0], rexp[0]);
    rows_load<NB, VEC>(X, ldx, m, n, clamp(t0 + 1), va);
  }
  lds_sync();
  // tile t: products from (F, rx), split of tile t + 1 (in va) into (Fn, rxn)
  auto body = [&](int64_t tile, const char* F, const int* rx, char* Fn, int* rxn) {
    constexpr int NIT = RT::NR * KSW;  // (R, ks) products, in NLD chunks
    const int64_t nt = clamp(tile + 1);
    f16v z[RT::NR];
#pragma unroll
    for (int R = 0; R < RT::NR; R++) z[R] = f16v{};
#pragma unroll
    for (int q = 0; q < RT::NLD; q++) {
      if (own) {
#pragma unroll
        for (int it = q * NIT / RT::NLD; it < (q + 1) * NIT / RT::NLD; it++) {
          const int R = it / KSW, ks = it % KSW;
          if (ks < kend) {  // wave-uniform
            z[R] = mfma(frag(F, (R * KSW + ks) * 2 + 0, RT::FRAGB, l), bw[ks], z[R]);
            z[R] = mfma(frag(F, (R * KSW + ks) * 2 + 1, RT::FRAGB, l), bw[ks], z[R]);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      split_row<NB, CENTER>(va[q], q, m, n, mean, var, nt, Fn, rxn);
      __builtin_amdgcn_sched_barrier(0);
    }
    rows_load<NB, VEC>(X, ldx, m, n, clamp(tile + 2), va);
    float zmx = 0.0f;
    if (own) {
#pragma unroll
      for (int R = 0; R < RT::NR; R++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int rt = 32 * R + 8 * (r >> 2) + 4 * hl + (r & 3);
          const float zz = ldexpf(z[R][r], -rx[rt]) * isc;
          z[R][r] = zz;
          zmx = fmaxf(zmx, fabsf(zz));
        }
    }
    zmx = group_max(zmx, 64);
    if (l == 0) red[w] = zmx;
    // also publishes the split of tile t + 1
    lds_sync();
    float tmax = red[0];
#pragma unroll
    for (int q = 1; q < NW2; q++) tmax = fmaxf(tmax, red[q]);
    e_run = update_scale<S>(tmax, e_run, acc);
    const float sc = ldexpf(1.0f, e_run == EUN ? 0 : e_run);
    if (own) {
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
    lds_sync();
    // F2 is next written after the next tile's first barrier, which every
    // wave reaches only once its reads here are done
#pragma unroll
    for (int ks = 0; ks < KSG; ks++) gram_step<S>(F2 + ks * NBLK * 2 * 64, sI, sJ, sv, acc, l);
  };
  for (int64_t tile = t0; tile < t1; tile++) {
    const int pb = int(tile - t0) & 1;
    body(tile, Fs[pb], rexp[pb], Fs[pb ^ 1], rexp[pb ^ 1]);
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
// (load mode, CENTER) as compile-time constants; mode 0 scalar, 1 float4,
// 2 float2 loads
template <class F>
int dispatch_vc(int vec, int center, F&& f) {
  if (center < 0 || center > 2) return -1;
  auto c = [&](auto vv) {
    if (center == 0) return f(vv, std::integral_constant<int, 0>{});
    if (center == 1) return f(vv, std::integral_constant<int, 1>{});
    return f(vv, std::integral_constant<int, 2>{});
  };
  if (vec == 1) return c(std::integral_constant<int, 1>{});
  if (vec == 2) return c(std::integral_constant<int, 2>{});
  return c(std::integral_constant<int, 0>{});
}

// float4 path: 16-byte aligned rows and n a multiple of 4 (a row's float4 is
// then either wholly inside or wholly outside the n valid columns); float2
// path: 8-byte aligned rows and n even.  zx is the byte stride between the
// problems of a batched launch.
int load_mode(const float* X, int64_t ldx, int n, int64_t zx) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(X);
  if ((a & 15) == 0 && (ldx & 3) == 0 && (n & 3) == 0 && zx % 16 == 0) return 1;
  if ((a & 7) == 0 && (ldx & 1) == 0 && (n & 1) == 0 && zx % 8 == 0) return 2;
  return 0;
}

}  // namespace

int mx2_nb(int n) { return n <= 32 ? 32 : n <= 64 ? 64 : n <= 128 ? 128 : n <= 256 ? 256 : -1; }
int mx2_gram_rows(int n) {
  return dispatch_nb(n, [](auto nbc) {
    constexpr int NB = decltype(nbc)::value, V = NB >= 128 ? 4 : NB / 32;
    return 16 * (NW2 / (NB / (32 * V)));
  });
}
int mx2_tile_rows(int n) {
  return dispatch_nb(n, [](auto nbc) { return RowTile<decltype(nbc)::value>::TR; });
}

int launch_gram_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, double* mean_io,
                    const double* var, int chunks, int64_t tpc, double* part, const PassBatch& bz) {
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    const int mode = load_mode(X, ldx, n, bz.zx);
    return dispatch_vc(mode == 1 ? 1 : 0, center, [&](auto vv, auto cc) {
      hipLaunchKernelGGL((k_gram_mx2<NB, decltype(vv)::value, decltype(cc)::value>),
                         dim3(unsigned(chunks), 1, unsigned(bz.nz)), dim3(NT2), 0, st, X, ldx, m, n, mean_io, var, tpc,
                         part, bz.zx, bz.zb);
      return 0;
    });
  });
}

int launch_zgram_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean,
                     const double* var, const _Float16* Wf, const float* wscale, int chunks, int64_t tpc,
                     double* part, const PassBatch& bz, bool upper) {
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    return dispatch_vc(load_mode(X, ldx, n, bz.zx), center, [&](auto vv, auto cc) {
      hipLaunchKernelGGL((k_zgram_mx2<NB, decltype(vv)::value, decltype(cc)::value>),
                         dim3(unsigned(chunks), 1, unsigned(bz.nz)), dim3(NT2), 0, st, X, ldx, m, n, mean, var,
                         (const h8*)Wf, wscale, tpc, part, upper ? 1 : 0, bz.zx, bz.zb);
      return 0;
    });
  });
}

// the W operand is packed NB x NB (nc <= n output columns, padded like the input)
int launch_gemm_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean,
                    const double* var, const _Float16* Wh, const _Float16* Wl, const float* wscale, int nc,
                    int chunks, int64_t tpc, float* C, int64_t ldc, const PassBatch& bz) {
  if (mx2_nb(nc) < 0 || mx2_nb(nc) > mx2_nb(n)) return -1;
  return dispatch_nb(n, [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    return dispatch_vc(load_mode(X, ldx, n, bz.zx), center, [&](auto vv, auto cc) {
      constexpr int VV = decltype(vv)::value, CC = decltype(cc)::value;
      if constexpr (VV == 1 && NB == 256 && CC != 2) {
        // LDS-staged pipeline
        {
          // it counts its own vector memory operations: a register spill would
          // add scratch traffic it does not count, so a build that spills is
          // refused rather than run
          static const bool spills = [] {
            hipFuncAttributes fa{};
            return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_gemm_mx2<NB, VV, CC, true>)) !=
                       hipSuccess ||
                   fa.localSizeBytes != 0;
          }();
          if (spills) return -2;
          hipLaunchKernelGGL((k_gemm_mx2<NB, VV, CC, true>), dim3(unsigned(chunks), 1, unsigned(bz.nz)), dim3(NT2), 0,
                             st, X, ldx, m, n, mean, var, (const h8*)Wh, (const h8*)Wl, wscale, nc, tpc, C, ldc, bz.zx,
                             bz.zb, bz.zc);
          return 0;
        }
      }
      hipLaunchKernelGGL((k_gemm_mx2<NB, decltype(vv)::value, decltype(cc)::value>),
                         dim3(unsigned(chunks), 1, unsigned(bz.nz)), dim3(NT2), 0, st, X, ldx, m, n, mean, var,
                         (const h8*)Wh, (const h8*)Wl, wscale, nc, tpc, C, ldc, bz.zx, bz.zb, bz.zc);
      return 0;
    });
  });
}
