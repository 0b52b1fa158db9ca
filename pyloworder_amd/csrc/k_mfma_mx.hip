// Operand packing of the mixed-precision (fp16 split) path: an FP64 matrix W
// becomes the B fragments of v_mfma_f32_32x32x16_f16 as two fp16 pieces,
// hi = fp16(w 2^e) and lo = fp16(w 2^e - hi), with one power-of-two scale per
// column (undone exactly by the streaming kernels of k_mfma_mx2.hip).
//
// MFMA lane layout of the B operand (wave64, verified on gfx950), with
// hl = lane / 32, c = lane % 32: lane holds B[8 hl + t][c], t = 0..7.
#include "common.hpp"
#include "kernels.hpp"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// exponent e with |x| * 2^e < 2^14 for every |x| <= mx (mx > 0)
PC_DEV int scale_exp(float mx) {
  int E;
  frexpf(mx, &E);  // mx = f 2^E, f in [0.5, 1)
  return 14 - E;
}

// ---------------------------------------------------------------------------
// W (FP64 NB x NCB, ld) -> B fragments (hi, lo) with per-column exponents;
// Wr = hi * 2^-f (the operand the hi-only products apply).  One workgroup
// per 32-column block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack_w(const double* __restrict__ W, int ldw, int NB, int NCB, h8* Wh, h8* Wl,
                                                float* wscale, double* Wr, int ldwr, int64_t zb = 0) {
  W = zoff(W, zb);
  Wh = zoff(Wh, zb);
  Wl = zoff(Wl, zb);
  wscale = zoff(wscale, zb);
  Wr = zoff(Wr, zb);
  int J = blockIdx.x;
  int ncb = NCB / 32;
  __shared__ float fsc[32];
  __shared__ double pmx[8][32];
  // column maxima: eight row groups per column, their loads all in flight
  {
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5, col = J * 32 + c;
    double mx = 0.0;
#pragma unroll 8
    for (int i = g; i < NB; i += 8) mx = fmax(mx, fabs(W[int64_t(i) * ldw + col]));
    pmx[g][c] = mx;
  }
  __syncthreads();
  if (threadIdx.x < 32) {
    int col = J * 32 + threadIdx.x;
    double mx = 0.0;
#pragma unroll
    for (int g = 0; g < 8; g++) mx = fmax(mx, pmx[g][threadIdx.x]);
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

}  // namespace

int launch_pack_w_mx(hipStream_t st, const double* W, int ldw, int n, int nc, int NB, int NCB, _Float16* Wh,
                     _Float16* Wl, float* wscale, double* Wr, int ldwr, int nz, int64_t zb) {
  (void)n; (void)nc;
  hipLaunchKernelGGL(k_pack_w, dim3(NCB / 32, 1, nz), dim3(256), 0, st, W, ldw, NB, NCB, (h8*)Wh, (h8*)Wl, wscale, Wr,
                     ldwr, zb);
  return 0;
}
