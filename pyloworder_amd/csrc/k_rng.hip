// Sketch matrix of the randomized QR drawn on the device.
//
// The reference draws Omega with np.random.seed(seed); np.random.rand(n, r)
// (pyLOM/vmmath/svd.py:131-133): numpy's legacy RandomState, i.e. MT19937
// (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998) seeded with init_genrand,
// and doubles built from two outputs as in genrand_res53,
// (a >> 5, b >> 6) -> (a 2^26 + b) 2^-53.  numpy keeps this stream frozen, so
// the same seed gives the reference's Omega bit for bit (then rounded to
// float32 for float32 data, as .astype(Ai.dtype)).
//
// One workgroup: thread t < 624 owns state word t.  The twist of the 624
// words runs in four dependency phases ([0,227) reads old words only,
// [227,454) and [454,623) read words of the previous phase, 623 reads words
// 0 and 396), each a read / barrier / write step; then every word is
// tempered and each pair becomes one double.
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int kMtN = 624;
constexpr int kMtM = 397;

PC_DEV uint32_t mt_step(const uint32_t* mt, int i) {
  const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % kMtN] & 0x7fffffffu);
  return mt[(i + kMtM) % kMtN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

template <class TO>
__global__ __launch_bounds__(1024) void k_mt_omega(uint32_t seed, int64_t count, TO* out) {
  __shared__ uint32_t mt[kMtN];
  __shared__ uint32_t tw[kMtN];
  const int t = threadIdx.x;
  if (t == 0) {
    uint32_t s = seed;
    for (int i = 0; i < kMtN; i++) {
      mt[i] = s;
      s = 1812433253u * (s ^ (s >> 30)) + uint32_t(i + 1);
    }
  }
  __syncthreads();
  const int lo[3] = {0, 227, 454}, hi[3] = {227, 454, 623};
  for (int64_t d0 = 0; d0 < count; d0 += kMtN / 2) {
#pragma unroll
    for (int ph = 0; ph < 3; ph++) {
      const bool mine = t >= lo[ph] && t < hi[ph];
      uint32_t v = 0;
      if (mine) v = mt_step(mt, t);
      __syncthreads();
      if (mine) mt[t] = v;
      __syncthreads();
    }
    if (t == kMtN - 1) mt[t] = mt_step(mt, t);
    __syncthreads();
    if (t < kMtN) {
      uint32_t y = mt[t];
      y ^= y >> 11;
      y ^= (y << 7) & 0x9d2c5680u;
      y ^= (y << 15) & 0xefc60000u;
      y ^= y >> 18;
      tw[t] = y;
    }
    __syncthreads();
    if (t < kMtN / 2 && d0 + t < count) {
      const double a = double(tw[2 * t] >> 5), b = double(tw[2 * t + 1] >> 6);
      out[d0 + t] = TO((a * 67108864.0 + b) * (1.0 / 9007199254740992.0));
    }
    __syncthreads();
  }
}

}  // namespace

// count uniform [0, 1) values of numpy's RandomState(seed).random_sample into
// out (float32 or float64), in stream order
int launch_mt_omega(hipStream_t st, uint32_t seed, int64_t count, pc_dtype dt, void* out) {
  if (dt == PC_F64) hipLaunchKernelGGL((k_mt_omega<double>), dim3(1), dim3(1024), 0, st, seed, count, (double*)out);
  else if (dt == PC_F32) hipLaunchKernelGGL((k_mt_omega<float>), dim3(1), dim3(1024), 0, st, seed, count, (float*)out);
  else return -1;
  return 0;
}
