// Mixed-precision (fp16 hi/lo split) MFMA streaming kernels — placeholder
// launchers until the kernels land; every entry reports "no variant".
#include "common.hpp"
#include "kernels.hpp"
int mx_nb(int n) { (void)n; return -1; }
int mx_gram_rows() { return 128; }
int mx_gemm_rows() { return 64; }
int launch_gram_mx(hipStream_t, const float*, int64_t, int64_t, int, int, double*, double*, int, int64_t, double*) { return -1; }
int launch_zgram_mx(hipStream_t, const float*, int64_t, int64_t, int, int, const double*, const double*, const _Float16*,
                    const float*, int, int64_t, double*) { return -1; }
int launch_gemm_mx(hipStream_t, const float*, int64_t, int64_t, int, int, const double*, const double*, const _Float16*,
                   const _Float16*, const float*, int, int, int64_t, float*, int64_t) { return -1; }
int launch_pack_w_mx(hipStream_t, const double*, int, int, int, int, int, _Float16*, _Float16*, float*, double*, int) { return -1; }
