// Host-side launchers of the device kernels (one declaration per kernel family).
#pragma once
#include "common.hpp"

// FP64 MFMA streaming kernels (k_mfma_f64.hip); return 0, or -1 when n is too large
int f64_nb(int n);
int f64_tile_rows();
// Batched streaming passes: nz problems in one launch (blockIdx.z), the input
// moving by zx bytes, the workspace operands by zb and the output by zc.
struct PassBatch {
  int nz = 1;
  int64_t zx = 0, zb = 0, zc = 0;
};
// W non-null: the second pass, Gram of Z = Y W; upper: W is upper triangular
// (zero-padded), so Z skips the k-steps where W is zero
int launch_gram_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int n, int center,
                    int stats_in, const double* mean_in, const double* var_in, double* mean_out, double* var_out,
                    const double* W, int chunks, int64_t tpc, double* part, const PassBatch& bz = PassBatch(),
                    bool upper = false);
int launch_gemm_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int n, int center,
                    const double* mean_in, const double* var_in, const double* W, int nc, int chunks, int64_t tpc,
                    void* C, int64_t ldc, const PassBatch& bz = PassBatch());
int launch_xgram_f64(hipStream_t st, pc_dtype dt, const void* A, int64_t lda, int p, const void* B, int64_t ldb, int q,
                     int64_t m, int chunks, int64_t tpc, double* part);

// mixed (fp16 split) MFMA streaming kernels (k_mfma_mx2.hip): 8 waves, W operand in registers
int mx2_nb(int n);
int mx2_gram_rows(int n);
int mx2_tile_rows(int n);
int launch_gram_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, double* mean_io,
                    const double* var, int chunks, int64_t tpc, double* part, const PassBatch& bz = PassBatch());
// upper: Wf is upper triangular (zero rows below each 32-column block's
// diagonal), so the Z = Ys W products skip the zero k-steps
int launch_zgram_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean,
                     const double* var, const _Float16* Wf, const float* wscale, int chunks, int64_t tpc,
                     double* part, const PassBatch& bz = PassBatch(), bool upper = false);
int launch_gemm_mx2(hipStream_t st, const float* X, int64_t ldx, int64_t m, int n, int center, const double* mean,
                    const double* var, const _Float16* Wh, const _Float16* Wl, const float* wscale, int nc,
                    int chunks, int64_t tpc, float* C, int64_t ldc, const PassBatch& bz = PassBatch());
// FP64 row mean / variance of float32 rows (k_stream.hip)
void launch_row_stats_f32(hipStream_t st, const float* X, int64_t ldx, int64_t m, int64_t n, double* mean,
                          double* var);
void launch_row_stats_f64(hipStream_t st, const double* X, int64_t ldx, int64_t m, int64_t n, double* mean,
                          double* var);
// pack an FP64 n x nc matrix (ld ldw) into fp16 hi (and lo) fragment order with
// per-column power-of-two scales; also writes the FP64 value of hi*2^-s (Wr)
// (k_mfma_mx.hip)
int launch_pack_w_mx(hipStream_t st, const double* W, int ldw, int n, int nc, int NB, int NCB, _Float16* Wh,
                     _Float16* Wl, float* wscale, double* Wr, int ldwr, int nz = 1, int64_t zb = 0);
// numpy RandomState(seed).random_sample stream (MT19937, genrand_res53) into
// count float32 / float64 values (k_rng.hip); -1 for other dtypes
int launch_mt_omega(hipStream_t st, uint32_t seed, int64_t count, pc_dtype dt, void* out);
// wide problems (256 < real columns <= 2048), FP64 MFMA (k_wide.hip): stripe-pair
// Gram into NBp x NBp slabs per row chunk, and the K-looped tall product
int wide_stripe();
int wide_tile_rows();
int launch_wgram_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int n, int center,
                     const double* mean, const double* var, int chunks, int64_t tpc, double* part, int NBp);
int launch_wgemm_f64(hipStream_t st, pc_dtype dt, const void* X, int64_t ldx, int64_t m, int K, int center,
                     const double* mean, const double* var, const double* W, int64_t ldw, int N, int chunks,
                     int64_t tpc, void* C, int64_t ldc, int out_f64);
int launch_wxgram_f64(hipStream_t st, pc_dtype dt, const void* A, int64_t lda, int p, const void* B, int64_t ldb,
                      int q, int64_t m, int chunks, int64_t tpc, double* part, int PBp, int QBp);
