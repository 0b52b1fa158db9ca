/*
 * podcore.h — C ABI of the MI355X POD core (tall-skinny SVD / QR, POD).
 *
 * Every entry point replaces one function of the reference's native or
 * Python path (pyLOM, ArnauMiro/pyLowOrder); the reference interface is cited
 * beside each declaration.  Conventions that differ from the reference:
 *   - pointers are DEVICE pointers (HBM) owned by the caller; the library never
 *     frees caller memory.  Library workspace lives in the context and is
 *     counted against its HBM budget;
 *   - sizes are int64_t and every matrix carries a leading dimension (row
 *     stride, in elements), so column views such as U[:, :N] pass uncopied;
 *     all matrices are row-major;
 *   - inputs are never modified (the reference's ?svd destroys its input,
 *     pyLOM/vmmath/svd.pyx:138-144);
 *   - the return value is PC_OK (0) or a negative pc_status; nothing aborts
 *     (the reference calls MPI_Abort through raiseError,
 *     pyLOM/utils/errors.py:22-28).  pc_last_error() gives the message;
 *   - work is ordered on the context's stream (pc_set_stream) and calls are
 *     asynchronous unless documented otherwise.  One context per host thread;
 *     concurrent calls on one context race.
 *
 * Element types: PC_F32 / PC_F64 real, PC_C64 / PC_C128 complex (interleaved
 * re, im).  Singular values are always real, stored as float for PC_F32 /
 * PC_C64 and as double for PC_F64 / PC_C128.
 *
 * Precision modes (argument `precision`):
 *   PC_PREC_FP64  — all arithmetic in FP64 (FP64 MFMA), whatever the storage;
 *   PC_PREC_MIXED — FP32/complex64 storage only: products on fp16 hi/lo
 *                   split MFMA with FP32 accumulation, every n x n step
 *                   (Gram reductions, Cholesky, small SVD) in FP64.
 */
#ifndef PODCORE_H
#define PODCORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pc_ctx pc_ctx;

typedef enum { PC_F32 = 0, PC_F64 = 1, PC_C64 = 2, PC_C128 = 3 } pc_dtype;
typedef enum { PC_PREC_FP64 = 0, PC_PREC_MIXED = 1 } pc_precision;
typedef enum {
  PC_CENTER_NONE = 0,  /* use the rows as given                                   */
  PC_CENTER_MEAN = 1,  /* subtract the row (temporal) mean                        */
  PC_CENTER_MEAN_VAR = 2 /* subtract the mean and divide by the row VARIANCE      */
} pc_center;

typedef enum {
  PC_OK = 0,
  PC_ERR_INVALID = -1,     /* null pointer, bad enum, bad leading dimension     */
  PC_ERR_SHAPE = -2,       /* unsupported or inconsistent sizes                 */
  PC_ERR_NOMEM = -3,       /* workspace would exceed the context's HBM budget   */
  PC_ERR_HIP = -4,         /* a HIP runtime call failed                         */
  PC_ERR_LINALG = -5,      /* non-finite input or a factorisation broke down    */
  PC_ERR_UNSUPPORTED = -6  /* dtype / precision combination not provided        */
} pc_status;

/* ---- context ------------------------------------------------------------
 * No reference counterpart: the reference keeps MPI_COMM_WORLD and the cupy
 * device implicit (pyLOM/utils/mpi.py:17-19, pyLOM/utils/gpu.py:18-23).     */
int pc_ctx_create(int device, int64_t hbm_budget_bytes, uint32_t flags, pc_ctx** out);
int pc_ctx_destroy(pc_ctx* ctx);
int pc_set_stream(pc_ctx* ctx, void* hip_stream);
const char* pc_last_error(const pc_ctx* ctx);
/* bytes of workspace currently held, and the most a solve of (m, n) needs.   */
int64_t pc_workspace_bytes(const pc_ctx* ctx);
int64_t pc_workspace_query(int op, pc_dtype dtype, int64_t m, int64_t n, int64_t r, int precision);
int pc_synchronize(pc_ctx* ctx);
/* Per-stage timing (no reference counterpart; diagnostics for measurement):
 * when enabled, HIP events bracket each streaming pass on the context stream.
 * pc_profile_read synchronises and returns cumulative milliseconds and launch
 * counts per stage: 0 Gram pass, 1 preconditioned Gram pass, 2 product pass. */
int pc_profile(pc_ctx* ctx, int enable);
int pc_profile_read(pc_ctx* ctx, double* ms, int64_t* counts, int nstages);
const char* pc_version(void);

/* ---- averaging -----------------------------------------------------------
 * temporal_mean   pyLOM/vmmath/averaging.py:17-29, src/averaging.c:10-46
 * temporal_variance pyLOM/vmmath/averaging.py:46-59, src/averaging.c:48-90
 * mean / var are length m, same real type as X (var may be NULL).           */
int pc_row_stats(pc_ctx* ctx, pc_dtype dtype, const void* X, int64_t ldx, int64_t m, int64_t n,
                 void* mean, void* var);
/* subtract_mean  pyLOM/vmmath/averaging.py:31-44, src/averaging.c:92-124
 * norm_variance  pyLOM/vmmath/averaging.py:61-75, src/averaging.c:126-157
 * Y = (X - mean 1^T) [./ var 1^T]; var may be NULL.  Y may alias X.         */
int pc_row_center(pc_ctx* ctx, pc_dtype dtype, const void* X, int64_t ldx, int64_t m, int64_t n,
                  const void* mean, const void* var, void* Y, int64_t ldy);

/* ---- SVD / QR ------------------------------------------------------------
 * tsqr_svd  pyLOM/vmmath/svd.py:227-252; src/svd.c:529-563 (s), 678-712 (d),
 *           827-861 (c), 976-1010 (z); int dtsqr_svd(double*Ui,double*S,
 *           double*VT,double*Ai,int m,int n) at src/svd.h:25.
 * With center != PC_CENTER_NONE the row centering of POD.run
 * (pyLOM/POD/wrapper.py:33-45) is fused into the passes over A.
 * U is m x n (ldu), S n, VT n x n (ldvt, rows = right singular vectors).   */
int pc_tsqr_svd(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n,
                int center, void* U, int64_t ldu, void* S, void* VT, int64_t ldvt, int precision);
/* No counterpart in the reference: pc_tsqr_svd for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * pc_tsqr_svd, but the workspace is the library's static device scratch of
 * 128 MiB per GPU (held until the process ends) instead of the context's
 * hipMalloc'd buffer; the call makes no hipMalloc / hipFree and never grows a
 * buffer.  A problem needing more scratch (pc_workspace_query(0, ...)
 * > 128 MiB: n <= 64 fits up to about 6 million rows, n = 128 about 390,000,
 * n = 256 about 29,000) returns PC_ERR_NOMEM with no work queued.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams, other threads, other
 * contexts) race; the scratch keeps the last call's contents.               */
int pc_tsqr_svd_noalloc(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n, int center,
                        void* U, int64_t ldu, void* S, void* VT, int64_t ldvt, int precision);
/* tsqr  pyLOM/vmmath/svd.py:49-118; src/svd.c:416-527 (s), 565-676 (d),
 *       714-825 (c), 863-974 (z).  Also serves qr (svd.py:28-36, svd.c:237-407).
 * Q m x n (ldq), R n x n upper triangular (ldr).                            */
int pc_tsqr(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n,
            void* Q, int64_t ldq, void* R, int64_t ldr, int precision);
/* svd (dense, thin)  pyLOM/vmmath/svd.py:38-47; src/svd.c:35-235.
 * A is m x n with min(m, n) <= 512; U m x k, S k, VT k x n, k = min(m, n).  */
int pc_svd(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n,
           void* U, int64_t ldu, void* S, void* VT, int64_t ldvt);
/* randomized_qr   pyLOM/vmmath/svd.py:120-144; src/svd.c:1012-1064, 1267-1319
 * randomized_svd  pyLOM/vmmath/svd.py:254-273; src/svd.c:1198-1265, 1453-1520
 * omega (n x r, device, ldo = r) is the caller's sketch matrix (the reference
 * draws it from numpy's seeded stream, svd.py:131-133; pass it for parity).
 * If omega is NULL it is drawn on the device from `seed` (uniform [0,1)).  */
int pc_randomized_qr(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n,
                     int64_t r, int q, uint64_t seed, const void* omega,
                     void* Q, int64_t ldq, void* B, int64_t ldb, int precision);
int pc_randomized_svd(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n,
                      int64_t r, int q, uint64_t seed, const void* omega,
                      void* U, int64_t ldu, void* S, void* VT, int64_t ldvt, int precision);

/* ---- products ------------------------------------------------------------
 * matmul   pyLOM/vmmath/maths.py:76-90; src/vector_matrix.c:158-242
 * matmulp  pyLOM/vmmath/maths.py:92-110; src/vector_matrix.c:330-384 (on one
 *          GPU the cross-rank allreduce is empty, the product is matmul).
 * C = alpha * op(A) op(B) + beta * C, op = identity or (conjugate) transpose
 * (trans = 0 none, 1 transpose, 2 conjugate transpose).  op(A) is m x k.   */
int pc_gemm(pc_ctx* ctx, pc_dtype dtype, int transA, int transB, int64_t m, int64_t n, int64_t k,
            double alpha, const void* A, int64_t lda, const void* B, int64_t ldb, double beta,
            void* C, int64_t ldc, int precision);
/* vecmat  pyLOM/vmmath/maths.py:112-129; src/vector_matrix.c:386-445:
 * C[i,:] = v[i] * A[i,:] (v real for real dtypes, same type otherwise).    */
int pc_vecmat(pc_ctx* ctx, pc_dtype dtype, const void* v, const void* A, int64_t lda, int64_t m,
              int64_t n, void* C, int64_t ldc);
/* Y += alpha X (m x n blocks).  No reference function of its own: it is the
 * sketch accumulation Yo += Yn of update_qr_streaming, pyLOM/vmmath/svd.py:217. */
int pc_axpy(pc_ctx* ctx, pc_dtype dtype, double alpha, const void* X, int64_t ldx, void* Y, int64_t ldy,
            int64_t m, int64_t n);
/* transpose  pyLOM/vmmath/maths.py:17-29; src/vector_matrix.c:30-52.       */
int pc_transpose(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, int64_t m, int64_t n,
                 void* B, int64_t ldb);

/* ---- POD -----------------------------------------------------------------
 * reconstruct  pyLOM/POD/wrapper.py:85-103; POD/wrapper.pyx:284-377:
 *   X = U diag(S) VT, U m x N, S N (real), VT N x n, X m x n.
 * project (no reference function; idiom matmulp(transpose(U), Y) at
 *   pyLOM/DMD/wrapper.py:85): C = U^H X, C N x n.                           */
int pc_reconstruct(pc_ctx* ctx, pc_dtype dtype, const void* U, int64_t ldu, const void* S,
                   const void* VT, int64_t ldvt, int64_t m, int64_t N, int64_t n, void* X, int64_t ldx,
                   int precision);
int pc_project(pc_ctx* ctx, pc_dtype dtype, const void* U, int64_t ldu, const void* X, int64_t ldx,
               int64_t m, int64_t N, int64_t n, void* C, int64_t ldc, int precision);

/* ---- statistics / truncation ---------------------------------------------
 * RMSE    pyLOM/vmmath/stats.py:16-34; src/stats.c:13-40
 * energy  pyLOM/vmmath/truncation.py:41-60; src/truncation.c:128-187
 * out[0] = sum |A-B|^2, out[1] = sum |A|^2 (host doubles; synchronous).     */
int pc_diff_norms(pc_ctx* ctx, pc_dtype dtype, const void* A, int64_t lda, const void* B, int64_t ldb,
                  int64_t m, int64_t n, double* out);
/* compute_truncation_residual  pyLOM/vmmath/truncation.py:17-39,
 * src/truncation.c:14-74 (host function; S host array, r<0 returns i+1).   */
int pc_truncation_residual(const double* S, int64_t n, double r, int64_t* N);

/* ---- batched (grouped) solves --------------------------------------------
 * The per-frequency loop of pyLOM/SPOD/wrapper.py:81-86 (tsqr_svd per block)
 * and any list of independent POD problems in one call.  Every problem is
 * validated first; the statuses are read once at the end, and an error names
 * the problem.  A uniform complex group (one shape, leading dimensions and
 * centring; A, U, S and VT at constant strides; n <= 32) runs every kernel as
 * one launch over all problems on the context stream, one workspace slice per
 * problem (falling back to the path below when the slices exceed the HBM
 * budget).  Other groups are taken from a queue, most expensive first, by
 * internal streams forked from and joined back to the context stream: one per
 * HIP hardware queue (GPU_MAX_HW_QUEUES, 4 when unset, at most 16 used).
 * Results are identical, bit for bit, to pc_tsqr_svd on each problem alone. */
typedef struct {
  const void* A; int64_t lda; int64_t m; int64_t n; int32_t center; int32_t pad_;
  void* U; int64_t ldu; void* S; void* VT; int64_t ldvt;
} pc_problem;
int pc_grouped_tsqr_svd(pc_ctx* ctx, pc_dtype dtype, const pc_problem* problems, int64_t count,
                        int precision);

/* Captured grouped solves (no reference counterpart: repeated SPOD / POD of
 * fixed shapes, e.g. new snapshots written into the same buffers).
 * pc_grouped_graph_create records pc_grouped_tsqr_svd of this problem list
 * (same pointers, shapes and precision) into a HIP graph that owns its own
 * workspaces; pc_grouped_graph_launch replays it on the context stream and
 * returns when the results are ready, naming a failing problem like the
 * direct call.  Inside a graph every block-Jacobi sweep is recorded (no host
 * check); converged sweeps return at once.  The buffers of the problems must
 * outlive the graph.                                                        */
typedef struct pc_graph pc_graph;
int pc_grouped_graph_create(pc_ctx* ctx, pc_dtype dtype, const pc_problem* problems, int64_t count,
                            int precision, pc_graph** out);
int pc_grouped_graph_launch(pc_ctx* ctx, pc_graph* graph);
int pc_graph_destroy(pc_graph* graph);

/* ---- SPOD front end --------------------------------------------------------
 * Welch blocks of pyLOM/SPOD/wrapper.py:39-82: per-row mean removal (FP64
 * means), Hamming window (pyLOM/vmmath/fft.py:17-27), real DFT of each block
 * of nDFT samples with step nDFT-nolap, scaled by winWeight/nDFT (x2 for the
 * interior frequencies) and 1/sqrt(nBlks).  X: M x N float32, rows are points,
 * columns time samples.  Q receives nf = ceil(nDFT/2)+1 complex64 matrices of
 * M x nBlks, frequency f at Q + f*M*nBlks, row-major (lda = nBlks): the
 * operands of the per-frequency SVD (pc_grouped_tsqr_svd with PC_C64).
 * Requires 2 <= nDFT <= min(N, 4096), 0 <= nolap < nDFT.                    */
int pc_spod_blocks(pc_ctx* ctx, const float* X, int64_t ldx, int64_t M, int64_t N, int64_t nDFT,
                   int64_t nolap, int remove_mean, void* Q);

/* SPOD modes (pyLOM/SPOD/wrapper.py:84 and the sort of wrapper.py:90-94):
 * P[(j*M + i)*nf + k] = Re U[order[k]][i][j], i.e. column k of the
 * (nBlks*M) x nf matrix P is the real part of frequency order[k]'s M x nBlks
 * left singular vectors flattened column-major.  U: nf complex64 M x nBlks
 * matrices, contiguous (as pc_spod_blocks lays them out); order: nf int64
 * indices on the device; P float32 (out_f64 = 0) or float64 (out_f64 = 1).
 * Queued on the context stream; nBlks <= 12288.                               */
int pc_spod_modes(pc_ctx* ctx, const void* U, int64_t M, int64_t nBlks, int64_t nf, const int64_t* order, void* P,
                  int out_f64);

/* ---- data assembly ---------------------------------------------------------
 * Dataset.X  pyLOM/dataset.py:252-292: the snapshot matrix of the given
 * fields, rows interleaved per point: row p * nc + c (nc = sum of ndim) is
 * component idim of field f for point p, i.e. row idim + ndim[f] * p of
 * field f (src[f]: ndim[f] * npoints rows of nt values, leading dimension
 * ld[f]).  X: npoints * nc rows of nt values (ldx).  Host arrays of device
 * pointers; at most 32 fields; float32 or float64.                          */
int pc_dataset_x(pc_ctx* ctx, pc_dtype dtype, int64_t nfields, const void* const* src, const int64_t* ndim,
                 const int64_t* ld, int64_t npoints, int64_t nt, void* X, int64_t ldx);

/* ---- DMD ---------------------------------------------------------------------
 * eigen  pyLOM/vmmath/maths.py:161-183 (numpy.linalg.eig = LAPACK ?geev):
 *   eigenvalues wr + i wi and right eigenvectors V (n x n complex128, ldv) of
 *   the real n x n matrix A (FP64, row-major).  Pairs come with the positive
 *   imaginary part first; vectors have unit 2-norm, a pair's largest
 *   component is real (?geev's normalisation).  The order of the eigenvalues
 *   follows the real Schur form this solver reaches, which can differ from
 *   LAPACK's.  Synchronous (status read).                                   */
int pc_eig(pc_ctx* ctx, const double* A, int64_t lda, int64_t n, double* wr, double* wi, void* V, int64_t ldv);
/* The projected operator of DMD.run (pyLOM/DMD/wrapper.py:85-88):
 *   Atilde = (U^T Y1) VT^T diag(1/S), U m x N, Y1 = Y[:, 1:] m x ns,
 *   S N, VT N x ns (the truncated SVD of Y[:, :-1]); Atilde N x N, FP64.    */
int pc_dmd_operator(pc_ctx* ctx, const double* U, int64_t ldu, const double* Y1, int64_t ldy, int64_t m, int64_t N,
                    int64_t ns, const double* S, const double* VT, int64_t ldvt, double* Atilde, int64_t ldat);
/* Modes, amplitudes and order of DMD.run (pyLOM/DMD/wrapper.py:90-117,
 * _order_modes :18-46): from (wr, wi, W) = eig(Atilde) (W N x N complex128,
 * ld N), the amplitudes b of Jovanovic et al. (2014) (P = (W^H W) o
 * conj(Vand Vand^H), q = conj(diag(Vand G^H W)), G = diag(S) VT, b = P^-1 q),
 * then everything ordered by decreasing |b| with each conjugate pair
 * positive-imaginary first: mu_re / mu_im (N), b (N complex128) and the modes
 * Phi = Y1 VT^T diag(1/S) W / mu (m x N complex128, ldphi), one pass over
 * Y1.  Synchronous (status read).                                         */
int pc_dmd_modes(pc_ctx* ctx, const double* Y1, int64_t ldy, int64_t m, int64_t ns, int64_t N, const double* S,
                 const double* VT, int64_t ldvt, const double* wr, const double* wi, const void* W, double* mu_re,
                 double* mu_im, void* Phi, int64_t ldphi, void* b);
/* reconstruction_jovanovic  pyLOM/DMD/wrapper.py:141-147:
 *   X = Re(Phi diag(b) Vand(t)), Vand[k][j] = mu_k^{t_j} (t: nt times on the
 *   device), X m x nt FP64 (ldx).                                          */
int pc_dmd_reconstruct(pc_ctx* ctx, const void* Phi, int64_t ldphi, int64_t m, int64_t N, const double* mu_re,
                       const double* mu_im, const void* b, const double* t, int64_t nt, double* X, int64_t ldx);

/* ---- launch tuning ------------------------------------------------------------
 * No reference counterpart (tile shapes of the GPU passes).  The product pass
 * U = Y W of pc_tsqr_svd / pc_tsqr / POD is row-local, so the number of
 * workgroups its row tiles are spread over changes its speed, never its
 * result.  pc_autotune(ctx, 1) makes the first such pass of a new (dtype, m,
 * n, precision) on ctx time every candidate chunk count twice on the real
 * operands and keep the fastest in a process-wide table; later solves of that
 * shape on any context use it (batched or graph-captured launches only read
 * it).  pc_autotune_query gives the tuned chunk count (returns 1) or 0 when
 * the shape is not tuned; pc_autotune_clear empties the table.             */
int pc_autotune(pc_ctx* ctx, int enable);
int pc_autotune_query(pc_dtype dtype, int64_t m, int64_t n, int precision, int64_t* chunks);
int pc_autotune_clear(void);

/* ---- serving -----------------------------------------------------------------
 * No reference counterpart (pyLOM runs one POD per script): a native request
 * loop for callers that serve many POD, projection and reconstruction
 * requests on one GPU.  Each request is one call of pc_tsqr_svd (with the
 * server's centring), pc_project or pc_reconstruct (both FP64) on a slot's own
 * context and stream, so its result is bit for bit the one that call gives
 * alone.  A request is admitted, first in first out with at most skip_ahead
 * POD requests passing a head that does not fit (projections and
 * reconstructions, which pin their basis from submission, may always pass),
 * as soon as a slot is idle and its output blocks fit in the output arena
 * beside the resident bases and the unreleased outputs; idle bases are
 * evicted least recently used first.  Unreleased outputs and pinned bases are
 * never evicted: a request that only fits once the caller releases or unpins
 * waits.  POD results stay resident as bases (U m x n, S n, VT n x n,
 * row-major, the request's dtype) until evicted.  budget_bytes is reserved at
 * creation: each slot's workspace for the largest POD (m_max x n_max) either
 * precision may run, and the rest as one device allocation (the output arena)
 * from which every output and basis block is carved (4 KiB granules), so
 * serving calls neither hipMalloc nor hipFree (PC_ERR_NOMEM when the
 * workspaces alone exceed budget_bytes or the device cannot hold it); a mixed server runs
 * float64 or n > 256 requests in FP64.  Input X is read on the device until the
 * request is done: the caller keeps it alive.  request->stream names the
 * stream X is produced on (NULL: the null stream): pc_serve_submit records an
 * event there and the request's slot waits for it before reading X, so X may
 * still be in flight at submission.  Calls are thread-safe.
 * pc_serve_wait returns PC_OK once the request is done (its outcome in
 * result->status), or PC_PENDING when timeout_s (>= 0) expires first;
 * timeout_s < 0 waits without limit.  A projection / reconstruction output
 * (result->out, device, out_rows x out_cols, row stride out_ld) belongs to the
 * server and is counted against the budget until pc_serve_release.
 * pc_serve_basis gives the device pointers of a resident basis and pins
 * (pin = 1) or unpins (pin = -1) it against eviction.  pc_serve_stats fills up
 * to 9 counters: completed, rejected, evictions, in-flight output bytes,
 * resident bytes (bases and unreleased outputs), queued, running, bases, and
 * the bytes left for outputs and bases once the workspaces are reserved.    */
#define PC_PENDING 1
typedef struct pc_server pc_server;
typedef enum { PC_REQ_POD = 0, PC_REQ_PROJECT = 1, PC_REQ_RECONSTRUCT = 2 } pc_req_kind;
typedef struct {
  int32_t kind;    /* pc_req_kind                                            */
  int32_t dtype;   /* PC_F32 / PC_F64 of X (pod, project)                    */
  const void* X;   /* pod: the snapshots m x n; project: m x n onto the basis */
  int64_t ldx, m, n;
  int64_t basis;   /* project / reconstruct: resident basis id               */
  int64_t modes;   /* project / reconstruct: leading modes used (0 = all)    */
  void* stream;    /* hipStream_t X is produced on (NULL: the null stream)   */
} pc_request;
typedef struct {
  int32_t status;  /* PC_OK or the request's pc_status                       */
  int32_t pad_;
  int64_t basis;   /* pod: id of the basis holding (U, S, VT)                */
  void* out;       /* project: modes x n; reconstruct: m x n                 */
  int64_t out_rows, out_cols, out_ld;
  double t_submit, t_admit, t_done;  /* seconds since the server was created */
} pc_result;
typedef struct {
  void* U; void* S; void* VT;
  int64_t m, n;
  int32_t dtype; int32_t pad_;
} pc_basis_view;
int pc_serve_create(int device, int slots, int64_t budget_bytes, int skip_ahead, int precision, int center,
                    int64_t m_max, int64_t n_max, pc_server** out);
int pc_serve_destroy(pc_server* s);
int pc_serve_submit(pc_server* s, const pc_request* req, int64_t* ticket);
int pc_serve_wait(pc_server* s, int64_t ticket, double timeout_s, pc_result* result);
int pc_serve_release(pc_server* s, int64_t ticket);
int pc_serve_drain(pc_server* s, double timeout_s);
int pc_serve_basis(pc_server* s, int64_t basis, int pin, pc_basis_view* view);
/* ids of the resident bases (up to cap of them, increasing id) and their count */
int pc_serve_bases(pc_server* s, int64_t* ids, int64_t cap, int64_t* count);
int pc_serve_stats(pc_server* s, int64_t* counters, int n);
const char* pc_serve_last_error(const pc_server* s);

#ifdef __cplusplus
}
#endif
#endif /* PODCORE_H */
