// FP64 dense kernels for real n x n matrices, n <= 256 (the small dense
// stage of the POD core).  Shifted Cholesky, pivoted Cholesky and Householder
// QR with explicit Q are blocked single-workgroup factorizations (1024
// threads): matrices live in global memory (L2-resident at these sizes),
// 32-column panels are staged in LDS and every rank-32 update runs on
// v_mfma_f64_16x16x4f64.  The triangular inverse is n independent column
// back-substitutions, one wavefront each.
//
// f64 MFMA 16x16x4 lane layout (wave64): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator register r holds D[4 r + l / 16][l % 16].
// Included by solve.hip after k_small.hip (uses NT, block_max, mfma64s, d4s).

namespace dense {

constexpr int PW = 32;           // panel width
constexpr int PLD = 257;         // LDS leading dimension of panels (odd: no bank conflicts)
constexpr int NMAX = 256;

// (I, J), I <= J, of the t-th upper tile of an nt x nt tile grid
PC_DEV void upper_tile(int t, int nt, int& I, int& J) {
  int i = 0;
  while (t >= nt - i) { t -= nt - i; i++; }
  I = i;
  J = i + t;
}

// DPP move of a double (both 32-bit halves), lanes that the row mask leaves
// out return `old`
template <int CTRL, int ROWS>
PC_DEV double dpp_f64(double x, double old) {
  const long long xi = __builtin_bit_cast(long long, x), oi = __builtin_bit_cast(long long, old);
  const int lo = __builtin_amdgcn_update_dpp(int(oi), int(xi), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(int(oi >> 32), int(xi >> 32), CTRL, ROWS, 0xF, false);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
PC_DEV double readlane63_f64(double x) {
  const long long xi = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane(int(xi), 63), hi = __builtin_amdgcn_readlane(int(xi >> 32), 63);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
// Sum over the 64 lanes of a wave, returned to every lane: within each row
// of 16 lanes two quad permutations and two rotations (by 4, by 8) leave the
// row's sum in all its lanes; row 0 is added into row 1 and row 2 into row 3
// (row_bcast:15), then row 1 into rows 2 and 3 (row_bcast:31), so lane 63
// holds every lane once.  Call with the whole wave active.
PC_DEV double wave_sum_dpp(double v) {
  v += dpp_f64<0xB1, 0xF>(v, 0.0);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xF>(v, 0.0);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x124, 0xF>(v, 0.0);  // row_ror:4
  v += dpp_f64<0x128, 0xF>(v, 0.0);  // row_ror:8
  v += dpp_f64<0x142, 0xA>(v, 0.0);  // row_bcast:15 into rows 1, 3
  v += dpp_f64<0x143, 0xC>(v, 0.0);  // row_bcast:31 into rows 2, 3
  return readlane63_f64(v);
}
// Maximum over the 64 lanes (NaN-free input), returned to every lane
PC_DEV double wave_max_dpp(double v) {
  v = fmax(v, dpp_f64<0xB1, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x4E, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x124, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x128, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x142, 0xA>(v, v));
  v = fmax(v, dpp_f64<0x143, 0xC>(v, v));
  return readlane63_f64(v);
}

}  // namespace dense

// ---------------------------------------------------------------------------
// Shifted Cholesky, real: R upper with R^T R = G + tau I, tau = shift_rel *
// max G_ii.  Right-looking with 32-wide panels; a pivot below tau/2 (rounding
// made G indefinite at that scale) restarts with tau *= 8.
// status[0] = 1 on a non-finite Gram or when 20 restarts did not suffice.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_chol_fast(const double* __restrict__ G, int ldg, int n, double shift_rel,
                                                  double* __restrict__ R, int ldr, int* status) {
  using namespace dense;
  __shared__ double D[PW * (PW + 1)];
  __shared__ double P[PW * PLD];
  __shared__ double sh[32];
  __shared__ int restart;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  double dmax = 0.0;
  for (int i = tid; i < n; i += NT) dmax = fmax(dmax, G[int64_t(i) * ldg + i]);
  dmax = block_max(dmax, sh);
  if (!isfinite(dmax)) {
    if (tid == 0) status[0] = 1;
    dmax = 1.0;
  }
  double tau = shift_rel * dmax;
  if (!(tau > 0.0)) tau = 1.0;
  for (int attempt = 0; attempt < 20; attempt++) {
    for (int i = w; i < n; i += NT / 64)
      for (int j = l; j < n; j += 64) {
        double v = i <= j ? G[int64_t(i) * ldg + j] : 0.0;
        if (i == j) v += tau;
        R[int64_t(i) * ldr + j] = v;
      }
    if (tid == 0) restart = 0;
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += PW) {
      const int bs = min(PW, n - k0);
      for (int idx = tid; idx < PW * PW; idx += NT) {
        int i = idx >> 5, j = idx & 31;
        D[i * (PW + 1) + j] = (i < bs && j < bs && i <= j) ? R[int64_t(k0 + i) * ldr + k0 + j] : 0.0;
      }
      __syncthreads();
      for (int j = 0; j < bs; j++) {
        if (tid == 0) {
          double d = D[j * (PW + 1) + j];
          if (!isfinite(d)) { status[0] = 1; d = tau; }
          else if (d < 0.5 * tau) { restart = 1; d = tau; }
          D[j * (PW + 1) + j] = sqrt(d);
        }
        __syncthreads();
        if (restart) break;
        const double r = D[j * (PW + 1) + j];
        if (tid > j && tid < bs) D[j * (PW + 1) + tid] /= r;
        __syncthreads();
        {
          int i = tid >> 5, c = tid & 31;
          if (i > j && i <= c && c < bs) D[i * (PW + 1) + c] -= D[j * (PW + 1) + i] * D[j * (PW + 1) + c];
        }
        __syncthreads();
      }
      if (restart) break;
      for (int idx = tid; idx < PW * PW; idx += NT) {
        int i = idx >> 5, j = idx & 31;
        if (i < bs && j < bs && i <= j) R[int64_t(k0 + i) * ldr + k0 + j] = D[i * (PW + 1) + j];
      }
      // panel: rows k0..k0+bs-1 of R right of the block, D^T X = A
      const int c1 = k0 + bs, nt = n - c1;
      // (each thread owns one column of the panel; its column of P is its scratch)
      for (int c = tid; c < nt; c += NT) {
        for (int i = 0; i < PW; i++) {
          double s = 0.0;
          if (i < bs) {
            s = R[int64_t(k0 + i) * ldr + c1 + c];
            for (int t = 0; t < i; t++) s -= D[t * (PW + 1) + i] * P[t * PLD + c];
            s /= D[i * (PW + 1) + i];
            R[int64_t(k0 + i) * ldr + c1 + c] = s;
          }
          P[i * PLD + c] = s;
        }
      }
      __syncthreads();
      // trailing update (upper tiles): A[c1:, c1:] -= P^T P
      const int ntl = (nt + 15) >> 4;
      for (int t = w; t < ntl * (ntl + 1) / 2; t += NT / 64) {
        int I, J;
        upper_tile(t, ntl, I, J);
        const int col = 16 * J + lr;
        d4s acc;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = 16 * I + 4 * r + lk;
          acc[r] = (row < nt && col < nt) ? -R[int64_t(c1 + row) * ldr + c1 + col] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PW; k += 4) acc = mfma64s(P[(k + lk) * PLD + 16 * I + lr], P[(k + lk) * PLD + col], acc);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = 16 * I + 4 * r + lk;
          if (row < nt && col < nt) R[int64_t(c1 + row) * ldr + c1 + col] = -acc[r];
        }
      }
      __syncthreads();
    }
    int again = restart;
    __syncthreads();
    if (!again) {
      // strictly lower part
      for (int i = w; i < n; i += NT / 64)
        for (int j = l; j < i; j += 64) R[int64_t(i) * ldr + j] = 0.0;
      return;
    }
    tau *= 8.0;
  }
  if (tid == 0) status[0] = 1;
}

// ---------------------------------------------------------------------------
// Pivoted Cholesky with rank detection, real (n <= 256): P^T G P = R^T R.
// The factorisation runs in physical (unpermuted) column indices, so nothing
// is swapped: row j of R picks the unpivoted index p with the largest Schur
// diagonal, R_phys[j][q] = (A[p][q] - sum_{t in panel, t<j} R_phys[t][p]
// R_phys[t][q]) / sqrt(d_p); earlier panels are folded into A by MFMA rank-32
// updates.  At the end R[t][c] = R_phys[t][perm[c]].  Stops when the largest
// remaining diagonal is <= tol * max(diag G) (rank); rows >= rank are zero and
// perm[rank..] lists the unpivoted indices in increasing order.
// A: n x n workspace (upper triangle used, physical indices).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_pchol_fast(const double* __restrict__ G, int ldg, int n, double tol,
                                                   double* __restrict__ A, double* __restrict__ R, int ldr,
                                                   int* __restrict__ perm, int* __restrict__ rank_out,
                                                   int* __restrict__ status) {
  using namespace dense;
  __shared__ double P[PW * PLD];
  __shared__ double dg[NMAX];
  __shared__ double wv_best[2 * (NT / 64)];
  __shared__ int wv_idx[2 * (NT / 64)];
  __shared__ double sh[32];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  for (int i = w; i < n; i += NT / 64)
    for (int j = l; j < n; j += 64) A[int64_t(i) * n + j] = i <= j ? G[int64_t(i) * ldg + j] : 0.0;
  for (int q = tid; q < n; q += NT) dg[q] = G[int64_t(q) * ldg + q];
  double dmax = 0.0;
  for (int q = tid; q < n; q += NT) dmax = fmax(dmax, G[int64_t(q) * ldg + q]);
  dmax = block_max(dmax, sh);
  if (!(dmax > 0.0) || !isfinite(dmax)) {
    for (int i = w; i < n; i += NT / 64)
      for (int j = l; j < n; j += 64) R[int64_t(i) * ldr + j] = 0.0;
    for (int q = tid; q < n; q += NT) perm[q] = q;
    if (tid == 0) { rank_out[0] = 0; if (!isfinite(dmax)) atomicOr(status, 16); }
    return;
  }
  const double thr = tol * dmax;
  int rank = n;
  // Each of the first n threads owns one index q = tid and keeps its Schur
  // diagonal dq in a register (-2 once pivoted, -3 for threads beyond n).
  // A pivot step is one barrier: the owners form their entry of the new row
  // and update dq, and every wave leaves its (max dq, lowest index at that
  // max) in one of two LDS slots, read by all threads after the barrier to
  // pick the next pivot (largest dq, ties to the lower index).
  const int nw = (n + 63) >> 6;
  const bool own = tid < n;
  double dq = own ? dg[tid] : -3.0;
  auto wave_argmax = [&](int slot) {
    const double m = wave_max_dpp(dq);
    const uint64_t at = __ballot(dq == m);
    if (l == 0) {
      wv_best[slot * (NT / 64) + w] = m;
      wv_idx[slot * (NT / 64) + w] = 64 * w + (at ? __builtin_ctzll(at) : 63);
    }
  };
  wave_argmax(0);
  __syncthreads();
  int slot = 0;
  for (int k0 = 0; k0 < n; k0 += PW) {
    const int bs = min(PW, n - k0);
    int jj = 0;
    for (; jj < bs; jj++) {
      const int j = k0 + jj;
      // all candidates are read at once (n <= 256: at most four waves own indices)
      double cb[NMAX / 64];
      int ci[NMAX / 64];
#pragma unroll
      for (int q = 0; q < NMAX / 64; q++) {
        cb[q] = wv_best[slot * (NT / 64) + q];
        ci[q] = wv_idx[slot * (NT / 64) + q];
      }
      double b = cb[0];
      int p = ci[0];
#pragma unroll
      for (int q = 1; q < NMAX / 64; q++)
        if (q < nw && cb[q] > b) { b = cb[q]; p = ci[q]; }
      if (!(b > thr) || p >= n) break;  // the same decision in every thread
      if (tid == 0) perm[j] = p;
      if (own) {
        const int q = tid;
        double v;
        const double r = sqrt(b);
        if (q == p) v = r;
        else if (dq < 0.0) v = 0.0;  // already pivoted
        else {
          double sacc = p <= q ? A[int64_t(p) * n + q] : A[int64_t(q) * n + p];
#pragma unroll 8
          for (int t = 0; t < jj; t++) sacc -= P[t * PLD + p] * P[t * PLD + q];
          v = sacc / r;
        }
        P[jj * PLD + q] = v;
        if (q == p) dq = -2.0;
        else if (dq >= 0.0) dq = fmax(dq - v * v, 0.0);
      }
      slot ^= 1;
      wave_argmax(slot);
      __syncthreads();
    }
    if (jj < bs) rank = k0 + jj;
    // rows of this panel (zero beyond the rows computed)
    for (int t = w; t < PW; t += NT / 64)
      for (int q = l; q < n; q += 64) {
        if (t >= jj) P[t * PLD + q] = 0.0;
        if (t < bs) R[int64_t(k0 + t) * ldr + q] = t < jj ? P[t * PLD + q] : 0.0;
      }
    __syncthreads();
    if (jj < bs) break;
    // fold this panel into A (upper tiles, physical indices)
    const int ntl = (n + 15) >> 4;
    for (int t = w; t < ntl * (ntl + 1) / 2; t += NT / 64) {
      int I, J;
      upper_tile(t, ntl, I, J);
      const int col = 16 * J + lr;
      d4s acc;
#pragma unroll
      for (int r4 = 0; r4 < 4; r4++) {
        int row = 16 * I + 4 * r4 + lk;
        acc[r4] = (row < n && col < n) ? -A[int64_t(row) * n + col] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < PW; k += 4) acc = mfma64s(P[(k + lk) * PLD + 16 * I + lr], P[(k + lk) * PLD + col], acc);
#pragma unroll
      for (int r4 = 0; r4 < 4; r4++) {
        int row = 16 * I + 4 * r4 + lk;
        if (row < n && col < n) A[int64_t(row) * n + col] = -acc[r4];
      }
    }
    __syncthreads();
  }
  // rows at and beyond the rank are zero; perm completes with the unpivoted indices
  for (int t = rank + w; t < n; t += NT / 64)
    for (int q = l; q < n; q += 64) R[int64_t(t) * ldr + q] = 0.0;
  if (own) dg[tid] = dq;
  __syncthreads();
  if (tid == 0) {
    int k = rank;
    for (int q = 0; q < n; q++)
      if (dg[q] >= 0.0 && k < n) perm[k++] = q;
    rank_out[0] = rank;
  }
  __syncthreads();
  // R[t][c] = R_phys[t][perm[c]] for c >= t, in chunks of 32 rows through LDS
  for (int t0 = 0; t0 < rank; t0 += PW) {
    for (int t = w; t < PW; t += NT / 64)
      for (int q = l; q < n; q += 64) P[t * PLD + q] = (t0 + t < rank) ? R[int64_t(t0 + t) * ldr + q] : 0.0;
    __syncthreads();
    for (int t = w; t < PW && t0 + t < rank; t += NT / 64)
      for (int c = l; c < n; c += 64) {
        int pc = perm[c];
        pc = pc < 0 ? 0 : (pc >= n ? n - 1 : pc);
        R[int64_t(t0 + t) * ldr + c] = c >= t0 + t ? P[t * PLD + pc] : 0.0;
      }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Householder QR of a real n x n matrix (n <= 256) in one launch on one
// workgroup: the same panels, block reflectors and MFMA sequence as the
// launch sequence below (k_hqr_panel / k_hqr_strip / k_hqr_rq), so Q and R
// are the same; it serves solves that share the GPU with other streams,
// where one launch costs less than the sequence's 3 npan launches.
// LAPACK conventions (beta = -sign(alpha) ||x||, tau = (beta - alpha) / beta).
// work: n*n (the factored matrix) + n (tau) + 32*32*ceil(n/32) (T).  The
// K-loops over global memory load eight k-steps ahead of their products.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_house_qr_fast(const double* __restrict__ A, int lda, int n,
                                                      double* __restrict__ Q, int ldq, double* __restrict__ R,
                                                      int ldr, double* __restrict__ work) {
  using namespace dense;
  __shared__ double Pn[PW * PLD];  // panel columns (column c at Pn[c * PLD], rows relative to the panel start)
  __shared__ double Xs[PW * PLD];  // X = V^T A_t (row i at Xs[i * PLD]); also V^T V
  __shared__ double Ts[PW * (PW + 1)];
  __shared__ double tau_s[PW], sc_s[PW], beta_s[PW];
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  double* Wm = work;
  double* tau = work + int64_t(n) * n;
  double* Tg = tau + n;
  for (int i = w; i < n; i += NW)
    for (int j = l; j < n; j += 64) Wm[int64_t(i) * n + j] = A[int64_t(i) * lda + j];
  __syncthreads();
  const int npan = (n + PW - 1) / PW;
  // V element (row i of the panel, column c), unit diagonal, zero above
  auto vget = [&](int i, int c) -> double { return i < c ? 0.0 : (i == c ? 1.0 : Pn[c * PLD + i]); };
  // X = V^T B for a K = mr tall B (row k, column col at B[k * ldb]), two
  // tile rows of 16 x ntl tiles, KD k-steps of loads ahead of the products
  constexpr int KD = 4;
  auto vt_times = [&](const double* B, int64_t ldb, int mr, int bs, int ncols, int ntl) {
    for (int t = w; t < 2 * ntl; t += NW) {
      const int I = t / ntl, J = t % ntl, col = 16 * J + lr, ca = 16 * I + lr;
      d4s acc = {0.0, 0.0, 0.0, 0.0};
      const double* Bc = B + col;
      for (int k0 = 0; k0 < mr; k0 += 4 * KD) {
        double bv[KD];
#pragma unroll
        for (int u = 0; u < KD; u++) {
          const int k = k0 + 4 * u + lk;
          bv[u] = (k < mr && col < ncols) ? Bc[int64_t(k) * ldb] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < KD; u++) {
          const int k = k0 + 4 * u + lk;
          acc = mfma64s((k < mr && ca < bs) ? vget(k, ca) : 0.0, bv[u], acc);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) Xs[(16 * I + 4 * r + lk) * PLD + col] = acc[r];
    }
  };
  // X <- T^T X (trans) or T X, on MFMA; every wave holds its (at most two)
  // result tiles in registers until all have read X
  auto t_times = [&](bool trans, int ntl) {
    d4s y[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int t = w + u * NW;
      y[u] = d4s{0.0, 0.0, 0.0, 0.0};
      if (t < 2 * ntl) {
        const int I = t / ntl, J = t % ntl, col = 16 * J + lr, i = 16 * I + lr;
#pragma unroll
        for (int k0 = 0; k0 < PW; k0 += 4) {
          const int k = k0 + lk;
          const double tv = trans ? Ts[k * (PW + 1) + i] : Ts[i * (PW + 1) + k];
          y[u] = mfma64s(tv, Xs[k * PLD + col], y[u]);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int t = w + u * NW;
      if (t < 2 * ntl) {
        const int I = t / ntl, J = t % ntl, col = 16 * J + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) Xs[(16 * I + 4 * r + lk) * PLD + col] = y[u][r];
      }
    }
    __syncthreads();
  };
  // C -= V X for an mr x ncols block C (row-major, ldc), two tiles per wave
  // at a time: both tiles' loads go out before their products
  auto minus_vx = [&](double* C, int64_t ldc, int mr, int bs, int ncols, int nml, int ntl) {
    const int ntiles = nml * ntl;
    for (int t = w; t < ntiles; t += 2 * NW) {
      d4s acc[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int tt = t + u * NW, I = tt / ntl, J = tt % ntl, col = 16 * J + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = 16 * I + 4 * r + lk;
          acc[u][r] = (tt < ntiles && row < mr && col < ncols) ? -C[int64_t(row) * ldc + col] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int tt = t + u * NW, I = tt / ntl, J = tt % ntl, col = 16 * J + lr, arow = 16 * I + lr;
        if (tt < ntiles) {
#pragma unroll
          for (int k0 = 0; k0 < PW; k0 += 4) {
            const int kc = k0 + lk;
            const double av = (arow < mr && kc < bs) ? vget(arow, kc) : 0.0;
            acc[u] = mfma64s(av, Xs[kc * PLD + col], acc[u]);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int tt = t + u * NW, I = tt / ntl, J = tt % ntl, col = 16 * J + lr;
        if (tt < ntiles) {
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int row = 16 * I + 4 * r + lk;
            if (row < mr && col < ncols) C[int64_t(row) * ldc + col] = -acc[u][r];
          }
        }
      }
    }
  };
  for (int pp = 0; pp < npan; pp++) {
    const int c0 = pp * PW, bs = min(PW, n - c0), mr = n - c0;
    for (int c = w; c < PW; c += NW)
      for (int i = l; i < PLD; i += 64) Pn[c * PLD + i] = (c < bs && i < mr) ? Wm[int64_t(c0 + i) * n + c0 + c] : 0.0;
    __syncthreads();
    for (int j = 0; j < bs; j++) {
      // every wave forms the reflector of column j itself (same arithmetic, same result)
      double s = 0.0;
      for (int i = j + 1 + l; i < mr; i += 64) s += Pn[j * PLD + i] * Pn[j * PLD + i];
      s = wave_sum_dpp(s);
      const double al = Pn[j * PLD + j];
      double tj, sc, beta;
      if (s == 0.0) {
        tj = 0.0; sc = 0.0; beta = al;
      } else {
        double anorm = sqrt(al * al + s);
        beta = al >= 0 ? -anorm : anorm;
        tj = (beta - al) / beta;
        sc = 1.0 / (al - beta);
      }
      if (tid == 0) { tau_s[j] = tj; sc_s[j] = sc; beta_s[j] = beta; }
      // apply H_j to the panel columns c > j (column j itself is read, not written)
      for (int c = j + 1 + w; c < bs; c += NW) {
        double d = 0.0;
        for (int i = j + l; i < mr; i += 64) {
          double v = i == j ? 1.0 : Pn[j * PLD + i] * sc;
          d += v * Pn[c * PLD + i];
        }
        d = wave_sum_dpp(d);
        double f = tj * d;
        for (int i = j + l; i < mr; i += 64) {
          double v = i == j ? 1.0 : Pn[j * PLD + i] * sc;
          Pn[c * PLD + i] -= v * f;
        }
      }
      __syncthreads();
    }
    // the panel's reflectors in place: v below the diagonal, beta on it
    for (int c = w; c < bs; c += NW) {
      const double sc = sc_s[c];
      for (int i = c + 1 + l; i < mr; i += 64) Pn[c * PLD + i] *= sc;
      if (l == 0) Pn[c * PLD + c] = beta_s[c];
    }
    __syncthreads();
    // V^T V (bs x bs) into Xs: four 16 x 16 tiles, one wave each, two
    // accumulators over alternating k-steps
    if (w < 4) {
      const int I = w >> 1, J = w & 1, ca = 16 * I + lr, cb = 16 * J + lr;
      d4s a0 = {0.0, 0.0, 0.0, 0.0}, a1 = {0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < mr; k0 += 8) {
        const int k = k0 + lk, k2 = k0 + 4 + lk;
        const double x0 = (k < mr && ca < bs) ? vget(k, ca) : 0.0;
        const double y0 = (k < mr && cb < bs) ? vget(k, cb) : 0.0;
        const double x1 = (k2 < mr && ca < bs) ? vget(k2, ca) : 0.0;
        const double y1 = (k2 < mr && cb < bs) ? vget(k2, cb) : 0.0;
        a0 = mfma64s(x0, y0, a0);
        a1 = mfma64s(x1, y1, a1);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) Xs[(16 * I + 4 * r + lk) * PLD + cb] = a0[r] + a1[r];
    }
    __syncthreads();
    // T of the compact WY form.  T^-1 = diag(1 / tau) + strictly upper
    // (V^T V) (Puglisi 1992), so in 16 x 16 blocks T11 and T22 come from back
    // substitution, one lane per column in registers: T_jj = tau_j,
    // T_ij = -tau_i sum_{k = i+1..j} (V^T V)_ik T_kj; multiplying by tau_i
    // (not dividing by 1 / tau_i) keeps a trivial reflector (tau_i = 0) exact,
    // its row and column of T zero.  Then T12 = -T11 (V^T V)_12 T22 on MFMA.
    if (w == 0) {
      auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      };
      {
        // lane jl of half-wave `base / 16` owns column jl of block (base, base);
        // its column lives in Ts (row stride PW + 1), rows from the bottom up
        const int base = 16 * ((l >> 4) & 1), jl = l & 15;
        double* tc = Ts + base * (PW + 1) + base + jl;
#pragma unroll 1
        for (int i = 15; i >= 0; i--) {
          double acc = 0.0;
#pragma unroll
          for (int k = 0; k < 16; k++)
            if (k > i && k <= jl) acc += Xs[(base + i) * PLD + base + k] * tc[k * (PW + 1)];
          const double ti = base + i < bs ? tau_s[base + i] : 0.0;
          const double v = i == jl ? ti : (i < jl ? -ti * acc : 0.0);
          if (l < 32) tc[i * (PW + 1)] = v;
        }
        if (l < 16)
          for (int i = 0; i < 16; i++) Ts[(16 + i) * (PW + 1) + jl] = 0.0;
      }
      wave_sync();
      // W = (V^T V)_12 T22 into the lower-left block of Xs (not read again)
      d4s acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = mfma64s(Xs[lr * PLD + 16 + k0 + lk], Ts[(16 + k0 + lk) * (PW + 1) + 16 + lr], acc);
#pragma unroll
      for (int r = 0; r < 4; r++) Xs[(16 + 4 * r + lk) * PLD + lr] = acc[r];
      wave_sync();
      acc = d4s{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = mfma64s(Ts[lr * (PW + 1) + k0 + lk], Xs[(16 + k0 + lk) * PLD + lr], acc);
#pragma unroll
      for (int r = 0; r < 4; r++) Ts[(4 * r + lk) * (PW + 1) + 16 + lr] = -acc[r];
    }
    __syncthreads();
    // store the panel (R part and V) and T
    for (int c = w; c < bs; c += NW)
      for (int i = l; i < mr; i += 64) Wm[int64_t(c0 + i) * n + c0 + c] = Pn[c * PLD + i];
    for (int idx = tid; idx < PW * PW; idx += NT) Tg[int64_t(pp) * PW * PW + idx] = Ts[(idx >> 5) * (PW + 1) + (idx & 31)];
    if (tid < bs) tau[c0 + tid] = tau_s[tid];
    // trailing update of columns c1.. (rows c0..)
    const int c1 = c0 + bs, nt = n - c1;
    if (nt > 0) {
      const int ntl = (nt + 15) >> 4, nml = (mr + 15) >> 4;
      vt_times(Wm + int64_t(c0) * n + c1, n, mr, PW, nt, ntl);
      __syncthreads();
      t_times(true, ntl);
      minus_vx(Wm + int64_t(c0) * n + c1, n, mr, PW, nt, nml, ntl);
    }
    __syncthreads();
  }
  // R
  for (int i = w; i < n; i += NW)
    for (int j = l; j < n; j += 64) R[int64_t(i) * ldr + j] = i <= j ? Wm[int64_t(i) * n + j] : 0.0;
  // Q = H_0 ... H_{n-1} I, accumulated backwards over the panels
  for (int i = w; i < n; i += NW)
    for (int j = l; j < n; j += 64) Q[int64_t(i) * ldq + j] = i == j ? 1.0 : 0.0;
  __syncthreads();
  for (int pp = npan - 1; pp >= 0; pp--) {
    const int c0 = pp * PW, bs = min(PW, n - c0), mr = n - c0, nq = n - c0;
    for (int c = w; c < PW; c += NW)
      for (int i = l; i < PLD; i += 64)
        Pn[c * PLD + i] = (c < bs && i < mr && i > c) ? Wm[int64_t(c0 + i) * n + c0 + c] : 0.0;
    for (int idx = tid; idx < PW * PW; idx += NT) Ts[(idx >> 5) * (PW + 1) + (idx & 31)] = Tg[int64_t(pp) * PW * PW + idx];
    __syncthreads();
    const int ntl = (nq + 15) >> 4, nml = (mr + 15) >> 4;
    double* Qs = Q + int64_t(c0) * ldq + c0;
    vt_times(Qs, ldq, mr, bs, nq, ntl);
    __syncthreads();
    t_times(false, ntl);
    minus_vx(Qs, ldq, mr, bs, nq, nml, ntl);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Householder QR of a real n x n matrix (n <= 256), A = Q R, Q explicit.
// LAPACK conventions (beta = -sign(alpha) ||x||, tau = (beta - alpha) / beta).
// Panels of 32 columns are factored in LDS; the block reflector
// I - V T V^T (compact WY, T upper) updates the trailing columns with MFMA
// (X = V^T A_t, X <- T^T X, A_t -= V X); Q is accumulated backwards the
// same way.  Latency is what costs in the panel (one workgroup): wave sums
// use DPP, a column of the panel costs one barrier (its scaling waits for
// the panel's end), V^T V and T's off-diagonal block run on MFMA.
// The factorisation runs in launches that spread its FP64 MFMA work: the
// panels are factored on one workgroup (k_hqr_panel: the column loop, V^T V,
// T), while the block-reflector updates, of the trailing columns after each
// panel and of Q (accumulated backwards) afterwards, run one workgroup per
// 32-column strip (k_hqr_strip, the strip of B staged in LDS).  A panel's
// update is ~3,800 FP64 MFMAs at 64 cycles each per SIMD, about 25 us on one
// CU.  The strips form the same tiles with the same MFMA sequence as one
// workgroup doing the whole update would, so Q and R do not depend on the
// split.  work: n*n (the factored matrix) + n (tau) + 32*32*ceil(n/32) (T).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_hqr_panel(const double* __restrict__ A, int lda, int n, int pp,
                                                  double* __restrict__ work) {
  using namespace dense;
  __shared__ double Pn[PW * PLD];  // panel columns (column c at Pn[c * PLD], rows relative to the panel start)
  __shared__ double Xs[PW * PLD];  // V^T V
  __shared__ double Ts[PW * (PW + 1)];
  __shared__ double tau_s[PW], sc_s[PW], beta_s[PW];
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  double* Wm = work;
  double* tau = work + int64_t(n) * n;
  double* Tg = tau + n;
  if (pp == 0) {
    for (int i = w; i < n; i += NW)
      for (int j = l; j < n; j += 64) Wm[int64_t(i) * n + j] = A[int64_t(i) * lda + j];
    __syncthreads();
  }
  auto vget = [&](int i, int c) -> double { return i < c ? 0.0 : (i == c ? 1.0 : Pn[c * PLD + i]); };
  {
    const int c0 = pp * PW, bs = min(PW, n - c0), mr = n - c0;
    for (int c = w; c < PW; c += NW)
      for (int i = l; i < PLD; i += 64) Pn[c * PLD + i] = (c < bs && i < mr) ? Wm[int64_t(c0 + i) * n + c0 + c] : 0.0;
    __syncthreads();
    for (int j = 0; j < bs; j++) {
      // every wave forms the reflector of column j itself (same arithmetic, same result)
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NMAX / 64; q++) {  // rows j + 1 + l + 64 q < mr: the reads go out together
        const int i = j + 1 + l + 64 * q;
        if (i < mr) s += Pn[j * PLD + i] * Pn[j * PLD + i];
      }
      s = wave_sum_dpp(s);
      const double al = Pn[j * PLD + j];
      double tj, sc, beta;
      if (s == 0.0) {
        tj = 0.0; sc = 0.0; beta = al;
      } else {
        double anorm = sqrt(al * al + s);
        beta = al >= 0 ? -anorm : anorm;
        tj = (beta - al) / beta;
        sc = 1.0 / (al - beta);
      }
      if (tid == 0) { tau_s[j] = tj; sc_s[j] = sc; beta_s[j] = beta; }
      // apply H_j to the panel columns c > j (column j itself is read, not written)
      for (int c = j + 1 + w; c < bs; c += NW) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < NMAX / 64; q++) {
          const int i = j + l + 64 * q;
          if (i < mr) {
            double v = i == j ? 1.0 : Pn[j * PLD + i] * sc;
            d += v * Pn[c * PLD + i];
          }
        }
        d = wave_sum_dpp(d);
        double f = tj * d;
#pragma unroll
        for (int q = 0; q < NMAX / 64; q++) {
          const int i = j + l + 64 * q;
          if (i < mr) {
            double v = i == j ? 1.0 : Pn[j * PLD + i] * sc;
            Pn[c * PLD + i] -= v * f;
          }
        }
      }
      __syncthreads();
    }
    // the panel's reflectors in place: v below the diagonal, beta on it
    for (int c = w; c < bs; c += NW) {
      const double sc = sc_s[c];
      for (int i = c + 1 + l; i < mr; i += 64) Pn[c * PLD + i] *= sc;
      if (l == 0) Pn[c * PLD + c] = beta_s[c];
    }
    __syncthreads();
    // V^T V (bs x bs) into Xs: four 16 x 16 tiles, one wave each, two
    // accumulators over alternating k-steps
    if (w < 4) {
      const int I = w >> 1, J = w & 1, ca = 16 * I + lr, cb = 16 * J + lr;
      d4s a0 = {0.0, 0.0, 0.0, 0.0}, a1 = {0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < mr; k0 += 8) {
        const int k = k0 + lk, k2 = k0 + 4 + lk;
        const double x0 = (k < mr && ca < bs) ? vget(k, ca) : 0.0;
        const double y0 = (k < mr && cb < bs) ? vget(k, cb) : 0.0;
        const double x1 = (k2 < mr && ca < bs) ? vget(k2, ca) : 0.0;
        const double y1 = (k2 < mr && cb < bs) ? vget(k2, cb) : 0.0;
        a0 = mfma64s(x0, y0, a0);
        a1 = mfma64s(x1, y1, a1);
      }
  #pragma unroll
      for (int r = 0; r < 4; r++) Xs[(16 * I + 4 * r + lk) * PLD + cb] = a0[r] + a1[r];
    }
    __syncthreads();
    // T of the compact WY form.  T^-1 = diag(1 / tau) + strictly upper
    // (V^T V) (Puglisi 1992), so in 16 x 16 blocks T11 and T22 come from back
    // substitution, one lane per column in registers: T_jj = tau_j,
    // T_ij = -tau_i sum_{k = i+1..j} (V^T V)_ik T_kj; multiplying by tau_i
    // (not dividing by 1 / tau_i) keeps a trivial reflector (tau_i = 0) exact,
    // its row and column of T zero.  Then T12 = -T11 (V^T V)_12 T22 on MFMA.
    if (w == 0) {
      auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      };
      {
        // lane jl of half-wave `base / 16` owns column jl of block (base, base);
        // its column lives in Ts (row stride PW + 1), rows from the bottom up
        const int base = 16 * ((l >> 4) & 1), jl = l & 15;
        double* tc = Ts + base * (PW + 1) + base + jl;
  #pragma unroll 1
        for (int i = 15; i >= 0; i--) {
          double acc = 0.0;
  #pragma unroll
          for (int k = 0; k < 16; k++)
            if (k > i && k <= jl) acc += Xs[(base + i) * PLD + base + k] * tc[k * (PW + 1)];
          const double ti = base + i < bs ? tau_s[base + i] : 0.0;
          const double v = i == jl ? ti : (i < jl ? -ti * acc : 0.0);
          if (l < 32) tc[i * (PW + 1)] = v;
        }
        if (l < 16)
          for (int i = 0; i < 16; i++) Ts[(16 + i) * (PW + 1) + jl] = 0.0;
      }
      wave_sync();
      // W = (V^T V)_12 T22 into the lower-left block of Xs (not read again)
      d4s acc = {0.0, 0.0, 0.0, 0.0};
  #pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = mfma64s(Xs[lr * PLD + 16 + k0 + lk], Ts[(16 + k0 + lk) * (PW + 1) + 16 + lr], acc);
  #pragma unroll
      for (int r = 0; r < 4; r++) Xs[(16 + 4 * r + lk) * PLD + lr] = acc[r];
      wave_sync();
      acc = d4s{0.0, 0.0, 0.0, 0.0};
  #pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = mfma64s(Ts[lr * (PW + 1) + k0 + lk], Xs[(16 + k0 + lk) * PLD + lr], acc);
  #pragma unroll
      for (int r = 0; r < 4; r++) Ts[(4 * r + lk) * (PW + 1) + 16 + lr] = -acc[r];
    }
    __syncthreads();
    // store the panel (R part and V) and T
    for (int c = w; c < bs; c += NW)
      for (int i = l; i < mr; i += 64) Wm[int64_t(c0 + i) * n + c0 + c] = Pn[c * PLD + i];
    for (int idx = tid; idx < PW * PW; idx += NT) Tg[int64_t(pp) * PW * PW + idx] = Ts[(idx >> 5) * (PW + 1) + (idx & 31)];
    if (tid < bs) tau[c0 + tid] = tau_s[tid];
  }
}

// one workgroup per 32-column strip: qmode 0 updates the trailing columns of
// Wm after panel pp, qmode 1 applies panel pp's block reflector to Q
__global__ __launch_bounds__(256) void k_hqr_strip(int n, int pp, int qmode, double* __restrict__ work,
                                                   double* __restrict__ Q, int ldq) {
  using namespace dense;
  constexpr int NWS = 4, XLD = PW + 1, BLD = PW + 1;
  __shared__ double Pn[PW * PLD];
  __shared__ double Bs[NMAX * BLD];  // the strip of B (rows relative to the panel start)
  __shared__ double Xs[PW * XLD];
  __shared__ double Ts[PW * (PW + 1)];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lk = l >> 4;
  double* Wm = work;
  const double* Tg = work + int64_t(n) * n + n;
  const int c0 = pp * PW, bs = min(PW, n - c0), mr = n - c0;
  const int cst = qmode ? c0 : c0 + bs;        // first column of the updated block
  const int ncols = n - cst, s0 = 32 * blockIdx.x;
  if (s0 >= ncols) return;
  double* B = qmode ? Q + int64_t(c0) * ldq + cst : Wm + int64_t(c0) * n + cst;
  const int64_t ldb = qmode ? ldq : n;
  const int bsv = qmode ? bs : PW;             // reflector columns the products read
  // stage V (below the diagonal), T and the strip of B: every load of a
  // thread is issued before its first LDS store
  {
    constexpr int PER = PW * NMAX / 256;  // 32 elements of V and of B per thread
    double vv[PER], bb[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = tid + 256 * q, i = idx >> 5, c = idx & 31;  // row i, column c
      vv[q] = (c < bs && i < mr && i > c) ? Wm[int64_t(c0 + i) * n + c0 + c] : 0.0;
      bb[q] = (i < mr && s0 + c < ncols) ? B[int64_t(i) * ldb + s0 + c] : 0.0;
    }
    for (int idx = tid; idx < PW * PW; idx += 256) Ts[(idx >> 5) * (PW + 1) + (idx & 31)] = Tg[int64_t(pp) * PW * PW + idx];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = tid + 256 * q, i = idx >> 5, c = idx & 31;
      Pn[c * PLD + i] = vv[q];
      Bs[i * BLD + c] = bb[q];
    }
  }
  __syncthreads();
  auto vget = [&](int i, int c) -> double { return i < c ? 0.0 : (i == c ? 1.0 : Pn[c * PLD + i]); };
  // X = V^T B for the strip: wave w forms tile (I, Jl) = (w >> 1, w & 1)
  {
    const int I = w >> 1, Jl = w & 1, cl = 16 * Jl + lr, ca = 16 * I + lr;
    d4s acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < mr; k0 += 4) {
      const int k = k0 + lk;
      acc = mfma64s((k < mr && ca < bsv) ? vget(k, ca) : 0.0, k < mr ? Bs[k * BLD + cl] : 0.0, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) Xs[(16 * I + 4 * r + lk) * XLD + cl] = acc[r];
  }
  __syncthreads();
  // X <- T^T X (trailing) or T X (Q)
  {
    const int I = w >> 1, Jl = w & 1, i = 16 * I + lr;
    d4s y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < PW; k0 += 4) {
      const int k = k0 + lk;
      const double tv = qmode ? Ts[i * (PW + 1) + k] : Ts[k * (PW + 1) + i];
      y = mfma64s(tv, Xs[k * XLD + 16 * Jl + lr], y);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) Xs[(16 * I + 4 * r + lk) * XLD + 16 * Jl + lr] = y[r];
  }
  __syncthreads();
  // B -= V X: tiles (row block I, column block Jl) of the strip, B from LDS
  const int nml = (mr + 15) >> 4, ntiles = nml * 2;
  for (int t = w; t < ntiles; t += NWS) {
    const int I = t >> 1, Jl = t & 1, cl = 16 * Jl + lr, col = s0 + cl, arow = 16 * I + lr;
    d4s acc;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 16 * I + 4 * r + lk;
      acc[r] = row < mr ? -Bs[row * BLD + cl] : 0.0;
    }
#pragma unroll
    for (int k0 = 0; k0 < PW; k0 += 4) {
      const int kc = k0 + lk;
      const double av = (arow < mr && kc < bsv) ? vget(arow, kc) : 0.0;
      acc = mfma64s(av, Xs[kc * XLD + cl], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 16 * I + 4 * r + lk;
      if (row < mr && col < ncols) B[int64_t(row) * ldb + col] = -acc[r];
    }
  }
}

// R (upper part of the factored matrix) and Q = I before the backward accumulation
__global__ __launch_bounds__(256) void k_hqr_rq(int n, const double* __restrict__ work, double* __restrict__ R,
                                                int ldr, double* __restrict__ Q, int ldq) {
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= int64_t(n) * n) return;
  const int i = int(idx / n), j = int(idx % n);
  R[int64_t(i) * ldr + j] = i <= j ? work[idx] : 0.0;
  Q[int64_t(i) * ldq + j] = i == j ? 1.0 : 0.0;
}

// W = R^-1 for an upper triangular R (n <= 256, row-major), one wavefront per
// column j of W: the back substitution R w = e_j runs from i = j down to 0
// with the right-hand side r in registers (lane l owns rows l + 64 q):
//   w_i = r_i / R_ii,   r_k -= R_ki w_i  (k < i).
// The columns of R a stretch of 64 steps needs are staged in LDS, rows
// padded to 65 doubles so that the lanes of a wave read one column without
// bank conflicts.  The n columns are independent solves, one workgroup each.
__global__ __launch_bounds__(64) void k_trinv_cols(const double* __restrict__ R, int ldr, int n,
                                                   double* __restrict__ W, int ldw) {
  __shared__ double Rs[dense::NMAX * 65];
  const int j = blockIdx.x, l = threadIdx.x;
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  double r[4], w[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    r[q] = (l + 64 * q == j) ? 1.0 : 0.0;
    w[q] = 0.0;
  }
  for (int g = j >> 6; g >= 0; --g) {
    const int c0 = 64 * g;
    const int cl = j - c0 < 63 ? j - c0 : 63;  // last column of the stretch
    const int rows = c0 + cl + 1;              // rows 0 .. c0 + cl hold the upper triangle
    wave_sync();  // the previous stretch has been read
    if (l <= cl) {
#pragma unroll 8
      for (int k = 0; k < rows; k++) Rs[k * 65 + l] = R[int64_t(k) * ldr + c0 + l];
    }
    wave_sync();
    // The stretch's diagonal and its reciprocals, one division per lane in
    // parallel.  A step's quotient is then q0 = r y refined by one residual
    // step, q0 + (r - d q0) y (the last step of Newton-Markstein division),
    // so the division leaves the step's dependency chain.
    const double dgl = l <= cl ? Rs[(c0 + l) * 65 + l] : 1.0;
    const double dl = 1.0 / dgl;
    auto lane_read = [](double v, int t) {  // t is wave-uniform: a lane read, not a shuffle
      const long long b = __builtin_bit_cast(long long, v);
      return __builtin_bit_cast(double, (long long)(unsigned)__builtin_amdgcn_readlane(int(b), t) |
                                            ((long long)__builtin_amdgcn_readlane(int(b >> 32), t) << 32));
    };
    for (int t = cl; t >= 0; --t) {
      const int i = c0 + t;
      double rv = r[0];
#pragma unroll
      for (int q = 1; q < 4; q++) rv = g == q ? r[q] : rv;
      const double rt = lane_read(rv, t), d = lane_read(dgl, t), y = lane_read(dl, t);
      const double q0 = rt * y;
      const double wi = fma(fma(-d, q0, rt), y, q0);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (g == q && l == t) w[q] = wi;
        const int k = l + 64 * q;
        if (k < i) r[q] -= Rs[k * 65 + t] * wi;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int k = l + 64 * q;
    if (k < n) W[int64_t(k) * ldw + j] = k <= j ? w[q] : 0.0;
  }
}
