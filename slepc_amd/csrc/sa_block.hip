// Block BV kernels for gfx950 (MI355X): BVDot (M = Y^T X on MFMA, per-
// workgroup partial tiles reduced in a fixed order) and BVMatMult (Y = A X,
// a vector-CSR SpMM that reads each nonzero once for up to SPMM_NC columns).
// Reference: bvglobal.c:86-120 (BVDot), bvops.c:692-722 (BVMatMult).
#include <algorithm>

#include "sa_gs_device.h"

namespace sa {

// ---------------------------------------------------------------------------
// BVDot: M(i, j) = Y(:, i) . X(:, j), i < my <= 64, j < mx <= 64.  A
// workgroup walks its rows in 32-row chunks, stages the chunk of both
// blocks in LDS (stride 33: the 16 lanes of a tile row read 16 columns of
// one row without bank conflicts) and accumulates 16x16 tiles of M with
// MFMA 16x16x4 (operand A = Y^T, operand B = X); wave w owns tiles w, w+4,
// w+8, w+12.  The tile values of every workgroup go to part[(i + j*my) * G
// + b], summed afterwards in a fixed order (deterministic results).
// ---------------------------------------------------------------------------
constexpr int BD_ROWS = 32, BD_LDP = BD_ROWS + 1, BD_MAXC = 64, BD_MAXG = 1024;

template <typename T, bool NT, int PC>
__global__ __launch_bounds__(BLK) void bvdot_kernel(int64_t n, int64_t rpb, const T *__restrict__ X, int64_t ldx,
                                                    int mx, const T *__restrict__ Y, int64_t ldy, int my,
                                                    double *__restrict__ part, int G) {
  typedef typename Vec4<T>::type V4;
  __shared__ T Ys[PC * (BLK / BD_ROWS) * BD_LDP];   // tiles sized for the columns handled
  __shared__ T Xs[PC * (BLK / BD_ROWS) * BD_LDP];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, g = lane >> 4, c16 = lane & 15;
  const int nty = (my + 15) / 16, ntx = (mx + 15) / 16, ntile = nty * ntx;
  V4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; q++) acc[q] = (V4){0, 0, 0, 0};
  const int64_t rb = (int64_t)blockIdx.x * rpb, re = min(n, rb + rpb);
  const int row = t & (BD_ROWS - 1), cg = t / BD_ROWS;      // 8 column groups
  static_assert(PC * (BLK / BD_ROWS) <= BD_MAXC, "columns per thread and block");
  // register double buffer: the next chunk's loads are in flight while the
  // MFMA steps of the current one run (two chunks ahead measured slower: the
  // extra registers halve the occupancy)
  T py[PC], px[PC];
  auto fetch = [&](int64_t r0) {
    const bool in = r0 + row < re;
#pragma unroll
    for (int q = 0; q < PC; q++) {
      const int i = cg + q * (BLK / BD_ROWS);
      py[q] = (in && i < my) ? tile_load<NT>(Y + (size_t)i * ldy + r0 + row) : (T)0;
      px[q] = (in && i < mx) ? tile_load<NT>(X + (size_t)i * ldx + r0 + row) : (T)0;
    }
  };
  if (rb < re) fetch(rb);
  for (int64_t r0 = rb; r0 < re; r0 += BD_ROWS) {
#pragma unroll
    for (int q = 0; q < PC; q++) {
      const int i = cg + q * (BLK / BD_ROWS);
      if (i < my) Ys[i * BD_LDP + row] = py[q];
      if (i < mx) Xs[i * BD_LDP + row] = px[q];
    }
    __syncthreads();
    if (r0 + BD_ROWS < re) fetch(r0 + BD_ROWS);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tile = w + 4 * q;
      if (tile < ntile) {
        const int yi = (tile / ntx) * 16 + c16, xj = (tile % ntx) * 16 + c16;
#pragma unroll
        for (int st = 0; st < BD_ROWS / 4; st++) {
          const int kk = 4 * st + g;
          const T a = yi < my ? Ys[yi * BD_LDP + kk] : (T)0;
          const T b = xj < mx ? Xs[xj * BD_LDP + kk] : (T)0;
          acc[q] = mfma16x16x4(a, b, acc[q]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int tile = w + 4 * q;
    if (tile < ntile) {
      const int j = (tile % ntx) * 16 + c16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = (tile / ntx) * 16 + mfma_row<T>(g, v);
        if (i < my && j < mx) part[((size_t)i + (size_t)j * my) * G + blockIdx.x] = (double)acc[q][v];
      }
    }
  }
}

// out[col] = sum over the G partials of entry col, one workgroup per entry.
__global__ __launch_bounds__(BLK) void bvdot_sum_kernel(const double *part, int G, double *out) {
  gs_sum_column(part, G, (int)blockIdx.x, out);
}

template <typename T>
static int bvdot_block(const BV &X, int lx, int mx, const BV &Y, int ly, int my, double *part, int G, int64_t rpb,
                       double *out, hipStream_t s) {
  const T *xp = static_cast<const T *>(X.data) + (size_t)lx * X.ld;
  const T *yp = static_cast<const T *>(Y.data) + (size_t)ly * Y.ld;
  // blocks read once from HBM: non-temporal loads (as the Gram-Schmidt tiles)
  const bool nt = X.n * (int64_t)(mx + my) * (int64_t)sizeof(T) > GS_NT_BYTES;
  // registers for 4 columns per thread and block when both blocks have at
  // most 32 columns (the occupancy the loads in flight need), else 8
  const bool narrow = mx <= 4 * (BLK / BD_ROWS) && my <= 4 * (BLK / BD_ROWS);
  if (nt && narrow)
    hipLaunchKernelGGL((bvdot_kernel<T, true, 4>), dim3(G), dim3(BLK), 0, s, X.n, rpb, xp, X.ld, mx, yp, Y.ld, my, part, G);
  else if (nt)
    hipLaunchKernelGGL((bvdot_kernel<T, true, 8>), dim3(G), dim3(BLK), 0, s, X.n, rpb, xp, X.ld, mx, yp, Y.ld, my, part, G);
  else if (narrow)
    hipLaunchKernelGGL((bvdot_kernel<T, false, 4>), dim3(G), dim3(BLK), 0, s, X.n, rpb, xp, X.ld, mx, yp, Y.ld, my, part, G);
  else
    hipLaunchKernelGGL((bvdot_kernel<T, false, 8>), dim3(G), dim3(BLK), 0, s, X.n, rpb, xp, X.ld, mx, yp, Y.ld, my, part, G);
  hipLaunchKernelGGL(bvdot_sum_kernel, dim3(my * mx), dim3(BLK), 0, s, (const double *)part, G, out);
  SA_HIP(hipGetLastError());
  return 0;
}

// Partial-sum workspace of BVDot, one per host thread (each thread has its
// own stream): kept between calls, grown when a call needs more, released
// with the process (at most 32 MB; freeing it from a thread-exit destructor
// could run after the HIP runtime has shut down).
struct BdWork {
  double *part = nullptr, *blk = nullptr;
  int G = 0;
};

int launch_bvdot(const BV &X, int lx, int kx, const BV &Y, int ly, int ky, double *M_dev, hipStream_t s) {
  const int mx = kx - lx, my = ky - ly;
  if (mx <= 0 || my <= 0) return 0;
  SA_CHECK(X.prec == Y.prec && X.n == Y.n, "BVDot: bases differ in length or precision");
  const int64_t chunks = std::max<int64_t>(1, (X.n + BD_ROWS - 1) / BD_ROWS);
  const int G = (int)std::min<int64_t>(BD_MAXG, chunks);
  const int64_t rpb = (chunks + G - 1) / G * BD_ROWS;
  thread_local BdWork W;
  if (W.G < G) {
    if (W.part) (void)hipFree(W.part);
    W.part = nullptr;
    W.G = 0;
    SA_HIP(hipMalloc(&W.part, sizeof(double) * (size_t)BD_MAXC * BD_MAXC * G));
    W.G = G;
  }
  if (!W.blk) SA_HIP(hipMalloc(&W.blk, sizeof(double) * BD_MAXC * BD_MAXC));
  double *part = W.part, *blk = W.blk;
  int rc = 0;
  // blocks of at most 64 x 64 entries; block (bi, bj) lands in M(bi:, bj:)
  for (int bj = 0; bj < mx && !rc; bj += BD_MAXC)
    for (int bi = 0; bi < my && !rc; bi += BD_MAXC) {
      const int nx = std::min(BD_MAXC, mx - bj), ny = std::min(BD_MAXC, my - bi);
      rc = X.prec == SA_FP64 ? bvdot_block<double>(X, lx + bj, nx, Y, ly + bi, ny, part, G, rpb, blk, s)
                             : bvdot_block<float>(X, lx + bj, nx, Y, ly + bi, ny, part, G, rpb, blk, s);
      if (!rc && hipMemcpy2DAsync(M_dev + bi + (size_t)bj * my, sizeof(double) * my, blk, sizeof(double) * ny,
                                  sizeof(double) * ny, nx, hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = 1;
    }
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = 1;
  if (rc == 1) set_error("BVDot: HIP error");
  return rc;
}

// ---------------------------------------------------------------------------
// BVMatMult: Y(:, j) = A X(:, j) for j < nc <= SPMM_NC.  L lanes per row;
// each nonzero's value and column index are loaded once and applied to the
// nc columns, so the operator is read once per SPMM_NC columns.  The sum of
// each column runs in the order of the vector-CSR SpMV.  Operators larger
// than the caches take the stream layout below instead.
// ---------------------------------------------------------------------------
constexpr int SPMM_NC = 8;

template <typename T, int L>
__global__ __launch_bounds__(BLK) void spmm_kernel(int n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                   const T *__restrict__ val, const T *__restrict__ X, int64_t ldx,
                                                   T *__restrict__ Y, int64_t ldy, int nc) {
  const int lg = threadIdx.x % L;
  const int64_t row = ((int64_t)blockIdx.x * BLK + threadIdx.x) / L;
  T s[SPMM_NC];
#pragma unroll
  for (int j = 0; j < SPMM_NC; j++) s[j] = (T)0;
  if (row < n) {
    const int32_t pe = rp[row + 1];
    for (int32_t p = rp[row] + lg; p < pe; p += L) {
      const T v = val[p];
      const int32_t c = ci[p];
#pragma unroll
      for (int j = 0; j < SPMM_NC; j++)
        if (j < nc) s[j] += v * X[(size_t)j * ldx + c];
    }
  }
  // every lane of the group ends with each column's sum; lane q keeps the
  // columns q, q+L, ... so the stores need SPMM_NC/L instructions, not SPMM_NC
  constexpr int NS = (SPMM_NC + L - 1) / L;
  T mine[NS];
#pragma unroll
  for (int u = 0; u < NS; u++) mine[u] = (T)0;
#pragma unroll
  for (int j = 0; j < SPMM_NC; j++) {
    T sj = s[j];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sj += __shfl_xor(sj, off, L);
    if (j % L == lg) mine[j / L] = sj;
  }
  if (row < n) {
#pragma unroll
    for (int u = 0; u < NS; u++) {
      const int j = u * L + lg;
      if (j < nc) Y[(size_t)j * ldy + row] = mine[u];
    }
  }
}

// Stream SpMM for operators larger than the caches: a workgroup owns RB rows
// (the stream SpMV's row blocks, at most SPMV_NZB nonzeros each) and stages
// their values and column indices in LDS with coalesced non-temporal loads.
// Thread t serves row t % RB for the columns t / RB + u * TPR (TPR = BLK / RB
// threads per row), so a wave covers 64 consecutive rows of the same columns:
// for a stencil its gathers of x read neighbouring addresses and its stores
// are coalesced.  Each staged nonzero is read from LDS once per thread and
// applied to all of the thread's columns.  Each column's sum is formed as the
// stream SpMV forms it (TPR strided partial sums of rounded products, added in
// the order of its shuffle tree), so a column of BVMatMult equals MatMult.
// One row of the stream SpMM: nonzeros b..e-1 of (vs, cs) applied to NCT
// columns, into TPR strided partial sums per column (nonzero p goes to sum
// (p - b) % TPR, in order).  The nonzeros are taken TPR at a time: their
// values and indices first (indices clamped to the row, so every load is in
// bounds), then all their gathers of x (TPR * NCT = SPMM_NC loads in flight),
// then the sums.  Batches of 2 or 4 times as many nonzeros measured slower at
// 128^3 on the 27-point operator (more x in flight than the L2 holds).
// Products are rounded before they are added, as in the stream SpMV.
constexpr int SPMM_UNR = 1;

template <typename T, int TPR, int NCT, typename VP, typename CP>
__device__ __forceinline__ void spmm_row(VP vs, CP cs, int b, int e, const T *const (&xc)[NCT],
                                         T (&acc)[NCT][TPR]) {
#pragma clang fp contract(off)
  constexpr int NB = TPR * SPMM_UNR;
#pragma unroll
  for (int u = 0; u < NCT; u++)
#pragma unroll
    for (int q = 0; q < TPR; q++) acc[u][q] = (T)0;
  for (int p = b; p < e; p += NB) {
    T v[NB], xv[NB][NCT];
    int32_t c[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const int pk = min(p + k, e - 1);
      v[k] = vs[pk];
      c[k] = cs[pk];
    }
#pragma unroll
    for (int k = 0; k < NB; k++)
#pragma unroll
      for (int u = 0; u < NCT; u++) xv[k][u] = xc[u][c[k]];
#pragma unroll
    for (int k = 0; k < NB; k++)
      if (p + k < e)
#pragma unroll
        for (int u = 0; u < NCT; u++) acc[u][k % TPR] += v[k] * xv[k][u];
  }
}

template <typename T, int RB>
__global__ __launch_bounds__(BLK) void spmm_stream_kernel(int n, const int32_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci, const T *__restrict__ val,
                                                          const T *__restrict__ X, int64_t ldx, T *__restrict__ Y,
                                                          int64_t ldy, int nc) {
  constexpr int TPR = BLK / RB;
  constexpr int NCT = SPMM_NC / TPR;   // columns per thread
  static_assert(RB <= BLK && BLK % RB == 0 && SPMM_NC % TPR == 0, "row block");
  __shared__ T sval[SPMV_NZB];
  __shared__ int32_t scol[SPMV_NZB];
  __shared__ int32_t srp[RB + 1];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * RB;
  const int rows = (int)min((int64_t)RB, (int64_t)n - r0);
  for (int i = t; i <= rows; i += BLK) srp[i] = rp[r0 + i];
  __syncthreads();
  const int32_t p0 = srp[0], nz = srp[rows] - p0;
  const bool staged = nz <= SPMV_NZB;    // uniform over the workgroup
  if (staged)
    for (int i = t; i < nz; i += BLK) {
      sval[i] = __builtin_nontemporal_load(val + p0 + i);
      scol[i] = __builtin_nontemporal_load(ci + p0 + i);
    }
  __syncthreads();
  const int r = t % RB, jg = t / RB;
  if (r >= rows || jg >= nc) return;
  const int b = srp[r] - p0, e = srp[r + 1] - p0;
  const T *xc[NCT];   // columns beyond nc repeat column nc - 1 (loaded, never stored)
#pragma unroll
  for (int u = 0; u < NCT; u++) xc[u] = X + (size_t)min(jg + u * TPR, nc - 1) * ldx;
  T acc[NCT][TPR];
  if (staged)
    spmm_row<T, TPR, NCT>(sval, scol, b, e, xc, acc);
  else
    spmm_row<T, TPR, NCT>(val + p0, ci + p0, b, e, xc, acc);
#pragma unroll
  for (int u = 0; u < NCT; u++) {
#pragma unroll
    for (int h = TPR / 2; h > 0; h >>= 1)
#pragma unroll
      for (int q = 0; q < h; q++) acc[u][q] = acc[u][q] + acc[u][q + h];
    if (jg + u * TPR < nc) Y[(size_t)(jg + u * TPR) * ldy + r0 + r] = acc[u][0];
  }
}

template <typename T>
static void spmm_t(const Mat &A, const T *X, int64_t ldx, T *Y, int64_t ldy, int nc, hipStream_t s) {
  if (A.spmv_rb > 0) {   // operators larger than the caches: the stream layout
    const int grid = (int)(((int64_t)A.n + A.spmv_rb - 1) / A.spmv_rb);
    const T *val = static_cast<const T *>(A.val);
    switch (A.spmv_rb) {
      case 256: hipLaunchKernelGGL((spmm_stream_kernel<T, 256>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
      case 128: hipLaunchKernelGGL((spmm_stream_kernel<T, 128>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
      case 64: hipLaunchKernelGGL((spmm_stream_kernel<T, 64>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
      default: hipLaunchKernelGGL((spmm_stream_kernel<T, 32>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
    }
    return;
  }
  const int grid = (int)(((int64_t)A.n * A.lanes + BLK - 1) / BLK);
  const T *val = static_cast<const T *>(A.val);
  switch (A.lanes) {
    case 4: hipLaunchKernelGGL((spmm_kernel<T, 4>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
    case 8: hipLaunchKernelGGL((spmm_kernel<T, 8>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
    case 16: hipLaunchKernelGGL((spmm_kernel<T, 16>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
    default: hipLaunchKernelGGL((spmm_kernel<T, 32>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, Y, ldy, nc); break;
  }
}

int launch_spmm(const Mat &A, const BV &X, int lx, const BV &Y, int ly, int ncols, hipStream_t s) {
  SA_CHECK(A.prec == X.prec && X.prec == Y.prec && A.n == X.n && A.n == Y.n,
           "BVMatMult: operator and bases differ in size or precision");
  if (A.n == 0) return 0;
  for (int j0 = 0; j0 < ncols; j0 += SPMM_NC) {
    const int nc = std::min(SPMM_NC, ncols - j0);
    if (X.prec == SA_FP64) {
      const double *xp = static_cast<const double *>(X.data) + (size_t)(lx + j0) * X.ld;
      double *yp = static_cast<double *>(Y.data) + (size_t)(ly + j0) * Y.ld;
      spmm_t<double>(A, xp, X.ld, yp, Y.ld, nc, s);
    } else {
      const float *xp = static_cast<const float *>(X.data) + (size_t)(lx + j0) * X.ld;
      float *yp = static_cast<float *>(Y.data) + (size_t)(ly + j0) * Y.ld;
      spmm_t<float>(A, xp, X.ld, yp, Y.ld, nc, s);
    }
  }
  SA_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// BVMult (bvops.c:46-108 -> bvmat.c:17-36): Y(:, 0:N) = beta Y + alpha X(:, 0:K) Q
// with Q (K x N, FP64, device) or, without Q, Y = beta Y + alpha X (N == K).
// A workgroup owns 256 rows and 16 output columns (blockIdx.y): the Q slice is
// staged in LDS, each thread accumulates its row's 16 outputs in FP64.  beta
// == 0 ignores Y's old contents, as gemm does.
// ---------------------------------------------------------------------------
constexpr int BVM_NJ = 16;

template <typename T>
__global__ __launch_bounds__(BLK) void bvmult_kernel(int64_t n, const T *__restrict__ X, int64_t ldx, int K,
                                                     T *__restrict__ Y, int64_t ldy, int N,
                                                     const double *__restrict__ Q, int ldq, double alpha,
                                                     double beta) {
  __shared__ double qs[GS_MAXC * BVM_NJ];
  const int j0 = blockIdx.y * BVM_NJ, nj = min(BVM_NJ, N - j0);
  if (Q)
    for (int idx = threadIdx.x; idx < K * BVM_NJ; idx += BLK) {
      const int i = idx / BVM_NJ, jj = idx % BVM_NJ;
      qs[idx] = jj < nj ? Q[i + (size_t)(j0 + jj) * ldq] : 0.0;
    }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (r >= n) return;
  double acc[BVM_NJ];
#pragma unroll
  for (int jj = 0; jj < BVM_NJ; jj++) acc[jj] = 0.0;
  if (Q) {
    for (int i = 0; i < K; i++) {
      const double x = (double)X[(size_t)i * ldx + r];
#pragma unroll
      for (int jj = 0; jj < BVM_NJ; jj++) acc[jj] += x * qs[i * BVM_NJ + jj];
    }
  } else {
#pragma unroll
    for (int jj = 0; jj < BVM_NJ; jj++)
      if (jj < nj) acc[jj] = (double)X[(size_t)(j0 + jj) * ldx + r];
  }
#pragma unroll
  for (int jj = 0; jj < BVM_NJ; jj++) {
    if (jj < nj) {
      T *y = Y + (size_t)(j0 + jj) * ldy + r;
      *y = (T)(beta == 0.0 ? alpha * acc[jj] : beta * (double)*y + alpha * acc[jj]);
    }
  }
}

int launch_bvmult(const BV &X, int lx, int kx, BV &Y, int ly, int ky, const double *Q_dev, int ldq, double alpha,
                  double beta, hipStream_t s) {
  const int K = kx - lx, N = ky - ly;
  SA_CHECK(K >= 0 && N >= 0 && K <= GS_MAXC, "BVMult: invalid active columns");
  SA_CHECK(Q_dev || K == N, "BVMult without Q needs equal numbers of active columns");
  SA_CHECK(X.n == Y.n && X.prec == Y.prec, "BVMult: X and Y differ in length or precision");
  if (N == 0 || X.n == 0) return 0;
  const dim3 grid((unsigned)((X.n + BLK - 1) / BLK), (unsigned)((N + BVM_NJ - 1) / BVM_NJ));
  if (X.prec == SA_FP64)
    hipLaunchKernelGGL(bvmult_kernel<double>, grid, dim3(BLK), 0, s, X.n, (const double *)col_ptr(X, lx), X.ld, K,
                       (double *)col_ptr(Y, ly), Y.ld, N, Q_dev, ldq, alpha, beta);
  else
    hipLaunchKernelGGL(bvmult_kernel<float>, grid, dim3(BLK), 0, s, X.n, (const float *)col_ptr(X, lx), X.ld, K,
                       (float *)col_ptr(Y, ly), Y.ld, N, Q_dev, ldq, alpha, beta);
  SA_HIP(hipGetLastError());
  return 0;
}

}  // namespace sa
