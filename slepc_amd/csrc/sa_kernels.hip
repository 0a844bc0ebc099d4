// HIP kernels for gfx950 (MI355X): CSR SpMV, the BV block kernels (dot
// products, linear combinations, norms, scaling, conversion) and the restart
// GEMM on MFMA.  Wave size 64; 256-thread workgroups.  The fused Gram-Schmidt
// step of the Arnoldi sweep lives in sa_gs.hip.
#include <algorithm>
#include <vector>
#include <cmath>

#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

// ---------------------------------------------------------------------------
// Cross-workgroup reductions use no fences and no atomics (an agent-scope
// release on MI355X writes back the XCD's L2, and atomics from 8 XCDs
// serialise): every workgroup stores its partial sums column-major,
// part[col * G + b]; the end of the kernel publishes them and a one-workgroup
// kernel adds them in a fixed order (deterministic results).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(RBLK) void colsum_kernel(const double *part, int G, int ncol, double *out) {
  sum_partials_block(part, G, ncol, out);
}

// ---------------------------------------------------------------------------
// Vector-CSR SpMV: L lanes per row, shuffle reduction (matrices whose row
// blocks are too dense for the stream kernel below).
// ---------------------------------------------------------------------------
template <typename T, int L>
__global__ __launch_bounds__(BLK) void spmv_kernel(int n, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ ci,
                                                   const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y) {
  const int lg = threadIdx.x % L;
  const int64_t row = ((int64_t)blockIdx.x * BLK + threadIdx.x) / L;
  T s = (T)0;
  if (row < n) {
    const int32_t pe = rp[row + 1];
    for (int32_t p = rp[row] + lg; p < pe; p += L) s += val[p] * x[ci[p]];
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
  if (lg == 0 && row < n) y[row] = s;
}

// Stream CSR SpMV for operators larger than the caches: RB rows per
// workgroup.  The workgroup's nonzeros (a contiguous range of val/col) are
// loaded with every lane active (SPMV_PT value/index loads per thread) and
// staged in LDS.  Thread t then forms the strided partial sum q = t / RB of
// row t % RB (TPR = BLK / RB partials per row: nonzeros q, q + TPR, ... of
// the row, in order), so a wave covers 64 consecutive rows and its gathers
// of x (for a stencil, the same neighbour of consecutive rows) coalesce.  The
// partials are added in a fixed tree order through LDS.  Values and indices
// are read once, so they are loaded non-temporally and do not evict x from
// the caches.  A block whose nonzeros do not fit the LDS buffer is computed
// straight from memory instead, so the kernel is correct for any CSR; the
// host picks RB so that every block fits.

// Partial sum over nonzeros b, b + TPR, ... < e of (vs, cs): four at a time,
// their indices first (clamped into the row, so every load is in bounds),
// then their gathers of x, then the additions in order.  Products are rounded
// before they are added (no fused multiply-add), as the stream SpMM forms them.
template <typename T, int TPR, typename VP, typename CP>
__device__ __forceinline__ T spmv_row_part(VP vs, CP cs, int b, int e, const T *__restrict__ x) {
#pragma clang fp contract(off)
  constexpr int U = 4;
  T s = (T)0;
  for (int p = b; p < e; p += U * TPR) {
    T v[U], xv[U];
    int32_t c[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int pk = min(p + u * TPR, e - 1);
      v[u] = vs[pk];
      c[u] = cs[pk];
    }
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < U; u++)
      if (p + u * TPR < e) s += v[u] * xv[u];
  }
  return s;
}

template <typename T, int RB>
__global__ __launch_bounds__(BLK) void spmv_stream_kernel(int n, const int32_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci,
                                                          const T *__restrict__ val, const T *__restrict__ x,
                                                          T *__restrict__ y) {
  constexpr int TPR = BLK / RB;
  static_assert(RB <= BLK && BLK % RB == 0, "rows per block must divide the workgroup");
  constexpr int NZB = BLK * SPMV_PT;
  static_assert(NZB == SPMV_NZB, "block capacity differs from the host's");
  __shared__ T sval[NZB];
  __shared__ int32_t scol[NZB];
  __shared__ int32_t srp[RB + 1];
  __shared__ T part[TPR > 1 ? BLK : 1];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  const int rows = (int)min((int64_t)RB, (int64_t)n - r0);
  for (int i = t; i <= rows; i += BLK) srp[i] = rp[r0 + i];
  __syncthreads();
  const int32_t p0 = srp[0], nz = srp[rows] - p0;
  const bool staged = nz <= NZB;         // uniform across the workgroup
  if (staged) {
    T v[SPMV_PT];
    int32_t c[SPMV_PT];
#pragma unroll
    for (int k = 0; k < SPMV_PT; k++) {
      const int i = k * BLK + t;
      if (i < nz) {
        v[k] = __builtin_nontemporal_load(val + p0 + i);
        c[k] = __builtin_nontemporal_load(ci + p0 + i);
      }
    }
#pragma unroll
    for (int k = 0; k < SPMV_PT; k++) {
      const int i = k * BLK + t;
      if (i < nz) {
        sval[i] = v[k];
        scol[i] = c[k];
      }
    }
  }
  __syncthreads();
  const int r = t % RB, q = t / RB;
  T s = (T)0;
  if (r < rows) {
    const int b = srp[r] - p0 + q, e = srp[r + 1] - p0;
    s = staged ? spmv_row_part<T, TPR>(sval, scol, b, e, x) : spmv_row_part<T, TPR>(val + p0, ci + p0, b, e, x);
  }
  if constexpr (TPR == 1) {
    if (r < rows) y[r0 + r] = s;
  } else {
    part[t] = s;                         // part[q * RB + r]
    __syncthreads();
    if (q == 0 && r < rows) {
      T a[TPR];
#pragma unroll
      for (int u = 0; u < TPR; u++) a[u] = part[u * RB + r];
#pragma unroll
      for (int h = TPR / 2; h > 0; h >>= 1)
#pragma unroll
        for (int u = 0; u < h; u++) a[u] = a[u] + a[u + h];
      y[r0 + r] = a[0];
    }
  }
}

template <typename T>
static int spmv_t(const Mat &A, const T *x, T *y, hipStream_t s) {
  if (A.spmv_rb > 0) {
    const int grid = (int)(((int64_t)A.n + A.spmv_rb - 1) / A.spmv_rb);
    const T *val = (const T *)A.val;
    switch (A.spmv_rb) {
      case 256: hipLaunchKernelGGL((spmv_stream_kernel<T, 256>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, x, y); break;
      case 128: hipLaunchKernelGGL((spmv_stream_kernel<T, 128>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, x, y); break;
      case 64: hipLaunchKernelGGL((spmv_stream_kernel<T, 64>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, x, y); break;
      default: hipLaunchKernelGGL((spmv_stream_kernel<T, 32>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, x, y); break;
    }
    SA_HIP(hipGetLastError());
    return 0;
  }
  const int64_t threads = (int64_t)A.n * A.lanes;
  const int grid = (int)((threads + BLK - 1) / BLK);
  switch (A.lanes) {
    case 4: hipLaunchKernelGGL((spmv_kernel<T, 4>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, (const T *)A.val, x, y); break;
    case 8: hipLaunchKernelGGL((spmv_kernel<T, 8>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, (const T *)A.val, x, y); break;
    case 16: hipLaunchKernelGGL((spmv_kernel<T, 16>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, (const T *)A.val, x, y); break;
    default: hipLaunchKernelGGL((spmv_kernel<T, 32>), dim3(grid), dim3(BLK), 0, s, A.n, A.rowptr, A.col, (const T *)A.val, x, y); break;
  }
  SA_HIP(hipGetLastError());
  return 0;
}

// MatNorm(A, NORM_INFINITY): part[b] = the largest absolute row sum of the
// rows that workgroup b strides over; the host takes the largest partial.
template <typename T>
__global__ __launch_bounds__(BLK) void row_abs_max_kernel(int n, const int32_t *__restrict__ rp,
                                                          const T *__restrict__ val, double *part) {
  __shared__ double sh[BLK / 64];
  double m = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double r = 0.0;
    for (int32_t q = rp[i]; q < rp[i + 1]; q++) r += fabs((double)val[q]);
    m = fmax(m, r);
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLK / 64; w++) m = fmax(m, sh[w]);
    part[blockIdx.x] = fmax(m, sh[0]);
  }
}

int mat_norm_inf(const Mat &A, double *out) {
  *out = 0.0;
  if (A.n == 0) return 0;
  const int G = (int)std::min<int64_t>(1024, ((int64_t)A.n + BLK - 1) / BLK);
  hipStream_t s = stream();
  double *part = nullptr;
  SA_HIP(hipMalloc(&part, sizeof(double) * G));
  if (A.prec == SA_FP64)
    hipLaunchKernelGGL(row_abs_max_kernel<double>, dim3(G), dim3(BLK), 0, s, A.n, A.rowptr, (const double *)A.val, part);
  else
    hipLaunchKernelGGL(row_abs_max_kernel<float>, dim3(G), dim3(BLK), 0, s, A.n, A.rowptr, (const float *)A.val, part);
  std::vector<double> h(G);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * G, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(part);
  SA_HIP(e);
  for (double v : h) *out = std::max(*out, v);
  return 0;
}

int launch_spmv(const Mat &A, const void *x, void *y, hipStream_t s) {
  if (A.n == 0) return 0;
  if (A.prec == SA_FP64) return spmv_t<double>(A, (const double *)x, (double *)y, s);
  return spmv_t<float>(A, (const float *)x, (float *)y, s);
}

// ---------------------------------------------------------------------------
// General BVDotVec: q = V(:,l:k)^T x  (partials per workgroup, then reduce).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BLK) void dotvec_kernel(int64_t n, int64_t ld, const T *__restrict__ V, int ncol,
                                                     const T *__restrict__ x, int rpb, double *part, int G) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *Vs = reinterpret_cast<T *>(smem_raw);
  T *xs = Vs + (size_t)ncol * LDP;
  __shared__ double s_acc[BLK];
  const int t = threadIdx.x;
  const int nsl = BLK / ncol, dcol = t % ncol, dsl = t / ncol;
  double acc = 0.0;
  const int64_t rb = (int64_t)blockIdx.x * rpb, re = min(n, rb + rpb);
  for (int64_t r0 = rb; r0 < re; r0 += CH) {
    const int rows = (int)min((int64_t)CH, re - r0);
    const int row = t & (CH - 1), cg = t >> 6;
    if (t < CH) xs[t] = (t < rows) ? x[r0 + t] : (T)0;
    for (int i = cg; i < ncol; i += 4) Vs[i * LDP + row] = (row < rows) ? V[(size_t)i * ld + r0 + row] : (T)0;
    __syncthreads();
    if (dsl < nsl) {
      T s = (T)0;
      for (int r = dsl; r < rows; r += nsl) s += Vs[dcol * LDP + r] * xs[r];
      acc += (double)s;
    }
    __syncthreads();
  }
  s_acc[t] = acc;
  __syncthreads();
  if (t < ncol) {
    double s = 0.0;
    for (int q = 0; q < nsl; q++) s += s_acc[q * ncol + t];
    part[(size_t)t * G + blockIdx.x] = s;
  }
}

int launch_dotvec(BV &V, int l, int k, const void *x, double *q_dev, hipStream_t s) {
  const int ncol = k - l;
  if (ncol <= 0) return 0;
  SA_CHECK(ncol <= GS_MAXC, "too many active columns for BVDotVec");
  const size_t es = prec_size(V.prec);
  const size_t shm = ((size_t)ncol * LDP + CH) * es;
  if (V.prec == SA_FP64)
    hipLaunchKernelGGL(dotvec_kernel<double>, dim3(V.G), dim3(BLK), shm, s, V.n, V.ld,
                       (const double *)col_ptr(V, l), ncol, (const double *)x, V.rows_per_block, V.ws.part, V.G);
  else
    hipLaunchKernelGGL(dotvec_kernel<float>, dim3(V.G), dim3(BLK), shm, s, V.n, V.ld,
                       (const float *)col_ptr(V, l), ncol, (const float *)x, V.rows_per_block, V.ws.part, V.G);
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(RBLK), 0, s, (const double *)V.ws.part, V.G, ncol, q_dev);
  SA_HIP(hipGetLastError());
  return 0;
}

// y = beta*y + alpha*V(:,0:ncol) q
template <typename T>
__global__ __launch_bounds__(BLK) void multvec_kernel(int64_t n, int64_t ld, const T *__restrict__ V, int ncol,
                                                      const double *__restrict__ q, double alpha,
                                                      double beta, T *__restrict__ y) {
  __shared__ T qs[GS_MAXC];
  for (int i = threadIdx.x; i < ncol; i += BLK) qs[i] = (T)q[i];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (r >= n) return;
  T s = (T)0;
  for (int i = 0; i < ncol; i++) s += V[(size_t)i * ld + r] * qs[i];
  y[r] = (beta == 0.0 ? (T)0 : (T)beta * y[r]) + (T)alpha * s;   // beta 0: y is not read (gemv)
}

int launch_multvec(BV &V, int l, int k, double alpha, const double *q_dev, double beta, void *y,
                   hipStream_t s) {
  const int ncol = std::max(0, k - l);
  SA_CHECK(ncol <= GS_MAXC, "too many active columns for BVMultVec");
  const int grid = (int)((V.n + BLK - 1) / BLK);
  if (V.n == 0) return 0;
  if (V.prec == SA_FP64)
    hipLaunchKernelGGL(multvec_kernel<double>, dim3(grid), dim3(BLK), 0, s, V.n, V.ld,
                       (const double *)col_ptr(V, l), ncol, q_dev, alpha, beta, (double *)y);
  else
    hipLaunchKernelGGL(multvec_kernel<float>, dim3(grid), dim3(BLK), 0, s, V.n, V.ld,
                       (const float *)col_ptr(V, l), ncol, q_dev, alpha, beta, (float *)y);
  SA_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Restart GEMM on MFMA: V(:,s:e) = V(:,l:k) * Q(l:k,s:e), in place.
// multinplace_kernel (for Q slices too large for LDS): a workgroup owns 64
// rows, stages V(rows, l:k) in LDS, then each wave computes a 16-row slab of
// the output with 16x16x4 MFMA tiles, reading Q from global memory.  The
// Q-in-LDS kernel below is the common path (2.6x faster at 2.1M rows).
// ---------------------------------------------------------------------------
constexpr int MIP_NT = 8;   // up to 8 output tiles of 16 columns (128 columns)

template <typename T>
__global__ __launch_bounds__(BLK) void multinplace_kernel(int64_t n, int64_t ld, T *V, int l, int K,
                                                          int s, int N, const double *__restrict__ Q,
                                                          int ldq) {
  typedef typename Vec4<T>::type V4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *As = reinterpret_cast<T *>(smem_raw);       // [K][LDP]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * CH;
  const int rows = (int)min((int64_t)CH, n - r0);
  {
    const int row = t & (CH - 1), cg = t >> 6;
    for (int i = cg; i < K; i += 4)
      As[i * LDP + row] = (row < rows) ? V[(size_t)(l + i) * ld + r0 + row] : (T)0;
  }
  __syncthreads();
  const int ntiles = (N + 15) / 16;
  V4 acc[MIP_NT];
#pragma unroll
  for (int q = 0; q < MIP_NT; q++) acc[q] = (V4){0, 0, 0, 0};
  const int ai = w * 16 + (lane & 15);     // A row within the block
  const int kq = lane >> 4;                // k index within the 4-wide step
  const int bj = lane & 15;                // B column within the tile
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int kk = k0 + kq;
    const T av = (kk < K) ? As[kk * LDP + ai] : (T)0;
#pragma unroll
    for (int q = 0; q < MIP_NT; q++) {
      if (q < ntiles) {
        const int j = q * 16 + bj;
        const T bv = (kk < K && j < N) ? (T)Q[(size_t)(l + kk) + (size_t)(s + j) * ldq] : (T)0;
        acc[q] = mfma16x16x4(av, bv, acc[q]);
      }
    }
  }
  // Output layout: column j = lane % 16.  f32 16x16x4: element v of lane
  // group g = lane/16 is row 4g + v; f64 16x16x4: row g + 4v.
  constexpr bool kF64 = sizeof(T) == 8;
#pragma unroll
  for (int q = 0; q < MIP_NT; q++) {
    if (q < ntiles) {
      const int j = q * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = w * 16 + (kF64 ? ((lane >> 4) + 4 * v) : (4 * (lane >> 4) + v));
        if (j < N && i < rows) V[(size_t)(s + j) * ld + r0 + i] = acc[q][v];
      }
    }
  }
}

// Restart GEMM with Q staged in LDS (used when Q(l:k,s:e) fits in MIP_QS_BYTES).
// The workgroup converts Q(l:k, s:e) to T in LDS once, then strides over
// 64-row blocks; wave w computes rows 16w..16w+15 of a block for all N
// columns.  The MFMA computes the transposed tile, C^T(j, i) = sum_k Q(k, j)
// V(i, k): operand A is Q from LDS, operand B the basis values, loaded
// straight from HBM (lanes 0-15 read 16 consecutive rows of one column) in
// batches of MIP_KB k-steps with the next batch in flight.  Output element v
// of a lane is column j = 4g+v (f32) or g+4v (f64) of the tile, g = lane/16,
// and row i = lane%16, so a store writes 128-byte runs of a column.  Every
// wave reads all K basis values of its rows before it writes them, and no
// other wave touches those rows, so the product is safe in place.  A basis
// larger than the last-level cache (NT) is read with non-temporal loads, each
// value being read once (1.2x faster at K = 32, n = 2.1M; at K = 96 the FP64
// MFMA work, 2 K N flops per row, is as long as the memory time).
constexpr int MIP_KB = 8;
constexpr size_t MIP_QS_BYTES = 64 * 1024;
constexpr int64_t MIP_GRID = 1024;             // workgroups (each stages Q once)

template <typename T, int NTL, bool NT>
__global__ __launch_bounds__(BLK) void multinplace_q_kernel(int64_t n, int64_t ld, T *V, int l, int K, int s,
                                                            int N, const double *__restrict__ Q, int ldq) {
  typedef typename Vec4<T>::type V4;
  constexpr int NP = NTL * 16 + 1;               // LDS row stride (odd: rows start in different banks)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *Qs = reinterpret_cast<T *>(smem_raw);        // [K4][NP]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int K4 = (K + 3) & ~3;
  for (int idx = t; idx < K4 * NTL * 16; idx += BLK) {
    const int kk = idx / (NTL * 16), j = idx % (NTL * 16);
    Qs[kk * NP + j] = (kk < K && j < N) ? (T)Q[(size_t)(l + kk) + (size_t)(s + j) * ldq] : (T)0;
  }
  __syncthreads();
  const int g = lane >> 4, c16 = lane & 15;
  const int nsteps = K4 / 4, nbatch = (nsteps + MIP_KB - 1) / MIP_KB;
  const int64_t nblk = (n + CH - 1) / CH;
  constexpr bool kF64 = sizeof(T) == 8;
  for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int64_t row = b * CH + w * 16 + c16;
    const bool rin = row < n;
    const T *src = V + (size_t)l * ld + row;
    V4 acc[NTL];
#pragma unroll
    for (int q = 0; q < NTL; q++) acc[q] = (V4){0, 0, 0, 0};
    T av[MIP_KB], an[MIP_KB];
#pragma unroll
    for (int u = 0; u < MIP_KB; u++) {
      const int kk = 4 * u + g;
      av[u] = (rin && kk < K) ? ld_stream<NT>(src + (size_t)kk * ld) : (T)0;
    }
    for (int kb = 0; kb < nbatch; kb++) {
      if (kb + 1 < nbatch) {
#pragma unroll
        for (int u = 0; u < MIP_KB; u++) {
          const int kk = 4 * ((kb + 1) * MIP_KB + u) + g;
          an[u] = (rin && kk < K) ? ld_stream<NT>(src + (size_t)kk * ld) : (T)0;
        }
      }
#pragma unroll
      for (int u = 0; u < MIP_KB; u++) {
        const int step = kb * MIP_KB + u;
        if (step < nsteps) {
          const T *qrow = Qs + (4 * step + g) * NP + c16;
#pragma unroll
          for (int q = 0; q < NTL; q++) acc[q] = mfma16x16x4(qrow[q * 16], av[u], acc[q]);
        }
      }
#pragma unroll
      for (int u = 0; u < MIP_KB; u++) av[u] = an[u];
    }
    if (rin) {
      T *dst = V + (size_t)s * ld + row;
#pragma unroll
      for (int q = 0; q < NTL; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int j = q * 16 + (kF64 ? (g + 4 * v) : (4 * g + v));
          if (j < N) dst[(size_t)j * ld] = acc[q][v];
        }
    }
  }
}

template <typename T, int NTL>
static void launch_mip_q(BV &V, int l, int K, int s_, int N, const double *Q_dev, int ldq, hipStream_t s) {
  const int64_t nblk = (V.n + CH - 1) / CH;
  const int grid = (int)std::min<int64_t>(nblk, MIP_GRID);
  const size_t shm = (size_t)((K + 3) & ~3) * (NTL * 16 + 1) * sizeof(T);
  if (V.n * (int64_t)K * (int64_t)sizeof(T) > GS_NT_BYTES)
    hipLaunchKernelGGL((multinplace_q_kernel<T, NTL, true>), dim3(grid), dim3(BLK), shm, s, V.n, V.ld, (T *)V.data,
                       l, K, s_, N, Q_dev, ldq);
  else
    hipLaunchKernelGGL((multinplace_q_kernel<T, NTL, false>), dim3(grid), dim3(BLK), shm, s, V.n, V.ld, (T *)V.data,
                       l, K, s_, N, Q_dev, ldq);
}

template <typename T>
static bool multinplace_q(BV &V, int l, int K, int s_, int N, const double *Q_dev, int ldq, hipStream_t s) {
  const int ntl = (N + 15) / 16;
  if ((size_t)((K + 3) & ~3) * (ntl * 16 + 1) * sizeof(T) > MIP_QS_BYTES) return false;
  switch (ntl) {
    case 1: launch_mip_q<T, 1>(V, l, K, s_, N, Q_dev, ldq, s); break;
    case 2: launch_mip_q<T, 2>(V, l, K, s_, N, Q_dev, ldq, s); break;
    case 3: launch_mip_q<T, 3>(V, l, K, s_, N, Q_dev, ldq, s); break;
    case 4: launch_mip_q<T, 4>(V, l, K, s_, N, Q_dev, ldq, s); break;
    case 5: launch_mip_q<T, 5>(V, l, K, s_, N, Q_dev, ldq, s); break;
    case 6: launch_mip_q<T, 6>(V, l, K, s_, N, Q_dev, ldq, s); break;
    case 7: launch_mip_q<T, 7>(V, l, K, s_, N, Q_dev, ldq, s); break;
    default: launch_mip_q<T, 8>(V, l, K, s_, N, Q_dev, ldq, s); break;
  }
  return true;
}

int launch_multinplace(BV &V, int l, int k, int s_, int e, const double *Q_dev, int ldq, hipStream_t s) {
  const int K = k - l, N = e - s_;
  if (K <= 0 || N <= 0 || V.n == 0) return 0;
  SA_CHECK(N <= MIP_NT * 16 && K <= GS_MAXC, "BVMultInPlace block too large");
  const bool done = V.prec == SA_FP64 ? multinplace_q<double>(V, l, K, s_, N, Q_dev, ldq, s)
                                      : multinplace_q<float>(V, l, K, s_, N, Q_dev, ldq, s);
  if (done) {
    SA_HIP(hipGetLastError());
    return 0;
  }
  const int grid = (int)((V.n + CH - 1) / CH);
  const size_t shm = (size_t)K * LDP * prec_size(V.prec);
  SA_CHECK(shm <= 65536, "BVMultInPlace block too large for the LDS tile");
  if (V.prec == SA_FP64)
    hipLaunchKernelGGL(multinplace_kernel<double>, dim3(grid), dim3(BLK), shm, s, V.n, V.ld, (double *)V.data,
                       l, K, s_, N, Q_dev, ldq);
  else
    hipLaunchKernelGGL(multinplace_kernel<float>, dim3(grid), dim3(BLK), shm, s, V.n, V.ld, (float *)V.data,
                       l, K, s_, N, Q_dev, ldq);
  SA_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Small elementwise / reduction kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void scale_kernel(T *x, int64_t n, T alpha) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i < n) x[i] *= alpha;
}

int launch_scale(int prec, void *x, int64_t n, double alpha, hipStream_t s) {
  if (n == 0) return 0;
  const int grid = (int)((n + BLK - 1) / BLK);
  if (prec == SA_FP64) hipLaunchKernelGGL(scale_kernel<double>, dim3(grid), dim3(BLK), 0, s, (double *)x, n, alpha);
  else hipLaunchKernelGGL(scale_kernel<float>, dim3(grid), dim3(BLK), 0, s, (float *)x, n, (float)alpha);
  SA_HIP(hipGetLastError());
  return 0;
}

// Sum of squares: block partials then a single-block finish (FP64 totals).
template <typename T>
__global__ __launch_bounds__(BLK) void sumsq_partial_kernel(const T *x, const T *y, double lambda, int64_t n,
                                                            double *part) {
  __shared__ double red[BLK];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    double v = y ? (double)x[i] - lambda * (double)y[i] : (double)x[i];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void sum_finish_kernel(const double *part, int np, double *out) {
  __shared__ double red[BLK];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += BLK) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

static int sumsq(int prec, const void *x, const void *y, double lambda, int64_t n, double *part, double *out,
                 hipStream_t s) {
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + BLK - 1) / BLK));
  if (prec == SA_FP64)
    hipLaunchKernelGGL(sumsq_partial_kernel<double>, dim3(grid), dim3(BLK), 0, s, (const double *)x,
                       (const double *)y, lambda, n, part);
  else
    hipLaunchKernelGGL(sumsq_partial_kernel<float>, dim3(grid), dim3(BLK), 0, s, (const float *)x,
                       (const float *)y, lambda, n, part);
  hipLaunchKernelGGL(sum_finish_kernel, dim3(1), dim3(BLK), 0, s, part, grid, out);
  SA_HIP(hipGetLastError());
  return 0;
}

int launch_dot_self(int prec, const void *x, int64_t n, double *out_dev, double *part, hipStream_t s) {
  return sumsq(prec, x, nullptr, 0.0, n, part, out_dev, s);
}

int launch_axpy_norm(int prec, const void *Ax, const void *x, double lambda, int64_t n, double *part,
                     double *out_dev, hipStream_t s) {
  return sumsq(prec, Ax, x, lambda, n, part, out_dev, s);
}

template <typename S, typename D>
__global__ void convert_kernel(const S *src, int64_t sld, D *dst, int64_t dld, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) dst[(size_t)j * dld + i] = (D)src[(size_t)j * sld + i];
}

int launch_convert(int sp, const void *src, int64_t sld, int dp, void *dst, int64_t dld, int64_t n, int ncols,
                   hipStream_t s) {
  if (n == 0 || ncols == 0) return 0;
  dim3 grid((unsigned)((n + BLK - 1) / BLK), (unsigned)ncols);
  if (sp == SA_FP64 && dp == SA_FP32)
    hipLaunchKernelGGL((convert_kernel<double, float>), grid, dim3(BLK), 0, s, (const double *)src, sld, (float *)dst, dld, n);
  else if (sp == SA_FP32 && dp == SA_FP64)
    hipLaunchKernelGGL((convert_kernel<float, double>), grid, dim3(BLK), 0, s, (const float *)src, sld, (double *)dst, dld, n);
  else if (sp == SA_FP64)
    hipLaunchKernelGGL((convert_kernel<double, double>), grid, dim3(BLK), 0, s, (const double *)src, sld, (double *)dst, dld, n);
  else
    hipLaunchKernelGGL((convert_kernel<float, float>), grid, dim3(BLK), 0, s, (const float *)src, sld, (float *)dst, dld, n);
  SA_HIP(hipGetLastError());
  return 0;
}

}  // namespace sa
