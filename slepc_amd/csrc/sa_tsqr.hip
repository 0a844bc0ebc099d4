// Tall-skinny QR (BVOrthogonalize with block method TSQR / TSQRCHOL) for
// gfx950 (MI355X).  Reference: bvorthog.c:615-659 and bvlapack.c:347-480,
// 510-560 (geqrf + orgqr on each process, a binary tree of QR factorizations
// of stacked R factors across processes).
//
// On one GPU the tree runs over tiles of rows instead of processes (256-row
// tiles factored by a workgroup, one row per thread, for 33..64 columns;
// tiles of 128 rows (17..32 columns) or 256 rows (up to 16) factored by a
// single wave, 2 or 4 rows per lane, below):
//   level 0: every tile's Householder QR of its rows
//            of V(:, l:k) (held in registers) is computed; the kernel writes the
//            Householder vectors over the tile and its k x k R into the
//            stacked matrix of level 1;
//   level i: the same kernel on the stacked R factors, until one tile is left,
//            whose R is the R of the whole block.
// Q is then formed top-down: each tile applies its reflectors to the k x k
// block of the level above that belongs to it (padded with zero rows), in
// place; the top tile forms its explicit factor (the orgqr of the tile).
//
// The R of a tree differs from the R of one flat Householder QR (what geqrf
// computes on one process) only in the signs of its rows.  Those signs follow
// from the first k rows of Q: Householder QR of A and of Q D (D the signs of
// diag(R), so that R D has a positive diagonal) use the same reflectors, and
// the sign LAPACK picks for column j is minus the sign of the pivot of an LU
// factorization of those k rows that subtracts each chosen sign from its
// pivot (Ballard, Demmel, Grigori, Jacquelin, Nguyen, Solomonik, "Reconstructing
// Householder vectors from tall-skinny QR", IPDPS 2014).  The host computes
// these k signs from a one-tile run of the level-0 kernel and the full level-0
// pass applies them as it stores Q, so Q and R equal geqrf/orgqr's up to
// rounding at no extra pass over V.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

constexpr int TS_ROWS = BLK;   // rows per tile, one per thread
constexpr int TS_MAXK = 64;    // most columns per call

template <typename T, int K>
struct TsShared {
  T sig[2][BLK / 64];     // per-wave partial column norms (two buffers: steps alternate)
  T alpha[2];             // diagonal entry of the current column
  T red[BLK / 64][K];     // per-wave partial products v^T X
  T w[K];                 // their sums
  T tau[K];               // Householder scalars of the tile
};

template <typename T>
__device__ __forceinline__ T ts_wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Reduce-scatter over one wave of the products v * x[OFF + c], c < KS = K - OFF
// (recursive halving: at mask m a lane keeps the half its bit m selects and
// adds its partner's copy of that half).  Lane l returns the wave sum for
// column OFF + (l & (KS - 1)); the sums are bitwise equal on every lane that
// holds the same column.
// The two halves pass through an empty asm statement before the lane-dependent
// choice between them: without it the compiler rewrites the choice as a
// dynamically indexed read of p and emits a compare/select chain over the
// whole array for every element.
template <typename T>
__device__ __forceinline__ T ts_keep_add(T a, T b, bool hi, int m) {
  asm volatile("" : "+v"(a), "+v"(b));
  return (hi ? b : a) + __shfl_xor(hi ? a : b, m, 64);
}

template <typename T, int K, int OFF>
__device__ __forceinline__ T ts_dots(const T (&x)[K], T v, int lane) {
  constexpr int KS = K - OFF, H = KS / 2;
  static_assert(KS >= 2 && (KS & (KS - 1)) == 0 && KS <= 64, "window must be a power of two up to 64");
  T p[H];
  {
    const bool hi = (lane & H) != 0;
#pragma unroll
    for (int i = 0; i < H; i++) p[i] = ts_keep_add(v * x[OFF + i], v * x[OFF + H + i], hi, H);
  }
#pragma unroll
  for (int m = H / 2; m >= 1; m >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < m; i++) p[i] = ts_keep_add(p[i], p[i + m], hi, m);
  }
#pragma unroll
  for (int m = KS; m < 64; m <<= 1) p[0] += __shfl_xor(p[0], m, 64);
  return p[0];
}

// w[c] = sum over the tile of v * x[c] for the columns of the window
// [OFF, K), left in sh.w (fixed summation order: waves 0, 1, 2, 3).
template <typename T, int K, int OFF>
__device__ __forceinline__ void ts_tile_dots(const T (&x)[K], T v, int t, TsShared<T, K> &sh) {
  constexpr int KS = K - OFF;
  const int lane = t & 63, wv = t >> 6;
  const T ps = ts_dots<T, K, OFF>(x, v, lane);
  if (lane < KS) sh.red[wv][OFF + lane] = ps;
  __syncthreads();
  if (t >= OFF && t < K) {
    T s = sh.red[0][t];
#pragma unroll
    for (int q = 1; q < BLK / 64; q++) s += sh.red[q][t];
    sh.w[t] = s;
  }
  __syncthreads();
}

// Householder steps j0 <= j < j1 (dgeqr2 / dlarfg order) on the tile whose
// row t is x; only columns OFF.. can still change.  Column j becomes the
// Householder vector below the diagonal and beta on it.
template <typename T, int K, int OFF>
__device__ __forceinline__ void ts_factor_phase(T (&x)[K], int j0, int j1, int t, TsShared<T, K> &sh) {
  const int lane = t & 63, wv = t >> 6;
  // one copy of the step body: an unrolled step loop outgrows the instruction cache
#pragma nounroll
  for (int j = j0; j < j1; j++) {
    T xj = x[OFF];
#pragma unroll
    for (int c = OFF + 1; c < K; c++) xj = c == j ? x[c] : xj;
    const T s = ts_wave_sum(t > j ? xj * xj : (T)0);
    const int b = j & 1;
    if (lane == 0) sh.sig[b][wv] = s;
    if (t == j) sh.alpha[b] = xj;
    __syncthreads();
    T sigma = sh.sig[b][0];
#pragma unroll
    for (int q = 1; q < BLK / 64; q++) sigma += sh.sig[b][q];
    const T alpha = sh.alpha[b];
    T tau = 0;
    if (sigma != (T)0) {               // else H = I, as dlarfg does for a zero column
      const T nrm = sqrt(alpha * alpha + sigma);
      const T beta = alpha >= (T)0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      const T scal = (T)1 / (alpha - beta);
      const T v = t > j ? xj * scal : (t == j ? (T)1 : (T)0);
      ts_tile_dots<T, K, OFF>(x, v, t, sh);
      const T tv = tau * v;
#pragma unroll
      for (int c = OFF; c < K; c++)
        if (c > j) x[c] -= tv * sh.w[c];
      const T nj = t > j ? v : (t == j ? beta : xj);
#pragma unroll
      for (int c = OFF; c < K; c++)
        if (c == j) x[c] = nj;
    }
    if (t == 0) sh.tau[j] = tau;
  }
}

// The phases of the factorization: the window [OFF, K) halves while the
// columns to its left are finished (64 -> 32 -> 16 -> 8 wide for K = 64).
template <typename T, int K, int OFF>
__device__ __forceinline__ void ts_factor_all(T (&x)[K], int kk, int t, TsShared<T, K> &sh) {
  constexpr int KS = K - OFF, JE = KS == 8 ? K : OFF + KS / 2;
  ts_factor_phase<T, K, OFF>(x, OFF, min(kk, JE), t, sh);
  if constexpr (KS > 8) ts_factor_all<T, K, JE>(x, kk, t, sh);
}

// One tree level: tile b = blockIdx.x covers rows [256 b, 256 b + nb) of the
// rows x k matrix A (column-major, lda).  Y (may equal A; nullptr: not kept)
// receives the Householder vectors and R in geqrf layout, tau[b k + j] the
// scalars and rows b k .. b k + k - 1 of R (ldr) the tile's k x k R (rows at
// or beyond nb zero).
template <typename T, int K>
__global__ __launch_bounds__(BLK) void tsqr_factor_kernel(const T *A, int64_t lda, int64_t rows, int k, T *Y,
                                                          int64_t ldy, T *__restrict__ tau, T *__restrict__ R,
                                                          int64_t ldr) {
  __shared__ TsShared<T, K> sh;
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * TS_ROWS;
  const int nb = (int)min((int64_t)TS_ROWS, rows - r0);
  const int kk = min(k, nb);
  T x[K];
#pragma unroll
  for (int c = 0; c < K; c++) x[c] = (c < k && t < nb) ? A[r0 + t + (int64_t)c * lda] : (T)0;
  ts_factor_all<T, K, 0>(x, kk, t, sh);
  __syncthreads();
  if (Y && t < nb) {
#pragma unroll
    for (int c = 0; c < K; c++)
      if (c < k) Y[r0 + t + (int64_t)c * ldy] = x[c];
  }
  if (t < k) {
    const int64_t rr = (int64_t)blockIdx.x * k + t;
#pragma unroll
    for (int c = 0; c < K; c++)
      if (c < k) R[rr + (int64_t)c * ldr] = (t < nb && c >= t) ? x[c] : (T)0;
    tau[rr] = t < kk ? sh.tau[t] : (T)0;
  }
}

// Applies H_j for j1 > j >= j0 (descending, dorg2r order) to the matrix
// whose row t is q; columns left of OFF must be ones these reflectors leave
// unchanged.
template <typename T, int K, int OFF>
__device__ __forceinline__ void ts_form_phase(T (&q)[K], int j1, int j0, int t, const T *yrow, int64_t ldy, bool inrow,
                                              const T *tau_b, TsShared<T, K> &sh) {
#pragma nounroll
  for (int j = j1 - 1; j >= j0; j--) {
    const T tj = tau_b[j];
    if (tj == (T)0) continue;
    const T v = t > j ? (inrow ? yrow[(int64_t)j * ldy] : (T)0) : (t == j ? (T)1 : (T)0);
    ts_tile_dots<T, K, OFF>(q, v, t, sh);
    const T tv = tj * v;
#pragma unroll
    for (int c = OFF; c < K; c++) q[c] -= tv * sh.w[c];
  }
}

template <typename T, int K, int OFF>
__device__ __forceinline__ void ts_form_all(T (&q)[K], int kk, int t, const T *yrow, int64_t ldy, bool inrow,
                                            const T *tau_b, TsShared<T, K> &sh) {
  constexpr int KS = K - OFF, JE = KS == 8 ? K : OFF + KS / 2;
  if constexpr (KS > 8) ts_form_all<T, K, JE>(q, kk, t, yrow, ldy, inrow, tau_b, sh);
  ts_form_phase<T, K, OFF>(q, min(kk, JE), OFF, t, yrow, ldy, inrow, tau_b, sh);
}

// Q of one tree level: tile b applies its reflectors (rows x k Y in geqrf
// layout, scalars tau[b k ..]) to [C_b; 0], C_b = rows b k .. b k + k - 1 of
// the Q of the level above (ldc), which gives its rows of Q; with C == nullptr
// (the top tile) to [I; 0], the explicit Householder factor (dorg2r), where
// the columns left of each reflector stay unit vectors and the windows shrink
// as in the factorization.  Row t of the result, times sgn[c] per column when
// sgn is given, goes to out (may equal Y: each thread reads only its own row
// of Y).
template <typename T, int K>
__global__ __launch_bounds__(BLK) void tsqr_form_kernel(const T *Y, int64_t ldy, int64_t rows, int k,
                                                        const T *__restrict__ tau, const T *__restrict__ C, int64_t ldc,
                                                        T *out, int64_t ldo, const T *__restrict__ sgn) {
  __shared__ TsShared<T, K> sh;
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * TS_ROWS;
  const int nb = (int)min((int64_t)TS_ROWS, rows - r0);
  const int kk = min(k, nb);
  const bool inrow = t < nb;
  const T *yrow = Y + r0 + t;
  const T *tau_b = tau + (int64_t)blockIdx.x * k;
  T q[K];
  if (C) {
    const bool crow = t < kk;   // rows of C_b beyond the tile's rows meet zero rows of its R
#pragma unroll
    for (int c = 0; c < K; c++) q[c] = (crow && c < k) ? C[(int64_t)blockIdx.x * k + t + (int64_t)c * ldc] : (T)0;
    ts_form_phase<T, K, 0>(q, kk, 0, t, yrow, ldy, inrow, tau_b, sh);
  } else {
#pragma unroll
    for (int c = 0; c < K; c++) q[c] = t == c ? (T)1 : (T)0;
    ts_form_all<T, K, 0>(q, kk, t, yrow, ldy, inrow, tau_b, sh);
  }
  if (inrow) {
#pragma unroll
    for (int c = 0; c < K; c++)
      if (c < k) out[r0 + t + (int64_t)c * ldo] = sgn ? q[c] * sgn[c] : q[c];
  }
}

// ---- one wave per tile (k <= 32) ----------------------------------------------
// A wave factors a tile of 64 RPL rows on its own: lane L holds rows L, L+64,
// ..., L + 64 (RPL-1) of the tile (each column load is coalesced), so every
// reduction is a wave reduction and no Householder step waits at a workgroup
// barrier.  The 4 waves of a workgroup take 4 consecutive tiles.  The steps
// are those of the workgroup kernels above (dgeqr2 / dorg2r order, the same
// windows); only the order of the additions inside a reduction differs.

// A value that is uniform across the wave, read from lane l (l uniform).
template <typename T>
__device__ __forceinline__ T ts_readlane(T v, int l) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  }
}

// Lane l returns sum over the tile of v * x[c] for column c = OFF + (l & (KS-1)):
// each lane first adds its RPL rows, then the wave reduce-scatters as ts_dots.
template <typename T, int K, int RPL, int OFF>
__device__ __forceinline__ T tw_dots(const T (&x)[RPL][K], const T (&v)[RPL], int lane) {
  constexpr int KS = K - OFF, H = KS / 2;
  static_assert(KS >= 2 && (KS & (KS - 1)) == 0 && KS <= 64, "window must be a power of two up to 64");
  T p[H];
  {
    const bool hi = (lane & H) != 0;
#pragma unroll
    for (int i = 0; i < H; i++) {
      T a = v[0] * x[0][OFF + i], b = v[0] * x[0][OFF + H + i];
#pragma unroll
      for (int r = 1; r < RPL; r++) {
        a += v[r] * x[r][OFF + i];
        b += v[r] * x[r][OFF + H + i];
      }
      p[i] = ts_keep_add(a, b, hi, H);
    }
  }
#pragma unroll
  for (int m = H / 2; m >= 1; m >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < m; i++) p[i] = ts_keep_add(p[i], p[i + m], hi, m);
  }
#pragma unroll
  for (int m = KS; m < 64; m <<= 1) p[0] += __shfl_xor(p[0], m, 64);
  return p[0];
}

template <typename T, int K, int RPL, int OFF>
__device__ __forceinline__ void tw_factor_phase(T (&x)[RPL][K], int j0, int j1, int lane, T &tau_l) {
#pragma nounroll
  for (int j = j0; j < j1; j++) {
    T xj[RPL];
    T s = 0;
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      T e = x[r][OFF];
#pragma unroll
      for (int c = OFF + 1; c < K; c++) e = c == j ? x[r][c] : e;
      xj[r] = e;
      s += (lane + 64 * r > j) ? e * e : (T)0;
    }
    const T sigma = ts_wave_sum(s);
    const T alpha = ts_readlane(xj[0], j);       // row j: lane j, first row slot (j < K <= 64)
    T tau = 0;
    if (sigma != (T)0) {                         // else H = I, as dlarfg does for a zero column
      const T nrm = sqrt(alpha * alpha + sigma);
      const T beta = alpha >= (T)0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      const T scal = (T)1 / (alpha - beta);
      T v[RPL];
#pragma unroll
      for (int r = 0; r < RPL; r++) {
        const int row = lane + 64 * r;
        v[r] = row > j ? xj[r] * scal : (row == j ? (T)1 : (T)0);
      }
      const T ws = tw_dots<T, K, RPL, OFF>(x, v, lane);
#pragma unroll
      for (int c = OFF; c < K; c++) {
        const T w = ts_readlane(ws, c - OFF);
        if (c > j) {
#pragma unroll
          for (int r = 0; r < RPL; r++) x[r][c] -= (tau * v[r]) * w;
        }
      }
#pragma unroll
      for (int r = 0; r < RPL; r++) {
        const int row = lane + 64 * r;
        const T nj = row > j ? v[r] : (row == j ? beta : xj[r]);
#pragma unroll
        for (int c = OFF; c < K; c++)
          if (c == j) x[r][c] = nj;
      }
    }
    if (lane == j) tau_l = tau;
  }
}

template <typename T, int K, int RPL, int OFF>
__device__ __forceinline__ void tw_factor_all(T (&x)[RPL][K], int kk, int lane, T &tau_l) {
  constexpr int KS = K - OFF, JE = KS == 8 ? K : OFF + KS / 2;
  tw_factor_phase<T, K, RPL, OFF>(x, OFF, min(kk, JE), lane, tau_l);
  if constexpr (KS > 8) tw_factor_all<T, K, RPL, JE>(x, kk, lane, tau_l);
}

// tsqr_factor_kernel with one wave per tile: G tiles of 64 RPL rows.
template <typename T, int K, int RPL>
__global__ __launch_bounds__(BLK) void tsqr_factor_wave_kernel(const T *A, int64_t lda, int64_t rows, int k, T *Y,
                                                               int64_t ldy, T *__restrict__ tau, T *__restrict__ R,
                                                               int64_t ldr, int64_t G) {
  constexpr int TR = 64 * RPL;
  static_assert(K <= 64, "row j of a tile must sit in lane j");
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * (BLK / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (b >= G) return;                           // the whole wave: no barrier follows
  const int64_t r0 = b * TR;
  const int nb = (int)min((int64_t)TR, rows - r0);
  const int kk = min(k, nb);
  T x[RPL][K];
#pragma unroll
  for (int r = 0; r < RPL; r++) {
    const int row = lane + 64 * r;
#pragma unroll
    for (int c = 0; c < K; c++) x[r][c] = (c < k && row < nb) ? A[r0 + row + (int64_t)c * lda] : (T)0;
  }
  T tau_l = 0;
  tw_factor_all<T, K, RPL, 0>(x, kk, lane, tau_l);
  if (Y) {
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      const int row = lane + 64 * r;
      if (row < nb) {
#pragma unroll
        for (int c = 0; c < K; c++)
          if (c < k) Y[r0 + row + (int64_t)c * ldy] = x[r][c];
      }
    }
  }
  if (lane < k) {                               // rows 0..k-1 of the tile: first row slot
    const int64_t rr = b * k + lane;
#pragma unroll
    for (int c = 0; c < K; c++)
      if (c < k) R[rr + (int64_t)c * ldr] = (lane < nb && c >= lane) ? x[0][c] : (T)0;
    tau[rr] = lane < kk ? tau_l : (T)0;
  }
}

// Applies H_j for j1 > j >= j0 (descending) to the rows q of the tile; the
// Householder vector of step j is read one step ahead.
template <typename T, int K, int RPL, int OFF>
__device__ __forceinline__ void tw_form_phase(T (&q)[RPL][K], int j1, int j0, int lane, const T *ytile, int64_t ldy,
                                              int nb, const T *tau_b) {
  T yn[RPL];
  if (j1 - 1 >= j0) {
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      const int row = lane + 64 * r;
      yn[r] = row < nb ? ytile[row + (int64_t)(j1 - 1) * ldy] : (T)0;
    }
  }
#pragma nounroll
  for (int j = j1 - 1; j >= j0; j--) {
    T y[RPL];
#pragma unroll
    for (int r = 0; r < RPL; r++) y[r] = yn[r];
    if (j - 1 >= j0) {
#pragma unroll
      for (int r = 0; r < RPL; r++) {
        const int row = lane + 64 * r;
        yn[r] = row < nb ? ytile[row + (int64_t)(j - 1) * ldy] : (T)0;
      }
    }
    const T tj = tau_b[j];
    if (tj == (T)0) continue;
    T v[RPL];
#pragma unroll
    for (int r = 0; r < RPL; r++) {
      const int row = lane + 64 * r;
      v[r] = row > j ? y[r] : (row == j ? (T)1 : (T)0);
    }
    const T ws = tw_dots<T, K, RPL, OFF>(q, v, lane);
#pragma unroll
    for (int c = OFF; c < K; c++) {
      const T w = ts_readlane(ws, c - OFF);
#pragma unroll
      for (int r = 0; r < RPL; r++) q[r][c] -= (tj * v[r]) * w;
    }
  }
}

template <typename T, int K, int RPL, int OFF>
__device__ __forceinline__ void tw_form_all(T (&q)[RPL][K], int kk, int lane, const T *ytile, int64_t ldy, int nb,
                                            const T *tau_b) {
  constexpr int KS = K - OFF, JE = KS == 8 ? K : OFF + KS / 2;
  if constexpr (KS > 8) tw_form_all<T, K, RPL, JE>(q, kk, lane, ytile, ldy, nb, tau_b);
  tw_form_phase<T, K, RPL, OFF>(q, min(kk, JE), OFF, lane, ytile, ldy, nb, tau_b);
}

// tsqr_form_kernel with one wave per tile: G tiles of 64 RPL rows.
template <typename T, int K, int RPL>
__global__ __launch_bounds__(BLK) void tsqr_form_wave_kernel(const T *Y, int64_t ldy, int64_t rows, int k,
                                                             const T *__restrict__ tau, const T *__restrict__ C,
                                                             int64_t ldc, T *out, int64_t ldo, const T *__restrict__ sgn,
                                                             int64_t G) {
  constexpr int TR = 64 * RPL;
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * (BLK / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (b >= G) return;
  const int64_t r0 = b * TR;
  const int nb = (int)min((int64_t)TR, rows - r0);
  const int kk = min(k, nb);
  const T *ytile = Y + r0;
  const T *tau_b = tau + b * k;
  T q[RPL][K];
  if (C) {
    // rows of C_b beyond the tile's rows meet zero rows of its R
#pragma unroll
    for (int r = 0; r < RPL; r++)
#pragma unroll
      for (int c = 0; c < K; c++)
        q[r][c] = (r == 0 && lane < kk && c < k) ? C[b * k + lane + (int64_t)c * ldc] : (T)0;
    tw_form_phase<T, K, RPL, 0>(q, kk, 0, lane, ytile, ldy, nb, tau_b);
  } else {
#pragma unroll
    for (int r = 0; r < RPL; r++)
#pragma unroll
      for (int c = 0; c < K; c++) q[r][c] = (lane + 64 * r == c) ? (T)1 : (T)0;
    tw_form_all<T, K, RPL, 0>(q, kk, lane, ytile, ldy, nb, tau_b);
  }
#pragma unroll
  for (int r = 0; r < RPL; r++) {
    const int row = lane + 64 * r;
    if (row < nb) {
#pragma unroll
      for (int c = 0; c < K; c++)
        if (c < k) out[r0 + row + (int64_t)c * ldo] = sgn ? q[r][c] * sgn[c] : q[r][c];
    }
  }
}

// ---- host side ----------------------------------------------------------------

void tsqr_signs(const double *Qtop, const double *R, int64_t n, int k, double *f) {
  std::vector<double> U(Qtop, Qtop + (size_t)k * k), D(k);
  for (int j = 0; j < k; j++) {
    D[j] = R[j + (size_t)j * k] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < k; i++) U[i + (size_t)j * k] *= D[j];
  }
  for (int j = 0; j < k; j++) {
    const double piv = U[j + (size_t)j * k];
    // dlarfg: beta = -sign(alpha) ||x||; a column with no rows below the
    // diagonal (the last of a square block) is left as it is (tau = 0)
    const double sj = (j == n - 1) ? (piv < 0.0 ? -1.0 : 1.0) : (piv < 0.0 ? 1.0 : -1.0);
    f[j] = D[j] * sj;
    const double d = piv - sj;
    U[j + (size_t)j * k] = d;
    for (int i = j + 1; i < k; i++) U[i + (size_t)j * k] /= d;
    for (int c = j + 1; c < k; c++)
      for (int i = j + 1; i < k; i++) U[i + (size_t)c * k] -= U[i + (size_t)j * k] * U[j + (size_t)c * k];
  }
}

namespace {

// Tree workspace, one per host thread (each has its own stream): kept between
// calls, grown when a call needs more, released with the process.
struct TsWork {
  void *buf = nullptr;
  size_t bytes = 0;
};

template <typename T>
struct TsLevel {
  T *M;
  int64_t ld, rows;
  int G;
  T *tau;
};

// Up to 32 columns the tree uses the one-wave tile kernels (128-row tiles
// for 17..32 columns, 256-row tiles below: RPL rows per lane in registers);
// wider blocks use the workgroup kernels on 256-row tiles.
template <typename T, int K>
int tsqr_t(BV &V, int l, int k, bool formq, double *R_host, double *f, hipStream_t s) {
  constexpr int RPL = K <= 16 ? 4 : 2;
  constexpr int TR = K <= 32 ? 64 * RPL : TS_ROWS;
  auto factor = [&](const T *A, int64_t lda, int64_t rows, int G, T *Y, int64_t ldy, T *tau, T *R, int64_t ldr) {
    if constexpr (K <= 32)
      tsqr_factor_wave_kernel<T, K, RPL><<<(G + 3) / 4, BLK, 0, s>>>(A, lda, rows, k - l, Y, ldy, tau, R, ldr, G);
    else
      tsqr_factor_kernel<T, K><<<G, BLK, 0, s>>>(A, lda, rows, k - l, Y, ldy, tau, R, ldr);
  };
  auto form = [&](const T *Y, int64_t ldy, int64_t rows, int G, const T *tau, const T *C, int64_t ldc, T *out,
                  int64_t ldo, const T *sg) {
    if constexpr (K <= 32)
      tsqr_form_wave_kernel<T, K, RPL><<<(G + 3) / 4, BLK, 0, s>>>(Y, ldy, rows, k - l, tau, C, ldc, out, ldo, sg, G);
    else
      tsqr_form_kernel<T, K><<<G, BLK, 0, s>>>(Y, ldy, rows, k - l, tau, C, ldc, out, ldo, sg);
  };
  const int nk = k - l;
  const int64_t n = V.n;
  // level sizes: rows_{i+1} = tiles_i * nk until one tile is left
  std::vector<TsLevel<T>> L;
  size_t need = 0;
  auto al = [](size_t e) { return (e + 63) / 64 * 64; };
  for (int64_t rows = n;;) {
    const int64_t G = (rows + TR - 1) / TR;
    SA_CHECK(G <= (int64_t)1 << 30, "TSQR: too many tiles");
    L.push_back({nullptr, 0, rows, (int)G, nullptr});
    need += al((size_t)G * nk);                             // tau
    if (G == 1) break;
    rows = G * nk;
    need += al((size_t)rows * nk);                          // next level's matrix
  }
  need += al((size_t)nk * nk) + al((size_t)TR * nk) + al(nk);   // R, probe tile, signs
  thread_local TsWork W;
  if (W.bytes < need * sizeof(T)) {
    if (W.buf) (void)hipFree(W.buf);
    W.buf = nullptr;
    W.bytes = 0;
    SA_HIP(hipMalloc(&W.buf, need * sizeof(T)));
    W.bytes = need * sizeof(T);
  }
  T *p = (T *)W.buf;
  L[0].M = (T *)V.data + (int64_t)l * V.ld;
  L[0].ld = V.ld;
  for (size_t i = 0; i < L.size(); i++) {
    L[i].tau = p;
    p += al((size_t)L[i].G * nk);
    if (i + 1 < L.size()) {
      L[i + 1].M = p;
      L[i + 1].ld = L[i + 1].rows;
      p += al((size_t)L[i + 1].rows * nk);
    }
  }
  T *Rtop = p;
  p += al((size_t)nk * nk);
  T *probe = p;
  p += al((size_t)TR * nk);
  T *sgn = p;
  const int top = (int)L.size() - 1;
  for (int i = 0; i <= top; i++) {
    T *Rout = i == top ? Rtop : L[i + 1].M;
    const int64_t ldr = i == top ? nk : L[i + 1].ld;
    factor(L[i].M, L[i].ld, L[i].rows, L[i].G, formq ? L[i].M : nullptr, L[i].ld, L[i].tau, Rout, ldr);
  }
  SA_HIP(hipGetLastError());
  std::vector<T> buf((size_t)nk * nk);
  SA_HIP(hipMemcpyAsync(buf.data(), Rtop, sizeof(T) * buf.size(), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  for (size_t e = 0; e < buf.size(); e++) R_host[e] = (double)buf[e];
  if (!formq) return 0;
  for (int i = top; i >= 1; i--)
    form(L[i].M, L[i].ld, L[i].rows, L[i].G, L[i].tau, i == top ? nullptr : L[i + 1].M, i == top ? 0 : L[i + 1].ld,
         L[i].M, L[i].ld, nullptr);
  const T *C0 = top == 0 ? nullptr : L[1].M;
  const int64_t ldc0 = top == 0 ? 0 : L[1].ld;
  // the first nk rows of Q (tile 0 holds them: n >= nk) for the signs
  form(L[0].M, L[0].ld, L[0].rows, 1, L[0].tau, C0, ldc0, probe, TR, nullptr);
  SA_HIP(hipGetLastError());
  SA_HIP(hipMemcpy2DAsync(buf.data(), sizeof(T) * nk, probe, sizeof(T) * TR, sizeof(T) * nk, nk,
                          hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  std::vector<double> Qtop(buf.begin(), buf.end());
  tsqr_signs(Qtop.data(), R_host, n, nk, f);
  for (int j = 0; j < nk; j++) buf[j] = (T)f[j];
  SA_HIP(hipMemcpyAsync(sgn, buf.data(), sizeof(T) * nk, hipMemcpyHostToDevice, s));
  form(L[0].M, L[0].ld, L[0].rows, L[0].G, L[0].tau, C0, ldc0, L[0].M, L[0].ld, sgn);
  SA_HIP(hipGetLastError());
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

template <typename T>
int tsqr_k(BV &V, int l, int k, bool formq, double *R_host, double *f, hipStream_t s) {
  const int nk = k - l;
  if (nk <= 8) return tsqr_t<T, 8>(V, l, k, formq, R_host, f, s);
  if (nk <= 16) return tsqr_t<T, 16>(V, l, k, formq, R_host, f, s);
  if (nk <= 32) return tsqr_t<T, 32>(V, l, k, formq, R_host, f, s);
  return tsqr_t<T, 64>(V, l, k, formq, R_host, f, s);
}

}  // namespace

int launch_tsqr(BV &V, int l, int k, bool formq, double *R_host, double *f, hipStream_t s) {
  const int nk = k - l;
  SA_CHECK(nk >= 1 && nk <= TS_MAXK, "TSQR: at most 64 active columns per call");
  SA_CHECK(V.n >= nk, "TSQR: fewer rows than active columns");
  return V.prec == SA_FP64 ? tsqr_k<double>(V, l, k, formq, R_host, f, s)
                           : tsqr_k<float>(V, l, k, formq, R_host, f, s);
}

int bv_top_block(const BV &V, int l, int k, double *out) {
  const int nk = k - l;
  const size_t es = V.prec == SA_FP64 ? sizeof(double) : sizeof(float);
  std::vector<char> buf(es * nk * nk);
  SA_HIP(hipMemcpy2D(buf.data(), es * nk, (const char *)V.data + es * (size_t)l * V.ld, es * V.ld, es * nk, nk,
                     hipMemcpyDeviceToHost));
  for (size_t e = 0; e < (size_t)nk * nk; e++)
    out[e] = V.prec == SA_FP64 ? ((const double *)buf.data())[e] : (double)((const float *)buf.data())[e];
  return 0;
}

}  // namespace sa
