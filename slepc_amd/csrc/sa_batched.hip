// Batched Krylov-Schur for many independent symmetric problems of uneven size.
//
// All problems live in one block-diagonal CSR operator and one basis V whose
// rows are the concatenation of the problems' rows (each problem padded to a
// multiple of 64 rows).  One restart phase is ONE launch of the sweep kernel:
// workgroup b runs the complete Arnoldi/Lanczos sweep (SpMV, classical
// Gram-Schmidt with DGKS refinement, normalisation) of problem order[b], so all
// synchronisation is inside the workgroup and no grid-wide barrier exists.
// Problems are ordered largest first so the hardware scheduler balances CUs.
//
// A second sweep ("wide", lock-step) runs each Gram-Schmidt slot of every
// active problem as ONE launch over the 64-row chunks of all those problems
// (the slot kernels of sa_gs_device.h with per-problem state), so problems
// with many rows are spread over many workgroups.  An autotuner times both
// sweeps on the first restart round of a solve and keeps the faster one.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "sa_ds.h"
#include "sa_gs_device.h"
#include "sa_internal.h"

namespace sa {


template <typename T>
struct BatchArgs {
  T *V;
  int64_t ldv;
  const int32_t *rp, *ci;
  const T *val;
  const int *order;
  const int64_t *off;
  const int *n;
  const int *k0, *nv;
  int spmv;
  double *dbuf;
  int m;
  double *beta;
  int *stop;
  int *passes;
  double eta;
};

__device__ __forceinline__ double bsafe_sqrt(double a) { return a < 0.0 ? 0.0 : sqrt(a); }

// Sums acc[0..ncol) over the workgroup into out[0..ncol) (FP64).
template <typename T, int CM, int NT>
__device__ __forceinline__ void block_sum(const T (&acc)[CM + 1], int ncol, double *red, double *out) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  constexpr int NW = NT / 64;
#pragma unroll
  for (int i = 0; i <= CM; i++) {
    if (i < ncol) {
      double v = (double)acc[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[w * (CM + 1) + i] = v;
    }
  }
  __syncthreads();
  for (int i = t; i < ncol; i += NT) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NW; q++) s += red[q * (CM + 1) + i];
    out[i] = s;
  }
  __syncthreads();
}

template <typename T, int CM, int NT, int L>
__global__ __launch_bounds__(NT) void batched_sweep_kernel(BatchArgs<T> a) {
  __shared__ double s_red[(NT / 64) * (CM + 1)];
  __shared__ double s_h[CM + 1];      // dots of the current pass (index c: v^T v)
  __shared__ double s_hacc[CM + 1];   // accumulated coefficients of the column
  __shared__ T s_coef[CM];
  __shared__ int s_mode;              // 0 final (mult+scale), 1 continue (mult+dots), 2 pending (mult+dots, norm check)
  __shared__ double s_scale, s_onrm, s_nrm;
  __shared__ int s_pass, s_done;
  const int t = threadIdx.x;
  const int p = a.order[blockIdx.x];
  const int n = a.n[p];
  const int64_t off = a.off[p];
  const int k0 = a.k0[p], nv = a.nv[p];
  const int cfirst = a.spmv ? k0 + 1 : k0, clast = a.spmv ? nv : k0;
  double *dbuf = a.dbuf + (size_t)p * a.m * a.m;
  if (cfirst > clast || a.stop[p] >= 0) return;
  for (int c = cfirst; c <= clast; c++) {
    T *vc = a.V + (size_t)c * a.ldv + off;
    // ---- SpMV: v_c = A v_{c-1} (global column indices of the block-diagonal operator)
    if (a.spmv) {
      const T *x = a.V + (size_t)(c - 1) * a.ldv;
      constexpr int RPS = NT / L;
      const int g = t / L, lg = t % L;
      for (int r0 = 0; r0 < n; r0 += RPS) {
        const int r = r0 + g;
        T s = (T)0;
        if (r < n) {
          const int64_t row = off + r;
          const int32_t pe = a.rp[row + 1];
          for (int32_t q = a.rp[row] + lg; q < pe; q += L) s += a.val[q] * x[a.ci[q]];
        }
#pragma unroll
        for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
        if (lg == 0 && r < n) vc[r] = s;
      }
      __syncthreads();
    }
    // ---- pass-1 dot products V(:,0..c)^T v_c
    {
      T acc[CM + 1];
#pragma unroll
      for (int i = 0; i <= CM; i++) acc[i] = (T)0;
      for (int r = t; r < n; r += NT) {
        const T v = vc[r];
#pragma unroll
        for (int i = 0; i < CM; i++)
          if (i < c) acc[i] += a.V[(size_t)i * a.ldv + off + r] * v;
        acc[CM] += v * v;
      }
      if (c < CM) acc[c] = acc[CM];
      block_sum<T, CM, NT>(acc, c + 1, s_red, s_h);
    }
    for (int i = t; i < c; i += NT) s_hacc[i] = 0.0;
    if (t == 0) { s_pass = 0; s_done = 0; }
    __syncthreads();
    // ---- CGS1 passes with DGKS refinement (bvorthog.c:91-217)
    while (true) {
      if (t == 0) {
        const double self = s_h[c];
        double sum = 0.0;
        for (int i = 0; i < c; i++) sum += s_h[i] * s_h[i];
        s_pass++;
        const double onrm = bsafe_sqrt(self);
        const double est = onrm * onrm - sum;
        s_onrm = onrm;
        if (est <= 0.0) {
          s_mode = 2;
          s_scale = 1.0;
        } else {
          const double nrm = sqrt(est);
          if (nrm != 0.0 && nrm < a.eta * onrm && s_pass < 3) {
            s_mode = 1;
            s_scale = 1.0;
          } else {
            s_mode = 0;
            s_nrm = nrm;
            s_scale = (nrm != 0.0 && nrm != 1.0) ? 1.0 / nrm : 1.0;
          }
        }
      }
      __syncthreads();
      for (int i = t; i < c; i += NT) {
        s_hacc[i] += s_h[i];
        s_coef[i] = (T)s_h[i];
      }
      __syncthreads();
      const int mode = s_mode;
      const T scl = (T)s_scale;
      T acc[CM + 1];
#pragma unroll
      for (int i = 0; i <= CM; i++) acc[i] = (T)0;
      for (int r = t; r < n; r += NT) {
        T s = (T)0;
#pragma unroll
        for (int i = 0; i < CM; i++)
          if (i < c) s += a.V[(size_t)i * a.ldv + off + r] * s_coef[i];
        const T vn = (vc[r] - s) * scl;
        vc[r] = vn;
        if (mode != 0) {
#pragma unroll
          for (int i = 0; i < CM; i++)
            if (i < c) acc[i] += a.V[(size_t)i * a.ldv + off + r] * vn;
          acc[CM] += vn * vn;
        }
      }
      if (mode == 0) break;
      if (c < CM) acc[c] = acc[CM];
      __syncthreads();
      block_sum<T, CM, NT>(acc, c + 1, s_red, s_h);
      if (mode == 2) {   // the estimate failed: the new v^T v is the true norm
        if (t == 0) {
          const double nrm = bsafe_sqrt(s_h[c]);
          if (!(nrm != 0.0 && nrm < a.eta * s_onrm && s_pass < 3)) {
            s_done = 1;
            s_nrm = nrm;
            s_scale = (nrm != 0.0 && nrm != 1.0) ? 1.0 / nrm : 1.0;
          }
        }
        __syncthreads();
        if (s_done) {
          const T sc = (T)s_scale;
          for (int r = t; r < n; r += NT) vc[r] *= sc;
          break;
        }
      }
    }
    __syncthreads();
    // ---- store coefficients and the norm (bvorthog.c:209-215), breakdown test
    int lind = 0;
    {
      const double nrm = s_nrm, onrm = s_onrm;
      lind = !(nrm != 0.0 && fabs(nrm) >= a.eta * fabs(onrm));
      for (int i = t; i < c; i += NT) dbuf[(size_t)c * a.m + i] = s_hacc[i];
      if (t == 0) {
        dbuf[(size_t)c * a.m + c] = lind ? 0.0 : nrm;
        a.beta[p] = nrm;
        a.passes[p] += s_pass;
        if (lind) a.stop[p] = c;
      }
    }
    __syncthreads();
    if (lind) break;
  }
}

// In-place restart GEMM per problem: V_p(:,s:e) = V_p(:,l:k) Q_p(l:k,s:e), one
// workgroup per (problem, 64-row chunk); FP64/FP32 via plain FMA (the block is
// at most 128 x 128 and the kernel is bound by the V traffic).
template <typename T>
__global__ __launch_bounds__(256) void batched_mip_kernel(T *V, int64_t ldv, const int *chunk_prob,
                                                          const int *chunk_row, const int64_t *off, const int *n,
                                                          const int *ml, const int *mk, const int *ms, const int *me,
                                                          const double *Q, int m) {
  __shared__ T As[64][65];   // K <= 64 columns (ncv < 64)
  const int g = blockIdx.x, p = chunk_prob[g], r0 = chunk_row[g];
  const int l = ml[p], k = mk[p], s = ms[p], e = me[p];
  const int K = k - l, N = e - s;
  if (K <= 0 || N <= 0) return;
  const int rows = min(64, n[p] - r0);
  T *Vp = V + off[p] + r0;
  const int t = threadIdx.x, row = t & 63, cg = t >> 6;
  for (int i = cg; i < K; i += 4) As[i][row] = (row < rows) ? Vp[(size_t)(l + i) * ldv + row] : (T)0;
  __syncthreads();
  const double *Qp = Q + (size_t)p * m * m;
  for (int j = cg; j < N; j += 4) {
    T acc = (T)0;
    for (int i = 0; i < K; i++) acc += As[i][row] * (T)Qp[(size_t)(l + i) + (size_t)(s + j) * m];
    if (row < rows) Vp[(size_t)(s + j) * ldv + row] = acc;
  }
}

template <typename T>
__global__ void batched_copy_kernel(T *V, int64_t ldv, const int *chunk_prob, const int *chunk_row, const int64_t *off,
                                    const int *n, const int *src, const int *dst) {
  const int g = blockIdx.x, p = chunk_prob[g], r = chunk_row[g] + (int)threadIdx.x;
  if (src[p] < 0 || r >= n[p]) return;
  V[(size_t)dst[p] * ldv + off[p] + r] = V[(size_t)src[p] * ldv + off[p] + r];
}

template <typename T>
__global__ void batched_scale_kernel(T *V, int64_t ldv, const int *chunk_prob, const int *chunk_row,
                                     const int64_t *off, const int *n, const int *col, const double *alpha) {
  const int g = blockIdx.x, p = chunk_prob[g], r = chunk_row[g] + (int)threadIdx.x;
  if (col[p] < 0 || r >= n[p]) return;
  V[(size_t)col[p] * ldv + off[p] + r] *= (T)alpha[p];
}

// ---------------------------------------------------------------------------
// True residuals of candidate Ritz vectors (EPSSetTrueResidual,
// epskrylov.c:225-265) for every problem at once: slot s of problem p holds
// the coefficients z of one candidate (a column of the projected problem's
// eigenvectors) and its eigenvalue.  batched_ritz_kernel forms x = V_p(:,0:nv_p) z
// into column s of X (the problem's rows), batched_tres_kernel takes
// ||A x - lambda x||^2 over each 64-row chunk of the block-diagonal operator
// (the problem's rows only), and batched_tres_sum_kernel adds a problem's
// chunk partials in a fixed order.
// ---------------------------------------------------------------------------
constexpr int TR_SLOTS = 4;

template <typename T>
__global__ __launch_bounds__(64) void batched_ritz_kernel(const T *V, int64_t ldv, const int *chunk_prob,
                                                          const int *chunk_row, const int64_t *off, const int *n,
                                                          const int *tnv, const int *tact, const double *Z, int m,
                                                          T *X) {
  const int g = blockIdx.x, sl = blockIdx.y, p = chunk_prob[g];
  if (!tact[p * TR_SLOTS + sl]) return;
  const int r = chunk_row[g] + (int)threadIdx.x;
  if (r >= n[p]) return;
  const double *z = Z + ((size_t)p * TR_SLOTS + sl) * m;
  const T *vp = V + off[p] + r;
  double acc = 0.0;
  for (int i = 0; i < tnv[p]; i++) acc += (double)vp[(size_t)i * ldv] * z[i];
  X[(size_t)sl * ldv + off[p] + r] = (T)acc;
}

template <typename T, int L>
__global__ __launch_bounds__(256) void batched_tres_kernel(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                           const T *__restrict__ val, const T *X, int64_t ldv,
                                                           const int *chunk_prob, const int *chunk_row,
                                                           const int64_t *off, const int *n, const int *tact,
                                                           const double *lam, double *part, int nchunks) {
  __shared__ double red[4];
  const int g = blockIdx.x, sl = blockIdx.y, p = chunk_prob[g];
  if (!tact[p * TR_SLOTS + sl]) return;   // uniform over the workgroup
  const T *x = X + (size_t)sl * ldv;
  const double lm = lam[p * TR_SLOTS + sl];
  const int r0 = chunk_row[g], rows = min(64, n[p] - r0), lg = threadIdx.x % L;
  double acc = 0.0;
  for (int rr = threadIdx.x / L; rr < 64; rr += 256 / L) {
    const int64_t row = off[p] + r0 + rr;
    T sum = (T)0;
    if (rr < rows)
      for (int32_t q = rp[row] + lg; q < rp[row + 1]; q += L) sum += val[q] * x[ci[q]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, L);
    if (lg == 0 && rr < rows) {
      const double rv = (double)sum - lm * (double)x[row];
      acc += rv * rv;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)sl * nchunks + g] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(64) void batched_tres_sum_kernel(const double *part, int nchunks, const int *pchunk0,
                                                              const int *pnchunk, const int *tact, double *out) {
  const int p = blockIdx.x, sl = blockIdx.y;
  if (!tact[p * TR_SLOTS + sl]) return;
  const double *pp = part + (size_t)sl * nchunks + pchunk0[p];
  double acc = 0.0;
  for (int i = threadIdx.x; i < pnchunk[p]; i += 64) acc += pp[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (threadIdx.x == 0) out[p * TR_SLOTS + sl] = acc;
}

// ---------------------------------------------------------------------------
// Lock-step ("wide") sweep.  Step j orthogonalises column c_p = k0[p]+1+j of
// every listed problem p (c_p = k0[p] without SpMV); problems whose sweep is
// shorter keep their state chain and skip the work.  One workgroup per active
// 64-row chunk; per problem, partial sums are [c+1][G_p] at chunk offset
// pc0[p] * m, reduced dots are [m] at p * m, states are GS_NSTATE at p.
// ---------------------------------------------------------------------------
template <typename T>
struct BwArgs {
  T *V;
  int64_t ldv;
  const int32_t *rp, *ci;
  const T *val;
  const int64_t *off;
  const int *n, *k0, *nv;
  const int *cprob, *crow;    // active chunks: problem, first local row
  const int *pc0, *pnc;       // per problem: first active chunk, number of chunks
  int j, spmv, slot, m;
  GsState *st;
  double *dbuf;
  double *part_out;
  const double *dots_in;
  double eta;
};

__device__ __forceinline__ int bw_col(const int *k0, int p, int spmv, int j) { return k0[p] + (spmv ? 1 + j : 0); }

template <typename T, int L, bool NT>
__global__ __launch_bounds__(BLK) void bw_slot_kernel(BwArgs<T> b) {
  const int g = blockIdx.x, p = b.cprob[g];
  const int c = bw_col(b.k0, p, b.spmv, b.j);
  const int bidx = g - b.pc0[p];
  GsState *st = b.st + (size_t)GS_NSTATE * p;
  if (c > b.nv[p]) {           // this problem's sweep ended: carry its state forward
    if (bidx == 0 && threadIdx.x == 0) {
      GsState S = st[(b.slot + 1) % GS_NSTATE];
      S.done = 1;
      S.dots_ready = 0;
      st[b.slot - 1] = S;
    }
    return;
  }
  GsArgs<T> a;
  a.n = b.n[p];
  a.ld = b.ldv;
  a.V = b.V + b.off[p];
  a.c = c;
  a.slot = b.slot;
  const bool sp = b.spmv && b.slot == 1;
  a.rp = sp ? b.rp + b.off[p] : nullptr;      // row pointers of the problem's rows (global nnz index)
  a.ci = sp ? b.ci : nullptr;                 // global column indices of the block-diagonal operator
  a.val = sp ? b.val : nullptr;
  a.xin = sp ? b.V + (size_t)(c - 1) * b.ldv : nullptr;
  a.st = st;
  a.dbuf = b.dbuf + (size_t)p * b.m * b.m;
  a.m = b.m;
  a.dots_in = b.dots_in + (size_t)p * b.m;
  a.part_out = b.part_out + (size_t)b.pc0[p] * b.m;
  a.G = b.pnc[p];
  a.rpb = CH;
  a.eta = b.eta;
  a.refine = GS_REFINE_IFNEEDED;
  const int64_t rb = b.crow[g];
  gs_slot<T, L, NT>(a, rb, min(a.n, rb + CH), bidx);
}

__global__ __launch_bounds__(BLK) void bw_sum_kernel(const int *list, const int *k0, const int *nv, int spmv, int j,
                                                     int slot, const GsState *st, const double *part, const int *pc0,
                                                     const int *pnc, int m, double *dots) {
  const int p = list[blockIdx.y], col = blockIdx.x;
  const int c = bw_col(k0, p, spmv, j);
  if (c > nv[p] || col > c) return;
  // (summed even when the slot produced no dots: the next slot then ignores them)
  gs_sum_column(part + (size_t)pc0[p] * m, pnc[p], col, dots + (size_t)p * m);
}

__global__ void bw_reset_kernel(const int *list, int nl, GsState *st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  GsState &S = st[(size_t)GS_NSTATE * list[i] + 2];
  S.stop = 0; S.stop_col = 0; S.pending = 0; S.lindep = 0; S.pass = 0;
  S.done = 0; S.dots_ready = 0; S.beta = 0.0;
}

// Results of a wide sweep in the persistent sweep's output arrays.
__global__ void bw_export_kernel(const int *list, int nl, GsState *st, int *stop, double *beta, int *passes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const int p = list[i];
  GsState &S = st[(size_t)GS_NSTATE * p + 2];
  stop[p] = S.stop ? S.stop_col : -1;
  beta[p] = S.beta;
  passes[p] += (int)S.passes_total;
  S.passes_total = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// A lock-step sweep replayed from a HIP graph: the signature fixes every
// launch of the sweep (grids, LDS sizes, step indices, schedule); the data it
// reads (problem list, k0, nv, chunk lists) lives in device buffers whose
// addresses the graph keeps, refreshed by copies before each launch.
struct SweepGraph {
  std::vector<int> sig;
  hipGraphExec_t exec = nullptr;
  uint64_t used = 0;
};
constexpr size_t SWEEP_GRAPHS_MAX = 256;

// Sweep schedule: kind 1 one workgroup per problem (persistent sweep), 2
// lock-step; for lock-step, nt selects non-temporal, wave-uniform tile loads
// and lanes the lanes per row of the SpMV fused into slot 1.
struct Sched {
  int kind = 2, nt = 0, lanes = 8;
};

struct BatchDev {
  int P = 0, m = 0, prec = SA_FP64;
  int64_t ldv = 0, nrows = 0, nnz = 0;
  void *V = nullptr;
  int32_t *rp = nullptr, *ci = nullptr;
  void *val = nullptr;
  int lanes = 4;
  int64_t *off = nullptr;
  int *n = nullptr, *order = nullptr, *k0 = nullptr, *nv = nullptr, *stop = nullptr, *passes = nullptr;
  double *dbuf = nullptr, *beta = nullptr, *Q = nullptr, *alpha = nullptr;
  int *chunk_prob = nullptr, *chunk_row = nullptr, nchunks = 0;
  int *ml = nullptr, *mk = nullptr, *ms = nullptr, *me = nullptr, *src = nullptr, *dst = nullptr;
  // lock-step sweep
  GsState *wst = nullptr;
  double *wpart = nullptr, *wdots = nullptr;   // 2 x [nchunks][m], 2 x [P][m]
  int *cprob = nullptr, *crow = nullptr, *pc0 = nullptr, *pnc = nullptr;
  int wm = 0;                                  // basis width the wide buffers were sized for
  // true residuals: candidate columns, coefficients, eigenvalues, partials
  void *tx = nullptr;                          // TR_SLOTS columns of ldv
  double *tz = nullptr, *tlam = nullptr, *tpart = nullptr, *tout = nullptr;
  int *tnv = nullptr, *tact = nullptr, *pchunk0 = nullptr, *pnchunk = nullptr;
  int *passes_save = nullptr;                  // copy of passes while the autotuner reruns a sweep
  // lock-step sweeps captured as HIP graphs, keyed by their launch signature
  std::vector<SweepGraph> graphs;
  bool graphs_on = true;
  uint64_t graph_clock = 0, graph_captures = 0, graph_replays = 0;
};

static void clear_graphs(BatchDev &B) {
  for (auto &g : B.graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  B.graphs.clear();
}

template <typename T, int CM, int NT>
static void launch_sweep_t(const BatchDev &B, int nact, int spmv, double eta, hipStream_t s) {
  BatchArgs<T> a;
  a.V = (T *)B.V; a.ldv = B.ldv; a.rp = B.rp; a.ci = B.ci; a.val = (const T *)B.val;
  a.order = B.order; a.off = B.off; a.n = B.n; a.k0 = B.k0; a.nv = B.nv; a.spmv = spmv;
  a.dbuf = B.dbuf; a.m = B.m; a.beta = B.beta; a.stop = B.stop; a.passes = B.passes; a.eta = eta;
  switch (B.lanes) {
    case 4: hipLaunchKernelGGL((batched_sweep_kernel<T, CM, NT, 4>), dim3(nact), dim3(NT), 0, s, a); break;
    case 8: hipLaunchKernelGGL((batched_sweep_kernel<T, CM, NT, 8>), dim3(nact), dim3(NT), 0, s, a); break;
    case 16: hipLaunchKernelGGL((batched_sweep_kernel<T, CM, NT, 16>), dim3(nact), dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((batched_sweep_kernel<T, CM, NT, 32>), dim3(nact), dim3(NT), 0, s, a); break;
  }
}

static int launch_sweep(const BatchDev &B, int nact, int spmv, double eta, hipStream_t s) {
  if (nact == 0) return 0;
  const int cmax = B.m - 1;   // largest column index
  if (B.prec == SA_FP64) {
    if (cmax < 32) launch_sweep_t<double, 32, 512>(B, nact, spmv, eta, s);
    else if (cmax < 64) launch_sweep_t<double, 64, 256>(B, nact, spmv, eta, s);
    else return set_error("batched basis wider than 64 columns in FP64"), 2;
  } else {
    if (cmax < 32) launch_sweep_t<float, 32, 512>(B, nact, spmv, eta, s);
    else if (cmax < 64) launch_sweep_t<float, 64, 512>(B, nact, spmv, eta, s);
    else return set_error("batched basis wider than 64 columns in FP32"), 2;
  }
  SA_HIP(hipGetLastError());
  return 0;
}

// The one-workgroup-per-problem sweep keeps acc[c] in registers up to 64 columns.
static bool persistent_fits(const BatchDev &B) { return B.m - 1 < 64; }

template <typename T, bool NT>
static void launch_bw_slot_nt(const BwArgs<T> &a, int lanes, int grid, size_t shm, hipStream_t s) {
  switch (lanes) {
    case 4: hipLaunchKernelGGL((bw_slot_kernel<T, 4, NT>), dim3(grid), dim3(BLK), shm, s, a); break;
    case 8: hipLaunchKernelGGL((bw_slot_kernel<T, 8, NT>), dim3(grid), dim3(BLK), shm, s, a); break;
    case 16: hipLaunchKernelGGL((bw_slot_kernel<T, 16, NT>), dim3(grid), dim3(BLK), shm, s, a); break;
    default: hipLaunchKernelGGL((bw_slot_kernel<T, 32, NT>), dim3(grid), dim3(BLK), shm, s, a); break;
  }
}

template <typename T>
static void launch_bw_slot(const BwArgs<T> &a, int lanes, bool nt, int grid, size_t shm, hipStream_t s) {
  if (nt) launch_bw_slot_nt<T, true>(a, lanes, grid, shm, s);
  else launch_bw_slot_nt<T, false>(a, lanes, grid, shm, s);
}

template <typename T>
static int wide_sweep_t(BatchDev &B, const Sched &sc, const std::vector<int> &list, const std::vector<int> &k0,
                        const std::vector<int> &nv, int nch, int spmv, double eta, hipStream_t s) {
  const int P = B.P, m = B.m, nl = (int)list.size();
  const bool nt = sc.nt != 0;
  int steps = 1;
  if (spmv) {
    steps = 0;
    for (int p : list) steps = std::max(steps, nv[p] - k0[p]);
  }
  // largest column of each step over the problems still sweeping
  std::vector<int> cmax(steps, 0);
  for (int j = 0; j < steps; j++) {
    for (int p : list) {
      const int c = k0[p] + (spmv ? 1 + j : 0);
      if (c <= nv[p]) cmax[j] = std::max(cmax[j], c);
    }
    SA_CHECK(gs_lds_bytes(cmax[j], sizeof(T)) + 4096 + 8 * BLK <= 65536,
             "basis too wide for the LDS tile of the Gram-Schmidt kernel");
  }
  double *pa = B.wpart, *pb = B.wpart + (size_t)B.nchunks * m;
  double *da = B.wdots, *db = B.wdots + (size_t)P * m;
  BwArgs<T> a;
  a.V = (T *)B.V; a.ldv = B.ldv; a.rp = B.rp; a.ci = B.ci; a.val = (const T *)B.val;
  a.off = B.off; a.n = B.n; a.k0 = B.k0; a.nv = B.nv;
  a.cprob = B.cprob; a.crow = B.crow; a.pc0 = B.pc0; a.pnc = B.pnc;
  a.spmv = spmv; a.m = m; a.st = B.wst; a.dbuf = B.dbuf; a.eta = eta;
  auto enqueue = [&]() {
    hipLaunchKernelGGL(bw_reset_kernel, dim3((nl + 255) / 256), dim3(256), 0, s, B.order, nl, B.wst);
    for (int j = 0; j < steps; j++) {
      const size_t shm = gs_lds_bytes(cmax[j], sizeof(T));
      a.j = j;
      for (int slot = 1; slot <= 3; slot++) {
        a.slot = slot;
        a.part_out = slot == 1 ? pa : pb;
        a.dots_in = slot == 2 ? da : db;
        launch_bw_slot<T>(a, sc.lanes, nt, nch, shm, s);
        if (slot < 3)
          hipLaunchKernelGGL(bw_sum_kernel, dim3(cmax[j] + 1, nl), dim3(BLK), 0, s, (const int *)B.order,
                             (const int *)B.k0, (const int *)B.nv, spmv, j, slot, (const GsState *)B.wst,
                             (const double *)a.part_out, (const int *)B.pc0, (const int *)B.pnc, m,
                             slot == 1 ? da : db);
      }
    }
    hipLaunchKernelGGL(bw_export_kernel, dim3((nl + 255) / 256), dim3(256), 0, s, (const int *)B.order, nl, B.wst,
                       B.stop, B.beta, B.passes);
  };
  if (!B.graphs_on) {
    enqueue();
    SA_HIP(hipGetLastError());
    return 0;
  }
  // replay the captured sweep of this launch signature, capturing it first
  std::vector<int> sig = {spmv, sc.nt, sc.lanes, nch, nl, steps};
  sig.insert(sig.end(), cmax.begin(), cmax.end());
  SweepGraph *g = nullptr;
  for (auto &e : B.graphs)
    if (e.sig == sig) { g = &e; break; }
  if (!g) {
    SA_HIP(hipGetLastError());
    SA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue();
    const hipError_t le = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (le != hipSuccess || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      SA_HIP(le);
      SA_HIP(ec);
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SA_HIP(ei);
    if (B.graphs.size() >= SWEEP_GRAPHS_MAX) {   // evict the least recently used
      auto lru = std::min_element(B.graphs.begin(), B.graphs.end(),
                                  [](const SweepGraph &x, const SweepGraph &y) { return x.used < y.used; });
      (void)hipGraphExecDestroy(lru->exec);
      B.graphs.erase(lru);
    }
    B.graphs.push_back(SweepGraph{sig, exec, 0});
    g = &B.graphs.back();
    B.graph_captures++;
  } else {
    B.graph_replays++;
  }
  g->used = ++B.graph_clock;
  SA_HIP(hipGraphLaunch(g->exec, s));
  return 0;
}

// Lock-step sweep of the listed problems; their k0/nv (host copies given) and
// the list itself (B.order) are already on the device.  Buffers are allocated
// on first use for the current basis width.
static int wide_sweep(BatchDev &B, const Sched &sc, const std::vector<int> &nrows, const std::vector<int> &list,
                      const std::vector<int> &k0, const std::vector<int> &nv, int spmv, double eta, hipStream_t s) {
  if (list.empty()) return 0;
  const int P = B.P, m = B.m;
  if (B.wm != m) {
    clear_graphs(B);   // the captured sweeps point at the buffers replaced here
    for (double **q : {&B.wpart, &B.wdots})
      if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (!B.wst) {
      SA_HIP(hipMalloc(&B.wst, sizeof(GsState) * GS_NSTATE * P));
      SA_HIP(hipMemsetAsync(B.wst, 0, sizeof(GsState) * GS_NSTATE * P, s));
      for (int **q : {&B.cprob, &B.crow}) SA_HIP(hipMalloc(q, sizeof(int) * std::max(1, B.nchunks)));
      for (int **q : {&B.pc0, &B.pnc}) SA_HIP(hipMalloc(q, sizeof(int) * P));
    }
    SA_HIP(hipMalloc(&B.wpart, sizeof(double) * 2 * (size_t)std::max(1, B.nchunks) * m));
    SA_HIP(hipMalloc(&B.wdots, sizeof(double) * 2 * (size_t)P * m));
    SA_HIP(hipMemsetAsync(B.wpart, 0, sizeof(double) * 2 * (size_t)std::max(1, B.nchunks) * m, s));
    SA_HIP(hipMemsetAsync(B.wdots, 0, sizeof(double) * 2 * (size_t)P * m, s));
    B.wm = m;
  }
  std::vector<int> cp, cr, pc0(P, 0), pnc(P, 0);
  for (int p : list) {
    pc0[p] = (int)cp.size();
    for (int r = 0; r < nrows[p]; r += CH) {
      cp.push_back(p);
      cr.push_back(r);
    }
    pnc[p] = (int)cp.size() - pc0[p];
  }
  const int nch = (int)cp.size();
  SA_CHECK(nch <= B.nchunks, "chunk list larger than the batch");
  SA_HIP(hipMemcpyAsync(B.cprob, cp.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.crow, cr.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.pc0, pc0.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.pnc, pnc.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  int rc = B.prec == SA_FP64 ? wide_sweep_t<double>(B, sc, list, k0, nv, nch, spmv, eta, s)
                             : wide_sweep_t<float>(B, sc, list, k0, nv, nch, spmv, eta, s);
  // the host vectors must outlive the asynchronous copies
  SA_HIP(hipStreamSynchronize(s));
  return rc;
}

// One sweep of the listed problems with schedule sc (kind 1: one workgroup per
// problem, 2: lock-step).
static int run_sweep(BatchDev &B, const Sched &sc, const std::vector<int> &nrows, const std::vector<int> &list,
                     const std::vector<int> &k0, const std::vector<int> &nv, int spmv, double eta, hipStream_t s) {
  if (sc.kind == 1) return launch_sweep(B, (int)list.size(), spmv, eta, s);
  return wide_sweep(B, sc, nrows, list, k0, nv, spmv, eta, s);
}

// The lock-step schedule without tuning: non-temporal tiles for a basis larger
// than the last-level cache, lanes per row from the operator's mean row length.
static Sched default_sched(const BatchDev &B) {
  Sched sc;
  sc.kind = 2;
  sc.nt = B.ldv * B.m * (int64_t)prec_size(B.prec) > GS_NT_BYTES ? 1 : 0;
  sc.lanes = B.lanes;
  return sc;
}

}  // namespace sa

using namespace sa;

struct sa_batch_s {
  int P = 0;
  int sweep_req = 0;                    // 0 autotune, 1 persistent, 2 lock-step
  int trueres = 0;                      // EPSSetTrueResidual for every problem
  int sweep_used = 0;                   // mode of the last solve's FP64 phase
  double ms_persistent = -1.0, ms_wide = -1.0;   // autotuner timings of one restart sweep
  int tuned_m[2] = {0, 0};              // basis width the schedule autotuner ran at, per precision (FP64, FP32)
  Sched tuned[2];                       // its choice
  Sched sched_used;                     // schedule of the last solve's last phase
  double tune_ms[4] = {-1, -1, -1, -1}; // candidate timings: persistent, lock-step, other NT, doubled lanes
  std::vector<int> pncv;                // per-problem ncv chosen by sa_batch_autotune_ncv (empty: none)
  std::vector<double> ncv_scores;       // its seconds per converged pair, [class][candidate]
  std::vector<int> ncv_cand, class_hi;  // its candidates and the largest n of each size class
  bool graphs = true;                   // capture the lock-step sweeps as HIP graphs
  std::vector<int> ncv_used;            // ncv of each problem in the last solve
  std::vector<int> n;
  std::vector<int64_t> off;
  int64_t nrows = 0, nnz = 0;
  // host CSR of the block-diagonal operator (FP64), kept for both precisions
  std::vector<int32_t> rp, ci;
  std::vector<double> val;
  BatchDev dev64, dev32;
  // results
  std::vector<std::vector<double>> eig, err;
  std::vector<std::vector<int>> perm;
  std::vector<int> nconv, its, reason;
  int64_t matvecs = 0, matvecs32 = 0, passes = 0, passes32 = 0;
  double seconds = 0, seconds32 = 0, seconds_ds = 0;
  int its_max = 0, its_max32 = 0, solved = 0;
};

static void free_dev(BatchDev &B) {
  clear_graphs(B);
  void *ptrs[] = {B.V,   B.rp,    B.ci,         B.val,        B.off,  B.n,    B.order, B.k0,  B.nv,
                  B.stop, B.passes, B.dbuf,     B.beta,       B.Q,    B.alpha, B.chunk_prob, B.chunk_row,
                  B.ml,  B.mk,    B.ms,         B.me,         B.src,  B.dst,  B.wst,   B.wpart, B.wdots,
                  B.cprob, B.crow, B.pc0,       B.pnc,        B.tx,   B.tz,   B.tlam,  B.tpart, B.tout,
                  B.tnv, B.tact, B.pchunk0, B.pnchunk, B.passes_save};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  B = BatchDev();
}

static int alloc_dev_impl(sa_batch_s *b, BatchDev &B, int prec, int m);

// Device state of one precision for a basis of m columns.  A failed
// allocation releases everything, so a later call starts from scratch
// instead of finding a basis of the right width with its operator missing.
static int alloc_dev(sa_batch_s *b, BatchDev &B, int prec, int m) {
  const int rc = alloc_dev_impl(b, B, prec, m);
  if (rc) free_dev(B);
  return rc;
}

static int alloc_dev_impl(sa_batch_s *b, BatchDev &B, int prec, int m) {
  if (B.V && B.m == m && B.prec == prec) return 0;
  const bool keep_op = B.rp && B.prec == prec;
  if (!keep_op) free_dev(B);
  else if (B.V) {
    clear_graphs(B);   // they point at the basis replaced here
    (void)hipFree(B.V);
    B.V = nullptr;
    if (B.tz) (void)hipFree(B.tz);   // sized by the basis width
    B.tz = nullptr;
  }
  hipStream_t s = stream();
  B.P = b->P;
  B.m = m;
  B.prec = prec;
  B.ldv = b->nrows;
  B.nrows = b->nrows;
  B.nnz = b->nnz;
  const size_t es = prec_size(prec);
  SA_HIP(hipMalloc(&B.V, es * (size_t)B.ldv * m));
  SA_HIP(hipMemsetAsync(B.V, 0, es * (size_t)B.ldv * m, s));
  if (keep_op) return 0;
  const int P = b->P;
  SA_HIP(hipMalloc(&B.rp, sizeof(int32_t) * (b->nrows + 1)));
  SA_HIP(hipMalloc(&B.ci, sizeof(int32_t) * std::max<int64_t>(1, b->nnz)));
  SA_HIP(hipMalloc(&B.val, es * std::max<int64_t>(1, b->nnz)));
  SA_HIP(hipMemcpyAsync(B.rp, b->rp.data(), sizeof(int32_t) * (b->nrows + 1), hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.ci, b->ci.data(), sizeof(int32_t) * b->nnz, hipMemcpyHostToDevice, s));
  if (prec == SA_FP64) {
    SA_HIP(hipMemcpyAsync(B.val, b->val.data(), sizeof(double) * b->nnz, hipMemcpyHostToDevice, s));
  } else {
    std::vector<float> f(b->val.begin(), b->val.end());
    SA_HIP(hipMemcpyAsync(B.val, f.data(), sizeof(float) * b->nnz, hipMemcpyHostToDevice, s));
    SA_HIP(hipStreamSynchronize(s));
  }
  const double avg = b->nrows ? (double)b->nnz / b->nrows : 1.0;
  B.lanes = avg <= 4.5 ? 4 : (avg <= 9 ? 8 : (avg <= 18 ? 16 : 32));
  SA_HIP(hipMalloc(&B.off, sizeof(int64_t) * P));
  SA_HIP(hipMalloc(&B.n, sizeof(int) * P));
  for (int **q : {&B.order, &B.k0, &B.nv, &B.stop, &B.passes, &B.ml, &B.mk, &B.ms, &B.me, &B.src, &B.dst,
                  &B.passes_save})
    SA_HIP(hipMalloc(q, sizeof(int) * P));
  SA_HIP(hipMalloc(&B.dbuf, sizeof(double) * (size_t)P * 128 * 128));
  SA_HIP(hipMalloc(&B.Q, sizeof(double) * (size_t)P * 128 * 128));
  SA_HIP(hipMalloc(&B.beta, sizeof(double) * P));
  SA_HIP(hipMalloc(&B.alpha, sizeof(double) * P));
  SA_HIP(hipMemcpyAsync(B.off, b->off.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.n, b->n.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemsetAsync(B.passes, 0, sizeof(int) * P, s));
  std::vector<int> cp, cr;
  for (int p = 0; p < P; p++)
    for (int r = 0; r < b->n[p]; r += 64) {
      cp.push_back(p);
      cr.push_back(r);
    }
  B.nchunks = (int)cp.size();
  SA_HIP(hipMalloc(&B.chunk_prob, sizeof(int) * std::max(1, B.nchunks)));
  SA_HIP(hipMalloc(&B.chunk_row, sizeof(int) * std::max(1, B.nchunks)));
  SA_HIP(hipMemcpyAsync(B.chunk_prob, cp.data(), sizeof(int) * B.nchunks, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.chunk_row, cr.data(), sizeof(int) * B.nchunks, hipMemcpyHostToDevice, s));
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

// Buffers of the true-residual rounds (allocated on the first solve that asks for them).
static int ensure_tres(sa_batch_s *b, BatchDev &B) {
  const int P = b->P, m = B.m;
  hipStream_t s = stream();
  if (!B.tx) {
    SA_HIP(hipMalloc(&B.tx, prec_size(B.prec) * (size_t)B.ldv * TR_SLOTS));
    SA_HIP(hipMalloc(&B.tlam, sizeof(double) * (size_t)P * TR_SLOTS));
    SA_HIP(hipMalloc(&B.tout, sizeof(double) * (size_t)P * TR_SLOTS));
    SA_HIP(hipMalloc(&B.tpart, sizeof(double) * (size_t)TR_SLOTS * std::max(1, B.nchunks)));
    SA_HIP(hipMalloc(&B.tnv, sizeof(int) * P));
    SA_HIP(hipMalloc(&B.tact, sizeof(int) * (size_t)P * TR_SLOTS));
    SA_HIP(hipMalloc(&B.pchunk0, sizeof(int) * P));
    SA_HIP(hipMalloc(&B.pnchunk, sizeof(int) * P));
    std::vector<int> c0(P), nc(P);
    int acc = 0;
    for (int q = 0; q < P; q++) {    // chunks are numbered problem by problem
      c0[q] = acc;
      nc[q] = (b->n[q] + 63) / 64;
      acc += nc[q];
    }
    SA_CHECK(acc == B.nchunks, "internal: chunk numbering");
    SA_HIP(hipMemcpyAsync(B.pchunk0, c0.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.pnchunk, nc.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipStreamSynchronize(s));
  }
  if (!B.tz) SA_HIP(hipMalloc(&B.tz, sizeof(double) * (size_t)P * TR_SLOTS * m));
  return 0;
}

// One round: the candidates in Z/lam/act/nv (host, P x TR_SLOTS) -> squared
// residual norms in out (host, P x TR_SLOTS).
template <typename T>
static void tres_launch(const BatchDev &B, hipStream_t s) {
  const dim3 grid((unsigned)B.nchunks, TR_SLOTS);
  hipLaunchKernelGGL(batched_ritz_kernel<T>, grid, dim3(64), 0, s, (const T *)B.V, B.ldv, B.chunk_prob, B.chunk_row,
                     B.off, B.n, B.tnv, B.tact, B.tz, B.m, (T *)B.tx);
  const T *val = (const T *)B.val;
  switch (B.lanes) {
    case 4: hipLaunchKernelGGL((batched_tres_kernel<T, 4>), grid, dim3(256), 0, s, B.rp, B.ci, val, (const T *)B.tx, B.ldv, B.chunk_prob, B.chunk_row, B.off, B.n, B.tact, B.tlam, B.tpart, B.nchunks); break;
    case 8: hipLaunchKernelGGL((batched_tres_kernel<T, 8>), grid, dim3(256), 0, s, B.rp, B.ci, val, (const T *)B.tx, B.ldv, B.chunk_prob, B.chunk_row, B.off, B.n, B.tact, B.tlam, B.tpart, B.nchunks); break;
    case 16: hipLaunchKernelGGL((batched_tres_kernel<T, 16>), grid, dim3(256), 0, s, B.rp, B.ci, val, (const T *)B.tx, B.ldv, B.chunk_prob, B.chunk_row, B.off, B.n, B.tact, B.tlam, B.tpart, B.nchunks); break;
    default: hipLaunchKernelGGL((batched_tres_kernel<T, 32>), grid, dim3(256), 0, s, B.rp, B.ci, val, (const T *)B.tx, B.ldv, B.chunk_prob, B.chunk_row, B.off, B.n, B.tact, B.tlam, B.tpart, B.nchunks); break;
  }
  hipLaunchKernelGGL(batched_tres_sum_kernel, dim3((unsigned)B.P, TR_SLOTS), dim3(64), 0, s, B.tpart, B.nchunks,
                     B.pchunk0, B.pnchunk, B.tact, B.tout);
}

static int tres_round(BatchDev &B, const std::vector<double> &Z, const std::vector<double> &lam,
                      const std::vector<int> &act, const std::vector<int> &nv, std::vector<double> &out,
                      hipStream_t s) {
  const int P = B.P;
  SA_HIP(hipMemcpyAsync(B.tz, Z.data(), sizeof(double) * (size_t)P * TR_SLOTS * B.m, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.tlam, lam.data(), sizeof(double) * (size_t)P * TR_SLOTS, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.tact, act.data(), sizeof(int) * (size_t)P * TR_SLOTS, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.tnv, nv.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  if (B.prec == SA_FP64) tres_launch<double>(B, s);
  else tres_launch<float>(B, s);
  SA_HIP(hipGetLastError());
  out.resize((size_t)P * TR_SLOTS);
  SA_HIP(hipMemcpyAsync(out.data(), B.tout, sizeof(double) * (size_t)P * TR_SLOTS, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

namespace {
struct ProbState {
  DenseProblem ds;
  std::vector<double> eigr, eigi, errest;
  int nconv = 0, l = 0, its = 0, reason = 0, k = 0, nv = 0, k0 = 0;
  int tnext = 0;           // true residuals: next candidate index, or -1 once decided
  int breakdown = 0;
  bool active = true, restart_random = false;
  double beta = 0.0;
  int stop = -1;
};

template <typename F>
void parallel_for(int n, F f) {
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  if (n < 8 || nt == 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int w = 0; w < nt; w++)
    th.emplace_back([&, w]() {
      for (int i = w; i < n; i += nt) f(i);
    });
  for (auto &x : th) x.join();
}
double nowb() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
}  // namespace

// One batched Krylov-Schur solve (symmetric problems, Lanczos + DSHEP on each)
// in basis precision prec; starts from column 0 of V.  Mirrors
// EPSSolve_KrylovSchur_Default (krylovschur.c:227-348) per problem.
// pncv: basis size of each problem (B.m = the largest + 1); subset: when
// given, only the problems marked nonzero are solved (the others stay
// inactive with no converged pairs); pbytes: when given, receives each
// problem's algorithmic HBM bytes (two Gram-Schmidt passes per step, the
// SpMV, the restart GEMM and the true-residual checks).
static int batch_ks(sa_batch_s *b, BatchDev &B, int which, int nev, const std::vector<int> &pncv, double tol,
                    int max_it, bool trueres, std::vector<ProbState> &ps, int64_t &matvecs, int64_t &passes_out,
                    int &its_max, double &tds, const std::vector<char> *subset = nullptr,
                    std::vector<double> *pbytes = nullptr) {
  const int P = b->P, m = B.m, ld = m;
  const double eta = 0.7071, keep = 0.5;
  B.graphs_on = b->graphs;
  hipStream_t s = stream();
  SA_CHECK(lapack_ready(), "no LAPACK library found for the projected problem");
  if (trueres) SA_TRY(ensure_tres(b, B));
  std::vector<double> tZ, tlam, tout;
  std::vector<int> tact, tnvv;
  ps.assign(P, ProbState());
  for (int p = 0; p < P; p++) {
    ps[p].ds.allocate(ld, true, which);
    ps[p].eigr.assign(ld, 0.0);
    ps[p].eigi.assign(ld, 0.0);
    ps[p].errest.assign(ld, 0.0);
  }
  std::vector<int> order, k0(P), nv(P), stop(P), passes0(P), passes1(P), zero(P, 0);
  std::vector<double> beta(P), dbuf, Qh((size_t)P * m * m, 0.0);
  std::vector<int> ml(P), mk(P), msv(P), mev(P), src(P), dst(P);
  // problems by decreasing size: the scheduler hands the largest sweeps out first
  for (int p = 0; p < P; p++) {
    if (subset && !(*subset)[p]) ps[p].active = false;
    else order.push_back(p);
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->n[x] > b->n[y]; });
  SA_CHECK(!order.empty(), "no problem to solve");
  SA_HIP(hipMemcpyAsync(B.passes, zero.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  // orthonormalise the start vectors (column 0): EPSGetStartVector
  for (int p = 0; p < P; p++) { k0[p] = 0; nv[p] = 0; stop[p] = -1; }
  SA_HIP(hipMemcpyAsync(B.order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.k0, k0.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.nv, nv.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  SA_HIP(hipMemcpyAsync(B.stop, stop.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  // sweep schedule: requested, tuned by an earlier solve of this precision
  // and basis width, or decided by timing the candidates on the first
  // restart round (autotuner)
  const int pi = B.prec == SA_FP64 ? 0 : 1;
  Sched sc = default_sched(B);
  bool tune = false;
  if (b->sweep_req == 1) {
    SA_CHECK(persistent_fits(B), "the one-workgroup-per-problem sweep handles ncv < 64");
    sc.kind = 1;
  } else if (b->sweep_req == 0) {
    if (b->tuned_m[pi] == B.m) sc = b->tuned[pi];
    else tune = true;
  }
  SA_TRY(run_sweep(B, sc, b->n, order, k0, nv, 0, eta, s));
  SA_HIP(hipMemcpyAsync(stop.data(), B.stop, sizeof(int) * P, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  for (int p : order) SA_CHECK(stop[p] < 0, "a start vector is zero");
  its_max = 0;
  matvecs = 0;
  const double es = (double)prec_size(B.prec);
  auto spmv_bytes_p = [&](int p) {
    const double nz = (double)(b->rp[b->off[p] + b->n[p]] - b->rp[b->off[p]]);
    return nz * (es + 4.0) + 4.0 * (b->n[p] + 1) + 2.0 * es * b->n[p];
  };
  if (pbytes) pbytes->assign(P, 0.0);
  while (true) {
    // ---- set up the sweep of every active problem
    std::vector<int> act;
    for (size_t i = 0; i < order.size(); i++) {
      const int p = order[i];
      ProbState &S = ps[p];
      if (!S.active) continue;
      S.its++;
      S.nv = std::min(S.nconv + pncv[p], pncv[p]);
      S.k0 = S.nconv + S.l;
      S.ds.set_dimensions(S.nv, S.nconv, S.k0);
      act.push_back(p);
    }
    if (act.empty()) break;
    for (int p = 0; p < P; p++) {
      k0[p] = ps[p].k0;
      nv[p] = ps[p].active ? ps[p].nv : ps[p].k0;
      stop[p] = -1;
    }
    SA_HIP(hipMemcpyAsync(B.order, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.k0, k0.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.nv, nv.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.stop, stop.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    if (tune) {
      // autotune: the same sweep (it reads only columns <= k0 and rewrites
      // the later ones) with each candidate schedule, timed with events; the
      // fastest is run once more last so that the round's results are its
      // own, and the pass counters are restored before that run
      // candidates: 0 persistent, 1 lock-step, 2 lock-step with the other
      // tile-load kind, 3 lock-step with twice the SpMV lanes per row
      std::vector<Sched> cand;
      std::vector<int> cslot;
      if (persistent_fits(B)) { Sched c = sc; c.kind = 1; cand.push_back(c); cslot.push_back(0); }
      cand.push_back(default_sched(B));
      cslot.push_back(1);
      { Sched c = default_sched(B); c.nt = 1 - c.nt; cand.push_back(c); cslot.push_back(2); }
      if (B.lanes < 32) { Sched c = default_sched(B); c.lanes = 2 * B.lanes; cand.push_back(c); cslot.push_back(3); }
      SA_HIP(hipMemcpyAsync(B.passes_save, B.passes, sizeof(int) * P, hipMemcpyDeviceToDevice, s));
      std::vector<float> ms(cand.size(), 0.f);
      B.graphs_on = false;   // time the kernels, not the first capture of a graph
      hipEvent_t ev[2];
      for (auto &e : ev) SA_HIP(hipEventCreate(&e));
      for (size_t c = 0; c < cand.size(); c++) {
        SA_HIP(hipMemcpyAsync(B.stop, stop.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
        SA_HIP(hipEventRecord(ev[0], s));
        SA_TRY(run_sweep(B, cand[c], b->n, act, k0, nv, 1, eta, s));
        SA_HIP(hipEventRecord(ev[1], s));
        SA_HIP(hipEventSynchronize(ev[1]));
        SA_HIP(hipEventElapsedTime(&ms[c], ev[0], ev[1]));
      }
      for (auto &e : ev) (void)hipEventDestroy(e);
      const size_t best = std::min_element(ms.begin(), ms.end()) - ms.begin();
      for (int q = 0; q < 4; q++) b->tune_ms[q] = -1.0;
      for (size_t c = 0; c < cand.size(); c++) b->tune_ms[cslot[c]] = ms[c];
      b->ms_persistent = b->tune_ms[0];
      b->ms_wide = b->tune_ms[1];
      sc = cand[best];
      B.graphs_on = b->graphs;
      SA_HIP(hipMemcpyAsync(B.passes, B.passes_save, sizeof(int) * P, hipMemcpyDeviceToDevice, s));
      SA_HIP(hipMemcpyAsync(B.stop, stop.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
      SA_TRY(run_sweep(B, sc, b->n, act, k0, nv, 1, eta, s));
      b->tuned_m[pi] = B.m;
      b->tuned[pi] = sc;
      tune = false;
    } else {
      SA_TRY(run_sweep(B, sc, b->n, act, k0, nv, 1, eta, s));
    }
    b->sweep_used = sc.kind;
    b->sched_used = sc;
    // coefficients of the problems from the first to the last active one
    const int plo = *std::min_element(act.begin(), act.end()), phi = *std::max_element(act.begin(), act.end());
    dbuf.resize((size_t)P * m * m);
    SA_HIP(hipMemcpyAsync(dbuf.data() + (size_t)plo * m * m, B.dbuf + (size_t)plo * m * m,
                          sizeof(double) * (size_t)(phi - plo + 1) * m * m, hipMemcpyDeviceToHost, s));
    SA_HIP(hipMemcpyAsync(beta.data(), B.beta, sizeof(double) * P, hipMemcpyDeviceToHost, s));
    SA_HIP(hipMemcpyAsync(stop.data(), B.stop, sizeof(int) * P, hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    const double t0 = nowb();
    its_max++;
    // ---- host: projected problems of all active problems in parallel
    std::vector<int> fail(act.size(), 0);
    parallel_for((int)act.size(), [&](int ia) {
      const int p = act[ia];
      ProbState &S = ps[p];
      int nvp = S.nv;
      const int breakdown = stop[p] >= 0;
      if (breakdown) nvp = stop[p];
      const double bet = beta[p];
      const double *Bp = dbuf.data() + (size_t)p * m * m;
      double *d = S.ds.d(), *e = S.ds.e();
      for (int j = S.k0; j < nvp; j++) {
        d[j] = Bp[(size_t)(j + 1) * m + j];
        e[j] = Bp[(size_t)(j + 1) * m + j + 1];
      }
      S.nv = nvp;
      S.ds.set_dimensions(nvp, S.nconv, S.k0);
      S.ds.set_state(S.l ? DS_RAW : DS_INTERMEDIATE);
      if (S.ds.solve(S.eigr.data(), S.eigi.data())) { fail[ia] = 1; return; }
      S.ds.sort(S.eigr.data(), S.eigi.data());
      S.ds.update_extra_row();
      S.breakdown = breakdown;
      if (trueres) {          // decided by the device rounds below
        S.tnext = S.nconv;
        return;
      }
      int k, marker = -1;
      for (k = S.nconv; k < nvp; k++) {
        double rn = 0.0;
        int cp = 0;
        S.ds.ritz_residual(k, &rn, &cp);
        rn *= bet;
        const double w = std::fabs(S.eigr[k]);
        S.errest[k] = w != 0.0 ? rn / w : 1.79e308;
        if (marker == -1 && S.errest[k] >= tol) marker = k;
        if (marker != -1) break;
      }
      if (marker != -1) k = marker;
      S.k = k;
    });
    for (size_t ia = 0; ia < act.size(); ia++) SA_CHECK(!fail[ia], "LAPACK failure in a batched projected problem");
    if (trueres) {
      // EPSKrylovConvergence with true residuals (epskrylov.c:225-265): per
      // problem, the candidates from nconv on in order until the first whose
      // explicit residual fails; a round evaluates up to TR_SLOTS candidates
      // of every undecided problem (those whose estimate passes, and the
      // first that fails, which usually ends the walk).
      tZ.assign((size_t)P * TR_SLOTS * m, 0.0);
      tlam.assign((size_t)P * TR_SLOTS, 0.0);
      tact.assign((size_t)P * TR_SLOTS, 0);
      tnvv.assign(P, 0);
      while (true) {
        bool any = false;
        std::fill(tact.begin(), tact.end(), 0);
        for (int p : act) {
          ProbState &S = ps[p];
          if (S.tnext < 0) continue;
          if (S.tnext >= S.nv) {        // every candidate passed
            S.k = S.nv;
            S.tnext = -1;
            continue;
          }
          tnvv[p] = S.nv;
          for (int sl = 0; sl < TR_SLOTS && S.tnext + sl < S.nv; sl++) {
            const int k = S.tnext + sl;
            const double *Qk = S.ds.Q() + (size_t)k * ld;
            std::memcpy(tZ.data() + ((size_t)p * TR_SLOTS + sl) * m, Qk, sizeof(double) * S.nv);
            tlam[(size_t)p * TR_SLOTS + sl] = std::fabs(S.eigr[k]) > 2.220446049250313e-16 ? S.eigr[k] : 0.0;
            tact[(size_t)p * TR_SLOTS + sl] = 1;
            any = true;
            double rn = 0.0;
            int cp = 0;
            S.ds.ritz_residual(k, &rn, &cp);
            const double w = std::fabs(S.eigr[k]);
            if ((w != 0.0 ? rn * beta[p] / w : 1.79e308) >= tol) break;   // the estimate fails: likely the end
          }
        }
        if (!any) break;
        SA_TRY(tres_round(B, tZ, tlam, tact, tnvv, tout, s));
        for (int p : act) {
          ProbState &S = ps[p];
          if (S.tnext < 0) continue;
          int sl = 0;
          for (; sl < TR_SLOTS && tact[(size_t)p * TR_SLOTS + sl]; sl++) {
            const int k = S.tnext + sl;
            const double w = std::fabs(S.eigr[k]);
            S.errest[k] = w != 0.0 ? std::sqrt(tout[(size_t)p * TR_SLOTS + sl]) / w : 1.79e308;
            if (S.errest[k] >= tol) {
              S.k = k;
              S.tnext = -1;
              break;
            }
          }
          if (S.tnext >= 0) S.tnext += sl;
        }
      }
    }
    parallel_for((int)act.size(), [&](int ia) {
      const int p = act[ia];
      ProbState &S = ps[p];
      const int nvp = S.nv, breakdown = S.breakdown;
      const int k = S.k;
      if (k >= nev) S.reason = 1;
      else if (S.its >= max_it) S.reason = -1;
      int newl = 0;
      if (!(S.reason != 0 || breakdown || k == nvp)) newl = std::max(1, (int)((nvp - k) * keep));
      S.restart_random = false;
      if (S.reason == 0) {
        if (breakdown || k == nvp) {
          if (k < nev) S.restart_random = true;
        } else {
          S.ds.truncate(k + newl);
        }
      }
      // restart GEMM parameters: V(:,nconv:k+l) = V(:,nconv:nv) Q
      std::memcpy(Qh.data() + (size_t)p * m * m, S.ds.Q(), sizeof(double) * m * m);
      ml[p] = S.nconv;
      mk[p] = nvp;
      msv[p] = S.nconv;
      mev[p] = k + newl;
      src[p] = (S.reason == 0 && !breakdown) ? nvp : -1;
      dst[p] = k + newl;
      S.k = k;
      S.l = newl;
    });
    for (size_t ia = 0; ia < act.size(); ia++) SA_CHECK(!fail[ia], "LAPACK failure in a batched projected problem");
    for (int p = 0; p < P; p++)
      if (!ps[p].active) { ml[p] = mk[p] = msv[p] = mev[p] = 0; src[p] = -1; dst[p] = 0; }
    for (int p : act) matvecs += ps[p].nv - ps[p].k0;
    if (pbytes)
      for (int p : act) {
        const ProbState &S = ps[p];
        double bb = 0.0;
        for (int c = S.k0 + 1; c <= S.nv; c++) bb += 2.0 * (2.0 * c + 2.0) * es * b->n[p] + spmv_bytes_p(p);
        bb += ((double)(S.nv - S.nconv) + (double)(S.k + S.l - S.nconv)) * es * b->n[p];
        if (trueres) bb += (double)(S.k - S.nconv + 1) * (((double)S.nv + 1.0) * es * b->n[p] + spmv_bytes_p(p));
        (*pbytes)[p] += bb;
      }
    tds += nowb() - t0;
    SA_HIP(hipMemcpyAsync(B.Q + (size_t)plo * m * m, Qh.data() + (size_t)plo * m * m,
                          sizeof(double) * (size_t)(phi - plo + 1) * m * m, hipMemcpyHostToDevice, s));
    for (auto pr : {std::make_pair(B.ml, &ml), std::make_pair(B.mk, &mk), std::make_pair(B.ms, &msv),
                    std::make_pair(B.me, &mev), std::make_pair(B.src, &src), std::make_pair(B.dst, &dst)})
      SA_HIP(hipMemcpyAsync(pr.first, pr.second->data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    if (B.prec == SA_FP64) {
      hipLaunchKernelGGL(batched_mip_kernel<double>, dim3(B.nchunks), dim3(256), 0, s, (double *)B.V, B.ldv,
                         B.chunk_prob, B.chunk_row, B.off, B.n, B.ml, B.mk, B.ms, B.me, B.Q, m);
      hipLaunchKernelGGL(batched_copy_kernel<double>, dim3(B.nchunks), dim3(64), 0, s, (double *)B.V, B.ldv,
                         B.chunk_prob, B.chunk_row, B.off, B.n, B.src, B.dst);
    } else {
      hipLaunchKernelGGL(batched_mip_kernel<float>, dim3(B.nchunks), dim3(256), 0, s, (float *)B.V, B.ldv,
                         B.chunk_prob, B.chunk_row, B.off, B.n, B.ml, B.mk, B.ms, B.me, B.Q, m);
      hipLaunchKernelGGL(batched_copy_kernel<float>, dim3(B.nchunks), dim3(64), 0, s, (float *)B.V, B.ldv,
                         B.chunk_prob, B.chunk_row, B.off, B.n, B.src, B.dst);
    }
    SA_HIP(hipGetLastError());
    // breakdown restarts (EPSGetStartVector(k)): random column, orthonormalised
    std::vector<int> rs;
    for (int p : act)
      if (ps[p].restart_random) rs.push_back(p);
    if (!rs.empty()) {
      SA_HIP(hipStreamSynchronize(s));
      for (int p : rs) {
        std::vector<double> x(b->n[p]);
        sa_rand_uniform(0x5EEDULL + (uint64_t)ps[p].k, b->n[p], x.data());
        const size_t es = prec_size(B.prec);
        char *dstp = (char *)B.V + es * ((size_t)ps[p].k * B.ldv + b->off[p]);
        if (B.prec == SA_FP64) {
          SA_HIP(hipMemcpy(dstp, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice));
        } else {
          std::vector<float> f(x.begin(), x.end());
          SA_HIP(hipMemcpy(dstp, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
        }
        k0[p] = ps[p].k;
      }
      std::vector<int> st2(P, 0);
      for (int p = 0; p < P; p++) st2[p] = -1;
      SA_HIP(hipMemcpyAsync(B.order, rs.data(), sizeof(int) * rs.size(), hipMemcpyHostToDevice, s));
      SA_HIP(hipMemcpyAsync(B.k0, k0.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
      SA_HIP(hipMemcpyAsync(B.stop, st2.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
      SA_TRY(run_sweep(B, sc, b->n, rs, k0, nv, 0, eta, s));
      SA_HIP(hipMemcpyAsync(st2.data(), B.stop, sizeof(int) * P, hipMemcpyDeviceToHost, s));
      SA_HIP(hipStreamSynchronize(s));
      for (int p : rs)
        if (st2[p] >= 0) ps[p].reason = -2;
    }
    for (int p : act) {
      ProbState &S = ps[p];
      S.nconv = S.k;
      if (S.reason != 0) S.active = false;
    }
  }
  std::vector<int> pz(P);
  SA_HIP(hipMemcpyAsync(pz.data(), B.passes, sizeof(int) * P, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  passes_out = 0;
  for (int p = 0; p < P; p++) passes_out += pz[p];
  return 0;
}

static void sort_perm(int which, const std::vector<double> &eigr, int nconv, std::vector<int> &perm) {
  perm.resize(nconv);
  std::iota(perm.begin(), perm.end(), 0);
  // SlepcSortEigenvalues for real eigenvalues (slepcsc.c:89-140)
  for (int i = nconv - 1; i >= 0; i--) {
    const double re = eigr[perm[i]];
    int j = i + 1;
    while (j < nconv) {
      if (sc_compare(which, re, 0.0, eigr[perm[j]], 0.0) <= 0) break;
      std::swap(perm[j - 1], perm[j]);
      j++;
    }
  }
}

extern "C" {

int sa_batch_create(int P, const int32_t *n, const int32_t *rowptr, const int32_t *col, const double *val,
                    sa_batch *out) {
  SA_TRY(ensure_init());
  SA_CHECK(P > 0 && n && rowptr && col && val && out, "invalid batch arguments");
  // every problem's CSR is checked as sa_mat_create_csr checks one operator,
  // before anything is allocated
  int64_t rows = 0, nnz = 0, rp_off = 0;
  for (int p = 0; p < P; p++) {
    SA_CHECK(n[p] > 0, "empty problem in batch");
    const int32_t *rpp = rowptr + rp_off;
    SA_CHECK(rpp[0] == 0, "row pointer of a batch problem must start at 0");
    for (int r = 0; r < n[p]; r++) {
      SA_CHECK(rpp[r + 1] >= rpp[r], "row pointer not monotone in batch");
      for (int32_t q = rpp[r]; q < rpp[r + 1]; q++) {
        SA_CHECK(col[nnz + q] >= 0 && col[nnz + q] < n[p], "column index out of range in batch");
        SA_CHECK(q == rpp[r] || col[nnz + q] > col[nnz + q - 1], "column indices must ascend within a row");
      }
    }
    rows += (n[p] + 63) / 64 * 64;
    nnz += rpp[n[p]];
    rp_off += n[p] + 1;
  }
  SA_CHECK(nnz < INT32_MAX && rows < INT32_MAX, "batch too large for int32 CSR indices");
  auto *b = new sa_batch_s();
  b->P = P;
  if (const char *e = std::getenv("SLEPC_AMD_BATCH_SWEEP"))
    b->sweep_req = !std::strcmp(e, "persistent") ? 1 : (!std::strcmp(e, "wide") ? 2 : 0);
  b->n.assign(n, n + P);
  b->off.resize(P);
  int64_t r0 = 0;
  for (int p = 0; p < P; p++) {   // problems start at 64-row boundaries
    b->off[p] = r0;
    r0 += (n[p] + 63) / 64 * 64;
  }
  b->nrows = rows;
  b->nnz = nnz;
  b->rp.assign(rows + 1, 0);
  b->ci.resize(nnz);
  b->val.resize(nnz);
  int64_t z = 0, vo = 0;
  rp_off = 0;
  for (int p = 0; p < P; p++) {
    const int32_t *rpp = rowptr + rp_off;
    const int64_t base = b->off[p];
    for (int r = 0; r < n[p]; r++) {
      b->rp[base + r] = (int32_t)z;
      for (int32_t q = rpp[r]; q < rpp[r + 1]; q++) {
        b->ci[z] = (int32_t)(base + col[vo + q]);
        b->val[z] = val[vo + q];
        z++;
      }
    }
    const int64_t padded_end = (p + 1 < P) ? b->off[p + 1] : rows;
    for (int64_t r = base + n[p]; r < padded_end; r++) b->rp[r] = (int32_t)z;
    vo += rpp[n[p]];
    rp_off += n[p] + 1;
  }
  b->rp[rows] = (int32_t)z;
  *out = b;
  return 0;
}

int sa_batch_destroy(sa_batch b) {
  if (!b) return 0;
  free_dev(b->dev64);
  free_dev(b->dev32);
  delete b;
  return 0;
}

// Solves every problem of the batch: nev eigenpairs of `which`, ncv basis
// vectors, tolerance tol.  mixed != 0: FP32 bulk to tol32, then FP64 polish from
// the sum of the FP32 Ritz vectors.  v0 (optional): concatenated start vectors
// (problem p at offset sum_{q<p} n_q); NULL draws U[-1,1) with seed
// sa_sub_seed(seed, p).
int sa_batch_solve(sa_batch b, int which, int nev, int ncv, double tol, int max_it, int mixed, double tol32,
                   const double *v0, uint64_t seed) {
  SA_TRY(ensure_init());
  SA_CHECK(b && nev > 0 && ncv >= 0 && tol > 0, "invalid batch solve arguments");
  const int P = b->P;
  // basis size of each problem: ncv, or (ncv == 0) the autotuner's choice,
  // or EPS's default max(2 nev, nev + 15) capped at n
  std::vector<int> pncv(P);
  for (int p = 0; p < P; p++) {
    pncv[p] = ncv > 0 ? ncv : (!b->pncv.empty() ? b->pncv[p] : std::min(b->n[p], std::max(2 * nev, nev + 15)));
    SA_CHECK(pncv[p] > nev && pncv[p] <= 64 && pncv[p] <= b->n[p], "invalid ncv: need nev < ncv <= min(64, n)");
  }
  b->ncv_used = pncv;
  const int m = *std::max_element(pncv.begin(), pncv.end()) + 1;
  hipStream_t s = stream();
  if (max_it <= 0) max_it = 100000;
  const double t0 = nowb();
  // start vectors into column 0 of the working basis
  std::vector<double> x0((size_t)b->nrows, 0.0);
  {
    int64_t vo = 0;
    for (int p = 0; p < P; p++) {
      if (v0) std::memcpy(x0.data() + b->off[p], v0 + vo, sizeof(double) * b->n[p]);
      else sa_rand_uniform(sa_sub_seed(seed, p), b->n[p], x0.data() + b->off[p]);
      vo += b->n[p];
    }
  }
  SA_TRY(alloc_dev(b, b->dev64, SA_FP64, m));
  std::vector<ProbState> ps;
  b->matvecs32 = b->passes32 = 0;
  b->its_max32 = 0;
  b->seconds32 = 0;
  double tds = 0.0;
  if (mixed) {
    SA_TRY(alloc_dev(b, b->dev32, SA_FP32, m));
    std::vector<float> f(x0.begin(), x0.end());
    SA_HIP(hipMemcpyAsync(b->dev32.V, f.data(), sizeof(float) * b->nrows, hipMemcpyHostToDevice, s));
    SA_HIP(hipStreamSynchronize(s));
    const double t1 = nowb();
    SA_TRY(batch_ks(b, b->dev32, which, nev, pncv, tol32, max_it, b->trueres != 0, ps, b->matvecs32, b->passes32,
                    b->its_max32, tds));
    // polish seeds: sum of the first nev Ritz vectors of each problem, in FP64
    SA_TRY(launch_convert(SA_FP32, b->dev32.V, b->dev32.ldv, SA_FP64, b->dev64.V, b->dev64.ldv, b->nrows, m, s));
    std::vector<double> Qh((size_t)P * m * m, 0.0);
    std::vector<int> ml(P, 0), mk(P, 0), msv(P, 0), mev(P, 0);
    for (int p = 0; p < P; p++) {
      std::vector<int> perm;
      sort_perm(which, ps[p].eigr, ps[p].nconv, perm);
      const int nr = std::min(nev, ps[p].nconv);
      for (int i = 0; i < nr; i++) Qh[(size_t)p * m * m + perm[i]] = 1.0;   // column 0 of Q
      if (nr == 0) Qh[(size_t)p * m * m] = 1.0;                             // keep the start vector
      ml[p] = 0; mk[p] = std::max(1, ps[p].nconv); msv[p] = 0; mev[p] = 1;
    }
    BatchDev &B = b->dev64;
    SA_HIP(hipMemcpyAsync(B.Q, Qh.data(), sizeof(double) * (size_t)P * m * m, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.ml, ml.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.mk, mk.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.ms, msv.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(B.me, mev.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(batched_mip_kernel<double>, dim3(B.nchunks), dim3(256), 0, s, (double *)B.V, B.ldv,
                       B.chunk_prob, B.chunk_row, B.off, B.n, B.ml, B.mk, B.ms, B.me, B.Q, m);
    SA_HIP(hipGetLastError());
    SA_HIP(hipStreamSynchronize(s));
    b->seconds32 = nowb() - t1;
  } else {
    SA_HIP(hipMemcpyAsync(b->dev64.V, x0.data(), sizeof(double) * b->nrows, hipMemcpyHostToDevice, s));
    SA_HIP(hipStreamSynchronize(s));
  }
  SA_TRY(batch_ks(b, b->dev64, which, nev, pncv, tol, max_it, b->trueres != 0, ps, b->matvecs, b->passes, b->its_max,
                  tds));
  b->nconv.resize(P);
  b->its.resize(P);
  b->reason.resize(P);
  b->eig.resize(P);
  b->err.resize(P);
  b->perm.resize(P);
  for (int p = 0; p < P; p++) {
    b->nconv[p] = ps[p].nconv;
    b->its[p] = ps[p].its;
    b->reason[p] = ps[p].reason;
    b->eig[p] = ps[p].eigr;
    b->err[p] = ps[p].errest;
    sort_perm(which, ps[p].eigr, ps[p].nconv, b->perm[p]);
  }
  b->seconds = nowb() - t0;
  b->seconds_ds = tds;
  b->solved = 1;
  return 0;
}

int sa_batch_get_converged(sa_batch b, int p, int *nconv, int *its, int *reason) {
  SA_CHECK(b && b->solved && p >= 0 && p < b->P, "invalid problem index");
  if (nconv) *nconv = b->nconv[p];
  if (its) *its = b->its[p];
  if (reason) *reason = b->reason[p];
  return 0;
}

int sa_batch_get_eigenpair(sa_batch b, int p, int i, double *re, double *x_host) {
  SA_CHECK(b && b->solved && p >= 0 && p < b->P && i >= 0 && i < b->nconv[p], "invalid eigenpair index");
  const int k = b->perm[p][i];
  if (re) *re = b->eig[p][k];
  if (x_host) {
    const BatchDev &B = b->dev64;
    SA_HIP(hipMemcpy(x_host, (double *)B.V + (size_t)k * B.ldv + b->off[p], sizeof(double) * b->n[p],
                     hipMemcpyDeviceToHost));
  }
  return 0;
}

// Relative residual ||A x - lambda x|| / |lambda| of every requested pair on
// the device in FP64 (EPSComputeError, epssolve.c:742-815); err[p*nev + i].
int sa_batch_compute_errors(sa_batch b, int nev, double *err) {
  SA_TRY(ensure_init());
  SA_CHECK(b && b->solved && err, "batch not solved");
  const BatchDev &B = b->dev64;
  hipStream_t s = stream();
  double *y = nullptr, *part = nullptr, *out = nullptr;
  SA_HIP(hipMalloc(&y, sizeof(double) * b->nrows));
  SA_HIP(hipMalloc(&part, sizeof(double) * 1024));
  SA_HIP(hipMalloc(&out, sizeof(double)));
  Mat A;
  A.n = (int)b->nrows;
  A.nnz = b->nnz;
  A.prec = SA_FP64;
  A.rowptr = B.rp;
  A.col = B.ci;
  A.val = B.val;
  A.lanes = B.lanes;
  int rc = 0;
  for (int p = 0; p < b->P && !rc; p++) {
    for (int i = 0; i < nev; i++) {
      if (i >= b->nconv[p]) { err[(size_t)p * nev + i] = 1e300; continue; }
      const int k = b->perm[p][i];
      const double lam = b->eig[p][k];
      const double *x = (const double *)B.V + (size_t)k * B.ldv;
      Mat sub = A;   // rows of problem p only (column indices are global)
      sub.n = b->n[p];
      sub.rowptr = B.rp + b->off[p];
      if ((rc = launch_spmv(sub, x, y + b->off[p], s))) break;
      if ((rc = launch_axpy_norm(SA_FP64, y + b->off[p], x + b->off[p], lam, b->n[p], part, out, s))) break;
      double h = 0.0;
      if (hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) { rc = 1; break; }
      err[(size_t)p * nev + i] = std::sqrt(h) / std::fabs(lam);
    }
  }
  (void)hipFree(y);
  (void)hipFree(part);
  (void)hipFree(out);
  return rc;
}

// matvecs, Gram-Schmidt passes, restart rounds (FP64 then FP32 phase), seconds.
int sa_batch_get_stats(sa_batch b, int64_t *counters6, double *seconds3) {
  SA_CHECK(b && b->solved, "batch not solved");
  if (counters6) {
    counters6[0] = b->matvecs; counters6[1] = b->passes; counters6[2] = b->its_max;
    counters6[3] = b->matvecs32; counters6[4] = b->passes32; counters6[5] = b->its_max32;
  }
  if (seconds3) {
    seconds3[0] = b->seconds; seconds3[1] = b->seconds32; seconds3[2] = b->seconds_ds;
  }
  return 0;
}

// Sweep kernel choice: 0 autotune (time both on the first restart round of
// the first solve phase of each precision and basis width, keep the faster
// for later solves of the batch), 1 one workgroup per problem, 2 lock-step.
// Setting a mode forgets earlier autotuner choices.
// EPSSetTrueResidual for every problem of the batch: the convergence test
// takes explicit residuals of the candidate Ritz vectors, formed and
// evaluated on the device for all problems at once.
int sa_batch_set_true_residual(sa_batch b, int trueres) {
  SA_CHECK(b, "null batch");
  b->trueres = trueres != 0;
  return 0;
}

int sa_batch_set_sweep(sa_batch b, int mode) {
  SA_CHECK(b && mode >= 0 && mode <= 2, "invalid sweep mode");
  b->sweep_req = mode;
  b->tuned_m[0] = b->tuned_m[1] = 0;
  return 0;
}

int sa_batch_get_sweep(sa_batch b, int *mode, double *ms_persistent, double *ms_wide) {
  SA_CHECK(b, "null batch");
  if (mode) *mode = b->sweep_used;
  if (ms_persistent) *ms_persistent = b->ms_persistent;
  if (ms_wide) *ms_wide = b->ms_wide;
  return 0;
}

int sa_batch_get_schedule(sa_batch b, int *kind, int *nt, int *lanes, double *ms4) {
  SA_CHECK(b, "null batch");
  if (kind) *kind = b->sched_used.kind;
  if (nt) *nt = b->sched_used.nt;
  if (lanes) *lanes = b->sched_used.lanes;
  if (ms4)
    for (int q = 0; q < 4; q++) ms4[q] = b->tune_ms[q];
  return 0;
}

int sa_batch_set_graphs(sa_batch b, int on) {
  SA_CHECK(b, "null batch");
  b->graphs = on != 0;
  return 0;
}

int sa_batch_get_graph_stats(sa_batch b, int64_t *captures, int64_t *replays, int *cached) {
  SA_CHECK(b, "null batch");
  const BatchDev *Bs[2] = {&b->dev64, &b->dev32};
  int64_t c = 0, r = 0;
  int k = 0;
  for (const BatchDev *B : Bs) {
    c += (int64_t)B->graph_captures;
    r += (int64_t)B->graph_replays;
    k += (int)B->graphs.size();
  }
  if (captures) *captures = c;
  if (replays) *replays = r;
  if (cached) *cached = k;
  return 0;
}

int sa_batch_get_ncv(sa_batch b, int *ncv) {
  SA_CHECK(b && ncv && b->solved, "batch not solved");
  std::copy(b->ncv_used.begin(), b->ncv_used.end(), ncv);
  return 0;
}

// Autotuner of the basis size: problems sorted by size are split into nclass
// classes of (nearly) equal count and nsample problems spread over each class
// form one stratified sample.  The sample is solved as a sub-batch with every
// candidate ncv (FP64, lock-step sweeps, this batch's true-residual setting,
// start vectors U[-1,1) with seed sa_sub_seed(seed, p)); each candidate's
// wall time is split over the classes in proportion to their problems'
// algorithmic bytes (lock-step kernels run all problems' chunks together, so
// a problem's share of the time is its share of the traffic), and the
// candidate with the least time per converged pair becomes the ncv of every
// problem of the class.  Later sa_batch_solve calls with ncv == 0 use it.
int sa_batch_autotune_ncv(sa_batch b, int which, int nev, double tol, const int *cand, int ncand, int nclass,
                          int nsample, uint64_t seed, int *ncv_out) {
  SA_TRY(ensure_init());
  SA_CHECK(b && cand && ncand > 0 && ncand <= 16 && nclass > 0 && nsample > 0 && nev > 0 && tol > 0,
           "invalid autotune arguments");
  const int P = b->P;
  nclass = std::min(nclass, P);
  int maxc = 0;
  for (int i = 0; i < ncand; i++) {
    SA_CHECK(cand[i] > nev && cand[i] <= 64, "ncv candidates must satisfy nev < ncv <= 64");
    maxc = std::max(maxc, cand[i]);
  }
  std::vector<int> idx(P);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return b->n[x] < b->n[y]; });
  // classes and the stratified sample
  std::vector<int> cls(P, 0), hi(nclass, 0);
  std::vector<char> sub(P, 0);
  int nmin = INT32_MAX;
  for (int c = 0; c < nclass; c++) {
    const int lo = (int)((int64_t)c * P / nclass), up = (int)((int64_t)(c + 1) * P / nclass), cnt = up - lo;
    hi[c] = b->n[idx[up - 1]];
    for (int i = lo; i < up; i++) cls[idx[i]] = c;
    const int ns = std::min(nsample, cnt);
    for (int i = 0; i < ns; i++) {
      const int p = idx[lo + (int)(((int64_t)2 * i + 1) * cnt / (2 * ns))];
      sub[p] = 1;
      nmin = std::min(nmin, b->n[p]);
    }
  }
  hipStream_t s = stream();
  SA_TRY(alloc_dev(b, b->dev64, SA_FP64, maxc + 1));
  BatchDev &B = b->dev64;
  std::vector<double> x0((size_t)b->nrows, 0.0);
  for (int p = 0; p < P; p++) sa_rand_uniform(sa_sub_seed(seed, p), b->n[p], x0.data() + b->off[p]);
  const int sweep_req = b->sweep_req;
  const bool graphs = b->graphs;
  b->sweep_req = 2;
  b->graphs = false;   // time the kernels, not first captures
  std::vector<double> scores((size_t)nclass * ncand, 1e300), pbytes;
  std::vector<ProbState> ps;
  int rc = 0;
  for (int q = 0; q < ncand && !rc; q++) {
    if (cand[q] > nmin) continue;
    std::vector<int> pncv(P, cand[q]);
    if (hipMemcpyAsync(B.V, x0.data(), sizeof(double) * b->nrows, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("autotune: start vector upload failed");
      rc = 1;
      break;
    }
    int64_t mv = 0, pa = 0;
    int itm = 0;
    double tds = 0.0;
    const double t0 = nowb();
    rc = batch_ks(b, B, which, nev, pncv, tol, 100000, b->trueres != 0, ps, mv, pa, itm, tds, &sub, &pbytes);
    const double dt = nowb() - t0;
    if (rc) break;
    double tot = 0.0;
    std::vector<double> cb(nclass, 0.0);
    std::vector<int> cpairs(nclass, 0), cwant(nclass, 0);
    for (int p = 0; p < P; p++)
      if (sub[p]) {
        tot += pbytes[p];
        cb[cls[p]] += pbytes[p];
        cpairs[cls[p]] += std::min(nev, ps[p].nconv);
        cwant[cls[p]] += nev;
      }
    for (int c = 0; c < nclass; c++)
      if (cwant[c] > 0 && cpairs[c] == cwant[c] && tot > 0.0)
        scores[(size_t)c * ncand + q] = dt * cb[c] / tot / cpairs[c];
  }
  b->sweep_req = sweep_req;
  b->graphs = graphs;
  if (rc) return rc;
  b->pncv.assign(P, 0);
  for (int c = 0; c < nclass; c++) {
    int best = 0;
    for (int q = 1; q < ncand; q++)
      if (scores[(size_t)c * ncand + q] < scores[(size_t)c * ncand + best]) best = q;
    SA_CHECK(scores[(size_t)c * ncand + best] < 1e300, "autotune: no candidate converged on a size class");
    for (int p = 0; p < P; p++)
      if (cls[p] == c) b->pncv[p] = std::min(cand[best], b->n[p]);
  }
  b->ncv_scores = scores;
  b->ncv_cand.assign(cand, cand + ncand);
  b->class_hi = hi;
  if (ncv_out) std::copy(b->pncv.begin(), b->pncv.end(), ncv_out);
  return 0;
}

// The autotuner's measurements: for class c (problems with n <= class_hi[c]),
// scores[c * ncand + q] = seconds per converged pair with candidate q (1e300:
// not run or not all pairs converged).
int sa_batch_get_ncv_tuning(sa_batch b, int *nclass, int *ncand, int *class_hi, double *scores) {
  SA_CHECK(b && !b->pncv.empty(), "the ncv autotuner has not run");
  const int nc = (int)b->class_hi.size(), nq = (int)b->ncv_cand.size();
  if (nclass) *nclass = nc;
  if (ncand) *ncand = nq;
  if (class_hi) std::copy(b->class_hi.begin(), b->class_hi.end(), class_hi);
  if (scores) std::copy(b->ncv_scores.begin(), b->ncv_scores.end(), scores);
  return 0;
}

}  // extern "C"
