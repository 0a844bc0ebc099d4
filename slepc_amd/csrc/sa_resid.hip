// Residual norms ||A x - a x + b z|| of approximate eigenpairs, the SpMV fused
// with the axpy and the squared-norm partial sums (EPSComputeResidualNorm_Private,
// src/eps/interface/epssolve.c:666-718: a real pair takes z = 0; the real and
// imaginary parts of a complex pair each take the other part as z).  Used by
// the true-residual convergence test (epskrylov.c:225-263), by EPSComputeError
// and by the sa_mat_residual_norms entry points.
//
// One launch handles up to RESID_MAXVEC vectors (blockIdx.y); each of G
// workgroups strides over row groups of 256/L rows with L lanes per row
// (vector CSR, shuffle reduction) and stores the sum of its squared residual
// entries at part[v * G + b].  A one-workgroup kernel then adds the G partials
// of each vector in a fixed order, so equal inputs give equal bits.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

template <typename T, int L>
__global__ __launch_bounds__(BLK) void resid_partial_kernel(int n, const int32_t *__restrict__ rp,
                                                            const int32_t *__restrict__ ci,
                                                            const T *__restrict__ val, const T *__restrict__ X,
                                                            int64_t ldx, ResidArgs args, double *__restrict__ part) {
  constexpr int RPW = BLK / L;                  // rows per workgroup step
  __shared__ double red[BLK / 64];
  const int v = blockIdx.y;
  const T *x = X + (size_t)args.xcol[v] * ldx;
  const T *z = args.zcol[v] >= 0 ? X + (size_t)args.zcol[v] * ldx : nullptr;
  const double a = args.a[v], b = args.b[v];
  const int lg = threadIdx.x % L;
  double acc = 0.0;
  for (int64_t r0 = (int64_t)blockIdx.x * RPW; r0 < n; r0 += (int64_t)gridDim.x * RPW) {
    const int64_t row = r0 + threadIdx.x / L;
    T s = (T)0;
    if (row < n) {
      const int32_t pe = rp[row + 1];
      for (int32_t p = rp[row] + lg; p < pe; p += L) s += val[p] * x[ci[p]];
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
    if (lg == 0 && row < n) {
      double r = (double)s - a * (double)x[row];
      if (z) r += b * (double)z[row];
      acc += r * r;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < BLK / 64; w++) t += red[w];
    part[(size_t)v * gridDim.x + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(RBLK) void resid_finish_kernel(const double *part, int G, int nvec, double *out) {
  sum_partials_block(part, G, nvec, out);
}

int resid_grid(int64_t n, int lanes) {
  const int64_t rows_per = BLK / std::max(1, lanes);
  return (int)std::max<int64_t>(1, std::min<int64_t>(RESID_MAXG, (n + rows_per - 1) / rows_per));
}

template <typename T>
static void resid_launch(const Mat &A, const T *X, int64_t ldx, const ResidArgs &args, int G, double *part,
                         hipStream_t s) {
  const dim3 grid((unsigned)G, (unsigned)args.nvec);
  const T *val = (const T *)A.val;
  switch (A.lanes) {
    case 4: hipLaunchKernelGGL((resid_partial_kernel<T, 4>), grid, dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, args, part); break;
    case 8: hipLaunchKernelGGL((resid_partial_kernel<T, 8>), grid, dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, args, part); break;
    case 16: hipLaunchKernelGGL((resid_partial_kernel<T, 16>), grid, dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, args, part); break;
    default: hipLaunchKernelGGL((resid_partial_kernel<T, 32>), grid, dim3(BLK), 0, s, A.n, A.rowptr, A.col, val, X, ldx, args, part); break;
  }
}

int launch_resid_norms(const Mat &A, const void *X, int64_t ldx, const ResidArgs &args, double *part, double *out,
                       hipStream_t s) {
  SA_CHECK(args.nvec >= 1 && args.nvec <= RESID_MAXVEC, "residual norms: 1 to 16 vectors per launch");
  SA_CHECK(A.n > 0, "residual norms of an empty operator");
  const int G = resid_grid(A.n, A.lanes);
  if (A.prec == SA_FP64)
    resid_launch<double>(A, (const double *)X, ldx, args, G, part, s);
  else
    resid_launch<float>(A, (const float *)X, ldx, args, G, part, s);
  hipLaunchKernelGGL(resid_finish_kernel, dim3(1), dim3(RBLK), 0, s, part, G, args.nvec, out);
  SA_HIP(hipGetLastError());
  return 0;
}

// Pairs (real, or the real and imaginary parts of a complex pair in columns
// c, c + 1) turned into the launch records of EPSComputeResidualNorm_Private:
// a pair whose imaginary part is below |re| eps is real; |lambda| <= eps drops
// the lambda terms.  Host state lives in fixed arrays (no allocation).
int residual_norms(const Mat &A, const void *X, int64_t ldx, int k, const double *lr, const double *li,
                   double *res, double *part, double *out, hipStream_t s) {
  constexpr double EPS = 2.220446049250313e-16;
  ResidArgs args{};
  int owner[RESID_MAXVEC];   // first column of the pair each launch record belongs to
  double h[RESID_MAXVEC];
  auto flush = [&]() -> int {
    if (args.nvec == 0) return 0;
    SA_TRY(launch_resid_norms(A, X, ldx, args, part, out, s));
    SA_HIP(hipMemcpyAsync(h, out, sizeof(double) * args.nvec, hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    for (int v = 0; v < args.nvec;) {
      const int c = owner[v];
      const bool pair = v + 1 < args.nvec && owner[v + 1] == c;
      const double nrm = pair ? std::hypot(std::sqrt(h[v]), std::sqrt(h[v + 1])) : std::sqrt(h[v]);
      res[c] = nrm;
      if (pair) {
        res[c + 1] = nrm;
        v += 2;
      } else {
        v += 1;
      }
    }
    args.nvec = 0;
    return 0;
  };
  for (int i = 0; i < k;) {
    const double kr = lr[i], ki = li ? li[i] : 0.0;
    const bool cplx = ki != 0.0 && std::fabs(ki) >= std::fabs(kr * EPS);
    if (!cplx) {
      if (args.nvec + 1 > RESID_MAXVEC) SA_TRY(flush());
      const int v = args.nvec++;
      args.xcol[v] = i;
      args.zcol[v] = -1;
      args.a[v] = std::fabs(kr) > EPS ? kr : 0.0;
      args.b[v] = 0.0;
      owner[v] = i;
      i += 1;
      continue;
    }
    SA_CHECK(i + 1 < k, "a complex pair needs its imaginary part in the next column");
    if (args.nvec + 2 > RESID_MAXVEC) SA_TRY(flush());
    const bool big = std::hypot(kr, ki) > EPS;
    const int v = args.nvec;
    args.xcol[v] = i;          // A xr - kr xr + ki xi
    args.zcol[v] = big ? i + 1 : -1;
    args.a[v] = big ? kr : 0.0;
    args.b[v] = big ? ki : 0.0;
    args.xcol[v + 1] = i + 1;  // A xi - kr xi - ki xr
    args.zcol[v + 1] = big ? i : -1;
    args.a[v + 1] = big ? kr : 0.0;
    args.b[v + 1] = big ? -ki : 0.0;
    owner[v] = owner[v + 1] = i;
    args.nvec += 2;
    i += 2;
  }
  return flush();
}

// ---- scratch of the no-allocation entry point ------------------------------
// The partial sums and their totals of one launch (RESID_MAXVEC vectors of at
// most RESID_MAXG partials: 131,200 bytes) in a file-scope device array, so
// that sa_mat_residual_norms_noalloc allocates nothing.  Each GPU holds its
// own copy for the life of the process; its device address is taken once
// per device at initialisation (the array's name in host code is only a
// host-side handle).  Every partial and total is written by the launch that
// reads it, so nothing needs clearing between calls; calls that overlap on
// one GPU share the array and race.
constexpr size_t NOALLOC_DOUBLES = (size_t)RESID_MAXVEC * RESID_MAXG + RESID_MAXVEC;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAXDEV = 64;
static double *g_scratch_dev[NOALLOC_MAXDEV] = {};
static std::mutex g_scratch_mu;

int noalloc_scratch_setup(int device) {
  SA_CHECK(device >= 0 && device < NOALLOC_MAXDEV, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return 0;
  void *p = nullptr;
  SA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  SA_CHECK(p != nullptr, "no device address for the residual scratch");
  g_scratch_dev[device] = static_cast<double *>(p);
  return 0;
}

double *noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAXDEV) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_dev[device];
}

}  // namespace sa
