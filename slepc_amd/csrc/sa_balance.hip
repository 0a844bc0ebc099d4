// Balancing of non-Hermitian problems (EPSSetBalance; EPSBuildBalance_Krylov,
// src/eps/interface/epsdefault.c:443-504, applied through STSetBalanceMatrix
// and undone in EPSComputeVectors_Schur, epsdefault.c:128-139).  The diagonal
// D is built on the device from random +-1 vectors (the library's uniform
// generator, seeds 0xBA1A + j, stands in for PetscRandom); the solve then runs
// on the explicitly scaled operator B = D A D^-1, whose values are those of A
// times D_row / D_col (one SpMV per application instead of scale, SpMV,
// scale: the same operator up to rounding), and the eigenvectors are divided
// by D and normalised, complex pairs together.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

namespace {

constexpr int BAL_MAXG = 1024;

__global__ __launch_bounds__(BLK) void bal_div_kernel(const double *x, const double *D, double *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i < n) y[i] = x[i] / D[i];
}

__global__ __launch_bounds__(BLK) void bal_mul_kernel(double *x, const double *D, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i < n) x[i] *= D[i];
}

// p = |p| in place, part[b] = the block's largest |p|
__global__ __launch_bounds__(BLK) void bal_absmax_kernel(double *p, int64_t n, double *part) {
  __shared__ double red[BLK];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const double a = fabs(p[i]);
    p[i] = a;
    m = fmax(m, a);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// oneside: D /= |p| where p != 0; twoside: r /= D, then D *= sqrt(|r / p|)
// where |p| > cut and r != 0 (epsdefault.c:484-498)
__global__ __launch_bounds__(BLK) void bal_update_kernel(const double *p, double *r, double *D, int64_t n,
                                                         int twoside, double cut) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  if (twoside) {
    const double rr = r[i] / D[i];
    r[i] = rr;
    if (fabs(p[i]) > cut && rr != 0.0) D[i] *= sqrt(fabs(rr / p[i]));
  } else if (p[i] != 0.0) {
    D[i] /= fabs(p[i]);
  }
}

// val(q) *= D(row) / D(col(q)), one thread per row
__global__ __launch_bounds__(BLK) void bal_scale_csr_kernel(const int32_t *rp, const int32_t *ci, double *val,
                                                            const double *D, int n) {
  const int row = blockIdx.x * BLK + threadIdx.x;
  if (row >= n) return;
  const double dr = D[row];
  for (int32_t q = rp[row]; q < rp[row + 1]; q++) val[q] = val[q] * dr / D[ci[q]];
}

}  // namespace

int eps_build_balance(const Mat &A, const Mat *AT, bool twoside, int its, double cutoff, double *D) {
  SA_CHECK(A.prec == SA_FP64, "balancing runs in FP64");
  SA_CHECK(!twoside || (AT && AT->n == A.n), "two-sided balancing needs the transposed operator");
  hipStream_t s = stream();
  const int64_t n = A.n;
  const unsigned nb = (unsigned)std::max<int64_t>(1, (n + BLK - 1) / BLK);
  const int G = (int)std::max<int64_t>(1, std::min<int64_t>(BAL_MAXG, (n + BLK - 1) / BLK));
  double *buf = nullptr;
  SA_HIP(hipMalloc(&buf, sizeof(double) * (size_t)(3 * n + BAL_MAXG)));
  struct Free {
    double *p;
    ~Free() { (void)hipFree(p); }
  } guard{buf};
  double *z = buf, *r = buf + n, *p = buf + 2 * n, *part = buf + 3 * n;
  std::vector<double> h(n, 1.0), hp(G);
  SA_HIP(hipMemcpyAsync(D, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  double norma = 0.0;
  for (int j = 0; j < its; j++) {
    sa_rand_uniform(0xBA1AULL + (uint64_t)j, n, h.data());
    // a [0,1) draw (PetscRandom's default interval) from the U[-1,1) stream, signed at 0.5
    for (int64_t i = 0; i < n; i++) h[i] = 0.5 * (h[i] + 1.0) < 0.5 ? -1.0 : 1.0;
    SA_HIP(hipMemcpyAsync(z, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bal_div_kernel, dim3(nb), dim3(BLK), 0, s, z, D, r, n);   // r = z ./ D
    SA_TRY(launch_spmv(A, r, p, s));                                              // p = A r
    hipLaunchKernelGGL(bal_mul_kernel, dim3(nb), dim3(BLK), 0, s, p, D, n);       // p .*= D
    if (twoside) {
      if (j == 0) {   // matrix inf-norm estimate (p = |p|)
        hipLaunchKernelGGL(bal_absmax_kernel, dim3(G), dim3(BLK), 0, s, p, n, part);
        SA_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * G, hipMemcpyDeviceToHost, s));
        SA_HIP(hipStreamSynchronize(s));
        norma = *std::max_element(hp.begin(), hp.end());
      }
      hipLaunchKernelGGL(bal_mul_kernel, dim3(nb), dim3(BLK), 0, s, z, D, n);     // z .*= D
      SA_TRY(launch_spmv(*AT, z, r, s));                                          // r = A^T z
    }
    hipLaunchKernelGGL(bal_update_kernel, dim3(nb), dim3(BLK), 0, s, p, r, D, n, twoside ? 1 : 0,
                       cutoff * norma);
    SA_HIP(hipGetLastError());
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int balance_scale_mat(Mat &B, const double *D) {
  SA_CHECK(B.prec == SA_FP64, "balancing runs in FP64");
  hipStream_t s = stream();
  const unsigned nb = (unsigned)std::max(1, (B.n + BLK - 1) / BLK);
  hipLaunchKernelGGL(bal_scale_csr_kernel, dim3(nb), dim3(BLK), 0, s, B.rowptr, B.col, (double *)B.val, D, B.n);
  SA_HIP(hipGetLastError());
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int balance_unscale_vectors(BV &V, const double *D, int nconv, const double *eigi) {
  SA_CHECK(V.prec == SA_FP64 && nconv <= V.m, "balancing runs in FP64");
  hipStream_t s = stream();
  const unsigned nb = (unsigned)std::max<int64_t>(1, (V.n + BLK - 1) / BLK);
  std::vector<double> nrm(nconv, 0.0);
  for (int i = 0; i < nconv; i++) {
    hipLaunchKernelGGL(bal_div_kernel, dim3(nb), dim3(BLK), 0, s, (const double *)col_ptr(V, i), D,
                       (double *)col_ptr(V, i), V.n);
    SA_TRY(launch_dot_self(V.prec, col_ptr(V, i), V.n, V.ws.dots, V.ws.part, s));
    SA_HIP(hipMemcpyAsync(&nrm[i], V.ws.dots, sizeof(double), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
  }
  for (int i = 0; i < nconv; i++) {   // BVNormalize with eigi: a pair shares one norm
    if (eigi[i] != 0.0 && i + 1 < nconv) {
      const double t = std::sqrt(nrm[i] + nrm[i + 1]);
      SA_TRY(launch_scale(V.prec, col_ptr(V, i), V.n, 1.0 / t, s));
      SA_TRY(launch_scale(V.prec, col_ptr(V, i + 1), V.n, 1.0 / t, s));
      i++;
    } else {
      SA_TRY(launch_scale(V.prec, col_ptr(V, i), V.n, 1.0 / std::sqrt(nrm[i]), s));
    }
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace sa
