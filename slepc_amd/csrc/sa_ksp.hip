// Linear solves for the shift-and-invert spectral transformation
// (STSINVERT: src/sys/classes/st/impls/sinvert/sinvert.c, STMatSolve in
// st/interface/stsles.c): x = (A - sigma I)^-1 b by banded LU (the default,
// PREONLY + LU), by BiCGStab or by GMRES(m), the iterative ones with right
// Jacobi (or no) preconditioning, in FP64.  BiCGStab keeps every scalar in
// device memory so that one iteration is a chain of stream-ordered kernels
// with no host round trip; the host reads the residual norm every KSP_CHECK
// iterations.  Dot products are per-workgroup partials summed by one
// workgroup in a fixed order (deterministic).
#include <algorithm>
#include <cmath>
#include <vector>

#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

namespace {

constexpr int KSP_MAXG = 1024;   // partial-sum workgroups
constexpr int KSP_CHECK = 8;     // iterations between host residual checks

// scalar slots in device memory
enum { S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_RR, S_FLAG, S_NSCAL };

__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += sh[q];
  return s;
}

__global__ __launch_bounds__(BLK) void ksp_init_kernel(int64_t n, const double *__restrict__ b, double *r,
                                                       double *rhat, double *p, double *v, double *x, double *scal) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i < n) {
    r[i] = b[i];
    rhat[i] = b[i];
    p[i] = 0.0;
    v[i] = 0.0;
    x[i] = 0.0;
  }
  if (i == 0) {
    scal[S_RHO] = 1.0;
    scal[S_ALPHA] = 1.0;
    scal[S_OMEGA] = 1.0;
    scal[S_BETA] = 0.0;
    scal[S_RR] = 0.0;
    scal[S_FLAG] = 0.0;
  }
}

// part[g] = partial a.b, part[G + g] = partial c.d (c may be null)
__global__ __launch_bounds__(BLK) void ksp_dot2_kernel(int64_t n, const double *__restrict__ a,
                                                       const double *__restrict__ b, const double *__restrict__ c,
                                                       const double *__restrict__ d, double *part, int G) {
  __shared__ double sh[BLK / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)G * BLK) {
    s1 += a[i] * b[i];
    if (c) s2 += c[i] * d[i];
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[G + blockIdx.x] = s2;
  }
}

// y -= sigma x in place, then the partials of w.y (and of y.z when z given)
__global__ __launch_bounds__(BLK) void ksp_shift_dot_kernel(int64_t n, double *y, const double *__restrict__ x,
                                                            double sigma, const double *__restrict__ w,
                                                            const double *__restrict__ z, double *part, int G) {
  __shared__ double sh[BLK / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)G * BLK) {
    const double yi = y[i] - sigma * x[i];
    y[i] = yi;
    s1 += w[i] * yi;
    if (z) s2 += yi * z[i];
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[G + blockIdx.x] = s2;
  }
}

// one workgroup: d0, d1 = sums of the partials (fixed order), then
// op 0: rho' = d0, rr = d1, beta = (rho'/rho)(alpha/omega), rho = rho'
// op 1: alpha = rho / d0
// op 2: omega = d0 / d1 (t.s / t.t)
// op 3: scal[0] = d0, scal[1] = d1 (plain sums, for the host checks)
__global__ __launch_bounds__(BLK) void ksp_scalar_kernel(const double *__restrict__ part, int G, int op,
                                                         double *scal) {
  __shared__ double sh[BLK / 64];
  double a = 0.0, b = 0.0;
  for (int g = threadIdx.x; g < G; g += BLK) {
    a += part[g];
    b += part[G + g];
  }
  const double d0 = block_sum(a, sh), d1 = block_sum(b, sh);
  if (threadIdx.x == 0) {
    if (op == 0) {
      const double rho = scal[S_RHO], alpha = scal[S_ALPHA], omega = scal[S_OMEGA];
      const bool bad = d0 == 0.0 || rho == 0.0 || omega == 0.0;
      scal[S_BETA] = bad ? 0.0 : (d0 / rho) * (alpha / omega);
      scal[S_RHO] = d0;
      scal[S_RR] = d1;
      if (bad && d1 != 0.0) scal[S_FLAG] = 1.0;   // breakdown with a nonzero residual
    } else if (op == 1) {
      scal[S_ALPHA] = d0 != 0.0 ? scal[S_RHO] / d0 : 0.0;
      if (d0 == 0.0) scal[S_FLAG] = 1.0;
    } else if (op == 2) {
      scal[S_OMEGA] = d1 != 0.0 ? d0 / d1 : 0.0;
    } else {
      scal[0] = d0;
      scal[1] = d1;
    }
  }
}

// p = r + beta (p - omega v); phat = dinv .* p
__global__ __launch_bounds__(BLK) void ksp_p_kernel(int64_t n, const double *__restrict__ r, double *p,
                                                    const double *__restrict__ v, const double *__restrict__ dinv,
                                                    double *phat, const double *__restrict__ scal) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const double beta = scal[S_BETA], omega = scal[S_OMEGA];
  const double pi = r[i] + beta * (p[i] - omega * v[i]);
  p[i] = pi;
  phat[i] = dinv ? dinv[i] * pi : pi;
}

// s = r - alpha v; shat = dinv .* s
__global__ __launch_bounds__(BLK) void ksp_s_kernel(int64_t n, const double *__restrict__ r,
                                                    const double *__restrict__ v, const double *__restrict__ dinv,
                                                    double *s, double *shat, const double *__restrict__ scal) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const double si = r[i] - scal[S_ALPHA] * v[i];
  s[i] = si;
  shat[i] = dinv ? dinv[i] * si : si;
}

// x += alpha phat + omega shat; r = s - omega t
__global__ __launch_bounds__(BLK) void ksp_x_kernel(int64_t n, double *x, const double *__restrict__ phat,
                                                    const double *__restrict__ shat, double *r,
                                                    const double *__restrict__ s, const double *__restrict__ t,
                                                    const double *__restrict__ scal) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const double alpha = scal[S_ALPHA], omega = scal[S_OMEGA];
  x[i] += alpha * phat[i] + omega * shat[i];
  r[i] = s[i] - omega * t[i];
}

// dinv = 1 / (a_ii - sigma) (PCJACOBI on A - sigma I; 1 for a zero diagonal)
__global__ __launch_bounds__(BLK) void ksp_jacobi_kernel(int n, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ ci,
                                                         const double *__restrict__ val, double sigma, double *dinv) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  double d = -sigma;
  for (int32_t q = rp[i]; q < rp[i + 1]; q++)
    if (ci[q] == i) d += val[q];
  dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
}

// ---------------------------------------------------------------------------
// GMRES(m) with right preconditioning (KSPGMRES, PETSc's default KSP type,
// restart 30): the Arnoldi basis is a BV orthonormalised by the fused
// Gram-Schmidt kernels with classical Gram-Schmidt and refinement never (the
// defaults of KSPGMRES), which leave the Hessenberg column in the coefficient
// buffer; these kernels do the rest of each step.
// ---------------------------------------------------------------------------

// dst = dinv .* src (src copied when dinv is null)
__global__ __launch_bounds__(BLK) void gm_prec_kernel(int64_t n, const double *__restrict__ dinv,
                                                      const double *__restrict__ src, double *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  dst[i] = dinv ? dinv[i] * src[i] : src[i];
}

// y -= sigma z
__global__ __launch_bounds__(BLK) void gm_shift_kernel(int64_t n, double *y, const double *__restrict__ z,
                                                       double sigma) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  y[i] -= sigma * z[i];
}

// r = b - (t - sigma x), t = A x
__global__ __launch_bounds__(BLK) void gm_resid_kernel(int64_t n, const double *__restrict__ b,
                                                       const double *__restrict__ t, const double *__restrict__ x,
                                                       double sigma, double *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  r[i] = b[i] - (t[i] - sigma * x[i]);
}

// x += dinv .* z (z when dinv is null)
__global__ __launch_bounds__(BLK) void gm_update_kernel(int64_t n, double *x, const double *__restrict__ dinv,
                                                        const double *__restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  x[i] += dinv ? dinv[i] * z[i] : z[i];
}

// ---------------------------------------------------------------------------
// Direct solves (KSPPREONLY + PCLU, the reference's default for STSINVERT,
// stsles.c:53-54): banded LU with partial pivoting of A - sigma I in LAPACK's
// band layout (dgbtf2/dgbtrs): element (r, c) at ab[(kv + r - c) + c * ldab],
// kv = kl + ku, ldab = 2 kl + ku + 1, the rows above ku holding the fill-in.
// One workgroup runs the column loop (the pivot search, the row swap, the
// multipliers and the rank-1 update of the band each a parallel phase); the
// solves run the two substitutions the same way.
// ---------------------------------------------------------------------------
constexpr int LU_T = 1024;

__global__ void band_width_kernel(int n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci, int *klku) {
  const int r = blockIdx.x * BLK + threadIdx.x;
  if (r >= n) return;
  int lo = 0, hi = 0;
  for (int32_t q = rp[r]; q < rp[r + 1]; q++) {
    lo = max(lo, r - ci[q]);
    hi = max(hi, ci[q] - r);
  }
  atomicMax(klku, lo);
  atomicMax(klku + 1, hi);
}

__global__ void band_fill_kernel(int n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                 const double *__restrict__ val, double sigma, int kv, int64_t ldab, double *ab) {
  const int r = blockIdx.x * BLK + threadIdx.x;
  if (r >= n) return;
  ab[kv + (int64_t)r * ldab] = -sigma;   // the shift, also where the pattern has no diagonal entry
  for (int32_t q = rp[r]; q < rp[r + 1]; q++) {
    const int c = ci[q];
    ab[(kv + r - c) + (int64_t)c * ldab] = val[q] - (c == r ? sigma : 0.0);
  }
}

// dgbtf2: info[0] = first zero pivot + 1, or 0
__global__ __launch_bounds__(LU_T) void band_lu_kernel(int n, int kl, int ku, int64_t ldab, double *ab, int *ipiv,
                                                       int *info) {
  __shared__ double smax[LU_T / 64];
  __shared__ int sidx[LU_T / 64];
  __shared__ int s_jp, s_ju;
  const int t = threadIdx.x, kv = kl + ku;
  if (t == 0) { s_ju = 0; info[0] = 0; }
  __syncthreads();
  for (int j = 0; j < n; j++) {
    double *col = ab + (int64_t)j * ldab;
    const int km = min(kl, n - 1 - j);
    // pivot: first index of the largest |a| among rows j..j+km (idamax)
    double best = -1.0;
    int bi = 0;
    for (int i = t; i <= km; i += LU_T) {
      const double a = fabs(col[kv + i]);
      if (a > best) { best = a; bi = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if ((t & 63) == 0) { smax[t >> 6] = best; sidx[t >> 6] = bi; }
    __syncthreads();
    if (t == 0) {
      double b = smax[0];
      int i0 = sidx[0];
      for (int w = 1; w < LU_T / 64; w++)
        if (smax[w] > b || (smax[w] == b && sidx[w] < i0)) { b = smax[w]; i0 = sidx[w]; }
      s_jp = i0;
      ipiv[j] = j + i0;
      if (col[kv + i0] != 0.0) s_ju = max(s_ju, min(j + ku + i0, n - 1));
      else if (info[0] == 0) info[0] = j + 1;
    }
    __syncthreads();
    const int jp = s_jp, ju = s_ju;
    if (col[kv + jp] == 0.0) continue;          // uniform: singular column, nothing to eliminate
    if (jp != 0)                                // swap rows j and j+jp over columns j..ju
      for (int c = j + t; c <= ju; c += LU_T) {
        double *a = ab + (kv + j - c) + (int64_t)c * ldab, *b = a + jp;
        const double x = *a;
        *a = *b;
        *b = x;
      }
    __syncthreads();
    const double rp_ = 1.0 / col[kv];
    for (int i = 1 + t; i <= km; i += LU_T) col[kv + i] *= rp_;
    __syncthreads();
    const int nc = ju - j;                       // rank-1 update of rows j+1..j+km, columns j+1..ju
    for (int e = t; e < nc * km; e += LU_T) {
      const int cc = e / km, i = e % km + 1, c = j + 1 + cc;
      double *cb = ab + (int64_t)c * ldab + (kv + j - c);
      cb[i] -= col[kv + i] * cb[0];
    }
    __syncthreads();
  }
}

// dgbtrs ('N') on one right-hand side held in x
__global__ __launch_bounds__(LU_T) void band_solve_kernel(int n, int kl, int ku, int64_t ldab, const double *ab,
                                                          const int *ipiv, double *x) {
  const int t = threadIdx.x, kv = kl + ku;
  for (int j = 0; j < n - 1; j++) {             // L: swaps and the multipliers of column j
    const int l = ipiv[j];
    if (t == 0 && l != j) {
      const double tmp = x[l];
      x[l] = x[j];
      x[j] = tmp;
    }
    __syncthreads();
    const int lm = min(kl, n - 1 - j);
    const double xj = x[j];
    const double *col = ab + (int64_t)j * ldab + kv;
    for (int i = 1 + t; i <= lm; i += LU_T) x[j + i] -= col[i] * xj;
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; j--) {            // U, column by column
    const double *col = ab + (int64_t)j * ldab;
    __shared__ double s_xj;
    if (t == 0) {
      s_xj = x[j] / col[kv];
      x[j] = s_xj;
    }
    __syncthreads();
    const double xj = s_xj;
    const int i0 = max(0, j - kv);
    for (int i = i0 + t; i < j; i += LU_T) x[i] -= col[kv + i - j] * xj;
    __syncthreads();
  }
}

}  // namespace

// Limits of the device banded LU: band storage and the work of the
// one-workgroup factorization (n kl (kl + ku) multiply-adds).
constexpr int64_t LU_MAX_BAND_DOUBLES = 1ll << 27;   // 1 GiB
constexpr double LU_MAX_WORK = 2e9;

static int band_lu_setup(KspWork &w, const Mat &A, double sigma) {
  const int n = A.n;
  hipStream_t s = stream();
  int *klku = nullptr;
  SA_HIP(hipMalloc(&klku, sizeof(int) * 4));
  struct F { int *p; ~F() { (void)hipFree(p); } } fk{klku};
  SA_HIP(hipMemsetAsync(klku, 0, sizeof(int) * 4, s));
  const unsigned nb = (unsigned)std::max(1, (n + BLK - 1) / BLK);
  hipLaunchKernelGGL(band_width_kernel, dim3(nb), dim3(BLK), 0, s, n, A.rowptr, A.col, klku);
  int h[2] = {0, 0};
  SA_HIP(hipMemcpyAsync(h, klku, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  const int kl = h[0], ku = h[1];
  const int64_t ldab = 2 * (int64_t)kl + ku + 1;
  SA_CHECK(ldab * n <= LU_MAX_BAND_DOUBLES && (double)n * kl * (kl + ku) <= LU_MAX_WORK,
           "the band of A - sigma I is too wide for the device LU (use the BiCGStab solver, "
           "sa_eps_set_st_ksp with SA_PC_JACOBI or SA_PC_NONE)");
  if (w.ab_n != ldab * n) {
    if (w.ab) (void)hipFree(w.ab);
    if (w.ipiv) (void)hipFree(w.ipiv);
    w.ab = nullptr;
    w.ipiv = nullptr;
    SA_HIP(hipMalloc(&w.ab, sizeof(double) * (size_t)std::max<int64_t>(1, ldab * n)));
    SA_HIP(hipMalloc(&w.ipiv, sizeof(int) * (size_t)std::max(1, n) + sizeof(int)));
    w.ab_n = ldab * n;
  }
  w.kl = kl;
  w.ku = ku;
  w.ldab = ldab;
  SA_HIP(hipMemsetAsync(w.ab, 0, sizeof(double) * (size_t)std::max<int64_t>(1, ldab * n), s));
  hipLaunchKernelGGL(band_fill_kernel, dim3(nb), dim3(BLK), 0, s, n, A.rowptr, A.col, (const double *)A.val, sigma,
                     kl + ku, ldab, w.ab);
  int *info = w.ipiv + n;
  hipLaunchKernelGGL(band_lu_kernel, dim3(1), dim3(LU_T), 0, s, n, kl, ku, ldab, w.ab, w.ipiv, info);
  SA_HIP(hipGetLastError());
  int hinfo = 0;
  SA_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  SA_CHECK(hinfo == 0, "shift-and-invert: A - sigma I is exactly singular (zero pivot in the LU)");
  return 0;
}

int ksp_alloc(KspWork &w, const Mat &A, double sigma, bool jacobi, bool direct, bool gmres, int restart) {
  SA_CHECK(A.prec == SA_FP64, "the shift-and-invert solves run in FP64");
  SA_CHECK(!gmres || (restart >= 1 && restart <= GS_MAXC - 2), "GMRES restart must be in [1, 126]");
  const int64_t n = A.n;
  w.direct = direct;
  w.gmres = gmres && !direct;
  w.restart = restart;
  if (direct) {
    w.n = n;
    w.sigma = sigma;
    return band_lu_setup(w, A, sigma);
  }
  if (w.n != n || !w.buf) {
    ksp_free(w);
    SA_HIP(hipMalloc(&w.buf, sizeof(double) * ((size_t)9 * std::max<int64_t>(n, 1) + 2 * KSP_MAXG + 8)));
    w.n = n;
  }
  w.G = (int)std::min<int64_t>(KSP_MAXG, std::max<int64_t>(1, (n + 4 * BLK - 1) / (4 * BLK)));
  w.sigma = sigma;
  w.jacobi = jacobi;
  if (jacobi && n > 0) {
    hipLaunchKernelGGL(ksp_jacobi_kernel, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, stream(), A.n,
                       A.rowptr, A.col, (const double *)A.val, sigma, w.buf + 8 * n);
    SA_HIP(hipGetLastError());
  }
  if (w.gmres && (w.gv.m != restart + 1 || w.gv.n != n)) {
    bv_free(w.gv);
    SA_TRY(bv_alloc(w.gv, n, restart + 1, SA_FP64));
    w.gv.refine = GS_REFINE_NEVER;   // KSP_GMRES_CGS_REFINE_NEVER, PETSc's default for KSPGMRES
  }
  return 0;
}

void ksp_free(KspWork &w) {
  bv_free(w.gv);
  if (w.buf) (void)hipFree(w.buf);
  if (w.ab) (void)hipFree(w.ab);
  if (w.ipiv) (void)hipFree(w.ipiv);
  w.buf = nullptr;
  w.ab = nullptr;
  w.ipiv = nullptr;
  w.ab_n = 0;
  w.n = 0;
}

namespace {

// GMRES(m), right preconditioned: each cycle starts from r = b - (A - sigma I) x
// in column 0 of the basis; step j builds column j+1 = (A - sigma I) M^-1 v_j,
// orthonormalises it (the Hessenberg column in the coefficient buffer), and
// Givens rotations on the host turn the least-squares residual into the
// estimate |g_{j+1}|, tested against rtol ||b|| (KSPConvergedDefault); the
// cycle ends by x += M^-1 V y.  A zero norm (exact breakdown) ends the solve.
int gmres_solve(const Mat &A, KspWork &w, const double *b, double *x, double rtol, int max_it, int *its_out) {
  const int64_t n = w.n;
  hipStream_t s = stream();
  BV &V = w.gv;
  const int m = w.restart;
  SA_CHECK(V.data && V.m == m + 1 && V.n == n, "GMRES workspace does not match the operator");
  double *t = w.buf, *z = w.buf + n, *dinv = w.jacobi ? w.buf + 8 * n : nullptr;
  double *part = w.buf + 9 * n, *spare = part + 2 * KSP_MAXG + S_NSCAL;
  const int G = w.G;
  const unsigned nb = (unsigned)std::max<int64_t>(1, (n + BLK - 1) / BLK);
  const double sigma = w.sigma;
  hipLaunchKernelGGL(ksp_dot2_kernel, dim3(G), dim3(BLK), 0, s, n, b, b, nullptr, nullptr, part, G);
  hipLaunchKernelGGL(ksp_scalar_kernel, dim3(1), dim3(BLK), 0, s, part, G, 3, spare);
  SA_HIP(hipGetLastError());
  double hb = 0.0;
  SA_HIP(hipMemcpyAsync(&hb, spare, sizeof(double), hipMemcpyDeviceToHost, s));
  SA_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  SA_HIP(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(std::max(0.0, hb)), target = rtol * bnorm;
  int its = 0;
  if (bnorm == 0.0) {
    if (its_out) *its_out = 0;
    return 0;
  }
  std::vector<double> H((size_t)m * m), cs(m), sn(m), g(m + 1), h(m + 2), y(m);
  for (bool first = true;; first = false) {
    double *v0 = (double *)col_ptr(V, 0);
    if (first) {
      SA_HIP(hipMemcpyAsync(v0, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    } else {
      SA_TRY(launch_spmv(A, x, t, s));
      hipLaunchKernelGGL(gm_resid_kernel, dim3(nb), dim3(BLK), 0, s, n, b, t, x, sigma, v0);
    }
    SA_TRY(launch_gs_reset(V, s));
    SA_TRY(launch_gs_column(V, nullptr, 0, 0.7071, s));
    SA_HIP(hipMemcpyAsync(V.st_host, V.st + 2, sizeof(GsState), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    const double beta = V.st_host->nrm;
    SA_CHECK(std::isfinite(beta), "shift-and-invert: GMRES residual is not finite in (A - sigma I) x = b");
    if (beta <= target) break;                         // the true residual meets the test
    SA_CHECK(its < max_it, "shift-and-invert: the linear solver did not converge within max_it iterations");
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;                                         // columns of V in this cycle's update
    bool done = false;
    for (int j = 0; j < m && its < max_it; j++) {
      its++;
      double *vj = (double *)col_ptr(V, j), *vn = (double *)col_ptr(V, j + 1);
      hipLaunchKernelGGL(gm_prec_kernel, dim3(nb), dim3(BLK), 0, s, n, dinv, vj, z);
      SA_TRY(launch_spmv(A, z, vn, s));
      hipLaunchKernelGGL(gm_shift_kernel, dim3(nb), dim3(BLK), 0, s, n, vn, z, sigma);
      SA_TRY(launch_gs_column(V, nullptr, j + 1, 0.7071, s));
      SA_HIP(hipMemcpyAsync(h.data(), V.dbuf + (size_t)(j + 1) * V.m, sizeof(double) * (j + 2),
                            hipMemcpyDeviceToHost, s));
      SA_HIP(hipMemcpyAsync(V.st_host, V.st + 2, sizeof(GsState), hipMemcpyDeviceToHost, s));
      SA_HIP(hipStreamSynchronize(s));
      const bool stop = V.st_host->stop != 0;
      if (stop) h[j + 1] = 0.0;                        // the Krylov space is invariant
      for (int i = 0; i < j; i++) {                    // earlier rotations
        const double a = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = a;
      }
      const double r = std::hypot(h[j], h[j + 1]);
      SA_CHECK(r != 0.0, "shift-and-invert: GMRES breakdown in (A - sigma I) x = b (singular system)");
      cs[j] = h[j] / r;
      sn[j] = h[j + 1] / r;
      h[j] = r;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; i++) H[i + (size_t)j * m] = h[i];
      k = j + 1;
      if (std::fabs(g[j + 1]) <= target || stop) {
        done = true;
        break;
      }
    }
    for (int i = k - 1; i >= 0; i--) {                // H(0:k,0:k) y = g(0:k)
      double v = g[i];
      for (int c = i + 1; c < k; c++) v -= H[i + (size_t)c * m] * y[c];
      y[i] = v / H[i + (size_t)i * m];
    }
    SA_HIP(hipMemcpyAsync(V.dq, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    SA_TRY(launch_multvec(V, 0, k, 1.0, V.dq, 0.0, z, s));
    hipLaunchKernelGGL(gm_update_kernel, dim3(nb), dim3(BLK), 0, s, n, x, dinv, z);
    SA_HIP(hipGetLastError());
    SA_HIP(hipStreamSynchronize(s));                   // y is rewritten by the next cycle
    if (done) break;
  }
  if (its_out) *its_out = its;
  return 0;
}

}  // namespace

int ksp_solve(const Mat &A, KspWork &w, const double *b, double *x, double rtol, int max_it, int *its_out) {
  const int64_t n = w.n;
  hipStream_t s = stream();
  if (w.gmres && !w.direct) {
    SA_CHECK(n == A.n && w.buf, "solver workspace does not match the operator");
    if (n == 0) return 0;
    return gmres_solve(A, w, b, x, rtol, max_it, its_out);
  }
  if (w.direct) {        // PREONLY + LU: x = U^-1 L^-1 P b
    SA_CHECK(n == A.n && w.ab, "solver workspace does not match the operator");
    if (n == 0) return 0;
    SA_HIP(hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(band_solve_kernel, dim3(1), dim3(LU_T), 0, s, (int)n, w.kl, w.ku, w.ldab, w.ab, w.ipiv, x);
    SA_HIP(hipGetLastError());
    if (its_out) *its_out = 1;
    return 0;
  }
  SA_CHECK(n == A.n && w.buf, "solver workspace does not match the operator");
  double *r = w.buf, *rhat = r + n, *p = rhat + n, *v = p + n, *sv = v + n, *t = sv + n, *phat = t + n,
         *shat = phat + n, *dinv = w.jacobi ? shat + n : nullptr;
  double *part = w.buf + 9 * n, *scal = part + 2 * KSP_MAXG;
  const int G = w.G;
  const unsigned nb = (unsigned)std::max<int64_t>(1, (n + BLK - 1) / BLK);
  const double sigma = w.sigma;
  hipLaunchKernelGGL(ksp_init_kernel, dim3(nb), dim3(BLK), 0, s, n, b, r, rhat, p, v, x, scal);
  // ||b|| for the relative test
  double *spare = scal + S_NSCAL;   // two slots for the host checks
  hipLaunchKernelGGL(ksp_dot2_kernel, dim3(G), dim3(BLK), 0, s, n, b, b, nullptr, nullptr, part, G);
  hipLaunchKernelGGL(ksp_scalar_kernel, dim3(1), dim3(BLK), 0, s, part, G, 3, spare);
  SA_HIP(hipGetLastError());
  double hs[S_NSCAL + 2];
  SA_HIP(hipMemcpyAsync(hs, spare, sizeof(double), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(std::max(0.0, hs[0]));
  int its = 0;
  if (bnorm == 0.0) {   // x = 0 is the solution (set by the init kernel)
    if (its_out) *its_out = 0;
    return 0;
  }
  bool conv = false;
  while (!conv) {
    for (int q = 0; q < KSP_CHECK && its < max_it; q++, its++) {
      hipLaunchKernelGGL(ksp_dot2_kernel, dim3(G), dim3(BLK), 0, s, n, rhat, r, r, r, part, G);
      hipLaunchKernelGGL(ksp_scalar_kernel, dim3(1), dim3(BLK), 0, s, part, G, 0, scal);
      hipLaunchKernelGGL(ksp_p_kernel, dim3(nb), dim3(BLK), 0, s, n, r, p, v, dinv, phat, scal);
      SA_TRY(launch_spmv(A, phat, v, s));
      hipLaunchKernelGGL(ksp_shift_dot_kernel, dim3(G), dim3(BLK), 0, s, n, v, phat, sigma, rhat, nullptr, part, G);
      hipLaunchKernelGGL(ksp_scalar_kernel, dim3(1), dim3(BLK), 0, s, part, G, 1, scal);
      hipLaunchKernelGGL(ksp_s_kernel, dim3(nb), dim3(BLK), 0, s, n, r, v, dinv, sv, shat, scal);
      SA_TRY(launch_spmv(A, shat, t, s));
      hipLaunchKernelGGL(ksp_shift_dot_kernel, dim3(G), dim3(BLK), 0, s, n, t, shat, sigma, sv, t, part, G);
      hipLaunchKernelGGL(ksp_scalar_kernel, dim3(1), dim3(BLK), 0, s, part, G, 2, scal);
      hipLaunchKernelGGL(ksp_x_kernel, dim3(nb), dim3(BLK), 0, s, n, x, phat, shat, r, sv, t, scal);
    }
    SA_HIP(hipGetLastError());
    // the residual norm of the current x: one more r.r
    hipLaunchKernelGGL(ksp_dot2_kernel, dim3(G), dim3(BLK), 0, s, n, r, r, nullptr, nullptr, part, G);
    hipLaunchKernelGGL(ksp_scalar_kernel, dim3(1), dim3(BLK), 0, s, part, G, 3, spare);
    SA_HIP(hipMemcpyAsync(hs, scal, sizeof(double) * (S_NSCAL + 2), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    const double rnorm = std::sqrt(std::max(0.0, hs[S_NSCAL]));
    if (rnorm <= rtol * bnorm) conv = true;
    else if (hs[S_FLAG] != 0.0 || !std::isfinite(rnorm)) {
      set_error("shift-and-invert: BiCGStab breakdown in (A - sigma I) x = b");
      return 1;
    } else if (its >= max_it) {
      set_error("shift-and-invert: the linear solver did not converge within max_it iterations");
      return 1;
    }
  }
  if (its_out) *its_out = its;
  return 0;
}

}  // namespace sa
