// Modified Gram-Schmidt of one basis column against a chosen set of columns
// (BV_ORTHOG_MGS: BVOrthogonalizeMGS1, src/sys/classes/bv/interface/
// bvorthog.c:52-85, with refinement IFNEEDED, bvorthog.c:145-215), as used by
// EPSLanczos with local reorthogonalization (lanczos.c:58-62,89-118): each
// step needs the coefficient of the previous column before it can start, so
// one column costs two stream-ordered launches, a fused "subtract the previous
// column, partial dot with the next" pass over n rows and a one-workgroup
// finish that leaves the coefficient in device memory for the next pass.  The
// host reads the two norms once per pass.  Partial sums are reduced in a
// fixed order, so the result does not depend on timing.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

namespace {

__device__ __forceinline__ double block_total(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

// w -= (*coef) vp (when vp is set), then part[b] = partial sum of w . vn, or
// of w . w when vn is null.  The dot uses the stored (rounded) w.
template <typename T>
__global__ __launch_bounds__(BLK) void mgs_step_kernel(T *w, const T *vp, const double *coef, const T *vn,
                                                       int64_t n, double *part) {
  __shared__ double red[BLK];
  const double c = vp ? *coef : 0.0;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    T x = w[i];
    if (vp) {
      x = (T)((double)x - c * (double)vp[i]);
      w[i] = x;
    }
    const double xd = (double)x;
    acc += vn ? xd * (double)vn[i] : xd * xd;
  }
  const double s = block_total(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// s = sum of part[0..np); dot mode (h set): h += s and *slot = s (the
// coefficient of the next subtraction); norm mode: *slot = sqrt(s), and
// hn = sqrt(s) when hn is set.
__global__ __launch_bounds__(BLK) void mgs_finish_kernel(const double *part, int np, double *h, double *slot,
                                                         double *hn) {
  __shared__ double red[BLK];
  double a = 0.0;
  for (int i = threadIdx.x; i < np; i += BLK) a += part[i];
  const double s = block_total(a, red);
  if (threadIdx.x == 0) {
    if (h) {
      *h += s;
      *slot = s;
    } else {
      const double r = safe_sqrt(s);
      *slot = r;
      if (hn) *hn = r;
    }
  }
}

}  // namespace

// One MGS pass: the vector w (W.n rows, W's precision) against the columns
// cols[0..nc) of W in order, each coefficient added into h[cols[q]]; onrm
// (device) gets ||w|| before, nrm (device) ||w|| after; hn (device, or null)
// also gets ||w|| after.  Partial sums in W's workspace.
static int mgs_pass(const BV &W, void *w, const int *cols, int nc, double *h, double *slot, double *onrm,
                    double *nrm, double *hn, hipStream_t s) {
  const int64_t n = W.n;
  const int64_t cap = 2 * (int64_t)W.ws.maxG * (GS_MAXC + 1);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>({1024, (n + BLK - 1) / BLK, cap}));
  double *part = W.ws.part;
  auto step = [&](int prev, int next) {
    if (W.prec == SA_FP64) {
      const double *vp = prev >= 0 ? (const double *)col_ptr(W, prev) : nullptr;
      const double *vn = next >= 0 ? (const double *)col_ptr(W, next) : nullptr;
      hipLaunchKernelGGL(mgs_step_kernel<double>, dim3(grid), dim3(BLK), 0, s, (double *)w, vp, slot, vn, n, part);
    } else {
      const float *vp = prev >= 0 ? (const float *)col_ptr(W, prev) : nullptr;
      const float *vn = next >= 0 ? (const float *)col_ptr(W, next) : nullptr;
      hipLaunchKernelGGL(mgs_step_kernel<float>, dim3(grid), dim3(BLK), 0, s, (float *)w, vp, slot, vn, n, part);
    }
  };
  if (onrm) {
    step(-1, -1);
    hipLaunchKernelGGL(mgs_finish_kernel, dim3(1), dim3(BLK), 0, s, part, grid, nullptr, onrm, nullptr);
  }
  int prev = -1;
  for (int q = 0; q < nc; q++) {
    step(prev, cols[q]);
    hipLaunchKernelGGL(mgs_finish_kernel, dim3(1), dim3(BLK), 0, s, part, grid, h + cols[q], slot, nullptr);
    prev = cols[q];
  }
  step(prev, -1);
  hipLaunchKernelGGL(mgs_finish_kernel, dim3(1), dim3(BLK), 0, s, part, grid, nullptr, nrm, hn);
  SA_HIP(hipGetLastError());
  return 0;
}

// BVOrthogonalizeGS with MGS (bvorthog.c:145-217) and the basis's refinement:
// IFNEEDED passes while ||w|| < eta ||w_before||, at most three; NEVER one
// pass, lindep only for a zero norm; ALWAYS two passes, the second measured.
// Returns the final norm and the lindep test.
static int mgs_refined(const BV &W, void *w, const int *cols, int nc, double *h, double eta, double *norm,
                       int *lindep, int64_t *passes) {
  hipStream_t s = stream();
  double *sc = W.dq + (size_t)W.m * W.m;              // scalar slots past the staging area
  double *slot = sc, *nrm_d = sc + 1, *onrm_d = sc + 2;
  double hs[2];
  if (W.refine == GS_REFINE_ALWAYS) {                 // bvorthog.c:195-199: a first pass without norms
    SA_TRY(mgs_pass(W, w, cols, nc, h, slot, nullptr, nrm_d, nullptr, s));
    (*passes)++;
  }
  SA_TRY(mgs_pass(W, w, cols, nc, h, slot, onrm_d, nrm_d, nullptr, s));
  SA_HIP(hipMemcpyAsync(hs, nrm_d, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  (*passes)++;
  double nrm = hs[0], onrm = hs[1];
  if (W.refine == GS_REFINE_NEVER) {                  // bvorthog.c:187-193: lindep for a zero norm only
    *norm = nrm;
    *lindep = !(nrm != 0.0);
    return 0;
  }
  int l = W.refine == GS_REFINE_ALWAYS ? 3 : 1;       // ALWAYS: no further pass
  while (l < 3 && nrm != 0.0 && std::fabs(nrm) < eta * std::fabs(onrm)) {
    l++;
    onrm = nrm;
    SA_TRY(mgs_pass(W, w, cols, nc, h, slot, nullptr, nrm_d, nullptr, s));
    SA_HIP(hipMemcpyAsync(hs, nrm_d, sizeof(double), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    (*passes)++;
    nrm = hs[0];
  }
  *norm = nrm;
  *lindep = !(nrm != 0.0 && std::fabs(nrm) >= eta * std::fabs(onrm));
  return 0;
}

int bv_mgs_vec(BV &W, int nc, void *w, double eta, double *norm, int *lindep) {
  SA_CHECK(nc >= 0 && nc <= W.m, "column count out of range");
  std::vector<int> cols(nc);
  for (int q = 0; q < nc; q++) cols[q] = q;
  SA_HIP(hipMemsetAsync(W.dbuf, 0, sizeof(double) * (size_t)std::max(nc, 1), stream()));
  return mgs_refined(W, w, cols.data(), nc, W.dbuf, eta, norm, lindep, &W.passes_done);
}

int bv_mgs_column(BV &V, int j, const int *cols, int nc, double eta, double *norm, int *lindep) {
  SA_CHECK(j >= 0 && j < V.m && nc >= 0, "column index out of range");
  for (int q = 0; q < nc; q++) SA_CHECK(cols[q] >= 0 && cols[q] < j, "orthogonalization column out of range");
  hipStream_t s = stream();
  double *h = V.dbuf + (size_t)j * V.m;               // coefficients of column j
  SA_HIP(hipMemsetAsync(h, 0, sizeof(double) * (size_t)(j + 1), s));   // BV_CleanCoefficients
  double nrm = 0.0;
  int ld = 0;
  SA_TRY(mgs_refined(V, col_ptr(V, j), cols, nc, h, eta, &nrm, &ld, &V.passes_done));
  const double hj = ld ? 0.0 : nrm;
  SA_HIP(hipMemcpyAsync(h + j, &hj, sizeof(double), hipMemcpyHostToDevice, s));
  SA_HIP(hipStreamSynchronize(s));
  if (norm) *norm = nrm;
  if (lindep) *lindep = ld;
  return 0;
}

}  // namespace sa
