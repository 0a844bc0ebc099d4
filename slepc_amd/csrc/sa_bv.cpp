// BV (basis vectors) host side: allocation and the Arnoldi/Lanczos sweep that
// chains the fused SpMV + Gram-Schmidt kernels without host round trips.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "sa_internal.h"

namespace sa {

int launch_gs_reset(BV &V, hipStream_t s);

int bv_alloc(BV &V, int64_t n, int m, int prec) {
  SA_CHECK(n >= 0 && m > 0, "invalid BV sizes");
  SA_CHECK(prec == SA_FP64 || prec == SA_FP32, "invalid precision");
  V.n = n;
  V.m = m;
  V.prec = prec;
  V.ld = std::max<int64_t>(64, (n + 63) / 64 * 64);
  V.l = 0;
  V.k = m;
  const size_t es = prec_size(prec);
  SA_HIP(hipMalloc(&V.data, (size_t)V.ld * m * es));
  SA_HIP(hipMemsetAsync(V.data, 0, (size_t)V.ld * m * es, stream()));
  SA_HIP(hipMalloc(&V.dbuf, sizeof(double) * (size_t)m * m));
  SA_HIP(hipMemsetAsync(V.dbuf, 0, sizeof(double) * (size_t)m * m, stream()));
  SA_HIP(hipMalloc(&V.st, sizeof(GsState) * GS_NSTATE));
  SA_HIP(hipMemsetAsync(V.st, 0, sizeof(GsState) * GS_NSTATE, stream()));
  SA_HIP(hipHostMalloc((void **)&V.st_host, sizeof(GsState)));
  // Partition rows into workgroup ranges of whole 64-row chunks.
  const int64_t chunks = std::max<int64_t>(1, (n + 63) / 64);
  // at most gs_max_blocks() workgroups (1024, or 4096 for a basis larger than
  // the last-level cache, whose slot kernels then fit more workgroups per CU)
  // and at least gs_min_chunks() chunks each, so the reductions of the
  // partial sums stay short
  const bool big = V.ld * m * (int64_t)es > GS_NT_BYTES;   // the slot kernels' NT flag
  const int64_t maxG = gs_max_blocks(big), minc = gs_min_chunks();
  int64_t G = std::min(maxG, std::max<int64_t>(1, (chunks + minc - 1) / minc));
  const int64_t cpb = (chunks + G - 1) / G;
  G = (chunks + cpb - 1) / cpb;
  V.G = (int)G;
  V.rows_per_block = (int)(cpb * 64);
  V.ws.maxG = (int)G;
  SA_HIP(hipMalloc(&V.ws.part, sizeof(double) * 2 * (size_t)G * (GS_MAXC + 1)));
  SA_HIP(hipMalloc(&V.ws.dots, sizeof(double) * 2 * (GS_MAXC + 1)));
  // the reductions read every partial row even after a slot that produced none
  SA_HIP(hipMemsetAsync(V.ws.part, 0, sizeof(double) * 2 * (size_t)G * (GS_MAXC + 1), stream()));
  SA_HIP(hipMemsetAsync(V.ws.dots, 0, sizeof(double) * 2 * (GS_MAXC + 1), stream()));
  SA_HIP(hipMalloc(&V.dq, sizeof(double) * (size_t)m * m + 64 * sizeof(double)));
  SA_HIP(hipHostMalloc((void **)&V.hq, sizeof(double) * (size_t)m * m + 64 * sizeof(double)));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

void bv_free(BV &V) {
  if (V.data) (void)hipFree(V.data);
  if (V.dbuf) (void)hipFree(V.dbuf);
  if (V.st) (void)hipFree(V.st);
  if (V.st_host) (void)hipHostFree(V.st_host);
  if (V.ws.part) (void)hipFree(V.ws.part);
  if (V.ws.dots) (void)hipFree(V.ws.dots);
  if (V.dq) (void)hipFree(V.dq);
  if (V.hq) (void)hipHostFree(V.hq);
  V = BV();
}

// BVMatArnoldi/BVMatLanczos (bvkrylov.c:56-113,165-226): for j = k..m-1,
// V(:,j+1) = A V(:,j) followed by BVOrthonormalizeColumn(V,j+1).  All steps
// are enqueued back to back; the device state machine applies the DGKS rule
// and stops the sweep at a breakdown.  One synchronisation at the end reads
// the state (breakdown column and the last norm).  (Replaying the sweep as a
// captured HIP graph measured no faster: the kernels run back to back with
// no gaps, so the time is inside them.)
int bv_sweep(BV &V, const Mat &A, int k, int *m, double eta, double *beta, int *breakdown) {
  SA_CHECK(k >= 0 && *m > k && *m < V.m, "invalid Arnoldi range");
  SA_CHECK(A.prec == V.prec && A.n == V.n, "matrix and basis sizes or precisions differ");
  hipStream_t s = stream();
  SA_TRY(launch_gs_reset(V, s));
  for (int j = k; j < *m; j++) SA_TRY(launch_gs_column(V, &A, j + 1, eta, s));
  SA_HIP(hipMemcpyAsync(V.st_host, V.st + 2, sizeof(GsState), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  const GsState &st = *V.st_host;
  if (st.stop) {
    *m = st.stop_col;
    *breakdown = 1;
  } else {
    *breakdown = 0;
  }
  *beta = st.beta;
  V.passes_done = st.passes_total;
  return 0;
}

// BVOrthonormalizeColumn(V,j,PETSC_FALSE,&norm,&lindep) (bvorthog.c:380-427)
int bv_orthonormalize(BV &V, int j, double eta, double *norm, int *lindep) {
  hipStream_t s = stream();
  SA_TRY(launch_gs_reset(V, s));
  SA_TRY(launch_gs_column(V, nullptr, j, eta, s));
  SA_HIP(hipMemcpyAsync(V.st_host, V.st + 2, sizeof(GsState), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  if (norm) *norm = V.st_host->nrm;
  if (lindep) *lindep = V.st_host->lindep;
  V.passes_done = V.st_host->passes_total;
  return 0;
}

}  // namespace sa
