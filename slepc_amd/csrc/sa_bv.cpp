// BV (basis vectors) host side: allocation and the Arnoldi/Lanczos sweep that
// chains the fused SpMV + Gram-Schmidt kernels without host round trips.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sa_ds.h"
#include "sa_internal.h"

namespace sa {


static int bv_alloc_impl(BV &V, int64_t n, int m, int prec);

// A basis of m columns of length n.  A failed allocation leaves V empty
// (BV()), so no caller can mistake it for a basis of the requested size.
int bv_alloc(BV &V, int64_t n, int m, int prec) {
  const int rc = bv_alloc_impl(V, n, m, prec);
  if (rc) bv_free(V);
  return rc;
}

static int bv_alloc_impl(BV &V, int64_t n, int m, int prec) {
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
  const int64_t maxG = gs_max_blocks(big), minc = gs_min_chunks(big);
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

// BVMatArnoldi with the shift-and-invert operator (STApply of STSINVERT:
// one linear solve per column, stsolve.c:16-60): V(:,j+1) = (A - sigma I)^-1
// V(:,j) by the device BiCGStab, then the Gram-Schmidt slots without SpMV.
int bv_sweep_sinvert(BV &V, const Mat &A, KspWork &w, double rtol, int max_it, int k, int *m, double eta,
                     double *beta, int *breakdown, int64_t *ksp_its) {
  SA_CHECK(k >= 0 && *m > k && *m < V.m, "invalid Arnoldi range");
  SA_CHECK(A.prec == SA_FP64 && V.prec == SA_FP64 && A.n == V.n, "shift-and-invert needs an FP64 operator and basis");
  hipStream_t s = stream();
  SA_TRY(launch_gs_reset(V, s));
  int stop = 0;
  for (int j = k; j < *m && !stop; j++) {
    int its = 0;
    SA_TRY(ksp_solve(A, w, (const double *)col_ptr(V, j), (double *)col_ptr(V, j + 1), rtol, max_it, &its));
    if (ksp_its) *ksp_its += its;
    SA_TRY(launch_gs_column(V, nullptr, j + 1, eta, s));
    SA_HIP(hipMemcpyAsync(V.st_host, V.st + 2, sizeof(GsState), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    stop = V.st_host->stop;
  }
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

// x -= C C^T x with classical Gram-Schmidt, `passes` times.  The reference
// keeps the deflation space as constraint columns of V (BVInsertConstraints,
// epssetup.c:412-418) that every BVOrthogonalizeColumn runs against together
// with the basis; one pass leaves components of order eps ||x|| /
// ||(I - CC^T) x||, the second removes them.
int bv_project_out(BV &C, int nc, void *x, hipStream_t s, int passes) {
  if (nc <= 0) return 0;
  SA_CHECK(C.prec == SA_FP64 && nc <= C.m, "invalid deflation space");
  for (int pass = 0; pass < passes; pass++) {
    SA_TRY(launch_dotvec(C, 0, nc, x, C.dq, s));
    SA_TRY(launch_multvec(C, 0, nc, -1.0, C.dq, 1.0, x, s));
  }
  return 0;
}

// BVMatArnoldi (bvkrylov.c:56-113) with constraints: each new column is the
// operator applied to the previous one, orthogonalized against the deflation
// space and then against V(:, 0:j+1) by the Gram-Schmidt slots (no SpMV in
// them).  The slots subtract V h, and V carries deflation components of
// order eps; divided by the new column's norm beta they would grow from
// column to column when beta is small, so one more pass against the space
// follows the slots (the reference's joint refinement over constraints and
// basis does the same).  With w the operator is the shift-and-invert solve,
// as in bv_sweep_sinvert.
int bv_sweep_deflated(BV &V, const Mat &A, KspWork *w, double rtol, int max_it, BV &C, int nc, int k, int *m,
                      double eta, double *beta, int *breakdown, int64_t *ksp_its) {
  SA_CHECK(k >= 0 && *m > k && *m < V.m, "invalid Arnoldi range");
  SA_CHECK(A.prec == SA_FP64 && V.prec == SA_FP64 && A.n == V.n && C.n == V.n,
           "the deflation space needs an FP64 operator and basis of its length");
  hipStream_t s = stream();
  SA_TRY(launch_gs_reset(V, s));
  int stop = 0;
  for (int j = k; j < *m && !stop; j++) {
    if (w) {
      int its = 0;
      SA_TRY(ksp_solve(A, *w, (const double *)col_ptr(V, j), (double *)col_ptr(V, j + 1), rtol, max_it, &its));
      if (ksp_its) *ksp_its += its;
    } else {
      SA_TRY(launch_spmv(A, col_ptr(V, j), col_ptr(V, j + 1), s));
    }
    SA_TRY(bv_project_out(C, nc, col_ptr(V, j + 1), s));
    SA_TRY(launch_gs_column(V, nullptr, j + 1, eta, s));
    SA_TRY(bv_project_out(C, nc, col_ptr(V, j + 1), s, 1));
    SA_HIP(hipMemcpyAsync(V.st_host, V.st + 2, sizeof(GsState), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    stop = V.st_host->stop;
  }
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

// ---- BVOrthogonalize (bvorthog.c:489-767, bvlapack.c:136-330) --------------
namespace {

// V(:, s:e) = V(:, lq:kq) Q(lq:kq, s:e) with Q in V.hq (leading dimension
// ldq): the restart GEMM, in place.
int apply_q(BV &V, int lq, int kq, int s_, int e, int ldq, hipStream_t s) {
  SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * (size_t)ldq * e, hipMemcpyHostToDevice, s));
  SA_TRY(launch_multinplace(V, lq, kq, s_, e, V.dq, ldq, s));
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

// BVOrthogonalize_BlockGS (bvorthog.c:489-502): R12 = V(:,0:l)^T V(:,l:k),
// V(:,l:k) -= V(:,0:l) R12, the update as the restart GEMM with Q = [-R12; I].
int block_gs(BV &V, int l, int k, std::vector<double> &R12, double *M_dev, hipStream_t s) {
  const int nk = k - l;
  SA_TRY(launch_bvdot(V, l, k, V, 0, l, M_dev, s));
  R12.assign((size_t)l * nk, 0.0);
  SA_HIP(hipMemcpy(R12.data(), M_dev, sizeof(double) * (size_t)l * nk, hipMemcpyDeviceToHost));
  std::fill(V.hq, V.hq + (size_t)k * k, 0.0);
  for (int j = 0; j < nk; j++) {
    for (int i = 0; i < l; i++) V.hq[i + (size_t)(l + j) * k] = -R12[i + (size_t)j * l];
    V.hq[(l + j) + (size_t)(l + j) * k] = 1.0;
  }
  return apply_q(V, 0, k, l, k, k, s);
}

// TSQR of more than TS_MAXK columns: panels of 64 columns, each first
// orthogonalised twice against the previous panels (block classical
// Gram-Schmidt, BVDot + the restart GEMM), then factored by the tile tree;
// at the end one sign fix from the first rows of the whole Q makes Q and R
// those of geqrf/orgqr on the whole block.  R_host: (k-l) x (k-l).
int tsqr_panels(BV &V, int l, int k, double *Rf, double *M_dev, hipStream_t s) {
  constexpr int P = 64;
  const int nk = k - l;
  std::fill(Rf, Rf + (size_t)nk * nk, 0.0);
  std::vector<double> R12, Rp, fp;
  for (int c0 = l; c0 < k; c0 += P) {
    const int w = std::min(P, k - c0), off = c0 - l;
    for (int pass = 0; pass < 2 && off > 0; pass++) {
      SA_TRY(launch_bvdot(V, c0, c0 + w, V, l, c0, M_dev, s));   // V(:,l:c0)^T V(:,c0:c0+w)
      R12.resize((size_t)off * w);
      SA_HIP(hipMemcpy(R12.data(), M_dev, sizeof(double) * R12.size(), hipMemcpyDeviceToHost));
      std::fill(V.hq, V.hq + (size_t)k * k, 0.0);
      for (int j = 0; j < w; j++) {
        for (int i = 0; i < off; i++) {
          V.hq[(l + i) + (size_t)(c0 + j) * k] = -R12[i + (size_t)j * off];
          Rf[i + (size_t)(off + j) * nk] += R12[i + (size_t)j * off];
        }
        V.hq[(c0 + j) + (size_t)(c0 + j) * k] = 1.0;
      }
      SA_TRY(apply_q(V, l, c0 + w, c0, c0 + w, k, s));
    }
    Rp.assign((size_t)w * w, 0.0);
    fp.assign(w, 1.0);
    SA_TRY(launch_tsqr(V, c0, c0 + w, true, Rp.data(), fp.data(), s));
    for (int j = 0; j < w; j++)
      for (int i = 0; i <= j; i++) Rf[(off + i) + (size_t)(off + j) * nk] = fp[i] * Rp[i + (size_t)j * w];
  }
  std::vector<double> Qt((size_t)nk * nk), f(nk);
  SA_TRY(bv_top_block(V, l, k, Qt.data()));
  tsqr_signs(Qt.data(), Rf, V.n, nk, f.data());
  for (int j = 0; j < nk; j++) {
    if (f[j] < 0.0) SA_TRY(launch_scale(V.prec, col_ptr(V, l + j), V.n, -1.0, s));
    for (int c = 0; c < nk; c++) Rf[j + (size_t)c * nk] *= f[j];
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

int bv_orthogonalize(BV &V, int method, double *R, int ldr) {
  SA_CHECK(method >= 0 && method <= 4,
           "unknown block orthogonalization method (0 gs, 1 chol, 2 svqb, 3 tsqr, 4 tsqrchol)");
  const int l = V.l, k = V.k, nk = k - l;
  SA_CHECK(!R || ldr >= k, "R is smaller than the number of active columns");
  if (nk <= 0) return 0;
  hipStream_t s = stream();
  if (method == 0) {   // BVOrthogonalize_GS (bvorthog.c:507-548): column by column
    std::vector<double> h(V.m);
    for (int j = l; j < k; j++) {
      double nrm = 0.0;
      int lind = 0;
      SA_TRY(bv_orthonormalize(V, j, 0.7071, &nrm, &lind));
      SA_CHECK(nrm != 0.0, "Breakdown in BVOrthogonalize due to a linearly dependent column");
      if (R) {
        SA_HIP(hipMemcpy(h.data(), V.dbuf + (size_t)j * V.m, sizeof(double) * V.m, hipMemcpyDeviceToHost));
        for (int i = 0; i < k; i++) R[i + (size_t)j * ldr] = i < j ? h[i] : (i == j ? nrm : 0.0);
      }
    }
    return 0;
  }
  SA_CHECK(lapack_ready(), "no LAPACK library found for the block orthogonalization");
  double *M_dev = nullptr;
  SA_HIP(hipMalloc(&M_dev, sizeof(double) * (size_t)k * k));
  std::vector<double> R12, G((size_t)nk * nk);
  int rc = l ? block_gs(V, l, k, R12, M_dev, s) : 0;
  if (method >= 3) {   // BVOrthogonalize_TSQR / _TSQRCHOL (bvorthog.c:615-659)
    std::vector<double> Rt(G.size()), f(nk, 1.0);
    if (!rc && nk > 64) {   // panels; TSQRCHOL's Q = V2 inv(R) is the same Q, formed here directly
      rc = tsqr_panels(V, l, k, Rt.data(), M_dev, s);
      (void)hipFree(M_dev);
      if (rc) return rc;
      if (R)
        for (int j = 0; j < nk; j++)
          for (int i = 0; i < k; i++) R[i + (size_t)(l + j) * ldr] = i < l ? R12[i + (size_t)j * l] : Rt[(i - l) + (size_t)j * nk];
      return 0;
    }
    (void)hipFree(M_dev);
    if (rc) return rc;
    SA_TRY(launch_tsqr(V, l, k, method == 3, Rt.data(), f.data(), s));
    if (method == 4) {   // R only (bvlapack.c:510-560), then V2 = V2 inv(R) (bvorthog.c:654-655)
      std::vector<double> Si(Rt), Vt(G.size()), Qt(G.size(), 0.0);
      SA_CHECK(la_trtri_upper(nk, Si.data(), nk) == 0, "BVOrthogonalize: triangular inverse (trtri) failed");
      SA_TRY(bv_top_block(V, l, k, Vt.data()));
      for (int j = 0; j < nk; j++)           // first rows of V2 inv(R), for the signs
        for (int a = 0; a <= j; a++)
          for (int i = 0; i < nk; i++) Qt[i + (size_t)j * nk] += Vt[i + (size_t)a * nk] * Si[a + (size_t)j * nk];
      tsqr_signs(Qt.data(), Rt.data(), V.n, nk, f.data());
      std::fill(V.hq, V.hq + (size_t)k * k, 0.0);
      for (int j = 0; j < nk; j++)
        for (int i = 0; i < nk; i++) V.hq[(l + i) + (size_t)(l + j) * k] = Si[i + (size_t)j * nk] * f[j];
      SA_TRY(apply_q(V, l, k, l, k, k, s));
    }
    if (R)
      for (int j = 0; j < nk; j++)
        for (int i = 0; i < k; i++)
          R[i + (size_t)(l + j) * ldr] = i < l ? R12[i + (size_t)j * l] : f[i - l] * Rt[(i - l) + (size_t)j * nk];
    return 0;
  }
  if (!rc) rc = launch_bvdot(V, l, k, V, l, k, M_dev, s);            // G = V2^T V2
  if (!rc && hipMemcpy(G.data(), M_dev, sizeof(double) * G.size(), hipMemcpyDeviceToHost) != hipSuccess) rc = 1;
  (void)hipFree(M_dev);
  if (rc == 1) set_error("HIP error in the block orthogonalization");
  if (rc) return rc;
  std::vector<double> Rf(G.size(), 0.0), S(G.size(), 0.0);
  if (method == 1) {   // BVOrthogonalize_Chol: R = chol(G), V2 = V2 inv(R) (bvlapack.c:136-203)
    Rf = G;
    int info = la_potrf_upper(nk, Rf.data(), nk);
    if (info) {        // retry on the diagonally perturbed matrix
      Rf = G;
      for (int i = 0; i < nk; i++) Rf[i + (size_t)i * nk] += 50.0 * DBL_EPSILON;
      info = la_potrf_upper(nk, Rf.data(), nk);
    }
    SA_CHECK(info == 0, "BVOrthogonalize: Cholesky factorization (potrf) failed");
    for (int j = 0; j < nk; j++)
      for (int i = j + 1; i < nk; i++) Rf[i + (size_t)j * nk] = 0.0;
    S = Rf;
    info = la_trtri_upper(nk, S.data(), nk);
    SA_CHECK(info == 0, "BVOrthogonalize: triangular inverse (trtri) failed");
  } else {             // BVOrthogonalize_SVQB (bvlapack.c:261-330)
    std::vector<double> D(nk), w(nk), U(G.size());
    for (int i = 0; i < nk; i++) D[i] = 1.0 / std::sqrt(G[i + (size_t)i * nk]);
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < nk; i++) U[i + (size_t)j * nk] = G[i + (size_t)j * nk] * D[i] * D[j];
    const int info = la_syev_lower(nk, U.data(), nk, w.data());
    SA_CHECK(info == 0, "BVOrthogonalize: symmetric eigensolver (syev) failed");
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < nk; i++) {
        S[i + (size_t)j * nk] = U[i + (size_t)j * nk] * D[i] / std::sqrt(w[j]);      // D U Lambda^{-1/2}
        Rf[i + (size_t)j * nk] = U[j + (size_t)i * nk] * std::sqrt(w[i]) / D[j];    // its inverse
      }
  }
  std::fill(V.hq, V.hq + (size_t)k * k, 0.0);
  for (int j = 0; j < nk; j++)
    for (int i = 0; i < nk; i++) V.hq[(l + i) + (size_t)(l + j) * k] = S[i + (size_t)j * nk];
  SA_TRY(apply_q(V, l, k, l, k, k, s));
  if (R)
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < k; i++)
        R[i + (size_t)(l + j) * ldr] = i < l ? R12[i + (size_t)j * l] : Rf[(i - l) + (size_t)j * nk];
  return 0;
}

}  // namespace sa
