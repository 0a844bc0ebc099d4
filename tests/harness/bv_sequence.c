/*
 * C harness for the drop-in boundary: drives libslepc_amd.so through its C ABI
 * (include/slepc_amd.h) with the call sequences SLEPc's own BV and Mat code
 * issues through the ops table of a registered BV type, and checks every
 * result against plain C loops on the host.  No Python, no ctypes.
 *
 *  1. BVOrthonormalizeColumn (src/sys/classes/bv/interface/bvorthog.c:380-427)
 *     as BVOrthogonalizeGS (bvorthog.c:145-217) runs it with the gramschmidt
 *     slot (bvorthog.c:134): coefficients cleaned, one CGS1 pass per hook
 *     call (sa_bv_gs_pass), the DGKS "if needed" loop (eta 0.7071, at most 3
 *     passes) and the lindep rule on the host, the norm stored at row j,
 *     then BVScaleColumn.  Compared with the fused kernels
 *     (sa_bv_orthonormalize_column) on a copy of the same columns, and with
 *     Q^T Q = I and V0 = Q R on the host; a planted dependent last column
 *     must leave only rounding noise on both paths.
 *  2. ops->dotvec / ops->multvec with q == NULL (the coefficient-buffer
 *     convention, bvmat.c:38-53,110-130): dots into the buffer, then
 *     y = y - V c from the buffer.
 *  3. ops->scale over the active columns (j < 0).
 *  4. ops->multinplace with a host Q whose leading dimension exceeds the
 *     active columns (MATSEQDENSE lda, bvmat.c:55-71).
 *  5. BVMult (with Q and as BVAXPY), BVCopy, BVResize (copy), BVDot,
 *     BVMatMult, MatMult and BVMatArnoldi's relation A V = V H + beta v e^T.
 *
 * Exit status 0 when every check passes; each check prints one line.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/slepc_amd.h"

static int failures = 0;

#define CALL(x)                                                                      \
  do {                                                                               \
    if ((x) != 0) {                                                                  \
      fprintf(stderr, "FAILED call %s: %s\n", #x, sa_last_error());                  \
      exit(2);                                                                       \
    }                                                                                \
  } while (0)

static void check(const char *name, double err, double tol) {
  int ok = err <= tol;
  printf("%s %-44s err %.3e (tol %.1e)\n", ok ? "ok  " : "FAIL", name, err, tol);
  if (!ok) failures++;
}

/* deterministic inputs: a 64-bit LCG mapped to [-1, 1) */
static uint64_t lcg = 0x243F6A8885A308D3ULL;
static double urand(void) {
  lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(lcg >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

static void upload(sa_bv V, int j, const double *x) { CALL(sa_bv_set_column(V, j, x)); }
static void download(sa_bv V, int j, double *x) { CALL(sa_bv_get_column(V, j, x)); }

static double maxabs_diff(const double *a, const double *b, int64_t n) {
  double e = 0.0;
  for (int64_t i = 0; i < n; i++) e = fmax(e, fabs(a[i] - b[i]));
  return e;
}

/* 2-D 5-point Laplacian, row-major grid, ascending columns */
static int64_t lap2d(int nx, int ny, int32_t *rp, int32_t *ci, double *v) {
  int64_t nz = 0;
  for (int r = 0; r < nx * ny; r++) {
    int i = r / nx, j = r % nx;
    rp[r] = (int32_t)nz;
    if (i > 0) { ci[nz] = r - nx; v[nz++] = -1.0; }
    if (j > 0) { ci[nz] = r - 1; v[nz++] = -1.0; }
    ci[nz] = r; v[nz++] = 4.0;
    if (j < nx - 1) { ci[nz] = r + 1; v[nz++] = -1.0; }
    if (i < ny - 1) { ci[nz] = r + nx; v[nz++] = -1.0; }
  }
  rp[nx * ny] = (int32_t)nz;
  return nz;
}

static void spmv_host(int n, const int32_t *rp, const int32_t *ci, const double *v, const double *x, double *y) {
  for (int r = 0; r < n; r++) {
    double s = 0.0;
    for (int32_t p = rp[r]; p < rp[r + 1]; p++) s += v[p] * x[ci[p]];
    y[r] = s;
  }
}

int main(void) {
  const int n = 5003, m = 12;
  const double eta = 0.7071;
  CALL(sa_init(0));
  double *V0 = malloc(sizeof(double) * n * m), *Q = malloc(sizeof(double) * n * m);
  double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < (int64_t)n * m; i++) V0[i] = urand();
  /* the last column planted as a combination of columns 0..6 */
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int c = 0; c < 7; c++) s += (c + 1) * 0.25 * V0[i + (int64_t)c * n];
    V0[i + (int64_t)(m - 1) * n] = s;
  }

  /* ---- 1. BVOrthonormalizeColumn through the gramschmidt slot ---------- */
  sa_bv A, B;
  CALL(sa_bv_create(n, m, SA_FP64, &A));
  CALL(sa_bv_create(n, m, SA_FP64, &B));
  for (int j = 0; j < m; j++) { upload(A, j, V0 + (int64_t)j * n); upload(B, j, V0 + (int64_t)j * n); }
  double *Hs = calloc((size_t)m * m, sizeof(double)), *Hf = calloc((size_t)m * m, sizeof(double));
  double c[128];
  int lindep_s[64], lindep_f[64];
  double norm_s[64], norm_f[64];
  CALL(sa_bv_set_active_columns(A, 0, m));
  for (int j = 0; j < m; j++) {
    double *h = Hs + (int64_t)j * m, onrm = 0.0, nrm = 0.0;
    for (int i = 0; i <= j; i++) h[i] = 0.0;                 /* BV_CleanCoefficients */
    CALL(sa_bv_gs_pass(A, j, h, c, &onrm, &nrm));             /* BVOrthogonalizeGS1 */
    int l = 1;
    while (l < 3 && nrm != 0.0 && fabs(nrm) < eta * fabs(onrm)) {
      l++;
      CALL(sa_bv_gs_pass(A, j, h, c, &onrm, &nrm));
    }
    int lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
    h[j] = lindep ? 0.0 : nrm;                                /* BV_SetValue */
    lindep_s[j] = lindep;
    norm_s[j] = nrm;
    if (!lindep) CALL(sa_bv_scale_column(A, j, 1.0 / nrm)); /* BVScaleColumn */
    CALL(sa_bv_orthonormalize_column(B, j, eta, &norm_f[j], &lindep_f[j]));
  }
  CALL(sa_bv_get_coefficients(B, 0, m, Hf));
  double eh = 0.0, en = 0.0;
  int lmis = 0;
  for (int j = 0; j < m; j++) {
    lmis += lindep_s[j] != lindep_f[j];
    if (lindep_s[j] || j == m - 1) continue;   /* the planted column's norm is rounding noise */
    en = fmax(en, fabs(norm_s[j] - norm_f[j]) / norm_f[j]);
    for (int i = 0; i <= j; i++)   /* rows of a dependent column differ by its (tiny) leftover */
      if (!lindep_s[i] || i == j) eh = fmax(eh, fabs(Hs[i + (int64_t)j * m] - Hf[i + (int64_t)j * m]));
  }
  check("orthonormalize: hook passes vs fused, coeffs", eh, 1e-11);
  check("orthonormalize: hook passes vs fused, norms", en, 1e-12);
  check("orthonormalize: lindep flags agree", (double)lmis, 0.0);
  /* the planted combination leaves a residual at rounding level; with the
     relative DGKS criterion the last pass compares that residual with its
     own norm, so lindep depends on the noise: only its size is checked */
  {
    double v0n = 0.0;
    for (int i = 0; i < n; i++) v0n += V0[i + (int64_t)(m - 1) * n] * V0[i + (int64_t)(m - 1) * n];
    v0n = sqrt(v0n);
    check("orthonormalize: planted column leaves noise (hook)", norm_s[m - 1] / v0n, 1e-13);
    check("orthonormalize: planted column leaves noise (fused)", norm_f[m - 1] / v0n, 1e-13);
  }
  for (int j = 0; j < m; j++) download(A, j, Q + (int64_t)j * n);
  double eo = 0.0, er = 0.0;
  for (int a = 0; a < m; a++) {
    if (lindep_s[a]) continue;
    for (int b = 0; b < m; b++) {
      if (lindep_s[b]) continue;
      double s = 0.0;
      for (int i = 0; i < n; i++) s += Q[i + (int64_t)a * n] * Q[i + (int64_t)b * n];
      eo = fmax(eo, fabs(s - (a == b ? 1.0 : 0.0)));
    }
  }
  for (int j = 0; j < m; j++) {            /* V0(:,j) = sum_i H(i,j) Q(:,i) for independent j */
    if (lindep_s[j]) continue;
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int q = 0; q <= j; q++)
        if (!lindep_s[q]) s += Hs[q + (int64_t)j * m] * Q[i + (int64_t)q * n];
      er = fmax(er, fabs(s - V0[i + (int64_t)j * n]));
    }
  }
  check("orthonormalize: ||Q^T Q - I||_max", eo, 1e-13);
  check("orthonormalize: V0 = Q R", er, 1e-12);

  /* replace the dependent column by a fresh orthonormal one for the rest */
  for (int i = 0; i < n; i++) x[i] = urand();
  upload(A, m - 1, x);
  { double nn; int ld; CALL(sa_bv_orthonormalize_column(A, m - 1, eta, &nn, &ld)); check("replacement column independent", (double)ld, 0.0); }
  for (int j = 0; j < m; j++) download(A, j, Q + (int64_t)j * n);

  /* ---- 2. dotvec / multvec with q == NULL (coefficient buffer) ---------- */
  double *xd = NULL, *yd = NULL;
  CALL(sa_malloc((void **)&xd, sizeof(double) * n));
  CALL(sa_malloc((void **)&yd, sizeof(double) * n));
  for (int i = 0; i < n; i++) { x[i] = urand(); y[i] = x[i]; }
  CALL(sa_memcpy_h2d(xd, x, sizeof(double) * n));
  CALL(sa_memcpy_h2d(yd, y, sizeof(double) * n));
  const int l0 = 2, k0 = 9;
  CALL(sa_bv_set_active_columns(A, l0, k0));
  CALL(sa_bv_dotvec(A, xd, NULL));
  CALL(sa_bv_multvec(A, -1.0, 1.0, yd, NULL));
  double buf0[128];
  CALL(sa_bv_get_coefficients(A, 0, 1, buf0));
  double ed = 0.0;
  for (int q = l0; q < k0; q++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += Q[i + (int64_t)q * n] * x[i];
    ed = fmax(ed, fabs(s - buf0[q - l0]));
    for (int i = 0; i < n; i++) y[i] -= s * Q[i + (int64_t)q * n];
  }
  CALL(sa_memcpy_d2h(z, yd, sizeof(double) * n));
  check("dotvec q=NULL: buffer column 0 (from index 0)", ed, 1e-12);
  check("multvec q=NULL: y - V c", maxabs_diff(z, y, n), 1e-12);

  /* ---- 3. scale of the active columns ----------------------------------- */
  CALL(sa_bv_scale_column(A, -1, -2.5));
  double es = 0.0;
  for (int q = 0; q < m; q++) {
    download(A, q, z);
    double f = (q >= l0 && q < k0) ? -2.5 : 1.0;
    for (int i = 0; i < n; i++) es = fmax(es, fabs(z[i] - f * Q[i + (int64_t)q * n]));
  }
  check("scale j<0: active columns only", es, 1e-14);
  CALL(sa_bv_scale_column(A, -1, -0.4));                   /* back to Q */

  /* ---- 4. multinplace with a host Q of leading dimension k + 3 ---------- */
  const int ldq = m + 3, s0 = 3, e0 = 8;
  double *Qm = malloc(sizeof(double) * ldq * m);
  for (int i = 0; i < ldq * m; i++) Qm[i] = urand();
  CALL(sa_bv_set_active_columns(A, 1, 10));
  for (int j = 0; j < m; j++) download(A, j, Q + (int64_t)j * n);
  CALL(sa_bv_multinplace(A, Qm, ldq, s0, e0));
  double em = 0.0;
  for (int j = s0; j < e0; j++) {
    download(A, j, z);
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int q = 1; q < 10; q++) s += Q[i + (int64_t)q * n] * Qm[q + (int64_t)j * ldq];
      em = fmax(em, fabs(s - z[i]));
    }
  }
  check("multinplace: V(:,s:e) = V(:,l:k) Q(l:k,s:e), ldq > k", em, 1e-12);
  for (int j = 0; j < m; j++) download(A, j, Q + (int64_t)j * n);

  /* ---- 5. BVMult, BVCopy, BVResize, BVDot, BVMatMult, MatMult, Arnoldi -- */
  sa_bv Y;
  CALL(sa_bv_create(n, m, SA_FP64, &Y));
  double *Y0 = malloc(sizeof(double) * n * m);
  for (int64_t i = 0; i < (int64_t)n * m; i++) Y0[i] = urand();
  for (int j = 0; j < m; j++) upload(Y, j, Y0 + (int64_t)j * n);
  CALL(sa_bv_set_active_columns(A, 2, 7));                /* X.l = 2, X.k = 7 */
  CALL(sa_bv_set_active_columns(Y, 4, 9));                /* Y.l = 4, Y.k = 9 */
  const int ldm = m + 1;
  double *Mq = malloc(sizeof(double) * ldm * m);
  for (int i = 0; i < ldm * m; i++) Mq[i] = urand();
  CALL(sa_bv_mult(Y, 0.5, -1.5, A, Mq, ldm));
  double eb = 0.0;
  for (int j = 4; j < 9; j++) {
    download(Y, j, z);
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int q = 2; q < 7; q++) s += Q[i + (int64_t)q * n] * Mq[q + (int64_t)j * ldm];
      eb = fmax(eb, fabs(-1.5 * Y0[i + (int64_t)j * n] + 0.5 * s - z[i]));
    }
  }
  check("BVMult: Y = beta Y + alpha X Q(X.l:X.k, Y.l:Y.k)", eb, 1e-12);
  CALL(sa_bv_mult(Y, 2.0, 0.0, A, NULL, 0));             /* BVAXPY, beta 0 */
  double ea = 0.0;
  for (int j = 4; j < 9; j++) {
    download(Y, j, z);
    for (int i = 0; i < n; i++) ea = fmax(ea, fabs(2.0 * Q[i + (int64_t)(j - 2) * n] - z[i]));
  }
  check("BVMult without Q: Y = alpha X", ea, 1e-14);
  CALL(sa_bv_copy(A, Y));
  double ec = 0.0;
  for (int j = 4; j < 9; j++) {
    download(Y, j, z);
    ec = fmax(ec, maxabs_diff(z, Q + (int64_t)(j - 2) * n, n));
  }
  check("BVCopy: Y(:,Y.l+j) = X(:,X.l+j)", ec, 0.0);
  CALL(sa_bv_resize(Y, m + 5, 1));
  { int64_t nn; int mm, pr; int64_t ld; CALL(sa_bv_get_info(Y, &nn, &mm, &ld, &pr)); check("BVResize: new column count", (double)abs(mm - (m + 5)), 0.0); }
  download(Y, 5, z);
  check("BVResize(copy): old columns kept", maxabs_diff(z, Q + 3 * (int64_t)n, n), 0.0);
  CALL(sa_bv_set_active_columns(Y, 0, 3));
  CALL(sa_bv_set_active_columns(A, 0, 4));
  double Md[4 * 4 + 8];
  const int ldd = 5;
  double Mh[5 * 4];
  CALL(sa_bv_dot(A, Y, Mh, ldd));
  double edot = 0.0;
  for (int j = 0; j < 4; j++) {
    for (int i = 0; i < 3; i++) {
      download(Y, i, z);
      double s = 0.0;
      for (int r = 0; r < n; r++) s += z[r] * Q[r + (int64_t)j * n];
      edot = fmax(edot, fabs(s - Mh[i + j * ldd]));
    }
  }
  (void)Md;
  check("BVDot: M(Y.l+i, X.l+j) = Y_i^T X_j", edot, 1e-12);

  const int nx = 71, ny = 70, N = nx * ny;
  int32_t *rp = malloc(sizeof(int32_t) * (N + 1)), *ci = malloc(sizeof(int32_t) * 5 * N);
  double *va = malloc(sizeof(double) * 5 * N);
  int64_t nnz = lap2d(nx, ny, rp, ci, va);
  sa_mat Am;
  CALL(sa_mat_create_csr(N, nnz, rp, ci, va, SA_FP64, &Am));
  double *xl = malloc(sizeof(double) * N), *yl = malloc(sizeof(double) * N), *yh = malloc(sizeof(double) * N);
  for (int i = 0; i < N; i++) xl[i] = (double)((i * 7) % 11) - 5.0;
  double *xld = NULL, *yld = NULL;
  CALL(sa_malloc((void **)&xld, sizeof(double) * N));
  CALL(sa_malloc((void **)&yld, sizeof(double) * N));
  CALL(sa_memcpy_h2d(xld, xl, sizeof(double) * N));
  CALL(sa_mat_mult(Am, xld, yld));
  CALL(sa_memcpy_d2h(yl, yld, sizeof(double) * N));
  spmv_host(N, rp, ci, va, xl, yh);
  check("MatMult: integer stencil, exact", maxabs_diff(yl, yh, N), 0.0);

  sa_bv W, U;
  const int mw = 9;
  CALL(sa_bv_create(N, mw, SA_FP64, &W));
  CALL(sa_bv_create(N, mw, SA_FP64, &U));
  for (int j = 0; j < mw; j++) {
    for (int i = 0; i < N; i++) xl[i] = urand();
    upload(W, j, xl);
  }
  CALL(sa_bv_set_active_columns(W, 1, 5));
  CALL(sa_bv_set_active_columns(U, 3, 7));
  CALL(sa_bv_matmult(W, Am, U));
  double emm = 0.0;
  for (int j = 0; j < 4; j++) {
    download(W, 1 + j, xl);
    download(U, 3 + j, yl);
    spmv_host(N, rp, ci, va, xl, yh);
    emm = fmax(emm, maxabs_diff(yl, yh, N));
  }
  check("BVMatMult: U(:,U.l+j) = A W(:,W.l+j)", emm, 1e-13);

  /* BVMatArnoldi: A V(:,0:k) = V(:,0:k+1) H (bvkrylov.c:36-37) */
  const int ka = 7;
  double nrm0;
  int ld0;
  CALL(sa_bv_set_active_columns(W, 0, mw));
  CALL(sa_bv_orthonormalize_column(W, 0, eta, &nrm0, &ld0));   /* normalises column 0 */
  int mm = ka, bd = 0;
  double beta = 0.0;
  CALL(sa_bv_mat_arnoldi(W, Am, 0, &mm, eta, &beta, &bd));
  double *H = malloc(sizeof(double) * mw * mw), *Vw = malloc(sizeof(double) * (size_t)N * (ka + 1));
  CALL(sa_bv_get_coefficients(W, 0, mw, H));
  for (int j = 0; j <= ka; j++) download(W, j, Vw + (int64_t)j * N);
  double ear = 0.0;
  for (int j = 0; j < ka; j++) {
    spmv_host(N, rp, ci, va, Vw + (int64_t)j * N, yh);
    for (int i = 0; i < N; i++) {
      double s = 0.0;
      for (int q = 0; q <= j + 1; q++) s += Vw[i + (int64_t)q * N] * H[q + (int64_t)(j + 1) * mw];
      ear = fmax(ear, fabs(yh[i] - s));
    }
  }
  check("BVMatArnoldi: A V - V H", ear, 1e-12);
  check("BVMatArnoldi: no breakdown, full sweep", (double)(bd != 0 || mm != ka), 0.0);

  sa_bv_destroy(A); sa_bv_destroy(B); sa_bv_destroy(Y); sa_bv_destroy(W); sa_bv_destroy(U);
  sa_mat_destroy(Am);
  sa_free(xd); sa_free(yd); sa_free(xld); sa_free(yld);
  printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
