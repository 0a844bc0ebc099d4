/*
 * oracle.c -- CPU FP64 restatement of the SLEPc 3.22.2 Krylov-Schur hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product in slepc_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Structure follows the reference call stack (paths relative to the SLEPc tree):
 *   EPSSolve_KrylovSchur_Default   src/eps/impls/krylov/krylovschur/krylovschur.c:227-348
 *   BVMatLanczos / BVMatArnoldi    src/sys/classes/bv/interface/bvkrylov.c:56-113,165-226
 *   BVOrthonormalizeColumn         src/sys/classes/bv/interface/bvorthog.c:91-217,380-427
 *   DS HEP                         src/sys/classes/ds/impls/hep/dshep.c:137-426,643-670
 *   DS NHEP                        src/sys/classes/ds/impls/dsutil.c:21-166, nhep/dsnhep.c:101-425
 *   EPSKrylovConvergence           src/eps/impls/krylov/epskrylov.c:207-295
 *   EPSConvergedRelative           src/eps/interface/epsdefault.c:224-232
 *   EPSStoppingBasic               src/eps/interface/epsdefault.c:290-302
 *   SlepcSortEigenvalues           src/sys/slepcsc.c:89-140
 *   EPSArnoldi / EPSLanczos        src/eps/impls/krylov/arnoldi/arnoldi.c:31-131, lanczos/lanczos.c:518-697
 *   EPSLanczos local (MGS)         lanczos/lanczos.c:58-62,89-118; bvorthog.c:52-85,145-215
 *   EPSLanczos periodic/partial    lanczos/lanczos.c:260-441
 *   EPSLanczos selective           lanczos/lanczos.c:139-258 (LAPACK stevr)
 *   EPSLanczos delayed             epskrylov.c:24-114, lanczos/lanczos.c:495-512
 *   EPSSetBalance                  src/eps/interface/epsdefault.c:128-139,443-504
 *   BVDot / BVMatMult              src/sys/classes/bv/interface/bvglobal.c:86-120, bvops.c:692-722
 *   BVOrthogonalize (gs/chol/svqb) src/sys/classes/bv/interface/bvorthog.c:489-767, bvlapack.c:136-330
 * The dense projected problems call host LAPACK exactly where the reference does
 * (steqr, gehrd, orghr, hseqr, trexc, trevc, lartg, rot); the LAPACK library is
 * resolved at run time (scipy's bundled OpenBLAS or MKL).
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <dlfcn.h>
#include <glob.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================== */
/* Random numbers: counter-based splitmix64 so CPU and GPU draw equal bits.   */
/* ======================================================================== */
uint64_t orc_splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double orc_u01(uint64_t seed, uint64_t i)
{
  uint64_t z = orc_splitmix64(orc_splitmix64(seed) + i);
  return (double)(z >> 11) * 0x1.0p-53;
}
void orc_rand_uniform(uint64_t seed, int64_t n, double *out)
{
  for (int64_t i = 0; i < n; i++) out[i] = 2.0 * orc_u01(seed, (uint64_t)i) - 1.0;
}
double orc_gauss(uint64_t seed, uint64_t i)
{
  double u1 = 1.0 - orc_u01(seed, 2 * i), u2 = orc_u01(seed, 2 * i + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
uint64_t orc_sub_seed(uint64_t seed, uint64_t p)
{
  return orc_splitmix64(seed ^ (0xD1B54A32D192ED03ULL * (p + 1)));
}

/* ======================================================================== */
/* CSR generators.  Columns of every row ascending, as PETSc AIJ stores them  */
/* after MatAssemblyEnd (INSERT_VALUES, no duplicates).                       */
/* ======================================================================== */

/* 2-D 5-point Laplacian, row II = i*nx + j: src/eps/tutorials/ex2.c:46-53 */
int64_t orc_gen_lap2d(int nx, int ny, int32_t *rp, int32_t *ci, double *v)
{
  return orc_gen_lap2d_pot(nx, ny, 0, 0.0, rp, ci, v);
}

/* Same stencil plus diagonal potential scale*U[0,1) (batched family). */
int64_t orc_gen_lap2d_pot(int nx, int ny, uint64_t seed, double scale, int32_t *rp, int32_t *ci, double *v)
{
  int64_t N = (int64_t)nx * ny, nz = 0;
  for (int64_t II = 0; II < N; II++) {
    int64_t i = II / nx, j = II - i * nx;
    if (rp) rp[II] = (int32_t)nz;
#define PUT(c, x) do { if (ci) { ci[nz] = (int32_t)(c); v[nz] = (x); } nz++; } while (0)
    if (i > 0) PUT(II - nx, -1.0);
    if (j > 0) PUT(II - 1, -1.0);
    PUT(II, scale != 0.0 ? 4.0 + scale * orc_u01(seed, (uint64_t)II) : 4.0);
    if (j < nx - 1) PUT(II + 1, -1.0);
    if (i < ny - 1) PUT(II + nx, -1.0);
  }
  if (rp) rp[N] = (int32_t)nz;
  return nz;
}

/* 3-D 7-point convection-diffusion: diag 6, neighbour at -e_d: -1-0.1*b_d,
   at +e_d: -1+0.1*b_d, b = (1, 0.5, 0.25); row r = i + nx*(j + ny*k). */
int64_t orc_gen_convdiff3d(int nx, int ny, int nz_, int32_t *rp, int32_t *ci, double *v)
{
  const double b[3] = {1.0, 0.5, 0.25};
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) { lo[d] = -1.0 - 0.1 * b[d]; hi[d] = -1.0 + 0.1 * b[d]; }
  int64_t sxy = (int64_t)nx * ny, N = sxy * nz_, nz = 0;
  for (int64_t r = 0; r < N; r++) {
    int64_t k = r / sxy, rem = r - k * sxy, j = rem / nx, i = rem - j * nx;
    if (rp) rp[r] = (int32_t)nz;
    if (k > 0) PUT(r - sxy, lo[2]);
    if (j > 0) PUT(r - nx, lo[1]);
    if (i > 0) PUT(r - 1, lo[0]);
    PUT(r, 6.0);
    if (i < nx - 1) PUT(r + 1, hi[0]);
    if (j < ny - 1) PUT(r + nx, hi[1]);
    if (k < nz_ - 1) PUT(r + sxy, hi[2]);
  }
  if (rp) rp[N] = (int32_t)nz;
  return nz;
}

/* 3-D 27-point stencil: off-diagonal -1, diagonal 26 + U[0,1)(seed, r). */
int64_t orc_gen_stencil27(int nx, int ny, int nz_, uint64_t seed, int32_t *rp, int32_t *ci, double *v)
{
  int64_t sxy = (int64_t)nx * ny, N = sxy * nz_, nz = 0;
  for (int64_t r = 0; r < N; r++) {
    int64_t k = r / sxy, rem = r - k * sxy, j = rem / nx, i = rem - j * nx;
    if (rp) rp[r] = (int32_t)nz;
    for (int dk = -1; dk <= 1; dk++) {
      if (k + dk < 0 || k + dk >= nz_) continue;
      for (int dj = -1; dj <= 1; dj++) {
        if (j + dj < 0 || j + dj >= ny) continue;
        for (int di = -1; di <= 1; di++) {
          if (i + di < 0 || i + di >= nx) continue;
          int64_t c = r + di + (int64_t)nx * dj + sxy * dk;
          if (c == r) PUT(c, 26.0 + orc_u01(seed, (uint64_t)r));
          else PUT(c, -1.0);
        }
      }
    }
  }
  if (rp) rp[N] = (int32_t)nz;
  return nz;
}

/* Parameterised finite-volume diffusion -div(k grad u) on the unit square,
   nx x ny cells, homogeneous Dirichlet boundary, harmonic-mean face
   conductivities.  log k = sum_m theta_m sqrt(lambda_m) phi_p(x) phi_q(y) over
   the 8 leading separable cosine modes (p,q) of a Gaussian covariance with
   correlation length 0.2.  Returns sigma I - L with sigma the Gershgorin
   bound max_r sum_c |L_rc|, so the smallest eigenvalues of L become the
   largest of the returned matrix. */
static const int FV_P[8] = {1, 0, 1, 2, 0, 2, 1, 2};
static const int FV_Q[8] = {0, 1, 1, 0, 2, 1, 2, 2};
static double fv_lambda1(int p)
{
  const double ell = 0.2, PI = 3.141592653589793;
  double a = p * PI * ell;
  return sqrt(2.0 * PI) * ell * exp(-0.5 * a * a);
}
static double fv_phi(int p, double x)
{
  const double PI = 3.141592653589793;
  return p == 0 ? 1.0 : sqrt(2.0) * cos(p * PI * x);
}
void orc_fv2d_theta(uint64_t seed, int sample, double *theta8)
{
  for (int m = 0; m < 8; m++) theta8[m] = orc_gauss(seed, (uint64_t)sample * 8 + m);
}
int64_t orc_gen_fv2d(int nx, int ny, const double *theta, double *sigma_out, int32_t *rp, int32_t *ci, double *v)
{
  int64_t N = (int64_t)nx * ny, nz = 0;
  double *kc = (double *)malloc(sizeof(double) * N);
  double sq[8];
  for (int m = 0; m < 8; m++) sq[m] = theta[m] * sqrt(fv_lambda1(FV_P[m]) * fv_lambda1(FV_Q[m]));
  for (int64_t r = 0; r < N; r++) {
    int64_t j = r / nx, i = r - j * nx;
    double x = (i + 0.5) / nx, y = (j + 0.5) / ny, lk = 0.0;
    for (int m = 0; m < 8; m++) lk += sq[m] * fv_phi(FV_P[m], x) * fv_phi(FV_Q[m], y);
    kc[r] = exp(lk);
  }
  /* pass 1: Gershgorin bound of L (diag + |off| = 2*diag - boundary share) */
  double sigma = 0.0;
  for (int64_t r = 0; r < N; r++) {
    int64_t j = r / nx, i = r - j * nx;
    double k0 = kc[r], t[4], diag = 0.0, off = 0.0;
    int has[4] = {j > 0, i > 0, i < nx - 1, j < ny - 1};
    int64_t nb[4] = {r - nx, r - 1, r + 1, r + nx};
    for (int f = 0; f < 4; f++) {
      if (has[f]) { double k1 = kc[nb[f]]; t[f] = (2.0 * k0 * k1) / (k0 + k1); off += t[f]; }
      else t[f] = 2.0 * k0;
      diag += t[f];
    }
    double s = diag + off;
    if (s > sigma) sigma = s;
  }
  if (sigma_out) *sigma_out = sigma;
  for (int64_t r = 0; r < N; r++) {
    int64_t j = r / nx, i = r - j * nx;
    double k0 = kc[r], t[4], diag = 0.0;
    int has[4] = {j > 0, i > 0, i < nx - 1, j < ny - 1};
    int64_t nb[4] = {r - nx, r - 1, r + 1, r + nx};
    for (int f = 0; f < 4; f++) {
      if (has[f]) { double k1 = kc[nb[f]]; t[f] = (2.0 * k0 * k1) / (k0 + k1); }
      else t[f] = 2.0 * k0;
      diag += t[f];
    }
    if (rp) rp[r] = (int32_t)nz;
    if (has[0]) PUT(nb[0], t[0]);
    if (has[1]) PUT(nb[1], t[1]);
    PUT(r, sigma - diag);
    if (has[2]) PUT(nb[2], t[2]);
    if (has[3]) PUT(nb[3], t[3]);
  }
  if (rp) rp[N] = (int32_t)nz;
  free(kc);
  return nz;
}

/* Markov walk on a triangular grid, src/eps/tutorials/ex5.c:146-178
   (row ix-1 holds the transition probabilities into state ix-1). */
int64_t orc_gen_markov(int m, int32_t *rp, int32_t *ci, double *v)
{
  const double cst = 0.5 / (double)(m - 1);
  int64_t N = (int64_t)m * (m + 1) / 2, nz = 0, ix = 0;
  for (int i = 1; i <= m; i++) {
    int jmax = m - i + 1;
    for (int j = 1; j <= jmax; j++) {
      ix++;
      int64_t row = ix - 1;
      int64_t c[4]; double x[4]; int cnt = 0;
      double pu = 0.5 - cst * (double)(i + j - 3);
      if (i > 1) { c[cnt] = ix - jmax - 2; x[cnt++] = pu; }          /* west  */
      if (j > 1) { c[cnt] = ix - 2; x[cnt++] = pu; }                 /* south */
      if (j != jmax) {
        double pd = cst * (double)(i + j - 1);
        c[cnt] = ix; x[cnt++] = (i == 1) ? 2 * pd : pd;               /* north */
        c[cnt] = ix + jmax - 1; x[cnt++] = (j == 1) ? 2 * pd : pd;    /* east  */
      }
      /* sort ascending (at most 4 entries) */
      for (int a = 1; a < cnt; a++)
        for (int b2 = a; b2 > 0 && c[b2 - 1] > c[b2]; b2--) {
          int64_t tc = c[b2]; c[b2] = c[b2 - 1]; c[b2 - 1] = tc;
          double tx = x[b2]; x[b2] = x[b2 - 1]; x[b2 - 1] = tx;
        }
      if (rp) rp[row] = (int32_t)nz;
      for (int a = 0; a < cnt; a++) PUT(c[a], x[a]);
    }
  }
  if (rp) rp[N] = (int32_t)nz;
  return nz;
}
#undef PUT

/* ======================================================================== */
/* LAPACK / BLAS resolved at run time                                        */
/* ======================================================================== */
typedef void (*f_lartg)(const double *, const double *, double *, double *, double *);
typedef void (*f_rot)(const int *, double *, const int *, double *, const int *, const double *, const double *);
typedef void (*f_steqr)(const char *, const int *, double *, double *, double *, const int *, double *, int *, size_t);
typedef void (*f_stevr)(const char *, const char *, const int *, double *, double *, const double *, const double *,
                        const int *, const int *, const double *, int *, double *, double *, const int *, int *,
                        double *, const int *, int *, const int *, int *, size_t, size_t);
typedef void (*f_gehrd)(const int *, const int *, const int *, double *, const int *, double *, double *, const int *, int *);
typedef void (*f_orghr)(const int *, const int *, const int *, double *, const int *, const double *, double *, const int *, int *);
typedef void (*f_hseqr)(const char *, const char *, const int *, const int *, const int *, double *, const int *, double *, double *,
                        double *, const int *, double *, const int *, int *, size_t, size_t);
typedef void (*f_trexc)(const char *, const int *, double *, const int *, double *, const int *, int *, int *, double *, int *, size_t);
typedef void (*f_trevc)(const char *, const char *, int *, const int *, const double *, const int *, double *, const int *, double *,
                        const int *, const int *, int *, double *, int *, size_t, size_t);
typedef void (*f_potrf)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*f_trtri)(const char *, const char *, const int *, double *, const int *, int *, size_t, size_t);
typedef void (*f_syev)(const char *, const char *, const int *, double *, const int *, double *, double *, const int *, int *,
                       size_t, size_t);
typedef void (*f_getrf)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*f_getrs)(const char *, const int *, const int *, const double *, const int *, const int *, double *,
                        const int *, int *, size_t);
typedef void (*f_geqrf)(const int *, const int *, double *, const int *, double *, double *, const int *, int *);
typedef void (*f_orgqr)(const int *, const int *, const int *, double *, const int *, const double *, double *, const int *,
                        int *);
static struct {
  int tried, ok;
  char path[512];
  f_stevr stevr;   /* optional: selective Lanczos */
  f_lartg lartg; f_rot rot; f_steqr steqr; f_gehrd gehrd; f_orghr orghr; f_hseqr hseqr; f_trexc trexc; f_trevc trevc;
  f_potrf potrf; f_trtri trtri; f_syev syev; f_geqrf geqrf; f_orgqr orgqr;   /* block orthogonalisation only */
  f_getrf getrf; f_getrs getrs;                                               /* harmonic extraction only */
} LP;

static void *lp_sym(void *h, const char *name)
{
  char buf[64];
  snprintf(buf, sizeof buf, "%s_", name);
  void *p = dlsym(h, buf);
  if (!p) { snprintf(buf, sizeof buf, "scipy_%s_", name); p = dlsym(h, buf); }
  return p;
}
static int lp_try(const char *path)
{
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return 0;
  LP.lartg = (f_lartg)lp_sym(h, "dlartg"); LP.rot = (f_rot)lp_sym(h, "drot");
  LP.steqr = (f_steqr)lp_sym(h, "dsteqr"); LP.stevr = (f_stevr)lp_sym(h, "dstevr"); LP.gehrd = (f_gehrd)lp_sym(h, "dgehrd");
  LP.orghr = (f_orghr)lp_sym(h, "dorghr"); LP.hseqr = (f_hseqr)lp_sym(h, "dhseqr");
  LP.trexc = (f_trexc)lp_sym(h, "dtrexc"); LP.trevc = (f_trevc)lp_sym(h, "dtrevc");
  LP.potrf = (f_potrf)lp_sym(h, "dpotrf"); LP.trtri = (f_trtri)lp_sym(h, "dtrtri");
  LP.syev = (f_syev)lp_sym(h, "dsyev");
  LP.geqrf = (f_geqrf)lp_sym(h, "dgeqrf"); LP.orgqr = (f_orgqr)lp_sym(h, "dorgqr");
  LP.getrf = (f_getrf)lp_sym(h, "dgetrf"); LP.getrs = (f_getrs)lp_sym(h, "dgetrs");
  if (!(LP.lartg && LP.rot && LP.steqr && LP.gehrd && LP.orghr && LP.hseqr && LP.trexc && LP.trevc)) { dlclose(h); return 0; }
  void (*setlayer)(int) = (void (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
  if (setlayer) setlayer(1); /* MKL_THREADING_SEQUENTIAL */
  /* the projected problems are <= 97 x 97: one BLAS thread (a pool sized to
     every core of a large host only adds wake-up latency) */
  const char *nts[] = {"scipy_openblas_set_num_threads", "openblas_set_num_threads", NULL};
  for (int i = 0; nts[i]; i++) {
    void (*setnt)(int) = (void (*)(int))dlsym(h, nts[i]);
    if (setnt) { setnt(1); break; }
  }
  snprintf(LP.path, sizeof LP.path, "%s", path);
  return 1;
}
static int lp_load(void)
{
  if (LP.tried) return LP.ok;
  LP.tried = 1;
  const char *env = getenv("SLEPC_AMD_LAPACK");
  if (env && lp_try(env)) return LP.ok = 1;
  glob_t g;
  if (glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so", 0, NULL, &g) == 0) {
    for (size_t i = 0; i < g.gl_pathc; i++)
      if (lp_try(g.gl_pathv[i])) { globfree(&g); return LP.ok = 1; }
    globfree(&g);
  }
  const char *cands[] = {"/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", "libopenblas.so.0", "liblapack.so.3", NULL};
  for (int i = 0; cands[i]; i++) if (lp_try(cands[i])) return LP.ok = 1;
  return LP.ok = 0;
}
int orc_lapack_status(char *buf, int len)
{
  int ok = lp_load();
  if (buf && len > 0) snprintf(buf, (size_t)len, "%s", ok ? LP.path : "not found");
  return ok;
}

/* ======================================================================== */
/* SpMV and BV kernels (bvblas.c / bvmat.c semantics, OpenMP over rows)      */
/* ======================================================================== */
void orc_spmv(int n, const int32_t *rp, const int32_t *ci, const double *val, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; r++) {
    double s = 0.0;
    for (int32_t p = rp[r]; p < rp[r + 1]; p++) s += val[p] * x[ci[p]];
    y[r] = s;
  }
}

/* q = V(:,0:k)^T x : BVDotVec_BLAS_Private (bvblas.c:240-261) */
void orc_bv_dotvec(int64_t n, int k, const double *V, int64_t ld, const double *x, double *q)
{
  if (k <= 0) return;
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  double *part = (double *)calloc((size_t)nt * k, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    int64_t r0 = n * t / nt, r1 = n * (t + 1) / nt;
    double *p = part + (size_t)t * k;
    for (int c = 0; c < k; c++) {
      const double *vc = V + (int64_t)c * ld;
      double s = 0.0;
      for (int64_t r = r0; r < r1; r++) s += vc[r] * x[r];
      p[c] = s;
    }
  }
  for (int c = 0; c < k; c++) {
    double s = 0.0;
    for (int t = 0; t < nt; t++) s += part[(size_t)t * k + c];
    q[c] = s;
  }
  free(part);
}

/* y = beta*y + alpha*V(:,0:k) q : BVMultVec_BLAS_Private (bvblas.c:56-67) */
void orc_bv_multvec(int64_t n, int k, double alpha, const double *V, int64_t ld, const double *q, double beta, double *y)
{
  /* rows in blocks, columns streamed one after the other inside a block (the
     same additions per row, in the same order, as a row-by-row loop) */
  const int64_t RB = 1024;
#pragma omp parallel for schedule(static)
  for (int64_t r0 = 0; r0 < n; r0 += RB) {
    const int64_t r1 = r0 + RB < n ? r0 + RB : n;
    double s[1024];
    for (int64_t r = r0; r < r1; r++) s[r - r0] = 0.0;
    for (int c = 0; c < k; c++) {
      const double *vc = V + (int64_t)c * ld;
      const double qc = q[c];
      for (int64_t r = r0; r < r1; r++) s[r - r0] += vc[r] * qc;
    }
    /* beta == 0: y is not read (gemv semantics, bvblas.c:56-67) */
    for (int64_t r = r0; r < r1; r++) y[r] = (beta == 0.0 ? 0.0 : beta * y[r]) + alpha * s[r - r0];
  }
}

/* V(:,s:e) = V(:,l:k) * Q(l:k, s:e) : BVMultInPlace_BLAS_Private (bvblas.c:74-106) */
void orc_bv_multinplace(int64_t n, double *V, int64_t ld, int l, int k, int s, int e, const double *Q, int ldq)
{
  int ncol = e - s, kk = k - l;
  if (ncol <= 0 || n == 0) return;
  const int BS = 64;
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * BS * ncol);
#pragma omp for schedule(static)
    for (int64_t r0 = 0; r0 < n; r0 += BS) {
      int rows = (int)((n - r0) < BS ? (n - r0) : BS);
      for (int j = 0; j < ncol; j++)
        for (int r = 0; r < rows; r++) {
          double acc = 0.0;
          for (int c = 0; c < kk; c++) acc += V[r0 + r + (int64_t)(l + c) * ld] * Q[(l + c) + (int64_t)(s + j) * ldq];
          tmp[r + j * BS] = acc;
        }
      for (int j = 0; j < ncol; j++)
        for (int r = 0; r < rows; r++) V[r0 + r + (int64_t)(s + j) * ld] = tmp[r + j * BS];
    }
    free(tmp);
  }
}

void orc_bv_scale(int64_t n, double *x, double alpha)
{
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) x[r] *= alpha;
}

double orc_bv_norm2(int64_t n, const double *x)
{
  double s;
  orc_bv_dotvec(n, 1, x, n, x, &s);
  return sqrt(s);
}

/* ======================================================================== */
/* Gram-Schmidt: BVOrthogonalizeCGS1 (bvorthog.c:91-132) and                  */
/* BVOrthogonalizeGS REFINE_IFNEEDED (bvorthog.c:145-217).  The coefficient   */
/* buffer is m x m column-major; column 0 is the scratch vector c and column  */
/* j accumulates h_j with the norm stored at row j (bvimpl.h:289-415).        */
/* ======================================================================== */
static double safe_sqrt(double a) { return a < 0.0 ? 0.0 : sqrt(a); } /* BV_SafeSqrt, bvimpl.h:121-141 */

static void cgs1(int64_t n, double *V, int64_t ld, int j, double *buf, int m, double *onorm, double *norm, int64_t *passes)
{
  double *c = buf;               /* scratch column 0 */
  double *h = buf + (int64_t)j * m;
  double *vj = V + (int64_t)j * ld;
  (*passes)++;
  orc_bv_dotvec(n, j + 1, V, ld, vj, c);           /* BVDotColumnInc: includes v_j^T v_j */
  double beta = safe_sqrt(c[j]);                   /* BV_SquareRoot */
  orc_bv_multvec(n, j, -1.0, V, ld, c, 1.0, vj);    /* BVMultColumn(-1,1) */
  if (onorm) *onorm = beta;
  if (norm) {
    double sum = 0.0;
    for (int i = 0; i < j; i++) sum += c[i] * c[i];  /* BV_SquareSum */
    double nn = beta * beta - sum;
    if (nn <= 0.0) *norm = orc_bv_norm2(n, vj);
    else *norm = sqrt(nn);
  }
  for (int i = 0; i < j; i++) h[i] += c[i];        /* BV_AddCoefficients */
}

/* refine (BVSetOrthogonalization, bvorthog.c:172-199): ORC_REFINE_IFNEEDED
   repeats the pass while ||v'|| < eta ||v||, at most three passes;
   ORC_REFINE_NEVER takes one pass and the true norm after it (lindep only
   for a zero norm); ORC_REFINE_ALWAYS takes two passes, the second with the
   norm estimate. */
static void orthogonalize_gs(int64_t n, double *V, int64_t ld, int j, double *buf, int m, double eta, int refine,
                             double *norm, int *lindep, int64_t *passes)
{
  double onrm = 0.0, nrm = 0.0;
  double *h = buf + (int64_t)j * m;
  for (int i = 0; i < j; i++) h[i] = 0.0;           /* BV_CleanCoefficients */
  if (refine == ORC_REFINE_NEVER) {
    cgs1(n, V, ld, j, buf, m, NULL, NULL, passes);
    nrm = orc_bv_norm2(n, V + (int64_t)j * ld);     /* BV_NormVecOrColumn */
    *lindep = !(nrm != 0.0);
  } else if (refine == ORC_REFINE_ALWAYS) {
    cgs1(n, V, ld, j, buf, m, NULL, NULL, passes);
    cgs1(n, V, ld, j, buf, m, &onrm, &nrm, passes);
    *lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
  } else {
    cgs1(n, V, ld, j, buf, m, &onrm, &nrm, passes);
    int l = 1;
    while (l < 3 && nrm != 0.0 && fabs(nrm) < eta * fabs(onrm)) {
      l++;
      cgs1(n, V, ld, j, buf, m, &onrm, &nrm, passes);
    }
    *lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
  }
  *norm = nrm;
  h[j] = *lindep ? 0.0 : nrm;                       /* BV_SetValue(k,k,h,nrm) */
}

/* One pass of modified Gram-Schmidt (BVOrthogonalizeMGS1, bvorthog.c:52-85) of
   column j against the columns i < j with which[i] set (all when which is
   NULL), in increasing order; coefficients added into h. */
static void mgs1(int64_t n, double *V, int64_t ld, int j, const char *which, double *h, double *onrm, double *nrm)
{
  double *w = V + (int64_t)j * ld;
  if (onrm) *onrm = orc_bv_norm2(n, w);
  for (int i = 0; i < j; i++) {
    if (which && !which[i]) continue;
    const double *vi = V + (int64_t)i * ld;
    double dot = 0.0;
    for (int64_t r = 0; r < n; r++) dot += w[r] * vi[r];
    for (int64_t r = 0; r < n; r++) w[r] -= dot * vi[r];
    h[i] += dot;
  }
  if (nrm) *nrm = orc_bv_norm2(n, w);
}

/* BVOrthogonalizeGS (bvorthog.c:145-215) with BV_ORTHOG_MGS and refinement
   IFNEEDED: repeat while ||v|| < eta ||v_before||, at most three passes. */
static void orthogonalize_mgs(int64_t n, double *V, int64_t ld, int j, const char *which, double *buf, int m,
                              double eta, int refine, double *norm, int *lindep, int64_t *passes)
{
  double onrm = 0.0, nrm = 0.0;
  double *h = buf + (int64_t)j * m;
  for (int i = 0; i < j; i++) h[i] = 0.0;
  if (refine == ORC_REFINE_NEVER) {                 /* bvorthog.c:187-193 */
    mgs1(n, V, ld, j, which, h, NULL, NULL);
    (*passes)++;
    nrm = orc_bv_norm2(n, V + (int64_t)j * ld);
    *lindep = !(nrm != 0.0);
  } else if (refine == ORC_REFINE_ALWAYS) {         /* bvorthog.c:195-199 */
    mgs1(n, V, ld, j, which, h, NULL, NULL);
    mgs1(n, V, ld, j, which, h, &onrm, &nrm);
    (*passes) += 2;
    *lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
  } else {
    mgs1(n, V, ld, j, which, h, &onrm, &nrm);
    (*passes)++;
    int l = 1;
    while (l < 3 && nrm != 0.0 && fabs(nrm) < eta * fabs(onrm)) {
      l++;
      onrm = nrm;
      mgs1(n, V, ld, j, which, h, NULL, &nrm);
      (*passes)++;
    }
    *lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
  }
  *norm = nrm;
  h[j] = *lindep ? 0.0 : nrm;
}

/* BVOrthogonalizeSomeColumn (bvorthog.c:317-378, MGS): column j against the
   columns i < j with which[i] set, with refinement refine; coefficients into
   column j of the m x m buffer, the norm at row j. */
int orc_bv_orthogonalize_some_column(int64_t n, double *V, int64_t ld, int j, const char *which, double *buffer,
                                     int m, int refine, double eta, double *norm, int *lindep)
{
  int64_t passes = 0;
  double nrm = 0.0;
  int ld_ = 0;
  orthogonalize_mgs(n, V, ld, j, which, buffer, m, eta, refine, &nrm, &ld_, &passes);
  if (norm) *norm = nrm;
  if (lindep) *lindep = ld_;
  return (int)passes;
}

int orc_bv_orthonormalize_column(int64_t n, double *V, int64_t ld, int j, double *buffer, int m, double eta,
                                 double *norm, int *lindep)
{
  int64_t passes = 0;
  double nrm;
  int ld_ = 0;
  orthogonalize_gs(n, V, ld, j, buffer, m, eta, ORC_REFINE_IFNEEDED, &nrm, &ld_, &passes);
  if (nrm != 1.0 && nrm != 0.0) orc_bv_scale(n, V + (int64_t)j * ld, 1.0 / nrm);
  if (norm) *norm = nrm;
  if (lindep) *lindep = ld_;
  return (int)passes;
}

/* ======================================================================== */
/* Block BV operations                                                        */
/* ======================================================================== */
/* BVDot (bvglobal.c:86-120 -> BVDot_BLAS_Private bvblas.c:199-233, gemm "C","N"):
   M(i, j) = Y(:,i) . X(:,j), i < my, j < mx, M leading dimension ldm. */
void orc_bv_dot(int64_t n, const double *X, int64_t ldx, int mx, const double *Y, int64_t ldy, int my, double *M, int ldm)
{
  for (int j = 0; j < mx; j++)
    for (int i = 0; i < my; i++) {
      double s = 0.0;
      for (int64_t r = 0; r < n; r++) s += Y[r + (int64_t)i * ldy] * X[r + (int64_t)j * ldx];
      M[i + (int64_t)j * ldm] = s;
    }
}

/* BVMatMult (bvops.c:692-722, column by column as BV_MATMULT_VECS):
   Y(:, j) = A X(:, j), j < nc. */
void orc_bv_matmult(int n, const int32_t *rp, const int32_t *ci, const double *val, const double *X, int64_t ldx,
                    double *Y, int64_t ldy, int nc)
{
  for (int j = 0; j < nc; j++) orc_spmv(n, rp, ci, val, X + (int64_t)j * ldx, Y + (int64_t)j * ldy);
}

/* BVOrthogonalize (bvorthog.c:741-767) of columns l..k-1 of V (n x m, ld),
   the leading ones orthonormal.  method 0: BVOrthogonalize_GS (:507-548),
   column by column; 1: BVOrthogonalize_Chol (:588-603) with
   BVMatCholInv_LAPACK_Private (bvlapack.c:136-203); 2: BVOrthogonalize_SVQB
   (:671-686) with BVMatSVQB_LAPACK_Private (bvlapack.c:261-330); 3:
   BVOrthogonalize_TSQR (:615-631), 4: BVOrthogonalize_TSQRCHOL (:638-659),
   both as on one process (geqrf/orgqr of the whole block).  Methods 1-4
   first run BVOrthogonalize_BlockGS (:489-502) against the leading columns.
   R (ldr >= k) receives columns l..k-1 of the factor (zeros below the
   diagonal of the triangular ones).  Returns 0, or -1 on breakdown or a
   LAPACK failure. */
int orc_bv_orthogonalize(int64_t n, double *V, int64_t ld, int m, int l, int k, int method, double *R, int ldr)
{
  const int nk = k - l;
  if (nk <= 0) return 0;
  if (method == 0) {
    double *buf = (double *)calloc((size_t)m * m, sizeof(double));
    for (int j = l; j < k; j++) {
      double nrm;
      int lind;
      orc_bv_orthonormalize_column(n, V, ld, j, buf, m, 0.7071, &nrm, &lind);
      if (nrm == 0.0) { free(buf); return -1; }
      if (R)
        for (int i = 0; i < k; i++) R[i + (int64_t)j * ldr] = i < j ? buf[i + (int64_t)j * m] : (i == j ? nrm : 0.0);
    }
    free(buf);
    return 0;
  }
  if (!lp_load() || !LP.potrf || !LP.trtri || !LP.syev || !LP.geqrf || !LP.orgqr) return -1;
  double *R12 = (double *)calloc((size_t)(l > 0 ? l : 1) * nk, sizeof(double));
  double *G = (double *)malloc(sizeof(double) * nk * nk), *Rf = (double *)calloc((size_t)nk * nk, sizeof(double));
  double *S = (double *)calloc((size_t)nk * nk, sizeof(double)), *Q = (double *)calloc((size_t)k * k, sizeof(double));
  int rc = 0, info = 0;
  if (l) {                                /* V2 -= V1 (V1^T V2) */
    orc_bv_dot(n, V + (int64_t)l * ld, ld, nk, V, ld, l, R12, l);
    for (int j = 0; j < nk; j++) {
      for (int i = 0; i < l; i++) Q[i + (int64_t)(l + j) * k] = -R12[i + (int64_t)j * l];
      Q[(l + j) + (int64_t)(l + j) * k] = 1.0;
    }
    orc_bv_multinplace(n, V, ld, 0, k, l, k, Q, k);
  }
  if (method >= 3) {
    /* BVOrthogonalize_TSQR / _TSQRCHOL on one process (bvorthog.c:615-659,
       bvlapack.c:347-398,510-540 with nlevels = 0 and size = 1): geqrf of V2;
       R is its upper triangle; Q from orgqr (tsqr) or V2 inv(R) (tsqrchol). */
    const int nn = (int)n;
    double *A = (double *)malloc(sizeof(double) * (size_t)n * nk), *tau = (double *)malloc(sizeof(double) * nk), q;
    for (int j = 0; j < nk; j++) memcpy(A + (size_t)j * n, V + (int64_t)(l + j) * ld, sizeof(double) * n);
    int lwork = -1;
    LP.geqrf(&nn, &nk, A, &nn, tau, &q, &lwork, &info);
    lwork = (int)q > nk ? (int)q : nk;
    double *work = (double *)malloc(sizeof(double) * lwork);
    LP.geqrf(&nn, &nk, A, &nn, tau, work, &lwork, &info);
    if (info) rc = -1;
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < nk; i++) Rf[i + (int64_t)j * nk] = i <= j ? A[i + (size_t)j * n] : 0.0;
    if (!rc && method == 3) {
      LP.orgqr(&nn, &nk, &nk, A, &nn, tau, work, &lwork, &info);
      if (info) rc = -1;
      for (int j = 0; j < nk && !rc; j++) memcpy(V + (int64_t)(l + j) * ld, A + (size_t)j * n, sizeof(double) * n);
    } else if (!rc) {
      memcpy(S, Rf, sizeof(double) * nk * nk);
      LP.trtri("U", "N", &nk, S, &nk, &info, 1, 1);
      if (info) rc = -1;
      memset(Q, 0, sizeof(double) * k * k);
      for (int j = 0; j < nk; j++)
        for (int i = 0; i < nk; i++) Q[(l + i) + (int64_t)(l + j) * k] = S[i + (int64_t)j * nk];
      if (!rc) orc_bv_multinplace(n, V, ld, l, k, l, k, Q, k);
    }
    free(A); free(tau); free(work);
    if (!rc && R)
      for (int j = 0; j < nk; j++)
        for (int i = 0; i < k; i++) R[i + (int64_t)(l + j) * ldr] = i < l ? R12[i + (int64_t)j * l] : Rf[(i - l) + (int64_t)j * nk];
    free(R12); free(G); free(Rf); free(S); free(Q);
    return rc;
  }
  orc_bv_dot(n, V + (int64_t)l * ld, ld, nk, V + (int64_t)l * ld, ld, nk, G, nk);
  if (method == 1) {
    memcpy(Rf, G, sizeof(double) * nk * nk);
    LP.potrf("U", &nk, Rf, &nk, &info, 1);
    if (info) {                           /* retry on the diagonally perturbed matrix */
      memcpy(Rf, G, sizeof(double) * nk * nk);
      for (int i = 0; i < nk; i++) Rf[i + (int64_t)i * nk] += 50.0 * 2.220446049250313e-16;
      LP.potrf("U", &nk, Rf, &nk, &info, 1);
    }
    if (info) rc = -1;
    for (int j = 0; j < nk; j++)
      for (int i = j + 1; i < nk; i++) Rf[i + (int64_t)j * nk] = 0.0;
    memcpy(S, Rf, sizeof(double) * nk * nk);
    if (!rc) {
      LP.trtri("U", "N", &nk, S, &nk, &info, 1, 1);
      if (info) rc = -1;
    }
  } else {
    double *D = (double *)malloc(sizeof(double) * nk), *w = (double *)malloc(sizeof(double) * nk);
    double *U = (double *)malloc(sizeof(double) * nk * nk), q;
    int lwork = -1;
    for (int i = 0; i < nk; i++) D[i] = 1.0 / sqrt(G[i + (int64_t)i * nk]);
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < nk; i++) U[i + (int64_t)j * nk] = G[i + (int64_t)j * nk] * D[i] * D[j];
    LP.syev("V", "L", &nk, U, &nk, w, &q, &lwork, &info, 1, 1);
    lwork = (int)q > 1 ? (int)q : 1;
    double *work = (double *)malloc(sizeof(double) * lwork);
    LP.syev("V", "L", &nk, U, &nk, w, work, &lwork, &info, 1, 1);
    if (info) rc = -1;
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < nk; i++) {
        S[i + (int64_t)j * nk] = U[i + (int64_t)j * nk] * D[i] / sqrt(w[j]);
        Rf[i + (int64_t)j * nk] = U[j + (int64_t)i * nk] * sqrt(w[i]) / D[j];
      }
    free(D); free(w); free(U); free(work);
  }
  if (!rc) {                              /* V2 = V2 S */
    memset(Q, 0, sizeof(double) * k * k);
    for (int j = 0; j < nk; j++)
      for (int i = 0; i < nk; i++) Q[(l + i) + (int64_t)(l + j) * k] = S[i + (int64_t)j * nk];
    orc_bv_multinplace(n, V, ld, l, k, l, k, Q, k);
    if (R)
      for (int j = 0; j < nk; j++)
        for (int i = 0; i < k; i++) R[i + (int64_t)(l + j) * ldr] = i < l ? R12[i + (int64_t)j * l] : Rf[(i - l) + (int64_t)j * nk];
  }
  free(R12); free(G); free(Rf); free(S); free(Q);
  return rc;
}

/* ======================================================================== */
/* Dense projected problem (DS)                                               */
/* ======================================================================== */
enum { ST_RAW = 0, ST_INTERMEDIATE = 1, ST_CONDENSED = 2, ST_TRUNCATED = 3 };
typedef struct {
  int ld, n, l, k, t, state, hermitian, which;
  double *A, *Q, *T, *X, *work;
  int *perm, *iwork;
} DSo;

/* SlepcCompare* (slepcsc.c:152-267): result>0 means (ar,ai) goes after (br,bi);
   the target ones compare distances to the target. */
/* EPSSetEigenvalueComparison: the user's comparison for ORC_USER */
static orc_compare_fn user_cmp;
static void *user_cmp_ctx;
void orc_set_user_comparison(orc_compare_fn fn, void *ctx) { user_cmp = fn; user_cmp_ctx = ctx; }

static int sc_compare_t(int which, double ar, double ai, double br, double bi, double target)
{
  double a, b;
  switch (which) {
    case ORC_USER: {
      int r = 0;
      if (user_cmp) user_cmp(ar, ai, br, bi, &r, user_cmp_ctx);
      return r > 0 ? 1 : (r < 0 ? -1 : 0);
    }
    case ORC_LM: a = hypot(ar, ai); b = hypot(br, bi); return a < b ? 1 : (a > b ? -1 : 0);
    case ORC_SM: a = hypot(ar, ai); b = hypot(br, bi); return a > b ? 1 : (a < b ? -1 : 0);
    case ORC_LR: a = ar; b = br; return a < b ? 1 : (a > b ? -1 : 0);
    case ORC_SR: a = ar; b = br; return a > b ? 1 : (a < b ? -1 : 0);
    case ORC_LI: a = fabs(ai); b = fabs(bi); return a < b ? 1 : (a > b ? -1 : 0);
    case ORC_SI: a = fabs(ai); b = fabs(bi); return a > b ? 1 : (a < b ? -1 : 0);
    case ORC_TM: a = hypot(ar - target, ai); b = hypot(br - target, bi); return a > b ? 1 : (a < b ? -1 : 0);
    case ORC_TR: a = fabs(ar - target); b = fabs(br - target); return a > b ? 1 : (a < b ? -1 : 0);
  }
  return 0;
}
#define sc_compare(w, ar, ai, br, bi) sc_compare_t((w), (ar), (ai), (br), (bi), cmp_target)
static _Thread_local double cmp_target;   /* EPSSetTarget of the solve in progress on this thread */
static _Thread_local int cmp_sinv;        /* shift-and-invert: the DS compares back-transformed values */
static _Thread_local double cmp_sigma;

/* STBackTransform_Sinvert (sinvert.c:16-38) */
static void sinv_map(double *re, double *im)
{
  if (*im == 0.0) *re = 1.0 / *re + cmp_sigma;
  else { double t = *re * *re + *im * *im; *re = *re / t + cmp_sigma; *im = -*im / t; }
}

/* SlepcSCCompare of the DS (slepcsc.c:40-63): the ST map, then the criterion */
static int ds_compare(int w, double ar, double ai, double br, double bi)
{
  if (cmp_sinv) { sinv_map(&ar, &ai); sinv_map(&br, &bi); }
  return sc_compare_t(w, ar, ai, br, bi, cmp_target);
}

/* DSTranslateHarmonic_NHEP (dsnhep.c:466-537).  Forward: g = (A - tau I)^-T
   beta e_n, A(:,n-1) += beta g.  Recover (ncol = l + k): gh = -Q(:,0:ncol)^T g,
   A(i,j) += beta gh_i Q(n-1,j) for i < ncol, l <= j < ncol; g += Q(:,0:ncol) gh;
   extra row of columns l..ncol-1 times gamma.  gamma = sqrt(1 + ||g||^2). */
static int ds_translate_harmonic(DSo *ds, double tau, double beta, int recover, double *g, double *gamma)
{
  const int ld = ds->ld, n = ds->n;
  double *A = ds->A;
  if (!recover) {
    double *W = (double *)malloc(sizeof(double) * (size_t)ld * n);
    for (int j = 0; j < n; j++) memcpy(W + (int64_t)j * ld, A + (int64_t)j * ld, sizeof(double) * n);
    for (int i = 0; i < n; i++) { g[i] = 0.0; W[i + (int64_t)i * ld] -= tau; }
    g[n - 1] = beta;
    int nn = n, info = 0, one = 1;
    LP.getrf(&nn, &nn, W, &ld, ds->iwork, &info);
    if (!info) LP.getrs("C", &nn, &one, W, &ld, ds->iwork, g, &ld, &info, 1);
    free(W);
    if (info) return info;
    for (int i = 0; i < n; i++) A[i + (int64_t)(n - 1) * ld] += g[i] * beta;
  } else {
    const int ncol = ds->l + ds->k;
    const double *Q = ds->Q;
    double *gh = (double *)malloc(sizeof(double) * (ncol > 0 ? ncol : 1));
    for (int j = 0; j < ncol; j++) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += Q[i + (int64_t)j * ld] * g[i];
      gh[j] = -s;
    }
    for (int i = 0; i < ncol; i++)
      for (int j = ds->l; j < ncol; j++) A[i + (int64_t)j * ld] += gh[i] * Q[n - 1 + (int64_t)j * ld] * beta;
    for (int i = 0; i < n; i++) {
      double s = g[i];
      for (int j = 0; j < ncol; j++) s += Q[i + (int64_t)j * ld] * gh[j];
      g[i] = s;
    }
    free(gh);
  }
  double ss = 0.0;
  for (int i = 0; i < n; i++) ss += g[i] * g[i];
  *gamma = sqrt(1.0 + ss);
  if (recover)
    for (int j = ds->l; j < ds->l + ds->k; j++) A[n + (int64_t)j * ld] *= *gamma;
  return 0;
}

/* DSArrowTridiag (dshep.c:221-265), Rutishauser's Givens chase */
void orc_arrow_tridiag(int n, double *d, double *e, double *Q, int ld)
{
  int one = 1;
  double c, s, p, off, temp;
  if (n <= 2) return;
  lp_load();
  for (int j = 0; j < n - 2; j++) {
    temp = e[j + 1];
    LP.lartg(&temp, &e[j], &c, &s, &e[j + 1]);
    s = -s;
    temp = d[j + 1];
    e[j] = c * s * (temp - d[j]);
    d[j + 1] = s * s * d[j] + c * c * temp;
    d[j] = c * c * d[j] + s * s * temp;
    int j2 = j + 2;
    LP.rot(&j2, Q + (int64_t)j * ld, &one, Q + (int64_t)(j + 1) * ld, &one, &c, &s);
    for (int i = j - 1; i >= 0; i--) {
      off = -s * e[i];
      e[i] = c * e[i];
      temp = e[i + 1];
      LP.lartg(&temp, &off, &c, &s, &e[i + 1]);
      s = -s;
      temp = (d[i] - d[i + 1]) * s - 2.0 * c * e[i];
      p = s * temp;
      d[i + 1] += p;
      d[i] -= p;
      e[i] = -e[i] - c * temp;
      LP.rot(&j2, Q + (int64_t)i * ld, &one, Q + (int64_t)(i + 1) * ld, &one, &c, &s);
    }
  }
}

/* DSSetIdentity (dspriv.c:407-426) */
static void ds_set_identity(DSo *ds, double *M)
{
  memset(M + (int64_t)ds->l * ds->ld, 0, sizeof(double) * (size_t)ds->ld * (ds->n - ds->l));
  for (int i = ds->l; i < ds->n; i++) M[i + (int64_t)i * ds->ld] = 1.0;
}

/* DSSolve_HEP_QR with DSIntermediate_HEP, compact storage (dshep.c:267-321,383-426) */
static int ds_solve_hep(DSo *ds, double *wr, double *wi)
{
  int n = ds->n, l = ds->l, ld = ds->ld, n1 = n - l, info = 0;
  double *d = ds->T, *e = ds->T + ld;
  int na = ds->k - l + 1; if (na < 0) na = 0;
  ds_set_identity(ds, ds->Q);
  int off = l + l * ld;
  if (ds->state < ST_INTERMEDIATE) orc_arrow_tridiag(na, d + l, e + l, ds->Q + off, ld);
  for (int i = 0; i < l; i++) wr[i] = d[i];
  if (n1 > 0) LP.steqr("V", &n1, d + l, e + l, ds->Q + off, &ld, ds->work, &info, 1);
  for (int i = l; i < n; i++) wr[i] = d[i];
  for (int i = 0; i < n - 1; i++) e[i] = 0.0;
  if (wi) for (int i = l; i < n; i++) wi[i] = 0.0;
  ds->state = ST_CONDENSED;
  return info;
}

/* DSSortEigenvaluesReal_Private (dspriv.c:254-273) + DSPermuteColumns_Private (:278-300) */
static void ds_sort_hep(DSo *ds, double *wr)
{
  int n = ds->n, l = ds->l, ld = ds->ld, *perm = ds->perm;
  double *d = ds->T;
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int i = l + 1; i < ds->t; i++) {
    double re = d[perm[i]];
    int j = i - 1;
    int res = ds_compare(ds->which, re, 0.0, d[perm[j]], 0.0);
    while (res < 0 && j >= l) {
      int tmp = perm[j]; perm[j] = perm[j + 1]; perm[j + 1] = tmp; j--;
      if (j >= l) res = ds_compare(ds->which, re, 0.0, d[perm[j]], 0.0);
    }
  }
  for (int i = l; i < n; i++) wr[i] = d[perm[i]];
  for (int i = l; i < n; i++) {
    int p = perm[i];
    if (p != i) {
      int j = i + 1;
      while (perm[j] != i) j++;
      perm[j] = p; perm[i] = i;
      for (int k = 0; k < n; k++) { double t = ds->Q[k + p * ld]; ds->Q[k + p * ld] = ds->Q[k + i * ld]; ds->Q[k + i * ld] = t; }
    }
  }
  for (int i = l; i < n; i++) d[i] = wr[i];
}

/* DSUpdateExtraRow_HEP compact (dshep.c:349-381) */
static void ds_update_extra_row_hep(DSo *ds)
{
  int n = ds->n, ld = ds->ld;
  double *e = ds->T + ld, beta = e[n - 1];
  for (int i = 0; i < n; i++) e[i] = beta * ds->Q[n - 1 + i * ld];
  ds->k = n;
}

/* DSSolve_NHEP_Private (dsutil.c:21-90) */
static int ds_solve_nhep(DSo *ds, double *wr, double *wi)
{
  int n = ds->n, ld = ds->ld, ilo = ds->l + 1, info = 0, lwork = 6 * ld;
  double *A = ds->A, *Q = ds->Q, *tau = ds->work, *work = ds->work + ld;
  memset(Q, 0, sizeof(double) * (size_t)ld * ld);
  for (int i = 0; i < n; i++) Q[i + i * ld] = 1.0;
  if (n == 1) { wr[0] = A[0]; wi[0] = 0.0; ds->state = ST_CONDENSED; return 0; }
  if (ds->state < ST_INTERMEDIATE) {
    LP.gehrd(&n, &ilo, &n, A, &ld, tau, work, &lwork, &info);
    if (info) return info;
    for (int j = 0; j < n - 1; j++)
      for (int i = j + 2; i < n; i++) { Q[i + j * ld] = A[i + j * ld]; A[i + j * ld] = 0.0; }
    LP.orghr(&n, &ilo, &n, Q, &ld, tau, work, &lwork, &info);
    if (info) return info;
  }
  LP.hseqr("S", "V", &n, &ilo, &n, A, &ld, wr, wi, Q, &ld, work, &lwork, &info, 1, 1);
  for (int j = 0; j < ds->l; j++) {
    if (j == n - 1 || A[j + 1 + j * ld] == 0.0) { wr[j] = A[j + j * ld]; wi[j] = 0.0; }
    else {
      wr[j] = A[j + j * ld]; wr[j + 1] = A[j + j * ld];
      wi[j] = sqrt(fabs(A[j + 1 + j * ld])) * sqrt(fabs(A[j + (j + 1) * ld]));
      wi[j + 1] = -wi[j];
      j++;
    }
  }
  ds->state = ST_CONDENSED;
  return info;
}

/* DSSort_NHEP_Total (dsutil.c:93-166): selection sort with trexc */
static int ds_sort_nhep(DSo *ds, double *wr, double *wi)
{
  int n = ds->n, ld = ds->ld, info = 0;
  double *T = ds->A;
  for (int i = ds->l; i < n - 1; i++) {
    double re = wr[i], im = wi[i];
    int pos = 0, j = i + 1;
    if (im != 0) j = i + 2;
    for (; j < n; j++) {
      if (ds_compare(ds->which, re, im, wr[j], wi[j]) > 0) { re = wr[j]; im = wi[j]; pos = j; }
      if (wi[j] != 0) j++;
    }
    if (pos) {
      int ifst = pos + 1, ilst = i + 1;
      LP.trexc("V", &n, T, &ld, ds->Q, &ld, &ifst, &ilst, ds->work, &info, 1);
      if (info) return info;
      for (j = i; j < n; j++) {
        wr[j] = T[j + j * ld];
        if (j < n - 1 && T[j + 1 + j * ld] != 0.0) {
          wi[j] = sqrt(fabs(T[j + 1 + j * ld])) * sqrt(fabs(T[j + (j + 1) * ld]));
          wr[j + 1] = wr[j]; wi[j + 1] = -wi[j]; j++;
        } else wi[j] = 0.0;
      }
    }
    if (wi[i] != 0) i++;
  }
  return 0;
}

/* DSUpdateExtraRow_NHEP (dsnhep.c:318-340): b := Q^T b */
static void ds_update_extra_row_nhep(DSo *ds)
{
  int n = ds->n, ld = ds->ld;
  double *x = ds->work, *y = ds->work + ld;
  for (int i = 0; i < n; i++) x[i] = ds->A[n + i * ld];
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int r = 0; r < n; r++) s += ds->Q[r + i * ld] * x[r];
    y[i] = s;
  }
  for (int i = 0; i < n; i++) ds->A[n + i * ld] = y[i];
  ds->k = n;
}

/* DSVectors_NHEP_Eigen_Some (dsnhep.c:101-166): returns |last comp| and flag for complex pair */
static int ds_vector_nhep_some(DSo *ds, int k, double *rnorm, int *iscomplex)
{
  int n = ds->n, ld = ds->ld, mm, mout, info = 0;
  int *select = ds->iwork;
  for (int i = 0; i < n; i++) select[i] = 0;
  double *Y = ds->X + (int64_t)k * ld;
  select[k] = 1;
  *iscomplex = (k < n - 1 && ds->A[k + 1 + k * ld] != 0.0);
  mm = *iscomplex ? 2 : 1;
  if (*iscomplex) select[k + 1] = 1;
  LP.trevc("R", "S", select, &n, ds->A, &ld, Y, &ld, Y, &ld, &mm, &mout, ds->work, &info, 1, 1);
  if (info) return info;
  double *tmp = ds->work;
  memcpy(tmp, Y, sizeof(double) * (size_t)mout * ld);
  for (int c = 0; c < mout; c++)
    for (int r = 0; r < n; r++) {
      double s = 0.0;
      for (int q = 0; q < n; q++) s += ds->Q[r + q * ld] * tmp[q + c * ld];
      Y[r + c * ld] = s;
    }
  double nr = 0.0, ni = 0.0;
  for (int r = 0; r < n; r++) nr += Y[r] * Y[r];
  nr = sqrt(nr);
  double norm = nr;
  if (*iscomplex) { for (int r = 0; r < n; r++) ni += Y[r + ld] * Y[r + ld]; norm = hypot(nr, sqrt(ni)); }
  for (int c = 0; c < mout; c++) for (int r = 0; r < n; r++) Y[r + c * ld] /= norm;
  *rnorm = *iscomplex ? hypot(Y[n - 1], Y[n - 1 + ld]) : fabs(Y[n - 1]);
  return 0;
}

/* DSVectors_NHEP_Eigen_All (dsnhep.c:168-226): X = eigvecs(T), back-transformed
   with Q only while the DS is condensed (a trimmed DS is RAW: dsops.c:250) */
static int ds_vectors_nhep_all(DSo *ds, int back_q)
{
  int n = ds->n, ld = ds->ld, mout, info = 0;
  if (back_q) memcpy(ds->X, ds->Q, sizeof(double) * (size_t)ld * ld);
  LP.trevc("R", back_q ? "B" : "A", NULL, &n, ds->A, &ld, NULL, &ld, ds->X, &ld, &n, &mout, ds->work, &info, 1, 1);
  if (info) return info;
  for (int i = 0; i < n; i++) {
    int cplx = (i < n - 1 && ds->A[i + 1 + i * ld] != 0.0);
    double s = 0.0, s2 = 0.0;
    for (int r = 0; r < n; r++) s += ds->X[r + i * ld] * ds->X[r + i * ld];
    double norm = sqrt(s);
    if (cplx) { for (int r = 0; r < n; r++) s2 += ds->X[r + (i + 1) * ld] * ds->X[r + (i + 1) * ld]; norm = hypot(norm, sqrt(s2)); }
    for (int c = 0; c < (cplx ? 2 : 1); c++) for (int r = 0; r < n; r++) ds->X[r + (i + c) * ld] /= norm;
    if (cplx) i++;
  }
  return 0;
}

/* ======================================================================== */
/* Krylov-Schur driver                                                        */
/* ======================================================================== */
static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

typedef struct {
  int n;
  const int32_t *rp, *ci;
  const double *val;
  double *V; int64_t ld; int m;  /* BV: n x m, m = ncv+1 */
  double *buf;                   /* m x m coefficient buffer */
  double eta;
  int64_t matvecs, passes;
  uint64_t seed;                 /* random start vectors: seed + i (0: 0x5EED) */
  const double *lu;              /* shift-and-invert: LU of A - sigma I (n x n), or NULL */
  const int *piv;
  int mgs;                       /* BV_ORTHOG_MGS (EPSLanczos other than full, lanczos.c:58-62) */
  char *which;                   /* column mask of the local Lanczos sweep, m entries */
  int refine;                    /* BVSetOrthogonalization refinement: ORC_REFINE_* */
} KSo;

/* Orthonormalise column i against the columns before it (BVOrthonormalizeColumn,
   bvorthog.c:380-426, without replacement): lindep, norm; scaled unless the
   norm is 0 or 1. */
static void ks_orthonormalize(KSo *ks, int i, double *norm, int *lindep)
{
  if (ks->mgs) orthogonalize_mgs(ks->n, ks->V, ks->ld, i, NULL, ks->buf, ks->m, ks->eta, ks->refine, norm, lindep, &ks->passes);
  else orthogonalize_gs(ks->n, ks->V, ks->ld, i, ks->buf, ks->m, ks->eta, ks->refine, norm, lindep, &ks->passes);
}

/* EPSGetStartVector (epssolve.c:841-873) with a caller-provided vector for i==0
   or a deterministic random vector (seed 0x5EED + i) otherwise. */
static int ks_start_vector(KSo *ks, int i, int nini, int *breakdown)
{
  double *vi = ks->V + (int64_t)i * ks->ld;
  if (i > 0 || nini == 0) orc_rand_uniform((ks->seed ? ks->seed : 0x5EEDULL) + (uint64_t)i, ks->n, vi);
  double norm; int lindep;
  ks_orthonormalize(ks, i, &norm, &lindep);
  if (breakdown) *breakdown = lindep;
  else if (lindep || norm == 0.0) return -1;
  orc_bv_scale(ks->n, vi, 1.0 / norm);
  return 0;
}

/* EPSSetTrueResidual branch of EPSKrylovConvergence (epskrylov.c:256-264):
   EPSComputeRitzVector (epsdefault.c:386-398) x = V(:,0:nv) z, the imaginary
   part V(:,0:nv) z2 for a pair, then ||A x - lambda x|| with the
   back-transformed lambda. */
double orc_residual_abs(int n, const int32_t *rp, const int32_t *ci, const double *val, double kr, double ki,
                        const double *xr, const double *xi);
static double ks_true_residual(const KSo *ks, int nv, const double *z, const double *z2, double re, double im)
{
  const int n = ks->n;
  double *x = (double *)malloc(sizeof(double) * (size_t)n * 2);
  orc_bv_multvec(n, nv, 1.0, ks->V, ks->ld, z, 0.0, x);
  if (z2) orc_bv_multvec(n, nv, 1.0, ks->V, ks->ld, z2, 0.0, x + n);
  const double r = orc_residual_abs(n, ks->rp, ks->ci, ks->val, re, z2 ? im : 0.0, x, z2 ? x + n : NULL);
  free(x);
  return r;
}

/* BVMatArnoldi / BVMatLanczos (bvkrylov.c:56-113,165-226) */
/* y = Op x: A x, or (A - sigma I)^-1 x for STSINVERT with PREONLY + LU */
static void ks_op(KSo *ks, const double *x, double *y)
{
  if (ks->lu) {
    int nn = ks->n, one = 1, info = 0;
    memcpy(y, x, sizeof(double) * (size_t)nn);
    LP.getrs("N", &nn, &one, ks->lu, &nn, ks->piv, y, &nn, &info, 1);
  } else {
    orc_spmv(ks->n, ks->rp, ks->ci, ks->val, x, y);
  }
  ks->matvecs++;
}

/* EPSLocalLanczos (lanczos.c:89-118): column j+1 = Op v_j orthogonalised by
   MGS against the columns before k and v_{j-1}, v_j only; the coefficient
   buffer gets alpha_j (row j) and beta_j = the norm (row j+1, also on
   breakdown) of column j+1. */
static void ks_expand_local(KSo *ks, int k, int *m, double *beta, int *breakdown)
{
  int lindep = 0;
  char *which = ks->which;
  for (int i = 0; i < k; i++) which[i] = 1;
  for (int j = k; j < *m; j++) {
    ks_op(ks, ks->V + (int64_t)j * ks->ld, ks->V + (int64_t)(j + 1) * ks->ld);
    which[j] = 1;
    if (j - 2 >= k) which[j - 2] = 0;
    double nrm;
    orthogonalize_mgs(ks->n, ks->V, ks->ld, j + 1, which, ks->buf, ks->m, ks->eta, ks->refine, &nrm, &lindep, &ks->passes);
    ks->buf[j + 1 + (int64_t)(j + 1) * ks->m] = nrm;
    *beta = nrm;
    if (lindep) { *m = j + 1; break; }
    orc_bv_scale(ks->n, ks->V + (int64_t)(j + 1) * ks->ld, 1.0 / nrm);
  }
  *breakdown = lindep;
}

/* update_omega (lanczos.c:260-293): the omega recurrence of partial
   reorthogonalization.  B(x) is the reference's beta argument, which the
   caller passes shifted by one (lanczos.c:398), so B(x) = beta[x - 1]; alpha
   and omega are indexed from 0 whatever the number of locked columns.  The
   final exchange of the reference copies omega_old into omega and leaves
   omega_old as it is (its second assignment reads the value just written). */
static void update_omega(double *omega, double *omega_old, int j, const double *alpha, const double *beta,
                         double eps1, double anorm)
{
#define B(x) beta[(x) - 1]
  const double T = eps1 * anorm, binv = 1.0 / B(j + 1);
  omega_old[0] = B(1) * omega[1] + (alpha[0] - alpha[j]) * omega[0] - B(j) * omega_old[0];
  omega_old[0] = omega_old[0] > 0 ? binv * (omega_old[0] + T) : binv * (omega_old[0] - T);
  for (int q = 1; q < j - 1; q++) {
    omega_old[q] = B(q + 1) * omega[q + 1] + (alpha[q] - alpha[j]) * omega[q] + B(q) * omega[q - 1] - B(j) * omega_old[q];
    omega_old[q] = omega_old[q] > 0 ? binv * (omega_old[q] + T) : binv * (omega_old[q] - T);
  }
  omega_old[j - 1] = binv * T;
  for (int q = 0; q < j; q++) omega[q] = omega_old[q];
  omega[j] = eps1;
#undef B
}

/* compute_int (lanczos.c:295-333): which[i] for |mu_i| >= delta (else the
   largest), widened left and right while |mu| >= eta. */
static void compute_int(char *which, const double *mu, int j, double delta, double eta)
{
  int found = 0, maxpos = 0;
  double mx = 0.0;
  for (int i = 0; i < j; i++) {
    if (fabs(mu[i]) >= delta) { which[i] = 1; found = 1; }
    else which[i] = 0;
    if (fabs(mu[i]) > mx) { maxpos = i; mx = fabs(mu[i]); }
  }
  if (!found) which[maxpos] = 1;
  for (int i = 0; i < j; i++) {
    if (which[i]) {
      for (int q = i; q >= 0; q--) {
        if (fabs(mu[q]) < eta || which[q]) break;
        which[q] = 1;
      }
      for (int q = i + 1; q < j; q++) {
        if (fabs(mu[q]) < eta || which[q]) break;
        which[q] = 1;
      }
    }
  }
}

/* EPSPartialLanczos (lanczos.c:338-441), periodic or partial: a local step
   (MGS against the locked columns and v_{j-1}, v_j), the omega estimate of the
   loss of orthogonality, and when it exceeds delta a second MGS against the
   columns k..j (periodic: all of them, and the next step too; partial: the
   intervals of compute_int); once norm * delta < anorm * eps1 the rest of
   the sweep orthogonalises against every column.  alpha, beta: the projected
   problem's diagonal and off-diagonal (locked entries included, as the
   recurrence reads them).  anorm < 0: estimated from the recurrence. */
static void ks_expand_partial(KSo *ks, int k, int *M, double *betam, int *breakdown, double anorm, int periodic,
                              double *alpha, double *beta)
{
  const int m = *M, n = ks->n, ncv = ks->m - 1, sz = (m > 100 ? m : 100) + 1;
  const double meps = 2.220446049250313e-16;
  const double sqeps = 1.490116119384766e-08;   /* PETSc's PETSC_SQRT_MACHINE_EPSILON (double) */
  const double eps1 = sqrt((double)n) * meps / 2, delta = sqeps / sqrt((double)ncv);
  const double eta = pow(meps, 0.75) / sqrt((double)ncv);
  double *omega = (double *)calloc(sz, sizeof(double)), *omega_old = (double *)calloc(sz, sizeof(double));
  char *which = (char *)calloc(sz, 1), *which2 = (char *)calloc(sz, 1);
  int estimate = 0, force = 0, fro = 0, lindep = 0;
  if (anorm < 0.0) { anorm = 1.0; estimate = 1; }
  for (int i = 0; i < k; i++) which[i] = 1;
  double norm = 0.0;
  for (int j = k; j < m; j++) {
    ks_op(ks, ks->V + (int64_t)j * ks->ld, ks->V + (int64_t)(j + 1) * ks->ld);
    const double *h = ks->buf + (int64_t)(j + 1) * ks->m;
    if (fro) {
      orthogonalize_mgs(n, ks->V, ks->ld, j + 1, NULL, ks->buf, ks->m, ks->eta, ks->refine, &norm, &lindep, &ks->passes);
      alpha[j] = h[j];
    } else {
      which[j] = 1;
      if (j - 2 >= k) which[j - 2] = 0;
      orthogonalize_mgs(n, ks->V, ks->ld, j + 1, which, ks->buf, ks->m, ks->eta, ks->refine, &norm, &lindep, &ks->passes);
      alpha[j] = h[j];
      beta[j] = norm;
      if (estimate) {
        if (j > k) anorm = fmax(anorm, fabs(alpha[j]) + norm + beta[j - 1]);
        else anorm = fmax(anorm, fabs(alpha[j]) + norm);
      }
      int reorth = 0;
      if (j > k) {
        update_omega(omega, omega_old, j, alpha, beta, eps1, anorm);
        for (int i = 0; i < j - k; i++) if (fabs(omega[i]) > delta) reorth = 1;
      }
      if (reorth || force) {
        for (int i = 0; i < k; i++) which2[i] = 0;
        for (int i = k; i <= j; i++) which2[i] = 1;
        if (periodic) {
          force = !force;
          for (int i = 0; i < j - k; i++) omega[i] = eps1;
        } else if (force) {
          force = 0;
        } else {
          force = 1;
          compute_int(which2 + k, omega, j - k, delta, eta);
          for (int i = 0; i < j - k; i++) if (which2[i + k]) omega[i] = eps1;
        }
        orthogonalize_mgs(n, ks->V, ks->ld, j + 1, which2, ks->buf, ks->m, ks->eta, ks->refine, &norm, &lindep, &ks->passes);
      }
    }
    if (lindep || norm < n * anorm * meps) { *M = j + 1; break; }
    if (!fro && norm * delta < anorm * eps1) fro = 1;
    beta[j] = norm;
    orc_bv_scale(n, ks->V + (int64_t)(j + 1) * ks->ld, 1.0 / norm);
  }
  *betam = beta[*M - 1];
  *breakdown = lindep;
  free(omega); free(omega_old); free(which); free(which2);
}

/* BVOrthogonalizeVec (bvorthog.c:145-215 with a vector) of w against the
   columns 0..nc-1 of W, MGS with refinement IFNEEDED. */
static void mgs_vec(int64_t n, const double *W, int64_t ld, int nc, double *w, double eta, double *norm,
                    int *lindep)
{
  double onrm = orc_bv_norm2(n, w), nrm;
  for (int pass = 0; pass < 3; pass++) {
    if (pass > 0) {
      if (!(nrm != 0.0 && fabs(nrm) < eta * fabs(onrm))) break;
      onrm = nrm;
    }
    for (int i = 0; i < nc; i++) {
      const double *vi = W + (int64_t)i * ld;
      double dot = 0.0;
      for (int64_t r = 0; r < n; r++) dot += w[r] * vi[r];
      for (int64_t r = 0; r < n; r++) w[r] -= dot * vi[r];
    }
    nrm = orc_bv_norm2(n, w);
  }
  *lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
  *norm = nrm;
}

/* EPSSelectiveLanczos (lanczos.c:181-258): a local step, then the
   eigen-decomposition (stevr, DenseTridiagonal lanczos.c:139-179) of the
   tridiagonal block k..j; Ritz vectors whose bound norm |y_last| is below
   sqrt(eps) ||A|| are formed in AV (again only when their number grew), and
   the new vector is orthogonalised against them.  beta_j keeps the norm of
   the local step; the vector is scaled by the final norm. */
static int ks_expand_selective(KSo *ks, int k, int *M, double *betam, int *breakdown, double anorm, double *alpha,
                               double *beta, double *AV)
{
  const int m = *M, n = ks->n;
  const double sqeps = 1.490116119384766e-08;   /* PETSc's PETSC_SQRT_MACHINE_EPSILON (double) */
  double *d = (double *)calloc(m + 1, sizeof(double)), *e = (double *)calloc(m + 1, sizeof(double));
  double *ritz = (double *)calloc(m + 1, sizeof(double)), *Y = (double *)calloc((size_t)m * m + 1, sizeof(double));
  double *work = (double *)calloc(20 * (size_t)m + 1, sizeof(double));
  int *iwork = (int *)calloc(10 * (size_t)m + 1, sizeof(int)), *isuppz = (int *)calloc(2 * (size_t)m + 2, sizeof(int));
  char *which = (char *)calloc(m + 1, 1);
  int nritz = 0, lindep = 0, rc = 0;
  double norm = 0.0;
  for (int i = 0; i < k; i++) which[i] = 1;
  for (int j = k; j < m; j++) {
    double *vj1 = ks->V + (int64_t)(j + 1) * ks->ld;
    ks_op(ks, ks->V + (int64_t)j * ks->ld, vj1);
    which[j] = 1;
    if (j - 2 >= k) which[j - 2] = 0;
    orthogonalize_mgs(n, ks->V, ks->ld, j + 1, which, ks->buf, ks->m, ks->eta, ks->refine, &norm, &lindep, &ks->passes);
    alpha[j] = ks->buf[j + (int64_t)(j + 1) * ks->m];
    beta[j] = norm;
    if (lindep) { *M = j + 1; break; }
    int nn = j - k + 1, mf = 0, info = 0, lw = 20 * nn, liw = 10 * nn, il = 0, iu = 0;
    const double vl = 0.0, vu = 0.0, abstol = 0.0;
    for (int i = 0; i < nn; i++) { d[i] = alpha[i + k]; e[i] = beta[i + k]; }
    LP.stevr("V", "A", &nn, d, e, &vl, &vu, &il, &iu, &abstol, &mf, ritz, Y, &nn, isuppz, work, &lw, iwork, &liw,
             &info, 1, 1);
    if (info) { rc = -1; break; }
    for (int i = 0; i < nn; i++) if (fabs(ritz[i]) > anorm) anorm = fabs(ritz[i]);
    int nritzo = 0;
    for (int i = 0; i < nn; i++) if (norm * fabs(Y[(int64_t)i * nn + nn - 1]) < sqeps * anorm) nritzo++;
    if (nritzo > nritz) {
      nritz = 0;
      for (int i = 0; i < nn; i++)
        if (norm * fabs(Y[(int64_t)i * nn + nn - 1]) < sqeps * anorm) {
          orc_bv_multvec(n, nn, 1.0, ks->V + (int64_t)k * ks->ld, ks->ld, Y + (int64_t)i * nn, 0.0,
                         AV + (int64_t)nritz * ks->ld);
          nritz++;
        }
    }
    if (nritz > 0) {
      mgs_vec(n, AV, ks->ld, nritz, vj1, ks->eta, &norm, &lindep);
      ks->passes++;
      if (lindep) { *M = j + 1; break; }
    }
    orc_bv_scale(n, vj1, 1.0 / norm);
  }
  *betam = beta[*M - 1];
  *breakdown = lindep;
  free(d); free(e); free(ritz); free(Y); free(work); free(iwork); free(isuppz); free(which);
  return rc;
}

/* EPSDelayedArnoldi (epskrylov.c:24-114; lanczos.c:495-512 with refinement
   other than NEVER): one CGS pass per step whose normalisation is delayed by
   one step and whose reorthogonalisation by two.  At step j the dot products
   of v_{j+1} and v_j with the earlier columns are taken before v_j is
   normalised, the coefficients rescaled; the unnormalised v_j (t) is
   reorthogonalised against v_0..v_{j-1} and, one step later, normalised into
   column j.  H: leading dimension ldh >= m (Lanczos: a scratch T; Arnoldi:
   the projected problem's A). */
static void ks_expand_delayed(KSo *ks, int k, int m, double *H, int ldh, double *betam)
{
  const int64_t N = ks->n, ld = ks->ld;
  double *V = ks->V;
  double *u = (double *)calloc((size_t)N, sizeof(double)), *t = (double *)calloc((size_t)N, sizeof(double));
  double *lhh = (double *)calloc((size_t)m + 1, sizeof(double));
  double norm1 = 0.0, norm2 = 1.0;
#define COL(c) (V + (int64_t)(c) * ld)
  for (int j = k; j < m; j++) {
    ks_op(ks, COL(j), COL(j + 1));
    orc_bv_dotvec(N, j + 1, V, ld, COL(j + 1), H + (int64_t)ldh * j);   /* BVDotColumn(V, j+1) */
    if (j > k) {
      orc_bv_dotvec(N, j, V, ld, COL(j), lhh);                          /* BVDotColumn(V, j) */
      double dot = 0.0;
      for (int64_t r = 0; r < N; r++) dot += COL(j)[r] * COL(j)[r];
      if (j > k + 1) norm2 = orc_bv_norm2(N, u);
      norm1 = sqrt(dot);
      for (int i = 0; i < j; i++) H[(int64_t)ldh * j + i] /= norm1;
      H[(int64_t)ldh * j + j] /= dot;
      memcpy(t, COL(j), sizeof(double) * (size_t)N);
      orc_bv_scale(N, COL(j), 1.0 / norm1);
      orc_bv_scale(N, COL(j + 1), 1.0 / norm1);
    }
    orc_bv_multvec(N, j + 1, -1.0, V, ld, H + (int64_t)ldh * j, 1.0, COL(j + 1));   /* BVMultColumn(V, j+1) */
    if (j > k) {
      orc_bv_multvec(N, j, -1.0, V, ld, lhh, 1.0, t);
      for (int i = 0; i < j; i++) H[(int64_t)ldh * (j - 1) + i] += lhh[i];
    }
    if (j > k + 1) {
      memcpy(COL(j - 1), u, sizeof(double) * (size_t)N);
      orc_bv_scale(N, COL(j - 1), 1.0 / norm2);
      H[(int64_t)ldh * (j - 2) + j - 1] = norm2;
    }
    if (j < m - 1) memcpy(u, t, sizeof(double) * (size_t)N);
  }
  norm2 = orc_bv_norm2(N, t);
  orc_bv_scale(N, t, 1.0 / norm2);
  memcpy(COL(m - 1), t, sizeof(double) * (size_t)N);
  H[(int64_t)ldh * (m - 2) + m - 1] = norm2;
  orc_bv_dotvec(N, m, V, ld, COL(m), lhh);
  orc_bv_multvec(N, m, -1.0, V, ld, lhh, 1.0, COL(m));
  for (int i = 0; i < m; i++) H[(int64_t)ldh * (m - 1) + i] += lhh[i];
  *betam = orc_bv_norm2(N, COL(m));
  orc_bv_scale(N, COL(m), 1.0 / *betam);
#undef COL
  free(u); free(t); free(lhh);
}

/* EPSDelayedArnoldi1 (epskrylov.c:120-157; arnoldi.c:84-87 and lanczos.c:500
   with refinement NEVER): EPSDelayedArnoldi without reorthogonalisation, the
   normalisation alone delayed by one step.  At step j > k the dot products of
   v_{j+1} with v_0..v_j are taken before v_j is normalised and rescaled by its
   norm (by its square for the diagonal entry); v_j and v_{j+1} are scaled by
   1/||v_j||, H(j,j-1) = ||v_j|| and beta = ||v_j||; v_{j+1} -= V(:,0:j+1) H(:,j).
   The last column is left unnormalised and no breakdown is reported. */
static void ks_expand_delayed1(KSo *ks, int k, int m, double *H, int ldh, double *betam)
{
  const int64_t N = ks->n, ld = ks->ld;
  double *V = ks->V;
#define COL(c) (V + (int64_t)(c) * ld)
  for (int j = k; j < m; j++) {
    ks_op(ks, COL(j), COL(j + 1));
    double *Hj = H + (int64_t)ldh * j;
    orc_bv_dotvec(N, j + 1, V, ld, COL(j + 1), Hj);                  /* BVDotColumn(V, j+1) */
    if (j > k) {
      double dot = 0.0;
      for (int64_t r = 0; r < N; r++) dot += COL(j)[r] * COL(j)[r];    /* VecDot(vj, vj) */
      const double norm = sqrt(dot);
      orc_bv_scale(N, COL(j), 1.0 / norm);
      H[(int64_t)ldh * (j - 1) + j] = norm;
      for (int i = 0; i < j; i++) Hj[i] /= norm;
      Hj[j] /= dot;
      orc_bv_scale(N, COL(j + 1), 1.0 / norm);
      *betam = norm;
    }
    orc_bv_multvec(N, j + 1, -1.0, V, ld, Hj, 1.0, COL(j + 1));      /* BVMultColumn(V, j+1) */
  }
#undef COL
}

/* One local Lanczos sweep from v0 (test entry: exposes ks_expand_local):
   V (n x (m+1)) gets the basis, alpha/beta (m entries) the tridiagonal.
   Returns the number of steps taken. */
int orc_lanczos_local_sweep(int n, const int32_t *rp, const int32_t *ci, const double *val, int m, const double *v0,
                            double *V, double *alpha, double *beta)
{
  KSo ks = {0};
  ks.n = n; ks.rp = rp; ks.ci = ci; ks.val = val; ks.m = m + 1; ks.ld = n; ks.eta = 0.7071; ks.mgs = 1;
  ks.V = V;
  ks.buf = (double *)calloc((size_t)(m + 1) * (m + 1), sizeof(double));
  ks.which = (char *)calloc((size_t)m + 2, 1);
  memcpy(V, v0, sizeof(double) * (size_t)n);
  orc_bv_scale(n, V, 1.0 / orc_bv_norm2(n, V));
  int mm = m, bd = 0;
  double b = 0.0;
  ks_expand_local(&ks, 0, &mm, &b, &bd);
  for (int j = 0; j < mm; j++) {
    alpha[j] = ks.buf[j + (int64_t)(j + 1) * ks.m];
    beta[j] = ks.buf[j + 1 + (int64_t)(j + 1) * ks.m];
  }
  free(ks.buf); free(ks.which);
  return mm;
}

static void ks_expand(KSo *ks, int k, int *m, double *beta, int *breakdown)
{
  int lindep = 0;
  for (int j = k; j < *m; j++) {
    ks_op(ks, ks->V + (int64_t)j * ks->ld, ks->V + (int64_t)(j + 1) * ks->ld);
    double nrm;
    orthogonalize_gs(ks->n, ks->V, ks->ld, j + 1, ks->buf, ks->m, ks->eta, ks->refine, &nrm, &lindep, &ks->passes);
    if (nrm != 1.0 && nrm != 0.0) orc_bv_scale(ks->n, ks->V + (int64_t)(j + 1) * ks->ld, 1.0 / nrm);
    *beta = nrm;
    if (lindep) { *m = j + 1; break; }
  }
  *breakdown = lindep;
}

/* EPSBuildBalance_Krylov (epsdefault.c:443-504): the diagonal D for which
   D A D^-1 is balanced, from its random +-1 vectors z (here the oracle's
   uniform generator, seed 0xBA1A + j, mapped to [0,1), stands in for
   PetscRandom; a draw below 0.5 gives -1):
   p = D A (z ./ D); oneside: D ./= |p|; twoside: r = (A^T (z .* D)) ./ D and
   D .*= sqrt(|r ./ p|) where |p| > cutoff ||p_0||_inf and r != 0. */
static void build_balance(int n, const int32_t *rp, const int32_t *ci, const double *val, int twoside, int its,
                          double cutoff, double *D)
{
  double *z = (double *)malloc(sizeof(double) * (size_t)n), *r = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)malloc(sizeof(double) * (size_t)n), norma = 0.0;
  for (int i = 0; i < n; i++) D[i] = 1.0;
  for (int j = 0; j < its; j++) {
    orc_rand_uniform(0xBA1AULL + (uint64_t)j, n, z);
    /* a [0,1) draw (PetscRandom's default interval) from the U[-1,1) stream, signed at 0.5 */
    for (int i = 0; i < n; i++) z[i] = 0.5 * (z[i] + 1.0) < 0.5 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) r[i] = z[i] / D[i];
    orc_spmv(n, rp, ci, val, r, p);
    for (int i = 0; i < n; i++) p[i] *= D[i];
    if (twoside) {
      if (j == 0) { norma = 0.0; for (int i = 0; i < n; i++) { p[i] = fabs(p[i]); if (p[i] > norma) norma = p[i]; } }
      for (int i = 0; i < n; i++) z[i] *= D[i];
      for (int i = 0; i < n; i++) r[i] = 0.0;
      for (int row = 0; row < n; row++)                  /* r = A^T z */
        for (int32_t q = rp[row]; q < rp[row + 1]; q++) r[ci[q]] += val[q] * z[row];
      for (int i = 0; i < n; i++) r[i] /= D[i];
    }
    for (int i = 0; i < n; i++) {
      if (twoside) {
        if (fabs(p[i]) > cutoff * norma && r[i] != 0.0) D[i] *= sqrt(fabs(r[i] / p[i]));
      } else if (p[i] != 0.0) D[i] /= fabs(p[i]);
    }
  }
  free(z); free(r); free(p);
}

int orc_build_balance(int n, const int32_t *rp, const int32_t *ci, const double *val, int twoside, int its,
                      double cutoff, double *D)
{
  build_balance(n, rp, ci, val, twoside, its, cutoff, D);
  return 0;
}

/* EPSConvergedRelative / Absolute / Norm (epsdefault.c:224-256); for a
   standard problem ||B|| stays 0 (epssetup.c:351-359 computes it only with
   two matrices), so the norm test is ||r|| / ||A||_inf. */
static double conv_errest(int conv, double nrma, double re, double im, double res)
{
  const double w = hypot(re, im);
  if (conv == 1) return res;
  if (conv == 2) return res / (nrma + w * 0.0);
  return w != 0.0 ? res / w : 1.79e308;
}

/* MatNorm(A, NORM_INFINITY): the largest absolute row sum */
static double csr_norm_inf(int n, const int32_t *rp, const double *val)
{
  double m = 0.0;
  for (int r = 0; r < n; r++) {
    double t = 0.0;
    for (int32_t q = rp[r]; q < rp[r + 1]; q++) t += fabs(val[q]);
    if (t > m) m = t;
  }
  return m;
}

int orc_eps_solve(int n, const int32_t *rp, const int32_t *ci, const double *val, const orc_opts *o, const double *v0, orc_result *res)
{
  if (!lp_load()) return -10;
#ifdef _OPENMP
  int saved_threads = omp_get_max_threads();
  if (o->nthreads > 0) omp_set_num_threads(o->nthreads);
#endif
  int nev = o->nev > 0 ? o->nev : 1;
  if (nev > n) nev = n;                       /* small-problem safeguard, epssetup.c:342-345 */
  int ncv = o->ncv > 0 ? o->ncv : (nev + 15 > 2 * nev ? nev + 15 : 2 * nev);
  if (ncv > n) ncv = n;
  int mpd = o->mpd > 0 ? o->mpd : ncv;
  /* EPSSetDimensions_Default (epssetup.c:679) and krylovschur.c:88 */
  if (!(ncv >= nev + 1 || (ncv == nev && ncv == n)) || ncv > nev + mpd) return -16;
  int max_it = o->max_it > 0 ? o->max_it : (100 > 2 * n / ncv ? 100 : 2 * n / ncv);
  double tol = o->tol > 0 ? o->tol : 1e-8, keep = o->keep > 0 ? o->keep : 0.5;
  const double nrma = o->conv == 2 ? csr_norm_inf(n, rp, val) : 0.0;   /* epssetup.c:351-355 */
  int herm = o->hermitian, which = o->which ? o->which : ORC_LM, lock = o->lock;
  /* harmonic extraction (krylovschur.c:238-240; arnoldi.c:65): NHEP projected problem */
  const int harm = o->extraction == 1;
  if (harm) herm = 0;
  cmp_target = o->target;
  double *gh = (double *)calloc(ncv + 2, sizeof(double)), gamma = 1.0;
  /* shift-and-invert (sinvert.c, epssetup.c:119): dense LU of A - sigma I */
  const int sinv = o->st == 1;
  if (sinv && (which != ORC_TM && which != ORC_TR && which != ORC_USER)) return -12;
  if (sinv && harm) return -13;
  cmp_sinv = sinv; cmp_sigma = o->sigma;
  /* balancing (epssetup.c:399-405): non-Hermitian only; the balanced
     operator is formed explicitly, B = D A D^-1 */
  double *balD = NULL, *balval = NULL;
  if (o->balance && !herm) {
    if (sinv || harm || o->trueres) return -15;
    balD = (double *)malloc(sizeof(double) * (size_t)n);
    balval = (double *)malloc(sizeof(double) * (size_t)(rp[n] > 0 ? rp[n] : 1));
    build_balance(n, rp, ci, val, o->balance == 2, o->balance_its > 0 ? o->balance_its : 5,
                  o->balance_cutoff > 0 ? o->balance_cutoff : 1e-8, balD);
    for (int row = 0; row < n; row++)
      for (int32_t q = rp[row]; q < rp[row + 1]; q++) balval[q] = val[q] * balD[row] / balD[ci[q]];
    val = balval;
  }
  double *lu = NULL; int *piv = NULL;
  if (sinv) {
    if (!lp_load() || !LP.getrf || !LP.getrs) return -10;
    lu = (double *)calloc((size_t)n * n, sizeof(double)); piv = (int *)calloc(n, sizeof(int));
    for (int r = 0; r < n; r++) {
      for (int32_t q = rp[r]; q < rp[r + 1]; q++) lu[r + (int64_t)ci[q] * n] = val[q];
      lu[r + (int64_t)r * n] -= o->sigma;
    }
    int nn = n, info = 0;
    LP.getrf(&nn, &nn, lu, &nn, piv, &info);
    if (info) { free(lu); free(piv); return -14; }
  }
  /* EPSARNOLDI (arnoldi.c:52-131): explicit restart, always DSNHEP (:45);
     ncv <= nev + mpd (:181) */
  const int arn = o->solver == 1;
  if (arn) {
    if (ncv > nev + mpd) return -11;
    herm = 0;
  }
  /* EPSLANCZOS (lanczos.c:518-697): explicit restart from the most wanted
     unconverged Ritz vector, DSHEP; the sweep follows the reorthogonalization
     type (ks_expand for full, ks_expand_local / _partial / _selective /
     _delayed for the others) */
  const int lanc = o->solver == 2;
  if (lanc && (ncv > nev + mpd || !herm || harm)) return -11;
  if (lanc && o->lanczos_partial == 3 && !LP.stevr) return -10;
  double anorm = -1.0;
  double *bnd = (double *)calloc(ncv + 1, sizeof(double)), *yr = (double *)calloc(ncv + 1, sizeof(double));
  char *cv = (char *)calloc(ncv + 1, 1);
  int *lp = (int *)calloc(ncv + 1, sizeof(int));
  KSo ks = {0};
  ks.n = n; ks.rp = rp; ks.ci = ci; ks.val = val; ks.m = ncv + 1; ks.ld = n;
  ks.eta = o->eta > 0 ? o->eta : 0.7071;
  ks.refine = o->refine;
  ks.lu = lu; ks.piv = piv;
  ks.mgs = lanc && (o->lanczos_local || o->lanczos_partial);
  /* selective: the Ritz vectors' basis AV (BVDuplicate of V, lanczos.c:63-71) */
  double *AVs = NULL;
  if (lanc && o->lanczos_partial == 3) AVs = (double *)calloc((size_t)n * (ncv + 1), sizeof(double));
  ks.which = (char *)calloc(ks.m + 1, 1);
  ks.V = (double *)calloc((size_t)n * ks.m, sizeof(double));
  ks.buf = (double *)calloc((size_t)ks.m * ks.m, sizeof(double));
  int ld = ncv + 1;
  DSo ds = {0};
  ds.ld = ld; ds.hermitian = herm; ds.which = which;
  ds.A = (double *)calloc((size_t)ld * ld, sizeof(double));
  ds.Q = (double *)calloc((size_t)ld * ld, sizeof(double));
  ds.X = (double *)calloc((size_t)ld * ld, sizeof(double));
  ds.T = (double *)calloc((size_t)3 * ld, sizeof(double));
  ds.work = (double *)calloc((size_t)8 * ld * ld + 16 * ld, sizeof(double));
  ds.perm = (int *)calloc(ld, sizeof(int));
  ds.iwork = (int *)calloc(ld, sizeof(int));
  double *eigr = (double *)calloc(ld, sizeof(double)), *eigi = (double *)calloc(ld, sizeof(double));
  double *errest = (double *)calloc(ld, sizeof(double)), *lockd = (double *)calloc(ld, sizeof(double));
  int *perm = (int *)calloc(ld, sizeof(int));
  for (int i = 0; i < ncv; i++) perm[i] = i;
  int rc = 0, nini = 0;
  double t0 = now_s();

  /* EPSSetInitialSpace -> BVInsertVecs(V,0,&k,IS,TRUE) (epssetup.c:421-427, bvfunc.c:331-374) */
  if (v0) {
    memcpy(ks.V, v0, sizeof(double) * (size_t)n);
    double norm; int lindep;
    ks_orthonormalize(&ks, 0, &norm, &lindep);
    if (norm == 0.0 || lindep) { rc = -3; goto done; }
    orc_bv_scale(n, ks.V, 1.0 / norm);
    nini = 1;
  }
  if (ks_start_vector(&ks, 0, nini, NULL)) { rc = -3; goto done; }

  int l = 0, nconv = 0, its = 0, reason = 0;
  double *d = ds.T, *e = ds.T + ld;
  while (reason == 0) {
    its++;
    int nv = nconv + mpd < ncv ? nconv + mpd : ncv;
    ds.n = nv; ds.t = nv; ds.l = nconv; ds.k = (arn || lanc) ? 0 : nconv + l;   /* DSSetDimensions */
    double beta = 0.0; int breakdown = 0, k0 = nconv + l;
    const int partial = lanc ? o->lanczos_partial : 0;
    if (partial == 3) {
      if (ks_expand_selective(&ks, k0, &nv, &beta, &breakdown, anorm, d, e, AVs)) { rc = -20; goto done; }
    } else if (partial == 4) {                          /* lanczos.c:495-512 */
      if (nv - k0 < 2) { rc = -14; goto done; }
      double *Td = (double *)calloc((size_t)nv * nv, sizeof(double));
      if (o->refine == ORC_REFINE_NEVER) ks_expand_delayed1(&ks, k0, nv, Td, nv, &beta);   /* lanczos.c:500 */
      else ks_expand_delayed(&ks, k0, nv, Td, nv, &beta);
      breakdown = 0;
      for (int i = k0; i < nv - 1; i++) { d[i] = Td[(int64_t)nv * i + i]; e[i] = Td[(int64_t)nv * i + i + 1]; }
      d[nv - 1] = Td[(int64_t)nv * (nv - 1) + nv - 1];
      e[nv - 1] = beta;
      free(Td);
    } else if (partial) ks_expand_partial(&ks, k0, &nv, &beta, &breakdown, anorm, o->lanczos_partial == 1, d, e);
    else if (arn && o->arnoldi_delayed) {                /* arnoldi.c:88-91 */
      if (nv - k0 < 2) { rc = -14; goto done; }
      if (o->refine == ORC_REFINE_NEVER) ks_expand_delayed1(&ks, k0, nv, ds.A, ld, &beta);   /* arnoldi.c:84-87 */
      else ks_expand_delayed(&ks, k0, nv, ds.A, ld, &beta);
      breakdown = 0;
    }
    else if (ks.mgs) ks_expand_local(&ks, k0, &nv, &beta, &breakdown);
    else ks_expand(&ks, k0, &nv, &beta, &breakdown);
    int bm = ks.m;
    if (partial || (arn && o->arnoldi_delayed)) {
      /* the sweep wrote the projected problem itself */
    } else if (herm) {
      for (int j = k0; j < nv; j++) { d[j] = ks.buf[j + (int64_t)(j + 1) * bm]; e[j] = ks.buf[j + 1 + (int64_t)(j + 1) * bm]; }
    } else {
      double *H = ds.A;
      for (int j = k0; j < nv - 1; j++) memcpy(H + (int64_t)j * ld, ks.buf + (int64_t)(j + 1) * bm, sizeof(double) * (j + 2));
      memcpy(H + (int64_t)(nv - 1) * ld, ks.buf + (int64_t)nv * bm, sizeof(double) * nv);
      if (ld > nv) H[nv + (nv - 1) * ld] = ks.buf[nv + (int64_t)nv * bm];
    }
    ds.n = nv; ds.t = nv; ds.l = nconv; ds.k = (arn || lanc) ? 0 : nconv + l;
    ds.state = l ? ST_RAW : ST_INTERMEDIATE;
    if (harm && ds_translate_harmonic(&ds, o->target, beta, 0, gh, &gamma)) { rc = -70; goto done; }
    /* EPSLanczos keeps the projected eigenvalues (ritz) apart from eps->eigr
       and stores only the entries from nconv on (lanczos.c:663-666): the
       leading diagonal of the projected problem is not that of the locked
       values once the converged values are moved first */
    if (lanc) memcpy(lockd, eigr, sizeof(double) * (size_t)nconv);
    int info = herm ? ds_solve_hep(&ds, eigr, eigi) : ds_solve_nhep(&ds, eigr, eigi);
    if (info) { rc = -20 - info; goto done; }
    if (lanc) {
      ds_sort_hep(&ds, eigr);
      double *Y = ds.Q;
      for (int i = nconv; i < nv; i++) if (fabs(eigr[i]) > anorm) anorm = fabs(eigr[i]);
      for (int i = nconv; i < nv; i++) {                     /* beta |y_m| + eps ||A||, relative */
        double res = beta * fabs(Y[nv - 1 + (int64_t)i * ld]) + 2.220446049250313e-16 * anorm;
        bnd[i] = conv_errest(o->conv, nrma, eigr[i], 0.0, res);
        cv[i] = bnd[i] < tol ? 'C' : 'N';
      }
      if (o->lanczos_local)                                  /* purge repeated Ritz values (:570-575) */
        for (int i = nconv + 1; i < nv; i++)
          if (cv[i] == 'C' && fabs((eigr[i] - eigr[i - 1]) / eigr[i]) < tol) cv[i] = 'R';
      if (!breakdown) {                                      /* restart vector (lanczos.c:577-593) */
        int rs = nconv;
        while (rs < nv && cv[rs] != 'N') rs++;
        if (rs >= nv) breakdown = 1;
        else {
          for (int i = rs + 1; i < nv; i++)
            if (cv[i] == 'N' && sc_compare(which, eigr[rs], 0.0, eigr[i], 0.0) > 0) rs = i;
          for (int i = nconv; i < nv; i++) yr[i - nconv] = Y[i + (int64_t)rs * ld];
          orc_bv_multvec(n, nv - nconv, 1.0, ks.V + (int64_t)nconv * ks.ld, ks.ld, yr, 0.0, ks.V + (int64_t)nv * ks.ld);
        }
      }
      int k;                                                 /* converged first (:780-808) */
      for (int i = nconv; i < nv; i++) lp[i] = i;
      for (k = nconv; k < nv; k++) {
        if (cv[lp[k]] != 'C') {
          int j = k + 1;
          while (j < nv && cv[lp[j]] != 'C') j++;
          if (j >= nv) break;
          int t = lp[k]; lp[k] = lp[j]; lp[j] = t;
        }
      }
      for (int i = nconv; i < k; i++) {
        int x = lp[i];
        if (x != i) {
          int j = i + 1;
          while (lp[j] != i) j++;
          double td = eigr[x]; eigr[x] = eigr[i]; eigr[i] = td;
          td = bnd[x]; bnd[x] = bnd[i]; bnd[i] = td;
          char tc = cv[x]; cv[x] = cv[i]; cv[i] = tc;
          lp[j] = x; lp[i] = i;
          for (int q = 0; q < nv; q++) { td = Y[q + (int64_t)x * ld]; Y[q + (int64_t)x * ld] = Y[q + (int64_t)i * ld]; Y[q + (int64_t)i * ld] = td; }
        }
      }
      if (k > nconv) orc_bv_multinplace(n, ks.V, ks.ld, nconv, nv, nconv, k, Y, ld);
      if (o->lanczos_local) {                                /* purge spurious Ritz values (:633-660) */
        double *w = (double *)malloc(sizeof(double) * (size_t)n);
        for (int i = nconv; i < k; i++) {
          double *vi = ks.V + (int64_t)i * ks.ld;
          orc_bv_scale(n, vi, 1.0 / orc_bv_norm2(n, vi));
          ks_op(&ks, vi, w);                                 /* STApply: the operator of the ST */
          for (int64_t q = 0; q < n; q++) w[q] -= eigr[i] * vi[q];
          double nr = orc_bv_norm2(n, w);
          bnd[i] = conv_errest(o->conv, nrma, eigr[i], 0.0, nr);
          if (bnd[i] >= tol) cv[i] = 'S';
        }
        int i;
        for (i = nconv; i < k; i++) {
          if (cv[i] != 'C') {
            int j = i + 1;
            while (j < k && cv[j] != 'C') j++;
            if (j >= k) break;
            double td = eigr[j]; eigr[j] = eigr[i]; eigr[i] = td;
            td = bnd[j]; bnd[j] = bnd[i]; bnd[i] = td;
            char tc = cv[j]; cv[j] = cv[i]; cv[i] = tc;
            double *a = ks.V + (int64_t)i * ks.ld, *b = ks.V + (int64_t)j * ks.ld;
            for (int64_t q = 0; q < n; q++) { td = a[q]; a[q] = b[q]; b[q] = td; }
          }
        }
        k = i;
        free(w);
      }
      for (int i = nconv; i < nv; i++) errest[i] = bnd[i];
      memcpy(eigr, lockd, sizeof(double) * (size_t)nconv);
      nconv = k;
      if (res->trace_nconv && its <= res->trace_cap) { res->trace_nconv[its - 1] = k; res->trace_eig0[its - 1] = eigr[0]; }
      if (k >= nev) reason = 1;
      else if (its >= max_it) reason = -1;
      if (reason == 0) {                                     /* BVCopyColumn(V,nv,nconv) (:826) */
        memcpy(ks.V + (int64_t)nconv * ks.ld, ks.V + (int64_t)nv * ks.ld, sizeof(double) * (size_t)n);
        if (o->lanczos_local && !breakdown) {                /* reorthonormalise the restart vector (:678-681) */
          double nrm; int lind;
          ks_orthonormalize(&ks, nconv, &nrm, &lind);
          if (nrm != 1.0 && nrm != 0.0) orc_bv_scale(n, ks.V + (int64_t)nconv * ks.ld, 1.0 / nrm);
          if (lind) breakdown = 1;
        }
        if (breakdown) {
          int bd = 0;
          ks_start_vector(&ks, nconv, nini, &bd);
          if (bd) reason = -2;
        }
      }
      continue;
    }
    if (herm) { ds_sort_hep(&ds, eigr); ds_update_extra_row_hep(&ds); }
    else {
      if ((info = ds_sort_nhep(&ds, eigr, eigi))) { rc = -40; goto done; }
      ds_update_extra_row_nhep(&ds);
    }
    /* EPSKrylovConvergence (epskrylov.c:207-295) */
    int k, marker = -1;
    for (k = nconv; k < nv; k++) {
      double re = eigr[k], im = eigi[k], resnorm;
      int cplx = 0;
      if (herm) resnorm = fabs(ds.Q[nv - 1 + k * ld]);
      else if ((info = ds_vector_nhep_some(&ds, k, &resnorm, &cplx))) { rc = -50; goto done; }
      /* STBackTransform (sinvert.c:16-38) for the true residual and the norm
         test; the other tests take the transformed value (epskrylov.c:245-251) */
      if (o->st == 1 && (o->trueres || o->conv == 2)) {
        if (im == 0.0) re = 1.0 / re + o->sigma;
        else { double t = re * re + im * im; re = re / t + o->sigma; im = -im / t; }
      }
      if (o->trueres) {
        const double *z = herm ? ds.Q + (int64_t)k * ld : ds.X + (int64_t)k * ld;
        resnorm = ks_true_residual(&ks, nv, z, cplx ? z + ld : NULL, re, im);
      } else resnorm *= beta * gamma;
      errest[k] = conv_errest(o->conv, nrma, re, im, resnorm);
      if (marker == -1 && errest[k] >= tol) marker = k;
      if (cplx) { errest[k + 1] = errest[k]; k++; }
      if (marker != -1) break;
    }
    if (marker != -1) k = marker;
    if (arn) {
      /* BVMultInPlace(V,Q,nconv,min(k+1,nv)) (arnoldi.c:109-115); the
         (k+1)-th Schur vector is the next start vector */
      int e = k + 1 < nv ? k + 1 : nv;
      orc_bv_multinplace(n, ks.V, ks.ld, nconv, nv, nconv, e, ds.Q, ld);
      if (res->trace_nconv && its <= res->trace_cap) { res->trace_nconv[its - 1] = k; res->trace_eig0[its - 1] = eigr[0]; }
      if (k >= nev) reason = 1;
      else if (its >= max_it) reason = -1;
      if (reason == 0 && breakdown) {                          /* arnoldi.c:118-125 */
        int bd = 0;
        ks_start_vector(&ks, k, nini, &bd);
        if (bd) reason = -2;
      }
      nconv = k;
      continue;
    }
    /* EPSStoppingBasic (epsdefault.c:290-302) */
    if (res->trace_nconv && its <= res->trace_cap) { res->trace_nconv[its - 1] = k; res->trace_eig0[its - 1] = eigr[0]; }
    if (k >= nev) reason = 1;
    else if (its >= max_it) reason = -1;
    int newl;
    if (reason != 0 || breakdown || k == nv) newl = 0;
    else {
      newl = (int)((nv - k) * keep); if (newl < 1) newl = 1;
      if (!herm) {                                           /* DSGetTruncateSize_Default (dsops.c:329-344) */
        if (ds.A[(k + newl) + (k + newl - 1) * ld] != 0.0) { if (k + newl < nv - 1) newl++; else newl--; }
      }
    }
    if (!lock && newl > 0) { newl += k; k = 0; }
    l = newl;
    if (reason == 0) {
      if (breakdown || k == nv) {
        if (k < nev) {
          int bd = 0;
          ks_start_vector(&ks, k, nini, &bd);
          if (bd) reason = -2;
        }
      } else {                                                /* DSTruncate(k+l,FALSE) */
        if (harm) {              /* undo the translation, gamma u^ = u - U g~ (krylovschur.c:311-320) */
          ds.n = nv; ds.t = nv; ds.l = k; ds.k = l;
          ds_translate_harmonic(&ds, 0.0, beta, 1, gh, &gamma);
          double *u = ks.V + (int64_t)nv * ks.ld;
          for (int64_t r = 0; r < n; r++) {
            double acc = 0.0;
            for (int j = 0; j < nv; j++) acc += ks.V[r + (int64_t)j * ks.ld] * gh[j];
            u[r] = (u[r] - acc) / gamma;
          }
          ds.n = nv; ds.t = nv; ds.l = k; ds.k = nv;
        }
        int nn = k + l;
        if (!herm && ds.k == ds.n) {
          for (int i = ds.l; i < nn; i++) ds.A[nn + i * ld] = ds.A[ds.n + i * ld];
          for (int i = ds.l; i < ds.n; i++) ds.A[ds.n + i * ld] = 0.0;
        }
        ds.k = nn; ds.t = ds.n; ds.n = nn; ds.state = ST_TRUNCATED;
      }
    }
    /* BVMultInPlace(V,Q,nconv,k+l) with active columns [nconv,nv) */
    orc_bv_multinplace(n, ks.V, ks.ld, nconv, nv, nconv, k + l, ds.Q, ld);
    if (reason == 0 && !breakdown) memcpy(ks.V + (int64_t)(k + l) * ks.ld, ks.V + (int64_t)nv * ks.ld, sizeof(double) * (size_t)n);
    nconv = k;
  }
  /* DSTruncate(nconv,TRUE); EPSComputeVectors; SlepcSortEigenvalues */
  if (!herm) for (int i = ds.l; i < ds.n; i++) ds.A[ds.n + i * ld] = 0.0; /* clean extra row */
  ds.n = nconv; ds.l = 0; ds.k = 0; ds.t = nconv;
  if (!herm && nconv > 0) {
    if (ds_vectors_nhep_all(&ds, 0)) { rc = -60; goto done; }
    orc_bv_multinplace(n, ks.V, ks.ld, 0, nconv, 0, nconv, ds.X, ld);
  }
  /* EPSComputeValues: STBackTransform, then the EPS sorting criterion (no map) */
  if (sinv) for (int i = 0; i < nconv; i++) sinv_map(&eigr[i], &eigi[i]);
  cmp_sinv = 0;
  /* reorder conjugate pairs, positive imaginary part first (epssolve.c:159-176) */
  for (int i = 0; i < nconv - 1; i++) {
    if (eigi[i] != 0) {
      if (eigi[i] < 0) {
        eigi[i] = -eigi[i]; eigi[i + 1] = -eigi[i + 1];
        orc_bv_scale(n, ks.V + (int64_t)(i + 1) * ks.ld, -1.0);
      }
      i++;
    }
  }
  for (int i = 0; i < nconv; i++) perm[i] = i;
  for (int i = nconv - 1; i >= 0; i--) {                     /* SlepcSortEigenvalues (slepcsc.c:89-140) */
    double re = eigr[perm[i]], im = eigi[perm[i]];
    int j = i + 1;
    if (im != 0) { i--; im = eigi[perm[i]]; }
    while (j < nconv) {
      if (sc_compare(which, re, im, eigr[perm[j]], eigi[perm[j]]) <= 0) break;
      int tmp;
      if (!im) {
        if (eigi[perm[j]] == 0.0) { tmp = perm[j - 1]; perm[j - 1] = perm[j]; perm[j] = tmp; j++; }
        else { tmp = perm[j - 1]; perm[j - 1] = perm[j]; perm[j] = perm[j + 1]; perm[j + 1] = tmp; j += 2; }
      } else {
        if (eigi[perm[j]] == 0.0) { tmp = perm[j - 2]; perm[j - 2] = perm[j]; perm[j] = perm[j - 1]; perm[j - 1] = tmp; j++; }
        else {
          tmp = perm[j - 2]; perm[j - 2] = perm[j]; perm[j] = tmp;
          tmp = perm[j - 1]; perm[j - 1] = perm[j + 1]; perm[j + 1] = tmp; j += 2;
        }
      }
    }
  }
  if (balD) {   /* EPSComputeVectors_Schur (epsdefault.c:128-139): x ./= D, then BVNormalize with eigi */
    for (int i = 0; i < nconv; i++)
      for (int q = 0; q < n; q++) ks.V[q + (int64_t)i * ks.ld] /= balD[q];
    for (int i = 0; i < nconv; i++) {
      double *a = ks.V + (int64_t)i * ks.ld;
      if (eigi[i] != 0.0 && i + 1 < nconv) {
        double *b = a + ks.ld, nr = orc_bv_norm2(n, a), ni = orc_bv_norm2(n, b), t = sqrt(nr * nr + ni * ni);
        orc_bv_scale(n, a, 1.0 / t); orc_bv_scale(n, b, 1.0 / t);
        i++;
      } else orc_bv_scale(n, a, 1.0 / orc_bv_norm2(n, a));
    }
  }
  res->seconds = now_s() - t0;
  res->nconv = nconv; res->its = its; res->reason = reason; res->ncv_used = ncv;
  res->matvecs = ks.matvecs; res->orth_passes = ks.passes;
  for (int i = 0; i < nconv; i++) {
    if (res->eigr) res->eigr[i] = eigr[perm[i]];
    if (res->eigi) res->eigi[i] = eigi[perm[i]];
    if (res->errest) res->errest[i] = errest[perm[i]];
  }
  if (res->vecs) {
    int nv = res->nvec < nconv ? res->nvec : nconv;
    for (int i = 0; i < nv; i++) memcpy(res->vecs + (int64_t)i * n, ks.V + (int64_t)perm[i] * ks.ld, sizeof(double) * (size_t)n);
  }
done:
#ifdef _OPENMP
  omp_set_num_threads(saved_threads);
#endif
  free(gh); free(lu); free(piv); free(balD); free(balval);
  cmp_sinv = 0;
  free(AVs); free(lockd); free(ks.V); free(ks.buf); free(ks.which); free(ds.A); free(ds.Q); free(ds.X); free(ds.T); free(ds.work); free(ds.perm); free(ds.iwork);
  free(eigr); free(eigi); free(errest); free(perm);
  free(bnd); free(yr); free(cv); free(lp);
  return rc;
}

/* ======================================================================== */
/* Two-sided Krylov-Schur (ks-twosided.c:126-241) with DSNHEPTS (dsnhepts.c)   */
/* ======================================================================== */
static void ds_swap_sides(DSo *ds, double **B, double **Z, double **Y)
{
  double *t = ds->A; ds->A = *B; *B = t;
  t = ds->Q; ds->Q = *Z; *Z = t;
  t = ds->X; ds->X = *Y; *Y = t;
}

/* DSSortWithPermutation_NHEP_Private (dsutil.c:177-226) */
static int ds_sort_with_permutation(DSo *ds, int *perm, double *wr, double *wi)
{
  int n = ds->n, ld = ds->ld, info = 0;
  double *T = ds->A;
  for (int i = ds->l; i < n - 1; i++) {
    int pos = perm[i];
    int inc = (pos < n - 1 && T[pos + 1 + pos * ld] != 0.0) ? 2 : 1;
    if (pos != i) {
      if (!((T[pos + (pos - 1) * ld] == 0.0 || perm[i + 1] == pos - 1) && (T[pos + 1 + pos * ld] == 0.0 || perm[i + 1] == pos + 1)))
        return -200;
      int ifst = pos + 1, ilst = i + 1;
      LP.trexc("V", &n, T, &ld, ds->Q, &ld, &ifst, &ilst, ds->work, &info, 1);
      if (info) return info;
      for (int j = i + 1; j < n; j++) if (perm[j] >= i && perm[j] < pos) perm[j] += inc;
      perm[i] = i;
      if (inc == 2) perm[i + 1] = i + 1;
    }
    if (inc == 2) i++;
  }
  for (int j = ds->l; j < n; j++) {
    wr[j] = T[j + j * ld];
    if (j < n - 1 && T[j + 1 + j * ld] != 0.0) {
      wi[j] = sqrt(fabs(T[j + 1 + j * ld])) * sqrt(fabs(T[j + (j + 1) * ld]));
      wr[j + 1] = wr[j]; wi[j + 1] = -wi[j]; j++;
    } else wi[j] = 0.0;
  }
  return 0;
}

/* Hessenberg columns k0..nv-1 (and the extra row) from the coefficient buffer */
static void ts_fill(const KSo *ks, int k0, int nv, double *H, int ld)
{
  int bm = ks->m;
  for (int j = k0; j < nv - 1; j++) memcpy(H + (int64_t)j * ld, ks->buf + (int64_t)(j + 1) * bm, sizeof(double) * (j + 2));
  memcpy(H + (int64_t)(nv - 1) * ld, ks->buf + (int64_t)nv * bm, sizeof(double) * nv);
  if (ld > nv) H[nv + (nv - 1) * ld] = ks->buf[nv + (int64_t)nv * bm];
}

/* EPSComputeVectors_Schur, two-sided part (epsdefault.c:118-167): V = V X,
   W = W Y, W normalised (a pair by its joint norm), second column of each
   pair with positive imaginary part negated */
static void ts_vectors(int n, int nconv, double *V, double *W, const double *X, const double *Y, int ld, const double *eigi)
{
  if (nconv <= 0) return;
  orc_bv_multinplace(n, V, n, 0, nconv, 0, nconv, X, ld);
  orc_bv_multinplace(n, W, n, 0, nconv, 0, nconv, Y, ld);
  for (int q = 0; q < nconv; q++) {
    int pair = eigi[q] != 0.0 && q + 1 < nconv;
    double nr = orc_bv_norm2(n, W + (int64_t)q * n);
    if (pair) nr = hypot(nr, orc_bv_norm2(n, W + (int64_t)(q + 1) * n));
    orc_bv_scale(n, W + (int64_t)q * n, 1.0 / nr);
    if (pair) { orc_bv_scale(n, W + (int64_t)(q + 1) * n, (eigi[q] > 0.0 ? -1.0 : 1.0) / nr); q++; }
  }
}

int orc_eps_solve_twosided(int n, const int32_t *rp, const int32_t *ci, const double *val, const orc_opts *o,
                           const double *v0, orc_result *res, double *lvecs)
{
  if (!lp_load() || !LP.getrf || !LP.getrs) return -10;
  int nev = o->nev > 0 ? o->nev : 1;
  if (nev > n) nev = n;                       /* small-problem safeguard, epssetup.c:342-345 */
  int ncv = o->ncv > 0 ? o->ncv : (nev + 15 > 2 * nev ? nev + 15 : 2 * nev);
  if (ncv > n) ncv = n;
  int mpd = o->mpd > 0 ? o->mpd : ncv;
  /* EPSSetDimensions_Default (epssetup.c:679) and krylovschur.c:88 */
  if (!(ncv >= nev + 1 || (ncv == nev && ncv == n)) || ncv > nev + mpd) return -16;
  int max_it = o->max_it > 0 ? o->max_it : (100 > 2 * n / ncv ? 100 : 2 * n / ncv);
  double tol = o->tol > 0 ? o->tol : 1e-8, keep = o->keep > 0 ? o->keep : 0.5;
  const double nrma = o->conv == 2 ? csr_norm_inf(n, rp, val) : 0.0;   /* epssetup.c:351-355 */
  if (o->hermitian || o->solver != 0 || o->extraction != 0) return -11;
  const int which = o->which ? o->which : ORC_LM;
  cmp_target = o->target;
  const int64_t nnz = rp[n];
  /* A^T in CSR, columns ascending */
  int32_t *trp = (int32_t *)calloc(n + 1, sizeof(int32_t)), *tci = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  int32_t *nxt = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
  double *tva = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  for (int64_t q = 0; q < nnz; q++) trp[ci[q] + 1]++;
  for (int r = 0; r < n; r++) trp[r + 1] += trp[r];
  for (int r = 0; r <= n; r++) nxt[r] = trp[r];
  for (int r = 0; r < n; r++)
    for (int32_t q = rp[r]; q < rp[r + 1]; q++) { int32_t d = nxt[ci[q]]++; tci[d] = r; tva[d] = val[q]; }
  free(nxt);
  const int ld = ncv + 1;
  KSo ks = {0}, kw = {0};
  ks.n = n; ks.rp = rp; ks.ci = ci; ks.val = val; ks.m = ncv + 1; ks.ld = n; ks.eta = o->eta > 0 ? o->eta : 0.7071; ks.refine = o->refine;
  kw = ks;
  kw.rp = trp; kw.ci = tci; kw.val = tva; kw.seed = 0x1EF7ULL;
  ks.V = (double *)calloc((size_t)n * ks.m, sizeof(double)); ks.buf = (double *)calloc((size_t)ks.m * ks.m, sizeof(double));
  kw.V = (double *)calloc((size_t)n * kw.m, sizeof(double)); kw.buf = (double *)calloc((size_t)kw.m * kw.m, sizeof(double));
  DSo ds = {0};
  ds.ld = ld; ds.hermitian = 0; ds.which = which;
  ds.A = (double *)calloc((size_t)ld * ld, sizeof(double)); ds.Q = (double *)calloc((size_t)ld * ld, sizeof(double));
  ds.X = (double *)calloc((size_t)ld * ld, sizeof(double)); ds.T = (double *)calloc((size_t)3 * ld, sizeof(double));
  ds.work = (double *)calloc((size_t)8 * ld * ld + 16 * ld, sizeof(double));
  ds.perm = (int *)calloc(ld, sizeof(int)); ds.iwork = (int *)calloc(ld, sizeof(int));
  double *B = (double *)calloc((size_t)ld * ld, sizeof(double)), *Z = (double *)calloc((size_t)ld * ld, sizeof(double));
  double *Y = (double *)calloc((size_t)ld * ld, sizeof(double));
  double *wr2 = (double *)calloc(ld, sizeof(double)), *wi2 = (double *)calloc(ld, sizeof(double));
  double *eigr = (double *)calloc(ld, sizeof(double)), *eigi = (double *)calloc(ld, sizeof(double));
  double *errest = (double *)calloc(ld, sizeof(double));
  double *M = (double *)calloc((size_t)ncv * ncv, sizeof(double)), *Mlu = (double *)malloc(sizeof(double) * (size_t)ncv * ncv);
  double *w = (double *)calloc(ld, sizeof(double)), *T1 = (double *)calloc((size_t)ld * ld, sizeof(double));
  double *blk = (double *)calloc((size_t)ld * ld, sizeof(double));
  int *ipiv = (int *)calloc(ld, sizeof(int)), *perm = (int *)calloc(ld, sizeof(int));
  int *idx = (int *)calloc(ld, sizeof(int)), *idx2 = (int *)calloc(ld, sizeof(int)), *pp = (int *)calloc(ld, sizeof(int));
  int rc = 0, nini = 0, info = 0;
  double t0 = now_s();
  if (v0) {
    memcpy(ks.V, v0, sizeof(double) * (size_t)n);
    double norm; int lindep;
    orthogonalize_gs(n, ks.V, ks.ld, 0, ks.buf, ks.m, ks.eta, ks.refine, &norm, &lindep, &ks.passes);
    if (norm == 0.0 || lindep) { rc = -3; goto done; }
    orc_bv_scale(n, ks.V, 1.0 / norm);
    nini = 1;
  }
  if (ks_start_vector(&ks, 0, nini, NULL) || ks_start_vector(&kw, 0, 0, NULL)) { rc = -3; goto done; }
  int l = 0, nconv = 0, its = 0, reason = 0;
  while (reason == 0) {
    its++;
    int nv = nconv + mpd < ncv ? nconv + mpd : ncv, k0 = nconv + l;
    ds.n = nv; ds.t = nv; ds.l = nconv; ds.k = k0;
    double beta = 0.0, betat = 0.0; int bd = 0, bdt = 0;
    ks_expand(&ks, k0, &nv, &beta, &bd);
    int nvt = nv;
    ks_expand(&kw, k0, &nvt, &betat, &bdt);
    if (nvt < nv) nv = nvt;
    ts_fill(&ks, k0, nv, ds.A, ld);
    ts_fill(&kw, k0, nv, B, ld);
    ds.n = nv; ds.t = nv; ds.l = nconv; ds.k = k0;
    ds.state = l == 0 ? ST_INTERMEDIATE : ST_RAW;
    int breakdown = bd || bdt;
    /* BVMatProject(V,NULL,W,M) (bvglobal.c:1013-1049) */
    if (k0 > 0 && k0 < nv) {
      orc_bv_dot(n, ks.V + (int64_t)k0 * n, n, nv - k0, kw.V, n, k0, blk, k0);
      for (int j = k0; j < nv; j++) for (int i = 0; i < k0; i++) M[i + j * ncv] = blk[i + (j - k0) * k0];
    }
    if (k0 < nv) {
      orc_bv_dot(n, ks.V, n, nv, kw.V + (int64_t)k0 * n, n, nv - k0, blk, nv - k0);
      for (int j = 0; j < nv; j++) for (int i = 0; i < nv - k0; i++) M[(k0 + i) + j * ncv] = blk[i + j * (nv - k0)];
    }
    /* EPSTwoSidedRQUpdate1 (ks-twosided.c:27-73) */
    memcpy(Mlu, M, sizeof(double) * (size_t)ncv * ncv);
    int one = 1;
    LP.getrf(&nv, &nv, Mlu, &ncv, ipiv, &info);
    if (info) { rc = -71; goto done; }
    orc_bv_dotvec(n, nv, kw.V, n, ks.V + (int64_t)nv * n, w);
    LP.getrs("N", &nv, &one, Mlu, &ncv, ipiv, w, &ncv, &info, 1);
    orc_bv_multvec(n, nv, -1.0, ks.V, n, w, 1.0, ks.V + (int64_t)nv * n);
    for (int i = 0; i < nv; i++) ds.A[(nv - 1) * ld + i] += beta * w[i];
    orc_bv_dotvec(n, nv, ks.V, n, kw.V + (int64_t)nv * n, w);
    LP.getrs("C", &nv, &one, Mlu, &ncv, ipiv, w, &ncv, &info, 1);
    orc_bv_multvec(n, nv, -1.0, kw.V, n, w, 1.0, kw.V + (int64_t)nv * n);
    for (int i = 0; i < nv; i++) B[(nv - 1) * ld + i] += betat * w[i];
    /* DSSolve_NHEPTS, DSSort_NHEPTS, DSUpdateExtraRow_NHEPTS */
    int st = ds.state;
    if ((info = ds_solve_nhep(&ds, eigr, eigi))) { rc = -20; goto done; }
    ds_swap_sides(&ds, &B, &Z, &Y);
    ds.state = st;
    info = ds_solve_nhep(&ds, wr2, wi2);
    ds_swap_sides(&ds, &B, &Z, &Y);
    if (info) { rc = -21; goto done; }
    if ((info = ds_sort_nhep(&ds, eigr, eigi))) { rc = -40; goto done; }
    ds_swap_sides(&ds, &B, &Z, &Y);
    info = ds_sort_nhep(&ds, wr2, wi2);
    if (!info) {
      int cont = 0;
      for (int i = 0; i < ds.n; i++) {
        if (hypot(wr2[i] - eigr[i], wi2[i] - eigi[i]) > 1.4901161193847656e-08) { idx2[cont] = i; idx[cont++] = i; }
        pp[i] = -1;
      }
      if (cont) {
        int id = 0;
        for (int i = 0; i < cont; i++) {
          double t = 1.79e308, sd;
          for (int j = 0; j < cont; j++)
            if (idx2[j] != -1 && (sd = hypot(wr2[idx[j]] - eigr[idx[i]], wi2[idx[j]] - eigi[idx[i]])) < t) { id = j; t = sd; }
          pp[idx[i]] = idx[id];
          idx2[id] = -1;
        }
        for (int i = 0; i < ds.n; i++) if (pp[i] == -1) pp[i] = i;
        info = ds_sort_with_permutation(&ds, pp, wr2, wi2);
      }
    }
    ds_swap_sides(&ds, &B, &Z, &Y);
    if (info) { rc = -41; goto done; }
    ds_update_extra_row_nhep(&ds);
    ds_swap_sides(&ds, &B, &Z, &Y);
    ds_update_extra_row_nhep(&ds);
    ds_swap_sides(&ds, &B, &Z, &Y);
    ds.k = ds.n;
    /* EPSKrylovConvergence with right and left bounds */
    double norm = orc_bv_norm2(n, ks.V + (int64_t)nv * n), norm2 = orc_bv_norm2(n, kw.V + (int64_t)nv * n);
    int k, marker = -1;
    for (k = nconv; k < nv; k++) {
      double rn, rl; int cplx = 0, cplx2 = 0;
      if (ds_vector_nhep_some(&ds, k, &rn, &cplx)) { rc = -50; goto done; }
      if (o->trueres)                                        /* right residual from the Ritz vector */
        rn = ks_true_residual(&ks, nv, ds.X + (int64_t)k * ld, cplx ? ds.X + (int64_t)(k + 1) * ld : NULL,
                              eigr[k], eigi[k]);
      else rn = rn * beta * norm;
      errest[k] = conv_errest(o->conv, nrma, eigr[k], eigi[k], rn);
      if (marker == -1 && errest[k] >= tol) marker = k;
      ds_swap_sides(&ds, &B, &Z, &Y);
      info = ds_vector_nhep_some(&ds, k, &rl, &cplx2);
      ds_swap_sides(&ds, &B, &Z, &Y);
      if (info) { rc = -51; goto done; }
      double le = conv_errest(o->conv, nrma, eigr[k], eigi[k], rl * betat * norm2);
      if (le > errest[k]) errest[k] = le;
      if (marker == -1 && le >= tol) marker = k;
      if (cplx) { errest[k + 1] = errest[k]; k++; }
      if (marker != -1) break;
    }
    if (marker != -1) k = marker;
    if (k >= nev) reason = 1;
    else if (its >= max_it) reason = -1;
    int newl = 0;
    if (!(reason != 0 || breakdown || k == nv)) {
      newl = (int)((nv - k) * keep); if (newl < 1) newl = 1;
      if (ds.A[(k + newl) + (k + newl - 1) * ld] != 0.0 || B[(k + newl) + (k + newl - 1) * ld] != 0.0) {
        if (k + newl < nv - 1) newl++; else newl--;
      }
    }
    l = newl;
    orc_bv_multinplace(n, ks.V, n, nconv, nv, nconv, k + l, ds.Q, ld);
    orc_bv_multinplace(n, kw.V, n, nconv, nv, nconv, k + l, Z, ld);
    if (reason == 0 && !breakdown) {
      memcpy(ks.V + (int64_t)(k + l) * n, ks.V + (int64_t)nv * n, sizeof(double) * (size_t)n);
      memcpy(kw.V + (int64_t)(k + l) * n, kw.V + (int64_t)nv * n, sizeof(double) * (size_t)n);
    }
    if (reason == 0) {
      if (breakdown || k == nv) {
        if (k < nev) {
          int b1 = 0, b2 = 0;
          ks_start_vector(&ks, k, nini, &b1);
          ks_start_vector(&kw, k, 0, &b2);
          if (b1 || b2) reason = -2;
        }
      } else {                                       /* DSSetDimensions(n,k,dsk); DSTruncate(k+l,FALSE) */
        ds.l = k;
        int nn = k + l;
        if (ds.k == ds.n) {
          for (int i = ds.l; i < nn; i++) { ds.A[nn + i * ld] = ds.A[ds.n + i * ld]; B[nn + i * ld] = B[ds.n + i * ld]; }
          for (int i = ds.l; i < ds.n; i++) { ds.A[ds.n + i * ld] = 0.0; B[ds.n + i * ld] = 0.0; }
        }
        ds.k = nn; ds.t = ds.n; ds.n = nn; ds.state = ST_TRUNCATED;
      }
      /* EPSTwoSidedRQUpdate2 (ks-twosided.c:75-124) */
      int kk = k + l;
      KSo *side[2] = {&ks, &kw};
      double *H[2] = {ds.A, B};
      for (int sd = 0; sd < 2; sd++) {
        double nrm; int lind;
        orthogonalize_gs(n, side[sd]->V, n, kk, side[sd]->buf, side[sd]->m, side[sd]->eta, side[sd]->refine, &nrm, &lind, &side[sd]->passes);
        orc_bv_scale(n, side[sd]->V + (int64_t)kk * n, 1.0 / nrm);
        const double *c = side[sd]->buf + (int64_t)kk * side[sd]->m;
        for (int j = nconv; j < kk; j++) {
          for (int i = 0; i < kk; i++) H[sd][i + j * ld] += c[i] * H[sd][kk + j * ld];
          H[sd][kk + j * ld] *= nrm;
        }
      }
      for (int j = 0; j < nv; j++)
        for (int i = 0; i < nv; i++) {
          double acc = 0.0;
          for (int q = 0; q < nv; q++) acc += M[i + q * ncv] * ds.Q[q + j * ld];
          T1[i + j * ld] = acc;
        }
      for (int j = 0; j < nv; j++)
        for (int i = 0; i < nv; i++) {
          double acc = 0.0;
          for (int q = 0; q < nv; q++) acc += Z[q + i * ld] * T1[q + j * ld];
          M[i + j * ncv] = acc;
        }
    }
    nconv = k;
  }
  /* DSTruncate(nconv,TRUE); right and left eigenvectors (epsdefault.c:105-169) */
  for (int i = ds.l; i < ds.n; i++) { ds.A[ds.n + i * ld] = 0.0; B[ds.n + i * ld] = 0.0; }
  ds.n = nconv; ds.l = 0; ds.k = 0; ds.t = nconv; ds.state = ST_RAW;
  if (nconv > 0) {
    if (ds_vectors_nhep_all(&ds, 0)) { rc = -60; goto done; }
    ds_swap_sides(&ds, &B, &Z, &Y);
    info = ds_vectors_nhep_all(&ds, 0);
    ds_swap_sides(&ds, &B, &Z, &Y);
    if (info) { rc = -61; goto done; }
  }
  /* EPSSolve computes the vectors at the first conjugate pair it reorders
     (epssolve.c:159-176), so that pair already counts as positive */
  int computed = 0;
  for (int i = 0; i < nconv - 1; i++) {
    if (eigi[i] != 0) {
      if (eigi[i] < 0) {
        eigi[i] = -eigi[i]; eigi[i + 1] = -eigi[i + 1];
        if (!computed) { ts_vectors(n, nconv, ks.V, kw.V, ds.X, Y, ld, eigi); computed = 1; }
        orc_bv_scale(n, ks.V + (int64_t)(i + 1) * n, -1.0);
      }
      i++;
    }
  }
  if (!computed) ts_vectors(n, nconv, ks.V, kw.V, ds.X, Y, ld, eigi);
  /* SlepcSortEigenvalues */
  for (int i = 0; i < nconv; i++) perm[i] = i;
  for (int i = nconv - 1; i >= 0; i--) {
    double re = eigr[perm[i]], im = eigi[perm[i]];
    int j = i + 1, tmp;
    if (im != 0) { i--; im = eigi[perm[i]]; }
    while (j < nconv) {
      if (sc_compare(which, re, im, eigr[perm[j]], eigi[perm[j]]) <= 0) break;
      if (!im) {
        if (eigi[perm[j]] == 0.0) { tmp = perm[j - 1]; perm[j - 1] = perm[j]; perm[j] = tmp; j++; }
        else { tmp = perm[j - 1]; perm[j - 1] = perm[j]; perm[j] = perm[j + 1]; perm[j + 1] = tmp; j += 2; }
      } else {
        if (eigi[perm[j]] == 0.0) { tmp = perm[j - 2]; perm[j - 2] = perm[j]; perm[j] = perm[j - 1]; perm[j - 1] = tmp; j++; }
        else {
          tmp = perm[j - 2]; perm[j - 2] = perm[j]; perm[j] = tmp;
          tmp = perm[j - 1]; perm[j - 1] = perm[j + 1]; perm[j + 1] = tmp; j += 2;
        }
      }
    }
  }
  res->seconds = now_s() - t0;
  res->nconv = nconv; res->its = its; res->reason = reason; res->ncv_used = ncv;
  res->matvecs = ks.matvecs + kw.matvecs; res->orth_passes = ks.passes + kw.passes;
  for (int i = 0; i < nconv; i++) {
    if (res->eigr) res->eigr[i] = eigr[perm[i]];
    if (res->eigi) res->eigi[i] = eigi[perm[i]];
    if (res->errest) res->errest[i] = errest[perm[i]];
  }
  {
    int nvv = res->nvec < nconv ? res->nvec : nconv;
    for (int i = 0; i < nvv; i++) {
      if (res->vecs) memcpy(res->vecs + (int64_t)i * n, ks.V + (int64_t)perm[i] * n, sizeof(double) * (size_t)n);
      if (lvecs) memcpy(lvecs + (int64_t)i * n, kw.V + (int64_t)perm[i] * n, sizeof(double) * (size_t)n);
    }
  }
done:
  free(trp); free(tci); free(tva);
  free(ks.V); free(ks.buf); free(kw.V); free(kw.buf);
  free(ds.A); free(ds.Q); free(ds.X); free(ds.T); free(ds.work); free(ds.perm); free(ds.iwork);
  free(B); free(Z); free(Y); free(wr2); free(wi2); free(eigr); free(eigi); free(errest); free(M); free(Mlu);
  free(w); free(T1); free(blk); free(ipiv); free(perm); free(idx); free(idx2); free(pp);
  return rc;
}

/* EPSComputeResidualNorm_Private + EPS_ERROR_RELATIVE (epssolve.c:666-815) */
/* EPSComputeResidualNorm_Private (epssolve.c:666-718), trans = FALSE, B = I:
   ||A x - lambda x|| with x = xr + i xi; the lambda terms are dropped when
   |lambda| <= eps, and a pair whose imaginary part is below |kr| eps is real. */
double orc_residual_abs(int n, const int32_t *rp, const int32_t *ci, const double *val, double kr, double ki,
                        const double *xr, const double *xi)
{
  double *u = (double *)malloc(sizeof(double) * n);
  double nr, ni = 0.0;
  const double eps = 2.220446049250313e-16;
  if (ki == 0 || fabs(ki) < fabs(kr * eps) || !xi) {
    orc_spmv(n, rp, ci, val, xr, u);
    if (fabs(kr) > eps) for (int i = 0; i < n; i++) u[i] -= kr * xr[i];
    nr = orc_bv_norm2(n, u);
    free(u);
    return nr;
  }
  const int big = hypot(kr, ki) > eps;
  orc_spmv(n, rp, ci, val, xr, u);
  if (big) for (int i = 0; i < n; i++) u[i] += -kr * xr[i] + ki * xi[i];
  nr = orc_bv_norm2(n, u);
  orc_spmv(n, rp, ci, val, xi, u);
  if (big) for (int i = 0; i < n; i++) u[i] += -kr * xi[i] - ki * xr[i];
  ni = orc_bv_norm2(n, u);
  free(u);
  return hypot(nr, ni);
}

double orc_residual_rel(int n, const int32_t *rp, const int32_t *ci, const double *val, double kr, double ki,
                        const double *xr, const double *xi)
{
  const double eps = 2.220446049250313e-16;
  const int real = (ki == 0 || fabs(ki) < fabs(kr * eps) || !xi);
  return orc_residual_abs(n, rp, ci, val, kr, ki, xr, xi) / (real ? fabs(kr) : hypot(kr, ki));
}

/* DSSolve + DSSort on a compact HEP problem (arrowhead rows l..k, tridiagonal
   below), as src/sys/classes/ds/tests/test3.c drives it; returns wr[0..n) and
   Q (ld x ld, ld = n+1). */
int orc_ds_hep_compact(int n, int l, int k, const double *d, const double *e, int which, double *wr, double *Q)
{
  if (!lp_load()) return -10;
  int ld = n + 1;
  DSo ds = {0};
  ds.ld = ld; ds.hermitian = 1; ds.which = which;
  ds.Q = (double *)calloc((size_t)ld * ld, sizeof(double));
  ds.T = (double *)calloc((size_t)3 * ld, sizeof(double));
  ds.work = (double *)calloc((size_t)8 * ld * ld + 16 * ld, sizeof(double));
  ds.perm = (int *)calloc(ld, sizeof(int));
  memcpy(ds.T, d, sizeof(double) * n);
  memcpy(ds.T + ld, e, sizeof(double) * (n - 1));
  ds.n = n; ds.l = l; ds.k = k; ds.t = n;
  ds.state = (l == 0 && k == 0) ? ST_INTERMEDIATE : ST_RAW;
  double *wi = (double *)calloc(ld, sizeof(double));
  int info = ds_solve_hep(&ds, wr, wi);
  if (!info) ds_sort_hep(&ds, wr);
  if (Q) memcpy(Q, ds.Q, sizeof(double) * (size_t)ld * ld);
  free(ds.Q); free(ds.T); free(ds.work); free(ds.perm); free(wi);
  return info;
}
