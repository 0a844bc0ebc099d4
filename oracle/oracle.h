/*
 * oracle.h -- CPU FP64 restatement of the SLEPc 3.22.2 Krylov-Schur hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in slepc_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the SLEPc source tree).  Problem generators follow the stencil
 * definitions of the workloads (2-D/3-D Laplacians, convection-diffusion,
 * 27-point stencil, finite-volume diffusion, Markov walk of ex5.c).
 */
#ifndef SLEPC_AMD_ORACLE_H
#define SLEPC_AMD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- deterministic random numbers (counter based, splitmix64) ---------- */
uint64_t orc_splitmix64(uint64_t z);
double   orc_u01(uint64_t seed, uint64_t i);           /* uniform [0,1)  */
void     orc_rand_uniform(uint64_t seed, int64_t n, double *out); /* U[-1,1) */
double   orc_gauss(uint64_t seed, uint64_t i);         /* N(0,1) Box-Muller */
uint64_t orc_sub_seed(uint64_t seed, uint64_t p);

/* ---- CSR generators: return nnz; fill arrays when non-NULL ------------- */
int64_t orc_gen_lap2d(int nx, int ny, int32_t *rowptr, int32_t *col, double *val);
int64_t orc_gen_lap2d_pot(int nx, int ny, uint64_t seed, double scale,
                          int32_t *rowptr, int32_t *col, double *val);
int64_t orc_gen_convdiff3d(int nx, int ny, int nz, int32_t *rowptr, int32_t *col, double *val);
int64_t orc_gen_stencil27(int nx, int ny, int nz, uint64_t seed,
                          int32_t *rowptr, int32_t *col, double *val);
int64_t orc_gen_fv2d(int nx, int ny, const double *theta, double *sigma_out,
                     int32_t *rowptr, int32_t *col, double *val);
int64_t orc_gen_markov(int m, int32_t *rowptr, int32_t *col, double *val);
void    orc_fv2d_theta(uint64_t seed, int sample, double *theta8);

/* ---- SpMV and BV kernels ----------------------------------------------- */
void orc_spmv(int n, const int32_t *rowptr, const int32_t *col, const double *val,
              const double *x, double *y);
void orc_bv_dotvec(int64_t n, int k, const double *V, int64_t ld, const double *x, double *q);
void orc_bv_multvec(int64_t n, int k, double alpha, const double *V, int64_t ld,
                    const double *q, double beta, double *y);
void orc_bv_multinplace(int64_t n, double *V, int64_t ld, int l, int k, int s, int e,
                        const double *Q, int ldq);
void orc_bv_scale(int64_t n, double *x, double alpha);
double orc_bv_norm2(int64_t n, const double *x);
/* BVOrthonormalizeColumn with CGS + DGKS(if needed): buffer is m x m */
int  orc_bv_orthonormalize_column(int64_t n, double *V, int64_t ld, int j, double *buffer,
                                  int m, double eta, double *norm, int *lindep);

/* ---- block BV operations (BVDot, BVMatMult, BVOrthogonalize) ----------- */
void orc_bv_dot(int64_t n, const double *X, int64_t ldx, int mx, const double *Y, int64_t ldy, int my,
                double *M, int ldm);
void orc_bv_matmult(int n, const int32_t *rp, const int32_t *ci, const double *val, const double *X,
                    int64_t ldx, double *Y, int64_t ldy, int nc);
int  orc_bv_orthogonalize(int64_t n, double *V, int64_t ld, int m, int l, int k, int method, double *R, int ldr);

/* ---- dense projected problems (host LAPACK) ---------------------------- */
int  orc_lapack_status(char *buf, int len);
void orc_arrow_tridiag(int n, double *d, double *e, double *Q, int ld);
int  orc_ds_hep_compact(int n, int l, int k, const double *d, const double *e, int which,
                        double *wr, double *Q);

/* ---- Krylov-Schur and Arnoldi solves ----------------------------------- */
enum { ORC_LM = 1, ORC_SM = 2, ORC_LR = 3, ORC_SR = 4, ORC_LI = 5, ORC_SI = 6, ORC_TM = 7, ORC_TR = 8, ORC_USER = 11 };
/* EPSSetEigenvalueComparison: the comparison for ORC_USER (res < 0: a first). */
typedef int (*orc_compare_fn)(double ar, double ai, double br, double bi, int *res, void *ctx);
void orc_set_user_comparison(orc_compare_fn fn, void *ctx);
typedef struct {
  int     hermitian;   /* 1: Lanczos + DSHEP, 0: Arnoldi + DSNHEP */
  int     which;       /* ORC_LM ... */
  int     nev, ncv, mpd, max_it;
  double  tol;
  double  keep;        /* restart keep ratio, default 0.5 */
  double  eta;         /* DGKS threshold, default 0.7071 */
  int     lock;        /* locking variant (default 1) */
  int     nthreads;    /* OpenMP threads, 0 = default */
  int     solver;      /* 0 Krylov-Schur, 1 Arnoldi with explicit restart (arnoldi.c),
                          2 Lanczos with full reorthogonalization (lanczos.c) */
  int     extraction;  /* 0 Ritz, 1 harmonic Ritz (EPS_HARMONIC; Krylov-Schur and Arnoldi) */
  double  target;      /* EPSSetTarget: which 7/8 and the harmonic shift */
  int     st;          /* 0 shift (sigma = 0), 1 shift-and-invert (STSINVERT) */
  double  sigma;       /* shift of the shift-and-invert operator */
  int     lanczos_local; /* Lanczos: 1 -eps_lanczos_reorthog local, 0 full */
  int     arnoldi_delayed; /* EPSArnoldiSetDelayed */
  int     balance;       /* EPSSetBalance: 0 none, 1 oneside, 2 twoside */
  int     balance_its;   /* 0: 5 */
  double  balance_cutoff;  /* 0: 1e-8 */
  int     lanczos_partial; /* Lanczos: 1 periodic, 2 partial, 3 selective, 4 delayed (lanczos_local 0) */
  int     trueres;         /* EPSSetTrueResidual: explicit residuals in the convergence test
                              (Krylov-Schur, two-sided right residuals, Arnoldi) */
  int     refine;          /* BVSetOrthogonalization refinement: ORC_REFINE_IFNEEDED (default),
                              ORC_REFINE_NEVER, ORC_REFINE_ALWAYS */
  int     conv;            /* EPSSetConvergenceTest: 0 relative (default), 1 absolute, 2 norm-relative */
} orc_opts;
enum { ORC_REFINE_IFNEEDED = 0, ORC_REFINE_NEVER = 1, ORC_REFINE_ALWAYS = 2 };
/* BVOrthogonalizeSomeColumn (MGS against the columns with which[i] set) with a
   refinement type; returns the number of passes. */
int orc_bv_orthogonalize_some_column(int64_t n, double *V, int64_t ld, int j, const char *which, double *buffer,
                                     int m, int refine, double eta, double *norm, int *lindep);

typedef struct {
  int     nconv, its, reason;   /* reason: 1 converged, -1 max_it, -2 breakdown */
  int     ncv_used;
  double  *eigr, *eigi;         /* [ncv] sorted as EPSGetEigenpair returns them */
  double  *errest;              /* [ncv] */
  double  *vecs;                /* [n * nvec] eigenvectors (real parts), may be NULL */
  int     nvec;                 /* number of vectors to return */
  int64_t matvecs, orth_passes; /* counters */
  double  seconds;              /* wall time of the solve loop */
  /* per-iteration trace for the monitor comparison (EPSMonitor, epsmon.c:39-46),
     optional: entry its-1 receives the converged count the stopping test saw
     and the first Ritz value of that iteration, up to trace_cap iterations */
  int     *trace_nconv;
  double  *trace_eig0;
  int     trace_cap;
} orc_result;

int orc_lanczos_local_sweep(int n, const int32_t *rp, const int32_t *ci, const double *val, int m, const double *v0,
                            double *V, double *alpha, double *beta);
/* Returns 0, or a negative code: -16 for dimensions the reference rejects at
   setup (ncv < nev+1 unless ncv = nev = n, or ncv > nev+mpd), after nev and
   ncv are clamped to n as its small-problem safeguard does. */
int orc_eps_solve(int n, const int32_t *rowptr, const int32_t *col, const double *val,
                  const orc_opts *o, const double *v0, orc_result *r);
/* Two-sided Krylov-Schur (ks-twosided.c:126-241): as orc_eps_solve, plus the
   left eigenvectors (real parts) in lvecs [n * nvec], may be NULL. */
int orc_eps_solve_twosided(int n, const int32_t *rowptr, const int32_t *col, const double *val,
                           const orc_opts *o, const double *v0, orc_result *r, double *lvecs);
/* EPSBuildBalance_Krylov (epsdefault.c:443-504): D of the balanced operator. */
int orc_build_balance(int n, const int32_t *rp, const int32_t *ci, const double *val, int twoside, int its,
                      double cutoff, double *D);
double orc_residual_rel(int n, const int32_t *rowptr, const int32_t *col, const double *val,
                        double lr, double li, const double *xr, const double *xi);
double orc_residual_abs(int n, const int32_t *rowptr, const int32_t *col, const double *val,
                        double lr, double li, const double *xr, const double *xi);

#ifdef __cplusplus
}
#endif
#endif
