// Krylov-Schur driver interface shared by the C API.
#pragma once
#include <vector>

#include "sa_internal.h"

namespace sa {

// Per-iteration hooks set through the C ABI: EPSMonitor (epsmon.c:39-46),
// called at the end of every outer iteration with the Ritz values and error
// estimates of that iteration, and the stopping test (EPSSetStoppingTestFunction,
// epsopts.c; EPSStoppingBasic epsdefault.c:290-302 when unset).  owner is
// passed back to the trampolines (the sa_eps handle).  A nonzero return from
// either ends the solve with an error, as PetscCall does.
struct KsHooks {
  int (*monitor)(void *owner, int its, int nconv, const double *eigr, const double *eigi, const double *errest,
                 int nest) = nullptr;
  int (*stopping)(void *owner, int its, int max_it, int nconv, int nev, int *reason) = nullptr;
  // EPSSetConvergenceTestFunction: errest from the eigenvalue and residual norm
  int (*converged)(void *owner, double re, double im, double res, double *errest) = nullptr;
  void *owner = nullptr;
};

struct KsParams {
  bool hermitian = true;
  int which = SA_LARGEST_MAGNITUDE;
  int nev = 1, ncv = 16, mpd = 16, max_it = 100;
  double tol = 1e-8, keep = 0.5, eta = 0.7071;
  int conv = SA_CONV_REL;       // EPSSetConvergenceTest: SA_CONV_ABS, _REL, _NORM, _USER
  double nrma = 0.0, nrmb = 0.0;  // ||A||_inf and ||B||_inf of EPS_CONV_NORM (nrmb 0 unless set)
  int solver = 0;              // 0 Krylov-Schur, 1 Arnoldi with explicit restart (EPSARNOLDI), 2 Lanczos
  int extraction = 0;          // 0 Ritz, 1 harmonic Ritz (EPS_HARMONIC)
  double target = 0.0;         // EPSSetTarget: comparison point of which 7/8 and the harmonic shift
  int st_type = 0;             // 0 shift (sigma = 0), 1 shift-and-invert (STSINVERT)
  double st_shift = 0.0;       // sigma of the shift-and-invert operator
  bool ksp_direct = true;      // KSPPREONLY + PCLU (banded LU on the device); else BiCGStab
  bool ksp_jacobi = true;      // PCJACOBI for the BiCGStab solves (else none)
  bool ksp_gmres = false;      // GMRES(ksp_restart) instead of BiCGStab for the iterative solves
  int ksp_restart = 30;
  double ksp_rtol = 1e-12;     // relative residual of each solve
  int ksp_max_it = 10000;
  bool lanczos_local = false;  // EPSLanczosSetReorthog: local (else full)
  bool arnoldi_delayed = false;  // EPSArnoldiSetDelayed
  bool lock = true;              // EPSKrylovSchurSetLocking
  bool trueres = false;          // EPSSetTrueResidual (Krylov-Schur and Arnoldi)
  int lanczos_partial = 0;     // EPSLanczosSetReorthog: 1 periodic, 2 partial, 3 selective, 4 delayed
  int refine = 0;              // BVSetOrthogonalization refinement of the basis: GS_REFINE_*
  BV *defl = nullptr;          // EPSSetDeflationSpace: ndefl orthonormal columns the basis is kept orthogonal to
  int ndefl = 0;
  KsHooks hooks;
};

struct KsResult {
  int nconv = 0, its = 0, reason = 0;
  int64_t matvecs = 0, gs_passes = 0;
  double seconds = 0.0, seconds_ds = 0.0;
  double bytes_spmv = 0.0, bytes_bv = 0.0, bytes_restart = 0.0;
  int64_t ksp_its = 0;             // linear-solver iterations (shift-and-invert)
  int64_t true_residuals = 0;      // explicit residuals of the convergence test
  std::vector<int> perm;           // sorted position -> column of V
  std::vector<double> eigr, eigi, errest;
};

// The error estimate of a pair from its eigenvalue (re, im) and residual norm
// res by the convergence test of p (EPSConvergedAbsolute / Relative / Norm,
// epsdefault.c:224-256, or the user's function).
int eps_converged(const KsParams &p, double re, double im, double res, double *errest);
// The stopping test of iteration its with nconv converged pairs: the hook,
// or EPSStoppingBasic (epsdefault.c:290-302): 1 (converged) when nconv >=
// nev, -1 (diverged, iterations) when its >= max_it, else 0.
int eps_stopping(const KsParams &p, int its, int nconv, int nev, int *reason);
// EPSMonitor: every monitor hook in turn (nothing when none is set).
int eps_monitor(const KsParams &p, int its, int nconv, const double *eigr, const double *eigi, const double *errest,
                int nest);

double spmv_bytes(const Mat &A);
int bv_upload_column(BV &V, int j, const double *x);
int bv_download_column(BV &V, int j, double *x);
// Runs Krylov-Schur (or, with p.solver == 1, explicitly restarted Arnoldi) on
// A with basis V (V.m == ncv+1).  If have_initial, column 0 of V holds the
// initial vector (EPSSetInitialSpace).
int ks_solve(const KsParams &p, const Mat &A, BV &V, bool have_initial, KsResult &r);

// EPSGetStartVector (epssolve.c:841-873): column i of V from the initial
// vector (i == 0 and have_initial) or a uniform random vector with seed
// seed + i, orthonormalised against columns 0..i-1.
constexpr uint64_t EPS_SEED_RIGHT = 0x5EEDULL, EPS_SEED_LEFT = 0x1EF7ULL;
int eps_start_vector(BV &V, int i, bool have_initial, double eta, int *breakdown, std::vector<double> &scratch,
                     uint64_t seed = EPS_SEED_RIGHT, BV *defl = nullptr, int ndefl = 0);

// EPSComputeRitzVector + EPSComputeResidualNorm_Private (epsdefault.c:386-398,
// epssolve.c:666-718): x = V(:,0:nv) Z(:,0) (and the imaginary part
// V(:,0:nv) Z(:,1) when nx == 2) into work (two columns of V's leading
// dimension), *resnorm = ||A x - (re + i im) x||.  part: RESID_MAXVEC *
// RESID_MAXG + RESID_MAXVEC doubles.  Uses V's Q staging buffers.
int ritz_true_residual(BV &V, const Mat &A, int nv, const double *Z, int ldz, int nx, double re, double im,
                       void *work, double *part, double *resnorm);

// Two-sided Krylov-Schur (ks-twosided.c:126-241): right basis V of A and left
// basis W of A^T (AT), DSNHEPTS projected problem; on return the first nconv
// columns of V and W hold right and left eigenvectors.
int ks_twosided(const KsParams &p, const Mat &A, const Mat &AT, BV &V, BV &W, bool have_initial, KsResult &r);

}  // namespace sa
