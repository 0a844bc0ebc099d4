// Eigensolver driver: the restart loop of EPSSolve_KrylovSchur_Default
// (src/eps/impls/krylov/krylovschur/krylovschur.c:227-348), and the explicit
// restarts of EPSARNOLDI (arnoldi.c:52-131) and EPSLANCZOS with every
// reorthogonalization type (lanczos.c:36-697), over device-resident BV/Mat
// objects, with the projected problem on the host.
#include "sa_eps.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "sa_ds.h"

namespace sa {

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

// Algorithmic HBM bytes of one SpMV: values + column indices + row pointers + x + y.
int eps_converged(const KsParams &p, double re, double im, double res, double *errest) {
  const double w = std::hypot(re, im);
  switch (p.conv) {
    case SA_CONV_ABS: *errest = res; return 0;
    case SA_CONV_NORM: *errest = res / (p.nrma + w * p.nrmb); return 0;
    case SA_CONV_USER: {
      SA_CHECK(p.hooks.converged, "the user convergence test is not set");
      const int rc = p.hooks.converged(p.hooks.owner, re, im, res, errest);
      SA_CHECK(rc == 0, "the convergence test returned an error code");
      return 0;
    }
    default: *errest = w != 0.0 ? res / w : 1.79e308; return 0;
  }
}

int eps_stopping(const KsParams &p, int its, int nconv, int nev, int *reason) {
  if (p.hooks.stopping) {
    int r = 0;
    const int rc = p.hooks.stopping(p.hooks.owner, its, p.max_it, nconv, nev, &r);
    SA_CHECK(rc == 0, "the stopping test returned an error code");
    *reason = r;
    return 0;
  }
  *reason = nconv >= nev ? 1 : (its >= p.max_it ? -1 : 0);
  return 0;
}

int eps_monitor(const KsParams &p, int its, int nconv, const double *eigr, const double *eigi, const double *errest,
                int nest) {
  if (!p.hooks.monitor) return 0;
  const int rc = p.hooks.monitor(p.hooks.owner, its, nconv, eigr, eigi, errest, nest);
  SA_CHECK(rc == 0, "a monitor returned an error code");
  return 0;
}

double spmv_bytes(const Mat &A) {
  const double s = (double)prec_size(A.prec);
  return (double)A.nnz * (s + 4.0) + 4.0 * (A.n + 1) + 2.0 * A.n * s;
}

// Uploads host vector x (FP64) into column j, converting to the BV precision.
int bv_upload_column(BV &V, int j, const double *x) {
  hipStream_t s = stream();
  if (V.prec == SA_FP64) {
    SA_HIP(hipMemcpyAsync(col_ptr(V, j), x, sizeof(double) * V.n, hipMemcpyHostToDevice, s));
  } else {
    std::vector<float> f(x, x + V.n);
    SA_HIP(hipMemcpyAsync(col_ptr(V, j), f.data(), sizeof(float) * V.n, hipMemcpyHostToDevice, s));
    SA_HIP(hipStreamSynchronize(s));
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int bv_download_column(BV &V, int j, double *x) {
  hipStream_t s = stream();
  if (V.prec == SA_FP64) {
    SA_HIP(hipMemcpyAsync(x, col_ptr(V, j), sizeof(double) * V.n, hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    return 0;
  }
  std::vector<float> f(V.n);
  SA_HIP(hipMemcpyAsync(f.data(), col_ptr(V, j), sizeof(float) * V.n, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < V.n; i++) x[i] = f[i];
  return 0;
}

// EPSGetStartVector (epssolve.c:841-873): column i is a deterministic random
// vector unless an initial vector is already in place (i == 0 && nini > 0).
int eps_start_vector(BV &V, int i, bool have_initial, double eta, int *breakdown, std::vector<double> &scratch,
                     uint64_t seed, BV *defl, int ndefl) {
  if (i > 0 || !have_initial) {
    scratch.resize(V.n);
    sa_rand_uniform(seed + (uint64_t)i, V.n, scratch.data());
    SA_TRY(bv_upload_column(V, i, scratch.data()));
  }
  if (defl) SA_TRY(bv_project_out(*defl, ndefl, col_ptr(V, i), stream()));   // the constraints come first
  double norm = 0.0;
  int lindep = 0;
  SA_TRY(bv_orthonormalize(V, i, eta, &norm, &lindep));
  if (defl) SA_TRY(bv_project_out(*defl, ndefl, col_ptr(V, i), stream(), 1));   // as after each Arnoldi column
  if (breakdown) *breakdown = lindep;
  else SA_CHECK(!(lindep || norm == 0.0), "Initial vector is zero or belongs to the deflation space");
  return 0;
}

// PETSc's PETSC_SQRT_MACHINE_EPSILON for double: a decimal literal one ulp
// above 2^-26, which the thresholds below compare against.
constexpr double LZ_SQRT_EPS = 1.490116119384766e-08;

// ---- EPSPartialLanczos (lanczos.c:338-441): periodic and partial ----------
// The omega recurrence (update_omega, lanczos.c:260-293) estimates the loss of
// orthogonality of the next Lanczos vector against each earlier one; its
// beta argument is the caller's shifted by one, so B(x) = beta[x - 1], and
// alpha / omega are indexed from 0 whatever the number of locked columns.  The
// reference's closing exchange copies omega_old into omega and leaves
// omega_old unchanged; kept as it is.
static void lz_update_omega(double *omega, double *omega_old, int j, const double *alpha, const double *beta,
                            double eps1, double anorm) {
  auto B = [&](int x) { return beta[x - 1]; };
  const double T = eps1 * anorm, binv = 1.0 / B(j + 1);
  omega_old[0] = B(1) * omega[1] + (alpha[0] - alpha[j]) * omega[0] - B(j) * omega_old[0];
  omega_old[0] = omega_old[0] > 0 ? binv * (omega_old[0] + T) : binv * (omega_old[0] - T);
  for (int q = 1; q < j - 1; q++) {
    omega_old[q] = B(q + 1) * omega[q + 1] + (alpha[q] - alpha[j]) * omega[q] + B(q) * omega[q - 1] -
                   B(j) * omega_old[q];
    omega_old[q] = omega_old[q] > 0 ? binv * (omega_old[q] + T) : binv * (omega_old[q] - T);
  }
  omega_old[j - 1] = binv * T;
  for (int q = 0; q < j; q++) omega[q] = omega_old[q];
  omega[j] = eps1;
}

// compute_int (lanczos.c:295-333): columns with |mu| >= delta (else the
// largest), widened on both sides while |mu| >= eta.
static void lz_compute_int(char *which, const double *mu, int j, double delta, double eta) {
  bool found = false;
  int maxpos = 0;
  double mx = 0.0;
  for (int i = 0; i < j; i++) {
    which[i] = std::fabs(mu[i]) >= delta;
    found = found || which[i];
    if (std::fabs(mu[i]) > mx) {
      maxpos = i;
      mx = std::fabs(mu[i]);
    }
  }
  if (!found) which[maxpos] = 1;
  for (int i = 0; i < j; i++) {
    if (!which[i]) continue;
    for (int q = i; q >= 0; q--) {
      if (std::fabs(mu[q]) < eta || which[q]) break;
      which[q] = 1;
    }
    for (int q = i + 1; q < j; q++) {
      if (std::fabs(mu[q]) < eta || which[q]) break;
      which[q] = 1;
    }
  }
}

// The sweep: a local step (MGS against the locked columns and v_{j-1}, v_j),
// the omega update, and when some |omega| exceeds delta a second MGS against
// columns k..j (periodic: all, and again at the next step; partial: the
// intervals of compute_int, then all at the next step); after
// norm * delta < anorm * eps1 every later step orthogonalises against all
// columns.  alpha / beta: the projected problem's diagonal and off-diagonal,
// written in place (the recurrence also reads the locked entries).
static int lz_sweep_partial(BV &V, const std::function<int(const void *, void *)> &op, int k, int *M,
                            double *betam, int *breakdown, double anorm, bool periodic, int ncv, double eta_bv,
                            double *alpha, double *beta, KsResult &r) {
  hipStream_t s = stream();
  const int m = *M;
  const double n = (double)V.n, meps = DBL_EPSILON;
  const double eps1 = std::sqrt(n) * meps / 2, delta = LZ_SQRT_EPS / std::sqrt((double)ncv);
  const double eta = std::pow(meps, 0.75) / std::sqrt((double)ncv);
  const size_t sz = (size_t)std::max(m, 100) + 1;
  std::vector<double> omega(sz, 0.0), omega_old(sz, 0.0);
  std::vector<char> which(sz, 0), which2(sz, 0);
  const bool estimate = anorm < 0.0;
  if (estimate) anorm = 1.0;
  for (int i = 0; i < k; i++) which[i] = 1;
  bool force = false, fro = false;
  int lindep = 0;
  double norm = 0.0;
  std::vector<int> cols;
  auto mgs = [&](int c, const char *mask) -> int {   // column c against the masked columns < c
    cols.clear();
    for (int i = 0; i < c; i++)
      if (!mask || mask[i]) cols.push_back(i);
    SA_TRY(bv_mgs_column(V, c, cols.data(), (int)cols.size(), eta_bv, &norm, &lindep));
    r.bytes_bv += (4.0 * cols.size() + 4.0) * V.n * prec_size(V.prec);
    return 0;
  };
  auto coef = [&](int c, int row, double *out) -> int {   // coefficient buffer entry (row, column c)
    SA_HIP(hipMemcpyAsync(out, V.dbuf + (size_t)c * V.m + row, sizeof(double), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    return 0;
  };
  for (int j = k; j < m; j++) {
    SA_TRY(op(col_ptr(V, j), col_ptr(V, j + 1)));
    if (fro) {
      SA_TRY(mgs(j + 1, nullptr));
      SA_TRY(coef(j + 1, j, &alpha[j]));
    } else {
      which[j] = 1;
      if (j - 2 >= k) which[j - 2] = 0;
      SA_TRY(mgs(j + 1, which.data()));
      SA_TRY(coef(j + 1, j, &alpha[j]));
      beta[j] = norm;
      if (estimate) anorm = std::max(anorm, std::fabs(alpha[j]) + norm + (j > k ? beta[j - 1] : 0.0));
      bool reorth = false;
      if (j > k) {
        lz_update_omega(omega.data(), omega_old.data(), j, alpha, beta, eps1, anorm);
        for (int i = 0; i < j - k; i++)
          if (std::fabs(omega[i]) > delta) reorth = true;
      }
      if (reorth || force) {
        for (int i = 0; i < k; i++) which2[i] = 0;
        for (int i = k; i <= j; i++) which2[i] = 1;
        if (periodic) {
          force = !force;
          for (int i = 0; i < j - k; i++) omega[i] = eps1;
        } else if (force) {
          force = false;
        } else {
          force = true;
          lz_compute_int(which2.data() + k, omega.data(), j - k, delta, eta);
          for (int i = 0; i < j - k; i++)
            if (which2[i + k]) omega[i] = eps1;
        }
        SA_TRY(mgs(j + 1, which2.data()));
      }
    }
    if (lindep || norm < n * anorm * meps) {
      *M = j + 1;
      break;
    }
    if (!fro && norm * delta < anorm * eps1) fro = true;
    beta[j] = norm;
    SA_TRY(launch_scale(V.prec, col_ptr(V, j + 1), V.n, 1.0 / norm, s));
  }
  *betam = beta[*M - 1];
  *breakdown = lindep;
  return 0;
}

// ---- EPSSelectiveLanczos (lanczos.c:181-258) --------------------------------
// A local step, then the eigen-decomposition (stevr, DenseTridiagonal
// lanczos.c:139-179) of the tridiagonal block k..j; the Ritz vectors whose
// bound norm |y_last| is below sqrt(eps) ||A|| are formed in AV (again only
// when their number grew) and the new vector is orthogonalised against them
// (BVOrthogonalizeVec, MGS).  beta_j keeps the norm of the local step while
// the vector is scaled by the final norm, as in the reference.
static int lz_sweep_selective(BV &V, BV &AV, const std::function<int(const void *, void *)> &op, int k, int *M,
                              double *betam, int *breakdown, double anorm, double eta_bv, double *alpha,
                              double *beta, KsResult &r) {
  hipStream_t s = stream();
  const int m = *M;
  SA_CHECK(AV.n == V.n && AV.prec == V.prec && AV.m >= m, "selective Lanczos workspace does not match the basis");
  std::vector<double> d(m + 1), e(m + 1), ritz(m + 1), Y((size_t)m * m + 1);
  std::vector<int> cols;
  std::vector<char> which(m + 1, 0);
  for (int i = 0; i < k; i++) which[i] = 1;
  int nritz = 0, lindep = 0;
  double norm = 0.0;
  const double es = (double)prec_size(V.prec);
  for (int j = k; j < m; j++) {
    void *vj1 = col_ptr(V, j + 1);
    SA_TRY(op(col_ptr(V, j), vj1));
    which[j] = 1;
    if (j - 2 >= k) which[j - 2] = 0;
    cols.clear();
    for (int i = 0; i <= j; i++)
      if (which[i]) cols.push_back(i);
    SA_TRY(bv_mgs_column(V, j + 1, cols.data(), (int)cols.size(), eta_bv, &norm, &lindep));
    r.bytes_bv += (4.0 * cols.size() + 4.0) * V.n * es;
    SA_HIP(hipMemcpyAsync(&alpha[j], V.dbuf + (size_t)(j + 1) * V.m + j, sizeof(double), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    beta[j] = norm;
    if (lindep) {
      *M = j + 1;
      break;
    }
    const int nn = j - k + 1;
    for (int i = 0; i < nn; i++) {
      d[i] = alpha[i + k];
      e[i] = beta[i + k];
    }
    const int info = la_stevr(nn, d.data(), e.data(), ritz.data(), Y.data());
    SA_CHECK(info == 0, "LAPACK failure in the selective Lanczos step (stevr)");
    for (int i = 0; i < nn; i++) anorm = std::max(anorm, std::fabs(ritz[i]));
    auto near = [&](int i) { return norm * std::fabs(Y[(size_t)i * nn + nn - 1]) < LZ_SQRT_EPS * anorm; };
    int nritzo = 0;
    for (int i = 0; i < nn; i++) nritzo += near(i);
    if (nritzo > nritz) {
      nritz = 0;
      for (int i = 0; i < nn; i++) {
        if (!near(i)) continue;
        SA_HIP(hipMemcpyAsync(AV.dq, &Y[(size_t)i * nn], sizeof(double) * nn, hipMemcpyHostToDevice, s));
        SA_TRY(launch_multvec(V, k, k + nn, 1.0, AV.dq, 0.0, col_ptr(AV, nritz), s));
        SA_HIP(hipStreamSynchronize(s));
        r.bytes_bv += (nn + 1.0) * V.n * es;
        nritz++;
      }
    }
    if (nritz > 0) {
      SA_TRY(bv_mgs_vec(AV, nritz, vj1, eta_bv, &norm, &lindep));
      r.bytes_bv += (4.0 * nritz + 4.0) * V.n * es;
      if (lindep) {
        *M = j + 1;
        break;
      }
    }
    SA_TRY(launch_scale(V.prec, vj1, V.n, 1.0 / norm, s));
  }
  *betam = beta[*M - 1];
  *breakdown = lindep;
  return 0;
}

// EPSDelayedArnoldi1 (epskrylov.c:120-157; arnoldi.c:84-87 and lanczos.c:500
// with refinement NEVER): the delayed sweep without reorthogonalisation, the
// normalisation alone delayed by one step.  At step j > k the dot products of
// v_{j+1} with v_0..v_j and ||v_j||^2 are taken before v_j is normalised; the
// coefficients are rescaled by ||v_j|| (the diagonal one by its square), v_j
// and v_{j+1} scaled by 1/||v_j||, H(j,j-1) = beta = ||v_j||; then v_{j+1} -=
// V(:,0:j+1) H(:,j).  The last column stays unnormalised and no breakdown is
// reported, as in the reference.
static int lz_sweep_delayed1(BV &V, const std::function<int(const void *, void *)> &op, int k, int m, double *H,
                             int ldh, double *betam, KsResult &r) {
  hipStream_t s = stream();
  const double es = (double)prec_size(V.prec);
  double *qd = V.dq;
  for (int j = k; j < m; j++) {
    SA_TRY(op(col_ptr(V, j), col_ptr(V, j + 1)));
    double *Hj = H + (size_t)ldh * j;
    SA_TRY(launch_dotvec(V, 0, j + 1, col_ptr(V, j + 1), qd, s));       // BVDotColumn(V, j+1)
    SA_HIP(hipMemcpyAsync(Hj, qd, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, s));
    r.bytes_bv += (j + 2.0) * V.n * es;
    double dot = 0.0;
    if (j > k) {                                                          // VecDot(vj, vj)
      SA_TRY(launch_dot_self(V.prec, col_ptr(V, j), V.n, V.ws.dots, V.ws.part, s));
      SA_HIP(hipMemcpyAsync(&dot, V.ws.dots, sizeof(double), hipMemcpyDeviceToHost, s));
      r.bytes_bv += V.n * es;
    }
    SA_HIP(hipStreamSynchronize(s));
    if (j > k) {
      const double norm = std::sqrt(dot);
      SA_TRY(launch_scale(V.prec, col_ptr(V, j), V.n, 1.0 / norm, s));
      H[(size_t)ldh * (j - 1) + j] = norm;
      for (int i = 0; i < j; i++) Hj[i] /= norm;
      Hj[j] /= dot;
      SA_TRY(launch_scale(V.prec, col_ptr(V, j + 1), V.n, 1.0 / norm, s));
      r.bytes_bv += 4.0 * V.n * es;
      *betam = norm;
    }
    SA_HIP(hipMemcpyAsync(qd, Hj, sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));   // BVMultColumn
    SA_TRY(launch_multvec(V, 0, j + 1, -1.0, qd, 1.0, col_ptr(V, j + 1), s));
    SA_HIP(hipStreamSynchronize(s));
    r.bytes_bv += (j + 3.0) * V.n * es;
  }
  return 0;
}

// ---- EPSDelayedArnoldi (epskrylov.c:24-114; lanczos.c:495-512) ------------
// One classical Gram-Schmidt pass per step whose normalisation is delayed by
// one step and whose reorthogonalisation by two: at step j the dot products of
// v_{j+1} and v_j with the earlier columns are taken before v_j is normalised
// and the coefficients rescaled; the unnormalised v_j (t) is reorthogonalised
// against v_0..v_{j-1} and, one step later, normalised into column j.  H: host,
// leading dimension ldh >= m (Lanczos: a scratch T; Arnoldi, arnoldi.c:88-91:
// the projected problem's A).  u, t: device vectors of V.n entries.
static int lz_sweep_delayed(BV &V, const std::function<int(const void *, void *)> &op, int k, int m, double *H,
                            int ldh, double *betam, void *u, void *t, KsResult &r) {
  hipStream_t s = stream();
  SA_CHECK(m - k >= 2, "the delayed Lanczos sweep needs at least two steps (ncv - nconv >= 2)");
  const size_t vb = prec_size(V.prec) * (size_t)V.n;
  const double es = (double)prec_size(V.prec);
  std::vector<double> lhh(m + 1, 0.0);
  double norm1 = 0.0, norm2 = 1.0, h = 0.0;
  double *qd = V.dq;   // device staging of one coefficient vector
  auto dotcol = [&](int c, int nc, double *out) -> int {   // out = V(:, 0:nc)^T V(:, c)
    SA_TRY(launch_dotvec(V, 0, nc, col_ptr(V, c), qd, s));
    SA_HIP(hipMemcpyAsync(out, qd, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    r.bytes_bv += (nc + 1.0) * V.n * es;
    return 0;
  };
  auto multcol = [&](int nc, const double *q, void *y) -> int {   // y -= V(:, 0:nc) q
    SA_HIP(hipMemcpyAsync(qd, q, sizeof(double) * nc, hipMemcpyHostToDevice, s));
    SA_TRY(launch_multvec(V, 0, nc, -1.0, qd, 1.0, y, s));
    SA_HIP(hipStreamSynchronize(s));
    r.bytes_bv += (nc + 2.0) * V.n * es;
    return 0;
  };
  auto sumsq = [&](const void *x, double *out) -> int {
    SA_TRY(launch_dot_self(V.prec, x, V.n, V.ws.dots, V.ws.part, s));
    SA_HIP(hipMemcpyAsync(out, V.ws.dots, sizeof(double), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    return 0;
  };
  for (int j = k; j < m; j++) {
    SA_TRY(op(col_ptr(V, j), col_ptr(V, j + 1)));
    double *Hj = H + (size_t)ldh * j;
    SA_TRY(dotcol(j + 1, j + 1, Hj));
    if (j > k) {
      SA_TRY(dotcol(j, j, lhh.data()));
      double dot = 0.0;
      SA_TRY(sumsq(col_ptr(V, j), &dot));
      if (j > k + 1) {
        SA_TRY(sumsq(u, &h));
        norm2 = std::sqrt(h);
      }
      norm1 = std::sqrt(dot);
      for (int i = 0; i < j; i++) Hj[i] /= norm1;
      Hj[j] /= dot;
      SA_HIP(hipMemcpyAsync(t, col_ptr(V, j), vb, hipMemcpyDeviceToDevice, s));
      SA_TRY(launch_scale(V.prec, col_ptr(V, j), V.n, 1.0 / norm1, s));
      SA_TRY(launch_scale(V.prec, col_ptr(V, j + 1), V.n, 1.0 / norm1, s));
    }
    SA_TRY(multcol(j + 1, Hj, col_ptr(V, j + 1)));
    if (j > k) {
      SA_TRY(multcol(j, lhh.data(), t));
      for (int i = 0; i < j; i++) H[(size_t)ldh * (j - 1) + i] += lhh[i];
    }
    if (j > k + 1) {
      SA_HIP(hipMemcpyAsync(col_ptr(V, j - 1), u, vb, hipMemcpyDeviceToDevice, s));
      SA_TRY(launch_scale(V.prec, col_ptr(V, j - 1), V.n, 1.0 / norm2, s));
      H[(size_t)ldh * (j - 2) + j - 1] = norm2;
    }
    if (j < m - 1) SA_HIP(hipMemcpyAsync(u, t, vb, hipMemcpyDeviceToDevice, s));
  }
  SA_TRY(sumsq(t, &h));
  norm2 = std::sqrt(h);
  SA_TRY(launch_scale(V.prec, t, V.n, 1.0 / norm2, s));
  SA_HIP(hipMemcpyAsync(col_ptr(V, m - 1), t, vb, hipMemcpyDeviceToDevice, s));
  H[(size_t)ldh * (m - 2) + m - 1] = norm2;
  SA_TRY(dotcol(m, m, lhh.data()));
  SA_TRY(multcol(m, lhh.data(), col_ptr(V, m)));
  for (int i = 0; i < m; i++) H[(size_t)ldh * (m - 1) + i] += lhh[i];
  SA_TRY(sumsq(col_ptr(V, m), &h));
  *betam = std::sqrt(h);
  SA_TRY(launch_scale(V.prec, col_ptr(V, m), V.n, 1.0 / *betam, s));
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

// The restart of EPSSolve_Lanczos (lanczos.c:555-690): residual bounds
// beta |y_m| + eps ||A||, the most wanted unconverged Ritz vector as the next
// start vector (built in column nv, then copied to the new column nconv),
// converged pairs moved to the front, V(:,nconv:k) = V(:,nconv:nv)
// Y(nconv:nv, nconv:k), nconv = k.  Local reorthogonalization
// (p.lanczos_local, -eps_lanczos_reorthog local; its sweep is sweep_local in
// ks_solve) adds the restart steps of lanczos.c:570-575 (repeated Ritz
// values), 633-660 (spurious Ritz values: true residual of each converged
// vector through the operator op) and 678-681 (reorthonormalised restart
// vector, MGS).  wtmp: one device vector.
static int lanczos_restart(const KsParams &p, BV &V, const DenseProblem &ds, int &nconv, int nv, double beta,
                           bool breakdown, std::vector<double> &eigr, std::vector<double> &errest, double &anorm,
                           int its, int &reason, std::vector<double> &scratch, std::vector<double> &bnd,
                           std::vector<char> &conv, std::vector<int> &perm, std::vector<double> &Y,
                           std::vector<double> &yrst, KsResult &r,
                           const std::function<int(const void *, void *)> &op, void *wtmp) {
  hipStream_t s = stream();
  const int ld = ds.ld();
  const double es = (double)prec_size(V.prec);
  bnd.assign(ld, 0.0);
  conv.assign(ld, 'N');
  Y.assign(ds.Q(), ds.Q() + (size_t)ld * ld);
  for (int i = nconv; i < nv; i++) anorm = std::max(anorm, std::fabs(eigr[i]));
  for (int i = nconv; i < nv; i++) {
    const double res = beta * std::fabs(Y[nv - 1 + (size_t)i * ld]) + DBL_EPSILON * anorm;
    SA_TRY(eps_converged(p, eigr[i], 0.0, res, &bnd[i]));   // (*eps->converged), lanczos.c:564
    conv[i] = bnd[i] < p.tol ? 'C' : 'N';
  }
  if (p.lanczos_local)   // purge repeated Ritz values
    for (int i = nconv + 1; i < nv; i++)
      if (conv[i] == 'C' && std::fabs((eigr[i] - eigr[i - 1]) / eigr[i]) < p.tol) conv[i] = 'R';
  if (!breakdown) {
    int rs = nconv;
    while (rs < nv && conv[rs] != 'N') rs++;
    if (rs >= nv) {
      breakdown = true;
    } else {
      for (int i = rs + 1; i < nv; i++)
        if (conv[i] == 'N' && sc_compare(p.which, eigr[rs], 0.0, eigr[i], 0.0, p.target) > 0) rs = i;
      // BVMultColumn(V,1,0,nv,y): V(:,nv) = V(:,nconv:nv) Y(nconv:nv, rs)
      yrst.assign(Y.begin() + (size_t)rs * ld + nconv, Y.begin() + (size_t)rs * ld + nv);
      SA_HIP(hipMemcpy(V.ws.dots, yrst.data(), sizeof(double) * yrst.size(), hipMemcpyHostToDevice));
      SA_TRY(launch_multvec(V, nconv, nv, 1.0, V.ws.dots, 0.0, col_ptr(V, nv), s));
      r.bytes_bv += ((double)(nv - nconv) + 1.0) * V.n * es;
    }
  }
  // count converged pairs and put them first
  perm.resize(nv);
  for (int i = nconv; i < nv; i++) perm[i] = i;
  int k;
  for (k = nconv; k < nv; k++) {
    if (conv[perm[k]] != 'C') {
      int j = k + 1;
      while (j < nv && conv[perm[j]] != 'C') j++;
      if (j >= nv) break;
      std::swap(perm[k], perm[j]);
    }
  }
  for (int i = nconv; i < k; i++) {
    const int x = perm[i];
    if (x != i) {
      int j = i + 1;
      while (perm[j] != i) j++;
      std::swap(eigr[x], eigr[i]);
      std::swap(bnd[x], bnd[i]);
      std::swap(conv[x], conv[i]);
      perm[j] = x;
      perm[i] = i;
      for (int q = 0; q < nv; q++) std::swap(Y[q + (size_t)x * ld], Y[q + (size_t)i * ld]);
    }
  }
  if (k > nconv) {   // the converged Ritz vectors
    std::memcpy(V.hq, Y.data(), sizeof(double) * (size_t)ld * ld);
    SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * (size_t)ld * ld, hipMemcpyHostToDevice, s));
    SA_TRY(launch_multinplace(V, nconv, nv, nconv, k, V.dq, ld, s));
    SA_HIP(hipStreamSynchronize(s));
    r.bytes_restart += ((double)(nv - nconv) + (double)(k - nconv)) * V.n * es;
  }
  if (p.lanczos_local && k > nconv) {   // purge spurious Ritz values
    double h = 0.0;
    for (int i = nconv; i < k; i++) {
      void *vi = col_ptr(V, i);
      SA_TRY(launch_dot_self(V.prec, vi, V.n, V.ws.dots, V.ws.part, s));
      SA_HIP(hipMemcpyAsync(&h, V.ws.dots, sizeof(double), hipMemcpyDeviceToHost, s));
      SA_HIP(hipStreamSynchronize(s));
      SA_TRY(launch_scale(V.prec, vi, V.n, 1.0 / std::sqrt(h), s));
      SA_TRY(op(vi, wtmp));
      SA_TRY(launch_axpy_norm(V.prec, wtmp, vi, eigr[i], V.n, V.ws.part, V.ws.dots, s));
      SA_HIP(hipMemcpyAsync(&h, V.ws.dots, sizeof(double), hipMemcpyDeviceToHost, s));
      SA_HIP(hipStreamSynchronize(s));
      const double nr = std::sqrt(h);
      SA_TRY(eps_converged(p, eigr[i], 0.0, nr, &bnd[i]));   // lanczos.c:644
      if (bnd[i] >= p.tol) conv[i] = 'S';
    }
    int i;
    for (i = nconv; i < k; i++) {
      if (conv[i] != 'C') {
        int j = i + 1;
        while (j < k && conv[j] != 'C') j++;
        if (j >= k) break;
        std::swap(eigr[j], eigr[i]);
        std::swap(bnd[j], bnd[i]);
        std::swap(conv[j], conv[i]);
        SA_HIP(hipMemcpyAsync(wtmp, col_ptr(V, i), es * V.n, hipMemcpyDeviceToDevice, s));
        SA_HIP(hipMemcpyAsync(col_ptr(V, i), col_ptr(V, j), es * V.n, hipMemcpyDeviceToDevice, s));
        SA_HIP(hipMemcpyAsync(col_ptr(V, j), wtmp, es * V.n, hipMemcpyDeviceToDevice, s));
      }
    }
    k = i;
  }
  for (int i = nconv; i < nv; i++) errest[i] = bnd[i];
  nconv = k;
  SA_TRY(eps_stopping(p, its, k, p.nev, &reason));
  if (reason == 0) {   // BVCopyColumn(V,nv,nconv); a random vector after a breakdown
    SA_HIP(hipMemcpyAsync(col_ptr(V, nconv), col_ptr(V, nv), es * V.n, hipMemcpyDeviceToDevice, s));
    r.bytes_bv += 2.0 * V.n * es;
    if (p.lanczos_local && !breakdown) {   // reorthonormalise the restart vector (MGS, all columns)
      std::vector<int> cols(nconv);
      for (int i = 0; i < nconv; i++) cols[i] = i;
      double nrm = 0.0;
      int lind = 0;
      SA_TRY(bv_mgs_column(V, nconv, cols.data(), nconv, p.eta, &nrm, &lind));
      if (nrm != 1.0 && nrm != 0.0) SA_TRY(launch_scale(V.prec, col_ptr(V, nconv), V.n, 1.0 / nrm, s));
      if (lind) breakdown = true;
    }
    if (breakdown) {
      int bd = 0;
      SA_TRY(eps_start_vector(V, nconv, false, p.eta, &bd, scratch, EPS_SEED_RIGHT, p.defl, p.ndefl));
      if (bd) reason = -2;
    }
  }
  return 0;
}

int ritz_true_residual(BV &V, const Mat &A, int nv, const double *Z, int ldz, int nx, double re, double im,
                       void *work, double *part, double *resnorm) {
  hipStream_t s = stream();
  for (int c = 0; c < nx; c++) std::memcpy(V.hq + (size_t)c * ldz, Z + (size_t)c * ldz, sizeof(double) * nv);
  SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * (size_t)ldz * nx, hipMemcpyHostToDevice, s));
  const size_t cb = prec_size(V.prec) * (size_t)V.ld;
  for (int c = 0; c < nx; c++)
    SA_TRY(launch_multvec(V, 0, nv, 1.0, V.dq + (size_t)c * ldz, 0.0, static_cast<char *>(work) + c * cb, s));
  const double lr2[2] = {re, re}, li2[2] = {nx == 2 ? im : 0.0, nx == 2 ? -im : 0.0};
  double rn[2] = {0.0, 0.0};
  SA_TRY(residual_norms(A, work, V.ld, nx, lr2, li2, rn, part, part + (size_t)RESID_MAXVEC * RESID_MAXG, s));
  *resnorm = rn[0];
  return 0;
}

int ks_solve(const KsParams &p, const Mat &A, BV &V, bool have_initial, KsResult &r) {
  SA_CHECK(lapack_ready(), "no LAPACK library found for the projected problem");
  const int n = A.n, ncv = p.ncv, nev = p.nev, mpd = p.mpd, ld = ncv + 1;
  SA_CHECK(V.m == ncv + 1 && V.n == n && V.prec == A.prec, "basis does not match the operator/ncv");
  SA_CHECK(ncv + 1 < GS_MAXC, "ncv too large for the fused Gram-Schmidt kernels");
  SA_CHECK(p.refine >= GS_REFINE_IFNEEDED && p.refine <= GS_REFINE_ALWAYS, "invalid refinement type");
  V.refine = p.refine;   // BVSetOrthogonalization of the solver's basis
  // EPSARNOLDI (arnoldi.c:52-131, setup :31-50): explicit restart and
  // always a DSNHEP projected problem; ncv <= nev + mpd
  const bool arn = p.solver == 1;
  // EPSLANCZOS (lanczos.c:518-697, setup :36-78): explicit restart from the
  // most wanted unconverged Ritz vector, DSHEP, converged pairs moved to the
  // front and locked; the sweep follows the reorthogonalization type (full:
  // bv_sweep; local, periodic, partial, selective, delayed: the lz_ sweeps)
  const bool lanc = p.solver == 2;
  // Harmonic extraction (krylovschur.c:238-240,270-271,311-320; arnoldi.c:65,
  // 96-97): the projected matrix is translated by g b^T (DSTranslateHarmonic)
  // before each solve, always as a non-symmetric problem, and residual bounds
  // are scaled by gamma; Krylov-Schur undoes the translation at the restart.
  const bool harm = p.extraction == 1;
  const bool dsherm = p.hermitian && !arn && !harm;
  SA_CHECK(!((arn || lanc) && ncv > nev + mpd), "The value of ncv must not be larger than nev+mpd");
  SA_CHECK(!lanc || p.hermitian, "EPSLANCZOS needs a symmetric problem");
  SA_CHECK(!(lanc && harm), "EPSLANCZOS does not support harmonic extraction");
  std::vector<double> g(ld, 0.0);
  double gamma = 1.0;
  // Shift-and-invert (sinvert.c; epssetup.c:119): each operator application
  // is a solve with A - sigma I; the DS compares back-transformed values and
  // the eigenvalues are mapped back (EPSComputeValues) before the final sort.
  const bool sinv = p.st_type == 1;
  SA_CHECK(!sinv || (p.which == SA_TARGET_MAGNITUDE || p.which == SA_TARGET_REAL || p.which == SA_WHICH_USER),
           "Shift-and-invert requires a target 'which' (target magnitude, target real or user)");
  SA_CHECK(!(sinv && harm), "Cannot use a spectral transformation combined with harmonic extraction");
  KspWork kw;
  struct KspFree {
    KspWork &w;
    ~KspFree() { ksp_free(w); }
  } kw_free{kw};
  if (sinv) SA_TRY(ksp_alloc(kw, A, p.st_shift, p.ksp_jacobi, p.ksp_direct, p.ksp_gmres, p.ksp_restart));
  // y = Op x: the operator of the spectral transformation (local Lanczos)
  std::function<int(const void *, void *)> op_apply = [&](const void *x, void *y) -> int {
    if (!sinv) return launch_spmv(A, x, y, stream());
    int its = 0;
    const int rc = ksp_solve(A, kw, (const double *)x, (double *)y, p.ksp_rtol, p.ksp_max_it, &its);
    r.ksp_its += its;
    return rc;
  };
  struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
  } wtmp;
  const bool delayed = (lanc && p.lanczos_partial == 4) || (arn && p.arnoldi_delayed);
  if ((lanc && p.lanczos_local) || delayed)   // local: one vector; delayed: u and t
    SA_HIP(hipMalloc(&wtmp.p, prec_size(V.prec) * (size_t)std::max(1, n) * (delayed ? 2 : 1)));
  // selective: the basis of nearly converged Ritz vectors (BVDuplicate of V, lanczos.c:63-71)
  struct BVHold {
    BV b;
    ~BVHold() { bv_free(b); }
  } avh;
  BV &AV = avh.b;
  if (lanc && p.lanczos_partial == 3) SA_TRY(bv_alloc(AV, V.n, V.m, V.prec));
  // EPSSetTrueResidual (epsopts.c:1333-1345) for Krylov-Schur and Arnoldi: the
  // convergence test forms each candidate Ritz vector x = V(:,0:nv) X(:,k)
  // (EPSComputeRitzVector, epsdefault.c:386-398) and takes ||A x - lambda x||
  // with the back-transformed value (EPSComputeResidualNorm_Private), in
  // place of beta |y_last| (epskrylov.c:225-265).  Work: two columns for the
  // real and imaginary parts, the residual partials and their sums.
  struct TrueWork {
    void *x = nullptr;
    double *part = nullptr;
    ~TrueWork() {
      if (x) (void)hipFree(x);
      if (part) (void)hipFree(part);
    }
  } trw;
  const bool trueres = p.trueres && !lanc;
  if (trueres) {
    SA_HIP(hipMalloc(&trw.x, prec_size(V.prec) * (size_t)V.ld * 2));
    SA_HIP(hipMalloc((void **)&trw.part, sizeof(double) * ((size_t)RESID_MAXVEC * RESID_MAXG + RESID_MAXVEC)));
  }
  // EPSLocalLanczos (lanczos.c:89-118): V(:,j+1) = Op V(:,j) orthogonalised
  // by MGS against the columns before k and V(:,j-1), V(:,j) only; the
  // coefficient buffer gets alpha_j (row j) and beta_j = the norm (row j+1,
  // also on breakdown), as the full sweep leaves them.
  auto sweep_local = [&](int k, int *m, double *beta, int *bd) -> int {
    std::vector<int> cols;
    int lindep = 0;
    for (int j = k; j < *m; j++) {
      SA_TRY(op_apply(col_ptr(V, j), col_ptr(V, j + 1)));
      cols.clear();
      for (int i = 0; i < k; i++) cols.push_back(i);
      if (j - 1 >= k) cols.push_back(j - 1);
      cols.push_back(j);
      double nrm = 0.0;
      SA_TRY(bv_mgs_column(V, j + 1, cols.data(), (int)cols.size(), p.eta, &nrm, &lindep));
      SA_HIP(hipMemcpyAsync(V.dbuf + (size_t)(j + 1) * V.m + j + 1, &nrm, sizeof(double), hipMemcpyHostToDevice,
                            stream()));
      SA_HIP(hipStreamSynchronize(stream()));
      r.bytes_bv += (3.0 * cols.size() + 3.0) * V.n * prec_size(V.prec);
      *beta = nrm;
      if (lindep) {
        *m = j + 1;
        break;
      }
      SA_TRY(launch_scale(V.prec, col_ptr(V, j + 1), V.n, 1.0 / nrm, stream()));
    }
    *bd = lindep;
    return 0;
  };
  auto sweep = [&](int k0, int *nv, double *beta, int *bd) -> int {
    if (p.defl)
      return bv_sweep_deflated(V, A, sinv ? &kw : nullptr, p.ksp_rtol, p.ksp_max_it, *p.defl, p.ndefl, k0, nv, p.eta,
                               beta, bd, &r.ksp_its);
    if (lanc && p.lanczos_local) return sweep_local(k0, nv, beta, bd);
    if (!sinv) return bv_sweep(V, A, k0, nv, p.eta, beta, bd);
    return bv_sweep_sinvert(V, A, kw, p.ksp_rtol, p.ksp_max_it, k0, nv, p.eta, beta, bd, &r.ksp_its);
  };
  double anorm = -1.0;              // running estimate of ||A|| (Lanczos)
  std::vector<double> bnd, Yl, yrst;
  std::vector<char> conv;
  std::vector<int> lperm;
  hipStream_t s = stream();
  DenseProblem ds;
  ds.allocate(ld, dsherm, p.which);
  ds.set_target(p.target);
  ds.set_sinvert(sinv, p.st_shift);
  std::vector<double> eigr(ld, 0.0), eigi(ld, 0.0), errest(ld, 0.0), buf, scratch, locked;
  const double t0 = now();
  double tds = 0.0;
  r = KsResult();
  const int64_t passes0 = V.passes_done;
  const double bsp = spmv_bytes(A), es = (double)prec_size(V.prec);

  if (have_initial) {  // BVInsertVecs(V,0,&k,IS,TRUE) (bvfunc.c:331-374)
    double norm = 0.0;
    int lindep = 0;
    if (p.defl) SA_TRY(bv_project_out(*p.defl, p.ndefl, col_ptr(V, 0), s));
    SA_TRY(bv_orthonormalize(V, 0, p.eta, &norm, &lindep));
    SA_CHECK(!(norm == 0.0 || lindep), "initial vector is linearly dependent");
    r.bytes_bv += 3.0 * V.n * es;
  }
  SA_TRY(eps_start_vector(V, 0, have_initial, p.eta, nullptr, scratch, EPS_SEED_RIGHT, p.defl, p.ndefl));
  r.bytes_bv += 3.0 * V.n * es;

  int l = 0, nconv = 0, its = 0, reason = 0;
  while (reason == 0) {
    its++;
    int nv = std::min(nconv + mpd, ncv);
    const int k0 = nconv + l;
    ds.set_dimensions(nv, nconv, (arn || lanc) ? 0 : k0);
    double beta = 0.0;
    int breakdown = 0;
    const int64_t pbefore = V.passes_done;
    const bool partial = lanc && p.lanczos_partial != 0;
    if (partial && p.lanczos_partial == 3) {
      SA_TRY(lz_sweep_selective(V, AV, op_apply, k0, &nv, &beta, &breakdown, anorm, p.eta, ds.d(), ds.e(), r));
    } else if (partial && p.lanczos_partial == 4) {   // lanczos.c:495-512
      std::vector<double> Td((size_t)nv * nv, 0.0);
      const size_t vb = prec_size(V.prec) * (size_t)n;
      if (p.refine == GS_REFINE_NEVER) SA_TRY(lz_sweep_delayed1(V, op_apply, k0, nv, Td.data(), nv, &beta, r));
      else SA_TRY(lz_sweep_delayed(V, op_apply, k0, nv, Td.data(), nv, &beta, wtmp.p,
                                   static_cast<char *>(wtmp.p) + vb, r));
      breakdown = 0;
      double *d = ds.d(), *e = ds.e();
      for (int i = k0; i < nv - 1; i++) {
        d[i] = Td[(size_t)nv * i + i];
        e[i] = Td[(size_t)nv * i + i + 1];
      }
      d[nv - 1] = Td[(size_t)nv * (nv - 1) + nv - 1];
      e[nv - 1] = beta;
    } else if (arn && p.arnoldi_delayed) {   // arnoldi.c:88-91
      if (p.refine == GS_REFINE_NEVER) SA_TRY(lz_sweep_delayed1(V, op_apply, k0, nv, ds.A(), ld, &beta, r));
      else SA_TRY(lz_sweep_delayed(V, op_apply, k0, nv, ds.A(), ld, &beta, wtmp.p,
                                   static_cast<char *>(wtmp.p) + prec_size(V.prec) * (size_t)n, r));
      breakdown = 0;
    } else if (partial)
      SA_TRY(lz_sweep_partial(V, op_apply, k0, &nv, &beta, &breakdown, anorm, p.lanczos_partial == 1, ncv, p.eta,
                              ds.d(), ds.e(), r));
    else
      SA_TRY(sweep(k0, &nv, &beta, &breakdown));
    r.matvecs += nv - k0;
    // Algorithmic bytes of the sweep: SpMV, then per Gram-Schmidt pass the dot
    // products read columns 0..c-1 and the update reads them again plus v and
    // writes v, i.e. (2c+2) n s per pass; passes averaged over the sweep.
    if (nv > k0 && !(lanc && (p.lanczos_local || partial)) && !(arn && p.arnoldi_delayed)) {
      // (the other sweeps count their own bytes)
      const double avgp = (double)(V.passes_done - pbefore) / (double)(nv - k0);
      for (int c = k0 + 1; c <= nv; c++) r.bytes_bv += avgp * (2.0 * c + 2.0) * V.n * es;
    }
    r.bytes_spmv += (nv - k0) * bsp;
    // coefficients of columns k0+1..nv (BVMatArnoldi/BVMatLanczos copy-out, bvkrylov.c:100-109,210-222)
    const int m = V.m;
    buf.resize((size_t)m * (nv - k0));
    SA_HIP(hipMemcpyAsync(buf.data(), V.dbuf + (size_t)(k0 + 1) * m, sizeof(double) * m * (nv - k0),
                          hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    const double td0 = now();
    auto B = [&](int row, int colc) { return buf[(size_t)(colc - k0 - 1) * m + row]; };
    if (partial || (arn && p.arnoldi_delayed)) {
      // the sweep wrote the projected problem itself
    } else if (dsherm) {
      double *d = ds.d(), *e = ds.e();
      for (int j = k0; j < nv; j++) {
        d[j] = B(j, j + 1);
        e[j] = B(j + 1, j + 1);
      }
    } else {
      double *H = ds.A();
      for (int j = k0; j < nv - 1; j++)
        for (int i = 0; i < j + 2; i++) H[i + (size_t)j * ld] = B(i, j + 1);
      for (int i = 0; i < nv; i++) H[i + (size_t)(nv - 1) * ld] = B(i, nv);
      if (ld > nv) H[nv + (size_t)(nv - 1) * ld] = B(nv, nv);
    }
    ds.set_dimensions(nv, nconv, (arn || lanc) ? 0 : k0);
    ds.set_state(l ? DS_RAW : DS_INTERMEDIATE);
    if (harm) {
      const int hinfo = ds.translate_harmonic(p.target, beta, false, g.data(), &gamma);
      SA_CHECK(hinfo == 0, "LAPACK failure in the harmonic translation (getrf/getrs)");
    }
    // EPSLanczos keeps the projected eigenvalues apart from eps->eigr and
    // stores only the entries from nconv on (lanczos.c:663-666): the leading
    // diagonal of the projected problem is not that of the locked values once
    // the converged values have been moved first.
    if (lanc) locked.assign(eigr.begin(), eigr.begin() + nconv);
    int info = ds.solve(eigr.data(), eigi.data());
    SA_CHECK(info == 0, "LAPACK failure in the projected problem (solve)");
    info = ds.sort(eigr.data(), eigi.data());
    SA_CHECK(info == 0, "LAPACK failure in the projected problem (sort)");
    if (lanc) {
      SA_TRY(lanczos_restart(p, V, ds, nconv, nv, beta, breakdown != 0, eigr, errest, anorm, its, reason, scratch,
                             bnd, conv, lperm, Yl, yrst, r, op_apply, wtmp.p));
      std::copy(locked.begin(), locked.end(), eigr.begin());
      tds += now() - td0;
      SA_TRY(eps_monitor(p, its, nconv, eigr.data(), eigi.data(), errest.data(), nv));
      continue;
    }
    ds.update_extra_row();
    // EPSKrylovConvergence (epskrylov.c:207-295) with EPSConvergedRelative (epsdefault.c:224-232)
    int k = nconv, marker = -1;
    for (k = nconv; k < nv; k++) {
      double resnorm = 0.0;
      int cplx = 0;
      info = ds.ritz_residual(k, &resnorm, &cplx);
      SA_CHECK(info == 0, "LAPACK failure in the projected problem (vectors)");
      double re = eigr[k], im = eigi[k];
      // STBackTransform (sinvert.c:16-38) for the true residual and for the
      // norm-relative test; the other tests take the transformed value
      // (epskrylov.c:245-251)
      if (sinv && (trueres || p.conv == SA_CONV_NORM)) {
        if (im == 0.0) {
          re = 1.0 / re + p.st_shift;
        } else {
          const double t = re * re + im * im;
          re = re / t + p.st_shift;
          im = -im / t;
        }
      }
      if (trueres) {
        const double *Xk = dsherm ? ds.Q() + (size_t)k * ld : ds.X() + (size_t)k * ld;
        const int nx = cplx ? 2 : 1;
        SA_TRY(ritz_true_residual(V, A, nv, Xk, ld, nx, re, im, trw.x, trw.part, &resnorm));
        r.true_residuals++;
        r.bytes_bv += nx * ((double)nv + 1.0) * V.n * es;
        r.bytes_spmv += nx * bsp;
      } else {
        resnorm *= beta * gamma;
      }
      SA_TRY(eps_converged(p, re, im, resnorm, &errest[k]));
      if (marker == -1 && errest[k] >= p.tol) marker = k;
      if (cplx) {
        errest[k + 1] = errest[k];
        k++;
      }
      if (marker != -1) break;
    }
    if (marker != -1) k = marker;
    if (arn) {
      // BVMultInPlace(V,Q,nconv,min(k+1,nv)) (arnoldi.c:109-115): the
      // (k+1)-th Schur vector starts the next sweep
      tds += now() - td0;
      const int e = std::min(k + 1, nv);
      if (e > nconv) {
        std::memcpy(V.hq, ds.Q(), sizeof(double) * (size_t)ld * ld);
        SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * (size_t)ld * ld, hipMemcpyHostToDevice, s));
        SA_TRY(launch_multinplace(V, nconv, nv, nconv, e, V.dq, ld, s));
        SA_HIP(hipStreamSynchronize(s));   // the staging buffer is rewritten next iteration
        r.bytes_restart += ((double)(nv - nconv) + (double)(e - nconv)) * V.n * es;
      }
      SA_TRY(eps_stopping(p, its, k, nev, &reason));
      if (reason == 0 && breakdown) {    // arnoldi.c:118-125
        int bd = 0;
        SA_TRY(eps_start_vector(V, k, false, p.eta, &bd, scratch, EPS_SEED_RIGHT, p.defl, p.ndefl));
        if (bd) reason = -2;
      }
      nconv = k;
      SA_TRY(eps_monitor(p, its, nconv, eigr.data(), eigi.data(), errest.data(), nv));
      continue;
    }
    // the stopping test (krylovschur.c:284; EPSStoppingBasic by default)
    SA_TRY(eps_stopping(p, its, k, nev, &reason));
    int newl = 0;
    if (!(reason != 0 || breakdown || k == nv)) {
      newl = std::max(1, (int)((nv - k) * p.keep));
      if (!p.hermitian) newl = ds.truncate_size(k, nv, newl);
    }
    l = newl;
    const int kconv = k;
    if (!p.lock && l > 0) {   // non-locking variant (krylovschur.c:295): reset the converged count
      l += k;
      k = 0;
    }
    tds += now() - td0;
    if (reason == 0) {
      if (breakdown || k == nv) {
        if (k < nev) {
          int bd = 0;
          SA_TRY(eps_start_vector(V, k, false, p.eta, &bd, scratch, EPS_SEED_RIGHT, p.defl, p.ndefl));
          if (bd) reason = -2;
        }
      } else {
        if (harm) {   // undo the translation: gamma u^ = u - U g~ (krylovschur.c:311-320)
          ds.set_dimensions(nv, k, l);
          const int hinfo = ds.translate_harmonic(0.0, beta, true, g.data(), &gamma);
          SA_CHECK(hinfo == 0, "harmonic translation failed");
          std::memcpy(V.hq, g.data(), sizeof(double) * nv);
          SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * nv, hipMemcpyHostToDevice, s));
          SA_TRY(launch_multvec(V, 0, nv, -1.0, V.dq, 1.0, col_ptr(V, nv), s));
          SA_TRY(launch_scale(V.prec, col_ptr(V, nv), V.n, 1.0 / gamma, s));
          SA_HIP(hipStreamSynchronize(s));   // the staging buffer takes Q next
          r.bytes_bv += ((double)nv + 4.0) * V.n * es;
          ds.set_dimensions(nv, k, nv);
        }
        ds.truncate(k + l);
      }
    }
    // BVMultInPlace(V,Q,nconv,k+l) over active columns [nconv,nv) (bvops.c:220-245)
    if (k + l > nconv) {
      std::memcpy(V.hq, ds.Q(), sizeof(double) * (size_t)ld * ld);
      SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * (size_t)ld * ld, hipMemcpyHostToDevice, s));
      SA_TRY(launch_multinplace(V, nconv, nv, nconv, k + l, V.dq, ld, s));
      r.bytes_restart += ((double)(nv - nconv) + (double)(k + l - nconv)) * V.n * es;
    }
    if (reason == 0 && !breakdown) {  // BVCopyColumn(V,nv,k+l)
      SA_HIP(hipMemcpyAsync(col_ptr(V, k + l), col_ptr(V, nv), es * V.n, hipMemcpyDeviceToDevice, s));
      r.bytes_bv += 2.0 * V.n * es;
    }
    nconv = k;
    SA_TRY(eps_monitor(p, its, kconv, eigr.data(), eigi.data(), errest.data(), nv));   // krylovschur.c:344
  }
  // DSTruncate(nconv,TRUE); EPSComputeVectors (Schur: V <- V X).  Lanczos
  // leaves Ritz vectors in the leading columns already.
  const double td1 = now();
  if (!lanc) {
    int info = ds.trim_and_vectors(nconv);
    SA_CHECK(info == 0, "LAPACK failure computing eigenvectors of the projected problem");
  }
  tds += now() - td1;
  if (!dsherm && nconv > 0) {
    // the last restart may still be reading the pinned staging buffer
    SA_HIP(hipStreamSynchronize(s));
    std::memcpy(V.hq, ds.X(), sizeof(double) * (size_t)ld * ld);
    SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * (size_t)ld * ld, hipMemcpyHostToDevice, s));
    SA_TRY(launch_multinplace(V, 0, nconv, 0, nconv, V.dq, ld, s));
  }
  // EPSComputeValues: STBackTransform_Sinvert (sinvert.c:16-38)
  if (sinv)
    for (int i = 0; i < nconv; i++) {
      if (eigi[i] == 0.0) {
        eigr[i] = 1.0 / eigr[i] + p.st_shift;
      } else {
        const double t = eigr[i] * eigr[i] + eigi[i] * eigi[i];
        eigr[i] = eigr[i] / t + p.st_shift;
        eigi[i] = -eigi[i] / t;
      }
    }
  // reorder conjugate pairs, positive imaginary part first (epssolve.c:159-176)
  for (int i = 0; i < nconv - 1; i++) {
    if (eigi[i] != 0) {
      if (eigi[i] < 0) {
        eigi[i] = -eigi[i];
        eigi[i + 1] = -eigi[i + 1];
        SA_TRY(launch_scale(V.prec, col_ptr(V, i + 1), V.n, -1.0, s));
      }
      i++;
    }
  }
  // SlepcSortEigenvalues (slepcsc.c:89-140)
  std::vector<int> perm(nconv);
  for (int i = 0; i < nconv; i++) perm[i] = i;
  for (int i = nconv - 1; i >= 0; i--) {
    const double re = eigr[perm[i]];
    double im = eigi[perm[i]];
    int j = i + 1;
    if (im != 0) {
      i--;
      im = eigi[perm[i]];
    }
    while (j < nconv) {
      if (sc_compare(p.which, re, im, eigr[perm[j]], eigi[perm[j]], p.target) <= 0) break;
      if (!im) {
        if (eigi[perm[j]] == 0.0) {
          std::swap(perm[j - 1], perm[j]);
          j++;
        } else {
          const int tmp = perm[j - 1];
          perm[j - 1] = perm[j];
          perm[j] = perm[j + 1];
          perm[j + 1] = tmp;
          j += 2;
        }
      } else {
        if (eigi[perm[j]] == 0.0) {
          const int tmp = perm[j - 2];
          perm[j - 2] = perm[j];
          perm[j] = perm[j - 1];
          perm[j - 1] = tmp;
          j++;
        } else {
          std::swap(perm[j - 2], perm[j]);
          std::swap(perm[j - 1], perm[j + 1]);
          j += 2;
        }
      }
    }
  }
  SA_HIP(hipStreamSynchronize(s));
  r.seconds = now() - t0;
  r.seconds_ds = tds;
  r.nconv = nconv;
  r.its = its;
  r.reason = reason;
  r.gs_passes = V.passes_done - passes0;
  r.perm = perm;
  r.eigr.assign(eigr.begin(), eigr.begin() + nconv);
  r.eigi.assign(eigi.begin(), eigi.begin() + nconv);
  r.errest.assign(errest.begin(), errest.begin() + nconv);
  return 0;
}

}  // namespace sa
