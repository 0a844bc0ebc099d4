// Host FP64 projected problem with run-time resolved LAPACK.
// Follows src/sys/classes/ds/impls/hep/dshep.c, nhep/dsnhep.c and dsutil.c.
#include "sa_ds.h"

#include <dlfcn.h>
#include <glob.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace sa {

namespace {
typedef void (*f_lartg)(const double *, const double *, double *, double *, double *);
typedef void (*f_rot)(const int *, double *, const int *, double *, const int *, const double *, const double *);
typedef void (*f_steqr)(const char *, const int *, double *, double *, double *, const int *, double *, int *, size_t);
typedef void (*f_gehrd)(const int *, const int *, const int *, double *, const int *, double *, double *, const int *, int *);
typedef void (*f_orghr)(const int *, const int *, const int *, double *, const int *, const double *, double *, const int *,
                        int *);
typedef void (*f_hseqr)(const char *, const char *, const int *, const int *, const int *, double *, const int *, double *,
                        double *, double *, const int *, double *, const int *, int *, size_t, size_t);
typedef void (*f_trexc)(const char *, const int *, double *, const int *, double *, const int *, int *, int *, double *, int *,
                        size_t);
typedef void (*f_trevc)(const char *, const char *, int *, const int *, const double *, const int *, double *, const int *,
                        double *, const int *, const int *, int *, double *, int *, size_t, size_t);
typedef void (*f_potrf)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*f_stevr)(const char *, const char *, const int *, double *, double *, const double *,
                        const double *, const int *, const int *, const double *, int *, double *, double *,
                        const int *, int *, double *, const int *, int *, const int *, int *, size_t, size_t);
typedef void (*f_trtri)(const char *, const char *, const int *, double *, const int *, int *, size_t, size_t);
typedef void (*f_syev)(const char *, const char *, const int *, double *, const int *, double *, double *, const int *,
                       int *, size_t, size_t);
typedef void (*f_getrf)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*f_getrs)(const char *, const int *, const int *, const double *, const int *, const int *, double *,
                        const int *, int *, size_t);

struct Lapack {
  bool ok = false;
  std::string path;
  f_lartg lartg = nullptr;
  f_rot rot = nullptr;
  f_steqr steqr = nullptr;
  f_gehrd gehrd = nullptr;
  f_orghr orghr = nullptr;
  f_hseqr hseqr = nullptr;
  f_trexc trexc = nullptr;
  f_trevc trevc = nullptr;
  f_potrf potrf = nullptr;   // block orthogonalisation (BVOrthogonalize)
  f_stevr stevr = nullptr;   // selective Lanczos (optional)
  f_trtri trtri = nullptr;
  f_syev syev = nullptr;
  f_getrf getrf = nullptr;   // harmonic extraction (DSTranslateHarmonic)
  f_getrs getrs = nullptr;
};
Lapack LA;
std::once_flag la_once;

void *sym(void *h, const char *name) {
  std::string a = std::string(name) + "_";
  void *p = dlsym(h, a.c_str());
  if (!p) p = dlsym(h, ("scipy_" + a).c_str());
  return p;
}

bool try_lib(const char *path) {
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return false;
  Lapack L;
  L.lartg = (f_lartg)sym(h, "dlartg");
  L.rot = (f_rot)sym(h, "drot");
  L.steqr = (f_steqr)sym(h, "dsteqr");
  L.gehrd = (f_gehrd)sym(h, "dgehrd");
  L.orghr = (f_orghr)sym(h, "dorghr");
  L.hseqr = (f_hseqr)sym(h, "dhseqr");
  L.trexc = (f_trexc)sym(h, "dtrexc");
  L.trevc = (f_trevc)sym(h, "dtrevc");
  L.potrf = (f_potrf)sym(h, "dpotrf");
  L.trtri = (f_trtri)sym(h, "dtrtri");
  L.syev = (f_syev)sym(h, "dsyev");
  L.getrf = (f_getrf)sym(h, "dgetrf");
  L.getrs = (f_getrs)sym(h, "dgetrs");
  L.stevr = (f_stevr)sym(h, "dstevr");
  if (!(L.lartg && L.rot && L.steqr && L.gehrd && L.orghr && L.hseqr && L.trexc && L.trevc && L.potrf && L.trtri &&
        L.syev && L.getrf && L.getrs)) {
    dlclose(h);
    return false;
  }
  if (auto layer = (void (*)(int))dlsym(h, "MKL_Set_Threading_Layer")) layer(1);  // sequential
  // Projected problems are <= 129 x 129: one BLAS thread per caller (an
  // OpenBLAS pool sized to every core of the host only adds wake-up latency,
  // and concurrent solves run their DS on their own host threads).
  for (const char *nm : {"scipy_openblas_set_num_threads", "openblas_set_num_threads"})
    if (auto setnt = (void (*)(int))dlsym(h, nm)) {
      setnt(1);
      break;
    }
  L.ok = true;
  L.path = path;
  LA = L;
  return true;
}

void load() {
  if (const char *env = std::getenv("SLEPC_AMD_LAPACK"))
    if (try_lib(env)) return;
  glob_t g;
  if (glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so", 0, nullptr, &g) == 0) {
    for (size_t i = 0; i < g.gl_pathc; i++)
      if (try_lib(g.gl_pathv[i])) { globfree(&g); return; }
    globfree(&g);
  }
  const char *c[] = {"/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", "libopenblas.so.0",
                     "liblapack.so.3", nullptr};
  for (int i = 0; c[i]; i++)
    if (try_lib(c[i])) return;
}
}  // namespace

bool lapack_ready(std::string *path) {
  std::call_once(la_once, load);
  if (path) *path = LA.ok ? LA.path : std::string("not found");
  return LA.ok;
}

int la_potrf_upper(int n, double *a, int lda) {
  if (!lapack_ready()) return -1000;
  int info = 0;
  LA.potrf("U", &n, a, &lda, &info, 1);
  return info;
}

int la_trtri_upper(int n, double *a, int lda) {
  if (!lapack_ready()) return -1000;
  int info = 0;
  LA.trtri("U", "N", &n, a, &lda, &info, 1, 1);
  return info;
}

int la_stevr(int n, double *d, double *e, double *w, double *z) {
  if (!lapack_ready() || !LA.stevr) return -100;
  if (n <= 0) return 0;
  const int lwork = 20 * n, liwork = 10 * n;
  std::vector<double> work(lwork);
  std::vector<int> iwork(liwork), isuppz(2 * n);
  const double vl = 0.0, vu = 0.0, abstol = 0.0;
  const int il = 0, iu = 0;
  int m = 0, info = 0;
  LA.stevr("V", "A", &n, d, e, &vl, &vu, &il, &iu, &abstol, &m, w, z, &n, isuppz.data(), work.data(), &lwork,
           iwork.data(), &liwork, &info, 1, 1);
  return info;
}

int la_getrf(int n, double *a, int lda, int *ipiv) {
  if (!lapack_ready()) return -1000;
  int info = 0;
  LA.getrf(&n, &n, a, &lda, ipiv, &info);
  return info;
}

int la_getrs(char trans, int n, const double *a, int lda, const int *ipiv, double *b) {
  if (!lapack_ready()) return -1000;
  int info = 0, one = 1;
  const char t[2] = {trans, 0};
  LA.getrs(t, &n, &one, a, &lda, ipiv, b, &lda, &info, 1);
  return info;
}

int la_syev_lower(int n, double *a, int lda, double *w) {
  if (!lapack_ready()) return -1000;
  int info = 0, lwork = -1;
  double q = 0.0;
  LA.syev("V", "L", &n, a, &lda, w, &q, &lwork, &info, 1, 1);
  if (info) return info;
  lwork = std::max(1, (int)q);
  std::vector<double> work(lwork);
  LA.syev("V", "L", &n, a, &lda, w, work.data(), &lwork, &info, 1, 1);
  return info;
}

thread_local ScUser *t_sc_user = nullptr;

int sc_compare(int which, double ar, double ai, double br, double bi, double target) {
  double a, b;
  switch (which) {
    case 11: {   // SlepcSCCompare with the user's comparison (slepcsc.c:40-63)
      if (!t_sc_user || !t_sc_user->fn) return 0;
      int r = 0;
      if (t_sc_user->fn(ar, ai, br, bi, &r, t_sc_user->ctx)) t_sc_user->err = 1;
      return r > 0 ? 1 : (r < 0 ? -1 : 0);
    }
    case 1: a = std::hypot(ar, ai); b = std::hypot(br, bi); return a < b ? 1 : (a > b ? -1 : 0);
    case 2: a = std::hypot(ar, ai); b = std::hypot(br, bi); return a > b ? 1 : (a < b ? -1 : 0);
    case 3: return ar < br ? 1 : (ar > br ? -1 : 0);
    case 4: return ar > br ? 1 : (ar < br ? -1 : 0);
    case 5: a = std::fabs(ai); b = std::fabs(bi); return a < b ? 1 : (a > b ? -1 : 0);   // slepcsc.c:204-220
    case 6: a = std::fabs(ai); b = std::fabs(bi); return a > b ? 1 : (a < b ? -1 : 0);   // :222-238
    case 7: a = std::hypot(ar - target, ai); b = std::hypot(br - target, bi);            // :240-253
      return a > b ? 1 : (a < b ? -1 : 0);
    case 8: a = std::fabs(ar - target); b = std::fabs(br - target); return a > b ? 1 : (a < b ? -1 : 0);   // :255-267
  }
  return 0;
}

// SlepcSCCompare (slepcsc.c:40-63) with STBackTransform_Sinvert
// (sinvert.c:16-38) as the map when shift-and-invert is on.
int DenseProblem::cmp(double ar, double ai, double br, double bi) const {
  if (sinv_) {
    auto map = [&](double &re, double &im) {
      if (im == 0.0) {
        re = 1.0 / re + sigma_;
      } else {
        const double t = re * re + im * im;
        re = re / t + sigma_;
        im = -im / t;
      }
    };
    map(ar, ai);
    map(br, bi);
  }
  return sc_compare(which_, ar, ai, br, bi, target_);
}

// DSTranslateHarmonic_NHEP (dsnhep.c:466-537).  Forward (recover = false):
// g = (A - tau I)^-T beta e_n and A(:, n-1) += beta g.  Recover: with
// ncol = l + k, gh = -Q(:, 0:ncol)^T g, A(i, j) += beta gh_i Q(n-1, j) for
// i < ncol, l <= j < ncol, g += Q(:, 0:ncol) gh, and the extra row of columns
// l..ncol-1 scaled by gamma.  gamma = |(1, ||g||)| in both directions.
int DenseProblem::translate_harmonic(double tau, double beta, bool recover, double *g, double *gamma) {
  const int ld = ld_, n = n_;
  double *A = A_.data();
  if (!recover) {
    std::vector<double> W((size_t)ld * n);
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) W[i + (size_t)j * ld] = A[i + (size_t)j * ld];
    for (int i = 0; i < n; i++) g[i] = 0.0;
    g[n - 1] = beta;
    for (int i = 0; i < n; i++) W[i + (size_t)i * ld] -= tau;
    int nn = n, info = 0, one = 1;
    LA.getrf(&nn, &nn, W.data(), &ld, iwork_.data(), &info);
    if (info) return info;
    LA.getrs("C", &nn, &one, W.data(), &ld, iwork_.data(), g, &ld, &info, 1);
    if (info) return info;
    for (int i = 0; i < n; i++) A[i + (size_t)(n - 1) * ld] += g[i] * beta;
  } else {
    const int ncol = l_ + k_;
    const double *Q = Q_.data();
    std::vector<double> gh(ncol);
    for (int j = 0; j < ncol; j++) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += Q[i + (size_t)j * ld] * g[i];
      gh[j] = -s;
    }
    for (int i = 0; i < ncol; i++)
      for (int j = l_; j < ncol; j++) A[i + (size_t)j * ld] += gh[i] * Q[n - 1 + (size_t)j * ld] * beta;
    for (int i = 0; i < n; i++) {
      double s = g[i];
      for (int j = 0; j < ncol; j++) s += Q[i + (size_t)j * ld] * gh[j];
      g[i] = s;
    }
  }
  double nrm = 0.0;
  for (int i = 0; i < n; i++) nrm = std::hypot(nrm, g[i]);
  *gamma = std::hypot(1.0, nrm);
  if (recover)   // the NHEP problem of Krylov-Schur always has the extra row
    for (int j = l_; j < l_ + k_; j++) A[n + (size_t)j * ld] *= *gamma;
  return 0;
}

void DenseProblem::allocate(int ld, bool hermitian, int which) {
  ld_ = ld;
  herm_ = hermitian;
  which_ = which;
  A_.assign((size_t)ld * ld, 0.0);
  Q_.assign((size_t)ld * ld, 0.0);
  X_.assign((size_t)ld * ld, 0.0);
  T_.assign((size_t)3 * ld, 0.0);
  work_.assign((size_t)8 * ld * ld + 16 * ld, 0.0);
  perm_.assign(ld, 0);
  iwork_.assign(ld, 0);
  n_ = l_ = k_ = t_ = 0;
  state_ = DS_RAW;
}

// DSArrowTridiag (dshep.c:221-265): Givens reduction of the arrowhead block.
static void arrow_tridiag(int n, double *d, double *e, double *Q, int ld) {
  if (n <= 2) return;
  const int one = 1;
  double c, s, p, off, temp;
  for (int j = 0; j < n - 2; j++) {
    temp = e[j + 1];
    LA.lartg(&temp, &e[j], &c, &s, &e[j + 1]);
    s = -s;
    temp = d[j + 1];
    e[j] = c * s * (temp - d[j]);
    d[j + 1] = s * s * d[j] + c * c * temp;
    d[j] = c * c * d[j] + s * s * temp;
    const int j2 = j + 2;
    LA.rot(&j2, Q + (size_t)j * ld, &one, Q + (size_t)(j + 1) * ld, &one, &c, &s);
    for (int i = j - 1; i >= 0; i--) {
      off = -s * e[i];
      e[i] = c * e[i];
      temp = e[i + 1];
      LA.lartg(&temp, &off, &c, &s, &e[i + 1]);
      s = -s;
      temp = (d[i] - d[i + 1]) * s - 2.0 * c * e[i];
      p = s * temp;
      d[i + 1] += p;
      d[i] -= p;
      e[i] = -e[i] - c * temp;
      LA.rot(&j2, Q + (size_t)i * ld, &one, Q + (size_t)(i + 1) * ld, &one, &c, &s);
    }
  }
}

int DenseProblem::solve(double *wr, double *wi) {
  if (!lapack_ready()) return -100;
  const int ld = ld_, n = n_, l = l_;
  int info = 0;
  if (herm_) {
    // DSSolve_HEP_QR + DSIntermediate_HEP, compact storage (dshep.c:267-321,383-426)
    double *d = T_.data(), *e = T_.data() + ld;
    std::memset(Q_.data() + (size_t)l * ld, 0, sizeof(double) * (size_t)ld * (n - l));
    for (int i = l; i < n; i++) Q_[i + (size_t)i * ld] = 1.0;
    const int na = std::max(0, k_ - l + 1), off = l + l * ld, n1 = n - l;
    if (state_ < DS_INTERMEDIATE) arrow_tridiag(na, d + l, e + l, Q_.data() + off, ld);
    for (int i = 0; i < l; i++) wr[i] = d[i];
    if (n1 > 0) LA.steqr("V", &n1, d + l, e + l, Q_.data() + off, &ld, work_.data(), &info, 1);
    for (int i = l; i < n; i++) wr[i] = d[i];
    for (int i = 0; i < n - 1; i++) e[i] = 0.0;
    if (wi) for (int i = l; i < n; i++) wi[i] = 0.0;
    state_ = DS_CONDENSED;
    return info;
  }
  // DSSolve_NHEP_Private (dsutil.c:21-90)
  double *A = A_.data(), *Q = Q_.data(), *tau = work_.data(), *work = work_.data() + ld;
  int ilo = l + 1, lwork = 6 * ld, nn = n;
  std::memset(Q, 0, sizeof(double) * (size_t)ld * ld);
  for (int i = 0; i < n; i++) Q[i + (size_t)i * ld] = 1.0;
  if (n == 1) {
    wr[0] = A[0];
    wi[0] = 0.0;
    state_ = DS_CONDENSED;
    return 0;
  }
  if (state_ < DS_INTERMEDIATE) {
    LA.gehrd(&nn, &ilo, &nn, A, &ld, tau, work, &lwork, &info);
    if (info) return info;
    for (int j = 0; j < n - 1; j++)
      for (int i = j + 2; i < n; i++) {
        Q[i + (size_t)j * ld] = A[i + (size_t)j * ld];
        A[i + (size_t)j * ld] = 0.0;
      }
    LA.orghr(&nn, &ilo, &nn, Q, &ld, tau, work, &lwork, &info);
    if (info) return info;
  }
  LA.hseqr("S", "V", &nn, &ilo, &nn, A, &ld, wr, wi, Q, &ld, work, &lwork, &info, 1, 1);
  for (int j = 0; j < l; j++) {
    if (j == n - 1 || A[j + 1 + (size_t)j * ld] == 0.0) {
      wr[j] = A[j + (size_t)j * ld];
      wi[j] = 0.0;
    } else {
      wr[j] = wr[j + 1] = A[j + (size_t)j * ld];
      wi[j] = std::sqrt(std::fabs(A[j + 1 + (size_t)j * ld])) * std::sqrt(std::fabs(A[j + (size_t)(j + 1) * ld]));
      wi[j + 1] = -wi[j];
      j++;
    }
  }
  state_ = DS_CONDENSED;
  return info;
}

int DenseProblem::sort(double *wr, double *wi) {
  const int ld = ld_, n = n_, l = l_;
  if (herm_) {
    // DSSort_HEP (dshep.c:323-347): insertion sort of the permutation, then permute Q
    double *d = T_.data();
    int *perm = perm_.data();
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = l + 1; i < t_; i++) {
      const double re = d[perm[i]];
      int j = i - 1;
      int res = cmp(re, 0.0, d[perm[j]], 0.0);
      while (res < 0 && j >= l) {
        std::swap(perm[j], perm[j + 1]);
        j--;
        if (j >= l) res = cmp(re, 0.0, d[perm[j]], 0.0);
      }
    }
    for (int i = l; i < n; i++) wr[i] = d[perm[i]];
    for (int i = l; i < n; i++) {
      const int p = perm[i];
      if (p != i) {
        int j = i + 1;
        while (perm[j] != i) j++;
        perm[j] = p;
        perm[i] = i;
        for (int r = 0; r < n; r++) std::swap(Q_[r + (size_t)p * ld], Q_[r + (size_t)i * ld]);
      }
    }
    for (int i = l; i < n; i++) d[i] = wr[i];
    return 0;
  }
  // DSSort_NHEP_Total (dsutil.c:93-166): selection sort by block swaps
  double *T = A_.data();
  int nn = n, info = 0;
  for (int i = l; i < n - 1; i++) {
    double re = wr[i], im = wi[i];
    int pos = 0, j = (im != 0) ? i + 2 : i + 1;
    for (; j < n; j++) {
      if (cmp(re, im, wr[j], wi[j]) > 0) {
        re = wr[j];
        im = wi[j];
        pos = j;
      }
      if (wi[j] != 0) j++;
    }
    if (pos) {
      int ifst = pos + 1, ilst = i + 1;
      LA.trexc("V", &nn, T, &ld, Q_.data(), &ld, &ifst, &ilst, work_.data(), &info, 1);
      if (info) return info;
      for (j = i; j < n; j++) {
        wr[j] = T[j + (size_t)j * ld];
        if (j < n - 1 && T[j + 1 + (size_t)j * ld] != 0.0) {
          wi[j] = std::sqrt(std::fabs(T[j + 1 + (size_t)j * ld])) * std::sqrt(std::fabs(T[j + (size_t)(j + 1) * ld]));
          wr[j + 1] = wr[j];
          wi[j + 1] = -wi[j];
          j++;
        } else {
          wi[j] = 0.0;
        }
      }
    }
    if (wi[i] != 0) i++;
  }
  return 0;
}

void DenseProblem::update_extra_row() {
  const int ld = ld_, n = n_;
  if (herm_) {  // DSUpdateExtraRow_HEP compact (dshep.c:349-381)
    double *e = T_.data() + ld;
    const double beta = e[n - 1];
    for (int i = 0; i < n; i++) e[i] = beta * Q_[n - 1 + (size_t)i * ld];
  } else {      // DSUpdateExtraRow_NHEP (dsnhep.c:318-340): b <- Q^T b
    double *x = work_.data(), *y = work_.data() + ld;
    for (int i = 0; i < n; i++) x[i] = A_[n + (size_t)i * ld];
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int r = 0; r < n; r++) s += Q_[r + (size_t)i * ld] * x[r];
      y[i] = s;
    }
    for (int i = 0; i < n; i++) A_[n + (size_t)i * ld] = y[i];
  }
  k_ = n;
}

int DenseProblem::ritz_residual(int i, double *rnorm, int *cplx) {
  const int ld = ld_, n = n_;
  *cplx = 0;
  if (herm_) {  // DSVectors_HEP (dshep.c:137-172)
    *rnorm = std::fabs(Q_[n - 1 + (size_t)i * ld]);
    return 0;
  }
  // DSVectors_NHEP_Eigen_Some (dsnhep.c:101-166)
  int *select = iwork_.data();
  for (int q = 0; q < n; q++) select[q] = 0;
  double *Y = X_.data() + (size_t)i * ld;
  select[i] = 1;
  const bool iscomplex = (i < n - 1 && A_[i + 1 + (size_t)i * ld] != 0.0);
  int mm = iscomplex ? 2 : 1, mout = 0, info = 0, nn = n;
  if (iscomplex) select[i + 1] = 1;
  LA.trevc("R", "S", select, &nn, A_.data(), &ld, Y, &ld, Y, &ld, &mm, &mout, work_.data(), &info, 1, 1);
  if (info) return info;
  double *tmp = work_.data();
  std::memcpy(tmp, Y, sizeof(double) * (size_t)mout * ld);
  for (int c = 0; c < mout; c++)
    for (int r = 0; r < n; r++) {
      double s = 0.0;
      for (int q = 0; q < n; q++) s += Q_[r + (size_t)q * ld] * tmp[q + (size_t)c * ld];
      Y[r + (size_t)c * ld] = s;
    }
  double nr = 0.0;
  for (int r = 0; r < n; r++) nr += Y[r] * Y[r];
  double norm = std::sqrt(nr);
  if (iscomplex) {
    double ni = 0.0;
    for (int r = 0; r < n; r++) ni += Y[r + ld] * Y[r + ld];
    norm = std::hypot(norm, std::sqrt(ni));
  }
  for (int c = 0; c < mout; c++)
    for (int r = 0; r < n; r++) Y[r + (size_t)c * ld] /= norm;
  *rnorm = iscomplex ? std::hypot(Y[n - 1], Y[n - 1 + ld]) : std::fabs(Y[n - 1]);
  *cplx = iscomplex ? 1 : 0;
  return 0;
}

int DenseProblem::truncate_size(int k, int nv, int l) const {
  // DSGetTruncateSize_Default (dsops.c:329-344): do not split a 2x2 block
  if (herm_) return l;
  if (A_[(k + l) + (size_t)(k + l - 1) * ld_] != 0.0) {
    if (k + l < nv - 1) l++;
    else l--;
  }
  return l;
}

void DenseProblem::truncate(int n) {
  // DSTruncate_HEP / DSTruncate_NHEP with trim = FALSE (dshep.c:643-670, dsnhep.c:394-425)
  if (!herm_ && k_ == n_) {
    for (int i = l_; i < n; i++) A_[n + (size_t)i * ld_] = A_[n_ + (size_t)i * ld_];
    for (int i = l_; i < n_; i++) A_[n_ + (size_t)i * ld_] = 0.0;
  }
  k_ = n;
  t_ = n_;
  n_ = n;
  state_ = DS_TRUNCATED;
}

int DenseProblem::trim_and_vectors(int nconv) {
  // DSTruncate(nconv,TRUE) leaves the DS RAW, so DSVectors_NHEP_Eigen_All
  // (dsnhep.c:168-226) returns the eigenvectors of T itself.
  if (!herm_)
    for (int i = l_; i < n_; i++) A_[n_ + (size_t)i * ld_] = 0.0;
  n_ = nconv;
  l_ = k_ = 0;
  t_ = nconv;
  state_ = DS_RAW;
  if (herm_ || nconv == 0) return 0;
  const int ld = ld_, n = nconv;
  int mout = 0, info = 0, nn = n;
  LA.trevc("R", "A", nullptr, &nn, A_.data(), &ld, nullptr, &ld, X_.data(), &ld, &nn, &mout, work_.data(), &info, 1, 1);
  if (info) return info;
  for (int i = 0; i < n; i++) {
    const bool cp = (i < n - 1 && A_[i + 1 + (size_t)i * ld] != 0.0);
    double s = 0.0;
    for (int r = 0; r < n; r++) s += X_[r + (size_t)i * ld] * X_[r + (size_t)i * ld];
    double norm = std::sqrt(s);
    if (cp) {
      double s2 = 0.0;
      for (int r = 0; r < n; r++) s2 += X_[r + (size_t)(i + 1) * ld] * X_[r + (size_t)(i + 1) * ld];
      norm = std::hypot(norm, std::sqrt(s2));
    }
    for (int c = 0; c < (cp ? 2 : 1); c++)
      for (int r = 0; r < n; r++) X_[r + (size_t)(i + c) * ld] /= norm;
    if (cp) i++;
  }
  return 0;
}


// ---- DSNHEPTS (dsnhepts.c) -------------------------------------------------
void DenseProblem::enable_two_sided() {
  B_.assign((size_t)ld_ * ld_, 0.0);
  Z_.assign((size_t)ld_ * ld_, 0.0);
  Y_.assign((size_t)ld_ * ld_, 0.0);
  wr2_.assign(ld_, 0.0);
  wi2_.assign(ld_, 0.0);
}

void DenseProblem::swap_sides() {
  std::swap(A_, B_);
  std::swap(Q_, Z_);
  std::swap(X_, Y_);
}

// DSSolve_NHEPTS (dsnhepts.c:275-287): both matrices from the same state
int DenseProblem::solve_ts(double *wr, double *wi) {
  const int st = state_;
  int info = solve(wr, wi);
  if (info) return info;
  swap_sides();
  state_ = st;
  info = solve(wr2_.data(), wi2_.data());
  swap_sides();
  return info;
}

// DSSortWithPermutation_NHEP_Private (dsutil.c:177-226) on (A, Q)
int DenseProblem::sort_with_permutation(int *perm, double *wr, double *wi) {
  const int ld = ld_, n = n_;
  double *T = A_.data();
  int nn = n, info = 0;
  for (int i = l_; i < n - 1; i++) {
    const int pos = perm[i];
    const int inc = (pos < n - 1 && T[pos + 1 + (size_t)pos * ld] != 0.0) ? 2 : 1;
    if (pos != i) {
      if (!((T[pos + (size_t)(pos - 1) * ld] == 0.0 || perm[i + 1] == pos - 1) &&
            (T[pos + 1 + (size_t)pos * ld] == 0.0 || perm[i + 1] == pos + 1)))
        return -200;   // a 2x2 block would be split
      int ifst = pos + 1, ilst = i + 1;
      LA.trexc("V", &nn, T, &ld, Q_.data(), &ld, &ifst, &ilst, work_.data(), &info, 1);
      if (info) return info;
      for (int j = i + 1; j < n; j++)
        if (perm[j] >= i && perm[j] < pos) perm[j] += inc;
      perm[i] = i;
      if (inc == 2) perm[i + 1] = i + 1;
    }
    if (inc == 2) i++;
  }
  for (int j = l_; j < n; j++) {
    wr[j] = T[j + (size_t)j * ld];
    if (j < n - 1 && T[j + 1 + (size_t)j * ld] != 0.0) {
      wi[j] = std::sqrt(std::fabs(T[j + 1 + (size_t)j * ld])) * std::sqrt(std::fabs(T[j + (size_t)(j + 1) * ld]));
      wr[j + 1] = wr[j];
      wi[j + 1] = -wi[j];
      j++;
    } else {
      wi[j] = 0.0;
    }
  }
  return 0;
}

// DSSort_NHEPTS (dsnhepts.c:193-242): both Schur forms sorted by the
// criterion, then B's eigenvalues matched to A's where they differ by more
// than sqrt(eps) and B reordered to follow A.
int DenseProblem::sort_ts(double *wr, double *wi) {
  int info = sort(wr, wi);
  if (info) return info;
  swap_sides();
  info = sort(wr2_.data(), wi2_.data());
  if (info) { swap_sides(); return info; }
  const int n = n_;
  std::vector<int> idx(n), idx2(n), p(n);
  int cont = 0;
  const double seps = std::sqrt(2.220446049250313e-16);
  for (int i = 0; i < n; i++) {
    if (std::hypot(wr2_[i] - wr[i], wi2_[i] - wi[i]) > seps) {
      idx2[cont] = i;
      idx[cont++] = i;
    }
    p[i] = -1;
  }
  if (cont) {
    int id = 0;
    for (int i = 0; i < cont; i++) {
      double t = 1.79e308, sd;
      for (int j = 0; j < cont; j++)
        if (idx2[j] != -1 && (sd = std::hypot(wr2_[idx[j]] - wr[idx[i]], wi2_[idx[j]] - wi[idx[i]])) < t) {
          id = j;
          t = sd;
        }
      p[idx[i]] = idx[id];
      idx2[id] = -1;
    }
    for (int i = 0; i < n; i++)
      if (p[i] == -1) p[i] = i;
    info = sort_with_permutation(p.data(), wr2_.data(), wi2_.data());
  }
  swap_sides();
  return info;
}

void DenseProblem::update_extra_row_ts() {   // dsnhepts.c:244-273
  update_extra_row();
  swap_sides();
  update_extra_row();
  swap_sides();
}

int DenseProblem::left_residual(int i, double *rnorm, int *cplx) {   // dsnhepts.c:54-119, left
  swap_sides();
  const int info = ritz_residual(i, rnorm, cplx);
  swap_sides();
  return info;
}

int DenseProblem::truncate_size_ts(int k, int nv, int l) const {   // dsnhepts.c:353-371
  const int ld = ld_;
  if (A_[k + l + (size_t)(k + l - 1) * ld] != 0.0 || B_[k + l + (size_t)(k + l - 1) * ld] != 0.0) {
    if (k + l < nv - 1) l++;
    else l--;
  }
  return l;
}

void DenseProblem::truncate_ts(int n) {   // dsnhepts.c:373-406, trim = FALSE
  if (k_ == n_) {
    for (int i = l_; i < n; i++) {
      A_[n + (size_t)i * ld_] = A_[n_ + (size_t)i * ld_];
      B_[n + (size_t)i * ld_] = B_[n_ + (size_t)i * ld_];
    }
    for (int i = l_; i < n_; i++) {
      A_[n_ + (size_t)i * ld_] = 0.0;
      B_[n_ + (size_t)i * ld_] = 0.0;
    }
  }
  k_ = n;
  t_ = n_;
  n_ = n;
  state_ = DS_TRUNCATED;
}

int DenseProblem::trim_and_vectors_ts(int nconv) {
  const int n = n_, l = l_, k = k_, t = t_;
  int info = trim_and_vectors(nconv);
  if (info) return info;
  n_ = n; l_ = l; k_ = k; t_ = t;
  swap_sides();
  info = trim_and_vectors(nconv);
  swap_sides();
  return info;
}

}  // namespace sa
