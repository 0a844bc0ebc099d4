// Projected dense eigenproblem (the DS object of the reference) on the host in
// FP64 with LAPACK: DSHEP (compact arrowhead/tridiagonal, steqr) and DSNHEP
// (Hessenberg, hseqr/trexc/trevc).
#pragma once
#include <string>
#include <vector>

namespace sa {

bool lapack_ready(std::string *path = nullptr);

// LAPACK routines of the block orthogonalisation (bvlapack.c:136-330); each
// returns LAPACK's info (-1000 when no LAPACK was found).
int la_potrf_upper(int n, double *a, int lda);          // dpotrf("U")
int la_trtri_upper(int n, double *a, int lda);          // dtrtri("U", "N")
int la_syev_lower(int n, double *a, int lda, double *w);  // dsyev("V", "L"), ascending w
int la_getrf(int n, double *a, int lda, int *ipiv);       // dgetrf
// dstevr("V", "A"): eigenvalues w (ascending) and eigenvectors z (n x n,
// leading dimension n) of the tridiagonal (d, e); d, e are overwritten.
int la_stevr(int n, double *d, double *e, double *w, double *z);
int la_getrs(char trans, int n, const double *a, int lda, const int *ipiv, double *b);   // dgetrs, one rhs

enum DsState { DS_RAW = 0, DS_INTERMEDIATE = 1, DS_CONDENSED = 2, DS_TRUNCATED = 3 };

// SlepcCompare* (src/sys/slepcsc.c:152-267): >0 when (ar,ai) must follow (br,bi).
// which as EPSWhich: 1 LM, 2 SM, 3 LR, 4 SR, 5 LI, 6 SI, 7 target magnitude,
// 8 target real (distance to target).
// EPSSetEigenvalueComparison (EPS_WHICH_USER, which = 11): the user comparator
// of the solve running on this thread; a nonzero return of fn sets err.
struct ScUser {
  int (*fn)(double ar, double ai, double br, double bi, int *res, void *ctx) = nullptr;
  void *ctx = nullptr;
  int err = 0;
};
extern thread_local ScUser *t_sc_user;
int sc_compare(int which, double ar, double ai, double br, double bi, double target = 0.0);

class DenseProblem {
 public:
  void allocate(int ld, bool hermitian, int which);
  // DSSetDimensions (dsops.c:130-165)
  void set_dimensions(int n, int l, int k) { n_ = n; l_ = l; k_ = k; t_ = n; }
  void set_state(int s) { state_ = s; }
  void set_target(double t) { target_ = t; }      // comparison context of which 7/8
  // The DS sorting criterion compares back-transformed values (SlepcMap_ST,
  // epssetup.c:237): with shift-and-invert, theta -> sigma + 1/theta.
  void set_sinvert(bool on, double sigma) { sinv_ = on; sigma_ = sigma; }
  // DSTranslateHarmonic (NHEP): g has n entries; returns LAPACK's info
  int translate_harmonic(double tau, double beta, bool recover, double *g, double *gamma);
  int n() const { return n_; }
  int k() const { return k_; }
  int ld() const { return ld_; }
  double *d() { return T_.data(); }          // HEP diagonal
  double *e() { return T_.data() + ld_; }    // HEP off-diagonal / arrow
  double *A() { return A_.data(); }          // NHEP matrix (ld x ld, extra row at n)
  const double *Q() const { return Q_.data(); }
  const double *X() const { return X_.data(); }
  int solve(double *wr, double *wi);         // DSSolve
  int sort(double *wr, double *wi);          // DSSort
  void update_extra_row();                   // DSUpdateExtraRow
  // Residual weight of Ritz pair i: |last row of the eigenvector| (DSVectors with
  // index); sets *cplx when i starts a complex pair.
  int ritz_residual(int i, double *rnorm, int *cplx);
  int truncate_size(int k, int nv, int l) const;   // DSGetTruncateSize_Default
  void truncate(int n);                            // DSTruncate(n,FALSE)
  int trim_and_vectors(int nconv);                 // DSTruncate(n,TRUE) + DSVectors(X) (NHEP)

  // Two-sided problem (DSNHEPTS, dsnhepts.c): a second NHEP matrix B (the
  // projection for the left space) with Schur vectors Z and eigenvectors Y.
  void enable_two_sided();
  double *B() { return B_.data(); }
  const double *Z() const { return Z_.data(); }
  const double *Y() const { return Y_.data(); }
  int solve_ts(double *wr, double *wi);            // DSSolve_NHEPTS
  int sort_ts(double *wr, double *wi);             // DSSort_NHEPTS: B follows A's order
  void update_extra_row_ts();                      // DSUpdateExtraRow_NHEPTS
  int left_residual(int i, double *rnorm, int *cplx);   // DSVectors(Y, i)
  int truncate_size_ts(int k, int nv, int l) const;    // DSGetTruncateSize_NHEPTS
  void truncate_ts(int n);                             // DSTruncate_NHEPTS(n,FALSE)
  int trim_and_vectors_ts(int nconv);                  // DSTruncate(n,TRUE) + X and Y
 private:
  void swap_sides();   // exchange (A, Q, X) with (B, Z, Y)
  int sort_with_permutation(int *perm, double *wr, double *wi);   // on (A, Q)
  int ld_ = 0, n_ = 0, l_ = 0, k_ = 0, t_ = 0, state_ = DS_RAW, which_ = 1;
  double target_ = 0.0;
  bool sinv_ = false;
  double sigma_ = 0.0;
  int cmp(double ar, double ai, double br, double bi) const;   // SlepcSCCompare with the map
  bool herm_ = true;
  std::vector<double> A_, Q_, X_, T_, work_;
  std::vector<double> B_, Z_, Y_, wr2_, wi2_;     // two-sided only
  std::vector<int> perm_, iwork_;
};

}  // namespace sa
