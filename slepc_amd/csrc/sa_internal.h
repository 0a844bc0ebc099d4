// Internal declarations shared by the HIP kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/slepc_amd.h"

namespace sa {

// ---- errors ---------------------------------------------------------------
void set_error(const std::string &msg);
// A failed HIP call is reported once: the runtime's last-error state is cleared
// too, or the next launch check (hipGetLastError) would report it again.
#define SA_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::sa::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      (void)hipGetLastError();                                                    \
      return 1;                                                                   \
    }                                                                             \
  } while (0)
#define SA_CHECK(cond, msg)                                                       \
  do {                                                                            \
    if (!(cond)) { ::sa::set_error(msg); return 2; }                              \
  } while (0)
#define SA_TRY(call)                                                              \
  do { int r_ = (call); if (r_) return r_; } while (0)

// Owns an object under construction: destroys it if an error path returns
// before release() hands it to the caller.
template <typename H, int (*Destroy)(H)>
struct BuildGuard {
  H h;
  ~BuildGuard() { if (h) Destroy(h); }
  H release() { H r = h; h = nullptr; return r; }
};

hipStream_t stream();
int ensure_init();
int num_cus();
int lds_max_bytes();   // LDS a workgroup may allocate (device attribute)

// ---- Gram-Schmidt step machine ---------------------------------------------
// Column c is orthogonalised against columns 0..c-1 by three kernels (slots
// 1..3).  Slot s reads the state st[(s+1)%3] left by the previous kernel and
// workgroup 0 writes st[s-1]; the state after a column is st[2].
// BVSetOrthogonalization refinement types (BVOrthogRefineType)
enum GsRefine : int { GS_REFINE_IFNEEDED = 0, GS_REFINE_NEVER = 1, GS_REFINE_ALWAYS = 2 };
constexpr int GS_MAXC = 128;        // largest column index + 1 handled
constexpr int GS_NSTATE = 3;
constexpr int64_t GS_NT_BYTES = 256ll << 20;   // bases above this size use non-temporal tile loads
constexpr int SPMV_PT = 8;          // nonzeros per thread of the stream SpMV
constexpr int SPMV_NZB = 256 * SPMV_PT;   // nonzeros per block it stages in LDS (256-thread workgroups)

struct GsState {
  int pass;         // Gram-Schmidt passes done for the current column
  int pending;      // estimate was <= 0: next decision evaluates the true norm
  int lindep;       // final linear-dependence flag of the current column
  int stop;         // breakdown happened: later steps of the sweep are no-ops
  int stop_col;
  int done;         // current column finished (normalised)
  int dots_ready;   // the kernel that wrote this state produced dot products
  int slow_total;   // columns finished by the one-workgroup path (third pass)
  int64_t passes_total;
  double onrm, nrm, beta;
};

struct Workspace {           // device scratch for partial sums of one BV
  double *part = nullptr;    // 2 x [GS_MAXC+1][maxG], column-major partial sums
  double *dots = nullptr;    // 2 x [GS_MAXC+1] reduced dot products
  int maxG = 0;
};

struct BV {
  int64_t n = 0, ld = 0;
  int m = 0, prec = SA_FP64;
  int l = 0, k = 0;
  void *data = nullptr;        // T[ld * m]
  double *dbuf = nullptr;      // coefficient buffer m x m (FP64) on device
  GsState *st = nullptr;       // device, GS_NSTATE states (ping-pong)
  GsState *st_host = nullptr;  // pinned host mirror of st[2]
  Workspace ws;
  double *dq = nullptr;        // device staging for Q / q vectors (m*m doubles)
  double *hq = nullptr;        // pinned host staging
  int G = 0, rows_per_block = 0;
  int64_t passes_done = 0;     // host-side counter of orthogonalisation passes
  int refine = 0;              // BVSetOrthogonalization refinement (0 if needed, 1 never, 2 always)
};

struct Mat {
  int n = 0;
  int64_t nnz = 0;
  int prec = SA_FP64;
  int32_t *rowptr = nullptr, *col = nullptr;
  void *val = nullptr;
  int lanes = 8;               // lanes per row of the vector-CSR kernels
  int spmv_rb = 0;             // rows per workgroup of the stream SpMV (0: vector-CSR kernel);
                               // set for operators larger than the caches
};

// Rows per workgroup of the stream SpMV for the rows [0, n) of rp (rp[0] need
// not be 0): the largest RB of {256, 128, 64, 32} whose every RB-row block
// holds at most SPMV_NZB nonzeros and which gives at least SPMV_MIN_BLOCKS
// workgroups (else the smallest RB that fits); 0 when none fits.
constexpr int64_t SPMV_MIN_BLOCKS = 1024;
int spmv_pick_rows(const int32_t *rp, int64_t n);

// ---- kernel launchers (sa_kernels.hip) -------------------------------------
int launch_spmv(const Mat &A, const void *x, void *y, hipStream_t s);
// MatNorm(A, NORM_INFINITY): the largest absolute row sum, computed on the device.
int mat_norm_inf(const Mat &A, double *out);
// One Arnoldi/orthonormalisation step for column c of V; spmv_from >= 0 first
// computes V(:,c) = A V(:,spmv_from).  eta: DGKS threshold.
int launch_gs_column(BV &V, const Mat *A, int c, double eta, hipStream_t s);
// Clears the Gram-Schmidt state machine of V before a sweep.
int launch_gs_reset(BV &V, hipStream_t s);
int gs_min_chunks(bool big);   // fewest 64-row chunks per workgroup of the GS kernels (big: above GS_NT_BYTES)
int concurrent_solves();      // the application's hint (sa_set_concurrent_solves)
int gs_max_blocks(bool big);  // most workgroups of the GS kernels per basis (big: above GS_NT_BYTES)
int launch_dotvec(BV &V, int l, int k, const void *x, double *q_dev, hipStream_t s);
int launch_multvec(BV &V, int l, int k, double alpha, const double *q_dev, double beta, void *y,
                   hipStream_t s);
int launch_multinplace(BV &V, int l, int k, int s_, int e, const double *Q_dev, int ldq,
                       hipStream_t s);
int launch_scale(int prec, void *x, int64_t n, double alpha, hipStream_t s);
int launch_dot_self(int prec, const void *x, int64_t n, double *out_dev, double *part, hipStream_t s);
int launch_convert(int src_prec, const void *src, int64_t src_ld, int dst_prec, void *dst,
                   int64_t dst_ld, int64_t n, int ncols, hipStream_t s);
int launch_axpy_norm(int prec, const void *Ax, const void *x, double lambda, int64_t n,
                     double *part, double *out_dev, hipStream_t s);
int launch_fill_random(int prec, void *x, int64_t n, uint64_t seed, hipStream_t s);
// BVDot: M_dev(i + j*(ky-ly)) = Y(:,ly+i) . X(:,lx+j) (sa_block.hip); synchronises s
int launch_bvdot(const BV &X, int lx, int kx, const BV &Y, int ly, int ky, double *M_dev, hipStream_t s);
// BVMult: Y(:, ly:ky) = beta Y + alpha X(:, lx:kx) Q (Q_dev: (kx-lx) x (ky-ly),
// leading dimension ldq) or, with Q_dev null, beta Y + alpha X (sa_block.hip)
int launch_bvmult(const BV &X, int lx, int kx, BV &Y, int ly, int ky, const double *Q_dev, int ldq, double alpha,
                  double beta, hipStream_t s);
// BVMatMult: Y(:, ly+j) = A X(:, lx+j), j < ncols (sa_block.hip)
int launch_spmm(const Mat &A, const BV &X, int lx, const BV &Y, int ly, int ncols, hipStream_t s);
// Residual norms of approximate eigenpairs with the SpMV fused in
// (sa_resid.hip): vector v of a launch is ||A X(:,xcol) - a X(:,xcol) + b X(:,zcol)||^2
// (zcol < 0: no z term), X in the operator's precision with leading
// dimension ldx.  part holds RESID_MAXVEC * RESID_MAXG doubles, out
// RESID_MAXVEC; out[v] receives the squared norm of vector v.
constexpr int RESID_MAXVEC = 16;
constexpr int RESID_MAXG = 1024;
struct ResidArgs {
  int nvec;
  int xcol[RESID_MAXVEC], zcol[RESID_MAXVEC];
  double a[RESID_MAXVEC], b[RESID_MAXVEC];
};
int resid_grid(int64_t n, int lanes);
int launch_resid_norms(const Mat &A, const void *X, int64_t ldx, const ResidArgs &args, double *part, double *out,
                       hipStream_t s);
// ||A x - lambda x|| (absolute) of k columns of X with eigenvalues lr + i li
// by the rules of EPSComputeResidualNorm_Private (epssolve.c:666-718): when
// li[c] != 0 (beyond |lr[c]| eps) columns c, c+1 hold the real and imaginary
// parts of a pair and res[c] = res[c+1] = hypot of the two norms.  Blocks on
// the stream; part/out as for launch_resid_norms.
int residual_norms(const Mat &A, const void *X, int64_t ldx, int k, const double *lr, const double *li,
                   double *res, double *part, double *out, hipStream_t s);
// Device address of the residual scratch array on a device (sa_resid.hip):
// set up once per device by sa_init with that device current; null before.
int noalloc_scratch_setup(int device);
double *noalloc_scratch(int device);
// Shift-and-invert solves x = (A - sigma I)^-1 b (BiCGStab, FP64; optional
// Jacobi preconditioning), workspace of 9 vectors + partials in one buffer.
// With direct, a banded LU of A - sigma I (KSPPREONLY + PCLU) factored on the
// device at set-up, and each solve is its two substitutions.
struct KspWork {
  double *buf = nullptr;
  int64_t n = 0;
  int G = 1;
  double sigma = 0.0;
  bool jacobi = true;
  bool direct = false;
  bool gmres = false;          // GMRES(restart) instead of BiCGStab for the iterative solves
  int restart = 30;
  BV gv;                       // GMRES Arnoldi basis, restart + 1 columns
  double *ab = nullptr;        // band storage, ldab x n
  int *ipiv = nullptr;         // n pivots + info
  int64_t ab_n = 0, ldab = 0;
  int kl = 0, ku = 0;
};
int ksp_alloc(KspWork &w, const Mat &A, double sigma, bool jacobi, bool direct, bool gmres = false, int restart = 30);
void ksp_free(KspWork &w);
// rtol relative to ||b||; error when max_it is reached or on breakdown
int ksp_solve(const Mat &A, KspWork &w, const double *b_dev, double *x_dev, double rtol, int max_it, int *its);

int bv_sweep_sinvert(BV &V, const Mat &A, KspWork &w, double rtol, int max_it, int k, int *m, double eta,
                     double *beta, int *breakdown, int64_t *ksp_its);
// x -= C(:,0:nc) (C(:,0:nc)^T x), `passes` times; C has orthonormal columns.
int bv_project_out(BV &C, int nc, void *x, hipStream_t s, int passes = 2);
// The Arnoldi sweep with a deflation space C (nc columns): the operator
// (SpMV, or the shift-and-invert solve when w is given), C projected out,
// then the Gram-Schmidt slots without SpMV.
int bv_sweep_deflated(BV &V, const Mat &A, KspWork *w, double rtol, int max_it, BV &C, int nc, int k, int *m,
                      double eta, double *beta, int *breakdown, int64_t *ksp_its);

// TSQR of V(:, l:k) (k - l <= 64, V.n >= k - l): R_host ((k-l) x (k-l)) gets
// the R of the tile tree; with formq, V(:, l:k) is overwritten by Q times the
// signs f (k - l entries) that make Q and diag(f) R equal geqrf/orgqr's.
int launch_tsqr(BV &V, int l, int k, bool formq, double *R_host, double *f, hipStream_t s);
// The signs f that turn a QR factorization (Q, R), k x k block Qtop = the
// first k rows of Q, into the one flat Householder QR (geqrf) computes.
void tsqr_signs(const double *Qtop, const double *R, int64_t n, int k, double *f);
// Rows 0..k-l-1 of V(:, l:k) into out ((k-l) x (k-l), column-major, FP64).
int bv_top_block(const BV &V, int l, int k, double *out);

// ---- generators (sa_problems.cpp) ------------------------------------------
uint64_t splitmix64(uint64_t z);
double u01(uint64_t seed, uint64_t i);

// ---- BV helpers (sa_bv.cpp) ---------------------------------------------------
int bv_alloc(BV &V, int64_t n, int m, int prec);
void bv_free(BV &V);
int bv_sweep(BV &V, const Mat &A, int k, int *m, double eta, double *beta, int *breakdown);
int bv_orthonormalize(BV &V, int j, double eta, double *norm, int *lindep);
// Modified Gram-Schmidt of column j against the columns cols[0..nc) (each < j,
// in that order), refinement IFNEEDED (sa_mgs.hip): coefficients in the
// coefficient buffer column j, row j = the norm (0 on lindep); not scaled.
int bv_mgs_column(BV &V, int j, const int *cols, int nc, double eta, double *norm, int *lindep);
// Balancing (sa_balance.hip): D (device, A.n entries) from EPSBuildBalance_Krylov
// (oneside, or twoside with AT = A^T); B's values scaled to D B D^-1 in place;
// eigenvectors in V(:, 0:nconv) divided by D and normalised (pairs by eigi).
int eps_build_balance(const Mat &A, const Mat *AT, bool twoside, int its, double cutoff, double *D);
int balance_scale_mat(Mat &B, const double *D);
int balance_unscale_vectors(BV &V, const double *D, int nconv, const double *eigi);
// BVOrthogonalizeVec with MGS: the device vector w (W.n rows, W's precision)
// against the columns 0..nc-1 of W, refinement IFNEEDED; not scaled.
int bv_mgs_vec(BV &W, int nc, void *w, double eta, double *norm, int *lindep);
// BVOrthogonalize of the active columns (method 0 gs, 1 chol, 2 svqb); R: host
// column-major, leading dimension ldr >= V.k, or null
int bv_orthogonalize(BV &V, int method, double *R, int ldr);
inline size_t prec_size(int prec) { return prec == SA_FP64 ? 8 : 4; }
inline char *col_ptr(const BV &V, int j) {
  return static_cast<char *>(V.data) + (size_t)j * V.ld * prec_size(V.prec);
}

}  // namespace sa
