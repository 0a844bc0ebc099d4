// extern "C" entry points of libslepc_amd.so (declared in include/slepc_amd.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "sa_ds.h"
#include "sa_eps.h"
#include "sa_internal.h"

namespace sa {

namespace {
thread_local std::string g_err;
thread_local hipStream_t t_stream = nullptr;   // one stream per host thread
std::atomic<int> g_device{-1};
int g_cus = 0;
int g_lds_max = 65536;   // LDS bytes a workgroup may use (160 KB on gfx950)
std::mutex g_init_mu;
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
// Each host thread works on its own non-blocking stream, so independent solves
// issued from several threads overlap on the GPU.
hipStream_t stream() {
  if (!t_stream && g_device >= 0) {
    if (hipSetDevice(g_device) != hipSuccess ||
        hipStreamCreateWithFlags(&t_stream, hipStreamNonBlocking) != hipSuccess)
      t_stream = nullptr;
  }
  return t_stream;
}
int num_cus() { return g_cus; }
int lds_max_bytes() { return g_lds_max; }

// Called on entry to every C ABI function that touches the device.  It also
// clears the runtime's last-error state, which a HIP error handled in an
// earlier call (a failed allocation, say) leaves set, so that the launch
// checks of this call report only this call's errors.
int ensure_init() {
  (void)hipGetLastError();
  if (g_device >= 0 && stream()) return 0;
  return sa_init(0);
}

int spmv_pick_rows(const int32_t *rp, int64_t n) {
  int best = 0;
  for (int rb : {256, 128, 64, 32}) {
    bool fits = true;
    for (int64_t r = 0; r < n && fits; r += rb) fits = rp[std::min<int64_t>(n, r + rb)] - rp[r] <= SPMV_NZB;
    if (!fits) continue;
    best = rb;
    if ((n + rb - 1) / rb >= SPMV_MIN_BLOCKS) break;   // enough workgroups to fill the GPU
  }
  return best;
}

// Value/index bytes above which an operator uses the stream SpMV (settable
// with sa_set_spmv_stream_bytes).
std::atomic<int64_t> g_stream_bytes{64ll << 20};
// Number of solves the application runs concurrently on the device (partition
// hint for bases created afterwards, sa_set_concurrent_solves).
std::atomic<int> g_concurrent{1};
int concurrent_solves() { return g_concurrent.load(); }

}  // namespace sa

using namespace sa;

struct sa_mat_s {
  Mat A;
};
struct sa_bv_s {
  BV V;
};
struct sa_eps_s {
  sa_mat A = nullptr, A32 = nullptr;
  KsParams p;
  bool ncv_set = false, mpd_set = false, maxit_set = false;
  std::vector<double> init;   // initial space, nini x n
  int nini = 0;
  std::vector<double> defl;   // deflation space, nds x n (EPSSetDeflationSpace), cleared by each solve
  int nds = 0;
  int64_t defl_n = 0;
  BV D;                       // its orthonormalized columns on the device
  bool mixed = false;
  double tol32 = 1e-5;
  BV V;                        // FP64 basis holding the solution
  bool twosided = false;       // EPSSetTwoSided
  bool shift_set = false;      // STSetShift given (else sigma = target, STSetDefaultShift)
  sa_mat AT = nullptr;         // A^T, built at the first two-sided or two-side-balanced solve
  int balance = SA_BALANCE_NONE, balance_its = 5;   // EPSSetBalance
  double balance_cutoff = 1e-8;
  const void *AT_of = nullptr; // the operator AT was built from
  BV W;                        // left basis (two-sided)
  KsResult res, res32;
  sa_eps_stats st{};
  bool solved = false;
  double *rpart = nullptr;     // residual partial sums of EPSComputeError (allocated on first use)
  // EPSMonitorSet (epsmon.c:96-122, at most MAXEPSMONITORS = 5) and
  // EPSSetStoppingTestFunction (epsopts.c:863-885)
  std::vector<std::pair<sa_eps_monitor_fn, void *>> monitors;
  sa_eps_stopping_fn stopping = nullptr;
  void *stopping_ctx = nullptr;
  // EPSSetConvergenceTestFunction (epsopts.c:751-770)
  sa_eps_converged_fn converged = nullptr;
  void *converged_ctx = nullptr;
  ScUser cmp;                      // EPSSetEigenvalueComparison (EPS_WHICH_USER)
  double nrma = 0.0, nrmb = 0.0;   // ||A||_inf, ||B||_inf (0 until computed; reset with the operators)
};

#define API_INIT() SA_TRY(ensure_init())

extern "C" {

int sa_init(int device) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_device == device && t_stream) return 0;
  int ndev = 0;
  SA_HIP(hipGetDeviceCount(&ndev));
  SA_CHECK(ndev > 0, "no HIP device found");
  SA_CHECK(device >= 0 && device < ndev, "device index out of range");
  SA_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  SA_HIP(hipGetDeviceProperties(&prop, device));
  SA_CHECK(std::string(prop.gcnArchName).rfind("gfx950", 0) == 0,
           std::string("libslepc_amd is built for gfx950, found ") + prop.gcnArchName);
  g_cus = prop.multiProcessorCount;
  g_lds_max = (int)prop.sharedMemPerBlock;
  SA_TRY(noalloc_scratch_setup(device));
  if (t_stream) (void)hipStreamDestroy(t_stream);
  t_stream = nullptr;
  g_device = device;
  SA_CHECK(stream() != nullptr, "could not create the HIP stream");
  return 0;
}

const char *sa_last_error(void) { return g_err.c_str(); }
void *sa_stream(void) { return (void *)stream(); }
int sa_synchronize(void) {
  API_INIT();
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}
int sa_malloc(void **p, int64_t bytes) {
  API_INIT();
  SA_HIP(hipMalloc(p, (size_t)std::max<int64_t>(bytes, 1)));
  return 0;
}
int sa_free(void *p) {
  SA_HIP(hipFree(p));
  return 0;
}
int sa_memcpy_h2d(void *dst, const void *src, int64_t bytes) {
  API_INIT();
  SA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}
int sa_memcpy_d2h(void *dst, const void *src, int64_t bytes) {
  API_INIT();
  SA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// ---------------------------------------------------------------- Mat
int sa_mat_create_csr(int n, int64_t nnz, const int32_t *rp, const int32_t *ci, const double *val, int prec,
                      sa_mat *out) {
  API_INIT();
  SA_CHECK(out && n >= 0 && nnz >= 0, "invalid arguments to sa_mat_create_csr");
  SA_CHECK(prec == SA_FP64 || prec == SA_FP32, "invalid precision");
  SA_CHECK(rp && (nnz == 0 || (ci && val)), "null CSR arrays");
  SA_CHECK(rp[0] == 0 && rp[n] == nnz, "row pointer does not match nnz");
  for (int r = 0; r < n; r++) {
    SA_CHECK(rp[r + 1] >= rp[r], "row pointer not monotone");
    for (int32_t p = rp[r]; p < rp[r + 1]; p++) {
      SA_CHECK(ci[p] >= 0 && ci[p] < n, "column index out of range");
      SA_CHECK(p == rp[r] || ci[p] > ci[p - 1], "column indices must ascend within a row");
    }
  }
  BuildGuard<sa_mat, sa_mat_destroy> guard{new sa_mat_s()};
  sa_mat M = guard.h;
  Mat &A = M->A;
  A.n = n;
  A.nnz = nnz;
  A.prec = prec;
  const double avg = n ? (double)nnz / n : 1.0;
  A.lanes = avg <= 4.5 ? 4 : (avg <= 9 ? 8 : (avg <= 18 ? 16 : 32));
  const size_t es = prec_size(prec);
  // operators larger than the caches: the stream SpMV (small ones keep the
  // vector-CSR kernel, whose many short workgroups win when the operator is
  // cache-resident)
  A.spmv_rb = nnz * (int64_t)(es + 4) > g_stream_bytes.load() ? spmv_pick_rows(rp, n) : 0;
  hipStream_t s = stream();
  if (hipMalloc(&A.rowptr, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&A.col, sizeof(int32_t) * std::max<int64_t>(1, nnz)) != hipSuccess ||
      hipMalloc(&A.val, es * std::max<int64_t>(1, nnz)) != hipSuccess) {
    set_error("hipMalloc failed for the CSR arrays");
    return 1;
  }
  SA_HIP(hipMemcpyAsync(A.rowptr, rp, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
  if (nnz) {
    SA_HIP(hipMemcpyAsync(A.col, ci, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    if (prec == SA_FP64) {
      SA_HIP(hipMemcpyAsync(A.val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    } else {
      std::vector<float> f(val, val + nnz);
      SA_HIP(hipMemcpyAsync(A.val, f.data(), sizeof(float) * nnz, hipMemcpyHostToDevice, s));
      SA_HIP(hipStreamSynchronize(s));
    }
  }
  SA_HIP(hipStreamSynchronize(s));
  *out = guard.release();
  return 0;
}

int64_t sa_set_spmv_stream_bytes(int64_t bytes) { return g_stream_bytes.exchange(std::max<int64_t>(0, bytes)); }
int sa_set_concurrent_solves(int nsolves) { return g_concurrent.exchange(std::max(1, nsolves)); }

int sa_mat_destroy(sa_mat M) {
  if (!M) return 0;
  if (M->A.rowptr) (void)hipFree(M->A.rowptr);
  if (M->A.col) (void)hipFree(M->A.col);
  if (M->A.val) (void)hipFree(M->A.val);
  delete M;
  return 0;
}

int sa_mat_get_info(sa_mat M, int *n, int64_t *nnz, int *prec) {
  SA_CHECK(M, "null matrix");
  if (n) *n = M->A.n;
  if (nnz) *nnz = M->A.nnz;
  if (prec) *prec = M->A.prec;
  return 0;
}

int sa_mat_mult(sa_mat M, const void *x, void *y) {
  API_INIT();
  SA_CHECK(M && x && y, "null argument to sa_mat_mult");
  return launch_spmv(M->A, x, y, stream());
}

// ---------------------------------------------------------------- BV
int sa_bv_create(int64_t n, int m, int prec, sa_bv *out) {
  API_INIT();
  SA_CHECK(out, "null output");
  SA_CHECK(m < GS_MAXC, "too many columns (the fused kernels handle up to 127)");
  auto *B = new sa_bv_s();
  int rc = bv_alloc(B->V, n, m, prec);
  if (rc) {
    bv_free(B->V);
    delete B;
    return rc;
  }
  *out = B;
  return 0;
}

int sa_bv_destroy(sa_bv B) {
  if (!B) return 0;
  bv_free(B->V);
  delete B;
  return 0;
}

int sa_bv_get_info(sa_bv B, int64_t *n, int *m, int64_t *ld, int *prec) {
  SA_CHECK(B, "null BV");
  if (n) *n = B->V.n;
  if (m) *m = B->V.m;
  if (ld) *ld = B->V.ld;
  if (prec) *prec = B->V.prec;
  return 0;
}

int sa_bv_get_array(sa_bv B, void **data) {
  SA_CHECK(B && data, "null argument");
  *data = B->V.data;
  return 0;
}

int sa_bv_set_active_columns(sa_bv B, int l, int k) {
  SA_CHECK(B && l >= 0 && l <= k && k <= B->V.m, "invalid active columns");
  B->V.l = l;
  B->V.k = k;
  return 0;
}

int sa_bv_get_active_columns(sa_bv B, int *l, int *k) {
  SA_CHECK(B, "null BV");
  if (l) *l = B->V.l;
  if (k) *k = B->V.k;
  return 0;
}

int sa_bv_set_column(sa_bv B, int j, const double *x) {
  API_INIT();
  SA_CHECK(B && x && j >= 0 && j < B->V.m, "invalid column");
  return bv_upload_column(B->V, j, x);
}

int sa_bv_get_column(sa_bv B, int j, double *x) {
  API_INIT();
  SA_CHECK(B && x && j >= 0 && j < B->V.m, "invalid column");
  return bv_download_column(B->V, j, x);
}

int sa_bv_dotvec(sa_bv B, const void *x, double *q) {
  API_INIT();
  SA_CHECK(B && x, "null argument");
  BV &V = B->V;
  const int nc = V.k - V.l;
  double *qd = q ? V.dq : V.dbuf;   // q == NULL: column 0 of the coefficient buffer
  SA_TRY(launch_dotvec(V, V.l, V.k, x, qd, stream()));
  if (q && nc > 0) SA_HIP(hipMemcpyAsync(q, V.dq, sizeof(double) * nc, hipMemcpyDeviceToHost, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_multvec(sa_bv B, double alpha, double beta, void *y, const double *q) {
  API_INIT();
  SA_CHECK(B && y, "null argument");
  BV &V = B->V;
  const int nc = V.k - V.l;
  const double *qd = V.dbuf;
  if (q) {
    if (nc > 0) {
      std::memcpy(V.hq, q, sizeof(double) * nc);
      SA_HIP(hipMemcpyAsync(V.dq, V.hq, sizeof(double) * nc, hipMemcpyHostToDevice, stream()));
    }
    qd = V.dq;
  }
  SA_TRY(launch_multvec(V, V.l, V.k, alpha, qd, beta, y, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_multinplace(sa_bv B, const double *Q, int ldq, int s_, int e) {
  API_INIT();
  SA_CHECK(B && Q, "null argument");
  BV &V = B->V;
  SA_CHECK(s_ >= V.l && e <= V.k && s_ <= e, "columns out of the active range");
  SA_CHECK(ldq >= V.k, "leading dimension of Q too small");
  const int64_t bytes = sizeof(double) * (int64_t)ldq * e;   // Q holds at least e columns
  SA_CHECK(bytes <= (int64_t)sizeof(double) * V.m * V.m + 64 * (int64_t)sizeof(double), "Q too large");
  std::memcpy(V.hq, Q, bytes);
  SA_HIP(hipMemcpyAsync(V.dq, V.hq, bytes, hipMemcpyHostToDevice, stream()));
  SA_TRY(launch_multinplace(V, V.l, V.k, s_, e, V.dq, ldq, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_scale_column(sa_bv B, int j, double alpha) {
  API_INIT();
  SA_CHECK(B && j < B->V.m, "invalid column");
  BV &V = B->V;
  if (j >= 0) {
    SA_TRY(launch_scale(V.prec, col_ptr(V, j), V.n, alpha, stream()));
  } else {
    for (int c = V.l; c < V.k; c++) SA_TRY(launch_scale(V.prec, col_ptr(V, c), V.n, alpha, stream()));
  }
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_norm_column(sa_bv B, int j, double *nrm) {
  API_INIT();
  SA_CHECK(B && nrm && j >= 0 && j < B->V.m, "invalid column");
  BV &V = B->V;
  double *out = V.dq;
  SA_TRY(launch_dot_self(V.prec, col_ptr(V, j), V.n, out, V.ws.part, stream()));
  double h = 0.0;
  SA_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  *nrm = std::sqrt(h);
  return 0;
}

int sa_bv_norm(sa_bv B, double *nrm) {
  API_INIT();
  SA_CHECK(B && nrm, "null argument");
  BV &V = B->V;
  double tot = 0.0;
  for (int c = V.l; c < V.k; c++) {
    double v = 0.0;
    SA_TRY(sa_bv_norm_column(B, c, &v));
    tot += v * v;
  }
  *nrm = std::sqrt(tot);
  return 0;
}

int sa_bv_copy_column(sa_bv B, int src, int dst) {
  API_INIT();
  SA_CHECK(B && src >= 0 && dst >= 0 && src < B->V.m && dst < B->V.m, "invalid column");
  BV &V = B->V;
  SA_HIP(hipMemcpyAsync(col_ptr(V, dst), col_ptr(V, src), prec_size(V.prec) * V.n, hipMemcpyDeviceToDevice,
                        stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_mat_mult_column(sa_bv B, sa_mat M, int j) {
  API_INIT();
  SA_CHECK(B && M && j >= 0 && j + 1 < B->V.m, "invalid column");
  SA_CHECK(M->A.prec == B->V.prec && M->A.n == B->V.n, "matrix and basis differ in size or precision");
  SA_TRY(launch_spmv(M->A, col_ptr(B->V, j), col_ptr(B->V, j + 1), stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_orthonormalize_column(sa_bv B, int j, double eta, double *norm, int *lindep) {
  API_INIT();
  SA_CHECK(B && j >= 0 && j < B->V.m && j < GS_MAXC, "invalid column");
  return bv_orthonormalize(B->V, j, eta > 0 ? eta : 0.7071, norm, lindep);
}

int sa_bv_set_orthogonalization(sa_bv B, int refine) {
  SA_CHECK(B && refine >= SA_REFINE_IFNEEDED && refine <= SA_REFINE_ALWAYS, "invalid refinement type");
  B->V.refine = refine;
  return 0;
}

int sa_bv_orthogonalize_some_column(sa_bv B, int j, const int *which, double eta, double *h, double *norm,
                                    int *lindep) {
  API_INIT();
  SA_CHECK(B && which && j >= 0 && j < B->V.m, "invalid column");
  std::vector<int> cols;
  for (int i = 0; i < j; i++)
    if (which[i]) cols.push_back(i);
  double nrm = 0.0;
  int ld = 0;
  SA_TRY(bv_mgs_column(B->V, j, cols.data(), (int)cols.size(), eta > 0 ? eta : 0.7071, &nrm, &ld));
  if (h) {
    SA_HIP(hipMemcpy(h, B->V.dbuf + (size_t)j * B->V.m, sizeof(double) * (j + 1), hipMemcpyDeviceToHost));
  }
  if (norm) *norm = nrm;
  if (lindep) *lindep = ld;
  return 0;
}

int sa_bv_get_coefficients(sa_bv B, int c0, int c1, double *buf) {
  API_INIT();
  SA_CHECK(B && buf && c0 >= 0 && c1 <= B->V.m && c0 <= c1, "invalid coefficient range");
  const int m = B->V.m;
  if (c1 > c0)
    SA_HIP(hipMemcpyAsync(buf, B->V.dbuf + (size_t)c0 * m, sizeof(double) * m * (c1 - c0), hipMemcpyDeviceToHost,
                          stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_dot(sa_bv X, sa_bv Y, double *M, int ldm) {
  API_INIT();
  SA_CHECK(X && Y && M, "null argument");
  const BV &x = X->V, &y = Y->V;
  SA_CHECK(ldm >= y.k, "M has fewer rows than the active columns of Y");
  SA_CHECK(x.n == y.n && x.prec == y.prec, "X and Y differ in length or precision");
  const int mx = x.k - x.l, my = y.k - y.l;
  if (mx <= 0 || my <= 0) return 0;
  // the result goes through X's coefficient staging (m x m doubles >= mx x my
  // when Y has no more columns than X; else a temporary)
  const bool stage = (int64_t)mx * my <= (int64_t)x.m * x.m;
  double *dev = stage ? x.dq : nullptr;
  if (!stage) SA_HIP(hipMalloc(&dev, sizeof(double) * (size_t)mx * my));
  std::vector<double> h((size_t)mx * my);
  int rc = launch_bvdot(x, x.l, x.k, y, y.l, y.k, dev, stream());
  if (!rc && hipMemcpy(h.data(), dev, sizeof(double) * h.size(), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("HIP error in BVDot");
    rc = 1;
  }
  if (!stage) (void)hipFree(dev);
  if (rc) return rc;
  for (int j = 0; j < mx; j++)
    for (int i = 0; i < my; i++) M[(size_t)(y.l + i) + (size_t)(x.l + j) * ldm] = h[i + (size_t)j * my];
  return 0;
}

int sa_bv_matmult(sa_bv X, sa_mat A, sa_bv Y) {
  API_INIT();
  SA_CHECK(X && A && Y, "null argument");
  const BV &x = X->V, &y = Y->V;
  SA_CHECK(x.k - x.l == y.k - y.l, "Y and X have different numbers of active columns");
  SA_CHECK(A->A.n == x.n && A->A.n == y.n && A->A.prec == x.prec && x.prec == y.prec,
           "operator and bases differ in size or precision");
  SA_TRY(launch_spmm(A->A, x, x.l, y, y.l, x.k - x.l, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_mult(sa_bv Yb, double alpha, double beta, sa_bv Xb, const double *Q, int ldq) {
  API_INIT();
  SA_CHECK(Xb && Yb && Xb != Yb, "BVMult needs two distinct bases");
  const BV &X = Xb->V;
  BV &Y = Yb->V;
  const int K = X.k - X.l, N = Y.k - Y.l;
  SA_CHECK(Q == nullptr || ldq >= X.k, "leading dimension of Q smaller than the active columns of X");
  if (N <= 0) return 0;
  hipStream_t s = stream();
  if (!Q) {
    SA_TRY(launch_bvmult(X, X.l, X.k, Y, Y.l, Y.k, nullptr, 0, alpha, beta, s));
    SA_HIP(hipStreamSynchronize(s));
    return 0;
  }
  // Q(X.l:X.k, Y.l:Y.k) compacted to K x N on the device
  std::vector<double> q((size_t)std::max(1, K) * N, 0.0);
  for (int j = 0; j < N; j++)
    for (int i = 0; i < K; i++) q[i + (size_t)j * K] = Q[(X.l + i) + (size_t)(Y.l + j) * ldq];
  double *qd = nullptr;
  SA_HIP(hipMalloc((void **)&qd, sizeof(double) * q.size()));
  int rc = 0;
  if (hipMemcpy(qd, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("hipMemcpy of Q failed in BVMult");
    rc = 1;
  }
  if (!rc) rc = launch_bvmult(X, X.l, X.k, Y, Y.l, Y.k, qd, std::max(1, K), alpha, beta, s);
  if (!rc && hipStreamSynchronize(s) != hipSuccess) {
    set_error("HIP error in BVMult");
    rc = 1;
  }
  (void)hipFree(qd);
  return rc;
}

int sa_bv_copy(sa_bv Xb, sa_bv Yb) {
  API_INIT();
  SA_CHECK(Xb && Yb, "null argument");
  const BV &X = Xb->V;
  BV &Y = Yb->V;
  const int nc = X.k - X.l;
  SA_CHECK(X.n == Y.n && X.prec == Y.prec, "BVCopy: bases differ in length or precision");
  SA_CHECK(Y.l + nc <= Y.m, "BVCopy: the target has too few columns");
  if (nc <= 0 || X.n == 0) return 0;
  const size_t es = prec_size(X.prec);
  SA_HIP(hipMemcpy2DAsync(col_ptr(Y, Y.l), es * Y.ld, col_ptr(X, X.l), es * X.ld, es * X.n, nc,
                          hipMemcpyDeviceToDevice, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_bv_resize(sa_bv B, int m, int copy) {
  API_INIT();
  SA_CHECK(B && m > 0 && m < GS_MAXC, "invalid number of columns for BVResize");
  BV &V = B->V;
  if (V.m == m) return 0;
  BV W;
  int rc = bv_alloc(W, V.n, m, V.prec);
  if (!rc && copy && V.n > 0) {
    const size_t es = prec_size(V.prec);
    if (hipMemcpy2D(W.data, es * W.ld, V.data, es * V.ld, es * V.n, std::min(m, V.m), hipMemcpyDeviceToDevice) !=
        hipSuccess) {
      set_error("hipMemcpy2D failed in BVResize");
      rc = 1;
    }
  }
  if (rc) {
    bv_free(W);
    return rc;
  }
  W.l = std::min(V.l, m);
  W.k = std::min(V.k, m);
  W.passes_done = V.passes_done;
  bv_free(V);
  V = W;
  return 0;
}

// BVOrthogonalizeCGS1 (bvorthog.c:91-132): dots of column j with V(:, l:j+1)
// in one device reduction, v -= V(:, l:j) c, ||v|| from the self dot and the
// estimate ||v'|| = sqrt(||v||^2 - sum c^2) (the true norm when that is <= 0).
int sa_bv_gs_pass(sa_bv B, int j, double *h, double *c, double *onorm, double *norm) {
  API_INIT();
  SA_CHECK(B && c, "null argument");
  BV &V = B->V;
  SA_CHECK(j >= V.l && j < V.m && j - V.l < GS_MAXC, "invalid column for a Gram-Schmidt pass");
  hipStream_t s = stream();
  const int nc = j - V.l;
  SA_TRY(launch_dotvec(V, V.l, j + 1, col_ptr(V, j), V.dq, s));
  std::vector<double> d(nc + 1);
  SA_HIP(hipMemcpyAsync(d.data(), V.dq, sizeof(double) * (nc + 1), hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  SA_CHECK(d[nc] > -1e-8 * (1.0 + std::fabs(d[nc])), "Invalid inner product (negative squared norm)");
  const double beta = d[nc] < 0.0 ? 0.0 : std::sqrt(d[nc]);   // BV_SafeSqrt
  if (nc > 0) SA_TRY(launch_multvec(V, V.l, j, -1.0, V.dq, 1.0, col_ptr(V, j), s));
  double sum = 0.0;
  for (int i = 0; i < nc; i++) {
    c[i] = d[i];
    sum += d[i] * d[i];
    if (h) h[i] += d[i];
  }
  if (onorm) *onorm = beta;
  if (norm) {
    double nn = beta * beta - sum;
    if (nn <= 0.0) {
      SA_TRY(launch_dot_self(V.prec, col_ptr(V, j), V.n, V.dq, V.ws.part, s));
      SA_HIP(hipMemcpyAsync(&nn, V.dq, sizeof(double), hipMemcpyDeviceToHost, s));
      SA_HIP(hipStreamSynchronize(s));
      *norm = std::sqrt(nn);
    } else {
      *norm = std::sqrt(nn);
    }
  }
  SA_HIP(hipStreamSynchronize(s));
  V.passes_done++;
  return 0;
}

int sa_bv_orthogonalize(sa_bv V, int method, double *R, int ldr) {
  API_INIT();
  SA_CHECK(V, "null argument");
  return bv_orthogonalize(V->V, method, R, ldr);
}

int sa_bv_mat_arnoldi(sa_bv B, sa_mat M, int k, int *m, double eta, double *beta, int *breakdown) {
  API_INIT();
  SA_CHECK(B && M && m && beta && breakdown, "null argument");
  SA_CHECK(M->A.prec == B->V.prec && M->A.n == B->V.n, "matrix and basis differ in size or precision");
  return bv_sweep(B->V, M->A, k, m, eta > 0 ? eta : 0.7071, beta, breakdown);
}

// ---------------------------------------------------------------- EPS
int sa_eps_create(sa_eps *out) {
  SA_CHECK(out, "null output");
  *out = new sa_eps_s();
  return 0;
}

int sa_eps_destroy(sa_eps E) {
  if (!E) return 0;
  if (E->rpart) (void)hipFree(E->rpart);
  bv_free(E->V);
  bv_free(E->W);
  bv_free(E->D);
  if (E->AT) sa_mat_destroy(E->AT);
  delete E;
  return 0;
}

int sa_eps_set_operators(sa_eps E, sa_mat A) {
  SA_CHECK(E && A, "null argument");
  SA_CHECK(A->A.prec == SA_FP64, "the operator of EPS must be FP64 (use sa_eps_set_operators_fp32 for the bulk phase)");
  E->A = A;
  E->solved = false;
  E->nrma = E->nrmb = 0.0;   // epssetup.c:491-492
  // a cached A^T belongs to the previous operator (an operator allocated later
  // at the same address must not find it)
  if (E->AT) sa_mat_destroy(E->AT);
  E->AT = nullptr;
  E->AT_of = nullptr;
  return 0;
}

int sa_eps_set_true_residual(sa_eps E, int trueres) {
  SA_CHECK(E, "null EPS");
  E->p.trueres = trueres != 0;
  E->solved = false;
  return 0;
}

int sa_eps_get_true_residual(sa_eps E, int *trueres) {
  SA_CHECK(E && trueres, "null argument");
  *trueres = E->p.trueres ? 1 : 0;
  return 0;
}

int sa_eps_set_operators_fp32(sa_eps E, sa_mat A32) {
  SA_CHECK(E && A32 && A32->A.prec == SA_FP32, "expected an FP32 matrix");
  E->A32 = A32;
  return 0;
}

int sa_eps_set_problem_type(sa_eps E, int hermitian) {
  SA_CHECK(E, "null EPS");
  E->p.hermitian = hermitian != 0;
  return 0;
}

int sa_eps_set_type(sa_eps E, int type) {
  SA_CHECK(E && (type == SA_EPS_KRYLOVSCHUR || type == SA_EPS_ARNOLDI || type == SA_EPS_LANCZOS), "invalid EPS type");
  E->p.solver = type;
  return 0;
}

int sa_eps_set_which(sa_eps E, int which) {
  SA_CHECK(E && ((which >= SA_LARGEST_MAGNITUDE && which <= SA_TARGET_REAL) || which == SA_WHICH_USER),
           "invalid which");
  E->p.which = which;
  return 0;
}

int sa_eps_set_eigenvalue_comparison(sa_eps E, sa_eps_compare_fn cmp, void *ctx) {
  SA_CHECK(E, "null EPS");
  E->cmp.fn = cmp;
  E->cmp.ctx = cmp ? ctx : nullptr;
  E->p.which = cmp ? SA_WHICH_USER : SA_LARGEST_MAGNITUDE;
  E->solved = false;
  return 0;
}

int sa_eps_set_st(sa_eps E, int st_type, int shift_given, double shift) {
  SA_CHECK(E && (st_type == SA_ST_SHIFT || st_type == SA_ST_SINVERT), "unsupported spectral transformation");
  SA_CHECK(st_type == SA_ST_SINVERT || !shift_given || shift == 0.0, "the shift transformation runs with sigma = 0");
  E->p.st_type = st_type;
  E->shift_set = shift_given != 0;
  E->p.st_shift = shift_given ? shift : 0.0;
  E->solved = false;
  return 0;
}

int sa_eps_set_st_ksp(sa_eps E, int pc_type, double rtol, int max_it) {
  SA_CHECK(E && (pc_type == SA_PC_NONE || pc_type == SA_PC_JACOBI || pc_type == SA_PC_LU), "unsupported preconditioner");
  SA_CHECK(rtol > 0.0 && rtol < 1.0 && max_it > 0, "invalid linear solver tolerances");
  E->p.ksp_jacobi = pc_type == SA_PC_JACOBI;
  E->p.ksp_direct = pc_type == SA_PC_LU;
  E->p.ksp_rtol = rtol;
  E->p.ksp_max_it = max_it;
  return 0;
}

int sa_eps_set_st_ksp_type(sa_eps E, int ksp_type, int restart) {
  SA_CHECK(E && (ksp_type == SA_KSP_BCGS || ksp_type == SA_KSP_GMRES), "unsupported KSP type");
  SA_CHECK(ksp_type != SA_KSP_GMRES || (restart >= 1 && restart <= GS_MAXC - 2), "GMRES restart must be in [1, 126]");
  E->p.ksp_gmres = ksp_type == SA_KSP_GMRES;
  if (ksp_type == SA_KSP_GMRES) E->p.ksp_restart = restart;
  E->solved = false;
  return 0;
}

int sa_eps_krylovschur_set_restart(sa_eps E, double keep) {
  SA_CHECK(E, "null EPS");
  SA_CHECK(keep >= 0.1 && keep <= 0.9, "The keep argument must be in the range [0.1,0.9]");
  E->p.keep = keep;
  E->solved = false;
  return 0;
}

int sa_eps_krylovschur_get_restart(sa_eps E, double *keep) {
  SA_CHECK(E && keep, "null argument");
  *keep = E->p.keep;
  return 0;
}

int sa_eps_krylovschur_set_locking(sa_eps E, int lock) {
  SA_CHECK(E, "null EPS");
  E->p.lock = lock != 0;
  E->solved = false;
  return 0;
}

int sa_eps_krylovschur_get_locking(sa_eps E, int *lock) {
  SA_CHECK(E && lock, "null argument");
  *lock = E->p.lock ? 1 : 0;
  return 0;
}

int sa_eps_arnoldi_set_delayed(sa_eps E, int delayed) {
  SA_CHECK(E, "null EPS");
  E->p.arnoldi_delayed = delayed != 0;
  E->solved = false;
  return 0;
}

int sa_eps_arnoldi_get_delayed(sa_eps E, int *delayed) {
  SA_CHECK(E && delayed, "null argument");
  *delayed = E->p.arnoldi_delayed ? 1 : 0;
  return 0;
}

int sa_eps_lanczos_set_reorthog(sa_eps E, int reorthog) {
  SA_CHECK(E, "null EPS");
  SA_CHECK(reorthog >= SA_LANCZOS_REORTHOG_LOCAL && reorthog <= SA_LANCZOS_REORTHOG_DELAYED,
           "unsupported Lanczos reorthogonalization type");
  E->p.lanczos_local = reorthog == SA_LANCZOS_REORTHOG_LOCAL;
  E->p.lanczos_partial = reorthog == SA_LANCZOS_REORTHOG_PERIODIC    ? 1
                         : reorthog == SA_LANCZOS_REORTHOG_PARTIAL   ? 2
                         : reorthog == SA_LANCZOS_REORTHOG_SELECTIVE ? 3
                         : reorthog == SA_LANCZOS_REORTHOG_DELAYED   ? 4
                                                                     : 0;
  E->solved = false;
  return 0;
}

int sa_eps_lanczos_get_reorthog(sa_eps E, int *reorthog) {
  SA_CHECK(E && reorthog, "null argument");
  *reorthog = E->p.lanczos_local          ? SA_LANCZOS_REORTHOG_LOCAL
              : E->p.lanczos_partial == 1 ? SA_LANCZOS_REORTHOG_PERIODIC
              : E->p.lanczos_partial == 2 ? SA_LANCZOS_REORTHOG_PARTIAL
              : E->p.lanczos_partial == 3 ? SA_LANCZOS_REORTHOG_SELECTIVE
              : E->p.lanczos_partial == 4 ? SA_LANCZOS_REORTHOG_DELAYED
                                          : SA_LANCZOS_REORTHOG_FULL;
  return 0;
}

int sa_eps_set_two_sided(sa_eps E, int twosided) {
  SA_CHECK(E, "null EPS");
  E->twosided = twosided != 0;
  E->solved = false;
  return 0;
}

int sa_eps_set_target(sa_eps E, double target) {
  SA_CHECK(E, "null EPS");
  E->p.target = target;
  return 0;
}

int sa_eps_set_extraction(sa_eps E, int extraction) {
  SA_CHECK(E && (extraction == SA_EPS_RITZ || extraction == SA_EPS_HARMONIC), "unsupported extraction type");
  E->p.extraction = extraction;
  return 0;
}

int sa_eps_set_dimensions(sa_eps E, int nev, int ncv, int mpd) {
  SA_CHECK(E && nev > 0, "nev must be positive");
  E->p.nev = nev;
  E->ncv_set = ncv > 0;
  E->mpd_set = mpd > 0;
  if (ncv > 0) E->p.ncv = ncv;
  if (mpd > 0) E->p.mpd = mpd;
  return 0;
}

int sa_eps_set_tolerances(sa_eps E, double tol, int max_it) {
  SA_CHECK(E, "null EPS");
  if (tol > 0) E->p.tol = tol;
  E->maxit_set = max_it > 0;
  if (max_it > 0) E->p.max_it = max_it;
  return 0;
}

int sa_eps_set_initial_space(sa_eps E, int nvec, const double *vecs) {
  SA_CHECK(E && E->A, "set the operator first");
  SA_CHECK(nvec >= 0 && (nvec == 0 || vecs), "invalid initial space");
  const int n = E->A->A.n;
  E->init.assign(vecs, vecs + (size_t)nvec * n);
  E->nini = nvec;
  return 0;
}

int sa_eps_set_deflation_space(sa_eps E, int nvec, const double *vecs) {
  SA_CHECK(E && E->A, "set the operator first");
  SA_CHECK(nvec >= 0 && (nvec == 0 || vecs), "invalid deflation space");   // epssetup.c:575
  SA_CHECK(nvec <= GS_MAXC, "a deflation space of at most 128 vectors is supported");
  const int64_t n = E->A->A.n;
  E->defl.assign(vecs, vecs + (size_t)nvec * n);
  E->nds = nvec;
  E->defl_n = n;
  return 0;
}

int sa_eps_set_mixed(sa_eps E, int mixed, double tol32) {
  SA_CHECK(E, "null EPS");
  E->mixed = mixed != 0;
  if (tol32 > 0) E->tol32 = tol32;
  return 0;
}

// EPSSetUp + EPSSetDimensions_Default (epssetup.c:669-694, krylovschur.c:93-120)
// Trampolines from the driver's hooks to the user's functions (EPSMonitor,
// epsmon.c:39-46: every monitor in the order set).
static int eps_monitor_hook(void *owner, int its, int nconv, const double *eigr, const double *eigi,
                            const double *errest, int nest) {
  sa_eps E = static_cast<sa_eps>(owner);
  for (auto &m : E->monitors) {
    const int rc = m.first(E, its, nconv, eigr, eigi, errest, nest, m.second);
    if (rc) return rc;
  }
  return 0;
}

static int eps_stopping_hook(void *owner, int its, int max_it, int nconv, int nev, int *reason) {
  sa_eps E = static_cast<sa_eps>(owner);
  return E->stopping(E, its, max_it, nconv, nev, reason, E->stopping_ctx);
}

static int eps_converged_hook(void *owner, double re, double im, double res, double *errest) {
  sa_eps E = static_cast<sa_eps>(owner);
  return E->converged(E, re, im, res, errest, E->converged_ctx);
}

static int eps_setup_dims(sa_eps E, KsParams &p) {
  const int n = E->A->A.n;
  p = E->p;
  p.hooks.owner = E;
  p.hooks.monitor = E->monitors.empty() ? nullptr : eps_monitor_hook;
  p.hooks.stopping = E->stopping ? eps_stopping_hook : nullptr;
  p.hooks.converged = E->converged ? eps_converged_hook : nullptr;
  // initialization of the matrix norms (epssetup.c:351-359): ||A||_inf for
  // the norm-relative test; ||B|| is not computed for a standard problem
  if (p.conv == SA_CONV_NORM && E->nrma == 0.0) SA_TRY(mat_norm_inf(E->A->A, &E->nrma));
  p.nrma = E->nrma;
  p.nrmb = E->nrmb;
  // safeguard for small problems (epssetup.c:342-345): nev and a set ncv at most n
  if (p.nev > n) p.nev = n;
  if (E->ncv_set && p.ncv > n) p.ncv = n;
  if (!E->ncv_set && !E->mpd_set) p.ncv = std::min(n, std::max(2 * p.nev, p.nev + 15));
  else if (!E->ncv_set) p.ncv = std::min(n, p.nev + p.mpd);
  if (!E->mpd_set) p.mpd = p.ncv;
  SA_CHECK(p.ncv >= p.nev + 1 || (p.ncv == p.nev && p.ncv == n), "The value of ncv must be at least nev+1");
  SA_CHECK(p.ncv <= p.nev + p.mpd, "The value of ncv must not be larger than nev+mpd");
  SA_CHECK(E->nini <= p.ncv, "The number of initial vectors is larger than ncv");   // epssetup.c:423
  SA_CHECK(p.ncv <= GS_MAXC - 2, "ncv above 126 is beyond the fused Gram-Schmidt kernels (127 basis columns)");
  if (!E->maxit_set) p.max_it = std::max(100, 2 * n / p.ncv);
  if (p.st_type == SA_ST_SINVERT && !E->shift_set) p.st_shift = p.target;   // STSetDefaultShift
  return 0;
}

}  // extern "C"

// The operator's CSR arrays on the host.
static int download_csr(const Mat &A, std::vector<int32_t> &rp, std::vector<int32_t> &ci, std::vector<double> &va) {
  const int n = A.n;
  const int64_t nnz = A.nnz;
  SA_CHECK(A.prec == SA_FP64, "the operator must be FP64");
  rp.resize(n + 1);
  ci.resize(nnz);
  va.resize(nnz);
  SA_HIP(hipMemcpy(rp.data(), A.rowptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  if (nnz) {
    SA_HIP(hipMemcpy(ci.data(), A.col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    SA_HIP(hipMemcpy(va.data(), A.val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
  }
  return 0;
}

// A^T as CSR (host transpose of the device arrays), built once per operator.
// A^T of a device CSR operator as a new operator (columns ascending per row).
static int build_transpose(const Mat &A, sa_mat *out) {
  const int n = A.n;
  const int64_t nnz = A.nnz;
  std::vector<int32_t> rp, ci, trp(n + 1, 0), tci(nnz);
  std::vector<double> va, tva(nnz);
  SA_TRY(download_csr(A, rp, ci, va));
  for (int64_t q = 0; q < nnz; q++) trp[ci[q] + 1]++;
  for (int r = 0; r < n; r++) trp[r + 1] += trp[r];
  std::vector<int32_t> next(trp.begin(), trp.end() - 1);
  for (int r = 0; r < n; r++)          // rows in order: columns of A^T ascend
    for (int32_t q = rp[r]; q < rp[r + 1]; q++) {
      const int32_t dst = next[ci[q]]++;
      tci[dst] = r;
      tva[dst] = va[q];
    }
  return sa_mat_create_csr(n, nnz, trp.data(), tci.data(), tva.data(), SA_FP64, out);
}

static int ensure_transpose(sa_eps E) {
  if (E->AT && E->AT_of == (const void *)E->A) return 0;
  if (E->AT) sa_mat_destroy(E->AT);
  E->AT = nullptr;
  SA_TRY(build_transpose(E->A->A, &E->AT));
  E->AT_of = (const void *)E->A;
  return 0;
}

extern "C" {

int sa_mat_build_balance(sa_mat M, int twoside, int its, double cutoff, double *D_host) {
  API_INIT();
  SA_CHECK(M && D_host && its > 0 && M->A.prec == SA_FP64, "invalid balance arguments (FP64 operator, its > 0)");
  struct Hold {
    sa_mat AT = nullptr;
    double *D = nullptr;
    ~Hold() {
      if (AT) sa_mat_destroy(AT);
      if (D) (void)hipFree(D);
    }
  } h;
  if (twoside) SA_TRY(build_transpose(M->A, &h.AT));
  SA_HIP(hipMalloc(&h.D, sizeof(double) * (size_t)std::max(1, M->A.n)));
  SA_TRY(eps_build_balance(M->A, twoside ? &h.AT->A : nullptr, twoside != 0, its, cutoff > 0 ? cutoff : 1e-8, h.D));
  SA_HIP(hipMemcpy(D_host, h.D, sizeof(double) * (size_t)M->A.n, hipMemcpyDeviceToHost));
  return 0;
}

int sa_eps_set_balance(sa_eps E, int balance, int its, double cutoff) {
  SA_CHECK(E, "null EPS");
  SA_CHECK(balance == SA_BALANCE_NONE || balance == SA_BALANCE_ONESIDE || balance == SA_BALANCE_TWOSIDE,
           "unsupported balancing type (none, oneside and twoside are available)");
  E->balance = balance;
  E->balance_its = its > 0 ? its : 5;
  E->balance_cutoff = cutoff > 0.0 ? cutoff : 1e-8;
  E->solved = false;
  return 0;
}

static int eps_solve_impl(sa_eps E);

int sa_eps_solve(sa_eps E) {
  API_INIT();
  SA_CHECK(E && E->A, "EPS has no operator");
  SA_CHECK(E->p.which != SA_WHICH_USER || E->cmp.fn, "EPS_WHICH_USER needs EPSSetEigenvalueComparison");
  // the user's comparison serves the sorts of this solve on this thread
  E->cmp.err = 0;
  struct Install {
    ScUser *prev;
    explicit Install(ScUser *u) : prev(t_sc_user) { t_sc_user = u; }
    ~Install() { t_sc_user = prev; }
  } install(E->p.which == SA_WHICH_USER ? &E->cmp : nullptr);
  const int rc = eps_solve_impl(E);
  // EPSSolve removes the deflation space at its end (epssolve.c:201-205)
  E->defl.clear();
  E->nds = 0;
  if (E->cmp.err) {
    E->solved = false;
    set_error("the eigenvalue comparison returned an error code");
    return 2;
  }
  return rc;
}

static int eps_solve_impl(sa_eps E) {
  KsParams p;
  SA_TRY(eps_setup_dims(E, p));
  const Mat &A = E->A->A;
  if (E->nds > 0) {
    // EPSSetUp (epssetup.c:412-418): BVInsertConstraints orthonormalizes the
    // deflation vectors in order and drops each one that is linearly
    // dependent on those before it (BVInsertVecs, bvfunc.c:359-374).
    SA_CHECK(E->defl_n == A.n, "the deflation vectors do not have the operator's length");
    SA_CHECK(!E->mixed && !E->twosided && p.solver != 2 && !p.arnoldi_delayed &&
                 !(E->balance != SA_BALANCE_NONE && !p.hermitian),
             "the deflation space runs with one-sided FP64 Krylov-Schur or Arnoldi (no delayed "
             "reorthogonalization, no balancing)");
    if (E->D.m != E->nds || E->D.n != A.n) {
      bv_free(E->D);
      SA_TRY(bv_alloc(E->D, A.n, E->nds, SA_FP64));
    }
    int kd = 0;
    for (int i = 0; i < E->nds; i++) {
      SA_TRY(bv_upload_column(E->D, kd, E->defl.data() + (size_t)i * A.n));
      double norm = 0.0;
      int lindep = 0;
      SA_TRY(bv_orthonormalize(E->D, kd, p.eta, &norm, &lindep));
      if (!(norm == 0.0 || lindep)) kd++;
    }
    if (kd > 0) {
      p.defl = &E->D;
      p.ndefl = kd;
    }
  }
  if (E->V.m != p.ncv + 1 || E->V.n != A.n || E->V.prec != SA_FP64) {
    bv_free(E->V);
    SA_TRY(bv_alloc(E->V, A.n, p.ncv + 1, SA_FP64));
  }
  SA_CHECK(!(E->balance != SA_BALANCE_NONE && !p.hermitian && (E->mixed || E->twosided)),
           "balancing runs with the FP64 one-sided solvers");
  SA_CHECK(!(p.trueres && E->balance != SA_BALANCE_NONE && !p.hermitian),
           "the true residual is not available with balancing");
  E->st = sa_eps_stats{};
  E->res32 = KsResult();
  const double t0 = now();
  bool have_init = E->nini > 0;
  if (E->mixed) {
    // Bulk phase in FP32 (basis, SpMV and BV products), projected problem FP64.
    SA_CHECK(E->A32 && E->A32->A.n == A.n, "mixed precision needs the FP32 operator (sa_eps_set_operators_fp32)");
    BV V32;
    SA_TRY(bv_alloc(V32, A.n, p.ncv + 1, SA_FP32));
    KsParams p32 = p;
    p32.tol = E->tol32;
    int rc = 0;
    if (have_init) rc = bv_upload_column(V32, 0, E->init.data());
    if (!rc) rc = ks_solve(p32, E->A32->A, V32, have_init, E->res32);
    if (!rc) {
      // Seed for the FP64 polish: the sum of the first nev FP32 Ritz vectors,
      // each of unit norm, promoted to FP64 (EPSSetInitialSpace of the polish).
      const int nr = std::min(E->res32.nconv, p.nev);
      SA_CHECK(nr > 0 || E->res32.nconv == 0, "internal");
      if (nr > 0) {
        std::vector<double> q(p.ncv, 0.0);
        for (int i = 0; i < nr; i++) q[E->res32.perm[i]] = 1.0;
        // V64(:,0) = V32(:,perm[0..nr)) * ones, computed in FP64 after promotion.
        rc = launch_convert(SA_FP32, V32.data, V32.ld, SA_FP64, E->V.data, E->V.ld, A.n, p.ncv + 1, stream());
        if (!rc) {
          std::memcpy(E->V.hq, q.data(), sizeof(double) * p.ncv);
          if (hipMemcpyAsync(E->V.dq, E->V.hq, sizeof(double) * p.ncv, hipMemcpyHostToDevice, stream()) !=
              hipSuccess)
            rc = 1;
        }
        // V(:,ncv) = V(:,0:ncv) q, then moved to column 0
        if (!rc) rc = launch_multvec(E->V, 0, p.ncv, 1.0, E->V.dq, 0.0, col_ptr(E->V, p.ncv), stream());
        if (!rc && hipMemcpyAsync(col_ptr(E->V, 0), col_ptr(E->V, p.ncv), sizeof(double) * A.n,
                                  hipMemcpyDeviceToDevice, stream()) != hipSuccess)
          rc = 1;
        if (!rc && hipStreamSynchronize(stream()) != hipSuccess) rc = 1;
        have_init = true;
      }
    }
    bv_free(V32);
    if (rc) return rc;
  } else if (have_init) {
    SA_TRY(bv_upload_column(E->V, 0, E->init.data()));
  }
  SA_CHECK(!(p.st_type == SA_ST_SINVERT && (E->mixed || E->twosided)),
           "shift-and-invert runs in FP64 with the one-sided solvers");
  if (E->twosided) {
    SA_CHECK(!E->mixed, "the two-sided solver runs in FP64 only");
    SA_TRY(ensure_transpose(E));
    if (E->W.m != p.ncv + 1 || E->W.n != A.n) {
      bv_free(E->W);
      SA_TRY(bv_alloc(E->W, A.n, p.ncv + 1, SA_FP64));
    }
    SA_TRY(ks_twosided(p, A, E->AT->A, E->V, E->W, have_init, E->res));
  } else if (E->balance != SA_BALANCE_NONE && !p.hermitian) {
    // EPSSetUp (epssetup.c:399-405): D from EPSBuildBalance_Krylov, the
    // solve on B = D A D^-1, eigenvectors x = D^-1 y normalised
    // (epsdefault.c:128-139); errors are computed with A itself.
    SA_CHECK(p.st_type == SA_ST_SHIFT && p.extraction == 0,
             "balancing runs without a spectral transformation and with Ritz extraction");
    const bool two = E->balance == SA_BALANCE_TWOSIDE;
    if (two) SA_TRY(ensure_transpose(E));
    struct DevD {
      double *p = nullptr;
      sa_mat B = nullptr;
      ~DevD() {
        if (p) (void)hipFree(p);
        if (B) sa_mat_destroy(B);
      }
    } bal;
    SA_HIP(hipMalloc(&bal.p, sizeof(double) * (size_t)std::max(1, A.n)));
    SA_TRY(eps_build_balance(A, two ? &E->AT->A : nullptr, two, E->balance_its, E->balance_cutoff, bal.p));
    std::vector<int32_t> rp, ci;
    std::vector<double> va;
    SA_TRY(download_csr(A, rp, ci, va));
    SA_TRY(sa_mat_create_csr(A.n, A.nnz, rp.data(), ci.data(), va.data(), SA_FP64, &bal.B));
    SA_TRY(balance_scale_mat(bal.B->A, bal.p));
    SA_TRY(ks_solve(p, bal.B->A, E->V, have_init, E->res));
    SA_TRY(balance_unscale_vectors(E->V, bal.p, E->res.nconv, E->res.eigi.data()));
  } else {
    SA_TRY(ks_solve(p, A, E->V, have_init, E->res));
  }
  E->solved = true;
  sa_eps_stats &s = E->st;
  s.matvecs = E->res.matvecs;
  s.gs_passes = E->res.gs_passes;
  s.matvecs32 = E->res32.matvecs;
  s.gs_passes32 = E->res32.gs_passes;
  s.its = E->res.its;
  s.its32 = E->res32.its;
  s.nconv = E->res.nconv;
  s.reason = E->res.reason;
  s.seconds = now() - t0;
  s.seconds32 = E->res32.seconds;
  s.seconds_ds = E->res.seconds_ds + E->res32.seconds_ds;
  s.bytes_spmv = E->res.bytes_spmv + E->res32.bytes_spmv;
  s.bytes_bv = E->res.bytes_bv + E->res32.bytes_bv;
  s.bytes_restart = E->res.bytes_restart + E->res32.bytes_restart;
  s.true_residuals = E->res.true_residuals + E->res32.true_residuals;
  s.ksp_its = E->res.ksp_its + E->res32.ksp_its;
  return 0;
}

int sa_eps_get_converged(sa_eps E, int *nconv) {
  SA_CHECK(E && nconv && E->solved, "EPS not solved");
  *nconv = E->res.nconv;
  return 0;
}

int sa_eps_get_iteration_number(sa_eps E, int *its) {
  SA_CHECK(E && its && E->solved, "EPS not solved");
  *its = E->res.its;
  return 0;
}

int sa_eps_get_converged_reason(sa_eps E, int *reason) {
  SA_CHECK(E && reason && E->solved, "EPS not solved");
  *reason = E->res.reason;
  return 0;
}

int sa_eps_set_orthogonalization(sa_eps E, int refine, double eta) {
  SA_CHECK(E && refine >= SA_REFINE_IFNEEDED && refine <= SA_REFINE_ALWAYS, "invalid refinement type");
  E->p.refine = refine;
  if (eta > 0) E->p.eta = eta;
  return 0;
}

int sa_eps_monitor_set(sa_eps E, sa_eps_monitor_fn fn, void *ctx) {
  SA_CHECK(E && fn, "null EPS or monitor");
  SA_CHECK(E->monitors.size() < 5, "Too many EPS monitors set");
  E->monitors.emplace_back(fn, ctx);
  return 0;
}

int sa_eps_monitor_cancel(sa_eps E) {
  SA_CHECK(E, "null EPS");
  E->monitors.clear();
  return 0;
}

int sa_eps_set_stopping_test_function(sa_eps E, sa_eps_stopping_fn fn, void *ctx) {
  SA_CHECK(E, "null EPS");
  E->stopping = fn;
  E->stopping_ctx = ctx;
  return 0;
}

}  // extern "C"

// EPS_GetEigenvector (epssolve.c): real part in column k, imaginary part in
// column k+1 for a pair (k of the sorted index i)
static int basis_eigvec(BV &B, const KsResult &r, int i, double *xr, double *xi) {
  const int k = r.perm[i];
  if (r.eigi[k] == 0.0) {
    if (xr) SA_TRY(bv_download_column(B, k, xr));
    if (xi) std::memset(xi, 0, sizeof(double) * B.n);
  } else if (r.eigi[k] > 0) {
    if (xr) SA_TRY(bv_download_column(B, k, xr));
    if (xi) SA_TRY(bv_download_column(B, k + 1, xi));
  } else {
    if (xr) SA_TRY(bv_download_column(B, k - 1, xr));
    if (xi) {
      SA_TRY(bv_download_column(B, k, xi));
      for (int64_t q = 0; q < B.n; q++) xi[q] = -xi[q];
    }
  }
  return 0;
}

extern "C" {

int sa_eps_get_eigenpair(sa_eps E, int i, double *re, double *im, double *xr, double *xi) {
  API_INIT();
  SA_CHECK(E && E->solved && i >= 0 && i < E->res.nconv, "eigenpair index out of range");
  const KsResult &r = E->res;
  const int k = r.perm[i];
  if (re) *re = r.eigr[k];
  if (im) *im = r.eigi[k];
  if (xr || xi) SA_TRY(basis_eigvec(E->V, r, i, xr, xi));
  return 0;
}

int sa_eps_get_left_eigenvector(sa_eps E, int i, double *yr, double *yi) {
  API_INIT();
  SA_CHECK(E && E->solved && i >= 0 && i < E->res.nconv, "eigenpair index out of range");
  SA_CHECK(E->twosided, "Must request left vectors with sa_eps_set_two_sided");
  return basis_eigvec(E->W, E->res, i, yr, yi);
}

int sa_eps_get_error_estimate(sa_eps E, int i, double *errest) {
  SA_CHECK(E && E->solved && errest && i >= 0 && i < E->res.nconv, "eigenpair index out of range");
  *errest = E->res.errest[E->res.perm[i]];
  return 0;
}

// EPSComputeError (epssolve.c:742-815): ||A x - lambda x|| / |lambda| on the
// device in FP64, by the fused SpMV residual kernel on the basis columns that
// hold the vector (real part, then imaginary part for a complex pair).
int sa_eps_compute_error(sa_eps E, int i, int type, double *error) {
  API_INIT();
  SA_CHECK(E && E->solved && error && i >= 0 && i < E->res.nconv, "eigenpair index out of range");
  SA_CHECK(type == SA_ERROR_ABSOLUTE || type == SA_ERROR_RELATIVE || type == SA_ERROR_BACKWARD,
           "unsupported error type");
  const KsResult &r = E->res;
  const int k = r.perm[i];
  hipStream_t s = stream();
  if (!E->rpart)
    SA_HIP(hipMalloc((void **)&E->rpart, sizeof(double) * ((size_t)RESID_MAXVEC * RESID_MAXG + RESID_MAXVEC)));
  double *part = E->rpart, *out = E->rpart + (size_t)RESID_MAXVEC * RESID_MAXG;
  const double kr = r.eigr[k], ki = r.eigi[k];
  // a pair lives in columns (c, c+1) with the positive imaginary part at c
  const int c = (ki < 0.0) ? k - 1 : k;
  const double lr[2] = {kr, kr}, li[2] = {std::fabs(ki), -std::fabs(ki)};
  const int nx = ki != 0.0 ? 2 : 1;
  double rn[2] = {0.0, 0.0};
  SA_TRY(residual_norms(E->A->A, col_ptr(E->V, c), E->V.ld, nx, lr, ki != 0.0 ? li : nullptr, rn, part, out, s));
  double res = rn[0];
  // two-sided: the left residual ||A^T y - conj(lambda) y|| too, and the
  // larger of the two (epssolve.c:777-782, EPSComputeResidualNorm_Private
  // with trans = TRUE: the conjugate value on A^T)
  if (E->twosided) {
    const double lil[2] = {-std::fabs(ki), std::fabs(ki)};
    SA_TRY(residual_norms(E->AT->A, col_ptr(E->W, c), E->W.ld, nx, lr, ki != 0.0 ? lil : nullptr, rn, part, out, s));
    res = std::max(res, rn[0]);
  }
  if (type == SA_ERROR_RELATIVE) res /= std::hypot(kr, ki);
  if (type == SA_ERROR_BACKWARD) {   // epssolve.c:790-809: ||r|| / (||A|| + |lambda| ||B||), B = I here
    if (E->nrma == 0.0) SA_TRY(mat_norm_inf(E->A->A, &E->nrma));
    E->nrmb = 1.0;
    res /= E->nrma + std::hypot(kr, ki) * E->nrmb;
  }
  *error = res;
  return 0;
}

int sa_eps_set_convergence_test(sa_eps E, int conv) {
  SA_CHECK(E && conv >= SA_CONV_ABS && conv <= SA_CONV_USER, "invalid convergence test");
  SA_CHECK(conv != SA_CONV_USER || E->converged, "Must call EPSSetConvergenceTestFunction() first");
  E->p.conv = conv;
  E->solved = false;
  return 0;
}

int sa_eps_set_convergence_test_function(sa_eps E, sa_eps_converged_fn fn, void *ctx) {
  SA_CHECK(E, "null EPS");
  E->converged = fn;
  E->converged_ctx = fn ? ctx : nullptr;
  E->p.conv = fn ? SA_CONV_USER : SA_CONV_REL;
  E->solved = false;
  return 0;
}

int sa_eps_get_matrix_norms(sa_eps E, double *nrma, double *nrmb) {
  SA_CHECK(E, "null EPS");
  if (nrma) *nrma = E->nrma;
  if (nrmb) *nrmb = E->nrmb;
  return 0;
}

// EPSComputeResidualNorm_Private over k device columns (residual partials in
// memory allocated for the call).
int sa_mat_residual_norms(sa_mat M, int k, const void *X, int64_t ldx, const double *lr, const double *li,
                          double *res) {
  API_INIT();
  SA_CHECK(M && X && lr && res && k >= 0, "null argument to sa_mat_residual_norms");
  SA_CHECK(ldx >= M->A.n, "leading dimension of X smaller than the matrix size");
  if (k == 0) return 0;
  double *part = nullptr;
  SA_HIP(hipMalloc((void **)&part, sizeof(double) * ((size_t)RESID_MAXVEC * RESID_MAXG + RESID_MAXVEC)));
  const int rc = residual_norms(M->A, X, ldx, k, lr, li, res, part, part + (size_t)RESID_MAXVEC * RESID_MAXG, stream());
  (void)hipFree(part);
  return rc;
}

// The same with the partial sums in the library's static device scratch: no
// allocation; at most RESID_MAXVEC vectors (one launch) per call.
int sa_mat_residual_norms_noalloc(sa_mat M, int k, const void *X, int64_t ldx, const double *lr, const double *li,
                                  double *res) {
  API_INIT();
  SA_CHECK(M && X && lr && res && k >= 0, "null argument to sa_mat_residual_norms_noalloc");
  SA_CHECK(ldx >= M->A.n, "leading dimension of X smaller than the matrix size");
  SA_CHECK(k <= RESID_MAXVEC, "sa_mat_residual_norms_noalloc takes at most 16 vectors per call");
  if (k == 0) return 0;
  double *part = noalloc_scratch(g_device);
  SA_CHECK(part != nullptr, "the residual scratch of this device is not set up (sa_init)");
  return residual_norms(M->A, X, ldx, k, lr, li, res, part, part + (size_t)RESID_MAXVEC * RESID_MAXG, stream());
}

int sa_eps_get_stats(sa_eps E, sa_eps_stats *st) {
  SA_CHECK(E && st && E->solved, "EPS not solved");
  *st = E->st;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- timing
namespace {
thread_local hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
}
extern "C" {
// HIP events recorded on the library stream around a timed region.
int sa_timer_start(void) {
  API_INIT();
  if (!g_ev0) {
    SA_HIP(hipEventCreate(&g_ev0));
    SA_HIP(hipEventCreate(&g_ev1));
  }
  SA_HIP(hipEventRecord(g_ev0, stream()));
  return 0;
}
int sa_timer_stop(double *ms) {
  API_INIT();
  SA_CHECK(g_ev0 && ms, "timer not started");
  SA_HIP(hipEventRecord(g_ev1, stream()));
  SA_HIP(hipEventSynchronize(g_ev1));
  float f = 0.f;
  SA_HIP(hipEventElapsedTime(&f, g_ev0, g_ev1));
  *ms = f;
  return 0;
}
}  // extern "C"

// ---------------------------------------------------------------- DS entry
extern "C" {
// DSSolve + DSSort of a compact DSHEP problem (dshep.c:221-347,383-426): diagonal
// d[0..n), off-diagonal/arrow e[0..n-1), locked rows l, arrowhead up to row k.
int sa_ds_hep_compact(int n, int l, int k, const double *d, const double *e, int which, double *wr,
                      double *Q) {
  SA_CHECK(n > 0 && l >= 0 && l <= n && k >= 0 && k <= n && d && e && wr, "invalid DS arguments");
  SA_CHECK(lapack_ready(), "no LAPACK library found for the projected problem");
  DenseProblem ds;
  const int ld = n + 1;
  ds.allocate(ld, true, which);
  std::memcpy(ds.d(), d, sizeof(double) * n);
  std::memcpy(ds.e(), e, sizeof(double) * (n - 1));
  ds.set_dimensions(n, l, k);
  ds.set_state((l == 0 && k == 0) ? DS_INTERMEDIATE : DS_RAW);
  std::vector<double> wi(ld, 0.0);
  SA_CHECK(ds.solve(wr, wi.data()) == 0, "steqr failed");
  ds.sort(wr, wi.data());
  if (Q) std::memcpy(Q, ds.Q(), sizeof(double) * (size_t)ld * ld);
  return 0;
}
// Which LAPACK library the projected problem uses (path into buf).
int sa_lapack_info(char *buf, int len) {
  std::string p;
  const bool ok = lapack_ready(&p);
  if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", p.c_str());
  return ok ? 0 : 1;
}
}  // extern "C"

extern "C" {
// Total Gram-Schmidt passes run on this basis so far (device counter).
int sa_bv_get_pass_count(sa_bv B, int64_t *passes) {
  SA_CHECK(B && passes, "null argument");
  *passes = B->V.passes_done;
  return 0;
}
}  // extern "C"
