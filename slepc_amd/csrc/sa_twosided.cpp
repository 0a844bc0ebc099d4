// Two-sided Krylov-Schur (src/eps/impls/krylov/krylovschur/ks-twosided.c:
// 126-241; Zwaan and Hochstenbach, SIAM J. Matrix Anal. Appl. 38(2), 2017):
// Arnoldi sweeps for A on the right basis V and for A^T on the left basis W,
// both on the fused sweep kernels; the projected pencil kept in the DSNHEPTS
// problem (A for V, B for W) and the oblique projection M = W^T V on the
// host, updated by BVDot blocks on the device.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "sa_ds.h"
#include "sa_eps.h"

namespace sa {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// q_host[0:nv] = B(:, 0:nv)^T x (BVDotVec with active columns 0..nv)
int dotvec_host(BV &B, int nv, const void *x, std::vector<double> &q, hipStream_t s) {
  SA_TRY(launch_dotvec(B, 0, nv, x, B.dq, s));
  q.resize(nv);
  SA_HIP(hipMemcpyAsync(q.data(), B.dq, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

// B(:, j) -= B(:, 0:nv) q (BVMultColumn(B, -1, 1, j, q))
int multcol_host(BV &B, int nv, int j, const std::vector<double> &q, hipStream_t s) {
  std::memcpy(B.hq, q.data(), sizeof(double) * nv);
  SA_HIP(hipMemcpyAsync(B.dq, B.hq, sizeof(double) * nv, hipMemcpyHostToDevice, s));
  SA_TRY(launch_multvec(B, 0, nv, -1.0, B.dq, 1.0, col_ptr(B, j), s));
  SA_HIP(hipStreamSynchronize(s));   // the staging buffer is reused next
  return 0;
}

double column_norm(BV &B, int j, hipStream_t s) {
  double h = 0.0;
  if (launch_dot_self(B.prec, col_ptr(B, j), B.n, B.dq, B.ws.part, s) != 0) return -1.0;
  if (hipMemcpyAsync(&h, B.dq, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return -1.0;
  if (hipStreamSynchronize(s) != hipSuccess) return -1.0;
  return std::sqrt(h);
}

// The Hessenberg matrix of columns k0..nv-1 of a sweep (with the extra row)
// from the coefficient buffer of B (BVMatArnoldi copy-out, bvkrylov.c:100-109).
int fill_hessenberg(BV &B, int k0, int nv, double *H, int ld, std::vector<double> &buf, hipStream_t s) {
  if (nv <= k0) return 0;
  const int m = B.m;
  buf.resize((size_t)m * (nv - k0));
  SA_HIP(hipMemcpyAsync(buf.data(), B.dbuf + (size_t)(k0 + 1) * m, sizeof(double) * m * (nv - k0),
                        hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  auto C = [&](int row, int colc) { return buf[(size_t)(colc - k0 - 1) * m + row]; };
  for (int j = k0; j < nv - 1; j++)
    for (int i = 0; i < j + 2; i++) H[i + (size_t)j * ld] = C(i, j + 1);
  for (int i = 0; i < nv; i++) H[i + (size_t)(nv - 1) * ld] = C(i, nv);
  if (ld > nv) H[nv + (size_t)(nv - 1) * ld] = C(nv, nv);
  return 0;
}

// V(:, s:e) = V(:, nconv:nv) U(nconv:nv, s:e) with U in the staging buffer
int restart_gemm(BV &B, const double *U, int ld, int nconv, int nv, int e, hipStream_t s) {
  if (e <= nconv) return 0;
  std::memcpy(B.hq, U, sizeof(double) * (size_t)ld * ld);
  SA_HIP(hipMemcpyAsync(B.dq, B.hq, sizeof(double) * (size_t)ld * ld, hipMemcpyHostToDevice, s));
  SA_TRY(launch_multinplace(B, nconv, nv, nconv, e, B.dq, ld, s));
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

int ks_twosided(const KsParams &p, const Mat &A, const Mat &AT, BV &V, BV &W, bool have_initial, KsResult &r) {
  SA_CHECK(lapack_ready(), "no LAPACK library found for the projected problem");
  const int n = A.n, ncv = p.ncv, nev = p.nev, mpd = p.mpd, ld = ncv + 1;
  SA_CHECK(V.m == ncv + 1 && W.m == ncv + 1 && V.n == n && W.n == n && AT.n == n, "bases do not match the operator");
  SA_CHECK(V.prec == SA_FP64 && W.prec == SA_FP64 && A.prec == SA_FP64 && AT.prec == SA_FP64,
           "the two-sided solver runs in FP64");
  SA_CHECK(!p.hermitian, "Two-sided methods are not intended for symmetric problems");
  SA_CHECK(p.solver == 0 && p.extraction == 0, "two-sided runs with Krylov-Schur and Ritz extraction");
  SA_CHECK(ncv + 1 < GS_MAXC, "ncv too large for the fused Gram-Schmidt kernels");
  V.refine = W.refine = p.refine;   // BVSetOrthogonalization of both bases
  hipStream_t s = stream();
  DenseProblem ds;
  ds.allocate(ld, false, p.which);
  ds.set_target(p.target);
  ds.enable_two_sided();
  std::vector<double> eigr(ld, 0.0), eigi(ld, 0.0), errest(ld, 0.0), buf, scratch, w, c, Mlu;
  std::vector<double> M((size_t)ncv * ncv, 0.0);   // W^T V (BVMatProject), leading dimension ncv
  std::vector<int> ipiv(ld);
  double *Mdev = nullptr;
  SA_HIP(hipMalloc(&Mdev, sizeof(double) * (size_t)ncv * (ncv + 1)));
  struct Free {
    double *p;
    ~Free() { (void)hipFree(p); }
  } free_m{Mdev};
  struct TrueWork {   // EPSSetTrueResidual: Ritz vector columns and residual partials
    void *x = nullptr;
    double *part = nullptr;
    ~TrueWork() {
      if (x) (void)hipFree(x);
      if (part) (void)hipFree(part);
    }
  } trw;
  if (p.trueres) {
    SA_HIP(hipMalloc(&trw.x, sizeof(double) * (size_t)V.ld * 2));
    SA_HIP(hipMalloc((void **)&trw.part, sizeof(double) * ((size_t)RESID_MAXVEC * RESID_MAXG + RESID_MAXVEC)));
  }
  const double t0 = now_s();
  double tds = 0.0;
  r = KsResult();
  const int64_t passes0 = V.passes_done + W.passes_done;

  if (have_initial) {   // BVInsertVecs(V,0,&k,IS,TRUE)
    double norm = 0.0;
    int lindep = 0;
    SA_TRY(bv_orthonormalize(V, 0, p.eta, &norm, &lindep));
    SA_CHECK(!(norm == 0.0 || lindep), "initial vector is linearly dependent");
  }
  SA_TRY(eps_start_vector(V, 0, have_initial, p.eta, nullptr, scratch, EPS_SEED_RIGHT));
  SA_TRY(eps_start_vector(W, 0, false, p.eta, nullptr, scratch, EPS_SEED_LEFT));   // EPSGetLeftStartVector

  int l = 0, nconv = 0, its = 0, reason = 0;
  while (reason == 0) {
    its++;
    int nv = std::min(nconv + mpd, ncv);
    const int k0 = nconv + l;
    ds.set_dimensions(nv, nconv, k0);
    double beta = 0.0, betat = 0.0;
    int bd = 0, bdt = 0;
    SA_TRY(bv_sweep(V, A, k0, &nv, p.eta, &beta, &bd));
    int nvt = nv;
    SA_TRY(bv_sweep(W, AT, k0, &nvt, p.eta, &betat, &bdt));
    r.matvecs += (nv - k0) + (nvt - k0);
    nv = std::min(nv, nvt);
    const double td0 = now_s();
    SA_TRY(fill_hessenberg(V, k0, nv, ds.A(), ld, buf, s));
    SA_TRY(fill_hessenberg(W, k0, nv, ds.B(), ld, buf, s));
    ds.set_dimensions(nv, nconv, k0);
    ds.set_state(l == 0 ? DS_INTERMEDIATE : DS_RAW);
    const bool breakdown = bd || bdt;

    // BVMatProject(V,NULL,W,M) with active columns [k0, nv) (bvglobal.c:1013-1049)
    if (k0 > 0 && k0 < nv) {   // top right: W(:,0:k0)^T V(:,k0:nv)
      SA_TRY(launch_bvdot(V, k0, nv, W, 0, k0, Mdev, s));
      buf.resize((size_t)k0 * (nv - k0));
      SA_HIP(hipMemcpy(buf.data(), Mdev, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
      for (int j = k0; j < nv; j++)
        for (int i = 0; i < k0; i++) M[i + (size_t)j * ncv] = buf[i + (size_t)(j - k0) * k0];
    }
    if (nv > 0 && k0 < nv) {   // bottom: W(:,k0:nv)^T V(:,0:nv)
      SA_TRY(launch_bvdot(V, 0, nv, W, k0, nv, Mdev, s));
      const int my = nv - k0;
      buf.resize((size_t)my * nv);
      SA_HIP(hipMemcpy(buf.data(), Mdev, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
      for (int j = 0; j < nv; j++)
        for (int i = 0; i < my; i++) M[(k0 + i) + (size_t)j * ncv] = buf[i + (size_t)j * my];
    }

    // EPSTwoSidedRQUpdate1 (ks-twosided.c:27-73): oblique corrections of the
    // residual vectors and of the last columns of the Rayleigh quotients
    Mlu.assign(M.begin(), M.end());
    SA_CHECK(la_getrf(nv, Mlu.data(), ncv, ipiv.data()) == 0, "LAPACK getrf failed on W^T V");
    SA_TRY(dotvec_host(W, nv, col_ptr(V, nv), w, s));
    SA_CHECK(la_getrs('N', nv, Mlu.data(), ncv, ipiv.data(), w.data()) == 0, "LAPACK getrs failed");
    SA_TRY(multcol_host(V, nv, nv, w, s));
    for (int i = 0; i < nv; i++) ds.A()[(size_t)(nv - 1) * ld + i] += beta * w[i];
    SA_TRY(dotvec_host(V, nv, col_ptr(W, nv), w, s));
    SA_CHECK(la_getrs('C', nv, Mlu.data(), ncv, ipiv.data(), w.data()) == 0, "LAPACK getrs failed");
    SA_TRY(multcol_host(W, nv, nv, w, s));
    for (int i = 0; i < nv; i++) ds.B()[(size_t)(nv - 1) * ld + i] += betat * w[i];

    int info = ds.solve_ts(eigr.data(), eigi.data());
    SA_CHECK(info == 0, "LAPACK failure in the projected problem (solve)");
    info = ds.sort_ts(eigr.data(), eigi.data());
    SA_CHECK(info == 0, "LAPACK failure in the projected problem (sort)");
    ds.update_extra_row_ts();

    // EPSKrylovConvergence with right and left bounds (epskrylov.c:207-295)
    const double norm = column_norm(V, nv, s), norm2 = column_norm(W, nv, s);
    SA_CHECK(norm >= 0.0 && norm2 >= 0.0, "HIP error in a column norm");
    int k = nconv, marker = -1;
    for (k = nconv; k < nv; k++) {
      double rn = 0.0, rl = 0.0;
      int cplx = 0, cplx2 = 0;
      SA_CHECK(ds.ritz_residual(k, &rn, &cplx) == 0, "LAPACK failure in the projected problem (vectors)");
      if (p.trueres) {   // the right residual from the Ritz vector (epskrylov.c:256-264)
        SA_TRY(ritz_true_residual(V, A, nv, ds.X() + (size_t)k * ld, ld, cplx ? 2 : 1, eigr[k], eigi[k], trw.x, trw.part,
                                  &rn));
        r.true_residuals++;
      } else {
        rn = rn * beta * norm;
      }
      SA_TRY(eps_converged(p, eigr[k], eigi[k], rn, &errest[k]));
      if (marker == -1 && errest[k] >= p.tol) marker = k;
      SA_CHECK(ds.left_residual(k, &rl, &cplx2) == 0, "LAPACK failure in the projected problem (left vectors)");
      double lerr = 0.0;
      SA_TRY(eps_converged(p, eigr[k], eigi[k], rl * betat * norm2, &lerr));
      errest[k] = std::max(errest[k], lerr);
      if (marker == -1 && lerr >= p.tol) marker = k;
      if (cplx) {
        errest[k + 1] = errest[k];
        k++;
      }
      if (marker != -1) break;
    }
    if (marker != -1) k = marker;
    SA_TRY(eps_stopping(p, its, k, nev, &reason));
    int newl = 0;
    if (!(reason != 0 || breakdown || k == nv)) {
      newl = std::max(1, (int)((nv - k) * p.keep));
      newl = ds.truncate_size_ts(k, nv, newl);
    }
    l = newl;
    if (!p.lock && l > 0) {   // non-locking variant (ks-twosided.c:195)
      l += k;
      k = 0;
    }
    tds += now_s() - td0;

    // V(:,nconv:k+l) = V Q, W(:,nconv:k+l) = W Z
    SA_TRY(restart_gemm(V, ds.Q(), ld, nconv, nv, k + l, s));
    SA_TRY(restart_gemm(W, ds.Z(), ld, nconv, nv, k + l, s));
    r.bytes_restart += 2.0 * ((double)(nv - nconv) + (double)std::max(0, k + l - nconv)) * V.n * 8.0;
    if (reason == 0 && !breakdown) {
      SA_HIP(hipMemcpyAsync(col_ptr(V, k + l), col_ptr(V, nv), 8 * V.n, hipMemcpyDeviceToDevice, s));
      SA_HIP(hipMemcpyAsync(col_ptr(W, k + l), col_ptr(W, nv), 8 * W.n, hipMemcpyDeviceToDevice, s));
    }
    if (reason == 0) {
      if (breakdown || k == nv) {
        if (k < nev) {
          int b1 = 0, b2 = 0;
          SA_TRY(eps_start_vector(V, k, false, p.eta, &b1, scratch, EPS_SEED_RIGHT));
          SA_TRY(eps_start_vector(W, k, false, p.eta, &b2, scratch, EPS_SEED_LEFT));
          if (b1 || b2) reason = -2;
        }
      } else {
        ds.set_dimensions(ds.n(), k, ds.k());
        ds.truncate_ts(k + l);
      }
      // EPSTwoSidedRQUpdate2 (ks-twosided.c:75-124): orthonormalise the new
      // residual columns and fold their coefficients into the extra rows,
      // then M = Z^T M Q
      const int kk = k + l;
      double *Ab[2] = {ds.A(), ds.B()};
      BV *Bs[2] = {&V, &W};
      for (int side = 0; side < 2; side++) {
        BV &B = *Bs[side];
        double nrm = 0.0;
        int lind = 0;
        SA_TRY(bv_orthonormalize(B, kk, p.eta, &nrm, &lind));
        c.resize(B.m);
        SA_HIP(hipMemcpy(c.data(), B.dbuf + (size_t)kk * B.m, sizeof(double) * B.m, hipMemcpyDeviceToHost));
        double *H = Ab[side];
        for (int j = nconv; j < kk; j++) {
          for (int i = 0; i < kk; i++) H[i + (size_t)j * ld] += c[i] * H[kk + (size_t)j * ld];
          H[kk + (size_t)j * ld] *= nrm;
        }
      }
      std::vector<double> T1((size_t)nv * nv, 0.0);
      const double *Q = ds.Q(), *Z = ds.Z();
      for (int j = 0; j < nv; j++)          // T1 = M Q
        for (int q = 0; q < nv; q++) {
          const double qv = Q[q + (size_t)j * ld];
          if (qv == 0.0) continue;
          for (int i = 0; i < nv; i++) T1[i + (size_t)j * nv] += M[i + (size_t)q * ncv] * qv;
        }
      for (int j = 0; j < nv; j++)          // M = Z^T T1
        for (int i = 0; i < nv; i++) {
          double acc = 0.0;
          for (int q = 0; q < nv; q++) acc += Z[q + (size_t)i * ld] * T1[q + (size_t)j * nv];
          M[i + (size_t)j * ncv] = acc;
        }
    }
    nconv = k;
    SA_TRY(eps_monitor(p, its, nconv, eigr.data(), eigi.data(), errest.data(), nv));
  }

  // DSTruncate(nconv,TRUE); EPSComputeVectors_Schur (epsdefault.c:105-169):
  // V = V X, W = W Y, W normalised (a conjugate pair by its joint norm) and
  // the second column of each pair with positive imaginary part negated.
  // EPSSolve computes them at the first pair it reorders (epssolve.c:159-176),
  // so that pair's sign already counts as positive.
  const double td1 = now_s();
  SA_CHECK(ds.trim_and_vectors_ts(nconv) == 0, "LAPACK failure computing eigenvectors of the projected problem");
  tds += now_s() - td1;
  bool computed = false;
  auto compute_vectors = [&]() -> int {
    if (nconv <= 0) return 0;
    SA_TRY(restart_gemm(V, ds.X(), ld, 0, nconv, nconv, s));
    SA_TRY(restart_gemm(W, ds.Y(), ld, 0, nconv, nconv, s));
    for (int i = 0; i < nconv; i++) {
      const bool pair = eigi[i] != 0.0 && i + 1 < nconv;
      double nr = column_norm(W, i, s);
      if (pair) nr = std::hypot(nr, column_norm(W, i + 1, s));
      SA_CHECK(nr > 0.0, "zero left eigenvector");
      SA_TRY(launch_scale(SA_FP64, col_ptr(W, i), W.n, 1.0 / nr, s));
      if (pair) {
        SA_TRY(launch_scale(SA_FP64, col_ptr(W, i + 1), W.n, (eigi[i] > 0.0 ? -1.0 : 1.0) / nr, s));
        i++;
      }
    }
    computed = true;
    return 0;
  };
  for (int i = 0; i < nconv - 1; i++) {
    if (eigi[i] != 0) {
      if (eigi[i] < 0) {
        eigi[i] = -eigi[i];
        eigi[i + 1] = -eigi[i + 1];
        if (!computed) SA_TRY(compute_vectors());
        SA_TRY(launch_scale(V.prec, col_ptr(V, i + 1), V.n, -1.0, s));
      }
      i++;
    }
  }
  if (!computed) SA_TRY(compute_vectors());
  // SlepcSortEigenvalues (slepcsc.c:89-140): real values and pairs kept together
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
  r.seconds = now_s() - t0;
  r.seconds_ds = tds;
  r.nconv = nconv;
  r.its = its;
  r.reason = reason;
  r.gs_passes = V.passes_done + W.passes_done - passes0;
  r.perm = perm;
  r.eigr.assign(eigr.begin(), eigr.begin() + nconv);
  r.eigi.assign(eigi.begin(), eigi.begin() + nconv);
  r.errest.assign(errest.begin(), errest.begin() + nconv);
  return 0;
}

}  // namespace sa
