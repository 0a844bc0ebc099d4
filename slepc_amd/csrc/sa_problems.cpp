// CSR assembly (MatSetValues + MatAssemblyEnd semantics) and the stencil
// families of the workloads, assembled directly in AIJ order: per row the
// column indices ascend and each (row,col) appears once.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "sa_internal.h"

namespace sa {

uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double u01(uint64_t seed, uint64_t i) {
  return (double)(splitmix64(splitmix64(seed) + i) >> 11) * 0x1.0p-53;
}

namespace {

// Row-by-row emitter: count pass (arrays null) or fill pass.
struct Emit {
  int32_t *rp, *ci;
  double *v;
  int64_t nz = 0;
  void row(int64_t r) { if (rp) rp[r] = (int32_t)nz; }
  void put(int64_t c, double x) {
    if (ci) { ci[nz] = (int32_t)c; v[nz] = x; }
    nz++;
  }
  void end(int64_t n) { if (rp) rp[n] = (int32_t)nz; }
};

// Separable cosine Karhunen-Loeve modes of a Gaussian covariance (length 0.2)
// on the unit square; the eight leading (p,q) pairs other than (0,0).
const int KL_P[8] = {1, 0, 1, 2, 0, 2, 1, 2};
const int KL_Q[8] = {0, 1, 1, 0, 2, 1, 2, 2};
double kl_eig1(int p) {
  const double ell = 0.2, pi = 3.141592653589793;
  const double a = p * pi * ell;
  return std::sqrt(2.0 * pi) * ell * std::exp(-0.5 * a * a);
}
double kl_mode1(int p, double x) {
  const double pi = 3.141592653589793;
  return p == 0 ? 1.0 : std::sqrt(2.0) * std::cos(p * pi * x);
}

}  // namespace
}  // namespace sa

using namespace sa;

extern "C" {

double sa_u01(uint64_t seed, uint64_t i) { return u01(seed, i); }

uint64_t sa_sub_seed(uint64_t seed, uint64_t p) { return splitmix64(seed ^ (0xD1B54A32D192ED03ULL * (p + 1))); }

void sa_rand_uniform(uint64_t seed, int64_t n, double *out) {
  for (int64_t i = 0; i < n; i++) out[i] = 2.0 * u01(seed, (uint64_t)i) - 1.0;
}

void sa_fv2d_theta(uint64_t seed, int sample, double *theta8) {
  for (int m = 0; m < 8; m++) {
    const uint64_t i = (uint64_t)sample * 8 + m;
    const double a = 1.0 - u01(seed, 2 * i), b = u01(seed, 2 * i + 1);
    theta8[m] = std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
  }
}

int64_t sa_csr_assemble(int n, int64_t ntrip, const int32_t *rows, const int32_t *cols, const double *vals,
                        int32_t *rowptr, int32_t *col, double *val) {
  // Stable sort by (row, col): among duplicates the last given wins (INSERT_VALUES).
  std::vector<int64_t> idx(ntrip);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
  });
  int64_t nz = 0;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t q = 0; q < ntrip; q++) {
    const int64_t i = idx[q];
    if (rows[i] < 0 || rows[i] >= n || cols[i] < 0 || cols[i] >= n) return -1;
    const bool last_of_key = (q + 1 == ntrip) || rows[idx[q + 1]] != rows[i] || cols[idx[q + 1]] != cols[i];
    if (!last_of_key) continue;
    if (col) { col[nz] = cols[i]; val[nz] = vals[i]; }
    cnt[rows[i] + 1]++;
    nz++;
  }
  if (rowptr) {
    rowptr[0] = 0;
    for (int r = 0; r < n; r++) rowptr[r + 1] = (int32_t)(rowptr[r] + cnt[r + 1]);
  }
  return nz;
}

// 2-D 5-point Laplacian (ex2.c:46-53 numbering II = i*nx + j) with an optional
// diagonal potential pot_scale * U[0,1)(seed, II).
int64_t sa_gen_lap2d(int nx, int ny, double pot_scale, uint64_t seed, int32_t *rp, int32_t *ci, double *v) {
  Emit e{rp, ci, v};
  const int64_t N = (int64_t)nx * ny;
  for (int64_t r = 0; r < N; r++) {
    const int64_t i = r / nx, j = r % nx;
    e.row(r);
    if (i > 0) e.put(r - nx, -1.0);
    if (j > 0) e.put(r - 1, -1.0);
    e.put(r, pot_scale != 0.0 ? 4.0 + pot_scale * u01(seed, (uint64_t)r) : 4.0);
    if (j < nx - 1) e.put(r + 1, -1.0);
    if (i < ny - 1) e.put(r + nx, -1.0);
  }
  e.end(N);
  return e.nz;
}

// 3-D 7-point convection-diffusion: diagonal 6; the neighbour at -e_d has
// coefficient -1 - 0.1 b_d and the one at +e_d -1 + 0.1 b_d, b = (1, .5, .25).
int64_t sa_gen_convdiff3d(int nx, int ny, int nz, int32_t *rp, int32_t *ci, double *v) {
  const double b[3] = {1.0, 0.5, 0.25};
  double dn[3], up[3];
  for (int d = 0; d < 3; d++) {
    dn[d] = -1.0 - 0.1 * b[d];
    up[d] = -1.0 + 0.1 * b[d];
  }
  Emit e{rp, ci, v};
  const int64_t pl = (int64_t)nx * ny, N = pl * nz;
  for (int64_t r = 0; r < N; r++) {
    const int64_t z = r / pl, y = (r % pl) / nx, x = r % nx;
    e.row(r);
    if (z > 0) e.put(r - pl, dn[2]);
    if (y > 0) e.put(r - nx, dn[1]);
    if (x > 0) e.put(r - 1, dn[0]);
    e.put(r, 6.0);
    if (x < nx - 1) e.put(r + 1, up[0]);
    if (y < ny - 1) e.put(r + nx, up[1]);
    if (z < nz - 1) e.put(r + pl, up[2]);
  }
  e.end(N);
  return e.nz;
}

// 3-D 27-point stencil: -1 off the diagonal, 26 + U[0,1)(seed, r) on it.
int64_t sa_gen_stencil27(int nx, int ny, int nz, uint64_t seed, int32_t *rp, int32_t *ci, double *v) {
  Emit e{rp, ci, v};
  const int64_t pl = (int64_t)nx * ny, N = pl * nz;
  for (int64_t r = 0; r < N; r++) {
    const int64_t z = r / pl, y = (r % pl) / nx, x = r % nx;
    e.row(r);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (z + dz < 0 || z + dz >= nz || y + dy < 0 || y + dy >= ny || x + dx < 0 || x + dx >= nx) continue;
          const int64_t c = r + dx + (int64_t)nx * dy + pl * dz;
          e.put(c, c == r ? 26.0 + u01(seed, (uint64_t)r) : -1.0);
        }
  }
  e.end(N);
  return e.nz;
}

// Finite-volume -div(k grad u) with harmonic-mean face conductivities, homogeneous
// Dirichlet walls (half-cell distance: 2 k at the wall), cells row-major
// r = i + nx*j; returns sigma I - L with sigma = max_r sum_c |L_rc|.
int64_t sa_gen_fv2d(int nx, int ny, const double *theta, double *sigma_out, int32_t *rp, int32_t *ci, double *v) {
  const int64_t N = (int64_t)nx * ny;
  std::vector<double> kc(N);
  double w[8];
  for (int m = 0; m < 8; m++) w[m] = theta[m] * std::sqrt(kl_eig1(KL_P[m]) * kl_eig1(KL_Q[m]));
  for (int64_t r = 0; r < N; r++) {
    const int64_t j = r / nx, i = r % nx;
    const double x = (i + 0.5) / nx, y = (j + 0.5) / ny;
    double s = 0.0;
    for (int m = 0; m < 8; m++) s += w[m] * kl_mode1(KL_P[m], x) * kl_mode1(KL_Q[m], y);
    kc[r] = std::exp(s);
  }
  // face transmissibilities in the order (down, left, right, up)
  auto faces = [&](int64_t r, double t[4], bool has[4]) {
    const int64_t j = r / nx, i = r % nx;
    has[0] = j > 0; has[1] = i > 0; has[2] = i < nx - 1; has[3] = j < ny - 1;
    const int64_t nb[4] = {r - nx, r - 1, r + 1, r + nx};
    const double k0 = kc[r];
    for (int f = 0; f < 4; f++) {
      if (has[f]) {
        const double k1 = kc[nb[f]];
        t[f] = (2.0 * k0 * k1) / (k0 + k1);
      } else {
        t[f] = 2.0 * k0;
      }
    }
  };
  double sigma = 0.0;
  for (int64_t r = 0; r < N; r++) {
    double t[4];
    bool has[4];
    faces(r, t, has);
    double diag = 0.0, off = 0.0;
    for (int f = 0; f < 4; f++) {
      if (has[f]) off += t[f];
      diag += t[f];
    }
    sigma = std::max(sigma, diag + off);
  }
  if (sigma_out) *sigma_out = sigma;
  Emit e{rp, ci, v};
  for (int64_t r = 0; r < N; r++) {
    double t[4];
    bool has[4];
    faces(r, t, has);
    const double diag = ((t[0] + t[1]) + t[2]) + t[3];
    e.row(r);
    if (has[0]) e.put(r - nx, t[0]);
    if (has[1]) e.put(r - 1, t[1]);
    e.put(r, sigma - diag);
    if (has[2]) e.put(r + 1, t[2]);
    if (has[3]) e.put(r + nx, t[3]);
  }
  e.end(N);
  return e.nz;
}

// Markov walk of ex5.c:146-178 assembled through the COO path.
int64_t sa_gen_markov(int m, int32_t *rp, int32_t *ci, double *v) {
  const double cst = 0.5 / (double)(m - 1);
  const int n = m * (m + 1) / 2;
  std::vector<int32_t> R, C;
  std::vector<double> X;
  int ix = 0;
  for (int i = 1; i <= m; i++) {
    const int jmax = m - i + 1;
    for (int j = 1; j <= jmax; j++) {
      ix++;
      if (j != jmax) {
        const double pd = cst * (double)(i + j - 1);
        R.push_back(ix - 1); C.push_back(ix); X.push_back(i == 1 ? 2 * pd : pd);
        R.push_back(ix - 1); C.push_back(ix + jmax - 1); X.push_back(j == 1 ? 2 * pd : pd);
      }
      const double pu = 0.5 - cst * (double)(i + j - 3);
      if (j > 1) { R.push_back(ix - 1); C.push_back(ix - 2); X.push_back(pu); }
      if (i > 1) { R.push_back(ix - 1); C.push_back(ix - jmax - 2); X.push_back(pu); }
    }
  }
  return sa_csr_assemble(n, (int64_t)R.size(), R.data(), C.data(), X.data(), rp, ci, v);
}

}  // extern "C"
