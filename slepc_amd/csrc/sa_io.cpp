// File formats at the boundary: PETSc binary matrices and vectors (what
// MatLoad/VecLoad read and VecView writes with a binary viewer) and the
// text forms of EPSValuesView (src/eps/interface/epsview.c:530-640).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sa_internal.h"

namespace {

constexpr int32_t MAT_FILE_CLASSID = 1211216;
constexpr int32_t VEC_FILE_CLASSID = 1211214;

// PETSc binary files are big-endian.
int32_t be32(const unsigned char *p) {
  return (int32_t)((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3]);
}
double be64f(const unsigned char *p) {
  uint64_t u = 0;
  for (int i = 0; i < 8; i++) u = u << 8 | p[i];
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
void put32(std::vector<unsigned char> &b, int32_t v) {
  for (int i = 3; i >= 0; i--) b.push_back((unsigned char)((uint32_t)v >> (8 * i)));
}
void put64f(std::vector<unsigned char> &b, double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  for (int i = 7; i >= 0; i--) b.push_back((unsigned char)(u >> (8 * i)));
}

bool read_all(const char *path, std::vector<unsigned char> &buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  unsigned char tmp[1 << 16];
  size_t r;
  while ((r = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
  std::fclose(f);
  return true;
}

}  // namespace

using namespace sa;

// MatLoad of a PETSc binary AIJ file (the loader behind ex4.c and the
// datafiles/matrices/*.petsc fixtures): header {classid, M, N, nz}, M row
// lengths, nz column indices, nz values.  Rows keep the file's column order.
extern "C" int sa_mat_load_petsc_binary(const char *path, int *nrows, int *ncols, int64_t *nnz, int32_t *rowptr,
                                        int32_t *col, double *val) {
  SA_CHECK(path && nrows && ncols && nnz, "null argument");
  std::vector<unsigned char> b;
  SA_CHECK(read_all(path, b), std::string("cannot open ") + path);
  SA_CHECK(b.size() >= 16, "file too short for a PETSc binary header");
  SA_CHECK(be32(&b[0]) == MAT_FILE_CLASSID, "not a PETSc binary matrix (class id)");
  const int32_t M = be32(&b[4]), N = be32(&b[8]), nz = be32(&b[12]);
  SA_CHECK(M >= 0 && N >= 0 && nz >= 0, "negative sizes in the header");
  const size_t need = 16 + 4 * ((size_t)M + (size_t)nz) + 8 * (size_t)nz;
  SA_CHECK(b.size() >= need, "truncated PETSc binary matrix");
  *nrows = M;
  *ncols = N;
  *nnz = nz;
  if (!rowptr && !col && !val) return 0;
  SA_CHECK(rowptr && col && val, "rowptr, col and val must all be given");
  const unsigned char *p = &b[16];
  int64_t acc = 0;
  rowptr[0] = 0;
  for (int32_t i = 0; i < M; i++) {
    const int32_t len = be32(p + 4 * (size_t)i);
    SA_CHECK(len >= 0, "negative row length");
    acc += len;
    SA_CHECK(acc <= nz, "row lengths exceed nz");
    rowptr[i + 1] = (int32_t)acc;
  }
  SA_CHECK(acc == nz, "row lengths do not add up to nz");
  p += 4 * (size_t)M;
  for (int32_t k = 0; k < nz; k++) {
    const int32_t c = be32(p + 4 * (size_t)k);
    SA_CHECK(c >= 0 && c < N, "column index out of range");
    col[k] = c;
  }
  p += 4 * (size_t)nz;
  for (int32_t k = 0; k < nz; k++) val[k] = be64f(p + 8 * (size_t)k);
  return 0;
}

// VecView with a binary viewer: {classid, n} then n values.  append != 0 adds
// the vector after those already in the file (EPSVectorsView writes one
// vector after the other).
extern "C" int sa_vec_write_petsc_binary(const char *path, int64_t n, const double *x, int append) {
  SA_CHECK(path && (x || n == 0) && n >= 0 && n <= INT32_MAX, "invalid arguments");
  std::vector<unsigned char> b;
  b.reserve(8 + 8 * (size_t)n);
  put32(b, VEC_FILE_CLASSID);
  put32(b, (int32_t)n);
  for (int64_t i = 0; i < n; i++) put64f(b, x[i]);
  FILE *f = std::fopen(path, append ? "ab" : "wb");
  SA_CHECK(f, std::string("cannot open ") + path);
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
  SA_CHECK(w == b.size(), "short write");
  return 0;
}

// VecLoad of the index-th vector of a PETSc binary file; with x == NULL only
// *n is returned.
extern "C" int sa_vec_load_petsc_binary(const char *path, int64_t index, int64_t *n, double *x) {
  SA_CHECK(path && n && index >= 0, "invalid arguments");
  std::vector<unsigned char> b;
  SA_CHECK(read_all(path, b), std::string("cannot open ") + path);
  size_t off = 0;
  for (int64_t v = 0;; v++) {
    SA_CHECK(b.size() >= off + 8, "vector index beyond the end of the file");
    SA_CHECK(be32(&b[off]) == VEC_FILE_CLASSID, "not a PETSc binary vector (class id)");
    const int32_t len = be32(&b[off + 4]);
    SA_CHECK(len >= 0 && b.size() >= off + 8 + 8 * (size_t)len, "truncated PETSc binary vector");
    if (v == index) {
      *n = len;
      if (x)
        for (int32_t i = 0; i < len; i++) x[i] = be64f(&b[off + 8 + 8 * (size_t)i]);
      return 0;
    }
    off += 8 + 8 * (size_t)len;
  }
}

// The text of EPSValuesView for eigenvalues (re[i], im[i]):
//   format 0 (ASCII, epsview.c:602-618 with SlepcPrintEigenvalueASCII,
//   slepcimpl.h:92-112): "Eigenvalues = " then "   %.5f" or "   %.5f%+.5fi"
//   per value (a part negligible against the other, ratio below 1e-10 in
//   double precision builds, prints as zero) and an empty line;
//   format 1 (MATLAB, epsview.c:620-640): "Lambda_<name> = [", "%18.16e" or
//   "%18.16e%+18.16ei" per value, "];".
// Returns the length of the full text; writes at most len-1 characters and a
// terminating zero into buf when buf is not NULL.
extern "C" int64_t sa_format_eigenvalues(int n, const double *re, const double *im, int format, const char *name,
                                         char *buf, int64_t len) {
  if (n < 0 || (n > 0 && !re) || (format != 0 && format != 1)) {
    set_error("invalid arguments");
    return -1;
  }
  std::string s;
  char line[128];
  const double small = 1e-10;    // PETSC_SMALL in double precision
  if (format == 0) {
    s += "Eigenvalues = \n";
    for (int i = 0; i < n; i++) {
      double r = re[i], m = im ? im[i] : 0.0;
      if (std::fabs(m) != 0.0 && std::fabs(r) / std::fabs(m) < small) r = 0.0;
      if (std::fabs(r) != 0.0 && std::fabs(m) / std::fabs(r) < small) m = 0.0;
      if (m != 0.0) std::snprintf(line, sizeof line, "   %.5f%+.5fi\n", r, m);
      else std::snprintf(line, sizeof line, "   %.5f\n", r);
      s += line;
    }
    s += "\n";
  } else {
    s += "Lambda_";
    s += name ? name : "";
    s += " = [\n";
    for (int i = 0; i < n; i++) {
      const double r = re[i], m = im ? im[i] : 0.0;
      if (m != 0.0) std::snprintf(line, sizeof line, "%18.16e%+18.16ei\n", r, m);
      else std::snprintf(line, sizeof line, "%18.16e\n", r);
      s += line;
    }
    s += "];\n";
  }
  if (buf && len > 0) {
    const size_t k = std::min<size_t>(s.size(), (size_t)(len - 1));
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return (int64_t)s.size();
}
