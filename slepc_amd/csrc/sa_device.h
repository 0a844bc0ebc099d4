// Device-side constants and helpers shared by the HIP sources (gfx950,
// wave size 64, 256-thread workgroups).
#pragma once
#include <hip/hip_runtime.h>

namespace sa {

constexpr int BLK = 256;
constexpr int CH = 64;          // rows per chunk (one row per lane of a wave)
constexpr int LDP = CH + 1;     // padded LDS column stride
constexpr int RBLK = 1024;      // threads of the one-workgroup reductions

__device__ __forceinline__ double safe_sqrt(double a) { return a < 0.0 ? 0.0 : sqrt(a); }

// A load that is non-temporal when NT (data read once, larger than the
// caches: streaming it past the last-level cache keeps the caches for reused
// data and reads faster) and ordinary otherwise.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// The dispatcher hands workgroups to the 8 XCDs round-robin by index, so
// consecutive blocks land on different L2s.  xcd_block maps the dispatch
// index b of a G-block grid to a logical block such that each XCD takes one
// contiguous range of about G/8 blocks (a bijection for any G): row-blocked
// kernels whose rows share neighbours (stencil planes) then find them in
// their own XCD's L2.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int q = G / N_XCD, r = G % N_XCD, x = b % N_XCD, i = b / N_XCD;
  return x * q + min(x, r) + i;
}

// out[col] = sum_b part[col * G + b] for col < ncol.  Each wave reduces four
// columns at a time (columns w, w+nw, w+2nw, w+3nw) with four loads per column
// in flight, then a shuffle reduction; the order of the additions is fixed, so
// every caller reducing the same partials obtains the same bits.
__device__ inline void sum_partials_block(const double *part, int G, int ncol, double *out) {
  constexpr int NB = 4;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6;
  for (int c0 = w; c0 < ncol; c0 += NB * nw) {
    double s[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) s[k] = 0.0;
    int b = lane;
    for (; b + 192 < G; b += 256) {
#pragma unroll
      for (int k = 0; k < NB; k++) {
        const int col = c0 + k * nw;
        if (col < ncol) {
          const double *p = part + (size_t)col * G;
          s[k] += (p[b] + p[b + 64]) + (p[b + 128] + p[b + 192]);
        }
      }
    }
    for (; b < G; b += 64) {
#pragma unroll
      for (int k = 0; k < NB; k++) {
        const int col = c0 + k * nw;
        if (col < ncol) s[k] += part[(size_t)col * G + b];
      }
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
      double v = s[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      const int col = c0 + k * nw;
      if (lane == 0 && col < ncol) out[col] = v;
    }
  }
}

// MFMA 16x16x4 in FP64 and FP32: C(16x16) += A(16x4) B(4x16).  Lane l
// supplies A(l%16, l/16) and B(l/16, l%16); element v of its result is
// C(g+4v, l%16) in FP64 and C(4g+v, l%16) in FP32, g = l/16.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma16x16x4(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef dbl4 type; };
template <> struct Vec4<float> { typedef flt4 type; };

// Row of the MFMA 16x16x4 result held in element v by lane group g.
template <typename T>
__device__ __forceinline__ int mfma_row(int g, int v) {
  return sizeof(T) == 8 ? g + 4 * v : 4 * g + v;
}

}  // namespace sa
