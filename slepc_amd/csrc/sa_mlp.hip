// Learned warm start: a 3-layer ReLU MLP (din -> hidden -> hidden -> dout) with
// BF16 weights and activations on MFMA (v_mfma_f32_16x16x32_bf16, gfx950) and
// FP32 accumulation, FP32 master weights and Adam moments.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <vector>

#include "sa_internal.h"

namespace sa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// ---------------------------------------------------------------------------
// C[M x N] (fp32, row-major, ldc) = sum_k A(i,k) B(k,j)  with generic strides
//   A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj] (bf16)
// Block tile 64x64, 4 waves of 32x32, K step 32 staged through LDS.  When a
// stride is 1 the tiles are staged with 16-byte loads (8 bf16 per thread per
// tile): along k straight into the [row][k] LDS layout, along i or j with the
// 8 values scattered into it.  Long K (the backward product dY W3, K = dout)
// is split over gridDim.z slices whose FP32 partials a second kernel adds in a
// fixed order; otherwise the epilogue runs in place: optional bias[j] (fp32),
// optional ReLU, optional bf16 copy, beta.
// ---------------------------------------------------------------------------
struct GemmArgs {
  int M, N, K;
  const __bf16 *A;
  long sai, sak;
  const __bf16 *B;
  long sbk, sbj;
  float *C;
  long ldc;
  __bf16 *Cb;        // optional bf16 copy of the output (ld = ldc)
  const float *bias; // optional, length N
  int relu;
  float beta;        // C = beta*C + AB (beta 0 ignores the old C)
};

constexpr int GT = 64, GK = 32;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// staging modes: 0 generic, 1 contiguous along k, 2 contiguous along the row/column index
template <int MODE>
__device__ __forceinline__ void stage_tile(__bf16 (*S)[GK + 8], const __bf16 *P, long sr, long sk, int r0, int k0,
                                           int R, int K, int t) {
  if (MODE == 1) {            // 64 rows x 4 vectors of 8 along k
    const int r = t >> 2, kv = (t & 3) * 8;
    const int gr = r0 + r, gk = k0 + kv;
    u32x4 v = (u32x4){0u, 0u, 0u, 0u};
    if (gr < R && gk < K) v = *reinterpret_cast<const u32x4 *>(P + (long)gr * sr + gk);
    *reinterpret_cast<u32x4 *>(&S[r][kv]) = v;
  } else if (MODE == 2) {     // 32 k x 8 vectors of 8 along the row index
    const int k = t >> 3, rv = (t & 7) * 8;
    const int gr = r0 + rv, gk = k0 + k;
    union { u32x4 v; __bf16 h[8]; } u;
    u.v = (u32x4){0u, 0u, 0u, 0u};
    if (gk < K && gr < R) u.v = *reinterpret_cast<const u32x4 *>(P + (long)gk * sk + gr);
#pragma unroll
    for (int q = 0; q < 8; q++) S[rv + q][k] = u.h[q];
  } else {
    for (int e = t; e < GT * GK; e += 256) {
      const int r = e / GK, k = e % GK;
      const int gr = r0 + r, gk = k0 + k;
      S[r][k] = (gr < R && gk < K) ? P[(long)gr * sr + (long)gk * sk] : to_bf16(0.f);
    }
  }
}

template <int AM, int BM>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g, int kchunk, float *part) {
  __shared__ __bf16 As[GT][GK + 8];   // [i][k]
  __shared__ __bf16 Bs[GT][GK + 8];   // [j][k]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  const int wi = (w >> 1) * 32, wj = (w & 1) * 32;
  const int kb = blockIdx.z * kchunk, ke = min(g.K, kb + kchunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += GK) {
    stage_tile<AM>(As, g.A, g.sai, g.sak, i0, k0, g.M, ke, t);
    stage_tile<BM>(Bs, g.B, g.sbj, g.sbk, j0, k0, g.N, ke, t);
    __syncthreads();
    // fragments: A lane l -> row l%16, k = 8*(l/16)..+7 ; B lane l -> col l%16, same k
    const int fr = lane & 15, fk = (lane >> 4) * 8;
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&As[wi + a * 16 + fr][fk]);
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&Bs[wj + b * 16 + fr][fk]);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // output layout: column = lane%16, row = 4*(lane/16) + v
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = i0 + wi + a * 16 + 4 * (lane >> 4) + v;
        const int j = j0 + wj + b * 16 + (lane & 15);
        if (i < g.M && j < g.N) {
          float x = acc[a][b][v];
          if (part) {                 // split K: raw partial of slice z
            part[((long)blockIdx.z * g.M + i) * g.N + j] = x;
            continue;
          }
          if (g.bias) x += g.bias[j];
          if (g.relu) x = fmaxf(x, 0.f);
          float *cp = g.C + (long)i * g.ldc + j;
          if (g.beta != 0.f) x += g.beta * (*cp);
          *cp = x;
          if (g.Cb) g.Cb[(long)i * g.ldc + j] = to_bf16(x);
        }
      }
}

// Sum of the split-K slices in slice order, then the epilogue.
__global__ void gemm_splitk_sum_kernel(GemmArgs g, int nz, const float *part) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)g.M * g.N) return;
  const int i = (int)(e / g.N), j = (int)(e % g.N);
  float x = 0.f;
  for (int z = 0; z < nz; z++) x += part[(long)z * g.M * g.N + e];
  if (g.bias) x += g.bias[j];
  if (g.relu) x = fmaxf(x, 0.f);
  float *cp = g.C + (long)i * g.ldc + j;
  if (g.beta != 0.f) x += g.beta * (*cp);
  *cp = x;
  if (g.Cb) g.Cb[(long)i * g.ldc + j] = to_bf16(x);
}

// staging mode of an operand: 1 when k is contiguous, 2 when the row/column
// index is, with 16-byte aligned vectors; else 0
static int stage_mode(const __bf16 *P, long sr, long sk, int R, int K) {
  const bool al = (reinterpret_cast<uintptr_t>(P) & 15) == 0;
  if (al && sk == 1 && sr % 8 == 0 && K % 8 == 0) return 1;
  if (al && sr == 1 && sk % 8 == 0 && R % 8 == 0) return 2;
  return 0;
}

template <int AM>
static void gemm_launch_b(int bm, dim3 grid, const GemmArgs &g, int kchunk, float *part, hipStream_t s) {
  if (bm == 1) hipLaunchKernelGGL((gemm_bf16_kernel<AM, 1>), grid, dim3(256), 0, s, g, kchunk, part);
  else if (bm == 2) hipLaunchKernelGGL((gemm_bf16_kernel<AM, 2>), grid, dim3(256), 0, s, g, kchunk, part);
  else hipLaunchKernelGGL((gemm_bf16_kernel<AM, 0>), grid, dim3(256), 0, s, g, kchunk, part);
}

// split: workspace of at least GEMM_SPLIT_MAX * M * N floats for products whose
// output tile grid is small and K long, else null
constexpr int GEMM_SPLIT_MAX = 64;
static int gemm(const GemmArgs &g, hipStream_t s, float *split = nullptr) {
  if (g.M == 0 || g.N == 0) return 0;
  const int am = stage_mode(g.A, g.sai, g.sak, g.M, g.K), bm = stage_mode(g.B, g.sbj, g.sbk, g.N, g.K);
  const int tiles = ((g.N + GT - 1) / GT) * ((g.M + GT - 1) / GT);
  int nz = 1;
  if (split && tiles < 256 && g.K >= 4096) nz = std::min(GEMM_SPLIT_MAX, std::max(1, 512 / tiles));
  const int kchunk = ((g.K + nz - 1) / nz + GK - 1) / GK * GK;
  nz = (g.K + kchunk - 1) / kchunk;
  dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT, nz);
  float *part = nz > 1 ? split : nullptr;
  if (am == 1) gemm_launch_b<1>(bm, grid, g, kchunk, part, s);
  else if (am == 2) gemm_launch_b<2>(bm, grid, g, kchunk, part, s);
  else gemm_launch_b<0>(bm, grid, g, kchunk, part, s);
  if (part)
    hipLaunchKernelGGL(gemm_splitk_sum_kernel, dim3((unsigned)(((long)g.M * g.N + 255) / 256)), dim3(256), 0, s, g,
                       nz, (const float *)part);
  SA_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
__global__ void f32_to_bf16_kernel(const float *x, __bf16 *y, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}
// input rows padded from din to dpad columns
__global__ void pad_input_kernel(const float *x, __bf16 *y, int B, int din, int dpad) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)B * dpad) return;
  const int r = (int)(i / dpad), c = (int)(i % dpad);
  y[i] = to_bf16(c < din ? x[(long)r * din + c] : 0.f);
}
// dY = 2 (Y - T) / (B*dout) ; loss partials
__global__ void mse_grad_kernel(const float *Y, const float *T, float *dY, __bf16 *dYb, long n, float scale,
                                double *part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const float d = Y[i] - T[i];
    s += (double)d * d;
    const float gval = 2.f * d * scale;
    dY[i] = gval;
    dYb[i] = to_bf16(gval);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// dZ = dH * (Hpre > 0), bf16 copy
__global__ void relu_back_kernel(float *dH, const float *H, __bf16 *dHb, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const float v = H[i] > 0.f ? dH[i] : 0.f;
    dH[i] = v;
    dHb[i] = to_bf16(v);
  }
}
// column sums over B rows: db[j] = sum_r dZ[r*N + j]
__global__ void colsum_kernel(const float *dZ, int B, int N, float *db) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int r = 0; r < B; r++) s += dZ[(long)r * N + j];
  db[j] = s;
}
// Adam (Kingma & Ba 2015) on fp32 master weights, bf16 copy refreshed
__global__ void adam_kernel(float *w, const float *gr, float *m, float *v, __bf16 *wb, long n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = to_bf16(wi);
}
// Bilinear prolongation of g x g node-centred modes to nx x ny cell centres.
__global__ void prolong_kernel(const float *modes, int nm, int g, int nx, int ny, double *out) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long per = (long)nx * ny;
  if (idx >= per * nm) return;
  const int q = (int)(idx / per);
  const long r = idx % per;
  const int iy = (int)(r / nx), ix = (int)(r % nx);
  // cell centre in coarse-cell units
  const float x = ((ix + 0.5f) / nx) * g - 0.5f, y = ((iy + 0.5f) / ny) * g - 0.5f;
  int x0 = (int)floorf(x), y0 = (int)floorf(y);
  float fx = x - x0, fy = y - y0;
  int x1 = x0 + 1, y1 = y0 + 1;
  x0 = max(0, min(g - 1, x0)); x1 = max(0, min(g - 1, x1));
  y0 = max(0, min(g - 1, y0)); y1 = max(0, min(g - 1, y1));
  const float *m = modes + (long)q * g * g;
  const float v = (1.f - fy) * ((1.f - fx) * m[y0 * g + x0] + fx * m[y0 * g + x1]) +
                  fy * ((1.f - fx) * m[y1 * g + x0] + fx * m[y1 * g + x1]);
  out[idx] = v;
}

static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace sa

using namespace sa;

struct sa_mlp_s {
  int din = 0, dpad = 32, hid = 0, dout = 0, maxB = 0;
  // parameters: fp32 master, bf16 copy, Adam moments, gradients
  float *W[3] = {}, *b[3] = {}, *mW[3] = {}, *vW[3] = {}, *mb[3] = {}, *vb[3] = {}, *gW[3] = {}, *gb[3] = {};
  __bf16 *Wb[3] = {};
  long nW[3] = {}, nb[3] = {};
  // activations for batch maxB
  __bf16 *Xb = nullptr, *H1b = nullptr, *H2b = nullptr, *dYb = nullptr, *dZ2b = nullptr, *dZ1b = nullptr;
  float *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *T = nullptr, *dY = nullptr, *dH2 = nullptr, *dH1 = nullptr,
        *Xf = nullptr;
  double *part = nullptr;
  float *split = nullptr;     // split-K partials: GEMM_SPLIT_MAX x maxB x hid
  long step = 0;
};

static int mlp_ensure_batch(sa_mlp net, int B) {
  if (B <= net->maxB) return 0;
  // freed pointers are cleared, so a failed allocation below leaves nothing
  // that sa_mlp_destroy would free twice
  auto fr = [](auto *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  net->maxB = 0;
  fr(net->Xb); fr(net->H1b); fr(net->H2b); fr(net->dYb); fr(net->dZ2b); fr(net->dZ1b);
  fr(net->H1); fr(net->H2); fr(net->Y); fr(net->T); fr(net->dY); fr(net->dH2); fr(net->dH1); fr(net->Xf);
  const long h = net->hid, o = net->dout;
  SA_HIP(hipMalloc(&net->Xf, sizeof(float) * B * net->din));
  SA_HIP(hipMalloc(&net->Xb, sizeof(__bf16) * B * net->dpad));
  SA_HIP(hipMalloc(&net->H1b, sizeof(__bf16) * B * h));
  SA_HIP(hipMalloc(&net->H2b, sizeof(__bf16) * B * h));
  SA_HIP(hipMalloc(&net->dYb, sizeof(__bf16) * B * o));
  SA_HIP(hipMalloc(&net->dZ2b, sizeof(__bf16) * B * h));
  SA_HIP(hipMalloc(&net->dZ1b, sizeof(__bf16) * B * h));
  SA_HIP(hipMalloc(&net->H1, sizeof(float) * B * h));
  SA_HIP(hipMalloc(&net->H2, sizeof(float) * B * h));
  SA_HIP(hipMalloc(&net->Y, sizeof(float) * B * o));
  SA_HIP(hipMalloc(&net->T, sizeof(float) * B * o));
  SA_HIP(hipMalloc(&net->dY, sizeof(float) * B * o));
  SA_HIP(hipMalloc(&net->dH2, sizeof(float) * B * h));
  SA_HIP(hipMalloc(&net->dH1, sizeof(float) * B * h));
  fr(net->split);
  SA_HIP(hipMalloc(&net->split, sizeof(float) * GEMM_SPLIT_MAX * B * h));
  net->maxB = B;
  return 0;
}

static int mlp_forward_dev(sa_mlp net, int B, const float *X_host) {
  hipStream_t s = stream();
  SA_TRY(mlp_ensure_batch(net, B));
  SA_HIP(hipMemcpyAsync(net->Xf, X_host, sizeof(float) * B * net->din, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(pad_input_kernel, dim3(nblk((long)B * net->dpad)), dim3(256), 0, s, net->Xf, net->Xb, B, net->din,
                     net->dpad);
  const long h = net->hid, o = net->dout, dp = net->dpad;
  // H1 = relu(X W1^T + b1): A = X (B x dp), B(k,j) = W1[j*dp + k]
  GemmArgs g{B, (int)h, (int)dp, net->Xb, dp, 1, net->Wb[0], 1, dp, net->H1, h, net->H1b, net->b[0], 1, 0.f};
  SA_TRY(gemm(g, s));
  GemmArgs g2{B, (int)h, (int)h, net->H1b, h, 1, net->Wb[1], 1, h, net->H2, h, net->H2b, net->b[1], 1, 0.f};
  SA_TRY(gemm(g2, s));
  GemmArgs g3{B, (int)o, (int)h, net->H2b, h, 1, net->Wb[2], 1, h, net->Y, o, nullptr, net->b[2], 0, 0.f};
  SA_TRY(gemm(g3, s));
  return 0;
}

static int mlp_backward_dev(sa_mlp net, int B, const float *T_host, float *loss) {
  hipStream_t s = stream();
  const long h = net->hid, o = net->dout, dp = net->dpad;
  SA_HIP(hipMemcpyAsync(net->T, T_host, sizeof(float) * B * o, hipMemcpyHostToDevice, s));
  const int nbk = 512;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(nbk), dim3(256), 0, s, net->Y, net->T, net->dY, net->dYb, (long)B * o,
                     1.f / (float)(B * o), net->part);
  // dW3 = dY^T H2  (o x h): A(i,k) = dY[k*o + i], B(k,j) = H2[k*h + j]
  GemmArgs a{(int)o, (int)h, B, net->dYb, 1, o, net->H2b, h, 1, net->gW[2], h, nullptr, nullptr, 0, 0.f};
  SA_TRY(gemm(a, s));
  hipLaunchKernelGGL(colsum_kernel, dim3(nblk(o)), dim3(256), 0, s, net->dY, B, (int)o, net->gb[2]);
  // dH2 = dY W3 (B x h): A(i,k) = dY[i*o + k], B(k,j) = W3[k*h + j]
  GemmArgs b2{B, (int)h, (int)o, net->dYb, o, 1, net->Wb[2], h, 1, net->dH2, h, nullptr, nullptr, 0, 0.f};
  SA_TRY(gemm(b2, s, net->split));
  hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((long)B * h)), dim3(256), 0, s, net->dH2, net->H2, net->dZ2b, (long)B * h);
  hipLaunchKernelGGL(colsum_kernel, dim3(nblk(h)), dim3(256), 0, s, net->dH2, B, (int)h, net->gb[1]);
  // dW2 = dZ2^T H1 (h x h)
  GemmArgs c{(int)h, (int)h, B, net->dZ2b, 1, h, net->H1b, h, 1, net->gW[1], h, nullptr, nullptr, 0, 0.f};
  SA_TRY(gemm(c, s));
  // dH1 = dZ2 W2 (B x h)
  GemmArgs d{B, (int)h, (int)h, net->dZ2b, h, 1, net->Wb[1], h, 1, net->dH1, h, nullptr, nullptr, 0, 0.f};
  SA_TRY(gemm(d, s));
  hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((long)B * h)), dim3(256), 0, s, net->dH1, net->H1, net->dZ1b, (long)B * h);
  hipLaunchKernelGGL(colsum_kernel, dim3(nblk(h)), dim3(256), 0, s, net->dH1, B, (int)h, net->gb[0]);
  // dW1 = dZ1^T X (h x dp)
  GemmArgs e{(int)h, (int)dp, B, net->dZ1b, 1, h, net->Xb, dp, 1, net->gW[0], dp, nullptr, nullptr, 0, 0.f};
  SA_TRY(gemm(e, s));
  std::vector<double> hp(nbk);
  SA_HIP(hipMemcpyAsync(hp.data(), net->part, sizeof(double) * nbk, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  double tot = 0.0;
  for (double v : hp) tot += v;
  if (loss) *loss = (float)(tot / ((double)B * o));
  return 0;
}

extern "C" {

int sa_mlp_create(int din, int hidden, int dout, uint64_t seed, sa_mlp *out) {
  SA_TRY(ensure_init());
  SA_CHECK(out && din > 0 && din <= 32 && hidden > 0 && dout > 0, "invalid MLP sizes (din <= 32)");
  BuildGuard<sa_mlp, sa_mlp_destroy> guard{new sa_mlp_s()};
  sa_mlp net = guard.h;
  net->din = din;
  net->hid = hidden;
  net->dout = dout;
  const long shapes[3][2] = {{hidden, net->dpad}, {hidden, hidden}, {dout, hidden}};
  hipStream_t s = stream();
  for (int l = 0; l < 3; l++) {
    const long nW = shapes[l][0] * shapes[l][1], nb_ = shapes[l][0];
    net->nW[l] = nW;
    net->nb[l] = nb_;
    // He-uniform initialisation from the counter-based generator; padded input columns zero
    const long fan_in = (l == 0) ? din : shapes[l][1];
    const float lim = std::sqrt(6.f / (float)fan_in);
    std::vector<float> w(nW), bz(nb_, 0.f);
    for (long i = 0; i < nW; i++) {
      const long col = i % shapes[l][1];
      w[i] = (l == 0 && col >= din) ? 0.f : (float)((2.0 * u01(seed + 101 * l, (uint64_t)i) - 1.0) * lim);
    }
    SA_HIP(hipMalloc(&net->W[l], sizeof(float) * nW));
    SA_HIP(hipMalloc(&net->Wb[l], sizeof(__bf16) * nW));
    SA_HIP(hipMalloc(&net->gW[l], sizeof(float) * nW));
    SA_HIP(hipMalloc(&net->mW[l], sizeof(float) * nW));
    SA_HIP(hipMalloc(&net->vW[l], sizeof(float) * nW));
    SA_HIP(hipMalloc(&net->b[l], sizeof(float) * nb_));
    SA_HIP(hipMalloc(&net->gb[l], sizeof(float) * nb_));
    SA_HIP(hipMalloc(&net->mb[l], sizeof(float) * nb_));
    SA_HIP(hipMalloc(&net->vb[l], sizeof(float) * nb_));
    SA_HIP(hipMemcpyAsync(net->W[l], w.data(), sizeof(float) * nW, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemcpyAsync(net->b[l], bz.data(), sizeof(float) * nb_, hipMemcpyHostToDevice, s));
    SA_HIP(hipMemsetAsync(net->mW[l], 0, sizeof(float) * nW, s));
    SA_HIP(hipMemsetAsync(net->vW[l], 0, sizeof(float) * nW, s));
    SA_HIP(hipMemsetAsync(net->mb[l], 0, sizeof(float) * nb_, s));
    SA_HIP(hipMemsetAsync(net->vb[l], 0, sizeof(float) * nb_, s));
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(nblk(nW)), dim3(256), 0, s, net->W[l], net->Wb[l], nW);
    SA_HIP(hipStreamSynchronize(s));
  }
  SA_HIP(hipMalloc(&net->part, sizeof(double) * 512));
  *out = guard.release();
  return 0;
}

int sa_mlp_destroy(sa_mlp net) {
  if (!net) return 0;
  auto fr = [](void *p) { if (p) (void)hipFree(p); };
  for (int l = 0; l < 3; l++) {
    fr(net->W[l]); fr(net->Wb[l]); fr(net->gW[l]); fr(net->mW[l]); fr(net->vW[l]);
    fr(net->b[l]); fr(net->gb[l]); fr(net->mb[l]); fr(net->vb[l]);
  }
  fr(net->Xb); fr(net->H1b); fr(net->H2b); fr(net->dYb); fr(net->dZ2b); fr(net->dZ1b);
  fr(net->H1); fr(net->H2); fr(net->Y); fr(net->T); fr(net->dY); fr(net->dH2); fr(net->dH1); fr(net->Xf);
  fr(net->part);
  fr(net->split);
  delete net;
  return 0;
}

int sa_mlp_forward(sa_mlp net, int B, const float *X, float *Y) {
  SA_CHECK(net && X && Y && B > 0, "invalid arguments");
  SA_TRY(mlp_forward_dev(net, B, X));
  SA_HIP(hipMemcpyAsync(Y, net->Y, sizeof(float) * B * net->dout, hipMemcpyDeviceToHost, stream()));
  SA_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sa_mlp_gradients(sa_mlp net, int B, const float *X, const float *T, float *gW1, float *gb1, float *gW2,
                     float *gb2, float *gW3, float *gb3, float *loss) {
  SA_CHECK(net && X && T && B > 0, "invalid arguments");
  SA_TRY(mlp_forward_dev(net, B, X));
  SA_TRY(mlp_backward_dev(net, B, T, loss));
  hipStream_t s = stream();
  float *dst[6] = {gW1, gb1, gW2, gb2, gW3, gb3};
  for (int l = 0; l < 3; l++) {
    if (dst[2 * l]) SA_HIP(hipMemcpyAsync(dst[2 * l], net->gW[l], sizeof(float) * net->nW[l], hipMemcpyDeviceToHost, s));
    if (dst[2 * l + 1]) SA_HIP(hipMemcpyAsync(dst[2 * l + 1], net->gb[l], sizeof(float) * net->nb[l], hipMemcpyDeviceToHost, s));
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int sa_mlp_train_step(sa_mlp net, int B, const float *X, const float *T, float lr, float *loss) {
  SA_CHECK(net && X && T && B > 0 && lr > 0, "invalid arguments");
  SA_TRY(mlp_forward_dev(net, B, X));
  SA_TRY(mlp_backward_dev(net, B, T, loss));
  hipStream_t s = stream();
  net->step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)net->step), c2 = 1.f - std::pow(b2, (float)net->step);
  for (int l = 0; l < 3; l++) {
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(net->nW[l])), dim3(256), 0, s, net->W[l], net->gW[l], net->mW[l],
                       net->vW[l], net->Wb[l], net->nW[l], lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(net->nb[l])), dim3(256), 0, s, net->b[l], net->gb[l], net->mb[l],
                       net->vb[l], (__bf16 *)nullptr, net->nb[l], lr, b1, b2, eps, c1, c2);
  }
  SA_HIP(hipGetLastError());
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int sa_mlp_get_params(sa_mlp net, float *W1, float *b1, float *W2, float *b2, float *W3, float *b3) {
  SA_CHECK(net, "null network");
  hipStream_t s = stream();
  float *dst[6] = {W1, b1, W2, b2, W3, b3};
  for (int l = 0; l < 3; l++) {
    if (dst[2 * l]) SA_HIP(hipMemcpyAsync(dst[2 * l], net->W[l], sizeof(float) * net->nW[l], hipMemcpyDeviceToHost, s));
    if (dst[2 * l + 1]) SA_HIP(hipMemcpyAsync(dst[2 * l + 1], net->b[l], sizeof(float) * net->nb[l], hipMemcpyDeviceToHost, s));
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int sa_mlp_set_params(sa_mlp net, const float *W1, const float *b1, const float *W2, const float *b2, const float *W3,
                      const float *b3) {
  SA_CHECK(net, "null network");
  hipStream_t s = stream();
  const float *src[6] = {W1, b1, W2, b2, W3, b3};
  for (int l = 0; l < 3; l++) {
    if (src[2 * l]) {
      SA_HIP(hipMemcpyAsync(net->W[l], src[2 * l], sizeof(float) * net->nW[l], hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(nblk(net->nW[l])), dim3(256), 0, s, net->W[l], net->Wb[l], net->nW[l]);
    }
    if (src[2 * l + 1])
      SA_HIP(hipMemcpyAsync(net->b[l], src[2 * l + 1], sizeof(float) * net->nb[l], hipMemcpyHostToDevice, s));
  }
  SA_HIP(hipStreamSynchronize(s));
  return 0;
}

int sa_prolong_bilinear(int nmodes, int g, int nx, int ny, const float *modes_host, double *out_host) {
  SA_TRY(ensure_init());
  SA_CHECK(nmodes > 0 && g > 1 && nx > 0 && ny > 0 && modes_host && out_host, "invalid arguments");
  hipStream_t s = stream();
  float *dm = nullptr;
  double *dout = nullptr;
  const long per = (long)nx * ny;
  SA_HIP(hipMalloc(&dm, sizeof(float) * nmodes * g * g));
  SA_HIP(hipMalloc(&dout, sizeof(double) * nmodes * per));
  SA_HIP(hipMemcpyAsync(dm, modes_host, sizeof(float) * nmodes * g * g, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(prolong_kernel, dim3(nblk(per * nmodes)), dim3(256), 0, s, dm, nmodes, g, nx, ny, dout);
  SA_HIP(hipMemcpyAsync(out_host, dout, sizeof(double) * nmodes * per, hipMemcpyDeviceToHost, s));
  SA_HIP(hipStreamSynchronize(s));
  (void)hipFree(dm);
  (void)hipFree(dout);
  return 0;
}

}  // extern "C"
