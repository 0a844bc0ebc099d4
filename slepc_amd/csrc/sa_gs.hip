// Fused Gram-Schmidt step of the Arnoldi/Lanczos sweep on one basis, for
// gfx950 (MI355X): the three slot kernels of sa_gs_device.h over the row
// blocks of the basis, the per-column reductions between them, and the host
// launch sequence.  Every decision of the reference's rules
// (bvorthog.c:91-217, 380-427) is taken on the device.
#include <cstdlib>
#include <mutex>

#include "sa_gs_device.h"

namespace sa {

template <typename T, int L, bool NT>
__global__ __launch_bounds__(BLK) void gs_kernel(GsArgs<T> a) {
  const int64_t rb = (int64_t)blockIdx.x * a.rpb;
  gs_slot<T, L, NT>(a, rb, min(a.n, rb + a.rpb), (int)blockIdx.x);
}

// Bases wide enough that the FP64 tile exceeds 64 KB of LDS (beyond column
// 107) need the kernels' dynamic-LDS limit raised to what the device allows
// a workgroup (160 KB on gfx950); done once per instantiation.
template <typename T, bool NT>
static int allow_wide_lds() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const int bytes = lds_max_bytes() - 8192;   // room for the kernels' static LDS
    for (const void *f : {(const void *)gs_kernel<T, 4, NT>, (const void *)gs_kernel<T, 8, NT>,
                          (const void *)gs_kernel<T, 16, NT>, (const void *)gs_kernel<T, 32, NT>})
      if (err == hipSuccess) err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  });
  SA_HIP(err);
  return 0;
}

template <typename T, bool NT>
static void launch_gs_slot(const GsArgs<T> &a, int lanes, int G, size_t shm, hipStream_t s) {
  switch (lanes) {
    case 4: hipLaunchKernelGGL((gs_kernel<T, 4, NT>), dim3(G), dim3(BLK), shm, s, a); break;
    case 8: hipLaunchKernelGGL((gs_kernel<T, 8, NT>), dim3(G), dim3(BLK), shm, s, a); break;
    case 16: hipLaunchKernelGGL((gs_kernel<T, 16, NT>), dim3(G), dim3(BLK), shm, s, a); break;
    default: hipLaunchKernelGGL((gs_kernel<T, 32, NT>), dim3(G), dim3(BLK), shm, s, a); break;
  }
}

// Reduction of a slot's partial sums before the next slot: one workgroup per
// column.  It runs even when the slot produced no dot products (the next slot
// then ignores them), which saves the round trip of reading the state first.
__global__ __launch_bounds__(BLK) void gs_sum_kernel(const double *part, int G, double *dots) {
  gs_sum_column(part, G, (int)blockIdx.x, dots);
}

__global__ void gs_reset_kernel(GsState *st) {
  if (threadIdx.x == 0) {
    GsState S = st[2];
    S.stop = 0; S.stop_col = 0; S.pending = 0; S.lindep = 0; S.pass = 0;
    S.done = 0; S.dots_ready = 0; S.beta = 0.0;
    st[2] = S;
  }
}

// Fewest 64-row chunks per workgroup: SLEPC_AMD_GS_MIN_CHUNKS when set, else
// 2 for a cache-resident basis while the application runs several solves
// concurrently (sa_set_concurrent_solves), else 1.  One chunk per workgroup
// is fastest for one solve alone (1.84 ms per 64-step sweep at n = 65,536,
// against 2.34 ms with two); with 8 solves in flight two chunks let the slot
// kernels of different solves co-reside (92.7 against 78.5 eigenpairs/s on
// eight held-out 256x256 FV problems; four chunks: 83.7).
int gs_min_chunks(bool big) {
  static const int env = [] {
    const char *e = std::getenv("SLEPC_AMD_GS_MIN_CHUNKS");
    const int k = e ? std::atoi(e) : 0;
    return k < 0 ? 0 : k;
  }();
  if (env > 0) return env;
  return (!big && concurrent_solves() >= 2) ? 2 : 1;
}

int gs_max_blocks(bool big) {
  static const int v = [] {
    const char *e = std::getenv("SLEPC_AMD_GS_MAX_BLOCKS");
    const int k = e ? std::atoi(e) : 0;
    return k < 0 ? 0 : k;
  }();
  return v > 0 ? v : (big ? 4096 : 1024);
}

template <typename T>
static int gs_launch(BV &V, const Mat *A, int c, double eta, hipStream_t s) {
  GsArgs<T> a{};
  a.n = V.n; a.ld = V.ld; a.V = static_cast<T *>(V.data); a.c = c;
  a.st = V.st; a.dbuf = V.dbuf; a.m = V.m; a.eta = eta;
  a.G = V.G; a.rpb = V.rows_per_block; a.refine = V.refine;
  const size_t shm = gs_lds_bytes(c, sizeof(T));
  SA_CHECK(shm + 4096 + 8 * BLK <= (size_t)lds_max_bytes(), "basis too wide for the LDS tile of the Gram-Schmidt kernel");
  const bool nt = (int64_t)V.ld * V.m * (int64_t)sizeof(T) > GS_NT_BYTES;
  if (shm + 4096 + 8 * BLK > 65536) SA_TRY((nt ? allow_wide_lds<T, true>() : allow_wide_lds<T, false>()));
  double *pa = V.ws.part, *pb = V.ws.part + (size_t)V.G * (GS_MAXC + 1);
  double *da = V.ws.dots, *db = V.ws.dots + (GS_MAXC + 1);
  // Operators larger than the caches run the stream SpMV as a kernel of its
  // own, V(:,c) = A V(:,c-1), and slot 1 then reads v like slots 2-3: the
  // fused SpMV gathers row by row inside the slot and costs 2-3x the stream
  // kernel there.  Small operators stay fused (one launch fewer per step).
  const bool fused = A && A->spmv_rb == 0;
  if (A && !fused) SA_TRY(launch_spmv(*A, a.V + (size_t)(c - 1) * a.ld, a.V + (size_t)c * a.ld, s));
  const int lanes = fused ? A->lanes : 8;
  for (int slot = 1; slot <= 3; slot++) {
    a.slot = slot;
    const bool spmv = slot == 1 && fused;
    a.rp = spmv ? A->rowptr : nullptr;
    a.ci = spmv ? A->col : nullptr;
    a.val = spmv ? static_cast<const T *>(A->val) : nullptr;
    a.xin = spmv ? a.V + (size_t)(c - 1) * a.ld : nullptr;
    a.part_out = slot == 1 ? pa : pb;
    a.dots_in = slot == 2 ? da : db;
    if (nt) launch_gs_slot<T, true>(a, lanes, V.G, shm, s);
    else launch_gs_slot<T, false>(a, lanes, V.G, shm, s);
    if (slot < 3)
      hipLaunchKernelGGL(gs_sum_kernel, dim3(c + 1), dim3(BLK), 0, s, (const double *)a.part_out, V.G,
                         slot == 1 ? da : db);
  }
  SA_HIP(hipGetLastError());
  return 0;
}

int launch_gs_column(BV &V, const Mat *A, int c, double eta, hipStream_t s) {
  SA_CHECK(c >= 0 && c < GS_MAXC && c < V.m, "column index out of range for the Gram-Schmidt kernel");
  SA_CHECK(!A || (A->prec == V.prec && A->n == V.n), "matrix and basis sizes or precisions differ");
  SA_CHECK(!A || c > 0, "SpMV step needs a previous column");
  if (V.prec == SA_FP64) return gs_launch<double>(V, A, c, eta, s);
  return gs_launch<float>(V, A, c, eta, s);
}

int launch_gs_reset(BV &V, hipStream_t s) {
  hipLaunchKernelGGL(gs_reset_kernel, dim3(1), dim3(64), 0, s, V.st);
  SA_HIP(hipGetLastError());
  return 0;
}

}  // namespace sa
