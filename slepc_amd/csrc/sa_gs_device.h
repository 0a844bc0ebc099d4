// Device code of the fused Gram-Schmidt step (SpMV, classical Gram-Schmidt,
// DGKS refinement "if needed", normalisation), shared by the single-basis
// sweep (sa_gs.hip) and the lock-step batched sweep (sa_batched.hip).
//
// A column c is orthogonalised by three kernels ("slots"):
//   slot 1: v = A V(:,c-1) (when a matrix is given) or v = V(:,c); partial
//           dot products V(:,0:c+1)^T v per workgroup (column c is v itself,
//           so the last entry is v^T v)
//   slot 2: dots of pass 1 -> decision -> v = (v - V h) [* 1/norm], and the
//           partial dots of the updated v when a second pass is needed
//   slot 3: dots of pass 2 -> decision -> v = (v - V h) * 1/norm; the rare
//           third pass or failed norm estimate is finished by the lead
//           workgroup alone, over all rows, with in-workgroup reductions
// Between the slots a reduction kernel adds the partial sums in a fixed order
// (no fence or atomic: an agent-scope release on MI355X writes back the XCD's
// L2).  Every workgroup of slots 2-3 reads the same sums and takes the same
// decision; the lead workgroup records the coefficients, the norm and the
// state.  Slot s reads the state st[(s+1)%3] and writes st[s-1], so no
// workgroup reads a state that is being written.
#pragma once
#include "sa_device.h"
#include "sa_internal.h"

namespace sa {

enum GsAct : int { ACT_MULT_SCALE = 1, ACT_MULT_DOTS = 2, ACT_SCALE_ONLY = 3 };

template <typename T>
struct GsArgs {
  int64_t n, ld;            // rows of this basis, leading dimension
  T *V;                     // row 0 of this basis (column 0)
  int c, slot;
  const int32_t *rp, *ci;   // CSR rows of this basis (slot 1 of a sweep step), else null
  const T *val;
  const T *xin;             // SpMV input vector in the column-index space of ci
  GsState *st;              // GS_NSTATE states
  double *dbuf;             // coefficient buffer: column c receives h and the norm
  int m;
  const double *dots_in;    // dot products of the previous slot, reduced
  double *part_out;         // this slot's partial sums, [c+1][G]
  int G, rpb;
  double eta;
  int refine;               // BVSetOrthogonalization refinement: GS_REFINE_*
};

// The rule of BVOrthogonalizeGS (bvorthog.c:145-217) applied to the dots of
// one pass: dots[0:c] = V^T v, dots[c] = v^T v, sumsq = sum dots[i]^2.
// Refinement "if needed" repeats the pass while the estimated norm falls
// below eta times the norm before it (at most three passes); "never" takes
// one pass and then the true norm of the result (lindep only for a zero
// norm); "always" takes two passes, the second with the norm estimate.  An
// estimate that fails (<= 0) is replaced by the true norm after the update
// (bvorthog.c:124-127).  Updates S; returns the action for v.
__device__ inline int gs_rule(GsState &S, const double *dots, double sumsq, int c, double eta, int refine,
                              double *scale, int *record) {
  const double self = dots[c];
  int fin = 0, act = ACT_MULT_DOTS;
  double nrm_f = 0.0, onrm_f = 0.0;
  *record = 0;
  if (S.pending) {                       // true norm of the previous update
    const double nrm = safe_sqrt(self);
    S.pending = 0;
    if (refine != GS_REFINE_IFNEEDED || !(nrm != 0.0 && nrm < eta * S.onrm && S.pass < 3)) {
      fin = 1; nrm_f = nrm; onrm_f = S.onrm; act = ACT_SCALE_ONLY;
    }
  }
  if (!fin) {                            // one CGS1 pass (bvorthog.c:91-132)
    S.pass++;
    S.passes_total++;
    *record = 1;
    const double onrm = safe_sqrt(self);
    if (refine == GS_REFINE_NEVER) {     // the true norm after the update (bvorthog.c:187-193)
      S.pending = 1; S.onrm = onrm; act = ACT_MULT_DOTS;
    } else if (refine == GS_REFINE_ALWAYS && S.pass == 1) {   // a second pass follows
      act = ACT_MULT_DOTS;
    } else {
      const double est = onrm * onrm - sumsq;
      if (est <= 0.0) {                  // estimate failed: measure after the update
        S.pending = 1; S.onrm = onrm; act = ACT_MULT_DOTS;
      } else {
        const double nrm = sqrt(est);
        if (refine == GS_REFINE_IFNEEDED && nrm != 0.0 && nrm < eta * onrm && S.pass < 3) {
          S.onrm = onrm; act = ACT_MULT_DOTS;
        } else {
          fin = 1; nrm_f = nrm; onrm_f = onrm; act = ACT_MULT_SCALE;
        }
      }
    }
  }
  *scale = 1.0;
  if (fin) {                             // bvorthog.c:175-216 and 417-422
    const int lind = refine == GS_REFINE_NEVER ? !(nrm_f != 0.0)
                                               : !(nrm_f != 0.0 && fabs(nrm_f) >= eta * fabs(onrm_f));
    S.lindep = lind;
    S.nrm = nrm_f;
    S.beta = nrm_f;
    S.done = 1;
    *scale = (nrm_f != 0.0 && nrm_f != 1.0) ? 1.0 / nrm_f : 1.0;
    if (lind) { S.stop = 1; S.stop_col = c; }
  }
  S.dots_ready = (act == ACT_MULT_DOTS);
  return act;
}

// Thread 0 applies the rule to s_dots; the lead workgroup then records the
// coefficients (BV_CleanCoefficients + BV_AddCoefficients, bvimpl.h:308-322),
// the norm at row c and the new state.
template <typename T>
__device__ void gs_decide_block(const GsArgs<T> &a, const double *s_dots, GsState *s_st, int *s_act, int *s_rec,
                                double *s_scale, bool lead) {
  const int t = threadIdx.x, c = a.c;
  if (t == 0) {
    double sumsq = 0.0;
    for (int i = 0; i < c; i++) sumsq += s_dots[i] * s_dots[i];
    GsState S = *s_st;
    double scale;
    int rec;
    *s_act = gs_rule(S, s_dots, sumsq, c, a.eta, a.refine, &scale, &rec);
    *s_st = S;
    *s_scale = scale;
    *s_rec = rec;
  }
  __syncthreads();
  if (lead) {
    double *h = a.dbuf + (size_t)c * a.m;
    if (*s_rec) {
      const bool first = s_st->pass == 1;
      for (int i = t; i < c; i += BLK) h[i] = (first ? 0.0 : h[i]) + s_dots[i];
    }
    if (t == 0) {
      if (s_st->done) h[c] = s_st->lindep ? 0.0 : s_st->nrm;
      a.st[a.slot - 1] = *s_st;
    }
  }
}

// Rows [rb, re) in chunks of CH rows.  Per chunk: the vector chunk (SpMV into
// v when spmv, else loaded from V(:,c)), the tile V(rows, 0:c) in LDS, the
// update v = (v - V(:,0:c) cf) * scl when upd (the product only when mult),
// the store of v, and per-thread partial dot products of V(:,0:c+1) with v
// when dots.  The tile is loaded with GS_UT loads per thread in flight (a wave
// reads 512 B column segments).  Called by all threads of the workgroup with
// uniform arguments.
constexpr int GS_UT = 8;

// A basis tile load.  NT: non-temporal, for bases larger than the last-level
// cache, whose columns are not re-read before they are evicted anyway (about
// 3% faster at 2.1M rows x 65-97 columns; 4% slower on a cache-resident basis).
// The same flag selects wave-uniform column addressing in gs_rows.
template <bool NT, typename T>
__device__ __forceinline__ T tile_load(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <typename T, int L, bool NT>
__device__ double gs_rows(const GsArgs<T> &a, int64_t rb, int64_t re, bool spmv, bool mult, bool upd, bool dots,
                          T scl, T *Vs, T *vs, const T *cf, T *rt) {
  const int t = threadIdx.x, c = a.c;
  const int ncol = c + 1, nsl = BLK / ncol;
  const int dcol = t % ncol, dsl = t / ncol;
  const bool dact = dots && dsl < nsl;
  const bool need_tile = (mult || dots) && c > 0;
  // NT (a basis larger than the last-level cache): the wave's column group is
  // marked wave-uniform, so the column bases of the tile loads are scalar and
  // a lane keeps only its row offset (61 instead of 111 VGPRs: room for more
  // workgroups per CU); on a cache-resident basis the scalar address
  // arithmetic measured slower, so it keeps the per-lane form there
  const int prow = t & (CH - 1), pcg = NT ? __builtin_amdgcn_readfirstlane(t >> 6) : (t >> 6);
  T *vc = a.V + (size_t)c * a.ld;
  double acc = 0.0;
  for (int64_t r0 = rb; r0 < re; r0 += CH) {
    const int rows = (int)min((int64_t)CH, re - r0);
    const bool in = prow < rows;
    const T *src = a.V + r0 + prow;
    // 1. the first GS_UT tile loads, issued before the SpMV's dependent
    //    gathers so that their latencies overlap
    T x[GS_UT];
    if (need_tile) {
#pragma unroll
      for (int u = 0; u < GS_UT; u++) {
        const int i = pcg + 4 * u;
        x[u] = (in && i < c) ? tile_load<NT>(src + (size_t)i * a.ld) : (T)0;
      }
    }
    // 2. the vector chunk
    if (spmv) {
      constexpr int RPS = BLK / L;
      const int g = t / L, lg = t % L;
      for (int rr = g; rr < CH; rr += RPS) {
        T s = (T)0;
        if (rr < rows) {
          const int64_t row = r0 + rr;
          const int32_t pe = a.rp[row + 1];
          for (int32_t p = a.rp[row] + lg; p < pe; p += L) s += a.val[p] * a.xin[a.ci[p]];
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lg == 0) {
          vs[rr] = (rr < rows) ? s : (T)0;
          if (rr < rows) vc[r0 + rr] = s;
        }
      }
    } else if (t < CH) {
      vs[t] = (t < rows) ? vc[r0 + t] : (T)0;
    }
    // 3. the tile V(r0:r0+CH, 0:c) into LDS: the first batch from registers,
    //    the rest loaded GS_UT at a time (coalesced along rows)
    if (need_tile) {
#pragma unroll
      for (int u = 0; u < GS_UT; u++) {
        const int i = pcg + 4 * u;
        if (i < c) Vs[i * LDP + prow] = x[u];
      }
      for (int i0 = 4 * GS_UT; i0 < c; i0 += 4 * GS_UT) {
#pragma unroll
        for (int u = 0; u < GS_UT; u++) {
          const int i = i0 + pcg + 4 * u;
          x[u] = (in && i < c) ? tile_load<NT>(src + (size_t)i * a.ld) : (T)0;
        }
#pragma unroll
        for (int u = 0; u < GS_UT; u++) {
          const int i = i0 + pcg + 4 * u;
          if (i < c) Vs[i * LDP + prow] = x[u];
        }
      }
    }
    __syncthreads();
    // 4. update v = (v - V cf) * scl, four threads per row
    if (upd) {
      const int row = t & (CH - 1), cg = t >> 6;
      T s = (T)0;
      if (mult)
        for (int i = cg; i < c; i += 4) s += Vs[i * LDP + row] * cf[i];
      rt[t] = s;
      __syncthreads();
      if (t < CH) {
        const T tot = (rt[t] + rt[t + 64]) + (rt[t + 128] + rt[t + 192]);
        const T vn = (vs[t] - tot) * scl;
        vs[t] = (t < rows) ? vn : (T)0;
        if (t < rows) vc[r0 + t] = vn;
      }
      __syncthreads();
    }
    // 5. dot products with the (updated) vector, nsl threads per column
    if (dact) {
      const T *colp = (dcol < c) ? (Vs + dcol * LDP) : vs;
      T s = (T)0;
      for (int r = dsl; r < rows; r += nsl) s += colp[r] * vs[r];
      acc += (double)s;
    }
    __syncthreads();
  }
  return acc;
}

// Adds the per-thread partials of gs_rows into one value per column: into
// out[col] (LDS) or into part[col * G + bidx].
__device__ inline void gs_combine(double acc, int ncol, double *s_acc, double *out, double *part, int G, int bidx) {
  const int t = threadIdx.x, nsl = BLK / ncol;
  s_acc[t] = acc;
  __syncthreads();
  if (t < ncol) {
    double s = 0.0;
    for (int q = 0; q < nsl; q++) s += s_acc[q * ncol + t];
    if (out) out[t] = s;
    else part[(size_t)t * G + bidx] = s;
  }
  __syncthreads();
}

// Bytes of dynamic LDS a slot needs for column index c (tile, vector chunk,
// coefficients, update partials).
inline size_t gs_lds_bytes(int c, size_t es) { return ((size_t)c * LDP + CH + c + BLK) * es; }

// One slot of column a.c for the rows [rb, re) of workgroup bidx of this
// basis (bidx 0 is the lead).  Every workgroup of the basis calls it.
template <typename T, int L, bool NT>
__device__ void gs_slot(const GsArgs<T> &a, int64_t rb, int64_t re, int bidx) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int c = a.c;
  T *Vs = reinterpret_cast<T *>(smem_raw);   // [c][LDP]
  T *vs = Vs + (size_t)c * LDP;              // [CH]
  T *cf = vs + CH;                           // [c]
  T *rt = cf + c;                            // [BLK]
  __shared__ double s_dots[GS_MAXC + 1];
  __shared__ double s_acc[BLK];
  __shared__ GsState s_st;
  __shared__ int s_act, s_rec;
  __shared__ double s_scale;
  const int t = threadIdx.x;
  const bool lead = bidx == 0;
  const int ncol = c + 1;
  // the state and (slots 2-3) the reduced dots, loaded together
  if (t == 0) s_st = a.st[(a.slot + 1) % GS_NSTATE];
  if (a.slot > 1)
    for (int i = t; i < ncol; i += BLK) s_dots[i] = a.dots_in[i];
  __syncthreads();

  if (a.slot == 1) {
    const bool stopped = s_st.stop != 0;
    if (lead && t == 0) {
      GsState S = s_st;
      if (stopped) {
        S.done = 1;
        S.dots_ready = 0;
      } else {
        S.pass = 0; S.pending = 0; S.lindep = 0; S.done = 0; S.dots_ready = 1;
      }
      a.st[0] = S;
    }
    if (stopped) return;
    const double acc =
        gs_rows<T, L, NT>(a, rb, re, a.rp != nullptr, false, false, true, (T)1, Vs, vs, cf, rt);
    gs_combine(acc, ncol, s_acc, nullptr, a.part_out, a.G, bidx);
    return;
  }

  // slots 2 and 3
  if (s_st.done || !s_st.dots_ready) {        // nothing left for this column
    if (lead && t == 0) {
      GsState S = s_st;
      S.dots_ready = 0;
      a.st[a.slot - 1] = S;
    }
    return;
  }
  gs_decide_block(a, s_dots, &s_st, &s_act, &s_rec, &s_scale, lead);
  int act = s_act;

  if (act == ACT_MULT_DOTS && a.slot == 3) {
    // Third pass or failed estimate: the lead workgroup finishes the column
    // over all rows (the others took the same decision and leave v alone).
    if (!lead) return;
    if (t == 0) s_st.slow_total++;
    while (act == ACT_MULT_DOTS) {             // ends: S.pass < 3 bounds the passes
      for (int i = t; i < c; i += BLK) cf[i] = (T)s_dots[i];
      __syncthreads();
      const double acc = gs_rows<T, L, NT>(a, 0, a.n, false, true, true, true, (T)1, Vs, vs, cf, rt);
      gs_combine(acc, ncol, s_acc, s_dots, nullptr, 0, 0);
      gs_decide_block(a, s_dots, &s_st, &s_act, &s_rec, &s_scale, true);
      act = s_act;
    }
    const bool mult = act == ACT_MULT_SCALE;
    if (mult)
      for (int i = t; i < c; i += BLK) cf[i] = (T)s_dots[i];
    __syncthreads();
    gs_rows<T, L, NT>(a, 0, a.n, false, mult, true, false, (T)s_scale, Vs, vs, cf, rt);
    return;
  }

  const bool mult = act == ACT_MULT_SCALE || act == ACT_MULT_DOTS;
  const bool dots = act == ACT_MULT_DOTS;
  if (mult)
    for (int i = t; i < c; i += BLK) cf[i] = (T)s_dots[i];
  __syncthreads();
  const double acc = gs_rows<T, L, NT>(a, rb, re, false, mult, true, dots, (T)s_scale, Vs, vs, cf, rt);
  if (dots) gs_combine(acc, ncol, s_acc, nullptr, a.part_out, a.G, bidx);
}

// dots[col] = sum of the G partials part[col * G + b], b in a fixed order
// (one workgroup of BLK threads per column; one memory round trip).
__device__ inline void gs_sum_column(const double *part, int G, int col, double *dots) {
  __shared__ double red[BLK / 64];
  const int t = threadIdx.x;
  const double *p = part + (size_t)col * G;
  double s = 0.0;
  for (int b = t; b < G; b += BLK) s += p[b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
    for (int w = 0; w < BLK / 64; w++) tot += red[w];
    dots[col] = tot;
  }
}

}  // namespace sa
