// cn_hsc.hip -- apply of the hierarchical Schur preconditioner.
//
// With level l's rows split into the eliminated set F and the kept set C,
// A_l = [D_F A_FC; A_CF A_CC] and the hierarchy's next matrix approximates
// S = A_CC - A_CF D_F^-1 A_FC.  The block factorisation gives
//   down: r_{l+1} = r_C + R r_F        (R = -A_CF D_F^-1)
//   coarse: z_L = A_L^-1 r_L          (dense inverse, FP64 MFMA GEMM)
//   up:   z_F = D_F^-1 r_F + P z_C    (P = -D_F^-1 A_FC = R^T),  z_C = z_{l+1}
// R rows are the coarse vertices and gather their fine (F) neighbours: the
// per-coarse restriction of SlangCodegen/HierarchicalSparsify.lean:64-93 in
// its CSR-by-coarse form, with the Schur weights instead of 1.
//
// Every level is addressed in level-0 rows, so the whole apply works in
// place in the output z (n0 x k): the down pass rewrites the kept rows of a
// level (reading the eliminated ones, which it never writes), the up pass
// writes only the eliminated rows (the kept rows already hold their z from
// the deeper levels).  Level 0 reads the input r instead of z.  Only the
// coarse rows are gathered into a small buffer for the GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "cn_internal.h"
#include "cn_kernels.h"
#include "cn_prof.h"

namespace cn {

size_t coarse_product_work(int m, int k);
bool coarse_bf3_applies(int m, int k);
void coarse_split_bf16(int m, int ld, const double *a, uint16_t *hi, uint16_t *lo,
                       hipStream_t s);
int coarse_product_bf3(int m, int ld, int k, const uint16_t *Mh, const uint16_t *Ml,
                       const int32_t *rows, const float *S, float *Z, const int32_t *ctl,
                       double *part, hipStream_t s, double *rz);
int coarse_rz_rows(int m);
template <typename T>
int coarse_product(int m, int k, const double *M, const float *M32, const int32_t *rows,
                   const T *S, T *Z, const int32_t *ctl, double *part, hipStream_t s,
                   double *rz);

// ---- wide form: LPR lanes per (row, LPR VEC column tile), 64 / LPR rows
// per wavefront -----------------------------------------------------------
// Z[Fg[f]] = dinv[f] src[Fg[f]] + sum_e pval[e] Z[pcol[e]]   (src = r or Z)
// rz != nullptr: the block also writes its partial sums of d_f r_f^2 per
// column (row blockIdx.x of rz): over the levels these and the coarse
// rows' r_c . z_c add up to r . M^-1 r, since the apply is the symmetric
// block factorisation M^-1 = L^T diag(D_F^-1 ..., S_c^-1) L whose L r the
// down sweep leaves in the eliminated rows (the up pass's src).
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_hsc_up(int32_t nF, const int32_t *__restrict__ Fg, const T *__restrict__ dinv,
          const int32_t *__restrict__ pptr, const int32_t *__restrict__ pcol,
          const T *__restrict__ pval, int64_t k, const T *src, T *Z,
          const int32_t *__restrict__ ctl, double *__restrict__ rz, int32_t rz_add) {
  if (ctl && ctl_idle(ctl)) return;
  using P = Pack<LPR>;
  const int64_t c0 = P::col0(VEC);
  const bool valid = c0 < k;
  // loads and their uses are unconditional (lanes past k read the last VEC
  // columns), only the store is not: a load consumed under a branch makes
  // the compiler wait for the previous row's store before the next row
  const int64_t cl = valid ? c0 : k - VEC;
  double pd[1][VEC] = {};
  // the row of rz this block adds to, loaded before the block's rows (at
  // the end its latency would idle the block)
  double pre[VEC] = {};
  if (rz && (int32_t)blockIdx.x < rz_add) {
#pragma unroll
    for (int m = 0; m < VEC; ++m) pre[m] = rz[(int64_t)blockIdx.x * k + min(c0 + m, k - 1)];
  }
  const RowIter ri = row_iter(0, nF, P::kSlots, P::slot());
  for (int32_t row = ri.first; row < ri.end; row += ri.stride) {
    const int32_t f = P::u(row);
    const int64_t g = (int64_t)P::u(Fg[f]) * k;
    const T d = dinv[f];
    T acc[VEC];
    VecIO<T, VEC>::ld(src + g + cl, acc);
#pragma unroll
    for (int m = 0; m < VEC; ++m) {
      pd[0][m] += (double)d * (double)acc[m] * (double)acc[m];
      acc[m] = d * acc[m];
    }
    gather_wide<T, VEC>(P::u(pptr[f]), P::u(pptr[f + 1]), pcol, pval, Z, k, cl, true, acc);
    if (valid) VecIO<T, VEC>::st(Z + g + c0, acc);
  }
  if (rz) block_partials_wide<VEC, 1, LPR>(pd, rz, k, c0, 0, rz_add, pre);
}

// ---- the down sweep (wide form) -----------------------------------------
// Level l's pass r_{l+1}[c] = r_l[c] + (R_l r_l[F_l])[c] rewrites every kept
// row at every level.  Instead the levels go in groups [a, e): within a
// group, a row eliminated at level m (a < m < e) is taken straight from its
// value at level a to r_m[v] = r_a[v] + sum_{a<=l<m} (R_l r_l[F_l])[v], and
// the rows of level e to r_e.  Each such chain gathers only rows eliminated
// at levels a .. m-1, whose values are final when its launch runs, so a
// group is one launch per level of elimination and writes each row once;
// longer groups write less but gather rows of lower levels again in every
// launch (on the 100k capsule union at k = 768, groups of 1 / 2 / 3 / 4 /
// 17 levels have 2.17 / 1.53 / 1.44 / 1.48 / 2.05 GB of algorithmic traffic
// a sweep: rows read and written, distinct gathered rows once a launch).  r_a is r at
// a = 0 and Z otherwise; a group of one level is the in-place level pass.
// The sum is the level passes' -- r_a[v], then the R entries level by level,
// each in order -- with the weight-0 padding of the gathers only at the end
// of the two source segments.
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_hsc_chain(int32_t nrows, const int32_t *__restrict__ rows, int32_t a, int32_t m,
             const int32_t *__restrict__ seg_ptr, const int32_t *__restrict__ seg,
             const int32_t *__restrict__ ecol, const T *__restrict__ eval, int64_t k,
             const T *srcA, T *Z, const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  using P = Pack<LPR>;
  const int64_t c0 = P::col0(VEC);
  const bool valid = c0 < k;
  // loads and their uses are unconditional (lanes past k read the last VEC
  // columns), only the store is not
  const int64_t cl = valid ? c0 : k - VEC;
  const RowIter ri = row_iter(0, nrows, P::kSlots, P::slot());
  for (int32_t i = ri.first; i < ri.end; i += ri.stride) {
    const int32_t v = P::u(rows[P::u(i)]);
    const int32_t sp = P::u(seg_ptr[v]);
    const int32_t e0 = P::u(seg[sp + a]), e1 = P::u(seg[sp + a + 1]), e2 = P::u(seg[sp + m]);
    T acc[VEC];
    VecIO<T, VEC>::ld(srcA + (int64_t)v * k + cl, acc);
    // level a (a kept row has ~2-3 eliminated neighbours: chunks of 4),
    // then levels a+1 .. m-1
    gather_wide<T, VEC, 4>(e0, e1, ecol, eval, srcA, k, cl, true, acc);
    gather_wide<T, VEC, 8>(e1, e2, ecol, eval, Z, k, cl, true, acc);
    if (valid) VecIO<T, VEC>::st(Z + (int64_t)v * k + c0, acc);
  }
}

// The same launch over prebuilt row records {v, e0, e1, e2}: one 16-byte
// load per row instead of the dependent chain rows[i] -> seg_ptr[v] ->
// seg[...] before the gathers can start.
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_hsc_chain_rec(int32_t nrows, const int4 *__restrict__ rec,
                 const int32_t *__restrict__ ecol, const T *__restrict__ eval, int64_t k,
                 const T *srcA, T *Z, const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  using P = Pack<LPR>;
  const int64_t c0 = P::col0(VEC);
  const bool valid = c0 < k;
  const int64_t cl = valid ? c0 : k - VEC;
  const RowIter ri = row_iter(0, nrows, P::kSlots, P::slot());
  for (int32_t i = ri.first; i < ri.end; i += ri.stride) {
    const int4 r = rec[P::u(i)];
    const int32_t v = P::u(r.x), e0 = P::u(r.y), e1 = P::u(r.z), e2 = P::u(r.w);
    T acc[VEC];
    VecIO<T, VEC>::ld(srcA + (int64_t)v * k + cl, acc);
    gather_wide<T, VEC, 4>(e0, e1, ecol, eval, srcA, k, cl, true, acc);
    gather_wide<T, VEC, 8>(e1, e2, ecol, eval, Z, k, cl, true, acc);
    if (valid) VecIO<T, VEC>::st(Z + (int64_t)v * k + c0, acc);
  }
}

// rec[i] = {v, seg[sp + a], seg[sp + a + 1], seg[sp + m]}, v = rows[i]
__global__ void k_chain_rec(int32_t nrows, const int32_t *__restrict__ rows, int32_t a,
                            int32_t m, const int32_t *__restrict__ seg_ptr,
                            const int32_t *__restrict__ seg, int4 *__restrict__ rec) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int32_t v = rows[i], sp = seg_ptr[v];
  rec[i] = make_int4(v, seg[sp + a], seg[sp + a + 1], seg[sp + m]);
}

// ---- narrow form (k <= 4): one thread per row, its K columns in registers
// (the row's index entries are read once for all columns) ---------------
// acc[c] += sum_{e in [a,b)} val[e] X[col[e] K + c], CH entries at a time:
// a chunk's indices, then all its gathers, are issued back to back; the
// tail of a chunk re-reads the last entry with weight 0 (fma(0, x, acc) is
// acc), so the sum is the in-order one of the plain loop.
template <typename T, int K, int CH>
__device__ __forceinline__ void gather_rows_n(int32_t a, int32_t b,
                                              const int32_t *__restrict__ col,
                                              const T *__restrict__ val, const T *X,
                                              T (&acc)[K]) {
  for (int32_t e = a; e < b; e += CH) {
    int32_t j[CH];
    T v[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int32_t q = min(e + u, b - 1);
      j[u] = col[q];
      v[u] = (e + u < b) ? val[q] : T(0);
    }
    T x[CH][K];
#pragma unroll
    for (int u = 0; u < CH; ++u)
#pragma unroll
      for (int c = 0; c < K; ++c) x[u][c] = X[(int64_t)j[u] * K + c];
#pragma unroll
    for (int u = 0; u < CH; ++u)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] = fma(v[u], x[u][c], acc[c]);
  }
}

template <typename T, int K>
__global__ void __launch_bounds__(256)
k_hsc_down_n(int32_t nC, const int32_t *__restrict__ Cg, const int32_t *__restrict__ rptr,
             const int32_t *__restrict__ rcol, const T *__restrict__ rval, const T *src, T *Z,
             const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nC) return;
  const int64_t o = (int64_t)Cg[r] * K;
  T acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = src[o + c];
  gather_rows_n<T, K, 4>(rptr[r], rptr[r + 1], rcol, rval, src, acc);
#pragma unroll
  for (int c = 0; c < K; ++c) Z[o + c] = acc[c];
}

template <typename T, int K>
__global__ void __launch_bounds__(256)
k_hsc_up_n(int32_t nF, const int32_t *__restrict__ Fg, const T *__restrict__ dinv,
           const int32_t *__restrict__ pptr, const int32_t *__restrict__ pcol,
           const T *__restrict__ pval, const T *src, T *Z, const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  const int64_t o = (int64_t)Fg[f] * K;
  const T d = dinv[f];
  T acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = d * src[o + c];
  gather_rows_n<T, K, 8>(pptr[f], pptr[f + 1], pcol, pval, Z, acc);
#pragma unroll
  for (int c = 0; c < K; ++c) Z[o + c] = acc[c];
}

// z[rows[i]] = sum_j inv[i][j] W[j] for K <= 4 columns: wavefront i of the
// grid takes coarse row i, lane l the entries j = l, l + 64, ... (summed in
// FP64 over the FP64 inverse or its FP32 rounding, MT), and the lane sums
// meet in a fixed butterfly
template <typename T, typename MT, int K>
__global__ void __launch_bounds__(kBlock)
k_coarse_gemv(int32_t nc, const MT *__restrict__ inv, const T *__restrict__ W,
              const int32_t *__restrict__ rows, T *__restrict__ z,
              const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  const int32_t i = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= nc) return;
  double acc[K] = {};
  const MT *a = inv + (int64_t)i * nc;
  // the lane's entries j = lane, lane + 64, ... in order, 4 of them loaded
  // at a time (clamped into range, the clamped ones multiplied by 0)
  for (int32_t j0 = lane; j0 < nc; j0 += 4 * 64) {
    double m[4];
    T w[4][K];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t j = min(j0 + u * 64, nc - 1);
      const double mj = (double)a[j];
      m[u] = j0 + u * 64 < nc ? mj : 0.0;
#pragma unroll
      for (int c = 0; c < K; ++c) w[u][c] = W[(int64_t)j * K + c];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] += m[u] * (double)w[u][c];
  }
#pragma unroll
  for (int c = 0; c < K; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
  }
  if (lane < K) {
    double v = acc[0];
#pragma unroll
    for (int c = 1; c < K; ++c) v = lane == c ? acc[c] : v;
    z[(int64_t)rows[i] * K + lane] = (T)v;
  }
}

// W[t] = src[rows[t]] (the coarse rows, gathered for the GEMM)
template <typename T>
__global__ void k_gather_rows(int32_t m, const int32_t *__restrict__ rows, int64_t k,
                              const T *__restrict__ src, T *__restrict__ W,
                              const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)m * k) return;
  const int32_t t = (int32_t)(id / k);
  W[id] = src[(int64_t)rows[t] * k + (id - (int64_t)t * k)];
}

template <typename T>
static int hsc_vec(int64_t k) {
  if (k <= 4) return 0;  // scalar form
  if (sizeof(T) == 4) return (k % 4 == 0) ? 4 : ((k % 2 == 0) ? 2 : 1);
  return (k % 2 == 0) ? 2 : 1;
}

// Grid of a level pass: up to 16384 blocks over the column tiles (at least
// 1024 a tile).  A row is a chain of dependent loads, so the passes need
// many waves in flight: raising the cap from 1024 blocks a tile took the
// down pass from 0.57 to 0.62 of peak at k = 768 (3 tiles) and the
// step of the 2.46M-row icosphere from 54.8 to 61.4 poses/s at k = 48.  Totals of
// 8192, 32768 and 65536 blocks measured within 0.3 % of 16384.
static dim3 hsc_grid(int64_t rows, int64_t k, int v, int lpr, int32_t level_cap) {
  const int slots = kWaves * (64 / lpr);
  int64_t g = (rows + slots - 1) / slots;
  const int64_t tiles = (k + (int64_t)lpr * v - 1) / ((int64_t)lpr * v);
  const int64_t total = level_cap > 0 ? std::min<int64_t>(std::max(level_cap, 1024), 65536)
                                      : 16384;
  const int64_t cap = std::max<int64_t>(1024, total / tiles);
  g = std::max<int64_t>(1, std::min<int64_t>(g, cap));
  if (g >= 64) g -= g % 8;
  return dim3((unsigned)g, (unsigned)((k + (int64_t)lpr * v - 1) / ((int64_t)lpr * v)), 1);
}

// launch WIDE(V, L) for the lanes per row L of vector width V
#define HSC_LPR(lpr, WIDE, V)                                                 \
  do {                                                                        \
    if ((lpr) == 12) { WIDE(V, 12); }                                         \
    else if ((lpr) == 16) { WIDE(V, 16); }                                    \
    else if ((lpr) == 32) { WIDE(V, 32); }                                    \
    else { WIDE(V, 64); }                                                     \
  } while (0)

template <typename T> struct LvVals;
template <> struct LvVals<double> {
  static const double *r(const cn_hsc_t::Level &L) { return L.r_val.get(); }
  static const double *p(const cn_hsc_t::Level &L) { return L.p_val.get(); }
  static const double *d(const cn_hsc_t::Level &L) { return L.dinv.get(); }
};
template <> struct LvVals<float> {
  static const float *r(const cn_hsc_t::Level &L) { return L.r_val32.get(); }
  static const float *p(const cn_hsc_t::Level &L) { return L.p_val32.get(); }
  static const float *d(const cn_hsc_t::Level &L) { return L.dinv32.get(); }
};

// narrow form (k <= 4): level l's in-place pass over its kept rows
template <typename T>
static void level_down_n(const cn_hsc_t::Level &L, int64_t k, const T *src, T *Z,
                         const int32_t *ctl, hipStream_t s) {
  if (L.nC == 0) return;
#define N_(KK) k_hsc_down_n<T, KK><<<grid_for(L.nC, 256), 256, 0, s>>>(L.nC, L.Cg.get(), L.r_ptr.get(), L.r_colg.get(), LvVals<T>::r(L), src, Z, ctl)
  switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
}

// rz != nullptr: block b adds its r . z partials to row b of rz for
// b < rz_rows (rows earlier launches wrote) and stores them past it.
// Returns the rows of rz written so far (rz_rows, or the grid's blocks a
// column tile if more).  The narrow form (k <= 4) takes rz = nullptr.
template <typename T>
static int32_t level_up(const cn_hsc_t::Level &L, int64_t k, const T *src, T *Z,
                        const int32_t *ctl, hipStream_t s, int32_t level_cap,
                        double *rz = nullptr, int32_t rz_rows = 0) {
  if (L.nF == 0) return rz_rows;
  const int v = hsc_vec<T>(k);
  if (v == 0) {
#define N_(KK) k_hsc_up_n<T, KK><<<grid_for(L.nF, 256), 256, 0, s>>>(L.nF, L.Fg.get(), LvVals<T>::d(L), L.p_ptr.get(), L.p_colg.get(), LvVals<T>::p(L), src, Z, ctl)
    switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
    return 0;
  }
  const int lpr = pack_lpr12((k + v - 1) / v);
  const dim3 g = hsc_grid(L.nF, k, v, lpr, level_cap);
#define U_(VV, LL) kw_hsc_up<T, VV, LL><<<g, kBlock, 0, s>>>(L.nF, L.Fg.get(), LvVals<T>::d(L), L.p_ptr.get(), L.p_colg.get(), LvVals<T>::p(L), k, src, Z, ctl, rz, rz_rows)
  const int32_t rows = rz ? std::max(rz_rows, (int32_t)g.x) : 0;
  if constexpr (sizeof(T) == 4) {
    if (v == 4) { HSC_LPR(lpr, U_, 4); return rows; }
  }
  if (v >= 2) { HSC_LPR(lpr, U_, 2); } else { HSC_LPR(lpr, U_, 1); }
#undef U_
  return rows;
}

// Partial rows an apply of k columns writes for r . M^-1 r (hsc_apply with
// rz_part), or -1 where it computes none (the narrow form k <= 4, or the
// fused tail).
template <typename T>
int32_t hsc_rz_rows(const cn_hsc_t *H, int64_t k, const HscShape &sh);

// the rows `rows` (level-0 numbers) from level a to the ends of their
// chains at level m
template <typename T>
static void level_chain(const cn_hsc_t *H, const int32_t *rows, int32_t nrows, int32_t a,
                        int32_t m, int64_t k, const T *srcA, T *Z, const int32_t *ctl,
                        hipStream_t s, int32_t level_cap) {
  if (nrows == 0) return;
  const int v = hsc_vec<T>(k);
  const int lpr = pack_lpr12((k + v - 1) / v);
  const dim3 g = hsc_grid(nrows, k, v, lpr, level_cap);
  const T *ev;
  if constexpr (sizeof(T) == 4) ev = H->ch_val32.get(); else ev = H->ch_val.get();
#define C_(VV, LL) kw_hsc_chain<T, VV, LL><<<g, kBlock, 0, s>>>(nrows, rows, a, m, H->ch_seg_ptr.get(), H->ch_seg.get(), H->ch_col.get(), ev, k, srcA, Z, ctl)
  if constexpr (sizeof(T) == 4) {
    if (v == 4) { HSC_LPR(lpr, C_, 4); return; }
  }
  if (v >= 2) { HSC_LPR(lpr, C_, 2); } else { HSC_LPR(lpr, C_, 1); }
#undef C_
}

constexpr int32_t kChainLevels = 3;  // levels a group of down-sweep launches chains through

// the launches of a down sweep to level t in groups of G levels, in order:
// (a, m, rows, row count)
struct ChainLaunch { int32_t a, m; const int32_t *rows; int32_t nrows; };
static std::vector<ChainLaunch> chain_launches(const cn_hsc_t *H, int32_t t, int32_t G) {
  std::vector<ChainLaunch> out;
  for (int32_t a = 0; a < t;) {
    const int32_t e = std::min(a + G, t);
    for (int32_t m = a + 1; m < e; ++m)
      out.push_back({a, m, H->lv[m].Fg.get(), H->lv[m].nF});
    out.push_back({a, e, H->lv[e - 1].Cg.get(), H->lv[e - 1].nC});
    a = e;
  }
  return out;
}

// The row records of the sweep (G, t), built on first use (nullptr when
// may_alloc is false and they are not there yet).
static const cn_hsc_t::ChainRecs *chain_recs(cn_hsc_t *H, int32_t t, int32_t G,
                                             bool may_alloc, hipStream_t s) {
  for (const auto &c : H->ch_recs)
    if (c.G == G && c.t == t) return &c;
  if (!may_alloc) return nullptr;
  const std::vector<ChainLaunch> ls = chain_launches(H, t, G);
  cn_hsc_t::ChainRecs c;
  c.G = G;
  c.t = t;
  c.off.assign(1, 0);
  for (const ChainLaunch &l : ls) c.off.push_back(c.off.back() + l.nrows);
  c.rec.alloc((size_t)std::max<int64_t>(c.off.back(), 1));
  for (size_t i = 0; i < ls.size(); ++i)
    if (ls[i].nrows > 0)
      k_chain_rec<<<grid_for(ls[i].nrows, 256), 256, 0, s>>>(
          ls[i].nrows, ls[i].rows, ls[i].a, ls[i].m, H->ch_seg_ptr.get(), H->ch_seg.get(),
          c.rec.get() + c.off[i]);
  CN_HIP(hipGetLastError());
  H->ch_recs.push_back(std::move(c));
  return &H->ch_recs.back();
}

// one down-sweep launch over row records
template <typename T>
static void level_chain_rec(const cn_hsc_t *H, const int4 *rec, int32_t nrows, int64_t k,
                            const T *srcA, T *Z, const int32_t *ctl, hipStream_t s,
                            int32_t level_cap) {
  if (nrows == 0) return;
  const int v = hsc_vec<T>(k);
  const int lpr = pack_lpr12((k + v - 1) / v);
  const dim3 g = hsc_grid(nrows, k, v, lpr, level_cap);
  const T *ev;
  if constexpr (sizeof(T) == 4) ev = H->ch_val32.get(); else ev = H->ch_val.get();
#define C_(VV, LL) kw_hsc_chain_rec<T, VV, LL><<<g, kBlock, 0, s>>>(nrows, rec, H->ch_col.get(), ev, k, srcA, Z, ctl)
  if constexpr (sizeof(T) == 4) {
    if (v == 4) { HSC_LPR(lpr, C_, 4); return; }
  }
  if (v >= 2) { HSC_LPR(lpr, C_, 2); } else { HSC_LPR(lpr, C_, 1); }
#undef C_
}

// The down sweep up to level t (the coarse level, or the tail's first), in
// groups of chain_levels levels.
template <typename T>
static void down_sweep(cn_hsc_t *H, int32_t t, int64_t k, const T *r, T *z, const int32_t *ctl,
                       hipStream_t s, const HscShape &sh, bool may_alloc) {
  if (hsc_vec<T>(k) == 0) {
    for (int32_t l = 0; l < t; ++l) level_down_n<T>(H->lv[l], k, l == 0 ? r : z, z, ctl, s);
    return;
  }
  if (t == 0) return;
  if (!H->ch_ready) throw HipError{hipErrorNotReady, "hsc: restriction chains not built"};
  const int32_t G = sh.chain_levels > 0 ? sh.chain_levels : kChainLevels;
  if (const cn_hsc_t::ChainRecs *cr = chain_recs(H, t, G, may_alloc, s)) {
    const std::vector<ChainLaunch> ls = chain_launches(H, t, G);
    // (the down sweep's launches take their own grid cap: a group's last
    // launch gathers whole chains and wants more waves in flight than the
    // up passes, whose blocks end with the r.z partial sums)
    const int32_t cap = sh.down_cap > 0 ? sh.down_cap : sh.level_cap;
    for (size_t i = 0; i < ls.size(); ++i)
      level_chain_rec<T>(H, cr->rec.get() + cr->off[i], ls[i].nrows, k,
                         ls[i].a == 0 ? r : z, z, ctl, s, cap);
    return;
  }
  for (int32_t a = 0; a < t;) {
    const int32_t e = std::min(a + G, t);
    const T *srcA = a == 0 ? r : z;
    for (int32_t m = a + 1; m < e; ++m)
      level_chain<T>(H, H->lv[m].Fg.get(), H->lv[m].nF, a, m, k, srcA, z, ctl, s,
                     sh.level_cap);
    level_chain<T>(H, H->lv[e - 1].Cg.get(), H->lv[e - 1].nC, a, e, k, srcA, z, ctl, s,
                   sh.level_cap);
    a = e;
  }
}

// z = M^-1 r (n0 x k), r and z distinct.
// ---- fused tail: the small levels and the coarse solve in one launch ----
// The apply is independent per column, so one workgroup can own a slice of
// VEC columns and run every small level's down pass, the coarse product and
// the up passes for it, with a workgroup barrier between levels instead of
// a kernel boundary (the per-level kernels of levels of a few thousand rows
// are bound by launch and memory latency, not bandwidth).  Slices are dealt
// to XCDs in contiguous column ranges, so an XCD's L2 holds its columns of
// the small levels' rows.  Arithmetic per row is the level kernels': the
// row's own value, then fma over its R (or P) entries in order; the coarse
// product sums the symmetric inverse's column in FP64 in order.
template <typename T>
struct TailLevel {
  int32_t nC, nF;
  const int32_t *Cg, *r_ptr, *r_colg;
  const T *r_val;
  const int32_t *Fg;
  const T *dinv;
  const int32_t *p_ptr, *p_colg;
  const T *p_val;
};

constexpr int kTailBlock = 1024;
// Levels of at most this many rows go to the tail.  A workgroup walks a
// level's rows 1024 at a time, each a chain of three dependent loads, so the
// tail pays off only where a level is a round or two: measured on the 100k
// capsule union at k = 768, thresholds 0 / 1500 / 3000 / 6000 / 12288 rows
// gave tails of 36 / 64 / 104 / 185 / 389 us an apply against 51 / 88 / 120 /
// 155 / 187 us for the same levels and coarse solve run as separate kernels.
constexpr int32_t kTailRows = 1536;
constexpr int32_t kTailCoarse = 1024; // coarse rows the tail's LDS holds
constexpr int kTailSlices = 128;      // fewest column slices (workgroups) worth a tail

template <typename T, int VEC>
__global__ void __launch_bounds__(kTailBlock)
k_hsc_tail(int32_t l0, int32_t l1, const TailLevel<T> *__restrict__ lv, int32_t nc,
           const int32_t *__restrict__ crow, const double *__restrict__ cinv, int64_t k,
           const T *r, T *Z, const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  __shared__ double sw[kTailCoarse * VEC];
  const int32_t nsl = (int32_t)(k / VEC), b = blockIdx.x;
  const int32_t slice = nsl % kXcd == 0 ? (b % kXcd) * (nsl / kXcd) + b / kXcd : b;
  const int64_t c0 = (int64_t)slice * VEC;
  const int t = threadIdx.x;
  for (int32_t l = l0; l < l1; ++l) {
    const TailLevel<T> L = lv[l];
    const T *src = l == 0 ? r : Z;
    for (int32_t c = t; c < L.nC; c += kTailBlock) {
      const int64_t g = (int64_t)L.Cg[c] * k + c0;
      T acc[VEC];
      VecIO<T, VEC>::ld(src + g, acc);
      for (int32_t e = L.r_ptr[c]; e < L.r_ptr[c + 1]; ++e) {
        T x[VEC];
        VecIO<T, VEC>::ld(src + (int64_t)L.r_colg[e] * k + c0, x);
        const T w = L.r_val[e];
#pragma unroll
        for (int m = 0; m < VEC; ++m) acc[m] = fma(w, x[m], acc[m]);
      }
      VecIO<T, VEC>::st(Z + g, acc);
    }
    __syncthreads();
  }
  // coarse: z_c = A_c^-1 z_c on the coarse rows (the inverse is symmetric,
  // so row i is read as column i: coalesced over i)
  const T *csrc = l1 == 0 ? r : Z;
  for (int32_t j = t; j < nc; j += kTailBlock) {
    T x[VEC];
    VecIO<T, VEC>::ld(csrc + (int64_t)crow[j] * k + c0, x);
#pragma unroll
    for (int m = 0; m < VEC; ++m) sw[j * VEC + m] = (double)x[m];
  }
  __syncthreads();
  for (int32_t i = t; i < nc; i += kTailBlock) {
    double acc[VEC] = {};
    int32_t j = 0;
    // 8 column entries in flight at a time (in order, as one sum)
    for (; j + 8 <= nc; j += 8) {
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = cinv[(int64_t)(j + u) * nc + i];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int m = 0; m < VEC; ++m) acc[m] = fma(a[u], sw[(j + u) * VEC + m], acc[m]);
    }
    for (; j < nc; ++j) {
      const double a = cinv[(int64_t)j * nc + i];
#pragma unroll
      for (int m = 0; m < VEC; ++m) acc[m] = fma(a, sw[j * VEC + m], acc[m]);
    }
    T o[VEC];
#pragma unroll
    for (int m = 0; m < VEC; ++m) o[m] = (T)acc[m];
    VecIO<T, VEC>::st(Z + (int64_t)crow[i] * k + c0, o);
  }
  __syncthreads();
  for (int32_t l = l1 - 1; l >= l0; --l) {
    const TailLevel<T> L = lv[l];
    const T *src = l == 0 ? r : Z;
    for (int32_t f = t; f < L.nF; f += kTailBlock) {
      const int64_t g = (int64_t)L.Fg[f] * k + c0;
      T acc[VEC];
      VecIO<T, VEC>::ld(src + g, acc);
      const T d = L.dinv[f];
#pragma unroll
      for (int m = 0; m < VEC; ++m) acc[m] = d * acc[m];
      for (int32_t e = L.p_ptr[f]; e < L.p_ptr[f + 1]; ++e) {
        T x[VEC];
        VecIO<T, VEC>::ld(Z + (int64_t)L.p_colg[e] * k + c0, x);
        const T w = L.p_val[e];
#pragma unroll
        for (int m = 0; m < VEC; ++m) acc[m] = fma(w, x[m], acc[m]);
      }
      VecIO<T, VEC>::st(Z + g, acc);
    }
    __syncthreads();
  }
}

// First level the fused tail takes for an apply with k columns (levels
// [l0, L) and the coarse solve), or -1 when the per-level kernels run.
template <typename T>
static int32_t tail_level(const cn_hsc_t *H, int64_t k, int32_t tail_rows = 0) {
  const int v = hsc_vec<T>(k);
  if (tail_rows < 0 || v == 0 || k / v < kTailSlices || H->nc > kTailCoarse) return -1;
  const int32_t rows = tail_rows > 0 ? std::min(tail_rows, 16384) : kTailRows;
  const int32_t L = (int32_t)H->lv.size();
  int32_t l0 = L;
  while (l0 > 0 && H->lv[l0 - 1].n <= rows) --l0;
  return l0;
}

template <typename T> struct TailIdx;
template <> struct TailIdx<float> { static constexpr int i = 0; };
template <> struct TailIdx<double> { static constexpr int i = 1; };

template <typename T>
static const TailLevel<T> *tail_desc(cn_hsc_t *H, hipStream_t s) {
  DevBuf<unsigned char> &d = H->tail_desc[TailIdx<T>::i];
  if (!d.get()) {
    std::vector<TailLevel<T>> h(H->lv.size() + 1);
    for (size_t l = 0; l < H->lv.size(); ++l) {
      const cn_hsc_t::Level &L = H->lv[l];
      h[l] = TailLevel<T>{L.nC, L.nF, L.Cg.get(), L.r_ptr.get(), L.r_colg.get(),
                          LvVals<T>::r(L), L.Fg.get(), LvVals<T>::d(L), L.p_ptr.get(),
                          L.p_colg.get(), LvVals<T>::p(L)};
    }
    d.alloc(sizeof(TailLevel<T>) * h.size());
    CN_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(TailLevel<T>) * h.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
  }
  return reinterpret_cast<const TailLevel<T> *>(d.get());
}

// Workspace bytes of an apply of k columns: the gathered coarse rows of the
// narrow coarse product (k <= 4), or the per-slice FP64 partial sums of the
// matrix-core product.
template <typename T>
static size_t apply_work_bytes(const cn_hsc_t *H, int64_t k) {
  if (hsc_vec<T>(k) == 0) return (sizeof(T) * (size_t)H->nc * k + 255) / 256 * 256;
  return sizeof(double) * coarse_product_work(H->nc, (int)k) + 256;
}

// ext_work: a caller's workspace of at least apply_work_bytes (the
// hierarchy's own buffer, grown on demand, otherwise)
// rz_part != nullptr (only where hsc_rz_rows is not -1): the apply also
// writes the FP64 partial sums of r . z per column, hsc_rz_rows rows of k:
// the coarse rows' r_c z_c, then every up-pass block's d_f r_f^2 added
// into row blockIdx.x, levels from the coarsest down.
template <typename T>
int32_t hsc_rz_rows(const cn_hsc_t *H, int64_t k, const HscShape &sh) {
  const int v = hsc_vec<T>(k);
  if (v == 0 || tail_level<T>(H, k, sh.tail_rows) >= 0) return -1;
  int32_t rows = H->nc > 0 ? coarse_rz_rows(H->nc) : 0;
  const int lpr = pack_lpr12((k + v - 1) / v);
  for (const cn_hsc_t::Level &L : H->lv)
    if (L.nF > 0) rows = std::max(rows, (int32_t)hsc_grid(L.nF, k, v, lpr, sh.level_cap).x);
  return rows;
}
template int32_t hsc_rz_rows<float>(const cn_hsc_t *, int64_t, const HscShape &);
template int32_t hsc_rz_rows<double>(const cn_hsc_t *, int64_t, const HscShape &);

template <typename T>
void hsc_apply(cn_hsc_t *H, int64_t k, const T *r, T *z, hipStream_t s, const int32_t *ctl,
               const HscShape &sh, unsigned char *ext_work = nullptr,
               double *rz_part = nullptr) {
  const int32_t L = (int32_t)H->lv.size();
  const size_t bytes = apply_work_bytes<T>(H, k);
  unsigned char *work = ext_work;
  if (!work) {
    if (H->work_bytes < bytes) {
      H->work.alloc(bytes);
      H->work_bytes = bytes;
    }
    work = H->work.get();
  }
  T *W = reinterpret_cast<T *>(work);
  const int32_t l0 = tail_level<T>(H, k, sh.tail_rows);
  if (l0 >= 0 || hsc_vec<T>(k) == 0) rz_part = nullptr;  // hsc_rz_rows is -1 there
  int32_t rzr = 0;  // rows of rz_part written so far
  if (l0 >= 0) {
    const TailLevel<T> *desc = tail_desc<T>(H, s);
    {
      ProfScope ps("hsc_down", s);
      down_sweep<T>(H, l0, k, r, z, ctl, s, sh, ext_work == nullptr);
    }
    {
      ProfScope ps("hsc_tail", s);
      const int v = hsc_vec<T>(k);
      const unsigned g = (unsigned)(k / v);
#define T_(VV) k_hsc_tail<T, VV><<<g, kTailBlock, 0, s>>>(l0, L, desc, H->nc, H->coarse_rows.get(), H->coarse_inv.get(), k, r, z, ctl)
      if constexpr (sizeof(T) == 4) {
        if (v == 4) T_(4); else if (v == 2) T_(2); else T_(1);
      } else {
        if (v == 2) T_(2); else T_(1);
      }
#undef T_
    }
    {
      ProfScope ps("hsc_up", s);
      for (int32_t l = l0 - 1; l >= 0; --l)
        level_up<T>(H->lv[l], k, l == 0 ? r : z, z, ctl, s, sh.level_cap);
    }
    return;
  }
  {
    ProfScope ps("hsc_down", s);
    down_sweep<T>(H, L, k, r, z, ctl, s, sh, ext_work == nullptr);
  }
  if (H->nc > 0) {
    ProfScope ps("hsc_coarse", s);
    if (k <= 4) {
      // a few columns: one wavefront per coarse row beats padded MFMA tiles
      const int64_t tot = (int64_t)H->nc * k;
      k_gather_rows<T><<<grid_for(tot, 256), 256, 0, s>>>(H->nc, H->coarse_rows.get(), k,
                                                        L == 0 ? r : z, W, ctl);
      // (coarse_prec 2: the FP32 rounding of the inverse in FP32 applies,
      // half the bytes of the product's one pass over it)
      const unsigned gg = (unsigned)((H->nc + kWaves - 1) / kWaves);
      if (sizeof(T) == 4 && sh.coarse_prec >= 2) {
#define G_(KK)                                                                  \
  k_coarse_gemv<T, float, KK><<<gg, kBlock, 0, s>>>(H->nc, H->coarse_inv32.get(), W, \
                                                    H->coarse_rows.get(), z, ctl)
        switch (k) { case 1: G_(1); break; case 2: G_(2); break; case 3: G_(3); break; default: G_(4); }
#undef G_
      } else {
#define G_(KK)                                                                   \
  k_coarse_gemv<T, double, KK><<<gg, kBlock, 0, s>>>(H->nc, H->coarse_inv.get(), W, \
                                                     H->coarse_rows.get(), z, ctl)
        switch (k) { case 1: G_(1); break; case 2: G_(2); break; case 3: G_(3); break; default: G_(4); }
#undef G_
      }
    } else {
      // coarse_prec 3: the split-bfloat16 product where it runs, else FP32
      const bool bf3 = sh.coarse_prec == 3 && H->coarse_bh.get() &&
                       coarse_bf3_applies(H->nc, (int)k);
      if constexpr (sizeof(T) == 4) {
        if (bf3) {
          rzr = coarse_product_bf3(H->nc, H->coarse_ld, (int)k, H->coarse_bh.get(),
                                   H->coarse_bl.get(), H->coarse_rows.get(), L == 0 ? r : z,
                                   z, ctl, reinterpret_cast<double *>(work), s, rz_part);
        }
      }
      if (sizeof(T) != 4 || !bf3) {
        const float *M32 =
            sizeof(T) == 4 && sh.coarse_prec >= 2 ? H->coarse_inv32.get() : nullptr;
        rzr = coarse_product<T>(H->nc, (int)k, H->coarse_inv.get(), M32,
                                H->coarse_rows.get(), L == 0 ? r : z, z, ctl,
                                reinterpret_cast<double *>(work), s, rz_part);
      }
    }
  }
  {
    ProfScope ps("hsc_up", s);
    for (int32_t l = L - 1; l >= 0; --l)
      rzr = level_up<T>(H->lv[l], k, l == 0 ? r : z, z, ctl, s, sh.level_cap, rz_part, rzr);
  }
}
template void hsc_apply<float>(cn_hsc_t *, int64_t, const float *, float *, hipStream_t,
                               const int32_t *, const HscShape &, unsigned char *, double *);
template void hsc_apply<double>(cn_hsc_t *, int64_t, const double *, double *, hipStream_t,
                                const int32_t *, const HscShape &, unsigned char *,
                                double *);

// ---- workspace of the apply that allocates nothing (cn_hsc_apply_noalloc)
// A fixed array in device memory, one copy per GPU the library runs on,
// held until the process ends.  Its host-side name is not a device
// address: kernels get the address hipGetSymbolAddress returns for the
// current device, looked up once per device (at hierarchy build, or at the
// first call on a device) and kept.
static __device__ __align__(256) unsigned char scratch[CN_HSC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch` (nullptr and the error set on failure)
unsigned char *hsc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("hsc scratch: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hsc scratch: hipGetSymbolAddress failed (%d: %s)", (int)e,
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  return g_scratch_ptr[dev];
}

// Everything an apply reads besides its workspace, built at set-up: the
// fused tail's level descriptors (FP32 and FP64), the restriction chains
// and this device's scratch address, so cn_hsc_apply_noalloc never
// allocates.
// The restriction chains (cn_hsc_t::ch_*), from the levels' R on the device.
static void build_chains(cn_hsc_t *H, hipStream_t s) {
  const int32_t L = (int32_t)H->lv.size(), n0 = H->n0;
  if (L == 0 || n0 == 0) return;
  auto down = [&](const auto &d, size_t count, auto &h) {
    h.resize(count);
    if (count)
      CN_HIP(hipMemcpyAsync(h.data(), d.get(), sizeof(h[0]) * count, hipMemcpyDeviceToHost, s));
  };
  std::vector<std::vector<int32_t>> Fg(L), Cg(L), rp(L), rc(L);
  std::vector<std::vector<double>> rv(L);
  for (int32_t l = 0; l < L; ++l) {
    const cn_hsc_t::Level &Lv = H->lv[l];
    down(Lv.Fg, (size_t)Lv.nF, Fg[l]);
    down(Lv.Cg, (size_t)Lv.nC, Cg[l]);
    down(Lv.r_ptr, (size_t)Lv.nC + 1, rp[l]);
  }
  CN_HIP(hipStreamSynchronize(s));
  for (int32_t l = 0; l < L; ++l) {
    down(H->lv[l].r_colg, (size_t)rp[l][H->lv[l].nC], rc[l]);
    down(H->lv[l].r_val, (size_t)rp[l][H->lv[l].nC], rv[l]);
  }
  CN_HIP(hipStreamSynchronize(s));
  // level of elimination (L: kept to the coarse level) and entry counts
  std::vector<int32_t> cls((size_t)n0, L);
  for (int32_t l = 0; l < L; ++l)
    for (int32_t v : Fg[l]) cls[v] = l;
  std::vector<int64_t> cnt((size_t)n0, 0);
  for (int32_t l = 0; l < L; ++l)
    for (int32_t c = 0; c < H->lv[l].nC; ++c) cnt[Cg[l][c]] += rp[l][c + 1] - rp[l][c];
  // storage order: by level of elimination, then in that level's F order
  // (the coarse rows in the last level's C order)
  std::vector<int32_t> order;
  order.reserve((size_t)n0);
  for (int32_t l = 1; l < L; ++l) order.insert(order.end(), Fg[l].begin(), Fg[l].end());
  order.insert(order.end(), Cg[L - 1].begin(), Cg[L - 1].end());
  std::vector<int32_t> seg_ptr((size_t)n0, 0);
  std::vector<int64_t> start((size_t)n0, 0);
  int64_t nseg = 0, nent = 0;
  for (int32_t v : order) {
    seg_ptr[v] = (int32_t)nseg;
    start[v] = nent;
    nseg += cls[v] + 1;
    nent += cnt[v];
  }
  if (nseg > INT32_MAX || nent > INT32_MAX)
    throw HipError{hipErrorInvalidValue, "hsc chains exceed int32"};
  std::vector<int32_t> seg((size_t)nseg + 1, 0), col((size_t)nent + 1, 0);
  std::vector<double> val((size_t)nent + 1, 0.0);
  std::vector<int64_t> cur = start;
  for (int32_t l = 0; l < L; ++l)
    for (int32_t c = 0; c < H->lv[l].nC; ++c) {
      const int32_t v = Cg[l][c];
      seg[(size_t)seg_ptr[v] + l] = (int32_t)cur[v];
      for (int32_t e = rp[l][c]; e < rp[l][c + 1]; ++e) {
        col[(size_t)cur[v]] = rc[l][e];
        val[(size_t)cur[v]] = rv[l][e];
        ++cur[v];
      }
    }
  for (int32_t v : order) seg[(size_t)seg_ptr[v] + cls[v]] = (int32_t)cur[v];
  std::vector<float> val32(val.size());
  for (size_t i = 0; i < val.size(); ++i) val32[i] = (float)val[i];
  auto up = [&](auto &d, const auto &h) {
    d.alloc(h.size());
    CN_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(h[0]) * h.size(), hipMemcpyHostToDevice, s));
  };
  up(H->ch_seg_ptr, seg_ptr);
  up(H->ch_seg, seg);
  up(H->ch_col, col);
  up(H->ch_val, val);
  up(H->ch_val32, val32);
  CN_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  H->ch_ready = true;
}

__global__ void k_round_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

void hsc_prepare(cn_hsc_t *H, hipStream_t s) {
  if (H->nc > 0) {
    const int64_t nn = (int64_t)H->nc * H->nc;
    H->coarse_inv32.alloc((size_t)nn);
    k_round_f32<<<grid_for(nn, 256), 256, 0, s>>>(nn, H->coarse_inv.get(),
                                                 H->coarse_inv32.get());
    CN_HIP(hipGetLastError());
    if (H->nc >= 3072) {
      H->coarse_ld = (H->nc + 7) / 8 * 8;
      const size_t nb = (size_t)H->nc * H->coarse_ld;
      H->coarse_bh.alloc(nb);
      H->coarse_bl.alloc(nb);
      coarse_split_bf16(H->nc, H->coarse_ld, H->coarse_inv.get(), H->coarse_bh.get(),
                        H->coarse_bl.get(), s);
    }
  }
  (void)tail_desc<float>(H, s);
  (void)tail_desc<double>(H, s);
  build_chains(H, s);
  (void)hsc_scratch();
}

}  // namespace cn

using namespace cn;

extern "C" {

int32_t cn_hsc_tail_level(cn_hsc H, int32_t k, int32_t fp32) {
  if (!H || k <= 0) return -1;
  return fp32 ? tail_level<float>(H, k) : tail_level<double>(H, k);
}

cn_status cn_hsc_apply_noalloc(cn_hsc H, int32_t k, const double *d_r, double *d_z,
                               void *stream) {
  CN_API_BEGIN
  if (!H || k <= 0 || !d_r || !d_z || d_r == d_z) {
    set_error("cn_hsc_apply_noalloc: bad arguments");
    return CN_BAD_ARG;
  }
  if (H->n0 == 0) return CN_OK;
  const size_t need = apply_work_bytes<double>(H, k);
  if (need > CN_HSC_SCRATCH_BYTES) {
    set_error("cn_hsc_apply_noalloc: %d columns need %zu bytes of workspace, more than the "
              "%u-byte scratch (coarse rows %d); use cn_hsc_apply or fewer columns",
              (int)k, need, (unsigned)CN_HSC_SCRATCH_BYTES, (int)H->nc);
    return CN_BAD_ARG;
  }
  if (!H->tail_desc[1].get()) {
    set_error("cn_hsc_apply_noalloc: the hierarchy's tail descriptors are missing");
    return CN_BAD_ARG;
  }
  unsigned char *w = hsc_scratch();
  if (!w) return CN_HIP_ERROR;
  hsc_apply<double>(H, k, d_r, d_z, as_stream(stream), nullptr, HscShape{}, w);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_hsc_apply(cn_hsc H, int32_t k, const double *d_r, double *d_z, void *stream) {
  CN_API_BEGIN
  if (!H || k <= 0 || !d_r || !d_z || d_r == d_z) {
    set_error("cn_hsc_apply: bad arguments");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (H->n0 > 0) hsc_apply<double>(H, k, d_r, d_z, s, nullptr, HscShape{});
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

}  // extern "C"
