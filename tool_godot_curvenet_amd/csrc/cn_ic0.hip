// cn_ic0.hip -- no-fill incomplete Cholesky on gfx950.
//
// The factor follows IncompleteCholesky.factor3's recurrences
// (IncompleteCholesky.lean:49-69) in CSR form (:26-33):
//   L_ij = (A_ij - sum_{k<j} L_ik L_jk) / L_jj,  L_ii = sqrt(A_ii - sum L_ik^2)
// with each sum taken over k ascending, FP64, no fused multiply-add, so a
// factor computed here equals the oracle's bit for bit on the same ordering.
//
// Scheduling: rows are renumbered so that every dependency level is a
// contiguous row range.  CN_ORDER_COLOR uses GraphColoring.greedyColor
// (GraphColoring.lean:41-56) -- no two rows of a colour are adjacent, so a
// colour is a level and the whole triangular solve is 2*C passes (C ~ 4-7 on
// triangle meshes), the colour-pass schedule the reference planned for its
// IncompleteCholesky kernel (SlangCodegen/IncompleteCholesky.lean:10-12).
// CN_ORDER_NATURAL keeps the natural order and groups rows by dependency
// depth (more passes, stronger preconditioner).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "cn_internal.h"

#pragma clang fp contract(off)

namespace cn {

__global__ void k_perm_count(int32_t n, const int32_t *perm, const int32_t *ptr,
                             int32_t *cnt) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t o = perm[i];
  cnt[i] = ptr[o + 1] - ptr[o];
}

__global__ void k_perm_fill(int32_t n, const int32_t *perm, const int32_t *iperm,
                            const int32_t *ptr, const int32_t *col,
                            const double *val, const int32_t *optr,
                            int32_t *ocol, double *oval) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t o = perm[i], base = optr[i], m = ptr[o + 1] - ptr[o];
  for (int32_t t = 0; t < m; ++t) {
    int32_t c = iperm[col[ptr[o] + t]];
    double v = val[ptr[o] + t];
    int32_t q = base + t - 1;
    while (q >= base && ocol[q] > c) { ocol[q + 1] = ocol[q]; oval[q + 1] = oval[q]; --q; }
    ocol[q + 1] = c;
    oval[q + 1] = v;
  }
}

__global__ void k_tri_count(int32_t n, const int32_t *ptr, const int32_t *col,
                            int32_t *lc, int32_t *uc, int32_t *nodiag) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t l = 0, u = 0, d = 0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    int32_t c = col[p];
    if (c < i) ++l; else if (c > i) ++u; else d = 1;
  }
  lc[i] = l;
  uc[i] = u;
  if (!d) atomicMin(nodiag, i);
}

__global__ void k_tri_fill(int32_t n, const int32_t *ptr, const int32_t *col,
                           const double *val, const int32_t *lp, int32_t *lcol,
                           double *lval, double *adiag, const int32_t *up,
                           int32_t *ucol) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t a = lp[i], b = up[i];
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    int32_t c = col[p];
    if (c < i) { lcol[a] = c; lval[a] = val[p]; ++a; }
    else if (c > i) { ucol[b++] = c; }
    else adiag[i] = val[p];
  }
}

// u_map[e] = index in L of entry (row = ucol[e], col = i)
__global__ void k_u_map(int32_t n, const int32_t *up, const int32_t *ucol,
                        const int32_t *lp, const int32_t *lcol, int32_t *umap,
                        int32_t *asym) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t e = up[i]; e < up[i + 1]; ++e) {
    int32_t r = ucol[e];
    int32_t a = lp[r], b = lp[r + 1], end = b;
    while (a < b) {
      int32_t m = (a + b) >> 1;
      if (lcol[m] < i) a = m + 1; else b = m;
    }
    if (a >= end || lcol[a] != i) {  // pattern not symmetric
      atomicOr(asym, 1);
      a = lp[r];
    }
    umap[e] = a;
  }
}

// One level of the factorisation: the blocks [b0, b1) depend only on earlier
// levels and on earlier rows of themselves; one thread per block walks its
// rows in order.
__global__ void k_ic0_blocks(int32_t b0, int32_t b1, const int32_t *blk_ptr,
                             const int32_t *lp, const int32_t *lcol, double *lval,
                             const double *adiag, double *ldiag, double shift,
                             int32_t *fail) {
  int32_t b = b0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= b1) return;
  for (int32_t i = blk_ptr[b]; i < blk_ptr[b + 1]; ++i) {
    int32_t a0 = lp[i], a1 = lp[i + 1];
    for (int32_t a = a0; a < a1; ++a) {
      int32_t j = lcol[a];
      double t = lval[a];
      int32_t c = a0, q = lp[j], q1 = lp[j + 1];
      while (c < a && q < q1) {
        int32_t kc = lcol[c], kq = lcol[q];
        if (kc == kq) { t = t - lval[c] * lval[q]; ++c; ++q; }
        else if (kc < kq) ++c;
        else ++q;
      }
      lval[a] = t / ldiag[j];
    }
    double sd = adiag[i];
    if (shift != 0.0) sd = sd + shift * adiag[i];
    for (int32_t a = a0; a < a1; ++a) sd = sd - lval[a] * lval[a];
    if (!(sd > 0.0)) {
      atomicMin(fail, i);
      ldiag[i] = 1.0;
    } else {
      ldiag[i] = sqrt(sd);
    }
  }
}

__global__ void k_gather_u(int64_t nnz, const int32_t *umap, const double *lval,
                           double *uval) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) uval[e] = lval[umap[e]];
}

__global__ void k_to_f32(int64_t n, const double *x, float *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

__global__ void k_rdiag(int32_t n, const double *d, float *r32, double *r64) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = 1.0 / d[i];
  r32[i] = (float)r;
  r64[i] = r;
}

__global__ void k_renumber(int64_t nnz, const int32_t *perm, const int32_t *col,
                           int32_t *out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) out[e] = perm[col[e]];
}

// Exact (spec-order) sweeps for cn_ic0_apply: thread per (block, column),
// the block's rows in order (forward) or reverse order (backward).
__global__ void k_fwd_exact(int32_t b0, int32_t b1, int32_t k, const int32_t *blk_ptr,
                            const int32_t *lp, const int32_t *lcol, const double *lval,
                            const double *ldiag, const double *b, double *y) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)(b1 - b0) * k) return;
  int32_t blk = b0 + (int32_t)(id / k), c = (int32_t)(id % k);
  for (int32_t i = blk_ptr[blk]; i < blk_ptr[blk + 1]; ++i) {
    double sum = b[(int64_t)i * k + c];
    for (int32_t p = lp[i]; p < lp[i + 1]; ++p)
      sum = sum - lval[p] * y[(int64_t)lcol[p] * k + c];
    y[(int64_t)i * k + c] = sum / ldiag[i];
  }
}

__global__ void k_bwd_exact(int32_t b0, int32_t b1, int32_t k, const int32_t *blk_ptr,
                            const int32_t *up, const int32_t *ucol, const double *uval,
                            const double *ldiag, const double *y, double *x) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)(b1 - b0) * k) return;
  int32_t blk = b0 + (int32_t)(id / k), c = (int32_t)(id % k);
  for (int32_t i = blk_ptr[blk + 1] - 1; i >= blk_ptr[blk]; --i) {
    double sum = y[(int64_t)i * k + c];
    for (int32_t p = up[i]; p < up[i + 1]; ++p)
      sum = sum - uval[p] * x[(int64_t)ucol[p] * k + c];
    x[(int64_t)i * k + c] = sum / ldiag[i];
  }
}

__global__ void k_gather_rows(int32_t n, int32_t k, const int32_t *perm,
                              const double *src, double *dst) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * k) return;
  int64_t i = id / k, c = id % k;
  dst[id] = src[(int64_t)perm[i] * k + c];
}

__global__ void k_scatter_rows(int32_t n, int32_t k, const int32_t *perm,
                               const double *src, double *dst) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * k) return;
  int64_t i = id / k, c = id % k;
  dst[(int64_t)perm[i] * k + c] = src[id];
}

// GraphColoring.greedyColor (GraphColoring.lean:41-56): vertex order,
// smallest colour not used by an earlier neighbour.
int32_t greedy_color(int32_t n, const std::vector<int32_t> &ptr,
                     const std::vector<int32_t> &col, std::vector<int32_t> &color) {
  color.assign(n, 0);
  std::vector<int32_t> mark;
  int32_t ncol = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t deg = ptr[i + 1] - ptr[i];
    if ((int32_t)mark.size() < deg + 2) mark.resize(deg + 2, -1);
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      int32_t j = col[p];
      if (j < i && color[j] <= deg) mark[color[j]] = i;
    }
    int32_t c = 0;
    while (c <= deg && mark[c] == i) ++c;
    color[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

static int32_t natural_levels(int32_t n, const std::vector<int32_t> &ptr,
                              const std::vector<int32_t> &col,
                              std::vector<int32_t> &level) {
  level.assign(n, 0);
  int32_t nl = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t l = 0;
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      int32_t j = col[p];
      if (j < i) l = std::max(l, level[j] + 1);
    }
    level[i] = l;
    nl = std::max(nl, l + 1);
  }
  return nl;
}

// Block multicolour ordering: compact BFS patches of up to `bs` rows grown
// from the lowest unassigned row, patches coloured with greedyColor
// (GraphColoring.lean:41-56) on the patch adjacency graph.
static int32_t block_levels(int32_t n, const std::vector<int32_t> &ptr,
                            const std::vector<int32_t> &col, int32_t bs,
                            std::vector<int32_t> &blk, std::vector<int32_t> &bcolor) {
  blk.assign(n, -1);
  std::vector<int32_t> queue;
  queue.reserve(bs);
  int32_t nb = 0;
  for (int32_t s0 = 0; s0 < n; ++s0) {
    if (blk[s0] >= 0) continue;
    queue.clear();
    queue.push_back(s0);
    blk[s0] = nb;
    int32_t cnt = 1;
    for (size_t h = 0; h < queue.size() && cnt < bs; ++h) {
      int32_t u = queue[h];
      for (int32_t p = ptr[u]; p < ptr[u + 1] && cnt < bs; ++p) {
        int32_t w = col[p];
        if (blk[w] < 0) { blk[w] = nb; ++cnt; queue.push_back(w); }
      }
    }
    ++nb;
  }
  // patch graph (CSR, sorted, no self loops)
  std::vector<int64_t> keys;
  keys.reserve(ptr[n]);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      int32_t a = blk[i], b = blk[col[p]];
      if (a != b) keys.push_back(((int64_t)a << 32) | (uint32_t)b);
    }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  std::vector<int32_t> gptr(nb + 1, 0), gcol(keys.size());
  for (size_t e = 0; e < keys.size(); ++e) {
    gptr[(keys[e] >> 32) + 1]++;
    gcol[e] = (int32_t)(keys[e] & 0xffffffff);
  }
  for (int32_t b = 0; b < nb; ++b) gptr[b + 1] += gptr[b];
  return greedy_color(nb, gptr, gcol, bcolor);
}

// ---- triangles prepared for the solver's block sweeps -------------------
// position -> its block and the block's [first, end) positions
__global__ void k_pos_block(int32_t nblk, const int32_t *__restrict__ blk_ptr,
                            int32_t *__restrict__ pblk, int32_t *__restrict__ pb0,
                            int32_t *__restrict__ pb1) {
  int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int32_t p0 = blk_ptr[b], p1 = blk_ptr[b + 1];
  for (int32_t p = p0; p < p1; ++p) {
    pblk[p] = b;
    pb0[p] = p0;
    pb1[p] = p1;
  }
}

// Columns are positions.  An L entry (column < row) is inside the row's
// block when its column >= the block's first position; a U entry (column >
// row) when its column < the block's end.
__device__ __forceinline__ bool in_block(bool upper, int32_t q, int32_t p0, int32_t p1) {
  return upper ? (q < p1) : (q >= p0);
}

// entries in other blocks beyond the fixed-width table
__global__ void k_sweep_count(int32_t n, bool upper, const int32_t *__restrict__ tp,
                              const int32_t *__restrict__ tc, const int32_t *__restrict__ pb0,
                              const int32_t *__restrict__ pb1, int32_t *__restrict__ ovn) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t nx = 0;
  for (int32_t e = tp[i]; e < tp[i + 1]; ++e)
    if (!in_block(upper, tc[e], pb0[i], pb1[i])) ++nx;
  ovn[i] = nx > cn_ic0_t::kSweepEll ? nx - cn_ic0_t::kSweepEll : 0;
}

__global__ void k_sweep_fill(int32_t n, bool upper, int32_t mb, const int32_t *__restrict__ tp,
                             const int32_t *__restrict__ tc, const double *__restrict__ tv,
                             const int32_t *__restrict__ perm, const int32_t *__restrict__ pblk,
                             const int32_t *__restrict__ pb0, const int32_t *__restrict__ pb1,
                             const double *__restrict__ rdiag, int32_t *__restrict__ ell_col,
                             int32_t *__restrict__ ell_cnt, double *__restrict__ ell_val,
                             float *__restrict__ ell_val32, const int32_t *__restrict__ ov_ptr,
                             int32_t *__restrict__ ov_col, double *__restrict__ ov_val,
                             float *__restrict__ ov_val32, double *__restrict__ dense,
                             float *__restrict__ dense32) {
  constexpr int W = cn_ic0_t::kSweepEll;
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t p0 = pb0[i], p1 = pb1[i], t = i - p0;
  const int64_t base = ((int64_t)pblk[i] * mb + t) * mb;
  int32_t nx = 0, ov = ov_ptr[i];
  for (int32_t e = tp[i]; e < tp[i + 1]; ++e) {
    const int32_t q = tc[e];
    const double v = tv[e];
    if (in_block(upper, q, p0, p1)) {
      dense[base + (q - p0)] = v;
      dense32[base + (q - p0)] = (float)v;
    } else if (nx < W) {
      ell_col[(int64_t)i * W + nx] = perm[q];
      ell_val[(int64_t)i * W + nx] = v;
      ell_val32[(int64_t)i * W + nx] = (float)v;
      ++nx;
    } else {
      ov_col[ov] = perm[q];
      ov_val[ov] = v;
      ov_val32[ov] = (float)v;
      ++ov;
    }
  }
  ell_cnt[i] = nx;
  const int32_t pad = nx > 0 ? ell_col[(int64_t)i * W + nx - 1] : 0;
  for (int32_t u = nx; u < W; ++u) {
    ell_col[(int64_t)i * W + u] = pad;
    ell_val[(int64_t)i * W + u] = 0.0;
    ell_val32[(int64_t)i * W + u] = 0.0f;
  }
  dense[base + t] = rdiag[i];
  dense32[base + t] = (float)rdiag[i];
}

static void build_sweep(cn_ic0_t::Sweep &S, int32_t n, int32_t nblk, int32_t mb, bool upper,
                        const int32_t *tp, const int32_t *tc, const double *tv,
                        const int32_t *perm, const int32_t *pblk, const int32_t *pb0,
                        const int32_t *pb1, const double *rdiag, hipStream_t s) {
  constexpr int W = cn_ic0_t::kSweepEll;
  S.ov_ptr.alloc(n + 1);
  CN_HIP(hipMemsetAsync(S.ov_ptr.get(), 0, sizeof(int32_t) * (n + 1), s));
  if (n > 0)
    k_sweep_count<<<grid_for(n, 256), 256, 0, s>>>(n, upper, tp, tc, pb0, pb1,
                                                  S.ov_ptr.get());
  const int64_t nov = exclusive_scan_i32(S.ov_ptr.get(), (int64_t)n + 1, s);
  S.ell_col.alloc((size_t)n * W + 1);
  S.ell_cnt.alloc(n + 1);
  S.ell_val.alloc((size_t)n * W + 1);
  S.ell_val32.alloc((size_t)n * W + 1);
  S.ov_col.alloc(nov + 1);
  S.ov_val.alloc(nov + 1);
  S.ov_val32.alloc(nov + 1);
  const size_t nd = (size_t)nblk * mb * mb;
  S.dense.alloc(nd + 1);
  S.dense32.alloc(nd + 1);
  CN_HIP(hipMemsetAsync(S.dense.get(), 0, sizeof(double) * (nd + 1), s));
  CN_HIP(hipMemsetAsync(S.dense32.get(), 0, sizeof(float) * (nd + 1), s));
  if (n > 0)
    k_sweep_fill<<<grid_for(n, 256), 256, 0, s>>>(
        n, upper, mb, tp, tc, tv, perm, pblk, pb0, pb1, rdiag, S.ell_col.get(),
        S.ell_cnt.get(), S.ell_val.get(), S.ell_val32.get(), S.ov_ptr.get(), S.ov_col.get(),
        S.ov_val.get(), S.ov_val32.get(), S.dense.get(), S.dense32.get());
  CN_HIP(hipGetLastError());
}

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_ic0_factor(cn_system S, int32_t ordering, int32_t allow_shift,
                        cn_ic0 *out, void *stream) {
  cn_ic0_opts o;
  o.ordering = ordering;
  o.block_size = 16;
  o.allow_shift = allow_shift;
  return cn_ic0_factor_ex(S, &o, out, stream);
}

cn_status cn_ic0_factor_ex(cn_system S, const cn_ic0_opts *opt, cn_ic0 *out,
                           void *stream) {
  CN_API_BEGIN
  if (!S || !out || !opt ||
      (opt->ordering != CN_ORDER_COLOR && opt->ordering != CN_ORDER_NATURAL &&
       opt->ordering != CN_ORDER_BLOCK) ||
      (opt->ordering == CN_ORDER_BLOCK && (opt->block_size < 1 || opt->block_size > 4096))) {
    set_error("cn_ic0_factor: bad arguments");
    return CN_BAD_ARG;
  }
  const int32_t ordering = opt->ordering;
  const int32_t allow_shift = opt->allow_shift;
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  const int32_t n = S->n_free;
  std::vector<int32_t> ptr(n + 1), col(S->nnz_ff);
  CN_HIP(hipMemcpyAsync(ptr.data(), S->ff_ptr.get(), sizeof(int32_t) * (n + 1),
                        hipMemcpyDeviceToHost, s));
  if (S->nnz_ff > 0)
    CN_HIP(hipMemcpyAsync(col.data(), S->ff_col.get(), sizeof(int32_t) * S->nnz_ff,
                          hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  // per row: level and block id; rows sort by (level, block, row)
  std::vector<int32_t> lev, blk;
  int32_t nl = 0;
  if (ordering == CN_ORDER_BLOCK) {
    std::vector<int32_t> bcolor;
    nl = block_levels(n, ptr, col, opt->block_size, blk, bcolor);
    lev.resize(n);
    for (int32_t i = 0; i < n; ++i) lev[i] = bcolor[blk[i]];
  } else {
    nl = ordering == CN_ORDER_COLOR ? greedy_color(n, ptr, col, lev)
                                    : natural_levels(n, ptr, col, lev);
    blk.resize(n);
    for (int32_t i = 0; i < n; ++i) blk[i] = i;
  }
  cn_ic0_t *F = new cn_ic0_t();
  try {
    F->n = n;
    F->ordering = ordering;
    F->ncolors = nl;
    F->block_size = ordering == CN_ORDER_BLOCK ? opt->block_size : 1;
    std::vector<int32_t> perm(n);
    for (int32_t i = 0; i < n; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
      if (lev[a] != lev[b]) return lev[a] < lev[b];
      if (blk[a] != blk[b]) return blk[a] < blk[b];
      return a < b;
    });
    F->level_ptr.assign(nl + 1, 0);
    F->level_blk.assign(nl + 1, 0);
    std::vector<int32_t> bptr;
    bptr.reserve(n + 1);
    for (int32_t q = 0; q < n; ++q) {
      int32_t r = perm[q];
      F->level_ptr[lev[r] + 1]++;
      if (q == 0 || blk[perm[q - 1]] != blk[r]) {
        bptr.push_back(q);
        F->level_blk[lev[r] + 1]++;
      }
    }
    bptr.push_back(n);
    for (int32_t c = 0; c < nl; ++c) {
      F->level_ptr[c + 1] += F->level_ptr[c];
      F->level_blk[c + 1] += F->level_blk[c];
    }
    const int32_t nblk = (int32_t)bptr.size() - 1;
    for (int32_t b = 0; b < nblk; ++b) F->max_block = std::max(F->max_block, bptr[b + 1] - bptr[b]);
    F->blk_ptr.alloc((size_t)nblk + 1);
    CN_HIP(hipMemcpyAsync(F->blk_ptr.get(), bptr.data(), sizeof(int32_t) * (nblk + 1),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> iperm(n);
    for (int32_t i = 0; i < n; ++i) iperm[perm[i]] = i;
    F->perm.alloc(n + 1);
    F->iperm.alloc(n + 1);
    CN_HIP(hipMemcpyAsync(F->perm.get(), perm.data(), sizeof(int32_t) * n,
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipMemcpyAsync(F->iperm.get(), iperm.data(), sizeof(int32_t) * n,
                          hipMemcpyHostToDevice, s));
    const int64_t nnz = S->nnz_ff;
    F->a_ptr.alloc(n + 1);
    F->a_col.alloc(nnz + 1);
    F->a_val.alloc(nnz + 1);
    F->a_val32.alloc(nnz + 1);
    CN_HIP(hipMemsetAsync(F->a_ptr.get(), 0, sizeof(int32_t) * (n + 1), s));
    if (n > 0)
      k_perm_count<<<grid_for(n, 256), 256, 0, s>>>(n, F->perm.get(),
                                                   S->ff_ptr.get(), F->a_ptr.get());
    exclusive_scan_i32(F->a_ptr.get(), (int64_t)n + 1, s);
    if (n > 0)
      k_perm_fill<<<grid_for(n, 128), 128, 0, s>>>(
          n, F->perm.get(), F->iperm.get(), S->ff_ptr.get(), S->ff_col.get(),
          S->ff_val.get(), F->a_ptr.get(), F->a_col.get(), F->a_val.get());
    if (nnz > 0)
      k_to_f32<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, F->a_val.get(), F->a_val32.get());
    // split into L (strict lower) / U (strict upper) / diagonal
    F->l_ptr.alloc(n + 1);
    F->u_ptr.alloc(n + 1);
    DevBuf<int32_t> nodiag(1);
    int32_t big = INT_MAX;
    CN_HIP(hipMemcpyAsync(nodiag.get(), &big, sizeof(int32_t), hipMemcpyHostToDevice, s));
    CN_HIP(hipMemsetAsync(F->l_ptr.get(), 0, sizeof(int32_t) * (n + 1), s));
    CN_HIP(hipMemsetAsync(F->u_ptr.get(), 0, sizeof(int32_t) * (n + 1), s));
    if (n > 0)
      k_tri_count<<<grid_for(n, 256), 256, 0, s>>>(n, F->a_ptr.get(), F->a_col.get(),
                                                  F->l_ptr.get(), F->u_ptr.get(),
                                                  nodiag.get());
    int64_t nl_nz = exclusive_scan_i32(F->l_ptr.get(), (int64_t)n + 1, s);
    int64_t nu_nz = exclusive_scan_i32(F->u_ptr.get(), (int64_t)n + 1, s);
    int32_t hnd = 0;
    CN_HIP(hipMemcpyAsync(&hnd, nodiag.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    if (hnd != INT_MAX) {
      delete F;
      set_error("cn_ic0_factor: row %d has no diagonal entry", hnd);
      return CN_NOT_SPD;
    }
    F->nnz_l = nl_nz + n;
    F->l_col.alloc(nl_nz + 1);
    F->l_val.alloc(nl_nz + 1);
    F->l_val32.alloc(nl_nz + 1);
    F->u_col.alloc(nu_nz + 1);
    F->u_val.alloc(nu_nz + 1);
    F->u_val32.alloc(nu_nz + 1);
    F->u_map.alloc(nu_nz + 1);
    F->l_diag.alloc(n + 1);
    F->l_rdiag32.alloc(n + 1);
    DevBuf<double> adiag(n + 1), lval0(nl_nz + 1);
    DevBuf<int32_t> asymf(1);
    CN_HIP(hipMemsetAsync(asymf.get(), 0, sizeof(int32_t), s));
    if (n > 0) {
      k_tri_fill<<<grid_for(n, 256), 256, 0, s>>>(
          n, F->a_ptr.get(), F->a_col.get(), F->a_val.get(), F->l_ptr.get(),
          F->l_col.get(), lval0.get(), adiag.get(), F->u_ptr.get(), F->u_col.get());
      k_u_map<<<grid_for(n, 256), 256, 0, s>>>(n, F->u_ptr.get(), F->u_col.get(),
                                              F->l_ptr.get(), F->l_col.get(),
                                              F->u_map.get(), asymf.get());
    }
    CN_HIP(hipGetLastError());
    {
      int32_t asym = 0;
      CN_HIP(hipMemcpyAsync(&asym, asymf.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
      CN_HIP(hipStreamSynchronize(s));
      if (asym != 0) {
        delete F;
        set_error("cn_ic0_factor: matrix pattern is not symmetric");
        return CN_BAD_ARG;
      }
    }
    const double shifts[] = {0.0, 1e-4, 1e-3, 1e-2, 1e-1, 1.0};
    const int nshift = allow_shift ? 6 : 1;
    DevBuf<int32_t> fail(1);
    int32_t hf = INT_MAX;
    for (int si = 0; si < nshift; ++si) {
      CN_HIP(hipMemcpyAsync(F->l_val.get(), lval0.get(), sizeof(double) * (nl_nz + 1),
                            hipMemcpyDeviceToDevice, s));
      CN_HIP(hipMemcpyAsync(fail.get(), &big, sizeof(int32_t), hipMemcpyHostToDevice, s));
      for (int32_t c = 0; c < nl; ++c) {
        int32_t b0 = F->level_blk[c], b1 = F->level_blk[c + 1];
        if (b1 > b0)
          k_ic0_blocks<<<grid_for(b1 - b0, 64), 64, 0, s>>>(
              b0, b1, F->blk_ptr.get(), F->l_ptr.get(), F->l_col.get(), F->l_val.get(),
              adiag.get(), F->l_diag.get(), shifts[si], fail.get());
      }
      CN_HIP(hipGetLastError());
      CN_HIP(hipMemcpyAsync(&hf, fail.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
      CN_HIP(hipStreamSynchronize(s));
      if (hf == INT_MAX) { F->shift = shifts[si]; break; }
    }
    if (hf != INT_MAX) {
      delete F;
      set_error("cn_ic0_factor: non-positive pivot at permuted row %d", hf);
      return CN_NOT_SPD;
    }
    if (nu_nz > 0)
      k_gather_u<<<grid_for(nu_nz, 256), 256, 0, s>>>(nu_nz, F->u_map.get(),
                                                     F->l_val.get(), F->u_val.get());
    if (nl_nz > 0)
      k_to_f32<<<grid_for(nl_nz, 256), 256, 0, s>>>(nl_nz, F->l_val.get(), F->l_val32.get());
    if (nu_nz > 0)
      k_to_f32<<<grid_for(nu_nz, 256), 256, 0, s>>>(nu_nz, F->u_val.get(), F->u_val32.get());
    F->l_rdiag64.alloc(n + 1);
    F->l_col_nat.alloc(nl_nz + 1);
    F->u_col_nat.alloc(nu_nz + 1);
    if (n > 0) {
      k_rdiag<<<grid_for(n, 256), 256, 0, s>>>(n, F->l_diag.get(), F->l_rdiag32.get(),
                                              F->l_rdiag64.get());
    }
    if (nl_nz > 0)
      k_renumber<<<grid_for(nl_nz, 256), 256, 0, s>>>(nl_nz, F->perm.get(), F->l_col.get(),
                                                     F->l_col_nat.get());
    if (nu_nz > 0)
      k_renumber<<<grid_for(nu_nz, 256), 256, 0, s>>>(nu_nz, F->perm.get(), F->u_col.get(),
                                                     F->u_col_nat.get());
    CN_HIP(hipGetLastError());
    F->sweep_mb = 0;
    if (F->max_block > 1 && F->max_block <= 32)
      F->sweep_mb = F->max_block <= 8 ? 8 : (F->max_block <= 16 ? 16 : 32);
    if (F->sweep_mb > 0) {
      DevBuf<int32_t> pblk(n + 1), pb0(n + 1), pb1(n + 1);
      if (nblk > 0)
        k_pos_block<<<grid_for(nblk, 256), 256, 0, s>>>(nblk, F->blk_ptr.get(), pblk.get(),
                                                       pb0.get(), pb1.get());
      build_sweep(F->lsw, n, nblk, F->sweep_mb, false, F->l_ptr.get(), F->l_col.get(),
                  F->l_val.get(), F->perm.get(), pblk.get(), pb0.get(), pb1.get(),
                  F->l_rdiag64.get(), s);
      build_sweep(F->usw, n, nblk, F->sweep_mb, true, F->u_ptr.get(), F->u_col.get(),
                  F->u_val.get(), F->perm.get(), pblk.get(), pb0.get(), pb1.get(),
                  F->l_rdiag64.get(), s);
      CN_HIP(hipStreamSynchronize(s));
    }
    CN_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete F;
    throw;
  }
  *out = F;
  return CN_OK;
  CN_API_END
}

cn_status cn_ic0_get_info(cn_ic0 F, cn_ic0_info *info) {
  if (!F || !info) { set_error("null handle"); return CN_BAD_ARG; }
  info->n = F->n;
  info->nlevels = F->ncolors;
  info->nblocks = F->level_blk.empty() ? 0 : F->level_blk.back();
  info->max_block = F->max_block;
  info->ordering = F->ordering;
  info->nnz_l = F->nnz_l;
  info->shift = F->shift;
  info->d_perm = F->perm.get();
  info->d_a_ptr = F->a_ptr.get();
  info->d_a_col = F->a_col.get();
  info->d_a_val = F->a_val.get();
  info->d_l_ptr = F->l_ptr.get();
  info->d_l_col = F->l_col.get();
  info->d_l_val = F->l_val.get();
  info->d_l_diag = F->l_diag.get();
  return CN_OK;
}

cn_status cn_ic0_apply(cn_ic0 F, int32_t k, const double *d_r, double *d_z,
                       void *stream) {
  CN_API_BEGIN
  if (!F || k <= 0 || !d_r || !d_z) { set_error("cn_ic0_apply: bad arguments"); return CN_BAD_ARG; }
  hipStream_t s = as_stream(stream);
  const int32_t n = F->n;
  int64_t tot = (int64_t)n * k;
  if (tot == 0) return CN_OK;
  DevBuf<double> rp(tot), y(tot), x(tot);
  k_gather_rows<<<grid_for(tot, 256), 256, 0, s>>>(n, k, F->perm.get(), d_r, rp.get());
  for (int32_t c = 0; c < F->ncolors; ++c) {
    int32_t b0 = F->level_blk[c], b1 = F->level_blk[c + 1];
    int64_t t = (int64_t)(b1 - b0) * k;
    if (t > 0)
      k_fwd_exact<<<grid_for(t, 256), 256, 0, s>>>(b0, b1, k, F->blk_ptr.get(),
                                                  F->l_ptr.get(), F->l_col.get(),
                                                  F->l_val.get(), F->l_diag.get(),
                                                  rp.get(), y.get());
  }
  for (int32_t c = F->ncolors - 1; c >= 0; --c) {
    int32_t b0 = F->level_blk[c], b1 = F->level_blk[c + 1];
    int64_t t = (int64_t)(b1 - b0) * k;
    if (t > 0)
      k_bwd_exact<<<grid_for(t, 256), 256, 0, s>>>(b0, b1, k, F->blk_ptr.get(),
                                                  F->u_ptr.get(), F->u_col.get(),
                                                  F->u_val.get(), F->l_diag.get(),
                                                  y.get(), x.get());
  }
  k_scatter_rows<<<grid_for(tot, 256), 256, 0, s>>>(n, k, F->perm.get(), x.get(), d_z);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

cn_status cn_ic0_destroy(cn_ic0 F) {
  delete F;
  return CN_OK;
}

}  // extern "C"
