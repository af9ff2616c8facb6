// cn_assembly.hip -- cut-mesh cotangent Laplacian VtLhV as CSR on gfx950.
//
// Compiled with -ffp-contract=off: every FP64 value here must equal the
// spec's (CutMeshLaplacian.lean:56-93, PolygonLaplacian.lean:32-93) bit for
// bit, so no multiply-add may be fused.
//
// Layout: per face 9 doubles (the triangle cot Laplacian of the face loop
// that starts at halfedge 3f, whose targets are t1, t2, t0); per vertex the
// ascending list of incoming interior halfedges; per free row the sorted
// column list.  The values are a segmented reduction: every CSR entry is
// the sum of its terms (face_m entries) in the spec's order, one lane per
// entry over a plan built once from the pattern, so each entry is summed in
// exactly the spec's order without atomics.
#include <hip/hip_runtime.h>

#include <climits>

#include "cn_internal.h"
#include "cn_kernels.h"
#include "cn_prof.h"

#pragma clang fp contract(off)

namespace cn {

__device__ __forceinline__ int32_t tgt3(const int32_t *tris, int32_t h) {
  int32_t f = h / 3, l = h - 3 * f;
  return tris[3 * f + (l == 2 ? 0 : l + 1)];
}

// PolygonLaplacian.cotAtVertex (:32-40)
__device__ __forceinline__ double cot_at(const double *ap, const double *o1,
                                         const double *o2) {
  double ax = o1[0] - ap[0], ay = o1[1] - ap[1], az = o1[2] - ap[2];
  double bx = o2[0] - ap[0], by = o2[1] - ap[1], bz = o2[2] - ap[2];
  double dot = ax * bx + ay * by + az * bz;
  double cx = ay * bz - az * by;
  double cy = az * bx - ax * bz;
  double cz = ax * by - ay * bx;
  double cl = sqrt(cx * cx + cy * cy + cz * cz);
  return dot / cl;
}

// triangleCotLaplacian (:56-71) on the face loop (t1, t2, t0)
__global__ void k_face_cot(int32_t nf, const int32_t *tris, const double *pos,
                           double *m) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double a[3], b[3], c[3];
  int32_t ia = tris[3 * f + 1], ib = tris[3 * f + 2], ic = tris[3 * f];
  for (int d = 0; d < 3; ++d) {
    a[d] = pos[3 * ia + d];
    b[d] = pos[3 * ib + d];
    c[d] = pos[3 * ic + d];
  }
  double cotA = cot_at(a, b, c), cotB = cot_at(b, a, c), cotC = cot_at(c, a, b);
  double w_bc = cotA * 0.5, w_ac = cotB * 0.5, w_ab = cotC * 0.5;
  double *o = m + 9 * (int64_t)f;
  o[0] = w_ab + w_ac; o[1] = -w_ab;       o[2] = -w_ac;
  o[3] = -w_ab;       o[4] = w_ab + w_bc; o[5] = -w_bc;
  o[6] = -w_ac;       o[7] = -w_bc;       o[8] = w_ac + w_bc;
}

// RobustLaplacian.robustTriangleCotLaplacian (RobustLaplacian.lean:136-140)
// on the same corners as k_face_cot: edge lengths (:63-70), the per-triangle
// mollification eps = max(0, max over corners of delta + opposite - the two
// adjacent) (:83-89), cot by the law of cosines over Heron's area
// (:106-109), the 3x3 block of triangleCotLaplacian (:114-131).
__device__ __forceinline__ double max_f(double a, double b) { return a <= b ? b : a; }
__device__ __forceinline__ double edge_len(const double *u, const double *v) {
  double dx = u[0] - v[0], dy = u[1] - v[1], dz = u[2] - v[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}
__device__ __forceinline__ double cot_from_lengths(double eo, double e1, double e2) {
  double s = (eo + e1 + e2) * 0.5;
  double area = sqrt(s * (s - eo) * (s - e1) * (s - e2));
  return (e1 * e1 + e2 * e2 - eo * eo) / (4.0 * area);
}

__global__ void k_face_robust(int32_t nf, const int32_t *tris, const double *pos,
                              double delta, double *m) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double a[3], b[3], c[3];
  int32_t ia = tris[3 * f + 1], ib = tris[3 * f + 2], ic = tris[3 * f];
  for (int d = 0; d < 3; ++d) {
    a[d] = pos[3 * ia + d];
    b[d] = pos[3 * ib + d];
    c[d] = pos[3 * ic + d];
  }
  double eA = edge_len(b, c), eB = edge_len(a, c), eC = edge_len(a, b);
  double ra = delta + eA - eB - eC;
  double rb = delta + eB - eA - eC;
  double rc = delta + eC - eA - eB;
  double eps = max_f(0.0, max_f(ra, max_f(rb, rc)));
  eA = eA + eps;
  eB = eB + eps;
  eC = eC + eps;
  double cotA = cot_from_lengths(eA, eB, eC), cotB = cot_from_lengths(eB, eA, eC),
         cotC = cot_from_lengths(eC, eA, eB);
  double w_bc = cotA * 0.5, w_ac = cotB * 0.5, w_ab = cotC * 0.5;
  double *o = m + 9 * (int64_t)f;
  o[0] = w_ab + w_ac; o[1] = -w_ab;       o[2] = -w_ac;
  o[3] = -w_ab;       o[4] = w_ab + w_bc; o[5] = -w_bc;
  o[6] = -w_ac;       o[7] = -w_bc;       o[8] = w_ac + w_bc;
}

__global__ void k_flags(int32_t nv, const uint8_t *constrained, int32_t *fr,
                        int32_t *co) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  fr[v] = constrained[v] ? 0 : 1;
  co[v] = constrained[v] ? 1 : 0;
}

__global__ void k_maps(int32_t nv, const uint8_t *constrained, int32_t *free_of,
                       int32_t *con_of, int32_t *free_vertex,
                       int32_t *con_vertex) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  if (constrained[v]) {
    con_vertex[con_of[v]] = v;
    free_of[v] = -1;
  } else {
    free_vertex[free_of[v]] = v;
    con_of[v] = -1;
  }
}

__global__ void k_cand_count(int32_t nv, const int32_t *iptr, int32_t *cnt) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  cnt[v] = 1 + 2 * (iptr[v + 1] - iptr[v]);
}

// Candidate columns of row v: v and the other corners of every face around
// v (the faces of its incoming halfedges); sorted and de-duplicated in
// place.  Counts go to the free row's slot.
__global__ void k_cand_fill(int32_t nv, const int32_t *tris,
                            const int32_t *iptr, const int32_t *ihe,
                            const int32_t *cptr, int32_t *cand,
                            const uint8_t *constrained, const int32_t *free_of,
                            int32_t *ucnt, int32_t *cnt_ff, int32_t *cnt_fc) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int32_t *c = cand + cptr[v];
  int32_t n = 0;
  c[n++] = v;
  for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
    int32_t p = ihe[a], f = p / 3;
    for (int l = 0; l < 3; ++l) {
      int32_t w = tris[3 * f + l];
      if (w != v) c[n++] = w;
    }
  }
  for (int32_t i = 1; i < n; ++i) {
    int32_t x = c[i], j = i - 1;
    while (j >= 0 && c[j] > x) { c[j + 1] = c[j]; --j; }
    c[j + 1] = x;
  }
  int32_t u = 0, nff = 0, nfc = 0;
  for (int32_t i = 0; i < n; ++i)
    if (i == 0 || c[i] != c[i - 1]) {
      c[u++] = c[i];
      if (constrained[c[i]]) ++nfc; else ++nff;
    }
  ucnt[v] = u;
  int32_t r = free_of[v];
  if (r >= 0) { cnt_ff[r] = nff; cnt_fc[r] = nfc; }
}

__global__ void k_pattern_fill(int32_t n_free, const int32_t *free_vertex,
                               const int32_t *cptr, const int32_t *cand,
                               const int32_t *ucnt, const uint8_t *constrained,
                               const int32_t *free_of, const int32_t *con_of,
                               const int32_t *ff_ptr, const int32_t *fc_ptr,
                               int32_t *ff_col, int32_t *fc_col) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_free) return;
  int32_t v = free_vertex[r];
  const int32_t *c = cand + cptr[v];
  int32_t a = ff_ptr[r], b = fc_ptr[r];
  for (int32_t i = 0; i < ucnt[v]; ++i) {
    int32_t w = c[i];
    if (constrained[w]) fc_col[b++] = con_of[w];
    else ff_col[a++] = free_of[w];
  }
}

__device__ __forceinline__ int32_t find_col(const int32_t *col, int32_t a,
                                            int32_t b, int32_t key) {
  while (a < b) {
    int32_t m = (a + b) >> 1;
    if (col[m] < key) a = m + 1; else b = m;
  }
  return a;
}

// Row values (the spec's order): for p in inc(v) ascending, for each
// distinct target j of p's face: t = sum over q (ascending) targeting j of
// Lh[p][q], from 0; entry(j) += t, from 0.  Each CSR entry is a segment of
// that walk: the face_m positions of its terms in that order, a term that
// starts a new t flagged in bit 31.  The plan depends only on the pattern,
// so it is built once per system (two passes of the walk, one thread per
// row: counts, then the terms); every assembly then sums each segment on
// one lane.
constexpr uint32_t kGroupStart = 0x80000000u;

template <bool FILL>
__global__ void k_asm_plan(int32_t n_free, const int32_t *__restrict__ free_vertex,
                           const int32_t *__restrict__ tris, const int32_t *__restrict__ iptr,
                           const int32_t *__restrict__ ihe,
                           const uint8_t *__restrict__ constrained,
                           const int32_t *__restrict__ free_of,
                           const int32_t *__restrict__ con_of,
                           const int32_t *__restrict__ ff_ptr,
                           const int32_t *__restrict__ ff_col,
                           const int32_t *__restrict__ fc_ptr,
                           const int32_t *__restrict__ fc_col, int64_t nnz_ff, int32_t *cnt,
                           const int32_t *__restrict__ term_ptr, uint32_t *terms,
                           int32_t *diag_pos) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_free) return;
  const int32_t v = free_vertex[r];
  const int32_t a0 = ff_ptr[r], a1 = ff_ptr[r + 1], b0 = fc_ptr[r], b1 = fc_ptr[r + 1];
  // cnt: the count pass counts every entry's terms, the fill pass uses it
  // as the entry's write cursor; only this row's thread touches its entries
  for (int32_t i = a0; i < a1; ++i) cnt[i] = FILL ? term_ptr[i] : 0;
  for (int32_t i = b0; i < b1; ++i) cnt[nnz_ff + i] = FILL ? term_ptr[nnz_ff + i] : 0;
  if (FILL) {
    const int32_t d = find_col(ff_col, a0, a1, r);
    diag_pos[r] = (d < a1 && ff_col[d] == r) ? d : -1;
  }
  for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
    const int32_t p = ihe[a], f = p / 3, lp = p - 3 * f;
    const int32_t tj[3] = {tgt3(tris, 3 * f), tgt3(tris, 3 * f + 1), tgt3(tris, 3 * f + 2)};
    for (int lq = 0; lq < 3; ++lq) {
      const int32_t j = tj[lq];
      bool seen = false;
      for (int e = 0; e < lq; ++e) seen |= (tj[e] == j);
      if (seen) continue;
      int64_t pos;
      if (constrained[j]) {
        const int32_t q = find_col(fc_col, b0, b1, con_of[j]);
        if (!(q < b1 && fc_col[q] == con_of[j])) continue;
        pos = nnz_ff + q;
      } else {
        const int32_t q = find_col(ff_col, a0, a1, free_of[j]);
        if (!(q < a1 && ff_col[q] == free_of[j])) continue;
        pos = q;
      }
      for (int e = lq; e < 3; ++e) {
        if (tj[e] != j) continue;
        if (FILL)
          terms[cnt[pos]++] = (uint32_t)(9 * f + 3 * lp + e) | (e == lq ? kGroupStart : 0u);
        else
          ++cnt[pos];
      }
    }
  }
}

// Every entry of A_ff and A_fc (A_fc's after A_ff's): the in-order sum of
// its segment, written in FP64 and FP32.  Lanes take consecutive entries,
// so the plan is read and the values written in whole cache lines; the
// face_m terms of a row's entries come from the few faces around it.
__global__ void __launch_bounds__(256)
k_asm_values(int64_t nnz_tot, int64_t nnz_ff, const int32_t *__restrict__ term_ptr,
             const uint32_t *__restrict__ terms, const double *__restrict__ m,
             double *__restrict__ ff_val, float *__restrict__ ff_val32,
             double *__restrict__ fc_val, float *__restrict__ fc_val32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz_tot) return;
  double acc = 0.0, t = 0.0;
  bool any = false;
  const int32_t q1 = term_ptr[e + 1];
  for (int32_t q = term_ptr[e]; q < q1; ++q) {
    const uint32_t u = terms[q];
    if ((u & kGroupStart) && any) {
      acc = acc + t;
      t = 0.0;
    }
    t = t + (0.0 + m[u & ~kGroupStart]);
    any = true;
  }
  if (any) acc = acc + t;
  if (e < nnz_ff) {
    ff_val[e] = acc;
    ff_val32[e] = (float)acc;
  } else {
    fc_val[e - nnz_ff] = acc;
    fc_val32[e - nnz_ff] = (float)acc;
  }
}

__global__ void k_diag_at(int32_t n, const int32_t *__restrict__ diag_pos,
                          const double *__restrict__ val, double *__restrict__ d) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t q = diag_pos[i];
  d[i] = q >= 0 ? val[q] : 0.0;
}

__global__ void k_mirror(int64_t n, const double *x, float *y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

__global__ void k_diag(int32_t n, const int32_t *ptr, const int32_t *col,
                       const double *val, double *d) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = 0.0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p)
    if (col[p] == i) { x = val[p]; break; }
  d[i] = x;
}

// RHS of HarmonicSolve.solveMulti on the free rows (spec summation order):
// for free row r of vertex v, s = sum over v's incoming interior halfedges
// p (ascending) of t_p, t_p = sum over the three halfedges q of p's face of
// (0 + Lh[p][q]) Xc[col of q] (constrained targets only); rhs = -s.
// Computed over the right-hand side table: one row's column tile per
// LPR lanes, the row's groups and entries read once for all of its columns.
// s = sum over groups of (sum over entries of (0 + m) x), from 0 in order;
// the groups and entries the table leaves out would add exact zeros.
template <int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
k_harmonic_rhs_w(int32_t n_free, int64_t k, const int32_t *__restrict__ gptr,
                 const int32_t *__restrict__ eptr, const int32_t *__restrict__ cc,
                 const int32_t *__restrict__ midx, const double *__restrict__ m,
                 const double *__restrict__ Xc, double *__restrict__ rhs) {
  using P = Pack<LPR>;
  const int64_t c0 = P::col0(VEC);
  const bool valid = c0 < k;
  const int64_t cl = valid ? c0 : k - VEC;
  const RowIter ri = row_iter(0, n_free, P::kSlots, P::slot());
  for (int32_t r0 = ri.first; r0 < ri.end; r0 += ri.stride) {
    const int32_t r = P::u(r0);
    double acc[VEC];
#pragma unroll
    for (int q = 0; q < VEC; ++q) acc[q] = 0.0;
    const int32_t g1 = P::u(gptr[r + 1]);
    for (int32_t g = P::u(gptr[r]); g < g1; ++g) {
      double t[VEC];
#pragma unroll
      for (int q = 0; q < VEC; ++q) t[q] = 0.0;
      const int32_t e1 = P::u(eptr[g + 1]);
      for (int32_t e = P::u(eptr[g]); e < e1; ++e) {
        const double w = 0.0 + m[midx[e]];
        double x[VEC];
        VecIO<double, VEC>::ld(Xc + (int64_t)cc[e] * k + cl, x);
#pragma unroll
        for (int q = 0; q < VEC; ++q) t[q] = t[q] + w * x[q];
      }
#pragma unroll
      for (int q = 0; q < VEC; ++q) acc[q] = acc[q] + t[q];
    }
    if (valid) {
      double o[VEC];
#pragma unroll
      for (int q = 0; q < VEC; ++q) o[q] = -acc[q];
      VecIO<double, VEC>::st(rhs + (int64_t)r * k + c0, o);
    }
  }
}

// the right-hand side table of a system (host build from the incidence
// lists, once per system: the pattern does not depend on positions)
static void ensure_rhs_table(cn_system_t *S, hipStream_t s) {
  if (S->rhs_ready) return;
  const int32_t n = S->n_free;
  std::vector<int32_t> fv(n), tris((size_t)3 * S->nf), iptr((size_t)S->nv + 1), con(S->nv);
  CN_HIP(hipMemcpyAsync(fv.data(), S->free_vertex.get(), sizeof(int32_t) * n,
                        hipMemcpyDeviceToHost, s));
  CN_HIP(hipMemcpyAsync(tris.data(), S->tris.get(), sizeof(int32_t) * tris.size(),
                        hipMemcpyDeviceToHost, s));
  CN_HIP(hipMemcpyAsync(iptr.data(), S->inc.ptr.get(), sizeof(int32_t) * iptr.size(),
                        hipMemcpyDeviceToHost, s));
  CN_HIP(hipMemcpyAsync(con.data(), S->con_of.get(), sizeof(int32_t) * S->nv,
                        hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> ihe(iptr[S->nv]);
  if (!ihe.empty())
    CN_HIP(hipMemcpy(ihe.data(), S->inc.he.get(), sizeof(int32_t) * ihe.size(),
                     hipMemcpyDeviceToHost));
  std::vector<int32_t> gptr(1, 0), eptr(1, 0), cc, midx;
  gptr.reserve((size_t)n + 1);
  for (int32_t r = 0; r < n; ++r) {
    const int32_t v = fv[r];
    for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
      const int32_t p = ihe[a], f = p / 3, lp = p - 3 * f;
      const size_t before = cc.size();
      for (int lq = 0; lq < 3; ++lq) {
        const int32_t h = 3 * f + lq, l = h - 3 * f;
        const int32_t c = con[tris[3 * (size_t)f + (l == 2 ? 0 : l + 1)]];
        if (c >= 0) {
          cc.push_back(c);
          midx.push_back(9 * f + 3 * lp + lq);
        }
      }
      if (cc.size() > before) eptr.push_back((int32_t)cc.size());
    }
    gptr.push_back((int32_t)eptr.size() - 1);
  }
  auto up = [&](DevBuf<int32_t> &d, const std::vector<int32_t> &h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty())
      CN_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(int32_t) * h.size(),
                            hipMemcpyHostToDevice, s));
  };
  up(S->rhs_gptr, gptr);
  up(S->rhs_eptr, eptr);
  up(S->rhs_cc, cc);
  up(S->rhs_midx, midx);
  CN_HIP(hipStreamSynchronize(s));
  S->rhs_ready = true;
}

void batch_rhs_table(cn_system_t *B, const cn_system_t *const *members, int32_t nsys,
                     hipStream_t s) {
  std::vector<int32_t> gptr(1, 0), eptr(1, 0), cc, midx;
  int64_t faces = 0;
  for (int32_t i = 0; i < nsys; ++i) faces += members[i]->nf;
  if (9 * faces > INT32_MAX) throw HipError{hipErrorInvalidValue, "batch face table > int32"};
  B->face_m.alloc((size_t)std::max<int64_t>(9 * faces, 1));
  int32_t con0 = 0;
  int64_t face0 = 0;
  for (int32_t i = 0; i < nsys; ++i) {
    cn_system_t *A = const_cast<cn_system_t *>(members[i]);
    const int32_t n = A->n_free;
    if (n > 0 && A->nv > 0) ensure_rhs_table(A, s);
    std::vector<int32_t> g((size_t)n + 1, 0);
    if (n > 0 && A->rhs_ready)
      CN_HIP(hipMemcpyAsync(g.data(), A->rhs_gptr.get(), sizeof(int32_t) * g.size(),
                            hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> e((size_t)g[n] + 1, 0);
    if (g[n] > 0)
      CN_HIP(hipMemcpyAsync(e.data(), A->rhs_eptr.get(), sizeof(int32_t) * e.size(),
                            hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> c((size_t)e[g[n]]), m((size_t)e[g[n]]);
    if (!c.empty()) {
      CN_HIP(hipMemcpyAsync(c.data(), A->rhs_cc.get(), sizeof(int32_t) * c.size(),
                            hipMemcpyDeviceToHost, s));
      CN_HIP(hipMemcpyAsync(m.data(), A->rhs_midx.get(), sizeof(int32_t) * m.size(),
                            hipMemcpyDeviceToHost, s));
    }
    if (A->nf > 0)
      CN_HIP(hipMemcpyAsync(B->face_m.get() + 9 * face0, A->face_m.get(),
                            sizeof(double) * 9 * (size_t)A->nf, hipMemcpyDeviceToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
    const int32_t gbase = (int32_t)eptr.size() - 1, ebase = (int32_t)cc.size();
    for (int32_t r = 0; r < n; ++r) gptr.push_back(gbase + g[r + 1]);
    for (int32_t q = 1; q <= g[n]; ++q) eptr.push_back(ebase + e[q]);
    for (size_t q = 0; q < c.size(); ++q) {
      cc.push_back(con0 + c[q]);
      midx.push_back((int32_t)(9 * face0) + m[q]);
    }
    con0 += A->n_con;
    face0 += A->nf;
  }
  auto up = [&](DevBuf<int32_t> &d, const std::vector<int32_t> &h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty())
      CN_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(int32_t) * h.size(),
                            hipMemcpyHostToDevice, s));
  };
  up(B->rhs_gptr, gptr);
  up(B->rhs_eptr, eptr);
  up(B->rhs_cc, cc);
  up(B->rhs_midx, midx);
  CN_HIP(hipStreamSynchronize(s));
  B->rhs_ncon = con0;
  B->rhs_ready = true;
}

static void launch_harmonic_rhs(cn_system_t *S, int32_t k, const double *Xc, double *rhs,
                                hipStream_t s) {
  const int v = (k % 2 == 0) ? 2 : 1;
  const int lpr = pack_lpr((k + v - 1) / v);
  const int slots = kWaves * (64 / lpr);
  int64_t g = ((int64_t)S->n_free + slots - 1) / slots;
  g = std::max<int64_t>(1, std::min<int64_t>(g, 2048));
  if (g >= 64) g -= g % 8;
  const dim3 grid((unsigned)g, (unsigned)((k + (int64_t)lpr * v - 1) / ((int64_t)lpr * v)), 1);
#define H_(VV, LL)                                                                       \
  k_harmonic_rhs_w<VV, LL><<<grid, kBlock, 0, s>>>(S->n_free, k, S->rhs_gptr.get(),    \
                                                    S->rhs_eptr.get(), S->rhs_cc.get(),  \
                                                    S->rhs_midx.get(), S->face_m.get(),  \
                                                    Xc, rhs)
#define H_L(VV)                                                                          \
  do {                                                                                   \
    if (lpr == 16) H_(VV, 16);                                                           \
    else if (lpr == 32) H_(VV, 32);                                                      \
    else H_(VV, 64);                                                                     \
  } while (0)
  if (v == 2) H_L(2); else H_L(1);
#undef H_L
#undef H_
}

// the assembly plan of a system (once: it depends only on the pattern)
static void build_asm_plan(cn_system_t *S, hipStream_t s) {
  const int64_t tot = S->nnz_ff + S->nnz_fc;
  if ((int64_t)9 * S->nf >= (int64_t)kGroupStart)
    throw HipError{hipErrorInvalidValue, "face table exceeds 2^31 entries"};
  S->asm_ptr.alloc((size_t)tot + 1);
  S->diag_pos.alloc((size_t)S->n_free + 1);
  // zeroed whole: the count pass leaves the last slot, which the scan reads
  CN_HIP(hipMemsetAsync(S->asm_ptr.get(), 0, sizeof(int32_t) * ((size_t)tot + 1), s));
  if (S->n_free == 0 || tot == 0) {
    S->asm_terms.alloc(1);
    return;
  }
  // at most 9 terms per face, so the int32 scan cannot overflow
  const dim3 g(grid_for(S->n_free, 128)), b(128);
  k_asm_plan<false><<<g, b, 0, s>>>(
      S->n_free, S->free_vertex.get(), S->tris.get(), S->inc.ptr.get(), S->inc.he.get(),
      S->constrained.get(), S->free_of.get(), S->con_of.get(), S->ff_ptr.get(),
      S->ff_col.get(), S->fc_ptr.get(), S->fc_col.get(), S->nnz_ff, S->asm_ptr.get(), nullptr,
      nullptr, nullptr);
  CN_HIP(hipGetLastError());
  const int64_t nterms = exclusive_scan_i32(S->asm_ptr.get(), tot + 1, s);
  S->asm_terms.alloc((size_t)nterms + 1);
  DevBuf<int32_t> cursor((size_t)tot + 1);
  k_asm_plan<true><<<g, b, 0, s>>>(
      S->n_free, S->free_vertex.get(), S->tris.get(), S->inc.ptr.get(), S->inc.he.get(),
      S->constrained.get(), S->free_of.get(), S->con_of.get(), S->ff_ptr.get(),
      S->ff_col.get(), S->fc_ptr.get(), S->fc_col.get(), S->nnz_ff, cursor.get(),
      S->asm_ptr.get(), S->asm_terms.get(), S->diag_pos.get());
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));  // cursor goes out of scope
}

static void assemble_values(cn_system_t *S, hipStream_t s) {
  if (S->nf > 0) {
    ProfScope pf("asm_face", s);
    if (S->lap_kind == CN_LAP_ROBUST)
      k_face_robust<<<grid_for(S->nf, 256), 256, 0, s>>>(S->nf, S->tris.get(), S->pos.get(),
                                                        S->lap_delta, S->face_m.get());
    else
      k_face_cot<<<grid_for(S->nf, 256), 256, 0, s>>>(S->nf, S->tris.get(), S->pos.get(),
                                                     S->face_m.get());
  }
  const int64_t tot = S->nnz_ff + S->nnz_fc;
  if (tot > 0) {
    ProfScope pf("asm_values", s);
    k_asm_values<<<grid_for(tot, 256), 256, 0, s>>>(
        tot, S->nnz_ff, S->asm_ptr.get(), S->asm_terms.get(), S->face_m.get(),
        S->ff_val.get(), S->ff_val32.get(), S->fc_val.get(), S->fc_val32.get());
  }
  if (S->n_free > 0) {
    ProfScope pf("asm_diag", s);
    k_diag_at<<<grid_for(S->n_free, 256), 256, 0, s>>>(S->n_free, S->diag_pos.get(),
                                                       S->ff_val.get(), S->diag.get());
  }
  ++S->vals_gen;  // the solver's row tables refill from the new values
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
}

}  // namespace cn

using namespace cn;

extern "C" {

static cn_status system_create(int32_t nv, int32_t nf, const int32_t *d_tris,
                               const uint8_t *d_constrained, const double *d_pos,
                               int32_t lap_kind, double lap_delta, cn_system *out,
                               void *stream) {
  CN_API_BEGIN
  if (!out || nv <= 0 || nf < 0 || !d_constrained || !d_pos || (nf > 0 && !d_tris) ||
      nf > INT_MAX / 6 || (lap_kind != CN_LAP_COT && lap_kind != CN_LAP_ROBUST) ||
      !(lap_delta >= 0.0)) {
    set_error("cn_system_create: bad arguments");
    return CN_BAD_ARG;
  }
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  cn_system_t *S = new cn_system_t();
  try {
    S->nv = nv;
    S->nf = nf;
    S->lap_kind = lap_kind;
    S->lap_delta = lap_delta;
    S->tris.alloc((size_t)3 * nf + 1);
    S->constrained.alloc((size_t)nv);
    S->pos.alloc((size_t)3 * nv);
    if (nf > 0)
      CN_HIP(hipMemcpyAsync(S->tris.get(), d_tris, sizeof(int32_t) * 3 * (size_t)nf,
                            hipMemcpyDeviceToDevice, s));
    CN_HIP(hipMemcpyAsync(S->constrained.get(), d_constrained, (size_t)nv,
                          hipMemcpyDeviceToDevice, s));
    CN_HIP(hipMemcpyAsync(S->pos.get(), d_pos, sizeof(double) * 3 * (size_t)nv,
                          hipMemcpyDeviceToDevice, s));
    {  // index range check
      DevBuf<int32_t> dummy_t((size_t)6 * nf + 1), dummy_w((size_t)6 * nf + 1),
          dummy_n((size_t)6 * nf + 1), dummy_f((size_t)6 * nf + 1);
      int32_t nh = 0;
      cn_status st = cn_halfedge_from_triangles(nv, nf, S->tris.get(), dummy_t.get(),
                                                dummy_w.get(), dummy_n.get(),
                                                dummy_f.get(), &nh, stream);
      if (st != CN_OK) { delete S; return st; }
      S->he.nv = nv;
      S->he.nf = nf;
      S->he.nh = nh;
      S->he.target = std::move(dummy_t);
      S->he.twin = std::move(dummy_w);
      S->he.next = std::move(dummy_n);
      S->he.face = std::move(dummy_f);
    }
    build_incoming(nv, nf, S->tris.get(), S->inc, s);
    // free / constrained maps
    S->free_of.alloc((size_t)nv + 1);
    S->con_of.alloc((size_t)nv + 1);
    k_flags<<<grid_for(nv, 256), 256, 0, s>>>(nv, S->constrained.get(),
                                             S->free_of.get(), S->con_of.get());
    CN_HIP(hipGetLastError());
    S->n_free = (int32_t)exclusive_scan_i32(S->free_of.get(), nv, s);
    S->n_con = (int32_t)exclusive_scan_i32(S->con_of.get(), nv, s);
    S->free_vertex.alloc((size_t)S->n_free + 1);
    S->con_vertex.alloc((size_t)S->n_con + 1);
    k_maps<<<grid_for(nv, 256), 256, 0, s>>>(nv, S->constrained.get(),
                                            S->free_of.get(), S->con_of.get(),
                                            S->free_vertex.get(),
                                            S->con_vertex.get());
    CN_HIP(hipGetLastError());
    // pattern
    DevBuf<int32_t> cptr((size_t)nv + 1), ucnt((size_t)nv + 1);
    k_cand_count<<<grid_for(nv, 256), 256, 0, s>>>(nv, S->inc.ptr.get(), cptr.get());
    int64_t ncand = exclusive_scan_i32(cptr.get(), nv, s);
    DevBuf<int32_t> cand((size_t)ncand + 1);
    S->ff_ptr.alloc((size_t)S->n_free + 1);
    S->fc_ptr.alloc((size_t)S->n_free + 1);
    CN_HIP(hipMemsetAsync(S->ff_ptr.get(), 0, sizeof(int32_t) * ((size_t)S->n_free + 1), s));
    CN_HIP(hipMemsetAsync(S->fc_ptr.get(), 0, sizeof(int32_t) * ((size_t)S->n_free + 1), s));
    k_cand_fill<<<grid_for(nv, 128), 128, 0, s>>>(
        nv, S->tris.get(), S->inc.ptr.get(), S->inc.he.get(), cptr.get(),
        cand.get(), S->constrained.get(), S->free_of.get(), ucnt.get(),
        S->ff_ptr.get(), S->fc_ptr.get());
    CN_HIP(hipGetLastError());
    S->nnz_ff = exclusive_scan_i32(S->ff_ptr.get(), (int64_t)S->n_free + 1, s);
    S->nnz_fc = exclusive_scan_i32(S->fc_ptr.get(), (int64_t)S->n_free + 1, s);
    if (S->nnz_ff > INT_MAX || S->nnz_fc > INT_MAX) {
      delete S;
      set_error("nnz exceeds int32");
      return CN_BAD_ARG;
    }
    S->ff_col.alloc((size_t)S->nnz_ff + 1);
    S->fc_col.alloc((size_t)S->nnz_fc + 1);
    S->ff_val.alloc((size_t)S->nnz_ff + 1);
    S->fc_val.alloc((size_t)S->nnz_fc + 1);
    S->ff_val32.alloc((size_t)S->nnz_ff + 1);
    S->fc_val32.alloc((size_t)S->nnz_fc + 1);
    S->diag.alloc((size_t)S->n_free + 1);
    S->face_m.alloc((size_t)9 * nf + 1);
    if (S->n_free > 0)
      k_pattern_fill<<<grid_for(S->n_free, 256), 256, 0, s>>>(
          S->n_free, S->free_vertex.get(), cptr.get(), cand.get(), ucnt.get(),
          S->constrained.get(), S->free_of.get(), S->con_of.get(),
          S->ff_ptr.get(), S->fc_ptr.get(), S->ff_col.get(), S->fc_col.get());
    CN_HIP(hipGetLastError());
    build_asm_plan(S, s);
    assemble_values(S, s);
  } catch (...) {
    delete S;
    throw;
  }
  *out = S;
  return CN_OK;
  CN_API_END
}

cn_status cn_system_create(int32_t nv, int32_t nf, const int32_t *d_tris,
                           const uint8_t *d_constrained, const double *d_pos,
                           cn_system *out, void *stream) {
  return system_create(nv, nf, d_tris, d_constrained, d_pos, CN_LAP_COT, 0.0, out, stream);
}

void cn_system_default_opts(cn_system_opts *opts) {
  if (!opts) return;
  opts->laplacian = CN_LAP_COT;
  opts->delta = 0.0;
}

cn_status cn_system_create_ex(int32_t nv, int32_t nf, const int32_t *d_tris,
                              const uint8_t *d_constrained, const double *d_pos,
                              const cn_system_opts *opts, cn_system *out, void *stream) {
  if (!opts) {
    set_error("cn_system_create_ex: no options");
    return CN_BAD_ARG;
  }
  return system_create(nv, nf, d_tris, d_constrained, d_pos, opts->laplacian, opts->delta, out,
                       stream);
}

cn_status cn_system_from_csr(int32_t n, const int32_t *d_ptr, const int32_t *d_col,
                             const double *d_val, cn_system *out, void *stream) {
  CN_API_BEGIN
  if (!out || n < 0 || (n > 0 && (!d_ptr || !d_col || !d_val))) {
    set_error("cn_system_from_csr: bad arguments");
    return CN_BAD_ARG;
  }
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  int32_t nnz = 0;
  if (n > 0) {
    CN_HIP(hipMemcpyAsync(&nnz, d_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
  }
  if (nnz < 0) { set_error("cn_system_from_csr: negative nnz"); return CN_BAD_ARG; }
  cn_system_t *S = new cn_system_t();
  try {
    S->n_free = n;
    S->nnz_ff = nnz;
    S->ff_ptr.alloc((size_t)n + 1);
    S->ff_col.alloc((size_t)nnz + 1);
    S->ff_val.alloc((size_t)nnz + 1);
    S->ff_val32.alloc((size_t)nnz + 1);
    S->fc_ptr.alloc((size_t)n + 1);
    S->fc_col.alloc(1);
    S->fc_val.alloc(1);
    S->fc_val32.alloc(1);
    S->diag.alloc((size_t)n + 1);
    S->free_vertex.alloc((size_t)n + 1);
    if (n > 0) {
      CN_HIP(hipMemcpyAsync(S->ff_ptr.get(), d_ptr, sizeof(int32_t) * ((size_t)n + 1),
                            hipMemcpyDeviceToDevice, s));
      CN_HIP(hipMemsetAsync(S->fc_ptr.get(), 0, sizeof(int32_t) * ((size_t)n + 1), s));
    }
    if (nnz > 0) {
      CN_HIP(hipMemcpyAsync(S->ff_col.get(), d_col, sizeof(int32_t) * (size_t)nnz,
                            hipMemcpyDeviceToDevice, s));
      CN_HIP(hipMemcpyAsync(S->ff_val.get(), d_val, sizeof(double) * (size_t)nnz,
                            hipMemcpyDeviceToDevice, s));
      k_mirror<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, S->ff_val.get(), S->ff_val32.get());
    }
    if (n > 0)
      k_diag<<<grid_for(n, 256), 256, 0, s>>>(n, S->ff_ptr.get(), S->ff_col.get(),
                                             S->ff_val.get(), S->diag.get());
    CN_HIP(hipGetLastError());
    CN_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete S;
    throw;
  }
  *out = S;
  return CN_OK;
  CN_API_END
}

cn_status cn_system_get_info(cn_system S, cn_system_info *info) {
  if (!S || !info) { set_error("null handle"); return CN_BAD_ARG; }
  info->nv = S->nv;
  info->nf = S->nf;
  info->nh = S->he.nh;
  info->n_free = S->n_free;
  info->n_con = S->n_con;
  info->nnz_ff = S->nnz_ff;
  info->nnz_fc = S->nnz_fc;
  info->d_ff_ptr = S->ff_ptr.get();
  info->d_ff_col = S->ff_col.get();
  info->d_ff_val = S->ff_val.get();
  info->d_ff_val32 = S->ff_val32.get();
  info->d_fc_ptr = S->fc_ptr.get();
  info->d_fc_col = S->fc_col.get();
  info->d_fc_val = S->fc_val.get();
  info->d_free_of = S->free_of.get();
  info->d_con_of = S->con_of.get();
  info->d_free_vertex = S->free_vertex.get();
  info->d_con_vertex = S->con_vertex.get();
  info->d_he_target = S->he.target.get();
  info->d_he_twin = S->he.twin.get();
  info->d_he_next = S->he.next.get();
  info->d_he_face = S->he.face.get();
  return CN_OK;
}

cn_status cn_system_set_positions(cn_system S, const double *d_pos, void *stream) {
  CN_API_BEGIN
  if (!S || !d_pos || S->nv == 0) { set_error("bad arguments (or system without a mesh)"); return CN_BAD_ARG; }
  hipStream_t s = as_stream(stream);
  CN_HIP(hipMemcpyAsync(S->pos.get(), d_pos, sizeof(double) * 3 * (size_t)S->nv,
                        hipMemcpyDeviceToDevice, s));
  prof_reset();
  assemble_values(S, s);
  prof_collect();
  return CN_OK;
  CN_API_END
}

cn_status cn_system_destroy(cn_system S) {
  delete S;
  return CN_OK;
}

cn_status cn_harmonic_rhs(cn_system S, int32_t k, const double *d_Xc,
                          double *d_rhs, void *stream) {
  CN_API_BEGIN
  // a batch (cn_system_batch) takes its members' constrained values
  // stacked, through the table built with it
  const bool batch = S && S->nseg > 1;
  const int32_t ncon = S ? (batch ? S->rhs_ncon : S->n_con) : 0;
  if (!S || k <= 0 || (ncon > 0 && !d_Xc) || !d_rhs ||
      (!batch && S->nv == 0 && S->n_free > 0) || (batch && !S->rhs_ready)) {
    set_error("cn_harmonic_rhs: bad arguments");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  int64_t tot = (int64_t)S->n_free * k;
  if (tot > 0) {
    if (!batch) ensure_rhs_table(S, s);
    launch_harmonic_rhs(S, k, d_Xc, d_rhs, s);
  }
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

}  // extern "C"
