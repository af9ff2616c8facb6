// cn_hsc_build.hip -- construction of the hierarchical Schur preconditioner
// (host C++, FP64, built with -ffp-contract=off so the hierarchy equals the
// oracle's restatement bit for bit).
//
// After Krishnan, Fattal, Szeliski 2013 (HSC).  The per-vertex step is the
// spec's HSC.eliminateVertex (HierarchicalSparsifyCompensate.lean:96-125):
// eliminating v adds w_vi w_vj / deg(v) to every pair of v's neighbours.
// On an SPD matrix the pivot is the diagonal a_vv (degree plus the weight to
// constrained vertices) and the update is s_ij -= a_iv a_vj / a_vv.  A level
// eliminates a whole independent set F at once, which is exact because F's
// rows do not couple: S = A_CC - A_CF D_F^-1 A_FC.  Fill is then sparsified
// (drop small new entries, compensate on the diagonal) to keep the next
// level as sparse as the mesh.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cn_internal.h"

#pragma clang fp contract(off)

namespace cn {

int dense_cholesky(int n, double *A, int64_t lda, hipStream_t s);
void dense_cholesky_solve(int n, int k, const double *L, int64_t lda, double *B, int64_t ldb,
                          hipStream_t s);

struct HostCsr {
  std::vector<int32_t> ptr, col;
  std::vector<double> val;
  int32_t n() const { return (int32_t)ptr.size() - 1; }
  double at(int32_t i, int32_t j) const {  // entry (i, j); rows sorted by column
    auto b = col.begin() + ptr[i], e = col.begin() + ptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (it != e && *it == j) ? val[it - col.begin()] : 0.0;
  }
};

struct HostLevel {
  int32_t n = 0;
  std::vector<int32_t> F, C;
  std::vector<double> dinv;
  std::vector<int32_t> r_ptr, r_col, p_ptr, p_col;  // r_col: level-l rows of F vertices
  std::vector<double> r_val, p_val;
};

// One level: returns false when the level should not be built (A is the
// coarse matrix then).  On success `next` is the sparsified Schur complement.
static bool build_level(const HostCsr &A, const cn_hsc_opts &o, HostLevel &L, HostCsr &next) {
  const int32_t n = A.n();
  std::vector<int32_t> deg(n);
  for (int32_t i = 0; i < n; ++i) {
    int32_t d = 0;
    for (int32_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e) d += A.col[e] != i;
    deg[i] = d;
  }
  std::vector<int32_t> order(n);
  for (int32_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
  });
  std::vector<char> isF(n, 0), blocked(n, 0);
  for (int32_t v : order) {
    if (blocked[v]) continue;
    isF[v] = 1;
    blocked[v] = 1;
    for (int32_t e = A.ptr[v]; e < A.ptr[v + 1]; ++e) blocked[A.col[e]] = 1;
  }
  std::vector<int32_t> F, C, cidx(n, -1), fidx(n, -1);
  for (int32_t i = 0; i < n; ++i) {
    if (isF[i]) { fidx[i] = (int32_t)F.size(); F.push_back(i); }
    else { cidx[i] = (int32_t)C.size(); C.push_back(i); }
  }
  if ((double)F.size() < o.min_frac * n || F.empty()) return false;
  const int32_t nC = (int32_t)C.size();

  // S, row by row.  Row c starts as A_CC's row; then for every F neighbour f
  // of c (ascending), for every neighbour c2 of f (ascending):
  // S[c][c2] -= (a_cf * a_fc2) / a_ff.  Each entry sees its terms in
  // ascending f, the order of the f-major loop of the restatement.
  std::vector<double> acc(nC, 0.0);
  std::vector<char> touched(nC, 0), orig(nC, 0);
  std::vector<int32_t> list;
  std::vector<std::vector<int32_t>> scol(nC), sorig(nC);
  std::vector<std::vector<double>> sval(nC);
  for (int32_t t = 0; t < nC; ++t) {
    const int32_t c = C[t];
    list.clear();
    for (int32_t e = A.ptr[c]; e < A.ptr[c + 1]; ++e) {
      const int32_t j = A.col[e];
      if (isF[j]) continue;
      const int32_t u = cidx[j];
      acc[u] = A.val[e];
      touched[u] = 1;
      orig[u] = 1;
      list.push_back(u);
    }
    for (int32_t e = A.ptr[c]; e < A.ptr[c + 1]; ++e) {
      const int32_t f = A.col[e];
      if (!isF[f]) continue;
      const double acf = A.val[e];
      const double d = A.at(f, f);
      for (int32_t q = A.ptr[f]; q < A.ptr[f + 1]; ++q) {
        const int32_t c2 = A.col[q];
        if (c2 == f) continue;
        const int32_t u = cidx[c2];
        if (!touched[u]) { touched[u] = 1; acc[u] = 0.0; list.push_back(u); }
        acc[u] = acc[u] - acf * A.val[q] / d;
      }
    }
    std::sort(list.begin(), list.end());
    scol[t] = list;
    sval[t].resize(list.size());
    sorig[t].resize(list.size());
    for (size_t m = 0; m < list.size(); ++m) {
      sval[t][m] = acc[list[m]];
      sorig[t][m] = orig[list[m]];
      touched[list[m]] = 0;
      orig[list[m]] = 0;
    }
  }
  // diagonal of S before any compensation
  std::vector<double> sdiag(nC, 0.0);
  for (int32_t t = 0; t < nC; ++t)
    for (size_t m = 0; m < scol[t].size(); ++m)
      if (scol[t][m] == t) sdiag[t] = sval[t][m];
  // drops: fill entries (i < j) small against sqrt(s_ii s_jj), (i, j) ascending
  std::vector<std::vector<char>> dropped(nC);
  for (int32_t t = 0; t < nC; ++t) dropped[t].assign(scol[t].size(), 0);
  std::vector<double> ndiag = sdiag;
  for (int32_t i = 0; i < nC; ++i) {
    for (size_t m = 0; m < scol[i].size(); ++m) {
      const int32_t j = scol[i][m];
      if (j <= i || sorig[i][m]) continue;
      const double s = sval[i][m];
      if (!(std::fabs(s) < o.theta * std::sqrt(sdiag[i] * sdiag[j]))) continue;
      const double w = o.compensate == CN_HSC_DIAG ? std::fabs(s) : s;
      ndiag[i] = ndiag[i] + w;
      ndiag[j] = ndiag[j] + w;
      dropped[i][m] = 1;
      auto it = std::lower_bound(scol[j].begin(), scol[j].end(), i);
      dropped[j][it - scol[j].begin()] = 1;
    }
  }
  next.ptr.assign(1, 0);
  next.col.clear();
  next.val.clear();
  for (int32_t t = 0; t < nC; ++t) {
    for (size_t m = 0; m < scol[t].size(); ++m) {
      if (dropped[t][m]) continue;
      next.col.push_back(scol[t][m]);
      next.val.push_back(scol[t][m] == t ? ndiag[t] : sval[t][m]);
    }
    next.ptr.push_back((int32_t)next.col.size());
  }

  L.n = n;
  L.F = F;
  L.C = C;
  L.dinv.resize(F.size());
  for (size_t f = 0; f < F.size(); ++f) L.dinv[f] = 1.0 / A.at(F[f], F[f]);
  L.r_ptr.assign(1, 0);
  for (int32_t c : C) {
    for (int32_t e = A.ptr[c]; e < A.ptr[c + 1]; ++e) {
      const int32_t j = A.col[e];
      if (!isF[j]) continue;
      L.r_col.push_back(j);
      L.r_val.push_back(-A.val[e] / A.at(j, j));
    }
    L.r_ptr.push_back((int32_t)L.r_col.size());
  }
  L.p_ptr.assign(1, 0);
  for (int32_t f : F) {
    const double d = A.at(f, f);
    for (int32_t e = A.ptr[f]; e < A.ptr[f + 1]; ++e) {
      const int32_t j = A.col[e];
      if (j == f) continue;
      L.p_col.push_back(cidx[j]);
      L.p_val.push_back(-A.val[e] / d);
    }
    L.p_ptr.push_back((int32_t)L.p_col.size());
  }
  return true;
}

template <typename T>
static void upload(DevBuf<T> &d, const std::vector<T> &h, hipStream_t s) {
  d.alloc(h.size() + 1);
  if (!h.empty())
    CN_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

static std::vector<float> to_f32(const std::vector<double> &v) {
  std::vector<float> r(v.size());
  for (size_t i = 0; i < v.size(); ++i) r[i] = (float)v[i];
  return r;
}

__global__ void k_symmetrize(int n, double *A) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * n) return;
  const int i = (int)(id / n), j = (int)(id % n);
  if (j >= i) return;
  const double a = A[(int64_t)i * n + j], b = A[(int64_t)j * n + i];
  const double m = 0.5 * (a + b);
  A[(int64_t)i * n + j] = m;
  A[(int64_t)j * n + i] = m;
}

__global__ void k_identity(int n, double *A) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * n) return;
  A[id] = (id / n == id % n) ? 1.0 : 0.0;
}

}  // namespace cn

using namespace cn;

extern "C" {

void cn_hsc_default_opts(cn_hsc_opts *o) {
  if (!o) return;
  o->theta = 0.015;
  o->coarse_max = 512;
  o->max_levels = 24;
  o->min_frac = 0.05;
  o->compensate = CN_HSC_LUMP;
}

cn_status cn_hsc_build(cn_system S, const cn_hsc_opts *opts, cn_hsc *out, void *stream) {
  CN_API_BEGIN
  if (!S || !out) {
    set_error("cn_hsc_build: bad arguments");
    return CN_BAD_ARG;
  }
  cn_hsc_opts o;
  cn_hsc_default_opts(&o);
  if (opts) o = *opts;
  if (o.coarse_max < 1 || o.max_levels < 0 || !(o.theta >= 0.0) ||
      (o.compensate != CN_HSC_LUMP && o.compensate != CN_HSC_DIAG)) {
    set_error("cn_hsc_build: bad options");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  const int32_t n = S->n_free;
  HostCsr A;
  A.ptr.resize((size_t)n + 1);
  A.col.resize(S->nnz_ff);
  A.val.resize(S->nnz_ff);
  CN_HIP(hipMemcpyAsync(A.ptr.data(), S->ff_ptr.get(), sizeof(int32_t) * (n + 1),
                        hipMemcpyDeviceToHost, s));
  if (S->nnz_ff > 0) {
    CN_HIP(hipMemcpyAsync(A.col.data(), S->ff_col.get(), sizeof(int32_t) * S->nnz_ff,
                          hipMemcpyDeviceToHost, s));
    CN_HIP(hipMemcpyAsync(A.val.data(), S->ff_val.get(), sizeof(double) * S->nnz_ff,
                          hipMemcpyDeviceToHost, s));
  }
  CN_HIP(hipStreamSynchronize(s));
  cn_hsc_t *H = new cn_hsc_t();
  try {
    H->n0 = n;
    H->opts = o;
    std::vector<int32_t> g(n);  // level-0 row of every row of the current level
    for (int32_t i = 0; i < n; ++i) g[i] = i;
    while (A.n() > o.coarse_max && (int32_t)H->lv.size() < o.max_levels) {
      HostLevel hl;
      HostCsr next;
      if (!build_level(A, o, hl, next)) break;
      H->lv.emplace_back();
      cn_hsc_t::Level &L = H->lv.back();
      L.n = hl.n;
      L.nF = (int32_t)hl.F.size();
      L.nC = (int32_t)hl.C.size();
      std::vector<int32_t> Fg(L.nF), Cg(L.nC), rg(hl.r_col.size()), pg(hl.p_col.size());
      for (int32_t f = 0; f < L.nF; ++f) Fg[f] = g[hl.F[f]];
      for (int32_t c = 0; c < L.nC; ++c) Cg[c] = g[hl.C[c]];
      for (size_t e = 0; e < rg.size(); ++e) rg[e] = g[hl.r_col[e]];
      for (size_t e = 0; e < pg.size(); ++e) pg[e] = Cg[hl.p_col[e]];
      upload(L.F, hl.F, s);
      upload(L.C, hl.C, s);
      upload(L.Fg, Fg, s);
      upload(L.Cg, Cg, s);
      upload(L.r_colg, rg, s);
      upload(L.p_colg, pg, s);
      upload(L.dinv, hl.dinv, s);
      upload(L.dinv32, to_f32(hl.dinv), s);
      upload(L.r_ptr, hl.r_ptr, s);
      upload(L.r_col, hl.r_col, s);
      upload(L.r_val, hl.r_val, s);
      upload(L.r_val32, to_f32(hl.r_val), s);
      upload(L.p_ptr, hl.p_ptr, s);
      upload(L.p_col, hl.p_col, s);
      upload(L.p_val, hl.p_val, s);
      upload(L.p_val32, to_f32(hl.p_val), s);
      H->nnz += (int64_t)hl.r_col.size() + (int64_t)hl.p_col.size();
      CN_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
      g = std::move(Cg);
      A = std::move(next);
    }
    // coarse: dense, factored and inverted on the matrix cores
    const int32_t nc = A.n();
    H->nc = nc;
    upload(H->coarse_rows, g, s);
    if (nc > 0) {
      std::vector<double> dense((size_t)nc * nc, 0.0);
      for (int32_t i = 0; i < nc; ++i)
        for (int32_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e)
          dense[(size_t)i * nc + A.col[e]] = A.val[e];
      upload(H->coarse_L, dense, s);
      const int f = dense_cholesky(nc, H->coarse_L.get(), nc, s);
      if (f >= 0) {
        delete H;
        set_error("cn_hsc_build: coarse matrix (%d rows) not positive definite at column %d; "
                  "use a smaller theta or CN_HSC_DIAG", nc, f);
        return CN_NOT_SPD;
      }
      H->coarse_inv.alloc((size_t)nc * nc);
      const int64_t nn = (int64_t)nc * nc;
      k_identity<<<grid_for(nn, 256), 256, 0, s>>>(nc, H->coarse_inv.get());
      dense_cholesky_solve(nc, nc, H->coarse_L.get(), nc, H->coarse_inv.get(), nc, s);
      k_symmetrize<<<grid_for(nn, 256), 256, 0, s>>>(nc, H->coarse_inv.get());
      CN_HIP(hipGetLastError());
      CN_HIP(hipStreamSynchronize(s));
    }
    hsc_prepare(H, s);
  } catch (...) {
    delete H;
    throw;
  }
  *out = H;
  return CN_OK;
  CN_API_END
}

// A hierarchy from its persisted arrays (cn_hsc_import): the same device
// layout as cn_hsc_build gives, the coarse factor and inverse taken as
// stored, every index checked on the host first.
cn_status cn_hsc_import(cn_system S, const cn_hsc_opts *opts, int32_t nlevels,
                        const cn_hsc_level_in *levels, int32_t n_coarse,
                        const double *coarse_L, const double *coarse_inv, cn_hsc *out,
                        void *stream) {
  CN_API_BEGIN
  if (!S || !out || nlevels < 0 || (nlevels > 0 && !levels) || n_coarse < 0 ||
      (n_coarse > 0 && (!coarse_L || !coarse_inv))) {
    set_error("cn_hsc_import: bad arguments");
    return CN_BAD_ARG;
  }
  // structure: level 0 has n_free rows, level l + 1 the C rows of level l;
  // F and C split [0, n) in ascending order; R columns are F rows, P
  // columns level-(l + 1) rows; CSR pointers start at 0 and never decrease
  int32_t n = S->n_free;
  for (int32_t l = 0; l < nlevels; ++l) {
    const cn_hsc_level_in &L = levels[l];
    if (L.n != n || L.nF < 0 || L.nC < 0 || (int64_t)L.nF + L.nC != n ||
        (L.nF > 0 && (!L.F || !L.dinv || !L.p_ptr)) || (L.nC > 0 && !L.C) || !L.r_ptr ||
        !L.p_ptr) {
      set_error("cn_hsc_import: level %d has inconsistent sizes", l);
      return CN_BAD_ARG;
    }
    std::vector<int8_t> kind((size_t)n, 0);  // 1 = F, 2 = C
    for (int32_t f = 0; f < L.nF; ++f) {
      const int32_t v = L.F[f];
      if (v < 0 || v >= n || kind[v] || (f > 0 && v <= L.F[f - 1])) {
        set_error("cn_hsc_import: level %d: F is not an ascending set of rows", l);
        return CN_BAD_ARG;
      }
      kind[v] = 1;
    }
    for (int32_t c = 0; c < L.nC; ++c) {
      const int32_t v = L.C[c];
      if (v < 0 || v >= n || kind[v] || (c > 0 && v <= L.C[c - 1])) {
        set_error("cn_hsc_import: level %d: C is not an ascending set of rows outside F", l);
        return CN_BAD_ARG;
      }
      kind[v] = 2;
    }
    auto check_csr = [&](const int32_t *ptr, int32_t rows, const int32_t *col, int32_t lim,
                         bool f_cols) {
      if (ptr[0] != 0) return false;
      for (int32_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r]) return false;
      if (ptr[rows] > 0 && !col) return false;
      for (int32_t e = 0; e < ptr[rows]; ++e) {
        if (col[e] < 0 || col[e] >= lim) return false;
        if (f_cols && kind[col[e]] != 1) return false;
      }
      return true;
    };
    if (!check_csr(L.r_ptr, L.nC, L.r_col, n, true) ||
        !check_csr(L.p_ptr, L.nF, L.p_col, L.nC, false) ||
        (L.r_ptr[L.nC] > 0 && !L.r_val) || (L.p_ptr[L.nF] > 0 && !L.p_val)) {
      set_error("cn_hsc_import: level %d: R or P is not a valid CSR of this level", l);
      return CN_BAD_ARG;
    }
    n = L.nC;
  }
  if (n != n_coarse) {
    set_error("cn_hsc_import: %d coarse rows given, the levels leave %d", n_coarse, n);
    return CN_BAD_ARG;
  }
  // the apply reads the coarse inverse by columns: it must be symmetric, as
  // cn_hsc_build leaves it
  for (int32_t i = 0; i < n_coarse; ++i)
    for (int32_t j = 0; j < i; ++j)
      if (!(coarse_inv[(size_t)i * n_coarse + j] == coarse_inv[(size_t)j * n_coarse + i])) {
        set_error("cn_hsc_import: the coarse inverse is not symmetric");
        return CN_BAD_ARG;
      }
  hipStream_t s = as_stream(stream);
  cn_hsc_t *H = new cn_hsc_t();
  try {
    H->n0 = S->n_free;
    cn_hsc_default_opts(&H->opts);
    if (opts) H->opts = *opts;
    std::vector<int32_t> g((size_t)S->n_free);
    for (int32_t i = 0; i < S->n_free; ++i) g[i] = i;
    for (int32_t l = 0; l < nlevels; ++l) {
      const cn_hsc_level_in &in = levels[l];
      H->lv.emplace_back();
      cn_hsc_t::Level &L = H->lv.back();
      L.n = in.n;
      L.nF = in.nF;
      L.nC = in.nC;
      const int32_t nr = in.r_ptr[in.nC], np = in.p_ptr[in.nF];
      std::vector<int32_t> F(in.F, in.F + in.nF), Cv(in.C, in.C + in.nC);
      std::vector<int32_t> rp(in.r_ptr, in.r_ptr + in.nC + 1), rc(in.r_col, in.r_col + nr);
      std::vector<int32_t> pp(in.p_ptr, in.p_ptr + in.nF + 1), pc(in.p_col, in.p_col + np);
      std::vector<double> dv(in.dinv, in.dinv + in.nF), rv(in.r_val, in.r_val + nr),
          pv(in.p_val, in.p_val + np);
      std::vector<int32_t> Fg(in.nF), Cg(in.nC), rg(nr), pg(np);
      for (int32_t f = 0; f < in.nF; ++f) Fg[f] = g[F[f]];
      for (int32_t c = 0; c < in.nC; ++c) Cg[c] = g[Cv[c]];
      for (int32_t e = 0; e < nr; ++e) rg[e] = g[rc[e]];
      for (int32_t e = 0; e < np; ++e) pg[e] = Cg[pc[e]];
      upload(L.F, F, s);
      upload(L.C, Cv, s);
      upload(L.Fg, Fg, s);
      upload(L.Cg, Cg, s);
      upload(L.r_colg, rg, s);
      upload(L.p_colg, pg, s);
      upload(L.dinv, dv, s);
      upload(L.dinv32, to_f32(dv), s);
      upload(L.r_ptr, rp, s);
      upload(L.r_col, rc, s);
      upload(L.r_val, rv, s);
      upload(L.r_val32, to_f32(rv), s);
      upload(L.p_ptr, pp, s);
      upload(L.p_col, pc, s);
      upload(L.p_val, pv, s);
      upload(L.p_val32, to_f32(pv), s);
      H->nnz += (int64_t)nr + np;
      CN_HIP(hipStreamSynchronize(s));
      g = std::move(Cg);
    }
    H->nc = n_coarse;
    upload(H->coarse_rows, g, s);
    if (n_coarse > 0) {
      const size_t nn = (size_t)n_coarse * n_coarse;
      upload(H->coarse_L, std::vector<double>(coarse_L, coarse_L + nn), s);
      upload(H->coarse_inv, std::vector<double>(coarse_inv, coarse_inv + nn), s);
    }
    CN_HIP(hipStreamSynchronize(s));
    hsc_prepare(H, s);
  } catch (...) {
    delete H;
    throw;
  }
  *out = H;
  return CN_OK;
  CN_API_END
}

cn_status cn_hsc_get_info(cn_hsc H, cn_hsc_info *info) {
  if (!H || !info) return CN_BAD_ARG;
  info->nlevels = (int32_t)H->lv.size();
  info->n_coarse = H->nc;
  info->nnz = H->nnz;
  info->d_coarse_L = H->coarse_L.get();
  info->d_coarse_inv = H->coarse_inv.get();
  return CN_OK;
}

cn_status cn_hsc_get_level(cn_hsc H, int32_t l, cn_hsc_level *out) {
  if (!H || !out || l < 0 || l >= (int32_t)H->lv.size()) return CN_BAD_ARG;
  const cn_hsc_t::Level &L = H->lv[l];
  out->n = L.n;
  out->nF = L.nF;
  out->nC = L.nC;
  out->d_F = L.F.get();
  out->d_C = L.C.get();
  out->d_dinv = L.dinv.get();
  out->d_r_ptr = L.r_ptr.get();
  out->d_r_col = L.r_col.get();
  out->d_r_val = L.r_val.get();
  out->d_p_ptr = L.p_ptr.get();
  out->d_p_col = L.p_col.get();
  out->d_p_val = L.p_val.get();
  return CN_OK;
}

cn_status cn_hsc_destroy(cn_hsc H) {
  delete H;
  return CN_OK;
}

}  // extern "C"
