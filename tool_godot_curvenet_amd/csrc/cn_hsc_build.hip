// cn_hsc_build.hip -- construction of the hierarchical Schur preconditioner
// (FP64, built with -ffp-contract=off so the hierarchy equals the oracle's
// restatement bit for bit).  Each level is built on host threads
// (cn_hsc_host.cpp: independent set, exact Schur complement, sparsified
// fill with diagonal compensation) and uploaded; the dense coarse matrix is
// factored and inverted on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cn_hsc_host.h"
#include "cn_internal.h"

#pragma clang fp contract(off)

namespace cn {

int dense_cholesky(int n, double *A, int64_t lda, hipStream_t s);
void dense_cholesky_solve(int n, int k, const double *L, int64_t lda, double *B, int64_t ldb,
                          hipStream_t s);

template <typename T, typename Al>
static void upload(DevBuf<T> &d, const std::vector<T, Al> &h, hipStream_t s) {
  d.alloc(h.size() + 1);
  if (!h.empty())
    CN_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

template <typename Al>
static std::vector<float> to_f32(const std::vector<double, Al> &v) {
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
      if (!hsc_build_level(A, o, hl, next)) break;
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
