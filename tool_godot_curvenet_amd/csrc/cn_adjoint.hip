// cn_adjoint.hip -- adjoint of the harmonic solve.
//
// The forward problem is A_ff x_f = -A_fc x_c (HarmonicSolve.solveMulti,
// HarmonicSolve.lean:51-66, with A = V^T L_h V from
// CutMeshLaplacian.assembleVtLhV).  For a loss L(x_f) with gradient
// g = dL/dx_f:
//   lambda = A_ff^-1 g                      (A_ff symmetric)
//   dL/dx_c = -A_fc^T lambda
//   dL/dw_ij = -(lambda_i - lambda_j)(x_i - x_j)
// where w_ij is the weight of the edge an off-diagonal entry couples
// (A_ii += w, A_jj += w, A_ij = A_ji -= w) and lambda is 0 on constrained
// vertices.  The solve reuses the PCG path and its IC(0) factor; the
// products are FP64 in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_internal.h"

namespace cn {

__global__ void k_cf_count(int32_t n, const int32_t *__restrict__ fc_ptr,
                           const int32_t *__restrict__ fc_col, int32_t *__restrict__ cnt) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t e = fc_ptr[i]; e < fc_ptr[i + 1]; ++e) atomicAdd(&cnt[fc_col[e]], 1);
}

__global__ void k_cf_fill(int32_t n, const int32_t *__restrict__ fc_ptr,
                          const int32_t *__restrict__ fc_col,
                          const int32_t *__restrict__ cf_ptr, int32_t *__restrict__ cursor,
                          int32_t *__restrict__ cf_row, int32_t *__restrict__ cf_map) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t e = fc_ptr[i]; e < fc_ptr[i + 1]; ++e) {
    const int32_t c = fc_col[e];
    const int32_t slot = cf_ptr[c] + atomicAdd(&cursor[c], 1);
    cf_row[slot] = i;
    cf_map[slot] = e;
  }
}

// order each transposed row by free row (the fill order above is racy)
__global__ void k_cf_sort(int32_t ncon, const int32_t *__restrict__ cf_ptr,
                          int32_t *__restrict__ cf_row, int32_t *__restrict__ cf_map) {
  int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncon) return;
  const int32_t a = cf_ptr[c], b = cf_ptr[c + 1];
  for (int32_t p = a + 1; p < b; ++p) {
    const int32_t r = cf_row[p], m = cf_map[p];
    int32_t q = p - 1;
    while (q >= a && cf_row[q] > r) {
      cf_row[q + 1] = cf_row[q];
      cf_map[q + 1] = cf_map[q];
      --q;
    }
    cf_row[q + 1] = r;
    cf_map[q + 1] = m;
  }
}

// gc[c][j] = -sum_s A_fc[row_s][c] lambda[row_s][j], rows ascending
__global__ void k_adj_gc(int32_t ncon, int32_t k, const int32_t *__restrict__ cf_ptr,
                         const int32_t *__restrict__ cf_row, const int32_t *__restrict__ cf_map,
                         const double *__restrict__ fc_val, const double *__restrict__ lam,
                         double *__restrict__ gc) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)ncon * k) return;
  const int32_t c = (int32_t)(id / k), j = (int32_t)(id % k);
  double acc = 0.0;
  for (int32_t s = cf_ptr[c]; s < cf_ptr[c + 1]; ++s)
    acc += fc_val[cf_map[s]] * lam[(int64_t)cf_row[s] * k + j];
  gc[id] = -acc;
}

// gw[e] for every stored entry of A_ff (then A_fc, offset nnz_ff); one
// thread per free row, columns summed in order
__global__ void k_adj_gw(int32_t n, int32_t k, const int32_t *__restrict__ ff_ptr,
                         const int32_t *__restrict__ ff_col, const int32_t *__restrict__ fc_ptr,
                         const int32_t *__restrict__ fc_col, int64_t nnz_ff,
                         const double *__restrict__ lam, const double *__restrict__ x,
                         const double *__restrict__ xc, double *__restrict__ gw) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *li = lam + (int64_t)i * k, *xi = x + (int64_t)i * k;
  for (int32_t e = ff_ptr[i]; e < ff_ptr[i + 1]; ++e) {
    const int32_t j = ff_col[e];
    double acc = 0.0;
    if (j != i) {
      const double *lj = lam + (int64_t)j * k, *xj = x + (int64_t)j * k;
      for (int32_t c = 0; c < k; ++c) acc += (li[c] - lj[c]) * (xi[c] - xj[c]);
    }
    gw[e] = -acc;
  }
  if (!fc_ptr) return;
  for (int32_t e = fc_ptr[i]; e < fc_ptr[i + 1]; ++e) {
    const double *xcj = xc + (int64_t)fc_col[e] * k;
    double acc = 0.0;
    for (int32_t c = 0; c < k; ++c) acc += li[c] * (xi[c] - xcj[c]);
    gw[nnz_ff + e] = -acc;
  }
}

static void ensure_cf(cn_system_t *S, hipStream_t s) {
  if (S->cf_ready) return;
  const int32_t n = S->n_free, nc = S->n_con;
  const int64_t nnz = S->nnz_fc;
  S->cf_ptr.alloc((size_t)nc + 1);
  S->cf_row.alloc(nnz + 1);
  S->cf_map.alloc(nnz + 1);
  DevBuf<int32_t> cursor((size_t)nc + 1);
  CN_HIP(hipMemsetAsync(S->cf_ptr.get(), 0, sizeof(int32_t) * (nc + 1), s));
  CN_HIP(hipMemsetAsync(cursor.get(), 0, sizeof(int32_t) * (nc + 1), s));
  if (n > 0)
    k_cf_count<<<grid_for(n, 256), 256, 0, s>>>(n, S->fc_ptr.get(), S->fc_col.get(),
                                               S->cf_ptr.get());
  exclusive_scan_i32(S->cf_ptr.get(), (int64_t)nc + 1, s);
  if (n > 0)
    k_cf_fill<<<grid_for(n, 256), 256, 0, s>>>(n, S->fc_ptr.get(), S->fc_col.get(),
                                              S->cf_ptr.get(), cursor.get(), S->cf_row.get(),
                                              S->cf_map.get());
  if (nc > 0)
    k_cf_sort<<<grid_for(nc, 256), 256, 0, s>>>(nc, S->cf_ptr.get(), S->cf_row.get(),
                                               S->cf_map.get());
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  S->cf_ready = true;
}

}  // namespace cn

namespace cn {
// The adjoint after its solve: d_lambda holds A_ff^-1 g (status st).
static cn_status adjoint_gradients(cn_system_t *S, int32_t k, const double *d_x,
                                   const double *d_xc, const double *d_lambda, double *d_gc,
                                   double *d_gw, hipStream_t s) {
  if (d_gc && S->n_con > 0) {
    ensure_cf(S, s);
    const int64_t tot = (int64_t)S->n_con * k;
    k_adj_gc<<<grid_for(tot, 256), 256, 0, s>>>(S->n_con, k, S->cf_ptr.get(), S->cf_row.get(),
                                               S->cf_map.get(), S->fc_val.get(), d_lambda,
                                               d_gc);
  }
  if (d_gw && S->n_free > 0) {
    const bool fc = S->n_con > 0 && S->nnz_fc > 0;
    k_adj_gw<<<grid_for(S->n_free, 128), 128, 0, s>>>(
        S->n_free, k, S->ff_ptr.get(), S->ff_col.get(), fc ? S->fc_ptr.get() : nullptr,
        fc ? S->fc_col.get() : nullptr, S->nnz_ff, d_lambda, d_x, d_xc, d_gw);
  }
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
}

static bool adjoint_args(cn_system S, const cn_pcg_opts *o, int32_t k, const double *d_g,
                         const double *d_x, const double *d_xc, const double *d_lambda,
                         const double *d_gw, const char *who) {
  if (!S || !o || k <= 0 || !d_g || !d_lambda) {
    set_error("%s: bad arguments", who);
    return false;
  }
  if (d_gw && (!d_x || (S->n_con > 0 && !d_xc))) {
    set_error("%s: the weight gradient needs x (and x_c)", who);
    return false;
  }
  return true;
}
}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_adjoint_solve(cn_system S, cn_ic0 F, const cn_pcg_opts *o, int32_t k,
                           const double *d_g, const double *d_x, const double *d_xc,
                           double *d_lambda, double *d_gc, double *d_gw,
                           cn_solve_info *info, void *stream) {
  CN_API_BEGIN
  if (!adjoint_args(S, o, k, d_g, d_x, d_xc, d_lambda, d_gw, "cn_adjoint_solve"))
    return CN_BAD_ARG;
  cn_pcg_opts oo = *o;
  oo.warm_start = 0;
  const cn_status st = cn_pcg_solve(S, F, &oo, k, d_g, d_lambda, info, stream);
  if (st != CN_OK && st != CN_MAXIT) return st;
  adjoint_gradients(S, k, d_x, d_xc, d_lambda, d_gc, d_gw, as_stream(stream));
  return st;
  CN_API_END
}

cn_status cn_adjoint_solve_hsc(cn_system S, cn_hsc H, const cn_pcg_opts *o, int32_t k,
                               const double *d_g, const double *d_x, const double *d_xc,
                               double *d_lambda, double *d_gc, double *d_gw,
                               cn_solve_info *info, void *stream) {
  CN_API_BEGIN
  if (!adjoint_args(S, o, k, d_g, d_x, d_xc, d_lambda, d_gw, "cn_adjoint_solve_hsc"))
    return CN_BAD_ARG;
  cn_pcg_opts oo = *o;
  oo.warm_start = 0;
  const cn_status st = cn_pcg_solve_hsc(S, H, &oo, k, d_g, d_lambda, info, stream);
  if (st != CN_OK && st != CN_MAXIT) return st;
  adjoint_gradients(S, k, d_x, d_xc, d_lambda, d_gc, d_gw, as_stream(stream));
  return st;
  CN_API_END
}

}  // extern "C"
