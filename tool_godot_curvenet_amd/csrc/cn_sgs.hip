// cn_sgs.hip -- colour-parallel symmetric Gauss-Seidel and the coarse-vertex
// restriction on gfx950.
//
// SgsColor (SlangCodegen/SgsColor.lean:87-130): the reference runs one
// thread per row of the active colour, one dispatch per colour, colours
// 0..C-1 and back.  Here a thread owns one (row, column) of an n x k
// row-major block, so the threads of a row's k columns read each neighbour
// row as one coalesced segment.  Rows of one colour never neighbour each
// other, so the colour's rows update in parallel exactly as the spec's
// sequential loop would.  The file is compiled without floating-point
// contraction: every + - * / is one correctly rounded operation, as in the
// spec's loop (and in the oracle's restatement of it).
//
// HierarchicalSparsify restriction (SlangCodegen/HierarchicalSparsify.lean:
// 64-93): the spec scans all fine vertices per coarse vertex; the
// CSR-by-coarse form it names as the production form visits only the
// children, in the same ascending order, so the sums are the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "cn_internal.h"
#include "cn_kernels.h"
#include "cn_prof.h"

namespace cn {

// x_i = (b_i - (sum_e val_e x_{col_e} - d x_i)) / d for the rows of one
// colour; d = A_ii (rows with d == 0 are left alone, SgsColor.lean:107-108)
template <typename T>
__global__ void __launch_bounds__(256)
k_sgs_rows(int32_t nrows, const int32_t *__restrict__ rows, const int32_t *__restrict__ ptr,
           const int32_t *__restrict__ col, const T *__restrict__ val,
           const double *__restrict__ diag, int64_t k, const T *__restrict__ b, T *x,
           const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)nrows * k) return;
  const int32_t t = (int32_t)(id / k);
  const int64_t c = id - (int64_t)t * k;
  const int32_t i = rows[t];
  const T d = (T)diag[i];
  if (d == T(0)) return;
  T s = T(0);
  for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) s = s + val[q] * x[(int64_t)col[q] * k + c];
  const int64_t o = (int64_t)i * k + c;
  x[o] = (b[o] - (s - d * x[o])) / d;
}

template <typename T>
__global__ void k_sgs_zero(int64_t total, T *z, const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = T(0);
}

// coarse[c] = sum over the children of c, ascending, from 0
template <typename T>
__global__ void k_restrict(int32_t nc, const int32_t *__restrict__ cptr,
                           const int32_t *__restrict__ child, int64_t k,
                           const T *__restrict__ fine, T *__restrict__ coarse) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)nc * k) return;
  const int32_t c = (int32_t)(id / k);
  const int64_t j = id - (int64_t)c * k;
  T acc = T(0);
  for (int32_t t = cptr[c]; t < cptr[c + 1]; ++t) acc = acc + fine[(int64_t)child[t] * k + j];
  coarse[id] = acc;
}

template <typename T> static const T *sys_vals(const cn_system_t *S);
template <> const float *sys_vals<float>(const cn_system_t *S) { return S->ff_val32.get(); }
template <> const double *sys_vals<double>(const cn_system_t *S) { return S->ff_val.get(); }

template <typename T>
static void sgs_sweeps(const cn_sgs_t *G, int64_t k, const T *b, T *x, int32_t sweeps,
                       hipStream_t s, const int32_t *ctl) {
  const cn_system_t *S = G->sys;
  const int32_t C = G->ncolors;
  for (int32_t sw = 0; sw < sweeps; ++sw)
    for (int32_t pass = 0; pass < 2 * C; ++pass) {
      const int32_t c = pass < C ? pass : 2 * C - 1 - pass;
      const int32_t r0 = G->color_ptr[c], nr = G->color_ptr[c + 1] - r0;
      if (nr == 0) continue;
      k_sgs_rows<T><<<grid_for((int64_t)nr * k, 256), 256, 0, s>>>(
          nr, G->color_rows.get() + r0, S->ff_ptr.get(), S->ff_col.get(), sys_vals<T>(S),
          S->diag.get(), k, b, x, ctl);
    }
  CN_HIP(hipGetLastError());
}

// z = M^-1 r: one symmetric sweep from z = 0 (the solver's preconditioner)
template <typename T>
void sgs_apply(const cn_sgs_t *G, int64_t k, const T *r, T *z, hipStream_t s,
               const int32_t *ctl) {
  ProfScope ps("sgs", s);
  const int64_t total = (int64_t)G->n * k;
  k_sgs_zero<T><<<(unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192)),
                  256, 0, s>>>(total, z, ctl);
  sgs_sweeps<T>(G, k, r, z, 1, s, ctl);
}
template void sgs_apply<float>(const cn_sgs_t *, int64_t, const float *, float *, hipStream_t,
                               const int32_t *);
template void sgs_apply<double>(const cn_sgs_t *, int64_t, const double *, double *,
                                hipStream_t, const int32_t *);

template <typename T>
static cn_status sgs_sweep_api(cn_sgs G, int32_t k, const T *d_b, T *d_x, int32_t sweeps,
                               void *stream, const char *who) {
  CN_API_BEGIN
  if (!G || k <= 0 || sweeps < 0 || (G->n > 0 && (!d_b || !d_x))) {
    set_error("%s: bad arguments", who);
    return CN_BAD_ARG;
  }
  if (G->sys->n_free != G->n) {
    set_error("%s: the system no longer matches the colouring", who);
    return CN_BAD_ARG;
  }
  if (G->n == 0 || sweeps == 0) return CN_OK;
  hipStream_t s = as_stream(stream);
  sgs_sweeps<T>(G, k, d_b, d_x, sweeps, s, nullptr);
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

template <typename T>
static cn_status restrict_api(cn_restrict R, int32_t k, const T *d_fine, T *d_coarse,
                              void *stream, const char *who) {
  CN_API_BEGIN
  if (!R || k <= 0 || (R->n_fine > 0 && !d_fine) || (R->n_coarse > 0 && !d_coarse)) {
    set_error("%s: bad arguments", who);
    return CN_BAD_ARG;
  }
  if (R->n_coarse == 0) return CN_OK;
  hipStream_t s = as_stream(stream);
  const int64_t total = (int64_t)R->n_coarse * k;
  k_restrict<T><<<grid_for(total, 256), 256, 0, s>>>(R->n_coarse, R->cptr.get(),
                                                     R->child.get(), k, d_fine, d_coarse);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_sgs_build(cn_system S, cn_sgs *out, void *stream) {
  CN_API_BEGIN
  if (!S || !out) {
    set_error("cn_sgs_build: bad arguments");
    return CN_BAD_ARG;
  }
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  const int32_t n = S->n_free;
  std::vector<int32_t> ptr((size_t)n + 1, 0), col, color;
  if (n > 0) {
    CN_HIP(hipMemcpyAsync(ptr.data(), S->ff_ptr.get(), sizeof(int32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
  }
  col.resize((size_t)std::max(ptr[n], 1));
  if (ptr[n] > 0) {
    CN_HIP(hipMemcpyAsync(col.data(), S->ff_col.get(), sizeof(int32_t) * (size_t)ptr[n],
                          hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
  }
  cn_sgs_t *G = new cn_sgs_t();
  try {
    G->sys = S;
    G->n = n;
    G->ncolors = n > 0 ? greedy_color(n, ptr, col, color) : 0;
    // rows of each colour, ascending (stable counting sort)
    G->color_ptr.assign((size_t)G->ncolors + 1, 0);
    for (int32_t i = 0; i < n; ++i) G->color_ptr[color[i] + 1]++;
    for (int32_t c = 0; c < G->ncolors; ++c) G->color_ptr[c + 1] += G->color_ptr[c];
    std::vector<int32_t> rows((size_t)std::max(n, 1)), fill(G->color_ptr.begin(), G->color_ptr.end());
    for (int32_t i = 0; i < n; ++i) rows[fill[color[i]]++] = i;
    G->color_rows.alloc((size_t)std::max(n, 1));
    G->color.alloc((size_t)std::max(n, 1));
    if (n > 0) {
      CN_HIP(hipMemcpyAsync(G->color_rows.get(), rows.data(), sizeof(int32_t) * n,
                            hipMemcpyHostToDevice, s));
      CN_HIP(hipMemcpyAsync(G->color.get(), color.data(), sizeof(int32_t) * n,
                            hipMemcpyHostToDevice, s));
      CN_HIP(hipStreamSynchronize(s));
    }
  } catch (...) {
    delete G;
    throw;
  }
  *out = G;
  return CN_OK;
  CN_API_END
}

cn_status cn_sgs_get_info(cn_sgs G, cn_sgs_info *info) {
  if (!G || !info) {
    set_error("cn_sgs_get_info: bad arguments");
    return CN_BAD_ARG;
  }
  info->n = G->n;
  info->ncolors = G->ncolors;
  info->h_color_ptr = G->color_ptr.data();
  info->d_color_rows = G->color_rows.get();
  info->d_color = G->color.get();
  return CN_OK;
}

cn_status cn_sgs_sweep_f64(cn_sgs G, int32_t k, const double *d_b, double *d_x,
                           int32_t sweeps, void *stream) {
  return sgs_sweep_api<double>(G, k, d_b, d_x, sweeps, stream, "cn_sgs_sweep_f64");
}

cn_status cn_sgs_sweep_f32(cn_sgs G, int32_t k, const float *d_b, float *d_x,
                           int32_t sweeps, void *stream) {
  return sgs_sweep_api<float>(G, k, d_b, d_x, sweeps, stream, "cn_sgs_sweep_f32");
}

cn_status cn_sgs_apply(cn_sgs G, int32_t k, const double *d_r, double *d_z, void *stream) {
  CN_API_BEGIN
  if (!G || k <= 0 || (G->n > 0 && (!d_r || !d_z))) {
    set_error("cn_sgs_apply: bad arguments");
    return CN_BAD_ARG;
  }
  if (G->n == 0) return CN_OK;
  hipStream_t s = as_stream(stream);
  sgs_apply<double>(G, k, d_r, d_z, s, nullptr);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

cn_status cn_sgs_destroy(cn_sgs G) {
  delete G;
  return CN_OK;
}

cn_status cn_restrict_build(int32_t n_fine, const int32_t *d_parent, int32_t n_coarse,
                            cn_restrict *out, void *stream) {
  CN_API_BEGIN
  if (!out || n_fine < 0 || n_coarse < 0 || (n_fine > 0 && !d_parent)) {
    set_error("cn_restrict_build: bad arguments");
    return CN_BAD_ARG;
  }
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  std::vector<int32_t> parent((size_t)std::max(n_fine, 1));
  if (n_fine > 0) {
    CN_HIP(hipMemcpyAsync(parent.data(), d_parent, sizeof(int32_t) * n_fine,
                          hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
  }
  std::vector<int32_t> cptr((size_t)n_coarse + 1, 0);
  for (int32_t i = 0; i < n_fine; ++i)
    if (parent[i] >= 0 && parent[i] < n_coarse) cptr[parent[i] + 1]++;
  for (int32_t c = 0; c < n_coarse; ++c) cptr[c + 1] += cptr[c];
  std::vector<int32_t> child((size_t)std::max(cptr[n_coarse], 1)), fill(cptr.begin(), cptr.end());
  for (int32_t i = 0; i < n_fine; ++i)
    if (parent[i] >= 0 && parent[i] < n_coarse) child[fill[parent[i]]++] = i;
  cn_restrict_t *R = new cn_restrict_t();
  try {
    R->n_fine = n_fine;
    R->n_coarse = n_coarse;
    R->cptr.alloc((size_t)n_coarse + 1);
    R->child.alloc(child.size());
    CN_HIP(hipMemcpyAsync(R->cptr.get(), cptr.data(), sizeof(int32_t) * cptr.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipMemcpyAsync(R->child.get(), child.data(), sizeof(int32_t) * child.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete R;
    throw;
  }
  *out = R;
  return CN_OK;
  CN_API_END
}

cn_status cn_restrict_sum_f64(cn_restrict R, int32_t k, const double *d_fine,
                              double *d_coarse, void *stream) {
  return restrict_api<double>(R, k, d_fine, d_coarse, stream, "cn_restrict_sum_f64");
}

cn_status cn_restrict_sum_f32(cn_restrict R, int32_t k, const float *d_fine, float *d_coarse,
                              void *stream) {
  return restrict_api<float>(R, k, d_fine, d_coarse, stream, "cn_restrict_sum_f32");
}

cn_status cn_restrict_destroy(cn_restrict R) {
  delete R;
  return CN_OK;
}

}  // extern "C"
