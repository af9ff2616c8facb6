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
// its CSR-by-coarse form, with the Schur weights instead of 1.  Levels >= 1
// keep r and z in one buffer: the up pass overwrites r_l with z_l (each row
// reads only its own r).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_internal.h"
#include "cn_kernels.h"

namespace cn {

template <typename TB, typename TC>
void dense_gemm_nn(int m, int k, const double *M, const TB *X, int64_t ldx, TC *Z, int64_t ldz,
                   hipStream_t s);

// ---- wide form: one wavefront per (row, 64 VEC column tile) ------------
template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
kw_hsc_down(int32_t nC, const int32_t *__restrict__ C, const int32_t *__restrict__ rptr,
            const int32_t *__restrict__ rcol, const T *__restrict__ rval, int64_t k,
            const T *in, T *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  const RowIter ri = row_iter(0, nC, kWaves, threadIdx.x >> 6);
  for (int32_t row = ri.first; row < ri.end; row += ri.stride) {
    const int32_t r = uniform(row);
    const int32_t src = uniform(C[r]);
    T acc[VEC] = {};
    if (valid) VecIO<T, VEC>::ld(in + (int64_t)src * k + c0, acc);
    gather_wide<T, VEC>(uniform(rptr[r]), uniform(rptr[r + 1]), rcol, rval, in, k, c0, valid, acc);
    if (valid) VecIO<T, VEC>::st(out + (int64_t)r * k + c0, acc);
  }
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
kw_hsc_up(int32_t n, const int32_t *__restrict__ role, const T *__restrict__ dinv,
          const int32_t *__restrict__ pptr, const int32_t *__restrict__ pcol,
          const T *__restrict__ pval, int64_t k, const T *rin, const T *zin, T *out) {
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  const RowIter ri = row_iter(0, n, kWaves, threadIdx.x >> 6);
  for (int32_t row = ri.first; row < ri.end; row += ri.stride) {
    const int32_t i = uniform(row);
    const int32_t ro = uniform(role[i]);
    T acc[VEC] = {};
    if (ro < 0) {
      if (valid) VecIO<T, VEC>::ld(zin + (int64_t)(-1 - ro) * k + c0, acc);
    } else {
      const T d = dinv[ro];
      if (valid) {
        VecIO<T, VEC>::ld(rin + (int64_t)i * k + c0, acc);
#pragma unroll
        for (int m = 0; m < VEC; ++m) acc[m] = d * acc[m];
      }
      gather_wide<T, VEC>(uniform(pptr[ro]), uniform(pptr[ro + 1]), pcol, pval, zin, k, c0,
                          valid, acc);
    }
    if (valid) VecIO<T, VEC>::st(out + (int64_t)i * k + c0, acc);
  }
}

// ---- scalar form: one thread per (row, column) --------------------------
template <typename T>
__global__ void k_hsc_down_s(int32_t nC, const int32_t *__restrict__ C,
                             const int32_t *__restrict__ rptr, const int32_t *__restrict__ rcol,
                             const T *__restrict__ rval, int64_t k, const T *in,
                             T *__restrict__ out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)nC * k) return;
  const int32_t r = (int32_t)(id / k);
  const int64_t c = id - (int64_t)r * k;
  T acc = in[(int64_t)C[r] * k + c];
  for (int32_t e = rptr[r]; e < rptr[r + 1]; ++e) acc = fma(rval[e], in[(int64_t)rcol[e] * k + c], acc);
  out[id] = acc;
}

template <typename T>
__global__ void k_hsc_up_s(int32_t n, const int32_t *__restrict__ role, const T *__restrict__ dinv,
                           const int32_t *__restrict__ pptr, const int32_t *__restrict__ pcol,
                           const T *__restrict__ pval, int64_t k, const T *rin, const T *zin,
                           T *out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * k) return;
  const int32_t i = (int32_t)(id / k);
  const int64_t c = id - (int64_t)i * k;
  const int32_t ro = role[i];
  T acc;
  if (ro < 0) {
    acc = zin[(int64_t)(-1 - ro) * k + c];
  } else {
    acc = dinv[ro] * rin[id];
    for (int32_t e = pptr[ro]; e < pptr[ro + 1]; ++e)
      acc = fma(pval[e], zin[(int64_t)pcol[e] * k + c], acc);
  }
  out[id] = acc;
}

template <typename T>
static int hsc_vec(int64_t k) {
  if (k <= 4) return 0;  // scalar form
  if (sizeof(T) == 4) return (k % 4 == 0) ? 4 : ((k % 2 == 0) ? 2 : 1);
  return (k % 2 == 0) ? 2 : 1;
}

static dim3 hsc_grid(int64_t rows, int64_t k, int v) {
  int64_t g = (rows + kWaves - 1) / kWaves;
  g = std::max<int64_t>(1, std::min<int64_t>(g, 1024));
  if (g >= 64) g -= g % 8;
  return dim3((unsigned)g, (unsigned)((k + 64 * v - 1) / (64 * v)), 1);
}

template <typename T>
static void level_down(const cn_hsc_t::Level &L, int64_t k, const T *in, T *out, hipStream_t s);
template <typename T>
static void level_up(const cn_hsc_t::Level &L, int64_t k, const T *rin, const T *zin, T *out,
                     hipStream_t s);

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

template <typename T>
static void level_down(const cn_hsc_t::Level &L, int64_t k, const T *in, T *out, hipStream_t s) {
  if (L.nC == 0) return;
  const int v = hsc_vec<T>(k);
  if (v == 0) {
    const int64_t tot = (int64_t)L.nC * k;
    k_hsc_down_s<T><<<grid_for(tot, 256), 256, 0, s>>>(L.nC, L.C.get(), L.r_ptr.get(),
                                                     L.r_col.get(), LvVals<T>::r(L), k, in, out);
    return;
  }
  const dim3 g = hsc_grid(L.nC, k, v);
#define D_(VV) kw_hsc_down<T, VV><<<g, kBlock, 0, s>>>(L.nC, L.C.get(), L.r_ptr.get(), L.r_col.get(), LvVals<T>::r(L), k, in, out)
  if constexpr (sizeof(T) == 4) {
    if (v == 4) { D_(4); return; }
  }
  if (v >= 2) { D_(2); } else { D_(1); }
#undef D_
}

template <typename T>
static void level_up(const cn_hsc_t::Level &L, int64_t k, const T *rin, const T *zin, T *out,
                     hipStream_t s) {
  const int v = hsc_vec<T>(k);
  if (v == 0) {
    const int64_t tot = (int64_t)L.n * k;
    k_hsc_up_s<T><<<grid_for(tot, 256), 256, 0, s>>>(L.n, L.role.get(), LvVals<T>::d(L),
                                                   L.p_ptr.get(), L.p_col.get(), LvVals<T>::p(L),
                                                   k, rin, zin, out);
    return;
  }
  const dim3 g = hsc_grid(L.n, k, v);
#define U_(VV) kw_hsc_up<T, VV><<<g, kBlock, 0, s>>>(L.n, L.role.get(), LvVals<T>::d(L), L.p_ptr.get(), L.p_col.get(), LvVals<T>::p(L), k, rin, zin, out)
  if constexpr (sizeof(T) == 4) {
    if (v == 4) { U_(4); return; }
  }
  if (v >= 2) { U_(2); } else { U_(1); }
#undef U_
}

// z0 = M^-1 r0 (n0 x k); level buffers live in H->work.
template <typename T>
void hsc_apply(cn_hsc_t *H, int64_t k, const T *r0, T *z0, hipStream_t s) {
  const int32_t L = (int32_t)H->lv.size();
  // buffers: levels 1..L (n_l rows; level L is the coarse input) + coarse z
  std::vector<size_t> off(L + 2, 0);
  size_t tot = 0;
  for (int32_t l = 1; l <= L; ++l) {
    off[l] = tot;
    const int64_t nl = l < L ? H->lv[l].n : H->nc;
    tot += (size_t)nl * k;
  }
  off[L + 1] = tot;
  tot += (size_t)H->nc * k;
  const size_t bytes = tot * sizeof(T) + 256;
  if (H->work_bytes < bytes) {
    H->work.alloc(bytes);
    H->work_bytes = bytes;
  }
  T *base = reinterpret_cast<T *>(H->work.get());
  auto buf = [&](int32_t l) { return base + off[l]; };
  for (int32_t l = 0; l < L; ++l) level_down<T>(H->lv[l], k, l == 0 ? r0 : buf(l), buf(l + 1), s);
  const T *rc = L == 0 ? r0 : buf(L);
  T *zc = L == 0 ? z0 : buf(L + 1);
  if (H->nc > 0) dense_gemm_nn<T, T>(H->nc, (int)k, H->coarse_inv.get(), rc, k, zc, k, s);
  for (int32_t l = L - 1; l >= 0; --l) {
    const T *zin = l == L - 1 ? zc : buf(l + 1);
    const T *rin = l == 0 ? r0 : buf(l);
    T *out = l == 0 ? z0 : buf(l);
    level_up<T>(H->lv[l], k, rin, zin, out, s);
  }
}
template void hsc_apply<float>(cn_hsc_t *, int64_t, const float *, float *, hipStream_t);
template void hsc_apply<double>(cn_hsc_t *, int64_t, const double *, double *, hipStream_t);

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_hsc_apply(cn_hsc H, int32_t k, const double *d_r, double *d_z, void *stream) {
  CN_API_BEGIN
  if (!H || k <= 0 || !d_r || !d_z || d_r == d_z) {
    set_error("cn_hsc_apply: bad arguments");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (H->n0 > 0) hsc_apply<double>(H, k, d_r, d_z, s);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

}  // extern "C"
