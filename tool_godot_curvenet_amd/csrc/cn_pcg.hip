// cn_pcg.hip -- preconditioned conjugate gradients on gfx950.
//
// SparseLinAlg.cg / cgWithGuess (SparseLinAlg.lean:104-163) for k
// right-hand sides in lockstep (per-column alpha/beta, as the reference's
// *Multi kernels, SlangCodegen/AxpyMulti.lean:68-69).
//
// CN_PREC_MIXED: an FP64 iterative-refinement loop
//     r = b - A x (FP64 values, FP64 vectors)      -- the true residual
//     solve A d = r / ||r|| with FP32 PCG            -- SpMM, IC(0), p/z/q in FP32
//     x += ||r|| d (FP64)
// FP32 is used for everything that streams (matrix values, factor, search
// directions); dots, alpha/beta and the stop test are FP64.
//
// One inner iteration is (C = number of IC(0) levels):
//     SpMM q = A p  (+ partial p.q)     -> reduce: alpha = rz / p.q
//     C forward passes, each fused with d += alpha p, r -= alpha q (+ r.r)
//     C backward passes (+ r.z)         -> reduce: beta, stop flags
//     p = z + beta p
// All scalars live on the device; the host only reads the active-column
// counter every `check_every` iterations.  Kernels read that counter first
// and return at once when it is zero.
//
// CN_PREC_FP64 runs the same kernels in FP64.  CN_PREC_FP64_PLAIN is the
// unfused baseline: separate SpMV / dot / axpy / preconditioner / saxpby
// launches with alpha, the stop test and beta computed on the host
// (three device->host round trips per iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>

#include "cn_internal.h"
#include "cn_kernels.h"

namespace cn {

// ================================================================ kernels ==
#define NARROW_LOOP(row, n0, n1)                                              \
  for (int64_t row = (int64_t)(n0) + (int64_t)blockIdx.x * (kBlock / kGroup) + \
                     threadIdx.x / kGroup;                                     \
       row < (n1); row += (int64_t)gridDim.x * (kBlock / kGroup))
#define WIDE_LOOP(row, n0, n1)                                                \
  for (int32_t row = (n0) + blockIdx.x * kWaves + (threadIdx.x >> 6);        \
       row < (n1); row += gridDim.x * kWaves)

// ---- SpMM with partial p.q -------------------------------------------
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
kn_spmm_dot(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
            const T *__restrict__ val, const T *__restrict__ P, T *__restrict__ Q,
            double *__restrict__ part, const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int g = threadIdx.x % kGroup;
  double pd[1] = {0.0};
  NARROW_LOOP(row, 0, n) {
    T acc[K] = {};
    gather_narrow<T, K>(ptr[row], ptr[row + 1], g, col, val, P, acc);
    if (g < K) {
      T q = pick<T, K>(acc, g);
      Q[row * K + g] = q;
      pd[0] += (double)P[row * K + g] * (double)q;
    }
  }
  block_partials_narrow<K, 1>(pd, part, 0);
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
kw_spmm_dot(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
            const T *__restrict__ val, int64_t k, const T *__restrict__ P,
            T *__restrict__ Q, double *__restrict__ part,
            const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  double pd[1][VEC] = {};
  WIDE_LOOP(row, 0, n) {
    const int32_t r = uniform(row);
    const int32_t a = uniform(ptr[r]), b = uniform(ptr[r + 1]);
    T acc[VEC] = {};
    gather_wide<T, VEC>(a, b, col, val, P, k, c0, valid, acc);
    if (valid) {
      T p[VEC];
      VecIO<T, VEC>::ld(P + (int64_t)r * k + c0, p);
      VecIO<T, VEC>::st(Q + (int64_t)r * k + c0, acc);
#pragma unroll
      for (int m = 0; m < VEC; ++m) pd[0][m] += (double)p[m] * (double)acc[m];
    }
  }
  block_partials_wide<VEC, 1>(pd, part, k, c0, 0);
}

// ---- FP64 residual r = b - A x with partial r.r ------------------------
template <int K>
__global__ void __launch_bounds__(kBlock)
kn_resid(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
         const double *__restrict__ val, const double *__restrict__ X,
         const double *__restrict__ B, double *__restrict__ R,
         double *__restrict__ part) {
  const int g = threadIdx.x % kGroup;
  double pd[1] = {0.0};
  NARROW_LOOP(row, 0, n) {
    double acc[K] = {};
    gather_narrow<double, K>(ptr[row], ptr[row + 1], g, col, val, X, acc);
    if (g < K) {
      double r = B[row * K + g] - pick<double, K>(acc, g);
      R[row * K + g] = r;
      pd[0] += r * r;
    }
  }
  block_partials_narrow<K, 1>(pd, part, 0);
}

template <int VEC>
__global__ void __launch_bounds__(kBlock)
kw_resid(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
         const double *__restrict__ val, int64_t k, const double *__restrict__ X,
         const double *__restrict__ B, double *__restrict__ R,
         double *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  double pd[1][VEC] = {};
  WIDE_LOOP(row, 0, n) {
    const int32_t r = uniform(row);
    const int32_t a = uniform(ptr[r]), b = uniform(ptr[r + 1]);
    double acc[VEC] = {};
    gather_wide<double, VEC>(a, b, col, val, X, k, c0, valid, acc);
    if (valid) {
      double bb[VEC];
      VecIO<double, VEC>::ld(B + (int64_t)r * k + c0, bb);
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        acc[m] = bb[m] - acc[m];
        pd[0][m] += acc[m] * acc[m];
      }
      VecIO<double, VEC>::st(R + (int64_t)r * k + c0, acc);
    }
  }
  block_partials_wide<VEC, 1>(pd, part, k, c0, 0);
}

// ---- IC(0) forward pass over one level, optionally fused with the CG
//      update d += alpha p, r -= alpha q; partial r.r -----------------------
template <typename T, int K, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kn_fwd(int32_t r0, int32_t r1, const int32_t *__restrict__ lp,
       const int32_t *__restrict__ lcol, const T *__restrict__ lval,
       const T *__restrict__ rdiag, const double *__restrict__ alpha,
       const T *__restrict__ P, const T *__restrict__ Q, T *__restrict__ D,
       T *__restrict__ R, T *__restrict__ Y, double *__restrict__ part,
       const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int g = threadIdx.x % kGroup;
  double pd[1] = {0.0};
  const T al = (FUSED && g < K) ? (T)alpha[g] : T(0);
  NARROW_LOOP(row, r0, r1) {
    T acc[K] = {};
    gather_narrow<T, K>(lp[row], lp[row + 1], g, lcol, lval, Y, acc);
    if (g < K) {
      const int64_t o = row * K + g;
      T r = R[o];
      if (FUSED) {
        D[o] = fma(al, P[o], D[o]);
        r = fma(-al, Q[o], r);
        R[o] = r;
      }
      pd[0] += (double)r * (double)r;
      Y[o] = (r - pick<T, K>(acc, g)) * rdiag[row];
    }
  }
  block_partials_narrow<K, 1>(pd, part, 0);
}

template <typename T, int VEC, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kw_fwd(int32_t r0, int32_t r1, const int32_t *__restrict__ lp,
       const int32_t *__restrict__ lcol, const T *__restrict__ lval,
       const T *__restrict__ rdiag, const double *__restrict__ alpha, int64_t k,
       const T *__restrict__ P, const T *__restrict__ Q, T *__restrict__ D,
       T *__restrict__ R, T *__restrict__ Y, double *__restrict__ part,
       const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  double pd[1][VEC] = {};
  T al[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) al[m] = (FUSED && valid) ? (T)alpha[c0 + m] : T(0);
  WIDE_LOOP(row, r0, r1) {
    const int32_t r = uniform(row);
    const int32_t a = uniform(lp[r]), b = uniform(lp[r + 1]);
    T acc[VEC] = {};
    gather_wide<T, VEC>(a, b, lcol, lval, Y, k, c0, valid, acc);
    if (valid) {
      const int64_t o = (int64_t)r * k + c0;
      T rv[VEC];
      VecIO<T, VEC>::ld(R + o, rv);
      if (FUSED) {
        T pv[VEC], qv[VEC], dv[VEC];
        VecIO<T, VEC>::ld(P + o, pv);
        VecIO<T, VEC>::ld(Q + o, qv);
        VecIO<T, VEC>::ld(D + o, dv);
#pragma unroll
        for (int m = 0; m < VEC; ++m) {
          dv[m] = fma(al[m], pv[m], dv[m]);
          rv[m] = fma(-al[m], qv[m], rv[m]);
        }
        VecIO<T, VEC>::st(D + o, dv);
        VecIO<T, VEC>::st(R + o, rv);
      }
      const T rd = rdiag[r];
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        pd[0][m] += (double)rv[m] * (double)rv[m];
        acc[m] = (rv[m] - acc[m]) * rd;
      }
      VecIO<T, VEC>::st(Y + o, acc);
    }
  }
  block_partials_wide<VEC, 1>(pd, part, k, c0, 0);
}

// ---- IC(0) backward pass over one level (in place on Z), partial r.z ------
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
kn_bwd(int32_t r0, int32_t r1, const int32_t *__restrict__ up,
       const int32_t *__restrict__ ucol, const T *__restrict__ uval,
       const T *__restrict__ rdiag, const T *__restrict__ R, T *__restrict__ Z,
       double *__restrict__ part, const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int g = threadIdx.x % kGroup;
  double pd[1] = {0.0};
  NARROW_LOOP(row, r0, r1) {
    T acc[K] = {};
    gather_narrow<T, K>(up[row], up[row + 1], g, ucol, uval, Z, acc);
    if (g < K) {
      const int64_t o = row * K + g;
      T z = (Z[o] - pick<T, K>(acc, g)) * rdiag[row];
      Z[o] = z;
      pd[0] += (double)R[o] * (double)z;
    }
  }
  block_partials_narrow<K, 1>(pd, part, 0);
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
kw_bwd(int32_t r0, int32_t r1, const int32_t *__restrict__ up,
       const int32_t *__restrict__ ucol, const T *__restrict__ uval,
       const T *__restrict__ rdiag, int64_t k, const T *__restrict__ R,
       T *__restrict__ Z, double *__restrict__ part,
       const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  double pd[1][VEC] = {};
  WIDE_LOOP(row, r0, r1) {
    const int32_t r = uniform(row);
    const int32_t a = uniform(up[r]), b = uniform(up[r + 1]);
    T acc[VEC] = {};
    gather_wide<T, VEC>(a, b, ucol, uval, Z, k, c0, valid, acc);
    if (valid) {
      const int64_t o = (int64_t)r * k + c0;
      T yv[VEC], rv[VEC];
      VecIO<T, VEC>::ld(Z + o, yv);
      VecIO<T, VEC>::ld(R + o, rv);
      const T rd = rdiag[r];
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        acc[m] = (yv[m] - acc[m]) * rd;
        pd[0][m] += (double)rv[m] * (double)acc[m];
      }
      VecIO<T, VEC>::st(Z + o, acc);
    }
  }
  block_partials_wide<VEC, 1>(pd, part, k, c0, 0);
}

// ---- Jacobi (SparseLinAlg.applyJacobi, zero diagonal -> 0), fused update,
//      partials r.r and r.z ------------------------------------------------
template <typename T, int K, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kn_jacobi(int32_t n, const T *__restrict__ dinv, const double *__restrict__ alpha,
          const T *__restrict__ P, const T *__restrict__ Q, T *__restrict__ D,
          T *__restrict__ R, T *__restrict__ Z, double *__restrict__ part,
          int64_t stride, const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int g = threadIdx.x % kGroup;
  double pd[2] = {0.0, 0.0};
  const T al = (FUSED && g < K) ? (T)alpha[g] : T(0);
  NARROW_LOOP(row, 0, n) {
    if (g < K) {
      const int64_t o = row * K + g;
      T r = R[o];
      if (FUSED) {
        D[o] = fma(al, P[o], D[o]);
        r = fma(-al, Q[o], r);
        R[o] = r;
      }
      T z = r * dinv[row];
      Z[o] = z;
      pd[0] += (double)r * (double)r;
      pd[1] += (double)r * (double)z;
    }
  }
  block_partials_narrow<K, 2>(pd, part, stride);
}

template <typename T, int VEC, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kw_jacobi(int32_t n, const T *__restrict__ dinv, const double *__restrict__ alpha,
          int64_t k, const T *__restrict__ P, const T *__restrict__ Q,
          T *__restrict__ D, T *__restrict__ R, T *__restrict__ Z,
          double *__restrict__ part, int64_t stride,
          const int32_t *__restrict__ n_active) {
  if (*n_active == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  double pd[2][VEC] = {};
  T al[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) al[m] = (FUSED && valid) ? (T)alpha[c0 + m] : T(0);
  WIDE_LOOP(row, 0, n) {
    if (valid) {
      const int64_t o = (int64_t)row * k + c0;
      T rv[VEC], zv[VEC];
      VecIO<T, VEC>::ld(R + o, rv);
      if (FUSED) {
        T pv[VEC], qv[VEC], dv[VEC];
        VecIO<T, VEC>::ld(P + o, pv);
        VecIO<T, VEC>::ld(Q + o, qv);
        VecIO<T, VEC>::ld(D + o, dv);
#pragma unroll
        for (int m = 0; m < VEC; ++m) {
          dv[m] = fma(al[m], pv[m], dv[m]);
          rv[m] = fma(-al[m], qv[m], rv[m]);
        }
        VecIO<T, VEC>::st(D + o, dv);
        VecIO<T, VEC>::st(R + o, rv);
      }
      const T di = dinv[row];
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        zv[m] = rv[m] * di;
        pd[0][m] += (double)rv[m] * (double)rv[m];
        pd[1][m] += (double)rv[m] * (double)zv[m];
      }
      VecIO<T, VEC>::st(Z + o, zv);
    }
  }
  block_partials_wide<VEC, 2>(pd, part, k, c0, stride);
}

// ---- generic two-vector column dot (plain path): partial a.b -------------
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
kn_dot(int32_t n, const T *__restrict__ A, const T *__restrict__ B,
       double *__restrict__ part) {
  const int g = threadIdx.x % kGroup;
  double pd[1] = {0.0};
  NARROW_LOOP(row, 0, n) {
    if (g < K) pd[0] += (double)A[row * K + g] * (double)B[row * K + g];
  }
  block_partials_narrow<K, 1>(pd, part, 0);
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
kw_dot(int32_t n, int64_t k, const T *__restrict__ A, const T *__restrict__ B,
       double *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  double pd[1][VEC] = {};
  WIDE_LOOP(row, 0, n) {
    if (valid) {
      T a[VEC], b[VEC];
      VecIO<T, VEC>::ld(A + (int64_t)row * k + c0, a);
      VecIO<T, VEC>::ld(B + (int64_t)row * k + c0, b);
#pragma unroll
      for (int m = 0; m < VEC; ++m) pd[0][m] += (double)a[m] * (double)b[m];
    }
  }
  block_partials_wide<VEC, 1>(pd, part, k, c0, 0);
}

// ---- column reductions of the [rows][k] partial buffers (fixed order) ----
__device__ __forceinline__ double col_reduce(const double *__restrict__ part,
                                             int32_t nrows, int64_t k, int64_t c,
                                             double (*sh)[64]) {
  const int s = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double v = 0.0;
  if (c < k)
    for (int32_t b = s; b < nrows; b += 4) v += part[(int64_t)b * k + c];
  sh[s][lane] = v;
  __syncthreads();
  double t = sh[0][lane];
  t += sh[1][lane];
  t += sh[2][lane];
  t += sh[3][lane];
  __syncthreads();
  return t;
}

enum { ST_ALPHA = 0, ST_BETA, ST_RZ, ST_TOL2, ST_NORM, ST_SCALE, ST_BB, ST_RR, ST_TMP, ST_N };

// alpha = rz / p.q  (breakdown p.q == 0 deactivates the column)
__global__ void __launch_bounds__(kBlock)
k_reduce_alpha(int32_t nrows, int64_t k, const double *__restrict__ part,
               double *__restrict__ st, int32_t *__restrict__ active,
               int32_t *__restrict__ ctl) {
  if (ctl[0] == 0) return;
  __shared__ double sh[4][64];
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  double pq = col_reduce(part, nrows, k, c, sh);
  if (threadIdx.x < 64 && c < k) {
    double a = 0.0;
    if (active[c]) {
      if (pq != 0.0 && isfinite(pq)) {
        a = st[ST_RZ * k + c] / pq;
      } else {
        active[c] = 0;
        atomicSub(&ctl[0], 1);
        atomicOr(&ctl[1], 1);
      }
    }
    st[ST_ALPHA * k + c] = a;
  }
}

// after z = M^-1 r: rr and rz partials; mode 0 = first (sets rz and the
// inner tolerance), mode 1 = iteration (beta, stop test)
__global__ void __launch_bounds__(kBlock)
k_reduce_beta(int32_t nrows, int64_t k, const double *__restrict__ prr,
              const double *__restrict__ prz, double *__restrict__ st,
              int32_t *__restrict__ active, int32_t *__restrict__ ctl, int mode,
              double inner_tol2) {
  if (ctl[0] == 0) return;
  __shared__ double sh[4][64];
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  double rr = col_reduce(prr, nrows, k, c, sh);
  double rz = col_reduce(prz, nrows, k, c, sh);
  if (threadIdx.x < 64 && c < k) {
    double beta = 0.0;
    if (active[c]) {
      if (mode == 0) {
        if (rr > 0.0 && isfinite(rr)) {
          st[ST_TOL2 * k + c] = inner_tol2 * rr;
          st[ST_RZ * k + c] = rz;
        } else {
          active[c] = 0;
          atomicSub(&ctl[0], 1);
        }
      } else if (!(rr > st[ST_TOL2 * k + c]) || !isfinite(rr)) {
        active[c] = 0;
        atomicSub(&ctl[0], 1);
        if (!isfinite(rr)) atomicOr(&ctl[1], 1);
      } else {
        double old = st[ST_RZ * k + c];
        beta = old == 0.0 ? 0.0 : rz / old;
        st[ST_RZ * k + c] = rz;
      }
    }
    st[ST_BETA * k + c] = beta;
  }
}

// plain reduction into st[slot]
__global__ void __launch_bounds__(kBlock)
k_reduce_plain(int32_t nrows, int64_t k, const double *__restrict__ part,
               double *__restrict__ out) {
  __shared__ double sh[4][64];
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  double v = col_reduce(part, nrows, k, c, sh);
  if (threadIdx.x < 64 && c < k) out[c] = v;
}

// ---- elementwise -----------------------------------------------------------
// P = Z + beta P
template <typename T>
__global__ void k_p_update(int64_t total, int64_t k, const double *__restrict__ beta,
                           const T *__restrict__ Z, T *__restrict__ P,
                           const int32_t *__restrict__ ctl) {
  if (ctl[0] == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = i % k;
    P[i] = fma((T)beta[c], P[i], Z[i]);
  }
}

// Y = (T)(X * scale_c)
template <typename T>
__global__ void k_scale_in(int64_t total, int64_t k, const double *__restrict__ scale,
                           const double *__restrict__ X, T *__restrict__ Y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    Y[i] = (T)(X[i] * scale[i % k]);
}

// X += norm_c * D
template <typename T>
__global__ void k_x_update(int64_t total, int64_t k, const double *__restrict__ norm,
                           const T *__restrict__ D, double *__restrict__ X) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    X[i] = fma(norm[i % k], (double)D[i], X[i]);
}

// y += a_c x  (plain axpy with per-column scalar, SlangCodegen/AxpyMulti.lean)
template <typename T>
__global__ void k_axpy(int64_t total, int64_t k, const double *__restrict__ a,
                       double sign, const T *__restrict__ x, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + (T)(sign * a[i % k]) * x[i];
}

__global__ void k_perm_gather(int32_t n, int64_t k, const int32_t *__restrict__ perm,
                              const double *__restrict__ src, double *__restrict__ dst) {
  int64_t tot = (int64_t)n * k;
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < tot;
       id += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = id / k, c = id - i * k;
    dst[id] = src[(int64_t)(perm ? perm[i] : i) * k + c];
  }
}

__global__ void k_perm_scatter(int32_t n, int64_t k, const int32_t *__restrict__ perm,
                               const double *__restrict__ src, double *__restrict__ dst) {
  int64_t tot = (int64_t)n * k;
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < tot;
       id += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = id / k, c = id - i * k;
    dst[(int64_t)(perm ? perm[i] : i) * k + c] = src[id];
  }
}

__global__ void k_dinv(int32_t n, const double *__restrict__ d, double *__restrict__ d64,
                       float *__restrict__ d32) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = d[i] == 0.0 ? 0.0 : 1.0 / d[i];
  d64[i] = v;
  d32[i] = (float)v;
}

__global__ void k_rdiag64(int32_t n, const double *__restrict__ d, double *__restrict__ r) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = 1.0 / d[i];
}

// ============================================================== profiling ==
struct Prof {
  bool on = false;
  std::vector<hipEvent_t> pool;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> recs;
  size_t used = 0;
  std::map<std::string, std::pair<int, double>> result;
};
static thread_local Prof g_prof;

struct ProfScope {
  Prof::Rec *rec = nullptr;
  hipStream_t s;
  ProfScope(const char *name, hipStream_t st) : s(st) {
    Prof &P = g_prof;
    if (!P.on) return;
    if (P.used + 2 > P.pool.size()) {
      if (P.pool.size() >= 65536) return;
      size_t add = std::max<size_t>(256, P.pool.size());
      for (size_t i = 0; i < add; ++i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        P.pool.push_back(e);
      }
    }
    hipEvent_t a = P.pool[P.used++], b = P.pool[P.used++];
    P.recs.push_back({name, a, b});
    rec = &P.recs.back();
    (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (rec) (void)hipEventRecord(rec->b, s);
  }
};

static void prof_reset() {
  g_prof.recs.clear();
  g_prof.recs.reserve(32768);
  g_prof.used = 0;
  g_prof.result.clear();
}

static void prof_collect() {
  Prof &P = g_prof;
  if (!P.on) return;
  for (auto &r : P.recs) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess &&
        hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto &e = P.result[r.name];
      e.first += 1;
      e.second += ms;
    }
  }
  P.recs.clear();
  P.used = 0;
}

// ============================================================== workspace ==
struct SolverWork {
  int32_t n = 0, k = 0, part_rows = 0;
  DevBuf<double> xp, rp, bp;          // FP64 n x k: x, true residual, b
  DevBuf<double> buf[5];              // inner D, R, Z, P, Q (8 bytes/elem)
  DevBuf<double> part;                // 3 x part_rows x k
  DevBuf<double> st;                  // ST_N x k per-column scalars
  DevBuf<int32_t> active;             // k
  DevBuf<int32_t> ctl;                // [0] active count, [1] breakdown, [2] always 1
  DevBuf<double> dinv64;
  DevBuf<float> dinv32;
  int32_t *h_ctl = nullptr;           // pinned
  double *h_st = nullptr;             // pinned, ST_N x k
  ~SolverWork() {
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_st) (void)hipHostFree(h_st);
  }
};

void free_solver_work(SolverWork *w) { delete w; }

static SolverWork *ensure_work(cn_system_t *S, int32_t k, int32_t part_rows,
                               hipStream_t s) {
  SolverWork *W = S->work;
  const int32_t n = S->n_free;
  if (W && W->n == n && W->k >= k && W->part_rows >= part_rows) return W;
  int32_t kk = std::max(k, W ? W->k : 0);
  int32_t pr = std::max(part_rows, W ? W->part_rows : 0);
  delete W;
  S->work = nullptr;
  W = new SolverWork();
  W->n = n;
  W->k = kk;
  W->part_rows = pr;
  size_t nk = (size_t)n * kk + 4;
  W->xp.alloc(nk);
  W->rp.alloc(nk);
  W->bp.alloc(nk);
  for (auto &b : W->buf) b.alloc(nk);
  W->part.alloc((size_t)3 * pr * kk + 4);
  W->st.alloc((size_t)ST_N * kk + 4);
  W->active.alloc((size_t)kk + 4);
  W->ctl.alloc(8);
  W->dinv64.alloc((size_t)n + 1);
  W->dinv32.alloc((size_t)n + 1);
  CN_HIP(hipHostMalloc((void **)&W->h_ctl, sizeof(int32_t) * 8, hipHostMallocDefault));
  CN_HIP(hipHostMalloc((void **)&W->h_st, sizeof(double) * ((size_t)ST_N * kk + 4),
                       hipHostMallocDefault));
  if (n > 0)
    k_dinv<<<grid_for(n, 256), 256, 0, s>>>(n, S->diag.get(), W->dinv64.get(),
                                           W->dinv32.get());
  CN_HIP(hipGetLastError());
  S->work = W;
  return W;
}

// ================================================================ launches ==
static unsigned blocks_for_rows(int64_t rows, int rows_per_block, unsigned cap) {
  int64_t g = (rows + rows_per_block - 1) / rows_per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// grid for the row mappings; rows per block-iteration: narrow 32, wide 4
struct Grid {
  dim3 g;
  int32_t rows;  // partial rows written (= g.x)
};

template <typename T>
static int wide_vec(int64_t k) {
  if (sizeof(T) == 4) return (k % 4 == 0) ? 4 : ((k % 2 == 0) ? 2 : 1);
  return (k % 2 == 0) ? 2 : 1;
}

template <typename T>
static Grid grid_rows(int64_t rows, int64_t k, unsigned cap) {
  Grid G;
  if (k <= 4) {
    G.g = dim3(blocks_for_rows(rows, kBlock / kGroup, cap), 1, 1);
  } else {
    int v = wide_vec<T>(k);
    G.g = dim3(blocks_for_rows(rows, kWaves, cap),
               (unsigned)((k + 64 * v - 1) / (64 * v)), 1);
  }
  G.rows = (int32_t)G.g.x;
  return G;
}

static const unsigned kCapSpmm = 1024;
static const unsigned kCapLevel = 512;

#define DISPATCH_NARROW(K_, KERNEL, ...)                       \
  switch (K_) {                                                \
    case 1: KERNEL<1> __VA_ARGS__; break;                      \
    case 2: KERNEL<2> __VA_ARGS__; break;                      \
    case 3: KERNEL<3> __VA_ARGS__; break;                      \
    default: KERNEL<4> __VA_ARGS__; break;                     \
  }

// Matrix + preconditioner view of one solve (permuted if IC(0)).
template <typename T>
struct SysView {
  int32_t n;
  const int32_t *ptr, *col;
  const T *val;
  const double *val64;
  // IC(0)
  const cn_ic0_t *ic0 = nullptr;
  const int32_t *lp, *lcol, *up, *ucol;
  const T *lval, *uval, *rdiag;
  // Jacobi
  const T *dinv;
};

template <typename T>
static void launch_spmm_dot(const SysView<T> &V, int64_t k, const T *P, T *Q,
                            double *part, const int32_t *ctl, hipStream_t s,
                            int32_t *rows_out) {
  Grid G = grid_rows<T>(V.n, k, kCapSpmm);
  *rows_out = G.rows;
  ProfScope ps("spmm", s);
  if (k <= 4) {
#define KN_SPMM(KK) kn_spmm_dot<T, KK><<<G.g, kBlock, 0, s>>>(V.n, V.ptr, V.col, V.val, P, Q, part, ctl)
    switch (k) { case 1: KN_SPMM(1); break; case 2: KN_SPMM(2); break; case 3: KN_SPMM(3); break; default: KN_SPMM(4); }
#undef KN_SPMM
  } else {
    int v = wide_vec<T>(k);
    if (v == 4) kw_spmm_dot<T, (sizeof(T) == 4 ? 4 : 2)><<<G.g, kBlock, 0, s>>>(V.n, V.ptr, V.col, V.val, k, P, Q, part, ctl);
    else if (v == 2) kw_spmm_dot<T, 2><<<G.g, kBlock, 0, s>>>(V.n, V.ptr, V.col, V.val, k, P, Q, part, ctl);
    else kw_spmm_dot<T, 1><<<G.g, kBlock, 0, s>>>(V.n, V.ptr, V.col, V.val, k, P, Q, part, ctl);
  }
}

static void launch_resid(int32_t n, const int32_t *ptr, const int32_t *col,
                         const double *val, int64_t k, const double *X,
                         const double *B, double *R, double *part, hipStream_t s,
                         int32_t *rows_out) {
  Grid G = grid_rows<double>(n, k, kCapSpmm);
  *rows_out = G.rows;
  ProfScope ps("resid", s);
  if (k <= 4) {
#define KN_RES(KK) kn_resid<KK><<<G.g, kBlock, 0, s>>>(n, ptr, col, val, X, B, R, part)
    switch (k) { case 1: KN_RES(1); break; case 2: KN_RES(2); break; case 3: KN_RES(3); break; default: KN_RES(4); }
#undef KN_RES
  } else if (k % 2 == 0) {
    kw_resid<2><<<G.g, kBlock, 0, s>>>(n, ptr, col, val, k, X, B, R, part);
  } else {
    kw_resid<1><<<G.g, kBlock, 0, s>>>(n, ptr, col, val, k, X, B, R, part);
  }
}

// z = M^-1 r (fused with the CG update when FUSED); partials rr into
// part[0 .. rows), rz into part[stride ...].  Returns partial rows.
template <typename T, bool FUSED>
static int32_t launch_precond(const SysView<T> &V, int64_t k, const double *alpha,
                              const T *P, const T *Q, T *D, T *R, T *Z,
                              double *part, int64_t stride, const int32_t *ctl,
                              hipStream_t s) {
  const int v = wide_vec<T>(k);
  if (!V.ic0) {
    Grid G = grid_rows<T>(V.n, k, kCapSpmm);
    ProfScope ps("jacobi", s);
    if (k <= 4) {
#define KN_J(KK) kn_jacobi<T, KK, FUSED><<<G.g, kBlock, 0, s>>>(V.n, V.dinv, alpha, P, Q, D, R, Z, part, stride, ctl)
      switch (k) { case 1: KN_J(1); break; case 2: KN_J(2); break; case 3: KN_J(3); break; default: KN_J(4); }
#undef KN_J
    } else if (v == 4) {
      kw_jacobi<T, (sizeof(T) == 4 ? 4 : 2), FUSED><<<G.g, kBlock, 0, s>>>(V.n, V.dinv, alpha, k, P, Q, D, R, Z, part, stride, ctl);
    } else if (v == 2) {
      kw_jacobi<T, 2, FUSED><<<G.g, kBlock, 0, s>>>(V.n, V.dinv, alpha, k, P, Q, D, R, Z, part, stride, ctl);
    } else {
      kw_jacobi<T, 1, FUSED><<<G.g, kBlock, 0, s>>>(V.n, V.dinv, alpha, k, P, Q, D, R, Z, part, stride, ctl);
    }
    return G.rows;
  }
  const std::vector<int32_t> &lv = V.ic0->level_ptr;
  const int32_t nl = (int32_t)lv.size() - 1;
  int32_t off = 0;
  {
    ProfScope ps("fwd", s);
    for (int32_t c = 0; c < nl; ++c) {
      int32_t r0 = lv[c], r1 = lv[c + 1];
      if (r1 <= r0) continue;
      Grid G = grid_rows<T>(r1 - r0, k, kCapLevel);
      double *pp = part + (int64_t)off * k;
      if (k <= 4) {
#define KN_F(KK) kn_fwd<T, KK, FUSED><<<G.g, kBlock, 0, s>>>(r0, r1, V.lp, V.lcol, V.lval, V.rdiag, alpha, P, Q, D, R, Z, pp, ctl)
        switch (k) { case 1: KN_F(1); break; case 2: KN_F(2); break; case 3: KN_F(3); break; default: KN_F(4); }
#undef KN_F
      } else if (v == 4) {
        kw_fwd<T, (sizeof(T) == 4 ? 4 : 2), FUSED><<<G.g, kBlock, 0, s>>>(r0, r1, V.lp, V.lcol, V.lval, V.rdiag, alpha, k, P, Q, D, R, Z, pp, ctl);
      } else if (v == 2) {
        kw_fwd<T, 2, FUSED><<<G.g, kBlock, 0, s>>>(r0, r1, V.lp, V.lcol, V.lval, V.rdiag, alpha, k, P, Q, D, R, Z, pp, ctl);
      } else {
        kw_fwd<T, 1, FUSED><<<G.g, kBlock, 0, s>>>(r0, r1, V.lp, V.lcol, V.lval, V.rdiag, alpha, k, P, Q, D, R, Z, pp, ctl);
      }
      off += G.rows;
    }
  }
  int32_t off2 = 0;
  {
    ProfScope ps("bwd", s);
    for (int32_t c = nl - 1; c >= 0; --c) {
      int32_t r0 = lv[c], r1 = lv[c + 1];
      if (r1 <= r0) continue;
      Grid G = grid_rows<T>(r1 - r0, k, kCapLevel);
      double *pp = part + stride + (int64_t)off2 * k;
      if (k <= 4) {
#define KN_B(KK) kn_bwd<T, KK><<<G.g, kBlock, 0, s>>>(r0, r1, V.up, V.ucol, V.uval, V.rdiag, R, Z, pp, ctl)
        switch (k) { case 1: KN_B(1); break; case 2: KN_B(2); break; case 3: KN_B(3); break; default: KN_B(4); }
#undef KN_B
      } else if (v == 4) {
        kw_bwd<T, (sizeof(T) == 4 ? 4 : 2)><<<G.g, kBlock, 0, s>>>(r0, r1, V.up, V.ucol, V.uval, V.rdiag, k, R, Z, pp, ctl);
      } else if (v == 2) {
        kw_bwd<T, 2><<<G.g, kBlock, 0, s>>>(r0, r1, V.up, V.ucol, V.uval, V.rdiag, k, R, Z, pp, ctl);
      } else {
        kw_bwd<T, 1><<<G.g, kBlock, 0, s>>>(r0, r1, V.up, V.ucol, V.uval, V.rdiag, k, R, Z, pp, ctl);
      }
      off2 += G.rows;
    }
  }
  return std::max(off, off2);  // both passes write the same row count
}

// partial rows the preconditioner needs (for workspace sizing)
template <typename T>
static int32_t precond_rows(const SysView<T> &V, int64_t k) {
  if (!V.ic0) return grid_rows<T>(V.n, k, kCapSpmm).rows;
  const std::vector<int32_t> &lv = V.ic0->level_ptr;
  int32_t off = 0;
  for (size_t c = 0; c + 1 < lv.size(); ++c)
    if (lv[c + 1] > lv[c]) off += grid_rows<T>(lv[c + 1] - lv[c], k, kCapLevel).rows;
  return off;
}

template <typename T>
static void launch_dot(int32_t n, int64_t k, const T *A, const T *B, double *part,
                       hipStream_t s, int32_t *rows_out) {
  Grid G = grid_rows<T>(n, k, kCapSpmm);
  *rows_out = G.rows;
  if (k <= 4) {
#define KN_D(KK) kn_dot<T, KK><<<G.g, kBlock, 0, s>>>(n, A, B, part)
    switch (k) { case 1: KN_D(1); break; case 2: KN_D(2); break; case 3: KN_D(3); break; default: KN_D(4); }
#undef KN_D
  } else {
    int v = wide_vec<T>(k);
    if (v == 4) kw_dot<T, (sizeof(T) == 4 ? 4 : 2)><<<G.g, kBlock, 0, s>>>(n, k, A, B, part);
    else if (v == 2) kw_dot<T, 2><<<G.g, kBlock, 0, s>>>(n, k, A, B, part);
    else kw_dot<T, 1><<<G.g, kBlock, 0, s>>>(n, k, A, B, part);
  }
}

static unsigned ew_blocks(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ========================================================= inner PCG solve ==
// Solves A D = R (R given, D zeroed) for the active columns.  Returns the
// number of iterations run.  `active`/ctl[0] must be set by the caller.
template <typename T>
static int32_t inner_pcg(const SysView<T> &V, SolverWork *W, int64_t k,
                         int32_t max_it, double inner_tol, int32_t check_every,
                         hipStream_t s) {
  const int32_t n = V.n;
  const int64_t total = (int64_t)n * k;
  T *D = reinterpret_cast<T *>(W->buf[0].get());
  T *R = reinterpret_cast<T *>(W->buf[1].get());
  T *Z = reinterpret_cast<T *>(W->buf[2].get());
  T *P = reinterpret_cast<T *>(W->buf[3].get());
  T *Q = reinterpret_cast<T *>(W->buf[4].get());
  double *st = W->st.get();
  int32_t *ctl = W->ctl.get();
  const int64_t stride = (int64_t)W->part_rows * k;
  double *p0 = W->part.get(), *p1 = p0 + stride, *p2 = p1 + stride;
  const unsigned red_blocks = (unsigned)((k + 63) / 64);

  // z0 = M^-1 r0, rz0, rr0
  int32_t pr = launch_precond<T, false>(V, k, st + ST_ALPHA * k, P, Q, D, R, Z, p1,
                                        stride, ctl, s);
  k_reduce_beta<<<red_blocks, kBlock, 0, s>>>(pr, k, p1, p1 + stride, st,
                                              W->active.get(), ctl, 0,
                                              inner_tol * inner_tol);
  CN_HIP(hipMemcpyAsync(P, Z, sizeof(T) * total, hipMemcpyDeviceToDevice, s));
  int32_t it = 0;
  while (it < max_it) {
    int32_t rows = 0;
    launch_spmm_dot<T>(V, k, P, Q, p0, ctl, s, &rows);
    k_reduce_alpha<<<red_blocks, kBlock, 0, s>>>(rows, k, p0, st, W->active.get(), ctl);
    pr = launch_precond<T, true>(V, k, st + ST_ALPHA * k, P, Q, D, R, Z, p1, stride,
                                 ctl, s);
    k_reduce_beta<<<red_blocks, kBlock, 0, s>>>(pr, k, p1, p1 + stride, st,
                                                W->active.get(), ctl, 1, 0.0);
    k_p_update<T><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_BETA * k, Z, P, ctl);
    ++it;
    if (it % check_every == 0 || it == max_it) {
      CN_HIP(hipMemcpyAsync(W->h_ctl, ctl, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
      CN_HIP(hipStreamSynchronize(s));
      if (W->h_ctl[0] == 0) break;
    }
  }
  (void)p2;
  CN_HIP(hipGetLastError());
  return it;
}

}  // namespace cn

using namespace cn;

// ============================================================ entry point ==
extern "C" {

void cn_pcg_default_opts(cn_pcg_opts *o) {
  if (!o) return;
  o->precond = CN_PRECOND_IC0;
  o->precision = CN_PREC_MIXED;
  o->tol_rel = 1e-9;
  o->max_iter = 20000;
  o->inner_tol = 1e-5;
  o->max_refine = 30;
  o->check_every = 8;
  o->warm_start = 0;
}

void cn_profile_enable(int32_t on) {
  g_prof.on = on != 0;
  prof_reset();
}

int32_t cn_profile_query(const char *name, double *ms_total) {
  auto it = g_prof.result.find(name ? name : "");
  if (it == g_prof.result.end()) {
    if (ms_total) *ms_total = 0.0;
    return 0;
  }
  if (ms_total) *ms_total = it->second.second;
  return it->second.first;
}

}  // extern "C"

namespace cn {

template <typename T>
static SysView<T> make_view(cn_system_t *S, cn_ic0_t *F, SolverWork *W);

template <>
SysView<float> make_view<float>(cn_system_t *S, cn_ic0_t *F, SolverWork *W) {
  SysView<float> V{};
  V.n = S->n_free;
  if (F) {
    V.ptr = F->a_ptr.get(); V.col = F->a_col.get(); V.val = F->a_val32.get();
    V.val64 = F->a_val.get();
    V.ic0 = F;
    V.lp = F->l_ptr.get(); V.lcol = F->l_col.get(); V.lval = F->l_val32.get();
    V.up = F->u_ptr.get(); V.ucol = F->u_col.get(); V.uval = F->u_val32.get();
    V.rdiag = F->l_rdiag32.get();
  } else {
    V.ptr = S->ff_ptr.get(); V.col = S->ff_col.get(); V.val = S->ff_val32.get();
    V.val64 = S->ff_val.get();
  }
  V.dinv = W->dinv32.get();
  return V;
}

template <>
SysView<double> make_view<double>(cn_system_t *S, cn_ic0_t *F, SolverWork *W) {
  SysView<double> V{};
  V.n = S->n_free;
  if (F) {
    V.ptr = F->a_ptr.get(); V.col = F->a_col.get(); V.val = F->a_val.get();
    V.val64 = F->a_val.get();
    V.ic0 = F;
    V.lp = F->l_ptr.get(); V.lcol = F->l_col.get(); V.lval = F->l_val.get();
    V.up = F->u_ptr.get(); V.ucol = F->u_col.get(); V.uval = F->u_val.get();
    V.rdiag = F->l_rdiag64.get();
  } else {
    V.ptr = S->ff_ptr.get(); V.col = S->ff_col.get(); V.val = S->ff_val.get();
    V.val64 = S->ff_val.get();
  }
  V.dinv = W->dinv64.get();
  return V;
}

// refinement loop shared by MIXED (T = float) and FP64 (T = double)
template <typename T>
static cn_status refine_solve(cn_system_t *S, cn_ic0_t *F, const cn_pcg_opts *o,
                              int32_t k, const double *d_b, double *d_x,
                              cn_solve_info *info, hipStream_t s) {
  const int32_t n = S->n_free;
  const int64_t total = (int64_t)n * k;
  SysView<T> V0{};
  V0.n = n;
  V0.ic0 = F;
  int32_t need = std::max<int32_t>(precond_rows<T>(V0, k),
                                   grid_rows<T>(n, k, kCapSpmm).rows);
  need = std::max<int32_t>(need, grid_rows<double>(n, k, kCapSpmm).rows);
  SolverWork *W = ensure_work(S, k, need, s);
  SysView<T> V = make_view<T>(S, F, W);
  const int32_t *perm = F ? F->perm.get() : nullptr;
  double *st = W->st.get();
  double *hs = W->h_st;
  const int64_t stride = (int64_t)W->part_rows * k;
  const unsigned red_blocks = (unsigned)((k + 63) / 64);
  int32_t rows = 0;
  k_dinv<<<grid_for(n, 256), 256, 0, s>>>(n, S->diag.get(), W->dinv64.get(),
                                         W->dinv32.get());

  k_perm_gather<<<ew_blocks(total), 256, 0, s>>>(n, k, perm, d_b, W->bp.get());
  if (o->warm_start)
    k_perm_gather<<<ew_blocks(total), 256, 0, s>>>(n, k, perm, d_x, W->xp.get());
  else
    CN_HIP(hipMemsetAsync(W->xp.get(), 0, sizeof(double) * total, s));
  launch_dot<double>(n, k, W->bp.get(), W->bp.get(), W->part.get(), s, &rows);
  k_reduce_plain<<<red_blocks, kBlock, 0, s>>>(rows, k, W->part.get(), st + ST_BB * k);
  launch_resid(n, V.ptr, V.col, V.val64, k, W->xp.get(), W->bp.get(), W->rp.get(),
               W->part.get(), s, &rows);
  k_reduce_plain<<<red_blocks, kBlock, 0, s>>>(rows, k, W->part.get(), st + ST_RR * k);
  CN_HIP(hipGetLastError());

  int32_t iters = 0, refine = 0;
  cn_status status = CN_OK;
  double relmax = 0.0;
  std::vector<int32_t> act(k);
  std::vector<double> scale(k), norm(k);
  bool first = true;
  for (;;) {
    CN_HIP(hipMemcpyAsync(hs + ST_BB * k, st + ST_BB * k, sizeof(double) * 2 * k,
                          hipMemcpyDeviceToHost, s));  // BB and RR are adjacent
    CN_HIP(hipStreamSynchronize(s));
    relmax = 0.0;
    double relmax_active = 0.0;
    int32_t nact = 0;
    for (int32_t c = 0; c < k; ++c) {
      double bb = hs[ST_BB * k + c], rr = hs[ST_RR * k + c];
      double rel = bb > 0.0 ? std::sqrt(rr / bb) : (rr > 0.0 ? INFINITY : 0.0);
      if (!std::isfinite(rel)) rel = INFINITY;
      relmax = std::max(relmax, rel);
      bool a = rel > o->tol_rel;
      act[c] = a ? 1 : 0;
      scale[c] = a ? 1.0 / std::sqrt(rr) : 0.0;
      norm[c] = a ? std::sqrt(rr) : 0.0;
      if (a) { ++nact; relmax_active = std::max(relmax_active, rel); }
    }
    if (nact == 0) { status = CN_OK; break; }
    if (!std::isfinite(relmax)) { status = CN_BREAKDOWN; break; }
    if (refine >= o->max_refine || iters >= o->max_iter) { status = CN_MAXIT; break; }
    // inner relative target: enough to reach tol with a 10x margin, never
    // tighter than FP32 can deliver (MIXED)
    double inner = 0.1 * o->tol_rel / relmax_active;
    if (sizeof(T) == 4) inner = std::max(inner, o->inner_tol);
    else inner = std::max(inner, 1e-15);
    CN_HIP(hipMemcpyAsync(W->active.get(), act.data(), sizeof(int32_t) * k,
                          hipMemcpyHostToDevice, s));
    W->h_ctl[0] = nact;
    W->h_ctl[1] = 0;
    W->h_ctl[2] = 1;
    CN_HIP(hipMemcpyAsync(W->ctl.get(), W->h_ctl, sizeof(int32_t) * 3,
                          hipMemcpyHostToDevice, s));
    std::memcpy(hs + ST_SCALE * k, scale.data(), sizeof(double) * k);
    std::memcpy(hs + ST_NORM * k, norm.data(), sizeof(double) * k);
    CN_HIP(hipMemcpyAsync(st + ST_NORM * k, hs + ST_NORM * k, sizeof(double) * 2 * k,
                          hipMemcpyHostToDevice, s));  // NORM, SCALE adjacent
    T *D = reinterpret_cast<T *>(W->buf[0].get());
    T *R = reinterpret_cast<T *>(W->buf[1].get());
    CN_HIP(hipMemsetAsync(D, 0, sizeof(T) * total, s));
    k_scale_in<T><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_SCALE * k,
                                                   W->rp.get(), R);
    int32_t budget = o->max_iter - iters;
    int32_t it = inner_pcg<T>(V, W, k, budget, inner, std::max(1, o->check_every), s);
    iters += it;
    k_x_update<T><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_NORM * k, D,
                                                   W->xp.get());
    launch_resid(n, V.ptr, V.col, V.val64, k, W->xp.get(), W->bp.get(), W->rp.get(),
                 W->part.get(), s, &rows);
    k_reduce_plain<<<red_blocks, kBlock, 0, s>>>(rows, k, W->part.get(), st + ST_RR * k);
    CN_HIP(hipGetLastError());
    ++refine;
    if (it <= 1 && W->h_ctl[1]) { status = CN_BREAKDOWN; break; }
  }
  k_perm_scatter<<<ew_blocks(total), 256, 0, s>>>(n, k, perm, W->xp.get(), d_x);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  (void)stride;
  if (info) {
    info->status = status;
    info->iters = iters;
    info->refinements = refine;
    info->relres_max = relmax;
  }
  return status;
}

// Unfused FP64 baseline: one launch per step, scalars on the host.
static cn_status plain_solve(cn_system_t *S, cn_ic0_t *F, const cn_pcg_opts *o,
                             int32_t k, const double *d_b, double *d_x,
                             cn_solve_info *info, hipStream_t s) {
  const int32_t n = S->n_free;
  const int64_t total = (int64_t)n * k;
  SysView<double> V0{};
  V0.n = n;
  V0.ic0 = F;
  int32_t need = std::max<int32_t>(precond_rows<double>(V0, k),
                                   grid_rows<double>(n, k, kCapSpmm).rows);
  SolverWork *W = ensure_work(S, k, need, s);
  SysView<double> V = make_view<double>(S, F, W);
  k_dinv<<<grid_for(n, 256), 256, 0, s>>>(n, S->diag.get(), W->dinv64.get(),
                                         W->dinv32.get());
  const int32_t *perm = F ? F->perm.get() : nullptr;
  double *st = W->st.get();
  double *hs = W->h_st;
  const int64_t stride = (int64_t)W->part_rows * k;
  const unsigned red_blocks = (unsigned)((k + 63) / 64);
  double *X = W->xp.get(), *B = W->bp.get(), *R = W->rp.get();
  double *Z = W->buf[2].get(), *P = W->buf[3].get(), *Q = W->buf[4].get();
  int32_t *ctl = W->ctl.get();
  W->h_ctl[0] = 1; W->h_ctl[1] = 0; W->h_ctl[2] = 1;
  CN_HIP(hipMemcpyAsync(ctl, W->h_ctl, sizeof(int32_t) * 3, hipMemcpyHostToDevice, s));
  int32_t rows = 0;
  auto dot_to_host = [&](const double *a, const double *b, double *h) {
    launch_dot<double>(n, k, a, b, W->part.get(), s, &rows);
    k_reduce_plain<<<red_blocks, kBlock, 0, s>>>(rows, k, W->part.get(), st + ST_TMP * k);
    CN_HIP(hipMemcpyAsync(h, st + ST_TMP * k, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
  };
  auto apply_m = [&]() {  // Z = M^-1 R (no fused update)
    launch_precond<double, false>(V, k, st + ST_ALPHA * k, P, Q, X, R, Z, W->part.get(),
                                  stride, ctl, s);
  };
  k_perm_gather<<<ew_blocks(total), 256, 0, s>>>(n, k, perm, d_b, B);
  if (o->warm_start) {
    k_perm_gather<<<ew_blocks(total), 256, 0, s>>>(n, k, perm, d_x, X);
    launch_resid(n, V.ptr, V.col, V.val64, k, X, B, R, W->part.get(), s, &rows);
  } else {
    CN_HIP(hipMemsetAsync(X, 0, sizeof(double) * total, s));
    CN_HIP(hipMemcpyAsync(R, B, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
  }
  std::vector<double> bb(k), rz(k), tmp(k), alpha(k), beta(k), rr(k);
  dot_to_host(B, B, bb.data());
  apply_m();
  CN_HIP(hipMemcpyAsync(P, Z, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
  dot_to_host(R, Z, rz.data());
  std::vector<int32_t> act(k, 1);
  dot_to_host(R, R, rr.data());
  for (int32_t c = 0; c < k; ++c)
    if (!(rr[c] > o->tol_rel * o->tol_rel * bb[c])) act[c] = 0;
  int32_t it = 0;
  cn_status status = CN_MAXIT;
  bool any = false;
  for (int32_t c = 0; c < k; ++c) any |= act[c] != 0;
  if (!any) status = CN_OK;
  while (any && it < o->max_iter) {
    launch_spmm_dot<double>(V, k, P, Q, W->part.get(), ctl, s, &rows);   // SpMV
    dot_to_host(P, Q, tmp.data());                                        // p.Ap
    for (int32_t c = 0; c < k; ++c) {
      if (act[c] && tmp[c] == 0.0) { act[c] = 0; status = CN_BREAKDOWN; }
      alpha[c] = act[c] ? rz[c] / tmp[c] : 0.0;
    }
    CN_HIP(hipMemcpyAsync(st + ST_ALPHA * k, alpha.data(), sizeof(double) * k,
                          hipMemcpyHostToDevice, s));
    k_axpy<double><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_ALPHA * k, 1.0, P, X);
    k_axpy<double><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_ALPHA * k, -1.0, Q, R);
    ++it;
    dot_to_host(R, R, rr.data());                                         // stop test
    any = false;
    for (int32_t c = 0; c < k; ++c) {
      if (act[c] && !(rr[c] > o->tol_rel * o->tol_rel * bb[c])) act[c] = 0;
      any |= act[c] != 0;
    }
    if (!any) { if (status == CN_MAXIT) status = CN_OK; break; }
    apply_m();                                                            // z = M^-1 r
    dot_to_host(R, Z, tmp.data());                                        // r.z
    for (int32_t c = 0; c < k; ++c) {
      beta[c] = act[c] ? (rz[c] == 0.0 ? 0.0 : tmp[c] / rz[c]) : 0.0;
      if (act[c]) rz[c] = tmp[c];
    }
    CN_HIP(hipMemcpyAsync(st + ST_BETA * k, beta.data(), sizeof(double) * k,
                          hipMemcpyHostToDevice, s));
    k_p_update<double><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_BETA * k, Z, P, ctl);
  }
  // true residual
  launch_resid(n, V.ptr, V.col, V.val64, k, X, B, R, W->part.get(), s, &rows);
  k_reduce_plain<<<red_blocks, kBlock, 0, s>>>(rows, k, W->part.get(), st + ST_TMP * k);
  CN_HIP(hipMemcpyAsync(tmp.data(), st + ST_TMP * k, sizeof(double) * k,
                        hipMemcpyDeviceToHost, s));
  k_perm_scatter<<<ew_blocks(total), 256, 0, s>>>(n, k, perm, X, d_x);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  double relmax = 0.0;
  for (int32_t c = 0; c < k; ++c) {
    double rel = bb[c] > 0.0 ? std::sqrt(tmp[c] / bb[c]) : 0.0;
    relmax = std::max(relmax, rel);
  }
  if (status == CN_OK && relmax > o->tol_rel * 10.0) status = CN_MAXIT;
  if (info) {
    info->status = status;
    info->iters = it;
    info->refinements = 0;
    info->relres_max = relmax;
  }
  (void)hs;
  return status;
}

}  // namespace cn

extern "C" cn_status cn_pcg_solve(cn_system S, cn_ic0 F, const cn_pcg_opts *o,
                                  int32_t k, const double *d_b, double *d_x,
                                  cn_solve_info *info, void *stream) {
  CN_API_BEGIN
  if (!S || !o || k <= 0 || !d_b || !d_x) {
    set_error("cn_pcg_solve: bad arguments");
    return CN_BAD_ARG;
  }
  if (o->precond == CN_PRECOND_IC0 && !F) {
    set_error("cn_pcg_solve: IC0 preconditioner needs a factor");
    return CN_BAD_ARG;
  }
  if (F && F->n != S->n_free) {
    set_error("cn_pcg_solve: factor does not match the system");
    return CN_BAD_ARG;
  }
  if (o->precond != CN_PRECOND_IC0) F = nullptr;
  hipStream_t s = as_stream(stream);
  if (S->n_free == 0) {
    if (info) { info->status = CN_OK; info->iters = 0; info->refinements = 0; info->relres_max = 0.0; }
    return CN_OK;
  }
  if (F && !F->l_rdiag64.get()) {
    F->l_rdiag64.alloc((size_t)F->n + 1);
    k_rdiag64<<<grid_for(F->n, 256), 256, 0, s>>>(F->n, F->l_diag.get(), F->l_rdiag64.get());
    CN_HIP(hipGetLastError());
  }
  prof_reset();
  cn_status st;
  switch (o->precision) {
    case CN_PREC_FP64_PLAIN: st = plain_solve(S, F, o, k, d_b, d_x, info, s); break;
    case CN_PREC_FP64: st = refine_solve<double>(S, F, o, k, d_b, d_x, info, s); break;
    case CN_PREC_MIXED: st = refine_solve<float>(S, F, o, k, d_b, d_x, info, s); break;
    default:
      set_error("cn_pcg_solve: unknown precision %d", o->precision);
      return CN_BAD_ARG;
  }
  prof_collect();
  return st;
  CN_API_END
}
