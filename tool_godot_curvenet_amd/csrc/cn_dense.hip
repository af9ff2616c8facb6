// cn_dense.hip -- dense FP64 Cholesky factor and multi-RHS solve on the
// matrix cores (DenseLinAlg.choleskyFactor / forwardSolve / backwardSolve /
// solveWithCholesky, DenseLinAlg.lean:255-286), used for the coarsest level
// of the hierarchical Schur preconditioner.
//
// Right-looking blocked Cholesky, panel width 32: the diagonal block is
// factored by one workgroup in LDS, the panel below it is solved against
// L11^T (thread per row), and the trailing lower triangle is updated with
// the FP64 MFMA GEMM below (A22 -= L21 L21^T).  The solve runs block
// forward / backward substitution with the same GEMM for the off-diagonal
// blocks.  FP64 throughout; the accumulation order differs from the spec's
// left-looking loops, so results agree to rounding, not bit for bit.
//
// MFMA operand layout (v_mfma_f64_16x16x4f64, checked on gfx950): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] of a 16x4 by 4x16 step
// and holds D[l / 16 + 4 v][l % 16], v = 0..3, of the 16x16 result.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "cn_internal.h"
#include "cn_kernels.h"

namespace cn {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kGT = 32;  // GEMM block tile (rows and columns)
constexpr int kGK = 16;  // GEMM k chunk staged in LDS
constexpr int kNB = 32;  // Cholesky panel width

template <typename T>
__device__ __forceinline__ double to_d(T v) { return (double)v; }

// C[m x n] = beta C + alpha op(A)[m x kk] op(B)[kk x n], FP64 MFMA.
// op(A)(i, t) = AT ? A[t * lda + i] : A[i * lda + t]
// op(B)(t, j) = BT ? B[j * ldb + t] : B[t * ldb + j]
// LOWER: only tiles on or below the diagonal, only entries j <= i written.
// Block 256 threads = 4 waves; wave w owns the 16x16 subtile (w / 2, w % 2).
// C = alpha A B + beta C.  With part != nullptr (split K), block z sums the
// k-range [z kchunk, (z + 1) kchunk) and writes alpha times its sum to
// part[z][i][j] instead of C.
template <bool AT, bool BT, bool LOWER, typename TB, typename TC>
__global__ void __launch_bounds__(256)
k_gemm_f64(int m, int n, int kk, double alpha, const double *__restrict__ A, int64_t lda,
           const TB *__restrict__ B, int64_t ldb, double beta, TC *__restrict__ C,
           int64_t ldc, const int32_t *__restrict__ crow = nullptr,
           const int32_t *__restrict__ ctl = nullptr, double *__restrict__ part = nullptr,
           int kchunk = 0) {
  if (LOWER && blockIdx.x > blockIdx.y) return;
  if (ctl && ctl_idle(ctl)) return;
  const int kbeg = part ? (int)blockIdx.z * kchunk : 0;
  const int kend = part ? min(kk, kbeg + kchunk) : kk;
  __shared__ double As[kGT][kGK + 1];
  __shared__ double Bs[kGK][kGT + 1];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int bi = blockIdx.y * kGT, bj = blockIdx.x * kGT;
  const int wi = (wave >> 1) * 16, wj = (wave & 1) * 16;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = kbeg; k0 < kend; k0 += kGK) {
    // stage A (32 x 16) and B (16 x 32): two elements of each per thread
    for (int e = tid; e < kGT * kGK; e += 256) {
      const int r = e / kGK, t = e % kGK;
      const int gi = bi + r, gt = k0 + t;
      double a = 0.0;
      if (gi < m && gt < kend) a = AT ? A[(int64_t)gt * lda + gi] : A[(int64_t)gi * lda + gt];
      As[r][t] = a;
      const int t2 = e / kGT, c = e % kGT;
      const int gt2 = k0 + t2, gj = bj + c;
      double b = 0.0;
      if (gj < n && gt2 < kend) b = to_d(BT ? B[(int64_t)gj * ldb + gt2] : B[(int64_t)gt2 * ldb + gj]);
      Bs[t2][c] = b;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kGK; t += 4) {
      const double a = As[wi + (l & 15)][t + (l >> 4)];
      const double b = Bs[t + (l >> 4)][wj + (l & 15)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = bi + wi + (l >> 4) + 4 * v, j = bj + wj + (l & 15);
    if (part) {
      if (i < m && j < n) part[((int64_t)blockIdx.z * m + i) * n + j] = alpha * acc[v];
      continue;
    }
    if (i < m && j < n && (!LOWER || j <= i)) {
      TC *cp = C + (int64_t)(crow ? crow[i] : i) * ldc + j;
      const double old = beta == 0.0 ? 0.0 : to_d(*cp);
      *cp = (TC)(beta * old + alpha * acc[v]);
    }
  }
}

// Unblocked Cholesky of the b x b diagonal block at (j0, j0), in LDS.  A
// non-positive (or non-finite) pivot records its column in *fail (min).
__global__ void __launch_bounds__(64)
k_potrf_diag(double *A, int64_t lda, int j0, int b, int *fail) {
  __shared__ double s[kNB][kNB + 1];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  for (int e = t; e < b * b; e += 64) {
    const int i = e / b, j = e % b;
    s[i][j] = j <= i ? A[(int64_t)(j0 + i) * lda + j0 + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    if (t == 0) {
      const double d = s[j][j];
      if (!(d > 0.0) || !isfinite(d)) {
        bad = 1;
        atomicMin(fail, j0 + j);
        s[j][j] = 1.0;
      } else {
        s[j][j] = sqrt(d);
      }
    }
    __syncthreads();
    if (t > j && t < b) s[t][j] = s[t][j] / s[j][j];
    __syncthreads();
    for (int e = t; e < b * b; e += 64) {
      const int i = e / b, c = e % b;
      if (c > j && c <= i) s[i][c] -= s[i][j] * s[c][j];
    }
    __syncthreads();
  }
  for (int e = t; e < b * b; e += 64) {
    const int i = e / b, j = e % b;
    if (j <= i) A[(int64_t)(j0 + i) * lda + j0 + j] = s[i][j];
  }
}

// Rows [j0 + b, n) of the panel: X <- X L11^-T (forward substitution per row).
__global__ void __launch_bounds__(64)
k_trsm_panel(double *A, int64_t lda, int j0, int b, int n) {
  __shared__ double L[kNB][kNB + 1];
  for (int e = threadIdx.x; e < b * b; e += 64) {
    const int i = e / b, j = e % b;
    L[i][j] = j <= i ? A[(int64_t)(j0 + i) * lda + j0 + j] : 0.0;
  }
  __syncthreads();
  const int r = j0 + b + blockIdx.x * 64 + threadIdx.x;
  if (r >= n) return;
  double x[kNB];
  double *row = A + (int64_t)r * lda + j0;
  for (int t = 0; t < b; ++t) x[t] = row[t];
  for (int t = 0; t < b; ++t) {
    double v = x[t];
    for (int u = 0; u < t; ++u) v -= x[u] * L[t][u];
    x[t] = v / L[t][t];
  }
  for (int t = 0; t < b; ++t) row[t] = x[t];
}

__global__ void k_zero_upper(int n, double *A, int64_t lda) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * n) return;
  const int i = (int)(id / n), j = (int)(id % n);
  if (j > i) A[(int64_t)i * lda + j] = 0.0;
}

// B[j0 .. j0 + b) <- L_bb^-1 B (UPPER = false) or L_bb^-T B (UPPER = true),
// one thread per right-hand-side column.
template <bool UPPER>
__global__ void __launch_bounds__(64)
k_trsv_block(const double *L, int64_t lda, int j0, int b, int k, double *B, int64_t ldb) {
  __shared__ double s[kNB][kNB + 1];
  for (int e = threadIdx.x; e < b * b; e += 64) {
    const int i = e / b, j = e % b;
    s[i][j] = j <= i ? L[(int64_t)(j0 + i) * lda + j0 + j] : 0.0;
  }
  __syncthreads();
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= k) return;
  double x[kNB];
  for (int t = 0; t < b; ++t) x[t] = B[(int64_t)(j0 + t) * ldb + c];
  if (!UPPER) {
    for (int t = 0; t < b; ++t) {
      double v = x[t];
      for (int u = 0; u < t; ++u) v -= s[t][u] * x[u];
      x[t] = v / s[t][t];
    }
  } else {
    for (int t = b - 1; t >= 0; --t) {
      double v = x[t];
      for (int u = t + 1; u < b; ++u) v -= s[u][t] * x[u];
      x[t] = v / s[t][t];
    }
  }
  for (int t = 0; t < b; ++t) B[(int64_t)(j0 + t) * ldb + c] = x[t];
}

static dim3 gemm_grid(int m, int n) {
  return dim3((unsigned)((n + kGT - 1) / kGT), (unsigned)((m + kGT - 1) / kGT), 1);
}

// In-place lower Cholesky; returns the first failing column or -1.
int dense_cholesky(int n, double *A, int64_t lda, hipStream_t s) {
  DevBuf<int> fail(1);
  const int big = 0x7fffffff;
  CN_HIP(hipMemcpyAsync(fail.get(), &big, sizeof(int), hipMemcpyHostToDevice, s));
  for (int j0 = 0; j0 < n; j0 += kNB) {
    const int b = std::min(kNB, n - j0), rem = n - j0 - b;
    k_potrf_diag<<<1, 64, 0, s>>>(A, lda, j0, b, fail.get());
    if (rem > 0) {
      k_trsm_panel<<<(rem + 63) / 64, 64, 0, s>>>(A, lda, j0, b, n);
      double *P = A + (int64_t)(j0 + b) * lda + j0;
      double *A22 = A + (int64_t)(j0 + b) * lda + j0 + b;
      k_gemm_f64<false, true, true, double, double><<<gemm_grid(rem, rem), 256, 0, s>>>(
          rem, rem, b, -1.0, P, lda, P, lda, 1.0, A22, lda);
    }
  }
  if (n > 0) k_zero_upper<<<grid_for((int64_t)n * n, 256), 256, 0, s>>>(n, A, lda);
  CN_HIP(hipGetLastError());
  int h = big;
  CN_HIP(hipMemcpyAsync(&h, fail.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  return h == big ? -1 : h;
}

// B <- (L L^T)^-1 B for B n x k (ldb), L lower (lda).
void dense_cholesky_solve(int n, int k, const double *L, int64_t lda, double *B, int64_t ldb,
                          hipStream_t s) {
  for (int j0 = 0; j0 < n; j0 += kNB) {
    const int b = std::min(kNB, n - j0);
    if (j0 > 0)
      k_gemm_f64<false, false, false, double, double><<<gemm_grid(b, k), 256, 0, s>>>(
          b, k, j0, -1.0, L + (int64_t)j0 * lda, lda, B, ldb, 1.0, B + (int64_t)j0 * ldb, ldb);
    k_trsv_block<false><<<(k + 63) / 64, 64, 0, s>>>(L, lda, j0, b, k, B, ldb);
  }
  const int nblk = (n + kNB - 1) / kNB;
  for (int q = nblk - 1; q >= 0; --q) {
    const int j0 = q * kNB, b = std::min(kNB, n - j0), rest = n - j0 - b;
    if (rest > 0)
      k_gemm_f64<true, false, false, double, double><<<gemm_grid(b, k), 256, 0, s>>>(
          b, k, rest, -1.0, L + (int64_t)(j0 + b) * lda + j0, lda, B + (int64_t)(j0 + b) * ldb,
          ldb, 1.0, B + (int64_t)j0 * ldb, ldb);
    k_trsv_block<true><<<(k + 63) / 64, 64, 0, s>>>(L, lda, j0, b, k, B, ldb);
  }
  CN_HIP(hipGetLastError());
}

// Z (m x k, TC) = M (m x m, FP64) X (m x k, TB) on the matrix cores.
template <typename TB, typename TC>
void dense_gemm_nn(int m, int k, const double *M, const TB *X, int64_t ldx, TC *Z, int64_t ldz,
                   hipStream_t s) {
  k_gemm_f64<false, false, false, TB, TC><<<gemm_grid(m, k), 256, 0, s>>>(
      m, k, m, 1.0, M, m, X, ldx, 0.0, Z, ldz);
}
template void dense_gemm_nn<double, double>(int, int, const double *, const double *, int64_t,
                                            double *, int64_t, hipStream_t);
template void dense_gemm_nn<float, float>(int, int, const double *, const float *, int64_t,
                                          float *, int64_t, hipStream_t);

// Same product written to rows zrows[i] of Z.
// Split of the K range for an m x m by m x k product: a grid of a few
// hundred 32 x 32 tiles walks K one 16-wide step at a time, each step a
// global load, an LDS round and a barrier; with few tiles (moderate k) the
// K walk is split over up to 16 blocks a tile and the partial sums are
// added in order by k_gemm_reduce.
static int gemm_splits(int m, int k) {
  const int tiles = ((m + kGT - 1) / kGT) * ((k + kGT - 1) / kGT);
  int sp = 1;
  while (sp < 16 && tiles * sp < 1024 && m / (2 * sp) >= 4 * kGK) sp *= 2;
  return sp;
}

size_t dense_gemm_rows_work(int m, int k) {
  const int sp = gemm_splits(m, k);
  return sp > 1 ? (size_t)sp * m * k : 0;
}

template <typename TC>
__global__ void k_gemm_reduce(int m, int n, int sp, const double *__restrict__ part,
                              TC *__restrict__ C, int64_t ldc,
                              const int32_t *__restrict__ crow,
                              const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)m * n) return;
  const int i = (int)(id / n), j = (int)(id - (int64_t)i * n);
  double s = part[id];
  for (int z = 1; z < sp; ++z) s += part[(int64_t)z * m * n + id];
  C[(int64_t)(crow ? crow[i] : i) * ldc + j] = (TC)s;
}

template <typename TB, typename TC>
void dense_gemm_nn_rows(int m, int k, const double *M, const TB *X, int64_t ldx, TC *Z,
                        int64_t ldz, const int32_t *zrows, const int32_t *ctl, double *work,
                        hipStream_t s) {
  const int sp = work ? gemm_splits(m, k) : 1;
  if (sp <= 1) {
    k_gemm_f64<false, false, false, TB, TC><<<gemm_grid(m, k), 256, 0, s>>>(
        m, k, m, 1.0, M, m, X, ldx, 0.0, Z, ldz, zrows, ctl);
    return;
  }
  const int chunk = ((m + sp - 1) / sp + kGK - 1) / kGK * kGK;
  dim3 g = gemm_grid(m, k);
  g.z = (unsigned)sp;
  k_gemm_f64<false, false, false, TB, TC><<<g, 256, 0, s>>>(
      m, k, m, 1.0, M, m, X, ldx, 0.0, Z, ldz, zrows, ctl, work, chunk);
  k_gemm_reduce<TC><<<grid_for((int64_t)m * k, 256), 256, 0, s>>>(m, k, sp, work, Z, ldz,
                                                                  zrows, ctl);
}
template void dense_gemm_nn_rows<double, double>(int, int, const double *, const double *,
                                                 int64_t, double *, int64_t, const int32_t *,
                                                 const int32_t *, double *, hipStream_t);
template void dense_gemm_nn_rows<float, float>(int, int, const double *, const float *, int64_t,
                                               float *, int64_t, const int32_t *, const int32_t *,
                                               double *, hipStream_t);

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_dense_cholesky_f64(int32_t n, double *d_A, int32_t lda, void *stream) {
  CN_API_BEGIN
  if (n < 0 || (n > 0 && (!d_A || lda < n))) {
    set_error("cn_dense_cholesky_f64: bad arguments");
    return CN_BAD_ARG;
  }
  const int f = dense_cholesky(n, d_A, lda, as_stream(stream));
  if (f >= 0) {
    set_error("cn_dense_cholesky_f64: non-positive pivot at column %d", f);
    return CN_NOT_SPD;
  }
  return CN_OK;
  CN_API_END
}

cn_status cn_dense_cholesky_solve_f64(int32_t n, int32_t k, const double *d_L, int32_t lda,
                                      double *d_B, int32_t ldb, void *stream) {
  CN_API_BEGIN
  if (n < 0 || k < 0 || (n > 0 && k > 0 && (!d_L || !d_B || lda < n || ldb < k))) {
    set_error("cn_dense_cholesky_solve_f64: bad arguments");
    return CN_BAD_ARG;
  }
  if (n == 0 || k == 0) return CN_OK;
  hipStream_t s = as_stream(stream);
  dense_cholesky_solve(n, k, d_L, lda, d_B, ldb, s);
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

}  // extern "C"
