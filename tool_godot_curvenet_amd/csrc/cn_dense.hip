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
typedef float v4f __attribute__((ext_vector_type(4)));

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
template <bool AT, bool BT, bool LOWER, typename TB, typename TC>
__global__ void __launch_bounds__(256)
k_gemm_f64(int m, int n, int kk, double alpha, const double *__restrict__ A, int64_t lda,
           const TB *__restrict__ B, int64_t ldb, double beta, TC *__restrict__ C,
           int64_t ldc) {
  if (LOWER && blockIdx.x > blockIdx.y) return;
  const int kbeg = 0, kend = kk;
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
    if (i < m && j < n && (!LOWER || j <= i)) {
      TC *cp = C + (int64_t)i * ldc + j;
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

// ---- the coarse product of the hierarchical Schur apply -----------------
// Z[rows[i]] = sum_j M[i][j] S[rows[j]] over the m coarse rows and k
// columns: M is the dense FP64 coarse inverse (m x m, row-major), S the
// apply's n x k block (FP32 or FP64), whose coarse rows are read where they
// lie (no gathered copy).  A block of WM x WN waves owns a BM x BN tile of Z
// and one of sp slices of the j range; each wave owns a (BM/WM) x (BN/WN)
// part of the tile as 16 x 16 accumulators of v_mfma_f64_16x16x4f64.  The
// j range goes in chunks of 16 through two LDS stages: chunk c multiplies
// out of one stage while chunk c + 1, whose loads were issued before chunk
// c's MFMAs, is stored to the other, with one barrier a chunk (one stage and
// two barriers: 254 against 236 us a launch in FP32 at m = 3,834, k = 768).
// A scheduling barrier keeps the compiler from sinking the loads below the
// MFMAs, which would leave their latency exposed, and they are
// unconditional: indices
// are clamped into range, so no load sits under a branch that would make
// the compiler wait for it.  Rows past m and columns past k compute values
// nobody stores; for j past the slice the M side is stored as 0 and the S
// side is a real (finite) row, so those terms add exactly 0.  The slice's
// coarse row numbers are staged in LDS once, so forming a chunk's S
// addresses waits on LDS, never on the chunk loads in flight.  Built with
// the VGPR form of the MFMA (accumulators in VGPRs): with the AGPR form the
// compiler copied the accumulators out of and back into AGPRs every chunk.
// Workgroups go to the XCDs round-robin, so consecutive tiles are dealt to
// each XCD in contiguous runs: the column tiles of one row tile then share
// that XCD's L2 for their M rows (at m = 3,834, k = 768 the fetch beyond L2
// falls from 1.09 GB to 0.40 GB a launch).
// Every slice writes its FP64 partial sums; k_coarse_sum adds the slices in
// slice order and stores the rows (blocks storing rows of z that other
// blocks still read would race).  Measured with the sum: 64 x 64 tiles
// (4 waves) at m = 2,043, k = 768, 2 slices: 153 us (42 TFLOP/s); 128 x 128
// tiles (8 waves) at m = 3,834, 4 slices: 452 us (50 TFLOP/s, 0.64 of the
// FP64 matrix peak), against 505 us with 64 x 64 tiles.
// The FP32 apply can take M rounded to FP32 (MT = float): the same tiles
// on v_mfma_f32_16x16x4f32, at twice the FP64 matrix rate, FP32 partial
// sums (a slice of j), added in FP64 by k_coarse_sum.
// Chunks of kCK = 16.  For FP32, chunks of 32, or a second chunk of
// look-ahead in registers, took the kernel from 7 to 4-5 waves a SIMD
// (registers, LDS) and ran 3-10 % slower.
constexpr int kCK = 16;
constexpr int kCMaxChunk = 2048;  // longest slice (LDS row-number stage)

// The 16 x 16 x 4 MFMA of each matrix type: A (16 x 4) lane l holds
// A[l % 16][l / 16], B (4 x 16) lane l holds B[l / 16][l % 16]; of the
// 16 x 16 result lane l holds column l % 16, rows row(l, v), v = 0..3.
template <typename MT> struct CoarseMfma;
template <> struct CoarseMfma<double> {
  using acc_t = v4d;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int v) { return (l >> 4) + 4 * v; }
};
template <> struct CoarseMfma<float> {
  using acc_t = v4f;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
};

template <typename T, typename MT, int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
k_coarse_mm(int m, int k, int kchunk, const MT *__restrict__ M,
            const int32_t *__restrict__ rows, const T *__restrict__ S,
            MT *__restrict__ part, const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;  // accumulators a wave
  constexpr int AE = BM * kCK / NT;  // M entries a thread stages per chunk
  constexpr int BE = kCK * BN / NT;  // S values a thread stages per chunk
  static_assert(AE >= 1 && BE >= 1 && kCK % AE == 0 && BN % BE == 0, "tile shape");
  using F = CoarseMfma<MT>;
  // row strides that spread a wave's operand reads (16 rows x 4 k of A,
  // 4 k x 16 columns of B) over distinct LDS banks
  __shared__ MT As[2][BM][kCK + (sizeof(MT) == 4 ? 4 : 1)];
  __shared__ T Bs[2][kCK][BN + 16];
  // chunks loaded ahead in registers: the 64 x 64 tiles (small k, or a
  // small coarse block) have few MFMAs a chunk to hide a load behind and
  // take 4 (at m = 12,048, k = 48: FP32 302.6 -> 295.7 us, FP64 407.4 ->
  // 378.1 us; the occupancy is set by LDS either way); the 128 x 128 tiles
  // are bound by the matrix cores and keep one (more costs them occupancy)
  constexpr int NS = BM == 64 ? 4 : 1;
  __shared__ int32_t Rs[kCMaxChunk + (NS + 1) * kCK];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  // XCD-contiguous tile order (x fastest, then y, then the slice z)
  const int nx = gridDim.x, ny = gridDim.y, nb = nx * ny * gridDim.z;
  const int hw = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const int lid = hw < nb - nb % 8 ? (hw % 8) * (nb / 8) + hw / 8 : hw;
  const int tx = lid % nx, ty = (lid / nx) % ny, tz = lid / (nx * ny);
  const int bi = ty * BM, bj = tx * BN;
  const int wi = (wave / WN) * (BM / WM), wj = (wave % WN) * (BN / WN);
  const int jb = tz * kchunk, je = min(m, jb + kchunk);
  const int ar = tid / (kCK / AE), ac = (tid % (kCK / AE)) * AE;
  const int br = tid / (BN / BE), bc = (tid % (BN / BE)) * BE;
  const MT *Mrow = M + (int64_t)min(bi + ar, m - 1) * m;
  const int cb = bj + bc;  // the thread's first S column (clamped below)
  for (int q = tid; q < kchunk + (NS + 1) * kCK; q += NT) Rs[q] = rows[min(jb + q, m - 1)];
  __syncthreads();
  MT pa[AE];
  T pb[BE];
  auto fetch = [&](int j0) {
#pragma unroll
    for (int u = 0; u < AE; ++u) pa[u] = Mrow[min(j0 + ac + u, m - 1)];
    const T *Srow = S + (int64_t)Rs[j0 - jb + br] * k;
#pragma unroll
    for (int u = 0; u < BE; ++u) pb[u] = Srow[min(cb + u, k - 1)];
  };
  typename F::acc_t acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = typename F::acc_t{0, 0, 0, 0};
  // two LDS stages: chunk c multiplies out of stage c & 1 while chunk c + 1,
  // loaded during chunk c - 1, is stored to the other stage; one barrier a
  // chunk
  auto stage = [&](int j0, int sb) {
#pragma unroll
    for (int u = 0; u < AE; ++u) As[sb][ar][ac + u] = j0 + ac + u < je ? pa[u] : MT(0);
#pragma unroll
    for (int u = 0; u < BE; ++u) Bs[sb][br][bc + u] = pb[u];
  };
  if constexpr (NS == 1) {
    fetch(jb);
    stage(jb, 0);
    __syncthreads();
    fetch(jb + kCK);
    int sb = 0;
    for (int j0 = jb; j0 < je; j0 += kCK, sb ^= 1) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < kCK; t += 4) {
        MT a[TM], b[TN];
#pragma unroll
        for (int x = 0; x < TM; ++x) a[x] = As[sb][wi + 16 * x + (l & 15)][t + (l >> 4)];
#pragma unroll
        for (int y = 0; y < TN; ++y) b[y] = (MT)Bs[sb][t + (l >> 4)][wj + 16 * y + (l & 15)];
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TN; ++y) acc[x][y] = F::mma(a[x], b[y], acc[x][y]);
      }
      stage(j0 + kCK, sb ^ 1);
      __syncthreads();  // stage sb read by every wave, stage sb ^ 1 written
      fetch(j0 + 2 * kCK);
    }
  } else {
    // register slot q holds chunk c + 1 at chunk c = q (mod NS)
    MT ra[NS][AE];
    T rb[NS][BE];
    auto fetch_to = [&](int j0, MT *qa, T *qb) {
#pragma unroll
      for (int u = 0; u < AE; ++u) qa[u] = Mrow[min(j0 + ac + u, m - 1)];
      const T *Srow = S + (int64_t)Rs[j0 - jb + br] * k;
#pragma unroll
      for (int u = 0; u < BE; ++u) qb[u] = Srow[min(cb + u, k - 1)];
    };
    auto stage_from = [&](int j0, int sb, const MT *qa, const T *qb) {
#pragma unroll
      for (int u = 0; u < AE; ++u) As[sb][ar][ac + u] = j0 + ac + u < je ? qa[u] : MT(0);
#pragma unroll
      for (int u = 0; u < BE; ++u) Bs[sb][br][bc + u] = qb[u];
    };
    fetch_to(jb, ra[0], rb[0]);
    stage_from(jb, 0, ra[0], rb[0]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NS; ++q) fetch_to(jb + (q + 1) * kCK, ra[q], rb[q]);
    int sb = 0;
    for (int jg = jb; jg < je; jg += NS * kCK) {
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const int j0 = jg + q * kCK;
        if (j0 >= je) break;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < kCK; t += 4) {
          MT a[TM], b[TN];
#pragma unroll
          for (int x = 0; x < TM; ++x) a[x] = As[sb][wi + 16 * x + (l & 15)][t + (l >> 4)];
#pragma unroll
          for (int y = 0; y < TN; ++y) b[y] = (MT)Bs[sb][t + (l >> 4)][wj + 16 * y + (l & 15)];
#pragma unroll
          for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int y = 0; y < TN; ++y) acc[x][y] = F::mma(a[x], b[y], acc[x][y]);
        }
        stage_from(j0 + kCK, sb ^ 1, ra[q], rb[q]);
        __syncthreads();
        fetch_to(j0 + (NS + 1) * kCK, ra[q], rb[q]);
        sb ^= 1;
      }
    }
  }
  MT *P = part + (int64_t)tz * m * k;
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = bi + wi + 16 * x + F::row(l, v), j = bj + wj + 16 * y + (l & 15);
        if (i < m && j < k) P[(int64_t)i * k + j] = acc[x][y][v];
      }
}

// bfloat16 (the top 16 bits of an FP32) nearest to a finite f, ties to even
__device__ __forceinline__ uint16_t bf16_rn(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef short s4 __attribute__((ext_vector_type(4)));

// The coarse inverse as two bfloat16 matrices, a = hi + lo with hi the
// nearest bfloat16 to a and lo the nearest to a - hi (FP64 difference), each
// m rows of ld >= m entries, zero past column m.
__global__ void k_split_bf16(int m, int ld, const double *__restrict__ a,
                             uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)m * ld) return;
  const int i = (int)(t / ld), j = (int)(t % ld);
  const double v = j < m ? a[(int64_t)i * m + j] : 0.0;
  const uint16_t h = bf16_rn((float)v);
  hi[t] = h;
  lo[t] = bf16_rn((float)(v - (double)bf16_f(h)));
}

// The coarse product on bfloat16 matrix cores (v_mfma_f32_16x16x16bf16_1k)
// to about FP32 accuracy: with a = a_h + a_l from the split inverse and each
// staged coarse value b split the same way in registers, a b is taken as
// a_h b_h + a_h b_l + a_l b_h (the dropped a_l b_l and the roundings of the
// low parts are 2^-16 of a b), three MFMAs into one FP32 accumulator.  The
// tile, slice and staging scheme is k_coarse_mm's with chunks of 16 j and
// both halves of the inverse staged in LDS; A lane l holds rows l % 16 and
// j = 4 (l / 16) .. + 3, B lane l the same j of column l % 16, and the
// result the layout of the 16 x 16 x 4 FP32 MFMA.  (gfx950's 16 x 16 x 32
// bfloat16 MFMA with chunks of 32 needs 150 VGPRs and 78 KB of LDS, one
// block a CU, and made the solve 4 % slower; six waves a SIMD, to hold all
// 720 blocks of the 3,834 x 768 product at once, spilled 7 VGPRs and took
// 205 against 179 us.)
template <typename T, int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
k_coarse_mm_bf3(int m, int ld, int k, int kchunk, const uint16_t *__restrict__ Mh,
                const uint16_t *__restrict__ Ml, const int32_t *__restrict__ rows,
                const T *__restrict__ S, float *__restrict__ part,
                const int32_t *__restrict__ ctl) {
  if (ctl && ctl_idle(ctl)) return;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int AE = BM * kCK / NT;  // inverse entries a thread stages per chunk
  constexpr int BE = kCK * BN / NT;  // S values a thread stages per chunk
  static_assert(kCK == 16 && AE == 4 && BN % BE == 0, "tile shape");
  __shared__ __align__(16) uint16_t Ah[2][BM][kCK + 8];
  __shared__ __align__(16) uint16_t Al[2][BM][kCK + 8];
  // B lane l reads rows 4 (l / 16) + u: a row stride of 4 mod 16 words puts
  // the 4 row groups on distinct banks
  __shared__ T Bs[2][kCK][BN + 4];
  __shared__ int32_t Rs[kCMaxChunk + 2 * kCK];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int nx = gridDim.x, ny = gridDim.y, nb = nx * ny * gridDim.z;
  const int hw = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const int lid = hw < nb - nb % 8 ? (hw % 8) * (nb / 8) + hw / 8 : hw;
  const int tx = lid % nx, ty = (lid / nx) % ny, tz = lid / (nx * ny);
  const int bi = ty * BM, bj = tx * BN;
  const int wi = (wave / WN) * (BM / WM), wj = (wave % WN) * (BN / WN);
  const int jb = tz * kchunk, je = min(m, jb + kchunk);
  const int ar = tid / (kCK / AE), ac = (tid % (kCK / AE)) * AE;
  const int br = tid / (BN / BE), bc = (tid % (BN / BE)) * BE;
  const int64_t arow = (int64_t)min(bi + ar, m - 1) * ld;
  const int cb = bj + bc;
  for (int q = tid; q < kchunk + 2 * kCK; q += NT) Rs[q] = rows[min(jb + q, m - 1)];
  __syncthreads();
  uint2 ph, pl;
  T pb[BE];
  // j0 + ac is a multiple of 4 and, where any of its 4 entries is < m,
  // at most ld - 4; past the slice the entries are zeroed when staged
  auto fetch = [&](int j0) {
    const int jj = min(j0 + ac, ld - 4);
    ph = *reinterpret_cast<const uint2 *>(Mh + arow + jj);
    pl = *reinterpret_cast<const uint2 *>(Ml + arow + jj);
    const T *Srow = S + (int64_t)Rs[j0 - jb + br] * k;
#pragma unroll
    for (int u = 0; u < BE; ++u) pb[u] = Srow[min(cb + u, k - 1)];
  };
  auto stage = [&](int j0, int sb) {
    uint2 h = ph, lo = pl;
    const int n_in = je - (j0 + ac);  // entries of the 4 inside the slice
    if (n_in < 4) {
      const uint64_t keep = n_in <= 0 ? 0ull : (~0ull >> (64 - 16 * n_in));
      uint64_t hv = ((uint64_t)h.y << 32 | h.x) & keep, lv = ((uint64_t)lo.y << 32 | lo.x) & keep;
      h = make_uint2((uint32_t)hv, (uint32_t)(hv >> 32));
      lo = make_uint2((uint32_t)lv, (uint32_t)(lv >> 32));
    }
    *reinterpret_cast<uint2 *>(&Ah[sb][ar][ac]) = h;
    *reinterpret_cast<uint2 *>(&Al[sb][ar][ac]) = lo;
#pragma unroll
    for (int u = 0; u < BE; ++u) Bs[sb][br][bc + u] = pb[u];
  };
  v4f acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = v4f{0, 0, 0, 0};
  fetch(jb);
  stage(jb, 0);
  __syncthreads();
  fetch(jb + kCK);
  int sb = 0;
  for (int j0 = jb; j0 < je; j0 += kCK, sb ^= 1) {
    __builtin_amdgcn_sched_barrier(0);
    s4 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int x = 0; x < TM; ++x) {
      ah[x] = *reinterpret_cast<const s4 *>(&Ah[sb][wi + 16 * x + (l & 15)][4 * (l >> 4)]);
      al[x] = *reinterpret_cast<const s4 *>(&Al[sb][wi + 16 * x + (l & 15)][4 * (l >> 4)]);
    }
#pragma unroll
    for (int y = 0; y < TN; ++y) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float b = (float)Bs[sb][4 * (l >> 4) + u][wj + 16 * y + (l & 15)];
        const uint16_t h = bf16_rn(b);
        bh[y][u] = (short)h;
        bl[y][u] = (short)bf16_rn(b - bf16_f(h));
      }
    }
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TN; ++y) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[x], bh[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[x], bl[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[x], bh[y], acc[x][y], 0, 0, 0);
      }
    stage(j0 + kCK, sb ^ 1);
    __syncthreads();  // stage sb read by every wave, stage sb ^ 1 written
    fetch(j0 + 2 * kCK);
  }
  float *P = part + (int64_t)tz * m * k;
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = bi + wi + 16 * x + 4 * (l >> 4) + v, j = bj + wj + 16 * y + (l & 15);
        if (i < m && j < k) P[(int64_t)i * k + j] = acc[x][y][v];
      }
}

// Z[rows[i]] = the sum of the slices' partials, in slice order.  A block
// takes 256 columns of kSumRows rows; a thread's rows are loaded together,
// slice by slice.  rz != nullptr: it also writes, per column, the sum over
// its rows of r_c z_c (the row's value in the source Sv before the store,
// times the stored value) as row blockIdx.y of rz -- the coarse term of
// r . M^-1 r.
constexpr int kSumRows = 4;

template <typename T, typename PT>
__global__ void __launch_bounds__(256)
k_coarse_sum(int m, int k, int sp, const PT *__restrict__ part,
             const int32_t *__restrict__ rows, const T *Sv, T *Z,
             const int32_t *__restrict__ ctl, double *__restrict__ rz) {
  if (ctl && ctl_idle(ctl)) return;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= k) return;
  const int64_t mk = (int64_t)m * k;
  const int i0 = blockIdx.y * kSumRows;
  // rows past m repeat row m - 1 and are not stored
  int64_t id[kSumRows], o[kSumRows];
  double sum[kSumRows];
#pragma unroll
  for (int u = 0; u < kSumRows; ++u) {
    const int i = min(i0 + u, m - 1);
    id[u] = (int64_t)i * k + j;
    o[u] = (int64_t)rows[i] * k + j;
    sum[u] = (double)part[id[u]];
  }
  for (int z = 1; z < sp; ++z)
#pragma unroll
    for (int u = 0; u < kSumRows; ++u) sum[u] += (double)part[(int64_t)z * mk + id[u]];
  double e = 0.0;
#pragma unroll
  for (int u = 0; u < kSumRows; ++u) {
    if (i0 + u >= m) break;
    const T zc = (T)sum[u];
    if (rz) e += (double)Sv[o[u]] * (double)zc;
    Z[o[u]] = zc;
  }
  if (rz) rz[(int64_t)blockIdx.y * k + j] = e;
}

// Tile shape: 128 x 128 (8 waves) for a large product, 64 x 64 (4 waves)
// otherwise.
static bool coarse_big_tiles(int m, int k) { return m >= 3072 && k >= 512; }

// Slices of the j range: the fewest that give about 3 blocks of 4 waves a
// CU (more slices cost more partial sums to add), at most 16, each at least
// two chunks and at most kCMaxChunk long.  Measured at m = 2,043 and
// k = 768: 1 / 2 / 4 / 8 slices 198 / 156 / 150 / 160 us with the sum.  At
// m = 12,048, k = 48 (FP32) 8 slices take 296 us, and 13 slices, one block
// a CU more than fits at once in two full rounds, 304 us: that product is
// bound by neither the chunk count nor the matrix cores.
static int coarse_splits(int m, int k) {
  const bool big = coarse_big_tiles(m, k);
  const int bm = big ? 128 : 64, bn = big ? 128 : 64;
  const int want = big ? 384 : 768;
  const int tiles = ((m + bm - 1) / bm) * ((k + bn - 1) / bn);
  int sp = 1;
  while (sp < 16 && tiles * sp < want && m / (2 * sp) >= 2 * kCK) sp *= 2;
  while ((m + sp - 1) / sp > kCMaxChunk - kCK) ++sp;
  return sp;
}

size_t coarse_product_work(int m, int k) {
  return (size_t)coarse_splits(m, k) * m * k;
}

template <typename T, typename MT>
static int coarse_launch(int m, int k, const MT *M, const int32_t *rows, const T *S, T *Z,
                         const int32_t *ctl, MT *part, hipStream_t s, double *rz) {
  const int sp = coarse_splits(m, k);
  // a slice is at most kCMaxChunk - kCK rows, so the rounded chunk plus one
  // chunk of look-ahead fits the kernel's row-number stage
  const int chunk = ((m + sp - 1) / sp + kCK - 1) / kCK * kCK;
  if (coarse_big_tiles(m, k)) {
    const dim3 g((unsigned)((k + 127) / 128), (unsigned)((m + 127) / 128), (unsigned)sp);
    k_coarse_mm<T, MT, 128, 128, 4, 2><<<g, 512, 0, s>>>(m, k, chunk, M, rows, S, part, ctl);
  } else {
    const dim3 g((unsigned)((k + 63) / 64), (unsigned)((m + 63) / 64), (unsigned)sp);
    k_coarse_mm<T, MT, 64, 64, 2, 2><<<g, 256, 0, s>>>(m, k, chunk, M, rows, S, part, ctl);
  }
  const dim3 gs((unsigned)((k + 255) / 256), (unsigned)((m + kSumRows - 1) / kSumRows));
  k_coarse_sum<T, MT><<<gs, 256, 0, s>>>(m, k, sp, part, rows, S, Z, ctl, rz);
  return rz ? (int)gs.y : 0;
}

// M32 != nullptr (FP32 applies only): the product takes M rounded to FP32.
// Returns the partial rows written to rz (0 with rz = nullptr).
template <typename T>
int coarse_product(int m, int k, const double *M, const float *M32, const int32_t *rows,
                   const T *S, T *Z, const int32_t *ctl, double *part, hipStream_t s,
                   double *rz) {
  if (m <= 0 || k <= 0) return 0;
  if constexpr (sizeof(T) == 4) {
    if (M32)
      return coarse_launch<T, float>(m, k, M32, rows, S, Z, ctl,
                                     reinterpret_cast<float *>(part), s, rz);
  }
  return coarse_launch<T, double>(m, k, M, rows, S, Z, ctl, part, s, rz);
}

int coarse_rz_rows(int m) { return (m + kSumRows - 1) / kSumRows; }

// The split-bfloat16 product (k_coarse_mm_bf3) runs on hierarchies that
// keep the split inverse (at least 3,072 coarse rows), in the tiles the
// FP32 product would take there.
bool coarse_bf3_applies(int m, int k) { return m >= 3072 && k > 4; }

void coarse_split_bf16(int m, int ld, const double *a, uint16_t *hi, uint16_t *lo,
                       hipStream_t s) {
  const int64_t n = (int64_t)m * ld;
  if (n == 0) return;
  k_split_bf16<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(m, ld, a, hi, lo);
  CN_HIP(hipGetLastError());
}

// z[rows] = S_c^-1 s[rows] on the split inverse (hi, lo: m rows of ld);
// FP32 applies where coarse_bf3_applies.  Returns the partial rows written
// to rz (0 with rz = nullptr).
int coarse_product_bf3(int m, int ld, int k, const uint16_t *Mh, const uint16_t *Ml,
                       const int32_t *rows, const float *S, float *Z, const int32_t *ctl,
                       double *part, hipStream_t s, double *rz) {
  if (m <= 0 || k <= 0) return 0;
  const int sp = coarse_splits(m, k);
  const int chunk = ((m + sp - 1) / sp + kCK - 1) / kCK * kCK;
  float *pf = reinterpret_cast<float *>(part);
  if (coarse_big_tiles(m, k)) {
    const dim3 g((unsigned)((k + 127) / 128), (unsigned)((m + 127) / 128), (unsigned)sp);
    k_coarse_mm_bf3<float, 128, 128, 4, 2><<<g, 512, 0, s>>>(m, ld, k, chunk, Mh, Ml, rows,
                                                             S, pf, ctl);
  } else {
    const dim3 g((unsigned)((k + 63) / 64), (unsigned)((m + 63) / 64), (unsigned)sp);
    k_coarse_mm_bf3<float, 64, 64, 2, 2><<<g, 256, 0, s>>>(m, ld, k, chunk, Mh, Ml, rows, S,
                                                           pf, ctl);
  }
  const dim3 gs((unsigned)((k + 255) / 256), (unsigned)((m + kSumRows - 1) / kSumRows));
  k_coarse_sum<float, float><<<gs, 256, 0, s>>>(m, k, sp, pf, rows, S, Z, ctl, rz);
  return rz ? (int)gs.y : 0;
}

template int coarse_product<float>(int, int, const double *, const float *, const int32_t *,
                                   const float *, float *, const int32_t *, double *,
                                   hipStream_t, double *);
template int coarse_product<double>(int, int, const double *, const float *, const int32_t *,
                                    const double *, double *, const int32_t *, double *,
                                    hipStream_t, double *);

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
