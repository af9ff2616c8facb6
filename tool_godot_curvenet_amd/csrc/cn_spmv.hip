// cn_spmv.hip -- CSR SpMV / SpMM entry points.
//
// FP64 forms reproduce SparseLinAlg.spmv exactly (SparseLinAlg.lean:44-53):
// each output is summed left to right from 0.0 with separate multiply and
// add (this file is built with -ffp-contract=off).  One thread per
// (row, column), walking its row's entries one at a time: a wavefront's
// val/col loads of one step land in 64 different rows (a strided pattern;
// the lines are used in full only across the row's later steps, through the
// cache).  This form is for exactness, not speed.
//
// FP32 forms use fused multiply-adds in entry order: a thread per row for
// k <= 4, a wavefront per (row, column tile) above (see cn_kernels.h).
#include <hip/hip_runtime.h>

#include "cn_internal.h"
#include "cn_kernels.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace cn {

__global__ void k_spmm_exact(int32_t rows, int32_t k, const int32_t *ptr,
                             const int32_t *col, const double *val,
                             const double *X, double *Y) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)rows * k) return;
  int32_t i = (int32_t)(id / k), c = (int32_t)(id - (int64_t)i * k);
  double s = 0.0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p)
    s = s + val[p] * X[(int64_t)col[p] * k + c];
  Y[id] = s;
}

// Thread-per-row form (k <= 4): one thread owns a row and its K columns.
// The row's entries are taken 8 at a time: the chunk's columns and weights,
// then all its gathers, are issued back to back (the tail of a chunk
// re-reads the row's last entry with weight 0, so every load is
// unconditional and the sum is the plain in-order one); the row pointers of
// the thread's next row are loaded while the current row's gathers are in
// flight.  A wavefront reads 64 consecutive rows' index data, which covers
// whole cache lines across its 8 loads.
template <typename T, int K>
__global__ void __launch_bounds__(kBlock) kt_spmm(int32_t n, const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ col,
                                                  const T *__restrict__ val,
                                                  const T *__restrict__ X, T *__restrict__ Y) {
  constexpr int CH = 8;
  const int32_t stride = (int32_t)(gridDim.x * blockDim.x);
  int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= n) return;
  int32_t a = ptr[r], b = ptr[r + 1];
  while (true) {
    const int32_t rn = r + stride;
    const bool more = rn < n;
    int32_t an = 0, bn = 0;
    if (more) { an = ptr[rn]; bn = ptr[rn + 1]; }
    T acc[K] = {};
    for (int32_t e = a; e < b; e += CH) {
      int32_t j[CH];
      T v[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int32_t q = min(e + u, b - 1);
        j[u] = col[q];
        v[u] = (e + u < b) ? val[q] : T(0);
      }
      T x[CH][K];
#pragma unroll
      for (int u = 0; u < CH; ++u)
#pragma unroll
        for (int c = 0; c < K; ++c) x[u][c] = X[(int64_t)j[u] * K + c];
#pragma unroll
      for (int u = 0; u < CH; ++u)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] = fma(v[u], x[u][c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < K; ++c) Y[(int64_t)r * K + c] = acc[c];
    if (!more) break;
    r = rn;
    a = an;
    b = bn;
  }
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock) kw_spmm(int32_t n, const int32_t *ptr,
                                                  const int32_t *col, const T *val,
                                                  int64_t k, const T *X, T *Y) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c0 = (int64_t)blockIdx.y * 64 * VEC + lane * VEC;
  const bool valid = c0 < k;
  for (int32_t row = blockIdx.x * kWaves + w; row < n; row += gridDim.x * kWaves) {
    const int32_t r = uniform(row);
    const int32_t a = uniform(ptr[r]), b = uniform(ptr[r + 1]);
    T acc[VEC] = {};
    gather_wide<T, VEC>(a, b, col, val, X, k, c0, valid, acc);
    if (valid) VecIO<T, VEC>::st(Y + (int64_t)r * k + c0, acc);
  }
}

static unsigned row_blocks(int64_t rows, int rows_per_block) {
  int64_t g = (rows + rows_per_block - 1) / rows_per_block;
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  return (unsigned)g;
}

template <typename T>
static void launch_spmm_fast(int32_t n, const int32_t *ptr, const int32_t *col,
                             const T *val, int32_t k, const T *X, T *Y,
                             hipStream_t s) {
  if (n <= 0) return;
  if (k <= 4) {
    // a thread per row, one wave of resident blocks at most (the threads
    // walk on by the grid stride)
    const unsigned gx = (unsigned)std::min<int64_t>(((int64_t)n + kBlock - 1) / kBlock, 8192);
    switch (k) {
      case 1: kt_spmm<T, 1><<<gx, kBlock, 0, s>>>(n, ptr, col, val, X, Y); break;
      case 2: kt_spmm<T, 2><<<gx, kBlock, 0, s>>>(n, ptr, col, val, X, Y); break;
      case 3: kt_spmm<T, 3><<<gx, kBlock, 0, s>>>(n, ptr, col, val, X, Y); break;
      default: kt_spmm<T, 4><<<gx, kBlock, 0, s>>>(n, ptr, col, val, X, Y); break;
    }
  } else {
    unsigned gx = row_blocks(n, kWaves);
    constexpr int V = sizeof(T) == 4 ? 4 : 2;
    if (k % V == 0) {
      dim3 g(gx, (unsigned)((k + 64 * V - 1) / (64 * V)));
      kw_spmm<T, V><<<g, kBlock, 0, s>>>(n, ptr, col, val, k, X, Y);
    } else {
      dim3 g(gx, (unsigned)((k + 63) / 64));
      kw_spmm<T, 1><<<g, kBlock, 0, s>>>(n, ptr, col, val, k, X, Y);
    }
  }
}

template void launch_spmm_fast<float>(int32_t, const int32_t *, const int32_t *,
                                      const float *, int32_t, const float *,
                                      float *, hipStream_t);
template void launch_spmm_fast<double>(int32_t, const int32_t *, const int32_t *,
                                       const double *, int32_t, const double *,
                                       double *, hipStream_t);

}  // namespace cn

using namespace cn;

extern "C" {

static cn_status spmm_args(int32_t rows, const void *ptr, const void *col,
                           const void *val, int32_t k, const void *X, void *Y) {
  if (rows < 0 || k <= 0 || (rows > 0 && (!ptr || !col || !val || !X || !Y))) {
    set_error("spmv/spmm: bad arguments");
    return CN_BAD_ARG;
  }
  return CN_OK;
}

cn_status cn_spmv_f64(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const double *d_val, const double *d_x, double *d_y,
                      void *stream) {
  return cn_spmm_f64(rows, d_ptr, d_col, d_val, 1, d_x, d_y, stream);
}

cn_status cn_spmm_f64(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const double *d_val, int32_t k, const double *d_X,
                      double *d_Y, void *stream) {
  CN_API_BEGIN
  cn_status st = spmm_args(rows, d_ptr, d_col, d_val, k, d_X, d_Y);
  if (st != CN_OK) return st;
  hipStream_t s = as_stream(stream);
  int64_t tot = (int64_t)rows * k;
  if (tot > 0)
    k_spmm_exact<<<grid_for(tot, 256), 256, 0, s>>>(rows, k, d_ptr, d_col, d_val,
                                                   d_X, d_Y);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_spmv_f32(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const float *d_val, const float *d_x, float *d_y,
                      void *stream) {
  return cn_spmm_f32(rows, d_ptr, d_col, d_val, 1, d_x, d_y, stream);
}

cn_status cn_spmm_f32(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const float *d_val, int32_t k, const float *d_X,
                      float *d_Y, void *stream) {
  CN_API_BEGIN
  cn_status st = spmm_args(rows, d_ptr, d_col, d_val, k, d_X, d_Y);
  if (st != CN_OK) return st;
  launch_spmm_fast<float>(rows, d_ptr, d_col, d_val, k, d_X, d_Y, as_stream(stream));
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

}  // extern "C"
