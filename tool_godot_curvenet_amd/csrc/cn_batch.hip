// cn_batch.hip -- many independent systems as one block-diagonal system.
//
// The deform path solves one mesh per call (HarmonicSolve.solveMulti,
// HarmonicSolve.lean:51-66).  A set of uneven meshes solved one call each
// leaves the GPU idle on the small ones (every launch a few microseconds of
// latency); as one block-diagonal system every solver launch covers all of
// them.  The solver keeps one CG column per (segment, column) pair, so each
// segment follows its own recurrence and stops on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cn_internal.h"

namespace cn {

// dst[row0 + i] = src[i] + nnz0, seg_of_row[row0 + i] = seg for i < n
__global__ void k_batch_ptr(int32_t n, const int32_t *__restrict__ src, int32_t row0,
                            int32_t nnz0, int32_t seg, int32_t *__restrict__ dst,
                            int32_t *__restrict__ seg_of_row) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[row0 + i] = src[i] + nnz0;
  seg_of_row[row0 + i] = seg;
}

__global__ void k_batch_col(int64_t nnz, const int32_t *__restrict__ src, int32_t off,
                            int32_t *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) dst[i] = src[i] + off;
}

// Dot-grid blocks per segment: about one block per 512 rows overall (at
// most 4096 blocks), proportional to the segment's rows, at least one.
static void segment_blocks(const std::vector<int32_t> &seg_ptr, std::vector<int32_t> &bptr,
                           std::vector<int32_t> &bseg) {
  const int32_t nseg = (int32_t)seg_ptr.size() - 1;
  const int64_t N = seg_ptr[nseg];
  const int64_t B = std::max<int64_t>(1, std::min<int64_t>(4096, (N + 511) / 512));
  bptr.assign((size_t)nseg + 1, 0);
  for (int32_t s = 0; s < nseg; ++s) {
    const int64_t ns = seg_ptr[s + 1] - seg_ptr[s];
    const int64_t b = N > 0 ? std::max<int64_t>(1, (B * ns + N / 2) / N) : 1;
    bptr[s + 1] = bptr[s] + (int32_t)b;
  }
  bseg.resize((size_t)bptr[nseg]);
  for (int32_t s = 0; s < nseg; ++s)
    for (int32_t b = bptr[s]; b < bptr[s + 1]; ++b) bseg[b] = s;
}

// a single system: one segment covering every row
void ensure_segments(cn_system_t *S, hipStream_t s) {
  if (!S->seg_ptr.empty()) return;
  S->nseg = 1;
  S->seg_ptr = {0, S->n_free};
  std::vector<int32_t> bptr, bseg;
  segment_blocks(S->seg_ptr, bptr, bseg);
  S->seg_blocks = bptr.back();
  S->d_seg_ptr.alloc(2);
  S->blk_ptr.alloc(bptr.size());
  S->blk_seg.alloc(bseg.size());
  CN_HIP(hipMemcpyAsync(S->d_seg_ptr.get(), S->seg_ptr.data(), sizeof(int32_t) * 2,
                        hipMemcpyHostToDevice, s));
  CN_HIP(hipMemcpyAsync(S->blk_ptr.get(), bptr.data(), sizeof(int32_t) * bptr.size(),
                        hipMemcpyHostToDevice, s));
  CN_HIP(hipMemcpyAsync(S->blk_seg.get(), bseg.data(), sizeof(int32_t) * bseg.size(),
                        hipMemcpyHostToDevice, s));
  CN_HIP(hipStreamSynchronize(s));
}

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_system_batch(int32_t nsys, const cn_system *systems, cn_system *out,
                          void *stream) {
  CN_API_BEGIN
  if (!out || nsys <= 0 || !systems) {
    set_error("cn_system_batch: bad arguments");
    return CN_BAD_ARG;
  }
  *out = nullptr;
  int64_t n = 0, nnz = 0;
  for (int32_t i = 0; i < nsys; ++i) {
    if (!systems[i]) {
      set_error("cn_system_batch: system %d is null", i);
      return CN_BAD_ARG;
    }
    if (systems[i]->nseg > 1) {
      set_error("cn_system_batch: system %d is itself a batch", i);
      return CN_BAD_ARG;
    }
    n += systems[i]->n_free;
    nnz += systems[i]->nnz_ff;
  }
  if (n > INT32_MAX - 1 || nnz > INT32_MAX - 1) {
    set_error("cn_system_batch: %lld rows / %lld nonzeros exceed int32 indexing",
              (long long)n, (long long)nnz);
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  cn_system_t *B = new cn_system_t();
  try {
    B->n_free = (int32_t)n;
    B->nnz_ff = nnz;
    B->nseg = nsys;
    B->ff_ptr.alloc((size_t)n + 1);
    B->ff_col.alloc((size_t)nnz + 1);
    B->ff_val.alloc((size_t)nnz + 1);
    B->ff_val32.alloc((size_t)nnz + 1);
    B->fc_ptr.alloc((size_t)n + 1);
    B->fc_col.alloc(1);
    B->fc_val.alloc(1);
    B->fc_val32.alloc(1);
    B->diag.alloc((size_t)n + 1);
    B->seg_of_row.alloc((size_t)n + 1);
    B->seg_ptr.assign((size_t)nsys + 1, 0);
    int32_t row0 = 0, nnz0 = 0;
    for (int32_t i = 0; i < nsys; ++i) {
      const cn_system_t *A = systems[i];
      const int32_t ni = A->n_free, zi = (int32_t)A->nnz_ff;
      if (ni > 0) {
        k_batch_ptr<<<grid_for(ni, 256), 256, 0, s>>>(ni, A->ff_ptr.get(), row0, nnz0, i,
                                                      B->ff_ptr.get(), B->seg_of_row.get());
        CN_HIP(hipMemcpyAsync(B->diag.get() + row0, A->diag.get(), sizeof(double) * ni,
                              hipMemcpyDeviceToDevice, s));
      }
      if (zi > 0) {
        k_batch_col<<<grid_for(zi, 256), 256, 0, s>>>(zi, A->ff_col.get(), row0,
                                                      B->ff_col.get() + nnz0);
        CN_HIP(hipMemcpyAsync(B->ff_val.get() + nnz0, A->ff_val.get(), sizeof(double) * zi,
                              hipMemcpyDeviceToDevice, s));
        CN_HIP(hipMemcpyAsync(B->ff_val32.get() + nnz0, A->ff_val32.get(),
                              sizeof(float) * zi, hipMemcpyDeviceToDevice, s));
      }
      row0 += ni;
      nnz0 += zi;
      B->seg_ptr[i + 1] = row0;
    }
    // closing row pointer; no constrained columns (right-hand sides come
    // from the member systems)
    CN_HIP(hipMemcpyAsync(B->ff_ptr.get() + n, &nnz0, sizeof(int32_t), hipMemcpyHostToDevice, s));
    CN_HIP(hipMemsetAsync(B->fc_ptr.get(), 0, sizeof(int32_t) * ((size_t)n + 1), s));
    std::vector<int32_t> bptr, bseg;
    segment_blocks(B->seg_ptr, bptr, bseg);
    B->seg_blocks = bptr.back();
    B->d_seg_ptr.alloc(B->seg_ptr.size());
    B->blk_ptr.alloc(bptr.size());
    B->blk_seg.alloc(bseg.size());
    CN_HIP(hipMemcpyAsync(B->d_seg_ptr.get(), B->seg_ptr.data(),
                          sizeof(int32_t) * B->seg_ptr.size(), hipMemcpyHostToDevice, s));
    CN_HIP(hipMemcpyAsync(B->blk_ptr.get(), bptr.data(), sizeof(int32_t) * bptr.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipMemcpyAsync(B->blk_seg.get(), bseg.data(), sizeof(int32_t) * bseg.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipGetLastError());
    CN_HIP(hipStreamSynchronize(s));
    batch_rhs_table(B, systems, nsys, s);
  } catch (...) {
    delete B;
    throw;
  }
  *out = B;
  return CN_OK;
  CN_API_END
}

cn_status cn_system_segments(cn_system S, int32_t *nseg, const int32_t **h_seg_ptr) {
  CN_API_BEGIN
  if (!S || !nseg || !h_seg_ptr) {
    set_error("cn_system_segments: bad arguments");
    return CN_BAD_ARG;
  }
  ensure_segments(S, nullptr);
  *nseg = S->nseg;
  *h_seg_ptr = S->seg_ptr.data();
  return CN_OK;
  CN_API_END
}

}  // extern "C"
