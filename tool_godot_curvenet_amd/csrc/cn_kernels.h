// cn_kernels.h -- device building blocks shared by the SpMV/SpMM, IC(0) and
// PCG kernels.
//
// Two row mappings cover all right-hand-side counts k:
//   * narrow (k <= 4): a group of 8 lanes per CSR row; lanes stride over the
//     row's entries (coalesced col/val reads), gather k values of x per entry
//     and combine with a butterfly over the group.  8 rows per wavefront.
//   * wide (k > 4): one wavefront per (row, column tile).  The row's col/val
//     stream is wave-uniform (scalar loads, readfirstlane), every lane owns
//     VEC consecutive columns, so each entry costs one coalesced 64*VEC-wide
//     gather of the n x k row-major block (SpmvMulti.lean:113 layout).
// Partial dot products are accumulated in FP64 per lane, reduced across the
// 4 waves of a block through LDS in a fixed order, and written as one row of
// a [block][k] partial buffer; a second kernel sums the partials in a fixed
// order.  Results are therefore deterministic from run to run.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cn {

constexpr int kBlock = 256;    // 4 wavefronts
constexpr int kWaves = 4;
constexpr int kGroup = 8;      // lanes per row in the narrow mapping
constexpr int kMaxPartBlocks = 1024;

__device__ __forceinline__ int32_t uniform(int32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its
// own 4 MB L2.  Row kernels therefore give each XCD one contiguous eighth of
// the rows, and the XCD's row slots (its waves or 8-lane groups) interleave
// over that eighth: the rows in flight on an XCD form one compact window
// that slides forward, so a row's neighbours -- the x rows it gathers, a
// few ring-lengths away in mesh order -- are still in that XCD's L2.
constexpr int kXcd = 8;
// width of the solver SpMM's row table (diagonal + 6 neighbours covers a
// regular triangle mesh)
constexpr int kEll = 7;
// Solver control block (device), read by every kernel of the PCG loop.
// CTL_ITERS counts the iterations in which some column was still active.
// CTL_PAUSE stops the iteration until the host's next residual replacement:
// set when a column's iterated residual has fallen delta^4 below the last
// true one (far past the delta^2 trigger), so a fast-converging solve does
// not run on in FP32 between two host checks.
// CTL_DFRESH (the fused iteration): the round's d holds nothing yet, so its
// first d += alpha p reads no d and the residual replacement adds no d; the
// first iteration's r.z reduction clears it (d is never zeroed in memory).
enum { CTL_ACTIVE = 0, CTL_BREAKDOWN = 1, CTL_ONE = 2, CTL_TRIGGER = 3, CTL_JUST = 4,
       CTL_ITERS = 5, CTL_PAUSE = 6, CTL_DFRESH = 7, CTL_N = 8 };
__device__ __forceinline__ bool ctl_idle(const int32_t *ctl) {
  return ctl[CTL_ACTIVE] == 0 || ctl[CTL_PAUSE] != 0;
}

struct RowIter {
  int32_t first, end, stride;
};
// slots: row slots per block (4 waves, or 32 groups); slot: this thread's
__device__ __forceinline__ RowIter row_iter(int32_t n0, int32_t n1, int32_t slots,
                                            int32_t slot) {
  const int32_t nb = gridDim.x;
  if (nb % kXcd != 0) return RowIter{n0 + (int32_t)blockIdx.x * slots + slot, n1, nb * slots};
  const int32_t x = blockIdx.x % kXcd, lb = blockIdx.x / kXcd, nbx = nb / kXcd;
  const int32_t chunk = (n1 - n0 + kXcd - 1) / kXcd;
  const int32_t cs = n0 + x * chunk;
  return RowIter{cs + lb * slots + slot, min(n1, cs + chunk), nbx * slots};
}

// row_iter with each XCD's share cut into `gens` stripes: the grid holds
// gens * (resident blocks) blocks, generation g of an XCD's blocks walks
// stripe g, and generations start in dispatch order.
__device__ __forceinline__ RowIter row_iter_gen(int32_t n0, int32_t n1, int32_t slots,
                                                int32_t slot, int32_t gens) {
  const int32_t nb = gridDim.x;
  if (gens <= 1 || nb % (kXcd * gens) != 0) return row_iter(n0, n1, slots, slot);
  const int32_t x = blockIdx.x % kXcd, lb = blockIdx.x / kXcd, nbx = nb / kXcd;
  const int32_t F = nbx / gens, gi = lb / F, lf = lb % F;
  const int32_t chunk = (n1 - n0 + kXcd - 1) / kXcd;
  const int32_t cs = n0 + x * chunk, ce = min(n1, cs + chunk);
  const int32_t stripe = (chunk + gens - 1) / gens;
  const int32_t ss = cs + gi * stripe;
  return RowIter{ss + lf * slots + slot, min(ce, ss + stripe), F * slots};
}

template <typename T, int VEC>
struct VecIO;
template <>
struct VecIO<float, 4> {
  static __device__ __forceinline__ void ld(const float *p, float (&v)[4]) {
    float4 t = *reinterpret_cast<const float4 *>(p);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
  }
  static __device__ __forceinline__ void st(float *p, const float (&v)[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
};
template <>
struct VecIO<float, 2> {
  static __device__ __forceinline__ void ld(const float *p, float (&v)[2]) {
    float2 t = *reinterpret_cast<const float2 *>(p);
    v[0] = t.x; v[1] = t.y;
  }
  static __device__ __forceinline__ void st(float *p, const float (&v)[2]) {
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  }
};
template <>
struct VecIO<double, 2> {
  static __device__ __forceinline__ void ld(const double *p, double (&v)[2]) {
    double2 t = *reinterpret_cast<const double2 *>(p);
    v[0] = t.x; v[1] = t.y;
  }
  static __device__ __forceinline__ void st(double *p, const double (&v)[2]) {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
  }
};
template <>
struct VecIO<double, 4> {
  static __device__ __forceinline__ void ld(const double *p, double (&v)[4]) {
    double2 t = *reinterpret_cast<const double2 *>(p);
    double2 u = *reinterpret_cast<const double2 *>(p + 2);
    v[0] = t.x; v[1] = t.y; v[2] = u.x; v[3] = u.y;
  }
  static __device__ __forceinline__ void st(double *p, const double (&v)[4]) {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
  }
};
template <typename T>
struct VecIO<T, 1> {
  static __device__ __forceinline__ void ld(const T *p, T (&v)[1]) { v[0] = *p; }
  static __device__ __forceinline__ void st(T *p, const T (&v)[1]) { *p = v[0]; }
};

// acc[m] += sum_{e in [a,b)} val[e] * X[col[e]*k + c0 + m]   (wide mapping)
// Entries are taken CH at a time; the tail of a chunk re-reads the row's last
// entry with weight 0, so every gather of a chunk is issued back to back
// (one round-trip of memory latency per chunk instead of one per entry).
template <typename T, int VEC, int CH = 8>
__device__ __forceinline__ void gather_wide(int32_t a, int32_t b,
                                            const int32_t *__restrict__ col,
                                            const T *__restrict__ val,
                                            const T *X, int64_t k,
                                            int64_t c0, bool valid,
                                            T (&acc)[VEC]) {
  for (int32_t e = a; e < b; e += CH) {
    int32_t j[CH];
    T v[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int32_t q = min(e + u, b - 1);
      j[u] = col[q];
      v[u] = (e + u < b) ? val[q] : T(0);
    }
    if (valid) {
      T x[CH][VEC];
#pragma unroll
      for (int u = 0; u < CH; ++u) VecIO<T, VEC>::ld(X + (int64_t)j[u] * k + c0, x[u]);
#pragma unroll
      for (int u = 0; u < CH; ++u)
#pragma unroll
        for (int m = 0; m < VEC; ++m) acc[m] = fma(v[u], x[u][m], acc[m]);
    }
  }
}

// narrow mapping: lane g of the row's 8-lane group
template <typename T, int K>
__device__ __forceinline__ void gather_narrow(int32_t a, int32_t b, int g,
                                              const int32_t *__restrict__ col,
                                              const T *__restrict__ val,
                                              const T *X,
                                              T (&acc)[K]) {
  for (int32_t e = a + g; e < b; e += kGroup) {
    int32_t j = col[e];
    T v = val[e];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = fma(v, X[(int64_t)j * K + c], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < K; ++c) {
    acc[c] += __shfl_xor(acc[c], 4, kGroup);
    acc[c] += __shfl_xor(acc[c], 2, kGroup);
    acc[c] += __shfl_xor(acc[c], 1, kGroup);
  }
}

// Lane packing of the wide mapping: LPR lanes (16, 32 or 64) own one row's
// column tile, so a wavefront walks 64 / LPR rows at once.  A moderate k
// (k / VEC of 9..32 lanes) would otherwise leave most of a 64-lane tile
// idle.  With LPR < 64 a row index is no longer wave-uniform and row data
// is read with vector loads; LPR = 64 is the one-row-per-wave form.
// LPR = 12: 5 rows a wavefront, lanes 60-63 idle.  The idle lanes take the
// wave's last row slot and a column past every k, so in a kernel that
// guards its column accesses with `valid` (or clamps them to k - VEC) they
// load in bounds and store nothing, and block_partials_wide leaves their
// partials out.
template <int LPR>
struct Pack {
  static_assert(LPR == 12 || LPR == 16 || LPR == 32 || LPR == 64, "lanes per row");
  static constexpr int G = 64 / LPR;        // rows per wavefront
  static constexpr int kSlots = kWaves * G;  // rows per block
  // row slot of this thread: the wave's row groups are consecutive slots
  static __device__ __forceinline__ int32_t slot() {
    const int g = (threadIdx.x & 63) / LPR;
    return (int32_t)((threadIdx.x >> 6) * G + (g < G ? g : G - 1));
  }
  static __device__ __forceinline__ int64_t col0(int vec) {
    const int lane = threadIdx.x & 63;
    if (G * LPR < 64 && lane >= G * LPR) return INT64_MAX / 2;
    return (int64_t)blockIdx.y * LPR * vec + (int64_t)(lane % LPR) * vec;
  }
  // wave-uniform only when the wave holds a single row
  static __device__ __forceinline__ int32_t u(int32_t x) {
    if constexpr (LPR == 64) return uniform(x);
    else return x;
  }
};

// Lanes per row for a row's column work of `lanes` lanes: of 64, 32 and 16,
// the width that leaves the fewest lanes of its column tiles idle (ties go
// to the widest, whose fewer tiles re-read each row's index data less often).
inline int pack_lpr(int64_t lanes) {
  int best = 64;
  int64_t idle = (lanes + 63) / 64 * 64 - lanes;
  for (int l : {32, 16}) {
    const int64_t w = (lanes + l - 1) / l * l - lanes;
    if (w < idle) { best = l; idle = w; }
  }
  return best;
}

// pack_lpr's width l, or 12 when that puts more of the lanes to work
// without more column tiles: a row of `lanes` lanes of work uses
// lanes / (tiles * l) of its lanes at width l and 5 lanes / (64 tiles) at
// width 12 (for 12 lanes, k = 48 in FP32: 0.75 at 16, 0.94 at 12).  More
// tiles re-read each row's index data and start more row chains: 24 lanes
// (k = 48 in FP64) in two 12-lane tiles ran the FP64 residual 15 % slower
// than in one 32-lane tile.
inline int pack_lpr12(int64_t lanes) {
  const int l = pack_lpr(lanes);
  const int64_t tl = (lanes + l - 1) / l, t12 = (lanes + 11) / 12;
  return t12 <= tl && 5 * tl * l > 64 * t12 ? 12 : l;
}

// Block reduction of per-lane FP64 partials for the wide mapping:
// part[(lane % LPR)*VEC+m] of each row group summed over the 4 waves and
// the wave's 64 / LPR row groups (waves outer, groups inner: a fixed order;
// for LPR = 12 the idle lanes 60-63 form no group).
// `sh` holds kWaves * NP * 64 * VEC doubles of shared memory.
// Blocks blockIdx.x < add_below add their sums to the row instead of
// storing them (the row holds an earlier launch's partials: launches in
// stream order accumulate into one set of rows, in launch order); pre, if
// given, holds the thread's values of that row, loaded earlier by the
// caller (the writing threads' columns c0 + m, NP x VEC of them).
template <int VEC, int NP, int LPR = 64>
__device__ __forceinline__ void block_partials_wide_in(double (&p)[NP][VEC],
                                                       double *__restrict__ out,
                                                       int64_t k, int64_t c0,
                                                       int64_t stride_np, double *sh,
                                                       int32_t add_below = 0,
                                                       const double *pre = nullptr) {
  constexpr int TW = 64 * VEC, G = 64 / LPR;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int m = 0; m < VEC; ++m) sh[(w * NP + q) * TW + lane * VEC + m] = p[q][m];
  __syncthreads();
  if (w == 0 && lane < LPR) {
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        int64_t c = c0 + m;
        const int o = q * TW + lane * VEC + m;
        double s = sh[o];
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww)
#pragma unroll
          for (int g = 0; g < G; ++g)
            if (ww + g > 0) s += sh[ww * NP * TW + g * LPR * VEC + o];
        if (c < k) {
          double *o_ = out + q * stride_np + (int64_t)blockIdx.x * k + c;
          *o_ = (int32_t)blockIdx.x < add_below ? (pre ? pre[q * VEC + m] : *o_) + s : s;
        }
      }
  }
}

template <int VEC, int NP, int LPR = 64>
__device__ __forceinline__ void block_partials_wide(double (&p)[NP][VEC],
                                                    double *__restrict__ out,
                                                    int64_t k, int64_t c0,
                                                    int64_t stride_np,
                                                    int32_t add_below = 0,
                                                    const double *pre = nullptr) {
  __shared__ double sh[kWaves * NP * 64 * VEC];
  block_partials_wide_in<VEC, NP, LPR>(p, out, k, c0, stride_np, sh, add_below, pre);
}

// narrow: lane g < K of every 8-lane group holds the partials of column g;
// the 32 groups of the block are summed in order.
template <int K, int NP>
__device__ __forceinline__ void block_partials_narrow(const double (&p)[NP],
                                                      double *__restrict__ out,
                                                      int64_t stride_np) {
  __shared__ double sh[NP][kBlock / kGroup][K];
  const int grp = threadIdx.x / kGroup, g = threadIdx.x % kGroup;
  if (g < K)
#pragma unroll
    for (int q = 0; q < NP; ++q) sh[q][grp][g] = p[q];
  __syncthreads();
  if (threadIdx.x < NP * K) {
    int q = threadIdx.x / K, c = threadIdx.x % K;
    double s = 0.0;
    for (int i = 0; i < kBlock / kGroup; ++i) s += sh[q][i][c];
    out[q * stride_np + (int64_t)blockIdx.x * K + c] = s;
  }
}

// Per-thread partial reduction (thread t holds K columns) for the
// thread-per-row kernels: fixed-order sum over the block's 256 threads.
template <int K, int NP>
__device__ __forceinline__ void block_partials_thread(const double (&p)[NP][K],
                                                      double *__restrict__ out,
                                                      int64_t stride_np) {
  // fixed order in two stages: runs of 16 threads, then the 16 runs (a
  // serial walk over 256 threads costs microseconds of LDS latency)
  constexpr int RUNS = kBlock / 16;
  __shared__ double sh[NP][K][kBlock];
  __shared__ double s2[NP][K][RUNS];
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int c = 0; c < K; ++c) sh[q][c][threadIdx.x] = p[q][c];
  __syncthreads();
  if (threadIdx.x < NP * K * RUNS) {
    const int qc = threadIdx.x / RUNS, m = threadIdx.x % RUNS;
    const int q = qc / K, c = qc % K;
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) t += sh[q][c][m * 16 + u];
    s2[q][c][m] = t;
  }
  __syncthreads();
  if (threadIdx.x < NP * K) {
    const int q = threadIdx.x / K, c = threadIdx.x % K;
    double t = 0.0;
    for (int m = 0; m < RUNS; ++m) t += s2[q][c][m];
    out[q * stride_np + (int64_t)blockIdx.x * K + c] = t;
  }
}

// register-array element by a lane-varying index (K <= 4)
template <typename T, int K>
__device__ __forceinline__ T pick(const T (&a)[K], int g) {
  T r = a[0];
#pragma unroll
  for (int c = 1; c < K; ++c) r = (g == c) ? a[c] : r;
  return r;
}

}  // namespace cn
