// cn_ddm.hip -- Direct Delta Mush on gfx950: the runtime skinning matvec and
// the bind-time weight pipeline.
//
// Runtime (DirectDeltaMush.lbsMatvec, DirectDeltaMush.lean:85-98; GPU form
// SlangCodegen/DirectDeltaMush.lean:33-87): pos[v] = sum over the vertex's
// influences (i, w) of w * (T_i (rest_v, 1))[0..2].  One thread per vertex
// walks its CSR influence row for a run of poses, so the rest position and
// the influence row are read once per run; the per-handle transforms are
// small and stay in the caches.
//
// Bind time (DirectDeltaMushBind.lean:9-27):
//   1. harvest: unit-perturbation harmonic solves, one column per handle
//      (Features/KnownUnknowns/SparseBindHarvest.lean:3-5 -- the multi-RHS
//      PCG replaces a dense solve).  cn_bind_rhs writes the right-hand sides
//      -A_fc e_h of a block of handles, the solver runs on them, and
//      cn_bind_weights scatters the solution and the unit rows of the
//      sample vertices into the vertex x handle weight matrix;
//   2. smoothWeights (:76-97): nu damped passes toward the neighbour mean;
//   3. sparsifyTopK (:107-136): the K largest |w| of each row by selection
//      sort with swaps, renormalised to sum 1.
//
// The file is compiled without floating-point contraction: the FP64 forms
// reproduce the spec's operation order bit for bit.  The FP32 skinning
// kernel uses explicit fma().
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_internal.h"
#include "cn_kernels.h"

namespace cn {

constexpr int kTopKMax = 64;  // kept pairs per row of the top-K kernel (one per lane)

// ---- runtime matvec --------------------------------------------------------
// rows 0-2 of a 4x4 row-major transform; `in` = false gives the zero matrix
template <typename T>
struct Xf;
template <>
struct Xf<double> {
  // applyTransform4x4 row r: ((T0 x + T1 y) + T2 z) + T3, then acc + w t
  static __device__ __forceinline__ void apply(const double *M, bool in, double x, double y,
                                               double z, double w, double (&acc)[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
      if (in) {
        a = *reinterpret_cast<const double2 *>(M + 4 * r);
        b = *reinterpret_cast<const double2 *>(M + 4 * r + 2);
      }
      const double t = a.x * x + a.y * y + b.x * z + b.y;
      acc[r] = acc[r] + w * t;
    }
  }
};
template <>
struct Xf<float> {
  static __device__ __forceinline__ void apply(const float *M, bool in, float x, float y,
                                               float z, float w, float (&acc)[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
      if (in) a = *reinterpret_cast<const float4 *>(M + 4 * r);
      const float t = fmaf(a.x, x, fmaf(a.y, y, fmaf(a.z, z, a.w)));
      acc[r] = fmaf(w, t, acc[r]);
    }
  }
};

// Skinning kernel.  A block owns a tile of 256 consecutive vertices (one per
// thread) and a run of poses.  The tile's vertices share most of their
// handles, so the block first collects the distinct in-range handle indices
// of its influences (an LDS hash table, compacted in slot order) and, for
// every pose, stages those handles' transform rows in LDS once; each vertex
// then reads its matrices from LDS instead of gathering them from L2.  The
// arithmetic per vertex is the same in either path (influences in CSR order,
// an out-of-range handle as the zero matrix), so a tile that does not fit --
// more than kLbsRow influences on a vertex, more than half the hash table in
// influences, or more than LbsSmem<T>::kCap distinct handles -- falls back
// to direct loads with identical results.  The blend reads 8 x 48 bytes of
// LDS a vertex and pose, which bounds the kernel (LDS, not HBM).
constexpr int kLbsTile = 256;
constexpr int kLbsHashBits = 12;
constexpr int kLbsHash = 1 << kLbsHashBits;
constexpr int kLbsRow = 8;

__device__ __forceinline__ uint32_t lbs_hash(int32_t i) {
  return ((uint32_t)i * 2654435761u) >> (32 - kLbsHashBits);
}

template <typename T>
__device__ __forceinline__ void lbs_store(T *out, int64_t nv, int32_t p, int32_t v,
                                          const T (&acc)[3]) {
  T *o = out + ((int64_t)p * nv + v) * 3;
  o[0] = acc[0];
  o[1] = acc[1];
  o[2] = acc[2];
}

// four consecutive values of a transform row (16-byte aligned)
template <typename T>
__device__ __forceinline__ void ld4(const T *p, T (&v)[4]);
template <>
__device__ __forceinline__ void ld4<float>(const float *p, float (&v)[4]) {
  const float4 a = *reinterpret_cast<const float4 *>(p);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
template <>
__device__ __forceinline__ void ld4<double>(const double *p, double (&v)[4]) {
  const double2 a = *reinterpret_cast<const double2 *>(p);
  const double2 b = *reinterpret_cast<const double2 *>(p + 2);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <typename T>
__device__ __forceinline__ void st4(T *p, const T (&v)[4]);
template <>
__device__ __forceinline__ void st4<float>(float *p, const float (&v)[4]) {
  *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void st4<double>(double *p, const double (&v)[4]) {
  *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
  *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
}

// LDS of the tile kernel: the hash table (keys, then compacted ids) during
// set-up, the staged transform rows afterwards
template <typename T>
struct LbsSmem {
  // distinct handles a tile stages (12 values each: 24 KB)
  static constexpr int kCap = sizeof(T) == 4 ? 512 : 256;
  static constexpr size_t kHash = kLbsHash * (sizeof(int32_t) + sizeof(int16_t));
  static constexpr size_t kRows = kCap * 12 * sizeof(T);
  // waves per SIMD the register budget is held to
  static constexpr int kWavesPerEu = sizeof(T) == 4 ? 4 : 2;
  static constexpr size_t kBytes = kHash > kRows ? kHash : kRows;
};

template <typename T>
__global__ void __launch_bounds__(kLbsTile, LbsSmem<T>::kWavesPerEu)
k_lbs(int32_t nv, int32_t nt, int32_t P, int32_t pch, const T *__restrict__ rest,
      const T *__restrict__ Tm, const int32_t *__restrict__ start,
      const int32_t *__restrict__ idx, const T *__restrict__ w, T *__restrict__ out) {
  __shared__ __align__(16) unsigned char smem[LbsSmem<T>::kBytes];
  constexpr int kLbsCap = LbsSmem<T>::kCap;
  __shared__ int32_t hlist[kLbsCap];
  __shared__ int32_t wsum[kLbsTile / 64];
  int32_t *hkey = reinterpret_cast<int32_t *>(smem);
  int16_t *hid = reinterpret_cast<int16_t *>(smem + kLbsHash * sizeof(int32_t));
  T *Ts = reinterpret_cast<T *>(smem);
  const int t = threadIdx.x;
  const int32_t v0 = blockIdx.x * kLbsTile, v = v0 + t;
  const bool live = v < nv;
  const int32_t p0 = blockIdx.y * pch, p1 = min(P, p0 + pch);
  const int32_t vend = min(v0 + kLbsTile, nv);
  const int32_t a = live ? start[v] : 0, len = live ? start[v + 1] - a : 0;
  T x = T(0), y = T(0), z = T(0);
  if (live) {
    x = rest[3 * (int64_t)v];
    y = rest[3 * (int64_t)v + 1];
    z = rest[3 * (int64_t)v + 2];
  }
  const int32_t tot = start[vend] - start[v0];
  bool fb = tot > kLbsHash / 2;
  fb = __syncthreads_or(fb || len > kLbsRow) != 0;
  if (fb) {  // direct loads
    if (!live) return;
    for (int32_t p = p0; p < p1; ++p) {
      const T *Tp = Tm + (int64_t)p * nt * 16;
      T acc[3] = {T(0), T(0), T(0)};
      for (int32_t q = a; q < a + len; ++q) {
        const int32_t i = idx[q];
        const bool in = i >= 0 && i < nt;
        Xf<T>::apply(Tp + 16 * (int64_t)(in ? i : 0), in, x, y, z, w[q], acc);
      }
      lbs_store(out, nv, p, v, acc);
    }
    return;
  }
  // 1. distinct in-range handles of the tile (the table holds at most half
  //    its slots in keys, so linear probing always finds a place)
  for (int j = t; j < kLbsHash; j += kLbsTile) hkey[j] = -1;
  __syncthreads();
  int32_t li[kLbsRow];
  T lw[kLbsRow];
#pragma unroll
  for (int u = 0; u < kLbsRow; ++u) {
    li[u] = -1;
    lw[u] = T(0);
    if (u < len) {
      li[u] = idx[a + u];
      lw[u] = w[a + u];
      if (li[u] >= 0 && li[u] < nt) {
        uint32_t h = lbs_hash(li[u]);
        while (true) {
          const int32_t prev = atomicCAS(&hkey[h], -1, li[u]);
          if (prev == -1 || prev == li[u]) break;
          h = (h + 1) & (kLbsHash - 1);
        }
      } else {
        li[u] = -1;
      }
    }
  }
  __syncthreads();
  // 2. compact the occupied slots in slot order: thread t owns 16 slots
  constexpr int kPer = kLbsHash / kLbsTile;
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) cnt += hkey[t * kPer + j] >= 0;
  const int lane = t & 63, wv = t >> 6;
  int incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  int base = incl - cnt, nU = 0;
  for (int j = 0; j < kLbsTile / 64; ++j) {
    if (j < wv) base += wsum[j];
    nU += wsum[j];
  }
  if (nU > kLbsCap) {  // uniform: every thread read the same totals
    if (!live) return;
    for (int32_t p = p0; p < p1; ++p) {
      const T *Tp = Tm + (int64_t)p * nt * 16;
      T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int u = 0; u < kLbsRow; ++u)
        if (u < len) {
          const bool in = li[u] >= 0;
          Xf<T>::apply(Tp + 16 * (int64_t)(in ? li[u] : 0), in, x, y, z, lw[u], acc);
        }
      lbs_store(out, nv, p, v, acc);
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int h = t * kPer + j;
    if (hkey[h] >= 0) {
      hid[h] = (int16_t)base;
      hlist[base] = hkey[h];
      ++base;
    }
  }
  __syncthreads();
  // 3. each influence's LDS slot (-1: out of range, the zero matrix)
  int32_t ls[kLbsRow];
#pragma unroll
  for (int u = 0; u < kLbsRow; ++u) {
    ls[u] = -1;
    if (li[u] >= 0) {
      uint32_t h = lbs_hash(li[u]);
      while (hkey[h] != li[u]) h = (h + 1) & (kLbsHash - 1);
      ls[u] = hid[h];
    }
  }
  __syncthreads();  // the hash table's LDS becomes the staged rows
  // 4. poses: the rows 0-2 of the tile's transforms for pose p + 1 are
  //    fetched into registers while pose p is blended from LDS
  constexpr int kStage = (kLbsCap * 3 + kLbsTile - 1) / kLbsTile;
  T sreg[kStage][4];
  const int nrow = nU * 3;
  auto fetch = [&](int32_t p) {
    const T *Tp = Tm + (int64_t)p * nt * 16;
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int j = t + q * kLbsTile;
      if (j < nrow) {
        const int hh = hlist[j / 3], r = j - 3 * (j / 3);
        ld4<T>(Tp + 16 * (int64_t)hh + 4 * r, sreg[q]);
      }
    }
  };
  if (p0 < p1) fetch(p0);
  for (int32_t p = p0; p < p1; ++p) {
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int j = t + q * kLbsTile;
      if (j < nrow) st4<T>(Ts + 4 * j, sreg[q]);
    }
    __syncthreads();
    if (p + 1 < p1) fetch(p + 1);
    if (live) {
      T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int u = 0; u < kLbsRow; ++u)
        if (u < len) {
          const bool in = ls[u] >= 0;
          Xf<T>::apply(Ts + 12 * (in ? ls[u] : 0), in, x, y, z, lw[u], acc);
        }
      lbs_store(out, nv, p, v, acc);
    }
    __syncthreads();
  }
}

template <typename T>
static void launch_lbs(int32_t nv, int32_t nt, int32_t P, const T *rest, const T *Tm,
                       const int32_t *start, const int32_t *idx, const T *w, T *out,
                       hipStream_t s) {
  const unsigned bx = (unsigned)grid_for(nv, kLbsTile);
  // enough blocks to fill the device (~2048); otherwise the fewest runs
  // per tile, each paying the tile's handle collection once
  int64_t ny = std::max<int64_t>(1, std::min<int64_t>(P, (2048 + bx - 1) / bx));
  const int32_t pch = (int32_t)((P + ny - 1) / ny);
  ny = (P + pch - 1) / pch;
  k_lbs<T><<<dim3(bx, (unsigned)ny, 1), kLbsTile, 0, s>>>(nv, nt, P, pch, rest, Tm, start,
                                                          idx, w, out);
}

// ---- smoothWeights: one pass src -> dst ------------------------------------
// LPR lanes per row over VEC-column tiles; the neighbour sum runs from 0 in
// adjacency order (out-of-range neighbours add 0), is divided by the count
// when there are neighbours, and is blended as (1 - omega) w + omega avg.
template <int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
k_smooth(int32_t n, int64_t nh, const int32_t *__restrict__ adj_ptr,
         const int32_t *__restrict__ adj, const double *__restrict__ src,
         double *__restrict__ dst, double omega) {
  using P = Pack<LPR>;
  const int64_t c0 = P::col0(VEC);
  const bool valid = c0 < nh;
  const int64_t cl = valid ? c0 : nh - VEC;
  const double keep = 1.0 - omega;
  const RowIter ri = row_iter(0, n, P::kSlots, P::slot());
  for (int32_t r = ri.first; r < ri.end; r += ri.stride) {
    const int32_t a = adj_ptr[r], b = adj_ptr[r + 1];
    double nb[VEC], own[VEC];
#pragma unroll
    for (int m = 0; m < VEC; ++m) nb[m] = 0.0;
    for (int32_t q = a; q < b; ++q) {
      const int32_t u = adj[q];
      const bool in = u >= 0 && u < n;
      double x[VEC];
      VecIO<double, VEC>::ld(src + (int64_t)(in ? u : r) * nh + cl, x);
#pragma unroll
      for (int m = 0; m < VEC; ++m) nb[m] = nb[m] + (in ? x[m] : 0.0);
    }
    if (b - a > 0) {
      const double cnt = (double)(b - a);
#pragma unroll
      for (int m = 0; m < VEC; ++m) nb[m] = nb[m] / cnt;
    }
    VecIO<double, VEC>::ld(src + (int64_t)r * nh + cl, own);
#pragma unroll
    for (int m = 0; m < VEC; ++m) own[m] = keep * own[m] + omega * nb[m];
    if (valid) VecIO<double, VEC>::st(dst + (int64_t)r * nh + c0, own);
  }
}

static void launch_smooth(int32_t n, int64_t nh, const int32_t *adj_ptr, const int32_t *adj,
                          const double *src, double *dst, double omega, hipStream_t s) {
  const int v = (nh % 2 == 0) ? 2 : 1;
  const int lpr = pack_lpr((nh + v - 1) / v);
  const int slots = kWaves * (64 / lpr);
  int64_t g = (n + slots - 1) / slots;
  g = std::max<int64_t>(1, std::min<int64_t>(g, 2048));
  if (g >= 64) g -= g % 8;
  const dim3 grid((unsigned)g, (unsigned)((nh + (int64_t)lpr * v - 1) / ((int64_t)lpr * v)), 1);
#define S_(VV, LL) k_smooth<VV, LL><<<grid, kBlock, 0, s>>>(n, nh, adj_ptr, adj, src, dst, omega)
#define S_L(VV)                                                                 \
  do {                                                                          \
    if (lpr == 16) S_(VV, 16);                                                  \
    else if (lpr == 32) S_(VV, 32);                                             \
    else S_(VV, 64);                                                            \
  } while (0)
  if (v == 2) S_L(2); else S_L(1);
#undef S_L
#undef S_
}

// ---- sparsifyTopK ------------------------------------------------------------
// One wavefront per row.  The spec's selection sort swaps the selected pair
// into place, which moves the pair it displaces to the selected position;
// that order decides ties, so the kernel keeps the displaced positions as a
// list (entry j in lane j: position, original index) instead of a copy of the
// row.  The pair at position q >= k is the list's entry for q, else pair q.
// Step k picks the first position of the largest |w| among positions >= k
// (strict >: NaN never replaces, a NaN at position k stays).
__device__ __forceinline__ int32_t topk_content(int32_t q, int32_t nd, int32_t dpos,
                                                int32_t dval) {
  int32_t e = q;
  for (int32_t j = 0; j < nd; ++j) {
    const int32_t pj = __builtin_amdgcn_readlane(dpos, j);
    const int32_t ej = __builtin_amdgcn_readlane(dval, j);
    if (pj == q) e = ej;
  }
  return e;
}

__global__ void __launch_bounds__(kBlock)
k_topk(int32_t n, int32_t nh, const double *__restrict__ W, int32_t take,
       int32_t *__restrict__ oidx, double *__restrict__ ow) {
  const int lane = threadIdx.x & 63;
  const int32_t nw = gridDim.x * kWaves;
  for (int32_t row = blockIdx.x * kWaves + (threadIdx.x >> 6); row < n; row += nw) {
    const double *Wr = W + (int64_t)row * nh;
    int32_t dpos = -1, dval = 0, nd = 0;
    int32_t my_idx = 0;
    double my_w = 0.0;
    for (int32_t k = 0; k < take; ++k) {
      const int32_t ek = topk_content(k, nd, dpos, dval);
      const double wk = Wr[ek];
      int32_t bpos = k;
      if (wk == wk) {
        double bkey = fabs(wk);
        for (int32_t base = k + 1; base < nh; base += 64) {
          const int32_t q = min(base + lane, nh - 1);
          const int32_t e = topk_content(q, nd, dpos, dval);
          const double av = fabs(Wr[e]);
          if (base + lane < nh && av > bkey) {
            bkey = av;
            bpos = q;
          }
        }
        // largest key, ties to the smaller position
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const double ok = __shfl_xor(bkey, off);
          const int32_t op = __shfl_xor(bpos, off);
          if (ok > bkey || (ok == bkey && op < bpos)) {
            bkey = ok;
            bpos = op;
          }
        }
        bpos = __builtin_amdgcn_readfirstlane(bpos);
      }
      const int32_t eb = topk_content(bpos, nd, dpos, dval);
      if (lane == k) {
        my_idx = eb;
        my_w = Wr[eb];
      }
      // swap: position k leaves the scan range, position bpos takes ek
      if (lane < nd && dpos == k) dpos = -1;
      if (bpos != k) {
        const bool found = __ballot(lane < nd && dpos == bpos) != 0;
        if (found) {
          if (lane < nd && dpos == bpos) dval = ek;
        } else {
          if (lane == nd) {
            dpos = bpos;
            dval = ek;
          }
          ++nd;
        }
      }
    }
    double sum = 0.0;
    for (int32_t j = 0; j < take; ++j) sum = sum + __shfl(my_w, j);
    if (lane < take) {
      oidx[(int64_t)row * take + lane] = my_idx;
      ow[(int64_t)row * take + lane] = sum != 0.0 ? my_w / sum : my_w;
    }
  }
}

// ---- bind harvest --------------------------------------------------------------
// rhs[r][c] = -A_fc[r][h0 + c] (-0 where the row has no such entry): the
// harmonic right-hand side of the unit constrained vector e_{h0+c}
__global__ void k_bind_rhs(int32_t n_free, int32_t nb, int32_t h0,
                           const int32_t *__restrict__ fc_ptr, const int32_t *__restrict__ fc_col,
                           const double *__restrict__ fc_val, double *__restrict__ rhs) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n_free * nb) return;
  const int32_t r = (int32_t)(id / nb), c = (int32_t)(id - (int64_t)r * nb);
  double v = 0.0;
  for (int32_t q = fc_ptr[r]; q < fc_ptr[r + 1]; ++q)
    if (fc_col[q] == h0 + c) v = fc_val[q];
  rhs[id] = -v;
}

// W[v][h0 + c]: the free row's solution, or 1 / 0 for a sample vertex
__global__ void k_bind_weights(int32_t nv, int32_t nb, int32_t h0,
                               const int32_t *__restrict__ free_of,
                               const int32_t *__restrict__ con_of, const double *__restrict__ X,
                               double *__restrict__ W, int64_t ldw) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)nv * nb) return;
  const int32_t v = (int32_t)(id / nb), c = (int32_t)(id - (int64_t)v * nb);
  const int32_t f = free_of[v];
  W[(int64_t)v * ldw + h0 + c] =
      f >= 0 ? X[(int64_t)f * nb + c] : (con_of[v] == h0 + c ? 1.0 : 0.0);
}

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_lbs_matvec_f64(int32_t nv, int32_t nt, int32_t P, const double *d_rest,
                            const double *d_T, const int32_t *d_start, const int32_t *d_idx,
                            const double *d_w, double *d_out, void *stream) {
  CN_API_BEGIN
  if (nv < 0 || nt < 0 || P < 0 ||
      (nv > 0 && P > 0 && (!d_rest || !d_start || !d_out || (nt > 0 && !d_T)))) {
    set_error("cn_lbs_matvec_f64: bad arguments");
    return CN_BAD_ARG;
  }
  if (nv > 0 && P > 0)
    launch_lbs<double>(nv, nt, P, d_rest, d_T, d_start, d_idx, d_w, d_out, as_stream(stream));
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_lbs_matvec_f32(int32_t nv, int32_t nt, int32_t P, const float *d_rest,
                            const float *d_T, const int32_t *d_start, const int32_t *d_idx,
                            const float *d_w, float *d_out, void *stream) {
  CN_API_BEGIN
  if (nv < 0 || nt < 0 || P < 0 ||
      (nv > 0 && P > 0 && (!d_rest || !d_start || !d_out || (nt > 0 && !d_T)))) {
    set_error("cn_lbs_matvec_f32: bad arguments");
    return CN_BAD_ARG;
  }
  if (nv > 0 && P > 0)
    launch_lbs<float>(nv, nt, P, d_rest, d_T, d_start, d_idx, d_w, d_out, as_stream(stream));
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_smooth_weights_f64(int32_t n, int32_t nh, const int32_t *d_adj_ptr,
                                const int32_t *d_adj, double *d_W, double *d_tmp, int32_t nu,
                                double omega, void *stream) {
  CN_API_BEGIN
  if (n < 0 || nh < 0 || nu < 0 ||
      (n > 0 && nh > 0 && nu > 0 && (!d_adj_ptr || !d_W || !d_tmp || d_tmp == d_W))) {
    set_error("cn_smooth_weights_f64: bad arguments");
    return CN_BAD_ARG;
  }
  if (n == 0 || nh == 0 || nu == 0) return CN_OK;
  hipStream_t s = as_stream(stream);
  double *a = d_W, *b = d_tmp;
  for (int32_t pass = 0; pass < nu; ++pass) {
    launch_smooth(n, nh, d_adj_ptr, d_adj, a, b, omega, s);
    std::swap(a, b);
  }
  if (a != d_W)
    CN_HIP(hipMemcpyAsync(d_W, a, sizeof(double) * (size_t)n * nh, hipMemcpyDeviceToDevice, s));
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_sparsify_topk_f64(int32_t n, int32_t nh, const double *d_W, int32_t K,
                               int32_t *d_idx, double *d_w, void *stream) {
  CN_API_BEGIN
  if (n < 0 || nh < 0 || K < 0) {
    set_error("cn_sparsify_topk_f64: bad arguments");
    return CN_BAD_ARG;
  }
  const int32_t take = std::min(K, nh);
  if (take > kTopKMax) {
    set_error("cn_sparsify_topk_f64: min(K, nh) = %d exceeds %d kept pairs per row", take,
              kTopKMax);
    return CN_BAD_ARG;
  }
  if (n == 0 || take == 0) return CN_OK;
  if (!d_W || !d_idx || !d_w) {
    set_error("cn_sparsify_topk_f64: null buffer");
    return CN_BAD_ARG;
  }
  const int64_t g = std::min<int64_t>((n + kWaves - 1) / kWaves, 4096);
  k_topk<<<(unsigned)g, kBlock, 0, as_stream(stream)>>>(n, nh, d_W, take, d_idx, d_w);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_bind_rhs(cn_system S, int32_t h0, int32_t nb, double *d_rhs, void *stream) {
  CN_API_BEGIN
  if (!S || S->nseg != 1 || nb < 0 || h0 < 0 || (int64_t)h0 + nb > S->n_con ||
      (nb > 0 && !d_rhs)) {
    set_error("cn_bind_rhs: bad arguments (handles [%d, %d) of %d)", h0, h0 + nb,
              S ? S->n_con : 0);
    return CN_BAD_ARG;
  }
  const int64_t tot = (int64_t)S->n_free * nb;
  if (tot > 0)
    k_bind_rhs<<<grid_for(tot, 256), 256, 0, as_stream(stream)>>>(
        S->n_free, nb, h0, S->fc_ptr.get(), S->fc_col.get(), S->fc_val.get(), d_rhs);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_bind_weights(cn_system S, int32_t h0, int32_t nb, const double *d_x, double *d_W,
                          int64_t ldw, void *stream) {
  CN_API_BEGIN
  if (!S || S->nseg != 1 || nb < 0 || h0 < 0 || (int64_t)h0 + nb > ldw ||
      (nb > 0 && (!d_W || (S->n_free > 0 && !d_x)))) {
    set_error("cn_bind_weights: bad arguments");
    return CN_BAD_ARG;
  }
  const int64_t tot = (int64_t)S->nv * nb;
  if (tot > 0)
    k_bind_weights<<<grid_for(tot, 256), 256, 0, as_stream(stream)>>>(
        S->nv, nb, h0, S->free_of.get(), S->con_of.get(), d_x, d_W, ldw);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

}  // extern "C"
