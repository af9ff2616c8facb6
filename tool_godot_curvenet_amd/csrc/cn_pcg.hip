// cn_pcg.hip -- preconditioned conjugate gradients on gfx950.
//
// SparseLinAlg.cg / cgWithGuess (SparseLinAlg.lean:104-163) for k right-hand
// sides in lockstep (per-column alpha/beta, as the reference's *Multi
// kernels, SlangCodegen/AxpyMulti.lean:68-69).  All vectors are n x k
// row-major in the system's vertex order; only the IC(0) passes walk rows in
// level (colour) order through `perm`.
//
// One iteration (C = IC(0) levels), all scalars on the device:
//   SpMM   q = A p              (+ partial p.q)   -> alpha = rz / p.q
//   C fwd  r -= alpha q; y = L^-1 r  (+ partial r.r, y.y)
//                                                 -> rz' = y.y (= r.M^-1 r),
//                                                    beta, stop test
//   C bwd  z = L^-T y; d += alpha p; p = z + beta p
// so the vectors move 12 times per iteration (SpMM 2, fwd 4, bwd 6) and
// there are 2C + 3 launches.  The host reads a 4-int control block every
// `check_every` iterations; kernels return at once when no column is active.
//
// CN_PREC_MIXED keeps FP32 vectors (r, y/z, p, q, d) and FP32 matrix/factor
// values, with FP64 dots and scalars, plus "reliable updates": when a
// column's iterated residual has dropped 10x since the last update, the host
// folds d into the FP64 solution x, recomputes r = b - A x in FP64 and
// continues CG with that residual and the same search direction.  The
// Krylov space is kept (no restart), so the iteration count stays near the
// FP64 count, while the stop test and x are FP64.  CN_PREC_FP64 runs the
// same loop in FP64.  CN_PREC_FP64_PLAIN is the unfused baseline: one launch
// per BLAS-1 / SpMV / preconditioner step with alpha, the stop test and beta
// computed on the host (three device->host round trips per iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "cn_internal.h"
#include "cn_kernels.h"
#include "cn_prof.h"

namespace cn {

// per-column scalar slots (st[slot * k + c])
enum {
  ST_ALPHA = 0, ST_BETA, ST_GAMMA, ST_RZ, ST_TOL2, ST_RREF, ST_S, ST_INVS,
  ST_BB, ST_RR, ST_TMP, ST_N
};
// control block
// CTL_JUST counts columns stopped (or paused) by this iteration's
// k_reduce_beta: the backward pass that follows still has to apply their
// last d += alpha p.
__device__ __forceinline__ bool bwd_idle(const int32_t *ctl) {
  return ctl_idle(ctl) && ctl[CTL_JUST] == 0;
}

// The CG scalars of a batch (cn_system_batch) are kept per (segment,
// column): element (row, c) of an n x k block uses virtual column
// seg(row) * k + c of kk = nseg * k.  A single system has of_row = nullptr
// and kk = k.
struct SegMap {
  const int32_t *of_row;
  int64_t kk;
  __device__ __forceinline__ int64_t vc(int64_t row, int64_t k, int64_t c) const {
    return of_row ? (int64_t)of_row[row] * k + c : c;
  }
};

// ================================================================ kernels ==
// rows of an XCD's contiguous eighth, interleaved over its row slots
#define NARROW_LOOP(row, n0, n1)                                              \
  const RowIter ri_ = row_iter((n0), (n1), kBlock / kGroup, threadIdx.x / kGroup); \
  for (int64_t row = ri_.first; row < ri_.end; row += ri_.stride)
#define WIDE_LOOP(row, n0, n1)                                                \
  const RowIter ri_ = row_iter((n0), (n1), kWaves, threadIdx.x >> 6);         \
  for (int32_t row = ri_.first; row < ri_.end; row += ri_.stride)
#define WIDE_COLS(VEC)                                                        \
  const int lane = threadIdx.x & 63;                                          \
  const int64_t c0 = (int64_t)blockIdx.y * 64 * (VEC) + lane * (VEC);         \
  const bool valid = c0 < k;
// packed wide mapping: LPR lanes per row, 64 / LPR rows per wavefront
#define WIDE_LOOP_P(row, n0, n1, LPR)                                         \
  const RowIter ri_ = row_iter((n0), (n1), Pack<LPR>::kSlots, Pack<LPR>::slot()); \
  for (int32_t row = ri_.first; row < ri_.end; row += ri_.stride)
#define WIDE_COLS_P(VEC, LPR)                                                 \
  const int64_t c0 = Pack<LPR>::col0(VEC);                                    \
  const bool valid = c0 < k;

// ---- SpMM with partial p.q -------------------------------------------
// Narrow form (k <= 4): one thread per row of the row table, its K columns
// in registers; slot 0 is the row itself, so its gather also serves p.q.
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
kn_spmm_dot(int32_t n, const int32_t *__restrict__ ecol, const T *__restrict__ eval,
            const int32_t *__restrict__ ovp, const int32_t *__restrict__ ovc,
            const T *__restrict__ ovv, const T *__restrict__ P, T *__restrict__ Q,
            double *__restrict__ part, const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  double pd[1][K] = {};
  const RowIter it = row_iter(0, n, kBlock, threadIdx.x);
  for (int32_t r = it.first; r < it.end; r += it.stride) {
    const int64_t e0 = (int64_t)r * kEll;
    int32_t j[kEll];
    T w[kEll];
#pragma unroll
    for (int u = 0; u < kEll; ++u) {
      j[u] = ecol[e0 + u];
      w[u] = eval[e0 + u];
    }
    T x[kEll][K];
#pragma unroll
    for (int u = 0; u < kEll; ++u)
#pragma unroll
      for (int c = 0; c < K; ++c) x[u][c] = P[(int64_t)j[u] * K + c];
    T acc[K] = {};
#pragma unroll
    for (int u = 0; u < kEll; ++u)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] = fma(w[u], x[u][c], acc[c]);
    if (ovp) {
      const int32_t e1 = ovp[r + 1];
      for (int32_t e = ovp[r]; e < e1; ++e) {
        const T v = ovv[e];
        const int64_t q = (int64_t)ovc[e] * K;
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] = fma(v, P[q + c], acc[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      Q[(int64_t)r * K + c] = acc[c];
      pd[0][c] += (double)x[0][c] * (double)acc[c];
    }
  }
  block_partials_thread<K, 1>(pd, part, 0);
}

// Wide form over the row table: each wave walks its rows with the next
// row's kEll columns and weights (scalar loads) fetched while the current
// row's gathers are in flight, so a row costs one memory round trip.  Slot 0
// of every row is its own column, so the gather of P[r] also serves p.q.
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_spmm_dot(int32_t n, const int32_t *__restrict__ ecol, const T *__restrict__ eval,
            const int32_t *__restrict__ ovp, const int32_t *__restrict__ ovc,
            const T *__restrict__ ovv, int64_t k, const T *__restrict__ P,
            T *__restrict__ Q, double *__restrict__ part,
            const int32_t *__restrict__ ctl, int32_t gens) {
  if (ctl_idle(ctl)) return;
  WIDE_COLS_P(VEC, LPR)
  double pd[1][VEC] = {};
  const RowIter it = row_iter_gen(0, n, Pack<LPR>::kSlots, Pack<LPR>::slot(), gens);
  const int32_t end = it.end, stride = uniform(it.stride);
  int32_t r = Pack<LPR>::u(it.first);
  int32_t j[kEll], oa = 0, ob = 0;
  T v[kEll];
  if (r < end) {
#pragma unroll
    for (int u = 0; u < kEll; ++u) {
      j[u] = ecol[(int64_t)r * kEll + u];
      v[u] = eval[(int64_t)r * kEll + u];
    }
    if (ovp) { oa = ovp[r]; ob = ovp[r + 1]; }
  }
  // Loads and their uses are unconditional (lanes past k read the last VEC
  // columns); only the store depends on `valid`.  A load consumed under a
  // branch makes the compiler wait for the previous row's store at the top
  // of every iteration.
  const int64_t cl = valid ? c0 : k - VEC;
  while (r < end) {
    T x[kEll][VEC];
#pragma unroll
    for (int u = 0; u < kEll; ++u) VecIO<T, VEC>::ld(P + (int64_t)j[u] * k + cl, x[u]);
    const int32_t rn = r + stride;
    int32_t jn[kEll], oan = 0, obn = 0;
    T vn[kEll];
    if (rn < end) {
#pragma unroll
      for (int u = 0; u < kEll; ++u) {
        jn[u] = ecol[(int64_t)rn * kEll + u];
        vn[u] = eval[(int64_t)rn * kEll + u];
      }
      if (ovp) { oan = ovp[rn]; obn = ovp[rn + 1]; }
    }
    T acc[VEC] = {};
#pragma unroll
    for (int u = 0; u < kEll; ++u)
#pragma unroll
      for (int m = 0; m < VEC; ++m) acc[m] = fma(v[u], x[u][m], acc[m]);
    if (oa < ob) gather_wide<T, VEC, 2>(oa, ob, ovc, ovv, P, k, cl, true, acc);
    // lanes past k accumulate too; their partials are never written
#pragma unroll
    for (int m = 0; m < VEC; ++m) pd[0][m] += (double)x[0][m] * (double)acc[m];
    if (valid) VecIO<T, VEC>::st(Q + (int64_t)r * k + c0, acc);
#pragma unroll
    for (int u = 0; u < kEll; ++u) { j[u] = jn[u]; v[u] = vn[u]; }
    oa = oan;
    ob = obn;
    r = rn;
  }
  block_partials_wide<VEC, 1, LPR>(pd, part, k, c0, 0);
}

// ---- LDS-staged SpMM (wide k) -------------------------------------------
// A block owns one contiguous range of rows and one 256-byte column tile
// (64 FP32 / 32 FP64 columns, 16 lanes a row, 32 rows a step).  It streams
// the tile's rows of P through a ring of 12 chunks of 32 rows in LDS, so a
// row is read from memory once per launch however many rows gather it: the
// ring holds the 5 chunks behind and the 5 ahead of the chunk being
// computed (every neighbour within 160 rows of its row), and the loads of
// the next D chunks are in flight in registers.  Neighbours outside that
// window ("far" rows, a few per range on banded meshes) are read into a
// small LDS cache once at the start.
// The row table has 8 slots a row (diagonal first, then the row's entries
// in column order, padding with the last column and weight 0).  A slot
// holds the LDS row the neighbour sits in while its row is computed: ring
// slot (chunk mod 12) * 32 + row in chunk, or 12 * 32 + f for far row f.
// That is fixed by the range, so the kernel does no index arithmetic per
// entry.  The FMAs run in the order of the 7-slot table plus its overflow,
// so Q is the same as kw_spmm_dot's.
// Index data (LDS rows, weights) is staged through LDS with the same
// pipeline, so the only vector-memory operations of a step are the
// prefetches and the Q store and no wait drains the pipeline.
constexpr int kLdsRows = 32;    // rows per chunk (512 threads, 16 lanes a row)
constexpr int kLdsSlots = 12;   // ring of chunks
constexpr int kLdsBack = 5;     // chunks kept behind the computed one
constexpr int kLdsAhead = 5;    // chunks landed ahead of it
constexpr int kLdsFar = 160;    // far rows a range may cache
constexpr int kWinEll = 8;      // slots of the window row table
constexpr int kWinEnt = kLdsRows * kWinEll;  // index entries of a chunk
static_assert(kLdsSlots >= kLdsBack + kLdsAhead + 2, "ring too small for the window");

// bytes of a chunk's packed index block: kWinEnt int32 columns, then
// kWinEnt weights
template <typename T>
constexpr int win_block_bytes() { return kWinEnt * (4 + (int)sizeof(T)); }

// Ranges alternate direction: even ranges walk their rows up, odd ones
// down, so two neighbouring ranges reach their common boundary at the same
// time (both at the start or both at the end) and the rows either side
// read as halo are fetched once.  A range's walking position p is row
// rs + p (up) or re - 1 - p (down); the table is laid out by (range,
// position).  Blocks are mapped to (range, tile) pairs in contiguous runs
// per XCD (workgroups go round-robin over the 8 XCDs), so the tiles of one
// range share an XCD's L2 for the index data.
// Every load of the main loop is unconditional (rows clamped into the
// matrix, chunks clamped into the table): a load issued under a branch makes
// the compiler wait for all outstanding loads at the next use of any of
// them, which would drain the pipeline every step.  Clamped rows land in
// ring rows no neighbour index maps to.
// Three forms share the pipeline (MODE):
//  KL_PLAIN  Q = A P, partial p.q;
//  KL_DIR    the direction update folded into the product: every row that
//            lands in LDS is p = beta p_old + gamma z (the arithmetic of
//            kw_pd_update), the block's own rows are written to Pout (a
//            buffer other than Pin, so a block's halo rows read p_old even
//            after their owner has written p_new), then q = A p with the
//            partial p.q and no Q stored;
//  KL_UPD    q = A p recomputed in the same order (so the same bits as
//            KL_DIR's), then r -= alpha q and d += alpha p on the block's
//            own rows (kw_rupd's and kw_pd_update's arithmetic), partial r.r;
//  KL_RES    (FP64) the x update and residual of a residual replacement:
//            every row that lands in LDS is x = x_old + d / s (P is x_old, Df
//            the FP32 d; kw_x_update's arithmetic; no x_old read when x0 = 0,
//            no d when the round's d holds nothing, *dnone), the block's own
//            rows are written to Pout (another x buffer), then r = b - A x
//            (Z is b, own rows, loaded like KL_UPD's r), r scaled by the
//            column scale into Rf (FP32), partial r.r of the unscaled r
//            (kw_resid's arithmetic, the product in the window's order).
// With KL_DIR then KL_UPD an iteration's SpMM, residual and iterate updates
// and direction update move P in, P out, then P, R, D in and R, D out:
// 8 n x k passes instead of the 10 of the SpMM, kw_rupd and kw_pd_update.
// The own-row R and D values KL_UPD needs are loaded D steps ahead in
// registers (the thread that loads a row is the one that updates it).
enum { KL_PLAIN = 0, KL_DIR = 1, KL_UPD = 2, KL_RES = 3 };

template <typename T, int D, int MODE>
__global__ void __launch_bounds__(512)
kl_spmm(int32_t n, int32_t rpr, const unsigned char *__restrict__ wpk,
        const int32_t *__restrict__ far_ptr, const int32_t *__restrict__ far_rows, int64_t k,
        const T *__restrict__ P, const T *__restrict__ Z, T *__restrict__ Pout,
        T *__restrict__ Q, T *__restrict__ Rv, T *__restrict__ Dv,
        const double *__restrict__ st, double *__restrict__ part,
        const int32_t *__restrict__ ctl, float *__restrict__ Rf = nullptr,
        const float *__restrict__ Df = nullptr, const int32_t *__restrict__ dnone = nullptr,
        int xinit = 0) {
  if (ctl && ctl_idle(ctl)) return;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int W = 16 * VEC;          // columns of a tile (256 bytes)
  constexpr int CH = kLdsRows * W;     // elements of a chunk
  constexpr int IB = win_block_bytes<T>();
  // index words a thread stages (one T-sized word each; threads past the
  // block re-load its last word)
  using IW = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
  constexpr int NIW = IB / (int)sizeof(IW);
  static_assert(NIW <= 512, "index block larger than one word a thread");
  // (range, tile) of this block
  const int32_t R = (n + rpr - 1) / rpr;
  const int32_t tiles = (int32_t)((k + W - 1) / W);
  const int32_t pairs = R * tiles, per = (pairs + kXcd - 1) / kXcd;
  const int32_t pair = (int32_t)(blockIdx.x % kXcd) * per + (int32_t)(blockIdx.x / kXcd);
  if (pair >= pairs) return;
  const int32_t x = pair / tiles, y = pair - x * tiles;
  // the ring of chunks, then the far rows
  __shared__ __align__(16) T ring[kLdsSlots * CH + kLdsFar * W];
  T *const farc = ring + kLdsSlots * CH;
  __shared__ __align__(16) IW stage[2][NIW];
  const int tid = threadIdx.x;
  const int rsub = tid >> 4;
  const int lq = (tid & 15) * VEC;
  const int64_t c0 = (int64_t)y * W + lq;
  const bool valid = c0 < k;
  const int64_t cl = valid ? c0 : k - VEC;   // in-bounds loads for lanes past k
  const int32_t rs = x * rpr;
  const int32_t re = min(n, rs + rpr);
  const bool down = (x & 1) != 0;
  const int32_t len = re - rs;
  const int32_t nch = (len + kLdsRows - 1) / kLdsRows;
  const int32_t cbase = x * (rpr / kLdsRows);                // first index block
  const int32_t clast = R * (rpr / kLdsRows) - 1;            // last index block
  const int iw = min(tid, NIW - 1);
  auto row_of = [&](int32_t p) { return down ? re - 1 - p : rs + p; };
  // the column scalars (single system: slot c of k)
  T al[VEC], be[VEC], ga[VEC];
  double sc[VEC], iv[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) {
    sc[m] = (MODE == KL_RES && valid) ? st[ST_S * k + c0 + m] : 0.0;
    iv[m] = (MODE == KL_RES && valid) ? st[ST_INVS * k + c0 + m] : 0.0;
    al[m] = (MODE == KL_UPD && valid) ? (T)st[ST_ALPHA * k + c0 + m] : T(0);
    be[m] = (MODE == KL_DIR && valid) ? (T)st[ST_BETA * k + c0 + m] : T(0);
    ga[m] = (MODE == KL_DIR && valid) ? (T)st[ST_GAMMA * k + c0 + m] : T(0);
  }
  auto dir = [&](T (&v)[VEC], const T (&zv)[VEC]) {
#pragma unroll
    for (int m = 0; m < VEC; ++m) v[m] = fma(be[m], v[m], ga[m] * zv[m]);
  };
  // KL_RES: x_old is not read from x0 = 0, d not when it holds nothing
  const bool xskip = MODE == KL_RES && xinit != 0;
  const bool dskip = MODE == KL_RES && dnone && *dnone != 0;
  auto ld_x = [&](int32_t row, T (&v)[VEC]) {
    if (xskip) {
#pragma unroll
      for (int m = 0; m < VEC; ++m) v[m] = T(0);
    } else {
      VecIO<T, VEC>::ld(P + (int64_t)row * k + cl, v);
    }
  };
  auto ld_d = [&](int32_t row, float (&v)[VEC]) {
    if (dskip) {
#pragma unroll
      for (int m = 0; m < VEC; ++m) v[m] = 0.0f;
    } else {
      VecIO<float, VEC>::ld(Df + (int64_t)row * k + cl, v);
    }
  };
  auto xupd = [&](T (&v)[VEC], const float (&dv)[VEC]) {
#pragma unroll
    for (int m = 0; m < VEC; ++m) v[m] = fma(iv[m], (double)dv[m], (double)v[m]);
  };
  // far rows of the range, once
  {
    const int32_t fa = far_ptr[x], fb = far_ptr[x + 1];
    for (int32_t f = fa + rsub; f < fb; f += kLdsRows) {
      T v[VEC];
      if constexpr (MODE == KL_RES) {
        float dv[VEC];
        ld_x(far_rows[f], v);
        ld_d(far_rows[f], dv);
        xupd(v, dv);
      } else {
        VecIO<T, VEC>::ld(P + (int64_t)far_rows[f] * k + cl, v);
      }
      if constexpr (MODE == KL_DIR) {
        T zv[VEC];
        VecIO<T, VEC>::ld(Z + (int64_t)far_rows[f] * k + cl, zv);
        dir(v, zv);
      }
      VecIO<T, VEC>::st(farc + (f - fa) * W + lq, v);
    }
  }
  constexpr int DZ = MODE == KL_DIR ? D : 1;   // z of the landing rows
  constexpr int DR = (MODE == KL_UPD || MODE == KL_RES) ? D : 1;  // own-row r, d (b)
  constexpr int DX = MODE == KL_RES ? D : 1;   // d of the landing rows (FP32)
  T pb[D][VEC], zb[DZ][VEC], rb[DR][VEC], db[DR][VEC];
  float xd[DX][VEC];
  IW ib[D];
  auto issue = [&](int32_t t, int u) {
    const int32_t g = t + kLdsAhead + D;                 // P chunk
    const int32_t row = min(max(row_of(g * kLdsRows + rsub), 0), n - 1);
    if constexpr (MODE == KL_RES) {
      ld_x(row, pb[u]);
      ld_d(row, xd[u]);
    } else {
      VecIO<T, VEC>::ld(P + (int64_t)row * k + cl, pb[u]);
    }
    if constexpr (MODE == KL_DIR) VecIO<T, VEC>::ld(Z + (int64_t)row * k + cl, zb[u]);
    const int32_t gi = min(max(cbase + t + D, 0), clast);  // index chunk
    ib[u] = reinterpret_cast<const IW *>(wpk + (int64_t)gi * IB)[iw];
  };
  // own-row r and d of compute chunk t + D, after compute(t) read the buffer
  // (d is 0 in a round's first iteration, which reads none)
  const bool dfresh = MODE == KL_UPD && ctl[CTL_DFRESH] != 0;
  auto issue_rd = [&](int32_t t, int u) {
    const int32_t row = min(max(row_of((t + D) * kLdsRows + rsub), 0), n - 1);
    if constexpr (MODE == KL_RES) {
      VecIO<T, VEC>::ld(Z + (int64_t)row * k + cl, rb[u]);  // b
      return;
    }
    VecIO<T, VEC>::ld(Rv + (int64_t)row * k + cl, rb[u]);
    if (dfresh) {
#pragma unroll
      for (int m = 0; m < VEC; ++m) db[u][m] = T(0);
    } else {
      VecIO<T, VEC>::ld(Dv + (int64_t)row * k + cl, db[u]);
    }
  };
  auto land = [&](int32_t t, int u) {
    const int32_t g = t + kLdsAhead;
    if constexpr (MODE == KL_DIR || MODE == KL_RES) {
      if constexpr (MODE == KL_DIR) dir(pb[u], zb[u]);
      else xupd(pb[u], xd[u]);
      // the block's own rows go out once (halo rows belong to another range)
      const int32_t p = g * kLdsRows + rsub;
      if (valid && p >= 0 && p < len)
        VecIO<T, VEC>::st(Pout + (int64_t)row_of(p) * k + c0, pb[u]);
    }
    VecIO<T, VEC>::st(ring + ((g + kLdsSlots) % kLdsSlots) * CH + rsub * W + lq, pb[u]);
    stage[t & 1][iw] = ib[u];
  };
  double pd[VEC] = {};
  auto compute = [&](int32_t t, int u) {
    const int32_t p = t * kLdsRows + rsub;
    if (p >= len) return;
    const int32_t row = row_of(p);
    const int32_t *jc = reinterpret_cast<const int32_t *>(&stage[t & 1][0]) + rsub * kWinEll;
    const T *wv = reinterpret_cast<const T *>(reinterpret_cast<const unsigned char *>(
                      &stage[t & 1][0]) + 4 * kWinEnt) + rsub * kWinEll;
    T acc[VEC] = {}, x0[VEC];
#pragma unroll
    for (int e = 0; e < kWinEll; ++e) {
      T v[VEC];
      VecIO<T, VEC>::ld(ring + jc[e] * W + lq, v);
      const T w = wv[e];
#pragma unroll
      for (int m = 0; m < VEC; ++m) acc[m] = fma(w, v[m], acc[m]);
      if (e == 0) {
#pragma unroll
        for (int m = 0; m < VEC; ++m) x0[m] = v[m];
      }
    }
    if constexpr (MODE == KL_UPD) {
      T rv[VEC], dv[VEC];
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        rv[m] = fma(-al[m], acc[m], rb[u][m]);
        dv[m] = fma(al[m], x0[m], db[u][m]);
        pd[m] += (double)rv[m] * (double)rv[m];
      }
      if (valid) {
        VecIO<T, VEC>::st(Rv + (int64_t)row * k + c0, rv);
        VecIO<T, VEC>::st(Dv + (int64_t)row * k + c0, dv);
      }
    } else if constexpr (MODE == KL_RES) {
      float out[VEC];
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        const double rv = (double)rb[u][m] - (double)acc[m];
        pd[m] += rv * rv;
        out[m] = (float)(rv * sc[m]);
      }
      if (valid) VecIO<float, VEC>::st(Rf + (int64_t)row * k + c0, out);
    } else {
#pragma unroll
      for (int m = 0; m < VEC; ++m) pd[m] += (double)x0[m] * (double)acc[m];
      if (MODE == KL_PLAIN && valid) VecIO<T, VEC>::st(Q + (int64_t)row * k + c0, acc);
    }
  };
  const int32_t t0 = -kLdsBack - kLdsAhead;
  // prologue: the first D steps' loads
#pragma unroll
  for (int u = 0; u < D; ++u) issue(t0 - D + u, u);
  __syncthreads();  // the far cache
  // steps t0 .. nch - 1, rounded up to whole groups of D: every group runs
  // all its steps (the ones past nch compute nothing), so the compiler sees
  // the same loads in flight on every path and waits only for the oldest
  const int32_t groups = (nch - t0 + D - 1) / D;
  for (int32_t gi = 0; gi < groups; ++gi) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int32_t t = t0 + gi * D + u;
      land(t, u);
      issue(t, u);
      __syncthreads();
      if (t >= 0 && t < nch) compute(t, u);
      if constexpr (MODE == KL_UPD || MODE == KL_RES) issue_rd(t, u);
    }
  }
  // block partials (p.q, or r.r for KL_UPD): the 32 row slots of each
  // column, in order
  __syncthreads();
  double *sh = reinterpret_cast<double *>(ring);
#pragma unroll
  for (int m = 0; m < VEC; ++m) sh[rsub * W + lq + m] = pd[m];
  __syncthreads();
  if (tid < W) {
    double sum = 0.0;
    for (int q = 0; q < kLdsRows; ++q) sum += sh[q * W + tid];
    const int64_t c = (int64_t)y * W + tid;
    if (c < k) part[(int64_t)x * k + c] = sum;
  }
}

// The packed index blocks of the window row table from the CSR values:
// block c holds the LDS rows of its 32 walking positions' kWinEll slots,
// then their weights (q < 0: weight 0).
template <typename T>
__global__ void k_win_pack(int64_t total, const int32_t *__restrict__ wcol,
                           const int32_t *__restrict__ wq, const double *__restrict__ val,
                           unsigned char *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int64_t c = e / kWinEnt;
  const int i = (int)(e - c * kWinEnt);
  const int32_t q = wq[e];
  unsigned char *b = out + c * win_block_bytes<T>();
  reinterpret_cast<int32_t *>(b)[i] = wcol[e];
  reinterpret_cast<T *>(b + 4 * kWinEnt)[i] = q >= 0 ? (T)val[q] : T(0);
}

// Row table of A_ff.  Row r lists its diagonal first (column r, weight 0
// if the row stores none), then its other entries in column order; slots
// past the row's length repeat its last column with weight 0, and entries
// past kEll go to the overflow CSR in the same order.
__global__ void k_ell_fill(int32_t n, const int32_t *__restrict__ ptr,
                           const int32_t *__restrict__ col, const double *__restrict__ val,
                           int32_t *__restrict__ ecol, float *__restrict__ ev32,
                           double *__restrict__ ev64, const int32_t *__restrict__ ovp,
                           int32_t *__restrict__ ovc, float *__restrict__ ov32,
                           double *__restrict__ ov64) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t a = ptr[r], b = ptr[r + 1];
  double d = 0.0;
  for (int32_t q = a; q < b; ++q)
    if (col[q] == r) d = val[q];
  const int64_t o = (int64_t)r * kEll;
  ecol[o] = r;
  ev64[o] = d;
  ev32[o] = (float)d;
  int32_t u = 1, last = r, ov = ovp ? ovp[r] : 0;
  for (int32_t q = a; q < b; ++q) {
    if (col[q] == r) continue;
    if (u < kEll) {
      ecol[o + u] = col[q];
      ev64[o + u] = val[q];
      ev32[o + u] = (float)val[q];
      ++u;
    } else {
      ovc[ov] = col[q];
      ov64[ov] = val[q];
      ov32[ov] = (float)val[q];
      ++ov;
    }
    last = col[q];
  }
  for (; u < kEll; ++u) {
    ecol[o + u] = last;
    ev64[o + u] = 0.0;
    ev32[o + u] = 0.0f;
  }
}

// ---- FP64 residual r = b - A x with partial r.r ------------------------
// R = b - A x, or (scale != nullptr) R = scale[col] (b - A x) in the
// inner-loop type TO; partial r.r of the unscaled residual.
// Thread form (k <= 4): one thread per row, its K columns in registers; a
// chunk of 8 entries' indices, then all their gathers, are issued at once
// (the tail of a chunk re-reads the last entry with weight 0), so each
// column's sum is the in-order one.  Block partials: runs of 16 threads,
// then the 16 runs, in a fixed order.
template <int K, typename TO>
__global__ void __launch_bounds__(kBlock)
kt_resid(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
         const double *__restrict__ val, const double *__restrict__ X,
         const double *__restrict__ B, TO *__restrict__ R, const double *__restrict__ scale,
         double *__restrict__ part) {
  constexpr int CH = 8;
  double pd[K] = {};
  double sc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) sc[c] = scale ? scale[c] : 1.0;
  for (int32_t row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
    double acc[K] = {};
    const int32_t a = ptr[row], b = ptr[row + 1];
    for (int32_t e = a; e < b; e += CH) {
      int32_t j[CH];
      double v[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int32_t q = min(e + u, b - 1);
        j[u] = col[q];
        v[u] = (e + u < b) ? val[q] : 0.0;
      }
      double x[CH][K];
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
    for (int c = 0; c < K; ++c) {
      const double r = B[(int64_t)row * K + c] - acc[c];
      R[(int64_t)row * K + c] = (TO)(scale ? r * sc[c] : r);
      pd[c] += r * r;
    }
  }
  __shared__ double sh[K][kBlock];
  __shared__ double s2[K][kBlock / 16];
#pragma unroll
  for (int c = 0; c < K; ++c) sh[c][threadIdx.x] = pd[c];
  __syncthreads();
  if (threadIdx.x < K * (kBlock / 16)) {
    const int c = threadIdx.x / (kBlock / 16), m = threadIdx.x % (kBlock / 16);
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) t += sh[c][m * 16 + u];
    s2[c][m] = t;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
    for (int m = 0; m < kBlock / 16; ++m) t += s2[threadIdx.x][m];
    part[(int64_t)blockIdx.x * K + threadIdx.x] = t;
  }
}

template <int VEC, typename TO, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_resid(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
         const double *__restrict__ val, int64_t k, const double *__restrict__ X,
         const double *__restrict__ B, TO *__restrict__ R, const double *__restrict__ scale,
         double *__restrict__ part) {
  WIDE_COLS_P(VEC, LPR)
  double pd[1][VEC] = {};
  double sc[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) sc[m] = (scale && valid) ? scale[c0 + m] : 1.0;
  WIDE_LOOP_P(row, 0, n, LPR) {
    const int32_t r = Pack<LPR>::u(row);
    const int32_t a = Pack<LPR>::u(ptr[r]), b = Pack<LPR>::u(ptr[r + 1]);
    double acc[VEC] = {};
    gather_wide<double, VEC>(a, b, col, val, X, k, c0, valid, acc);
    if (valid) {
      double bb[VEC];
      TO out[VEC];
      VecIO<double, VEC>::ld(B + (int64_t)r * k + c0, bb);
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        acc[m] = bb[m] - acc[m];
        pd[0][m] += acc[m] * acc[m];
        out[m] = (TO)(scale ? acc[m] * sc[m] : acc[m]);
      }
      VecIO<TO, VEC>::st(R + (int64_t)r * k + c0, out);
    }
  }
  block_partials_wide<VEC, 1, LPR>(pd, part, k, c0, 0);
}

// ---- IC(0) forward pass over the blocks [b0, b1) of one level:
//      (FUSED) r -= alpha q;  y = (r - sum L y) / L_ii;  partials r.r, y.y
// A block's rows are walked in order by the same lanes that own the
// columns they write, so reads of rows written earlier in the block are
// same-thread reads.
template <typename T, int K, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kn_fwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, const int32_t *__restrict__ lp,
       const int32_t *__restrict__ lcol, const T *__restrict__ lval,
       const T *__restrict__ rdiag, const double *__restrict__ alpha,
       const T *__restrict__ Q, T *__restrict__ R, T *Y, double *__restrict__ part,
       int64_t stride, const int32_t *__restrict__ ctl) {
  // one row per block: 8-lane groups, lanes stride over the row's entries
  if (ctl_idle(ctl)) return;
  const int g = threadIdx.x % kGroup;
  double pd[2] = {0.0, 0.0};
  const T al = (FUSED && g < K) ? (T)alpha[g] : T(0);
  NARROW_LOOP(bi, b0, b1) {
    const int32_t pos = blk_ptr[bi];
    const int64_t row = rowmap[pos];
    T acc[K] = {};
    gather_narrow<T, K>(lp[pos], lp[pos + 1], g, lcol, lval, Y, acc);
    if (g < K) {
      const int64_t o = row * K + g;
      T r = R[o];
      if (FUSED) {
        r = fma(-al, Q[o], r);
        R[o] = r;
      }
      T y = (r - pick<T, K>(acc, g)) * rdiag[pos];
      Y[o] = y;
      pd[0] += (double)r * (double)r;
      pd[1] += (double)y * (double)y;
    }
  }
  block_partials_narrow<K, 2>(pd, part, stride);
}

// multi-row blocks, k <= 4: one thread per block, all K columns
template <typename T, int K, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kt_fwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, const int32_t *__restrict__ lp,
       const int32_t *__restrict__ lcol, const T *__restrict__ lval,
       const T *__restrict__ rdiag, const double *__restrict__ alpha,
       const T *__restrict__ Q, T *__restrict__ R, T *Y, double *__restrict__ part,
       int64_t stride, const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  double pd[2][K] = {};
  T al[K];
#pragma unroll
  for (int c = 0; c < K; ++c) al[c] = FUSED ? (T)alpha[c] : T(0);
  const RowIter ri = row_iter(b0, b1, kBlock, threadIdx.x);
  for (int32_t bi = ri.first; bi < ri.end; bi += ri.stride) {
    for (int32_t pos = blk_ptr[bi]; pos < blk_ptr[bi + 1]; ++pos) {
      const int64_t row = rowmap[pos];
      T acc[K] = {};
      for (int32_t e = lp[pos]; e < lp[pos + 1]; ++e) {
        const int64_t j = lcol[e];
        const T v = lval[e];
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] = fma(v, Y[j * K + c], acc[c]);
      }
      const T rd = rdiag[pos];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const int64_t o = row * K + c;
        T r = R[o];
        if (FUSED) {
          r = fma(-al[c], Q[o], r);
          R[o] = r;
        }
        T y = (r - acc[c]) * rd;
        Y[o] = y;
        pd[0][c] += (double)r * (double)r;
        pd[1][c] += (double)y * (double)y;
      }
    }
  }
  block_partials_thread<K, 2>(pd, part, stride);
}

template <typename T, int VEC, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kw_fwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, const int32_t *__restrict__ lp,
       const int32_t *__restrict__ lcol, const T *__restrict__ lval,
       const T *__restrict__ rdiag, const double *__restrict__ alpha, int64_t k,
       const T *__restrict__ Q, T *__restrict__ R, T *Y, double *__restrict__ part,
       int64_t stride, const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  WIDE_COLS(VEC)
  double pd[2][VEC] = {};
  T al[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) al[m] = (FUSED && valid) ? (T)alpha[c0 + m] : T(0);
  WIDE_LOOP(bi, b0, b1) {
    const int32_t bq = uniform(bi);
    const int32_t p0 = uniform(blk_ptr[bq]), p1 = uniform(blk_ptr[bq + 1]);
    for (int32_t ps = p0; ps < p1; ++ps) {
      const int32_t r = uniform(rowmap[ps]);
      const int32_t a = uniform(lp[ps]), b = uniform(lp[ps + 1]);
      T acc[VEC] = {};
      gather_wide<T, VEC, 4>(a, b, lcol, lval, Y, k, c0, valid, acc);
      if (valid) {
        const int64_t o = (int64_t)r * k + c0;
        T rv[VEC];
        VecIO<T, VEC>::ld(R + o, rv);
        if (FUSED) {
          T qv[VEC];
          VecIO<T, VEC>::ld(Q + o, qv);
#pragma unroll
          for (int m = 0; m < VEC; ++m) rv[m] = fma(-al[m], qv[m], rv[m]);
          VecIO<T, VEC>::st(R + o, rv);
        }
        const T rd = rdiag[ps];
#pragma unroll
        for (int m = 0; m < VEC; ++m) {
          acc[m] = (rv[m] - acc[m]) * rd;
          pd[0][m] += (double)rv[m] * (double)rv[m];
          pd[1][m] += (double)acc[m] * (double)acc[m];
        }
        VecIO<T, VEC>::st(Y + o, acc);
      }
    }
  }
  block_partials_wide<VEC, 2>(pd, part, k, c0, stride);
}

// ---- IC(0) backward pass over the blocks [b0, b1) of one level, in place
//      on Z (which holds y), block rows in reverse:
//      z = (y - sum U z) / L_ii;  d += alpha p;  p = gamma z + beta p
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
kn_bwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, const int32_t *__restrict__ up,
       const int32_t *__restrict__ ucol, const T *__restrict__ uval,
       const T *__restrict__ rdiag, const double *__restrict__ st, int64_t k, T *Z,
       T *__restrict__ P, T *__restrict__ D, const int32_t *__restrict__ ctl) {
  if (bwd_idle(ctl)) return;
  const int g = threadIdx.x % kGroup;
  const int gc = g < K ? g : 0;
  const T al = (T)st[ST_ALPHA * k + gc], be = (T)st[ST_BETA * k + gc],
          ga = (T)st[ST_GAMMA * k + gc];
  NARROW_LOOP(bi, b0, b1) {
    const int32_t pos = blk_ptr[bi];
    const int64_t row = rowmap[pos];
    T acc[K] = {};
    gather_narrow<T, K>(up[pos], up[pos + 1], g, ucol, uval, Z, acc);
    if (g < K) {
      const int64_t o = row * K + g;
      T z = (Z[o] - pick<T, K>(acc, g)) * rdiag[pos];
      Z[o] = z;
      T p = P[o];
      D[o] = fma(al, p, D[o]);
      P[o] = fma(be, p, ga * z);
    }
  }
}

template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
kt_bwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, const int32_t *__restrict__ up,
       const int32_t *__restrict__ ucol, const T *__restrict__ uval,
       const T *__restrict__ rdiag, const double *__restrict__ st, int64_t k, T *Z,
       T *__restrict__ P, T *__restrict__ D, const int32_t *__restrict__ ctl) {
  if (bwd_idle(ctl)) return;
  T al[K], be[K], ga[K];
#pragma unroll
  for (int c = 0; c < K; ++c) {
    al[c] = (T)st[ST_ALPHA * k + c];
    be[c] = (T)st[ST_BETA * k + c];
    ga[c] = (T)st[ST_GAMMA * k + c];
  }
  const RowIter ri = row_iter(b0, b1, kBlock, threadIdx.x);
  for (int32_t bi = ri.first; bi < ri.end; bi += ri.stride) {
    for (int32_t pos = blk_ptr[bi + 1] - 1; pos >= blk_ptr[bi]; --pos) {
      const int64_t row = rowmap[pos];
      T acc[K] = {};
      for (int32_t e = up[pos]; e < up[pos + 1]; ++e) {
        const int64_t j = ucol[e];
        const T v = uval[e];
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] = fma(v, Z[j * K + c], acc[c]);
      }
      const T rd = rdiag[pos];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const int64_t o = row * K + c;
        T z = (Z[o] - acc[c]) * rd;
        Z[o] = z;
        T p = P[o];
        D[o] = fma(al[c], p, D[o]);
        P[o] = fma(be[c], p, ga[c] * z);
      }
    }
  }
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
kw_bwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, const int32_t *__restrict__ up,
       const int32_t *__restrict__ ucol, const T *__restrict__ uval,
       const T *__restrict__ rdiag, const double *__restrict__ st, int64_t k, T *Z,
       T *__restrict__ P, T *__restrict__ D, const int32_t *__restrict__ ctl) {
  if (bwd_idle(ctl)) return;
  WIDE_COLS(VEC)
  T al[VEC], be[VEC], ga[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) {
    al[m] = valid ? (T)st[ST_ALPHA * k + c0 + m] : T(0);
    be[m] = valid ? (T)st[ST_BETA * k + c0 + m] : T(0);
    ga[m] = valid ? (T)st[ST_GAMMA * k + c0 + m] : T(0);
  }
  WIDE_LOOP(bi, b0, b1) {
    const int32_t bq = uniform(bi);
    const int32_t p0 = uniform(blk_ptr[bq]), p1 = uniform(blk_ptr[bq + 1]);
    for (int32_t ps = p1 - 1; ps >= p0; --ps) {
      const int32_t r = uniform(rowmap[ps]);
      const int32_t a = uniform(up[ps]), b = uniform(up[ps + 1]);
      T acc[VEC] = {};
      gather_wide<T, VEC, 4>(a, b, ucol, uval, Z, k, c0, valid, acc);
      if (valid) {
        const int64_t o = (int64_t)r * k + c0;
        T yv[VEC], pv[VEC], dv[VEC];
        VecIO<T, VEC>::ld(Z + o, yv);
        VecIO<T, VEC>::ld(P + o, pv);
        VecIO<T, VEC>::ld(D + o, dv);
        const T rd = rdiag[ps];
#pragma unroll
        for (int m = 0; m < VEC; ++m) {
          acc[m] = (yv[m] - acc[m]) * rd;
          dv[m] = fma(al[m], pv[m], dv[m]);
          pv[m] = fma(be[m], pv[m], ga[m] * acc[m]);
        }
        VecIO<T, VEC>::st(Z + o, acc);
        VecIO<T, VEC>::st(P + o, pv);
        VecIO<T, VEC>::st(D + o, dv);
      }
    }
  }
}

// ---- Block sweeps: one workgroup per (block, column tile) ---------------
// A block's rows depend on earlier levels (final when the level starts) and
// on earlier rows of the same block.  Phase 1 gathers the first kind for
// the block's rows (wave w takes rows w, w+4, ..., up to 4 rows at once so
// their loads are in flight together; the first kSweepEll entries of a row
// come from a fixed-width table, so no row-pointer load precedes them);
// phase 2 solves the dense in-block triangle, one thread per column, with
// the column in registers and the triangle in scalar registers; phase 3
// writes the rows back.
template <typename T>
struct SweepView {
  const int32_t *ell_col, *ell_cnt, *ov_ptr, *ov_col;
  const T *ell_val, *ov_val, *dense;
};

#define SWEEP_PARAMS(T)                                                        \
  const int32_t *__restrict__ ell_col, const int32_t *__restrict__ ell_cnt,   \
      const int32_t *__restrict__ ov_ptr, const int32_t *__restrict__ ov_col, \
      const T *__restrict__ ell_val, const T *__restrict__ ov_val,            \
      const T *__restrict__ dense
#define SWEEP_ARGS(sw) sw.ell_col, sw.ell_cnt, sw.ov_ptr, sw.ov_col, sw.ell_val, sw.ov_val, sw.dense

template <typename T, int VEC, int MB>
struct BlockShared {
  static constexpr int TW = 64 * VEC;
  static constexpr size_t kLoc = sizeof(T) * MB * TW;
  static constexpr size_t kPart = sizeof(double) * kWaves * 2 * TW;
  static constexpr size_t kBytes = kLoc > kPart ? kLoc : kPart;
};

// acc[g] += sum over row ps[g]'s entries in other blocks of val * X[col]
template <typename T, int VEC, int G>
__device__ __forceinline__ void sweep_gather(const int32_t (&ps)[G], const bool (&on)[G],
                                             const int32_t *__restrict__ ell_col,
                                             const int32_t *__restrict__ ell_cnt,
                                             const int32_t *__restrict__ ov_ptr,
                                             const int32_t *__restrict__ ov_col,
                                             const T *__restrict__ ell_val,
                                             const T *__restrict__ ov_val, const T *X,
                                             int64_t k, int64_t c0, bool valid,
                                             T (&acc)[G][VEC]) {
  constexpr int W = cn_ic0_t::kSweepEll;
  int32_t cnt[G], oa[G], ob[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    cnt[g] = on[g] ? uniform(ell_cnt[ps[g]]) : 0;
    oa[g] = uniform(ov_ptr[ps[g]]);
    ob[g] = on[g] ? uniform(ov_ptr[ps[g] + 1]) : oa[g];
  }
  int32_t jc[G][W];
  T vv[G][W];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int u = 0; u < W; ++u) {
      jc[g][u] = uniform(ell_col[(int64_t)ps[g] * W + u]);
      vv[g][u] = ell_val[(int64_t)ps[g] * W + u];
    }
  T x[G][W][VEC];
#pragma unroll
  for (int g = 0; g < G; ++g)
    if (cnt[g] > 0 && valid) {
#pragma unroll
      for (int u = 0; u < W; ++u) VecIO<T, VEC>::ld(X + (int64_t)jc[g][u] * k + c0, x[g][u]);
    }
#pragma unroll
  for (int g = 0; g < G; ++g)
    if (cnt[g] > 0 && valid) {
#pragma unroll
      for (int u = 0; u < W; ++u)
#pragma unroll
        for (int m = 0; m < VEC; ++m) acc[g][m] = fma(vv[g][u], x[g][u][m], acc[g][m]);
    }
#pragma unroll
  for (int g = 0; g < G; ++g)
    if (ob[g] > oa[g]) gather_wide<T, VEC, 4>(oa[g], ob[g], ov_col, ov_val, X, k, c0, valid, acc[g]);
}

// In-block triangle of block bq for the column `c` of the tile in `loc`
// (rows [0, nb)): forward y_t = (s_t - sum_{u<t} L_tu y_u) / L_tt, or
// backward z_t = (s_t - sum_{u>t} U_tu z_u) / L_tt.
template <typename T, int MB, bool UPPER>
__device__ __forceinline__ void block_solve(T *loc, int TW, int c, int32_t nb,
                                            const T *__restrict__ tri) {
  T y[MB];
#pragma unroll
  for (int t = 0; t < MB; ++t) y[t] = t < nb ? loc[t * TW + c] : T(0);
  if (!UPPER) {
#pragma unroll
    for (int t = 0; t < MB; ++t) {
      T v = y[t];
#pragma unroll
      for (int u = 0; u < t; ++u) v = fma(-tri[t * MB + u], y[u], v);
      y[t] = v * tri[t * MB + t];
    }
  } else {
#pragma unroll
    for (int t = MB - 1; t >= 0; --t) {
      T v = y[t];
#pragma unroll
      for (int u = t + 1; u < MB; ++u) v = fma(-tri[t * MB + u], y[u], v);
      y[t] = v * tri[t * MB + t];
    }
  }
#pragma unroll
  for (int t = 0; t < MB; ++t)
    if (t < nb) loc[t * TW + c] = y[t];
}

// forward: (FUSED: r -= alpha q) then y = L^-1 r over the blocks [b0, b1);
// partials r.r and y.y
template <typename T, int VEC, int MB, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kb_fwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, SWEEP_PARAMS(T),
       const double *__restrict__ alpha, int64_t k, const T *__restrict__ Q,
       T *__restrict__ R, T *Y, double *__restrict__ part, int64_t stride,
       const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  using SH = BlockShared<T, VEC, MB>;
  constexpr int TW = SH::TW, RPW = MB / kWaves, G = RPW < 4 ? RPW : 4;
  __shared__ __align__(16) unsigned char smem[SH::kBytes];
  T *loc = reinterpret_cast<T *>(smem);
  WIDE_COLS(VEC)
  const int wave = threadIdx.x >> 6;
  double pd[2][VEC] = {};
  T al[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) al[m] = (FUSED && valid) ? (T)alpha[c0 + m] : T(0);
  const RowIter ri = row_iter(b0, b1, 1, 0);
  for (int32_t bq = ri.first; bq < ri.end; bq += ri.stride) {
    const int32_t p0 = uniform(blk_ptr[bq]), nb = uniform(blk_ptr[bq + 1]) - p0;
#pragma unroll 1
    for (int g0 = 0; g0 < RPW; g0 += G) {
      if (wave + kWaves * g0 >= nb) break;
      int32_t ps[G], rows[G];
      bool on[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int t = wave + kWaves * (g0 + g);
        on[g] = t < nb;
        ps[g] = p0 + (on[g] ? t : 0);
        rows[g] = uniform(rowmap[ps[g]]);
      }
      T acc[G][VEC] = {};
      sweep_gather<T, VEC, G>(ps, on, ell_col, ell_cnt, ov_ptr, ov_col, ell_val, ov_val, Y, k,
                              c0, valid, acc);
      T rv[G][VEC] = {}, qv[G][VEC] = {};
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (on[g] && valid) {
          const int64_t o = (int64_t)rows[g] * k + c0;
          VecIO<T, VEC>::ld(R + o, rv[g]);
          if (FUSED) VecIO<T, VEC>::ld(Q + o, qv[g]);
        }
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (on[g]) {
          if (FUSED) {
#pragma unroll
            for (int m = 0; m < VEC; ++m) rv[g][m] = fma(-al[m], qv[g][m], rv[g][m]);
            if (valid) VecIO<T, VEC>::st(R + (int64_t)rows[g] * k + c0, rv[g]);
          }
#pragma unroll
          for (int m = 0; m < VEC; ++m) {
            pd[0][m] += (double)rv[g][m] * (double)rv[g][m];
            acc[g][m] = rv[g][m] - acc[g][m];
          }
          VecIO<T, VEC>::st(loc + (wave + kWaves * (g0 + g)) * TW + lane * VEC, acc[g]);
        }
    }
    __syncthreads();
    if (threadIdx.x < TW)
      block_solve<T, MB, false>(loc, TW, threadIdx.x, nb, dense + (int64_t)bq * MB * MB);
    __syncthreads();
#pragma unroll 1
    for (int t = wave; t < nb; t += kWaves) {
      const int32_t r = uniform(rowmap[p0 + t]);
      T y[VEC];
      VecIO<T, VEC>::ld(loc + t * TW + lane * VEC, y);
      if (valid) VecIO<T, VEC>::st(Y + (int64_t)r * k + c0, y);
#pragma unroll
      for (int m = 0; m < VEC; ++m) pd[1][m] += (double)y[m] * (double)y[m];
    }
    __syncthreads();
  }
  block_partials_wide_in<VEC, 2>(pd, part, k, c0, stride, reinterpret_cast<double *>(smem));
}

// backward, in place on Z (holding y): z = L^-T y over the blocks [b0, b1);
// d += alpha p;  p = gamma z + beta p
template <typename T, int VEC, int MB>
__global__ void __launch_bounds__(kBlock)
kb_bwd(int32_t b0, int32_t b1, const int32_t *__restrict__ blk_ptr,
       const int32_t *__restrict__ rowmap, SWEEP_PARAMS(T),
       const double *__restrict__ st, int64_t k, T *Z, T *__restrict__ P,
       T *__restrict__ D, const int32_t *__restrict__ ctl) {
  if (bwd_idle(ctl)) return;
  using SH = BlockShared<T, VEC, MB>;
  constexpr int TW = SH::TW, RPW = MB / kWaves, G = RPW < 4 ? RPW : 4;
  __shared__ __align__(16) unsigned char smem[SH::kBytes];
  T *loc = reinterpret_cast<T *>(smem);
  WIDE_COLS(VEC)
  const int wave = threadIdx.x >> 6;
  T al[VEC], be[VEC], ga[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) {
    al[m] = valid ? (T)st[ST_ALPHA * k + c0 + m] : T(0);
    be[m] = valid ? (T)st[ST_BETA * k + c0 + m] : T(0);
    ga[m] = valid ? (T)st[ST_GAMMA * k + c0 + m] : T(0);
  }
  const RowIter ri = row_iter(b0, b1, 1, 0);
  for (int32_t bq = ri.first; bq < ri.end; bq += ri.stride) {
    const int32_t p0 = uniform(blk_ptr[bq]), nb = uniform(blk_ptr[bq + 1]) - p0;
#pragma unroll 1
    for (int g0 = 0; g0 < RPW; g0 += G) {
      if (wave + kWaves * g0 >= nb) break;
      int32_t ps[G], rows[G];
      bool on[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int t = wave + kWaves * (g0 + g);
        on[g] = t < nb;
        ps[g] = p0 + (on[g] ? t : 0);
        rows[g] = uniform(rowmap[ps[g]]);
      }
      T acc[G][VEC] = {};
      sweep_gather<T, VEC, G>(ps, on, ell_col, ell_cnt, ov_ptr, ov_col, ell_val, ov_val, Z, k,
                              c0, valid, acc);
      T yv[G][VEC] = {};
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (on[g] && valid) VecIO<T, VEC>::ld(Z + (int64_t)rows[g] * k + c0, yv[g]);
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (on[g]) {
#pragma unroll
          for (int m = 0; m < VEC; ++m) acc[g][m] = yv[g][m] - acc[g][m];
          VecIO<T, VEC>::st(loc + (wave + kWaves * (g0 + g)) * TW + lane * VEC, acc[g]);
        }
    }
    __syncthreads();
    if (threadIdx.x < TW)
      block_solve<T, MB, true>(loc, TW, threadIdx.x, nb, dense + (int64_t)bq * MB * MB);
    __syncthreads();
#pragma unroll 1
    for (int g0 = 0; g0 < RPW; g0 += G) {
      if (wave + kWaves * g0 >= nb) break;
      int32_t rows[G];
      bool on[G];
      T z[G][VEC] = {}, pv[G][VEC] = {}, dv[G][VEC] = {};
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int t = wave + kWaves * (g0 + g);
        on[g] = t < nb;
        rows[g] = uniform(rowmap[p0 + (on[g] ? t : 0)]);
        if (on[g]) {
          VecIO<T, VEC>::ld(loc + t * TW + lane * VEC, z[g]);
          if (valid) {
            const int64_t o = (int64_t)rows[g] * k + c0;
            VecIO<T, VEC>::ld(P + o, pv[g]);
            VecIO<T, VEC>::ld(D + o, dv[g]);
          }
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (on[g] && valid) {
          const int64_t o = (int64_t)rows[g] * k + c0;
#pragma unroll
          for (int m = 0; m < VEC; ++m) {
            dv[g][m] = fma(al[m], pv[g][m], dv[g][m]);
            pv[g][m] = fma(be[m], pv[g][m], ga[m] * z[g][m]);
          }
          VecIO<T, VEC>::st(Z + o, z[g]);
          VecIO<T, VEC>::st(P + o, pv[g]);
          VecIO<T, VEC>::st(D + o, dv[g]);
        }
    }
    __syncthreads();
  }
}

// ---- (FUSED) r -= alpha q with partial r.r: the first half of a
//      preconditioner applied as a whole (HSC), whose r.z comes after it.
// Flat form for k <= 4 (kf_rupd), wide form for k > 4 (kw_rupd).
// ---- flat form of the narrow streaming kernels (k <= 4): the n x K array
// is walked as one flat range, lane after lane, so a wavefront reads 64
// consecutive values; a block of kFlat(K) threads (a multiple of K) and a
// grid stride that is a multiple of K keep each thread on one column.  The
// grid is the narrow row grid's block count, so the partial rows match.
__host__ __device__ constexpr int kFlat(int K) { return K == 3 ? 768 : 1024; }

// per-block column sums of the threads' partials in a fixed two-stage
// order: runs of 16 threads of a column, then the runs in order (a serial
// walk over the 256-1024 threads of a column would cost microseconds of
// LDS latency a launch)
template <int K, int NP>
__device__ __forceinline__ void flat_partials(const double (&p)[NP], double *__restrict__ out,
                                              int64_t stride_np) {
  constexpr int B = kFlat(K), M = B / K, R = 16, M2 = M / R;
  __shared__ double sh[NP][B];
  __shared__ double s2[NP][K * M2];
#pragma unroll
  for (int q = 0; q < NP; ++q) sh[q][threadIdx.x] = p[q];
  __syncthreads();
  if (threadIdx.x < NP * K * M2) {
    const int q = threadIdx.x / (K * M2), r = threadIdx.x % (K * M2);
    const int c = r % K, m2 = r / K;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < R; ++u) s += sh[q][c + K * (m2 * R + u)];
    s2[q][r] = s;
  }
  __syncthreads();
  if (threadIdx.x < NP * K) {
    const int q = threadIdx.x / K, c = threadIdx.x % K;
    double s = 0.0;
    for (int m2 = 0; m2 < M2; ++m2) s += s2[q][c + K * m2];
    out[q * stride_np + (int64_t)blockIdx.x * K + c] = s;
  }
}

// (FUSED) r -= alpha q; partial r.r
template <typename T, int K, bool FUSED>
__global__ void __launch_bounds__(1024)
kf_rupd(int32_t n, const double *__restrict__ alpha, SegMap sm, const T *__restrict__ Q,
        T *__restrict__ R, double *__restrict__ part, const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  constexpr int B = kFlat(K);
  const int c = threadIdx.x % K;
  const int64_t total = (int64_t)n * K, S = (int64_t)gridDim.x * B;
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x, row = i / K;
  const T al0 = FUSED ? (T)alpha[c] : T(0);
  double pd[1] = {0.0};
  for (; i < total; i += S, row += S / K) {
    T r = R[i];
    if (FUSED) {
      const T al = sm.of_row ? (T)alpha[sm.vc(row, K, c)] : al0;
      r = fma(-al, Q[i], r);
      R[i] = r;
    }
    pd[0] += (double)r * (double)r;
  }
  flat_partials<K, 1>(pd, part, 0);
}

// partial a.b (k <= 4)
template <typename T, int K>
__global__ void __launch_bounds__(1024)
kf_dot(int32_t n, const T *__restrict__ A, const T *__restrict__ Bv, double *__restrict__ part) {
  constexpr int B = kFlat(K);
  const int64_t total = (int64_t)n * K, S = (int64_t)gridDim.x * B;
  double pd[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * B + threadIdx.x; i < total; i += S)
    pd[0] += (double)A[i] * (double)Bv[i];
  flat_partials<K, 1>(pd, part, 0);
}

template <typename T, int VEC, bool FUSED, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_rupd(int32_t n, const double *__restrict__ alpha, SegMap sm, int64_t k,
        const T *__restrict__ Q, T *__restrict__ R, double *__restrict__ part,
        const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  WIDE_COLS_P(VEC, LPR)
  double pd[1][VEC] = {};
  T al[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) al[m] = (FUSED && valid) ? (T)alpha[c0 + m] : T(0);
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      const int64_t o = (int64_t)row * k + c0;
      T rv[VEC];
      VecIO<T, VEC>::ld(R + o, rv);
      if (FUSED) {
        if (sm.of_row) {
          const int64_t so = sm.vc(Pack<LPR>::u(row), k, c0);
#pragma unroll
          for (int m = 0; m < VEC; ++m) al[m] = (T)alpha[so + m];
        }
        T qv[VEC];
        VecIO<T, VEC>::ld(Q + o, qv);
#pragma unroll
        for (int m = 0; m < VEC; ++m) rv[m] = fma(-al[m], qv[m], rv[m]);
        VecIO<T, VEC>::st(R + o, rv);
      }
#pragma unroll
      for (int m = 0; m < VEC; ++m) pd[0][m] += (double)rv[m] * (double)rv[m];
    }
  }
  block_partials_wide<VEC, 1, LPR>(pd, part, k, c0, 0);
}

// ---- Jacobi (SparseLinAlg.applyJacobi, zero diagonal -> 0): z = r / d,
//      (FUSED) r -= alpha q first; partials r.r and r.z -----------------------
template <typename T, int K, bool FUSED>
__global__ void __launch_bounds__(kBlock)
kn_jac(int32_t n, const T *__restrict__ dinv, const double *__restrict__ alpha,
       const T *__restrict__ Q, T *__restrict__ R, T *__restrict__ Z,
       double *__restrict__ part, int64_t stride, const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  const int g = threadIdx.x % kGroup;
  double pd[2] = {0.0, 0.0};
  const T al = (FUSED && g < K) ? (T)alpha[g] : T(0);
  NARROW_LOOP(row, 0, n) {
    if (g < K) {
      const int64_t o = row * K + g;
      T r = R[o];
      if (FUSED) {
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

template <typename T, int VEC, bool FUSED, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_jac(int32_t n, const T *__restrict__ dinv, const double *__restrict__ alpha,
       int64_t k, const T *__restrict__ Q, T *__restrict__ R, T *__restrict__ Z,
       double *__restrict__ part, int64_t stride, const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  WIDE_COLS_P(VEC, LPR)
  double pd[2][VEC] = {};
  T al[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) al[m] = (FUSED && valid) ? (T)alpha[c0 + m] : T(0);
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      const int64_t o = (int64_t)row * k + c0;
      T rv[VEC], zv[VEC];
      VecIO<T, VEC>::ld(R + o, rv);
      if (FUSED) {
        T qv[VEC];
        VecIO<T, VEC>::ld(Q + o, qv);
#pragma unroll
        for (int m = 0; m < VEC; ++m) rv[m] = fma(-al[m], qv[m], rv[m]);
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
  block_partials_wide<VEC, 2, LPR>(pd, part, k, c0, stride);
}

// ---- generic two-vector column dot: partial a.b -------------------------
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_dot(int32_t n, int64_t k, const T *__restrict__ A, const T *__restrict__ B,
       double *__restrict__ part) {
  WIDE_COLS_P(VEC, LPR)
  double pd[1][VEC] = {};
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      T a[VEC], b[VEC];
      VecIO<T, VEC>::ld(A + (int64_t)row * k + c0, a);
      VecIO<T, VEC>::ld(B + (int64_t)row * k + c0, b);
#pragma unroll
      for (int m = 0; m < VEC; ++m) pd[0][m] += (double)a[m] * (double)b[m];
    }
  }
  block_partials_wide<VEC, 1, LPR>(pd, part, k, c0, 0);
}

// partial b.b of FP64 b with b stored as TO into Y (kw_dot<double>'s sums):
// a solve from x0 = 0 takes ||b|| and its unscaled first residual in one
// pass over b
template <typename TO, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_dot_copy(int32_t n, int64_t k, const double *__restrict__ B, TO *__restrict__ Y,
            double *__restrict__ part) {
  WIDE_COLS_P(VEC, LPR)
  double pd[1][VEC] = {};
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      double b[VEC];
      TO y[VEC];
      VecIO<double, VEC>::ld(B + (int64_t)row * k + c0, b);
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        pd[0][m] += b[m] * b[m];
        y[m] = (TO)b[m];
      }
      VecIO<TO, VEC>::st(Y + (int64_t)row * k + c0, y);
    }
  }
  block_partials_wide<VEC, 1, LPR>(pd, part, k, c0, 0);
}

// ---- column reductions of the [rows][k] partial buffers (fixed order) ----
// v + rows b0, b0 + step, ... < b1 of column c, added in that order; the
// rows are loaded 8 at a time (clamped into range, the clamped ones added
// as 0), so 8 loads are in flight instead of one per add.
__device__ __forceinline__ double strided_sum(const double *__restrict__ part, int32_t b0,
                                              int32_t b1, int32_t step, int64_t k, int64_t c,
                                              double v) {
  if (b0 >= b1) return v;
  for (int32_t b = b0; b < b1; b += 8 * step) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = part[(int64_t)min(b + u * step, b1 - 1) * k + c];
#pragma unroll
    for (int u = 0; u < 8; ++u) v += b + u * step < b1 ? x[u] : 0.0;
  }
  return v;
}

__device__ __forceinline__ double col_reduce(const double *__restrict__ part,
                                             int32_t nrows, int64_t k, int64_t c,
                                             double (*sh)[64]) {
  const int s = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double v = 0.0;
  if (c < k) v = strided_sum(part, s, nrows, 4, k, c, 0.0);
  sh[s][lane] = v;
  __syncthreads();
  double t = sh[0][lane];
  t += sh[1][lane];
  t += sh[2][lane];
  t += sh[3][lane];
  __syncthreads();
  return t;
}

// First stage of the partial reduction: block (x, y) sums rows
// [y*chunk, (y+1)*chunk) of `np` quantities for 64 columns, in a fixed
// order, so that the second stage reads at most kRedRows rows.
constexpr int kRedRows = 64;
__global__ void __launch_bounds__(kBlock)
k_reduce_stage1(int32_t nrows, int32_t chunk, int64_t k, const double *__restrict__ part,
                int64_t stride, int np, double *__restrict__ out, int64_t ostride) {
  __shared__ double sh[4][64];
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int s = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int32_t b0 = blockIdx.y * chunk, b1 = min(nrows, b0 + chunk);
  for (int q = 0; q < np; ++q) {
    double v = 0.0;
    if (c < k) v = strided_sum(part + q * stride, b0 + s, b1, 4, k, c, 0.0);
    sh[s][lane] = v;
    __syncthreads();
    if (s == 0 && c < k) {
      double t = sh[0][lane];
      t += sh[1][lane];
      t += sh[2][lane];
      t += sh[3][lane];
      out[q * ostride + (int64_t)blockIdx.y * k + c] = t;
    }
    __syncthreads();
  }
}

// alpha = rz / p.q for active columns (p.q == 0 or non-finite: breakdown,
// SparseLinAlg.lean:118 -- the column stops)
__global__ void __launch_bounds__(kBlock)
k_reduce_alpha(int32_t nrows, int64_t k, const double *__restrict__ part,
               double *__restrict__ st, int32_t *__restrict__ active,
               int32_t *__restrict__ ctl) {
  const bool idle = ctl_idle(ctl);
  __syncthreads();  // every thread has read ctl before block 0 clears JUST
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl[CTL_JUST] = 0;
    if (!idle) ctl[CTL_ITERS] += 1;
  }
  if (idle) return;
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
        atomicSub(&ctl[CTL_ACTIVE], 1);
        atomicOr(&ctl[CTL_BREAKDOWN], 1);
      }
    }
    st[ST_ALPHA * k + c] = a;
  }
}

// After the forward pass: r.r and r.M^-1.r partials.
//   mode 0 (iteration): stop test rr <= tol2 or rr < floor2 * rref (the
//          FP32 recurrence at rounding noise), reliable-update trigger
//          rr < delta2 * rref, pause below delta2^2 * rref, beta = rz'/rz,
//          gamma = 1 (inactive: beta 1, gamma 0 -> p, d frozen)
//   mode 1 (after a residual replacement / start): rz = rz', rref = rr;
//          continuing columns keep p (beta 1, gamma 0), restarted ones take
//          p = z (beta 0, gamma 1)
//   mode 2: mode 1 with rref as the host set it (no r.r partials)
__global__ void __launch_bounds__(kBlock)
k_reduce_beta(int32_t nrows, int32_t nrows_rz, int64_t k, const double *__restrict__ prr,
              const double *__restrict__ prz, double *__restrict__ st,
              int32_t *__restrict__ active, const int32_t *__restrict__ restart,
              int32_t *__restrict__ ctl, int mode, double delta2, double floor2) {
  if (ctl_idle(ctl)) return;
  // an iteration has run its d update: the round's d holds it from now on
  if (mode == 0 && blockIdx.x == 0 && threadIdx.x == 0) ctl[CTL_DFRESH] = 0;
  __shared__ double sh[4][64];
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  double rr = col_reduce(prr, nrows, k, c, sh);
  double rz = col_reduce(prz, nrows_rz, k, c, sh);
  if (threadIdx.x < 64 && c < k) {
    double beta = 1.0, gamma = 0.0;
    if (active[c]) {
      if (mode >= 1) {
        st[ST_RZ * k + c] = rz;
        if (mode == 1) st[ST_RREF * k + c] = rr;
        if (restart[c]) { beta = 0.0; gamma = 1.0; }
      } else if (!(rr > st[ST_TOL2 * k + c]) || !(rr >= floor2 * st[ST_RREF * k + c]) ||
                 !isfinite(rr) || !isfinite(rz)) {
        active[c] = 0;
        atomicSub(&ctl[CTL_ACTIVE], 1);
        atomicAdd(&ctl[CTL_JUST], 1);
        if (!isfinite(rr) || !isfinite(rz)) atomicOr(&ctl[CTL_BREAKDOWN], 1);
      } else {
        double old = st[ST_RZ * k + c];
        beta = old == 0.0 ? 0.0 : rz / old;
        gamma = 1.0;
        st[ST_RZ * k + c] = rz;
        if (rr < delta2 * st[ST_RREF * k + c]) atomicOr(&ctl[CTL_TRIGGER], 1);
        if (rr < delta2 * delta2 * st[ST_RREF * k + c]) {
          atomicOr(&ctl[CTL_PAUSE], 1);
          atomicAdd(&ctl[CTL_JUST], 1);
        }
      }
    }
    st[ST_BETA * k + c] = beta;
    st[ST_GAMMA * k + c] = gamma;
  }
}

__global__ void __launch_bounds__(kBlock)
k_reduce_plain(int32_t nrows, int64_t k, const double *__restrict__ part,
               double *__restrict__ out) {
  __shared__ double sh[4][64];
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  double v = col_reduce(part, nrows, k, c, sh);
  if (threadIdx.x < 64 && c < k) out[c] = v;
}

// ---- per-segment dots of a batch -------------------------------------------
// Blocks [blk_ptr[s], blk_ptr[s+1]) of the grid walk the rows of segment s
// only, so every block's partial row belongs to one segment.
__device__ __forceinline__ RowIter seg_rows(const int32_t *__restrict__ seg_ptr,
                                            const int32_t *__restrict__ blk_ptr,
                                            const int32_t *__restrict__ blk_seg,
                                            int32_t slots, int32_t slot) {
  const int32_t b = blockIdx.x, sg = blk_seg[b];
  const int32_t lb = b - blk_ptr[sg], nb = blk_ptr[sg + 1] - blk_ptr[sg];
  return RowIter{seg_ptr[sg] + lb * slots + slot, seg_ptr[sg + 1], nb * slots};
}

// partials of a.b (and, NP = 2, of c.d at part + stride).  Flat form for
// k <= 4: the block's share of its segment's n_s x K range, walked lane
// after lane with a stride of (blocks of the segment) x kFlat(K).
template <typename TA, typename TB, int K, int NP>
__global__ void __launch_bounds__(1024)
kf_segdot(const int32_t *__restrict__ seg_ptr, const int32_t *__restrict__ blk_ptr,
          const int32_t *__restrict__ blk_seg, const TA *__restrict__ a,
          const TB *__restrict__ b, const TA *__restrict__ c, const TB *__restrict__ d,
          double *__restrict__ part, int64_t stride) {
  constexpr int B = kFlat(K);
  const int32_t sg = blk_seg[blockIdx.x];
  const int32_t lb = blockIdx.x - blk_ptr[sg], nb = blk_ptr[sg + 1] - blk_ptr[sg];
  const int64_t end = (int64_t)seg_ptr[sg + 1] * K, S = (int64_t)nb * B;
  double pd[NP] = {};
  for (int64_t i = (int64_t)seg_ptr[sg] * K + (int64_t)lb * B + threadIdx.x; i < end; i += S) {
    pd[0] += (double)a[i] * (double)b[i];
    if (NP == 2) pd[NP - 1] += (double)c[i] * (double)d[i];
  }
  flat_partials<K, NP>(pd, part, stride);
}

template <typename TA, typename TB, int NP>
__global__ void __launch_bounds__(kBlock)
kw_segdot(const int32_t *__restrict__ seg_ptr, const int32_t *__restrict__ blk_ptr,
          const int32_t *__restrict__ blk_seg, int64_t k, const TA *__restrict__ a,
          const TB *__restrict__ b, const TA *__restrict__ c, const TB *__restrict__ d,
          double *__restrict__ part, int64_t stride) {
  WIDE_COLS(1)
  double pd[NP][1] = {};
  const RowIter it = seg_rows(seg_ptr, blk_ptr, blk_seg, kWaves, threadIdx.x >> 6);
  for (int32_t row = it.first; row < it.end; row += it.stride)
    if (valid) {
      const int64_t o = (int64_t)row * k + c0;
      pd[0][0] += (double)a[o] * (double)b[o];
      if (NP == 2) pd[NP - 1][0] += (double)c[o] * (double)d[o];
    }
  block_partials_wide<1, NP>(pd, part, k, c0, stride);
}

// out[q * kk + s * k + c] = sum over the blocks of segment s of their
// partial (q, c): one wavefront per (segment, quantity), lane l takes
// blocks l, l + 64, ..., and the lane sums meet in a fixed butterfly
__global__ void __launch_bounds__(64)
k_segsum(int32_t nseg, const int32_t *__restrict__ blk_ptr, int64_t k,
         const double *__restrict__ part, int64_t stride, double *__restrict__ out) {
  const int32_t sg = blockIdx.x;
  const int q = blockIdx.y, lane = threadIdx.x;
  const int32_t b0 = blk_ptr[sg], b1 = blk_ptr[sg + 1];
  for (int64_t c = 0; c < k; ++c) {
    double t = strided_sum(part + q * stride, b0 + lane, b1, 64, k, c, 0.0);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) out[(int64_t)q * nseg * k + (int64_t)sg * k + c] = t;
  }
}

// Jacobi as a whole apply (batches): z = dinv r, zero diagonal -> 0
template <typename T>
__global__ void k_diag_apply(int64_t total, int64_t k, const T *__restrict__ dinv,
                             const T *__restrict__ R, T *__restrict__ Z,
                             const int32_t *__restrict__ ctl) {
  if (ctl_idle(ctl)) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    Z[i] = dinv[i / k] * R[i];
}

// ---- elementwise -----------------------------------------------------------
// plain path: P = Z + beta P
template <typename T>
__global__ void k_p_update(int64_t total, int64_t k, const double *__restrict__ beta,
                           const T *__restrict__ Z, T *__restrict__ P) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    P[i] = fma((T)beta[i % k], P[i], Z[i]);
}

// d += alpha p; p = gamma z + beta p (gamma 0 and beta 1 freeze a finished
// column).  Flat form for k <= 4: each thread on one column, the row tracked
// across the grid stride.
template <typename T, int K>
__global__ void __launch_bounds__(1024)
kf_pd_update(int32_t n, SegMap sm, const double *__restrict__ st, const T *__restrict__ Z,
             T *__restrict__ P, T *__restrict__ D, const int32_t *__restrict__ ctl) {
  if (bwd_idle(ctl)) return;
  constexpr int B = kFlat(K);
  const int c = threadIdx.x % K;
  const int64_t total = (int64_t)n * K, S = (int64_t)gridDim.x * B;
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x, row = i / K;
  T al = (T)st[ST_ALPHA * sm.kk + c], be = (T)st[ST_BETA * sm.kk + c],
    ga = (T)st[ST_GAMMA * sm.kk + c];
  for (; i < total; i += S, row += S / K) {
    if (sm.of_row) {
      const int64_t v = sm.vc(row, K, c);
      al = (T)st[ST_ALPHA * sm.kk + v];
      be = (T)st[ST_BETA * sm.kk + v];
      ga = (T)st[ST_GAMMA * sm.kk + v];
    }
    const T p = P[i];
    D[i] = fma(al, p, D[i]);
    P[i] = fma(be, p, ga * Z[i]);
  }
}

// Wide form of the update (same arithmetic) for k > 4: LPR lanes per row, VEC
// columns a lane, the column scalars held in registers across rows (per row
// only for a batch, whose scalars depend on the row's system).
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_pd_update(int32_t n, int64_t k, SegMap sm, const double *__restrict__ st,
             const T *__restrict__ Z, T *__restrict__ P, T *__restrict__ D,
             const int32_t *__restrict__ ctl) {
  if (bwd_idle(ctl)) return;
  WIDE_COLS_P(VEC, LPR)
  T al[VEC], be[VEC], ga[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) {
    al[m] = valid ? (T)st[ST_ALPHA * sm.kk + c0 + m] : T(0);
    be[m] = valid ? (T)st[ST_BETA * sm.kk + c0 + m] : T(0);
    ga[m] = valid ? (T)st[ST_GAMMA * sm.kk + c0 + m] : T(0);
  }
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      if (sm.of_row) {
        const int64_t so = sm.vc(Pack<LPR>::u(row), k, c0);
#pragma unroll
        for (int m = 0; m < VEC; ++m) {
          al[m] = (T)st[ST_ALPHA * sm.kk + so + m];
          be[m] = (T)st[ST_BETA * sm.kk + so + m];
          ga[m] = (T)st[ST_GAMMA * sm.kk + so + m];
        }
      }
      const int64_t o = (int64_t)row * k + c0;
      T pv[VEC], dv[VEC], zv[VEC];
      VecIO<T, VEC>::ld(P + o, pv);
      VecIO<T, VEC>::ld(D + o, dv);
      VecIO<T, VEC>::ld(Z + o, zv);
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        dv[m] = fma(al[m], pv[m], dv[m]);
        pv[m] = fma(be[m], pv[m], ga[m] * zv[m]);
      }
      VecIO<T, VEC>::st(D + o, dv);
      VecIO<T, VEC>::st(P + o, pv);
    }
  }
}

// Y = (T)(X * s_c)
template <typename T>
__global__ void k_scale_in(int64_t total, int64_t k, SegMap sm, const double *__restrict__ s,
                           const double *__restrict__ X, T *__restrict__ Y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / k;
    Y[i] = (T)(X[i] * s[sm.vc(row, k, i - row * k)]);
  }
}

// X += invs_c * D; D = 0.  init: X is not read (the first update of a
// solve from x0 = 0, whose X was never written), the same fma from 0.
template <typename T>
// dfresh != nullptr (the fused iteration): d is left as it is, and a round
// whose d holds nothing yet (*dfresh != 0) adds none.
__global__ void k_x_update(int64_t total, int64_t k, SegMap sm, const double *__restrict__ invs,
                           T *__restrict__ D, double *__restrict__ X, int init,
                           const int32_t *__restrict__ dfresh) {
  const bool skip = dfresh && *dfresh != 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / k;
    const double d = skip ? 0.0 : (double)D[i];
    X[i] = fma(invs[sm.vc(row, k, i - row * k)], d, init ? 0.0 : X[i]);
    if (!dfresh) D[i] = T(0);
  }
}

// Wide forms of k_scale_in and k_x_update (same arithmetic), mapped like
// kw_pd_update.
template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_scale_in(int32_t n, int64_t k, SegMap sm, const double *__restrict__ sc,
            const double *__restrict__ X, T *__restrict__ Y) {
  WIDE_COLS_P(VEC, LPR)
  double s[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) s[m] = valid ? sc[c0 + m] : 0.0;
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      if (sm.of_row) {
        const int64_t so = sm.vc(Pack<LPR>::u(row), k, c0);
#pragma unroll
        for (int m = 0; m < VEC; ++m) s[m] = sc[so + m];
      }
      const int64_t o = (int64_t)row * k + c0;
      double xv[VEC];
      T yv[VEC];
      VecIO<double, VEC>::ld(X + o, xv);
#pragma unroll
      for (int m = 0; m < VEC; ++m) yv[m] = (T)(xv[m] * s[m]);
      VecIO<T, VEC>::st(Y + o, yv);
    }
  }
}

template <typename T, int VEC, int LPR>
__global__ void __launch_bounds__(kBlock)
kw_x_update(int32_t n, int64_t k, SegMap sm, const double *__restrict__ invs,
            T *__restrict__ D, double *__restrict__ X, int init,
            const int32_t *__restrict__ dfresh) {
  const bool skip = dfresh && *dfresh != 0;
  WIDE_COLS_P(VEC, LPR)
  double iv[VEC];
#pragma unroll
  for (int m = 0; m < VEC; ++m) iv[m] = valid ? invs[c0 + m] : 0.0;
  WIDE_LOOP_P(row, 0, n, LPR) {
    if (valid) {
      if (sm.of_row) {
        const int64_t so = sm.vc(Pack<LPR>::u(row), k, c0);
#pragma unroll
        for (int m = 0; m < VEC; ++m) iv[m] = invs[so + m];
      }
      const int64_t o = (int64_t)row * k + c0;
      double xv[VEC];
      T dv[VEC];
      if (init) {
#pragma unroll
        for (int m = 0; m < VEC; ++m) xv[m] = 0.0;
      } else {
        VecIO<double, VEC>::ld(X + o, xv);
      }
      if (skip) {
#pragma unroll
        for (int m = 0; m < VEC; ++m) dv[m] = T(0);
      } else {
        VecIO<T, VEC>::ld(D + o, dv);
      }
#pragma unroll
      for (int m = 0; m < VEC; ++m) {
        xv[m] = fma(iv[m], (double)dv[m], xv[m]);
        dv[m] = T(0);
      }
      VecIO<double, VEC>::st(X + o, xv);
      if (!dfresh) VecIO<T, VEC>::st(D + o, dv);
    }
  }
}

template <typename T>
__global__ void k_axpy(int64_t total, int64_t k, const double *__restrict__ a,
                       double sign, const T *__restrict__ x, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + (T)(sign * a[i % k]) * x[i];
}

__global__ void k_dinv(int32_t n, const double *__restrict__ d, double *__restrict__ d64,
                       float *__restrict__ d32) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = d[i] == 0.0 ? 0.0 : 1.0 / d[i];
  d64[i] = v;
  d32[i] = (float)v;
}

static uint64_t bits_of(double v) {
  uint64_t u;
  std::memcpy(&u, &v, sizeof u);
  return u;
}

// ============================================================== profiling ==
thread_local Prof g_prof;

void prof_reset() {
  g_prof.recs.clear();
  g_prof.used = 0;
  g_prof.result.clear();
}

void prof_collect() {
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
  int64_t nvc = 0;                    // scalar columns (nseg x k for a batch)
  DevBuf<double> seg;                 // 2 x nvc per-segment dot sums
  DevBuf<double> xp, rp;              // FP64 n x k: x, true residual
  DevBuf<double> buf[5];              // inner D, R, Z, P, Q (8 bytes/elem)
  DevBuf<double> part;                // 3 x part_rows x k
  DevBuf<double> red;                 // 2 x kRedRows x k, first-stage sums
  DevBuf<double> st;                  // ST_N x nvc per-column scalars
  // the control block and the per-column flags in one allocation,
  // [ctl: CTL_N | active: nvc + 4 | restart: nvc + 4], so a round uploads
  // them in one copy from the pinned mirror h_flags (same layout)
  struct View32 {
    int32_t *p = nullptr;
    int32_t *get() const { return p; }
  };
  DevBuf<int32_t> flags;
  View32 ctl, active, restart;
  DevBuf<double> dinv64;
  DevBuf<float> dinv32;
  int32_t *h_flags = nullptr;         // pinned, the layout of flags
  int32_t *h_ctl = nullptr;           // = h_flags
  int32_t *h_dact = nullptr;          // pinned, nvc + 4: active read back
  double *h_st = nullptr;             // pinned, ST_N x nvc
  // the captured block of inner iterations (fused solve), replayed while
  // the key -- everything its launches depend on -- is unchanged (a few
  // keys kept); captured on its own stream (a graph cannot be captured on
  // the legacy stream)
  hipStream_t cap = nullptr;
  std::vector<std::pair<std::vector<uint64_t>, hipGraphExec_t>> graphs;
  ~SolverWork() {
    for (auto &e : graphs) (void)hipGraphExecDestroy(e.second);
    if (cap) (void)hipStreamDestroy(cap);
    if (h_flags) (void)hipHostFree(h_flags);
    if (h_dact) (void)hipHostFree(h_dact);
    if (h_st) (void)hipHostFree(h_st);
  }
};

void free_solver_work(SolverWork *w) { delete w; }

static SolverWork *ensure_work(cn_system_t *S, int32_t k, int32_t part_rows) {
  SolverWork *W = S->work;
  const int32_t n = S->n_free;
  const int64_t nvc_need = (int64_t)std::max(S->nseg, 1) * k;
  if (W && W->n == n && W->k >= k && W->part_rows >= part_rows && W->nvc >= nvc_need)
    return W;
  int32_t kk = std::max(k, W ? W->k : 0);
  int32_t pr = std::max(part_rows, W ? W->part_rows : 0);
  const int64_t nvc = std::max<int64_t>(std::max<int64_t>(nvc_need, kk), W ? W->nvc : 0);
  delete W;
  S->work = nullptr;
  W = new SolverWork();
  W->n = n;
  W->k = kk;
  W->part_rows = pr;
  W->nvc = nvc;
  W->seg.alloc((size_t)2 * nvc + 4);
  size_t nk = (size_t)n * kk + 4;
  W->xp.alloc(nk);
  W->rp.alloc(nk);
  for (auto &b : W->buf) b.alloc(nk);
  W->part.alloc((size_t)3 * pr * kk + 4);
  W->red.alloc((size_t)2 * kRedRows * kk + 4);
  W->st.alloc((size_t)ST_N * nvc + 4);
  const size_t nflags = CTL_N + 2 * ((size_t)nvc + 4);
  W->flags.alloc(nflags);
  W->ctl.p = W->flags.get();
  W->active.p = W->ctl.p + CTL_N;
  W->restart.p = W->active.p + nvc + 4;
  W->dinv64.alloc((size_t)n + 1);
  W->dinv32.alloc((size_t)n + 1);
  CN_HIP(hipHostMalloc((void **)&W->h_flags, sizeof(int32_t) * nflags, hipHostMallocDefault));
  W->h_ctl = W->h_flags;
  CN_HIP(hipHostMalloc((void **)&W->h_dact, sizeof(int32_t) * ((size_t)nvc + 4),
                       hipHostMallocDefault));
  CN_HIP(hipHostMalloc((void **)&W->h_st, sizeof(double) * ((size_t)ST_N * nvc + 4),
                       hipHostMallocDefault));
  S->work = W;
  return W;
}

// ================================================================ launches ==
// blocks for a row kernel: multiples of 8 (one share per XCD) once there
// are enough rows, capped
static unsigned blocks_for_rows(int64_t rows, int rows_per_block, unsigned cap) {
  int64_t g = (rows + rows_per_block - 1) / rows_per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  if (g >= 64) g -= g % 8;
  return (unsigned)g;
}

struct Grid {
  dim3 g;
  int32_t rows;  // partial rows written (= g.x)
  int lpr = 64;  // lanes per row of the wide mapping
};

template <typename T>
static int wide_vec(int64_t k) {
  if (sizeof(T) == 4) return (k % 4 == 0) ? 4 : ((k % 2 == 0) ? 2 : 1);
  return (k % 2 == 0) ? 2 : 1;
}

// Blocks of kBlock threads the device holds at once (8 per CU).  A row
// sweep whose grid exceeds this runs its last column tiles as a second,
// half-empty wave of blocks; capping the grid at one full wave avoids it.
static unsigned resident_blocks() {
  static unsigned cached = 0;
  if (cached == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached = (unsigned)cus * 8;
  }
  return cached;
}

// lanes per row of the packed wide mapping
template <typename T>
static int wide_lpr(int64_t k) {
  return pack_lpr12((k + wide_vec<T>(k) - 1) / wide_vec<T>(k));
}

template <typename T>
static Grid grid_rows(int64_t rows, int64_t k, unsigned cap) {
  Grid G;
  if (k <= 4) {
    G.g = dim3(blocks_for_rows(rows, kBlock / kGroup, cap), 1, 1);
  } else {
    const int v = wide_vec<T>(k), l = wide_lpr<T>(k);
    const unsigned tiles = (unsigned)((k + (int64_t)l * v - 1) / ((int64_t)l * v));
    const unsigned fit = std::max(8u, resident_blocks() / tiles);
    G.g = dim3(blocks_for_rows(rows, kWaves * (64 / l), std::min(cap, fit)), tiles, 1);
    G.lpr = l;
  }
  G.rows = (int32_t)G.g.x;
  return G;
}

// Column-tile width for the IC(0) passes.  Multi-row blocks are walked
// sequentially, so a level needs enough (block, tile) wavefronts to hide the
// latency of that walk: narrow the tiles until there are ~4096 of them.
template <typename T>
static int sweep_vec(int64_t k, int64_t nblocks, bool multi) {
  int v = wide_vec<T>(k);
  if (!multi) return v;
  while (v > 1 && nblocks * ((k + 64 * v - 1) / (64 * v)) < 4096) v /= 2;
  return v;
}

template <typename T>
static Grid grid_vec(int64_t rows, int64_t k, int v, unsigned cap) {
  Grid G;
  G.g = dim3(blocks_for_rows(rows, kWaves, cap), (unsigned)((k + 64 * v - 1) / (64 * v)), 1);
  G.rows = (int32_t)G.g.x;
  return G;
}

static const unsigned kCapSpmm = 1024;
static const unsigned kCapLevel = 512;

// dispatch: NARROW(K) for k <= 4, WIDE(V, L) with V the vector width and
// L the lanes per row
#define WIDE_LPR(lpr, WIDE, V)                                                \
  do {                                                                        \
    if ((lpr) == 12) { WIDE(V, 12); }                                         \
    else if ((lpr) == 16) { WIDE(V, 16); }                                    \
    else if ((lpr) == 32) { WIDE(V, 32); }                                    \
    else { WIDE(V, 64); }                                                     \
  } while (0)
#define DISPATCH_K(k, T, lpr, NARROW, WIDE)                                   \
  do {                                                                        \
    if ((k) <= 4) {                                                           \
      switch (k) {                                                            \
        case 1: NARROW(1); break;                                             \
        case 2: NARROW(2); break;                                             \
        case 3: NARROW(3); break;                                             \
        default: NARROW(4); break;                                            \
      }                                                                       \
    } else {                                                                  \
      int v_ = wide_vec<T>(k);                                                \
      if (v_ == 4) { WIDE_LPR(lpr, WIDE, 4); }                                \
      else if (v_ == 2) { WIDE_LPR(lpr, WIDE, 2); }                           \
      else { WIDE_LPR(lpr, WIDE, 1); }                                        \
    }                                                                         \
  } while (0)

// the FP64 path never selects V = 4; map it to an existing instantiation
template <typename T, int V> struct WideV { static constexpr int v = V; };
template <> struct WideV<double, 4> { static constexpr int v = 2; };

// Matrix + preconditioner view of one solve.
template <typename T>
struct SysView {
  int32_t n;
  const int32_t *ptr, *col;
  const T *val;
  const double *val64;
  const cn_ic0_t *ic0 = nullptr;
  const int32_t *rowmap, *blk, *lp, *lcol, *up, *ucol;
  const T *lval, *uval, *rdiag;
  const T *dinv;
  cn_hsc_t *hsc = nullptr;  // whole-apply preconditioners (replace ic0 / dinv)
  const cn_sgs_t *sgs = nullptr;
  // per-(segment, column) scalars of a batch (of_row = nullptr: one system)
  SegMap sm{nullptr, 0};
  // row table of A_ff (SpMM); ovp = nullptr when no row overflows it
  const int32_t *ecol = nullptr, *ovp = nullptr, *ovc = nullptr;
  const T *eval = nullptr, *ovv = nullptr;
  // window row table of the LDS-staged SpMM; rpr = 0 when not in use
  int32_t rpr = 0;
  int32_t gens = 4;  // row stripes per XCD share of the row-table SpMM
  HscShape hsh{};    // HSC apply launch shapes
  const unsigned char *wpk = nullptr;
  const int32_t *far_ptr = nullptr, *far_rows = nullptr;
  // block sweeps (forward: L, backward: U); mb = 0 when not available
  int32_t mb = 0;
  SweepView<T> lsw{}, usw{};
};

// Row stripes per XCD share in the wide row-table SpMM (1, 2 or 4; 4 by
// default): the grid holds this many generations of resident blocks, so the
// rows in flight on an XCD stay within one stripe and its L2 serves the
// neighbour gathers (measured: L2 misses 705 MB -> 431 MB per launch on the
// 100k capsule mesh at k = 768).  4 is also the most the partial-sum
// buffers are sized for.
static const unsigned kSpmmGens = 4;

template <typename T>
static Grid grid_spmm(int64_t n, int64_t k, unsigned gens = kSpmmGens) {
  if (k <= 4) {  // one thread per row
    Grid G;
    G.g = dim3(blocks_for_rows(n, kBlock, kCapSpmm), 1, 1);
    G.rows = (int32_t)G.g.x;
    return G;
  }
  Grid G = grid_rows<T>(n, k, kCapSpmm);
  if (k > 4 && G.g.x % kXcd == 0) {
    G.g.x *= std::max(1u, std::min(gens, kSpmmGens));
    G.rows = (int32_t)G.g.x;
  }
  return G;
}

// Rows per range of the LDS-staged SpMM: one block per CU over the
// (range, column tile) grid, ranges a whole number of 32-row chunks.
template <typename T>
static int32_t win_rpr(int64_t n, int64_t k) {
  const int64_t W = 256 / sizeof(T);
  const int64_t tiles = (k + W - 1) / W;
  const int64_t cus = std::max<int64_t>(1, resident_blocks() / 8);
  const int64_t R = std::max<int64_t>(1, cus / tiles);
  const int64_t per = (n + R - 1) / R;
  return (int32_t)std::max<int64_t>(kLdsRows, (per + kLdsRows - 1) / kLdsRows * kLdsRows);
}

// the LDS-staged SpMM applies to wide k whose rows are 16-byte aligned
template <typename T>
static bool win_applies(int64_t n, int64_t k) {
  const int vec = 16 / (int)sizeof(T);
  return k > 4 && k % vec == 0 && n >= 4 * kLdsRows;
}

template <typename T>
static int32_t launch_spmm_dot(const SysView<T> &V, int64_t k, const T *P, T *Q,
                               double *part, const int32_t *ctl, hipStream_t s) {
  if (V.rpr > 0 && win_applies<T>(V.n, k)) {
    const int64_t W = 256 / sizeof(T);
    const int64_t R = (V.n + V.rpr - 1) / V.rpr, pairs = R * ((k + W - 1) / W);
    const dim3 g((unsigned)(kXcd * ((pairs + kXcd - 1) / kXcd)));
    ProfScope ps("spmm_lds", s);
    kl_spmm<T, 6, KL_PLAIN><<<g, 512, 0, s>>>(V.n, V.rpr, V.wpk, V.far_ptr, V.far_rows, k, P,
                                              nullptr, nullptr, Q, nullptr, nullptr, nullptr,
                                              part, ctl);
    return (int32_t)R;
  }
  Grid G = grid_spmm<T>(V.n, k, (unsigned)V.gens);
  const int32_t gens = (int32_t)(G.g.x / std::max(1u, grid_rows<T>(V.n, k, kCapSpmm).g.x));
  ProfScope ps("spmm", s);
#define N_(KK) kn_spmm_dot<T, KK><<<G.g, kBlock, 0, s>>>(V.n, V.ecol, V.eval, V.ovp, V.ovc, V.ovv, P, Q, part, ctl)
#define W_(VV, LL) kw_spmm_dot<T, WideV<T, VV>::v, LL><<<G.g, kBlock, 0, s>>>(V.n, V.ecol, V.eval, V.ovp, V.ovc, V.ovv, k, P, Q, part, ctl, gens)
  DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
  return G.rows;
}

// The direction update + SpMM (KL_DIR: Pin -> Pout, partial p.q) and the
// SpMM + residual and iterate update (KL_UPD: partial r.r) of the fused
// iteration; partial rows written = row ranges.
template <typename T, int MODE>
static int32_t launch_kl(const SysView<T> &V, int64_t k, const T *P, const T *Z, T *Pout,
                         T *R, T *D, const double *st, double *part, const int32_t *ctl,
                         hipStream_t s) {
  const int64_t W = 256 / sizeof(T);
  const int64_t nr = (V.n + V.rpr - 1) / V.rpr, pairs = nr * ((k + W - 1) / W);
  const dim3 g((unsigned)(kXcd * ((pairs + kXcd - 1) / kXcd)));
  ProfScope ps(MODE == KL_DIR ? "spmm_dir" : "spmm_upd", s);
  kl_spmm<T, 6, MODE><<<g, 512, 0, s>>>(V.n, V.rpr, V.wpk, V.far_ptr, V.far_rows, k, P, Z,
                                        Pout, nullptr, R, D, st, part, ctl);
  return (int32_t)nr;
}

// The x update and FP64 residual of a replacement with the LDS-staged
// product, over the system's window (the FP32 product's row ranges, FP64
// weights): Xout = X + d / s (X not read when xinit, d not when *dfresh),
// R = (b - A Xout) s (FP32), partial r.r a range.  Returns the partial rows.
static int32_t launch_resid_kl(const cn_system_t *S, int32_t n, int32_t rpr, int64_t k,
                               const double *X, const double *B, float *R,
                               const double *st, double *part, hipStream_t s,
                               const float *D, double *Xout, const int32_t *dfresh,
                               bool xinit) {
  const int64_t W = 256 / sizeof(double);
  const int64_t nr = (n + rpr - 1) / rpr, pairs = nr * ((k + W - 1) / W);
  const dim3 g((unsigned)(kXcd * ((pairs + kXcd - 1) / kXcd)));
  ProfScope ps("resid", s);
  kl_spmm<double, 6, KL_RES><<<g, 512, 0, s>>>(n, rpr, S->win_pk64.get(), S->win_far_ptr.get(),
                                               S->win_far_rows.get(), k, X, B, Xout,
                                               nullptr, nullptr, nullptr, st, part, nullptr,
                                               R, D, dfresh, xinit ? 1 : 0);
  return (int32_t)nr;
}

template <typename TO>
static int32_t launch_resid(int32_t n, const int32_t *ptr, const int32_t *col,
                            const double *val, int64_t k, const double *X,
                            const double *B, TO *R, const double *scale, double *part,
                            hipStream_t s) {
  Grid G = grid_rows<double>(n, k, kCapSpmm);
  ProfScope ps("resid", s);
  const unsigned gt = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, G.g.x);
#define N_(KK) kt_resid<KK, TO><<<std::max(1u, gt), kBlock, 0, s>>>(n, ptr, col, val, X, B, R, scale, part)
#define W_(VV, LL) kw_resid<WideV<double, VV>::v, TO, LL><<<G.g, kBlock, 0, s>>>(n, ptr, col, val, k, X, B, R, scale, part)
  DISPATCH_K(k, double, G.lpr, N_, W_);
#undef N_
#undef W_
  return k <= 4 ? (int32_t)std::max(1u, gt) : G.rows;
}

template <typename T>
void hsc_apply(cn_hsc_t *H, int64_t k, const T *r0, T *z0, hipStream_t s, const int32_t *ctl,
               const HscShape &sh, unsigned char *ext_work = nullptr,
               double *rz_part = nullptr);
template <typename T>
int32_t hsc_rz_rows(const cn_hsc_t *H, int64_t k, const HscShape &sh);
template <typename T>
void sgs_apply(const cn_sgs_t *G, int64_t k, const T *r, T *z, hipStream_t s,
               const int32_t *ctl);
template <typename T>
static int32_t launch_dot(int32_t n, int64_t k, const T *A, const T *B, double *part,
                          hipStream_t s);

// How one level of an IC(0) sweep is launched.
enum SweepKind { SW_NARROW, SW_THREAD, SW_WIDE, SW_BLOCK };
struct LevelGrid {
  SweepKind kind;
  dim3 g;
  int v;        // vector width (SW_WIDE, SW_BLOCK)
  int32_t rows; // partial rows written by the forward pass
};

// Vector width of a block sweep: the widest that fits k without leaving the
// tile mostly empty.
template <typename T>
static int block_vec(int64_t k) {
  int v = wide_vec<T>(k);
  while (v > 1 && 32 * v >= k) v /= 2;
  return v;
}

// (vector width, dense block size) dispatch of the block sweeps
#define KB_MB(v, MBV, CALL)                                                   \
  do {                                                                        \
    if ((v) == 4) { CALL(4, MBV); }                                           \
    else if ((v) == 2) { CALL(2, MBV); }                                      \
    else { CALL(1, MBV); }                                                    \
  } while (0)
#define KB_DISPATCH(v, mb, CALL)                                              \
  do {                                                                        \
    if ((mb) == 8) KB_MB(v, 8, CALL);                                         \
    else if ((mb) == 16) KB_MB(v, 16, CALL);                                  \
    else KB_MB(v, 32, CALL);                                                  \
  } while (0)

// Workgroups of one block-sweep kernel resident per CU (registers and
// shared memory), cached per kernel.
static unsigned wg_per_cu(const void *fn) {
  static std::map<const void *, unsigned> cache;
  auto it = cache.find(fn);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kBlock, 0) != hipSuccess || nb <= 0)
    nb = 1;
  (void)hipGetLastError();
  cache[fn] = (unsigned)nb;
  return (unsigned)nb;
}

template <typename T>
static unsigned kb_per_cu(int v, int mb, bool bwd) {
  const void *fn = nullptr;
#define F_(VV, MM)                                                                     \
  fn = bwd ? reinterpret_cast<const void *>(&kb_bwd<T, WideV<T, VV>::v, MM>)         \
           : reinterpret_cast<const void *>(&kb_fwd<T, WideV<T, VV>::v, MM, true>)
  KB_DISPATCH(v, mb, F_);
#undef F_
  return wg_per_cu(fn);
}

template <typename T>
static LevelGrid level_grid(const SysView<T> &V, int64_t k, int32_t nb, bool bwd) {
  LevelGrid L{};
  const bool multi = V.ic0->max_block > 1;
  if (k <= 4) {
    if (multi) {
      L.kind = SW_THREAD;
      L.g = dim3(blocks_for_rows(nb, kBlock, kCapLevel));
    } else {
      L.kind = SW_NARROW;
      L.g = dim3(blocks_for_rows(nb, kBlock / kGroup, kCapLevel));
    }
    L.rows = (int32_t)L.g.x;
    return L;
  }
  if (multi && V.mb > 0) {
    const int v = block_vec<T>(k);
    L.kind = SW_BLOCK;
    L.v = v;
    const unsigned tiles = (unsigned)((k + 64 * v - 1) / (64 * v));
    const unsigned per_cu = kb_per_cu<T>(v, V.mb, bwd);
    const unsigned fit = std::max(1u, (resident_blocks() / 8) * per_cu / tiles);
    // every workgroup takes the same number of blocks
    const int64_t iters = (nb + fit - 1) / fit;
    int64_t x = (nb + iters - 1) / iters;
    if (x >= 64) x = (x + 7) / 8 * 8;
    L.g = dim3((unsigned)x, tiles, 1);
    L.rows = (int32_t)x;
    return L;
  }
  L.kind = SW_WIDE;
  L.v = sweep_vec<T>(k, nb, multi);
  Grid G = grid_vec<T>(nb, k, L.v, kCapLevel);
  L.g = G.g;
  L.rows = G.rows;
  return L;
}

// First half of z = M^-1 r: (FUSED: r -= alpha q) then y = L^-1 r (IC(0))
// or z = r / d (Jacobi); partials r.r -> part[0..rows) and r.M^-1.r ->
// part[stride..].  Returns the partial rows written.
// no_rupd: the residual update and its r.r partials were done already (the
// LDS-staged KL_UPD product); only the preconditioner and r.z run, and the
// return value is the r.z partial rows alone.  rz_rows != nullptr: the r.z
// partials may have a row count of their own, returned there.  In both cases
// the hierarchical apply forms r.z from its own passes (r . M^-1 r = sum over
// the eliminated rows of d_f (L r)_f^2 plus the coarse rows' r_c z_c)
// instead of reading r and z again.
template <typename T, bool FUSED>
static int32_t launch_fwd(const SysView<T> &V, int64_t k, const double *alpha,
                          const T *Q, T *R, T *Z, double *part, int64_t stride,
                          const int32_t *ctl, hipStream_t s, bool no_rupd = false,
                          int32_t *rz_rows = nullptr) {
  if (rz_rows) *rz_rows = -1;  // -1: the r.z partials have the returned rows
  if (V.hsc || V.sgs || V.sm.of_row) {
    // (no scope of its own: each event pair around a launch group leaves a
    // few microseconds idle, so only the parts inside are timed)
    Grid G = grid_rows<T>(V.n, k, kCapSpmm);
#define N_(KK) kf_rupd<T, KK, FUSED><<<G.g, kFlat(KK), 0, s>>>(V.n, alpha, V.sm, Q, R, part, ctl)
#define W_(VV, LL) kw_rupd<T, WideV<T, VV>::v, FUSED, LL><<<G.g, kBlock, 0, s>>>(V.n, alpha, V.sm, k, Q, R, part, ctl)
    if (!no_rupd) DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
    if (V.hsc) {
      const int32_t er =
          (no_rupd || rz_rows) && !V.sm.of_row ? hsc_rz_rows<T>(V.hsc, k, V.hsh) : -1;
      hsc_apply<T>(V.hsc, k, R, Z, s, ctl, V.hsh, nullptr, er >= 0 ? part + stride : nullptr);
      if (er >= 0) {
        if (no_rupd) return er;
        *rz_rows = er;
        return G.rows;
      }
    } else if (V.sgs) {
      sgs_apply<T>(V.sgs, k, R, Z, s, ctl);
    } else {
      const int64_t total = (int64_t)V.n * k;
      k_diag_apply<T><<<(unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192)),
                        256, 0, s>>>(total, k, V.dinv, R, Z, ctl);
    }
    // a batch takes its per-segment dots separately (launch_segdot)
    if (!V.sm.of_row) {
      ProfScope pd("dot", s);
      launch_dot<T>(V.n, k, R, Z, part + stride, s);
    }
    return G.rows;
  }
  if (!V.ic0) {
    Grid G = grid_rows<T>(V.n, k, kCapSpmm);
    ProfScope ps("fwd", s);
#define N_(KK) kn_jac<T, KK, FUSED><<<G.g, kBlock, 0, s>>>(V.n, V.dinv, alpha, Q, R, Z, part, stride, ctl)
#define W_(VV, LL) kw_jac<T, WideV<T, VV>::v, FUSED, LL><<<G.g, kBlock, 0, s>>>(V.n, V.dinv, alpha, k, Q, R, Z, part, stride, ctl)
    DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
    return G.rows;
  }
  const std::vector<int32_t> &lb = V.ic0->level_blk;
  const int32_t nl = (int32_t)lb.size() - 1;
  int32_t off = 0;
  ProfScope ps("fwd", s);
  for (int32_t c = 0; c < nl; ++c) {
    int32_t b0 = lb[c], b1 = lb[c + 1];
    if (b1 <= b0) continue;
    double *pp = part + (int64_t)off * k;
    const LevelGrid L = level_grid<T>(V, k, b1 - b0, false);
    switch (L.kind) {
      case SW_THREAD:
#define N_(KK) kt_fwd<T, KK, FUSED><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, V.lp, V.lcol, V.lval, V.rdiag, alpha, Q, R, Z, pp, stride, ctl)
        switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
        break;
      case SW_NARROW:
#define N_(KK) kn_fwd<T, KK, FUSED><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, V.lp, V.lcol, V.lval, V.rdiag, alpha, Q, R, Z, pp, stride, ctl)
        switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
        break;
      case SW_BLOCK:
#define B_(VV, MM) kb_fwd<T, WideV<T, VV>::v, MM, FUSED><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, SWEEP_ARGS(V.lsw), alpha, k, Q, R, Z, pp, stride, ctl)
        KB_DISPATCH(L.v, V.mb, B_);
#undef B_
        break;
      case SW_WIDE:
#define W_(VV) kw_fwd<T, WideV<T, VV>::v, FUSED><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, V.lp, V.lcol, V.lval, V.rdiag, alpha, k, Q, R, Z, pp, stride, ctl)
        if (L.v == 4) { W_(4); } else if (L.v == 2) { W_(2); } else { W_(1); }
#undef W_
        break;
    }
    off += L.rows;
  }
  return off;
}

// Second half: z = L^-T y (IC(0)); d += alpha p; p = gamma z + beta p.
template <typename T>
static void launch_bwd(const SysView<T> &V, int64_t k, const double *st, T *Z, T *P,
                       T *D, const int32_t *ctl, hipStream_t s) {
  if (!V.ic0) {
    ProfScope ps("bwd", s);
    const Grid G = grid_rows<T>(V.n, k, kCapSpmm);
#define N_(KK) kf_pd_update<T, KK><<<G.g, kFlat(KK), 0, s>>>(V.n, V.sm, st, Z, P, D, ctl)
#define W_(VV, LL) kw_pd_update<T, WideV<T, VV>::v, LL><<<G.g, kBlock, 0, s>>>(V.n, k, V.sm, st, Z, P, D, ctl)
    DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
    return;
  }
  const std::vector<int32_t> &lb = V.ic0->level_blk;
  const int32_t nl = (int32_t)lb.size() - 1;
  ProfScope ps("bwd", s);
  for (int32_t c = nl - 1; c >= 0; --c) {
    int32_t b0 = lb[c], b1 = lb[c + 1];
    if (b1 <= b0) continue;
    const LevelGrid L = level_grid<T>(V, k, b1 - b0, true);
    switch (L.kind) {
      case SW_THREAD:
#define N_(KK) kt_bwd<T, KK><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, V.up, V.ucol, V.uval, V.rdiag, st, k, Z, P, D, ctl)
        switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
        break;
      case SW_NARROW:
#define N_(KK) kn_bwd<T, KK><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, V.up, V.ucol, V.uval, V.rdiag, st, k, Z, P, D, ctl)
        switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
        break;
      case SW_BLOCK:
#define B_(VV, MM) kb_bwd<T, WideV<T, VV>::v, MM><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, SWEEP_ARGS(V.usw), st, k, Z, P, D, ctl)
        KB_DISPATCH(L.v, V.mb, B_);
#undef B_
        break;
      case SW_WIDE:
#define W_(VV) kw_bwd<T, WideV<T, VV>::v><<<L.g, kBlock, 0, s>>>(b0, b1, V.blk, V.rowmap, V.up, V.ucol, V.uval, V.rdiag, st, k, Z, P, D, ctl)
        if (L.v == 4) { W_(4); } else if (L.v == 2) { W_(2); } else { W_(1); }
#undef W_
        break;
    }
  }
}

// Partial rows the forward pass writes (sizes the partials buffer).
template <typename T>
static int32_t fwd_rows(const SysView<T> &V, int64_t k) {
  if (!V.ic0) return grid_rows<T>(V.n, k, kCapSpmm).rows;
  const std::vector<int32_t> &lb = V.ic0->level_blk;
  int32_t off = 0;
  for (size_t c = 0; c + 1 < lb.size(); ++c) {
    int32_t nb = lb[c + 1] - lb[c];
    if (nb <= 0) continue;
    off += level_grid<T>(V, k, nb, false).rows;
  }
  return off;
}

template <typename T>
static int32_t launch_dot(int32_t n, int64_t k, const T *A, const T *B, double *part,
                          hipStream_t s) {
  Grid G = grid_rows<T>(n, k, kCapSpmm);
#define N_(KK) kf_dot<T, KK><<<G.g, kFlat(KK), 0, s>>>(n, A, B, part)
#define W_(VV, LL) kw_dot<T, WideV<T, VV>::v, LL><<<G.g, kBlock, 0, s>>>(n, k, A, B, part)
  DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
  return G.rows;
}

// k > 4 only (the wide mapping)
template <typename T>
static int32_t launch_dot_copy(int32_t n, int64_t k, const double *B, T *Y, double *part,
                               hipStream_t s) {
  Grid G = grid_rows<double>(n, k, kCapSpmm);
#define N_(KK) (void)0
#define W_(VV, LL) kw_dot_copy<T, WideV<double, VV>::v, LL><<<G.g, kBlock, 0, s>>>(n, k, B, Y, part)
  DISPATCH_K(k, double, G.lpr, N_, W_);
#undef N_
#undef W_
  return G.rows;
}

// Per-(segment, column) dots a.b (and c.d when c != nullptr) of a batch
// into W->seg (then W->seg + nvc): one partial row per segment-aligned
// block, summed per segment in block order.
template <typename TA, typename TB>
static void launch_segdot(const cn_system_t *S, SolverWork *W, int64_t k, const TA *a,
                          const TB *b, const TA *c, const TB *d, hipStream_t s) {
  const int np = c ? 2 : 1;
  double *part = W->part.get();
  const int64_t stride = (int64_t)W->part_rows * k;
  const dim3 g((unsigned)S->seg_blocks, k <= 4 ? 1u : (unsigned)((k + 63) / 64), 1);
  const int32_t *sp = S->d_seg_ptr.get(), *bp = S->blk_ptr.get(), *bs = S->blk_seg.get();
  if (k <= 4) {
#define N_(KK)                                                                              \
  if (np == 2)                                                                              \
    kf_segdot<TA, TB, KK, 2><<<g, kFlat(KK), 0, s>>>(sp, bp, bs, a, b, c, d, part, stride); \
  else                                                                                      \
    kf_segdot<TA, TB, KK, 1><<<g, kFlat(KK), 0, s>>>(sp, bp, bs, a, b, c, d, part, stride);
    switch (k) { case 1: N_(1); break; case 2: N_(2); break; case 3: N_(3); break; default: N_(4); }
#undef N_
  } else if (np == 2) {
    kw_segdot<TA, TB, 2><<<g, kBlock, 0, s>>>(sp, bp, bs, k, a, b, c, d, part, stride);
  } else {
    kw_segdot<TA, TB, 1><<<g, kBlock, 0, s>>>(sp, bp, bs, k, a, b, c, d, part, stride);
  }
  k_segsum<<<dim3((unsigned)S->nseg, (unsigned)np, 1), 64, 0, s>>>(S->nseg, bp, k, part,
                                                                    stride, W->seg.get());
  CN_HIP(hipGetLastError());
}

static unsigned ew_blocks(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Y = (T)(X * s_c) and X += invs_c D, D = 0 over n x k
template <typename T>
static void launch_scale_in(const SysView<T> &V, int64_t k, const double *sc, const double *X,
                            T *Y, hipStream_t s) {
  const int64_t total = (int64_t)V.n * k;
  const Grid G = grid_rows<T>(V.n, k, kCapSpmm);
#define N_(KK) k_scale_in<T><<<ew_blocks(total), 256, 0, s>>>(total, k, V.sm, sc, X, Y)
#define W_(VV, LL) kw_scale_in<T, WideV<T, VV>::v, LL><<<G.g, kBlock, 0, s>>>(V.n, k, V.sm, sc, X, Y)
  DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
}

template <typename T>
static void launch_x_update(const SysView<T> &V, int64_t k, const double *invs, T *D,
                            double *X, hipStream_t s, bool init = false,
                            const int32_t *dfresh = nullptr) {
  const int64_t total = (int64_t)V.n * k;
  const Grid G = grid_rows<T>(V.n, k, kCapSpmm);
  const int ini = init ? 1 : 0;
#define N_(KK) k_x_update<T><<<ew_blocks(total), 256, 0, s>>>(total, k, V.sm, invs, D, X, ini, dfresh)
#define W_(VV, LL) kw_x_update<T, WideV<T, VV>::v, LL><<<G.g, kBlock, 0, s>>>(V.n, k, V.sm, invs, D, X, ini, dfresh)
  DISPATCH_K(k, T, G.lpr, N_, W_);
#undef N_
#undef W_
}

// Two-stage column reduction: returns the buffer and row count the final
// k_reduce_* kernel should read (stride between quantities = rows * k).
struct Red {
  const double *p;
  int32_t rows;
  int64_t stride;
};
// slot 1 writes the first-stage sums to the second half of the buffer (a
// second reduction whose result is needed together with slot 0's)
static Red reduce_rows(SolverWork *W, const double *part, int32_t nrows, int64_t k,
                       int64_t stride, int np, hipStream_t s, int slot = 0) {
  // every partial-writing grid is sized into part_rows when the workspace
  // is set up; a grid past it would have written out of the buffer
  if (nrows > W->part_rows) throw HipError{hipErrorInvalidValue, "partial rows > workspace"};
  if (nrows <= kRedRows) return Red{part, nrows, stride};
  int32_t S = kRedRows;
  int32_t chunk = (nrows + S - 1) / S;
  S = (nrows + chunk - 1) / chunk;
  dim3 g((unsigned)((k + 63) / 64), (unsigned)S);
  const int64_t ostride = (int64_t)kRedRows * k;
  double *out = W->red.get() + (int64_t)slot * ostride;
  k_reduce_stage1<<<g, kBlock, 0, s>>>(nrows, chunk, k, part, stride, np, out, ostride);
  return Red{out, S, ostride};
}

// Row table of A_ff for the SpMM: pattern once per system, values at the
// first solve after an assembly (cn_system_set_positions changes them).
static void ensure_ell(cn_system_t *S, hipStream_t s) {
  const int32_t n = S->n_free;
  if (n <= 0) return;
  if (S->ov_nnz < 0) {
    std::vector<int32_t> ptr((size_t)n + 1), op((size_t)n + 1, 0);
    CN_HIP(hipMemcpyAsync(ptr.data(), S->ff_ptr.get(), sizeof(int32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> col((size_t)std::max(ptr[n], 1));
    if (ptr[n] > 0) {
      CN_HIP(hipMemcpyAsync(col.data(), S->ff_col.get(), sizeof(int32_t) * (size_t)ptr[n],
                            hipMemcpyDeviceToHost, s));
      CN_HIP(hipStreamSynchronize(s));
    }
    // slot 0 holds the diagonal whether or not the row stores one; the
    // row's other entries fill slots 1.. and then the overflow
    for (int32_t i = 0; i < n; ++i) {
      int32_t off = 0;
      for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) off += col[q] != i;
      op[i + 1] = op[i] + std::max(0, off - (kEll - 1));
    }
    S->ell_col.alloc((size_t)n * kEll);
    S->ell_val32.alloc((size_t)n * kEll);
    S->ell_val64.alloc((size_t)n * kEll);
    if (op[n] > 0) {
      S->ov_ptr.alloc((size_t)n + 1);
      S->ov_col.alloc((size_t)op[n]);
      S->ov_val32.alloc((size_t)op[n]);
      S->ov_val64.alloc((size_t)op[n]);
      CN_HIP(hipMemcpyAsync(S->ov_ptr.get(), op.data(), sizeof(int32_t) * ((size_t)n + 1),
                            hipMemcpyHostToDevice, s));
      CN_HIP(hipStreamSynchronize(s));
    }
    S->ov_nnz = op[n];
  }
  if (S->ell_gen == S->vals_gen) return;  // the table holds these values
  S->ell_gen = S->vals_gen;
  const bool ov = S->ov_nnz > 0;
  k_ell_fill<<<grid_for(n, 256), 256, 0, s>>>(
      n, S->ff_ptr.get(), S->ff_col.get(), S->ff_val.get(), S->ell_col.get(),
      S->ell_val32.get(), S->ell_val64.get(), ov ? S->ov_ptr.get() : nullptr,
      S->ov_col.get(), S->ov_val32.get(), S->ov_val64.get());
  CN_HIP(hipGetLastError());
}

// Window row table for ranges of rpr rows (see kl_spmm_dot); refills the
// weights at the first call after an assembly.  Returns false when the matrix does not fit the
// window (a row of more than kWinEll entries, or more than kLdsFar far rows
// in a range).
static bool ensure_win(cn_system_t *S, int32_t rpr, hipStream_t s) {
  const int32_t n = S->n_free;
  if (n <= 0 || rpr <= 0) return false;
  if (S->win_rpr != rpr) {
    S->win_rpr = rpr;
    S->win_ok = -1;
    std::vector<int32_t> ptr((size_t)n + 1);
    CN_HIP(hipMemcpyAsync(ptr.data(), S->ff_ptr.get(), sizeof(int32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> col((size_t)std::max(ptr[n], 1));
    if (ptr[n] > 0) {
      CN_HIP(hipMemcpyAsync(col.data(), S->ff_col.get(), sizeof(int32_t) * (size_t)ptr[n],
                            hipMemcpyDeviceToHost, s));
      CN_HIP(hipStreamSynchronize(s));
    }
    const int32_t R = (n + rpr - 1) / rpr;
    // entries by (range, walking position); positions past a range's rows
    // point at LDS row 0 with weight 0
    std::vector<int32_t> wc((size_t)R * rpr * kWinEll, 0), wq((size_t)R * rpr * kWinEll, -1);
    std::vector<int32_t> fp(R + 1, 0), fr;
    bool ok = true;
    for (int32_t x = 0; x < R && ok; ++x) {
      const int32_t rs = x * rpr, re = std::min(n, rs + rpr);
      const bool down = (x & 1) != 0;
      auto pos = [&](int32_t j) { return down ? re - 1 - j : j - rs; };  // walking position
      auto chunk = [&](int32_t j) {  // floor(pos / kLdsRows)
        const int32_t d = pos(j);
        return d >= 0 ? d / kLdsRows : -((-d + kLdsRows - 1) / kLdsRows);
      };
      std::vector<int32_t> far;  // this range's far rows, sorted
      for (int32_t r = rs; r < re; ++r) {
        const int32_t tr = chunk(r);
        for (int32_t q = ptr[r]; q < ptr[r + 1]; ++q) {
          const int32_t gj = chunk(col[q]);
          if (gj < tr - kLdsBack || gj > tr + kLdsAhead) far.push_back(col[q]);
        }
      }
      std::sort(far.begin(), far.end());
      far.erase(std::unique(far.begin(), far.end()), far.end());
      if ((int32_t)far.size() > kLdsFar) { ok = false; break; }
      for (int32_t r = rs; r < re && ok; ++r) {
        const int32_t tr = chunk(r);
        // the LDS row of neighbour j while row r is computed
        auto enc = [&](int32_t j) {
          const int32_t gj = chunk(j);
          if (gj >= tr - kLdsBack && gj <= tr + kLdsAhead)
            return ((gj + kLdsSlots) % kLdsSlots) * kLdsRows + (pos(j) & (kLdsRows - 1));
          const int32_t f = (int32_t)(std::lower_bound(far.begin(), far.end(), j) - far.begin());
          return kLdsSlots * kLdsRows + f;
        };
        const size_t o = ((size_t)x * rpr + (size_t)pos(r)) * kWinEll;
        int32_t *c = &wc[o], *qq = &wq[o];
        int32_t dq = -1, u = 1, last = r;
        for (int32_t q = ptr[r]; q < ptr[r + 1]; ++q) {
          if (col[q] == r) { dq = q; continue; }
          if (u >= kWinEll) { ok = false; break; }
          c[u] = enc(col[q]);
          qq[u] = q;
          last = col[q];
          ++u;
        }
        c[0] = enc(r);
        qq[0] = dq;
        for (; u < kWinEll; ++u) { c[u] = enc(last); qq[u] = -1; }
      }
      fp[x + 1] = fp[x] + (int32_t)far.size();
      fr.insert(fr.end(), far.begin(), far.end());
    }
    if (!ok) return false;
    S->win_col.alloc(wc.size());
    S->win_q.alloc(wq.size());
    S->win_pk32.release();
    S->win_pk64.release();
    S->win_gen = 0;
    S->win_far_ptr.alloc(fp.size());
    S->win_far_rows.alloc(std::max<size_t>(fr.size(), 1));
    CN_HIP(hipMemcpyAsync(S->win_col.get(), wc.data(), sizeof(int32_t) * wc.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipMemcpyAsync(S->win_q.get(), wq.data(), sizeof(int32_t) * wq.size(),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipMemcpyAsync(S->win_far_ptr.get(), fp.data(), sizeof(int32_t) * fp.size(),
                          hipMemcpyHostToDevice, s));
    if (!fr.empty())
      CN_HIP(hipMemcpyAsync(S->win_far_rows.get(), fr.data(), sizeof(int32_t) * fr.size(),
                            hipMemcpyHostToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
    S->win_ok = 1;
  }
  if (S->win_ok != 1) return false;
  const int64_t total = (int64_t)S->win_col.n;  // (R * rpr) positions x kWinEll
  const int64_t nblocks = total / kWinEnt;
  if (!S->win_pk32.get()) {
    S->win_pk32.alloc((size_t)nblocks * win_block_bytes<float>());
    S->win_pk64.alloc((size_t)nblocks * win_block_bytes<double>());
  }
  if (S->win_gen == S->vals_gen) return true;  // the packs hold these values
  S->win_gen = S->vals_gen;
  k_win_pack<float><<<grid_for(total, 256), 256, 0, s>>>(total, S->win_col.get(),
                                                        S->win_q.get(), S->ff_val.get(),
                                                        S->win_pk32.get());
  k_win_pack<double><<<grid_for(total, 256), 256, 0, s>>>(total, S->win_col.get(),
                                                         S->win_q.get(), S->ff_val.get(),
                                                         S->win_pk64.get());
  CN_HIP(hipGetLastError());
  return true;
}

// The system's launch shapes into the view; the LDS-staged SpMM unless the
// tune asks for the row table (or it does not apply).
template <typename T>
static void use_win(cn_system_t *S, SysView<T> &V, int64_t k, hipStream_t s);
template <typename T>
static void use_shapes(cn_system_t *S, SysView<T> &V, int64_t k, hipStream_t s) {
  V.gens = S->tune.spmm_gens > 0 ? S->tune.spmm_gens : (int32_t)kSpmmGens;
  V.hsh.level_cap = S->tune.level_cap;
  V.hsh.tail_rows = S->tune.tail_rows;
  V.hsh.chain_levels = S->tune.chain_levels;
  V.hsh.coarse_prec = S->tune.coarse_prec;
  V.hsh.down_cap = S->tune.down_cap;
  if (S->tune.spmm_form != 1) use_win<T>(S, V, k, s);
}
template <>
void use_win<float>(cn_system_t *S, SysView<float> &V, int64_t k, hipStream_t s) {
  if (!win_applies<float>(S->n_free, k)) return;
  if (!ensure_win(S, win_rpr<float>(S->n_free, k), s)) return;
  V.rpr = S->win_rpr;
  V.wpk = S->win_pk32.get();
  V.far_ptr = S->win_far_ptr.get(); V.far_rows = S->win_far_rows.get();
}
template <>
void use_win<double>(cn_system_t *S, SysView<double> &V, int64_t k, hipStream_t s) {
  if (!win_applies<double>(S->n_free, k)) return;
  if (!ensure_win(S, win_rpr<double>(S->n_free, k), s)) return;
  V.rpr = S->win_rpr;
  V.wpk = S->win_pk64.get();
  V.far_ptr = S->win_far_ptr.get(); V.far_rows = S->win_far_rows.get();
}

template <typename T>
static SysView<T> make_view(cn_system_t *S, cn_ic0_t *F, SolverWork *W);

template <>
SysView<float> make_view<float>(cn_system_t *S, cn_ic0_t *F, SolverWork *W) {
  SysView<float> V{};
  V.n = S->n_free;
  V.ptr = S->ff_ptr.get(); V.col = S->ff_col.get(); V.val = S->ff_val32.get();
  V.val64 = S->ff_val.get();
  V.ecol = S->ell_col.get(); V.eval = S->ell_val32.get();
  if (S->ov_nnz > 0) { V.ovp = S->ov_ptr.get(); V.ovc = S->ov_col.get(); V.ovv = S->ov_val32.get(); }
  if (F) {
    V.ic0 = F;
    V.rowmap = F->perm.get();
    V.blk = F->blk_ptr.get();
    V.lp = F->l_ptr.get(); V.lcol = F->l_col_nat.get(); V.lval = F->l_val32.get();
    V.up = F->u_ptr.get(); V.ucol = F->u_col_nat.get(); V.uval = F->u_val32.get();
    V.rdiag = F->l_rdiag32.get();
    V.mb = F->sweep_mb;
    V.lsw = {F->lsw.ell_col.get(), F->lsw.ell_cnt.get(), F->lsw.ov_ptr.get(), F->lsw.ov_col.get(),
             F->lsw.ell_val32.get(), F->lsw.ov_val32.get(), F->lsw.dense32.get()};
    V.usw = {F->usw.ell_col.get(), F->usw.ell_cnt.get(), F->usw.ov_ptr.get(), F->usw.ov_col.get(),
             F->usw.ell_val32.get(), F->usw.ov_val32.get(), F->usw.dense32.get()};
  }
  V.dinv = W->dinv32.get();
  return V;
}

template <>
SysView<double> make_view<double>(cn_system_t *S, cn_ic0_t *F, SolverWork *W) {
  SysView<double> V{};
  V.n = S->n_free;
  V.ptr = S->ff_ptr.get(); V.col = S->ff_col.get(); V.val = S->ff_val.get();
  V.val64 = S->ff_val.get();
  V.ecol = S->ell_col.get(); V.eval = S->ell_val64.get();
  if (S->ov_nnz > 0) { V.ovp = S->ov_ptr.get(); V.ovc = S->ov_col.get(); V.ovv = S->ov_val64.get(); }
  if (F) {
    V.ic0 = F;
    V.rowmap = F->perm.get();
    V.blk = F->blk_ptr.get();
    V.lp = F->l_ptr.get(); V.lcol = F->l_col_nat.get(); V.lval = F->l_val.get();
    V.up = F->u_ptr.get(); V.ucol = F->u_col_nat.get(); V.uval = F->u_val.get();
    V.rdiag = F->l_rdiag64.get();
    V.mb = F->sweep_mb;
    V.lsw = {F->lsw.ell_col.get(), F->lsw.ell_cnt.get(), F->lsw.ov_ptr.get(), F->lsw.ov_col.get(),
             F->lsw.ell_val.get(), F->lsw.ov_val.get(), F->lsw.dense.get()};
    V.usw = {F->usw.ell_col.get(), F->usw.ell_cnt.get(), F->usw.ov_ptr.get(), F->usw.ov_col.get(),
             F->usw.ell_val.get(), F->usw.ov_val.get(), F->usw.dense.get()};
  }
  V.dinv = W->dinv64.get();
  return V;
}

// ========================================= fused solve (MIXED and FP64) ==
template <typename T>
static cn_status fused_solve(cn_system_t *S, cn_ic0_t *F, cn_hsc_t *H, const cn_sgs_t *G,
                             const cn_pcg_opts *o,
                             int32_t k, const double *d_b, double *d_x,
                             cn_solve_info *info, hipStream_t s) {
  const int32_t n = S->n_free;
  const int64_t total = (int64_t)n * k;
  SysView<T> V0{};
  V0.n = n;
  V0.ic0 = F;
  if (F) V0.mb = F->sweep_mb;
  ensure_segments(S, s);
  // a batch keeps its scalars per (segment, column): kk virtual columns
  const bool batch = S->nseg > 1;
  const int64_t kk = (int64_t)S->nseg * k;
  // partial rows: the widest of every row grid that writes partials (the
  // SpMM's own, the row kernels' -- rupd, dot, residual -- and the sweeps')
  int32_t need = std::max<int32_t>(fwd_rows<T>(V0, k), grid_spmm<T>(n, k).rows);
  need = std::max<int32_t>(need, grid_spmm<double>(n, k).rows);
  need = std::max<int32_t>(need, grid_rows<T>(n, k, kCapSpmm).rows);
  need = std::max<int32_t>(need, grid_rows<double>(n, k, kCapSpmm).rows);
  need = std::max<int32_t>(need, S->seg_blocks);
  if (H) need = std::max<int32_t>(need, hsc_rz_rows<T>(H, k, HscShape{S->tune.level_cap,
                                                                     S->tune.tail_rows,
                                                                     S->tune.chain_levels}));
  SolverWork *W = ensure_work(S, k, need);
  k_dinv<<<grid_for(n, 256), 256, 0, s>>>(n, S->diag.get(), W->dinv64.get(),
                                         W->dinv32.get());
  ensure_ell(S, s);
  SysView<T> V = make_view<T>(S, F, W);
  use_shapes<T>(S, V, k, s);
  V.hsc = H;
  V.sgs = G;
  V.sm = SegMap{batch ? S->seg_of_row.get() : nullptr, kk};
  double *st = W->st.get();
  double *sg0 = W->seg.get(), *sg1 = sg0 + kk;  // per-segment dot sums
  double *hs = W->h_st;
  int32_t *ctl = W->ctl.get();
  int32_t *hc = W->h_ctl;
  const int64_t stride = (int64_t)W->part_rows * k;
  double *p0 = W->part.get(), *p1 = p0 + stride;
  const unsigned red = (unsigned)((kk + 63) / 64);
  T *D = reinterpret_cast<T *>(W->buf[0].get());
  T *R = reinterpret_cast<T *>(W->buf[1].get());
  T *Z = reinterpret_cast<T *>(W->buf[2].get());
  T *P = reinterpret_cast<T *>(W->buf[3].get());
  T *Q = reinterpret_cast<T *>(W->buf[4].get());
  // The fused iteration (a single system preconditioned by HSC or SGS, with
  // the LDS-staged SpMM): the direction update rides on the product
  // (KL_DIR) and the residual and iterate updates on a second product
  // (KL_UPD), so there is no Q; its buffer holds the other of two P buffers
  // (KL_DIR reads p_old from one and writes p_new to the other).  The
  // direction update of an iteration is applied by the next iteration's
  // KL_DIR ("pending"), or, before a round that re-applies the
  // preconditioner, by a plain update (the round's new r.z overwrites beta).
  const bool klf = !batch && !F && (H || G) && V.rpr > 0 && win_applies<T>(n, k);
  T *Pb[2] = {P, Q};
  int cur = 0;
  bool pending = false;
  // x accumulates in the caller's d_x (no copy in or out) unless d_x is d_b,
  // which is read until the end
  double *X = d_x != d_b ? d_x : W->xp.get(), *R64 = W->rp.get();
  // the two x buffers of the fused x update + residual (KL_RES): the
  // caller's x (or xp) and xp (or rp, free once the first round has scaled
  // its residual into R); the solve copies x into d_x at the end if needed
  double *const Xa = X, *const Xb = X == W->xp.get() ? W->rp.get() : W->xp.get();
  const double theta = 0.5;  // iterated-residual target below tol
  // reliable update: a residual replacement is due once a column's iterated
  // ||r|| has dropped by inner_tol (MIXED) or 1e-4 (FP64) since the last one
  const double delta2 = sizeof(T) == 4 ? o->inner_tol * o->inner_tol : 1e-8;
  // A column whose iterated r.r falls below floor2 times its value at the
  // last replacement stops, like a converged one: in FP32 past 1e-12 (||r||
  // by 1e-6) the recurrence carries rounding noise only, unrelated to the
  // true residual, so continuing its direction after the replacement would
  // not converge.  The next round restarts it (p = z of the true residual)
  // unless its FP64 residual is already at the tolerance.  A near-exact
  // preconditioner (a hierarchy whose coarse block is the whole matrix)
  // gets there in one iteration.
  const double floor2 = sizeof(T) == 4 ? 1e-12 : 0.0;

  // x0, true residual, ||b||
  if (o->warm_start) {
    if (X != d_x)
      CN_HIP(hipMemcpyAsync(X, d_x, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
  }
  // from x0 = 0, X is first written by the first x update (which then does
  // not read it); a solve that ends before any update zeroes it at the end
  bool x_unwritten = !o->warm_start;
  int32_t rows;
  Red rd;
  // the FP64 residual the first round scales into R: from x0 = 0 it is b
  // itself (b - A 0 = b exactly), so no SpMM, and ||r|| is ||b||
  const double *rsrc = o->warm_start ? R64 : d_b;
  // from x0 = 0 (one system, k > 4) the pass that forms ||b|| also stores b
  // into R, unscaled: a column whose ||b|| is within [1e-15, 1e15] then runs
  // unscaled (scale 1), far from the FP32 range's ends, and needs no
  // scaling pass
  const bool b_in_r = !batch && !o->warm_start && k > 4;
  if (batch) {
    launch_segdot<double, double>(S, W, k, d_b, d_b, nullptr, nullptr, s);
    k_reduce_plain<<<red, kBlock, 0, s>>>(1, kk, sg0, st + ST_BB * kk);
    if (o->warm_start) {
      launch_resid<double>(n, V.ptr, V.col, V.val64, k, X, d_b, R64, nullptr, p0, s);
      launch_segdot<double, double>(S, W, k, R64, R64, nullptr, nullptr, s);
    }
    k_reduce_plain<<<red, kBlock, 0, s>>>(1, kk, sg0, st + ST_RR * kk);
  } else {
    rows = b_in_r ? launch_dot_copy<T>(n, k, d_b, R, p0, s)
                  : launch_dot<double>(n, k, d_b, d_b, p0, s);
    rd = reduce_rows(W, p0, rows, k, stride, 1, s);
    k_reduce_plain<<<red, kBlock, 0, s>>>(rd.rows, k, rd.p, st + ST_BB * k);
    if (o->warm_start) {
      rows = launch_resid<double>(n, V.ptr, V.col, V.val64, k, X, d_b, R64, nullptr, p0, s);
      rd = reduce_rows(W, p0, rows, k, stride, 1, s);
    }
    k_reduce_plain<<<red, kBlock, 0, s>>>(rd.rows, k, rd.p, st + ST_RR * k);
  }
  // (the fused iteration's d needs no zeros: CTL_DFRESH)
  if (!klf) CN_HIP(hipMemsetAsync(D, 0, sizeof(T) * total, s));
  CN_HIP(hipMemsetAsync(P, 0, sizeof(T) * total, s));
  CN_HIP(hipGetLastError());

  // act: the columns the host started the round with; dact: those the
  // device still iterated at the round's end (its stop tests clear them).
  // act and rst live in the pinned mirror of the flags block.
  int32_t *act = W->h_flags + CTL_N;
  int32_t *rst = act + W->nvc + 4;
  int32_t *dact = W->h_dact;
  std::fill(act, act + kk, 0);
  std::fill(rst, rst + kk, 0);
  int32_t iters = 0, launched = 0, updates = 0;
  bool r_scaled = false;  // R already holds the scaled residual of this round
  cn_status status = CN_OK;
  double relmax = 0.0;
  bool first = true;
  for (;;) {
    // ---- read the true FP64 residual, decide which columns continue
    CN_HIP(hipMemcpyAsync(hs + ST_BB * kk, st + ST_BB * kk, sizeof(double) * 2 * kk,
                          hipMemcpyDeviceToHost, s));  // BB and RR adjacent
    if (!first)
      CN_HIP(hipMemcpyAsync(dact, W->active.get(), sizeof(int32_t) * kk,
                            hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    relmax = 0.0;
    int32_t nact = 0;
    for (int64_t c = 0; c < kk; ++c) {
      double bb = hs[ST_BB * kk + c], rr = hs[ST_RR * kk + c];
      double rel = bb > 0.0 ? std::sqrt(rr / bb) : (rr > 0.0 ? INFINITY : 0.0);
      if (!std::isfinite(rel)) rel = INFINITY;
      relmax = std::max(relmax, rel);
      bool a = rel > o->tol_rel && bb > 0.0;
      // a column the device stopped (its iterated residual reached 0.5 tol
      // or FP32 rounding noise), but whose true residual is not there yet,
      // restarts (p = z)
      rst[c] = (a && (first || !dact[c])) ? 1 : 0;
      act[c] = a ? 1 : 0;
      nact += a ? 1 : 0;
    }
    if (nact == 0) { status = CN_OK; break; }
    if (!std::isfinite(relmax)) { status = CN_BREAKDOWN; break; }
    if (launched >= o->max_iter || updates >= o->max_refine * 8) { status = CN_MAXIT; break; }
    // ---- per-column scale s = 1/||b|| (fixed for the solve), or 1 for a
    // column whose unscaled b is already in R
    if (first) {
      bool rescale = false;
      for (int64_t c = 0; c < kk; ++c) {
        double bb = hs[ST_BB * kk + c];
        // (b = 0: the column never runs, its R column is 0 either way)
        const bool one = b_in_r && (bb == 0.0 || (bb >= 1e-30 && bb <= 1e30));
        rescale = rescale || !one;
        hs[ST_S * kk + c] = one ? 1.0 : (bb > 0.0 ? 1.0 / std::sqrt(bb) : 0.0);
        hs[ST_INVS * kk + c] = one ? 1.0 : (bb > 0.0 ? std::sqrt(bb) : 0.0);
        // the stop test on the scaled residual: ||r s|| <= theta tol ||b|| s
        double t = theta * o->tol_rel;
        hs[ST_TOL2 * kk + c] = one ? t * t * bb : t * t;
        // the reference r.r of the round: the true one, scaled
        hs[ST_RREF * kk + c] = hs[ST_RR * kk + c] * hs[ST_S * kk + c] * hs[ST_S * kk + c];
      }
      CN_HIP(hipMemcpyAsync(st + ST_TOL2 * kk, hs + ST_TOL2 * kk, sizeof(double) * 4 * kk,
                            hipMemcpyHostToDevice, s));  // TOL2, RREF, S, INVS
      // R holds b: scaled already unless some column needs 1/||b||
      if (b_in_r && !rescale) r_scaled = true;
    }
    // Continuing columns keep p and the iterated r.z (van der Vorst-Ye
    // reliable update): only a (re)start needs z = M^-1 r of the new r.
    bool any_rst = false;
    for (int64_t c = 0; c < kk; ++c) any_rst = any_rst || rst[c];
    const bool apply_m = first || any_rst;
    const bool round0 = first;
    first = false;
    hc[CTL_ACTIVE] = nact;
    hc[CTL_BREAKDOWN] = 0;
    hc[CTL_ONE] = 1;
    hc[CTL_TRIGGER] = 0;
    hc[CTL_JUST] = 0;
    hc[CTL_ITERS] = iters;
    hc[CTL_PAUSE] = 0;
    hc[CTL_DFRESH] = klf ? 1 : 0;
    // control block, active and restart flags in one copy
    CN_HIP(hipMemcpyAsync(ctl, hc, sizeof(int32_t) * (CTL_N + 2 * ((size_t)W->nvc + 4)),
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipMemsetAsync(st + ST_ALPHA * kk, 0, sizeof(double) * kk, s));
    // after the first round the residual kernel writes the scaled R itself
    if (!r_scaled)
      launch_scale_in<T>(V, k, st + ST_S * kk, rsrc, R, s);
    if (apply_m) {
      // the deferred direction update first (alpha is 0 here: no d update),
      // while z and beta are still the last iteration's
      if (klf && pending) launch_bwd<T>(V, k, st, Z, Pb[cur], D, ctl, s);
      pending = false;
      // z = M^-1 r; rz; continuing columns keep p, restarted ones take p = z
      if (batch) {
        rows = launch_fwd<T, false>(V, k, st + ST_ALPHA * kk, Q, R, Z, p1, stride, ctl, s);
        launch_segdot<T, T>(S, W, k, R, R, R, Z, s);
        k_reduce_beta<<<red, kBlock, 0, s>>>(1, 1, kk, sg0, sg1, st, W->active.get(),
                                             W->restart.get(), ctl, 1, delta2, floor2);
      } else {
        // the reference r.r is the true one, scaled (the first round
        // uploaded it with the scales): no pass over r for it
        if (!round0) {
          for (int64_t c = 0; c < kk; ++c) {
            const double sc = hs[ST_S * kk + c];
            hs[ST_RREF * kk + c] = hs[ST_RR * kk + c] * sc * sc;
          }
          CN_HIP(hipMemcpyAsync(st + ST_RREF * kk, hs + ST_RREF * kk, sizeof(double) * kk,
                                hipMemcpyHostToDevice, s));
        }
        rows = launch_fwd<T, false>(V, k, st + ST_ALPHA * kk, Q, R, Z, p1, stride, ctl, s,
                                    true);
        Red b2 = reduce_rows(W, p1 + stride, rows, k, stride, 1, s, 1);
        k_reduce_beta<<<red, kBlock, 0, s>>>(0, b2.rows, k, p1, b2.p, st, W->active.get(),
                                             W->restart.get(), ctl, 2, delta2, floor2);
      }
      if (klf) pending = true;   // p = gamma z + beta p in the first KL_DIR
      else launch_bwd<T>(V, k, st, Z, P, D, ctl, s);
    } else {
      // new reference for the replacement trigger: the scaled true r.r
      for (int64_t c = 0; c < kk; ++c) {
        const double sc = hs[ST_S * kk + c];
        hs[ST_RREF * kk + c] = hs[ST_RR * kk + c] * sc * sc;
      }
      CN_HIP(hipMemcpyAsync(st + ST_RREF * kk, hs + ST_RREF * kk, sizeof(double) * kk,
                            hipMemcpyHostToDevice, s));
    }
    CN_HIP(hipGetLastError());
    // ---- CG iterations until every column stops or an update is due
    int32_t it = 0;
    const int32_t budget = o->max_iter - launched;
    // a whole-apply preconditioner converges in a few iterations per
    // residual replacement: read the flags often enough not to idle long
    const int32_t every = std::max(1, H ? std::min(o->check_every, 2) : o->check_every);
    // one CG iteration: every launch reads its scalars and flags from the
    // device, so the same launches serve every iteration
    auto iterate = [&](hipStream_t q) {
      if (klf) {
        const int32_t r0 = launch_kl<T, KL_DIR>(V, k, Pb[cur], Z, Pb[cur ^ 1], nullptr,
                                                nullptr, st, p0, ctl, q);
        Red a1 = reduce_rows(W, p0, r0, k, stride, 1, q);
        k_reduce_alpha<<<red, kBlock, 0, q>>>(a1.rows, k, a1.p, st, W->active.get(), ctl);
        const int32_t r1 = launch_kl<T, KL_UPD>(V, k, Pb[cur ^ 1], nullptr, nullptr, R, D, st,
                                                p1, ctl, q);
        const int32_t r2 = launch_fwd<T, true>(V, k, st + ST_ALPHA * kk, nullptr, R, Z, p1,
                                               stride, ctl, q, true);
        Red b1 = reduce_rows(W, p1, r1, k, stride, 1, q, 0);
        Red b2 = reduce_rows(W, p1 + stride, r2, k, stride, 1, q, 1);
        k_reduce_beta<<<red, kBlock, 0, q>>>(b1.rows, b2.rows, k, b1.p, b2.p, st,
                                             W->active.get(), W->restart.get(), ctl, 0, delta2, floor2);
        cur ^= 1;
        return;
      }
      int32_t rw = launch_spmm_dot<T>(V, k, P, Q, p0, ctl, q);
      if (batch) {
        launch_segdot<T, T>(S, W, k, P, Q, nullptr, nullptr, q);
        k_reduce_alpha<<<red, kBlock, 0, q>>>(1, kk, sg0, st, W->active.get(), ctl);
      } else {
        Red r1 = reduce_rows(W, p0, rw, k, stride, 1, q);
        k_reduce_alpha<<<red, kBlock, 0, q>>>(r1.rows, k, r1.p, st, W->active.get(), ctl);
      }
      int32_t rzr = 0;
      rw = launch_fwd<T, true>(V, k, st + ST_ALPHA * kk, Q, R, Z, p1, stride, ctl, q, false,
                               batch ? nullptr : &rzr);
      if (batch) {
        launch_segdot<T, T>(S, W, k, R, R, R, Z, q);
        k_reduce_beta<<<red, kBlock, 0, q>>>(1, 1, kk, sg0, sg1, st, W->active.get(),
                                             W->restart.get(), ctl, 0, delta2, floor2);
      } else {
        Red b1 = reduce_rows(W, p1, rw, k, stride, 1, q, 0);
        Red b2 = reduce_rows(W, p1 + stride, rzr >= 0 ? rzr : rw, k, stride, 1, q, 1);
        k_reduce_beta<<<red, kBlock, 0, q>>>(b1.rows, b2.rows, k, b1.p, b2.p, st,
                                             W->active.get(), W->restart.get(), ctl, 0, delta2,
                                             floor2);
      }
      launch_bwd<T>(V, k, st, Z, P, D, ctl, q);
    };
    const bool use_graph = o->graph && !g_prof.on && every > 1;
    // the captured block for the current P parity: everything its launches
    // depend on besides W's own buffers is in the key
    auto block_graph = [&]() -> hipGraphExec_t {
      const std::vector<uint64_t> key = {
          (uint64_t)k, sizeof(T), (uint64_t)every, (uint64_t)kk,
          H ? H->serial : 0, H ? (uint64_t)(uintptr_t)H->work.get() : 0,
          F ? F->serial : 0, G ? G->serial : 0, (uint64_t)V.rpr,
          (uint64_t)(uintptr_t)V.wpk, (uint64_t)(uintptr_t)V.ecol, (uint64_t)V.gens,
          (uint64_t)(uint32_t)V.hsh.level_cap, (uint64_t)(uint32_t)V.hsh.tail_rows,
          (uint64_t)(uint32_t)V.hsh.chain_levels, (uint64_t)(uint32_t)V.hsh.coarse_prec,
          (uint64_t)(uint32_t)V.hsh.down_cap,
          (uint64_t)klf, (uint64_t)cur,
          bits_of(delta2), bits_of(floor2)};
      for (auto &e : W->graphs)
        if (e.first == key) return e.second;
      if (W->graphs.size() >= 4) {
        CN_HIP(hipGraphExecDestroy(W->graphs.front().second));
        W->graphs.erase(W->graphs.begin());
      }
      if (!W->cap) CN_HIP(hipStreamCreateWithFlags(&W->cap, hipStreamNonBlocking));
      CN_HIP(hipStreamBeginCapture(W->cap, hipStreamCaptureModeThreadLocal));
      hipGraph_t gr = nullptr;
      const int cur0 = cur;
      try {
        for (int32_t e = 0; e < every; ++e) iterate(W->cap);
      } catch (...) {
        cur = cur0;
        (void)hipStreamEndCapture(W->cap, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      cur = cur0;  // captured, not run
      CN_HIP(hipStreamEndCapture(W->cap, &gr));
      hipGraphExec_t ex = nullptr;
      const hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      CN_HIP(ie);
      W->graphs.emplace_back(key, ex);
      return ex;
    };
    while (it < budget) {
      if (use_graph && budget - it >= every) {
        CN_HIP(hipGraphLaunch(block_graph(), s));
        if (klf && (every & 1)) cur ^= 1;
        it += every;
        if (klf) pending = true;
      } else {
        iterate(s);
        ++it;
        if (klf) pending = true;
      }
      if (it % every == 0 || it == budget) {
        CN_HIP(hipMemcpyAsync(hc, ctl, sizeof(int32_t) * 6, hipMemcpyDeviceToHost, s));
        CN_HIP(hipStreamSynchronize(s));
        if (hc[CTL_ACTIVE] == 0 || hc[CTL_TRIGGER]) break;
      }
    }
    CN_HIP(hipGetLastError());
    launched += it;
    iters = hc[CTL_ITERS];
    // ---- reliable update: x += d / s (FP64), d = 0; r = b - A x (FP64).
    // The fused iteration with the LDS-staged product does both in one
    // pass (KL_RES: x lands in LDS as x + d / s and goes to the other x
    // buffer); otherwise two kernels.  (One row-table pass that formed
    // x + d / s for the gathered neighbours as well was 0.4 % slower a step
    // at k = 768 than the two.)
    bool kl_res = false;
    if constexpr (sizeof(T) == 4) kl_res = klf && win_applies<double>(n, k);
    if (kl_res) {
      double *Xo = X == Xa ? Xb : Xa;
      rows = launch_resid_kl(S, n, V.rpr, k, X, d_b, reinterpret_cast<float *>(R), st, p0, s,
                             reinterpret_cast<const float *>(D), Xo, ctl + CTL_DFRESH,
                             x_unwritten);
      X = Xo;
      x_unwritten = false;
      r_scaled = true;
      rd = reduce_rows(W, p0, rows, k, stride, 1, s);
      k_reduce_plain<<<red, kBlock, 0, s>>>(rd.rows, k, rd.p, st + ST_RR * k);
    } else {
      launch_x_update<T>(V, k, st + ST_INVS * kk, D, X, s, x_unwritten,
                         klf ? ctl + CTL_DFRESH : nullptr);
      x_unwritten = false;
      if (batch) {
        // FP64 residual, per-segment r.r, scaled into R at the top of the loop
        launch_resid<double>(n, V.ptr, V.col, V.val64, k, X, d_b, R64, nullptr, p0, s);
        rsrc = R64;
        launch_segdot<double, double>(S, W, k, R64, R64, nullptr, nullptr, s);
        k_reduce_plain<<<red, kBlock, 0, s>>>(1, kk, sg0, st + ST_RR * kk);
      } else {
        rows = launch_resid<T>(n, V.ptr, V.col, V.val64, k, X, d_b, R, st + ST_S * k, p0, s);
        r_scaled = true;
        rd = reduce_rows(W, p0, rows, k, stride, 1, s);
        k_reduce_plain<<<red, kBlock, 0, s>>>(rd.rows, k, rd.p, st + ST_RR * k);
      }
    }
    CN_HIP(hipGetLastError());
    ++updates;
    if (it == 0) { status = CN_MAXIT; break; }
  }
  if (x_unwritten) CN_HIP(hipMemsetAsync(X, 0, sizeof(double) * total, s));
  if (X != d_x)
    CN_HIP(hipMemcpyAsync(d_x, X, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
  CN_HIP(hipStreamSynchronize(s));
  if (info) {
    info->status = status;
    info->iters = iters;
    info->refinements = updates;
    info->relres_max = relmax;
  }
  return status;
}

// ================================ unfused FP64 baseline, scalars on host ==
static cn_status plain_solve(cn_system_t *S, cn_ic0_t *F, cn_hsc_t *H, const cn_sgs_t *G,
                             const cn_pcg_opts *o,
                             int32_t k, const double *d_b, double *d_x,
                             cn_solve_info *info, hipStream_t s) {
  const int32_t n = S->n_free;
  const int64_t total = (int64_t)n * k;
  SysView<double> V0{};
  V0.n = n;
  V0.ic0 = F;
  if (F) V0.mb = F->sweep_mb;
  int32_t need = std::max<int32_t>(fwd_rows<double>(V0, k), grid_spmm<double>(n, k).rows);
  need = std::max<int32_t>(need, grid_rows<double>(n, k, kCapSpmm).rows);
  SolverWork *W = ensure_work(S, k, need);
  k_dinv<<<grid_for(n, 256), 256, 0, s>>>(n, S->diag.get(), W->dinv64.get(),
                                         W->dinv32.get());
  ensure_ell(S, s);
  SysView<double> V = make_view<double>(S, F, W);
  V.hsc = H;
  V.sgs = G;
  V.sm = SegMap{nullptr, k};
  double *st = W->st.get();
  const int64_t stride = (int64_t)W->part_rows * k;
  const unsigned red = (unsigned)((k + 63) / 64);
  double *X = W->xp.get(), *R = W->rp.get();
  double *Z = W->buf[2].get(), *P = W->buf[3].get(), *Q = W->buf[4].get();
  double *Dd = W->buf[0].get();
  int32_t *ctl = W->ctl.get();
  W->h_ctl[CTL_ACTIVE] = 1; W->h_ctl[CTL_BREAKDOWN] = 0; W->h_ctl[CTL_ONE] = 1;
  W->h_ctl[CTL_TRIGGER] = 0;
  W->h_ctl[CTL_JUST] = 0;
  W->h_ctl[CTL_ITERS] = 0;
  W->h_ctl[CTL_PAUSE] = 0;
  W->h_ctl[CTL_DFRESH] = 0;
  CN_HIP(hipMemcpyAsync(ctl, W->h_ctl, sizeof(int32_t) * 7, hipMemcpyHostToDevice, s));
  // preconditioner-only coefficients: alpha 0, beta 1, gamma 0 (p, d untouched)
  {
    std::vector<double> h(3 * (size_t)k);
    for (int32_t c = 0; c < k; ++c) { h[c] = 0.0; h[k + c] = 1.0; h[2 * k + c] = 0.0; }
    CN_HIP(hipMemcpyAsync(st + ST_ALPHA * k, h.data(), sizeof(double) * 3 * k,
                          hipMemcpyHostToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
  }
  auto dot_to_host = [&](const double *a, const double *b, double *h) {
    int32_t rows = launch_dot<double>(n, k, a, b, W->part.get(), s);
    Red rd = reduce_rows(W, W->part.get(), rows, k, stride, 1, s);
    k_reduce_plain<<<red, kBlock, 0, s>>>(rd.rows, k, rd.p, st + ST_TMP * k);
    CN_HIP(hipMemcpyAsync(h, st + ST_TMP * k, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
  };
  auto apply_m = [&]() {  // Z = M^-1 R
    launch_fwd<double, false>(V, k, st + ST_ALPHA * k, Q, R, Z, W->part.get(), stride,
                              ctl, s);
    if (F) launch_bwd<double>(V, k, st, Z, P, Dd, ctl, s);
  };
  if (o->warm_start) {
    CN_HIP(hipMemcpyAsync(X, d_x, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
    launch_resid<double>(n, V.ptr, V.col, V.val64, k, X, d_b, R, nullptr, W->part.get(), s);
  } else {
    CN_HIP(hipMemsetAsync(X, 0, sizeof(double) * total, s));
    CN_HIP(hipMemcpyAsync(R, d_b, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
  }
  CN_HIP(hipMemsetAsync(P, 0, sizeof(double) * total, s));
  std::vector<double> bb(k), rz(k), tmp(k), alpha(k), beta(k), rr(k);
  dot_to_host(d_b, d_b, bb.data());
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
    launch_spmm_dot<double>(V, k, P, Q, W->part.get(), ctl, s);      // SpMV
    dot_to_host(P, Q, tmp.data());                                   // p.Ap
    for (int32_t c = 0; c < k; ++c) {
      if (act[c] && tmp[c] == 0.0) { act[c] = 0; status = CN_BREAKDOWN; }
      alpha[c] = act[c] ? rz[c] / tmp[c] : 0.0;
    }
    CN_HIP(hipMemcpyAsync(st + ST_TMP * k, alpha.data(), sizeof(double) * k,
                          hipMemcpyHostToDevice, s));
    k_axpy<double><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_TMP * k, 1.0, P, X);
    k_axpy<double><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_TMP * k, -1.0, Q, R);
    ++it;
    dot_to_host(R, R, rr.data());                                    // stop test
    any = false;
    for (int32_t c = 0; c < k; ++c) {
      if (act[c] && !(rr[c] > o->tol_rel * o->tol_rel * bb[c])) act[c] = 0;
      any |= act[c] != 0;
    }
    if (!any) { if (status == CN_MAXIT) status = CN_OK; break; }
    apply_m();                                                       // z = M^-1 r
    dot_to_host(R, Z, tmp.data());                                   // r.z
    for (int32_t c = 0; c < k; ++c) {
      beta[c] = act[c] ? (rz[c] == 0.0 ? 0.0 : tmp[c] / rz[c]) : 0.0;
      if (act[c]) rz[c] = tmp[c];
    }
    CN_HIP(hipMemcpyAsync(st + ST_TMP * k, beta.data(), sizeof(double) * k,
                          hipMemcpyHostToDevice, s));
    k_p_update<double><<<ew_blocks(total), 256, 0, s>>>(total, k, st + ST_TMP * k, Z, P);
  }
  int32_t rows = launch_resid<double>(n, V.ptr, V.col, V.val64, k, X, d_b, R, nullptr,
                                      W->part.get(), s);
  Red rdf = reduce_rows(W, W->part.get(), rows, k, stride, 1, s);
  k_reduce_plain<<<red, kBlock, 0, s>>>(rdf.rows, k, rdf.p, st + ST_TMP * k);
  CN_HIP(hipMemcpyAsync(tmp.data(), st + ST_TMP * k, sizeof(double) * k,
                        hipMemcpyDeviceToHost, s));
  CN_HIP(hipMemcpyAsync(d_x, X, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
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
  return status;
}

}  // namespace cn

using namespace cn;

// ============================================================ entry points ==
extern "C" {

cn_status cn_system_set_tune(cn_system S, const cn_tune *t) {
  if (!S || !t || t->spmm_form < 0 || t->spmm_form > 2 || t->spmm_gens < 0 ||
      t->spmm_gens > (int32_t)kSpmmGens || t->spmm_gens == 3 || t->level_cap < 0 ||
      t->level_cap > 65536 || t->tail_rows < -1 || t->tail_rows > 16384 ||
      t->chain_levels < 0 || t->chain_levels > 24 || t->coarse_prec < 0 ||
      t->coarse_prec > 3 || t->down_cap < 0 || t->down_cap > 65536) {
    set_error("cn_system_set_tune: bad arguments");
    return CN_BAD_ARG;
  }
  S->tune = *t;
  return CN_OK;
}

cn_status cn_system_get_tune(cn_system S, cn_tune *t) {
  if (!S || !t) {
    set_error("cn_system_get_tune: bad arguments");
    return CN_BAD_ARG;
  }
  *t = S->tune;
  return CN_OK;
}

void cn_pcg_default_opts(cn_pcg_opts *o) {
  if (!o) return;
  o->precond = CN_PRECOND_IC0;
  o->precision = CN_PREC_MIXED;
  o->tol_rel = 1e-9;
  o->max_iter = 20000;
  o->inner_tol = 0.0316;  // rr falls by 1e-3 between replacements
  o->max_refine = 30;
  o->check_every = 8;
  o->warm_start = 0;
  o->graph = 1;
}

void cn_profile_enable(int32_t on) {
  g_prof.on = on != 0;
  prof_reset();
}

void cn_profile_only(const char *name) { g_prof.only = name ? name : ""; }

int32_t cn_profile_query(const char *name, double *ms_total) {
  auto it = g_prof.result.find(name ? name : "");
  if (it == g_prof.result.end()) {
    if (ms_total) *ms_total = 0.0;
    return 0;
  }
  if (ms_total) *ms_total = it->second.second;
  return it->second.first;
}

// fewest columns a fused solve pads to the vector width (below this the
// padded columns would cost more than the vector loads save)
constexpr int32_t kPadMinK = 32;

static cn_status solve_impl(cn_system S, cn_ic0 F, cn_hsc H, cn_sgs G, const cn_pcg_opts *o,
                            int32_t k,
                            const double *d_b, double *d_x, cn_solve_info *info, void *stream,
                            const char *who) {
  if (!S || !o || k <= 0 || !d_b || !d_x) {
    set_error("%s: bad arguments", who);
    return CN_BAD_ARG;
  }
  if (F && F->n != S->n_free) {
    set_error("%s: factor does not match the system", who);
    return CN_BAD_ARG;
  }
  if (H && H->n0 != S->n_free) {
    set_error("%s: hierarchy does not match the system", who);
    return CN_BAD_ARG;
  }
  if (G && (G->sys != S || G->n != S->n_free)) {
    set_error("%s: colouring does not belong to the system", who);
    return CN_BAD_ARG;
  }
  if (o->precision == CN_PREC_MIXED && !(o->inner_tol > 0.0 && o->inner_tol < 1.0)) {
    set_error("%s: inner_tol must lie in (0, 1)", who);
    return CN_BAD_ARG;
  }
  if (S->nseg > 1 && (F || o->precision == CN_PREC_FP64_PLAIN)) {
    set_error("%s: a batch of systems solves with Jacobi, HSC or SGS in CN_PREC_FP64 or "
              "CN_PREC_MIXED", who);
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (S->n_free == 0) {
    if (info) { info->status = CN_OK; info->iters = 0; info->refinements = 0; info->relres_max = 0.0; }
    return CN_OK;
  }
  if (o->precision != CN_PREC_FP64_PLAIN && o->precision != CN_PREC_FP64 &&
      o->precision != CN_PREC_MIXED) {
    set_error("%s: unknown precision %d", who, o->precision);
    return CN_BAD_ARG;
  }
  prof_reset();
  // The wide kernels move VEC columns per lane (4 FP32, 2 FP64) when k is a
  // multiple of VEC, one otherwise, and the LDS-staged products need it too:
  // at k = 765 the mixed solve took 33.5 ms against 22.3 ms at k = 768.  So
  // a fused solve of more than kPadMinK columns that are not a multiple of
  // VEC runs on a copy padded with zero columns; those have b = 0 and stay
  // inactive from the start, and every column's arithmetic is its own.
  const int32_t vec = o->precision == CN_PREC_MIXED ? 4 : 2;
  if (o->precision != CN_PREC_FP64_PLAIN && k > kPadMinK && k % vec != 0) {
    const int32_t kp = (k + vec - 1) / vec * vec;
    const size_t n = (size_t)S->n_free, need = n * kp;
    if (S->pad_b.n < need) S->pad_b.alloc(need);
    if (S->pad_x.n < need) S->pad_x.alloc(need);
    double *pb = S->pad_b.get(), *px = S->pad_x.get();
    const size_t w = sizeof(double) * k, wp = sizeof(double) * kp;
    CN_HIP(hipMemcpy2DAsync(pb, wp, d_b, w, w, n, hipMemcpyDeviceToDevice, s));
    CN_HIP(hipMemset2DAsync(pb + k, wp, 0, wp - w, n, s));
    if (o->warm_start) {
      CN_HIP(hipMemcpy2DAsync(px, wp, d_x, w, w, n, hipMemcpyDeviceToDevice, s));
      CN_HIP(hipMemset2DAsync(px + k, wp, 0, wp - w, n, s));
    }
    cn_status st = o->precision == CN_PREC_MIXED
                       ? fused_solve<float>(S, F, H, G, o, kp, pb, px, info, s)
                       : fused_solve<double>(S, F, H, G, o, kp, pb, px, info, s);
    CN_HIP(hipMemcpy2DAsync(d_x, w, px, wp, w, n, hipMemcpyDeviceToDevice, s));
    CN_HIP(hipStreamSynchronize(s));
    prof_collect();
    return st;
  }
  cn_status st;
  switch (o->precision) {
    case CN_PREC_FP64_PLAIN: st = plain_solve(S, F, H, G, o, k, d_b, d_x, info, s); break;
    case CN_PREC_FP64: st = fused_solve<double>(S, F, H, G, o, k, d_b, d_x, info, s); break;
    default: st = fused_solve<float>(S, F, H, G, o, k, d_b, d_x, info, s); break;
  }
  prof_collect();
  return st;
}

cn_status cn_pcg_solve(cn_system S, cn_ic0 F, const cn_pcg_opts *o, int32_t k,
                       const double *d_b, double *d_x, cn_solve_info *info,
                       void *stream) {
  CN_API_BEGIN
  if (o && o->precond == CN_PRECOND_IC0 && !F) {
    set_error("cn_pcg_solve: IC0 preconditioner needs a factor");
    return CN_BAD_ARG;
  }
  if (o && o->precond == CN_PRECOND_HSC) {
    set_error("cn_pcg_solve: the HSC preconditioner is applied by cn_pcg_solve_hsc");
    return CN_BAD_ARG;
  }
  if (o && o->precond == CN_PRECOND_SGS) {
    set_error("cn_pcg_solve: the SGS preconditioner is applied by cn_pcg_solve_sgs");
    return CN_BAD_ARG;
  }
  if (o && o->precond != CN_PRECOND_IC0) F = nullptr;
  return solve_impl(S, F, nullptr, nullptr, o, k, d_b, d_x, info, stream, "cn_pcg_solve");
  CN_API_END
}

cn_status cn_pcg_solve_hsc(cn_system S, cn_hsc H, const cn_pcg_opts *o, int32_t k,
                           const double *d_b, double *d_x, cn_solve_info *info,
                           void *stream) {
  CN_API_BEGIN
  if (!H) {
    set_error("cn_pcg_solve_hsc: no hierarchy");
    return CN_BAD_ARG;
  }
  return solve_impl(S, nullptr, H, nullptr, o, k, d_b, d_x, info, stream, "cn_pcg_solve_hsc");
  CN_API_END
}

cn_status cn_pcg_solve_sgs(cn_system S, cn_sgs G, const cn_pcg_opts *o, int32_t k,
                           const double *d_b, double *d_x, cn_solve_info *info,
                           void *stream) {
  CN_API_BEGIN
  if (!G) {
    set_error("cn_pcg_solve_sgs: no colouring");
    return CN_BAD_ARG;
  }
  return solve_impl(S, nullptr, nullptr, G, o, k, d_b, d_x, info, stream, "cn_pcg_solve_sgs");
  CN_API_END
}

}  // extern "C"
