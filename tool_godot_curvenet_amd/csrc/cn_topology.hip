// cn_topology.hip -- halfedge build, seam cut and knot projection on gfx950.
//
// All of this is integer / index work, bound by HBM traffic and atomics
// rather than arithmetic: one thread per halfedge, corner or vertex, with
// per-vertex buckets (counting sort + in-bucket sort) instead of the spec's
// quadratic scans.  Outputs are index-identical to the spec definitions.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdarg>
#include <cstring>

#include "cn_internal.h"

namespace cn {

static thread_local char g_err[512];

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ------------------------------------------------------------------ scan --
// Exclusive scan: each block scans 1024 items (4 per thread), block totals
// are scanned recursively and added back.
template <typename T>
__global__ void __launch_bounds__(256) k_scan_blocks(T *d, int64_t n, T *sums) {
  __shared__ T sh[256];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * 1024 + 4 * t;
  T v[4];
  T run = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = (base + i < n) ? d[base + i] : T(0);
    T x = v[i];
    v[i] = run;
    run += x;
  }
  sh[t] = run;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    T add = (t >= off) ? sh[t - off] : T(0);
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  T excl = sh[t] - run;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (base + i < n) d[base + i] = v[i] + excl;
  if (t == 255) sums[blockIdx.x] = sh[255];
}

template <typename T>
__global__ void k_scan_add(T *d, int64_t n, const T *sums) {
  int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  T off = sums[blockIdx.x];
  for (int r = 0; r < 4; ++r, i += 256)
    if (i < n) d[i] += off;
}

template <typename T>
static void scan_rec(T *d, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t nb = (n + 1023) / 1024;
  DevBuf<T> sums(nb);
  k_scan_blocks<T><<<(unsigned)nb, 256, 0, s>>>(d, n, sums.get());
  CN_HIP(hipGetLastError());
  if (nb > 1) {
    scan_rec<T>(sums.get(), nb, s);
    k_scan_add<T><<<(unsigned)nb, 256, 0, s>>>(d, n, sums.get());
    CN_HIP(hipGetLastError());
  }
  CN_HIP(hipStreamSynchronize(s));  // `sums` is freed on return
}

template <typename T>
static int64_t scan_total(T *d, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  T last = 0, lastx = 0;
  CN_HIP(hipMemcpyAsync(&last, d + n - 1, sizeof(T), hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  scan_rec<T>(d, n, s);
  CN_HIP(hipMemcpyAsync(&lastx, d + n - 1, sizeof(T), hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  return (int64_t)lastx + (int64_t)last;
}

int64_t exclusive_scan_i32(int32_t *d, int64_t n, hipStream_t s) {
  return scan_total<int32_t>(d, n, s);
}
int64_t exclusive_scan_i64(int64_t *d, int64_t n, hipStream_t s) {
  return scan_total<int64_t>(d, n, s);
}

// -------------------------------------------------------------- buckets --
__device__ __forceinline__ int32_t src_of(const int32_t *tris, int32_t h) {
  return tris[h];  // slot h = 3f+l holds t_l, the source of halfedge 3f+l
}
__device__ __forceinline__ int32_t tgt_of(const int32_t *tris, int32_t h) {
  int32_t f = h / 3, l = h - 3 * f;
  return tris[3 * f + (l == 2 ? 0 : l + 1)];
}

__global__ void k_bucket_count(int32_t nint, const int32_t *tris, int by_target,
                               int32_t *cnt) {
  int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nint) return;
  int32_t v = by_target ? tgt_of(tris, h) : src_of(tris, h);
  atomicAdd(&cnt[v], 1);
}

__global__ void k_bucket_fill(int32_t nint, const int32_t *tris, int by_target,
                              int32_t *fill, int32_t *he) {
  int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nint) return;
  int32_t v = by_target ? tgt_of(tris, h) : src_of(tris, h);
  he[atomicAdd(&fill[v], 1)] = h;
}

// atomics leave each bucket unordered: sort it (buckets are one-ring sized)
__global__ void k_bucket_sort(int32_t nv, const int32_t *ptr, int32_t *he) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int32_t a = ptr[v], b = ptr[v + 1];
  for (int32_t i = a + 1; i < b; ++i) {
    int32_t x = he[i], j = i - 1;
    while (j >= a && he[j] > x) { he[j + 1] = he[j]; --j; }
    he[j + 1] = x;
  }
}

static void build_buckets(int32_t nv, int32_t nf, const int32_t *tris,
                          int by_target, Buckets &b, hipStream_t s) {
  int32_t nint = 3 * nf;
  b.ptr.alloc((size_t)nv + 1);
  b.he.alloc((size_t)nint + 1);
  CN_HIP(hipMemsetAsync(b.ptr.get(), 0, sizeof(int32_t) * ((size_t)nv + 1), s));
  k_bucket_count<<<grid_for(nint, 256), 256, 0, s>>>(nint, tris, by_target,
                                                     b.ptr.get());
  CN_HIP(hipGetLastError());
  exclusive_scan_i32(b.ptr.get(), (int64_t)nv + 1, s);
  DevBuf<int32_t> fill((size_t)nv + 1);
  CN_HIP(hipMemcpyAsync(fill.get(), b.ptr.get(), sizeof(int32_t) * ((size_t)nv + 1),
                        hipMemcpyDeviceToDevice, s));
  k_bucket_fill<<<grid_for(nint, 256), 256, 0, s>>>(nint, tris, by_target,
                                                    fill.get(), b.he.get());
  k_bucket_sort<<<grid_for(nv, 256), 256, 0, s>>>(nv, b.ptr.get(), b.he.get());
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
}

void build_incoming(int32_t nv, int32_t nf, const int32_t *d_tris, Buckets &b,
                    hipStream_t s) {
  build_buckets(nv, nf, d_tris, 1, b, s);
}
void build_outgoing(int32_t nv, int32_t nf, const int32_t *d_tris, Buckets &b,
                    hipStream_t s) {
  build_buckets(nv, nf, d_tris, 0, b, s);
}

// ------------------------------------------------------------ halfedges --
__global__ void k_check_tris(int64_t n3, const int32_t *tris, int32_t nv,
                             int32_t *bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  int32_t v = tris[i];
  if (v < 0 || v >= nv) atomicOr(bad, 1);
}

// step 1 of fromTriangles (HalfedgeBuilder.lean:39-45)
__global__ void k_he_interior(int32_t nf, const int32_t *tris, int32_t *target,
                              int32_t *twin, int32_t *next, int32_t *face) {
  int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= 3 * nf) return;
  int32_t f = h / 3, l = h - 3 * f;
  target[h] = tris[3 * f + (l == 2 ? 0 : l + 1)];
  twin[h] = -1;
  next[h] = (l == 2) ? 3 * f : h + 1;
  face[h] = f;
}

// step 2 (:46-62): first interior j with target = src(h), source = tgt(h).
// Candidates are exactly the bucket out(tgt(h)), scanned in ascending order.
__global__ void k_he_twins(int32_t nint, const int32_t *tris,
                           const int32_t *optr, const int32_t *ohe,
                           int32_t *twin, int32_t *orphan) {
  int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nint) return;
  int32_t s = src_of(tris, h), t = tgt_of(tris, h);
  int32_t found = -1;
  for (int32_t a = optr[t], b = optr[t + 1]; a < b; ++a) {
    int32_t j = ohe[a];
    if (tgt_of(tris, j) == s) { found = j; break; }
  }
  twin[h] = found;
  orphan[h] = found < 0 ? 1 : 0;
}

// step 3 (:63-71): boundary halfedge per orphan, in orphan order
__global__ void k_he_boundary(int32_t nint, const int32_t *tris,
                              const int32_t *orphan_pos, const int32_t *twin_in,
                              int32_t *target, int32_t *twin, int32_t *next,
                              int32_t *face, int32_t *first_by_src) {
  int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nint) return;
  if (twin_in[h] >= 0) return;
  int32_t bh = nint + orphan_pos[h];
  target[bh] = src_of(tris, h);
  twin[bh] = h;
  next[bh] = bh;
  face[bh] = -1;
  twin[h] = bh;
  // source(bh) = target(h)
  atomicMin(&first_by_src[tgt_of(tris, h)], bh);
}

// step 4 (:72-89): next = first boundary halfedge whose source is my target
__global__ void k_he_boundary_next(int32_t nint, int32_t nh,
                                   const int32_t *target,
                                   const int32_t *first_by_src, int32_t *next) {
  int32_t bh = nint + blockIdx.x * blockDim.x + threadIdx.x;
  if (bh >= nh) return;
  int32_t nx = first_by_src[target[bh]];
  next[bh] = (nx == INT_MAX) ? bh : nx;
}

void build_halfedges(int32_t nv, int32_t nf, const int32_t *d_tris,
                     int32_t *target, int32_t *twin, int32_t *next,
                     int32_t *face, int32_t *nh_out, hipStream_t s) {
  int32_t nint = 3 * nf;
  Buckets out;
  build_outgoing(nv, nf, d_tris, out, s);
  k_he_interior<<<grid_for(nint, 256), 256, 0, s>>>(nf, d_tris, target, twin,
                                                   next, face);
  DevBuf<int32_t> orphan((size_t)nint + 1), tw0((size_t)nint + 1);
  k_he_twins<<<grid_for(nint, 256), 256, 0, s>>>(nint, d_tris, out.ptr.get(),
                                                out.he.get(), tw0.get(),
                                                orphan.get());
  CN_HIP(hipGetLastError());
  CN_HIP(hipMemcpyAsync(twin, tw0.get(), sizeof(int32_t) * (size_t)nint,
                        hipMemcpyDeviceToDevice, s));
  int64_t nb = exclusive_scan_i32(orphan.get(), nint, s);
  int32_t nh = nint + (int32_t)nb;
  DevBuf<int32_t> first((size_t)nv + 1);
  std::vector<int32_t> big((size_t)nv + 1, INT_MAX);
  CN_HIP(hipMemcpyAsync(first.get(), big.data(), sizeof(int32_t) * ((size_t)nv + 1),
                        hipMemcpyHostToDevice, s));
  k_he_boundary<<<grid_for(nint, 256), 256, 0, s>>>(
      nint, d_tris, orphan.get(), tw0.get(), target, twin, next, face,
      first.get());
  if (nb > 0)
    k_he_boundary_next<<<grid_for(nb, 256), 256, 0, s>>>(nint, nh, target,
                                                        first.get(), next);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  *nh_out = nh;
}

// -------------------------------------------------------------- seam cut --
__device__ __forceinline__ int32_t prev_in_tri(int32_t h) {
  int32_t f = h / 3, l = h - 3 * f;
  return 3 * f + (l == 0 ? 2 : l - 1);
}
// slot (3f+l) of the corner at which interior halfedge p arrives
__device__ __forceinline__ int32_t corner_slot(int32_t p) {
  int32_t f = p / 3, l = p - 3 * f;
  return 3 * f + (l == 2 ? 0 : l + 1);
}

__global__ void k_mark_cut(int32_t ncut, const int32_t *edges, int32_t nv,
                           const int32_t *tris, const int32_t *optr,
                           const int32_t *ohe, uint8_t *is_cut,
                           uint8_t *on_cut, int32_t *bad) {
  int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ncut) return;
  int32_t u = edges[2 * e], v = edges[2 * e + 1];
  if (u < 0 || v < 0 || u >= nv || v >= nv || u == v) { atomicOr(bad, 1); return; }
  int found = 0;
  for (int dir = 0; dir < 2; ++dir) {
    int32_t a = dir ? v : u, b = dir ? u : v;
    for (int32_t q = optr[a], qe = optr[a + 1]; q < qe; ++q) {
      int32_t h = ohe[q];
      if (tgt_of(tris, h) == b) { is_cut[h] = 1; found = 1; }
    }
  }
  if (!found) { atomicOr(bad, 2); return; }
  on_cut[u] = 1;
  on_cut[v] = 1;
}

// Rotation about the corner vertex.  Backward step from incoming p crosses
// edge (source p, v); forward step crosses edge (v, target next p).  A step
// stops at a boundary edge or a cut edge.
__device__ __forceinline__ int32_t rot_back(int32_t p, const int32_t *twin,
                                            const int32_t *face,
                                            const uint8_t *is_cut) {
  int32_t t = twin[p];
  if (t < 0 || face[t] < 0 || is_cut[p]) return -1;
  return prev_in_tri(t);
}
__device__ __forceinline__ int32_t rot_fwd(int32_t p, const int32_t *twin,
                                           const int32_t *next,
                                           const int32_t *face,
                                           const uint8_t *is_cut) {
  int32_t o = next[p];
  int32_t t = twin[o];
  if (t < 0 || face[t] < 0 || is_cut[o]) return -1;
  return t;
}

// wedge representative = corner where the backward walk stops (for a full
// cycle: the smallest halfedge of the cycle); also the wedge's min slot.
__global__ void k_wedges(int32_t nint, const int32_t *tris,
                         const int32_t *twin, const int32_t *next,
                         const int32_t *face, const uint8_t *is_cut,
                         const uint8_t *on_cut, const int32_t *iptr,
                         int32_t *wstart, int32_t *wmin) {
  int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nint) return;
  int32_t v = tgt_of(tris, p);
  if (!on_cut[v]) { wstart[p] = p; wmin[p] = corner_slot(p); return; }
  int32_t deg = iptr[v + 1] - iptr[v];
  int32_t cur = p, start = -1, cyc_min = p;
  for (int32_t step = 0; step <= deg; ++step) {
    int32_t b = rot_back(cur, twin, face, is_cut);
    if (b < 0) { start = cur; break; }
    cur = b;
    if (cur < cyc_min) cyc_min = cur;
    if (cur == p) { start = cyc_min; break; }
  }
  if (start < 0) start = cyc_min;  // non-manifold fan: degrade gracefully
  int32_t mn = corner_slot(start);
  cur = start;
  for (int32_t step = 0; step <= deg; ++step) {
    int32_t f = rot_fwd(cur, twin, next, face, is_cut);
    if (f < 0 || f == start) break;
    cur = f;
    int32_t sl = corner_slot(cur);
    if (sl < mn) mn = sl;
  }
  wstart[p] = start;
  wmin[p] = mn;
}

__global__ void k_count_wedges(int32_t nv, const uint8_t *on_cut,
                               const int32_t *iptr, const int32_t *ihe,
                               const int32_t *wstart, int32_t *extra) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int32_t c = 0;
  if (on_cut[v])
    for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a)
      if (wstart[ihe[a]] == ihe[a]) ++c;
  extra[v] = c > 1 ? c - 1 : 0;
}

__global__ void k_assign_copies(int32_t nint, int32_t nv, const int32_t *tris,
                                const uint8_t *on_cut, const int32_t *iptr,
                                const int32_t *ihe, const int32_t *wstart,
                                const int32_t *wmin, const int32_t *off,
                                int32_t *new_tris, int32_t *origin,
                                uint8_t *constrained) {
  int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nint) return;
  int32_t v = tgt_of(tris, p);
  int32_t slot = corner_slot(p);
  if (!on_cut[v]) { new_tris[slot] = v; return; }
  int32_t mine = wmin[p], rank = 0;
  for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
    int32_t q = ihe[a];
    if (wstart[q] == q && wmin[q] < mine) ++rank;
  }
  int32_t id = rank == 0 ? v : nv + off[v] + rank - 1;
  new_tris[slot] = id;
  if (rank > 0) { origin[id] = v; constrained[id] = 1; }
}

__global__ void k_init_origin(int32_t nv, const uint8_t *on_cut,
                              int32_t *origin, uint8_t *constrained) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  origin[v] = v;
  constrained[v] = on_cut[v];
}

// ------------------------------------------------------------ projection --
// SurfaceProjection.projectToVertices (SurfaceProjection.lean:36-64)
__global__ void k_project(int32_t nv, const double *pos, int32_t nq,
                          const double *q, int32_t *out) {
#pragma clang fp contract(off)
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  double kx = q[3 * i], ky = q[3 * i + 1], kz = q[3 * i + 2];
  double best = 1e308;
  int32_t bi = -1;
  for (int32_t v = 0; v < nv; ++v) {
    double dx = pos[3 * v] - kx, dy = pos[3 * v + 1] - ky, dz = pos[3 * v + 2] - kz;
    double d = dx * dx + dy * dy + dz * dz;
    if (d < best) { best = d; bi = v; }
  }
  out[i] = bi;
}


// SurfaceProjection.promoteVertexSamples (SurfaceProjection.lean:64-80).
// The spec walks the knots in order; here the same columns come from
// (1) the first knot of every vertex (atomicMin of the knot index),
// (2) an exclusive scan of the "first knot" flags -- a first knot's column is
// the number of first knots before it, which is the spec's running
// `nextCol` -- and (3) every other knot copying its vertex's first column.
__global__ void k_pv_init(int32_t nv, int32_t *vtc) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) vtc[v] = INT_MAX;
}

__global__ void k_pv_first(int32_t nq, const int32_t *idx, int32_t nv, int32_t *vtc,
                           int32_t *bad) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int32_t v = idx[i];
  if (v >= nv) *bad = 1;  // every writer stores 1
  else if (v >= 0) atomicMin(&vtc[v], i);
}

__global__ void k_pv_flag(int32_t nq, const int32_t *idx, int32_t nv, const int32_t *vtc,
                          int32_t *ktc) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int32_t v = idx[i];
  ktc[i] = (v >= 0 && v < nv && vtc[v] == i) ? 1 : 0;
}

// after the scan a first knot holds its column; the others take their
// vertex's (only non-first knots are written, only first knots are read)
__global__ void k_pv_share(int32_t nq, const int32_t *idx, int32_t nv, const int32_t *vtc,
                           int32_t *ktc) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int32_t v = idx[i];
  if (v < 0 || v >= nv) { ktc[i] = -1; return; }
  int32_t f = vtc[v];
  if (f != i) ktc[i] = ktc[f];
}

__global__ void k_pv_vertex_col(int32_t nv, int32_t *vtc, const int32_t *ktc) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int32_t f = vtc[v];
  vtc[v] = (f == INT_MAX) ? -1 : ktc[f];
}

}  // namespace cn

using namespace cn;

extern "C" {

const char *cn_last_error(void) { return cn::g_err; }
int32_t cn_version(void) { return 100; }

cn_status cn_device_check(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_error("no HIP device visible");
    return CN_NO_DEVICE;
  }
  return CN_OK;
}

static cn_status check_tris(int32_t nv, int32_t nf, const int32_t *d_tris,
                            hipStream_t s) {
  DevBuf<int32_t> bad(1);
  CN_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
  if (nf > 0)
    k_check_tris<<<grid_for(3ll * nf, 256), 256, 0, s>>>(3ll * nf, d_tris, nv,
                                                        bad.get());
  int32_t h = 0;
  CN_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  if (h) {
    set_error("triangle vertex index out of range [0, %d)", nv);
    return CN_BAD_ARG;
  }
  return CN_OK;
}

cn_status cn_halfedge_from_triangles(int32_t nv, int32_t nf,
                                     const int32_t *d_tris, int32_t *d_target,
                                     int32_t *d_twin, int32_t *d_next,
                                     int32_t *d_face, int32_t *h_nh,
                                     void *stream) {
  CN_API_BEGIN
  if (nv < 0 || nf < 0 || (nf > 0 && (!d_tris || !d_target || !d_twin || !d_next || !d_face)) || !h_nh) {
    set_error("cn_halfedge_from_triangles: bad arguments");
    return CN_BAD_ARG;
  }
  if (nf > (INT_MAX / 6)) { set_error("too many faces"); return CN_BAD_ARG; }
  hipStream_t s = as_stream(stream);
  cn_status st = check_tris(nv, nf, d_tris, s);
  if (st != CN_OK) return st;
  build_halfedges(nv, nf, d_tris, d_target, d_twin, d_next, d_face, h_nh, s);
  return CN_OK;
  CN_API_END
}

cn_status cn_cut_seams(int32_t nv, int32_t nf, const int32_t *d_tris,
                       int32_t ncut, const int32_t *d_cut_edges,
                       int32_t *d_new_tris, int32_t *d_origin,
                       uint8_t *d_constrained, int32_t *h_nv_out,
                       void *stream) {
  CN_API_BEGIN
  if (nv < 0 || nf < 0 || ncut < 0 || !h_nv_out || (nf > 0 && (!d_tris || !d_new_tris)) ||
      (ncut > 0 && !d_cut_edges) || (nv > 0 && (!d_origin || !d_constrained))) {
    set_error("cn_cut_seams: bad arguments");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  cn_status st = check_tris(nv, nf, d_tris, s);
  if (st != CN_OK) return st;
  int32_t nint = 3 * nf;
  DevBuf<int32_t> tg((size_t)2 * nint + 1), tw((size_t)2 * nint + 1),
      nx((size_t)2 * nint + 1), fc((size_t)2 * nint + 1);
  int32_t nh = 0;
  build_halfedges(nv, nf, d_tris, tg.get(), tw.get(), nx.get(), fc.get(), &nh, s);
  Buckets out, inc;
  build_outgoing(nv, nf, d_tris, out, s);
  build_incoming(nv, nf, d_tris, inc, s);
  DevBuf<uint8_t> is_cut((size_t)nint + 1), on_cut((size_t)nv + 1);
  DevBuf<int32_t> bad(1);
  CN_HIP(hipMemsetAsync(is_cut.get(), 0, (size_t)nint + 1, s));
  CN_HIP(hipMemsetAsync(on_cut.get(), 0, (size_t)nv + 1, s));
  CN_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
  if (ncut > 0)
    k_mark_cut<<<grid_for(ncut, 256), 256, 0, s>>>(
        ncut, d_cut_edges, nv, d_tris, out.ptr.get(), out.he.get(),
        is_cut.get(), on_cut.get(), bad.get());
  int32_t hbad = 0;
  CN_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CN_HIP(hipStreamSynchronize(s));
  if (hbad) {
    set_error("cn_cut_seams: cut edge %s", (hbad & 1) ? "index invalid" : "is not a mesh edge");
    return CN_BAD_ARG;
  }
  DevBuf<int32_t> wstart((size_t)nint + 1), wmin((size_t)nint + 1),
      extra((size_t)nv + 1);
  if (nint > 0)
    k_wedges<<<grid_for(nint, 256), 256, 0, s>>>(
        nint, d_tris, tw.get(), nx.get(), fc.get(), is_cut.get(), on_cut.get(),
        inc.ptr.get(), wstart.get(), wmin.get());
  if (nv > 0)
    k_count_wedges<<<grid_for(nv, 256), 256, 0, s>>>(
        nv, on_cut.get(), inc.ptr.get(), inc.he.get(), wstart.get(), extra.get());
  CN_HIP(hipGetLastError());
  int64_t nextra = exclusive_scan_i32(extra.get(), nv, s);
  if (nextra > 2ll * ncut) {  // only a non-manifold fan can split this much
    set_error("cn_cut_seams: %lld wedge copies exceed 2*ncut (non-manifold vertex?)",
              (long long)nextra);
    return CN_NOT_MANIFOLD;
  }
  if (nv > 0)
    k_init_origin<<<grid_for(nv, 256), 256, 0, s>>>(nv, on_cut.get(), d_origin,
                                                   d_constrained);
  if (nint > 0)
    k_assign_copies<<<grid_for(nint, 256), 256, 0, s>>>(
        nint, nv, d_tris, on_cut.get(), inc.ptr.get(), inc.he.get(),
        wstart.get(), wmin.get(), extra.get(), d_new_tris, d_origin,
        d_constrained);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  *h_nv_out = nv + (int32_t)nextra;
  return CN_OK;
  CN_API_END
}

cn_status cn_project_to_vertices(int32_t nv, const double *d_pos, int32_t nq,
                                 const double *d_q, int32_t *d_index,
                                 void *stream) {
  CN_API_BEGIN
  if (nv < 0 || nq < 0 || (nq > 0 && (!d_q || !d_index)) || (nv > 0 && !d_pos)) {
    set_error("cn_project_to_vertices: bad arguments");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  if (nq > 0)
    k_project<<<grid_for(nq, 64), 64, 0, s>>>(nv, d_pos, nq, d_q, d_index);
  CN_HIP(hipGetLastError());
  CN_HIP(hipStreamSynchronize(s));
  return CN_OK;
  CN_API_END
}

cn_status cn_promote_vertex_samples(int32_t nv, int32_t nq, const int32_t *d_index,
                                    int32_t *d_knot_to_col, int32_t *d_vertex_to_col,
                                    int32_t *h_ncols, void *stream) {
  CN_API_BEGIN
  if (nv < 0 || nq < 0 || !h_ncols || (nq > 0 && (!d_index || !d_knot_to_col)) ||
      (nv > 0 && !d_vertex_to_col)) {
    set_error("cn_promote_vertex_samples: bad arguments");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  *h_ncols = 0;
  if (nv > 0) k_pv_init<<<grid_for(nv, 256), 256, 0, s>>>(nv, d_vertex_to_col);
  CN_HIP(hipGetLastError());
  if (nq > 0) {
    DevBuf<int32_t> bad(1);
    CN_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), s));
    k_pv_first<<<grid_for(nq, 256), 256, 0, s>>>(nq, d_index, nv, d_vertex_to_col, bad.get());
    k_pv_flag<<<grid_for(nq, 256), 256, 0, s>>>(nq, d_index, nv, d_vertex_to_col,
                                                d_knot_to_col);
    CN_HIP(hipGetLastError());
    int64_t ncols = exclusive_scan_i32(d_knot_to_col, nq, s);  // synchronises
    k_pv_share<<<grid_for(nq, 256), 256, 0, s>>>(nq, d_index, nv, d_vertex_to_col,
                                                 d_knot_to_col);
    CN_HIP(hipGetLastError());
    if (nv > 0)
      k_pv_vertex_col<<<grid_for(nv, 256), 256, 0, s>>>(nv, d_vertex_to_col, d_knot_to_col);
    CN_HIP(hipGetLastError());
    int32_t hbad = 0;
    CN_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CN_HIP(hipStreamSynchronize(s));
    if (hbad) {
      set_error("cn_promote_vertex_samples: a knot projects to a vertex index >= nv");
      return CN_BAD_ARG;
    }
    *h_ncols = (int32_t)ncols;
  } else if (nv > 0) {
    k_pv_vertex_col<<<grid_for(nv, 256), 256, 0, s>>>(nv, d_vertex_to_col, nullptr);
    CN_HIP(hipGetLastError());
    CN_HIP(hipStreamSynchronize(s));
  }
  return CN_OK;
  CN_API_END
}

// ---- host-side surgery (CutAlgorithm.lean) ----------------------------
cn_status cn_cut_subdivide_edge(int32_t *h_nv, int32_t *h_nf, int32_t *h_nh,
                                int32_t h, int32_t *target, int32_t *twin,
                                int32_t *next, int32_t *face) {
  (void)h_nf;
  if (!h_nv || !h_nh || h < 0 || h >= *h_nh) { set_error("bad halfedge"); return CN_BAD_ARG; }
  int32_t nh = *h_nh;
  int32_t ot = target[h], onx = next[h], ofc = face[h];
  int32_t ti = twin[h] >= 0 ? twin[h] : h;
  int32_t tt = target[ti], tnx = next[ti], tfc = face[ti];
  int32_t c = *h_nv, hn = nh, tn = nh + 1;
  target[h] = c; twin[h] = tn; next[h] = hn; face[h] = ofc;
  target[ti] = c; twin[ti] = hn; next[ti] = tn; face[ti] = tfc;
  target[hn] = ot; twin[hn] = ti; next[hn] = onx; face[hn] = ofc;
  target[tn] = tt; twin[tn] = h; next[tn] = tnx; face[tn] = tfc;
  *h_nv += 1;
  *h_nh = nh + 2;
  return CN_OK;
}

cn_status cn_cut_split_face(int32_t *h_nv, int32_t *h_nf, int32_t *h_nh,
                            int32_t a, int32_t b, int32_t *target,
                            int32_t *twin, int32_t *next, int32_t *face) {
  (void)h_nv;
  if (!h_nf || !h_nh || a < 0 || b < 0 || a >= *h_nh || b >= *h_nh || a == b) {
    set_error("bad halfedges");
    return CN_BAD_ARG;
  }
  int32_t nh = *h_nh, fn = *h_nf, hab = nh, hba = nh + 1;
  int32_t at = target[a], an = next[a], af = face[a];
  int32_t bt = target[b], bn = next[b];
  int32_t cur = bn, steps = 0;
  while (cur != a && steps < nh) {
    int32_t n2 = next[cur];
    face[cur] = fn;
    cur = n2;
    ++steps;
  }
  face[a] = fn; next[a] = hab;
  next[b] = hba;
  target[hab] = bt; twin[hab] = hba; next[hab] = bn; face[hab] = fn;
  target[hba] = at; twin[hba] = hab; next[hba] = an; face[hba] = af;
  *h_nf = fn + 1;
  *h_nh = nh + 2;
  return CN_OK;
}

cn_status cn_cut_insert_crack(int32_t *h_nv, int32_t *h_nf, int32_t *h_nh,
                              int32_t h_anchor, int32_t *target, int32_t *twin,
                              int32_t *next, int32_t *face) {
  (void)h_nf;
  if (!h_nv || !h_nh || h_anchor < 0 || h_anchor >= *h_nh) { set_error("bad halfedge"); return CN_BAD_ARG; }
  int32_t nh = *h_nh;
  int32_t vb = target[h_anchor], on = next[h_anchor], fa = face[h_anchor];
  int32_t c = *h_nv, hout = nh, hin = nh + 1;
  next[h_anchor] = hout;
  target[hout] = c; twin[hout] = hin; next[hout] = hin; face[hout] = fa;
  target[hin] = vb; twin[hin] = hout; next[hin] = on; face[hin] = fa;
  *h_nv += 1;
  *h_nh = nh + 2;
  return CN_OK;
}

}  // extern "C"
