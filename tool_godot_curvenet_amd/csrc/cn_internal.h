// cn_internal.h -- shared internals of the gfx950 curvenet linear-algebra
// library: error plumbing, device buffers, handle structs, kernel helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/curvenet_la.h"

namespace cn {

// Last error message of the calling thread (cn_last_error()).
void set_error(const char *fmt, ...);

struct HipError {
  hipError_t err;
  const char *what;
};

#define CN_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) throw ::cn::HipError{e_, #call};                  \
  } while (0)

// Converts exceptions thrown inside an entry point into a cn_status.
#define CN_API_BEGIN try {
#define CN_API_END                                                          \
  }                                                                         \
  catch (const ::cn::HipError &e) {                                         \
    ::cn::set_error("HIP error %d (%s) in %s", (int)e.err,                 \
                    hipGetErrorString(e.err), e.what);                      \
    return CN_HIP_ERROR;                                                    \
  }                                                                         \
  catch (const std::bad_alloc &) {                                          \
    ::cn::set_error("host allocation failed");                             \
    return CN_OOM;                                                          \
  }

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Owning device buffer (hipMalloc / hipFree).  Setup-time objects only; the
// solver's per-iteration work never allocates.
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) CN_HIP(hipMalloc(&p, sizeof(T) * count));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T *get() const { return p; }
};

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 0x7fffffff) g = 0x7fffffff;
  return (unsigned)g;
}

// ---------------------------------------------------------------- scans --
// Exclusive prefix sum of int32 counts in place; returns the total.  `tmp`
// is scratch (grown as needed).  Synchronises the stream to read the total.
int64_t exclusive_scan_i32(int32_t *d, int64_t n, hipStream_t s);
int64_t exclusive_scan_i64(int64_t *d, int64_t n, hipStream_t s);

// HSC apply launch shapes (from a system's cn_tune; 0 = built-in)
struct HscShape {
  int32_t level_cap = 0;
  int32_t tail_rows = 0;
  int32_t chain_levels = 0;
  int32_t coarse_prec = 0;  // 2: FP32 coarse product in FP32 applies, 3: split bfloat16
  int32_t down_cap = 0;     // the down sweep's grid cap (0: level_cap)
};

// ----------------------------------------------------------- topology --
// Device halfedge mesh in SoA form (the spec's Halfedge record,
// Halfedge.lean:30-35, with -1 for `none`).
struct HeMesh {
  int32_t nv = 0, nf = 0, nh = 0;
  DevBuf<int32_t> target, twin, next, face;
};

// Build the halfedge mesh of a triangle list on the device
// (HalfedgeBuilder.fromTriangles semantics, index-identical).
void build_halfedges(int32_t nv, int32_t nf, const int32_t *d_tris,
                     int32_t *target, int32_t *twin, int32_t *next,
                     int32_t *face, int32_t *nh_out, hipStream_t s);

// Vertex buckets: for every vertex, the ascending list of interior
// halfedges whose target (or source) is the vertex.
struct Buckets {
  DevBuf<int32_t> ptr;  // nv + 1
  DevBuf<int32_t> he;   // 3 nf
};
void build_incoming(int32_t nv, int32_t nf, const int32_t *d_tris, Buckets &b,
                    hipStream_t s);  // by target (interior halfedges)
void build_outgoing(int32_t nv, int32_t nf, const int32_t *d_tris, Buckets &b,
                    hipStream_t s);  // by source (interior halfedges)

struct SolverWork;
void free_solver_work(SolverWork *w);

// A serial number for every preconditioner handle, never reused: cached
// captured graphs name the handles they were captured with by serial, since
// a new handle may get the address of a destroyed one.
inline uint64_t next_serial() {
  static std::atomic<uint64_t> c{1};
  return c.fetch_add(1);
}

}  // namespace cn

// ------------------------------------------------------------ handles --
// The reduced cut-Laplacian system A_ff (free x free) and the coupling
// block A_fc (free x constrained), plus the cut mesh they came from.
namespace cn {
// GraphColoring.greedyColor (GraphColoring.lean:41-56) of a CSR pattern:
// vertex order, smallest colour not used by an earlier neighbour.  Returns
// the number of colours.
int32_t greedy_color(int32_t n, const std::vector<int32_t> &ptr,
                     const std::vector<int32_t> &col, std::vector<int32_t> &color);
}  // namespace cn
struct cn_system_t;
namespace cn {
// one segment over every row of a single system, if it has no segments yet
void ensure_segments(cn_system_t *S, hipStream_t s);
// the right-hand side table of batch B from its members' (cn_harmonic_rhs
// over the batch then builds every member's right-hand side in one launch)
void batch_rhs_table(cn_system_t *B, const cn_system_t *const *members, int32_t nsys,
                     hipStream_t s);
}  // namespace cn

struct cn_system_t {
  int32_t nv = 0, nf = 0, n_free = 0, n_con = 0;
  int32_t lap_kind = 0;             // CN_LAP_COT or CN_LAP_ROBUST
  double lap_delta = 0.0;           // robust margin
  int64_t nnz_ff = 0, nnz_fc = 0;
  cn::DevBuf<int32_t> tris;         // cut-mesh triangles (3 nf)
  cn::DevBuf<uint8_t> constrained;  // nv
  cn::DevBuf<double> pos;           // 3 nv
  cn::HeMesh he;                    // halfedges of the cut mesh
  cn::Buckets inc;                  // incoming interior halfedges per vertex
  cn::DevBuf<int32_t> free_of, con_of, free_vertex, con_vertex;
  cn::DevBuf<double> face_m;        // 9 per face: triangle cot Laplacian
  cn::DevBuf<int32_t> ff_ptr, ff_col, fc_ptr, fc_col;
  cn::DevBuf<double> ff_val, fc_val;
  cn::DevBuf<float> ff_val32, fc_val32;
  cn::DevBuf<double> diag;          // diagonal of A_ff
  // assembly plan: for every entry of A_ff then A_fc, the face_m positions
  // of its terms in summation order (bit 31: the term starts a new partial
  // sum), and the CSR position of each free row's diagonal (-1: none)
  cn::DevBuf<int32_t> asm_ptr, diag_pos;
  cn::DevBuf<uint32_t> asm_terms;
  // A_fc transposed (rows = constrained columns), built on first use: the
  // free row of each entry and its index in fc_col / fc_val, rows sorted by
  // free row, so the values are read from fc_val and never go stale
  cn::DevBuf<int32_t> cf_ptr, cf_row, cf_map;
  bool cf_ready = false;
  // harmonic right-hand side table, built on first use: for every free row
  // the groups (incoming interior halfedges, ascending) that touch a
  // constrained vertex, each with its (face_m index, constrained column)
  // entries in face order -- the spec's summation order without the terms
  // that are structurally zero
  cn::DevBuf<int32_t> rhs_gptr, rhs_eptr, rhs_cc, rhs_midx;
  bool rhs_ready = false;
  // a batch's right-hand side table is its members' tables end to end (rows,
  // constrained columns and face blocks offset per member, face blocks
  // copied into face_m): the constrained columns it takes, stacked
  int32_t rhs_ncon = 0;
  // A_ff as a fixed-width row table for the solver's multi-column SpMM:
  // the first kEll entries of every row (short rows padded with their last
  // column and weight 0), the rest in an overflow CSR.  The pattern is built
  // on first use (ov_nnz = -1 until then), the values at the first solve
  // after they change (vals_gen counts the value assemblies; ell_gen and
  // win_gen are the ones the tables hold).
  cn::DevBuf<int32_t> ell_col, ov_ptr, ov_col;
  cn::DevBuf<float> ell_val32, ov_val32;
  cn::DevBuf<double> ell_val64, ov_val64;
  int64_t ov_nnz = -1;
  // A_ff as the 8-slot window row table of the LDS-staged SpMM, for ranges
  // of win_rpr rows: a slot holds the neighbour's row, or -(f + 1) for far
  // row f of its range (far rows: win_far_rows[win_far_ptr[x] ..]), and the
  // CSR position its weight is read from (-1: weight 0).  win_ok = 1 when
  // built and usable, -1 when the matrix does not fit the window.
  cn::DevBuf<int32_t> win_col, win_q, win_far_ptr, win_far_rows;
  // the table packed per 32-row chunk for the kernel (columns, then FP32 or
  // FP64 weights), refilled from the values at every solve
  cn::DevBuf<unsigned char> win_pk32, win_pk64;
  int32_t win_rpr = 0, win_ok = 0;
  uint64_t vals_gen = 1, ell_gen = 0, win_gen = 0;
  cn_tune tune{};                   // launch shapes (cn_system_set_tune); 0 = built-in
  // A batch of independent systems (cn_system_batch) is one block-diagonal
  // A_ff: segment s owns rows [seg_ptr[s], seg_ptr[s+1]).  The solver keeps
  // its CG scalars per (segment, column), and takes its dots over blocks of
  // rows that never cross a segment: blocks [blk_ptr[s], blk_ptr[s+1]) of
  // the dot grid cover segment s.  nseg = 1 for a single system.
  int32_t nseg = 1, seg_blocks = 0;
  std::vector<int32_t> seg_ptr;      // host, nseg + 1
  cn::DevBuf<int32_t> d_seg_ptr, seg_of_row, blk_ptr, blk_seg;
  cn::DevBuf<double> deform_ff;     // per-face F of cn_deform_rhs, grown on demand
  cn::SolverWork *work = nullptr;   // solver workspace, grown on demand
  // right-hand sides and solutions of a solve whose column count is padded
  // to the kernels' vector width (n x k_padded, grown on demand)
  cn::DevBuf<double> pad_b, pad_x;
  ~cn_system_t() { cn::free_solver_work(work); }
};

// Colour classes of A_ff for the SgsColor sweeps (pattern only: the
// values and the diagonal are read from the system at every sweep).
struct cn_sgs_t {
  uint64_t serial = cn::next_serial();
  cn_system_t *sys = nullptr;
  int32_t n = 0, ncolors = 0;
  std::vector<int32_t> color_ptr;   // host: rows of colour c are
                                    // color_rows[color_ptr[c] .. color_ptr[c+1])
  cn::DevBuf<int32_t> color_rows, color;
};

// CSR-by-coarse form of a parent map: the fine vertices of coarse vertex c,
// ascending, are child[cptr[c] .. cptr[c+1]).
struct cn_restrict_t {
  int32_t n_fine = 0, n_coarse = 0;
  cn::DevBuf<int32_t> cptr, child;
};

struct cn_ic0_t {
  uint64_t serial = cn::next_serial();
  int32_t n = 0, ncolors = 0, ordering = 0, block_size = 1, max_block = 1;
  double shift = 0.0;
  // Rows are renumbered into levels (colours), each level into blocks, each
  // block a run of consecutive positions.  Rows of a block depend on earlier
  // rows of the same block and on earlier levels only, so one thread or
  // wavefront walks a block in order while all blocks of a level run in
  // parallel.  Point orderings have one row per block.
  std::vector<int32_t> level_ptr;   // host: positions of level c are
                                    // [level_ptr[c], level_ptr[c+1])
  std::vector<int32_t> level_blk;   // host: blocks of level c are
                                    // [level_blk[c], level_blk[c+1])
  cn::DevBuf<int32_t> blk_ptr;      // positions of block b: [blk_ptr[b], blk_ptr[b+1])
  cn::DevBuf<int32_t> perm, iperm;  // perm[new] = old
  // permuted A (FP64 + FP32) for the solver
  cn::DevBuf<int32_t> a_ptr, a_col;
  cn::DevBuf<double> a_val;
  cn::DevBuf<float> a_val32;
  // L strictly lower (CSR rows) and U = L^T strictly upper (CSR rows)
  cn::DevBuf<int32_t> l_ptr, l_col, u_ptr, u_col, u_map;
  cn::DevBuf<double> l_val, u_val, l_diag;
  cn::DevBuf<float> l_val32, u_val32, l_rdiag32;
  cn::DevBuf<double> l_rdiag64;     // 1 / L_ii by level position
  // solver copies: column indices of L / U in the original vertex numbering,
  // so the solver keeps its vectors in natural (mesh) order and only the
  // triangular passes walk rows through `perm`
  cn::DevBuf<int32_t> l_col_nat, u_col_nat;
  int64_t nnz_l = 0;
  // L (forward) and U (backward) prepared for the solver's block sweeps.
  // Row `ps` of a triangle is split into its entries in other blocks (final
  // when the block's level runs) and its entries inside its own block:
  //  - other blocks: the first kSweepEll in a fixed-width table indexed by
  //    position (natural column; unused slots repeat the row's last entry
  //    with weight 0), the rest in an overflow CSR;
  //  - own block: a dense mb x mb tile per block, [t][u] = entry of block
  //    row t in block column u, with 1 / L_tt on the diagonal.
  // mb = 0 when there are no multi-row blocks or a block exceeds 32 rows.
  static constexpr int kSweepEll = 4;
  struct Sweep {
    cn::DevBuf<int32_t> ell_col, ell_cnt, ov_ptr, ov_col;
    cn::DevBuf<double> ell_val, ov_val, dense;
    cn::DevBuf<float> ell_val32, ov_val32, dense32;
  };
  int32_t sweep_mb = 0;
  Sweep lsw, usw;
};

// Hierarchical Schur preconditioner (see cn_hsc_build.hip).
struct cn_hsc_t {
  uint64_t serial = cn::next_serial();
  int32_t n0 = 0;  // rows of level 0 (= the system's n_free)
  cn_hsc_opts opts{};
  struct Level {
    int32_t n = 0, nF = 0, nC = 0;
    // level-local numbering (the hierarchy as built; cn_hsc_get_level)
    cn::DevBuf<int32_t> F, C;
    cn::DevBuf<double> dinv;
    cn::DevBuf<float> dinv32;
    cn::DevBuf<int32_t> r_ptr, r_col, p_ptr, p_col;
    cn::DevBuf<double> r_val, p_val;
    cn::DevBuf<float> r_val32, p_val32;
    // the same rows and columns as level-0 rows: the apply works in place
    // in one n0 x k buffer
    cn::DevBuf<int32_t> Fg, Cg, r_colg, p_colg;
  };
  std::vector<Level> lv;
  int32_t nc = 0;                   // coarse rows
  cn::DevBuf<int32_t> coarse_rows;  // their level-0 rows
  cn::DevBuf<double> coarse_L, coarse_inv;
  cn::DevBuf<float> coarse_inv32;   // coarse_inv rounded to FP32
  // coarse_inv as hi + lo bfloat16 matrices (rows of coarse_ld entries),
  // for hierarchies with at least 3,072 coarse rows
  cn::DevBuf<uint16_t> coarse_bh, coarse_bl;
  int32_t coarse_ld = 0;
  int64_t nnz = 0;
  // apply workspace: the gathered coarse rows (nc x k), grown on demand
  cn::DevBuf<unsigned char> work;
  size_t work_bytes = 0;
  // level descriptors of the fused tail (the small levels and the coarse
  // solve in one launch), FP32 and FP64 forms, built on first use
  cn::DevBuf<unsigned char> tail_desc[2];
  // restriction chains of the down sweep: every row eliminated at level m
  // (or kept to the coarse level, m = L) holds its R entries of levels
  // 0 .. m-1 end to end, by level -- columns as level-0 rows, FP64 and FP32
  // weights.  ch_seg[ch_seg_ptr[v] + l] is the first entry of row v's
  // level-l segment (l = 0 .. m; l = m: the end).  Rows are stored by level
  // of elimination, then in that level's F order.
  cn::DevBuf<int32_t> ch_seg_ptr, ch_seg, ch_col;
  cn::DevBuf<double> ch_val;
  cn::DevBuf<float> ch_val32;
  bool ch_ready = false;
  // the down sweep's launches for one (group length, end level): per row of
  // every launch, in launch order, {level-0 row, first entry of its level-a
  // segment, end of that segment, end of its chain at the launch's level},
  // so a row's chain costs one 16-byte load before its gathers; built on
  // first use of that sweep shape
  struct ChainRecs {
    int32_t G = 0, t = 0;
    cn::DevBuf<int4> rec;
    std::vector<int64_t> off;  // launch i's records: [off[i], off[i + 1])
  };
  std::vector<ChainRecs> ch_recs;
};

namespace cn {
// the tail descriptors and this device's scratch address of a hierarchy,
// built at set-up (cn_hsc_build, cn_hsc_import)
void hsc_prepare(cn_hsc_t *H, hipStream_t s);
}  // namespace cn
