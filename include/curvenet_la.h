/*
 * curvenet_la.h -- C ABI of the MI355X (gfx950) profile-curves linear-algebra
 * core: cut-mesh build, cotangent-Laplacian assembly, SpMV/SpMM, IC(0) and
 * preconditioned conjugate gradients.
 *
 * The reference (V-Sekai-fire/TOOL_godot_curvenet) specifies this path in Lean
 * (the lean/Curvenet/ spec files) and states that its C++ runtime `src/curvenet/*.h`
 * "mirrors this 1:1" (SparseLinAlg.lean:24, HalfedgeBuilder.lean:4).  Each
 * entry point below names the spec function it replaces.
 *
 * Conventions
 *   - plain C types only; indices are int32, "none" is -1
 *     (SlangCodegen/HalfedgeBuilder.lean:22);
 *   - pointers named d_* are device pointers (hipMalloc'd or torch CUDA
 *     memory); h_* are host pointers;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *   - calls are synchronous with respect to the host when they return sizes
 *     or status computed on the device; otherwise they are stream-ordered;
 *   - multi-column blocks are n x k row-major ("interleaved", x[i*k + c]),
 *     the layout of SlangCodegen/SpmvMulti.lean:113;
 *   - CSR is row_ptr[rows+1], col[nnz] ascending per row, val[nnz]
 *     (SparseLinAlg.lean:35-41);
 *   - handles are not thread-safe; use one stream per handle.
 */
#ifndef CURVENET_LA_H
#define CURVENET_LA_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum cn_status {
  CN_OK = 0,
  CN_MAXIT = 1,        /* iteration cap reached before the tolerance        */
  CN_BREAKDOWN = 2,    /* p^T A p == 0 (SparseLinAlg.lean:118)               */
  CN_NOT_SPD = 3,      /* non-positive IC(0) pivot (IncompleteCholesky.lean:53-57) */
  CN_BAD_ARG = -1,
  CN_HIP_ERROR = -2,
  CN_NO_DEVICE = -3,
  CN_OOM = -4,
  CN_NOT_MANIFOLD = -5
} cn_status;

/* Human-readable text of the last error on the calling thread. */
const char *cn_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int32_t cn_version(void);
/* CN_OK when a HIP device is visible, CN_NO_DEVICE otherwise. */
cn_status cn_device_check(void);

/* ------------------------------------------------------------------------
 * Topology (L1)
 * ---------------------------------------------------------------------- */

/* HalfedgeBuilder.fromTriangles (HalfedgeBuilder.lean:34-90): interior
 * halfedges 3f+{0,1,2} = t0->t1, t1->t2, t2->t0; twin = first reverse match;
 * one boundary halfedge per orphan in orphan order; boundary `next` = first
 * boundary halfedge leaving the target, self-loop otherwise.
 * d_target/d_twin/d_next/d_face need capacity 6*nf.  *h_nh receives nh. */
cn_status cn_halfedge_from_triangles(int32_t nv, int32_t nf,
                                     const int32_t *d_tris, int32_t *d_target,
                                     int32_t *d_twin, int32_t *d_next,
                                     int32_t *d_face, int32_t *h_nh,
                                     void *stream);

/* CutAlgorithm.subdivideEdge / splitFace / insertCrack
 * (CutAlgorithm.lean:49-94, 115-163, 212-243).  Host arrays with room for
 * two more halfedges; returns the new halfedge count in *h_nh and updates
 * *h_nv / *h_nf.  O(1) pointer surgery (splitFace walks one loop). */
cn_status cn_cut_subdivide_edge(int32_t *h_nv, int32_t *h_nf, int32_t *h_nh,
                                int32_t h, int32_t *h_target, int32_t *h_twin,
                                int32_t *h_next, int32_t *h_face);
cn_status cn_cut_split_face(int32_t *h_nv, int32_t *h_nf, int32_t *h_nh,
                            int32_t a, int32_t b, int32_t *h_target,
                            int32_t *h_twin, int32_t *h_next, int32_t *h_face);
cn_status cn_cut_insert_crack(int32_t *h_nv, int32_t *h_nf, int32_t *h_nh,
                              int32_t h_anchor, int32_t *h_target,
                              int32_t *h_twin, int32_t *h_next,
                              int32_t *h_face);

/* Seam duplication.  The reference records the curve->seam tracer as an
 * open axiom (Features/KnownUnknowns/CurveSegmentTracing.lean:3-7) and its
 * C++ `cut_algorithm::insert_crack_at_endpoint` is absent, so the cut rule is
 * this library's: the mesh is cut open along the given undirected edges.
 * Around every vertex touching a cut edge the face corners split into
 * wedges (corners joined across uncut interior edges).  The wedge holding
 * the smallest corner slot 3f+l keeps the vertex index; the others get new
 * vertices nv, nv+1, ... in (vertex, wedge-min-slot) order, appended like
 * CutAlgorithm's new vertices (CutAlgorithm.lean:40-47).  Every copy of a
 * vertex on a cut edge becomes a constrained (sample) vertex
 * (CutVertexKind.sample, CutMesh.lean:39-43).
 *   d_new_tris[3 nf]; d_origin / d_constrained capacity nv + 2*ncut.
 *   *h_nv_out = vertex count after the cut. */
cn_status cn_cut_seams(int32_t nv, int32_t nf, const int32_t *d_tris,
                       int32_t ncut, const int32_t *d_cut_edges,
                       int32_t *d_new_tris, int32_t *d_origin,
                       uint8_t *d_constrained, int32_t *h_nv_out, void *stream);

/* SurfaceProjection.projectToVertices (SurfaceProjection.lean:46-64):
 * nearest vertex by squared distance, strict <, lowest index wins ties. */
cn_status cn_project_to_vertices(int32_t nv, const double *d_pos, int32_t nq,
                                 const double *d_q, int32_t *d_index,
                                 void *stream);

/* SurfaceProjection.promoteVertexSamples (SurfaceProjection.lean:64-80):
 * the column-assignment pass over projected knots.  Walking the knots in
 * order, the first knot that lands on a vertex gives that vertex the next
 * sample column; later knots on the same vertex share it; a knot with
 * index < 0 (no vertex) gets -1.
 *   d_index[nq]: vertex per knot (cn_project_to_vertices' output);
 *   d_knot_to_col[nq]: column per knot (out);
 *   d_vertex_to_col[nv]: column per vertex, -1 where no knot landed (out);
 *   *h_ncols: number of columns.
 * An index >= nv is CN_BAD_ARG (the spec's `vertexToCol[v]!` would panic).
 * Synchronous. */
cn_status cn_promote_vertex_samples(int32_t nv, int32_t nq, const int32_t *d_index,
                                    int32_t *d_knot_to_col, int32_t *d_vertex_to_col,
                                    int32_t *h_ncols, void *stream);

/* ------------------------------------------------------------------------
 * Cut Laplacian system (L2): VtLhV of CutMeshLaplacian.assembleVtLhV
 * (CutMeshLaplacian.lean:56-93), stored sparse and split into the free
 * block A_ff (rows/cols = mesh vertices) and A_fc (cols = sample vertices;
 * the spec's sample rows and columns of VtLhV are zero, HarmonicSolve.lean:14-16).
 * Per entry the sum runs over incoming halfedges p in ascending order, then
 * over q in p's face in ascending order -- the order of the spec's two dense
 * matMul loops -- so FP64 values match the spec bit for bit.
 * ---------------------------------------------------------------------- */
typedef struct cn_system_t *cn_system;

typedef struct cn_system_info {
  int32_t nv, nf, nh, n_free, n_con;
  int64_t nnz_ff, nnz_fc;
  const int32_t *d_ff_ptr, *d_ff_col;   /* A_ff, CSR */
  const double *d_ff_val;
  const float *d_ff_val32;              /* FP32 mirror */
  const int32_t *d_fc_ptr, *d_fc_col;   /* A_fc, CSR (cols = con index) */
  const double *d_fc_val;
  const int32_t *d_free_of, *d_con_of;  /* per vertex: row / col, or -1 */
  const int32_t *d_free_vertex, *d_con_vertex; /* inverse maps */
  const int32_t *d_he_target, *d_he_twin, *d_he_next, *d_he_face;
} cn_system_info;

/* Build the system of a (cut) triangle mesh.  d_tris[3 nf], d_constrained[nv]
 * (1 = sample vertex), d_pos[3 nv] FP64.  Copies its inputs. */
cn_status cn_system_create(int32_t nv, int32_t nf, const int32_t *d_tris,
                           const uint8_t *d_constrained, const double *d_pos,
                           cn_system *out, void *stream);

/* Triangle Laplacian of the assembly.  CN_LAP_COT is the embedded cot of
 * PolygonLaplacian.triangleCotLaplacian (PolygonLaplacian.lean:56-71), what
 * CutMeshLaplacian assembles (CutMeshLaplacian.lean:64); CN_LAP_ROBUST is
 * RobustLaplacian.robustTriangleCotLaplacian (RobustLaplacian.lean:136-140):
 * edge lengths mollified per triangle by the smallest eps >= 0 giving every
 * corner a + b > c + delta, cot by the law of cosines over Heron's area --
 * finite on near-degenerate triangles. */
enum { CN_LAP_COT = 0, CN_LAP_ROBUST = 1 };
typedef struct cn_system_opts {
  int32_t laplacian;    /* CN_LAP_*                        */
  double delta;         /* CN_LAP_ROBUST margin (>= 0)     */
} cn_system_opts;
void cn_system_default_opts(cn_system_opts *opts);
/* cn_system_create with the triangle Laplacian chosen by opts; the choice
 * holds for cn_system_set_positions too. */
cn_status cn_system_create_ex(int32_t nv, int32_t nf, const int32_t *d_tris,
                              const uint8_t *d_constrained, const double *d_pos,
                              const cn_system_opts *opts, cn_system *out, void *stream);
/* A system holding a caller-assembled SPD matrix as its A_ff (no mesh, no
 * A_fc): the `SparseMatrixCSR` argument of SparseLinAlg.cg
 * (SparseLinAlg.lean:35-41, 104).  Copies its inputs. */
cn_status cn_system_from_csr(int32_t n, const int32_t *d_ptr,
                             const int32_t *d_col, const double *d_val,
                             cn_system *out, void *stream);
cn_status cn_system_get_info(cn_system sys, cn_system_info *info);
/* Re-assemble values for new vertex positions (same topology). */
cn_status cn_system_set_positions(cn_system sys, const double *d_pos,
                                  void *stream);
cn_status cn_system_destroy(cn_system sys);

/* Many independent systems as one batch: a block-diagonal copy of their
 * A_ff (segment s = systems[s], rows in order).  No counterpart in the
 * reference, which solves one mesh per call (HarmonicSolve.solveMulti,
 * HarmonicSolve.lean:51-66); it is the many-uneven-meshes launch of the
 * deform path.  Every solver entry point accepts a batch (IC(0) and
 * CN_PREC_FP64_PLAIN excepted): right-hand sides and solutions are
 * (sum of n_free) x k, segment s in rows [seg_ptr[s], seg_ptr[s+1]), and
 * every (segment, column) pair is its own CG column -- its own alpha, beta,
 * stop test and residual replacement -- so a segment's solution is the
 * solution of its system.  The batch copies the values: after
 * cn_system_set_positions on a member, rebuild it.  The right-hand side of
 * a segment comes from cn_harmonic_rhs on its member system, or every
 * segment's at once from cn_harmonic_rhs on the batch, which takes the
 * members' constrained values stacked (sum of n_con) x k (one launch;
 * the same bits as the per-member calls). */
cn_status cn_system_batch(int32_t nsys, const cn_system *systems, cn_system *out,
                          void *stream);
/* Launch shapes of a system's solver kernels.  No counterpart in the
 * reference: the autotuner (tool_godot_curvenet_amd/autotune.py) times
 * candidates per (system size, k, precision, preconditioner) and keeps the
 * fastest in a JSON cache.  0 in a field keeps the built-in choice; the
 * results are the same to rounding whatever the shapes (the SpMM forms and
 * the tail compute the same products; only partial sums are grouped
 * differently).
 *   spmm_form  1: row-table SpMM; 2 (or 0): LDS-staged SpMM where it
 *              applies (wide k, 16-byte rows, a banded matrix);
 *   spmm_gens  row-table SpMM: row stripes per XCD share, 1, 2 or 4 (0: 4);
 *   level_cap  HSC level passes: grid cap over all column tiles, 1024 to
 *              65536 blocks (0: 16384);
 *   tail_rows  HSC: levels of at most this many rows (and the coarse solve)
 *              run in one fused launch when the apply has enough columns,
 *              up to 16384 (0: 1536; -1: no fused tail);
 *   chain_levels  HSC down sweep (k > 4): levels one group of launches
 *              chains through, each row written once a group, 1 to 24
 *              (0: 3; 1: one in-place pass per level);
 *   down_cap   HSC down sweep: its launches' grid cap, as level_cap (0: the
 *              level_cap); the up passes keep level_cap;
 *   coarse_prec  HSC coarse product of the mixed-precision (FP32) apply:
 *              1 = the FP64 coarse inverse (on FP64 MFMA for k > 4), 2 = its
 *              FP32 rounding (on FP32 MFMA for k > 4, twice the rate; for
 *              k <= 4 half the bytes read; a preconditioner perturbed at
 *              FP32 rounding), 3 = the inverse split into two bfloat16
 *              matrices, each product a b as a_h b_h + a_h b_l + a_l b_h on
 *              bfloat16 MFMA (2^-16 of a b dropped; hierarchies of at least
 *              3,072 coarse rows and k > 4, else as 2) (0: 1).  FP64
 *              applies and the fused tail always take the FP64 inverse. */
typedef struct cn_tune {
  int32_t spmm_form;
  int32_t spmm_gens;
  int32_t level_cap;
  int32_t tail_rows;
  int32_t chain_levels;
  int32_t coarse_prec;
  int32_t down_cap;
} cn_tune;
cn_status cn_system_set_tune(cn_system sys, const cn_tune *tune);
cn_status cn_system_get_tune(cn_system sys, cn_tune *tune);

/* Segments of a system: *nseg = 1 and h_seg_ptr = {0, n_free} for a
 * single system.  h_seg_ptr (host, nseg + 1) lives as long as the system. */
cn_status cn_system_segments(cn_system sys, int32_t *nseg, const int32_t **h_seg_ptr);

/* HarmonicSolve.solveMulti right-hand side (HarmonicSolve.lean:35-45,57-64)
 * restricted to free rows: d_rhs[n_free*k] = -(Vt Lh C Fc), with
 * d_Xc[n_con*k] the sample values per constrained column, summed in the
 * spec's order (FP64, bit-exact). */
cn_status cn_harmonic_rhs(cn_system sys, int32_t k, const double *d_Xc,
                          double *d_rhs, void *stream);

/* ------------------------------------------------------------------------
 * SpMV / SpMM (L3): SparseLinAlg.spmv (SparseLinAlg.lean:44-53) and the
 * SpmvMulti kernel (SlangCodegen/SpmvMulti.lean:78-119) without its k<=16
 * cap.  The f64 forms sum each row left to right with separate multiply and
 * add, like the spec, so they are bit-exact against it.
 * ---------------------------------------------------------------------- */
cn_status cn_spmv_f64(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const double *d_val, const double *d_x, double *d_y,
                      void *stream);
cn_status cn_spmv_f32(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const float *d_val, const float *d_x, float *d_y,
                      void *stream);
cn_status cn_spmm_f64(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const double *d_val, int32_t k, const double *d_X,
                      double *d_Y, void *stream);
cn_status cn_spmm_f32(int32_t rows, const int32_t *d_ptr, const int32_t *d_col,
                      const float *d_val, int32_t k, const float *d_X,
                      float *d_Y, void *stream);

/* ------------------------------------------------------------------------
 * IC(0) (IncompleteCholesky.lean:49-101; CSR no-fill form :26-33)
 * ---------------------------------------------------------------------- */
typedef struct cn_ic0_t *cn_ic0;

enum {
  CN_ORDER_COLOR = 0,   /* GraphColoring.greedyColor order (GraphColoring.lean:41-56):
                           rows grouped by colour, the colour-pass schedule of
                           SlangCodegen/IncompleteCholesky.lean:10-12 */
  CN_ORDER_NATURAL = 1, /* natural order, rows grouped by dependency level */
  CN_ORDER_BLOCK = 2    /* block multicolour: compact BFS patches of
                           `block_size` rows (natural order inside a patch),
                           patches coloured by greedyColor on the patch
                           graph; a patch is walked sequentially, patches of
                           a colour in parallel */
};

typedef struct cn_ic0_opts {
  int32_t ordering;     /* CN_ORDER_*                                  */
  int32_t block_size;   /* CN_ORDER_BLOCK patch size (default 16)      */
  int32_t allow_shift;  /* see cn_ic0_factor                           */
} cn_ic0_opts;

typedef struct cn_ic0_info {
  int32_t n, nlevels, ordering, nblocks, max_block;
  int64_t nnz_l;        /* incl. diagonal */
  double shift;         /* diagonal shift used (0 unless a pivot failed) */
  const int32_t *d_perm;            /* perm[new] = old */
  const int32_t *d_a_ptr, *d_a_col; /* permuted A */
  const double *d_a_val;
  const int32_t *d_l_ptr, *d_l_col; /* strictly lower L, CSR */
  const double *d_l_val, *d_l_diag;
} cn_ic0_info;

/* Factor A_ff (permuted by `ordering`).  Pivots are FP64 and follow the
 * spec's recurrences in its summation order.  A non-positive pivot returns
 * CN_NOT_SPD unless `allow_shift` is set, in which case the factor of
 * A + s*diag(A) is taken for the smallest s in {1e-4,1e-3,...,1} that works
 * (reported in info.shift). */
cn_status cn_ic0_factor(cn_system sys, int32_t ordering, int32_t allow_shift,
                        cn_ic0 *out, void *stream);
/* cn_ic0_factor with an options struct (block ordering and its size). */
cn_status cn_ic0_factor_ex(cn_system sys, const cn_ic0_opts *opts, cn_ic0 *out,
                           void *stream);
cn_status cn_ic0_get_info(cn_ic0 f, cn_ic0_info *info);
/* z = (L L^T)^{-1} r in the system's vertex order (FP64), d_r/d_z n x k. */
cn_status cn_ic0_apply(cn_ic0 f, int32_t k, const double *d_r, double *d_z,
                       void *stream);
cn_status cn_ic0_destroy(cn_ic0 f);

/* ------------------------------------------------------------------------
 * Dense FP64 Cholesky on the matrix cores (DenseLinAlg.choleskyFactor /
 * forwardSolve / backwardSolve / solveWithCholesky, DenseLinAlg.lean:
 * 255-286).  Row-major, device memory.
 * ---------------------------------------------------------------------- */

/* In place: the lower triangle of A (n x n, leading dimension lda) becomes L
 * with A = L L^T, the strict upper triangle is zeroed (the spec returns L
 * with a zero upper triangle).  A non-positive pivot gives CN_NOT_SPD (the
 * spec produces NaN there). */
cn_status cn_dense_cholesky_f64(int32_t n, double *d_A, int32_t lda, void *stream);

/* B (n x k, leading dimension ldb) <- (L L^T)^-1 B for the factor of
 * cn_dense_cholesky_f64 (solveWithCholesky, every column at once). */
cn_status cn_dense_cholesky_solve_f64(int32_t n, int32_t k, const double *d_L,
                                      int32_t lda, double *d_B, int32_t ldb,
                                      void *stream);

/* ------------------------------------------------------------------------
 * Hierarchical Schur preconditioner (HSC, after Krishnan-Fattal-Szeliski
 * 2013; the per-vertex elimination rule is HSC.eliminateVertex,
 * HierarchicalSparsifyCompensate.lean:96-125, the coarse solve the dense
 * Cholesky above).  Level l eliminates a greedy independent set F (vertices
 * by ascending off-diagonal count, then index): the remaining C vertices get
 * the exact Schur complement S = A_CC - A_CF D_F^-1 A_FC; fill entries of S
 * with |s_ij| < theta sqrt(s_ii s_jj) are dropped and compensated on the
 * diagonals.  Levels stop at coarse_max rows; the coarse matrix is factored
 * and inverted on the matrix cores.  Apply: r_C += R r_F level by level
 * down, z_c = A_c^-1 r_c, z_F = D_F^-1 r_F + P z_C back up.
 * ---------------------------------------------------------------------- */
typedef struct cn_hsc_t *cn_hsc;

enum {
  CN_HSC_LUMP = 0,  /* a dropped s_ij is added to s_ii and s_jj: row sums
                       (the action on constants) are kept */
  CN_HSC_DIAG = 1   /* |s_ij| is added: S only grows, always definite */
};

typedef struct cn_hsc_opts {
  double theta;         /* drop threshold (default 0.015)                    */
  int32_t coarse_max;   /* stop at this many rows (default 512)              */
  int32_t max_levels;   /* default 24                                         */
  double min_frac;      /* stop when |F| < min_frac n (default 0.05)          */
  int32_t compensate;   /* CN_HSC_LUMP (default) or CN_HSC_DIAG              */
} cn_hsc_opts;

typedef struct cn_hsc_info {
  int32_t nlevels;      /* elimination levels                                */
  int32_t n_coarse;     /* rows of the coarse matrix                         */
  int64_t nnz;          /* stored R + P entries over all levels              */
  const double *d_coarse_L;    /* coarse Cholesky factor (n_coarse^2, lower) */
  const double *d_coarse_inv;  /* its inverse (n_coarse^2, symmetric)        */
} cn_hsc_info;

/* Level l of the hierarchy (device pointers, level-local row numbering):
 * F (nF rows eliminated, ascending), C (nC kept, ascending; C[c] is row c of
 * level l + 1), dinv = 1 / a_ff, R (nC rows; columns are level-l rows of F
 * vertices, values -a_cf / a_ff) and P (nF rows; columns are level-(l + 1)
 * rows, values -a_fc / a_ff), all CSR with ascending columns. */
typedef struct cn_hsc_level {
  int32_t n, nF, nC;
  const int32_t *d_F, *d_C;
  const double *d_dinv;
  const int32_t *d_r_ptr, *d_r_col;
  const double *d_r_val;
  const int32_t *d_p_ptr, *d_p_col;
  const double *d_p_val;
} cn_hsc_level;

/* One level of a persisted hierarchy (host pointers, the layout of
 * cn_hsc_level; r_col has r_ptr[nC] entries, p_col p_ptr[nF]). */
typedef struct cn_hsc_level_in {
  int32_t n, nF, nC;
  const int32_t *F, *C;
  const double *dinv;
  const int32_t *r_ptr, *r_col;
  const double *r_val;
  const int32_t *p_ptr, *p_col;
  const double *p_val;
} cn_hsc_level_in;

void cn_hsc_default_opts(cn_hsc_opts *opts);
cn_status cn_hsc_build(cn_system sys, const cn_hsc_opts *opts, cn_hsc *out,
                       void *stream);
cn_status cn_hsc_get_info(cn_hsc h, cn_hsc_info *info);
cn_status cn_hsc_get_level(cn_hsc h, int32_t level, cn_hsc_level *out);
/* Levels an apply of k columns runs in the fused tail kernel: the first
 * such level (levels [level, nlevels) and the coarse solve run in one
 * launch, one workgroup per column slice), or -1 when every level runs its
 * own kernels.  fp32 = 1 for the mixed-precision inner loop, 0 for FP64. */
int32_t cn_hsc_tail_level(cn_hsc h, int32_t k, int32_t fp32);
/* z = M^-1 r (FP64), r and z n_free x k, distinct. */
cn_status cn_hsc_apply(cn_hsc h, int32_t k, const double *d_r, double *d_z,
                       void *stream);
/* cn_hsc_apply that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results as cn_hsc_apply, with the apply's workspace (the
 * gathered coarse rows and the coarse product's split-K partial sums) in a
 * fixed device array of CN_HSC_SCRATCH_BYTES per GPU, held until the
 * process ends, instead of the hierarchy's own buffer.  An apply whose
 * workspace exceeds it (about n_coarse x k x 8 bytes, more with split-K
 * partials) is CN_BAD_ARG.  Every call on one GPU shares that array:
 * calls that overlap on a GPU (other streams, other threads) race, so
 * serialise them.  Stream-ordered (no synchronisation). */
#define CN_HSC_SCRATCH_BYTES (16u << 20)
cn_status cn_hsc_apply_noalloc(cn_hsc h, int32_t k, const double *d_r, double *d_z,
                               void *stream);
cn_status cn_hsc_destroy(cn_hsc h);
/* A hierarchy of `sys` from persisted arrays (a sidecar written from
 * cn_hsc_get_level / cn_hsc_get_info): no elimination and no coarse
 * factorization run, so applies are bit-identical to the saved hierarchy's.
 * Checks on the host that level 0 has n_free rows, every level's F and C
 * split its rows in ascending order, R / P are CSR of that level, and
 * n_coarse is what the levels leave; CN_BAD_ARG otherwise.  Whether the
 * arrays belong to this system's values is the caller's to check (the
 * sidecar stores a fingerprint).  This is the persisted-cache form of
 * Features/KnownUnknowns/BindCachePersistence.lean:3-5. */
cn_status cn_hsc_import(cn_system sys, const cn_hsc_opts *opts, int32_t nlevels,
                        const cn_hsc_level_in *levels, int32_t n_coarse,
                        const double *coarse_L, const double *coarse_inv, cn_hsc *out,
                        void *stream);

/* ------------------------------------------------------------------------
 * The deformation solve (DeformSolve.solveDeformation, DeformSolve.lean:
 * 110-135) on this system: stage 1 solves the deformation gradients F
 * (9 columns a pose: cn_harmonic_rhs of F_c, then a solve), stage 2 the
 * positions (3 columns a pose) with the right-hand side below, then a
 * solve.  Poses are column blocks: F is n x 9P (pose p in columns
 * 9p..9p+8, a 3x3 row-major), positions n x 3P.
 * ---------------------------------------------------------------------- */

/* Stage-2 right-hand side -V^T Lh (C X_c - y_h) (DeformSolve.lean:118-133)
 * for the free rows, with the bridge folded in: f_h = F of the halfedge's
 * target (computeFh, :38-52; d_Ffree for free vertices, d_Fc for
 * constrained ones), F_f = the mean over the face loop (averageOverFaces,
 * :56-69), y_h = X_target . F_f^T (computeYh, :74-96).  Summed in the
 * assembly's order, bit-identical to the dense spec form. */
cn_status cn_deform_rhs(cn_system sys, int32_t P, const double *d_Ffree,
                        const double *d_Fc, const double *d_Xc, double *d_rhs,
                        void *stream);

/* Curve inputs of the solve, one thread per evaluation, the spec's
 * operation order (FP64):
 *   cn_bezier3: EndToEndExample.bezier3 (EndToEndExample.lean:52-60),
 *     d_ctrl ncurves x 4 x 3, d_t nt, d_out ncurves x nt x 3;
 *   cn_segment_gradients: ScaledFrames.isolatedSegmentGradient
 *     (ScaledFrames.lean:90-95), F = (l / l_rest) R(t_rest, t), n x 3 points
 *     in, n x 9 out;
 *   cn_deformation_gradients: ScaledFrames.deformationGradient
 *     (ScaledFrames.lean:147-150), F = posed inv(rest), n x 9 each. */
cn_status cn_bezier3(int32_t ncurves, const double *d_ctrl, int32_t nt, const double *d_t,
                     double *d_out, void *stream);
cn_status cn_segment_gradients(int32_t n, const double *d_rest_p, const double *d_rest_q,
                               const double *d_posed_p, const double *d_posed_q, double *d_F,
                               void *stream);
cn_status cn_deformation_gradients(int32_t n, const double *d_rest, const double *d_posed,
                                   double *d_F, void *stream);

/* ------------------------------------------------------------------------
 * Direct Delta Mush.
 *
 * Runtime: DirectDeltaMush.lbsMatvec (DirectDeltaMush.lean:85-98) over every
 * vertex, in the buffers of the SlangCodegen/DirectDeltaMush.lean:33-87
 * shader: d_rest nv x 3, d_T P x nt x 16 (4x4 row-major, rows 0-2 used),
 * influences CSR d_start[nv + 1], d_idx, d_w; d_out P x nv x 3.  A handle
 * index outside [0, nt) reads the zero matrix.  The f64 form keeps the
 * spec's operation order (bit-exact); the f32 form is the shader's float
 * arithmetic with fused multiply-adds.  The caller guarantees a consistent
 * CSR (d_start[0] == 0, non-decreasing, d_start[nv] == entries of d_idx and
 * d_w): the kernels read d_idx/d_w at every position in [start[v],
 * start[v+1]) and do not re-check it (the Python layer does, on the host).
 * ---------------------------------------------------------------------- */
cn_status cn_lbs_matvec_f64(int32_t nv, int32_t nt, int32_t P, const double *d_rest,
                            const double *d_T, const int32_t *d_start, const int32_t *d_idx,
                            const double *d_w, double *d_out, void *stream);
cn_status cn_lbs_matvec_f32(int32_t nv, int32_t nt, int32_t P, const float *d_rest,
                            const float *d_T, const int32_t *d_start, const int32_t *d_idx,
                            const float *d_w, float *d_out, void *stream);

/* DirectDeltaMushBind.smoothWeights (DirectDeltaMushBind.lean:76-97): nu
 * passes over the n x nh row-major d_W, in place, with d_tmp (same size,
 * distinct) as the other buffer.  Adjacency CSR d_adj_ptr[n + 1], d_adj in
 * the caller's neighbour order; a neighbour outside [0, n) adds 0.
 * Bit-exact against the spec. */
cn_status cn_smooth_weights_f64(int32_t n, int32_t nh, const int32_t *d_adj_ptr,
                                const int32_t *d_adj, double *d_W, double *d_tmp, int32_t nu,
                                double omega, void *stream);

/* DirectDeltaMushBind.sparsifyTopK (DirectDeltaMushBind.lean:107-136):
 * take = min(K, nh) (index, weight) pairs per row into d_idx / d_w
 * (n x take), in the spec's selection order, renormalised by their sum when
 * it is nonzero.  take <= 64. */
cn_status cn_sparsify_topk_f64(int32_t n, int32_t nh, const double *d_W, int32_t K,
                               int32_t *d_idx, double *d_w, void *stream);

/* Bind harvest (DirectDeltaMushBind.lean:9-12; the sparse multi-RHS form of
 * Features/KnownUnknowns/SparseBindHarvest.lean:3-5).  Handle h is
 * constrained column h; its weights are the harmonic solution for the unit
 * constrained vector e_h.
 *   cn_bind_rhs: d_rhs n_free x nb, column c = -A_fc e_{h0+c}, equal to
 *     cn_harmonic_rhs of the unit block;
 *   cn_bind_weights: writes columns [h0, h0 + nb) of the nv x ldw vertex
 *     weight matrix d_W: free vertices take their row of the solution d_x
 *     (n_free x nb), sample vertices 1 on their own column, else 0.
 * Single systems only (not batches). */
cn_status cn_bind_rhs(cn_system sys, int32_t h0, int32_t nb, double *d_rhs, void *stream);
cn_status cn_bind_weights(cn_system sys, int32_t h0, int32_t nb, const double *d_x,
                          double *d_W, int64_t ldw, void *stream);

/* ------------------------------------------------------------------------
 * Colour-parallel symmetric Gauss-Seidel (SlangCodegen/SgsColor.lean:
 * 87-130) on A_ff.  Rows are grouped by GraphColoring.greedyColor
 * (GraphColoring.lean:41-56); one sweep runs the SgsColor row update over
 * colours 0..C-1 and then C-1..0 (SgsColor.lean:12-14), one launch per
 * colour.  Row update, column c of an n x k row-major block: d = A_ii (rows
 * with d == 0 are left alone), s = sum of the row's entries times x in
 * storage order from 0, x_i = (b_i - (s - d x_i)) / d, with no fused
 * multiply-add, so the FP64 sweep is bit-identical to the spec's loop and
 * the FP32 one to its float transcription.
 * ---------------------------------------------------------------------- */
typedef struct cn_sgs_t *cn_sgs;

typedef struct cn_sgs_info {
  int32_t n, ncolors;
  const int32_t *h_color_ptr;   /* host, ncolors + 1                         */
  const int32_t *d_color_rows;  /* rows of each colour, ascending            */
  const int32_t *d_color;       /* colour of each row                        */
} cn_sgs_info;

/* Colour the pattern of A_ff.  The handle keeps the system: sweeps read its
 * current values (cn_system_set_positions needs no rebuild). */
cn_status cn_sgs_build(cn_system sys, cn_sgs *out, void *stream);
cn_status cn_sgs_get_info(cn_sgs h, cn_sgs_info *info);
/* `sweeps` symmetric sweeps of A_ff x = b in place on d_x (n_free x k). */
cn_status cn_sgs_sweep_f64(cn_sgs h, int32_t k, const double *d_b, double *d_x,
                           int32_t sweeps, void *stream);
/* The same in FP32 (values and diagonal rounded from the FP64 matrix). */
cn_status cn_sgs_sweep_f32(cn_sgs h, int32_t k, const float *d_b, float *d_x,
                           int32_t sweeps, void *stream);
/* The SGS preconditioner z = M^-1 r, M = (D+L) D^-1 (D+U) in colour order:
 * one symmetric sweep from z = 0 (FP64). */
cn_status cn_sgs_apply(cn_sgs h, int32_t k, const double *d_r, double *d_z, void *stream);
cn_status cn_sgs_destroy(cn_sgs h);

/* ------------------------------------------------------------------------
 * Coarse-vertex restriction of a parent map (SlangCodegen/
 * HierarchicalSparsify.lean:64-93): coarse[c] = sum of fine[i] over fine
 * vertices i with parent[i] = c, added in ascending i from 0 (bit-identical
 * to the spec's scan).  Built once into the CSR-by-coarse form the spec
 * names as the production form; parents outside [0, n_coarse) match no
 * coarse vertex.  Blocks are row-major (n x k).
 * ---------------------------------------------------------------------- */
typedef struct cn_restrict_t *cn_restrict;

cn_status cn_restrict_build(int32_t n_fine, const int32_t *d_parent, int32_t n_coarse,
                            cn_restrict *out, void *stream);
cn_status cn_restrict_sum_f64(cn_restrict h, int32_t k, const double *d_fine,
                              double *d_coarse, void *stream);
cn_status cn_restrict_sum_f32(cn_restrict h, int32_t k, const float *d_fine,
                              float *d_coarse, void *stream);
cn_status cn_restrict_destroy(cn_restrict h);

/* ------------------------------------------------------------------------
 * Preconditioned conjugate gradients (SparseLinAlg.cg / cgWithGuess,
 * SparseLinAlg.lean:104-163) on A_ff, k right-hand sides in lockstep.
 * ---------------------------------------------------------------------- */
enum { CN_PRECOND_JACOBI = 0, CN_PRECOND_IC0 = 1, CN_PRECOND_HSC = 2, CN_PRECOND_SGS = 3 };
enum {
  CN_PREC_FP64_PLAIN = 0, /* FP64; one kernel per BLAS-1/SpMV step, alpha/beta
                             and the stop test on the host (3 host round
                             trips per iteration, the dispatch chain of the
                             reference's GPU CG) */
  CN_PREC_FP64 = 1,       /* FP64; scalars on the device, fused kernels */
  CN_PREC_MIXED = 2       /* FP32 SpMM / IC(0) / search directions inside an
                             FP64 iterative-refinement loop: FP64 residual
                             b - A x, FP64 dots and alpha/beta, FP64 update */
};

typedef struct cn_pcg_opts {
  int32_t precond;        /* CN_PRECOND_* (IC0 needs a factor)               */
  int32_t precision;      /* CN_PREC_*                                        */
  double tol_rel;         /* stop when ||b - A x||_2 <= tol_rel ||b||_2 per
                             column, recomputed in FP64                      */
  int32_t max_iter;       /* total inner iterations cap                       */
  double inner_tol;       /* MIXED: a residual replacement (FP64 x update and
                             recomputed b - A x) is due once a column's
                             iterated ||r|| has fallen by this factor since
                             the last one; in (0, 1), default 0.0316       */
  int32_t max_refine;     /* MIXED: refinement steps cap                      */
  int32_t check_every;    /* device-side stop flags are read every n iters    */
  int32_t warm_start;     /* 1: d_x holds x0 (cgWithGuess), 0: x0 = 0 (cg)    */
  int32_t graph;          /* 1: the fused solvers run the iterations between
                             two reads of the stop flags as one captured
                             hipGraph, replayed (captured once per system,
                             preconditioner and k); 0: launch each kernel.
                             The results are the same bits either way.
                             Not used while cn_profile_enable is on.        */
} cn_pcg_opts;

typedef struct cn_solve_info {
  int32_t status;         /* cn_status                                        */
  int32_t iters;          /* inner iterations performed                       */
  int32_t refinements;    /* FP64 residual recomputations (MIXED)             */
  double relres_max;      /* max over columns of ||b-Ax||/||b|| (FP64)        */
} cn_solve_info;

/* Fill opts with defaults (IC0, MIXED, tol 1e-9, 20000 iters, inner 1e-5,
 * graph replay on). */
void cn_pcg_default_opts(cn_pcg_opts *opts);

/* Solve A_ff X = B for k columns.  d_b / d_x are n_free x k FP64; d_x
 * holds the iterate during the solve and may be d_b itself (in place).
 * `ic0` may be NULL for Jacobi.  Promoted (sample) rows are not part of
 * A_ff; the caller overlays their values (DeformSolve.lean:108-109). */
cn_status cn_pcg_solve(cn_system sys, cn_ic0 ic0, const cn_pcg_opts *opts,
                       int32_t k, const double *d_b, double *d_x,
                       cn_solve_info *info, void *stream);

/* The same solve preconditioned by an HSC hierarchy of this system
 * (opts->precond is ignored). */
cn_status cn_pcg_solve_hsc(cn_system sys, cn_hsc hsc, const cn_pcg_opts *opts,
                           int32_t k, const double *d_b, double *d_x,
                           cn_solve_info *info, void *stream);

/* The same solve preconditioned by one symmetric SgsColor sweep from zero
 * per application (cn_sgs_apply; opts->precond is ignored). */
cn_status cn_pcg_solve_sgs(cn_system sys, cn_sgs sgs, const cn_pcg_opts *opts,
                           int32_t k, const double *d_b, double *d_x,
                           cn_solve_info *info, void *stream);

/* Adjoint of the harmonic solve A_ff x_f = -A_fc x_c.  No counterpart in
 * the reference (the deform path HarmonicSolve.solveMulti,
 * HarmonicSolve.lean:51-66, is only ever run forward); it is the backward
 * pass a gradient-based caller (an autograd function) needs.  For the
 * gradient g = dL/dx_f (n_free x k, FP64, device):
 *   d_lambda = A_ff^-1 g, solved like cn_pcg_solve (same opts and factor,
 *              always from x0 = 0);
 *   d_gc     = dL/dx_c = -A_fc^T lambda (n_con x k), skipped if NULL;
 *   d_gw     = dL/dw for every stored entry of A_ff (nnz_ff values) then of
 *              A_fc (nnz_fc): w is the weight of the edge (i, j) the entry
 *              couples (A_ii += w, A_jj += w, A_ij = A_ji -= w), so
 *              dL/dw = -sum_cols (lambda_i - lambda_j)(x_i - x_j) with
 *              lambda = 0 on constrained vertices; 0 on the diagonal.
 *              Needs d_x (n_free x k) and d_xc (n_con x k); skipped if NULL.
 * Returns the solve's status. */
cn_status cn_adjoint_solve(cn_system sys, cn_ic0 ic0, const cn_pcg_opts *opts,
                           int32_t k, const double *d_g, const double *d_x,
                           const double *d_xc, double *d_lambda, double *d_gc,
                           double *d_gw, cn_solve_info *info, void *stream);

/* cn_adjoint_solve with lambda solved by the HSC-preconditioned PCG
 * (cn_pcg_solve_hsc, same hierarchy as the forward solve). */
cn_status cn_adjoint_solve_hsc(cn_system sys, cn_hsc hsc, const cn_pcg_opts *opts,
                               int32_t k, const double *d_g, const double *d_x,
                               const double *d_xc, double *d_lambda, double *d_gc,
                               double *d_gw, cn_solve_info *info, void *stream);

/* ------------------------------------------------------------------------
 * safetensors container (host code; LeanSafetensors Header.lean,
 * Reader.lean, Writer.lean).  Used for mesh ingest and for the hierarchy
 * sidecar.  dtype codes follow Header.lean:36-41: F32 F64 F16 BF16 U8 U16
 * U32 U64 I8 I16 I32 I64 BOOL = 0..12.
 *   cn_st_parse / cn_st_read_file: SafeTensors.parse + validate
 *     (Reader.lean:33-61).  Tensors and metadata come back in key order (the
 *     reference decodes JSON objects into key-sorted maps; a repeated key
 *     keeps its last value); non-string metadata values as their JSON text.
 *   cn_st_find: SafeTensors.get (:64-69), index or -1;
 *   cn_st_metadata_get: metadataAt (:72-74), NULL when absent;
 *   cn_st_encode / cn_st_write_file: Writer.layout + encode
 *     (Writer.lean:20-58), tensors contiguous in the given order,
 *     "__metadata__" first, header space-padded to an 8-byte boundary.
 *     cn_st_encode with buf = NULL returns the size only.
 * A cn_st owns its bytes; pointers it hands out live until cn_st_free.
 * ---------------------------------------------------------------------- */
typedef struct cn_st_t *cn_st;
typedef struct cn_st_tensor {
  const char *name;
  int32_t dtype;
  int32_t ndim;
  const int64_t *shape;
  const void *data;
  int64_t nbytes;
} cn_st_tensor;

cn_status cn_st_parse(const void *bytes, int64_t size, cn_st *out);
cn_status cn_st_read_file(const char *path, cn_st *out);
int32_t cn_st_count(cn_st st);
cn_status cn_st_tensor_at(cn_st st, int32_t i, cn_st_tensor *out);
int32_t cn_st_find(cn_st st, const char *name);
int32_t cn_st_metadata_count(cn_st st);
cn_status cn_st_metadata_at(cn_st st, int32_t i, const char **key, const char **value);
const char *cn_st_metadata_get(cn_st st, const char *key);
void cn_st_free(cn_st st);
cn_status cn_st_encode(int32_t n, const cn_st_tensor *tensors, int32_t nmeta,
                       const char *const *keys, const char *const *values, void *buf,
                       int64_t cap, int64_t *size);
cn_status cn_st_write_file(const char *path, int32_t n, const cn_st_tensor *tensors,
                           int32_t nmeta, const char *const *keys, const char *const *values);

/* Per-kernel timing of the last cn_pcg_solve on this thread, gathered with
 * HIP events around every launch of the named kernel when enabled.
 * name: "spmm" (FP32 SpMM of the inner loop), "fwd" / "bwd" (forward and
 * backward preconditioner sweeps, or r -= alpha q + M^-1 r + r.z and the
 * p / d update for HSC), "resid" (FP64 residual), "hsc_down" / "hsc_up" /
 * "hsc_coarse" (the level passes and coarse product of every HSC apply).
 * Returns the scopes timed; *ms_total gets their summed duration. */
void cn_profile_enable(int32_t on);
int32_t cn_profile_query(const char *name, double *ms_total);
/* Record events only around the scopes of this name (NULL or "": all).
 * Each recorded event pair leaves the GPU idle a few microseconds, so a
 * timed run that needs one kernel's duration records only that kernel. */
void cn_profile_only(const char *name);

#ifdef __cplusplus
}
#endif
#endif
