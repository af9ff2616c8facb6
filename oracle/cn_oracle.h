/*
 * cn_oracle.h -- FP64 CPU restatement of the profile-curves linear-algebra
 * spec (V-Sekai-fire/TOOL_godot_curvenet, lean/Curvenet/ *.lean files).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library never
 * links or calls it.
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math so every + - * / sqrt is a
 * single correctly rounded IEEE binary64 operation, the same arithmetic Lean's
 * Float performs.  Loops keep the spec's iteration and summation order, so a
 * result computed here is what the Lean definition would return bit for bit.
 *
 * Conventions shared by all functions:
 *   - int32 indices, -1 = "none" (Option Nat = none in the spec);
 *   - CSR: row_ptr[rows+1], col[nnz], val[nnz], columns ascending per row;
 *   - multi-column blocks are n x k row-major ("interleaved", x[i*k + c]);
 *   - return value >= 0 is a count, < 0 is an error code.
 *
 * Every function names the spec lines it follows.  Builder-defined parts
 * (no counterpart in the spec) say so.
 */
#ifndef CN_ORACLE_H
#define CN_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- Halfedge topology (Halfedge.lean, HalfedgeBuilder.lean) ---------- */

/* HalfedgeBuilder.fromTriangles, literal O((3F)^2) transcription
 * (HalfedgeBuilder.lean:34-90).  Arrays sized >= 6*nf.  Returns nh. */
int32_t orc_he_from_triangles_spec(int32_t nv, int32_t nf, const int32_t *tris,
                                   int32_t *target, int32_t *twin,
                                   int32_t *next, int32_t *face);

/* Same output, sorted-key implementation for meshes too large for the
 * quadratic scan (index-identical: first match wins, orphan order kept). */
int32_t orc_he_from_triangles(int32_t nv, int32_t nf, const int32_t *tris,
                              int32_t *target, int32_t *twin, int32_t *next,
                              int32_t *face);

/* HalfedgeMesh.manifold? (Halfedge.lean:42-128).  Returns a bitmask:
 * 1 indicesInRange, 2 twinInvolutive, 4 twinsAreOpposite, 8 faceLoopsClose.
 * 15 = manifold. */
int32_t orc_he_manifold(int32_t nv, int32_t nf, int32_t nh,
                        const int32_t *target, const int32_t *twin,
                        const int32_t *next, const int32_t *face);

/* ---- Cut surgery (CutAlgorithm.lean) ----------------------------------
 * In-place on caller arrays with capacity for the 2 new halfedges.
 * Each returns the new halfedge count; *nv / *nf are updated. */
int32_t orc_subdivide_edge(int32_t *nv, int32_t *nf, int32_t nh, int32_t h,
                           int32_t *target, int32_t *twin, int32_t *next,
                           int32_t *face);                     /* :49-94   */
int32_t orc_split_face(int32_t *nv, int32_t *nf, int32_t nh, int32_t a,
                       int32_t b, int32_t *target, int32_t *twin,
                       int32_t *next, int32_t *face);          /* :115-163 */
int32_t orc_insert_crack(int32_t *nv, int32_t *nf, int32_t nh, int32_t h_anchor,
                         int32_t *target, int32_t *twin, int32_t *next,
                         int32_t *face);                       /* :212-243 */

/* ---- Seam cut (builder-defined; the spec's tracer is an axiom) --------
 * Cut a triangle mesh open along a set of undirected edges.  For every
 * vertex v incident to a cut edge, the corners (face, slot) around v are
 * grouped into wedges: two corners are in one wedge when their faces share
 * an uncut interior edge at v.  The wedge holding the smallest corner slot
 * (3*f + l) keeps index v; every other wedge gets a new vertex, numbered
 * nv, nv+1, ... in (v ascending, wedge-min-slot ascending) order.  All
 * copies of a vertex on a cut edge are constrained (sample vertices).
 *   new_tris[3*nf], origin[cap], constrained[cap] (cap >= nv + 2*ncut).
 * Returns nv' (vertex count after the cut) or < 0 on error. */
int32_t orc_cut_edges(int32_t nv, int32_t nf, const int32_t *tris,
                      int32_t ncut, const int32_t *cut_edges,
                      int32_t *new_tris, int32_t *origin,
                      uint8_t *constrained);

/* ---- Cotangent Laplacian (PolygonLaplacian.lean, CutMeshLaplacian.lean) */

/* PolygonLaplacian.cotAtVertex (:32-40). */
double orc_cot_at_vertex(const double *apex, const double *o1,
                         const double *o2);
/* triangleCotLaplacian (:56-71): m[9] row-major. */
void orc_triangle_cot_laplacian(const double *a, const double *b,
                                const double *c, double *m);
double orc_mollify_triangle(double delta, double eA, double eB, double eC);
double orc_cot_from_lengths(double eo, double e1, double e2);
void orc_robust_triangle_cot_laplacian(double delta, const double *a, const double *b,
                                       const double *c, double *m);
void orc_set_laplacian(int32_t kind, double delta);
/* polygonCotLaplacian (:77-93): poly[3*n], m[n*n]. */
void orc_polygon_cot_laplacian(int32_t n, const double *poly, double *m);

/* Dense CutMeshLaplacian.assembleVtLhV (:86-93) on a halfedge mesh with
 * vertex kinds given by `is_sample` (1 = sample / promoted vertex).  Literal
 * dense transcription: Lh (nh x nh), V (nh x nv), Vt, two matMul.  out[nv*nv].
 * For small meshes only (O(nh^2 nv)). */
void orc_assemble_vtlhv_dense(int32_t nv, int32_t nf, int32_t nh,
                              const int32_t *target, const int32_t *next,
                              const int32_t *face, const uint8_t *is_sample,
                              const double *pos, double *out);

/* Dense RHS of HarmonicSolve.solveMulti (HarmonicSolve.lean:51-66):
 * rhs = -Vt Lh C Fc, nv x k.  col_of_vertex[v] = C column of a sample
 * vertex (-1 for mesh vertices).  Small meshes only. */
void orc_harmonic_rhs_dense(int32_t nv, int32_t nf, int32_t nh,
                            const int32_t *target, const int32_t *next,
                            const int32_t *face, const int32_t *col_of_vertex,
                            const double *pos, int32_t k, const double *Fc,
                            double *rhs);

/* Sparse assembly of the same matrix with the same per-entry summation
 * order (ascending halfedge p targeting the row, then ascending q in p's
 * face), split into the free-free block A_ff and the free-constrained block
 * A_fc.  free_of[v] = row index of a free vertex, -1 for constrained;
 * con_of[v] = column of a constrained vertex in A_fc, -1 for free.
 * Capacities: *_col / *_val >= sum over free rows of (1 + 2*deg_face).
 * Output counts in nnz_out[2] = {nnz_ff, nnz_fc}.  Returns n_free. */
int32_t orc_cut_laplacian_csr(int32_t nv, int32_t nf, int32_t nh,
                              const int32_t *target, const int32_t *next,
                              const int32_t *face, const uint8_t *constrained,
                              const double *pos, int32_t *free_of,
                              int32_t *con_of, int32_t *ff_ptr,
                              int32_t *ff_col, double *ff_val,
                              int32_t *fc_ptr, int32_t *fc_col,
                              double *fc_val, int64_t cap, int64_t *nnz_out);

/* Sparse harmonic RHS, spec order (see orc_harmonic_rhs_dense):
 * rhs[n_free*k] for the free rows; Xc[n_con*k] per constrained column. */
void orc_harmonic_rhs_csr(int32_t nv, int32_t nh, const int32_t *target,
                          const int32_t *next, const int32_t *face,
                          const int32_t *free_of, const int32_t *con_of,
                          const double *pos, int32_t k, const double *Xc,
                          double *rhs);
void orc_bezier3(const double *P, double t, double *out);
void orc_isolated_segment_gradient(const double *rp, const double *rq, const double *pp,
                                   const double *pq, double *F);
double orc_mat3_det(const double *m);
void orc_mat3_inv(const double *m, double *o);
void orc_deformation_gradient(const double *rest, const double *posed, double *F);
void orc_deform_rhs_dense(int32_t nv, int32_t nf, int32_t nh, const int32_t *target,
                          const int32_t *next, const int32_t *face,
                          const int32_t *col_of_vertex, const double *pos, int32_t P,
                          const double *F, const double *Xc, double *rhs);
void orc_deform_rhs_csr(int32_t nv, int32_t nh, const int32_t *target, const int32_t *next,
                        const int32_t *face, const int32_t *free_of, const int32_t *con_of,
                        const double *pos, int32_t P, const double *Ffree, const double *Fc,
                        const double *Xc, double *rhs);


/* ---- Sparse linear algebra (SparseLinAlg.lean) ------------------------ */

void orc_spmv(int32_t rows, const int32_t *ptr, const int32_t *col,
              const double *val, const double *x, double *y);      /* :44-53 */
void orc_spmm(int32_t rows, const int32_t *ptr, const int32_t *col,
              const double *val, int32_t k, const double *X, double *Y);
double orc_dot(int32_t n, const double *a, const double *b);        /* :56-60 */
void orc_diagonal(int32_t rows, const int32_t *ptr, const int32_t *col,
                  const double *val, double *d);                   /* :78-87 */
void orc_apply_jacobi(int32_t n, const double *d, const double *b,
                      double *y);                                  /* :92-97 */

/* SparseLinAlg.cg / cgWithGuess (:104-163) literal: absolute tol on ||r||,
 * x0 == NULL means cold start.  Returns the number of completed loop
 * iterations (the spec returns x only; the count is extra). */
int32_t orc_cg(int32_t n, const int32_t *ptr, const int32_t *col,
               const double *val, const double *b, const double *x0,
               int32_t max_iter, double tol, double *x);

/* ---- IC(0) (IncompleteCholesky.lean) ---------------------------------- */

/* factor3 (:49-69) on a dense 3x3 with keep-pattern. */
void orc_ic0_factor3(const double *A, const double *pattern, double *L);

/* No-fill IC(0) on a CSR SPD matrix (:26-33 describe the CSR form): L has
 * the pattern of tril(A) incl. diagonal, CSR rows, columns ascending,
 *   L_ij = (A_ij - sum_{k<j} L_ik L_jk) / L_jj   (k ascending, :60-68)
 *   L_jj = sqrt(A_jj - sum_{k<j} L_jk^2)          (:53-58)
 * l_ptr[n+1], l_col / l_val sized nnz_lower.  Returns 0, or -(row+1) at the
 * first row whose radicand is <= 0 (not SPD under IC(0)). */
int32_t orc_ic0_factor(int32_t n, const int32_t *ptr, const int32_t *col,
                       const double *val, int32_t *l_ptr, int32_t *l_col,
                       double *l_val);
/* forwardSub / backwardSub (:83-101) on the CSR factor. */
void orc_ic0_solve(int32_t n, const int32_t *l_ptr, const int32_t *l_col,
                   const double *l_val, const double *b, double *x);

/* Preconditioned CG, builder variant used as the CPU baseline: same loop as
 * SparseLinAlg.cg with the preconditioner swapped (precond 0 = Jacobi,
 * 1 = IC(0) with the given factor) and the stop rule ||r|| < tol_abs.
 * Returns iterations. */
int32_t orc_pcg(int32_t n, const int32_t *ptr, const int32_t *col,
                const double *val, int32_t precond, const int32_t *l_ptr,
                const int32_t *l_col, const double *l_val, const double *b,
                const double *x0, int32_t max_iter, double tol_abs, double *x);
int64_t orc_pcg_many(int32_t n, const int32_t *ptr, const int32_t *col, const double *val,
                     int32_t precond, const int32_t *l_ptr, const int32_t *l_col,
                     const double *l_val, int32_t k, const double *B, int32_t max_iter,
                     double tol_rel, double *X, int32_t nthreads);

/* Symmetric permutation B = P A P^T with perm[new] = old; columns sorted. */
void orc_permute_csr(int32_t n, const int32_t *ptr, const int32_t *col,
                     const double *val, const int32_t *perm, int32_t *optr,
                     int32_t *ocol, double *oval);

/* ---- Graph colouring (GraphColoring.lean:41-56) ------------------------
 * Adjacency given as a CSR pattern; the diagonal is ignored.  Returns the
 * number of colours. */
int32_t orc_greedy_color(int32_t n, const int32_t *ptr, const int32_t *col,
                         int32_t *color);
void orc_sgs_sweep_f64(int32_t n, const int32_t *ptr, const int32_t *col, const double *val,
                       const double *diag, int32_t ncolors, const int32_t *color_ptr,
                       const int32_t *color_rows, int32_t k, const double *b, double *x,
                       int32_t sweeps);
void orc_sgs_sweep_f32(int32_t n, const int32_t *ptr, const int32_t *col, const float *val,
                       const float *diag, int32_t ncolors, const int32_t *color_ptr,
                       const int32_t *color_rows, int32_t k, const float *b, float *x,
                       int32_t sweeps);
void orc_restrict_sum_f64(int32_t n_fine, const int32_t *parent, int32_t n_coarse, int32_t k,
                          const double *fine, double *coarse);
void orc_restrict_sum_f32(int32_t n_fine, const int32_t *parent, int32_t n_coarse, int32_t k,
                          const float *fine, float *coarse);

/* ---- Dense linear algebra (DenseLinAlg.lean) -------------------------- */
void orc_dense_solve(int32_t n, const double *A, const double *b,
                     double *x);                                   /* :70-111 */
void orc_dense_solve_multi(int32_t n, int32_t k, const double *A,
                           const double *B, double *X);            /* :122-131 */
void orc_lu_factor(int32_t n, const double *A, double *LU,
                   int32_t *piv);                                  /* :139-166 */
void orc_lu_solve(int32_t n, const double *LU, const int32_t *piv,
                  const double *b, double *x);                     /* :169-190 */
void orc_cholesky_factor(int32_t n, const double *A, double *L);   /* :226-241 */
void orc_cholesky_solve(int32_t n, const double *L, const double *b,
                        double *x);                                /* :244-275 */

/* ---- HSC (HierarchicalSparsifyCompensate.lean:54-125) -----------------
 * Edge list (u < v, w).  eliminate writes at most ne + deg^2/2 edges and
 * returns the new edge count. */
double orc_hsc_degree_weight(int32_t ne, const int32_t *eu, const int32_t *ev,
                             const double *ew, int32_t v);
double orc_hsc_edge_weight(int32_t ne, const int32_t *eu, const int32_t *ev,
                           const double *ew, int32_t a, int32_t b);
int32_t orc_hsc_eliminate_vertex(int32_t ne, const int32_t *eu,
                                 const int32_t *ev, const double *ew,
                                 int32_t v, int32_t *ou, int32_t *ov,
                                 double *ow);

/* ---- Direct Delta Mush (DirectDeltaMush.lean, DirectDeltaMushBind.lean) --
 * Transforms are 4x4 row-major (16 doubles each).  Influences are CSR:
 * start[nv+1], idx[], w[].  Weight matrices are dense row-major n x nh. */
void orc_apply_transform4x4(const double *T, double x, double y, double z,
                            double *out);                          /* :69-73 */
void orc_lbs_matvec(int32_t nv, const double *rest, const double *T, int32_t nt,
                    const int32_t *start, const int32_t *idx, const double *w,
                    double *out);                                  /* :85-98 */
void orc_smooth_weights(int32_t n, int32_t nh, const int32_t *adj_ptr,
                        const int32_t *adj, const double *W, int32_t nu, double omega,
                        double *out);                              /* Bind :76-97 */
/* take = min(K, nh) pairs per row: idx[n*take], w[n*take] */
void orc_sparsify_topk(int32_t n, int32_t nh, const double *W, int32_t K, int32_t *idx,
                       double *w);                                 /* Bind :107-136 */

#ifdef __cplusplus
}
#endif
#endif
