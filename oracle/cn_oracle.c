/*
 * cn_oracle.c -- FP64 CPU restatement of the profile-curves spec.
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 * See cn_oracle.h for conventions.  All citations are to
 * V-Sekai-fire/TOOL_godot_curvenet, lean/Curvenet/<file>.lean:<line>.
 *
 * Compile with: -O2 -ffp-contract=off -fno-fast-math
 */
#include "cn_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Halfedge builder
 * ==================================================================== */

/* HalfedgeBuilder.sourceOfInterior (HalfedgeBuilder.lean:25-29). */
static int32_t src_of_interior(const int32_t *tris, int32_t h) {
  int32_t f = h / 3, l = h % 3;
  return tris[3 * f + l];
}

/* HalfedgeBuilder.fromTriangles (HalfedgeBuilder.lean:34-90), verbatim
 * loop structure including the first-match twin scan (:46-62), orphan
 * order (:63-71) and the boundary `next` scan with self-loop fallback
 * (:72-89). */
int32_t orc_he_from_triangles_spec(int32_t nv, int32_t nf, const int32_t *tris,
                                   int32_t *target, int32_t *twin,
                                   int32_t *next, int32_t *face) {
  (void)nv;
  int32_t nint = 3 * nf;
  for (int32_t f = 0; f < nf; ++f) {                    /* step 1, :39-45 */
    int32_t t0 = tris[3 * f], t1 = tris[3 * f + 1], t2 = tris[3 * f + 2];
    int32_t b = 3 * f;
    target[b] = t1;     twin[b] = -1;     next[b] = b + 1;     face[b] = f;
    target[b + 1] = t2; twin[b + 1] = -1; next[b + 1] = b + 2; face[b + 1] = f;
    target[b + 2] = t0; twin[b + 2] = -1; next[b + 2] = b;     face[b + 2] = f;
  }
  int32_t *need = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nint + 1));
  int32_t nneed = 0;
  for (int32_t i = 0; i < nint; ++i) {                  /* step 2, :46-62 */
    int32_t tgt = target[i], src = src_of_interior(tris, i);
    int32_t found = -1;
    for (int32_t j = 0; j < nint; ++j) {
      if (target[j] == src && src_of_interior(tris, j) == tgt) {
        found = j;
        break;
      }
    }
    if (found >= 0) twin[i] = found;
    else need[nneed++] = i;
  }
  int32_t bstart = nint, n = nint;
  for (int32_t kk = 0; kk < nneed; ++kk) {              /* step 3, :63-71 */
    int32_t i = need[kk];
    int32_t bh = n++;
    target[bh] = src_of_interior(tris, i);
    twin[bh] = i;
    next[bh] = 0;
    face[bh] = -1;
    twin[i] = bh;
  }
  for (int32_t i = bstart; i < n; ++i) {                /* step 4, :72-89 */
    int32_t my = target[i], nx = i;
    for (int32_t j = bstart; j < n; ++j) {
      if (twin[j] >= 0 && target[twin[j]] == my) {
        nx = j;
        break;
      }
    }
    next[i] = nx;
  }
  free(need);
  return n;
}

typedef struct { int64_t key; int32_t idx; } keyidx;

static int cmp_keyidx(const void *a, const void *b) {
  const keyidx *x = (const keyidx *)a, *y = (const keyidx *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* first element with key == k in a (key, idx)-sorted array, or -1 */
static int32_t first_with_key(const keyidx *a, int32_t n, int64_t k) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    int32_t mid = lo + (hi - lo) / 2;
    if (a[mid].key < k) lo = mid + 1; else hi = mid;
  }
  return (lo < n && a[lo].key == k) ? a[lo].idx : -1;
}

/* Index-identical to the spec builder: "first j" in a scan over j is the
 * smallest j with the key, which is what a (key, index) sort yields. */
int32_t orc_he_from_triangles(int32_t nv, int32_t nf, const int32_t *tris,
                              int32_t *target, int32_t *twin, int32_t *next,
                              int32_t *face) {
  (void)nv;
  int32_t nint = 3 * nf;
  for (int32_t f = 0; f < nf; ++f) {
    int32_t t0 = tris[3 * f], t1 = tris[3 * f + 1], t2 = tris[3 * f + 2];
    int32_t b = 3 * f;
    target[b] = t1;     twin[b] = -1;     next[b] = b + 1;     face[b] = f;
    target[b + 1] = t2; twin[b + 1] = -1; next[b + 1] = b + 2; face[b + 1] = f;
    target[b + 2] = t0; twin[b + 2] = -1; next[b + 2] = b;     face[b + 2] = f;
  }
  keyidx *ks = (keyidx *)malloc(sizeof(keyidx) * (size_t)(nint + 1));
  for (int32_t j = 0; j < nint; ++j) {
    ks[j].key = ((int64_t)src_of_interior(tris, j) << 32) | (uint32_t)target[j];
    ks[j].idx = j;
  }
  qsort(ks, (size_t)nint, sizeof(keyidx), cmp_keyidx);
  int32_t n = nint;
  for (int32_t i = 0; i < nint; ++i) {
    int64_t want = ((int64_t)target[i] << 32) | (uint32_t)src_of_interior(tris, i);
    int32_t j = first_with_key(ks, nint, want);
    if (j >= 0) twin[i] = j;
  }
  for (int32_t i = 0; i < nint; ++i) {
    if (twin[i] < 0) {
      int32_t bh = n++;
      target[bh] = src_of_interior(tris, i);
      twin[bh] = i;
      next[bh] = 0;
      face[bh] = -1;
      twin[i] = bh;
    }
  }
  int32_t nb = n - nint;
  keyidx *bk = (keyidx *)malloc(sizeof(keyidx) * (size_t)(nb + 1));
  for (int32_t j = 0; j < nb; ++j) {
    bk[j].key = target[twin[nint + j]];
    bk[j].idx = nint + j;
  }
  qsort(bk, (size_t)nb, sizeof(keyidx), cmp_keyidx);
  for (int32_t i = nint; i < n; ++i) {
    int32_t j = first_with_key(bk, nb, target[i]);
    next[i] = j >= 0 ? j : i;
  }
  free(ks);
  free(bk);
  return n;
}

/* HalfedgeMesh.manifold? (Halfedge.lean:42-128). */
int32_t orc_he_manifold(int32_t nv, int32_t nf, int32_t nh,
                        const int32_t *target, const int32_t *twin,
                        const int32_t *next, const int32_t *face) {
  int32_t mask = 0, ok = 1;
  for (int32_t i = 0; i < nh && ok; ++i) {              /* :45-57 */
    if (target[i] < 0 || target[i] >= nv) ok = 0;
    if (next[i] < 0 || next[i] >= nh) ok = 0;
    if (twin[i] >= nh) ok = 0;
    if (face[i] >= nf) ok = 0;
  }
  if (ok) mask |= 1;
  if (!ok) return mask; /* later checks index arrays */
  ok = 1;
  for (int32_t i = 0; i < nh && ok; ++i) {              /* :60-70 */
    int32_t t = twin[i];
    if (t >= 0 && twin[t] != i) ok = 0;
  }
  if (ok) mask |= 2;
  ok = 1;
  {                                                     /* :78-94 */
    int32_t *prev = (int32_t *)calloc((size_t)nh + 1, sizeof(int32_t));
    for (int32_t i = 0; i < nh; ++i) prev[next[i]] = i;
    for (int32_t i = 0; i < nh && ok; ++i) {
      int32_t t = twin[i];
      if (t < 0) continue;
      if (target[prev[i]] != target[t]) ok = 0;
    }
    free(prev);
  }
  if (ok) mask |= 4;
  ok = 1;
  for (int32_t s = 0; s < nh && ok; ++s) {              /* :99-118 */
    int32_t f0 = face[s];
    if (f0 < 0) continue;
    int32_t cur = next[s], steps = 0, closed = 0;
    while (steps < nh) {
      if (cur == s) { closed = 1; break; }
      if (face[cur] != f0) { ok = 0; break; }
      cur = next[cur];
      ++steps;
    }
    if (!closed) ok = 0;
  }
  if (ok) mask |= 8;
  return mask;
}

/* ======================================================================
 * Cut surgery (CutAlgorithm.lean)
 * ==================================================================== */

/* subdivideEdge (CutAlgorithm.lean:49-94). */
int32_t orc_subdivide_edge(int32_t *nv, int32_t *nf, int32_t nh, int32_t h,
                           int32_t *target, int32_t *twin, int32_t *next,
                           int32_t *face) {
  (void)nf;
  int32_t ot = target[h], otw = twin[h], onx = next[h], ofc = face[h];
  int32_t ti = otw >= 0 ? otw : h;
  int32_t tt = target[ti], ttw = twin[ti], tnx = next[ti], tfc = face[ti];
  (void)ttw;
  int32_t c = *nv, hn = nh, tn = nh + 1;
  target[h] = c; twin[h] = tn; next[h] = hn; face[h] = ofc;
  target[ti] = c; twin[ti] = hn; next[ti] = tn; face[ti] = tfc;
  target[hn] = ot; twin[hn] = ti; next[hn] = onx; face[hn] = ofc;
  target[tn] = tt; twin[tn] = h; next[tn] = tnx; face[tn] = tfc;
  *nv += 1;
  return nh + 2;
}

/* splitFace (CutAlgorithm.lean:115-163). */
int32_t orc_split_face(int32_t *nv, int32_t *nf, int32_t nh, int32_t a,
                       int32_t b, int32_t *target, int32_t *twin,
                       int32_t *next, int32_t *face) {
  (void)nv;
  int32_t at = target[a], aw = twin[a], an = next[a], af = face[a];
  int32_t bt = target[b], bw = twin[b], bn = next[b], bf = face[b];
  int32_t fn = *nf, hab = nh, hba = nh + 1;
  int32_t cur = bn, steps = 0;
  while (cur != a && steps < nh) {
    int32_t nx = next[cur];
    face[cur] = fn;
    cur = nx;
    ++steps;
  }
  target[a] = at; twin[a] = aw; face[a] = fn; next[a] = hab;
  target[b] = bt; twin[b] = bw; face[b] = bf; next[b] = hba;
  target[hab] = bt; twin[hab] = hba; next[hab] = bn; face[hab] = fn;
  target[hba] = at; twin[hba] = hab; next[hba] = an; face[hba] = af;
  *nf += 1;
  return nh + 2;
}

/* insertCrack (CutAlgorithm.lean:212-243). */
int32_t orc_insert_crack(int32_t *nv, int32_t *nf, int32_t nh, int32_t h_anchor,
                         int32_t *target, int32_t *twin, int32_t *next,
                         int32_t *face) {
  (void)nf;
  int32_t vb = target[h_anchor], on = next[h_anchor], fa = face[h_anchor];
  int32_t c = *nv, hout = nh, hin = nh + 1;
  next[h_anchor] = hout;
  target[hout] = c; twin[hout] = hin; next[hout] = hin; face[hout] = fa;
  target[hin] = vb; twin[hin] = hout; next[hin] = on; face[hin] = fa;
  *nv += 1;
  return nh + 2;
}

/* ======================================================================
 * Seam cut (builder-defined; see header)
 * ==================================================================== */

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : (x > y);
}

static int64_t ekey(int32_t u, int32_t v) {
  int32_t a = u < v ? u : v, b = u < v ? v : u;
  return ((int64_t)a << 32) | (uint32_t)b;
}

static int has_key(const int64_t *a, int32_t n, int64_t k) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    int32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo < n && a[lo] == k;
}

static int32_t uf_find(int32_t *p, int32_t x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

int32_t orc_cut_edges(int32_t nv, int32_t nf, const int32_t *tris,
                      int32_t ncut, const int32_t *cut_edges,
                      int32_t *new_tris, int32_t *origin,
                      uint8_t *constrained) {
  int32_t cap = 6 * nf + 1;
  int32_t *tg = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  int32_t *tw = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  int32_t *nx = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  int32_t *fc = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  orc_he_from_triangles(nv, nf, tris, tg, tw, nx, fc);

  /* mesh edge set, to validate the cut edges */
  int64_t *mk = (int64_t *)malloc(sizeof(int64_t) * (size_t)(3 * nf + 1));
  for (int32_t h = 0; h < 3 * nf; ++h)
    mk[h] = ekey(src_of_interior(tris, h), tg[h]);
  qsort(mk, (size_t)(3 * nf), sizeof(int64_t), cmp_i64);
  int64_t *ck = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncut + 1));
  uint8_t *on_cut = (uint8_t *)calloc((size_t)nv + 1, 1);
  for (int32_t e = 0; e < ncut; ++e) {
    int32_t u = cut_edges[2 * e], v = cut_edges[2 * e + 1];
    if (u < 0 || v < 0 || u >= nv || v >= nv || u == v) goto bad;
    ck[e] = ekey(u, v);
    if (!has_key(mk, 3 * nf, ck[e])) goto bad;
    on_cut[u] = on_cut[v] = 1;
  }
  qsort(ck, (size_t)ncut, sizeof(int64_t), cmp_i64);

  /* corners of each vertex, ascending slot (counting sort) */
  int32_t *cptr = (int32_t *)calloc((size_t)nv + 1, sizeof(int32_t));
  for (int32_t s = 0; s < 3 * nf; ++s) cptr[tris[s] + 1]++;
  for (int32_t v = 0; v < nv; ++v) cptr[v + 1] += cptr[v];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  memcpy(fill, cptr, sizeof(int32_t) * (size_t)nv);
  int32_t *cl = (int32_t *)malloc(sizeof(int32_t) * (size_t)(3 * nf + 1));
  for (int32_t s = 0; s < 3 * nf; ++s) cl[fill[tris[s]]++] = s;

  /* union-find over corner slots */
  int32_t *par = (int32_t *)malloc(sizeof(int32_t) * (size_t)(3 * nf + 1));
  for (int32_t s = 0; s < 3 * nf; ++s) par[s] = s;
  for (int32_t s = 0; s < 3 * nf; ++s) {
    int32_t v = tris[s];
    if (!on_cut[v]) continue;
    int32_t f = s / 3, l = s % 3;
    /* incoming halfedge into v inside face f */
    int32_t p = 3 * f + (l + 2) % 3;
    int32_t o = nx[p];            /* v -> w in face f */
    int32_t w = tg[o];
    int32_t t = tw[o];            /* w -> v in the neighbouring face */
    if (t < 0 || fc[t] < 0) continue;
    if (has_key(ck, ncut, ekey(v, w))) continue;
    int32_t f2 = fc[t];
    int32_t s2 = -1;
    for (int32_t m = 0; m < 3; ++m)
      if (tris[3 * f2 + m] == v) s2 = 3 * f2 + m;
    if (s2 < 0) continue;
    int32_t ra = uf_find(par, s), rb = uf_find(par, s2);
    if (ra != rb) {
      if (ra < rb) par[rb] = ra; else par[ra] = rb; /* root = min slot */
    }
  }
  for (int32_t v = 0; v < nv; ++v) { origin[v] = v; constrained[v] = on_cut[v]; }
  memcpy(new_tris, tris, sizeof(int32_t) * (size_t)(3 * nf));
  int32_t nvn = nv;
  int32_t *roots = (int32_t *)malloc(sizeof(int32_t) * (size_t)(3 * nf + 1));
  int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(3 * nf + 1));
  for (int32_t v = 0; v < nv; ++v) {
    if (!on_cut[v]) continue;
    int32_t nr = 0;
    /* distinct wedge roots in ascending order (root = min slot) */
    for (int32_t c = cptr[v]; c < cptr[v + 1]; ++c) {
      int32_t r = uf_find(par, cl[c]);
      int32_t seen = 0;
      for (int32_t q = 0; q < nr; ++q) if (roots[q] == r) { seen = 1; break; }
      if (!seen) roots[nr++] = r;
    }
    for (int32_t a = 1; a < nr; ++a) {               /* insertion sort */
      int32_t x = roots[a], b = a - 1;
      while (b >= 0 && roots[b] > x) { roots[b + 1] = roots[b]; --b; }
      roots[b + 1] = x;
    }
    for (int32_t q = 0; q < nr; ++q) {
      if (q == 0) ids[q] = v;
      else {
        ids[q] = nvn;
        origin[nvn] = v;
        constrained[nvn] = 1;
        ++nvn;
      }
    }
    for (int32_t c = cptr[v]; c < cptr[v + 1]; ++c) {
      int32_t r = uf_find(par, cl[c]);
      for (int32_t q = 0; q < nr; ++q)
        if (roots[q] == r) { new_tris[cl[c]] = ids[q]; break; }
    }
  }
  free(roots); free(ids); free(par); free(cl); free(fill); free(cptr);
  free(on_cut); free(ck); free(mk); free(tg); free(tw); free(nx); free(fc);
  return nvn;
bad:
  free(on_cut); free(ck); free(mk); free(tg); free(tw); free(nx); free(fc);
  return -2;
}

/* ======================================================================
 * Cotangent Laplacian
 * ==================================================================== */

/* PolygonLaplacian.cotAtVertex (PolygonLaplacian.lean:32-40). */
double orc_cot_at_vertex(const double *apex, const double *o1,
                         const double *o2) {
  double ax = o1[0] - apex[0], ay = o1[1] - apex[1], az = o1[2] - apex[2];
  double bx = o2[0] - apex[0], by = o2[1] - apex[1], bz = o2[2] - apex[2];
  double dot = ax * bx + ay * by + az * bz;
  double cx = ay * bz - az * by;
  double cy = az * bx - ax * bz;
  double cz = ax * by - ay * bx;
  double cl = sqrt(cx * cx + cy * cy + cz * cz);
  return dot / cl;
}

/* triangleCotLaplacian (PolygonLaplacian.lean:56-71). */
void orc_triangle_cot_laplacian(const double *a, const double *b,
                                const double *c, double *m) {
  double cotA = orc_cot_at_vertex(a, b, c);
  double cotB = orc_cot_at_vertex(b, a, c);
  double cotC = orc_cot_at_vertex(c, a, b);
  double w_bc = cotA * 0.5, w_ac = cotB * 0.5, w_ab = cotC * 0.5;
  m[1] = -w_ab; m[3] = -w_ab;
  m[2] = -w_ac; m[6] = -w_ac;
  m[5] = -w_bc; m[7] = -w_bc;
  m[0] = w_ab + w_ac;
  m[4] = w_ab + w_bc;
  m[8] = w_ac + w_bc;
}

/* RobustLaplacian (RobustLaplacian.lean:58-140): Sharp & Crane intrinsic
 * mollification.  Edge lengths from the embedding (:63-70), the smallest
 * eps >= 0 that makes every corner satisfy a + b > c + delta (:83-89, per
 * triangle, as the code does), cot from the mollified lengths by the law of
 * cosines over Heron's area (:106-109), the 3x3 block as
 * triangleCotLaplacian (:114-131). */
static double max_f(double a, double b) { return a <= b ? b : a; }

static double edge_len(const double *u, const double *v) {
  double dx = u[0] - v[0], dy = u[1] - v[1], dz = u[2] - v[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

double orc_mollify_triangle(double delta, double eA, double eB, double eC) {
  double ra = delta + eA - eB - eC;
  double rb = delta + eB - eA - eC;
  double rc = delta + eC - eA - eB;
  return max_f(0.0, max_f(ra, max_f(rb, rc)));
}

double orc_cot_from_lengths(double eo, double e1, double e2) {
  double s = (eo + e1 + e2) * 0.5;
  double area = sqrt(s * (s - eo) * (s - e1) * (s - e2));
  return (e1 * e1 + e2 * e2 - eo * eo) / (4.0 * area);
}

void orc_robust_triangle_cot_laplacian(double delta, const double *a, const double *b,
                                       const double *c, double *m) {
  double eA = edge_len(b, c), eB = edge_len(a, c), eC = edge_len(a, b);
  double eps = orc_mollify_triangle(delta, eA, eB, eC);
  eA = eA + eps;
  eB = eB + eps;
  eC = eC + eps;
  double cotA = orc_cot_from_lengths(eA, eB, eC);
  double cotB = orc_cot_from_lengths(eB, eA, eC);
  double cotC = orc_cot_from_lengths(eC, eA, eB);
  double w_bc = cotA * 0.5, w_ac = cotB * 0.5, w_ab = cotC * 0.5;
  m[1] = -w_ab; m[3] = -w_ab;
  m[2] = -w_ac; m[6] = -w_ac;
  m[5] = -w_bc; m[7] = -w_bc;
  m[0] = w_ab + w_ac;
  m[4] = w_ab + w_bc;
  m[8] = w_ac + w_bc;
}

/* Which triangle Laplacian the assembly uses: 0 = embedded cot
 * (triangleCotLaplacian), 1 = robust with margin delta.  A setting of the
 * checker, so that the assembly functions keep their spec signatures. */
static int g_lap_kind = 0;
static double g_lap_delta = 0.0;
void orc_set_laplacian(int32_t kind, double delta) {
  g_lap_kind = kind;
  g_lap_delta = delta;
}

/* polygonCotLaplacian (PolygonLaplacian.lean:77-93): fan from vertex 0. */
void orc_polygon_cot_laplacian(int32_t n, const double *poly, double *m) {
  for (int32_t i = 0; i < n * n; ++i) m[i] = 0.0;
  if (n < 3) return;
  double tri[9];
  for (int32_t i = 1; i < n - 1; ++i) {
    if (g_lap_kind == 1)
      orc_robust_triangle_cot_laplacian(g_lap_delta, poly, poly + 3 * i, poly + 3 * (i + 1), tri);
    else
      orc_triangle_cot_laplacian(poly, poly + 3 * i, poly + 3 * (i + 1), tri);
    int32_t idx[3] = {0, i, i + 1};
    for (int32_t li = 0; li < 3; ++li)
      for (int32_t lj = 0; lj < 3; ++lj)
        m[idx[li] * n + idx[lj]] = m[idx[li] * n + idx[lj]] + tri[li * 3 + lj];
  }
}

/* CutMeshLaplacian.faceLoop (CutMeshLaplacian.lean:29-46): halfedges of
 * face f in `next` order starting from the lowest-index halfedge of f.
 * `first` is the precomputed lowest halfedge per face (-1 = none). */
static int32_t face_loop(int32_t nh, const int32_t *next, int32_t start,
                         int32_t *out) {
  if (start < 0) return 0;
  int32_t n = 0;
  out[n++] = start;
  int32_t cur = next[start], steps = 0;
  while (cur != start && steps < nh) {
    out[n++] = cur;
    cur = next[cur];
    ++steps;
  }
  return n;
}

typedef struct {
  int32_t *first;     /* lowest halfedge per face */
  int32_t *loc;       /* local index of halfedge in its face loop, -1 */
  int32_t *loop_ptr;  /* per face: offset into loop_he */
  int32_t *loop_he;   /* face loops */
  int64_t *mat_ptr;   /* per face: offset into mat (n_f^2 entries) */
  double *mat;        /* per-face polygon Laplacian, 0 size when n_f < 3 */
} face_tables;

static void build_face_tables(int32_t nf, int32_t nh, const int32_t *target,
                              const int32_t *next, const int32_t *face,
                              const double *pos, face_tables *T) {
  T->first = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nf + 1));
  for (int32_t f = 0; f < nf; ++f) T->first[f] = -1;
  for (int32_t h = 0; h < nh; ++h) {
    int32_t f = face[h];
    if (f >= 0 && T->first[f] < 0) T->first[f] = h;
  }
  T->loc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  for (int32_t h = 0; h < nh; ++h) T->loc[h] = -1;
  T->loop_ptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nf + 1));
  T->loop_he = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1) * 2);
  T->mat_ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nf + 1));
  int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  int32_t off = 0;
  int64_t moff = 0;
  for (int32_t f = 0; f < nf; ++f) {
    int32_t n = face_loop(nh, next, T->first[f], tmp);
    T->loop_ptr[f] = off;
    memcpy(T->loop_he + off, tmp, sizeof(int32_t) * (size_t)n);
    off += n;
    T->mat_ptr[f] = moff;
    if (n >= 3) moff += (int64_t)n * n;
  }
  T->loop_ptr[nf] = off;
  T->mat_ptr[nf] = moff;
  T->mat = (double *)malloc(sizeof(double) * (size_t)(moff + 1));
  double *poly = (double *)malloc(sizeof(double) * 3 * (size_t)(nh + 1));
  for (int32_t f = 0; f < nf; ++f) {
    int32_t a = T->loop_ptr[f], n = T->loop_ptr[f + 1] - a;
    for (int32_t i = 0; i < n; ++i) {
      int32_t h = T->loop_he[a + i];
      if (face[h] == f && T->loc[h] < 0) T->loc[h] = i;
      for (int32_t d = 0; d < 3; ++d) poly[3 * i + d] = pos[3 * target[h] + d];
    }
    if (n >= 3) orc_polygon_cot_laplacian(n, poly, T->mat + T->mat_ptr[f]);
  }
  free(poly);
  free(tmp);
}

static void free_face_tables(face_tables *T) {
  free(T->first); free(T->loc); free(T->loop_ptr); free(T->loop_he);
  free(T->mat_ptr); free(T->mat);
}

/* Lh[p][q] per assembleLh (CutMeshLaplacian.lean:56-71): the face polygon
 * Laplacian entry when p and q share a face with >= 3 corners, else 0. */
static double lh_entry(const face_tables *T, const int32_t *face, int32_t p,
                       int32_t q) {
  int32_t f = face[p];
  if (f < 0 || face[q] != f) return 0.0;
  int32_t n = T->loop_ptr[f + 1] - T->loop_ptr[f];
  if (n < 3) return 0.0;
  int32_t li = T->loc[p], lj = T->loc[q];
  if (li < 0 || lj < 0) return 0.0;
  return 0.0 + T->mat[T->mat_ptr[f] + (int64_t)li * n + lj];
}

/* Dense assembleVtLhV (CutMeshLaplacian.lean:56-93) with DenseLinAlg.matMul
 * (DenseLinAlg.lean:38-46) and transpose (:49-54), transcribed. */
void orc_assemble_vtlhv_dense(int32_t nv, int32_t nf, int32_t nh,
                              const int32_t *target, const int32_t *next,
                              const int32_t *face, const uint8_t *is_sample,
                              const double *pos, double *out) {
  face_tables T;
  build_face_tables(nf, nh, target, next, face, pos, &T);
  double *L = (double *)calloc((size_t)nh * nh, sizeof(double));
  for (int32_t f = 0; f < nf; ++f) {                         /* assembleLh */
    int32_t a = T.loop_ptr[f], n = T.loop_ptr[f + 1] - a;
    if (n < 3) continue;
    const double *Lf = T.mat + T.mat_ptr[f];
    for (int32_t li = 0; li < n; ++li)
      for (int32_t lj = 0; lj < n; ++lj) {
        int32_t gi = T.loop_he[a + li], gj = T.loop_he[a + lj];
        L[(size_t)gi * nh + gj] = L[(size_t)gi * nh + gj] + Lf[li * n + lj];
      }
  }
  double *V = (double *)calloc((size_t)nh * nv, sizeof(double)); /* assembleV */
  for (int32_t h = 0; h < nh; ++h)
    if (!is_sample[target[h]]) V[(size_t)h * nv + target[h]] = 1.0;
  double *LV = (double *)calloc((size_t)nh * nv, sizeof(double));
  for (int32_t i = 0; i < nh; ++i)
    for (int32_t j = 0; j < nv; ++j) {
      double s = 0.0;
      for (int32_t p = 0; p < nh; ++p) s = s + L[(size_t)i * nh + p] * V[(size_t)p * nv + j];
      LV[(size_t)i * nv + j] = s;
    }
  for (int32_t i = 0; i < nv; ++i)
    for (int32_t j = 0; j < nv; ++j) {
      double s = 0.0;
      for (int32_t p = 0; p < nh; ++p) s = s + V[(size_t)p * nv + i] * LV[(size_t)p * nv + j];
      out[(size_t)i * nv + j] = s;
    }
  free(L); free(V); free(LV);
  free_face_tables(&T);
}

/* HarmonicSolve.solveMulti RHS (HarmonicSolve.lean:35-45, 57-64). */
void orc_harmonic_rhs_dense(int32_t nv, int32_t nf, int32_t nh,
                            const int32_t *target, const int32_t *next,
                            const int32_t *face, const int32_t *col_of_vertex,
                            const double *pos, int32_t k, const double *Fc,
                            double *rhs) {
  face_tables T;
  build_face_tables(nf, nh, target, next, face, pos, &T);
  double *L = (double *)calloc((size_t)nh * nh, sizeof(double));
  for (int32_t f = 0; f < nf; ++f) {
    int32_t a = T.loop_ptr[f], n = T.loop_ptr[f + 1] - a;
    if (n < 3) continue;
    const double *Lf = T.mat + T.mat_ptr[f];
    for (int32_t li = 0; li < n; ++li)
      for (int32_t lj = 0; lj < n; ++lj) {
        int32_t gi = T.loop_he[a + li], gj = T.loop_he[a + lj];
        L[(size_t)gi * nh + gj] = L[(size_t)gi * nh + gj] + Lf[li * n + lj];
      }
  }
  double *CF = (double *)calloc((size_t)nh * k, sizeof(double)); /* computeCFc */
  for (int32_t h = 0; h < nh; ++h) {
    int32_t c = col_of_vertex[target[h]];
    if (c >= 0)
      for (int32_t j = 0; j < k; ++j) CF[(size_t)h * k + j] = Fc[(size_t)c * k + j];
  }
  double *LCF = (double *)calloc((size_t)nh * k, sizeof(double));
  for (int32_t i = 0; i < nh; ++i)
    for (int32_t j = 0; j < k; ++j) {
      double s = 0.0;
      for (int32_t p = 0; p < nh; ++p) s = s + L[(size_t)i * nh + p] * CF[(size_t)p * k + j];
      LCF[(size_t)i * k + j] = s;
    }
  for (int32_t i = 0; i < nv; ++i)
    for (int32_t j = 0; j < k; ++j) {
      double s = 0.0;
      for (int32_t p = 0; p < nh; ++p) {
        double vt = (col_of_vertex[target[p]] < 0 && target[p] == i) ? 1.0 : 0.0;
        s = s + vt * LCF[(size_t)p * k + j];
      }
      rhs[(size_t)i * k + j] = -s;
    }
  free(L); free(CF); free(LCF);
  free_face_tables(&T);
}

/* incoming halfedges per vertex, ascending halfedge index */
static void incoming_lists(int32_t nv, int32_t nh, const int32_t *target,
                           int32_t **ptr_out, int32_t **he_out) {
  int32_t *ptr = (int32_t *)calloc((size_t)nv + 2, sizeof(int32_t));
  for (int32_t h = 0; h < nh; ++h) ptr[target[h] + 1]++;
  for (int32_t v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  memcpy(fill, ptr, sizeof(int32_t) * (size_t)nv);
  int32_t *he = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  for (int32_t h = 0; h < nh; ++h) he[fill[target[h]]++] = h;
  free(fill);
  *ptr_out = ptr;
  *he_out = he;
}

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return x < y ? -1 : (x > y);
}

/* Sparse VtLhV: entry (i,j) = sum_{p in inc(i) asc} sum_{q in face(p) asc,
 * target q = j} Lh[p][q], which is exactly the value the two dense matMul
 * loops produce (the skipped terms are products with 0 added to a sum that
 * starts at +0.0 and is never -0.0). */
int32_t orc_cut_laplacian_csr(int32_t nv, int32_t nf, int32_t nh,
                              const int32_t *target, const int32_t *next,
                              const int32_t *face, const uint8_t *constrained,
                              const double *pos, int32_t *free_of,
                              int32_t *con_of, int32_t *ff_ptr,
                              int32_t *ff_col, double *ff_val,
                              int32_t *fc_ptr, int32_t *fc_col,
                              double *fc_val, int64_t cap, int64_t *nnz_out) {
  face_tables T;
  build_face_tables(nf, nh, target, next, face, pos, &T);
  int32_t *iptr, *ihe;
  incoming_lists(nv, nh, target, &iptr, &ihe);
  int32_t nfree = 0, ncon = 0;
  for (int32_t v = 0; v < nv; ++v) {
    if (constrained[v]) { free_of[v] = -1; con_of[v] = ncon++; }
    else { free_of[v] = nfree++; con_of[v] = -1; }
  }
  int32_t *cols = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 8));
  int64_t nff = 0, nfc = 0;
  int32_t row = 0;
  ff_ptr[0] = 0;
  fc_ptr[0] = 0;
  int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  for (int32_t v = 0; v < nv; ++v) {
    if (constrained[v]) continue;
    int32_t nc = 0;
    cols[nc++] = v;
    for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
      int32_t p = ihe[a], f = face[p];
      if (f < 0) continue;
      int32_t b = T.loop_ptr[f], n = T.loop_ptr[f + 1] - b;
      if (n < 3) continue;
      for (int32_t t = 0; t < n; ++t) cols[nc++] = target[T.loop_he[b + t]];
    }
    qsort(cols, (size_t)nc, sizeof(int32_t), cmp_i32);
    int32_t u = 0;
    for (int32_t t = 0; t < nc; ++t)
      if (t == 0 || cols[t] != cols[t - 1]) cols[u++] = cols[t];
    for (int32_t t = 0; t < u; ++t) {
      int32_t j = cols[t];
      double s = 0.0;
      for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
        int32_t p = ihe[a], f = face[p];
        if (f < 0) { s = s + 0.0; continue; }
        int32_t b = T.loop_ptr[f], n = T.loop_ptr[f + 1] - b;
        /* q in face(p), ascending halfedge index */
        memcpy(sorted, T.loop_he + b, sizeof(int32_t) * (size_t)n);
        qsort(sorted, (size_t)n, sizeof(int32_t), cmp_i32);
        double tsum = 0.0;
        for (int32_t e = 0; e < n; ++e)
          if (target[sorted[e]] == j) tsum = tsum + lh_entry(&T, face, p, sorted[e]);
        s = s + tsum;
      }
      if (!constrained[j]) {
        if (nff >= cap) goto overflow;
        ff_col[nff] = free_of[j];
        ff_val[nff] = s;
        ++nff;
      } else {
        if (nfc >= cap) goto overflow;
        fc_col[nfc] = con_of[j];
        fc_val[nfc] = s;
        ++nfc;
      }
    }
    ++row;
    ff_ptr[row] = (int32_t)nff;
    fc_ptr[row] = (int32_t)nfc;
  }
  nnz_out[0] = nff;
  nnz_out[1] = nfc;
  free(sorted); free(cols); free(iptr); free(ihe);
  free_face_tables(&T);
  return nfree;
overflow:
  free(sorted); free(cols); free(iptr); free(ihe);
  free_face_tables(&T);
  return -3;
}

/* RHS rows of HarmonicSolve.solveMulti restricted to free vertices:
 * rhs[i][c] = -( sum_{p in inc(i) asc} sum_{q in face(p) asc, target q
 * constrained} Lh[p][q] * Xc[con(target q)][c] ). */
void orc_harmonic_rhs_csr(int32_t nv, int32_t nh, const int32_t *target,
                          const int32_t *next, const int32_t *face,
                          const int32_t *free_of, const int32_t *con_of,
                          const double *pos, int32_t k, const double *Xc,
                          double *rhs) {
  int32_t nf = 0;
  for (int32_t h = 0; h < nh; ++h) if (face[h] + 1 > nf) nf = face[h] + 1;
  face_tables T;
  build_face_tables(nf, nh, target, next, face, pos, &T);
  int32_t *iptr, *ihe;
  incoming_lists(nv, nh, target, &iptr, &ihe);
  int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  for (int32_t v = 0; v < nv; ++v) {
    int32_t i = free_of[v];
    if (i < 0) continue;
    for (int32_t c = 0; c < k; ++c) {
      double s = 0.0;
      for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
        int32_t p = ihe[a], f = face[p];
        if (f < 0) continue;
        int32_t b = T.loop_ptr[f], n = T.loop_ptr[f + 1] - b;
        memcpy(sorted, T.loop_he + b, sizeof(int32_t) * (size_t)n);
        qsort(sorted, (size_t)n, sizeof(int32_t), cmp_i32);
        double tsum = 0.0;
        for (int32_t e = 0; e < n; ++e) {
          int32_t q = sorted[e], cc = con_of[target[q]];
          if (cc >= 0) tsum = tsum + lh_entry(&T, face, p, q) * Xc[(size_t)cc * k + c];
        }
        s = s + tsum;
      }
      rhs[(size_t)i * k + c] = -s;
    }
  }
  free(sorted); free(iptr); free(ihe);
  free_face_tables(&T);
}

/* ======================================================================
 * Deformation solve (DeformSolve.lean) and its curve inputs
 * (ScaledFrames.lean, EndToEndExample.lean)
 * ==================================================================== */

/* EndToEndExample.bezier3 (EndToEndExample.lean:52-60). */
void orc_bezier3(const double *P, double t, double *out) {
  double u = 1.0 - t;
  double c0 = u * u * u, c1 = 3.0 * u * u * t, c2 = 3.0 * u * t * t, c3 = t * t * t;
  for (int d = 0; d < 3; ++d)
    out[d] = c0 * P[d] + c1 * P[3 + d] + c2 * P[6 + d] + c3 * P[9 + d];
}

/* ScaledFrames.tangentLength (ScaledFrames.lean:39-44). */
static double tangent_length(const double *p, const double *q, double *t) {
  double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
  double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double inv = l == 0.0 ? 0.0 : 1.0 / l;
  for (int i = 0; i < 3; ++i) t[i] = d[i] * inv;
  return l;
}

/* mat3Mul (ScaledFrames.lean:57-64): row i . column j, three terms left to
 * right. */
static void mat3_mul(const double *a, const double *b, double *c) {
  for (int k = 0; k < 9; ++k) {
    int i = k / 3, j = k % 3;
    c[k] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
  }
}

/* smallestRotation (ScaledFrames.lean:74-86): R = I + K + K^2 / (1 + cos). */
static void smallest_rotation(const double *f, const double *t, double *R) {
  double cx = f[1] * t[2] - f[2] * t[1];
  double cy = f[2] * t[0] - f[0] * t[2];
  double cz = f[0] * t[1] - f[1] * t[0];
  double cosT = f[0] * t[0] + f[1] * t[1] + f[2] * t[2];
  double K[9] = {0.0, -cz, cy, cz, 0.0, -cx, -cy, cx, 0.0}, K2[9];
  mat3_mul(K, K, K2);
  double denom = 1.0 + cosT;
  double inv = denom == 0.0 ? 0.0 : 1.0 / denom;
  static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int k = 0; k < 9; ++k) R[k] = (I[k] + K[k]) + inv * K2[k];
}

/* isolatedSegmentGradient (ScaledFrames.lean:90-95): F = (l/l0) R(t0, t). */
void orc_isolated_segment_gradient(const double *rp, const double *rq, const double *pp,
                                   const double *pq, double *F) {
  double t0[3], t1[3], R[9];
  double l0 = tangent_length(rp, rq, t0);
  double l1 = tangent_length(pp, pq, t1);
  double r = l0 == 0.0 ? 0.0 : l1 / l0;
  smallest_rotation(t0, t1, R);
  for (int k = 0; k < 9; ++k) F[k] = r * R[k];
}

/* mat3Det / mat3Inv / deformationGradient (ScaledFrames.lean:108-150). */
double orc_mat3_det(const double *m) {
  double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
         i = m[8];
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

void orc_mat3_inv(const double *m, double *o) {
  double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
         i = m[8];
  double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  double inv = det == 0.0 ? 0.0 : 1.0 / det;
  o[0] = inv * (e * i - f * h); o[1] = inv * (c * h - b * i); o[2] = inv * (b * f - c * e);
  o[3] = inv * (f * g - d * i); o[4] = inv * (a * i - c * g); o[5] = inv * (c * d - a * f);
  o[6] = inv * (d * h - e * g); o[7] = inv * (b * g - a * h); o[8] = inv * (a * e - b * d);
}

void orc_deformation_gradient(const double *rest, const double *posed, double *F) {
  double ri[9];
  orc_mat3_inv(rest, ri);
  mat3_mul(posed, ri, F);
}

/* Per-face average of a per-vertex 9-column F over the face loop
 * (averageOverFaces, DeformSolve.lean:56-69): sum from 0 in loop order,
 * times 1/n. */
static void face_average(const face_tables *T, const int32_t *target, int32_t nf,
                         const double *F, int32_t k, double *Ff) {
  for (int32_t f = 0; f < nf; ++f) {
    int32_t a = T->loop_ptr[f], n = T->loop_ptr[f + 1] - a;
    for (int32_t j = 0; j < k; ++j) Ff[(size_t)f * k + j] = 0.0;
    if (n == 0) continue;
    double invN = 1.0 / (double)n;
    for (int32_t j = 0; j < k; ++j) {
      double s = 0.0;
      for (int32_t i = 0; i < n; ++i) s = s + F[(size_t)target[T->loop_he[a + i]] * k + j];
      Ff[(size_t)f * k + j] = s * invN;
    }
  }
}

/* computeYh (DeformSolve.lean:74-96): y_h = pos[target] . row of F_f,
 * one 9-column block of Ff per pose. */
static double yh_entry(const double *pos, const double *Ff9, int32_t i) {
  return pos[0] * Ff9[3 * i] + pos[1] * Ff9[3 * i + 1] + pos[2] * Ff9[3 * i + 2];
}

/* The stage-2 right-hand side -V^T Lh (C Xc - y_h) in the dense spec form
 * (DeformSolve.lean:118-133), P poses: F is nv x 9P (per vertex; samples
 * hold their F_c row, i.e. computeFh folded in), Xc nc x 3P. */
void orc_deform_rhs_dense(int32_t nv, int32_t nf, int32_t nh, const int32_t *target,
                          const int32_t *next, const int32_t *face,
                          const int32_t *col_of_vertex, const double *pos, int32_t P,
                          const double *F, const double *Xc, double *rhs) {
  face_tables T;
  build_face_tables(nf, nh, target, next, face, pos, &T);
  const int32_t k9 = 9 * P, k3 = 3 * P;
  double *Ff = (double *)calloc((size_t)nf * k9 + 1, sizeof(double));
  face_average(&T, target, nf, F, k9, Ff);
  double *L = (double *)calloc((size_t)nh * nh, sizeof(double));
  for (int32_t f = 0; f < nf; ++f) {
    int32_t a = T.loop_ptr[f], n = T.loop_ptr[f + 1] - a;
    if (n < 3) continue;
    const double *Lf = T.mat + T.mat_ptr[f];
    for (int32_t li = 0; li < n; ++li)
      for (int32_t lj = 0; lj < n; ++lj) {
        int32_t gi = T.loop_he[a + li], gj = T.loop_he[a + lj];
        L[(size_t)gi * nh + gj] = L[(size_t)gi * nh + gj] + Lf[li * n + lj];
      }
  }
  double *diff = (double *)calloc((size_t)nh * k3 + 1, sizeof(double));
  for (int32_t h = 0; h < nh; ++h) {
    int32_t c = col_of_vertex[target[h]], f = face[h];
    for (int32_t j = 0; j < k3; ++j) {
      double cx = c >= 0 ? Xc[(size_t)c * k3 + j] : 0.0;
      double y = f >= 0 ? yh_entry(pos + 3 * target[h], Ff + (size_t)f * k9 + 9 * (j / 3), j % 3)
                        : 0.0;
      diff[(size_t)h * k3 + j] = cx - y;
    }
  }
  double *LD = (double *)calloc((size_t)nh * k3 + 1, sizeof(double));
  for (int32_t i = 0; i < nh; ++i)
    for (int32_t j = 0; j < k3; ++j) {
      double t = 0.0;
      for (int32_t p = 0; p < nh; ++p) t = t + L[(size_t)i * nh + p] * diff[(size_t)p * k3 + j];
      LD[(size_t)i * k3 + j] = t;
    }
  for (int32_t i = 0; i < nv; ++i)
    for (int32_t j = 0; j < k3; ++j) {
      double t = 0.0;
      for (int32_t p = 0; p < nh; ++p) {
        double vt = (col_of_vertex[target[p]] < 0 && target[p] == i) ? 1.0 : 0.0;
        t = t + vt * LD[(size_t)p * k3 + j];
      }
      rhs[(size_t)i * k3 + j] = -t;
    }
  free(Ff); free(L); free(diff); free(LD);
  free_face_tables(&T);
}

/* The same right-hand side in the sparse assembly order of
 * orc_harmonic_rhs_csr (incoming halfedges of a free vertex ascending, the
 * face's halfedges ascending), for the free rows; F split into its free
 * rows (Ffree, nfree x 9P) and constrained rows (Fc, nc x 9P). */
void orc_deform_rhs_csr(int32_t nv, int32_t nh, const int32_t *target, const int32_t *next,
                        const int32_t *face, const int32_t *free_of, const int32_t *con_of,
                        const double *pos, int32_t P, const double *Ffree, const double *Fc,
                        const double *Xc, double *rhs) {
  int32_t nf = 0;
  for (int32_t h = 0; h < nh; ++h) if (face[h] + 1 > nf) nf = face[h] + 1;
  face_tables T;
  build_face_tables(nf, nh, target, next, face, pos, &T);
  const int32_t k9 = 9 * P, k3 = 3 * P;
  double *F = (double *)calloc((size_t)nv * k9 + 1, sizeof(double));
  for (int32_t v = 0; v < nv; ++v) {
    const double *src = free_of[v] >= 0 ? Ffree + (size_t)free_of[v] * k9
                                        : (con_of[v] >= 0 ? Fc + (size_t)con_of[v] * k9 : NULL);
    if (src) memcpy(F + (size_t)v * k9, src, sizeof(double) * (size_t)k9);
  }
  double *Ff = (double *)calloc((size_t)nf * k9 + 1, sizeof(double));
  face_average(&T, target, nf, F, k9, Ff);
  int32_t *iptr, *ihe;
  incoming_lists(nv, nh, target, &iptr, &ihe);
  int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  for (int32_t v = 0; v < nv; ++v) {
    int32_t i = free_of[v];
    if (i < 0) continue;
    for (int32_t j = 0; j < k3; ++j) {
      double s = 0.0;
      for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
        int32_t p = ihe[a], f = face[p];
        if (f < 0) continue;
        int32_t b = T.loop_ptr[f], n = T.loop_ptr[f + 1] - b;
        memcpy(sorted, T.loop_he + b, sizeof(int32_t) * (size_t)n);
        qsort(sorted, (size_t)n, sizeof(int32_t), cmp_i32);
        double tsum = 0.0;
        for (int32_t e = 0; e < n; ++e) {
          int32_t q = sorted[e], cc = con_of[target[q]];
          double cx = cc >= 0 ? Xc[(size_t)cc * k3 + j] : 0.0;
          double y = yh_entry(pos + 3 * target[q], Ff + (size_t)f * k9 + 9 * (j / 3), j % 3);
          tsum = tsum + lh_entry(&T, face, p, q) * (cx - y);
        }
        s = s + tsum;
      }
      rhs[(size_t)i * k3 + j] = -s;
    }
  }
  free(sorted); free(iptr); free(ihe); free(F); free(Ff);
  free_face_tables(&T);
}

/* ======================================================================
 * Sparse linear algebra (SparseLinAlg.lean)
 * ==================================================================== */

void orc_spmv(int32_t rows, const int32_t *ptr, const int32_t *col,
              const double *val, const double *x, double *y) {
  for (int32_t i = 0; i < rows; ++i) {
    double s = 0.0;
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) s = s + val[p] * x[col[p]];
    y[i] = s;
  }
}

void orc_spmm(int32_t rows, const int32_t *ptr, const int32_t *col,
              const double *val, int32_t k, const double *X, double *Y) {
  for (int32_t i = 0; i < rows; ++i)
    for (int32_t c = 0; c < k; ++c) {
      double s = 0.0;
      for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p)
        s = s + val[p] * X[(size_t)col[p] * k + c];
      Y[(size_t)i * k + c] = s;
    }
}

double orc_dot(int32_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) s = s + a[i] * b[i];
  return s;
}

void orc_diagonal(int32_t rows, const int32_t *ptr, const int32_t *col,
                  const double *val, double *d) {
  for (int32_t i = 0; i < rows; ++i) {
    d[i] = 0.0;
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p)
      if (col[p] == i) { d[i] = val[p]; break; }
  }
}

void orc_apply_jacobi(int32_t n, const double *d, const double *b, double *y) {
  for (int32_t i = 0; i < n; ++i) y[i] = (d[i] == 0.0) ? 0.0 : b[i] / d[i];
}

/* transpose of the strictly-lower part of L (for the backward sweep) */
typedef struct { int32_t *ptr, *row; double *val; } csc_t;

static void lower_transpose(int32_t n, const int32_t *lp, const int32_t *lc,
                            const double *lv, csc_t *T) {
  T->ptr = (int32_t *)calloc((size_t)n + 2, sizeof(int32_t));
  int32_t nnz = lp[n];
  for (int32_t p = 0; p < nnz; ++p) T->ptr[lc[p] + 1]++;
  for (int32_t i = 0; i < n; ++i) T->ptr[i + 1] += T->ptr[i];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  memcpy(fill, T->ptr, sizeof(int32_t) * (size_t)n);
  T->row = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
  T->val = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
  for (int32_t i = 0; i < n; ++i)          /* rows ascending => kk ascending */
    for (int32_t p = lp[i]; p < lp[i + 1]; ++p) {
      int32_t c = lc[p];
      T->row[fill[c]] = i;
      T->val[fill[c]] = lv[p];
      fill[c]++;
    }
  free(fill);
}

/* IncompleteCholesky.forwardSub3 / backwardSub3 (:83-101) on CSR. */
static void ic0_solve_t(int32_t n, const int32_t *lp, const int32_t *lc,
                        const double *lv, const csc_t *T, const double *b,
                        double *x, double *y) {
  for (int32_t i = 0; i < n; ++i) {
    double s = b[i];
    int32_t e = lp[i + 1] - 1;                   /* diagonal is last */
    for (int32_t p = lp[i]; p < e; ++p) s = s - lv[p] * y[lc[p]];
    y[i] = s / lv[e];
  }
  for (int32_t jj = 0; jj < n; ++jj) {
    int32_t i = n - 1 - jj;
    double s = y[i];
    int32_t d = lp[i + 1] - 1;
    for (int32_t p = T->ptr[i]; p < T->ptr[i + 1]; ++p) {
      int32_t kk = T->row[p];
      if (kk > i) s = s - T->val[p] * x[kk];
    }
    x[i] = s / lv[d];
  }
}

void orc_ic0_solve(int32_t n, const int32_t *l_ptr, const int32_t *l_col,
                   const double *l_val, const double *b, double *x) {
  csc_t T;
  lower_transpose(n, l_ptr, l_col, l_val, &T);
  double *y = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  ic0_solve_t(n, l_ptr, l_col, l_val, &T, b, x, y);
  free(y); free(T.ptr); free(T.row); free(T.val);
}

/* SparseLinAlg.cg / cgWithGuess (SparseLinAlg.lean:104-163) with the
 * preconditioner as a parameter. */
static int32_t pcg_core(int32_t n, const int32_t *ptr, const int32_t *col,
                        const double *val, int32_t precond, const int32_t *lp,
                        const int32_t *lc, const double *lv, const double *b,
                        const double *x0, int32_t max_iter, double tol,
                        double *x) {
  double *d = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *r = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *z = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *p = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *Ap = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *y = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  csc_t T = {0, 0, 0};
  if (precond == 1) lower_transpose(n, lp, lc, lv, &T);
  orc_diagonal(n, ptr, col, val, d);
  if (x0) {
    memcpy(x, x0, sizeof(double) * (size_t)n);
    orc_spmv(n, ptr, col, val, x0, Ap);
    for (int32_t i = 0; i < n; ++i) r[i] = 1.0 * b[i] + -1.0 * Ap[i]; /* saxpby */
  } else {
    for (int32_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; }
  }
  if (precond == 1) ic0_solve_t(n, lp, lc, lv, &T, r, z, y);
  else orc_apply_jacobi(n, d, r, z);
  memcpy(p, z, sizeof(double) * (size_t)n);
  double rz = orc_dot(n, r, z);
  double tol2 = tol * tol;
  int32_t it = 0;
  for (; it < max_iter; ) {
    orc_spmv(n, ptr, col, val, p, Ap);
    double pAp = orc_dot(n, p, Ap);
    if (pAp == 0.0) break;
    double alpha = rz / pAp;
    for (int32_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
    for (int32_t i = 0; i < n; ++i) r[i] = r[i] + (-alpha) * Ap[i];
    ++it;
    double rr = orc_dot(n, r, r);
    if (rr < tol2) break;
    if (precond == 1) ic0_solve_t(n, lp, lc, lv, &T, r, z, y);
    else orc_apply_jacobi(n, d, r, z);
    double rzn = orc_dot(n, r, z);
    double beta = (rz == 0.0) ? 0.0 : rzn / rz;
    for (int32_t i = 0; i < n; ++i) p[i] = 1.0 * z[i] + beta * p[i];
    rz = rzn;
  }
  free(d); free(r); free(z); free(p); free(Ap); free(y);
  if (precond == 1) { free(T.ptr); free(T.row); free(T.val); }
  return it;
}

int32_t orc_cg(int32_t n, const int32_t *ptr, const int32_t *col,
               const double *val, const double *b, const double *x0,
               int32_t max_iter, double tol, double *x) {
  return pcg_core(n, ptr, col, val, 0, 0, 0, 0, b, x0, max_iter, tol, x);
}

int32_t orc_pcg(int32_t n, const int32_t *ptr, const int32_t *col,
                const double *val, int32_t precond, const int32_t *l_ptr,
                const int32_t *l_col, const double *l_val, const double *b,
                const double *x0, int32_t max_iter, double tol_abs, double *x) {
  return pcg_core(n, ptr, col, val, precond, l_ptr, l_col, l_val, b, x0,
                  max_iter, tol_abs, x);
}

/* Many independent right-hand sides (columns of a row-major n x k block),
 * each solved by orc_pcg with tolerance tol_rel * ||b_c||, spread over
 * `nthreads` POSIX threads (column c on thread c mod nthreads).  Every
 * column's arithmetic is exactly the single-column solve's.  Returns the
 * total iteration count. */
typedef struct {
  int32_t n, k, c0, step, precond, max_iter;
  const int32_t *ptr, *col, *l_ptr, *l_col;
  const double *val, *l_val, *B;
  double tol_rel, *X;
  int64_t iters;
} pcg_job;

static void *pcg_worker(void *arg) {
  pcg_job *J = (pcg_job *)arg;
  double *b = (double *)malloc(sizeof(double) * (size_t)J->n + 8);
  double *x = (double *)malloc(sizeof(double) * (size_t)J->n + 8);
  for (int32_t c = J->c0; c < J->k; c += J->step) {
    double bb = 0.0;
    for (int32_t i = 0; i < J->n; ++i) {
      b[i] = J->B[(size_t)i * J->k + c];
      bb = bb + b[i] * b[i];
    }
    J->iters += pcg_core(J->n, J->ptr, J->col, J->val, J->precond, J->l_ptr, J->l_col,
                         J->l_val, b, NULL, J->max_iter, J->tol_rel * sqrt(bb), x);
    for (int32_t i = 0; i < J->n; ++i) J->X[(size_t)i * J->k + c] = x[i];
  }
  free(b);
  free(x);
  return NULL;
}

int64_t orc_pcg_many(int32_t n, const int32_t *ptr, const int32_t *col, const double *val,
                     int32_t precond, const int32_t *l_ptr, const int32_t *l_col,
                     const double *l_val, int32_t k, const double *B, int32_t max_iter,
                     double tol_rel, double *X, int32_t nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > k) nthreads = k;
  pcg_job *jobs = (pcg_job *)calloc((size_t)nthreads, sizeof(pcg_job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int32_t t = 0; t < nthreads; ++t) {
    pcg_job J = {n, k, t, nthreads, precond, max_iter, ptr, col, l_ptr, l_col,
                 val, l_val, B, tol_rel, X, 0};
    jobs[t] = J;
  }
  for (int32_t t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, pcg_worker, &jobs[t]);
  pcg_worker(&jobs[0]);
  int64_t it = jobs[0].iters;
  for (int32_t t = 1; t < nthreads; ++t) {
    pthread_join(th[t], NULL);
    it += jobs[t].iters;
  }
  free(jobs);
  free(th);
  return it;
}

/* ======================================================================
 * IC(0) (IncompleteCholesky.lean)
 * ==================================================================== */

/* factor3 (IncompleteCholesky.lean:49-69). */
void orc_ic0_factor3(const double *A, const double *pattern, double *L) {
  for (int32_t i = 0; i < 9; ++i) L[i] = 0.0;
  for (int32_t j = 0; j < 3; ++j) {
    double s = A[3 * j + j];
    for (int32_t k = 0; k < j; ++k) {
      double ljk = L[3 * j + k];
      s = s - ljk * ljk;
    }
    L[3 * j + j] = sqrt(s);
    for (int32_t i = j + 1; i < 3; ++i) {
      int keep = pattern[3 * i + j] != 0.0 || i == j;
      if (keep) {
        double t = A[3 * i + j];
        for (int32_t k = 0; k < j; ++k) t = t - L[3 * i + k] * L[3 * j + k];
        L[3 * i + j] = t / L[3 * j + j];
      }
    }
  }
}

/* CSR no-fill IC(0): row-by-row evaluation of the same recurrences as
 * factor3 (column j's values only read rows < j, so rows can be finished in
 * order).  Each product sum runs over k ascending, as in the spec. */
int32_t orc_ic0_factor(int32_t n, const int32_t *ptr, const int32_t *col,
                       const double *val, int32_t *l_ptr, int32_t *l_col,
                       double *l_val) {
  int32_t nz = 0;
  l_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p)
      if (col[p] <= i) { l_col[nz] = col[p]; l_val[nz] = val[p]; ++nz; }
    l_ptr[i + 1] = nz;
    if (nz == l_ptr[i] || l_col[nz - 1] != i) return -(i + 1); /* no diag */
  }
  for (int32_t i = 0; i < n; ++i) {
    int32_t a0 = l_ptr[i], a1 = l_ptr[i + 1] - 1;   /* a1 = diagonal slot */
    for (int32_t a = a0; a < a1; ++a) {
      int32_t j = l_col[a];
      double t = l_val[a];
      int32_t b = l_ptr[j], b1 = l_ptr[j + 1] - 1;
      int32_t c = a0;
      while (c < a && b < b1) {                       /* k < j, ascending */
        if (l_col[c] == l_col[b]) { t = t - l_val[c] * l_val[b]; ++c; ++b; }
        else if (l_col[c] < l_col[b]) ++c;
        else ++b;
      }
      l_val[a] = t / l_val[b1];
    }
    double s = l_val[a1];
    for (int32_t a = a0; a < a1; ++a) s = s - l_val[a] * l_val[a];
    if (!(s > 0.0)) return -(i + 1);
    l_val[a1] = sqrt(s);
  }
  return 0;
}

void orc_permute_csr(int32_t n, const int32_t *ptr, const int32_t *col,
                     const double *val, const int32_t *perm, int32_t *optr,
                     int32_t *ocol, double *oval) {
  int32_t *inv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  for (int32_t i = 0; i < n; ++i) inv[perm[i]] = i;
  optr[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t o = perm[i];
    int32_t base = optr[i], m = ptr[o + 1] - ptr[o];
    for (int32_t t = 0; t < m; ++t) {
      int32_t c = inv[col[ptr[o] + t]];
      double v = val[ptr[o] + t];
      int32_t q = base + t - 1;                    /* insertion by column */
      while (q >= base && ocol[q] > c) { ocol[q + 1] = ocol[q]; oval[q + 1] = oval[q]; --q; }
      ocol[q + 1] = c;
      oval[q + 1] = v;
    }
    optr[i + 1] = base + m;
  }
  free(inv);
}

/* ======================================================================
 * Graph colouring (GraphColoring.lean:41-56)
 * ==================================================================== */
int32_t orc_greedy_color(int32_t n, const int32_t *ptr, const int32_t *col,
                         int32_t *color) {
  int32_t maxdeg = 0;
  for (int32_t i = 0; i < n; ++i)
    if (ptr[i + 1] - ptr[i] > maxdeg) maxdeg = ptr[i + 1] - ptr[i];
  uint8_t *used = (uint8_t *)calloc((size_t)maxdeg + 2, 1);
  int32_t ncol = 0;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      int32_t j = col[p];
      if (j < i && color[j] <= maxdeg) used[color[j]] = 1;
    }
    int32_t c = 0;
    while (c <= maxdeg && used[c]) ++c;
    color[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      int32_t j = col[p];
      if (j < i && color[j] <= maxdeg) used[color[j]] = 0;
    }
  }
  free(used);
  return ncol;
}

/* ======================================================================
 * Colour-parallel symmetric Gauss-Seidel (SlangCodegen/SgsColor.lean:87-130)
 * and the coarse-vertex restriction (SlangCodegen/HierarchicalSparsify.lean:
 * 64-93)
 * ==================================================================== */

/* One SgsColor row update (SgsColor.lean:104-125), column c of an n x k
 * row-major block: d = A_ii; a row with d == 0 is left alone; s_full =
 * sum over the row's entries in storage order, from 0; x_i = (b_i - (s_full
 * - d x_i)) / d. */
static void sgs_row_f64(int32_t i, const int32_t *ptr, const int32_t *col, const double *val,
                        const double *diag, int32_t k, int32_t c, const double *b, double *x) {
  double d = diag[i];
  if (d == 0.0) return;
  double s = 0.0;
  for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) s = s + val[q] * x[(size_t)col[q] * k + c];
  double off = s - d * x[(size_t)i * k + c];
  x[(size_t)i * k + c] = (b[(size_t)i * k + c] - off) / d;
}

static void sgs_row_f32(int32_t i, const int32_t *ptr, const int32_t *col, const float *val,
                        const float *diag, int32_t k, int32_t c, const float *b, float *x) {
  float d = diag[i];
  if (d == 0.0f) return;
  float s = 0.0f;
  for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) s = s + val[q] * x[(size_t)col[q] * k + c];
  float off = s - d * x[(size_t)i * k + c];
  x[(size_t)i * k + c] = (b[(size_t)i * k + c] - off) / d;
}

/* `sweeps` symmetric sweeps: colours 0..C-1 (forward), then C-1..0
 * (backward), the host schedule of SgsColor.lean:12-14.  Rows of one colour
 * are visited in ascending order; the colouring makes them independent. */
void orc_sgs_sweep_f64(int32_t n, const int32_t *ptr, const int32_t *col, const double *val,
                       const double *diag, int32_t ncolors, const int32_t *color_ptr,
                       const int32_t *color_rows, int32_t k, const double *b, double *x,
                       int32_t sweeps) {
  (void)n;
  for (int32_t sw = 0; sw < sweeps; ++sw)
    for (int32_t pass = 0; pass < 2 * ncolors; ++pass) {
      int32_t c = pass < ncolors ? pass : 2 * ncolors - 1 - pass;
      for (int32_t t = color_ptr[c]; t < color_ptr[c + 1]; ++t)
        for (int32_t j = 0; j < k; ++j)
          sgs_row_f64(color_rows[t], ptr, col, val, diag, k, j, b, x);
    }
}

void orc_sgs_sweep_f32(int32_t n, const int32_t *ptr, const int32_t *col, const float *val,
                       const float *diag, int32_t ncolors, const int32_t *color_ptr,
                       const int32_t *color_rows, int32_t k, const float *b, float *x,
                       int32_t sweeps) {
  (void)n;
  for (int32_t sw = 0; sw < sweeps; ++sw)
    for (int32_t pass = 0; pass < 2 * ncolors; ++pass) {
      int32_t c = pass < ncolors ? pass : 2 * ncolors - 1 - pass;
      for (int32_t t = color_ptr[c]; t < color_ptr[c + 1]; ++t)
        for (int32_t j = 0; j < k; ++j)
          sgs_row_f32(color_rows[t], ptr, col, val, diag, k, j, b, x);
    }
}

/* coarse[c] = sum over fine i with parent[i] = c of fine[i], scanning i in
 * ascending order from 0 (HierarchicalSparsify.lean:70-84), for k columns. */
void orc_restrict_sum_f64(int32_t n_fine, const int32_t *parent, int32_t n_coarse, int32_t k,
                          const double *fine, double *coarse) {
  for (int32_t c = 0; c < n_coarse; ++c)
    for (int32_t j = 0; j < k; ++j) {
      double acc = 0.0;
      for (int32_t i = 0; i < n_fine; ++i)
        if (parent[i] == c) acc = acc + fine[(size_t)i * k + j];
      coarse[(size_t)c * k + j] = acc;
    }
}

void orc_restrict_sum_f32(int32_t n_fine, const int32_t *parent, int32_t n_coarse, int32_t k,
                          const float *fine, float *coarse) {
  for (int32_t c = 0; c < n_coarse; ++c)
    for (int32_t j = 0; j < k; ++j) {
      float acc = 0.0f;
      for (int32_t i = 0; i < n_fine; ++i)
        if (parent[i] == c) acc = acc + fine[(size_t)i * k + j];
      coarse[(size_t)c * k + j] = acc;
    }
}

/* ======================================================================
 * Dense linear algebra (DenseLinAlg.lean)
 * ==================================================================== */

/* solve (DenseLinAlg.lean:70-111). */
void orc_dense_solve(int32_t n, const double *A, const double *b, double *x) {
  int32_t w = n + 1;
  double *g = (double *)malloc(sizeof(double) * (size_t)n * w + 8);
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t j = 0; j < n; ++j) g[(size_t)i * w + j] = A[(size_t)i * n + j];
    g[(size_t)i * w + n] = b[i];
  }
  for (int32_t k = 0; k < n; ++k) {
    int32_t piv = k;
    double best = fabs(g[(size_t)k * w + k]);
    for (int32_t r = k + 1; r < n; ++r) {
      double v = fabs(g[(size_t)r * w + k]);
      if (v > best) { best = v; piv = r; }
    }
    if (piv != k)
      for (int32_t j = 0; j < w; ++j) {
        double t = g[(size_t)k * w + j];
        g[(size_t)k * w + j] = g[(size_t)piv * w + j];
        g[(size_t)piv * w + j] = t;
      }
    double pivot = g[(size_t)k * w + k];
    if (pivot == 0.0) continue;
    for (int32_t r = k + 1; r < n; ++r) {
      double factor = g[(size_t)r * w + k] / pivot;
      for (int32_t j = k; j < w; ++j)
        g[(size_t)r * w + j] = g[(size_t)r * w + j] - factor * g[(size_t)k * w + j];
    }
  }
  for (int32_t i = 0; i < n; ++i) {
    int32_t row = n - 1 - i;
    double s = g[(size_t)row * w + n];
    for (int32_t j = row + 1; j < n; ++j) s = s - g[(size_t)row * w + j] * x[j];
    double pivot = g[(size_t)row * w + row];
    x[row] = (pivot == 0.0) ? 0.0 : s / pivot;
  }
  free(g);
}

/* solveMulti (DenseLinAlg.lean:122-131): one elimination per column. */
void orc_dense_solve_multi(int32_t n, int32_t k, const double *A,
                           const double *B, double *X) {
  double *b = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *x = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  for (int32_t c = 0; c < k; ++c) {
    for (int32_t i = 0; i < n; ++i) b[i] = B[(size_t)i * k + c];
    orc_dense_solve(n, A, b, x);
    for (int32_t i = 0; i < n; ++i) X[(size_t)i * k + c] = x[i];
  }
  free(b); free(x);
}

/* factorizeLU (DenseLinAlg.lean:139-166). */
void orc_lu_factor(int32_t n, const double *A, double *a, int32_t *piv) {
  memcpy(a, A, sizeof(double) * (size_t)n * n);
  for (int32_t i = 0; i < n; ++i) piv[i] = i;
  for (int32_t k = 0; k < n; ++k) {
    int32_t pr = k;
    double best = fabs(a[(size_t)k * n + k]);
    for (int32_t r = k + 1; r < n; ++r) {
      double v = fabs(a[(size_t)r * n + k]);
      if (v > best) { best = v; pr = r; }
    }
    if (pr != k) {
      for (int32_t j = 0; j < n; ++j) {
        double t = a[(size_t)k * n + j];
        a[(size_t)k * n + j] = a[(size_t)pr * n + j];
        a[(size_t)pr * n + j] = t;
      }
      int32_t tp = piv[k]; piv[k] = piv[pr]; piv[pr] = tp;
    }
    double pivot = a[(size_t)k * n + k];
    if (pivot == 0.0) continue;
    for (int32_t r = k + 1; r < n; ++r) {
      double factor = a[(size_t)r * n + k] / pivot;
      a[(size_t)r * n + k] = factor;
      for (int32_t j = k + 1; j < n; ++j)
        a[(size_t)r * n + j] = a[(size_t)r * n + j] - factor * a[(size_t)k * n + j];
    }
  }
}

/* solveWithLU (DenseLinAlg.lean:169-190). */
void orc_lu_solve(int32_t n, const double *a, const int32_t *piv,
                  const double *b, double *x) {
  double *y = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  for (int32_t i = 0; i < n; ++i) y[i] = b[piv[i]];
  for (int32_t i = 0; i < n; ++i) {
    double s = y[i];
    for (int32_t j = 0; j < i; ++j) s = s - a[(size_t)i * n + j] * y[j];
    y[i] = s;
  }
  for (int32_t i = 0; i < n; ++i) {
    int32_t row = n - 1 - i;
    double s = y[row];
    for (int32_t j = row + 1; j < n; ++j) s = s - a[(size_t)row * n + j] * x[j];
    double pivot = a[(size_t)row * n + row];
    x[row] = (pivot == 0.0) ? 0.0 : s / pivot;
  }
  free(y);
}

/* choleskyFactor (DenseLinAlg.lean:226-241). */
void orc_cholesky_factor(int32_t n, const double *A, double *L) {
  for (int64_t i = 0; i < (int64_t)n * n; ++i) L[i] = 0.0;
  for (int32_t j = 0; j < n; ++j) {
    double s = 0.0;
    for (int32_t k = 0; k < j; ++k) {
      double v = L[(size_t)j * n + k];
      s = s + v * v;
    }
    double diag = sqrt(A[(size_t)j * n + j] - s);
    L[(size_t)j * n + j] = diag;
    for (int32_t i = j + 1; i < n; ++i) {
      double t = 0.0;
      for (int32_t k = 0; k < j; ++k) t = t + L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = (A[(size_t)i * n + j] - t) / diag;
    }
  }
}

/* forwardSolve / backwardSolve / solveWithCholesky (DenseLinAlg.lean:244-280). */
void orc_cholesky_solve(int32_t n, const double *L, const double *b, double *x) {
  double *y = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t j = 0; j < i; ++j) s = s + L[(size_t)i * n + j] * y[j];
    y[i] = (b[i] - s) / L[(size_t)i * n + i];
  }
  for (int32_t ii = 0; ii < n; ++ii) {
    int32_t i = n - 1 - ii;
    double s = 0.0;
    for (int32_t j = i + 1; j < n; ++j) s = s + L[(size_t)j * n + i] * x[j];
    x[i] = (y[i] - s) / L[(size_t)i * n + i];
  }
  free(y);
}

/* ======================================================================
 * HSC (HierarchicalSparsifyCompensate.lean:54-125)
 * ==================================================================== */

double orc_hsc_degree_weight(int32_t ne, const int32_t *eu, const int32_t *ev,
                             const double *ew, int32_t v) {
  double s = 0.0;
  for (int32_t e = 0; e < ne; ++e)
    if (eu[e] == v || ev[e] == v) s = s + ew[e];
  return s;
}

double orc_hsc_edge_weight(int32_t ne, const int32_t *eu, const int32_t *ev,
                           const double *ew, int32_t a, int32_t b) {
  int32_t lo = a < b ? a : b, hi = a < b ? b : a;
  for (int32_t e = 0; e < ne; ++e)
    if (eu[e] == lo && ev[e] == hi) return ew[e];
  return 0.0;
}

int32_t orc_hsc_eliminate_vertex(int32_t ne, const int32_t *eu,
                                 const int32_t *ev, const double *ew,
                                 int32_t v, int32_t *ou, int32_t *ov,
                                 double *ow) {
  /* neighborsOf (:80-88) */
  int32_t nn = 0;
  int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne + 1));
  double *nw = (double *)malloc(sizeof(double) * (size_t)(ne + 1));
  for (int32_t e = 0; e < ne; ++e) {
    if (eu[e] == v) { nb[nn] = ev[e]; nw[nn] = ew[e]; ++nn; }
    else if (ev[e] == v) { nb[nn] = eu[e]; nw[nn] = ew[e]; ++nn; }
  }
  double dv = orc_hsc_degree_weight(ne, eu, ev, ew, v);
  int32_t m = 0;
  for (int32_t e = 0; e < ne; ++e)
    if (eu[e] != v && ev[e] != v) { ou[m] = eu[e]; ov[m] = ev[e]; ow[m] = ew[e]; ++m; }
  if (dv > 0.0) {
    for (int32_t i = 0; i < nn; ++i)
      for (int32_t j = i + 1; j < nn; ++j) {
        int32_t ni = nb[i], nj = nb[j];
        int32_t lo = ni < nj ? ni : nj, hi = ni < nj ? nj : ni;
        double comp = nw[i] * nw[j] / dv;
        int32_t found = 0;
        for (int32_t k = 0; k < m; ++k)
          if (ou[k] == lo && ov[k] == hi) { ow[k] = ow[k] + comp; found = 1; break; }
        if (!found) { ou[m] = lo; ov[m] = hi; ow[m] = comp; ++m; }
      }
  }
  free(nb); free(nw);
  return m;
}

/* ======================================================================
 * Direct Delta Mush: runtime matvec and bind-time weight post-processing
 * ==================================================================== */

/* DirectDeltaMush.applyTransform4x4 (DirectDeltaMush.lean:69-73): rows 0-2
 * of T (x, y, z, 1), each summed left to right. */
void orc_apply_transform4x4(const double *T, double x, double y, double z, double *out) {
  for (int r = 0; r < 3; ++r) out[r] = T[4 * r] * x + T[4 * r + 1] * y + T[4 * r + 2] * z + T[4 * r + 3];
}

/* DirectDeltaMush.lbsMatvec (DirectDeltaMush.lean:85-98), one vertex per CSR
 * row of influences, accumulated from 0 in list order.  A handle index past
 * the transform array reads the default (empty) matrix, whose entries are 0. */
void orc_lbs_matvec(int32_t nv, const double *rest, const double *T, int32_t nt,
                    const int32_t *start, const int32_t *idx, const double *w, double *out) {
  static const double zero16[16] = {0};
  for (int32_t v = 0; v < nv; ++v) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int32_t q = start[v]; q < start[v + 1]; ++q) {
      const double *M = (idx[q] >= 0 && idx[q] < nt) ? T + 16 * (int64_t)idx[q] : zero16;
      double t[3];
      orc_apply_transform4x4(M, rest[3 * v], rest[3 * v + 1], rest[3 * v + 2], t);
      for (int d = 0; d < 3; ++d) acc[d] = acc[d] + w[q] * t[d];
    }
    for (int d = 0; d < 3; ++d) out[3 * v + d] = acc[d];
  }
}

/* DirectDeltaMushBind.smoothWeights (DirectDeltaMushBind.lean:76-97): nu
 * passes of w <- (1 - omega) w + omega * (sum of neighbour weights / count),
 * the neighbour sum from 0 in adjacency order (0 for a vertex without
 * neighbours; a neighbour index outside [0, n) reads 0). */
void orc_smooth_weights(int32_t n, int32_t nh, const int32_t *adj_ptr, const int32_t *adj,
                        const double *W, int32_t nu, double omega, double *out) {
  const size_t sz = (size_t)n * (size_t)nh;
  double *cur = (double *)malloc(sizeof(double) * (sz ? sz : 1));
  double *nxt = (double *)malloc(sizeof(double) * (sz ? sz : 1));
  memcpy(cur, W, sizeof(double) * sz);
  for (int32_t pass = 0; pass < nu; ++pass) {
    for (int32_t v = 0; v < n; ++v) {
      const int32_t a = adj_ptr[v], b = adj_ptr[v + 1], count = b - a;
      for (int32_t i = 0; i < nh; ++i) {
        double nb = 0.0;
        if (count > 0) {
          for (int32_t q = a; q < b; ++q) {
            const int32_t u = adj[q];
            nb = nb + ((u >= 0 && u < n) ? cur[(size_t)u * nh + i] : 0.0);
          }
          nb = nb / (double)count;
        }
        nxt[(size_t)v * nh + i] = (1.0 - omega) * cur[(size_t)v * nh + i] + omega * nb;
      }
    }
    double *t = cur;
    cur = nxt;
    nxt = t;
  }
  memcpy(out, cur, sizeof(double) * sz);
  free(cur);
  free(nxt);
}

/* DirectDeltaMushBind.sparsifyTopK (DirectDeltaMushBind.lean:107-136):
 * selection sort of the row's (index, weight) pairs for the first
 * take = min(K, nh) positions -- the first position of the largest |w|
 * (strict >) is swapped into place -- then the kept weights are divided by
 * their sum (accumulated from 0 in kept order) when that sum is not 0. */
void orc_sparsify_topk(int32_t n, int32_t nh, const double *W, int32_t K, int32_t *idx,
                       double *w) {
  const int32_t take = K < nh ? K : nh;
  int32_t *pi = (int32_t *)malloc(sizeof(int32_t) * (nh ? nh : 1));
  double *pw = (double *)malloc(sizeof(double) * (nh ? nh : 1));
  for (int32_t v = 0; v < n; ++v) {
    for (int32_t i = 0; i < nh; ++i) {
      pi[i] = i;
      pw[i] = W[(size_t)v * nh + i];
    }
    for (int32_t k = 0; k < take; ++k) {
      int32_t best = k;
      for (int32_t j = k + 1; j < nh; ++j)
        if (fabs(pw[j]) > fabs(pw[best])) best = j;
      const int32_t ti = pi[k];
      const double tw = pw[k];
      pi[k] = pi[best];
      pw[k] = pw[best];
      pi[best] = ti;
      pw[best] = tw;
    }
    double sum = 0.0;
    for (int32_t k = 0; k < take; ++k) sum = sum + pw[k];
    for (int32_t k = 0; k < take; ++k) {
      idx[(size_t)v * take + k] = pi[k];
      w[(size_t)v * take + k] = sum != 0.0 ? pw[k] / sum : pw[k];
    }
  }
  free(pi);
  free(pw);
}
