// cn_deform.hip -- the deformation solve's bridge and its curve inputs on
// gfx950.
//
// DeformSolve.solveDeformation (DeformSolve.lean:110-135) runs two harmonic
// solves on the same A_ff: stage 1 for the deformation gradients F (9
// columns a pose), stage 2 for the positions (3 columns a pose) with the
// right-hand side -V^T Lh (C X_c - y_h).  The bridge between them:
//   f_h = V f_v + C f_c        (computeFh, :38-52): F of the halfedge target
//   F_f = mean of f_h over the face loop (averageOverFaces, :56-69)
//   y_h = X_target . F_f^T     (computeYh, :74-96)
// Here the bridge is folded into the stage-2 right-hand side kernel, which
// sums in the assembly's order (incoming interior halfedges of a free
// vertex ascending, the face's three halfedges ascending), so it is
// bit-identical to the oracle's sparse restatement and, through it, to the
// dense spec.  The file is compiled without floating-point contraction.
//
// Curve inputs: EndToEndExample.bezier3 (EndToEndExample.lean:52-60),
// ScaledFrames.isolatedSegmentGradient (ScaledFrames.lean:90-95) and
// ScaledFrames.deformationGradient (ScaledFrames.lean:147-150), one thread
// per evaluation, in the spec's operation order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_internal.h"

namespace cn {

__device__ __forceinline__ int32_t dtgt3(const int32_t *tris, int32_t h) {
  const int32_t f = h / 3, l = h - 3 * f;
  return tris[3 * f + (l == 2 ? 0 : l + 1)];
}

// F of vertex v: its free row, else its constrained row
__device__ __forceinline__ const double *vertex_F(int32_t v, const int32_t *free_of,
                                                  const int32_t *con_of, const double *Ffree,
                                                  const double *Fc, int64_t k9) {
  const int32_t fr = free_of[v];
  return fr >= 0 ? Ffree + (int64_t)fr * k9 : Fc + (int64_t)con_of[v] * k9;
}

// F_f: the face loop of face f is its halfedges 3f, 3f+1, 3f+2 (targets
// t1, t2, t0); sum from 0 in loop order, times 1/3
__global__ void k_face_F(int32_t nf, const int32_t *__restrict__ tris,
                         const int32_t *__restrict__ free_of, const int32_t *__restrict__ con_of,
                         const double *__restrict__ Ffree, const double *__restrict__ Fc,
                         int64_t k9, double *__restrict__ Ff) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)nf * k9) return;
  const int32_t f = (int32_t)(id / k9);
  const int64_t j = id - (int64_t)f * k9;
  const double invN = 1.0 / 3.0;
  double s = 0.0;
  for (int l = 0; l < 3; ++l)
    s = s + vertex_F(dtgt3(tris, 3 * f + l), free_of, con_of, Ffree, Fc, k9)[j];
  Ff[id] = s * invN;
}

// rhs[r][j] = -sum_p sum_q Lh[p][q] (CXc[q] - y_h[q])[j] for free row r
__global__ void k_deform_rhs(int32_t n_free, int64_t k3, const int32_t *__restrict__ free_vertex,
                             const int32_t *__restrict__ tris, const int32_t *__restrict__ iptr,
                             const int32_t *__restrict__ ihe, const double *__restrict__ m,
                             const int32_t *__restrict__ con_of, const double *__restrict__ pos,
                             const double *__restrict__ Ff, const double *__restrict__ Xc,
                             double *__restrict__ rhs) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n_free * k3) return;
  const int32_t r = (int32_t)(id / k3);
  const int64_t j = id - (int64_t)r * k3;
  const int64_t k9 = 3 * k3, pose = j / 3;
  const int i = (int)(j - 3 * pose);
  const int32_t v = free_vertex[r];
  double s = 0.0;
  for (int32_t a = iptr[v]; a < iptr[v + 1]; ++a) {
    const int32_t p = ihe[a], f = p / 3, lp = p - 3 * f;
    const double *mf = m + 9 * (int64_t)f + 3 * lp;
    const double *F9 = Ff + (int64_t)f * k9 + 9 * pose + 3 * i;
    double t = 0.0;
    for (int lq = 0; lq < 3; ++lq) {
      const int32_t tq = dtgt3(tris, 3 * f + lq), cc = con_of[tq];
      const double cx = cc >= 0 ? Xc[(int64_t)cc * k3 + j] : 0.0;
      const double *x = pos + 3 * (int64_t)tq;
      const double y = x[0] * F9[0] + x[1] * F9[1] + x[2] * F9[2];
      t = t + (0.0 + mf[lq]) * (cx - y);
    }
    s = s + t;
  }
  rhs[id] = -s;
}

__global__ void k_bezier3(int32_t ncurves, const double *__restrict__ ctrl, int32_t nt,
                          const double *__restrict__ tv, double *__restrict__ out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)ncurves * nt) return;
  const int32_t c = (int32_t)(id / nt);
  const double *P = ctrl + 12 * (int64_t)c;
  const double t = tv[id - (int64_t)c * nt];
  const double u = 1.0 - t;
  const double c0 = u * u * u, c1 = 3.0 * u * u * t, c2 = 3.0 * u * t * t, c3 = t * t * t;
  for (int d = 0; d < 3; ++d)
    out[3 * id + d] = c0 * P[d] + c1 * P[3 + d] + c2 * P[6 + d] + c3 * P[9 + d];
}

__device__ __forceinline__ double tangent_length(const double *p, const double *q, double *t) {
  const double d0 = q[0] - p[0], d1 = q[1] - p[1], d2 = q[2] - p[2];
  const double l = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double inv = l == 0.0 ? 0.0 : 1.0 / l;
  t[0] = d0 * inv;
  t[1] = d1 * inv;
  t[2] = d2 * inv;
  return l;
}

__device__ __forceinline__ void mat3_mul(const double *a, const double *b, double *c) {
  for (int k = 0; k < 9; ++k) {
    const int i = k / 3, j = k % 3;
    c[k] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
  }
}

__global__ void k_segment_gradients(int32_t n, const double *__restrict__ rp,
                                    const double *__restrict__ rq,
                                    const double *__restrict__ pp,
                                    const double *__restrict__ pq, double *__restrict__ F) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double t0[3], t1[3];
  const double l0 = tangent_length(rp + 3 * (int64_t)s, rq + 3 * (int64_t)s, t0);
  const double l1 = tangent_length(pp + 3 * (int64_t)s, pq + 3 * (int64_t)s, t1);
  const double r = l0 == 0.0 ? 0.0 : l1 / l0;
  // smallestRotation: R = I + K + K^2 / (1 + cos)
  const double cx = t0[1] * t1[2] - t0[2] * t1[1];
  const double cy = t0[2] * t1[0] - t0[0] * t1[2];
  const double cz = t0[0] * t1[1] - t0[1] * t1[0];
  const double cosT = t0[0] * t1[0] + t0[1] * t1[1] + t0[2] * t1[2];
  const double K[9] = {0.0, -cz, cy, cz, 0.0, -cx, -cy, cx, 0.0};
  double K2[9];
  mat3_mul(K, K, K2);
  const double denom = 1.0 + cosT;
  const double inv = denom == 0.0 ? 0.0 : 1.0 / denom;
  for (int k = 0; k < 9; ++k) {
    const double I = (k % 4 == 0) ? 1.0 : 0.0;
    F[9 * (int64_t)s + k] = r * ((I + K[k]) + inv * K2[k]);
  }
}

__global__ void k_deformation_gradients(int32_t n, const double *__restrict__ rest,
                                        const double *__restrict__ posed,
                                        double *__restrict__ F) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double *m = rest + 9 * (int64_t)s;
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6],
               h = m[7], i = m[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  const double inv = det == 0.0 ? 0.0 : 1.0 / det;
  const double ri[9] = {inv * (e * i - f * h), inv * (c * h - b * i), inv * (b * f - c * e),
                        inv * (f * g - d * i), inv * (a * i - c * g), inv * (c * d - a * f),
                        inv * (d * h - e * g), inv * (b * g - a * h), inv * (a * e - b * d)};
  mat3_mul(posed + 9 * (int64_t)s, ri, F + 9 * (int64_t)s);
}

}  // namespace cn

using namespace cn;

extern "C" {

cn_status cn_deform_rhs(cn_system S, int32_t P, const double *d_Ffree, const double *d_Fc,
                        const double *d_Xc, double *d_rhs, void *stream) {
  CN_API_BEGIN
  if (!S || P <= 0 || !d_rhs || (S->n_free > 0 && !d_Ffree) ||
      (S->n_con > 0 && (!d_Fc || !d_Xc)) || (S->nv == 0 && S->n_free > 0)) {
    set_error("cn_deform_rhs: bad arguments (a system with a mesh is needed)");
    return CN_BAD_ARG;
  }
  hipStream_t s = as_stream(stream);
  const int64_t k9 = 9 * (int64_t)P, k3 = 3 * (int64_t)P;
  const size_t need = (size_t)std::max(S->nf, 1) * (size_t)k9;
  if (S->deform_ff.n < need) S->deform_ff.alloc(need);
  if (S->nf > 0)
    k_face_F<<<grid_for((int64_t)S->nf * k9, 256), 256, 0, s>>>(
        S->nf, S->tris.get(), S->free_of.get(), S->con_of.get(), d_Ffree, d_Fc, k9,
        S->deform_ff.get());
  if (S->n_free > 0)
    k_deform_rhs<<<grid_for((int64_t)S->n_free * k3, 256), 256, 0, s>>>(
        S->n_free, k3, S->free_vertex.get(), S->tris.get(), S->inc.ptr.get(), S->inc.he.get(),
        S->face_m.get(), S->con_of.get(), S->pos.get(), S->deform_ff.get(), d_Xc, d_rhs);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_bezier3(int32_t ncurves, const double *d_ctrl, int32_t nt, const double *d_t,
                     double *d_out, void *stream) {
  CN_API_BEGIN
  if (ncurves < 0 || nt < 0 || ((int64_t)ncurves * nt > 0 && (!d_ctrl || !d_t || !d_out))) {
    set_error("cn_bezier3: bad arguments");
    return CN_BAD_ARG;
  }
  const int64_t tot = (int64_t)ncurves * nt;
  if (tot > 0)
    k_bezier3<<<grid_for(tot, 256), 256, 0, as_stream(stream)>>>(ncurves, d_ctrl, nt, d_t,
                                                                  d_out);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_segment_gradients(int32_t n, const double *d_rest_p, const double *d_rest_q,
                               const double *d_posed_p, const double *d_posed_q, double *d_F,
                               void *stream) {
  CN_API_BEGIN
  if (n < 0 || (n > 0 && (!d_rest_p || !d_rest_q || !d_posed_p || !d_posed_q || !d_F))) {
    set_error("cn_segment_gradients: bad arguments");
    return CN_BAD_ARG;
  }
  if (n > 0)
    k_segment_gradients<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(
        n, d_rest_p, d_rest_q, d_posed_p, d_posed_q, d_F);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

cn_status cn_deformation_gradients(int32_t n, const double *d_rest, const double *d_posed,
                                   double *d_F, void *stream) {
  CN_API_BEGIN
  if (n < 0 || (n > 0 && (!d_rest || !d_posed || !d_F))) {
    set_error("cn_deformation_gradients: bad arguments");
    return CN_BAD_ARG;
  }
  if (n > 0)
    k_deformation_gradients<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(n, d_rest,
                                                                             d_posed, d_F);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_API_END
}

}  // extern "C"
