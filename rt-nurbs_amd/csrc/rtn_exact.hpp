// Reference-order FP32 arithmetic for the exact (drop-in) path.
//
// Each function performs the same IEEE FP32 operations, in the same order,
// as the reference function it names; translation units that include this
// header are compiled with -ffp-contract=off and correctly rounded division
// and square root, so device results equal the host reference bit for bit.
// Citations are Frostbite16/rt-nurbs file:line.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define RTN_HD __host__ __device__ __forceinline__

namespace rtn {
namespace ex {

struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };

RTN_HD v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RTN_HD v3 add(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RTN_HD v3 mul(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// dotProduct (vector_ops.h:52-55): (x*x' + y*y') + z*z'
RTN_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// cross (vector_ops.h:115-119)
RTN_HD v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// reciprocal_length, host form (vector_ops.h:102-106): correctly rounded 1/sqrt
RTN_HD float rlen(v3 a) { return 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
RTN_HD v3 normalize(v3 a) { return mul(a, rlen(a)); }
RTN_HD float len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }

RTN_HD v4 ld4(const float4& f) { return {f.x, f.y, f.z, f.w}; }
// sum(multiply(1-u, a), multiply(u, b)) (bezier_patch.h:63)
RTN_HD v4 lerp4(v4 a, v4 b, float u) {
  float om = 1.0f - u;
  return {a.x * om + b.x * u, a.y * om + b.y * u, a.z * om + b.z * u, a.w * om + b.w * u};
}
// multiply(n, subtract(b, a)) (bezier_patch.h:75)
RTN_HD v4 dscale(v4 a, v4 b, float n) { return {(b.x - a.x) * n, (b.y - a.y) * n, (b.z - a.z) * n, (b.w - a.w) * n}; }

// dehom (bezier_patch.h:32-35): multiply by 1/w
RTN_HD v3 dehom(v4 c) {
  float iw = 1.f / c.w;
  return {c.x * iw, c.y * iw, c.z * iw};
}

// ----------------------------------------------------------- de Casteljau
// deCasteljau1 (bezier_patch.h:57-65) on a register array, compile-time degree.
template <int N>
RTN_HD v4 decas1(const v4* P, float u) {
  v4 Q[N + 1];
#pragma unroll
  for (int i = 0; i <= N; i++) Q[i] = P[i];
#pragma unroll
  for (int k = 1; k <= N; k++)
#pragma unroll
    for (int i = 0; i <= N - k; i++) Q[i] = lerp4(Q[i], Q[i + 1], u);
  return Q[0];
}
// deCasteljauDers (bezier_patch.h:72-80), compile-time degree.
template <int N>
RTN_HD v4 decas_der(const v4* P, float u) {
  v4 Q[N];
#pragma unroll
  for (int i = 0; i < N; i++) Q[i] = dscale(P[i], P[i + 1], (float)N);
#pragma unroll
  for (int k = 1; k < N; k++)
#pragma unroll
    for (int i = 0; i <= N - 1 - k; i++) Q[i] = lerp4(Q[i], Q[i + 1], u);
  return Q[0];
}

// Runtime-degree forms (degree n <= MAXD), fully unrolled with guards so the
// arrays stay in registers; the operations applied for a given n are those of
// the compile-time forms.
template <int MAXD>
RTN_HD v4 decas1_rt(const v4* P, int n, float u) {
  v4 Q[MAXD + 1];
#pragma unroll
  for (int i = 0; i <= MAXD; i++)
    if (i <= n) Q[i] = P[i];
#pragma unroll
  for (int k = 1; k <= MAXD; k++)
#pragma unroll
    for (int i = 0; i <= MAXD - k; i++)
      if (i <= n - k) Q[i] = lerp4(Q[i], Q[i + 1], u);
  return Q[0];
}
template <int MAXD>
RTN_HD v4 decas_der_rt(const v4* P, int n, float u) {
  v4 Q[MAXD];
  const float fn = (float)n;
#pragma unroll
  for (int i = 0; i < MAXD; i++)
    if (i < n) Q[i] = dscale(P[i], P[i + 1], fn);
#pragma unroll
  for (int k = 1; k < MAXD; k++)
#pragma unroll
    for (int i = 0; i <= MAXD - 1 - k; i++)
      if (i <= n - 1 - k) Q[i] = lerp4(Q[i], Q[i + 1], u);
  return Q[0];
}

// BezierPatchEval (bezier_patch.h:173-191), compile-time degrees; the net is
// read from memory row-major [i*(Q+1)+j].
template <int P, int Q>
__device__ __forceinline__ void patch_eval(const float4* __restrict__ cp, float u, float v, v4& du, v4& dv, v4& pt) {
  v4 curve[(P > Q ? P : Q) + 1];
#pragma unroll
  for (int i = 0; i <= P; i++) {
    v4 row[Q + 1];
#pragma unroll
    for (int j = 0; j <= Q; j++) row[j] = ld4(cp[i * (Q + 1) + j]);
    curve[i] = decas1<Q>(row, v);
  }
  du = decas_der<P>(curve, u);
  pt = decas1<P>(curve, u);
#pragma unroll
  for (int j = 0; j <= Q; j++) {
    v4 col[P + 1];
#pragma unroll
    for (int i = 0; i <= P; i++) col[i] = ld4(cp[i * (Q + 1) + j]);
    curve[j] = decas1<P>(col, u);
  }
  dv = decas_der<Q>(curve, v);
}

template <int MAXD>
__device__ __forceinline__ void patch_eval_rt(const float4* __restrict__ cp, int p, int q, float u, float v, v4& du,
                                              v4& dv, v4& pt) {
  v4 curve[MAXD + 1];
#pragma unroll
  for (int i = 0; i <= MAXD; i++) {
    if (i <= p) {
      v4 row[MAXD + 1];
#pragma unroll
      for (int j = 0; j <= MAXD; j++)
        if (j <= q) row[j] = ld4(cp[i * (q + 1) + j]);
      curve[i] = decas1_rt<MAXD>(row, q, v);
    }
  }
  du = decas_der_rt<MAXD>(curve, p, u);
  pt = decas1_rt<MAXD>(curve, p, u);
#pragma unroll
  for (int j = 0; j <= MAXD; j++) {
    if (j <= q) {
      v4 col[MAXD + 1];
#pragma unroll
      for (int i = 0; i <= MAXD; i++)
        if (i <= p) col[i] = ld4(cp[i * (q + 1) + j]);
      curve[j] = decas1_rt<MAXD>(col, p, u);
    }
  }
  dv = decas_der_rt<MAXD>(curve, q, v);
}

// ------------------------------------------------------------ ray helpers
struct RayPlane { v3 n1, n2; float d1, d2; };

// getRayPlanes (math_utils.h:33-54) with the host reciprocal_length
RTN_HD RayPlane ray_planes(v3 o, v3 d) {
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  bool xd = (ax > ay) && (ax > az);
  v3 base = xd ? v3{d.x, d.y, 0.0f} : v3{d.y, d.z, 0.0f};
  float den = rlen(base);
  v3 n1 = xd ? mul(v3{d.y, -d.x, 0.0f}, den) : mul(v3{0.0f, d.z, -d.y}, den);
  v3 n2 = cross(n1, mul(d, 1.0f / len(d)));
  return {n1, n2, -dot(n1, o), -dot(n2, o)};
}

// intersectBound (bounds_ops.h:83-114)
RTN_HD float slab(float lx, float hx, float ly, float hy, float lz, float hz, v3 o, v3 rd) {
  float tmin = 0, tmax = INFINITY;
  float tx0 = (lx - o.x) * rd.x, tx1 = (hx - o.x) * rd.x;
  float ty0 = (ly - o.y) * rd.y, ty1 = (hy - o.y) * rd.y;
  float tz0 = (lz - o.z) * rd.z, tz1 = (hz - o.z) * rd.z;
  float nx = fminf(tx0, tx1), fx = fmaxf(tx0, tx1);
  float ny = fminf(ty0, ty1), fy = fmaxf(ty0, ty1);
  float nz = fminf(tz0, tz1), fz = fmaxf(tz0, tz1);
  tmin = fmaxf(nz, fmaxf(ny, fmaxf(nx, tmin)));
  tmax = fminf(fz, fminf(fy, fminf(fx, tmax)));
  return (tmin <= tmax && tmax > 0) ? tmin : INFINITY;
}

// biliniearRay (bezier_patch.h:273-327), as written (GARP's n = cross(n, pa)
// is absent in the reference; its accept test therefore runs on rounding
// noise, and the seeds here follow it exactly).
RTN_HD void bilinear_seed(v3 q00, v3 q10, v3 q11, v3 q01, v3 orig, v3 dir, float& su, float& sv) {
  v3 e10 = sub(q10, q00);
  v3 e11 = sub(q11, q10);
  v3 e00 = sub(q01, q00);
  v3 qn = cross(e10, sub(q01, q11));
  q00 = sub(q00, orig);
  q10 = sub(q10, orig);
  float a = dot(cross(q00, dir), e00);
  float c = dot(qn, dir);
  float b = dot(cross(q10, dir), e11);
  b -= a + c;
  float det = b * b - 4.0f * a * c;
  su = 0.5f;
  sv = 0.5f;
  if (det < 0) return;
  det = sqrtf(det);
  float ux0, ux1;
  if (c == 0) {
    ux0 = -a / b;
    ux1 = -1;
  } else {
    ux0 = (-b - copysignf(det, b)) / 2.0f;
    ux1 = a / ux0;
    ux0 /= c;
  }
  float best_t = INFINITY;
#pragma unroll
  for (int i = 0; i < 2; i++) {
    float ux = i == 0 ? ux0 : ux1;
    if (ux < 0 || ux > 1) continue;
    float om = 1.0f - ux;
    v3 pb = {e00.x * om + e11.x * ux, e00.y * om + e11.y * ux, e00.z * om + e11.z * ux};
    v3 n = cross(dir, pb);
    float dd = dot(n, n);
    float tx = dot(n, pb);
    float vx = dot(n, dir);
    if (tx > 0 && 0 <= vx && vx <= dd && tx < best_t) {
      su = ux;
      sv = vx / dd;
      best_t = tx;
    }
  }
}

// pcg_hash (math_utils.h:12-17) and randf (math_utils.h:70-74)
RTN_HD uint32_t pcg_hash(uint32_t x) {
  x = x * 747796405u + 2891336453u;
  uint32_t w = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
  return (w >> 22u) ^ w;
}
RTN_HD float randf(uint32_t& seed) {
  seed = pcg_hash(seed);
  return (float)seed * 2.3283064e-10f;
}

// Raster jitter of ray r, sample axis k (0: x, 1: y): 24 random bits scaled
// to [0,1) exactly.  Not in the reference (its rays sit at pixel centres).
RTN_HD float jitter(uint32_t r, uint32_t seed, uint32_t k) {
  uint32_t h = pcg_hash(pcg_hash(r * 2u + k) ^ (seed * 0x9E3779B9u));
  return (float)(h >> 8) * 5.9604645e-08f;
}

// S, Su, Sv from a homogeneous evaluation (bezier_patch.h:211-222)
RTN_HD void dehom_ders(v4 pt, v4 du, v4 dv, v3& S, v3& Su, v3& Sv) {
  float invW = 1.0f / pt.w;
  S = {pt.x * invW, pt.y * invW, pt.z * invW};
  Su = {(du.x - S.x * du.w) * invW, (du.y - S.y * du.w) * invW, (du.z - S.z * du.w) * invW};
  Sv = {(dv.x - S.x * dv.w) * invW, (dv.y - S.y * dv.w) * invW, (dv.z - S.z * dv.w) * invW};
}

struct NewtonRes {
  bool hit;
  v3 S, Su, Sv;
  float u, v;
  int evals;
};

// newtonRhapson (bezier_patch.h:198-266); EvalFn(u, v, du, dv, pt) evaluates
// the patch.  2x2 determinant/inverse/multiply of matrix_ops.h:14-32.
template <class EvalFn>
__device__ __forceinline__ NewtonRes newton(const RayPlane& rp, EvalFn eval, float u0, float v0, float thr_det,
                                            float thr_res, uint32_t seed, int max_iter) {
  float u = u0, v = v0;
  NewtonRes r;
  r.hit = false;
  r.evals = 0;
  for (int it = 1; it <= max_iter; it++) {
    v4 du, dv, pt;
    eval(u, v, du, dv, pt);
    r.evals = it;
    v3 S, Su, Sv;
    dehom_ders(pt, du, dv, S, Su, Sv);
    float r0 = dot(rp.n1, S) + rp.d1;
    float r1 = dot(rp.n2, S) + rp.d2;
    float err = fabsf(r0) + fabsf(r1);
    if (err < thr_res) {
      r.hit = true;
      r.S = S; r.Su = Su; r.Sv = Sv; r.u = u; r.v = v;
      return r;
    }
    float m0 = dot(rp.n1, Su), m1 = dot(rp.n2, Su), m2 = dot(rp.n1, Sv), m3 = dot(rp.n2, Sv);
    float rand1 = randf(seed);
    float rand2 = randf(seed);
    float det = m0 * m3 - m2 * m1;
    if (fabsf(det) < thr_det) {
      u = u + 0.1f * ((u0 - u) * rand1);
      v = v + 0.1f * ((v0 - v) * rand2);
    } else {
      float i0 = m3 / det, i1 = -m1 / det, i2 = -m2 / det, i3 = m0 / det;
      float ju = i0 * r0 + i2 * r1;
      float jv = i1 * r0 + i3 * r1;
      u = u - ju;
      v = v - jv;
    }
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
  }
  return r;
}

}  // namespace ex
}  // namespace rtn
