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

template <class T> struct v3 { T x, y, z; };
template <class T> struct v4 { T x, y, z, w; };

// scalar helpers dispatching to the FP32 or FP64 correctly rounded routines
RTN_HD float xsqrt(float a) { return sqrtf(a); }
RTN_HD double xsqrt(double a) { return sqrt(a); }
RTN_HD float xabs(float a) { return fabsf(a); }
RTN_HD double xabs(double a) { return fabs(a); }
RTN_HD float xmin(float a, float b) { return fminf(a, b); }
RTN_HD double xmin(double a, double b) { return fmin(a, b); }
RTN_HD float xmax(float a, float b) { return fmaxf(a, b); }
RTN_HD double xmax(double a, double b) { return fmax(a, b); }
RTN_HD float xcopysign(float a, float b) { return copysignf(a, b); }
RTN_HD double xcopysign(double a, double b) { return copysign(a, b); }

template <class T> RTN_HD v3<T> sub(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> RTN_HD v3<T> add(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> RTN_HD v3<T> mul(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
// dotProduct (vector_ops.h:52-55): (x*x' + y*y') + z*z'
template <class T> RTN_HD T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// cross (vector_ops.h:115-119)
template <class T> RTN_HD v3<T> cross(v3<T> a, v3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// reciprocal_length, host form (vector_ops.h:102-106): correctly rounded 1/sqrt
template <class T> RTN_HD T rlen(v3<T> a) { return T(1) / xsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
template <class T> RTN_HD v3<T> normalize(v3<T> a) { return mul(a, rlen(a)); }
template <class T> RTN_HD T len(v3<T> a) { return xsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

template <class T> RTN_HD v4<T> ld4(const float4& f) { return {T(f.x), T(f.y), T(f.z), T(f.w)}; }
// sum(multiply(1-u, a), multiply(u, b)) (bezier_patch.h:63)
template <class T> RTN_HD v4<T> lerp4(v4<T> a, v4<T> b, T u) {
  T om = T(1) - u;
  return {a.x * om + b.x * u, a.y * om + b.y * u, a.z * om + b.z * u, a.w * om + b.w * u};
}
// multiply(n, subtract(b, a)) (bezier_patch.h:75)
template <class T> RTN_HD v4<T> dscale(v4<T> a, v4<T> b, T n) {
  return {(b.x - a.x) * n, (b.y - a.y) * n, (b.z - a.z) * n, (b.w - a.w) * n};
}

// dehom (bezier_patch.h:32-35): multiply by 1/w
template <class T> RTN_HD v3<T> dehom(v4<T> c) {
  T iw = T(1) / c.w;
  return {c.x * iw, c.y * iw, c.z * iw};
}

// ----------------------------------------------------------- de Casteljau
// deCasteljau1 (bezier_patch.h:57-65) on a register array, compile-time degree.
template <int N, class T>
RTN_HD v4<T> decas1(const v4<T>* P, T u) {
  v4<T> Q[N + 1];
#pragma unroll
  for (int i = 0; i <= N; i++) Q[i] = P[i];
#pragma unroll
  for (int k = 1; k <= N; k++)
#pragma unroll
    for (int i = 0; i <= N - k; i++) Q[i] = lerp4(Q[i], Q[i + 1], u);
  return Q[0];
}
// deCasteljauDers (bezier_patch.h:72-80), compile-time degree.
template <int N, class T>
RTN_HD v4<T> decas_der(const v4<T>* P, T u) {
  v4<T> Q[N];
#pragma unroll
  for (int i = 0; i < N; i++) Q[i] = dscale(P[i], P[i + 1], T(N));
#pragma unroll
  for (int k = 1; k < N; k++)
#pragma unroll
    for (int i = 0; i <= N - 1 - k; i++) Q[i] = lerp4(Q[i], Q[i + 1], u);
  return Q[0];
}

// Runtime-degree forms (degree n <= MAXD), fully unrolled with guards so the
// arrays stay in registers; the operations applied for a given n are those of
// the compile-time forms.
template <int MAXD, class T>
RTN_HD v4<T> decas1_rt(const v4<T>* P, int n, T u) {
  v4<T> Q[MAXD + 1];
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
template <int MAXD, class T>
RTN_HD v4<T> decas_der_rt(const v4<T>* P, int n, T u) {
  v4<T> Q[MAXD];
  const T fn = (T)n;
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
template <int P, int Q, class T>
__device__ __forceinline__ void patch_eval(const float4* __restrict__ cp, T u, T v, v4<T>& du, v4<T>& dv,
                                           v4<T>& pt) {
  v4<T> curve[(P > Q ? P : Q) + 1];
#pragma unroll
  for (int i = 0; i <= P; i++) {
    v4<T> row[Q + 1];
#pragma unroll
    for (int j = 0; j <= Q; j++) row[j] = ld4<T>(cp[i * (Q + 1) + j]);
    curve[i] = decas1<Q>(row, v);
  }
  du = decas_der<P>(curve, u);
  pt = decas1<P>(curve, u);
#pragma unroll
  for (int j = 0; j <= Q; j++) {
    v4<T> col[P + 1];
#pragma unroll
    for (int i = 0; i <= P; i++) col[i] = ld4<T>(cp[i * (Q + 1) + j]);
    curve[j] = decas1<P>(col, u);
  }
  dv = decas_der<Q>(curve, v);
}

template <int MAXD, class T>
__device__ __forceinline__ void patch_eval_rt(const float4* __restrict__ cp, int p, int q, T u, T v, v4<T>& du,
                                              v4<T>& dv, v4<T>& pt) {
  v4<T> curve[MAXD + 1];
#pragma unroll
  for (int i = 0; i <= MAXD; i++) {
    if (i <= p) {
      v4<T> row[MAXD + 1];
#pragma unroll
      for (int j = 0; j <= MAXD; j++)
        if (j <= q) row[j] = ld4<T>(cp[i * (q + 1) + j]);
      curve[i] = decas1_rt<MAXD>(row, q, v);
    }
  }
  du = decas_der_rt<MAXD>(curve, p, u);
  pt = decas1_rt<MAXD>(curve, p, u);
#pragma unroll
  for (int j = 0; j <= MAXD; j++) {
    if (j <= q) {
      v4<T> col[MAXD + 1];
#pragma unroll
      for (int i = 0; i <= MAXD; i++)
        if (i <= p) col[i] = ld4<T>(cp[i * (q + 1) + j]);
      curve[j] = decas1_rt<MAXD>(col, p, u);
    }
  }
  dv = decas_der_rt<MAXD>(curve, q, v);
}

// ------------------------------------------------------------ ray helpers
template <class T> struct RayPlane { v3<T> n1, n2; T d1, d2; };

// getRayPlanes (math_utils.h:33-54) with the host reciprocal_length
template <class T> RTN_HD RayPlane<T> ray_planes(v3<T> o, v3<T> d) {
  T ax = xabs(d.x), ay = xabs(d.y), az = xabs(d.z);
  bool xd = (ax > ay) && (ax > az);
  v3<T> base = xd ? v3<T>{d.x, d.y, T(0)} : v3<T>{d.y, d.z, T(0)};
  T den = rlen(base);
  v3<T> n1 = xd ? mul(v3<T>{d.y, -d.x, T(0)}, den) : mul(v3<T>{T(0), d.z, -d.y}, den);
  v3<T> n2 = cross(n1, mul(d, T(1) / len(d)));
  return {n1, n2, -dot(n1, o), -dot(n2, o)};
}

// intersectBound (bounds_ops.h:83-114)
template <class T>
RTN_HD T slab(T lx, T hx, T ly, T hy, T lz, T hz, v3<T> o, v3<T> rd) {
  T tmin = 0, tmax = T(INFINITY);
  T tx0 = (lx - o.x) * rd.x, tx1 = (hx - o.x) * rd.x;
  T ty0 = (ly - o.y) * rd.y, ty1 = (hy - o.y) * rd.y;
  T tz0 = (lz - o.z) * rd.z, tz1 = (hz - o.z) * rd.z;
  T nx = xmin(tx0, tx1), fx = xmax(tx0, tx1);
  T ny = xmin(ty0, ty1), fy = xmax(ty0, ty1);
  T nz = xmin(tz0, tz1), fz = xmax(tz0, tz1);
  tmin = xmax(nz, xmax(ny, xmax(nx, tmin)));
  tmax = xmin(fz, xmin(fy, xmin(fx, tmax)));
  return (tmin <= tmax && tmax > 0) ? tmin : T(INFINITY);
}

// biliniearRay (bezier_patch.h:273-327), as written (GARP's n = cross(n, pa)
// is absent in the reference; its accept test therefore runs on rounding
// noise, and the seeds here follow it exactly).  Split in two: the terms
// that depend on the patch alone (its corners' differences), which a scene
// may compute once per patch, and the per-ray solve; bilinear_seed runs both,
// and either route performs the same operations in the same order.
template <class T> struct SeedTerms { v3<T> q00, q10, e00, e11, qn; };

template <class T>
RTN_HD SeedTerms<T> seed_terms(v3<T> q00, v3<T> q10, v3<T> q11, v3<T> q01) {
  SeedTerms<T> s;
  const v3<T> e10 = sub(q10, q00);
  s.e11 = sub(q11, q10);
  s.e00 = sub(q01, q00);
  s.qn = cross(e10, sub(q01, q11));
  s.q00 = q00;
  s.q10 = q10;
  return s;
}

template <class T>
RTN_HD void seed_solve(const SeedTerms<T>& s, v3<T> orig, v3<T> dir, T& su, T& sv) {
  const v3<T> e00 = s.e00, e11 = s.e11;
  const v3<T> q00 = sub(s.q00, orig);
  const v3<T> q10 = sub(s.q10, orig);
  T a = dot(cross(q00, dir), e00);
  T c = dot(s.qn, dir);
  T b = dot(cross(q10, dir), e11);
  b -= a + c;
  T det = b * b - T(4) * a * c;
  su = T(0.5f);
  sv = T(0.5f);
  if (det < 0) return;
  det = xsqrt(det);
  T ux0, ux1;
  if (c == 0) {
    ux0 = -a / b;
    ux1 = -1;
  } else {
    ux0 = (-b - xcopysign(det, b)) / T(2);
    ux1 = a / ux0;
    ux0 /= c;
  }
  T best_t = T(INFINITY);
#pragma unroll
  for (int i = 0; i < 2; i++) {
    T ux = i == 0 ? ux0 : ux1;
    if (ux < 0 || ux > 1) continue;
    T om = T(1) - ux;
    v3<T> pb = {e00.x * om + e11.x * ux, e00.y * om + e11.y * ux, e00.z * om + e11.z * ux};
    v3<T> n = cross(dir, pb);
    T dd = dot(n, n);
    T tx = dot(n, pb);
    T vx = dot(n, dir);
    if (tx > 0 && 0 <= vx && vx <= dd && tx < best_t) {
      su = ux;
      sv = vx / dd;
      best_t = tx;
    }
  }
}

template <class T>
RTN_HD void bilinear_seed(v3<T> q00, v3<T> q10, v3<T> q11, v3<T> q01, v3<T> orig, v3<T> dir, T& su, T& sv) {
  seed_solve(seed_terms(q00, q10, q11, q01), orig, dir, su, sv);
}

// pcg_hash (math_utils.h:12-17) and randf (math_utils.h:70-74)
RTN_HD uint32_t pcg_hash(uint32_t x) {
  x = x * 747796405u + 2891336453u;
  uint32_t w = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
  return (w >> 22u) ^ w;
}
template <class T> RTN_HD T randf(uint32_t& seed) {
  seed = pcg_hash(seed);
  return (T)seed * T(2.3283064e-10f);
}

// Raster jitter of ray r, sample axis k (0: x, 1: y): 24 random bits scaled
// to [0,1) exactly.  Not in the reference (its rays sit at pixel centres).
RTN_HD float jitter(uint32_t r, uint32_t seed, uint32_t k) {
  uint32_t h = pcg_hash(pcg_hash(r * 2u + k) ^ (seed * 0x9E3779B9u));
  return (float)(h >> 8) * 5.9604645e-08f;
}

// S, Su, Sv from a homogeneous evaluation (bezier_patch.h:211-222)
template <class T>
RTN_HD void dehom_ders(v4<T> pt, v4<T> du, v4<T> dv, v3<T>& S, v3<T>& Su, v3<T>& Sv) {
  T invW = T(1) / pt.w;
  S = {pt.x * invW, pt.y * invW, pt.z * invW};
  Su = {(du.x - S.x * du.w) * invW, (du.y - S.y * du.w) * invW, (du.z - S.z * du.w) * invW};
  Sv = {(dv.x - S.x * dv.w) * invW, (dv.y - S.y * dv.w) * invW, (dv.z - S.z * dv.w) * invW};
}

template <class T> struct NewtonRes {
  bool hit;
  v3<T> S, Su, Sv;
  T u, v;
  int evals;
};

// newtonRhapson (bezier_patch.h:198-266); EvalFn(u, v, du, dv, pt) evaluates
// the patch.  2x2 determinant/inverse/multiply of matrix_ops.h:14-32.
template <class T, class EvalFn>
__device__ __forceinline__ NewtonRes<T> newton(const RayPlane<T>& rp, EvalFn eval, T u0, T v0, T thr_det,
                                               T thr_res, uint32_t seed, int max_iter) {
  T u = u0, v = v0;
  NewtonRes<T> r;
  r.hit = false;
  r.evals = 0;
  for (int it = 1; it <= max_iter; it++) {
    v4<T> du, dv, pt;
    eval(u, v, du, dv, pt);
    r.evals = it;
    v3<T> S, Su, Sv;
    dehom_ders(pt, du, dv, S, Su, Sv);
    T r0 = dot(rp.n1, S) + rp.d1;
    T r1 = dot(rp.n2, S) + rp.d2;
    T err = xabs(r0) + xabs(r1);
    if (err < thr_res) {
      r.hit = true;
      r.S = S; r.Su = Su; r.Sv = Sv; r.u = u; r.v = v;
      return r;
    }
    T m0 = dot(rp.n1, Su), m1 = dot(rp.n2, Su), m2 = dot(rp.n1, Sv), m3 = dot(rp.n2, Sv);
    T rand1 = randf<T>(seed);
    T rand2 = randf<T>(seed);
    T det = m0 * m3 - m2 * m1;
    if (xabs(det) < thr_det) {
      u = u + T(0.1f) * ((u0 - u) * rand1);
      v = v + T(0.1f) * ((v0 - v) * rand2);
    } else {
      T i0 = m3 / det, i1 = -m1 / det, i2 = -m2 / det, i3 = m0 / det;
      T ju = i0 * r0 + i2 * r1;
      T jv = i1 * r0 + i3 * r1;
      u = u - ju;
      v = v - jv;
    }
    u = xmin(xmax(u, T(0)), T(1));
    v = xmin(xmax(v, T(0)), T(1));
  }
  return r;
}

// Newton flop price of one surface evaluation + update at degree (p,q): the
// de Casteljau lerp counts 12 FP32 operations (4 components x 2 mul + 1 add).
RTN_HD int eval_flops(int p, int q) {
  return 12 * ((p + 1) * q * (q + 1) / 2 + (q + 1) * p * (p + 1) / 2 + p * (p + 1) / 2 + p * (p - 1) / 2 +
               q * (q - 1) / 2) + 8 * (p + q) + 100;
}

}  // namespace ex
}  // namespace rtn
