// Arithmetic of the fast trace forms: explicit FMA, tensor-product Bernstein
// evaluation with control-net rows streamed from memory, and the Newton
// iteration of the plain method (same seed handling, PCG perturbation on a
// singular Jacobian, clamping) in that arithmetic; exact rejection tests; and
// the FP64 pieces of mixed precision (the FP64 seed, the error estimate that
// selects hits for the FP64 step, and the step itself).  Results agree with
// the reference-order path to rounding; they are checked against the FP64
// reference, not bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "rtn_exact.hpp"
#include "rtn_mem.hpp"

namespace rtn {
namespace fast {

using ex::pcg_hash;

struct f3 { float x, y, z; };


__device__ __forceinline__ float dot3(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// Bernstein basis of degree N at t and its derivative, by the scalar
// de Casteljau triangle: B^{N-1} first, then dB^N_i = N (B^{N-1}_{i-1} - B^{N-1}_i)
// and B^N from one more step.
template <int N>
__device__ __forceinline__ void bernstein(float t, float (&b)[N + 1], float (&db)[N + 1]) {
  const float s = 1.0f - t;
  float c[N + 1];
  c[0] = 1.0f;
#pragma unroll
  for (int k = 1; k < N; k++) {
    c[k] = t * c[k - 1];
#pragma unroll
    for (int i = k - 1; i >= 1; i--) c[i] = fmaf(s, c[i], t * c[i - 1]);
    c[0] = s * c[0];
  }
  // c = B^{N-1}
#pragma unroll
  for (int i = 0; i <= N; i++) {
    const float lo = i > 0 ? c[i - 1] : 0.0f;
    const float hi = i < N ? c[i] : 0.0f;
    db[i] = (float)N * (lo - hi);
    b[i] = fmaf(s, hi, t * lo);
  }
}

struct Plane2 { f3 n1, n2; float d1, d2; };

// Ray planes (math_utils.h:33-54) in FMA arithmetic.
__device__ __forceinline__ Plane2 planes(f3 o, f3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  const bool xd = (ax > ay) && (ax > az);
  const float b0 = xd ? d.x : d.y, b1 = xd ? d.y : d.z;
  const float den = 1.0f / sqrtf(fmaf(b0, b0, b1 * b1));
  f3 n1 = xd ? f3{d.y * den, -d.x * den, 0.f} : f3{0.f, d.z * den, -d.y * den};
  const float il = 1.0f / sqrtf(dot3(d, d));
  const f3 dn{d.x * il, d.y * il, d.z * il};
  f3 n2{fmaf(n1.y, dn.z, -n1.z * dn.y), fmaf(n1.z, dn.x, -n1.x * dn.z), fmaf(n1.x, dn.y, -n1.y * dn.x)};
  return {n1, n2, -dot3(n1, o), -dot3(n2, o)};
}

struct Hit {
  bool hit;
  float u, v;
  f3 S, Su, Sv;
  float m0, m1, m2, m3;  // Jacobian at the returned iterate
  float r0, r1;          // residual (ray-plane distances) at the returned iterate
  int evals;
};

// Dehomogenised point and partials.  The reciprocal is the hardware one
// (within 1 ulp, one instruction against the ten of a correctly rounded
// divide): Newton corrects its own rounding, and a 1-ulp weight error moves a
// hit point by about 6e-8 of its size.
__device__ __forceinline__ void dehom(float4 pt, float4 du, float4 dv, f3& S, f3& Su, f3& Sv) {
  const float iw = __builtin_amdgcn_rcpf(pt.w);
  S = {pt.x * iw, pt.y * iw, pt.z * iw};
  Su = {fmaf(-S.x, du.w, du.x) * iw, fmaf(-S.y, du.w, du.y) * iw, fmaf(-S.z, du.w, du.z) * iw};
  Sv = {fmaf(-S.x, dv.w, dv.x) * iw, fmaf(-S.y, dv.w, dv.y) * iw, fmaf(-S.z, dv.w, dv.z) * iw};
}

// 2-norm condition number of [m0 m2; m1 m3] from its Frobenius norm and
// determinant: cond = (F^2 + sqrt(F^4 - 4 det^2)) / (2 |det|).
__device__ __forceinline__ float cond2(float m0, float m1, float m2, float m3) {
  const float f2 = fmaf(m0, m0, fmaf(m1, m1, fmaf(m2, m2, m3 * m3)));
  const float det = fmaf(m0, m3, -m2 * m1);
  const float disc = fmaxf(fmaf(f2, f2, -4.0f * det * det), 0.0f);
  return (f2 + sqrtf(disc)) / (2.0f * fabsf(det));
}

}  // namespace fast
}  // namespace rtn

namespace rtn {
namespace fast {

struct D3 { double x, y, z; };

// Distance change along the ray that one more Newton step from an accepted
// iterate would make: delta = J^-1 r in (u,v), then d . (S_u du + S_v dv).
// The accepted (pre-update) iterate is off the root by about this much, so
// |dt| / t predicts the FP32 hit's relative t error; infinite for a singular J.
__device__ __forceinline__ float predicted_dt(const Hit& h, f3 d) {
  const float det = fmaf(h.m0, h.m3, -h.m2 * h.m1);
  if (det == 0.0f) return INFINITY;
  const float id = 1.0f / det;
  const float du = fmaf(h.m3, h.r0, -h.m2 * h.r1) * id;
  const float dv = fmaf(-h.m1, h.r0, h.m0 * h.r1) * id;
  const f3 dS{fmaf(h.Su.x, du, h.Sv.x * dv), fmaf(h.Su.y, du, h.Sv.y * dv), fmaf(h.Su.z, du, h.Sv.z * dv)};
  return fabsf(dot3(d, dS));
}

}  // namespace fast
}  // namespace rtn

namespace rtn {
namespace fast {

constexpr int MAXD = 5;

// Bernstein basis of runtime degree n <= MAXD at t (and its derivative) into
// b[0..MAXD], db[0..MAXD]; entries above n are zero.  Guarded unrolled loops:
// for a wave-uniform n the guards are scalar branches.
__device__ __forceinline__ void bernstein_rt(int n, float t, float (&b)[MAXD + 1], float (&db)[MAXD + 1]) {
  const float s = 1.0f - t;
  float c[MAXD + 1];
#pragma unroll
  for (int i = 0; i <= MAXD; i++) c[i] = 0.0f;
  c[0] = 1.0f;
#pragma unroll
  for (int k = 1; k < MAXD; k++) {
    if (k < n) {
      c[k] = t * c[k - 1];
#pragma unroll
      for (int i = k - 1; i >= 1; i--) c[i] = fmaf(s, c[i], t * c[i - 1]);
      c[0] = s * c[0];
    }
  }
  const float fn = (float)n;
#pragma unroll
  for (int i = 0; i <= MAXD; i++) {
    const float lo = i > 0 ? c[i - 1] : 0.0f;
    const float hi = i < n ? c[i] : 0.0f;
    db[i] = i <= n ? fn * (lo - hi) : 0.0f;
    b[i] = i <= n ? fmaf(s, hi, t * lo) : 0.0f;
  }
}

// Bernstein basis of a wave-uniform runtime degree n (1..MAXD): a scalar
// branch to the compile-time basis, entries above n zero.  The same values as
// bernstein_rt, with no per-entry selects.
template <int N>
__device__ __forceinline__ void basis_into(float t, float (&b)[MAXD + 1], float (&db)[MAXD + 1]) {
  float x[N + 1], y[N + 1];
  bernstein<N>(t, x, y);
#pragma unroll
  for (int i = 0; i <= N; i++) {
    b[i] = x[i];
    db[i] = y[i];
  }
}

__device__ __forceinline__ void bernstein_uniform(int n, float t, float (&b)[MAXD + 1], float (&db)[MAXD + 1]) {
#pragma unroll
  for (int i = 0; i <= MAXD; i++) {
    b[i] = 0.0f;
    db[i] = 0.0f;
  }
  switch (n) {
    case 1: basis_into<1>(t, b, db); break;
    case 2: basis_into<2>(t, b, db); break;
    case 3: basis_into<3>(t, b, db); break;
    case 4: basis_into<4>(t, b, db); break;
    default: basis_into<5>(t, b, db); break;
  }
}

// Point and partials of a patch with runtime row count P+1 and compile-time
// row length Q+1: rows stream from memory; the u-direction weights rotate
// through registers (weight of row i is always in bu[0]).  A control net
// read through scalar loads (every lane on the same patch) has a wave-uniform
// P: its basis takes a scalar branch and its rows are unrolled behind
// scalar guards, so the next row's loads can issue early.
template <int Q, class NP = const float4*>
__device__ __forceinline__ void eval_q(NP net, int P, float u, float v, float4& pt,
                                       float4& du, float4& dv) {
  float bu[MAXD + 1], dbu[MAXD + 1], bv[Q + 1], dbv[Q + 1];
  if constexpr (std::is_same<NP, cptr<float4>>::value) {
    bernstein_uniform(P, u, bu, dbu);
    bernstein<Q>(v, bv, dbv);
    pt = make_float4(0.f, 0.f, 0.f, 0.f);
    du = pt;
    dv = pt;
#pragma unroll
    for (int i = 0; i <= MAXD; i++) {
      if (i > P) break;
      const NP row = net + i * (Q + 1);
      float4 r = make_float4(0.f, 0.f, 0.f, 0.f), rv = r;
#pragma unroll
      for (int j = 0; j <= Q; j++) {
        const float4 c = ld4(row, j);
        r.x = fmaf(bv[j], c.x, r.x); r.y = fmaf(bv[j], c.y, r.y); r.z = fmaf(bv[j], c.z, r.z); r.w = fmaf(bv[j], c.w, r.w);
        rv.x = fmaf(dbv[j], c.x, rv.x); rv.y = fmaf(dbv[j], c.y, rv.y);
        rv.z = fmaf(dbv[j], c.z, rv.z); rv.w = fmaf(dbv[j], c.w, rv.w);
      }
      const float b = bu[i], db = dbu[i];
      pt.x = fmaf(b, r.x, pt.x); pt.y = fmaf(b, r.y, pt.y); pt.z = fmaf(b, r.z, pt.z); pt.w = fmaf(b, r.w, pt.w);
      du.x = fmaf(db, r.x, du.x); du.y = fmaf(db, r.y, du.y); du.z = fmaf(db, r.z, du.z); du.w = fmaf(db, r.w, du.w);
      dv.x = fmaf(b, rv.x, dv.x); dv.y = fmaf(b, rv.y, dv.y); dv.z = fmaf(b, rv.z, dv.z); dv.w = fmaf(b, rv.w, dv.w);
    }
    return;
  }
  bernstein_rt(P, u, bu, dbu);
  bernstein<Q>(v, bv, dbv);
  pt = make_float4(0.f, 0.f, 0.f, 0.f);
  du = pt;
  dv = pt;
#pragma unroll 1
  for (int i = 0; i <= P; i++) {
    const NP row = net + i * (Q + 1);
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f), rv = r;
#pragma unroll
    for (int j = 0; j <= Q; j++) {
      const float4 c = ld4(row, j);
      r.x = fmaf(bv[j], c.x, r.x); r.y = fmaf(bv[j], c.y, r.y); r.z = fmaf(bv[j], c.z, r.z); r.w = fmaf(bv[j], c.w, r.w);
      rv.x = fmaf(dbv[j], c.x, rv.x); rv.y = fmaf(dbv[j], c.y, rv.y);
      rv.z = fmaf(dbv[j], c.z, rv.z); rv.w = fmaf(dbv[j], c.w, rv.w);
    }
    const float b = bu[0], db = dbu[0];
    pt.x = fmaf(b, r.x, pt.x); pt.y = fmaf(b, r.y, pt.y); pt.z = fmaf(b, r.z, pt.z); pt.w = fmaf(b, r.w, pt.w);
    du.x = fmaf(db, r.x, du.x); du.y = fmaf(db, r.y, du.y); du.z = fmaf(db, r.z, du.z); du.w = fmaf(db, r.w, du.w);
    dv.x = fmaf(b, rv.x, dv.x); dv.y = fmaf(b, rv.y, dv.y); dv.z = fmaf(b, rv.z, dv.z); dv.w = fmaf(b, rv.w, dv.w);
#pragma unroll
    for (int k = 0; k < MAXD; k++) {
      bu[k] = bu[k + 1];
      dbu[k] = dbu[k + 1];
    }
  }
}

// Newton on one patch (compile-time row length Q+1).  With `stall`, a solve
// is abandoned as a miss from its 4th evaluation on when two consecutive
// updates leave the parameter domain on the same side(s) and the residual fell
// by less than 10%: the iterate is stuck against the patch boundary.
template <int Q, class NP = const float4*>
__device__ __forceinline__ Hit newton_q(NP net, int P, const Plane2& rp, float u0, float v0,
                                        float thr_det, float thr_res, uint32_t seed, int max_iter, bool stall) {
  float u = u0, v = v0;
  Hit h;
  h.hit = false;
  h.evals = 0;
  int prev_side = 0;
  float prev_err = INFINITY;
  int pend = 0;  // PCG draws owed: taken only on the rare singular step
#pragma unroll 1
  for (int it = 1; it <= max_iter; it++) {
    float4 pt, du, dv;
    eval_q<Q>(net, P, u, v, pt, du, dv);
    h.evals = it;
    f3 S, Su, Sv;
    dehom(pt, du, dv, S, Su, Sv);
    const float r0 = dot3(rp.n1, S) + rp.d1;
    const float r1 = dot3(rp.n2, S) + rp.d2;
    const float m0 = dot3(rp.n1, Su), m1 = dot3(rp.n2, Su), m2 = dot3(rp.n1, Sv), m3 = dot3(rp.n2, Sv);
    const float err = fabsf(r0) + fabsf(r1);
    if (err < thr_res) {
      h.hit = true;
      h.u = u; h.v = v; h.S = S; h.Su = Su; h.Sv = Sv;
      h.m0 = m0; h.m1 = m1; h.m2 = m2; h.m3 = m3;
      h.r0 = r0; h.r1 = r1;
      return h;
    }
    // two PCG draws per iteration, as the plain method; the state is only
    // advanced when a singular step uses them (same values, same results)
    pend += 2;
    const float det = fmaf(m0, m3, -m2 * m1);
    if (fabsf(det) < thr_det) {
      for (; pend > 2; pend--) seed = pcg_hash(seed);
      seed = pcg_hash(seed);
      const float rand1 = (float)seed * 2.3283064e-10f;
      seed = pcg_hash(seed);
      const float rand2 = (float)seed * 2.3283064e-10f;
      pend = 0;
      u = fmaf(0.1f, (u0 - u) * rand1, u);
      v = fmaf(0.1f, (v0 - v) * rand2, v);
    } else {
      const float id = __builtin_amdgcn_rcpf(det);  // |det| >= thr_det: a normal number
      u = u - fmaf(m3, r0, -m2 * r1) * id;
      v = v - fmaf(-m1, r0, m0 * r1) * id;
    }
    if (stall) {
      const int side = (u < 0.0f) | (u > 1.0f) << 1 | (v < 0.0f) << 2 | (v > 1.0f) << 3;
      if (it >= 4 && side != 0 && side == prev_side && err > 0.9f * prev_err) return h;
      prev_side = side;
      prev_err = err;
    }
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
  }
  return h;
}

// Bernstein basis of runtime degree n <= MAXD at t and its derivative in FP64
// (the de Casteljau triangle of bernstein_rt), entries above n zero.
__device__ __forceinline__ void bernstein64_rt(int n, double t, double (&b)[MAXD + 1], double (&db)[MAXD + 1]) {
  const double s = 1.0 - t;
  double c[MAXD + 1];
#pragma unroll
  for (int i = 0; i <= MAXD; i++) c[i] = 0.0;
  c[0] = 1.0;
#pragma unroll
  for (int k = 1; k < MAXD; k++) {
    if (k < n) {
      c[k] = t * c[k - 1];
#pragma unroll
      for (int i = k - 1; i >= 1; i--) c[i] = fma(s, c[i], t * c[i - 1]);
      c[0] = s * c[0];
    }
  }
  const double dn = (double)n;
#pragma unroll
  for (int i = 0; i <= MAXD; i++) {
    const double lo = i > 0 ? c[i - 1] : 0.0;
    const double hi = i < n ? c[i] : 0.0;
    db[i] = i <= n ? dn * (lo - hi) : 0.0;
    b[i] = i <= n ? fma(s, hi, t * lo) : 0.0;
  }
}

// FP64 point and partials, runtime degrees: the tensor-product Bernstein sum
// with both bases formed once per evaluation.
template <class NP = const float4*>
__device__ __forceinline__ void eval64_rt(NP net, int P, int Q, double u, double v, D3& S,
                                          D3& Su, D3& Sv) {
  double bu[MAXD + 1], dbu[MAXD + 1], bv[MAXD + 1], dbv[MAXD + 1];
  bernstein64_rt(P, u, bu, dbu);
  bernstein64_rt(Q, v, bv, dbv);
  double pt[4] = {0, 0, 0, 0}, du[4] = {0, 0, 0, 0}, dv[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i <= MAXD; i++) {
    if (i > P) break;
    // the row's control points are loaded together (guarded, not a loop of
    // dependent loads), then accumulated; the basis is zero past the degree
    float4 row[MAXD + 1];
#pragma unroll
    for (int j = 0; j <= MAXD; j++) row[j] = j <= Q ? ld4(net, i * (Q + 1) + j) : make_float4(0.f, 0.f, 0.f, 0.f);
    double r[4] = {0, 0, 0, 0}, rv[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j <= MAXD; j++) {
      const double cc[4] = {row[j].x, row[j].y, row[j].z, row[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        r[k] = fma(bv[j], cc[k], r[k]);
        rv[k] = fma(dbv[j], cc[k], rv[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pt[k] = fma(bu[i], r[k], pt[k]);
      du[k] = fma(dbu[i], r[k], du[k]);
      dv[k] = fma(bu[i], rv[k], dv[k]);
    }
  }
  const double iw = 1.0 / pt[3];
  S = {pt[0] * iw, pt[1] * iw, pt[2] * iw};
  Su = {(du[0] - S.x * du[3]) * iw, (du[1] - S.y * du[3]) * iw, (du[2] - S.z * du[3]) * iw};
  Sv = {(dv[0] - S.x * dv[3]) * iw, (dv[1] - S.y * dv[3]) * iw, (dv[2] - S.z * dv[3]) * iw};
}

// The FP64 correction of a flagged FP32 hit: Newton steps in FP64 from the
// FP32 (u,v) on FP64 ray planes (at most kFix64Steps; one from the usual FP32
// accept, a few more from the far accepts of grazing rays, where the t-interval
// within the residual tolerance is ~ thr / tan(angle)), then the corrected
// point.  A step is kept only when it stays in the patch's domain [0,1]^2 (the
// plain method clamps its iterates there) and lowers the FP64 residual
// |r0| + |r1|; the iteration ends at the first step that does not, or once
// the residual is below 1e-12 of the point's magnitude.  Returns false (the FP32 hit stands)
// when not even the first step is kept.
constexpr int kFix64Steps = 4;

template <class NP = const float4*>
__device__ __forceinline__ bool correct64_rt(NP net, int P, int Q, f3 o, f3 d, float& uo,
                                             float& vo, f3& So, f3& Suo, f3& Svo, float& t_out) {
  const double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
  const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
  const bool xd = (ax > ay) && (ax > az);
  const double b0 = xd ? dx : dy, b1 = xd ? dy : dz;
  const double den = 1.0 / sqrt(b0 * b0 + b1 * b1);
  const D3 n1 = xd ? D3{dy * den, -dx * den, 0.0} : D3{0.0, dz * den, -dy * den};
  const double il = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
  const D3 dn{dx * il, dy * il, dz * il};
  const D3 n2{n1.y * dn.z - n1.z * dn.y, n1.z * dn.x - n1.x * dn.z, n1.x * dn.y - n1.y * dn.x};
  const double d1 = -(n1.x * ox + n1.y * oy + n1.z * oz), d2 = -(n2.x * ox + n2.y * oy + n2.z * oz);
  double u = uo, v = vo;
  D3 S, Su, Sv;
  eval64_rt(net, P, Q, u, v, S, Su, Sv);
  double r0 = n1.x * S.x + n1.y * S.y + n1.z * S.z + d1;
  double r1 = n2.x * S.x + n2.y * S.y + n2.z * S.z + d2;
  double err = fabs(r0) + fabs(r1);
  // the FP64 reference's own convergence level (thr_res 1e-10), relative to the scene's scale
  const double floor_err = 1e-12 * (1.0 + fabs(S.x) + fabs(S.y) + fabs(S.z));
  bool moved = false;
  for (int step = 0; step < kFix64Steps && err > floor_err; step++) {
    const double m0 = n1.x * Su.x + n1.y * Su.y + n1.z * Su.z, m1 = n2.x * Su.x + n2.y * Su.y + n2.z * Su.z;
    const double m2 = n1.x * Sv.x + n1.y * Sv.y + n1.z * Sv.z, m3 = n2.x * Sv.x + n2.y * Sv.y + n2.z * Sv.z;
    const double det = m0 * m3 - m2 * m1;
    if (det == 0.0) break;
    const double un = u - (m3 * r0 - m2 * r1) / det;
    const double vn = v - (-m1 * r0 + m0 * r1) / det;
    if (!(un >= 0.0 && un <= 1.0 && vn >= 0.0 && vn <= 1.0)) break;
    D3 Sn, Sun, Svn;
    eval64_rt(net, P, Q, un, vn, Sn, Sun, Svn);
    const double s0 = n1.x * Sn.x + n1.y * Sn.y + n1.z * Sn.z + d1;
    const double s1 = n2.x * Sn.x + n2.y * Sn.y + n2.z * Sn.z + d2;
    const double en = fabs(s0) + fabs(s1);
    if (!(en < err)) break;
    u = un; v = vn; S = Sn; Su = Sun; Sv = Svn; r0 = s0; r1 = s1; err = en;
    moved = true;
  }
  if (!moved) return false;
  uo = (float)u;
  vo = (float)v;
  So = {(float)S.x, (float)S.y, (float)S.z};
  Suo = {(float)Su.x, (float)Su.y, (float)Su.z};
  Svo = {(float)Sv.x, (float)Sv.y, (float)Sv.z};
  t_out = (float)((S.x - ox) * dx + (S.y - oy) * dy + (S.z - oz) * dz);
  return true;
}

// The plain bilinear seed (bezier_patch.h:38-44, :273-327) evaluated in FP64,
// exactly as the FP64 reference trace computes it: the patch's FP64 terms
// (corners dehomogenised in FP64, their differences; formed once per patch at
// scene build, ex::seed_terms) and the per-ray solve of the reference
// (ex::seed_solve, T = double), rounded to FP32.  The seed's accept test runs
// on rounding noise (its GARP cross product is missing), so an FP32
// evaluation starts Newton from (0.5, 0.5) or (u_root, ~0) by the sign of
// FP32 noise, and FP64 by the sign of FP64 noise; mixed precision starts
// where the FP64 reference starts.
template <class DP = const double*>
__device__ __forceinline__ void seed64(DP t, f3 o, f3 d, float& su, float& sv) {
  ex::SeedTerms<double> s;
  s.q00 = {t[0], t[1], t[2]};
  s.q10 = {t[3], t[4], t[5]};
  s.e00 = {t[6], t[7], t[8]};
  s.e11 = {t[9], t[10], t[11]};
  s.qn = {t[12], t[13], t[14]};
  double gu, gv;
  ex::seed_solve<double>(s, {(double)o.x, (double)o.y, (double)o.z}, {(double)d.x, (double)d.y, (double)d.z}, gu,
                         gv);
  su = (float)gu;
  sv = (float)gv;
}

}  // namespace fast
}  // namespace rtn

namespace rtn {
namespace fast {

// Exact rejection of a (ray, patch) pair from the patch's oriented box (the
// box holds the dehomogenised control points, hence the surface, when the
// weights are positive).  Newton converges only where |r0| + |r1| < thr, and
// |r0| + |r1| >= the distance from the ray's line, so a line that misses the
// box widened by the margin cannot converge on the patch.  Line (not ray)
// against box, slab test in the box frame with hardware reciprocals; the
// widening also covers the FP32 rounding of the transform.
template <class NP = const float4*>
__device__ __forceinline__ bool obb_culled(NP box, f3 o, f3 d, float margin) {
  const float4 a0 = ld4(box, 0), a1 = ld4(box, 1), a2 = ld4(box, 2), c = ld4(box, 3);
  const float ox = o.x - c.x, oy = o.y - c.y, oz = o.z - c.z;
  const float l0 = fmaf(a0.x, ox, fmaf(a0.y, oy, a0.z * oz)), l1 = fmaf(a1.x, ox, fmaf(a1.y, oy, a1.z * oz)),
              l2 = fmaf(a2.x, ox, fmaf(a2.y, oy, a2.z * oz));
  const float d0 = fmaf(a0.x, d.x, fmaf(a0.y, d.y, a0.z * d.z)), d1 = fmaf(a1.x, d.x, fmaf(a1.y, d.y, a1.z * d.z)),
              d2 = fmaf(a2.x, d.x, fmaf(a2.y, d.y, a2.z * d.z));
  const float slack = margin + 1e-6f * (fabsf(o.x) + fabsf(o.y) + fabsf(o.z) + fabsf(c.x) + fabsf(c.y) + fabsf(c.z));
  const float e0 = a0.w + slack, e1 = a1.w + slack, e2 = a2.w + slack;
  const float i0 = __builtin_amdgcn_rcpf(fabsf(d0) > 1e-30f ? d0 : copysignf(1e-30f, d0));
  const float i1 = __builtin_amdgcn_rcpf(fabsf(d1) > 1e-30f ? d1 : copysignf(1e-30f, d1));
  const float i2 = __builtin_amdgcn_rcpf(fabsf(d2) > 1e-30f ? d2 : copysignf(1e-30f, d2));
  const float p0 = (-e0 - l0) * i0, q0 = (e0 - l0) * i0;
  const float p1 = (-e1 - l1) * i1, q1 = (e1 - l1) * i1;
  const float p2 = (-e2 - l2) * i2, q2 = (e2 - l2) * i2;
  const float tn = fmaxf(fmaxf(fminf(p0, q0), fminf(p1, q1)), fminf(p2, q2));
  const float tf = fminf(fminf(fmaxf(p0, q0), fmaxf(p1, q1)), fmaxf(p2, q2));
  return tn > tf + 1e-5f * (fabsf(tn) + fabsf(tf));
}

// Exact rejection of a (ray, patch) pair from its control points projected on
// the ray planes.  With positive weights, the residual vector (r0, r1) of any
// surface point is a convex combination of the projected control points
// (c1/w, c2/w), c = n.P.xyz + d*P.w; Newton can only converge where
// |r0| + |r1| < thr, i.e. inside the L1 ball of that radius.  The pair is
// rejected when a line separates the projected points from the ball: tested
// along the eight directions (+-1,0), (0,+-1), (+-1,+-1), whose L1-ball support
// is the margin itself, with x in {c1, c2, c1+c2, c1-c2}.  The exact test
// min(x - m*w) > 0 (or max(x + m*w) < 0) is relaxed to min(x) > m*max(w) (or
// max(x) < -m*max(w)), which implies it (w > 0): a pair it rejects cannot
// converge, and it costs two comparisons per point and direction fewer.
// Without `diagonals`, the four axis directions only (the per-plane sign
// test).  Runtime degrees; no divisions.
template <class NP = const float4*>
__device__ __forceinline__ bool hull2d_culled(NP net, int n_cp, const Plane2& rp,
                                              float margin, bool diagonals) {
  float lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY, lo3 = INFINITY;
  float hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY, hi3 = -INFINITY;
  float wmax = 0.0f;
#pragma unroll 4
  for (int k = 0; k < n_cp; k++) {
    const float4 c = ld4(net, k);
    const float c1 = fmaf(rp.n1.x, c.x, fmaf(rp.n1.y, c.y, fmaf(rp.n1.z, c.z, rp.d1 * c.w)));
    const float c2 = fmaf(rp.n2.x, c.x, fmaf(rp.n2.y, c.y, fmaf(rp.n2.z, c.z, rp.d2 * c.w)));
    const float s = c1 + c2, d = c1 - c2;
    lo0 = fminf(lo0, c1); hi0 = fmaxf(hi0, c1);
    lo1 = fminf(lo1, c2); hi1 = fmaxf(hi1, c2);
    lo2 = fminf(lo2, s); hi2 = fmaxf(hi2, s);
    lo3 = fminf(lo3, d); hi3 = fmaxf(hi3, d);
    wmax = fmaxf(wmax, c.w);
  }
  if (!diagonals) {
    lo2 = lo3 = -INFINITY;
    hi2 = hi3 = INFINITY;
  }
  const float mw = margin * wmax;
  return fmaxf(fmaxf(lo0, lo1), fmaxf(lo2, lo3)) > mw || fminf(fminf(hi0, hi1), fminf(hi2, hi3)) < -mw;
}

}  // namespace fast
}  // namespace rtn
