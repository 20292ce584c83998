// Scene assembly on the GPU: NURBS -> rational Bezier decomposition (A5.7),
// patch bounds / seed corners / flatness, and the exhaustive-SAH binary BVH of
// the reference, rebuilt level-synchronously on the device, followed by the
// collapse to 4-wide structure-of-arrays nodes used by the fast tracers.
//
// Reference behaviour reproduced bit for bit (FP32, -ffp-contract=off):
//   segs / getPatch          bezier_patch.h:46-53, :87-139
//   decomposeSurface (A5.7)  nurbs.h:144-250
//   bounds                   bezier_patch.h:154-169
//   patchDelta               main.cpp:18-43
//   buildBVH / subdivide     bvh.h:47-179 (candidate order, strict '<', default
//                            axis 1, Hoare partition, depth-first node numbering)
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "rtn_exact.hpp"
#include "rtn_internal.hpp"

using namespace rtn;

namespace {

constexpr int MAXD = RTN_MAX_DEGREE;

struct SurfDesc {      // per-surface description on the device
  int n, p, m, q;
  int u_off, v_off;    // knot offsets (floats)
  int64_t pw_off;      // float4 offset of the input net
  int64_t strip_off;   // float4 offset of the strips
  int64_t patch_cp_off;// float4 offset of the surface's first patch net
  int patch_base;      // index of the surface's first patch
  int nbU, nbV;
};

// ------------------------------------------------------------ A5.7, 1 curve
// Knot insertion to multiplicity p at every interior knot of one curve
// (nurbs.h:152-198 restricted to one row).  Input point i at P[i*ps], i in
// [0,n]; output point i of segment b at Q[b*qs + i*qi].  span[b] receives the
// knot index a at which segment b starts.  Works in place in Q as the
// reference does, with the reference's blend order alpha*Q[k] + (1-alpha)*Q[k-1].
__device__ void decompose_curve(const float4* __restrict__ P, int64_t ps, int n, int p, const float* __restrict__ K,
                                float4* Q, int64_t qs, int64_t qi, int* span) {
  float alphas[MAXD];
  int a = p, b = p + 1, nb = 0;
  for (int i = 0; i <= p; i++) Q[i * qi] = P[i * ps];
  const int mu = n + p + 1;
  if (span) span[0] = a;
  while (b < mu) {
    const int i0 = b;
    while (b < mu && K[b + 1] == K[b]) b++;
    const int mult = b - i0 + 1;
    if (mult < p) {
      const float numer = K[b] - K[a];
      for (int j = p; j > mult; j--) alphas[j - mult - 1] = numer / (K[a + j] - K[a]);
      const int r = p - mult;
      for (int j = 1; j <= r; j++) {
        const int save = r - j, s = mult + j;
        for (int k = p; k >= s; k--) {
          const float alpha = alphas[k - s];
          const float om = 1.0f - alpha;
          const float4 A = Q[nb * qs + k * qi], B = Q[nb * qs + (k - 1) * qi];
          Q[nb * qs + k * qi] = make_float4(A.x * alpha + B.x * om, A.y * alpha + B.y * om, A.z * alpha + B.z * om,
                                            A.w * alpha + B.w * om);
        }
        if (b < mu) Q[(nb + 1) * qs + save * qi] = Q[nb * qs + p * qi];
      }
    }
    nb = nb + 1;
    if (b < mu) {
      for (int i = p - mult; i <= p; i++) Q[nb * qs + i * qi] = P[(int64_t)(b - p + i) * ps];
      a = b;
      b = b + 1;
      if (span) span[nb] = a;
    }
  }
}

// segs (bezier_patch.h:46-53): breakpoints into out[0..nb], returns nb.
__device__ int segs(const float* K, int last, int p, float* out) {
  int nb = 0;
  out[nb++] = K[p];
  for (int i = p + 1; i <= last - p - 1; ++i)
    if (K[i] != K[i - 1]) out[nb++] = K[i];
  out[nb] = K[last - p];
  return nb;
}

// One thread per surface: segment counts and breakpoints.
__global__ void k_count(SurfDesc* sd, int ns, const float* knots, float* brk) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  SurfDesc& d = sd[s];
  d.nbU = segs(knots + d.u_off, d.n + d.p + 1, d.p, brk + d.u_off);
  d.nbV = segs(knots + d.v_off, d.m + d.q + 1, d.q, brk + d.v_off);
}

// U pass: one block per surface, one thread per row of the net.
__global__ void __launch_bounds__(256) k_decomp_u(const SurfDesc* sd, const float* knots, const float4* pw,
                                                  float4* strips, int* span_u) {
  const SurfDesc d = sd[blockIdx.x];
  for (int r = threadIdx.x; r <= d.m; r += blockDim.x)
    decompose_curve(pw + d.pw_off + r, d.m + 1, d.n, d.p, knots + d.u_off, strips + d.strip_off + r,
                    (int64_t)(d.p + 1) * (d.m + 1), d.m + 1, r == 0 ? span_u + d.u_off : nullptr);
}

// V pass: one block per surface, one thread per (strip, row).  The reference
// calls decomposeSurface(V_DIRECTION) per strip with n := p (bezier_patch.h:108-112).
__global__ void __launch_bounds__(256) k_decomp_v(const SurfDesc* sd, const float* knots, const float4* strips,
                                                  float4* patches, int* span_v) {
  const SurfDesc d = sd[blockIdx.x];
  const int rows = d.p + 1;
  const int64_t ss = (int64_t)(d.p + 1) * (d.m + 1), ps = (int64_t)(d.p + 1) * (d.q + 1);
  for (int t = threadIdx.x; t < d.nbU * rows; t += blockDim.x) {
    const int s = t / rows, r = t % rows;
    decompose_curve(strips + d.strip_off + s * ss + (int64_t)r * (d.m + 1), 1, d.m, d.q, knots + d.v_off,
                    patches + d.patch_cp_off + (int64_t)s * d.nbV * ps + (int64_t)r * (d.q + 1), ps, 1,
                    (t == 0) ? span_v + d.v_off : nullptr);
  }
}

// Patch headers in getPatch order (bezier_patch.h:116-131), one thread per patch.
__global__ void k_headers(const SurfDesc* sd, int ns, int np, const float* brk, const int* span_u, const int* span_v,
                          int* cp_off, int* pq, float4* uv, int* pid, int2* kidx) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  int lo = 0, hi = ns - 1;  // surface containing patch g
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (sd[mid].patch_base <= g) lo = mid; else hi = mid - 1;
  }
  const SurfDesc& d = sd[lo];
  const int k = g - d.patch_base, i = k / d.nbV, j = k % d.nbV;
  const int ps = (d.p + 1) * (d.q + 1);
  cp_off[g] = (int)(d.patch_cp_off + (int64_t)k * ps);
  pq[g] = d.p | (d.q << 8);
  uv[g] = make_float4(brk[d.u_off + i], brk[d.u_off + i + 1], brk[d.v_off + j], brk[d.v_off + j + 1]);
  pid[g] = g;
  kidx[g] = make_int2(span_u[d.u_off + i], span_v[d.v_off + j]);
}

// Nets repacked in BVH order: patch i copies the net of src = perm[i].
__global__ void k_gather_nets(const float4* __restrict__ cp_in, const int* __restrict__ off_in,
                              const int* __restrict__ pq_in, const uint32_t* __restrict__ perm,
                              const int* __restrict__ off_out, int np, float4* __restrict__ cp_out) {
  const int i = blockIdx.x;
  if (i >= np) return;
  const int src = (int)perm[i];
  const int len = ((pq_in[src] & 0xff) + 1) * (((pq_in[src] >> 8) & 0xff) + 1);
  for (int k = threadIdx.x; k < len; k += blockDim.x) cp_out[off_out[i] + k] = cp_in[off_in[src] + k];
}

// bounds (bezier_patch.h:154-169): per-coordinate true division by w.
__global__ void k_bounds(const float4* cp, const int* cp_off, const int* pq, int np, float* bounds) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  const int p = pq[g] & 0xff, q = (pq[g] >> 8) & 0xff;
  const float4* c = cp + cp_off[g];
  float mn0 = INFINITY, mn1 = INFINITY, mn2 = INFINITY, mx0 = -INFINITY, mx1 = -INFINITY, mx2 = -INFINITY;
  for (int i = 0; i < (p + 1) * (q + 1); i++) {
    const float4 v = c[i];
    float x = v.x / v.w, y = v.y / v.w, z = v.z / v.w;
    mn0 = fminf(mn0, x); mx0 = fmaxf(mx0, x);
    mn1 = fminf(mn1, y); mx1 = fmaxf(mx1, y);
    mn2 = fminf(mn2, z); mx2 = fmaxf(mx2, z);
  }
  float* b = bounds + 6 * (int64_t)g;
  b[0] = mn0; b[1] = mn1; b[2] = mn2; b[3] = mx0; b[4] = mx1; b[5] = mx2;
}

// The patch terms of the bilinear seed (bezier_patch.h:273-286) in FP64, from
// the corners dehomogenised in FP64 (bezier_patch.h:32-44), exactly as the
// FP64 reference trace forms them: mixed precision starts Newton from the
// FP64 reference's seed and needs only the per-ray part of it per test.
// Layout per patch: q00, q10, e00, e11, qn (15 doubles) and one pad word.
__global__ void k_seed_terms(const float4* cp, const int* cp_off, const int* pq, int np, double* out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  const int p = pq[g] & 0xff, q = (pq[g] >> 8) & 0xff, row = q + 1;
  const float4* c = cp + cp_off[g];
  using namespace ex;
  const SeedTerms<double> t = seed_terms(dehom(ld4<double>(c[0])), dehom(ld4<double>(c[p * row])),
                                         dehom(ld4<double>(c[p * row + q])), dehom(ld4<double>(c[q])));
  const double w[kSeedTermWords] = {t.q00.x, t.q00.y, t.q00.z, t.q10.x, t.q10.y, t.q10.z, t.e00.x, t.e00.y,
                                    t.e00.z, t.e11.x, t.e11.y, t.e11.z, t.qn.x,  t.qn.y,  t.qn.z,  0.0};
  for (int k = 0; k < kSeedTermWords; k++) out[(int64_t)kSeedTermWords * g + k] = w[k];
}

// Seed corners (bezier_patch.h:38-44) and patchDelta (main.cpp:18-43); sets
// bit 16 of pq when every weight of the net is positive (the condition under
// which the projected-hull rejection of the fast forms is exact).
__global__ void k_prep(const float4* cp, const int* cp_off, int* pq, int np, float4* corners, float* delta) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  const int p = pq[g] & 0xff, q = (pq[g] >> 8) & 0xff, row = q + 1;
  const float4* c = cp + cp_off[g];
  using namespace ex;
  v3<float> c00 = dehom(ld4<float>(c[0])), c10 = dehom(ld4<float>(c[p * row]));
  v3<float> c11 = dehom(ld4<float>(c[p * row + q])), c01 = dehom(ld4<float>(c[q]));
  corners[3 * g + 0] = make_float4(c00.x, c00.y, c00.z, c10.x);
  corners[3 * g + 1] = make_float4(c10.y, c10.z, c11.x, c11.y);
  corners[3 * g + 2] = make_float4(c11.z, c01.x, c01.y, c01.z);
  float d = 0.0f;
  for (int i = 0; i <= p; ++i)
    for (int j = 0; j <= q; ++j) {
      float u = (float)i / p, v = (float)j / q;
      float w00 = (1 - u) * (1 - v), w10 = u * (1 - v), w01 = (1 - u) * v, w11 = u * v;
      float Bx = w00 * c00.x + w10 * c10.x + w01 * c01.x + w11 * c11.x;
      float By = w00 * c00.y + w10 * c10.y + w01 * c01.y + w11 * c11.y;
      float Bz = w00 * c00.z + w10 * c10.z + w01 * c01.z + w11 * c11.z;
      v3<float> P = dehom(ld4<float>(c[i * row + j]));
      float dx = P.x - Bx, dy = P.y - By, dz = P.z - Bz;
      d = fmaxf(d, sqrtf(dx * dx + dy * dy + dz * dz));
    }
  delta[g] = d;
  bool pos = true;
  for (int i = 0; i < (p + 1) * (q + 1); i++) pos = pos && c[i].w > 0.0f;
  pq[g] = (pq[g] & 0xffff) | (pos ? 0x10000 : 0);
}

// Oriented bounding box of each patch's dehomogenised control points (they
// bound the surface when the weights are positive): the normal axis w from the
// cross product of the corner diagonals, u along the mean of the two u-edges
// made perpendicular to w, v = w x u (the coordinate axes when a patch's
// corners are degenerate).  Stored as (u, e_u), (v, e_v), (w, e_w), (centre, 0)
// with half extents widened by a relative slack that covers the FP32 rounding
// of this projection and of the ray's transform into the box frame.
__global__ void k_obb(const float4* cp, const int* cp_off, const int* pq, int np, float4* obb) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  const int p = pq[g] & 0xff, q = (pq[g] >> 8) & 0xff, row = q + 1, n = (p + 1) * row;
  const float4* c = cp + cp_off[g];
  auto dh = [](float4 a) { return make_float3(a.x / a.w, a.y / a.w, a.z / a.w); };
  auto sub = [](float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); };
  auto dot = [](float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
  auto cross = [](float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
  };
  const float3 q00 = dh(c[0]), q10 = dh(c[p * row]), q11 = dh(c[p * row + q]), q01 = dh(c[q]);
  float3 w = cross(sub(q11, q00), sub(q01, q10));
  float wl = sqrtf(dot(w, w));
  float3 ax[3] = {make_float3(1.f, 0.f, 0.f), make_float3(0.f, 1.f, 0.f), make_float3(0.f, 0.f, 1.f)};
  if (wl > 1e-30f && wl < INFINITY) {
    w = make_float3(w.x / wl, w.y / wl, w.z / wl);
    float3 u = make_float3(q10.x - q00.x + q11.x - q01.x, q10.y - q00.y + q11.y - q01.y, q10.z - q00.z + q11.z - q01.z);
    const float uw = dot(u, w);
    u = make_float3(u.x - w.x * uw, u.y - w.y * uw, u.z - w.z * uw);
    float ul = sqrtf(dot(u, u));
    if (!(ul > 1e-30f && ul < INFINITY)) {  // any unit vector perpendicular to w
      u = fabsf(w.x) < 0.9f ? make_float3(0.f, w.z, -w.y) : make_float3(-w.z, 0.f, w.x);
      ul = sqrtf(dot(u, u));
    }
    u = make_float3(u.x / ul, u.y / ul, u.z / ul);
    ax[0] = u;
    ax[1] = cross(w, u);
    ax[2] = w;
  }
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float mag = 0.0f;
  for (int i = 0; i < n; i++) {
    const float3 x = dh(c[i]);
    mag = fmaxf(mag, fabsf(x.x) + fabsf(x.y) + fabsf(x.z));
    for (int k = 0; k < 3; k++) {
      const float y = dot(ax[k], x);
      lo[k] = fminf(lo[k], y);
      hi[k] = fmaxf(hi[k], y);
    }
  }
  float3 ctr = make_float3(0.f, 0.f, 0.f);
  float e[3];
  for (int k = 0; k < 3; k++) {
    const float m = 0.5f * (lo[k] + hi[k]);
    ctr = make_float3(ctr.x + ax[k].x * m, ctr.y + ax[k].y * m, ctr.z + ax[k].z * m);
    e[k] = 0.5f * (hi[k] - lo[k]) + 4e-6f * mag + 1e-6f;
  }
  for (int k = 0; k < 3; k++) obb[4 * g + k] = make_float4(ax[k].x, ax[k].y, ax[k].z, e[k]);
  obb[4 * g + 3] = make_float4(ctr.x, ctr.y, ctr.z, 0.0f);
}

// Gather an array of T by permutation: out[i] = in[perm[i]] (n items of `w` elements).
template <class T>
__global__ void k_gather(const T* in, const uint32_t* perm, int n, int w, T* out) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n * w) return;
  int i = (int)(g / w), k = (int)(g % w);
  out[g] = in[(int64_t)perm[i] * w + k];
}

// ------------------------------------------------------------ SAH kernels
__device__ __forceinline__ float area6(float lx, float ly, float lz, float hx, float hy, float hz) {
  float ex = hx - lx, ey = hy - ly, ez = hz - lz;  // bound_area (bounds_ops.h:72-77)
  return ex * ey + ey * ez + ez * ex;
}

struct SahTask { int first, count, node, pad; };

constexpr int SAH_BLOCK = 64;

// evaluateSAH (bvh.h:47-69) for every candidate (axis, i) of every task:
// block = (task, chunk of 64 candidates); each lane owns one candidate and
// sweeps the node's primitives staged through LDS.  The minimum of
// (cost bits, candidate index) per task equals the reference's first strict
// minimum over axis-major, primitive-minor candidate order.
__global__ void __launch_bounds__(SAH_BLOCK) k_sah_eval(const SahTask* tasks, const int2* blocks, const uint32_t* idx,
                                                        const float* pb, const float4* cent,
                                                        unsigned long long* best) {
  __shared__ float s_b[SAH_BLOCK][6];
  __shared__ float s_c[SAH_BLOCK][3];
  const int2 bt = blocks[blockIdx.x];
  const SahTask t = tasks[bt.x];
  const int cand = bt.y * SAH_BLOCK + threadIdx.x;
  const bool active = cand < 3 * t.count;
  int axis = 0, ci = 0;
  float pos = 0.f;
  if (active) {
    axis = cand / t.count;
    ci = cand % t.count;
    float4 c = cent[idx[t.first + ci]];
    pos = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
  }
  float l0 = INFINITY, l1 = INFINITY, l2 = INFINITY, l3 = -INFINITY, l4 = -INFINITY, l5 = -INFINITY;
  float r0 = INFINITY, r1 = INFINITY, r2 = INFINITY, r3 = -INFINITY, r4 = -INFINITY, r5 = -INFINITY;
  int lc = 0, rc = 0;
  for (int base = 0; base < t.count; base += SAH_BLOCK) {
    const int j = base + threadIdx.x;
    if (j < t.count) {
      const uint32_t k = idx[t.first + j];
#pragma unroll
      for (int e = 0; e < 6; e++) s_b[threadIdx.x][e] = pb[6 * (int64_t)k + e];
      float4 c = cent[k];
      s_c[threadIdx.x][0] = c.x; s_c[threadIdx.x][1] = c.y; s_c[threadIdx.x][2] = c.z;
    }
    __syncthreads();
    const int lim = min(SAH_BLOCK, t.count - base);
    if (active) {
      for (int jj = 0; jj < lim; jj++) {
        const float cv = s_c[jj][axis];
        const float* b = s_b[jj];
        if (cv < pos) {
          lc++;
          l0 = fminf(l0, b[0]); l1 = fminf(l1, b[1]); l2 = fminf(l2, b[2]);
          l3 = fmaxf(l3, b[3]); l4 = fmaxf(l4, b[4]); l5 = fmaxf(l5, b[5]);
        } else {
          rc++;
          r0 = fminf(r0, b[0]); r1 = fminf(r1, b[1]); r2 = fminf(r2, b[2]);
          r3 = fmaxf(r3, b[3]); r4 = fmaxf(r4, b[4]); r5 = fmaxf(r5, b[5]);
        }
      }
    }
    __syncthreads();
  }
  unsigned long long key = ~0ull;
  if (active) {
    float cost = (float)lc * area6(l0, l1, l2, l3, l4, l5) + (float)rc * area6(r0, r1, r2, r3, r4, r5);
    cost = cost > 0 ? cost : 1e30f;
    key = ((unsigned long long)__float_as_uint(cost) << 32) | (unsigned)cand;
  }
  // wave-wide minimum (64 lanes = one block), then one atomic per block
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(key, off);
    key = o < key ? o : key;
  }
  if (threadIdx.x == 0 && key != ~0ull) atomicMin(best + bt.x, key);
}

// subdivide's partition (bvh.h:103-125): one thread per task runs the
// reference's Hoare-style swap loop on idx[first, first+count).
__global__ void k_partition(const SahTask* tasks, int nt, const unsigned long long* best, uint32_t* idx,
                            const float4* cent, int* left_count) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const SahTask tk = tasks[t];
  int axis = 1;
  float pos = 0.0f;
  const unsigned long long key = best[t];
  if (key != ~0ull && __uint_as_float((unsigned)(key >> 32)) < 1e30f) {
    const int cand = (int)(key & 0xffffffffu);
    axis = cand / tk.count;
    const float4 c = cent[idx[tk.first + cand % tk.count]];
    pos = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
  }
  int i = tk.first, j = i + tk.count - 1;
  while (i <= j) {
    const float4 c = cent[idx[i]];
    const float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
    if (cv < pos) {
      i++;
    } else {
      uint32_t tmp = idx[i];
      idx[i] = idx[j];
      idx[j] = tmp;
      j--;
    }
  }
  left_count[t] = i - tk.first;
}

// centroid (bounds_ops.h:27-33)
__global__ void k_centroids(const float* pb, int n, float4* cent) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float* b = pb + 6 * (int64_t)g;
  cent[g] = make_float4(.5f * (b[0] + b[3]), .5f * (b[1] + b[4]), .5f * (b[2] + b[5]), 0.f);
}

// updateNodeBounds (bvh.h:72-82): union over the node's primitive range; one
// wave per node.  Min/max are exact, so the lane split does not change bits.
__global__ void __launch_bounds__(64) k_node_bounds(const int2* ranges, int nn, const uint32_t* idx, const float* pb,
                                                    float* out) {
  const int nd = blockIdx.x;
  if (nd >= nn) return;
  const int2 r = ranges[nd];
  float b0 = INFINITY, b1 = INFINITY, b2 = INFINITY, b3 = -INFINITY, b4 = -INFINITY, b5 = -INFINITY;
  for (int i = threadIdx.x; i < r.y; i += 64) {
    const float* b = pb + 6 * (int64_t)idx[r.x + i];
    b0 = fminf(b0, b[0]); b1 = fminf(b1, b[1]); b2 = fminf(b2, b[2]);
    b3 = fmaxf(b3, b[3]); b4 = fmaxf(b4, b[4]); b5 = fmaxf(b5, b[5]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    b0 = fminf(b0, __shfl_xor(b0, off)); b1 = fminf(b1, __shfl_xor(b1, off)); b2 = fminf(b2, __shfl_xor(b2, off));
    b3 = fmaxf(b3, __shfl_xor(b3, off)); b4 = fmaxf(b4, __shfl_xor(b4, off)); b5 = fmaxf(b5, __shfl_xor(b5, off));
  }
  if (threadIdx.x == 0) {
    float* o = out + 6 * (int64_t)nd;
    o[0] = b0; o[1] = b1; o[2] = b2; o[3] = b3; o[4] = b4; o[5] = b5;
  }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace rtn {

// Exhaustive-SAH build of the reference (bvh.h:148-179) over device prim
// bounds pb[6n].  Returns reference-layout nodes (depth-first numbering) and
// the final primitive order perm (device).
int build_sah_device(const float* d_pb, int n, std::vector<rtn_bvh_node>& nodes_out, DevBuf<uint32_t>& perm,
                     int64_t* max_depth, int64_t* n_leaves) {
  RTN_CHECK(n >= 1, RTN_EINVAL, "build_bvh: no primitives");
  DevBuf<float4> cent;
  RTN_TRY(cent.alloc(n));
  hipLaunchKernelGGL(k_centroids, dim3(nblk(n, 256)), dim3(256), 0, nullptr, d_pb, n, cent.ptr);
  std::vector<uint32_t> idx0(n);
  std::iota(idx0.begin(), idx0.end(), 0u);
  RTN_TRY(perm.upload(idx0.data(), n));

  // BFS tree mirror on the host: (first, count, left child or -1)
  struct BNode { int first, count, left, depth; };
  std::vector<BNode> bt;
  bt.reserve(2 * (size_t)n);
  bt.push_back({0, n, -1, 0});
  std::vector<int> level = {};
  if (n >= 2) level.push_back(0);
  DevBuf<SahTask> d_tasks;
  DevBuf<int2> d_blocks;
  DevBuf<unsigned long long> d_best;
  DevBuf<int> d_left;
  std::vector<SahTask> tasks;
  std::vector<int2> blocks;
  std::vector<int> left;
  int64_t depth = 0;
  while (!level.empty()) {
    tasks.clear();
    blocks.clear();
    for (size_t k = 0; k < level.size(); k++) {
      const BNode& b = bt[level[k]];
      tasks.push_back({b.first, b.count, level[k], 0});
      const int nb = (3 * b.count + SAH_BLOCK - 1) / SAH_BLOCK;
      for (int c = 0; c < nb; c++) blocks.push_back(make_int2((int)k, c));
    }
    const int nt = (int)tasks.size();
    RTN_TRY(d_tasks.upload(tasks.data(), nt));
    RTN_TRY(d_blocks.upload(blocks.data(), blocks.size()));
    RTN_TRY(d_best.alloc(nt));
    RTN_HIP(hipMemset(d_best.ptr, 0xff, nt * sizeof(unsigned long long)));
    RTN_TRY(d_left.alloc(nt));
    hipLaunchKernelGGL(k_sah_eval, dim3((unsigned)blocks.size()), dim3(SAH_BLOCK), 0, nullptr, d_tasks.ptr,
                       d_blocks.ptr, perm.ptr, d_pb, cent.ptr, d_best.ptr);
    RTN_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_partition, dim3(nblk(nt, 64)), dim3(64), 0, nullptr, d_tasks.ptr, nt, d_best.ptr, perm.ptr,
                       cent.ptr, d_left.ptr);
    RTN_HIP(hipGetLastError());
    left.resize(nt);
    RTN_TRY(d_left.download(left.data(), nt));
    std::vector<int> next;
    for (int k = 0; k < nt; k++) {
      const int id = level[k];
      const int f = bt[id].first, c = bt[id].count, lc = left[k];
      if (lc == 0 || lc == c) continue;  // leaf (bvh.h:125)
      const int l = (int)bt.size();
      bt[id].left = l;
      const int d = bt[id].depth + 1;
      depth = std::max<int64_t>(depth, d);
      bt.push_back({f, lc, -1, d});
      bt.push_back({f + lc, c - lc, -1, d});
      if (lc >= 2) next.push_back(l);
      if (c - lc >= 2) next.push_back(l + 1);
    }
    level.swap(next);
  }
  // node bounds on the device, BFS order
  const int nn = (int)bt.size();
  std::vector<int2> ranges(nn);
  for (int i = 0; i < nn; i++) ranges[i] = make_int2(bt[i].first, bt[i].count);
  DevBuf<int2> d_ranges;
  DevBuf<float> d_nb;
  RTN_TRY(d_ranges.upload(ranges.data(), nn));
  RTN_TRY(d_nb.alloc(6 * (size_t)nn));
  hipLaunchKernelGGL(k_node_bounds, dim3(nn), dim3(64), 0, nullptr, d_ranges.ptr, nn, perm.ptr, d_pb, d_nb.ptr);
  RTN_HIP(hipGetLastError());
  std::vector<float> nbh(6 * (size_t)nn);
  RTN_TRY(d_nb.download(nbh.data(), nbh.size()));
  // depth-first numbering (bvh.h:127-142): a node's two children take the
  // next two indices when it is split; the left subtree is then numbered
  // completely before the right one.
  nodes_out.assign(nn, rtn_bvh_node{});
  std::vector<int> dfs(nn, -1);
  dfs[0] = 0;
  int counter = 1;
  std::vector<int> stack = {0};
  int64_t leaves = 0;
  while (!stack.empty()) {
    const int id = stack.back();
    stack.pop_back();
    const BNode& b = bt[id];
    rtn_bvh_node& o = nodes_out[dfs[id]];
    const float* bb = &nbh[6 * (size_t)id];
    o.p_min_x = bb[0]; o.p_min_y = bb[1]; o.p_min_z = bb[2];
    o.p_max_x = bb[3]; o.p_max_y = bb[4]; o.p_max_z = bb[5];
    if (b.left < 0) {
      o.left_first = (uint32_t)b.first;
      o.prim_count = (uint32_t)b.count;
      leaves++;
    } else {
      dfs[b.left] = counter;
      dfs[b.left + 1] = counter + 1;
      counter += 2;
      o.left_first = (uint32_t)dfs[b.left];
      o.prim_count = 0;
      stack.push_back(b.left + 1);  // processed after the left subtree
      stack.push_back(b.left);
    }
  }
  if (max_depth) *max_depth = depth;
  if (n_leaves) *n_leaves = leaves;
  return RTN_OK;
}

// Binned-SAH binary BVH on the host, for scenes too large for the exhaustive
// split search (which is quadratic in the primitives of a node).  Per node:
// centroid bounds, `bins` equal bins on each axis, the cheapest plane by the
// surface-area heuristic (half areas, as bounds_ops.h:72-77), in-place
// partition of the primitive order.  A node becomes a leaf when it holds one
// primitive, or at most max_leaf primitives and no split beats testing them
// all; a node whose centroids coincide is split at its median.  Output uses
// the reference node layout (children adjacent, left_first = left child).
int build_binned_host(const std::vector<float>& pb, int n, int bins, int max_leaf, std::vector<rtn_bvh_node>& nodes,
                      std::vector<uint32_t>& order, int64_t* max_depth, int64_t* n_leaves) {
  bins = std::max(2, std::min(bins, 256));
  max_leaf = std::max(1, std::min(max_leaf, 128));
  order.resize(n);
  std::vector<float> cen(3 * (size_t)n);
  for (int i = 0; i < n; i++) {
    order[i] = (uint32_t)i;
    for (int a = 0; a < 3; a++) cen[3 * (size_t)i + a] = 0.5f * (pb[6 * (size_t)i + a] + pb[6 * (size_t)i + 3 + a]);
  }
  nodes.clear();
  nodes.reserve(2 * (size_t)n);
  nodes.push_back(rtn_bvh_node{});
  struct Task { int node, begin, end, depth; };
  std::vector<Task> stack{{0, 0, n, 0}};
  int64_t depth = 0, leaves = 0;
  auto half_area = [](const float* b) {
    const float ex = b[3] - b[0], ey = b[4] - b[1], ez = b[5] - b[2];
    return ex * ey + ey * ez + ez * ex;
  };
  std::vector<float> bb(6 * bins), rb(6 * bins);
  std::vector<int> bc(bins);
  while (!stack.empty()) {
    const Task t = stack.back();
    stack.pop_back();
    depth = std::max<int64_t>(depth, t.depth);
    float box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    float cb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int k = t.begin; k < t.end; k++) {
      const size_t i = order[k];
      for (int a = 0; a < 3; a++) {
        box[a] = std::min(box[a], pb[6 * i + a]);
        box[3 + a] = std::max(box[3 + a], pb[6 * i + 3 + a]);
        cb[a] = std::min(cb[a], cen[3 * i + a]);
        cb[3 + a] = std::max(cb[3 + a], cen[3 * i + a]);
      }
    }
    rtn_bvh_node& nd = nodes[t.node];
    nd.p_min_x = box[0]; nd.p_min_y = box[1]; nd.p_min_z = box[2];
    nd.p_max_x = box[3]; nd.p_max_y = box[4]; nd.p_max_z = box[5];
    const int cnt = t.end - t.begin;
    auto make_leaf = [&]() {
      nodes[t.node].left_first = (uint32_t)t.begin;
      nodes[t.node].prim_count = (uint32_t)cnt;
      leaves++;
    };
    if (cnt == 1) { make_leaf(); continue; }
    // best binned split over the three axes
    float best_cost = INFINITY;
    int best_axis = -1, best_bin = 0;
    for (int a = 0; a < 3; a++) {
      const float lo = cb[a], ext = cb[3 + a] - cb[a];
      if (!(ext > 0.f)) continue;
      const float scale = (float)bins / ext;
      std::fill(bc.begin(), bc.end(), 0);
      for (int b = 0; b < bins; b++) {
        bb[6 * b] = bb[6 * b + 1] = bb[6 * b + 2] = INFINITY;
        bb[6 * b + 3] = bb[6 * b + 4] = bb[6 * b + 5] = -INFINITY;
      }
      for (int k = t.begin; k < t.end; k++) {
        const size_t i = order[k];
        const int b = std::min(bins - 1, (int)((cen[3 * i + a] - lo) * scale));
        bc[b]++;
        for (int c = 0; c < 3; c++) {
          bb[6 * b + c] = std::min(bb[6 * b + c], pb[6 * i + c]);
          bb[6 * b + 3 + c] = std::max(bb[6 * b + 3 + c], pb[6 * i + 3 + c]);
        }
      }
      // right-to-left sweep of accumulated boxes
      float acc[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      std::vector<int> rc(bins);
      int rn = 0;
      for (int b = bins - 1; b >= 1; b--) {
        for (int c = 0; c < 3; c++) {
          acc[c] = std::min(acc[c], bb[6 * b + c]);
          acc[3 + c] = std::max(acc[3 + c], bb[6 * b + 3 + c]);
        }
        rn += bc[b];
        rc[b] = rn;
        for (int c = 0; c < 6; c++) rb[6 * b + c] = acc[c];
      }
      float lacc[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      int ln = 0;
      for (int b = 0; b < bins - 1; b++) {  // plane between bin b and b+1
        for (int c = 0; c < 3; c++) {
          lacc[c] = std::min(lacc[c], bb[6 * b + c]);
          lacc[3 + c] = std::max(lacc[3 + c], bb[6 * b + 3 + c]);
        }
        ln += bc[b];
        const int rn2 = rc[b + 1];
        if (ln == 0 || rn2 == 0) continue;
        const float cost = ln * half_area(lacc) + rn2 * half_area(&rb[6 * (b + 1)]);
        if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = b; }
      }
    }
    int mid;
    if (best_axis < 0) {  // coincident centroids: median split
      if (cnt <= max_leaf) { make_leaf(); continue; }
      mid = t.begin + cnt / 2;
    } else {
      if (cnt <= max_leaf && best_cost >= cnt * half_area(box)) { make_leaf(); continue; }
      const float lo = cb[best_axis], scale = (float)bins / (cb[3 + best_axis] - cb[best_axis]);
      auto it = std::partition(order.begin() + t.begin, order.begin() + t.end, [&](uint32_t i) {
        return std::min(bins - 1, (int)((cen[3 * (size_t)i + best_axis] - lo) * scale)) <= best_bin;
      });
      mid = (int)(it - order.begin());
      if (mid == t.begin || mid == t.end) mid = t.begin + cnt / 2;
    }
    const int left = (int)nodes.size();
    nodes.push_back(rtn_bvh_node{});
    nodes.push_back(rtn_bvh_node{});
    nodes[t.node].left_first = (uint32_t)left;
    nodes[t.node].prim_count = 0;
    stack.push_back({left + 1, mid, t.end, t.depth + 1});
    stack.push_back({left, t.begin, mid, t.depth + 1});
  }
  if (max_depth) *max_depth = depth;
  if (n_leaves) *n_leaves = leaves;
  return RTN_OK;
}

// Collapse of the binary BVH into 4-wide nodes: each wide node takes a binary
// interior node and repeatedly opens its largest-area interior child until it
// has four children.  Leaves keep the binary leaf's primitive range.
int collapse_wide(const std::vector<rtn_bvh_node>& bn, std::vector<WideNode>& out) {
  out.clear();
  if (bn.empty()) return RTN_OK;
  auto area = [&](int i) {
    const rtn_bvh_node& b = bn[i];
    float ex = b.p_max_x - b.p_min_x, ey = b.p_max_y - b.p_min_y, ez = b.p_max_z - b.p_min_z;
    return ex * ey + ey * ez + ez * ex;
  };
  auto leaf_code = [&](int i, int32_t& code) -> int {
    const rtn_bvh_node& b = bn[i];
    RTN_CHECK(b.prim_count >= 1 && b.prim_count <= 128 && b.left_first < (1u << 24), RTN_EUNSUPPORTED,
              "wide BVH: leaf with %u primitives at %u outside the packed range", b.prim_count, b.left_first);
    code = (int32_t)(0x80000000u | ((b.prim_count - 1) << 24) | b.left_first);
    return RTN_OK;
  };
  // a root that is itself a leaf becomes a wide node with one child
  std::vector<std::pair<int, int>> work;  // (binary node, wide index)
  out.push_back(WideNode{});
  work.push_back({0, 0});
  while (!work.empty()) {
    auto [bi, wi] = work.back();
    work.pop_back();
    int kids[4];
    int nk = 0;
    if (bn[bi].prim_count > 0) {
      kids[nk++] = bi;
    } else {
      kids[nk++] = (int)bn[bi].left_first;
      kids[nk++] = (int)bn[bi].left_first + 1;
      while (nk < 4) {
        int best = -1;
        float ba = -1.f;
        for (int k = 0; k < nk; k++)
          if (bn[kids[k]].prim_count == 0 && area(kids[k]) > ba) { ba = area(kids[k]); best = k; }
        if (best < 0) break;
        int c = kids[best];
        kids[best] = (int)bn[c].left_first;
        kids[nk++] = (int)bn[c].left_first + 1;
      }
    }
    WideNode w;
    for (int k = 0; k < 4; k++) {
      if (k < nk) {
        const rtn_bvh_node& b = bn[kids[k]];
        w.lo_x[k] = b.p_min_x; w.lo_y[k] = b.p_min_y; w.lo_z[k] = b.p_min_z;
        w.hi_x[k] = b.p_max_x; w.hi_y[k] = b.p_max_y; w.hi_z[k] = b.p_max_z;
        if (b.prim_count > 0) {
          RTN_TRY(leaf_code(kids[k], w.child[k]));
        } else {
          const int ni = (int)out.size();
          out.push_back(WideNode{});
          w.child[k] = ni;
          work.push_back({kids[k], ni});
        }
      } else {  // empty slot: an inverted box that no ray enters
        w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = INFINITY;
        w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = -INFINITY;
        w.child[k] = (int32_t)0x7fffffff;
      }
      w.pad[k] = 0;
    }
    out[wi] = w;
  }
  return RTN_OK;
}

}  // namespace rtn

// ------------------------------------------------------------------- C ABI
extern "C" int rtn_scene_create(rtn_scene** scene) {
  RTN_CHECK(scene, RTN_EINVAL, "rtn_scene_create: null out pointer");
  *scene = new rtn_scene();
  return RTN_OK;
}

extern "C" int rtn_scene_destroy(rtn_scene* scene) {
  delete scene;
  return RTN_OK;
}

extern "C" int rtn_scene_add_nurbs(rtn_scene* s, int n, int p, const float* U, int m, int q, const float* V,
                                   const float* Pw) {
  RTN_CHECK(s && U && V && Pw, RTN_EINVAL, "rtn_scene_add_nurbs: null argument");
  RTN_CHECK(!s->built, RTN_EINVAL, "rtn_scene_add_nurbs: scene already built");
  RTN_CHECK(p >= 1 && q >= 1, RTN_EINVAL, "rtn_scene_add_nurbs: degree must be >= 1");
  RTN_CHECK(p <= RTN_MAX_DEGREE && q <= RTN_MAX_DEGREE, RTN_EUNSUPPORTED,
            "rtn_scene_add_nurbs: degree (%d,%d) above %d", p, q, RTN_MAX_DEGREE);
  RTN_CHECK(n >= p && m >= q, RTN_EINVAL, "rtn_scene_add_nurbs: need n >= p and m >= q");
  // interior multiplicity must not exceed the degree (A5.7 reads P[b-p+i])
  for (int dir = 0; dir < 2; dir++) {
    const float* K = dir ? V : U;
    const int nn = dir ? m : n, dd = dir ? q : p;
    for (int i = 1; i < nn + dd + 2; i++)
      RTN_CHECK(K[i] >= K[i - 1], RTN_EINVAL, "rtn_scene_add_nurbs: knots not sorted");
    // clamped ends and interior knots strictly inside: then the segment count
    // of segs() equals the number of segments A5.7 writes
    for (int i = 0; i <= dd; i++)
      RTN_CHECK(K[i] == K[0] && K[nn + 1 + i] == K[nn + 1], RTN_EINVAL, "rtn_scene_add_nurbs: knots not clamped");
    for (int i = dd + 1; i <= nn; i++)
      RTN_CHECK(K[i] > K[dd] && K[i] < K[nn + 1], RTN_EINVAL, "rtn_scene_add_nurbs: interior knot on an end value");
    for (int i = dd + 1; i <= nn;) {
      int j = i;
      while (j + 1 <= nn && K[j + 1] == K[i]) j++;
      RTN_CHECK(j - i + 1 <= dd, RTN_EINVAL, "rtn_scene_add_nurbs: interior multiplicity %d above degree %d",
                j - i + 1, dd);
      i = j + 1;
    }
    RTN_CHECK(K[dd] < K[nn + 1], RTN_EINVAL, "rtn_scene_add_nurbs: empty parameter range");
  }
  HostSurface hs;
  hs.n = n; hs.p = p; hs.m = m; hs.q = q;
  hs.U.assign(U, U + n + p + 2);
  hs.V.assign(V, V + m + q + 2);
  hs.Pw.assign(Pw, Pw + 4 * (size_t)(n + 1) * (m + 1));
  s->surfaces.push_back(std::move(hs));
  return RTN_OK;
}

extern "C" int rtn_scene_add_patches(rtn_scene* s, int np, int p, int q, const float* cp) {
  RTN_CHECK(s && cp && np > 0, RTN_EINVAL, "rtn_scene_add_patches: bad arguments");
  RTN_CHECK(!s->built, RTN_EINVAL, "rtn_scene_add_patches: scene already built");
  RTN_CHECK(p >= 1 && q >= 1 && p <= RTN_MAX_DEGREE && q <= RTN_MAX_DEGREE, RTN_EUNSUPPORTED,
            "rtn_scene_add_patches: degree (%d,%d) outside 1..%d", p, q, RTN_MAX_DEGREE);
  rtn_scene::PatchGroup g{np, p, q, std::vector<float>(cp, cp + 4 * (size_t)np * (p + 1) * (q + 1))};
  s->groups.push_back(std::move(g));
  return RTN_OK;
}

namespace rtn {
int scene_finish(rtn_scene* s);  // defined below
}

extern "C" void rtn_build_params_default(rtn_build_params* bp) {
  if (!bp) return;
  memset(bp, 0, sizeof(*bp));
  bp->bvh = RTN_BVH_EXACT_SAH;
  bp->bins = 32;
  bp->max_leaf = 4;
}

extern "C" int rtn_scene_build(rtn_scene* s) {
  rtn_build_params bp;
  rtn_build_params_default(&bp);
  return rtn_scene_build_ex(s, &bp);
}

extern "C" int rtn_scene_build_ex(rtn_scene* s, const rtn_build_params* bp) {
  RTN_CHECK(s && bp, RTN_EINVAL, "rtn_scene_build: null argument");
  RTN_CHECK(bp->bvh == RTN_BVH_EXACT_SAH || bp->bvh == RTN_BVH_BINNED_SAH, RTN_EINVAL,
            "rtn_scene_build: unknown BVH builder %d", bp->bvh);
  RTN_CHECK(!s->built, RTN_EINVAL, "rtn_scene_build: scene already built");
  RTN_CHECK(!s->surfaces.empty() || !s->groups.empty(), RTN_EINVAL, "rtn_scene_build: empty scene");
  auto t0 = std::chrono::steady_clock::now();
  const int ns = (int)s->surfaces.size();
  // ---- decomposition of the NURBS surfaces
  std::vector<SurfDesc> sd(ns);
  std::vector<float> knots;
  std::vector<float4> pw;
  for (int i = 0; i < ns; i++) {
    const HostSurface& h = s->surfaces[i];
    SurfDesc& d = sd[i];
    d.n = h.n; d.p = h.p; d.m = h.m; d.q = h.q;
    d.u_off = (int)knots.size();
    knots.insert(knots.end(), h.U.begin(), h.U.end());
    d.v_off = (int)knots.size();
    knots.insert(knots.end(), h.V.begin(), h.V.end());
    d.pw_off = (int64_t)pw.size();
    for (size_t k = 0; k < h.Pw.size() / 4; k++)
      pw.push_back(make_float4(h.Pw[4 * k], h.Pw[4 * k + 1], h.Pw[4 * k + 2], h.Pw[4 * k + 3]));
  }
  DevBuf<SurfDesc> d_sd;
  DevBuf<float> d_knots, d_brk;
  DevBuf<float4> d_pw, d_strips, d_cp;
  DevBuf<int> d_span_u, d_span_v;
  int64_t total_cp = 0, total_patches = 0;
  std::vector<int> group_base, group_cp;
  if (ns > 0) {
    RTN_TRY(d_sd.upload(sd.data(), ns));
    RTN_TRY(d_knots.upload(knots.data(), knots.size()));
    RTN_TRY(d_brk.alloc(knots.size()));
    RTN_TRY(d_pw.upload(pw.data(), pw.size()));
    hipLaunchKernelGGL(k_count, dim3(nblk(ns, 64)), dim3(64), 0, nullptr, d_sd.ptr, ns, d_knots.ptr, d_brk.ptr);
    RTN_HIP(hipGetLastError());
    RTN_TRY(d_sd.download(sd.data(), ns));
    int64_t strips = 0;
    for (int i = 0; i < ns; i++) {
      SurfDesc& d = sd[i];
      d.strip_off = strips;
      strips += (int64_t)d.nbU * (d.p + 1) * (d.m + 1);
      d.patch_cp_off = total_cp;
      d.patch_base = (int)total_patches;
      total_cp += (int64_t)d.nbU * d.nbV * (d.p + 1) * (d.q + 1);
      total_patches += (int64_t)d.nbU * d.nbV;
    }
    RTN_TRY(d_sd.upload(sd.data(), ns));
    RTN_TRY(d_strips.alloc(strips));
    RTN_TRY(d_span_u.alloc(knots.size()));
    RTN_TRY(d_span_v.alloc(knots.size()));
  }
  // pre-decomposed groups follow the surfaces
  for (auto& g : s->groups) {
    group_base.push_back((int)total_patches);
    group_cp.push_back((int)total_cp);
    total_patches += g.np;
    total_cp += (int64_t)g.np * (g.p + 1) * (g.q + 1);
  }
  RTN_CHECK(total_cp < (int64_t)1 << 31 && total_patches < (1 << 24), RTN_EUNSUPPORTED,
            "rtn_scene_build: scene too large (%lld patches)", (long long)total_patches);
  const int np = (int)total_patches;
  RTN_TRY(d_cp.alloc(total_cp));
  DevBuf<int> o_cp_off, o_pq, o_pid;
  DevBuf<float4> o_uv;
  DevBuf<int2> o_kidx;
  RTN_TRY(o_cp_off.alloc(np));
  RTN_TRY(o_pq.alloc(np));
  RTN_TRY(o_pid.alloc(np));
  RTN_TRY(o_uv.alloc(np));
  RTN_TRY(o_kidx.alloc(np));
  if (ns > 0) {
    hipLaunchKernelGGL(k_decomp_u, dim3(ns), dim3(256), 0, nullptr, d_sd.ptr, d_knots.ptr, d_pw.ptr, d_strips.ptr,
                       d_span_u.ptr);
    hipLaunchKernelGGL(k_decomp_v, dim3(ns), dim3(256), 0, nullptr, d_sd.ptr, d_knots.ptr, d_strips.ptr, d_cp.ptr,
                       d_span_v.ptr);
    const int nps = sd.back().patch_base + sd.back().nbU * sd.back().nbV;
    hipLaunchKernelGGL(k_headers, dim3(nblk(nps, 256)), dim3(256), 0, nullptr, d_sd.ptr, ns, nps, d_brk.ptr,
                       d_span_u.ptr, d_span_v.ptr, o_cp_off.ptr, o_pq.ptr, o_uv.ptr, o_pid.ptr, o_kidx.ptr);
    RTN_HIP(hipGetLastError());
  }
  for (size_t gi = 0; gi < s->groups.size(); gi++) {
    auto& g = s->groups[gi];
    const int ps = (g.p + 1) * (g.q + 1);
    RTN_HIP(hipMemcpy(d_cp.ptr + group_cp[gi], g.cp.data(), g.cp.size() * 4, hipMemcpyHostToDevice));
    std::vector<int> off(g.np), pq(g.np, g.p | (g.q << 8)), pid(g.np);
    std::vector<float4> uv(g.np, make_float4(0.f, 1.f, 0.f, 1.f));
    std::vector<int2> kx(g.np, make_int2(-1, -1));
    for (int k = 0; k < g.np; k++) { off[k] = group_cp[gi] + k * ps; pid[k] = group_base[gi] + k; }
    RTN_HIP(hipMemcpy(o_cp_off.ptr + group_base[gi], off.data(), g.np * 4, hipMemcpyHostToDevice));
    RTN_HIP(hipMemcpy(o_pq.ptr + group_base[gi], pq.data(), g.np * 4, hipMemcpyHostToDevice));
    RTN_HIP(hipMemcpy(o_pid.ptr + group_base[gi], pid.data(), g.np * 4, hipMemcpyHostToDevice));
    RTN_HIP(hipMemcpy(o_uv.ptr + group_base[gi], uv.data(), g.np * 16, hipMemcpyHostToDevice));
    RTN_HIP(hipMemcpy(o_kidx.ptr + group_base[gi], kx.data(), g.np * 8, hipMemcpyHostToDevice));
  }
  DevBuf<float> d_pb;
  RTN_TRY(d_pb.alloc(6 * (size_t)np));
  hipLaunchKernelGGL(k_bounds, dim3(nblk(np, 256)), dim3(256), 0, nullptr, d_cp.ptr, o_cp_off.ptr, o_pq.ptr, np,
                     d_pb.ptr);
  RTN_HIP(hipGetLastError());
  RTN_HIP(hipDeviceSynchronize());
  auto t1 = std::chrono::steady_clock::now();
  // ---- BVH (reference exhaustive SAH, or binned SAH), then reorder patches to BVH order
  std::vector<rtn_bvh_node> nodes;
  DevBuf<uint32_t> perm;
  if (bp->bvh == RTN_BVH_EXACT_SAH) {
    RTN_TRY(build_sah_device(d_pb.ptr, np, nodes, perm, &s->max_depth, &s->n_leaves));
  } else {
    std::vector<float> pb_h(6 * (size_t)np);
    RTN_TRY(d_pb.download(pb_h.data(), pb_h.size()));
    std::vector<uint32_t> ord;
    RTN_TRY(build_binned_host(pb_h, np, bp->bins, bp->max_leaf, nodes, ord, &s->max_depth, &s->n_leaves));
    RTN_TRY(perm.upload(ord.data(), np));
  }
  // BVH-order patch arrays; nets are repacked contiguously in BVH order
  std::vector<int> pq_h(np);
  std::vector<uint32_t> perm_h(np);
  RTN_TRY(perm.download(perm_h.data(), np));
  RTN_TRY(o_pq.download(pq_h.data(), np));
  std::vector<int> new_off(np);
  std::vector<int64_t> hist((RTN_MAX_DEGREE + 1) * (RTN_MAX_DEGREE + 1), 0);
  int64_t acc = 0;
  for (int i = 0; i < np; i++) {
    const int src = (int)perm_h[i];
    const int p = pq_h[src] & 0xff, q = (pq_h[src] >> 8) & 0xff;
    new_off[i] = (int)acc;
    acc += (p + 1) * (q + 1);
    hist[p * (RTN_MAX_DEGREE + 1) + q]++;
  }
  RTN_TRY(s->cp.alloc(total_cp));
  RTN_TRY(s->cp_off.upload(new_off.data(), np));
  hipLaunchKernelGGL(k_gather_nets, dim3(np), dim3(64), 0, nullptr, d_cp.ptr, o_cp_off.ptr, o_pq.ptr, perm.ptr,
                     s->cp_off.ptr, np, s->cp.ptr);
  RTN_HIP(hipGetLastError());
  RTN_TRY(s->pq.alloc(np));
  RTN_TRY(s->uv.alloc(np));
  RTN_TRY(s->patch_id.alloc(np));
  RTN_TRY(s->knot_idx.alloc(np));
  RTN_TRY(s->bounds.alloc(6 * (size_t)np));
  hipLaunchKernelGGL(k_gather<int>, dim3(nblk(np, 256)), dim3(256), 0, nullptr, o_pq.ptr, perm.ptr, np, 1, s->pq.ptr);
  hipLaunchKernelGGL(k_gather<float4>, dim3(nblk(np, 256)), dim3(256), 0, nullptr, o_uv.ptr, perm.ptr, np, 1,
                     s->uv.ptr);
  hipLaunchKernelGGL(k_gather<int>, dim3(nblk(np, 256)), dim3(256), 0, nullptr, o_pid.ptr, perm.ptr, np, 1,
                     s->patch_id.ptr);
  hipLaunchKernelGGL(k_gather<int2>, dim3(nblk(np, 256)), dim3(256), 0, nullptr, o_kidx.ptr, perm.ptr, np, 1,
                     s->knot_idx.ptr);
  hipLaunchKernelGGL(k_gather<float>, dim3(nblk(6 * (int64_t)np, 256)), dim3(256), 0, nullptr, d_pb.ptr, perm.ptr,
                     np, 6, s->bounds.ptr);
  RTN_HIP(hipGetLastError());
  RTN_TRY(s->corners.alloc(3 * (size_t)np));
  RTN_TRY(s->delta.alloc(np));
  hipLaunchKernelGGL(k_prep, dim3(nblk(np, 256)), dim3(256), 0, nullptr, s->cp.ptr, s->cp_off.ptr, s->pq.ptr, np,
                     s->corners.ptr, s->delta.ptr);
  RTN_HIP(hipGetLastError());
  RTN_TRY(s->seed64.alloc(kSeedTermWords * (size_t)np));
  hipLaunchKernelGGL(k_seed_terms, dim3(nblk(np, 256)), dim3(256), 0, nullptr, s->cp.ptr, s->cp_off.ptr, s->pq.ptr,
                     np, s->seed64.ptr);
  RTN_HIP(hipGetLastError());
  RTN_TRY(s->obb.alloc(4 * (size_t)np));
  hipLaunchKernelGGL(k_obb, dim3(nblk(np, 256)), dim3(256), 0, nullptr, s->cp.ptr, s->cp_off.ptr, s->pq.ptr, np,
                     s->obb.ptr);
  RTN_HIP(hipGetLastError());
  RTN_TRY(s->nodes.upload(nodes.data(), nodes.size()));
  RTN_HIP(hipDeviceSynchronize());
  auto t2 = std::chrono::steady_clock::now();
  std::vector<WideNode> wide;
  RTN_TRY(collapse_wide(nodes, wide));
  RTN_TRY(s->wide.upload(wide.data(), wide.size()));
  auto t3 = std::chrono::steady_clock::now();
  s->n_patches = np;
  s->n_cp = total_cp;
  s->n_nodes = (int64_t)nodes.size();
  s->n_wide = (int64_t)wide.size();
  s->deg_hist = hist;
  s->ms_decomp = std::chrono::duration<double, std::milli>(t1 - t0).count();
  s->ms_bvh = std::chrono::duration<double, std::milli>(t2 - t1).count();
  s->ms_wide = std::chrono::duration<double, std::milli>(t3 - t2).count();
  if (!s->ctl.ptr) RTN_TRY(s->ctl.alloc(1));
  s->built = true;
  return RTN_OK;
}

extern "C" int rtn_scene_stats_get(const rtn_scene* s, rtn_scene_stats* o) {
  RTN_CHECK(s && o, RTN_EINVAL, "rtn_scene_stats_get: null argument");
  memset(o, 0, sizeof(*o));
  o->n_surfaces = (int64_t)s->surfaces.size();
  o->n_patches = s->n_patches;
  o->n_cp = s->n_cp;
  o->n_nodes = s->n_nodes;
  o->n_wide_nodes = s->n_wide;
  o->max_depth = s->max_depth;
  o->n_leaves = s->n_leaves;
  for (int p = 0; p <= RTN_MAX_DEGREE; p++)
    for (int q = 0; q <= RTN_MAX_DEGREE; q++)
      o->deg_hist[p][q] = s->deg_hist.empty() ? 0 : s->deg_hist[p * (RTN_MAX_DEGREE + 1) + q];
  o->build_ms_decompose = s->ms_decomp;
  o->build_ms_bvh = s->ms_bvh;
  o->build_ms_wide = s->ms_wide;
  return RTN_OK;
}

extern "C" int rtn_scene_download(const rtn_scene* s, float* patch_cp, int* patch_off, int* patch_pq,
                                  float* patch_uv, int* patch_id, int* knot_idx, rtn_bvh_node* nodes, float* delta) {
  RTN_CHECK(s && s->built, RTN_EINVAL, "rtn_scene_download: scene not built");
  const size_t np = (size_t)s->n_patches;
  if (patch_cp) RTN_HIP(hipMemcpy(patch_cp, s->cp.ptr, s->n_cp * sizeof(float4), hipMemcpyDeviceToHost));
  if (patch_off) RTN_HIP(hipMemcpy(patch_off, s->cp_off.ptr, np * 4, hipMemcpyDeviceToHost));
  if (patch_pq) {
    std::vector<int> pq(np);
    RTN_HIP(hipMemcpy(pq.data(), s->pq.ptr, np * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < np; i++) { patch_pq[2 * i] = pq[i] & 0xff; patch_pq[2 * i + 1] = (pq[i] >> 8) & 0xff; }
  }
  if (patch_uv) RTN_HIP(hipMemcpy(patch_uv, s->uv.ptr, np * 16, hipMemcpyDeviceToHost));
  if (patch_id) RTN_HIP(hipMemcpy(patch_id, s->patch_id.ptr, np * 4, hipMemcpyDeviceToHost));
  if (knot_idx) RTN_HIP(hipMemcpy(knot_idx, s->knot_idx.ptr, np * 8, hipMemcpyDeviceToHost));
  if (nodes) RTN_HIP(hipMemcpy(nodes, s->nodes.ptr, s->n_nodes * sizeof(rtn_bvh_node), hipMemcpyDeviceToHost));
  if (delta) RTN_HIP(hipMemcpy(delta, s->delta.ptr, np * 4, hipMemcpyDeviceToHost));
  return RTN_OK;
}

// ------------------------------------------------ reference-shaped helpers
extern "C" int rtn_get_patch(int n, int p, const float* U, int m, int q, const float* V, const float* Pw, int* nbU,
                             int* nbV, float* cp, float* uv_range) {
  RTN_CHECK(nbU && nbV, RTN_EINVAL, "rtn_get_patch: null count pointers");
  rtn_scene sc;
  RTN_TRY(rtn_scene_add_nurbs(&sc, n, p, U, m, q, V, Pw));
  // decomposition only: run the count + A5.7 kernels without the BVH
  SurfDesc d{};
  d.n = n; d.p = p; d.m = m; d.q = q;
  d.u_off = 0;
  d.v_off = n + p + 2;
  std::vector<float> knots(U, U + n + p + 2);
  knots.insert(knots.end(), V, V + m + q + 2);
  DevBuf<SurfDesc> d_sd;
  DevBuf<float> d_knots, d_brk;
  RTN_TRY(d_sd.upload(&d, 1));
  RTN_TRY(d_knots.upload(knots.data(), knots.size()));
  RTN_TRY(d_brk.alloc(knots.size()));
  hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, nullptr, d_sd.ptr, 1, d_knots.ptr, d_brk.ptr);
  RTN_HIP(hipGetLastError());
  RTN_TRY(d_sd.download(&d, 1));
  *nbU = d.nbU;
  *nbV = d.nbV;
  if (!cp && !uv_range) return RTN_OK;
  const int64_t nps = (int64_t)d.nbU * d.nbV, ps = (int64_t)(p + 1) * (q + 1);
  d.strip_off = 0;
  d.patch_cp_off = 0;
  d.patch_base = 0;
  std::vector<float4> pw((size_t)(n + 1) * (m + 1));
  memcpy(pw.data(), Pw, pw.size() * 16);
  DevBuf<float4> d_pw, d_strips, d_cp;
  DevBuf<int> d_su, d_sv;
  RTN_TRY(d_sd.upload(&d, 1));
  RTN_TRY(d_pw.upload(pw.data(), pw.size()));
  RTN_TRY(d_strips.alloc((size_t)d.nbU * (p + 1) * (m + 1)));
  RTN_TRY(d_cp.alloc(nps * ps));
  RTN_TRY(d_su.alloc(knots.size()));
  RTN_TRY(d_sv.alloc(knots.size()));
  hipLaunchKernelGGL(k_decomp_u, dim3(1), dim3(256), 0, nullptr, d_sd.ptr, d_knots.ptr, d_pw.ptr, d_strips.ptr,
                     d_su.ptr);
  hipLaunchKernelGGL(k_decomp_v, dim3(1), dim3(256), 0, nullptr, d_sd.ptr, d_knots.ptr, d_strips.ptr, d_cp.ptr,
                     d_sv.ptr);
  RTN_HIP(hipGetLastError());
  if (cp) RTN_HIP(hipMemcpy(cp, d_cp.ptr, nps * ps * sizeof(float4), hipMemcpyDeviceToHost));
  if (uv_range) {
    std::vector<float> brk(knots.size());
    RTN_TRY(d_brk.download(brk.data(), brk.size()));
    for (int i = 0; i < d.nbU; i++)
      for (int j = 0; j < d.nbV; j++) {
        float* r = uv_range + 4 * ((size_t)i * d.nbV + j);
        r[0] = brk[i]; r[1] = brk[i + 1]; r[2] = brk[d.v_off + j]; r[3] = brk[d.v_off + j + 1];
      }
  }
  return RTN_OK;
}

extern "C" int rtn_patch_bounds(const float* cp, int np, int p, int q, float* bounds) {
  RTN_CHECK(cp && bounds && np > 0 && p >= 1 && q >= 1, RTN_EINVAL, "rtn_patch_bounds: bad arguments");
  const int ps = (p + 1) * (q + 1);
  DevBuf<float4> d_cp;
  std::vector<float4> h((size_t)np * ps);
  memcpy(h.data(), cp, h.size() * 16);
  RTN_TRY(d_cp.upload(h.data(), h.size()));
  std::vector<int> off(np), pq(np, p | (q << 8));
  for (int i = 0; i < np; i++) off[i] = i * ps;
  DevBuf<int> d_off, d_pq;
  DevBuf<float> d_b;
  RTN_TRY(d_off.upload(off.data(), np));
  RTN_TRY(d_pq.upload(pq.data(), np));
  RTN_TRY(d_b.alloc(6 * (size_t)np));
  hipLaunchKernelGGL(k_bounds, dim3(nblk(np, 256)), dim3(256), 0, nullptr, d_cp.ptr, d_off.ptr, d_pq.ptr, np, d_b.ptr);
  RTN_HIP(hipGetLastError());
  return d_b.download(bounds, 6 * (size_t)np);
}

extern "C" int rtn_build_bvh(const float* bounds, int np, rtn_bvh_node* nodes, uint32_t* perm, int* nodes_used) {
  RTN_CHECK(bounds && nodes && perm && np > 0, RTN_EINVAL, "rtn_build_bvh: bad arguments");
  DevBuf<float> d_pb;
  RTN_TRY(d_pb.upload(bounds, 6 * (size_t)np));
  std::vector<rtn_bvh_node> out;
  DevBuf<uint32_t> d_perm;
  RTN_TRY(build_sah_device(d_pb.ptr, np, out, d_perm, nullptr, nullptr));
  memset(nodes, 0, sizeof(rtn_bvh_node) * (size_t)(2 * np - 1));
  memcpy(nodes, out.data(), out.size() * sizeof(rtn_bvh_node));
  if (nodes_used) *nodes_used = (int)out.size();
  return d_perm.download(perm, np);
}
