// Curvature-driven knot refinement of every NURBS surface of a scene on the
// GPU: curvatureRefinementU then curvatureRefinementV (nurbs.h:290-486) for
// all surfaces in two launches per direction.
//
//   k_refine_count   one workgroup per surface; one work item per (span,
//                    column curve): knotsToAdd (nurbs.h:255-281) on the
//                    dehomogenised curve, max over the columns by atomicMax.
//   k_refine_insert  one workgroup per surface; a lane per column curve
//                    inserts the span's knots one at a time with A5.1
//                    (nurbs.h:97-140, r = 1) in place; the knot vector is
//                    shared by the workgroup and advanced between insertions.
//
// The arithmetic follows the host implementation (refine.cpp) operation by
// operation: this file is compiled without FP contraction and with correctly
// rounded division and square root, so every step matches the host's IEEE
// single-precision result.  The one exception is Δ^1.5 in knotsToAdd, where
// the host calls powf: here it is sqrt(Δ)·Δ in double, rounded once to float
// (the correctly rounded x^1.5).  glibc's powf is not correctly rounded on
// about 0.07% of inputs; there the two differ by one ulp, which changes a
// span's knot count only if C·|A|·Δ^1.5/√(ΣV/p) lies within that ulp of an
// integer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "rtn_internal.hpp"

namespace {

// Per-surface description of one direction's refinement.
struct RefSurf {
  int nn;        // last control point index along the refined direction
  int p;         // degree along it
  int other;     // last column index (the other direction)
  int stride_a;  // float4 step between consecutive points of a column curve in the input net
  int stride_c;  // float4 step between column curves in the input net
  int64_t net;   // float4 offset of the input net
  int64_t knots; // float offset of the input knot vector
  int64_t span;  // offset of the per-span counts (nn - p + 1 of them)
  int64_t work;  // float4 offset of the column curves, (other + 1) x (nfin + 1)
  int64_t kout;  // float offset of the refined knot vector (nfin + p + 2)
  int nfin;      // last control point index after refinement
};

__host__ __device__ __forceinline__ float3 dehom(float4 w) { return make_float3(w.x / w.w, w.y / w.w, w.z / w.w); }

__host__ __device__ __forceinline__ float len3(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }

// x^1.5 for x >= 0, correctly rounded to float.
__host__ __device__ __forceinline__ float pow15(float x) {
  const double d = (double)x;
  return (float)(sqrt(d) * d);
}

// knotsToAdd for span i of one column curve (points P[a] at net + a * stride).
__host__ __device__ int knots_to_add_dev(const float4* __restrict__ net, int stride, const float* __restrict__ U, int i,
                                float C, int p) {
  float vsum = 0.0f;
  float3 Vj[RTN_MAX_DEGREE + 1];
  const int base = i - p + 1;
  for (int j = i - p + 1; j <= i; j++) {
    const float d = U[j + p] - U[j];
    float3 v = make_float3(0.f, 0.f, 0.f);
    if (d > 0.0f) {
      const float s = (float)p / d;
      const float3 a = dehom(net[(int64_t)j * stride]), b = dehom(net[(int64_t)(j - 1) * stride]);
      v = make_float3((a.x - b.x) * s, (a.y - b.y) * s, (a.z - b.z) * s);
    }
    Vj[j - base] = v;
    vsum += len3(v);
  }
  float amod = 0.0f;
  for (int j = i - p + 2; j <= i; j++) {
    const float d = U[j + p - 1] - U[j];
    float3 A = make_float3(0.f, 0.f, 0.f);
    if (d > 0.0f) {
      const float s = (float)(p - 1) / d;
      const float3 a = Vj[j - base], b = Vj[j - 1 - base];
      A = make_float3((a.x - b.x) * s, (a.y - b.y) * s, (a.z - b.z) * s);
    }
    amod = fmaxf(amod, len3(A));
  }
  const float num = amod * pow15(U[i + 1] - U[i]);
  const float den = sqrtf(vsum / (float)p);
  if (!(den > 1e-9f)) return 0;
  const float nf = C * num / den;
  if (!(nf > 1e-9f)) return 0;
  return (int)nf;
}

__global__ void k_refine_count(const RefSurf* __restrict__ S, const float4* __restrict__ nets,
                               const float* __restrict__ knots, float C, int* __restrict__ add) {
  const RefSurf s = S[blockIdx.x];
  const int64_t spans = s.nn - s.p + 1, cols = s.other + 1;
  for (int64_t w = threadIdx.x; w < spans * cols; w += blockDim.x) {
    const int i = s.p + (int)(w / cols), c = (int)(w % cols);
    const int k = knots_to_add_dev(nets + s.net + (int64_t)c * s.stride_c, s.stride_a, knots + s.knots, i, C, s.p);
    if (k > 0) atomicMax(add + s.span + (i - s.p), k);
  }
}

__host__ __device__ int find_span_dev(int n, int p, float u, const float* U) {
  if (u == U[n + 1]) return n;
  int low = p, high = n + 1, mid = (low + high) / 2;
  while (u < U[mid] || u >= U[mid + 1]) {
    if (u < U[mid]) high = mid; else low = mid;
    mid = (low + high) / 2;
  }
  return mid;
}

// A5.1 with r = 1 in place on one column curve P[0..ncur] (room for one more
// point): u enters span k, whose knot has multiplicity sm < p; U is the knot
// vector before the insertion.
__host__ __device__ inline void insert_in_column(float4* P, int ncur, int p, const float* U, float u, int k, int sm) {
  const int L = k - p + 1;
  float4 R[RTN_MAX_DEGREE + 1];
  for (int t = 0; t <= p - sm; t++) R[t] = P[k - p + t];
  for (int t = ncur; t >= k - sm; t--) P[t + 1] = P[t];
  for (int t = 0; t <= p - 1 - sm; t++) {
    const float alpha = (u - U[L + t]) / (U[t + k + 1] - U[L + t]);
    const float om = 1.0f - alpha;
    R[t] = make_float4(R[t + 1].x * alpha + R[t].x * om, R[t + 1].y * alpha + R[t].y * om,
                       R[t + 1].z * alpha + R[t].z * om, R[t + 1].w * alpha + R[t].w * om);
  }
  for (int t = 0; t <= p - 1 - sm; t++) P[L + t] = R[t];
}

// The knot vector U[0..ncur+p+1] gains u after index k.
__host__ __device__ inline void insert_knot(float* U, int ncur, int p, float u, int k) {
  for (int t = ncur + p + 1; t >= k + 1; t--) U[t + 1] = U[t];
  U[k + 1] = u;
}

__global__ void k_refine_insert(const RefSurf* __restrict__ S, const float4* __restrict__ nets,
                                const float* __restrict__ knots, const int* __restrict__ add,
                                float4* __restrict__ work, float* __restrict__ kout, int* __restrict__ bad) {
  const RefSurf s = S[blockIdx.x];
  const int p = s.p, cols = s.other + 1, ld = s.nfin + 1;
  const float* U0 = knots + s.knots;
  float* Us = kout + s.kout;  // the current knot vector, advanced by lane 0
  float4* W = work + s.work;
  for (int i = threadIdx.x; i < s.nn + p + 2; i += blockDim.x) Us[i] = U0[i];
  for (int c = threadIdx.x; c < cols; c += blockDim.x)
    for (int a = 0; a <= s.nn; a++) W[(int64_t)c * ld + a] = nets[s.net + (int64_t)c * s.stride_c + (int64_t)a * s.stride_a];
  __syncthreads();
  int ncur = s.nn;
  for (int i = p; i <= s.nn; i++) {
    const int cnt = add[s.span + (i - p)];
    const float lo = U0[i], hi = U0[i + 1];
    for (int j = 1; j <= cnt; j++) {
      const float u = lo + (hi - lo) * (float)j / (float)(cnt + 1);
      const int k = find_span_dev(ncur, p, u, Us);
      int sm = 0;
      for (int t = k; t >= 0 && Us[t] == u; --t) ++sm;
      if (1 + sm > p) {  // A5.1 cannot insert beyond the degree (the reference reports an error)
        if (threadIdx.x == 0) atomicOr(bad, 1);
        return;  // every lane takes the same branch: no barrier is left waiting
      }
      for (int c = threadIdx.x; c < cols; c += blockDim.x) insert_in_column(W + (int64_t)c * ld, ncur, p, Us, u, k, sm);
      __syncthreads();  // every lane has read the old knots
      if (threadIdx.x == 0) insert_knot(Us, ncur, p, u, k);
      __syncthreads();
      ncur++;
    }
  }
}

}  // namespace

extern "C" int rtn_scene_refine(rtn_scene* sc, float C_u, float C_v) {
  RTN_CHECK(sc, RTN_EINVAL, "rtn_scene_refine: null scene");
  RTN_CHECK(!sc->built, RTN_EINVAL, "rtn_scene_refine: scene already built");
  RTN_CHECK(!(C_u < 0.0f) && !(C_v < 0.0f), RTN_EINVAL, "rtn_scene_refine: negative constant");
  const int ns = (int)sc->surfaces.size();
  if (ns == 0) return RTN_OK;
  for (int dir = 0; dir < 2; dir++) {
    const float C = dir == 0 ? C_u : C_v;
    if (!(C > 0.0f)) continue;
    std::vector<RefSurf> meta(ns);
    std::vector<float4> nets;
    std::vector<float> knots;
    int64_t spans = 0;
    for (int k = 0; k < ns; k++) {
      const rtn::HostSurface& h = sc->surfaces[k];
      RefSurf& r = meta[k];
      r.nn = dir == 0 ? h.n : h.m;
      r.p = dir == 0 ? h.p : h.q;
      r.other = dir == 0 ? h.m : h.n;
      // the net is row-major in u: P[a][c] at a * (m + 1) + c
      r.stride_a = dir == 0 ? h.m + 1 : 1;
      r.stride_c = dir == 0 ? 1 : h.m + 1;
      r.net = (int64_t)nets.size();
      r.knots = (int64_t)knots.size();
      r.span = spans;
      spans += r.nn - r.p + 1;
      const size_t n4 = h.Pw.size() / 4;
      const float4* src = reinterpret_cast<const float4*>(h.Pw.data());
      nets.insert(nets.end(), src, src + n4);
      const std::vector<float>& K = dir == 0 ? h.U : h.V;
      knots.insert(knots.end(), K.begin(), K.end());
    }
    rtn::DevBuf<RefSurf> d_meta;
    rtn::DevBuf<float4> d_nets;
    rtn::DevBuf<float> d_knots;
    rtn::DevBuf<int> d_add;
    RTN_TRY(d_meta.upload(meta.data(), meta.size()));
    RTN_TRY(d_nets.upload(nets.data(), nets.size()));
    RTN_TRY(d_knots.upload(knots.data(), knots.size()));
    RTN_TRY(d_add.alloc((size_t)spans));
    RTN_HIP(hipMemset(d_add.ptr, 0, sizeof(int) * (size_t)spans));
    hipLaunchKernelGGL(k_refine_count, dim3(ns), dim3(256), 0, 0, d_meta.ptr, d_nets.ptr, d_knots.ptr, C, d_add.ptr);
    RTN_HIP(hipGetLastError());
    std::vector<int> add((size_t)spans);
    RTN_TRY(d_add.download(add.data(), add.size()));
    int64_t work = 0, kout = 0;
    for (int k = 0; k < ns; k++) {
      RefSurf& r = meta[k];
      int64_t tot = 0;
      for (int i = 0; i <= r.nn - r.p; i++) tot += add[(size_t)(r.span + i)];
      RTN_CHECK(r.nn + tot < (1 << 24), RTN_EUNSUPPORTED, "rtn_scene_refine: surface %d would get %lld knots", k,
                (long long)tot);
      r.nfin = r.nn + (int)tot;
      r.work = work;
      r.kout = kout;
      work += (int64_t)(r.other + 1) * (r.nfin + 1);
      kout += r.nfin + r.p + 2;
    }
    rtn::DevBuf<float4> d_work;
    rtn::DevBuf<float> d_kout;
    rtn::DevBuf<int> d_bad;
    RTN_TRY(d_meta.upload(meta.data(), meta.size()));
    RTN_TRY(d_work.alloc((size_t)work));
    RTN_TRY(d_kout.alloc((size_t)kout));
    RTN_TRY(d_bad.alloc(1));
    RTN_HIP(hipMemset(d_bad.ptr, 0, sizeof(int)));
    hipLaunchKernelGGL(k_refine_insert, dim3(ns), dim3(64), 0, 0, d_meta.ptr, d_nets.ptr, d_knots.ptr, d_add.ptr,
                       d_work.ptr, d_kout.ptr, d_bad.ptr);
    RTN_HIP(hipGetLastError());
    int bad = 0;
    RTN_TRY(d_bad.download(&bad, 1));
    RTN_CHECK(bad == 0, RTN_EINVAL, "rtn_scene_refine: a knot insertion would exceed the degree");
    std::vector<float4> wk((size_t)work);
    std::vector<float> ko((size_t)kout);
    RTN_TRY(d_work.download(wk.data(), wk.size()));
    RTN_TRY(d_kout.download(ko.data(), ko.size()));
    for (int k = 0; k < ns; k++) {
      rtn::HostSurface& h = sc->surfaces[k];
      const RefSurf& r = meta[k];
      const int ld = r.nfin + 1;
      std::vector<float> K(ko.begin() + r.kout, ko.begin() + r.kout + r.nfin + r.p + 2);
      std::vector<float> Pw((size_t)4 * (r.other + 1) * ld);
      float4* o = reinterpret_cast<float4*>(Pw.data());
      for (int c = 0; c <= r.other; c++)
        for (int a = 0; a <= r.nfin; a++) {
          const float4 v = wk[(size_t)(r.work + (int64_t)c * ld + a)];
          if (dir == 0) o[(size_t)a * (r.other + 1) + c] = v;  // new n = nfin, m unchanged
          else o[(size_t)c * ld + a] = v;                      // n unchanged, new m = nfin
        }
      if (dir == 0) { h.n = r.nfin; h.U = std::move(K); }
      else { h.m = r.nfin; h.V = std::move(K); }
      h.Pw = std::move(Pw);
    }
  }
  return RTN_OK;
}
