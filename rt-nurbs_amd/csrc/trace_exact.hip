// Plain method, exact form: calculateIntersect (intersect.cu:20-194) for gfx950.
//
// One lane per ray over the reference's binary BVH (32-byte nodes fetched as
// two 16-byte loads), the traversal stack in LDS (one column per lane, so the
// 64 lanes of a wave hit 64 distinct banks), the bilinear seed and the Newton
// solve in the reference's operation order.  Degree (2,2) and (3,3) patches
// run compile-time-degree evaluators whose de Casteljau triangles stay in
// VGPRs; other degrees use a guarded unrolled evaluator (degree <= 5).
//
// T = float is the plain FP32 method; this translation unit is compiled with
// -ffp-contract=off, so its results equal the host reference bit for bit.
// T = double is the same traversal in FP64, used as the FP64 reference for
// error reports (with its own iteration cap and tolerance).
#include <climits>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "rtn_exact.hpp"
#include "rtn_internal.hpp"

using namespace rtn;
using namespace rtn::ex;

namespace {

constexpr int TB = 128;  // lanes per block (2 waves)

template <class T>
struct ExactArgs {
  const float4* nodes;  // rtn_bvh_node viewed as 2 x float4
  const float *ox, *oy, *oz, *dx, *dy, *dz;
  int64_t n_rays;
  int W, H;             // > 0: 2-D pixel grid (drop-in), 16 x 8 pixel tiles
  const float* delta;
  float flat, cam_k;
  int max_iter;
  T thr_det, thr_res;
  T prune_slack;        // boxes entered beyond t_max * slack are skipped (1 = reference)
  int grid;             // FP64 closest-root reference: > 0 adds grid x grid Newton starts per tested patch
  uint32_t seed_base;
  // scene-form outputs
  int* prim;
  T *t, *u, *v, *dpdu, *dpdv;
  T* t2;                // second-nearest converged hit on another patch (FP64 reference)
  uint32_t* counters;
  // drop-in outputs (surfaceInteraction AoS + idx_interact)
  rtn_surface_interaction* inter;
  int* idx_interact;
  unsigned long long* totals;  // 8 totals, or null
};

struct DropInPatches {  // bezierPatch AoS (bezier_patch.h:10-16) with device pointers
  const rtn_bezier_patch* p;
  __device__ __forceinline__ void get(int i, const float4*& cp, int& P, int& Q) const {
    cp = reinterpret_cast<const float4*>(p[i].control_point);
    P = p[i].p;
    Q = p[i].q;
  }
};
struct ScenePatches {
  const float4* cp;
  const int* off;
  const int* pq;
  __device__ __forceinline__ void get(int i, const float4*& c, int& P, int& Q) const {
    c = cp + off[i];
    const int v = pq[i];
    P = v & 0xff;
    Q = (v >> 8) & 0xff;
  }
};

__device__ __forceinline__ void load_node(const float4* nodes, int i, float b[6], uint32_t& lf, uint32_t& pc) {
  const float4 a = nodes[2 * i], c = nodes[2 * i + 1];
  b[0] = a.x; b[1] = a.y; b[2] = a.z; b[3] = a.w; b[4] = c.x; b[5] = c.y;
  lf = __float_as_uint(c.z);
  pc = __float_as_uint(c.w);
}

template <class T>
__device__ __forceinline__ T slab6(const float b[6], v3<T> o, v3<T> rd) {
  return slab<T>(T(b[0]), T(b[3]), T(b[1]), T(b[4]), T(b[2]), T(b[5]), o, rd);
}

template <class T>
__device__ __forceinline__ NewtonRes<T> newton_dispatch(const RayPlane<T>& rp, const float4* cp, int P, int Q, T gu,
                                                        T gv, const ExactArgs<T>& A, uint32_t seed) {
  if (P == 3 && Q == 3)
    return newton<T>(rp, [&](T uu, T vv, v4<T>& du, v4<T>& dv, v4<T>& pt) { patch_eval<3, 3>(cp, uu, vv, du, dv, pt); },
                     gu, gv, A.thr_det, A.thr_res, seed, A.max_iter);
  if (P == 2 && Q == 2)
    return newton<T>(rp, [&](T uu, T vv, v4<T>& du, v4<T>& dv, v4<T>& pt) { patch_eval<2, 2>(cp, uu, vv, du, dv, pt); },
                     gu, gv, A.thr_det, A.thr_res, seed, A.max_iter);
  return newton<T>(rp,
                   [&](T uu, T vv, v4<T>& du, v4<T>& dv, v4<T>& pt) {
                     patch_eval_rt<RTN_MAX_DEGREE>(cp, P, Q, uu, vv, du, dv, pt);
                   },
                   gu, gv, A.thr_det, A.thr_res, seed, A.max_iter);
}

template <int STACK, class T, class Patches>
__device__ __forceinline__ void trace_one(const ExactArgs<T>& A, const Patches& PT, int64_t idx, int tid,
                                          int (*stk)[TB], uint32_t cnt[8]) {
  const v3<T> o{T(A.ox[idx]), T(A.oy[idx]), T(A.oz[idx])};
  const v3<T> d{T(A.dx[idx]), T(A.dy[idx]), T(A.dz[idx])};
  uint32_t nc = 0, ne = 0, sc = 0, niv = 0, nlv = 0, npt = 0, ovf = 0, nfl = 0;
  const v3<T> rd{T(1) / d.x, T(1) / d.y, T(1) / d.z};
  const RayPlane<T> rp = ray_planes(o, d);
  int off = 0;
  T t_max = T(INFINITY), t_second = T(INFINITY);
  int best = -1;
  T bt = 0, bu = 0, bv = 0;
  v3<T> bsu{0, 0, 0}, bsv{0, 0, 0};
  float cb[6];
  uint32_t lf, pc;
  load_node(A.nodes, 0, cb, lf, pc);
  T it = slab6<T>(cb, o, rd);
  bool go = it < t_max;
  const uint32_t seed = (uint32_t)idx + A.seed_base;
  const T slack = A.prune_slack;
  while (go) {
    if (pc > 0) {
      nlv++;
      it = slab6<T>(cb, o, rd);
      for (uint32_t i = 0; i < pc; i++) {
        const int pi = (int)(lf + i);
        bool do_newton = true;
        npt++;
        const float4* cp;
        int P, Q;
        PT.get(pi, cp, P, Q);
        const v3<T> q00 = dehom(ld4<T>(cp[0])), q10 = dehom(ld4<T>(cp[P * (Q + 1)]));
        const v3<T> q11 = dehom(ld4<T>(cp[P * (Q + 1) + Q])), q01 = dehom(ld4<T>(cp[Q]));
        T gu, gv;
        bilinear_seed(q00, q10, q11, q01, o, d, gu, gv);
        if (A.delta && T(A.delta[pi]) * T(A.cam_k) / it < T(A.flat)) {  // flat-skip (intersect.cu:104-123)
          v4<T> du, dv, pt;
          patch_eval_rt<RTN_MAX_DEGREE>(cp, P, Q, gu, gv, du, dv, pt);
          v3<T> S, Su, Sv;
          dehom_ders(pt, du, dv, S, Su, Sv);
          const T resid = xabs(dot(rp.n1, S) + rp.d1) + xabs(dot(rp.n2, S) + rp.d2);
          if (resid * T(A.cam_k) / it < T(A.flat)) {
            sc++;
            const T t = dot(sub(S, o), d);
            if (0 < t && t < t_max) {
              t_second = t_max;
              t_max = t; best = pi; bt = t; bu = gu; bv = gv; bsu = Su; bsv = Sv;
              do_newton = false;
            }
          }
        }
        if (do_newton) {
          nc++;
          NewtonRes<T> nr = newton_dispatch<T>(rp, cp, P, Q, gu, gv, A, seed);
          ne += nr.evals;
          nfl += nr.evals * eval_flops(P, Q);
          T t = nr.hit ? (nr.S.x - o.x) * d.x + (nr.S.y - o.y) * d.y + (nr.S.z - o.z) * d.z : T(0);
          if constexpr (std::is_same<T, double>::value) {
            // closest-root reference: Newton also from the centres of a grid x
            // grid partition of the domain; the patch's hit is its nearest
            // converged root in front of the origin over all starts
            for (int k = 0; k < A.grid * A.grid; k++) {
              const T su = (T(k / A.grid) + T(0.5)) / T(A.grid), sv = (T(k % A.grid) + T(0.5)) / T(A.grid);
              const NewtonRes<T> g = newton_dispatch<T>(rp, cp, P, Q, su, sv, A, seed);
              ne += g.evals;
              nfl += g.evals * eval_flops(P, Q);
              if (!g.hit) continue;
              const T tg = (g.S.x - o.x) * d.x + (g.S.y - o.y) * d.y + (g.S.z - o.z) * d.z;
              if (0 < tg && (!nr.hit || !(0 < t) || tg < t)) {
                nr = g;
                t = tg;
              }
            }
          }
          if (nr.hit) {
            if (0 < t && t < t_max) {
              t_second = t_max;
              best = pi; bt = t; bu = nr.u; bv = nr.v; bsu = nr.Su; bsv = nr.Sv;
              t_max = t;
            } else if (0 < t && t < t_second) {
              t_second = t;
            }
          }
        }
      }
      if (off == 0) break;
      load_node(A.nodes, stk[--off][tid], cb, lf, pc);
    } else {
      niv++;
      float lb[6], rb[6];
      uint32_t llf, lpc, rlf, rpc;
      load_node(A.nodes, (int)lf, lb, llf, lpc);
      load_node(A.nodes, (int)lf + 1, rb, rlf, rpc);
      const T lt = slab6<T>(lb, o, rd);
      const T rt = slab6<T>(rb, o, rd);
      it = xmin(lt, rt);
      const T lim = t_max * slack;
      if (it >= lim) {
        if (off == 0) break;
        load_node(A.nodes, stk[--off][tid], cb, lf, pc);
      } else if (lt <= it) {
        if (rt < lim) {
          if (off < STACK) stk[off++][tid] = (int)lf + 1; else ovf = 1;
        }
#pragma unroll
        for (int k = 0; k < 6; k++) cb[k] = lb[k];
        lf = llf; pc = lpc;
      } else {
        if (lt < lim) {
          if (off < STACK) stk[off++][tid] = (int)lf; else ovf = 1;
        }
#pragma unroll
        for (int k = 0; k < 6; k++) cb[k] = rb[k];
        lf = rlf; pc = rpc;
      }
    }
  }
  if (A.inter) {  // drop-in record (intersect.cu:114-119, :133-138, :192-193)
    rtn_surface_interaction* si = A.inter + idx;
    int ii = A.idx_interact[idx];
    if (best >= 0) {
      si->t_hit = (float)bt;
      si->uv[0] = (float)bu; si->uv[1] = (float)bv;
      si->dpdu[0] = (float)bsu.x; si->dpdu[1] = (float)bsu.y; si->dpdu[2] = (float)bsu.z;
      si->dpdv[0] = (float)bsv.x; si->dpdv[1] = (float)bsv.y; si->dpdv[2] = (float)bsv.z;
      ii = (int)idx;
      A.idx_interact[idx] = ii;
    }
    if (ii != -1) {
      si->wo[0] = (float)d.x * -1.0f; si->wo[1] = (float)d.y * -1.0f; si->wo[2] = (float)d.z * -1.0f;
    }
  }
  if (A.prim) {
    A.prim[idx] = best;
    if (A.t2) A.t2[idx] = t_second;
    if (best >= 0) {
      if (A.t) A.t[idx] = bt;
      if (A.u) A.u[idx] = bu;
      if (A.v) A.v[idx] = bv;
      if (A.dpdu) { A.dpdu[3 * idx] = bsu.x; A.dpdu[3 * idx + 1] = bsu.y; A.dpdu[3 * idx + 2] = bsu.z; }
      if (A.dpdv) { A.dpdv[3 * idx] = bsv.x; A.dpdv[3 * idx + 1] = bsv.y; A.dpdv[3 * idx + 2] = bsv.z; }
    }
  }
  cnt[0] = nc; cnt[1] = ne; cnt[2] = sc; cnt[3] = niv; cnt[4] = nlv; cnt[5] = npt; cnt[6] = ovf; cnt[7] = nfl;
  if (A.counters) {
    uint32_t* c = A.counters + 8 * idx;
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = cnt[k];
  }
}

template <int STACK, bool GRID2D, class T, class Patches>
__global__ void __launch_bounds__(TB) k_trace_exact(ExactArgs<T> A, Patches PT) {
  __shared__ int stk[STACK][TB];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  int64_t idx;
  bool valid;
  if (GRID2D) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    valid = x < A.W && y < A.H;
    idx = (int64_t)y * A.W + x;
  } else {
    idx = (int64_t)blockIdx.x * TB + tid;
    valid = idx < A.n_rays;
  }
  uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (valid) trace_one<STACK, T>(A, PT, idx, tid, stk, cnt);
  if (A.totals) {  // every lane of the wave takes part in the reduction
#pragma unroll
    for (int k = 0; k < 8; k++) {
      unsigned long long x = cnt[k];
      for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
      if ((tid & 63) == 0 && x) atomicAdd(A.totals + k, x);
    }
  }
}

template <class T>
int launch_scene(const rtn_scene* s, const ExactArgs<T>& A, int64_t n, hipStream_t stream) {
  ScenePatches P{s->cp.ptr, s->cp_off.ptr, s->pq.ptr};
  const unsigned blocks = (unsigned)((n + TB - 1) / TB);
  if (s->max_depth < 32)
    hipLaunchKernelGGL((k_trace_exact<32, false, T, ScenePatches>), dim3(blocks), dim3(TB), 0, stream, A, P);
  else
    hipLaunchKernelGGL((k_trace_exact<64, false, T, ScenePatches>), dim3(blocks), dim3(TB), 0, stream, A, P);
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

}  // namespace

namespace rtn {

// Launches the exact FP32 kernel over a 1-D ray array on scene buffers.
int trace_exact_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                      const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream) {
  ExactArgs<float> A{};
  A.nodes = reinterpret_cast<const float4*>(s->nodes.ptr);
  A.ox = rays[0]; A.oy = rays[1]; A.oz = rays[2]; A.dx = rays[3]; A.dy = rays[4]; A.dz = rays[5];
  A.n_rays = n;
  A.delta = prm->flat_threshold > 0.f ? s->delta.ptr : nullptr;
  A.flat = prm->flat_threshold;
  A.cam_k = prm->cam_k;
  A.max_iter = prm->max_iter;
  A.thr_det = prm->thr_det;
  A.thr_res = prm->thr_res;
  A.prune_slack = 1.0f;
  A.seed_base = prm->seed_base;
  A.prim = hits->prim; A.t = hits->t; A.u = hits->u; A.v = hits->v; A.dpdu = hits->dpdu; A.dpdv = hits->dpdv;
  A.counters = prm->count ? hits->counters : nullptr;
  A.totals = d_totals;
  return launch_scene<float>(s, A, n, stream);
}

}  // namespace rtn

// FP64 reference trace: the plain method's traversal and seed in FP64 with
// its own Newton cap and tolerances; boxes are pruned only beyond
// t_max * (1 + 4e-5) so the second-nearest hit of near-tie pixels is found.
// grid > 0: the closest-root reference (grid x grid extra starts per patch).
extern "C" int rtn_trace_f64_roots(const rtn_scene* s, const float* const rays[6], int64_t n, int max_iter,
                                   double thr_det, double thr_res, int grid, uint32_t seed_base, int* prim, double* t,
                                   double* u, double* v, double* dpdu, double* dpdv, double* t_second,
                                   uint32_t* counters, void* stream) {
  RTN_CHECK(s && s->built && rays && prim, RTN_EINVAL, "rtn_trace_f64: bad arguments");
  RTN_CHECK(max_iter >= 1 && max_iter <= 64, RTN_EINVAL, "rtn_trace_f64: max_iter outside 1..64");
  RTN_CHECK(grid >= 0 && grid <= 16, RTN_EINVAL, "rtn_trace_f64_roots: grid outside 0..16");
  RTN_CHECK(n >= 0 && n < ((int64_t)1 << 40), RTN_EINVAL, "rtn_trace_f64: bad ray count");
  if (n == 0) return RTN_OK;
  ExactArgs<double> A{};
  A.nodes = reinterpret_cast<const float4*>(s->nodes.ptr);
  A.ox = rays[0]; A.oy = rays[1]; A.oz = rays[2]; A.dx = rays[3]; A.dy = rays[4]; A.dz = rays[5];
  A.n_rays = n;
  A.max_iter = max_iter;
  A.thr_det = thr_det;
  A.thr_res = thr_res;
  A.prune_slack = 1.0 + 4e-5;
  A.grid = grid;
  A.seed_base = seed_base;
  A.prim = prim; A.t = t; A.u = u; A.v = v; A.dpdu = dpdu; A.dpdv = dpdv; A.t2 = t_second;
  A.counters = counters;
  RTN_TRY(launch_scene<double>(s, A, n, (hipStream_t)stream));
  if (!stream) RTN_HIP(hipStreamSynchronize(nullptr));
  return RTN_OK;
}

extern "C" int rtn_trace_f64(const rtn_scene* s, const float* const rays[6], int64_t n, int max_iter, double thr_det,
                             double thr_res, uint32_t seed_base, int* prim, double* t, double* u, double* v,
                             double* dpdu, double* dpdv, double* t_second, uint32_t* counters, void* stream) {
  return rtn_trace_f64_roots(s, rays, n, max_iter, thr_det, thr_res, 0, seed_base, prim, t, u, v, dpdu, dpdv,
                             t_second, counters, stream);
}

namespace {
// The drop-in launch on reference-layout buffers with the caller's totals
// block tot[8] (zeroed here by a kernel), then the reference's synchronous
// wrap-up: counters added to, stack overflow reported.
__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

int calculate_intersect(const rtn_ray_queue* rays, const rtn_bvh_node* nodes, const rtn_bezier_patch* patches,
                        rtn_surface_interaction* interact, int W, int H, float flat_threshold, float v_fov,
                        const float* delta, unsigned long long* newton_count, unsigned long long* skip_count,
                        unsigned long long* tot) {
  ExactArgs<float> A{};
  A.nodes = reinterpret_cast<const float4*>(nodes);
  A.ox = rays->origin_x; A.oy = rays->origin_y; A.oz = rays->origin_z;
  A.dx = rays->direction_x; A.dy = rays->direction_y; A.dz = rays->direction_z;
  A.n_rays = (int64_t)W * H;
  A.W = W;
  A.H = H;
  A.delta = delta;
  A.flat = flat_threshold;
  A.cam_k = (float)(0.5 * (float)H / std::tan((double)(v_fov / 2.0f)));  // intersect.cu:212, in double
  A.max_iter = 7;        // NEWTON_MAX_ITERATIONS (constants.h:40)
  A.thr_det = 1e-5f;     // intersect.cu:127
  A.thr_res = 1e-4f;
  A.prune_slack = 1.0f;
  A.seed_base = 0;
  A.inter = interact;
  A.idx_interact = rays->idx_interact;
  A.totals = tot;  // the totals also carry the stack-overflow count
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, nullptr, tot, 8);
  DropInPatches P{patches};
  dim3 block(16, 8), grid((W + 15) / 16, (H + 7) / 8);
  hipLaunchKernelGGL((k_trace_exact<32, true, float, DropInPatches>), grid, block, 0, nullptr, A, P);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "calculateIntersect launch", __FILE__, __LINE__);
  unsigned long long h[8];
  RTN_HIP(hipMemcpy(h, tot, sizeof(h), hipMemcpyDeviceToHost));
  // counters are added to, as the reference's atomicAdd does (intersect.cu:189-190)
  unsigned long long* outs[2] = {newton_count, skip_count};
  const unsigned long long adds[2] = {h[0], h[2]};
  for (int k = 0; k < 2; k++) {
    if (!outs[k]) continue;
    unsigned long long cur;
    RTN_HIP(hipMemcpy(&cur, outs[k], 8, hipMemcpyDefault));
    cur += adds[k];
    RTN_HIP(hipMemcpy(outs[k], &cur, 8, hipMemcpyDefault));
  }
  if (h[6]) {
    set_error("calculateIntersect: traversal stack (32 entries) overflowed on %llu rays", h[6]);
    return RTN_EOVERFLOW;
  }
  return RTN_OK;
}

int check_dropin_args(const rtn_ray_queue* rays, const rtn_bvh_node* nodes, const rtn_bezier_patch* patches,
                      rtn_surface_interaction* interact, int W, int H, const char* who) {
  RTN_CHECK(rays && nodes && patches && interact && W > 0 && H > 0, RTN_EINVAL, "%s: bad arguments", who);
  RTN_CHECK(rays->idx_interact, RTN_EINVAL, "%s: idx_interact is required", who);
  return RTN_OK;
}

// Scratch of rtn_calculate_intersect_noalloc: its totals block.  A file-scope
// device array, one copy per GPU, held until the process ends; host code uses
// the device address that hipGetSymbolAddress gives for the current device.
constexpr int kNoAllocWords = 8;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned long long scratch[kNoAllocWords];

namespace {
// The calling thread's current device's address of `scratch`, looked up once
// per device.
int scratch_address(unsigned long long** out) {
  int dev = 0;
  RTN_HIP(hipGetDevice(&dev));
  RTN_CHECK(dev >= 0 && dev < kMaxDevices, RTN_EUNSUPPORTED, "scratch: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    RTN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch[dev];
  return RTN_OK;
}
}  // namespace

extern "C" int rtn_calculate_intersect(const rtn_ray_queue* rays, const rtn_bvh_node* nodes,
                                       const rtn_bezier_patch* patches, rtn_surface_interaction* interact, int W,
                                       int H, float flat_threshold, float v_fov, const float* delta,
                                       unsigned long long* newton_count, unsigned long long* skip_count) {
  RTN_TRY(check_dropin_args(rays, nodes, patches, interact, W, H, "rtn_calculate_intersect"));
  unsigned long long* tot = nullptr;
  RTN_HIP(hipMalloc(&tot, 8 * sizeof(unsigned long long)));
  const int rc = calculate_intersect(rays, nodes, patches, interact, W, H, flat_threshold, v_fov, delta, newton_count,
                                     skip_count, tot);
  (void)hipFree(tot);
  return rc;
}

extern "C" int rtn_calculate_intersect_noalloc(const rtn_ray_queue* rays, const rtn_bvh_node* nodes,
                                               const rtn_bezier_patch* patches, rtn_surface_interaction* interact,
                                               int W, int H, float flat_threshold, float v_fov, const float* delta,
                                               unsigned long long* newton_count, unsigned long long* skip_count) {
  RTN_TRY(check_dropin_args(rays, nodes, patches, interact, W, H, "rtn_calculate_intersect_noalloc"));
  RTN_CHECK((int64_t)W * H <= INT32_MAX, RTN_EINVAL,
            "rtn_calculate_intersect_noalloc: %d x %d pixels exceed the int idx_interact range", W, H);
  unsigned long long* tot = nullptr;
  RTN_TRY(scratch_address(&tot));
  return calculate_intersect(rays, nodes, patches, interact, W, H, flat_threshold, v_fov, delta, newton_count,
                             skip_count, tot);
}

// ------------------------------------------------------------ pair solves
// The plain method's Newton solve (bezier_patch.h:198-266, reference
// operation order) on given (ray, patch) pairs, from the plain bilinear seed
// or from given seeds: isolates what a seed does to the iteration.
namespace {
__global__ void k_solve_pairs(const float4* cp, const int* cp_off, const int* pq, const float4* corners, int np,
                              const float* ox, const float* oy, const float* oz, const float* dx, const float* dy,
                              const float* dz, int64_t n_rays, const int* pair_ray, const int* pair_patch,
                              const float* seed_uv, int64_t n, int max_iter, float thr_det, float thr_res,
                              uint32_t seed_base, int grid, int* iters, int* converged, float* uv_out, float* t_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = pair_ray[i], k = pair_patch[i];
  if (r < 0 || r >= n_rays || k < 0 || k >= np) {
    iters[i] = -1;
    converged[i] = 0;
    return;
  }
  const v3<float> o{ox[r], oy[r], oz[r]}, d{dx[r], dy[r], dz[r]};
  const RayPlane<float> rp = ray_planes(o, d);
  const float4* net = cp + cp_off[k];
  const int P = pq[k] & 0xff, Q = (pq[k] >> 8) & 0xff;
  float u0, v0;
  if (seed_uv) {
    u0 = seed_uv[2 * i];
    v0 = seed_uv[2 * i + 1];
  } else {
    const float4 c0 = corners[3 * k], c1 = corners[3 * k + 1], c2 = corners[3 * k + 2];
    bilinear_seed<float>({c0.x, c0.y, c0.z}, {c0.w, c1.x, c1.y}, {c1.z, c1.w, c2.x}, {c2.y, c2.z, c2.w}, o, d, u0, v0);
  }
  const auto eval = [&](float uu, float vv, v4<float>& du, v4<float>& dv, v4<float>& pt) {
    patch_eval_rt<RTN_MAX_DEGREE>(net, P, Q, uu, vv, du, dv, pt);
  };
  NewtonRes<float> nr = newton<float>(rp, eval, u0, v0, thr_det, thr_res, (uint32_t)r + seed_base, max_iter);
  int evals = nr.evals;
  float t = nr.hit ? (nr.S.x - o.x) * d.x + (nr.S.y - o.y) * d.y + (nr.S.z - o.z) * d.z : -1.0f;
  bool hit = nr.hit && t > 0.0f;
  // closest root: the same solve also from the centres of a grid x grid
  // partition of the domain; the nearest converged root with t > 0 is kept
  for (int k = 0; k < grid * grid; k++) {
    const float su = ((float)(k / grid) + 0.5f) / (float)grid, sv = ((float)(k % grid) + 0.5f) / (float)grid;
    const NewtonRes<float> g = newton<float>(rp, eval, su, sv, thr_det, thr_res, (uint32_t)r + seed_base, max_iter);
    evals += g.evals;
    if (!g.hit) continue;
    const float tg = (g.S.x - o.x) * d.x + (g.S.y - o.y) * d.y + (g.S.z - o.z) * d.z;
    if (tg > 0.0f && (!hit || tg < t)) {
      nr = g;
      t = tg;
      hit = true;
    }
  }
  if (grid == 0) hit = nr.hit;  // the single solve reports convergence whatever the sign of t
  iters[i] = evals;
  converged[i] = hit ? 1 : 0;
  if (uv_out) {
    uv_out[2 * i] = hit ? nr.u : u0;
    uv_out[2 * i + 1] = hit ? nr.v : v0;
  }
  if (t_out) t_out[i] = hit ? t : -1.0f;
}
}  // namespace

extern "C" int rtn_solve_pairs_roots(const rtn_scene* s, const float* const rays[6], int64_t n_rays,
                                     const int* pair_ray, const int* pair_patch, const float* seed_uv, int64_t n,
                                     const rtn_trace_params* prm, int grid, int* iters, int* converged, float* uv,
                                     float* t, void* stream) {
  RTN_CHECK(s && s->built && rays && pair_ray && pair_patch && prm && iters && converged && n >= 0, RTN_EINVAL,
            "rtn_solve_pairs: bad arguments");
  RTN_CHECK(grid >= 0 && grid <= 16, RTN_EINVAL, "rtn_solve_pairs_roots: grid outside 0..16");
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_solve_pairs: null ray component %d", k);
  RTN_CHECK(prm->max_iter >= 1 && prm->max_iter <= 64, RTN_EINVAL, "rtn_solve_pairs: max_iter outside 1..64");
  if (n == 0) return RTN_OK;
  hipLaunchKernelGGL(k_solve_pairs, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, (hipStream_t)stream, s->cp.ptr,
                     s->cp_off.ptr, s->pq.ptr, s->corners.ptr, (int)s->n_patches, rays[0], rays[1], rays[2], rays[3],
                     rays[4], rays[5], n_rays, pair_ray, pair_patch, seed_uv, n, prm->max_iter, prm->thr_det,
                     prm->thr_res, prm->seed_base, grid, iters, converged, uv, t);
  RTN_HIP(hipGetLastError());
  if (!stream) RTN_HIP(hipStreamSynchronize(nullptr));
  return RTN_OK;
}

extern "C" int rtn_solve_pairs(const rtn_scene* s, const float* const rays[6], int64_t n_rays, const int* pair_ray,
                               const int* pair_patch, const float* seed_uv, int64_t n, const rtn_trace_params* prm,
                               int* iters, int* converged, float* uv, float* t, void* stream) {
  return rtn_solve_pairs_roots(s, rays, n_rays, pair_ray, pair_patch, seed_uv, n, prm, 0, iters, converged, uv, t,
                               stream);
}
