// Plain method, exact form: calculateIntersect (intersect.cu:20-194) for gfx950.
//
// One lane per ray over the reference's binary BVH (32-byte nodes fetched as
// two 16-byte loads), the traversal stack in LDS (one column per lane, so the
// 64 lanes of a wave hit 64 distinct banks), the bilinear seed and the FP32
// Newton solve in the reference's operation order.  Degree (2,2) and (3,3)
// patches run compile-time-degree evaluators whose de Casteljau triangles
// stay in VGPRs; other degrees use a guarded unrolled evaluator (degree <= 5).
// This translation unit is compiled with -ffp-contract=off, so results equal
// the host reference bit for bit.
#include <cmath>

#include "rtn_exact.hpp"
#include "rtn_internal.hpp"

using namespace rtn;
using namespace rtn::ex;

namespace {

constexpr int TB = 128;  // lanes per block (2 waves)

struct ExactArgs {
  const float4* nodes;  // rtn_bvh_node viewed as 2 x float4
  const float *ox, *oy, *oz, *dx, *dy, *dz;
  int64_t n_rays;
  int W, H;             // > 0: 2-D pixel grid (drop-in), 16 x 8 pixel tiles
  const float* delta;
  float flat, cam_k;
  int max_iter;
  float thr_det, thr_res;
  uint32_t seed_base;
  // scene-form outputs
  int* prim;
  float *t, *u, *v, *dpdu, *dpdv;
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
    Q = v >> 8;
  }
};

__device__ __forceinline__ void load_node(const float4* nodes, int i, float b[6], uint32_t& lf, uint32_t& pc) {
  const float4 a = nodes[2 * i], c = nodes[2 * i + 1];
  b[0] = a.x; b[1] = a.y; b[2] = a.z; b[3] = a.w; b[4] = c.x; b[5] = c.y;
  lf = __float_as_uint(c.z);
  pc = __float_as_uint(c.w);
}

__device__ __forceinline__ float slab6(const float b[6], v3 o, v3 rd) {
  return slab(b[0], b[3], b[1], b[4], b[2], b[5], o, rd);
}

template <int STACK, class Patches>
__device__ __forceinline__ void trace_one(const ExactArgs& A, const Patches& PT, int64_t idx, int tid,
                                          int (*stk)[TB], uint32_t cnt[7]) {
  const v3 o{A.ox[idx], A.oy[idx], A.oz[idx]};
  const v3 d{A.dx[idx], A.dy[idx], A.dz[idx]};
  uint32_t nc = 0, ne = 0, sc = 0, niv = 0, nlv = 0, npt = 0, ovf = 0;
  const v3 rd{1 / d.x, 1 / d.y, 1 / d.z};
  const RayPlane rp = ray_planes(o, d);
  int off = 0;
  float t_max = INFINITY;
  int best = -1;
  float bt = 0.f, bu = 0.f, bv = 0.f;
  v3 bsu{0.f, 0.f, 0.f}, bsv{0.f, 0.f, 0.f};
  float cb[6];
  uint32_t lf, pc;
  load_node(A.nodes, 0, cb, lf, pc);
  float it = slab6(cb, o, rd);
  bool go = it < t_max;
  const uint32_t seed = (uint32_t)idx + A.seed_base;
  while (go) {
    if (pc > 0) {
      nlv++;
      it = slab6(cb, o, rd);
      for (uint32_t i = 0; i < pc; i++) {
        const int pi = (int)(lf + i);
        bool do_newton = true;
        npt++;
        const float4* cp;
        int P, Q;
        PT.get(pi, cp, P, Q);
        const v3 q00 = dehom(ld4(cp[0])), q10 = dehom(ld4(cp[P * (Q + 1)]));
        const v3 q11 = dehom(ld4(cp[P * (Q + 1) + Q])), q01 = dehom(ld4(cp[Q]));
        float gu, gv;
        bilinear_seed(q00, q10, q11, q01, o, d, gu, gv);
        if (A.delta && A.delta[pi] * A.cam_k / it < A.flat) {  // flat-skip (intersect.cu:104-123)
          v4 du, dv, pt;
          patch_eval_rt<RTN_MAX_DEGREE>(cp, P, Q, gu, gv, du, dv, pt);
          v3 S, Su, Sv;
          dehom_ders(pt, du, dv, S, Su, Sv);
          const float resid = fabsf(dot(rp.n1, S) + rp.d1) + fabsf(dot(rp.n2, S) + rp.d2);
          if (resid * A.cam_k / it < A.flat) {
            sc++;
            const float t = dot(sub(S, o), d);
            if (0 < t && t < t_max) {
              t_max = t; best = pi; bt = t; bu = gu; bv = gv; bsu = Su; bsv = Sv;
              do_newton = false;
            }
          }
        }
        if (do_newton) {
          nc++;
          NewtonRes nr;
          if (P == 3 && Q == 3) {
            nr = newton(rp, [&](float uu, float vv, v4& du, v4& dv, v4& pt) { patch_eval<3, 3>(cp, uu, vv, du, dv, pt); },
                        gu, gv, A.thr_det, A.thr_res, seed, A.max_iter);
          } else if (P == 2 && Q == 2) {
            nr = newton(rp, [&](float uu, float vv, v4& du, v4& dv, v4& pt) { patch_eval<2, 2>(cp, uu, vv, du, dv, pt); },
                        gu, gv, A.thr_det, A.thr_res, seed, A.max_iter);
          } else {
            nr = newton(rp,
                        [&](float uu, float vv, v4& du, v4& dv, v4& pt) {
                          patch_eval_rt<RTN_MAX_DEGREE>(cp, P, Q, uu, vv, du, dv, pt);
                        },
                        gu, gv, A.thr_det, A.thr_res, seed, A.max_iter);
          }
          ne += nr.evals;
          if (nr.hit) {
            const float t = (nr.S.x - o.x) * d.x + (nr.S.y - o.y) * d.y + (nr.S.z - o.z) * d.z;
            if (0 < t && t < t_max) {
              best = pi; bt = t; bu = nr.u; bv = nr.v; bsu = nr.Su; bsv = nr.Sv;
              t_max = t;
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
      const float lt = slab6(lb, o, rd);
      const float rt = slab6(rb, o, rd);
      it = fminf(lt, rt);
      if (it >= t_max) {
        if (off == 0) break;
        load_node(A.nodes, stk[--off][tid], cb, lf, pc);
      } else if (lt <= it) {
        if (rt < t_max) {
          if (off < STACK) stk[off++][tid] = (int)lf + 1; else ovf = 1;
        }
#pragma unroll
        for (int k = 0; k < 6; k++) cb[k] = lb[k];
        lf = llf; pc = lpc;
      } else {
        if (lt < t_max) {
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
      si->t_hit = bt;
      si->uv[0] = bu; si->uv[1] = bv;
      si->dpdu[0] = bsu.x; si->dpdu[1] = bsu.y; si->dpdu[2] = bsu.z;
      si->dpdv[0] = bsv.x; si->dpdv[1] = bsv.y; si->dpdv[2] = bsv.z;
      ii = (int)idx;
      A.idx_interact[idx] = ii;
    }
    if (ii != -1) {
      si->wo[0] = d.x * -1.0f; si->wo[1] = d.y * -1.0f; si->wo[2] = d.z * -1.0f;
    }
  }
  if (A.prim) {
    A.prim[idx] = best;
    if (best >= 0) {
      if (A.t) A.t[idx] = bt;
      if (A.u) A.u[idx] = bu;
      if (A.v) A.v[idx] = bv;
      if (A.dpdu) { A.dpdu[3 * idx] = bsu.x; A.dpdu[3 * idx + 1] = bsu.y; A.dpdu[3 * idx + 2] = bsu.z; }
      if (A.dpdv) { A.dpdv[3 * idx] = bsv.x; A.dpdv[3 * idx + 1] = bsv.y; A.dpdv[3 * idx + 2] = bsv.z; }
    }
  }
  if (A.counters) {
    uint32_t* c = A.counters + 8 * idx;
    c[0] = nc; c[1] = ne; c[2] = sc; c[3] = niv; c[4] = nlv; c[5] = npt; c[6] = ovf; c[7] = 0;
  }
  cnt[0] = nc; cnt[1] = ne; cnt[2] = sc; cnt[3] = niv; cnt[4] = nlv; cnt[5] = npt; cnt[6] = ovf;
}

template <int STACK, bool GRID2D, class Patches>
__global__ void __launch_bounds__(TB) k_trace_exact(ExactArgs A, Patches PT) {
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
  uint32_t cnt[7] = {0, 0, 0, 0, 0, 0, 0};
  if (valid) trace_one<STACK>(A, PT, idx, tid, stk, cnt);
  if (A.totals) {  // every lane of the wave takes part in the reduction
#pragma unroll
    for (int k = 0; k < 7; k++) {
      unsigned long long x = cnt[k];
      for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
      if ((tid & 63) == 0 && x) atomicAdd(A.totals + k, x);
    }
  }
}

}  // namespace

namespace rtn {

// Launches the exact kernel over a 1-D ray array on scene buffers.
int trace_exact_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                      const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream) {
  ExactArgs A{};
  A.nodes = reinterpret_cast<const float4*>(s->nodes.ptr);
  A.ox = rays[0]; A.oy = rays[1]; A.oz = rays[2]; A.dx = rays[3]; A.dy = rays[4]; A.dz = rays[5];
  A.n_rays = n;
  A.delta = prm->flat_threshold > 0.f ? s->delta.ptr : nullptr;
  A.flat = prm->flat_threshold;
  A.cam_k = prm->cam_k;
  A.max_iter = prm->max_iter;
  A.thr_det = prm->thr_det;
  A.thr_res = prm->thr_res;
  A.seed_base = prm->seed_base;
  A.prim = hits->prim; A.t = hits->t; A.u = hits->u; A.v = hits->v; A.dpdu = hits->dpdu; A.dpdv = hits->dpdv;
  A.counters = prm->count ? hits->counters : nullptr;
  A.totals = d_totals;
  ScenePatches P{s->cp.ptr, s->cp_off.ptr, s->pq.ptr};
  const unsigned blocks = (unsigned)((n + TB - 1) / TB);
  if (s->max_depth < 32)
    hipLaunchKernelGGL((k_trace_exact<32, false, ScenePatches>), dim3(blocks), dim3(TB), 0, stream, A, P);
  else
    hipLaunchKernelGGL((k_trace_exact<64, false, ScenePatches>), dim3(blocks), dim3(TB), 0, stream, A, P);
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

}  // namespace rtn

extern "C" int rtn_calculate_intersect(const rtn_ray_queue* rays, const rtn_bvh_node* nodes,
                                       const rtn_bezier_patch* patches, rtn_surface_interaction* interact, int W,
                                       int H, float flat_threshold, float v_fov, const float* delta,
                                       unsigned long long* newton_count, unsigned long long* skip_count) {
  RTN_CHECK(rays && nodes && patches && interact && W > 0 && H > 0, RTN_EINVAL,
            "rtn_calculate_intersect: bad arguments");
  RTN_CHECK(rays->idx_interact, RTN_EINVAL, "rtn_calculate_intersect: idx_interact is required");
  ExactArgs A{};
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
  A.seed_base = 0;
  A.inter = interact;
  A.idx_interact = rays->idx_interact;
  unsigned long long* tot = nullptr;  // totals also carry the stack-overflow flag
  RTN_HIP(hipMalloc(&tot, 8 * sizeof(unsigned long long)));
  RTN_HIP(hipMemset(tot, 0, 8 * sizeof(unsigned long long)));
  A.totals = tot;
  DropInPatches P{patches};
  dim3 block(16, 8), grid((W + 15) / 16, (H + 7) / 8);
  hipLaunchKernelGGL((k_trace_exact<32, true, DropInPatches>), grid, block, 0, nullptr, A, P);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  int rc = RTN_OK;
  if (e != hipSuccess) rc = hip_fail(e, "calculateIntersect launch", __FILE__, __LINE__);
  if (tot) {
    unsigned long long h[8];
    if (rc == RTN_OK && hipMemcpy(h, tot, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess) {
      // counters are added to, as the reference's atomicAdd does (intersect.cu:189-190)
      if (newton_count) {
        unsigned long long cur;
        if (hipMemcpy(&cur, newton_count, 8, hipMemcpyDefault) == hipSuccess) {
          cur += h[0];
          (void)hipMemcpy(newton_count, &cur, 8, hipMemcpyDefault);
        }
      }
      if (skip_count) {
        unsigned long long cur;
        if (hipMemcpy(&cur, skip_count, 8, hipMemcpyDefault) == hipSuccess) {
          cur += h[2];
          (void)hipMemcpy(skip_count, &cur, 8, hipMemcpyDefault);
        }
      }
      if (h[6]) {
        set_error("calculateIntersect: traversal stack (32 entries) overflowed on %llu rays", h[6]);
        rc = RTN_EOVERFLOW;
      }
    }
    (void)hipFree(tot);
  }
  return rc;
}
