// rtn_trace: dispatch over the trace forms (exact / wide / persistent / packet).
#include <cstring>

#include "rtn_internal.hpp"
#include "trace_wide.hpp"

namespace rtn {
int trace_exact_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                      const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream);
int trace_wide_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                     const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream, const PairOut* pairs);
int trace_packet_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                       const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream);
}  // namespace rtn

using namespace rtn;

extern "C" void rtn_trace_params_default(rtn_trace_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->mode = RTN_MODE_EXACT;
  p->max_iter = 7;     // NEWTON_MAX_ITERATIONS (constants.h:40)
  p->thr_det = 1e-5f;  // intersect.cu:127
  p->thr_res = 1e-4f;
  p->flat_threshold = 0.f;
  p->cam_k = 0.f;
  p->seed_base = 0;
  p->seed_mode = RTN_SEED_PLAIN;
  p->precision = RTN_PREC_FP32;
  p->cond_thr = 1e4f;
  p->hull_cull = 0;
  p->count = 0;
}

extern "C" int rtn_trace(const rtn_scene* s, const float* const rays[6], int64_t n_rays, const rtn_trace_params* prm,
                         const rtn_hits* hits, rtn_trace_totals* totals, void* stream) {
  RTN_CHECK(s && s->built, RTN_EINVAL, "rtn_trace: scene not built");
  RTN_CHECK(rays && prm && hits && hits->prim, RTN_EINVAL, "rtn_trace: null argument");
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_trace: null ray component %d", k);
  RTN_CHECK(n_rays >= 0 && n_rays < ((int64_t)1 << 32), RTN_EINVAL, "rtn_trace: bad ray count");
  RTN_CHECK(prm->max_iter >= 1 && prm->max_iter <= 64, RTN_EINVAL, "rtn_trace: max_iter outside 1..64");
  RTN_CHECK(prm->hull_cull >= 0 && prm->hull_cull <= 2, RTN_EINVAL, "rtn_trace: hull_cull outside 0..2");
  RTN_CHECK(prm->refill_min >= 0 && prm->refill_min <= 64, RTN_EINVAL, "rtn_trace: refill_min outside 0..64");
  RTN_CHECK(!prm->count || hits->counters, RTN_EINVAL, "rtn_trace: count requested without a counter buffer");
  if (n_rays == 0) return RTN_OK;
  hipStream_t st = (hipStream_t)stream;
  // the scene's control block: work queues and totals, zeroed by a kernel;
  // the kernels always add their stack-overflow count to totals[6]
  TraceCtl* ctl = s->ctl.ptr;
  RTN_TRY(ctl_reset(ctl, st));
  unsigned long long* d_tot = ctl->totals;
  int rc;
  switch (prm->mode) {
    case RTN_MODE_EXACT:
      RTN_CHECK(prm->seed_mode == RTN_SEED_PLAIN && prm->precision == RTN_PREC_FP32 && !prm->hull_cull &&
                    !prm->any_hit && !prm->skip_patch,
                RTN_EUNSUPPORTED, "rtn_trace: the exact form runs the plain method only");
      // the binary traversal pushes at most one entry per level, so with the
      // 32-entry stack (trees of depth < 32) it cannot overflow and the plain
      // kernel runs without the totals reduction unless they are asked for
      rc = trace_exact_scene(s, rays, n_rays, prm, hits, totals || s->max_depth >= 32 ? d_tot : nullptr, st);
      break;
    case RTN_MODE_WIDE:
    case RTN_MODE_PERSISTENT:
      rc = trace_wide_scene(s, rays, n_rays, prm, hits, d_tot, st, nullptr);
      break;
    case RTN_MODE_PACKET:
    case RTN_MODE_PERSISTENT_PACKET:
      rc = trace_packet_scene(s, rays, n_rays, prm, hits, d_tot, st);
      break;
    default:
      set_error("rtn_trace: unknown mode %d", prm->mode);
      rc = RTN_EINVAL;
  }
  if (rc != RTN_OK) return rc;
  // totals requested or a synchronous call: read the totals back and report a
  // stack overflow; an asynchronous call leaves that to rtn_scene_trace_overflow
  if (totals || !st) {
    unsigned long long h[12];
    RTN_HIP(hipMemcpyAsync(h, d_tot, sizeof(h), hipMemcpyDeviceToHost, st));
    RTN_HIP(hipStreamSynchronize(st));
    if (totals) memcpy(totals, h, sizeof(h));
    if (h[6]) {
      set_error("rtn_trace: traversal stack overflowed on %llu rays", h[6]);
      rc = RTN_EOVERFLOW;
    }
  }
  return rc;
}

extern "C" int rtn_scene_trace_overflow(const rtn_scene* s, unsigned long long* count) {
  RTN_CHECK(s && s->built && count, RTN_EINVAL, "rtn_scene_trace_overflow: bad arguments");
  RTN_HIP(hipDeviceSynchronize());
  RTN_HIP(hipMemcpy(count, &s->ctl.ptr->totals[6], sizeof(unsigned long long), hipMemcpyDeviceToHost));
  return RTN_OK;
}

extern "C" int rtn_trace_pairs(const rtn_scene* s, const float* const rays[6], int64_t n_rays,
                               const rtn_trace_params* prm, int all, int* pair_ray, int* pair_patch,
                               float* pair_uv, int64_t cap, int64_t* count_host, void* stream) {
  RTN_CHECK(s && s->built, RTN_EINVAL, "rtn_trace_pairs: scene not built");
  RTN_CHECK(rays && prm && pair_ray && pair_patch && pair_uv && count_host && cap >= 0, RTN_EINVAL,
            "rtn_trace_pairs: bad arguments");
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_trace_pairs: null ray component %d", k);
  RTN_CHECK(n_rays >= 0 && n_rays < ((int64_t)1 << 31), RTN_EINVAL, "rtn_trace_pairs: bad ray count");
  RTN_CHECK(prm->mode == RTN_MODE_WIDE || prm->mode == RTN_MODE_PERSISTENT, RTN_EUNSUPPORTED,
            "rtn_trace_pairs: runs the wide or persistent form");
  RTN_CHECK(prm->seed_mode == RTN_SEED_PLAIN && !prm->count, RTN_EINVAL, "rtn_trace_pairs: plain seeds, no counters");
  RTN_CHECK(prm->max_iter >= 1 && prm->max_iter <= 64, RTN_EINVAL, "rtn_trace_pairs: max_iter outside 1..64");
  *count_host = 0;
  if (n_rays == 0) return RTN_OK;
  hipStream_t st = (hipStream_t)stream;
  DevBuf<int> prim;
  DevBuf<float> tuv;  // t, u, v of the closest hits: the FP64 correction pass of the mixed form needs them
  DevBuf<unsigned long long> cnt;
  RTN_TRY(prim.alloc((size_t)n_rays));
  RTN_TRY(cnt.alloc(1));
  RTN_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), st));
  rtn_hits h{};
  h.prim = prim.ptr;
  if (prm->precision == RTN_PREC_MIXED) {
    RTN_TRY(tuv.alloc(3 * (size_t)n_rays));
    h.t = tuv.ptr;
    h.u = tuv.ptr + n_rays;
    h.v = tuv.ptr + 2 * n_rays;
  }
  PairOut po{pair_ray, pair_patch, reinterpret_cast<float2*>(pair_uv), cnt.ptr, cap, all ? 1 : 0};
  RTN_TRY(ctl_reset(s->ctl.ptr, st));
  RTN_TRY(trace_wide_scene(s, rays, n_rays, prm, &h, nullptr, st, &po));
  unsigned long long c = 0;
  RTN_HIP(hipMemcpyAsync(&c, cnt.ptr, sizeof(c), hipMemcpyDeviceToHost, st));
  RTN_HIP(hipStreamSynchronize(st));
  *count_host = (int64_t)c;
  return RTN_OK;
}
