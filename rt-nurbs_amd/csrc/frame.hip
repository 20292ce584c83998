// Multi-scene batches and frames (no reference counterpart; the reference
// traces one scene, one primary ray per pixel).
//
// A batch merges K built scenes into one set of device arrays: nets, patch
// headers and 4-wide nodes are concatenated, child links relocated (interior
// links by the scene's node base, leaf primitive ranges by its patch base),
// and each scene keeps its root.  A ray carries its scene index, so rays of
// all scenes share one persistent launch and its work queue.
//
// A frame is a fixed sequence over a batch with every buffer allocated when
// it is created: counter reset (a fill kernel), persistent trace of the
// primary rays, shadow + mirror generation from the hits (fixed slots per
// primary ray, empty for a miss), persistent trace of the secondaries
// (shadow rays end at their first hit, empty slots are skipped).  The
// sequence allocates nothing and copies nothing from the host, so it can be
// captured into a hipGraph once and replayed.  The autotuner times the
// frame over the built launch shapes and keeps the fastest.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rtn_internal.hpp"
#include "trace_wide.hpp"

using namespace rtn;

namespace rtn {
void launch_secondaries(const float* const rays[6], const rtn_hits* hits, int64_t n, const float ldir[3], float eps,
                        int want_shadow, int want_mirror, float* const out[6], int* parent, int* kind,
                        unsigned long long* count, const int* ray_scene, int* out_scene, hipStream_t st);
}

struct rtn_batch {
  rtn_scene merged;
  DevBuf<int> roots;
  std::vector<int64_t> patch_base;  // K+1
  std::vector<int> root_h;
  int k = 0;
};

namespace {

__global__ void k_reloc_wide(const WideNode* __restrict__ in, int n, int node_base, int patch_base,
                             WideNode* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  WideNode w = in[i];
  for (int k = 0; k < 4; k++) {
    const int32_t c = w.child[k];
    if (c == (int32_t)0x7fffffff) continue;
    if (c < 0) {
      const uint32_t u = (uint32_t)c;
      w.child[k] = (int32_t)((u & 0xff000000u) | ((u & 0xffffffu) + (uint32_t)patch_base));
    } else {
      w.child[k] = c + node_base;
    }
  }
  out[i] = w;
}

__global__ void k_add_i32(const int* __restrict__ in, int n, int add, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] + add;
}

template <class T>
int copy_into(DevBuf<T>& dst, size_t at, const DevBuf<T>& src, size_t n) {
  if (n) RTN_HIP(hipMemcpy(dst.ptr + at, src.ptr, n * sizeof(T), hipMemcpyDeviceToDevice));
  return RTN_OK;
}

// Device control block of a frame.
struct FrameCtl {
  unsigned int work1[kWorkWords], work2[kWorkWords];  // persistent work queues
  unsigned long long sec_count;   // secondaries written
  unsigned long long totals[12];  // trace totals (overflow at [6])
  unsigned long long count1[12], count2[12];  // work totals of a counting run (rtn_frame_count)
};

__global__ void k_frame_reset(FrameCtl* c) {
  const int t = threadIdx.x;
  if (t < 12) {
    c->totals[t] = 0;
    c->count1[t] = 0;
    c->count2[t] = 0;
  }
  if (t == 12) c->sec_count = 0;
  for (int i = t; i < kWorkWords; i += blockDim.x) {
    c->work1[i] = 0;
    c->work2[i] = 0;
  }
}

__global__ void k_check_scene_ids(const int* __restrict__ ids, int64_t n, int k, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (ids[i] < 0 || ids[i] >= k)) atomicExch(bad, 1);
}

// Every ray's scene index must name a scene of the batch (the kernels index
// the root table with it).  `bad` is a zeroed device word.
int check_scene_ids(const rtn_batch* b, const int* ids, int64_t n, int* bad, hipStream_t st) {
  if (n <= 0) return RTN_OK;
  hipLaunchKernelGGL(k_check_scene_ids, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ids, n, b->k, bad);
  RTN_HIP(hipGetLastError());
  int h = 0;
  RTN_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  RTN_HIP(hipStreamSynchronize(st));
  RTN_CHECK(!h, RTN_EINVAL, "batch: a ray's scene index is outside 0..%d", b->k - 1);
  return RTN_OK;
}

int batch_args(const rtn_batch* b, const float* const rays[6], const int* ray_scene, int64_t n,
               const rtn_trace_params* prm, const rtn_hits* hits, unsigned long long* d_totals, WArgs& A) {
  RTN_CHECK(prm->mode == RTN_MODE_WIDE || prm->mode == RTN_MODE_PERSISTENT || prm->mode == RTN_MODE_PACKET ||
                prm->mode == RTN_MODE_PERSISTENT_PACKET,
            RTN_EUNSUPPORTED, "batch trace: runs the wide, persistent or packet forms");
  RTN_TRY(fill_wide_args(&b->merged, rays, n, prm, hits, d_totals, nullptr, A));
  A.bounds = b->merged.bounds.ptr;
  A.scene_root = b->roots.ptr;
  A.ray_scene = ray_scene;
  return RTN_OK;
}

}  // namespace

extern "C" int rtn_batch_create(rtn_batch** out, rtn_scene* const* scenes, int k) {
  RTN_CHECK(out && scenes && k >= 1, RTN_EINVAL, "rtn_batch_create: bad arguments");
  *out = nullptr;
  int64_t np = 0, ncp = 0, nw = 0;
  for (int i = 0; i < k; i++) {
    RTN_CHECK(scenes[i] && scenes[i]->built, RTN_EINVAL, "rtn_batch_create: scene %d is not built", i);
    np += scenes[i]->n_patches;
    ncp += scenes[i]->n_cp;
    nw += scenes[i]->n_wide;
  }
  RTN_CHECK(np < (1 << 24) && ncp < ((int64_t)1 << 31) && nw < ((int64_t)1 << 31), RTN_EUNSUPPORTED,
            "rtn_batch_create: batch too large (%lld patches)", (long long)np);
  rtn_batch* b = new rtn_batch();
  rtn_scene& m = b->merged;
  auto fail = [&](int rc) { delete b; return rc; };
  int rc;
  if ((rc = m.cp.alloc(ncp)) || (rc = m.cp_off.alloc(np)) || (rc = m.pq.alloc(np)) || (rc = m.bounds.alloc(6 * np)) ||
      (rc = m.delta.alloc(np)) || (rc = m.corners.alloc(3 * np)) || (rc = m.obb.alloc(4 * np)) ||
      (rc = m.seed64.alloc(kSeedTermWords * np)) ||
      (rc = m.wide.alloc(nw)))
    return fail(rc);
  int64_t pb = 0, cb = 0, wb = 0;
  b->patch_base.push_back(0);
  for (int i = 0; i < k; i++) {
    const rtn_scene* s = scenes[i];
    const int spn = (int)s->n_patches, sw = (int)s->n_wide;
    if ((rc = copy_into(m.cp, cb, s->cp, s->n_cp)) || (rc = copy_into(m.pq, pb, s->pq, spn)) ||
        (rc = copy_into(m.bounds, 6 * pb, s->bounds, 6 * (size_t)spn)) || (rc = copy_into(m.delta, pb, s->delta, spn)) ||
        (rc = copy_into(m.corners, 3 * pb, s->corners, 3 * (size_t)spn)) ||
        (rc = copy_into(m.obb, 4 * pb, s->obb, 4 * (size_t)spn)) ||
        (rc = copy_into(m.seed64, kSeedTermWords * pb, s->seed64, kSeedTermWords * (size_t)spn)))
      return fail(rc);
    hipLaunchKernelGGL(k_add_i32, dim3((spn + 255) / 256), dim3(256), 0, nullptr, s->cp_off.ptr, spn, (int)cb,
                       m.cp_off.ptr + pb);
    hipLaunchKernelGGL(k_reloc_wide, dim3((sw + 255) / 256), dim3(256), 0, nullptr, s->wide.ptr, sw, (int)wb, (int)pb,
                       m.wide.ptr + wb);
    if (hipGetLastError() != hipSuccess) return fail(RTN_EHIP);
    b->root_h.push_back((int)wb);
    pb += spn;
    cb += s->n_cp;
    wb += sw;
    b->patch_base.push_back(pb);
    m.max_depth = std::max(m.max_depth, s->max_depth);
  }
  if ((rc = b->roots.upload(b->root_h.data(), k)) || (rc = m.ctl.alloc(1))) return fail(rc);
  if (hipDeviceSynchronize() != hipSuccess) return fail(RTN_EHIP);
  m.n_patches = np;
  m.n_cp = ncp;
  m.n_wide = nw;
  m.built = true;
  b->k = k;
  *out = b;
  return RTN_OK;
}

extern "C" void rtn_batch_destroy(rtn_batch* b) { delete b; }

extern "C" int rtn_batch_patch_base(const rtn_batch* b, int64_t* base) {
  RTN_CHECK(b && base, RTN_EINVAL, "rtn_batch_patch_base: null argument");
  for (int i = 0; i <= b->k; i++) base[i] = b->patch_base[i];
  return RTN_OK;
}

extern "C" int rtn_batch_trace(const rtn_batch* b, const float* const rays[6], const int* ray_scene, int64_t n,
                               const rtn_trace_params* prm, const rtn_hits* hits, rtn_trace_totals* totals,
                               void* stream) {
  RTN_CHECK(b && rays && ray_scene && prm && hits && hits->prim, RTN_EINVAL, "rtn_batch_trace: null argument");
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_batch_trace: null ray component %d", k);
  RTN_CHECK(n >= 0 && n < ((int64_t)1 << 32), RTN_EINVAL, "rtn_batch_trace: bad ray count");
  RTN_CHECK(prm->max_iter >= 1 && prm->max_iter <= 64, RTN_EINVAL, "rtn_batch_trace: max_iter outside 1..64");
  RTN_CHECK(!prm->count || hits->counters, RTN_EINVAL, "rtn_batch_trace: count requested without a counter buffer");
  if (n == 0) return RTN_OK;
  hipStream_t st = (hipStream_t)stream;
  // the batch's control block (allocated with the batch): work queues, totals, flag
  TraceCtl* d = b->merged.ctl.ptr;
  RTN_TRY(ctl_reset(d, st));
  RTN_TRY(check_scene_ids(b, ray_scene, n, &d->flag, st));
  WArgs a;
  int rc = batch_args(b, rays, ray_scene, n, prm, hits, &d->totals[0], a);
  if (rc == RTN_OK) {
    a.work = d->work;
    const bool packet = prm->mode == RTN_MODE_PACKET || prm->mode == RTN_MODE_PERSISTENT_PACKET;
    rc = packet ? launch_packet_kernel(prm->mode == RTN_MODE_PERSISTENT_PACKET, prm->count != 0,
                                       prm->seed_mode == RTN_SEED_LEARNED, n, st, a)
                : launch_wide_kernel(kDefaultShape, prm->mode == RTN_MODE_PERSISTENT, prm->count != 0,
                                     prm->seed_mode == RTN_SEED_LEARNED, n, st, a);
  }
  unsigned long long tot[12] = {0};
  if (rc == RTN_OK) {
    RTN_HIP(hipMemcpyAsync(tot, d->totals, sizeof(tot), hipMemcpyDeviceToHost, st));
    RTN_HIP(hipStreamSynchronize(st));
    if (totals) memcpy(totals, tot, sizeof(tot));
    if (tot[6]) {
      set_error("rtn_batch_trace: traversal stack overflowed on %llu rays", tot[6]);
      rc = RTN_EOVERFLOW;
    }
  }
  return rc;
}

// ------------------------------------------------------------------ frames
struct rtn_frame {
  const rtn_batch* b = nullptr;
  rtn_frame_params fp{};
  const float* rays[6] = {};
  const int* ray_scene = nullptr;
  int64_t n = 0, cap2 = 0;
  DevBuf<int> prim, s_prim, s_scene, s_parent, s_kind;
  DevBuf<float> t, u, v, dpdu, dpdv, s_t, s_u, s_v, s_dpdu, s_dpdv;
  DevBuf<float> s_ray[6];
  DevBuf<FrameCtl> ctl;
  WArgs a1{}, a2{};  // kernel arguments of the primary and secondary launches
  LaunchShape shape = kDefaultShape;
  int refill = 64;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~rtn_frame() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  rtn_hits hits1() const { return rtn_hits{prim.ptr, t.ptr, u.ptr, v.ptr, dpdu.ptr, dpdv.ptr, nullptr}; }
  rtn_hits hits2() const { return rtn_hits{s_prim.ptr, s_t.ptr, s_u.ptr, s_v.ptr, s_dpdu.ptr, s_dpdv.ptr, nullptr}; }
};

namespace {

bool frame_learned(const rtn_frame* f) { return f->fp.trace.seed_mode == RTN_SEED_LEARNED; }

// Enqueues the frame (no allocation, no host copies).  count: the counting
// builds of the same launches (default shape), their work totals going to
// count1 / count2 of the control block instead of the overflow word.
int frame_enqueue(rtn_frame* f, hipStream_t st, bool count = false) {
  FrameCtl* c = f->ctl.ptr;
  hipLaunchKernelGGL(k_frame_reset, dim3(1), dim3(64), 0, st, c);
  RTN_HIP(hipGetLastError());
  const bool learned = frame_learned(f);
  WArgs a1 = f->a1, a2 = f->a2;
  if (count) {
    a1.totals = c->count1;
    a2.totals = c->count2;
    a1.counters = a2.counters = nullptr;
  }
  const LaunchShape sh = count ? kDefaultShape : f->shape;
  if (f->fp.packet_primary) RTN_TRY(launch_packet_kernel(false, count, learned, f->n, st, a1));
  else RTN_TRY(launch_wide_kernel(sh, true, count, learned, f->n, st, a1));
  if (f->cap2 > 0) {
    const rtn_hits h1 = f->hits1();
    float* outs[6];
    for (int k = 0; k < 6; k++) outs[k] = f->s_ray[k].ptr;
    launch_secondaries(f->rays, &h1, f->n, f->fp.light, f->fp.eps, f->fp.shadow, f->fp.mirror, outs,
                       f->s_parent.ptr, f->s_kind.ptr, &c->sec_count, f->ray_scene, f->s_scene.ptr, st);
    RTN_HIP(hipGetLastError());
    if (f->fp.packet_secondary) RTN_TRY(launch_packet_kernel(false, count, learned, f->cap2, st, a2));
    else RTN_TRY(launch_wide_kernel(sh, true, count, learned, f->cap2, st, a2));
  }
  return RTN_OK;
}

int frame_set_refill(rtn_frame* f, int refill) {
  const int v = refill > 0 ? refill : 64;
  f->a1.refill_min = v;  // launch arguments: a captured graph must be captured again
  f->a2.refill_min = v;
  f->refill = v;
  return RTN_OK;
}

void drop_graph(rtn_frame* f) {
  if (f->exec) (void)hipGraphExecDestroy(f->exec);
  if (f->graph) (void)hipGraphDestroy(f->graph);
  f->exec = nullptr;
  f->graph = nullptr;
}

}  // namespace

extern "C" void rtn_frame_params_default(rtn_frame_params* fp) {
  if (!fp) return;
  memset(fp, 0, sizeof(*fp));
  rtn_trace_params_default(&fp->trace);
  fp->trace.mode = RTN_MODE_PERSISTENT;
  fp->trace.max_iter = 8;
  fp->trace.thr_res = 1e-5f;
  fp->trace.thr_det = 1e-5f;
  fp->trace.hull_cull = 2;
  fp->light[0] = 0.3f;
  fp->light[1] = -0.4f;
  fp->light[2] = 1.0f;
  fp->eps = 1e-4f;
  fp->shadow = 1;
  fp->mirror = 1;
  fp->skip_origin = 1;
}

extern "C" int rtn_frame_create(rtn_frame** out, const rtn_batch* b, const float* const rays[6], const int* ray_scene,
                                int64_t n, const rtn_frame_params* fp) {
  RTN_CHECK(out && b && rays && ray_scene && fp && n > 0, RTN_EINVAL, "rtn_frame_create: bad arguments");
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_frame_create: null ray component %d", k);
  RTN_CHECK(n < ((int64_t)1 << 31), RTN_EINVAL, "rtn_frame_create: too many rays");
  RTN_CHECK(fp->trace.mode == RTN_MODE_PERSISTENT, RTN_EINVAL, "rtn_frame_create: frames run the persistent form");
  RTN_CHECK(!fp->trace.count && !fp->trace.any_hit, RTN_EINVAL, "rtn_frame_create: no counters, closest-hit primaries");
  RTN_CHECK((fp->shadow == 0 || fp->shadow == 1) && (fp->mirror == 0 || fp->mirror == 1), RTN_EINVAL,
            "rtn_frame_create: shadow and mirror are 0 or 1");
  RTN_CHECK((fp->packet_primary == 0 || fp->packet_primary == 1) &&
                (fp->packet_secondary == 0 || fp->packet_secondary == 1),
            RTN_EINVAL, "rtn_frame_create: packet_primary and packet_secondary are 0 or 1");
  RTN_CHECK(fp->skip_origin == 0 || fp->skip_origin == 1, RTN_EINVAL, "rtn_frame_create: skip_origin is 0 or 1");
  RTN_CHECK(!fp->trace.skip_patch, RTN_EINVAL,
            "rtn_frame_create: trace.skip_patch is not used by frames (skip_origin sets the secondary rays' skip)");
  const float ll = std::sqrt(fp->light[0] * fp->light[0] + fp->light[1] * fp->light[1] + fp->light[2] * fp->light[2]);
  RTN_CHECK(ll > 0.f, RTN_EINVAL, "rtn_frame_create: zero light direction");
  *out = nullptr;
  RTN_TRY(ctl_reset(b->merged.ctl.ptr, nullptr));
  RTN_TRY(check_scene_ids(b, ray_scene, n, &b->merged.ctl.ptr->flag, nullptr));
  rtn_frame* f = new rtn_frame();
  f->b = b;
  f->fp = *fp;
  for (int k = 0; k < 3; k++) f->fp.light[k] = fp->light[k] / ll;
  for (int k = 0; k < 6; k++) f->rays[k] = rays[k];
  f->ray_scene = ray_scene;
  f->n = n;
  f->cap2 = (int64_t)(fp->shadow + fp->mirror) * n;
  auto fail = [&](int rc) { delete f; return rc; };
  int rc;
  if ((rc = f->prim.alloc(n)) || (rc = f->t.alloc(n)) || (rc = f->u.alloc(n)) || (rc = f->v.alloc(n)) ||
      (rc = f->dpdu.alloc(3 * n)) || (rc = f->dpdv.alloc(3 * n)) || (rc = f->ctl.alloc(1)))
    return fail(rc);
  if (f->cap2 > 0) {
    const int64_t m = f->cap2;
    for (int k = 0; k < 6; k++)
      if ((rc = f->s_ray[k].alloc(m))) return fail(rc);
    if ((rc = f->s_prim.alloc(m)) || (rc = f->s_scene.alloc(m)) || (rc = f->s_parent.alloc(m)) ||
        (rc = f->s_kind.alloc(m)) || (rc = f->s_t.alloc(m)) || (rc = f->s_u.alloc(m)) || (rc = f->s_v.alloc(m)) ||
        (rc = f->s_dpdu.alloc(3 * m)) || (rc = f->s_dpdv.alloc(3 * m)))
      return fail(rc);
  }
  FrameCtl h{};
  const rtn_hits h1 = f->hits1();
  if ((rc = batch_args(b, f->rays, ray_scene, n, &f->fp.trace, &h1, f->ctl.ptr->totals, f->a1))) return fail(rc);
  f->a1.work = f->ctl.ptr->work1;
  if (f->cap2 > 0) {
    const rtn_hits h2 = f->hits2();
    const float* srays[6];
    for (int k = 0; k < 6; k++) srays[k] = f->s_ray[k].ptr;
    if ((rc = batch_args(b, srays, f->s_scene.ptr, f->cap2, &f->fp.trace, &h2, f->ctl.ptr->totals, f->a2)))
      return fail(rc);
    f->a2.ray_kind = f->s_kind.ptr;
    if (fp->skip_origin) {  // a secondary ray skips its primary ray's hit patch
      f->a2.skip_prim = f->prim.ptr;
      f->a2.skip_parent = f->s_parent.ptr;
    }
    f->a2.work = f->ctl.ptr->work2;
  }
  if (hipMemcpy(f->ctl.ptr, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) return fail(RTN_EHIP);
  *out = f;
  return RTN_OK;
}

extern "C" void rtn_frame_destroy(rtn_frame* f) { delete f; }

extern "C" int rtn_frame_run(rtn_frame* f, void* stream) {
  RTN_CHECK(f, RTN_EINVAL, "rtn_frame_run: null frame");
  hipStream_t st = (hipStream_t)stream;
  if (f->exec) {
    RTN_HIP(hipGraphLaunch(f->exec, st));
    return RTN_OK;
  }
  return frame_enqueue(f, st);
}

extern "C" int rtn_frame_capture(rtn_frame* f, void* stream) {
  RTN_CHECK(f && stream, RTN_EINVAL, "rtn_frame_capture: needs a frame and a non-default stream");
  hipStream_t st = (hipStream_t)stream;
  drop_graph(f);
  RTN_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
  const int rc = frame_enqueue(f, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  RTN_TRY(rc);
  RTN_HIP(e);
  f->graph = g;
  RTN_HIP(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
  return RTN_OK;
}

extern "C" int rtn_frame_status(const rtn_frame* f, rtn_frame_status_t* out) {
  RTN_CHECK(f && out, RTN_EINVAL, "rtn_frame_status: null argument");
  RTN_HIP(hipDeviceSynchronize());
  FrameCtl h;
  RTN_HIP(hipMemcpy(&h, f->ctl.ptr, sizeof(h), hipMemcpyDeviceToHost));
  memset(out, 0, sizeof(*out));
  out->n_primary = f->n;
  out->n_secondary = (int64_t)h.sec_count;
  out->n_secondary_slots = f->cap2;
  out->overflow = h.totals[6];
  out->shape_lanes = f->shape.wb;
  out->shape_stack = f->shape.wstk;
  out->shape_blocks_per_cu = f->shape.blocks_per_cu;
  out->captured = f->exec != nullptr;
  out->refill_min = f->refill;
  out->packet = (f->fp.packet_primary ? 1 : 0) | (f->fp.packet_secondary ? 2 : 0);
  return RTN_OK;
}

extern "C" int rtn_frame_count(rtn_frame* f, rtn_trace_totals* primary, rtn_trace_totals* secondary) {
  RTN_CHECK(f && primary && secondary, RTN_EINVAL, "rtn_frame_count: null argument");
  RTN_TRY(frame_enqueue(f, nullptr, true));
  RTN_HIP(hipDeviceSynchronize());
  FrameCtl h;
  RTN_HIP(hipMemcpy(&h, f->ctl.ptr, sizeof(h), hipMemcpyDeviceToHost));
  static_assert(sizeof(rtn_trace_totals) == sizeof(h.count1), "totals layout");
  memcpy(primary, h.count1, sizeof(h.count1));
  memcpy(secondary, h.count2, sizeof(h.count2));
  RTN_CHECK(!h.count1[6] && !h.count2[6], RTN_EOVERFLOW, "rtn_frame_count: traversal stack overflowed");
  return RTN_OK;
}

extern "C" int rtn_frame_results(const rtn_frame* f, rtn_hits* primary, rtn_hits* secondary, const int** parent,
                                 const int** kind, const float** secondary_rays6) {
  RTN_CHECK(f, RTN_EINVAL, "rtn_frame_results: null frame");
  if (primary) *primary = f->hits1();
  if (secondary) *secondary = f->hits2();
  if (parent) *parent = f->s_parent.ptr;
  if (kind) *kind = f->s_kind.ptr;
  if (secondary_rays6)
    for (int k = 0; k < 6; k++) secondary_rays6[k] = f->s_ray[k].ptr;
  return RTN_OK;
}

extern "C" int rtn_frame_set_shape(rtn_frame* f, int lanes, int stack, int blocks_per_cu) {
  RTN_CHECK(f && blocks_per_cu >= 1 && blocks_per_cu <= 32, RTN_EINVAL, "rtn_frame_set_shape: bad arguments");
  const LaunchShape sh{lanes, stack, blocks_per_cu};
  RTN_CHECK(shape_available(sh, true, false, frame_learned(f)), RTN_EUNSUPPORTED,
            "rtn_frame_set_shape: %d lanes x %d stack entries is not built for this form", lanes, stack);
  f->shape = sh;
  drop_graph(f);
  return RTN_OK;
}

extern "C" int rtn_frame_set_refill(rtn_frame* f, int refill_min) {
  RTN_CHECK(f && refill_min >= 0 && refill_min <= 64, RTN_EINVAL, "rtn_frame_set_refill: bad arguments");
  drop_graph(f);
  return frame_set_refill(f, refill_min);
}

extern "C" int rtn_frame_set_packet(rtn_frame* f, int packet_primary, int packet_secondary) {
  RTN_CHECK(f && (packet_primary == 0 || packet_primary == 1) && (packet_secondary == 0 || packet_secondary == 1),
            RTN_EINVAL, "rtn_frame_set_packet: bad arguments");
  drop_graph(f);
  f->fp.packet_primary = packet_primary;
  f->fp.packet_secondary = packet_secondary;
  return RTN_OK;
}

// Autotuner: every built launch shape (lanes per block x stack entries) at
// 2, 4, 8 and 12 resident blocks per compute unit with whole-wave refill, then
// the best shape at refill thresholds 1, 16 and 32, then the packet form for
// the primary launch and for both launches; each candidate runs the
// frame once to warm up and `reps` times under HIP events; candidates whose
// stack overflows are rejected.  The fastest is kept (and any graph dropped).
extern "C" int rtn_frame_autotune(rtn_frame* f, int reps, rtn_tune_entry* table, int table_cap, int* n_tried,
                                  void* stream) {
  RTN_CHECK(f && reps >= 1, RTN_EINVAL, "rtn_frame_autotune: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  drop_graph(f);
  struct Cand { LaunchShape sh; int refill; int pp, ps; };
  std::vector<Cand> cand;
  const int bpcs[4] = {2, 4, 8, 12};  // capped at the resident blocks when launched
  const bool learned = frame_learned(f);
  for (int b = 0; b < 4; b++) {
    cand.push_back({{kDefaultShape.wb, kDefaultShape.wstk, bpcs[b]}, 64, 0, 0});
    if (!learned)
      for (int i = 0; i < kNumTunedShapes; i++)
        cand.push_back({{kTunedShapes[i].wb, kTunedShapes[i].wstk, bpcs[b]}, 64, 0, 0});
  }
  hipEvent_t e0, e1;
  RTN_HIP(hipEventCreate(&e0));
  RTN_HIP(hipEventCreate(&e1));
  Cand best{f->shape, f->refill, f->fp.packet_primary, f->fp.packet_secondary};
  float best_ms = INFINITY;
  int tried = 0;
  int rc = RTN_OK;
  auto run_one = [&](const Cand& c) -> float {
    f->shape = c.sh;
    f->fp.packet_primary = c.pp;
    f->fp.packet_secondary = c.ps;
    if ((rc = frame_set_refill(f, c.refill))) return -2.f;
    if ((rc = frame_enqueue(f, st))) return -2.f;
    FrameCtl h;
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(&h, f->ctl.ptr, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = RTN_EHIP;
      return -2.f;
    }
    if (h.totals[6] != 0) return -1.f;  // stack overflow: rejected
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && rc == RTN_OK; r++) rc = frame_enqueue(f, st);
    (void)hipEventRecord(e1, st);
    if (rc || hipEventSynchronize(e1) != hipSuccess) {
      rc = rc ? rc : RTN_EHIP;
      return -2.f;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms / (float)reps;
  };
  auto record = [&](const Cand& c, float ms) {
    if (table && tried < table_cap) {
      table[tried].lanes = c.sh.wb;
      table[tried].stack = c.sh.wstk;
      table[tried].blocks_per_cu = c.sh.blocks_per_cu;
      table[tried].refill_min = c.refill;
      table[tried].ms = ms;
      table[tried].packet_primary = c.pp;
      table[tried].packet_secondary = c.ps;
    }
    tried++;
    if (ms > 0.f && ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  };
  for (const Cand& c : cand) {
    const float ms = run_one(c);
    if (ms < -1.5f) break;
    record(c, ms);
  }
  if (rc == RTN_OK && best_ms < INFINITY) {
    const LaunchShape sh = best.sh;
    for (int r : {1, 16, 32}) {
      const Cand c{sh, r, 0, 0};
      const float ms = run_one(c);
      if (ms < -1.5f) break;
      record(c, ms);
    }
  }
  // then the packet form for the primary launch, and for both launches
  if (rc == RTN_OK && best_ms < INFINITY) {
    const Cand base = best;
    for (int ps = 0; ps <= 1; ps++) {
      const Cand c{base.sh, base.refill, 1, ps};
      const float ms = run_one(c);
      if (ms < -1.5f) break;
      record(c, ms);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  f->shape = best.sh;
  f->fp.packet_primary = best.pp;
  f->fp.packet_secondary = best.ps;
  const int rc2 = frame_set_refill(f, best.refill);
  if (n_tried) *n_tried = tried;
  RTN_TRY(rc);
  RTN_TRY(rc2);
  RTN_CHECK(best_ms < INFINITY, RTN_EOVERFLOW, "rtn_frame_autotune: every launch shape overflowed its stack");
  return RTN_OK;
}
