// Arguments and launch interface of the fast trace kernels (wide and
// persistent forms), shared by the single-scene entry points and the
// multi-scene frame (batched persistent launch, graph capture, autotuning).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"
#include "rtn_internal.hpp"

namespace rtn {

constexpr int WB_DEFAULT = 128;   // lanes per block (2 waves) of the persistent form
constexpr int WSTK_DEFAULT = 20;  // stack entries per lane: 20 KB of LDS per 128-lane block
// The one-ray-per-lane form runs single-wave blocks: a finished wave frees its
// slot at once instead of waiting for the other waves of its block (+5%).
constexpr int WB_ONE_RAY = 64;

// Work queues of the persistent forms.  One counter would serialise every
// wave's atomic on one address (the limit, not the tracing, at tens of
// thousands of waves); instead kQueues counters, 64 bytes apart, each hand
// out the 64-ray chunks c with c mod kQueues == its index, in order: position
// p of queue q is ray ((q + kQueues * (p / 64)) * 64 + p mod 64).  A wave
// starts on queue blockIdx mod kQueues and moves on when its queue runs past
// the last ray.  args.work holds kWorkWords zeroed words.
constexpr int kQueues = 8;
constexpr int kQueueStride = 16;
constexpr int kWorkWords = kQueues * kQueueStride;
static_assert(kWorkWords <= kCtlWorkWords, "the work queues fit the control block");
__host__ __device__ inline int64_t queue_ray(int q, uint32_t p) {
  return ((int64_t)q + (int64_t)kQueues * (p >> 6)) * 64 + (p & 63);
}

// Kernel arguments, passed by value.
struct WArgs {
  const WideNode* nodes;
  const float4* cp;
  const int* cp_off;
  const int* pq;
  const float4* corners;  // dehomogenised corners q00,q10,q11,q01 (12 floats per patch)
  const float4* obb;      // oriented box of each net (4 float4 per patch)
  const double* seed64;   // FP64 patch terms of the bilinear seed (kSeedTermWords per patch)
  const float *ox, *oy, *oz, *dx, *dy, *dz;
  int64_t n_rays;
  int max_iter;
  float thr_det, thr_res;
  uint32_t seed_base;
  int hull_cull;
  float cull_margin;
  int mixed;
  int any_hit;
  int stall_exit;         // abandon Newton solves stuck against the patch boundary
  int refill_min;         // persistent: lanes idle before a wave takes new rays
  float cond_thr;
  float t_tol;            // mixed: predicted relative t error that also selects the FP64 step
  // learned form: BF16 MFMA tiles and biases of the seed network, patch boxes
  const mlp::bf16x8* mlp_tiles;
  const float* mlp_bias;
  const float* bounds;
  int learned_iters;
  int fallback_iters;
  int seed_share;
  int mlp_kind;           // the network's feature kind (mlp::FeatureKind)
  float hit_gate;         // below this hit probability: the plain seed, no learned attempt
  float hit_skip;         // below this hit probability: the pair is not solved
  int learned_min_lanes;  // fewer lanes with a pair: no network, plain seeds
  // training pairs: every converged solve with t > 0 is appended (pair_all:
  // every solve, uv = (-1, -1) for the others)
  int* pair_ray;
  int* pair_patch;
  float2* pair_uv;
  unsigned long long* pair_count;
  int64_t pair_cap;
  int pair_all;
  // multi-scene batch: root node of each scene and the scene of each ray
  const int* scene_root;
  const int* ray_scene;
  const int* ray_kind;    // per ray, 0 = occlusion (any-hit) ray; null = all as any_hit says
  // per ray, a patch not tested: skip_prim[skip_parent[r]] (skip_parent null:
  // skip_prim[r]); skip_prim null = none
  const int* skip_prim;
  const int* skip_parent;
  int* prim;
  float *t, *u, *v, *dpdu, *dpdv;
  uint32_t* counters;
  unsigned long long* totals;
  unsigned int* work;     // persistent: the work queues (kWorkWords words)
};

// Launch shape: lanes per block (64, 128 or 256: 1, 2 or 4 waves), traversal
// stack entries per lane (LDS = 8 bytes x entries x lanes per block) and, for
// the persistent form, resident blocks per compute unit.
struct LaunchShape {
  int wb, wstk, blocks_per_cu;
};

extern const LaunchShape kDefaultShape;
extern const LaunchShape kTunedShapes[];
extern const int kNumTunedShapes;

bool shape_available(const LaunchShape& sh, bool persistent, bool count, bool learned);
int fill_wide_args(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                   const rtn_hits* hits, unsigned long long* d_totals, const PairOut* pairs, WArgs& A);
// n_max bounds the rays the launch may see (grid size).
int launch_wide_kernel(const LaunchShape& sh, bool persistent, bool count, bool learned, int64_t n_max,
                       hipStream_t stream, const WArgs& args);
// The packet form (one wave traverses for its 64 rays; persistent: 64-ray
// chunks from args.work); args.bounds must be set.
int launch_packet_kernel(bool persistent, bool count, bool learned, int64_t n_max, hipStream_t stream,
                         const WArgs& args);
// Both launchers end with the mixed-precision correction pass (k_fix64) when
// args.mixed is set: the FP64 step of every ray whose closest hit is
// ill-conditioned.
int launch_fix64(int64_t n_max, hipStream_t stream, const WArgs& args);
// Zeroes a trace control block with a kernel (no memset node when captured).
int ctl_reset(TraceCtl* ctl, hipStream_t stream);

}  // namespace rtn
