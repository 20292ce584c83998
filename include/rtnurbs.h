/*
 * rtnurbs.h -- C ABI of the MI355X (gfx950) NURBS ray-tracing path.
 *
 * Plain C: POD structs, raw pointers, sizes and int status codes.  Every
 * entry point names the reference interface (Frostbite16/rt-nurbs, file:line)
 * whose behaviour it provides.  The C++ drop-in layer with the reference's own
 * names and signatures is include/rtnurbs.hpp, built on these functions.
 *
 * Conventions
 *   - Status: RTN_OK (0) or a negative RTN_E* code; rtn_last_error() gives the
 *     message of the calling thread's last failure.
 *   - "device" pointers are HIP device (or managed) allocations, e.g. from
 *     rtn_malloc; "host" pointers are ordinary host memory.
 *   - stream arguments are hipStream_t passed as void*; NULL is the default
 *     stream.  Functions without a stream argument are synchronous, like the
 *     reference wrappers (intersect.cu:218, generateRaysKernel.cu:61,66).
 */
#ifndef RTNURBS_H
#define RTNURBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RTN_OK = 0,
  RTN_EINVAL = -1,      /* invalid argument (sizes, degrees, null pointers)  */
  RTN_EHIP = -2,        /* HIP runtime error                                 */
  RTN_ENOMEM = -3,      /* allocation failure                                */
  RTN_EUNSUPPORTED = -4,/* outside the supported envelope (e.g. degree > 5)  */
  RTN_EOVERFLOW = -5    /* traversal stack overflowed on some ray            */
};

#define RTN_MAX_DEGREE 5   /* reference: NURBS_MAX_DEGREE = 4 (constants.h:31) */

/* ------------------------------------------------------------------ layouts
 * Bit-compatible with the reference's structs so reference-side buffers can
 * be passed unchanged. */
typedef struct { float v[4]; } rtn_vec4;                 /* MathTypes::vector4f, math_types.h:36 */
typedef struct {                                         /* BVHNode, bvh.h:13-17 (32 bytes)       */
  float p_min_x, p_min_y, p_min_z, p_max_x, p_max_y, p_max_z;
  uint32_t left_first, prim_count;
} rtn_bvh_node;
typedef struct {                                         /* bezierPatch, bezier_patch.h:10-16     */
  rtn_vec4* control_point;
  float u0, u1, v0, v1;
  int p, q;
  rtn_vec4* heap_pointer;
} rtn_bezier_patch;
typedef struct {                                         /* surfaceInteraction, gpu/interact.h:6-21 */
  float t_hit;
  float wo[3];
  float uv[2];
  float dpdu[3], dpdv[3], dndu[3], dndv[3];
} rtn_surface_interaction;
typedef struct {                                         /* SoAQueue::ray_queue, soa.h:6-14       */
  float *origin_x, *origin_y, *origin_z;
  float *direction_x, *direction_y, *direction_z;
  int* idx_interact;
} rtn_ray_queue;
typedef struct {                                         /* ParametricSurf::nurbsSurf, nurbs.h:17-22 */
  float *u_knots, *v_knots;                              /* n+p+2 and m+q+2 knots                  */
  rtn_vec4* control_point;                               /* (n+1)(m+1) homogeneous, u-major        */
  int n, m;                                              /* indices of the last control point      */
  int p, q;
} rtn_nurbs_surf;
typedef struct { float m[16]; } rtn_mat4;                /* MathTypes::matrix4x4f, column major    */
typedef struct {                                         /* camera_config, camera.h:13-27 (472 B)  */
  rtn_mat4 world_from_camera;
  rtn_mat4 screen_from_camera;
  rtn_mat4 NDC_from_screen;
  rtn_mat4 raster_from_NDC;
  rtn_mat4 raster_from_screen;
  rtn_mat4 screen_from_raster;
  rtn_mat4 camera_from_raster;
  int screen_resolution_x, screen_resolution_y;
  float screen_window[4];                                /* MathExtra::bounds2f: min x, min y,
                                                            max x, max y                           */
} rtn_camera_config;
enum { RTN_PROJ_ORTHOGRAPHIC = 0, RTN_PROJ_PERSPECTIVE = 1 }; /* enum projection, camera.h:37       */

/* Error message of the calling thread's last failure, and the library
 * version (no reference counterpart: the reference prints errors and goes on,
 * memory.h:32-57). */
const char* rtn_last_error(void);
int rtn_version(void);

/* ------------------------------------------------------------------ memory
 * CudaDevAlloc / CudaAllocManaged / CudaFree (memory.h:32-57).  Unlike the
 * reference these return a status instead of printing and continuing. */
int rtn_malloc(void** ptr, size_t bytes);
int rtn_malloc_managed(void** ptr, size_t bytes);
int rtn_free(void* ptr);
/* CudaHostAlloc / CudaCpuFree (memory.h:23-28, :60-65): pinned host memory. */
int rtn_malloc_host(void** ptr, size_t bytes);
int rtn_free_host(void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device, 3 inferred */
int rtn_memcpy(void* dst, const void* src, size_t bytes, int kind);
int rtn_memset(void* dst, int value, size_t bytes);
int rtn_device_synchronize(void);
int rtn_stream_create(void** stream);
int rtn_stream_destroy(void* stream);
int rtn_stream_synchronize(void* stream);
/* HIP events for timing kernels on a stream (the reference times with host
 * std::chrono around launch + sync, benchmark.cpp:238-244). */
int rtn_event_create(void** event);
int rtn_event_destroy(void* event);
int rtn_event_record(void* event, void* stream);
int rtn_event_elapsed_ms(void* start, void* stop, float* ms);

/* ------------------------------------------------------------------ camera
 * Transform::lookAt (transform.h:46-60) + CameraOps::init_camera and
 * getPerspectiveCamera (camera.h:63-141), host-side as in the reference.
 * fov_deg in degrees (camera.h:69 applies radians()). */
int rtn_camera_perspective(const float eye[3], const float look[3], const float up[3], int W, int H,
                           float fov_deg, float znear, float zfar, float world_from_camera[16],
                           float camera_from_raster[16]);

/* The reference's camera set-up on its full camera_config, host-side and in
 * its operation order: Transform::lookAt (transform.h:46-60),
 * CameraOps::init_camera (camera.h:123-141; it calls initCameraTransforms,
 * :89-119, and leaves screen_from_camera and camera_from_raster zero),
 * getPerspectiveCamera (:63-85; fov in degrees, radians() applied) and
 * getOrthographicCamera (:42-59). */
int rtn_lookat(const float pos[3], const float look[3], const float up[3], float world_from_camera[16]);
int rtn_camera_init(const float world_from_camera[16], int res_x, int res_y, rtn_camera_config* out);
int rtn_camera_get_perspective(const rtn_camera_config* in, float fov_deg, float z_near, float z_far,
                               rtn_camera_config* out);
int rtn_camera_get_orthographic(const rtn_camera_config* in, float z_near, float z_far, rtn_camera_config* out);

/* calculateRaysWrapper(ray_queue&, projection, camera_config&)
 * (generateRaysKernel.h:6, .cu:39-67): one ray through the centre of every
 * pixel of the camera's screen_resolution_x x screen_resolution_y raster,
 * ray index y*W + x, idx_interact = -1.  RTN_PROJ_PERSPECTIVE:
 * calculateRay<perspective_projection> (camera.h:162-174); RTN_PROJ_ORTHOGRAPHIC:
 * calculateRay<orthographic_projection> (camera.h:150-160): origin the pixel's
 * camera-space point taken to world space, direction the camera's z axis
 * (world_from_camera.m[8..10]).  Synchronous when stream is NULL. */
int rtn_rays_generate(const rtn_ray_queue* rays, int projection, const rtn_camera_config* camera, void* stream);

/* calculateRaysWrapper<perspective> (generateRaysKernel.cu:39-67; ray math
 * camera.h:162-174).  Writes W*H*spp rays into the device SoA queue and
 * idx_interact = -1.  flags 0: pixel centres (x+0.5, y+0.5), ray index
 * (y*W + x)*spp + s, the reference's rays.  RTN_RAYS_JITTER: per-ray raster
 * offsets from a seeded hash of the ray index (multi-sample configurations).
 * RTN_RAYS_TILED: ray index = (tiled position of the pixel)*spp + s (see the
 * enum), the same rays in a layout that keeps a wave's rays together. */
enum {
  RTN_RAYS_JITTER = 1,  /* jittered sample positions (else pixel centres)                        */
  RTN_RAYS_TILED = 2    /* ray layout by 8x8-pixel tiles (row-major tiles, Morton order inside),
                           so a wave's 64 rays cover a compact pixel block; W, H multiples of 8   */
};
int rtn_rays_perspective(const float world_from_camera[16], const float camera_from_raster[16], int W, int H,
                         int spp, int flags, uint32_t jitter_seed, const rtn_ray_queue* rays, void* stream);

/* ----------------------------------------------------------- scene (GPU)
 * Scene assembly for the trace path: NURBS surfaces are decomposed on the GPU
 * into rational Bezier patches (getPatch, bezier_patch.h:87-139, A5.7 of
 * nurbs.h:144-250), bounded (bezier_patch.h:154-169) and indexed by the
 * exhaustive-SAH BVH of buildBVH (bvh.h:148-179), built on the GPU with the
 * reference's node numbering and primitive order. */
typedef struct rtn_scene rtn_scene;

int rtn_scene_create(rtn_scene** scene);
int rtn_scene_destroy(rtn_scene* scene);
/* Copies one nurbsSurf (nurbs.h:17-22): U[n+p+2], V[m+q+2], Pw[(n+1)(m+1)]
 * homogeneous, u-major.  Degrees 1..RTN_MAX_DEGREE. */
int rtn_scene_add_nurbs(rtn_scene* scene, int n, int p, const float* U, int m, int q, const float* V,
                        const float* Pw);
/* Adds already-decomposed patches of one degree: cp[np*(p+1)*(q+1)*4], each
 * net row-major homogeneous as getPatch writes it (bezier_patch.h:119-134).
 * No reference counterpart: the reference decomposes one surface per BVH. */
int rtn_scene_add_patches(rtn_scene* scene, int np, int p, int q, const float* cp);
/* Runs decomposition, bounds and the BVH build on the device. */
int rtn_scene_build(rtn_scene* scene);

/* Scene files (no reference counterpart: the reference builds its scenes in
 * code, main.cpp:70-98).  A file holds the surfaces and patch groups added to
 * an unbuilt scene, exactly as rtn_scene_add_nurbs / rtn_scene_add_patches
 * took them, and optionally a perspective camera; little-endian binary,
 * magic "RTNSCENE", version 1 (layout in rt-nurbs_amd/csrc/scene_io.hip).
 * Host-only calls.  rtn_scene_load returns a new, unbuilt scene; every count is
 * checked against the file size and every surface through rtn_scene_add_nurbs. */
typedef struct {
  float eye[3], look[3], up[3];
  float fov_deg;
  int width, height;
} rtn_camera_desc;
int rtn_scene_save(const rtn_scene* scene, const rtn_camera_desc* camera_or_null, const char* path);
int rtn_scene_load(const char* path, rtn_scene** scene, rtn_camera_desc* camera_or_null, int* has_camera);

/* Build options.  RTN_BVH_EXACT_SAH is buildBVH's exhaustive SAH (bvh.h:85-143,
 * reference numbering; quadratic in the primitives of a node); RTN_BVH_BINNED_SAH
 * is a binned-SAH build for large scenes (no reference counterpart): `bins`
 * planes per axis, leaves of up to max_leaf primitives where splitting does not
 * pay.  Both feed the same node layout and the same trace forms. */
enum { RTN_BVH_EXACT_SAH = 0, RTN_BVH_BINNED_SAH = 1 };
typedef struct {
  int bvh;                /* RTN_BVH_*                                          */
  int bins;               /* binned builder: planes per axis (default 32)       */
  int max_leaf;           /* binned builder: largest leaf (default 4, <= 128)   */
  int reserved[5];
} rtn_build_params;
void rtn_build_params_default(rtn_build_params* bp);
int rtn_scene_build_ex(rtn_scene* scene, const rtn_build_params* bp);

typedef struct {
  int64_t n_surfaces, n_patches, n_cp, n_nodes, n_wide_nodes, max_depth, n_leaves;
  int64_t deg_hist[RTN_MAX_DEGREE + 1][RTN_MAX_DEGREE + 1]; /* patches per (p,q) */
  double build_ms_decompose, build_ms_bvh, build_ms_wide;
} rtn_scene_stats;
int rtn_scene_stats_get(const rtn_scene* scene, rtn_scene_stats* out);

/* Copies scene arrays to host memory (any pointer may be NULL):
 *   patch_cp   float[n_cp*4]     nets in BVH order (patch-major, row-major)
 *   patch_off  int[n_patches]    float4 offset of each net
 *   patch_pq   int[2*n_patches]  (p,q)
 *   patch_uv   float[4*n_patches] (u0,u1,v0,v1) of the patch on its surface
 *   patch_id   int[n_patches]    original id: surface offset + i*nbV + j
 *   knot_idx   int[2*n_patches]  span-start knot index (u,v) of the patch
 *   nodes      rtn_bvh_node[n_nodes] reference layout (bvh.h:13-17)
 *   delta      float[n_patches]  patchDelta flatness (main.cpp:18-43), BVH order */
int rtn_scene_download(const rtn_scene* scene, float* patch_cp, int* patch_off, int* patch_pq, float* patch_uv,
                       int* patch_id, int* knot_idx, rtn_bvh_node* nodes, float* delta);

/* Silhouette rays (no reference counterpart; the mixed-precision workload).
 * Ray i aims at a random point of a random patch of the built scene: patch k
 * and (u,v) uniform, S its point, N = unit(S_u x S_v), T a random unit
 * tangent, theta ~ U(0, max_deg) degrees; d = unit(T cos theta - N sin theta),
 * origin S - t0 d with t0 ~ U(t0_lo, t0_hi).  Deterministic in (i, seed).
 * aim_patch[n] (BVH position) and aim[4n] = (u, v, theta radians, t0) are
 * optional device outputs. */
int rtn_rays_silhouette(const rtn_scene* scene, int64_t n, uint32_t seed, float max_deg, float t0_lo, float t0_hi,
                        const rtn_ray_queue* rays, int* aim_patch, float* aim, void* stream);

/* --------------------------------------------------------------- tracing
 * Plain method (intersect.cu:33-193): per-ray ordered closest-hit traversal,
 * bilinear seed (bezier_patch.h:273-327), FP32 Newton (bezier_patch.h:198-266).
 * With RTN_MODE_EXACT the FP32 operation order is the reference's (no FMA),
 * so results equal the host reference bit for bit. */
enum {
  RTN_MODE_EXACT = 0,     /* reference binary BVH, reference op order            */
  RTN_MODE_WIDE = 1,      /* 4-wide SoA BVH, FMA arithmetic, degree-binned Newton */
  RTN_MODE_PERSISTENT = 2,/* persistent work-queue kernel over the wide BVH       */
  RTN_MODE_PACKET = 3,    /* one wave traverses the wide BVH for its 64 rays together
                             (coherent rays: primary rays in the tiled layout);
                             single scenes                                        */
  RTN_MODE_PERSISTENT_PACKET = 4 /* persistent waves taking 64-ray chunks from a work
                             queue, each chunk traced as a packet                 */
};
enum {
  RTN_SEED_PLAIN = 0,     /* the reference's bilinear seed                      */
  RTN_SEED_LEARNED = 1    /* MLP seed, plain seed as fallback                    */
};
enum {
  RTN_PREC_FP32 = 0,      /* FP32 Newton                                         */
  RTN_PREC_MIXED = 1      /* FP32 Newton started from the plain seed evaluated in FP64 (the FP64
                             reference's seed; the FP32 seed's accept test runs on rounding
                             noise, see rtn_fast.hpp seed64); after the trace, the FP64
                             correction (up to four FP64 Newton steps, each kept only if it stays
                             in [0,1]^2 and lowers the residual) on each ray's closest hit whose
                             cond(J) > cond_thr or whose predicted relative t error (the change
                             one more Newton step would make) exceeds t_tol.  Fast forms; needs
                             the t, u and v outputs                                      */
};

typedef struct {
  int mode;               /* RTN_MODE_*                                          */
  int max_iter;           /* reference NEWTON_MAX_ITERATIONS = 7 (constants.h:40)*/
  float thr_det;          /* reference 1e-5 (intersect.cu:127)                   */
  float thr_res;          /* reference 1e-4 (intersect.cu:127)                   */
  float flat_threshold;   /* reference flat-skip threshold in pixels (0 = off)   */
  float cam_k;            /* 0.5*H/tan(vfov/2) (intersect.cu:212)                */
  uint32_t seed_base;     /* Newton RNG seed of ray r = r + seed_base            */
  int seed_mode;          /* RTN_SEED_*                                          */
  int precision;          /* RTN_PREC_*                                          */
  float cond_thr;         /* condition-number threshold of the FP64 step         */
  int hull_cull;          /* exact rejection at leaf entry, before Newton (fast forms): 0 off; 1 the
                             ray's line against the patch's oriented box, then a sign test of the net
                             on each ray plane; 2 the box, then the projected-hull test (8
                             directions); both need positive weights: nets with a weight <= 0 are
                             never rejected                                       */
  int count;              /* 1: fill the per-ray counters                        */
  int learned_iters;      /* learned seed: Newton iterations from the network's seed before the
                             plain-seed fallback (0 = max_iter)                  */
  int any_hit;            /* fast forms: a ray ends at its first converged hit (shadow rays); such an
                             occlusion query (and a frame's shadow ray) takes no FP64 step under
                             mixed precision: the step cannot change hit or miss   */
  int refill_min;         /* persistent form: a wave takes new rays once this many of its 64 lanes
                             are idle (0 = 64: whole-wave refill, the fastest for coherent
                             primary rays; smaller values suit incoherent rays)     */
  int fallback_iters;     /* learned seed: Newton iterations of the plain-seed fallback attempt
                             (0 = max_iter; max_iter - learned_iters keeps the per-pair budget
                             of the plain method)                                 */
  int seed_share;         /* learned seed: 1 = the four lanes of an aligned lane quad (four consecutive
                             rays: a pixel's samples, or neighbouring pixels) that test the same patch
                             in the same pass start from one network evaluation, the quad's first
                             ray's; 0 = one evaluation per (ray, patch) pair     */
  int stall_exit;         /* fast forms: 1 = abandon a Newton solve as a miss from its 4th evaluation
                             on when two consecutive updates leave the (u,v) domain on the same
                             side(s) and the residual fell by less than 10% (stuck against the
                             patch boundary); 0 = always run the full iteration count. The
                             reference's Newton carries a disabled divergence exit
                             (bezier_patch.h:231-232)                             */
  const struct rtn_mlp* seed_net; /* RTN_SEED_LEARNED: the seed network          */
  const int* skip_patch;  /* fast forms: per ray (device), a patch (BVH order, as rtn_hits.prim) the ray
                             does not test, or -1; NULL = none.  For a secondary ray, the patch it
                             leaves: a solve on it converges to the ray's own origin, t <= 0, which
                             is rejected anyway, so the skip changes only the rare ray that a solve
                             on its own patch would send to a second crossing of that patch */
  float t_tol;            /* RTN_PREC_MIXED: a closest hit whose predicted relative t error exceeds
                             this also gets the FP64 step (0 = 2.5e-6, a quarter of the 1e-5 bar) */
  float hit_gate;         /* learned seed: a pair whose predicted hit probability is below this
                             starts from the plain seed with max_iter iterations (no learned
                             attempt); 0 = every pair takes the learned attempt */
  float hit_skip;         /* learned seed: a pair whose predicted hit probability is below this is
                             not solved at all (counted in learned_skips); 0 = off.  Trades hits
                             the network misjudges for fewer Newton solves */
  int learned_min_lanes;  /* learned seed: the network runs for a wave's pairs on a patch (packet
                             forms) or for its pending pairs (one-ray forms) only when at least
                             this many lanes have one; otherwise they start from the plain seed
                             (0 = always) */
} rtn_trace_params;

void rtn_trace_params_default(rtn_trace_params* p);

/* Device outputs, n_rays each (any may be NULL except prim):
 *   prim   int     hit patch position in BVH order, -1 on a miss
 *   t,u,v  float   hit distance (S-o).d and local patch (u,v)
 *   dpdu, dpdv float[3*n]  surface derivatives at the hit
 *   counters uint32[8*n]: newton calls, surface evaluations, flat skips,
 *            interior visits, leaf visits, patch tests, stack overflow,
 *            Newton FP32 operations (evaluations priced per degree) */
typedef struct {
  int* prim;
  float *t, *u, *v;
  float *dpdu, *dpdv;
  uint32_t* counters;
} rtn_hits;

/* Totals accumulated over a trace (device atomics; host-readable after sync). */
typedef struct {
  unsigned long long newton_calls, newton_evals, skips, interior, leaves, patch_tests, overflow, newton_flops;
  unsigned long long fp64_steps, hull_culled, learned_fallbacks, learned_skips;
} rtn_trace_totals;

/* The trace of calculateIntersectWrapper -> calculateIntersect
 * (intersect.cu:20-219) over a built scene, writing SoA hit buffers: the
 * exact mode is the reference's method and operation order (bit-exact), the
 * wide and persistent modes the same method in the fast forms.  Traces n_rays
 * (0 <= n_rays < 2^32; the persistent form takes at most 2^32 - 2^24, the
 * range of its ray queue) in the form prm->mode selects. */
int rtn_trace(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays, const rtn_trace_params* prm,
              const rtn_hits* hits, rtn_trace_totals* totals_host, void* stream);
/* Stack overflow.  Every form counts the rays whose traversal stack
 * overflowed (their closest hit may be wrong).  rtn_trace returns
 * RTN_EOVERFLOW when that count is non-zero if it is synchronous (stream NULL)
 * or totals_host is given; an asynchronous call without totals returns at
 * launch, and rtn_scene_trace_overflow then gives the last trace's count on
 * the scene (it synchronises the device).  rtn_trace allocates nothing: its
 * work queues and totals live in a control block the scene allocates when it
 * is built, so traces of one scene that overlap in time (on several streams)
 * share that block and race; trace one scene from one stream at a time. */
int rtn_scene_trace_overflow(const rtn_scene* scene, unsigned long long* count);

/* Newton solves of a trace as (ray, patch, u, v) records (training pairs of
 * the learned seed; no reference counterpart).  Runs the wide or persistent
 * form with plain seeds (FP32, or the FP64 seed with RTN_PREC_MIXED; stall_exit
 * as given) and appends to device arrays of capacity cap every converged
 * Newton solve with t > 0 (not only the closest hit) and, with all = 1, every
 * other solve the trace runs as (u, v) = (-1, -1): the pairs the trace tests
 * and whether each converges.  *count_host = number of records (may exceed
 * cap; only cap are stored). */
int rtn_trace_pairs(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays,
                    const rtn_trace_params* prm, int all, int* pair_ray, int* pair_patch, float* pair_uv,
                    int64_t cap, int64_t* count_host, void* stream);

/* The plain method's Newton solve (bezier_patch.h:198-266, reference order)
 * on given (ray, patch) pairs: from the plain bilinear seed (seed_uv NULL)
 * or from seed_uv[2n] (e.g. the learned seed), max_iter / thr_* of prm, RNG
 * seeded by ray index + prm->seed_base.  Device outputs: iters[n]
 * (evaluations, -1 for an invalid pair), converged[n] (0/1), optional uv[2n]
 * (solution, or the seed when not converged) and t[n] (-1 when not). */
int rtn_solve_pairs(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays, const int* pair_ray,
                    const int* pair_patch, const float* seed_uv, int64_t n, const rtn_trace_params* prm, int* iters,
                    int* converged, float* uv, float* t, void* stream);
/* rtn_solve_pairs that also solves each pair from the centres
 * ((i + 1/2) / grid, (j + 1/2) / grid), i, j < grid, of the domain and
 * reports the nearest converged root with t > 0 (converged = 1 when any start
 * reaches one; iters = the evaluations of all starts): the closest-root
 * training targets of the learned seed.  No reference counterpart (the
 * reference solves once per pair).  grid in 0..16; 0 is rtn_solve_pairs. */
int rtn_solve_pairs_roots(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays,
                          const int* pair_ray, const int* pair_patch, const float* seed_uv, int64_t n,
                          const rtn_trace_params* prm, int grid, int* iters, int* converged, float* uv, float* t,
                          void* stream);

/* ------------------------------------------------------------ learned seed
 * No reference counterpart.  A 20 -> 64 -> 64 -> 64 -> 3 perceptron (GELU,
 * tanh form; sigmoid outputs) predicting, for a (ray, patch) pair, Newton's
 * starting (u,v) and the probability that Newton converges on the pair (a
 * hit).  Inputs, by the network's feature kind:
 *   RTN_MLP_FEAT_PROJECTED  a 3 x 3 sample of the control net projected on
 *                           the ray's two planes (the ray at the origin), in
 *                           units of the patch box's half extent, and the degrees;
 *   RTN_MLP_FEAT_BOX        the ray (its entry point into the patch box, and its
 *                           unit direction), the four corner control points
 *                           and the degrees, in the box's frame (centre,
 *                           largest half extent).
 * Inference and training run BF16 MFMA with FP32 accumulation; FP32 master
 * parameters, Adam.  Parameter layout (RTN_MLP_NPARAM floats): W1[64][20]
 * b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3], weights
 * row-major (out x in); output rows u, v, hit. */
#define RTN_MLP_NIN 20
#define RTN_MLP_HIDDEN 64
#define RTN_MLP_NOUT 3
#define RTN_MLP_NPARAM 9859
enum { RTN_MLP_FEAT_PROJECTED = 0, RTN_MLP_FEAT_BOX = 1 };
typedef struct rtn_mlp rtn_mlp;
typedef struct {
  int epochs;             /* passes over the pairs, reshuffled each pass          */
  int batch;              /* pairs per Adam step                                  */
  float lr, lr_final;     /* cosine decay from lr to lr_final over all steps      */
  float beta1, beta2, eps;
  float hit_weight;       /* weight of the hit output's cross-entropy in the loss */
  uint64_t seed;          /* shuffle seed                                         */
} rtn_mlp_train_params;

/* Glorot-uniform weights, zero biases; kind = RTN_MLP_FEAT_*. */
int rtn_mlp_create(rtn_mlp** net, uint64_t seed, int kind);
void rtn_mlp_destroy(rtn_mlp* net);
int rtn_mlp_kind(const rtn_mlp* net); /* the feature kind, -1 for NULL */
int rtn_mlp_set_params(rtn_mlp* net, const float* params_host); /* resets Adam state */
int rtn_mlp_get_params(const rtn_mlp* net, float* params_host);
/* Features of a kind (device float[20*n]) of device pairs (ray index, patch
 * position in BVH order). */
int rtn_mlp_pair_features(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays,
                          const int* pair_ray, const int* pair_patch, int64_t n, int kind, float* feats,
                          void* stream);
/* Inference: device feats[20*n] -> device uv[2*n] and, if hit is not NULL,
 * hit[n] (hit probability).  valid (device int[n], may be NULL): pairs with
 * valid[i] == 0 are skipped and their outputs left untouched; the remaining
 * pairs of each 64-pair group are packed before the matrix products, as in
 * the trace kernels. */
int rtn_mlp_forward(const rtn_mlp* net, const float* feats, const int* valid, int64_t n, float* uv, float* hit,
                    void* stream);
/* Gradient (host float[RTN_MLP_NPARAM]) of the training loss over device
 * feats[20*n] and targets[3*n] = (u, v, h) per pair, h = 1 for a hit (a
 * converged solve at (u,v)) and 0 for a miss:
 *   mean_n [ h_n |uv_n - (u,v)_n|^2 + hit_weight * BCE(hit_n, h_n) ],
 * BCE(p, h) = -(h log p + (1 - h) log(1 - p)); *loss_host = that mean.
 * Synchronous. */
int rtn_mlp_gradient(rtn_mlp* net, const float* feats, const float* targets, int64_t n, float hit_weight,
                     float* grad_host, float* loss_host);
/* Defaults: 4 epochs, batch 32768 (one resident round of single-wave blocks at
 * two per compute unit), lr 6e-3 decaying to 1e-4, beta 0.9 / 0.999, eps 1e-8,
 * hit_weight 0.05. */
void rtn_mlp_train_params_default(rtn_mlp_train_params* tp);
/* Adam training on device pairs (targets as rtn_mlp_gradient); the pairs are
 * reshuffled on the device each epoch (a keyed pseudo-random permutation);
 * epoch_loss_host[epochs] = mean loss per epoch (may be NULL). */
int rtn_mlp_train(rtn_mlp* net, const float* feats, const float* targets, int64_t n,
                  const rtn_mlp_train_params* tp, float* epoch_loss_host, void* stream);

/* Secondary rays of primary hits (no reference counterpart): per hit, in
 * this order, one shadow ray towards the directional light (direction
 * `light`, normalised) and one mirror ray, both from the hit point offset by
 * eps * (1 + max|S|) along the normal facing the incoming ray.  Slots are
 * kind-major: primary ray r owns slot j*n + r for the j-th enabled kind
 * (shadow first; per = want_shadow + want_mirror kinds), so the layout is
 * deterministic and a wave's 64 consecutive slots hold one kind from 64
 * neighbouring primaries; parent[i] = r, kind[i] = 0 shadow, 1 mirror, -1
 * empty (a miss).  The trace entry points skip rays whose kind is negative
 * when given the kind array (rtn_frame does).  cap >= per * n.
 * *count_host = valid (non-empty) rays. */
int rtn_rays_secondary(const float* const rays_soa[6], const rtn_hits* hits, int64_t n, const float light[3],
                       float eps, int want_shadow, int want_mirror, const rtn_ray_queue* out, int* parent, int* kind,
                       int64_t cap, int64_t* count_host, void* stream);

/* Lambert shading with one directional light (replaces the reference's debug
 * shaders, testShading.cu:9-55): pixel = mean over its spp consecutive rays of
 * albedo * max(0, N.L), N = unit(dpdu x dpdv) facing the ray (benchmark.cpp:91-92),
 * 0 on a miss or where shadow_hit[r] != 0 (optional).  image: device float[W*H]. */
int rtn_shade_lambert(const float* const rays_soa[6], const rtn_hits* hits, int W, int H, int spp, int ray_flags,
                      const float light[3], float albedo, const int* shadow_hit, float* image, void* stream);

/* ---------------------------------------------- multi-scene batches, frames
 * No reference counterpart (the reference traces one scene).  A batch merges
 * built scenes into one set of device arrays with one BVH root per scene;
 * rays carry their scene index (ray_scene[i] in 0..k-1), so rays of all
 * scenes share one launch.  Hit prim values are batch-wide patch positions:
 * scene j owns [base[j], base[j+1]) of rtn_batch_patch_base. */
typedef struct rtn_batch rtn_batch;
int rtn_batch_create(rtn_batch** batch, rtn_scene* const* scenes, int k);
void rtn_batch_destroy(rtn_batch* batch);
int rtn_batch_patch_base(const rtn_batch* batch, int64_t* base /* k+1 */);
/* rtn_trace over a batch (wide or persistent form). */
int rtn_batch_trace(const rtn_batch* batch, const float* const rays_soa[6], const int* ray_scene, int64_t n_rays,
                    const rtn_trace_params* prm, const rtn_hits* hits, rtn_trace_totals* totals_host, void* stream);

/* A frame: primary rays of a batch traced by the persistent kernel, then one
 * shadow ray (any-hit, towards the light) and one mirror ray per hit, in the
 * fixed (kind-major) slots of rtn_rays_secondary, traced by a second launch
 * (empty slots are skipped).  All buffers are allocated by rtn_frame_create;
 * rtn_frame_run enqueues the whole sequence without allocating or copying
 * from the host, so rtn_frame_capture can record it into a hipGraph that
 * later rtn_frame_run calls replay.  Calls on one frame must not overlap. */
typedef struct rtn_frame rtn_frame;
typedef struct {
  rtn_trace_params trace; /* mode RTN_MODE_PERSISTENT; count 0; any_hit 0      */
  float light[3];         /* direction towards the light                        */
  float eps;              /* secondary origin offset, eps * (1 + max|S|)        */
  int shadow, mirror;     /* 0/1: generate shadow / mirror rays                 */
  int packet_primary;     /* 1: trace the primary rays with the packet form (one wave per 64
                             coherent rays) instead of the persistent per-lane form  */
  int packet_secondary;   /* 1: the same for the shadow and mirror rays         */
  int skip_origin;        /* 1 (default): a secondary ray does not test the patch it leaves (its
                             primary ray's hit patch; see rtn_trace_params.skip_patch)  */
  int reserved;
} rtn_frame_params;
typedef struct {
  int64_t n_primary, n_secondary;  /* rays; n_secondary counts non-empty slots   */
  int64_t n_secondary_slots;       /* (shadow + mirror) x n_primary              */
  unsigned long long overflow;  /* rays whose traversal stack overflowed        */
  int shape_lanes, shape_stack, shape_blocks_per_cu, captured, refill_min;
  int packet;                      /* 1: packet primary launch, 2: packet secondary launch */
} rtn_frame_status_t;
typedef struct { int lanes, stack, blocks_per_cu, refill_min; float ms; int packet_primary, packet_secondary; } rtn_tune_entry;

void rtn_frame_params_default(rtn_frame_params* fp);
int rtn_frame_create(rtn_frame** frame, const rtn_batch* batch, const float* const rays_soa[6], const int* ray_scene,
                     int64_t n_rays, const rtn_frame_params* fp);
void rtn_frame_destroy(rtn_frame* frame);
int rtn_frame_run(rtn_frame* frame, void* stream);
int rtn_frame_capture(rtn_frame* frame, void* stream /* non-default */);
int rtn_frame_status(const rtn_frame* frame, rtn_frame_status_t* out); /* synchronises the device */
/* Work totals of the frame's two launches (for pricing their algorithmic
 * work): runs the frame once, eagerly, with the counting builds of the same
 * launches (default launch shape: same results), and returns the totals of
 * the primary and of the secondary launch.  Synchronous; leaves a captured
 * graph in place. */
int rtn_frame_count(rtn_frame* frame, rtn_trace_totals* primary, rtn_trace_totals* secondary);
/* Device pointers of the frame's outputs (any argument may be NULL). */
int rtn_frame_results(const rtn_frame* frame, rtn_hits* primary, rtn_hits* secondary, const int** parent,
                      const int** kind, const float** secondary_rays_soa6);
/* Launch shape: lanes per block (64/128/256 = 1/2/4 waves), traversal stack
 * entries per lane (16/20/24/32; LDS = 8 B x entries x lanes), resident blocks
 * per compute unit.  Drops a captured graph. */
int rtn_frame_set_shape(rtn_frame* frame, int lanes, int stack, int blocks_per_cu);
/* Refill threshold of both persistent launches (see rtn_trace_params.refill_min). */
int rtn_frame_set_refill(rtn_frame* frame, int refill_min);
/* Packet form for the primary and/or the secondary launch (see
 * rtn_frame_params.packet_primary).  Drops a captured graph. */
int rtn_frame_set_packet(rtn_frame* frame, int packet_primary, int packet_secondary);
/* Autotuner: times the frame over every built launch shape at 2, 4, 8 and 12
 * blocks per CU (capped at the blocks that stay resident; whole-wave refill;
 * reps timed runs each; stack overflow rejects a shape), then the best shape
 * at refill thresholds 1, 16 and 32 idle lanes, then the packet form for the
 * primary launch and for both launches, and keeps the fastest (shape, refill
 * threshold and packet flags).  table (optional) receives each candidate,
 * ms < 0 for rejected ones.  Drops a captured graph. */
int rtn_frame_autotune(rtn_frame* frame, int reps, rtn_tune_entry* table, int table_cap, int* n_tried,
                       void* stream);

/* FP64 reference trace (no reference counterpart: the reference is FP32
 * only).  The plain method's traversal, seed and Newton iteration evaluated
 * in FP64 with its own cap and tolerances, used as the error reference of
 * the FP32 forms.  Outputs are device FP64 arrays of n (dpdu/dpdv 3n);
 * t_second = nearest converged hit on another patch (tie detection);
 * counters as rtn_hits (may be NULL).  Synchronous when stream is NULL. */
int rtn_trace_f64(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays, int max_iter,
                  double thr_det, double thr_res, uint32_t seed_base, int* prim, double* t, double* u, double* v,
                  double* dpdu, double* dpdv, double* t_second, uint32_t* counters, void* stream);
/* FP64 closest-root reference (no reference counterpart: the reference runs
 * one Newton solve per patch from its bilinear seed, intersect.cu:126).
 * rtn_trace_f64 in which every tested patch also runs Newton from the
 * centres ((i + 1/2) / grid, (j + 1/2) / grid), i, j < grid, of its domain and
 * keeps its nearest converged root with t > 0: the pixel's hit is then the
 * nearest root the searches find, the reference against which a form that
 * finds roots the plain seed misses is judged.  grid in 0..16 (0 is
 * rtn_trace_f64); the same outputs. */
int rtn_trace_f64_roots(const rtn_scene* scene, const float* const rays_soa[6], int64_t n_rays, int max_iter,
                        double thr_det, double thr_res, int grid, uint32_t seed_base, int* prim, double* t,
                        double* u, double* v, double* dpdu, double* dpdv, double* t_second, uint32_t* counters,
                        void* stream);

/* calculateIntersectWrapper (intersect.cu:199-219; declared intersect.h:9-13)
 * on reference-layout buffers: nodes[2N-1], patches[N] whose control_point
 * pointers are device-visible, delta[N] in BVH order, interact[W*H].  As in
 * the reference the 8th argument is the vertical field of view in radians,
 * the call is synchronous, record fields are written only on a hit, and
 * idx_interact[idx] = idx marks hits.  Counters (may be NULL) are added to. */
int rtn_calculate_intersect(const rtn_ray_queue* rays, const rtn_bvh_node* nodes, const rtn_bezier_patch* patches,
                            rtn_surface_interaction* interact, int W, int H, float flat_threshold, float v_fov,
                            const float* delta, unsigned long long* newton_count, unsigned long long* skip_count);

/* rtn_calculate_intersect without any device allocation (no counterpart in the
 * reference, whose wrapper has no such form): for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results, synchronous behaviour and errors as
 * rtn_calculate_intersect; W*H must fit the int idx_interact range.  Its
 * scratch (the 8-word totals block: Newton and skip counts, stack overflow)
 * is a file-scope device array of 64 bytes, one copy per GPU kept until the
 * process ends, whatever the input size; the call zeroes it first.  Calls
 * that overlap on one GPU (several host threads) share that array and race. */
int rtn_calculate_intersect_noalloc(const rtn_ray_queue* rays, const rtn_bvh_node* nodes,
                                    const rtn_bezier_patch* patches, rtn_surface_interaction* interact, int W, int H,
                                    float flat_threshold, float v_fov, const float* delta,
                                    unsigned long long* newton_count, unsigned long long* skip_count);

/* getPatch (bezier_patch.h:87-139) with the A5.7 passes run on the GPU.
 * Host in, host out.  First call with cp == NULL to get nbU, nbV. */
int rtn_get_patch(int n, int p, const float* U, int m, int q, const float* V, const float* Pw, int* nbU, int* nbV,
                  float* cp, float* uv_range);

/* buildBVH (bvh.h:148-179) on the GPU.  bounds: host float[6*np] (min xyz,
 * max xyz per primitive, as BezierPatchOps::bounds).  nodes: host
 * rtn_bvh_node[2np-1]; perm[i] = primitive stored at BVH position i. */
int rtn_build_bvh(const float* bounds, int np, rtn_bvh_node* nodes, uint32_t* perm, int* nodes_used);

/* BezierPatchOps::bounds (bezier_patch.h:154-169) for patches of one degree,
 * on the GPU: cp host float[np*(p+1)(q+1)*4] -> bounds host float[6*np]. */
int rtn_patch_bounds(const float* cp, int np, int p, int q, float* bounds);

/* NurbsOps::curvatureRefinementU/V (nurbs.h:290-486), host-side as in the
 * reference.  dir 0 refines U, 1 refines V.  Call with knots_out == NULL to
 * get the number of inserted knots; then with arrays sized for the result. */
int rtn_refine(int dir, int n, int p, const float* U, int m, const float* Pw, float C, int* inserted,
               float* knots_out, float* Pw_out);

/* NurbsOps::curvatureRefinementU then curvatureRefinementV (nurbs.h:290-486)
 * of every NURBS surface added to the scene, on the GPU: one launch per
 * direction counts the knots of all spans of all surfaces, one more inserts
 * them (A5.1, nurbs.h:97-140), each column curve on its own lane.  The
 * scene's surfaces are replaced by the refined ones; call before
 * rtn_scene_build.  A constant <= 0 skips its direction.  Synchronous.
 * Results equal rtn_refine's bit for bit, except that Δ^1.5 in knotsToAdd is
 * the correctly rounded power where the host calls powf (a one-ulp difference
 * on rare inputs, which moves a knot count only when the count's real value
 * lies within that ulp of an integer). */
int rtn_scene_refine(rtn_scene* scene, float C_u, float C_v);

#ifdef __cplusplus
}
#endif
#endif /* RTNURBS_H */
