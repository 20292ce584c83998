/* rtnurbs.hpp - the reference's C++ API of the trace path, header-only, over
 * the C ABI of rtnurbs.h (librtnurbs.so, gfx950).
 *
 * It declares the reference's own types and namespaces with its names and
 * layouts (MathTypes, MathExtra, ParametricSurf::nurbsSurf, bezierPatch,
 * BVHNode, surfaceInteraction, SoAQueue::ray_queue, camera_config,
 * projection) and its entry points with their signatures:
 *
 *   Transform::lookAt                         transform.h:46-60
 *   MathUtils::radians                        math_utils.h:20-23
 *   CameraOps::init_camera                    camera.h:123-141
 *   CameraOps::getPerspectiveCamera           camera.h:63-85
 *   CameraOps::getOrthographicCamera          camera.h:42-59
 *   calculateRaysWrapper                      generateRaysKernel.h:6
 *   BezierPatchOps::getPatch / freePatches    bezier_patch.h:87-139, :143-149
 *   BvhOps::buildBVH                          bvh.h:148-179
 *   NurbsOps::curvatureRefinementU / V        nurbs.h:290-486
 *   calculateIntersectWrapper                 intersect.h:9-13
 *   CudaDevAlloc / CudaAllocManaged / CudaHostAlloc / CudaFree / CudaCpuFree
 *                                             memory.h:23-65
 *
 * so a driver written against the reference (main.cpp's shape) switches by
 * including this header instead of the reference's and linking
 * librtnurbs.so.  Each type is layout-identical to its rtn_* twin in
 * rtnurbs.h (checked below), so buffers pass to the C ABI unchanged.
 *
 * Same ownership and (void-returning, print-and-continue) error behaviour as
 * the reference; every call is synchronous, like the reference wrappers.
 * The memory names allocate with HIP (hipMalloc, hipMallocManaged,
 * hipHostMalloc inside the library), not with CUDA. */
#ifndef RTNURBS_HPP
#define RTNURBS_HPP

#include <sys/types.h>

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rtnurbs.h"

namespace rtnurbs {
/* Print-and-continue, as the reference's CUDA_CHECK does (memory.h:9-20). */
inline void report(int rc, const char* where) {
  if (rc != RTN_OK) std::fprintf(stderr, "%s: %s\n", where, rtn_last_error());
}
}  // namespace rtnurbs

/* ------------------------------------------------------------------ types */
namespace MathTypes {  // math_types.h:5-57
struct matrix4x4f { float m[16]; };  // column major
struct vector4f { float v[4]; };
struct vector3f { float v[3]; };
struct vector2f { float v[2]; };
struct point4f { float v[4]; };
struct point3f { float v[3]; };
struct point2f { float v[2]; };
}  // namespace MathTypes

namespace MathExtra {  // math_types.h:60-80
struct bounds2f { float p_min_x, p_min_y, p_max_x, p_max_y; };
struct bounds3f { float p_min_x, p_min_y, p_min_z, p_max_x, p_max_y, p_max_z; };
}  // namespace MathExtra

namespace ParametricSurf {  // nurbs.h:17-22
struct nurbsSurf {
  float *u_knots, *v_knots;            // n+p+2 / m+q+2 knots
  MathTypes::vector4f* control_point;  // (n+1)(m+1), homogeneous, u-major
  int n, m;                            // indices of the last control point
  int p, q;
};
}  // namespace ParametricSurf

struct bezierPatch {  // bezier_patch.h:10-16
  MathTypes::vector4f* control_point;
  float u0, u1, v0, v1;
  int p, q;
  MathTypes::vector4f* heap_pointer;  // the one control-point block of the getPatch call
};

struct BVHNode {  // bvh.h:13-17
  MathExtra::bounds3f node_bound;
  uint left_first;
  uint prim_count;
};

struct surfaceInteraction {  // gpu/interact.h:6-21
  float t_hit;
  MathTypes::vector3f wo;
  MathTypes::point2f uv;
  MathTypes::vector3f dpdu;
  MathTypes::vector3f dpdv;
  MathTypes::vector3f dndu;
  MathTypes::vector3f dndv;
};

namespace SoAQueue {  // soa.h:6-27
struct ray_queue {
  float* origin_x;
  float* origin_y;
  float* origin_z;
  float* direction_x;
  float* direction_y;
  float* direction_z;
  int* idx_interact;
};
struct frame_buffer {
  float* r;
  float* g;
  float* b;
};
}  // namespace SoAQueue

struct camera_config {  // camera.h:13-27
  MathTypes::matrix4x4f world_from_camera;
  MathTypes::matrix4x4f screen_from_camera;
  MathTypes::matrix4x4f NDC_from_screen;
  MathTypes::matrix4x4f raster_from_NDC;
  MathTypes::matrix4x4f raster_from_screen;
  MathTypes::matrix4x4f screen_from_raster;
  MathTypes::matrix4x4f camera_from_raster;
  int screen_resolution_x;
  int screen_resolution_y;
  MathExtra::bounds2f screen_window;
};

enum projection { orthographic, perspective };  // camera.h:37

/* Layout twins: each reference type above is the rtn_* type of the C ABI. */
#define RTN_SAME_FIELD(A, a, B, b) static_assert(offsetof(A, a) == offsetof(B, b), #A "." #a " is " #B "." #b)
static_assert(sizeof(MathTypes::vector4f) == sizeof(rtn_vec4), "vector4f");
static_assert(sizeof(ParametricSurf::nurbsSurf) == sizeof(rtn_nurbs_surf), "nurbsSurf");
RTN_SAME_FIELD(ParametricSurf::nurbsSurf, control_point, rtn_nurbs_surf, control_point);
RTN_SAME_FIELD(ParametricSurf::nurbsSurf, n, rtn_nurbs_surf, n);
RTN_SAME_FIELD(ParametricSurf::nurbsSurf, q, rtn_nurbs_surf, q);
static_assert(sizeof(bezierPatch) == sizeof(rtn_bezier_patch), "bezierPatch");
RTN_SAME_FIELD(bezierPatch, u0, rtn_bezier_patch, u0);
RTN_SAME_FIELD(bezierPatch, p, rtn_bezier_patch, p);
RTN_SAME_FIELD(bezierPatch, heap_pointer, rtn_bezier_patch, heap_pointer);
static_assert(sizeof(BVHNode) == sizeof(rtn_bvh_node), "BVHNode");
RTN_SAME_FIELD(BVHNode, left_first, rtn_bvh_node, left_first);
RTN_SAME_FIELD(BVHNode, prim_count, rtn_bvh_node, prim_count);
static_assert(sizeof(surfaceInteraction) == sizeof(rtn_surface_interaction), "surfaceInteraction");
RTN_SAME_FIELD(surfaceInteraction, uv, rtn_surface_interaction, uv);
RTN_SAME_FIELD(surfaceInteraction, dpdu, rtn_surface_interaction, dpdu);
RTN_SAME_FIELD(surfaceInteraction, dndv, rtn_surface_interaction, dndv);
static_assert(sizeof(SoAQueue::ray_queue) == sizeof(rtn_ray_queue), "ray_queue");
RTN_SAME_FIELD(SoAQueue::ray_queue, idx_interact, rtn_ray_queue, idx_interact);
static_assert(sizeof(camera_config) == sizeof(rtn_camera_config), "camera_config");
RTN_SAME_FIELD(camera_config, camera_from_raster, rtn_camera_config, camera_from_raster);
RTN_SAME_FIELD(camera_config, screen_resolution_y, rtn_camera_config, screen_resolution_y);
RTN_SAME_FIELD(camera_config, screen_window, rtn_camera_config, screen_window);
static_assert((int)orthographic == RTN_PROJ_ORTHOGRAPHIC && (int)perspective == RTN_PROJ_PERSPECTIVE, "projection");
#undef RTN_SAME_FIELD

/* ----------------------------------------------------------------- memory
 * memory.h:23-65: allocation of `count` T, printing on failure and
 * continuing (a null pointer), as CUDA_CHECK does; CudaDevAlloc aborts on a
 * byte-count overflow as the reference does. */
template <typename T>
T* CudaDevAlloc(size_t count) {
  if (count > SIZE_MAX / sizeof(T)) std::abort();
  void* p = nullptr;
  rtnurbs::report(rtn_malloc(&p, sizeof(T) * count), "CudaDevAlloc");
  return static_cast<T*>(p);
}
template <typename T>
T* CudaAllocManaged(size_t count) {
  void* p = nullptr;
  rtnurbs::report(rtn_malloc_managed(&p, sizeof(T) * count), "CudaAllocManaged");
  return static_cast<T*>(p);
}
template <typename T>
T* CudaHostAlloc(size_t count) {
  void* p = nullptr;
  rtnurbs::report(rtn_malloc_host(&p, sizeof(T) * count), "CudaHostAlloc");
  return static_cast<T*>(p);
}
inline void CudaFree(void* addr) {
  if (addr) rtnurbs::report(rtn_free(addr), "CudaFree");
}
inline void CudaCpuFree(void* addr) {
  if (addr) rtnurbs::report(rtn_free_host(addr), "CudaCpuFree");
}

/* --------------------------------------------------------------- camera */
namespace MathUtils {
/* radians (math_utils.h:20-23) with the reference's PI (constants.h:18). */
inline float radians(float degree) { return (3.14159265f / 180) * degree; }
}  // namespace MathUtils

namespace Transform {
/* lookAt (transform.h:46-60): camera-to-world transform. */
inline MathTypes::matrix4x4f lookAt(const MathTypes::point3f& pos, const MathTypes::point3f& look,
                                    const MathTypes::vector3f& up) {
  MathTypes::matrix4x4f m{};
  rtnurbs::report(rtn_lookat(pos.v, look.v, up.v, m.m), "lookAt");
  return m;
}
}  // namespace Transform

namespace CameraOps {
inline camera_config init_camera(const MathTypes::matrix4x4f world_from_camera, int res_x, int res_y) {
  camera_config c{};
  rtnurbs::report(rtn_camera_init(world_from_camera.m, res_x, res_y, reinterpret_cast<rtn_camera_config*>(&c)),
                  "init_camera");
  return c;
}
inline camera_config getPerspectiveCamera(const camera_config& initial_config, float fov, float n, float f) {
  camera_config c{};
  rtnurbs::report(rtn_camera_get_perspective(reinterpret_cast<const rtn_camera_config*>(&initial_config), fov, n, f,
                                             reinterpret_cast<rtn_camera_config*>(&c)),
                  "getPerspectiveCamera");
  return c;
}
inline camera_config getOrthographicCamera(const camera_config& initial_config, float z_near, float z_far) {
  camera_config c{};
  rtnurbs::report(rtn_camera_get_orthographic(reinterpret_cast<const rtn_camera_config*>(&initial_config), z_near,
                                              z_far, reinterpret_cast<rtn_camera_config*>(&c)),
                  "getOrthographicCamera");
  return c;
}
}  // namespace CameraOps

/* calculateRaysWrapper (generateRaysKernel.h:6, .cu:39-67): one ray per pixel
 * of the camera's raster into the device queue, idx_interact = -1. */
inline void calculateRaysWrapper(SoAQueue::ray_queue& ray_queue, projection proj, camera_config& camera) {
  rtnurbs::report(rtn_rays_generate(reinterpret_cast<const rtn_ray_queue*>(&ray_queue), (int)proj,
                                    reinterpret_cast<const rtn_camera_config*>(&camera), nullptr),
                  "calculateRaysWrapper");
}

/* ---------------------------------------------------------------- scene */
namespace BezierPatchOps {
/* getPatch (bezier_patch.h:87-139): the rational Bezier patches of one NURBS
 * surface, u-major (patch i*nbV + j), decomposed on the GPU.  Returns a new[]
 * header array whose control points live in one new[] block; every patch's
 * heap_pointer is that block, as in the reference (bezier_patch.h:128), so it
 * survives buildBVH's permutation.  Release with freePatches.  NULL on error. */
inline bezierPatch* getPatch(const ParametricSurf::nurbsSurf& s, int& nbU, int& nbV) {
  int rc = rtn_get_patch(s.n, s.p, s.u_knots, s.m, s.q, s.v_knots, reinterpret_cast<const float*>(s.control_point),
                         &nbU, &nbV, nullptr, nullptr);
  if (rc != RTN_OK) {
    rtnurbs::report(rc, "getPatch");
    return nullptr;
  }
  const int np = nbU * nbV, ps = (s.p + 1) * (s.q + 1);
  MathTypes::vector4f* block = new MathTypes::vector4f[(size_t)np * ps];
  std::vector<float> uv(4 * (size_t)np);
  rc = rtn_get_patch(s.n, s.p, s.u_knots, s.m, s.q, s.v_knots, reinterpret_cast<const float*>(s.control_point), &nbU,
                     &nbV, reinterpret_cast<float*>(block), uv.data());
  if (rc != RTN_OK) {
    rtnurbs::report(rc, "getPatch");
    delete[] block;
    return nullptr;
  }
  bezierPatch* out = new bezierPatch[np];
  for (int k = 0; k < np; k++) {
    out[k].control_point = block + (size_t)k * ps;
    out[k].u0 = uv[4 * k];
    out[k].u1 = uv[4 * k + 1];
    out[k].v0 = uv[4 * k + 2];
    out[k].v1 = uv[4 * k + 3];
    out[k].p = s.p;
    out[k].q = s.q;
    out[k].heap_pointer = block;
  }
  return out;
}

/* freePatches (bezier_patch.h:143-149): the control-point block, then the headers. */
inline void freePatches(bezierPatch* patch) {
  if (!patch) return;
  delete[] patch[0].heap_pointer;
  delete[] patch;
}
}  // namespace BezierPatchOps

namespace BvhOps {
/* buildBVH (bvh.h:148-179): the reference's exhaustive-SAH BVH over
 * primitive_size patches of one degree (host headers whose control_point
 * pointers are host readable), built on the GPU with the reference's node
 * numbering; nodes (2n-1, caller-allocated, host-visible) are written and
 * `patches` is permuted in place into BVH order.  Like the reference it
 * always builds from node 0 (bvh.h:158-159). */
inline void buildBVH(BVHNode* nodes, uint node_root_idx, bezierPatch*& patches, int primitive_size) {
  (void)node_root_idx;
  const int n = primitive_size;
  if (n <= 0) return;
  const int p = patches[0].p, q = patches[0].q, ps = (p + 1) * (q + 1);
  std::vector<float> cp((size_t)n * ps * 4), bounds(6 * (size_t)n);
  for (int k = 0; k < n; k++)
    std::memcpy(&cp[(size_t)k * ps * 4], patches[k].control_point, sizeof(float) * 4 * ps);
  int rc = rtn_patch_bounds(cp.data(), n, p, q, bounds.data());
  std::vector<uint32_t> perm(n);
  int used = 0;
  if (rc == RTN_OK) rc = rtn_build_bvh(bounds.data(), n, reinterpret_cast<rtn_bvh_node*>(nodes), perm.data(), &used);
  if (rc != RTN_OK) {
    rtnurbs::report(rc, "buildBVH");
    return;
  }
  std::vector<bezierPatch> tmp(patches, patches + n);
  for (int i = 0; i < n; i++) patches[i] = tmp[perm[i]];
}
}  // namespace BvhOps

namespace NurbsOps {
namespace detail {
/* One direction of curvature refinement; like the reference it takes
 * ownership of the caller's new[] knot and control-point arrays and replaces
 * them (and the count) with new[] arrays of the refined surface. */
inline void refine(int dir, float*& K, int deg, int& n_dir, int n_other, float C, MathTypes::vector4f*& P) {
  int inserted = 0;
  const int n = dir == 0 ? n_dir : n_other, m = dir == 0 ? n_other : n_dir;
  int rc = rtn_refine(dir, n, deg, K, m, reinterpret_cast<const float*>(P), C, &inserted, nullptr, nullptr);
  if (rc != RTN_OK || inserted == 0) {
    rtnurbs::report(rc, "curvatureRefinement");
    return;
  }
  float* K2 = new float[(size_t)n_dir + inserted + deg + 2];
  const size_t ncp = (size_t)(n + (dir == 0 ? inserted : 0) + 1) * (m + (dir == 1 ? inserted : 0) + 1);
  MathTypes::vector4f* P2 = new MathTypes::vector4f[ncp];
  rc = rtn_refine(dir, n, deg, K, m, reinterpret_cast<const float*>(P), C, &inserted, K2,
                  reinterpret_cast<float*>(P2));
  if (rc != RTN_OK) {
    rtnurbs::report(rc, "curvatureRefinement");
    delete[] K2;
    delete[] P2;
    return;
  }
  delete[] K;
  delete[] P;
  K = K2;
  P = P2;
  n_dir += inserted;
}
}  // namespace detail
/* curvatureRefinementU / V (nurbs.h:290, :384). */
inline void curvatureRefinementU(float*& U, int p, int& n, int& m, float C, MathTypes::vector4f*& P) {
  detail::refine(0, U, p, n, m, C, P);
}
inline void curvatureRefinementV(float*& V, int q, int& m, int n, float C, MathTypes::vector4f*& P) {
  detail::refine(1, V, q, m, n, C, P);
}
}  // namespace NurbsOps

/* ---------------------------------------------------------------- trace */
/* calculateIntersectWrapper (intersect.h:9-13, intersect.cu:199-219).  The
 * eighth argument is the vertical field of view in radians (the reference
 * header calls it cam_k; its definition converts it, intersect.cu:202, :212).
 * Buffers are device or managed memory; patches' control_point pointers must
 * be device-visible; delta is in BVH order; counters are added to. */
inline void calculateIntersectWrapper(SoAQueue::ray_queue& ray_queue, BVHNode* nodes, bezierPatch* patches,
                                      surfaceInteraction* interact, int screen_res_x, int screen_res_y,
                                      float flat_threshold, float cam_k, float* delta,
                                      unsigned long long* newton_count = nullptr,
                                      unsigned long long* skip_count = nullptr) {
  rtnurbs::report(rtn_calculate_intersect(reinterpret_cast<const rtn_ray_queue*>(&ray_queue),
                                          reinterpret_cast<const rtn_bvh_node*>(nodes),
                                          reinterpret_cast<const rtn_bezier_patch*>(patches),
                                          reinterpret_cast<rtn_surface_interaction*>(interact), screen_res_x,
                                          screen_res_y, flat_threshold, cam_k, delta, newton_count, skip_count),
                  "calculateIntersectWrapper");
}

#endif /* RTNURBS_HPP */
