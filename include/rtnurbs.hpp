/* rtnurbs.hpp - the reference's C++ entry points of the trace path, as
 * header-only shims over the C ABI of rtnurbs.h (librtnurbs.so, gfx950).
 *
 * Same names, argument meaning, ownership and (void-returning, print-and-
 * continue) error behaviour as the reference functions they replace, so a
 * caller of the reference (main.cpp / benchmark.cpp style drivers) switches by
 * including this header and linking librtnurbs.so.  The structs are the
 * layout-compatible rtn_* types of rtnurbs.h (rtn_bvh_node = BVHNode,
 * rtn_bezier_patch = bezierPatch, rtn_surface_interaction =
 * surfaceInteraction, rtn_ray_queue = SoAQueue::ray_queue).
 *
 * Every shim is synchronous, like the reference wrappers. */
#ifndef RTNURBS_HPP
#define RTNURBS_HPP

#include <cstdio>
#include <cstring>
#include <vector>

#include "rtnurbs.h"

namespace rtnurbs {

/* Print-and-continue, as CUDA_CHECK does (memory.h:9-20). */
inline void report(int rc, const char* where) {
  if (rc != RTN_OK) std::fprintf(stderr, "%s: %s\n", where, rtn_last_error());
}

/* calculateIntersectWrapper (intersect.h:9-13, intersect.cu:199-219).  The
 * eighth argument is the vertical field of view in radians (the reference
 * header calls it cam_k; its definition converts it, intersect.cu:202, :212).
 * Buffers are device or managed memory; patches' control_point pointers must
 * be device-visible; delta is in BVH order; counters are added to. */
inline void calculateIntersectWrapper(rtn_ray_queue& rays, rtn_bvh_node* nodes, rtn_bezier_patch* patches,
                                      rtn_surface_interaction* interact, int W, int H, float flat_threshold,
                                      float cam_k, float* delta, unsigned long long* newton_count = nullptr,
                                      unsigned long long* skip_count = nullptr) {
  report(rtn_calculate_intersect(&rays, nodes, patches, interact, W, H, flat_threshold, cam_k, delta, newton_count,
                                 skip_count),
         "calculateIntersectWrapper");
}

/* calculateRaysWrapper (generateRaysKernel.h:6, .cu:39-67) for the
 * perspective projection: the camera enters as its world_from_camera and
 * camera_from_raster matrices (the payload the reference copies to
 * __constant__ memory, generateRaysKernel.cu:42), resolution W x H. */
inline void calculateRaysWrapper(rtn_ray_queue& rays, const float world_from_camera[16],
                                 const float camera_from_raster[16], int W, int H) {
  report(rtn_rays_perspective(world_from_camera, camera_from_raster, W, H, 1, 0, 0u, &rays, nullptr),
         "calculateRaysWrapper");
}

/* CameraOps::init_camera + getPerspectiveCamera + Transform::lookAt
 * (camera.h:63-141, transform.h:46): the two matrices above. */
inline void perspectiveCamera(const float eye[3], const float look[3], const float up[3], int W, int H,
                              float fov_deg, float znear, float zfar, float world_from_camera[16],
                              float camera_from_raster[16]) {
  report(rtn_camera_perspective(eye, look, up, W, H, fov_deg, znear, zfar, world_from_camera, camera_from_raster),
         "perspectiveCamera");
}

/* BezierPatchOps::getPatch (bezier_patch.h:87-139): Bezier patches of one
 * NURBS surface (nurbsSurf, nurbs.h:17-22: n, p, U[n+p+2], m, q, V[m+q+2],
 * Pw[(n+1)(m+1)] homogeneous u-major).  Returns a new[] header array of
 * nbU*nbV patches (u-major) whose control points live in one new[] block
 * (heap_pointer of patch 0); release with freePatches.  NULL on error. */
inline rtn_bezier_patch* getPatch(int n, int p, const float* U, int m, int q, const float* V, const float* Pw,
                                  int& nbU, int& nbV) {
  int rc = rtn_get_patch(n, p, U, m, q, V, Pw, &nbU, &nbV, nullptr, nullptr);
  if (rc != RTN_OK) {
    report(rc, "getPatch");
    return nullptr;
  }
  const int np = nbU * nbV, ps = (p + 1) * (q + 1);
  rtn_vec4* block = new rtn_vec4[(size_t)np * ps];
  std::vector<float> uv(4 * (size_t)np);
  rc = rtn_get_patch(n, p, U, m, q, V, Pw, &nbU, &nbV, reinterpret_cast<float*>(block), uv.data());
  if (rc != RTN_OK) {
    report(rc, "getPatch");
    delete[] block;
    return nullptr;
  }
  rtn_bezier_patch* out = new rtn_bezier_patch[np];
  for (int k = 0; k < np; k++) {
    out[k].control_point = block + (size_t)k * ps;
    out[k].u0 = uv[4 * k];
    out[k].u1 = uv[4 * k + 1];
    out[k].v0 = uv[4 * k + 2];
    out[k].v1 = uv[4 * k + 3];
    out[k].p = p;
    out[k].q = q;
    out[k].heap_pointer = k == 0 ? block : nullptr;
  }
  return out;
}

/* BezierPatchOps::freePatches (bezier_patch.h:143-149). */
inline void freePatches(rtn_bezier_patch* patches) {
  if (!patches) return;
  delete[] static_cast<rtn_vec4*>(patches[0].heap_pointer);
  delete[] patches;
}

/* BvhOps::buildBVH (bvh.h:148-179): the reference's exhaustive-SAH BVH over n
 * patches of one degree (host headers whose control_point pointers are host
 * readable), built on the GPU with the reference's node numbering; nodes
 * (2n-1, caller-allocated) are written and `patches` is permuted in place
 * into BVH order, as the reference does. */
inline void buildBVH(rtn_bvh_node* nodes, unsigned root, rtn_bezier_patch*& patches, int n) {
  (void)root;  // the reference always builds from node 0 (bvh.h:158-159)
  if (n <= 0) return;
  const int p = patches[0].p, q = patches[0].q, ps = (p + 1) * (q + 1);
  std::vector<float> cp((size_t)n * ps * 4), bounds(6 * (size_t)n);
  for (int k = 0; k < n; k++)
    std::memcpy(&cp[(size_t)k * ps * 4], patches[k].control_point, sizeof(float) * 4 * ps);
  int rc = rtn_patch_bounds(cp.data(), n, p, q, bounds.data());
  std::vector<uint32_t> perm(n);
  int used = 0;
  if (rc == RTN_OK) rc = rtn_build_bvh(bounds.data(), n, nodes, perm.data(), &used);
  if (rc != RTN_OK) {
    report(rc, "buildBVH");
    return;
  }
  std::vector<rtn_bezier_patch> tmp(patches, patches + n);
  for (int i = 0; i < n; i++) patches[i] = tmp[perm[i]];
}

/* NurbsOps::curvatureRefinementU / V (nurbs.h:290-486): inserts the
 * curvature-driven knots of one direction; like the reference it takes
 * ownership of the caller's new[] knot and control-point arrays and replaces
 * them (and the count) with new[] arrays of the refined surface. */
inline void curvatureRefinement(int dir, float*& K, int deg, int& n_dir, int n_other, float C, rtn_vec4*& P) {
  int inserted = 0;
  const int n = dir == 0 ? n_dir : n_other, m = dir == 0 ? n_other : n_dir;
  int rc = rtn_refine(dir, n, deg, K, m, reinterpret_cast<const float*>(P), C, &inserted, nullptr, nullptr);
  if (rc != RTN_OK || inserted == 0) {
    report(rc, "curvatureRefinement");
    return;
  }
  float* K2 = new float[(size_t)n_dir + inserted + deg + 2];
  const size_t ncp = (size_t)(n + (dir == 0 ? inserted : 0) + 1) * (m + (dir == 1 ? inserted : 0) + 1);
  rtn_vec4* P2 = new rtn_vec4[ncp];
  rc = rtn_refine(dir, n, deg, K, m, reinterpret_cast<const float*>(P), C, &inserted, K2,
                  reinterpret_cast<float*>(P2));
  if (rc != RTN_OK) {
    report(rc, "curvatureRefinement");
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
inline void curvatureRefinementU(float*& U, int p, int& n, int& m, float C, rtn_vec4*& P) {
  curvatureRefinement(0, U, p, n, m, C, P);
}
inline void curvatureRefinementV(float*& V, int q, int& m, int n, float C, rtn_vec4*& P) {
  curvatureRefinement(1, V, q, m, n, C, P);
}

}  // namespace rtnurbs

#endif /* RTNURBS_HPP */
