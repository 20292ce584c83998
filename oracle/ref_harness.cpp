// oracle/ref_harness.cpp -- ctypes harness around the reference's own host-
// compilable headers (Frostbite16/rt-nurbs src/, compiled where they lie).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/build_ref.sh into oracle/_ref/
// (never committed) and used to validate oracle/oracle.cpp and to write the
// golden fixtures under tests/golden/.  Every arithmetic function called here
// is the reference's; only the per-pixel loop of intersect.cu:33-193 (a CUDA
// kernel body, not host-compilable) and the ray-generation loop of
// generateRaysKernel.cu:17-34 are restated as host control flow.

#include "core/bvh.h"
#include "core/camera.h"
#include "gpu/interact.h"
#include "gpu/soa.h"
#include "math/bezier_patch.h"
#include "math/bounds_ops.h"
#include "math/nurbs.h"

#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

void ref_lookat(const float* eye, const float* look, const float* up, float* out16) {
  MathTypes::matrix4x4f m = Transform::lookAt(MathTypes::point3f{eye[0], eye[1], eye[2]},
                                              MathTypes::point3f{look[0], look[1], look[2]},
                                              MathTypes::vector3f{up[0], up[1], up[2]});
  std::memcpy(out16, m.m, sizeof(m.m));
}

void ref_camera(const float* wfc16, int W, int H, float fov_deg, float znear, float zfar, float* cfr16) {
  MathTypes::matrix4x4f wfc;
  std::memcpy(wfc.m, wfc16, sizeof(wfc.m));
  camera_config cam = CameraOps::init_camera(wfc, W, H);
  cam = CameraOps::getPerspectiveCamera(cam, fov_deg, znear, zfar);
  std::memcpy(cfr16, cam.camera_from_raster.m, sizeof(cam.camera_from_raster.m));
}

// The whole camera_config (camera.h:13-27) of init_camera (camera.h:123-141),
// then getPerspectiveCamera (proj 1, fov in degrees) or getOrthographicCamera
// (proj 0) (camera.h:42-85), copied out as its bytes; proj < 0 stops after
// init_camera.
void ref_camera_config(const float* wfc16, int W, int H, int proj, float fov_deg, float znear, float zfar,
                       void* out) {
  MathTypes::matrix4x4f wfc;
  std::memcpy(wfc.m, wfc16, sizeof(wfc.m));
  camera_config cam = CameraOps::init_camera(wfc, W, H);
  if (proj == 1) cam = CameraOps::getPerspectiveCamera(cam, fov_deg, znear, zfar);
  if (proj == 0) cam = CameraOps::getOrthographicCamera(cam, znear, zfar);
  std::memcpy(out, &cam, sizeof(cam));
}
int ref_camera_config_size() { return (int)sizeof(camera_config); }

// generateRaysKernel.cu:17-34 loop with calculateRay<orthographic_projection>
void ref_gen_rays_ortho(const float* wfc16, const float* cfr16, int W, int H, float* ox, float* oy, float* oz,
                        float* dx, float* dy, float* dz) {
  camera_gpu_payload pay;
  std::memcpy(pay.world_from_camera.m, wfc16, 64);
  std::memcpy(pay.camera_from_raster.m, cfr16, 64);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      MathExtra::ray r = CameraOps::calculateRay<orthographic_projection>(pay, x, y);
      long i = (long)y * W + x;
      ox[i] = r.origin.v[0]; oy[i] = r.origin.v[1]; oz[i] = r.origin.v[2];
      dx[i] = r.direction.v[0]; dy[i] = r.direction.v[1]; dz[i] = r.direction.v[2];
    }
}

// generateRaysKernel.cu:17-34 loop, calling calculateRay<perspective_projection>
void ref_gen_rays(const float* wfc16, const float* cfr16, int W, int H, float* ox, float* oy, float* oz, float* dx,
                  float* dy, float* dz) {
  camera_gpu_payload pay;
  std::memcpy(pay.world_from_camera.m, wfc16, 64);
  std::memcpy(pay.camera_from_raster.m, cfr16, 64);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      MathExtra::ray r = CameraOps::calculateRay<perspective_projection>(pay, x, y);
      long i = (long)y * W + x;
      ox[i] = r.origin.v[0]; oy[i] = r.origin.v[1]; oz[i] = r.origin.v[2];
      dx[i] = r.direction.v[0]; dy[i] = r.direction.v[1]; dz[i] = r.direction.v[2];
    }
}

// getPatch (bezier_patch.h:87-139): returns nbU*nbV; fills cp (patch-major) and
// (u0,u1,v0,v1) per patch when the output pointers are non-null.
int ref_get_patch(int n, int p, const float* U, int m, int q, const float* V, const float* Pw, float* cp,
                  float* uv_range, int* nbU_out, int* nbV_out) {
  std::vector<float> u(U, U + n + p + 2), v(V, V + m + q + 2);
  std::vector<MathTypes::vector4f> P((size_t)(n + 1) * (m + 1));
  std::memcpy(P.data(), Pw, P.size() * sizeof(MathTypes::vector4f));
  ParametricSurf::nurbsSurf s{u.data(), v.data(), P.data(), n, m, p, q};
  int nbU, nbV;
  bezierPatch* pt = BezierPatchOps::getPatch(s, nbU, nbV);
  int np = nbU * nbV, ps = (p + 1) * (q + 1);
  if (cp) std::memcpy(cp, pt[0].heap_pointer, (size_t)np * ps * sizeof(MathTypes::vector4f));
  if (uv_range)
    for (int i = 0; i < np; i++) {
      uv_range[4 * i] = pt[i].u0; uv_range[4 * i + 1] = pt[i].u1;
      uv_range[4 * i + 2] = pt[i].v0; uv_range[4 * i + 3] = pt[i].v1;
    }
  *nbU_out = nbU;
  *nbV_out = nbV;
  BezierPatchOps::freePatches(pt);
  return np;
}

// curvatureRefinementU/V (nurbs.h:290-486).  In/out arrays are caller-sized
// for the refined result; returns the new last index (n or m).
int ref_refine_u(float* U, int p, int n, int m, float C, float* Pw, float* U_out, float* Pw_out) {
  float* Uh = new float[n + p + 2];
  std::memcpy(Uh, U, (n + p + 2) * sizeof(float));
  MathTypes::vector4f* Ph = new MathTypes::vector4f[(size_t)(n + 1) * (m + 1)];
  std::memcpy(Ph, Pw, (size_t)(n + 1) * (m + 1) * 16);
  int nn = n, mm = m;
  NurbsOps::curvatureRefinementU(Uh, p, nn, mm, C, Ph);
  if (U_out) std::memcpy(U_out, Uh, (nn + p + 2) * sizeof(float));
  if (Pw_out) std::memcpy(Pw_out, Ph, (size_t)(nn + 1) * (mm + 1) * 16);
  delete[] Uh;
  delete[] Ph;
  return nn;
}
int ref_refine_v(float* V, int q, int m, int n, float C, float* Pw, float* V_out, float* Pw_out) {
  float* Vh = new float[m + q + 2];
  std::memcpy(Vh, V, (m + q + 2) * sizeof(float));
  MathTypes::vector4f* Ph = new MathTypes::vector4f[(size_t)(n + 1) * (m + 1)];
  std::memcpy(Ph, Pw, (size_t)(n + 1) * (m + 1) * 16);
  int mm = m;
  NurbsOps::curvatureRefinementV(Vh, q, mm, n, C, Ph);
  if (V_out) std::memcpy(V_out, Vh, (mm + q + 2) * sizeof(float));
  if (Pw_out) std::memcpy(Pw_out, Ph, (size_t)(n + 1) * (mm + 1) * 16);
  delete[] Vh;
  delete[] Ph;
  return mm;
}

// buildBVH (bvh.h:148-179) over np patches of one degree; patch i's net at
// cp + i*(p+1)*(q+1)*4.  nodes_out: (2np-1) x 32 B.  perm[i] = original index of
// the patch at BVH position i (recovered from the permuted control pointers).
int ref_build_bvh(const float* cp, int np, int p, int q, void* nodes_out, uint32_t* perm) {
  int ps = (p + 1) * (q + 1);
  std::vector<MathTypes::vector4f> block((size_t)np * ps);
  std::memcpy(block.data(), cp, block.size() * 16);
  bezierPatch* pt = new bezierPatch[np];
  for (int i = 0; i < np; i++) {
    pt[i].control_point = block.data() + (size_t)i * ps;
    pt[i].p = p; pt[i].q = q;
    pt[i].u0 = pt[i].u1 = pt[i].v0 = pt[i].v1 = 0;
    pt[i].heap_pointer = nullptr;
  }
  BVHNode* nodes = reinterpret_cast<BVHNode*>(nodes_out);
  std::memset(nodes, 0, sizeof(BVHNode) * (size_t)(2 * np - 1));
  BvhOps::buildBVH(nodes, 0, pt, np);
  for (int i = 0; i < np; i++) perm[i] = (uint32_t)((pt[i].control_point - block.data()) / ps);
  delete[] pt;
  return (int)sizeof(BVHNode);
}

void ref_patch_bounds(const float* cp, int p, int q, float* out6) {
  bezierPatch pt{};
  pt.control_point = (MathTypes::vector4f*)cp;
  pt.p = p; pt.q = q;
  MathExtra::bounds3f b = BezierPatchOps::bounds(pt);
  out6[0] = b.p_min_x; out6[1] = b.p_min_y; out6[2] = b.p_min_z;
  out6[3] = b.p_max_x; out6[4] = b.p_max_y; out6[5] = b.p_max_z;
}

// intersect.cu:33-193 restated as host control flow; every arithmetic call is
// the reference's.  Patches of one degree (p,q), in BVH order, nets at
// cp + i*(p+1)(q+1)*4.  Outputs per ray: hit_prim (BVH position or -1), t, u, v,
// dpdu, dpdv; newton[r] = number of Newton calls.
void ref_trace(const void* nodes_in, const float* cp, int np, int p, int q, const float* delta, float flat_threshold,
               float cam_k, long n_rays, const float* ox, const float* oy, const float* oz, const float* dx,
               const float* dy, const float* dz, int* hit_prim, float* t_out, float* u_out, float* v_out,
               float* dpdu, float* dpdv, int* newton) {
  const BVHNode* nodes = reinterpret_cast<const BVHNode*>(nodes_in);
  int ps = (p + 1) * (q + 1);
  std::vector<bezierPatch> patches(np);
  for (int i = 0; i < np; i++) {
    patches[i].control_point = (MathTypes::vector4f*)(cp + (size_t)i * ps * 4);
    patches[i].p = p; patches[i].q = q;
    patches[i].u0 = patches[i].u1 = patches[i].v0 = patches[i].v1 = 0;
    patches[i].heap_pointer = nullptr;
  }
  using namespace VectorOps;
  for (long idx = 0; idx < n_rays; idx++) {
    MathExtra::ray ray{MathTypes::vector3f{ox[idx], oy[idx], oz[idx]}, MathTypes::vector3f{dx[idx], dy[idx], dz[idx]}};
    int nc = 0;
    MathTypes::vector3f rdr{1 / ray.direction.v[0], 1 / ray.direction.v[1], 1 / ray.direction.v[2]};
    MathExtra::ray_plane rp = MathUtils::getRayPlanes(ray);
    int stack[64];
    int off = 0;
    float t_max = INFINITY;
    int best = -1;
    BVHNode cur = nodes[0];
    const MathExtra::bounds3f* b = &cur.node_bound;
    float it = BoundsOps::intersectBound(b->p_min_x, b->p_max_x, b->p_min_y, b->p_max_y, b->p_min_z, b->p_max_z,
                                         ray.origin.v[0], ray.origin.v[1], ray.origin.v[2], rdr.v[0], rdr.v[1],
                                         rdr.v[2]);
    int go = it < t_max;
    while (go) {
      if (cur.prim_count > 0) {
        it = BoundsOps::intersectBound(cur.node_bound.p_min_x, cur.node_bound.p_max_x, cur.node_bound.p_min_y,
                                       cur.node_bound.p_max_y, cur.node_bound.p_min_z, cur.node_bound.p_max_z,
                                       ray.origin.v[0], ray.origin.v[1], ray.origin.v[2], rdr.v[0], rdr.v[1], rdr.v[2]);
        for (int i = 0; i < (int)cur.prim_count; i++) {
          bool do_newton = true;
          bezierPatch patch = patches[cur.left_first + i];
          MathTypes::vector3f q00, q10, q11, q01;
          corners(patch, q00, q10, q11, q01);
          MathTypes::vector2f g = BezierPatchOps::biliniearRay(q00, q10, q11, q01, ray.origin, ray.direction);
          if (delta && delta[cur.left_first + i] * cam_k / it < flat_threshold) {
            MathTypes::vector3f S, Su, Sv;
            BezierPatchOps::evalPointNormal(patch, g.v[0], g.v[1], ray.direction, S, Su, Sv);
            float resid = fabsf(dotProduct(rp.n1, S) + rp.d1) + fabsf(dotProduct(rp.n2, S) + rp.d2);
            if (resid * cam_k / it < flat_threshold) {
              float t = dotProduct(subtract(S, ray.origin), ray.direction);
              if (0 < t && t < t_max) {
                t_max = t; best = cur.left_first + i;
                t_out[idx] = t; u_out[idx] = g.v[0]; v_out[idx] = g.v[1];
                std::memcpy(dpdu + 3 * idx, Su.v, 12); std::memcpy(dpdv + 3 * idx, Sv.v, 12);
                do_newton = false;
              }
            }
          }
          if (do_newton) {
            nc++;
            newtonRhapsonPayload nr = BezierPatchOps::newtonRhapson(rp, patch, g.v[0], g.v[1], 1e-5f, 1e-4f, idx);
            if (nr.hit) {
              float t = (nr.s.v[0] - ray.origin.v[0]) * ray.direction.v[0] +
                        (nr.s.v[1] - ray.origin.v[1]) * ray.direction.v[1] +
                        (nr.s.v[2] - ray.origin.v[2]) * ray.direction.v[2];
              if (0 < t && t < t_max) {
                std::memcpy(dpdu + 3 * idx, nr.su.v, 12); std::memcpy(dpdv + 3 * idx, nr.sv.v, 12);
                u_out[idx] = nr.u; v_out[idx] = nr.v; t_out[idx] = t;
                best = cur.left_first + i;
                t_max = t;
              }
            }
          }
        }
        if (off == 0) break;
        cur = nodes[stack[--off]];
      } else {
        MathExtra::bounds3f lb = nodes[cur.left_first].node_bound, rb = nodes[cur.left_first + 1].node_bound;
        float lt = BoundsOps::intersectBound(lb.p_min_x, lb.p_max_x, lb.p_min_y, lb.p_max_y, lb.p_min_z, lb.p_max_z,
                                             ray.origin.v[0], ray.origin.v[1], ray.origin.v[2], rdr.v[0], rdr.v[1],
                                             rdr.v[2]);
        float rt = BoundsOps::intersectBound(rb.p_min_x, rb.p_max_x, rb.p_min_y, rb.p_max_y, rb.p_min_z, rb.p_max_z,
                                             ray.origin.v[0], ray.origin.v[1], ray.origin.v[2], rdr.v[0], rdr.v[1],
                                             rdr.v[2]);
        it = fminf(lt, rt);
        if (it >= t_max) {
          if (off == 0) break;
          cur = nodes[stack[--off]];
        } else if (lt <= it) {
          if (rt < t_max && off < 64) stack[off++] = cur.left_first + 1;
          cur = nodes[cur.left_first];
        } else {
          if (lt < t_max && off < 64) stack[off++] = cur.left_first;
          cur = nodes[cur.left_first + 1];
        }
      }
    }
    hit_prim[idx] = best;
    newton[idx] = nc;
  }
}

// Single-call wrappers of reference device functions, for unit comparison.
void ref_bilinear_seed(const float* cp, int p, int q, const float* o3, const float* d3, float* out2) {
  bezierPatch pt{};
  pt.control_point = (MathTypes::vector4f*)cp;
  pt.p = p; pt.q = q;
  MathTypes::vector3f q00, q10, q11, q01;
  corners(pt, q00, q10, q11, q01);
  MathTypes::vector2f g = BezierPatchOps::biliniearRay(q00, q10, q11, q01, MathTypes::vector3f{o3[0], o3[1], o3[2]},
                                                       MathTypes::vector3f{d3[0], d3[1], d3[2]});
  out2[0] = g.v[0]; out2[1] = g.v[1];
}
int ref_newton(const float* cp, int p, int q, const float* o3, const float* d3, float u0, float v0, uint32_t seed,
               float* out11) {
  bezierPatch pt{};
  pt.control_point = (MathTypes::vector4f*)cp;
  pt.p = p; pt.q = q;
  MathExtra::ray ray{MathTypes::vector3f{o3[0], o3[1], o3[2]}, MathTypes::vector3f{d3[0], d3[1], d3[2]}};
  MathExtra::ray_plane rp = MathUtils::getRayPlanes(ray);
  newtonRhapsonPayload r = BezierPatchOps::newtonRhapson(rp, pt, u0, v0, 1e-5f, 1e-4f, seed);
  float o[11] = {r.s.v[0], r.s.v[1], r.s.v[2], r.su.v[0], r.su.v[1], r.su.v[2], r.sv.v[0], r.sv.v[1], r.sv.v[2], r.u, r.v};
  std::memcpy(out11, o, sizeof(o));
  return r.hit ? 1 : 0;
}
void ref_ray_planes(const float* o3, const float* d3, float* out8) {
  MathExtra::ray ray{MathTypes::vector3f{o3[0], o3[1], o3[2]}, MathTypes::vector3f{d3[0], d3[1], d3[2]}};
  MathExtra::ray_plane rp = MathUtils::getRayPlanes(ray);
  float o[8] = {rp.n1.v[0], rp.n1.v[1], rp.n1.v[2], rp.n2.v[0], rp.n2.v[1], rp.n2.v[2], rp.d1, rp.d2};
  std::memcpy(out8, o, sizeof(o));
}
void ref_patch_eval(const float* cp, int p, int q, float u, float v, float* out12) {
  MathTypes::vector4f ders[2], point;
  BezierPatchOps::BezierPatchEval((const MathTypes::vector4f*)cp, p, q, u, v, ders, point);
  float o[12] = {point.v[0], point.v[1], point.v[2], point.v[3], ders[0].v[0], ders[0].v[1],
                 ders[0].v[2], ders[0].v[3], ders[1].v[0], ders[1].v[1], ders[1].v[2], ders[1].v[3]};
  std::memcpy(out12, o, sizeof(o));
}
uint32_t ref_pcg_hash(uint32_t x) { return pcg_hash(x); }

}  // extern "C"
