// A driver in the shape of the reference's main.cpp (main.cpp:45-204), written
// against the reference's own API names and signatures as include/rtnurbs.hpp
// declares them: Transform::lookAt, CameraOps::init_camera /
// getPerspectiveCamera / getOrthographicCamera, CudaDevAlloc /
// CudaAllocManaged / CudaFree, NurbsOps::curvatureRefinementU/V,
// BezierPatchOps::getPatch / freePatches, BvhOps::buildBVH (patches permuted
// in place, control points re-based from patches[0].heap_pointer into one
// managed block, main.cpp:121-134), calculateRaysWrapper(ray_queue,
// projection, camera) and calculateIntersectWrapper with patchDelta per patch
// (main.cpp:18-43).  Only the include and the link line differ from a build
// against the reference.  The surface and camera come from a file and the
// hits go to one, instead of the reference's built-in torus and PPM images.
//
// input  (little endian): int32 n,p,m,q,W,H,proj (0 orthographic, 1 perspective);
//        float U[n+p+2], V[m+q+2], Pw[4(n+1)(m+1)], eye[3], look[3], up[3],
//        fov_deg, refine_C (<= 0: none), flat_threshold
// output: int32 W,H,np; then per pixel float ox,oy,oz,dx,dy,dz,t_hit,u,v and
//         int32 idx_interact
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rtnurbs.hpp"

// patchDelta (main.cpp:18-43): largest distance of a control point from the
// bilinear interpolant of the patch's four corners at its node (i/p, j/q).
static float patchDelta(const bezierPatch& P) {
  auto deh = [&](int idx) {
    MathTypes::vector4f c = P.control_point[idx];
    float iw = 1.0f / c.v[3];
    return MathTypes::vector3f{c.v[0] * iw, c.v[1] * iw, c.v[2] * iw};
  };
  int p = P.p, q = P.q, row = q + 1;
  MathTypes::vector3f c00 = deh(0), c10 = deh(p * row), c01 = deh(q), c11 = deh(p * row + q);
  float d = 0.0f;
  for (int i = 0; i <= p; ++i)
    for (int j = 0; j <= q; ++j) {
      float u = (float)i / p, v = (float)j / q;
      float B[3];
      for (int k = 0; k < 3; k++)
        B[k] = (1 - u) * (1 - v) * c00.v[k] + u * (1 - v) * c10.v[k] + (1 - u) * v * c01.v[k] + u * v * c11.v[k];
      MathTypes::vector3f Pij = deh(i * row + j);
      float dx = Pij.v[0] - B[0], dy = Pij.v[1] - B[1], dz = Pij.v[2] - B[2];
      d = fmaxf(d, sqrtf(dx * dx + dy * dy + dz * dz));
    }
  return d;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[7];
  if (std::fread(hdr, 4, 7, f) != 7) return 2;
  int n = hdr[0], p = hdr[1], m = hdr[2], q = hdr[3];
  const int res_x = hdr[4], res_y = hdr[5], proj = hdr[6];
  float* U = new float[n + p + 2];
  float* V = new float[m + q + 2];
  MathTypes::vector4f* control_points = new MathTypes::vector4f[(size_t)(n + 1) * (m + 1)];
  float cam[12];
  if (std::fread(U, 4, n + p + 2, f) != (size_t)(n + p + 2) || std::fread(V, 4, m + q + 2, f) != (size_t)(m + q + 2) ||
      std::fread(control_points, 16, (size_t)(n + 1) * (m + 1), f) != (size_t)(n + 1) * (m + 1) ||
      std::fread(cam, 4, 12, f) != 12)
    return 2;
  std::fclose(f);

  // camera (main.cpp:46-61)
  MathTypes::matrix4x4f world_from_camera = Transform::lookAt(MathTypes::point3f{cam[0], cam[1], cam[2]},
                                                              MathTypes::point3f{cam[3], cam[4], cam[5]},
                                                              MathTypes::vector3f{cam[6], cam[7], cam[8]});
  camera_config main_cam = CameraOps::init_camera(world_from_camera, res_x, res_y);
  const int max_pixels = res_x * res_y;
  const float d_fov = cam[9];
  if (proj == 1)
    main_cam = CameraOps::getPerspectiveCamera(main_cam, d_fov, 0.01f, 1000.0f);
  else
    main_cam = CameraOps::getOrthographicCamera(main_cam, 0, 1000);
  SoAQueue::ray_queue ray_queue = {CudaDevAlloc<float>(max_pixels), CudaDevAlloc<float>(max_pixels),
                                   CudaDevAlloc<float>(max_pixels), CudaDevAlloc<float>(max_pixels),
                                   CudaDevAlloc<float>(max_pixels), CudaDevAlloc<float>(max_pixels),
                                   CudaDevAlloc<int>(max_pixels)};

  // surface (main.cpp:100-104)
  const float C = cam[10];
  if (C > 0) {
    NurbsOps::curvatureRefinementU(U, p, n, m, C, control_points);
    NurbsOps::curvatureRefinementV(V, q, m, n, C, control_points);
  }
  ParametricSurf::nurbsSurf nurbs_prim = {U, V, control_points, n, m, p, q};

  // patches, BVH, device copies (main.cpp:108-136)
  int nbU, nbV;
  bezierPatch* patches = BezierPatchOps::getPatch(nurbs_prim, nbU, nbV);
  if (!patches) return 1;
  int primitive_size = nbU * nbV;
  int psz = (patches[0].p + 1) * (patches[0].q + 1);
  int block_size = primitive_size * psz;
  MathTypes::vector4f* gpu_control_points = CudaAllocManaged<MathTypes::vector4f>(block_size);
  BVHNode* nodes = CudaAllocManaged<BVHNode>(2 * (nbU * nbV) - 1);
  bezierPatch* gpu_Patches = CudaAllocManaged<bezierPatch>(nbU * nbV);
  BvhOps::buildBVH(nodes, 0, patches, primitive_size);
  MathTypes::vector4f* base_pointer = patches[0].heap_pointer;
  std::memcpy(gpu_control_points, base_pointer, block_size * sizeof(MathTypes::vector4f));
  for (int i = 0; i < primitive_size; i++) {
    gpu_Patches[i] = patches[i];
    gpu_Patches[i].control_point = gpu_control_points + (patches[i].control_point - base_pointer);
    gpu_Patches[i].heap_pointer = patches[i].heap_pointer;
  }
  BezierPatchOps::freePatches(patches);

  // rays and hits (main.cpp:140-156)
  calculateRaysWrapper(ray_queue, proj == 1 ? perspective : orthographic, main_cam);
  surfaceInteraction* surfaceInteract = CudaAllocManaged<surfaceInteraction>(max_pixels);
  for (int i = 0; i < max_pixels; i++) surfaceInteract[i] = surfaceInteraction{};
  float* delta = CudaAllocManaged<float>(nbU * nbV);
  for (int i = 0; i < nbU * nbV; i++) delta[i] = patchDelta(gpu_Patches[i]);
  float flat_threshold = cam[11];
  unsigned long long* newton = CudaAllocManaged<unsigned long long>(1);
  unsigned long long* skips = CudaAllocManaged<unsigned long long>(1);
  *newton = 0;
  *skips = 0;
  calculateIntersectWrapper(ray_queue, nodes, gpu_Patches, surfaceInteract, res_x, res_y, flat_threshold,
                            MathUtils::radians(d_fov), delta, newton, skips);

  // output: rays (host copies of the device queue) and records
  std::vector<float> rays(6 * (size_t)max_pixels);
  std::vector<int> idx(max_pixels);
  float* comp[6] = {ray_queue.origin_x, ray_queue.origin_y, ray_queue.origin_z,
                    ray_queue.direction_x, ray_queue.direction_y, ray_queue.direction_z};
  for (int k = 0; k < 6; k++)
    rtnurbs::report(rtn_memcpy(&rays[(size_t)k * max_pixels], comp[k], sizeof(float) * max_pixels, 1), "copy");
  rtnurbs::report(rtn_memcpy(idx.data(), ray_queue.idx_interact, sizeof(int) * max_pixels, 1), "copy");
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  const int head[3] = {res_x, res_y, primitive_size};
  std::fwrite(head, 4, 3, o);
  for (int i = 0; i < max_pixels; i++) {
    float rec[9];
    for (int k = 0; k < 6; k++) rec[k] = rays[(size_t)k * max_pixels + i];
    rec[6] = surfaceInteract[i].t_hit;
    rec[7] = surfaceInteract[i].uv.v[0];
    rec[8] = surfaceInteract[i].uv.v[1];
    std::fwrite(rec, 4, 9, o);
    std::fwrite(&idx[i], 4, 1, o);
  }
  std::fclose(o);
  std::printf("patches %d newton_calls %llu skips %llu\n", primitive_size, *newton, *skips);

  CudaFree(nodes);
  CudaFree(gpu_control_points);
  CudaFree(gpu_Patches);
  CudaFree(delta);
  CudaFree(newton);
  CudaFree(skips);
  CudaFree(surfaceInteract);
  CudaFree(ray_queue.direction_x);
  CudaFree(ray_queue.direction_y);
  CudaFree(ray_queue.direction_z);
  CudaFree(ray_queue.origin_x);
  CudaFree(ray_queue.origin_y);
  CudaFree(ray_queue.origin_z);
  CudaFree(ray_queue.idx_interact);
  delete[] U;
  delete[] V;
  delete[] control_points;
  return 0;
}
