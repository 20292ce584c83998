// A reference-style driver (main.cpp:45-204 shape) written against the
// reference-named shims of rtnurbs.hpp: getPatch, buildBVH,
// calculateRaysWrapper, calculateIntersectWrapper, with managed memory as
// the reference drivers use.  Reads one NURBS surface and a camera from a
// binary file, writes t_hit, uv and idx_interact per pixel.
//
// input  (little endian): int32 n,p,m,q,W,H; float U[n+p+2], V[m+q+2],
//        Pw[4(n+1)(m+1)], eye[3], look[3], up[3], fov_deg
// output: int32 W,H; then per pixel float t_hit, u, v and int32 idx_interact
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rtnurbs.hpp"

template <class T>
static T* managed(size_t n) {
  void* p = nullptr;
  if (rtn_malloc_managed(&p, n * sizeof(T)) != RTN_OK) {
    std::fprintf(stderr, "managed allocation failed: %s\n", rtn_last_error());
    std::exit(2);
  }
  return static_cast<T*>(p);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[6];
  if (std::fread(hdr, 4, 6, f) != 6) return 2;
  const int n = hdr[0], p = hdr[1], m = hdr[2], q = hdr[3], W = hdr[4], H = hdr[5];
  std::vector<float> U(n + p + 2), V(m + q + 2), Pw(4 * (size_t)(n + 1) * (m + 1)), cam(10);
  if (std::fread(U.data(), 4, U.size(), f) != U.size() || std::fread(V.data(), 4, V.size(), f) != V.size() ||
      std::fread(Pw.data(), 4, Pw.size(), f) != Pw.size() || std::fread(cam.data(), 4, 10, f) != 10)
    return 2;
  std::fclose(f);

  // scene: Bezier patches, BVH (patches permuted into BVH order)
  int nbU = 0, nbV = 0;
  rtn_bezier_patch* patches = rtnurbs::getPatch(n, p, U.data(), m, q, V.data(), Pw.data(), nbU, nbV);
  if (!patches) return 1;
  const int np = nbU * nbV, ps = (p + 1) * (q + 1);
  rtn_bvh_node* nodes = managed<rtn_bvh_node>(2 * (size_t)np - 1);
  rtnurbs::buildBVH(nodes, 0, patches, np);
  // device-visible copies: one managed control-point block, re-based headers (main.cpp:128-134)
  rtn_vec4* cp = managed<rtn_vec4>((size_t)np * ps);
  rtn_bezier_patch* dpatch = managed<rtn_bezier_patch>(np);
  for (int k = 0; k < np; k++) {
    for (int j = 0; j < ps; j++) cp[(size_t)k * ps + j] = patches[k].control_point[j];
    dpatch[k] = patches[k];
    dpatch[k].control_point = cp + (size_t)k * ps;
    dpatch[k].heap_pointer = nullptr;
  }
  // rays and hits
  float wfc[16], cfr[16];
  rtnurbs::perspectiveCamera(&cam[0], &cam[3], &cam[6], W, H, cam[9], 0.01f, 1000.0f, wfc, cfr);
  const size_t nr = (size_t)W * H;
  rtn_ray_queue rays{managed<float>(nr), managed<float>(nr), managed<float>(nr), managed<float>(nr),
                     managed<float>(nr), managed<float>(nr), managed<int>(nr)};
  rtnurbs::calculateRaysWrapper(rays, wfc, cfr, W, H);
  rtn_surface_interaction* inter = managed<rtn_surface_interaction>(nr);
  for (size_t i = 0; i < nr; i++) inter[i] = rtn_surface_interaction{};
  unsigned long long newton = 0, skips = 0;
  rtnurbs::calculateIntersectWrapper(rays, nodes, dpatch, inter, W, H, 0.0f, 60.0f * 3.14159265f / 180.0f,
                                     nullptr, &newton, &skips);
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  const int wh[2] = {W, H};
  std::fwrite(wh, 4, 2, o);
  for (size_t i = 0; i < nr; i++) {
    const float rec[3] = {inter[i].t_hit, inter[i].uv[0], inter[i].uv[1]};
    std::fwrite(rec, 4, 3, o);
    std::fwrite(&rays.idx_interact[i], 4, 1, o);
  }
  std::fclose(o);
  std::printf("patches %d newton_calls %llu\n", np, newton);
  rtnurbs::freePatches(patches);
  return 0;
}
