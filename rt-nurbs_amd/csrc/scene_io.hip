// Scene files: the surfaces (and pre-decomposed patch groups) of an unbuilt
// scene plus an optional camera, in one little-endian binary file.  The
// reference has no scene format (its scenes are built in code, main.cpp:70-98);
// this one stores exactly what rtn_scene_add_nurbs / rtn_scene_add_patches take.
//
// Layout (all 4-byte fields, little endian):
//   char magic[8] = "RTNSCENE"; u32 version = 1; u32 n_surfaces; u32 n_groups; u32 has_camera
//   camera (has_camera = 1): f32 eye[3], look[3], up[3], fov_deg; i32 width, height
//   n_surfaces x { i32 n, p, m, q; f32 U[n+p+2]; f32 V[m+q+2]; f32 Pw[4(n+1)(m+1)] }
//   n_groups   x { i32 np, p, q; f32 cp[4 np (p+1)(q+1)] }
// Loading checks every count against the bytes left in the file before it
// allocates, and every surface through rtn_scene_add_nurbs's own validation.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "rtn_internal.hpp"

namespace {

const char kMagic[8] = {'R', 'T', 'N', 'S', 'C', 'E', 'N', 'E'};
constexpr uint32_t kVersion = 1;

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

bool put(FILE* f, const void* p, size_t bytes) { return bytes == 0 || fwrite(p, 1, bytes, f) == bytes; }

struct Reader {
  FILE* f;
  int64_t left;  // bytes not yet read
  bool get(void* p, int64_t bytes) {
    if (bytes < 0 || bytes > left) return false;
    if (bytes && fread(p, 1, (size_t)bytes, f) != (size_t)bytes) return false;
    left -= bytes;
    return true;
  }
};

}  // namespace

extern "C" int rtn_scene_save(const rtn_scene* s, const rtn_camera_desc* cam, const char* path) {
  RTN_CHECK(s && path, RTN_EINVAL, "rtn_scene_save: null argument");
  File f(fopen(path, "wb"));
  RTN_CHECK(f, RTN_EINVAL, "rtn_scene_save: cannot open %s for writing", path);
  const uint32_t head[4] = {kVersion, (uint32_t)s->surfaces.size(), (uint32_t)s->groups.size(), cam ? 1u : 0u};
  bool ok = put(f.get(), kMagic, 8) && put(f.get(), head, sizeof(head));
  if (ok && cam) {
    const float cf[10] = {cam->eye[0], cam->eye[1], cam->eye[2], cam->look[0], cam->look[1],
                          cam->look[2], cam->up[0],  cam->up[1],  cam->up[2],  cam->fov_deg};
    const int32_t wh[2] = {cam->width, cam->height};
    ok = put(f.get(), cf, sizeof(cf)) && put(f.get(), wh, sizeof(wh));
  }
  for (const rtn::HostSurface& h : s->surfaces) {
    if (!ok) break;
    const int32_t dims[4] = {h.n, h.p, h.m, h.q};
    ok = put(f.get(), dims, sizeof(dims)) && put(f.get(), h.U.data(), 4 * h.U.size()) &&
         put(f.get(), h.V.data(), 4 * h.V.size()) && put(f.get(), h.Pw.data(), 4 * h.Pw.size());
  }
  for (const rtn_scene::PatchGroup& g : s->groups) {
    if (!ok) break;
    const int32_t dims[3] = {g.np, g.p, g.q};
    ok = put(f.get(), dims, sizeof(dims)) && put(f.get(), g.cp.data(), 4 * g.cp.size());
  }
  ok = ok && fflush(f.get()) == 0;
  RTN_CHECK(ok, RTN_EINVAL, "rtn_scene_save: write to %s failed", path);
  return RTN_OK;
}

extern "C" int rtn_scene_load(const char* path, rtn_scene** out, rtn_camera_desc* cam, int* has_camera) {
  RTN_CHECK(path && out, RTN_EINVAL, "rtn_scene_load: null argument");
  *out = nullptr;
  if (has_camera) *has_camera = 0;
  File f(fopen(path, "rb"));
  RTN_CHECK(f, RTN_EINVAL, "rtn_scene_load: cannot open %s", path);
  RTN_CHECK(fseek(f.get(), 0, SEEK_END) == 0, RTN_EINVAL, "rtn_scene_load: cannot seek %s", path);
  const long size = ftell(f.get());
  RTN_CHECK(size >= 0 && fseek(f.get(), 0, SEEK_SET) == 0, RTN_EINVAL, "rtn_scene_load: cannot seek %s", path);
  Reader r{f.get(), (int64_t)size};
  char magic[8];
  uint32_t head[4];
  RTN_CHECK(r.get(magic, 8) && memcmp(magic, kMagic, 8) == 0 && r.get(head, sizeof(head)), RTN_EINVAL,
            "rtn_scene_load: %s is not a scene file", path);
  RTN_CHECK(head[0] == kVersion, RTN_EUNSUPPORTED, "rtn_scene_load: version %u not supported", head[0]);
  RTN_CHECK(head[3] <= 1, RTN_EINVAL, "rtn_scene_load: bad camera flag");
  rtn_camera_desc c{};
  if (head[3]) {
    float cf[10];
    int32_t wh[2];
    RTN_CHECK(r.get(cf, sizeof(cf)) && r.get(wh, sizeof(wh)), RTN_EINVAL, "rtn_scene_load: truncated camera");
    for (int k = 0; k < 3; k++) {
      c.eye[k] = cf[k];
      c.look[k] = cf[3 + k];
      c.up[k] = cf[6 + k];
    }
    c.fov_deg = cf[9];
    c.width = wh[0];
    c.height = wh[1];
  }
  std::unique_ptr<rtn_scene> s(new rtn_scene());
  std::vector<float> U, V, P;
  for (uint32_t i = 0; i < head[1]; i++) {
    int32_t d[4];
    RTN_CHECK(r.get(d, sizeof(d)), RTN_EINVAL, "rtn_scene_load: truncated surface %u", i);
    const int n = d[0], p = d[1], m = d[2], q = d[3];
    RTN_CHECK(n >= 0 && m >= 0 && p >= 0 && q >= 0 && n < (1 << 24) && m < (1 << 24) && p < 64 && q < 64,
              RTN_EINVAL, "rtn_scene_load: surface %u has bad dimensions", i);
    const int64_t nu = (int64_t)n + p + 2, nv = (int64_t)m + q + 2, np = 4 * (int64_t)(n + 1) * (m + 1);
    RTN_CHECK(4 * (nu + nv + np) <= r.left, RTN_EINVAL, "rtn_scene_load: truncated surface %u", i);
    U.resize(nu);
    V.resize(nv);
    P.resize(np);
    RTN_CHECK(r.get(U.data(), 4 * nu) && r.get(V.data(), 4 * nv) && r.get(P.data(), 4 * np), RTN_EINVAL,
              "rtn_scene_load: truncated surface %u", i);
    RTN_TRY(rtn_scene_add_nurbs(s.get(), n, p, U.data(), m, q, V.data(), P.data()));
  }
  for (uint32_t i = 0; i < head[2]; i++) {
    int32_t d[3];
    RTN_CHECK(r.get(d, sizeof(d)), RTN_EINVAL, "rtn_scene_load: truncated patch group %u", i);
    const int np = d[0], p = d[1], q = d[2];
    RTN_CHECK(np > 0 && p >= 1 && q >= 1 && p <= RTN_MAX_DEGREE && q <= RTN_MAX_DEGREE, RTN_EINVAL,
              "rtn_scene_load: patch group %u has bad dimensions", i);
    const int64_t ncp = 4 * (int64_t)np * (p + 1) * (q + 1);
    RTN_CHECK(4 * ncp <= r.left, RTN_EINVAL, "rtn_scene_load: truncated patch group %u", i);
    P.resize(ncp);
    RTN_CHECK(r.get(P.data(), 4 * ncp), RTN_EINVAL, "rtn_scene_load: truncated patch group %u", i);
    RTN_TRY(rtn_scene_add_patches(s.get(), np, p, q, P.data()));
  }
  RTN_CHECK(r.left == 0, RTN_EINVAL, "rtn_scene_load: %lld trailing bytes in %s", (long long)r.left, path);
  if (cam) *cam = c;
  if (has_camera) *has_camera = (int)head[3];
  *out = s.release();
  return RTN_OK;
}
