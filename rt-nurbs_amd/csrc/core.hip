// Core C-ABI: error state, memory, streams, camera and primary-ray generation.
#include <cmath>
#include <cstring>

#include "rtn_exact.hpp"
#include "rtn_internal.hpp"

namespace rtn {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s:%d %s: %s", file, line, what, hipGetErrorString(e));
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? RTN_ENOMEM : RTN_EHIP;
}

}  // namespace rtn

using namespace rtn;

extern "C" const char* rtn_last_error(void) { return g_err.c_str(); }
extern "C" int rtn_version(void) { return 1; }

// ----------------------------------------------------------------- memory
extern "C" int rtn_malloc(void** ptr, size_t bytes) {
  RTN_CHECK(ptr, RTN_EINVAL, "rtn_malloc: null out pointer");
  RTN_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return RTN_OK;
}
extern "C" int rtn_malloc_managed(void** ptr, size_t bytes) {
  RTN_CHECK(ptr, RTN_EINVAL, "rtn_malloc_managed: null out pointer");
  RTN_HIP(hipMallocManaged(ptr, bytes ? bytes : 1, hipMemAttachGlobal));
  return RTN_OK;
}
extern "C" int rtn_free(void* ptr) {
  if (ptr) RTN_HIP(hipFree(ptr));
  return RTN_OK;
}
extern "C" int rtn_malloc_host(void** ptr, size_t bytes) {
  RTN_CHECK(ptr, RTN_EINVAL, "rtn_malloc_host: null out pointer");
  RTN_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return RTN_OK;
}
extern "C" int rtn_free_host(void* ptr) {
  if (ptr) RTN_HIP(hipHostFree(ptr));
  return RTN_OK;
}
extern "C" int rtn_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  static const hipMemcpyKind kinds[4] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                                         hipMemcpyDefault};
  RTN_CHECK(kind >= 0 && kind < 4, RTN_EINVAL, "rtn_memcpy: bad kind %d", kind);
  if (bytes) RTN_HIP(hipMemcpy(dst, src, bytes, kinds[kind]));
  return RTN_OK;
}
extern "C" int rtn_memset(void* dst, int value, size_t bytes) {
  if (bytes) RTN_HIP(hipMemset(dst, value, bytes));
  return RTN_OK;
}
extern "C" int rtn_device_synchronize(void) {
  RTN_HIP(hipDeviceSynchronize());
  return RTN_OK;
}
extern "C" int rtn_stream_create(void** stream) {
  RTN_CHECK(stream, RTN_EINVAL, "rtn_stream_create: null out pointer");
  hipStream_t s;
  RTN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return RTN_OK;
}
extern "C" int rtn_stream_destroy(void* stream) {
  if (stream) RTN_HIP(hipStreamDestroy((hipStream_t)stream));
  return RTN_OK;
}
extern "C" int rtn_stream_synchronize(void* stream) {
  RTN_HIP(hipStreamSynchronize((hipStream_t)stream));
  return RTN_OK;
}

// ----------------------------------------------------------------- events
extern "C" int rtn_event_create(void** ev) {
  RTN_CHECK(ev, RTN_EINVAL, "rtn_event_create: null out pointer");
  hipEvent_t e;
  RTN_HIP(hipEventCreate(&e));
  *ev = (void*)e;
  return RTN_OK;
}
extern "C" int rtn_event_destroy(void* ev) {
  if (ev) RTN_HIP(hipEventDestroy((hipEvent_t)ev));
  return RTN_OK;
}
extern "C" int rtn_event_record(void* ev, void* stream) {
  RTN_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return RTN_OK;
}
extern "C" int rtn_event_elapsed_ms(void* start, void* stop, float* ms) {
  RTN_CHECK(ms, RTN_EINVAL, "rtn_event_elapsed_ms: null out pointer");
  RTN_HIP(hipEventSynchronize((hipEvent_t)stop));
  RTN_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return RTN_OK;
}

// ----------------------------------------------------------------- camera
// Column-major 4x4 helpers in the reference's order (matrix_ops.h:101-171,
// transform.h:14-38).
namespace {
void mat_mul(const float* a, const float* b, float* r) {
  float out[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      float sum = 0;
      for (int k = 0; k < 4; k++) sum += a[k * 4 + i] * b[j * 4 + k];
      out[j * 4 + i] = sum;
    }
  memcpy(r, out, sizeof(out));
}
float det3(float m0, float m1, float m2, float m3, float m4, float m5, float m6, float m7, float m8) {
  return m0 * m4 * m8 + m3 * m7 * m2 + m6 * m1 * m5 - m6 * m4 * m2 - m3 * m1 * m8 - m0 * m7 * m5;
}
void mat_inverse(const float* m, float* r) {
  float a[16];
  a[0] = det3(m[5], m[6], m[7], m[9], m[10], m[11], m[13], m[14], m[15]);
  a[4] = -det3(m[4], m[6], m[7], m[8], m[10], m[11], m[12], m[14], m[15]);
  a[8] = det3(m[4], m[5], m[7], m[8], m[9], m[11], m[12], m[13], m[15]);
  a[12] = -det3(m[4], m[5], m[6], m[8], m[9], m[10], m[12], m[13], m[14]);
  a[1] = -det3(m[1], m[2], m[3], m[9], m[10], m[11], m[13], m[14], m[15]);
  a[5] = det3(m[0], m[2], m[3], m[8], m[10], m[11], m[12], m[14], m[15]);
  a[9] = -det3(m[0], m[1], m[3], m[8], m[9], m[11], m[12], m[13], m[15]);
  a[13] = det3(m[0], m[1], m[2], m[8], m[9], m[10], m[12], m[13], m[14]);
  a[2] = det3(m[1], m[2], m[3], m[5], m[6], m[7], m[13], m[14], m[15]);
  a[6] = -det3(m[0], m[2], m[3], m[4], m[6], m[7], m[12], m[14], m[15]);
  a[10] = det3(m[0], m[1], m[3], m[4], m[5], m[7], m[12], m[13], m[15]);
  a[14] = -det3(m[0], m[1], m[2], m[4], m[5], m[6], m[12], m[13], m[14]);
  a[3] = -det3(m[1], m[2], m[3], m[5], m[6], m[7], m[9], m[10], m[11]);
  a[7] = det3(m[0], m[2], m[3], m[4], m[6], m[7], m[8], m[10], m[11]);
  a[11] = -det3(m[0], m[1], m[3], m[4], m[5], m[7], m[8], m[9], m[11]);
  a[15] = det3(m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]);
  float det = m[0] * a[0] + m[1] * a[4] + m[2] * a[8] + m[3] * a[12];
  float ad = det < 0.0f ? -det : det;
  if (ad < 1e-8f) {  // identity fallback (matrix_ops.h:159-161)
    for (int i = 0; i < 16; i++) r[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    return;
  }
  float di = 1.0f / det;
  for (int i = 0; i < 16; i++) r[i] = a[i] * di;
}
void mat_diag(float sx, float sy, float sz, float* m) {
  memset(m, 0, 64);
  m[0] = sx; m[5] = sy; m[10] = sz; m[15] = 1.0f;
}
void mat_translate(float tx, float ty, float tz, float* m) {
  mat_diag(1.0f, 1.0f, 1.0f, m);
  m[12] = tx; m[13] = ty; m[14] = tz;
}

// Transform::lookAt (transform.h:46-60)
void lookat(const float* eye, const float* look, const float* up, float* out) {
  using namespace rtn::ex;
  using V = v3<float>;
  V pos{eye[0], eye[1], eye[2]};
  V dir = normalize(sub(V{look[0], look[1], look[2]}, pos));
  V right = normalize(cross(V{up[0], up[1], up[2]}, dir));
  V nu = cross(dir, right);
  const float m[16] = {right.x, right.y, right.z, 0.0f, nu.x, nu.y, nu.z, 0.0f,
                       dir.x,   dir.y,   dir.z,   0.0f, pos.x, pos.y, pos.z, 1.0f};
  memcpy(out, m, sizeof(m));
}

// init_camera (camera.h:123-141) with initCameraTransforms (:89-119): a zeroed
// configuration, the screen window from the aspect ratio, then the raster <->
// screen transforms; screen_from_camera and camera_from_raster stay zero.
void camera_init(const float* wfc, int W, int H, rtn_camera_config* c) {
  memset(c, 0, sizeof(*c));
  memcpy(c->world_from_camera.m, wfc, 64);
  c->screen_resolution_x = W;
  c->screen_resolution_y = H;
  const float aspect = float(W) / float(H);
  float* win = c->screen_window;  // min x, min y, max x, max y
  if (aspect > 1.0f) { win[0] = -aspect; win[1] = -1.0f; win[2] = aspect; win[3] = 1.0f; }
  else { win[0] = -1.0f; win[1] = -1.0f / aspect; win[2] = 1.0f; win[3] = 1.0f / aspect; }
  float sc[16], tr[16];
  mat_diag(1 / (win[2] - win[0]), 1 / (win[3] - win[1]), 1, sc);
  mat_translate(-win[0], -win[3], 0, tr);
  mat_mul(sc, tr, c->NDC_from_screen.m);
  mat_diag(float(W), float(-H), 1, c->raster_from_NDC.m);
  mat_mul(c->raster_from_NDC.m, c->NDC_from_screen.m, c->raster_from_screen.m);
  mat_inverse(c->raster_from_screen.m, c->screen_from_raster.m);
}

// getPerspectiveCamera (camera.h:63-85): pbrt's perspective transform scaled
// by 1/tan(fov/2), fov in degrees through MathUtils::radians.
void camera_perspective(const rtn_camera_config* in, float fov, float n, float f, rtn_camera_config* out) {
  const rtn_camera_config c = *in;
  const float pt[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, f / (f - n), 1, 0, 0, -f * n / (f - n), 0};
  const float PI = 3.14159265f;
  const float inv_tan = 1.0f / tanf(((PI / 180) * fov) / 2);
  float sc[16], inv[16];
  *out = c;
  mat_diag(inv_tan, inv_tan, 1, sc);
  mat_mul(sc, pt, out->screen_from_camera.m);
  mat_inverse(out->screen_from_camera.m, inv);
  mat_mul(inv, c.screen_from_raster.m, out->camera_from_raster.m);
}

// getOrthographicCamera (camera.h:42-59): screen_from_camera maps z_near..z_far to 0..1.
void camera_orthographic(const rtn_camera_config* in, float zn, float zf, rtn_camera_config* out) {
  const rtn_camera_config c = *in;
  float sc[16], tr[16], inv[16];
  *out = c;
  mat_diag(1, 1, 1 / (zf - zn), sc);
  mat_translate(0, 0, -zn, tr);
  mat_mul(sc, tr, out->screen_from_camera.m);
  mat_inverse(out->screen_from_camera.m, inv);
  mat_mul(inv, c.screen_from_raster.m, out->camera_from_raster.m);
}
}  // namespace

extern "C" int rtn_lookat(const float pos[3], const float look[3], const float up[3], float wfc[16]) {
  RTN_CHECK(pos && look && up && wfc, RTN_EINVAL, "rtn_lookat: null argument");
  lookat(pos, look, up, wfc);
  return RTN_OK;
}

extern "C" int rtn_camera_init(const float wfc[16], int res_x, int res_y, rtn_camera_config* out) {
  RTN_CHECK(wfc && out && res_x > 0 && res_y > 0, RTN_EINVAL, "rtn_camera_init: bad arguments");
  camera_init(wfc, res_x, res_y, out);
  return RTN_OK;
}

extern "C" int rtn_camera_get_perspective(const rtn_camera_config* in, float fov_deg, float z_near, float z_far,
                                          rtn_camera_config* out) {
  RTN_CHECK(in && out, RTN_EINVAL, "rtn_camera_get_perspective: null argument");
  camera_perspective(in, fov_deg, z_near, z_far, out);
  return RTN_OK;
}

extern "C" int rtn_camera_get_orthographic(const rtn_camera_config* in, float z_near, float z_far,
                                           rtn_camera_config* out) {
  RTN_CHECK(in && out, RTN_EINVAL, "rtn_camera_get_orthographic: null argument");
  camera_orthographic(in, z_near, z_far, out);
  return RTN_OK;
}

extern "C" int rtn_camera_perspective(const float eye[3], const float look[3], const float up[3], int W, int H,
                                      float fov_deg, float znear, float zfar, float wfc[16], float cfr[16]) {
  RTN_CHECK(eye && look && up && wfc && cfr && W > 0 && H > 0, RTN_EINVAL, "rtn_camera_perspective: bad arguments");
  lookat(eye, look, up, wfc);
  rtn_camera_config c;
  camera_init(wfc, W, H, &c);
  camera_perspective(&c, fov_deg, znear, zfar, &c);
  memcpy(cfr, c.camera_from_raster.m, 64);
  return RTN_OK;
}

// ------------------------------------------------------------------ rays
struct CamPayload {  // camera_gpu_payload (camera.h:32-35)
  float wfc[16];
  float cfr[16];
};

// calculateRays<perspective_projection> (generateRaysKernel.cu:15-35,
// camera.h:162-174).  One thread per ray, 1D over W*H*spp; rays of a pixel are
// consecutive.  The camera travels as a kernel argument (SGPR-resident)
// instead of a process-global __constant__ symbol, so calls are re-entrant.
__global__ void __launch_bounds__(256) k_rays_perspective(CamPayload cam, int W, int H, int spp, int jit, int tiled,
                                                          uint32_t seed, rtn_ray_queue q) {
  const int64_t n = (int64_t)W * H * spp;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t pix = r / spp;
  int x, y;
  if (tiled) {
    rtn::tiled_pixel(pix, W, x, y);
  } else {
    x = (int)(pix % W);
    y = (int)(pix / W);
  }
  const float ox = jit ? ex::jitter((uint32_t)r, seed, 0) : 0.5f;
  const float oy = jit ? ex::jitter((uint32_t)r, seed, 1) : 0.5f;
  const float e0 = float(x) + ox, e1 = float(y) + oy, e2 = 0.0f, e3 = 1.0f;
  const float* c = cam.cfr;
  float px = c[0] * e0 + c[4] * e1 + c[8] * e2 + c[12] * e3;
  float py = c[1] * e0 + c[5] * e1 + c[9] * e2 + c[13] * e3;
  float pz = c[2] * e0 + c[6] * e1 + c[10] * e2 + c[14] * e3;
  ex::v3<float> pc = ex::normalize(ex::v3<float>{px, py, pz});
  const float* w = cam.wfc;
  float dx = w[0] * pc.x + w[4] * pc.y + w[8] * pc.z + w[12] * 0.0f;
  float dy = w[1] * pc.x + w[5] * pc.y + w[9] * pc.z + w[13] * 0.0f;
  float dz = w[2] * pc.x + w[6] * pc.y + w[10] * pc.z + w[14] * 0.0f;
  q.origin_x[r] = w[12];
  q.origin_y[r] = w[13];
  q.origin_z[r] = w[14];
  q.direction_x[r] = dx;
  q.direction_y[r] = dy;
  q.direction_z[r] = dz;
  if (q.idx_interact) q.idx_interact[r] = -1;
}

extern "C" int rtn_rays_perspective(const float wfc[16], const float cfr[16], int W, int H, int spp, int flags,
                                    uint32_t jitter_seed, const rtn_ray_queue* rays, void* stream) {
  RTN_CHECK(wfc && cfr && rays && W > 0 && H > 0 && spp > 0, RTN_EINVAL, "rtn_rays_perspective: bad arguments");
  RTN_CHECK((flags & ~(RTN_RAYS_JITTER | RTN_RAYS_TILED)) == 0, RTN_EINVAL, "rtn_rays_perspective: unknown flags");
  const int jitter = (flags & RTN_RAYS_JITTER) ? 1 : 0, tiled = (flags & RTN_RAYS_TILED) ? 1 : 0;
  RTN_CHECK(!tiled || (W % 8 == 0 && H % 8 == 0), RTN_EINVAL,
            "rtn_rays_perspective: the tiled layout needs W and H multiples of 8");
  RTN_CHECK(rays->origin_x && rays->origin_y && rays->origin_z && rays->direction_x && rays->direction_y &&
                rays->direction_z,
            RTN_EINVAL, "rtn_rays_perspective: null ray buffer");
  const int64_t n = (int64_t)W * H * spp;
  RTN_CHECK(n < (int64_t)1 << 32, RTN_EINVAL, "rtn_rays_perspective: too many rays");
  CamPayload cam;
  memcpy(cam.wfc, wfc, 64);
  memcpy(cam.cfr, cfr, 64);
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_rays_perspective, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, cam, W, H, spp,
                     jitter, tiled, jitter_seed, *rays);
  RTN_HIP(hipGetLastError());
  if (!stream) RTN_HIP(hipStreamSynchronize(nullptr));
  return RTN_OK;
}

// calculateRays<orthographic_projection> (generateRaysKernel.cu:15-35,
// camera.h:150-160): the pixel centre's camera-space point taken to world
// space is the origin, the camera's z axis the direction.  Row-major pixels.
__global__ void __launch_bounds__(256) k_rays_orthographic(CamPayload cam, int W, int H, rtn_ray_queue q) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (int64_t)W * H) return;
  const int x = (int)(r % W), y = (int)(r / W);
  const float e0 = float(x) + 0.5f, e1 = float(y) + 0.5f, e2 = 0.0f, e3 = 1.0f;
  const float* c = cam.cfr;
  const float p0 = c[0] * e0 + c[4] * e1 + c[8] * e2 + c[12] * e3;
  const float p1 = c[1] * e0 + c[5] * e1 + c[9] * e2 + c[13] * e3;
  const float p2 = c[2] * e0 + c[6] * e1 + c[10] * e2 + c[14] * e3;
  const float p3 = c[3] * e0 + c[7] * e1 + c[11] * e2 + c[15] * e3;
  const float* w = cam.wfc;
  q.origin_x[r] = w[0] * p0 + w[4] * p1 + w[8] * p2 + w[12] * p3;
  q.origin_y[r] = w[1] * p0 + w[5] * p1 + w[9] * p2 + w[13] * p3;
  q.origin_z[r] = w[2] * p0 + w[6] * p1 + w[10] * p2 + w[14] * p3;
  q.direction_x[r] = w[8];
  q.direction_y[r] = w[9];
  q.direction_z[r] = w[10];
  if (q.idx_interact) q.idx_interact[r] = -1;
}

extern "C" int rtn_rays_generate(const rtn_ray_queue* rays, int projection, const rtn_camera_config* cam,
                                 void* stream) {
  RTN_CHECK(rays && cam, RTN_EINVAL, "rtn_rays_generate: null argument");
  RTN_CHECK(projection == RTN_PROJ_ORTHOGRAPHIC || projection == RTN_PROJ_PERSPECTIVE, RTN_EINVAL,
            "rtn_rays_generate: unknown projection %d", projection);
  const int W = cam->screen_resolution_x, H = cam->screen_resolution_y;
  RTN_CHECK(W > 0 && H > 0, RTN_EINVAL, "rtn_rays_generate: camera resolution %d x %d", W, H);
  if (projection == RTN_PROJ_PERSPECTIVE)
    return rtn_rays_perspective(cam->world_from_camera.m, cam->camera_from_raster.m, W, H, 1, 0, 0u, rays, stream);
  RTN_CHECK(rays->origin_x && rays->origin_y && rays->origin_z && rays->direction_x && rays->direction_y &&
                rays->direction_z,
            RTN_EINVAL, "rtn_rays_generate: null ray buffer");
  const int64_t n = (int64_t)W * H;
  RTN_CHECK(n < (int64_t)1 << 32, RTN_EINVAL, "rtn_rays_generate: too many rays");
  CamPayload pay;
  memcpy(pay.wfc, cam->world_from_camera.m, 64);
  memcpy(pay.cfr, cam->camera_from_raster.m, 64);
  hipLaunchKernelGGL(k_rays_orthographic, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, pay,
                     W, H, *rays);
  RTN_HIP(hipGetLastError());
  if (!stream) RTN_HIP(hipStreamSynchronize(nullptr));
  return RTN_OK;
}
