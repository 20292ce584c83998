// Ray producers beyond the reference's pinhole camera, and shading:
//   silhouette rays  grazing rays aimed at random surface points (mixed-
//                    precision workload; no reference counterpart)
//   secondaries      one shadow ray (towards a directional light) and one
//                    mirror ray per primary hit, in fixed kind-major slots
//                    (no reference counterpart; the reference has primary
//                    rays only)
//   Lambert shading  one directional light, spp averaged, replacing the
//                    reference's debug shaders (testShading.cu:9-55)
#include <algorithm>
#include <cmath>

#include "rtn_exact.hpp"

#include "rtn_internal.hpp"

using namespace rtn;

namespace {

// Uniform float in [0,1) from a ray index, a draw index and a seed (PCG
// chain; the same construction as the pixel jitter).
__device__ __forceinline__ float draw(uint32_t i, uint32_t k, uint32_t seed) {
  const uint32_t h = ex::pcg_hash(ex::pcg_hash(i * 8u + k) ^ (seed * 0x9E3779B9u));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ ex::v3<float> unit(ex::v3<float> a) {
  const float il = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x * il, a.y * il, a.z * il};
}

// Silhouette ray i: patch k and (u,v) uniform, S, N = unit(S_u x S_v), a
// random unit tangent T, grazing angle theta ~ U(0, max_deg), direction
// d = unit(T cos(theta) - N sin(theta)), origin S - t0 d with t0 ~ U(t0_lo, t0_hi).
__global__ void k_rays_silhouette(const float4* __restrict__ cp, const int* __restrict__ cp_off,
                                  const int* __restrict__ pq, int np, int64_t n, uint32_t seed, float max_deg,
                                  float t0_lo, float t0_hi, float* ox, float* oy, float* oz, float* dx, float* dy,
                                  float* dz, int* aim_patch, float* aim) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = (uint32_t)i;
  const int k = min(np - 1, (int)(draw(r, 0, seed) * (float)np));
  const float u = draw(r, 1, seed), v = draw(r, 2, seed);
  const float phi = 6.2831853f * draw(r, 3, seed);
  const float theta = max_deg * 0.017453292f * draw(r, 4, seed);
  const float t0 = t0_lo + (t0_hi - t0_lo) * draw(r, 5, seed);
  const int pqv = pq[k];
  ex::v4<float> pt, du, dv;
  ex::patch_eval_rt<RTN_MAX_DEGREE>(cp + cp_off[k], pqv & 0xff, (pqv >> 8) & 0xff, u, v, du, dv, pt);
  ex::v3<float> S, Su, Sv;
  ex::dehom_ders(pt, du, dv, S, Su, Sv);
  const ex::v3<float> N = unit(ex::cross(Su, Sv));
  const ex::v3<float> a = unit(Su);
  const ex::v3<float> b = ex::cross(N, a);
  const float cp_ = cosf(phi), sp = sinf(phi), ct = cosf(theta), st = sinf(theta);
  const ex::v3<float> T{cp_ * a.x + sp * b.x, cp_ * a.y + sp * b.y, cp_ * a.z + sp * b.z};
  const ex::v3<float> d = unit(ex::v3<float>{T.x * ct - N.x * st, T.y * ct - N.y * st, T.z * ct - N.z * st});
  ox[i] = S.x - t0 * d.x;
  oy[i] = S.y - t0 * d.y;
  oz[i] = S.z - t0 * d.z;
  dx[i] = d.x;
  dy[i] = d.y;
  dz[i] = d.z;
  if (aim_patch) aim_patch[i] = k;
  if (aim) {
    aim[4 * i] = u;
    aim[4 * i + 1] = v;
    aim[4 * i + 2] = theta;
    aim[4 * i + 3] = t0;
  }
}

// Unit geometric normal of a hit, facing the incoming ray (benchmark.cpp:91-92).
__device__ __forceinline__ ex::v3<float> facing_normal(const float* dpdu, const float* dpdv, int64_t r,
                                                       ex::v3<float> d) {
  const ex::v3<float> su{dpdu[3 * r], dpdu[3 * r + 1], dpdu[3 * r + 2]};
  const ex::v3<float> sv{dpdv[3 * r], dpdv[3 * r + 1], dpdv[3 * r + 2]};
  ex::v3<float> N = unit(ex::cross(su, sv));
  if (N.x * d.x + N.y * d.y + N.z * d.z > 0.0f) N = {-N.x, -N.y, -N.z};
  return N;
}

// Secondary rays of primary ray r in slots j*n + r, j the index of the kind
// among the enabled ones (shadow first): shadow ray towards the light and
// mirror ray, origins offset along the facing normal by eps * (1 + max|S|).
// Kind-major slots keep a wave's 64 secondaries of one kind and from 64
// neighbouring primaries, so they trace well as a packet.  A miss leaves its
// slots empty (kind -1), so the layout, and the Newton seeds that follow the
// ray index, are the same on every run.  *count += valid rays.  Grid-stride loop over a
// grid sized to the device: each lane sums its count and a wave adds it with
// one atomic at the end (an atomic per wave of rays serialised on the one
// counter and cost several times the kernel's memory traffic).
__global__ void k_secondaries(const float* ox, const float* oy, const float* oz, const float* dx, const float* dy,
                              const float* dz, const int* prim, const float* t, const float* dpdu,
                              const float* dpdv, int64_t n, float lx, float ly, float lz, float eps, int want_shadow,
                              int want_mirror, float* sox, float* soy, float* soz, float* sdx, float* sdy, float* sdz,
                              int* parent, int* kind, unsigned long long* count, const int* ray_scene,
                              int* out_scene) {
  const int per = want_shadow + want_mirror;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long valid = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    int64_t slot = r;
    if (prim[r] < 0) {
      for (int k = 0; k < per; k++, slot += n) {
        sox[slot] = soy[slot] = soz[slot] = 0.0f;
        sdx[slot] = sdy[slot] = sdz[slot] = 0.0f;
        parent[slot] = (int)r;
        kind[slot] = -1;
        if (out_scene) out_scene[slot] = ray_scene[r];
      }
      continue;
    }
    valid += per;
    const ex::v3<float> d{dx[r], dy[r], dz[r]};
    const float th = t[r];
    const ex::v3<float> S{ox[r] + th * d.x, oy[r] + th * d.y, oz[r] + th * d.z};
    const ex::v3<float> N = facing_normal(dpdu, dpdv, r, d);
    const float off = eps * (1.0f + fmaxf(fabsf(S.x), fmaxf(fabsf(S.y), fabsf(S.z))));
    const ex::v3<float> O{S.x + off * N.x, S.y + off * N.y, S.z + off * N.z};
    if (want_shadow) {
      sox[slot] = O.x; soy[slot] = O.y; soz[slot] = O.z;
      sdx[slot] = lx; sdy[slot] = ly; sdz[slot] = lz;
      parent[slot] = (int)r;
      kind[slot] = 0;
      if (out_scene) out_scene[slot] = ray_scene[r];
      slot += n;
    }
    if (want_mirror) {
      const float dn = 2.0f * (d.x * N.x + d.y * N.y + d.z * N.z);
      const ex::v3<float> m = unit(ex::v3<float>{d.x - dn * N.x, d.y - dn * N.y, d.z - dn * N.z});
      sox[slot] = O.x; soy[slot] = O.y; soz[slot] = O.z;
      sdx[slot] = m.x; sdy[slot] = m.y; sdz[slot] = m.z;
      parent[slot] = (int)r;
      kind[slot] = 1;
      if (out_scene) out_scene[slot] = ray_scene[r];
    }
  }
  for (int sft = 32; sft > 0; sft >>= 1) valid += __shfl_xor(valid, sft);
  if ((threadIdx.x & 63) == 0 && valid) atomicAdd(count, valid);
}

// Lambert shading of W*H pixels with spp consecutive rays each:
// albedo * max(0, N.L) on hits (N facing the ray), 0 on misses; spp mean.
__global__ void k_shade_lambert(const float* dx, const float* dy, const float* dz, const int* prim,
                                const float* dpdu, const float* dpdv, int64_t n_pix, int W, int spp, int tiled,
                                float lx, float ly, float lz, float albedo, const int* shadow_hit, float* out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pix) return;
  int64_t pix = p;  // image index of ray group p
  if (tiled) {
    int x, y;
    rtn::tiled_pixel(p, W, x, y);
    pix = (int64_t)y * W + x;
  }
  float acc = 0.0f;
  for (int s = 0; s < spp; s++) {
    const int64_t r = p * spp + s;
    if (prim[r] < 0) continue;
    if (shadow_hit && shadow_hit[r]) continue;
    const ex::v3<float> N = facing_normal(dpdu, dpdv, r, ex::v3<float>{dx[r], dy[r], dz[r]});
    acc += albedo * fmaxf(0.0f, N.x * lx + N.y * ly + N.z * lz);
  }
  out[pix] = acc / (float)spp;
}

}  // namespace

namespace rtn {
// Secondary generation for the frame pipeline: counts into *count (device),
// propagates the scene of the parent ray when ray_scene is given.
void launch_secondaries(const float* const rays[6], const rtn_hits* hits, int64_t n, const float ldir[3], float eps,
                        int want_shadow, int want_mirror, float* const out[6], int* parent, int* kind,
                        unsigned long long* count, const int* ray_scene, int* out_scene, hipStream_t st) {
  if (n <= 0) return;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)cus * 16));
  hipLaunchKernelGGL(k_secondaries, dim3((unsigned)blocks), dim3(256), 0, st, rays[0], rays[1], rays[2],
                     rays[3], rays[4], rays[5], hits->prim, hits->t, hits->dpdu, hits->dpdv, n, ldir[0], ldir[1],
                     ldir[2], eps, want_shadow, want_mirror, out[0], out[1], out[2], out[3], out[4], out[5], parent,
                     kind, count, ray_scene, out_scene);
}
}  // namespace rtn

extern "C" int rtn_rays_silhouette(const rtn_scene* s, int64_t n, uint32_t seed, float max_deg, float t0_lo,
                                   float t0_hi, const rtn_ray_queue* rays, int* aim_patch, float* aim, void* stream) {
  RTN_CHECK(s && s->built && rays && n >= 0, RTN_EINVAL, "rtn_rays_silhouette: bad arguments");
  RTN_CHECK(rays->origin_x && rays->origin_y && rays->origin_z && rays->direction_x && rays->direction_y &&
                rays->direction_z,
            RTN_EINVAL, "rtn_rays_silhouette: null ray buffer");
  RTN_CHECK(max_deg >= 0.f && max_deg < 90.f && t0_lo >= 0.f && t0_hi >= t0_lo, RTN_EINVAL,
            "rtn_rays_silhouette: bad angle or distance range");
  RTN_CHECK(n < ((int64_t)1 << 31), RTN_EINVAL, "rtn_rays_silhouette: too many rays");
  if (n == 0) return RTN_OK;
  hipLaunchKernelGGL(k_rays_silhouette, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     s->cp.ptr, s->cp_off.ptr, s->pq.ptr, (int)s->n_patches, n, seed, max_deg, t0_lo, t0_hi,
                     rays->origin_x, rays->origin_y, rays->origin_z, rays->direction_x, rays->direction_y,
                     rays->direction_z, aim_patch, aim);
  RTN_HIP(hipGetLastError());
  if (!stream) RTN_HIP(hipStreamSynchronize(nullptr));
  return RTN_OK;
}

extern "C" int rtn_rays_secondary(const float* const rays[6], const rtn_hits* hits, int64_t n, const float light[3],
                                  float eps, int want_shadow, int want_mirror, const rtn_ray_queue* out, int* parent,
                                  int* kind, int64_t cap, int64_t* count_host, void* stream) {
  RTN_CHECK(rays && hits && hits->prim && hits->t && hits->dpdu && hits->dpdv && light && out && parent && kind &&
                count_host && n >= 0,
            RTN_EINVAL, "rtn_rays_secondary: bad arguments");
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_rays_secondary: null ray component %d", k);
  RTN_CHECK(want_shadow == 0 || want_shadow == 1, RTN_EINVAL, "rtn_rays_secondary: want_shadow is 0 or 1");
  RTN_CHECK(want_mirror == 0 || want_mirror == 1, RTN_EINVAL, "rtn_rays_secondary: want_mirror is 0 or 1");
  RTN_CHECK(cap >= (int64_t)(want_shadow + want_mirror) * n, RTN_EINVAL,
            "rtn_rays_secondary: capacity below (shadow + mirror) x n");
  *count_host = 0;
  if (n == 0 || want_shadow + want_mirror == 0) return RTN_OK;
  const float ll = std::sqrt(light[0] * light[0] + light[1] * light[1] + light[2] * light[2]);
  RTN_CHECK(ll > 0.f, RTN_EINVAL, "rtn_rays_secondary: zero light direction");
  hipStream_t st = (hipStream_t)stream;
  DevBuf<unsigned long long> cnt;
  RTN_TRY(cnt.alloc(1));
  RTN_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), st));
  hipLaunchKernelGGL(k_secondaries, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rays[0], rays[1], rays[2],
                     rays[3], rays[4], rays[5], hits->prim, hits->t, hits->dpdu, hits->dpdv, n, light[0] / ll,
                     light[1] / ll, light[2] / ll, eps, want_shadow, want_mirror, out->origin_x, out->origin_y,
                     out->origin_z, out->direction_x, out->direction_y, out->direction_z, parent, kind, cnt.ptr,
                     nullptr, nullptr);
  RTN_HIP(hipGetLastError());
  unsigned long long c = 0;
  RTN_HIP(hipMemcpyAsync(&c, cnt.ptr, sizeof(c), hipMemcpyDeviceToHost, st));
  RTN_HIP(hipStreamSynchronize(st));
  *count_host = (int64_t)c;
  return RTN_OK;
}

extern "C" int rtn_shade_lambert(const float* const rays[6], const rtn_hits* hits, int W, int H, int spp,
                                 int ray_flags, const float light[3], float albedo, const int* shadow_hit,
                                 float* image, void* stream) {
  RTN_CHECK(rays && hits && hits->prim && hits->dpdu && hits->dpdv && light && image && W > 0 && H > 0 && spp > 0,
            RTN_EINVAL, "rtn_shade_lambert: bad arguments");
  const int tiled = (ray_flags & RTN_RAYS_TILED) ? 1 : 0;
  RTN_CHECK(!tiled || (W % 8 == 0 && H % 8 == 0), RTN_EINVAL, "rtn_shade_lambert: tiled layout needs W, H % 8 == 0");
  const float ll = std::sqrt(light[0] * light[0] + light[1] * light[1] + light[2] * light[2]);
  RTN_CHECK(ll > 0.f, RTN_EINVAL, "rtn_shade_lambert: zero light direction");
  const int64_t np = (int64_t)W * H;
  hipLaunchKernelGGL(k_shade_lambert, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, (hipStream_t)stream, rays[3],
                     rays[4], rays[5], hits->prim, hits->dpdu, hits->dpdv, np, W, spp, tiled, light[0] / ll,
                     light[1] / ll, light[2] / ll, albedo, shadow_hit, image);
  RTN_HIP(hipGetLastError());
  if (!stream) RTN_HIP(hipStreamSynchronize(nullptr));
  return RTN_OK;
}
