// Internal plumbing shared by the library's translation units: error state,
// HIP checks, device buffers and the scene object.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/rtnurbs.h"

namespace rtn {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define RTN_HIP(expr)                                                     \
  do {                                                                    \
    hipError_t e_ = (expr);                                               \
    if (e_ != hipSuccess) return ::rtn::hip_fail(e_, #expr, __FILE__, __LINE__); \
  } while (0)

#define RTN_CHECK(cond, code, ...)         \
  do {                                     \
    if (!(cond)) {                         \
      ::rtn::set_error(__VA_ARGS__);       \
      return code;                         \
    }                                      \
  } while (0)

#define RTN_TRY(expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != RTN_OK) return rc_; \
  } while (0)

// Owning device buffer (hipMalloc); moves only.
template <class T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&ptr, count * sizeof(T));
    if (e != hipSuccess) { ptr = nullptr; return hip_fail(e, "hipMalloc", __FILE__, __LINE__); }
    n = count;
    return RTN_OK;
  }
  int upload(const T* h, size_t count) {
    RTN_TRY(alloc(count));
    if (count) RTN_HIP(hipMemcpy(ptr, h, count * sizeof(T), hipMemcpyHostToDevice));
    return RTN_OK;
  }
  int download(T* h, size_t count) const {
    if (count) RTN_HIP(hipMemcpy(h, ptr, count * sizeof(T), hipMemcpyDeviceToHost));
    return RTN_OK;
  }
};

// One input surface as added by the caller (host copy until build).
struct HostSurface {
  int n, p, m, q;
  std::vector<float> U, V, Pw;
};

// 4-wide BVH node, structure of arrays inside the node: child boxes as
// lo/hi planes per axis, children as signed indices (>=0 wide node, <0 leaf
// ~first with count in the high bits).  128 bytes = one cache line pair.
struct alignas(16) WideNode {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  int32_t child[4];
  int32_t pad[4];
};
static_assert(sizeof(WideNode) == 128, "wide node is 128 bytes");

// Pixel of position p in the tiled ray layout: 8x8-pixel tiles in row-major
// tile order, Morton (Z) order inside a tile, so 16 consecutive pixels form a
// 4x4 block (W a multiple of 8).
__host__ __device__ inline void tiled_pixel(int64_t p, int W, int& x, int& y) {
  const int64_t tile = p >> 6;
  const int in = (int)(p & 63);
  const int tiles_x = W >> 3;
  const int mx = (in & 1) | ((in >> 1) & 2) | ((in >> 2) & 4);
  const int my = ((in >> 1) & 1) | ((in >> 2) & 2) | ((in >> 3) & 4);
  x = (int)(tile % tiles_x) * 8 + mx;
  y = (int)(tile / tiles_x) * 8 + my;
}

// FP64 patch terms of the bilinear seed per patch (q00, q10, e00, e11, qn, pad).
constexpr int kSeedTermWords = 16;

// Device control words of one trace: the persistent forms' work queues, the
// trace totals (stack-overflow count at [6]) and a flag word.  A scene and a
// batch each own one, allocated when they are built, so a trace allocates
// nothing; traces that overlap on one scene or batch share it and race.
constexpr int kCtlWorkWords = 128;
struct TraceCtl {
  unsigned int work[kCtlWorkWords];
  unsigned long long totals[12];
  int flag;
  int pad[3];
};

// Training-pair output of the wide kernels (converged plain-method solves).
struct PairOut {
  int* ray;
  int* patch;
  float2* uv;
  unsigned long long* count;
  int64_t cap;
  int all;  // 1: also the solves that do not converge (or converge at t <= 0), uv = (-1, -1)
};

}  // namespace rtn

// Scene object behind the opaque C handle.
struct rtn_scene {
  std::vector<rtn::HostSurface> surfaces;
  // pre-decomposed patch groups (rtn_scene_add_patches)
  struct PatchGroup { int np, p, q; std::vector<float> cp; };
  std::vector<PatchGroup> groups;
  bool built = false;

  int64_t n_patches = 0, n_cp = 0, n_nodes = 0, n_wide = 0, max_depth = 0, n_leaves = 0;
  double ms_decomp = 0, ms_bvh = 0, ms_wide = 0;

  // device arrays, BVH order
  rtn::DevBuf<float4> cp;          // all nets, patch-major
  rtn::DevBuf<int> cp_off;         // float4 offset of each net
  rtn::DevBuf<int> pq;             // p | (q << 8)
  rtn::DevBuf<float4> uv;          // (u0,u1,v0,v1)
  rtn::DevBuf<int> patch_id;       // original id
  rtn::DevBuf<int2> knot_idx;      // span start (u,v)
  rtn::DevBuf<float> bounds;       // 6 per patch, BVH order
  rtn::DevBuf<float> delta;        // patchDelta, BVH order
  rtn::DevBuf<float4> corners;     // dehomogenised corners q00,q10,q11,q01 (xyz) packed 3 float4
  rtn::DevBuf<float4> obb;         // oriented box of the dehomogenised net: (axis k, half extent k) x3, (centre, 0)
  rtn::DevBuf<double> seed64;      // FP64 patch terms of the bilinear seed (kSeedTermWords per patch)
  rtn::DevBuf<rtn_bvh_node> nodes; // reference layout
  rtn::DevBuf<rtn::WideNode> wide; // 4-wide SoA nodes
  rtn::DevBuf<rtn::TraceCtl> ctl;  // control words of its traces (allocated at build)
  std::vector<int64_t> deg_hist;   // (RTN_MAX_DEGREE+1)^2
};
